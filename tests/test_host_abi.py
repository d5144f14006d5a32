"""CPU-side checks of the product: the C-ABI library loads and exports every
function include/fmm2d.h declares (no compute calls: there may be no GPU),
the snapshot reader/writer, and the seeded workload generators."""
import ctypes
import os
import re
import struct
import subprocess

import numpy as np
import pytest

from conftest import ROOT, sfile, workloads

HEADER = os.path.join(ROOT, "include", "fmm2d.h")
LIB = os.path.join(ROOT, "2dfmm_amd", "libfmm2d.so")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"^\s*(?:void|int|const char\s*\*)\s*\*?\s*(\w+)\s*\(", text,
                                 flags=re.M)))


def test_header_declares_reference_abi():
    fns = declared_functions()
    assert "make_tree" in fns and "biot_tree" in fns
    assert len(fns) >= 20
    text = open(HEADER).read()
    assert re.search(r"void make_tree\(int nvor_b, const double \*vor_b, const double \*cir_b,"
                     r"\s*const double \*cor_b\);", text)
    assert "void biot_tree(double x, double y, double *u, double *v);" in text


def _ensure_lib():
    if not os.path.exists(LIB):
        subprocess.run(["python", os.path.join(ROOT, "2dfmm_amd", "_build.py")], check=True)


def test_library_exports_every_declared_symbol():
    _ensure_lib()
    lib = ctypes.CDLL(LIB)  # loading does not touch the GPU
    for name in declared_functions():
        assert hasattr(lib, name), name
    out = subprocess.run(["nm", "-D", "--defined-only", LIB], capture_output=True, text=True,
                         check=True).stdout
    exported = {line.split()[-1] for line in out.splitlines() if line.strip()}
    for name in declared_functions():
        assert name in exported, name
    # the two reference symbols are unmangled, as bind(c) without name= expects
    assert "make_tree" in exported and "biot_tree" in exported


def test_library_has_gfx950_code():
    _ensure_lib()
    out = subprocess.run(["/opt/rocm/bin/roc-obj-ls", LIB], capture_output=True, text=True)
    if out.returncode != 0:
        pytest.skip("roc-obj-ls not available")
    assert "gfx950" in out.stdout


def _fake_snapshot(path, n, seed):
    rng = np.random.default_rng(seed)
    recs = [struct.pack("<2d", 3000.0, 0.015), struct.pack("<d", 0.5), struct.pack("<i", 0)]
    xy = rng.random((n, 2))
    vel = rng.random((n, 3, 2))
    cir = rng.standard_normal(n)
    cor = rng.random(n) * 1e-2
    ids = np.arange(n, dtype=np.int32) + 7
    recs.append(struct.pack("<i", n))
    for j in range(n):
        recs += [struct.pack("<2d", *xy[j])] + [struct.pack("<2d", *vel[j, k]) for k in range(3)]
        recs += [struct.pack("<d", cir[j]), struct.pack("<d", cor[j]), struct.pack("<i", ids[j])]
    recs += [struct.pack("<i", 2), struct.pack("<2d", 1.0, 2.0)]
    with open(path, "wb") as f:
        for r in recs:
            m = struct.pack("<i", len(r))
            f.write(m + r + m)
    return xy, vel, cir, cor, ids


def test_snapshot_read_write_roundtrip(tmp_path):
    sf = sfile()
    p = tmp_path / "s_00001_step.dat"
    xy, vel, cir, cor, ids = _fake_snapshot(str(p), 50, 1)
    s = sf.read_snapshot(str(p))
    assert np.array_equal(s.xy, xy) and np.array_equal(s.vel, vel)
    assert np.array_equal(s.cir, cir) and np.array_equal(s.cor, cor) and np.array_equal(s.blob_id, ids)
    s.vel[:, 0] = 1.25
    q = tmp_path / "s_00002_step.dat"
    sf.write_snapshot(str(q), s)
    t = sf.read_snapshot(str(q))
    assert (t.vel[:, 0] == 1.25).all() and np.array_equal(t.xy, xy)
    assert open(p, "rb").read()[:40] == open(q, "rb").read()[:40]
    assert sf.read_records(str(q))[-1] == struct.pack("<2d", 1.0, 2.0)


def test_snapshot_reader_on_reference_file():
    path = "/root/reference/ver-c/s_file/s_00198_step.dat"
    if not os.path.exists(path):
        pytest.skip("reference snapshot not present")
    from conftest import load_golden
    s = sfile().read_snapshot(path)
    g = load_golden("s00198_refparams")
    assert np.array_equal(s.xy, g["xy"]) and np.array_equal(s.cir, g["cir"])


def test_workloads_deterministic_and_shaped():
    wl = workloads()
    a = wl.gaussian_mixture(1000, seed=3)
    b = wl.gaussian_mixture(1000, seed=3)
    for x, y in zip(a, b):
        assert np.array_equal(x, y)
    xy, cir, cor = wl.uniform_square(5000, seed=1)
    assert xy.shape == (5000, 2) and (xy >= 0).all() and (xy < 1).all()
    assert abs(cir.std() - 1e-3) < 1e-4 and (cor == 0.01).all()
    xy, cir, cor = wl.sheet(10000, seed=2)
    assert np.abs(xy[:, 1] - 0.01 * np.sin(2 * np.pi * xy[:, 0])).max() < 1e-3
    assert np.allclose(cir, 1e-4)
    u = wl.uniform(5, 6, 100000)
    assert abs(u.mean() - 0.5) < 0.01 and u.min() >= 0 and u.max() < 1
    z = wl.normal(5, 6, 100000)
    assert abs(z.mean()) < 0.02 and abs(z.std() - 1) < 0.02
    s = wl.sample_targets(1000, 100, seed=1)
    assert len(np.unique(s)) == 100 and (np.diff(s) > 0).all()
    systems = wl.batch_systems(8, seed=4, scale=0.01)
    assert len(systems) == 8 and all(x[0].shape[0] >= 1 for x in systems)
    xy, cir, _ = wl.elliptic_patch(20000)
    assert abs(cir.sum() - 1.0) < 1e-12
    assert ((xy[:, 0] - 0.5) ** 2 / 0.09 + (xy[:, 1] - 0.5) ** 2 / 0.0225 <= 1 + 1e-12).all()
