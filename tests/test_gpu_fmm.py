"""The accurate modes (FP64 and mixed precision) against the FP64 direct sum.

These modes are not the reference's algorithm; their judge is the regularised
direct sum of fmm.c:121-142 over all sources (cut-off 1e-10), computed by the
oracle on the CPU for small N and by the GPU direct kernel (itself checked
against the oracle in test_gpu_parity.py) for large N."""
import sys

import numpy as np
import pytest

from conftest import rel_err, workloads, fmm_module

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def fmm():
    m = fmm_module()
    yield m
    m.set_mode(m.MODE_TREECODE)
    m.set_fmm_params()


def _say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.mark.parametrize("fp32,tol", [(False, 1e-12), (True, 1e-4)])
def test_mfma_lane_layout(fmm, fp32, tol):
    assert fmm.selftest_mfma(fp32) <= tol


def _run(fmm, mode, xy, cir, cor, p=16, leaf=64, theta=0.5, cut=4.0):
    fmm.set_mode(mode)
    fmm.set_fmm_params(p, leaf, theta, cut)
    fmm.make_tree(xy, cir, cor)
    return fmm.eval_sources()


@pytest.mark.parametrize("kind,n,seed", [("uniform", 20000, 51), ("mixture", 40000, 52),
                                         ("sheet", 30000, 53)])
def test_fmm64_vs_direct(fmm, oracle, kind, n, seed):
    wl = workloads()
    xy, cir, cor = {"uniform": wl.uniform_square, "sheet": wl.sheet,
                    "mixture": wl.gaussian_mixture}[kind](n, seed=seed)
    idx = wl.sample_targets(n, 1500, seed=seed)
    d = oracle.direct(xy, cir, cor, xy[idx], style=1)
    _say(f"[fmm64] {kind} N={n}")
    uv = _run(fmm, fmm.MODE_FMM64, xy, cir, cor)
    l2, mx = rel_err(uv[idx], d)
    _say(f"  FP64 FMM p=16: rel_L2={l2:.2e} rel_max={mx:.2e}")
    # the 4-core-radius cut leaves exp(-16) = 1.1e-7 of each far pair's core term
    assert l2 <= 5e-7
    uvm = _run(fmm, fmm.MODE_MIXED, xy, cir, cor)
    l2m, mxm = rel_err(uvm[idx], d)
    _say(f"  mixed FMM p=16: rel_L2={l2m:.2e} rel_max={mxm:.2e}")
    assert l2m <= 1e-6


def test_fmm_converges_in_p(fmm, oracle):
    wl = workloads()
    xy, cir, cor = wl.gaussian_mixture(30000, seed=54)
    idx = wl.sample_targets(30000, 1000, seed=54)
    d = oracle.direct(xy, cir, cor, xy[idx], style=1)
    errs = []
    for p in (4, 8, 12, 16):
        uv = _run(fmm, fmm.MODE_FMM64, xy, cir, cor, p=p)
        errs.append(rel_err(uv[idx], d)[0])
    _say(f"  rel_L2 by p 4/8/12/16: {['%.2e' % e for e in errs]}")
    assert errs[0] > errs[1] > errs[2] and errs[3] <= errs[2]
    assert errs[3] < 5e-7


def test_fmm_cut_radius_matters(fmm, oracle):
    """Without the core cut-off the far field is the singular kernel next to
    the target (as in the treecode) and the error against the regularised
    direct sum is large; with 4 core radii it is small."""
    wl = workloads()
    xy, cir, cor = wl.uniform_square(20000, seed=55)
    cor = np.full(20000, 0.05)
    idx = wl.sample_targets(20000, 800, seed=55)
    d = oracle.direct(xy, cir, cor, xy[idx], style=1)
    e0 = rel_err(_run(fmm, fmm.MODE_FMM64, xy, cir, cor, leaf=16, cut=0.0)[idx], d)[0]
    e4 = rel_err(_run(fmm, fmm.MODE_FMM64, xy, cir, cor, leaf=16, cut=4.0)[idx], d)[0]
    _say(f"  cut 0: {e0:.2e}   cut 4 sigma: {e4:.2e}")
    assert e0 > 1e-3 and e4 < 1e-6


def test_fmm_large_vs_gpu_direct(fmm):
    wl = workloads()
    n = 400_000
    xy, cir, cor = wl.gaussian_mixture(n, seed=56)
    idx = wl.sample_targets(n, 3000, seed=56)
    d = fmm.direct(xy, cir, cor, xy[idx], style=1)
    for mode, tol in ((fmm.MODE_FMM64, 5e-7), (fmm.MODE_MIXED, 1e-6)):
        uv = _run(fmm, mode, xy, cir, cor)
        l2, mx = rel_err(uv[idx], d)
        _say(f"  400k mixture mode {mode}: rel_L2={l2:.2e} rel_max={mx:.2e}")
        assert l2 <= tol
    fmm.set_mode(fmm.MODE_TREECODE)


def test_fmm_mode_biot_tree(fmm):
    """make_tree's prefetch uses the accurate mode; biot_tree at blobs returns it."""
    wl = workloads()
    xy, cir, cor = wl.uniform_square(5000, seed=57)
    uv = _run(fmm, fmm.MODE_MIXED, xy, cir, cor)
    for i in range(0, 5000, 499):
        assert fmm.biot_tree(*xy[i]) == tuple(uv[i])
    fmm.set_mode(fmm.MODE_TREECODE)


def test_fmm_varying_core_radii(fmm, oracle):
    """Core radii that differ between blobs use the general FP32 encoding
    (w = log2(e)/sigma^2 per source); equal radii use coordinates scaled by
    sqrt(w).  Both against the direct sum, and the two encodings against each
    other on one layout (one radius nudged by 1e-12 selects the general one)."""
    wl = workloads()
    n = 30000
    xy, cir, _ = wl.gaussian_mixture(n, seed=58)
    rng = np.random.default_rng(58)
    cor = 5e-4 + 1.5e-3 * rng.random(n)
    idx = wl.sample_targets(n, 1200, seed=58)
    d = oracle.direct(xy, cir, cor, xy[idx], style=1)
    for mode, tol in ((fmm.MODE_FMM64, 5e-7), (fmm.MODE_MIXED, 1e-6)):
        l2, mx = rel_err(_run(fmm, mode, xy, cir, cor)[idx], d)
        _say(f"  varying radii mode {mode}: rel_L2={l2:.2e} rel_max={mx:.2e}")
        assert l2 <= tol
    cu = np.full(n, 1e-3)
    cn = cu.copy()
    cn[0] *= 1.0 + 1e-12
    uu = _run(fmm, fmm.MODE_MIXED, xy, cir, cu)
    un = _run(fmm, fmm.MODE_MIXED, xy, cir, cn)
    diff = np.abs(uu - un).max() / np.abs(uu).max()
    _say(f"  equal-radius vs general FP32 encoding: max diff {diff:.2e}")
    assert diff <= 1e-5
    du = oracle.direct(xy, cir, cu, xy[idx], style=1)
    assert rel_err(uu[idx], du)[0] <= 1e-6
    fmm.set_mode(fmm.MODE_TREECODE)


@pytest.mark.parametrize("mode", [1, 2])
def test_points_off_blobs_vs_direct(fmm, oracle, mode):
    """Accurate modes at points that are not blobs (the solver's sheet centres,
    e_velo_mpi.f90:59-70; points in empty space; points outside the root
    square) against the FP64 direct sum; at blob positions the point walk and
    the target-leaf kernel agree to the accuracy of both."""
    wl = workloads()
    n = 40000
    xy, cir, cor = wl.sheet(n, seed=71)
    mxy, mcir, mcor = wl.gaussian_mixture(n, seed=72)
    xy = np.concatenate([xy, mxy])
    cir = np.concatenate([cir, mcir])
    cor = np.concatenate([cor, np.full(n, 1e-3)])
    centres = 0.5 * (xy[:n - 1][::40] + xy[1:n][::40])
    rng = np.random.default_rng(73)
    empty = rng.random((500, 2))
    outside = np.array([[-0.3, 0.5], [1.7, -0.2], [0.5, 2.5], [-1.0, -1.0]])
    pts = np.concatenate([centres, empty, outside])
    fmm.set_mode(mode)
    fmm.set_fmm_params(16, 64, 0.5, 4.0)
    fmm.make_tree(xy, cir, cor)
    uv_src = fmm.eval_sources()
    got = fmm.eval_batch(pts)
    d = oracle.direct(xy, cir, cor, pts, style=1)
    l2, mx = rel_err(got, d)
    _say(f"[points] mode {mode}: {len(pts)} off-blob points vs FP64 direct: rel_L2={l2:.2e} "
         f"rel_max={mx:.2e}")
    assert l2 <= 5e-7 and mx <= 5e-6
    for j in range(0, len(pts), 97):
        assert fmm.biot_tree(*pts[j]) == tuple(got[j])
    idx = wl.sample_targets(2 * n, 500, seed=74)
    at_blobs = fmm.eval_batch(xy[idx])
    l2b, _ = rel_err(at_blobs, uv_src[idx])
    _say(f"  point walk vs target-leaf kernel at 500 blobs: rel_L2={l2b:.2e}")
    assert l2b <= 1e-6
    fmm.set_mode(fmm.MODE_TREECODE)


@pytest.mark.parametrize("mode", [1, 2])
def test_accurate_biot_tree_without_prefetch(fmm, mode):
    """With the prefetch off, and for trees built from device arrays or with
    mirror images, biot_tree at a blob returns the accurate all-blob value
    (never the treecode rule)."""
    import torch

    wl = workloads()
    n = 6000
    xy, cir, cor = wl.gaussian_mixture(n, seed=75)
    ref = _run(fmm, mode, xy, cir, cor)
    fmm.set_prefetch(False)
    try:
        fmm.make_tree(xy, cir, cor)
        for i in range(0, n, 251):
            assert fmm.biot_tree(*xy[i]) == tuple(ref[i])
        dev = torch.device("cuda:0")
        d_xy = torch.from_numpy(xy).to(dev)
        d_cir = torch.from_numpy(cir).to(dev)
        d_cor = torch.from_numpy(cor).to(dev)
        torch.cuda.synchronize()
        fmm.make_tree_device(n, d_xy.data_ptr(), d_cir.data_ptr(), d_cor.data_ptr())
        rng = np.random.default_rng(76)
        for i in rng.permutation(n)[:40]:
            assert fmm.biot_tree(*xy[i]) == tuple(ref[i])
        # mirror images: the real blobs' values, looked up by their positions
        fmm.make_tree_mirror_device(n, d_xy.data_ptr(), d_cir.data_ptr(), d_cor.data_ptr(), -0.5)
        m_uv = fmm.eval_sources()
        for i in range(0, n, 601):
            assert fmm.biot_tree(*xy[i]) == tuple(m_uv[i])
    finally:
        fmm.set_prefetch(True)
        fmm.set_mode(fmm.MODE_TREECODE)


@pytest.mark.parametrize("mode", [1, 2])
def test_l2l_parent_expansions(fmm, oracle, mode):
    """Local expansions inherited from the parent: with them on, the target
    leaves contract far fewer cells (the parents take the cells far from
    them once), the accuracy against the FP64 direct sum stays where it was,
    and the results do not depend on the launch shape (grid, persistent,
    graph replay)."""
    wl = workloads()
    n = 60000
    xy, cir, cor = wl.gaussian_mixture(n, seed=81)
    idx = wl.sample_targets(n, 1500, seed=81)
    d = oracle.direct(xy, cir, cor, xy[idx], style=1)
    out = {}
    for on in (False, True):
        fmm.set_l2l(on)
        uv = _run(fmm, mode, xy, cir, cor)
        w = fmm.count_work_detail()
        out[on] = (rel_err(uv[idx], d), w, uv)
        _say(f"[l2l] mode {mode} l2l={on}: rel_L2={out[on][0][0]:.2e} rel_max={out[on][0][1]:.2e} "
             f"leaf far={w['far']} parent far={w['parent_far']} parents={w['parents']} "
             f"shifts={w['l2l_shifts']}")
    fmm.set_l2l(True)
    e0, w0, _ = out[False]
    e1, w1, u1 = out[True]
    assert w0["parent_far"] == 0 and w1["parents"] > 0 and w1["l2l_shifts"] >= 2 * w1["parents"]
    assert w1["far"] <= 0.7 * w0["far"]
    assert w1["far"] + w1["parent_far"] < w0["far"]
    assert w1["pairs"] == w0["pairs"] and w1["smooth"] == w0["smooth"]
    assert e1[0] <= max(1.5 * e0[0], 2e-7) and e1[1] <= max(1.5 * e0[1], 5e-7)
    for blocks, graph in ((256, False), (256, True)):
        fmm.set_persistent(blocks, graph)
        fmm.make_tree(xy, cir, cor)
        got = fmm.eval_sources()
        assert np.array_equal(got.view(np.uint64), u1.view(np.uint64)), (blocks, graph)
    fmm.set_persistent(0, False)
    fmm.set_mode(fmm.MODE_TREECODE)


@pytest.mark.parametrize("mode", [1, 2])
def test_start_lists_same_interactions(fmm, oracle, mode):
    """The children of a parent expansion walk from the parent's start list
    (default) or from the root: both must meet the same far cells and near
    items, so the velocities agree to the rounding of a different summation
    order (FP64: 1e-12; mixed: FP32 pair sums grouped differently), with
    heavy leaves split into continuation chunks (small split budget) and on
    the persistent path.  A missing or repeated item would move some
    targets by far more."""
    wl = workloads()
    n = 60000
    xy, cir, cor = wl.gaussian_mixture(n, seed=82)
    idx = wl.sample_targets(n, 1500, seed=82)
    d = oracle.direct(xy, cir, cor, xy[idx], style=1)
    tol = 1e-12 if mode == 1 else 1e-6
    try:
        fmm.set_split(256)
        out = {}
        # off, lists of at most 12 entries (most parents fall back to the
        # root walk, the rest keep their lists), and the full lists
        for cap in (0, 12, 128):
            fmm.set_start_lists(cap)
            out[cap] = _run(fmm, mode, xy, cir, cor)
            e = rel_err(out[cap][idx], d)
            _say(f"[start lists] mode {mode} cap={cap}: rel_L2={e[0]:.2e} rel_max={e[1]:.2e}")
            assert e[0] <= 5e-7
        scale = np.abs(out[0]).max()
        for cap in (12, 128):
            diff = np.abs(out[cap] - out[0]).max() / scale
            _say(f"[start lists] mode {mode}: max |cap {cap} - off| / max |u| = {diff:.2e}")
            assert diff <= tol
        out[True] = out[128]
        fmm.set_split(0)
        whole = _run(fmm, mode, xy, cir, cor)
        assert np.abs(whole - out[True]).max() / scale <= tol
        fmm.set_split(256)
        fmm.set_persistent(256, False)
        fmm.make_tree(xy, cir, cor)
        got = fmm.eval_sources()
        assert np.array_equal(got.view(np.uint64), out[True].view(np.uint64))
    finally:
        fmm.set_start_lists(True)
        fmm.set_persistent(0, False)
        fmm.set_split(8192)
        fmm.set_mode(fmm.MODE_TREECODE)


