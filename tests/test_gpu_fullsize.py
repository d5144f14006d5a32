"""The benchmark configurations at their full sizes, through the C ABI.

* 4M Gaussian mixture (leaf 64): the whole tree bit-exact against the oracle
  (ids, levels, centres, regions, counts, leaf flags, leaf member lists), the
  multipole coefficients of sampled nodes to 1e-12, and the error ladder of
  the mixed mode over p = 8, 12, 16 against the FP64 direct sum;
* 1M vortex sheet (p = 16, leaf 64): the tree bit-exact against the oracle,
  treecode velocities to 1e-12 of the oracle, and the plain FP64 accurate mode
  against the FP64 direct sum on 10,000 sampled targets;
* 256 systems of 2e3..2e5 blobs in one forest (mixed, p = 12): per-system
  velocities equal the single-system path bit for bit, and their error
  against each system's FP64 direct sum is bounded;
* the 2M elliptic patch of the learned form: a few forward-Euler steps with
  plain-FMM velocities, one labelled step, and the shipped network within
  1e-6 of the FP64 direct sum on a later (held-out) step.
"""
import importlib
import os
import sys

import numpy as np
import pytest

from conftest import ROOT, fmm_module, rel_err, workloads

pytestmark = pytest.mark.gpu


def _say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def fmm():
    m = fmm_module()
    m.set_prefetch(False)
    yield m
    m.set_prefetch(True)
    m.set_learned(None)
    m.set_mode(m.MODE_TREECODE)
    m.set_fmm_params()


def _leaf_runs(lo, count, members, is_leaf):
    """Concatenated member lists of the leaves, in node order."""
    lo = np.asarray(lo, np.int64)[:-1][is_leaf == 1]
    c = np.asarray(count, np.int64)[is_leaf == 1]
    start = np.repeat(lo - np.concatenate([[0], np.cumsum(c)[:-1]]), c)
    return members[start + np.arange(int(c.sum()))]


def _tree_vs_oracle(fmm, oracle, xy, cir, cor, nterms, num_p, theta, coef_sample=4000):
    n = xy.shape[0]
    fmm.set_mode(fmm.MODE_TREECODE)
    fmm.set_params(nterms, num_p, theta)
    fmm.make_tree(xy, cir, cor)
    t = fmm.dump_tree(n, nterms)
    T = oracle.Tree(xy, cir, cor, nterms, num_p, theta)
    nd = T.nodes()
    _say(f"  nodes={len(nd['gid'])} depth={T.depth} (GPU {fmm.tree_size()['depth']})")
    assert len(t["gid"]) == len(nd["gid"])
    for k in ("gid", "level", "count", "is_leaf"):
        assert np.array_equal(t[k], nd[k]), k
    for k in ("cx", "cy", "region"):
        assert np.array_equal(t[k].view(np.uint64), nd[k].view(np.uint64)), k
    leaf = nd["is_leaf"]
    got = _leaf_runs(t["leaf_lo"], t["count"], t["leaf_members"], leaf)
    want = _leaf_runs(nd["lo"], nd["count"], nd["members"], leaf)
    assert np.array_equal(got, want)
    assert np.array_equal(np.sort(got), np.arange(n))
    # coefficients a_k = sum_j cir_j (z_j - c)^k of sampled nodes: the GPU sums
    # in another order than the oracle, so the difference is measured against
    # the sum's own magnitude, sum_j |cir_j| |z_j - c|^k (each node's blobs
    # are one run of the oracle's member array)
    rng = np.random.default_rng(n)
    pick = rng.choice(len(nd["gid"]), min(coef_sample, len(nd["gid"])), replace=False)
    ca, cb = t["coeffs"][pick], T.coeffs()[pick]
    k = np.arange(nterms)
    err = 0.0
    for r, i in enumerate(pick):
        ids = nd["members"][nd["lo"][i]:nd["lo"][i] + nd["count"][i]]
        if len(ids) == 0:
            assert np.all(ca[r] == 0)
            continue
        d = np.hypot(xy[ids, 0] - nd["cx"][i], xy[ids, 1] - nd["cy"][i])
        mag = (np.abs(cir[ids])[:, None] * d[:, None] ** k[None, :]).sum(axis=0)
        mag[mag == 0] = 1.0
        err = max(err, float((np.abs(ca[r] - cb[r]) / mag).max()))
    _say(f"  coefficients of {len(pick)} sampled nodes: max |diff| / sum|terms| {err:.2e}")
    assert err <= 1e-12
    return T


def test_mixture_4m_tree_bitexact(fmm, oracle):
    """Headline size: the 4M-blob mixture tree (leaf 64, p = 16) equals the
    oracle's bit for bit."""
    n = 4_000_000
    _say(f"[fullsize] 4M mixture tree vs oracle: N={n} p=16 leaf 64")
    xy, cir, cor = workloads().gaussian_mixture(n, seed=3)
    _tree_vs_oracle(fmm, oracle, xy, cir, cor, 16, 64, 0.5)


def test_sheet_1m_tree_and_fmm64(fmm, oracle):
    """Vortex-sheet configuration (1M, p = 16, leaf 64): tree and treecode
    velocities against the oracle; the plain FP64 accurate mode against the
    FP64 direct sum on 10,000 sampled targets."""
    wl = workloads()
    n = 1_000_000
    _say(f"[fullsize] 1M sheet: N={n} p=16 leaf 64")
    xy, cir, cor = wl.sheet(n, seed=2)
    T = _tree_vs_oracle(fmm, oracle, xy, cir, cor, 16, 64, 0.5)
    idx = wl.sample_targets(n, 10_000, seed=5)
    uv = fmm.eval_sources()
    want = T.eval(xy[idx[:1000]])
    l2, mx = rel_err(uv[idx[:1000]], want)
    _say(f"  treecode vs oracle on 1000 targets: rel_L2={l2:.2e} rel_max={mx:.2e}")
    assert mx <= 1e-12
    ref = fmm.direct(xy, cir, cor, xy[idx], style=1)
    out = {}
    for mode, name in ((fmm.MODE_FMM64, "fmm64"), (fmm.MODE_MIXED, "mixed")):
        fmm.set_mode(mode)
        fmm.set_fmm_params(16, 64, 0.5, 4.0)
        fmm.make_tree(xy, cir, cor)
        out[name] = rel_err(fmm.eval_sources()[idx], ref)
        _say(f"  {name} vs FP64 direct on {len(idx)} targets: rel_L2={out[name][0]:.2e} "
             f"rel_max={out[name][1]:.2e}")
    # the cut at 4 core radii leaves exp(-16) = 1.1e-7 of each far pair's core term
    assert out["fmm64"][0] <= 5e-7 and out["fmm64"][1] <= 1e-6
    assert out["mixed"][0] <= 1e-6


def test_mixture_4m_error_ladder(fmm):
    """Precision form at 4M: the mixed mode's error against the FP64 direct sum
    on 10,000 sampled targets for p = 8, 12, 16 falls with p, and p = 16 meets
    1e-6; the plain FP64 mode at p = 16 is below the mixed one."""
    wl = workloads()
    n = 4_000_000
    _say(f"[fullsize] 4M mixture error ladder: N={n}")
    xy, cir, cor = wl.gaussian_mixture(n, seed=3)
    idx = wl.sample_targets(n, 10_000, seed=7)
    ref = fmm.direct(xy, cir, cor, xy[idx], style=1)
    err = {}
    fmm.set_mode(fmm.MODE_MIXED)
    for p in (8, 12, 16):
        fmm.set_fmm_params(p, 64, 0.5, 4.0)
        fmm.make_tree(xy, cir, cor)
        err[p] = rel_err(fmm.eval_sources()[idx], ref)
        _say(f"  mixed p={p}: rel_L2={err[p][0]:.3e} rel_max={err[p][1]:.3e}")
    fmm.set_mode(fmm.MODE_FMM64)
    fmm.set_fmm_params(16, 64, 0.5, 4.0)
    fmm.make_tree(xy, cir, cor)
    e64 = rel_err(fmm.eval_sources()[idx], ref)
    _say(f"  fmm64 p=16: rel_L2={e64[0]:.3e} rel_max={e64[1]:.3e}")
    assert err[8][0] > err[12][0] > err[16][0]
    assert err[16][0] <= 1e-6 and err[16][1] <= 1e-5
    assert e64[0] <= err[16][0]


def test_batch_256_systems(fmm):
    """Batching form at its configured size: 256 systems, N log-uniform in
    [2e3, 2e5], uniform or 2-cluster layouts, core 1e-3, mixed mode, p = 12."""
    wl = workloads()
    systems = wl.batch_systems(256, seed=4)
    fmm.set_mode(fmm.MODE_MIXED)
    fmm.set_fmm_params(12, 64, 0.5, 4.0)
    off = fmm.make_forest(systems)
    sizes = np.diff(off)
    _say(f"[fullsize] 256 systems: N={off[-1]} (min {sizes.min()}, max {sizes.max()})")
    uv = fmm.eval_sources()
    # bit for bit against the single-system path on 8 systems (smallest,
    # largest and six spread between)
    order = np.argsort(sizes)
    pick = sorted(set([int(order[0]), int(order[-1])] + [int(order[k]) for k in
                                                        np.linspace(1, 254, 6).astype(int)]))
    for s in pick:
        xy, cir, cor = systems[s]
        fmm.make_tree(xy, cir, cor)
        one = fmm.eval_sources()
        assert np.array_equal(uv[off[s]:off[s + 1]].view(np.uint64), one.view(np.uint64)), s
    errs = []
    for s in range(0, 256, 8):
        xy, cir, cor = systems[s]
        t = np.arange(0, xy.shape[0], max(1, xy.shape[0] // 300))
        ref = fmm.direct(xy, cir, cor, xy[t], style=1)
        errs.append(rel_err(uv[off[s] + t], ref)[0])
    errs = np.array(errs)
    _say(f"  per-system rel_L2 vs FP64 direct (32 systems): median {np.median(errs):.2e} "
         f"worst {errs.max():.2e}")
    assert np.median(errs) <= 1e-6 and errs.max() <= 5e-6


def test_patch_2m_learned_pipeline(fmm):
    """Learned form: the 2M elliptic patch advected 3 forward-Euler steps
    (dt = 1e-3) with plain-FMM velocities; step 1 labelled (the smallest cut,
    then p, meeting 1e-6 per leaf, for 200 sampled leaves); the shipped
    network evaluated on step 2, which it has not seen, within 1e-6 of the
    FP64 direct sum."""
    learned = importlib.import_module("2dfmm_amd.learned")
    wl = workloads()
    n = 2_000_000
    _say(f"[fullsize] 2M elliptic patch: N={n}")
    xy, cir, cor = wl.elliptic_patch(n)
    with np.load(os.path.join(ROOT, "2dfmm_amd", "learned_weights.npz")) as z:
        w, pm, cm = z["weights"], float(z["p_margin"]), float(z["cut_margin"])
    for step in range(3):
        if step == 1:
            f, lab, leaves = learned.label_tree(fmm, xy, cir, cor, 1e-6, 200, seed=1)
            _say(f"  step 1 labelled: {len(f)} leaves, mean p* {lab[:, 0].mean():.2f}, "
                 f"mean cut* {lab[:, 1].mean():.2f}")
            assert f.shape == (len(leaves), 8) and np.isfinite(f).all()
            assert set(np.unique(lab[:, 0])) <= set(learned.PS)
            assert set(np.unique(lab[:, 1])) <= set(learned.CUTS)
        if step == 2:
            r = learned.evaluate(fmm, xy, cir, cor, w, pm, cm, judge=4000, seed=2, reps=1)
            _say(f"  step 2 (held out): fixed {r['fixed']}  learned {r['learned']}")
            assert r["learned"]["rel_L2"] <= 1e-6
            assert r["learned"]["rel_max"] <= 1e-5
            break
        fmm.set_mode(fmm.MODE_FMM64)
        fmm.set_fmm_params(16, 64, 0.5, 4.0)
        fmm.set_learned(None)
        fmm.make_tree(xy, cir, cor)
        xy = xy + 1e-3 * fmm.eval_sources()
