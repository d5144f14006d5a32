"""Splitting of target leaves with much near work (accurate modes).

A leaf whose near work exceeds the budget is summed in chunks on other waves
and the chunk sums are added in chunk order.  Every near leaf is still summed
exactly once with the same per-leaf arithmetic, so split and unsplit results
differ only by the order of the FP64 additions; a split result does not depend
on scheduling (grid, persistent, graph) nor on whether the checkpoint buffer
had room (without room the fused wave runs its chunks itself with the same
arithmetic as the summing pass)."""
import sys

import numpy as np
import pytest

from conftest import rel_err, workloads, fmm_module

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def fmm():
    m = fmm_module()
    yield m
    m.set_split(8192)
    m.set_persistent(0, False)
    m.set_mode(m.MODE_TREECODE)
    m.set_fmm_params()


def _say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _dense(n=60000, seed=61):
    wl = workloads()
    # two tight clusters and a core radius several times the blob spacing:
    # near lists of thousands of sources per target leaf
    return wl.gaussian_mixture(n, seed=seed, k=2, std_lo=2e-3, std_hi=8e-3, core=1e-3)


def _eval(fmm, mode, budget, xy, cir, cor, blocks=0, graph=False):
    fmm.set_mode(mode)
    fmm.set_fmm_params(16, 64, 0.5, 4.0)
    fmm.set_split(budget)
    fmm.set_persistent(blocks, graph)
    fmm.make_tree(xy, cir, cor)
    return fmm.eval_sources()


@pytest.mark.parametrize("mode", [1, 2])
def test_split_matches_unsplit(fmm, mode):
    xy, cir, cor = _dense()
    _say(f"[split] mode {mode} N={len(xy)}: budgets 0, 256, 3")
    u0 = _eval(fmm, mode, 0, xy, cir, cor)
    scale = np.abs(u0).max()
    for budget in (256, 3):
        u = _eval(fmm, mode, budget, xy, cir, cor)
        d = np.abs(u - u0).max() / scale
        _say(f"  budget {budget}: max |split - unsplit| / max|u| = {d:.2e}")
        assert np.all(np.isfinite(u))
        assert d <= 1e-12, (budget, d)
    fmm.set_split(8192)


@pytest.mark.parametrize("mode", [1, 2])
def test_split_deterministic_across_paths(fmm, mode):
    """Budget 3 makes more checkpoints than the record buffer holds, so some
    leaves run their chunks in the fused wave; which ones depends on timing,
    the result must not."""
    xy, cir, cor = _dense(seed=62)
    cus = fmm.device_cus()
    runs = [("grid", 0, False), ("grid again", 0, False), ("persistent", 2 * cus, False),
            ("graph", 2 * cus, True), ("graph replay", 2 * cus, True)]
    ref = None
    for name, blocks, graph in runs:
        _say(f"[split] mode {mode} budget 3 {name}")
        u = _eval(fmm, mode, 3, xy, cir, cor, blocks, graph)
        if ref is None:
            ref = u
        else:
            assert np.array_equal(u, ref), name
    fmm.set_persistent(0, False)
    fmm.set_split(8192)


def test_split_accuracy_vs_direct(fmm, oracle):
    wl = workloads()
    xy, cir, cor = _dense(n=40000, seed=63)
    idx = wl.sample_targets(len(xy), 1200, seed=63)
    d = oracle.direct(xy, cir, cor, xy[idx], style=1)
    for mode, tol in ((1, 5e-7), (2, 1e-6)):
        u = _eval(fmm, mode, 512, xy, cir, cor)
        l2, mx = rel_err(u[idx], d)
        _say(f"[split] mode {mode} budget 512: rel_L2={l2:.2e} rel_max={mx:.2e}")
        assert l2 <= tol
    fmm.set_split(8192)
