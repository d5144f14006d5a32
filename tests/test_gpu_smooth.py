"""Smooth cells of the mixed mode (one core radius).

A near node whose half side is at most eta core radii is summed for every
target through a degree-12 Taylor expansion of the regularised kernel about its
centre (coefficients from `smooth_coef_kernel`) instead of pair by pair.  The
expansion converges for any target position, inside the cell too, so the
result must stay within the mixed mode's tolerance of the FP64 direct sum;
with eta 0 the library sums every near pair as before."""
import sys

import numpy as np
import pytest

from conftest import rel_err, workloads, fmm_module

pytestmark = pytest.mark.gpu

# the library's default: nodes of half side <= 0.5 core radii
ETA = 0.5


@pytest.fixture(scope="module")
def fmm():
    m = fmm_module()
    yield m
    m.set_smooth(ETA, 48)
    m.set_split(8192)
    m.set_persistent(0, False)
    m.set_mode(m.MODE_TREECODE)
    m.set_fmm_params()


def _say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _run(fmm, xy, cir, cor, eta, mode=2, budget=8192, blocks=0, graph=False):
    fmm.set_smooth(eta, 16)
    fmm.set_mode(mode)
    fmm.set_fmm_params(16, 64, 0.5, 4.0)
    fmm.set_split(budget)
    fmm.set_persistent(blocks, graph)
    fmm.make_tree(xy, cir, cor)
    uv = fmm.eval_sources()
    return uv, fmm.count_work()


@pytest.mark.parametrize("n,seed", [(400_000, 71), (1_000_000, 72)])
def test_smooth_vs_direct(fmm, n, seed):
    wl = workloads()
    xy, cir, cor = wl.gaussian_mixture(n, seed=seed)
    idx = wl.sample_targets(n, 3000, seed=seed)
    d = fmm.direct(xy, cir, cor, xy[idx], style=1)
    _say(f"[smooth] mixture N={n} seed={seed}")
    u1, (f1, p1, s1) = _run(fmm, xy, cir, cor, ETA)
    u0, (f0, p0, s0) = _run(fmm, xy, cir, cor, 0.0)
    l1 = rel_err(u1[idx], d)[0]
    l0 = rel_err(u0[idx], d)[0]
    _say(f"  smooth on: rel_L2={l1:.2e} far={f1} pairs={p1} smooth={s1}")
    _say(f"  smooth off: rel_L2={l0:.2e} far={f0} pairs={p0} smooth={s0}")
    assert s0 == 0 and s1 > 0
    assert p1 < p0
    assert l1 <= 5e-7 and l0 <= 5e-7
    # the two differ by the expansion error and FP32 rounding only
    assert rel_err(u1, u0)[0] <= 3e-7


def test_smooth_inside_dense_cluster(fmm):
    """One tight cluster: every target lies inside or next to smooth cells
    (the series branch of phi near the cell centre)."""
    wl = workloads()
    n = 200_000
    xy, cir, cor = wl.gaussian_mixture(n, seed=73, k=1, std_lo=1e-3, std_hi=1e-3, core=5e-4)
    idx = wl.sample_targets(n, 2000, seed=73)
    d = fmm.direct(xy, cir, cor, xy[idx], style=1)
    u, (far, pairs, sm) = _run(fmm, xy, cir, cor, ETA)
    l2, mx = rel_err(u[idx], d)
    _say(f"[smooth] one cluster N={n}: rel_L2={l2:.2e} rel_max={mx:.2e} pairs={pairs} smooth={sm}")
    assert sm > 20 * pairs / 1000
    assert l2 <= 5e-7 and mx <= 2e-6


def test_smooth_needs_one_core_radius(fmm):
    """Different core radii: no smooth cells, every near pair summed."""
    wl = workloads()
    n = 100_000
    xy, cir, cor = wl.gaussian_mixture(n, seed=74)
    cor = cor.copy()
    cor[0] *= 1.0 + 1e-9
    _, (_, _, sm) = _run(fmm, xy, cir, cor, ETA)
    assert sm == 0
    _, (_, _, sm) = _run(fmm, xy, cir, cor, ETA, mode=1)
    assert sm == 0


def test_smooth_split_deterministic(fmm):
    """Smooth cells in split leaves: budgets and scheduling paths change how
    cells are grouped into batches, never the value a cell adds (each cell is
    evaluated alone in FP32), so only FP64 addition order differs, and results
    with one budget do not depend on the path."""
    wl = workloads()
    xy, cir, cor = wl.gaussian_mixture(60000, seed=75, k=2, std_lo=2e-3, std_hi=8e-3, core=1e-3)
    u0, (_, _, sm) = _run(fmm, xy, cir, cor, ETA, budget=0)
    assert sm > 0
    scale = np.abs(u0).max()
    cus = fmm.device_cus()
    ref = None
    for name, blocks, graph in (("grid", 0, False), ("persistent", 2 * cus, False),
                                ("graph", 2 * cus, True), ("graph replay", 2 * cus, True)):
        u, _ = _run(fmm, xy, cir, cor, ETA, budget=3, blocks=blocks, graph=graph)
        dd = np.abs(u - u0).max() / scale
        _say(f"[smooth] budget 3 {name}: vs unsplit {dd:.2e}")
        assert dd <= 1e-12
        if ref is None:
            ref = u
        else:
            assert np.array_equal(u, ref), name
