"""Batching form: many independent systems of uneven size in one forest,
evaluated by one persistent launch that can be replayed from a HIP graph.

Pins: every system's velocities equal, bit for bit, those of the
single-system path with the same parameters; the persistent and graph-replayed
evaluations equal the one-wave-per-leaf launch bit for bit."""
import importlib
import sys

import numpy as np
import pytest

from conftest import fmm_module, rel_err, workloads

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def fmm():
    m = fmm_module()
    yield m
    m.set_mode(m.MODE_TREECODE)
    m.set_fmm_params()
    m.set_persistent(0, False)


def _say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _systems(count, seed, scale):
    wl = workloads()
    sy = wl.batch_systems(count, seed=seed, scale=scale)
    # an empty system and a one-blob system ride along
    sy.insert(3, (np.zeros((0, 2)), np.zeros(0), np.zeros(0)))
    sy.insert(5, (np.array([[0.25, 0.75]]), np.array([0.3]), np.array([1e-3])))
    return sy


@pytest.mark.parametrize("mode", [0, 1, 2])
def test_forest_equals_single_systems(fmm, mode):
    systems = _systems(14, 61, 0.05)
    fmm.set_mode(mode)
    fmm.set_params(12, 32, 0.5)
    fmm.set_fmm_params(12, 48, 0.5, 4.0)
    off = fmm.make_forest(systems)
    uv = fmm.eval_sources()
    _say(f"[forest] mode {mode}: {len(systems)} systems, N={off[-1]}")
    for s, (xy, cir, cor) in enumerate(systems):
        if xy.shape[0] == 0:
            continue
        fmm.make_tree(xy, cir, cor)
        one = fmm.eval_sources()
        got = uv[off[s]:off[s + 1]]
        assert np.array_equal(got.view(np.uint64), one.view(np.uint64)), s


@pytest.mark.parametrize("mixed", [False, True])
def test_persistent_and_graph_equal_plain(fmm, mixed):
    wl = workloads()
    fmm.set_mode(fmm.MODE_MIXED if mixed else fmm.MODE_FMM64)
    fmm.set_fmm_params(12, 64, 0.5, 4.0)
    systems = _systems(20, 62, 0.1)
    fmm.set_persistent(0, False)
    fmm.make_forest(systems)
    ref = fmm.eval_sources()
    cus = fmm.device_cus()
    assert cus > 0
    fmm.set_persistent(2 * cus, False)
    fmm.make_forest(systems)
    assert np.array_equal(fmm.eval_sources().view(np.uint64), ref.view(np.uint64))
    fmm.set_persistent(2 * cus, True)
    # graph replays across rebuilt trees with circulations scaled by 2^k (exact
    # in FP32 and FP64, so the velocities scale exactly), then the original
    for k in range(3):
        scaled = [(xy, cir * 2.0 ** k, cor) for xy, cir, cor in systems]
        fmm.make_forest(scaled)
        got = fmm.eval_sources()
        assert np.array_equal(got.view(np.uint64), (ref * 2.0 ** k).view(np.uint64)), k
    fmm.make_forest(systems)
    assert np.array_equal(fmm.eval_sources().view(np.uint64), ref.view(np.uint64))
    fmm.set_persistent(0, False)


def test_forest_accuracy(fmm, oracle):
    """Each system of the forest against its own FP64 direct sum."""
    systems = _systems(10, 63, 0.1)
    fmm.set_mode(fmm.MODE_MIXED)
    fmm.set_fmm_params(12, 64, 0.5, 4.0)
    off = fmm.make_forest(systems)
    uv = fmm.eval_sources()
    worst = 0.0
    for s, (xy, cir, cor) in enumerate(systems):
        if xy.shape[0] < 2:
            continue
        t = xy[:: max(1, xy.shape[0] // 200)]
        d = oracle.direct(xy, cir, cor, t, style=1)
        got = uv[off[s]:off[s + 1]][:: max(1, xy.shape[0] // 200)]
        worst = max(worst, rel_err(got, d)[0])
    _say(f"  worst per-system rel_L2 (mixed, p=12): {worst:.2e}")
    # p = 12 at theta = 0.5: truncation (theta / (2 - theta))^12 = 1.9e-6 per far cell
    assert worst <= 5e-6


def test_autotuner_knobs(fmm):
    """The autotuner's knobs (leaf, grid, graph, chunk size, far-field burst)
    each keep the accuracy; it returns the fastest configuration it timed and
    leaves it applied."""
    tune = importlib.import_module("2dfmm_amd.autotune")
    import torch

    systems = _systems(12, 63, 0.05)
    off, xy, cir, cor = fmm.concat_systems(systems)
    n = int(off[-1])
    dev = torch.device("cuda:0")
    d = [torch.from_numpy(a).to(dev) for a in (xy, cir, cor)]
    d_uv = torch.empty((n, 2), dtype=torch.float64, device=dev)
    torch.cuda.synchronize()
    fmm.set_mode(fmm.MODE_MIXED)
    best, table = tune.tune_forest(fmm, off, d[0].data_ptr(), d[1].data_ptr(), d[2].data_ptr(),
                                   d_uv.data_ptr(), p=12, leaves=(32, 64), blocks_per_cu=(0, 2),
                                   splits=(512, 8192), bursts=(64, 128), reps=1,
                                   sync=torch.cuda.synchronize)
    _say(f"[autotune] best {best} of {len(table)}")
    assert best == min(table, key=lambda t: t[1])[0]
    assert {c["burst"] for c, _ in table} == {64, 128} and {c["split"] for c, _ in table} == {512, 8192}
    s = 1 if off[2] > off[1] else 2
    a, b = int(off[s]), int(off[s + 1])
    ref = fmm.direct(xy[a:b], cir[a:b], cor[a:b], xy[a:b], style=1)
    for cfg, _ in table:
        tune.apply(fmm, cfg)
        fmm.make_forest(systems)
        got = fmm.eval_sources()[a:b]
        assert rel_err(got, ref)[0] <= 5e-6, cfg
    tune.reset(fmm)
