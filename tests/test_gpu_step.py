"""Time stepping around the evaluation and mirror images, on the GPU.

The stepping formulas are the solver's (e_drift.f90:29-46, e_core.f90:15-19,
e_velo_update.f90:17-38, e_velo_mpi.f90:44-47); they are checked against the
same arithmetic in numpy (FP64, no contraction on either side: bit for bit),
against the exactness of third-order Adams-Bashforth on quadratic velocity
histories, and against closed forms.  Mirror images are checked against the
point-vortex image in closed form and against the direct sum over blobs and
images."""
import math

import numpy as np
import pytest
import torch

from conftest import fmm_module, rel_err, workloads

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def fmm():
    m = fmm_module()
    yield m
    m.set_mode(m.MODE_TREECODE)


def _dev(a):
    return torch.from_numpy(np.ascontiguousarray(a)).to("cuda:0")


def test_drift_history_store_core(fmm):
    rng = np.random.default_rng(1)
    n = 10007
    xy = rng.random((n, 2))
    hist = rng.standard_normal((n, 3, 2))
    cor = rng.random(n) * 1e-2 + 1e-3
    uv = rng.standard_normal((n, 2))
    dxy, dh, dc, du = _dev(xy), _dev(hist), _dev(cor), _dev(uv)
    dt, dm = 0.015, 0.4
    fmm.drift_ab3_device(n, dxy.data_ptr(), dh.data_ptr(), dt, True, dm)
    fmm.synchronize()
    vm = 23.0 / 12.0 * hist[:, 0] - 4.0 / 3.0 * hist[:, 1] + 5.0 / 12.0 * hist[:, 2]
    want = xy + vm * dt
    low = want[:, 1] <= dm
    want[low, 1] = 2.0 * dm - want[low, 1]
    assert np.array_equal(dxy.cpu().numpy(), want)
    fmm.core_spread_device(n, dc.data_ptr(), dt, 3000.0)
    fmm.synchronize()
    assert np.array_equal(dc.cpu().numpy(), np.sqrt(cor ** 2 + (2.2418 ** 2) * dt / 3000.0))
    fmm.store_velocity_device(n, dxy.data_ptr(), du.data_ptr(), dh.data_ptr(), 1.0, 0.0, 0.1)
    fmm.synchronize()
    p = dxy.cpu().numpy()
    h = dh.cpu().numpy()
    assert np.array_equal(h[:, 0, 0], (1.0 + (-0.1 * p[:, 1])) + uv[:, 0])
    assert np.array_equal(h[:, 0, 1], (0.0 + (0.1 * p[:, 0])) + uv[:, 1])
    fmm.history_shift_device(n, dh.data_ptr(), False)
    fmm.synchronize()
    h2 = dh.cpu().numpy()
    assert np.array_equal(h2[:, 1], h[:, 0]) and np.array_equal(h2[:, 2], h[:, 1])
    fmm.history_shift_device(n, dh.data_ptr(), True)
    fmm.synchronize()
    h3 = dh.cpu().numpy()
    assert np.array_equal(h3[:, 1], h2[:, 0]) and np.array_equal(h3[:, 2], h2[:, 0])


def test_ab3_exact_for_quadratic_histories(fmm):
    """x(t+dt) = x(t) + integral of v; AB3 is exact when v is quadratic in t."""
    n = 64
    rng = np.random.default_rng(2)
    a, b, c = rng.standard_normal((3, n, 2))
    dt = 0.01
    v = lambda t: a + b * t + c * t * t  # noqa: E731
    hist = np.stack([v(0.0), v(-dt), v(-2 * dt)], axis=1)
    # integral of the quadratic extrapolation from 0 to dt (what AB3 integrates)
    x0 = rng.random((n, 2))
    dxy, dh = _dev(x0), _dev(hist)
    fmm.drift_ab3_device(n, dxy.data_ptr(), dh.data_ptr(), dt)
    fmm.synchronize()
    exact = x0 + a * dt + b * dt ** 2 / 2 + c * dt ** 3 / 3
    assert np.abs(dxy.cpu().numpy() - exact).max() < 1e-14


def test_mirror_images_closed_form(fmm, oracle):
    g, s, dm = 0.37, 0.05, 0.2
    xy = np.array([[0.3, 0.5]])
    fmm.set_mode(fmm.MODE_TREECODE)
    fmm.set_params(8, 4, 0.01)  # tiny opening angle: everything summed directly
    d = [_dev(xy), _dev(np.array([g])), _dev(np.array([s]))]
    fmm.make_tree_mirror_device(1, *[t.data_ptr() for t in d], dm)
    uv = fmm.eval_sources()
    assert uv.shape == (1, 2)
    # at the blob itself only the image acts: a point vortex -g at (0.3, -0.1)
    rx, ry = 0.0, 0.5 - (2 * dm - 0.5)
    f = -g / (2 * math.pi * (rx * rx + ry * ry))
    assert abs(uv[0, 0] - (-f * ry)) <= 1e-12 * abs(f * ry) and abs(uv[0, 1]) <= 1e-15


@pytest.mark.parametrize("mode", [0, 1, 2])
def test_mirror_images_vs_direct(fmm, oracle, mode):
    wl = workloads()
    n = 20000
    xy, cir, cor = wl.uniform_square(n, seed=81)
    xy[:, 1] = 0.3 + 0.7 * xy[:, 1]
    dm = 0.25
    fmm.set_mode(mode)
    fmm.set_params(16, 32, 0.3)
    fmm.set_fmm_params(16, 64, 0.5, 4.0)
    d = [_dev(xy), _dev(cir), _dev(cor)]
    fmm.make_tree_mirror_device(n, *[t.data_ptr() for t in d], dm)
    uv = fmm.eval_sources()
    idx = wl.sample_targets(n, 400, seed=81)
    ixy = np.stack([xy[:, 0], 2 * dm - xy[:, 1]], 1)
    allxy = np.concatenate([xy, ixy])
    allc = np.concatenate([cir, -cir])
    allr = np.concatenate([cor, np.full(n, 1e-300)])
    ref = oracle.direct(allxy, allc, allr, xy[idx], style=1)
    l2, _ = rel_err(uv[idx], ref)
    assert l2 < (1e-6 if mode else 1e-3), l2
