"""The numpy MLP used to check the GPU MLP kernels, pinned by central finite
differences of its own FP64 loss."""
import numpy as np

import mlp_ref


def test_reference_gradient_matches_finite_differences():
    rng = np.random.default_rng(0)
    w = mlp_ref.init(11).astype(np.float64)
    w[-66:] += 0.05
    x = rng.standard_normal((64, 8))
    y = rng.standard_normal((64, 2))
    g = mlp_ref.grad(w, x, y)
    for k in rng.choice(mlp_ref.NW, 80, replace=False):
        wp, wm = w.copy(), w.copy()
        wp[k] += 1e-6
        wm[k] -= 1e-6
        fd = (mlp_ref.loss(wp, x, y) - mlp_ref.loss(wm, x, y)) / 2e-6
        assert abs(fd - g[k]) <= 1e-6 * (1 + abs(fd)), k
    assert mlp_ref.NW == 4866
