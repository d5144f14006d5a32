"""The allocation-free direct sum (fmm_direct_device_noalloc) against
fmm_direct_device, through the C ABI: bit-equal results on two calls in a row
(the second smaller, so stale partial sums in the shared scratch array would
show), on the null stream and on a caller's stream, and a clear error past the
scratch array's limit (raised on the host, nothing launched)."""
import numpy as np
import pytest

from conftest import fmm_module, workloads


def _inputs(torch, n, nt, seed):
    wl = workloads()
    xy, cir, cor = wl.gaussian_mixture(n, seed=seed)
    rng = np.random.default_rng(seed)
    txy = np.ascontiguousarray(xy[rng.choice(n, nt, replace=False)] + 1e-4 * rng.standard_normal((nt, 2)))
    dev = torch.device("cuda:0")
    return [torch.from_numpy(np.ascontiguousarray(a)).to(dev) for a in (xy, cir, cor, txy)]


@pytest.mark.gpu
@pytest.mark.parametrize("style", [0, 1])
def test_noalloc_direct_matches_direct_device(style):
    import torch

    fmm = fmm_module()
    for case, (n, nt, seed) in enumerate([(60_000, 20_000, 1), (9_000, 700, 2)]):
        print(f"noalloc direct: style={style} call={case} ns={n} nt={nt}", flush=True)
        d_xy, d_cir, d_cor, d_t = _inputs(torch, n, nt, seed)
        ref = torch.empty((nt, 2), dtype=torch.float64, device="cuda:0")
        got = torch.full((nt, 2), float("nan"), dtype=torch.float64, device="cuda:0")
        torch.cuda.synchronize()
        fmm.direct_device(n, d_xy.data_ptr(), d_cir.data_ptr(), d_cor.data_ptr(), nt, d_t.data_ptr(),
                          ref.data_ptr(), style)
        fmm.synchronize()
        fmm.direct_device_noalloc(n, d_xy.data_ptr(), d_cir.data_ptr(), d_cor.data_ptr(), nt,
                                  d_t.data_ptr(), got.data_ptr(), style, 0)
        torch.cuda.synchronize()
        a, b = got.cpu().numpy(), ref.cpu().numpy()
        assert np.isfinite(a).all()
        assert np.array_equal(a, b)


@pytest.mark.gpu
def test_noalloc_direct_on_caller_stream():
    import torch

    fmm = fmm_module()
    n, nt = 30_000, 5_000
    print(f"noalloc direct on a caller stream: ns={n} nt={nt}", flush=True)
    d_xy, d_cir, d_cor, d_t = _inputs(torch, n, nt, 3)
    ref = torch.empty((nt, 2), dtype=torch.float64, device="cuda:0")
    got = torch.empty((nt, 2), dtype=torch.float64, device="cuda:0")
    torch.cuda.synchronize()
    fmm.direct_device(n, d_xy.data_ptr(), d_cir.data_ptr(), d_cor.data_ptr(), nt, d_t.data_ptr(),
                      ref.data_ptr(), 1)
    fmm.synchronize()
    st = torch.cuda.Stream()
    with torch.cuda.stream(st):
        fmm.direct_device_noalloc(n, d_xy.data_ptr(), d_cir.data_ptr(), d_cor.data_ptr(), nt,
                                  d_t.data_ptr(), got.data_ptr(), 1, st.cuda_stream)
    st.synchronize()
    assert np.array_equal(got.cpu().numpy(), ref.cpu().numpy())


@pytest.mark.gpu
def test_noalloc_direct_rejects_too_many_targets():
    import torch

    fmm = fmm_module()
    nt = fmm.NOALLOC_MAX_TARGETS + 1
    print(f"noalloc direct: nt={nt} beyond the limit", flush=True)
    buf = torch.zeros(2 * nt, dtype=torch.float64, device="cuda:0")
    out = torch.zeros(2 * nt, dtype=torch.float64, device="cuda:0")
    torch.cuda.synchronize()
    with pytest.raises(fmm.FmmError, match="scratch"):
        fmm.direct_device_noalloc(1, buf.data_ptr(), buf.data_ptr(), buf.data_ptr(), nt,
                                  buf.data_ptr(), out.data_ptr(), 1, 0)
    with pytest.raises(fmm.FmmError, match="style"):
        fmm.direct_device_noalloc(1, buf.data_ptr(), buf.data_ptr(), buf.data_ptr(), 4,
                                  buf.data_ptr(), out.data_ptr(), 2, 0)
