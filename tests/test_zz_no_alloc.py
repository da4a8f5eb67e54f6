"""frh_prefilter_noalloc (no device allocation; scratch in a static device
array) against frh_prefilter_ex with caller scratch, through the C ABI:
two calls in a row, the second with a smaller input, bit for bit; inputs
beyond the static array's size are rejected before any device work."""

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _weights(sigma, truncate=4.0):
    from pyflowreg_amd.preprocess import gaussian_weights
    return gaussian_weights(sigma, truncate)


def _both(L, raw, f64_in, out64, semantics, sy, sx, lo, den):
    import torch
    from pyflowreg_amd._lib import check, ptr, scratch, stream_handle
    T, H, W, C = raw.shape
    ry, wy = _weights(sy)
    rx, wx = _weights(sx)
    x = torch.from_numpy(raw).cuda()
    odt = torch.float64 if out64 else torch.float32
    a = torch.full((T, C, H, W), float("nan"), dtype=odt, device="cuda")
    b = torch.full((T, C, H, W), float("nan"), dtype=odt, device="cuda")
    nb = L.frh_prefilter_scratch_bytes(T, H, W, C, 0)
    work = scratch(nb)
    check(L.frh_prefilter_ex(ptr(x), int(f64_in), T, H, W, C, ptr(lo), ptr(den), ptr(wy), ry, ptr(wx), rx,
                             ptr(a), int(out64), semantics, ptr(work), nb, stream_handle()), "frh_prefilter_ex")
    check(L.frh_prefilter_noalloc(ptr(x), int(f64_in), T, H, W, C, ptr(lo), ptr(den), ptr(wy), ry, ptr(wx), rx,
                                  ptr(b), int(out64), semantics, stream_handle()), "frh_prefilter_noalloc")
    torch.cuda.synchronize()
    return a.cpu().numpy(), b.cpu().numpy()


def test_noalloc_prefilter_matches_scratch_prefilter_twice(gpu):
    from pyflowreg_amd._lib import lib
    L = lib()
    assert L.frh_prefilter_noalloc_max_bytes() == 32 << 20
    rng = np.random.default_rng(71)
    # first call: 3 frames of 256 x 320 x 2, float32 in, float32 out
    raw = (rng.random((3, 256, 320, 2)) * 900.0).astype(np.float32)
    lo = np.array([1.5, -2.0])
    den = np.array([880.0, 910.0])
    a, b = _both(L, raw, False, False, 0, 1.0, 1.0, lo, den)
    assert np.isfinite(a).all()
    np.testing.assert_array_equal(b, a)
    # second call, smaller: 1 frame of 100 x 120 x 1, float64 in, float64
    # out, other sigmas (the first call's scratch contents are stale)
    raw2 = rng.random((1, 100, 120, 1)) * 3.0
    a2, b2 = _both(L, raw2, True, True, 0, 1.7, 0.8, np.array([0.25]), np.array([2.5]))
    np.testing.assert_array_equal(b2, a2)
    # float32 semantics flags (FlowRegLive's float32 reference)
    raw3 = (rng.random((1, 64, 48, 2)) * 7.0).astype(np.float32)
    a3, b3 = _both(L, raw3, False, True, 3, 1.0, 1.0, np.array([0.5, 0.5]), np.array([6.0, 6.0]))
    np.testing.assert_array_equal(b3, a3)


def test_noalloc_prefilter_rejects_oversize_input(gpu):
    import torch
    from pyflowreg_amd._lib import lib, ptr, stream_handle
    L = lib()
    T, H, W, C = 3, 1024, 1024, 2        # 48 MiB of float64 scratch needed
    assert L.frh_prefilter_scratch_bytes(T, H, W, C, 0) > L.frh_prefilter_noalloc_max_bytes()
    x = torch.zeros((1,), device="cuda")
    w, lo, den = np.ones(1), np.zeros(2), np.ones(2)
    rc = L.frh_prefilter_noalloc(ptr(x), 0, T, H, W, C, ptr(lo), ptr(den), ptr(w), 0, ptr(w), 0, ptr(x), 0, 0,
                                 stream_handle())
    assert rc == -1
    assert b"static scratch holds 33554432" in L.frh_last_error()
