"""GPU checks of the linear registration warp and the cross-correlation
prealignment (cc_initialization) against the oracle, and the reference's
own prealignment tests (tests/util/test_xcorr_prealignment.py) on the GPU
path."""

import numpy as np
import pytest
from scipy.ndimage import shift as ndi_shift

from oracle import xcorr_oracle as XO

pytestmark = pytest.mark.gpu


# ------------------------------------------------------------ linear warp
@pytest.mark.parametrize("dt", [np.float32, np.float64])
def test_linear_warp_bit_exact(gpu, oracle, dt):
    from pyflowreg_amd.optical_flow import imregister_wrapper
    rng = np.random.default_rng(40)
    H, W, C = 37, 45, 2
    f2 = (rng.random((H, W, C)) * 10).astype(np.float32).astype(dt)
    f1 = (rng.random((H, W, C)) * 10).astype(np.float32).astype(dt)
    u = rng.uniform(-4, 4, (H, W))
    v = rng.uniform(-4, 4, (H, W))
    out = imregister_wrapper(f2, u, v, f1, interpolation_method="linear")
    ref = oracle.imregister_wrapper(f2, u, v, f1, "linear")
    np.testing.assert_array_equal(out, ref)
    out1 = imregister_wrapper(f2[..., 0], u, v, f1[..., 0], interpolation_method="linear")
    np.testing.assert_array_equal(out1, ref[..., 0])


def test_linear_warp_batch_bit_exact(gpu, oracle):
    import torch
    from pyflowreg_amd.executor import warp_frames
    rng = np.random.default_rng(41)
    T, H, W, C = 3, 29, 33, 2
    frames = rng.random((T, H, W, C)).astype(np.float32)
    fill = rng.random((H, W, C)).astype(np.float32)
    flow = rng.uniform(-3, 3, (T, H, W, 2)).astype(np.float32)
    dev = torch.device("cuda")
    out = warp_frames(torch.from_numpy(frames).to(dev), torch.from_numpy(flow).to(dev),
                      torch.from_numpy(fill).to(dev), "linear").cpu().numpy()
    for t in range(T):
        ref = oracle.imregister_wrapper(frames[t], flow[t, ..., 0].astype(np.float64),
                                        flow[t, ..., 1].astype(np.float64), fill, "linear")
        np.testing.assert_array_equal(out[t], ref)


# ---------------------------------------------------------------- resize
@pytest.mark.parametrize("shape", [(128, 128, 64, 64), (512, 512, 256, 256), (100, 70, 37, 64)])
def test_lanczos_resize_bit_exact(gpu, shape):
    import torch
    from pyflowreg_amd.xcorr import lanczos4_tables, resize_lanczos4
    h, w, oh, ow = shape
    rng = np.random.default_rng(42)
    img = rng.random((2, h, w)).astype(np.float32)
    out = resize_lanczos4(torch.from_numpy(img).cuda(), oh, ow).cpu().numpy()
    ix, wx = lanczos4_tables(w, ow)
    iy, wy = lanczos4_tables(h, oh)
    for n in range(2):
        np.testing.assert_array_equal(out[n], XO.resize_with_tables(img[n], oh, ow, (ix, wx, iy, wy)))


# -------------------------------------------------------------- estimates
def _tables_for(H, W, target):
    from pyflowreg_amd.xcorr import lanczos4_tables
    Th, Tw = min(H, target[0]), min(W, target[1])
    if (Th, Tw) == (H, W):
        return None
    ix, wx = lanczos4_tables(W, Th)
    iy, wy = lanczos4_tables(H, Tw)
    return ix, wx, iy, wy


def _both(ref, mov, **kw):
    from pyflowreg_amd.xcorr import estimate_rigid_xcorr_2d
    g = estimate_rigid_xcorr_2d(ref, mov, **kw)
    H, W = ref.shape[:2]
    t = kw.get("target_hw", (256, 256))
    t = (t, t) if isinstance(t, int) else t
    o = XO.estimate_rigid_xcorr_2d(ref, mov, tables=_tables_for(H, W, t), **kw)
    return g, o


def test_reference_basic_rigid_shift(gpu):
    """tests/util/test_xcorr_prealignment.py:20-43."""
    H, W = 128, 128
    ref = np.zeros((H, W), dtype=np.float32)
    ref[40:60, 40:60] = 1.0
    ref[70:80, 70:80] = 0.8
    true = np.array([5.0, -3.0], np.float32)
    mov = ndi_shift(ref, shift=(true[1], true[0]), order=1, mode="constant")
    g, o = _both(ref, mov, target_hw=(64, 64), up=1)
    assert np.all(np.abs(g - true) <= 1.0)
    np.testing.assert_array_equal(g, o)
    assert g.dtype == np.float32 and g.shape == (2,)


def test_reference_subpixel_accuracy(gpu):
    """tests/util/test_xcorr_prealignment.py:45-75."""
    H, W = 128, 128
    ref = np.random.default_rng(43).random((H, W)).astype(np.float32) * 0.1
    y, x = np.ogrid[:H, :W]
    ref += np.exp(-((y - 30) ** 2 + (x - 40) ** 2) / 50) * 2.0
    ref += np.exp(-((y - 80) ** 2 + (x - 90) ** 2) / 40) * 1.5
    ref += np.exp(-((y - 60) ** 2 + (x - 60) ** 2) / 60) * 1.8
    ref[45:55, 70:85] = 1.2
    ref[90:100, 20:30] = 0.8
    true = np.array([2.3, -1.7], np.float32)
    mov = ndi_shift(ref, shift=(true[1], true[0]), order=3, mode="constant")
    g, o = _both(ref, mov, target_hw=(128, 128), up=10)
    assert np.all(np.abs(g - true) < 0.5)
    np.testing.assert_array_equal(g, o)


def test_reference_multichannel_and_channel_weights(gpu):
    """tests/util/test_xcorr_prealignment.py:77-107, plus 1-D channel
    weights."""
    H, W, C = 64, 64, 2
    ref = np.zeros((H, W, C), dtype=np.float32)
    y, x = np.ogrid[:H, :W]
    ref[..., 0] = np.exp(-((y - 20) ** 2 + (x - 30) ** 2) / 50)
    ref[..., 1] = np.exp(-((y - 40) ** 2 + (x - 35) ** 2) / 40)
    ref[15:25, 40:50, 0] = 1.0
    ref[35:45, 10:20, 1] = 0.8
    ref += (np.random.default_rng(44).standard_normal((H, W, C)) * 0.05).astype(np.float32)
    true = np.array([3.0, -2.0], np.float32)
    mov = np.stack([ndi_shift(ref[..., c], shift=(true[1], true[0]), order=1, mode="constant")
                    for c in range(C)], -1)
    for kw in ({}, {"weight": np.array([0.7, 0.3])}):
        g, o = _both(ref, mov, target_hw=(32, 32), up=1, **kw)
        assert np.all(np.abs(g - true) <= 1.0)
        np.testing.assert_array_equal(g, o)
    from pyflowreg_amd.xcorr import estimate_rigid_xcorr_2d
    with pytest.raises(ValueError):
        estimate_rigid_xcorr_2d(ref, mov, target_hw=(32, 32), weight=np.ones((H, W, C)))


def test_reference_weighted_estimation(gpu):
    """tests/util/test_xcorr_prealignment.py:109-137."""
    H, W = 128, 128
    ref = np.zeros((H, W), dtype=np.float32)
    ref[30:50, 30:50] = 1.0
    ref[80:90, 80:90] = 0.3
    weight = np.zeros((H, W), dtype=np.float32)
    weight[20:60, 20:60] = 1.0
    true = np.array([4.0, -3.0], np.float32)
    mov = ndi_shift(ref, shift=(true[1], true[0]), order=1, mode="constant")
    g, o = _both(ref, mov, weight=weight, target_hw=(64, 64), up=1)
    assert np.all(np.abs(g - true) <= 1.0)
    np.testing.assert_array_equal(g, o)


def test_batch_matches_per_frame_and_oracle(gpu):
    """A batch of 6 translated frames (512x512 downsampled to 256x256,
    subpixel refinement and disambiguation on) equals the per-frame oracle
    estimates."""
    import torch
    from pyflowreg_amd.xcorr import estimate_rigid_xcorr_batch
    rng = np.random.default_rng(45)
    H, W = 512, 512
    from scipy.ndimage import gaussian_filter
    ref = gaussian_filter(rng.random((H, W)), 3.0).astype(np.float32)
    shifts = [(7.5, -3.25), (-20.0, 11.0), (0.0, 0.0), (3.3, 4.4), (-60.0, -45.0), (1.0, -90.0)]
    movs = np.stack([ndi_shift(ref, (dy, dx), order=3, mode="nearest") for dx, dy in shifts]).astype(np.float32)
    g = estimate_rigid_xcorr_batch(ref, movs, target_hw=256, up=4).cpu().numpy()
    tabs = _tables_for(H, W, (256, 256))
    for t, (dx, dy) in enumerate(shifts):
        o = XO.estimate_rigid_xcorr_2d(ref, movs[t], target_hw=256, up=4, tables=tabs)
        np.testing.assert_array_equal(g[t], o)
        assert abs(g[t, 0] - dx) <= 0.75 and abs(g[t, 1] - dy) <= 0.75
    assert isinstance(estimate_rigid_xcorr_batch(ref, movs[:1], target_hw=256), torch.Tensor)


def test_disambiguation_batch_matches_oracle(gpu):
    """Zero-filled translations beyond half the image: the batched GPU
    disambiguation picks the same aliases as the oracle (integer and
    subpixel refinement)."""
    import torch
    from pyflowreg_amd.xcorr import phase_cross_correlation
    from test_xcorr_oracle import _zero_filled_translation
    rng = np.random.default_rng(48)
    ref = rng.random((64, 64)).astype(np.float32)
    movs = np.stack([_zero_filled_translation(ref, dy, dx) for dy, dx in
                     [(40, -45), (-36, 10), (5, 50), (20, -25), (0, 0)]])
    for up in (1, 4):
        g = phase_cross_correlation(torch.from_numpy(ref).cuda(), torch.from_numpy(movs).cuda(),
                                    upsample_factor=up, disambiguate=True)
        for t in range(movs.shape[0]):
            np.testing.assert_array_equal(g[t], XO.phase_cross_correlation(ref, movs[t], upsample_factor=up,
                                                                             disambiguate=True))


# ------------------------------------------------------- cc in the executor
def _rigid_video(T, H, W, seed):
    """tests/util/test_xcorr_prealignment.py:142-180's generator (seeded)."""
    rng = np.random.default_rng(seed)
    reference = rng.random((H, W)).astype(np.float32) * 0.2
    y, x = np.ogrid[:H, :W]
    reference += (10 * np.exp(-((y - H / 2) ** 2 + (x - W / 2) ** 2) / 200)).astype(np.float32)
    reference[30:50, 30:50] = 5.0
    reference[70:90, 70:90] = 4.0
    reference[40:60, 80:100] = 3.0
    shifts = np.array([[10.0, -5.0], [-8.0, 7.0], [6.0, -4.0], [-5.0, -3.0], [4.0, 6.0]], np.float32)[:T]
    frames = np.stack([ndi_shift(reference, shift=(s[1], s[0]), order=1, mode="nearest")
                       + (rng.standard_normal((H, W)) * 0.1).astype(np.float32) for s in shifts])
    return reference.astype(np.float32), frames.astype(np.float32), shifts


def test_executor_cc_matches_sequential_oracle(gpu, oracle):
    """parallelization/sequential.py:97-169 with cc_initialization=True,
    restated with the oracle frame by frame: identical rigid estimates,
    flows and registered frames."""
    from pyflowreg_amd.executor import HipBatchedExecutor
    H = W = 128
    ref, frames, _ = _rigid_video(5, H, W, 46)
    ref_raw = ref[..., None].astype(np.float64)
    ref_proc = oracle.preprocess(ref_raw).astype(np.float32)
    batch = frames[..., None]
    bp = oracle.preprocess(batch, ref_raw).astype(np.float32)
    w_init = np.zeros((H, W, 2), np.float32)
    w_init[..., 0] = 0.5
    weight = np.ones((H, W, 1))
    fp = dict(alpha=(1.0, 1.0), weight=weight, levels=10, min_level=0, eta=0.8, iterations=20,
              update_lag=5, a_data=0.45, a_smooth=1.0, cc_initialization=True, cc_hw=64, cc_up=2)
    ex = HipBatchedExecutor()
    reg, w = ex.process_batch(batch, bp, ref_raw, ref_proc, w_init, flow_params=fp,
                              interpolation_method="linear")
    tabs = _tables_for(H, W, (64, 64))
    kw = {k: v for k, v in fp.items() if k not in ("cc_initialization", "cc_hw", "cc_up", "weight")}
    for t in range(batch.shape[0]):
        partial = oracle.imregister_wrapper(bp[t], w_init[..., 0].astype(np.float64),
                                            w_init[..., 1].astype(np.float64), ref_proc, "linear")
        wc = XO.estimate_rigid_xcorr_2d(ref_proc[..., 0], partial[..., 0], target_hw=(64, 64), up=2,
                                        weight=weight, tables=tabs)
        comb = w_init.copy()
        comb[..., 0] += wc[0]
        comb[..., 1] += wc[1]
        aligned = oracle.imregister_wrapper(bp[t], comb[..., 0].astype(np.float64),
                                            comb[..., 1].astype(np.float64), ref_proc, "linear")
        res = oracle.get_displacement(ref_proc, aligned, uv=np.zeros_like(w_init), weight=weight, **kw)
        flow = (comb + res).astype(np.float32)
        np.testing.assert_array_equal(w[t], flow)
        # base.py:63-73: float32 raw frames and reference, cv2's 32F path
        oreg = oracle.imregister_wrapper(batch[t], flow[..., 0].astype(np.float64),
                                         flow[..., 1].astype(np.float64), ref_raw.astype(np.float32), "linear")
        np.testing.assert_array_equal(reg[t], oreg)


def test_compensate_cc_improves_registration(gpu):
    """tests/util/test_xcorr_prealignment.py:182-221 on the GPU path:
    with few iterations the prealignment does not make the registration
    worse and the error stays small."""
    from pyflowreg_amd.compensate import RegistrationOptions, compensate_arr
    reference, frames, _ = _rigid_video(5, 128, 128, 47)
    base = dict(alpha=(1.0, 1.0), levels=5, eta=0.8, iterations=5, quality_setting="fast")
    reg0, _ = compensate_arr(frames, reference, RegistrationOptions(**base))
    reg1, w1 = compensate_arr(frames, reference, RegistrationOptions(cc_initialization=True, cc_hw=64, cc_up=2,
                                                                     **base))
    e0 = np.mean(np.abs(reg0 - reference[None]))
    e1 = np.mean(np.abs(reg1 - reference[None]))
    assert e1 <= e0 * 1.05, (e0, e1)
    assert e1 < 0.5
    assert w1.shape == (5, 128, 128, 2) and w1.dtype == np.float32

