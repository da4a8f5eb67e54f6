"""CPU pins of the prealignment oracle (oracle/xcorr_oracle.py) and of the
linear warp in the C oracle, plus the host-side Lanczos tables of the
package against the oracle's independent ones."""

import numpy as np
import pytest
import torch
from scipy.ndimage import shift as ndi_shift

from oracle import xcorr_oracle as X


# ---------------------------------------------------------------- Lanczos
def test_lanczos_kernel_closed_forms():
    t = np.array([0.0, 1.0, -2.0, 3.0, 0.5, -1.5, 2.25])
    k = X.lanczos4_kernel(t)
    assert k[0] == 1.0 and k[1] == 0.0 and k[2] == 0.0 and k[3] == 0.0
    for tt, kk in zip(t[4:], k[4:]):
        ref = (np.sin(np.pi * tt) / (np.pi * tt)) * (np.sin(np.pi * tt / 4) / (np.pi * tt / 4))
        assert abs(kk - ref) < 1e-15


@pytest.mark.parametrize("n_in,n_out", [(64, 64), (128, 64), (512, 256), (100, 37), (37, 100)])
def test_lanczos_tables(n_in, n_out):
    idx, wt = X.lanczos4_tables(n_in, n_out)
    assert np.abs(wt.astype(np.float64).sum(1) - 1).max() < 4e-7
    assert idx.min() >= 0 and idx.max() <= n_in - 1
    if n_in == n_out:  # scale 1: every output is its source sample
        np.testing.assert_array_equal(wt, np.tile(np.eye(8, dtype=np.float32)[3], (n_out, 1)))
        np.testing.assert_array_equal(idx[:, 3], np.arange(n_out))
    if n_in == 2 * n_out:  # half-pixel offsets: symmetric taps
        np.testing.assert_array_equal(wt, wt[:, ::-1])


@pytest.mark.parametrize("n_in,n_out", [(128, 64), (512, 256), (100, 37), (256, 64)])
def test_package_lanczos_tables_match_oracle(n_in, n_out):
    from pyflowreg_amd.xcorr import lanczos4_tables
    i1, w1 = lanczos4_tables(n_in, n_out)
    i2, w2 = X.lanczos4_tables(n_in, n_out)
    np.testing.assert_array_equal(i1, i2)
    # the package scales by the float32 reciprocal of the sum, the oracle
    # divides by the sum: a few float32 ulps apart
    ulp = np.spacing(np.abs(w2).astype(np.float32))
    assert np.all(np.abs(w1 - w2) <= 4 * ulp)


def test_lanczos_resize_of_constant_and_identity():
    img = np.full((40, 30), 3.25, np.float32)
    out = X.resize_lanczos4(img, 17, 11)
    assert np.abs(out - 3.25).max() < 2e-6
    rng = np.random.default_rng(0)
    img = rng.random((24, 18)).astype(np.float32)
    np.testing.assert_array_equal(X.resize_lanczos4(img, 24, 18), img)


# ------------------------------------------------------- phase correlation
@pytest.mark.parametrize("dy,dx", [(3, -5), (-7, 11), (0, 0), (16, -23), (-15, 1)])
@pytest.mark.parametrize("norm", ["phase", None])
def test_integer_circular_shift_recovered(dy, dx, norm):
    rng = np.random.default_rng(1)
    ref = rng.random((32, 48)).astype(np.float32)
    mov = np.roll(ref, (dy, dx), axis=(0, 1))
    s = X.phase_cross_correlation(ref, mov, normalization=norm)
    wrap = lambda a, n: ((a + n // 2 - 1) % n) - n // 2 + 1  # noqa: E731  (-n/2, n/2]
    assert tuple(s) == (wrap(-dy, 32), wrap(-dx, 48))
    s10 = X.phase_cross_correlation(ref, mov, upsample_factor=10, normalization=norm)
    np.testing.assert_allclose(s10, s, atol=1e-6)


def test_upsampled_dft_equals_inverse_fft_on_integer_points():
    rng = np.random.default_rng(2)
    R = rng.standard_normal((12, 10)) + 1j * rng.standard_normal((12, 10))
    out = X.upsampled_dft(R, 7, 1.0, (0.0, 0.0))
    ref = np.fft.ifft2(R) * R.size
    np.testing.assert_allclose(out, ref[:7, :7], atol=1e-10)


@pytest.mark.parametrize("sy,sx,n", [(2.3, -1.7, 64), (-0.4, 3.9, 63), (5.0, 0.6, 40)])
def test_upsampled_dft_peak_of_fractional_ramp(sy, sx, n):
    """A pure phase ramp exp(-2 pi i f s / n) is the spectrum of a shift by
    s: its inverse DFT has |.| = n*n exactly at s and less elsewhere."""
    up = 10.0
    f = np.fft.fftfreq(n, 1.0 / n)
    R = np.exp(-2j * np.pi * (f[:, None] * sy + f[None, :] * sx) / n)
    rounded = np.round(np.array([sy, sx]) * up) / up
    region = int(np.ceil(up * 1.5))
    dftshift = np.fix(region / 2.0)
    off = dftshift - rounded * up
    mag = np.abs(X.upsampled_dft(R, region, up, off))
    pk = np.array(np.unravel_index(np.argmax(mag), mag.shape)) - dftshift
    np.testing.assert_allclose(rounded + pk / up, [sy, sx], atol=1e-9)
    assert abs(mag.max() - n * n) < 1e-6 * n * n


def _zero_filled_translation(ref, dy, dx):
    h, w = ref.shape
    mov = np.zeros_like(ref)
    mov[max(0, dy):h + min(0, dy), max(0, dx):w + min(0, dx)] = \
        ref[max(0, -dy):h - max(0, dy), max(0, -dx):w - max(0, dx)]
    return mov


@pytest.mark.parametrize("dy,dx", [(40, -45), (-36, 10), (5, 50), (20, -25)])
def test_disambiguation_picks_the_overlapping_alias(dy, dx):
    """A non-periodic translation (zero fill) of a 64x64 image: the periodic
    estimate wraps into (-32, 32]; the disambiguated shift is the true one,
    also beyond half the image."""
    rng = np.random.default_rng(3)
    ref = rng.random((64, 64)).astype(np.float32)
    mov = _zero_filled_translation(ref, dy, dx)
    s = X.phase_cross_correlation(ref, mov, disambiguate=True)
    assert tuple(s) == (-dy, -dx)
    if abs(dy) > 32 or abs(dx) > 32:
        assert tuple(X.phase_cross_correlation(ref, mov)) != (-dy, -dx)


# --------------------------------- reference tests (test_xcorr_prealignment)
def _basic():
    H, W = 128, 128
    ref = np.zeros((H, W), dtype=np.float32)
    ref[40:60, 40:60] = 1.0
    ref[70:80, 70:80] = 0.8
    return ref


def test_reference_basic_rigid_shift():
    """tests/util/test_xcorr_prealignment.py:20-43."""
    ref = _basic()
    true = np.array([5.0, -3.0], np.float32)
    mov = ndi_shift(ref, shift=(true[1], true[0]), order=1, mode="constant")
    est = X.estimate_rigid_xcorr_2d(ref, mov, target_hw=(64, 64), up=1)
    assert np.all(np.abs(est - true) <= 1.0)


def _subpixel_ref(seed):
    H, W = 128, 128
    ref = np.random.default_rng(seed).random((H, W)).astype(np.float32) * 0.1
    y, x = np.ogrid[:H, :W]
    ref += np.exp(-((y - 30) ** 2 + (x - 40) ** 2) / 50) * 2.0
    ref += np.exp(-((y - 80) ** 2 + (x - 90) ** 2) / 40) * 1.5
    ref += np.exp(-((y - 60) ** 2 + (x - 60) ** 2) / 60) * 1.8
    ref[45:55, 70:85] = 1.2
    ref[90:100, 20:30] = 0.8
    return ref


def test_reference_subpixel_accuracy():
    """tests/util/test_xcorr_prealignment.py:45-75."""
    ref = _subpixel_ref(4)
    true = np.array([2.3, -1.7], np.float32)
    mov = ndi_shift(ref, shift=(true[1], true[0]), order=3, mode="constant")
    est = X.estimate_rigid_xcorr_2d(ref, mov, target_hw=(128, 128), up=10)
    assert np.all(np.abs(est - true) < 0.5)


def _multichannel(seed):
    H, W, C = 64, 64, 2
    ref = np.zeros((H, W, C), dtype=np.float32)
    y, x = np.ogrid[:H, :W]
    ref[..., 0] = np.exp(-((y - 20) ** 2 + (x - 30) ** 2) / 50)
    ref[..., 1] = np.exp(-((y - 40) ** 2 + (x - 35) ** 2) / 40)
    ref[15:25, 40:50, 0] = 1.0
    ref[35:45, 10:20, 1] = 0.8
    ref += (np.random.default_rng(seed).standard_normal((H, W, C)) * 0.05).astype(np.float32)
    return ref


def test_reference_multichannel():
    """tests/util/test_xcorr_prealignment.py:77-107."""
    ref = _multichannel(5)
    true = np.array([3.0, -2.0], np.float32)
    mov = np.stack([ndi_shift(ref[..., c], shift=(true[1], true[0]), order=1, mode="constant")
                    for c in range(2)], -1)
    est = X.estimate_rigid_xcorr_2d(ref, mov, target_hw=(32, 32), up=1)
    assert np.all(np.abs(est - true) <= 1.0)
    est_w = X.estimate_rigid_xcorr_2d(ref, mov, target_hw=(32, 32), up=1, weight=np.array([0.7, 0.3]))
    assert np.all(np.abs(est_w - true) <= 1.0)


def test_reference_weighted_estimation():
    """tests/util/test_xcorr_prealignment.py:109-137 (a 2-D image: the
    weight map is not used)."""
    H, W = 128, 128
    ref = np.zeros((H, W), dtype=np.float32)
    ref[30:50, 30:50] = 1.0
    ref[80:90, 80:90] = 0.3
    weight = np.zeros((H, W), dtype=np.float32)
    weight[20:60, 20:60] = 1.0
    true = np.array([4.0, -3.0], np.float32)
    mov = ndi_shift(ref, shift=(true[1], true[0]), order=1, mode="constant")
    est = X.estimate_rigid_xcorr_2d(ref, mov, weight=weight, target_hw=(64, 64), up=1)
    assert np.all(np.abs(est - true) <= 1.0)


def test_full_weight_map_with_channels_raises():
    """The executors pass the (H, W, C) weight map; with C > 1 numpy's
    tensordot over the channel axis rejects its H*W*C entries."""
    ref = _multichannel(6)
    with pytest.raises(ValueError):
        X.estimate_rigid_xcorr_2d(ref, ref, target_hw=(32, 32), weight=np.ones((64, 64, 2)))


# ------------------------------------------------------------ linear warp
def test_linear_warp_reference_tests(oracle):
    """tests/core/test_warping.py:22-33,250-360 with interpolation_method=
    "linear": zero shift is the identity, integer shifts copy, pixels mapped
    outside take the fixed frame (float32 and float64 images)."""
    rng = np.random.default_rng(7)
    H, W = 30, 34
    for dt in (np.float32, np.float64):
        f2 = rng.random((H, W, 2)).astype(np.float32).astype(dt)
        f1 = rng.random((H, W, 2)).astype(np.float32).astype(dt)
        z = np.zeros((H, W))
        np.testing.assert_array_equal(oracle.imregister_wrapper(f2, z, z, f1, "linear"), f2.astype(np.float32))
        out = oracle.imregister_wrapper(f2, z + 4.0, z - 3.0, f1, "linear")
        np.testing.assert_array_equal(out[3:, :-4], f2[:-3, 4:].astype(np.float32))
        np.testing.assert_array_equal(out[:3], f1[:3].astype(np.float32))
        np.testing.assert_array_equal(out[:, -4:], f1[:, -4:].astype(np.float32))
    with pytest.raises(ValueError):
        oracle.imregister_wrapper(f2, z, z, f1, "nearest")


def test_linear_warp_matches_torch_bilinear_at_quantized_positions(oracle):
    """Bilinear at the 1/32-px rounded map equals torch's bilinear
    grid_sample (align_corners, border padding) there, and a 0.3 px shift
    acts as 0.3125 px."""
    rng = np.random.default_rng(8)
    H, W = 26, 31
    f2 = rng.random((H, W)).astype(np.float32)
    u = rng.uniform(-1.5, 1.5, (H, W))
    v = rng.uniform(-1.5, 1.5, (H, W))
    out = oracle.imregister_wrapper(f2, u, v, f2, "linear")
    yy, xx = np.meshgrid(np.arange(H), np.arange(W), indexing="ij")
    mx = (xx + u).astype(np.float32)
    my = (yy + v).astype(np.float32)
    inside = (mx >= 0) & (mx < W) & (my >= 0) & (my < H)
    qx = np.round(np.clip(mx, 0, W - 1) * 32) / 32
    qy = np.round(np.clip(my, 0, H - 1) * 32) / 32
    grid = torch.from_numpy(np.stack([2 * qx / (W - 1) - 1, 2 * qy / (H - 1) - 1], -1).astype(np.float64))[None]
    ref = torch.nn.functional.grid_sample(torch.from_numpy(f2.astype(np.float64))[None, None], grid,
                                          mode="bilinear", padding_mode="border", align_corners=True)[0, 0].numpy()
    np.testing.assert_allclose(out[inside], ref[inside], atol=2e-6)
    ramp = np.tile(np.arange(W, dtype=np.float32), (H, 1))
    o = oracle.imregister_wrapper(ramp, np.full((H, W), 0.3), np.zeros((H, W)), ramp, "linear")
    np.testing.assert_allclose(o[:, :-2], ramp[:, :-2] + 0.3125, atol=1e-5)


def test_linear_warp64_sums_in_float64(oracle):
    """64F path: float weights, float64 products and sums, one rounding to
    float32; checked with numpy float64 at 1/32-px positions."""
    rng = np.random.default_rng(9)
    H, W = 18, 22
    f2 = (rng.random((H, W)) * 100).astype(np.float32).astype(np.float64)
    ku = rng.integers(-40, 40, (H, W))
    kv = rng.integers(-40, 40, (H, W))
    out = oracle.imregister_wrapper(f2, ku / 32.0, kv / 32.0, f2, "linear")
    checked = 0
    for y in range(H):
        for x in range(W):
            X_, Y_ = x * 32 + ku[y, x], y * 32 + kv[y, x]
            if not (0 <= X_ <= (W - 1) * 32 and 0 <= Y_ <= (H - 1) * 32):
                continue
            sx, sy = X_ >> 5, Y_ >> 5
            fx, fy = np.float32((X_ & 31) / 32.0), np.float32((Y_ & 31) / 32.0)
            w = [np.float32(1 - fy) * np.float32(1 - fx), np.float32(1 - fy) * fx, fy * np.float32(1 - fx), fy * fx]
            x1, y1 = min(sx + 1, W - 1), min(sy + 1, H - 1)
            s = (f2[sy, sx] * np.float64(w[0]) + f2[sy, x1] * np.float64(w[1]) + f2[y1, sx] * np.float64(w[2]) +
                 f2[y1, x1] * np.float64(w[3]))
            assert out[y, x] == np.float32(s)
            checked += 1
    assert checked > H * W // 2



def test_cc_parameter_validation():
    """tests/util/test_xcorr_prealignment.py:368-386."""
    from pyflowreg_amd.compensate import RegistrationOptions
    assert RegistrationOptions(cc_initialization=True, cc_hw=128).cc_hw == 128
    assert RegistrationOptions(cc_initialization=True, cc_hw=(128, 256)).cc_hw == (128, 256)
    with pytest.raises(ValueError):
        RegistrationOptions(cc_initialization=True, cc_up=0)
    o = RegistrationOptions()
    assert o.cc_initialization is False and o.cc_hw == 256 and o.cc_up == 1
