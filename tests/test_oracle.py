"""Pins of the CPU oracle (oracle/) against the reference's own known-answer
tests, closed forms and independent library routines.  CPU only.

Each test names the oracle function it pins.  A check that merely re-typed
the oracle's formula would pin nothing, so the expected values here come from
a different route: torch's bicubic resampler (Keys, A = -0.75), scipy's
Gaussian filter, analytic derivatives of polynomials, dense linear algebra,
and the numbers the reference's tests hold.
"""

import math

import numpy as np
import pytest
import torch
from scipy.ndimage import gaussian_filter1d


# ---------------------------------------------------------------- geometry
@pytest.mark.parametrize("size,n_levels", [((512, 512), 19), ((256, 256), 16),
                                           ((796, 512), 21), ((1024, 1024), 22)])
def test_level_counts(oracle, size, n_levels):
    plan, ml = oracle.level_plan(*size, 0.8, 50, 0)
    assert len(plan) == n_levels and ml == 0
    assert plan[-1][1] == size


def test_level_sizes_512(oracle):
    plan, _ = oracle.level_plan(512, 512, 0.8, 50, 0)
    sides = [s[0] for _, s in plan]
    assert sides == [9, 12, 14, 18, 23, 28, 35, 44, 55, 69, 86, 107, 134, 168, 210, 262, 328,
                     410, 512]
    # closed form: round(512 * 0.8^k) >= 10 for k < depth, < 10 at depth
    depth = oracle.warping_depth(0.8, 50, 512, 512)
    assert round(512 * 0.8 ** (depth - 1)) >= 10 > round(512 * 0.8 ** depth)


# ------------------------------------------------------------------ resize
def _torch_bicubic(img, oh, ow):
    t = torch.from_numpy(img.astype(np.float64))[None, None]
    return torch.nn.functional.interpolate(t, size=(oh, ow), mode="bicubic",
                                           align_corners=False)[0, 0].numpy()


def test_resize_identity_at_scale_one(oracle):
    img = np.random.default_rng(0).random((23, 31)).astype(np.float32)
    np.testing.assert_array_equal(oracle.resize(img, (23, 31)), img)


def test_resize_tables_partition_of_unity(oracle):
    for (i, o, s) in [(512, 9, 0.6 * 512 / 9), (100, 37, 0.6 * 100 / 37), (13, 29, 0.0)]:
        idx, wt = oracle.resize_tables(i, o, s)
        np.testing.assert_allclose(wt.astype(np.float64).sum(axis=1), 1.0, atol=2e-6)
        assert idx.min() >= 0 and idx.max() < i


def _torch_keys_matrix(n, o):
    """T[i, j] = torch's bicubic (Keys, A = -0.75) weight of input j for
    output i, from resizing unit impulses."""
    eye = torch.eye(n, dtype=torch.float64)[:, None, None, :]
    out = torch.nn.functional.interpolate(eye, size=(1, o), mode="bicubic", align_corners=False)
    return out[:, 0, 0, :].numpy().T


def test_resize_upsample_uses_reference_tap_window(oracle):
    """sigma = 0 (flow upsampling): util/resize_util.py:199-213 takes taps
    j = floor(x-2) ... floor(x)+1 and renormalises, i.e. the Keys tap at
    floor(x)+2 is dropped.  Expected values use torch's Keys weights on the
    three remaining non-zero taps."""
    n, o = 20, 37
    f = np.random.default_rng(1).random((3, n)).astype(np.float32)
    out = oracle.resize(f, (3, o))
    T = _torch_keys_matrix(n, o)
    checked = 0
    for i in range(o):
        x = (i + 0.5) / (o / n) - 0.5
        fl = int(np.floor(x))
        if fl - 2 < 0 or fl + 2 >= n:
            continue
        keep = [fl - 1, fl, fl + 1]
        w = T[i, keep]
        np.testing.assert_allclose(out[:, i], f[:, keep].astype(np.float64) @ w / w.sum(),
                                   atol=1e-6)
        checked += 1
    assert checked > 25
    # ... which differs from plain bicubic by far more than rounding
    full = f.astype(np.float64) @ T.T
    assert np.abs(out[:, 5:-5] - full[:, 5:-5]).max() > 1e-3


def test_resize_downsample_is_gaussian_then_cubic(oracle):
    """1-D downsample: the fused kernel is the float32 Gaussian g followed by
    Keys-cubic resampling (torch), minus the one dropped tap
    j* = floor(x) + 2 + R whose weight is g[0] k(frac - 2), renormalised."""
    n, o = 200, 61
    f = np.random.default_rng(2).random((1, n)).astype(np.float32)
    out = oracle.resize(f, (1, o))[0].astype(np.float64)
    sig = 0.6 / (o / n)
    R, g = oracle.gauss_g(sig)
    g = g.astype(np.float64)
    gf = np.convolve(np.pad(f[0].astype(np.float64), R, mode="symmetric"), g, mode="valid")
    T = _torch_keys_matrix(n, o)
    checked = 0
    for i in range(o):
        x = (i + 0.5) / (o / n) - 0.5
        fl = int(np.floor(x))
        if fl - 3 - 2 * R < 0 or fl + 3 + 2 * R >= n:
            continue
        wmiss = g[0] * T[i, fl + 2]
        want = (T[i] @ gf - wmiss * f[0, fl + 2 + R]) / (1 - wmiss)
        assert abs(out[i] - want) < 2e-6
        checked += 1
    assert checked > 20


# ------------------------------------------- cv2.resize INTER_CUBIC (min_level > 0)
def _torch_keys_at(n_in, coords):
    """K[i, j] = torch's bicubic (Keys, A = -0.75, taps clamped to the edge)
    weight of input j at source coordinate coords[i], from grid_sample of
    unit impulses (align_corners=False: coordinate c <-> grid (2c + 1)/n - 1)."""
    eye = torch.eye(n_in, dtype=torch.float64)[:, None, None, :]
    g = (2.0 * np.asarray(coords, np.float64) + 1.0) / n_in - 1.0
    grid = torch.zeros((n_in, 1, len(g), 2), dtype=torch.float64)
    grid[..., 0] = torch.from_numpy(g)
    out = torch.nn.functional.grid_sample(eye, grid, mode="bicubic", padding_mode="border",
                                          align_corners=False)
    return out[:, 0, 0, :].numpy().T


def _cv_source_coords(n_in, n_out):
    """OpenCV's resize source coordinate, formed in double and rounded to
    float32: float((i + 0.5) * (1 / (n_out / n_in)) - 0.5)."""
    return ((np.arange(n_out) + 0.5) * (1.0 / (n_out / n_in)) - 0.5).astype(np.float32)


def _keys_resize(w, m, n, coords_y, coords_x):
    Ky = _torch_keys_at(w.shape[0], coords_y)
    Kx = _torch_keys_at(w.shape[1], coords_x)
    return np.stack([Ky @ w[..., c] @ Kx.T for c in range(w.shape[2])], axis=-1)


@pytest.mark.parametrize("size,min_level", [((512, 512), 4), ((512, 512), 6),
                                            ((256, 256), 4), ((256, 256), 6),
                                            ((796, 512), 4)])
def test_cubic_resize_cv_final_upsample_matches_torch_bicubic(oracle, size, min_level):
    """or_cubic_resize_cv: the final upsample of core/optical_flow.py:612-617
    from the level that min_level stops at, against torch's Keys weights at
    OpenCV's float32 source coordinates.  The remaining difference is the
    float32 rounding of OpenCV's coefficients: at most 5.6e-6 px for flows of
    a few pixels.  At 796 rows the scale 326/796 is not dyadic, so there a
    coordinate kept in float64 misses by 1.7e-4; align_corners or a transposed
    axis miss by more than 1e-3."""
    m, n = size
    plan, ml = oracle.level_plan(m, n, 0.8, 50, min_level)
    assert ml == min_level
    lh, lw = plan[-1][1]
    assert (lh, lw) != (m, n)
    rng = np.random.default_rng(min_level + m)
    w = rng.uniform(-4.0, 4.0, (lh, lw, 2))
    got = oracle.cubic_resize_cv(w, m, n)
    assert got.shape == (m, n, 2)
    want = _keys_resize(w, m, n, _cv_source_coords(lh, m), _cv_source_coords(lw, n))
    d = np.abs(got - want).max()
    assert d < 1e-5, d
    # torch's own resize forms the coordinate in float64: equal where the scale
    # is dyadic (512 -> 210, 256 -> 105 ...), about 1e-4 apart at 326/796
    t = torch.from_numpy(np.ascontiguousarray(w.transpose(2, 0, 1)))[None]
    plain = torch.nn.functional.interpolate(t, size=(m, n), mode="bicubic",
                                            align_corners=False)[0].numpy().transpose(1, 2, 0)
    if m == 796:
        assert np.abs(got - plain).max() > 1e-4
    else:
        assert np.abs(got - plain).max() < 1e-5
    aligned = torch.nn.functional.interpolate(t, size=(m, n), mode="bicubic",
                                              align_corners=True)[0].numpy().transpose(1, 2, 0)
    assert np.abs(got - aligned).max() > 1e-3
    if m != n:
        swapped = _keys_resize(w, m, n, _cv_source_coords(lh, m)[::-1].copy(),
                               _cv_source_coords(lw, n))
        assert np.abs(got - swapped).max() > 1e-3


def test_cubic_resize_cv_constants_and_a(oracle):
    """Constants pass through (OpenCV's last coefficient is 1 - the other
    three, so the weights sum to one up to float32 rounding), and the kernel
    is Keys' A = -0.75, not Catmull-Rom's A = -0.5: A = -0.5 reproduces a
    linear ramp exactly, A = -0.75 does not, and the restatement follows the
    latter by the amount torch's A = -0.75 weights predict."""
    const = np.full((23, 31, 2), 2.5)
    np.testing.assert_allclose(oracle.cubic_resize_cv(const, 57, 64), 2.5, rtol=0, atol=1e-6)
    h, w, m, n = 41, 67, 96, 160
    x = np.arange(w, dtype=np.float64)
    field = np.ascontiguousarray(np.broadcast_to((0.25 * x + 3.0)[None, :, None], (h, w, 1)))
    got = oracle.cubic_resize_cv(field, m, n)[:, :, 0]
    fx = _cv_source_coords(w, n).astype(np.float64)
    ci = (fx >= 2) & (fx < w - 3)
    ramp = 0.25 * fx[ci] + 3.0
    keys = (_torch_keys_at(w, fx) @ (0.25 * x + 3.0))[ci]
    assert np.abs(got[0, ci] - ramp).max() > 1e-3
    np.testing.assert_allclose(got[:, ci], np.broadcast_to(keys, (m, ci.sum())), rtol=0, atol=2e-6)


def test_cubic_resize_cv_identity_at_scale_one(oracle):
    """At scale 1 every source coordinate is an integer and the Keys
    coefficients are (0, 1, 0, 0): the resize returns its input bit for bit."""
    w = np.random.default_rng(7).uniform(-3, 3, (37, 53, 2))
    np.testing.assert_array_equal(oracle.cubic_resize_cv(w, 37, 53), w)


# -------------------------------------------------------------------- warp
def test_warp_zero_displacement_is_identity(oracle):
    """tests/core/test_warping.py:22-33 (zero shift preserves frames)."""
    rng = np.random.default_rng(3)
    f2 = rng.random((32, 40, 2)).astype(np.float32)
    f1 = rng.random((32, 40, 2)).astype(np.float32)
    z = np.zeros((32, 40))
    np.testing.assert_array_equal(oracle.imregister_wrapper(f2, z, z, f1), f2)


@pytest.mark.parametrize("sx,sy", [(10, 5), (-8, -6), (3, 0)])
def test_warp_integer_shift_content_and_oob(oracle, sx, sy):
    """tests/core/test_warping.py:250-360: warped[y, x] = f2[y+sy, x+sx];
    pixels mapped outside take the fixed frame's value."""
    H, W = 50, 64
    yy, xx = np.meshgrid(np.arange(H), np.arange(W), indexing="ij")
    f2 = (yy * 1000 + xx).astype(np.float32)
    f1 = np.full((H, W), -7.0, np.float32)
    u = np.full((H, W), float(sx))
    v = np.full((H, W), float(sy))
    out = oracle.imregister_wrapper(f2, u, v, f1)
    ys = slice(max(0, -sy), H - max(0, sy))
    xs = slice(max(0, -sx), W - max(0, sx))
    np.testing.assert_array_equal(out[ys, xs],
                                  f2[ys.start + sy:ys.stop + sy, xs.start + sx:xs.stop + sx])
    oob = (xx + sx < 0) | (xx + sx >= W) | (yy + sy < 0) | (yy + sy >= H)
    assert oob.any()
    np.testing.assert_array_equal(out[oob], -7.0)


def _torch_warp(f2, mx, my):
    H, W = f2.shape
    grid = torch.from_numpy(np.stack([2 * mx / (W - 1) - 1, 2 * my / (H - 1) - 1], -1).astype(np.float64))[None]
    return torch.nn.functional.grid_sample(torch.from_numpy(f2.astype(np.float64))[None, None],
                                           grid, mode="bicubic", padding_mode="border",
                                           align_corners=True)[0, 0].numpy()


def test_warp_quantizes_positions_to_1_32_px(oracle):
    """A 0.3 px shift is applied as round(0.3 * 32) / 32 = 0.3125 px (OpenCV's
    INTER_BITS = 5): the result equals torch's bicubic at 0.3125, not 0.3."""
    H, W = 24, 40
    f2 = np.random.default_rng(11).random((H, W)).astype(np.float32)
    yy, xx = np.meshgrid(np.arange(H, dtype=np.float64), np.arange(W, dtype=np.float64),
                         indexing="ij")
    out = oracle.imregister_wrapper(f2, np.full((H, W), 0.3), np.zeros((H, W)), f2)
    q = _torch_warp(f2, xx + 0.3125, yy)
    c = _torch_warp(f2, xx + 0.3, yy)
    inner = (slice(3, -3), slice(3, -3))
    np.testing.assert_allclose(out[inner], q[inner], atol=1e-5)
    assert np.abs(out[inner] - c[inner]).max() > 1e-3


def test_warp_matches_torch_bicubic_at_quantized_positions(oracle):
    rng = np.random.default_rng(4)
    H, W = 30, 36
    f2 = rng.random((H, W)).astype(np.float32)
    u = rng.uniform(-0.9, 0.9, (H, W))
    v = rng.uniform(-0.9, 0.9, (H, W))
    out = oracle.imregister_wrapper(f2, u, v, f2)
    yy, xx = np.meshgrid(np.arange(H), np.arange(W), indexing="ij")
    mx = np.round((xx + u).astype(np.float32) * 32) / 32
    my = np.round((yy + v).astype(np.float32) * 32) / 32
    ref = _torch_warp(f2, mx, my)
    np.testing.assert_allclose(out[3:-3, 3:-3], ref[3:-3, 3:-3], atol=1e-5)


def _cv_coeffs(x):
    """OpenCV's interpolateCubic (A = -0.75) in float32, from its public
    source: the 4 weights at fraction x."""
    A = np.float32(-0.75)
    x = np.float32(x)
    one = np.float32(1)
    c0 = ((A * (x + one) - np.float32(5) * A) * (x + one) + np.float32(8) * A) * (x + one) - np.float32(4) * A
    c1 = ((A + np.float32(2)) * x - (A + np.float32(3))) * x * x + one
    c2 = ((A + np.float32(2)) * (one - x) - (A + np.float32(3))) * (one - x) * (one - x) + one
    c3 = one - c0 - c1 - c2
    return [c0, c1, c2, c3]


def test_warp64_reference_warp_tests(oracle):
    """tests/core/test_warping.py:22-33,250-360 on float64 images (the level
    images of get_displacement): zero shift is the identity, integer shifts
    copy, and out-of-bounds pixels take the fixed frame rounded to float32."""
    rng = np.random.default_rng(21)
    H, W = 30, 34
    f2 = rng.random((H, W, 2)).astype(np.float32).astype(np.float64)
    f1 = rng.random((H, W, 2))  # not float32-valued
    z = np.zeros((H, W))
    np.testing.assert_array_equal(oracle.imregister_wrapper(f2, z, z, f1), f2.astype(np.float32))
    out = oracle.imregister_wrapper(f2, z + 4.0, z - 3.0, f1)
    np.testing.assert_array_equal(out[3:, :-4], f2[:-3, 4:].astype(np.float32))
    np.testing.assert_array_equal(out[:3], f1[:3].astype(np.float32))
    np.testing.assert_array_equal(out[:, -4:], f1[:, -4:].astype(np.float32))


def test_warp64_sums_in_float64(oracle):
    """cv2.remap's 64F path: float weights (products of the float 1-D
    tables), float64 products and sums (row sums added in turn), one
    rounding to float32 at the end; evaluated here with numpy float64 at
    1/32-px positions of interior pixels.  The 32F path (float sums) gives
    different bits on the same data."""
    rng = np.random.default_rng(22)
    H, W = 20, 26
    f2 = (rng.random((H, W)) * 100).astype(np.float32).astype(np.float64)
    ku = rng.integers(-40, 40, (H, W))
    kv = rng.integers(-40, 40, (H, W))
    u, v = ku / 32.0, kv / 32.0
    out = oracle.imregister_wrapper(f2, u, v, f2)
    out32 = oracle.imregister_wrapper(f2.astype(np.float32), u, v, f2.astype(np.float32))
    checked = 0
    for y in range(H):
        for x in range(W):
            X, Y = x * 32 + ku[y, x], y * 32 + kv[y, x]
            sx, sy = (X >> 5) - 1, (Y >> 5) - 1
            if not (0 <= sx < W - 3 and 0 <= sy < H - 3):
                continue
            wx, wy = _cv_coeffs((X & 31) / 32.0), _cv_coeffs((Y & 31) / 32.0)
            total = None
            for i in range(4):
                row = None
                for j in range(4):
                    term = f2[sy + i, sx + j] * np.float64(np.float32(wy[i] * wx[j]))
                    row = term if row is None else row + term
                total = row if total is None else total + row
            assert out[y, x] == np.float32(total), (y, x)
            checked += 1
    assert checked > 200
    assert np.any(out != out32)


@pytest.mark.parametrize("dt", [np.uint16, np.int16])
@pytest.mark.parametrize("method", ["cubic", "linear"])
def test_warp_int16_identity(oracle, dt, method):
    """16-bit raw frames: zero displacement reproduces every sample exactly
    (the weights at fraction 0 are 0, 1, 0, 0), including the type's
    extremes."""
    info = np.iinfo(dt)
    rng = np.random.default_rng(23)
    f2 = rng.integers(info.min, info.max, (14, 17, 2), endpoint=True).astype(dt)
    f2[0, 0, 0], f2[0, 1, 0] = info.min, info.max
    z = np.zeros((14, 17))
    out = oracle.imregister_wrapper(f2, z, z, f2.astype(np.float64), method)
    assert out.dtype == np.float32
    np.testing.assert_array_equal(out, f2.astype(np.float32))


def test_warp_uint16_rounds_half_to_even(oracle):
    """A half-pixel shift with linear taps gives (a + b) / 2, exact in float;
    cv2's saturate_cast rounds it half to even (1.5 -> 2, 2.5 -> 2), where a
    truncating cast would give 1 and 2 and rounding half up 2 and 3.  The
    last column's map (W - 0.5) is clipped to W - 1 and reads that sample."""
    H, W = 5, 8
    row = np.array([1, 2, 2, 3, 7, 8, 100, 101], dtype=np.uint16)
    f2 = np.tile(row, (H, 1))
    u, v = np.full((H, W), 0.5), np.zeros((H, W))
    out = oracle.imregister_wrapper(f2, u, v, np.zeros((H, W)), "linear")
    exp = np.array([2, 2, 2, 5, 8, 54, 100, 101], dtype=np.float32)  # 1.5, 2, 2.5, 5, 7.5, 54, 100.5, clip
    np.testing.assert_array_equal(out, np.tile(exp, (H, 1)))


@pytest.mark.parametrize("dt", [np.uint16, np.int16])
def test_warp_int16_cubic_saturates(oracle, dt):
    """Keys' cubic (A = -0.75) at fraction 1/2 has weights -3/32, 19/32,
    19/32, -3/32: next to a step from lo to hi it overshoots by 3/32 of the
    step on both sides, which the 16-bit cast clamps to the type's range;
    away from the step it reproduces lo, hi and their midpoint."""
    info = np.iinfo(dt)
    lo, hi = int(info.min), int(info.max)
    H, W = 8, 12
    row = np.array([lo] * 6 + [hi] * 6, dtype=dt)
    f2 = np.tile(row, (H, 1))
    u, v = np.full((H, W), 0.5), np.zeros((H, W))
    out = oracle.imregister_wrapper(f2, u, v, np.zeros((H, W)), "cubic")
    ys = slice(1, 5)  # interior rows: sy = y - 1 in [0, H - 3)
    assert np.all(out[ys, 4] == lo)   # taps lo, lo, lo, hi: below lo, clamped
    assert np.all(out[ys, 6] == hi)   # taps lo, hi, hi, hi: above hi, clamped
    mid = np.float32(np.rint(np.float32(lo) * np.float32(0.5) + np.float32(hi) * np.float32(0.5)))
    assert np.all(out[ys, 5] == mid)  # taps lo, lo, hi, hi
    assert np.all(out[ys, 2] == lo) and np.all(out[ys, 8] == hi)
    # the float32 path on the same samples keeps the overshoot
    f32 = oracle.imregister_wrapper(f2.astype(np.float32), u, v, np.zeros((H, W)), "cubic")
    assert np.all(f32[ys, 4] < lo) and np.all(f32[ys, 6] > hi)


# ------------------------------------------------------------------ tensor
def test_motion_tensor_reference_known_answer(oracle):
    """tests/core/test_optical_flow.py:532-551: f = 1e-3 x^2 with hx = 2 gives
    J11 = 0.8 at interior pixels; identical frames give J33 = J13 = 0."""
    hy, hx = 1.0, 2.0
    H, W = 17, 23
    x = np.arange(W, dtype=np.float64) * hx
    f = np.tile(1e-3 * x ** 2, (H, 1)).astype(np.float32)
    J11, J22, J33, J12, J13, J23 = oracle.get_motion_tensor_gc(f, f, hy, hx)
    c = (H // 2, W // 2)
    np.testing.assert_allclose(J11[c], 0.8, atol=0.01)
    np.testing.assert_allclose(J33[c], 0.0, atol=1e-12)
    np.testing.assert_allclose(J13[c], 0.0, atol=1e-12)


def test_motion_tensor_quadratic_closed_form(oracle):
    """f1 = a x^2 + b x y + c y^2 and f2 = f1 + d x on a grid with spacings
    (hx, hy): central differences are exact, so fxx = 2a, fyy = 2c, fxy = b,
    fxt = d, fyt = 0 and J follows in closed form at interior pixels."""
    a, b, c, d = 0.003, -0.002, 0.0015, 0.01
    hy, hx = 1.5, 0.75
    H, W = 15, 19
    yy, xx = np.meshgrid(np.arange(H) * hy, np.arange(W) * hx, indexing="ij")
    f1 = (a * xx ** 2 + b * xx * yy + c * yy ** 2 + 0.3).astype(np.float32)
    f2 = (f1.astype(np.float64) + d * xx).astype(np.float32)
    J = oracle.get_motion_tensor_gc(f1, f2, hy, hx)
    fxx, fyy, fxy, fxt, fyt = 2 * a, 2 * c, b, d, 0.0
    rx = 1 / (fxx ** 2 + fxy ** 2 + 1e-6)
    ry = 1 / (fxy ** 2 + fyy ** 2 + 1e-6)
    want = [rx * fxx ** 2 + ry * fxy ** 2, rx * fxy ** 2 + ry * fyy ** 2,
            rx * fxt ** 2 + ry * fyt ** 2, rx * fxx * fxy + ry * fxy * fyy,
            rx * fxx * fxt + ry * fxy * fyt, rx * fxy * fxt + ry * fyy * fyt]
    inner = (slice(4, -4), slice(4, -4))
    for got, w in zip(J, want):
        np.testing.assert_allclose(got[inner], w, rtol=2e-3, atol=2e-4)
        assert np.all(got[0, :] == 0) and np.all(got[:, -1] == 0)


# ------------------------------------------------------------------ solver
def _random_level(oracle, rng, h, w, K, zero_data=False):
    from scipy.ndimage import gaussian_filter
    f1 = gaussian_filter(rng.random((h, w)), 1.2).astype(np.float32)
    f2 = f1 if zero_data else gaussian_filter(rng.random((h, w)), 1.2).astype(np.float32)
    J = np.zeros((6, h + 2, w + 2, K))
    for k in range(K):
        Jk = oracle.get_motion_tensor_gc(f1, f2, 1.0, 1.0)
        for t in range(6):
            J[t, :, :, k] = Jk[t] * (1 + 0.3 * k)
    W = np.zeros((h + 2, w + 2, K))
    W[1:-1, 1:-1] = rng.uniform(0.3, 1.0, (h, w, K))
    u = np.pad(rng.standard_normal((h, w)) * 0.2, 1, mode="edge")
    v = np.pad(rng.standard_normal((h, w)) * 0.2, 1, mode="edge")
    return J, W, u, v


def _dense_system(J, W, u, v, ax, ay, psi):
    """Linear system of one lag block in unknown order (pixel column-major,
    du before dv): returns A (true Neumann operator), Dp (the solver's
    diagonal incl. out-of-level weights) and b."""
    m, n, K = J[0].shape
    h, w = m - 2, n - 2
    N = h * w

    def idx(r, c):
        return c * h + r
    A = np.zeros((2 * N, 2 * N))
    Dp = np.zeros(2 * N)
    b = np.zeros(2 * N)
    for c in range(w):
        for r in range(h):
            p = idx(r, c)
            jr, jc = r + 1, c + 1
            wp = W[jr, jc] * psi[jr, jc]
            a11 = (wp * J[0][jr, jc]).sum()
            a22 = (wp * J[1][jr, jc]).sum()
            a12 = (wp * J[3][jr, jc]).sum()
            b1 = -(wp * J[4][jr, jc]).sum()
            b2 = -(wp * J[5][jr, jc]).sum()
            ws = 0.0
            for (dr, dc, wt) in [(0, -1, ax), (0, 1, ax), (1, 0, ay), (-1, 0, ay)]:
                rr, cc = r + dr, c + dc
                ws += wt
                b1 += wt * (u[jr + dr, jc + dc] - u[jr, jc])
                b2 += wt * (v[jr + dr, jc + dc] - v[jr, jc])
                if 0 <= rr < h and 0 <= cc < w:
                    q = idx(rr, cc)
                    A[2 * p, 2 * q] -= wt
                    A[2 * p + 1, 2 * q + 1] -= wt
                    A[2 * p, 2 * p] += wt
                    A[2 * p + 1, 2 * p + 1] += wt
            A[2 * p, 2 * p] += a11
            A[2 * p + 1, 2 * p + 1] += a22
            A[2 * p, 2 * p + 1] = a12
            A[2 * p + 1, 2 * p] = a12
            Dp[2 * p] = ws + a11
            Dp[2 * p + 1] = ws + a22
            b[2 * p], b[2 * p + 1] = b1, b2
    return A, Dp, b


def _to_field(x, h, w):
    du = x[0::2].reshape(w, h).T
    dv = x[1::2].reshape(w, h).T
    return du, dv


def test_solver_zero_flow_for_identical_frames(oracle):
    rng = np.random.default_rng(5)
    J, W, _, _ = _random_level(oracle, rng, 11, 13, 2, zero_data=True)
    u = np.full((13, 15), 0.7)
    v = np.full((13, 15), -0.2)
    f = oracle.compute_flow(*J, W, u, v, 2, 2, 50, 5, [0.45, 0.45], 1.0, 1.0, 1.0)
    assert np.all(f == 0)


def test_solver_converges_to_dense_solve(oracle):
    """a_data = 1 makes psi = 1: after many sweeps the SOR iterate equals the
    solution of the 2N x 2N linear system."""
    rng = np.random.default_rng(6)
    h, w = 6, 7
    J, W, u, v = _random_level(oracle, rng, h, w, 2)
    ax, ay = 2.0 / 1.3 ** 2, 2.5 / 0.8 ** 2
    f = oracle.compute_flow(*J, W, u, v, 2.0, 2.5, 3000, 10 ** 6, [1.0, 1.0], 1.0, 1.3, 0.8)
    A, _, b = _dense_system(J, W, u, v, ax, ay, np.ones_like(W))
    du, dv = _to_field(np.linalg.solve(A, b), h, w)
    np.testing.assert_allclose(f[1:-1, 1:-1, 0], du, atol=1e-9)
    np.testing.assert_allclose(f[1:-1, 1:-1, 1], dv, atol=1e-9)


@pytest.mark.parametrize("sweeps", [1, 2, 3])
def test_solver_sweeps_equal_matrix_form_sor(oracle, sweeps):
    """x <- x + M^-1 (b - A x), M = Dp/omega + strictly-lower(A) in column-major
    order with du before dv: this is lexicographic SOR; red-black or
    row-major orderings fail it."""
    rng = np.random.default_rng(7)
    h, w = 5, 6
    J, W, u, v = _random_level(oracle, rng, h, w, 1)
    ax, ay = 2.0 / 1.1 ** 2, 2.0 / 0.9 ** 2
    A, Dp, b = _dense_system(J, W, u, v, ax, ay, np.ones_like(W))
    M = np.tril(A, -1) + np.diag(Dp / 1.95)
    x = np.zeros(2 * h * w)
    for _ in range(sweeps):
        x = x + np.linalg.solve(M, b - A @ x)
    f = oracle.compute_flow(*J, W, u, v, 2.0, 2.0, sweeps, 10 ** 6, [1.0], 1.0, 1.1, 0.9)
    du, dv = _to_field(x, h, w)
    np.testing.assert_allclose(f[1:-1, 1:-1, 0], du, atol=1e-12)
    np.testing.assert_allclose(f[1:-1, 1:-1, 1], dv, atol=1e-12)
    if sweeps == 2:
        rb = oracle.compute_flow(*J, W, u, v, 2.0, 2.0, sweeps, 10 ** 6, [1.0], 1.0, 1.1, 0.9,
                                 order=1)
        assert np.abs(rb[1:-1, 1:-1, 0] - du).max() > 1e-6


def test_solver_psi_refresh_matches_charbonnier(oracle):
    """update_lag = 1 refreshes psi before sweep 0 from du = dv = 0:
    psi_k = a (J33_k + 1e-5)^(a - 1), then one matrix-form SOR sweep."""
    rng = np.random.default_rng(8)
    h, w = 5, 5
    J, W, u, v = _random_level(oracle, rng, h, w, 2)
    a = 0.45
    psi = a * (np.maximum(J[2], 0) + 1e-5) ** (a - 1)
    ax = ay = 2.0
    A, Dp, b = _dense_system(J, W, u, v, ax, ay, psi)
    M = np.tril(A, -1) + np.diag(Dp / 1.95)
    x = np.linalg.solve(M, b)
    f = oracle.compute_flow(*J, W, u, v, 2.0, 2.0, 1, 1, [a, a], 1.0, 1.0, 1.0)
    du, dv = _to_field(x, h, w)
    np.testing.assert_allclose(f[1:-1, 1:-1, 0], du, rtol=1e-10, atol=1e-12)
    np.testing.assert_allclose(f[1:-1, 1:-1, 1], dv, rtol=1e-10, atol=1e-12)


def test_solver_literal_reference_order_agrees(oracle):
    """The lag-block coefficient form and the reference's per-sweep expression
    order (order=2, libm pow) agree to rounding on a smooth problem."""
    rng = np.random.default_rng(9)
    J, W, u, v = _random_level(oracle, rng, 16, 19, 2)
    args = (*J, W, u, v, 2.0, 2.0, 50, 5, [0.45, 0.45], 1.0, 1.2, 0.9)
    a = oracle.compute_flow(*args)
    b = oracle.compute_flow(*args, order=2)
    np.testing.assert_allclose(a, b, atol=1e-12)


def _load_detpow():
    """Compile include/flowreg_detmath.h alone (gcc, no contraction)."""
    import ctypes
    import os
    import subprocess
    import tempfile
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    src = ('#include "%s/include/flowreg_detmath.h"\n'
           "double dp(double x, double b){return frh_det_pow(x,b);}\n"
           "float dpf(float x, float b){return frh_det_powf(x,b);}\n" % root)
    d = tempfile.mkdtemp()
    c, so = os.path.join(d, "dp.c"), os.path.join(d, "dp.so")
    open(c, "w").write(src)
    subprocess.run(["gcc", "-O2", "-fPIC", "-shared", "-ffp-contract=off", "-o", so, c, "-lm"],
                   check=True)
    L = ctypes.CDLL(so)
    L.dp.restype, L.dp.argtypes = ctypes.c_double, [ctypes.c_double, ctypes.c_double]
    L.dpf.restype, L.dpf.argtypes = ctypes.c_float, [ctypes.c_float, ctypes.c_float]
    return L


def test_det_pow_accuracy():
    """frh_det_pow against numpy's pow: relative error below 4e-15."""
    L = _load_detpow()
    xs = np.geomspace(1e-5, 1e4, 3000)
    for bexp in (-0.55, -0.9, 0.3, -0.1, 0.999):
        got = np.array([L.dp(x, bexp) for x in xs])
        assert np.max(np.abs(got / np.power(xs, bexp) - 1)) < 4e-15


def test_det_powf_accuracy():
    L = _load_detpow()
    xs = np.geomspace(1e-5, 1e4, 500).astype(np.float32)
    for bexp in (-0.55, 0.3):
        got = np.array([L.dpf(float(x), bexp) for x in xs], dtype=np.float64)
        assert np.max(np.abs(got / np.power(xs.astype(np.float64), bexp) - 1)) < 2e-6


# ------------------------------------------------------- get_displacement
def test_driver_binds_geometric_spacings(oracle):
    """tests/core/test_optical_flow.py:471-500: the solver receives
    hx = n / cols and hy = m / rows (32 x 96 input)."""
    rng = np.random.default_rng(7)
    m, n = 32, 96
    recs = []

    def solver(*args):
        J11 = args[0]
        recs.append(dict(rows=J11.shape[0] - 2, cols=J11.shape[1] - 2, hx=args[15], hy=args[16]))
        return np.zeros_like(args[7]), np.zeros_like(args[8])
    flow = oracle.get_displacement(rng.random((m, n)), rng.random((m, n)), iterations=1,
                                   update_lag=1, level_solver_backend=solver)
    assert flow.shape == (m, n, 2) and len(recs) > 1
    for r in recs:
        assert math.isclose(r["hx"], n / r["cols"], rel_tol=1e-12)
        assert math.isclose(r["hy"], m / r["rows"], rel_tol=1e-12)
    assert any(not math.isclose(r["hx"], r["hy"]) for r in recs)


def test_driver_recovers_constant_shift(oracle):
    """Loose end-to-end sanity (cf. tests/session/test_session_integration.py:
    165-240): a constant shift of the scene is recovered on average."""
    from pyflowreg_amd import synth
    H = W = 96
    scene = synth.blob_scene(H, W, 200, seed=3)
    d = torch.zeros((1, 2, H, W), dtype=torch.float64)
    d[:, 0] = 1.5
    d[:, 1] = -0.75
    mov = synth.warp_scene(scene[None], d)[0, 0].numpy().astype(np.float32)
    fix = scene.numpy().astype(np.float32)
    ref = oracle.preprocess(fix[:, :, None].astype(np.float64)).astype(np.float32)
    mv = oracle.preprocess(mov[:, :, None], fix[:, :, None].astype(np.float64)).astype(np.float32)
    w = oracle.get_displacement(ref, mv)
    inner = w[16:-16, 16:-16]
    assert abs(inner[..., 0].mean() - 1.5) < 0.1
    assert abs(inner[..., 1].mean() + 0.75) < 0.1


def test_preprocess_normalizes_with_reference_range(oracle):
    """util/image_processing.py:90-98: (x - min(ref)) / (max(ref) - min(ref) + eps)."""
    ref = np.array([[[2.0], [6.0]]])
    arr = np.array([[[4.0], [8.0]]])
    np.testing.assert_allclose(oracle.normalize(arr, ref), (arr - 2) / (4 + 1e-8))


# ------------------------------------------- nonlinear smoothness (a_smooth != 1)
def test_psi_smooth_matches_numpy_gradient(oracle):
    """core/level_solver.py:69-155: one-sided differences on the padded edge
    and central ones inside are numpy's np.gradient (edge_order 1); psi_s =
    a (t + 1e-5)^(a - 1) with numpy's pow."""
    rng = np.random.default_rng(30)
    m, n = 9, 12
    u, du, v, dv = (rng.standard_normal((m, n)) for _ in range(4))
    a, hx, hy = 0.6, 1.3, 0.7
    ps = oracle.psi_smooth(u, du, v, dv, a, hx, hy)
    uy, ux = np.gradient(u + du, hy, hx)
    vy, vx = np.gradient(v + dv, hy, hx)
    ref = a * np.power(ux * ux + uy * uy + vx * vx + vy * vy + 1e-5, a - 1)
    np.testing.assert_allclose(ps, ref, rtol=4e-15)


def _dense_system_nl(J, W, u, v, ax, ay, psi, psi_s):
    """_dense_system with the edge weights 0.5 (psi_s[c] + psi_s[nb]) w of
    core/level_solver.py:277-316."""
    m, n, K = J[0].shape
    h, w = m - 2, n - 2
    N = h * w
    A = np.zeros((2 * N, 2 * N))
    Dp = np.zeros(2 * N)
    b = np.zeros(2 * N)
    for c in range(w):
        for r in range(h):
            p = c * h + r
            jr, jc = r + 1, c + 1
            wp = W[jr, jc] * psi[jr, jc]
            a11 = (wp * J[0][jr, jc]).sum()
            a22 = (wp * J[1][jr, jc]).sum()
            a12 = (wp * J[3][jr, jc]).sum()
            b1 = -(wp * J[4][jr, jc]).sum()
            b2 = -(wp * J[5][jr, jc]).sum()
            ws = 0.0
            for (dr, dc, wd) in [(0, -1, ax), (0, 1, ax), (1, 0, ay), (-1, 0, ay)]:
                wt = 0.5 * (psi_s[jr, jc] + psi_s[jr + dr, jc + dc]) * wd
                rr, cc = r + dr, c + dc
                ws += wt
                b1 += wt * (u[jr + dr, jc + dc] - u[jr, jc])
                b2 += wt * (v[jr + dr, jc + dc] - v[jr, jc])
                if 0 <= rr < h and 0 <= cc < w:
                    q = cc * h + rr
                    A[2 * p, 2 * q] -= wt
                    A[2 * p + 1, 2 * q + 1] -= wt
                    A[2 * p, 2 * p] += wt
                    A[2 * p + 1, 2 * p + 1] += wt
            A[2 * p, 2 * p] += a11
            A[2 * p + 1, 2 * p + 1] += a22
            A[2 * p, 2 * p + 1] = a12
            A[2 * p + 1, 2 * p] = a12
            Dp[2 * p] = ws + a11
            Dp[2 * p + 1] = ws + a22
            b[2 * p] = b1
            b[2 * p + 1] = b2
    return A, Dp, b


@pytest.mark.parametrize("sweeps", [1, 2])
def test_solver_nonlinear_smoothness_matrix_form(oracle, sweeps):
    """update_lag = 1 with a_smooth = 0.6: psi_s is refreshed before sweep 0
    from du = 0 (psi_smooth of u, v alone) and every later sweep refreshes
    again; the first sweep equals one matrix-form SOR step with the edge
    weights 0.5 (psi_s[c] + psi_s[nb]) alpha / h^2."""
    rng = np.random.default_rng(31)
    h, w = 5, 6
    J, W, u, v = _random_level(oracle, rng, h, w, 1)
    a_s, hx, hy = 0.6, 1.1, 0.9
    ax, ay = 2.0 / hx ** 2, 2.0 / hy ** 2
    psi = np.ones_like(W)
    z = np.zeros_like(u)
    ps = oracle.psi_smooth(u, z, v, z, a_s, hx, hy)
    A, Dp, b = _dense_system_nl(J, W, u, v, ax, ay, psi, ps)
    M = np.tril(A, -1) + np.diag(Dp / 1.95)
    x = np.linalg.solve(M, b)
    f = oracle.compute_flow(*J, W, u, v, 2.0, 2.0, 1, 1, [1.0], a_s, hx, hy)
    du, dv = _to_field(x, h, w)
    np.testing.assert_allclose(f[1:-1, 1:-1, 0], du, rtol=1e-10, atol=1e-12)
    np.testing.assert_allclose(f[1:-1, 1:-1, 1], dv, rtol=1e-10, atol=1e-12)
    if sweeps == 2:
        # the second refresh sees the padded ring that set_boundary_2d copied
        # before sweep 0 (zeros) and the interior after sweep 0
        du_p = np.zeros_like(u)
        dv_p = np.zeros_like(u)
        du_p[1:-1, 1:-1], dv_p[1:-1, 1:-1] = du, dv
        ps2 = oracle.psi_smooth(u, du_p, v, dv_p, a_s, hx, hy)
        A2, Dp2, b2 = _dense_system_nl(J, W, u, v, ax, ay, psi, ps2)
        M2 = np.tril(A2, -1) + np.diag(Dp2 / 1.95)
        x2 = x + np.linalg.solve(M2, b2 - A2 @ x)
        f2 = oracle.compute_flow(*J, W, u, v, 2.0, 2.0, 2, 1, [1.0], a_s, hx, hy)
        du2, dv2 = _to_field(x2, h, w)
        np.testing.assert_allclose(f2[1:-1, 1:-1, 0], du2, rtol=1e-9, atol=1e-12)
        np.testing.assert_allclose(f2[1:-1, 1:-1, 1], dv2, rtol=1e-9, atol=1e-12)


def test_solver_nonlinear_smoothness_literal_order_agrees(oracle):
    """With a_smooth = 0.5 and lag 5 (several psi_s refreshes over stale
    set_boundary rings) the lag-block coefficient form equals the reference's
    per-sweep expression order to rounding."""
    rng = np.random.default_rng(32)
    J, W, u, v = _random_level(oracle, rng, 14, 17, 2)
    args = (*J, W, u, v, 2.0, 2.0, 50, 5, [0.45, 0.45], 0.5, 1.2, 0.9)
    a = oracle.compute_flow(*args)
    b = oracle.compute_flow(*args, order=2)
    np.testing.assert_allclose(a, b, atol=1e-11)


# ------------------------------------------------ 8-bit frames (cv2's 8U remap)
def test_remap_u8_weight_tables(oracle):
    """or_remap_u8_weights: every table cell sums to exactly 2^15; the
    bilinear cells are the exact products (32 - fx)(32 - fy) * 32 except cell
    (0, 0), whose weight 1 saturates to 32767 and gets the missing unit at
    (1, 1); the bicubic cells are torch's Keys (A = -0.75) weight products
    times 2^15 to within 2 units (float32 table, rounding, the sum fix)."""
    for fy in range(32):
        for fx in range(32):
            lin = oracle.remap_u8_weights(True, fx, fy)
            assert lin.sum() == 32768
            want = np.outer([32 - fy, fy], [32 - fx, fx]) * 32
            if fx == 0 and fy == 0:
                np.testing.assert_array_equal(lin, [[32767, 0], [0, 1]])
            else:
                np.testing.assert_array_equal(lin, want)
    kx = _torch_keys_at(8, 3.0 + np.arange(32) / 32.0)[:, 2:6]   # taps 2..5 of source 3 + k/32
    fixed = 0
    for fy in range(32):
        for fx in range(32):
            cub = oracle.remap_u8_weights(False, fx, fy)
            assert cub.sum() == 32768
            if fx == 0 and fy == 0:
                continue                        # the saturated cell, below
            off = np.abs(cub - np.outer(kx[fy], kx[fx]) * 32768)
            # rounding leaves every cell within half a unit (plus the float32
            # table's error); the sum fix moves one cell of the lower-right
            # central window by the rounding residue
            big = np.argwhere(off > 0.52)
            assert len(big) <= 1
            if len(big):
                assert tuple(big[0]) in {(2, 2), (2, 3), (3, 2), (3, 3)} and off.max() < 4
                fixed += 1
    assert fixed > 50
    np.testing.assert_array_equal(oracle.remap_u8_weights(False, 0, 0),
                                  [[0, 0, 0, 0], [0, 32767, 0, 0], [0, 0, 1, 0], [0, 0, 0, 0]])


@pytest.mark.parametrize("method", ["cubic", "linear"])
def test_warp_u8_identity_and_integer_shift(oracle, method):
    """Reference warp tests (tests/core/test_warping.py:232-447) on uint8:
    zero shift returns every sample, including 0 and 255; an integer shift
    copies the content; positions mapped outside take the fixed frame."""
    rng = np.random.default_rng(21)
    H, W = 20, 24
    f2 = rng.integers(0, 256, (H, W, 2), dtype=np.uint8)
    f2[0, 0] = 0
    f2[1, 1] = 255
    f1 = rng.integers(0, 256, (H, W, 2), dtype=np.uint8).astype(np.float32)
    z = np.zeros((H, W))
    np.testing.assert_array_equal(oracle.imregister_wrapper(f2, z, z, f1, method), f2.astype(np.float32))
    out = oracle.imregister_wrapper(f2, z + 3.0, z - 2.0, f1, method)
    np.testing.assert_array_equal(out[2:, :W - 3], f2[:H - 2, 3:].astype(np.float32))
    np.testing.assert_array_equal(out[:2], f1[:2])
    np.testing.assert_array_equal(out[:, W - 3:], f1[:, W - 3:])


def test_warp_u8_linear_half_pixel_rounds_half_up(oracle):
    """A half-pixel linear shift weighs both taps 2^14: the 8U path stores
    (a + b + 1) >> 1, i.e. halves round up (2.5 -> 3), where the 16-bit
    path rounds half to even (2.5 -> 2) and the float path keeps 2.5."""
    H, W = 4, 8
    f2 = np.tile(np.array([2, 3, 1, 2, 254, 255, 7, 7], np.uint8), (H, 1))
    u = np.full((H, W), 0.5)
    out = oracle.imregister_wrapper(f2, u, np.zeros((H, W)), f2, "linear")
    want = (f2[:, :-1].astype(np.int64) + f2[:, 1:] + 1) // 2
    np.testing.assert_array_equal(out[:, :-1], want.astype(np.float32))
    assert out[0, 0] == 3 and out[0, 4] == 255


def test_warp_u8_cubic_saturates_and_tracks_torch(oracle):
    """Keys' cubic next to a 0 -> 255 step overshoots on both sides; the 8U
    path clamps to [0, 255] instead of wrapping, and elsewhere it is torch's
    bicubic at the 1/32-px-rounded positions to within half a count."""
    H, W = 16, 32
    f2 = np.zeros((H, W), np.uint8)
    f2[:, W // 2:] = 255
    u = np.full((H, W), 0.4)
    out = oracle.imregister_wrapper(f2, u, np.zeros((H, W)), f2, "cubic")
    assert out.min() >= 0 and out.max() <= 255
    f32 = oracle.imregister_wrapper(f2.astype(np.float32), u, np.zeros((H, W)), f2.astype(np.float32), "cubic")
    assert f32.min() < -1 and f32.max() > 256          # the float path keeps the overshoot
    inner = (slice(3, -3), slice(3, -4))
    np.testing.assert_array_equal(out[inner], np.clip(np.floor(f32[inner] + 0.5), 0, 255))
    rng = np.random.default_rng(5)
    img = rng.integers(0, 256, (24, 28), dtype=np.uint8)
    uu = rng.uniform(-0.9, 0.9, (24, 28))
    vv = rng.uniform(-0.9, 0.9, (24, 28))
    got = oracle.imregister_wrapper(img, uu, vv, img, "cubic")
    yy, xx = np.meshgrid(np.arange(24), np.arange(28), indexing="ij")
    mx = np.round((xx + uu).astype(np.float32) * 32) / 32
    my = np.round((yy + vv).astype(np.float32) * 32) / 32
    ref = np.clip(_torch_warp(img.astype(np.float32), mx, my), 0, 255)
    assert np.abs(got[3:-3, 3:-3] - ref[3:-3, 3:-3]).max() <= 0.5 + 0.02


@pytest.mark.parametrize("dt", [np.int8, np.int32, np.uint32])
def test_warp_rejects_dtypes_cv2_remap_lacks(oracle, dt):
    f2 = np.zeros((8, 8), dt)
    with pytest.raises(ValueError):
        oracle.imregister_wrapper(f2, np.zeros((8, 8)), np.zeros((8, 8)), f2)


def test_full_pyramid_literal_reference_order_agrees(oracle):
    """get_displacement end to end (a 64x64x2 synthetic pair, 13 levels):
    the lag-block coefficient form that the GPU kernels share and the
    reference's literal per-sweep expression order with libm pow
    (core/level_solver.py:229-364) agree to 1e-12 px over the whole pyramid,
    i.e. the choice of summation order never moves a warp across a 1/32-px
    grid line here."""
    from pyflowreg_amd import synth
    frames, _ = synth.video(2, 64, 64, n_channels=2, max_disp=2.0, sigma_d=12.0, seed=1000)
    f = frames.numpy()
    ref_raw = f[0].astype(np.float64)
    ref = oracle.preprocess(ref_raw).astype(np.float32)
    mov = oracle.preprocess(f[1], ref_raw).astype(np.float32)
    a = oracle.get_displacement(ref, mov, weight=[0.5, 0.5])
    b = oracle.get_displacement(ref, mov, weight=[0.5, 0.5], order=2)
    assert np.abs(a).max() > 0.1
    d = np.abs(a - b).max()
    assert d <= 1e-12, d
