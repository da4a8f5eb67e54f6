"""Pins of the oracle's alternative data terms (gray value, census), of its
float32 handling in the motion tensors, and of the GNC solver and stage
loop.  CPU only.

Expected values come from numpy's own np.gradient / np.pad on arrays of the
dtype the reference passes (the library routines the reference calls), from
invariants the reference's tests state for the census tensor
(tests/core/test_optical_flow.py:373-446) and the GNC solver
(tests/core/test_level_solver_gnc.py:63-106), and from call counts of the
stage loop (tests/core/test_optical_flow.py:64-296).
"""

import numpy as np
import pytest


def _images(shape=(9, 11), seed=0):
    from scipy.ndimage import gaussian_filter
    rng = np.random.default_rng(seed)
    f1 = gaussian_filter(rng.random(shape), 1.2).astype(np.float32)
    f2 = gaussian_filter(rng.random(shape), 1.2).astype(np.float32)
    return f1, f2


def _numpy_gray(f1, f2, hy, hx):
    """get_motion_tensor_gray's steps with numpy on the given dtypes."""
    f1p = np.pad(f1, 1, mode="symmetric")
    f2p = np.pad(f2, 1, mode="symmetric")
    fy1, fx1 = np.gradient(f1p, hy, hx)
    fy2, fx2 = np.gradient(f2p, hy, hx)
    fx = 0.5 * (fx1 + fx2)
    fy = 0.5 * (fy1 + fy2)
    ft = f2p - f1p
    return [a[1:-1, 1:-1] for a in (fx * fx, fy * fy, ft * ft, fx * fy, fx * ft, fy * ft)]


@pytest.mark.parametrize("dt1,dt2", [(np.float64, np.float32), (np.float32, np.float32),
                                     (np.float64, np.float64)])
def test_gray_tensor_follows_numpy_dtypes(oracle, dt1, dt2):
    f1, f2 = _images()
    a, b = f1.astype(dt1), f2.astype(dt2)
    J = oracle.get_motion_tensor_gray(a, b, 1.3, 0.7)
    for got, want in zip(J, _numpy_gray(a, b, 1.3, 0.7)):
        np.testing.assert_array_equal(got[1:-1, 1:-1], want.astype(np.float64))
        assert np.all(got[0] == 0) and np.all(got[:, -1] == 0)


def test_gc_tensor_float32_warped_image_differs_from_float64(oracle):
    """Below the top level the warped image is float32 and numpy takes its
    derivatives in float32; the result is not the float64 one."""
    f1, f2 = _images((16, 18), 3)
    mixed = oracle.get_motion_tensor_gc(f1.astype(np.float64), f2, 1.1, 0.9)
    full = oracle.get_motion_tensor_gc(f1.astype(np.float64), f2.astype(np.float64), 1.1, 0.9)
    assert any(not np.array_equal(x, y) for x, y in zip(mixed, full))
    for x, y in zip(mixed, full):
        np.testing.assert_allclose(x, y, rtol=1e-4, atol=1e-6)


def test_gc_tensor_fx_term_follows_numpy(oracle):
    """With f1 = f2 shifted by a constant, ft is constant, so J33 = J13 = J23
    = 0 and J11 = reg_x fxx^2 + reg_y fxy^2 where fxx, fxy come from numpy's
    float32 second differences of f2 averaged with float64 ones of f1."""
    f1, _ = _images((12, 13), 4)
    a = f1.astype(np.float64)
    J = oracle.get_motion_tensor_gc(a, f1, 1.0, 1.0)
    np.testing.assert_array_equal(J[2], 0)
    np.testing.assert_array_equal(J[4], 0)
    f1p = np.pad(a, 1, mode="symmetric")
    f2p = np.pad(f1, 1, mode="symmetric")
    fxx1 = (f1p[1:-1, 0:-2] - 2 * f1p[1:-1, 1:-1] + f1p[1:-1, 2:]) / (1.0 ** 2)
    fxx2 = (f2p[1:-1, 0:-2] - 2 * f2p[1:-1, 1:-1] + f2p[1:-1, 2:]) / (1.0 ** 2)
    assert fxx2.dtype == np.float32
    fxx = 0.5 * (fxx1 + fxx2)
    _, fx1 = np.gradient(f1p, 1.0, 1.0)
    _, fx2 = np.gradient(f2p, 1.0, 1.0)
    fx = np.pad((0.5 * (fx1 + fx2))[1:-1, 1:-1], 1, mode="symmetric")
    fxy = np.gradient(fx, 1.0, 1.0)[0][1:-1, 1:-1]
    reg_x = 1.0 / ((np.sqrt(fxx ** 2 + fxy ** 2) ** 2) + 1e-6)
    fyy1 = (f1p[0:-2, 1:-1] - 2 * f1p[1:-1, 1:-1] + f1p[2:, 1:-1]) / 1.0
    fyy2 = (f2p[0:-2, 1:-1] - 2 * f2p[1:-1, 1:-1] + f2p[2:, 1:-1]) / 1.0
    fyy = 0.5 * (fyy1 + fyy2)
    reg_y = 1.0 / ((np.sqrt(fxy ** 2 + fyy ** 2) ** 2) + 1e-6)
    np.testing.assert_array_equal(J[0][1:-1, 1:-1], reg_x * fxx ** 2 + reg_y * fxy ** 2)


# ------------------------------------------------------------------ census

def test_census_shape_and_zero_border(oracle):
    f1, f2 = _images()
    J = oracle.get_motion_tensor_cs(f1.astype(np.float64), f2.astype(np.float64), 1.0, 1.0)
    for t in J:
        assert t.shape == (11, 13)
        assert np.all(t[0] == 0) and np.all(t[-1] == 0) and np.all(t[:, 0] == 0) and np.all(t[:, -1] == 0)


def test_census_constant_images_near_zero(oracle):
    J = oracle.get_motion_tensor_cs(np.full((7, 6), 0.25), np.full((7, 6), 0.75), 1.0, 1.0)
    for t in J:
        np.testing.assert_allclose(t, 0.0, atol=1e-14)


def test_census_additive_shift_invariance(oracle):
    f1, f2 = _images()
    a, b = f1.astype(np.float64), f2.astype(np.float64)
    base = oracle.get_motion_tensor_cs(a, b, 1.0, 1.0)
    shifted = oracle.get_motion_tensor_cs(a + 4.0, b + 4.0, 1.0, 1.0)
    for x, y in zip(base, shifted):
        np.testing.assert_allclose(y, x, atol=1e-12)


def test_census_default_eps(oracle):
    f1, f2 = _images()
    a, b = f1.astype(np.float64), f2.astype(np.float64)
    for x, y in zip(oracle.get_motion_tensor_cs(a, b, 1.0, 1.0),
                    oracle.get_motion_tensor_cs(a, b, 1.0, 1.0, eps=0.1 / 255.0)):
        np.testing.assert_array_equal(x, y)


def test_census_no_wraparound(oracle):
    f1, f2 = _images((9, 11))
    a, b = f1.astype(np.float64), f2.astype(np.float64)
    a2, b2 = a.copy(), b.copy()
    a2[-1, :] += 0.5
    b2[-1, :] += 0.3
    a2[:, -1] += 0.2
    b2[:, -1] += 0.4
    for x, y in zip(oracle.get_motion_tensor_cs(a, b, 1.0, 1.0), oracle.get_motion_tensor_cs(a2, b2, 1.0, 1.0)):
        np.testing.assert_array_equal(x[1:4, 1:4], y[1:4, 1:4])


def test_census_anisotropic_spacing_finite(oracle):
    f1, f2 = _images((6, 10))
    for t in oracle.get_motion_tensor_cs(f1.astype(np.float64), f2, 0.5, 2.0):
        assert t.shape == (8, 12) and np.all(np.isfinite(t))


def test_census_single_offset_known_answer(oracle):
    """A horizontal ramp f1 = x (hx = 1) with f2 = f1: every directional
    difference is r = dx / |d|, so away from the first and last column (where
    the symmetric pad repeats a sample) the census images are constant per
    offset and all derivatives vanish; next to those columns they do not."""
    x = np.tile(np.arange(8, dtype=np.float64), (6, 1))
    J = oracle.get_motion_tensor_cs(x, x.copy(), 1.0, 1.0)
    for t in J:
        np.testing.assert_allclose(t[1:-1, 3:-3], 0.0, atol=1e-15)
    assert np.all(J[0][1:-1, 2] > 0)


# ------------------------------------------------------------- GNC solver

def _data_case(a_data=0.45):
    shape = (3, 3, 1)
    J = [np.zeros(shape) for _ in range(6)]
    J[0][1, 1, 0] = 1.0
    J[4][1, 1, 0] = -1.0
    J[2][1, 1, 0] = 4.0
    return J, np.ones(shape), np.zeros((3, 3)), np.zeros((3, 3)), np.array([a_data])


def _gnc(oracle, J, W, u, v, ad, lag, beta, a_smooth=1.0, its=1):
    return oracle.level_solver(*J, W, u, v, [1.0, 1.0], its, lag, 0, ad, a_smooth, 1.0, 1.0, beta)


def test_gnc_data_weights_follow_lag_cadence(oracle):
    J, W, u, v, ad = _data_case()
    q0 = _gnc(oracle, J, W, u, v, ad, 2, 0.0)
    q1 = _gnc(oracle, J, W, u, v, ad, 2, 1.0)
    np.testing.assert_allclose(q0, q1)
    a = _gnc(oracle, J, W, u, v, ad, 1, 0.0)[0][1, 1]
    b = _gnc(oracle, J, W, u, v, ad, 1, 0.5)[0][1, 1]
    c = _gnc(oracle, J, W, u, v, ad, 1, 1.0)[0][1, 1]
    assert a > b > c


def test_gnc_quadratic_data_term_is_beta_invariant(oracle):
    J, W, u, v, ad = _data_case(1.0)
    np.testing.assert_allclose(_gnc(oracle, J, W, u, v, ad, 1, 0.0), _gnc(oracle, J, W, u, v, ad, 1, 1.0))


def test_gnc_beta_one_equals_compute_flow(oracle):
    J, W, u, v, ad = _data_case()
    np.testing.assert_array_equal(_gnc(oracle, J, W, u, v, ad, 1, 1.0), _gnc(oracle, J, W, u, v, ad, 1, None))


def _smooth_case():
    J = [np.zeros((5, 5, 1)) for _ in range(6)]
    u = np.zeros((5, 5))
    u[2, 2] = 1.0
    return J, np.ones((5, 5, 1)), u, np.zeros((5, 5)), np.array([1.0])


def test_gnc_blends_sublinear_smoothness(oracle):
    J, W, u, v, ad = _smooth_case()
    a = _gnc(oracle, J, W, u, v, ad, 1, 0.0, a_smooth=0.5)
    b = _gnc(oracle, J, W, u, v, ad, 1, 1.0, a_smooth=0.5)
    assert not np.allclose(a, b)


def test_gnc_quadratic_smoothness_is_beta_invariant(oracle):
    J, W, u, v, ad = _smooth_case()
    np.testing.assert_allclose(_gnc(oracle, J, W, u, v, ad, 1, 0.0), _gnc(oracle, J, W, u, v, ad, 1, 1.0))


# ------------------------------------------------------------ stage loop

def _recorder(calls):
    def solver(*args):
        u, v = args[7], args[8]
        calls.append({"shape": u.shape, "beta": args[17] if len(args) > 17 else None})
        return np.zeros_like(u), np.zeros_like(v)
    return solver


def test_stage_loop_call_counts(oracle):
    f = np.zeros((24, 24))
    n_levels = len(oracle.level_plan(24, 24, 0.8, 50, 0)[0])
    calls = []
    oracle.get_displacement(f, f, level_solver_backend=_recorder(calls), warping_steps=4)
    assert len(calls) == n_levels and all(c["beta"] is None for c in calls)
    calls = []
    oracle.get_displacement(f, f, level_solver_backend=_recorder(calls), gnc_schedule=(0.0, 0.5, 1.0),
                            warping_steps=2)
    assert len(calls) == 3 * n_levels * 2
    assert [c["beta"] for c in calls[:: n_levels * 2]] == [0.0, 0.5, 1.0]
    calls = []
    oracle.get_displacement(f, f, level_solver_backend=_recorder(calls), gnc_schedule=(0.0, 1.0))
    assert len(calls) == 2 * n_levels * 10


def test_stage_loop_rejects_bad_options(oracle):
    f = np.zeros((16, 16))
    with pytest.raises(ValueError, match="Unknown constancy assumption"):
        oracle.get_displacement(f, f, const_assumption="invalid")
    for bad in ([0.5, 1.0], [0.0], [0.0, 0.7, 0.5, 1.0], [0.0, 0.9]):
        with pytest.raises(ValueError):
            oracle.get_displacement(f, f, gnc_schedule=bad)
    with pytest.raises(ValueError):
        oracle.get_displacement(f, f, gnc_schedule=[0.0, 1.0], warping_steps=0)


@pytest.mark.parametrize("kind", ["gray", "brightness", "cs", "census", "gradient"])
def test_stage_loop_recovers_constant_shift(oracle, kind):
    """Every data term recovers a constant sub-pixel shift of a smooth scene."""
    from scipy.ndimage import gaussian_filter, shift
    rng = np.random.default_rng(9)
    scene = gaussian_filter(rng.random((48, 48)), 2.5)
    scene = (scene - scene.min()) / (scene.max() - scene.min())
    moving = shift(scene, (0.6, -0.4), order=3, mode="nearest")
    w = oracle.get_displacement(scene.astype(np.float32), moving.astype(np.float32),
                                const_assumption=kind)
    inner = w[12:-12, 12:-12]
    # moving(y, x) = scene(y - 0.6, x + 0.4); moving(x + w) = scene(x) gives
    # u = -0.4 (columns) and v = +0.6 (rows)
    um, vm = inner[..., 0].mean(), inner[..., 1].mean()
    if kind in ("cs", "census"):
        # with eps = 0.1/255 the smoothed census of a smooth scene is nearly
        # piecewise constant, so its data term is weak: direction and order
        # of magnitude only
        assert -0.5 < um < -0.1 and 0.15 < vm < 0.7
    else:
        np.testing.assert_allclose(um, -0.4, atol=0.12)
        np.testing.assert_allclose(vm, 0.6, atol=0.12)
