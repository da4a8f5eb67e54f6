"""CPU pins of the torch autograd restatement (tests/grad_reference.py) that
the GPU gradient tests compare against: each stage against the oracle (and
through it the reference's semantics), and the refresh schedule."""

import numpy as np
import pytest
import torch
from scipy.ndimage import gaussian_filter, median_filter

import grad_reference as R

F64 = torch.float64


def _scene(rng, H, W, C):
    return np.stack([gaussian_filter(rng.random((H, W)), 2.0) for _ in range(C)], -1) \
        .astype(np.float32).astype(np.float64)


def test_lag_schedule():
    from pyflowreg_amd.autograd import lag_blocks
    assert lag_blocks(10, 5) == [(0, 4, True), (4, 9, False), (9, 10, False)]
    assert lag_blocks(3, 1) == [(0, 1, False), (1, 2, False), (2, 3, False)]
    assert lag_blocks(50, 5)[-1] == (49, 50, False) and len(lag_blocks(50, 5)) == 11
    assert lag_blocks(7, 5) == R.lag_blocks(7, 5)
    with pytest.raises(ValueError):
        lag_blocks(0, 5)


def test_motion_tensor_matches_oracle(oracle):
    rng = np.random.default_rng(1)
    f1, f2 = _scene(rng, 14, 17, 2), _scene(rng, 14, 17, 2)
    J = R.motion_tensor_gc(torch.from_numpy(f1).permute(2, 0, 1), torch.from_numpy(f2).permute(2, 0, 1)[None],
                           1.3, 0.9)[0]
    for c in range(2):
        Jo = oracle.get_motion_tensor_gc(f1[:, :, c], f2[:, :, c], 1.3, 0.9)
        for e in range(6):
            ref = Jo[e][1:-1, 1:-1]
            assert np.abs(J[e, c].numpy() - ref).max() <= 1e-13 * np.abs(ref).max()


@pytest.mark.parametrize("lag", [1, 5])
def test_level_solve_matches_oracle_compute_flow(oracle, lag):
    """The diagonal-order sweep, the refresh schedule and the smoothness
    terms of a nonzero initial flow against the oracle's compute_flow."""
    rng = np.random.default_rng(2)
    H, W, C = 13, 16, 2
    f1, f2 = _scene(rng, H, W, C), _scene(rng, H, W, C)
    u = gaussian_filter(rng.standard_normal((H, W)), 2.0)
    v = gaussian_filter(rng.standard_normal((H, W)), 2.0)
    J = R.motion_tensor_gc(torch.from_numpy(f1).permute(2, 0, 1), torch.from_numpy(f2).permute(2, 0, 1)[None],
                           1.0, 1.0)
    Jo = np.zeros((6, H + 2, W + 2, C))
    Jo[:, 1:-1, 1:-1, :] = J[0].permute(0, 2, 3, 1).numpy()
    wl = np.zeros((H + 2, W + 2, C))
    wl[1:-1, 1:-1] = 0.5
    a = np.array([0.45, 0.6])
    du, dv = oracle.level_solver(*Jo, wl, np.pad(u, 1, mode="edge"), np.pad(v, 1, mode="edge"), [2.0, 1.5],
                                 12, lag, 0, a, 1.0, 1.0, 1.0)
    u0 = torch.from_numpy(np.stack([u, v]))[None]
    x = torch.zeros_like(u0)
    for s, e, psi_one in R.lag_blocks(12, lag):
        coef = R.coefficients(x, u0, J, torch.full((C, H, W), 0.5, dtype=F64), 2.0, 1.5,
                              torch.from_numpy(a), psi_one)
        for _ in range(s, e):
            x = R.sor_sweep(coef, x, 2.0, 1.5)
    scale = max(np.abs(du).max(), np.abs(dv).max())
    assert np.abs(x[0, 0].numpy() - du[1:-1, 1:-1]).max() <= 1e-12 * scale
    assert np.abs(x[0, 1].numpy() - dv[1:-1, 1:-1]).max() <= 1e-12 * scale


def test_median_matches_scipy():
    rng = np.random.default_rng(3)
    x = rng.standard_normal((2, 9, 12))
    y = R.median5(torch.from_numpy(x)).numpy()
    for k in range(2):
        np.testing.assert_array_equal(y[k], median_filter(x[k], size=(5, 5), mode="mirror"))


def test_continuous_warp_matches_oracle_without_quantisation(oracle):
    """The oracle's warp with its 1/32-px quantisation switched off computes
    the continuous map in float32; the torch warp in float64 agrees to
    float32 rounding, including out-of-bounds fill and border clipping."""
    rng = np.random.default_rng(4)
    H, W = 15, 19
    f2, f1 = _scene(rng, H, W, 1), _scene(rng, H, W, 1)
    u = (rng.random((H, W)) - 0.5) * 5
    v = (rng.random((H, W)) - 0.5) * 5
    L = oracle.lib()
    L.or_set_warp_quantize(0)
    try:
        ref = oracle.imregister_wrapper(f2[:, :, 0], u, v, f1[:, :, 0])
    finally:
        L.or_set_warp_quantize(1)
    out = R.warp(torch.from_numpy(f2).permute(2, 0, 1)[None], torch.from_numpy(u)[None],
                 torch.from_numpy(v)[None], torch.from_numpy(f1).permute(2, 0, 1))
    ref = np.asarray(ref, dtype=np.float64).reshape(H, W)
    assert np.abs(out[0, 0].numpy() - ref).max() < 5e-6


def test_single_level_displacement_matches_oracle(oracle):
    rng = np.random.default_rng(5)
    fixed = _scene(rng, 20, 23, 2)
    moving = np.stack([_scene(rng, 20, 23, 2) for _ in range(2)])
    w = R.get_displacement(torch.from_numpy(fixed).permute(2, 0, 1), torch.from_numpy(moving).permute(0, 3, 1, 2),
                           [2.0, 2.0], 0.45, oracle, levels=0)
    for t in range(2):
        ref = oracle.get_displacement(fixed, moving[t], levels=0)
        assert np.abs(w[t].permute(1, 2, 0).numpy() - ref).max() < 1e-11 * max(1.0, np.abs(ref).max())


def test_pyramid_close_to_quantised_reference(oracle):
    """With the pyramid, the continuous float64 warp departs from the
    reference's 1/32-px float32 warp only slightly: the flows agree to a few
    hundredths of a pixel."""
    from pyflowreg_amd import synth
    vid, _ = synth.video(2, 40, 48, 2, seed=31, max_disp=2.0)
    raw = vid.double().numpy()
    fixed = oracle.preprocess(raw[0]).astype(np.float32).astype(np.float64)
    moving = oracle.preprocess(raw[1:], raw[0]).astype(np.float32).astype(np.float64)
    w = R.get_displacement(torch.from_numpy(fixed).permute(2, 0, 1), torch.from_numpy(moving).permute(0, 3, 1, 2),
                           [2.0, 2.0], 0.45, oracle, iterations=20)
    ref = oracle.get_displacement(fixed, moving[0], iterations=20)
    d = np.abs(w[0].permute(1, 2, 0).numpy() - ref)
    print("max", d.max(), "mean", d.mean())
    assert d.mean() < 0.04 and d.max() < 0.3
