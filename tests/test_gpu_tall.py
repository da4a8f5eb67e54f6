"""Levels of more than 1024 rows (band mode: sweeps in bands of 512 rows)
against the oracle, bit for bit: the stand-alone level solver (one launch
per band from the top; constant and nonlinear smoothness, band boundaries at
rows 512 and 1024), get_displacement on a tall narrow frame (two levels over
1024 rows), the batched solver, and a 2048x2048x1 pair (four levels over 1024
rows).  The batched solver runs a lag block's sweeps of such a level as a
chain of work-groups (sweeps x bands) when they fit the device: the batch of
30 splits each block into launches of 2 sweeps, and nonlinear smoothness
runs its nine-coefficient form with the border ring."""

import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _smooth(rng, h, w, K):
    from scipy.ndimage import gaussian_filter
    img = rng.random((h, w, K))
    for k in range(K):
        img[:, :, k] = gaussian_filter(img[:, :, k], 1.5)
    return img.astype(np.float32)


def _level_inputs(oracle, rng, h, w, K):
    f1 = _smooth(rng, h, w, K)
    f2 = _smooth(rng, h, w, K)
    J = np.zeros((6, h + 2, w + 2, K))
    for k in range(K):
        Jk = oracle.get_motion_tensor_gc(f1[:, :, k], f2[:, :, k], 1.1, 0.9)
        for t in range(6):
            J[t, :, :, k] = Jk[t]
    W = np.zeros((h + 2, w + 2, K))
    W[1:-1, 1:-1] = rng.uniform(0.2, 1.0, (h, w, K)).astype(np.float32)
    u = np.pad(rng.standard_normal((h, w)).astype(np.float32).astype(np.float64), 1, mode="edge")
    v = np.pad(rng.standard_normal((h, w)).astype(np.float32).astype(np.float64), 1, mode="edge")
    return J, W, u, v


@pytest.mark.parametrize("h,w,K,its,lag,a_s", [(1100, 23, 1, 7, 5, 1.0), (1537, 9, 2, 4, 2, 1.0),
                                                (1030, 14, 1, 6, 5, 0.6)])
def test_tall_level_solver_bit_exact(gpu, oracle, h, w, K, its, lag, a_s):
    rng = np.random.default_rng(h)
    J, W, u, v = _level_inputs(oracle, rng, h, w, K)
    ad = np.full(K, 0.45)
    args = (J[0], J[1], J[2], J[3], J[4], J[5], W, u, v, [2.0, 2.5], its, lag, 0, ad, a_s, 1.3, 0.8)
    du, dv = gpu.level_solver(*args)
    ou, ov = oracle.level_solver(*args)
    np.testing.assert_array_equal(du, ou)
    np.testing.assert_array_equal(dv, ov)
    assert np.abs(ou[1:-1, 1:-1]).max() > 0


def _pair(H, W, C, seed):
    from pyflowreg_amd import synth
    frames, _ = synth.video(2, H, W, n_channels=C, max_disp=3.0, sigma_d=32.0, seed=seed)
    return frames[0].numpy(), frames[1].numpy()


def test_tall_frame_get_displacement_and_batch_bit_exact(gpu, oracle):
    """1300 x 300 x 1: levels of 1300 and 1040 rows run in band mode."""
    import torch
    from pyflowreg_amd.optical_flow import DisplacementSolver
    f, m = _pair(1300, 300, 1, 5101)
    _, m2 = _pair(1300, 300, 1, 5102)
    ref = oracle.preprocess(f.astype(np.float64)).astype(np.float32)
    movs = oracle.preprocess(np.stack([m, m2]), f.astype(np.float64)).astype(np.float32)
    _say("1300x300: GPU")
    a = gpu.get_displacement(ref, movs[0])
    s = DisplacementSolver(1300, 300, 1, max_batch=2)
    s.set_reference(ref, None, None)
    out = s.run(torch.from_numpy(movs).cuda(), out_dtype=torch.float64).cpu().numpy()
    s.close()
    for t in range(2):
        _say(f"1300x300: oracle frame {t}")
        b = oracle.get_displacement(ref, movs[t])
        if t == 0:
            d = np.abs(a - b)
            assert d.max() == 0.0, f"get_displacement: max |dflow| {d.max():.3e}"
        d = np.abs(out[t] - b)
        assert d.max() == 0.0, f"batched frame {t}: max |dflow| {d.max():.3e}"


@pytest.mark.timeout(900)
def test_2048x2048x1_bit_exact(gpu, oracle):
    f, m = _pair(2048, 2048, 1, 5103)
    ref = oracle.preprocess(f.astype(np.float64)).astype(np.float32)
    mov = oracle.preprocess(m, f.astype(np.float64)).astype(np.float32)
    _say("2048x2048: GPU")
    a = gpu.get_displacement(ref, mov)
    _say("2048x2048: oracle")
    b = oracle.get_displacement(ref, mov)
    d = np.abs(a - b)
    assert d.max() == 0.0, f"2048x2048x1: max |dflow| {d.max():.3e}"
    assert np.abs(b).max() > 0.5


def test_tall_batch_chains_bit_exact(gpu, oracle):
    """30 frames of 1300 x 20 x 1: the 1300- and 1040-row levels' lag blocks
    run as chains of 3 bands x 2 sweeps per launch (a 5-sweep block takes 3
    launches, the first of a level with zero start values)."""
    import torch
    from pyflowreg_amd import synth
    from pyflowreg_amd.optical_flow import DisplacementSolver
    T, H, W = 30, 1300, 20
    frames, _ = synth.video(T + 1, H, W, n_channels=1, max_disp=3.0, sigma_d=32.0, seed=5111)
    f = frames[0].numpy()
    ref = oracle.preprocess(f.astype(np.float64)).astype(np.float32)
    movs = oracle.preprocess(frames[1:].numpy(), f.astype(np.float64)).astype(np.float32)
    _say("1300x20 x 30: GPU")
    s = DisplacementSolver(H, W, 1, max_batch=T)
    s.set_reference(ref, None, None)
    out = s.run(torch.from_numpy(movs).cuda(), out_dtype=torch.float64).cpu().numpy()
    s.close()
    for t in (0, 13, 29):
        _say(f"1300x20: oracle frame {t}")
        d = np.abs(out[t] - oracle.get_displacement(ref, movs[t]))
        assert d.max() == 0.0, f"frame {t}: max |dflow| {d.max():.3e}"


def test_tall_batch_nonlinear_smoothness_bit_exact(gpu, oracle):
    """a_smooth = 0.7 on 2 frames of 1100 x 40 x 2: the 1100-row level's
    chains carry nine coefficients per cell and record the border ring that
    the next psi_s refresh reads."""
    import torch
    from pyflowreg_amd.optical_flow import DisplacementSolver
    f, m = _pair(1100, 40, 2, 5121)
    _, m2 = _pair(1100, 40, 2, 5122)
    ref = oracle.preprocess(f.astype(np.float64)).astype(np.float32)
    movs = oracle.preprocess(np.stack([m, m2]), f.astype(np.float64)).astype(np.float32)
    _say("1100x40x2 a_smooth 0.7: GPU")
    s = DisplacementSolver(1100, 40, 2, max_batch=2, a_smooth=0.7)
    s.set_reference(ref, [0.5, 0.5], None)
    out = s.run(torch.from_numpy(movs).cuda(), out_dtype=torch.float64).cpu().numpy()
    s.close()
    for t in range(2):
        _say(f"1100x40x2: oracle frame {t}")
        b = oracle.get_displacement(ref, movs[t], weight=[0.5, 0.5], a_smooth=0.7)
        d = np.abs(out[t] - b)
        assert d.max() == 0.0, f"frame {t}: max |dflow| {d.max():.3e}"
