"""Alternative data terms, float32 handling in the motion tensors and GNC on
the GPU vs the CPU oracle on identical seeded inputs.  Bar: bit-exact."""

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _img(rng, h, w):
    from scipy.ndimage import gaussian_filter
    return gaussian_filter(rng.random((h, w)), 1.3).astype(np.float32)


@pytest.mark.parametrize("kind", ["gc", "gray", "cs"])
@pytest.mark.parametrize("dt1,dt2", [(np.float64, np.float32), (np.float32, np.float32),
                                     (np.float64, np.float64), (np.float32, np.float64)])
def test_tensor_bit_exact(gpu, oracle, kind, dt1, dt2):
    rng = np.random.default_rng(11)
    f1 = _img(rng, 21, 17).astype(dt1)
    f2 = _img(rng, 21, 17).astype(dt2)
    g = {"gc": gpu.get_motion_tensor_gc, "gray": gpu.get_motion_tensor_gray,
         "cs": gpu.get_motion_tensor_cs}[kind]
    o = {"gc": oracle.get_motion_tensor_gc, "gray": oracle.get_motion_tensor_gray,
         "cs": oracle.get_motion_tensor_cs}[kind]
    for a, b in zip(g(f1, f2, 1.25, 0.8), o(f1, f2, 1.25, 0.8)):
        np.testing.assert_array_equal(a, b)


def test_census_eps_argument(gpu, oracle):
    rng = np.random.default_rng(12)
    f1, f2 = _img(rng, 12, 15).astype(np.float64), _img(rng, 12, 15)
    for a, b in zip(gpu.get_motion_tensor_cs(f1, f2, 1.0, 1.0, eps=0.05),
                    oracle.get_motion_tensor_cs(f1, f2, 1.0, 1.0, eps=0.05)):
        np.testing.assert_array_equal(a, b)


@pytest.mark.parametrize("beta", [0.0, 0.5, 1.0])
@pytest.mark.parametrize("lag", [1, 5])
def test_gnc_level_solver_bit_exact(gpu, oracle, beta, lag):
    rng = np.random.default_rng(13)
    h, w, K = 18, 23, 2
    J = np.zeros((6, h + 2, w + 2, K))
    for k in range(K):
        J[:, :, :, k] = np.stack(oracle.get_motion_tensor_gc(_img(rng, h, w).astype(np.float64),
                                                             _img(rng, h, w), 1.1, 0.9))
    W = np.zeros((h + 2, w + 2, K))
    W[1:-1, 1:-1] = rng.uniform(0.2, 1.0, (h, w, K)).astype(np.float32)
    u = np.pad(rng.standard_normal((h, w)).astype(np.float32).astype(np.float64), 1, mode="edge")
    v = np.pad(rng.standard_normal((h, w)).astype(np.float32).astype(np.float64), 1, mode="edge")
    args = (*J, W, u, v, [2.0, 1.5], 12, lag, 0, np.array([0.45, 0.6]), 1.0, 1.3, 0.8, beta)
    du, dv = gpu.level_solver(*args)
    ou, ov = oracle.level_solver(*args)
    np.testing.assert_array_equal(du, ou)
    np.testing.assert_array_equal(dv, ov)


def _pair(H, W, C, seed):
    from pyflowreg_amd import synth
    frames, _ = synth.video(2, H, W, n_channels=C, max_disp=2.0, sigma_d=H / 6, seed=seed)
    f, m = frames[0].numpy(), frames[1].numpy()
    return f, m


@pytest.mark.parametrize("term", ["gray", "brightness", "cs", "census"])
def test_displacement_data_terms_bit_exact(gpu, oracle, term):
    f, m = _pair(40, 48, 2, 21)
    ref = oracle.preprocess(f.astype(np.float64)).astype(np.float32)
    mov = oracle.preprocess(m, f.astype(np.float64)).astype(np.float32)
    a = gpu.get_displacement(ref, mov, weight=[0.5, 0.5], const_assumption=term)
    b = oracle.get_displacement(ref, mov, weight=[0.5, 0.5], const_assumption=term)
    np.testing.assert_array_equal(a, b)


def test_batched_gray_equals_single(gpu):
    from pyflowreg_amd import synth
    frames, _ = synth.video(3, 48, 40, n_channels=1, max_disp=2.0, sigma_d=8, seed=5)
    fr = frames.numpy()
    s = gpu.DisplacementSolver(48, 40, 1, max_batch=2, const_assumption="gray")
    s.set_reference(fr[0])
    wb = s.run(fr[1:]).cpu().numpy()
    for t in range(2):
        np.testing.assert_array_equal(wb[t], gpu.get_displacement(fr[0], fr[1 + t], const_assumption="gray"))


@pytest.mark.parametrize("uv_dtype", [None, np.float32, np.float64])
def test_gnc_displacement_bit_exact(gpu, oracle, uv_dtype):
    f, m = _pair(24, 28, 1, 31)
    ref = oracle.preprocess(f.astype(np.float64)).astype(np.float32)
    mov = oracle.preprocess(m, f.astype(np.float64)).astype(np.float32)
    kw = dict(gnc_schedule=(0.0, 0.5, 1.0), warping_steps=2, iterations=10)
    if uv_dtype is not None:
        uv = np.zeros((24, 28, 2), uv_dtype)
        uv[..., 0], uv[..., 1] = 0.3, -0.15
        kw["uv"] = uv
    a = gpu.get_displacement(ref, mov, **kw)
    b = oracle.get_displacement(ref, mov, **kw)
    np.testing.assert_array_equal(a, b)


def test_hooked_solver_with_census(gpu, oracle):
    f, m = _pair(32, 36, 1, 41)
    ref = oracle.preprocess(f.astype(np.float64)).astype(np.float32)
    mov = oracle.preprocess(m, f.astype(np.float64)).astype(np.float32)
    a = gpu.get_displacement(ref, mov, const_assumption="cs", level_solver_backend=oracle.level_solver)
    b = oracle.get_displacement(ref, mov, const_assumption="cs")
    np.testing.assert_array_equal(a, b)


def test_uneven_pairs_bit_exact(gpu, oracle):
    """Independent pairs of uneven sizes and channel counts (the uneven-batch
    workload at small sizes): grouped, batched solves equal per-pair oracle
    runs; each pair has its own reference."""
    rng = np.random.default_rng(17)
    sizes = [(32, 32, 1), (40, 48, 2), (32, 32, 1), (56, 40, 1), (40, 48, 2), (32, 32, 2)]
    fixed, moving = [], []
    for k, (H, W, C) in enumerate(sizes):
        f, m = _pair(H, W, C, 100 + k)
        f = f[..., :C]
        m = m[..., :C]
        fixed.append(oracle.preprocess(f.astype(np.float64)).astype(np.float32))
        moving.append(oracle.preprocess(m, f.astype(np.float64)).astype(np.float32))
    ws = gpu.get_displacements(fixed, moving)
    for k in range(len(sizes)):
        np.testing.assert_array_equal(ws[k], oracle.get_displacement(fixed[k], moving[k]))
