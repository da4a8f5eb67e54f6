"""The hot path at the benchmark's real sizes: 512x512x2 frames in one
batched launch against the oracle, and size-independent properties at that
size (identical frames give zero flow; a frame's flow does not depend on
the other frames of its batch).  796x512 (a size of the uneven workload)
puts more than 512 rows into the 1024-lane sweep work-groups."""

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _frames(T, H, W, C, seed):
    from pyflowreg_amd import synth
    v, _ = synth.video(T + 1, H, W, n_channels=C, max_disp=3.0, sigma_d=48.0, seed=seed)
    return v.numpy()


def _prep(oracle, v):
    ref_raw = v[0].astype(np.float64)
    ref = oracle.preprocess(ref_raw).astype(np.float32)
    movs = oracle.preprocess(v[1:], ref_raw).astype(np.float32)
    return ref, movs


def test_batched_512x512x2_bit_exact(gpu, oracle):
    import torch
    from pyflowreg_amd.optical_flow import DisplacementSolver
    v = _frames(3, 512, 512, 2, 60)
    ref, movs = _prep(oracle, v)
    s = DisplacementSolver(512, 512, 2, max_batch=3)
    s.set_reference(ref, np.array([0.5, 0.5]), None)
    out = s.run(torch.from_numpy(movs).cuda(), out_dtype=torch.float64).cpu().numpy()
    for t in (0, 2):
        exp = oracle.get_displacement(ref, movs[t], weight=[0.5, 0.5])
        d = np.abs(out[t] - exp)
        assert d.max() == 0.0, f"frame {t}: max |dflow| {d.max():.3e}"
    s.close()


def test_512x512x2_identical_frames_give_zero_flow(gpu):
    """f1 = f2 makes J13 = J23 = J33 = 0, so every update is exactly 0."""
    import torch
    from pyflowreg_amd.optical_flow import DisplacementSolver
    v = _frames(1, 512, 512, 2, 61)
    ref = v[0].astype(np.float32)
    s = DisplacementSolver(512, 512, 2, max_batch=2)
    s.set_reference(ref, np.array([0.5, 0.5]), None)
    out = s.run(torch.from_numpy(np.stack([ref, ref])).cuda(), out_dtype=torch.float64).cpu().numpy()
    assert np.all(out == 0.0)
    s.close()


def test_512x512x2_batch_order_does_not_matter(gpu):
    import torch
    from pyflowreg_amd.optical_flow import DisplacementSolver
    v = _frames(4, 512, 512, 2, 62).astype(np.float32)
    ref, movs = v[0], v[1:]
    s = DisplacementSolver(512, 512, 2, max_batch=4)
    s.set_reference(ref, np.array([0.5, 0.5]), None)
    a = s.run(torch.from_numpy(movs).cuda(), out_dtype=torch.float64).cpu().numpy()
    perm = [2, 0, 3, 1]
    b = s.run(torch.from_numpy(movs[perm]).cuda(), out_dtype=torch.float64).cpu().numpy()
    np.testing.assert_array_equal(b, a[perm])
    s.close()


def test_796x512x1_bit_exact(gpu, oracle):
    """796 rows: the sweeps run in 1024-lane work-groups (at most 2 sweeps
    per pass) on the finest levels."""
    from pyflowreg_amd.optical_flow import get_displacement
    v = _frames(1, 796, 512, 1, 63)
    ref, movs = _prep(oracle, v)
    a = get_displacement(ref, movs[0])
    b = oracle.get_displacement(ref, movs[0])
    d = np.abs(a - b)
    assert d.max() == 0.0, f"max |dflow| {d.max():.3e}"
