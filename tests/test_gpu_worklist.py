"""Uneven pairs through one work list (frh_worklist_*: one launch per stage
and level step for all frame sizes, channel counts and pyramid depths) vs
per-pair CPU-oracle runs and vs the per-size grouped launches.  Bar:
bit-exact."""

import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, file=sys.stderr, flush=True)


def _pair(oracle, H, W, C, seed):
    from pyflowreg_amd import synth
    frames, _ = synth.video(2, H, W, n_channels=C, max_disp=2.0, sigma_d=max(H, W) / 6, seed=seed)
    f = frames[0].numpy()
    m = frames[1].numpy()
    ref = oracle.preprocess(f.astype(np.float64)).astype(np.float32)
    mov = oracle.preprocess(m, f.astype(np.float64)).astype(np.float32)
    return ref, mov


# sizes with different pyramid depths and channel counts; 600 rows puts the
# top level into the 1024-lane form (on the work list's second stream)
SIZES = [(32, 32, 1), (40, 48, 2), (600, 40, 1), (32, 32, 1), (56, 40, 2), (96, 64, 1)]


@pytest.fixture(scope="module")
def uneven(oracle):
    pairs = [_pair(oracle, H, W, C, 500 + k) for k, (H, W, C) in enumerate(SIZES)]
    return [p[0] for p in pairs], [p[1] for p in pairs]


def test_worklist_bit_exact_vs_oracle(gpu, oracle, uneven):
    fixed, moving = uneven
    _say("worklist: GPU")
    got = gpu.get_displacements(fixed, moving, worklist=True)
    for k, (f, m) in enumerate(zip(fixed, moving)):
        _say(f"worklist: oracle pair {k} {f.shape}")
        np.testing.assert_array_equal(got[k], oracle.get_displacement(f, m), err_msg=f"pair {k} {f.shape}")


def test_worklist_equals_grouped_launches(gpu, uneven):
    fixed, moving = uneven
    a = gpu.get_displacements(fixed, moving, worklist=True)
    b = gpu.get_displacements(fixed, moving, worklist=False)
    for k in range(len(a)):
        np.testing.assert_array_equal(a[k], b[k], err_msg=f"pair {k}")


def test_worklist_min_level_and_gray(gpu, oracle, uneven):
    """min_level > 0 (coarse flow, then the final cubic resize per solver)
    and the gray-value data term through the work list."""
    fixed, moving = uneven
    sel = [0, 1, 4, 5]
    f = [fixed[k] for k in sel]
    m = [moving[k] for k in sel]
    got = gpu.get_displacements(f, m, min_level=2, const_assumption="gray", worklist=True)
    for k in range(len(sel)):
        exp = oracle.get_displacement(f[k], m[k], min_level=2, const_assumption="gray")
        np.testing.assert_array_equal(got[k], exp, err_msg=f"pair {sel[k]}")


def test_worklist_reuse_smaller_batch(gpu, oracle):
    """One work list run twice, the second time with fewer pairs per solver:
    both runs equal per-pair oracle results (descriptors rebuilt per run)."""
    import torch
    from pyflowreg_amd.optical_flow import DisplacementSolver, WorkList
    specs = [(48, 40, 1, 3), (72, 56, 2, 2)]
    solvers, data = [], []
    for k, (H, W, C, T) in enumerate(specs):
        s = DisplacementSolver(H, W, C, max_batch=T)
        s.set_reference(None, None)
        pairs = [_pair(oracle, H, W, C, 700 + 10 * k + t) for t in range(T)]
        fx = np.stack([p[0].reshape(H, W, C) for p in pairs])
        mv = np.stack([p[1].reshape(H, W, C) for p in pairs])
        solvers.append(s)
        data.append((fx, mv))
    wl = WorkList(solvers)
    try:
        for rep, take in enumerate([None, 1]):
            fx = [d[0][:take] for d in data]
            mv = [d[1][:take] for d in data]
            outs = wl.run(fx, mv, out_dtype=torch.float64)
            for k, o in enumerate(outs):
                o = o.cpu().numpy()
                assert o.shape[0] == fx[k].shape[0]
                for t in range(o.shape[0]):
                    exp = oracle.get_displacement(fx[k][t], mv[k][t])
                    np.testing.assert_array_equal(o[t], exp, err_msg=f"run {rep} solver {k} pair {t}")
    finally:
        wl.close()
        for s in solvers:
            s.close()


def test_worklist_rejects_nonlinear_smoothness(gpu):
    from pyflowreg_amd.optical_flow import DisplacementSolver, WorkList
    s = DisplacementSolver(32, 32, 1, max_batch=1, a_smooth=1.5)
    try:
        with pytest.raises(ValueError):
            WorkList([s])
    finally:
        s.close()


def test_worklist_float32_output(gpu, uneven):
    """float32 flows are the float64 flows rounded (the median's output cast)."""
    import torch
    from pyflowreg_amd.optical_flow import DisplacementSolver, WorkList
    fixed, moving = uneven
    sel = [1, 2]
    solvers = []
    for k in sel:
        H, W, C = SIZES[k]
        s = DisplacementSolver(H, W, C, max_batch=1)
        s.set_reference(None, None)
        solvers.append(s)
    wl = WorkList(solvers)
    try:
        fx = [fixed[k].reshape(1, *SIZES[k]) for k in sel]
        mv = [moving[k].reshape(1, *SIZES[k]) for k in sel]
        o64 = [o.cpu().numpy() for o in wl.run(fx, mv, out_dtype=torch.float64)]
        o32 = [o.cpu().numpy() for o in wl.run(fx, mv, out_dtype=torch.float32)]
        for a, b in zip(o64, o32):
            assert b.dtype == np.float32
            np.testing.assert_array_equal(a.astype(np.float32), b)
    finally:
        wl.close()
        for s in solvers:
            s.close()
