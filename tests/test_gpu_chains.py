"""Chains of work-groups in the batched solver (a pass's sweeps, and on
levels of more than 512 rows its 512-row bands, as one work-group each per
frame, handing results on under progress flags): bit-identical to the
one-work-group-per-frame passes (frh_solver_set_chains(s, 0)) and to the
oracle, for a batch whose top level has 600 rows (2 bands: band chains) and
whose next levels have 256+ anti-diagonals (sweep chains)."""

import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, file=sys.stderr, flush=True)


def test_chains_equal_single_work_group_passes(gpu, oracle):
    import torch
    from pyflowreg_amd import synth
    from pyflowreg_amd.optical_flow import DisplacementSolver
    T, H, W, C = 3, 600, 200, 2
    frames, _ = synth.video(T + 1, H, W, n_channels=C, max_disp=3.0, sigma_d=40.0, seed=5301)
    f = frames[0].numpy()
    ref = oracle.preprocess(f.astype(np.float64)).astype(np.float32)
    movs = oracle.preprocess(frames[1:].numpy(), f.astype(np.float64)).astype(np.float32)
    s = DisplacementSolver(H, W, C, max_batch=T)
    s.set_reference(ref, np.array([0.5, 0.5]), None)
    x = torch.from_numpy(movs).cuda()
    _say("600x200x2: chains")
    a = s.run(x, out_dtype=torch.float64).cpu().numpy()
    s.set_chains(False)
    _say("600x200x2: one work-group per frame")
    b = s.run(x, out_dtype=torch.float64).cpu().numpy()
    s.close()
    np.testing.assert_array_equal(a, b)
    _say("600x200x2: oracle frame 1")
    np.testing.assert_array_equal(a[1], oracle.get_displacement(ref, movs[1], weight=[0.5, 0.5]))


@pytest.mark.parametrize("T", [30, 50])
def test_multi_sweep_chain_stages_equal_pipelined_passes(gpu, oracle, T):
    """Chain stages of several pipelined sweeps: on a 256-CU device 30 frames
    run the 5-sweep lag blocks of the 200 x 100 top level (299 anti-diagonals)
    as three stages of two sweeps (the last one sweep short) and 50 frames
    as two stages of three sweeps (the fewest sweeps per stage whose stages
    fill at most half the compute units); bit-identical to the pipelined
    passes and, for one frame, to the oracle."""
    import torch
    from pyflowreg_amd import synth
    from pyflowreg_amd.optical_flow import DisplacementSolver
    H, W, C = 200, 100, 1
    frames, _ = synth.video(T + 1, H, W, n_channels=C, max_disp=3.0, sigma_d=30.0, seed=5302 + T)
    f = frames[0].numpy()
    ref = oracle.preprocess(f.astype(np.float64)).astype(np.float32)
    movs = oracle.preprocess(frames[1:].numpy(), f.astype(np.float64)).astype(np.float32)
    s = DisplacementSolver(H, W, C, max_batch=T)
    s.set_reference(ref, np.array([1.0]), None)
    x = torch.from_numpy(movs).cuda()
    _say(f"200x100x1, {T} frames: chains")
    a = s.run(x, out_dtype=torch.float64).cpu().numpy()
    s.set_chains(False)
    _say(f"200x100x1, {T} frames: pipelined passes")
    b = s.run(x, out_dtype=torch.float64).cpu().numpy()
    s.close()
    np.testing.assert_array_equal(a, b)
    k = T - 1
    _say(f"200x100x1, {T} frames: oracle frame {k}")
    np.testing.assert_array_equal(a[k], oracle.get_displacement(ref, movs[k], weight=[1.0]))
