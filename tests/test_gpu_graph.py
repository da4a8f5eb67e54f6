"""Graph replay of the solver's level loop (frh_solver_set_graph): every
replay equals the eager launches bit for bit, across batch-size and output
type changes (re-capture), with eager work between replays, and in the
independent-pairs mode."""

import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


def _frames(T, H, W, C, seed):
    from pyflowreg_amd import synth
    vid, _ = synth.video(T + 1, H, W, C, seed=seed, max_disp=2.0)
    return vid[0].cuda(), vid[1:].cuda()


def _solvers(H, W, C, fixed, **kw):
    import pyflowreg_amd as fr
    eager = fr.DisplacementSolver(H, W, C, max_batch=4, **kw)
    graph = fr.DisplacementSolver(H, W, C, max_batch=4, graph=True, **kw)
    for s in (eager, graph):
        s.set_reference(fixed, [0.5] * C if C > 1 else None)
    return eager, graph


def test_graph_replays_equal_eager_runs(gpu):
    H, W, C = 48, 40, 2
    fixed, moving = _frames(4, H, W, C, seed=51)
    eager, graph = _solvers(H, W, C, fixed, iterations=20)
    cases = [(moving[:3], torch.float64), (moving[:3], torch.float64), (moving[1:3], torch.float64),
             (moving[:3], torch.float32), (moving[:4], torch.float64), (moving[:3], torch.float64)]
    for k, (mv, dt) in enumerate(cases):
        ref = eager.run(mv, out_dtype=dt)
        junk = torch.randn(1 << 20, device="cuda").cumsum(0)  # eager work between replays
        out = graph.run(mv, out_dtype=dt)
        torch.cuda.synchronize()
        assert torch.equal(out, ref), k
        del junk


def test_graph_follows_reference_updates(gpu):
    """The graph reads the solver's reference buffers: a new reference set
    after the capture is used by the next replay."""
    H, W, C = 40, 44, 1
    f1, moving = _frames(2, H, W, C, seed=52)
    f2, _ = _frames(1, H, W, C, seed=53)
    eager, graph = _solvers(H, W, C, f1, iterations=10)
    assert torch.equal(graph.run(moving), eager.run(moving))
    for s in (eager, graph):
        s.set_reference(f2)
    out = graph.run(moving)
    assert torch.equal(out, eager.run(moving))


def test_graph_pairs_equal_eager(gpu):
    H, W, C = 36, 52, 2
    fa, ma = _frames(3, H, W, C, seed=54)
    _, mb = _frames(3, H, W, C, seed=55)
    import pyflowreg_amd as fr
    eager = fr.DisplacementSolver(H, W, C, max_batch=3, iterations=15)
    graph = fr.DisplacementSolver(H, W, C, max_batch=3, iterations=15, graph=True)
    for s in (eager, graph):
        s.set_reference(None, [0.5, 0.5])
    fixed = torch.stack([fa, mb[0], mb[1]])
    for _ in range(3):
        assert torch.equal(graph.run_pairs(fixed, ma), eager.run_pairs(fixed, ma))
