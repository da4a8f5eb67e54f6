"""SOR wavefront pass limits and their autotuner: the flows do not depend on
the sweeps per pass, the tuned plan is cached and re-applied, and limits
outside a level's range are rejected."""

import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


def _setup(H, W, C, T, seed, **kw):
    import pyflowreg_amd as fr
    from pyflowreg_amd import synth
    v, _ = synth.video(T + 1, H, W, C, seed=seed, max_disp=2.0)
    s = fr.DisplacementSolver(H, W, C, max_batch=T, **kw)
    s.set_reference(v[0].cuda(), [0.5] * C if C > 1 else None)
    return s, v[1:].cuda()


def test_pass_limits_do_not_change_flows(gpu):
    from pyflowreg_amd import autotune as at
    s, mv = _setup(70, 52, 2, 3, 71, iterations=20)
    ref = s.run(mv)
    info = at.level_info(s)
    assert all(p == mx for (_, _, p, mx, _) in info)
    for lim in (1, 3):
        at.apply_plan(s, [min(lim, mx) for (_, _, _, mx, _) in info])
        assert torch.equal(s.run(mv), ref), lim


def test_pass_limit_bounds(gpu):
    from pyflowreg_amd import _lib
    s, _ = _setup(40, 36, 1, 1, 72)
    L = _lib.lib()
    assert L.frh_solver_set_pass_limit(s._h, 0, 0) != 0
    assert L.frh_solver_set_pass_limit(s._h, 0, 6) != 0
    assert L.frh_solver_set_pass_limit(s._h, 99, 2) != 0
    assert L.frh_solver_set_pass_limit(s._h, 0, 2) == 0


def test_autotune_plan_is_cached_and_reapplied(gpu, tmp_path):
    from pyflowreg_amd import autotune as at
    s, mv = _setup(64, 56, 2, 4, 73, iterations=10)
    ref = s.run(mv)
    cache = str(tmp_path / "tune.json")
    res = at.autotune(s, mv, cache=cache, candidates=(2, 5), repeats=1)
    assert not res["cached"] and len(res["limits"]) == len(s.plan.levels)
    assert set(res["limits"]) <= {2, 5}
    assert [li[2] for li in at.level_info(s)] == res["limits"]
    assert torch.equal(s.run(mv), ref)
    s2, _ = _setup(64, 56, 2, 4, 73, iterations=10)
    res2 = at.autotune(s2, mv, cache=cache)
    assert res2["cached"] and res2["limits"] == res["limits"]
    assert [li[2] for li in at.level_info(s2)] == res["limits"]
