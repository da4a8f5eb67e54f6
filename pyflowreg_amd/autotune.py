"""Autotuning of the SOR wavefront's sweeps per pass, per pyramid level.

Between coefficient refreshes a level's sweeps run as wavefront passes of at
most ``pass_limit`` sweeps (``frh_solver_set_pass_limit``): more sweeps per
pass read and write the state fewer times but hold more registers per lane
and lengthen each pass by two diagonals per extra sweep.  The flows do not
depend on the choice, only the time does, so the tuner times the solver on
a sample batch with every candidate limit on every level at once (one
profiled run per candidate, per-level sweep + assembly time from HIP
events), keeps each level's fastest, and caches the plan as JSON keyed by
device name, frame geometry, batch size and solver schedule.
"""

from __future__ import annotations

import json
import os
from typing import Dict, Optional

import numpy as np
import torch

from ._lib import check, lib


def _cache_path(path: Optional[str] = None) -> str:
    if path:
        return path
    base = os.environ.get("FLOWREG_HIP_CACHE", os.path.join(os.path.expanduser("~"), ".cache", "flowreg_hip"))
    return os.path.join(base, "autotune.json")


def plan_key(solver, batch: int) -> str:
    dev = torch.cuda.get_device_name(torch.cuda.current_device())
    return (f"{dev}|{solver.H}x{solver.W}x{solver.C}|T{batch}|it{solver.iterations}|lag{solver.update_lag}|"
            f"levels{len(solver.plan.levels)}")


def level_info(solver):
    """[(h, w, pass_limit, pass_limit_max, last_ms)] per level, coarsest first."""
    out = []
    for lvl in range(len(solver.plan.levels)):
        v = [np.zeros(1, np.int32) for _ in range(4)] + [np.zeros(1, np.float64)]
        check(lib().frh_solver_level_info(solver._h, lvl, *[a.ctypes.data for a in v]), "frh_solver_level_info")
        out.append((int(v[0][0]), int(v[1][0]), int(v[2][0]), int(v[3][0]), float(v[4][0])))
    return out


def apply_plan(solver, limits) -> None:
    for lvl, p in enumerate(limits):
        check(lib().frh_solver_set_pass_limit(solver._h, lvl, int(p)), "frh_solver_set_pass_limit")


def tune(solver, moving, candidates=(1, 2, 3, 4, 5), repeats: int = 2) -> Dict:
    """Times ``solver.run(moving)`` (reference already set) with each
    candidate limit on every level (clamped to the level's maximum), and
    applies and returns the fastest per level:
    {"limits": [...], "level_ms": {candidate: [...]}, "levels": [[h, w], ...]}."""
    info = level_info(solver)
    n = len(info)
    times = {}
    for p in candidates:
        limits = [min(p, mx) for (_, _, _, mx, _) in info]
        apply_plan(solver, limits)
        solver.run(moving, out_dtype=torch.float32)  # warm-up
        acc = np.zeros(n)
        for _ in range(repeats):
            solver.run(moving, out_dtype=torch.float32, profile=True)
            acc += np.array([li[4] for li in level_info(solver)])
        times[p] = (acc / repeats).tolist()
    best = []
    for lvl in range(n):
        mx = info[lvl][3]
        cand = [(times[p][lvl], p) for p in candidates if p <= mx]
        best.append(min(cand)[1])
    apply_plan(solver, best)
    return {"limits": best, "level_ms": {str(p): t for p, t in times.items()},
            "levels": [[h, w] for (h, w, _, _, _) in info]}


def autotune(solver, moving, cache: Optional[str] = None, force: bool = False, **kw) -> Dict:
    """Applies the cached plan for this solver and batch size, or tunes on
    ``moving`` and stores the plan in the JSON cache."""
    path = _cache_path(cache)
    T = moving.shape[0] if moving.ndim == 4 else 1
    key = plan_key(solver, T)
    db = {}
    if os.path.exists(path):
        with open(path) as f:
            db = json.load(f)
    if not force and key in db and len(db[key]["limits"]) == len(solver.plan.levels):
        apply_plan(solver, db[key]["limits"])
        return dict(db[key], cached=True)
    res = tune(solver, moving, **kw)
    db[key] = res
    os.makedirs(os.path.dirname(path), exist_ok=True)
    tmp = path + ".tmp"
    with open(tmp, "w") as f:
        json.dump(db, f, indent=1)
    os.replace(tmp, path)
    return dict(res, cached=False)
