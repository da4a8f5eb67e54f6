"""Pyramid loop for ``get_displacement(..., level_solver_backend=f)``.

When a caller supplies its own level solver (pyflowreg's tests do this with
recording solvers, tests/core/test_optical_flow.py:22-33,449-468), the level
loop must hand that callable the reference's positional arguments as numpy
arrays.  This module runs core/optical_flow.py:476-617 level by level: the
resize, warp, motion tensor and median stages run on the GPU, the callable
runs wherever it runs.
"""

from __future__ import annotations

import numpy as np

from . import optical_flow as of


def solve_with_level_solver(fixed, moving, alpha, update_lag, iterations, min_level, levels, eta,
                            a_smooth, a_data_arr, uv, weight, level_solver_backend):
    fixed = np.asarray(fixed, dtype=np.float64)
    moving = np.asarray(moving, dtype=np.float64)
    if fixed.ndim == 2:
        fixed = fixed[:, :, None]
        moving = moving[:, :, None]
    m, n, C = fixed.shape
    weight = of.expand_weight(weight, m, n, C)
    plan, min_level = of.level_plan(m, n, eta, levels, min_level)
    top = plan[0][0]
    if uv is not None:
        u_init, v_init = np.asarray(uv)[:, :, 0], np.asarray(uv)[:, :, 1]
    else:
        u_init = np.zeros((m, n))
        v_init = np.zeros((m, n))
    u = v = None
    for i, level_size in plan:
        f1 = of.imresize_fused_gauss_cubic(fixed, level_size)
        f2 = of.imresize_fused_gauss_cubic(moving, level_size)
        h_row = float(m) / f1.shape[0]
        h_col = float(n) / f1.shape[1]
        if i == top:
            u = np.pad(of.imresize_fused_gauss_cubic(u_init, level_size), 1, mode="edge")
            v = np.pad(of.imresize_fused_gauss_cubic(v_init, level_size), 1, mode="edge")
            tmp = f2.astype(np.float32)
        else:
            u = np.pad(of.imresize_fused_gauss_cubic(u[1:-1, 1:-1], level_size), 1, mode="edge")
            v = np.pad(of.imresize_fused_gauss_cubic(v[1:-1, 1:-1], level_size), 1, mode="edge")
            tmp = of.imregister_wrapper(f2.astype(np.float32), u[1:-1, 1:-1] / h_col,
                                        v[1:-1, 1:-1] / h_row, f1.astype(np.float32))
            if tmp.ndim == 2:
                tmp = tmp[:, :, None]
        lh, lw = level_size
        J = np.zeros((6, lh + 2, lw + 2, C))
        for ch in range(C):
            Jc = of.get_motion_tensor_gc(f1[:, :, ch], tmp[:, :, ch], h_row, h_col)
            for t in range(6):
                J[t, :, :, ch] = Jc[t]
        wl = of.imresize_fused_gauss_cubic(weight, (lh, lw))
        weight_level = np.zeros((lh + 2, lw + 2, C))
        weight_level[1:-1, 1:-1] = wl.reshape(lh, lw, C)
        alpha_scaling = 1 if i == min_level else eta ** (-0.5 * i)
        alpha_tmp = [alpha_scaling * alpha[j] for j in range(len(alpha))]
        du, dv = level_solver_backend(
            np.ascontiguousarray(J[0]), np.ascontiguousarray(J[1]), np.ascontiguousarray(J[2]),
            np.ascontiguousarray(J[3]), np.ascontiguousarray(J[4]), np.ascontiguousarray(J[5]),
            np.ascontiguousarray(weight_level), np.ascontiguousarray(u), np.ascontiguousarray(v),
            alpha_tmp, iterations, update_lag, 0, a_data_arr, a_smooth, h_col, h_row, None)
        du = np.array(du, dtype=np.float64)
        dv = np.array(dv, dtype=np.float64)
        if min(level_size) > 5:
            du[1:-1, 1:-1] = of.median5_mirror(du[1:-1, 1:-1])
            dv[1:-1, 1:-1] = of.median5_mirror(dv[1:-1, 1:-1])
        u = u + du
        v = v + dv
    w = np.zeros((u.shape[0] - 2, u.shape[1] - 2, 2))
    w[:, :, 0] = u[1:-1, 1:-1]
    w[:, :, 1] = v[1:-1, 1:-1]
    if min_level > 0:
        w = of.cv_resize_cubic(w, m, n)
    return w
