"""Host-driven pyramid loops over the GPU stage entry points.

Two callers need the reference's level loop with numpy arrays between the
stages rather than the fused device-resident solver:

* ``get_displacement(..., level_solver_backend=f)``: the caller's level
  solver receives the reference's positional arguments as numpy arrays
  (pyflowreg's tests do this with recording solvers,
  tests/core/test_optical_flow.py:22-33,449-468);
* GNC schedules (``gnc_schedule``): core/optical_flow.py:620-770 warps every
  level (the top level too), relinearises ``warping_steps`` times per level
  and takes the median of u itself after each solve.

``solve_stage`` restates core/optical_flow.py:476-617 and
``solve_stage_gnc`` :620-770.  The resize, warp, motion tensor, level solver
(unless the caller supplies one) and median run on the GPU.  Arrays keep
numpy's dtypes exactly as the reference's do, because the motion tensor's
arithmetic follows them: level images are float64, the top level's "warped"
image is the resized level itself (float64), the warp returns float32, and a
float32 initial flow stays float32 until the first update.
"""

from __future__ import annotations

import numpy as np

from . import optical_flow as of


def _level_tensor(tensor_fn, f1, tmp, h_row, h_col, C):
    lh, lw = f1.shape[:2]
    J = np.zeros((6, lh + 2, lw + 2, C))
    for ch in range(C):
        Jc = tensor_fn(f1[:, :, ch], tmp[:, :, ch], h_row, h_col)
        for t in range(6):
            J[t, :, :, ch] = Jc[t]
    return J


def _weight_level(weight, lh, lw, C):
    wl = of.imresize_fused_gauss_cubic(weight, (lh, lw))
    out = np.zeros((lh + 2, lw + 2, C))
    out[1:-1, 1:-1] = wl.reshape(lh, lw, C)
    return out


def _call_solver(solver_func, J, weight_level, u, v, alpha_tmp, iterations, update_lag,
                 a_data_arr, a_smooth, h_col, h_row, gnc_beta):
    du, dv = solver_func(
        np.ascontiguousarray(J[0]), np.ascontiguousarray(J[1]), np.ascontiguousarray(J[2]),
        np.ascontiguousarray(J[3]), np.ascontiguousarray(J[4]), np.ascontiguousarray(J[5]),
        np.ascontiguousarray(weight_level), np.ascontiguousarray(u), np.ascontiguousarray(v),
        alpha_tmp, iterations, update_lag, 0, a_data_arr, a_smooth, h_col, h_row, gnc_beta)
    return np.array(du, dtype=np.float64), np.array(dv, dtype=np.float64)


def _prepare(fixed, moving, weight, uv):
    fixed = np.asarray(fixed, dtype=np.float64)
    moving = np.asarray(moving, dtype=np.float64)
    if fixed.ndim == 2:
        fixed = fixed[:, :, None]
        moving = moving[:, :, None]
    m, n, C = fixed.shape
    weight = of.expand_weight(weight, m, n, C)
    if uv is not None:
        uv = np.asarray(uv)
        u_init, v_init = uv[:, :, 0], uv[:, :, 1]
    else:
        u_init = np.zeros((m, n))
        v_init = np.zeros((m, n))
    return fixed, moving, weight, u_init, v_init


def _finish(u, v, m, n, min_level):
    w = np.zeros((u.shape[0] - 2, u.shape[1] - 2, 2))
    w[:, :, 0] = u[1:-1, 1:-1]
    w[:, :, 1] = v[1:-1, 1:-1]
    if min_level > 0:
        w = of.cv_resize_cubic(w, m, n)
    return w


def solve_stage(fixed, moving, alpha, update_lag, iterations, min_level, levels, eta, a_smooth,
                a_data_arr, uv, weight, solver_func, tensor_fn):
    """core/optical_flow.py:476-617 with ``solver_func`` as level solver."""
    fixed, moving, weight, u_init, v_init = _prepare(fixed, moving, weight, uv)
    m, n, C = fixed.shape
    plan, min_level = of.level_plan(m, n, eta, levels, min_level)
    top = plan[0][0]
    u = v = None
    for i, level_size in plan:
        f1 = of.imresize_fused_gauss_cubic(fixed, level_size)
        f2 = of.imresize_fused_gauss_cubic(moving, level_size)
        h_row = float(m) / f1.shape[0]
        h_col = float(n) / f1.shape[1]
        if i == top:
            u = np.pad(of.imresize_fused_gauss_cubic(u_init, level_size), 1, mode="edge")
            v = np.pad(of.imresize_fused_gauss_cubic(v_init, level_size), 1, mode="edge")
            tmp = f2.copy()
        else:
            u = np.pad(of.imresize_fused_gauss_cubic(u[1:-1, 1:-1], level_size), 1, mode="edge")
            v = np.pad(of.imresize_fused_gauss_cubic(v[1:-1, 1:-1], level_size), 1, mode="edge")
            tmp = of.imregister_wrapper(f2, u[1:-1, 1:-1] / h_col, v[1:-1, 1:-1] / h_row, f1)
            if tmp.ndim == 2:
                tmp = tmp[:, :, None]
        lh, lw = level_size
        J = _level_tensor(tensor_fn, f1, tmp, h_row, h_col, C)
        weight_level = _weight_level(weight, lh, lw, C)
        alpha_scaling = 1 if i == min_level else eta ** (-0.5 * i)
        alpha_tmp = [alpha_scaling * alpha[j] for j in range(len(alpha))]
        du, dv = _call_solver(solver_func, J, weight_level, u, v, alpha_tmp, iterations, update_lag,
                              a_data_arr, a_smooth, h_col, h_row, None)
        if min(level_size) > 5:
            du[1:-1, 1:-1] = of.median5_mirror(du[1:-1, 1:-1])
            dv[1:-1, 1:-1] = of.median5_mirror(dv[1:-1, 1:-1])
        u = u + du
        v = v + dv
    return _finish(u, v, m, n, min_level)


def solve_stage_gnc(fixed, moving, alpha, update_lag, iterations, min_level, levels, eta,
                    a_smooth, a_data_arr, uv, weight, solver_func, tensor_fn, gnc_beta,
                    warping_steps):
    """core/optical_flow.py:620-770: one GNC stage."""
    fixed, moving, weight, u_init, v_init = _prepare(fixed, moving, weight, uv)
    m, n, C = fixed.shape
    plan, min_level = of.level_plan(m, n, eta, levels, min_level)
    top = plan[0][0]
    u = v = None
    for i, level_size in plan:
        f1 = of.imresize_fused_gauss_cubic(fixed, level_size)
        f2 = of.imresize_fused_gauss_cubic(moving, level_size)
        h_row = float(m) / f1.shape[0]
        h_col = float(n) / f1.shape[1]
        if i == top:
            u = np.pad(of.imresize_fused_gauss_cubic(u_init, level_size), 1, mode="edge")
            v = np.pad(of.imresize_fused_gauss_cubic(v_init, level_size), 1, mode="edge")
        else:
            u = np.pad(of.imresize_fused_gauss_cubic(u[1:-1, 1:-1], level_size), 1, mode="edge")
            v = np.pad(of.imresize_fused_gauss_cubic(v[1:-1, 1:-1], level_size), 1, mode="edge")
        lh, lw = level_size
        weight_level = _weight_level(weight, lh, lw, C)
        alpha_scaling = 1 if i == min_level else eta ** (-0.5 * i)
        alpha_tmp = [alpha_scaling * alpha[j] for j in range(len(alpha))]
        for _ in range(warping_steps):
            # u / h_col keeps u's dtype (float32 for a float32 initial flow)
            tmp = of.imregister_wrapper(f2, u[1:-1, 1:-1] / h_col, v[1:-1, 1:-1] / h_row, f1)
            if tmp.ndim == 2:
                tmp = tmp[:, :, None]
            J = _level_tensor(tensor_fn, f1, tmp, h_row, h_col, C)
            du, dv = _call_solver(solver_func, J, weight_level, u, v, alpha_tmp, iterations,
                                  update_lag, a_data_arr, a_smooth, h_col, h_row, gnc_beta)
            u = u + du
            v = v + dv
            if min(level_size) > 5:
                u = np.pad(of.median5_mirror(u[1:-1, 1:-1]), 1, mode="edge")
                v = np.pad(of.median5_mirror(v[1:-1, 1:-1]), 1, mode="edge")
    return _finish(u, v, m, n, min_level)
