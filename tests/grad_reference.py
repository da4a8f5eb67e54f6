"""Pure-torch (CPU, float64) restatement of the differentiable registration
map, differentiated by torch.autograd: the reference gradient for the HIP
backward kernels.  Test infrastructure only.

Written with vectorised torch ops, independently of the HIP kernels:
* motion tensor: core/optical_flow.py:73-160 on replicate-padded planes,
  fx / ft re-padded before their second differences (reg = 1 / (a^2 + b^2 +
  1e-6), the reference's sqrt-then-square folded);
* coefficients: core/level_solver.py:229-255,317-353 (a_smooth = 1);
* SOR: core/level_solver.py:257-364 in column-major lexicographic order,
  evaluated one anti-diagonal at a time (a cell reads its left and upper
  neighbours' new values and its right and lower neighbours' old values);
* median: scipy.ndimage.median_filter(size=5, mode="mirror") as a sort of
  reflect-padded 5x5 windows;
* warp: core/warping.py:160-234 with a continuous map;
* resize: util/resize_util.py:194-254 as dense matrices of the oracle's tap
  tables.
"""

from __future__ import annotations

import numpy as np
import torch
import torch.nn.functional as F

OMEGA = 1.95


def _rep(x):
    """replicate-pad the last two dims by 1 (x: (..., h, w))."""
    shp = x.shape
    y = F.pad(x.reshape(-1, 1, shp[-2], shp[-1]), (1, 1, 1, 1), mode="replicate")
    return y.reshape(*shp[:-2], shp[-2] + 2, shp[-1] + 2)


def motion_tensor_gc(f1, f2, hy, hx):
    """f1 (C, h, w), f2 (T, C, h, w) -> J (T, 6, C, h, w)."""
    F1, F2 = _rep(f1), _rep(f2)

    def ddx(P):
        return (P[..., 1:-1, 2:] - P[..., 1:-1, :-2]) / (2.0 * hx)

    def dxx(P, h):
        return (P[..., 1:-1, :-2] - 2.0 * P[..., 1:-1, 1:-1] + P[..., 1:-1, 2:]) / (h * h)

    def dyy(P, h):
        return (P[..., :-2, 1:-1] - 2.0 * P[..., 1:-1, 1:-1] + P[..., 2:, 1:-1]) / (h * h)

    fx = 0.5 * (ddx(F1) + ddx(F2))
    ft = f2 - f1
    FX, FT = _rep(fx), _rep(ft)
    fxy = (FX[..., 2:, 1:-1] - FX[..., :-2, 1:-1]) / (2.0 * hy)
    fxt = (FT[..., 1:-1, 2:] - FT[..., 1:-1, :-2]) / (2.0 * hx)
    fyt = (FT[..., 2:, 1:-1] - FT[..., :-2, 1:-1]) / (2.0 * hy)
    fxx = 0.5 * (dxx(F1, hx) + dxx(F2, hx))
    fyy = 0.5 * (dyy(F1, hy) + dyy(F2, hy))
    rx = 1.0 / (fxx * fxx + fxy * fxy + 1e-6)
    ry = 1.0 / (fxy * fxy + fyy * fyy + 1e-6)
    J = [rx * fxx * fxx + ry * fxy * fxy,
         rx * fxy * fxy + ry * fyy * fyy,
         rx * fxt * fxt + ry * fyt * fyt,
         rx * fxx * fxy + ry * fxy * fyy,
         rx * fxx * fxt + ry * fxy * fyt,
         rx * fxy * fxt + ry * fyy * fyt]
    return torch.stack(J, dim=1)


def _nb_diffs(f):
    """(sum over in-level left/right neighbours of f_nb - f_c, same up/down)
    for f (..., h, w) with edge padding."""
    P = _rep(f)
    sx = (P[..., 1:-1, 2:] - f) + (P[..., 1:-1, :-2] - f)
    sy = (P[..., 2:, 1:-1] - f) + (P[..., :-2, 1:-1] - f)
    return sx, sy


def coefficients(x, u0, J, W, ax, ay, a_data, psi_one):
    """Lag-block coefficients (T, 6, h, w) = A12, Bu, Bv, omega/Du,
    omega/Dv, b.  x, u0 (T, 2, h, w); J (T, 6, C, h, w); W (C, h, w);
    a_data (C,)."""
    T, _, C, h, w = J.shape
    du, dv = x[:, 0:1], x[:, 1:2]  # (T, 1, h, w), broadcast over channels
    j11, j22, j33, j12, j13, j23 = (J[:, e] for e in range(6))
    if psi_one:
        psi = torch.ones_like(j11)
    else:
        E = j11 * du * du + j22 * dv * dv + 2.0 * (j23 * dv) + 2.0 * (j12 * du * dv) + 2.0 * (j13 * du) + j33
        a = a_data.reshape(1, C, 1, 1)
        psi = a * (torch.clamp(E, min=0.0) + 1e-5) ** (a - 1.0)
    wp = W.unsqueeze(0) * psi
    sxu, syu = _nb_diffs(u0[:, 0])
    sxv, syv = _nb_diffs(u0[:, 1])
    ws = 2.0 * ax + 2.0 * ay
    Du = ws + (wp * j11).sum(1)
    Dv = ws + (wp * j22).sum(1)
    A12 = (wp * j12).sum(1)
    Bu = ax * sxu + ay * syu - (wp * j13).sum(1)
    Bv = ax * sxv + ay * syv - (wp * j23).sum(1)
    cols = torch.arange(w, dtype=torch.float64)
    rows = torch.arange(h, dtype=torch.float64)[:, None]
    nlr = ((cols == 0).double() + (cols == w - 1).double()).expand(h, w)
    nud = ((rows == 0).double() + (rows == h - 1).double()).expand(h, w)
    b = (nlr * ax + nud * ay).expand(T, h, w)
    return torch.stack([A12, Bu, Bv, OMEGA / Du, OMEGA / Dv, b], dim=1)


_DIAG_CACHE = {}


def _diagonals(h, w):
    key = (h, w)
    if key not in _DIAG_CACHE:
        diags, order = [], []
        for d in range(h + w - 1):
            r0, r1 = max(0, d - w + 1), min(h - 1, d)
            rs = torch.arange(r0, r1 + 1)
            diags.append((r0, rs, d - rs))
            order.append(rs * w + (d - rs))
        order = torch.cat(order)
        inv = torch.empty_like(order)
        inv[order] = torch.arange(order.numel())
        _DIAG_CACHE[key] = (diags, inv)
    return _DIAG_CACHE[key]


def sor_sweep(coef, x, ax, ay):
    """One column-major lexicographic SOR sweep x (T, 2, h, w) -> new."""
    T, _, h, w = x.shape
    diags, inv = _diagonals(h, w)
    oup = F.pad(x[:, 0], (0, 1, 0, 1))
    ovp = F.pad(x[:, 1], (0, 1, 0, 1))
    outs_u, outs_v = [], []
    prev = None
    for r0, rs, cs in diags:
        A12, Bu, Bv, wu, wv, b = (coef[:, e, rs, cs] for e in range(6))
        ou, ov = x[:, 0, rs, cs], x[:, 1, rs, cs]
        ru, dd_u = oup[:, rs, cs + 1], oup[:, rs + 1, cs]
        rv, dd_v = ovp[:, rs, cs + 1], ovp[:, rs + 1, cs]
        if prev is None:
            lu = uu = lv = uv = torch.zeros_like(ou)
        else:
            pr0, pu, pv = prev
            pu, pv = F.pad(pu, (1, 1)), F.pad(pv, (1, 1))
            kl, ku = rs - pr0 + 1, rs - pr0
            lu, uu, lv, uv = pu[:, kl], pu[:, ku], pv[:, kl], pv[:, ku]
        nu = Bu + ax * (lu + ru) + ay * (dd_u + uu) - A12 * ov
        un = wu * nu + (1.0 - OMEGA + wu * b) * ou
        nv = Bv + ax * (lv + rv) + ay * (dd_v + uv) - A12 * un
        vn = wv * nv + (1.0 - OMEGA + wv * b) * ov
        outs_u.append(un)
        outs_v.append(vn)
        prev = (r0, un, vn)
    U = torch.cat(outs_u, dim=1)[:, inv].reshape(T, h, w)
    V = torch.cat(outs_v, dim=1)[:, inv].reshape(T, h, w)
    return torch.stack([U, V], dim=1)


def median5(x):
    """5x5 mirror median of (N, h, w) planes (h, w >= 3)."""
    N, h, w = x.shape
    P = F.pad(x.unsqueeze(1), (2, 2, 2, 2), mode="reflect")[:, 0]
    win = P.unfold(1, 5, 1).unfold(2, 5, 1).reshape(N, h, w, 25)
    return win.median(dim=-1).values


def lag_blocks(iterations, update_lag):
    starts = [0] + [it for it in range(1, iterations) if (it + 1) % update_lag == 0]
    ends = starts[1:] + [iterations]
    return [(s, e, s == 0 and update_lag != 1) for s, e in zip(starts, ends)]


def level_solve(f1, f2w, u0, ax, ay, a_data, W, iterations, update_lag, hx, hy, median):
    J = motion_tensor_gc(f1, f2w, hy, hx)
    x = torch.zeros_like(u0)
    for s, e, psi_one in lag_blocks(iterations, update_lag):
        coef = coefficients(x, u0, J, W, ax, ay, a_data, psi_one)
        for _ in range(s, e):
            x = sor_sweep(coef, x, ax, ay)
    if median:
        T, _, h, w = x.shape
        x = median5(x.reshape(2 * T, h, w)).reshape(T, 2, h, w)
    return u0 + x


def _keys_in(a):
    A = -0.75
    return ((A + 2) * a - (A + 3)) * a * a + 1


def _keys_out(a):
    A = -0.75
    return ((A * a - 5 * A) * a + 8 * A) * a - 4 * A


def warp(img, u, v, fill):
    """img (T, C, H, W); u, v (T, H, W); fill (C, H, W)."""
    T, C, H, W = img.shape
    xs = torch.arange(W, dtype=torch.float64)
    ys = torch.arange(H, dtype=torch.float64)[:, None]
    mx, my = xs + u, ys + v
    oob = (mx < 0) | (mx >= W) | (my < 0) | (my >= H)
    cx, cy = mx.clamp(0, W - 1), my.clamp(0, H - 1)
    x0, y0 = torch.floor(cx).detach(), torch.floor(cy).detach()
    tx, ty = cx - x0, cy - y0
    wx = [_keys_out(1 + tx), _keys_in(tx), _keys_in(1 - tx), _keys_out(2 - tx)]
    wy = [_keys_out(1 + ty), _keys_in(ty), _keys_in(1 - ty), _keys_out(2 - ty)]
    tt = torch.arange(T)[:, None, None].expand(T, H, W)
    out = 0.0
    for i in range(4):
        yi = (y0.long() - 1 + i).clamp(0, H - 1)
        for j in range(4):
            xj = (x0.long() - 1 + j).clamp(0, W - 1)
            S = img[tt, :, yi, xj]  # (T, H, W, C)
            out = out + (wy[i] * wx[j]).unsqueeze(-1) * S
    out = out.permute(0, 3, 1, 2)
    return torch.where(oob.unsqueeze(1), fill.unsqueeze(0).expand_as(out), out)


def _dense(in_len, out_len, sig, oracle):
    idx, wt = oracle.resize_tables(in_len, out_len, sig)
    M = np.zeros((out_len, in_len))
    np.add.at(M, (np.repeat(np.arange(out_len), idx.shape[1]), idx.ravel()), wt.ravel().astype(np.float64))
    return torch.from_numpy(M)


def resize(x, size, oracle):
    h, w = x.shape[-2:]
    oh, ow = size
    s = min(oh / h, ow / w)
    sig = (0.6 / s) if s < 1.0 else 0.0
    return _dense(h, oh, sig, oracle) @ x @ _dense(w, ow, sig, oracle).T


def get_displacement(fixed, moving, alpha, a_data, oracle, update_lag=5, iterations=50, levels=50,
                     eta=0.8):
    """fixed (C, H, W), moving (T, C, H, W), alpha (2,), a_data (C,) ->
    flows (T, 2, H, W); one-pass pyramid, continuous warp, uniform 1/C
    weights, min_level 0."""
    T, C, H, W = moving.shape
    alpha = torch.as_tensor(alpha, dtype=torch.float64)
    a_data = torch.as_tensor(a_data, dtype=torch.float64).reshape(-1).expand(C)
    plan, min_level = oracle.level_plan(H, W, eta, levels, 0)
    assert min_level == 0
    top = plan[0][0]
    wmap = torch.full((C, H, W), 1.0 / C, dtype=torch.float64)
    u = None
    for i, (lh, lw) in plan:
        f1 = resize(fixed, (lh, lw), oracle)
        f2 = resize(moving, (lh, lw), oracle)
        wl = resize(wmap, (lh, lw), oracle)
        hy, hx = H / lh, W / lw
        if i == top:
            u0 = torch.zeros((T, 2, lh, lw), dtype=torch.float64)
            f2w = f2
        else:
            u0 = resize(u, (lh, lw), oracle)
            f2w = warp(f2, u0[:, 0] / hx, u0[:, 1] / hy, f1)
        scale = 1.0 if i == min_level else eta ** (-0.5 * i)
        ax = alpha[0] * scale / (hx * hx)
        ay = alpha[1] * scale / (hy * hy)
        u = level_solve(f1, f2w, u0, ax, ay, a_data, wl, iterations, update_lag, hx, hy, min(lh, lw) > 5)
    return u


def loss(fixed, moving, flow, mask=None):
    r = warp(moving, flow[:, 0], flow[:, 1], fixed) - fixed
    sq = r * r
    if mask is not None:
        sq = sq * mask
    return sq.sum()
