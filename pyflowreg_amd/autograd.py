"""Differentiable registration: torch.autograd Functions whose forward and
backward passes run the hand-written HIP kernels of csrc/frh_grad.hip.

``get_displacement`` restates one pass of core/optical_flow.py:476-617
(gradient-constancy data term, a_smooth = 1, no GNC) for a batch of moving
frames against one fixed frame, with two changes that make it differentiable:

* the warp between levels is continuous: map = x + u, without the 1/32-px
  fixed-point map of cv2.remap (core/warping.py:160-234) and without rounding
  the warped image to float32;
* level images, resized flows and weights stay float64 (the reference rounds
  them to float32 in util/resize_util.py:232-254).

With ``levels=0`` (one level, no warp) and float32-valued inputs the forward
equals the reference's ``get_displacement(..., levels=0)`` up to rounding.

Gradients reach ``alpha`` (2), ``a_data`` (per channel) and the moving frames
(through the level images, the warps and the motion tensors); ``fixed`` and
``weight`` are treated as constants.  The level solve's backward is the exact
reverse of its unrolled, lag-refreshed SOR sweeps and medians -- not an
implicit adjoint of a converged system, which would not match the 50 sweeps
the forward actually runs: the state is checkpointed at every coefficient
refresh and each lag block is recomputed from its checkpoint in the backward
pass.

Layouts (device, float64, contiguous): images (T, C, h, w), flows
(T, 2, h, w) with plane 0 = u (x) and plane 1 = v (y); inside a level solve
the sweeps' states and coefficients use the skewed layout (``skew_plane``).
"""

from __future__ import annotations

import numpy as np
import torch

from ._lib import check, lib, ptr, stream_handle
from .optical_flow import _resize_sigma, level_plan, resize_tables

_F64 = torch.float64


def _c(t: torch.Tensor) -> torch.Tensor:
    if t.dtype != _F64 or not t.is_cuda:
        raise TypeError("flowreg_hip autograd: expected float64 CUDA tensors")
    return t.contiguous()


def _chain_scratch(L, nmax, T, dev, beside=1):
    """Flags (zeroed) for a lag block's sweeps as a chain of work-groups, one
    per sweep and frame, when those fit on the device at once, beside
    `beside` - 1 other chains of the same size (batches that would leave
    most compute units idle); None: one work-group per frame."""
    if beside * T * nmax > torch.cuda.get_device_properties(dev).multi_processor_count:
        return None
    return torch.zeros(L.frh_grad_sweeps_scratch_bytes(nmax, T), dtype=torch.uint8, device=dev)


def _scratch_args(scr):
    return (ptr(scr), scr.numel()) if scr is not None else (None, 0)


def _check_chain(scr, what):
    """The failure word of a chain's scratch: set when a stage gave up waiting
    for its predecessor (its results are invalid)."""
    if scr is not None and int(scr[:4].view(torch.int32)[0]) != 0:
        raise RuntimeError(f"{what}: a chained stage gave up waiting for its predecessor")


def lag_blocks(iterations: int, update_lag: int):
    """Coefficient refresh schedule of core/level_solver.py:229-255: a block
    starts at iteration 0 and at every iteration with (it + 1) % update_lag
    == 0; the first block uses psi = 1 unless iteration 0 is itself a
    refresh.  Returns [(start, stop, psi_one)]."""
    if iterations < 1 or update_lag < 1:
        raise ValueError("iterations and update_lag must be >= 1")
    starts = [0] + [it for it in range(1, iterations) if (it + 1) % update_lag == 0]
    ends = starts[1:] + [iterations]
    return [(s, e, s == 0 and update_lag != 1) for s, e in zip(starts, ends)]


# ------------------------------------------------------------------- warp
class ContinuousWarp(torch.autograd.Function):
    """out = img sampled at (x + u, y + v) with Keys cubic (A = -0.75),
    replicated borders and the unclipped map tested against the image (out
    of bounds -> fill); core/warping.py:160-234 with a continuous map.

    img (T, C, H, W); fill (C, H, W) shared or (T, C, H, W); u, v (T, H, W)
    in pixels.  Gradients: img, fill (where out of bounds), u, v."""

    @staticmethod
    def forward(ctx, img, fill, u, v):
        img, fill, u, v = _c(img), _c(fill), _c(u), _c(v)
        T, C, H, W = img.shape
        if u.shape != (T, H, W) or v.shape != (T, H, W) or fill.shape[-3:] != (C, H, W) or \
                (fill.dim() == 4 and fill.shape[0] != T) or fill.dim() not in (3, 4):
            raise ValueError("ContinuousWarp: img (T, C, H, W), u and v (T, H, W), fill (C, H, W) "
                             "or (T, C, H, W)")
        per_frame = int(fill.dim() == 4)
        out = torch.empty_like(img)
        check(lib().frh_grad_warp(ptr(img), ptr(fill), per_frame, ptr(u), ptr(v), T, C, H, W, ptr(out),
                                  None, None, None, None, stream_handle()), "frh_grad_warp")
        ctx.save_for_backward(img, fill, u, v)
        return out

    @staticmethod
    def backward(ctx, g_out):
        img, fill, u, v = ctx.saved_tensors
        T, C, H, W = img.shape
        g_out = _c(g_out)
        g_img = torch.zeros_like(img) if ctx.needs_input_grad[0] else None
        g_u = torch.empty_like(u)
        g_v = torch.empty_like(v)
        check(lib().frh_grad_warp(ptr(img), ptr(fill), int(fill.dim() == 4), ptr(u), ptr(v), T, C, H, W,
                                  None, ptr(g_out), ptr(g_img), ptr(g_u), ptr(g_v), stream_handle()),
              "frh_grad_warp")
        g_fill = None
        if ctx.needs_input_grad[1]:
            xs = torch.arange(W, device=u.device, dtype=_F64)
            ys = torch.arange(H, device=u.device, dtype=_F64)[:, None]
            mx, my = xs + u, ys + v
            oob = ((mx < 0) | (mx >= W) | (my < 0) | (my >= H)).unsqueeze(1)
            g_fill = g_out * oob
            if fill.dim() == 3:
                g_fill = g_fill.sum(0)
        return g_img, g_fill, g_u, g_v


def warp(img, u, v, fill):
    """Continuous warp (see ContinuousWarp)."""
    return ContinuousWarp.apply(img, fill, u, v)


# ------------------------------------------------------------ level solve
def skew_plane(h: int, w: int) -> int:
    """Doubles per plane of the solver's skewed layout: cell (r, c) at slot
    (r + c) * hs + r, hs = 64 * ceil(h / 64)."""
    return (h + w - 1) * 64 * ((h + 63) // 64)


def to_skew(x: torch.Tensor) -> torch.Tensor:
    """(..., K, h, w) row-major planes -> (..., plane, K) skewed, each cell's K
    values interleaved in its slot (padding = 0)."""
    K, h, w = x.shape[-3:]
    lead = x.shape[:-3]
    src = _c(x)
    out = torch.empty(lead + (skew_plane(h, w), K), dtype=_F64, device=x.device)
    n = int(np.prod(lead)) if lead else 1
    check(lib().frh_grad_skew(ptr(src), n, K, h, w, ptr(out), 1, stream_handle()), "frh_grad_skew")
    return out


def from_skew(x: torch.Tensor, h: int, w: int) -> torch.Tensor:
    """(..., plane, K) skewed -> (..., K, h, w) row-major."""
    lead, K = x.shape[:-2], x.shape[-1]
    src = _c(x)
    out = torch.empty(lead + (K, h, w), dtype=_F64, device=x.device)
    n = int(np.prod(lead)) if lead else 1
    check(lib().frh_grad_skew(ptr(src), n, K, h, w, ptr(out), 0, stream_handle()), "frh_grad_skew")
    return out


class LevelSolve(torch.autograd.Function):
    """One pyramid level of core/optical_flow.py:582-613:
    J = get_motion_tensor_gc(f1, f2w); (du, dv) = compute_flow(J, u0) with
    ``iterations`` column-major SOR sweeps and coefficient refreshes every
    ``update_lag`` sweeps; the 5x5 mirror median of du, dv when the level is
    larger than 5 in both directions; returns u0 + (du, dv).

    f1 (C, h, w) fixed level (constant); f2w (T, C, h, w) warped moving
    level; u0 (T, 2, h, w) the level's initial flow; ax = alpha_x / hx^2 and
    ay = alpha_y / hy^2 (0-dim); a_data (C,); weight (C, h, w) (constant).
    The sweeps' states, coefficients and adjoints are kept in the skewed
    layout (``skew_plane``)."""

    @staticmethod
    def forward(ctx, f1, f2w, u0, ax, ay, a_data, weight, iterations, update_lag, hx, hy, median):
        f1, f2w, u0, a_data, weight = _c(f1), _c(f2w), _c(u0), _c(a_data), _c(weight)
        T, C, h, w = f2w.shape
        if (f1.shape != (C, h, w) or u0.shape != (T, 2, h, w) or weight.shape != (C, h, w)
                or a_data.shape != (C,)):
            raise ValueError("LevelSolve: f1 (C, h, w), f2w (T, C, h, w), u0 (T, 2, h, w), "
                             "weight (C, h, w) and a_data (C,) must agree")
        if C > 2 or h > 1024:
            raise ValueError("LevelSolve: at most 2 channels and 1024 rows")
        L, st = lib(), stream_handle()
        axf, ayf = float(ax), float(ay)
        P = skew_plane(h, w)
        dev = f2w.device
        J = torch.empty((T, 6, C, h, w), dtype=_F64, device=dev)
        check(L.frh_grad_tensor(ptr(f1), ptr(f2w), T, C, h, w, hy, hx, ptr(J), None, None, st),
              "frh_grad_tensor")
        blocks = lag_blocks(iterations, update_lag)
        # ckpt[b]: the state entering lag block b (the checkpoints the
        # backward recomputes from); each block's sweeps write the next one
        ckpt = torch.empty((len(blocks) + 1, T, P, 2), dtype=_F64, device=dev)
        ckpt[0].zero_()
        coef = torch.empty((T, P, 6), dtype=_F64, device=dev)
        scr = _chain_scratch(L, max(e - s for s, e, _ in blocks), T, dev)
        for b, (s, e, psi_one) in enumerate(blocks):
            check(L.frh_grad_coef(ptr(ckpt[b]), ptr(u0), ptr(J), ptr(weight), T, C, h, w, axf, ayf, ptr(a_data),
                                  int(psi_one), ptr(coef), None, None, None, None, None, st),
                  "frh_grad_coef")
            # the block's sweeps share the coefficients: one chain of
            # work-groups or pipelined launches
            check(L.frh_grad_sweeps(ptr(coef), ptr(ckpt[b]), ptr(ckpt[b + 1]), 0, e - s, 0, T, h, w, axf, ayf,
                                    *_scratch_args(scr), st), "frh_grad_sweeps")
        _check_chain(scr, "LevelSolve forward")
        du = from_skew(ckpt[len(blocks)], h, w)
        if median:
            dm = torch.empty_like(du)
            check(L.frh_grad_median(ptr(du), 2 * T, h, w, ptr(dm), None, None, st), "frh_grad_median")
        else:
            dm = du
        ctx.save_for_backward(f1, f2w, u0, a_data, weight, J, ckpt, du)
        ctx.cfg = (iterations, update_lag, hx, hy, median, axf, ayf)
        return u0 + dm

    @staticmethod
    def backward(ctx, g_u):
        f1, f2w, u0, a_data, weight, J, ckpt, du = ctx.saved_tensors
        iterations, update_lag, hx, hy, median, axf, ayf = ctx.cfg
        T, C, h, w = f2w.shape
        L, st = lib(), stream_handle()
        P = skew_plane(h, w)
        dev = du.device
        g_u = _c(g_u)
        g_u0 = g_u.clone()
        if median:
            g_du = torch.zeros_like(du)
            check(L.frh_grad_median(ptr(du), 2 * T, h, w, None, ptr(g_u), ptr(g_du), st), "frh_grad_median")
        else:
            g_du = g_u
        lam = to_skew(g_du)
        lam_in = torch.empty_like(lam)
        g_J = torch.zeros_like(J)
        g_par = torch.zeros((T, 2 + C, P), dtype=_F64, device=dev)
        g_axy = torch.zeros((T, 2, P), dtype=_F64, device=dev)
        g_coef = torch.empty((T, P, 6), dtype=_F64, device=dev)
        blocks = lag_blocks(iterations, update_lag)
        nmax = max(e - s for s, e, _ in blocks)
        # Lag blocks run backwards.  Recomputing a block's sweeps from its
        # checkpoint does not depend on the later blocks' adjoints, so the
        # recompute of the next block runs on a side stream while the main
        # stream runs the current block's adjoint (each uses one work-group
        # per frame); two state / coefficient buffers alternate.
        bufs = [torch.empty((nmax, T, P, 2), dtype=_F64, device=dev) for _ in range(2)]
        chain = torch.zeros(L.frh_grad_sweeps_adjoint_scratch_bytes(nmax, T, h, w), dtype=torch.uint8, device=dev)
        # the recompute (side stream) runs beside the adjoint chain: chained
        # only when both fit (32 frames of 512^2: 93.7 frames/s with the
        # recompute pipelined against 91.0 chained)
        scr = _chain_scratch(L, nmax, T, dev, beside=2)
        coefs = [torch.empty((T, P, 6), dtype=_F64, device=dev) for _ in range(2)]
        main = torch.cuda.current_stream(dev)
        side = torch.cuda.Stream(device=dev)
        side.wait_stream(main)
        hs = side.cuda_stream

        def recompute(b, slot):
            s, e, psi_one = blocks[b]
            check(L.frh_grad_coef(ptr(ckpt[b]), ptr(u0), ptr(J), ptr(weight), T, C, h, w, axf, ayf, ptr(a_data),
                                  int(psi_one), ptr(coefs[slot]), None, None, None, None, None, hs),
                  "frh_grad_coef")
            check(L.frh_grad_sweeps(ptr(coefs[slot]), ptr(ckpt[b]), ptr(bufs[slot]), bufs[slot][0].numel(), e - s,
                                    1, T, h, w, axf, ayf, *_scratch_args(scr), hs), "frh_grad_sweeps")
            ev = torch.cuda.Event()
            ev.record(side)
            return ev

        order = list(range(len(blocks) - 1, -1, -1))
        ev_rec = recompute(order[0], 0)
        ev_adj = None
        for i, b in enumerate(order):
            slot = i % 2
            ev_next = None
            if i + 1 < len(order):
                if ev_adj is not None:  # the other buffers served the previous block's adjoint
                    side.wait_event(ev_adj)
                ev_next = recompute(order[i + 1], 1 - slot)
            main.wait_event(ev_rec)
            s, e, psi_one = blocks[b]
            # the block's adjoint sweeps as one chain of work-groups per frame;
            # the block's last sweep (its first adjoint) sets g_coef, the
            # others add to it
            check(L.frh_grad_sweeps_adjoint(ptr(coefs[slot]), ptr(ckpt[b]), ptr(bufs[slot]), bufs[slot][0].numel(),
                                            e - s, ptr(lam), ptr(lam_in), ptr(g_coef), 1, ptr(g_axy), T, h, w,
                                            axf, ayf, ptr(chain), chain.numel(), st), "frh_grad_sweeps_adjoint")
            lam, lam_in = lam_in, lam
            # the block's coefficients depend on its start state, J, u0 and
            # the parameters; lam becomes the adjoint of the start state
            check(L.frh_grad_coef(ptr(ckpt[b]), ptr(u0), ptr(J), ptr(weight), T, C, h, w, axf, ayf, ptr(a_data),
                                  int(psi_one), None, ptr(g_coef), ptr(lam), ptr(g_u0), ptr(g_J), ptr(g_par),
                                  st), "frh_grad_coef")
            ev_adj = torch.cuda.Event()
            ev_adj.record(main)
            ev_rec = ev_next
        main.wait_stream(side)
        _check_chain(chain, "LevelSolve backward")
        _check_chain(scr, "LevelSolve backward (recompute)")
        # the solve starts from du = dv = 0: lam (its adjoint) is dropped
        g_f2w = None
        if ctx.needs_input_grad[1]:
            g_f2w = torch.zeros_like(f2w)
            check(L.frh_grad_tensor(ptr(f1), ptr(f2w), T, C, h, w, hy, hx, None, ptr(g_J), ptr(g_f2w), st),
                  "frh_grad_tensor")
        g_ax = (g_par[:, 0].sum() + g_axy[:, 0].sum()).reshape(())
        g_ay = (g_par[:, 1].sum() + g_axy[:, 1].sum()).reshape(())
        g_a = g_par[:, 2:].sum(dim=(0, 2))
        return None, g_f2w, g_u0, g_ax, g_ay, g_a, None, None, None, None, None, None


# ----------------------------------------------------------------- resize
def _resize_matrix(in_len, out_len, sigma, device):
    """Dense (out_len, in_len) matrix of util/resize_util.py:194-213's fused
    Gaussian+cubic taps (reflected indices summed)."""
    idx, wt = resize_tables(in_len, out_len, sigma)
    M = np.zeros((out_len, in_len))
    rows = np.repeat(np.arange(out_len), idx.shape[1])
    np.add.at(M, (rows, idx.ravel()), wt.ravel().astype(np.float64))
    return torch.from_numpy(M).to(device)


def resize(x, size, device=None):
    """imresize_fused_gauss_cubic (util/resize_util.py:232-254) of
    (..., h, w) planes to size, as two dense products (differentiable)."""
    h, w = x.shape[-2:]
    oh, ow = size
    sig = _resize_sigma(h, w, oh, ow)
    Ry = _resize_matrix(h, oh, sig, x.device)
    Rx = _resize_matrix(w, ow, sig, x.device)
    return Ry @ x @ Rx.transpose(0, 1)


# ------------------------------------------------------ get_displacement
def _planes(a, lead):
    """(H, W[, C]) -> (C, H, W); (T, H, W[, C]) -> (T, C, H, W)."""
    if a.dim() == lead + 2:
        return a.unsqueeze(-3)
    return a.movedim(-1, -3)


def get_displacement(fixed, moving, alpha=(2.0, 2.0), update_lag=5, iterations=50, min_level=0,
                     levels=50, eta=0.8, a_data=0.45, weight=None):
    """Differentiable get_displacement for T moving frames against one fixed
    frame (gradient constancy, a_smooth = 1).

    fixed (H, W[, C]), moving (T, H, W[, C]) float64 CUDA tensors; alpha a
    2-element tensor or pair; a_data a scalar or (C,) tensor; weight None
    (1/C per channel), a (C,) vector or an (H, W[, C]) map.  Returns the
    flows (T, H, W, 2) of core/optical_flow.py:476-617 with the continuous
    warp between levels."""
    fixed = _c(fixed)
    moving = _c(moving)
    dev = moving.device
    f = _planes(fixed, 0)
    m_ = _planes(moving, 1)
    T, C, H, W = m_.shape
    if f.shape != (C, H, W):
        raise ValueError(f"fixed {tuple(fixed.shape)} does not match moving {tuple(moving.shape)}")
    if C > 2:
        raise ValueError("flowreg_hip autograd: at most 2 channels")
    alpha = torch.as_tensor(alpha, dtype=_F64, device=dev).reshape(-1)
    if alpha.numel() == 1:
        alpha = alpha.expand(2)
    a_data = torch.as_tensor(a_data, dtype=_F64, device=dev).reshape(-1)
    if a_data.numel() == 1:
        a_data = a_data.expand(C)
    a_data = a_data.contiguous()
    if weight is None:
        wmap = torch.full((C, H, W), 1.0 / C, dtype=_F64, device=dev)
    else:
        wt = torch.as_tensor(weight, dtype=_F64, device=dev)
        if wt.dim() == 1:
            wv = torch.full((C,), 1.0 / C, dtype=_F64, device=dev)
            k = min(C, wt.numel())
            wv[:k] = wt[:k]
            wv = wv / wv.sum()
            wmap = wv[:, None, None].expand(C, H, W).contiguous()
        elif wt.dim() == 2:
            wmap = wt.expand(C, H, W).contiguous()
        else:
            wmap = wt.movedim(-1, 0).contiguous()
    plan, eff_min = level_plan(H, W, eta, levels, min_level)
    if eff_min != 0:
        raise ValueError("flowreg_hip autograd: min_level > 0 (a final cv2 resize) is not differentiable here")
    top = plan[0][0]
    u = None
    for i, (lh, lw) in plan:
        if lh > 1024:
            raise ValueError("flowreg_hip autograd: levels of at most 1024 rows")
        f1 = resize(f, (lh, lw)).contiguous()
        f2 = resize(m_, (lh, lw)).contiguous()
        wl = resize(wmap, (lh, lw)).contiguous()
        hy = float(H) / lh
        hx = float(W) / lw
        if i == top:
            u0 = torch.zeros((T, 2, lh, lw), dtype=_F64, device=dev)
            f2w = f2
        else:
            u0 = resize(u, (lh, lw)).contiguous()
            f2w = warp(f2, u0[:, 0] / hx, u0[:, 1] / hy, f1)
        scale = 1.0 if i == eff_min else eta ** (-0.5 * i)
        ax = alpha[0] * scale / (hx * hx)
        ay = alpha[1] * scale / (hy * hy)
        u = LevelSolve.apply(f1, f2w, u0, ax, ay, a_data, wl, int(iterations), int(update_lag), hx, hy,
                             min(lh, lw) > 5)
    return u.permute(0, 2, 3, 1)


def registration_loss(fixed, moving, flow, mask=None):
    """Masked sum of squared differences between the moving frames warped by
    ``flow`` (T, H, W, 2) (continuous warp, out of bounds -> fixed) and the
    fixed frame."""
    f = _planes(_c(fixed), 0)
    m_ = _planes(_c(moving), 1)
    r = warp(m_, flow[..., 0], flow[..., 1], f) - f
    sq = r * r
    if mask is not None:
        mk = torch.as_tensor(mask, dtype=_F64, device=sq.device)
        sq = sq * (mk if mk.dim() != 2 else mk[None, None])
    return sq.sum()
