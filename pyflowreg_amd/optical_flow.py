"""Drop-in ``get_displacement`` for pyflowreg on MI355X.

Mirrors ``core/optical_flow.py`` of pyflowreg: the same function names,
positional/keyword arguments, defaults and error behaviour, with every
numeric stage executed by libflowreg_hip.so on the GPU.  The host side only
plans the pyramid (level sizes, grid spacings, alpha scaling and the resize
tap tables, computed exactly as the reference computes them) and moves
arrays.

Entry points
  get_displacement(fixed, moving, **kw)         core/optical_flow.py:773-929
  level_solver(J11, ..., hx, hy, gnc_beta)       core/optical_flow.py:348-411
  get_motion_tensor_gc(f1, f2, hy, hx)           core/optical_flow.py:73-160
  imregister_wrapper(f2, u, v, f1, method)       core/warping.py:160-234
  imresize_fused_gauss_cubic(img, size)          util/resize_util.py:232-254
  warpingDepth(eta, levels, m, n)                core/warping.py:237-277
  DisplacementSolver                             batched form of get_displacement
"""

from __future__ import annotations

import functools
from dataclasses import dataclass

import numpy as np
import torch

from . import _lib
from ._lib import check, lib, ptr, scratch, stream_handle

_SUPPORTED_DATA_TERMS = {
    "gc": "gc", "gradient": "gc",
    "gray": "gray", "brightness": "gray",
    "cs": "cs", "census": "cs",
}
# FRH_TENSOR_* of include/flowreg_hip.h
_TENSOR_KIND = {"gc": 0, "gray": 1, "cs": 2}
_CENSUS_EPS = 0.1 / 255.0


def _device() -> torch.device:
    if not torch.cuda.is_available():
        raise _lib.FlowregHipError("flowreg_hip needs a ROCm GPU; none is visible")
    return torch.device("cuda", torch.cuda.current_device())


# ------------------------------------------------------------------ geometry
def warpingDepth(eta, levels, m, n):  # noqa: N802  (reference name)
    """core/warping.py:237-277."""
    return int(lib().frh_warping_depth(float(eta), int(levels), int(m), int(n)))


def level_plan(m: int, n: int, eta: float, levels: int, min_level: int):
    """core/optical_flow.py:497-518: [(i, (h, w)), ...] from coarse to fine
    and the effective min_level."""
    mly = warpingDepth(eta, levels, m, m)
    mlx = warpingDepth(eta, levels, n, n)
    max_level = min(mlx, mly) * 4
    mly = min(mly, max_level)
    mlx = min(mlx, max_level)
    top = max(mlx, mly)
    if top <= min_level:
        min_level = top - 1
    if min_level < 0:
        min_level = 0
    plan = [(i, (int(round(m * eta ** min(i, mly))), int(round(n * eta ** min(i, mlx)))))
            for i in range(top, min_level - 1, -1)]
    return plan, min_level


def _gauss_taps(sigma: float):
    """util/resize_util.py:216-225 with numpy float32 arithmetic."""
    if sigma <= 0.0:
        return 0, np.array([1.0], dtype=np.float32)
    R = int(np.ceil(2.0 * sigma))
    x = np.arange(-R, R + 1, dtype=np.float32)
    g = np.exp(-0.5 * (x / sigma) ** 2).astype(np.float32)
    g /= g.sum()
    return R, np.ascontiguousarray(g)


@functools.lru_cache(maxsize=512)
def resize_tables(in_len: int, out_len: int, sigma: float):
    """Tap indices (int32) and weights (float32), shape (out_len, 2R+4)."""
    R, g = _gauss_taps(sigma)
    P = 2 * R + 4
    idx = np.empty((out_len, P), np.int32)
    wt = np.empty((out_len, P), np.float32)
    check(lib().frh_fill_resize_tables(in_len, out_len, R, ptr(g), ptr(idx), ptr(wt)),
          "frh_fill_resize_tables")
    idx.flags.writeable = False
    wt.flags.writeable = False
    return idx, wt


def _resize_sigma(ih, iw, oh, ow, sigma_coeff=0.6):
    """util/resize_util.py:235-242 (per_axis=False)."""
    sy = oh / ih
    sx = ow / iw
    s = sy if sy < sx else sx
    return (sigma_coeff / s) if s < 1.0 else 0.0


@dataclass
class PyramidPlan:
    H: int
    W: int
    levels: list          # [(i, h, w, h_row, h_col, alpha_x, alpha_y)]
    min_level: int
    img_tables: list      # per level: (ix, wx, iy, wy)
    up_tables: list       # per level: None at top, else (ix, wx, iy, wy)


@functools.lru_cache(maxsize=64)
def pyramid_plan(H, W, eta, levels, min_level, alpha):
    plan, min_level = level_plan(H, W, eta, levels, min_level)
    lv, img, up = [], [], []
    prev = None
    for i, (h, w) in plan:
        h_row = float(H) / h
        h_col = float(W) / w
        alpha_scaling = 1 if i == min_level else eta ** (-0.5 * i)
        ax, ay = [alpha_scaling * alpha[j] for j in range(len(alpha))][:2]
        lv.append((i, h, w, h_row, h_col, ax, ay))
        sig = _resize_sigma(H, W, h, w)
        ix, wx = resize_tables(W, w, sig)
        iy, wy = resize_tables(H, h, sig)
        img.append((ix, wx, iy, wy))
        if prev is None:
            up.append(None)
        else:
            ph, pw = prev
            sig_u = _resize_sigma(ph, pw, h, w)
            ux, uwx = resize_tables(pw, w, sig_u)
            uy, uwy = resize_tables(ph, h, sig_u)
            up.append((ux, uwx, uy, uwy))
        prev = (h, w)
    return PyramidPlan(H, W, lv, min_level, img, up)


# ---------------------------------------------------------- argument helpers
def _resolve_const_assumption(const_assumption) -> str:
    """core/optical_flow.py:446-473."""
    if hasattr(const_assumption, "value"):
        const_assumption = const_assumption.value
    key = str(const_assumption).strip().lower()
    if key not in _SUPPORTED_DATA_TERMS:
        supported = "', '".join(sorted(_SUPPORTED_DATA_TERMS))
        raise ValueError(f"Unknown constancy assumption: '{const_assumption}'. "
                         f"Supported values are: '{supported}'.")
    return _SUPPORTED_DATA_TERMS[key]


def normalize_gnc_schedule(gnc_schedule):
    """core/optical_flow.py:414-432: validated float64 stage weights or None."""
    if gnc_schedule is None:
        return None
    sch = np.asarray(gnc_schedule, dtype=np.float64)
    if sch.ndim != 1:
        raise ValueError("gnc_schedule must be a 1D sequence of stage weights")
    if sch.size < 2:
        raise ValueError("gnc_schedule must contain at least two stages")
    if np.any(sch < 0.0) or np.any(sch > 1.0):
        raise ValueError("gnc_schedule entries must lie in [0, 1]")
    if not np.all(np.diff(sch) >= 0.0):
        raise ValueError("gnc_schedule must be monotone nondecreasing")
    if not np.isclose(sch[0], 0.0):
        raise ValueError("gnc_schedule must start at 0.0")
    if not np.isclose(sch[-1], 1.0):
        raise ValueError("gnc_schedule must end at 1.0")
    return np.ascontiguousarray(sch)


def normalize_warping_steps(warping_steps):
    """core/optical_flow.py:435-443."""
    if warping_steps is None:
        return None
    steps = int(warping_steps)
    if steps < 1:
        raise ValueError("warping_steps must be a positive integer")
    return steps


def expand_weight(weight, m, n, C) -> np.ndarray:
    """core/optical_flow.py:856-879: (m, n, C) float64 weight map."""
    if weight is None:
        return np.ones((m, n, C), dtype=np.float64) / C
    weight = np.asarray(weight).astype(np.float64)
    if weight.ndim < 3:
        if weight.ndim == 1:
            if len(weight) < C:
                we = np.full(C, 1.0 / C, dtype=np.float64)
                we[: len(weight)] = weight
                weight = we
            elif len(weight) > C:
                weight = weight[:C]
            weight = weight / weight.sum()
            weight = np.ones((m, n, C), dtype=np.float64) * weight.reshape(1, 1, -1)
        else:
            weight = np.ones((m, n, C), dtype=np.float64) * weight[..., np.newaxis]
    return weight


def _as_planes(a, dtype=torch.float32) -> torch.Tensor:
    """(H, W[, C]) or (T, H, W, C) array/tensor -> contiguous (.., C, H, W)
    device planes."""
    dev = _device()
    t = a if isinstance(a, torch.Tensor) else torch.from_numpy(np.ascontiguousarray(a))
    t = t.to(device=dev, dtype=dtype)
    if t.ndim == 2:
        t = t[None]
    elif t.ndim == 3:
        t = t.permute(2, 0, 1)
    else:
        t = t.permute(0, 3, 1, 2)
    return t.contiguous()


# ------------------------------------------------------------ batched solver
class DisplacementSolver:
    """Device-resident form of ``get_displacement`` for a batch of moving
    frames that share one reference, weight and initial flow (the situation
    of one BatchMotionCorrector batch).  Wraps frh_solver_*."""

    def __init__(self, H, W, C, *, alpha=(2, 2), update_lag=5, iterations=50, min_level=0,
                 levels=50, eta=0.8, a_smooth=1.0, a_data=0.45, max_batch=1,
                 const_assumption="gc", graph=False):
        if np.isscalar(alpha):
            alpha = (float(alpha), float(alpha))
        alpha = tuple(float(a) for a in alpha)
        self.H, self.W, self.C = int(H), int(W), int(C)
        self.max_batch = int(max_batch)
        self.iterations, self.update_lag = int(iterations), int(update_lag)
        self.a_smooth = float(a_smooth)
        self.plan = pyramid_plan(self.H, self.W, float(eta), int(levels), int(min_level), alpha)
        ad = np.broadcast_to(np.asarray(a_data, dtype=np.float64), (self.C,))
        self._a_data = np.ascontiguousarray(ad)
        keep = []
        descs = (_lib.LevelDesc * len(self.plan.levels))()
        for k, (i, h, w, h_row, h_col, ax, ay) in enumerate(self.plan.levels):
            d = descs[k]
            d.h, d.w, d.h_row, d.h_col, d.alpha_x, d.alpha_y = h, w, h_row, h_col, ax, ay
            ix, wx, iy, wy = self.plan.img_tables[k]
            d.P_img = ix.shape[1]
            d.img_ix, d.img_wx, d.img_iy, d.img_wy = ptr(ix), ptr(wx), ptr(iy), ptr(wy)
            keep += [ix, wx, iy, wy]
            if self.plan.up_tables[k] is not None:
                ux, uwx, uy, uwy = self.plan.up_tables[k]
                d.P_up = ux.shape[1]
                d.up_ix, d.up_wx, d.up_iy, d.up_wy = ptr(ux), ptr(uwx), ptr(uy), ptr(uwy)
                keep += [ux, uwx, uy, uwy]
        sd = _lib.SolverDesc()
        sd.H, sd.W, sd.C = self.H, self.W, self.C
        sd.n_levels = len(self.plan.levels)
        sd.levels = descs
        sd.iterations, sd.update_lag = int(iterations), int(update_lag)
        sd.a_smooth = float(a_smooth)
        sd.a_data = ptr(self._a_data)
        sd.max_batch = self.max_batch
        sd.const_assumption = _TENSOR_KIND[_resolve_const_assumption(const_assumption)]
        _device()
        h = _lib.c_void_p()
        check(lib().frh_solver_create(_lib.ctypes.byref(h), _lib.ctypes.byref(sd)),
              "frh_solver_create")
        self._h = h
        self._ref = None
        self._wcache = None  # (host float32 weight map, its device planes)
        if graph:
            self.set_graph(True)

    def set_graph(self, enable=True):
        """Replay the level loop as a hipGraph (captured on the first run with
        a given batch size and output dtype; bit-identical to eager runs)."""
        check(lib().frh_solver_set_graph(self._h, int(bool(enable))), "frh_solver_set_graph")

    def set_chains(self, enable=True):
        """Run passes of small batches as chains of work-groups (one per sweep,
        band and frame; the default, bit-identical); disable when other work
        shares the device."""
        check(lib().frh_solver_set_chains(self._h, int(bool(enable))), "frh_solver_set_chains")
        self._chains = bool(enable)

    def close(self):
        if getattr(self, "_h", None) is not None and self._h.value:
            lib().frh_solver_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_reference(self, fixed, weight=None, uv=None):
        """fixed: (H, W[, C]), or None before run_pairs only; weight as
        get_displacement accepts; uv (H, W, 2)."""
        f = None
        if fixed is not None:
            f = _as_planes(fixed)
            if f.shape != (self.C, self.H, self.W):
                raise ValueError(f"fixed has shape {tuple(f.shape)}, expected C,H,W="
                                 f"{(self.C, self.H, self.W)}")
        wmap = expand_weight(weight, self.H, self.W, self.C).astype(np.float32)
        # an unchanged weight map (every batch of a recording) keeps its
        # device planes: the upload from pageable memory would wait for the
        # work already queued on the stream
        wc = self._wcache
        if wc is not None and wc[0].shape == wmap.shape and np.array_equal(wc[0], wmap):
            w = wc[1]
        else:
            w = _as_planes(wmap)
            self._wcache = (wmap, w)
        u = None
        if uv is not None:
            u = uv if isinstance(uv, torch.Tensor) else torch.from_numpy(np.ascontiguousarray(uv))
            u = u.to(device=w.device, dtype=torch.float32).permute(2, 0, 1).contiguous()
        self._ref = (f, w, u)
        check(lib().frh_solver_set_reference(self._h, ptr(f), ptr(w), ptr(u), stream_handle()),
              "frh_solver_set_reference")

    def run(self, moving, out_dtype=torch.float64, profile=False) -> torch.Tensor:
        """moving: (T, H, W, C) or (H, W[, C]) -> flow (T, H, W, 2) on the GPU."""
        mv = moving if isinstance(moving, torch.Tensor) else torch.from_numpy(np.ascontiguousarray(moving))
        single = mv.ndim < 4
        if mv.ndim == 2:
            mv = mv[None, :, :, None]
        elif mv.ndim == 3:
            mv = mv[None]
        planes = _as_planes(mv)
        T = planes.shape[0]
        if planes.shape[1:] != (self.C, self.H, self.W):
            raise ValueError(f"moving frames have shape {tuple(planes.shape[1:])}, expected "
                             f"{(self.C, self.H, self.W)}")
        f64 = out_dtype == torch.float64 or self.plan.min_level > 0
        out = torch.empty((T, self.H, self.W, 2), dtype=torch.float64 if f64 else out_dtype,
                          device=planes.device)
        check(lib().frh_solver_profile(self._h, 1 if profile else 0), "frh_solver_profile")
        for s in range(0, T, self.max_batch):
            e = min(T, s + self.max_batch)
            check(lib().frh_solver_run(self._h, ptr(planes[s:e]), e - s, ptr(out[s:e]),
                                       1 if f64 else 0, stream_handle()),
                  "frh_solver_run")
        out = out.to(out_dtype)
        return out[0] if single else out

    def run_pairs(self, fixed, moving, out_dtype=torch.float64, profile=False) -> torch.Tensor:
        """Independent pairs: fixed, moving (T, H, W, C) -> flow (T, H, W, 2),
        each pair solved as its own get_displacement call (shared weight and
        initial flow from set_reference)."""
        fp, mp = self._pair_planes(fixed), self._pair_planes(moving)
        if fp.shape[0] != mp.shape[0]:
            raise ValueError("fixed and moving hold different numbers of frames")
        T = fp.shape[0]
        f64 = out_dtype == torch.float64 or self.plan.min_level > 0
        out = torch.empty((T, self.H, self.W, 2), dtype=torch.float64 if f64 else out_dtype,
                          device=mp.device)
        check(lib().frh_solver_profile(self._h, 1 if profile else 0), "frh_solver_profile")
        for s in range(0, T, self.max_batch):
            e = min(T, s + self.max_batch)
            check(lib().frh_solver_run_pairs(self._h, ptr(fp[s:e]), ptr(mp[s:e]), e - s, ptr(out[s:e]),
                                             1 if f64 else 0, stream_handle()), "frh_solver_run_pairs")
        return out.to(out_dtype)

    def _pair_planes(self, x):
        """(T, H, W[, C]) frames -> (T, C, H, W) float32 device planes."""
        t = x if isinstance(x, torch.Tensor) else torch.from_numpy(np.ascontiguousarray(x))
        if t.ndim == 3:
            t = t[..., None]
        p = _as_planes(t)
        if p.shape[1:] != (self.C, self.H, self.W):
            raise ValueError(f"frames have shape {tuple(p.shape[1:])}, expected {(self.C, self.H, self.W)}")
        return p

    def timings(self):
        """(ms, counts) per category of the last profiled run."""
        ms = np.zeros(7, np.float64)
        cnt = np.zeros(7, np.int32)
        lib().frh_solver_timings(self._h, ptr(ms), ptr(cnt), 7)
        names = ["total", "sweep", "assemble", "tensor", "resize", "warp", "median"]
        return {n: (float(ms[k]), int(cnt[k])) for k, n in enumerate(names)}


@functools.lru_cache(maxsize=8)
def _cached_solver(H, W, C, alpha, update_lag, iterations, min_level, levels, eta, a_smooth,
                   a_data, max_batch, term="gc"):
    return DisplacementSolver(H, W, C, alpha=alpha, update_lag=update_lag, iterations=iterations,
                              min_level=min_level, levels=levels, eta=eta, a_smooth=a_smooth,
                              a_data=np.array(a_data), max_batch=max_batch, const_assumption=term)


# ------------------------------------------------------------ get_displacement
def get_displacement(fixed, moving, alpha=(2, 2), update_lag=5, iterations=50, min_level=0,
                     levels=50, eta=0.8, a_smooth=1.0, a_data=0.45, const_assumption="gc",
                     uv=None, weight=None, level_solver_backend=None, gnc_schedule=None,
                     warping_steps=None):
    """core/optical_flow.py:773-929 on the GPU.

    Returns (H, W, 2) float64 numpy ([..., 0] = u, [..., 1] = v) for numpy
    inputs, or a float64 CUDA tensor when ``fixed`` is a tensor.  Without a
    GNC schedule or a caller-supplied level solver the whole pyramid runs on
    the device (DisplacementSolver); otherwise the reference's level loop
    runs on the host over the GPU stages (pyflowreg_amd.stage_loop).
    """
    assert fixed.ndim == moving.ndim, (
        f"Fixed and moving must have same dimensions: fixed.shape={fixed.shape}, "
        f"moving.shape={moving.shape}")
    term = _resolve_const_assumption(const_assumption)
    sched = normalize_gnc_schedule(gnc_schedule)
    steps = normalize_warping_steps(warping_steps)
    as_tensor = isinstance(fixed, torch.Tensor)
    if fixed.ndim == 2:
        m, n = fixed.shape
        C = 1
    else:
        m, n, C = fixed.shape
    if np.isscalar(alpha):
        alpha = (alpha, alpha)
    alpha = tuple(float(a) for a in alpha)
    ad = tuple(np.broadcast_to(np.asarray(a_data, np.float64), (C,)).tolist())
    if sched is not None or level_solver_backend is not None:
        from . import stage_loop
        tensor_fn = _TENSOR_FUNCS[term]
        solver_func = level_solver_backend if level_solver_backend is not None else level_solver
        f = fixed.cpu().numpy() if as_tensor else np.asarray(fixed)
        mv = moving.cpu().numpy() if isinstance(moving, torch.Tensor) else np.asarray(moving)
        uvn = uv.cpu().numpy() if isinstance(uv, torch.Tensor) else uv
        if sched is None:
            w = stage_loop.solve_stage(f, mv, alpha, update_lag, iterations, min_level, levels, eta,
                                       a_smooth, np.array(ad), uvn, weight, solver_func, tensor_fn)
        else:
            w = uvn
            for beta in sched:
                w = stage_loop.solve_stage_gnc(f, mv, alpha, update_lag, iterations, min_level,
                                               levels, eta, a_smooth, np.array(ad), w, weight,
                                               solver_func, tensor_fn, float(beta),
                                               10 if steps is None else steps)
        return torch.from_numpy(w).to(_device()) if as_tensor else w
    solver = _cached_solver(m, n, C, alpha, int(update_lag), int(iterations), int(min_level),
                            int(levels), float(eta), float(a_smooth), ad, 1, term)
    solver.set_reference(fixed, weight, uv)
    w = solver.run(moving, out_dtype=torch.float64)
    if as_tensor:
        return w
    return w.cpu().numpy()


def worklist_eligible(solvers) -> bool:
    """Whether frh_worklist accepts these solvers: at most 32, a_smooth = 1,
    at most 4 channels, levels of at most 1024 rows."""
    if not 1 <= len(solvers) <= 32:
        return False
    for s in solvers:
        if s.a_smooth != 1.0 or s.C > 4 or any(h > 1024 for _, h, *_ in s.plan.levels):
            return False
    return True


class WorkList:
    """Independent pairs of uneven sizes, channel counts and pyramid depths,
    one DisplacementSolver per (H, W, C), solved together: every stage of
    every level step is one launch over the frames of all solvers
    (frh_worklist_*; pyramids aligned at the finest level).  Each pair's
    flow equals its own get_displacement call bit for bit.  The solvers'
    weights and initial flows come from their set_reference calls."""

    def __init__(self, solvers):
        self.solvers = list(solvers)
        n = len(self.solvers)
        if not worklist_eligible(self.solvers):
            raise ValueError("frh_worklist needs 1-32 solvers with a_smooth = 1, at most 4 channels "
                             "and levels of at most 1024 rows")
        hs = (_lib.c_void_p * n)(*[s._h.value for s in self.solvers])
        h = _lib.c_void_p()
        check(lib().frh_worklist_create(_lib.ctypes.byref(h), hs, n), "frh_worklist_create")
        self._h = h

    def close(self):
        if getattr(self, "_h", None) is not None and self._h.value:
            lib().frh_worklist_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def run(self, fixed, moving, out_dtype=torch.float64):
        """fixed, moving: one (T_k, H_k, W_k, C_k) array or tensor per solver
        -> list of (T_k, H_k, W_k, 2) flows on the GPU."""
        n = len(self.solvers)
        if len(fixed) != n or len(moving) != n:
            raise ValueError(f"expected {n} fixed and moving batches")
        fps, mps, outs, Ts = [], [], [], []
        f64 = out_dtype == torch.float64 or any(s.plan.min_level > 0 for s in self.solvers)
        for s, f, m in zip(self.solvers, fixed, moving):
            fp, mp = s._pair_planes(f), s._pair_planes(m)
            if fp.shape[0] != mp.shape[0]:
                raise ValueError("fixed and moving hold different numbers of frames")
            if not 1 <= fp.shape[0] <= s.max_batch:
                raise ValueError(f"{fp.shape[0]} pairs for a solver of max_batch {s.max_batch}")
            fps.append(fp)
            mps.append(mp)
            Ts.append(fp.shape[0])
            outs.append(torch.empty((fp.shape[0], s.H, s.W, 2), dtype=torch.float64 if f64 else out_dtype,
                                    device=mp.device))
        arr = lambda xs: (_lib.c_void_p * n)(*[ptr(x) for x in xs])  # noqa: E731
        check(lib().frh_worklist_run(self._h, arr(fps), arr(mps), (_lib.c_int * n)(*Ts), arr(outs),
                                     1 if f64 else 0, stream_handle()), "frh_worklist_run")
        return [o.to(out_dtype) for o in outs]


def get_displacements(fixed_seq, moving_seq, alpha=(2, 2), update_lag=5, iterations=50,
                      min_level=0, levels=50, eta=0.8, a_smooth=1.0, a_data=0.45,
                      const_assumption="gc", weight=None, max_batch=256, worklist=None):
    """get_displacement for many independent (fixed, moving) pairs of uneven
    sizes and channel counts.  Pairs are grouped by (H, W, C), one solver
    per group; with worklist=None (where frh_worklist accepts the groups) or
    True all groups run as one work list (one launch per stage and level
    step), with worklist=False each group runs its own batched launches on
    one of 4 streams.  Returns a list of (H, W, 2) float64 numpy flows in
    input order; each equals the corresponding get_displacement(fixed,
    moving, ...) call."""
    term = _resolve_const_assumption(const_assumption)
    if np.isscalar(alpha):
        alpha = (alpha, alpha)
    alpha = tuple(float(a) for a in alpha)
    groups = {}
    pairs = list(zip(fixed_seq, moving_seq))
    for k, (f, m) in enumerate(pairs):
        f = np.asarray(f)
        key = f.shape if f.ndim == 3 else f.shape + (1,)
        groups.setdefault(key, []).append(k)
    out = [None] * len(pairs)
    jobs = []
    wl = None
    try:
        for (H, W, C), idx in groups.items():
            ad = tuple(np.broadcast_to(np.asarray(a_data, np.float64), (C,)).tolist())
            for s0 in range(0, len(idx), max_batch):
                part = idx[s0:s0 + max_batch]
                solver = DisplacementSolver(H, W, C, alpha=alpha, update_lag=update_lag,
                                            iterations=iterations, min_level=min_level, levels=levels,
                                            eta=eta, a_smooth=a_smooth, a_data=np.array(ad),
                                            max_batch=len(part), const_assumption=term)
                jobs.append((solver, None, None, part))
                solver.set_reference(None, weight)
                fx = np.stack([np.asarray(pairs[k][0], np.float32).reshape(H, W, C) for k in part])
                mv = np.stack([np.asarray(pairs[k][1], np.float32).reshape(H, W, C) for k in part])
                jobs[-1] = (solver, fx, mv, part)
        solvers = [j[0] for j in jobs]
        use_wl = worklist_eligible(solvers) if worklist is None else bool(worklist)
        if use_wl:
            wl = WorkList(solvers)
            results = wl.run([j[1] for j in jobs], [j[2] for j in jobs])
        else:
            results = run_groups([(s, fx, mv) for s, fx, mv, _ in jobs])
        for (_, _, _, idx), w in zip(jobs, results):
            w = w.cpu().numpy()
            for j, k in enumerate(idx):
                out[k] = w[j]
    finally:
        if wl is not None:
            wl.close()
        for s, *_ in jobs:
            s.close()
    return out


def run_groups(groups, n_streams=4, out_dtype=torch.float64):
    """Runs DisplacementSolver.run_pairs for several (solver, fixed, moving)
    groups on up to n_streams streams (default priority), so that the
    latency-bound sweeps of small groups share the device; returns the flows
    in group order after the caller's stream has waited for all of them."""
    dev = _device()
    cur = torch.cuda.current_stream(dev)
    streams = [torch.cuda.Stream(dev) for _ in range(max(1, min(n_streams, len(groups))))]
    # largest groups first, dealt round-robin
    order = sorted(range(len(groups)), key=lambda g: -groups[g][0].H * groups[g][0].W * groups[g][0].C)
    res = [None] * len(groups)
    # chains assume a device of their own: off while groups share it
    shared = len(streams) > 1
    prev = [getattr(g[0], "_chains", True) for g in groups]
    if shared:
        for g in groups:
            g[0].set_chains(False)
    try:
        for j, g in enumerate(order):
            s, fx, mv = groups[g]
            st = streams[j % len(streams)]
            st.wait_stream(cur)
            with torch.cuda.stream(st):
                res[g] = s.run_pairs(fx, mv, out_dtype=out_dtype)
    finally:
        if shared:
            for g, p in zip(groups, prev):
                g[0].set_chains(p)
    for st in streams:
        cur.wait_stream(st)
    return res


# ---------------------------------------------------------------- stages
def imresize_fused_gauss_cubic(img, size, sigma_coeff=0.6, per_axis=False):
    """util/resize_util.py:232-254 (per_axis=False) on the GPU."""
    if per_axis:
        raise ValueError("flowreg_hip: per_axis resize is not implemented")
    oh, ow = size[:2]
    src = np.asarray(img) if not isinstance(img, torch.Tensor) else img
    in_dtype = src.dtype
    h, w = src.shape[:2]
    sig = _resize_sigma(h, w, oh, ow, sigma_coeff)
    ix, wx = resize_tables(w, ow, sig)
    iy, wy = resize_tables(h, oh, sig)
    planes = _as_planes(src)
    N = planes.shape[0]
    dev = planes.device
    d = [torch.from_numpy(np.array(a)).to(dev) for a in (ix, wx, iy, wy)]
    dst = torch.empty((N, oh, ow), dtype=torch.float32, device=dev)
    tmp = torch.empty((N, h, ow), dtype=torch.float32, device=dev)
    check(lib().frh_resize(ptr(planes), N, h, w, ptr(dst), oh, ow, ptr(d[0]), ptr(d[1]),
                           ptr(d[2]), ptr(d[3]), ix.shape[1], ptr(tmp), stream_handle()),
          "frh_resize")
    out = dst.permute(1, 2, 0)
    if src.ndim == 2:
        out = out[:, :, 0]
    if isinstance(img, torch.Tensor):
        return out.contiguous()
    res = out.cpu().numpy()
    return res.astype(in_dtype, copy=False)


def remap_depth(raw_dtype) -> str:
    """The cv2.remap code path (core/warping.py:219-227 calls it on the raw
    frame) for a frame dtype: "f32", "f64", "u8", "u16" or "s16".  cv2.remap
    has no path for other integer depths (8-bit signed, 32-bit, 64-bit) or
    half floats, where the reference's call raises; so does this one
    (ValueError)."""
    d = np.dtype(raw_dtype)
    table = {np.dtype(np.float32): "f32", np.dtype(np.float64): "f64", np.dtype(np.uint8): "u8",
             np.dtype(np.uint16): "u16", np.dtype(np.int16): "s16"}
    if d not in table:
        raise ValueError(f"imregister_wrapper: cv2.remap has no {d} path (uint8, uint16, int16, float32 "
                         "and float64 frames are registered)")
    return table[d]


def imregister_wrapper(f2_level, u, v, f1_level, interpolation_method="cubic"):
    """core/warping.py:160-234 (cubic or linear) on the GPU; numpy in, numpy
    (float32) out.  u, v are the displacements in level pixels (the
    reference's u / h_col, v / h_row) of any dtype: the map float32(x + u) is
    formed in float64, as numpy forms grid + u.  cv2.remap's arithmetic
    follows the image dtype: float64 images (get_displacement's level images,
    the preprocessed frames of update_reference) sum in float64, float32
    images in float32; uint8, uint16 and int16 images take OpenCV's 8U and
    16U / 16S paths (frh_warp_u8, frh_warp_int16); other dtypes raise
    ValueError, as cv2.remap does."""
    method = interpolation_method.lower()
    if method not in ("cubic", "linear"):
        raise ValueError("Unsupported interpolation method. Use 'linear' or 'cubic'.")
    squeeze = np.asarray(f2_level).ndim == 2 if not isinstance(f2_level, torch.Tensor) else f2_level.ndim == 2
    raw = (torch.empty((), dtype=f2_level.dtype).numpy().dtype if isinstance(f2_level, torch.Tensor)
           else np.asarray(f2_level).dtype)
    depth = remap_depth(raw)
    f64 = depth == "f64"
    dt = torch.float64 if f64 else torch.float32
    f2 = _as_planes(f2_level, dt)
    f1 = _as_planes(f1_level, dt)
    C, H, W = f2.shape
    dev = f2.device

    def disp(a):
        if isinstance(a, torch.Tensor):
            return a.to(device=dev, dtype=torch.float64).contiguous()
        return torch.as_tensor(np.asarray(a).astype(np.float64), device=dev).contiguous()
    uu, vv = disp(u), disp(v)
    out = torch.empty((C, H, W), dtype=torch.float32, device=dev)
    lin = int(method == "linear")
    if depth == "u8":
        check(lib().frh_warp_u8(lin, ptr(f2), ptr(f1), 0, 1, C, H, W, ptr(uu), ptr(vv), ptr(out),
                                stream_handle()), "frh_warp_u8")
    elif depth in ("u16", "s16"):
        check(lib().frh_warp_int16(lin, int(depth == "s16"), ptr(f2), ptr(f1), 0, 1, C, H, W, ptr(uu), ptr(vv),
                                   ptr(out), stream_handle()), "frh_warp_int16")
    else:
        name = "frh_warp_cubic_ex" if method == "cubic" else "frh_warp_linear_ex"
        check(getattr(lib(), name)(ptr(f2), ptr(f1), int(f64), int(f64), 0, 1, C, H, W, ptr(uu), ptr(vv),
                                   ptr(out), stream_handle()), name)
    res = out.permute(1, 2, 0)
    res = res[:, :, 0] if squeeze else res
    if isinstance(f2_level, torch.Tensor):
        return res.contiguous()
    return res.cpu().numpy()


def _motion_tensor(kind, f1, f2, hy, hx, eps):
    """One channel on the GPU.  A float32 input makes numpy evaluate that
    image's derivatives in float32 (the reference's warped image below the
    top level); float64 inputs must hold float32-representable values (as
    every image on the reference's path does)."""
    dev = _device()
    flags = []
    planes = []
    for f in (f1, f2):
        t = f if isinstance(f, torch.Tensor) else torch.from_numpy(np.ascontiguousarray(f))
        flags.append(1 if t.dtype == torch.float32 else 0)
        t32 = t.to(device=dev, dtype=torch.float32).contiguous()
        if t.dtype != torch.float32 and not torch.equal(t32.double(), t.to(device=dev, dtype=torch.float64)):
            raise ValueError("flowreg_hip motion tensors take float32-valued images")
        planes.append(t32)
    h, w = planes[0].shape
    J = [torch.empty((h + 2, w + 2), dtype=torch.float64, device=dev) for _ in range(6)]
    check(lib().frh_motion_tensor(kind, ptr(planes[0]), ptr(planes[1]), h, w, float(hy), float(hx),
                                  float(eps), flags[0], flags[1], *[ptr(j) for j in J],
                                  stream_handle()), "frh_motion_tensor")
    return tuple(j.cpu().numpy() for j in J)


def get_motion_tensor_gc(f1, f2, hy, hx):
    """core/optical_flow.py:73-160 on the GPU (one channel); float64 outputs
    of shape (h+2, w+2)."""
    return _motion_tensor(0, f1, f2, hy, hx, _CENSUS_EPS)


def get_motion_tensor_gray(f1, f2, hy, hx):
    """core/optical_flow.py:163-228 on the GPU (one channel)."""
    return _motion_tensor(1, f1, f2, hy, hx, _CENSUS_EPS)


def get_motion_tensor_cs(f1, f2, hy, hx, eps=None):
    """core/optical_flow.py:231-345 on the GPU (one channel); eps defaults to
    0.1 / 255."""
    return _motion_tensor(2, f1, f2, hy, hx, _CENSUS_EPS if eps is None else float(eps))


_TENSOR_FUNCS = {"gc": get_motion_tensor_gc, "gray": get_motion_tensor_gray,
                 "cs": get_motion_tensor_cs}


def level_solver(J11, J22, J33, J12, J13, J23, weight, u, v, alpha, iterations, update_lag,
                 verbose, a_data, a_smooth, hx, hy, gnc_beta=None):
    """core/optical_flow.py:348-411 on the GPU (compute_flow, or
    compute_flow_gnc with gnc_beta): returns (du, dv) (m, n) float64."""
    dev = _device()

    def d(a):
        return torch.as_tensor(np.ascontiguousarray(a, dtype=np.float64), device=dev)
    Js = [d(j) for j in (J11, J22, J33, J12, J13, J23)]
    m, n, K = Js[0].shape
    W_ = d(weight).reshape(m, n, -1)
    uu, vv = d(u), d(v)
    ad = np.ascontiguousarray(np.broadcast_to(np.asarray(a_data, np.float64), (K,)))
    du = torch.empty((m, n), dtype=torch.float64, device=dev)
    dv = torch.empty((m, n), dtype=torch.float64, device=dev)
    nb = lib().frh_level_solver_scratch_bytes(m, n, K, float(a_smooth))
    work = scratch(nb, dev)
    check(lib().frh_level_solver(*[ptr(j) for j in Js], ptr(W_), ptr(uu), ptr(vv), m, n, K,
                                 float(alpha[0]), float(alpha[1]), int(iterations),
                                 int(update_lag), ptr(ad), float(a_smooth), float(hx), float(hy),
                                 -1.0 if gnc_beta is None else float(gnc_beta),
                                 ptr(du), ptr(dv), ptr(work), nb, stream_handle()), "frh_level_solver")
    return du.cpu().numpy(), dv.cpu().numpy()


def median5_mirror(f):
    """scipy.ndimage.median_filter(f, size=(5,5), mode='mirror') on the GPU
    for a 2-D float64 array (core/optical_flow.py:607-609)."""
    dev = _device()
    t = torch.as_tensor(np.ascontiguousarray(f, dtype=np.float64), device=dev).clone()
    h, w = t.shape
    nb = lib().frh_median5_mirror_scratch_bytes(1, h, w)
    work = scratch(nb, dev)
    check(lib().frh_median5_mirror(ptr(t), 1, h, w, ptr(work), nb, stream_handle()), "frh_median5_mirror")
    return t.cpu().numpy()


def cv_resize_cubic(w, m, n):
    """cv2.resize(w, (n, m), interpolation=cv2.INTER_CUBIC) for an (h, w, C)
    float64 flow (core/optical_flow.py:615-616) on the GPU."""
    dev = _device()
    src = torch.as_tensor(np.ascontiguousarray(w, dtype=np.float64), device=dev)
    h, ww, C = src.shape
    dst = torch.empty((m, n, C), dtype=torch.float64, device=dev)
    nb = lib().frh_resize_cubic_cv_scratch_bytes(1, h, n, C)
    work = scratch(nb, dev)
    check(lib().frh_resize_cubic_cv(ptr(src), 1, h, ww, C, ptr(dst), m, n, ptr(work), nb, stream_handle()),
          "frh_resize_cubic_cv")
    return dst.cpu().numpy()
