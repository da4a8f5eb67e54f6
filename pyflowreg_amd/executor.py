"""Batched GPU executor with pyflowreg's executor contract.

``HipBatchedExecutor.process_batch`` follows
motion_correction/parallelization/base.py:76-110 and the sequential
executor's semantics (parallelization/sequential.py:27-169): every frame of
the batch is registered against ``reference_proc`` starting from ``w_init``,
then the raw frame is warped onto ``reference_raw`` with the flow.  Instead of
one ``get_displacement`` call per frame, all frames of the batch are solved
together by one ``DisplacementSolver`` run (one launch per kernel and pyramid
level for the whole batch), and the registration warp is one launch.
"""

from __future__ import annotations

from typing import Callable, Optional, Tuple

import numpy as np
import torch

from . import _lib
from ._lib import check, lib, ptr, stream_handle
from .optical_flow import DisplacementSolver, _resolve_const_assumption, remap_depth

_EXECUTORS = {}


def register_parallelization_executor(name: str, cls) -> None:
    """Mirror of RuntimeContext.register_parallelization_executor
    (pyflowreg/_runtime.py:151-171) for environments without pyflowreg."""
    _EXECUTORS[name] = cls


def get_parallelization_executor(name: str):
    return _EXECUTORS[name]


def _device():
    if not torch.cuda.is_available():
        raise _lib.FlowregHipError("flowreg_hip needs a ROCm GPU; none is visible")
    return torch.device("cuda", torch.cuda.current_device())


def warp_frames(frames: torch.Tensor, flow: torch.Tensor, fill: torch.Tensor,
                interpolation_method: str = "cubic", raw_dtype=None) -> torch.Tensor:
    """core/warping.py:160-234 (cubic or linear) for a batch on the GPU.
    frames (T, H, W, C) float32 (cv2.remap's 32F path) or float64 (its 64F
    path, double sums), flow (T, H, W, 2) of any float dtype (the map
    float32(x + u) is formed in float64), fill (H, W, C) -> (T, H, W, C)
    float32.  raw_dtype: the frames' dtype before staging.  uint16 / int16:
    frames hold those integer samples (as float32) and cv2.remap works in
    their depth, rounding each in-bounds result half to even and clamping
    it to the type's range (frh_warp_int16).  uint8: its 8U path, 15-bit
    fixed-point weights and results rounded and saturated to [0, 255]
    (frh_warp_u8).  Dtypes cv2.remap has no path for raise ValueError."""
    method = str(interpolation_method).lower()
    if method not in ("cubic", "linear"):
        raise ValueError("Unsupported interpolation method. Use 'linear' or 'cubic'.")
    depth = remap_depth(raw_dtype) if raw_dtype is not None else None
    integer = depth in ("u8", "u16", "s16")
    if integer and frames.dtype != torch.float32:
        raise ValueError("warp_frames: integer raw frames are passed as float32")
    T, H, W, C = frames.shape
    f64 = frames.dtype == torch.float64
    f2 = frames.permute(0, 3, 1, 2).contiguous()
    f1 = fill.to(frames.dtype).permute(2, 0, 1).contiguous()
    u = flow[..., 0].to(torch.float64).contiguous()
    v = flow[..., 1].to(torch.float64).contiguous()
    out = torch.empty(f2.shape, dtype=torch.float32, device=f2.device)
    if depth == "u8":
        check(lib().frh_warp_u8(int(method == "linear"), ptr(f2), ptr(f1), 0, T, C, H, W, ptr(u), ptr(v),
                                ptr(out), stream_handle()), "frh_warp_u8")
        return out.permute(0, 2, 3, 1)
    if integer:
        signed = int(depth == "s16")
        check(lib().frh_warp_int16(int(method == "linear"), signed, ptr(f2), ptr(f1), 0, T, C, H, W, ptr(u),
                                   ptr(v), ptr(out), stream_handle()), "frh_warp_int16")
        return out.permute(0, 2, 3, 1)
    name = "frh_warp_cubic_ex" if method == "cubic" else "frh_warp_linear_ex"
    check(getattr(lib(), name)(ptr(f2), ptr(f1), int(f64), int(f64), 0, T, C, H, W, ptr(u), ptr(v), ptr(out),
                               stream_handle()), name)
    return out.permute(0, 2, 3, 1)


def mean_frames(w: torch.Tensor) -> torch.Tensor:
    """np.mean(w, axis=0) of a float32 (n, ...) device tensor, bit for bit
    (frh_mean_frames_f32: the frames summed in order in float32, then
    divided by n)."""
    if w.dtype != torch.float32 or w.shape[0] < 1:
        raise ValueError("mean_frames: a float32 tensor with at least one frame")
    w = w.contiguous()
    out = torch.empty(w.shape[1:], dtype=torch.float32, device=w.device)
    check(lib().frh_mean_frames_f32(ptr(w), int(w.shape[0]), int(out.numel()), ptr(out), stream_handle()),
          "frh_mean_frames_f32")
    return out


def compute_batch_valid_masks(w) -> np.ndarray:
    """core/warping.py:14-60,128-157: uint8 masks (0/255) of pixels whose
    backward-warp position x + u, y + v stays inside the frame."""
    t = torch.as_tensor(w, device=_device())
    T, H, W = t.shape[:3]
    gy = torch.arange(H, device=t.device, dtype=torch.float64)[None, :, None]
    gx = torch.arange(W, device=t.device, dtype=torch.float64)[None, None, :]
    mx = gx + t[..., 0].double()
    my = gy + t[..., 1].double()
    valid = (mx >= 0) & (mx < W) & (my >= 0) & (my < H)
    return (valid.to(torch.uint8) * 255).cpu().numpy()


def _to_numpy32(a):
    if isinstance(a, torch.Tensor):
        return a.to(torch.float32).cpu().numpy()
    return np.ascontiguousarray(np.asarray(a).astype(np.float32, copy=False))


def _to_device32(a):
    if isinstance(a, torch.Tensor):
        return a.to(torch.float32)
    return torch.from_numpy(np.ascontiguousarray(np.asarray(a).astype(np.float32, copy=False)))


class HipBatchedExecutor:
    """Executor 'hip_batched': whole batches per GPU launch."""

    name = "hip_batched"

    def __init__(self, n_workers: Optional[int] = None, max_batch: int = 256):
        self.n_workers = 1
        self.max_batch = int(max_batch)
        self.reserve = 0  # frames a new solver makes room for at least
        self._key = None
        self._solver = None

    @classmethod
    def register(cls):
        register_parallelization_executor(cls.name, cls)
        try:  # pyflowreg present: register there too
            from pyflowreg._runtime import RuntimeContext  # type: ignore
            RuntimeContext.register_parallelization_executor(cls.name, cls)
        except Exception:
            pass

    def setup(self):
        _device()

    def cleanup(self):
        if self._solver is not None:
            self._solver.close()
        self._solver = None
        self._key = None

    def __enter__(self):
        self.setup()
        return self

    def __exit__(self, exc_type, exc, tb):
        self.cleanup()
        return False

    def get_info(self) -> dict:
        return {"name": self.name, "n_workers": 1, "parallel": True,
                "description": "GPU batched frame registration (flowreg_hip)"}

    def _get_solver(self, H, W, C, T, fp, term):
        alpha = fp.get("alpha", (2, 2))
        if np.isscalar(alpha):
            alpha = (alpha, alpha)
        a_data = fp.get("a_data", 0.45)
        key = (H, W, C, tuple(float(a) for a in alpha), int(fp.get("update_lag", 5)),
               int(fp.get("iterations", 50)), int(fp.get("min_level", 0)),
               int(fp.get("levels", 50)), float(fp.get("eta", 0.8)),
               float(fp.get("a_smooth", 1.0)),
               tuple(np.broadcast_to(np.asarray(a_data, np.float64), (C,)).tolist()), term)
        # the cached solver is kept for any batch it has room for; a new one
        # gets room for max(T, reserve) frames (at most max_batch), so a
        # recording's first, short batch does not force a second allocation
        need = min(T, self.max_batch)
        if key != self._key or self._solver.max_batch < need:
            self.cleanup()
            nb = min(max(need, int(self.reserve)), self.max_batch)
            self._solver = DisplacementSolver(
                H, W, C, alpha=key[3], update_lag=key[4], iterations=key[5], min_level=key[6],
                levels=key[7], eta=key[8], a_smooth=key[9], a_data=np.array(key[10]),
                max_batch=nb, const_assumption=term)
            self._key = key
        return self._solver

    def _prealigned_flows(self, frames_proc, reference_proc, w_init, fp, term, cc):
        """The cc_initialization steps of parallelization/sequential.py:97-140
        for the whole batch: warp every frame by w_init (linear), estimate
        the rigid residual against the reference by phase correlation, warp
        the frames again by w_init + residual, solve the non-rigid residual
        from zero and return (w_init + residual + flow) in float32."""
        from .xcorr import estimate_rigid_xcorr_batch
        T, H, W, C = frames_proc.shape
        hw, up = cc
        rp = _to_device32(reference_proc).to(frames_proc.device).contiguous()
        wi = _to_device32(w_init).to(frames_proc.device).contiguous()
        partial = warp_frames(frames_proc, wi[None].expand(T, H, W, 2), rp, "linear")
        ref_cc = rp[..., 0] if C == 1 else rp
        mov_cc = partial[..., 0] if C == 1 else partial
        target = (hw, hw) if isinstance(hw, int) else hw
        w_cross = estimate_rigid_xcorr_batch(ref_cc, mov_cc, target_hw=target, up=up, weight=fp.get("weight"))
        w_comb = wi[None] + w_cross[:, None, None, :]
        aligned = warp_frames(frames_proc, w_comb, rp, "linear")
        solver = self._get_solver(H, W, C, T, fp, term)
        solver.set_reference(rp, fp.get("weight"), None)
        w_res = solver.run(aligned.contiguous(), out_dtype=torch.float64)
        return (w_comb.to(torch.float64) + w_res).to(torch.float32)

    def register_device(self, frames: torch.Tensor, frames_proc: torch.Tensor, reference_raw: torch.Tensor,
                        reference_proc, w_init, flow_params: dict, term: str = "gc",
                        interpolation_method: str = "cubic", cc=None, raw_dtype=None):
        """The batch's registration with every array on the device: frames
        (T, H, W, C) raw (float32, or float64 for cv2's 64F warp), frames_proc
        (T, H, W, C) float32 preprocessed, reference_raw (H, W, C) in the
        frames' dtype, reference_proc (H, W, C) and w_init (H, W, 2) (cast to
        float32 as the executors do).  cc = (cc_hw, cc_up) runs the
        cross-correlation prealignment first.  raw_dtype: the raw frames'
        dtype before staging (uint8 / uint16 / int16 frames warp in their
        integer depth, see warp_frames).  Returns (registered float32, flow float32)
        device tensors."""
        T, H, W, C = frames.shape
        fp = flow_params
        if cc is not None:
            if fp.get("gnc_schedule") is not None:
                raise ValueError("hip_batched: cc_initialization with a GNC schedule is not implemented")
            flow = self._prealigned_flows(frames_proc, reference_proc, w_init, fp, term, cc)
        elif fp.get("gnc_schedule") is not None:
            # GNC stages relinearise on the host loop (pyflowreg_amd.stage_loop),
            # one frame at a time, as the sequential executor does
            from .optical_flow import get_displacement
            kw = {k: v for k, v in fp.items() if k != "weight"}
            rp = _to_numpy32(reference_proc)
            wi = _to_numpy32(w_init)
            bp = frames_proc.cpu().numpy()
            flows_np = np.stack([get_displacement(rp, bp[t], weight=fp.get("weight"), uv=wi.copy(), **kw)
                                 for t in range(T)]).astype(np.float32)
            flow = torch.from_numpy(flows_np).to(frames.device)
        else:
            solver = self._get_solver(H, W, C, T, fp, term)
            solver.set_reference(_to_device32(reference_proc), fp.get("weight"), _to_device32(w_init))
            flow = solver.run(frames_proc, out_dtype=torch.float32)
        return warp_frames(frames, flow, reference_raw, interpolation_method, raw_dtype), flow

    def process_batch(self, batch: np.ndarray, batch_proc: np.ndarray, reference_raw: np.ndarray,
                      reference_proc: np.ndarray, w_init: np.ndarray,
                      get_displacement_func: Optional[Callable] = None,
                      imregister_func: Optional[Callable] = None,
                      interpolation_method: str = "cubic",
                      progress_callback: Optional[Callable[[int], None]] = None,
                      **kwargs) -> Tuple[np.ndarray, np.ndarray]:
        """parallelization/base.py:76-110.  Returns (registered (T, H, W, C)
        in batch's dtype, flow (T, H, W, 2) float32)."""
        flow_params_all = dict(kwargs.get("flow_params", {}))
        cc = None
        if bool(flow_params_all.get("cc_initialization", False)):
            cc = (flow_params_all.get("cc_hw", 256), int(flow_params_all.get("cc_up", 1)))
        fp = {k: v for k, v in flow_params_all.items()
              if k not in ("cc_initialization", "cc_hw", "cc_up")}
        term = _resolve_const_assumption(fp.get("const_assumption", "gc"))
        if str(interpolation_method).lower() not in ("cubic", "linear"):
            raise ValueError("Unsupported interpolation method. Use 'linear' or 'cubic'.")
        as4 = lambda x: x[..., None] if x.ndim == 3 else x  # noqa: E731
        as3 = lambda x: x[..., None] if x.ndim == 2 else x  # noqa: E731
        batch = np.ascontiguousarray(as4(np.asarray(batch)))
        remap_depth(batch.dtype)
        bp = np.ascontiguousarray(as4(np.asarray(batch_proc)).astype(np.float32, copy=False))
        rp = np.ascontiguousarray(as3(np.asarray(reference_proc)).astype(np.float32, copy=False))
        wi = np.ascontiguousarray(np.asarray(w_init).astype(np.float32, copy=False))
        dev = _device()
        # cv2.remap works in the raw frames' depth: float64 frames sum in float64
        wdt = np.float64 if batch.dtype == np.float64 else np.float32
        reg, flow = self.register_device(
            torch.from_numpy(batch.astype(wdt, copy=False)).to(dev), torch.from_numpy(bp).to(dev),
            torch.from_numpy(np.asarray(as3(np.asarray(reference_raw)), dtype=wdt)).to(dev), rp, wi, fp, term,
            interpolation_method, cc, batch.dtype)
        registered = reg.cpu().numpy().astype(batch.dtype, copy=False)
        flows = flow.cpu().numpy()
        if progress_callback is not None:
            progress_callback(batch.shape[0])
        return registered, flows


HipBatchedExecutor.register()
