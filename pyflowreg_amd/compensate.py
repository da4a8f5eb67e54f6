"""In-memory motion compensation on the GPU (``compensate_arr`` equivalent).

Mirrors pyflowreg's ``compensate_arr`` (motion_correction/compensate_arr.py:
16-231) driving ``BatchMotionCorrector`` (motion_correction/
compensate_recording.py):

* reference: the given raw frame, preprocessed with its own range
  (:491-519); per-channel weights normalised to sum 1 (:504-516);
* frames are read in batches of ``buffer_size`` (OF_options.py:457) and
  preprocessed with the raw reference's range (:521-561, :782-784);
* the first batch computes ``w_init`` as the float32 mean flow of its first
  min(22, T) frames registered from a zero initial flow (:645-678);
* every batch is registered with the current ``w_init`` (executor
  ``hip_batched``), then ``w_init`` becomes the float32 mean of the batch's
  last (up to) 20 flows (:786-806);
* per-frame statistics mean/max displacement, mean divergence and mean
  translation are accumulated (:808-821);
* batches are staged in pinned host memory and uploaded on a side stream
  while the previous batch is solved; preprocessing, solving and warping
  stay on the device; ``w_init`` (and, with ``update_reference``, the new
  reference) is formed on the device, so the next batch's GPU work is
  queued before the current batch's results are copied to the host on a
  side stream;
* with ``update_reference`` the processed reference becomes the float64
  mean of the batch's last (up to) 100 preprocessed float64 frames, each
  warped by its flow with the current reference as fill (:680-713,
  :866-867).

The ``w_init`` means are numpy's float32 axis-0 means, bit for bit
(``frh_mean_frames_f32``); the statistics are computed with numpy on the
host, exactly as the reference computes them.
"""

from __future__ import annotations

from dataclasses import dataclass, field
from typing import Callable, Optional, Sequence

import numpy as np
import torch

from .executor import HipBatchedExecutor, mean_frames, warp_frames
from .of_options import normalize_weight
from .optical_flow import _resolve_const_assumption
from .preprocess import _range, preprocess

_QUALITY_MIN_LEVEL = {"quality": 0, "balanced": 4, "fast": 6}
_BUILTIN_BACKENDS = ("flowreg", "flowreg_cuda", "flowreg_torch")
_OUTPUT_TYPES = {"single": np.float32, "double": np.float64, "uint8": np.uint8, "uint16": np.uint16,
                 "int16": np.int16, "int32": np.int32}


@dataclass
class RegistrationOptions:
    """The OFOptions fields the registration path uses, with OFOptions'
    defaults (motion_correction/OF_options.py:401-500)."""
    alpha: Sequence[float] = (1.5, 1.5)
    weight: Sequence[float] = (0.5, 0.5)
    levels: int = 100
    min_level: int = -1
    quality_setting: str = "quality"
    eta: float = 0.8
    update_lag: int = 5
    iterations: int = 50
    a_smooth: float = 1.0
    a_data: float = 0.45
    sigma: Sequence = ((1.0, 1.0, 0.1), (1.0, 1.0, 0.1))
    buffer_size: int = 400
    channel_normalization: str = "together"
    interpolation_method: str = "cubic"
    update_initialization_w: bool = True
    update_reference: bool = False
    constancy_assumption: str = "gc"
    cc_initialization: bool = False
    cc_hw: object = 256
    cc_up: int = 1
    gnc_schedule: Optional[Sequence[float]] = None
    warping_steps: Optional[int] = None
    output_typename: Optional[str] = "double"
    flow_backend: Optional[str] = None
    backend_params: dict = field(default_factory=dict)
    get_displacement_impl: Optional[Callable] = None
    get_displacement_factory: Optional[Callable] = None
    extra: dict = field(default_factory=dict)
    # the weight has had OFOptions' normalisation (from_any of an OFOptions)
    weight_normalized: bool = False

    def __post_init__(self):
        # OF_options.py's validation of the prealignment fields
        if int(self.cc_up) < 1:
            raise ValueError("cc_up must be >= 1")
        if not self.weight_normalized:
            self.weight = normalize_weight(self.weight)
            self.weight_normalized = True

    @property
    def effective_min_level(self) -> int:
        """OF_options.py:649-661."""
        if self.min_level >= 0:
            return self.min_level
        return _QUALITY_MIN_LEVEL.get(str(self.quality_setting).lower(), 0)

    def flow_params(self) -> dict:
        """OF_options.py:1084-1101 (to_dict) for get_displacement."""
        alpha = self.alpha if not np.isscalar(self.alpha) else (self.alpha, self.alpha)
        return {"alpha": tuple(alpha), "levels": self.levels,
                "min_level": self.effective_min_level, "eta": self.eta,
                "iterations": self.iterations, "update_lag": self.update_lag,
                "a_data": self.a_data, "a_smooth": self.a_smooth, "gnc_schedule": self.gnc_schedule,
                "warping_steps": self.warping_steps, "const_assumption": self.constancy_assumption}

    @classmethod
    def from_any(cls, options) -> "RegistrationOptions":
        if options is None:
            return cls()
        if isinstance(options, cls):
            return options
        kw = {}
        for f in ("alpha", "weight", "levels", "min_level", "eta", "update_lag", "iterations",
                  "a_smooth", "a_data", "sigma", "buffer_size", "update_initialization_w",
                  "update_reference", "cc_initialization", "cc_hw", "cc_up", "gnc_schedule",
                  "warping_steps"):
            if hasattr(options, f):
                kw[f] = getattr(options, f)
        for f in ("quality_setting", "channel_normalization", "interpolation_method",
                  "constancy_assumption"):
            if hasattr(options, f):
                v = getattr(options, f)
                kw[f] = getattr(v, "value", v)
        for f in ("output_typename", "get_displacement_impl", "get_displacement_factory"):
            if hasattr(options, f):
                kw[f] = getattr(options, f)
        # OFOptions normalises its weight when it is set (OF_options.py:547-584)
        kw["weight_normalized"] = hasattr(options, "get_weight_at")
        if getattr(options, "backend_params", None):
            kw["backend_params"] = dict(options.backend_params)
        # OFOptions' flow_backend names pyflowreg's own backends ("flowreg",
        # the default, and its CuPy / torch variants): this pipeline is the
        # flowreg_hip backend, so those names select it; any other name must
        # be registered in this package's registry
        fb = getattr(options, "flow_backend", None)
        fb = getattr(fb, "value", fb)
        if fb is not None and fb not in _BUILTIN_BACKENDS:
            kw["flow_backend"] = fb
        if kw.get("channel_normalization") in ("joint",):
            kw["channel_normalization"] = "together"
        # the options' own file and reader features this in-memory pipeline
        # does not have
        for f, default in (("bin_size", 1), ("n_references", 1), ("channel_idx", None)):
            if getattr(options, f, default) != default:
                raise ValueError(f"flowreg_hip compensate_arr: {f}={getattr(options, f)!r} is not implemented")
        return cls(**kw)


def _weight_maps(weight, H, W, C):
    """compensate_recording.py:504-516 with OFOptions.get_weight_at
    (OF_options.py:679-724) on an already normalised weight: the (H, W, C)
    float64 weight of each channel.  A single value applies to every
    channel; more values than channels are cut to C and renormalised; a
    missing channel gets 1/C; an (H, W) map is channel 0's (the others 1/C)
    and an (H, W, C) map gives each channel its plane."""
    w = np.asarray(weight, dtype=np.float64)
    if w.ndim <= 1:
        w = w.reshape(-1)
        if w.size == 1:
            vals = [float(w[0])] * C
        else:
            if w.size > C:
                w = w[:C] / w[:C].sum()
            vals = [float(w[c]) if c < w.size else 1.0 / C for c in range(C)]
        return np.ones((H, W, C)) * np.asarray(vals).reshape(1, 1, -1)
    out = np.empty((H, W, C))
    if w.ndim == 2:
        for c in range(C):
            out[:, :, c] = w if c == 0 else 1.0 / C
        return out
    if w.ndim == 3:
        for c in range(C):
            out[:, :, c] = w[:, :, c] if c < w.shape[2] else 1.0 / C
        return out
    raise ValueError(f"Unexpected weight array with {w.ndim} dimensions")


def _sigma_rows(sigma):
    s = np.asarray(sigma, dtype=np.float64)
    return s if s.ndim == 2 else s[None]


_STAT_WORKERS = 8
_STAT_POOL = None


def _stat_pool():
    global _STAT_POOL
    if _STAT_POOL is None:
        from concurrent.futures import ThreadPoolExecutor
        _STAT_POOL = ThreadPoolExecutor(max_workers=_STAT_WORKERS)
    return _STAT_POOL


def _parallel_copy(dst: np.ndarray, src: np.ndarray) -> None:
    """np.copyto(dst, src, casting="unsafe") in frame chunks on the worker
    threads (numpy releases the GIL while it copies)."""
    n = dst.shape[0]
    step = max(1, -(-n // _STAT_WORKERS))
    list(_stat_pool().map(lambda lo: np.copyto(dst[lo:lo + step], src[lo:lo + step], casting="unsafe"),
                          range(0, n, step)))


class BatchStager:
    """Pinned host staging of raw frame batches (the batches read_batch
    returns, util/io/_base.py:283-313) with asynchronous uploads on a side
    stream: two pinned-host / device buffer pairs alternate, so the host copy
    and upload of the next batch overlap the GPU work on the current one.
    Frames are staged as float64 when they are float64, else as float32."""

    def __init__(self, device):
        self.device = device
        self.stream = torch.cuda.Stream(device=device)
        self.slots = [None, None]
        self.turn = 0

    def upload(self, frames: np.ndarray) -> int:
        frames = np.asarray(frames)
        dt = np.float64 if frames.dtype == np.float64 else np.float32
        key = (frames.shape, dt)
        k = self.turn
        self.turn ^= 1
        slot = self.slots[k]
        if slot is None or slot["key"] != key:
            if slot is not None:  # nothing may still read or write the buffers it replaces
                for ev in (slot["up"], slot["done"]):
                    if ev is not None:
                        ev.synchronize()
            tdt = torch.float64 if dt == np.float64 else torch.float32
            # allocated from the side stream's pool: a block that kernels on the
            # main stream freed but may still be using is never handed out here
            with torch.cuda.stream(self.stream):
                dev = torch.empty(frames.shape, dtype=tdt, device=self.device)
            slot = {"key": key, "host": torch.empty(frames.shape, dtype=tdt, pin_memory=True),
                    "dev": dev, "up": None, "done": None}
            self.slots[k] = slot
        if slot["up"] is not None:
            slot["up"].synchronize()  # the last upload from this host buffer has left it
        np.copyto(slot["host"].numpy(), frames, casting="unsafe")
        with torch.cuda.stream(self.stream):
            if slot["done"] is not None:
                self.stream.wait_event(slot["done"])  # the last batch in this device buffer is used up
            slot["dev"].copy_(slot["host"], non_blocking=True)
            ev = torch.cuda.Event()
            ev.record(self.stream)
            slot["up"] = ev
        return k

    def acquire(self, k: int) -> torch.Tensor:
        slot = self.slots[k]
        torch.cuda.current_stream(self.device).wait_event(slot["up"])
        return slot["dev"]

    def release(self, k: int) -> None:
        ev = torch.cuda.Event()
        ev.record(torch.cuda.current_stream(self.device))
        self.slots[k]["done"] = ev


class MotionCorrector:
    """Device-resident BatchMotionCorrector state for one recording: a batch
    is staged and uploaded once, preprocessed, solved and warped on the GPU,
    and only its outputs come back to the host (where numpy forms the means
    and statistics exactly as the reference does)."""

    def __init__(self, reference, options=None, max_batch: int = 256, get_disp: Optional[Callable] = None):
        self.opt = RegistrationOptions.from_any(options)
        # None: this backend's batched solver; else a displacement callable
        # called frame by frame (see _register_custom)
        self.get_disp = get_disp
        ref = np.asarray(reference, dtype=np.float64)
        self.reference_raw = ref[..., None] if ref.ndim == 2 else ref
        H, W, C = self.reference_raw.shape
        self.shape = (H, W, C)
        self.weight = _weight_maps(self.opt.weight, H, W, C)
        self.sigma = _sigma_rows(self.opt.sigma)
        # float64, as the reference keeps it; the executor casts it to float32
        self.reference_proc = self._preprocess(self.reference_raw, None, torch.float64)[0].permute(1, 2, 0)
        self.executor = HipBatchedExecutor(max_batch=max_batch)
        self.executor.reserve = min(int(self.opt.buffer_size), max_batch)  # one solver for every batch
        self.device = self.reference_proc.device
        self.stager = BatchStager(self.device)
        self._ref_raw_dev = {}
        self.w_init = None  # (H, W, 2) float32 device tensor once the first batch is queued
        self.copy_stream = torch.cuda.Stream(device=self.device)
        self._pin = {}  # pinned host buffers of the outputs, reused batch after batch
        self._norm_bounds = None
        self.mean_disp, self.max_disp, self.mean_div, self.mean_translation = [], [], [], []
        if str(self.opt.interpolation_method).lower() not in ("cubic", "linear"):
            raise ValueError("Unsupported interpolation method. Use 'linear' or 'cubic'.")

    def _preprocess(self, frames, norm_ref, out_dtype=torch.float32):
        if norm_ref is self.reference_raw:
            # the raw reference's range, formed once per recording
            if self._norm_bounds is None:
                self._norm_bounds = _range(self.reference_raw, self.shape[2], self.opt.channel_normalization,
                                           1e-8)
            return preprocess(frames, sigma=self.sigma, channel_normalization=self.opt.channel_normalization,
                              out_dtype=out_dtype, norm_bounds=self._norm_bounds)
        return preprocess(frames, normalization_ref=norm_ref, sigma=self.sigma,
                          channel_normalization=self.opt.channel_normalization, out_dtype=out_dtype)

    def _flow_params(self):
        fp = self.opt.flow_params()
        H, W, C = self.shape
        fp["weight"] = self.weight
        return fp

    def _reference_raw(self, dtype):
        if dtype not in self._ref_raw_dev:
            self._ref_raw_dev[dtype] = torch.as_tensor(self.reference_raw, dtype=dtype, device=self.device)
        return self._ref_raw_dev[dtype]

    def _register(self, frames, frames_proc, w_init, prealign=True, raw_dtype=None, with_uv=True):
        fp = self._flow_params()
        if self.get_disp is not None:
            return self._register_custom(frames, frames_proc, w_init, fp, raw_dtype, with_uv)
        term = _resolve_const_assumption(fp.get("const_assumption", "gc"))
        cc = (self.opt.cc_hw, int(self.opt.cc_up)) if self.opt.cc_initialization and prealign else None
        return self.executor.register_device(frames, frames_proc, self._reference_raw(frames.dtype),
                                             self.reference_proc, w_init, fp, term,
                                             self.opt.interpolation_method, cc, raw_dtype)

    def _register_custom(self, frames, frames_proc, w_init, fp, raw_dtype, with_uv):
        """A displacement callable that is not this backend's solver
        (get_displacement / get_displacement_factory / another registered
        backend): called frame by frame on host arrays as the sequential
        executor calls it (parallelization/sequential.py:141-145, float32
        inputs, uv = a copy of w_init, the flow cast to float32); the
        registration warp stays on the GPU.  with_uv=False is the reference's
        serial w_init path for at most 4 frames (compensate_recording.py:
        664-671), which passes no uv."""
        if self.opt.cc_initialization:
            raise ValueError("compensate_arr: cc_initialization with a custom displacement callable is not "
                             "implemented by the flowreg_hip pipeline")
        rp = np.ascontiguousarray(self.reference_proc.to(torch.float32).cpu().numpy())
        bp = frames_proc.cpu().numpy()
        wi = w_init.cpu().numpy() if isinstance(w_init, torch.Tensor) else np.asarray(w_init)
        wi = np.ascontiguousarray(wi.astype(np.float32, copy=False))
        flows = []
        for t in range(bp.shape[0]):
            if with_uv:
                f = self.get_disp(rp, bp[t], uv=wi.copy(), **fp)
            else:
                f = self.get_disp(rp, bp[t], **fp)
            flows.append(np.asarray(f).astype(np.float32, copy=False))
        flow = torch.from_numpy(np.stack(flows)).to(frames.device)
        reg = warp_frames(frames, flow, self._reference_raw(frames.dtype), self.opt.interpolation_method,
                          raw_dtype)
        return reg, flow

    def _launch(self, frames, raw_dtype=None):
        """Queue the GPU work of one batch (T, H, W, C) on the device:
        preprocess, w_init on the first batch, the batch's flows and
        registered frames, then the next batch's w_init (and reference, with
        update_reference).  raw_dtype: the host batch's dtype (the staged
        frames are float32 or float64).  Returns (registered, flow, event
        recorded after the batch's work)."""
        if self.opt.update_reference:
            bp64 = self._preprocess(frames, self.reference_raw, torch.float64)
            bp = bp64.to(torch.float32)  # the same rounding as the float32 prefilter output
        else:
            bp64 = None
            bp = self._preprocess(frames, self.reference_raw)
        bp = bp.permute(0, 2, 3, 1)
        H, W, _ = self.shape
        if self.w_init is None:
            n_init = min(22, frames.shape[0])
            # compensate_recording.py:645-678: more than 4 frames go through
            # the executor (with the prealignment), fewer through
            # get_displacement directly
            _, w0 = self._register(frames[:n_init], bp[:n_init], np.zeros((H, W, 2)), prealign=n_init > 4,
                                   raw_dtype=raw_dtype, with_uv=n_init > 4)
            self.w_init = mean_frames(w0)
        cur = self.w_init if self.opt.update_initialization_w else torch.zeros_like(self.w_init)
        reg, flow = self._register(frames, bp, cur, raw_dtype=raw_dtype)
        # (T, H, W, C) in memory here, so that the copy to the host is one DMA
        reg = reg.contiguous()
        # compensate_recording.py:801-806 and :866-867, on the device
        if self.opt.update_initialization_w:
            self.w_init = mean_frames(flow[-20:] if flow.shape[0] > 20 else flow)
        if self.opt.update_reference:
            self._update_reference(bp64, flow)
        done = torch.cuda.Event()
        done.record(torch.cuda.current_stream(self.device))
        return reg, flow, done

    def _pinned(self, name, shape, dtype):
        key = (tuple(shape), dtype)
        buf = self._pin.get(name)
        if buf is None or buf[0] != key:
            buf = (key, torch.empty(tuple(shape), dtype=dtype, pin_memory=True))
            self._pin[name] = buf
        return buf[1]

    def _finish(self, dtype, reg, flow, done, out_reg=None, out_w=None):
        """Copy one batch's outputs to the host (into out_reg / out_w when
        given): DMA into pinned buffers on the copy stream after the batch's
        work (event done), while the next batch's work, if queued, keeps
        running; then host threads copy them into the result arrays (the
        registered frames cast to dtype with numpy's unsafe cast, as the
        reference's array assignment does)."""
        with torch.cuda.stream(self.copy_stream):
            self.copy_stream.wait_event(done)
            hr = self._pinned("reg", reg.shape, reg.dtype)
            hw = self._pinned("flow", flow.shape, flow.dtype)
            hr.copy_(reg, non_blocking=True)
            hw.copy_(flow, non_blocking=True)
            ev = torch.cuda.Event()
            ev.record(self.copy_stream)
        ev.synchronize()
        registered = out_reg if out_reg is not None else np.empty(tuple(reg.shape), dtype=dtype)
        w = out_w if out_w is not None else np.empty(tuple(flow.shape), dtype=np.float32)
        _parallel_copy(registered, hr.numpy())
        _parallel_copy(w, hw.numpy())
        return registered, w

    def _statistics(self, w):
        """compensate_recording.py:808-821, per frame with numpy (frame
        chunks on a thread pool: each frame's reductions are its own, so the
        values are those of the serial loop)."""
        def chunk(lo, hi):
            ww = w[lo:hi]
            mag = np.sqrt(ww[:, :, :, 0] ** 2 + ww[:, :, :, 1] ** 2)
            md = np.mean(mag, axis=(1, 2)).tolist()
            mx = np.max(mag, axis=(1, 2)).tolist()
            div, tr = [], []
            for t in range(ww.shape[0]):
                du_dx = np.gradient(ww[t, :, :, 0], axis=1)
                dv_dy = np.gradient(ww[t, :, :, 1], axis=0)
                div.append(float(np.mean(du_dx + dv_dy)))
                um = float(np.mean(ww[t, :, :, 0]))
                vm = float(np.mean(ww[t, :, :, 1]))
                tr.append(float(np.sqrt(um ** 2 + vm ** 2)))
            return md, mx, div, tr
        T = w.shape[0]
        step = max(1, -(-T // _STAT_WORKERS))
        parts = list(_stat_pool().map(lambda lo: chunk(lo, min(T, lo + step)), range(0, T, step)))
        for md, mx, div, tr in parts:
            self.mean_disp.extend(md)
            self.max_disp.extend(mx)
            self.mean_div.extend(div)
            self.mean_translation.extend(tr)

    def process(self, batch):
        """Register one batch (T, H, W, C); returns (registered, w)."""
        batch = np.asarray(batch)
        batch = batch[..., None] if batch.ndim == 3 else batch
        k = self.stager.upload(batch)
        state = self._launch(self.stager.acquire(k), batch.dtype)
        self.stager.release(k)
        registered, w = self._finish(batch.dtype, *state)
        self._statistics(w)
        return registered, w

    def _update_reference(self, bp64, w):
        """compensate_recording.py:680-713: the processed reference becomes
        the mean of the last min(100, T) preprocessed float64 frames, each
        warped (interpolation_method, float64 images: cv2.remap's 64F path,
        float32 result)
        by its flow with the current processed reference as fill; the mean
        is float64, summed frame by frame as numpy's axis-0 mean does."""
        T = bp64.shape[0]
        n = min(100, T)
        if n < 1:
            return
        frames = bp64[T - n:].permute(0, 2, 3, 1)
        flows = w[T - n:]
        reg = warp_frames(frames, flows, self.reference_proc, self.opt.interpolation_method)
        acc = reg[0].to(torch.float64)
        for k in range(1, n):
            acc = acc + reg[k].to(torch.float64)
        self.reference_proc = (acc / n).contiguous()

    def close(self):
        self.executor.cleanup()


def _resolve_displacement(opt: RegistrationOptions):
    """OFOptions.resolve_get_displacement (OF_options.py:1021-1082):
    get_displacement_impl, else get_displacement_factory(**backend_params),
    else the registered backend's factory(**backend_params).  Returns None
    when the result is this backend's own get_displacement (the batched
    GPU solver then registers whole batches), else the callable."""
    from .backend_registry import get_backend
    from .optical_flow import get_displacement as own
    if opt.get_displacement_impl is not None:
        fn = opt.get_displacement_impl
    elif opt.get_displacement_factory is not None:
        fn = opt.get_displacement_factory(**opt.backend_params)
    else:
        fn = get_backend(opt.flow_backend or "flowreg_hip")(**opt.backend_params)
    return None if fn is own else fn


def _resolve_executor(opt: RegistrationOptions, registration_config) -> str:
    """compensate_recording.py:200-267 for this pipeline: the requested
    parallelization must be one the backend supports, else a warning names
    the fallback; None selects hip_batched (the reference auto-selects
    among its built-in executors; flowreg_hip declares sequential and
    hip_batched, which give identical results here)."""
    import warnings
    from .backend_registry import get_backend_executors
    backend = opt.flow_backend or "flowreg_hip"
    supported = get_backend_executors(backend)
    name = getattr(registration_config, "parallelization", None)
    if name is None and isinstance(registration_config, dict):
        name = registration_config.get("parallelization")
    if name is None:
        return "hip_batched" if "hip_batched" in supported else "sequential"
    if name in supported and name in ("sequential", "hip_batched"):
        return name
    reason = (f"Backend '{backend}' does not support '{name}' executor" if name not in supported
              else f"Executor '{name}' is not available on this system")
    warnings.warn(f"{reason}. Supported executors: {sorted(supported)}. Falling back to 'sequential'.")
    return "sequential"


def _fire(callbacks, what, *args):
    """compensate_recording.py:823-867: a callback's exception becomes a
    warning and the run goes on."""
    import warnings
    for cb in callbacks:
        try:
            cb(*args)
        except Exception as e:  # noqa: BLE001 -- the reference catches every exception here
            warnings.warn(f"{what} callback error: {e}")


def compensate_arr(c1, c_ref, options=None,
                   progress_callback: Optional[Callable[[int, int], None]] = None, *,
                   flow_backend: Optional[str] = None,
                   backend_params: Optional[dict] = None,
                   get_displacement: Optional[Callable] = None,
                   get_displacement_factory: Optional[Callable[..., Callable]] = None,
                   w_callback: Optional[Callable[[np.ndarray, int, int], None]] = None,
                   registered_callback: Optional[Callable[[np.ndarray, int, int], None]] = None,
                   registration_config=None,
                   max_batch: int = 256):
    """compensate_arr (motion_correction/compensate_arr.py:16-231) on the
    GPU, with the reference's arguments:

    * shapes: (T, H, W, C); (T, H, W) with a 2-D c_ref (returned as
      (T, H, W)); (H, W) (returned as (H, W), the flow as (H, W, 2));
      (H, W, C) with a 3-D c_ref is one frame (returned as (1, H, W, C)),
      unless its last axis is longer than 4 (then (T, H, W), one channel);
    * the registered frames are cast per ``options.output_typename``
      (default "double": float64; "single", "uint8", "uint16", "int16",
      "int32"; other values keep the frames' dtype);
    * flow_backend / backend_params / get_displacement /
      get_displacement_factory override the options' displacement callable
      (resolved as OFOptions.resolve_get_displacement does).  This
      backend's own get_displacement runs batched on the GPU; any other
      callable is called frame by frame, with the warp on the GPU;
    * progress_callback(current, total) after each batch (per frame with
      registration_config.parallelization = "sequential"),
      registered_callback(batch, start, end) and w_callback(w, start, end)
      after it, in that order, as compensate_recording.py:786-867 fires
      them; a raising callback is turned into a warning;
    * registration_config: RegistrationConfig (or a dict) whose
      parallelization names the executor ("hip_batched", default, or
      "sequential"; others fall back to "sequential" with a warning).

    Returns (registered, w float32 (T, H, W, 2))."""
    c1 = np.asarray(c1)
    c_ref = np.asarray(c_ref)
    if c1.size == 0:
        raise ValueError("Input array cannot be empty")
    opt = RegistrationOptions.from_any(options)
    over = {}
    if flow_backend is not None:
        # the reference's own backend names select this backend, as in the options
        over["flow_backend"] = None if flow_backend in _BUILTIN_BACKENDS else flow_backend
    if backend_params is not None:
        over["backend_params"] = {**opt.backend_params, **dict(backend_params)}
    if get_displacement is not None:
        over["get_displacement_impl"] = get_displacement
    if get_displacement_factory is not None:
        over["get_displacement_factory"] = get_displacement_factory
    if over:
        from dataclasses import replace as _replace
        opt = _replace(opt, **over)
    get_disp = _resolve_displacement(opt)
    executor = _resolve_executor(opt, registration_config)
    orig_ndim = c1.ndim
    squeeze = None
    if c1.ndim == 3 and c_ref.ndim == 2:
        video, ref, squeeze = c1[..., None], c_ref[..., None], "channel"
    elif c1.ndim == 2:
        video, ref, squeeze = c1[None, :, :, None], (c_ref[..., None] if c_ref.ndim == 2 else c_ref), "all"
    elif c1.ndim == 3:
        video = c1[None] if c1.shape[-1] <= 4 else c1[..., None]
        ref = c_ref
    else:
        video, ref = c1, c_ref
    if video.ndim != 4:
        raise ValueError(f"compensate_arr: cannot read frames of shape {c1.shape}")
    mc = MotionCorrector(ref, opt, max_batch=max_batch, get_disp=get_disp)
    T = video.shape[0]
    bs = int(mc.opt.buffer_size)
    registered = np.empty(video.shape, dtype=video.dtype)
    flows = np.empty(video.shape[:3] + (2,), dtype=np.float32)
    progress = [progress_callback] if progress_callback is not None else []
    try:
        k = mc.stager.upload(video[0:bs])
        state = mc._launch(mc.stager.acquire(k), video.dtype)
        mc.stager.release(k)
        for s in range(0, T, bs):
            e = min(T, s + bs)
            nxt = None
            if e < T:
                # stage the next batch and queue its GPU work (its w_init is
                # already on the device) before this batch's outputs are
                # copied out and its statistics computed on the host
                k = mc.stager.upload(video[e:e + bs])
                nxt = mc._launch(mc.stager.acquire(k), video.dtype)
                mc.stager.release(k)
            _, w = mc._finish(video.dtype, *state, out_reg=registered[s:e], out_w=flows[s:e])
            state = nxt
            if executor == "sequential":
                for cur in range(s + 1, e + 1):
                    _fire(progress, "Progress", cur, T)
            else:
                _fire(progress, "Progress", e, T)
            mc._statistics(w)
            _fire([registered_callback] if registered_callback is not None else [], "Registered frames",
                  registered[s:e], s, e)
            _fire([w_callback] if w_callback is not None else [], "Displacement field", flows[s:e], s, e)
    finally:
        mc.close()
    if opt.output_typename in _OUTPUT_TYPES:
        registered = registered.astype(_OUTPUT_TYPES[opt.output_typename])
    if squeeze == "all":
        registered = np.squeeze(registered)
        flows = np.squeeze(flows, axis=0)
    elif squeeze == "channel" and orig_ndim == 3:
        registered = np.squeeze(registered, axis=-1)
    return registered, flows
