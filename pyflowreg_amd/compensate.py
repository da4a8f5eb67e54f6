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
* with ``update_reference`` the processed reference becomes the float64
  mean of the batch's last (up to) 100 preprocessed float64 frames, each
  warped by its flow with the current reference as fill (:680-713,
  :866-867).

Means and statistics are computed with numpy on the host, exactly as the
reference computes them, so that ``w_init`` is bit-identical.
"""

from __future__ import annotations

from dataclasses import dataclass, field
from typing import Callable, Optional, Sequence

import numpy as np
import torch

from .executor import HipBatchedExecutor, warp_frames
from .preprocess import preprocess

_QUALITY_MIN_LEVEL = {"quality": 0, "balanced": 4, "fast": 6}


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
    extra: dict = field(default_factory=dict)

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
                "a_data": self.a_data, "a_smooth": self.a_smooth, "gnc_schedule": None,
                "warping_steps": None, "const_assumption": self.constancy_assumption}

    @classmethod
    def from_any(cls, options) -> "RegistrationOptions":
        if options is None:
            return cls()
        if isinstance(options, cls):
            return options
        kw = {}
        for f in ("alpha", "weight", "levels", "min_level", "eta", "update_lag", "iterations",
                  "a_smooth", "a_data", "sigma", "buffer_size", "update_initialization_w",
                  "update_reference"):
            if hasattr(options, f):
                kw[f] = getattr(options, f)
        for f in ("quality_setting", "channel_normalization", "interpolation_method",
                  "constancy_assumption"):
            if hasattr(options, f):
                v = getattr(options, f)
                kw[f] = getattr(v, "value", v)
        if kw.get("channel_normalization") in ("joint",):
            kw["channel_normalization"] = "together"
        return cls(**kw)


def _channel_weights(weight, C):
    """compensate_recording.py:504-516: per-channel constants summing to 1."""
    w1 = np.asarray(weight, dtype=np.float64).reshape(-1)
    s = w1.sum()
    if s > 0:
        w1 = w1 / s
    return np.array([w1[c] if c < len(w1) else 1.0 / C for c in range(C)])


def _sigma_rows(sigma):
    s = np.asarray(sigma, dtype=np.float64)
    return s if s.ndim == 2 else s[None]


class MotionCorrector:
    """Device-resident BatchMotionCorrector state for one recording."""

    def __init__(self, reference, options=None, max_batch: int = 256):
        self.opt = RegistrationOptions.from_any(options)
        ref = np.asarray(reference, dtype=np.float64)
        self.reference_raw = ref[..., None] if ref.ndim == 2 else ref
        H, W, C = self.reference_raw.shape
        self.shape = (H, W, C)
        self.weight = _channel_weights(self.opt.weight, C)
        self.sigma = _sigma_rows(self.opt.sigma)
        # float64, as the reference keeps it; the executor casts it to float32
        self.reference_proc = self._preprocess(self.reference_raw, None, torch.float64)[0].permute(1, 2, 0)
        self.executor = HipBatchedExecutor(max_batch=max_batch)
        self.w_init = None
        self.mean_disp, self.max_disp, self.mean_div, self.mean_translation = [], [], [], []

    def _preprocess(self, frames, norm_ref, out_dtype=torch.float32):
        return preprocess(frames, normalization_ref=norm_ref, sigma=self.sigma[0],
                          channel_normalization=self.opt.channel_normalization, out_dtype=out_dtype)

    def _flow_params(self):
        fp = self.opt.flow_params()
        H, W, C = self.shape
        fp["weight"] = np.ones((H, W, C)) * self.weight.reshape(1, 1, -1)
        return fp

    def _register(self, batch, batch_proc, w_init):
        return self.executor.process_batch(
            batch, batch_proc.permute(0, 2, 3, 1).cpu().numpy(), self.reference_raw,
            self.reference_proc.cpu().numpy(), w_init, interpolation_method=self.opt.interpolation_method,
            flow_params=self._flow_params())

    def process(self, batch):
        """Register one batch (T, H, W, C); returns (registered, w)."""
        batch = np.asarray(batch)
        batch = batch[..., None] if batch.ndim == 3 else batch
        if self.opt.update_reference:
            bp64 = self._preprocess(batch, self.reference_raw, torch.float64)
            bp = bp64.to(torch.float32)  # the same rounding as the float32 prefilter output
        else:
            bp = self._preprocess(batch, self.reference_raw)
        H, W, _ = self.shape
        if self.w_init is None:
            n_init = min(22, batch.shape[0])
            _, w0 = self._register(batch[:n_init], bp[:n_init], np.zeros((H, W, 2)))
            self.w_init = np.mean(w0, axis=0)
        cur = self.w_init if self.opt.update_initialization_w else np.zeros_like(self.w_init)
        registered, w = self._register(batch, bp, cur)
        if self.opt.update_initialization_w:
            self.w_init = np.mean(w[-20:], axis=0) if w.shape[0] > 20 else np.mean(w, axis=0)
        mag = np.sqrt(w[:, :, :, 0] ** 2 + w[:, :, :, 1] ** 2)
        self.mean_disp.extend(np.mean(mag, axis=(1, 2)).tolist())
        self.max_disp.extend(np.max(mag, axis=(1, 2)).tolist())
        for t in range(w.shape[0]):
            du_dx = np.gradient(w[t, :, :, 0], axis=1)
            dv_dy = np.gradient(w[t, :, :, 1], axis=0)
            self.mean_div.append(float(np.mean(du_dx + dv_dy)))
            um = float(np.mean(w[t, :, :, 0]))
            vm = float(np.mean(w[t, :, :, 1]))
            self.mean_translation.append(float(np.sqrt(um ** 2 + vm ** 2)))
        if self.opt.update_reference:
            self._update_reference(bp64, w)
        return registered, w

    def _update_reference(self, bp64, w):
        """compensate_recording.py:680-713: the processed reference becomes
        the mean of the last min(100, T) preprocessed float64 frames, each
        warped (cubic, float64 images: cv2.remap's 64F path, float32 result)
        by its flow with the current processed reference as fill; the mean
        is float64, summed frame by frame as numpy's axis-0 mean does."""
        if str(self.opt.interpolation_method).lower() != "cubic":
            raise ValueError("flowreg_hip: update_reference supports cubic warps only")
        T = bp64.shape[0]
        n = min(100, T)
        if n < 1:
            return
        frames = bp64[T - n:].permute(0, 2, 3, 1)
        flows = torch.from_numpy(np.ascontiguousarray(w[T - n:])).to(frames.device)
        reg = warp_frames(frames, flows, self.reference_proc)
        acc = reg[0].to(torch.float64)
        for k in range(1, n):
            acc = acc + reg[k].to(torch.float64)
        self.reference_proc = (acc / n).contiguous()

    def close(self):
        self.executor.cleanup()


def compensate_arr(c1, c_ref, options=None,
                   progress_callback: Optional[Callable[[int, int], None]] = None,
                   max_batch: int = 256):
    """compensate_arr (motion_correction/compensate_arr.py:16-231) on the
    GPU: returns (registered (T, H, W[, C]) in c1's dtype, w (T, H, W, 2)
    float32)."""
    c1 = np.asarray(c1)
    squeeze = c1.ndim == 3 and np.asarray(c_ref).ndim == 2
    video = c1[..., None] if squeeze else c1
    if video.ndim == 3:
        video = video[None]
    mc = MotionCorrector(c_ref, options, max_batch=max_batch)
    T = video.shape[0]
    bs = int(mc.opt.buffer_size)
    regs, ws = [], []
    try:
        for s in range(0, T, bs):
            r, w = mc.process(video[s:s + bs])
            regs.append(r)
            ws.append(w)
            if progress_callback is not None:
                progress_callback(min(T, s + bs), T)
    finally:
        mc.close()
    registered = np.concatenate(regs)
    if squeeze:
        registered = registered[..., 0]
    return registered, np.concatenate(ws)
