"""Real-time registration on the GPU: ``FlowRegLive``
(motion_correction/flow_reg_live.py) with its state on the device.

Per frame (flow_reg_live.py:216-323):
* normalise with the filtered reference's min / max and apply the 2-D
  Gaussian (``preprocess`` with explicit bounds, float64 result);
* filter the newest frame over the temporal buffer with the causal half
  Gaussian (util/image_processing.py:168-228: float32 weights, float64
  sum from the newest frame back);
* solve get_displacement(reference_proc, frame_proc, uv=last_flow) with a
  one-frame ``DisplacementSolver`` replayed as a hipGraph;
* warp the raw frame with the raw reference as fill (cubic or linear, per
  ``interpolation_method``);
* every ``reference_update_interval`` frames, mix the warped processed frame
  into the processed reference and the warped raw frame into the raw
  reference with weight ``reference_update_weight`` (numpy's weak-scalar
  rule makes ``weight * warped`` a float32 product), then refresh the
  normalisation bounds.

The reference frame is set from one frame, or from several frames by
registering them to their mean with ``compensate_arr`` (balanced quality)
and averaging the registered frames (flow_reg_live.py:111-177).
"""

from __future__ import annotations

from collections import deque
from dataclasses import replace
from typing import Optional, Tuple

import numpy as np
import torch

from .compensate import RegistrationOptions, compensate_arr
from .executor import warp_frames
from .optical_flow import DisplacementSolver, _device, remap_depth
from .preprocess import PRE_FILT32, PRE_NORM32, preprocess


def half_kernel_weights(n_frames: int, sigma_t: float, truncate: float = 4.0) -> np.ndarray:
    """util/image_processing.py:210-217: the causal half Gaussian over the
    newest min(radius + 1, n_frames) frames, float32, newest first."""
    kernel_radius = int(truncate * sigma_t + 0.5)
    kernel_size = min(kernel_radius + 1, n_frames)
    x = np.arange(kernel_size, dtype=np.float32)
    kernel = np.exp(-0.5 * (x / sigma_t) ** 2)
    return kernel / kernel.sum()


class FlowRegLive:
    """flow_reg_live.py:22-360 on the GPU (same constructor arguments and
    methods; options as ``RegistrationOptions`` or an OFOptions-like
    object, with the fast quality setting forced as the reference does)."""

    def __init__(self, options=None, reference_buffer_size: int = 50, reference_update_interval: int = 20,
                 reference_update_weight: float = 0.2, truncate: float = 4.0, **kwargs):
        opt = RegistrationOptions.from_any(options)
        self.options = replace(opt, quality_setting="fast", **kwargs)
        self.truncate = truncate
        self.reference_buffer = deque(maxlen=reference_buffer_size)
        sig = np.asarray(self.options.sigma, dtype=np.float64)
        if sig.ndim == 1:
            self.sigma_2d = sig[:2]
            self.sigma_t = float(sig[2]) if len(sig) > 2 else 0.0
        else:
            self.sigma_2d = sig[0, :2]
            self.sigma_t = float(np.max(sig[:, 2])) if sig.shape[1] > 2 else 0.0
        self.temporal_buffer = deque(maxlen=max(1, int(self.truncate * self.sigma_t + 0.5) + 1))
        self.reference_raw = None    # (H, W, C) on the device, float32 or float64 as given
        self.reference_proc = None   # (H, W, C) float64 on the device
        self.reference_update_interval = reference_update_interval
        self.reference_update_weight = reference_update_weight
        self.frame_count = 0
        self.last_flow = None        # (H, W, 2) float64 on the device
        self.norm_min = None
        self.norm_max = None
        self._solver = None

    # ------------------------------------------------------------ reference
    def set_reference(self, frames: Optional[np.ndarray] = None):
        """flow_reg_live.py:111-177."""
        if frames is None:
            if len(self.reference_buffer) == 0:
                raise ValueError("Reference buffer is empty, cannot initialize reference")
            self.set_reference(np.array(list(self.reference_buffer)))
            return
        frames = np.asarray(frames)
        if frames.ndim == 2:
            ref = frames[..., None].copy()
        elif frames.ndim == 3:
            ref = frames.copy()
        else:
            ref_opt = replace(self.options, quality_setting="balanced")
            initial_ref = np.mean(frames, axis=0)
            registered, _ = compensate_arr(frames, initial_ref, ref_opt)
            ref = np.mean(registered, axis=0)
            ref = ref[..., None] if ref.ndim == 2 else ref
        # the raw reference keeps numpy's dtype: a float32 reference stays
        # float32 through normalize and the reference updates
        if ref.dtype != np.float32:
            ref = ref.astype(np.float64)
        dev = _device()
        self.reference_raw = torch.as_tensor(np.ascontiguousarray(ref), device=dev)
        # normalize (own range) + 2-D Gaussian, float64 result (:179-190)
        lo, den, sem = self._reference_bounds(ref)
        self.reference_proc = preprocess(ref, sigma=self._sigma3(), truncate=self.truncate,
                                         out_dtype=torch.float64, norm_bounds=(lo, den),
                                         semantics=sem)[0].permute(1, 2, 0).contiguous()
        self._update_norm()
        self.last_flow = None
        self.temporal_buffer.clear()

    def reset_reference(self, new_reference: np.ndarray):
        self.set_reference(new_reference)

    def _reference_bounds(self, ref):
        """normalize(reference_raw) (util/image_processing.py:12-98) in the
        reference's dtype: (lo, den) per channel and the float32 semantics
        flags.  A float32 reference is normalised in float32; with "together"
        the float32 result is filtered as a float32 array, with "separate" it
        is stored in a float64 array first."""
        C = ref.shape[2]
        eps = 1e-8
        sep = self.options.channel_normalization == "separate"
        if ref.dtype == np.float32:
            if sep:
                lo = np.array([ref[..., c].min() for c in range(C)], dtype=np.float32)
                hi = np.array([ref[..., c].max() for c in range(C)], dtype=np.float32)
            else:
                lo = np.full(C, ref.min(), dtype=np.float32)
                hi = np.full(C, ref.max(), dtype=np.float32)
            den = (hi - lo) + np.float32(eps)
            sem = PRE_NORM32 if sep else (PRE_NORM32 | PRE_FILT32)
            return lo.astype(np.float64), den.astype(np.float64), sem
        if sep:
            lo = np.array([ref[..., c].min() for c in range(C)])
            hi = np.array([ref[..., c].max() for c in range(C)])
        else:
            lo = np.full(C, ref.min())
            hi = np.full(C, ref.max())
        return lo, (hi - lo) + eps, 0

    def _sigma3(self):
        return (float(self.sigma_2d[0]), float(self.sigma_2d[1]), 0.0)

    def _update_norm(self):
        rp = self.reference_proc
        if self.options.channel_normalization == "separate":
            self.norm_min = [float(rp[..., c].min()) for c in range(rp.shape[2])]
            self.norm_max = [float(rp[..., c].max()) for c in range(rp.shape[2])]
        else:
            self.norm_min = float(rp.min())
            self.norm_max = float(rp.max())

    def _bounds(self, C):
        eps = 1e-8
        if self.options.channel_normalization == "separate":
            lo = np.array([self.norm_min[c] for c in range(C)])
            den = np.array([self.norm_max[c] - self.norm_min[c] + eps for c in range(C)])
        else:
            lo = np.full(C, self.norm_min)
            den = np.full(C, self.norm_max - self.norm_min + eps)
        return lo, den

    def _get_solver(self, H, W, C):
        if self._solver is None or (self._solver.H, self._solver.W, self._solver.C) != (H, W, C):
            fp = self.options.flow_params()
            kw = {k: fp[k] for k in ("alpha", "update_lag", "iterations", "min_level", "levels", "eta",
                                     "a_smooth", "a_data")}
            self._solver = DisplacementSolver(H, W, C, max_batch=1, graph=True,
                                              const_assumption=fp["const_assumption"], **kw)
        return self._solver

    # ---------------------------------------------------------------- frames
    def __call__(self, frame: np.ndarray, normalize: bool = True) -> Tuple[np.ndarray, np.ndarray]:
        """flow_reg_live.py:216-323: (registered (H, W, C) float32, flow
        (H, W, 2) float64)."""
        if self.reference_raw is None:
            self.reference_buffer.append(np.array(frame, copy=True))
            return frame, np.zeros(frame.shape[:2] + (2,), dtype=np.float32)
        frame = np.asarray(frame)
        remap_depth(frame.dtype)
        if frame.ndim == 2:
            frame = frame[..., None]
        self.frame_count += 1
        H, W, C = frame.shape
        dev = self.reference_raw.device
        fr = torch.as_tensor(frame if frame.dtype in (np.float32, np.float64) else frame.astype(np.float64),
                             device=dev)
        filt = preprocess(fr, sigma=self._sigma3(), truncate=self.truncate, out_dtype=torch.float64,
                          norm_bounds=self._bounds(C))[0].permute(1, 2, 0).contiguous()
        self.temporal_buffer.append(filt)
        if self.sigma_t > 0 and len(self.temporal_buffer) > 1:
            k = half_kernel_weights(len(self.temporal_buffer), self.sigma_t, self.truncate)
            proc = torch.zeros_like(filt)
            for i in range(len(k)):
                proc = proc + float(k[i]) * self.temporal_buffer[-(i + 1)]
        else:
            proc = filt
        solver = self._get_solver(H, W, C)
        weight = np.asarray(self.options.weight, dtype=np.float64)
        solver.set_reference(self.reference_proc, weight, self.last_flow)
        flow = solver.run(proc.to(torch.float32)[None], out_dtype=torch.float64)[0]
        self.last_flow = flow
        # cv2.remap keeps 8-bit and 16-bit frames in their depth (fixed-point
        # or float sums, rounded and saturated results) and has no path for
        # other integer dtypes (warp_frames raises for those)
        fwarp = fr.to(torch.float64) if frame.dtype == np.float64 else fr.to(torch.float32)
        interp = self.options.interpolation_method
        registered = warp_frames(fwarp[None], flow[None], self.reference_raw, interp, frame.dtype)[0]
        if self.frame_count % self.reference_update_interval == 0:
            reg_proc = warp_frames(proc[None], flow[None], self.reference_proc, interp)[0]
            wgt = self.reference_update_weight
            # numpy's rules, which torch shares: a Python float times a
            # float32 array is a float32 product, float64 + float32 is float64
            self.reference_proc = (1 - wgt) * self.reference_proc + wgt * reg_proc
            self.reference_raw = (1 - wgt) * self.reference_raw + wgt * registered
            self._update_norm()
        return registered.cpu().numpy(), flow.cpu().numpy()

    def register_frames(self, frames: np.ndarray) -> Tuple[np.ndarray, np.ndarray]:
        """flow_reg_live.py:325-341."""
        frames = np.asarray(frames)
        registered = np.empty_like(frames)
        flows = np.empty(frames.shape[:3] + (2,), dtype=np.float32)
        for i in range(frames.shape[0]):
            r, f = self(frames[i], False)
            registered[i] = r.reshape(registered[i].shape)
            flows[i] = f
        return registered, flows

    def get_current_flow(self) -> Optional[np.ndarray]:
        return self.last_flow.cpu().numpy().copy() if self.last_flow is not None else None

    def set_flow_init(self, w_init: np.ndarray):
        self.last_flow = torch.as_tensor(np.array(w_init, dtype=np.float64), device=_device())
