"""Frame preprocessing on the GPU: normalize, then Gaussian prefilter.

Mirrors ``BatchMotionCorrector._preprocess_frames``
(motion_correction/compensate_recording.py:521-561): ``normalize``
(util/image_processing.py:12-98) with the raw reference's range, then
``apply_gaussian_filter`` (util/image_processing.py:101-165, scipy
``gaussian_filter`` with mode "reflect", truncate 4), then the executors'
float32 cast (motion_correction/parallelization/base.py:63-73).
"""

from __future__ import annotations

import numpy as np
import torch

from ._lib import FlowregHipError, check, lib, ptr, scratch, stream_handle

PRE_NORM32 = 1  # FRH_PRE_NORM32
PRE_FILT32 = 2  # FRH_PRE_FILT32


def gaussian_weights(sigma: float, truncate: float = 4.0):
    """The 1-D kernel scipy.ndimage.gaussian_filter1d correlates with
    (order 0): radius int(truncate * sigma + 0.5), normalised exp."""
    sd = float(sigma)
    radius = int(truncate * sd + 0.5)
    x = np.arange(-radius, radius + 1)
    phi = np.exp(-0.5 / (sd * sd) * x ** 2)
    phi = phi / phi.sum()
    return radius, np.ascontiguousarray(phi[::-1], dtype=np.float64)


def _range(src, C, channel_normalization, eps):
    """Per-channel (lo, den) of normalize(): den = (max - min) + eps."""
    a = src if isinstance(src, torch.Tensor) else torch.as_tensor(np.asarray(src))
    lo = np.empty(C)
    den = np.empty(C)
    if channel_normalization == "separate" and a.ndim >= 3:
        for c in range(C):
            mn = float(a[..., c].min())
            mx = float(a[..., c].max())
            lo[c], den[c] = mn, (mx - mn) + eps
    else:
        mn = float(a.min())
        mx = float(a.max())
        lo[:] = mn
        den[:] = (mx - mn) + eps
    return lo, den


def preprocess(frames, normalization_ref=None, sigma=(1.0, 1.0, 0.1),
               channel_normalization="together", eps=1e-8, truncate=4.0,
               out_dtype=torch.float32, norm_bounds=None, semantics=0) -> torch.Tensor:
    """frames: (H, W, C) or (T, H, W, C) numpy/tensor (float32 or float64).
    Returns CUDA planes (T, C, H, W) (T = 1 for a single frame): float32, the
    executors' cast, or float64 (``out_dtype``), the preprocessed frames
    before that cast as update_reference and FlowRegLive use them.
    ``norm_bounds`` = (lo, den) per channel replaces the range of
    ``normalization_ref``: x -> (x - lo) / den (FlowRegLive normalises with
    its filtered reference's min and max, flow_reg_live.py:192-214);
    ``semantics`` takes FRH_PRE_NORM32 / FRH_PRE_FILT32 for a float32 frame
    normalised and filtered in float32 as numpy and scipy do."""
    if not torch.cuda.is_available():
        raise FlowregHipError("flowreg_hip needs a ROCm GPU; none is visible")
    dev = torch.device("cuda", torch.cuda.current_device())
    x = frames if isinstance(frames, torch.Tensor) else torch.from_numpy(np.ascontiguousarray(frames))
    spatial_only = x.ndim == 3
    if x.ndim == 3:
        x = x[None]
    if x.ndim != 4:
        raise ValueError("preprocess expects (H, W, C) or (T, H, W, C) frames")
    f64 = x.dtype == torch.float64
    x = x.to(device=dev, dtype=torch.float64 if f64 else torch.float32).contiguous()
    T, H, W, C = x.shape
    if norm_bounds is not None:
        lo = np.ascontiguousarray(np.broadcast_to(np.asarray(norm_bounds[0], dtype=np.float64), (C,)))
        den = np.ascontiguousarray(np.broadcast_to(np.asarray(norm_bounds[1], dtype=np.float64), (C,)))
    else:
        ref = normalization_ref if normalization_ref is not None else frames
        lo, den = _range(ref, C, channel_normalization, eps)
    sig = np.asarray(sigma, dtype=np.float64)
    out64 = out_dtype == torch.float64
    if sig.ndim == 2 and not np.all(sig == sig[0]):
        # per-channel sigmas (row min(c, rows - 1) for channel c): each
        # channel filtered on its own with its range
        out = torch.empty((T, C, H, W), dtype=torch.float64 if out64 else torch.float32, device=dev)
        for c in range(C):
            xc = x[..., c:c + 1].contiguous()
            out[:, c] = preprocess(xc if not spatial_only else xc[0], sigma=sig[min(c, len(sig) - 1)],
                                   truncate=truncate, out_dtype=out_dtype, norm_bounds=(lo[c:c + 1], den[c:c + 1]),
                                   semantics=semantics)[:, 0]
        return out
    if sig.ndim == 2:
        sig = sig[0]
    sy, sx = float(sig[0]), float(sig[1])
    st = float(sig[2]) if (len(sig) > 2 and not spatial_only) else 0.0
    ry, wy = gaussian_weights(sy, truncate) if sy > 1e-15 else (0, np.ones(1))
    rx, wx = gaussian_weights(sx, truncate) if sx > 1e-15 else (0, np.ones(1))
    rt, wt = gaussian_weights(st, truncate) if st > 1e-15 else (0, np.ones(1))
    out = torch.empty((T, C, H, W), dtype=torch.float64 if out64 else torch.float32, device=dev)
    if rt > 0:
        # scipy filters the frame axis first (apply_gaussian_filter's 4-D branch)
        if semantics:
            raise ValueError("flowreg_hip: float32 prefilter semantics with a temporal kernel are not implemented")
        nb = lib().frh_prefilter_scratch_bytes(T, H, W, C, rt)
        work = scratch(nb, dev)
        check(lib().frh_prefilter_3d(ptr(x), int(f64), T, H, W, C, ptr(lo), ptr(den), ptr(wt), rt, ptr(wy), ry,
                                     ptr(wx), rx, ptr(out), int(out64), ptr(work), nb, stream_handle()),
              "frh_prefilter_3d")
        return out
    nb = lib().frh_prefilter_scratch_bytes(T, H, W, C, 0)
    work = scratch(nb, dev)
    check(lib().frh_prefilter_ex(ptr(x), int(f64), T, H, W, C, ptr(lo), ptr(den), ptr(wy), ry, ptr(wx), rx,
                                 ptr(out), int(out64), int(semantics), ptr(work), nb, stream_handle()),
          "frh_prefilter_ex")
    return out
