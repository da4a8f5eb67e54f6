"""Rigid prealignment by phase cross-correlation on the GPU.

``estimate_rigid_xcorr_2d`` follows util/xcorr_prealignment.py:7-113 and
``estimate_rigid_xcorr_batch`` runs it for T moving frames against one
reference in one pass (batched FFTs, one resize launch per image set), which
is what the executors' ``cc_initialization`` step needs
(motion_correction/parallelization/sequential.py:97-130).

Per frame: combine channels (weighted by the normalised channel weights, or
the channel mean), downsample to ``target_hw`` when smaller than the frame,
float32, subtract the mean, multiply by the separable Hann window
(``np.hanning`` rows x columns), phase cross-correlation, and return the
forward shift [dx, dy] = [-tx * W / Tw, -ty * H / Th] in float32.

Third-party arithmetic on this path, restated from the published algorithms
(neither library is importable here, so parity with them is unpinned; the
reference's own tests, tests/util/test_xcorr_prealignment.py, pin the
recovered shifts, and tests/test_xcorr.py runs them):

* ``cv2.resize(..., INTER_LANCZOS4)`` (opencv-contrib-python>=4.10), called
  by util/resize_util.py:27-40 with dsize (size[1], size[0]) -- so the
  reference's (Tw, Th) request yields an array of Tw rows and Th columns,
  which is kept.  Source position (d + 0.5) * in / out - 0.5 (float32),
  8 taps from floor - 3 clamped to the image, Lanczos-4 weights
  sinc(t) sinc(t / 4) normalised to sum 1 as float32 tables.  The taps are
  applied with the separable table resize kernel of the pyramid (float32
  products, float64 sums, float32 result).
* ``skimage.registration.phase_cross_correlation`` (scikit-image; Guizar-
  Sicairos, Thurman and Fienup, Opt. Lett. 33, 156 (2008)): cross-power
  spectrum F_ref * conj(F_mov) (complex64 FFTs of the float32 images), with
  normalization="phase" divided by max(|.|, 100 eps_float32); the peak of
  |ifft2| gives the integer shift, wrapped to (-n/2, n/2].  With
  upsample_factor up > 1 the shift is rounded to the 1/up grid and refined on
  a ceil(1.5 up)-point grid around it by a matrix DFT of the cross-power
  spectrum (evaluated here in complex128).  disambiguate=True resolves the
  periodic ambiguity: the moving image is shifted periodically by the
  estimate, the frame is split at the positive shift on each axis, and the
  alias whose overlap tile has the highest Pearson correlation with the
  reference wins (a subpixel periodic shift is applied in the Fourier domain
  here, where scikit-image uses a cubic spline with grid-wrap; it only
  chooses among the 4 tiles).
"""

from __future__ import annotations

import math
import warnings
from typing import Optional

import numpy as np
import torch

from ._lib import check, lib, ptr, stream_handle


# ------------------------------------------------------------------ resize
def lanczos4_tables(n_in: int, n_out: int):
    """cv2 INTER_LANCZOS4 taps along one axis: indices (n_out, 8) int32
    clamped to [0, n_in - 1] and float32 weights (n_out, 8)."""
    scale = n_in / n_out
    idx = np.empty((n_out, 8), np.int32)
    wt = np.empty((n_out, 8), np.float32)
    k = np.arange(8)
    for d in range(n_out):
        f = np.float32((d + 0.5) * scale - 0.5)
        s = int(math.floor(f))
        f = float(np.float32(f - np.float32(s)))
        w = np.empty(8, np.float32)
        for i in range(8):
            t = f + 3 - i
            if abs(t) < 1e-6:
                w[i] = np.float32(1e30)  # the centre tap; the others vanish on normalising
            else:
                w[i] = np.float32(math.sin(math.pi * t) * math.sin(math.pi * t / 4) / (t * t))
        tot = np.float32(0)
        for i in range(8):
            tot = np.float32(tot + w[i])
        wt[d] = (w * np.float32(np.float32(1) / tot)).astype(np.float32)
        idx[d] = np.clip(s - 3 + k, 0, n_in - 1)
    return idx, wt


def resize_lanczos4(planes: torch.Tensor, oh: int, ow: int) -> torch.Tensor:
    """(N, h, w) float32 device planes -> (N, oh, ow) float32."""
    N, h, w = planes.shape
    ix, wx = lanczos4_tables(w, ow)
    iy, wy = lanczos4_tables(h, oh)
    dev = planes.device
    d = [torch.from_numpy(a).to(dev) for a in (ix, wx, iy, wy)]
    src = planes.contiguous()
    dst = torch.empty((N, oh, ow), dtype=torch.float32, device=dev)
    tmp = torch.empty((N, h, ow), dtype=torch.float32, device=dev)
    check(lib().frh_resize(ptr(src), N, h, w, ptr(dst), oh, ow, ptr(d[0]), ptr(d[1]), ptr(d[2]), ptr(d[3]), 8,
                           ptr(tmp), stream_handle()), "frh_resize")
    return dst


# ------------------------------------------------------- phase correlation
def _freqs(n: int, dev) -> torch.Tensor:
    """Signed integer frequency of each FFT bin: 0, 1, ..., -2, -1."""
    return torch.from_numpy(np.fft.fftfreq(n, 1.0 / n)).to(dev)


def _upsampled_cc(R: torch.Tensor, region: int, up: float, offsets: torch.Tensor) -> torch.Tensor:
    """|inverse DFT| of the cross-power spectra R (T, h, w) at the points
    (j - offset) / up, j = 0..region-1, per axis: (T, region, region)."""
    T, h, w = R.shape
    dev = R.device
    j = torch.arange(region, device=dev, dtype=torch.float64)
    Ey = torch.exp((2j * math.pi / (h * up)) * ((j[None, :, None] - offsets[:, 0, None, None]) *
                                                  _freqs(h, dev)[None, None, :]))
    Ex = torch.exp((2j * math.pi / (w * up)) * ((j[None, :, None] - offsets[:, 1, None, None]) *
                                                  _freqs(w, dev)[None, None, :]))
    Rc = R.to(torch.complex128).contiguous()
    cc = torch.bmm(torch.bmm(Ey.contiguous(), Rc), Ex.transpose(1, 2).contiguous())
    return cc.abs()


def _periodic_shift(imgs: torch.Tensor, shifts: np.ndarray) -> torch.Tensor:
    """imgs (T, h, w) each shifted by its shifts[t] = (dy, dx) with
    wrap-around: out[t, y, x] = imgs[t, y - dy, x - dx].  Integer shifts
    gather; a frame with a fractional shift takes the Fourier shift."""
    T, h, w = imgs.shape
    dev = imgs.device
    integer = np.all(shifts == np.round(shifts), axis=1)
    out = torch.empty_like(imgs)
    if integer.any():
        idx = torch.from_numpy(np.nonzero(integer)[0]).to(dev)
        s = torch.from_numpy(np.round(shifts[integer]).astype(np.int64)).to(dev)
        ys = (torch.arange(h, device=dev)[None, :] - s[:, 0:1]) % h
        xs = (torch.arange(w, device=dev)[None, :] - s[:, 1:2]) % w
        src = imgs[idx]
        out[idx] = src[torch.arange(len(idx), device=dev)[:, None, None], ys[:, :, None], xs[:, None, :]]
    if (~integer).any():
        idx = torch.from_numpy(np.nonzero(~integer)[0]).to(dev)
        s = torch.from_numpy(shifts[~integer].astype(np.float64)).to(dev)
        fy = _freqs(h, dev)[None, :, None] / h
        fx = _freqs(w, dev)[None, None, :] / w
        ramp = torch.exp(-2j * math.pi * (fy * s[:, 0, None, None] + fx * s[:, 1, None, None]))
        out[idx] = torch.fft.ifft2(torch.fft.fft2(imgs[idx].to(torch.float64)) * ramp).real.to(imgs.dtype)
    return out


def _disambiguate(ref: torch.Tensor, movs: torch.Tensor, shifts: np.ndarray) -> np.ndarray:
    """The periodic alias of each frame's shift whose overlap tile
    correlates best with the reference (Pearson, centred per tile); one
    host transfer for the whole batch."""
    T, h, w = movs.shape
    dev = movs.device
    pos = np.stack([shifts[:, 0] % h, shifts[:, 1] % w], axis=1)
    neg = pos - np.array([h, w], dtype=np.float64)
    split = np.round(pos).astype(np.int64)
    shifted = _periodic_shift(movs, shifts).to(torch.float64)
    a = ref.to(torch.float64)[None].expand(T, h, w)
    sy = torch.from_numpy(split[:, 0]).to(dev)[:, None, None]
    sx = torch.from_numpy(split[:, 1]).to(dev)[:, None, None]
    rows = torch.arange(h, device=dev)[None, :, None]
    cols = torch.arange(w, device=dev)[None, None, :]
    corr = torch.empty((T, 4), dtype=torch.float64, device=dev)
    count = torch.empty((T, 4), dtype=torch.float64, device=dev)
    # tiles in the order (rows before/after the split) x (columns before/after)
    for k, (ky, kx) in enumerate(((0, 0), (0, 1), (1, 0), (1, 1))):
        my = (rows >= sy) if ky else (rows < sy)
        mx = (cols >= sx) if kx else (cols < sx)
        m = (my & mx).to(torch.float64)
        n = m.sum(dim=(1, 2))
        na = n.clamp(min=1)
        ma = (a * m).sum(dim=(1, 2)) / na
        mb = (shifted * m).sum(dim=(1, 2)) / na
        da = (a - ma[:, None, None]) * m
        db = (shifted - mb[:, None, None]) * m
        corr[:, k] = (da * db).sum(dim=(1, 2)) / torch.sqrt((da * da).sum(dim=(1, 2)) * (db * db).sum(dim=(1, 2)))
        count[:, k] = n
    corr = corr.cpu().numpy()
    count = count.cpu().numpy()
    out = np.array(shifts, dtype=np.float64)
    for t in range(T):
        best, pick = -1.0, None
        for k in range(4):
            c = corr[t, k] if count[t, k] > 2 else -1.0
            if c > best:  # NaN (a constant tile) never wins
                best, pick = c, k
        if pick is None:
            warnings.warn("phase_cross_correlation: could not disambiguate the shift (no overlap tile "
                          "correlates); returning the periodic estimate")
            continue
        ky, kx = pick // 2, pick % 2
        out[t] = [pos[t, 0] if ky == 1 else neg[t, 0], pos[t, 1] if kx == 1 else neg[t, 1]]
    return out


def phase_cross_correlation(ref: torch.Tensor, movs: torch.Tensor, upsample_factor: int = 1,
                            normalization: Optional[str] = "phase", disambiguate: bool = False) -> np.ndarray:
    """Shifts (T, 2) = (rows, cols) that register each of movs (T, h, w) to
    ref (h, w), both float32 device tensors."""
    if normalization not in ("phase", None):
        raise ValueError("normalization must be either phase or None")
    T, h, w = movs.shape
    Fr = torch.fft.fft2(ref.to(torch.float32))
    Fm = torch.fft.fft2(movs.to(torch.float32))
    R = Fr[None] * Fm.conj()
    if normalization == "phase":
        eps = np.finfo(np.float32).eps
        R = R / torch.clamp(R.abs(), min=100 * eps)
    cc = torch.fft.ifft2(R).abs().reshape(T, -1)
    peak = torch.argmax(cc, dim=1).cpu().numpy()
    shifts = np.stack([peak // w, peak % w], axis=1).astype(np.float32)
    mid = np.array([np.fix(h / 2), np.fix(w / 2)], np.float32)
    size = np.array([h, w], np.float32)
    shifts = np.where(shifts > mid, shifts - size, shifts).astype(np.float32)
    if upsample_factor > 1:
        up = np.float32(upsample_factor)
        shifts = (np.round(shifts * up) / up).astype(np.float32)
        region = int(np.ceil(up * np.float32(1.5)))
        dftshift = np.float32(np.fix(region / 2.0))
        offsets = (dftshift - shifts * up).astype(np.float32)
        mag = _upsampled_cc(R, region, float(up),
                            torch.from_numpy(offsets.astype(np.float64)).to(R.device)).reshape(T, -1)
        pk = torch.argmax(mag, dim=1).cpu().numpy()
        maxima = np.stack([pk // region, pk % region], axis=1).astype(np.float32) - dftshift
        shifts = (shifts + maxima / up).astype(np.float32)
    out = shifts.astype(np.float64)
    if disambiguate:
        out = _disambiguate(ref, movs, out)
    return out


# ---------------------------------------------------------------- frontend
def _as_frames(x) -> torch.Tensor:
    if isinstance(x, torch.Tensor):
        return x
    return torch.from_numpy(np.ascontiguousarray(np.asarray(x)))


def _combine(img: torch.Tensor, weight, has_channels: bool) -> torch.Tensor:
    """(..., H, W, C) or (..., H, W) -> (..., H, W):
    util/xcorr_prealignment.py:41-55."""
    if has_channels and img.shape[-1] > 1:
        C = img.shape[-1]
        if weight is not None:
            w = np.asarray(weight).reshape(-1).astype(np.float32)
            if w.size != C:
                raise ValueError(f"shape-mismatch for sum: {w.size} channel weights for {C} channels")
            w = w / w.sum()
            wt = torch.from_numpy(w).to(img.device)
            acc = img[..., 0] * wt[0].to(img.dtype)
            for c in range(1, C):
                acc = acc + img[..., c] * wt[c].to(img.dtype)
            return acc
        return img.mean(dim=-1)
    if has_channels:
        return img[..., 0]
    return img


def estimate_rigid_xcorr_batch(ref_img, mov_imgs, target_hw=(256, 256), up: int = 1,
                               normalization: Optional[str] = "phase", disambiguate: bool = True,
                               weight=None) -> torch.Tensor:
    """estimate_rigid_xcorr_2d for every frame of mov_imgs (T, H, W[, C])
    against ref_img (H, W[, C]); returns (T, 2) float32 [dx, dy] on the
    device."""
    from .optical_flow import _device
    dev = _device()
    ref = _as_frames(ref_img).to(dev)
    movs = _as_frames(mov_imgs).to(dev)
    if ref.ndim not in (2, 3) or movs.ndim != ref.ndim + 1 or tuple(movs.shape[1:]) != tuple(ref.shape):
        raise ValueError(f"mov_imgs must hold frames shaped like ref_img {tuple(ref.shape)}, "
                         f"got {tuple(movs.shape)}")
    ref2 = _combine(ref, weight, ref.ndim == 3)
    mov2 = _combine(movs, weight, ref.ndim == 3)
    H, W = ref2.shape
    if target_hw is not None and isinstance(target_hw, (int, float)):
        target_hw = (int(target_hw), int(target_hw))
    Th = H if target_hw is None else min(H, int(target_hw[0]))
    Tw = W if target_hw is None else min(W, int(target_hw[1]))
    if up < 1:
        print(f"Warning: upsampling factor {up} < 1 is invalid. Setting to 1 (no upsampling).")
        up = 1
    sy = H / Th
    sx = W / Tw
    ref_s = ref2.to(torch.float32)
    mov_s = mov2.to(torch.float32)
    if (Th, Tw) != (H, W):
        # cv2.resize(img, (Th, Tw)): Tw rows, Th columns
        ref_s = resize_lanczos4(ref_s[None], Tw, Th)[0]
        mov_s = resize_lanczos4(mov_s, Tw, Th)
    ref_s = ref_s - ref_s.mean()
    mov_s = mov_s - mov_s.mean(dim=(1, 2), keepdim=True)
    hy = torch.from_numpy(np.hanning(ref_s.shape[0]).astype(np.float32)).to(dev)
    hx = torch.from_numpy(np.hanning(ref_s.shape[1]).astype(np.float32)).to(dev)
    win = hy[:, None] * hx[None, :]
    ref_s = ref_s * win
    mov_s = mov_s * win[None]
    shifts = phase_cross_correlation(ref_s, mov_s, upsample_factor=up, normalization=normalization,
                                     disambiguate=disambiguate)
    dx = -shifts[:, 1] * sx
    dy = -shifts[:, 0] * sy
    return torch.from_numpy(np.stack([dx, dy], axis=1).astype(np.float32)).to(dev)


def estimate_rigid_xcorr_2d(ref_img, mov_img, target_hw=(256, 256), up: int = 1, normalization="phase",
                            disambiguate: bool = True, weight=None) -> np.ndarray:
    """util/xcorr_prealignment.py:7-113: [dx, dy] float32 (the forward
    shift; negate it to register mov_img onto ref_img)."""
    mov = _as_frames(mov_img)
    out = estimate_rigid_xcorr_batch(ref_img, mov[None], target_hw=target_hw, up=up,
                                     normalization=normalization, disambiguate=disambiguate, weight=weight)
    return out[0].cpu().numpy()
