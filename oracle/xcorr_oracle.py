"""CPU restatement of the rigid cross-correlation prealignment
(util/xcorr_prealignment.py:7-113) for the tests: test infrastructure only,
never imported by the package.

Third-party pieces, neither installed here (parity with them is unpinned):
* cv2.resize(..., INTER_LANCZOS4) (opencv-contrib-python>=4.10), called by
  util/resize_util.py:27-40 with dsize (size[1], size[0]);
* skimage.registration.phase_cross_correlation (scikit-image), the
  published algorithm of Guizar-Sicairos, Thurman and Fienup, Opt. Lett. 33,
  156 (2008), with scikit-image's normalization="phase" and disambiguate
  options.

Pinned by (tests/test_xcorr_oracle.py): the reference's own tests of
estimate_rigid_xcorr_2d (tests/util/test_xcorr_prealignment.py:19-140, same
inputs and tolerances); exact recovery of circular integer shifts (the
phase-normalised cross-power spectrum of a periodic shift is a pure phase
ramp, whose inverse DFT is a delta); the matrix DFT against numpy's inverse
FFT at integer points and against the closed-form peak of a fractional
phase ramp; the Lanczos weights against sinc(t) sinc(t/4) at their
definition (partition of unity, identity at scale 1, symmetry at half-pixel
offsets).
"""

from __future__ import annotations

import itertools
import math

import numpy as np
import scipy.fft as sfft
from scipy import ndimage as ndi

from oracle import oracle as _oracle


# ------------------------------------------------------------------ resize
def lanczos4_kernel(t: np.ndarray) -> np.ndarray:
    """sinc(t) sinc(t / 4) with its exact values at the integers (1 at 0,
    0 elsewhere)."""
    t = np.asarray(t, dtype=np.float64)
    out = np.sinc(t) * np.sinc(t / 4)
    integer = t == np.round(t)
    out = np.where(integer, (t == 0).astype(np.float64), out)
    return out


def lanczos4_tables(n_in: int, n_out: int):
    """Taps of cv2's INTER_LANCZOS4 along one axis: source position
    (d + 0.5) * n_in / n_out - 0.5 in float32, 8 taps from floor - 3,
    indices clamped to the image, float32 weights normalised to sum 1."""
    scale = n_in / n_out
    idx = np.empty((n_out, 8), np.int32)
    wt = np.empty((n_out, 8), np.float32)
    for d in range(n_out):
        f = np.float32((d + 0.5) * scale - 0.5)
        s = math.floor(f)
        frac = float(np.float32(f - np.float32(s)))
        t = frac + 3 - np.arange(8)
        if np.any(np.abs(t) < 1e-6):  # on a source sample: the identity
            w = (np.abs(t) < 1e-6).astype(np.float32)
        else:
            w = lanczos4_kernel(t).astype(np.float32)
            w = (w / np.float32(w.sum(dtype=np.float32))).astype(np.float32)
        wt[d] = w
        idx[d] = np.clip(s - 3 + np.arange(8), 0, n_in - 1)
    return idx, wt


def resize_with_tables(img: np.ndarray, oh: int, ow: int, tables) -> np.ndarray:
    """Separable resize of a float32 image (h, w) with the given tap tables
    (ix, wx, iy, wy): float32 products, float64 sums, float32 result (the C
    oracle's table resize)."""
    ix, wx, iy, wy = tables
    src = np.ascontiguousarray(img, dtype=np.float32)
    h, w = src.shape
    out = np.empty((oh, ow), np.float32)
    _oracle.lib().or_resize_tab(src, h, w, 1, out, oh, ow, np.ascontiguousarray(ix), np.ascontiguousarray(wx),
                                np.ascontiguousarray(iy), np.ascontiguousarray(wy), ix.shape[1])
    return out


def resize_lanczos4(img: np.ndarray, oh: int, ow: int, tables=None) -> np.ndarray:
    if tables is None:
        ix, wx = lanczos4_tables(img.shape[1], ow)
        iy, wy = lanczos4_tables(img.shape[0], oh)
        tables = (ix, wx, iy, wy)
    return resize_with_tables(img, oh, ow, tables)


# ------------------------------------------------------- phase correlation
def upsampled_dft(R: np.ndarray, region: int, up: float, offsets) -> np.ndarray:
    """sum_k R[k] exp(+2 pi i f_k (j - offset) / (n up)) per axis, j in
    [0, region): the inverse DFT (without 1/n) of R sampled on a 1/up grid
    (complex128)."""
    h, w = R.shape
    fy = np.fft.fftfreq(h, 1.0 / h)
    fx = np.fft.fftfreq(w, 1.0 / w)
    j = np.arange(region, dtype=np.float64)
    Ey = np.exp(2j * np.pi / (h * up) * np.outer(j - offsets[0], fy))
    Ex = np.exp(2j * np.pi / (w * up) * np.outer(j - offsets[1], fx))
    return Ey @ R.astype(np.complex128) @ Ex.T


def _disambiguate(ref, mov, shift):
    h, w = ref.shape
    pos = [shift[0] % h, shift[1] % w]
    neg = [pos[0] - h, pos[1] - w]
    subpixel = any(s % 1 != 0 for s in shift)
    shifted = ndi.shift(mov, shift, mode="grid-wrap", order=3 if subpixel else 0)
    splits = [((0, int(round(p))), (int(round(p)), n)) for p, n in zip(pos, (h, w))]
    best, pick = -1.0, None
    for ky, kx in itertools.product(range(2), range(2)):
        (y0, y1), (x0, x1) = splits[0][ky], splits[1][kx]
        a = ref[y0:y1, x0:x1].ravel()
        b = shifted[y0:y1, x0:x1].ravel()
        corr = -1.0
        if a.size > 2:
            with np.errstate(invalid="ignore", divide="ignore"):
                corr = np.corrcoef(a, b)[0, 1]
        if corr > best:
            best, pick = corr, (ky, kx)
    if pick is None:
        return np.asarray(shift, dtype=np.float64)
    return np.array([pos[0] if pick[0] == 1 else neg[0], pos[1] if pick[1] == 1 else neg[1]])


def phase_cross_correlation(ref: np.ndarray, mov: np.ndarray, upsample_factor: int = 1,
                            normalization="phase", disambiguate: bool = False) -> np.ndarray:
    """(row, col) shift registering mov to ref (float32 images)."""
    F = sfft.fft2(ref)
    G = sfft.fft2(mov)
    R = F * G.conj()
    if normalization == "phase":
        eps = np.finfo(R.real.dtype).eps
        R = R / np.maximum(np.abs(R), 100 * eps)
    elif normalization is not None:
        raise ValueError("normalization must be either phase or None")
    cc = sfft.ifft2(R)
    peak = np.unravel_index(np.argmax(np.abs(cc)), cc.shape)
    shape = np.array(cc.shape, dtype=np.float32)
    shifts = np.array(peak, dtype=np.float32)
    mid = np.fix(shape / 2)
    shifts = np.where(shifts > mid, shifts - shape, shifts).astype(np.float32)
    if upsample_factor > 1:
        up = np.float32(upsample_factor)
        shifts = (np.round(shifts * up) / up).astype(np.float32)
        region = int(np.ceil(up * np.float32(1.5)))
        dftshift = np.float32(np.fix(region / 2.0))
        offsets = (dftshift - shifts * up).astype(np.float64)
        cu = np.abs(upsampled_dft(R, region, float(up), offsets))
        pk = np.array(np.unravel_index(np.argmax(cu), cu.shape), dtype=np.float32) - dftshift
        shifts = (shifts + pk / up).astype(np.float32)
    out = shifts.astype(np.float64)
    if disambiguate:
        out = _disambiguate(ref, mov, out)
    return out


def estimate_rigid_xcorr_2d(ref_img, mov_img, target_hw=(256, 256), up=1, normalization="phase",
                            disambiguate=True, weight=None, tables=None):
    """util/xcorr_prealignment.py:7-113; tables (ix, wx, iy, wy) replaces
    the Lanczos tables (to check the GPU kernel's arithmetic alone)."""
    ref = np.asarray(ref_img)
    mov = np.asarray(mov_img)
    if ref.ndim == 3 and ref.shape[2] > 1:
        if weight is not None:
            w = np.asarray(weight).reshape(-1).astype(np.float32)
            w /= w.sum()
            ref = np.tensordot(ref, w, axes=([2], [0]))
            mov = np.tensordot(mov, w, axes=([2], [0]))
        else:
            ref = ref.mean(axis=2)
            mov = mov.mean(axis=2)
    elif ref.ndim == 3:
        ref = ref[..., 0]
        mov = mov[..., 0]
    H, W = ref.shape
    if isinstance(target_hw, (int, float)):
        target_hw = (int(target_hw), int(target_hw))
    Th = H if target_hw is None else min(H, int(target_hw[0]))
    Tw = W if target_hw is None else min(W, int(target_hw[1]))
    up = max(int(up), 1)
    sy, sx = H / Th, W / Tw
    ref_s = ref.astype(np.float32)
    mov_s = mov.astype(np.float32)
    if (Th, Tw) != (H, W):
        ref_s = resize_lanczos4(ref_s, Tw, Th, tables)
        mov_s = resize_lanczos4(mov_s, Tw, Th, tables)
    ref_s = ref_s - ref_s.mean()
    mov_s = mov_s - mov_s.mean()
    win = np.hanning(ref_s.shape[0]).astype(np.float32)[:, None] * np.hanning(ref_s.shape[1]).astype(np.float32)[None]
    ref_s = ref_s * win
    mov_s = mov_s * win
    s = phase_cross_correlation(ref_s, mov_s, upsample_factor=up, normalization=normalization,
                                disambiguate=disambiguate)
    return np.array([-s[1] * sx, -s[0] * sy], dtype=np.float32)
