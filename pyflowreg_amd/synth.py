"""Synthetic two-photon-like recordings for tests and benchmarks.

Scenes are sums of Gaussian blobs on a 0.05 background.  A moving frame
samples the scene at x - d(x) with bicubic interpolation, so that the
registration convention ``registered(x) = moving(x + w(x))`` recovers
w ~= d.  Displacements are white noise low-passed by a Gaussian of width
``sigma_d`` px (periodic, in the Fourier domain) and scaled to a target
maximum norm; Poisson noise is applied at a peak count.

Everything runs in torch so a batch can be generated on the device that
will register it; seeds make every frame reproducible on any device.
"""

from __future__ import annotations

import math

import numpy as np
import torch


def _gen(seed: int, device) -> torch.Generator:
    g = torch.Generator(device=device)
    g.manual_seed(int(seed))
    return g


def blob_scene(H: int, W: int, n_blobs: int, seed: int, device="cpu", background=0.05):
    """Sum of ``n_blobs`` Gaussians (sigma U[1,4] px, amplitude U[0.2,1],
    uniform centres) on a constant background, float64 (H, W)."""
    g = _gen(seed, "cpu")
    sig = torch.rand(n_blobs, generator=g, dtype=torch.float64) * 3 + 1
    amp = torch.rand(n_blobs, generator=g, dtype=torch.float64) * 0.8 + 0.2
    cy = torch.rand(n_blobs, generator=g, dtype=torch.float64) * H
    cx = torch.rand(n_blobs, generator=g, dtype=torch.float64) * W
    y = torch.arange(H, dtype=torch.float64)
    x = torch.arange(W, dtype=torch.float64)
    gy = torch.exp(-0.5 * ((y[None, :] - cy[:, None]) / sig[:, None]) ** 2)  # (n, H)
    gx = torch.exp(-0.5 * ((x[None, :] - cx[:, None]) / sig[:, None]) ** 2)  # (n, W)
    img = (gy * amp[:, None]).T @ gx + background
    return img.to(device)


def smooth_field(T: int, H: int, W: int, sigma_d: float, max_disp, seed: int, device="cpu"):
    """(T, 2, H, W) float64 displacement fields (component 0 = x, 1 = y):
    N(0,1) noise, Gaussian low-pass of width sigma_d, scaled per frame so
    that max ||d|| equals ``max_disp`` (scalar or per-frame sequence)."""
    g = _gen(seed, device)
    noise = torch.randn((T, 2, H, W), generator=g, dtype=torch.float64, device=device)
    fy = torch.fft.fftfreq(H, dtype=torch.float64, device=device)
    fx = torch.fft.rfftfreq(W, dtype=torch.float64, device=device)
    transfer = torch.exp(-2 * math.pi ** 2 * sigma_d ** 2 * (fy[:, None] ** 2 + fx[None, :] ** 2))
    d = torch.fft.irfft2(torch.fft.rfft2(noise) * transfer, s=(H, W))
    norm = torch.sqrt(d[:, 0] ** 2 + d[:, 1] ** 2).amax(dim=(1, 2))
    target = torch.as_tensor(max_disp, dtype=torch.float64, device=device).reshape(-1)
    return d * (target / norm.clamp_min(1e-30))[:, None, None, None]


def warp_scene(scene: torch.Tensor, d: torch.Tensor):
    """Bicubic sample of ``scene`` (C, H, W) at x - d(x) for each frame of
    d (T, 2, H, W); returns (T, C, H, W) float64."""
    T = d.shape[0]
    C, H, W = scene.shape
    y = torch.arange(H, dtype=torch.float64, device=d.device)
    x = torch.arange(W, dtype=torch.float64, device=d.device)
    sx = x[None, None, :] - d[:, 0]
    sy = y[None, :, None] - d[:, 1]
    grid = torch.stack([2 * sx / (W - 1) - 1, 2 * sy / (H - 1) - 1], dim=-1)
    src = scene[None].expand(T, C, H, W)
    return torch.nn.functional.grid_sample(src, grid, mode="bicubic", padding_mode="border",
                                           align_corners=True)


def poisson(img: torch.Tensor, peak, seed: int):
    """Poisson noise at ``peak`` counts for the image maximum (scalar or
    per-frame (T,) values for a (T, ...) stack)."""
    g = _gen(seed, img.device)
    pk = torch.as_tensor(peak, dtype=img.dtype, device=img.device)
    if pk.ndim == 1:
        pk = pk.reshape(-1, *([1] * (img.ndim - 1)))
    return torch.poisson(img * pk, generator=g) / pk


def single_pair(H=256, W=256, n_blobs=400, sigma_d=32.0, max_disp=4.0, peak=200.0, seed=0,
                device="cpu"):
    """One noise-free fixed frame and a warped, Poisson-noisy moving frame,
    1 channel; returns numpy float32 (H, W) x2 and the true d (H, W, 2)."""
    scene = blob_scene(H, W, n_blobs, seed, device)
    d = smooth_field(1, H, W, sigma_d, max_disp, seed + 1, device)
    mov = poisson(warp_scene(scene[None], d)[0, 0], peak, seed + 2)
    gt = d[0].permute(1, 2, 0)
    return (scene.float().cpu().numpy(), mov.float().cpu().numpy(), gt.cpu().numpy())


def video(T: int, H=512, W=512, n_channels=2, n_blobs=None, drift_step=0.3, sigma_d=48.0,
          max_disp=3.0, peak=500.0, seed=1000, device="cpu"):
    """A (T, H, W, C) float32 recording: rigid random-walk drift (per-axis
    step N(0, drift_step^2)) plus a fresh smooth field per frame; channel 2
    is 0.5 x channel 1 plus an independent blob texture.  Returns the frames
    and the true displacement (T, H, W, 2) as torch tensors on ``device``."""
    if n_blobs is None:
        n_blobs = int(round(400 * H * W / (256 * 256)))
    s1 = blob_scene(H, W, n_blobs, seed, device)
    chans = [s1]
    if n_channels > 1:
        s2 = 0.5 * s1 + blob_scene(H, W, n_blobs, seed + 1, device)
        chans.append(s2)
    scene = torch.stack(chans[:n_channels])
    g = _gen(seed + 2, "cpu")
    steps = torch.randn((T, 2), generator=g, dtype=torch.float64) * drift_step
    drift = torch.cumsum(steps, dim=0).to(device)
    d = smooth_field(T, H, W, sigma_d, max_disp, seed + 3, device) + drift[:, :, None, None]
    frames = poisson(warp_scene(scene, d), peak, seed + 4)
    return frames.permute(0, 2, 3, 1).float().contiguous(), d.permute(0, 2, 3, 1).contiguous()


def uneven_pairs(n_pairs: int, seed=3000, sizes=((256, 256), (512, 512), (796, 512), (1024, 1024)),
                 device="cpu"):
    """Independent (fixed, moving) pairs with sizes drawn from ``sizes``,
    1-2 channels, max displacement U[2,8] px and Poisson peak U[100,1000];
    returns a list of (fixed, moving) float32 (H, W, C) tensors."""
    rng = np.random.default_rng(seed)
    out = []
    for k in range(n_pairs):
        H, W = sizes[int(rng.integers(len(sizes)))]
        C = int(rng.integers(1, 3))
        md = float(rng.uniform(2, 8))
        pk = float(rng.uniform(100, 1000))
        nb = int(round(400 * H * W / (256 * 256)))
        chans = [blob_scene(H, W, nb, seed + 10 * k + c, device) for c in range(C)]
        scene = torch.stack(chans)
        d = smooth_field(1, H, W, 32.0, md, seed + 10 * k + 5, device)
        mov = poisson(warp_scene(scene, d)[0], pk, seed + 10 * k + 6)
        out.append((scene.permute(1, 2, 0).float().contiguous(),
                    mov.permute(1, 2, 0).float().contiguous()))
    return out
