"""Write the golden fixtures of the optical-flow path from the CPU oracle.

Each fixture (tests/golden/<name>.npz) holds seeded inputs, the oracle's
outputs and, for cases that downsample, the float32 Gaussian taps the
oracle used (numpy's float32 exp may round differently on another host;
tests compare those taps first).  Inputs come from numpy's PCG64 generator
only, so they do not depend on torch.

Usage: python tests/golden/make_golden.py
"""

import json
import os
import sys

import numpy as np
from scipy.ndimage import gaussian_filter, map_coordinates

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from oracle import oracle as O  # noqa: E402


def blob_scene(rng, h, w, n_blobs):
    """Gaussian blobs (sigma U[1,4] px, amplitude U[0.2,1]) on a 0.05 background."""
    yy, xx = np.mgrid[0:h, 0:w].astype(np.float64)
    img = np.full((h, w), 0.05)
    for _ in range(n_blobs):
        cy, cx = rng.uniform(0, h), rng.uniform(0, w)
        s, a = rng.uniform(1, 4), rng.uniform(0.2, 1.0)
        img += a * np.exp(-((yy - cy) ** 2 + (xx - cx) ** 2) / (2 * s * s))
    return img


def moved(rng, scene, max_disp, sigma_d, peak):
    """Samples the scene at x - d(x) (cubic), then Poisson noise at ``peak``."""
    h, w = scene.shape[:2]
    d = np.stack([gaussian_filter(rng.standard_normal((h, w)), sigma_d, mode="wrap") for _ in range(2)])
    d *= max_disp / np.sqrt((d ** 2).sum(0)).max()
    yy, xx = np.mgrid[0:h, 0:w].astype(np.float64)
    out = np.stack([map_coordinates(scene[..., c], [yy - d[1], xx - d[0]], order=3, mode="nearest")
                    for c in range(scene.shape[2])], axis=-1)
    out = rng.poisson(np.clip(out, 0, None) * peak) / peak
    return out.astype(np.float32), np.moveaxis(d, 0, -1)


def pair(seed, h, w, C):
    rng = np.random.default_rng(seed)
    n = max(4, int(round(400 * h * w / 65536)))
    chans = [blob_scene(rng, h, w, n)]
    if C > 1:
        chans.append(0.5 * chans[0] + blob_scene(rng, h, w, n))
    scene = np.stack(chans[:C], axis=-1)
    mov, d = moved(rng, scene, 2.0, max(h, w) / 6, 500.0)
    return scene.astype(np.float32), mov, d


def level_sigmas(H, W, eta=0.8, levels=50, min_level=0):
    """Resize sigmas of the level images (0.6 / min(scale) below scale 1)."""
    out = []
    for (_, (h, w)) in O.level_plan(H, W, eta, levels, min_level)[0]:
        s = min(h / H, w / W)
        out.append(0.6 / s if s < 1 else 0.0)
    return out


def save(name, meta, **arrays):
    path = os.path.join(HERE, name + ".npz")
    np.savez_compressed(path, meta=np.array(json.dumps(meta)), **arrays)
    print(f"{path}: {os.path.getsize(path)} bytes")


def main():
    O.build()
    rng = np.random.default_rng(20240601)

    img = rng.random((24, 20, 2)).astype(np.float32)
    save("resize_down", {"size": [13, 11], "sigma": 0.6 / min(13 / 24, 11 / 20)}, img=img,
         out=O.resize(img, (13, 11)), gauss=O.gauss_g(0.6 / min(13 / 24, 11 / 20))[1])
    save("resize_up", {"size": [40, 36]}, img=img, out=O.resize(img, (40, 36)))

    f2 = rng.random((20, 24, 2)).astype(np.float32)
    f1 = rng.random((20, 24, 2)).astype(np.float32)
    u = (rng.standard_normal((20, 24)) * 2.0).astype(np.float32)
    v = (rng.standard_normal((20, 24)) * 2.0).astype(np.float32)
    save("warp", {}, f2=f2, f1=f1, u=u, v=v,
         out=O.imregister_wrapper(f2, u.astype(np.float64), v.astype(np.float64), f1))

    a = gaussian_filter(rng.random((18, 21)), 1.5).astype(np.float32)
    b = gaussian_filter(rng.random((18, 21)), 1.5).astype(np.float32)
    # float32 inputs: numpy computes the derivatives in float32
    J = O.get_motion_tensor_gc(a, b, 1.1, 0.9)
    save("tensor_gc", {"hy": 1.1, "hx": 0.9}, f1=a, f2=b, J=np.stack(J))
    # the pipeline's case below the top level: float64 level image, float32 warped image
    a64 = a.astype(np.float64)
    for kind, fn in (("gc", O.get_motion_tensor_gc), ("gray", O.get_motion_tensor_gray),
                     ("cs", O.get_motion_tensor_cs)):
        J = fn(a64, b, 1.1, 0.9)
        save(f"tensor_{kind}_mixed", {"hy": 1.1, "hx": 0.9, "kind": kind}, f1=a64, f2=b, J=np.stack(J))

    for name, (h, w, K, its, lag) in {"level_solver_a": (16, 14, 2, 12, 5),
                                      "level_solver_b": (9, 9, 1, 7, 3)}.items():
        f1s = np.stack([gaussian_filter(rng.random((h, w)), 1.5) for _ in range(K)], -1).astype(np.float32)
        f2s = np.stack([gaussian_filter(rng.random((h, w)), 1.5) for _ in range(K)], -1).astype(np.float32)
        Js = np.zeros((6, h + 2, w + 2, K))
        for k in range(K):
            Js[:, :, :, k] = np.stack(O.get_motion_tensor_gc(f1s[..., k], f2s[..., k], 1.1, 0.9))
        Wt = np.zeros((h + 2, w + 2, K))
        Wt[1:-1, 1:-1] = rng.uniform(0.2, 1.0, (h, w, K)).astype(np.float32)
        uu = np.pad(rng.standard_normal((h, w)).astype(np.float32).astype(np.float64), 1, mode="edge")
        vv = np.pad(rng.standard_normal((h, w)).astype(np.float32).astype(np.float64), 1, mode="edge")
        meta = {"alpha": [2.0, 2.5], "iterations": its, "update_lag": lag, "a_data": 0.45,
                "a_smooth": 1.0, "hx": 1.3, "hy": 0.8}
        du, dv = O.level_solver(*Js, Wt, uu, vv, meta["alpha"], its, lag, 0, np.full(K, 0.45), 1.0,
                                1.3, 0.8)
        save(name, meta, J=Js, weight=Wt, u=uu, v=vv, du=du, dv=dv)
        # compute_flow_gnc at a mid stage weight
        meta_g = dict(meta, gnc_beta=0.5)
        du, dv = O.level_solver(*Js, Wt, uu, vv, meta["alpha"], its, lag, 0, np.full(K, 0.45), 1.0,
                                1.3, 0.8, 0.5)
        save(name + "_gnc", meta_g, J=Js, weight=Wt, u=uu, v=vv, du=du, dv=dv)

    cases = {"displacement_32x32x1": (7, 32, 32, 1, None, {}),
             "displacement_40x48x2": (8, 40, 48, 2, (0.3, -0.2), {}),
             "displacement_gray_24x28x1": (9, 24, 28, 1, None, {"const_assumption": "gray"}),
             "displacement_cs_24x28x2": (10, 24, 28, 2, None, {"const_assumption": "cs"}),
             "displacement_gnc_20x24x1": (11, 20, 24, 1, (0.2, 0.1),
                                          {"gnc_schedule": [0.0, 0.5, 1.0], "warping_steps": 2,
                                           "iterations": 10})}
    for name, (seed, H, W, C, uv0, opts) in cases.items():
        fixed, mov, d = pair(seed, H, W, C)
        ref = O.preprocess(fixed.astype(np.float64)).astype(np.float32)
        mv = O.preprocess(mov, fixed.astype(np.float64)).astype(np.float32)
        kw = dict(opts)
        if C > 1:
            kw["weight"] = [0.5, 0.5]
        uv = None
        if uv0 is not None:
            uv = np.zeros((H, W, 2), np.float32)
            uv[..., 0], uv[..., 1] = uv0
            kw["uv"] = uv
        w = O.get_displacement(ref, mv, **kw)
        sig = level_sigmas(H, W)
        gauss = {f"gauss_{k}": O.gauss_g(s)[1] for k, s in enumerate(sig)}
        meta = {"seed": seed, "weight": kw.get("weight"), "sigmas": sig, "opts": opts,
                "options": "alpha=2, iterations=50, update_lag=5, eta=0.8, a_data=0.45 unless in opts"}
        save(name, meta, fixed_raw=fixed, moving_raw=mov, fixed=ref, moving=mv,
             uv=uv if uv is not None else np.zeros(0, np.float32), w=w, truth=d, **gauss)

    raw = (rng.random((20, 24, 2)) * 100).astype(np.float32)
    ref = (rng.random((20, 24, 2)) * 90 + 5).astype(np.float32)
    save("preprocess", {"sigma": [1.0, 1.0, 0.1]}, raw=raw, ref=ref,
         out=O.preprocess(raw, ref.astype(np.float64)).astype(np.float32))

    f = rng.standard_normal((12, 10))
    from scipy.ndimage import median_filter
    save("median", {}, f=f, out=median_filter(f, size=(5, 5), mode="mirror"))


if __name__ == "__main__":
    main()
