"""Ordering / precision study of the reference algorithm (test infrastructure).

Runs the CPU oracle on one synthetic frame pair and reports max-abs and RMS
flow differences (px) between:
  * the FP64 lexicographic solver (the parity oracle) and
    - the reference's literal per-sweep expression order with libm pow,
    - the same solver on inputs perturbed by ~1 float32 ulp,
    - the float32 build of the whole solver/tensor path,
    - red-black instead of lexicographic SOR ordering,
    - the continuous (unquantized) bicubic warp;
  * the ground-truth displacement of the generator (informational).

Usage: python oracle/precision_study.py [H] [channels]
"""

import os
import sys
import time

import numpy as np

sys.path.insert(0, os.path.dirname(os.path.dirname(os.path.abspath(__file__))))

from oracle import oracle as O  # noqa: E402
from pyflowreg_amd import synth  # noqa: E402


def stats(a, b):
    d = np.abs(a - b)
    return float(d.max()), float(np.sqrt((d ** 2).mean()))


def main():
    H = int(sys.argv[1]) if len(sys.argv) > 1 else 256
    C = int(sys.argv[2]) if len(sys.argv) > 2 else 1
    O.build()
    if C == 1:
        fixed, mov, gt = synth.single_pair(H, H, n_blobs=int(400 * H * H / 65536), seed=0)
        f = fixed[:, :, None].astype(np.float64)
        m = mov[:, :, None]
        weight = None
    else:
        frames, d = synth.video(2, H, H, n_channels=C, seed=1000)
        f = frames[0].numpy().astype(np.float64)
        m = frames[1].numpy()
        gt = (d[1] - 0).numpy()
        weight = [0.5] * C
    ref = O.preprocess(f).astype(np.float32)
    mv = O.preprocess(m, f).astype(np.float32)
    t = time.time()
    w64 = O.get_displacement(ref, mv, weight=weight)
    dt = time.time() - t
    rows = []
    rows.append(("literal reference order (libm pow)",
                 stats(O.get_displacement(ref, mv, weight=weight, order=2), w64)))
    rng = np.random.default_rng(5)
    mv2 = (mv * (1 + 1e-7 * rng.standard_normal(mv.shape))).astype(np.float32)
    rows.append(("input perturbed by ~1 float32 ulp", stats(O.get_displacement(ref, mv2, weight=weight), w64)))
    rows.append(("float32 solver and tensor", stats(O.get_displacement(ref, mv, weight=weight, precision="f32"), w64)))
    rows.append(("red-black SOR ordering", stats(O.get_displacement(ref, mv, weight=weight, order=1), w64)))
    O.lib("f64").or_set_warp_quantize(0)
    rows.append(("continuous warp (no 1/32 px grid)", stats(O.get_displacement(ref, mv, weight=weight), w64)))
    O.lib("f64").or_set_warp_quantize(1)
    rows.append(("ground truth displacement", stats(w64, gt)))
    print(f"# {H}x{H}x{C}, oracle FP64 time {dt:.2f} s")
    print(f"{'variant vs FP64 lexicographic oracle':42s} {'max |dflow| px':>15s} {'RMS px':>10s}")
    for name, (mx, rms) in rows:
        print(f"{name:42s} {mx:15.3e} {rms:10.3e}")


if __name__ == "__main__":
    main()
