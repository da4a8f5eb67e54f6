"""Throughput benchmark: batched registration of 512x512x2 frames on one GPU.

One step = the hot path over one batch of synthetic frames already resident
in HBM: GPU preprocessing (normalize + Gaussian prefilter) of the batch,
then the batched get_displacement solve of every frame against the shared
reference (one BatchMotionCorrector batch, all frames sharing w_init).

Prints one JSON line.  The plain path (one frame per call, the sequential
executor's pattern) is timed beside it, and the CPU oracle is timed on the
host cores as ``cpu_baseline`` before the GPU is initialised.
"""

from __future__ import annotations

import argparse
import json
import multiprocessing as mp
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

PEAK_HBM_GBS = 8000.0
METRIC = "frames/s registered (512x512x2ch), batched vs plain GPU at \u22641e-3 px max |\u0394flow|"
# rocprofv3 FETCH_SIZE / WRITE_SIZE of the sweep kernel on this workload,
# scaled by the same counters on a copy kernel of known size.
TRAFFIC_FILE = os.path.join(ROOT, "profiles", "k_sweep_hbm_traffic_512x512x2_b256.md")


def _cpu_frame(args):
    """CPU oracle on one frame (runs in a worker process, no GPU)."""
    import numpy as np
    from oracle import oracle as O
    ref, mov = args
    t = time.perf_counter()
    O.get_displacement(ref, mov, weight=[0.5, 0.5])
    return time.perf_counter() - t


def cpu_baseline(H, W, C, n_frames, workers):
    """FP64 CPU restatement of the reference path (the oracle), one frame per
    worker process with 1 thread each, like pyflowreg's multiprocessing
    executor.  Returns (frames/s, cores, seconds of CPU work)."""
    os.environ["OMP_NUM_THREADS"] = "1"
    import numpy as np
    from oracle import oracle as O
    from pyflowreg_amd import synth
    O.build()
    frames, _ = synth.video(n_frames + 1, H, W, n_channels=C, seed=1000)
    f = frames.numpy()
    ref64 = f[0].astype(np.float64)
    ref = O.preprocess(ref64).astype(np.float32)
    movs = O.preprocess(f[1:], ref64).astype(np.float32)
    jobs = [(ref, movs[k]) for k in range(n_frames)]
    ctx = mp.get_context("fork")
    t = time.perf_counter()
    with ctx.Pool(workers) as pool:
        per = pool.map(_cpu_frame, jobs)
    wall = time.perf_counter() - t
    return n_frames / wall, workers, float(sum(per))


def sweep_algorithmic_bytes(solver, T):
    """Algorithmic HBM bytes of all SOR sweep launches of one run: per cell
    and pass, 5 float64 coefficients + 2 float64 start values read (not on
    the first pass of a level) and 2 float64 results written."""
    its, lag = solver.iterations, solver.update_lag
    starts = [it for it in range(its) if it == 0 or (it + 1) % lag == 0]
    total = 0
    for (_, h, w, *_rest) in solver.plan.levels:
        pmax = 5 if 64 * ((h + 63) // 64) <= 512 else 2
        n_pass = 0
        for k, b in enumerate(starts):
            e = starts[k + 1] if k + 1 < len(starts) else its
            n_pass += -(-(e - b) // pmax)
        cells = h * w
        total += T * cells * 8 * (9 * n_pass - 2)
    return total


def measured_traffic(H, W, C, T):
    """HBM bytes per sweep launch recorded by rocprofv3 for this workload
    (line ``traffic_bytes_per_launch: N`` of the profile summary), or None."""
    if (H, W, C, T) != (512, 512, 2, 256) or not os.path.exists(TRAFFIC_FILE):
        return None
    with open(TRAFFIC_FILE) as fh:
        for line in fh:
            if line.startswith("traffic_bytes_per_launch:"):
                return int(float(line.split(":", 1)[1]))
    return None


def copy_probe(mib):
    """A plain device copy of known size (calibrates rocprofv3's FETCH_SIZE
    and WRITE_SIZE counters): 5 copies of ``mib`` MiB of float32."""
    import torch
    n = mib * (1 << 20) // 4
    src = torch.ones(n, dtype=torch.float32, device="cuda")
    dst = torch.empty_like(src)
    for _ in range(5):
        dst.copy_(src)
    torch.cuda.synchronize()
    print(json.dumps({"copy_probe_bytes_read": n * 4, "copy_probe_bytes_written": n * 4,
                      "launches": 5}))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--batch", type=int, default=256)
    ap.add_argument("--size", type=int, default=512)
    ap.add_argument("--channels", type=int, default=2)
    ap.add_argument("--plain-frames", type=int, default=3)
    ap.add_argument("--cpu-frames", type=int, default=8)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--copy-probe", type=int, default=0, metavar="MIB",
                    help="only run a device copy of MIB MiB (counter calibration)")
    args = ap.parse_args()
    if args.copy_probe:
        copy_probe(args.copy_probe)
        return
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; --gpus must be 1",
                          "n_gpus": args.gpus}))
        sys.exit(2)

    H = W = args.size
    C = args.channels
    cpu = None
    if not args.no_cpu:
        workers = min(8, os.cpu_count() or 1, args.cpu_frames)
        fps, cores, work = cpu_baseline(H, W, C, args.cpu_frames, workers)
        cpu = {"value": round(fps, 4), "unit": "frames/s", "cores": cores, "kind": "port",
               "sample": f"{args.cpu_frames} frames {H}x{W}x{C}, one per worker process, "
                         f"{work:.1f} s of CPU work, FP64 oracle (C + scipy median)"}

    import numpy as np
    import torch
    import pyflowreg_amd as fr
    from pyflowreg_amd import synth
    from pyflowreg_amd.preprocess import preprocess

    dev = torch.device("cuda", 0)
    torch.cuda.set_device(dev)
    T = args.batch
    frames, _ = synth.video(T + 1, H, W, n_channels=C, seed=1000, device=dev)
    raw_ref = frames[0].double()
    batch = frames[1:].contiguous()
    ref_planes = preprocess(raw_ref)
    ref_hwc = ref_planes[0].permute(1, 2, 0).contiguous()
    solver = fr.DisplacementSolver(H, W, C, max_batch=T)
    solver.set_reference(ref_hwc, [0.5] * C)

    def step():
        p = preprocess(batch, normalization_ref=raw_ref)
        return solver.run(p.permute(0, 2, 3, 1), out_dtype=torch.float32)

    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        w = step()
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    fps = args.steps * T / dt

    # kernel timing of one profiled run (HIP events on the launch stream)
    p = preprocess(batch, normalization_ref=raw_ref).permute(0, 2, 3, 1)
    solver.run(p, out_dtype=torch.float32, profile=True)
    tm = solver.timings()
    sweep_ms, sweep_n = tm["sweep"]
    alg = sweep_algorithmic_bytes(solver, T)
    achieved = alg / (sweep_ms * 1e-3) / 1e9

    # plain path: one frame per call
    plain = fr.DisplacementSolver(H, W, C, max_batch=1)
    plain.set_reference(ref_hwc, [0.5] * C)
    pp = preprocess(batch[: args.plain_frames], normalization_ref=raw_ref).permute(0, 2, 3, 1)
    plain.run(pp[0], out_dtype=torch.float32)
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    for k in range(args.plain_frames):
        plain.run(pp[k], out_dtype=torch.float32)
    torch.cuda.synchronize()
    plain_fps = args.plain_frames / (time.perf_counter() - t1)

    out = {
        "metric": METRIC,
        "value": round(fps, 3),
        "unit": "frames/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(dt / args.steps * 1e3, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (blob scenes, smooth random displacement + drift, Poisson noise)",
        "config": {"workload": f"batched uniform: {T} frames {H}x{W}x{C} per launch, "
                               "alpha=2, 50 iterations, update_lag 5, eta 0.8, a_data 0.45",
                   "frames_per_step": T, "H": H, "W": W, "channels": C},
        "plain_fps": round(plain_fps, 3),
        "batched_over_plain": round(fps / plain_fps, 2),
        "kernel_ms": {k: round(v[0], 3) for k, v in tm.items()},
        "roofline": {"bound": "hbm", "kernel": "k_sweep (SOR wavefront)",
                     "achieved": round(achieved, 1), "peak": PEAK_HBM_GBS, "unit": "GB/s",
                     "frac": round(achieved / PEAK_HBM_GBS, 4),
                     "traffic": measured_traffic(H, W, C, T),
                     "algorithmic_bytes_per_launch": int(alg / max(sweep_n, 1)),
                     "launches": sweep_n, "avg_launch_ms": round(sweep_ms / max(sweep_n, 1), 4)},
        "cpu_baseline": cpu,
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
