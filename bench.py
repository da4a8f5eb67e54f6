"""Throughput benchmark: batched registration of 512x512x2 frames on one GPU.

One step = the hot path over one batch of synthetic frames already resident
in HBM: GPU preprocessing (normalize + Gaussian prefilter) of the batch,
then the batched get_displacement solve of every frame against the shared
reference (one BatchMotionCorrector batch, all frames sharing w_init).

Prints one JSON line.  Beside the batched path it reports
* the plain path (one frame per solver call, the sequential executor's
  pattern) and the batched/plain ratio;
* parity: the CPU oracle's flows for the first frames of the same batch
  against the GPU's float64 flows;
* ``cpu_baseline``: the CPU oracle timed on the host cores (worker
  processes forked before the GPU is initialised);
* ``roofline`` for the dominant kernel (algorithmic bytes over HIP-event
  time on the launch stream; HBM traffic per launch from the rocprofv3
  FETCH_SIZE/WRITE_SIZE summary in profiles/, when recorded for this
  workload), with the measured bandwidth of a plain 1 GiB device copy on
  the same GPU beside the 8 TB/s peak (``measured_copy_gbs``,
  ``frac_of_measured_copy``).
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
METRIC = "frames/s registered (512x512x2ch), batched vs plain GPU at ≤1e-3 px max |Δflow|"
# rocprofv3 FETCH_SIZE / WRITE_SIZE per launch of the solver kernels on this
# workload, scaled by the same counters on a copy kernel of known size.
TRAFFIC_FILE = os.path.join(ROOT, "profiles", "hbm_traffic_512x512x2_b256.md")


def _cpu_frame(args):
    """CPU oracle on one frame (runs in a worker process, no GPU)."""
    from oracle import oracle as O
    ref, mov = args
    t = time.perf_counter()
    w = O.get_displacement(ref, mov, weight=[0.5, 0.5])
    return time.perf_counter() - t, w


def _cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(raw, n_frames, workers):
    """FP64 CPU restatement of the reference path (the oracle) on the first
    n_frames moving frames, one frame per worker process with 1 thread each,
    like pyflowreg's multiprocessing executor
    (motion_correction/parallelization/multiprocessing.py:38-45,322-335).
    Returns (frames/s, seconds of CPU work, flows)."""
    os.environ["OMP_NUM_THREADS"] = "1"
    import numpy as np
    from oracle import oracle as O
    O.build()
    ref64 = raw[0].astype(np.float64)
    ref = O.preprocess(ref64).astype(np.float32)
    movs = O.preprocess(raw[1:n_frames + 1], ref64).astype(np.float32)
    jobs = [(ref, movs[k]) for k in range(n_frames)]
    ctx = mp.get_context("fork")
    t = time.perf_counter()
    with ctx.Pool(workers) as pool:
        res = pool.map(_cpu_frame, jobs, chunksize=1)
    wall = time.perf_counter() - t
    return n_frames / wall, float(sum(r[0] for r in res)), np.stack([r[1] for r in res])


def cpu_latency(raw, frames_0):
    """Single-thread per-frame latency of the CPU restatement (in this
    process, 1 thread): the 256x256x1 single pair (preprocessed with the
    fixed frame's range) over frames_0 runs, and one 512x512x2 frame of the
    recording."""
    os.environ["OMP_NUM_THREADS"] = "1"
    import numpy as np
    from oracle import oracle as O
    from pyflowreg_amd import synth
    fixed, moving, _ = synth.single_pair(256, 256, seed=0)
    r64 = fixed[..., None].astype(np.float64)
    ref0 = O.preprocess(r64).astype(np.float32)
    mov0 = O.preprocess(moving[None, ..., None], r64).astype(np.float32)[0]
    t = time.perf_counter()
    for _ in range(frames_0):
        O.get_displacement(ref0, mov0)
    lat0 = (time.perf_counter() - t) / frames_0
    ref64 = raw[0].astype(np.float64)
    ref1 = O.preprocess(ref64).astype(np.float32)
    mov1 = O.preprocess(raw[1], ref64).astype(np.float32)
    t = time.perf_counter()
    O.get_displacement(ref1, mov1, weight=[0.5, 0.5])
    lat1 = time.perf_counter() - t
    return lat0, lat1


def _passes(solver, h):
    """Sweep counts and assemble flags of a level's passes (frh_api.hip
    plan_passes with the sweep kernel's pass-size limit)."""
    its, lag = solver.iterations, solver.update_lag
    hs = 64 * ((h + 63) // 64)
    pmax = 5 if hs <= 512 else (2 if hs <= 1024 else 1)
    starts = [it for it in range(its) if it == 0 or (it + 1) % lag == 0]
    out = []
    for k, b in enumerate(starts):
        e = starts[k + 1] if k + 1 < len(starts) else its
        first = True
        for s in range(b, e, pmax):
            out.append((min(pmax, e - s), first))
            first = False
    return out


def algorithmic_bytes(solver, T):
    """Algorithmic HBM bytes of one run's k_sweep and k_assemble launches.
    k_sweep, per cell and pass: 5 float64 coefficients and 2 float64 start
    values read (not on a level's first pass) and 2 float64 results written.
    k_assemble, per cell: 6C float64 J, C float32 weights and float64
    cu, cv read, + float64 du, dv after the level's first pass, and 5
    float64 coefficients written."""
    C = solver.C
    sweep = assemble = 0
    for (_, h, w, *_rest) in solver.plan.levels:
        cells = T * h * w
        for k, (n, asm) in enumerate(_passes(solver, h)):
            sweep += cells * 8 * (7 + 2 if k > 0 else 5 + 2)
            if asm:
                assemble += cells * (8 * 6 * C + 4 * C + 16 + (16 if k > 0 else 0) + 40)
    return {"sweep": sweep, "assemble": assemble}


def measured_traffic(H, W, C, T):
    """HBM bytes per launch recorded by rocprofv3 for this workload (lines
    ``<kernel> traffic_bytes_per_launch: N`` of the profile summary)."""
    out = {}
    if (H, W, C, T) != (512, 512, 2, 256) or not os.path.exists(TRAFFIC_FILE):
        return out
    with open(TRAFFIC_FILE) as fh:
        for line in fh:
            parts = line.split()
            if len(parts) == 3 and parts[1] == "traffic_bytes_per_launch:":
                out[parts[0]] = int(float(parts[2]))
    return out


def device_copy_gbs(mib=1024, reps=10):
    """Measured HBM bandwidth of a plain device copy (read + write bytes per
    second) on this GPU, timed with events around reps copies of mib MiB."""
    import torch
    n = mib * (1 << 20) // 4
    src = torch.ones(n, dtype=torch.float32, device="cuda")
    dst = torch.empty_like(src)
    dst.copy_(src)
    a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    a.record()
    for _ in range(reps):
        dst.copy_(src)
    b.record()
    b.synchronize()
    gbs = 2 * n * 4 * reps / (a.elapsed_time(b) * 1e-3) / 1e9
    del src, dst
    return round(gbs, 1)


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


def uneven(args):
    """Independent pairs with sizes drawn from {256^2, 512^2, 796x512,
    1024^2} and 1-2 channels, max displacement U[2, 8] px, Poisson peak
    U[100, 1000]: one solver per size and channel count, all solved as one
    work list (or, with --uneven-mode streams, each group's own batched
    launches on concurrent streams); one step = all pairs.  Prints one JSON
    line (pairs/s)."""
    import numpy as np
    import torch
    import pyflowreg_amd as fr
    from pyflowreg_amd import synth
    from pyflowreg_amd.preprocess import preprocess
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; --gpus must be 1"}))
        sys.exit(2)
    dev = torch.device("cuda", 0)
    pairs = synth.uneven_pairs(args.pairs, seed=3000, device=dev)
    groups = {}
    for f, m in pairs:
        groups.setdefault(tuple(f.shape), []).append((f, m))
    solvers = []
    for (H, W, C), items in sorted(groups.items()):
        fx = torch.stack([preprocess(f)[0] for f, _ in items])
        mv = torch.stack([preprocess(m[None], normalization_ref=f)[0] for f, m in items])
        s = fr.DisplacementSolver(H, W, C, max_batch=len(items))
        s.set_reference(None, None)
        solvers.append((s, fx.permute(0, 2, 3, 1).contiguous(), mv.permute(0, 2, 3, 1).contiguous()))

    wl = None
    if args.uneven_mode == "worklist":
        wl = fr.WorkList([s for s, _, _ in solvers])

        def step():
            return wl.run([fx for _, fx, _ in solvers], [mv for _, _, mv in solvers], out_dtype=torch.float32)
    else:
        def step():
            return fr.optical_flow.run_groups(solvers, n_streams=args.streams, out_dtype=torch.float32)
    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        step()
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    print(json.dumps({
        "metric": "pairs/s registered, uneven batch (256^2, 512^2, 796x512, 1024^2; 1-2 channels)",
        "value": round(args.steps * len(pairs) / dt, 3), "unit": "pairs/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(dt / args.steps * 1e3, 3),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64",
        "data": "synthetic pairs (blob scenes, smooth displacement max U[2,8] px, Poisson U[100,1000])",
        "config": {"workload": f"batched uneven: {len(pairs)} independent pairs per step, "
                               + (f"{len(groups)} size/channel groups in one work list (one launch per "
                                  "stage and level step)" if wl is not None else
                                  f"{len(groups)} size/channel groups on {args.streams} streams"),
                   "groups": {f"{H}x{W}x{C}": len(v) for (H, W, C), v in sorted(groups.items())}},
    }))


def grad(args):
    """Differentiable registration: 32 moving frames (512x512, 2 channels)
    against one fixed frame, full pyramid, continuous warp; one step =
    forward (get_displacement + masked SSE of the registered frames) and
    backward to alpha (2), a_data (2) and the moving frames.  Prints one
    JSON line (frames/s through forward + backward)."""
    import torch
    from pyflowreg_amd import autograd as AG
    from pyflowreg_amd import synth
    from pyflowreg_amd.preprocess import preprocess
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; --gpus must be 1"}))
        sys.exit(2)
    dev = torch.device("cuda", 0)
    H = W = args.size
    C = args.channels
    T = args.grad_frames
    frames, _ = synth.video(T + 1, H, W, n_channels=C, seed=1000, device=dev)
    raw_ref = frames[0].double()
    fixed = preprocess(raw_ref)[0].permute(1, 2, 0).double().contiguous()
    moving = preprocess(frames[1:], normalization_ref=raw_ref).permute(0, 2, 3, 1).double().contiguous()
    mask = torch.zeros((H, W), dtype=torch.float64, device=dev)
    mask[8:-8, 8:-8] = 1.0

    def step():
        alpha = torch.tensor([2.0, 2.0], dtype=torch.float64, device=dev, requires_grad=True)
        a_data = torch.tensor([0.45] * C, dtype=torch.float64, device=dev, requires_grad=True)
        mv = moving.clone().requires_grad_(True)
        flow = AG.get_displacement(fixed, mv, alpha=alpha, a_data=a_data)
        loss = AG.registration_loss(fixed, mv, flow, mask)
        loss.backward()
        return loss.detach(), alpha.grad, a_data.grad, mv.grad

    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        loss, g_alpha, g_a, g_m = step()
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    with torch.no_grad():
        AG.get_displacement(fixed, moving, alpha=torch.tensor([2.0, 2.0], dtype=torch.float64, device=dev))
    torch.cuda.synchronize()
    fwd = time.perf_counter() - t1
    print(json.dumps({
        "metric": "frames/s through differentiable registration (forward + backward), 512x512x2ch",
        "value": round(args.steps * T / dt, 3), "unit": "frames/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(dt / args.steps * 1e3, 3),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64",
        "data": "synthetic (blob scenes, smooth random displacement + drift, Poisson noise)",
        "config": {"workload": f"differentiable: {T} frames {H}x{W}x{C} against one fixed frame, full "
                               "pyramid, continuous warp, alpha=2, 50 iterations, update_lag 5, a_data 0.45; "
                               "loss = masked SSE of the registered frames; gradients to alpha, a_data, "
                               "moving frames", "frames_per_step": T},
        "forward_only_ms": round(fwd * 1e3, 3),
        "loss": float(loss), "grad_alpha": [float(x) for x in g_alpha],
        "grad_a_data": [float(x) for x in g_a], "grad_moving_rms": float(g_m.pow(2).mean().sqrt()),
    }))


def compensate(args):
    """compensate_arr end to end: a host recording of --frames frames
    (512x512x2 float32) registered in batches of --batch with w_init updates,
    batches staged through pinned memory; host arrays in and out.  Prints
    one JSON line (frames/s)."""
    import torch
    from pyflowreg_amd import synth
    from pyflowreg_amd.compensate import RegistrationOptions, compensate_arr
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; --gpus must be 1"}))
        sys.exit(2)
    dev = torch.device("cuda", 0)
    H = W = args.size
    C = args.channels
    frames, _ = synth.video(args.frames + 1, H, W, n_channels=C, seed=1000, device=dev)
    video = frames[1:].cpu().numpy()
    ref = frames[0].double().cpu().numpy()
    opt = RegistrationOptions(alpha=(2.0, 2.0), buffer_size=args.batch, sigma=(1.0, 1.0, 0.1))
    compensate_arr(video[: min(args.batch, 32)], ref, opt, max_batch=args.batch)  # warm-up
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        reg, w = compensate_arr(video, ref, opt, max_batch=args.batch)
    dt = time.perf_counter() - t0
    print(json.dumps({
        "metric": "frames/s compensate_arr end to end (512x512x2ch, host arrays in and out)",
        "value": round(args.steps * video.shape[0] / dt, 3), "unit": "frames/s", "n_gpus": 1,
        "steps": args.steps, "warmup": 1, "ms_per_step": round(dt / args.steps * 1e3, 3),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64",
        "data": "synthetic (blob scenes, smooth random displacement + drift, Poisson noise)",
        "config": {"workload": f"compensate_arr: {video.shape[0]} frames {H}x{W}x{C}, batches of {args.batch}, "
                               "w_init from the first 22 frames then the mean of each batch's last 20",
                   "frames": int(video.shape[0])},
    }))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--batch", type=int, default=256, help="frames per launch")
    ap.add_argument("--size", type=int, default=512)
    ap.add_argument("--channels", type=int, default=2)
    ap.add_argument("--plain-frames", type=int, default=3, help="0 skips the plain path")
    ap.add_argument("--graph", action="store_true", help="replay the batched level loop as a hipGraph")
    ap.add_argument("--autotune", action="store_true",
                    help="tune (or load from the JSON cache) the sweeps per wavefront pass of each level")
    ap.add_argument("--cpu-frames", type=int, default=32,
                    help="frames of the CPU throughput sample (at least twice the worker count)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--copy-probe", type=int, default=0, metavar="MIB",
                    help="only run a device copy of MIB MiB (counter calibration)")
    ap.add_argument("--frames", type=int, default=512, help="recording length for the compensate workload")
    ap.add_argument("--workload", choices=("uniform", "uneven", "grad", "compensate"), default="uniform",
                    help="uneven: 512 independent pairs of mixed sizes and channel counts; "
                         "grad: differentiable registration (forward + backward); "
                         "compensate: compensate_arr end to end")
    ap.add_argument("--grad-frames", type=int, default=32, help="moving frames for the grad workload")
    ap.add_argument("--pairs", type=int, default=512)
    ap.add_argument("--streams", type=int, default=4, help="streams for the uneven workload's groups")
    ap.add_argument("--uneven-mode", choices=("worklist", "streams"), default="worklist",
                    help="uneven workload: one work list over all groups, or each group's own launches "
                         "on --streams streams")
    args = ap.parse_args()
    if args.copy_probe:
        copy_probe(args.copy_probe)
        return
    if args.workload == "uneven":
        uneven(args)
        return
    if args.workload == "grad":
        grad(args)
        return
    if args.workload == "compensate":
        compensate(args)
        return
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; --gpus must be 1",
                          "n_gpus": args.gpus}))
        sys.exit(2)

    import numpy as np
    import torch
    from pyflowreg_amd import synth

    H = W = args.size
    C = args.channels
    T = args.batch
    # inputs are generated once on the host, so the oracle and the GPU see
    # the same frames (frame 0 is the reference)
    frames_h, truth_h = synth.video(T + 1, H, W, n_channels=C, seed=1000)
    raw = frames_h.numpy()
    cpu = None
    cpu_w = None
    if not args.no_cpu:
        n_host = os.cpu_count() or 1
        # one GPU's share of the host: at most 16 worker processes, whatever
        # os.cpu_count() reports on a shared multi-GPU host
        workers = min(16, n_host)
        n_cpu = min(max(args.cpu_frames, 2 * workers), T)
        fps_cpu, work, cpu_w = cpu_baseline(raw, n_cpu, workers)
        lat0, lat1 = cpu_latency(raw, 3) if H == 512 and C == 2 else (None, None)
        cpu = {"value": round(fps_cpu, 4), "unit": "frames/s", "cores": workers, "kind": "port",
               "sample": f"first {n_cpu} frames of the batch ({H}x{W}x{C}), {workers} worker processes "
                         f"with 1 thread each (one frame per task, as the multiprocessing executor), "
                         f"{work:.1f} s of CPU work, FP64 oracle (C + scipy median)",
               "os_cpu_count": n_host, "cpu_model": _cpu_model(),
               "latency_s_per_frame_1_thread": {
                   "256x256x1 single pair": None if lat0 is None else round(lat0, 4),
                   "512x512x2": None if lat1 is None else round(lat1, 4)}}

    import pyflowreg_amd as fr
    from pyflowreg_amd.preprocess import preprocess

    dev = torch.device("cuda", 0)
    torch.cuda.set_device(dev)
    frames = frames_h.to(dev)
    raw_ref = frames[0].double()
    batch = frames[1:].contiguous()
    ref_hwc = preprocess(raw_ref)[0].permute(1, 2, 0).contiguous()
    solver = fr.DisplacementSolver(H, W, C, max_batch=T, graph=args.graph)
    solver.set_reference(ref_hwc, [0.5] * C)
    tuned = None
    if args.autotune:
        from pyflowreg_amd import autotune as at
        p0 = preprocess(batch, normalization_ref=raw_ref).permute(0, 2, 3, 1)
        tuned = at.autotune(solver, p0)
        tuned = {"limits": tuned["limits"], "cached": tuned["cached"]}

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

    # kernel timing of one profiled run (HIP events on the launch stream),
    # float64 output for the parity check
    p = preprocess(batch, normalization_ref=raw_ref).permute(0, 2, 3, 1)
    w64 = solver.run(p, out_dtype=torch.float64, profile=True)
    tm = solver.timings()
    alg = algorithmic_bytes(solver, T)
    traffic = measured_traffic(H, W, C, T)
    kernels = {}
    for key, name in (("sweep", "k_sweep"), ("assemble", "k_assemble")):
        ms, n = tm[key]
        gbs = alg[key] / (ms * 1e-3) / 1e9
        kernels[name] = {"bound": "hbm", "achieved": round(gbs, 1), "peak": PEAK_HBM_GBS,
                         "unit": "GB/s", "frac": round(gbs / PEAK_HBM_GBS, 4),
                         "traffic": traffic.get(name),
                         "algorithmic_bytes_per_launch": int(alg[key] / max(n, 1)),
                         "launches": n, "avg_launch_ms": round(ms / max(n, 1), 4),
                         "total_ms": round(ms, 3)}
    dominant = max(kernels, key=lambda k: kernels[k]["total_ms"])
    copy_gbs = device_copy_gbs()
    for k in kernels.values():
        k["frac_of_measured_copy"] = round(k["achieved"] / copy_gbs, 4)
    roofline = dict(kernels[dominant], kernel=dominant, measured_copy_gbs=copy_gbs)

    parity = None
    if cpu_w is not None:
        g = w64[: cpu_w.shape[0]].cpu().numpy()
        d = np.abs(g - cpu_w)
        parity = {"frames": int(cpu_w.shape[0]), "max_abs_px": float(d.max()),
                  "rms_px": float(np.sqrt((d ** 2).mean())), "reference": "FP64 CPU oracle"}
    epe = float(torch.linalg.vector_norm(w64.cpu() - truth_h[1:], dim=-1).mean())

    plain_fps = plain_fps_eager = None
    if args.plain_frames > 0:
        # one frame per call; the level loop replayed as a hipGraph (and,
        # for comparison, launched eagerly)
        pp = preprocess(batch[: args.plain_frames], normalization_ref=raw_ref).permute(0, 2, 3, 1)
        rates = {}
        for use_graph in (True, False):
            plain = fr.DisplacementSolver(H, W, C, max_batch=1, graph=use_graph)
            plain.set_reference(ref_hwc, [0.5] * C)
            plain.run(pp[0], out_dtype=torch.float32)
            torch.cuda.synchronize()
            t1 = time.perf_counter()
            for k in range(args.plain_frames):
                plain.run(pp[k], out_dtype=torch.float32)
            torch.cuda.synchronize()
            rates[use_graph] = args.plain_frames / (time.perf_counter() - t1)
            plain.close()
        plain_fps, plain_fps_eager = rates[True], rates[False]

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
        "plain_fps": round(plain_fps, 3) if plain_fps else None,
        "plain_fps_eager": round(plain_fps_eager, 3) if plain_fps_eager else None,
        "batched_graph": bool(args.graph),
        "autotune": tuned,
        "batched_over_plain": round(fps / plain_fps, 2) if plain_fps else None,
        "parity": parity,
        "epe_vs_truth_px": round(epe, 4),
        "kernel_ms": {k: round(v[0], 3) for k, v in tm.items()},
        "roofline": roofline,
        "kernels": kernels,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
