"""Benchmark: law-S thin-disk geodesic tracing at 1920x1080, 450 steps (r_obs = 30 M, inclination
85 deg, FOV 60 deg, disk [6, 20] M, seeded per-pixel jitter), on one MI355X.

One step = one full frame (2,073,600 rays) marched through the C-ABI (bhs_render) with the FP32
march and the FP64 band near b_crit. Prints one JSON line (see --help for flags).

roofline: achieved FP32 TFLOP/s of the main march kernel = algorithmic flops of the FP32 rays of one
launch (blackholerender_amd.workload model, from the kernel's own step/crossing/escape counts) /
the kernel's average duration from HIP events recorded on its stream inside the timed region.
cpu_baseline: the CPU oracle's FP64 law-S march of the same frame on this host's cores.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)


def cpu_model() -> str:
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def host_threads() -> int:
    try:
        n = len(os.sched_getaffinity(0))
    except AttributeError:
        n = os.cpu_count() or 1
    return max(1, min(n, int(os.environ.get("BHS_CPU_THREADS", "16"))))


def run_cpu_baseline(cfg, max_seconds: float = 20.0) -> dict:
    """FP64 law-S march of the workload's rays by the CPU oracle (a strided sample if the whole
    frame would take longer than max_seconds)."""
    import numpy as np
    import oracle
    from blackholerender_amd.camera import PRESETS
    from blackholerender_amd.workload import camera_for
    cam = camera_for(cfg).as_dict()
    ss, dstep = PRESETS[cfg["max_steps"]]
    p = oracle.lawS_params(max_steps=cfg["max_steps"], step_scale=ss, disk_step=dstep,
                           disk_inner=cfg["disk_inner"], disk_outer=cfg["disk_outer"])
    n = cfg["width"] * cfg["height"]
    threads = host_threads()
    # probe: every 64th ray
    t0 = time.perf_counter()
    oracle.render(cam, p, mode="f64", stride=64, threads=threads)
    probe = time.perf_counter() - t0
    est_full = probe * 64
    stride = 1 if est_full <= max_seconds else int(np.ceil(est_full / max_seconds))
    t0 = time.perf_counter()
    out = oracle.render(cam, p, mode="f64", stride=stride, threads=threads)
    dt = time.perf_counter() - t0
    return {"value": round(len(out) / dt / 1e6, 4), "unit": "Mrays/s", "cores": threads,
            "kind": "port",
            "sample": (f"FP64 law-S oracle (oracle/bhs_oracle.c, gcc -O2 x86-64-v3) on "
                       f"{len(out)} of the {n} rays (every {stride}th pixel) of the same frame, "
                       f"{threads} threads, {dt:.2f} s, CPU {cpu_model()}")}


def main(argv=None) -> int:
    ap = argparse.ArgumentParser(description=__doc__)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--precision", default="band", choices=["f32", "band", "f64"])
    ap.add_argument("--no-tune", action="store_true", help="skip the autotuner (library default)")
    ap.add_argument("--tuning", default=None,
                    help="fixed launch shape 'mode,block,waves_per_simd,refill_min' (skips the "
                         "autotuner)")
    ap.add_argument("--copy-probe", action="store_true",
                    help="only run a 256 MiB device-to-device copy (calibrates the HBM counters "
                         "FETCH_SIZE / WRITE_SIZE against a known byte count)")
    args = ap.parse_args(argv)
    if args.copy_probe:
        import torch
        x = torch.ones(64 << 20, dtype=torch.float32, device="cuda:0")
        y = torch.empty_like(x)
        for _ in range(3):
            y.copy_(x)
        torch.cuda.synchronize()
        print(json.dumps({"copy_probe_bytes": x.numel() * 4, "dispatches": 3}), flush=True)
        return 0
    if args.gpus != 1:
        print("bench.py measures one GPU: this build has no multi-GPU path (every ray is "
              "independent; see DESIGN.md for how the frame would shard). Run with --gpus 1.",
              file=sys.stderr)
        return 2

    import torch
    import blackholerender_amd as bhs
    from blackholerender_amd import workload as W

    cfg = W.C2
    prec = {"f32": bhs.PREC_F32, "band": bhs.PREC_F32_BAND64, "f64": bhs.PREC_F64}[args.precision]
    tracer = bhs.Tracer(0)
    cam = W.camera_for(cfg)
    params = W.params_for(cfg, precision=prec)
    n = cam.n_rays
    out = bhs.RayBuffers(n, "cuda:0")
    stream = torch.cuda.current_stream()
    if args.tuning:
        tracer.set_tuning(*[int(v) for v in args.tuning.split(",")])
        tuned = tracer.get_tuning()
    elif args.no_tune:
        tuned = tracer.get_tuning()
    else:
        tuned = tracer.autotune(cam, params, out, reps=3)
    for _ in range(args.warmup):
        tracer.render(cam, params, out, stream)
    torch.cuda.synchronize()
    tracer.set_timing(True)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        tracer.render(cam, params, out, stream)
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    tm = tracer.timing()
    tracer.set_timing(False)
    stats = tracer.last_stats()
    wall = t1 - t0
    ms_per_step = wall / args.steps * 1e3

    # algorithmic work of the FP32 rays of one launch (the main kernel's rays)
    fl = out.flags.to(torch.int32)
    f32 = (fl & bhs.FLAG_BAND64) == 0
    steps = out.steps_i32()
    ncr = out.n_cross.to(torch.int32)
    rays32 = int(f32.sum())
    steps32 = int(steps[f32].sum())
    esc32 = int(((fl & bhs.FLAG_ESCAPED) != 0)[f32].sum())
    cross32 = int(ncr[f32].sum())
    rec32 = int(torch.clamp(ncr, max=2)[f32].sum())
    flops = W.march_flops(rays32, steps32, esc32, cross32, rec32)
    tflops = flops / (tm["march_ms"] * 1e-3) / 1e12 if tm["march_ms"] > 0 else 0.0

    res = {
        "metric": "Mrays/s at 1080p/450 steps: surrogate+fallback vs plain, <=1e-4 rad vs FP64",
        "value": round(n * args.steps / wall / 1e6, 2),
        "unit": "Mrays/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32" if prec != bhs.PREC_F64 else "f64",
        "data": "synthetic (seeded per-pixel jitter)",
        "config": {"workload": "plain law-S march, 1920x1080 (2,073,600 rays), 450 steps, "
                               "r_obs=30M, incl 85deg, FOV 60deg, thin disk [6,20]M, "
                               f"precision={args.precision}",
                   "path": "plain", "rays_per_step": n, "parallelism": "single GPU",
                   "launch": tuned},
        "roofline": {"bound": "valu", "achieved": round(tflops, 3),
                     "peak": W.PEAK_FP32_TFLOPS, "unit": "TFLOP/s",
                     "frac": round(tflops / W.PEAK_FP32_TFLOPS, 4), "traffic": None,
                     "kernel": "k_render_f32" if prec != bhs.PREC_F64 else "k_render_f64",
                     "kernel_ms": round(tm["march_ms"], 5), "call_ms": round(tm["total_ms"], 5),
                     "flops_per_launch": flops},
        "march": {"mean_steps": round(int(steps.sum()) / n, 3), "band64_rays": stats["band64"],
                  "escapes": stats["escapes"], "crossings": stats["crossings"]},
    }
    if not args.no_cpu:
        res["cpu_baseline"] = run_cpu_baseline(cfg)
    print(json.dumps(res), flush=True)
    tracer.close()
    return 0


if __name__ == "__main__":
    sys.exit(main())
