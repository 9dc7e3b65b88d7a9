"""Benchmark: Schwarzschild geodesic tracing at 1920x1080, 450 steps (r_obs = 30 M, inclination
85 deg, FOV 60 deg, thin disk [6, 20] M, seeded sub-pixel jitter per frame) on one MI355X.

One step = one call through the C-ABI over a batch of 32 jittered 1080p frames (66,355,200 rays;
--frames N for another count, 1 = one frame per call):
  surrogate leg (value): bhs_render_surrogate_batch with the committed model
      blackholerender_amd/models/lawS_450_disk6-20.bhsmlp in table mode (per pose: the network
      residuals and the exact Schwarzschild orbit tabulated over (b/b_crit, psi)), the plain
      integrator (FP32, FP64 band) for every ray the surrogate does not serve;
  plain leg: bhs_render_batch (persistent FP32 march + FP64 band; launch shape from the
      autotuner, then tuned on the batch itself) over the same frames, timed the same way.
single_frame: both legs again at one frame per call (latency-bound).
Accuracy (outside the timed region): the first frame's surrogate render against FP64 renders of
the same rays: escape-direction error of the served rays, class / crossing-count agreement, and
crossing-radius error against the FP64 march's chord and refined crossings.

roofline: the plain march kernel (the path's VALU-bound integrator): algorithmic FP32 flops of its
rays (blackholerender_amd.workload model, from the kernel's own step/crossing/escape counts) / its
average duration from HIP events recorded on its stream inside the timed region. traffic: HBM
bytes per launch from rocprofv3 FETCH_SIZE/WRITE_SIZE (profiles/traffic.json), if measured for the
workload.
footprint: the sky footprint post-pass (bhs_sky_footprint_batch: GLSL-quad derivatives of the
escape direction and makeFootprint, shaders/blackhole.frag:297-306, shaders/sky.glsl:27-51) over
one surrogate render of the step's frames, all output planes, timed by HIP events on its stream;
HBM-bound, algorithmic bytes 82 per pixel.
shade: the thin-disk colour and transparency pass (bhs_shade_thin_disk_batch, default disk settings)
over one surrogate render of the step's frames, and the accumulation of its colour (bhs_accumulate);
HBM-bound, algorithmic bytes 18 per ray plus 12 per recorded crossing (36 per pixel to accumulate).
sky: the sky pass (bhs_sample_sky_batch: the two star layers and a mip-mapped cube map backdrop,
sky.glsl:57-119) over the same render weighted by the shading's transmittance, with its VALU rate
and the oracle's pass over a sample of the first frame, checked bit for bit.
volumetric: the volumetric disk (bhs_render_volumetric_batch) at the High preset on 8 jittered 1080p
frames per call, with the VALU rate of its FP32 march and the oracle's march of a sample beside it.
gas_lens: the gas lens table (bhs_gas_lens_table, shaders/gas_lens.comp:11-44) at the reference's
size, 1024 fan rays x 1024 swept angles, beside the oracle's build of it on one thread; `batch`:
16 such tables for radii 8..1000 M in one launch (bhs_gas_lens_tables).
surrogate.roofline: the surrogate pass (table build + two k_surrogate_tab dispatches) against the
HBM peak (45-byte records written) and the VALU issue peak (wave-instructions per launch from
rocprofv3, profiles/valu.json).
per_ray: the surrogate with per-ray bf16 MFMA network inference on one 1080p frame per call, with
its MFMA fraction.
full_form: the full form (3840x2160, 8 jittered spp, 4 poses = 32 cameras, 265,420,800 rays per call,
700 steps, disk [6, 30]) through both legs, with its own accuracy block.
batch64: the 64-camera batch captured in one HIP graph and replayed per step, beside the eager call
and the plain march kernel's roofline on it.
cpu_baseline: the CPU paths built for this host (law T, the reference's trace() at its defaults,
and law S in FP64) on the 256x256 camera and a 1080p frame, all CPUs of this job and one thread,
best of 5 runs.
--config c5: the full form as the headline instead.
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


def host_cpus() -> dict:
    """CPUs this process may use: its affinity set, capped by the cgroup CPU quota when one is set
    (the host's share for this job), and at most 16, the CPU share of one GPU on the benchmark
    hosts; plus nproc and the CPU model for the record."""
    try:
        aff = len(os.sched_getaffinity(0))
    except AttributeError:
        aff = os.cpu_count() or 1
    quota = None
    try:
        with open("/sys/fs/cgroup/cpu.max") as f:
            q, per = f.read().split()[:2]
            if q != "max":
                quota = max(1, int(int(q) / int(per)))
    except (OSError, ValueError):
        pass
    use = min(aff, quota) if quota else aff
    return {"threads": max(1, min(use, 16)), "affinity": aff, "cgroup_cpus": quota,
            "nproc": os.cpu_count(), "cpu": cpu_model()}


def _best_of(fn, runs):
    """Warm-up call, then the best and median wall time of `runs` calls."""
    import statistics
    fn()
    ts = []
    for _ in range(runs):
        t0 = time.perf_counter()
        fn()
        ts.append(time.perf_counter() - t0)
    return min(ts), statistics.median(ts)


def _stride_for(n, per_ray_s, target_s):
    """Pixel stride that makes one timed run over a camera last about target_s."""
    import math
    return max(1, math.ceil(n * per_ray_s / target_s))


def run_cpu_baseline(runs: int = 5, target_s: float = 0.3) -> dict:
    """The CPU paths on the 256x256 camera and on one 1920x1080 benchmark frame (camera 0 of the
    step), each on all CPUs of this process's share and on one thread, best of `runs` timed runs
    after a warm-up, each run about target_s of wall time (a strided sample of the camera's pixels
    when the whole camera would take longer):
      law_T: the reference's CPU path, bhs::physics::trace (src/physics/Geodesic.h:107-140) at its
             own defaults (maxSteps 200000, stepScale 0.005, escTol 1e-4, rEscMin 30), FP64, on
             the camera's rays (the oracle's restatement, oracle_trace_T);
      law_S: the shader's thin-disk march (shaders/blackhole.frag:77-253) in FP64 with the
             camera's step settings (oracle_lawS_f64), its 1/sqrt as an IEEE square root and
             division, the fastest IEEE form on x86 (the oracle's exact form, the device's
             refined float value, takes about 1.4x as long on the CPU).
    Built for this host by the oracle's Makefile (make native: gcc -O3 -march=native
    -ffp-contract=off -DORACLE_F64_RSQRT_DIV); the prebuilt -O2 x86-64-v3 library, with the exact
    form, if that build fails. value = law_S on the 1080p frame, all
    CPUs."""
    import tempfile
    import numpy as np
    import oracle
    from blackholerender_amd import workload as W
    from blackholerender_amd.camera import PRESETS
    info = host_cpus()
    threads = info["threads"]
    tmp = tempfile.mkdtemp(prefix="bhs_cpu_baseline_")
    try:
        path = oracle.build_native(os.path.join(tmp, "libbhs_oracle_native.so"))
        build = ("gcc -O3 -march=native -ffp-contract=off -fno-fast-math "
                 "-DORACLE_F64_RSQRT_DIV (built on this host)")
    except Exception as e:  # the prebuilt oracle, named in the record
        path = oracle.build()
        build = f"prebuilt gcc -O2 -march=x86-64-v3 (native build failed: {type(e).__name__})"
    table = {}
    with oracle.using(path):
        for cfg in (W.C1, W.C2):
            cam = W.camera_for(cfg, sample=0).as_dict()
            n = cfg["width"] * cfg["height"]
            ss, dstep = PRESETS[cfg["max_steps"]]
            p = oracle.lawS_params(max_steps=cfg["max_steps"], step_scale=ss, disk_step=dstep,
                                   disk_inner=cfg["disk_inner"], disk_outer=cfg["disk_outer"])
            pos = np.asarray(cam["pos"], np.float64)
            for law in ("law_S", "law_T"):
                for th in (threads, 1):
                    probe_stride = max(1, n // (512 * th))
                    if law == "law_S":
                        run = lambda st: oracle.render(cam, p, mode="f64", stride=st, threads=th)
                    else:
                        def run(st, th=th):
                            d = oracle.camera_dirs(cam, 0, None, st).astype(np.float64)
                            return oracle.trace_batch(np.broadcast_to(pos, d.shape), d,
                                                      threads=th)
                    t0 = time.perf_counter()
                    k = len(run(probe_stride))
                    per_ray = (time.perf_counter() - t0) / max(k, 1)
                    stride = _stride_for(n, per_ray, target_s)
                    rays = (n + stride - 1) // stride
                    if law == "law_T":   # the camera's rays are generated outside the timed call
                        d = oracle.camera_dirs(cam, 0, None, stride).astype(np.float64)
                        pp = np.ascontiguousarray(np.broadcast_to(pos, d.shape))
                        fn = lambda: oracle.trace_batch(pp, d, threads=th)
                    else:
                        fn = lambda: oracle.render(cam, p, mode="f64", stride=stride, threads=th)
                    best, med = _best_of(fn, runs)
                    table.setdefault(law, {}).setdefault(cfg["name"], {})[
                        "all" if th == threads else "1_thread"] = {
                        "mrays_per_s": round(rays / best / 1e6, 4), "rays": rays, "stride": stride,
                        "threads": th, "best_s": round(best, 4), "median_s": round(med, 4)}
    try:
        os.remove(path) if path.startswith(tmp) else None
        os.rmdir(tmp)
    except OSError:
        pass
    head = table["law_S"][W.C2["name"]]["all"]
    return {"value": head["mrays_per_s"], "unit": "Mrays/s", "cores": threads, "kind": "port",
            "sample": (f"FP64 law-S march (oracle/bhs_oracle.c, 1/sqrt as IEEE sqrt and "
                       f"division) of {head['rays']} rays (every "
                       f"{head['stride']}th pixel) of the first 1920x1080 frame of the step, "
                       f"{threads} threads, best of {runs} runs; law T and the 256x256 camera "
                       f"beside it in `table`"),
            "build": build, "runs": runs, **{k: v for k, v in info.items() if k != "threads"},
            "table": table}


def _tune_batch(tracer, cams, params, base: dict) -> dict:
    """Launch shape of the plain march for a batch: the single-camera autotuner's choice and the
    persistent shape at 4-8 resident waves per SIMD (block 256, its refill threshold), each timed
    on the batch itself (main-kernel ms over three calls); keeps the fastest."""
    import torch
    import blackholerender_amd as bhs
    out = bhs.RayBuffers(sum(c.width * c.height for c in cams), "cuda:0")
    stream = torch.cuda.current_stream()
    first = (base["mode"], base["block"], base["waves_per_simd"], base["refill_min"])
    cands = [first] + [(1, 256, w, base["refill_min"]) for w in (4, 5, 6, 7, 8)]
    best = None
    for cand in dict.fromkeys(cands):
        tracer.set_tuning(*cand)
        tracer.render_batch(cams, params, out, stream)
        torch.cuda.synchronize()
        tracer.set_timing(True)
        for _ in range(3):
            tracer.render_batch(cams, params, out, stream)
        torch.cuda.synchronize()
        ms = tracer.timing()["march_ms"]
        tracer.set_timing(False)
        if best is None or ms < best[1]:
            best = (cand, ms)
    tracer.set_tuning(*best[0])
    del out
    return dict(mode=best[0][0], block=best[0][1], waves_per_simd=best[0][2],
                refill_min=best[0][3], ms=best[1], tuned_on="the batch")


def _timed(fn, steps, warmup):
    import torch
    for _ in range(warmup):
        fn()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        fn()
    torch.cuda.synchronize()
    return time.perf_counter() - t0


def _angle(a, b):
    import numpy as np
    a = a.astype(np.float64)
    b = b.astype(np.float64)
    return np.arctan2(np.linalg.norm(np.cross(a, b), axis=1), (a * b).sum(1))


def _radius_err(A, R, sel):
    """Relative error of the recorded crossing radii of rays `sel` (both records present)."""
    import numpy as np
    errs = []
    for k in range(2):
        m = sel & (A["n_cross"] > k) & (R["n_cross"] > k)
        if m.any():
            errs.append(np.abs(A["disk_r"][m, k].astype(np.float64) - R["disk_r"][m, k]) /
                        R["disk_r"][m, k])
    e = np.concatenate(errs) if errs else np.zeros(1)
    return {"max": float(e.max()), "p999": float(np.quantile(e, 0.999)),
            "p50": float(np.median(e)), "n": int(e.size)}


def _phi_g_err(A, R, sel):
    """Azimuth error (rad, wrapped to [-pi, pi]) and relative redshift error of the recorded
    crossings of rays `sel` (both records present)."""
    import numpy as np
    ep, eg = [], []
    for k in range(2):
        m = sel & (A["n_cross"] > k) & (R["n_cross"] > k)
        if m.any():
            d = A["disk_phi"][m, k].astype(np.float64) - R["disk_phi"][m, k]
            ep.append(np.abs((d + np.pi) % (2.0 * np.pi) - np.pi))
            eg.append(np.abs(A["disk_g"][m, k].astype(np.float64) / R["disk_g"][m, k] - 1.0))
    out = {}
    for name, e in (("phi_rad", ep), ("g_rel", eg)):
        e = np.concatenate(e) if e else np.zeros(1)
        out[name] = {"max": float(e.max()), "p999": float(np.quantile(e, 0.999)),
                     "p50": float(np.median(e)), "n": int(e.size)}
    return out


def _plain_vs_fp64(P, R):
    """The plain integrator's FP32 rays (outside the FP64 band) against the FP64 march of the same
    rays: class, crossing-count and step-count agreement, escape and crossing errors."""
    import numpy as np
    f32 = (P["flags"] & 4) == 0
    both = f32 & ((P["flags"] & 2) > 0) & ((R["flags"] & 2) > 0)
    e = _angle(P["esc"][both], R["esc"][both]) if both.any() else np.zeros(1)
    n = int(f32.sum())
    return {"rays_fp32": n,
            "class_mismatch": int((P["cls"][f32] != R["cls"][f32]).sum()),
            "n_cross_mismatch": int((P["n_cross"][f32] != R["n_cross"][f32]).sum()),
            "steps_mismatch_rate": round(float((P["steps"][f32] != R["steps"][f32]).mean()), 6),
            "steps_abs_diff_max": int(np.abs(P["steps"][f32] - R["steps"][f32]).max()),
            "escape_err_max_rad": float(e.max()), "escape_err_p50_rad": float(np.median(e)),
            "disk_r_rel_err": _radius_err(P, R, f32), **_phi_g_err(P, R, f32)}


def _accuracy(sur_out, plain_out, ref_out, ref_root_out, n):
    """Served-ray escape error, crossing-radius error and record agreement of a surrogate frame
    against FP64 (chord crossings, and refined crossings ref_root_out)."""
    import numpy as np
    import blackholerender_amd as bhs
    S, R = sur_out, ref_out
    sv = (S["flags"] & bhs.FLAG_SURROGATE) > 0
    esc = sv & ((S["flags"] & 2) > 0) & ((R["flags"] & 2) > 0)
    e = _angle(S["esc"][esc], R["esc"][esc]) if esc.any() else np.zeros(1)
    pe = ((plain_out["flags"] & 2) > 0) & ((R["flags"] & 2) > 0)
    ep = _angle(plain_out["esc"][pe], R["esc"][pe])
    fb = ~sv
    return {
        "served_fraction": round(float(sv.mean()), 5),
        "escape_err_max_rad": float(e.max()), "escape_err_p9999_rad": float(np.quantile(e, 0.9999)),
        "escape_err_p50_rad": float(np.median(e)),
        "class_mismatch_served": int((S["cls"][sv] != R["cls"][sv]).sum()),
        "n_cross_mismatch_served": int((S["n_cross"][sv] != R["n_cross"][sv]).sum()),
        "fallback_identical_to_plain": bool(all(np.array_equal(S[k][fb], plain_out[k][fb])
                                                for k in ("cls", "flags", "n_cross", "steps"))),
        "plain_escape_err_max_rad": float(ep.max()),
        "plain_class_mismatch": int((plain_out["cls"] != R["cls"]).sum()),
        "disk_r_rel_err_served_vs_fp64_refined": _radius_err(S, ref_root_out, sv),
        "disk_r_rel_err_served_vs_fp64_chord": _radius_err(S, R, sv),
        "disk_r_rel_err_plain_vs_fp64_chord": _radius_err(plain_out, R, sv),
        "disk_phi_g_err_served_vs_fp64_refined": _phi_g_err(S, ref_root_out, sv),
        "disk_phi_g_err_served_vs_fp64_chord": _phi_g_err(S, R, sv),
        "plain_fp32_vs_fp64": _plain_vs_fp64(plain_out, R),
        "rays": n,
    }


def _traffic():
    path = os.path.join(ROOT, "profiles", "traffic.json")
    try:
        with open(path) as f:
            return json.load(f)
    except (OSError, ValueError):
        return None


def main(argv=None) -> int:
    ap = argparse.ArgumentParser(description=__doc__)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--config", default="c2", choices=["c2", "c5"])
    ap.add_argument("--frames", type=int, default=32,
                    help="jittered 1080p frames per call (one step = one call)")
    ap.add_argument("--no-single", action="store_true",
                    help="skip the one-frame-per-call record (single_frame)")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--no-accuracy", action="store_true", help="skip the FP64 comparison")
    ap.add_argument("--plain-only", action="store_true", help="time only the plain integrator")
    ap.add_argument("--per-ray", action="store_true",
                    help="surrogate with per-ray network inference instead of the table mode")
    ap.add_argument("--table", default="1024x128",
                    help="table-mode grid N2xN3 over (b/b_crit, psi) features")
    ap.add_argument("--no-tune", action="store_true", help="skip the autotuner (library default)")
    ap.add_argument("--no-full-form", action="store_true", help="skip the full-form record")
    ap.add_argument("--no-c3", action="store_true", help="skip the 64-camera graph record")
    ap.add_argument("--no-per-ray", action="store_true", help="skip the per-ray network record")
    ap.add_argument("--no-volumetric", action="store_true", help="skip the volumetric disk record")
    ap.add_argument("--tuning", default=None,
                    help="fixed launch shape 'mode,block,waves_per_simd,refill_min' (skips the "
                         "autotuner)")
    ap.add_argument("--copy-probe", action="store_true",
                    help="only run a 256 MiB device-to-device copy (calibrates the HBM counters "
                         "FETCH_SIZE / WRITE_SIZE against a known byte count)")
    ap.add_argument("--valu-probe", action="store_true",
                    help="only measure the VALU issue rate per instruction class (bhs_valu_probe)")
    args = ap.parse_args(argv)
    if args.valu_probe:
        import torch
        import blackholerender_amd as bhs
        tr = bhs.Tracer(0)
        cus = torch.cuda.get_device_properties(0).multi_processor_count
        names = ["v_fma_f32", "v_mul_f32", "v_fma_f64", "v_xor_b32 + v_add_u32", "v_rsq_f32"]
        res = {}
        for kind, name in enumerate(names):
            for wps in (4, 8):
                r = tr.valu_probe(kind, wps, 4096)
                g = r["wave_instr"] / (r["ms"] * 1e-3) / 1e9
                res[f"{name} @ {wps} waves/SIMD"] = {
                    "G_wave_instr_per_s": round(g, 1),
                    "cycles_per_wave_instr_per_simd_at_2.4GHz": round(
                        cus * 4 * 2.4e9 / (g * 1e9), 3)}
        print(json.dumps({"valu_issue_probe": res, "cus": cus}), flush=True)
        return 0
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
        print("bench.py measures one GPU: the library has no multi-GPU path (every ray is "
              "independent, so a frame shards by 16x4-pixel tiles or whole cameras across GPUs, one "
              "process each). Run with --gpus 1.", file=sys.stderr)
        return 2

    import torch
    import blackholerender_amd as bhs
    from blackholerender_amd import workload as W
    from blackholerender_amd.surrogate import Surrogate

    tracer = bhs.Tracer(0)
    if args.config == "c5":
        cfg = W.C5
        cams = W.cameras_c5()
        model = "lawS_700_disk6-30.bhsmlp"
        workload = ("full form: 3840x2160, 8 jittered spp, 4 poses (32 cameras, 265,420,800 rays "
                    "per step), 700 steps, r_obs=15M, incl 80deg, FOV 60deg, thin disk [6,30]M")
    else:
        cfg = W.C2
        cams = [W.camera_for(cfg, sample=k) for k in range(args.frames)]
        model = "lawS_450_disk6-20.bhsmlp"
        workload = ("1920x1080 frames (2,073,600 rays each, seeded sub-pixel jitter per frame), "
                    "%d frame(s) per step = %s rays, 450 steps, r_obs=30M, incl 85deg, FOV 60deg, "
                    "thin disk [6,20]M" % (args.frames, format(2073600 * args.frames, ",")))
    params = W.params_for(cfg)
    tuned_single = None
    leg_tune = not (args.tuning or args.no_tune)   # the other batches tune their own shape too
    if args.tuning:
        tracer.set_tuning(*[int(v) for v in args.tuning.split(",")])
        tuned = tracer.get_tuning()
    elif args.no_tune:
        tuned = tracer.get_tuning()
    else:
        tuned = tracer.autotune(cams[0], params, reps=3)
        if len(cams) > 1:
            tuned_single = tuned
            tuned = _tune_batch(tracer, cams, params, tuned)
    sur = None
    tab = None
    if not args.plain_only:
        sur = Surrogate.load(tracer, os.path.join(ROOT, "blackholerender_amd", "models", model))
        if not args.per_ray:
            tab = tuple(int(v) for v in args.table.lower().split("x"))
            sur.set_table(*tab)

    plain_rec, sur_rec, roof = _legs(tracer, sur, cams, params, args.steps, args.warmup)
    res = {
        "metric": "Mrays/s at 1080p/450 steps: surrogate+fallback vs plain, <=1e-4 rad vs FP64",
        "unit": "Mrays/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32",
        "data": "synthetic (seeded per-pixel jitter; committed trained model)",
    }
    tr = _traffic()
    key = f"{args.config}x{len(cams)}"
    if tr and key in tr.get("workloads", {}):
        roof["traffic"] = tr["workloads"][key]["bytes_per_launch"]
        roof["traffic_source"] = tr["workloads"][key]["source"]
    n = sum(c.width * c.height for c in cams)
    config = {"workload": workload, "rays_per_step": n, "parallelism": "single GPU",
              "launch": tuned}
    if sur is None:
        res.update({"value": plain_rec["value"], "ms_per_step": plain_rec["ms_per_step"],
                    "config": dict(config, path="plain"), "roofline": roof, "plain": plain_rec})
    else:
        res.update({
            "value": sur_rec.pop("value"),
            "ms_per_step": sur_rec.pop("ms_per_step"),
            "speedup_vs_plain": sur_rec.pop("speedup_vs_plain"),
            "config": dict(config, path="surrogate+fallback", model=model,
                           surrogate=(f"table {tab[0]}x{tab[1]}" if tab else "per-ray network")),
            "roofline": roof, "plain": plain_rec, "surrogate": sur_rec,
        })
        res["footprint"] = _footprint_leg(tracer, sur, cams, params, args.steps, args.warmup,
                                          cpu=not args.no_cpu)
        res["shade"] = _shade_leg(tracer, sur, cams, params, args.steps, args.warmup,
                                  cpu=not args.no_cpu)
        if args.config == "c2":
            res["sky"] = _sky_leg(tracer, sur, cams, params, args.steps, args.warmup,
                                  cpu=not args.no_cpu)
        if args.config == "c2" and not args.no_volumetric:
            res["volumetric"] = _volumetric_leg(tracer, 8, max(3, args.steps // 4), 2,
                                                cpu=not args.no_cpu)
        if args.config == "c2":
            res["gas_lens"] = _gas_lens_leg(tracer, args.steps, cpu=not args.no_cpu)
        fkey = f"footprint_{args.config}x{len(cams)}"
        if tr and fkey in tr.get("workloads", {}):
            res["footprint"]["roofline"]["traffic"] = tr["workloads"][fkey]["bytes_per_launch"]
            res["footprint"]["roofline"]["traffic_source"] = tr["workloads"][fkey]["source"]
        if args.config == "c2" and len(cams) > 1 and not args.no_single:
            # the same workload one frame per call (latency-bound: the longest serial rays of a
            # frame, FP64 band and fallback marches, set its time)
            if tuned_single is not None:   # the single-camera autotuner's shape for one frame
                tracer.set_tuning(tuned_single["mode"], tuned_single["block"],
                                  tuned_single["waves_per_simd"], tuned_single["refill_min"])
            p1, s1, _ = _legs(tracer, sur, cams[:1], params, args.steps, args.warmup)
            res["single_frame"] = {"value": s1["value"], "ms_per_step": s1["ms_per_step"],
                                   "speedup_vs_plain": s1["speedup_vs_plain"],
                                   "plain_value": p1["value"],
                                   "plain_ms_per_step": p1["ms_per_step"],
                                   "network_part_ms": s1["network_part_ms"],
                                   "surrogate_call_ms": s1["call_ms"],
                                   "plain_kernel_ms": p1["kernel_ms"],
                                   "plain_call_ms": p1["call_ms"]}
            tracer.set_tuning(tuned["mode"], tuned["block"], tuned["waves_per_simd"],
                              tuned["refill_min"])   # back to the batch's shape
        if not args.no_accuracy:
            cam0 = cams[0]
            n0 = cam0.width * cam0.height
            p64 = W.params_for(cfg, precision=bhs.PREC_F64)
            ref = tracer.render(cam0, p64)
            ref_root = tracer.render(cam0, W.params_for(cfg, precision=bhs.PREC_F64, root64=1))
            pl0 = tracer.render(cam0, params)
            s0 = sur.render(cam0, params)
            torch.cuda.synchronize()
            res["accuracy"] = _accuracy(s0.numpy(), pl0.numpy(), ref.numpy(), ref_root.numpy(), n0)
            res["accuracy"]["frame"] = "camera 0 of the step vs FP64 law-S march of the same rays"
        if args.config == "c2":
            res["surrogate"]["roofline"] = _surrogate_roofline(
                {"network_part_ms": res["surrogate"]["network_part_ms"]}, n, key)
            if not args.no_per_ray:
                res["per_ray"] = _per_ray_leg(tracer, sur, cams[0], params,
                                              max(2, args.steps // 4), 2)
        sur.close()
        if args.config == "c2" and not args.no_full_form:
            res["full_form"] = _full_form_leg(tracer, max(3, args.steps // 4), 2,
                                              accuracy=not args.no_accuracy, tune=leg_tune)
    if args.config == "c2" and not args.no_c3:
        res["batch64"] = _c3_leg(tracer, max(3, args.steps // 4), 2, tune=leg_tune)
    if not args.no_cpu and args.config == "c2":
        res["cpu_baseline"] = run_cpu_baseline()
    print(json.dumps(res), flush=True)
    tracer.close()
    return 0


# Sky footprint post-pass, algorithmic HBM bytes per pixel: reads flags (1 B) and the escape
# direction (12 B); writes ddx, ddy, e1, e2, rinv (3 planes each), area, mu (17 floats) and the
# fallback byte.
FOOTPRINT_BYTES_PER_PIXEL = 13 + 17 * 4 + 1


def _footprint_leg(tracer, sur, cams, params, steps, warmup, cpu=True):
    """Times the sky footprint pass (all planes) over one surrogate render of the cameras: HIP
    events on the launch stream around `steps` calls (each call also uploads the camera table of
    a batch). cpu: also the CPU oracle's pass over the first camera (one thread)."""
    import torch
    import blackholerender_amd as bhs
    from blackholerender_amd import workload as W
    n = sum(c.width * c.height for c in cams)
    stream = torch.cuda.current_stream()
    rays = bhs.RayBuffers(n, "cuda:0")
    if len(cams) == 1:
        sur.render(cams[0], params, rays, stream)
        fp = lambda: tracer.sky_footprint(cams[0], rays, fout, stream=stream)
    else:
        sur.render_batch(cams, params, rays, stream)
        fp = lambda: tracer.sky_footprint_batch(cams, rays, fout, stream=stream)
    fout = bhs.FootprintBuffers(n, "cuda:0")
    for _ in range(max(warmup, 1)):
        fp()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record(stream)
    for _ in range(steps):
        fp()
    e1.record(stream)
    e1.synchronize()
    ms = e0.elapsed_time(e1) / steps
    gbs = n * FOOTPRINT_BYTES_PER_PIXEL / (ms * 1e-3) / 1e9
    fallback = float(fout.fallback.float().mean())
    rec = {"kernel": "k_footprint (sky footprint post-pass, all planes)",
           "call_ms": round(ms, 5), "Gpix_per_s": round(n / (ms * 1e-3) / 1e9, 2),
           "bytes_per_pixel": FOOTPRINT_BYTES_PER_PIXEL,
           "roofline": {"bound": "hbm", "achieved": round(gbs, 1), "peak": W.PEAK_HBM_GBS,
                        "unit": "GB/s", "frac": round(gbs / W.PEAK_HBM_GBS, 4)},
           "fallback_fraction": round(fallback, 5)}
    if cpu:
        import oracle as O
        cam0 = cams[0]
        n0 = cam0.width * cam0.height
        fl = rays.flags[:n0].cpu().numpy()
        esc = rays.esc[:, :n0].cpu().numpy().T.copy()
        t0 = time.perf_counter()
        O.sky_footprint(cam0.as_dict(), fl, esc)
        dt = time.perf_counter() - t0
        rec["cpu_oracle"] = {"value": round(n0 / dt / 1e6, 2), "unit": "Mpix/s", "cores": 1,
                             "kind": "port", "sample": f"one {cam0.width}x{cam0.height} frame"}
    del fout, rays
    return rec


# Shading pass, algorithmic HBM bytes: per ray the flags and crossing count (2 B) read and the
# colour and sky weight (16 B) written; per recorded crossing its r, phi, g (12 B) read.
SHADE_BYTES_PER_RAY = 18
SHADE_BYTES_PER_CROSSING = 12
ACCUM_BYTES_PER_PIXEL = 36    # current and running rgb read, running rgb written


def _shade_leg(tracer, sur, cams, params, steps, warmup, cpu=True):
    """Times the thin-disk shading pass (bhs_shade_thin_disk_batch at the reference's default disk
    settings) over one surrogate render of the step's frames, and the accumulation of its colour
    into a running average: HIP events on the launch stream. cpu: the oracle's shading of the
    first camera's records on one thread."""
    import torch
    import blackholerender_amd as bhs
    from blackholerender_amd import workload as W
    n = sum(c.width * c.height for c in cams)
    stream = torch.cuda.current_stream()
    rays = bhs.RayBuffers(n, "cuda:0")
    sur.render_batch(cams, params, rays, stream)
    sh = bhs.disk_shading()
    out = bhs.ShadeBuffers(n, "cuda:0")
    fn = lambda: tracer.shade_thin_disk_batch(cams, rays, params, sh, out, stream)
    for _ in range(max(warmup, 1)):
        fn()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record(stream)
    for _ in range(steps):
        fn()
    e1.record(stream)
    e1.synchronize()
    ms = e0.elapsed_time(e1) / steps
    rec_cross = int(torch.clamp(rays.n_cross.to(torch.int32), max=2).sum())
    many = int((rays.n_cross.to(torch.int32) > 2).sum())
    nbytes = n * SHADE_BYTES_PER_RAY + rec_cross * SHADE_BYTES_PER_CROSSING
    gbs = nbytes / (ms * 1e-3) / 1e9
    acc = torch.zeros_like(out.rgb)
    tracer.accumulate(out.rgb, acc, 0, stream=stream)
    a0, a1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    a0.record(stream)
    for k in range(steps):
        tracer.accumulate(out.rgb, acc, k + 1, stream=stream)
    a1.record(stream)
    a1.synchronize()
    ams = a0.elapsed_time(a1) / steps
    ags = n * ACCUM_BYTES_PER_PIXEL / (ams * 1e-3) / 1e9
    rec = {"kernel": "k_shade_planes + k_shade_complete (thin-disk colour and transparency, "
                     "default disk settings)",
           "call_ms": round(ms, 5), "Gpix_per_s": round(n / (ms * 1e-3) / 1e9, 2),
           "recorded_crossings": rec_cross, "rays_remarched_for_more_crossings": many,
           "roofline": {"bound": "hbm", "achieved": round(gbs, 1), "peak": W.PEAK_HBM_GBS,
                        "unit": "GB/s", "frac": round(gbs / W.PEAK_HBM_GBS, 4),
                        "bytes_per_launch": nbytes},
           "accumulate": {"kernel": "k_accumulate (running average of the rgb planes)",
                          "call_ms": round(ams, 5),
                          "roofline": {"bound": "hbm", "achieved": round(ags, 1),
                                       "peak": W.PEAK_HBM_GBS, "unit": "GB/s",
                                       "frac": round(ags / W.PEAK_HBM_GBS, 4)}}}
    if cpu:
        import oracle as O
        cam0 = cams[0]
        n0 = cam0.width * cam0.height
        d = rays.disk[:, :n0].cpu().numpy()
        r0 = dict(flags=rays.flags[:n0].cpu().numpy(), n_cross=rays.n_cross[:n0].cpu().numpy(),
                  disk_r=d[0:2].T.copy(), disk_phi=d[2:4].T.copy(), disk_g=d[4:6].T.copy())
        op = O.lawS_params(max_steps=params.max_steps, step_scale=params.step_scale,
                           disk_step=params.disk_step, disk_inner=params.disk_inner,
                           disk_outer=params.disk_outer)
        osh = O.shading(inner=params.disk_inner, outer=params.disk_outer)
        O.noise_volume()
        O.blackbody_lut()
        t0 = time.perf_counter()
        O.shade_rays(cam0.as_dict(), op, osh, r0)
        dt = time.perf_counter() - t0
        rec["cpu_oracle"] = {"value": round(n0 / dt / 1e6, 2), "unit": "Mrays/s", "cores": 1,
                             "kind": "port", "sample": f"one {cam0.width}x{cam0.height} frame"}
    del out, rays, acc
    return rec


# Sky pass, algorithmic HBM bytes per escaped pixel: flags (1) and escape direction (12) of the
# pixel (its quad's other lanes come from the same cache lines), the weight (4), the colour read
# and written (24).
SKY_BYTES_PER_PIXEL = 41


def synthetic_cube_faces(res=512, seed=11):
    """A backdrop for the sky pass: a bright band about a tilted great circle (a galaxy's disk)
    with dust-like modulation and a faint floor, on the six faces of a res^2 cube map."""
    import numpy as np
    rng = np.random.default_rng(seed)
    c = (np.arange(res) + 0.5) / res * 2 - 1
    U, V = np.meshgrid(c, c)
    dirs = [(np.ones_like(U), -V, -U), (-np.ones_like(U), -V, U), (U, np.ones_like(U), V),
            (U, -np.ones_like(U), -V), (U, -V, np.ones_like(U)), (-U, -V, -np.ones_like(U))]
    north = np.array([0.2, 0.9, 0.38]); north /= np.linalg.norm(north)
    phase = rng.uniform(0, 6.3, 3)
    faces = np.zeros((6, res, res, 3), np.float32)
    for f, (x, y, z) in enumerate(dirs):
        n = np.sqrt(x * x + y * y + z * z)
        x, y, z = x / n, y / n, z / n
        lat = x * north[0] + y * north[1] + z * north[2]
        band = np.exp(-(lat / 0.12) ** 2)
        dust = 0.6 + 0.4 * np.sin(9 * x + phase[0]) * np.sin(7 * y + phase[1]) * np.sin(8 * z + phase[2])
        faces[f, :, :, 0] = 0.004 + 0.05 * band * dust
        faces[f, :, :, 1] = 0.004 + 0.045 * band * dust
        faces[f, :, :, 2] = 0.006 + 0.04 * band
    return faces


def _sky_leg(tracer, sur, cams, params, steps, warmup, cpu=True):
    """Times the sky pass (bhs_sample_sky_batch: the two star layers and the backdrop cube map,
    512^2 faces with 10 box-filtered levels) over a surrogate render of the step's frames, weighted
    by the thin-disk shading's transmittance: HIP events on the launch stream; VALU
    wave-instructions per launch from rocprofv3 (profiles/valu.json). cpu: the oracle's sky pass
    over a strided sample of the first frame, checked bit for bit against the GPU."""
    import numpy as np
    import torch
    import blackholerender_amd as bhs
    from blackholerender_amd import workload as W
    n = sum(c.width * c.height for c in cams)
    stream = torch.cuda.current_stream()
    rays = bhs.RayBuffers(n, "cuda:0")
    sur.render_batch(cams, params, rays, stream)
    sh = tracer.shade_thin_disk_batch(cams, rays, params, bhs.disk_shading(), stream=stream)
    cube = bhs.SkyCube.from_faces(synthetic_cube_faces(), "cuda:0")
    k = bhs.sky_params()
    rgb = sh.rgb.clone()
    fn = lambda: tracer.sample_sky_batch(cams, rays, k, cube=cube, weight=sh.sky_weight, rgb=rgb,
                                         stream=stream)
    for _ in range(max(warmup, 1)):
        fn()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record(stream)
    for _ in range(steps):
        fn()
    e1.record(stream)
    e1.synchronize()
    ms = e0.elapsed_time(e1) / steps
    lit = int(((rays.flags & 2) != 0).logical_and(sh.sky_weight != 0).sum())
    gbs = lit * SKY_BYTES_PER_PIXEL / (ms * 1e-3) / 1e9
    rec = {"kernel": "k_sky (quad derivatives, footprint, two star layers, cube map backdrop)",
           "call_ms": round(ms, 5), "Gpix_per_s": round(n / (ms * 1e-3) / 1e9, 2),
           "sky_pixels": lit, "cube": "512^2 x 6 faces, 10 levels",
           "hbm": {"bound": "hbm", "achieved": round(gbs, 1), "peak": W.PEAK_HBM_GBS,
                   "unit": "GB/s", "frac": round(gbs / W.PEAK_HBM_GBS, 4),
                   "bytes_per_sky_pixel": SKY_BYTES_PER_PIXEL}}
    pm = _pmc(f"sky_c2x{len(cams)}")
    if pm:
        gi = pm["valu_insts_per_launch"] / (ms * 1e-3) / 1e9
        rec["roofline"] = {"bound": "valu", "achieved": round(gi, 2), "peak": PEAK_VALU_GINST,
                           "unit": "G wave-instructions/s", "frac": round(gi / PEAK_VALU_GINST, 4),
                           "kernel": "k_sky", "valu_insts_per_launch": pm["valu_insts_per_launch"],
                           "source": pm["source"]}
    if cpu:
        import oracle as O
        cam0 = cams[0]
        n0 = cam0.width * cam0.height
        th = host_cpus()["threads"]
        O.blackbody_lut()
        base = sh.rgb[:, :n0].cpu().numpy()
        got = tracer.sample_sky(cam0, _first_camera(rays, n0), k, cube=cube,
                                weight=sh.sky_weight[:n0].contiguous(),
                                rgb=sh.rgb[:, :n0].contiguous(), stream=stream)
        torch.cuda.synchronize()
        stride = 97
        flags0 = rays.flags[:n0].cpu().numpy()
        esc0 = rays.esc[:, :n0].cpu().numpy().T.copy()
        t0 = time.perf_counter()
        want = O.sky_pass(cam0.as_dict(), flags0, esc0, O.sky_params(), cube=cube.numpy(),
                          weight=sh.sky_weight[:n0].cpu().numpy(), rgb=base, stride=stride,
                          threads=th)
        dt = time.perf_counter() - t0
        m = np.count_nonzero(flags0[::stride] & 2)
        rec["cpu_oracle"] = {"value": round(m / dt / 1e6, 4), "unit": "Msky-pixels/s",
                             "cores": th, "kind": "port",
                             "sample": f"every {stride}th pixel of the first frame ({m} sky "
                                       f"pixels; the footprint of the whole frame first), "
                                       f"{dt:.2f} s",
                             "bit_exact_on_sample": bool(np.array_equal(
                                 got.cpu().numpy()[:, ::stride], want[:, ::stride]))}
    del rays, sh, rgb, cube
    return rec


def _first_camera(rays, n0):
    """RayBuffers view of the first n0 rays of a batch (contiguous copies of the planes)."""
    import blackholerender_amd as bhs
    r = bhs.RayBuffers.__new__(bhs.RayBuffers)
    r.n = n0
    for name in ("cls", "flags", "n_cross", "steps", "b"):
        t = getattr(rays, name)
        setattr(r, name, t[:n0].contiguous() if t is not None else None)
    for name in ("esc", "disk"):
        t = getattr(rays, name)
        setattr(r, name, t[:, :n0].contiguous() if t is not None else None)
    return r


def _volumetric_leg(tracer, frames, steps, warmup, cpu=True):
    """The volumetric disk at the High preset (450 steps, disk step 0.18, disk [6, 20], default
    disk settings) on `frames` jittered 1920x1080 frames of the benchmark camera per call
    (bhs_render_volumetric_batch; FP32 with the FP64 band): wall time per call and the FP32 march
    kernel's time (HIP events on its stream); VALU instructions per launch from rocprofv3
    (profiles/valu.json) over that time. cpu: the oracle's volumetric march of a strided sample of
    the first frame on this host's CPUs."""
    import numpy as np
    import torch
    import blackholerender_amd as bhs
    from blackholerender_amd import workload as W
    cams = [W.camera_for(W.C2, sample=k) for k in range(frames)]
    p = bhs.march_params(disk_mode=bhs.DISK_VOLUMETRIC)
    sh = bhs.disk_shading()
    n = sum(c.width * c.height for c in cams)
    stream = torch.cuda.current_stream()
    rays = bhs.RayBuffers(n, "cuda:0")
    out = bhs.ShadeBuffers(n, "cuda:0")
    fn = lambda: tracer.render_volumetric_batch(cams, p, sh, rays=rays, out=out, stream=stream)
    fn()
    torch.cuda.synchronize()
    tracer.set_timing(True)
    wall = _timed(fn, steps, warmup)
    tm = tracer.timing()
    tracer.set_timing(False)
    fn()
    st = tracer.last_stats()
    cls = rays.cls.to(torch.int32)
    rec = {"workload": (f"{frames} jittered 1920x1080 frame(s) per call ({n:,} rays), volumetric "
                        f"disk [6, 20], High preset (450 steps, disk step 0.18)"),
           "value": round(n * steps / wall / 1e6, 2), "unit": "Mrays/s",
           "ms_per_step": round(wall / steps * 1e3, 4), "steps": steps, "warmup": warmup,
           "kernel_ms": round(tm["march_ms"], 5), "call_ms": round(tm["total_ms"], 5),
           "mean_steps": round(st["steps"] / n, 3), "band64_rays": st["band64"],
           "absorbed_fraction": round(float((cls == bhs.CLASS_DISK).float().mean()), 5),
           "escaped_fraction": round(float((cls == bhs.CLASS_ESCAPE).float().mean()), 5)}
    pm = _pmc(f"volumetric_c2x{frames}")
    if pm and tm["march_ms"] > 0:
        gi = pm["valu_insts_per_launch"] / (tm["march_ms"] * 1e-3) / 1e9
        rec["roofline"] = {"bound": "valu", "achieved": round(gi, 2), "peak": PEAK_VALU_GINST,
                           "unit": "G wave-instructions/s", "frac": round(gi / PEAK_VALU_GINST, 4),
                           "kernel": "k_vol_f32", "valu_insts_per_launch":
                           pm["valu_insts_per_launch"], "source": pm["source"]}
    if cpu:
        import oracle as O
        cam0 = cams[0]
        th = host_cpus()["threads"]
        op = O.lawS_params(disk_mode=2)
        osh = O.shading()
        O.noise_volume()
        O.blackbody_lut()
        stride = 37
        t0 = time.perf_counter()
        ro = O.volumetric(cam0.as_dict(), op, osh, mode="band", stride=stride, threads=th)
        dt = time.perf_counter() - t0
        rec["cpu_oracle"] = {"value": round(len(ro["cls"]) / dt / 1e6, 4), "unit": "Mrays/s",
                             "cores": th, "kind": "port",
                             "sample": f"every {stride}th pixel of the first frame "
                                       f"({len(ro['cls'])} rays), {dt:.2f} s"}
        g0 = out.rgb[:, :cam0.width * cam0.height:stride].cpu().numpy().T
        rec["cpu_oracle"]["bit_exact_on_sample"] = bool(np.array_equal(g0, ro["rgb"]))
    del rays, out
    return rec


def _gas_lens_leg(tracer, steps, cpu=True, cam_r=30.0):
    """Times the gas lens table at the reference's size (1024 rays x 1024 swept angles, camera at
    r = 30 M): HIP events on the launch stream around `steps` builds; cpu: the oracle's build of
    the same table on one thread."""
    import torch
    import blackholerender_amd as bhs
    stream = torch.cuda.current_stream()
    ac = bhs.shadow_angular_radius_f32(cam_r)
    out = tracer.gas_lens_table(cam_r, ac, stream=stream)
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record(stream)
    for _ in range(steps):
        tracer.gas_lens_table(cam_r, ac, out=out, stream=stream)
    e1.record(stream)
    e1.synchronize()
    rec = {"kernel": "k_gas_lens (1024-ray fan, 1024 swept angles, r_cam 30)",
           "ms_per_table": round(e0.elapsed_time(e1) / steps, 4),
           "bound": "latency of the longest fan ray (16 waves on 256 CUs)"}
    # 16 tables of the same size in one launch (bhs_gas_lens_tables): radii log-spaced over
    # 8..1000 M, each with its own shadow angle
    radii = [float(8.0 * (1000.0 / 8.0) ** (k / 15.0)) for k in range(16)]
    acs = [bhs.shadow_angular_radius_f32(r) for r in radii]
    tabs = tracer.gas_lens_tables(radii, acs, stream=stream)
    e0.record(stream)
    for _ in range(steps):
        tracer.gas_lens_tables(radii, acs, out=tabs, stream=stream)
    e1.record(stream)
    e1.synchronize()
    ms = e0.elapsed_time(e1) / steps
    rec["batch"] = {"kernel": "k_gas_lens_batch (16 tables of 1024 x 1024, r_cam 8..1000 M)",
                    "ms_per_call": round(ms, 4), "ms_per_table": round(ms / len(radii), 4)}
    del tabs
    if cpu:
        import oracle as O
        t0 = time.perf_counter()
        O.gas_lens_table(cam_r, ac)
        rec["cpu_oracle_ms_per_table"] = round((time.perf_counter() - t0) * 1e3, 2)
        rec["cpu_cores"] = 1
    return rec


# Surrogate pass, algorithmic HBM bytes per ray: the record it writes (cls, flags, n_cross: 1 B
# each; steps 2; escape direction 12; two crossings' r, phi, g 24; b 4). It reads no per-ray input.
RECORD_BYTES_PER_RAY = 45
# MFMA work of one per-ray network evaluation, padded to the 16x16x32 tiles (the three hidden
# layers' 128x128 products with their 64-lane padding).
MLP_FLOPS_PADDED = 77_824
# VALU issue peak: 256 CUs x 4 SIMDs x 2.4 GHz, one wave64 instruction per SIMD every 4 cycles.
PEAK_VALU_GINST = 256 * 4 * 2.4 / 4


def _pmc(key):
    path = os.path.join(ROOT, "profiles", "valu.json")
    try:
        with open(path) as f:
            return json.load(f).get("workloads", {}).get(key)
    except (OSError, ValueError):
        return None


def _surrogate_roofline(sur_rec, n, key):
    """Rooflines of the surrogate pass (the table build and the two k_surrogate_tab dispatches,
    timed together by HIP events on the call's stream): HBM bytes of the records it writes, and
    VALU wave-instructions per launch from rocprofv3 (profiles/valu.json) over the same time."""
    ms = sur_rec["network_part_ms"]
    gbs = n * RECORD_BYTES_PER_RAY / (ms * 1e-3) / 1e9
    rec = {"kernel": "surrogate pass (k_tab_range, k_tab_phi, k_table, k_chunk_class, 2 x "
                     "k_surrogate_tab, k_stats_reduce)",
           "pass_ms": ms,
           "hbm": {"bound": "hbm", "achieved": round(gbs, 1), "peak": 8000.0, "unit": "GB/s",
                   "frac": round(gbs / 8000.0, 4), "bytes_per_ray": RECORD_BYTES_PER_RAY}}
    pm = _pmc(key)
    if pm:
        gi = pm["valu_insts_per_launch"] / (ms * 1e-3) / 1e9
        rec["valu"] = {"bound": "valu", "achieved": round(gi, 2), "peak": PEAK_VALU_GINST,
                       "unit": "G wave-instructions/s", "frac": round(gi / PEAK_VALU_GINST, 4),
                       "valu_insts_per_launch": pm["valu_insts_per_launch"],
                       "source": pm["source"]}
    return rec


def _full_form_leg(tracer, steps, warmup, accuracy=True, tune=True):
    """The full form (3840x2160, 8 jittered spp, 4 poses: 32 cameras, 265,420,800 rays per call;
    700 steps, disk [6, 30]) through both legs with the full-form model (table mode 1024x128), in
    the launch shape tuned on this batch (tune; else the caller's) and the caller's shape restored
    afterwards; accuracy on camera 0 (8,294,400 rays) against the FP64 march of the same rays."""
    import torch
    import blackholerender_amd as bhs
    from blackholerender_amd import workload as W
    from blackholerender_amd.surrogate import Surrogate
    cams = W.cameras_c5()
    params = W.params_for(W.C5)
    sur = Surrogate.load(tracer, os.path.join(ROOT, "blackholerender_amd", "models",
                                              "lawS_700_disk6-30.bhsmlp"))
    sur.set_table(1024, 128)
    prev = tracer.get_tuning()
    tuned = _tune_batch(tracer, cams, params, prev) if tune else dict(prev, tuned_on="the caller")
    plain_rec, sur_rec, roof = _legs(tracer, sur, cams, params, steps, warmup)
    tracer.set_tuning(prev["mode"], prev["block"], prev["waves_per_simd"], prev["refill_min"])
    n = sum(c.width * c.height for c in cams)
    rec = {"workload": ("full form: 3840x2160, 8 jittered spp, 4 poses (32 cameras, "
                        "265,420,800 rays per step), 700 steps, r_obs=15M, incl 80deg, FOV 60deg, "
                        "thin disk [6,30]M"),
           "model": "lawS_700_disk6-30.bhsmlp", "surrogate_mode": "table 1024x128",
           "launch": tuned, "value": sur_rec.pop("value"), "unit": "Mrays/s",
           "ms_per_step": sur_rec.pop("ms_per_step"),
           "speedup_vs_plain": sur_rec.pop("speedup_vs_plain"), "steps": steps,
           "warmup": warmup, "plain": plain_rec, "surrogate": sur_rec, "roofline": roof,
           "surrogate_roofline": _surrogate_roofline(sur_rec, n, f"c5x{len(cams)}")}
    if accuracy:
        cam0 = cams[0]
        ref = tracer.render(cam0, W.params_for(W.C5, precision=bhs.PREC_F64))
        ref_root = tracer.render(cam0, W.params_for(W.C5, precision=bhs.PREC_F64, root64=1))
        pl0 = tracer.render(cam0, params)
        s0 = sur.render(cam0, params)
        torch.cuda.synchronize()
        rec["accuracy"] = _accuracy(s0.numpy(), pl0.numpy(), ref.numpy(), ref_root.numpy(),
                                    cam0.width * cam0.height)
        rec["accuracy"]["frame"] = "camera 0 of the full form vs FP64 law-S march of the same rays"
        del ref, ref_root, pl0, s0
    sur.close()
    return rec


def _c3_leg(tracer, steps, warmup, tune=True):
    """The 64-camera batch (resolutions 64..1024 per axis, r_obs 8..1000 M, inclinations 0..90,
    FOV 20..90, 200-700 steps with their presets; seeded) captured in one HIP graph
    (bhs_graph_create) and replayed once per step; beside it the eager batch call with the plain
    march kernel's roofline. The launch shape is tuned on this batch (_tune_batch; tune false:
    the caller's) and the caller's shape restored afterwards."""
    import torch
    import blackholerender_amd as bhs
    from blackholerender_amd import workload as W
    cams, params = W.batch_c3()
    n = sum(c.width * c.height for c in cams)
    prev = tracer.get_tuning()
    tuned = _tune_batch(tracer, cams, params, prev) if tune else dict(prev, tuned_on="the caller")
    stream = torch.cuda.current_stream()
    out = bhs.RayBuffers(n, "cuda:0")
    eager = lambda: tracer.render_batch(cams, params, out, stream)
    eager()
    torch.cuda.synchronize()
    tracer.set_timing(True)
    wall_e = _timed(eager, steps, warmup)
    tm = tracer.timing()
    tracer.set_timing(False)
    eager()
    st = tracer.last_stats()
    fl = out.flags.to(torch.int32)
    f32 = (fl & bhs.FLAG_BAND64) == 0
    stp = out.steps_i32()
    ncr = out.n_cross.to(torch.int32)
    flops = W.march_flops(int(f32.sum()), int(stp[f32].sum()),
                          int(((fl & bhs.FLAG_ESCAPED) != 0)[f32].sum()), int(ncr[f32].sum()),
                          int(torch.clamp(ncr, max=2)[f32].sum()))
    tfl = flops / (tm["march_ms"] * 1e-3) / 1e12
    g = tracer.graph(cams, params, out)
    try:
        wall_g = _timed(lambda: g.launch(stream), steps, warmup)
        gst = g.stats()
    finally:
        g.close()
    del out
    tracer.set_tuning(prev["mode"], prev["block"], prev["waves_per_simd"], prev["refill_min"])
    return {"workload": (f"64 cameras in one call, {n:,} rays (seeded sizes 64..1024 per axis, "
                         f"r_obs 8..1000 M, incl 0..90 deg, FOV 20..90 deg, 200/300/450/700 steps)"),
            "launch": tuned,
            "graph_value": round(n * steps / wall_g / 1e6, 2), "unit": "Mrays/s",
            "graph_ms_per_step": round(wall_g / steps * 1e3, 4),
            "eager_value": round(n * steps / wall_e / 1e6, 2),
            "eager_ms_per_step": round(wall_e / steps * 1e3, 4),
            "kernel_ms": round(tm["march_ms"], 5), "band64_rays": st["band64"],
            "graph_stats_equal_eager": all(gst[k] == st[k] for k in ("rays", "steps", "band64")),
            "roofline": {"bound": "valu", "achieved": round(tfl, 3), "peak": W.PEAK_FP32_TFLOPS,
                         "unit": "TFLOP/s", "frac": round(tfl / W.PEAK_FP32_TFLOPS, 4),
                         "kernel": "k_persist (plain FP32 law-S march, batch form)",
                         "flops_per_launch": flops}}


def _per_ray_leg(tracer, sur, cam, params, steps, warmup):
    """The surrogate with per-ray network inference (every candidate ray through the bf16 MFMA
    network, 3->128->128->128->5) on one 1080p frame per call; MFMA fraction = padded network
    flops of the served rays (a lower bound of the evaluations) over the network pass's time
    (HIP events on the call's stream)."""
    import torch
    import blackholerender_amd as bhs
    n = cam.width * cam.height
    stream = torch.cuda.current_stream()
    out = bhs.RayBuffers(n, "cuda:0")
    sur.set_table(0, 0)
    fn = lambda: sur.render(cam, params, out, stream)
    fn()
    torch.cuda.synchronize()
    tracer.set_timing(True)
    wall = _timed(fn, steps, warmup)
    tm = tracer.timing()
    tracer.set_timing(False)
    fn()
    st = tracer.last_stats()
    tf = st["surrogate"] * MLP_FLOPS_PADDED / (tm["march_ms"] * 1e-3) / 1e12
    del out
    return {"value": round(n * steps / wall / 1e6, 2), "unit": "Mrays/s",
            "ms_per_step": round(wall / steps * 1e3, 4), "network_pass_ms": round(tm["march_ms"], 5),
            "served": st["surrogate"],
            "mfma": {"bound": "mfma", "achieved": round(tf, 2), "peak": 2500.0,
                     "unit": "TFLOP/s", "frac": round(tf / 2500.0, 5),
                     "flops_per_evaluation": MLP_FLOPS_PADDED,
                     "evaluations": "served rays (lower bound)"}}


def _legs(tracer, sur, cams, params, steps, warmup):
    """Times the plain integrator and (sur given) the surrogate on the cameras (one call per step:
    bhs_render / bhs_render_surrogate for one camera, the batch calls for several). Returns the
    plain record, the surrogate record (None without sur) and the plain kernel's roofline."""
    import torch
    import blackholerender_amd as bhs
    from blackholerender_amd import workload as W
    n = sum(c.width * c.height for c in cams)
    stream = torch.cuda.current_stream()
    single = len(cams) == 1
    pout = bhs.RayBuffers(n, "cuda:0")
    if single:
        plain = lambda: tracer.render(cams[0], params, pout, stream)
    else:
        plain = lambda: tracer.render_batch(cams, params, pout, stream)
    tracer.set_timing(False)
    plain()
    torch.cuda.synchronize()
    tracer.set_timing(True)
    wall_p = _timed(plain, steps, warmup)
    tm_p = tracer.timing()
    tracer.set_timing(False)
    plain()
    pstats = tracer.last_stats()
    # algorithmic work of the FP32 rays of one launch (the main kernel's rays)
    fl = pout.flags.to(torch.int32)
    f32 = (fl & bhs.FLAG_BAND64) == 0
    st = pout.steps_i32()
    ncr = pout.n_cross.to(torch.int32)
    flops = W.march_flops(int(f32.sum()), int(st[f32].sum()),
                          int(((fl & bhs.FLAG_ESCAPED) != 0)[f32].sum()), int(ncr[f32].sum()),
                          int(torch.clamp(ncr, max=2)[f32].sum()))
    tflops = flops / (tm_p["march_ms"] * 1e-3) / 1e12 if tm_p["march_ms"] > 0 else 0.0
    plain_rate = n * steps / wall_p / 1e6
    plain_rec = {"value": round(plain_rate, 2),
                 "ms_per_step": round(wall_p / steps * 1e3, 4),
                 "kernel_ms": round(tm_p["march_ms"], 5), "call_ms": round(tm_p["total_ms"], 5),
                 "mean_steps": round(int(st.sum()) / n, 3), "band64_rays": pstats["band64"]}
    roof = {"bound": "valu", "achieved": round(tflops, 3), "peak": W.PEAK_FP32_TFLOPS,
            "unit": "TFLOP/s", "frac": round(tflops / W.PEAK_FP32_TFLOPS, 4),
            "traffic": None, "kernel": "k_persist (plain FP32 law-S march)",
            "kernel_ms": round(tm_p["march_ms"], 5), "flops_per_launch": flops}
    del pout
    if sur is None:
        return plain_rec, None, roof
    sout = bhs.RayBuffers(n, "cuda:0")
    if single:
        surf = lambda: sur.render(cams[0], params, sout, stream)
    else:
        surf = lambda: sur.render_batch(cams, params, sout, stream)
    surf()
    torch.cuda.synchronize()
    tracer.set_timing(True)
    wall_s = _timed(surf, steps, warmup)
    tm_s = tracer.timing()
    tracer.set_timing(False)
    surf()
    sstats = tracer.last_stats()
    rate = n * steps / wall_s / 1e6
    sur_rec = {"value": round(rate, 2), "ms_per_step": round(wall_s / steps * 1e3, 4),
               "speedup_vs_plain": round(rate / plain_rate, 3),
               "network_part_ms": round(tm_s["march_ms"], 5),
               "call_ms": round(tm_s["total_ms"], 5),
               "served": sstats["surrogate"], "fallback": sstats["fallback"],
               "band64": sstats["band64"], "fallback_steps": sstats["steps"]}
    return plain_rec, sur_rec, roof


if __name__ == "__main__":
    sys.exit(main())
