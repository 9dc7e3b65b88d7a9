"""Named workloads and the algorithmic work model of the law-S march.

Work model (source-level count of lawS_device.hpp; fused multiply-add = 2 flops, add/sub/mul/div/
sqrt = 1, comparisons/abs/min/max/select = 0; only always-executed work is counted, so the figure is
a lower bound):
  per ray            113  ray generation (47) + setup (65) + band test (1)
  per iteration      145  |x| 6, 1/|v| 7, x.v 5, ds 2, rcyl^2 3, dt 1, RK4 121
                          (RK4: 4 accelerations 4 x (dot 5 + sqrt 1 + 2 mul + div 1 + 3 mul),
                          minus the first dot/sqrt reused from the iteration head, 12 fma-pairs for
                          the stages, dt/6, 2 x 3 x 7 for the x and v updates)
  per escape          50  far-field direction (geodesic.glsl:43-53)
  per crossing        12  chord + |hit.xz|; +29 for each of the first two (atan2 17, redshift 12)
Branch-dependent extras (slab clamp 3-6, escape test 6 per outgoing iteration past rEscMin) are not
counted.
"""
from __future__ import annotations

import math

FLOPS_PER_RAY = 113
FLOPS_PER_STEP = 145
FLOPS_PER_ESCAPE = 50
FLOPS_PER_CROSSING = 12
FLOPS_PER_RECORDED_CROSSING = 29

# MI355X published peaks.
PEAK_FP32_TFLOPS = 157.3       # vector FP32 (the non-packed VALU rate is half of it)
PEAK_FP64_TFLOPS = 78.6
PEAK_BF16_TFLOPS = 2500.0      # dense MFMA
PEAK_HBM_GBS = 8000.0


def march_flops(rays: int, steps: int, escapes: int, crossings: int, recorded_crossings: int) -> int:
    return (FLOPS_PER_RAY * rays + FLOPS_PER_STEP * steps + FLOPS_PER_ESCAPE * escapes +
            FLOPS_PER_CROSSING * crossings + FLOPS_PER_RECORDED_CROSSING * recorded_crossings)


# The reference's CPU-runnable camera and the 1080p benchmark camera: r_obs = 30 M, inclination
# 85 deg, vertical FOV 60 deg, High preset (450 steps, step scale 0.05, disk step 0.18), thin disk
# [6, 20] M with thickness 0.035.
C1 = dict(name="c1_256x256", r_obs=30.0, inclination_deg=85.0, fov_deg=60.0, width=256, height=256,
          max_steps=450, disk_inner=6.0, disk_outer=20.0, jitter_hashed=0, seed=0)
C2 = dict(name="c2_1920x1080", r_obs=30.0, inclination_deg=85.0, fov_deg=60.0, width=1920,
          height=1080, max_steps=450, disk_inner=6.0, disk_outer=20.0, jitter_hashed=1,
          seed=20260601)


# The full form: 3840x2160, 8 jittered samples per pixel, 4 simultaneous camera poses (azimuths 0,
# 90, 180, 270 deg), r_obs = 15 M, inclination 80 deg, vertical FOV 60 deg, Ultra preset (700 steps,
# step scale 0.035, disk step 0.10), thin disk [6, 30] M: 32 cameras, 265,420,800 rays.
C5 = dict(name="c5_3840x2160_8spp_4poses", r_obs=15.0, inclination_deg=80.0, fov_deg=60.0,
          width=3840, height=2160, max_steps=700, disk_inner=6.0, disk_outer=30.0,
          jitter_hashed=1, seed=20260605, spp=8, azimuths=(0.0, 90.0, 180.0, 270.0))


def cameras_c5(cfg=C5, width=None, height=None):
    """The 32 cameras of the full form (pose-major, sample-minor; camera_id = pose)."""
    from .camera import orbit_camera
    w = width or cfg["width"]
    h = height or cfg["height"]
    return [orbit_camera(cfg["r_obs"], cfg["inclination_deg"], cfg["fov_deg"], w, h,
                         jitter_hashed=1, seed=cfg["seed"], camera_id=p, sample=s,
                         azimuth_deg=az)
            for p, az in enumerate(cfg["azimuths"]) for s in range(cfg["spp"])]


def camera_for(cfg, **over):
    from .camera import orbit_camera
    c = dict(cfg)
    c.update(over)
    return orbit_camera(c["r_obs"], c["inclination_deg"], c["fov_deg"], c["width"], c["height"],
                        jitter_hashed=c["jitter_hashed"], seed=c["seed"],
                        camera_id=c.get("camera_id", 0), sample=c.get("sample", 0))


def params_for(cfg, precision=None, root64=0):
    from . import PREC_F32_BAND64, march_params
    return march_params(max_steps=cfg["max_steps"], disk_inner=cfg["disk_inner"],
                        disk_outer=cfg["disk_outer"],
                        precision=PREC_F32_BAND64 if precision is None else precision,
                        root64=root64)


def batch_c3(seed: int = 20260603, n: int = 64):
    """The 64-camera batch: resolutions uniform in {64, 128, ..., 1024} per axis, r_obs
    log-uniform in [8, 1000] M, inclination uniform in [0, 90] deg, vertical FOV uniform in
    [20, 90] deg, max steps uniform in {200, 300, 450, 700} with the matching quality preset, thin
    disk [6, 20] M, seeded per-pixel jitter. Returns (cameras, march params)."""
    import numpy as np
    from . import PREC_F32_BAND64, march_params
    from .camera import orbit_camera
    rng = np.random.Generator(np.random.PCG64(seed))
    cams, params = [], []
    for k in range(n):
        w = int(rng.integers(1, 17)) * 64
        h = int(rng.integers(1, 17)) * 64
        r_obs = float(np.exp(rng.uniform(np.log(8.0), np.log(1000.0))))
        inc = float(rng.uniform(0.0, 90.0))
        fov = float(rng.uniform(20.0, 90.0))
        steps = int(rng.choice([200, 300, 450, 700]))
        cams.append(orbit_camera(r_obs, inc, fov, w, h, jitter_hashed=1, seed=seed, camera_id=k))
        params.append(march_params(max_steps=steps, precision=PREC_F32_BAND64))
    return cams, params


def b_crit() -> float:
    return 3.0 * math.sqrt(3.0)
