"""CPU oracle of the Schwarzschild null-geodesic tracer (ctypes view of libbhs_oracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and bench.py's
cpu_baseline leg, always as the checker (or the timed CPU baseline), never by the product package.

Law T restates bhs::physics (src/physics/Geodesic.h:15-140) in FP64; law S restates the thin-disk
march of shaders/blackhole.frag:67-253 in float and double with one pinned operation order
(oracle/lawS_body.inc). Build: ``make -C oracle``.
"""
from __future__ import annotations

import contextlib
import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "libbhs_oracle.so")


def build(force: bool = False) -> str:
    """Compile the oracle library with its Makefile (gcc); returns the library path."""
    if force or not os.path.exists(_LIB_PATH):
        subprocess.run(["make", "-C", _HERE, "-s", "-B" if force else "all"], check=True)
    return _LIB_PATH


class TraceResult(C.Structure):
    _fields_ = [("status", C.c_int32), ("steps", C.c_int32), ("equator_crossings", C.c_int32),
                ("pad_", C.c_int32), ("dir", C.c_double * 3), ("min_r", C.c_double)]


class LawSParams(C.Structure):
    _fields_ = [("max_steps", C.c_int32), ("disk_mode", C.c_int32), ("root64", C.c_int32),
                ("cross_all", C.c_int32), ("step_scale", C.c_float), ("disk_step", C.c_float),
                ("disk_inner", C.c_float), ("disk_outer", C.c_float),
                ("disk_thickness", C.c_float), ("band_rel", C.c_float)]


class LawSRay(C.Structure):
    _fields_ = [("cls", C.c_int32), ("flags", C.c_int32), ("steps", C.c_int32),
                ("n_cross", C.c_int32), ("esc", C.c_float * 3), ("disk_r", C.c_float * 2),
                ("disk_phi", C.c_float * 2), ("disk_g", C.c_float * 2), ("b", C.c_float)]


class Footprint(C.Structure):
    _fields_ = [("e1", C.c_float * 3), ("e2", C.c_float * 3), ("rinv", C.c_float * 3),
                ("area", C.c_float), ("mu", C.c_float)]


class SkyParams(C.Structure):
    """oracle_sky_params (uSkyBrightness, uStarDensity, uStarBrightness, uDopplerStrength)."""
    _fields_ = [("sky_brightness", C.c_float), ("star_density", C.c_float),
                ("star_brightness", C.c_float), ("doppler", C.c_float)]


class Shading(C.Structure):
    """oracle_shading: the disk shading uniforms (src/Renderer.cpp:328-343) and the annulus."""
    _fields_ = [(k, C.c_float) for k in ("density", "opacity", "peak_temp", "brightness",
                                         "noise_scale", "noise_strength", "warmth", "doppler",
                                         "time", "inflow", "inner", "outer")]


LAWS_RAY_DTYPE = np.dtype([("cls", "<i4"), ("flags", "<i4"), ("steps", "<i4"), ("n_cross", "<i4"),
                           ("esc", "<f4", 3), ("disk_r", "<f4", 2), ("disk_phi", "<f4", 2),
                           ("disk_g", "<f4", 2), ("b", "<f4")])
assert LAWS_RAY_DTYPE.itemsize == C.sizeof(LawSRay)
TRACE_DTYPE = np.dtype([("status", "<i4"), ("steps", "<i4"), ("equator_crossings", "<i4"),
                        ("pad_", "<i4"), ("dir", "<f8", 3), ("min_r", "<f8")])
assert TRACE_DTYPE.itemsize == C.sizeof(TraceResult)

_lib = None
_D3 = C.c_double * 3
_F3 = C.c_float * 3


def build_native(out_path: str) -> str:
    """The same oracle source at -O3 -march=native (the timed CPU baseline, built on the host that
    times it) into out_path; returns out_path."""
    subprocess.run(["make", "-C", _HERE, "-s", "native", f"NATIVE_OUT={out_path}"], check=True)
    return out_path


@contextlib.contextmanager
def using(path: str):
    """Every oracle call inside the block goes to the library at `path` (same API)."""
    global _lib
    prev = _lib
    _lib = _bind(path)
    try:
        yield
    finally:
        _lib = prev


def lib():
    global _lib
    if _lib is None:
        build()
        _lib = _bind(_LIB_PATH)
    return _lib


def _bind(path: str):
    """ctypes view of the oracle library at `path`, with every function's argument types."""
    L = C.CDLL(path)
    dp, fp = C.POINTER(C.c_double), C.POINTER(C.c_float)
    L.oracle_critical_impact_parameter.restype = C.c_double
    L.oracle_accel.argtypes = [dp, C.c_double, dp]
    L.oracle_rk4_step.argtypes = [dp, dp, C.c_double, C.c_double]
    L.oracle_initial_velocity.argtypes = [dp, dp, dp]
    L.oracle_impact_parameter.argtypes = [dp, dp]
    L.oracle_impact_parameter.restype = C.c_double
    L.oracle_impact_parameter_about_axis.argtypes = [dp, dp, dp]
    L.oracle_impact_parameter_about_axis.restype = C.c_double
    L.oracle_far_field_direction.argtypes = [dp, dp, C.c_double, dp]
    L.oracle_shadow_angular_radius.argtypes = [C.c_double]
    L.oracle_shadow_angular_radius.restype = C.c_double
    L.oracle_kepler_omega.argtypes = [C.c_double]
    L.oracle_kepler_omega.restype = C.c_double
    L.oracle_redshift_factor.argtypes = [C.c_double, C.c_double, C.c_double]
    L.oracle_redshift_factor.restype = C.c_double
    L.oracle_trace_T.argtypes = [dp, dp, C.c_int32, C.c_double, C.c_double, C.c_double,
                                 C.POINTER(TraceResult)]
    L.oracle_trace_T_batch.argtypes = [dp, dp, C.c_int64, C.c_int32, C.c_double, C.c_double,
                                       C.c_double, C.c_void_p, C.c_int32]
    L.oracle_ray_dir.argtypes = [fp, fp, C.c_float, C.c_int32, C.c_int32, fp, C.c_int32,
                                 C.c_uint64, C.c_uint32, C.c_uint32, C.c_int64, fp]
    L.oracle_pcg.argtypes = [C.c_uint32]
    L.oracle_pcg.restype = C.c_uint32
    for name in ("oracle_lawS_f32", "oracle_lawS_f64", "oracle_lawS_band"):
        getattr(L, name).argtypes = [fp, fp, C.POINTER(LawSParams), C.POINTER(LawSRay)]
    L.oracle_lawS_h_f32.argtypes = [fp, fp]
    L.oracle_lawS_h_f32.restype = C.c_float
    L.oracle_lawS_b_f32.argtypes = [fp, fp]
    L.oracle_lawS_b_f32.restype = C.c_float
    L.oracle_render.argtypes = [fp, fp, C.c_float, C.c_int32, C.c_int32, fp, C.c_int32,
                                C.c_uint64, C.c_uint32, C.c_uint32, C.POINTER(LawSParams),
                                C.c_int32, C.c_int64, C.c_int64, C.c_int64, C.c_void_p,
                                C.c_int32]
    L.oracle_march_batch.argtypes = [fp, fp, C.c_int64, C.POINTER(LawSParams), C.c_int32,
                                     C.c_void_p, C.c_int32]
    L.oracle_ray_dir_at.argtypes = [fp, fp, C.c_float, C.c_int32, C.c_int32, fp, C.c_int32,
                                    C.c_uint64, C.c_uint32, C.c_uint32, C.c_int64, C.c_int64,
                                    C.c_int64, fp]
    L.oracle_make_footprint.argtypes = [fp, fp, fp, C.c_float, C.POINTER(Footprint)]
    vp = C.c_void_p
    L.oracle_sky_footprint.argtypes = [fp, fp, C.c_float, C.c_int32, C.c_int32, fp, C.c_int32,
                                       C.c_uint64, C.c_uint32, C.c_uint32, C.c_int32, vp, vp,
                                       vp, vp, vp, vp, vp, vp, vp, vp]
    L.oracle_gas_lens_table.argtypes = [C.c_float, C.c_float, C.c_int32, C.c_int32,
                                        C.c_int32, C.c_float, C.c_void_p]
    L.oracle_atan2f.argtypes = [C.c_float, C.c_float]
    L.oracle_atan2f.restype = C.c_float
    L.oracle_rsqrtf.argtypes = [C.c_float]
    L.oracle_rsqrtf.restype = C.c_float
    L.oracle_rsqrt_d.argtypes = [C.c_double]
    L.oracle_rsqrt_d.restype = C.c_double
    L.oracle_cube_face.argtypes = [fp, fp]
    L.oracle_cube_face.restype = C.c_int
    L.oracle_cube_sample.argtypes = [C.c_void_p, C.c_int32, C.c_int32, fp, fp, fp, fp, fp]
    L.oracle_star_layer.argtypes = [C.c_void_p, fp, C.POINTER(Footprint), C.c_float, C.c_float,
                                    C.c_float, C.c_float, C.c_float, C.c_uint32, C.c_float, fp]
    L.oracle_sample_sky.argtypes = [C.POINTER(SkyParams), C.c_void_p, C.c_int32, C.c_int32,
                                    C.c_void_p, C.c_float, fp, fp, fp, C.POINTER(Footprint),
                                    C.c_float, fp]
    L.oracle_sky_pass.argtypes = [fp, fp, C.c_float, C.c_int32, C.c_int32, fp, C.c_int32,
                                  C.c_uint64, C.c_uint32, C.c_uint32, C.c_int32, C.c_void_p,
                                  C.c_void_p, C.POINTER(SkyParams), C.c_void_p, C.c_int32,
                                  C.c_int32, C.c_void_p, C.c_void_p, C.c_void_p, C.c_int64,
                                  C.c_int64, C.c_int32]
    L.oracle_camera_dirs.argtypes = [fp, fp, C.c_float, C.c_int32, C.c_int32, fp, C.c_int32,
                                     C.c_uint64, C.c_uint32, C.c_uint32, C.c_int64, C.c_int64,
                                     C.c_int64, fp]
    f, vp = C.c_float, C.c_void_p
    for name in ("oracle_exp2f", "oracle_log2f", "oracle_logf", "oracle_expf"):
        getattr(L, name).argtypes = [f]
        getattr(L, name).restype = f
    L.oracle_powf.argtypes = [f, f]
    L.oracle_powf.restype = f
    L.oracle_f2h.argtypes = [f]
    L.oracle_f2h.restype = C.c_uint16
    L.oracle_h2f.argtypes = [C.c_uint16]
    L.oracle_h2f.restype = f
    L.oracle_grad3.argtypes = [C.c_int, C.c_int, C.c_int, C.c_uint32, fp]
    L.oracle_pgnoise3.argtypes = [f, f, f, C.c_int, C.c_uint32]
    L.oracle_pgnoise3.restype = f
    L.oracle_noise_texel.argtypes = [C.c_int, C.c_int, C.c_int, C.c_uint32]
    L.oracle_noise_texel.restype = f
    L.oracle_noise_bake.argtypes = [vp, C.c_uint32, C.c_int32]
    L.oracle_sample_noise.argtypes = [vp, f, f, f]
    L.oracle_sample_noise.restype = f
    L.oracle_blackbody_lut.argtypes = [vp]
    L.oracle_blackbody.argtypes = [vp, f, fp]
    sp = C.POINTER(Shading)
    for name in ("oracle_temp_profile", "oracle_radial_profile"):
        getattr(L, name).argtypes = [sp, f]
        getattr(L, name).restype = f
    L.oracle_disk_noise.argtypes = [sp, vp, f, f, f]
    L.oracle_disk_noise.restype = f
    L.oracle_thin_disk_sample.argtypes = [sp, vp, vp, f, f, f, f, fp]
    L.oracle_shade_rays.argtypes = [fp, fp, f, C.c_int32, C.c_int32, fp, C.c_int32, C.c_uint64,
                                    C.c_uint32, C.c_uint32, C.POINTER(LawSParams), sp, vp, vp,
                                    C.c_int64, C.c_int64, C.c_int64, vp, vp, vp, vp, vp, vp, vp]
    L.oracle_shade_march.argtypes = [fp, fp, f, C.c_int32, C.c_int32, fp, C.c_int32, C.c_uint64,
                                     C.c_uint32, C.c_uint32, C.POINTER(LawSParams), sp, vp, vp,
                                     C.c_int64, C.c_int64, C.c_int64, C.c_int32, vp, vp]
    L.oracle_accumulate_weight.argtypes = [C.c_int32, C.c_int32]
    L.oracle_accumulate_weight.restype = f
    L.oracle_accumulate.argtypes = [vp, vp, C.c_int64, C.c_int32, C.c_int32]
    L.oracle_halton_jitter.argtypes = [C.c_int32, fp]
    L.oracle_volumetric.argtypes = [fp, fp, f, C.c_int32, C.c_int32, fp, C.c_int32, C.c_uint64,
                                    C.c_uint32, C.c_uint32, C.POINTER(LawSParams), sp, vp, vp,
                                    C.c_int64, C.c_int64, C.c_int64, C.c_int32, vp, vp, vp, vp,
                                    vp, vp, C.c_int32]
    L.oracle_vol_ray.argtypes = [fp, fp, C.POINTER(LawSParams), sp, vp, vp, C.c_int32, fp]
    L.oracle_vol_ray.restype = C.c_int32
    L.oracle_vol_sample.argtypes = [sp, f, f, vp, vp, f, f, f, f, f, f, fp, fp]
    L.oracle_hermite_f64.argtypes = [dp, dp, dp, dp, C.c_double, dp]
    return L


# ---------------------------------------------------------------------------------- law T helpers
def _d3(v):
    return _D3(*[float(t) for t in v])


def critical_impact_parameter() -> float:
    return lib().oracle_critical_impact_parameter()


def accel(x, h2):
    o = _D3()
    lib().oracle_accel(_d3(x), h2, o)
    return np.array(o[:])


def rk4_step(x, v, h2, dt):
    xx, vv = _d3(x), _d3(v)
    lib().oracle_rk4_step(xx, vv, h2, dt)
    return np.array(xx[:]), np.array(vv[:])


def initial_velocity(cam, d):
    o = _D3()
    lib().oracle_initial_velocity(_d3(cam), _d3(d), o)
    return np.array(o[:])


def impact_parameter(cam, d) -> float:
    return lib().oracle_impact_parameter(_d3(cam), _d3(d))


def impact_parameter_about_axis(cam, d, axis) -> float:
    return lib().oracle_impact_parameter_about_axis(_d3(cam), _d3(d), _d3(axis))


def far_field_direction(x, v, h2):
    o = _D3()
    lib().oracle_far_field_direction(_d3(x), _d3(v), h2, o)
    return np.array(o[:])


def shadow_angular_radius(r) -> float:
    return lib().oracle_shadow_angular_radius(r)


def kepler_omega(r) -> float:
    return lib().oracle_kepler_omega(r)


def redshift_factor(r_emit, b_axis, r_cam) -> float:
    return lib().oracle_redshift_factor(r_emit, b_axis, r_cam)


def trace(cam, d, max_steps=200000, step_scale=0.005, esc_tol=1e-4, r_esc_min=30.0) -> TraceResult:
    """bhs::physics::trace (src/physics/Geodesic.h:107-140) with its default arguments."""
    res = TraceResult()
    lib().oracle_trace_T(_d3(cam), _d3(d), max_steps, step_scale, esc_tol, r_esc_min, C.byref(res))
    return res


def trace_batch(pos3, dir3, max_steps=200000, step_scale=0.005, esc_tol=1e-4, r_esc_min=30.0,
                threads=8) -> np.ndarray:
    pos3 = np.ascontiguousarray(pos3, dtype=np.float64).reshape(-1, 3)
    dir3 = np.ascontiguousarray(dir3, dtype=np.float64).reshape(-1, 3)
    n = pos3.shape[0]
    out = np.zeros(n, dtype=TRACE_DTYPE)
    dp = C.POINTER(C.c_double)
    lib().oracle_trace_T_batch(pos3.ctypes.data_as(dp), dir3.ctypes.data_as(dp), n, max_steps,
                               step_scale, esc_tol, r_esc_min, out.ctypes.data, threads)
    return out


# ---------------------------------------------------------------------------------- law S
def lawS_params(max_steps=450, step_scale=0.05, disk_step=0.18, disk_inner=6.0, disk_outer=20.0,
                disk_thickness=0.035, disk_mode=1, root64=0, band_rel=0.02,
                cross_all=0) -> LawSParams:
    return LawSParams(max_steps, disk_mode, root64, cross_all, step_scale, disk_step, disk_inner,
                      disk_outer, disk_thickness, band_rel)


_MODES = {"f32": 0, "f64": 1, "band": 2}


def _f(a, n):
    a = np.ascontiguousarray(a, dtype=np.float32).reshape(n)
    return a, a.ctypes.data_as(C.POINTER(C.c_float))


def march(pos, d, params: LawSParams, mode="f32") -> LawSRay:
    """Law S on one ray (float position and direction)."""
    out = LawSRay()
    p, pp = _f(pos, 3)
    dd, dp_ = _f(d, 3)
    fn = {"f32": lib().oracle_lawS_f32, "f64": lib().oracle_lawS_f64,
          "band": lib().oracle_lawS_band}[mode]
    fn(pp, dp_, C.byref(params), C.byref(out))
    return out


def march_batch(pos3, dir3, params: LawSParams, mode="f32", threads=8) -> np.ndarray:
    """Law S on n explicit rays ((n,3) float32 positions and directions)."""
    pos3 = np.ascontiguousarray(pos3, dtype=np.float32).reshape(-1, 3)
    dir3 = np.ascontiguousarray(dir3, dtype=np.float32).reshape(-1, 3)
    n = pos3.shape[0]
    out = np.zeros(n, dtype=LAWS_RAY_DTYPE)
    fp = C.POINTER(C.c_float)
    lib().oracle_march_batch(pos3.ctypes.data_as(fp), dir3.ctypes.data_as(fp), n, C.byref(params),
                             _MODES[mode], out.ctypes.data, threads)
    return out


def lawS_h_f32(pos, d) -> float:
    p, pp = _f(pos, 3)
    dd, dp_ = _f(d, 3)
    return lib().oracle_lawS_h_f32(pp, dp_)


def lawS_b_f32(pos, d) -> float:
    p, pp = _f(pos, 3)
    dd, dp_ = _f(d, 3)
    return lib().oracle_lawS_b_f32(pp, dp_)


def atan2f(y, x) -> float:
    return lib().oracle_atan2f(y, x)


def pcg(v: int) -> int:
    return lib().oracle_pcg(v & 0xFFFFFFFF)


def ray_dir(cam, pixel: int):
    """Float view direction of `pixel` for a camera dict (see camera_dict)."""
    p, pp = _f(cam["pos"], 3)
    b, bp = _f(cam["basis"], 9)
    j, jp = _f(cam.get("jitter", (0.0, 0.0)), 2)
    o = _F3()
    lib().oracle_ray_dir(pp, bp, cam["tan_half_fov"], cam["width"], cam["height"], jp,
                         int(cam.get("jitter_hashed", 0)), int(cam.get("seed", 0)),
                         int(cam.get("camera_id", 0)), int(cam.get("sample", 0)), pixel, o)
    return np.array(o[:], dtype=np.float32)


def render(cam, params: LawSParams, mode="f32", first=0, count=None, stride=1,
           threads=8) -> np.ndarray:
    """Law S over pixels first, first+stride, ... of a camera dict; returns a LAWS_RAY_DTYPE array."""
    n = cam["width"] * cam["height"]
    if count is None:
        count = (n - first + stride - 1) // stride
    out = np.zeros(count, dtype=LAWS_RAY_DTYPE)
    p, pp = _f(cam["pos"], 3)
    b, bp = _f(cam["basis"], 9)
    j, jp = _f(cam.get("jitter", (0.0, 0.0)), 2)
    lib().oracle_render(pp, bp, cam["tan_half_fov"], cam["width"], cam["height"], jp,
                        int(cam.get("jitter_hashed", 0)), int(cam.get("seed", 0)),
                        int(cam.get("camera_id", 0)), int(cam.get("sample", 0)), C.byref(params),
                        _MODES[mode], first, count, stride, out.ctypes.data, threads)
    return out


def camera_dirs(cam, first=0, count=None, stride=1) -> np.ndarray:
    """(count, 3) float32 view directions of pixels first, first + stride, ... of a camera dict."""
    n = cam["width"] * cam["height"]
    if count is None:
        count = (n - first + stride - 1) // stride
    out = np.zeros((count, 3), np.float32)
    p, pp = _f(cam["pos"], 3)
    b, bp = _f(cam["basis"], 9)
    j, jp = _f(cam.get("jitter", (0.0, 0.0)), 2)
    lib().oracle_camera_dirs(pp, bp, cam["tan_half_fov"], cam["width"], cam["height"], jp,
                             int(cam.get("jitter_hashed", 0)), int(cam.get("seed", 0)),
                             int(cam.get("camera_id", 0)), int(cam.get("sample", 0)), first, count,
                             stride, out.ctypes.data_as(C.POINTER(C.c_float)))
    return out


def ray_dir_at(cam, col: int, row: int):
    """Float view direction of pixel (col, row) of a camera dict; col/row may lie outside the image."""
    p, pp = _f(cam["pos"], 3)
    b, bp = _f(cam["basis"], 9)
    j, jp = _f(cam.get("jitter", (0.0, 0.0)), 2)
    o = _F3()
    lib().oracle_ray_dir_at(pp, bp, cam["tan_half_fov"], cam["width"], cam["height"], jp,
                            int(cam.get("jitter_hashed", 0)), int(cam.get("seed", 0)),
                            int(cam.get("camera_id", 0)), int(cam.get("sample", 0)),
                            row * cam["width"] + col, col, row, o)
    return np.array(o[:], dtype=np.float32)


def make_footprint(d, a, b, omega0: float) -> dict:
    """makeFootprint (shaders/sky.glsl:27-51) of one escape direction and its derivatives."""
    f = Footprint()
    _, dp_ = _f(d, 3)
    _, ap = _f(a, 3)
    _, bp = _f(b, 3)
    lib().oracle_make_footprint(dp_, ap, bp, omega0, C.byref(f))
    return dict(e1=np.array(f.e1[:], np.float32), e2=np.array(f.e2[:], np.float32),
                rinv=np.array(f.rinv[:], np.float32), area=np.float32(f.area),
                mu=np.float32(f.mu))


def pixel_omega0(cam) -> np.float32:
    """uPixelOmega0 = (2 tan(fov/2) / h)^2 in float (src/Renderer.cpp:246-247,331)."""
    pa = np.float32(2.0) * np.float32(cam["tan_half_fov"]) / np.float32(cam["height"])
    return np.float32(pa * pa)


def sky_footprint(cam, flags, esc, coarse=False) -> dict:
    """Footprint image of one camera (shaders/blackhole.frag:297-306 + shaders/sky.glsl:27-51).

    flags: (n,) uint8 ray flags (bit 2 = escaped); esc: (n, 3) float escape directions. Returns
    ddx, ddy, e1, e2, rinv as (n, 3) float32, area and mu as (n,) float32, fallback as (n,) uint8.
    """
    n = cam["width"] * cam["height"]
    flags = np.ascontiguousarray(flags, dtype=np.uint8).reshape(n)
    esc_planes = np.ascontiguousarray(np.asarray(esc, dtype=np.float32).reshape(n, 3).T)
    out = {k: np.zeros((3, n), np.float32) for k in ("ddx", "ddy", "e1", "e2", "rinv")}
    out["area"] = np.zeros(n, np.float32)
    out["mu"] = np.zeros(n, np.float32)
    out["fallback"] = np.zeros(n, np.uint8)
    p, pp = _f(cam["pos"], 3)
    b, bp = _f(cam["basis"], 9)
    j, jp = _f(cam.get("jitter", (0.0, 0.0)), 2)
    lib().oracle_sky_footprint(pp, bp, cam["tan_half_fov"], cam["width"], cam["height"], jp,
                               int(cam.get("jitter_hashed", 0)), int(cam.get("seed", 0)),
                               int(cam.get("camera_id", 0)), int(cam.get("sample", 0)),
                               int(bool(coarse)), flags.ctypes.data, esc_planes.ctypes.data,
                               *(out[k].ctypes.data for k in ("ddx", "ddy", "e1", "e2", "rinv",
                                                             "area", "mu", "fallback")))
    for k in ("ddx", "ddy", "e1", "e2", "rinv"):
        out[k] = np.ascontiguousarray(out[k].T)
    return out


def gas_lens_table(cam_r: float, alpha_crit: float, n_rays=1024, n_phis=1024, max_steps=20000,
                   r_max=2.0e5) -> np.ndarray:
    """Gas lens table (shaders/gas_lens.comp:11-44): (n_phis, n_rays) float32 log r."""
    t = np.zeros((n_phis, n_rays), np.float32)
    lib().oracle_gas_lens_table(cam_r, alpha_crit, n_rays, n_phis, max_steps, r_max,
                                t.ctypes.data)
    return t


# ---------------------------------------------------------------------------------- shading
def shading(density=1.0, opacity=2.0, peak_temp=6000.0, brightness=5.5, noise_scale=1.0,
            noise_strength=0.85, warmth=0.4, doppler=0.0, time=0.0, inflow=0.0, inner=6.0,
            outer=20.0) -> Shading:
    """Shading uniforms; defaults of src/Settings.h:32-45 (thin disk [6, 20])."""
    return Shading(density, opacity, peak_temp, brightness, noise_scale, noise_strength, warmth,
                   doppler, time, inflow, inner, outer)


_tables = {}


def noise_volume(seed=7, threads=8) -> np.ndarray:
    """The 128^3 R16F noise volume (uint16 texels, index (z, y, x)) of noise_bake.comp (cached)."""
    key = ("noise", seed)
    if key not in _tables:
        v = np.zeros((128, 128, 128), np.uint16)
        lib().oracle_noise_bake(v.ctypes.data, seed, threads)
        _tables[key] = v
    return _tables[key]


def blackbody_lut() -> np.ndarray:
    """(256, 4) float32 blackbody table of DiskTextures::bakeBlackbodyLut."""
    if "lut" not in _tables:
        t = np.zeros((256, 4), np.float32)
        lib().oracle_blackbody_lut(t.ctypes.data)
        _tables["lut"] = t
    return _tables["lut"]


def blackbody(T, lut=None) -> np.ndarray:
    lut = blackbody_lut() if lut is None else np.ascontiguousarray(lut, np.float32)
    o = (C.c_float * 4)()
    lib().oracle_blackbody(lut.ctypes.data, T, o)
    return np.array(o[:], np.float32)


def thin_disk_sample(sh: Shading, r, phi, g, r_cam, vol=None, lut=None) -> np.ndarray:
    """thinDiskSample of one crossing: (r, g, b, alpha) float32."""
    vol = noise_volume() if vol is None else vol
    lut = blackbody_lut() if lut is None else lut
    l2 = np.float32(1.0) - np.float32(2.0) / np.float32(r_cam)
    lapse = np.float32(np.sqrt(max(l2, np.float32(1e-4)), dtype=np.float32))
    o = (C.c_float * 4)()
    lib().oracle_thin_disk_sample(C.byref(sh), vol.ctypes.data, lut.ctypes.data, float(lapse),
                                  r, phi, g, o)
    return np.array(o[:], np.float32)


def shade_rays(cam, params: LawSParams, sh: Shading, rec: dict, first=0, stride=1, vol=None,
               lut=None):
    """Thin-disk colour of a camera's rays from their march records (rec: the RayBuffers.numpy()
    layout of pixels first, first + stride, ...): (rgb (n, 3), sky_weight (n,)) float32."""
    vol = noise_volume() if vol is None else vol
    lut = blackbody_lut() if lut is None else lut
    n = len(rec["flags"])
    fl = np.ascontiguousarray(rec["flags"], np.uint8)
    nc = np.ascontiguousarray(np.minimum(rec["n_cross"], 255), np.uint8)
    dr = np.ascontiguousarray(rec["disk_r"], np.float32)
    dp = np.ascontiguousarray(rec["disk_phi"], np.float32)
    dg = np.ascontiguousarray(rec["disk_g"], np.float32)
    rgb = np.zeros((3, n), np.float32)
    sky = np.zeros(n, np.float32)
    p, pp = _f(cam["pos"], 3)
    b, bp = _f(cam["basis"], 9)
    j, jp = _f(cam.get("jitter", (0.0, 0.0)), 2)
    lib().oracle_shade_rays(pp, bp, cam["tan_half_fov"], cam["width"], cam["height"], jp,
                            int(cam.get("jitter_hashed", 0)), int(cam.get("seed", 0)),
                            int(cam.get("camera_id", 0)), int(cam.get("sample", 0)),
                            C.byref(params), C.byref(sh), vol.ctypes.data, lut.ctypes.data, first,
                            n, stride, fl.ctypes.data, nc.ctypes.data, dr.ctypes.data,
                            dp.ctypes.data, dg.ctypes.data, rgb.ctypes.data, sky.ctypes.data)
    return rgb.T.copy(), sky


def shade_march(cam, params: LawSParams, sh: Shading, mode="band", first=0, count=None, stride=1,
                vol=None, lut=None):
    """The same colour marched from scratch with every crossing shaded in the loop (the
    reference's structure): (rgb (n, 3), sky_weight (n,))."""
    vol = noise_volume() if vol is None else vol
    lut = blackbody_lut() if lut is None else lut
    n = cam["width"] * cam["height"]
    if count is None:
        count = (n - first + stride - 1) // stride
    rgb = np.zeros((3, count), np.float32)
    sky = np.zeros(count, np.float32)
    p, pp = _f(cam["pos"], 3)
    b, bp = _f(cam["basis"], 9)
    j, jp = _f(cam.get("jitter", (0.0, 0.0)), 2)
    lib().oracle_shade_march(pp, bp, cam["tan_half_fov"], cam["width"], cam["height"], jp,
                             int(cam.get("jitter_hashed", 0)), int(cam.get("seed", 0)),
                             int(cam.get("camera_id", 0)), int(cam.get("sample", 0)),
                             C.byref(params), C.byref(sh), vol.ctypes.data, lut.ctypes.data,
                             first, count, stride, _MODES[mode], rgb.ctypes.data, sky.ctypes.data)
    return rgb.T.copy(), sky


def accumulate(cur, acc, sample_index: int, cap: int) -> np.ndarray:
    """acc after accumulate.frag with Renderer::accumulatePass's weight (new array)."""
    cur = np.ascontiguousarray(cur, np.float32)
    out = np.ascontiguousarray(acc, np.float32).copy()
    lib().oracle_accumulate(cur.ctypes.data, out.ctypes.data, cur.size, sample_index, cap)
    return out


def accumulate_weight(sample_index: int, cap: int) -> float:
    return lib().oracle_accumulate_weight(sample_index, cap)


def halton_jitter(sample: int) -> np.ndarray:
    o = (C.c_float * 2)()
    lib().oracle_halton_jitter(sample, o)
    return np.array(o[:], np.float32)


def vol_ray(pos, d, params: LawSParams, sh: Shading, mode="f32", vol=None, lut=None) -> dict:
    """One volumetric ray: status, rgb, sky_weight, steps, esc."""
    vol = noise_volume() if vol is None else vol
    lut = blackbody_lut() if lut is None else lut
    o = (C.c_float * 8)()
    p_, pp = _f(pos, 3)
    d_, dp_ = _f(d, 3)
    st = lib().oracle_vol_ray(pp, dp_, C.byref(params), C.byref(sh), vol.ctypes.data,
                              lut.ctypes.data, 1 if mode == "f64" else 0, o)
    return dict(status=st, rgb=np.array(o[0:3], np.float32), sky_weight=np.float32(o[3]),
                steps=int(o[4]), esc=np.array(o[5:8], np.float32))


def vol_sample(sh: Shading, thickness, lapse, p, r, ds, b_axis, rgb_t, vol=None, lut=None):
    """One volumetricDiskStep onto the composite rgb_t = (r, g, b, T): returns (r, g, b, T)."""
    vol = noise_volume() if vol is None else vol
    lut = blackbody_lut() if lut is None else lut
    i = (C.c_float * 4)(*[float(v) for v in rgb_t])
    o = (C.c_float * 4)()
    lib().oracle_vol_sample(C.byref(sh), thickness, lapse, vol.ctypes.data, lut.ctypes.data,
                            float(p[0]), float(p[1]), float(p[2]), r, ds, b_axis, i, o)
    return np.array(o[:], np.float32)


def hermite(x0, m0, x1, m1, t) -> np.ndarray:
    o = _D3()
    lib().oracle_hermite_f64(_d3(x0), _d3(m0), _d3(x1), _d3(m1), t, o)
    return np.array(o[:])


def volumetric(cam, params: LawSParams, sh: Shading, mode="band", first=0, count=None, stride=1,
               vol=None, lut=None, threads=8) -> dict:
    """Volumetric disk render (blackhole.frag DISK_VOLUMETRIC, disk.glsl:81-110) of pixels first,
    first + stride, ... : rgb (n, 3), sky_weight, cls, flags, steps, esc (n, 3)."""
    vol = noise_volume() if vol is None else vol
    lut = blackbody_lut() if lut is None else lut
    n = cam["width"] * cam["height"]
    if count is None:
        count = (n - first + stride - 1) // stride
    o = dict(rgb=np.zeros((3, count), np.float32), sky_weight=np.zeros(count, np.float32),
             cls=np.zeros(count, np.uint8), flags=np.zeros(count, np.uint8),
             steps=np.zeros(count, np.int32), esc=np.zeros((3, count), np.float32))
    p, pp = _f(cam["pos"], 3)
    b, bp = _f(cam["basis"], 9)
    j, jp = _f(cam.get("jitter", (0.0, 0.0)), 2)
    lib().oracle_volumetric(pp, bp, cam["tan_half_fov"], cam["width"], cam["height"], jp,
                            int(cam.get("jitter_hashed", 0)), int(cam.get("seed", 0)),
                            int(cam.get("camera_id", 0)), int(cam.get("sample", 0)),
                            C.byref(params), C.byref(sh), vol.ctypes.data, lut.ctypes.data, first,
                            count, stride, _MODES[mode], *(o[k].ctypes.data for k in
                                                          ("rgb", "sky_weight", "cls", "flags",
                                                           "steps", "esc")), threads)
    o["rgb"] = o["rgb"].T.copy()
    o["esc"] = o["esc"].T.copy()
    return o


# ---------------------------------------------------------------------------------------- sky
def sky_params(sky_brightness=1.0, star_density=1.0, star_brightness=1.0, doppler=0.0) -> SkyParams:
    """The sky's uniforms (src/Settings.h:53-55 defaults; uDopplerStrength 0 = film look)."""
    return SkyParams(sky_brightness, star_density, star_brightness, doppler)


def pack_cube(levels) -> tuple:
    """A cube map for the sky: levels = [array (6, S_l, S_l, 3) per level] -> (texels (m, 4)
    float32, res, n_levels) in the library's layout (level-major, face, row j, column i)."""
    flat = []
    for a in levels:
        a = np.asarray(a, np.float32)
        t = np.zeros(a.shape[:3] + (4,), np.float32)
        t[..., :3] = a
        flat.append(t.reshape(-1, 4))
    return np.ascontiguousarray(np.concatenate(flat)), int(np.asarray(levels[0]).shape[1]), len(levels)


def cube_face(d):
    """Face index and (s, t) of direction d (OpenGL 4.6 Table 8.19)."""
    st = (C.c_float * 2)()
    _, dp_ = _f(d, 3)
    f = lib().oracle_cube_face(dp_, st)
    return f, float(st[0]), float(st[1])


def cube_sample(cube, d, dx, dy):
    """textureGrad on the cube (texels, res, levels) at d with derivatives dx, dy: (rgb, lod)."""
    texels, res, levels = cube
    o = (C.c_float * 3)()
    lod = C.c_float()
    _, dp_ = _f(d, 3)
    _, xp = _f(dx, 3)
    _, yp = _f(dy, 3)
    lib().oracle_cube_sample(texels.ctypes.data, res, levels, dp_, xp, yp, o, C.byref(lod))
    return np.array(o[:], np.float32), float(lod.value)


def _footprint_struct(fp: dict) -> Footprint:
    f = Footprint()
    for k in range(3):
        f.e1[k] = float(fp["e1"][k])
        f.e2[k] = float(fp["e2"][k])
        f.rinv[k] = float(fp["rinv"][k])
    f.area = float(fp["area"])
    f.mu = float(fp["mu"])
    return f


def star_layer(d, fp: dict, N, prob, flux, sig_star, omega0, salt, g_shift, lut=None):
    """starLayer (shaders/sky.glsl:57-106) of direction d with footprint fp (make_footprint)."""
    lut = blackbody_lut() if lut is None else lut
    o = (C.c_float * 3)()
    _, dp_ = _f(d, 3)
    f = _footprint_struct(fp)
    lib().oracle_star_layer(lut.ctypes.data, dp_, C.byref(f), N, prob, flux, sig_star, omega0,
                            salt, g_shift, o)
    return np.array(o[:], np.float32)


def sample_sky(k: SkyParams, d, dx, dy, fp: dict, omega0, rcam, cube=None, lut=None):
    """sampleSky (shaders/sky.glsl:108-119) of one escaped direction."""
    lut = blackbody_lut() if lut is None else lut
    texels, res, levels = cube if cube is not None else (None, 0, 0)
    o = (C.c_float * 3)()
    _, dp_ = _f(d, 3)
    _, xp = _f(dx, 3)
    _, yp = _f(dy, 3)
    f = _footprint_struct(fp)
    lib().oracle_sample_sky(C.byref(k), texels.ctypes.data if texels is not None else None, res,
                            levels, lut.ctypes.data, rcam, dp_, xp, yp, C.byref(f), omega0, o)
    return np.array(o[:], np.float32)


def sky_pass(cam, flags, esc, k: SkyParams, cube=None, weight=None, rgb=None, coarse=False,
             first=0, stride=1, threads=8, lut=None) -> np.ndarray:
    """The sky pass over one camera's planes: rgb (3, n) += weight * sampleSky for escaped pixels
    (every stride-th pixel from first); returns the new (3, n) float32 planes."""
    n = cam["width"] * cam["height"]
    lut = blackbody_lut() if lut is None else lut
    flags = np.ascontiguousarray(flags, dtype=np.uint8).reshape(n)
    esc_planes = np.ascontiguousarray(np.asarray(esc, dtype=np.float32).reshape(n, 3).T)
    out = (np.zeros((3, n), np.float32) if rgb is None
           else np.ascontiguousarray(np.array(rgb, dtype=np.float32).reshape(3, n)))
    w = None if weight is None else np.ascontiguousarray(weight, dtype=np.float32).reshape(n)
    texels, res, levels = cube if cube is not None else (None, 0, 0)
    p, pp = _f(cam["pos"], 3)
    b, bp = _f(cam["basis"], 9)
    j, jp = _f(cam.get("jitter", (0.0, 0.0)), 2)
    lib().oracle_sky_pass(pp, bp, cam["tan_half_fov"], cam["width"], cam["height"], jp,
                          int(cam.get("jitter_hashed", 0)), int(cam.get("seed", 0)),
                          int(cam.get("camera_id", 0)), int(cam.get("sample", 0)), int(bool(coarse)),
                          flags.ctypes.data, esc_planes.ctypes.data, C.byref(k),
                          texels.ctypes.data if texels is not None else None, res, levels,
                          lut.ctypes.data, w.ctypes.data if w is not None else None,
                          out.ctypes.data, first, stride, threads)
    return out
