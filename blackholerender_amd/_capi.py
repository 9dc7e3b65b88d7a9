"""ctypes view of the C-ABI in include/bhs_trace.h (libbhs_hip.so).

Loading the library needs no GPU; every compute entry point does. The library must be built
(`python -m blackholerender_amd.build` or __graft_entry__.build()); there is no fallback path.
"""
from __future__ import annotations

import ctypes as C
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libbhs_hip.so")

BHS_OK = 0
STATUS_NAMES = {0: "BHS_OK", -1: "BHS_E_INVALID", -2: "BHS_E_DEVICE", -3: "BHS_E_ALLOC",
                -4: "BHS_E_LAUNCH", -5: "BHS_E_IO", -6: "BHS_E_STATE"}

TRACE_ESCAPED, TRACE_CAPTURED, TRACE_MAXSTEPS = 0, 1, 2
CLASS_HORIZON, CLASS_DISK, CLASS_ESCAPE = 0, 1, 2
FLAG_MAXSTEPS, FLAG_ESCAPED, FLAG_BAND64, FLAG_ROOT64, FLAG_SURROGATE = 1, 2, 4, 8, 16
PREC_F32, PREC_F32_BAND64, PREC_F64 = 0, 1, 2
DISK_NONE, DISK_THIN, DISK_VOLUMETRIC = 0, 1, 2


class TraceResult(C.Structure):
    _fields_ = [("status", C.c_int32), ("steps", C.c_int32), ("equator_crossings", C.c_int32),
                ("pad_", C.c_int32), ("dir", C.c_double * 3), ("min_r", C.c_double)]


class Camera(C.Structure):
    _fields_ = [("pos", C.c_float * 3), ("basis", C.c_float * 9), ("tan_half_fov", C.c_float),
                ("width", C.c_int32), ("height", C.c_int32), ("jitter", C.c_float * 2),
                ("jitter_hashed", C.c_int32), ("camera_id", C.c_uint32), ("sample", C.c_uint32),
                ("pad_", C.c_uint32), ("seed", C.c_uint64)]


class MarchParams(C.Structure):
    _fields_ = [("max_steps", C.c_int32), ("disk_mode", C.c_int32), ("precision", C.c_int32),
                ("root64", C.c_int32), ("step_scale", C.c_float), ("disk_step", C.c_float),
                ("disk_inner", C.c_float), ("disk_outer", C.c_float),
                ("disk_thickness", C.c_float), ("band_rel", C.c_float),
                ("cross_all", C.c_int32)]


class RaySoA(C.Structure):
    _fields_ = [("cls_dev", C.c_void_p), ("flags_dev", C.c_void_p), ("n_cross_dev", C.c_void_p),
                ("steps_dev", C.c_void_p), ("esc_dev", C.c_void_p), ("disk_dev", C.c_void_p),
                ("b_dev", C.c_void_p)]


class FootprintSoA(C.Structure):
    _fields_ = [("ddx_dev", C.c_void_p), ("ddy_dev", C.c_void_p), ("e1_dev", C.c_void_p),
                ("e2_dev", C.c_void_p), ("rinv_dev", C.c_void_p), ("area_dev", C.c_void_p),
                ("mu_dev", C.c_void_p), ("fallback_dev", C.c_void_p)]


class DiskShading(C.Structure):
    _fields_ = [(k, C.c_float) for k in ("density", "opacity", "peak_temp", "brightness",
                                         "noise_scale", "noise_strength", "warmth", "doppler",
                                         "time", "inflow")]


class ShadeSoA(C.Structure):
    _fields_ = [("rgb_dev", C.c_void_p), ("sky_weight_dev", C.c_void_p)]


class SkyParams(C.Structure):
    _fields_ = [(k, C.c_float) for k in ("sky_brightness", "star_density", "star_brightness",
                                         "doppler")]


class SkyCube(C.Structure):
    _fields_ = [("texels_dev", C.c_void_p), ("res", C.c_int32), ("levels", C.c_int32)]


class Tuning(C.Structure):
    _fields_ = [("mode", C.c_int32), ("block", C.c_int32), ("waves_per_simd", C.c_int32),
                ("refill_min", C.c_int32)]


class Stats(C.Structure):
    _fields_ = [("rays", C.c_uint64), ("steps", C.c_uint64), ("crossings", C.c_uint64),
                ("escapes", C.c_uint64), ("band64", C.c_uint64), ("fallback", C.c_uint64),
                ("surrogate", C.c_uint64), ("pad_", C.c_uint64)]


MLP_PARAMS = 34181


class MlpInfo(C.Structure):
    _fields_ = [("trained_for", MarchParams), ("dataset_rays", C.c_int64),
                ("adam_steps", C.c_int64), ("tol_phi", C.c_float), ("radius_margin", C.c_float),
                ("logit_margin", C.c_float), ("steps_slack", C.c_float),
                ("steps_max", C.c_uint32 * 16), ("served_cells", C.c_int32), ("pad_", C.c_int32)]


class MlpValidation(C.Structure):
    _fields_ = [("rays", C.c_int64), ("in_domain", C.c_int64), ("served", C.c_int64),
                ("escaping", C.c_int64), ("mismatches", C.c_int64),
                ("phi_err_max", C.c_float), ("phi_err_p99", C.c_float),
                ("phi_err_p50", C.c_float), ("r_err_max", C.c_float), ("r_err_p99", C.c_float),
                ("r_err_p50", C.c_float), ("radius_margin", C.c_float),
                ("served_cells", C.c_int32)]


_lib = None


def load():
    """Load libbhs_hip.so and declare the signatures; raises if it is not built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(f"{LIB_PATH} is not built: run `python -m blackholerender_amd.build`")
    L = C.CDLL(LIB_PATH)
    vp, i32, i64, f64 = C.c_void_p, C.c_int32, C.c_int64, C.c_double
    L.bhs_version.restype = C.c_int
    L.bhs_create.argtypes = [C.c_int, C.POINTER(C.c_int)]
    L.bhs_create.restype = vp
    L.bhs_destroy.argtypes = [vp]
    L.bhs_destroy.restype = None
    L.bhs_last_error.argtypes = [vp]
    L.bhs_last_error.restype = C.c_char_p
    L.bhs_sync.argtypes = [vp]
    L.bhs_trace_rays.argtypes = [vp, vp, vp, i64, i32, f64, f64, f64, vp]
    L.bhs_render.argtypes = [vp, C.POINTER(Camera), C.POINTER(MarchParams), C.POINTER(RaySoA), vp]
    L.bhs_render_noalloc.argtypes = [vp, C.POINTER(Camera), C.POINTER(MarchParams),
                                     C.POINTER(RaySoA), vp]
    L.bhs_shade_thin_disk.argtypes = [vp, C.POINTER(Camera), C.POINTER(MarchParams),
                                      C.POINTER(DiskShading), C.POINTER(RaySoA),
                                      C.POINTER(ShadeSoA), vp]
    L.bhs_shade_thin_disk_batch.argtypes = [vp, vp, vp, C.c_int32, C.POINTER(DiskShading),
                                            C.POINTER(RaySoA), C.POINTER(ShadeSoA), vp]
    L.bhs_render_volumetric.argtypes = [vp, C.POINTER(Camera), C.POINTER(MarchParams),
                                        C.POINTER(DiskShading), vp, C.POINTER(ShadeSoA), vp]
    L.bhs_render_volumetric_batch.argtypes = [vp, vp, vp, C.c_int32, C.POINTER(DiskShading), vp,
                                              C.POINTER(ShadeSoA), vp]
    L.bhs_render_volumetric.restype = C.c_int
    L.bhs_render_volumetric_batch.restype = C.c_int
    L.bhs_accumulate.argtypes = [vp, vp, vp, C.c_int64, C.c_int32, C.c_int32, vp]
    L.bhs_accumulate_weight.argtypes = [C.c_int32, C.c_int32]
    L.bhs_accumulate_weight.restype = C.c_float
    L.bhs_halton_jitter.argtypes = [C.c_int32, C.POINTER(C.c_float)]
    L.bhs_halton_jitter.restype = None
    L.bhs_disk_textures.argtypes = [vp, vp, vp]
    L.bhs_sample_sky.argtypes = [vp, C.POINTER(Camera), C.POINTER(RaySoA), i32,
                                 C.POINTER(SkyParams), vp, vp, vp, vp]
    L.bhs_sample_sky_batch.argtypes = [vp, vp, i32, C.POINTER(RaySoA), i32, C.POINTER(SkyParams),
                                       vp, vp, vp, vp]
    L.bhs_sample_sky.restype = C.c_int
    L.bhs_sample_sky_batch.restype = C.c_int
    for name in ("bhs_shade_thin_disk", "bhs_shade_thin_disk_batch", "bhs_accumulate",
                 "bhs_disk_textures"):
        getattr(L, name).restype = C.c_int
    L.bhs_march_rays.argtypes = [vp, vp, vp, i64, C.POINTER(MarchParams), C.POINTER(RaySoA), vp]
    L.bhs_last_stats.argtypes = [vp, C.POINTER(Stats)]
    L.bhs_rsqrt_rn.argtypes = [vp, vp, vp, i64, vp]
    L.bhs_valu_probe.argtypes = [vp, i32, i32, i32, C.POINTER(C.c_double), C.POINTER(C.c_double)]
    L.bhs_set_timing.argtypes = [vp, C.c_int]
    L.bhs_timing.argtypes = [vp, C.POINTER(C.c_double), C.POINTER(C.c_double), C.POINTER(i32)]
    L.bhs_render_batch.argtypes = [vp, vp, vp, i32, C.POINTER(RaySoA), vp]
    L.bhs_sky_footprint.argtypes = [vp, C.POINTER(Camera), C.POINTER(RaySoA), i32,
                                    C.POINTER(FootprintSoA), vp]
    L.bhs_sky_footprint_batch.argtypes = [vp, vp, i32, C.POINTER(RaySoA), i32,
                                          C.POINTER(FootprintSoA), vp]
    L.bhs_gas_lens_table.argtypes = [vp, C.c_float, C.c_float, i32, i32, vp, vp]
    L.bhs_gas_lens_tables.argtypes = [vp, i32, C.POINTER(C.c_float), C.POINTER(C.c_float), i32,
                                      i32, vp, vp]
    L.bhs_graph_create.argtypes = [vp, vp, vp, i32, C.POINTER(RaySoA), C.POINTER(vp)]
    L.bhs_graph_update.argtypes = [vp, vp, vp, vp]
    L.bhs_graph_launch.argtypes = [vp, vp, vp]
    L.bhs_graph_stats.argtypes = [vp, vp, C.POINTER(Stats)]
    L.bhs_graph_destroy.argtypes = [vp]
    L.bhs_graph_destroy.restype = None
    L.bhs_set_tuning.argtypes = [vp, C.POINTER(Tuning)]
    L.bhs_get_tuning.argtypes = [vp, C.POINTER(Tuning)]
    L.bhs_autotune.argtypes = [vp, C.POINTER(Camera), C.POINTER(MarchParams), C.POINTER(RaySoA), i32,
                               C.POINTER(Tuning), C.POINTER(C.c_double)]
    u32, fp = C.c_uint32, C.POINTER(C.c_float)
    L.bhs_mlp_create.argtypes = [vp, C.POINTER(MarchParams), u32, C.POINTER(C.c_int)]
    L.bhs_mlp_create.restype = vp
    L.bhs_mlp_destroy.argtypes = [vp]
    L.bhs_mlp_destroy.restype = None
    L.bhs_mlp_make_dataset.argtypes = [vp, vp, i64, u32]
    L.bhs_mlp_train.argtypes = [vp, vp, i32, i32, C.c_float, C.c_float, vp]
    L.bhs_mlp_grads.argtypes = [vp, vp, vp, vp, vp, i32, vp, fp]
    L.bhs_mlp_forward.argtypes = [vp, vp, vp, i64, vp, vp]
    L.bhs_mlp_validate.argtypes = [vp, vp, i64, u32, C.c_float, C.POINTER(MlpValidation)]
    L.bhs_mlp_get_params.argtypes = [vp, vp, vp, i64]
    L.bhs_mlp_set_params.argtypes = [vp, vp, vp, i64]
    L.bhs_mlp_get_info.argtypes = [vp, vp, C.POINTER(MlpInfo)]
    L.bhs_mlp_save.argtypes = [vp, vp, C.c_char_p]
    L.bhs_mlp_load.argtypes = [vp, C.c_char_p, C.POINTER(C.c_int)]
    L.bhs_mlp_load.restype = vp
    L.bhs_mlp_set_table.argtypes = [vp, vp, i32, i32]
    L.bhs_render_surrogate_batch.argtypes = [vp, vp, vp, vp, i32, C.POINTER(RaySoA), vp]
    L.bhs_render_surrogate.argtypes = [vp, vp, C.POINTER(Camera), C.POINTER(MarchParams),
                                       C.POINTER(RaySoA), vp]
    for name in ("bhs_mlp_make_dataset", "bhs_mlp_train", "bhs_mlp_grads", "bhs_mlp_forward",
                 "bhs_mlp_validate", "bhs_mlp_get_params", "bhs_mlp_set_params",
                 "bhs_mlp_get_info", "bhs_mlp_save", "bhs_mlp_set_table", "bhs_render_surrogate",
                 "bhs_render_surrogate_batch"):
        getattr(L, name).restype = C.c_int
    for name in ("bhs_sync", "bhs_trace_rays", "bhs_render", "bhs_render_noalloc",
                 "bhs_march_rays", "bhs_last_stats", "bhs_rsqrt_rn", "bhs_valu_probe",
                 "bhs_set_timing", "bhs_timing", "bhs_set_tuning", "bhs_get_tuning",
                 "bhs_autotune", "bhs_render_batch", "bhs_graph_create", "bhs_graph_update",
                 "bhs_graph_launch", "bhs_graph_stats"):
        getattr(L, name).restype = C.c_int
    _lib = L
    return L


def header_symbols(header_path: str | None = None) -> list[str]:
    """Function names declared in include/bhs_trace.h."""
    import re
    if header_path is None:
        header_path = os.path.join(os.path.dirname(HERE), "include", "bhs_trace.h")
    text = open(header_path).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(bhs_[a-z0-9_]+)\s*\(", text)))
