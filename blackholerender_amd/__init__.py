"""MI355X-native Schwarzschild null-geodesic tracer (host side).

The compute runs in hand-written HIP kernels (libbhs_hip.so, gfx950) reached through the C-ABI of
include/bhs_trace.h; this module is the Python view of that ABI. PyTorch is used only to allocate
device memory and to hand over streams.

Reference interfaces mirrored:
  trace(...)        bhs::physics::trace (src/physics/Geodesic.h:107-140), same defaults and
                    TraceResult fields (status Escaped/Captured/MaxSteps, dir, steps, minR,
                    equatorCrossings); FP64 on the GPU.
  Tracer.gas_lens_table
                    the gas lens table of shaders/gas_lens.comp:11-44 (GasParticles' lens program).
  Tracer.sky_footprint
                    the sky footprint of shaders/blackhole.frag:297-306 and shaders/sky.glsl:27-51
                    as a post-pass over the ray planes of a render.
  Tracer.render     the thin-disk trace pass of shaders/blackhole.frag:67-253 driven by
                    Renderer::tracePass/setTraceUniforms (src/Renderer.cpp:242-387), returning per-ray
                    geometry (class, steps, escape direction, disk crossings r/phi/g) instead of colour.
"""
from __future__ import annotations

import ctypes as C
import threading
from dataclasses import dataclass

import numpy as np

from . import _capi
from ._capi import (CLASS_DISK, CLASS_ESCAPE, CLASS_HORIZON, DISK_NONE, DISK_THIN,
                    DISK_VOLUMETRIC, FLAG_BAND64,
                    FLAG_ESCAPED, FLAG_MAXSTEPS, FLAG_ROOT64, FLAG_SURROGATE,
                    PREC_F32, PREC_F32_BAND64, PREC_F64, TRACE_CAPTURED, TRACE_ESCAPED,
                    TRACE_MAXSTEPS)
from .camera import PRESETS, Camera, look_basis, orbit_camera

__all__ = ["Tracer", "BhsError", "TraceResult", "RayBuffers", "FootprintBuffers", "ShadeBuffers",
           "disk_shading", "halton_jitter", "march_params", "sky_params", "SkyCube",
           "trace", "Camera", "orbit_camera", "look_basis", "PRESETS", "critical_impact_parameter",
           "CLASS_HORIZON", "CLASS_DISK", "CLASS_ESCAPE", "FLAG_MAXSTEPS", "FLAG_ESCAPED",
           "FLAG_BAND64", "FLAG_ROOT64", "FLAG_SURROGATE", "PREC_F32",
           "PREC_F32_BAND64", "PREC_F64", "DISK_NONE", "DISK_THIN", "DISK_VOLUMETRIC",
           "TRACE_ESCAPED",
           "TRACE_CAPTURED", "TRACE_MAXSTEPS"]

kHorizon = 2.0
kPhotonSphere = 3.0
kIsco = 6.0


def critical_impact_parameter() -> float:
    """b_crit = 3 sqrt(3) (src/physics/Geodesic.h:19)."""
    return 3.0 * np.sqrt(3.0)


def shadow_angular_radius_f32(r: float) -> float:
    """The shadow's angular radius for a static observer at r, in float as the host computes
    uAlphaCrit (src/GasParticles.cpp:40-44)."""
    f = np.float32
    r = f(r)
    s = min(f(1.0), f(f(5.19615242) / r) * f(np.sqrt(max(f(1.0) - f(2.0) / r, f(0.0)))))
    a = f(np.arcsin(s))
    return float(a if r >= f(3.0) else f(3.14159265) - a)


class BhsError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"{_capi.STATUS_NAMES.get(code, code)}: {msg}")
        self.code = code


@dataclass
class TraceResult:
    """TraceResult of src/physics/Geodesic.h:97-103."""
    status: int
    dir: np.ndarray
    steps: int
    minR: float
    equatorCrossings: int

    Escaped, Captured, MaxSteps = TRACE_ESCAPED, TRACE_CAPTURED, TRACE_MAXSTEPS


def march_params(max_steps=450, step_scale=None, disk_step=None, disk_inner=6.0, disk_outer=20.0,
                 disk_thickness=0.035, disk_mode=DISK_THIN, precision=PREC_F32_BAND64, root64=0,
                 band_rel=0.02, cross_all=0) -> _capi.MarchParams:
    """bhs_march_params; step_scale/disk_step default to the quality preset of max_steps
    (src/Settings.cpp:15-30), else to the High preset."""
    ss, dstep = PRESETS.get(max_steps, PRESETS[450])
    return _capi.MarchParams(max_steps, disk_mode, precision, root64,
                             ss if step_scale is None else step_scale,
                             dstep if disk_step is None else disk_step, disk_inner, disk_outer,
                             disk_thickness, band_rel, cross_all)


def _camera_struct(cam: Camera) -> _capi.Camera:
    c = _capi.Camera()
    c.pos[:] = [float(v) for v in np.asarray(cam.pos, np.float32)]
    c.basis[:] = [float(v) for v in np.asarray(cam.basis, np.float32)]
    c.tan_half_fov = float(cam.tan_half_fov)
    c.width, c.height = int(cam.width), int(cam.height)
    c.jitter[:] = [float(cam.jitter[0]), float(cam.jitter[1])]
    c.jitter_hashed = int(cam.jitter_hashed)
    c.camera_id = int(cam.camera_id) & 0xFFFFFFFF
    c.sample = int(cam.sample) & 0xFFFFFFFF
    c.seed = int(cam.seed) & 0xFFFFFFFFFFFFFFFF
    return c


def _check_layout(t, dtype: str, shape: tuple, what: str):
    import torch
    if not isinstance(t, torch.Tensor):
        raise ValueError(f"{what}: expected a torch tensor, got {type(t).__name__}")
    if str(t.dtype) != f"torch.{dtype}":
        raise ValueError(f"{what}: dtype {t.dtype}, expected torch.{dtype}")
    if tuple(t.shape) != tuple(shape):
        raise ValueError(f"{what}: shape {tuple(t.shape)}, expected {tuple(shape)}")


def _check_place(t, device: int, what: str):
    if not t.is_contiguous():
        raise ValueError(f"{what}: tensor is not contiguous")
    if t.device.type != "cuda" or t.device.index != device:
        raise ValueError(f"{what}: tensor on {t.device}, expected cuda:{device}")


def _check_tensor(t, dtype: str, shape: tuple, device: int, what: str):
    """Raises ValueError unless t is a contiguous tensor of this dtype and shape on cuda:device."""
    _check_layout(t, dtype, shape, what)
    _check_place(t, device, what)


def _check_rays_like(buf, layout, n: int, device: int, what: str):
    if getattr(buf, "n", None) != n:
        raise ValueError(f"{what}: buffers hold {getattr(buf, 'n', None)} rays, the call needs {n}")
    present = [(name, getattr(buf, name), dtype, planes) for name, dtype, planes in layout
               if getattr(buf, name) is not None]
    for name, t, dtype, planes in present:
        _check_layout(t, dtype, (planes, n) if planes > 1 else (n,), f"{what}.{name}")
    for name, t, _, _ in present:
        _check_place(t, device, f"{what}.{name}")


class RayBuffers:
    """Device structure-of-arrays output (bhs_ray_soa) held as torch tensors."""

    def __init__(self, n: int, device):
        import torch
        self.n = n
        self.cls = torch.empty(n, dtype=torch.uint8, device=device)
        self.flags = torch.empty(n, dtype=torch.uint8, device=device)
        self.n_cross = torch.empty(n, dtype=torch.uint8, device=device)
        self.steps = torch.empty(n, dtype=torch.int16, device=device)   # uint16 bit pattern
        self.esc = torch.empty((3, n), dtype=torch.float32, device=device)
        self.disk = torch.empty((6, n), dtype=torch.float32, device=device)
        self.b = torch.empty(n, dtype=torch.float32, device=device)

    _LAYOUT = (("cls", "uint8", 1), ("flags", "uint8", 1), ("n_cross", "uint8", 1),
               ("steps", "int16", 1), ("esc", "float32", 3), ("disk", "float32", 6),
               ("b", "float32", 1))

    def check(self, n: int, device: int, what: str):
        """Raises ValueError unless these buffers hold exactly n rays, every plane a contiguous
        tensor of its dtype on cuda:device (the C side takes the plane stride from the camera
        sizes, so any other n would read or write out of bounds)."""
        _check_rays_like(self, self._LAYOUT, n, device, what)

    def soa(self) -> _capi.RaySoA:
        return _capi.RaySoA(self.cls.data_ptr(), self.flags.data_ptr(), self.n_cross.data_ptr(),
                            self.steps.data_ptr(), self.esc.data_ptr(), self.disk.data_ptr(),
                            self.b.data_ptr())

    def steps_i32(self):
        import torch
        return self.steps.to(torch.int32) & 0xFFFF

    def numpy(self) -> dict:
        """Host copies: cls, flags, n_cross, steps (int32), esc (n,3), disk_r/phi/g (n,2), b."""
        d = self.disk.cpu().numpy()
        return dict(cls=self.cls.cpu().numpy(), flags=self.flags.cpu().numpy(),
                    n_cross=self.n_cross.cpu().numpy(),
                    steps=self.steps.cpu().numpy().view(np.uint16).astype(np.int32),
                    esc=self.esc.cpu().numpy().T.copy(), disk_r=d[0:2].T.copy(),
                    disk_phi=d[2:4].T.copy(), disk_g=d[4:6].T.copy(), b=self.b.cpu().numpy())


class FootprintBuffers:
    """Device sky footprint planes (bhs_footprint_soa) held as torch tensors; `planes` names the
    planes to allocate (the others are passed as NULL and not written)."""

    ALL = ("ddx", "ddy", "e1", "e2", "rinv", "area", "mu", "fallback")

    def __init__(self, n: int, device, planes=ALL):
        import torch
        self.n = n
        for k in self.ALL:
            t = None
            if k in planes:
                if k in ("area", "mu"):
                    t = torch.empty(n, dtype=torch.float32, device=device)
                elif k == "fallback":
                    t = torch.empty(n, dtype=torch.uint8, device=device)
                else:
                    t = torch.empty((3, n), dtype=torch.float32, device=device)
            setattr(self, k, t)

    def check(self, n: int, device: int, what: str):
        """Raises ValueError unless the allocated planes hold exactly n pixels each, contiguous
        tensors of their dtypes on cuda:device."""
        layout = [(k, "float32" if k not in ("fallback",) else "uint8",
                   1 if k in ("area", "mu", "fallback") else 3) for k in self.ALL]
        _check_rays_like(self, layout, n, device, what)

    def soa(self) -> _capi.FootprintSoA:
        return _capi.FootprintSoA(*[getattr(self, k).data_ptr() if getattr(self, k) is not None
                                    else None for k in self.ALL])

    def numpy(self) -> dict:
        """Host copies of the allocated planes (three-component planes as (n, 3))."""
        out = {}
        for k in self.ALL:
            t = getattr(self, k)
            if t is not None:
                a = t.cpu().numpy()
                out[k] = a.T.copy() if a.ndim == 2 else a
        return out


def disk_shading(density=1.0, opacity=2.0, peak_temp=6000.0, brightness=5.5, noise_scale=1.0,
                 noise_strength=0.85, warmth=0.4, doppler=0.0, time=0.0,
                 inflow=0.0) -> _capi.DiskShading:
    """bhs_disk_shading; defaults of the reference's Settings (src/Settings.h:32-45)."""
    return _capi.DiskShading(density, opacity, peak_temp, brightness, noise_scale,
                             noise_strength, warmth, doppler, time, inflow)


def sky_params(sky_brightness=1.0, star_density=1.0, star_brightness=1.0,
               doppler=0.0) -> _capi.SkyParams:
    """bhs_sky_params; defaults of the reference's Settings (src/Settings.h:53-55) and the film
    look (uDopplerStrength 0)."""
    return _capi.SkyParams(sky_brightness, star_density, star_brightness, doppler)


class SkyCube:
    """Backdrop cube map for bhs_sample_sky: `levels` is a list of (6, S_l, S_l, 3) arrays (faces
    +X -X +Y -Y +Z -Z, row j at t = (j + 0.5) / S_l, S_l = max(1, res >> l)), packed as float4 texels
    in one device tensor in the library's layout."""

    def __init__(self, levels, device):
        import numpy as np
        import torch
        res = int(np.asarray(levels[0]).shape[1])
        if not 1 <= len(levels) <= 16:
            raise ValueError("SkyCube: 1..16 levels")
        flat = []
        for l, a in enumerate(levels):
            a = np.asarray(a, np.float32)
            S = max(1, res >> l)
            if a.shape != (6, S, S, 3):
                raise ValueError(f"SkyCube: level {l} must be (6, {S}, {S}, 3), got {a.shape}")
            t = np.zeros((6, S, S, 4), np.float32)
            t[..., :3] = a
            flat.append(t.reshape(-1, 4))
        self.res, self.levels = res, len(levels)
        self.texels = torch.from_numpy(np.concatenate(flat)).to(device)

    @classmethod
    def from_faces(cls, faces, device, levels: int | None = None):
        """Cube map from level-0 faces (6, res, res, 3) with box-filtered mip levels (as
        glGenerateMipmap; res a power of two)."""
        import numpy as np
        a = np.asarray(faces, np.float32)
        res = a.shape[1]
        n = levels if levels is not None else res.bit_length()
        out = [a]
        for _ in range(1, n):
            S = a.shape[1]
            if S == 1:
                break
            a = 0.25 * (a[:, 0::2, 0::2] + a[:, 1::2, 0::2] + a[:, 0::2, 1::2] + a[:, 1::2, 1::2])
            out.append(a.astype(np.float32))
        return cls(out, device)

    def struct(self) -> _capi.SkyCube:
        return _capi.SkyCube(self.texels.data_ptr(), self.res, self.levels)

    def numpy(self):
        """(texels (m, 4) float32, res, levels) on the host, the oracle's cube form."""
        return self.texels.cpu().numpy(), self.res, self.levels


def halton_jitter(sample: int):
    """Sub-pixel jitter of accumulation sample `sample` (jitterFor, src/Renderer.cpp:16-29)."""
    o = (C.c_float * 2)()
    _capi.load().bhs_halton_jitter(sample, o)
    return (o[0], o[1])


class ShadeBuffers:
    """Device shading planes (bhs_shade_soa): rgb (3, n) and sky_weight (n) float32."""

    def __init__(self, n: int, device):
        import torch
        self.n = n
        self.rgb = torch.empty((3, n), dtype=torch.float32, device=device)
        self.sky_weight = torch.empty(n, dtype=torch.float32, device=device)

    def check(self, n: int, device: int, what: str):
        _check_rays_like(self, (("rgb", "float32", 3), ("sky_weight", "float32", 1)), n, device,
                         what)

    def soa(self) -> _capi.ShadeSoA:
        return _capi.ShadeSoA(self.rgb.data_ptr(), self.sky_weight.data_ptr())

    def numpy(self) -> dict:
        return dict(rgb=self.rgb.cpu().numpy().T.copy(), sky_weight=self.sky_weight.cpu().numpy())


class Tracer:
    """One bhs_ctx on one HIP device."""

    def __init__(self, device: int = 0):
        self._lib = _capi.load()
        st = C.c_int(0)
        self._ctx = self._lib.bhs_create(device, C.byref(st))
        if not self._ctx:
            raise BhsError(st.value, f"bhs_create(device={device}) failed")
        self.device = device

    def close(self):
        if getattr(self, "_ctx", None):
            self._lib.bhs_destroy(self._ctx)
            self._ctx = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _check(self, rc: int, what: str):
        if rc != _capi.BHS_OK:
            msg = self._lib.bhs_last_error(self._ctx)
            raise BhsError(rc, f"{what}: {msg.decode() if msg else ''}")

    @staticmethod
    def _stream_ptr(stream):
        if stream is None:
            import torch
            return torch.cuda.current_stream().cuda_stream
        return getattr(stream, "cuda_stream", stream)

    # ---- law T: bhs::physics::trace --------------------------------------------------------
    def trace_batch(self, pos3, dir3, max_steps=200000, step_scale=0.005, esc_tol=1e-4,
                    r_esc_min=30.0) -> np.ndarray:
        pos3 = np.ascontiguousarray(pos3, dtype=np.float64).reshape(-1, 3)
        dir3 = np.ascontiguousarray(dir3, dtype=np.float64).reshape(-1, 3)
        if pos3.shape != dir3.shape:
            raise ValueError("pos3 and dir3 must have the same shape")
        n = pos3.shape[0]
        out = (_capi.TraceResult * max(n, 1))()
        rc = self._lib.bhs_trace_rays(self._ctx, pos3.ctypes.data, dir3.ctypes.data, n, max_steps,
                                      step_scale, esc_tol, r_esc_min, C.addressof(out))
        self._check(rc, "bhs_trace_rays")
        dt = np.dtype([("status", "<i4"), ("steps", "<i4"), ("equator_crossings", "<i4"),
                       ("pad_", "<i4"), ("dir", "<f8", 3), ("min_r", "<f8")])
        return np.frombuffer(bytes(out), dtype=dt, count=n).copy()

    def trace(self, cam_pos, d, max_steps=200000, step_scale=0.005, esc_tol=1e-4,
              r_esc_min=30.0) -> TraceResult:
        r = self.trace_batch([cam_pos], [d], max_steps, step_scale, esc_tol, r_esc_min)[0]
        return TraceResult(int(r["status"]), np.array(r["dir"]), int(r["steps"]), float(r["min_r"]),
                           int(r["equator_crossings"]))

    # ---- law S ---------------------------------------------------------------------------------
    def render(self, cam: Camera, params: _capi.MarchParams | None = None,
               out: RayBuffers | None = None, stream=None) -> RayBuffers:
        """Law-S render of one camera into device buffers (stream-ordered, asynchronous)."""
        if params is None:
            params = march_params()
        if out is None:
            out = RayBuffers(cam.width * cam.height, f"cuda:{self.device}")
        out.check(cam.width * cam.height, self.device, "render: out")
        cs = _camera_struct(cam)
        soa = out.soa()
        rc = self._lib.bhs_render(self._ctx, C.byref(cs), C.byref(params), C.byref(soa),
                                  C.c_void_p(self._stream_ptr(stream)))
        self._check(rc, "bhs_render")
        return out

    def render_noalloc(self, cam: Camera, params: _capi.MarchParams | None = None,
                       out: RayBuffers | None = None, stream=None) -> RayBuffers:
        """bhs_render_noalloc: render() whose scratch is the library's static device array (no
        device allocation in the call; renders that need more than BHS_NOALLOC_SCRATCH_BYTES of
        scratch are rejected). `out` must be given: the call allocates nothing."""
        if params is None:
            params = march_params()
        if out is None:
            raise ValueError("render_noalloc: out must be given (the call allocates nothing)")
        out.check(cam.width * cam.height, self.device, "render_noalloc: out")
        cs = _camera_struct(cam)
        soa = out.soa()
        rc = self._lib.bhs_render_noalloc(self._ctx, C.byref(cs), C.byref(params), C.byref(soa),
                                          C.c_void_p(self._stream_ptr(stream)))
        self._check(rc, "bhs_render_noalloc")
        return out

    @staticmethod
    def _batch_structs(cams, params):
        n = len(cams)
        if isinstance(params, _capi.MarchParams):
            params = [params] * n
        if len(params) != n:
            raise ValueError("one MarchParams per camera")
        ca = (_capi.Camera * n)(*[_camera_struct(c) for c in cams])
        pa = (_capi.MarchParams * n)(*params)
        return ca, pa

    @staticmethod
    def batch_offsets(cams) -> list:
        """First output index of each camera of a batch (prefix sums of width*height)."""
        offs, o = [], 0
        for c in cams:
            offs.append(o)
            o += c.width * c.height
        return offs + [o]

    def render_batch(self, cams, params, out: RayBuffers | None = None, stream=None) -> RayBuffers:
        """Law-S render of up to 64 cameras in one persistent launch (bhs_render_batch); camera k's
        rays start at batch_offsets(cams)[k]."""
        ca, pa = self._batch_structs(cams, params)
        if out is None:
            out = RayBuffers(self.batch_offsets(cams)[-1], f"cuda:{self.device}")
        out.check(self.batch_offsets(cams)[-1], self.device, "render_batch: out")
        soa = out.soa()
        rc = self._lib.bhs_render_batch(self._ctx, ca, pa, len(cams), C.byref(soa),
                                        C.c_void_p(self._stream_ptr(stream)))
        self._check(rc, "bhs_render_batch")
        return out

    def sky_footprint(self, cam: Camera, rays: RayBuffers, out: FootprintBuffers | None = None,
                      coarse: bool = False, stream=None) -> FootprintBuffers:
        """Sky footprint post-pass over the planes of a render of `cam` (bhs_sky_footprint):
        derivatives of escDir over 2x2 quads (shaders/blackhole.frag:297-306) and makeFootprint
        (shaders/sky.glsl:27-51)."""
        n = cam.width * cam.height
        rays.check(n, self.device, "sky_footprint: rays")
        if out is None:
            out = FootprintBuffers(n, f"cuda:{self.device}")
        out.check(n, self.device, "sky_footprint: out")
        cs = _camera_struct(cam)
        rs, fs = rays.soa(), out.soa()
        rc = self._lib.bhs_sky_footprint(self._ctx, C.byref(cs), C.byref(rs), int(bool(coarse)),
                                         C.byref(fs), C.c_void_p(self._stream_ptr(stream)))
        self._check(rc, "bhs_sky_footprint")
        return out

    def sky_footprint_batch(self, cams, rays: RayBuffers, out: FootprintBuffers | None = None,
                            coarse: bool = False, stream=None) -> FootprintBuffers:
        """The footprint pass over the planes of a batch render (bhs_sky_footprint_batch)."""
        n = len(cams)
        total = self.batch_offsets(cams)[-1]
        rays.check(total, self.device, "sky_footprint_batch: rays")
        if out is None:
            out = FootprintBuffers(total, f"cuda:{self.device}")
        out.check(total, self.device, "sky_footprint_batch: out")
        ca = (_capi.Camera * n)(*[_camera_struct(c) for c in cams])
        rs, fs = rays.soa(), out.soa()
        rc = self._lib.bhs_sky_footprint_batch(self._ctx, ca, n, C.byref(rs), int(bool(coarse)),
                                               C.byref(fs), C.c_void_p(self._stream_ptr(stream)))
        self._check(rc, "bhs_sky_footprint_batch")
        return out

    def gas_lens_table(self, cam_r: float, alpha_crit: float | None = None, n_rays: int = 1024,
                       n_phis: int = 1024, out=None, stream=None):
        """Gas lens table (bhs_gas_lens_table; shaders/gas_lens.comp:11-44): an (n_phis, n_rays)
        float32 tensor of log r per swept angle and fan ray. alpha_crit defaults to the shadow's
        angular radius at cam_r in float (src/GasParticles.cpp:40-44), as GasParticles passes it."""
        import torch
        if alpha_crit is None:
            alpha_crit = shadow_angular_radius_f32(cam_r)
        if out is None:
            out = torch.empty((n_phis, n_rays), dtype=torch.float32, device=f"cuda:{self.device}")
        _check_tensor(out, "float32", (n_phis, n_rays), self.device, "gas_lens_table: out")
        rc = self._lib.bhs_gas_lens_table(self._ctx, cam_r, alpha_crit, n_rays, n_phis,
                                          out.data_ptr(), C.c_void_p(self._stream_ptr(stream)))
        self._check(rc, "bhs_gas_lens_table")
        return out

    def gas_lens_tables(self, cam_rs, alpha_crits=None, n_rays: int = 1024, n_phis: int = 1024,
                        out=None, stream=None):
        """Several gas lens tables in one launch (bhs_gas_lens_tables): a (len(cam_rs), n_phis,
        n_rays) float32 tensor, table k equal to gas_lens_table(cam_rs[k], alpha_crits[k])."""
        import torch
        k = len(cam_rs)
        if not 1 <= k <= 64:
            raise ValueError("gas_lens_tables: 1..64 camera radii")
        if alpha_crits is None:
            alpha_crits = [shadow_angular_radius_f32(r) for r in cam_rs]
        if len(alpha_crits) != k:
            raise ValueError("gas_lens_tables: one alpha_crit per camera radius")
        if out is None:
            out = torch.empty((k, n_phis, n_rays), dtype=torch.float32,
                              device=f"cuda:{self.device}")
        _check_tensor(out, "float32", (k, n_phis, n_rays), self.device, "gas_lens_tables: out")
        ra = (C.c_float * k)(*[float(r) for r in cam_rs])
        aa = (C.c_float * k)(*[float(a) for a in alpha_crits])
        rc = self._lib.bhs_gas_lens_tables(self._ctx, k, ra, aa, n_rays, n_phis, out.data_ptr(),
                                           C.c_void_p(self._stream_ptr(stream)))
        self._check(rc, "bhs_gas_lens_tables")
        return out

    # ---- sky ----------------------------------------------------------------------------------
    def _sky_call(self, fn, first_args, n, rays, params, cube, weight, rgb, coarse, stream, what):
        import torch
        rays.check(n, self.device, f"{what}: rays")
        if rays.flags is None or rays.esc is None:
            raise ValueError(f"{what}: the rays need their flags and escape direction planes")
        if rgb is None:
            rgb = torch.zeros((3, n), dtype=torch.float32, device=f"cuda:{self.device}")
        _check_tensor(rgb, "float32", (3, n), self.device, f"{what}: rgb")
        if weight is not None:
            _check_tensor(weight, "float32", (n,), self.device, f"{what}: weight")
        if cube is not None and cube.texels.device != rgb.device:
            raise ValueError(f"{what}: the cube map lives on {cube.texels.device}")
        params = sky_params() if params is None else params
        cs = cube.struct() if cube is not None else None
        rs = rays.soa()
        rc = fn(self._ctx, *first_args, C.byref(rs), int(bool(coarse)), C.byref(params),
                C.byref(cs) if cs is not None else None,
                C.c_void_p(weight.data_ptr()) if weight is not None else None,
                C.c_void_p(rgb.data_ptr()), C.c_void_p(self._stream_ptr(stream)))
        self._check(rc, what)
        return rgb

    def sample_sky(self, cam: Camera, rays: RayBuffers, params: _capi.SkyParams | None = None,
                   cube: SkyCube | None = None, weight=None, rgb=None, coarse: bool = False,
                   stream=None):
        """The sky of the escaped rays of a render of `cam` (bhs_sample_sky; sky.glsl:57-119):
        rgb (3, n) += weight * sampleSky; weight (n,) is the shading's sky weight (None: 1)."""
        return self._sky_call(self._lib.bhs_sample_sky, (C.byref(_camera_struct(cam)),),
                              cam.width * cam.height, rays, params, cube, weight, rgb, coarse,
                              stream, "bhs_sample_sky")

    def sample_sky_batch(self, cams, rays: RayBuffers, params: _capi.SkyParams | None = None,
                         cube: SkyCube | None = None, weight=None, rgb=None, coarse: bool = False,
                         stream=None):
        """The sky pass over the planes of a batch render (bhs_sample_sky_batch)."""
        ca = (_capi.Camera * len(cams))(*[_camera_struct(c) for c in cams])
        return self._sky_call(self._lib.bhs_sample_sky_batch, (ca, len(cams)),
                              self.batch_offsets(cams)[-1], rays, params, cube, weight, rgb,
                              coarse, stream, "bhs_sample_sky_batch")

    # ---- thin-disk shading, accumulation --------------------------------------------------------
    def shade_thin_disk(self, cam: Camera, rays: RayBuffers, params: _capi.MarchParams,
                        shading: _capi.DiskShading | None = None, out: ShadeBuffers | None = None,
                        stream=None) -> ShadeBuffers:
        """Thin-disk colour and transparency (bhs_shade_thin_disk) of a render of `cam` with
        march parameters `params`: blackhole.frag:209-216,234 with thinDiskSample."""
        n = cam.width * cam.height
        rays.check(n, self.device, "shade_thin_disk: rays")
        if shading is None:
            shading = disk_shading()
        if out is None:
            out = ShadeBuffers(n, f"cuda:{self.device}")
        out.check(n, self.device, "shade_thin_disk: out")
        cs = _camera_struct(cam)
        rs, os_ = rays.soa(), out.soa()
        rc = self._lib.bhs_shade_thin_disk(self._ctx, C.byref(cs), C.byref(params),
                                           C.byref(shading), C.byref(rs), C.byref(os_),
                                           C.c_void_p(self._stream_ptr(stream)))
        self._check(rc, "bhs_shade_thin_disk")
        return out

    def shade_thin_disk_batch(self, cams, rays: RayBuffers, params,
                              shading: _capi.DiskShading | None = None,
                              out: ShadeBuffers | None = None, stream=None) -> ShadeBuffers:
        """The shading pass over a batch render (bhs_shade_thin_disk_batch)."""
        ca, pa = self._batch_structs(cams, params)
        total = self.batch_offsets(cams)[-1]
        rays.check(total, self.device, "shade_thin_disk_batch: rays")
        if shading is None:
            shading = disk_shading()
        if out is None:
            out = ShadeBuffers(total, f"cuda:{self.device}")
        out.check(total, self.device, "shade_thin_disk_batch: out")
        rs, os_ = rays.soa(), out.soa()
        rc = self._lib.bhs_shade_thin_disk_batch(self._ctx, ca, pa, len(cams), C.byref(shading),
                                                 C.byref(rs), C.byref(os_),
                                                 C.c_void_p(self._stream_ptr(stream)))
        self._check(rc, "bhs_shade_thin_disk_batch")
        return out

    def render_volumetric(self, cam: Camera, params: _capi.MarchParams | None = None,
                          shading: _capi.DiskShading | None = None,
                          rays: RayBuffers | None = None, out: ShadeBuffers | None = None,
                          stream=None):
        """Volumetric disk render (bhs_render_volumetric; blackhole.frag DISK_VOLUMETRIC,
        disk.glsl:81-110). params default to the High preset with disk_mode DISK_VOLUMETRIC.
        Returns (ShadeBuffers, RayBuffers or None)."""
        if params is None:
            params = march_params(disk_mode=DISK_VOLUMETRIC)
        if shading is None:
            shading = disk_shading()
        n = cam.width * cam.height
        if out is None:
            out = ShadeBuffers(n, f"cuda:{self.device}")
        out.check(n, self.device, "render_volumetric: out")
        if rays is not None:
            rays.check(n, self.device, "render_volumetric: rays")
        cs = _camera_struct(cam)
        os_ = out.soa()
        rs = C.byref(rays.soa()) if rays is not None else None
        rc = self._lib.bhs_render_volumetric(self._ctx, C.byref(cs), C.byref(params),
                                             C.byref(shading), rs, C.byref(os_),
                                             C.c_void_p(self._stream_ptr(stream)))
        self._check(rc, "bhs_render_volumetric")
        return out, rays

    def render_volumetric_batch(self, cams, params, shading: _capi.DiskShading | None = None,
                                rays: RayBuffers | None = None, out: ShadeBuffers | None = None,
                                stream=None):
        """The volumetric render of up to 64 cameras in one call (bhs_render_volumetric_batch)."""
        ca, pa = self._batch_structs(cams, params)
        if shading is None:
            shading = disk_shading()
        total = self.batch_offsets(cams)[-1]
        if out is None:
            out = ShadeBuffers(total, f"cuda:{self.device}")
        out.check(total, self.device, "render_volumetric_batch: out")
        if rays is not None:
            rays.check(total, self.device, "render_volumetric_batch: rays")
        os_ = out.soa()
        rs = C.byref(rays.soa()) if rays is not None else None
        rc = self._lib.bhs_render_volumetric_batch(self._ctx, ca, pa, len(cams), C.byref(shading),
                                                   rs, C.byref(os_),
                                                   C.c_void_p(self._stream_ptr(stream)))
        self._check(rc, "bhs_render_volumetric_batch")
        return out, rays

    def accumulate(self, cur_rgb, acc_rgb, sample_index: int, cap: int = 256, stream=None):
        """acc_rgb <- running average with cur_rgb (bhs_accumulate; (3, n) float32 tensors)."""
        n = cur_rgb.shape[-1]
        _check_tensor(cur_rgb, "float32", (3, n), self.device, "accumulate: cur_rgb")
        _check_tensor(acc_rgb, "float32", (3, n), self.device, "accumulate: acc_rgb")
        rc = self._lib.bhs_accumulate(self._ctx, cur_rgb.data_ptr(), acc_rgb.data_ptr(), n,
                                      sample_index, cap, C.c_void_p(self._stream_ptr(stream)))
        self._check(rc, "bhs_accumulate")
        return acc_rgb

    def disk_textures(self):
        """Host copies of the context's noise volume ((128,128,128) uint16 half texels, z y x)
        and blackbody table ((256, 4) float32)."""
        vol = np.zeros((128, 128, 128), np.uint16)
        lut = np.zeros((256, 4), np.float32)
        self._check(self._lib.bhs_disk_textures(self._ctx, vol.ctypes.data, lut.ctypes.data),
                    "bhs_disk_textures")
        return vol, lut

    def graph(self, cams, params, out: RayBuffers | None = None) -> "BatchGraph":
        """Captures the batch launch into a HIP graph (bhs_graph_create) for replay."""
        return BatchGraph(self, cams, params, out)

    def march_rays(self, pos3, dir3, params: _capi.MarchParams | None = None,
                   out: RayBuffers | None = None, stream=None) -> RayBuffers:
        """Law S on explicit rays; pos3/dir3 are (n,3) float32 CUDA tensors."""
        import torch
        if params is None:
            params = march_params()
        pos3 = pos3.contiguous()
        dir3 = dir3.contiguous()
        if pos3.dim() != 2 or pos3.shape[1] != 3:
            raise ValueError("march_rays: pos3 must have shape (n, 3)")
        n = pos3.shape[0]
        _check_tensor(pos3, "float32", (n, 3), self.device, "march_rays: pos3")
        _check_tensor(dir3, "float32", (n, 3), self.device, "march_rays: dir3")
        if out is None:
            out = RayBuffers(n, pos3.device)
        out.check(n, self.device, "march_rays: out")
        soa = out.soa()
        rc = self._lib.bhs_march_rays(self._ctx, pos3.data_ptr(), dir3.data_ptr(), n,
                                      C.byref(params), C.byref(soa),
                                      C.c_void_p(self._stream_ptr(stream)))
        self._check(rc, "bhs_march_rays")
        return out

    def last_stats(self) -> dict:
        s = _capi.Stats()
        self._check(self._lib.bhs_last_stats(self._ctx, C.byref(s)), "bhs_last_stats")
        return dict(rays=s.rays, steps=s.steps, crossings=s.crossings, escapes=s.escapes,
                    band64=s.band64, fallback=s.fallback, surrogate=s.surrogate)

    def rsqrt_rn(self, x, out=None, stream=None):
        """The float march's reciprocal square root (bhs_rsqrt_rn) of a float32 device tensor:
        1/sqrt(x) rounded once to float."""
        n = x.numel()
        _check_tensor(x, "float32", (n,), self.device, "rsqrt_rn: x")
        if out is None:
            import torch
            out = torch.empty_like(x)
        _check_tensor(out, "float32", (n,), self.device, "rsqrt_rn: out")
        self._check(self._lib.bhs_rsqrt_rn(self._ctx, C.c_void_p(x.data_ptr()),
                                           C.c_void_p(out.data_ptr()), n,
                                           C.c_void_p(self._stream_ptr(stream))), "bhs_rsqrt_rn")
        return out

    def valu_probe(self, kind: int, waves_per_simd: int = 8, iters: int = 4096) -> dict:
        """VALU issue-rate probe (bhs_valu_probe): kernel ms and wave64 instructions issued."""
        ms, wi = C.c_double(), C.c_double()
        self._check(self._lib.bhs_valu_probe(self._ctx, kind, waves_per_simd, iters, C.byref(ms),
                                             C.byref(wi)), "bhs_valu_probe")
        return {"ms": ms.value, "wave_instr": wi.value}

    def set_tuning(self, mode: int, block: int = 256, waves_per_simd: int = 8,
                   refill_min: int = 32):
        t = _capi.Tuning(mode, block, waves_per_simd, refill_min)
        self._check(self._lib.bhs_set_tuning(self._ctx, C.byref(t)), "bhs_set_tuning")

    def get_tuning(self) -> dict:
        t = _capi.Tuning()
        self._check(self._lib.bhs_get_tuning(self._ctx, C.byref(t)), "bhs_get_tuning")
        return dict(mode=t.mode, block=t.block, waves_per_simd=t.waves_per_simd,
                    refill_min=t.refill_min)

    def autotune(self, cam: Camera, params: _capi.MarchParams | None = None,
                 out: RayBuffers | None = None, reps: int = 5) -> dict:
        """Runs the library's autotuner (bhs_autotune) on this camera; keeps and returns the best
        launch shape with its median main-kernel ms."""
        if params is None:
            params = march_params()
        if out is None:
            out = RayBuffers(cam.width * cam.height, f"cuda:{self.device}")
        cs = _camera_struct(cam)
        soa = out.soa()
        best = _capi.Tuning()
        ms = C.c_double()
        self._check(self._lib.bhs_autotune(self._ctx, C.byref(cs), C.byref(params), C.byref(soa),
                                           reps, C.byref(best), C.byref(ms)), "bhs_autotune")
        return dict(mode=best.mode, block=best.block, waves_per_simd=best.waves_per_simd,
                    refill_min=best.refill_min, ms=ms.value)

    def set_timing(self, on: bool):
        self._check(self._lib.bhs_set_timing(self._ctx, 1 if on else 0), "bhs_set_timing")

    def timing(self) -> dict:
        """Average GPU ms of the main march kernel and of the whole call since the last call."""
        km, tm, n = C.c_double(), C.c_double(), C.c_int32()
        self._check(self._lib.bhs_timing(self._ctx, C.byref(km), C.byref(tm), C.byref(n)),
                    "bhs_timing")
        return dict(march_ms=km.value, total_ms=tm.value, launches=n.value)

    def sync(self):
        self._check(self._lib.bhs_sync(self._ctx), "bhs_sync")


class BatchGraph:
    """A captured batch launch (bhs_graph_*): replay with launch(), new poses with update()."""

    def __init__(self, tracer: Tracer, cams, params, out: RayBuffers | None = None):
        self.tracer = tracer
        self.cams = list(cams)
        ca, pa = tracer._batch_structs(self.cams, params)
        self.out = out if out is not None else RayBuffers(tracer.batch_offsets(self.cams)[-1],
                                                          f"cuda:{tracer.device}")
        self.out.check(tracer.batch_offsets(self.cams)[-1], tracer.device, "graph: out")
        self._soa = self.out.soa()
        h = C.c_void_p()
        rc = tracer._lib.bhs_graph_create(tracer._ctx, ca, pa, len(self.cams), C.byref(self._soa),
                                          C.byref(h))
        tracer._check(rc, "bhs_graph_create")
        self._g = h

    def launch(self, stream=None):
        rc = self.tracer._lib.bhs_graph_launch(self.tracer._ctx, self._g,
                                               C.c_void_p(self.tracer._stream_ptr(stream)))
        self.tracer._check(rc, "bhs_graph_launch")
        return self.out

    def update(self, cams, params):
        ca, pa = self.tracer._batch_structs(list(cams), params)
        self.tracer._check(self.tracer._lib.bhs_graph_update(self.tracer._ctx, self._g, ca, pa),
                           "bhs_graph_update")

    def stats(self) -> dict:
        s = _capi.Stats()
        self.tracer._check(self.tracer._lib.bhs_graph_stats(self.tracer._ctx, self._g, C.byref(s)),
                           "bhs_graph_stats")
        return dict(rays=s.rays, steps=s.steps, crossings=s.crossings, escapes=s.escapes,
                    band64=s.band64, fallback=s.fallback, surrogate=s.surrogate)

    def close(self):
        if getattr(self, "_g", None):
            self.tracer._lib.bhs_graph_destroy(self._g)
            self._g = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


_tls = threading.local()


def default_tracer() -> Tracer:
    t = getattr(_tls, "tracer", None)
    if t is None:
        t = Tracer(0)
        _tls.tracer = t
    return t


def trace(cam_pos, d, max_steps=200000, step_scale=0.005, esc_tol=1e-4,
          r_esc_min=30.0) -> TraceResult:
    """bhs::physics::trace (src/physics/Geodesic.h:107-140) on the GPU, same defaults."""
    return default_tracer().trace(cam_pos, d, max_steps, step_scale, esc_tol, r_esc_min)
