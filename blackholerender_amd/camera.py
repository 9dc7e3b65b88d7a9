"""Pinhole cameras for the law-S render entry points.

Mirrors the reference's ray-generation inputs: Camera::lookRotation (src/Camera.cpp:12-23) builds
the basis whose columns are right, up, back (the camera looks down -back); Renderer::setTraceUniforms
(src/Renderer.cpp:246-252) passes tan(fov/2) of the vertical field of view and the resolution.
The basis is computed in double and rounded to float once (the reference's quaternion round trip,
src/Camera.cpp:22 and :67, only adds float noise and is not reproduced).
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field

import numpy as np


def look_basis(pos, target, up=(0.0, 1.0, 0.0)) -> np.ndarray:
    """Column-major 3x3 basis (right, up, back) of src/Camera.cpp:12-23, as 9 float32."""
    pos = np.asarray(pos, dtype=np.float64)
    fd = np.asarray(target, dtype=np.float64) - pos
    ln = float(np.linalg.norm(fd))
    f = fd / ln if ln > 0.0 else np.array([0.0, 0.0, -1.0])
    u = np.asarray(up, dtype=np.float64)
    if abs(float(np.dot(f, u / np.linalg.norm(u)))) > 0.999:   # degenerate up: Camera.cpp:17
        u = np.array([0.0, 0.0, -1.0])
    r = np.cross(f, u)
    r = r / np.linalg.norm(r)
    uu = np.cross(r, f)
    return np.concatenate([r, uu, -f]).astype(np.float32)


@dataclass
class Camera:
    """Fields of the C-ABI struct bhs_camera (include/bhs_trace.h)."""
    pos: np.ndarray
    basis: np.ndarray
    tan_half_fov: float
    width: int
    height: int
    jitter: tuple = (0.0, 0.0)
    jitter_hashed: int = 0
    camera_id: int = 0
    sample: int = 0
    seed: int = 0
    meta: dict = field(default_factory=dict)

    @property
    def n_rays(self) -> int:
        return self.width * self.height

    def as_dict(self) -> dict:
        return dict(pos=np.asarray(self.pos, np.float32), basis=np.asarray(self.basis, np.float32),
                    tan_half_fov=float(np.float32(self.tan_half_fov)), width=self.width,
                    height=self.height, jitter=tuple(self.jitter), jitter_hashed=self.jitter_hashed,
                    camera_id=self.camera_id, sample=self.sample, seed=self.seed)


def orbit_camera(r_obs: float, inclination_deg: float, fov_deg: float, width: int, height: int,
                 *, jitter_hashed: int = 0, seed: int = 0, camera_id: int = 0, sample: int = 0,
                 target=(0.0, 0.0, 0.0), up=(0.0, 1.0, 0.0), azimuth_deg: float = 0.0) -> Camera:
    """Camera at r_obs * (sin i sin a, cos i, sin i cos a) looking at the hole, +y up
    (src/Camera.cpp:54-55; azimuth a = 0 is the reference's bookmark plane)."""
    inc = math.radians(inclination_deg)
    az = math.radians(azimuth_deg)
    pos = np.array([r_obs * math.sin(inc) * math.sin(az), r_obs * math.cos(inc),
                    r_obs * math.sin(inc) * math.cos(az)], dtype=np.float64)
    basis = look_basis(pos, target, up)
    tan_h = math.tan(math.radians(fov_deg) * 0.5)
    return Camera(pos=pos.astype(np.float32), basis=basis, tan_half_fov=float(np.float32(tan_h)),
                  width=width, height=height, jitter_hashed=jitter_hashed, seed=seed,
                  camera_id=camera_id, sample=sample,
                  meta=dict(r_obs=r_obs, inclination_deg=inclination_deg, fov_deg=fov_deg))


# Quality presets (src/Settings.cpp:15-30): max steps -> (step_scale, disk_step).
PRESETS = {200: (0.10, 0.35), 300: (0.07, 0.30), 450: (0.05, 0.18), 700: (0.035, 0.10)}
