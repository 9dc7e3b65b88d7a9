"""Host FP64 helpers of bhs::physics (src/physics/Geodesic.h:15-95), for callers that need the
closed forms and the single RK4 step on the CPU, as the reference's header-only functions are.

These are not the hot path: trace() and every batched march run on the GPU (see Tracer). The
operation order follows GLM's (dot = (ax bx + ay by) + az bz, normalize = v * (1/sqrt(dot)), vec /
scalar divides each component) so results match the reference's CPU functions bit for bit.
"""
from __future__ import annotations

import math

import numpy as np

kHorizon = 2.0
kPhotonSphere = 3.0
kIsco = 6.0


def _dot(a, b) -> float:
    return (a[0] * b[0] + a[1] * b[1]) + a[2] * b[2]


def _cross(x, y):
    return np.array([x[1] * y[2] - y[1] * x[2], x[2] * y[0] - y[2] * x[0], x[0] * y[1] - y[0] * x[1]])


def _length(a) -> float:
    return math.sqrt(_dot(a, a))


def _normalize(v):
    return np.asarray(v, float) * (1.0 / math.sqrt(_dot(v, v)))


def critical_impact_parameter() -> float:  # Geodesic.h:19
    return 3.0 * math.sqrt(3.0)


def accel(x, h2):  # Geodesic.h:21-25
    x = np.asarray(x, float)
    r2 = _dot(x, x)
    r = math.sqrt(r2)
    return x * (-3.0 * h2 / (r2 * r2 * r))


def rk4_step(x, v, h2, dt):  # Geodesic.h:27-34 (returns the new x, v)
    x = np.asarray(x, float)
    v = np.asarray(v, float)
    hd = 0.5 * dt
    k1v, k1x = accel(x, h2), v
    k2v, k2x = accel(x + hd * k1x, h2), v + hd * k1v
    k3v, k3x = accel(x + hd * k2x, h2), v + hd * k2v
    k4v, k4x = accel(x + dt * k3x, h2), v + dt * k3v
    d6 = dt / 6.0
    return (x + d6 * (((k1x + 2.0 * k2x) + 2.0 * k3x) + k4x),
            v + d6 * (((k1v + 2.0 * k2v) + 2.0 * k3v) + k4v))


def initial_velocity(cam_pos, d):  # Geodesic.h:40-47
    cam_pos = np.asarray(cam_pos, float)
    d = np.asarray(d, float)
    r = _length(cam_pos)
    rhat = cam_pos / r
    nr = _dot(d, rhat)
    nperp = d - nr * rhat
    f = math.sqrt(max(1.0 - 2.0 / r, 0.0))
    return _normalize((nr * f) * rhat + nperp)


def impact_parameter(cam_pos, d) -> float:  # Geodesic.h:50-54
    cam_pos = np.asarray(cam_pos, float)
    r = _length(cam_pos)
    sin_a = _length(_cross(cam_pos / r, np.asarray(d, float)))
    return r * sin_a / math.sqrt(max(1.0 - 2.0 / r, 1e-12))


def impact_parameter_about_axis(cam_pos, d, axis) -> float:  # Geodesic.h:58-61
    r = _length(cam_pos)
    return -_dot(_cross(np.asarray(cam_pos, float), np.asarray(d, float)), axis) / math.sqrt(
        max(1.0 - 2.0 / r, 1e-12))


def far_field_direction(x, v, h2):  # Geodesic.h:65-75
    x = np.asarray(x, float)
    v = np.asarray(v, float)
    vlen = _length(v)
    vh = v / vlen
    r = _length(x)
    p = _dot(x, vh)
    bvec = x - p * vh
    s = min(max(p / r, 0.0), 1.0)
    r3 = r * r * r
    dv = -(h2 / vlen) * (vh / r3 + bvec * ((2.0 + s) / (r3 * r * (1.0 + s) * (1.0 + s))))
    return _normalize(v + dv)


def shadow_angular_radius(r) -> float:  # Geodesic.h:78-82
    s = min(1.0, critical_impact_parameter() / r * math.sqrt(max(1.0 - 2.0 / r, 0.0)))
    a = math.asin(s)
    return a if r >= kPhotonSphere else math.pi - a


def kepler_omega(r) -> float:  # Geodesic.h:85
    return r ** -1.5


def redshift_factor(r_emit, b_axis, r_cam) -> float:  # Geodesic.h:90-95
    omega = kepler_omega(r_emit)
    ut = 1.0 / math.sqrt(max(1.0 - 3.0 / r_emit, 1e-6))
    lapse_cam = math.sqrt(max(1.0 - 2.0 / r_cam, 1e-6))
    return 1.0 / (lapse_cam * ut * (1.0 - omega * b_axis))
