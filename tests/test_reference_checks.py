"""The physics checks of the reference's tests/geodesic_tests.cpp (8 functions, 15 CHECKs),
restated against two backends:

  oracle : the CPU oracle's law T (oracle/bhs_oracle.c) — pins the oracle;
  gpu    : the product — trace() on the GPU through the C-ABI, helpers from
           blackholerender_amd.physics — the drop-in runs the reference's own tests.

Each test cites the reference lines it restates; thresholds are the reference's.
"""
import math

import numpy as np
import pytest

from helpers import angle_between, dir_for_impact_parameter

BACKENDS = ["oracle", pytest.param("gpu", marks=pytest.mark.gpu)]


class Oracle:
    def __init__(self):
        import oracle as O
        self.O = O

    def trace(self, cam, d, *a):
        r = self.O.trace(cam, d, *a)
        return dict(status=r.status, dir=np.array(r.dir[:]), steps=r.steps, minR=r.min_r,
                    equatorCrossings=r.equator_crossings)

    def __getattr__(self, k):
        return getattr(self.O, k)


class Gpu:
    def __init__(self, tracer):
        from blackholerender_amd import physics
        self.P = physics
        self.t = tracer

    def trace(self, cam, d, *a):
        r = self.t.trace(cam, d, *a)
        return dict(status=r.status, dir=r.dir, steps=r.steps, minR=r.minR,
                    equatorCrossings=r.equatorCrossings)

    def __getattr__(self, k):
        return getattr(self.P, k)


@pytest.fixture(params=BACKENDS)
def be(request):
    if request.param == "oracle":
        return Oracle()
    return Gpu(request.getfixturevalue("tracer"))


ESCAPED, CAPTURED = 0, 1


def test_impact_parameter_round_trip(be):  # geodesic_tests.cpp:36-48
    n = 0
    for D in (10.0, 25.0, 2000.0):
        for b in (1.0, 5.196, 8.0):
            if b * math.sqrt(1.0 - 2.0 / D) / D >= 1.0:
                continue
            bb = be.impact_parameter((0, 0, D), dir_for_impact_parameter(D, b))
            assert abs(bb - b) < 1e-9 * max(1.0, b), (D, b, bb)
            n += 1
    assert n == 9


@pytest.mark.parametrize("b", [50.0, 100.0, 200.0])
def test_weak_field_deflection(be, b):  # geodesic_tests.cpp:50-64
    D = 2000.0
    d = dir_for_impact_parameter(D, b)
    res = be.trace((0, 0, D), d, 400000, 0.002, 1e-6, 50.0)
    expected = 4.0 / b + 15.0 * math.pi / (4.0 * b * b)
    got = angle_between(d, res["dir"])
    assert res["status"] == ESCAPED
    assert abs(got - expected) / expected < 0.01


@pytest.mark.parametrize("D", [10.0, 25.0, 2000.0])
def test_capture_threshold(be, D):  # geodesic_tests.cpp:66-78
    cam = (0, 0, D)
    inside = be.trace(cam, dir_for_impact_parameter(D, 5.15), 400000, 0.002, 1e-4, 30.0)
    outside = be.trace(cam, dir_for_impact_parameter(D, 5.25), 400000, 0.002, 1e-4, 30.0)
    assert inside["status"] == CAPTURED
    assert outside["status"] == ESCAPED


@pytest.mark.parametrize("r", [4.0, 6.0, 25.0, 100.0])
def test_shadow_angular_size(be, r):  # geodesic_tests.cpp:80-94
    a = be.shadow_angular_radius(r)
    cam = (0, 0, r)
    da = 0.004
    d_in = (math.sin(a - da), 0.0, -math.cos(a - da))
    d_out = (math.sin(a + da), 0.0, -math.cos(a + da))
    assert be.trace(cam, d_in, 400000, 0.002, 1e-4, 30.0)["status"] == CAPTURED
    assert be.trace(cam, d_out, 400000, 0.002, 1e-4, 30.0)["status"] == ESCAPED


def test_photon_sphere(be):  # geodesic_tests.cpp:96-112 (host FP64 rk4Step)
    x = np.array([3.0, 0.0, 0.0])
    v = np.asarray(be.initial_velocity(x, (0.0, 0.0, 1.0)), float)
    hv = np.cross(x, v)
    h2 = float(np.dot(hv, hv))
    max_dev = 0.0
    ds = 0.01
    path_len = 2.0 * 2.0 * math.pi * 3.0
    s = 0.0
    while s < path_len:
        x, v = be.rk4_step(x, v, h2, ds / math.sqrt(float(np.dot(v, v))))
        max_dev = max(max_dev, abs(math.sqrt(float(np.dot(x, x))) - 3.0))
        s += ds
    assert max_dev < 1e-4


def test_redshift(be):  # geodesic_tests.cpp:114-143
    cam = np.array([0.0, 0.0, 30.0])
    axis = (0.0, 1.0, 0.0)
    d_away = (np.array([8.0, 0, 0]) - cam) / np.linalg.norm(np.array([8.0, 0, 0]) - cam)
    d_toward = (np.array([-8.0, 0, 0]) - cam) / np.linalg.norm(np.array([-8.0, 0, 0]) - cam)
    b_away = be.impact_parameter_about_axis(cam, d_away, axis)
    b_toward = be.impact_parameter_about_axis(cam, d_toward, axis)
    assert b_away < 0.0 and b_toward > 0.0
    assert abs(abs(b_away) - 8.0) < 0.15
    g_away = be.redshift_factor(8.0, b_away, 30.0)
    g_toward = be.redshift_factor(8.0, b_toward, 30.0)
    assert g_away < 1.0 < g_toward
    expect = 1.0 / (math.sqrt(1 - 2.0 / 30) / math.sqrt(1 - 3.0 / 8) * (1 - 8.0 ** -1.5 * b_away))
    assert abs(g_away - expect) < 1e-9
    cam_top = np.array([0.0, 30.0, 0.0])
    dt = np.array([8.0, 0, 0]) - cam_top
    b_top = be.impact_parameter_about_axis(cam_top, dt / np.linalg.norm(dt), axis)
    g_top = be.redshift_factor(8.0, b_top, 30.0)
    expect_top = math.sqrt(1 - 3.0 / 8) / math.sqrt(1 - 2.0 / 30)
    assert abs(b_top) < 1e-9 and abs(g_top - expect_top) < 1e-9


@pytest.mark.parametrize("b", [8.0, 20.0, 60.0])
def test_far_field_correction(be, b):  # geodesic_tests.cpp:145-159
    D = 2000.0
    d = dir_for_impact_parameter(D, b)
    early = be.trace((0, 0, D), d, 400000, 0.002, 1e-2, 12.0)
    late = be.trace((0, 0, D), d, 400000, 0.002, 1e-7, 12.0)
    assert early["status"] == ESCAPED and late["status"] == ESCAPED
    assert angle_between(early["dir"], late["dir"]) < 5e-4


def test_disk_images_above_and_below(be):  # geodesic_tests.cpp:161-179
    cam = np.array([0.0, 2.6, 24.9])
    rc = float(np.linalg.norm(cam))
    to_hole = -cam / rc
    up = np.array([0.0, 1.0, 0.0]) - to_hole * np.dot([0.0, 1.0, 0.0], to_hole)
    up /= np.linalg.norm(up)
    a = be.shadow_angular_radius(rc) + 0.01
    for sgn in (1.0, -1.0):
        d = to_hole * math.cos(a) + sgn * up * math.sin(a)
        d /= np.linalg.norm(d)
        res = be.trace(cam, d, 400000, 0.002, 1e-4, 30.0)
        assert res["status"] == ESCAPED and res["equatorCrossings"] >= 1
