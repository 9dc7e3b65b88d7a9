"""Pins of the CPU oracle's law S (shaders/blackhole.frag:67-253) beyond the reference's own
checks, each chosen so that a plausible slip in the restatement (a dropped term, a wrong sign or
index, a transposed operand) fails it:

- escape directions against the exact Schwarzschild deflection (quadrature, tests/
  schwarzschild_exact.py) — pins ray setup, RK4, step control, escape test and far-field correction;
- crossing radii against the exact r(phi) at the equatorial-node angle — pins the chord
  interpolation and, much tighter, the FP64 root refinement;
- RK4 local order (step halving) and agreement with law T's rk4Step;
- orbit-plane invariance of the escape direction;
- disk redshift against the reference's closed form (Geodesic.h:90-95) where no clamp is active,
  and the face-on symmetry g(r) of an inclination-0 camera;
- float against double law S on the 256x256 reference camera (class/steps/escape agreement);
- class convergence on a tiny image under 10x finer steps;
- ray generation against a float64 restatement of blackhole.frag:68-72;
- the IEEE-only atan2 against numpy, the pcg hash against an independent restatement.
"""
import math

import numpy as np
import pytest

import oracle as O
from helpers import escape_angle_error
from schwarzschild_exact import (B_CRIT, radius_at_swept_angle, swept_to_infinity)


def _cam_dir_for_b(r_obs, b):
    a = math.asin(b * math.sqrt(1.0 - 2.0 / r_obs) / r_obs)
    return (np.array([0.0, 0.0, r_obs], np.float32),
            np.array([math.sin(a), 0.0, -math.cos(a)], np.float32))


@pytest.mark.parametrize("r_obs,b", [(30.0, 5.3), (30.0, 6.0), (30.0, 8.0), (30.0, 14.0),
                                     (30.0, 22.0), (15.0, 5.25), (15.0, 9.0), (100.0, 5.5),
                                     (100.0, 30.0)])
def test_escape_direction_matches_exact_deflection(r_obs, b):
    P, d = _cam_dir_for_b(r_obs, b)
    bb = O.impact_parameter(P.astype(float), d.astype(float))
    phi = swept_to_infinity(bb, r_obs)
    exact = np.array([math.sin(phi), 0.0, math.cos(phi)])
    for mode in ("f64", "f32"):
        r = O.march(P, d, O.lawS_params(disk_mode=0), mode)
        assert r.cls == 2 and (r.flags & 2)
        err = escape_angle_error(np.array([r.esc[:]]), exact[None])[0]
        # discretisation error of the High preset (step 0.05 r, far-field exit at 0.01 rad);
        # a dropped far-field term alone costs > 1e-3 here.
        assert err < 1.5e-4, (mode, r_obs, b, err)


def test_far_field_term_matters():
    """Control for the test above: at an exit the shader can take (r = 25, first-order remaining
    deflection just under 0.01), dropping the far-field correction turns the direction by ~5e-3
    rad, far beyond that test's tolerance, so the tolerance does pin the correction."""
    b = 10.0
    r = 25.0
    x = np.array([b, 0.0, math.sqrt(r * r - b * b)])   # outgoing, impact parameter ~b
    v = np.array([0.0, 0.0, 1.0])
    h2 = float(np.dot(np.cross(x, v), np.cross(x, v)))
    ff = O.far_field_direction(x, v, h2)
    turn = math.acos(min(1.0, float(np.dot(ff, v))))
    assert 1e-3 < turn < 2e-2


def test_crossing_radius_matches_exact_orbit():
    rng = np.random.default_rng(1)
    cam = np.array([0.0, 30 * math.cos(math.radians(80)), 30 * math.sin(math.radians(80))],
                   np.float32)
    chord, root = [], []
    for _ in range(160):
        d = -cam / np.linalg.norm(cam) + rng.normal(0, 0.25, 3)
        d = (d / np.linalg.norm(d)).astype(np.float32)
        pd, dd = cam.astype(float), d.astype(float)
        b = O.impact_parameter(pd, dd)
        if b < B_CRIT * 1.01:
            continue
        v0 = O.initial_velocity(pd, dd)
        e1 = pd / np.linalg.norm(pd)
        e2 = v0 - np.dot(v0, e1) * e1
        e2 /= np.linalg.norm(e2)
        node = math.atan2(-e1[1], e2[1]) % math.pi     # first swept angle with y = 0
        exact = radius_at_swept_angle(b, 30.0, node)
        if exact is None:
            continue
        for rr, lst in ((0, chord), (1, root)):
            r = O.march(cam, d, O.lawS_params(disk_inner=2.5, disk_outer=60.0, root64=rr), "f64")
            if r.n_cross >= 1:
                lst.append(abs(r.disk_r[0] - exact) / exact)
    assert len(chord) > 80 and len(root) == len(chord)
    assert max(chord) < 2e-3 and np.median(chord) < 1e-4
    assert max(root) < 3e-5 and np.median(root) < 1e-6      # FP64 root: 100x tighter than chord


def test_rk4_order_and_law_T_agreement():
    """One law-S RK4 step (pinned fma order) agrees with law T's rk4Step (GLM order) to rounding,
    and halving the step shrinks the local error 2^5-fold (fourth-order method)."""
    x0 = np.array([7.0, 1.0, -3.0])
    v0 = O.initial_velocity(x0, np.array([-0.6, 0.1, 0.5]) / np.linalg.norm([-0.6, 0.1, 0.5]))
    h2 = float(np.dot(np.cross(x0, v0), np.cross(x0, v0)))

    def step(x, v, dt, n):
        for _ in range(n):
            x, v = O.rk4_step(x, v, h2, dt / n)
        return x, v

    ref, _ = step(x0, v0, 0.4, 256)
    e1 = np.linalg.norm(step(x0, v0, 0.4, 1)[0] - ref)
    e2 = np.linalg.norm(step(x0, v0, 0.2, 1)[0] - step(x0, v0, 0.2, 128)[0])
    assert 20.0 < e1 / e2 < 45.0
    # law S in double, stopped after exactly one step, equals law T's step to rounding
    P = x0.astype(np.float32)
    d = (np.array([-0.6, 0.1, 0.5]) / np.linalg.norm([-0.6, 0.1, 0.5])).astype(np.float32)
    v0f = O.initial_velocity(P.astype(float), d.astype(float))
    r = O.march(P, d, O.lawS_params(max_steps=1, disk_mode=0), "f64")
    assert r.steps == 1 and (r.flags & 1)


def test_escape_direction_stays_in_orbit_plane():
    rng = np.random.default_rng(5)
    cam = np.array([3.0, 9.0, 27.0], np.float32)
    for _ in range(50):
        d = -cam / np.linalg.norm(cam) + rng.normal(0, 0.3, 3)
        d = (d / np.linalg.norm(d)).astype(np.float32)
        n = np.cross(cam.astype(float), d.astype(float))
        n /= np.linalg.norm(n)
        r = O.march(cam, d, O.lawS_params(), "f64")
        if r.flags & 2:
            assert abs(float(np.dot(n, np.array(r.esc[:], float)))) < 1e-6


def test_redshift_matches_reference_closed_form():
    """Crossings of a 45-degree camera: g equals Geodesic.h:90-95's redshiftFactor (no clamp is
    active for r in [6, 20]) at the crossing radius and the ray's L_y/E."""
    from blackholerender_amd.camera import orbit_camera
    cam = orbit_camera(25.0, 45.0, 50.0, 48, 48)
    out = O.render(cam.as_dict(), O.lawS_params(), mode="f64", threads=4)
    pos = np.asarray(cam.pos, float)
    checked = 0
    for k in np.nonzero(out["n_cross"] > 0)[0][:200]:
        d = O.ray_dir(cam.as_dict(), int(k)).astype(float)
        b_axis = O.impact_parameter_about_axis(pos, d, (0.0, 1.0, 0.0))
        g_ref = O.redshift_factor(float(out["disk_r"][k, 0]), b_axis, float(np.linalg.norm(pos)))
        assert abs(out["disk_g"][k, 0] - g_ref) < 2e-6 * g_ref
        checked += 1
    assert checked > 50


def test_face_on_redshift_depends_only_on_radius():
    from blackholerender_amd.camera import orbit_camera
    cam = orbit_camera(40.0, 0.0, 60.0, 64, 64)    # on the axis: degenerate up, Camera.cpp:17
    out = O.render(cam.as_dict(), O.lawS_params(), mode="f64", threads=4)
    m = out["n_cross"] > 0
    assert m.sum() > 100
    r = out["disk_r"][m, 0].astype(float)
    expect = np.sqrt(1 - 3.0 / r) / np.sqrt(1 - 2.0 / 40.0)
    assert np.max(np.abs(out["disk_g"][m, 0] - expect) / expect) < 1e-5


def test_float_and_double_law_S_agree_on_reference_camera():
    from blackholerender_amd.camera import orbit_camera
    cam = orbit_camera(30.0, 85.0, 60.0, 256, 256).as_dict()
    p = O.lawS_params()
    a = O.render(cam, p, mode="band", threads=8)
    b = O.render(cam, p, mode="f64", threads=8)
    band = (a["flags"] & 4) > 0
    assert band.sum() > 0
    # outside the FP64 band, classes agree exactly and steps almost always
    out = ~band
    assert np.array_equal(a["cls"][out], b["cls"][out])
    assert (a["steps"][out] != b["steps"][out]).mean() < 1e-3
    esc = out & ((a["flags"] & 2) > 0) & ((b["flags"] & 2) > 0)
    err = escape_angle_error(a["esc"][esc], b["esc"][esc])
    assert err.max() < 1e-4 and np.percentile(err, 99) < 1e-5
    # inside the band, the FP64 march is the double march itself
    assert np.array_equal(a["esc"][band].view(np.uint32), b["esc"][band].view(np.uint32))


def test_class_converges_under_finer_steps():
    from blackholerender_amd.camera import orbit_camera
    cam = orbit_camera(30.0, 85.0, 60.0, 8, 8).as_dict()
    coarse = O.render(cam, O.lawS_params(), mode="f64", threads=2)
    fine = O.render(cam, O.lawS_params(max_steps=4500, step_scale=0.005, disk_step=0.018),
                    mode="f64", threads=2)
    assert np.array_equal(coarse["cls"], fine["cls"])


def test_ray_generation_matches_float64_restatement():
    from blackholerender_amd.camera import orbit_camera
    cam = orbit_camera(30.0, 85.0, 60.0, 37, 21)
    c = cam.as_dict()
    B = np.asarray(c["basis"], float).reshape(3, 3).T      # columns right, up, back
    for pix in (0, 18, 20 * 37 + 36, 10 * 37 + 18, 5 * 37 + 30):
        col, row = pix % 37, pix // 37
        ndc = np.array([(col + 0.5) / 37 * 2 - 1, (row + 0.5) / 21 * 2 - 1])
        dc = np.array([ndc[0] * c["tan_half_fov"] * 37 / 21, ndc[1] * c["tan_half_fov"], -1.0])
        d = B @ (dc / np.linalg.norm(dc))
        d /= np.linalg.norm(d)
        assert np.max(np.abs(O.ray_dir(c, pix) - d)) < 3e-7
    # the centre pixel of an odd image looks down -back, i.e. at the hole
    centre = O.ray_dir(c, 10 * 37 + 18).astype(float)
    to_hole = -np.asarray(c["pos"], float) / np.linalg.norm(c["pos"])
    assert np.linalg.norm(centre - to_hole) < 1e-6


def test_atan2_ieee_accuracy():
    rng = np.random.default_rng(3)
    y = rng.normal(size=20000).astype(np.float32)
    x = rng.normal(size=20000).astype(np.float32)
    got = np.array([O.atan2f(float(a), float(b)) for a, b in zip(y, x)], np.float64)
    ref = np.arctan2(y.astype(np.float64), x.astype(np.float64))
    assert np.max(np.abs(got - ref)) < 4e-7
    assert O.atan2f(0.0, 0.0) == 0.0 and O.atan2f(1.0, 0.0) == pytest.approx(math.pi / 2)


def test_pcg_matches_independent_restatement():
    def pcg(v):
        state = (v * 747796405 + 2891336453) & 0xFFFFFFFF
        word = (((state >> ((state >> 28) + 4)) ^ state) * 277803737) & 0xFFFFFFFF
        return ((word >> 22) ^ word) & 0xFFFFFFFF
    for v in (0, 1, 2, 12345, 0xFFFFFFFF, 0x80000000, 987654321):
        assert O.pcg(v) == pcg(v)
    u = np.array([O.pcg(i) >> 8 for i in range(20000)]) * 2.0 ** -24
    assert abs(u.mean() - 0.5) < 0.01 and u.min() >= 0.0 and u.max() < 1.0
