"""Pins of the oracle's volumetric disk march (oracle/vol_body.inc: shaders/blackhole.frag:142-253
with DISK_VOLUMETRIC, shaders/disk.glsl:81-110).

- the cubic Hermite sub-sampling (blackhole.frag:224-227) reproduces any cubic from its end values
  and tangents, and a straight segment exactly — a wrong basis coefficient fails;
- volumetricDiskStep obeys Beer-Lambert: two half-length steps at one point give the transmittance
  and colour of one whole step, colour = source (1 - T) from an empty composite — an absorption not
  exponential in the path length, or a colour not weighted by the running transmittance, fails;
- with no disk matter (density 0, or opacity 0) the march adds no colour and every escaping ray
  keeps its whole sky; the escape directions of rays running through the slab match the exact
  Schwarzschild deflection (quadrature) — pins the volumetric step control and the march;
- the float and double marches agree on class and colour on the 1080p camera's view.
"""
import math

import numpy as np
import pytest

import oracle as O
from helpers import escape_angle_error
from schwarzschild_exact import swept_to_infinity


def test_hermite_reproduces_cubics():
    rng = np.random.default_rng(3)
    for _ in range(20):
        c = rng.normal(size=(4, 3))                      # p(t) = c0 + c1 t + c2 t^2 + c3 t^3
        p = lambda t: c[0] + c[1] * t + c[2] * t * t + c[3] * t ** 3
        dp = lambda t: c[1] + 2 * c[2] * t + 3 * c[3] * t * t
        for t in (0.0, 0.125, 0.5, 0.8125, 1.0):
            np.testing.assert_allclose(O.hermite(p(0), dp(0), p(1), dp(1), t), p(t), atol=1e-12)
    x0, x1 = np.array([1.0, -2.0, 3.0]), np.array([4.0, 0.5, -1.0])
    for t in np.linspace(0, 1, 9):
        np.testing.assert_allclose(O.hermite(x0, x1 - x0, x1, x1 - x0, t), x0 + t * (x1 - x0),
                                   atol=1e-14)


def test_volume_step_is_beer_lambert():
    sh = O.shading(opacity=2.0, density=3.0)
    lapse = float(np.sqrt(1.0 - 2.0 / 30.0))
    p = (0.0, 0.02, 8.0)
    r = 8.0
    for ds in (0.05, 0.18, 0.6):
        one = O.vol_sample(sh, 0.035, lapse, p, r, ds, 2.0, (0, 0, 0, 1))
        half = O.vol_sample(sh, 0.035, lapse, p, r, ds / 2, 2.0, (0, 0, 0, 1))
        two = O.vol_sample(sh, 0.035, lapse, p, r, ds / 2, 2.0, half)
        assert 0.0 < one[3] < 1.0
        assert abs(two[3] - one[3]) < 2e-6 * max(1.0, one[3])
        np.testing.assert_allclose(two[:3], one[:3], rtol=2e-5)
        src = one[:3] / (1.0 - one[3])                  # colour = source (1 - T) from T = 1
        dim = O.vol_sample(sh, 0.035, lapse, p, r, ds, 2.0, (0.1, 0.2, 0.3, 0.25))
        np.testing.assert_allclose(dim[:3], np.array([0.1, 0.2, 0.3]) + 0.25 * (1 - one[3]) * src,
                                   rtol=2e-5)
        assert abs(dim[3] - 0.25 * one[3]) < 1e-7


def _dir_for_b_in_plane(r_obs, b):
    """Camera on +z in the disk plane (y = 0), ray in that plane with impact parameter b."""
    a = math.asin(b * math.sqrt(1.0 - 2.0 / r_obs) / r_obs)
    return (np.array([0.0, 0.0, r_obs], np.float32),
            np.array([math.sin(a), 0.0, -math.cos(a)], np.float32))


@pytest.mark.parametrize("r_obs,b", [(30.0, 6.0), (30.0, 8.0), (30.0, 14.0), (15.0, 9.0)])
def test_empty_disk_keeps_sky_and_exact_deflection(r_obs, b):
    P, d = _dir_for_b_in_plane(r_obs, b)
    phi = swept_to_infinity(O.impact_parameter(P.astype(float), d.astype(float)), r_obs)
    exact = np.array([math.sin(phi), 0.0, math.cos(phi)])
    p = O.lawS_params(disk_mode=2)
    for sh in (O.shading(density=0.0), O.shading(opacity=0.0)):
        for mode in ("f32", "f64"):
            o = O.vol_ray(P, d, p, sh, mode)
            assert o["status"] == 1 and o["sky_weight"] == 1.0 and np.all(o["rgb"] == 0.0)
            err = escape_angle_error(o["esc"][None].astype(np.float64), exact[None])[0]
            assert err < 1.5e-4, (mode, err)
    # inside the slab the volumetric rule caps steps at 6 disk_step: more steps than the thin
    # march of the same ray; with the default matter the in-plane ray turns opaque and glows
    clear = O.vol_ray(P, d, p, O.shading(opacity=0.0), "f32")
    thin_steps = O.march(P, d, O.lawS_params(disk_mode=1), "f32").steps
    assert clear["steps"] >= thin_steps
    if b >= 14.0:      # farMul 2.7: the law-S step exceeds the cap inside the slab
        assert clear["steps"] > thin_steps
    thick = O.vol_ray(P, d, p, O.shading(), "f32")
    assert thick["status"] == 3 and thick["sky_weight"] == 0.0 and thick["rgb"].sum() > 0


def test_float_and_double_marches_agree():
    from blackholerender_amd.camera import orbit_camera
    cam = orbit_camera(30.0, 85.0, 60.0, 64, 36, jitter_hashed=1, seed=4).as_dict()
    p = O.lawS_params(disk_mode=2)
    sh = O.shading(doppler=1.0)
    a = O.volumetric(cam, p, sh, mode="f32")
    b = O.volumetric(cam, p, sh, mode="f64")
    assert (a["cls"] == b["cls"]).mean() > 0.995
    la, lb = a["rgb"].sum(1), b["rgb"].sum(1)
    same = a["cls"] == b["cls"]
    rel = np.abs(la - lb)[same] / np.maximum(lb[same], 1e-3)
    assert np.median(rel) < 1e-4 and np.quantile(rel, 0.99) < 0.05
    assert (a["cls"] == 2).mean() > 0.5 and la.max() > 1.0
