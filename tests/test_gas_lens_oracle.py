"""CPU pins of the oracle's gas lens table (shaders/gas_lens.comp:11-44, gas_lens.glsl:13-23).

The reference holds no tests for the lens table; the restatement is pinned by:
  * the exact Schwarzschild orbit: for every fan ray that escapes (b > b_crit), r at each tabulated
    swept angle matches r(phi) from Darwin's quadrature (tests/schwarzschild_exact.py);
  * capture and escape: rays below b_crit end at log 2, rays above it at log r_max;
  * row 0 is log r_cam; the fan angle is lensAlpha's closed form (monotone, alpha_crit at u_c);
  * "for a fixed phi, r grows with alpha" (gas_lens.glsl:9) on the escaping rays.
"""
import math

import numpy as np

import oracle as O
from schwarzschild_exact import B_CRIT, radius_at_swept_angle, swept_to_infinity


def _alpha(i, n, ac):
    u = (i + 0.5) / n
    uc = ac / math.pi
    if u < uc:
        return ac - ac * ((uc - u) / uc) ** 3
    return ac + (math.pi - ac) * ((u - uc) / (1 - uc)) ** 3


def _b(cam_r, a):
    d = np.array([-math.cos(a), math.sin(a), 0.0], np.float32).astype(np.float64)
    return O.impact_parameter([cam_r, 0.0, 0.0], d)


def _sar(r):
    import blackholerender_amd as bhs
    return bhs.shadow_angular_radius_f32(r)


def test_lens_table_matches_the_exact_orbit():
    R, n_rays, n_phis = 30.0, 96, 257
    ac = _sar(R)
    t = O.gas_lens_table(R, ac, n_rays, n_phis)
    assert np.all(t[0] == np.float32(math.log(R)))
    dphi = 2 * math.pi / (n_phis - 1)
    errs = []
    for i in range(n_rays):
        a = _alpha(i, n_rays, ac)
        b = _b(R, a)
        if a >= math.pi / 2 or b < B_CRIT * 1.02:   # ingoing, escaping rays only
            continue
        for j in range(1, n_phis, 4):
            r = radius_at_swept_angle(b, R, j * dphi)
            if r is None or r > 1e3:   # beyond, the float sum of swept angles (~1e-6 rad) moves r
                break                  # by ~1e-6 r / b relative
            errs.append((abs(math.exp(float(t[j, i])) - r) / r, j * dphi))
    e = np.array([v for v, _ in errs])
    early = np.array([v for v, p in errs if p < math.pi / 2])
    assert len(errs) > 300 and len(early) > 100
    # the method's own truncation (steps of 0.4-1% of r) grows on near-critical rays' last turn
    assert e.max() < 2e-3 and np.median(e) < 1e-4, (e.max(), np.median(e))
    assert early.max() < 2e-4, early.max()


def test_capture_and_escape_fills():
    R, n_rays, n_phis = 12.0, 128, 129
    ac = _sar(R)
    t = O.gas_lens_table(R, ac, n_rays, n_phis)
    log_h, log_max = np.float32(math.log(2.0)), np.float32(math.log(2.0e5))
    checked = 0
    assert t.min() >= log_h and t.max() <= log_max + 1e-6
    for i in range(n_rays):
        a = _alpha(i, n_rays, ac)
        b = _b(R, a)
        if a < math.pi / 2 and b < B_CRIT * 0.99:
            assert t[-1, i] == log_h, i               # fell in
            k = int(np.argmax(t[:, i] == log_h))
            assert np.all(t[k:, i] == log_h)
        elif b > B_CRIT * 1.01:
            total = (swept_to_infinity(b, R) if a < math.pi / 2 else math.pi - a)
            if total < 2 * math.pi - 0.05:
                assert t[-1, i] == log_max, i         # escaped before a full turn
            else:
                assert t[-1, i] < log_max, i          # still winding at a full turn
            checked += 1
    assert checked > 60


def test_radius_grows_with_alpha_at_fixed_phi():
    R, n_rays, n_phis = 30.0, 256, 129
    ac = _sar(R)
    t = O.gas_lens_table(R, ac, n_rays, n_phis)
    first = next(i for i in range(n_rays) if _b(R, _alpha(i, n_rays, ac)) > B_CRIT * 1.001)
    esc = t[:, first:]
    assert np.all(np.diff(esc, axis=1) >= 0), np.argwhere(np.diff(esc, axis=1) < 0)[:5]


def test_captured_rays_fall_in_at_the_exact_swept_angle():
    """Rays inside the shadow (b < b_crit, lensAlpha's inner branch) reach r = 2.01, where the
    march stops, at the swept angle of the exact orbit: integral of du / sqrt(1/b^2 - u^2 (1 - 2u))
    from 1/r_cam to 1/2.01 (no turning point below b_crit)."""
    from scipy.integrate import quad
    R, n_rays, n_phis = 4.0, 64, 1025      # close in: the shadow spans a third of the fan
    ac = _sar(R)
    t = O.gas_lens_table(R, ac, n_rays, n_phis)
    log_h = np.float32(math.log(2.0))
    dphi = 2 * math.pi / (n_phis - 1)
    checked = 0
    for i in range(n_rays):
        a = _alpha(i, n_rays, ac)
        b = _b(R, a)
        if a >= math.pi / 2 or b > B_CRIT * 0.98:   # ingoing rays below b_crit only
            continue
        exact = quad(lambda u: 1.0 / math.sqrt(1.0 / b ** 2 - u * u * (1 - 2 * u)), 1.0 / R,
                     1.0 / 2.01, epsabs=1e-12, epsrel=1e-12)[0]
        first = int(np.argmax(t[:, i] == log_h)) * dphi
        # the fill starts at the first node after the step that crossed r = 2.01: up to one node
        # spacing plus one step's swept angle (ds = 0.008 at r = 2, ~0.01 rad) late
        assert exact - 1e-3 <= first <= exact + dphi + 0.012, (i, first, exact)
        checked += 1
    assert checked > 15
