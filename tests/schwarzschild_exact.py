"""Exact Schwarzschild null geodesics by one-dimensional quadrature (test reference).

The orbit equation (du/dphi)^2 = 1/b^2 - u^2 (1 - 2u), u = 1/r (M = 1), integrated in closed
quadrature form (Darwin 1959; Chandrasekhar, The Mathematical Theory of Black Holes, sec. 20):
swept angle from u_a to u_b on one branch = integral of du / sqrt(1/b^2 - u^2 (1 - 2u)).
With F(u) = u^2 (1 - 2u) and the turning point u0 (F(u0) = 1/b^2, the smallest positive root),
F(u0) - F(u) = (u0 - u) Q(u), Q(u) = (u0 + u) - 2 (u0^2 + u0 u + u^2); the substitution
u = u0 (1 - t^2) turns the integrand into 2 sqrt(u0) / sqrt(Q(u)), which is smooth on [0, 1].
"""
import math

from scipy.integrate import quad
from scipy.optimize import brentq

B_CRIT = 3.0 * math.sqrt(3.0)


def turning_u(b: float) -> float:
    """Smallest positive root u0 of u^2 (1 - 2u) = 1/b^2 (b > b_crit, so u0 < 1/3)."""
    return brentq(lambda u: u * u * (1.0 - 2.0 * u) - 1.0 / (b * b), 0.0, 1.0 / 3.0, xtol=1e-16,
                  rtol=1e-15, maxiter=500)


def _branch(u0: float, t_hi: float) -> float:
    """Swept angle from u = u0 (1 - t_hi^2) to the turning point u0."""
    su0 = math.sqrt(u0)

    def h(t):
        u = u0 * (1.0 - t * t)
        q = (u0 + u) - 2.0 * (u0 * u0 + u0 * u + u * u)
        return 2.0 * su0 / math.sqrt(q)

    return quad(h, 0.0, t_hi, epsabs=1e-13, epsrel=1e-13, limit=400)[0]


def swept_to_infinity(b: float, r_obs: float) -> float:
    """Total swept angle of an ingoing ray from r_obs to infinity (b > b_crit)."""
    u0 = turning_u(b)
    uo = 1.0 / r_obs
    return _branch(u0, math.sqrt(1.0 - uo / u0)) + _branch(u0, 1.0)


def radius_at_swept_angle(b: float, r_obs: float, phi: float):
    """r of an ingoing ray (b > b_crit) once it has swept `phi` from r_obs, or None if it escapes
    to infinity first."""
    u0 = turning_u(b)
    uo = 1.0 / r_obs
    t_obs = math.sqrt(1.0 - uo / u0)
    to_peri = _branch(u0, t_obs)
    if phi <= to_peri:   # ingoing branch: angle from u_obs to u is to_peri - branch(u)
        f = lambda t: (to_peri - _branch(u0, t)) - phi
        t = brentq(f, 0.0, t_obs, xtol=1e-15, rtol=1e-15)
    else:                # outgoing branch: angle beyond periapsis
        rest = phi - to_peri
        if rest >= _branch(u0, 1.0):
            return None
        t = brentq(lambda t: _branch(u0, t) - rest, 0.0, 1.0, xtol=1e-15, rtol=1e-15)
    return 1.0 / (u0 * (1.0 - t * t))
