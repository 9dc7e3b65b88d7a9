"""Recipes of the fitted constants compiled into the kernels and the oracle.

Two tables are fitted, not derived in closed form:

* the odd polynomial of the IEEE-only atan2 (``atan2_ieee`` in ``csrc/lawS_device.hpp``,
  ``oracle_atan2f`` in ``oracle/bhs_oracle.c``): atan(u) ~ u + u^3 P(u^2) on
  0 <= u <= tan(pi/8), P of degree 4, least squares of the absolute error at 1000 Chebyshev nodes
  of the first kind;
* the Chebyshev series of ahat (``kAhatC`` in ``csrc/mlp_device.hpp``): the total swept angle of
  a Schwarzschild null geodesic from infinity to infinity minus pi, as a function of
  s = log(beta - 1), beta = b / b_crit, on log 1e-3 <= s <= log 200; 25 terms, least squares at
  200 Chebyshev nodes of the first kind. The swept angle is Darwin's integral
  2 int_0^{u0} du / sqrt(1/b^2 - u^2 (1 - 2u)) with u = u0 (1 - t^2), which leaves a smooth
  integrand on [0, 1].

``python -m blackholerender_amd.fit_constants`` prints both tables as C initialisers (float32,
nine significant digits, which round-trip). tests/test_fitted_constants.py refits both and checks
them against the tables in the sources.
"""
from __future__ import annotations

import math

import numpy as np

TAN_PI_8 = math.tan(math.pi / 8.0)
ATAN_DEGREE = 4
ATAN_NODES = 1000

AHAT_TERMS = 25
AHAT_NODES = 200
AHAT_S0 = math.log(1e-3)
AHAT_S1 = math.log(200.0)
B_CRIT = 3.0 * math.sqrt(3.0)


def _cheb_nodes(n: int) -> np.ndarray:
    """Chebyshev nodes of the first kind on [-1, 1]."""
    return np.cos(np.pi * (np.arange(n) + 0.5) / n)


def fit_atan_poly() -> np.ndarray:
    """Coefficients p0..p4 (float64) of atan(u) ~ u + u^3 (p0 + p1 u^2 + ... + p4 u^8)."""
    u = 0.5 * TAN_PI_8 * (_cheb_nodes(ATAN_NODES) + 1.0)
    a = np.stack([u ** (3 + 2 * k) for k in range(ATAN_DEGREE + 1)], axis=1)
    return np.linalg.lstsq(a, np.arctan(u) - u, rcond=None)[0]


def atan_poly_error(coef, n: int = 200001) -> float:
    """Largest |u + u^3 P(u^2) - atan u| over a fine grid of [0, tan(pi/8)] (float64)."""
    u = np.linspace(0.0, TAN_PI_8, n)
    p = np.polynomial.polynomial.polyval(u * u, np.asarray(coef, np.float64))
    return float(np.max(np.abs(u + u ** 3 * p - np.arctan(u))))


def _turning_u(b: float) -> float:
    """Smallest positive root u0 of u^2 (1 - 2u) = 1/b^2 for b > b_crit (u0 < 1/3)."""
    from scipy.optimize import brentq
    return brentq(lambda u: u * u * (1.0 - 2.0 * u) - 1.0 / (b * b), 0.0, 1.0 / 3.0,
                  xtol=1e-16, rtol=1e-15, maxiter=500)


def total_sweep(b: float) -> float:
    """Swept angle of a null geodesic with impact parameter b > b_crit from infinity to infinity."""
    from scipy.integrate import quad
    u0 = _turning_u(b)
    su0 = math.sqrt(u0)

    def g(t):
        u = u0 * (1.0 - t * t)
        # u0^2 (1 - 2 u0) - u^2 (1 - 2u) = (u0 - u) [(u0 + u) - 2 (u0^2 + u0 u + u^2)]
        q = (u0 + u) - 2.0 * (u0 * u0 + u0 * u + u * u)
        return 2.0 * su0 / math.sqrt(q)

    return 2.0 * quad(g, 0.0, 1.0, epsabs=1e-13, epsrel=1e-13, limit=400)[0]


def ahat_exact(s: float) -> float:
    """ahat(s) = total_sweep(b_crit (1 + e^s)) - pi."""
    return total_sweep(B_CRIT * (1.0 + math.exp(s))) - math.pi


def _to_s(x):
    return 0.5 * (AHAT_S0 + AHAT_S1) + 0.5 * (AHAT_S1 - AHAT_S0) * np.asarray(x)


def fit_ahat_chebyshev() -> np.ndarray:
    """Chebyshev coefficients c0..c24 (float64) of ahat in t = (s - mid) / half on [-1, 1]."""
    x = _cheb_nodes(AHAT_NODES)
    y = np.array([ahat_exact(s) for s in _to_s(x)])
    return np.polynomial.chebyshev.chebfit(x, y, AHAT_TERMS - 1)


def ahat_series_error(coef, n: int = 301) -> float:
    """Largest |series - ahat| (float64 evaluation) over n points of the domain."""
    x = np.linspace(-1.0, 1.0, n)
    ser = np.polynomial.chebyshev.chebval(x, np.asarray(coef, np.float64))
    return float(np.max(np.abs(ser - np.array([ahat_exact(s) for s in _to_s(x)]))))


def c_floats(coef) -> str:
    """float32 C initialiser list with nine significant digits (round-trips float32)."""
    return ", ".join(f"{float(np.float32(c)):.9e}f" for c in coef)


def main() -> None:
    at = fit_atan_poly()
    print("// atan(u) ~ u + u^3 P(u^2), P coefficients p0..p4, max error "
          f"{atan_poly_error(at):.2e} before rounding")
    print(c_floats(at))
    ah = fit_ahat_chebyshev()
    print(f"// ahat Chebyshev c0..c{AHAT_TERMS - 1}, max error {ahat_series_error(ah):.2e} rad")
    print(c_floats(ah))


if __name__ == "__main__":
    main()
