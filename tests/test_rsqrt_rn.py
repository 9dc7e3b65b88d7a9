"""The march's reciprocal square root is correctly rounded on both sides.

The float march takes 1/sqrt(x) rounded once (rsqrt_rn on the device: the v_rsq_f32 estimate and
one refinement in double; oracle_rsqrtf in the oracle: 1.0 / sqrt((double)x) rounded to float).
tests/cpp/rsqrt_rn_check.c decides the correctly rounded value of every float in [1, 4) with
integer arithmetic and checks both sequences against it, the device's for every hardware estimate
within 2 ulps (the instruction is documented within 1), plus every positive denormal and the
special values; scaling by 4^k extends the result to every float. Here the oracle's exported
function is compared with that formula and with the integer decision on a sample, so the library
the tests use is the one checked.
"""
import os
import subprocess

import numpy as np

import oracle as O

HERE = os.path.dirname(os.path.abspath(__file__))


def test_exhaustive_check_passes():
    subprocess.run(["make", "-s", "-C", os.path.join(HERE, "cpp"), "rsqrt_rn_check"], check=True)
    r = subprocess.run([os.path.join(HERE, "cpp", "rsqrt_rn_check")], capture_output=True,
                       text=True, timeout=300)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "not correctly rounded: 0" in r.stdout and "within 2 ulps: 0" in r.stdout


def _cr_rsqrt_exact(x: float) -> float:
    """Correctly rounded 1/sqrt(x) by exact rational comparison at the rounding midpoints."""
    from fractions import Fraction
    y = np.float32(1.0 / np.sqrt(np.float64(x)))
    fx = Fraction(float(x))
    for cand in (np.nextafter(y, np.float32(0)), y, np.nextafter(y, np.float32(np.inf))):
        lo = Fraction(float(np.nextafter(cand, np.float32(0))))
        hi = Fraction(float(np.nextafter(cand, np.float32(np.inf))))
        c = Fraction(float(cand))
        m_lo, m_hi = (lo + c) / 2, (c + hi) / 2
        if fx * m_lo * m_lo < 1 < fx * m_hi * m_hi:
            return float(cand)
    raise AssertionError(x)


def test_oracle_export_is_correctly_rounded():
    L = O.lib()
    rng = np.random.default_rng(11)
    xs = np.concatenate([
        np.frombuffer(rng.integers(0x00800000, 0x7f800000, 2000, dtype=np.uint32).tobytes(),
                      np.float32),
        np.array([1.0, 2.0, 4.0, float.fromhex("0x1.7431c6p+1"), float.fromhex("0x1.fffffep+1")],
                 np.float32)])
    for x in xs:
        got = L.oracle_rsqrtf(float(x))
        assert np.float32(got) == np.float32(1.0 / np.sqrt(np.float64(x)))
        assert float(np.float32(got)) == _cr_rsqrt_exact(np.float32(x)), x
