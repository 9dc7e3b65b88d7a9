"""The device's reciprocal square root (bhs_rsqrt_rn) against the oracle's, bit for bit.

The oracle's value, (float)(1.0 / sqrt((double)x)), is the correctly rounded 1/sqrt(x) for every
float (tests/cpp/rsqrt_rn_check.c decides it exhaustively with integer arithmetic); numpy computes
the same formula here. Inputs: every float in [1, 4) (both exponent parities: 16,777,216 values),
every positive denormal, the two ends of the normal range, 4 million random positive normals over
the whole exponent range, and the special values. This runs the hardware v_rsq_f32 estimate through
the refinement on every significand the march can meet, on the GPU itself.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _expected(x: np.ndarray) -> np.ndarray:
    with np.errstate(divide="ignore", invalid="ignore"):
        return (1.0 / np.sqrt(x.astype(np.float64))).astype(np.float32)


def _device(tracer, x: np.ndarray) -> np.ndarray:
    import torch
    t = torch.from_numpy(x).to(f"cuda:{tracer.device}")
    y = tracer.rsqrt_rn(t)
    torch.cuda.synchronize()
    return y.cpu().numpy()


def _bits(a: np.ndarray) -> np.ndarray:
    return a.view(np.uint32)


@pytest.mark.parametrize("what", ["[1,4)", "denormals"])
def test_rsqrt_exhaustive_ranges(tracer, what):
    if what == "[1,4)":
        lo, hi = np.float32(1.0).view(np.uint32), np.float32(4.0).view(np.uint32)
        bits = np.arange(lo, hi, dtype=np.uint32)
    else:
        bits = np.arange(1, 0x800000, dtype=np.uint32)
    x = bits.view(np.float32)
    print(f"[rsqrt] {what}: {x.size} floats", flush=True)
    got, want = _device(tracer, x), _expected(x)
    bad = np.flatnonzero(_bits(got) != _bits(want))
    assert bad.size == 0, (f"{bad.size} mismatches, first x = {x[bad[0]]!r}: "
                           f"{got[bad[0]]!r} vs {want[bad[0]]!r}")


def test_rsqrt_random_normals_and_specials(tracer):
    rng = np.random.default_rng(5)
    bits = rng.integers(0x00800000, 0x7f800000, 4_000_000, dtype=np.uint32)
    x = np.concatenate([bits.view(np.float32),
                        np.array([np.finfo(np.float32).tiny, np.finfo(np.float32).max,
                                  0.0, np.inf], np.float32)])
    got, want = _device(tracer, x), _expected(x)
    np.testing.assert_array_equal(_bits(got), _bits(want))
    neg = _device(tracer, np.array([-0.0, -1.0, np.nan], np.float32))
    assert neg[0] == -np.inf and np.isnan(neg[1]) and np.isnan(neg[2])
