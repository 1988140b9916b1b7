"""The glass Fresnel power x^5 (cuda-path-tracer_amd/csrc/pt_libm.h) on the CPU.

The reference's host build evaluates pow(1 - cosine, 5) with the host libm's
pow(double, double) (src/renderer/transport.h:147).  The renderer computes x^5
correctly rounded with exact integer arithmetic and replaces the inputs where
the host libm differs, from a table the host builds.  Pinned here by:
- pow5_rounded against exact rational arithmetic (Python integers) on edge
  cases and random floats of every magnitude;
- the table-corrected value against Python's math.pow, which calls the same
  host libm pow, on every table entry and a large sample of the table domain
  (every float 1 - c for a float c in [0, 4]).
"""
import math
import os
import sys
from fractions import Fraction

import numpy as np

sys.path.insert(0, os.path.join(os.path.dirname(os.path.abspath(__file__)), "native"))
import wide_model as wm  # noqa: E402


def _exact_double(x):
    """The double nearest x^5 (ties to even), from exact rational arithmetic."""
    v = Fraction(float(x)) ** 5
    if v == 0:
        return math.copysign(0.0, float(x))
    sign = -1 if v < 0 else 1
    a = abs(v)
    e = a.numerator.bit_length() - a.denominator.bit_length()
    while Fraction(2) ** e > a:
        e -= 1
    while Fraction(2) ** (e + 1) <= a:
        e += 1
    scaled = a / Fraction(2) ** (e - 52)       # in [2^52, 2^53)
    q, r = divmod(scaled.numerator, scaled.denominator)
    if 2 * r > scaled.denominator or (2 * r == scaled.denominator and q % 2):
        q += 1
    return sign * math.ldexp(float(q), e - 52)


def _floats(n, seed):
    rng = np.random.default_rng(seed)
    bits = rng.integers(0, 1 << 32, n, dtype=np.uint64).astype(np.uint32)
    x = bits.view(np.float32)
    return x[np.isfinite(x)]


def test_pow5_rounded_is_correctly_rounded():
    edge = np.array([0.0, -0.0, 1.0, -1.0, 2.0, 0.5, 1.5, -0.5, 3.0, 1e-45, -1e-45, 1.17549435e-38, 3.4028235e38,
                     -3.4028235e38, 2.0 ** -24, 1 - 2.0 ** -24, float.fromhex("0x1.00c82cp-24"), float.fromhex("0x1.fffffep-1"), 0.7, 0.1, -2.5],
                    np.float32)
    x = np.concatenate([edge, _floats(20000, 1), np.random.default_rng(2).uniform(-3, 1, 20000).astype(np.float32)])
    exact, _, _ = wm.pow5(x)
    want = np.array([_exact_double(v) for v in x])
    assert (exact.view(np.int64) == want.view(np.int64)).all()


def test_table_reproduces_the_host_libm():
    rng = np.random.default_rng(3)
    # every float 1 - c for float c in [0, 4] is k * 2^-24, k in [-3 * 2^24, 2^24]
    k = rng.integers(-3 * (1 << 24), (1 << 24) + 1, 400000)
    x = (k.astype(np.float32) * np.float32(2.0 ** -24)).astype(np.float32)
    c = rng.uniform(0, 4, 200000).astype(np.float32)
    x = np.concatenate([x, (np.float32(1) - c).astype(np.float32)])
    exact, libm, entries = wm.pow5(x)
    host = np.array([math.pow(float(v), 5.0) for v in x])
    assert (libm.view(np.int64) == host.view(np.int64)).all()
    # glibc's pow is not correctly rounded everywhere: the table is not empty,
    # and without it some of these inputs would differ
    assert entries > 1000
    assert (exact != host).any()
