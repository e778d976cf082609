"""The stage kernels divide by C, D, r and rho0 through their reciprocals
(quot_rcp and div_group in solar_amd/csrc/solar_kernels.hip): q = RN(a y)
with y = RN(1/b), then two corrections q <- RN(q - r y) with the negated
remainder r = b q - a (FMA), numerators with a frexp exponent outside
[-900, 900] taken by the hardware division, and the divisors held to
[2^-60, 2^60] by the host.  These
CPU tests restate that sequence with an exact FMA (rational arithmetic,
rounded once) and check it against IEEE division, bit for bit, on the
divisors of real grids and on numerators across the whole double range:
zeros of both signs, subnormals, binade edges.  Variants with one of the
safeguards removed fail, so the tests can tell.
"""
import math
import struct
from fractions import Fraction

import numpy as np

from solar_amd import background, host_tables


def _bits(x):
    return struct.unpack("<Q", struct.pack("<d", x))[0]


def _fma(x, y, z):
    """RN(x*y + z) with one rounding, IEEE zero signs."""
    exact = Fraction(x) * Fraction(y) + Fraction(z)
    if exact != 0:
        return float(exact)
    prod_neg = (math.copysign(1.0, x) * math.copysign(1.0, y)) < 0
    prod_zero = x == 0.0 or y == 0.0
    if prod_zero and z == 0.0:  # a sum of two zeros: -0 only if both are -0
        return -0.0 if (prod_neg and math.copysign(1.0, z) < 0) else 0.0
    return 0.0  # exact cancellation of nonzero terms rounds to +0


def _div_by(a, b, y, negated=True, guard=True):
    if guard and not (-900 <= math.frexp(a)[1] <= 900):
        return a / b
    q0 = a * y
    if negated:
        r0 = _fma(b, q0, -a)
        q1 = _fma(-r0, y, q0)
        r1 = _fma(b, q1, -a)
        return _fma(-r1, y, q1)
    r0 = _fma(-b, q0, a)
    q1 = _fma(r0, y, q0)
    r1 = _fma(-b, q1, a)
    return _fma(r1, y, q1)


def _divisors():
    """C, D, r and rho0 of a few grids and both backgrounds."""
    out = []
    for n, nr, kind in ((32, 64, 0), (96, 192, 1), (256, 384, 1)):
        gr, c2, rho0, g0 = background(kind, nr)
        t = host_tables(n, nr, gr, c2, rho0, g0, max(1, round(0.08 * nr)), 0.1)
        metric_cd = t["n1"][4 * n:6 * n]  # metric_C, metric_D (the stage's C(i), D(j))
        out += list(metric_cd) + list(gr) + list(rho0)
    return np.unique(np.array(out))


def _numerators(rng, count):
    vals = [0.0, -0.0, 5e-324, -5e-324, 2.2250738585072014e-308, 1.0, -1.0, 2.0 ** -900, 2.0 ** 900,
            float.fromhex("0x1.fffffffffffffp-901")]
    for _ in range(count):
        e = int(rng.integers(-1074, 1024))
        m = float(rng.integers(0, 2 ** 52)) / 2.0 ** 52
        if rng.random() < 0.15:  # significands next to the binade edges
            m = float(rng.integers(0, 256)) / 2.0 ** 52 if rng.random() < 0.5 else 1.0 - float(rng.integers(1, 257)) / 2.0 ** 52
        v = math.ldexp(1.0 + m, e) if e >= -1022 else math.ldexp(m if m > 0 else 2.0 ** -52, -1022)
        vals.append(-v if rng.random() < 0.5 else v)
    return vals


def test_reciprocal_division_equals_ieee_division():
    rng = np.random.default_rng(7)
    divs = _divisors()
    assert divs.min() >= 2.0 ** -60 and divs.max() <= 2.0 ** 60  # the host's fast-path condition holds
    a_vals = _numerators(rng, 4000)
    bad = 0
    for i, a in enumerate(a_vals):
        for b in (float(divs[(7 * i) % len(divs)]), float(divs[(7 * i + 3) % len(divs)]),
                  math.ldexp(1.0 + rng.random(), int(rng.integers(-60, 60)))):
            if _bits(_div_by(a, b, 1.0 / b)) != _bits(a / b):
                bad += 1
    assert bad == 0


def test_division_safeguards_are_needed():
    """Without the negated remainder a -0 numerator gives +0; without the
    exponent guard subnormal numerators round differently."""
    b = 0.7853981633974483
    assert _bits(_div_by(-0.0, b, 1.0 / b)) == _bits(-0.0 / b)
    assert _bits(_div_by(-0.0, b, 1.0 / b, negated=False)) != _bits(-0.0 / b)
    rng = np.random.default_rng(3)
    mismatches = 0
    for _ in range(3000):  # numerators near and below the normal range (about 2% differ unguarded)
        a = math.ldexp(1.0 + rng.random(), int(rng.integers(-1074, -1000)))
        bb = math.ldexp(1.0 + rng.random(), int(rng.integers(-10, 10)))
        mismatches += _bits(_div_by(a, bb, 1.0 / bb, guard=False)) != _bits(a / bb)
        assert _bits(_div_by(a, bb, 1.0 / bb)) == _bits(a / bb)
    assert mismatches > 0
