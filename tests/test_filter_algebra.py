"""The filter and stage kernels (solar_amd/csrc/solar_kernels.hip) skip or
share floating-point operations whose results cannot change a stored bit:

  - FILTER_R (filter.f90:682-693): the eleven taps are symmetric, so each
    plane's FILTER_ETA value v enters the eleven pending outputs through six
    products filtr[m] * v, accumulated in tap order, instead of eleven
    products per output over a window of the last eleven planes;
  - FILTER_XI and FILTER_ETA sum from their first product instead of from
    0.0 (filter.f90:176-182, 448-455): that changes at most the sign of a
    zero result, a sum's value does not depend on the signs of zero terms,
    and the stored sum (FILTER_R's, from 0.0) never gives -0, so the output
    bits are the reference's;
  - the stage's xi/eta DRP stencils (derivative.f90:137-148, 456-467) form
    no product for the zero centre tap.

These CPU tests restate both forms in IEEE double (Python floats: every
multiply and add rounded on its own, as the kernels are compiled) and
compare bits on adversarial data -- zeros of both signs, subnormals, exact
cancellations -- and show that the sign-of-zero case the second rewrite
relies on does occur.
"""
import math
import random
import struct

# header:99-105 and header:123-133 (the reference's truncated decimals)
FILT = [0.015625, -0.09375, 0.234375, 0.6875, 0.234375, -0.09375, 0.015625]
FILTR = [0.0009765, -0.0097655, 0.043945, -0.1171875, 0.2050775, 0.753906,
         0.2050775, -0.1171875, 0.043945, -0.0097655, 0.0009765]
DRP = [-0.02651995, 0.18941314, -0.79926643, 0.0, 0.79926643, -0.18941314, 0.02651995]  # header:40-46


def _bits(x):
    return struct.unpack("<Q", struct.pack("<d", x))[0]


def _adversarial(rng, count):
    """Values that stress signs of zero, subnormals and cancellation."""
    pool = [0.0, -0.0, 5e-324, -5e-324, 2.2250738585072014e-308, -1e-310, 1.0, -1.0, 0.5, 3.0]
    out = []
    for _ in range(count):
        r = rng.random()
        if r < 0.3:
            out.append(rng.choice(pool))
        elif r < 0.5:
            out.append(math.ldexp(rng.uniform(-1, 1), rng.randint(-1074, -1000)))
        else:
            out.append(rng.uniform(-2.0, 2.0) * 10.0 ** rng.randint(-5, 5))
    return out


def _sum_from_zero(coeffs, vals):
    s = 0.0
    for c, v in zip(coeffs, vals):
        s = s + c * v
    return s


def _sum_from_first(coeffs, vals):
    s = coeffs[0] * vals[0]
    for c, v in zip(coeffs[1:], vals[1:]):
        s = s + c * v
    return s


def test_filter_taps_are_symmetric():
    assert all(FILTR[q] == FILTR[10 - q] for q in range(11))
    assert all(FILT[q] == FILT[6 - q] for q in range(7))
    assert DRP[3] == 0.0


def test_radial_products_shared_equal_window_sums():
    rng = random.Random(11)
    for trial in range(200):
        n = 30
        v = _adversarial(rng, n)
        if trial % 7 == 0:  # a run of products that are all -0 (and exact cancellations)
            v = [(-0.0 if FILTR[p % 11] > 0 else 0.0) for p in range(n)]
        # the reference: output k from 0.0 over the window k-5 .. k+5
        ref = {k: _sum_from_zero(FILTR, v[k - 5:k + 6]) for k in range(5, n - 5)}
        # the kernel: acc[d] is output p - 5 + d when plane p arrives
        acc = [0.0] * 11
        got = {}
        for p in range(n):
            pr = [FILTR[m] * v[p] for m in range(6)]
            for d in range(10):
                acc[d] = acc[d] + pr[d if d < 6 else 10 - d]
            acc[10] = 0.0 + pr[0]
            if p - 10 >= 0:
                got[p - 5] = acc[0]
            acc = acc[1:] + [acc[10]]
        for k in ref:
            assert _bits(got[k]) == _bits(ref[k]), (trial, k)


def test_first_product_sums_change_only_zero_signs_the_next_pass_ignores():
    rng = random.Random(5)
    sign_cases = 0
    for trial in range(3000):
        x = _adversarial(rng, 7)
        if trial % 10 == 0:  # every product -0: the sum from 0.0 gives +0, from the first product -0
            x = [(-0.0 if c > 0 else 0.0) for c in FILT]
        a = _sum_from_zero(FILT, x)
        b = _sum_from_first(FILT, x)
        assert a == b  # numerically equal always
        if _bits(a) != _bits(b):
            assert a == 0.0 and math.copysign(1.0, a) > 0 and math.copysign(1.0, b) < 0
            sign_cases += 1
        # the next pass (FILTER_ETA after FILTER_XI, FILTER_R after
        # FILTER_ETA) sums from 0.0 over several such values
        w = _adversarial(rng, 10)
        for pos in (0, 4, 10):
            va = w[:pos] + [a] + w[pos:]
            vb = w[:pos] + [b] + w[pos:]
            assert _bits(_sum_from_zero(FILTR, va)) == _bits(_sum_from_zero(FILTR, vb))
    assert sign_cases > 0  # the rewrite does change zero signs, which the next pass then hides


def test_zero_centre_tap_leaves_stencil_sums_unchanged():
    rng = random.Random(9)
    taps = [q for q in range(7) if q != 3]
    for _ in range(5000):
        u = _adversarial(rng, 7)
        full = _sum_from_zero(DRP, u)
        skip = 0.0
        for q in taps:
            skip = skip + DRP[q] * u[q]
        assert _bits(full) == _bits(skip)
    # a non-finite centre value is where the two differ (0 * inf = nan)
    u = [1.0, 2.0, 3.0, math.inf, 5.0, 6.0, 7.0]
    assert math.isnan(_sum_from_zero(DRP, u))


def test_filter_chain_equals_reference_bits():
    """FILTER_XI, FILTER_ETA and FILTER_R on 11 planes of 7 x 7 values: the
    kernel's forms (first-product xi and eta, shared radial products) against
    the reference's (every sum from 0.0, a window per output)."""
    rng = random.Random(21)
    for trial in range(300):
        if trial % 5 == 0:  # blocks of signed zeros whose products are all -0 in some rows
            u = [[[(-0.0 if FILT[i] > 0 else 0.0) if rng.random() < 0.8 else rng.choice([1e-310, -2.0, 0.5])
                   for i in range(7)] for _ in range(7)] for _ in range(11)]
        else:
            u = [[_adversarial(rng, 7) for _ in range(7)] for _ in range(11)]
        x_ref = [[_sum_from_zero(FILT, row) for row in plane] for plane in u]
        e_ref = [_sum_from_zero(FILT, xs) for xs in x_ref]
        out_ref = _sum_from_zero(FILTR, e_ref)
        x_k = [[_sum_from_first(FILT, row) for row in plane] for plane in u]
        e_k = [_sum_from_first(FILT, xs) for xs in x_k]
        acc = [0.0] * 11
        for p, v in enumerate(e_k):
            pr = [FILTR[m] * v for m in range(6)]
            for d in range(10):
                acc[d] = acc[d] + pr[d if d < 6 else 10 - d]
            acc[10] = 0.0 + pr[0]
            out_k = acc[0]
            acc = acc[1:] + [acc[10]]
        assert _bits(out_k) == _bits(out_ref), trial
