"""CPU checks of identities the stage kernels rely on, against the oracle
(the FP64 CPU restatement of the reference): the kernels evaluate the
limiter's half-slope in forms other than the reference's nested minmod, and
these tests show the forms agree on edge and random inputs."""
import numpy as np

import oracle_lib as O


def _edge_values():
    nz0 = np.copysign(0.0, -1.0)
    base = [0.0, nz0, 1.0, -1.0, 0.5, -0.5, 2.0, 3.0, -3.0, 1e-300, -1e-300, 7.25, 1.0 + 2 ** -52]
    return base


def test_half_slope_fp64_form_equals_reference_limiter():
    """The stage kernels' FP64 half-slope (sign * min(|a|, |c|, |b|/4), the sign
    of a, zero when a and c differ in sign bit) equals half of the reference's
    approximateDerivative on every triple of edge values (zeros of both
    signs, ties, tiny and ordinary magnitudes) and on random triples, up to
    the sign of a zero result."""
    def half_slope(x, y, z):
        a, b, c = y - x, z - x, z - y
        m = min(min(abs(a), abs(c)), abs(b) * 0.25)
        same = np.signbit(a) == np.signbit(c)
        return float(np.copysign(m, a)) if same else 0.0
    vals = _edge_values()
    rng = np.random.default_rng(4)
    triples = [(x, y, z) for x in vals for y in vals for z in vals]
    triples += [tuple(rng.standard_normal(3) * 10.0 ** rng.integers(-3, 3)) for _ in range(20000)]
    for x, y, z in triples:
        ref = O.lib().kto_approx_deriv(x, y, z) / 2
        got = half_slope(x, y, z)
        assert got == ref, (x, y, z, got, ref)


def test_half_slope_fp32_median_form():
    """The mixed form's FP32 half-slope med3(a, med3(b/4, c, 0), 0) equals the
    sign-class minmod of (a, b/4, c) evaluated in FP32 (the two-argument
    minmod is the median of its arguments and zero), on edge and random
    triples, up to the sign of a zero result."""
    f = np.float32

    def med3(p, q, r):
        return sorted([p, q, r])[1]

    def class_form(x, y, z):
        a, b, c = f(y) - f(x), f(z) - f(x), f(z) - f(y)
        m = min(min(abs(a), abs(c)), abs(b) * f(0.25))
        na, nc = a < 0, c < 0
        return (-m if na else m) if na == nc else f(0)
    vals = [f(v) for v in _edge_values() if abs(v) > 1e-30 or v == 0]
    rng = np.random.default_rng(5)
    triples = [(x, y, z) for x in vals for y in vals for z in vals]
    triples += [tuple(f(v) for v in rng.standard_normal(3) * 10.0 ** rng.integers(-3, 3))
                for _ in range(20000)]
    for x, y, z in triples:
        a, b, c = f(y) - f(x), f(z) - f(x), f(z) - f(y)
        got = med3(a, med3(b * f(0.25), c, f(0)), f(0))
        assert got == class_form(x, y, z), (x, y, z)
