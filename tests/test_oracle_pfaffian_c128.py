"""Pins of the oracle's complex Pfaffian / inverse (oracle/pfapack_c.h) and of
Julia's complex division and reciprocal it uses, by what the mathematics
fixes: closed forms, brute-force expansion, Pf^2 = det from an independent
LU (modulus and phase), inverse residuals, agreement of the two LTL orders and
of the complex path with the real one on real input (the checks of
crates/pfapack/tests/fixed_matrix_identities.rs and
crates/mvmc-core/src/pfaffian.rs:1806), on the reference's complex input
matrices (tests/fixtures/pfapack/complex_n*.orig.txt)."""
import math
import os
from fractions import Fraction

import numpy as np
import pytest

import oracle_py as O

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
EPS = np.finfo(float).eps


def _random_skew_c(rng, n):
    a = rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n))
    return a - a.T


def _pf_bruteforce(A):
    """Pfaffian by expansion along the first row (combinatorial definition)."""
    n = A.shape[0]
    if n == 0:
        return 1.0 + 0j
    total = 0j
    for j in range(1, n):
        idx = [k for k in range(n) if k not in (0, j)]
        total += (-1) ** (j + 1) * A[0, j] * _pf_bruteforce(A[np.ix_(idx, idx)])
    return total


def _planes(mats):
    # column-major planes: element (r, c) at r + c*n -> array [c][r] = A.T
    return np.stack([m.T.copy() for m in mats])


@pytest.mark.parametrize("order", [0, 1])
def test_complex_pfaffian_closed_forms(order):
    a = np.array([[0, 1.7 - 0.3j], [-1.7 + 0.3j, 0]])
    pf, inv, st = O.pfinv_c128(_planes([a]), order)
    assert st[0] == 0 and pf[0] == 1.7 - 0.3j
    # inverse of [[0, a], [-a, 0]] is [[0, -1/a], [1/a, 0]]
    assert abs(inv[0].T[0, 1] + 1 / a[0, 1]) <= 4 * EPS * abs(1 / a[0, 1])
    rng = np.random.default_rng(41)
    b = _random_skew_c(rng, 4)
    want = b[0, 1] * b[2, 3] - b[0, 2] * b[1, 3] + b[0, 3] * b[1, 2]
    pf, inv, st = O.pfinv_c128(_planes([b]), order)
    assert abs(pf[0] - want) <= 32 * EPS * (abs(want) + 1)


@pytest.mark.parametrize("order", [0, 1])
@pytest.mark.parametrize("n", [2, 4, 6, 8, 10])
def test_complex_pfaffian_matches_bruteforce(n, order):
    rng = np.random.default_rng(200 + n)
    for _ in range(3):
        a = _random_skew_c(rng, n)
        want = _pf_bruteforce(a)
        pf, inv, st = O.pfinv_c128(_planes([a]), order)
        assert st[0] == 0
        assert abs(pf[0] - want) <= 40 * n * EPS * max(1.0, abs(want))


def test_complex_fixture_inputs_identities():
    """Pf^2 = det (log-modulus and phase) and the inverse residual on the
    reference's complex inputs, for both LTL orders, which also agree with each
    other to 4 n eps relative."""
    data = np.load(os.path.join(GOLDEN, "pfapack_complex_inputs.npz"))
    assert len(data.files) >= 6
    for name in data.files:
        A = data[name]
        n = A.shape[0]
        assert np.array_equal(A, -A.T), name
        pfs = []
        for order in (0, 1):
            pf, inv, st = O.pfinv_c128(_planes([A]), order)
            assert st[0] == 0, name
            phase, logdet = np.linalg.slogdet(A)
            assert abs(2 * math.log(abs(pf[0])) - logdet) <= n * 1e-13 + 1e-12 * abs(logdet), name
            assert abs((pf[0] / abs(pf[0])) ** 2 - phase) <= n * 1e-13, name
            X = inv[0].T
            res = np.abs(A @ X - np.eye(n)).max() / (np.abs(A).max() * np.abs(X).max() * n)
            assert res <= 64 * EPS, (name, order, res)
            assert np.abs(X + X.T).max() <= 1e-12 * np.abs(X).max()
            pfs.append(pf[0])
        assert abs(pfs[0] - pfs[1]) <= 4 * n * EPS * abs(pfs[0]), name


def test_complex_path_on_real_input_matches_real_path():
    """Complex LTL on real input keeps zero imaginary parts and agrees with the
    real oracle (the complex update sums c + (x t1 - y t2), the real one
    (c + x t1) - y t2, so agreement is to rounding, not bitwise)."""
    data = np.load(os.path.join(GOLDEN, "pfapack_real_inputs.npz"))
    for name in data.files:
        A = data[name]
        n = A.shape[0]
        pf_r, inv_r, _ = O.pfinv(_planes([A]))
        for order in (0, 1):
            pf_c, inv_c, st = O.pfinv_c128(_planes([A.astype(np.complex128)]), order)
            assert st[0] == 0
            assert pf_c[0].imag == 0 and not inv_c.imag.any(), name
            assert abs(pf_c[0].real - pf_r[0]) <= 4 * n * EPS * abs(pf_r[0]), name
            assert np.abs(inv_c[0].real - inv_r[0]).max() <= 1e-10 * np.abs(inv_r[0]).max(), name


def test_complex_status_codes():
    z = np.zeros((2, 4, 4), dtype=np.complex128)
    z[1] = _random_skew_c(np.random.default_rng(5), 4).T
    z[1][0, 1] = complex(np.inf, 0)
    z[1][1, 0] = complex(-np.inf, 0)
    for order in (0, 1):
        pf, inv, st = O.pfinv_c128(z, order)
        assert st[0] > 0 and pf[0] == 0 and not inv[0].any()
        # non-finite Pfaffian: status -1, the computed value reported, zero inverse
        assert st[1] == -1 and not np.isfinite(pf[1]) and not inv[1].any()


def _exact_quotient(z, w):
    """z / w rounded once from exact rational arithmetic (None when it
    overflows the double range)."""
    a, b, c, d = (Fraction(float(v)) for v in (z.real, z.imag, w.real, w.imag))
    den = c * c + d * d
    re, im = (a * c + b * d) / den, (b * c - a * d) / den
    if max(abs(re), abs(im)) > Fraction(1, 2) * Fraction(np.finfo(float).max):
        return None
    return complex(float(re), float(im))


def _rel(x, y):
    """Normwise relative error |x - y| / |y| (Smith-type division is accurate
    relative to the quotient's modulus, not componentwise)."""
    return abs(complex(x) - y) / abs(y)


def test_julia_division_and_reciprocal():
    """Julia's division and reciprocal: within 4 eps (normwise) of the exactly
    rounded quotient on ordinary and on scaled (huge / tiny) operands, and the exact
    special values of their infinity branches."""
    rng = np.random.default_rng(11)
    w = rng.standard_normal(400) + 1j * rng.standard_normal(400)
    z = rng.standard_normal(400) + 1j * rng.standard_normal(400)
    scales = np.array([1.0, 1e300, 1e-300, 1e-310, 2.0 ** 600])
    w = np.concatenate([w * s for s in scales] + [np.array([3.0 + 0j, 0 + 2j, -5 + 0j])])
    z = np.concatenate([z * s for s in scales[::-1]] + [np.array([1 + 1j, 1 + 1j, 2 - 1j])])
    q = O.julia_complex(0, z, w)
    r = O.julia_complex(1, None, w)
    for zi, wi, qi, ri in zip(z, w, q, r):
        if np.isfinite(zi) and np.isfinite(wi):
            want = _exact_quotient(zi, wi)
            if want is not None and want != 0:
                assert _rel(qi, want) <= 4 * EPS, (zi, wi, qi, want)
            want = _exact_quotient(1 + 0j, wi)
            if want is not None:
                assert _rel(ri, want) <= 4 * EPS, (wi, ri, want)
    # special values: the real-denominator reciprocal keeps a signed zero
    # imaginary part, an infinite denominator gives signed zeros
    r = O.julia_complex(1, None, np.array([2.0 + 0j, complex(np.inf, 0), complex(-np.inf, 1.0)]))
    assert r[0] == 0.5 and math.copysign(1, r[0].imag) < 0
    assert r[1] == 0 and math.copysign(1, r[1].real) > 0 and math.copysign(1, r[1].imag) < 0
    assert r[2] == 0 and math.copysign(1, r[2].real) < 0
    q = O.julia_complex(0, np.array([1.0 + 1.0j]), np.array([complex(np.inf, 0)]))
    assert q[0] == 0
