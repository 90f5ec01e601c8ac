"""Pins of the CPU oracle's building blocks against the reference's golden
vectors and against what the mathematics fixes."""
import itertools
import json
import math
import os
import struct
from decimal import Decimal, getcontext

import numpy as np
import pytest

import oracle_py as O

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


# ---------------------------------------------------------------- SFMT ----
def test_sfmt_first_words_match_c_library():
    g = json.load(open(os.path.join(GOLDEN, "sfmt_golden.json")))
    for seed, words in g["first8_u32"].items():
        assert list(O.sfmt_words(int(seed), 8)) == words, seed


def test_sfmt_refill_boundaries_match_c_library():
    g = json.load(open(os.path.join(GOLDEN, "sfmt_golden.json")))
    words = O.sfmt_words(g["boundary_seed"], 2000)
    for idx, want in g["boundary"]:
        assert int(words[idx]) == want, idx
    w = O.sfmt_words(g["first8_u64_seed"], 16)
    u64 = [int(w[2 * i]) | (int(w[2 * i + 1]) << 32) for i in range(8)]
    assert u64 == g["first8_u64"]


# ---------------------------------------------------------- Pfaffian ------
def _random_skew(rng, n):
    a = rng.standard_normal((n, n))
    return a - a.T


def _pf_bruteforce(A):
    """Pfaffian by expansion along the first row (exact combinatorial definition)."""
    n = A.shape[0]
    if n == 0:
        return 1.0
    total = 0.0
    for j in range(1, n):
        idx = [k for k in range(n) if k not in (0, j)]
        sub = A[np.ix_(idx, idx)]
        total += (-1) ** (j + 1) * A[0, j] * _pf_bruteforce(sub)
    return total


def _col_major_planes(mats):
    # plane stored column-major: element (r, c) at r + c*n -> array [c][r] = A.T
    return np.stack([m.T.copy() for m in mats])


def test_pfaffian_closed_forms():
    a = np.array([[0.0, 1.7], [-1.7, 0.0]])
    pf, inv, st = O.pfinv(_col_major_planes([a]))
    assert st[0] == 0 and pf[0] == 1.7
    rng = np.random.default_rng(4)
    b = _random_skew(rng, 4)
    want = b[0, 1] * b[2, 3] - b[0, 2] * b[1, 3] + b[0, 3] * b[1, 2]
    pf, inv, st = O.pfinv(_col_major_planes([b]))
    assert abs(pf[0] - want) <= 16 * np.finfo(float).eps * (abs(want) + 1)


@pytest.mark.parametrize("n", [2, 4, 6, 8, 10])
def test_pfaffian_matches_bruteforce(n):
    rng = np.random.default_rng(100 + n)
    for _ in range(4):
        a = _random_skew(rng, n)
        want = _pf_bruteforce(a)
        pf, inv, st = O.pfinv(_col_major_planes([a]))
        assert st[0] == 0
        assert abs(pf[0] - want) <= 4 * n * np.finfo(float).eps * max(1.0, abs(want)) * 10


def test_pfaffian_sign_flip_and_transpose():
    rng = np.random.default_rng(7)
    a = _random_skew(rng, 8)
    pf1 = O.pfinv(_col_major_planes([a]))[0][0]
    pf2 = O.pfinv(_col_major_planes([-a]))[0][0]  # Pf(-A) = (-1)^(n/2) Pf(A)
    assert abs(pf1 - pf2) <= 1e-13 * abs(pf1)
    p = np.arange(8)
    p[[0, 1]] = p[[1, 0]]
    pf3 = O.pfinv(_col_major_planes([a[np.ix_(p, p)]]))[0][0]  # one transposition flips the sign
    assert abs(pf1 + pf3) <= 1e-13 * abs(pf1)


def test_pfapack_fixture_inputs_identities():
    """Reference inputs (tests/fixtures/pfapack/real_n*.orig.txt): Pf^2 = det (LU,
    independent) and the inverse residual, the identities of
    crates/mvmc-gpu/src/testkit.rs:167-189 and crates/pfapack/tests/fixed_matrix_identities.rs."""
    data = np.load(os.path.join(GOLDEN, "pfapack_real_inputs.npz"))
    for name in data.files:
        A = data[name]
        n = A.shape[0]
        pf, inv, st = O.pfinv(_col_major_planes([A]))
        assert st[0] == 0, name
        sign, logdet = np.linalg.slogdet(A)
        assert sign > 0
        assert abs(2 * math.log(abs(pf[0])) - logdet) <= n * 1e-13 + 1e-12 * abs(logdet), name
        X = inv[0].T  # back to row-major
        res = np.linalg.norm(A @ X - np.eye(n)) / (np.linalg.norm(A) * np.linalg.norm(X) * n)
        assert res <= 64 * np.finfo(float).eps, (name, res)
        # inverse of a skew matrix is skew to rounding
        assert np.linalg.norm(X + X.T) <= 1e-12 * np.linalg.norm(X)


def test_pfinv_status_codes():
    z = np.zeros((4, 4))
    pf, inv, st = O.pfinv(_col_major_planes([z]))
    assert st[0] > 0 and pf[0] == 0.0 and not inv.any()


def test_two_hop_bilinear_matches_c_bits():
    d = np.load(os.path.join(GOLDEN, "two_hop_bilinear.npz"))
    sizes, values, bits = d["sizes"], d["values"], d["bits"]
    off = 0
    import ctypes
    L = O.lib()
    L.oracle_two_hop_bilinear.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p]
    assert len(sizes) >= 100
    for n, b in zip(sizes, bits):
        n = int(n)
        block = values[off:off + n * n + 2 * n]
        off += n * n + 2 * n
        m = np.ascontiguousarray(block[:n * n])
        a = np.ascontiguousarray(block[n * n:n * n + n])
        bb = np.ascontiguousarray(block[n * n + n:])
        got = L.oracle_two_hop_bilinear(O.ptr(m), n, O.ptr(a), O.ptr(bb))
        assert struct.unpack("<Q", struct.pack("<d", got))[0] == int(b)


# ---------------------------------------------------------- exp / log -----
def test_correctly_rounded_exp_log_against_high_precision():
    getcontext().prec = 60
    rng = np.random.default_rng(11)
    L = O.lib()
    for x in np.concatenate([rng.uniform(-60, 8, 300), rng.uniform(-1e-3, 1e-3, 50)]):
        x = float(x)
        got = L.oracle_cr_exp(x)
        exact = Decimal(x).exp()
        lo, hi = np.nextafter(got, -np.inf), np.nextafter(got, np.inf)
        assert abs(Decimal(got) - exact) <= abs(Decimal(lo) - exact) and abs(Decimal(got) - exact) <= abs(Decimal(hi) - exact)
    for x in np.exp(rng.uniform(-300, 300, 300)):
        x = float(x)
        got = L.oracle_cr_log(x)
        exact = Decimal(x).ln()
        lo, hi = np.nextafter(got, -np.inf), np.nextafter(got, np.inf)
        assert abs(Decimal(got) - exact) <= abs(Decimal(lo) - exact) and abs(Decimal(got) - exact) <= abs(Decimal(hi) - exact)


def test_fast_paths_equal_accurate_paths():
    """The table-driven fast paths return the accurate double-double results
    on 2 million inputs (exp over the whole normal range and near 0, log over
    ~600 binades and near 1)."""
    rng = np.random.default_rng(17)
    xe = np.concatenate([rng.uniform(-745.0, 709.7, 400000), rng.uniform(-2, 2, 300000),
                         rng.uniform(-1e-3, 1e-3, 300000)])
    assert np.array_equal(O.cr_eval(0, xe), O.cr_eval(1, xe))
    xl = np.concatenate([np.exp(rng.uniform(-700, 700, 500000)), 1.0 + rng.uniform(-1e-4, 1e-4, 250000),
                         rng.uniform(0.5, 2.0, 250000)])
    assert np.array_equal(O.cr_eval(2, xl), O.cr_eval(3, xl))


def test_correctly_rounded_math_agrees_with_libm_nearly_always():
    """glibc exp/log are not correctly rounded in every case; the device uses
    the correctly rounded values.  Bound the disagreement rate (1 ulp), which
    bounds how often a Metropolis decision can differ from a glibc build."""
    rng = np.random.default_rng(5)
    L = O.lib()
    xs = rng.uniform(-40, 5, 200000)
    mism = sum(1 for x in xs if L.oracle_cr_exp(float(x)) != math.exp(float(x)))
    assert mism / xs.size < 2e-3
    ys = np.exp(rng.uniform(-200, 200, 200000))
    mism_l = sum(1 for y in ys if L.oracle_cr_log(float(y)) != math.log(float(y)))
    assert mism_l / ys.size < 1e-4


def test_metropolis_closed_forms():
    """metropolis_weight / metropolis_decision (metropolis.rs:42-83,92-128):
    w = exp(2((dproj + 0) + lip_new - lip_old)); a non-finite weight becomes
    -1; accept iff w > real2 with real2(word) = word 2^-32, so
    real2(u32::MAX) = 1 - 2^-32 still accepts w = 1."""
    acc, w = O.metropolis_word(0.1, 0.2, 0.05, 0)
    assert w == math.exp(2.0 * (((0.1 + 0.0) + 0.2) - 0.05)) and acc
    acc, w = O.metropolis_word(0.0, 0.3, 0.3, 0xFFFFFFFF)
    assert w == 1.0 and acc                      # 1 > 1 - 2^-32
    acc, w = O.metropolis_word(0.5 * math.log(0.5), 0.0, 0.0, 1 << 31)
    assert w == math.exp(math.log(0.5)) and acc == (w > 0.5)
    acc, w = O.metropolis_word(1000.0, 0.0, 0.0, 0)
    assert w == -1.0 and not acc                 # overflow: weight -1, rejected even at u = 0
    acc, w = O.metropolis_word(0.0, float("nan"), 0.0, 0)
    assert w == -1.0 and not acc
    acc, w = O.metropolis_word(0.0, -1000.0, 0.0, 0)
    assert w == 0.0 and not acc                  # underflow to 0 is finite: 0 > 0 fails


def _read_def_rows(path, ncols):
    rows = []
    for line in open(path):
        parts = line.split()
        if len(parts) != ncols:
            continue
        try:
            rows.append([float(x) for x in parts])
        except ValueError:
            continue
    return rows


def test_local_energy_matches_brute_force_overlaps():
    """Local energy of the 6-site Hubbard chain (U = 4, 24 transfer terms, 8 QP
    planes) against brute force: E_L(x) = U sum n_up n_dn - sum_k t_k
    psi(x'_k) / psi(x), every psi from full Pfaffians of the moved
    configuration (the electron keeps its label) times exp(sum_p v_p n_p(x)),
    with the Green-ratio formula of the oracle nowhere in the check
    (observables.rs:483-552,1638-1755)."""
    inputs = os.path.join(GOLDEN, "inputs", "hubbard_chain")
    run = O.Run(os.path.join(inputs, "namelist.def"), walkers=1)
    proj, _ = run.params()
    trans = [(int(a), int(b), int(c), int(d), t) for a, b, c, d, t, _ in
             _read_def_rows(os.path.join(inputs, "trans.def"), 6)]
    coul = {int(r): u for r, u in _read_def_rows(os.path.join(inputs, "coulombintra.def"), 2)}
    assert len(trans) == 24 and len(coul) == 6
    ns, ne = run.nsite, run.ne
    rng = np.random.default_rng(8)

    def psi(idx):
        st, ip, cnt = run.overlap_idx(idx)
        assert st == 0
        return ip * math.exp(float(np.dot(proj, cnt)))

    checked = 0
    for _ in range(25):
        idx = np.concatenate([rng.choice(ns, ne, replace=False), rng.choice(ns, ne, replace=False)]).astype(np.int32)
        e_fast, _, ok = run.measure_idx(idx)
        if not ok:
            continue
        occ = np.zeros((2, ns), dtype=int)
        for mi, r in enumerate(idx):
            occ[mi // ne, r] = 1
        e = sum(u * occ[0, r] * occ[1, r] for r, u in coul.items())
        p0 = psi(idx)
        for ri, s1, rj, s2, t in trans:
            if s1 == s2 and ri == rj:
                g = occ[s2, rj]
            elif s1 != s2 or occ[s1, ri] == 1 or occ[s2, rj] == 0:
                g = 0.0
            else:
                moved = idx.copy()
                mi = [k for k in range(s2 * ne, (s2 + 1) * ne) if idx[k] == rj][0]
                moved[mi] = ri
                g = psi(moved) / p0
            e -= t * g
        assert abs(e_fast - e) <= 1e-11 * max(1.0, abs(e)), (idx, e_fast, e)
        checked += 1
    assert checked >= 20
