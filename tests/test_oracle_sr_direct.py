"""Oracle pins of the direct SR path (NSRCG 0) and of the SrStages the C
order defines besides the CG step: S/g assembly, Cholesky solve, the
complex Gram and the composite S/g of the validation harness.

Pins: the C 1.4.0 dposv solutions of the negative-DSROptStepDt records
(tests/fixtures/negative_stepdt/c_sr.txt), the native-C full <OO>, <HO> and
rank states of the direct-SR Heisenberg cells (mpi_matrix_179 real-d0 /
real-d1), and closed forms: exact sums on integer / dyadic inputs, exact
Hermitian symmetry of the Gram, LAPACK info of a matrix whose third leading
minor is not positive definite."""
import json
import os
import shutil
from fractions import Fraction

import numpy as np
import pytest

import oracle_py as O

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
HEIS_DIR = os.path.join(GOLDEN, "inputs", "heisenberg_chain")

# The fixed operands of the negative-DSROptStepDt records (C c_toolbox
# program; the (n+1) x (n+1) layout of crates/mvmc-core/src/sr_negative_stepdt_tests.rs:10-17):
# <OO> with the constant component first, <HO>.
OO = np.array([[1.0, 0.3, -0.2, 0.1],
               [0.3, 2.09, 0.44, 0.13],
               [-0.2, 0.44, 3.04, 0.18],
               [0.1, 0.13, 0.18, 1.51]])
HO = np.array([-1.5, 0.7, -0.4, 0.25])


def close(a, b, atol, rtol):
    a, b = np.asarray(a), np.asarray(b)
    return bool(np.all(np.abs(a - b) <= atol + rtol * np.abs(b)))


def test_direct_solve_matches_c_dposv_records():
    """assemble_s_g + cholesky_solve on the C operands: every record's
    solution within 1e-15 + 1e-13 |x| (the reference's own bar: a small SPD
    system solved by two LAPACK builds), for positive, negative (normalised)
    and shifted steps."""
    cases = json.load(open(os.path.join(GOLDEN, "negative_stepdt_c_sr.json")))
    assert len(cases) == 5
    for c in cases:
        dt = abs(float(c["raw_dt"]))
        assert dt == c["dt"], c["name"]
        assert c["srflag"] == (1 if float(c["raw_dt"]) < 0 else 0)
        S, g = O.sr_assemble_s_g(OO.reshape(-1), HO, [0, 1, 2], 4, 1, float(c["sta_del"]), dt)
        # the entries as the reference forms them
        for si in range(3):
            for sj in range(3):
                v = OO[si + 1, sj + 1] - OO[0, si + 1] * OO[0, sj + 1]
                if si == sj:
                    v *= 1.0 + float(c["sta_del"])
                assert S[si, sj] == v
        info, x = O.cholesky_solve(S, g)
        assert info == c["info"] == 0
        assert close(x, c["solution"], 1e-15, 1e-13), (c["name"], x, c["solution"])


def test_negative_step_dt_is_normalised_by_the_reader(tmp_path):
    """readdef.c:959-975: a negative DSROptStepDt is negated (diagonalisation
    mode) and rejected together with SR-CG."""
    for nsrcg, ok in ((0, True), (1, False)):
        d = tmp_path / f"h{nsrcg}"
        shutil.copytree(HEIS_DIR, d)
        mp = d / "modpara.def"
        txt = mp.read_text()
        lines = []
        for line in txt.splitlines():
            k = line.split()[0] if line.split() else ""
            if k == "DSROptStepDt":
                line = "DSROptStepDt -0.02"
            elif k == "NSRCG":
                line = f"NSRCG {nsrcg}"
            lines.append(line)
        mp.write_text("\n".join(lines) + "\n")
        if ok:
            run = O.Run(str(d / "namelist.def"), walkers=1)
            assert run.step(do_sr=True) == 0
        else:
            with pytest.raises(RuntimeError, match="negative DSROptStepDt"):
                O.Run(str(d / "namelist.def"), walkers=1)


def _heis_direct(tmp_path, nstore):
    d = tmp_path / f"heis_d{nstore}"
    shutil.copytree(HEIS_DIR, d)
    mp = d / "modpara.def"
    out = []
    for line in mp.read_text().splitlines():
        k = line.split()[0] if line.split() else ""
        if k == "NSRCG":
            line = "NSRCG 0"
        elif k == "NStore":
            line = f"NStore {nstore}"
        out.append(line)
    mp.write_text("\n".join(out) + "\n")
    return str(d / "namelist.def")


@pytest.mark.parametrize("cell,nstore", [("real_d0_r2w1_s7", 0), ("real_d0_r4w1_s7", 0), ("real_d1_r2w1_s7", 1)])
def test_direct_sr_operands_match_c_records(tmp_path, cell, nstore):
    """Heisenberg chain, W = 2 or 4 ranks, 7 samples, direct SR: every rank's
    counters, EleIdx and SFMT state exact; the full <OO> (NStore 0: per-sample
    rank-1 sums; NStore 1: the Gram of the stored rows), <HO> and the energy
    within the records' 1e-13 + 1e-12 |b|."""
    fx = json.load(open(os.path.join(GOLDEN, f"mpi179_{cell}.json")))
    W = len(fx["ranks"])
    run = O.Run(_heis_direct(tmp_path, nstore), walkers=W)
    run.step(do_sr=False)
    for w in range(W):
        st = run.walker(w)
        rec = fx["ranks"][w]
        for s, vals in rec["eleidx"].items():
            assert list(st["ele_idx"][int(s)]) == vals, (w, s)
        assert st["rng_idx"] == rec["rng_idx"] and list(st["rng"]) == rec["rng"]
    ops = fx["operands"]
    size = run.size
    e = run.energy()
    assert abs(e[1] - ops["energy"][0]) <= 1e-13 + 1e-12 * abs(ops["energy"][0])
    oo, ho = run.operands()
    assert len(ops["oo"]) == 2 * size * size
    assert close(ho, ops["ho"][0::2], 1e-13, 1e-12)
    assert close(oo[:size * size], ops["oo"][0::2], 1e-13, 1e-12)


def test_direct_sr_step_solves_the_assembled_system(tmp_path):
    """A full direct-SR step of the oracle: the parameters that move are the
    active set, and their change x solves S x = g formed independently (numpy)
    from the oracle's averaged <OO>, <HO> (S = <OO> - <O><O>^T with the
    diagonal times 1 + DSROptStaDel, g = -2 dt (<HO> - <H><O>)): residual
    within cond(S) eps."""
    nl = _heis_direct(tmp_path, 0)
    mp = open(os.path.join(os.path.dirname(nl), "modpara.def")).read().split()
    sta_del = float(mp[mp.index("DSROptStaDel") + 1])
    dt = float(mp[mp.index("DSROptStepDt") + 1])
    run = O.Run(nl, walkers=2)
    run.step(do_sr=False)
    oo, ho = run.operands()
    size = run.size
    p0 = np.concatenate(run.params())
    run2 = O.Run(nl, walkers=2)
    assert run2.step(do_sr=True) == 0
    p1 = np.concatenate(run2.params())
    act = np.nonzero(p1 != p0)[0]
    assert len(act) > 0
    OOm = oo[:size * size].reshape(size, size)
    S = OOm[1:, 1:][np.ix_(act, act)] - np.outer(OOm[0, 1 + act], OOm[0, 1 + act])
    S[np.diag_indices_from(S)] *= 1.0 + sta_del
    g = -2.0 * dt * (ho[1 + act] - ho[0] * OOm[0, 1 + act])
    x = (p1 - p0)[act]
    res = np.abs(S @ x - g).max()
    assert res <= 1e-14 * np.linalg.cond(S) * np.abs(g).max() + 1e-16, res


def test_cholesky_info_and_residual():
    rng = np.random.default_rng(4)
    A = rng.standard_normal((40, 40))
    S = A @ A.T + 40 * np.eye(40)
    b = rng.standard_normal(40)
    info, x = O.cholesky_solve(S, b)
    assert info == 0
    assert np.abs(S @ x - b).max() <= 1e-12 * np.abs(b).max() * np.linalg.cond(S)
    # third leading minor not positive definite -> info 3
    M = np.eye(5)
    M[0, 2] = M[2, 0] = 2.0
    info, _ = O.cholesky_solve(M, np.ones(5))
    assert info == 3


def test_complex_gram_exact_on_integers_and_hermitian():
    rng = np.random.default_rng(6)
    st = rng.integers(-9, 10, (37, 11)) + 1j * rng.integers(-9, 10, (37, 11))
    G = O.sr_gram_complex(st)
    ref = np.einsum("si,sj->ij", st.astype(np.complex128), np.conj(st.astype(np.complex128)))
    assert np.array_equal(G, ref)   # integer sums: exact in any order
    st = rng.standard_normal((64, 23)) + 1j * rng.standard_normal((64, 23))
    G = O.sr_gram_complex(st)
    assert np.array_equal(G, np.conj(G.T))          # exact: same rounded products
    assert np.all(np.diag(G).imag == 0) and np.all(np.diag(G).real > 0)
    ref = st.T @ np.conj(st)
    assert np.abs(G - ref).max() <= 64 * 1e-16 * np.abs(st).max() ** 2 * 64


def test_sr_s_g_exact_on_dyadic_inputs():
    """Integer O and E with weights 1/8: every product and sum is exact, so
    S and g equal the rational weighted covariances."""
    rng = np.random.default_rng(8)
    ns, npar = 8, 5
    o = rng.integers(-6, 7, (ns, npar)).astype(float)
    e = rng.integers(-6, 7, ns).astype(float)
    w = np.full(ns, 1.0 / ns)
    S, g = O.sr_s_g(o, e, w)
    fo = [[Fraction(int(v)) for v in row] for row in o]
    fe = [Fraction(int(v)) for v in e]
    wf = Fraction(1, ns)
    mean = [sum(wf * fo[k][p] for k in range(ns)) for p in range(npar)]
    em = sum(wf * fe[k] for k in range(ns))
    for p in range(npar):
        for q in range(npar):
            ref = sum(wf * (fo[k][p] - mean[p]) * (fo[k][q] - mean[q]) for k in range(ns))
            assert S[p, q] == float(ref)
        assert g[p] == float(sum(wf * (fo[k][p] - mean[p]) * (fe[k] - em) for k in range(ns)))
    # random inputs against numpy
    o = rng.standard_normal((50, 7))
    e = rng.standard_normal(50)
    w = rng.random(50)
    w /= w.sum()
    S, g = O.sr_s_g(o, e, w)
    d = o - w @ o
    de = e - w @ e
    assert np.abs(S - (d * w[:, None]).T @ d).max() <= 1e-13
    assert np.abs(g - (d * w[:, None]).T @ de).max() <= 1e-13
