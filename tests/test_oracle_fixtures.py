"""The oracle against the reference's native-C golden records:
  * Hubbard chain step-1 SR operands (hop path, CoulombIntra, Gutzwiller,
    Jastrow, 8-point spin projection), tolerance 1e-14 abs + 1e-13 rel
    (crates/mvmc-core/tests/c_sr_operands_358.rs scale);
  * Heisenberg chain multi-rank state (exchange path, local spins, Hund,
    CoulombInter, Exchange): every saved EleIdx, the full SFMT state and the
    counters exact; operands 1e-13 + 1e-12 |b| (mpi_matrix_179/README.md)."""
import json
import os
import struct

import numpy as np
import pytest

import oracle_py as O

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def hexd(s):
    return struct.unpack(">d", bytes.fromhex(s))[0]


def close(a, b, abs_tol, rel_tol):
    a, b = np.asarray(a), np.asarray(b)
    return np.all(np.abs(a - b) <= abs_tol + rel_tol * np.abs(b))


@pytest.mark.parametrize("cr", [0, 1])
def test_hubbard_chain_step1_operands(cr):
    O.set_math(cr)
    try:
        ref = {k: np.array([hexd(x) for x in v]) for k, v in
               json.load(open(os.path.join(GOLDEN, "hubbard_cg_store0.json"))).items()}
        run = O.Run(os.path.join(GOLDEN, "inputs", "hubbard_chain", "namelist.def"), walkers=1)
        run.step(do_sr=False)
        e = run.energy()
        oo, ho = run.operands()
        size = run.size
        assert e[0] == 100.0
        assert e[1] == ref["energy"][0]
        assert close(ho, ref["ho"], 1e-14, 1e-13)
        assert close(oo[:2 * size], ref["oo"], 1e-14, 1e-13)
        assert close(run.o_last(0), ref["o"], 1e-14, 1e-13)
    finally:
        O.set_math(0)


@pytest.mark.parametrize("cell", ["real_cg_r2w1_s7", "real_cg_r4w1_s7"])
def test_heisenberg_multiwalker_state_exact(cell):
    fx = json.load(open(os.path.join(GOLDEN, f"mpi179_{cell}.json")))
    W = len(fx["ranks"])
    run = O.Run(os.path.join(GOLDEN, "inputs", "heisenberg_chain", "namelist.def"), walkers=W)
    run.step(do_sr=False)
    tot = np.zeros(10, dtype=np.int64)
    for w in range(W):
        st = run.walker(w)
        rec = fx["ranks"][w]
        for s, vals in rec["eleidx"].items():
            assert list(st["ele_idx"][int(s)]) == vals, (w, s)
        assert st["rng_idx"] == rec["rng_idx"]
        assert list(st["rng"]) == rec["rng"]
        tot += st["counter"]
        if w > 0:
            assert list(st["counter"][:6]) == rec["counter"]
    # rank 0 prints the ReduceCounter sum
    assert list(tot[:6]) == fx["ranks"][0]["counter"]
    ops = fx["operands"]
    e = run.energy()
    assert abs(e[1] - ops["energy"][0]) <= 1e-13 + 1e-12 * abs(ops["energy"][0])
    oo, ho = run.operands()
    size = run.size
    assert close(ho, ops["ho"][0::2], 1e-13, 1e-12)
    assert close(oo[:2 * size], ops["oo"][0::2], 1e-13, 1e-12)
    # rank 0's last-sample derivative (not reduced by C)
    assert close(run.o_last(0), ops["o"][0::2], 1e-13, 1e-12)


def _parse_cg(path):
    lines = [l for l in open(path).read().split("\n") if not l.startswith("#")]
    it = iter(lines)
    n, samples, cplx = map(int, next(it).split())
    p = lambda l: np.array([hexd(v) for v in l.split()])
    mean, diag, real, imag, g, z = p(next(it)), p(next(it)), p(next(it)), p(next(it)), p(next(it)), p(next(it))
    limits = []
    for line in it:
        if not line.strip():
            continue
        w = line.split(" ", 2)
        limits.append((int(w[0]), int(w[1]), p(w[2]), p(next(it)), p(next(it))))
    return n, samples, mean, diag, real, g, z, limits


@pytest.mark.parametrize("name,diag_scale", [("sr_cg_real.txt", False), ("sr_cg_real_diag_scale.txt", True)])
def test_sr_cg_iterates_match_c(name, diag_scale):
    """C 1.4.0 StochasticOptCG iterates (crates/mvmc-core/tests/sr_cg.rs:296-520
    policy): exact status at every limit, backward residual at every limit,
    forward iterates through limit 5 (the GEMV order is OpenBLAS there)."""
    n, samples, mean, diag, O_, g, z_ref, limits = _parse_cg(os.path.join(GOLDEN, name))
    eps = np.finfo(float).eps
    z = O.cg_apply(mean, diag, O_, g, 1.0 / samples, 1e-5)
    budget = 8.0 * (n + samples) * eps
    assert close(z, z_ref, budget, budget)
    Om = O_.reshape(samples, n)
    cov = Om.T @ Om / samples - np.outer(mean, mean) + np.diag(1e-5 * diag)
    for limit, status, x_ref, r_ref, d_ref in limits:
        st, x, r, d = O.cg_solve(mean, diag, O_, g, 1.0 / samples, 1e-5, 0.0, limit, diag_scale)
        assert st == status, (limit, st, status)
        for row in range(n):
            ax = cov[row] @ x
            scale = abs(g[row]) + np.abs(cov[row] * x).sum()
            assert abs(r[row] - (g[row] - ax)) <= 16.0 * (n + samples) * limit * eps * scale * 4
        if limit <= 5:
            b = 8.0 * (n + samples) * limit * eps
            for got, want in ((x, x_ref), (r, r_ref), (d, d_ref)):
                assert close(got, want, b, b), limit


def _parse_candidate_cases(path):
    cases, cur = [], None
    for line in open(path):
        parts = line.split()
        if not parts:
            continue
        if parts[0] == "case":
            cur = {"name": parts[1], "kind": parts[2]}
        elif parts[0] == "end":
            cases.append(cur)
            cur = None
        elif cur is not None:
            cur[parts[0]] = [int(x) for x in parts[1:]]
    return cases


def test_hopping_candidate_retry_cap_matches_c_tapes():
    """makeCandidate_hopping on the C retry-cap tapes
    (tests/fixtures/candidate/hopping_retry_cap_559.txt, C vmcmake.c verbatim on
    raw words): source and destination draws, the Nsite^2 cap tested after each
    destination draw, the source loop's itinerant-electron check, and the words
    consumed before and after the Metropolis draw."""
    cases = _parse_candidate_cases(os.path.join(GOLDEN, "candidate_hopping_retry_cap.txt"))
    hop = [c for c in cases if c["kind"] == "hopping"]
    assert len(hop) >= 7
    for c in hop:
        ns, ne = c["nsite"][0], c["nelec"][0]
        got = O.hop_candidate_tape(ns, ne, c["loc_spn"], c["ele_idx"], c["ele_cfg"], c["words"])
        assert got == c["result"], (c["name"], got, c["result"])


def zvo_row_text(e, e2, var, sz, sz2):
    """The zvo_out row layout (crates/mvmc-core/src/io.rs:166-179): C "% .18e"
    with two spaces before E^2 and "%.18e" for the Sz columns."""
    f = lambda x: "% .18e" % x
    return f"{f(e)} {f(0.0)}  {f(e2)} {f(var)} {'%.18e' % sz} {'%.18e' % sz2}"


def test_zvo_row_layout_matches_c_bytes():
    """Every row of the C 1.4.0 zvo_out.dat fixture re-rendered from its own
    values reproduces the C bytes (pins the row layout the library writes)."""
    for line in open(os.path.join(GOLDEN, "zvo_out_c_rows.dat")):
        if not line.strip():
            continue
        v = [float(x) for x in line.split()]
        assert zvo_row_text(v[0], v[2], v[3], v[4], v[5]) == line.rstrip("\n")


def test_fsz_candidates_match_c_tapes():
    """The FSZ siblings on the same C tapes: makeCandidate_hopping_fsz
    (target spin drawn, 2Sz = -1), makeCandidate_hopping_csz (spin kept) and
    makeCandidate_LocalSpinFlip_conduction, at their final and last permitted
    draws (candidate.rs:387-566).  The FSZ sampler that will use them is not
    built; the loops are pinned ahead of it."""
    cases = _parse_candidate_cases(os.path.join(GOLDEN, "candidate_hopping_retry_cap.txt"))
    kinds = {"hopping_fsz": 0, "hopping_csz": 1, "conduction": 2}
    fsz = [c for c in cases if c["kind"] in kinds]
    assert len(fsz) == 6
    for c in fsz:
        got = O.fsz_candidate_tape(kinds[c["kind"]], c["nsite"][0], c["loc_spn"], c["ele_idx"], c["ele_cfg"],
                                   c["ele_spn"], c["words"])
        assert got == c["result"], (c["name"], got, c["result"])
