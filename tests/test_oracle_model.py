"""The oracle's own Expert-mode reader and parameter tables (oracle/expert.cpp),
pinned without running product code:

* the Gutzwiller / Jastrow gauge gate against the C SetFlagShift records
  (tests/fixtures/optimization_flags/c_shift_flags.txt of the reference);
* the gauge sync by its invariants: after a shift the Gutzwiller + Jastrow
  values sum to zero and keep their differences, max |f| becomes 4, nothing
  else moves;
* the Gauss-Legendre points of the QP weights by their closed forms and by
  exactness on polynomials of degree <= 2n - 1;
* the Slater table by skew symmetry and, with one QP plane, by the orbital
  values read from the .def files in plain Python.
"""
import ctypes
import math
import os

import numpy as np
import pytest

import oracle_py as O

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
EPS = np.finfo(float).eps


def _ip(a):
    return ctypes.c_void_p(a.ctypes.data)


def gate(ng, nj, flags):
    f = np.ascontiguousarray(flags, dtype=np.int32)
    return O.lib().oracle_gj_shift_gate(ctypes.c_int(ng), ctypes.c_int(nj), _ip(f), ctypes.c_int(f.size))


def sync(ng, nj, flags, proj, slater):
    f = np.ascontiguousarray(flags, dtype=np.int32)
    p = np.array(proj, dtype=np.float64)
    s = np.array(slater, dtype=np.float64)
    O.lib().oracle_sync_parameters(ctypes.c_int(ng), ctypes.c_int(nj), _ip(f), ctypes.c_int(f.size), _ip(p),
                                   ctypes.c_int(s.size), _ip(s))
    return p, s


def shift_records():
    rows = [l.rstrip("\n") for l in open(os.path.join(GOLDEN, "c_shift_flags.txt")) if not l.startswith("#")]
    assert len(rows) == 86 * 3
    for k in range(0, len(rows), 3):
        head = rows[k].split()
        counts = [int(x) for x in head[1:]]
        flags = [int(x) for x in rows[k + 1].split()]
        gates = [int(x) for x in rows[k + 2].split()]
        yield head[0], counts, flags, gates


def test_gj_gate_matches_c_set_flag_shift():
    """The C gate of the Gutzwiller / Jastrow shift for every record: the G and
    J real component flags exactly 1 (-2, 0, 2, 3 in any real slot close it,
    imaginary slots do not count) and both families present.  The records'
    doublon-holon flags follow the G/J block and do not enter the G/J gate."""
    checked = 0
    for name, (ng, nj, _nd2, _nd4), flags, gates in shift_records():
        assert gate(ng, nj, flags) == gates[0], name
        checked += 1
    assert checked == 86


def test_sync_shift_only_where_the_c_gate_opens():
    rng = np.random.default_rng(3)
    opened = closed = 0
    for name, (ng, nj, _d2, _d4), flags, gates in shift_records():
        proj = rng.uniform(-0.5, 0.5, ng + nj)
        slater = rng.uniform(-1, 1, 5)
        p, s = sync(ng, nj, flags, proj, slater)
        if gates[0] == 1:
            opened += 1
            assert abs(p.sum()) <= 8 * EPS * (ng + nj) * np.abs(proj).max(), name
            # one common shift: differences to the first value kept to rounding
            assert np.allclose(p - p[0], proj - proj[0], rtol=0, atol=8 * EPS), name
            assert not np.array_equal(p, proj), name
        else:
            closed += 1
            assert np.array_equal(p, proj), name
    assert opened > 0 and closed > 0


def test_sync_rescales_slater_to_amplitude_four():
    rng = np.random.default_rng(4)
    for n in (1, 7, 256):
        s0 = rng.uniform(-0.3, 0.3, n)
        _, s = sync(0, 0, [], [], s0)
        assert abs(np.abs(s).max() - 4.0) <= 4 * EPS
        # one common factor 4 / max|f|: ratios unchanged
        k = int(np.argmax(np.abs(s0)))
        assert np.allclose(s / s[k], s0 / s0[k], rtol=4 * EPS, atol=0)
    _, s = sync(0, 0, [], [], np.zeros(4))
    assert np.array_equal(s, np.zeros(4))      # all-zero orbitals stay zero
    # projection values untouched without both families
    p, _ = sync(2, 0, [1, 0, 1, 0], [0.3, -0.1], [1.0])
    assert np.array_equal(p, [0.3, -0.1])


def gauss_leg(n, a, b):
    x = np.zeros(n)
    w = np.zeros(n)
    O.lib().oracle_gauss_leg(ctypes.c_int(n), ctypes.c_double(a), ctypes.c_double(b), _ip(x), _ip(w))
    return x, w


@pytest.mark.parametrize("n", [1, 2, 3, 5, 8, 12])
def test_gauss_legendre_exact_on_polynomials(n):
    x, w = gauss_leg(n, 0.0, math.pi)
    assert np.all(np.diff(x) > 0)
    assert np.allclose(x + x[::-1], math.pi, rtol=0, atol=8 * EPS)       # symmetric nodes
    for k in range(2 * n):
        exact = math.pi ** (k + 1) / (k + 1)
        assert abs(np.dot(w, x ** k) - exact) <= 1e-13 * exact, k


def test_gauss_legendre_two_points_closed_form():
    x, w = gauss_leg(2, 0.0, math.pi)
    d = math.pi / (2 * math.sqrt(3.0))
    assert np.allclose(x, [math.pi / 2 - d, math.pi / 2 + d], rtol=2 * EPS, atol=0)
    assert np.allclose(w, [math.pi / 2, math.pi / 2], rtol=2 * EPS, atol=0)


@pytest.mark.parametrize("nl,stot", [(1, 0), (2, 0), (8, 0), (4, 1)])
def test_spin_projection_weights_from_nodes(nl, stot):
    out = np.zeros(4 * nl)
    O.lib().oracle_spin_gl(ctypes.c_int(nl), ctypes.c_int(stot), _ip(out))
    cc, ss, cs, qw = out[:nl], out[nl:2 * nl], out[2 * nl:3 * nl], out[3 * nl:]
    if nl == 1:
        assert (cc[0], ss[0], cs[0], qw[0]) == (1.0, 0.0, 0.0, 1.0)
        return
    beta, wgl = gauss_leg(nl, 0.0, math.pi)
    legendre = np.polynomial.legendre.Legendre.basis(stot)
    assert np.allclose(cc, np.cos(beta / 2) ** 2, rtol=4 * EPS, atol=0)
    assert np.allclose(ss, np.sin(beta / 2) ** 2, rtol=4 * EPS, atol=0)
    assert np.allclose(cs, np.cos(beta / 2) * np.sin(beta / 2), rtol=4 * EPS, atol=0)
    assert np.allclose(cc + ss, 1.0, rtol=0, atol=4 * EPS)
    assert np.allclose(qw, 0.5 * np.sin(beta) * wgl * legendre(np.cos(beta)), rtol=1e-14, atol=1e-16)
    if stot == 0:
        # the S = 0 projector's weights integrate sin(b)/2 over [0, pi]: total 1
        assert abs(qw.sum() - 1.0) < (1e-3 if nl >= 8 else 0.2)


def _read_rows(path):
    lines = open(path).read().split("\n")
    return [l.split() for l in lines[5:] if l.split()]


def _slater_table(run):
    S = np.zeros(run.nqp * (2 * run.nsite) ** 2)
    O.lib().oracle_get_slater_table(run.h, _ip(S))
    return S.reshape(run.nqp, 2 * run.nsite, 2 * run.nsite)


@pytest.mark.parametrize("inputs", ["hubbard_chain", "heisenberg_chain"])
def test_slater_table_is_skew_symmetric(inputs):
    run = O.Run(os.path.join(GOLDEN, "inputs", inputs, "namelist.def"), walkers=1)
    run.step(do_sr=False)
    S = _slater_table(run)
    assert S.shape[0] == run.nqp
    assert np.array_equal(S, -S.transpose(0, 2, 1))
    assert np.abs(S).max() > 0


def test_slater_table_one_plane_equals_orbital_values(tmp_path):
    """One QP plane (NSPGaussLeg 1, identity translation): the up-down block
    S[ri][rj + Ns] is the orbital value f[idx(ri, rj)] times its sign, with the
    index table read here from orbitalidx.def by plain Python."""
    import lattice
    info = lattice.write_workload("chain16", str(tmp_path / "c16"))
    run = O.Run(info["namelist"], info["initial"], walkers=1)
    run.step(do_sr=False)
    proj, slater = run.params()
    S = _slater_table(run)
    ns = run.nsite
    rows = _read_rows(str(tmp_path / "c16" / "orbitalidx.def"))
    idx = np.zeros((ns, ns), dtype=int)
    for r in rows[:ns * ns]:
        idx[int(r[0]), int(r[1])] = int(r[2])
    assert run.nqp == 1
    assert np.array_equal(S[0][:ns, ns:], slater[idx])
    assert np.array_equal(S[0][ns:, :ns], -slater[idx].T)
    assert np.array_equal(S[0][:ns, :ns], np.zeros((ns, ns)))   # cos * sin = 0
    # the gauge: max |f| = 4
    assert abs(np.abs(slater).max() - 4.0) <= 4 * EPS
