"""GPU path (libmvmc_hip.so through its C ABI) on the benchmark configurations
against the oracle: the 16-site chain at its full 256 walkers over ten SR
steps, the exchange sampler on the 8x8 and 16x16 lattices, the in-place
batched Pfaffian, and the derivative form on the 8x8 lattice.  Bar:
bit-exact configurations, SFMT states, counters, zvo rows and parameters
(the oracle runs with the device's correctly rounded exp/log); derivatives
within the finite-difference tolerance stated in the test."""
import os
import sys

import numpy as np
import pytest

import oracle_py as O

pytestmark = pytest.mark.gpu


def log(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def M():
    # torch brings its HIP runtime up first; the library binds to it
    import torch
    torch.zeros(1, device="cuda")
    torch.cuda.synchronize()
    import mvmc_hip
    return mvmc_hip


def _walkers_equal(ens, run, W):
    for w in range(W):
        sg, so = ens.walker(w), run.walker(w)
        assert np.array_equal(sg["ele_idx"], so["ele_idx"]), w
        assert np.array_equal(sg["rng"], so["rng"]) and sg["rng_idx"] == so["rng_idx"], w
        assert list(sg["counter"]) == list(so["counter"][:4]), (w, sg["counter"], so["counter"][:4])


def test_chain16_full_walkers_ten_sr_steps(M, tmp_path):
    """The 16-site Hubbard chain (Nsize 16, U = 4, half filling, Gutzwiller +
    nearest-neighbour Jastrow + 2-site-translation orbitals, 1 QP plane) at
    its full 256 walkers, 16 sweeps per sample, ten SR-CG steps (shift 1e-3):
    after every step the zvo row, every parameter and every walker's saved
    configuration, SFMT state and counters equal the oracle's bit for bit."""
    import lattice
    info = lattice.write_workload("chain16", str(tmp_path / "c1"), sr_steps=10)
    W = info["walkers"]
    assert W == 256 and info["nsize"] == 16 and info["nqp"] == 1
    log(f"chain16 W={W}: 10 SR steps vs oracle")
    O.set_math(1)
    try:
        run = O.Run(info["namelist"], info["initial"], walkers=W, threads=8)
        ens = M.Ensemble(info["namelist"], info["initial"], walkers=W)
        for step in range(10):
            run.step(do_sr=True)
            row = ens.opt_step()
            ro = run.energies()[-1]
            log(f"  step {step}: E gpu {row[0]!r} oracle {ro[0]!r}")
            assert np.array_equal(row, ro), (step, row, ro)
            pg, sg = ens.params()
            po, so = run.params()
            assert np.array_equal(pg, po) and np.array_equal(sg, so), step
            _walkers_equal(ens, run, W)
        ens.close()
    finally:
        O.set_math(0)


@pytest.mark.parametrize("name,W,warmup", [("square8", 2, 4), ("square16", 1, 1)])
def test_exchange_path_benchmark_lattices(M, tmp_path, name, W, warmup):
    """The exchange sampler (NExUpdatePath 1: each proposal is a hop or, with
    probability 1/2, a spin exchange, candidate.rs:320-385) on the 8x8
    (Nsize 64) and 16x16 (Nsize 224) lattices, 8 QP planes: the two-electron
    ratio, the rank-2 two-electron update and the exchange LDS carve at
    n >= 64.  One SR-CG step: zvo row, parameters and every walker's state
    bit-exact against the oracle, with exchange moves proposed and accepted."""
    import lattice
    info = lattice.write_workload(name, str(tmp_path / name), sr_steps=1, warmup=warmup, ex_update_path=1)
    log(f"{name} W={W} NExUpdatePath 1 vs oracle")
    O.set_math(1)
    try:
        run = O.Run(info["namelist"], info["initial"], walkers=W)
        ens = M.Ensemble(info["namelist"], info["initial"], walkers=W)
        run.step(do_sr=True)
        row = ens.opt_step()
        ro = run.energies()[-1]
        assert np.array_equal(row, ro), (row, ro)
        _walkers_equal(ens, run, W)
        for w in range(W):
            c = ens.walker(w)["counter"]
            assert c[2] > 0 and c[3] > 0, ("no exchange proposed/accepted", w, c)
        pg, sg = ens.params()
        po, so = run.params()
        assert np.array_equal(pg, po) and np.array_equal(sg, so)
        ens.close()
    finally:
        O.set_math(0)


@pytest.mark.parametrize("ex_path", [0, 1])
def test_write_pass_accounting(M, tmp_path, ex_path):
    """The sampler writes a walker's planes back at every second accepted hop
    (the pending list of two is full) and when it flushes the list (before
    an exchange proposal, at the end of a block); a periodic rebuild drops
    the list.  Per segment of h accepted hops between flushes or rebuilds
    that makes floor((h - 1) / 2) list-full writes, so over the sample call
    hops/2 - segments <= writes - flushes <= hops/2, and every flush is a
    write.  These counts give the sampler's algorithmic bytes."""
    import lattice
    W, warmup = 32, 4
    info = lattice.write_workload("square8", str(tmp_path / "s8"), sr_steps=1, warmup=warmup, ex_update_path=ex_path)
    log(f"square8 W={W} NExUpdatePath {ex_path}: write-pass accounting")
    ens = M.Ensemble(info["namelist"], info["initial"], walkers=W)
    ens.opt_step()
    t = ens.timings()
    c = np.array([ens.walker(w)["counter"] for w in range(W)], dtype=np.int64)
    hops, ex_props = int(c[:, 1].sum()), int(c[:, 2].sum())
    writes, flushes = int(t["block_write_passes"]), int(t["block_flush_passes"])
    blocks = W * (warmup + 1)
    segments = blocks + int(t["rebuilt"]) + ex_props
    log(f"  hops {hops}, exchange proposals {ex_props}, writes {writes}, flushes {flushes}, rebuilt {t['rebuilt']}")
    ens.close()
    assert hops > 0 and 0 < flushes <= writes
    assert flushes <= blocks + ex_props
    assert hops / 2 - segments <= writes - flushes <= hops // 2


def _random_skew(rng, n):
    a = rng.standard_normal((n, n))
    return a - a.T


@pytest.mark.parametrize("n,count", [(64, 3000), (224, 300)])
def test_pfinv_inplace_many_planes(M, n, count):
    """mvmc_pfinv_f64_inplace (pfaffian_inverse_batch_in_place,
    stage_backend.rs:145-155): thousands of planes (hundreds of workgroups
    beyond one per CU), with a zero plane and a plane whose elimination meets
    a zero column inside the batch: Pfaffians, statuses and the in-place
    inverses bit-exact against the oracle, failed planes zero-filled and
    their neighbours unaffected."""
    import torch
    rng = np.random.default_rng(n)
    planes = np.stack([_random_skew(rng, n).T.copy() for _ in range(count)])
    z0 = count // 3
    planes[z0] = 0.0
    z1 = 2 * count // 3
    planes[z1] = 0.0
    for i in range(0, n, 2):     # block-diagonal 2x2 blocks, one of them zero
        if i != n // 2:
            planes[z1, i + 1, i], planes[z1, i, i + 1] = 1.0 + i, -(1.0 + i)
    log(f"pfinv in place n={n} count={count}")
    pf_o, inv_o, st_o = O.pfinv(planes)
    assert st_o[z0] > 0 and st_o[z1] > 0 and (st_o == 0).sum() == count - 2
    dp = torch.tensor(planes, device="cuda")
    pf = torch.zeros(count, dtype=torch.float64, device="cuda")
    st = torch.zeros(count, dtype=torch.int32, device="cuda")
    M.pfaffian_inverse_batched_inplace(dp, pf, st, n, count)
    torch.cuda.synchronize()
    assert np.array_equal(st.cpu().numpy(), st_o)
    assert np.array_equal(pf.cpu().numpy(), pf_o)
    out = dp.cpu().numpy()
    assert np.array_equal(out, inv_o)
    assert not out[z0].any() and not out[z1].any()


def test_derivative_form_square8_central_differences(M, tmp_path):
    """Differentiation form (3) on the 8x8 lattice (Nsize 64, 8 QP planes,
    Gutzwiller + distance Jastrow + 2x2-sublattice orbitals): the measured O
    rows d log|psi(x)| / d p of saved configurations against central finite
    differences of log|psi| of the same configurations (h = 1e-6; truncation
    O(h^2) and rounding eps |log psi| / h ~ 1e-9 are below the 1e-6 bar)."""
    import lattice
    info = lattice.write_workload("square8", str(tmp_path / "sq8"), sr_steps=1, warmup=4)
    W = 2
    ens = M.Ensemble(info["namelist"], info["initial"], walkers=W)
    ens.sample()
    ens.measure()
    Orows = np.concatenate([ens.o_store(w).reshape(ens.nsample, ens.size) for w in range(W)])
    proj, slater = ens.params()
    lp0 = ens.log_psi().reshape(-1)
    assert np.all(np.isfinite(lp0))
    h = 1e-6
    picks = [("proj", k) for k in (0, 1, 2, len(proj) - 1)] + \
            [("slater", k) for k in range(0, len(slater), max(1, len(slater) // 12))]
    log(f"square8 derivative form: {len(picks)} parameters, {Orows.shape[0]} samples")
    worst = 0.0
    for kind, k in picks:
        vals = []
        for sgn in (1.0, -1.0):
            p2, s2 = proj.copy(), slater.copy()
            (p2 if kind == "proj" else s2)[k] += sgn * h
            ens.set_params(p2, s2)
            vals.append(ens.log_psi().reshape(-1))
        ens.set_params(proj, slater)
        fd = (vals[0] - vals[1]) / (2 * h)
        col = 1 + k if kind == "proj" else 1 + len(proj) + k
        err = np.abs(fd - Orows[:, col]) / np.maximum(1.0, np.abs(Orows[:, col]))
        worst = max(worst, float(err.max()))
        assert np.all(err <= 1e-6), (kind, k, fd, Orows[:, col])
    log(f"  worst relative difference {worst:.2e}")
    assert np.array_equal(ens.log_psi().reshape(-1), lp0)
    ens.close()


def test_heisenberg_ten_steps_vs_c_zvo_rows(M):
    """Ten optimisation steps of the Heisenberg chain on the device (exchange
    sampler, local spins, 8 spin-projection points, direct SR with the MFMA
    Gram, NStore 1, 100 samples, seed 1): row 0 (before any update)
    bit-exact against the oracle (device exp/log) and within 4e-16 of C;
    every column of every row within the reference's 5e-10 of the C 1.4.0
    rows (phase4_zvo_gate.rs:58-66) and of the oracle (the direct-SR solve
    sums the Gram in MFMA order and factorises with rocSOLVER: last-bit
    differences from step 1 on, as between LAPACK builds)."""
    from test_oracle_zvo_heisenberg import NAMELIST, c_rows
    ref = c_rows()
    ens = M.Ensemble(NAMELIST, walkers=1)
    got = np.array([ens.opt_step() for _ in range(10)])
    ens.close()
    O.set_math(1)
    try:
        run = O.Run(NAMELIST, walkers=1)
        for s in range(10):
            assert run.step(do_sr=True) == 0
            ro = run.energies()[-1]
            log(f"  step {s}: E gpu {got[s, 0]!r} oracle {ro[0]!r} C {ref[s, 0]!r}")
            if s == 0:
                assert np.array_equal(got[s], ro), (got[s], ro)
            assert np.all(np.abs(got[s] - ro) <= np.maximum(5e-10, 5e-10 * np.abs(ro))), (s, got[s], ro)
    finally:
        O.set_math(0)
    assert abs(got[0, 0] - ref[0, 0]) <= 4e-16 * abs(ref[0, 0])
    tol = np.maximum(5e-10, 5e-10 * np.abs(ref))
    assert np.all(np.abs(got - ref) <= tol), np.abs(got - ref).max()


def test_precision_form_square16_replay(M, tmp_path):
    """Precision form (sampler inverse stored and updated in FP32) on the 16x16
    lattice, 64 walkers, from the same walker states, seeds and parameters as
    the FP64 run: one SR step each; the rows are finite and the energies per
    site agree within 5 combined Monte Carlo standard errors.  With the
    initial parameters most chains part within the step (a decision flips
    once the FP32 ratio error crosses a draw), so the difference is of the
    Monte Carlo error's size; the count of walkers whose saved samples stay
    identical is logged."""
    import lattice
    info = lattice.write_workload("square16", str(tmp_path / "sq16"), sr_steps=1, warmup=1)
    W = 64
    res = {}
    for mode in ("fp64", "mixed"):
        ens = M.Ensemble(info["namelist"], info["initial"], walkers=W)
        ens.set_precision(mode)
        row = ens.opt_step()
        res[mode] = (row, [ens.walker(w)["ele_idx"].copy() for w in range(W)], ens.nsample)
        ens.close()
    r64, c64, ns = res["fp64"]
    r32, c32, _ = res["mixed"]
    assert np.all(np.isfinite(r64)) and np.all(np.isfinite(r32))
    same = sum(np.array_equal(a, b) for a, b in zip(c64, c32))
    de = abs(r32[0] - r64[0]) / info["nsite"]
    se = np.sqrt((max(r64[2] - r64[0] ** 2, 0.0) + max(r32[2] - r32[0] ** 2, 0.0)) / (W * ns)) / info["nsite"]
    log(f"square16 precision form vs FP64 from identical states: |dE/site| {de:.2e} "
        f"(combined standard error {se:.2e}), {same}/{W} walkers' samples identical")
    assert de <= 5 * se
