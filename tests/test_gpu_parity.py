"""GPU path (libmvmc_hip.so through its C ABI) against the oracle and the
reference's golden records.  Bar: bit-exact configurations, SFMT states,
counters, Pfaffians and inverses (the device reproduces the C operation
order, -ffp-contract=off); floating-point operands bit-exact against the
oracle run with the same correctly rounded exp/log, and within the
reference's own tolerances (1e-14 + 1e-13 |b|) against the C records."""
import json
import math
import os
import struct
import sys

import numpy as np
import pytest

import oracle_py as O

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
HUB = os.path.join(GOLDEN, "inputs", "hubbard_chain", "namelist.def")
HEIS = os.path.join(GOLDEN, "inputs", "heisenberg_chain", "namelist.def")

pytestmark = pytest.mark.gpu


def log(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def M():
    # torch and libmvmc_hip.so each link a HIP runtime (same sonames); in a
    # process that uses both, torch brings its runtime up first and the
    # library binds to it, whichever test runs first
    import torch
    torch.zeros(1, device="cuda")
    torch.cuda.synchronize()
    import mvmc_hip
    return mvmc_hip


def hexd(s):
    return struct.unpack(">d", bytes.fromhex(s))[0]


def _random_skew(rng, n):
    a = rng.standard_normal((n, n))
    return a - a.T


def test_pfinv_bit_exact_vs_oracle(M):
    import torch
    rng = np.random.default_rng(3)
    data = np.load(os.path.join(GOLDEN, "pfapack_real_inputs.npz"))
    # 224: the benchmark plane order; 258 and 300 take more than one group of
    # 256 rows in the row-owner passes of the inverse
    cases = [data[k] for k in data.files] + [_random_skew(rng, 224) for _ in range(3)] + \
            [_random_skew(rng, 16) for _ in range(9)] + [_random_skew(rng, 258)] + \
            [_random_skew(rng, 300) for _ in range(2)]
    by_n = {}
    for A in cases:
        by_n.setdefault(A.shape[0], []).append(A)
    for n, mats in sorted(by_n.items()):
        log(f"pfinv n={n} count={len(mats)}")
        planes = np.stack([m.T.copy() for m in mats])  # column-major planes
        pf_o, inv_o, st_o = O.pfinv(planes)
        dp = torch.tensor(planes, device="cuda")
        pf = torch.zeros(len(mats), dtype=torch.float64, device="cuda")
        inv = torch.zeros_like(dp)
        st = torch.zeros(len(mats), dtype=torch.int32, device="cuda")
        M.pfaffian_inverse_batched(dp, pf, inv, st, n, len(mats))
        torch.cuda.synchronize()
        assert np.array_equal(st.cpu().numpy(), st_o)
        assert np.array_equal(pf.cpu().numpy(), pf_o), n
        assert np.array_equal(inv.cpu().numpy(), inv_o), n


def test_pfinv_zero_plane_status(M):
    """Zero plane -> positive INFO, pf = 0, zero-filled inverse; a plane whose
    LTL elimination meets a zero column (block diagonal) reports the same INFO
    as the sequential algorithm; failed planes do not disturb their neighbours."""
    import torch
    rng = np.random.default_rng(9)
    z = np.zeros((3, 4, 4))
    z[1, 0, 1], z[1, 1, 0], z[1, 2, 3], z[1, 3, 2] = 1.0, -1.0, 2.0, -2.0
    z[2] = _random_skew(rng, 4).T
    pf_o, inv_o, st_o = O.pfinv(z)
    dz = torch.tensor(z, device="cuda")
    pf = torch.zeros(3, dtype=torch.float64, device="cuda")
    inv = torch.zeros_like(dz)
    st = torch.zeros(3, dtype=torch.int32, device="cuda")
    M.pfaffian_inverse_batched(dz, pf, inv, st, 4, 3)
    torch.cuda.synchronize()
    s = st.cpu().numpy()
    assert s[0] > 0 and s[2] == 0
    assert np.array_equal(s, st_o)
    assert np.array_equal(pf.cpu().numpy(), pf_o)
    assert np.array_equal(inv.cpu().numpy(), inv_o)
    assert pf_o[0] == 0.0 and not inv_o[0].any()


def test_pfinv_rejects_odd_n(M):
    import torch
    p = torch.zeros((1, 3, 3), dtype=torch.float64, device="cuda")
    with pytest.raises(M.InvalidShape):
        M.pfaffian_inverse_batched(p, torch.zeros(1, dtype=torch.float64, device="cuda"), p,
                                   torch.zeros(1, dtype=torch.int32, device="cuda"), 3, 1)


def test_pfinv_nonfinite_plane_reports_value(M):
    """A plane whose Pfaffian overflows: status -1, the computed (infinite)
    Pfaffian reported as the oracle forms it, zero inverse; its neighbour is
    unaffected (crates/mvmc-gpu/src/lib.rs:12-15)."""
    import torch
    rng = np.random.default_rng(12)
    big = _random_skew(rng, 8) * 1e200
    ok = _random_skew(rng, 8)
    planes = np.stack([big.T.copy(), ok.T.copy()])
    pf_o, inv_o, st_o = O.pfinv(planes)
    assert st_o[0] == -1 and np.isinf(pf_o[0]) and st_o[1] == 0
    dp = torch.tensor(planes, device="cuda")
    pf = torch.zeros(2, dtype=torch.float64, device="cuda")
    inv = torch.zeros_like(dp)
    st = torch.zeros(2, dtype=torch.int32, device="cuda")
    M.pfaffian_inverse_batched(dp, pf, inv, st, 8, 2)
    torch.cuda.synchronize()
    assert np.array_equal(st.cpu().numpy(), st_o)
    assert np.array_equal(pf.cpu().numpy(), pf_o)
    assert np.array_equal(inv.cpu().numpy(), inv_o)


def _random_skew_c(rng, n):
    a = rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n))
    return a - a.T


def _run_c128(M, planes, n, order):
    import torch
    dp = torch.tensor(planes, dtype=torch.complex128, device="cuda")
    count = planes.shape[0]
    pf = torch.zeros(count, dtype=torch.complex128, device="cuda")
    inv = torch.zeros_like(dp)
    st = torch.zeros(count, dtype=torch.int32, device="cuda")
    M.pfaffian_inverse_batched_c128(dp, pf, inv, st, n, count, order)
    torch.cuda.synchronize()
    return pf.cpu().numpy(), inv.cpu().numpy(), st.cpu().numpy()


@pytest.mark.parametrize("order", [0, 1])
def test_pfinv_c128_bit_exact_vs_oracle(M, order):
    """Complex planes (BatchedEngine::run_c64): the reference's complex inputs
    and random planes on both sides of the 64-column product switch (n = 64,
    66), at the 12x12 flux lattice's n = 144 and above 256 rows (n = 258):
    Pfaffians, inverses and statuses bit-identical to the oracle, for the
    zsktf2 order (0) and the C order of the production sampler (1)."""
    rng = np.random.default_rng(31 + order)
    data = np.load(os.path.join(GOLDEN, "pfapack_complex_inputs.npz"))
    cases = [data[k] for k in data.files] + [_random_skew_c(rng, 2) for _ in range(3)] + \
            [_random_skew_c(rng, 10) for _ in range(5)] + [_random_skew_c(rng, 64)] + \
            [_random_skew_c(rng, 66) for _ in range(2)] + [_random_skew_c(rng, 144) for _ in range(3)] + \
            [_random_skew_c(rng, 258)]
    by_n = {}
    for A in cases:
        by_n.setdefault(A.shape[0], []).append(A)
    for n, mats in sorted(by_n.items()):
        log(f"pfinv_c128 order={order} n={n} count={len(mats)}")
        planes = np.stack([m.T.copy() for m in mats])
        pf_o, inv_o, st_o = O.pfinv_c128(planes, order)
        pf, inv, st = _run_c128(M, planes, n, order)
        assert np.array_equal(st, st_o), n
        assert (st == 0).all(), n
        assert np.array_equal(pf, pf_o), n
        assert np.array_equal(inv, inv_o), n


def test_pfinv_c128_persistent_grid_in_place(M):
    """More planes than the resident grid holds (each workgroup loops over
    several planes), factorised in place (inv == planes), with a zero plane
    inside the batch: every plane bit-identical to the oracle."""
    import torch
    rng = np.random.default_rng(17)
    count, n = 1500, 8
    mats = [_random_skew_c(rng, n) for _ in range(count)]
    mats[777] = np.zeros((n, n), dtype=np.complex128)
    planes = np.stack([m.T.copy() for m in mats])
    for order in (0, 1):
        log(f"pfinv_c128 in place order={order} count={count}")
        pf_o, inv_o, st_o = O.pfinv_c128(planes, order)
        dp = torch.tensor(planes, dtype=torch.complex128, device="cuda")
        pf = torch.zeros(count, dtype=torch.complex128, device="cuda")
        st = torch.zeros(count, dtype=torch.int32, device="cuda")
        M.pfaffian_inverse_batched_c128(dp, pf, dp, st, n, count, order)
        torch.cuda.synchronize()
        assert st_o[777] > 0 and (np.delete(st_o, 777) == 0).all()
        assert np.array_equal(st.cpu().numpy(), st_o)
        assert np.array_equal(pf.cpu().numpy(), pf_o)
        assert np.array_equal(dp.cpu().numpy(), inv_o)


def test_pfinv_c128_statuses_and_shapes(M):
    """Zero plane and zero-column plane: positive INFO, pf = 0, zero inverse;
    a non-finite Pfaffian: status -1 with its value; neighbours unaffected;
    odd n and unknown orders rejected."""
    import torch
    rng = np.random.default_rng(13)
    z = np.zeros((4, 4, 4), dtype=np.complex128)
    z[1][0, 1], z[1][1, 0], z[1][2, 3], z[1][3, 2] = 1 + 1j, -1 - 1j, 2j, -2j
    z[2] = (_random_skew_c(rng, 4) * 1e200).T
    z[3] = _random_skew_c(rng, 4).T
    for order in (0, 1):
        pf_o, inv_o, st_o = O.pfinv_c128(z, order)
        pf, inv, st = _run_c128(M, z, 4, order)
        assert st_o[0] > 0 and st_o[2] == -1 and st_o[3] == 0
        assert np.array_equal(st, st_o)
        # inf - inf in the overflowing product: NaN payloads are not compared
        assert not np.isfinite(pf[2]) and np.array_equal(pf, pf_o, equal_nan=True)
        assert np.array_equal(inv, inv_o)
    p = torch.zeros((1, 3, 3), dtype=torch.complex128, device="cuda")
    out = torch.zeros(1, dtype=torch.complex128, device="cuda")
    sti = torch.zeros(1, dtype=torch.int32, device="cuda")
    with pytest.raises(M.InvalidShape):
        M.pfaffian_inverse_batched_c128(p, out, p, sti, 3, 1)
    p4 = torch.zeros((1, 4, 4), dtype=torch.complex128, device="cuda")
    with pytest.raises(M.InvalidShape):
        M.pfaffian_inverse_batched_c128(p4, out, p4, sti, 4, 1, order=2)


def test_hubbard_step1_operands(M):
    log("hubbard step-1 sample+measure")
    O.set_math(1)
    try:
        ref = {k: np.array([hexd(x) for x in v]) for k, v in
               json.load(open(os.path.join(GOLDEN, "hubbard_cg_store0.json"))).items()}
        run = O.Run(HUB, walkers=1)
        run.step(do_sr=False)
        ens = M.Ensemble(HUB, walkers=1)
        assert np.array_equal(ens.params()[1], run.params()[1])
        ens.sample()
        st_g = ens.walker(0)
        st_o = run.walker(0)
        assert np.array_equal(st_g["ele_idx"], st_o["ele_idx"])
        assert np.array_equal(st_g["rng"], st_o["rng"]) and st_g["rng_idx"] == st_o["rng_idx"]
        assert list(st_g["counter"]) == list(st_o["counter"][:4])
        ens.measure()
        e_g, e_o = ens.energy(), run.energy()
        assert np.array_equal(e_g, e_o), (e_g, e_o)
        oo_g, ho_g = ens.operands()
        oo_o, ho_o = run.operands()
        assert np.array_equal(ho_g, ho_o)
        assert np.array_equal(oo_g, oo_o[:2 * run.size])
        assert np.array_equal(ens.o_store(0), run.o_store(0))
        size = run.size
        assert e_g[1] == ref["energy"][0]
        assert np.all(np.abs(ho_g - ref["ho"]) <= 1e-14 + 1e-13 * np.abs(ref["ho"]))
        assert np.all(np.abs(oo_g - ref["oo"]) <= 1e-14 + 1e-13 * np.abs(ref["oo"]))
        assert ens.status() == 0
    finally:
        O.set_math(0)


@pytest.mark.parametrize("cell", ["real_cg_r2w1_s7", "real_cg_r4w1_s7"])
def test_heisenberg_exchange_walkers_exact(M, cell):
    log(f"heisenberg {cell}")
    fx = json.load(open(os.path.join(GOLDEN, f"mpi179_{cell}.json")))
    W = len(fx["ranks"])
    ens = M.Ensemble(HEIS, walkers=W)
    ens.sample()
    tot = np.zeros(4, dtype=np.int64)
    for w in range(W):
        st = ens.walker(w)
        rec = fx["ranks"][w]
        for s, vals in rec["eleidx"].items():
            assert list(st["ele_idx"][int(s)]) == vals, (w, s)
        assert st["rng_idx"] == rec["rng_idx"]
        assert list(st["rng"]) == rec["rng"]
        tot += st["counter"]
        if w > 0:
            assert list(st["counter"]) == rec["counter"][:4]
    assert list(tot) == fx["ranks"][0]["counter"][:4]
    ens.measure()
    ops = fx["operands"]
    e = ens.energy()
    assert abs(e[1] - ops["energy"][0]) <= 1e-13 + 1e-12 * abs(ops["energy"][0])
    oo, ho = ens.operands()
    assert np.all(np.abs(ho - ops["ho"][0::2]) <= 1e-13 + 1e-12 * np.abs(ops["ho"][0::2]))
    assert np.all(np.abs(oo - ops["oo"][0::2]) <= 1e-13 + 1e-12 * np.abs(ops["oo"][0::2]))
    # against the oracle with the device arithmetic: bit-exact
    O.set_math(1)
    try:
        run = O.Run(HEIS, walkers=W)
        run.step(do_sr=False)
        assert np.array_equal(ens.energy(), run.energy())
        oo_o, ho_o = run.operands()
        assert np.array_equal(ho, ho_o)
        assert np.array_equal(oo, oo_o[:2 * run.size])
    finally:
        O.set_math(0)


def test_multistep_optimisation_matches_oracle(M):
    """Five SR-CG steps on 3 walkers: every zvo_out row and every parameter bit-exact."""
    log("multistep hubbard W=3")
    O.set_math(1)
    try:
        run = O.Run(HUB, walkers=3)
        ens = M.Ensemble(HUB, walkers=3)
        for step in range(5):
            rc = run.step(do_sr=True)
            row = ens.opt_step()
            ro = run.energies()[-1]
            log(f"  step {step}: gpu {row[0]!r} oracle {ro[0]!r}")
            assert np.array_equal(row, ro), step
            pg, sg = ens.params()
            po, so = run.params()
            assert np.array_equal(pg, po) and np.array_equal(sg, so), step
    finally:
        O.set_math(0)


@pytest.mark.parametrize("name,variant", [("sr_cg_real.txt", 0), ("sr_cg_real_diag_scale.txt", 1)])
def test_sr_cg_solver_c_abi(M, name, variant):
    import torch
    from test_oracle_fixtures import _parse_cg
    n, samples, mean, diag, Osamp, g, z_ref, limits = _parse_cg(os.path.join(GOLDEN, name))
    dev = lambda a: torch.tensor(np.ascontiguousarray(a), dtype=torch.float64, device="cuda")
    pack, dm, dd, dg = dev(Osamp), dev(mean), dev(diag), dev(g)
    for limit, status, x_ref, r_ref, d_ref in limits[:12]:
        x = torch.zeros(n, dtype=torch.float64, device="cuda")
        st = M.sr_cg_solve(pack, dm, dd, dg, x, n, 1, samples, 1.0 / samples, 1e-5, 0.0, limit, variant)
        sto, xo, _, _ = O.cg_solve(mean, diag, Osamp, g, 1.0 / samples, 1e-5, 0.0, limit, bool(variant))
        assert st == sto == status
        assert np.array_equal(x.cpu().numpy(), xo), limit


@pytest.mark.parametrize("namelist,W", [(HUB, 8), (HEIS, 4)])
def test_mixed_precision_form(M, namelist, W):
    """Precision form (1): the sampler's resident inverse and the SR-CG
    operand stored in FP32.  Against the FP64 path on the same seeds: the
    first step's Metropolis decisions do not flip on these chains, so every
    configuration and (the measurement being FP64) the first zvo row are
    identical; the SR update, formed from FP32 operands, stays within 1e-3 of
    the FP64 update (relative to its largest component; the FP32 rounding of
    O amplified by the conditioning of S); two more steps stay within 1e-4
    relative in energy; the FP32 inverse after a sampling call holds FP32
    values within 1e-3 (relative to the plane's largest element) of the FP64
    inverse: the drift of at most Nsite FP32 rank-2 updates, each dividing by
    its hop's ratio (measured up to 3e-4 here; a misapplied update is O(1))."""
    log(f"mixed precision {os.path.basename(os.path.dirname(namelist))} W={W}")
    e64 = M.Ensemble(namelist, walkers=W)
    e32 = M.Ensemble(namelist, walkers=W)
    e32.set_precision("mixed")
    p0 = np.concatenate(e64.params())
    r64, r32 = e64.opt_step(), e32.opt_step()
    for w in range(W):
        assert np.array_equal(e64.walker(w)["ele_idx"], e32.walker(w)["ele_idx"]), w
    assert np.array_equal(r64, r32)
    d64 = np.concatenate(e64.params()) - p0
    d32 = np.concatenate(e32.params()) - p0
    rel = np.abs(d32 - d64).max() / np.abs(d64).max()
    log(f"  SR update difference {rel:.2e} (relative to the largest FP64 component)")
    assert rel <= 1e-3
    for step in range(2):
        r64, r32 = e64.opt_step(), e32.opt_step()
        assert abs(r32[0] - r64[0]) <= 1e-4 * abs(r64[0]), (step, r32[0], r64[0])
    e64.sample()
    e32.sample()
    for w in range(W):
        t64, t32 = e64.tables(w)[0], e32.tables(w)[0]
        assert np.array_equal(t32, t32.astype(np.float32).astype(np.float64))
    with pytest.raises(M.InvalidShape):
        e32.set_precision(7)
    # the inverse drift, on identical parameters and configurations
    f64 = M.Ensemble(namelist, walkers=W)
    f32 = M.Ensemble(namelist, walkers=W)
    f32.set_precision("mixed")
    f64.sample()
    f32.sample()
    for w in range(W):
        assert np.array_equal(f64.walker(w)["ele_idx"], f32.walker(w)["ele_idx"]), w
        t64, t32 = f64.tables(w)[0], f32.tables(w)[0]
        scale = np.abs(t64).max()
        assert np.abs(t32 - t64).max() <= 1e-3 * scale, (w, np.abs(t32 - t64).max() / scale)


def test_derivative_form_vjp_and_finite_differences(M):
    """Differentiation form (3): the measurement's O rows are
    d log|psi(x)| / d p (projection counts, Slater adjoint).  Checked
    against central finite differences of log|psi| of the same saved
    configurations (mvmc_ens_log_psi, h = 1e-6: truncation and rounding both
    below 1e-9 here), and the VJP against the same serial sum over the
    downloaded O rows (bit-exact)."""
    log("derivative form: VJP + finite differences")
    W = 3
    ens = M.Ensemble(HUB, walkers=W)
    ens.sample()
    ens.measure()
    O = np.concatenate([ens.o_store(w).reshape(ens.nsample, ens.size) for w in range(W)])
    rows = O.shape[0]
    rng = np.random.default_rng(5)
    cot = rng.standard_normal(rows)
    ref = np.zeros(ens.size)
    for k in range(ens.size):
        acc = 0.0
        for r in range(rows):
            acc += float(cot[r]) * float(O[r, k])
        ref[k] = acc
    assert np.array_equal(ens.vjp(cot), ref)
    proj, slater = ens.params()
    lp0 = ens.log_psi().reshape(-1)
    assert np.all(np.isfinite(lp0))
    h = 1e-6
    picks = [("proj", k) for k in range(min(3, len(proj)))] + \
            [("slater", k) for k in range(0, len(slater), max(1, len(slater) // 6))]
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
        assert np.all(np.abs(fd - O[:, col]) <= 1e-6 * np.maximum(1.0, np.abs(O[:, col]))), (kind, k, fd, O[:, col])
    assert np.array_equal(ens.log_psi().reshape(-1), lp0)


def test_direct_sr_matches_dense_solve(M, tmp_path):
    """Direct SR (NSRCG 0; sr.rs:60-160): the device Gram (rocBLAS DGEMM), S/g
    assembly (sr_backend.rs:484-505: S = <OO> - <O><O>^T, diagonal times
    1 + DSROptStaDel; g = -2 dt (<HO> - <H><O>)) and Cholesky solve give the
    solution of the same system formed on the host from the downloaded O rows
    (numpy, independent summation order: tolerance scaled by cond(S)).
    Bit-level parity with the reference's OpenBLAS/LAPACK order: unpinned."""
    import shutil
    log("direct SR")
    dst = tmp_path / "hub_direct"
    shutil.copytree(os.path.dirname(HUB), dst)
    mp = dst / "modpara.def"
    txt = mp.read_text()
    assert "NSRCG 1" in txt
    mp.write_text(txt.replace("NSRCG 1", "NSRCG 0"))
    W = 4
    ens = M.Ensemble(str(dst / "namelist.def"), walkers=W)
    ens.sample()
    ens.measure()
    size = ens.size
    wc = ens.energy()[0]
    oo, ho = ens.operands()
    mean, var = oo[1:size], oo[size + 1:] - oo[1:size] ** 2
    act = np.nonzero(var >= var.max() * 1e-8)[0]          # DSROptRedCut 1e-8
    O = np.concatenate([ens.o_store(w).reshape(ens.nsample, size) for w in range(W)])
    P = O[:, 1 + act]
    S = P.T @ P / wc - np.outer(mean[act], mean[act])
    S[np.diag_indices_from(S)] *= 1.0 + 0.01              # DSROptStaDel
    g = -2.0 * 0.003 * (ho[1 + act] - ho[0] * mean[act])  # DSROptStepDt
    x_ref = np.linalg.solve(S, g)
    assert ens.sr_direct() == 0
    _, x = ens.cg()
    assert len(x) == len(act)
    tol = max(1e-12, 1e-14 * np.linalg.cond(S)) * np.abs(x_ref).max()
    assert np.abs(x - x_ref).max() <= tol, (np.abs(x - x_ref).max(), tol)
    # a full step through the NSRCG 0 path
    row = ens.opt_step()
    assert np.isfinite(row).all()


def test_uneven_lattices_side_by_side(M, tmp_path):
    """Batching form (2), lattices of uneven Nsize on one GPU: ensembles of a
    16-site chain (Nsize 16), a 24-site chain (Nsize 24) and the 8x8 square
    lattice (Nsize 64, 8 QP planes) live in one process and are stepped
    concurrently from host threads, one HIP stream each; every zvo row equals
    the one of the same ensemble run alone (the handles share no state; the
    kernels' LDS limits only grow)."""
    import threading
    import time
    import lattice
    log("uneven lattices side by side")
    cfgs = [("chain16", {}, 32), ("chain16", {"L": 24, "nup": 12}, 32), ("square8", {}, 64)]
    files = []
    for k, (name, ov, W) in enumerate(cfgs):
        info = lattice.write_workload(name, str(tmp_path / f"wl{k}"), sr_steps=2, warmup=2, overrides=ov)
        files.append((info["namelist"], info["initial"], W, info["nsize"]))
    assert sorted(f[3] for f in files) == [16, 24, 64]
    t0 = time.time()
    alone = []
    for nl, ini, W, _ in files:
        ens = M.Ensemble(nl, ini, walkers=W)
        alone.append([ens.opt_step() for _ in range(2)])
        ens.close()
    t_alone = time.time() - t0
    ens = [M.Ensemble(nl, ini, walkers=W) for nl, ini, W, _ in files]
    out = [None] * len(ens)

    def run(k):
        out[k] = [ens[k].opt_step() for _ in range(2)]

    t0 = time.time()
    th = [threading.Thread(target=run, args=(k,)) for k in range(len(ens))]
    for t in th:
        t.start()
    for t in th:
        t.join()
    t_side = time.time() - t0
    log(f"  alone {t_alone:.2f} s, side by side {t_side:.2f} s")
    for k in range(len(ens)):
        assert out[k] is not None
        for a, b in zip(alone[k], out[k]):
            assert np.array_equal(a, b), k
        ens[k].close()


@pytest.mark.parametrize("na,rows", [(1, 1), (300, 1100), (2500, 37), (9000, 70), (7784, 600)])
def test_sr_cg_local_product(M, na, rows):
    """SrStages::cg_local_product (mvmc_sr_cg_apply): z = S x of the sampled
    operator bit-identical to the oracle's restatement, for one column, for
    several 32-row chunks and 16-chunk superchunks, for every columns-per-lane
    width up to the wide path (na > 8192), and at the benchmark's 7784 active
    parameters."""
    import torch
    rng = np.random.default_rng(na + rows)
    O_ = rng.standard_normal((rows, na))
    mean = O_.mean(axis=0)
    diag = (O_ * O_).mean(axis=0) - mean * mean
    x = rng.standard_normal(na)
    inv_w, shift = 1.0 / rows, 1e-3
    z_o = O.cg_apply(mean, diag, O_, x, inv_w, shift)
    dev = lambda a: torch.tensor(np.ascontiguousarray(a), dtype=torch.float64, device="cuda")
    z = torch.zeros(na, dtype=torch.float64, device="cuda")
    M.sr_cg_apply(dev(O_), dev(mean), dev(diag), dev(x), z, na, rows, 1, inv_w, shift)
    torch.cuda.synchronize()
    assert np.array_equal(z.cpu().numpy(), z_o)


def test_sr_gram_and_cholesky_stages(M):
    """SrStages::gram_real and cholesky_solve (the direct-SR stages): the Gram
    of a [n, samples] store against exact rational sums to DGEMM rounding
    (bit-level parity with OpenBLAS's summation order is unpinned), the
    Cholesky solve against a dense host solve within cond(S) 1e-14; an
    indefinite matrix reports the factorisation's info."""
    import torch
    from fractions import Fraction
    rng = np.random.default_rng(77)
    n, samples = 96, 300
    store = rng.standard_normal((samples, n))          # rows = samples: column-major [n, samples]
    dev = lambda a: torch.tensor(np.ascontiguousarray(a), dtype=torch.float64, device="cuda")
    out = torch.zeros((n, n), dtype=torch.float64, device="cuda")
    M.sr_gram_real(dev(store), out, n, samples)
    torch.cuda.synchronize()
    G = out.cpu().numpy().T                            # column-major -> row-major
    ref = store.T @ store
    assert np.array_equal(G, G.T)
    for i, j in [(0, 0), (5, 17), (95, 3), (40, 40)]:
        exact = float(sum(Fraction(float(a)) * Fraction(float(b)) for a, b in zip(store[:, i], store[:, j])))
        assert abs(G[i, j] - exact) <= samples * 4 * np.finfo(float).eps * np.abs(store[:, i] * store[:, j]).sum()
    assert np.abs(G - ref).max() <= 1e-12 * np.abs(ref).max()
    S = ref / samples + 1e-3 * np.eye(n)
    b = rng.standard_normal(n)
    s_dev, rhs = dev(S.T), dev(b)
    info = M.sr_cholesky_solve(s_dev, rhs, n)
    torch.cuda.synchronize()
    assert info == 0
    x = rhs.cpu().numpy()
    want = np.linalg.solve(S, b)
    assert np.abs(x - want).max() <= np.linalg.cond(S) * 1e-14 * np.abs(want).max()
    bad = -np.eye(4)
    rhs4 = dev(np.ones(4))
    assert M.sr_cholesky_solve(dev(bad), rhs4, 4) > 0


@pytest.mark.parametrize("window", [2, 1])
def test_opt_data_output(M, tmp_path, window):
    """zqp_opt.dat after a 4-step run with a window of the last `window`
    steps (output_opt_data, io.rs:561-727): the file text equals the
    reference rules applied to the recorded energies and post-update
    parameters (C "% .18e"; pairs for one snapshot, mean / 0 / standard
    deviation triples and the Gutzwiller, Jastrow and orbital block files for
    more), and reading it back as an initial.def gives the window means to
    the rounding of the gauge fixing applied on creation."""
    import shutil
    src = os.path.join(GOLDEN, "inputs", "hubbard_chain")
    d = tmp_path / "in"
    shutil.copytree(src, d)
    mp = (d / "modpara.def").read_text().splitlines()
    out = []
    for line in mp:
        t = line.split()
        if t and t[0] == "NSROptItrStep":
            line = "NSROptItrStep 4"
        elif t and t[0] == "NSROptItrSmp":
            line = f"NSROptItrSmp {window}"
        out.append(line)
    (d / "modpara.def").write_text("\n".join(out) + "\n")
    ens = M.Ensemble(str(d / "namelist.def"), walkers=2)
    rows, pars = [], []
    for _ in range(4):
        rows.append(ens.opt_step())
        p_, s_ = ens.params()
        pars.append(np.concatenate([p_, s_]))
    ens.write_opt_data(tmp_path)
    f = lambda x: "% .18e" % x
    win = list(range(4 - window, 4))
    vals = [[rows[k][0], rows[k][2]] + list(pars[k]) for k in win]
    width = len(vals[0])
    if window == 1:
        want = "".join(f"{f(vals[0][i])} {f(0.0)} " for i in range(width)) + "\n"
        means = vals[0]
    else:
        means = []
        want = ""
        for i in range(width):
            s_ = 0.0
            for v in vals:
                s_ += v[i]
            mu = s_ / len(vals)
            var = 0.0
            for v in vals:
                var += (v[i] - mu) * (v[i] - mu)
            means.append(mu)
            want += f"{f(mu)} {f(0.0)} {f(math.sqrt(var / (len(vals) - 1.0)))} "
        want += "\n"
    got = (tmp_path / "zqp_opt.dat").read_text()
    assert got == want
    blocks = sorted(p.name for p in tmp_path.glob("zqp_*_opt.dat"))
    if window == 1:
        assert blocks == []
    else:
        assert "zqp_orbital_opt.dat" in blocks
        orb = (tmp_path / "zqp_orbital_opt.dat").read_text().splitlines()
        assert orb[1].split()[0] == "NOrbitalIdx" and int(orb[1].split()[1]) == ens.nslater
        first = orb[5].split()
        assert first[0] == "0" and first[1] == f(means[2 + ens.nproj]).strip()
    if window == 1:
        # C's one-snapshot form holds (value, 0) pairs, not initial.def triples
        with pytest.raises(M.BackendError):
            M.Ensemble(str(d / "namelist.def"), str(tmp_path / "zqp_opt.dat"), walkers=1)
        ens.close()
        return
    # read back as an initial.def: the means, up to the rounding of the gauge
    # fixing that creation re-applies (sync_modified_parameter)
    back = M.Ensemble(str(d / "namelist.def"), str(tmp_path / "zqp_opt.dat"), walkers=1)
    p_, s_ = back.params()
    m_ = np.array(means[2:])
    assert np.allclose(np.concatenate([p_, s_]), m_, rtol=1e-12, atol=1e-14 * np.abs(m_).max())
    back.close()
    ens.close()


def test_step_output_files(M, tmp_path):
    """Per-step output (output_data, io.rs:142-232): three optimisation steps
    of the Hubbard chain append zvo_out rows byte-identical to the C layout
    (the layout pinned on the C 1.4.0 rows in test_oracle_fixtures) with the
    variance by Julia's division, and zvo_var lines carrying the energies and
    the parameters each step sampled with (before its SR update)."""
    namelist = os.path.join(GOLDEN, "inputs", "hubbard_chain", "namelist.def")
    ens = M.Ensemble(namelist, walkers=2)
    out_dir = tmp_path / "created" / "by" / "writer"   # missing directories are created
    ens.set_output_dir(out_dir)
    f = lambda x: "% .18e" % x
    rows, pre = [], []
    for _ in range(3):
        p_, s_ = ens.params()
        pre.append(np.concatenate([p_, s_]))
        rows.append(ens.opt_step())
    out = (out_dir / "zvo_out.dat").read_text().splitlines()
    var = (out_dir / "zvo_var.dat").read_text().splitlines()
    assert len(out) == 3 and len(var) == 3
    for k, r in enumerate(rows):
        e, e2 = r[0], r[2]
        a, c = e2 - e * e, e * e
        v = (a + 0.0 * (0.0 / c)) * (1.0 / (c + 0.0 * (0.0 / c)))    # Julia's (a + 0i)/(c + 0i)
        want = f"{f(e)} {f(0.0)}  {f(e2)} {f(v)} {'%.18e' % r[4]} {'%.18e' % r[5]}"
        assert out[k] == want, (k, out[k], want)
        fields = var[k].split()
        assert len(fields) == 3 * (2 + len(pre[k]))
        assert float(fields[0]) == e and float(fields[3]) == e2
        got_p = np.array([float(x) for x in fields[6::3]])
        assert np.array_equal(got_p, pre[k])
    ens.set_output_dir(None)
    ens.close()


def test_measure_caller_configurations(M, tmp_path):
    """MeasurementStages::measure_sample on caller-supplied configurations
    (mvmc_ens_measure_configs): the 8x8 benchmark lattice's sampled
    configurations and random placements, over two saved-sample layers;
    local energy, O row and validity bit-identical to the oracle measuring
    the same EleIdx, and the sampled configurations' O rows identical to the
    ensemble's own measurement; malformed configurations rejected."""
    import lattice
    info = lattice.write_workload("square8", str(tmp_path / "sq8"), sr_steps=1, warmup=2, nvmc_sample=2)
    W = 12
    ens = M.Ensemble(info["namelist"], info["initial"], walkers=W)
    log(f"square8 W={W}: sample + measure")
    ens.sample()
    ens.measure()
    own = [ens.o_store(w) for w in range(W)]                      # [nsample][size] per walker
    sampled = [ens.walker(w)["ele_idx"][s] for w in range(W) for s in range(ens.nsample)]
    rng = np.random.default_rng(23)
    ns, ne = ens.nsite, ens.ne
    rand = [np.concatenate([rng.choice(ns, ne, replace=False), rng.choice(ns, ne, replace=False)])
            for _ in range(6)]
    configs = np.array(sampled[:14] + rand, dtype=np.int32)       # 20 > W: two layers
    e, o, ok = ens.measure_configs(configs)
    for c in range(14):
        w, s = divmod(c, ens.nsample)
        assert np.array_equal(o[c], own[w][s]), c
    O.set_math(1)
    try:
        run = O.Run(info["namelist"], info["initial"], walkers=1)
        for c, cfg in enumerate(configs):
            eo, oo, oko = run.measure_idx(cfg)
            assert ok[c] == oko, c
            assert e[c] == eo, (c, e[c], eo)
            assert np.array_equal(o[c], oo), c
    finally:
        O.set_math(0)
    bad = configs[:1].copy()
    bad[0, 0] = ns
    with pytest.raises(M.BackendError):
        ens.measure_configs(bad)
    bad = configs[:1].copy()
    bad[0, 1] = bad[0, 0]
    with pytest.raises(M.BackendError):
        ens.measure_configs(bad)
    with pytest.raises(M.InvalidShape):
        ens.measure_configs(np.zeros((W * ens.nsample + 1, ens.n), dtype=np.int32))
    ens.close()


@pytest.mark.parametrize("W,warmup", [(1100, 2), (16384, 0)])
def test_walkers_beyond_resident_grid_match_single_walker_oracle(M, tmp_path, W, warmup):
    """16x16 benchmark lattice with more than twice as many walkers as the
    sampler's resident grid holds (and at the benchmark's full 16384
    walkers), so walkers wait for slots, are handed to the rebuild kernel and
    resume in later launches: after the first sampling phase (thermalisation
    + one sample, no SR yet) every checked walker, from each scheduling
    round, has the configuration, SFMT state and counters of a one-walker
    oracle run seeded RndSeed + w (walkers are independent chains until the
    first SR update, multichain.rs:88)."""
    import lattice
    info = lattice.write_workload("square16", str(tmp_path / "sq16"), sr_steps=1, warmup=warmup)
    ens = M.Ensemble(info["namelist"], info["initial"], walkers=W)
    assert W > 2 * ens.grid, ens.grid
    log(f"square16 W={W} grid={ens.grid} NVMCWarmUp={warmup}: sampling")
    ens.sample()
    O.set_math(1)
    try:
        for w in sorted({0, ens.grid - 1, ens.grid, 2 * ens.grid + 1, W // 2 + 3, W - 1}):
            log(f"  walker {w} vs one-walker oracle")
            run = O.Run(info["namelist"], info["initial"], walkers=1, seed=11272 + w)
            run.sample()
            sg, so = ens.walker(w), run.walker(0)
            assert np.array_equal(sg["ele_idx"], so["ele_idx"]), w
            assert np.array_equal(sg["rng"], so["rng"]) and sg["rng_idx"] == so["rng_idx"], w
            assert list(sg["counter"]) == list(so["counter"][:4]), (w, sg["counter"], so["counter"][:4])
    finally:
        O.set_math(0)
        ens.close()


@pytest.mark.parametrize("name,W,warmup", [("square8", 2, 4), ("square16", 1, 2)])
def test_benchmark_lattices_match_oracle(M, tmp_path, name, W, warmup):
    """The benchmark's own lattices (8x8: Nsize 64; 16x16 t-t'-U at 1/8
    doping: Nsize 224; 8 QP planes each, sublattice orbitals, Gutzwiller and
    Jastrow), one SR-CG step against the oracle with the device's correctly
    rounded exp/log: the zvo row, every parameter after the update and every
    walker's saved configuration and SFMT state bit-exact."""
    import lattice
    log(f"{name} W={W} vs oracle")
    info = lattice.write_workload(name, str(tmp_path / name), sr_steps=1, warmup=warmup)
    O.set_math(1)
    try:
        run = O.Run(info["namelist"], info["initial"], walkers=W)
        ens = M.Ensemble(info["namelist"], info["initial"], walkers=W)
        run.step(do_sr=True)
        row = ens.opt_step()
        ro = run.energies()[-1]
        assert np.array_equal(row, ro), (row, ro)
        for w in range(W):
            sg, so = ens.walker(w), run.walker(w)
            assert np.array_equal(sg["ele_idx"], so["ele_idx"]), w
            assert np.array_equal(sg["rng"], so["rng"]) and sg["rng_idx"] == so["rng_idx"], w
        pg, sg_ = ens.params()
        po, so_ = run.params()
        assert np.array_equal(pg, po) and np.array_equal(sg_, so_)
    finally:
        O.set_math(0)
