"""SrStages on the device (libmvmc_hip.so through its C ABI) against the
oracle: assemble_s_g, gram_complex and sr_s_g bit-exact on random operands;
the direct step (assemble_s_g + cholesky_solve) on the C dposv records; and a
direct-SR (NSRCG 0) optimisation step of the ensemble against the oracle's
direct SR."""
import json
import os
import shutil
import sys

import numpy as np
import pytest

import oracle_py as O

pytestmark = pytest.mark.gpu
GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
HEIS_DIR = os.path.join(GOLDEN, "inputs", "heisenberg_chain")


def log(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def M():
    import torch
    torch.zeros(1, device="cuda")
    torch.cuda.synchronize()
    import mvmc_hip
    return mvmc_hip


def _dev(a, dtype=None):
    import torch
    return torch.tensor(np.ascontiguousarray(a), dtype=dtype, device="cuda")


@pytest.mark.parametrize("cplx", [False, True])
def test_assemble_s_g_bit_exact(M, cplx):
    import torch
    rng = np.random.default_rng(11 + cplx)
    size, nm = 40, 23
    oo = rng.standard_normal((size, size))
    oo = oo @ oo.T / size
    ho = rng.standard_normal(size)
    if cplx:   # real parts read through the complex view
        oo = oo + 1j * rng.standard_normal((size, size))
        ho = ho + 1j * rng.standard_normal(size)
    amap = np.sort(rng.choice(size - 1, nm, replace=False)).astype(np.int32)
    S_o, g_o = O.sr_assemble_s_g(oo.reshape(-1), ho, amap, size, 1, 0.02, 0.013)
    log(f"assemble_s_g complex={cplx} nm={nm}")
    s = torch.zeros(nm * nm, dtype=torch.float64, device="cuda")
    g = torch.zeros(nm, dtype=torch.float64, device="cuda")
    M.sr_assemble_s_g(_dev(oo.reshape(-1)), _dev(ho), _dev(amap), nm, size, 1, 0.02, 0.013, s, g,
                      oo_complex=cplx, ho_complex=cplx)
    torch.cuda.synchronize()
    assert np.array_equal(s.cpu().numpy().reshape(nm, nm).T, S_o)
    assert np.array_equal(g.cpu().numpy(), g_o)


def test_direct_step_on_c_dposv_records(M):
    """Device assemble_s_g + cholesky_solve (rocSOLVER) on the operands of the
    C negative-DSROptStepDt records: every solution within 1e-15 + 1e-13 |x|."""
    import torch
    from test_oracle_sr_direct import HO, OO
    cases = json.load(open(os.path.join(GOLDEN, "negative_stepdt_c_sr.json")))
    for c in cases:
        s = torch.zeros(9, dtype=torch.float64, device="cuda")
        g = torch.zeros(3, dtype=torch.float64, device="cuda")
        M.sr_assemble_s_g(_dev(OO.reshape(-1)), _dev(HO), _dev(np.array([0, 1, 2], dtype=np.int32)), 3, 4, 1,
                          float(c["sta_del"]), c["dt"], s, g)
        info = M.sr_cholesky_solve(s, g, 3)
        x = g.cpu().numpy()
        assert info == 0
        ref = np.array(c["solution"])
        assert np.all(np.abs(x - ref) <= 1e-15 + 1e-13 * np.abs(ref)), (c["name"], x, ref)


@pytest.mark.parametrize("n,samples", [(70, 45), (130, 200), (5, 1)])
def test_gram_complex_bit_exact(M, n, samples):
    import torch
    rng = np.random.default_rng(n)
    st = rng.standard_normal((samples, n)) + 1j * rng.standard_normal((samples, n))
    ref = O.sr_gram_complex(st)
    log(f"gram_complex n={n} samples={samples}")
    out = torch.zeros(n * n, dtype=torch.complex128, device="cuda")
    M.sr_gram_complex(_dev(st), out, n, samples)
    torch.cuda.synchronize()
    got = out.cpu().numpy().reshape(n, n).T
    assert np.array_equal(got, ref)
    assert np.array_equal(got, np.conj(got.T))


@pytest.mark.parametrize("ns,np_", [(50, 70), (129, 3)])
def test_sr_s_g_bit_exact(M, ns, np_):
    import torch
    rng = np.random.default_rng(ns)
    o = rng.standard_normal((ns, np_))
    e = rng.standard_normal(ns)
    w = rng.random(ns)
    w /= w.sum()
    S_o, g_o = O.sr_s_g(o, e, w)
    log(f"sr_s_g ns={ns} np={np_}")
    s = torch.zeros(np_ * np_, dtype=torch.float64, device="cuda")
    g = torch.zeros(np_, dtype=torch.float64, device="cuda")
    M.sr_s_g(_dev(o.T.copy()), _dev(e), _dev(w), s, g, ns, np_)
    torch.cuda.synchronize()
    assert np.array_equal(s.cpu().numpy().reshape(np_, np_).T, S_o)
    assert np.array_equal(g.cpu().numpy(), g_o)


def test_direct_sr_step_matches_oracle(M, tmp_path):
    """Heisenberg chain (exchange sampler), W = 2, NSRCG 0: the step's zvo row
    (formed before the solve) bit-exact against the oracle's direct SR, and
    the parameter update within 1e-14 cond(S) of it (the device Gram is one
    DGEMM of the stored rows and the solve rocSOLVER's; the oracle sums per
    sample and solves in the reference-BLAS order: last-bit differences,
    amplified by the conditioning of S, 14 samples for 12 parameters)."""
    d = tmp_path / "heis_d0"
    shutil.copytree(HEIS_DIR, d)
    mp = d / "modpara.def"
    mp.write_text("\n".join("NSRCG 0" if l.split()[:1] == ["NSRCG"] else l
                            for l in mp.read_text().splitlines()) + "\n")
    nl = str(d / "namelist.def")
    O.set_math(1)
    try:
        run = O.Run(nl, walkers=2)
        ens = M.Ensemble(nl, walkers=2)
        p0 = np.concatenate(run.params())
        assert run.step(do_sr=True) == 0
        row = ens.opt_step()
        assert np.array_equal(row, run.energies()[-1])
        pg, po = np.concatenate(ens.params()), np.concatenate(run.params())
        dg, do = pg - p0, po - p0
        assert np.array_equal(dg != 0, do != 0)
        # conditioning of the solved system, from the oracle's averaged operands
        oo, _ = run.operands()
        size = run.size
        OOm = oo[:size * size].reshape(size, size)
        act = np.nonzero(do)[0]
        S = OOm[1:, 1:][np.ix_(act, act)] - np.outer(OOm[0, 1 + act], OOm[0, 1 + act])
        tol = 1e-14 * np.linalg.cond(S) * np.abs(do).max()
        log(f"  direct SR update: max difference {np.abs(dg - do).max():.2e}, cond(S) {np.linalg.cond(S):.2e}")
        assert np.abs(dg - do).max() <= tol, (np.abs(dg - do).max(), tol)
        ens.close()
    finally:
        O.set_math(0)


@pytest.mark.parametrize("n,samples", [(1, 5), (63, 17), (130, 1000), (200, 33)])
def test_gram_mfma_shapes(M, n, samples):
    """SrStages::gram_real on the MFMA Gram kernel over ragged shapes (one
    partial tile, several tiles, rows not a multiple of the 16-row chunk):
    every entry within 4 samples eps of sum_r |O_ri O_rj| of the exact Gram
    (numpy in extended accumulation is within the same bound), both
    triangles written and exactly symmetric."""
    import torch
    rng = np.random.default_rng(n * 1000 + samples)
    store = rng.standard_normal((samples, n))
    out = torch.full((n, n), 7.0, dtype=torch.float64, device="cuda")
    M.sr_gram_real(_dev(store), out, n, samples)
    torch.cuda.synchronize()
    G = out.cpu().numpy().T
    assert np.array_equal(G, G.T)
    ref = store.T @ store
    bound = 4 * samples * np.finfo(float).eps * (np.abs(store).T @ np.abs(store))
    assert np.all(np.abs(G - ref) <= 2 * bound + 1e-300), np.abs(G - ref).max()


def test_direct_sr_fp32_mfma_gram(M, tmp_path):
    """Direct SR of one measurement twice: FP64 MFMA Gram, then (precision
    form) samples packed in FP32 with the FP32-MFMA Gram; the FP32 update
    within 1e-3 relative of the FP64 one (8x8 lattice, 512 samples, the
    shift 1 + DSROptStaDel on the diagonal)."""
    import lattice
    info = lattice.write_workload("square8", str(tmp_path / "sq8"), sr_steps=1, warmup=4)
    ens = M.Ensemble(info["namelist"], info["initial"], walkers=512)
    ens.sample()
    ens.measure()
    p0 = ens.params()
    assert ens.sr_direct() == 0
    p64 = np.concatenate(ens.params()) - np.concatenate(p0)
    ms64 = ens.timings()["cg_ms"]
    ens.set_params(*p0)
    ens.set_precision("mixed")
    assert ens.sr_direct() == 0
    p32 = np.concatenate(ens.params()) - np.concatenate(p0)
    rel = np.linalg.norm(p32 - p64) / np.linalg.norm(p64)
    log(f"direct SR square8 512 samples: FP32-MFMA update vs FP64 relative {rel:.2e} "
        f"(FP64 {ms64:.2f} ms, FP32 {ens.timings()['cg_ms']:.2f} ms)")
    assert 0 < rel <= 1e-3
    ens.close()


@pytest.mark.parametrize("n,samples", [(16, 4), (130, 1000), (200, 33)])
def test_gram_mfma_f32(M, n, samples):
    """Precision form of gram_real (FP32 store, FP32 MFMA, FP32
    accumulation): every entry within 4 (samples + 2) eps32 of sum_r |O_ri
    O_rj| of the exact Gram of the float32 store; small-integer stores
    (every product and partial sum exact in FP32) give the exact Gram."""
    import torch
    rng = np.random.default_rng(n + samples)
    ints = rng.integers(-3, 4, (samples, n)).astype(np.float32)
    out = torch.zeros((n, n), dtype=torch.float64, device="cuda")
    M.sr_gram_real_f32(_dev(ints, torch.float32), out, n, samples)
    torch.cuda.synchronize()
    G = out.cpu().numpy().T
    ref = ints.astype(np.float64).T @ ints.astype(np.float64)
    assert np.array_equal(G, ref)
    store = rng.standard_normal((samples, n)).astype(np.float32)
    M.sr_gram_real_f32(_dev(store, torch.float32), out, n, samples)
    torch.cuda.synchronize()
    G = out.cpu().numpy().T
    s64 = store.astype(np.float64)
    ref = s64.T @ s64
    bound = 4 * (samples + 2) * np.finfo(np.float32).eps * (np.abs(s64).T @ np.abs(s64))
    assert np.array_equal(G, G.T)
    assert np.all(np.abs(G - ref) <= bound), np.abs(G - ref).max()


def test_sr_cg_fp32_gram_product(M, tmp_path):
    """SR-CG of one measurement twice: the FP64 path (products as passes over
    the FP64 samples), then the precision form (samples packed in FP32, their
    Gram once on FP32 MFMA with FP32 accumulation, every product from the
    Gram); the two updates within 1e-3 relative (8x8 lattice, 512 samples,
    CG to its iteration limit)."""
    import lattice
    info = lattice.write_workload("square8", str(tmp_path / "sq8"), sr_steps=1, warmup=4)
    ens = M.Ensemble(info["namelist"], info["initial"], walkers=512)
    ens.sample()
    ens.measure()
    p0 = ens.params()
    ens.sr_cg()
    t64 = ens.timings()
    d64 = np.concatenate(ens.params()) - np.concatenate(p0)
    ens.set_params(*p0)
    ens.set_precision("mixed")
    ens.sr_cg()
    t32 = ens.timings()
    d32 = np.concatenate(ens.params()) - np.concatenate(p0)
    rel = np.linalg.norm(d32 - d64) / np.linalg.norm(d64)
    log(f"SR-CG square8 512 samples: FP32-Gram update vs FP64 relative {rel:.2e}; "
        f"{t64['cg_iter']:.0f} / {t32['cg_iter']:.0f} iterations, {t64['cg_ms']:.1f} / {t32['cg_ms']:.1f} ms")
    assert 0 < rel <= 1e-3
    ens.close()
