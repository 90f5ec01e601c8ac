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
