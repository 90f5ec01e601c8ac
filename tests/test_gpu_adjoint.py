"""Differentiation form by the adjoint kernels (mvmc_ens_vjp_adjoint): the
VJP sum_r cot[r] O[r] formed without per-sample O rows (inverse elements
scattered back onto the Slater table, the orbital map applied once to the
sums) against the VJP over the stored O rows (bit-exact against the oracle
and pinned on the native-C O records elsewhere) and against central finite
differences of sum_r cot[r] log|psi(x_r)|."""
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


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


@pytest.mark.parametrize("name,W,warmup", [("chain16", 64, 4), ("square8", 8, 4), ("square16", 2, 1)])
def test_adjoint_vjp_matches_stored_rows(M, tmp_path, name, W, warmup):
    """Random cotangents: adjoint VJP = sum over the stored O rows to
    rounding (different summation order and association: every component
    within 1e-11 of the largest sum_r |cot_r| |O_rk|), twice in a row (the
    accumulators are cleared per call)."""
    import lattice
    info = lattice.write_workload(name, str(tmp_path / name), sr_steps=1, warmup=warmup)
    ens = M.Ensemble(info["namelist"], info["initial"], walkers=W)
    ens.sample()
    ens.measure()
    rows = W * ens.nsample
    O = np.concatenate([ens.o_store(w).reshape(ens.nsample, ens.size) for w in range(W)])
    rng = np.random.default_rng(W)
    for call in range(2):
        cot = rng.standard_normal(rows)
        ref = ens.vjp(cot)
        got = ens.vjp_adjoint(cot)
        scale = np.abs(cot) @ np.abs(O)
        err = np.abs(got - ref)
        log(f"{name} W={W} call {call}: max |adjoint - rows| {err.max():.2e}, max scale {scale.max():.2e}, "
            f"adjoint {ens.timings()['adjoint_vjp_ms']:.2f} ms")
        assert err.max() <= 1e-11 * scale.max(), (int(np.argmax(err)), err.max(), scale.max())
        assert np.any(got[1 + len(ens.params()[0]):] != 0.0)
    ens.close()


def test_adjoint_vjp_central_differences_square8(M, tmp_path):
    """8x8 lattice, 8 QP planes: each picked component of the adjoint VJP
    against central differences (h = 1e-6) of F(p) = sum_r cot[r]
    log|psi_p(x_r)| on the same configurations."""
    import lattice
    info = lattice.write_workload("square8", str(tmp_path / "sq8"), sr_steps=1, warmup=4)
    W = 4
    ens = M.Ensemble(info["namelist"], info["initial"], walkers=W)
    ens.sample()
    ens.measure()
    rng = np.random.default_rng(5)
    cot = rng.standard_normal(W * ens.nsample)
    g = ens.vjp_adjoint(cot)
    proj, slater = ens.params()
    h = 1e-6
    picks = [("proj", k) for k in (0, 1, len(proj) - 1)] + \
            [("slater", k) for k in range(0, len(slater), max(1, len(slater) // 10))]
    worst = 0.0
    for kind, k in picks:
        f = []
        for sgn in (1.0, -1.0):
            p2, s2 = proj.copy(), slater.copy()
            (p2 if kind == "proj" else s2)[k] += sgn * h
            ens.set_params(p2, s2)
            f.append(float(cot @ ens.log_psi().reshape(-1)))
        ens.set_params(proj, slater)
        fd = (f[0] - f[1]) / (2 * h)
        col = 1 + k if kind == "proj" else 1 + len(proj) + k
        rel = abs(fd - g[col]) / max(1.0, abs(g[col]))
        worst = max(worst, rel)
        assert rel <= 1e-6, (kind, k, fd, g[col])
    log(f"square8 adjoint VJP vs central differences: {len(picks)} components, worst relative {worst:.2e}")
    ens.close()
