"""Register-resident batched Pfaffian + inverse through mvmc_pfinv_f64: the
panel form (64 < n <= 224, one plane per 1024-thread workgroup) and the full
form (n <= 64, one plane per 256-thread workgroup, the whole plane in
registers): Pfaffians, statuses and inverses bit-exact against the oracle
across the register-block sizes, more planes than workgroups, in place, and
with zero and non-finite planes inside the batch; and both forms as the
sampler's rebuild variant."""
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
    import torch
    torch.zeros(1, device="cuda")
    torch.cuda.synchronize()
    import mvmc_hip
    return mvmc_hip


def _skew(rng, count, n):
    a = rng.standard_normal((count, n, n))
    return np.ascontiguousarray(a - a.transpose(0, 2, 1))


@pytest.mark.parametrize("n", [66, 96, 98, 128, 130, 144, 160, 176, 192, 194, 200, 222, 224])
def test_rr_bit_exact_vs_oracle(M, n, monkeypatch):
    """The panel form at every n of its range (MVMC_PFINV_RR=2 routes the
    sizes where the panel-delayed kernel is the default to it too)."""
    import torch
    monkeypatch.setenv("MVMC_PFINV_RR", "2")
    rng = np.random.default_rng(1000 + n)
    count = 300 if n == 224 else 40
    planes = _skew(rng, count, n)
    planes[count // 3] = 0.0                 # zero pivot at once
    planes[count // 2] *= 1e300              # non-finite Pfaffian
    log(f"rr n={n} count={count}")
    pf_o, inv_o, st_o = O.pfinv(planes)
    assert st_o[count // 3] > 0 and st_o[count // 2] == -1
    dp = torch.tensor(planes, device="cuda")
    pf = torch.zeros(count, dtype=torch.float64, device="cuda")
    st = torch.zeros(count, dtype=torch.int32, device="cuda")
    inv = torch.full_like(dp, 5.0)
    M.pfaffian_inverse_batched(dp, pf, inv, st, n, count)
    torch.cuda.synchronize()
    assert np.array_equal(st.cpu().numpy(), st_o)
    pg = pf.cpu().numpy()
    assert np.array_equal(pg[np.isfinite(pf_o)], pf_o[np.isfinite(pf_o)])
    assert np.array_equal(np.isinf(pg), np.isinf(pf_o))
    assert np.array_equal(inv.cpu().numpy(), inv_o)
    # in place
    M.pfaffian_inverse_batched_inplace(dp, pf, st, n, count)
    torch.cuda.synchronize()
    assert np.array_equal(dp.cpu().numpy(), inv_o)


@pytest.mark.parametrize("n", [2, 4, 6, 16, 30, 32, 34, 48, 62, 64])
def test_rr_full_form_bit_exact_vs_oracle(M, n):
    """The full form (n <= 64: 16 x 16 threads, the whole plane in registers,
    utu2inv's small-matrix product over every element) through
    mvmc_pfinv_f64: Pfaffians, statuses and inverses bit-exact against the
    oracle, with a zero plane and a non-finite one in the batch, in place
    too."""
    import torch
    rng = np.random.default_rng(2000 + n)
    count = 600
    planes = _skew(rng, count, n)
    planes[count // 3] = 0.0
    planes[count // 2] *= 1e300
    log(f"rr full form n={n} count={count}")
    pf_o, inv_o, st_o = O.pfinv(planes)
    dp = torch.tensor(planes, device="cuda")
    pf = torch.zeros(count, dtype=torch.float64, device="cuda")
    st = torch.zeros(count, dtype=torch.int32, device="cuda")
    inv = torch.full_like(dp, 5.0)
    M.pfaffian_inverse_batched(dp, pf, inv, st, n, count)
    torch.cuda.synchronize()
    assert np.array_equal(st.cpu().numpy(), st_o)
    pg = pf.cpu().numpy()
    fin = np.isfinite(pf_o)
    assert np.array_equal(pg[fin], pf_o[fin])
    assert np.array_equal(inv.cpu().numpy(), inv_o)
    M.pfaffian_inverse_batched_inplace(dp, pf, st, n, count)
    torch.cuda.synchronize()
    assert np.array_equal(dp.cpu().numpy(), inv_o)


@pytest.mark.parametrize("case,W", [("chain16", 8), ("square8", 2), ("uneven10", 2), ("uneven12", 2),
                                    ("square16", 1)])
def test_rr_rebuilds_in_the_sampler_vs_oracle(M, tmp_path, case, W, monkeypatch):
    """The sampler's periodic rebuild, the measurement rebuild and, from the
    second step on, the rebuild of the walkers' burn configurations at the
    start of a sampling call (k_begin_rr) through the register-resident
    kernels (rebuild variant 2, fixed with MVMC_RB_VARIANT): the full form on
    the 16-site chain and the 8x8 lattice (Nsize 16, 64), the panel form on
    uneven-batch runs of side 10 and 12 (Nsize 90, 112) and the 16x16
    lattice (Nsize 224), two SR-CG steps: zvo rows, parameters and every
    walker's state bit-exact against the oracle after each step, with
    periodic rebuilds done."""
    import lattice
    if case in ("chain16", "square8", "square16"):
        info = lattice.write_workload(case, str(tmp_path / case), sr_steps=2, warmup=1 if case == "square16" else 4)
    else:
        side = int(case[len("uneven"):])
        info = [i for i in lattice.write_uneven_batch(str(tmp_path / "b"), walkers=W, runs_per_side=1, sr_steps=2,
                                                      warmup=2) if i["L"] == side][0]
        assert info["nsize"] > 64
    monkeypatch.setenv("MVMC_RB_VARIANT", "2")
    log(f"rr rebuild {case} W={W}")
    ens = M.Ensemble(info["namelist"], info["initial"], walkers=W)
    O.set_math(1)
    try:
        run = O.Run(info["namelist"], info["initial"], walkers=W)
        for step in range(2):
            row = ens.opt_step()
            t = ens.timings()
            log(f"  step {step}: rebuilt {t['rebuilt']}, variant {t['rebuild_variant']}")
            assert t["rebuild_variant"] == 2 and t["rebuilt"] > 0, t
            run.step(do_sr=True)
            assert np.array_equal(row, run.energies()[-1]), step
            assert np.array_equal(np.concatenate(ens.params()), np.concatenate(run.params())), step
            for w in range(W):
                sg, so = ens.walker(w), run.walker(w)
                assert np.array_equal(sg["ele_idx"], so["ele_idx"]), (step, w)
                assert np.array_equal(sg["rng"], so["rng"]) and sg["rng_idx"] == so["rng_idx"], (step, w)
    finally:
        O.set_math(0)
        ens.close()
