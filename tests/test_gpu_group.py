"""Batching form on the uneven batch (mvmc_group_*, one launch per kernel
class over independent runs of Nsize 28..144): every run's zvo rows,
parameters and walker states bit-exact against the same run stepped alone,
with and without graph-captured sweeps, and against the oracle at small W."""
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
    import torch
    torch.zeros(1, device="cuda")
    torch.cuda.synchronize()
    import mvmc_hip
    return mvmc_hip


def _walker_state(e, w):
    s = e.walker(w)
    return s["ele_idx"], s["rng"], s["rng_idx"], tuple(s["counter"])


@pytest.mark.parametrize("graph_rounds", [0, 4])
def test_group_matches_runs_alone(M, tmp_path, graph_rounds):
    """Eight runs (two per side L = 6, 8, 10, 12; U, filling and parameters
    drawn per run), 48 walkers each, two SR-CG steps as one group: each run's
    rows, parameters and every walker's configuration, SFMT state and
    counters equal those of the run stepped alone."""
    import lattice
    infos = lattice.write_uneven_batch(str(tmp_path / "batch"), walkers=48, runs_per_side=2, sr_steps=2, warmup=2)
    R = len(infos)
    assert sorted({i["L"] for i in infos}) == [6, 8, 10, 12]
    W = [48] * R
    log(f"group of {R} runs, Nsize {[i['nsize'] for i in infos]}, graph rounds {graph_rounds}")
    g = M.EnsembleGroup([i["namelist"] for i in infos], [i["initial"] for i in infos], W)
    g.set_graph(graph_rounds)
    rows_g = [g.opt_step() for _ in range(2)]
    t = g.timings()
    log(f"  group timings {t}")
    assert t["grid_sampler"] >= R
    if graph_rounds:
        assert t["graph_replays"] >= 1
    for r, info in enumerate(infos):
        e = M.Ensemble(info["namelist"], info["initial"], walkers=W[r])
        rows_a = [e.opt_step() for _ in range(2)]
        for k in range(2):
            assert np.array_equal(rows_g[k][r], rows_a[k]), (r, k, rows_g[k][r], rows_a[k])
        gp, ap = np.concatenate(g.runs[r].params()), np.concatenate(e.params())
        assert np.array_equal(gp, ap), r
        for w in (0, W[r] // 2, W[r] - 1):
            a, b = _walker_state(g.runs[r], w), _walker_state(e, w)
            assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1]) and a[2:] == b[2:], (r, w)
        e.close()
    g.close()


def test_group_exchange_sampler_matches_runs_alone(M, tmp_path):
    """The group sampler with exchange moves (NExUpdatePath 1: hops and spin
    exchanges, whose proposals flush the pending hop updates to the planes
    first): one run per side L = 6, 8, 10, 12, 16 walkers each, two SR-CG
    steps; each run's rows, parameters and walker states equal those of the
    run stepped alone, with exchange moves proposed and accepted."""
    import lattice
    infos = lattice.write_uneven_batch(str(tmp_path / "bx"), walkers=16, runs_per_side=1, sr_steps=2, warmup=2,
                                       ex_update_path=1)
    R = len(infos)
    W = [16] * R
    log(f"exchange group of {R} runs, Nsize {[i['nsize'] for i in infos]}")
    g = M.EnsembleGroup([i["namelist"] for i in infos], [i["initial"] for i in infos], W)
    rows_g = [g.opt_step() for _ in range(2)]
    for r, info in enumerate(infos):
        e = M.Ensemble(info["namelist"], info["initial"], walkers=W[r])
        rows_a = [e.opt_step() for _ in range(2)]
        for k in range(2):
            assert np.array_equal(rows_g[k][r], rows_a[k]), (r, k, rows_g[k][r], rows_a[k])
        assert np.array_equal(np.concatenate(g.runs[r].params()), np.concatenate(e.params())), r
        ex = 0
        for w in range(W[r]):
            a, b = _walker_state(g.runs[r], w), _walker_state(e, w)
            assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1]) and a[2:] == b[2:], (r, w)
            ex += a[3][3]
        assert ex > 0, ("no exchange accepted", r)
        e.close()
    g.close()


def test_group_matches_oracle(M, tmp_path):
    """One run per side (Nsize 28..144), 2 walkers each, one SR-CG step as a
    group against the oracle run by run (device exp/log): zvo rows,
    parameters and walker states bit-exact."""
    import lattice
    infos = lattice.write_uneven_batch(str(tmp_path / "b1"), walkers=2, runs_per_side=1, sr_steps=1, warmup=2)
    g = M.EnsembleGroup([i["namelist"] for i in infos], [i["initial"] for i in infos], [2] * len(infos))
    rows = g.opt_step()
    O.set_math(1)
    try:
        for r, info in enumerate(infos):
            log(f"  run {r}: L={info['L']} Nsize={info['nsize']} vs oracle")
            run = O.Run(info["namelist"], info["initial"], walkers=2)
            run.step(do_sr=True)
            assert np.array_equal(rows[r], run.energies()[-1]), r
            gp, op = np.concatenate(g.runs[r].params()), np.concatenate(run.params())
            assert np.array_equal(gp, op), r
            for w in range(2):
                sg, so = g.runs[r].walker(w), run.walker(w)
                assert np.array_equal(sg["ele_idx"], so["ele_idx"]), (r, w)
                assert np.array_equal(sg["rng"], so["rng"]) and sg["rng_idx"] == so["rng_idx"], (r, w)
    finally:
        O.set_math(0)
        g.close()
