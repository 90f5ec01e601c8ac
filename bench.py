#!/usr/bin/env python3
"""Benchmark: Metropolis samples/s and SR-step seconds of the sampling + SR
optimisation path on one MI355X, 16x16 t-t'-U Hubbard model (Nsize 224,
8 QP planes, 16384 walkers; mvmc-rs_amd/lattice.py workload "square16").

A step is one SR step in the steady state: every walker restores its burn
configuration, runs (NVMCSample + 1) x NVMCInterval x Nsite proposals,
saves NVMCSample samples, the samples are measured and reduced, and the
SR-CG solve updates the parameters.  --warmup untimed steps come first (the
first one thermalises with NVMCWarmUp x NVMCInterval x Nsite extra
proposals; --vmc-warmup sets NVMCWarmUp, default 2, which only changes the
length of that untimed step).

value  = saved samples / second of the whole timed SR steps (all walkers).
roofline: dominant kernel = the sampling kernel k_block; algorithmic bytes
per launch = proposals * NQP * 2n*8 (inverse row + Slater row) + accepted
moves * NQP * 2n^2*8 (inverse read + write), divided by the launch's
average device time (HIP events on the kernel's stream).
cpu_baseline: the FP64 oracle (same algorithm and operation order) on a
bounded walker subset of the same workload, one thread per walker, timed
over at least 10 s of steady SR steps.
precision_forms.mixed_fp32_inverse: the same run (inputs, seeds, steps) with
the sampler's resident inverse in FP32 (mvmc_ens_set_precision 1); its
samples/s, speed-up over the FP64 line and |dE/site| against the FP64 rows
(Monte Carlo size once a chain's decisions differ; the FP64 standard error
is reported beside it).
"""
import argparse
import json

import numpy as np
import os
import sys
import tempfile
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
for sub in ("mvmc-rs_amd", "oracle"):
    sys.path.insert(0, os.path.join(ROOT, sub))

HBM_PEAK_GBS = 8000.0


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def measured_traffic(workload, walkers, steps, warmup):
    """HBM bytes per k_block launch of the last timed step from the two PMC passes
    (FETCH_SIZE, WRITE_SIZE, each scaled by a 1 GiB copy) recorded under profiles/
    for this workload and schedule; None when no record matches this run."""
    path = os.path.join(ROOT, "profiles", f"k_block_{workload}_traffic.json")
    try:
        rec = json.load(open(path))
    except (OSError, ValueError):
        return None
    if (rec.get("walkers"), rec.get("steps"), rec.get("warmup")) != (walkers, steps, warmup):
        return None
    return rec.get("traffic_bytes_per_launch")


def cpu_baseline(workload_dir, walkers, threads, min_seconds=10.0, max_steps=60):
    """Oracle leg: one untimed warm-up SR step (thermalisation), then steady SR
    steps until at least min_seconds of CPU work are timed."""
    import oracle_py
    run = oracle_py.Run(os.path.join(workload_dir, "namelist.def"), os.path.join(workload_dir, "initial.def"),
                        walkers=walkers, threads=threads)
    t0 = time.time()
    run.step(do_sr=True)
    t1 = time.time()
    steps = 0
    while steps < max_steps and (steps == 0 or time.time() - t1 < min_seconds):
        run.step(do_sr=True)
        steps += 1
    t2 = time.time()
    samples = walkers * run.nsample * steps
    return {"value": samples / (t2 - t1), "unit": "samples/s", "cores": threads, "kind": "port",
            "sample": f"{walkers} walkers of the same workload, {steps} steady-state SR steps after 1 warm-up step "
                      f"(warm-up {t1 - t0:.1f} s, timed {t2 - t1:.1f} s), one thread per walker"}


def mixed_form(info, walkers, args, rows64, total64):
    """Warm-up and timed steps of the FP64 run repeated with the sampler's
    inverse in FP32 (same inputs and seeds); |dE/site| per timed step against
    the FP64 run's zvo rows."""
    import mvmc_hip
    ens = mvmc_hip.Ensemble(info["namelist"], info["initial"], walkers=walkers)
    ens.set_precision("mixed")
    for s in range(args.warmup):
        t = time.time()
        ens.opt_step()
        log(f"mixed warmup step {s}: {time.time() - t:.1f} s")
    tm0 = ens.timings()
    times, rows = [], []
    for s in range(args.steps):
        t = time.time()
        rows.append(ens.opt_step())
        times.append(time.time() - t)
        log(f"mixed step {s}: E/site = {rows[-1][0] / info['nsite']:.8f}  {times[-1]:.2f} s")
    tm1 = ens.timings()
    total = sum(times)
    ens.close()
    de = [abs(r[0] - r64[0]) / info["nsite"] for r, r64 in zip(rows, rows64)]
    # Monte Carlo standard error of E/site of the FP64 step (zvo row: Etot, ., Etot2, ...):
    # once a decision flips, the chains differ and dE is of this size
    se = [float(np.sqrt(max(r64[2] - r64[0] ** 2, 0.0) / (walkers * ens.nsample))) / info["nsite"] for r64 in rows64]
    return {"value": walkers * ens.nsample * args.steps / total, "unit": "samples/s",
            "speedup_vs_fp64": total64 / total, "sr_step_s": total / args.steps,
            "sampling_s_per_step": (tm1["sample_s"] - tm0["sample_s"]) / args.steps,
            "energy_per_site": [r[0] / info["nsite"] for r in rows], "abs_dE_per_site_vs_fp64": de,
            "fp64_stat_err_per_site": se,
            "arithmetic": "sampler inverse stored in FP32, its rank-2 updates in FP32; SR-CG samples stored in "
                          "FP32, accumulated in FP64; ratios, vec1 sums, Pfaffians, periodic rebuilds, "
                          "measurement, CG scalars and vectors in FP64"}


def complex_pfinv(n=144, count=8192, cpu_planes=16):
    """Batched complex Pfaffian + inverse (mvmc_pfinv_c128, the run_c64 seam)
    at the 12x12 flux lattice's plane size: device time of one call over
    `count` random skew planes already in HBM (HIP events on the library's
    stream), flops counted as 8 n^3 per plane (LTL, trtri and M^T C at 8 real
    flops per complex multiply-add), beside the oracle on `cpu_planes` planes
    on one host core."""
    import torch
    import mvmc_hip
    import oracle_py
    g = torch.Generator(device="cuda").manual_seed(5)
    a = torch.randn(count, n, n, dtype=torch.complex128, device="cuda", generator=g)
    a = (a - a.transpose(1, 2)).contiguous()
    pf = torch.zeros(count, dtype=torch.complex128, device="cuda")
    inv = torch.empty_like(a)
    st = torch.zeros(count, dtype=torch.int32, device="cuda")
    res = {"n": n, "planes": count}
    for order, name in ((0, "zsktf2"), (1, "c_order")):
        mvmc_hip.pfaffian_inverse_batched_c128(a, pf, inv, st, n, count, order)   # warm-up
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        mvmc_hip.pfaffian_inverse_batched_c128(a, pf, inv, st, n, count, order)
        e1.record()
        e1.synchronize()
        ms = e0.elapsed_time(e1)
        assert int((st != 0).sum()) == 0
        res[name] = {"ms": ms, "planes_per_s": count / (ms / 1e3),
                     "achieved_TFLOPs": 8.0 * n ** 3 * count / (ms / 1e3) / 1e12,
                     "frac_fp64": 8.0 * n ** 3 * count / (ms / 1e3) / 1e12 / 78.6}
    host = a[:cpu_planes].cpu().numpy()
    del a, inv
    t = time.perf_counter()
    oracle_py.pfinv_c128(host, 0)
    dt = time.perf_counter() - t
    res["cpu_oracle"] = {"planes_per_s": cpu_planes / dt, "cores": 1, "sample": f"{cpu_planes} planes, zsktf2 order"}
    return res


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", default="square16")
    ap.add_argument("--vmc-warmup", type=int, default=2,
                    help="NVMCWarmUp of the generated workload: thermalisation sweeps of the first (untimed) "
                         "SR step only; steady-state steps do not depend on it")
    ap.add_argument("--walkers", type=int, default=0, help="override (not a valid headline number)")
    ap.add_argument("--cpu-walkers", type=int, default=16)
    ap.add_argument("--cpu-threads", type=int, default=16)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-direct", action="store_true", help="skip the direct-SR measurement")
    ap.add_argument("--no-c128", action="store_true", help="skip the complex batched Pfaffian measurement")
    ap.add_argument("--c128-only", action="store_true",
                    help="run only the complex batched Pfaffian measurement and print its JSON (profiling aid; "
                         "not the benchmark line)")
    ap.add_argument("--no-mixed", action="store_true",
                    help="skip the precision-form run (sampler inverse in FP32) reported beside the FP64 line")
    ap.add_argument("--copy-probe-mib", type=int, default=0,
                    help="before the run, copy a buffer of this many MiB once (device-to-device), the known-size "
                         "kernel that scales FETCH_SIZE / WRITE_SIZE counters in a profiler run")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU only (asked for {args.gpus}); walkers shard across GPUs "
              f"by construction but no multi-GPU path is built", file=sys.stderr)
        sys.exit(2)

    if args.c128_only:
        import torch
        torch.cuda.init()
        print(json.dumps({"k_pfinv_c128": complex_pfinv()}), flush=True)
        return
    if args.copy_probe_mib > 0 or not args.no_c128:
        # torch ships its own HIP runtime under the same sonames: it has to
        # bring the device up before this library's first HIP call
        import torch
        torch.cuda.init()
    import lattice
    import mvmc_hip
    if args.copy_probe_mib > 0:
        import torch
        nel = args.copy_probe_mib * (1 << 20) // 8
        src = torch.ones(nel, dtype=torch.float64, device="cuda")
        dst = torch.empty_like(src)
        dst.copy_(src)
        torch.cuda.synchronize()
        log(f"copy probe: {args.copy_probe_mib} MiB read + {args.copy_probe_mib} MiB written")
        del src, dst
    wl_dir = os.path.join(tempfile.gettempdir(), f"mvmc_bench_{args.workload}")
    info = lattice.write_workload(args.workload, wl_dir, sr_steps=args.steps + args.warmup, warmup=args.vmc_warmup)
    walkers = args.walkers or info["walkers"]
    log(f"workload {args.workload}: Nsite={info['nsite']} Nsize={info['nsize']} NQP={info['nqp']} "
        f"NPara={info['npara']} walkers={walkers}")
    t = time.time()
    ens = mvmc_hip.Ensemble(info["namelist"], info["initial"], walkers=walkers)
    log(f"ensemble created in {time.time() - t:.1f} s (grid {ens.grid}, LDS {ens.lds_sampler} B)")
    for s in range(args.warmup):
        t = time.time()
        row = ens.opt_step()
        log(f"warmup step {s}: E/site = {row[0] / info['nsite']:.8f}  ({time.time() - t:.1f} s)")
    tm0 = ens.timings()
    step_times, rows, kernel = [], [], []
    for s in range(args.steps):
        before = ens.timings()
        t = time.time()
        row = ens.opt_step()
        dt = time.time() - t
        after = ens.timings()
        step_times.append(dt)
        rows.append(row)
        kernel.append((after["block_ms"], after["block_launches"], after["proposals"] - before["proposals"],
                       after["accepted"] - before["accepted"]))
        log(f"step {s}: E/site = {row[0] / info['nsite']:.8f}  {dt:.2f} s  (sampling {after['sample_s'] - before['sample_s']:.2f} s, "
            f"measure {after['measure_s'] - before['measure_s']:.2f} s, SR {after['sr_s'] - before['sr_s']:.2f} s)")
    tm1 = ens.timings()
    total = sum(step_times)
    samples = walkers * ens.nsample * args.steps
    n, nqp = ens.n, ens.nqp
    # roofline of the sampling kernel (last timed step's launches)
    block_ms, launches, props, accs = kernel[-1]
    bytes_total = props * nqp * 2 * n * 8 + accs * nqp * 2 * n * n * 8
    avg_launch_s = block_ms / 1e3 / max(launches, 1)
    achieved = bytes_total / max(launches, 1) / avg_launch_s / 1e9
    out = {
        "metric": "Metropolis samples/sec + SR-step s, 16x16 Hubbard Nsize 224; |dE/site| vs FP64",
        "value": samples / total,
        "unit": "samples/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": 1e3 * total / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": f"{args.workload}: {info['nsite']}-site t-t'-U Hubbard, Nsize {info['nsize']}, "
                               f"NQP {nqp}, NPara {info['npara']}, {walkers} walkers, 1 sample/walker, "
                               f"16 sweeps/sample, SR-CG",
                   "walkers": walkers, "nsize": n, "nqp": nqp, "npara": info["npara"], "parallelism": "walkers"},
        "sr_step_s": total / args.steps,
        "sampling_s_per_step": (tm1["sample_s"] - tm0["sample_s"]) / args.steps,
        "measure_s_per_step": (tm1["measure_s"] - tm0["measure_s"]) / args.steps,
        "sr_cg_s_per_step": (tm1["sr_s"] - tm0["sr_s"]) / args.steps,
        "proposals_per_s": (tm1["proposals"] - tm0["proposals"]) / total,
        "acceptance": accs / max(props, 1),
        "energy_per_site": [r[0] / info["nsite"] for r in rows],
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": measured_traffic(args.workload, walkers, args.steps, args.warmup),
                     "kernel": "k_block", "avg_launch_ms": avg_launch_s * 1e3, "launches": launches,
                     "algorithmic_bytes_per_launch": bytes_total / max(launches, 1)},
    }
    # the other rows' kernels in the timed step, against their algorithmic bytes
    # (HBM 8 TB/s) and flops (FP64 vector 78.6 TFLOP/s)
    tl = ens.timings()
    plane_b = 2 * n * n * 8 + 12            # read A, write invM (+ status, Pf)
    plane_f = 2 * n ** 3                   # LTL (2/3 n^3) + trtri (1/3) + trmm (1)
    sec = {}
    for key, cnt, ms in (("k_rebuild (a): periodic Pf + inverse", tl["rebuilt"], tl["rebuild_ms"]),
                         ("k_meas_rebuild (a): Pf + inverse per saved sample", tl["measured"], tl["meas_rebuild_ms"])):
        if ms > 0 and cnt > 0:
            sec[key] = {"planes": cnt * nqp, "ms": ms, "achieved_GBs": cnt * nqp * plane_b / (ms / 1e3) / 1e9,
                        "frac_hbm": cnt * nqp * plane_b / (ms / 1e3) / 1e9 / HBM_PEAK_GBS,
                        "achieved_TFLOPs": cnt * nqp * plane_f / (ms / 1e3) / 1e12,
                        "frac_fp64": cnt * nqp * plane_f / (ms / 1e3) / 1e12 / 78.6}
    if tl["measure_ms"] > 0:
        mb = tl["measured"] * nqp * (n * n + info["ntransfer"] * n) * 8
        sec["k_measure (c): Green ratios + O"] = {"samples": tl["measured"], "ms": tl["measure_ms"],
                                                   "achieved_GBs": mb / (tl["measure_ms"] / 1e3) / 1e9,
                                                   "frac_hbm": mb / (tl["measure_ms"] / 1e3) / 1e9 / HBM_PEAK_GBS}
    if tl["cg_ms"] > 0 and tl["cg_iter"] > 0:
        cb = tl["cg_iter"] * tl["cg_rows"] * tl["cg_na"] * 8   # one pass over the packed samples per S.x
        sec["k_cg_* (d): SR-CG"] = {"iterations": tl["cg_iter"], "ms": tl["cg_ms"],
                                    "achieved_GBs": cb / (tl["cg_ms"] / 1e3) / 1e9,
                                    "frac_hbm": cb / (tl["cg_ms"] / 1e3) / 1e9 / HBM_PEAK_GBS}
    # direct SR (NSRCG 0 path) on the same measured operands, after the timed
    # steps: Gram by DGEMM on the matrix cores + Cholesky
    if not args.no_direct:
        ens.sr_direct()   # first call: rocBLAS / rocSOLVER initialisation
        ens.sr_direct()
        td = ens.timings()
        na_, rows_ = td["cg_na"], td["cg_rows"]
        fl = 2.0 * na_ * na_ * rows_ + na_ ** 3 / 3.0 + 2.0 * na_ * na_   # DGEMM (full Gram) + Cholesky
        sec["direct SR (d, NSRCG 0): DGEMM + DPOTRF/DPOTRS"] = {
            "active_parameters": na_, "rows": rows_, "ms": td["cg_ms"],
            "achieved_TFLOPs": fl / (td["cg_ms"] / 1e3) / 1e12,
            "frac_fp64_mfma": fl / (td["cg_ms"] / 1e3) / 1e12 / 78.6}
    if not args.no_c128:
        sec["k_pfinv_c128 (a, complex planes of the 12x12 flux lattice)"] = complex_pfinv()
    out["secondary_rooflines"] = sec
    out["autotuned"] = {"rebuild_variant": ["3 workgroups/CU, 8-column close batches",
                                            "2 workgroups/CU, 16-column close batches",
                                            "not decided"][tl["rebuild_variant"] if tl["rebuild_variant"] >= 0 else 2],
                        "rebuild_ms_per_walker_measured": list(tl["rebuild_ms_per_walker"])}
    if not args.no_mixed:
        # precision form (1): the same workload, seeds and step count with the
        # sampler's resident inverse in FP32; energies against the FP64 rows
        ens.close()
        del ens
        out["precision_forms"] = {"mixed_fp32_inverse": mixed_form(info, walkers, args, rows, total)}
    if not args.no_cpu:
        try:
            out["cpu_baseline"] = cpu_baseline(wl_dir, args.cpu_walkers, args.cpu_threads)
        except Exception as e:  # the baseline is reported, never the target
            out["cpu_baseline"] = {"value": None, "error": str(e)}
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
