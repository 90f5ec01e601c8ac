#!/usr/bin/env python3
"""Benchmark: Metropolis samples/s and SR-step seconds of the sampling + SR
optimisation path on one MI355X, 16x16 t-t'-U Hubbard model (Nsize 224,
8 QP planes, 16384 walkers; mvmc-rs_amd/lattice.py workload "square16").

A step is one SR step in the steady state: every walker restores its burn
configuration, runs (NVMCSample + 1) x NVMCInterval x Nsite proposals,
saves NVMCSample samples, the samples are measured and reduced, and the
SR-CG solve updates the parameters.  Untimed warm-up steps come first (the
first one thermalises with NVMCWarmUp x NVMCInterval x Nsite extra
proposals; --vmc-warmup sets NVMCWarmUp, default 2, which only changes the
length of that untimed step).

The whole run keeps to a wall budget (--budget seconds from process start,
default 470): at least one warm-up and two timed steps always run, further
requested warm-up and timed steps only while the measured step time says
they fit beside the CPU leg and the precision form (one warm-up and two
timed steps).  `steps` / `warmup` in the JSON line are the
steps actually run; `steps_requested` / `warmup_requested` what was asked.
The headline line (value, roofline, cpu_baseline) is printed and flushed as
soon as the FP64 and CPU legs are done; the secondary legs (direct SR, the
complex batched Pfaffian, the precision form) run afterwards within what is
left of the budget, each marked skipped when it does not fit, and a final
line repeats the headline with their results merged in.

Precision form (sampler inverse stored and updated in FP32): its first step
starts from the same walker states (seeds, placement) and parameters as the
FP64 run's first step, so that step is a replay on identical inputs: the
|dE/site| of its zvo row and the relative difference of its SR update
against FP64 are reported as `replay_step0`; later steps run from parameters
that already differ and are reported for speed.

value  = saved samples / second of the whole timed SR steps (all walkers).
roofline: dominant kernel = the sampling kernel k_block; algorithmic bytes
per launch = proposals * NQP * 2n*8 (inverse row + Slater row) + (accepted
moves + flush passes) * NQP * n^2*8 (one read of the inverse planes each) +
write passes * NQP * n^2*8 (the planes written back: every second accepted
hop, and each flush of the pending updates before an exchange move or at
the end of a block), divided by the launch's average device time (HIP
events on the kernel's stream).  traffic = HBM
bytes per launch from the profiled counter passes recorded under profiles/
(bytes per proposal and per accepted move), applied to this run's counts.
cpu_baseline: the FP64 oracle (same algorithm and operation order) on a
bounded walker subset of the same workload, one thread per walker on up to
16 host cores, timed over at least 10 s of steady SR steps.
"""
import time

T_START = time.time()

import argparse  # noqa: E402
import json  # noqa: E402
import os  # noqa: E402
import sys  # noqa: E402
import tempfile  # noqa: E402

import numpy as np  # noqa: E402

ROOT = os.path.dirname(os.path.abspath(__file__))
for sub in ("mvmc-rs_amd", "oracle"):
    sys.path.insert(0, os.path.join(ROOT, sub))

HBM_PEAK_GBS = 8000.0
FP64_PEAK_TFLOPS = 78.6
FP32_MFMA_PEAK_TFLOPS = 157.3
METRIC = "Metropolis samples/sec + SR-step s, 16x16 Hubbard Nsize 224; |dE/site| vs FP64"


def log(msg):
    print(f"[{time.time() - T_START:7.1f} s] {msg}", file=sys.stderr, flush=True)


def elapsed():
    return time.time() - T_START


def host_info():
    model = ""
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    try:
        avail = len(os.sched_getaffinity(0))
    except AttributeError:
        avail = os.cpu_count() or 1
    return {"nproc": os.cpu_count(), "cores_available": avail, "cpu_model": model}


def traffic_model(workload):
    """HBM traffic model of k_block recorded under profiles/ from the two PMC
    passes (FETCH_SIZE, WRITE_SIZE scaled by a device copy of known size):
    bytes per proposal and per accepted move fitted over the profiled
    launches, or the profiled traffic / algorithmic ratio."""
    path = os.path.join(ROOT, "profiles", f"k_block_{workload}_traffic.json")
    try:
        return json.load(open(path))
    except (OSError, ValueError):
        return None


def traffic_per_launch(model, props, accs, launches, alg_bytes):
    if model is None or launches <= 0:
        return None
    if "bytes_per_proposal" in model and "bytes_per_accept" in model:
        return (model["bytes_per_proposal"] * props + model["bytes_per_accept"] * accs) / launches
    if "traffic_over_algorithmic" in model:
        return model["traffic_over_algorithmic"] * alg_bytes / launches
    return None


def cpu_baseline(workload_dir, walkers, threads, params, min_seconds=10.0, max_steps=60):
    """Oracle leg: one untimed warm-up SR step (thermalisation), then steady SR
    steps until at least min_seconds of CPU work are timed.  Every step starts
    from `params` (the GPU run's parameters at the start of its timed steps),
    so both legs sample at the same acceptance rate; the oracle's own SR
    update of a step is computed and then replaced by them."""
    import oracle_py
    run = oracle_py.Run(os.path.join(workload_dir, "namelist.def"), os.path.join(workload_dir, "initial.def"),
                        walkers=walkers, threads=threads)
    t0 = time.time()
    run.set_params(*params)
    run.step(do_sr=True)
    t1 = time.time()
    steps = 0
    while steps < max_steps and (steps == 0 or time.time() - t1 < min_seconds):
        run.set_params(*params)
        run.step(do_sr=True)
        steps += 1
    t2 = time.time()
    samples = walkers * run.nsample * steps
    hi = host_info()
    return {"value": samples / (t2 - t1), "unit": "samples/s", "cores": threads, "kind": "port",
            "sample": f"{walkers} walkers of the same workload, {steps} steady-state SR steps after 1 warm-up step "
                      f"(warm-up {t1 - t0:.1f} s, timed {t2 - t1:.1f} s) from the GPU run's parameters at its first timed "
                      f"step, one thread per walker on {threads} cores "
                      f"of {hi['cores_available']} available ({hi['nproc']} in the machine), {hi['cpu_model']}",
            "host": hi}


def mixed_form(info, walkers, p_init, step0, rows64, times64, deadline):
    """The FP64 run repeated with the sampler's inverse in FP32 (same inputs
    and seeds): one thermalising step, then timed steps (as many as the FP64
    run timed, stopping at `deadline` after at least one).  The first step
    starts from the same walker states and parameters as the FP64 run's
    first step (`step0` = its zvo row and parameters after it, `p_init` the
    parameters before it): |dE/site| of that row and the relative difference
    of the SR update are a replay on identical inputs.  The timed steps
    start from parameters that already differ (the SR updates differ in the
    last bits), so their |dE/site| against the FP64 rows is of the size of
    the Monte Carlo error once any decision flips.  times64: the FP64 run's
    step times from its first step on; the speedup compares steps of the same
    index (the step time grows with the acceptance as the parameters
    improve, so only steps at the same point of the optimisation compare)."""
    import mvmc_hip
    nsite = info["nsite"]
    ens = mvmc_hip.Ensemble(info["namelist"], info["initial"], walkers=walkers)
    ens.set_precision("mixed")
    t = time.time()
    row0 = ens.opt_step()
    p1 = ens.params()
    log(f"mixed first step: {time.time() - t:.1f} s")
    nsample = ens.nsample
    replay = None
    if step0 is not None:
        row64, p1_64 = step0
        d64 = np.concatenate(p1_64) - np.concatenate(p_init)
        d32 = np.concatenate(p1) - np.concatenate(p_init)
        replay = {"abs_dE_per_site": abs(row0[0] - row64[0]) / nsite,
                  "energy_per_site_fp64": row64[0] / nsite, "energy_per_site_mixed": row0[0] / nsite,
                  "fp64_stat_err_per_site": float(np.sqrt(max(row64[2] - row64[0] ** 2, 0.0) /
                                                          (walkers * nsample))) / nsite,
                  "sr_update_rel_diff": float(np.linalg.norm(d32 - d64) / max(np.linalg.norm(d64), 1e-300)),
                  "what": "first SR step of both runs from identical walker states, seeds and parameters: "
                          "the zvo energy and the SR parameter update of the precision form against FP64"}
    tm0 = ens.timings()
    times, rows = [], []
    for s in range(max(1, min(2, len(times64) - 1))):
        if s >= 1 and time.time() + max(times) > deadline:
            break
        t = time.time()
        rows.append(ens.opt_step())
        times.append(time.time() - t)
        log(f"mixed step {s}: E/site = {rows[-1][0] / nsite:.8f}  {times[-1]:.2f} s")
    tm1 = ens.timings()
    total = sum(times)
    ens.close()
    k = len(times)
    de = [abs(r[0] - r64[0]) / nsite for r, r64 in zip(rows, rows64[1:])]
    se = [float(np.sqrt(max(r64[2] - r64[0] ** 2, 0.0) / (walkers * nsample))) / nsite for r64 in rows64[1:k + 1]]
    t64 = times64[1:k + 1]
    return {"value": walkers * nsample * k / total, "unit": "samples/s", "steps": k,
            "speedup_vs_fp64": sum(t64) / total if len(t64) == k else None,
            "fp64_same_steps_s": t64, "sr_step_s": total / k, "sr_step_s_each": times,
            "sampling_s_per_step": (tm1["sample_s"] - tm0["sample_s"]) / k,
            "sr_s_per_step": (tm1["sr_s"] - tm0["sr_s"]) / k,
            "replay_step0": replay,
            "energy_per_site": [r[0] / nsite for r in rows], "abs_dE_per_site_vs_fp64_steps": de,
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
                     "frac_fp64": 8.0 * n ** 3 * count / (ms / 1e3) / 1e12 / FP64_PEAK_TFLOPS}
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
    ap.add_argument("--steps", type=int, default=2, help="timed SR steps requested (at least 2 always run)")
    ap.add_argument("--warmup", type=int, default=1, help="untimed SR steps requested (at least 1 always runs)")
    ap.add_argument("--budget", type=float, default=470.0,
                    help="wall seconds from process start for the whole run; optional steps and legs that do "
                         "not fit are left out")
    ap.add_argument("--workload", default="square16")
    ap.add_argument("--vmc-warmup", type=int, default=2,
                    help="NVMCWarmUp of the generated workload: thermalisation sweeps of the first (untimed) "
                         "SR step only; steady-state steps do not depend on it")
    ap.add_argument("--walkers", type=int, default=0, help="override (not a valid headline number)")
    ap.add_argument("--cpu-threads", type=int, default=0,
                    help="oracle threads (one walker each); default: the cores available, at most 16")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-direct", action="store_true", help="skip the direct-SR measurement")
    ap.add_argument("--no-c128", action="store_true", help="skip the complex batched Pfaffian measurement")
    ap.add_argument("--c128-only", action="store_true",
                    help="run only the complex batched Pfaffian measurement and print its JSON (profiling aid; "
                         "not the benchmark line)")
    ap.add_argument("--no-mixed", action="store_true",
                    help="skip the precision-form run (sampler inverse in FP32) reported beside the FP64 line")
    ap.add_argument("--launch-log", action="store_true",
                    help="print every k_block launch of the timed steps (device ms, proposals, accepted, walkers) "
                         "to stderr, for the per-launch traffic fit of a profiled run")
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
    # torch ships its own HIP runtime under the same sonames: it has to bring
    # the device up before this library's first HIP call
    import torch
    torch.cuda.init()
    import lattice
    import mvmc_hip
    if args.copy_probe_mib > 0:
        nel = args.copy_probe_mib * (1 << 20) // 8
        src = torch.ones(nel, dtype=torch.float64, device="cuda")
        dst = torch.empty_like(src)
        dst.copy_(src)
        torch.cuda.synchronize()
        log(f"copy probe: {args.copy_probe_mib} MiB read + {args.copy_probe_mib} MiB written")
        del src, dst
    hi = host_info()
    cpu_threads = args.cpu_threads or max(1, min(16, hi["cores_available"]))
    cpu_reserve = 0.0 if args.no_cpu else 45.0     # oracle warm-up step + >= 10 s timed + set-up
    margin = 15.0

    def mixed_reserve(first, est):
        # precision form: set-up, its first (thermalising) step and two timed steps
        return 0.0 if args.no_mixed else 15.0 + 0.95 * first + 2.0 * est
    deadline = T_START + args.budget

    wl_dir = os.path.join(tempfile.gettempdir(), f"mvmc_bench_{args.workload}")
    info = lattice.write_workload(args.workload, wl_dir, sr_steps=max(args.steps, 2) + max(args.warmup, 1),
                                  warmup=args.vmc_warmup)
    walkers = args.walkers or info["walkers"]
    log(f"workload {args.workload}: Nsite={info['nsite']} Nsize={info['nsize']} NQP={info['nqp']} "
        f"NPara={info['npara']} walkers={walkers}")
    t = time.time()
    ens = mvmc_hip.Ensemble(info["namelist"], info["initial"], walkers=walkers)
    log(f"ensemble created in {time.time() - t:.1f} s (grid {ens.grid}, LDS {ens.lds_sampler} B)")

    # ---- FP64 leg: warm-up steps, then timed steps, within the budget
    warm_times, warm_rows = [], []
    p_init = ens.params()
    step0 = None
    for s in range(max(args.warmup, 1)):
        est = warm_times[-1] if warm_times else 0.0
        if s >= 1 and time.time() + est * 3 + cpu_reserve + mixed_reserve(warm_times[0], est) + margin > deadline:
            log(f"warm-up stops after {s} steps (budget)")
            break
        t = time.time()
        row = ens.opt_step()
        warm_times.append(time.time() - t)
        warm_rows.append(row)
        if s == 0:
            step0 = (row, ens.params())
        log(f"warmup step {s}: E/site = {row[0] / info['nsite']:.8f}  ({warm_times[-1]:.1f} s)")
    first_step_s = warm_times[0]
    params_at_timed = ens.params()
    tm0 = ens.timings()
    step_times, rows, kernel, launch_logs = [], [], [], []
    for s in range(max(args.steps, 2)):
        est = step_times[-1] if step_times else (warm_times[-1] if len(warm_times) > 1 else first_step_s)
        if s >= 2 and time.time() + est + cpu_reserve + mixed_reserve(first_step_s, est) + margin > deadline:
            log(f"timed steps stop after {s} (budget)")
            break
        before = ens.timings()
        p_last = ens.params()
        t = time.time()
        row = ens.opt_step()
        dt = time.time() - t
        after = ens.timings()
        step_times.append(dt)
        rows.append(row)
        kernel.append((after["block_ms"], after["block_launches"], after["proposals"] - before["proposals"],
                       after["accepted"] - before["accepted"], after["block_write_passes"],
                       after["block_flush_passes"]))
        ll = ens.launch_log()
        launch_logs.append(ll)
        log(f"step {s}: E/site = {row[0] / info['nsite']:.8f}  {dt:.2f} s  (sampling {after['sample_s'] - before['sample_s']:.2f} s, "
            f"measure {after['measure_s'] - before['measure_s']:.2f} s, SR {after['sr_s'] - before['sr_s']:.2f} s; "
            f"k_block launches {len(ll)}, walkers per launch {[int(x) for x in ll[:, 3]]})")
        if args.launch_log:
            for i, r in enumerate(ll):
                log(f"launch_log step {s} launch {i}: ms {r[0]:.3f} proposals {int(r[1])} accepted {int(r[2])} "
                    f"walkers {int(r[3])}")
    tm1 = ens.timings()
    nsteps = len(step_times)
    total = sum(step_times)
    samples = walkers * ens.nsample * nsteps
    n, nqp = ens.n, ens.nqp
    # roofline of the sampling kernel over the timed steps' launches
    launches = sum(int(k[1]) for k in kernel)
    block_ms = sum(k[0] for k in kernel)
    props = sum(k[2] for k in kernel)
    accs = sum(k[3] for k in kernel)
    wpasses = sum(k[4] for k in kernel)
    fpasses = sum(k[5] for k in kernel)
    bytes_total = props * nqp * 2 * n * 8 + (accs + fpasses + wpasses) * nqp * n * n * 8
    avg_launch_s = block_ms / 1e3 / max(launches, 1)
    achieved = bytes_total / max(launches, 1) / avg_launch_s / 1e9
    tmodel = traffic_model(args.workload)
    out = {
        "metric": METRIC,
        "value": samples / total,
        "unit": "samples/s",
        "n_gpus": 1,
        "steps": nsteps,
        "warmup": len(warm_times),
        "steps_requested": args.steps,
        "warmup_requested": args.warmup,
        "ms_per_step": 1e3 * total / nsteps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": f"{args.workload}: {info['nsite']}-site t-t'-U Hubbard, Nsize {info['nsize']}, "
                               f"NQP {nqp}, NPara {info['npara']}, {walkers} walkers, 1 sample/walker, "
                               f"16 sweeps/sample, SR-CG",
                   "walkers": walkers, "nsize": n, "nqp": nqp, "npara": info["npara"], "parallelism": "walkers"},
        "sr_step_s": total / nsteps,
        "sr_step_s_each": step_times,
        "first_warmup_step_s": first_step_s,
        "sampling_s_per_step": (tm1["sample_s"] - tm0["sample_s"]) / nsteps,
        "measure_s_per_step": (tm1["measure_s"] - tm0["measure_s"]) / nsteps,
        "sr_cg_s_per_step": (tm1["sr_s"] - tm0["sr_s"]) / nsteps,
        "proposals_per_s": (tm1["proposals"] - tm0["proposals"]) / total,
        "acceptance": accs / max(props, 1),
        "energy_per_site": [r[0] / info["nsite"] for r in rows],
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS,
                     "traffic": traffic_per_launch(tmodel, props, accs, launches, bytes_total),
                     "kernel": "k_block", "avg_launch_ms": avg_launch_s * 1e3, "launches": launches,
                     "algorithmic_bytes_per_launch": bytes_total / max(launches, 1),
                     "accepted_moves": accs, "write_passes": wpasses, "flush_passes": fpasses,
                     # the same launches priced at the write-every-hop byte
                     # count (one read + one write of the planes per accepted
                     # move), for comparison with that form only
                     "write_every_hop_equivalent_GBs": (props * nqp * 2 * n * 8 + accs * nqp * 2 * n * n * 8)
                     / max(launches, 1) / avg_launch_s / 1e9,
                     "walkers_per_launch": [[int(x) for x in ll[:, 3]] for ll in launch_logs]},
    }
    if tmodel is not None:
        out["roofline"]["traffic_model"] = {k: tmodel[k] for k in ("bytes_per_proposal", "bytes_per_accept",
                                                                   "traffic_over_algorithmic", "source")
                                            if k in tmodel}
    # the other rows' kernels in the last timed step, against their algorithmic
    # bytes (HBM 8 TB/s) and flops (FP64 vector 78.6 TFLOP/s)
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
                        "frac_fp64": cnt * nqp * plane_f / (ms / 1e3) / 1e12 / FP64_PEAK_TFLOPS}
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
    variants = ["3 workgroups/CU, 8-column close batches", "2 workgroups/CU, 16-column close batches",
                "register-resident, one 1024-thread workgroup per CU"]
    rv = int(tl["rebuild_variant"])
    out["autotuned"] = {"rebuild_variant": variants[rv] if 0 <= rv < len(variants) else "not decided",
                        "rebuild_ms_per_walker_measured": list(tl["rebuild_ms_per_walker"])}
    out["secondary_rooflines"] = sec

    # ---- CPU leg (before the headline line, which carries it)
    if not args.no_cpu:
        try:
            out["cpu_baseline"] = cpu_baseline(wl_dir, cpu_threads, cpu_threads, params_at_timed)
        except Exception as e:  # the baseline is reported, never the target
            out["cpu_baseline"] = {"value": None, "error": str(e)}
        log(f"cpu baseline: {out['cpu_baseline'].get('value')} samples/s")
    out["wall_s_at_headline"] = elapsed()
    print(json.dumps(out), flush=True)

    # ---- secondary legs, each only if it fits in what is left of the budget
    skipped = {}
    if not args.no_mixed:
        # precision form (1): the same workload and seeds with the sampler's
        # resident inverse in FP32
        need = mixed_reserve(first_step_s, total / nsteps) + margin
        if time.time() + need < deadline:
            out["precision_forms"] = {"mixed_fp32_inverse": mixed_form(
                info, walkers, p_init, step0, warm_rows + rows, warm_times + step_times, deadline - margin)}
        else:
            skipped["precision_form"] = "budget"
    if time.time() + 20 + margin < deadline:
        # derivative form (3): VJP of the last measurement by the adjoint
        # kernels (no O rows) and over the stored O rows
        rng = np.random.default_rng(0)
        cot = rng.standard_normal(walkers * ens.nsample)
        ens.set_params(*p_last)   # the parameters the last measurement's O rows were formed with
        t = time.time()
        g_rows = ens.vjp(cot)
        t_rows = time.time() - t
        g_adj = ens.vjp_adjoint(cot)   # first call: accumulators and orbital map
        t = time.time()
        g_adj = ens.vjp_adjoint(cot)
        t_adj = time.time() - t
        ta = ens.timings()
        scale = float(np.max(np.abs(g_rows)))
        sec["derivative form (3): adjoint VJP"] = {
            "samples": walkers * ens.nsample, "parameters": ens.size - 1,
            "adjoint_wall_ms": 1e3 * t_adj, "adjoint_device_ms": ta["adjoint_vjp_ms"],
            "of_which_plane_rebuild_ms": ta["adjoint_vjp_rebuild_ms"],
            "stored_rows_vjp_wall_ms": 1e3 * t_rows,
            "max_abs_diff_vs_rows_over_max": float(np.max(np.abs(g_adj - g_rows))) / max(scale, 1e-300)}
    else:
        skipped["derivative_form"] = "budget"
    if not args.no_direct:
        if time.time() + 30 + margin < deadline:
            # direct SR (NSRCG 0 path) on the same measured operands: Gram by
            # DGEMM on the matrix cores + Cholesky
            ens.sr_direct()   # first call: rocSOLVER initialisation
            ens.sr_direct()
            td = ens.timings()
            na_, rows_ = td["cg_na"], td["cg_rows"]
            gram_fl = 1.0 * na_ * na_ * rows_            # upper triangle of the Gram (SYRK count)
            fl = gram_fl + na_ ** 3 / 3.0 + 2.0 * na_ * na_   # + Cholesky and the two triangular solves
            leg = {"active_parameters": na_, "rows": rows_, "ms": td["cg_ms"],
                   "achieved_TFLOPs": fl / (td["cg_ms"] / 1e3) / 1e12,
                   "gram_mfma_fp64": {"ms": td["gram_ms"], "achieved_TFLOPs": gram_fl / (td["gram_ms"] / 1e3) / 1e12,
                                      "frac_fp64_mfma_peak": gram_fl / (td["gram_ms"] / 1e3) / 1e12 / FP64_PEAK_TFLOPS}}
            ens.set_precision("mixed")   # samples packed in FP32, FP32-MFMA Gram
            ens.sr_direct()
            tf = ens.timings()
            leg["gram_mfma_fp32"] = {"ms": tf["gram_ms"], "achieved_TFLOPs": gram_fl / (tf["gram_ms"] / 1e3) / 1e12,
                                     "frac_fp32_mfma_peak": gram_fl / (tf["gram_ms"] / 1e3) / 1e12 / FP32_MFMA_PEAK_TFLOPS,
                                     "solve_ms": tf["cg_ms"]}
            ens.set_precision("fp64")
            sec["direct SR (d, NSRCG 0): MFMA Gram + DPOTRF/DPOTRS"] = leg
        else:
            skipped["direct_sr"] = "budget"
    ens.close()
    del ens
    if not args.no_c128:
        if time.time() + 30 + margin < deadline:
            sec["k_pfinv_c128 (a, complex planes of the 12x12 flux lattice)"] = complex_pfinv()
            torch.cuda.empty_cache()
        else:
            skipped["k_pfinv_c128"] = "budget"
    if skipped:
        out["skipped_legs"] = skipped
    out["wall_s"] = elapsed()
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
