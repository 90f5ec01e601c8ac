/* C ABI of libmvmc_hip.so: the MI355X sampling + SR-optimisation path of
 * mvmc-rs behind the reference's backend seams.  Plain C types only; device
 * buffers are caller-owned device pointers where noted; every function
 * returns 0 on success and a negative code on error (message via
 * mvmc_last_error).  Handles are single-owner and not thread-safe, like the
 * reference's resident sampler (gpu/mvmc-gpu-cuda/src/interval/resident.rs:104-130).
 * There is no CPU fallback: without a usable GPU every call fails.
 */
#ifndef MVMC_HIP_H
#define MVMC_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Error codes. */
#define MVMC_OK 0
#define MVMC_E_INVALID_SHAPE (-1)   /* mvmc_gpu::Error::InvalidShape (crates/mvmc-gpu/src/lib.rs:83-97) */
#define MVMC_E_UNAVAILABLE (-2)     /* mvmc_gpu::Error::BackendUnavailable: no GPU / HIP error at init */
#define MVMC_E_BACKEND (-3)         /* mvmc_gpu::Error::Backend / StageError::Failed: HIP runtime error */
#define MVMC_E_INPUT (-4)           /* input files / parameters rejected */
#define MVMC_E_NUMERIC (-5)         /* sticky device status set (see mvmc_ens_status) */

const char* mvmc_last_error(void);
int mvmc_abi_version(void);

/* (a) Batched Pfaffian + inverse of real skew-symmetric planes.
 * Replaces mvmc_gpu::BatchedEngine::run_f64 (crates/mvmc-gpu/src/lib.rs:104-117,
 * semantics lib.rs:1-80): planes [n,n,count] column-major (plane p = q + NQP*b),
 * n even >= 2; outputs pf[count], inv [n,n,count] = the TRUE inverse A^{-1}
 * (mVMC's invM = -inv is applied by the caller), status[count]: 0 ok,
 * k > 0 zero pivot at 1-based row k, -1 non-finite Pfaffian (PlaneStatus
 * wire codes, lib.rs:64-80); the inverse of a failed plane is zero-filled, a
 * zero-pivot plane reports pf = 0 and a non-finite plane its computed pf
 * (lib.rs:12-15).  The results are BatchedEngine::run_f64's BatchOutput<f64>
 * = (pf, inv, status).  All pointers are device pointers; `stream` is a
 * hipStream_t (NULL = default stream).  Asynchronous on `stream`.
 * n <= 64 and 152 <= n <= 224: one plane per workgroup held in registers
 * (256 threads for n <= 64, 1024 above); other n: the panel-delayed kernel
 * (several planes per workgroup through a global work buffer), whichever
 * measured faster at that size; MVMC_PFINV_RR=0 selects the latter for
 * every n.  Both give the same bits. */
int mvmc_pfinv_f64(const double* planes, int n, int count, double* pf, double* inv, int* status, void* stream);

/* In-place form (crates/mvmc-core/src/stage_backend.rs:145-155): planes are
 * overwritten with the true inverse. */
int mvmc_pfinv_f64_inplace(double* planes, int n, int count, double* pf, int* status, void* stream);

/* No counterpart in the reference: mvmc_pfinv_f64 for callers that cannot
 * allocate device memory (code under HIP graph capture, serving loops,
 * harnesses that forbid hipMalloc).  Same arguments, results and status
 * codes as mvmc_pfinv_f64 (inv == planes factorises in place), but the
 * elimination's work buffer is a 128 MiB array in the library's code
 * object on each device instead of a per-call allocation, so no allocation
 * call is made.  n must be even, 2 <= n <= 1024 (MVMC_E_INVALID_SHAPE
 * otherwise).  Calls that overlap on one device share that array and race:
 * serialise them (one stream, or host-side ordering).  Asynchronous on
 * `stream`. */
int mvmc_pfinv_f64_noalloc(const double* planes, int n, int count, double* pf, double* inv, int* status,
                           void* stream);

/* (a) Batched Pfaffian + inverse of complex skew-symmetric planes.
 * Replaces mvmc_gpu::BatchedEngine::run_c64 (crates/mvmc-gpu/src/lib.rs:109-115;
 * CPU reference crates/mvmc-gpu/src/scalar.rs:146-170): planes [n,n,count]
 * column-major Complex64 (re, im interleaved; 16-byte aligned), n even,
 * 2 <= n <= 512; pf[2*count] (re, im), inv [n,n,count] complex = the TRUE
 * inverse, status[count] as mvmc_pfinv_f64 (zero-pivot planes: pf = 0 and a
 * zero inverse; non-finite planes: the computed pf and a zero inverse).
 * order 0: zsktf2 (crates/pfapack/src/ltl.rs:39-43, the run_c64 order);
 * order 1: zsktf2_c_compat (ltl.rs:54-59, the order of the production
 * calc_m_all_complex, crates/mvmc-core/src/pfaffian.rs:383-398).
 * inv == planes factorises in place.  Device pointers; asynchronous on
 * `stream` (NULL = default stream).  n <= 160: one plane per workgroup held
 * in registers; larger n: one plane per 256-lane workgroup through a global
 * work plane (MVMC_PFINV_RR=0 selects the latter for every n); same bits. */
int mvmc_pfinv_c128(const double* planes, int n, int count, int order, double* pf, double* inv, int* status,
                    void* stream);

/* Walker ensemble: W independent chains (= W ungrouped C ranks, walker w
 * seeded RndSeed + w; crates/mvmc-core/src/multichain.rs:88-96) of one
 * Expert-mode model, resident in HBM.  Covers the SamplingEngine/RealPfStage
 * seam (crates/mvmc-core/src/sampling/{engine.rs:134-168,stage.rs:56-128}),
 * the MeasurementStages seam (crates/mvmc-core/src/stage_backend.rs:418-438)
 * and the SrStages::cg_step seam (crates/mvmc-core/src/sr_backend.rs:128-292). */
typedef struct mvmc_ens mvmc_ens;

/* Load namelist.def (+ optional initial.def, NULL = none: C init_parameter
 * draws) and allocate the ensemble.  seed_override is used when
 * use_seed_override != 0, else RndSeed from modpara.def. */
int mvmc_ens_create(const char* namelist, const char* initial_def, int walkers, long long seed_override,
                    int use_seed_override, mvmc_ens** out);
int mvmc_ens_destroy(mvmc_ens* e);

/* dims[16]: Nsite, Ne, Nsize, NQP, NProj, NSlater, SROptSize, NVMCSample,
 * walkers, NExUpdatePath, NVMCInterval, NVMCWarmUp, persistent grid, lds bytes
 * (sampler), lds bytes (measure), reserved. */
int mvmc_ens_dims(mvmc_ens* e, int* dims16);

/* Variational parameters (host arrays): projection [NProj], Slater [NSlater]. */
int mvmc_ens_get_params(mvmc_ens* e, double* proj, double* slater);
int mvmc_ens_set_params(mvmc_ens* e, const double* proj, const double* slater);

/* Precision form of the sampler (mode 0, default: FP64 throughout; mode 1,
 * mixed: the resident inverse the Metropolis sweep streams and updates is
 * stored in FP32 (rounded to nearest on every store; ratios, updates and
 * Pfaffians are computed in FP64), while the periodic from-scratch
 * Pfaffian/inverse, the measurement, the SR-CG and every accumulator stay
 * FP64).  Takes effect at the next mvmc_ens_sample (which recomputes the
 * inverse from the configuration).  Mode 1 allocates W*NQP*n^2 floats once.
 * No reference counterpart as an entry point: it is the precision form of
 * the sweep kernel (updates.rs:588-638,1184-1224) that the reference's CUDA
 * backend offers only as a delayed-update FP64 variant. */
int mvmc_ens_set_precision(mvmc_ens* e, int mode);

/* One SR step's sampling on every walker (vmc_make_sample,
 * crates/mvmc-core/src/sampling/driver.rs:175-458): first call warm-up
 * (NVMCWarmUp + NVMCSample outer steps), later calls restore the burn state
 * (NVMCSample + 1 outer steps); NVMCInterval*Nsite proposals per outer step. */
int mvmc_ens_sample(mvmc_ens* e);

/* Measurement of the saved samples + rank-order reductions
 * (crates/mvmc-core/src/run.rs:2937-3616, 1863-1903; average.rs:18-63). */
int mvmc_ens_measure(mvmc_ens* e);

/* (c)+(3) on caller-supplied configurations: MeasurementStages::measure_sample
 * batched over samples (crates/mvmc-core/src/stage_backend.rs:418-438, the
 * per-sample work of run.rs:2937-3616).  ele_idx [count][n] host array in the
 * reference's EleIdx layout (electron mi + s*Ne -> site), count <= walkers x
 * NVMCSample; outputs (host) e_loc[count] (0 for an invalid sample),
 * o[count][1 + NPara] = [1, projection counts, Slater derivatives] (may be
 * NULL), valid[count] (0 when calc_m_all fails or the overlap vanishes; may be
 * NULL).  Uses the current parameters.  Overwrites the ensemble's saved
 * samples and inverse tables: call mvmc_ens_sample before the next
 * mvmc_ens_measure.  Out-of-range sites or doubly occupied spin-sites return
 * MVMC_E_INPUT. */
int mvmc_ens_measure_configs(mvmc_ens* e, int count, const int* ele_idx, double* e_loc, double* o, int* valid);

/* SR-CG solve on the reduced operands, parameter update and gauge sync
 * (crates/mvmc-core/src/sr_cg.rs:201-484, sync.rs:18-104).  cg_info gets the C
 * status (iterations or -(iter+1)). */
int mvmc_ens_sr_cg(mvmc_ens* e, long long* cg_info);
/* Direct SR on the reduced operands (crates/mvmc-core/src/sr.rs:60-160,
 * stochastic_opt_real; assembly sr_backend.rs:484-516): upper triangle of
 * the Gram of the packed active samples by the library's MFMA kernel
 * (FP64 MFMA; in the precision form, mvmc_ens_set_precision 1, the samples
 * packed in FP32 and FP32 MFMA with FP32 accumulation), S and g as the
 * reference assembles them, Cholesky by rocSOLVER DPOTRF/DPOTRS, parameter update and
 * gauge sync.  The path NSRCG 0 selects in mvmc_ens_opt_step; skipped gets
 * 1 when the factorisation fails or the solution is not finite (no update),
 * as the reference's cholesky_solve. */
int mvmc_ens_sr_direct(mvmc_ens* e, int* skipped);

/* sample + measure + SR-CG; energy_row6 gets the zvo_out.dat row
 * (Etot, Im Etot, Etot2, (Etot2-Etot^2)/Etot^2, Sztot, Sztot2). */
int mvmc_ens_opt_step(mvmc_ens* e, double* energy_row6);

/* Run-loop parameter output: output_opt_data (crates/mvmc-core/src/io.rs:561-727).
 * mvmc_ens_opt_step records (E, E^2, parameters after the SR update) for the
 * steps [NSROptItrStep - NSROptItrSmp, NSROptItrStep) of the run
 * (run.rs:1258-1262,1457-1460); this writes {CParaFileHead}_opt.dat into
 * out_dir (NULL or "" = current directory) in C's "% .18e": one snapshot as
 * (value, 0) pairs, more as (mean, 0, sample standard deviation) triples, plus
 * the {head}_{gutzwiller,jastrow,orbital}_opt.dat blocks of the means.
 * MVMC_E_INPUT when the window is incomplete or a file cannot be written. */
int mvmc_ens_write_opt_data(mvmc_ens* e, const char* out_dir);

/* Per-step output of the run loop: output_data (crates/mvmc-core/src/io.rs:
 * 142-232, text form).  With a directory set (NULL switches it off; ""
 * = current directory), every mvmc_ens_opt_step appends, after its
 * measurement and before its SR update (run.rs:1376-1385), the zvo_out row
 * "E 0  E^2 var Sz Sz^2" (C "% .18e", Sz columns "%.18e"; var by Julia's
 * complex division) to {CDataFileHead}_out.dat and the energies plus every
 * parameter ("re im 0.0" triples) to {CDataFileHead}_var.dat; the first step
 * of the handle truncates both. */
int mvmc_ens_set_output_dir(mvmc_ens* e, const char* out_dir);

/* Results for parity checks (host arrays). */
int mvmc_ens_get_energy(mvmc_ens* e, double* wc_etot_etot2_sz_sz2);
int mvmc_ens_get_operands(mvmc_ens* e, double* oo_2size, double* ho_size);
int mvmc_ens_get_o_store(mvmc_ens* e, int walker, double* o_nsample_size);

/* Differentiation form.  The measurement's O rows are the per-sample
 * adjoints O[r] = [1, d log|psi(x_r)| / d p] (projection counts and the
 * Slater derivative, slater_derivative.rs:289-428,490-548,
 * observables.rs:202-366), r = walker * NVMCSample + sample.
 * mvmc_ens_vjp: grad[k] = sum_r cot[r] * O[r][k] (r ascending; host arrays
 * cot[W*NVMCSample], grad[1+NPara]) over the last mvmc_ens_measure.
 * mvmc_ens_log_psi: log|psi(x_r)| = log|sum_q w_q Pf_q| + sum_p proj_p
 * cnt_p(x_r) of every saved sample at the current parameters (NaN for an
 * invalid sample), the function the O rows differentiate (finite-difference
 * checks).  It rebuilds the measurement planes, so call it between
 * mvmc_ens_measure and the next mvmc_ens_sample only for checks.  No
 * reference entry points: they expose the derivative the reference computes
 * inside its measurement. */
int mvmc_ens_vjp(mvmc_ens* e, const double* cot, double* grad);
/* mvmc_ens_vjp_adjoint: the same VJP by the adjoint kernels, without the O
 * rows: for every saved sample (its planes rebuilt as for the measurement)
 * each inverse element, scaled by cot[r] w_q Pf_q / ip (d log Pf(A) =
 * tr(A^-1 dA) / 2), is added at the Slater-table position it was gathered
 * from; the orbital map of update_slater_elm (slater_update.rs:17-126:
 * translations, orbital signs, spin-projection factors) is then applied once
 * to the sums instead of once per sample and QP point as the per-parameter
 * assembly does (slater_derivative.rs:289-356,490-548).  Same host arrays
 * as mvmc_ens_vjp; sums in a fixed order (per-workgroup accumulators over
 * walkers in order, folded in workgroup order), not that of mvmc_ens_vjp
 * (agreement to rounding).  Uses the saved configurations of the last
 * sampling call and the current parameters, and leaves the measurement
 * planes rebuilt (as mvmc_ens_log_psi).  Device memory: up to 4 GiB of
 * per-workgroup accumulators, allocated on the first call. */
int mvmc_ens_vjp_adjoint(mvmc_ens* e, const double* cot, double* grad);
int mvmc_ens_log_psi(mvmc_ens* e, double* log_psi);
int mvmc_ens_get_walker(mvmc_ens* e, int walker, int* saved_ele_idx, long long* counters4, uint32_t* rng625);
int mvmc_ens_get_tables(mvmc_ens* e, int walker, double* inv, double* pf);
int mvmc_ens_get_config(mvmc_ens* e, int walker, int* ele_idx, int* ele_cfg, int* ele_num, int* proj_cnt);
int mvmc_ens_get_cg(mvmc_ens* e, double* solution, int* n_active);
/* timings[26]: sampling s, measurement s, SR s (wall, summed over calls),
 * proposals, accepted, device ms of the last sampling call's k_block
 * launches, their number, device ms of its periodic-rebuild launches,
 * walkers those rebuilt; device ms of the last measurement call's rebuilds
 * and of its measurement launches, samples it measured; device ms of the
 * last SR call's CG solve, its iterations, active parameters, sample rows
 * (HIP events on the library's stream); the rebuild-kernel variant in use
 * (-1 while it is being measured; 0: three workgroups per CU, 8-column
 * panel-close batches; 1: two per CU, 16 columns; 2: register-resident,
 * one 1024-thread workgroup per CU, 64 < n <= 224) and the device ms per
 * rebuilt walker measured for variants 0, 1 and 2; device ms of the last
 * mvmc_ens_vjp_adjoint call and of its plane rebuilds; device ms of the
 * last direct SR's Gram kernel; passes of the last sampling call's k_block
 * launches that wrote a walker's inverse planes back (every second accepted
 * hop, and the flushes), and of them the flushes: passes that accept no
 * move (before an exchange move, at the end of a block); one reserved.  An
 * accepted hop reads the planes once and writes them back every second
 * time. */
int mvmc_ens_timings(mvmc_ens* e, double* timings26);
/* Device phase clocks of the sampling kernel, summed over walkers (ticks of
 * the 100 MHz constant clock): 0 candidate, 1 projection deltas + ratio,
 * 2 inverse/Slater row staging, 3 Pfaffian ratio dots, 4 Metropolis,
 * 5 accepted-hop inverse pass, 6 periodic rebuild, 7 exchange path,
 * 8 block end; then of the periodic rebuild's Pfaffian/inverse kernel:
 * 16 pivot search, 17 deferred scaling + trailing catch-up, 18 swaps,
 * 19 panel record + eager columns, 20 panel close, 25 Pfaffian, 21 inverse
 * of the unit triangle, 22 triangular solve, 23 triangular product and
 * skew fill, 24 permutations.  clocks32 receives 32 values.  Only libraries
 * built with MVMC_PHASE_CLOCKS collect them; others return
 * MVMC_E_UNAVAILABLE.  No reference counterpart (profiling aid, like the
 * reference's timer sections in crates/mvmc-core/src/timer.rs). */
int mvmc_ens_phase_clocks(mvmc_ens* e, long long* clocks32);
/* Per-launch record of the sampling kernel over the last mvmc_ens_sample
 * call: 4 doubles per k_block launch (device ms by HIP events on the
 * library's stream, proposals, accepted moves, walkers that advanced in the
 * launch).  Writes min(rows, max_rows) rows to out and returns the number of
 * launches.  No reference counterpart (profiling aid: the launch tail and
 * the per-launch traffic fit of the benchmark). */
int mvmc_ens_launch_log(mvmc_ens* e, double* out, int max_rows);
/* Sticky device status word (0 ok; 1 init retries exhausted, 2 bad input,
 * 3 rebuild failure, 4 non-finite, 5 CG failure). */
int mvmc_ens_status(mvmc_ens* e);

/* Ensemble group: several independent runs (Expert-mode models of uneven
 * Nsize, each with its own walkers, seeds RndSeed + w, parameters and SR)
 * stepped together, one persistent launch per kernel class (placement,
 * Metropolis sweep, periodic rebuild, measurement rebuild, measurement) over
 * all runs' walkers: the grid is sized for the largest run's LDS and cut
 * into per-run slot ranges in proportion to each run's work.  Each run
 * gives exactly the results of the same ensemble stepped alone (its walkers,
 * order of operations and walker-order reductions are its own).  No single
 * reference counterpart: it runs the per-run step of
 * crates/mvmc-core/src/multichain.rs:647-689 for many models at once.  The
 * runs share NVMCSample and use the FP64 sampler.  Runs are owned by the
 * group (do not destroy them); mvmc_group_run gives a run's handle for the
 * mvmc_ens_* getters. */
typedef struct mvmc_group mvmc_group;
int mvmc_group_create(int nruns, const char* const* namelists, const char* const* initial_defs, const int* walkers,
                      mvmc_group** out);
int mvmc_group_destroy(mvmc_group* g);
int mvmc_group_run(mvmc_group* g, int r, mvmc_ens** out);
/* Sweeps captured as a HIP graph of `rounds` (sweep + rebuild) launches per
 * replay, replayed until no walker waits (0: launches issued directly). */
int mvmc_group_set_graph(mvmc_group* g, int rounds);
/* Slot ranges re-cut after every step in proportion to each run's measured
 * busy time per launch class (on by default; results do not depend on it). */
int mvmc_group_set_autotune(mvmc_group* g, int on);
/* One SR step of every run; rows[nruns][6] = each run's zvo_out row. */
int mvmc_group_opt_step(mvmc_group* g, double* rows);
/* The sampling phase of one SR step of every run (vmc_make_sample). */
int mvmc_group_sample(mvmc_group* g);
/* timings10: sampling s, measurement s, SR s of the last step, device ms of
 * its sweep/rebuild rounds, rounds launched, graph replays, workgroups of the
 * sweep, rebuild and measurement launches, sampler LDS bytes. */
int mvmc_group_timings(mvmc_group* g, double* timings10);

/* The standalone SR entry points below (mvmc_sr_cg_apply, mvmc_sr_gram_real,
 * mvmc_sr_cholesky_solve, mvmc_sr_cg_solve) keep one device workspace or
 * rocBLAS handle per process: calls that overlap from several host threads
 * race; serialise them (the ensemble's own SR uses per-handle state). */

/* SrStages::cg_local_product (crates/mvmc-core/src/sr_backend.rs:518-598),
 * one product of the sampled operator on device operands:
 * z = inv_w O^T (O x) - (mean . x) mean + shift diag * x (sr_cg.rs:952-1004),
 * O = samples packed [walkers*nsample][na], in the blocked order of the
 * solver below (row dots, 32-row chunks, 16-chunk superchunks folded left).
 * mean = 0, inv_w = 1, shift = 0 give cg_local_product's bare O^T (O x). */
int mvmc_sr_cg_apply(int na, int walkers, int nsample, const double* pack, const double* mean, const double* diag,
                     const double* x, double inv_w, double shift, double* z, void* stream);

/* SrStages::gram_real (sr_backend.rs:219-226, C order c_order_gram_real
 * sr_backend.rs:305-333): out = O O^T, the full n x n column-major Gram of
 * the [n, samples] column-major store, by the library's FP64 MFMA Gram
 * kernel (64 x 64 tiles of the upper triangle mirrored to the lower; 4 rows
 * per MFMA, the MFMA partial sums in row order: not the serial sum of the
 * reference's BLAS, agreement to rounding).  Device pointers, asynchronous
 * on `stream`. */
int mvmc_sr_gram_real(int n, int samples, const double* store, double* out, void* stream);
/* Precision form of gram_real: the store in FP32, FP32 MFMA with FP32
 * accumulation, the Gram written as doubles (the direct SR of the precision
 * form uses the same kernel).  No reference counterpart (the reference's
 * Gram is FP64 only). */
int mvmc_sr_gram_real_f32(int n, int samples, const float* store, double* out, void* stream);

/* SrStages::cholesky_solve (sr_backend.rs:243-247, crates/mvmc-core/src/sr.rs:535):
 * solves the symmetric positive-definite S x = rhs in place (rhs becomes x,
 * S is overwritten by its factor) with rocSOLVER DPOTRF/DPOTRS.  *info: 0 ok,
 * k > 0 the factorisation's info (leading minor k not positive definite),
 * -1 non-finite solution (the reference's Err(())).  Device pointers for s and
 * rhs; synchronises `stream`. */
int mvmc_sr_cholesky_solve(int n, double* s, double* rhs, int* info, void* stream);

/* SrStages::assemble_s_g (crates/mvmc-core/src/sr_backend.rs:249-256, C order
 * COrderSr::assemble_s_g :484-516, C stcopt.c:69): S column-major [nm][nm],
 * S[si + sj nm] = oo[(pi+off) ld + (pj+off)] - oo[pi+off] oo[pj+off] (pi =
 * map[si]), the diagonal times 1 + sta_del; g[si] = -2 step_dt (ho[pi+off] -
 * ho[0] oo[pi+off]).  oo / ho are real arrays, or complex arrays (re, im
 * interleaved) read through their real parts when *_complex != 0 (RealView,
 * sr_backend.rs:53-70).  Every element as the reference forms it (bit-exact).
 * Device pointers; asynchronous on `stream`. */
int mvmc_sr_assemble_s_g(const double* oo, int oo_complex, const double* ho, int ho_complex, const int* map, int nm,
                         long long ld, int offset, double sta_del, double step_dt, double* s, double* g,
                         void* stream);

/* SrStages::gram_complex (sr_backend.rs:228-234, C order :386-420):
 * out[i + j n] = sum_s O[i,s] conj(O[j,s]) over the complex [samples][n] store
 * (O[i,s] at store[s n + i], re, im interleaved), samples summed in order and
 * each product formed as the reference's (bit-exact).  Device pointers;
 * asynchronous on `stream`. */
int mvmc_sr_gram_complex(int n, int samples, const double* store, double* out, void* stream);

/* SrStages::sr_s_g (sr_backend.rs:206-216, C order :436-460 over centered(),
 * stage_backend.rs:167-184): o column-major [ns][np] (sample k of component
 * p at o[k + p ns]), e, w [ns]; S[p + q np] = sum_k w_k d_kp d_kq, g[q] =
 * sum_k w_k d_kq de_k with d = o - <o>_w, de = e - <e>_w, k in order
 * (bit-exact).  Device pointers; takes ns*np + ns + 1 doubles of stream-
 * ordered scratch; asynchronous on `stream`. */
int mvmc_sr_s_g(const double* o, int ns, int np, const double* e, const double* w, double* s, double* g,
                void* stream);

/* SrStages::cg_step on caller-provided device operands: samples packed
 * [nrows][na] in walker order (nrows = W*nsample), mean/diag/g [na];
 * solves (S + shift diag) x = g with S = O^T O*inv_w - mean mean^T
 * (sr_cg.rs:611-880); variant 0 plain, 1 diagonal scaling. */
int mvmc_sr_cg_solve(int na, int walkers, int nsample, const double* pack, const double* mean, const double* diag,
                     const double* g, double inv_w, double shift, double tol, int max_it, int variant, double* x,
                     long long* status, void* stream);

#ifdef __cplusplus
}
#endif
#endif
