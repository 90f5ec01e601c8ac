// Host side of libmvmc_hip.so: the C ABI of include/mvmc_hip.h, the walker
// ensemble's device memory, and the SR-step driver that sequences the
// kernels of ensemble.hip (crates/mvmc-core/src/run.rs:1294-1475 for the
// step order; multichain.rs:647-689 for the walker = rank mapping).
#include <hip/hip_runtime.h>
#include <rocblas/rocblas.h>
#include <rocsolver/rocsolver.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <filesystem>
#include <map>
#include <mutex>
#include <set>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/mvmc_hip.h"
#include "../host/model.h"
#include "ensemble.hip"
#include "pfinv_c128_dev.h"
#include "pfinv_rr_c128.h"
#include "sr_stages.h"

using namespace mvmc_hip;

namespace {

thread_local std::string g_err;

struct Fail : std::runtime_error {
  int code;
  Fail(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

#define HIPCHK(x)                                                                                 \
  do {                                                                                            \
    hipError_t e_ = (x);                                                                          \
    if (e_ != hipSuccess)                                                                         \
      throw Fail(MVMC_E_BACKEND, std::string(#x) + ": " + hipGetErrorString(e_) + " (" + __FILE__ + \
                                     ":" + std::to_string(__LINE__) + ")");                       \
  } while (0)
#define BLASCHK(x)                                                                                \
  do {                                                                                            \
    rocblas_status s_ = (x);                                                                      \
    if (s_ != rocblas_status_success)                                                             \
      throw Fail(MVMC_E_BACKEND, std::string(#x) + ": rocBLAS status " + std::to_string((int)s_));  \
  } while (0)

double now_s() {
  return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

template <class F>
int guarded(F f) {
  try {
    f();
    return MVMC_OK;
  } catch (Fail& e) {
    g_err = e.what();
    return e.code;
  } catch (std::exception& e) {
    g_err = e.what();
    return MVMC_E_INPUT;
  }
}

void require_gpu() {
  int count = 0;
  hipError_t e = hipGetDeviceCount(&count);
  if (e != hipSuccess || count == 0) throw Fail(MVMC_E_UNAVAILABLE, "no HIP device available");
}

// Device scratch of the SR-CG solver.
struct CgWork {
  double *x = nullptr, *r = nullptr, *d = nullptr, *q = nullptr, *g = nullptr;
  double *cpart = nullptr, *spart = nullptr;  // chunk / superchunk partials of S x
  double* sc = nullptr;  // scalars: 0 delta, 2 alpha, 4 beta, 5 coef
  int* ctl = nullptr;    // 0 done, 1..2 status (int64)
  int cap_n = 0, cap_rows = 0;
  void ensure(int na, int rows) {
    if (na <= cap_n && rows <= cap_rows) return;
    release();
    cap_n = std::max(na, 1);
    cap_rows = std::max(rows, 1);
    const size_t nch = (size_t)(cap_rows + mvmc_hip::CG_RCH - 1) / mvmc_hip::CG_RCH;
    const size_t nsup = (nch + mvmc_hip::CG_SCH - 1) / mvmc_hip::CG_SCH;
    HIPCHK(hipMalloc(&x, sizeof(double) * cap_n));
    HIPCHK(hipMalloc(&r, sizeof(double) * cap_n));
    HIPCHK(hipMalloc(&d, sizeof(double) * cap_n));
    HIPCHK(hipMalloc(&q, sizeof(double) * cap_n));
    HIPCHK(hipMalloc(&g, sizeof(double) * cap_n));
    HIPCHK(hipMalloc(&cpart, sizeof(double) * cap_n * nch));
    HIPCHK(hipMalloc(&spart, sizeof(double) * cap_n * nsup));
    HIPCHK(hipMalloc(&sc, sizeof(double) * 8));
    HIPCHK(hipMalloc(&ctl, sizeof(int) * 4));
  }
  void release() {
    for (void* p : {(void*)x, (void*)r, (void*)d, (void*)q, (void*)g, (void*)cpart, (void*)spart, (void*)sc,
                    (void*)ctl})
      if (p) (void)hipFree(p);
    x = r = d = q = g = cpart = spart = sc = nullptr;
    ctl = nullptr;
    cap_n = cap_rows = 0;
  }
};

}  // namespace

namespace mvmc_hip {

__global__ void k_fill(double* p, int n, double v) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}

}  // namespace mvmc_hip

namespace {

inline int nblk(long long n, int b = 256) { return (int)((n + b - 1) / b); }

// out = S x (or g - S x when g is given) with the sampled operator
// (sr_cg.rs:952-1004) in the blocked reduction order.
template <class P>
void cg_apply_dev(int na, int rows, const P* pack, const double* mean, const double* diag, const double* xv,
                  const double* g, double* out, double inv_w, double shift, CgWork& wk, hipStream_t st) {
  const int nch = (rows + CG_RCH - 1) / CG_RCH;
  const int nsup = (nch + CG_SCH - 1) / CG_SCH;
  const int cols = (na + 255) / 256;
  const dim3 grid(nch + 1), blk(256);
  double* coef = wk.sc + 5;
  if (cols <= 4)
    hipLaunchKernelGGL((k_cg_chunk<4, P>), grid, blk, 0, st, pack, rows, na, xv, mean, wk.cpart, coef, wk.ctl);
  else if (cols <= 8)
    hipLaunchKernelGGL((k_cg_chunk<8, P>), grid, blk, 0, st, pack, rows, na, xv, mean, wk.cpart, coef, wk.ctl);
  else if (cols <= 16)
    hipLaunchKernelGGL((k_cg_chunk<16, P>), grid, blk, 0, st, pack, rows, na, xv, mean, wk.cpart, coef, wk.ctl);
  else if (cols <= CG_CMAX)
    hipLaunchKernelGGL((k_cg_chunk<CG_CMAX, P>), grid, blk, 0, st, pack, rows, na, xv, mean, wk.cpart, coef,
                       wk.ctl);
  else
    hipLaunchKernelGGL(k_cg_chunk_wide<P>, grid, blk, 0, st, pack, rows, na, xv, mean, wk.cpart, coef, wk.ctl);
  if (nsup > 0)
    hipLaunchKernelGGL(k_cg_fold1, dim3(nblk(na), nsup), blk, 0, st, wk.cpart, nch, na, wk.spart, wk.ctl);
  hipLaunchKernelGGL(k_cg_fold2, dim3(nblk(na)), blk, 0, st, wk.spart, nsup, na, xv, mean, diag, coef, inv_w, shift,
                     g, out, wk.ctl);
}

// The same product from the explicit Gram G = O^T O (FP32, full symmetric
// na x na; the precision form).
void cg_apply_gram(int na, const float* G, const double* mean, const double* diag, const double* xv,
                   const double* g, double* out, double inv_w, double shift, CgWork& wk, hipStream_t st) {
  double* coef = wk.sc + 5;
  hipLaunchKernelGGL(k_gram_coef, dim3(1), dim3(256), 0, st, mean, xv, na, coef, wk.ctl);
  hipLaunchKernelGGL(k_gram_apply, dim3((na + 3) / 4), dim3(256), 0, st, G, na, xv, mean, diag, coef, inv_w, shift,
                     g, out, wk.ctl);
}

// C 1.4.0 StochasticOptCG_Main / PreconCG_DiagScale (sr_cg.rs:700-880).
// Iterations run on the device as a HIP graph of one residual-refresh period
// (CG_PERIOD iterations, the refresh S x at its last one), captured once per
// solve and replayed: the kernels read the period's first iteration from
// ctl[CTL_ITER] (k_cg_advance moves it on), and a done flag set by the head /
// guard kernels turns the rest into no-ops.  The host reads the flag once per
// period.  The graph holds kernel nodes only.
constexpr int CG_PERIOD = 20;

template <class P>
long long cg_solve_dev(int na, int W, int NS, const P* pack, const double* mean, const double* diag,
                       const double* g, double inv_w, double shift, double tol, int max_it, bool diag_scale,
                       double* x_out, CgWork& wk, hipStream_t st_in, const float* gram = nullptr) {
  const int rows = W * NS;
  wk.ensure(na, rows);
  // graph capture needs a created stream: the null stream's work is ordered
  // with a default-flag stream by the runtime
  hipStream_t own = nullptr;
  if (!st_in) HIPCHK(hipStreamCreate(&own));
  hipStream_t st = st_in ? st_in : own;
  const double threshold = tol * tol * (double)na * (double)na;
  const double pfac = 1.0 + shift;
  const dim3 blk(256);
  // S x: one pass over the packed samples, or (gram set) one over their Gram
  auto apply = [&](const double* xv, const double* gv, double* out) {
    if (gram) cg_apply_gram(na, gram, mean, diag, xv, gv, out, inv_w, shift, wk, st);
    else cg_apply_dev(na, rows, pack, mean, diag, xv, gv, out, inv_w, shift, wk, st);
  };
  HIPCHK(hipMemsetAsync(wk.ctl, 0, sizeof(int) * 4, st));
  HIPCHK(hipMemcpyAsync(wk.g, g, sizeof(double) * na, hipMemcpyDeviceToDevice, st));
  hipLaunchKernelGGL(k_fill, dim3(nblk(na)), blk, 0, st, wk.x, na, 0.0);
  HIPCHK(hipMemcpyAsync(wk.r, g, sizeof(double) * na, hipMemcpyDeviceToDevice, st));
  if (diag_scale)
    apply(wk.x, wk.g, wk.r);
  else
    HIPCHK(hipMemcpyAsync(wk.d, g, sizeof(double) * na, hipMemcpyDeviceToDevice, st));
  hipLaunchKernelGGL(k_cg_init, dim3(1), blk, 0, st, na, wk.r, wk.d, diag, pfac, diag_scale ? 1 : 0, wk.sc, wk.ctl);
  HIPCHK(hipGetLastError());
  // one refresh period of iterations (kernel launches only)
  auto issue_period = [&]() {
    for (int j = 0; j < CG_PERIOD; ++j) {
      hipLaunchKernelGGL(k_cg_head, dim3(1), dim3(64), 0, st, j, max_it, threshold, wk.sc, wk.ctl);
      apply(wk.d, nullptr, wk.q);
      hipLaunchKernelGGL(k_cg_alpha, dim3(1), blk, 0, st, j, na, wk.d, wk.q, wk.sc, wk.ctl);
      const int refresh = (j == CG_PERIOD - 1);  // (iter + 1) % 20 == 0 with periods starting at multiples of 20
      hipLaunchKernelGGL(k_cg_update, dim3(nblk(na)), blk, 0, st, na, wk.sc, wk.ctl, wk.d, wk.x, wk.q, wk.r,
                         refresh);
      if (refresh) apply(wk.x, wk.g, wk.r);
      hipLaunchKernelGGL(k_cg_beta, dim3(1), blk, 0, st, j, na, wk.r, wk.q, diag, pfac, diag_scale ? 1 : 0, wk.sc,
                         wk.ctl);
      hipLaunchKernelGGL(k_cg_dir, dim3(nblk(na)), blk, 0, st, na, wk.sc, wk.ctl, diag_scale ? wk.q : wk.r, wk.d);
    }
    hipLaunchKernelGGL(k_cg_advance, dim3(1), dim3(64), 0, st, wk.ctl, CG_PERIOD);
  };
  // MVMC_CG_GRAPH=0 issues the same launches without a graph (for tools that
  // do not follow graph launches)
  static const bool use_graph = [] {
    const char* e = std::getenv("MVMC_CG_GRAPH");
    return !(e && e[0] == '0');
  }();
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  if (use_graph) {
    HIPCHK(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
    issue_period();
    HIPCHK(hipStreamEndCapture(st, &graph));
    HIPCHK(hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0));
  }
  long long status = 0;
  bool done = false;
  for (int it0 = 0; it0 <= max_it; it0 += CG_PERIOD) {
    if (use_graph) HIPCHK(hipGraphLaunch(exec, st));
    else issue_period();
    HIPCHK(hipGetLastError());
    int ctl[4];
    HIPCHK(hipMemcpyAsync(ctl, wk.ctl, sizeof(int) * 4, hipMemcpyDeviceToHost, st));
    HIPCHK(hipStreamSynchronize(st));
    if (ctl[0]) {
      status = (long long)(((unsigned long long)(unsigned)ctl[2] << 32) | (unsigned)ctl[1]);
      done = true;
      break;
    }
  }
  if (exec) (void)hipGraphExecDestroy(exec);
  if (graph) (void)hipGraphDestroy(graph);
  if (!done) {
    if (own) (void)hipStreamDestroy(own);
    throw Fail(MVMC_E_BACKEND, "SR-CG control flag not set at the iteration limit");
  }
  HIPCHK(hipMemcpyAsync(x_out, wk.x, sizeof(double) * na, hipMemcpyDeviceToDevice, st));
  HIPCHK(hipStreamSynchronize(st));
  if (own) HIPCHK(hipStreamDestroy(own));
  return status;
}

// A kernel's dynamic-LDS limit is a process-wide attribute while every
// ensemble (and batched call) needs its own size: the limit only grows, so
// handles of different models can live and run side by side.
void raise_lds_limit(const void* fn, size_t bytes) {
  static std::mutex mu;
  static std::map<const void*, size_t> cur;
  std::lock_guard<std::mutex> lk(mu);
  size_t& c = cur[fn];
  if (bytes <= c) return;
  HIPCHK(hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes));
  c = bytes;
}

// Work array of mvmc_pfinv_f64_noalloc, in the code object (no allocation
// call): every device the library runs on holds its own copy, 128 MiB, until
// the process ends.  Host code never uses the name `scratch` as an address:
// noalloc_scratch() looks the copy of the current device up once with
// hipGetSymbolAddress and keeps it per device.  Calls that overlap on one
// device share the array and race; contents persist between calls (the
// kernel writes every scratch element before it reads it, so nothing needs
// clearing).
constexpr size_t NOALLOC_SCRATCH_DOUBLES = (128ull << 20) / sizeof(double);
constexpr int NOALLOC_MAX_N = 1024;       // one work slot (4 n^2 doubles = 32 MiB) fits
constexpr int NOALLOC_MAX_SLOTS = 2048;   // workgroups of one call
static __device__ __align__(256) double scratch[NOALLOC_SCRATCH_DOUBLES];

double* noalloc_scratch() {
  static std::mutex mu;
  static std::map<int, double*> per_device;
  int dev = 0;
  HIPCHK(hipGetDevice(&dev));
  std::lock_guard<std::mutex> lk(mu);
  auto it = per_device.find(dev);
  if (it != per_device.end()) return it->second;
  void* p = nullptr;
  HIPCHK(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
  if (!p) throw Fail(MVMC_E_BACKEND, "hipGetSymbolAddress(scratch) returned a null address");
  per_device[dev] = (double*)p;
  return (double*)p;
}

// The register-resident Pfaffian + inverse (full form n <= 64, panel form
// 64 < n <= RR_NMAX) is a rebuild variant of the sampler (chosen by
// measurement) and serves the batched entry points where it measured faster
// than the panel-delayed kernel: n <= 64 and RR_BATCHED_MIN <= n <= RR_NMAX
// (MI355X, random planes: 1.46-1.65x at n = 16..64; 0.81x at n = 96, 0.97x
// at 144, 1.10x at 160, 1.18x at 176, 1.29x at 192 and 208, 1.36x at 224).  MVMC_PFINV_RR=0 selects the
// panel-delayed kernels throughout, MVMC_PFINV_RR=2 the register-resident
// ones for every 2 <= n <= RR_NMAX (same results either way).
constexpr int RR_BATCHED_MIN = 152;
int rr_mode() {
  const char* e = std::getenv("MVMC_PFINV_RR");
  return e ? std::atoi(e) : 1;
}
bool rr_enabled() { return rr_mode() != 0; }
bool rr_batched(int n) {
  if (!rr_enabled() || n < 2 || n > RR_NMAX) return false;
  return rr_mode() == 2 || n <= RR_NFULL || n >= RR_BATCHED_MIN;
}

// Shape of the register-resident kernels for order n: the full form (16 x 16
// threads, R = ceil(n / 16) <= 4) for n <= 64, the panel form (32 x 32,
// R = ceil(n / 32) <= 7) above.
struct RrCfg {
  int t = 0, r = 0;
  bool full = false;
};
RrCfg rr_cfg(int n) {
  RrCfg c;
  c.full = n <= RR_NFULL;
  c.t = c.full ? RR_TS : RR_T;
  c.r = (n + c.t - 1) / c.t;
  return c;
}

// Calls f.template operator()<T, R, FULL>() for the configuration's
// compiled instance.
template <class F>
void rr_dispatch(const RrCfg& c, F&& f) {
  if (c.full) {
    switch (c.r) {
      case 1: f.template operator()<RR_TS, 1, true>(); return;
      case 2: f.template operator()<RR_TS, 2, true>(); return;
      case 3: f.template operator()<RR_TS, 3, true>(); return;
      case 4: f.template operator()<RR_TS, 4, true>(); return;
    }
  } else {
    switch (c.r) {
      case 3: f.template operator()<RR_T, 3, false>(); return;
      case 4: f.template operator()<RR_T, 4, false>(); return;
      case 5: f.template operator()<RR_T, 5, false>(); return;
      case 6: f.template operator()<RR_T, 6, false>(); return;
      case 7: f.template operator()<RR_T, 7, false>(); return;
    }
  }
  throw Fail(MVMC_E_INVALID_SHAPE, "register-resident path covers 2 <= n <= 224");
}

// LDS limits of the rebuild kernels and their resident workgroups per CU.
int rr_rebuild_prepare(const RrCfg& c, size_t lds) {
  int occ = 1;
  rr_dispatch(c, [&]<int T, int R, bool FULL>() {
    raise_lds_limit((const void*)k_rebuild_rr<T, R, FULL>, lds);
    raise_lds_limit((const void*)k_meas_rebuild_rr<T, R, FULL>, lds);
    raise_lds_limit((const void*)k_begin_rr<T, R, FULL>, lds);
    HIPCHK(hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, k_rebuild_rr<T, R, FULL>, T * T, lds));
  });
  return std::max(occ, 1);
}

// Periodic rebuild (s < 0) or measurement rebuild of saved sample s with the
// register-resident kernels, one T x T workgroup per slot.
void rr_launch_rebuild(const RrCfg& c, int grid, size_t lds, hipStream_t st, const EnsView& v, int s) {
  rr_dispatch(c, [&]<int T, int R, bool FULL>() {
    if (s < 0)
      hipLaunchKernelGGL((k_rebuild_rr<T, R, FULL>), dim3(grid), dim3(T * T), lds, st, v);
    else
      hipLaunchKernelGGL((k_meas_rebuild_rr<T, R, FULL>), dim3(grid), dim3(T * T), lds, st, v, s);
  });
}

template <class T>
T* dalloc(std::vector<void*>& list, size_t count) {
  void* p = nullptr;
  if (count == 0) count = 1;
  HIPCHK(hipMalloc(&p, sizeof(T) * count));
  list.push_back(p);
  return (T*)p;
}

template <class T>
T* dupload(std::vector<void*>& list, const std::vector<T>& h, hipStream_t st) {
  T* p = dalloc<T>(list, h.size());
  if (!h.empty()) {
    HIPCHK(hipMemcpyAsync(p, h.data(), sizeof(T) * h.size(), hipMemcpyHostToDevice, st));
    HIPCHK(hipStreamSynchronize(st));
  }
  return p;
}

}  // namespace

struct mvmc_ens {
  // per-step output (output_data): directory, or empty for none
  std::string out_dir;
  bool out_enabled = false;
  // optimisation window of the last NSROptItrSmp steps (store_opt_data,
  // crates/mvmc-core/src/io.rs:21-48): energy, energy^2 and the parameters
  // after the step's SR update
  int opt_steps = 0;
  std::vector<std::vector<double>> opt_win;
  std::vector<double> opt_win_e, opt_win_e2;
  std::vector<char> opt_win_set;
  mvmc::Model m;
  EnsView v{};
  std::vector<void*> allocs;
  hipStream_t st = nullptr;
  int grid_samp = 0, grid_meas = 0, grid_rebuild = 0, slots = 0;
  // rebuild-kernel variants (0: 3 workgroups per CU, 8-column close batches;
  // 1: 2 per CU, 16 columns; 2: register-resident, one 1024-thread workgroup
  // per CU, for 64 < n <= RR_NMAX), tuned on the first well-filled launches:
  // rb_variant -1 while measuring, then the fastest one
  int grid_rb[3] = {0, 0, 0};
  int rb_nvar = 2;
  int rb_variant = -1;
  double rb_ms[3] = {0, 0, 0}, rb_walkers[3] = {0, 0, 0};
  int rb_samples[3] = {0, 0, 0};
  size_t lds_rebuild = 0, lds_rr = 0;
  double rebuild_ms = 0, rebuild_launches = 0;
  size_t lds_samp = 0, lds_meas = 0;
  double* d_slater = nullptr;
  double* d_proj = nullptr;
  double* d_dg = nullptr;
  double* d_dj = nullptr;
  double* d_out5 = nullptr;
  double *d_ho = nullptr, *d_mean = nullptr, *d_diag = nullptr;
  double* d_gscr = nullptr;
  int gstride = 1;
  void* d_pack = nullptr;  // packed active samples: double, or float for the precision form's CG
  size_t pack_cap = 0;      // bytes
  int* d_map = nullptr;
  double *d_cmean = nullptr, *d_cdiag = nullptr, *d_cg = nullptr, *d_cx = nullptr;
  CgWork cg;
  // host results
  double wc = 0, etot = 0, etot2 = 0, sztot = 0, sztot2 = 0;
  std::vector<double> oo, ho;
  std::vector<double> cg_x;
  long long cg_info = 0;
  double t_sample = 0, t_measure = 0, t_sr = 0, proposals = 0, accepted = 0;
  float block_ms = 0;
  int block_launches = 0;
  // passes of the last sample call's k_block launches over a walker's planes
  // that write them back, and those of them that accept no move (flushes)
  double block_write_passes = 0, block_flush_passes = 0;
  // per k_block launch of the last sample call: device ms, proposals,
  // accepted moves and walkers that advanced in it
  std::vector<double> launch_log;
  // precision form: 0 FP64 throughout; 1 the sampler's inverse stored in
  // FP32 (inv32), everything else FP64
  int precision = 0;
  float* inv32 = nullptr;
  double* d_rows = nullptr;  // [W*nsample] cotangents / log|psi| of the saved samples
  // direct SR (NSRCG 0): rocBLAS handle on the library's stream, S [na][na], Cholesky info
  rocblas_handle blas = nullptr;
  double* d_S = nullptr;
  size_t S_cap = 0;
  // precision form: FP32 Gram of the packed samples for the CG product
  void* d_gram = nullptr;
  size_t gram_cap = 0;
  int* d_info = nullptr;
  double* d_grad = nullptr;  // [size]
  // adjoint VJP: per-slot Slater-table cotangents [adj_slots][nqp][n2^2] and
  // constant + projection sums [adj_slots][1 + nproj]; the orbital map as
  // CSR (per orbital: table positions in ascending order and their weights)
  int adj_slots = 0;
  double* d_sbar = nullptr;
  double* d_gproj = nullptr;
  int* d_adj_start = nullptr;
  long long* d_adj_ent = nullptr;
  double* d_adj_wt = nullptr;
  // device times of the last calls (HIP events on the library's stream)
  double rebuilt = 0, meas_rebuild_ms = 0, measure_ms = 0, measured = 0, cg_ms = 0, cg_iter = 0, cg_na = 0,
         cg_rows = 0, adj_ms = 0, adj_rebuild_ms = 0, gram_ms = 0;
  ~mvmc_ens() {
    cg.release();
    if (d_S) (void)hipFree(d_S);
    if (d_gram) (void)hipFree(d_gram);
    if (blas) (void)rocblas_destroy_handle(blas);
    for (void* p : allocs) (void)hipFree(p);
    if (st) (void)hipStreamDestroy(st);
  }
  void upload_tables() {
    std::vector<double> S, g, j;
    mvmc::build_slater_table(m, S);
    mvmc::direct_projection_tables(m, g, j);
    HIPCHK(hipMemcpyAsync(d_slater, S.data(), sizeof(double) * S.size(), hipMemcpyHostToDevice, st));
    if (!m.proj.empty())
      HIPCHK(hipMemcpyAsync(d_proj, m.proj.data(), sizeof(double) * m.proj.size(), hipMemcpyHostToDevice, st));
    HIPCHK(hipMemcpyAsync(d_dg, g.data(), sizeof(double) * g.size(), hipMemcpyHostToDevice, st));
    HIPCHK(hipMemcpyAsync(d_dj, j.data(), sizeof(double) * j.size(), hipMemcpyHostToDevice, st));
    HIPCHK(hipStreamSynchronize(st));
  }
  // Device-to-host copy ordered on the handle's own (non-blocking) stream:
  // never on the legacy stream, which would wait on (and break) another
  // handle's stream capture in the same process.
  void d2h(void* dst, const void* src, size_t bytes) {
    HIPCHK(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, st));
    HIPCHK(hipStreamSynchronize(st));
  }
  void check_status() {
    int s = 0;
    d2h(&s, v.status, sizeof(int));
    if (s == ST_BAD_INPUT || s == ST_INIT_RETRY)
      throw Fail(MVMC_E_NUMERIC, "device status " + std::to_string(s));
  }
};

namespace {

void create_impl(mvmc_ens* E, const char* namelist, const char* initial_def, int walkers, long long seed_override,
                 int use_seed) {
  require_gpu();
  if (walkers < 1) throw Fail(MVMC_E_INVALID_SHAPE, "walkers must be >= 1");
  mvmc::Model& m = E->m;
  m = mvmc::load_namelist(namelist);
  if (m.mp.vmc_calc_mode != 0) throw Fail(MVMC_E_INPUT, "only NVMCCalMode 0 (optimisation) is supported");
  if (m.ne() < 1 || m.ne() > m.nsite()) throw Fail(MVMC_E_INPUT, "Ne must be in [1, Nsite]");
  if (m.nqp() < 1) throw Fail(MVMC_E_INPUT, "no quantum-projection planes (NMPTrans/TransSym)");
  {
    // every translated site must be a site (the kernels index tables with it)
    std::vector<int> ts, tg;
    mvmc::translation_maps(m, ts, tg);
    for (int t : ts)
      if (t < 0 || t >= m.nsite()) throw Fail(MVMC_E_INPUT, "TransSym maps a site outside [0, Nsite)");
    // placement loops terminate: local spins fit, itinerant electrons find free sites
    int nloc = 0;
    for (int f : m.loc_spn) nloc += (f == 1);
    if (nloc > 2 * m.ne()) throw Fail(MVMC_E_INPUT, "more local spins than electrons");
    if (nloc < 2 * m.ne() && m.nsite() - nloc < m.ne())
      throw Fail(MVMC_E_INPUT, "not enough itinerant sites for the itinerant electrons");
    if (m.mp.nex_update_path < 0 || m.mp.nex_update_path > 2)
      throw Fail(MVMC_E_INPUT, "NExUpdatePath must be 0, 1 or 2 for the normal real sampler");
  }
  const long long seed = use_seed ? seed_override : m.mp.rnd_seed;
  // init_parameter (parameter_init.rs:63-160) on walker 0's stream; every
  // walker consumes the same number of draws before sampling.
  int n_active = 0;
  {
    SfmtLds r;
    sfmt_seed_host(r.w, (uint32_t)seed);
    r.idx = 624;
    for (int i = 0; i < m.n_slater; ++i) {
      const bool active = m.opt_flags[2 * (m.n_proj() + i)] > 0;
      m.slater[i] = 0.0;
      if (active) {
        m.slater[i] = 2.0 * (sfmt_real2(r) - 0.5);
        ++n_active;
      }
    }
    for (double& p : m.proj) p = 0.0;
  }
  if (initial_def && *initial_def) {
    if (!mvmc::load_initial_def(m, initial_def)) throw Fail(MVMC_E_INPUT, std::string("cannot read ") + initial_def);
  }
  mvmc::sync_modified_parameter(m, true);

  HIPCHK(hipStreamCreateWithFlags(&E->st, hipStreamNonBlocking));
  EnsView& v = E->v;
  auto& A = E->allocs;
  v.W = walkers;
  v.n = m.nsize(); v.ne = m.ne(); v.ns = m.nsite(); v.n2 = m.nsite2();
  v.nqp = m.nqp(); v.nleg = m.n_sp_gl(); v.nmp = m.n_mp_trans();
  v.nproj = m.n_proj(); v.ngutz = m.n_gutz; v.nslater = m.n_slater; v.size = m.sr_opt_size();
  v.nsample = m.mp.nvmc_sample; v.warmup = m.mp.nvmc_warmup; v.interval = m.mp.nvmc_interval;
  v.ex_path = m.mp.nex_update_path;
  if (v.nsample < 1) throw Fail(MVMC_E_INPUT, "NVMCSample must be >= 1");
  // orbital rows injective -> ordered parallel scatter in the Slater derivative
  int inj = 1;
  for (int a = 0; a < v.ns && inj; ++a) {
    std::set<int> seen;
    for (int b = 0; b < v.ns; ++b) {
      int o = m.orb_idx[(size_t)a * v.ns + b];
      if (o >= 0 && !seen.insert(o).second) { inj = 0; break; }
    }
  }
  v.orb_injective = inj;
  v.exp_cr = 1;
  // model tables
  E->d_slater = dalloc<double>(A, (size_t)v.nqp * v.n2 * v.n2);
  v.slater = E->d_slater;
  v.qpw = dupload(A, m.qp_full_weight, E->st);
  E->d_proj = dalloc<double>(A, std::max(v.nproj, 1));
  v.proj = E->d_proj;
  v.gutz_idx = dupload(A, m.gutz_idx, E->st);
  v.jast_idx = dupload(A, m.jast_idx, E->st);
  v.loc_spn = dupload(A, m.loc_spn, E->st);
  {
    std::vector<int> t;
    std::vector<double> tv;
    for (auto& x : m.transfer) {
      t.insert(t.end(), {x.site1, x.spin1, x.site2, x.spin2});
      tv.push_back(x.re);
    }
    v.ntrans = (int)m.transfer.size();
    v.trans = dupload(A, t, E->st);
    v.trans_v = dupload(A, tv, E->st);
    std::vector<int> is;
    std::vector<double> iv;
    for (auto& x : m.coulomb_intra) { is.push_back(x.first); iv.push_back(x.second); }
    v.nintra = (int)is.size();
    v.intra_site = dupload(A, is, E->st);
    v.intra_v = dupload(A, iv, E->st);
    auto pairs = [&](const std::vector<mvmc::PairTerm>& P, const int*& site, const double*& val, int& cnt) {
      std::vector<int> s;
      std::vector<double> vv;
      for (auto& x : P) { s.push_back(x.site1); s.push_back(x.site2); vv.push_back(x.value); }
      cnt = (int)P.size();
      site = dupload(A, s, E->st);
      val = dupload(A, vv, E->st);
    };
    pairs(m.coulomb_inter, v.inter_site, v.inter_v, v.ninter);
    pairs(m.hund, v.hund_site, v.hund_v, v.nhund);
    pairs(m.exchange, v.exch_site, v.exch_v, v.nexch);
  }
  E->d_dg = dalloc<double>(A, v.ns);
  E->d_dj = dalloc<double>(A, (size_t)v.ns * v.ns);
  v.dproj_g = E->d_dg;
  v.dproj_j = E->d_dj;
  {
    std::vector<int> ts, tg;
    mvmc::translation_maps(m, ts, tg);
    v.tsite = dupload(A, ts, E->st);
    v.tsgn = dupload(A, tg, E->st);
  }
  v.orb_idx = dupload(A, m.orb_idx, E->st);
  v.orb_sgn = dupload(A, m.orb_sgn, E->st);
  v.spgl_cs = dupload(A, m.spgl_cos_sin, E->st);
  v.spgl_cc = dupload(A, m.spgl_cos_cos, E->st);
  v.spgl_ss = dupload(A, m.spgl_sin_sin, E->st);
  // walker state
  const size_t W = (size_t)walkers;
  v.rng = dalloc<uint32_t>(A, W * 625);
  v.ele_idx = dalloc<int>(A, W * v.n);
  v.ele_cfg = dalloc<int>(A, W * v.n2);
  v.ele_num = dalloc<int>(A, W * v.n2);
  v.proj_cnt = dalloc<int>(A, W * std::max(v.nproj, 1));
  v.inv = dalloc<double>(A, W * v.nqp * v.n * v.n);
  v.pf = dalloc<double>(A, W * v.nqp);
  v.log_ip = dalloc<double>(A, W);
  v.wstate = dalloc<int>(A, W * 8);
  v.counter = dalloc<long long>(A, W * 4);
  const size_t WS = W * v.nsample;
  v.s_idx = dalloc<int>(A, WS * v.n);
  v.s_cfg = dalloc<int>(A, WS * v.n2);
  v.s_num = dalloc<int>(A, WS * v.n2);
  v.s_proj = dalloc<int>(A, WS * std::max(v.nproj, 1));
  v.e_loc = dalloc<double>(A, WS);
  v.s_valid = dalloc<int>(A, WS);
  v.sz = dalloc<double>(A, WS);
  v.o_store = dalloc<double>(A, WS * v.size);
  v.status = dalloc<int>(A, 1);
  HIPCHK(hipMemsetAsync(v.status, 0, sizeof(int), E->st));
  HIPCHK(hipMemsetAsync(v.wstate, 0, sizeof(int) * W * 8, E->st));
  HIPCHK(hipMemsetAsync(v.counter, 0, sizeof(long long) * W * 4, E->st));
  HIPCHK(hipMemsetAsync(v.ele_idx, 0, sizeof(int) * W * v.n, E->st));
  HIPCHK(hipMemsetAsync(v.ele_cfg, 0, sizeof(int) * W * v.n2, E->st));
  HIPCHK(hipMemsetAsync(v.ele_num, 0, sizeof(int) * W * v.n2, E->st));
  HIPCHK(hipMemsetAsync(v.proj_cnt, 0, sizeof(int) * W * std::max(v.nproj, 1), E->st));
  // persistent grids sized to the resident capacity
  E->lds_samp = lds_bytes_sampler(v, 256);
  E->lds_meas = lds_bytes_measure(v, 256);
  int dev = 0, ncu = 0, maxlds = 0;
  HIPCHK(hipGetDevice(&dev));
  HIPCHK(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev));
  HIPCHK(hipDeviceGetAttribute(&maxlds, hipDeviceAttributeMaxSharedMemoryPerBlock, dev));
  if (E->lds_samp > (size_t)maxlds || E->lds_meas > (size_t)maxlds)
    throw Fail(MVMC_E_INVALID_SHAPE, "model needs " + std::to_string(std::max(E->lds_samp, E->lds_meas)) +
                                         " B of LDS per workgroup, device allows " + std::to_string(maxlds));
  raise_lds_limit((const void*)k_block<double, false>, E->lds_samp);
  raise_lds_limit((const void*)k_block<float, false>, E->lds_samp);
  raise_lds_limit((const void*)k_begin<false>, E->lds_samp);
  raise_lds_limit((const void*)k_measure<false>, E->lds_meas);
  E->lds_rebuild = lds_bytes_rebuild(v);
  raise_lds_limit((const void*)k_rebuild<3, 8>, E->lds_rebuild);
  raise_lds_limit((const void*)k_rebuild<2, 16>, E->lds_rebuild);
  raise_lds_limit((const void*)k_meas_rebuild<3, 8, false>, E->lds_rebuild);
  raise_lds_limit((const void*)k_meas_rebuild<2, 16, false>, E->lds_rebuild);
  int occ_s = 0, occ_m = 0, occ_r = 0;
  {
    int occ_d = 0, occ_f = 0;
    HIPCHK(hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ_d, k_block<double, false>, 256, E->lds_samp));
    HIPCHK(hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ_f, k_block<float, false>, 256, E->lds_samp));
    occ_s = std::min(occ_d, occ_f);  // one grid serves both forms
  }
  HIPCHK(hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ_m, k_measure<false>, 256, E->lds_meas));
  int occ_r1 = 0;
  HIPCHK(hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ_r, k_rebuild<3, 8>, 256, E->lds_rebuild));
  HIPCHK(hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ_r1, k_rebuild<2, 16>, 256, E->lds_rebuild));
  occ_r1 = std::max(1, occ_r1);
  occ_s = std::max(1, occ_s);
  occ_m = std::max(1, occ_m);
  occ_r = std::max(1, occ_r);
  E->grid_samp = (int)std::min<long long>(walkers, (long long)occ_s * ncu);
  E->grid_meas = (int)std::min<long long>(walkers, (long long)occ_m * ncu);
  E->grid_rebuild = (int)std::min<long long>(walkers, (long long)occ_r * ncu);
  E->grid_rb[0] = E->grid_rebuild;
  E->grid_rb[1] = (int)std::min<long long>(walkers, (long long)occ_r1 * ncu);
  size_t work_per_slot = pfinv_work_doubles(v.n, v.nqp);
  if (rr_enabled() && v.n >= 2 && v.n <= RR_NMAX) {
    // ei, rs, pfn, then the register-resident carve; 2 n^2 doubles of work
    E->lds_rr = 2 * rnd16((size_t)v.n * 4) + rnd16((size_t)v.nqp * 8) + rr_lds_bytes(v.n);
    if (E->lds_rr <= (size_t)maxlds) {
      const int occ = rr_rebuild_prepare(rr_cfg(v.n), E->lds_rr);
      E->rb_nvar = 3;
      E->grid_rb[2] = (int)std::min<long long>(walkers, (long long)occ * ncu);
      work_per_slot = std::max(work_per_slot, (size_t)2 * v.n * v.n);
    }
  }
  // MVMC_RB_VARIANT=k fixes the rebuild variant instead of tuning it
  if (const char* f = std::getenv("MVMC_RB_VARIANT")) {
    const int k = std::atoi(f);
    if (k >= 0 && k < E->rb_nvar) E->rb_variant = k;
  }
  E->slots = std::max(std::max(E->grid_samp, E->grid_meas),
                      std::max(std::max(E->grid_rb[0], E->grid_rb[1]), E->grid_rb[2]));
  v.work = dalloc<double>(A, (size_t)E->slots * work_per_slot);
  v.qorb = dalloc<double>(A, (size_t)E->slots * v.nqp * std::max(v.nslater, 1));
  v.pend = dalloc<double>(A, (size_t)E->slots * 2 * PEND_MAX * v.nqp * v.n);
  v.nflag = dalloc<int>(A, NFLAGS);
  v.rlist = dalloc<int>(A, (size_t)walkers);
#ifdef MVMC_PHASE_CLOCKS
  v.clk = dalloc<long long>(A, (size_t)E->slots * 32);
  HIPCHK(hipMemsetAsync(v.clk, 0, sizeof(long long) * E->slots * 32, E->st));
#else
  v.clk = nullptr;
#endif
  E->gstride = std::max(1, std::max(v.ntrans, v.nexch));
  E->d_gscr = dalloc<double>(A, (size_t)E->slots * E->gstride);
  v.gscr = E->d_gscr;
  v.gstride = E->gstride;
  E->d_out5 = dalloc<double>(A, 8);
  E->d_ho = dalloc<double>(A, v.size);
  E->d_mean = dalloc<double>(A, v.size);
  E->d_diag = dalloc<double>(A, v.size);
  E->d_map = dalloc<int>(A, v.size);
  E->d_cmean = dalloc<double>(A, v.size);
  E->d_cdiag = dalloc<double>(A, v.size);
  E->d_cg = dalloc<double>(A, v.size);
  E->d_cx = dalloc<double>(A, v.size);
  hipLaunchKernelGGL(k_seed, dim3(nblk(walkers, 64)), dim3(64), 0, E->st, v, seed, n_active);
  HIPCHK(hipGetLastError());
  HIPCHK(hipStreamSynchronize(E->st));
}

void launch_rebuild(mvmc_ens* E, int var) {
  if (var == 2)
    rr_launch_rebuild(rr_cfg(E->v.n), E->grid_rb[2], E->lds_rr, E->st, E->v, -1);
  else if (var == 1)
    hipLaunchKernelGGL((k_rebuild<2, 16>), dim3(E->grid_rb[1]), dim3(256), E->lds_rebuild, E->st, E->v);
  else
    hipLaunchKernelGGL((k_rebuild<3, 8>), dim3(E->grid_rb[0]), dim3(256), E->lds_rebuild, E->st, E->v);
  HIPCHK(hipGetLastError());
}

// Start of a sampling call: with the register-resident rebuild chosen, the
// walkers resuming from their burn configuration are rebuilt by k_begin_rr
// and k_begin serves the rest (placement, failed or non-finite rebuilds);
// otherwise k_begin serves every walker.
void launch_begin(mvmc_ens* E) {
  EnsView& v = E->v;
  int marked = v.W;
  if (E->rb_variant == 2) {
    HIPCHK(hipMemsetAsync(v.nflag + 5, 0, sizeof(int), E->st));
    rr_dispatch(rr_cfg(v.n), [&]<int T, int R, bool FULL>() {
      hipLaunchKernelGGL((k_begin_rr<T, R, FULL>), dim3(E->grid_rb[2]), dim3(T * T), E->lds_rr, E->st, v);
    });
    HIPCHK(hipGetLastError());
    HIPCHK(hipMemcpyAsync(&marked, v.nflag + 5, sizeof(int), hipMemcpyDeviceToHost, E->st));
    HIPCHK(hipStreamSynchronize(E->st));
    if (marked == 0) return;
  }
  hipLaunchKernelGGL(k_begin<false>, dim3(E->grid_samp), dim3(256), E->lds_samp, E->st, v, GroupArgs{},
                     E->rb_variant == 2 ? 1 : 0);
  HIPCHK(hipGetLastError());
}

void launch_meas_rebuild(mvmc_ens* E, int s) {
  if (E->rb_variant == 2)
    rr_launch_rebuild(rr_cfg(E->v.n), E->grid_rb[2], E->lds_rr, E->st, E->v, s);
  else if (E->rb_variant == 1)
    hipLaunchKernelGGL((k_meas_rebuild<2, 16, false>), dim3(E->grid_rb[1]), dim3(256), E->lds_rebuild, E->st, E->v,
                       GroupArgs{}, s);
  else
    hipLaunchKernelGGL((k_meas_rebuild<3, 8, false>), dim3(E->grid_rb[0]), dim3(256), E->lds_rebuild, E->st, E->v,
                       GroupArgs{}, s);
  HIPCHK(hipGetLastError());
}

void sample_impl(mvmc_ens* E) {
  EnsView& v = E->v;
  const double t0 = now_s();
  E->upload_tables();
  launch_begin(E);
  std::vector<int> ws((size_t)v.W * 8);
  HIPCHK(hipMemcpyAsync(ws.data(), v.wstate, sizeof(int) * ws.size(), hipMemcpyDeviceToHost, E->st));
  HIPCHK(hipStreamSynchronize(E->st));
  E->check_status();
  int n_out = 0;
  for (int w = 0; w < v.W; ++w) n_out = std::max(n_out, ws[(size_t)w * 8 + 2]);
  const int n_in = v.interval * std::max(v.ns, 1);
  // Per outer step: k_block runs every walker until the step's end or its
  // next periodic rebuild; k_rebuild rebuilds the walkers handed over; repeat
  // until no walker is waiting.  Each k_block launch is timed with HIP
  // events on the library's stream.
  hipEvent_t e0, e1, r0, r1;
  HIPCHK(hipEventCreate(&e0));
  HIPCHK(hipEventCreate(&e1));
  HIPCHK(hipEventCreate(&r0));
  HIPCHK(hipEventCreate(&r1));
  double block_ms = 0, rebuild_ms = 0;
  int launches = 0, rebuilds = 0;
  long long rebuilt = 0, wpasses = 0, fpasses = 0;
  E->launch_log.clear();
  std::vector<long long> cnt_prev((size_t)v.W * 4), cnt_now((size_t)v.W * 4);
  E->d2h(cnt_prev.data(), v.counter, sizeof(long long) * cnt_prev.size());
  for (int out = 0; out < n_out; ++out) {
    hipLaunchKernelGGL(k_reset_resume, dim3(nblk(v.W)), dim3(256), 0, E->st, v);
    for (;;) {
      HIPCHK(hipMemsetAsync(v.nflag, 0, NFLAGS * sizeof(int), E->st));
      HIPCHK(hipEventRecord(e0, E->st));
      if (E->precision == 1)
        hipLaunchKernelGGL((k_block<float, false>), dim3(E->grid_samp), dim3(256), E->lds_samp, E->st, v, GroupArgs{},
                           out, n_in);
      else
        hipLaunchKernelGGL((k_block<double, false>), dim3(E->grid_samp), dim3(256), E->lds_samp, E->st, v, GroupArgs{},
                           out, n_in);
      HIPCHK(hipGetLastError());
      HIPCHK(hipEventRecord(e1, E->st));
      int flags[NFLAGS] = {};
      HIPCHK(hipMemcpyAsync(flags, v.nflag, sizeof(flags), hipMemcpyDeviceToHost, E->st));
      HIPCHK(hipStreamSynchronize(E->st));
      const int nf = flags[0];
      wpasses += flags[3];
      fpasses += flags[4];
      float ms = 0;
      HIPCHK(hipEventElapsedTime(&ms, e0, e1));
      block_ms += ms;
      ++launches;
      {
        E->d2h(cnt_now.data(), v.counter, sizeof(long long) * cnt_now.size());
        double props = 0, accs = 0, active = 0;
        for (int w = 0; w < v.W; ++w) {
          const size_t b = (size_t)w * 4;
          const long long dp = cnt_now[b] + cnt_now[b + 2] - cnt_prev[b] - cnt_prev[b + 2];
          props += (double)dp;
          accs += (double)(cnt_now[b + 1] + cnt_now[b + 3] - cnt_prev[b + 1] - cnt_prev[b + 3]);
          active += dp > 0 ? 1.0 : 0.0;
        }
        E->launch_log.insert(E->launch_log.end(), {(double)ms, props, accs, active});
        cnt_prev.swap(cnt_now);
      }
      if (nf == 0) break;
      rebuilt += nf;
      // variant: the tuned one, else alternate on launches that fill every
      // slot twice over (smaller ones run variant 0 and are not counted)
      const int nvar = E->rb_nvar;
      int gmax = 0, least = 0;
      for (int k = 0; k < nvar; ++k) {
        gmax = std::max(gmax, E->grid_rb[k]);
        if (E->rb_samples[k] < E->rb_samples[least]) least = k;
      }
      const bool tune = E->rb_variant < 0 && nf >= 2 * gmax;
      const int var = E->rb_variant >= 0 ? E->rb_variant : (tune ? least : 0);
      HIPCHK(hipEventRecord(r0, E->st));
      launch_rebuild(E, var);
      HIPCHK(hipEventRecord(r1, E->st));
      HIPCHK(hipEventSynchronize(r1));
      HIPCHK(hipEventElapsedTime(&ms, r0, r1));
      rebuild_ms += ms;
      ++rebuilds;
      if (tune) {
        E->rb_ms[var] += ms;
        E->rb_walkers[var] += nf;
        ++E->rb_samples[var];
        bool done = true;
        for (int k = 0; k < nvar; ++k) done &= E->rb_samples[k] >= 2;
        if (done) {
          int best = 0;
          for (int k = 1; k < nvar; ++k)
            if (E->rb_ms[k] / E->rb_walkers[k] < E->rb_ms[best] / E->rb_walkers[best]) best = k;
          E->rb_variant = best;
        }
      }
    }
  }
  for (hipEvent_t ev : {e0, e1, r0, r1}) (void)hipEventDestroy(ev);
  E->block_ms = block_ms;
  E->block_launches = launches;
  E->block_write_passes = (double)wpasses;
  E->block_flush_passes = (double)fpasses;
  E->rebuild_ms = rebuild_ms;
  E->rebuild_launches = rebuilds;
  E->rebuilt = (double)rebuilt;
  E->check_status();
  std::vector<long long> c((size_t)v.W * 4);
  E->d2h(c.data(), v.counter, sizeof(long long) * c.size());
  for (int w = 0; w < v.W; ++w) {
    E->proposals += (double)(c[(size_t)w * 4 + 0] + c[(size_t)w * 4 + 2]);
    E->accepted += (double)(c[(size_t)w * 4 + 1] + c[(size_t)w * 4 + 3]);
  }
  E->t_sample += now_s() - t0;
}

void measure_finish(mvmc_ens* E);

void measure_impl(mvmc_ens* E) {
  EnsView& v = E->v;
  const double t0 = now_s();
  hipEvent_t ev[3];
  for (auto& e : ev) HIPCHK(hipEventCreate(&e));
  double ms_rb = 0, ms_m = 0;
  for (int s = 0; s < v.nsample; ++s) {
    HIPCHK(hipEventRecord(ev[0], E->st));
    launch_meas_rebuild(E, s);
    HIPCHK(hipEventRecord(ev[1], E->st));
    hipLaunchKernelGGL(k_measure<false>, dim3(E->grid_meas), dim3(256), E->lds_meas, E->st, v, GroupArgs{}, 1,
                       s);
    HIPCHK(hipEventRecord(ev[2], E->st));
    HIPCHK(hipGetLastError());
    HIPCHK(hipEventSynchronize(ev[2]));
    float a = 0, b = 0;
    HIPCHK(hipEventElapsedTime(&a, ev[0], ev[1]));
    HIPCHK(hipEventElapsedTime(&b, ev[1], ev[2]));
    ms_rb += a;
    ms_m += b;
  }
  for (auto& e : ev) (void)hipEventDestroy(e);
  E->meas_rebuild_ms = ms_rb;
  E->measure_ms = ms_m;
  E->measured = (double)v.W * v.nsample;
  measure_finish(E);
  E->t_measure += now_s() - t0;
}

// Walker-order reductions of the measured samples and the averages
// (run.rs:1863-1903, average.rs:18-63), on the handle's stream.
void measure_finish(mvmc_ens* E) {
  EnsView& v = E->v;
  hipLaunchKernelGGL(k_reduce_energy, dim3(1), dim3(64), 0, E->st, v, E->d_out5);
  hipLaunchKernelGGL(k_reduce_ops, dim3(nblk(v.size)), dim3(256), 0, E->st, v, E->d_ho, E->d_mean, E->d_diag);
  HIPCHK(hipGetLastError());
  double o5[5];
  std::vector<double> ho(v.size), mean(v.size), diag(v.size);
  HIPCHK(hipMemcpyAsync(o5, E->d_out5, sizeof(double) * 5, hipMemcpyDeviceToHost, E->st));
  HIPCHK(hipMemcpyAsync(ho.data(), E->d_ho, sizeof(double) * v.size, hipMemcpyDeviceToHost, E->st));
  HIPCHK(hipMemcpyAsync(mean.data(), E->d_mean, sizeof(double) * v.size, hipMemcpyDeviceToHost, E->st));
  HIPCHK(hipMemcpyAsync(diag.data(), E->d_diag, sizeof(double) * v.size, hipMemcpyDeviceToHost, E->st));
  HIPCHK(hipStreamSynchronize(E->st));
  E->check_status();
  E->wc = o5[0]; E->etot = o5[1]; E->etot2 = o5[2]; E->sztot = o5[3]; E->sztot2 = o5[4];
  E->oo.assign(2 * (size_t)v.size, 0.0);
  for (int i = 0; i < v.size; ++i) { E->oo[i] = mean[i]; E->oo[v.size + i] = diag[i]; }
  E->ho = ho;
  // weight_average_we / weight_average_sr_opt_real (average.rs:18-63)
  if (std::fabs(E->wc) >= 1.0e-15) {
    const double inv_c = E->wc / (E->wc * E->wc + 0.0 * 0.0);
    E->etot *= inv_c; E->etot2 *= inv_c; E->sztot *= inv_c; E->sztot2 *= inv_c;
    const double inv = 1.0 / E->wc;
    for (auto& x : E->oo) x *= inv;
    for (auto& x : E->ho) x *= inv;
  }
}

// Measurement of caller-supplied configurations: MeasurementStages::
// measure_sample batched over samples (crates/mvmc-core/src/stage_backend.rs:
// 418-438; per sample the work of run.rs:2937-3616).  Configuration c goes to
// the saved-sample slot of walker c mod W in layer c / W; every layer is
// rebuilt (k_meas_rebuild) and measured (k_measure) by the kernels of
// mvmc_ens_measure with the current parameters.  EleCfg / EleNum and the
// projection counts are derived here (projection.rs:179-277).  The saved
// samples and the walkers' inverse tables are overwritten.
void measure_configs_impl(mvmc_ens* E, int count, const int* ele_idx, double* e_loc, double* o, int* valid) {
  EnsView& v = E->v;
  if (count < 0 || (long long)count > (long long)v.W * v.nsample)
    throw Fail(MVMC_E_INVALID_SHAPE, "count must be between 0 and walkers x NVMCSample");
  if (count == 0) return;
  if (!ele_idx || !e_loc) throw Fail(MVMC_E_INVALID_SHAPE, "null buffer");
  const mvmc::Model& m = E->m;
  const int n = v.n, ne = v.ne, ns = v.ns, n2 = v.n2, np = v.nproj, size = v.size, nsamp = v.nsample;
  const int npc = np > 0 ? np : 1;
  E->upload_tables();
  const int W0 = v.W;
  struct RestoreW {
    EnsView& view;
    int w;
    ~RestoreW() { view.W = w; }
  } restore{v, W0};
  hipStream_t st = E->st;
  for (int s = 0; (long long)s * W0 < count; ++s) {
    const int base = s * W0, wl = std::min(W0, count - base);
    std::vector<int> idx((size_t)wl * n), cfg((size_t)wl * n2, -1), num((size_t)wl * n2, 0), pc((size_t)wl * npc, 0);
    for (int w = 0; w < wl; ++w) {
      const int* e = ele_idx + (size_t)(base + w) * n;
      int* cf = cfg.data() + (size_t)w * n2;
      int* nm = num.data() + (size_t)w * n2;
      for (int mi = 0; mi < n; ++mi) {
        const int ri = e[mi];
        if (ri < 0 || ri >= ns)
          throw Fail(MVMC_E_INPUT, "configuration " + std::to_string(base + w) + ": site index out of range");
        const int slot = ri + (mi / ne) * ns;
        if (cf[slot] != -1)
          throw Fail(MVMC_E_INPUT, "configuration " + std::to_string(base + w) + ": two electrons of one spin on a site");
        cf[slot] = mi % ne;
        nm[slot] = 1;
        idx[(size_t)w * n + mi] = ri;
      }
      int* cnt = pc.data() + (size_t)w * npc;
      for (int ri = 0; ri < ns; ++ri) {
        const int g = m.gutz_idx[ri];
        if (g >= 0 && g < np) cnt[g] += nm[ri] * nm[ri + ns];
      }
      if (np > m.n_gutz)
        for (int ri = 0; ri < ns; ++ri) {
          const int xi = nm[ri] + nm[ri + ns] - 1;
          if (xi == 0) continue;
          for (int rj = ri + 1; rj < ns; ++rj) {
            const int j = m.jast_idx[(size_t)ri * ns + rj];
            if (j < 0) continue;
            cnt[m.n_gutz + j] += xi * (nm[rj] + nm[rj + ns] - 1);
          }
        }
    }
    auto put = [&](int* dst, const std::vector<int>& src, int width) {
      HIPCHK(hipMemcpy2DAsync(dst + (size_t)s * width, sizeof(int) * nsamp * width, src.data(), sizeof(int) * width,
                              sizeof(int) * width, wl, hipMemcpyHostToDevice, st));
    };
    put(v.s_idx, idx, n);
    put(v.s_cfg, cfg, n2);
    put(v.s_num, num, n2);
    if (np > 0) put(v.s_proj, pc, np);
    v.W = wl;
    launch_meas_rebuild(E, s);
    hipLaunchKernelGGL(k_measure<false>, dim3(E->grid_meas), dim3(256), E->lds_meas, st, v, GroupArgs{}, 1, s);
    HIPCHK(hipGetLastError());
    v.W = W0;
    HIPCHK(hipMemcpy2DAsync(e_loc + base, sizeof(double), v.e_loc + s, sizeof(double) * nsamp, sizeof(double), wl,
                            hipMemcpyDeviceToHost, st));
    if (o)
      HIPCHK(hipMemcpy2DAsync(o + (size_t)base * size, sizeof(double) * size, v.o_store + (size_t)s * size,
                              sizeof(double) * size * nsamp, sizeof(double) * size, wl, hipMemcpyDeviceToHost, st));
    if (valid)
      HIPCHK(hipMemcpy2DAsync(valid + base, sizeof(int), v.s_valid + s, sizeof(int) * nsamp, sizeof(int), wl,
                              hipMemcpyDeviceToHost, st));
    HIPCHK(hipStreamSynchronize(st));
  }
  E->check_status();
}

// Julia's ComplexF64 division (crates/pfapack/src/julia_complex.rs:9-77) on
// the host, for the zvo_out variance.
double julia_div_part_h(double a, double b, double c, double d, double r, double t) {
  if (r != 0.0) {
    const double br = b * r;
    return br != 0.0 ? (a + br) * t : a * t + (b * t) * r;
  }
  return (a + d * (b / c)) * t;
}
void julia_div_h(double a, double b, double c, double d, double& re, double& im) {
  const double ab = std::fabs(a) >= std::fabs(b) ? std::fabs(a) : std::fabs(b);
  const double cdm = std::fabs(c) >= std::fabs(d) ? std::fabs(c) : std::fabs(d);
  if (std::isinf(c) || std::isinf(d)) {
    auto sgn = [](double x) { return (x == 0.0 || std::isnan(x)) ? x : std::copysign(1.0, x); };
    if (std::isfinite(a) && std::isfinite(b)) {
      re = 0.0 * sgn(a) * sgn(c);
      im = -0.0 * sgn(b) * sgn(d);
    } else {
      re = im = NAN;
    }
    return;
  }
  const double halfov = 0x1.fffffffffffffp+1022, twouneps = 0x1p-969, bs = 0x1p+105;
  double s = 1.0;
  const bool scaled = ab >= halfov || ab <= twouneps || cdm >= halfov || cdm <= twouneps;
  if (scaled) {
    if (ab >= halfov) { a *= 0.5; b *= 0.5; s *= 2.0; }
    else if (ab <= twouneps) { a *= bs; b *= bs; s /= bs; }
    if (cdm >= halfov) { c *= 0.5; d *= 0.5; s *= 0.5; }
    else if (cdm <= twouneps) { c *= bs; d *= bs; s *= bs; }
  }
  bool swapped = false;
  if (!(std::fabs(d) <= std::fabs(c))) {
    std::swap(a, b);
    std::swap(c, d);
    swapped = true;
  }
  const double r = d / c, t = 1.0 / (c + d * r);
  re = julia_div_part_h(a, b, c, d, r, t);
  im = julia_div_part_h(b, -a, c, d, r, t);
  if (swapped) im = -im;
  if (scaled) { re *= s; im *= s; }
}

std::string cfmt18(double x, bool space) {
  char b[64];
  std::snprintf(b, sizeof b, space ? "% .18e" : "%.18e", x);
  return b;
}

// output_data (crates/mvmc-core/src/io.rs:142-232), text form: appends the
// step's {head}_out.dat row and {head}_var.dat line (step 0 truncates both);
// called after the measurement and before the SR update (run.rs:1376-1385).
void write_step_output(mvmc_ens* E, int step) {
  const mvmc::Model& m = E->m;
  const std::string dir = E->out_dir.empty() ? std::string() : E->out_dir + "/";
  if (!E->out_dir.empty()) {  // output_path creates the directory (io.rs:127-140)
    std::error_code ec;
    std::filesystem::create_directories(E->out_dir, ec);
  }
  const std::string head = m.mp.c_data_file_head.empty() ? "zvo" : m.mp.c_data_file_head;
  const double er = E->etot, ei = 0.0;
  // (etot2 - etot*etot) / (etot*etot) in complex arithmetic, Julia's division
  const double sq_re = er * er - ei * ei, sq_im = er * ei + ei * er;
  double vre = 0.0, vim = 0.0;
  if (std::hypot(er, ei) > 1.0e-14) julia_div_h(E->etot2 - sq_re, 0.0 - sq_im, sq_re, sq_im, vre, vim);
  const char* mode = step == 0 ? "w" : "a";
  const std::string op = dir + head + "_out.dat";
  FILE* f = std::fopen(op.c_str(), mode);
  if (!f) throw Fail(MVMC_E_INPUT, "cannot write " + op);
  std::fprintf(f, "%s %s  %s %s %s %s\n", cfmt18(er, true).c_str(), cfmt18(ei, true).c_str(),
               cfmt18(E->etot2, true).c_str(), cfmt18(vre, true).c_str(), cfmt18(E->sztot, false).c_str(),
               cfmt18(E->sztot2, false).c_str());
  std::fclose(f);
  const std::string vp = dir + head + "_var.dat";
  FILE* g = std::fopen(vp.c_str(), mode);
  if (!g) throw Fail(MVMC_E_INPUT, "cannot write " + vp);
  std::fprintf(g, "%s %s 0.0 %s %s 0.0 ", cfmt18(er, true).c_str(), cfmt18(ei, true).c_str(),
               cfmt18(E->etot2, true).c_str(), cfmt18(0.0, true).c_str());
  for (double x : m.proj) std::fprintf(g, "%s %s 0.0 ", cfmt18(x, true).c_str(), cfmt18(0.0, true).c_str());
  for (double x : m.slater) std::fprintf(g, "%s %s 0.0 ", cfmt18(x, true).c_str(), cfmt18(0.0, true).c_str());
  std::fprintf(g, "\n");
  std::fclose(g);
}

// Parameter update of an SR step (sr.rs:408-421, sync.rs): p[map[si]] += x[si].
void apply_sr_update(mvmc_ens* E, const std::vector<int>& map, const std::vector<double>& x) {
  mvmc::Model& m = E->m;
  for (size_t si = 0; si < map.size(); ++si) {
    const int pi = map[si];
    if (pi < m.n_proj()) m.proj[pi] += x[si];
    else m.slater[pi - m.n_proj()] += x[si];
  }
  mvmc::sync_modified_parameter(m, true);
}

// Direct SR (sr.rs:60-160 real path; sr_backend.rs:484-516 assembly): the
// Gram inv_w * P P^T of the packed active samples (P = [na][rows] column
// major) by rocBLAS DGEMM on the matrix cores, S = Gram - mean mean^T with
// the diagonal scaled by 1 + DSROptStaDel, then Cholesky (rocSOLVER DPOTRF,
// DPOTRS, upper) on g in place.  A failed factorisation or a non-finite
// solution skips the update (return 1), as the reference's cholesky_solve.
int direct_sr(mvmc_ens* E, const std::vector<int>& map, int na, size_t rows, double inv_w, double t0,
              long long* info_out, bool f32) {
  mvmc::Model& m = E->m;
  hipStream_t st = E->st;
  if (!E->blas) {
    BLASCHK(rocblas_create_handle(&E->blas));
    E->d_info = dalloc<int>(E->allocs, 4);
  }
  BLASCHK(rocblas_set_stream(E->blas, st));
  const size_t nn = (size_t)na * na;
  if (E->S_cap < nn) {
    if (E->d_S) (void)hipFree(E->d_S);
    E->d_S = nullptr;
    HIPCHK(hipMalloc(&E->d_S, sizeof(double) * nn));
    E->S_cap = nn;
  }
  hipEvent_t c0, c1;
  HIPCHK(hipEventCreate(&c0));
  HIPCHK(hipEventCreate(&c1));
  HIPCHK(hipEventRecord(c0, st));
  // upper triangle of the Gram inv_w P P^T on the matrix cores (FP64 MFMA;
  // FP32 MFMA with FP32 accumulation in the precision form)
  if (rows > (size_t)INT32_MAX) throw Fail(MVMC_E_INVALID_SHAPE, "too many sample rows for the Gram kernel");
  const unsigned tiles = (unsigned)gm_tiles(na);
  hipEvent_t g1;
  HIPCHK(hipEventCreate(&g1));
  if (f32)
    hipLaunchKernelGGL(k_gram_mfma<float>, dim3(tiles), dim3(256), 0, st, (const float*)E->d_pack, (int)rows, na,
                       inv_w, E->d_S, 0);
  else
    hipLaunchKernelGGL(k_gram_mfma<double>, dim3(tiles), dim3(256), 0, st, (const double*)E->d_pack, (int)rows, na,
                       inv_w, E->d_S, 0);
  HIPCHK(hipGetLastError());
  HIPCHK(hipEventRecord(g1, st));
  hipLaunchKernelGGL(k_sr_assemble, dim3(nblk((long long)nn)), dim3(256), 0, st, na, E->d_cmean,
                     1.0 + m.mp.dsr_opt_sta_del, E->d_S);
  HIPCHK(hipGetLastError());
  HIPCHK(hipMemcpyAsync(E->d_cx, E->d_cg, sizeof(double) * na, hipMemcpyDeviceToDevice, st));
  BLASCHK(rocsolver_dpotrf(E->blas, rocblas_fill_upper, na, E->d_S, na, E->d_info));
  int info = 0;
  HIPCHK(hipMemcpyAsync(&info, E->d_info, sizeof(int), hipMemcpyDeviceToHost, st));
  HIPCHK(hipStreamSynchronize(st));
  if (info == 0) BLASCHK(rocsolver_dpotrs(E->blas, rocblas_fill_upper, na, 1, E->d_S, na, E->d_cx, na));
  HIPCHK(hipEventRecord(c1, st));
  HIPCHK(hipEventSynchronize(c1));
  float ms = 0, gms = 0;
  HIPCHK(hipEventElapsedTime(&ms, c0, c1));
  HIPCHK(hipEventElapsedTime(&gms, c0, g1));
  (void)hipEventDestroy(c0);
  (void)hipEventDestroy(c1);
  (void)hipEventDestroy(g1);
  E->cg_ms = ms;
  E->gram_ms = gms;
  E->cg_iter = 0;
  E->cg_na = na;
  E->cg_rows = (double)rows;
  E->cg_info = 0;
  *info_out = 0;
  E->cg_x.clear();
  if (info != 0) {
    E->t_sr += now_s() - t0;
    return 1;
  }
  std::vector<double> x(na);
  E->d2h(x.data(), E->d_cx, sizeof(double) * na);
  E->cg_x = x;
  for (double xv : x)
    if (!std::isfinite(xv)) {
      E->t_sr += now_s() - t0;
      return 1;
    }
  apply_sr_update(E, map, x);
  E->t_sr += now_s() - t0;
  return 0;
}

// sr_cg.rs:201-484 (real parameters, offset 1, no RescaleSmat); NSRCG 0
// takes the direct path above.
int sr_impl(mvmc_ens* E, long long* info_out, bool force_direct = false) {
  mvmc::Model& m = E->m;
  EnsView& v = E->v;
  const double t0 = now_s();
  const int np = m.n_para(), size = m.sr_opt_size();
  if (m.mp.rescale_smat) throw Fail(MVMC_E_INPUT, "RescaleSmat is not supported");
  std::vector<double> var(np);
  bool nan = false;
  for (int pi = 0; pi < np; ++pi) {
    const int idx = pi + 1;
    var[pi] = E->oo[size + idx] - E->oo[idx] * E->oo[idx];
    nan = nan || std::isnan(var[pi]);
  }
  double vmax = -INFINITY;
  if (nan) vmax = NAN;
  else
    for (double x : var) vmax = std::fmax(vmax, x);
  const double thr = vmax * m.mp.dsr_opt_red_cut;
  std::vector<int> map;
  for (int pi = 0; pi < np; ++pi) {
    if (m.opt_flags[2 * pi] != 1) continue;
    if (var[pi] < thr) continue;
    map.push_back(pi);
  }
  E->cg_x.clear();
  if (map.empty()) { *info_out = 0; E->cg_info = 0; E->t_sr += now_s() - t0; return 0; }
  const int na = (int)map.size();
  std::vector<double> mean(na), diag(na), g(na);
  const double dt = 2.0 * m.mp.dsr_opt_step_dt;
  for (int si = 0; si < na; ++si) {
    const int idx = map[si] + 1;
    mean[si] = E->oo[idx];
    diag[si] = var[map[si]];
    g[si] = -dt * (E->ho[idx] - E->ho[0] * E->oo[idx]);
  }
  hipStream_t st = E->st;
  HIPCHK(hipMemcpyAsync(E->d_map, map.data(), sizeof(int) * na, hipMemcpyHostToDevice, st));
  HIPCHK(hipMemcpyAsync(E->d_cmean, mean.data(), sizeof(double) * na, hipMemcpyHostToDevice, st));
  HIPCHK(hipMemcpyAsync(E->d_cdiag, diag.data(), sizeof(double) * na, hipMemcpyHostToDevice, st));
  HIPCHK(hipMemcpyAsync(E->d_cg, g.data(), sizeof(double) * na, hipMemcpyHostToDevice, st));
  const size_t rows = (size_t)v.W * v.nsample;
  // the precision form packs the samples in FP32: the CG operand is
  // accumulated in FP64, the direct Gram in FP32 on the matrix cores; the
  // FP64 path packs doubles
  const bool direct = m.mp.nsrcg == 0 || force_direct;
  const bool f32 = E->precision == 1;
  const size_t pbytes = rows * na * (f32 ? sizeof(float) : sizeof(double));
  if (E->pack_cap < pbytes) {
    if (E->d_pack) (void)hipFree(E->d_pack);
    E->d_pack = nullptr;
    HIPCHK(hipMalloc(&E->d_pack, pbytes));
    E->pack_cap = pbytes;
  }
  if (f32)
    hipLaunchKernelGGL(k_cg_pack<float>, dim3(nblk((long long)(rows * na))), dim3(256), 0, st, v, E->d_map, na,
                       (float*)E->d_pack);
  else
    hipLaunchKernelGGL(k_cg_pack<double>, dim3(nblk((long long)(rows * na))), dim3(256), 0, st, v, E->d_map, na,
                       (double*)E->d_pack);
  HIPCHK(hipGetLastError());
  const double inv_w = 1.0 / E->wc;
  if (direct) return direct_sr(E, map, na, rows, inv_w, t0, info_out, f32);
  const int max_it = m.mp.nsr_opt_cg_max_iter > 0 ? m.mp.nsr_opt_cg_max_iter : na;
  hipEvent_t c0, c1;
  HIPCHK(hipEventCreate(&c0));
  HIPCHK(hipEventCreate(&c1));
  HIPCHK(hipEventRecord(c0, st));
  // precision form: the CG product from the FP32-MFMA Gram of the FP32
  // samples (one MFMA pass, then na^2 per product instead of rows * na),
  // when the Gram fits in 2 GiB
  const float* gram = nullptr;
  if (f32 && (size_t)na * na * sizeof(float) <= ((size_t)2 << 30)) {
    const size_t gb = (size_t)na * na * sizeof(float);
    if (E->gram_cap < gb) {
      if (E->d_gram) (void)hipFree(E->d_gram);
      E->d_gram = nullptr;
      HIPCHK(hipMalloc(&E->d_gram, gb));
      E->gram_cap = gb;
    }
    hipLaunchKernelGGL((k_gram_mfma<float, float>), dim3((unsigned)gm_tiles(na)), dim3(256), 0, st,
                       (const float*)E->d_pack, (int)rows, na, 1.0, (float*)E->d_gram, 1);
    HIPCHK(hipGetLastError());
    gram = (const float*)E->d_gram;
  }
  auto solve = [&](bool dscale) {
    return f32 ? cg_solve_dev(na, v.W, v.nsample, (const float*)E->d_pack, E->d_cmean, E->d_cdiag, E->d_cg, inv_w,
                              m.mp.dsr_opt_sta_del, m.mp.dsr_opt_cg_tol, max_it, dscale, E->d_cx, E->cg, st, gram)
               : cg_solve_dev(na, v.W, v.nsample, (const double*)E->d_pack, E->d_cmean, E->d_cdiag, E->d_cg, inv_w,
                              m.mp.dsr_opt_sta_del, m.mp.dsr_opt_cg_tol, max_it, dscale, E->d_cx, E->cg, st);
  };
  long long status = solve(m.mp.use_diag_scale != 0);
  HIPCHK(hipEventRecord(c1, st));
  HIPCHK(hipEventSynchronize(c1));
  {
    float ms = 0;
    HIPCHK(hipEventElapsedTime(&ms, c0, c1));
    E->cg_ms = ms;
    E->cg_iter = (double)(status >= 0 ? status : -status - 1);
    E->cg_na = na;
    E->cg_rows = (double)rows;
  }
  (void)hipEventDestroy(c0);
  (void)hipEventDestroy(c1);
  long long info = status;
  if (status < 0) {
    if (m.mp.nsrcg_fallback) {
      long long st2 = solve(m.mp.use_diag_scale == 0);
      info = -((-status - 1) + 1);
      if (st2 < 0 && m.mp.nsrcg_abort_on_fail) throw Fail(MVMC_E_NUMERIC, "SR-CG fallback failed");
    } else {
      if (m.mp.nsrcg_abort_on_fail) throw Fail(MVMC_E_NUMERIC, "SR-CG failed (NSRCGAbortOnFail)");
      info = -((-status - 1) + 1);
    }
  }
  std::vector<double> x(na);
  E->d2h(x.data(), E->d_cx, sizeof(double) * na);
  E->cg_x = x;
  E->cg_info = info;
  *info_out = info;
  for (double xv : x)
    if (!std::isfinite(xv)) {
      if (m.mp.nsrcg_abort_on_fail) throw Fail(MVMC_E_NUMERIC, "SR-CG produced non-finite updates");
      E->t_sr += now_s() - t0;
      return 1;
    }
  apply_sr_update(E, map, x);
  E->t_sr += now_s() - t0;
  return 0;
}

// The rest of one SR step after the measurement (run.rs:1294-1475): zvo row,
// per-step output, SR solve and update, the optimisation window.
void finish_opt_step(mvmc_ens* e, double* row) {
  if (row) {
    double var = 0.0;
    if (std::fabs(e->etot) > 1.0e-14) var = (e->etot2 - e->etot * e->etot) / (e->etot * e->etot);
    const double r6[6] = {e->etot, 0.0, e->etot2, var, e->sztot, e->sztot2};
    std::memcpy(row, r6, sizeof(r6));
  }
  if (e->out_enabled) write_step_output(e, e->opt_steps);
  long long info = 0;
  int rc = sr_impl(e, &info);
  if (rc != 0) throw Fail(MVMC_E_NUMERIC, "SR step skipped (non-finite update)");
  // store_opt_data for the window [NSROptItrStep - NSROptItrSmp, NSROptItrStep)
  // (crates/mvmc-core/src/run.rs:1258-1262,1457-1460)
  const mvmc::Model& m = e->m;
  const int step = e->opt_steps++;
  const int len = m.mp.nsr_opt_itr_smp, start = std::max(m.mp.nsr_opt_itr_step, 0) - len;
  if (len > 0 && step >= start && step - start < len) {
    const size_t k = (size_t)(step - start);
    if (e->opt_win.size() < (size_t)len) {
      e->opt_win.resize(len);
      e->opt_win_e.resize(len);
      e->opt_win_e2.resize(len);
      e->opt_win_set.assign(len, 0);
    }
    std::vector<double> p(m.proj);
    p.insert(p.end(), m.slater.begin(), m.slater.end());
    e->opt_win[k] = p;
    e->opt_win_e[k] = e->etot;
    e->opt_win_e2[k] = e->etot2;
    e->opt_win_set[k] = 1;
  }
}

// ---------------------------------------------------------------------------
// Ensemble group: several independent runs (models of uneven Nsize) stepped
// together, one launch per kernel class (walker = rank structure of
// crates/mvmc-core/src/multichain.rs:647-689 per run).  The grid of each
// launch is sized for the largest run's LDS and cut into per-run slot ranges
// (GroupArgs); a run's slot range serves its walkers exactly as its own
// launch would, so every run's results equal those of the run stepped alone.
// ---------------------------------------------------------------------------
struct GroupPart {
  std::vector<int> start;     // [nruns + 1]
  int* d = nullptr;           // device copy
  long long* time = nullptr;  // [capacity] busy ticks per workgroup of the last step
  int capacity = 0;           // resident workgroups of the launch class
  std::vector<double> weight; // per-run work estimate the ranges are cut by
};

// Complex planes: the same shapes, full form n <= 64, panel form n <= 160
// (two doubles per slot), each in both LTL orders.
template <int ORDER, class F>
void rr_dispatch_c(const RrCfg& c, F&& f) {
  if (c.full) {
    switch (c.r) {
      case 1: f.template operator()<RR_TS, 1, true, ORDER>(); return;
      case 2: f.template operator()<RR_TS, 2, true, ORDER>(); return;
      case 3: f.template operator()<RR_TS, 3, true, ORDER>(); return;
      case 4: f.template operator()<RR_TS, 4, true, ORDER>(); return;
    }
  } else {
    switch (c.r) {
      case 3: f.template operator()<RR_T, 3, false, ORDER>(); return;
      case 4: f.template operator()<RR_T, 4, false, ORDER>(); return;
      case 5: f.template operator()<RR_T, 5, false, ORDER>(); return;
    }
  }
  throw Fail(MVMC_E_INVALID_SHAPE, "complex register-resident path covers 2 <= n <= 160");
}

void launch_pfinv_rr_c128(const cz* in, cz* out, int n, int count, int order, cz* pf, int* status, hipStream_t st) {
  int dev = 0, ncu = 0;
  HIPCHK(hipGetDevice(&dev));
  HIPCHK(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev));
  const RrCfg c = rr_cfg(n);
  const size_t lds = rr_lds_bytes<cz>(n);
  auto go = [&]<int T, int R, bool FULL, int ORDER>() {
    raise_lds_limit((const void*)k_pfinv_rr_c128<T, R, FULL, ORDER>, lds);
    int occ = 1;
    HIPCHK(hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, k_pfinv_rr_c128<T, R, FULL, ORDER>, T * T, lds));
    const int grid = (int)std::min<long long>(count, (long long)std::max(occ, 1) * std::max(ncu, 1));
    cz* gs = nullptr;
    HIPCHK(hipMallocAsync((void**)&gs, sizeof(cz) * 2 * (size_t)n * n * grid, st));
    hipLaunchKernelGGL((k_pfinv_rr_c128<T, R, FULL, ORDER>), dim3(grid), dim3(T * T), lds, st, in, out, n, count, gs,
                       pf, status);
    HIPCHK(hipGetLastError());
    HIPCHK(hipFreeAsync(gs, st));
  };
  if (order == 0) rr_dispatch_c<0>(c, go);
  else rr_dispatch_c<1>(c, go);
}

void launch_pfinv_rr(const double* in, double* out, int n, int count, double* pf, int* status, hipStream_t st) {
  int dev = 0, ncu = 0;
  HIPCHK(hipGetDevice(&dev));
  HIPCHK(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev));
  const RrCfg c = rr_cfg(n);
  const size_t lds = rr_lds_bytes(n);
  rr_dispatch(c, [&]<int T, int R, bool FULL>() {
    raise_lds_limit((const void*)k_pfinv_rr<T, R, FULL>, lds);
    int occ = 1;
    HIPCHK(hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, k_pfinv_rr<T, R, FULL>, T * T, lds));
    const int grid = (int)std::min<long long>(count, (long long)std::max(occ, 1) * std::max(ncu, 1));
    double* gs = nullptr;
    HIPCHK(hipMallocAsync((void**)&gs, sizeof(double) * 2 * (size_t)n * n * grid, st));
    hipLaunchKernelGGL((k_pfinv_rr<T, R, FULL>), dim3(grid), dim3(T * T), lds, st, in, out, n, count, gs, pf, status);
    HIPCHK(hipGetLastError());
    HIPCHK(hipFreeAsync(gs, st));
  });
}

}  // namespace

struct mvmc_group {
  std::vector<mvmc_ens*> runs;
  hipStream_t st = nullptr;
  std::vector<void*> allocs;
  EnsView* d_views = nullptr;
  int* d_flags = nullptr;  // [nruns][NFLAGS]: the runs' walker counters (their v.nflag points here)
  size_t lds_samp = 0, lds_meas = 0, lds_rebuild = 0;
  GroupPart samp, meas, rb;  // slot ranges of the placement, measurement and measurement-rebuild launches
  // shared walker queue of the sweep and periodic-rebuild launches
  int* d_gq = nullptr;
  int* d_glist = nullptr;
  int* d_qstart = nullptr;
  int* d_qrun = nullptr;
  double* d_work = nullptr;
  double* d_pend = nullptr;
  long long work_stride = 0, pend_stride = 0;
  int graph_rounds = 0;      // k_block + k_rebuild rounds per captured graph (0: direct launches)
  int autotune = 1;          // re-cut the slot ranges from measured busy times after every step
  // timings of the last group step
  double t_sample = 0, t_measure = 0, t_sr = 0, block_ms = 0, rebuild_ms = 0;
  int launches = 0, graph_replays = 0;
  ~mvmc_group() {
    for (void* p : allocs) (void)hipFree(p);
    if (st) (void)hipStreamDestroy(st);
    for (mvmc_ens* e : runs) delete e;
  }
};

namespace {

// Slot ranges of one launch class: run r gets a share of `grid` slots
// proportional to weight[r] (at least 1, at most its own scratch slots and
// walkers).
void group_partition(mvmc_group* G, GroupPart& P, int grid, const std::vector<double>& weight) {
  const int R = (int)G->runs.size();
  P.capacity = grid;
  P.weight = weight;
  if (!P.time) P.time = dalloc<long long>(G->allocs, (size_t)grid);
  double tot = 0;
  for (double x : weight) tot += x;
  std::vector<int> cnt(R);
  int used = 0;
  for (int r = 0; r < R; ++r) {
    const mvmc_ens* E = G->runs[r];
    int c = (int)std::floor(grid * weight[r] / tot);
    c = std::max(1, std::min({c, E->slots, E->v.W}));
    cnt[r] = c;
    used += c;
  }
  // hand out what rounding left, largest weight first
  std::vector<int> order(R);
  for (int r = 0; r < R; ++r) order[r] = r;
  std::sort(order.begin(), order.end(), [&](int a, int b) { return weight[a] > weight[b]; });
  for (bool more = true; more && used < grid;) {
    more = false;
    for (int r : order) {
      const mvmc_ens* E = G->runs[r];
      if (used < grid && cnt[r] < std::min(E->slots, E->v.W)) { ++cnt[r]; ++used; more = true; }
    }
  }
  P.start.assign(R + 1, 0);
  for (int r = 0; r < R; ++r) P.start[r + 1] = P.start[r] + cnt[r];
  if (!P.d) P.d = dalloc<int>(G->allocs, R + 1);
  HIPCHK(hipMemcpyAsync(P.d, P.start.data(), sizeof(int) * (R + 1), hipMemcpyHostToDevice, G->st));
  HIPCHK(hipStreamSynchronize(G->st));
}

__global__ void k_fill_int(int* p, int n, int v) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}

void group_create_impl(mvmc_group* G, int nruns, const char* const* namelists, const char* const* initials,
                       const int* walkers) {
  require_gpu();
  if (nruns < 1) throw Fail(MVMC_E_INVALID_SHAPE, "a group needs at least one run");
  for (int r = 0; r < nruns; ++r) {
    mvmc_ens* E = new mvmc_ens();
    G->runs.push_back(E);
    create_impl(E, namelists[r], initials ? initials[r] : nullptr, walkers[r], 0, 0);
    if (E->v.nsample != G->runs[0]->v.nsample)
      throw Fail(MVMC_E_INPUT, "runs of a group must share NVMCSample");
  }
  HIPCHK(hipStreamCreateWithFlags(&G->st, hipStreamNonBlocking));
  G->d_flags = dalloc<int>(G->allocs, NFLAGS * (size_t)nruns);
  std::vector<EnsView> views;
  for (int r = 0; r < nruns; ++r) {
    mvmc_ens* E = G->runs[r];
    E->v.nflag = G->d_flags + NFLAGS * r;
    views.push_back(E->v);
    G->lds_samp = std::max(G->lds_samp, E->lds_samp);
    G->lds_meas = std::max(G->lds_meas, E->lds_meas);
    G->lds_rebuild = std::max(G->lds_rebuild, E->lds_rebuild);
  }
  G->d_views = dalloc<EnsView>(G->allocs, nruns);
  HIPCHK(hipMemcpyAsync(G->d_views, views.data(), sizeof(EnsView) * nruns, hipMemcpyHostToDevice, G->st));
  raise_lds_limit((const void*)k_begin<true>, G->lds_samp);
  raise_lds_limit((const void*)k_block<double, true>, G->lds_samp);
  raise_lds_limit((const void*)k_rebuild_q<3, 8>, G->lds_rebuild);
  raise_lds_limit((const void*)k_meas_rebuild<3, 8, true>, G->lds_rebuild);
  raise_lds_limit((const void*)k_measure<true>, G->lds_meas);
  int dev = 0, ncu = 0, occ_s = 0, occ_m = 0, occ_r = 0;
  HIPCHK(hipGetDevice(&dev));
  HIPCHK(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev));
  HIPCHK(hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ_s, k_block<double, true>, 256, G->lds_samp));
  HIPCHK(hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ_m, k_measure<true>, 256, G->lds_meas));
  HIPCHK(hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ_r, k_meas_rebuild<3, 8, true>, 256, G->lds_rebuild));
  // first work estimates (re-cut from measured busy times after every step):
  // a proposal costs a fixed serial chain plus row work ~ n NQP, a walker's
  // step ~ Nsite proposals; a rebuild ~ n^3 NQP on top of a fixed latency
  std::vector<double> ws, wr;
  for (mvmc_ens* E : G->runs) {
    const double W = E->v.W, n = E->v.n, q = E->v.nqp;
    ws.push_back(W * E->v.ns * (600.0 + n * q));
    wr.push_back(W * (2.0e5 + n * n * n * q / 16.0));
  }
  group_partition(G, G->samp, std::max(1, occ_s) * ncu, ws);
  group_partition(G, G->meas, std::max(1, occ_m) * ncu, wr);
  group_partition(G, G->rb, std::max(1, occ_r) * ncu, wr);
  // shared walker queue: runs in descending order of Nsize (longest walkers
  // first), scratch for the largest run in every workgroup of the sweep and
  // rebuild launches
  std::vector<int> order(nruns), qstart(nruns + 1, 0);
  for (int r = 0; r < nruns; ++r) order[r] = r;
  std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return G->runs[a]->v.n > G->runs[b]->v.n; });
  for (int k = 0; k < nruns; ++k) qstart[k + 1] = qstart[k] + G->runs[order[k]]->v.W;
  G->d_qstart = dalloc<int>(G->allocs, nruns + 1);
  G->d_qrun = dalloc<int>(G->allocs, nruns);
  G->d_gq = dalloc<int>(G->allocs, 4);
  G->d_glist = dalloc<int>(G->allocs, std::max(qstart[nruns], 1));
  HIPCHK(hipMemcpyAsync(G->d_qstart, qstart.data(), sizeof(int) * (nruns + 1), hipMemcpyHostToDevice, G->st));
  HIPCHK(hipMemcpyAsync(G->d_qrun, order.data(), sizeof(int) * nruns, hipMemcpyHostToDevice, G->st));
  for (mvmc_ens* E : G->runs) {
    G->work_stride = std::max<long long>(G->work_stride, (long long)pfinv_work_doubles(E->v.n, E->v.nqp));
    G->pend_stride = std::max<long long>(G->pend_stride, 2LL * PEND_MAX * E->v.nqp * E->v.n);
  }
  const size_t slots = (size_t)std::max(G->samp.capacity, G->rb.capacity);
  G->d_work = dalloc<double>(G->allocs, slots * G->work_stride);
  G->d_pend = dalloc<double>(G->allocs, slots * G->pend_stride);
  HIPCHK(hipStreamSynchronize(G->st));
}

GroupArgs group_args(mvmc_group* G, const GroupPart& P, bool timed = true) {
  GroupArgs a{};
  a.views = G->d_views;
  a.slot_start = P.d;
  a.nruns = (int)G->runs.size();
  a.slot_time = timed ? P.time : nullptr;
  a.gq = G->d_gq;
  a.glist = G->d_glist;
  a.qstart = G->d_qstart;
  a.qrun = G->d_qrun;
  a.work = G->d_work;
  a.pend = G->d_pend;
  a.work_stride = G->work_stride;
  a.pend_stride = G->pend_stride;
  return a;
}

// Re-cut a launch class after a step: each run's work is the busy time its
// workgroups summed (invariant to how many it had), so ranges proportional
// to it balance the runs' finishing times.  Runs without recorded work keep
// their previous weight.
void group_retune(mvmc_group* G, GroupPart& P) {
  const int R = (int)G->runs.size();
  std::vector<long long> t((size_t)P.start[R]);
  HIPCHK(hipMemcpyAsync(t.data(), P.time, sizeof(long long) * t.size(), hipMemcpyDeviceToHost, G->st));
  HIPCHK(hipStreamSynchronize(G->st));
  std::vector<double> w(R, 0.0);
  double tot = 0.0, told = 0.0;
  for (int r = 0; r < R; ++r) {
    for (int b = P.start[r]; b < P.start[r + 1]; ++b) w[r] += (double)t[b];
    tot += w[r];
    told += P.weight[r];
  }
  if (tot <= 0.0) return;
  for (int r = 0; r < R; ++r)
    if (w[r] <= 0.0) w[r] = P.weight[r] / told * tot;
  group_partition(G, P, P.capacity, w);
}

void group_clear_times(mvmc_group* G) {
  for (GroupPart* P : {&G->samp, &G->rb, &G->meas})
    HIPCHK(hipMemsetAsync(P->time, 0, sizeof(long long) * P->capacity, G->st));
}

// One round: clear the walker counters, run every walker until its step ends
// or its rebuild is due, rebuild the walkers handed over.
void group_round(mvmc_group* G, int out, int n_in_max) {
  hipLaunchKernelGGL(k_fill_int, dim3(1), dim3(64), 0, G->st, G->d_gq, 4, 0);
  hipLaunchKernelGGL((k_block<double, true>), dim3(G->samp.capacity), dim3(256), G->lds_samp, G->st, EnsView{},
                     group_args(G, G->samp, false), out, n_in_max);
  hipLaunchKernelGGL((k_rebuild_q<3, 8>), dim3(G->rb.capacity), dim3(256), G->lds_rebuild, G->st,
                     group_args(G, G->rb, false));
}

void group_sample(mvmc_group* G) {
  const int R = (int)G->runs.size();
  const double t0 = now_s();
  for (mvmc_ens* E : G->runs) {
    if (E->precision != 0) throw Fail(MVMC_E_INVALID_SHAPE, "group runs use the FP64 sampler");
    E->upload_tables();
  }
  group_clear_times(G);
  hipLaunchKernelGGL(k_begin<true>, dim3(G->samp.start[R]), dim3(256), G->lds_samp, G->st, EnsView{},
                     group_args(G, G->samp), 0);
  HIPCHK(hipGetLastError());
  int n_out = 0;
  for (mvmc_ens* E : G->runs) {
    std::vector<int> ws((size_t)E->v.W * 8);
    HIPCHK(hipMemcpyAsync(ws.data(), E->v.wstate, sizeof(int) * ws.size(), hipMemcpyDeviceToHost, G->st));
    HIPCHK(hipStreamSynchronize(G->st));
    for (int w = 0; w < E->v.W; ++w) n_out = std::max(n_out, ws[(size_t)w * 8 + 2]);
  }
  for (mvmc_ens* E : G->runs) E->check_status();
  hipEvent_t e0, e1;
  HIPCHK(hipEventCreate(&e0));
  HIPCHK(hipEventCreate(&e1));
  G->launches = 0;
  G->graph_replays = 0;
  double ms_total = 0;
  for (int out = 0; out < n_out; ++out) {
    for (mvmc_ens* E : G->runs)
      hipLaunchKernelGGL(k_reset_resume, dim3(nblk(E->v.W)), dim3(256), 0, G->st, E->v);
    // n_in differs per run (NVMCInterval x Nsite): each run's slots read it
    // from their view; the argument is ignored in group launches
    HIPCHK(hipEventRecord(e0, G->st));
    hipGraphExec_t exec = nullptr;
    if (G->graph_rounds > 0) {
      hipGraph_t graph = nullptr;
      HIPCHK(hipStreamBeginCapture(G->st, hipStreamCaptureModeThreadLocal));
      for (int k = 0; k < G->graph_rounds; ++k) group_round(G, out, 0);
      HIPCHK(hipStreamEndCapture(G->st, &graph));
      HIPCHK(hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0));
      (void)hipGraphDestroy(graph);
    }
    for (;;) {
      if (exec) {
        HIPCHK(hipGraphLaunch(exec, G->st));
        ++G->graph_replays;
        G->launches += G->graph_rounds;
      } else {
        group_round(G, out, 0);
        ++G->launches;
      }
      HIPCHK(hipGetLastError());
      int handed = 0;
      HIPCHK(hipMemcpyAsync(&handed, G->d_gq, sizeof(int), hipMemcpyDeviceToHost, G->st));
      HIPCHK(hipStreamSynchronize(G->st));
      if (handed == 0) break;
    }
    HIPCHK(hipEventRecord(e1, G->st));
    HIPCHK(hipEventSynchronize(e1));
    float ms = 0;
    HIPCHK(hipEventElapsedTime(&ms, e0, e1));
    ms_total += ms;
    if (exec) (void)hipGraphExecDestroy(exec);
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  G->block_ms = ms_total;
  for (mvmc_ens* E : G->runs) E->check_status();
  G->t_sample = now_s() - t0;
}

void group_measure(mvmc_group* G) {
  const int R = (int)G->runs.size();
  const double t0 = now_s();
  const int nsample = G->runs[0]->v.nsample;
  for (int s = 0; s < nsample; ++s) {
    hipLaunchKernelGGL((k_meas_rebuild<3, 8, true>), dim3(G->rb.start[R]), dim3(256), G->lds_rebuild, G->st,
                       EnsView{}, group_args(G, G->rb), s);
    hipLaunchKernelGGL(k_measure<true>, dim3(G->meas.start[R]), dim3(256), G->lds_meas, G->st, EnsView{},
                       group_args(G, G->meas), 1, s);
    HIPCHK(hipGetLastError());
  }
  HIPCHK(hipStreamSynchronize(G->st));
  for (mvmc_ens* E : G->runs) measure_finish(E);
  G->t_measure = now_s() - t0;
}

}  // namespace

extern "C" {

const char* mvmc_last_error(void) { return g_err.c_str(); }
int mvmc_abi_version(void) { return 1; }

int mvmc_pfinv_f64(const double* planes, int n, int count, double* pf, double* inv, int* status, void* stream) {
  return guarded([&] {
    require_gpu();
    if (n < 2 || (n % 2) != 0 || count < 0) throw Fail(MVMC_E_INVALID_SHAPE, "n must be even and >= 2");
    if (count == 0) return;
    if (!planes || !pf || !inv || !status) throw Fail(MVMC_E_INVALID_SHAPE, "null buffer");
    hipStream_t st = (hipStream_t)stream;
    if (rr_batched(n)) {
      launch_pfinv_rr(planes, inv, n, count, pf, status, st);
      return;
    }
    if ((const double*)inv != planes)
      HIPCHK(hipMemcpyAsync(inv, planes, sizeof(double) * n * n * count, hipMemcpyDeviceToDevice, st));
    const int np = n <= 32 ? 8 : 4;  // the inverse runs one wave per plane: 4 planes fill a workgroup
    const int blocks = (count + np - 1) / np;
    double* work = nullptr;
    HIPCHK(hipMallocAsync((void**)&work, sizeof(double) * (size_t)blocks * pfinv_work_doubles(n, np), st));
    size_t lds = pf_bytes(n, np) + 64;
    raise_lds_limit((const void*)k_pfinv_batched, lds);
    hipLaunchKernelGGL(k_pfinv_batched, dim3(blocks), dim3(256), lds, st, inv, work, n, count, np, blocks, pf,
                       status);
    HIPCHK(hipGetLastError());
    HIPCHK(hipFreeAsync(work, st));
  });
}

int mvmc_pfinv_f64_inplace(double* planes, int n, int count, double* pf, int* status, void* stream) {
  return mvmc_pfinv_f64(planes, n, count, pf, planes, status, stream);
}

int mvmc_pfinv_f64_noalloc(const double* planes, int n, int count, double* pf, double* inv, int* status,
                           void* stream) {
  return guarded([&] {
    require_gpu();
    if (n < 2 || (n % 2) != 0 || count < 0) throw Fail(MVMC_E_INVALID_SHAPE, "n must be even and >= 2");
    if (n > NOALLOC_MAX_N)
      throw Fail(MVMC_E_INVALID_SHAPE, "n = " + std::to_string(n) + " exceeds the no-allocation maximum " +
                                           std::to_string(NOALLOC_MAX_N));
    if (count == 0) return;
    if (!planes || !pf || !inv || !status) throw Fail(MVMC_E_INVALID_SHAPE, "null buffer");
    double* work = noalloc_scratch();
    hipStream_t st = (hipStream_t)stream;
    if ((const double*)inv != planes)
      HIPCHK(hipMemcpyAsync(inv, planes, sizeof(double) * n * n * count, hipMemcpyDeviceToDevice, st));
    const int np = n <= 32 ? 8 : 4;
    const int groups = (count + np - 1) / np;
    // workgroups = work slots that fit in the scratch array (at least one:
    // NOALLOC_MAX_N keeps a slot within it), at most one per plane group
    const size_t per_slot = pfinv_work_doubles(n, np);
    const int slots = (int)std::min<size_t>({(size_t)groups, NOALLOC_SCRATCH_DOUBLES / per_slot,
                                             (size_t)NOALLOC_MAX_SLOTS});
    size_t lds = pf_bytes(n, np) + 64;
    raise_lds_limit((const void*)k_pfinv_batched, lds);
    hipLaunchKernelGGL(k_pfinv_batched, dim3(slots), dim3(256), lds, st, inv, work, n, count, np, groups, pf,
                       status);
    HIPCHK(hipGetLastError());
  });
}

int mvmc_pfinv_c128(const double* planes, int n, int count, int order, double* pf, double* inv, int* status,
                    void* stream) {
  return guarded([&] {
    require_gpu();
    if (n < 2 || (n % 2) != 0 || n > PFC_NMAX || count < 0)
      throw Fail(MVMC_E_INVALID_SHAPE, "n must be even, 2 <= n <= 512");
    if (order != 0 && order != 1) throw Fail(MVMC_E_INVALID_SHAPE, "order must be 0 (zsktf2) or 1 (C order)");
    if (count == 0) return;
    if (!planes || !pf || !inv || !status) throw Fail(MVMC_E_INVALID_SHAPE, "null buffer");
    if (((uintptr_t)planes | (uintptr_t)pf | (uintptr_t)inv) % 16 != 0)
      throw Fail(MVMC_E_INVALID_SHAPE, "complex buffers must be 16-byte aligned");
    hipStream_t st = (hipStream_t)stream;
    const size_t nn = (size_t)n * n;
    if (rr_enabled() && n <= RR_C_NMAX) {
      launch_pfinv_rr_c128((const cz*)planes, (cz*)inv, n, count, order, (cz*)pf, status, st);
      return;
    }
    if ((const double*)inv != planes)
      HIPCHK(hipMemcpyAsync(inv, planes, sizeof(cz) * nn * count, hipMemcpyDeviceToDevice, st));
    int dev = 0, cus = 0;
    HIPCHK(hipGetDevice(&dev));
    HIPCHK(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
    const int grid = std::min(count, 4 * std::max(cus, 1));  // planes stride over a resident grid
    cz* work = nullptr;
    HIPCHK(hipMallocAsync((void**)&work, sizeof(cz) * 2 * nn * grid, st));
    hipLaunchKernelGGL(k_pfinv_c128, dim3(grid), dim3(256), 0, st, (cz*)inv, work, n, count, order, (cz*)pf,
                       status);
    HIPCHK(hipGetLastError());
    HIPCHK(hipFreeAsync(work, st));
  });
}

int mvmc_ens_create(const char* namelist, const char* initial_def, int walkers, long long seed_override,
                    int use_seed_override, mvmc_ens** out) {
  if (!out) return MVMC_E_INVALID_SHAPE;
  *out = nullptr;
  mvmc_ens* E = new mvmc_ens();
  int rc = guarded([&] { create_impl(E, namelist, initial_def, walkers, seed_override, use_seed_override); });
  if (rc != MVMC_OK) {
    delete E;
    return rc;
  }
  *out = E;
  return MVMC_OK;
}

int mvmc_ens_destroy(mvmc_ens* e) {
  delete e;
  return MVMC_OK;
}

int mvmc_ens_dims(mvmc_ens* e, int* d) {
  const EnsView& v = e->v;
  int vals[16] = {v.ns, v.ne, v.n, v.nqp, v.nproj, v.nslater, v.size, v.nsample, v.W, v.ex_path, v.interval,
                  v.warmup, e->grid_samp, (int)e->lds_samp, (int)e->lds_meas, e->grid_meas};
  std::memcpy(d, vals, sizeof(vals));
  return MVMC_OK;
}

int mvmc_ens_get_params(mvmc_ens* e, double* proj, double* slater) {
  std::copy(e->m.proj.begin(), e->m.proj.end(), proj);
  std::copy(e->m.slater.begin(), e->m.slater.end(), slater);
  return MVMC_OK;
}

int mvmc_ens_set_params(mvmc_ens* e, const double* proj, const double* slater) {
  std::copy(proj, proj + e->m.n_proj(), e->m.proj.begin());
  std::copy(slater, slater + e->m.n_slater, e->m.slater.begin());
  return MVMC_OK;
}

int mvmc_ens_sample(mvmc_ens* e) { return guarded([&] { sample_impl(e); }); }
int mvmc_ens_measure(mvmc_ens* e) { return guarded([&] { measure_impl(e); }); }
int mvmc_ens_measure_configs(mvmc_ens* e, int count, const int* ele_idx, double* e_loc, double* o, int* valid) {
  return guarded([&] { measure_configs_impl(e, count, ele_idx, e_loc, o, valid); });
}
int mvmc_ens_sr_cg(mvmc_ens* e, long long* info) {
  long long dummy = 0;
  return guarded([&] { sr_impl(e, info ? info : &dummy); });
}

int mvmc_ens_sr_direct(mvmc_ens* e, int* skipped) {
  return guarded([&] {
    long long info = 0;
    const int rc = sr_impl(e, &info, true);
    if (skipped) *skipped = rc;
  });
}

int mvmc_ens_opt_step(mvmc_ens* e, double* row) {
  return guarded([&] {
    sample_impl(e);
    measure_impl(e);
    finish_opt_step(e, row);
  });
}

// output_opt_data (crates/mvmc-core/src/io.rs:561-727): {head}_opt.dat from the
// stored window in C's "% .18e" -- one snapshot: (value, 0) pairs; more:
// (mean.re, mean.im, sample standard deviation) triples, mean summed in
// window order; then {head}_{gutzwiller,jastrow,orbital}_opt.dat blocks of
// the means.
int mvmc_ens_set_output_dir(mvmc_ens* e, const char* out_dir) {
  return guarded([&] {
    e->out_enabled = out_dir != nullptr;
    e->out_dir = out_dir ? out_dir : "";
  });
}

int mvmc_ens_write_opt_data(mvmc_ens* e, const char* out_dir) {
  return guarded([&] {
    const mvmc::Model& m = e->m;
    const size_t len = e->opt_win.size();
    bool complete = len > 0;
    for (char c : e->opt_win_set) complete = complete && c;
    if (!complete) throw Fail(MVMC_E_INPUT, "incomplete optimization window");
    const std::string dir = out_dir && *out_dir ? std::string(out_dir) + "/" : std::string();
    if (out_dir && *out_dir) {  // output_path creates the directory (io.rs:127-140)
      std::error_code ec;
      std::filesystem::create_directories(out_dir, ec);
    }
    const std::string head = m.mp.c_para_file_head.empty() ? "zqp" : m.mp.c_para_file_head;
    const size_t width = e->opt_win[0].size();
    auto value = [&](size_t w, size_t idx) {
      return idx == 0 ? e->opt_win_e[w] : idx == 1 ? e->opt_win_e2[w] : e->opt_win[w][idx - 2];
    };
    auto cfmt = [](double x) {
      char b[64];
      std::snprintf(b, sizeof b, "% .18e", x);
      return std::string(b);
    };
    const std::string path = dir + head + "_opt.dat";
    FILE* f = std::fopen(path.c_str(), "w");
    if (!f) throw Fail(MVMC_E_INPUT, "cannot write " + path);
    std::vector<double> mean(width + 2);
    if (len == 1) {
      for (size_t idx = 0; idx < width + 2; ++idx)
        std::fprintf(f, "%s %s ", cfmt(value(0, idx)).c_str(), cfmt(0.0).c_str());
    } else {
      for (size_t idx = 0; idx < width + 2; ++idx) {
        double s = 0.0;
        for (size_t w = 0; w < len; ++w) s += value(w, idx);
        const double mu = s / (double)len;
        double var = 0.0;
        for (size_t w = 0; w < len; ++w) {
          const double d = value(w, idx) - mu;
          var += d * d;
        }
        mean[idx] = mu;
        std::fprintf(f, "%s %s %s ", cfmt(mu).c_str(), cfmt(0.0).c_str(),
                     cfmt(std::sqrt(var / ((double)len - 1.0))).c_str());
      }
    }
    std::fprintf(f, "\n");
    std::fclose(f);
    if (len == 1) return;
    struct Block {
      const char* suffix;
      const char* label;
      size_t count;
    };
    const Block blocks[3] = {{"gutzwiller", "NGutzwillerIdx", (size_t)m.n_gutz},
                             {"jastrow", "NJastrowIdx", (size_t)m.n_jast},
                             {"orbital", "NOrbitalIdx", (size_t)m.n_slater}};
    size_t off = 2;
    for (const Block& b : blocks) {
      if (b.count > 0) {
        const std::string bp = dir + head + "_" + b.suffix + "_opt.dat";
        FILE* g = std::fopen(bp.c_str(), "w");
        if (!g) throw Fail(MVMC_E_INPUT, "cannot write " + bp);
        const char* sep = "======================";
        std::fprintf(g, "%s\n%s  %zu\n%s\n%s\n%s\n", sep, b.label, b.count, sep, sep, sep);
        for (size_t i = 0; i < b.count; ++i)
          std::fprintf(g, "%zu %s %s \n", i, cfmt(mean[off + i]).c_str(), cfmt(0.0).c_str());
        std::fclose(g);
      }
      off += b.count;
    }
  });
}

int mvmc_ens_get_energy(mvmc_ens* e, double* o) {
  o[0] = e->wc; o[1] = e->etot; o[2] = e->etot2; o[3] = e->sztot; o[4] = e->sztot2;
  return MVMC_OK;
}

int mvmc_ens_get_operands(mvmc_ens* e, double* oo, double* ho) {
  if (oo) std::copy(e->oo.begin(), e->oo.end(), oo);
  if (ho) std::copy(e->ho.begin(), e->ho.end(), ho);
  return MVMC_OK;
}

int mvmc_ens_get_o_store(mvmc_ens* e, int w, double* o) {
  return guarded([&] {
    const EnsView& v = e->v;
    if (w < 0 || w >= v.W) throw Fail(MVMC_E_INVALID_SHAPE, "walker out of range");
    e->d2h(o, v.o_store + (size_t)w * v.nsample * v.size, sizeof(double) * v.nsample * v.size);
  });
}

int mvmc_ens_get_walker(mvmc_ens* e, int w, int* saved, long long* counters, uint32_t* rng) {
  return guarded([&] {
    const EnsView& v = e->v;
    if (w < 0 || w >= v.W) throw Fail(MVMC_E_INVALID_SHAPE, "walker out of range");
    if (saved)
      e->d2h(saved, v.s_idx + (size_t)w * v.nsample * v.n, sizeof(int) * v.nsample * v.n);
    if (counters) e->d2h(counters, v.counter + (size_t)w * 4, sizeof(long long) * 4);
    if (rng) e->d2h(rng, v.rng + (size_t)w * 625, sizeof(uint32_t) * 625);
  });
}

int mvmc_ens_get_tables(mvmc_ens* e, int w, double* inv, double* pf) {
  return guarded([&] {
    const EnsView& v = e->v;
    if (w < 0 || w >= v.W) throw Fail(MVMC_E_INVALID_SHAPE, "walker out of range");
    const size_t nn = (size_t)v.n * v.n;
    if (inv && e->precision == 1) {  // the sampler's FP32 inverse, widened
      std::vector<float> h((size_t)v.nqp * nn);
      e->d2h(h.data(), e->inv32 + (size_t)w * v.nqp * nn, sizeof(float) * h.size());
      for (size_t k = 0; k < h.size(); ++k) inv[k] = h[k];
    } else if (inv) {
      e->d2h(inv, v.inv + (size_t)w * v.nqp * nn, sizeof(double) * v.nqp * nn);
    }
    if (pf) e->d2h(pf, v.pf + (size_t)w * v.nqp, sizeof(double) * v.nqp);
  });
}

int mvmc_ens_get_config(mvmc_ens* e, int w, int* ei, int* ec, int* en, int* pc) {
  return guarded([&] {
    const EnsView& v = e->v;
    if (w < 0 || w >= v.W) throw Fail(MVMC_E_INVALID_SHAPE, "walker out of range");
    if (ei) e->d2h(ei, v.ele_idx + (size_t)w * v.n, sizeof(int) * v.n);
    if (ec) e->d2h(ec, v.ele_cfg + (size_t)w * v.n2, sizeof(int) * v.n2);
    if (en) e->d2h(en, v.ele_num + (size_t)w * v.n2, sizeof(int) * v.n2);
    if (pc && v.nproj > 0)
      e->d2h(pc, v.proj_cnt + (size_t)w * v.nproj, sizeof(int) * v.nproj);
  });
}

int mvmc_ens_get_cg(mvmc_ens* e, double* x, int* n_active) {
  if (x) std::copy(e->cg_x.begin(), e->cg_x.end(), x);
  if (n_active) *n_active = (int)e->cg_x.size();
  return (int)e->cg_info;
}

int mvmc_ens_log_psi(mvmc_ens* e, double* out) {
  return guarded([&] {
    EnsView& v = e->v;
    const size_t rows = (size_t)v.W * v.nsample;
    if (!e->d_rows) e->d_rows = dalloc<double>(e->allocs, rows);
    e->upload_tables();
    for (int s = 0; s < v.nsample; ++s) {
      launch_meas_rebuild(e, s);
      hipLaunchKernelGGL(k_log_psi, dim3(nblk(v.W)), dim3(256), 0, e->st, v, s, e->d_rows);
    }
    HIPCHK(hipGetLastError());
    HIPCHK(hipMemcpyAsync(out, e->d_rows, sizeof(double) * rows, hipMemcpyDeviceToHost, e->st));
    HIPCHK(hipStreamSynchronize(e->st));
  });
}

int mvmc_ens_vjp(mvmc_ens* e, const double* cot, double* grad) {
  return guarded([&] {
    EnsView& v = e->v;
    const size_t rows = (size_t)v.W * v.nsample;
    if (!e->d_rows) e->d_rows = dalloc<double>(e->allocs, rows);
    if (!e->d_grad) e->d_grad = dalloc<double>(e->allocs, (size_t)v.size);
    HIPCHK(hipMemcpyAsync(e->d_rows, cot, sizeof(double) * rows, hipMemcpyHostToDevice, e->st));
    hipLaunchKernelGGL(k_vjp, dim3(nblk(v.size)), dim3(256), 0, e->st, v, e->d_rows, e->d_grad);
    HIPCHK(hipGetLastError());
    HIPCHK(hipMemcpyAsync(grad, e->d_grad, sizeof(double) * v.size, hipMemcpyDeviceToHost, e->st));
    HIPCHK(hipStreamSynchronize(e->st));
  });
}

__global__ void k_fill64(double* p, long long n, double v) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}

// Orbital map of the adjoint VJP, built once per handle: for every plane q
// (translation q / NSPGaussLeg, Gauss-Legendre point q % NSPGaussLeg) and
// Slater-table position (a, b) = (site + spin Nsite, site' + spin' Nsite)
// the orbital the position's element derives from and its weight, the same
// factors the per-sample assembly applies (translation signs, orbital sign,
// cos/sin products of the spin block), grouped by orbital in ascending
// (q, a, b) order.
void adj_prepare(mvmc_ens* e) {
  EnsView& v = e->v;
  if (e->d_sbar) return;
  int dev = 0, ncu = 0;
  HIPCHK(hipGetDevice(&dev));
  HIPCHK(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev));
  const size_t pl = (size_t)v.n2 * v.n2, per = (size_t)v.nqp * pl;
  const size_t cap = ((size_t)4 << 30) / sizeof(double);  // at most 4 GiB of slot accumulators
  e->adj_slots = (int)std::max<size_t>(1, std::min<size_t>({(size_t)v.W, (size_t)std::max(ncu, 1), cap / per}));
  e->d_sbar = dalloc<double>(e->allocs, (size_t)e->adj_slots * per);
  e->d_gproj = dalloc<double>(e->allocs, (size_t)e->adj_slots * (1 + v.nproj));
  const mvmc::Model& m = e->m;
  std::vector<int> ts, tg;
  mvmc::translation_maps(m, ts, tg);
  const int ns = v.ns, nleg = std::max(v.nleg, 1);
  std::vector<int> cnt((size_t)v.nslater + 1, 0);
  auto visit = [&](auto&& f) {
    for (int q = 0; q < v.nqp; ++q) {
      const int tr = q / nleg, sp = q % nleg;
      const double cs = m.spgl_cos_sin[sp], cc = m.spgl_cos_cos[sp], ss = m.spgl_sin_sin[sp];
      for (int a = 0; a < v.n2; ++a)
        for (int b = 0; b < v.n2; ++b) {
          const int sa = a % ns, sb = b % ns;
          const bool up_a = a < ns, up_b = b < ns;
          const int ti = ts[(size_t)tr * ns + sa], tj = ts[(size_t)tr * ns + sb];
          const int orb = m.orb_idx[(size_t)ti * ns + tj];
          if (orb < 0 || orb >= v.nslater) continue;
          const int sg = tg[(size_t)tr * ns + sa] * tg[(size_t)tr * ns + sb] * m.orb_sgn[(size_t)ti * ns + tj];
          const double fac = up_b ? (up_a ? cs : ss) : -(up_a ? cc : cs);
          f(orb, (long long)q * (long long)pl + (long long)a * v.n2 + b, fac * (double)sg);
        }
    }
  };
  visit([&](int orb, long long, double) { ++cnt[(size_t)orb + 1]; });
  for (int o = 0; o < v.nslater; ++o) cnt[(size_t)o + 1] += cnt[o];
  if (cnt[v.nslater] > INT32_MAX) throw Fail(MVMC_E_INVALID_SHAPE, "orbital map too large for the adjoint VJP");
  std::vector<long long> ent((size_t)std::max(cnt[v.nslater], 1));
  std::vector<double> wt(ent.size());
  std::vector<int> fill(cnt.begin(), cnt.end() - 1);
  visit([&](int orb, long long k, double w) {
    const int t = fill[orb]++;
    ent[t] = k;
    wt[t] = w;
  });
  e->d_adj_start = dupload(e->allocs, cnt, e->st);
  e->d_adj_ent = dupload(e->allocs, ent, e->st);
  e->d_adj_wt = dupload(e->allocs, wt, e->st);
}

int mvmc_ens_vjp_adjoint(mvmc_ens* e, const double* cot, double* grad) {
  return guarded([&] {
    require_gpu();
    EnsView& v = e->v;
    if (!cot || !grad) throw Fail(MVMC_E_INVALID_SHAPE, "null buffer");
    const size_t rows = (size_t)v.W * v.nsample;
    if (!e->d_rows) e->d_rows = dalloc<double>(e->allocs, rows);
    if (!e->d_grad) e->d_grad = dalloc<double>(e->allocs, (size_t)v.size);
    adj_prepare(e);
    e->upload_tables();
    HIPCHK(hipMemcpyAsync(e->d_rows, cot, sizeof(double) * rows, hipMemcpyHostToDevice, e->st));
    const long long len = (long long)v.nqp * v.n2 * v.n2, glen = 1 + v.nproj;
    const long long tot = len * e->adj_slots;
    hipLaunchKernelGGL(k_fill64, dim3((unsigned)((tot + 255) / 256)), dim3(256), 0, e->st, e->d_sbar, tot, 0.0);
    hipLaunchKernelGGL(k_fill64, dim3((unsigned)((glen * e->adj_slots + 255) / 256)), dim3(256), 0, e->st,
                       e->d_gproj, glen * e->adj_slots, 0.0);
    hipEvent_t ev[3];
    for (auto& x : ev) HIPCHK(hipEventCreate(&x));
    const size_t lds = 2 * rnd16((size_t)v.n * 4) + rnd16((size_t)v.nqp * 8);
    float rb = 0, tot_ms = 0;
    HIPCHK(hipEventRecord(ev[0], e->st));
    for (int s = 0; s < v.nsample; ++s) {
      HIPCHK(hipEventRecord(ev[1], e->st));
      launch_meas_rebuild(e, s);
      HIPCHK(hipEventRecord(ev[2], e->st));
      hipLaunchKernelGGL(k_adj_scatter, dim3(e->adj_slots), dim3(256), lds, e->st, v, s, e->d_rows, e->d_sbar,
                         e->d_gproj);
      HIPCHK(hipGetLastError());
      HIPCHK(hipEventSynchronize(ev[2]));
      float a = 0;
      HIPCHK(hipEventElapsedTime(&a, ev[1], ev[2]));
      rb += a;
    }
    hipLaunchKernelGGL(k_adj_fold, dim3((unsigned)((len + 255) / 256)), dim3(256), 0, e->st, e->d_sbar,
                       e->adj_slots, len);
    hipLaunchKernelGGL(k_adj_fold, dim3((unsigned)((glen + 255) / 256)), dim3(256), 0, e->st, e->d_gproj,
                       e->adj_slots, glen);
    hipLaunchKernelGGL(k_adj_orbital, dim3(nblk(v.size)), dim3(256), 0, e->st, e->d_sbar, e->d_adj_start,
                       e->d_adj_ent, e->d_adj_wt, e->d_gproj, v.nproj, v.nslater, e->d_grad);
    HIPCHK(hipGetLastError());
    HIPCHK(hipEventRecord(ev[1], e->st));
    HIPCHK(hipEventSynchronize(ev[1]));
    HIPCHK(hipEventElapsedTime(&tot_ms, ev[0], ev[1]));
    for (auto& x : ev) (void)hipEventDestroy(x);
    e->adj_ms = tot_ms;
    e->adj_rebuild_ms = rb;
    HIPCHK(hipMemcpyAsync(grad, e->d_grad, sizeof(double) * v.size, hipMemcpyDeviceToHost, e->st));
    HIPCHK(hipStreamSynchronize(e->st));
  });
}

int mvmc_ens_set_precision(mvmc_ens* e, int mode) {
  return guarded([&] {
    if (mode != 0 && mode != 1) throw Fail(MVMC_E_INVALID_SHAPE, "precision form must be 0 (FP64) or 1 (mixed)");
    if (mode == 1 && !e->inv32)
      e->inv32 = dalloc<float>(e->allocs, (size_t)e->v.W * e->v.nqp * e->v.n * e->v.n);
    e->precision = mode;
    e->v.inv32 = mode == 1 ? e->inv32 : nullptr;
  });
}

int mvmc_ens_phase_clocks(mvmc_ens* e, long long* out) {
  return guarded([&] {
    if (!e->v.clk) throw Fail(MVMC_E_UNAVAILABLE, "library built without MVMC_PHASE_CLOCKS");
    std::vector<long long> h((size_t)e->slots * 32);
    e->d2h(h.data(), e->v.clk, sizeof(long long) * h.size());
    for (int k = 0; k < 32; ++k) out[k] = 0;
    for (int sl = 0; sl < e->slots; ++sl)
      for (int k = 0; k < 32; ++k) out[k] += h[(size_t)sl * 32 + k];
  });
}

int mvmc_ens_timings(mvmc_ens* e, double* t) {
  t[0] = e->t_sample; t[1] = e->t_measure; t[2] = e->t_sr; t[3] = e->proposals; t[4] = e->accepted;
  t[5] = e->block_ms; t[6] = e->block_launches; t[7] = e->rebuild_ms;
  t[8] = e->rebuilt; t[9] = e->meas_rebuild_ms; t[10] = e->measure_ms; t[11] = e->measured;
  t[12] = e->cg_ms; t[13] = e->cg_iter; t[14] = e->cg_na; t[15] = e->cg_rows;
  t[16] = e->rb_variant;
  t[17] = e->rb_walkers[0] > 0 ? e->rb_ms[0] / e->rb_walkers[0] : 0.0;
  t[18] = e->rb_walkers[1] > 0 ? e->rb_ms[1] / e->rb_walkers[1] : 0.0;
  t[19] = e->rb_walkers[2] > 0 ? e->rb_ms[2] / e->rb_walkers[2] : 0.0;
  t[20] = e->adj_ms; t[21] = e->adj_rebuild_ms; t[22] = e->gram_ms; t[23] = e->block_write_passes;
  t[24] = e->block_flush_passes; t[25] = 0.0;
  return MVMC_OK;
}

int mvmc_ens_launch_log(mvmc_ens* e, double* out, int max_rows) {
  const int rows = (int)(e->launch_log.size() / 4);
  for (int i = 0; i < std::min(rows, max_rows) * 4; ++i) out[i] = e->launch_log[(size_t)i];
  return rows;
}

int mvmc_ens_status(mvmc_ens* e) {
  int s = -1;
  if (hipMemcpyAsync(&s, e->v.status, sizeof(int), hipMemcpyDeviceToHost, e->st) != hipSuccess ||
      hipStreamSynchronize(e->st) != hipSuccess)
    return -1;
  return s;
}

// One rocBLAS handle for the standalone SR stages, created on first use.
rocblas_handle sr_blas() {
  static std::mutex mu;
  static rocblas_handle h = nullptr;
  std::lock_guard<std::mutex> lock(mu);
  if (!h) BLASCHK(rocblas_create_handle(&h));
  return h;
}

int mvmc_sr_gram_real(int n, int samples, const double* store, double* out, void* stream) {
  return guarded([&] {
    require_gpu();
    if (n < 0 || samples < 0) throw Fail(MVMC_E_INVALID_SHAPE, "bad Gram dimensions");
    if (n == 0) return;
    if (!out || (!store && samples > 0)) throw Fail(MVMC_E_INVALID_SHAPE, "null buffer");
    hipLaunchKernelGGL(k_gram_mfma<double>, dim3((unsigned)gm_tiles(n)), dim3(256), 0, (hipStream_t)stream,
                       store, samples, n, 1.0, out, 1);
    HIPCHK(hipGetLastError());
  });
}

int mvmc_sr_gram_real_f32(int n, int samples, const float* store, double* out, void* stream) {
  return guarded([&] {
    require_gpu();
    if (n < 0 || samples < 0) throw Fail(MVMC_E_INVALID_SHAPE, "bad Gram dimensions");
    if (n == 0) return;
    if (!out || (!store && samples > 0)) throw Fail(MVMC_E_INVALID_SHAPE, "null buffer");
    hipLaunchKernelGGL(k_gram_mfma<float>, dim3((unsigned)gm_tiles(n)), dim3(256), 0, (hipStream_t)stream,
                       store, samples, n, 1.0, out, 1);
    HIPCHK(hipGetLastError());
  });
}

int mvmc_sr_cholesky_solve(int n, double* s, double* rhs, int* info, void* stream) {
  return guarded([&] {
    require_gpu();
    if (n < 0) throw Fail(MVMC_E_INVALID_SHAPE, "bad dimension");
    if (!info) throw Fail(MVMC_E_INVALID_SHAPE, "null info");
    *info = 0;
    if (n == 0) return;
    if (!s || !rhs) throw Fail(MVMC_E_INVALID_SHAPE, "null buffer");
    hipStream_t st = (hipStream_t)stream;
    rocblas_handle h = sr_blas();
    BLASCHK(rocblas_set_stream(h, st));
    int* d_info = nullptr;
    HIPCHK(hipMallocAsync((void**)&d_info, sizeof(int), st));
    BLASCHK(rocsolver_dpotrf(h, rocblas_fill_upper, n, s, n, d_info));
    int fi = 0;
    HIPCHK(hipMemcpyAsync(&fi, d_info, sizeof(int), hipMemcpyDeviceToHost, st));
    HIPCHK(hipFreeAsync(d_info, st));
    HIPCHK(hipStreamSynchronize(st));
    if (fi != 0) {
      *info = fi;
      return;
    }
    BLASCHK(rocsolver_dpotrs(h, rocblas_fill_upper, n, 1, s, n, rhs, n));
    std::vector<double> x(n);
    HIPCHK(hipMemcpyAsync(x.data(), rhs, sizeof(double) * n, hipMemcpyDeviceToHost, st));
    HIPCHK(hipStreamSynchronize(st));
    for (double v : x)
      if (!std::isfinite(v)) {
        *info = -1;
        return;
      }
  });
}

int mvmc_sr_assemble_s_g(const double* oo, int oo_complex, const double* ho, int ho_complex, const int* map, int nm,
                         long long ld, int offset, double sta_del, double step_dt, double* s, double* g,
                         void* stream) {
  return guarded([&] {
    require_gpu();
    if (nm < 0 || ld < 0 || offset < 0) throw Fail(MVMC_E_INVALID_SHAPE, "bad S/g dimensions");
    if (nm == 0) return;
    if (!oo || !ho || !map || !s || !g) throw Fail(MVMC_E_INVALID_SHAPE, "null buffer");
    const long long total = (long long)nm * nm + nm;
    hipLaunchKernelGGL(k_sr_assemble_stage, dim3(nblk(total)), dim3(256), 0, (hipStream_t)stream, oo,
                       oo_complex ? 1 : 0, ho, ho_complex ? 1 : 0, map, nm, ld, offset, 1.0 + sta_del,
                       -2.0 * step_dt, s, g);
    HIPCHK(hipGetLastError());
  });
}

int mvmc_sr_gram_complex(int n, int samples, const double* store, double* out, void* stream) {
  return guarded([&] {
    require_gpu();
    if (n < 0 || samples < 0) throw Fail(MVMC_E_INVALID_SHAPE, "bad Gram dimensions");
    if (n == 0) return;
    if (!out || (!store && samples > 0)) throw Fail(MVMC_E_INVALID_SHAPE, "null buffer");
    const int tiles = (n + GR_T - 1) / GR_T;
    hipLaunchKernelGGL(k_gram_complex, dim3(tiles, tiles), dim3(256), 0, (hipStream_t)stream, store, n, samples,
                       out);
    HIPCHK(hipGetLastError());
  });
}

int mvmc_sr_s_g(const double* o, int ns, int np, const double* e, const double* w, double* s, double* g,
                void* stream) {
  return guarded([&] {
    require_gpu();
    if (ns < 0 || np < 0) throw Fail(MVMC_E_INVALID_SHAPE, "bad S/g dimensions");
    if (np == 0) return;
    if (!o || !e || !w || !s || !g) throw Fail(MVMC_E_INVALID_SHAPE, "null buffer");
    hipStream_t st = (hipStream_t)stream;
    double* d = nullptr;
    HIPCHK(hipMallocAsync((void**)&d, sizeof(double) * ((size_t)ns * np + ns + 1), st));
    double* de = d + (size_t)ns * np;
    hipLaunchKernelGGL(k_ssg_center, dim3(nblk((long long)np + 1)), dim3(256), 0, st, o, ns, np, e, w, d, de);
    const int tiles = (np + GR_T - 1) / GR_T;
    hipLaunchKernelGGL(k_ssg_s, dim3(tiles, tiles), dim3(256), 0, st, d, w, ns, np, s);
    hipLaunchKernelGGL(k_ssg_g, dim3(nblk(np)), dim3(256), 0, st, d, w, de, ns, np, g);
    HIPCHK(hipGetLastError());
    HIPCHK(hipFreeAsync(d, st));
  });
}

int mvmc_sr_cg_apply(int na, int walkers, int nsample, const double* pack, const double* mean, const double* diag,
                     const double* x, double inv_w, double shift, double* z, void* stream) {
  static CgWork wk;
  return guarded([&] {
    require_gpu();
    if (na < 1 || walkers < 1 || nsample < 0) throw Fail(MVMC_E_INVALID_SHAPE, "bad CG dimensions");
    if ((!pack && nsample > 0) || !mean || !diag || !x || !z) throw Fail(MVMC_E_INVALID_SHAPE, "null buffer");
    hipStream_t st = (hipStream_t)stream;
    wk.ensure(na, walkers * nsample);
    HIPCHK(hipMemsetAsync(wk.ctl, 0, sizeof(int) * 4, st));
    cg_apply_dev(na, walkers * nsample, pack, mean, diag, x, (const double*)nullptr, z, inv_w, shift, wk, st);
    HIPCHK(hipGetLastError());
  });
}

int mvmc_sr_cg_solve(int na, int walkers, int nsample, const double* pack, const double* mean, const double* diag,
                     const double* g, double inv_w, double shift, double tol, int max_it, int variant, double* x,
                     long long* status, void* stream) {
  static CgWork wk;
  return guarded([&] {
    require_gpu();
    if (na < 1 || walkers < 1 || nsample < 0) throw Fail(MVMC_E_INVALID_SHAPE, "bad CG dimensions");
    *status = cg_solve_dev(na, walkers, nsample, pack, mean, diag, g, inv_w, shift, tol, max_it, variant != 0, x, wk,
                           (hipStream_t)stream);
  });
}

int mvmc_group_create(int nruns, const char* const* namelists, const char* const* initial_defs, const int* walkers,
                      mvmc_group** out) {
  if (!out || !namelists || !walkers) return MVMC_E_INVALID_SHAPE;
  *out = nullptr;
  mvmc_group* G = new mvmc_group();
  int rc = guarded([&] { group_create_impl(G, nruns, namelists, initial_defs, walkers); });
  if (rc != MVMC_OK) {
    delete G;
    return rc;
  }
  *out = G;
  return MVMC_OK;
}

int mvmc_group_destroy(mvmc_group* g) {
  delete g;
  return MVMC_OK;
}

int mvmc_group_run(mvmc_group* g, int r, mvmc_ens** out) {
  if (!g || !out || r < 0 || r >= (int)g->runs.size()) return MVMC_E_INVALID_SHAPE;
  *out = g->runs[r];
  return MVMC_OK;
}

int mvmc_group_set_autotune(mvmc_group* g, int on) {
  if (!g) return MVMC_E_INVALID_SHAPE;
  g->autotune = on ? 1 : 0;
  return MVMC_OK;
}

int mvmc_group_set_graph(mvmc_group* g, int rounds) {
  if (!g || rounds < 0) return MVMC_E_INVALID_SHAPE;
  g->graph_rounds = rounds;
  return MVMC_OK;
}

int mvmc_group_opt_step(mvmc_group* g, double* rows) {
  return guarded([&] {
    group_sample(g);
    group_measure(g);
    const double t0 = now_s();
    for (size_t r = 0; r < g->runs.size(); ++r) finish_opt_step(g->runs[r], rows ? rows + 6 * r : nullptr);
    g->t_sr = now_s() - t0;
    if (g->autotune) {
      group_retune(g, g->samp);
      group_retune(g, g->rb);
      group_retune(g, g->meas);
    }
  });
}

int mvmc_group_sample(mvmc_group* g) { return guarded([&] { group_sample(g); }); }

int mvmc_group_timings(mvmc_group* g, double* t) {
  const int R = (int)g->runs.size();
  t[0] = g->t_sample; t[1] = g->t_measure; t[2] = g->t_sr; t[3] = g->block_ms; t[4] = g->launches;
  t[5] = g->graph_replays; t[6] = g->samp.start[R]; t[7] = g->rb.start[R]; t[8] = g->meas.start[R];
  t[9] = (double)g->lds_samp;
  return MVMC_OK;
}

}  // extern "C"
