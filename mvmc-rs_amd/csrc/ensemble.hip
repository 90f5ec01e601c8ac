// Walker-ensemble kernels of the MI355X sampling + SR path.
//
// One workgroup advances one walker (an independent Markov chain with its
// own SFMT19937 stream); grids are persistent (grid = resident capacity,
// slot = blockIdx.x owns a scratch slot) and loop over walkers.  Inside a
// walker every floating-point value is produced by the operation sequence of
// the reference's C-order CPU path, so decisions, configurations and tables
// reproduce that path bit for bit:
//   candidates / Metropolis   crates/mvmc-core/src/sampling/{candidate.rs:137-385,metropolis.rs:42-83}
//   projection counters       crates/mvmc-core/src/sampling/projection.rs:278-416
//   Pfaffian ratios/updates   crates/mvmc-core/src/sampling/updates.rs:73-158,588-638,1184-1224,
//                             395-441,779-830,1429-1622
//   block loop / recompute    crates/mvmc-core/src/sampling/engine.rs:188-475
//   driver (burn, save)       crates/mvmc-core/src/sampling/driver.rs:175-458
//   local energy              crates/mvmc-core/src/observables.rs:483-552,1638-1768,558-880
//   Slater derivative         crates/mvmc-core/src/slater_derivative.rs:75-99,490-548
//   O store / reductions      crates/mvmc-core/src/observables.rs:202-366, run.rs:1863-1903
// Serial sums stay serial (one lane); parallelism is across independent
// elements, QP planes and walkers.
#include <hip/hip_runtime.h>

#include "common.h"
#include "ensemble.h"
#include "pfinv_dev.h"

namespace mvmc_hip {

// ---------------------------------------------------------------------------
// LDS carving (the host computes the same total through lds_bytes_*)
// ---------------------------------------------------------------------------
struct Carver {
  char* p;
  template <class T>
  __device__ T* take(size_t count) {
    T* r = (T*)p;
    size_t bytes = (count * sizeof(T) + 15) & ~(size_t)15;
    p += bytes;
    return r;
  }
};

__host__ __device__ inline size_t rnd16(size_t b) { return (b + 15) & ~(size_t)15; }
}  // namespace mvmc_hip
#include "pfinv_rr.h"
namespace mvmc_hip {

// Run of this workgroup in a group launch (slot_start is small and uniform).
__device__ inline int group_run(const GroupArgs& g) {
  int r = 0;
  while (r + 1 < g.nruns && (int)blockIdx.x >= g.slot_start[r + 1]) ++r;
  return r;
}

// Run and walker of queue entry g of a group's shared walker queue.
__device__ inline int queue_run(const GroupArgs& g, int e, int& w) {
  int k = 0;
  while (k + 1 < g.nruns && e >= g.qstart[k + 1]) ++k;
  w = e - g.qstart[k];
  return g.qrun[k];
}

// View of run r in a group launch.  The host writes the views before the
// launch and nothing writes them during it, so they are read through the
// constant address space: the compiler then holds the fields in scalar
// registers, reloading them when it needs the room, instead of keeping
// vector copies live across the walker loop.
__device__ inline EnsView group_view(const GroupArgs& g, int r) {
  static_assert(sizeof(EnsView) % 4 == 0, "EnsView is copied as 32-bit words");
  typedef const __attribute__((address_space(4))) uint32_t* ConstWords;
  const ConstWords src = (ConstWords)(g.views + r);
  EnsView v;
  uint32_t* dst = reinterpret_cast<uint32_t*>(&v);
#pragma unroll
  for (int i = 0; i < (int)(sizeof(EnsView) / 4); ++i) dst[i] = src[i];
  return v;
}

// Kernel prologue of the walker kernels: `v` is the ensemble this workgroup
// serves, `slot` its scratch slot and `nslots` the workgroups serving that
// ensemble (the launch's own grid outside a group).
// Adds the workgroup's lifetime (constant-clock ticks) to *out when it ends.
struct SlotTimer {
  long long* out;
  long long t0;
  __device__ explicit SlotTimer(long long* o) : out(o), t0(o ? wall_clock64() : 0) {}
  __device__ ~SlotTimer() {
    if (out) *out += wall_clock64() - t0;
  }
};

#define ENS_SLOT(VARG, GARG)                                                                               \
  SlotTimer slot_timer_(GROUP && threadIdx.x == 0 && (GARG).slot_time ? (GARG).slot_time + blockIdx.x     \
                                                                      : nullptr);                        \
  const int grp_run_ = GROUP ? group_run(GARG) : 0;                                                      \
  const EnsView v = GROUP ? group_view(GARG, grp_run_) : (VARG);                                         \
  const int slot = GROUP ? (int)blockIdx.x - (GARG).slot_start[grp_run_] : (int)blockIdx.x;             \
  const int nslots = GROUP ? (GARG).slot_start[grp_run_ + 1] - (GARG).slot_start[grp_run_] : (int)gridDim.x;

// Row stride of the sampler's hop staging arrays in LDS: odd, so that lanes
// reading the same element of different QP planes hit distinct banks.
__host__ __device__ inline int hop_ld(int n) { return n | 1; }

__host__ __device__ inline size_t union_doubles_sampler(const EnsView& v) {
  const size_t nq = (size_t)v.nqp * v.n;
  size_t hop = 3 * (size_t)v.nqp * hop_ld(v.n);
  size_t ex = v.ex_path != 0 ? 5 * nq + 16 * (size_t)v.nqp : 0;
  size_t m = hop > ex ? hop : ex;
  size_t pfb = (pf_bytes(v.n, v.nqp) + 7) / 8;
  return m > pfb ? m : pfb;
}

__host__ __device__ inline size_t lds_bytes_sampler(const EnsView& v, int nthreads) {
  size_t b = 0;
  b += rnd16(sizeof(SfmtLds));
  b += rnd16((size_t)v.n * 4) + 2 * rnd16((size_t)v.n2 * 4);
  b += rnd16(64 * 4);                      // misc ints
  b += 2 * rnd16((size_t)v.nqp * 8) + rnd16(32 * 8);  // pf, pf_new, dmisc
  b += rnd16((size_t)(v.nproj > 0 ? v.nproj : 1) * 4);  // delta
  b += rnd16((size_t)(4 * v.ns + 16) * 8);  // compacted terms
  b += rnd16((size_t)(nthreads + 64) * 4); // scan
  b += rnd16((size_t)v.n * 4);             // rs
  b += 2 * rnd16((size_t)v.ns * 4) + rnd16((size_t)v.nqp * 8);  // loc_spn, gutz_idx, qpw copies
  b += rnd16(union_doubles_sampler(v) * 8);
  return b;
}

__host__ __device__ inline size_t lds_bytes_measure(const EnsView& v, int nthreads) {
  size_t b = 0;
  b += rnd16((size_t)v.n * 4) + 2 * rnd16((size_t)v.n2 * 4);
  b += rnd16((size_t)(v.nproj > 0 ? v.nproj : 1) * 4);
  b += rnd16((size_t)v.nqp * 8) + rnd16(32 * 8);
  b += 2 * rnd16((size_t)v.n * 4);  // tri, sgi per QP plane handled in loop (n each)
  b += rnd16((size_t)v.n * 4);      // rs
  b += rnd16(pf_bytes(v.n, v.nqp));
  b += rnd16((size_t)v.nqp * v.n * 4) * 2;  // tri/sgi for all planes
  b += rnd16(64 * 4);
  return b;
}

// ---------------------------------------------------------------------------
// small block helpers
// ---------------------------------------------------------------------------
__device__ inline void set_status(int* st, int code) { atomicCAS(st, 0, code); }

// Exclusive block scan of one int per thread; returns the total via *total.
__device__ inline int block_exscan(int v, int* scan, int* total) {
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6, nw = (blockDim.x + 63) >> 6;
  int x = v;
  for (int off = 1; off < 64; off <<= 1) {
    int y = __shfl_up(x, off, 64);
    if (lane >= off) x += y;
  }
  if (lane == 63 || tid == (int)blockDim.x - 1) scan[wave] = x;
  __syncthreads();
  if (tid == 0) {
    int acc = 0;
    for (int k = 0; k < nw; ++k) { int t = scan[k]; scan[k] = acc; acc += t; }
    scan[nw] = acc;
  }
  __syncthreads();
  const int ex = scan[wave] + x - v;
  *total = scan[nw];
  __syncthreads();
  return ex;
}

// calc_m_all for one walker: assemble, LTL, Pfaffian, inverse, negate.
// native = true skips the all-zero guard (C's separate real setup).
// Returns 0 on success, 1 if any plane failed.
// clk: optional phase clocks of the elimination (profiling builds).
template <int JB = PF_JB>
__device__ __forceinline__ int wg_calc_m_all(const EnsView& v, const int* ele_idx_l, int* rs, double* planes, double* work,
                             double* pf_l, char* pfscratch, bool native, long long* clk = nullptr) {
  const int tid = threadIdx.x;
  __shared__ int bad_site;
  __shared__ double wave_max[16];
  if (tid == 0) bad_site = 0;
  __syncthreads();
  for (int i = tid; i < v.n; i += blockDim.x) {
    const int r = ele_idx_l[i];
    rs[i] = r + (i / v.ne) * v.ns;
    if (r < 0 || r >= v.ns) bad_site = 1;
  }
  __syncthreads();
  if (bad_site) {
    if (tid == 0) set_status(v.status, ST_BAD_INPUT);
    __syncthreads();
    return 1;
  }
  assemble_planes(planes, v.slater, rs, v.n, v.n2, v.nqp, 0);
  Carver c{pfscratch};
  PfinvShared sh;
  carve_pfinv(c, sh, v.n, v.nqp);
  __syncthreads();
  __shared__ int allzero_flag;
  if (!native) {
    // max|a|^2 < 1e-28 per plane (pfaffian.rs:884-890)
    if (tid == 0) allzero_flag = 0;
    __syncthreads();
    const int nn = v.n * v.n;
    for (int p = 0; p < v.nqp; ++p) {
      double mx = 0.0;
      for (int e = tid; e < nn; e += blockDim.x) {
        double a = planes[(size_t)p * nn + e];
        mx = fmax(mx, a * a);
      }
      for (int off = 32; off >= 1; off >>= 1) mx = fmax(mx, __shfl_xor(mx, off, 64));
      if ((tid & 63) == 0) wave_max[tid >> 6] = mx;
      __syncthreads();
      if (tid == 0) {
        double m = 0.0;
        for (int k = 0; k < (int)(blockDim.x + 63) / 64; ++k) m = fmax(m, wave_max[k]);
        if (m < 1.0e-28) allzero_flag = 1;
      }
      __syncthreads();
    }
    if (allzero_flag) return 1;
  }
  wg_pfinv<JB>(planes, work, v.n, v.nqp, sh, pf_l, -1.0, clk);
  int bad = 0;
  for (int p = 0; p < v.nqp; ++p) bad |= (sh.info[p] != 0);
  __syncthreads();
  return bad;
}

__device__ inline double ip_of(const EnsView& v, const double* pf) {
  double ip = 0.0;
  for (int q = 0; q < v.nqp; ++q) ip += v.qpw[q] * pf[q];
  return ip;
}

__device__ inline double ip_of_l(const double* qpw, const double* pf, int nqp) {
  double ip = 0.0;
  for (int q = 0; q < nqp; ++q) ip += qpw[q] * pf[q];
  return ip;
}

// ---------------------------------------------------------------------------
// seeding: init_gen_rand(seed + w), then `skip` words (init_parameter draws)
// ---------------------------------------------------------------------------
__global__ void k_seed(EnsView v, long long base_seed, int skip) {
  const int w = blockIdx.x * blockDim.x + threadIdx.x;
  if (w >= v.W) return;
  uint32_t* st = v.rng + (size_t)w * 625;
  uint32_t seed = (uint32_t)(base_seed + w);
  st[0] = seed;
  for (int i = 1; i < 624; ++i) st[i] = 1812433253u * (st[i - 1] ^ (st[i - 1] >> 30)) + (uint32_t)i;
  const uint32_t parity[4] = {0x00000001u, 0x00000000u, 0x00000000u, 0x13c9e684u};
  uint32_t inner = 0;
  for (int i = 0; i < 4; ++i) inner ^= st[i] & parity[i];
  for (int i = 16; i > 0; i >>= 1) inner ^= inner >> i;
  if (!(inner & 1)) {
    bool done = false;
    for (int i = 0; i < 4 && !done; ++i) {
      uint32_t work = 1;
      for (int j = 0; j < 32; ++j) {
        if (work & parity[i]) { st[i] ^= work; done = true; break; }
        work <<= 1;
      }
    }
  }
  int idx = 624;
  for (int k = 0; k < skip; ++k) {
    if (idx >= 624) {
      sfmt_refill_words(st);
      idx = 0;
    }
    ++idx;
  }
  st[624] = (uint32_t)idx;
  int* ws = v.wstate + (size_t)w * 8;
  ws[0] = 0;
  ws[1] = 0;
}

// ---------------------------------------------------------------------------
// sampler state load / store
// ---------------------------------------------------------------------------
struct SampLds {
  SfmtLds* rng;
  int *ele_idx, *ele_cfg, *ele_num, *misc, *delta, *scan, *rs;
  double *pf, *pf_new, *dmisc, *terms, *uni;
  // copies of small model tables read on the serial lane-0 path
  int *loc_spn, *gutz_idx;
  double* qpw;
};

__device__ inline SampLds carve_sampler(const EnsView& v, char* base) {
  Carver c{base};
  SampLds L;
  L.rng = c.take<SfmtLds>(1);
  L.ele_idx = c.take<int>(v.n);
  L.ele_cfg = c.take<int>(v.n2);
  L.ele_num = c.take<int>(v.n2);
  L.misc = c.take<int>(64);
  L.pf = c.take<double>(v.nqp);
  L.pf_new = c.take<double>(v.nqp);
  L.dmisc = c.take<double>(32);
  L.delta = c.take<int>(v.nproj > 0 ? v.nproj : 1);
  L.terms = c.take<double>(4 * v.ns + 16);
  L.scan = c.take<int>(blockDim.x + 64);
  L.rs = c.take<int>(v.n);
  L.loc_spn = c.take<int>(v.ns);
  L.gutz_idx = c.take<int>(v.ns);
  L.qpw = c.take<double>(v.nqp);
  L.uni = c.take<double>(union_doubles_sampler(v));
  return L;
}

// Fill the model-table copies of a freshly carved SampLds (all lanes; the
// caller synchronises before use).
__device__ inline void load_tables(const EnsView& v, SampLds& L) {
  for (int i = threadIdx.x; i < v.ns; i += blockDim.x) {
    L.loc_spn[i] = v.loc_spn[i];
    L.gutz_idx[i] = v.gutz_idx[i];
  }
  for (int q = threadIdx.x; q < v.nqp; q += blockDim.x) L.qpw[q] = v.qpw[q];
}

// Serial sum_k a[k] * b[k] (k ascending) over LDS arrays, with the operands
// of DOT_U terms read into registers ahead of their dependent adds.
constexpr int DOT_U = 16;
__device__ inline double lds_serial_dot(const double* a, const double* b, int n) {
  double r = 0.0;
  for (int k0 = 0; k0 < n; k0 += DOT_U) {
    double av[DOT_U], bv[DOT_U];
#pragma unroll
    for (int u = 0; u < DOT_U; ++u)
      if (k0 + u < n) { av[u] = a[k0 + u]; bv[u] = b[k0 + u]; }
#pragma unroll
    for (int u = 0; u < DOT_U; ++u)
      if (k0 + u < n) r += av[u] * bv[u];
  }
  return r;
}

// Serial r + sum_{k0 <= k < k1} a[k] (k ascending) over an LDS array, same
// prefetching; a sum split at any k is the unsplit sum.
__device__ inline double lds_serial_sum(const double* a, int k0, int k1, double r) {
  for (int kb = k0; kb < k1; kb += DOT_U) {
    double av[DOT_U];
#pragma unroll
    for (int u = 0; u < DOT_U; ++u)
      if (kb + u < k1) av[u] = a[kb + u];
#pragma unroll
    for (int u = 0; u < DOT_U; ++u)
      if (kb + u < k1) r += av[u];
  }
  return r;
}

__device__ inline void load_walker(const EnsView& v, int w, SampLds& L, bool with_rng) {
  const int tid = threadIdx.x;
  if (with_rng) {
    const uint32_t* st = v.rng + (size_t)w * 625;
    for (int i = tid; i < 624; i += blockDim.x) L.rng->w[i] = st[i];
    if (tid == 0) L.rng->idx = (int)st[624];
  }
  for (int i = tid; i < v.n; i += blockDim.x) L.ele_idx[i] = v.ele_idx[(size_t)w * v.n + i];
  for (int i = tid; i < v.n2; i += blockDim.x) {
    L.ele_cfg[i] = v.ele_cfg[(size_t)w * v.n2 + i];
    L.ele_num[i] = v.ele_num[(size_t)w * v.n2 + i];
  }
  for (int i = tid; i < v.nproj; i += blockDim.x) L.delta[i] = 0;
  __syncthreads();
}

__device__ inline void store_walker(const EnsView& v, int w, SampLds& L) {
  const int tid = threadIdx.x;
  uint32_t* st = v.rng + (size_t)w * 625;
  for (int i = tid; i < 624; i += blockDim.x) st[i] = L.rng->w[i];
  if (tid == 0) st[624] = (uint32_t)L.rng->idx;
  for (int i = tid; i < v.n; i += blockDim.x) v.ele_idx[(size_t)w * v.n + i] = L.ele_idx[i];
  for (int i = tid; i < v.n2; i += blockDim.x) {
    v.ele_cfg[(size_t)w * v.n2 + i] = L.ele_cfg[i];
    v.ele_num[(size_t)w * v.n2 + i] = L.ele_num[i];
  }
}

// make_initial_sample (initial.rs:77-141), lane 0.
__device__ void initial_placement(const EnsView& v, SampLds& L) {
  SfmtLds& R = *L.rng;
  const int ns = v.ns, ne = v.ne;
  for (int i = 0; i < v.n; ++i) L.ele_idx[i] = -1;
  for (int i = 0; i < v.n2; ++i) L.ele_cfg[i] = -1;
  for (int ri = 0; ri < ns; ++ri) {
    if (v.loc_spn[ri] != 1) continue;
    for (;;) {
      int mi = (int)sfmt_mod(R, (uint32_t)ne);
      double r = sfmt_real2(R);
      int si = r < 0.5 ? 0 : 1;
      if (L.ele_idx[mi + si * ne] == -1) {
        L.ele_cfg[ri + si * ns] = mi;
        L.ele_idx[mi + si * ne] = ri;
        break;
      }
    }
  }
  for (int si = 0; si < 2; ++si)
    for (int mi = 0; mi < ne; ++mi) {
      if (L.ele_idx[mi + si * ne] != -1) continue;
      for (;;) {
        int ri = (int)sfmt_mod(R, (uint32_t)ns);
        if (L.ele_cfg[ri + si * ns] == -1 && v.loc_spn[ri] != 1) {
          L.ele_cfg[ri + si * ns] = mi;
          L.ele_idx[mi + si * ne] = ri;
          break;
        }
      }
    }
  for (int i = 0; i < v.n2; ++i) L.ele_num[i] = L.ele_cfg[i] < 0 ? 0 : 1;
}

// Projection counts from scratch (projection.rs:179-277), written to HBM.
__device__ void make_proj_cnt(const EnsView& v, const int* ele_num, int* cnt) {
  const int ns = v.ns;
  for (int i = threadIdx.x; i < v.nproj; i += blockDim.x) cnt[i] = 0;
  __syncthreads();
  if (threadIdx.x == 0) {
    if (v.ngutz > 0)
      for (int ri = 0; ri < ns; ++ri) {
        int idx = v.gutz_idx[ri];
        if (idx >= 0 && idx < v.nproj) cnt[idx] += ele_num[ri] * ele_num[ri + ns];
      }
  }
  __syncthreads();
  if (v.nproj > v.ngutz) {
    for (int ri = threadIdx.x; ri < ns; ri += blockDim.x) {
      int xi = ele_num[ri] + ele_num[ri + ns] - 1;
      if (xi == 0) continue;
      for (int rj = ri + 1; rj < ns; ++rj) {
        int idx = v.jast_idx[(size_t)ri * ns + rj];
        if (idx < 0) continue;
        int xj = ele_num[rj] + ele_num[rj + ns] - 1;
        atomicAdd(&cnt[v.ngutz + idx], xi * xj);
      }
    }
  }
  __syncthreads();
}

// ---------------------------------------------------------------------------
// begin: burn restore or validated initial placement, initial tables, log|ip|
// (driver.rs:239-330)
// ---------------------------------------------------------------------------
// The mixed-precision form keeps the sampler's inverse in FP32: after a
// from-scratch FP64 calc_m_all, walker w's planes are copied rounded to
// nearest (n is even, so pairs never straddle the end).
__device__ inline void copy_inv32(const EnsView& v, int w) {
  if (!v.inv32) return;
  const size_t m2 = (size_t)v.nqp * v.n * v.n / 2;
  const double2* src = reinterpret_cast<const double2*>(v.inv + (size_t)w * 2 * m2);
  float2* dst = reinterpret_cast<float2*>(v.inv32 + (size_t)w * 2 * m2);
  for (size_t e = threadIdx.x; e < m2; e += blockDim.x) {
    const double2 x = src[e];
    dst[e] = make_float2((float)x.x, (float)x.y);
  }
}

// only_marked: serve only the walkers k_begin_rr left to it (wstate[7] set).
template <bool GROUP>
__global__ void __launch_bounds__(256) k_begin(EnsView v_arg, GroupArgs ga, int only_marked) {
  ENS_SLOT(v_arg, ga)
  extern __shared__ __align__(16) char lds[];
  SampLds L = carve_sampler(v, lds);
  const int tid = threadIdx.x;
  double* work = v.work + (size_t)slot * pfinv_work_doubles(v.n, v.nqp);
  char* pfscr = (char*)L.uni;
  for (int w = slot; w < v.W; w += nslots) {
    int* ws = v.wstate + (size_t)w * 8;
    if (only_marked && ws[7] == 0) continue;
    double* planes = v.inv + (size_t)w * v.nqp * v.n * v.n;
    int* cnt = v.proj_cnt + (size_t)w * v.nproj;
    load_walker(v, w, L, true);
    int burn = ws[0];
    // phase 0: fresh placement (up to 100 retries while calc_m_all fails),
    // 1: recompute from the burn configuration, 2: one fresh placement
    // after a non-finite log|<x|psi>|.  A single calc_m_all call site.
    int phase = burn ? 1 : 0, attempt = 0;
    __shared__ double lip;
    for (;;) {
      const bool placing = phase != 1;
      if (placing) {
        ++attempt;
        if (tid == 0) initial_placement(v, L);
        __syncthreads();
      }
      const int bad = wg_calc_m_all(v, L.ele_idx, L.rs, planes, work, L.pf, pfscr, true);
      if (placing) {
        if (attempt <= 100 && bad) continue;
        if (attempt > 100) set_status(v.status, ST_INIT_RETRY);
        make_proj_cnt(v, L.ele_num, cnt);
      }
      if (tid == 0) lip = ddm::cr_log(fabs(ip_of(v, L.pf)));
      __syncthreads();
      if (phase == 2 || lip - lip == 0.0) break;
      phase = 2;
      attempt = 0;
      burn = 0;
    }
    copy_inv32(v, w);
    for (int q = tid; q < v.nqp; q += blockDim.x) v.pf[(size_t)w * v.nqp + q] = L.pf[q];
    if (tid == 0) {
      v.log_ip[w] = lip;
      ws[1] = 0;
      ws[2] = burn ? v.nsample + 1 : v.warmup + v.nsample;
      ws[7] = 0;
      for (int k = 0; k < 4; ++k) v.counter[(size_t)w * 4 + k] = 0;
    }
    store_walker(v, w, L);
    __syncthreads();
  }
}

// ---------------------------------------------------------------------------
// Projection-count deltas of a hop ri -> rj (post-hop occupations) into
// LDS `delta` (integer adds commute, so lanes may add concurrently).
// ---------------------------------------------------------------------------
__device__ inline int jast_of(const EnsView& v, int a, int b) {
  if (a == b) return -1;
  // the table is stored in full (the loader checks its symmetry), so row a
  // is contiguous and lanes over b read it coalesced
  return v.jast_idx[(size_t)a * v.ns + b];
}

__device__ void hop_deltas(const EnsView& v, SampLds& L, int ri, int rj) {
  const int ns = v.ns;
  const int* n0 = L.ele_num;
  const int* n1 = L.ele_num + ns;
  if (ri == rj) return;
  if (threadIdx.x == 0 && v.ngutz > 0) {
    int a = L.gutz_idx[ri], b = L.gutz_idx[rj];
    if (a >= 0) atomicAdd(&L.delta[a], -(n0[ri] + n1[ri]));
    if (b >= 0) atomicAdd(&L.delta[b], n0[rj] * n1[rj]);
  }
  if (v.nproj > v.ngutz) {
    const int off = v.ngutz;
    if (threadIdx.x == 0) {
      int idx = jast_of(v, ri, rj);
      if (idx >= 0) atomicAdd(&L.delta[off + idx], (n0[ri] + n1[ri]) - (n0[rj] + n1[rj]) + 1);
    }
    for (int rk = threadIdx.x; rk < ns; rk += blockDim.x) {
      if (rk == ri || rk == rj) continue;
      int x = n0[rk] + n1[rk] - 1;
      if (x == 0) continue;
      int a = jast_of(v, ri, rk);
      if (a >= 0) atomicAdd(&L.delta[off + a], -x);
      int b = jast_of(v, rj, rk);
      if (b >= 0) atomicAdd(&L.delta[off + b], x);
    }
  }
}

// The terms of log_proj_ratio, proj[p] * delta[p] over the changed entries
// in ascending parameter order (zero terms leave a serial sum unchanged),
// compacted into L.terms; returns their count (all lanes).  The parameter
// loads of CB entries are issued together.
__device__ int proj_terms(const EnsView& v, SampLds& L) {
  const int tid = threadIdx.x, nt = blockDim.x;
  const int np = v.nproj;
  const int chunk = (np + nt - 1) / nt;
  const int p0 = tid * chunk, p1 = min(np, p0 + chunk);
  int cntl = 0;
  for (int p = p0; p < p1; ++p) cntl += (L.delta[p] != 0);
  int total;
  int off = block_exscan(cntl, L.scan, &total);
  const int cap = 4 * v.ns + 16;
  if (total > cap) {
    if (tid == 0) set_status(v.status, ST_BAD_INPUT);
    total = cap;
  }
  constexpr int CB = 8;
  for (int b = p0; b < p1; b += CB) {
    int d[CB];
    double pv[CB];
#pragma unroll
    for (int u = 0; u < CB; ++u) d[u] = b + u < p1 ? L.delta[b + u] : 0;
#pragma unroll
    for (int u = 0; u < CB; ++u)
      if (d[u] != 0) pv[u] = v.proj[b + u];
#pragma unroll
    for (int u = 0; u < CB; ++u)
      if (d[u] != 0) {
        if (off < cap) L.terms[off] = pv[u] * (double)d[u];
        ++off;
      }
  }
  __syncthreads();
  return total;
}

__device__ inline void apply_clear_deltas(const EnsView& v, SampLds& L, int* cnt, bool apply) {
  for (int p = threadIdx.x; p < v.nproj; p += blockDim.x) {
    int d = L.delta[p];
    if (d != 0) {
      if (apply) cnt[p] += d;
      L.delta[p] = 0;
    }
  }
}

// ---------------------------------------------------------------------------
// block of proposals (engine.rs:188-475) and sample saving (driver.rs:419-443)
// ---------------------------------------------------------------------------
constexpr int CG_RCH = 32, CG_SCH = 16, CG_CMAX = 32;  // SR-CG reduction blocking

// One not yet applied hop update (updates.rs:1184-1224: vec1, vec2 = scaled
// old row `pm`) on element (i, j), with the operation order of the in-place
// update so the delayed element equals the eagerly updated one bit for bit.
template <class CT>
__device__ inline CT apply_pend(CT x, CT v1i, CT v2i, CT v1j, CT v2j, int i, int j, int pm) {
  x += v1i * v2j - v1j * v2i;
  if (j == pm) x -= v2i;
  if (i == pm) x += v2j;
  return x;
}

// Raw buffer access to one walker's planes: a 32-bit per-lane byte offset
// plus a wave-uniform row offset in a scalar register, so a tile of loads
// costs one address register instead of one 64-bit pointer per load.
typedef unsigned int u32x4_t __attribute__((ext_vector_type(4)));
typedef unsigned int u32x2_t __attribute__((ext_vector_type(2)));
constexpr int BUF_WORD3 = 0x00020000;  // raw untyped buffer (gfx9 family)
// Cache policy of the stream (0: default; the non-temporal hint measured
// slower: the walker's inverse is re-read at its next accepted hop).
constexpr int BUF_NT = 0;

// Storage of the planes seen as 16-byte vectors of EPV elements: FP64
// (EPV = 2) or FP32, the mixed-precision form (EPV = 4; the rank-2 updates
// of the stream in FP32, everything else in FP64).  cnt = elements of the
// vector inside the row (EPV, or 2 for an FP32 vector at the end of a row
// whose length is 2 mod 4).
template <class T>
struct PlaneIO;
template <>
struct PlaneIO<double> {
  static constexpr int EPV = 2;
  static __device__ inline u32x4_t load(__amdgpu_buffer_rsrc_t r, int voff, int soff, int) {
    return __builtin_amdgcn_raw_buffer_load_b128(r, voff, soff, BUF_NT);
  }
  static __device__ inline void store(u32x4_t x, __amdgpu_buffer_rsrc_t r, int voff, int soff, int) {
    __builtin_amdgcn_raw_buffer_store_b128(x, r, voff, soff, BUF_NT);
  }
  static __device__ inline double get(u32x4_t x, int h) {
    const double2 d = __builtin_bit_cast(double2, x);
    return h == 0 ? d.x : d.y;
  }
  static __device__ inline void set(u32x4_t& x, int h, double v) {
    double2 d = __builtin_bit_cast(double2, x);
    if (h == 0) d.x = v; else d.y = v;
    x = __builtin_bit_cast(u32x4_t, d);
  }
};
template <>
struct PlaneIO<float> {
  static constexpr int EPV = 4;
  static __device__ inline u32x4_t load(__amdgpu_buffer_rsrc_t r, int voff, int soff, int cnt) {
    if (cnt == 4) return __builtin_amdgcn_raw_buffer_load_b128(r, voff, soff, BUF_NT);
    const u32x2_t h = __builtin_amdgcn_raw_buffer_load_b64(r, voff, soff, BUF_NT);
    u32x4_t x = {h.x, h.y, 0u, 0u};
    return x;
  }
  static __device__ inline void store(u32x4_t x, __amdgpu_buffer_rsrc_t r, int voff, int soff, int cnt) {
    if (cnt == 4) __builtin_amdgcn_raw_buffer_store_b128(x, r, voff, soff, BUF_NT);
    else __builtin_amdgcn_raw_buffer_store_b64(u32x2_t{x.x, x.y}, r, voff, soff, BUF_NT);
  }
  static __device__ inline float get(u32x4_t x, int h) {
    const unsigned u = h == 0 ? x.x : h == 1 ? x.y : h == 2 ? x.z : x.w;
    return __builtin_bit_cast(float, u);
  }
  static __device__ inline void set(u32x4_t& x, int h, float v) {
    const unsigned u = __builtin_bit_cast(unsigned, v);
    if (h == 0) x.x = u; else if (h == 1) x.y = u; else if (h == 2) x.z = u; else x.w = u;
  }
};

// Accepted hop updates not yet written back to the stored planes: up to
// PEND_MAX of them, in acceptance order.  Update k is the pair of vectors
// (vec1, vec2) of updates.rs:1184-1224 and the row pm[k] it replaces; the
// pairs live in the slot's scratch as pend[(q * n + i) * PEND_MAX + k] =
// {vec1[q][i], vec2[q][i]}, so one 16-byte load gives a row's two values.
// An element of the current inverse is its stored value with the pending
// updates applied in order (apply_pend), which is bit for bit the eagerly
// updated element; so the pass of an accepted hop, which has to read every
// element anyway, writes the planes back only when the list is full.
constexpr int PEND_MAX = 2;
// Storage of a pending pair: FP64; FP32 in the mixed-precision form, whose
// stream applies the updates in FP32 anyway (the row staging of a proposal
// then applies the FP32-rounded vectors in FP64).
template <class T>
struct PendPair;
template <>
struct PendPair<double> {
  using type = double2;
  static __device__ inline type make(double a, double b) { return make_double2(a, b); }
};
template <>
struct PendPair<float> {
  using type = float2;
  static __device__ inline type make(double a, double b) { return make_float2((float)a, (float)b); }
};
struct PendList {
  int np;             // pending updates (uniform over the workgroup)
  int pm[PEND_MAX];   // their rows
};

// One streaming pass over a walker's planes.  A lane owns CPL = 2 * EPV
// contiguous columns of one plane (q = tid / LP; columns
// LP*CPL*g + CPL*(tid % LP) + c), read and written as two 16-byte vectors
// per row, and issues the loads of STREAM_ROWS rows before using any of
// them: each lane keeps STREAM_ROWS rows of its columns in flight, which is
// what bounds a pass over HBM at this occupancy (the FP32 form therefore
// covers twice the columns per round trip).  The pending values of a tile's
// rows (cache hits) are loaded before the next tile's rows, so waiting for
// them does not wait for that tile.
//   P.np > 0 : apply the pending updates, in order, in the storage type's
//              arithmetic; store = write the result back;
//   ACC      : out[q][j] = sum_i -X[i][j] * b[q][i] in FP64, serial in i
//              (the update_one_real vec1 = -(X^T b)), X the current inverse;
//              rows of b and out are ldb doubles apart.
constexpr int STREAM_ROWS = 4;
template <bool ACC, class T>
__device__ void stream_planes(T* planes, int n, int nqp, const typename PendPair<T>::type* pend, const PendList& P,
                              bool store,
                              const double* b, double* out, int ldb = 0) {
  using IO = PlaneIO<T>;
  using PV = typename PendPair<T>::type;
  constexpr int EPV = IO::EPV, NV = 2, CPL = NV * EPV;
  constexpr int ES = (int)sizeof(T);
  const int nt = blockDim.x, tid = threadIdx.x;
  const int nn = n * n;
  const int PG = nqp < nt ? nqp : nt;
  const int LP = nt / PG;
  const int qo = tid / LP, jl = tid % LP;
  const int span = LP * CPL;  // columns per group
  const __amdgpu_buffer_rsrc_t rs =
      __builtin_amdgcn_make_buffer_rsrc(planes, 0, (int)((size_t)nqp * nn * ES), BUF_WORD3);
  const int rowb = n * ES;
  const int np = P.np;
  for (int q0 = 0; q0 < nqp; q0 += PG) {
    const int q = q0 + qo;
    const bool on = qo < PG && q < nqp;
    const int qoff = (on ? q : 0) * nn * ES;
    const PV* pq = pend + (size_t)(on ? q : 0) * n * PEND_MAX;
    const double* bq = b + (on ? q : 0) * ldb;
    for (int g0 = 0; g0 < n; g0 += span) {
      const int jb = g0 + CPL * jl;  // first column of the lane
      int cnt[NV];                   // elements of vector m inside the row
      double acc[CPL];
      T c1[PEND_MAX][CPL], c2[PEND_MAX][CPL];
#pragma unroll
      for (int m = 0; m < NV; ++m) cnt[m] = on ? max(0, min(EPV, n - (jb + EPV * m))) : 0;
#pragma unroll
      for (int c = 0; c < CPL; ++c) {
        const bool in = c % EPV < cnt[c / EPV];
        acc[c] = 0.0;
#pragma unroll
        for (int k = 0; k < PEND_MAX; ++k) {
          PV d = {0, 0};
          if (k < np && in) d = pq[(jb + c) * PEND_MAX + k];
          c1[k][c] = (T)d.x;
          c2[k][c] = (T)d.y;
        }
      }
      const int vbase = qoff + jb * ES;
      // Two tiles of plane rows in flight: the loads of tile i0 +
      // STREAM_ROWS are issued before tile i0 is updated and stored (a load
      // after a buffer store could not be moved above it by the compiler).
      u32x4_t x[STREAM_ROWS][NV], xn[STREAM_ROWS][NV];
      auto load_tile = [&](int i0, u32x4_t (&tx)[STREAM_ROWS][NV]) {
#pragma unroll
        for (int r = 0; r < STREAM_ROWS; ++r)
          if (i0 + r < n) {
#pragma unroll
            for (int m = 0; m < NV; ++m)
              if (cnt[m] > 0) tx[r][m] = IO::load(rs, vbase + 16 * m, (i0 + r) * rowb, cnt[m]);
          }
      };
      load_tile(0, x);
      for (int i0 = 0; i0 < n; i0 += STREAM_ROWS) {
        PV pr[STREAM_ROWS][PEND_MAX];
        double rsv[STREAM_ROWS];
#pragma unroll
        for (int r = 0; r < STREAM_ROWS; ++r)
          if (i0 + r < n) {
#pragma unroll
            for (int k = 0; k < PEND_MAX; ++k)
              if (k < np) pr[r][k] = pq[(i0 + r) * PEND_MAX + k];
            if (ACC) rsv[r] = bq[i0 + r];
          }
        if (i0 + STREAM_ROWS < n) load_tile(i0 + STREAM_ROWS, xn);
#pragma unroll
        for (int r = 0; r < STREAM_ROWS; ++r) {
          const int i = i0 + r;
          if (i < n) {
            if (np > 0) {
#pragma unroll
              for (int m = 0; m < NV; ++m)
                if (cnt[m] > 0) {
#pragma unroll
                  for (int h = 0; h < EPV; ++h)
                    if (h < cnt[m]) {
                      T e = IO::get(x[r][m], h);
#pragma unroll
                      for (int k = 0; k < PEND_MAX; ++k)
                        if (k < np)
                          e = apply_pend(e, (T)pr[r][k].x, (T)pr[r][k].y, c1[k][EPV * m + h], c2[k][EPV * m + h], i,
                                         jb + EPV * m + h, P.pm[k]);
                      IO::set(x[r][m], h, e);
                    }
                  if (store) IO::store(x[r][m], rs, vbase + 16 * m, i * rowb, cnt[m]);
                }
            }
            if (ACC) {
              const double s0 = rsv[r];
#pragma unroll
              for (int m = 0; m < NV; ++m)
#pragma unroll
                for (int h = 0; h < EPV; ++h)
                  if (h < cnt[m]) acc[EPV * m + h] += -(double)IO::get(x[r][m], h) * s0;
            }
          }
        }
#pragma unroll
        for (int r = 0; r < STREAM_ROWS; ++r)
#pragma unroll
          for (int m = 0; m < NV; ++m) x[r][m] = xn[r][m];
      }
      if (ACC) {
#pragma unroll
        for (int c = 0; c < CPL; ++c)
          if (c % EPV < cnt[c / EPV]) out[q * ldb + jb + c] = acc[c];
      }
    }
  }
}

// Apply every pending update to the stored planes and empty the list;
// passes (if not null) = {write passes, flush passes} counters.
template <class T>
__device__ inline void flush_pending(T* planes, int n, int nqp, const typename PendPair<T>::type* pend, PendList& P,
                                     int* passes) {
  if (P.np > 0) {
    stream_planes<false>(planes, n, nqp, pend, P, true, nullptr, nullptr);
    if (passes && threadIdx.x == 0) {
      atomicAdd(passes, 1);
      atomicAdd(passes + 1, 1);
    }
    P.np = 0;
  }
}

// Optional phase clocks (MVMC_PHASE_CLOCKS builds): lane 0 charges the time
// since the previous mark to phase k of its slot, after the barrier that
// ends the phase.
#ifdef MVMC_PHASE_CLOCKS
#define PHASE_INIT                   \
  long long clk_t0 = wall_clock64(); \
  long long* clk = v.clk + (size_t)slot * 32;
#define PHASE(k)                           \
  do {                                     \
    if (threadIdx.x == 0) {                \
      const long long t_ = wall_clock64(); \
      clk[k] += t_ - clk_t0;               \
      clk_t0 = t_;                         \
    }                                      \
  } while (0)
#else
#define PHASE_INIT
#define PHASE(k) \
  do {           \
  } while (0)
#endif

enum { M_TYPE = 0, M_REJ, M_MI, M_S, M_RI, M_RJ, M_MJ, M_T, M_ACC, M_NACC, M_NOUT, M_SKIP, M_RES, M_DONE };

// Elem: storage type of the sampler's inverse (double; float in the
// mixed-precision form, whose arithmetic stays FP64).
// In a group launch (GROUP) the workgroups take entries of the group's
// shared walker queue, so every run's walkers are spread over the whole grid
// as they come; the view, LDS carve and tables follow the entry's run.
template <class Elem, bool GROUP>
__global__ void __launch_bounds__(256, 2) k_block(EnsView v_arg, GroupArgs ga, int out_step, int n_in_arg) {
  SlotTimer slot_timer_(GROUP && threadIdx.x == 0 && ga.slot_time ? ga.slot_time + blockIdx.x : nullptr);
  const int slot = blockIdx.x;
  int cur_run = -1;
  EnsView v = v_arg;
  if (GROUP) {
    int w0 = 0;
    cur_run = queue_run(ga, 0, w0);
    v = group_view(ga, cur_run);
  }
  // proposals per outer step (NVMCInterval x Nsite): each run's own in a group
  int n_in = GROUP ? v.interval * (v.ns > 0 ? v.ns : 1) : n_in_arg;
  Elem* inv = sizeof(Elem) == sizeof(float) ? (Elem*)v.inv32 : (Elem*)v.inv;
  extern __shared__ __align__(16) char lds[];
  SampLds L = carve_sampler(v, lds);
  load_tables(v, L);
  const int tid = threadIdx.x, nt = blockDim.x;
  int n = v.n, ne = v.ne, ns = v.ns, n2 = v.n2, nqp = v.nqp, nn = n * n;
  double* work = GROUP ? ga.work + (size_t)slot * ga.work_stride : v.work + (size_t)slot * pfinv_work_doubles(n, nqp);
  // pending hop updates of the slot's walker
  using PV = typename PendPair<Elem>::type;
  const PV* pend_c = (const PV*)(GROUP ? ga.pend + (size_t)slot * ga.pend_stride
                                       : v.pend + (size_t)slot * 2 * PEND_MAX * nqp * n);
  PV* pend = const_cast<PV*>(pend_c);
  // write and flush passes of a single-ensemble launch (nflag[3], nflag[4]);
  // not counted in a group
  int* const wpasses = GROUP ? nullptr : v_arg.nflag + 3;
  PHASE_INIT
  __shared__ int next_w;
  for (;;) {
    // walkers are taken dynamically, so the launch ends when the work does
    // (a walker's results do not depend on the slot that runs it: the
    // slot's scratch is rewritten before it is read)
    if (tid == 0) next_w = atomicAdd(GROUP ? ga.gq + 1 : v.nflag + 1, 1);
    __syncthreads();
    const int qe = next_w;
    __syncthreads();
    int w = qe;
    if (GROUP) {
      if (qe >= ga.qstart[ga.nruns]) break;
      const int r = queue_run(ga, qe, w);
      if (r != cur_run) {
        cur_run = r;
        v = group_view(ga, r);
        n_in = v.interval * (v.ns > 0 ? v.ns : 1);
        inv = sizeof(Elem) == sizeof(float) ? (Elem*)v.inv32 : (Elem*)v.inv;
        n = v.n; ne = v.ne; ns = v.ns; n2 = v.n2; nqp = v.nqp; nn = n * n;
        L = carve_sampler(v, lds);
        load_tables(v, L);
        __syncthreads();
      }
    } else if (w >= v.W) {
      break;
    }
    int* ws = v.wstate + (size_t)w * 8;
    PendList P;  // hop updates not yet written back (uniform over the workgroup)
    P.np = 0;
    if (tid == 0) {
      L.misc[M_NOUT] = ws[2];
      L.misc[M_NACC] = ws[1];
      L.misc[M_RES] = ws[4];
      L.misc[M_DONE] = ws[6] | ws[5];
      L.dmisc[1] = v.log_ip[w];
    }
    __syncthreads();
    const int n_out = L.misc[M_NOUT];
    const int resume = L.misc[M_RES];
    if (out_step >= n_out || L.misc[M_DONE]) { __syncthreads(); continue; }
    load_walker(v, w, L, true);
    Elem* planes = inv + (size_t)w * nqp * nn;
    int* cnt = v.proj_cnt + (size_t)w * v.nproj;
    long long* ctr = v.counter + (size_t)w * 4;
    for (int q = tid; q < nqp; q += nt) L.pf[q] = v.pf[(size_t)w * nqp + q];
    __syncthreads();
    int stop_at = -1;  // proposal index to resume at after a rebuild
    for (int step = resume; step < n_in; ++step) {
      // ---- candidate (lane 0) ----
      if (tid == 0) {
        SfmtLds& R = *L.rng;
        int type;
        switch (v.ex_path) {
          case 0: type = 0; break;
          case 1: type = sfmt_real2(R) < 0.5 ? 1 : 0; break;
          case 2: type = 1; break;
          default: type = 3;
        }
        int rej = 0, mi = 0, s = 0, ri = 0, rj = 0, mj = 0, t = 0;
        if (type == 0) {
          ctr[0]++;
          const long long icnt_max = (long long)ns * ns;
          long long icnt_e = 0;
          for (;;) {
            if (icnt_e == icnt_max + 1) {
              bool any = false;
              for (int k = 0; k < 2 * ne; ++k) {
                int r = L.ele_idx[k] < 0 ? 0 : L.ele_idx[k];
                any = any || !(r >= ns || L.loc_spn[r] == 1);
              }
              if (!any) { rej = 1; break; }
            }
            ++icnt_e;
            mi = (int)sfmt_mod(R, (uint32_t)ne);
            s = sfmt_real2(R) < 0.5 ? 0 : 1;
            ri = L.ele_idx[mi + s * ne];
            if (!(ri < 0 || ri >= ns || L.loc_spn[ri] == 1)) break;
          }
          if (!rej) {
            long long icnt = 0;
            for (;;) {
              rj = (int)sfmt_mod(R, (uint32_t)ns);
              if (icnt > icnt_max) { rej = 1; break; }
              ++icnt;
              if (L.ele_cfg[rj + s * ns] == -1 && L.loc_spn[rj] != 1) break;
            }
          }
          if (!rej) {
            // update_ele_config
            L.ele_idx[mi + s * ne] = rj;
            L.ele_cfg[ri + s * ns] = -1;
            L.ele_cfg[rj + s * ns] = mi;
            L.ele_num[ri + s * ns] = 0;
            L.ele_num[rj + s * ns] = 1;
          }
        } else if (type == 1) {
          ctr[2]++;
          bool single = false;
          for (int r = 0; r < ns; ++r)
            if (L.ele_num[r] + L.ele_num[r + ns] == 1) { single = true; break; }
          if (!single) rej = 1;
          else {
            for (;;) {
              mi = (int)sfmt_mod(R, (uint32_t)ne);
              s = sfmt_real2(R) < 0.5 ? 0 : 1;
              ri = L.ele_idx[mi + s * ne];
              if (L.ele_cfg[ri + (1 - s) * ns] == -1) break;
            }
            t = 1 - s;
            for (;;) {
              mj = (int)sfmt_mod(R, (uint32_t)ne);
              rj = L.ele_idx[mj + t * ne];
              if (L.ele_cfg[rj + s * ns] == -1) break;
            }
            // first hop (mi: ri -> rj, spin s)
            L.ele_idx[mi + s * ne] = rj;
            L.ele_cfg[ri + s * ns] = -1;
            L.ele_cfg[rj + s * ns] = mi;
            L.ele_num[ri + s * ns] = 0;
            L.ele_num[rj + s * ns] = 1;
          }
        } else {
          rej = 2;  // no update: still runs the recompute guard
        }
        L.misc[M_TYPE] = type;
        L.misc[M_REJ] = rej;
        L.misc[M_MI] = mi; L.misc[M_S] = s; L.misc[M_RI] = ri; L.misc[M_RJ] = rj;
        L.misc[M_MJ] = mj; L.misc[M_T] = t;
      }
      __syncthreads();
      PHASE(0);
      const int rej = L.misc[M_REJ], type = L.misc[M_TYPE];
      if (rej == 1) {
        // `continue` in the reference skips the recompute guard too; the
        // barrier keeps lane 0 from overwriting the candidate fields before
        // every wave has read them.
        __syncthreads();
        continue;
      }
      if (rej == 0) {
        const int mi = L.misc[M_MI], s = L.misc[M_S], ri = L.misc[M_RI], rj = L.misc[M_RJ];
        const int mj = L.misc[M_MJ], t = L.misc[M_T];
        hop_deltas(v, L, ri, rj);
        __syncthreads();
        if (type == 1) {
          if (tid == 0) {  // second hop (mj: rj -> ri, spin t)
            L.ele_idx[mj + t * ne] = ri;
            L.ele_cfg[rj + t * ns] = -1;
            L.ele_cfg[ri + t * ns] = mj;
            L.ele_num[rj + t * ns] = 0;
            L.ele_num[ri + t * ns] = 1;
          }
          __syncthreads();
          hop_deltas(v, L, rj, ri);
          __syncthreads();
        }
        const int nterms = proj_terms(v, L);
        PHASE(1);
        // ---- Pfaffian ratios ----
        // The projection sum is one serial chain: on the hop path lane 0 of
        // wave 0 forms it (in two parts, carried in a register) while the
        // other waves stage the rows and form the Pfaffian-ratio dots.
        // Rows of the hop staging arrays are hop_ld(n) (odd) doubles apart,
        // so the ratio dots' lanes (one per QP plane, same k) read distinct
        // LDS banks.
        const int ld = hop_ld(n);
        double* stage = L.uni;              // [nqp][ld]
        double* slt = L.uni + nqp * ld;     // [nqp][ld]
        double* vec1 = L.uni + 2 * nqp * ld;
        if (type == 0) {
          const int msa = mi + s * ne;
          const int rsa = L.ele_idx[msa] + s * ns;
          const int ksplit = nterms / 2;
          double lsum = 0.0;
          if (tid == 0) lsum = lds_serial_sum(L.terms, 0, ksplit, 0.0);
          // row msa of the current inverse (stored row + pending updates)
          // and the Slater row (waves 1..), loads of STAGE_U elements per
          // lane issued together
          constexpr int STAGE_U = 4;
          const int st_id = tid - 64, st_n = nt - 64;
          const int np = P.np;
          for (int e0 = 0; e0 < nqp * n && tid >= 64; e0 += st_n * STAGE_U) {
            double xv[STAGE_U], sv[STAGE_U];
            PV pj[STAGE_U][PEND_MAX];
#pragma unroll
            for (int u = 0; u < STAGE_U; ++u) {
              const int e = e0 + st_id + u * st_n;
              if (e < nqp * n) {
                const int q = e / n, j = e % n;
                xv[u] = planes[(size_t)q * nn + (size_t)msa * n + j];
                const int rsj = L.ele_idx[j] + (j < ne ? 0 : ns);
                sv[u] = v.slater[((size_t)q * n2 + rsa) * n2 + rsj];
#pragma unroll
                for (int k = 0; k < PEND_MAX; ++k)
                  if (k < np) pj[u][k] = pend_c[(size_t)e * PEND_MAX + k];
              }
            }
#pragma unroll
            for (int u = 0; u < STAGE_U; ++u) {
              const int e = e0 + st_id + u * st_n;
              if (e < nqp * n) {
                const int q = e / n, j = e % n;
                double x = xv[u];
#pragma unroll
                for (int k = 0; k < PEND_MAX; ++k)
                  if (k < np) {
                    const PV pi = pend_c[((size_t)q * n + msa) * PEND_MAX + k];
                    x = apply_pend<double>(x, pi.x, pi.y, pj[u][k].x, pj[u][k].y, msa, j, P.pm[k]);
                  }
                stage[q * ld + j] = x;
                slt[q * ld + j] = sv[u];
              }
            }
          }
          __syncthreads();
          PHASE(2);
          if (tid == 0) L.dmisc[0] = lds_serial_sum(L.terms, ksplit, nterms, lsum);
          for (int q = tid - 64; q < nqp && tid >= 64; q += nt - 64) {
            const double ratio = lds_serial_dot(stage + q * ld, slt + q * ld, n);
            L.pf_new[q] = -ratio * L.pf[q];
          }
        } else {
          if (tid == 0) L.dmisc[0] = lds_serial_sum(L.terms, 0, nterms, 0.0);
          // two-electron ratio (updates.rs:395-441, 1539-1622) on the
          // fully updated inverse
          if (P.np > 0) {
            flush_pending(planes, n, nqp, pend_c, P, wpasses);
            __syncthreads();
          }
          const int msa = mi + s * ne, msb = mj + t * ne;
          double* va = L.uni;                 // [nqp][n]
          double* vb = L.uni + nqp * n;
          double* ia = L.uni + 2 * nqp * n;   // inv row msa
          double* ib = L.uni + 3 * nqp * n;   // inv row msb
          double* tmp = L.uni + 4 * nqp * n;  // row dots of the bilinear form
          const int rsa = L.ele_idx[msa] + s * ns, rsb = L.ele_idx[msb] + t * ns;
          for (int e = tid; e < nqp * n; e += nt) {
            const int q = e / n, i = e % n;
            const int rsi = L.ele_idx[i] + (i < ne ? 0 : ns);
            const double* S = v.slater + (size_t)q * n2 * n2;
            va[e] = S[(size_t)rsa * n2 + rsi];
            vb[e] = S[(size_t)rsb * n2 + rsi];
            ia[e] = planes[(size_t)q * nn + (size_t)msa * n + i];
            ib[e] = planes[(size_t)q * nn + (size_t)msb * n + i];
          }
          __syncthreads();
          for (int e = tid; e < nqp * n; e += nt) {
            const int q = e / n, i = e % n;
            const Elem* row = planes + (size_t)q * nn + (size_t)i * n;
            const double* a = va + q * n;
            double acc = 0.0;
            for (int j = 0; j < n; ++j) acc += row[j] * a[j];
            tmp[e] = acc;
          }
          __syncthreads();
          for (int q = tid; q < nqp; q += nt) {
            const double* a = va + q * n;
            const double* b = vb + q * n;
            const double* ra = ia + q * n;
            const double* rb = ib + q * n;
            double pa = 0.0, pb = 0.0, qa = 0.0, qb = 0.0;
            for (int i = 0; i < n; ++i) {
              pa += ra[i] * a[i];
              pb += rb[i] * a[i];
              qa += ra[i] * b[i];
              qb += rb[i] * b[i];
            }
            const double iab = ra[msb];
            double bma = 0.0;
            for (int i = 0; i < n; ++i) bma += b[i] * tmp[q * n + i];
            const double ratio = iab * b[msa] + iab * bma + pa * qb - pb * qa;
            L.pf_new[q] = ratio * L.pf[q];
          }
        }
        __syncthreads();
        PHASE(type == 0 ? 3 : 7);
        // ---- Metropolis (lane 0) ----
        if (tid == 0) {
          const double lpd = L.dmisc[0];
          const double lip_new = ddm::cr_log(fabs(ip_of_l(L.qpw, L.pf_new, nqp)));
          const double ex = 2.0 * (((lpd + 0.0) + lip_new) - L.dmisc[1]);
          double wgt = ddm::cr_exp(ex);
          if (!(wgt - wgt == 0.0)) wgt = -1.0;
          const double u = sfmt_real2(*L.rng);
          const int acc = wgt > u;
          L.misc[M_ACC] = acc;
          if (acc) {
            L.dmisc[1] = lip_new;
            L.misc[M_NACC] += 1;
            ctr[type == 0 ? 1 : 3]++;
          } else {
            if (type == 1) {
              L.ele_idx[mj + t * ne] = rj;
              L.ele_cfg[ri + t * ns] = -1;
              L.ele_cfg[rj + t * ns] = mj;
              L.ele_num[ri + t * ns] = 0;
              L.ele_num[rj + t * ns] = 1;
            }
            L.ele_idx[mi + s * ne] = ri;
            L.ele_cfg[ri + s * ns] = mi;
            L.ele_cfg[rj + s * ns] = -1;
            L.ele_num[ri + s * ns] = 1;
            L.ele_num[rj + s * ns] = 0;
          }
        }
        __syncthreads();
        PHASE(4);
        const int acc = L.misc[M_ACC];
        apply_clear_deltas(v, L, cnt, acc != 0);
        if (acc) {
          if (type == 0) {
            // update_one_real_scalar (updates.rs:1184-1224), delayed: one
            // pass forms vec1 = -(X^T slt) of this update from the current
            // inverse (pending updates applied in registers), writing the
            // planes back only when the pending list is full; this update
            // joins the list.
            const int msa = mi + s * ne;
            const bool full = P.np == PEND_MAX;
            stream_planes<true>(planes, n, nqp, pend_c, P, full, slt, vec1, ld);
            if (full && wpasses && tid == 0) atomicAdd(wpasses, 1);
            __syncthreads();
            if (full) P.np = 0;
            const int k = P.np;
            for (int e = tid; e < nqp * n; e += nt) {
              const int q = e / n;
              const int j = e - q * n;
              const double inv_a = -1.0 / vec1[q * ld + msa];
              pend[(size_t)e * PEND_MAX + k] = PendPair<Elem>::make(vec1[q * ld + j], stage[q * ld + j] * inv_a);
            }
#pragma unroll
            for (int kk = 0; kk < PEND_MAX; ++kk)
              if (kk == k) P.pm[kk] = msa;
            P.np = k + 1;
            for (int q = tid; q < nqp; q += nt) L.pf[q] = L.pf_new[q];
            __syncthreads();
            PHASE(5);
          } else {
            // update_two_real (updates.rs:1429-1503); pf *= ratio recomputed here
            const int msa = mi + s * ne, msb = mj + t * ne;
            const int rsa = L.ele_idx[msa] + s * ns, rsb = L.ele_idx[msb] + t * ns;
            const int rsa_old = ri + s * ns, rsb_old = rj + t * ns;
            double* P = L.uni;
            double* Q = L.uni + nqp * n;
            double* Sv = L.uni + 2 * nqp * n;
            double* T = L.uni + 3 * nqp * n;
            double* cf = L.uni + 5 * nqp * n;  // [nqp][16]
            for (int e = tid; e < nqp * n; e += nt) {
              const int q = e / n, i = e % n;
              const int rsi = L.ele_idx[i] + (i < ne ? 0 : ns);
              const double* S = v.slater + (size_t)q * n2 * n2;
              Sv[e] = (i == msb) ? S[(size_t)rsa_old * n2 + rsb_old] : S[(size_t)rsa * n2 + rsi];
              T[e] = S[(size_t)rsb * n2 + rsi];
            }
            __syncthreads();
            for (int e = tid; e < nqp * n; e += nt) {
              const int q = e / n, i = e % n;
              const Elem* row = planes + (size_t)q * nn + (size_t)i * n;
              double p = 0.0, qq = 0.0;
              for (int j = 0; j < n; ++j) {
                const double x = row[j];
                p += x * Sv[q * n + j];
                qq += x * T[q * n + j];
              }
              P[e] = p;
              Q[e] = qq;
            }
            __syncthreads();
            for (int q = tid; q < nqp; q += nt) {
              const double* Pq = P + q * n;
              const double* Qq = Q + q * n;
              const double* Tq = T + q * n;
              double bma = 0.0;
              for (int i = 0; i < n; ++i) bma += Tq[i] * Pq[i];
              const double iab = planes[(size_t)q * nn + (size_t)msa * n + msb];
              const double ratio = iab * Tq[msa] + iab * bma + Pq[msa] * Qq[msb] - Pq[msb] * Qq[msa];
              L.pf[q] *= ratio;
              const double a = -Pq[msa], b = Pq[msb], c = Qq[msa], d = -Qq[msb];
              const double ee = -bma - Tq[msa];
              const double f = iab;
              const double det = a * d - b * c - ee * f;
              double* k = cf + q * 16;
              k[0] = a; k[1] = b; k[2] = c; k[3] = d; k[4] = ee; k[5] = f; k[6] = det; k[7] = 1.0 / det;
            }
            __syncthreads();
            for (int e = tid; e < nqp * n; e += nt) {  // S, T rows of the old inverse, scaled
              const int q = e / n, i = e % n;
              const double idet = cf[q * 16 + 7];
              const double sa = planes[(size_t)q * nn + (size_t)msa * n + i];
              const double sb = planes[(size_t)q * nn + (size_t)msb * n + i];
              Sv[e] = idet * sa;
              T[e] = idet * sb;
            }
            __syncthreads();
            for (size_t e = tid; e < (size_t)nqp * nn; e += nt) {
              const int q = (int)(e / nn), r = (int)(e % nn);
              const int i = r / n, j = r % n;
              const double* k = cf + q * 16;
              const double a = k[0], b = k[1], c = k[2], d = k[3], ee = k[4], f = k[5], det = k[6], idet = k[7];
              const double pi = P[q * n + i], qi = Q[q * n + i], si = Sv[q * n + i], ti = T[q * n + i];
              const double pj = P[q * n + j], qj = Q[q * n + j], sj = Sv[q * n + j], tj = T[q * n + j];
              double x = planes[e];
              x += a * (qi * tj - qj * ti) + b * (qi * sj - qj * si) + c * (pi * tj - pj * ti) +
                   d * (pi * sj - pj * si) + ee * det * (si * tj - sj * ti) + f * idet * (pi * qj - qi * pj);
              if (j == msa) x += -c * ti - d * si - f * idet * qi;
              if (j == msb) x += -a * ti - b * si + f * idet * pi;
              if (i == msa) x += c * tj + d * sj + f * idet * qj;
              if (i == msb) x += a * tj + b * sj - f * idet * pj;
              if (i == msa && j == msb) x += f * idet;
              if (i == msb && j == msa) x -= f * idet;
              planes[e] = static_cast<Elem>(x);
            }
            __syncthreads();
          }
        } else {
          __syncthreads();
        }
      }
      // ---- periodic recompute (engine.rs:454-473): handed to k_rebuild ----
      const int nacc_now = L.misc[M_NACC];
      __syncthreads();
      if (nacc_now > ns) {
        stop_at = step + 1;
        break;
      }
    }
    if (stop_at >= 0) {
      // the planes are rebuilt from the configuration, so the pending
      // updates are dropped; the walker resumes at proposal stop_at
      store_walker(v, w, L);
      for (int q = tid; q < nqp; q += nt) v.pf[(size_t)w * nqp + q] = L.pf[q];
      if (tid == 0) {
        v.log_ip[w] = L.dmisc[1];
        ws[1] = L.misc[M_NACC];
        ws[4] = stop_at;
        ws[5] = 1;
        if (GROUP) ga.glist[atomicAdd(ga.gq, 1)] = qe;
        else v.rlist[atomicAdd(v.nflag, 1)] = w;
      }
      __syncthreads();
      PHASE(6);
      continue;
    }
    // ---- end of block: write the pending updates back, persist state, save sample ----
    flush_pending(planes, n, nqp, pend_c, P, wpasses);
    store_walker(v, w, L);
    for (int q = tid; q < nqp; q += nt) v.pf[(size_t)w * nqp + q] = L.pf[q];
    if (tid == 0) {
      v.log_ip[w] = L.dmisc[1];
      ws[1] = L.misc[M_NACC];
      ws[4] = n_in;
      ws[6] = 1;
      if (out_step == n_out - 1) ws[0] = 1;  // save_burn + counter[9] = 1
    }
    if (out_step + v.nsample >= n_out) {
      const int smp = out_step + v.nsample - n_out;
      if (smp < v.nsample) {
        const size_t ws_ = (size_t)w * v.nsample + smp;
        for (int i = tid; i < n; i += nt) v.s_idx[ws_ * n + i] = L.ele_idx[i];
        for (int i = tid; i < n2; i += nt) {
          v.s_cfg[ws_ * n2 + i] = L.ele_cfg[i];
          v.s_num[ws_ * n2 + i] = L.ele_num[i];
        }
        for (int i = tid; i < v.nproj; i += nt) v.s_proj[ws_ * v.nproj + i] = cnt[i];
      }
    }
    __syncthreads();
    PHASE(8);
  }
}

// ---------------------------------------------------------------------------
// Periodic rebuild of the walkers k_block handed over (engine.rs:454-473):
// calc_m_all from the configuration; on success the Pfaffians and
// log|<x|psi>| are replaced; the accept window restarts either way.  A
// kernel of its own so the elimination's register budget does not limit the
// sampler's occupancy.
// ---------------------------------------------------------------------------
__host__ __device__ inline size_t lds_bytes_rebuild(const EnsView& v) {
  return 2 * rnd16((size_t)v.n * 4) + rnd16((size_t)v.nqp * 8) + rnd16(pf_bytes(v.n, v.nqp));
}

// Two compiled variants, bit-identical in results (only the schedule of the
// elimination differs), picked at run time by measurement: MINB workgroups
// per CU, JB columns per batch of the panel close.
// One walker's periodic rebuild (engine.rs:454-473): calc_m_all from its
// configuration; on success the Pfaffians and log|<x|psi>| are replaced; the
// accept window restarts either way.
template <int JB>
__device__ __forceinline__ void rebuild_walker(const EnsView& v, int w, int* ei, int* rs, double* pfn, char* pfscr,
                                               double* work, long long* clk) {
  const int tid = threadIdx.x;
  int* ws = v.wstate + (size_t)w * 8;
  for (int i = tid; i < v.n; i += blockDim.x) ei[i] = v.ele_idx[(size_t)w * v.n + i];
  __syncthreads();
  double* planes = v.inv + (size_t)w * v.nqp * v.n * v.n;
  const int bad = wg_calc_m_all<JB>(v, ei, rs, planes, work, pfn, pfscr, false,
                                clk);
  copy_inv32(v, w);
  if (tid == 0) {
    if (!bad) {
      for (int q = 0; q < v.nqp; ++q) v.pf[(size_t)w * v.nqp + q] = pfn[q];
      v.log_ip[w] = ddm::cr_log(fabs(ip_of(v, pfn)));
    }
    ws[1] = 0;
    ws[5] = 0;
  }
  __syncthreads();
}

template <int MINB, int JB>
__global__ void __launch_bounds__(256, MINB) k_rebuild(EnsView v) {
  extern __shared__ __align__(16) char lds[];
  Carver c{lds};
  int* ei = c.take<int>(v.n);
  int* rs = c.take<int>(v.n);
  double* pfn = c.take<double>(v.nqp);
  char* pfscr = (char*)c.take<char>(pf_bytes(v.n, v.nqp));
  const int tid = threadIdx.x;
  double* work = v.work + (size_t)blockIdx.x * pfinv_work_doubles(v.n, v.nqp);
  const int nhand = *v.nflag;  // walkers handed over by the k_block launch before
  __shared__ int next_i;
  for (;;) {
    // hand-over entries are taken dynamically: every slot stays busy until
    // the list is done
    if (tid == 0) next_i = atomicAdd(v.nflag + 2, 1);
    __syncthreads();
    const int li = next_i;
    __syncthreads();
    if (li >= nhand) break;
    const int w = v.rlist[li];
    rebuild_walker<JB>(v, w, ei, rs, pfn, pfscr, work, v.clk ? v.clk + (size_t)blockIdx.x * 32 + 16 : nullptr);
  }
}

// The same over a group's shared hand-over list: entries of any run, the
// LDS carve following the entry's run, scratch slot blockIdx.x of the
// group's buffers.
template <int MINB, int JB>
__global__ void __launch_bounds__(256, MINB) k_rebuild_q(GroupArgs ga) {
  extern __shared__ __align__(16) char lds[];
  const int tid = threadIdx.x;
  double* work = ga.work + (size_t)blockIdx.x * ga.work_stride;
  const int nhand = *ga.gq;  // entries handed over by the k_block launch before
  __shared__ int next_i;
  for (;;) {
    if (tid == 0) next_i = atomicAdd(ga.gq + 2, 1);
    __syncthreads();
    const int li = next_i;
    __syncthreads();
    if (li >= nhand) break;
    int w = 0;
    const EnsView& v = ga.views[queue_run(ga, ga.glist[li], w)];
    Carver c{lds};
    int* ei = c.take<int>(v.n);
    int* rs = c.take<int>(v.n);
    double* pfn = c.take<double>(v.nqp);
    char* pfscr = (char*)c.take<char>(pf_bytes(v.n, v.nqp));
    rebuild_walker<JB>(v, w, ei, rs, pfn, pfscr, work, nullptr);
  }
}

// Start of an outer step: every walker resumes at proposal 0.
__global__ void k_reset_resume(EnsView v) {
  const int w = blockIdx.x * blockDim.x + threadIdx.x;
  if (w >= v.W) return;
  v.wstate[(size_t)w * 8 + 4] = 0;
  v.wstate[(size_t)w * 8 + 6] = 0;
}

// ---------------------------------------------------------------------------
// measurement of the saved samples (run.rs:2937-3616)
// ---------------------------------------------------------------------------
__device__ inline double julia_reciprocal_re(double c) {
  const double cd = fabs(c);
  if (1.0547686614863e-154 <= cd && cd <= 9.480751908109176e153) {
    const double norm = fma(cd, cd, 0.0 * 0.0);
    return c / norm;
  }
  return 1.0 / c;
}

// direct projection ratio of a hop source -> dest of `spin` (observables.rs:1638-1672)
__device__ double direct_proj_ratio_moved(const EnsView& v, int source, int dest, int spin, const int* en) {
  const int n = v.ns;
  const int src_slot = source + spin * n, dst_slot = dest + spin * n;
  auto occ = [&](int slot) { return slot == src_slot ? 0 : (slot == dst_slot ? 1 : en[slot]); };
  const double* row_dest = v.dproj_j + (size_t)dest * n;
  const double* row_source = v.dproj_j + (size_t)source * n;
  double z = 0.0;
  z -= v.dproj_g[source] * (double)(occ(source) + occ(n + source));
  z += v.dproj_g[dest] * (double)(occ(dest) * occ(n + dest));
  z += row_source[dest] * (double)((occ(source) + occ(n + source) - 1) - (occ(dest) + occ(n + dest) - 1) + 1);
  for (int site = 0; site < n; ++site)
    if (site != source && site != dest) z += (row_dest[site] - row_source[site]) * (double)(en[site] + en[n + site] - 1);
  return ddm::cr_exp(z);
}

// Rebuild for measurement: the inverse and Pfaffians of saved sample s of
// every walker into the walker's planes and pf; s_valid[w][s] = the sample
// is not the all-zero / all-negative placeholder and calc_m_all succeeded.
template <int MINB, int JB, bool GROUP>
__global__ void __launch_bounds__(256, MINB) k_meas_rebuild(EnsView v_arg, GroupArgs ga, int s) {
  ENS_SLOT(v_arg, ga)
  extern __shared__ __align__(16) char lds[];
  Carver c{lds};
  int* ei = c.take<int>(v.n);
  int* rs = c.take<int>(v.n);
  double* pfn = c.take<double>(v.nqp);
  char* pfscr = (char*)c.take<char>(pf_bytes(v.n, v.nqp));
  const int tid = threadIdx.x;
  double* work = v.work + (size_t)slot * pfinv_work_doubles(v.n, v.nqp);
  __shared__ int flag;
  for (int w = slot; w < v.W; w += nslots) {
    const size_t ws = (size_t)w * v.nsample + s;
    for (int i = tid; i < v.n; i += blockDim.x) ei[i] = v.s_idx[ws * v.n + i];
    __syncthreads();
    if (tid == 0) {
      int all0 = 1, allneg = 1;
      for (int k = 0; k < v.n; ++k) { all0 &= (ei[k] == 0); allneg &= (ei[k] < 0); }
      flag = (all0 || allneg) ? 0 : 1;
    }
    __syncthreads();
    int valid = flag;
    if (valid) {
      double* planes = v.inv + (size_t)w * v.nqp * v.n * v.n;
      valid = !wg_calc_m_all<JB>(v, ei, rs, planes, work, pfn, pfscr, false);
    }
    if (tid == 0) {
      v.s_valid[ws] = valid;
      if (valid)
        for (int q = 0; q < v.nqp; ++q) v.pf[(size_t)w * v.nqp + q] = pfn[q];
    }
    __syncthreads();
  }
}

template <bool GROUP>
__global__ void __launch_bounds__(256, 2) k_measure(EnsView v_arg, GroupArgs ga, int mode_exp, int s_only) {
  ENS_SLOT(v_arg, ga)
  extern __shared__ __align__(16) char lds[];
  Carver c{lds};
  int* ei = c.take<int>(v.n);
  int* ec = c.take<int>(v.n2);
  int* en = c.take<int>(v.n2);
  int* epc = c.take<int>(v.nproj > 0 ? v.nproj : 1);
  double* pf = c.take<double>(v.nqp);
  double* dm = c.take<double>(32);
  int* rs = c.take<int>(v.n);
  char* pfscr = (char*)c.take<char>(pf_bytes(v.n, v.nqp));
  int* tri = c.take<int>((size_t)v.nqp * v.n);
  int* sgi = c.take<int>((size_t)v.nqp * v.n);
  int* flags = c.take<int>(64);
  const int tid = threadIdx.x, nt = blockDim.x;
  const int n = v.n, ne = v.ne, ns = v.ns, n2 = v.n2, nqp = v.nqp, nn = n * n;
  double* work = v.work + (size_t)slot * pfinv_work_doubles(n, nqp);
  double* qorb = v.qorb + (size_t)slot * nqp * v.nslater;
  double* g1 = v.gscr + (size_t)slot * v.gstride;  // per term
  for (int w = slot; w < v.W; w += nslots) {
    double* planes = v.inv + (size_t)w * nqp * nn;
    for (int s = s_only; s == s_only; ++s) {
      const size_t ws = (size_t)w * v.nsample + s;
      double* o = v.o_store + ws * v.size;
      for (int i = tid; i < v.size; i += nt) o[i] = 0.0;
      for (int i = tid; i < n; i += nt) ei[i] = v.s_idx[ws * n + i];
      for (int i = tid; i < n2; i += nt) { ec[i] = v.s_cfg[ws * n2 + i]; en[i] = v.s_num[ws * n2 + i]; }
      for (int i = tid; i < v.nproj; i += nt) epc[i] = v.s_proj[ws * v.nproj + i];
      if (tid == 0) flags[0] = v.s_valid[ws];  // set by k_meas_rebuild
      for (int q = tid; q < nqp; q += nt) pf[q] = v.pf[(size_t)w * nqp + q];
      __syncthreads();
      int valid = flags[0];
      double ip = 0.0;
      if (valid) {
        if (tid == 0) dm[0] = ip_of(v, pf);
        __syncthreads();
        ip = dm[0];
        if (fabs(ip) < 1.0e-100) valid = 0;
      }
      double e = 0.0;
      if (valid) {
        // diagonal part (observables.rs:483-552)
        if (tid == 0) {
          double d = 0.0;
          for (int k = 0; k < v.nintra; ++k) { int r = v.intra_site[k]; d += v.intra_v[k] * (double)(en[r] * en[r + ns]); }
          for (int k = 0; k < v.ninter; ++k) {
            int a = v.inter_site[2 * k], b = v.inter_site[2 * k + 1];
            d += v.inter_v[k] * (double)(en[a] + en[a + ns]) * (double)(en[b] + en[b + ns]);
          }
          for (int k = 0; k < v.nhund; ++k) {
            int a = v.hund_site[2 * k], b = v.hund_site[2 * k + 1];
            d -= v.hund_v[k] * ((double)(en[a] * en[b]) + (double)(en[a + ns] * en[b + ns]));
          }
          dm[1] = d;
        }
        // transfer Green ratios, one term per lane (calh1_real_fast)
        if (v.ntrans > 0 && fabs(ip) > 0.0) {
          for (int k = tid; k < v.ntrans; k += nt) {
            const int ri = v.trans[4 * k], s1 = v.trans[4 * k + 1], rj = v.trans[4 * k + 2], s2 = v.trans[4 * k + 3];
            const int dst = ri + s1 * ns, src = rj + s2 * ns;
            double g;
            if (s1 == s2 && ri == rj) g = (double)en[src];
            else if (s1 != s2 || en[dst] == 1 || en[src] == 0 || ec[src] < 0) g = 0.0;
            else {
              const int mj = ec[src];
              const int msa = mj + s2 * ne;
              const double pr = direct_proj_ratio_moved(v, rj, ri, s2, en);
              const int rsa = ri + s2 * ns;
              double nip = 0.0;
              for (int q = 0; q < nqp; ++q) {
                const double* row = v.slater + ((size_t)q * n2 + rsa) * n2;
                const double* iv = planes + (size_t)q * nn + (size_t)msa * n;
                double ratio = 0.0;
                for (int msj = 0; msj < n; ++msj) {
                  const int site = (msj == msa) ? ri : ei[msj];
                  ratio += iv[msj] * row[site + (msj < ne ? 0 : ns)];
                }
                const double npf = -ratio * pf[q];
                nip += v.qpw[q] * npf;
              }
              g = pr * nip / ip;
            }
            g1[k] = g;
          }
          __syncthreads();
          if (tid == 0) {
            double te = 0.0;
            for (int k = 0; k < v.ntrans; ++k) te -= v.trans_v[k] * g1[k];
            dm[1] = dm[1] + te;
          }
        }
        __syncthreads();
        // exchange terms (green_func2 general case), one term per lane
        if (v.nexch > 0 && fabs(ip) > 0.0) {
          for (int k = tid; k < v.nexch; k += nt) {
            const int a = v.exch_site[2 * k], b = v.exch_site[2 * k + 1];
            double gsum[2] = {0.0, 0.0};
            for (int o2 = 0; o2 < 2; ++o2) {
              const int spin = o2 == 0 ? 0 : 1, spin_o = 1 - spin;
              const int ri = a, rj = b, rk = b, rl = a;
              double gv = 0.0;
              if (ri != rj && en[ri + spin * ns] == 0 && en[rj + spin * ns] == 1 && en[rk + spin_o * ns] == 0 &&
                  en[rl + spin_o * ns] == 1) {
                const int mjx = ec[rj + spin * ns], mix = ec[rl + spin_o * ns];
                if (mjx >= 0 && mix >= 0) {
                  const int msj = mjx + spin * ne, msi = mix + spin_o * ne;
                  // projection deltas of the two hops (rl->rk spin_o, then rj->ri spin)
                  double lpd = 0.0;
                  {
                    // count changes via a small local map: evaluate log_proj_ratio over all p
                    // in ascending order from the exact integer deltas.
                    // new occupation after both hops
                    auto num_after = [&](int slot) {
                      if (slot == rl + spin_o * ns) return 0;
                      if (slot == rk + spin_o * ns) return 1;
                      if (slot == rj + spin * ns) return 0;
                      if (slot == ri + spin * ns) return 1;
                      return en[slot];
                    };
                    auto num_mid = [&](int slot) {
                      if (slot == rl + spin_o * ns) return 0;
                      if (slot == rk + spin_o * ns) return 1;
                      return en[slot];
                    };
                    double z = 0.0;
                    for (int p = 0; p < v.nproj; ++p) {
                      int d = 0;
                      // Gutzwiller / Jastrow deltas of hop 1 (rl -> rk) with mid occupations
                      // and of hop 2 (rj -> ri) with final occupations
                      for (int hop = 0; hop < 2; ++hop) {
                        const int hri = hop == 0 ? rl : rj, hrj = hop == 0 ? rk : ri;
                        auto nm = [&](int slot) { return hop == 0 ? num_mid(slot) : num_after(slot); };
                        if (hri == hrj) continue;
                        if (p < v.ngutz) {
                          if (v.gutz_idx[hri] == p) d -= nm(hri) + nm(hri + ns);
                          if (v.gutz_idx[hrj] == p) d += nm(hrj) * nm(hrj + ns);
                        } else {
                          const int jp = p - v.ngutz;
                          if (jast_of(v, hri, hrj) == jp) d += (nm(hri) + nm(hri + ns)) - (nm(hrj) + nm(hrj + ns)) + 1;
                          for (int rkk = 0; rkk < ns; ++rkk) {
                            if (rkk == hri || rkk == hrj) continue;
                            const int x = nm(rkk) + nm(rkk + ns) - 1;
                            if (jast_of(v, hri, rkk) == jp) d -= x;
                            if (jast_of(v, hrj, rkk) == jp) d += x;
                          }
                        }
                      }
                      if (d != 0) z += v.proj[p] * (double)d;
                    }
                    lpd = z;
                  }
                  const double proj_ratio = ddm::cr_exp(lpd);
                  // two-electron ratio with mi (spin_o) moved to rk and mj (spin) to ri
                  const int msa = msi, msb = msj;
                  auto site_after = [&](int m) { return m == msi ? rk : (m == msj ? ri : ei[m]); };
                  const int rsa = rk + spin_o * ns, rsb = ri + spin * ns;
                  double nip = 0.0;
                  for (int q = 0; q < nqp; ++q) {
                    const double* S = v.slater + (size_t)q * n2 * n2;
                    const double* inv = planes + (size_t)q * nn;
                    double pa = 0, pb = 0, qa = 0, qb = 0, bma = 0;
                    for (int i = 0; i < n; ++i) {
                      const int rsi = site_after(i) + (i < ne ? 0 : ns);
                      const double vai = S[(size_t)rsa * n2 + rsi], vbi = S[(size_t)rsb * n2 + rsi];
                      const double xa = inv[(size_t)msa * n + i], xb = inv[(size_t)msb * n + i];
                      pa += xa * vai; pb += xb * vai; qa += xa * vbi; qb += xb * vbi;
                    }
                    for (int i = 0; i < n; ++i) {
                      double tmp = 0.0;
                      for (int j = 0; j < n; ++j) {
                        const int rsj = site_after(j) + (j < ne ? 0 : ns);
                        tmp += inv[(size_t)i * n + j] * S[(size_t)rsa * n2 + rsj];
                      }
                      const int rsi = site_after(i) + (i < ne ? 0 : ns);
                      bma += S[(size_t)rsb * n2 + rsi] * tmp;
                    }
                    const double vba = S[(size_t)rsb * n2 + (site_after(msa) + (msa < ne ? 0 : ns))];
                    const double iab = inv[(size_t)msa * n + msb];
                    const double ratio = iab * vba + iab * bma + pa * qb - pb * qa;
                    const double npf = ratio * pf[q];
                    nip += v.qpw[q] * npf;
                  }
                  gv = (proj_ratio * nip) / ip;
                }
              }
              gsum[o2] = gv;
            }
            g1[k] = gsum[0] + gsum[1];
          }
          __syncthreads();
          if (tid == 0) {
            double d = dm[1];
            for (int k = 0; k < v.nexch; ++k) d += v.exch_v[k] * g1[k];
            dm[1] = d;
          }
        }
        __syncthreads();
        e = dm[1];
        if (!(e + 0.0 - (e + 0.0) == 0.0)) valid = 0;
      }
      if (tid == 0) {
        v.s_valid[ws] = valid;
        v.e_loc[ws] = valid ? e : 0.0;
        long long szs = 0;
        for (int r = 0; r < ns; ++r) szs += en[r] - en[r + ns];
        v.sz[ws] = (double)szs;
      }
      if (valid) {
        // O = [1, projection counts, Slater derivative]
        if (tid == 0) o[0] = 1.0;
        for (int p = tid; p < v.nproj; p += nt) o[1 + p] = (double)epc[p];
        if (v.nslater > 0) {
          for (size_t k = tid; k < (size_t)nqp * v.nslater; k += nt) qorb[k] = 0.0;
          for (int e2 = tid; e2 < nqp * n; e2 += nt) {
            const int q = e2 / n, m = e2 % n;
            const int tr = q / v.nleg;
            tri[e2] = v.tsite[(size_t)tr * ns + ei[m]];
            sgi[e2] = v.tsgn[(size_t)tr * ns + ei[m]];
          }
          __syncthreads();
          auto contrib = [&](int q, int msi, int msj) {
            const int sp = q % v.nleg;
            const double p = pf[q];
            const double cs = p * v.spgl_cs[sp], cc = p * v.spgl_cc[sp], ss = p * v.spgl_ss[sp];
            const double first = msi < ne ? cs : ss;
            const double second = msi < ne ? cc : cs;
            const int ti = tri[q * n + msi], tj = tri[q * n + msj];
            const int orb = v.orb_idx[(size_t)ti * ns + tj];
            if (orb < 0 || orb >= v.nslater) return;
            const int sg = sgi[q * n + msi] * sgi[q * n + msj] * v.orb_sgn[(size_t)ti * ns + tj];
            const double x = planes[(size_t)q * nn + msj + (size_t)msi * n];
            const double value = msj < ne ? x * first : -x * second;
            qorb[(size_t)q * v.nslater + orb] += value * (double)sg;
          };
          if (v.orb_injective) {
            // ordered scatter: (msi, spin block of msj) phases, disjoint targets within a phase
            for (int msi = 0; msi < n; ++msi) {
              for (int blk = 0; blk < 2; ++blk) {
                for (int e2 = tid; e2 < nqp * ne; e2 += nt) {
                  const int q = e2 / ne, msj = e2 % ne + blk * ne;
                  contrib(q, msi, msj);
                }
                __syncthreads();
              }
            }
          } else {
            for (int q = tid; q < nqp; q += nt)
              for (int msi = 0; msi < n; ++msi)
                for (int msj = 0; msj < n; ++msj) contrib(q, msi, msj);
            __syncthreads();
          }
          const double inv_ip = julia_reciprocal_re(ip);
          for (int orb = tid; orb < v.nslater; orb += nt) {
            double acc = 0.0;
            for (int q = 0; q < nqp; ++q) acc += v.qpw[q] * qorb[(size_t)q * v.nslater + orb];
            o[1 + v.nproj + orb] = acc * inv_ip;
          }
        }
      }
      __syncthreads();
    }
  }
}

// ---------------------------------------------------------------------------
// reductions over walkers in walker (= rank) order (run.rs:1863-1903,
// multichain.rs:474-600): per walker the sum over its samples, then a left
// fold over walkers.
// ---------------------------------------------------------------------------
// Direct SR assembly on the upper triangle (sr_backend.rs:484-505):
// S[i][j] = Gram[i][j] - mean[i] * mean[j], the diagonal times ratio_diag.
__global__ void k_sr_assemble(int na, const double* mean, double ratio_diag, double* S) {
  const size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= (size_t)na * na) return;
  const int i = (int)(e % na), j = (int)(e / na);
  if (i > j) return;
  double x = S[e] - mean[i] * mean[j];
  if (i == j) x *= ratio_diag;
  S[e] = x;
}

// log|psi(x)| = log|<x|phi>| + sum_p proj_p cnt_p(x) of saved sample s of
// every walker at the current parameters (calculate_ip_real,
// observables.rs:128-145, and the projection factor), after
// k_meas_rebuild(s); NaN for an invalid sample.  One thread per walker.
__global__ void k_log_psi(EnsView v, int s, double* out) {
  const int w = blockIdx.x * blockDim.x + threadIdx.x;
  if (w >= v.W) return;
  const size_t ws = (size_t)w * v.nsample + s;
  double r = __builtin_nan("");
  if (v.s_valid[ws]) {
    const double ip = ip_of(v, v.pf + (size_t)w * v.nqp);
    double lp = 0.0;
    for (int p = 0; p < v.nproj; ++p) lp += v.proj[p] * (double)v.s_proj[ws * v.nproj + p];
    r = ddm::cr_log(fabs(ip)) + lp;
  }
  out[ws] = r;
}

// Vector-Jacobian product over the stored samples: grad[k] = sum_r cot[r] *
// O[r][k], r ascending (r = walker * NVMCSample + sample), O = [1,
// d log|psi| / d p] per sample as the measurement stored it.  One thread
// per component; consecutive threads read consecutive components.
__global__ void k_vjp(EnsView v, const double* cot, double* grad) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= v.size) return;
  const size_t rows = (size_t)v.W * v.nsample;
  double acc = 0.0;
  for (size_t r = 0; r < rows; ++r) acc += cot[r] * v.o_store[r * v.size + k];
  grad[k] = acc;
}

// ---------------------------------------------------------------------------
// Adjoint VJP (derivative form): grad = sum_r cot[r] O[r] without the
// per-sample O rows.  d log Pf(A) = tr(A^-1 dA) / 2, so the cotangent of the
// gathered plane is the inverse itself; by linearity of the Slater table in
// the orbital parameters (update_slater_elm, slater_update.rs:17-126) the
// orbital map is applied once to the sum over samples instead of once per
// sample (slater_derivative.rs:289-356 assembles per sample and per QP).
// ---------------------------------------------------------------------------
// Saved sample s of every walker (after k_meas_rebuild(s)): element (i, j)
// of inverse plane q goes back to the Slater-table position it was gathered
// from, sbar[q][rs_i][rs_j] += (cot w_q Pf_q / ip) invM_q[i][j] (rs_i =
// site + spin * Nsite of electron i: distinct electrons sit on distinct
// (site, spin), so one sample writes every position at most once per
// plane).  Workgroup `slot` owns sbar + slot * nqp * n2^2 and gproj +
// slot * (1 + nproj) and takes walkers slot, slot + nslots, ... in order,
// so every accumulator sums in a fixed order.  gproj gets cot (the O row's
// constant component) and cot * cnt_p (the projection components).
// Invalid samples (zero O rows) add nothing.
__global__ void __launch_bounds__(256) k_adj_scatter(EnsView v, int s, const double* cot, double* sbar,
                                                     double* gproj) {
  extern __shared__ __align__(16) char lds[];
  Carver c{lds};
  int* rs = c.take<int>(v.n);
  double* coef = c.take<double>(v.nqp);
  __shared__ int ok_s;
  const int slot = blockIdx.x, nslots = gridDim.x, tid = threadIdx.x, nt = blockDim.x;
  const int n = v.n, n2 = v.n2, nqp = v.nqp;
  const size_t nn = (size_t)n * n, pl = (size_t)n2 * n2;
  double* sb = sbar + (size_t)slot * nqp * pl;
  double* gp = gproj + (size_t)slot * (1 + v.nproj);
  for (int w = slot; w < v.W; w += nslots) {
    const size_t ws = (size_t)w * v.nsample + s;
    const double ct = cot[ws];
    if (tid == 0) {
      int ok = v.s_valid[ws];
      const double* pf = v.pf + (size_t)w * nqp;
      if (ok) {
        const double ip = ip_of(v, pf);
        if (fabs(ip) < 1.0e-100) ok = 0;
        const double inv_ip = julia_reciprocal_re(ip);
        for (int q = 0; q < nqp; ++q) coef[q] = ct * (v.qpw[q] * (pf[q] * inv_ip));
      }
      ok_s = ok;
    }
    for (int i = tid; i < n; i += nt) rs[i] = v.s_idx[ws * n + i] + (i / v.ne) * v.ns;
    __syncthreads();
    if (ok_s) {
      if (tid == 0) gp[0] += ct;
      for (int p = tid; p < v.nproj; p += nt) gp[1 + p] += ct * (double)v.s_proj[ws * v.nproj + p];
      if (v.nslater > 0) {
        const double* planes = v.inv + (size_t)w * nqp * nn;
        for (size_t e = tid; e < (size_t)nqp * nn; e += nt) {
          const int q = (int)(e / nn);
          const int r = (int)(e % nn);
          const int i = r / n, j = r % n;
          double& x = sb[q * pl + (size_t)rs[i] * n2 + rs[j]];
          x += coef[q] * planes[e];
        }
      }
    }
    __syncthreads();
  }
}

// buf[e] = sum over slots (ascending) of buf[slot * len + e].
__global__ void k_adj_fold(double* buf, int nslots, long long len) {
  const long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= len) return;
  double a = buf[e];
  for (int sl = 1; sl < nslots; ++sl) a += buf[(size_t)sl * len + e];
  buf[e] = a;
}

// grad[0 .. nproj] = gproj; grad[1 + nproj + orb] = sum over the orbital's
// table positions k = ent[start[orb] .. start[orb + 1]) (ascending (q, a,
// b)) of sbar[k] * wt[k], wt = translation signs x orbital sign x the
// Gauss-Legendre spin factor of the (spin a, spin b) block.
__global__ void k_adj_orbital(const double* sbar, const int* start, const long long* ent, const double* wt,
                              const double* gproj, int nproj, int nslater, double* grad) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k < 1 + nproj) {
    grad[k] = gproj[k];
    return;
  }
  const int orb = k - 1 - nproj;
  if (orb >= nslater) return;
  double a = 0.0;
  for (int t = start[orb]; t < start[orb + 1]; ++t) a += sbar[ent[t]] * wt[t];
  grad[k] = a;
}

__global__ void k_reduce_energy(EnsView v, double* out5) {
  if (blockIdx.x != 0 || threadIdx.x != 0) return;
  double tot[5] = {0, 0, 0, 0, 0};
  for (int w = 0; w < v.W; ++w) {
    double a[5] = {0, 0, 0, 0, 0};
    for (int s = 0; s < v.nsample; ++s) {
      const size_t ws = (size_t)w * v.nsample + s;
      if (!v.s_valid[ws]) continue;
      const double e = v.e_loc[ws], sz = v.sz[ws], wt = 1.0;
      a[0] += wt;
      a[1] += wt * e;
      a[2] += (wt * e) * e;
      a[3] += wt * sz;
      a[4] += wt * sz * sz;
    }
    for (int k = 0; k < 5; ++k) tot[k] = (w == 0) ? a[k] : tot[k] + a[k];
  }
  for (int k = 0; k < 5; ++k) out5[k] = tot[k];
}

// ho[i], mean[i], diag[i] (CG mode: finalize_oo_store_real diagonal_only)
__global__ void k_reduce_ops(EnsView v, double* ho, double* mean, double* diag) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= v.size) return;
  double th = 0, tm = 0, td = 0;
  for (int w = 0; w < v.W; ++w) {
    double h = 0.0, m = 0.0, d = 0.0;
    for (int s = 0; s < v.nsample; ++s) {
      const size_t ws = (size_t)w * v.nsample + s;
      const double o = v.o_store[ws * v.size + i];
      if (v.s_valid[ws]) h += (1.0 * v.e_loc[ws]) * o;
      m += o;
      d += o * o;
    }
    if (w == 0) { th = h; tm = m; td = d; }
    else { th += h; tm += m; td += d; }
  }
  ho[i] = th;
  mean[i] = tm;
  diag[i] = td;
}

// ---------------------------------------------------------------------------
// SR-CG (sr_cg.rs:700-880, 952-1004).  Samples are packed [rows][na],
// rows = W*nsample in walker order.  Reduction order (the reference's own
// products go through BLAS dgemv, whose order is the BLAS build's):
//   bdot: 256 strided lane partials (ascending index), then the pairwise
//         tree p_l += p_{l+h}, h = 128, 64, ..., 1;
//   S x:  y_r = bdot(O_r, x); rows cut into chunks of CG_RCH, per chunk
//         c_i = sum_r O_ri y_r (ascending r); chunks folded left within
//         superchunks of CG_SCH chunks, superchunks folded left.
// ---------------------------------------------------------------------------
template <class P>
__global__ void k_cg_pack(EnsView v, const int* map, int na, P* pack) {
  const size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  const size_t total = (size_t)v.W * v.nsample * na;
  if (e >= total) return;
  const size_t ws = e / na;
  const int si = (int)(e % na);
  pack[e] = static_cast<P>(v.o_store[ws * v.size + map[si] + 1]);
}

// ctl: 0 done, 1..2 status (int64), 3 first iteration of the current period
// of the replayed CG graph (kernels get their offset in the period)
enum { CTL_DONE = 0, CTL_STATUS_LO = 1, CTL_STATUS_HI = 2, CTL_ITER = 3 };

__device__ inline void cg_finish(int* ctl, long long status) {
  ctl[CTL_STATUS_LO] = (int)(status & 0xffffffffLL);
  ctl[CTL_STATUS_HI] = (int)(status >> 32);
  ctl[CTL_DONE] = 1;
}

// Pairwise tree over the 256 lane partials of a 256-thread block; result
// returned to every lane.  `red` is a 256-double LDS array.
__device__ double block_tree256(double p, double* red) {
  const int tid = threadIdx.x;
  red[tid] = p;
  __syncthreads();
  if (tid < 128) red[tid] += red[tid + 128];
  __syncthreads();
  if (tid < 64) {
    double v = red[tid] + red[tid + 64];
    for (int h = 32; h >= 1; h >>= 1) v += __shfl_down(v, h, 64);
    if (tid == 0) red[0] = v;
  }
  __syncthreads();
  const double r = red[0];
  __syncthreads();
  return r;
}

template <class TA>
__device__ inline double block_bdot(const TA* a, const double* b, int n, double* red) {
  double p = 0.0;
  for (int i = threadIdx.x; i < n; i += 256) p += (double)a[i] * b[i];
  return block_tree256(p, red);
}

// Row of n elements read through a raw buffer resource (stride 0, range
// n * sizeof(P) bytes, gfx9 dword3 0x00020000): an element offset past the
// range returns 0 without a memory access.
template <class P>
struct RowBuf;
template <>
struct RowBuf<double> {
  __amdgpu_buffer_rsrc_t rs;
  __device__ RowBuf(const double* p, int n)
      : rs(__builtin_amdgcn_make_buffer_rsrc((void*)p, (short)0, n * (int)sizeof(double), 0x00020000)) {}
  __device__ double load(int i) const {
    const auto u = __builtin_amdgcn_raw_buffer_load_b64(rs, i * (int)sizeof(double), 0, 0);
    return __builtin_bit_cast(double, u);
  }
};
template <>
struct RowBuf<float> {
  __amdgpu_buffer_rsrc_t rs;
  __device__ RowBuf(const float* p, int n)
      : rs(__builtin_amdgcn_make_buffer_rsrc((void*)p, (short)0, n * (int)sizeof(float), 0x00020000)) {}
  __device__ float load(int i) const {
    return __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(rs, i * (int)sizeof(float), 0, 0));
  }
};

// Chunk partials cpart[b][i] = sum_{r in chunk b} O_ri y_r with y_r formed
// from the same registers; block nch computes coef = bdot(mean, x).
// C = columns per lane (na <= 256*C).  P: element type of the packed samples
// (double; float in the precision form, accumulated in FP64).
template <int C, class P>
__global__ void __launch_bounds__(256, 2) k_cg_chunk(const P* pack, int rows, int na, const double* x,
                                                  const double* mean, double* cpart, double* coef, const int* ctl) {
  if (ctl[CTL_DONE]) return;
  __shared__ double red[256];
  const int tid = threadIdx.x;
  const int nch = (rows + CG_RCH - 1) / CG_RCH;
  const int b = blockIdx.x;
  if (b == nch) {
    const double c = block_bdot(mean, x, na, red);
    if (tid == 0) *coef = c;
    return;
  }
  double xv[C], acc[C];
#pragma unroll
  for (int k = 0; k < C; ++k) {
    const int i = tid + 256 * k;
    xv[k] = i < na ? x[i] : 0.0;
    acc[k] = 0.0;
  }
  const int r1 = min(rows, (b + 1) * CG_RCH);
  for (int r = b * CG_RCH; r < r1; ++r) {
    // one buffer resource per row: element i >= na reads 0 in hardware, so
    // the C loads carry no branch and are all in flight before the first use
    const RowBuf<P> row(pack + (size_t)r * na, na);
    P raw[C];
#pragma unroll
    for (int k = 0; k < C; ++k) raw[k] = row.load(tid + 256 * k);
    double o[C];
#pragma unroll
    for (int k = 0; k < C; ++k) o[k] = (double)raw[k];
    double p = 0.0;
#pragma unroll
    for (int k = 0; k < C; ++k)
      if (tid + 256 * k < na) p += o[k] * xv[k];
    const double y = block_tree256(p, red);
#pragma unroll
    for (int k = 0; k < C; ++k)
      if (tid + 256 * k < na) acc[k] += o[k] * y;
  }
#pragma unroll
  for (int k = 0; k < C; ++k) {
    const int i = tid + 256 * k;
    if (i < na) cpart[(size_t)b * na + i] = acc[k];
  }
}

// Same order for na > 256*CG_CMAX: the chunk accumulators live in cpart.
template <class P>
__global__ void __launch_bounds__(256) k_cg_chunk_wide(const P* pack, int rows, int na, const double* x,
                                                       const double* mean, double* cpart, double* coef,
                                                       const int* ctl) {
  if (ctl[CTL_DONE]) return;
  __shared__ double red[256];
  const int tid = threadIdx.x;
  const int nch = (rows + CG_RCH - 1) / CG_RCH;
  const int b = blockIdx.x;
  if (b == nch) {
    const double c = block_bdot(mean, x, na, red);
    if (tid == 0) *coef = c;
    return;
  }
  double* cp = cpart + (size_t)b * na;
  const int r0 = b * CG_RCH, r1 = min(rows, (b + 1) * CG_RCH);
  for (int r = r0; r < r1; ++r) {
    const P* row = pack + (size_t)r * na;
    const double y = block_bdot(row, x, na, red);
    for (int i = tid; i < na; i += 256) {
      double a = (r == r0) ? 0.0 : cp[i];
      a += (double)row[i] * y;
      cp[i] = a;
    }
  }
}

// Superchunk folds: spart[s][i] = c_{s*SCH} + ... (left fold)
__global__ void k_cg_fold1(const double* cpart, int nch, int na, double* spart, const int* ctl) {
  if (ctl[CTL_DONE]) return;
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const int s = blockIdx.y;
  if (i >= na) return;
  const int b0 = s * CG_SCH, b1 = min(nch, b0 + CG_SCH);
  double a = cpart[(size_t)b0 * na + i];
  for (int b = b0 + 1; b < b1; ++b) a += cpart[(size_t)b * na + i];
  spart[(size_t)s * na + i] = a;
}

// z = inv_w * fold_s(spart) - coef mean + shift diag x; resid: out = g - z.
__global__ void k_cg_fold2(const double* spart, int nsup, int na, const double* x, const double* mean,
                           const double* diag, const double* coef, double inv_w, double shift, const double* g,
                           double* out, const int* ctl) {
  if (ctl[CTL_DONE]) return;
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= na) return;
  double tot = nsup > 0 ? spart[i] : 0.0;
  for (int s = 1; s < nsup; ++s) tot += spart[(size_t)s * na + i];
  const double z = inv_w * tot - coef[0] * mean[i] + shift * diag[i] * x[i];
  out[i] = g ? g[i] - z : z;
}

__device__ inline double cg_precond(double pfac, double dg, double r) {
  const double EPS = 1.0e-12;
  double dq = pfac * dg;
  if (!(dq - dq == 0.0) || fabs(dq) < EPS) dq = dq < 0.0 ? -EPS : EPS;
  return r / dq;
}

// sc: 0 delta, 2 alpha, 4 beta, 5 coef.  ctl: done flag + 64-bit status.
// Start of the initial residual: diag-scale d = precond(r), delta = r.d;
// plain: delta = r.r.  Non-finite delta ends the solve with status -1.
__global__ void __launch_bounds__(256) k_cg_init(int na, const double* r, double* d, const double* diag, double pfac,
                                                 int diag_scale, double* sc, int* ctl) {
  __shared__ double red[256];
  double p = 0.0;
  for (int i = threadIdx.x; i < na; i += 256) {
    if (diag_scale) {
      const double q = cg_precond(pfac, diag[i], r[i]);
      d[i] = q;
      p += r[i] * q;
    } else {
      p += r[i] * r[i];
    }
  }
  const double delta = block_tree256(p, red);
  if (threadIdx.x == 0) {
    sc[0] = delta;
    ctl[CTL_DONE] = 0;
    if (!(delta - delta == 0.0)) cg_finish(ctl, -1);
  }
}

// Head of iteration `iter` (the loop condition of the C solver).
__global__ void k_cg_head(int j, int max_it, double threshold, const double* sc, int* ctl) {
  if (threadIdx.x || blockIdx.x || ctl[CTL_DONE]) return;
  const int iter = ctl[CTL_ITER] + j;
  const double delta = sc[0];
  if (iter >= max_it) {
    cg_finish(ctl, fabs(delta) >= threshold ? -(long long)(max_it + 1) : (long long)iter);
    return;
  }
  if (!(delta - delta == 0.0)) cg_finish(ctl, -(long long)(iter + 1));
  else if (fabs(delta) < threshold) cg_finish(ctl, iter);
}

// alpha = delta / (d.q) with the C guards
__global__ void __launch_bounds__(256) k_cg_alpha(int j, int na, const double* d, const double* q, double* sc,
                                                  int* ctl) {
  if (ctl[CTL_DONE]) return;
  const int iter = ctl[CTL_ITER] + j;
  __shared__ double red[256];
  const double dq = block_bdot(d, q, na, red);
  if (threadIdx.x == 0) {
    if (!(dq - dq == 0.0) || fabs(dq) < 1.0e-20) { cg_finish(ctl, -(long long)(iter + 1)); return; }
    const double a = sc[0] / dq;
    if (!(a - a == 0.0)) { cg_finish(ctl, -(long long)(iter + 1)); return; }
    sc[2] = a;
  }
}

// x += alpha d; r -= alpha q (not on refresh iterations)
__global__ void k_cg_update(int na, const double* sc, const int* ctl, const double* d, double* x, const double* q,
                            double* r, int refresh) {
  if (ctl[CTL_DONE]) return;
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= na) return;
  const double a = sc[2];
  x[i] += a * d[i];
  if (!refresh) r[i] -= a * q[i];
}

// dnew (diag-scale: q = precond(r), dnew = r.q; plain: r.r), beta, delta
__global__ void __launch_bounds__(256) k_cg_beta(int j, int na, const double* r, double* q, const double* diag,
                                                 double pfac, int diag_scale, double* sc, int* ctl) {
  if (ctl[CTL_DONE]) return;
  const int iter = ctl[CTL_ITER] + j;
  __shared__ double red[256];
  double p = 0.0;
  for (int i = threadIdx.x; i < na; i += 256) {
    if (diag_scale) {
      const double qq = cg_precond(pfac, diag[i], r[i]);
      q[i] = qq;
      p += r[i] * qq;
    } else {
      p += r[i] * r[i];
    }
  }
  const double dn = block_tree256(p, red);
  if (threadIdx.x == 0) {
    const double dl = sc[0];
    if (!(dn - dn == 0.0) || fabs(dl) < 1.0e-20) { cg_finish(ctl, -(long long)(iter + 1)); return; }
    const double b = dn / dl;
    if (!(b - b == 0.0)) { cg_finish(ctl, -(long long)(iter + 1)); return; }
    sc[4] = b;
    sc[0] = diag_scale ? dn : dl * b;
  }
}

// d = src + beta d
// End of a replayed period of the CG graph: the next period starts `period`
// iterations later.
__global__ void k_cg_advance(int* ctl, int period) {
  if (threadIdx.x == 0 && blockIdx.x == 0 && !ctl[CTL_DONE]) ctl[CTL_ITER] += period;
}

__global__ void k_cg_dir(int na, const double* sc, const int* ctl, const double* src, double* d) {
  if (ctl[CTL_DONE]) return;
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= na) return;
  d[i] = src[i] + sc[4] * d[i];
}

// ---------------------------------------------------------------------------
// Batched Pfaffian + inverse (BatchedEngine): planes [count][n][n]
// column-major; one workgroup per plane batch of np planes.
// ---------------------------------------------------------------------------
// Batched Pfaffian + inverse of n <= RR_NMAX planes, one plane per T x T
// workgroup at a time (planes blockIdx.x, + gridDim.x, ...), register-
// resident (pfinv_rr.h: the panel form, T = 32, for n > 64; the full form,
// T = 16, for n <= 64); gs: 2 n^2 doubles of scratch per workgroup.
// Results and statuses as k_pfinv_batched; in == out works in place (the
// plane is read whole before it is written).
template <int T, int R, bool FULL>
__global__ void __launch_bounds__(T * T) k_pfinv_rr(const double* in, double* out, int n, int count, double* gs_base,
                                                   double* pf, int* status) {
  extern __shared__ __align__(16) char lds[];
  Carver c{lds};
  RrLds L = carve_rr(c, n);
  double* gs = gs_base + (size_t)blockIdx.x * 2 * n * n;
  for (int p = blockIdx.x; p < count; p += gridDim.x) {
    const double* A = in + (size_t)p * n * n;
    double* O = out + (size_t)p * n * n;
    double pfv = 0.0;
    int info = 0;
    bool az = false;
    rr_pfinv<T, R, FULL>([&](int i, int j) { return A[i + (size_t)j * n]; }, n, O, 1.0, gs, L, pfv, info, false, az);
    if (threadIdx.x == 0) {
      status[p] = info;
      pf[p] = info > 0 ? 0.0 : pfv;
    }
    if (info != 0)
      for (int e = threadIdx.x; e < n * n; e += blockDim.x) O[e] = 0.0;
    __syncthreads();
  }
}

// Periodic rebuild (engine.rs:454-473) of the handed-over walkers with the
// register-resident Pfaffian + inverse, one walker per T x T workgroup, its
// NQP planes in turn, gathered straight from the Slater table
// (A(r, c) = -S[rs_c][rs_r], pfaffian.rs:1107-1188).  A plane that is all
// zero or fails stops the walker's rebuild (its Pfaffians are then not
// replaced; the accept window restarts either way).
template <int T, int R, bool FULL>
__device__ __forceinline__ int rr_calc_m_all(const EnsView& v, const int* ei, int* rs, double* planes, double* gs,
                                             RrLds& L, double* pfn) {
  const int tid = threadIdx.x;
  __shared__ int bad_site;
  if (tid == 0) bad_site = 0;
  __syncthreads();
  for (int i = tid; i < v.n; i += blockDim.x) {
    const int r = ei[i];
    rs[i] = r + (i / v.ne) * v.ns;
    if (r < 0 || r >= v.ns) bad_site = 1;
  }
  __syncthreads();
  if (bad_site) {
    if (tid == 0) set_status(v.status, ST_BAD_INPUT);
    __syncthreads();
    return 1;
  }
  const int n = v.n, n2 = v.n2;
  for (int q = 0; q < v.nqp; ++q) {
    const double* S = v.slater + (size_t)q * n2 * n2;
    double pfv = 0.0;
    int info = 0;
    bool az = false;
    rr_pfinv<T, R, FULL>([&](int i, int j) { return -S[(size_t)rs[j] * n2 + rs[i]]; }, n,
                         planes + (size_t)q * n * n, -1.0, gs, L, pfv, info, true, az);
    if (az || info != 0) return 1;
    if (tid == 0) pfn[q] = pfv;
  }
  __syncthreads();
  return 0;
}

template <int T, int R, bool FULL>
__global__ void __launch_bounds__(T * T) k_rebuild_rr(EnsView v) {
  extern __shared__ __align__(16) char lds[];
  Carver c{lds};
  int* ei = c.take<int>(v.n);
  int* rs = c.take<int>(v.n);
  double* pfn = c.take<double>(v.nqp);
  RrLds L = carve_rr(c, v.n);
  const int tid = threadIdx.x;
  double* gs = v.work + (size_t)blockIdx.x * 2 * v.n * v.n;
  const int nhand = *v.nflag;
  __shared__ int next_i;
  for (;;) {
    if (tid == 0) next_i = atomicAdd(v.nflag + 2, 1);
    __syncthreads();
    const int li = next_i;
    __syncthreads();
    if (li >= nhand) break;
    const int w = v.rlist[li];
    int* ws = v.wstate + (size_t)w * 8;
    for (int i = tid; i < v.n; i += blockDim.x) ei[i] = v.ele_idx[(size_t)w * v.n + i];
    __syncthreads();
    double* planes = v.inv + (size_t)w * v.nqp * v.n * v.n;
    const int bad = rr_calc_m_all<T, R, FULL>(v, ei, rs, planes, gs, L, pfn);
    copy_inv32(v, w);
    if (tid == 0) {
      if (!bad) {
        for (int q = 0; q < v.nqp; ++q) v.pf[(size_t)w * v.nqp + q] = pfn[q];
        v.log_ip[w] = ddm::cr_log(fabs(ip_of(v, pfn)));
      }
      ws[1] = 0;
      ws[5] = 0;
    }
    __syncthreads();
  }
}

// Start of a sampling call (k_begin) for the walkers that resume from their
// burn configuration, the case of every call after the first: its rebuild
// through the register-resident Pfaffian + inverse, walkers blockIdx.x,
// + gridDim.x, ...  That phase draws no random numbers and keeps the
// configuration, so a walker is done here when every plane factorises (no
// AllZero plane, info 0: the planes and Pfaffians then equal k_begin's bit
// for bit) and log|<x|psi>| is finite.  Any other walker is marked
// (wstate[7] = 1, counted in nflag[5]) and left to k_begin, which repeats
// the same rebuild and goes on from there.
template <int T, int R, bool FULL>
__global__ void __launch_bounds__(T * T) k_begin_rr(EnsView v) {
  extern __shared__ __align__(16) char lds[];
  Carver c{lds};
  int* ei = c.take<int>(v.n);
  int* rs = c.take<int>(v.n);
  double* pfn = c.take<double>(v.nqp);
  RrLds L = carve_rr(c, v.n);
  const int tid = threadIdx.x;
  double* gs = v.work + (size_t)blockIdx.x * 2 * v.n * v.n;
  __shared__ double lip;
  for (int w = blockIdx.x; w < v.W; w += gridDim.x) {
    int* ws = v.wstate + (size_t)w * 8;
    bool done = false;
    if (ws[0]) {
      for (int i = tid; i < v.n; i += blockDim.x) ei[i] = v.ele_idx[(size_t)w * v.n + i];
      __syncthreads();
      double* planes = v.inv + (size_t)w * v.nqp * v.n * v.n;
      const int bad = rr_calc_m_all<T, R, FULL>(v, ei, rs, planes, gs, L, pfn);
      if (tid == 0) lip = bad ? 0.0 : ddm::cr_log(fabs(ip_of(v, pfn)));
      __syncthreads();
      done = !bad && lip - lip == 0.0;
    }
    if (done) {
      copy_inv32(v, w);
      if (tid == 0) {
        for (int q = 0; q < v.nqp; ++q) v.pf[(size_t)w * v.nqp + q] = pfn[q];
        v.log_ip[w] = lip;
        ws[1] = 0;
        ws[2] = v.nsample + 1;
        ws[7] = 0;
        for (int k = 0; k < 4; ++k) v.counter[(size_t)w * 4 + k] = 0;
      }
    } else if (tid == 0) {
      ws[7] = 1;
      atomicAdd(v.nflag + 5, 1);
    }
    __syncthreads();
  }
}

// Measurement rebuild of saved sample s (as k_meas_rebuild) with the
// register-resident Pfaffian + inverse, walkers blockIdx.x, + gridDim.x, ...
// (T x T workgroups)
template <int T, int R, bool FULL>
__global__ void __launch_bounds__(T * T) k_meas_rebuild_rr(EnsView v, int s) {
  extern __shared__ __align__(16) char lds[];
  Carver c{lds};
  int* ei = c.take<int>(v.n);
  int* rs = c.take<int>(v.n);
  double* pfn = c.take<double>(v.nqp);
  RrLds L = carve_rr(c, v.n);
  const int tid = threadIdx.x;
  double* gs = v.work + (size_t)blockIdx.x * 2 * v.n * v.n;
  __shared__ int flag;
  for (int w = blockIdx.x; w < v.W; w += gridDim.x) {
    const size_t ws = (size_t)w * v.nsample + s;
    for (int i = tid; i < v.n; i += blockDim.x) ei[i] = v.s_idx[ws * v.n + i];
    __syncthreads();
    if (tid == 0) {
      int all0 = 1, allneg = 1;
      for (int k = 0; k < v.n; ++k) { all0 &= (ei[k] == 0); allneg &= (ei[k] < 0); }
      flag = (all0 || allneg) ? 0 : 1;
    }
    __syncthreads();
    int valid = flag;
    if (valid) {
      double* planes = v.inv + (size_t)w * v.nqp * v.n * v.n;
      valid = !rr_calc_m_all<T, R, FULL>(v, ei, rs, planes, gs, L, pfn);
    }
    if (tid == 0) {
      v.s_valid[ws] = valid;
      if (valid)
        for (int q = 0; q < v.nqp; ++q) v.pf[(size_t)w * v.nqp + q] = pfn[q];
    }
    __syncthreads();
  }
}

// Batched Pfaffian + inverse: plane group g = np consecutive planes; the
// workgroup at blockIdx.x takes groups blockIdx.x, + gridDim.x, ... of the
// ngroups, using work slot blockIdx.x (pfinv_work_doubles(n, np) doubles)
// for each in turn.
__global__ void __launch_bounds__(256) k_pfinv_batched(double* planes, double* work, int n, int count, int np,
                                                       int ngroups, double* pf, int* status) {
  extern __shared__ __align__(16) char lds[];
  double* wk = work + (size_t)blockIdx.x * pfinv_work_doubles(n, np);
  for (int grp = blockIdx.x; grp < ngroups; grp += gridDim.x) {
    const int first = grp * np;
    const int cnt = min(np, count - first);
    if (cnt <= 0) break;
    Carver c{lds};
    PfinvShared sh;
    carve_pfinv(c, sh, n, np);
    double* pl = planes + (size_t)first * n * n;
    wg_pfinv(pl, wk, n, cnt, sh, pf + first, 1.0);
    for (int p = threadIdx.x; p < cnt; p += blockDim.x) {
      const int info = sh.info[p];
      status[first + p] = info;
      if (info > 0) {
        pf[first + p] = 0.0;
      } else if (info < 0) {
        // non-finite Pfaffian: reported with its value (crates/mvmc-gpu/src/lib.rs:14-15),
        // recomputed as wg_pfinv forms it from the LTL form the plane still holds
        const double* a = pl + (size_t)p * n * n;
        double v = 1.0;
        for (int i = 0; i + 1 < n; i += 2) v *= a[i + (size_t)(i + 1) * n];
        bool neg = false;
        for (int k = 0; k < n; ++k)
          if (sh.P(p, k) != k + 1) neg = !neg;
        const double s = neg ? (0.0 - 1.0) : 1.0;
        pf[first + p] = s * v;
      }
    }
    __syncthreads();
    for (int p = 0; p < cnt; ++p)
      if (sh.info[p] != 0)
        for (int e = threadIdx.x; e < n * n; e += blockDim.x) pl[(size_t)p * n * n + e] = 0.0;
    __syncthreads();  // the next group re-carves the same LDS and reuses the work slot
  }
}

}  // namespace mvmc_hip
