// Workgroup-cooperative skew-symmetric Pfaffian + inverse of `np` planes at
// once (the QP planes of one walker, or a batch of planes).
//
// Each plane is column-major n x n (element (r,c) at r + c*n) in global
// memory; the workgroup walks the LTL^T elimination in lock step over all
// planes, so every synchronisation covers np planes of work.  Every matrix
// element receives exactly the operation sequence of the C PfaPack order
// (crates/pfapack/src/ltl.rs:61-191 and :683-705: DSKR2 (c + x*t1) - y*t2;
// crates/pfapack/src/utu2.rs:52-93 Pfaffian; :338-513 inverse with the scalar
// trtri/trmm of backend.rs:233-300), so results are bit-identical to a
// sequential CPU evaluation of that order; only independent elements run in
// parallel.  Requires -ffp-contract=off.
#pragma once
#include "common.h"

namespace mvmc_hip {

// Steps per panel of the delayed trailing update of the LTL^T elimination.
constexpr int LTL_K = 8;
// Columns per block of the row-owner passes of the inverse.
constexpr int TJ = 8;
// Most waves per workgroup wg_pfinv is called with (blockDim.x <= 256).
constexpr int PF_WAVES = 4;
// Columns per batch of the panel close: PF_JB by default, at most PF_JB_MAX
// (a template parameter of wg_pfinv; the rebuild kernels pick it by
// measurement).  The per-wave LDS broadcast buffer holds a batch's t1, t2
// (and a 64-value block in the inverse).
constexpr int PF_JB = 8, PF_JB_MAX = 16;
constexpr int PF_WB = 2 * PF_JB_MAX * LTL_K;

// Orders the LDS accesses of one wave around a lane-to-lane exchange (LDS
// operations of a wave execute in issue order; this keeps the compiler from
// moving them across).
__device__ inline void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Doubles of global scratch wg_pfinv needs for np planes of order n.
__host__ __device__ inline size_t pfinv_work_doubles(int n, int np) {
  return (size_t)np * n * n;
}

// LDS scratch of wg_pfinv, carved from dynamic shared memory by the caller:
// piv[np*n], colbuf[np*n], kp/info[np], colmax[np], ks[np*n] (panel step of
// the last swap through each index), act[np*LTL_K] (step applied per plane),
// zmax[2*np] (extent of the zeroed diagonal; whether the last x column
// awaits scaling), wb[PF_WAVES*PF_WB] (one broadcast buffer per wave).
struct PfinvShared {
  int* piv;
  double* colbuf;
  int* kp;
  int* info;
  double* colmax;
  int* ks;
  int* act;
  int* zmax;
  double* wb;
  __device__ int& P(int p, int i) { return piv[p * ld + i]; }
  __device__ double& C(int p, int i) { return colbuf[p * ld + i]; }
  __device__ int& KS(int p, int i) { return ks[p * ld + i]; }
  int ld;
};

__host__ __device__ inline size_t pf_rnd16(size_t b) { return (b + 15) & ~(size_t)15; }

// Bytes of LDS that carve_pfinv takes.
__host__ __device__ inline size_t pf_bytes(int n, int np) {
  return 2 * pf_rnd16((size_t)np * n * sizeof(int)) + pf_rnd16((size_t)np * n * sizeof(double)) +
         3 * pf_rnd16((size_t)np * 8) + pf_rnd16((size_t)np * LTL_K * sizeof(int)) + pf_rnd16(2 * (size_t)np * sizeof(int)) +
         pf_rnd16((size_t)PF_WAVES * PF_WB * sizeof(double));
}

template <class Carver>
__device__ inline void carve_pfinv(Carver& c, PfinvShared& sh, int n, int np) {
  sh.ld = n;
  sh.piv = c.template take<int>((size_t)np * n);
  sh.colbuf = c.template take<double>((size_t)np * n);
  sh.kp = c.template take<int>(np);
  sh.info = c.template take<int>(np);
  sh.colmax = c.template take<double>(np);
  sh.ks = c.template take<int>((size_t)np * n);
  sh.act = c.template take<int>((size_t)np * LTL_K);
  sh.zmax = c.template take<int>(2 * (size_t)np);
  sh.wb = c.template take<double>((size_t)PF_WAVES * PF_WB);
}

// Argmax of |x| over a column with C IDAMAX semantics: first index of the
// maximum; a NaN only wins at index 0 (the sequential scan starts from it).
__device__ inline void piv_key(double v, int j, double& kv) {
  if (v != v) kv = (j == 0) ? INFINITY : -1.0;
  else kv = v;
}

// Assemble -S[rs_i][rs_j] into the planes (pfaffian.rs:1107-1188).
// Returns (via *allzero) whether any plane has max|a|^2 < 1e-28.
__device__ inline void assemble_planes(double* planes, const double* slater, const int* rs, int n, int n2,
                                       int np, int q0) {
  const int nn = n * n;
  for (int e = threadIdx.x; e < np * nn; e += blockDim.x) {
    const int p = e / nn, r = e % nn;
    const int msi = r / n, msj = r % n;
    const double* S = slater + (size_t)(q0 + p) * n2 * n2;
    planes[(size_t)p * nn + msj + (size_t)msi * n] = -S[(size_t)rs[msi] * n2 + rs[msj]];
  }
}

// Core routine.  planes: np planes (stride n*n).  work: pfinv_work_doubles
// of scratch (panel vectors of the elimination, then M^T of every plane,
// then the permutation gather).  On exit, for every plane with info == 0,
// pf[p] holds the Pfaffian and the plane holds sign * A^{-1} (sign = -1
// gives mVMC's invM).  info[p]: 0 ok, k > 0 zero pivot at 1-based row k, -1
// non-finite Pfaffian.  All threads of the block (at most PF_WAVES waves)
// must call it.
//
// Every phase is written so that a lane issues the loads of all its items
// (PF_U per batch) before its first store: the work is bound by memory
// round trips, and a load after a store to the same buffer could not be
// hoisted by the compiler.
constexpr int PF_U = 8;
// Opaque copies of a wave-uniform value / pointer: the compiler cannot prove
// them equal to the original, so nothing derived from them is hoisted above
// this point (used at the entry of routines called in a loop, whose
// loop-invariant address arithmetic would otherwise stay live across the
// whole loop and cost occupancy).
__device__ inline int opaque_uniform(int x) {
  asm volatile("" : "+v"(x));
  return __builtin_amdgcn_readfirstlane(x);
}
template <class T>
__device__ inline T* opaque_uniform(T* p) {
  const unsigned long long u = (unsigned long long)p;
  const unsigned lo = (unsigned)opaque_uniform((int)(unsigned)u);
  const unsigned hi = (unsigned)opaque_uniform((int)(unsigned)(u >> 32));
  return (T*)(((unsigned long long)hi << 32) | lo);
}

template <int JB = PF_JB>
__device__ __forceinline__ void wg_pfinv(double* planes, double* work, int n, int np, PfinvShared& sh_in, double* pf, double sign,
                         long long* clk = nullptr) {
  int tid = threadIdx.x;
  asm volatile("" : "+v"(tid));
  int nthr = opaque_uniform((int)blockDim.x);
  n = opaque_uniform(n);
  np = opaque_uniform(np);
  planes = opaque_uniform(planes);
  work = opaque_uniform(work);
  pf = opaque_uniform(pf);
  PfinvShared sh = sh_in;
  sh.piv = opaque_uniform(sh.piv);
  sh.colbuf = opaque_uniform(sh.colbuf);
  sh.kp = opaque_uniform(sh.kp);
  sh.info = opaque_uniform(sh.info);
  sh.colmax = opaque_uniform(sh.colmax);
  sh.ks = opaque_uniform(sh.ks);
  sh.act = opaque_uniform(sh.act);
  sh.zmax = opaque_uniform(sh.zmax);
  sh.wb = opaque_uniform(sh.wb);
  sh.ld = opaque_uniform(sh.ld);
  // optional phase clocks (lane 0, after the barrier that ends a phase)
  long long ck0 = clk ? wall_clock64() : 0;
  auto PFC = [&](int k) {
    if (clk && tid == 0) {
      const long long t_ = wall_clock64();
      clk[k] += t_ - ck0;
      ck0 = t_;
    }
  };
  int lane = tid & 63, wave = tid >> 6, nwave = (nthr + 63) >> 6;
  int nn = n * n;
  // Fresh opaque copies at the start of each phase: expressions the phases
  // share (item decompositions, addresses) are then computed per phase
  // instead of once and kept live across the others.
#define PF_FRESH()                                 \
  do {                                             \
    asm volatile("" : "+v"(tid));                  \
    nthr = opaque_uniform(nthr);                   \
    n = opaque_uniform(n);                         \
    np = opaque_uniform(np);                       \
    planes = opaque_uniform(planes);               \
    work = opaque_uniform(work);                   \
    lane = tid & 63;                               \
    wave = tid >> 6;                               \
    nwave = (nthr + 63) >> 6;                      \
    nn = n * n;                                    \
  } while (0)
  for (int e = tid; e < np * n; e += nthr) sh.P(e / n, e % n) = e % n + 1;
  for (int p = tid; p < np; p += nthr) sh.info[p] = 0;
  __syncthreads();

  // ---------------- LTL^T (dsktf2) ----------------
  // Panels of `kw` steps.  The panel's own future y columns [m, kk0) take
  // every step's rank-2 update as it happens (eager); the trailing block
  // (columns j < m) takes them at the panel's end (delayed).  Every element
  // still receives its updates in step order, each as (c + x*t1) - y*t2 with
  // the x, y, t1, t2 of that step, at the position it held at that step:
  //  - before a swap, the trailing elements it moves (column kp, row kp) are
  //    brought up to date, so nothing pending is moved;
  //  - a trailing element at (i, j) has received the panel steps before
  //    max(ks[i], ks[j]) (ks[k]: the last panel step whose swap wrote row /
  //    column k); the others stay pending with the step's x, y rows i, j,
  //    which later swaps do not touch;
  //  - the scaling of the finished x column by alpha is applied in the next
  //    step, before its swap (a swap only exchanges rows that are both
  //    scaled, so the order does not matter).
  // Panel vectors (x, y, t1 = alpha*y, t2 = alpha*x per step) live in `work`.
  const int kw = n >= 4 * LTL_K ? LTL_K : (n >= 4 ? n / 4 : 1);
  auto PX = [&](int p, int t) { return work + ((size_t)(p * 4 + 0) * kw + t) * n; };
  auto PY = [&](int p, int t) { return work + ((size_t)(p * 4 + 1) * kw + t) * n; };
  auto PT1 = [&](int p, int t) { return work + ((size_t)(p * 4 + 2) * kw + t) * n; };
  auto PT2 = [&](int p, int t) { return work + ((size_t)(p * 4 + 3) * kw + t) * n; };
  // pending steps [s0, t_end) of plane p on c = element (i, j), i < j
  auto pending = [&](int p, int i, int j, int s0, int t_end, double c) {
    constexpr int CU_G = 4;  // pending steps whose loads are issued together
    for (int g = s0; g < t_end; g += CU_G) {
      double xv[CU_G], yv[CU_G], t1v[CU_G], t2v[CU_G];
#pragma unroll
      for (int u = 0; u < CU_G; ++u)
        if (g + u < t_end) {
          xv[u] = PX(p, g + u)[i];
          yv[u] = PY(p, g + u)[i];
          t1v[u] = PT1(p, g + u)[j];
          t2v[u] = PT2(p, g + u)[j];
        }
#pragma unroll
      for (int u = 0; u < CU_G; ++u)
        if (g + u < t_end && sh.act[p * LTL_K + g + u]) c = (c + xv[u] * t1v[u]) - yv[u] * t2v[u];
    }
    return c;
  };
  double* alph = sh.colbuf;  // alpha of the step whose x column awaits scaling, per plane (colbuf[p*n])
  for (int e = tid; e < np * n; e += nthr) sh.ks[e] = 0;
  for (int p = tid; p < 2 * np; p += nthr) sh.zmax[p] = 0;
  __syncthreads();
  int tp = 0;          // steps recorded in the open panel
  int m_pan = 0;       // first eager column of the open panel
  int scale_col = -1;  // x column of the previous step (awaiting scaling)
  auto scale_on = [&](int p) { return sh.zmax[np + p] != 0; };
  for (int k0 = n - 1; k0 >= 1; --k0) {
    const int kk0 = k0 - 1;
    PF_FRESH();
    // (A) pivot search on the (up to date) column k0: one wave per plane
    for (int p = wave; p < np; p += nwave) {
      const double* col = planes + (size_t)p * nn + (size_t)k0 * n;
      double bv = -2.0;
      int bj = 0x7fffffff;
      for (int j0 = 0; j0 <= kk0; j0 += 64 * 4) {
        double v4[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) {
          const int j = j0 + lane + 64 * u;
          if (j <= kk0) v4[u] = col[j];
        }
#pragma unroll
        for (int u = 0; u < 4; ++u) {
          const int j = j0 + lane + 64 * u;
          if (j <= kk0) {
            double kv;
            piv_key(fabs(v4[u]), j, kv);
            if (kv > bv) { bv = kv; bj = j; }
          }
        }
      }
      for (int off = 32; off >= 1; off >>= 1) {
        double ov = __shfl_xor(bv, off, 64);
        int oj = __shfl_xor(bj, off, 64);
        if (ov > bv || (ov == bv && oj < bj)) { bv = ov; bj = oj; }
      }
      if (lane == 0) {
        sh.kp[p] = bj;
        sh.colmax[p] = bv;  // -1 only for an all-NaN tail; 0 means a zero column
      }
    }
    __syncthreads();
    PFC(0);
    if (tp == 0) m_pan = max(0, kk0 - kw + 1);
    const int m = m_pan;  // trailing columns j < m; eager columns [m, kk0)
    PF_FRESH();
    // (B1) scale the previous x column (rows < k0)
    if (scale_col >= 0) {
      for (int b0 = 0; b0 < np * k0; b0 += nthr * PF_U) {
        double vv[PF_U];
#pragma unroll
        for (int u = 0; u < PF_U; ++u) {
          const int it = b0 + tid + nthr * u;
          if (it < np * k0 && scale_on(it / k0)) vv[u] = planes[(size_t)(it / k0) * nn + it % k0 + (size_t)scale_col * n];
        }
#pragma unroll
        for (int u = 0; u < PF_U; ++u) {
          const int it = b0 + tid + nthr * u;
          if (it < np * k0 && scale_on(it / k0))
            planes[(size_t)(it / k0) * nn + it % k0 + (size_t)scale_col * n] = vv[u] * alph[(it / k0) * n];
        }
      }
    }
    PF_FRESH();
    // (B2) bring the trailing elements a swap moves up to date: items
    // (plane, f), f < m - 1: f < kp: (f, kp), else (kp, f + 1)  [kp < m]
    if (tp > 0 && m >= 2) {
      const int per = m - 1;
      for (int it = tid; it < np * per; it += nthr) {
        const int p = it / per, f = it % per;
        const int kp = sh.kp[p];
        if (sh.colmax[p] != 0.0 && kp != kk0 && kp < m) {
          const int i = f < kp ? f : kp, j = f < kp ? kp : f + 1;
          const int s0 = max(sh.KS(p, i), sh.KS(p, j));
          if (s0 < tp) {
            double* e = planes + (size_t)p * nn + i + (size_t)j * n;
            *e = pending(p, i, j, s0, tp, *e);
          }
        }
      }
    }
    __syncthreads();
    PFC(1);
    PF_FRESH();
    // (C) swaps and sign flips (ltl.rs:131-158), items (plane, e), e < n - 1:
    //   e < kp: a[e][kk0] <-> a[e][kp];  e == kp: a[kp][kk0] = -a[kp][kk0];
    //   kp < e < kk0: a[e][kk0] <-> a[kp][e], both negated;
    //   e >= kk0: rows kk0 <-> kp of column e + 1 (finished columns)
    {
      // the item count is made opaque so the item decompositions are not
      // hoisted out of the step loop (they would stay live across phases)
      int nit = np * (n - 1);
      asm volatile("" : "+v"(nit));
      nit = __builtin_amdgcn_readfirstlane(nit);
      for (int b0 = 0; b0 < nit; b0 += nthr * 4) {
        double v0[4], v1[4];
        int o0[4], o1[4], kind[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) {
          const int it = b0 + tid + nthr * u;
          kind[u] = 0;
          if (it < nit) {
            const int p = it / (n - 1), e = it % (n - 1);
            const int kp = sh.kp[p];
            if (sh.colmax[p] != 0.0 && kp != kk0) {
              const int base = p * nn;
              if (e < kp) { kind[u] = 1; o0[u] = base + e + kk0 * n; o1[u] = base + e + kp * n; }
              else if (e == kp) { kind[u] = 2; o0[u] = base + kp + kk0 * n; o1[u] = o0[u]; }
              else if (e < kk0) { kind[u] = 3; o0[u] = base + e + kk0 * n; o1[u] = base + kp + e * n; }
              else { kind[u] = 1; o0[u] = base + kk0 + (e + 1) * n; o1[u] = base + kp + (e + 1) * n; }
              v0[u] = planes[o0[u]];
              v1[u] = planes[o1[u]];
            }
          }
        }
#pragma unroll
        for (int u = 0; u < 4; ++u) {
          if (kind[u] == 1) { planes[o0[u]] = v1[u]; planes[o1[u]] = v0[u]; }
          else if (kind[u] == 2) planes[o0[u]] = -v0[u];
          else if (kind[u] == 3) { planes[o0[u]] = -v1[u]; planes[o1[u]] = -v0[u]; }
        }
      }
    }
    PF_FRESH();
    if (tid < np) {
      const int p = tid;
      if (sh.colmax[p] == 0.0) {
        if (sh.info[p] == 0) sh.info[p] = kk0 + 1;
        sh.P(p, kk0) = kk0 + 1;
      } else {
        sh.P(p, kk0) = sh.kp[p] + 1;
        if (sh.kp[p] != kk0) sh.KS(p, sh.kp[p]) = tp;
      }
      const int on = (sh.colmax[p] != 0.0 && kk0 >= 1);
      sh.act[p * LTL_K + tp] = on;
      if (on && kk0 > sh.zmax[p]) sh.zmax[p] = kk0;
    }
    __syncthreads();
    PFC(2);
    PF_FRESH();
    // (D) record the step (x, y unscaled, t1 = alpha*y[j], t2 = alpha*x[j])
    // and apply it to the eager columns [m, kk0): a lane owns row i of plane
    // p, loads x_i, y_i and its eager elements together; x_j, y_j are
    // wave-shared loads.  t1 = alpha*y_j, t2 = alpha*x_j are formed by the
    // same expressions as the recorded ones.
    if (kk0 >= 1) {
      for (int b0 = 0; b0 < np * kk0; b0 += nthr) {
        const int it = b0 + tid;
        if (it < np * kk0) {
          const int p = it / kk0, i = it % kk0;
          if (sh.colmax[p] != 0.0) {
            double* a = planes + (size_t)p * nn;
            const double xi = a[i + (size_t)k0 * n], yi = a[i + (size_t)kk0 * n];
            const double den = a[kk0 + (size_t)k0 * n];
            const double alpha = 1.0 / den;
            PX(p, tp)[i] = xi;
            PY(p, tp)[i] = yi;
            PT1(p, tp)[i] = alpha * yi;
            PT2(p, tp)[i] = alpha * xi;
            if (i == 0) alph[p * n] = alpha;
            // eager columns in chunks of EC (their loads issued together)
            constexpr int EC = 4;
            for (int c0 = m; c0 < kk0; c0 += EC) {
              double cv[EC], xj[EC], yj[EC];
#pragma unroll
              for (int c = 0; c < EC; ++c) {
                const int j = c0 + c;
                if (j < kk0) {
                  xj[c] = a[j + (size_t)k0 * n];
                  yj[c] = a[j + (size_t)kk0 * n];
                  if (i < j) cv[c] = a[i + (size_t)j * n];
                }
              }
#pragma unroll
              for (int c = 0; c < EC; ++c) {
                const int j = c0 + c;
                if (j < kk0 && i < j) a[i + (size_t)j * n] = (cv[c] + xi * (alpha * yj[c])) - yi * (alpha * xj[c]);
              }
            }
          }
        }
      }
      __syncthreads();
      PFC(3);
      scale_col = k0;
    } else {
      scale_col = -1;
    }
    for (int p = tid; p < np; p += nthr) sh.zmax[np + p] = sh.colmax[p] != 0.0 && kk0 >= 1;  // scale_on
    ++tp;
    PF_FRESH();
    // (E) close the panel: every element of the trailing triangle (i < j < kk0)
    // takes its pending steps.
    if (tp == kw || kk0 <= 1) {
      __syncthreads();
      const int m = kk0;  // trailing columns j < m, rows i < m - 1
      if (m >= 2) {
        // work item = (plane, 64-row chunk) per wave; chunk index varies
        // slowest so every wave gets a mix of long and short chunks.  A
        // lane owns row i; JB columns at a time: their elements are loaded
        // together, their t1, t2 of every step are one wave load each,
        // broadcast through the wave's LDS buffer.
        static_assert(JB <= PF_JB_MAX, "the broadcast buffer holds PF_JB_MAX columns");
        constexpr int NH = JB * LTL_K / 64;       // wave loads per step-scalar array
        static_assert(NH * 64 == JB * LTL_K, "a batch's step scalars fill whole wave loads");
        double* wb = sh.wb + wave * PF_WB;
        const int chunks = (m - 1 + 63) / 64;
        for (int it = __builtin_amdgcn_readfirstlane(wave); it < np * chunks; it += nwave) {
          const int p = it % np, c = it / np;
          const int i = c * 64 + lane;
          const bool on = i < m - 1;
          double xr[LTL_K], yr[LTL_K];
          int ar[LTL_K];
#pragma unroll
          for (int t = 0; t < LTL_K; ++t) {
            const bool u = on && t < tp && sh.act[p * LTL_K + t];
            ar[t] = u;
            xr[t] = u ? PX(p, t)[i] : 0.0;
            yr[t] = u ? PY(p, t)[i] : 0.0;
          }
          const int ksi = on ? sh.KS(p, i) : tp;
          double* a = planes + (size_t)p * nn;
          const int tl = lane % LTL_K, jl = lane / LTL_K;
          for (int j0 = c * 64 + 1; j0 < m; j0 += JB) {
            double v1[NH], v2[NH];
#pragma unroll
            for (int h = 0; h < NH; ++h) {
              const int jc = jl + h * (64 / LTL_K);
              const bool lv = tl < tp && j0 + jc < m;
              v1[h] = lv ? PT1(p, tl)[j0 + jc] : 0.0;
              v2[h] = lv ? PT2(p, tl)[j0 + jc] : 0.0;
            }
            double cc[JB];
#pragma unroll
            for (int jj = 0; jj < JB; ++jj) {
              const int j = j0 + jj;
              cc[jj] = (on && j > i && j < m) ? (a + (size_t)j * n)[i] : 0.0;
            }
            wave_sync();
#pragma unroll
            for (int h = 0; h < NH; ++h) {
              wb[lane + 64 * h] = v1[h];                // wb[t + LTL_K*jj] = t1 of step t, column j0+jj
              wb[JB * LTL_K + lane + 64 * h] = v2[h];   // t2
            }
            wave_sync();
            double* ao = opaque_uniform(a);  // store addresses not shared with the loads above
#pragma unroll
            for (int jj = 0; jj < JB; ++jj) {
              const int j = j0 + jj;
              if (on && j > i && j < m) {
                const int s0 = max(ksi, sh.KS(p, j));
                if (s0 < tp) {
                  double x = cc[jj];
#pragma unroll
                  for (int t = 0; t < LTL_K; ++t)
                    if (t >= s0 && ar[t])
                      x = (x + xr[t] * wb[t + LTL_K * jj]) - yr[t] * wb[JB * LTL_K + t + LTL_K * jj];
                  (ao + (size_t)j * n)[i] = x;
                }
              }
            }
          }
          wave_sync();
        }
      }
      __syncthreads();
      for (int e = tid; e < np * n; e += nthr) sh.ks[e] = 0;
      tp = 0;
      __syncthreads();
      PFC(4);
    }
  }
  PF_FRESH();
  // the unblocked elimination zeroes the diagonal of every block it updated
  for (int e = tid; e < np * n; e += nthr) {
    const int p = e / n, j = e % n;
    if (j < sh.zmax[p]) planes[(size_t)p * nn + j + (size_t)j * n] = 0.0;
  }
  __syncthreads();
  PF_FRESH();
  // ---------------- Pfaffian (utu2pfa) ----------------
  if (tid < np) {
    const int p = tid;
    if (sh.info[p] == 0) {
      const double* a = planes + (size_t)p * nn;
      double v = 1.0;
      for (int i = 0; i + 1 < n; i += 2) v *= a[i + (size_t)(i + 1) * n];
      bool neg = false;
      for (int k = 0; k < n; ++k)
        if (sh.P(p, k) != k + 1) neg = !neg;
      const double s = neg ? (0.0 - 1.0) : 1.0;
      v = s * v;
      if (!(v - v == 0.0)) sh.info[p] = -1;
      else pf[p] = v;
    }
  }
  __syncthreads();
  PFC(9);
  PF_FRESH();
  // ---------------- inverse (utu2inv) ----------------
  // (1) trtri('U','U') on B(i,j) = A(i,j+1), 0 <= i < j < n-1, in place:
  //   Bn(i,j) = -(B(i,j) + sum_{k=i+1}^{j-1} Bn(i,k) B(k,j)), k ascending.
  // Rows are independent chains.  One wave per plane (no workgroup
  // barriers); a lane owns rows g0 + lane + 64 r (r < TR) of a block of TJ
  // columns.  The B values a row reads from other rows (B(k,j), k > i) are
  // all read before the wave writes the block, and row groups go in
  // ascending order, so no in-place write precedes a read of the value it
  // replaces.  The B block of each batch of KB steps is one coalesced wave
  // load, broadcast to the lanes through the wave's LDS buffer.
  {
    constexpr int TR = 4, KB = 4;  // KB * TJ <= 64
    double* wb = sh.wb + wave * PF_WB;
    for (int p = __builtin_amdgcn_readfirstlane(wave); p < np; p += nwave) {
      if (sh.info[p] != 0) continue;
      double* a = planes + (size_t)p * nn;
      for (int j0 = 1; j0 < n - 1; j0 += TJ) {
        const int j1 = min(n - 1, j0 + TJ);
        for (int g0 = 0; g0 < j1 - 1; g0 += 64 * TR) {
          int ir[TR];
          bool on[TR];
          double acc[TR][TJ];
#pragma unroll
          for (int r = 0; r < TR; ++r) {
            ir[r] = g0 + lane + 64 * r;
            on[r] = ir[r] < j1 - 1;
#pragma unroll
            for (int jj = 0; jj < TJ; ++jj)
              acc[r][jj] = (on[r] && j0 + jj < j1 && ir[r] < j0 + jj) ? (a + (size_t)(j0 + jj + 1) * n)[ir[r]] : 0.0;
          }
          // k < j0: finished columns of earlier blocks
          const int bk = lane % KB, bj = lane / KB;
          for (int k = g0 + 1; k < j0; k += KB) {
            const double bv = (bj < TJ && k + bk < j0 && j0 + bj < j1) ? (a + (size_t)(j0 + bj + 1) * n)[k + bk] : 0.0;
            double bn[TR][KB];
#pragma unroll
            for (int r = 0; r < TR; ++r)
#pragma unroll
              for (int u = 0; u < KB; ++u)
                bn[r][u] = (on[r] && k + u > ir[r] && k + u < j0) ? (a + (size_t)(k + u + 1) * n)[ir[r]] : 0.0;
            wave_sync();
            wb[lane] = bv;  // wb[u + KB*jj] = B(k+u, j0+jj)
            wave_sync();
#pragma unroll
            for (int u = 0; u < KB; ++u)
#pragma unroll
              for (int r = 0; r < TR; ++r)
                if (on[r] && k + u > ir[r] && k + u < j0) {
#pragma unroll
                  for (int jj = 0; jj < TJ; ++jj)
                    if (j0 + jj < j1) acc[r][jj] += bn[r][u] * wb[u + KB * jj];
                }
          }
          // k in [j0, j): the block's own columns, finished by each lane in order
          {
            const int kk = lane % TJ, jj = lane / TJ;  // wb[kk + TJ*jj] = B(j0+kk, j0+jj)
            const double bv = (j0 + jj < j1 && kk < jj) ? (a + (size_t)(j0 + jj + 1) * n)[j0 + kk] : 0.0;
            wave_sync();
            wb[lane] = bv;
            wave_sync();
          }
#pragma unroll
          for (int jj = 0; jj < TJ; ++jj) {
#pragma unroll
            for (int kk = 0; kk < jj; ++kk)
#pragma unroll
              for (int r = 0; r < TR; ++r)
                if (on[r] && j0 + jj < j1 && ir[r] < j0 + kk) acc[r][jj] += acc[r][kk] * wb[kk + TJ * jj];
#pragma unroll
            for (int r = 0; r < TR; ++r) acc[r][jj] = -acc[r][jj];
          }
          double* ao = opaque_uniform(a);  // store addresses not shared with the loads above
#pragma unroll
          for (int r = 0; r < TR; ++r)
#pragma unroll
            for (int jj = 0; jj < TJ; ++jj)
              if (on[r] && j0 + jj < j1 && ir[r] < j0 + jj) (ao + (size_t)(j0 + jj + 1) * n)[ir[r]] = acc[r][jj];
          wave_sync();
        }
      }
    }
  }
  __syncthreads();
  PFC(5);
  // (2) Mt = M^T of every plane into work (Mt(col,row) at col + row*n), with
  // M = I, M(ir, jr+1) = A(ir, jr+2) for ir <= jr < n-2; vt(i) = -A(i, i+1).
  PF_FRESH();
  for (int b0 = 0; b0 < np * nn; b0 += nthr * PF_U) {
    double mv[PF_U];
#pragma unroll
    for (int u = 0; u < PF_U; ++u) {
      const int e = b0 + tid + nthr * u;
      const int p = e / nn, r = e % nn;
      const int col = r % n, row = r / n;
      double v = (row == col) ? 1.0 : 0.0;
      if (e < np * nn && col >= 1 && col <= n - 2 && row <= col - 1 && sh.info[p] == 0)
        v = planes[(size_t)p * nn + row + (size_t)(col + 1) * n];
      mv[u] = v;
    }
#pragma unroll
    for (int u = 0; u < PF_U; ++u) {
      const int e = b0 + tid + nthr * u;
      if (e < np * nn) work[e] = mv[u];
    }
  }
  for (int e = tid; e < np * (n - 1); e += nthr) {
    const int p = e / (n - 1), i = e % (n - 1);
    sh.C(p, i) = -planes[(size_t)p * nn + i + (size_t)(i + 1) * n];
  }
  __syncthreads();
  // (3) skew-tridiagonal solve T C = M, C overwrites A; items (plane, column
  // j), the recurrence carries its last value in a register, right-hand
  // sides are loaded SU at a time ahead of it.
  PF_FRESH();
  for (int e = tid; e < np * n; e += nthr) {
    constexpr int SU = 8;
    const int p = e / n, j = e % n;
    if (sh.info[p] != 0) continue;
    const double* vt = &sh.C(p, 0);
    const double* mt = work + (size_t)p * nn;
    double* c = planes + (size_t)p * nn + (size_t)j * n;
    double prev = mt[j + 0 * (size_t)n] / -vt[0];
    c[1] = prev;
    for (int i0 = 2; i0 < n; i0 += 2 * SU) {
      double bb[SU];
#pragma unroll
      for (int u = 0; u < SU; ++u) bb[u] = (i0 + 2 * u < n) ? mt[j + (size_t)(i0 + 2 * u) * n] : 0.0;
#pragma unroll
      for (int u = 0; u < SU; ++u) {
        const int i = i0 + 2 * u;
        if (i < n) {
          prev = (bb[u] - prev * vt[i - 1]) / -vt[i];
          c[i + 1] = prev;
        }
      }
    }
    prev = mt[j + (size_t)(n - 1) * n] / vt[n - 2];
    c[n - 2] = prev;
    for (int i0 = n - 3; i0 >= 1; i0 -= 2 * SU) {
      double bb[SU];
#pragma unroll
      for (int u = 0; u < SU; ++u) bb[u] = (i0 - 2 * u >= 1) ? mt[j + (size_t)(i0 - 2 * u) * n] : 0.0;
#pragma unroll
      for (int u = 0; u < SU; ++u) {
        const int i = i0 - 2 * u;
        if (i >= 1) {
          prev = (bb[u] + prev * vt[i]) / vt[i - 1];
          c[i - 1] = prev;
        }
      }
    }
  }
  __syncthreads();
  PFC(6);
  // (4) T = M^T C in place: T(i,j) = C(i,j) + sum_{k<i} M(k,i) C(k,j), k
  // ascending.  For n > 64 the reference forms only the upper triangle
  // (i < j) and restores skew symmetry; for n <= 64 it forms every element.
  // One wave per plane, a lane owns rows g0 + lane + 64 r of a block of TJ
  // columns; row groups go in descending order (a row reads C of the rows
  // above it, which a group only overwrites after its own reads), and the
  // C block of each batch of KB steps is broadcast through the wave's LDS
  // buffer.
  PF_FRESH();
  {
    constexpr int TR = 4, KB = 4;  // KB * TJ <= 64
    double* wb = sh.wb + wave * PF_WB;
    const bool upper = n > 64;
    for (int p = __builtin_amdgcn_readfirstlane(wave); p < np; p += nwave) {
      if (sh.info[p] != 0) continue;
      double* a = planes + (size_t)p * nn;
      const double* mt = work + (size_t)p * nn;
      for (int j0 = 0; j0 < n; j0 += TJ) {
        const int j1 = min(n, j0 + TJ);
        const int rows = upper ? j1 - 1 : n;
        for (int gi = (rows + 64 * TR - 1) / (64 * TR) - 1; gi >= 0; --gi) {
          const int g0 = gi * 64 * TR;
          int ir[TR];
          bool on[TR];
          double acc[TR][TJ];
#pragma unroll
          for (int r = 0; r < TR; ++r) {
            ir[r] = g0 + lane + 64 * r;
            on[r] = ir[r] < rows;
#pragma unroll
            for (int jj = 0; jj < TJ; ++jj)
              acc[r][jj] = (on[r] && j0 + jj < j1) ? (a + (size_t)(j0 + jj) * n)[ir[r]] : 0.0;
          }
          const int kend = min(rows, g0 + 64 * TR) - 1;  // the group's largest row
          const int bk = lane % KB, bj = lane / KB;
          for (int k = 0; k < kend; k += KB) {
            const double bv = (bj < TJ && k + bk < kend && j0 + bj < j1) ? (a + (size_t)(j0 + bj) * n)[k + bk] : 0.0;
            double mv[TR][KB];
#pragma unroll
            for (int r = 0; r < TR; ++r)
#pragma unroll
              for (int u = 0; u < KB; ++u)
                mv[r][u] = (on[r] && k + u < ir[r]) ? (mt + (size_t)(k + u) * n)[ir[r]] : 0.0;
            wave_sync();
            wb[lane] = bv;  // wb[u + KB*jj] = C(k+u, j0+jj)
            wave_sync();
#pragma unroll
            for (int u = 0; u < KB; ++u)
#pragma unroll
              for (int r = 0; r < TR; ++r)
                if (on[r] && k + u < ir[r]) {
#pragma unroll
                  for (int jj = 0; jj < TJ; ++jj)
                    if (j0 + jj < j1) acc[r][jj] += mv[r][u] * wb[u + KB * jj];
                }
          }
          double* ao = opaque_uniform(a);  // store addresses not shared with the loads above
#pragma unroll
          for (int r = 0; r < TR; ++r)
#pragma unroll
            for (int jj = 0; jj < TJ; ++jj)
              if (on[r] && j0 + jj < j1 && (!upper || ir[r] < j0 + jj)) (ao + (size_t)(j0 + jj) * n)[ir[r]] = acc[r][jj];
          wave_sync();
        }
      }
    }
  }
  __syncthreads();
  if (n > 64) {
    // skew fill: lower triangle = -upper, zero diagonal
    PF_FRESH();
    for (int b0 = 0; b0 < np * nn; b0 += nthr * PF_U) {
      double vv[PF_U];
#pragma unroll
      for (int u = 0; u < PF_U; ++u) {
        const int e = b0 + tid + nthr * u;
        const int p = e / nn, r = e % nn;
        const int row = r % n, col = r / n;
        if (e < np * nn && row > col && sh.info[p] == 0) vv[u] = planes[(size_t)p * nn + col + (size_t)row * n];
      }
#pragma unroll
      for (int u = 0; u < PF_U; ++u) {
        const int e = b0 + tid + nthr * u;
        const int p = e / nn, r = e % nn;
        const int row = r % n, col = r / n;
        if (e < np * nn && sh.info[p] == 0) {
          if (row == col) planes[e] = 0.0;
          else if (row > col) planes[e] = -vv[u];
        }
      }
    }
    __syncthreads();
  }
  PFC(7);
  PF_FRESH();
  // column then row permutations: the sequential swaps (j <-> P[j]-1,
  // j ascending) compose into one index map g, applied to rows and columns
  // as a gather (pure data movement), then the sign (calc_m_all's final
  // DSCAL by -1).  g lives in ks, the gather goes through `work`.
  if (tid < np) {
    const int p = tid;
    int* g = &sh.KS(p, 0);
    for (int k = 0; k < n; ++k) g[k] = k;
    for (int j = 0; j < n; ++j) {
      const int t = sh.P(p, j) - 1;
      if (t != j) { const int x = g[j]; g[j] = g[t]; g[t] = x; }
    }
  }
  __syncthreads();
  for (int b0 = 0; b0 < np * nn; b0 += nthr * PF_U) {
    double vv[PF_U];
#pragma unroll
    for (int u = 0; u < PF_U; ++u) {
      const int e = b0 + tid + nthr * u;
      if (e < np * nn) {
        const int p = e / nn, r = e % nn;
        const int row = r % n, col = r / n;
        if (sh.info[p] == 0) vv[u] = planes[(size_t)p * nn + sh.KS(p, row) + (size_t)sh.KS(p, col) * n];
      }
    }
#pragma unroll
    for (int u = 0; u < PF_U; ++u) {
      const int e = b0 + tid + nthr * u;
      if (e < np * nn && sh.info[e / nn] == 0) work[e] = sign != 1.0 ? -vv[u] : vv[u];
    }
  }
  __syncthreads();
  for (int b0 = 0; b0 < np * nn; b0 += nthr * PF_U) {
    double vv[PF_U];
#pragma unroll
    for (int u = 0; u < PF_U; ++u) {
      const int e = b0 + tid + nthr * u;
      if (e < np * nn && sh.info[e / nn] == 0) vv[u] = work[e];
    }
#pragma unroll
    for (int u = 0; u < PF_U; ++u) {
      const int e = b0 + tid + nthr * u;
      if (e < np * nn && sh.info[e / nn] == 0) planes[e] = vv[u];
    }
  }
  __syncthreads();
  PFC(8);
#undef PF_FRESH
}

}  // namespace mvmc_hip
