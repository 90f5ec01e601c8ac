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

// Doubles of global scratch wg_pfinv needs for np planes of order n.
__host__ __device__ inline size_t pfinv_work_doubles(int n, int np) {
  return (size_t)np * n * n + (size_t)TJ * n;
}

// LDS scratch of wg_pfinv, carved from dynamic shared memory by the caller:
// piv[np*n], colbuf[np*n], kp/info[np], colmax[np], ks[np*n] (panel step of
// the last swap through each index), act[np*LTL_K] (step applied per plane).
struct PfinvShared {
  int* piv;
  double* colbuf;
  int* kp;
  int* info;
  double* colmax;
  int* ks;
  int* act;
  int* zmax;
  __device__ int& P(int p, int i) { return piv[p * ld + i]; }
  __device__ double& C(int p, int i) { return colbuf[p * ld + i]; }
  __device__ int& KS(int p, int i) { return ks[p * ld + i]; }
  int ld;
};

__host__ __device__ inline size_t pf_rnd16(size_t b) { return (b + 15) & ~(size_t)15; }

// Bytes of LDS that carve_pfinv takes.
__host__ __device__ inline size_t pf_bytes(int n, int np) {
  return 2 * pf_rnd16((size_t)np * n * sizeof(int)) + pf_rnd16((size_t)np * n * sizeof(double)) +
         3 * pf_rnd16((size_t)np * 8) + pf_rnd16((size_t)np * LTL_K * sizeof(int)) + pf_rnd16((size_t)np * sizeof(int));
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
  sh.zmax = c.template take<int>(np);
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
// of scratch (panel vectors of the elimination, then M^T and a staged column
// block).  On exit, for every plane with info == 0,
// pf[p] holds the Pfaffian and the plane holds sign * A^{-1}
// (sign = -1 gives mVMC's invM).  info[p]: 0 ok, k > 0 zero pivot at 1-based
// row k, -1 non-finite Pfaffian.  All threads of the block must call it.
__device__ void wg_pfinv(double* planes, double* work, int n, int np, PfinvShared& sh, double* pf, double sign) {
  const int tid = threadIdx.x, nthr = blockDim.x;
  const int lane = tid & 63, wave = tid >> 6, nwave = (nthr + 63) >> 6;
  const int nn = n * n;
  for (int e = tid; e < np * n; e += nthr) sh.P(e / n, e % n) = e % n + 1;
  for (int p = tid; p < np; p += nthr) sh.info[p] = 0;
  __syncthreads();

  // ---------------- LTL^T (dsktf2) ----------------
  // The rank-2 updates of the trailing block are delayed and applied per
  // panel of `kw` steps.  Every element still receives its updates in step
  // order, each as (c + x*t1) - y*t2 with the x, y, t1, t2 of that step, at
  // the position it held at that step:
  //  - before a swap, the elements it moves (columns kk0 and kp, row kp) and
  //    the new y column are brought up to date, so nothing pending is moved;
  //  - an element at (i, j) has received the steps before
  //    max(ks[i], ks[j]) (the last panel step whose swap wrote row/column i
  //    or j); the others stay pending with the step's x, y rows i, j, which
  //    later swaps do not touch.
  // Panel vectors (x, y, t1 = alpha*y, t2 = alpha*x per step) live in `work`.
  const int kw = n >= 4 * LTL_K ? LTL_K : (n >= 4 ? n / 4 : 1);
  auto PX = [&](int p, int t) { return work + ((size_t)(p * 4 + 0) * kw + t) * n; };
  auto PY = [&](int p, int t) { return work + ((size_t)(p * 4 + 1) * kw + t) * n; };
  auto PT1 = [&](int p, int t) { return work + ((size_t)(p * 4 + 2) * kw + t) * n; };
  auto PT2 = [&](int p, int t) { return work + ((size_t)(p * 4 + 3) * kw + t) * n; };
  // pending steps [s, t_end) of plane p on element (i, j), i < j
  // (the loads of CU_G pending steps are issued before their serial chain)
  constexpr int CU_G = 4;
  auto catch_up = [&](int p, int i, int j, int t_end) {
    const int s0 = max(sh.KS(p, i), sh.KS(p, j));
    if (s0 >= t_end) return;
    double* a = planes + (size_t)p * nn;
    double c = a[i + (size_t)j * n];
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
    a[i + (size_t)j * n] = c;
  };
  for (int e = tid; e < np * n; e += nthr) sh.ks[e] = 0;
  for (int p = tid; p < np; p += nthr) sh.zmax[p] = 0;
  __syncthreads();
  int tp = 0;  // steps recorded in the open panel
  for (int k0 = n - 1; k0 >= 1; --k0) {
    const int kk0 = k0 - 1;
    // pivot search on the (up to date) column k0: one wave per plane
    for (int p = wave; p < np; p += nwave) {
      const double* col = planes + (size_t)p * nn + (size_t)k0 * n;
      double bv = -2.0;
      int bj = 0x7fffffff;
      for (int j = lane; j <= kk0; j += 64) {
        double kv;
        piv_key(fabs(col[j]), j, kv);
        if (kv > bv) { bv = kv; bj = j; }
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
    // bring the y column (kk0) and, when swapping, column kp and row kp up to
    // date; items (plane, e) over all planes at once:
    //   e < kk0: (e, kk0);  then f = e - kk0: f < kp: (f, kp), else (kp, f + 1)
    if (tp > 0) {
      const int per = 2 * kk0 - 1;
      for (int it = tid; it < np * per; it += nthr) {
        const int p = it / per, e = it % per;
        const int kp = sh.kp[p];
        const bool sw = sh.colmax[p] != 0.0 && kp != kk0;
        if (e < kk0) catch_up(p, e, kk0, tp);
        else if (sw) {
          const int f = e - kk0;
          if (f < kp) catch_up(p, f, kp, tp);
          else catch_up(p, kp, f + 1, tp);
        }
      }
      __syncthreads();
    }
    // swaps and sign flips (ltl.rs:131-158), items (plane, e), e < n - 1:
    //   e < kp: a[e][kk0] <-> a[e][kp];  e == kp: a[kp][kk0] = -a[kp][kk0];
    //   kp < e < kk0: a[e][kk0] <-> a[kp][e], both negated;
    //   e >= kk0: rows kk0 <-> kp of column e + 1 (finished columns)
    for (int it = tid; it < np * (n - 1); it += nthr) {
      const int p = it / (n - 1), e = it % (n - 1);
      const int kp = sh.kp[p];
      if (sh.colmax[p] == 0.0 || kp == kk0) continue;
      double* a = planes + (size_t)p * nn;
      if (e < kp) {
        double t = a[e + (size_t)kk0 * n]; a[e + (size_t)kk0 * n] = a[e + (size_t)kp * n]; a[e + (size_t)kp * n] = t;
      } else if (e == kp) {
        a[kp + (size_t)kk0 * n] = -a[kp + (size_t)kk0 * n];
      } else if (e < kk0) {
        double t = a[e + (size_t)kk0 * n]; a[e + (size_t)kk0 * n] = -a[kp + (size_t)e * n]; a[kp + (size_t)e * n] = -t;
      } else {
        const int j = e + 1;
        double t = a[kk0 + (size_t)j * n]; a[kk0 + (size_t)j * n] = a[kp + (size_t)j * n]; a[kp + (size_t)j * n] = t;
      }
    }
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
    if (kk0 >= 1) {
      // record the step: x, y (unscaled), t1 = alpha*y[j], t2 = alpha*x[j]
      for (int e = tid; e < np * kk0; e += nthr) {
        const int p = e / kk0, i = e % kk0;
        if (sh.colmax[p] == 0.0) continue;
        const double* a = planes + (size_t)p * nn;
        const double alpha = 1.0 / a[kk0 + (size_t)k0 * n];
        const double x = a[i + (size_t)k0 * n], y = a[i + (size_t)kk0 * n];
        PX(p, tp)[i] = x;
        PY(p, tp)[i] = y;
        PT1(p, tp)[i] = alpha * y;
        PT2(p, tp)[i] = alpha * x;
      }
      __syncthreads();
      for (int e = tid; e < np * kk0; e += nthr) {
        const int p = e / kk0, i = e % kk0;
        if (sh.colmax[p] == 0.0) continue;
        double* a = planes + (size_t)p * nn;
        const double alpha = 1.0 / a[kk0 + (size_t)k0 * n];
        a[i + (size_t)k0 * n] *= alpha;
      }
    }
    ++tp;
    // close the panel: every element of the trailing triangle (i < j < kk0)
    // takes its pending steps; a lane owns row i of plane p and keeps the
    // row's x, y of all panel steps in registers.
    if (tp == kw || kk0 <= 1) {
      __syncthreads();
      const int m = kk0;  // trailing columns j < m, rows i < m - 1
      if (m >= 2) {
        // work item = (plane, 64-row chunk) per wave, so plane and column
        // are wave-uniform and the step scalars t1, t2 load once per wave
        const int chunks = (m - 1 + 63) / 64;
        for (int it = __builtin_amdgcn_readfirstlane(wave); it < np * chunks; it += nwave) {
          const int p = it / chunks, c = it % chunks;
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
          // JU columns per iteration, all their loads issued first
          constexpr int JU = 2;
          for (int j = c * 64 + 1; j < m; j += JU) {
            double t1v[JU][LTL_K], t2v[JU][LTL_K], cv[JU];
            int ksj[JU];
#pragma unroll
            for (int u = 0; u < JU; ++u) {
              const int jj = j + u;
              if (jj < m) {
                ksj[u] = sh.KS(p, jj);
#pragma unroll
                for (int t = 0; t < LTL_K; ++t)
                  if (t < tp) {
                    t1v[u][t] = PT1(p, t)[jj];
                    t2v[u][t] = PT2(p, t)[jj];
                  }
                if (on && jj > i) cv[u] = a[i + (size_t)jj * n];
              }
            }
#pragma unroll
            for (int u = 0; u < JU; ++u) {
              const int jj = j + u;
              if (jj < m && on && jj > i) {
                const int s0 = max(ksi, ksj[u]);
                if (s0 < tp) {
                  double cc = cv[u];
#pragma unroll
                  for (int t = 0; t < LTL_K; ++t)
                    if (t >= s0 && ar[t]) cc = (cc + xr[t] * t1v[u][t]) - yr[t] * t2v[u][t];
                  a[i + (size_t)jj * n] = cc;
                }
              }
            }
          }
        }
      }
      __syncthreads();
      for (int e = tid; e < np * n; e += nthr) sh.ks[e] = 0;
      tp = 0;
      __syncthreads();
    }
  }
  // the unblocked elimination zeroes the diagonal of every block it updated
  for (int e = tid; e < np * n; e += nthr) {
    const int p = e / n, j = e % n;
    if (j < sh.zmax[p]) planes[(size_t)p * nn + j + (size_t)j * n] = 0.0;
  }
  __syncthreads();
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
  // ---------------- inverse (utu2inv) ----------------
  // Planes one at a time; a lane owns a row and TJ columns of a column
  // block, whose original column values are staged in `stage` (global,
  // read by whole waves at one address) so results can be written in place.
  double* stage = work + (size_t)np * nn;
  for (int p = 0; p < np; ++p) {
    if (sh.info[p] != 0) continue;
    double* a = planes + (size_t)p * nn;
    // trtri('U','U') on B(i,j) = A(i,j+1), 0 <= i < j < n-1:
    // Bn(i,j) = -(B(i,j) + sum_{k=i+1}^{j-1} Bn(i,k) B(k,j)), k ascending.
    for (int j0 = 1; j0 < n - 1; j0 += TJ) {
      const int j1 = min(n - 1, j0 + TJ);
      for (int e = tid; e < TJ * n; e += nthr) {
        const int jj = e / n, k = e % n;
        stage[e] = (j0 + jj < j1 && k < j0 + jj) ? a[k + (size_t)(j0 + jj + 1) * n] : 0.0;
      }
      __syncthreads();
      for (int i0 = 0; i0 < j1 - 1; i0 += nthr) {
        const int i = i0 + tid;
        const bool on = i < j1 - 1;
        double acc[TJ];
#pragma unroll
        for (int jj = 0; jj < TJ; ++jj) acc[jj] = (on && j0 + jj < j1 && i < j0 + jj) ? stage[jj * n + i] : 0.0;
        // k < j0: finished columns of earlier blocks (KU loads issued together)
        const int kst = i0 + (tid & ~63) + 1;  // wave-uniform start
        constexpr int KU = 4;
        for (int k = kst; k < j0; k += KU) {
          double bn[KU];
#pragma unroll
          for (int u = 0; u < KU; ++u) {
            const int kk = k + u;
            bn[u] = (on && kk < j0 && kk > i) ? a[i + (size_t)(kk + 1) * n] : 0.0;
          }
#pragma unroll
          for (int u = 0; u < KU; ++u) {
            const int kk = k + u;
            if (on && kk < j0 && kk > i) {
#pragma unroll
              for (int jj = 0; jj < TJ; ++jj)
                if (j0 + jj < j1) acc[jj] += bn[u] * stage[jj * n + kk];
            }
          }
        }
        // k in [j0, j'): columns of this block, finished by this lane in order
#pragma unroll
        for (int jj = 0; jj < TJ; ++jj) {
#pragma unroll
          for (int kk = 0; kk < jj; ++kk)
            if (on && j0 + jj < j1 && i < j0 + kk) acc[jj] += acc[kk] * stage[jj * n + j0 + kk];
          acc[jj] = -acc[jj];
        }
        // in place: this block reads its own columns only from `stage`
#pragma unroll
        for (int jj = 0; jj < TJ; ++jj)
          if (on && j0 + jj < j1 && i < j0 + jj) a[i + (size_t)(j0 + jj + 1) * n] = acc[jj];
      }
      __syncthreads();
    }
    // Mt = M^T with M = I, M(ir, jr+1) = A(ir, jr+2) for ir <= jr < n-2;
    // vt(i) = -A(i, i+1)
    double* mt = work + (size_t)p * nn;
    for (int e = tid; e < nn; e += nthr) {
      const int col = e % n, row = e / n;  // Mt(col, row) = M(row, col) at col + row*n
      double v = (row == col) ? 1.0 : 0.0;
      if (col >= 1 && col <= n - 2 && row <= col - 1) v = a[row + (size_t)(col + 1) * n];
      mt[e] = v;
    }
    for (int i = tid; i < n - 1; i += nthr) sh.C(p, i) = -a[i + (size_t)(i + 1) * n];
    __syncthreads();
    // skew-tridiagonal solve T C = M, C overwrites A (column per lane; the
    // recurrence carries its last value in a register, right-hand sides are
    // loaded SU at a time ahead of it)
    for (int j = tid; j < n; j += nthr) {
      constexpr int SU = 8;
      const double* vt = &sh.C(p, 0);
      double* c = a + (size_t)j * n;
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
    // T = M^T C: T(i,j) = C(i,j) + sum_{k<i} M(k,i) C(k,j), k ascending.
    // For n > 64 the reference forms only the upper triangle and restores
    // skew symmetry; for n <= 64 it forms every element.
    const bool upper = n > 64;
    for (int j0 = 0; j0 < n; j0 += TJ) {
      const int j1 = min(n, j0 + TJ);
      for (int e = tid; e < TJ * n; e += nthr) {
        const int jj = e / n, k = e % n;
        stage[e] = (j0 + jj < j1) ? a[k + (size_t)(j0 + jj) * n] : 0.0;
      }
      __syncthreads();
      const int rows = upper ? j1 - 1 : n;
      for (int i0 = 0; i0 < rows; i0 += nthr) {
        const int i = i0 + tid;
        const bool on = i < rows;
        double acc[TJ];
#pragma unroll
        for (int jj = 0; jj < TJ; ++jj) acc[jj] = on ? stage[jj * n + (on ? i : 0)] : 0.0;
        const int kend = min(rows, i0 + (tid & ~63) + 64) - 1;  // wave's largest row
        constexpr int KU = 4;
        for (int k = 0; k < kend; k += KU) {
          double mv[KU];
#pragma unroll
          for (int u = 0; u < KU; ++u) mv[u] = (on && k + u < i) ? mt[i + (size_t)(k + u) * n] : 0.0;
#pragma unroll
          for (int u = 0; u < KU; ++u)
            if (on && k + u < i) {
#pragma unroll
              for (int jj = 0; jj < TJ; ++jj) acc[jj] += mv[u] * stage[jj * n + k + u];
            }
        }
#pragma unroll
        for (int jj = 0; jj < TJ; ++jj) {
          const int j = j0 + jj;
          if (on && j < j1 && (!upper || i < j)) a[i + (size_t)j * n] = acc[jj];
        }
      }
      __syncthreads();
    }
    if (upper) {
      for (int e = tid; e < nn; e += nthr) {
        const int row = e % n, col = e / n;
        if (row == col) a[e] = 0.0;
        else if (row > col) a[e] = -a[col + (size_t)row * n];
      }
      __syncthreads();
    }
  }
  // column then row permutations (sequential swap order per row / column),
  // then the sign (calc_m_all's final DSCAL by -1).
  for (int e = tid; e < np * n; e += nthr) {
    const int p = e / n, r = e % n;
    if (sh.info[p] != 0) continue;
    double* a = planes + (size_t)p * nn;
    for (int j = 0; j < n; ++j) {
      const int t = sh.P(p, j) - 1;
      if (t != j) {
        double x = a[r + (size_t)j * n];
        a[r + (size_t)j * n] = a[r + (size_t)t * n];
        a[r + (size_t)t * n] = x;
      }
    }
  }
  __syncthreads();
  for (int e = tid; e < np * n; e += nthr) {
    const int p = e / n, c = e % n;
    if (sh.info[p] != 0) continue;
    double* col = planes + (size_t)p * nn + (size_t)c * n;
    for (int i = 0; i < n; ++i) {
      const int t = sh.P(p, i) - 1;
      if (t != i) { double x = col[i]; col[i] = col[t]; col[t] = x; }
    }
    if (sign != 1.0)
      for (int i = 0; i < n; ++i) col[i] = -col[i];
  }
  __syncthreads();
}

}  // namespace mvmc_hip
