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

// LDS scratch of wg_pfinv, carved from dynamic shared memory by the caller:
// piv[np*n], colbuf[np*n], kp/info[np], colmax[np].
struct PfinvShared {
  int* piv;
  double* colbuf;
  int* kp;
  int* info;
  double* colmax;
  __device__ int& P(int p, int i) { return piv[p * ld + i]; }
  __device__ double& C(int p, int i) { return colbuf[p * ld + i]; }
  int ld;
};

__device__ inline size_t pfinv_lds_bytes(int n, int np) {
  return (size_t)np * n * (sizeof(int) + sizeof(double)) + (size_t)np * (2 * sizeof(int) + sizeof(double)) + 64;
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

// Core routine.  planes: np planes (stride n*n).  work: np planes of scratch
// (the unit upper triangular M).  On exit, for every plane with info == 0,
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
  for (int k0 = n - 1; k0 >= 1; --k0) {
    const int kk0 = k0 - 1;
    // pivot search: one wave per plane (round robin)
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
    // swaps (disjoint element sets, see ltl.rs:131-158)
    for (int p = 0; p < np; ++p) {
      const int kp = sh.kp[p];
      if (sh.colmax[p] == 0.0 || kp == kk0) continue;
      double* a = planes + (size_t)p * nn;
      for (int j = tid; j < kp; j += nthr) {
        double t = a[j + (size_t)kk0 * n]; a[j + (size_t)kk0 * n] = a[j + (size_t)kp * n]; a[j + (size_t)kp * n] = t;
      }
      for (int j = kp + 1 + tid; j < kk0; j += nthr) {
        double t = a[j + (size_t)kk0 * n]; a[j + (size_t)kk0 * n] = a[kp + (size_t)j * n]; a[kp + (size_t)j * n] = t;
      }
      for (int j = k0 + tid; j < n; j += nthr) {
        double t = a[kk0 + (size_t)j * n]; a[kk0 + (size_t)j * n] = a[kp + (size_t)j * n]; a[kp + (size_t)j * n] = t;
      }
    }
    __syncthreads();
    for (int p = 0; p < np; ++p) {
      const int kp = sh.kp[p];
      if (sh.colmax[p] == 0.0 || kp == kk0) continue;
      double* a = planes + (size_t)p * nn;
      for (int j = kp + tid; j < kk0; j += nthr) a[j + (size_t)kk0 * n] = -a[j + (size_t)kk0 * n];
      for (int j = kp + 1 + tid; j < kk0; j += nthr) a[kp + (size_t)j * n] = -a[kp + (size_t)j * n];
    }
    if (tid < np) {
      const int p = tid;
      if (sh.colmax[p] == 0.0) {
        if (sh.info[p] == 0) sh.info[p] = kk0 + 1;
        sh.P(p, kk0) = kk0 + 1;
      } else {
        sh.P(p, kk0) = sh.kp[p] + 1;
      }
    }
    __syncthreads();
    if (kk0 >= 1) {
      // rank-2 update of the strict upper triangle of A[0:kk0, 0:kk0];
      // work items: (plane, column j) per wave, rows over lanes.
      const int ncolwork = np * kk0;
      for (int cw = wave; cw < ncolwork; cw += nwave) {
        const int p = cw / kk0, j = cw % kk0;
        if (sh.colmax[p] == 0.0) continue;
        double* a = planes + (size_t)p * nn;
        const double* x = a + (size_t)k0 * n;
        const double* y = a + (size_t)kk0 * n;
        const double alpha = 1.0 / a[kk0 + (size_t)k0 * n];
        const double t1 = alpha * y[j];
        const double t2 = alpha * x[j];
        double* c = a + (size_t)j * n;
        for (int i = lane; i < j; i += 64) c[i] = (c[i] + x[i] * t1) - y[i] * t2;
        if (lane == 0) c[j] = 0.0;
      }
      __syncthreads();
      for (int e = tid; e < np * kk0; e += nthr) {
        const int p = e / kk0, i = e % kk0;
        if (sh.colmax[p] == 0.0) continue;
        double* a = planes + (size_t)p * nn;
        const double alpha = 1.0 / a[kk0 + (size_t)k0 * n];
        a[i + (size_t)k0 * n] *= alpha;
      }
      __syncthreads();
    }
  }
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
  // trtri('U','U') on B(i,j) = A(i,j+1), 0 <= i < j < n-1, column by column;
  // new column values are buffered so every element reads the original
  // B(k,j) (k > i) and the finished B(i,k) (k < j), as the nested C loop does.
  for (int j = 1; j < n - 1; ++j) {
    for (int e = tid; e < np * j; e += nthr) {
      const int p = e / j, i = e % j;
      if (sh.info[p] != 0) continue;
      const double* a = planes + (size_t)p * nn;
      const double* colj = a + (size_t)(j + 1) * n;
      double acc = colj[i];
      for (int k = i + 1; k < j; ++k) acc += a[i + (size_t)(k + 1) * n] * colj[k];
      sh.C(p, i) = -acc;
    }
    __syncthreads();
    for (int e = tid; e < np * j; e += nthr) {
      const int p = e / j, i = e % j;
      if (sh.info[p] != 0) continue;
      planes[(size_t)p * nn + i + (size_t)(j + 1) * n] = sh.C(p, i);
    }
    __syncthreads();
  }
  // M = I; M(ir, jr+1) = A(ir, jr+2) for ir <= jr < n-2; vt(i) = -A(i,i+1)
  for (size_t e = tid; e < (size_t)np * nn; e += nthr) {
    const int p = (int)(e / nn), r = (int)(e % nn);
    if (sh.info[p] != 0) continue;
    const int row = r % n, col = r / n;
    double v = (row == col) ? 1.0 : 0.0;
    if (col >= 1 && col <= n - 2 && row <= col - 1) v = planes[(size_t)p * nn + row + (size_t)(col + 1) * n];
    work[e] = v;
  }
  for (int e = tid; e < np * (n - 1); e += nthr) {
    const int p = e / (n - 1), i = e % (n - 1);
    if (sh.info[p] != 0) continue;
    sh.C(p, i) = -planes[(size_t)p * nn + i + (size_t)(i + 1) * n];
  }
  __syncthreads();
  // skew-tridiagonal solve T C = M, C overwrites A (column per thread)
  for (int e = tid; e < np * n; e += nthr) {
    const int p = e / n, j = e % n;
    if (sh.info[p] != 0) continue;
    const double* vt = &sh.C(p, 0);
    const double* b = work + (size_t)p * nn + (size_t)j * n;
    double* c = planes + (size_t)p * nn + (size_t)j * n;
    c[1] = b[0] / -vt[0];
    for (int i = 2; i < n; i += 2) c[i + 1] = (b[i] - c[i - 1] * vt[i - 1]) / -vt[i];
    c[n - 2] = b[n - 1] / vt[n - 2];
    for (int i = n - 3; i >= 1; i -= 2) c[i - 1] = (b[i] + c[i + 1] * vt[i]) / vt[i - 1];
  }
  __syncthreads();
  // T = M^T C per element (i,j): A(i,j) + sum_{k<i} M(k,i) A(k,j), k ascending,
  // on the unpermuted columns.  For n > 64 the reference forms only the upper
  // triangle and restores skew symmetry; for n <= 64 it forms every element.
  const bool panel = n > 64;
  for (int j = 0; j < n; ++j) {
    const int rows = panel ? j : n;  // i < j (upper) or all i
    for (int e = tid; e < np * rows; e += nthr) {
      const int p = e / rows, i = e % rows;
      if (sh.info[p] != 0) continue;
      const double* a = planes + (size_t)p * nn + (size_t)j * n;
      const double* mi = work + (size_t)p * nn + (size_t)i * n;
      double acc = a[i];
      for (int k = 0; k < i; ++k) acc += mi[k] * a[k];
      sh.C(p, i) = acc;
    }
    __syncthreads();
    for (int e = tid; e < np * rows; e += nthr) {
      const int p = e / rows, i = e % rows;
      if (sh.info[p] != 0) continue;
      planes[(size_t)p * nn + i + (size_t)j * n] = sh.C(p, i);
    }
    __syncthreads();
  }
  if (panel) {
    for (size_t e = tid; e < (size_t)np * nn; e += nthr) {
      const int p = (int)(e / nn), r = (int)(e % nn);
      if (sh.info[p] != 0) continue;
      const int row = r % n, col = r / n;
      if (row == col) planes[e] = 0.0;
      else if (row > col) planes[e] = -planes[(size_t)p * nn + col + (size_t)row * n];
    }
    __syncthreads();
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
