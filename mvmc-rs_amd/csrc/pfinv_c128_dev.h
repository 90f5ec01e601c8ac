// Batched Pfaffian + inverse of complex skew-symmetric planes on gfx950.
//
// Follows, per matrix element, the operation order of the reference's scalar
// pfapack build:
//   LTL^T      crates/pfapack/src/ltl.rs:61-191 with
//              order 0 = zsktf2 (ltl.rs:39-43; the BatchedEngine::run_c64
//              reference crates/mvmc-gpu/src/scalar.rs:146-170):
//                alpha = Julia 1/pivot (julia_complex.rs:36-77), column
//                scaling by complex products (backend.rs:351-355);
//              order 1 = zsktf2_c_compat (ltl.rs:54-59; production
//              calc_m_all_complex, crates/mvmc-core/src/pfaffian.rs:383-398):
//                alpha = (re, -im)/(re^2 + im^2) (ltl.rs:256-263), column
//                scaling with fused multiply-adds (ltl.rs:303-315);
//              both: rank-2 update c + (x t1 - y t2), t1 = alpha y_j,
//              t2 = alpha x_j (ltl.rs:707-818)
//   Pfaffian   utu2.rs:61-93
//   inverse    utu2.rs:399-515 (c_real = false): trtri 'U','U'
//              (backend.rs:231-251), Julia-style skew-tridiagonal solve with
//              Julia reciprocals (utu2.rs:110-282, julia_complex.rs:79-122),
//              n > 64: upper product + skew fill before the permutations
//              (backend.rs:265-300, utu2.rs:305-330), else column
//              permutation, M^T A (backend.rs:254-263), row permutation.
// Complex arithmetic is num_complex's ((a+ib)(c+id) = (ac - bd) + i(ad + bc),
// no contraction: the library is built with -ffp-contract=off).
//
// Layout: planes [count][n][n] complex (re, im interleaved), column-major
// per plane, A(i,j) at i + j*n.  One workgroup (256 lanes) per plane, a
// persistent grid striding over planes.  The plane is factorised in place;
// the workgroup's two work planes hold M^T (the unit upper triangular factor
// transposed so that lanes over rows read consecutive addresses in the solve
// and in the product) and the product M^T C.  Every stage is elementwise
// parallel except the ordered sums, which stay on one lane each.
#pragma once
#include <hip/hip_runtime.h>

namespace mvmc_hip {

constexpr int PFC_NMAX = 512;  // largest plane side (LDS vectors below)
constexpr int PFC_U = 4;       // elements a lane loads before it stores in the rank-2 update

struct alignas(16) cz {
  double re, im;
};
__device__ __forceinline__ cz cz_mul(cz a, cz b) { return {a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re}; }
__device__ __forceinline__ cz cz_add(cz a, cz b) { return {a.re + b.re, a.im + b.im}; }
__device__ __forceinline__ cz cz_sub(cz a, cz b) { return {a.re - b.re, a.im - b.im}; }
__device__ __forceinline__ cz cz_neg(cz a) { return {-a.re, -a.im}; }

// Julia's ComplexF64 reciprocal (crates/pfapack/src/julia_complex.rs:79-122).
__device__ inline cz julia_recip(cz w) {
  double c = w.re, d = w.im;
  const double absc = fabs(c), absd = fabs(d);
  const double cd = absc > absd ? absc : absd;
  const double dc = absc > absd ? absd : absc;
  // sqrt(DBL_MIN/2) <= cd <= sqrt(DBL_MAX/2)
  if (0x1.6a09e667f3bcdp-512 <= cd && cd <= 0x1.6a09e667f3bccp+511) {
    const double norm = fma(cd, cd, dc * dc);
    return {c / norm, -d / norm};
  }
  if (isinf(c) || isinf(d)) return {copysign(0.0, c), copysign(-0.0, -d)};
  const double bs = 0x1p+105;  // 2 / eps^2
  double scale = 1.0;
  if (cd >= 0x1.fffffffffffffp+1022) {  // DBL_MAX / 2
    c *= 0.5;
    d *= 0.5;
    scale = 0.5;
  } else if (cd <= 0x1p-969) {  // 2 DBL_MIN / eps
    c *= bs;
    d *= bs;
    scale = bs;
  }
  double p, q;
  if (absd <= absc) {
    const double r = d / c, z = fma(d, r, c);
    p = 1.0 / z;
    q = -r / z;
  } else {
    const double r = -c / -d, z = fma(-c, r, -d);
    q = 1.0 / z;
    p = -r / z;
  }
  return {p * scale, q * scale};
}

__device__ inline double julia_div_part(double a, double b, double c, double d, double r, double t) {
  if (r != 0.0) {
    const double br = b * r;
    return br != 0.0 ? (a + br) * t : a * t + (b * t) * r;
  }
  return (a + d * (b / c)) * t;
}

// Julia's scaled ComplexF64 division (julia_complex.rs:9-77).
__device__ inline cz julia_div(cz z, cz w) {
  double a = z.re, b = z.im, c = w.re, d = w.im;
  const double ab = fabs(a) >= fabs(b) ? fabs(a) : fabs(b);
  const double cdm = fabs(c) >= fabs(d) ? fabs(c) : fabs(d);
  if (isinf(c) || isinf(d)) {
    auto sgn = [](double x) { return (x == 0.0 || isnan(x)) ? x : copysign(1.0, x); };
    if (isfinite(a) && isfinite(b)) return {0.0 * sgn(a) * sgn(c), -0.0 * sgn(b) * sgn(d)};
    return {__builtin_nan(""), __builtin_nan("")};
  }
  const double halfov = 0x1.fffffffffffffp+1022, twouneps = 0x1p-969;
  double s = 1.0;
  const bool scaled = ab >= halfov || ab <= twouneps || cdm >= halfov || cdm <= twouneps;
  if (scaled) {
    const double bs = 0x1p+105;
    if (ab >= halfov) {
      a *= 0.5;
      b *= 0.5;
      s *= 2.0;
    } else if (ab <= twouneps) {
      a *= bs;
      b *= bs;
      s /= bs;
    }
    if (cdm >= halfov) {
      c *= 0.5;
      d *= 0.5;
      s *= 0.5;
    } else if (cdm <= twouneps) {
      c *= bs;
      d *= bs;
      s *= bs;
    }
  }
  bool swapped = false;
  if (!(fabs(d) <= fabs(c))) {
    double t0 = a;
    a = b;
    b = t0;
    t0 = c;
    c = d;
    d = t0;
    swapped = true;
  }
  const double r = d / c;
  const double t = 1.0 / (c + d * r);
  const double p = julia_div_part(a, b, c, d, r, t);
  double q = julia_div_part(b, -a, c, d, r, t);
  if (swapped) q = -q;
  if (scaled) return {p * s, q * s};
  return {p, q};
}

struct PfcShared {
  int piv[PFC_NMAX];
  int lab_c[PFC_NMAX];  // final column j holds column lab_c[j] (sequential swaps composed)
  int lab_r[PFC_NMAX];
  // four complex vectors of PFC_NMAX: during the LTL x, y, t1, t2 of the
  // current step; during the inverse vt, 1/(-vt), 1/vt and a trtri column
  cz buf[4 * PFC_NMAX];
  int kp, zero_col, info;
};

// One plane: a (in place, becomes the true inverse or zeros), mt / wk: work
// planes of n*n.  Writes *pf and *status (thread 0).
__device__ void wg_pfinv_c128(cz* a, cz* mt, cz* wk, int n, int order, PfcShared& sh, cz* pf, int* status) {
  const int tid = threadIdx.x;
  const int nthr = blockDim.x;
  auto A = [&](int i, int j) -> cz& { return a[i + (size_t)j * n]; };
  cz* const cx = sh.buf;                 // LTL step: column k0
  cz* const cy = sh.buf + PFC_NMAX;      //           column kk0
  cz* const t1v = sh.buf + 2 * PFC_NMAX; //           alpha y_j
  cz* const t2v = sh.buf + 3 * PFC_NMAX; //           alpha x_j
  cz* const vtv = sh.buf;                // inverse: vt
  cz* const rneg = sh.buf + PFC_NMAX;    //          Julia 1 / (-vt[i])
  cz* const rpos = sh.buf + 2 * PFC_NMAX;//          Julia 1 / vt[i]
  cz* const tcol = sh.buf + 3 * PFC_NMAX;//          trtri column
  for (int i = tid; i < n; i += nthr) sh.piv[i] = i + 1;
  if (tid == 0) sh.info = 0;
  __syncthreads();

  // ---------------- LTL^T (ltl.rs:61-191) ----------------
  for (int k0 = n - 1; k0 >= 1; --k0) {
    const int kk0 = k0 - 1;
    // first index of the largest |re| + |im| in A(0..kk0, k0) (strict >,
    // NaN never wins; a NaN at j = 0 keeps kp = 0)
    if (tid < 64) {
      double bv = -1.0;
      int bj = 0x7fffffff;
      bool nan0 = false;
      for (int j = tid; j <= kk0; j += 64) {
        const cz x = A(j, k0);
        const double v = fabs(x.re) + fabs(x.im);
        if (j == 0 && v != v) nan0 = true;
        if (v > bv) {
          bv = v;
          bj = j;
        }
      }
      for (int off = 32; off >= 1; off >>= 1) {
        const double ov = __shfl_xor(bv, off, 64);
        const int oj = __shfl_xor(bj, off, 64);
        if (ov > bv || (ov == bv && oj < bj)) {
          bv = ov;
          bj = oj;
        }
      }
      nan0 = __shfl(nan0 ? 1 : 0, 0, 64) != 0;
      if (tid == 0) {
        if (nan0) {
          sh.kp = 0;
          sh.zero_col = 0;
        } else {
          sh.kp = bj;
          sh.zero_col = (bv == 0.0) ? 1 : 0;
        }
      }
    }
    __syncthreads();
    const int kp = sh.kp;
    if (sh.zero_col) {
      if (tid == 0) {
        if (sh.info == 0) sh.info = kk0 + 1;
        sh.piv[kk0] = kk0 + 1;
      }
      __syncthreads();
      continue;
    }
    if (kp != kk0) {
      for (int t = tid; t < n; t += nthr) {
        if (t < kp) {
          const cz u = A(t, kk0);
          A(t, kk0) = A(t, kp);
          A(t, kp) = u;
        } else if (t > kp && t < kk0) {
          const cz u = A(t, kk0);
          A(t, kk0) = A(kp, t);
          A(kp, t) = u;
        } else if (t >= k0) {
          const cz u = A(kk0, t);
          A(kk0, t) = A(kp, t);
          A(kp, t) = u;
        }
      }
      __syncthreads();
      for (int t = tid; t < kk0; t += nthr) {
        if (t >= kp) A(t, kk0) = cz_neg(A(t, kk0));
        if (t > kp) A(kp, t) = cz_neg(A(kp, t));
      }
      __syncthreads();
    }
    if (tid == 0) sh.piv[kk0] = kp + 1;
    if (kk0 >= 1) {
      const cz pv = A(kk0, k0);
      cz alpha;
      if (order == 0) {
        alpha = julia_div(cz{1.0, 0.0}, pv);
      } else {
        const double den = pv.re * pv.re + pv.im * pv.im;
        alpha = cz{pv.re / den, -pv.im / den};
      }
      // x = A(0..kk0-1, k0), y = A(0..kk0-1, kk0) and the per-column
      // t1 = alpha y_j, t2 = alpha x_j into LDS: the update below then reads
      // and writes only its own elements
      for (int t = tid; t < kk0; t += nthr) {
        const cz x = A(t, k0), y = A(t, kk0);
        cx[t] = x;
        cy[t] = y;
        t1v[t] = cz_mul(alpha, y);
        t2v[t] = cz_mul(alpha, x);
      }
      __syncthreads();
      // strict upper triangle i < j < kk0 enumerated without gaps: column j
      // holds j elements starting at j(j-1)/2; a lane loads PFC_U elements
      // before it stores any.  The column scaling of k0 writes elements the
      // update neither reads nor writes.
      const int tot = kk0 * (kk0 - 1) / 2;
      for (int e0 = tid; e0 < tot; e0 += nthr * PFC_U) {
        int ii[PFC_U], jj[PFC_U];
        cz c[PFC_U];
#pragma unroll
        for (int u = 0; u < PFC_U; ++u) {
          const int e = e0 + u * nthr;
          int j = (int)((1.0 + sqrt(1.0 + 8.0 * (double)e)) * 0.5);
          while (j > 1 && j * (j - 1) / 2 > e) --j;
          while ((j + 1) * j / 2 <= e) ++j;
          ii[u] = e - j * (j - 1) / 2;
          jj[u] = j;
          if (e < tot) c[u] = A(ii[u], j);
        }
#pragma unroll
        for (int u = 0; u < PFC_U; ++u) {
          if (e0 + u * nthr < tot) {
            const int i = ii[u], j = jj[u];
            const cz p = cz_mul(cx[i], t1v[j]);
            const cz q = cz_mul(cy[i], t2v[j]);
            A(i, j) = cz_add(c[u], cz_sub(p, q));
          }
        }
      }
      for (int t = tid; t < kk0; t += nthr) {
        A(t, t) = cz{0.0, 0.0};
        const cz v = cx[t];
        A(t, k0) = order == 0 ? cz_mul(v, alpha)
                              : cz{fma(alpha.re, v.re, -(alpha.im * v.im)), fma(alpha.re, v.im, alpha.im * v.re)};
      }
    }
    __syncthreads();
  }

  // ---------------- Pfaffian (utu2.rs:61-93) ----------------
  if (tid == 0) {
    cz v{1.0, 0.0};
    for (int i = 0; i + 1 < n; i += 2) v = cz_mul(v, A(i, i + 1));
    bool neg = false;
    for (int k = 0; k < n; ++k)
      if (sh.piv[k] != k + 1) neg = !neg;
    const cz sign = neg ? cz{0.0 - 1.0, 0.0 - 0.0} : cz{1.0, 0.0};
    v = cz_mul(sign, v);
    int st = sh.info;
    if (st != 0) v = cz{0.0, 0.0};
    else if (!(isfinite(v.re) && isfinite(v.im))) st = -1;
    *pf = v;
    *status = st;
    sh.info = st;
  }
  __syncthreads();
  const size_t nn = (size_t)n * n;
  if (sh.info != 0) {
    for (size_t e = tid; e < nn; e += nthr) a[e] = cz{0.0, 0.0};
    __syncthreads();
    return;
  }

  // ---------------- inverse (utu2.rs:399-515) ----------------
  // (1) trtri('U','U') on B(i,j) = A(i,j+1): column j from the inverted
  //     columns before it and its own original entries, rows in parallel
  for (int j = 1; j < n - 1; ++j) {
    for (int i = tid; i < j; i += nthr) {
      cz acc = A(i, j + 1);
      for (int k = i + 1; k < j; ++k) acc = cz_add(acc, cz_mul(A(i, k + 1), A(k, j + 1)));
      tcol[i] = cz_neg(acc);
    }
    __syncthreads();
    for (int i = tid; i < j; i += nthr) A(i, j + 1) = tcol[i];
    __syncthreads();
  }
  // (2) M = I except M(k,i) = A(k,i+1) for 1 <= i <= n-2, k < i, stored
  //     transposed: mt(i,k) = M(k,i) at i + k*n
  for (size_t e = tid; e < nn; e += nthr) {
    const int i = (int)(e % n), k = (int)(e / n);  // mt(i,k) = M(k,i)
    cz m{0.0, 0.0};
    if (i >= 1 && i <= n - 2 && k <= i - 1) m = A(k, i + 1);
    else if (i == k) m = cz{1.0, 0.0};
    mt[e] = m;
  }
  for (int i = tid; i < n - 1; i += nthr) {
    const cz v = cz_neg(A(i, i + 1));
    vtv[i] = v;
    rneg[i] = julia_recip(cz_neg(v));
    rpos[i] = julia_recip(v);
  }
  if (tid == 0) {
    for (int j = 0; j < n; ++j) {
      sh.lab_c[j] = j;
      sh.lab_r[j] = j;
    }
    for (int j = 0; j < n; ++j) {
      const int t = sh.piv[j] - 1;
      if (t != j) {
        int u = sh.lab_c[j];
        sh.lab_c[j] = sh.lab_c[t];
        sh.lab_c[t] = u;
        u = sh.lab_r[j];
        sh.lab_r[j] = sh.lab_r[t];
        sh.lab_r[t] = u;
      }
    }
  }
  __syncthreads();
  // (3) skew-tridiagonal solve T C = M, one column per lane, C into A
  auto Mx = [&](int r, int j) -> cz { return mt[j + (size_t)r * n]; };  // M(r, j)
  for (int j = tid; j < n; j += nthr) {
    A(1, j) = cz_mul(Mx(0, j), rneg[0]);
    for (int ij = 2; ij < n; ij += 2)
      A(ij + 1, j) = cz_mul(cz_sub(Mx(ij, j), cz_mul(A(ij - 1, j), vtv[ij - 1])), rneg[ij]);
    A(n - 2, j) = cz_mul(Mx(n - 1, j), rpos[n - 2]);
    if (n >= 4)
      for (int i = n - 3; i >= 1; i -= 2)
        A(i - 1, j) = cz_mul(cz_add(Mx(i, j), cz_mul(A(i + 1, j), vtv[i])), rpos[i - 1]);
  }
  __syncthreads();
  // (4) product W = M^T C: n > 64, the strict upper triangle of M^T C (then
  //     skew fill and both permutations); else all of M^T (C P) with the
  //     column permutation folded into the read.  Sums ascend in k.
  const bool panel = n > 64;
  for (size_t e = tid; e < nn; e += nthr) {
    const int i = (int)(e % n), j = (int)(e / n);
    if (panel && i >= j) continue;
    const int cj = panel ? j : sh.lab_c[j];
    const cz* ccol = a + (size_t)cj * n;
    cz acc = ccol[i];
    for (int k = 0; k < i; ++k) acc = cz_add(acc, cz_mul(mt[i + (size_t)k * n], ccol[k]));
    wk[e] = acc;
  }
  __syncthreads();
  // (5) gather: out(i,j) = W(R[i], j) or, with the panel, F(R[i], L[j]),
  //     F = W's strict upper triangle, zero diagonal, skew lower triangle
  for (size_t e = tid; e < nn; e += nthr) {
    const int i = (int)(e % n), j = (int)(e / n);
    const int r = sh.lab_r[i];
    cz v;
    if (!panel) {
      v = wk[r + (size_t)j * n];
    } else {
      const int c = sh.lab_c[j];
      if (r < c) v = wk[r + (size_t)c * n];
      else if (r == c) v = cz{0.0, 0.0};
      else v = cz_neg(wk[c + (size_t)r * n]);
    }
    a[e] = v;
  }
  __syncthreads();
}

__global__ void __launch_bounds__(256) k_pfinv_c128(cz* planes, cz* work, int n, int count, int order, cz* pf,
                                                    int* status) {
  __shared__ PfcShared sh;
  const size_t nn = (size_t)n * n;
  cz* mt = work + (size_t)blockIdx.x * 2 * nn;
  cz* wk = mt + nn;
  for (int p = blockIdx.x; p < count; p += gridDim.x)
    wg_pfinv_c128(planes + (size_t)p * nn, mt, wk, n, order, sh, pf + p, status + p);
}

}  // namespace mvmc_hip
