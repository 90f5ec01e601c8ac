// ORACLE (test infrastructure only): complex skew-symmetric LTL^T
// factorisation, Pfaffian and inverse in the operation order of the
// reference's scalar (no BLAS) pfapack build.
//   zsktf2         crates/pfapack/src/ltl.rs:39-43 (order 0, the batched
//                  run_c64 reference crates/mvmc-gpu/src/scalar.rs:146-170):
//                  alpha = Julia 1/pivot, rank-2 update of
//                  update_upper_rank2_c64_scalar (ltl.rs:707-798), column
//                  scaling by num_complex products
//   zsktf2_c_compat ltl.rs:54-59 (order 1, the production calc_m_all_complex,
//                  crates/mvmc-core/src/pfaffian.rs:383-398,947-1011):
//                  alpha = (re, -im) / (re^2 + im^2) (ltl.rs:256-263), the
//                  same per-element update (ltl.rs:800-818), column scaling
//                  with fused multiply-adds (ltl.rs:303-315)
//   utu2pfa        crates/pfapack/src/utu2.rs:57-93
//   utu2inv        utu2.rs:349-360,399-515 (c_real = false): trtri and trmm
//                  scalar forms (backend.rs:231-263), Julia-style
//                  skew-tridiagonal solve with Julia reciprocals
//                  (utu2.rs:110-282, julia_complex.rs:79-122), for n > 64 the
//                  upper panel product and skew fill (backend.rs:265-300,
//                  utu2.rs:305-330) before the permutations.
// Complex arithmetic is num_complex's: (a+ib)(c+id) = (ac - bd) + i(ad + bc),
// sums and differences componentwise, no contraction (build with
// -ffp-contract=off).  Matrices are column-major n x n complex (re, im
// interleaved), pivots 1-based.
// Pinned by tests/test_oracle_pfaffian_c128.py: closed forms for n = 2 and 4,
// brute-force Pfaffian expansion for n <= 10, Pf^2 = det (independent LU,
// modulus and phase), inverse residual, on the reference's committed complex
// input matrices (tests/golden/pfapack_complex_inputs.npz); Julia's division
// and reciprocal against Python's complex arithmetic within 1 ulp and exact
// on the cases the scaling branches cover.  The production reference links
// OpenBLAS ztrtri / ztrmm (blas-backend), whose summation orders are not
// bit-reproducible, so against that build the inverse is pinned by residuals
// only.
#pragma once
#include <cfloat>
#include <cmath>
#include <utility>
#include <vector>

namespace oracle {

struct zc {
  double re, im;
};
inline zc zmul(zc a, zc b) { return {a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re}; }
inline zc zadd(zc a, zc b) { return {a.re + b.re, a.im + b.im}; }
inline zc zsub(zc a, zc b) { return {a.re - b.re, a.im - b.im}; }
inline zc zneg(zc a) { return {-a.re, -a.im}; }

// Julia's ComplexF64 reciprocal (julia_complex.rs:79-122).
inline zc julia_reciprocal(zc w) {
  double c = w.re, d = w.im;
  const double absc = std::fabs(c), absd = std::fabs(d);
  const double cd = absc > absd ? absc : absd;
  const double dc = absc > absd ? absd : absc;
  if (std::sqrt(DBL_MIN / 2.0) <= cd && cd <= std::sqrt(DBL_MAX / 2.0)) {
    const double norm = std::fma(cd, cd, dc * dc);
    return {c / norm, -d / norm};
  }
  if (std::isinf(c) || std::isinf(d)) return {std::copysign(0.0, c), std::copysign(-0.0, -d)};
  const double bs = 2.0 / (DBL_EPSILON * DBL_EPSILON);
  double scale = 1.0;
  if (cd >= DBL_MAX / 2.0) {
    c *= 0.5;
    d *= 0.5;
    scale = 0.5;
  } else if (cd <= 2.0 * DBL_MIN / DBL_EPSILON) {
    c *= bs;
    d *= bs;
    scale = bs;
  }
  auto robust = [](double c, double d, double& p, double& q) {
    const double r = d / c;
    const double z = std::fma(d, r, c);
    p = 1.0 / z;
    q = -r / z;
  };
  double p, q;
  if (absd <= absc) {
    robust(c, d, p, q);
  } else {
    robust(-d, -c, q, p);
  }
  return {p * scale, q * scale};
}

// Julia's scaled ComplexF64 division (julia_complex.rs:9-77).
inline double julia_div_component(double a, double b, double c, double d, double r, double t) {
  if (r != 0.0) {
    const double br = b * r;
    return br != 0.0 ? (a + br) * t : a * t + (b * t) * r;
  }
  return (a + d * (b / c)) * t;
}
inline zc julia_div_unscaled(double a, double b, double c, double d) {
  bool swapped = false;
  if (!(std::fabs(d) <= std::fabs(c))) {
    std::swap(a, b);
    std::swap(c, d);
    swapped = true;
  }
  const double r = d / c;
  const double t = 1.0 / (c + d * r);
  const double p = julia_div_component(a, b, c, d, r, t);
  const double q = julia_div_component(b, -a, c, d, r, t);
  return {p, swapped ? -q : q};
}
inline zc julia_divide(zc z, zc w) {
  double a = z.re, b = z.im, c = w.re, d = w.im;
  const double ab = std::fabs(a) >= std::fabs(b) ? std::fabs(a) : std::fabs(b);
  const double cd = std::fabs(c) >= std::fabs(d) ? std::fabs(c) : std::fabs(d);
  if (std::isinf(c) || std::isinf(d)) {
    // Rust's signum: NaN stays NaN, zero keeps its sign
    auto sgn = [](double x) { return (x == 0.0 || std::isnan(x)) ? x : std::copysign(1.0, x); };
    if (std::isfinite(a) && std::isfinite(b)) return {0.0 * sgn(a) * sgn(c), -0.0 * sgn(b) * sgn(d)};
    return {NAN, NAN};
  }
  const double halfov = 0.5 * DBL_MAX;
  const double twouneps = DBL_MIN * 2.0 / DBL_EPSILON;
  if (ab >= halfov || ab <= twouneps || cd >= halfov || cd <= twouneps) {
    const double bs = 2.0 / (DBL_EPSILON * DBL_EPSILON);
    double s = 1.0;
    if (ab >= halfov) {
      a *= 0.5;
      b *= 0.5;
      s *= 2.0;
    } else if (ab <= twouneps) {
      a *= bs;
      b *= bs;
      s /= bs;
    }
    if (cd >= halfov) {
      c *= 0.5;
      d *= 0.5;
      s *= 0.5;
    } else if (cd <= twouneps) {
      c *= bs;
      d *= bs;
      s *= bs;
    }
    const zc u = julia_div_unscaled(a, b, c, d);
    return zc{u.re * s, u.im * s};  // complex times real scalar
  }
  return julia_div_unscaled(a, b, c, d);
}

// order 0: zsktf2; order 1: zsktf2_c_compat.  Returns 0 or the 1-based row
// of the first zero pivot column.
inline int zsktf2(zc* a, int n, int* piv, int order) {
  int info = 0;
  for (int i = 0; i < n; ++i) piv[i] = i + 1;
  if (n < 2) return 0;
  auto A = [&](int i, int j) -> zc& { return a[i + (size_t)j * n]; };
  for (int k0 = n - 1; k0 >= 1; --k0) {
    const int kk0 = k0 - 1;
    int kp = 0;
    double colmax = 0.0;
    for (int j = 0; j <= kk0; ++j) {
      const double v = std::fabs(A(j, k0).re) + std::fabs(A(j, k0).im);
      if (j == 0 || v > colmax) {
        colmax = v;
        kp = j;
      }
    }
    if (colmax == 0.0) {
      if (info == 0) info = kk0 + 1;
      piv[kk0] = kk0 + 1;
      continue;
    }
    if (kp != kk0) {
      for (int j = 0; j < kp; ++j) std::swap(A(j, kk0), A(j, kp));
      for (int j = kp + 1; j < kk0; ++j) std::swap(A(j, kk0), A(kp, j));
      for (int j = k0; j < n; ++j) std::swap(A(kk0, j), A(kp, j));
      for (int j = kp; j < kk0; ++j) A(j, kk0) = zneg(A(j, kk0));
      for (int j = kp + 1; j < kk0; ++j) A(kp, j) = zneg(A(kp, j));
    }
    piv[kk0] = kp + 1;
    if (kk0 >= 1) {
      const zc pv = A(kk0, k0);
      zc alpha;
      if (order == 0) {
        alpha = julia_divide(zc{1.0, 0.0}, pv);
      } else {
        const double den = pv.re * pv.re + pv.im * pv.im;
        alpha = zc{pv.re / den, -pv.im / den};
      }
      for (int j = 0; j < kk0; ++j) {
        const zc t1 = zmul(alpha, A(j, kk0));
        const zc t2 = zmul(alpha, A(j, k0));
        for (int i = 0; i < j; ++i) {
          const zc p = zmul(A(i, k0), t1);
          const zc q = zmul(A(i, kk0), t2);
          A(i, j) = zadd(A(i, j), zsub(p, q));
        }
        A(j, j) = zc{0.0, 0.0};
      }
      for (int i = 0; i < kk0; ++i) {
        const zc v = A(i, k0);
        if (order == 0)
          A(i, k0) = zmul(v, alpha);
        else
          A(i, k0) = zc{std::fma(alpha.re, v.re, -(alpha.im * v.im)), std::fma(alpha.re, v.im, alpha.im * v.re)};
      }
    }
  }
  return info;
}

inline zc zutu2pfa(const zc* a, int n, const int* piv) {
  zc pf{1.0, 0.0};
  if (n == 0) return pf;
  for (int i = 0; i + 1 < n; i += 2) pf = zmul(pf, a[i + (size_t)(i + 1) * n]);
  bool neg = false;
  for (int k = 0; k < n; ++k)
    if (piv[k] != k + 1) neg = !neg;
  const zc sign = neg ? zc{0.0 - 1.0, 0.0 - 0.0} : zc{1.0, 0.0};
  return zmul(sign, pf);
}

// In place: on entry the LTL form from zsktf2, on exit A^{-1} (true inverse).
inline void zutu2inv(zc* a, int n, const int* piv) {
  std::vector<zc> m((size_t)n * n, zc{0.0, 0.0}), vt(n > 1 ? n - 1 : 1);
  auto A = [&](int i, int j) -> zc& { return a[i + (size_t)j * n]; };
  auto Mx = [&](int i, int j) -> zc& { return m[i + (size_t)j * n]; };
  for (int i = 0; i < n; ++i) Mx(i, i) = zc{1.0, 0.0};
  // trtri('U','U') on B(i,j) = A(i, j+1), size n-1 (backend.rs:231-251)
  for (int j = 1; j < n - 1; ++j)
    for (int i = 0; i < j; ++i) {
      zc acc = A(i, j + 1);
      for (int k = i + 1; k < j; ++k) acc = zadd(acc, zmul(A(i, k + 1), A(k, j + 1)));
      A(i, j + 1) = zneg(acc);
    }
  for (int jr = 0; jr + 2 < n; ++jr)
    for (int ir = 0; ir <= jr; ++ir) Mx(ir, jr + 1) = A(ir, jr + 2);
  for (int i = 0; i + 1 < n; ++i) vt[i] = zneg(A(i, i + 1));
  // Julia-style skew-tridiagonal solve T C = M, C overwrites A
  {
    const zc r1 = julia_reciprocal(zneg(vt[0]));
    for (int j = 0; j < n; ++j) A(1, j) = zmul(Mx(0, j), r1);
    for (int ij = 2; ij < n; ij += 2) {
      const zc r = julia_reciprocal(zneg(vt[ij]));
      const zc v = vt[ij - 1];
      for (int j = 0; j < n; ++j) A(ij + 1, j) = zmul(zsub(Mx(ij, j), zmul(A(ij - 1, j), v)), r);
    }
    const zc rn = julia_reciprocal(vt[n - 2]);
    for (int j = 0; j < n; ++j) A(n - 2, j) = zmul(Mx(n - 1, j), rn);
    if (n >= 4)
      for (int i = n - 3; i >= 1; i -= 2) {
        const zc r = julia_reciprocal(vt[i - 1]);
        const zc v = vt[i];
        for (int j = 0; j < n; ++j) A(i - 1, j) = zmul(zadd(Mx(i, j), zmul(A(i + 1, j), v)), r);
      }
  }
  const bool panel = n > 64;
  auto product = [&](int i, int j) {
    zc acc = A(i, j);
    for (int k = 0; k < i; ++k) acc = zadd(acc, zmul(Mx(k, i), A(k, j)));
    A(i, j) = acc;
  };
  if (panel) {
    // upper triangle by column panels (row order within a panel descending;
    // the per-element sums do not depend on the panel width), then skew fill
    for (int i = n - 1; i >= 0; --i)
      for (int j = i; j < n; ++j) product(i, j);
    for (int j = 0; j < n; ++j) {
      A(j, j) = zc{0.0, 0.0};
      for (int i = j + 1; i < n; ++i) A(i, j) = zneg(A(j, i));
    }
  }
  for (int j = 0; j < n; ++j) {
    const int t = piv[j] - 1;
    if (t != j)
      for (int r = 0; r < n; ++r) std::swap(A(r, j), A(r, t));
  }
  if (!panel)
    for (int i = n - 1; i >= 0; --i)
      for (int j = 0; j < n; ++j) product(i, j);
  for (int j = 0; j < n; ++j)
    for (int i = 0; i < n; ++i) {
      const int t = piv[i] - 1;
      if (t != i) std::swap(A(i, j), A(t, j));
    }
}

}  // namespace oracle
