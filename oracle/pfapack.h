// ORACLE (test infrastructure only): real skew-symmetric LTL^T factorisation,
// Pfaffian and inverse in the C PfaPack operation order.
//   dsktf2     crates/pfapack/src/ltl.rs:34-191 (+ :683-705, C DSKR2 order
//              (c + x*t1) - y*t2, strict upper triangle, zero diagonal)
//   utu2pfa    crates/pfapack/src/utu2.rs:52-93
//   utu2inv    crates/pfapack/src/utu2.rs:338-513 with the scalar forms of
//              backend.rs:233-300 (trtri 'U','U'; trmm 'L','U','T','U';
//              64-column upper panel product for n > 64, then skew fill).
// Matrices are column-major n x n with lda = n.  Pivots are 1-based.
// Pinned by tests/test_oracle_pfaffian.py: closed forms for n = 2 and 4,
// brute-force Pfaffian expansion for n <= 10, Pf^2 = det (LU), inverse
// residual, on the reference's committed input matrices
// (tests/golden/pfapack_inputs.npz).  The reference links OpenBLAS dtrtri /
// dtrmm here; those BLAS orders are not bit-reproducible, so the inverse is
// pinned by residuals only.
#pragma once
#include <cmath>
#include <utility>
#include <vector>

namespace oracle {

// Returns 0 on success or the 1-based row of the first zero pivot column.
inline int dsktf2(double* a, int n, int* piv) {
  int info = 0;
  for (int i = 0; i < n; ++i) piv[i] = i + 1;
  if (n < 2) return 0;
  for (int k0 = n - 1; k0 >= 1; --k0) {
    const int kk0 = k0 - 1;
    int kp = 0;
    double colmax = 0.0;
    for (int j = 0; j <= kk0; ++j) {
      double v = std::fabs(a[j + (size_t)k0 * n]);
      if (j == 0 || v > colmax) { colmax = v; kp = j; }
    }
    if (colmax == 0.0) {
      if (info == 0) info = kk0 + 1;
      piv[kk0] = kk0 + 1;
      continue;
    }
    if (kp != kk0) {
      for (int j = 0; j < kp; ++j) std::swap(a[j + (size_t)kk0 * n], a[j + (size_t)kp * n]);
      for (int j = kp + 1; j < kk0; ++j) std::swap(a[j + (size_t)kk0 * n], a[kp + (size_t)j * n]);
      for (int j = k0; j < n; ++j) std::swap(a[kk0 + (size_t)j * n], a[kp + (size_t)j * n]);
      for (int j = kp; j < kk0; ++j) a[j + (size_t)kk0 * n] = -a[j + (size_t)kk0 * n];
      for (int j = kp + 1; j < kk0; ++j) a[kp + (size_t)j * n] = -a[kp + (size_t)j * n];
    }
    piv[kk0] = kp + 1;
    if (kk0 >= 1) {
      const double alpha = 1.0 / a[kk0 + (size_t)k0 * n];
      const double* x = a + (size_t)k0 * n;   // column k0
      const double* y = a + (size_t)kk0 * n;  // column kk0
      for (int j = 0; j < kk0; ++j) {
        const double t1 = alpha * y[j];
        const double t2 = alpha * x[j];
        double* c = a + (size_t)j * n;
        for (int i = 0; i < j; ++i) c[i] = (c[i] + x[i] * t1) - y[i] * t2;
        c[j] = 0.0;
      }
      double* col = a + (size_t)k0 * n;
      for (int i = 0; i < kk0; ++i) col[i] *= alpha;
    }
  }
  return info;
}

inline double utu2pfa(const double* a, int n, const int* piv) {
  if (n == 0) return 1.0;
  double pf = 1.0;
  for (int i = 0; i + 1 < n; i += 2) pf *= a[i + (size_t)(i + 1) * n];
  bool neg = false;
  for (int k = 0; k < n; ++k)
    if (piv[k] != k + 1) neg = !neg;
  const double sign = neg ? (0.0 - 1.0) : 1.0;
  return sign * pf;
}

// In place: on entry the LTL form from dsktf2, on exit A^{-1} (true inverse).
inline void utu2inv(double* a, int n, const int* piv) {
  std::vector<double> m((size_t)n * n, 0.0), vt(n > 1 ? n - 1 : 1);
  for (int i = 0; i < n; ++i) m[i + (size_t)i * n] = 1.0;
  // trtri('U','U') on the (n-1)x(n-1) block B(i,j) = A(i, j+1).
  if (n - 1 >= 2) {
    auto B = [&](int i, int j) -> double& { return a[i + (size_t)(j + 1) * n]; };
    for (int j = 1; j < n - 1; ++j)
      for (int i = 0; i < j; ++i) {
        double acc = B(i, j);
        for (int k = i + 1; k < j; ++k) acc += B(i, k) * B(k, j);
        B(i, j) = -acc;
      }
  }
  for (int jr = 0; jr + 2 < n; ++jr)
    for (int ir = 0; ir <= jr; ++ir) m[ir + (size_t)(jr + 1) * n] = a[ir + (size_t)(jr + 2) * n];
  for (int i = 0; i + 1 < n; ++i) vt[i] = -a[i + (size_t)(i + 1) * n];
  // Skew-tridiagonal solve T C = M (C overwrites A), direct divisions.
  for (int j = 0; j < n; ++j) {
    const double* b = m.data() + (size_t)j * n;
    double* c = a + (size_t)j * n;
    c[1] = b[0] / -vt[0];
    for (int i = 2; i < n; i += 2) c[i + 1] = (b[i] - c[i - 1] * vt[i - 1]) / -vt[i];
    c[n - 2] = b[n - 1] / vt[n - 2];
    for (int i = n - 3; i >= 1; i -= 2) c[i - 1] = (b[i] + c[i + 1] * vt[i]) / vt[i - 1];
  }
  const bool panel = n > 64;
  if (panel) {
    const int P = 64;
    for (int pc = 0; pc < n; pc += P) {
      const int nloc = (P < n - pc) ? P : n - pc;
      const int rows = pc + nloc;
      for (int i = rows - 1; i >= 0; --i)
        for (int j = pc; j < pc + nloc; ++j) {
          double acc = a[i + (size_t)j * n];
          for (int k = 0; k < i; ++k) acc += m[k + (size_t)i * n] * a[k + (size_t)j * n];
          a[i + (size_t)j * n] = acc;
        }
    }
    for (int j = 0; j < n; ++j) {
      a[j + (size_t)j * n] = 0.0;
      for (int i = j + 1; i < n; ++i) a[i + (size_t)j * n] = -a[j + (size_t)i * n];
    }
  }
  for (int j = 0; j < n; ++j) {
    int t = piv[j] - 1;
    if (t != j)
      for (int r = 0; r < n; ++r) std::swap(a[r + (size_t)j * n], a[r + (size_t)t * n]);
  }
  if (!panel) {
    for (int i = n - 1; i >= 0; --i)
      for (int j = 0; j < n; ++j) {
        double acc = a[i + (size_t)j * n];
        for (int k = 0; k < i; ++k) acc += m[k + (size_t)i * n] * a[k + (size_t)j * n];
        a[i + (size_t)j * n] = acc;
      }
  }
  bool ident = true;
  for (int i = 0; i < n; ++i) ident = ident && (piv[i] - 1 == i);
  if (!ident)
    for (int j = 0; j < n; ++j) {
      double* col = a + (size_t)j * n;
      for (int i = 0; i < n; ++i) {
        int t = piv[i] - 1;
        if (t != i) std::swap(col[i], col[t]);
      }
    }
}

}  // namespace oracle
