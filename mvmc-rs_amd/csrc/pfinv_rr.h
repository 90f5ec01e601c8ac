// Register-resident skew-symmetric Pfaffian + inverse of one plane per
// workgroup, for n <= RR_NMAX: the C PfaPack order of dsktf2
// (crates/pfapack/src/ltl.rs:61-191, :683-705), utu2pfa (utu2.rs:52-93)
// and utu2inv (utu2.rs:338-513 with the scalar trtri / trmm forms of
// backend.rs:233-300), every element receiving exactly the operation
// sequence of that order (bit-identical to the sequential loops; built with
// -ffp-contract=off).
//
// Layout: T x T threads, thread (a, b) = (tid % T, tid / T) owns the
// elements (i, j) = (a + T ri, b + T cj) of the plane.  Panel form (n > 64,
// T = 32): only the strict upper triangle is ever needed (the elimination
// reads it, the inverse's panel path forms the upper triangle and fills the
// lower by skew symmetry), and (i, j) with cj > ri is always upper, cj < ri
// never, cj == ri iff b > a: a thread keeps the R(R+1)/2 slots cj >= ri in
// registers.  Full form (n <= 64, T = 16, the whole plane held on chip):
// utu2inv's small-matrix path multiplies the whole solved matrix, so a
// thread keeps all R^2 slots (the elimination still uses the upper ones).
// Steps exchange rows and columns through LDS vectors; the trtri result M
// and the solved C go to a per-workgroup global scratch (row-major) from
// which the product reads one row of each per step.  The matrix is read once
// (gathered from the source) and the inverse written once.
#pragma once
#include "common.h"

// needs Carver and rnd16 (ensemble.hip) before inclusion

namespace mvmc_hip {

constexpr int RR_T = 32;      // threads per side of the panel form (1024 per workgroup)
constexpr int RR_TS = 16;     // threads per side of the full form (256 per workgroup)
constexpr int RR_NMAX = 224;  // largest n (R = 7 slots per side)
constexpr int RR_NFULL = 64;  // largest n of the full form (utu2inv's small-matrix path)

// LDS of one plane: 11 vectors of n scalars (real or complex), 2 of n
// ints, scalars.
template <class S>
struct RrLdsT {
  S *v0, *v1, *v2, *v3, *buf0, *buf1, *buf2, *buf3;
  S* sup;          // the superdiagonal A(i, i+1) after the LTL
  S *rneg, *rpos;  // complex solve: reciprocals of -vt[i] and vt[i]
  int *piv, *perm;
  int* si;     // [16] ints: 2 info, 4..7 pivot results by step parity
  S* sa;       // [4] scalars: 0, 1 alpha by step parity, 2 the Pfaffian
  double* sd;  // [32] doubles: 16.. max |a|^2 per wave
};
using RrLds = RrLdsT<double>;

template <class S = double>
__host__ __device__ inline size_t rr_lds_bytes(int n) {
  return 11 * rnd16((size_t)n * sizeof(S)) + 2 * rnd16((size_t)n * 4) + 16 * 4 + rnd16(4 * sizeof(S)) + 32 * 8;
}

template <class S = double, class Carver>
__device__ inline RrLdsT<S> carve_rr(Carver& c, int n) {
  RrLdsT<S> L;
  L.v0 = c.template take<S>(n);
  L.v1 = c.template take<S>(n);
  L.v2 = c.template take<S>(n);
  L.v3 = c.template take<S>(n);
  L.buf0 = c.template take<S>(n);
  L.buf1 = c.template take<S>(n);
  L.buf2 = c.template take<S>(n);
  L.buf3 = c.template take<S>(n);
  L.sup = c.template take<S>(n);
  L.rneg = c.template take<S>(n);
  L.rpos = c.template take<S>(n);
  L.piv = c.template take<int>(n);
  L.perm = c.template take<int>(n);
  L.si = c.template take<int>(16);
  L.sa = c.template take<S>(4);
  L.sd = c.template take<double>(32);
  return L;
}

// Slot of block (ri, cj): packed upper block triangle (cj >= ri) in the
// panel form, row-major R x R in the full form.
template <int R, bool FULL>
__host__ __device__ constexpr int rr_slot(int ri, int cj) {
  return FULL ? ri * R + cj : ri * R - ri * (ri - 1) / 2 + (cj - ri);
}

// C IDAMAX order of the pivot search: first index of the largest |x|; a NaN
// wins only at index 0 (the sequential scan starts from it).
__device__ inline double rr_key(double v, int i) {
  if (v != v) return i == 0 ? INFINITY : -1.0;
  return v;
}

// Arithmetic of the real planes (dsktf2 / utu2 with real operands): every
// operation is the one the sequential loops perform.
struct RrReal {
  using S = double;
  static constexpr bool kRecip = false;  // the solve divides by vt
  static __device__ __forceinline__ S zero() { return 0.0; }
  static __device__ __forceinline__ S one() { return 1.0; }
  static __device__ __forceinline__ S neg(S x) { return -x; }
  static __device__ __forceinline__ S add(S x, S y) { return x + y; }
  static __device__ __forceinline__ S sub(S x, S y) { return x - y; }
  static __device__ __forceinline__ S mul(S x, S y) { return x * y; }
  static __device__ __forceinline__ double mag2(S x) { return x * x; }
  static __device__ __forceinline__ double key(S x, int i) { return rr_key(fabs(x), i); }
  static __device__ __forceinline__ S alpha(S pivot) { return 1.0 / pivot; }
  // rank-2 update of one element: (c + x_i t1) - y_i t2
  static __device__ __forceinline__ S update(S c, S xi, S yi, S t1, S t2) { return (c + xi * t1) - yi * t2; }
  static __device__ __forceinline__ S scale(S x, S alpha) { return x * alpha; }
  static __device__ __forceinline__ S sign(bool neg) { return neg ? (0.0 - 1.0) : 1.0; }
  static __device__ __forceinline__ bool finite(S x) { return isfinite(x); }
  // solve: num / -vt (forward), num / vt (backward)
  static __device__ __forceinline__ S div_negvt(S num, S vt, S) { return num / -vt; }
  static __device__ __forceinline__ S div_vt(S num, S vt, S) { return num / vt; }
  static __device__ __forceinline__ S recip(S x) { return 1.0 / x; }
};

// One plane.  src(i, j) gives A(i, j) (any i, j < n).  On exit: info (0 ok,
// k > 0 zero pivot at 1-based row k, -1 non-finite Pfaffian), pf (the
// Pfaffian; on a zero pivot not formed), and for info == 0 the plane `out`
// (column-major, element (r, c) at r + c n) holds sign * A^{-1}.  allzero
// (when check_zero) reports max |a|^2 < 1e-28 over the whole plane, in which
// case nothing else is done.  gs: 2 n^2 doubles of global scratch.  Every
// thread of the T x T workgroup (blockDim.x == T * T) calls it.  FULL: the
// full form (n <= 64), else the panel form (n > 64).
template <int T, int R, bool FULL, class Ar = RrReal, class Src>
__device__ __forceinline__ void rr_pfinv(Src src, int n, typename Ar::S* out, double sign, typename Ar::S* gs,
                                         RrLdsT<typename Ar::S>& L, typename Ar::S& pf, int& info, bool check_zero,
                                         bool& allzero) {
  using S = typename Ar::S;
  constexpr int NS = FULL ? R * R : R * (R + 1) / 2;
  const int tid = threadIdx.x, NT = T * T;
  const int a0 = tid % T, b0 = tid / T;
  // per-step copies of (a, b) the compiler cannot prove loop-invariant: the
  // per-slot row/column indices are recomputed in each step instead of being
  // kept live across the whole elimination (registers hold the slots)
  int a = a0, b = b0;
#define RR_FRESH() \
  do {             \
    a = a0;        \
    b = b0;        \
    asm volatile("" : "+v"(a), "+v"(b)); \
  } while (0)
  S c[NS];
#ifdef MVMC_PHASE_CLOCKS
  long long rr_t0 = wall_clock64(), rr_ph[8] = {0, 0, 0, 0, 0, 0, 0, 0};
  long long rr_s0 = 0, rr_sub[4] = {0, 0, 0, 0};
#define RR_SUB(k)                        \
  do {                                   \
    const long long t_ = wall_clock64(); \
    rr_sub[k] += t_ - rr_s0;             \
    rr_s0 = t_;                          \
  } while (0)
#define RR_PHASE(k)                           \
  do {                                        \
    const long long t_ = wall_clock64();      \
    rr_ph[k] += t_ - rr_t0;                   \
    rr_t0 = t_;                               \
  } while (0)
#else
#define RR_PHASE(k) \
  do {              \
  } while (0)
#define RR_SUB(k) \
  do {            \
  } while (0)
#endif
  // ---- gather (and the all-zero guard over every element, pfaffian.rs:884-890) ----
  double mx = 0.0;
#pragma unroll
  for (int ri = 0; ri < R; ++ri)
#pragma unroll
    for (int cj = ri; cj < R; ++cj) {
      const int i = a + T * ri, j = b + T * cj;
      S x = Ar::zero();
      if (i < n && j < n && i <= j) {
        x = src(i, j);
        if (check_zero) {
          const S y = src(j, i);
          mx = fmax(mx, fmax(Ar::mag2(x), Ar::mag2(y)));
        }
      }
      c[rr_slot<R, FULL>(ri, cj)] = x;
    }
  if (FULL) {
#pragma unroll
    for (int ri = 1; ri < R; ++ri)
#pragma unroll
      for (int cj = 0; cj < ri; ++cj) c[rr_slot<R, FULL>(ri, cj)] = Ar::zero();
  }
  for (int i = tid; i < n; i += NT) L.piv[i] = i + 1;
  if (tid == 0) L.si[2] = 0;
  if (check_zero) {
    for (int off = 32; off >= 1; off >>= 1) mx = fmax(mx, __shfl_xor(mx, off, 64));
    if ((tid & 63) == 0) L.sd[16 + (tid >> 6)] = mx;
    __syncthreads();
    double m = 0.0;
    for (int k = 0; k < NT / 64; ++k) m = fmax(m, L.sd[16 + k]);
    allzero = m < 1.0e-28;
    __syncthreads();
    if (allzero) return;
  } else {
    allzero = false;
    __syncthreads();
  }

  // Slots of one block column / block row: f(global row or column index,
  // slot) for the lanes that own that column (b == col % T) or row; the block
  // index is uniform, so the other blocks' code is branched over.
  auto for_col = [&](int col, auto f) {
    if (b != col % T) return;
    const int kc = col / T;
#pragma unroll
    for (int cj = 0; cj < R; ++cj)
      if (cj == kc) {
#pragma unroll
        for (int ri = 0; ri <= cj; ++ri) f(a + T * ri, c[rr_slot<R, FULL>(ri, cj)]);
      }
  };
  auto for_row = [&](int row, auto f) {
    if (a != row % T) return;
    const int kr = row / T;
#pragma unroll
    for (int ri = 0; ri < R; ++ri)
      if (ri == kr) {
#pragma unroll
        for (int cj = ri; cj < R; ++cj) f(b + T * cj, c[rr_slot<R, FULL>(ri, cj)]);
      }
  };

  RR_PHASE(0);
  // ---- dsktf2: k0 = n-1 .. 1 ----
  // The pivot column of step k0 - 1 is column kk0 as swapped in step k0,
  // which the rank-2 update of step k0 does not touch: its search runs
  // before that update (look-ahead), so a step is swap staging | swap
  // write-back + staging of x, y, alpha + next pivot | update.  Pivot
  // results, x / y and alpha alternate between two LDS sets by step parity.
  auto pivot_search = [&](int col, int par) {
    if (b != col % T) return;
    const int kk = col - 1;
    double best = -2.0;
    int bi = 0x7fffffff;
    const int kc = col / T;
#pragma unroll
    for (int cj = 0; cj < R; ++cj)
      if (cj == kc) {
#pragma unroll
        for (int ri = 0; ri <= cj; ++ri) {
          const int i = a + T * ri;
          if (i <= kk) {
            const double key = Ar::key(c[rr_slot<R, FULL>(ri, cj)], i);
            if (key > best) { best = key; bi = i; }
          }
        }
      }
#pragma unroll
    for (int off = T / 2; off >= 1; off >>= 1) {
      const double ob = __shfl_xor(best, off, 64);
      const int oi = __shfl_xor(bi, off, 64);
      if (ob > best || (ob == best && oi < bi)) { best = ob; bi = oi; }
    }
    if (a == 0) {
      L.si[4 + 2 * par] = bi;
      L.si[5 + 2 * par] = best == 0.0 ? 1 : 0;
    }
  };
  pivot_search(n - 1, (n - 1) & 1);
  __syncthreads();
  for (int k0 = n - 1; k0 >= 1; --k0) {
    const int kk0 = k0 - 1;
    const int par = k0 & 1;
    RR_FRESH();
#ifdef MVMC_PHASE_CLOCKS
    rr_s0 = wall_clock64();
#endif
    const int kp = L.si[4 + 2 * par];
    if (L.si[5 + 2 * par]) {
      // zero column: INFO records the first, the step is skipped
      if (tid == 0) {
        if (L.si[2] == 0) L.si[2] = kk0 + 1;
        L.piv[kk0] = kk0 + 1;
      }
      if (kk0 >= 1) pivot_search(kk0, par ^ 1);
      __syncthreads();
      continue;
    }
    if (kp != kk0) {
      // interchange rows/columns kk0 and kp in the upper triangle with the
      // sign twiddles: stage column kk0 (rows < kk0), column kp (rows < kp),
      // row kp (columns > kp), row kk0 (columns >= k0)
      for_col(kk0, [&](int i, S& x) { if (i < kk0) L.v0[i] = x; });
      for_col(kp, [&](int i, S& x) { if (i < kp) L.v1[i] = x; });
      for_row(kp, [&](int j, S& x) { if (j > kp && j < n) L.v2[j] = x; });
      for_row(kk0, [&](int j, S& x) { if (j >= k0 && j < n) L.v3[j] = x; });
      __syncthreads();
      RR_SUB(1);
      // write-back; the four groups target disjoint elements
      for_col(kk0, [&](int i, S& x) {
        if (i < kp) x = L.v1[i];
        else if (i == kp) x = Ar::neg(L.v0[kp]);
        else if (i < kk0) x = Ar::neg(L.v2[i]);
      });
      for_col(kp, [&](int i, S& x) { if (i < kp) x = L.v0[i]; });
      for_row(kp, [&](int j, S& x) {
        if (j > kp && j < kk0) x = Ar::neg(L.v0[j]);
        else if (j >= k0 && j < n) x = L.v3[j];
      });
      for_row(kk0, [&](int j, S& x) { if (j >= k0 && j < n) x = L.v2[j]; });
    }
    if (tid == 0) L.piv[kk0] = kp + 1;
    if (kk0 >= 1) {
      // rank-2 update of the leading kk0 x kk0 block with x = column k0,
      // y = column kk0 (as swapped), alpha = 1 / A(kk0, k0)
      S* xs = par ? L.buf0 : L.buf2;
      S* ys = par ? L.buf1 : L.buf3;
      for_col(k0, [&](int i, S& x) {
        if (i < kk0) xs[i] = x;
        else if (i == kk0) L.sa[par] = Ar::alpha(x);
      });
      for_col(kk0, [&](int i, S& x) { if (i < kk0) ys[i] = x; });
      pivot_search(kk0, par ^ 1);
      __syncthreads();
      RR_SUB(2);
      const S alpha = L.sa[par];
#pragma unroll
      for (int cj = 0; cj < R; ++cj)
        if (T * cj < kk0) {
          const int j = b + T * cj;
          if (j < kk0) {
            const S t1 = Ar::mul(alpha, ys[j]);
            const S t2 = Ar::mul(alpha, xs[j]);
#pragma unroll
            for (int ri = 0; ri <= cj; ++ri) {
              const int i = a + T * ri;
              S& x = c[rr_slot<R, FULL>(ri, cj)];
              if (i < j) x = Ar::update(x, xs[i], ys[i], t1, t2);
              else if (i == j) x = Ar::zero();
            }
          }
        }
      // scale column k0 (its pre-scale values were staged in xs)
      for_col(k0, [&](int i, S& x) { if (i < kk0) x = Ar::scale(x, alpha); });
      RR_SUB(3);
    }
  }
  __syncthreads();
  RR_PHASE(1);
  info = L.si[2];
  if (info != 0) {
    pf = Ar::zero();
    return;
  }
  // ---- utu2pfa: product of A(i, i+1), i even, sign by the count of moves ----
  // (the superdiagonal also gives vt = -A(i, i+1) of the solve: sup)
#pragma unroll
  for (int ri = 0; ri < R; ++ri)
#pragma unroll
    for (int cj = ri; cj < R; ++cj) {
      const int i = a + T * ri, j = b + T * cj;
      if (j == i + 1 && j < n) L.sup[i] = c[rr_slot<R, FULL>(ri, cj)];
    }
  __syncthreads();
  if (tid == 0) {
    S p = Ar::one();
    for (int i = 0; i + 1 < n; i += 2) p = Ar::mul(p, L.sup[i]);
    bool neg = false;
    for (int k = 0; k < n; ++k)
      if (L.piv[k] != k + 1) neg = !neg;
    L.sa[2] = Ar::mul(Ar::sign(neg), p);
  }
  if (Ar::kRecip)
    for (int i = tid; i + 1 < n; i += NT) {
      const S vt = Ar::neg(L.sup[i]);
      L.rneg[i] = Ar::recip(Ar::neg(vt));
      L.rpos[i] = Ar::recip(vt);
    }
  __syncthreads();
  pf = L.sa[2];
  if (!Ar::finite(pf)) {
    info = -1;
    return;
  }

  RR_PHASE(2);
  // ---- utu2inv (panel path for n > 64; the full form differs from the
  // solve on) ----
  // trtri('U','U') on B(i, j) = A(i, j+1), 0 <= i < j <= n-2: at step k the
  // column k of the inverse is final (Bn(i, k) = -acc(i, k), i < k) and
  // every acc(i, j), i < k < j, adds Bn(i, k) B(k, j) with row k of B still
  // as factored (no earlier step writes it): each element takes its k in
  // ascending order, as the row-chain loops do.
  for (int k = 1; k <= n - 2; ++k) {
    RR_FRESH();
    S* U = (k & 1) ? L.v0 : L.v2;   // column k of the inverse, rows < k
    S* Bk = (k & 1) ? L.v1 : L.v3;  // row k of B (columns > k)
    for_col(k + 1, [&](int i, S& x) {
      if (i < k) {
        x = Ar::neg(x);
        U[i] = x;
      }
    });
    for_row(k, [&](int jj, S& x) { if (jj >= k + 2 && jj < n) Bk[jj - 1] = x; });
    __syncthreads();
    if (k <= n - 3) {
#pragma unroll
      for (int ri = 0; ri < R; ++ri)
        if (T * ri < k) {
#pragma unroll
          for (int cj = ri; cj < R; ++cj)
            if (T * cj + T > k + 2) {
              const int i = a + T * ri, jj = b + T * cj;
              if (i < k && jj >= k + 2 && jj < n) {
                S& x = c[rr_slot<R, FULL>(ri, cj)];
                x = Ar::add(x, Ar::mul(U[i], Bk[jj - 1]));
              }
            }
        }
    }
  }
  __syncthreads();
  RR_PHASE(3);
  // M (unit upper, M(i, j) = Bn(i, j) for i < j <= n-2, M(i, n-1) = 0 for
  // i < n-1) to gs row-major: gm[i * n + j]
  S* gm = gs;
  S* gc = gs + (size_t)n * n;
  for (int e = tid; e < n * n; e += NT) gm[e] = (e / n == e % n) ? Ar::one() : Ar::zero();
  __syncthreads();
#pragma unroll
  for (int ri = 0; ri < R; ++ri)
#pragma unroll
    for (int cj = ri; cj < R; ++cj) {
      const int i = a + T * ri, jj = b + T * cj;
      if (i + 1 < jj && jj < n) gm[(size_t)i * n + (jj - 1)] = c[rr_slot<R, FULL>(ri, cj)];
    }
  __syncthreads();
  RR_PHASE(4);
  // skew-tridiagonal solve T C = M, all columns at once, row by row through
  // LDS (vt[i] = -A(i, i+1) = -sup[i]); every row of C goes to gc row-major
  // and into the registers (panel form: the strict upper part).
  auto keep_row = [&](int r, const S* row) {
#pragma unroll
    for (int ri = 0; ri < R; ++ri)
      if (a + T * ri == r) {
#pragma unroll
        for (int cj = FULL ? 0 : ri; cj < R; ++cj) {
          const int j = b + T * cj;
          if ((FULL || r < j) && j < n) c[rr_slot<R, FULL>(ri, cj)] = row[j];
        }
      }
  };
  {
    // thread j < n carries column j; the rows of M it reads do not depend on
    // the chain, so they are loaded two steps ahead
    const int j = tid;
    const bool col = j < n;
    // forward: c[1] = b[0] / -vt[0]; c[i+1] = (b[i] - c[i-1] vt[i-1]) / -vt[i]
    S m0 = col ? gm[j] : Ar::zero();
    S m1 = (col && 2 < n) ? gm[(size_t)2 * n + j] : Ar::zero();
    S m2 = (col && 4 < n) ? gm[(size_t)4 * n + j] : Ar::zero();
    if (col) {
      const S vt0 = Ar::neg(L.sup[0]);
      const S x = Ar::div_negvt(m0, vt0, L.rneg[0]);
      L.buf0[j] = x;
      gc[(size_t)1 * n + j] = x;
    }
    __syncthreads();
    keep_row(1, L.buf0);
    S* prev = L.buf0;
    S* next = L.buf1;
    for (int i = 2; i < n; i += 2) {
      const S mi = m1;
      m1 = m2;
      m2 = (col && i + 4 < n) ? gm[(size_t)(i + 4) * n + j] : Ar::zero();
      const S vtm = Ar::neg(L.sup[i - 1]), vti = Ar::neg(L.sup[i]);
      if (col) {
        const S x = Ar::div_negvt(Ar::sub(mi, Ar::mul(prev[j], vtm)), vti, L.rneg[i]);
        next[j] = x;
        gc[(size_t)(i + 1) * n + j] = x;
      }
      __syncthreads();
      RR_FRESH();
      keep_row(i + 1, next);
      S* t = prev; prev = next; next = t;
    }
    __syncthreads();
    // backward: c[n-2] = b[n-1] / vt[n-2]; c[i-1] = (b[i] + c[i+1] vt[i]) / vt[i-1]
    m0 = col ? gm[(size_t)(n - 1) * n + j] : Ar::zero();
    m1 = (col && n - 3 >= 1) ? gm[(size_t)(n - 3) * n + j] : Ar::zero();
    m2 = (col && n - 5 >= 1) ? gm[(size_t)(n - 5) * n + j] : Ar::zero();
    if (col) {
      const S vtl = Ar::neg(L.sup[n - 2]);
      const S x = Ar::div_vt(m0, vtl, L.rpos[n - 2]);
      prev[j] = x;
      gc[(size_t)(n - 2) * n + j] = x;
    }
    __syncthreads();
    keep_row(n - 2, prev);
    for (int i = n - 3; i >= 1; i -= 2) {
      const S mi = m1;
      m1 = m2;
      m2 = (col && i - 4 >= 1) ? gm[(size_t)(i - 4) * n + j] : Ar::zero();
      const S vti = Ar::neg(L.sup[i]), vtm = Ar::neg(L.sup[i - 1]);
      if (col) {
        const S x = Ar::div_vt(Ar::add(mi, Ar::mul(prev[j], vti)), vtm, L.rpos[i - 1]);
        next[j] = x;
        gc[(size_t)(i - 1) * n + j] = x;
      }
      __syncthreads();
      RR_FRESH();
      keep_row(i - 1, next);
      S* t = prev; prev = next; next = t;
    }
  }
  __syncthreads();
  RR_PHASE(5);
  // product C'(i, j) = C(i, j) + sum_{k < i} M(k, i) C(k, j) (k ascending,
  // C as solved): right-looking over k with rows k of M and C from the
  // scratch.  Panel form: the upper part (the lower follows by skew
  // symmetry).  Full form: every element (the small-matrix path multiplies
  // the column-permuted C; a column's result depends on that column alone,
  // so the permutation moves to the output gather).
  {
    // rows k of M and C, loaded two steps ahead by thread j < n
    const int j = tid;
    const bool col = j < n;
    S pm0 = col ? gm[j] : Ar::zero(), pc0 = col ? gc[j] : Ar::zero();
    S pm1 = (col && 1 < n) ? gm[(size_t)n + j] : Ar::zero(), pc1 = (col && 1 < n) ? gc[(size_t)n + j] : Ar::zero();
    for (int k = 0; k + 2 < n; ++k) {
      RR_FRESH();
      S* Mr = (k & 1) ? L.v0 : L.v2;
      S* Cr = (k & 1) ? L.v1 : L.v3;
      if (col) {
        Mr[j] = pm0;
        Cr[j] = pc0;
      }
      pm0 = pm1;
      pc0 = pc1;
      if (col && k + 2 < n) {
        pm1 = gm[(size_t)(k + 2) * n + j];
        pc1 = gc[(size_t)(k + 2) * n + j];
      }
      __syncthreads();
#pragma unroll
      for (int ri = 0; ri < R; ++ri)
        if (T * ri + T > k + 1) {
#pragma unroll
          for (int cj = FULL ? 0 : ri; cj < R; ++cj) {
            const int i = a + T * ri, jj = b + T * cj;
            if (k < i && (FULL || i < jj) && jj < n && i < n) {
              S& x = c[rr_slot<R, FULL>(ri, cj)];
              x = Ar::add(x, Ar::mul(Mr[i], Cr[jj]));
            }
          }
        }
    }
  }
  __syncthreads();
  RR_PHASE(6);
  // permutation of the column swaps (and the same row swaps): position r
  // holds index perm[r] after the swap sequence of utu2inv
  if (tid == 0) {
    for (int r = 0; r < n; ++r) L.perm[r] = r;
    for (int j = 0; j < n; ++j) {
      const int t = L.piv[j] - 1;
      if (t != j) { const int x = L.perm[j]; L.perm[j] = L.perm[t]; L.perm[t] = x; }
    }
  }
  // C' to the scratch (row-major: panel form the upper part, full form
  // all), then the output gather: out(r, c) = sign * F(perm[r], perm[c]),
  // F = C' (full form) or the skew completion of C' (panel form, 0 on the
  // diagonal)
#pragma unroll
  for (int ri = 0; ri < R; ++ri)
#pragma unroll
    for (int cj = FULL ? 0 : ri; cj < R; ++cj) {
      const int i = a + T * ri, j = b + T * cj;
      if ((FULL || i < j) && i < n && j < n) gc[(size_t)i * n + j] = c[rr_slot<R, FULL>(ri, cj)];
    }
  __syncthreads();
  for (int e = tid; e < n * n; e += NT) {
    const int r = e % n, cc = e / n;
    const int u = L.perm[r], v = L.perm[cc];
    S x = Ar::zero();
    if (FULL) x = gc[(size_t)u * n + v];
    else if (u < v) x = gc[(size_t)u * n + v];
    else if (u > v) x = Ar::neg(gc[(size_t)v * n + u]);
    out[e] = sign != 1.0 ? Ar::neg(x) : x;
  }
  __syncthreads();
  RR_PHASE(7);
#ifdef MVMC_PHASE_CLOCKS
  if (tid == 0 && blockIdx.x == 0)
    printf("rr n=%d phases(10ns): gather %lld ltl %lld pf %lld trtri %lld mwrite %lld solve %lld product %lld out %lld"
           " | ltl: pivot %lld swapstage %lld wb+stage %lld update %lld\n",
           n, rr_ph[0], rr_ph[1], rr_ph[2], rr_ph[3], rr_ph[4], rr_ph[5], rr_ph[6], rr_ph[7], rr_sub[0], rr_sub[1],
           rr_sub[2], rr_sub[3]);
#endif
#undef RR_PHASE
#undef RR_SUB
#undef RR_FRESH
}

}  // namespace mvmc_hip
