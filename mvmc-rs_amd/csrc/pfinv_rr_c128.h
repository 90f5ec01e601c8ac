// Register-resident batched Pfaffian + inverse of complex skew-symmetric
// planes: the kernel of pfinv_rr.h (one plane per workgroup held in
// registers; full form n <= 64, panel form n <= RR_C_NMAX) with the complex
// arithmetic of the reference's scalar pfapack build, element by element as
// wg_pfinv_c128 (pfinv_c128_dev.h) performs it:
//   LTL^T  (crates/pfapack/src/ltl.rs:61-191) pivot by the first largest
//          |re| + |im|; order 0 = zsktf2: alpha = Julia 1/pivot
//          (julia_complex.rs:36-77), column k0 scaled by complex products
//          (backend.rs:351-355); order 1 = zsktf2_c_compat: alpha =
//          (re, -im) / (re^2 + im^2) (ltl.rs:256-263), scaling with fused
//          multiply-adds (ltl.rs:303-315); rank-2 update c + (x t1 - y t2)
//          (ltl.rs:707-818)
//   inverse (utu2.rs:399-515): trtri 'U','U', the skew-tridiagonal solve
//          multiplying by Julia reciprocals of -vt and vt
//          (utu2.rs:110-282, julia_complex.rs:79-122), the product and the
//          permutations as the real kernel.
// Complex arithmetic is num_complex's; the library is built with
// -ffp-contract=off.  Two complex doubles per register slot: the panel form
// stops at R = 5 (n <= 160) to stay within the 128-VGPR budget of a
// 1024-thread workgroup.
#pragma once
#include "pfinv_c128_dev.h"
#include "pfinv_rr.h"

namespace mvmc_hip {

constexpr int RR_C_NMAX = 160;

template <int ORDER>
struct RrCplx {
  using S = cz;
  static constexpr bool kRecip = true;  // the solve multiplies by Julia reciprocals
  static __device__ __forceinline__ S zero() { return cz{0.0, 0.0}; }
  static __device__ __forceinline__ S one() { return cz{1.0, 0.0}; }
  static __device__ __forceinline__ S neg(S x) { return cz_neg(x); }
  static __device__ __forceinline__ S add(S x, S y) { return cz_add(x, y); }
  static __device__ __forceinline__ S sub(S x, S y) { return cz_sub(x, y); }
  static __device__ __forceinline__ S mul(S x, S y) { return cz_mul(x, y); }
  static __device__ __forceinline__ double mag2(S x) { return x.re * x.re + x.im * x.im; }
  // |re| + |im|: a NaN wins only at index 0, as the sequential scan
  static __device__ __forceinline__ double key(S x, int i) { return rr_key(fabs(x.re) + fabs(x.im), i); }
  static __device__ __forceinline__ S alpha(S pv) {
    if (ORDER == 0) return julia_div(cz{1.0, 0.0}, pv);
    const double den = pv.re * pv.re + pv.im * pv.im;
    return cz{pv.re / den, -pv.im / den};
  }
  static __device__ __forceinline__ S update(S c, S xi, S yi, S t1, S t2) {
    return cz_add(c, cz_sub(cz_mul(xi, t1), cz_mul(yi, t2)));
  }
  static __device__ __forceinline__ S scale(S v, S alpha) {
    if (ORDER == 0) return cz_mul(v, alpha);
    return cz{fma(alpha.re, v.re, -(alpha.im * v.im)), fma(alpha.re, v.im, alpha.im * v.re)};
  }
  static __device__ __forceinline__ S sign(bool neg) { return neg ? cz{0.0 - 1.0, 0.0 - 0.0} : cz{1.0, 0.0}; }
  static __device__ __forceinline__ bool finite(S x) { return isfinite(x.re) && isfinite(x.im); }
  static __device__ __forceinline__ S div_negvt(S num, S, S rneg) { return cz_mul(num, rneg); }
  static __device__ __forceinline__ S div_vt(S num, S, S rpos) { return cz_mul(num, rpos); }
  static __device__ __forceinline__ S recip(S x) { return julia_recip(x); }
};

// Planes blockIdx.x, + gridDim.x, ... of [count][n][n] complex column-major
// planes; gs: 2 n^2 complex of scratch per workgroup.  Results and statuses
// as k_pfinv_c128 (a failed plane's inverse zero-filled, pf 0 on a zero
// pivot, the computed pf when non-finite); in == out works in place.
template <int T, int R, bool FULL, int ORDER>
__global__ void __launch_bounds__(T * T) k_pfinv_rr_c128(const cz* in, cz* out, int n, int count, cz* gs_base,
                                                        cz* pf, int* status) {
  extern __shared__ __align__(16) char lds[];
  Carver c{lds};
  RrLdsT<cz> L = carve_rr<cz>(c, n);
  cz* gs = gs_base + (size_t)blockIdx.x * 2 * n * n;
  for (int p = blockIdx.x; p < count; p += gridDim.x) {
    const cz* A = in + (size_t)p * n * n;
    cz* O = out + (size_t)p * n * n;
    cz pfv{0.0, 0.0};
    int info = 0;
    bool az = false;
    rr_pfinv<T, R, FULL, RrCplx<ORDER>>([&](int i, int j) { return A[i + (size_t)j * n]; }, n, O, 1.0, gs, L, pfv,
                                        info, false, az);
    if (threadIdx.x == 0) {
      status[p] = info;
      pf[p] = info > 0 ? cz{0.0, 0.0} : pfv;
    }
    if (info != 0)
      for (int e = threadIdx.x; e < n * n; e += blockDim.x) O[e] = cz{0.0, 0.0};
    __syncthreads();
  }
}

}  // namespace mvmc_hip
