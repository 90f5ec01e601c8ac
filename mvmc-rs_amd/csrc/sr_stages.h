// SrStages of the C order besides the CG step, on device operands
// (crates/mvmc-core/src/sr_backend.rs:193-292, COrderSr :429-516):
//   k_sr_assemble_stage   assemble_s_g   (S/g of the direct solve, stcopt.c:69)
//   k_gram_complex        gram_complex   (G = O O^H, samples summed in order)
//   k_ssg_*               sr_s_g         (weighted covariance S and force g of
//                                          the validation harness, over centered())
// Every output element keeps the reference's operation sequence (products
// rounded separately, sums over samples in ascending order; built with
// -ffp-contract=off); only independent elements run in parallel.
#pragma once
#include <hip/hip_runtime.h>

namespace mvmc_hip {

// Real view of a real array or of the real parts of a complex one
// (sr_backend.rs:53-70).
__device__ __forceinline__ double rview(const double* p, int cplx, long long i) { return cplx ? p[2 * i] : p[i]; }

// assemble_s_g (sr_backend.rs:484-516): element e < nm*nm is S column sj =
// e / nm, row si = e % nm; e in [nm*nm, nm*nm + nm) is g[e - nm*nm].
// m2dt = -2 * step_dt (formed on the host, as the reference's left-to-right
// product does).
__global__ void k_sr_assemble_stage(const double* oo, int oo_c, const double* ho, int ho_c, const int* map, int nm,
                                    long long ld, int off, double ratio_diag, double m2dt, double* s, double* g) {
  const long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  const long long nn = (long long)nm * nm;
  if (e < nn) {
    const int sj = (int)(e / nm), si = (int)(e % nm);
    const long long pi = map[si], pj = map[sj];
    const double tmp = rview(oo, oo_c, pi + off);
    double v = rview(oo, oo_c, (pi + off) * ld + (pj + off)) - tmp * rview(oo, oo_c, pj + off);
    if (si == sj) v *= ratio_diag;
    s[e] = v;
  } else if (e < nn + nm) {
    const int si = (int)(e - nn);
    const long long pi = map[si];
    const double v = rview(ho, ho_c, pi + off) - rview(ho, ho_c, 0) * rview(oo, oo_c, pi + off);
    g[si] = m2dt * v;
  }
}

// gram_complex (sr_backend.rs:386-420): G[i + j n] = sum_s O[i,s] conj(O[j,s])
// over a [samples][n] complex store (re, im interleaved).  A workgroup of
// 16 x 16 threads owns a 64 x 64 tile of G, each thread a 4 x 4 block
// (rows i0 + ty + 16a, columns j0 + tx + 16b); samples stream through LDS in
// chunks of GR_KC, and every entry adds its products in sample order:
// pr = o_i.re c.re - o_i.im c.im, pi = o_i.re c.im + o_i.im c.re with
// c = conj(o_j) (the num_complex product).
constexpr int GR_T = 64, GR_KC = 16;
__global__ void __launch_bounds__(256) k_gram_complex(const double* store, int n, int samples, double* out) {
  __shared__ double2 ti[GR_KC][GR_T];
  __shared__ double2 tj[GR_KC][GR_T];
  const int tx = threadIdx.x % 16, ty = threadIdx.x / 16;
  const int i0 = blockIdx.x * GR_T, j0 = blockIdx.y * GR_T;
  const double2* o = reinterpret_cast<const double2*>(store);
  double ar[4][4], ai[4][4];
#pragma unroll
  for (int a = 0; a < 4; ++a)
#pragma unroll
    for (int b = 0; b < 4; ++b) ar[a][b] = ai[a][b] = 0.0;
  for (int s0 = 0; s0 < samples; s0 += GR_KC) {
    for (int e = threadIdx.x; e < GR_KC * GR_T; e += 256) {
      const int ks = e / GR_T, c = e % GR_T;
      const int s = s0 + ks;
      const bool in_s = s < samples;
      ti[ks][c] = (in_s && i0 + c < n) ? o[(size_t)s * n + i0 + c] : make_double2(0.0, 0.0);
      tj[ks][c] = (in_s && j0 + c < n) ? o[(size_t)s * n + j0 + c] : make_double2(0.0, 0.0);
    }
    __syncthreads();
    const int kmax = min(GR_KC, samples - s0);
    for (int ks = 0; ks < kmax; ++ks) {
      double2 vi[4], vj[4];
#pragma unroll
      for (int a = 0; a < 4; ++a) vi[a] = ti[ks][ty + 16 * a];
#pragma unroll
      for (int b = 0; b < 4; ++b) vj[b] = tj[ks][tx + 16 * b];
#pragma unroll
      for (int a = 0; a < 4; ++a)
#pragma unroll
        for (int b = 0; b < 4; ++b) {
          const double cr = vj[b].x, ci = -vj[b].y;
          const double pr = vi[a].x * cr - vi[a].y * ci;
          const double pim = vi[a].x * ci + vi[a].y * cr;
          ar[a][b] += pr;
          ai[a][b] += pim;
        }
    }
    __syncthreads();
  }
#pragma unroll
  for (int a = 0; a < 4; ++a)
#pragma unroll
    for (int b = 0; b < 4; ++b) {
      const int i = i0 + ty + 16 * a, j = j0 + tx + 16 * b;
      if (i < n && j < n) {
        out[2 * ((size_t)j * n + i)] = ar[a][b];
        out[2 * ((size_t)j * n + i) + 1] = ai[a][b];
      }
    }
}

// Gram of the stored samples on the matrix cores (COrderSr::gram_real,
// sr_backend.rs:312-333, and the Gram of the direct solve, sr.rs:60-160):
// G[i + j na] = scale * sum_r P[r][i] P[r][j] over a row-major [rows][na]
// store.  One 256-thread workgroup per 64 x 64 tile of the upper triangle
// (tiles (ti, tj), ti <= tj, enumerated column by column); each wave owns a
// 32 x 32 quarter as 2 x 2 blocks of 16 x 16 accumulators; rows stream
// through LDS GM_KC at a time (double-buffered: the next chunk's global
// loads are in flight while the current one feeds the MFMAs, one barrier
// per chunk) and every block consumes them 4 per MFMA
// (v_mfma_f64_16x16x4f64, or v_mfma_f32_16x16x4f32 with FP32 accumulation
// for a float store).  MFMA operand layout (16 x 16 x 4, one block): lane l
// supplies A[l % 16][l / 16] and B[l / 16][l % 16]; accumulator element e of
// lane l is D[row(l, e)][l % 16] with row = l / 16 + 4 e for the FP64 MFMA
// and 4 (l / 16) + e for the FP32 one (both measured on the device against
// a host Gram of small integer stores).  Diagonal tiles write both halves;
// full != 0 mirrors every off-diagonal tile into the lower triangle.
// Summation order: per MFMA 4 rows, MFMAs in row order (not a serial sum).
typedef double gm_v4d __attribute__((ext_vector_type(4)));
typedef float gm_v4f __attribute__((ext_vector_type(4)));
template <class P>
struct GramMfma;
template <>
struct GramMfma<double> {
  using acc_t = gm_v4d;
  static __device__ __forceinline__ acc_t op(double a, double b, acc_t c) {
    return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
  }
  static __device__ __forceinline__ int row(int lane, int e) { return (lane >> 4) + 4 * e; }
};
template <>
struct GramMfma<float> {
  using acc_t = gm_v4f;
  static __device__ __forceinline__ acc_t op(float a, float b, acc_t c) {
    return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
  }
  static __device__ __forceinline__ int row(int lane, int e) { return 4 * (lane >> 4) + e; }
};
constexpr int GM_T = 64, GM_KC = 16;
constexpr int GM_XCDS = 8;  // XCDs of an MI355X (workgroup dispatch order)
constexpr int GM_LD = 80;  // LDS row stride (elements): the 4 k-rows of one operand fetch fall in two bank halves

__host__ __device__ inline long long gm_tiles(int na) {
  const long long nt = (na + GM_T - 1) / GM_T;
  return nt * (nt + 1) / 2;
}

template <class P, class OutT = double>
__global__ void __launch_bounds__(256) k_gram_mfma(const P* pack, int rows, int na, double scale, OutT* G,
                                                   int full) {
  using Op = GramMfma<P>;
  using acc_t = typename Op::acc_t;
  // two LDS buffers: the chunk in use and the next one, stored from the
  // registers its global loads landed in while the MFMAs ran
  __shared__ P sA[2][GM_KC][GM_LD];
  __shared__ P sB[2][GM_KC][GM_LD];
  // XCD-aware tile order: workgroups are dealt round-robin over the 8 XCDs
  // (blockIdx.x % 8), so XCD x gets the x-th contiguous range of the
  // column-by-column tile order and its L2 serves the column tiles its
  // neighbours in that range share
  const long long ntiles = gridDim.x;
  const long long xcd = blockIdx.x % GM_XCDS, local = blockIdx.x / GM_XCDS;
  const long long base = ntiles / GM_XCDS, rem = ntiles % GM_XCDS;
  const long long t = xcd * base + (xcd < rem ? xcd : rem) + local;
  // tile (ti, tj), ti <= tj, from the column-by-column index t
  int tj = (int)((sqrt(8.0 * (double)t + 1.0) - 1.0) * 0.5);
  while ((long long)(tj + 1) * (tj + 2) / 2 <= t) ++tj;
  while ((long long)tj * (tj + 1) / 2 > t) --tj;
  const int ti = (int)(t - (long long)tj * (tj + 1) / 2);
  const int i0 = ti * GM_T, j0 = tj * GM_T;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int wi = wave >> 1, wj = wave & 1;
  acc_t acc[2][2];
#pragma unroll
  for (int a = 0; a < 2; ++a)
#pragma unroll
    for (int b = 0; b < 2; ++b) acc[a][b] = acc_t{0, 0, 0, 0};
  const int lr = tid >> 4, lc = (tid & 15) * 4;  // loader: row lr of the chunk, 4 columns from lc
  P ra[4], rb[4];
  auto load = [&](int r0) {
    const int r = r0 + lr;
    const bool rin = r < rows;
    const P* row = pack + (size_t)(rin ? r : 0) * na;
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      const int ci = i0 + lc + u, cj = j0 + lc + u;
      ra[u] = (rin && ci < na) ? row[ci] : (P)0;
      rb[u] = (rin && cj < na) ? row[cj] : (P)0;
    }
  };
  auto store = [&](int buf) {
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      sA[buf][lr][lc + u] = ra[u];
      sB[buf][lr][lc + u] = rb[u];
    }
  };
  load(0);
  store(0);
  __syncthreads();
  int buf = 0;
  for (int r0 = 0; r0 < rows; r0 += GM_KC) {
    const bool more = r0 + GM_KC < rows;
    if (more) load(r0 + GM_KC);
#pragma unroll
    for (int k0 = 0; k0 < GM_KC; k0 += 4) {
      const int k = k0 + (lane >> 4), m = lane & 15;
      P av[2], bv[2];
#pragma unroll
      for (int a = 0; a < 2; ++a) av[a] = sA[buf][k][32 * wi + 16 * a + m];
#pragma unroll
      for (int b = 0; b < 2; ++b) bv[b] = sB[buf][k][32 * wj + 16 * b + m];
#pragma unroll
      for (int a = 0; a < 2; ++a)
#pragma unroll
        for (int b = 0; b < 2; ++b) acc[a][b] = Op::op(av[a], bv[b], acc[a][b]);
    }
    // the other buffer was last read before the previous barrier
    if (more) store(buf ^ 1);
    __syncthreads();
    buf ^= 1;
  }
#pragma unroll
  for (int a = 0; a < 2; ++a)
#pragma unroll
    for (int b = 0; b < 2; ++b)
#pragma unroll
      for (int e = 0; e < 4; ++e) {
        const int i = i0 + 32 * wi + 16 * a + Op::row(lane, e);
        const int j = j0 + 32 * wj + 16 * b + (lane & 15);
        if (i < na && j < na) {
          const OutT x = (OutT)(scale * (double)acc[a][b][e]);
          G[i + (size_t)j * na] = x;
          if (full && ti != tj) G[j + (size_t)i * na] = x;
        }
      }
}

// S x of the sampled operator from its Gram (the precision form of the
// SR-CG product, Gram by FP32 MFMA): out_i = inv_w sum_j G[j + i na] x_j
// - coef mean_i + shift diag_i x_i (minus from g when g is set), G the full
// symmetric Gram (column i read as row i), one wave per row, FP64
// accumulation of the FP32 Gram entries; coef = mean . x (k_gram_coef).
// ctl[0] set: the solve has ended (as the other CG kernels).
__global__ void __launch_bounds__(256) k_gram_apply(const float* G, int na, const double* x, const double* mean,
                                                    const double* diag, const double* coef, double inv_w,
                                                    double shift, const double* g, double* out, const int* ctl) {
  if (ctl && ctl[0]) return;
  const int lane = threadIdx.x & 63;
  const int i = blockIdx.x * 4 + (threadIdx.x >> 6);
  if (i >= na) return;
  const float* col = G + (size_t)i * na;
  double acc = 0.0;
  for (int j = lane; j < na; j += 64) acc += (double)col[j] * x[j];
#pragma unroll
  for (int off = 32; off >= 1; off >>= 1) acc += __shfl_xor(acc, off, 64);
  if (lane == 0) {
    const double z = inv_w * acc - coef[0] * mean[i] + shift * diag[i] * x[i];
    out[i] = g ? g[i] - z : z;
  }
}

// coef = mean . x (one 256-thread block, FP64).
__global__ void __launch_bounds__(256) k_gram_coef(const double* mean, const double* x, int na, double* coef,
                                                   const int* ctl) {
  if (ctl && ctl[0]) return;
  __shared__ double red[256];
  double p = 0.0;
  for (int i = threadIdx.x; i < na; i += 256) p += mean[i] * x[i];
  red[threadIdx.x] = p;
  __syncthreads();
  for (int h = 128; h >= 1; h >>= 1) {
    if ((int)threadIdx.x < h) red[threadIdx.x] += red[threadIdx.x + h];
    __syncthreads();
  }
  if (threadIdx.x == 0) coef[0] = red[0];
}

// sr_s_g, centered() (stage_backend.rs:167-184): thread p < np forms the
// weighted mean of column p (k ascending) and writes d[., p] = o[., p] - mean;
// thread np forms <E>_w and writes de.  o, d: column-major [ns][np].
__global__ void k_ssg_center(const double* o, int ns, int np, const double* e, const double* w, double* d,
                             double* de) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p < np) {
    const double* col = o + (size_t)p * ns;
    double mean = 0.0;
    for (int k = 0; k < ns; ++k) mean += w[k] * col[k];
    for (int k = 0; k < ns; ++k) d[(size_t)p * ns + k] = col[k] - mean;
  } else if (p == np) {
    double em = 0.0;
    for (int k = 0; k < ns; ++k) em += w[k] * e[k];
    for (int k = 0; k < ns; ++k) de[k] = e[k] - em;
  }
}

// S[p + q np] = sum_k (w_k d_kp) d_kq (sr_backend.rs:447-452): a 64 x 64 tile
// per workgroup, 4 x 4 entries per thread, k streamed through LDS in order.
__global__ void __launch_bounds__(256) k_ssg_s(const double* d, const double* w, int ns, int np, double* s) {
  __shared__ double tp[GR_KC][GR_T];
  __shared__ double tq[GR_KC][GR_T];
  __shared__ double tw[GR_KC];
  const int tx = threadIdx.x % 16, ty = threadIdx.x / 16;
  const int p0 = blockIdx.x * GR_T, q0 = blockIdx.y * GR_T;
  double acc[4][4];
#pragma unroll
  for (int a = 0; a < 4; ++a)
#pragma unroll
    for (int b = 0; b < 4; ++b) acc[a][b] = 0.0;
  for (int k0 = 0; k0 < ns; k0 += GR_KC) {
    for (int e = threadIdx.x; e < GR_KC * GR_T; e += 256) {
      const int kk = e % GR_KC, c = e / GR_KC;   // k fastest: coalesced along a column
      const int k = k0 + kk;
      tp[kk][c] = (k < ns && p0 + c < np) ? d[(size_t)(p0 + c) * ns + k] : 0.0;
      tq[kk][c] = (k < ns && q0 + c < np) ? d[(size_t)(q0 + c) * ns + k] : 0.0;
    }
    if (threadIdx.x < GR_KC) tw[threadIdx.x] = k0 + (int)threadIdx.x < ns ? w[k0 + threadIdx.x] : 0.0;
    __syncthreads();
    const int kmax = min(GR_KC, ns - k0);
    for (int kk = 0; kk < kmax; ++kk) {
      const double wk = tw[kk];
      double vp[4], vq[4];
#pragma unroll
      for (int a = 0; a < 4; ++a) vp[a] = wk * tp[kk][ty + 16 * a];
#pragma unroll
      for (int b = 0; b < 4; ++b) vq[b] = tq[kk][tx + 16 * b];
#pragma unroll
      for (int a = 0; a < 4; ++a)
#pragma unroll
        for (int b = 0; b < 4; ++b) acc[a][b] += vp[a] * vq[b];
    }
    __syncthreads();
  }
#pragma unroll
  for (int a = 0; a < 4; ++a)
#pragma unroll
    for (int b = 0; b < 4; ++b) {
      const int p = p0 + ty + 16 * a, q = q0 + tx + 16 * b;
      if (p < np && q < np) s[p + (size_t)q * np] = acc[a][b];
    }
}

// g[q] = sum_k (w_k d_kq) de_k (sr_backend.rs:453-457), one thread per q.
__global__ void k_ssg_g(const double* d, const double* w, const double* de, int ns, int np, double* g) {
  const int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= np) return;
  const double* col = d + (size_t)q * ns;
  double acc = 0.0;
  for (int k = 0; k < ns; ++k) acc += w[k] * col[k] * de[k];
  g[q] = acc;
}

}  // namespace mvmc_hip
