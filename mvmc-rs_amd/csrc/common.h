// Shared device definitions of the MI355X walker-ensemble library.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "ddmath.h"

namespace mvmc_hip {

// Sticky status codes (first failing stage wins; 0 = ok).
enum Status : int {
  ST_OK = 0,
  ST_INIT_RETRY = 1,      // initial placement exhausted 101 validated attempts
  ST_BAD_INPUT = 2,       // invalid shape / index (host or kernel bounds check)
  ST_RECOMPUTE_FAIL = 3,  // a from-scratch Pfaffian rebuild failed (informational)
  ST_NONFINITE = 4,       // non-finite local energy / derivative
  ST_CG_FAIL = 5,         // SR-CG breakdown with abort requested
};

// ---------------------------------------------------------------------------
// SFMT19937 on the device.  State lives in LDS (624 words + cursor) while a
// walker is active and in HBM between launches; refills are done by one lane.
// Algorithm: Saito & Matsumoto, SFMT-19937:122-18-1-11-1 (parameters
// dfffffef-ddfecb7f-bffaffff-bffffff6), little-endian 128-bit words.
// ---------------------------------------------------------------------------
struct SfmtLds {
  uint32_t w[624];
  int idx;
};

__host__ __device__ inline void sfmt_recursion(uint32_t* r, const uint32_t* a, const uint32_t* b, const uint32_t* c,
                                      const uint32_t* d) {
  const uint64_t ah = ((uint64_t)a[3] << 32) | a[2], al = ((uint64_t)a[1] << 32) | a[0];
  const uint64_t xh = (ah << 8) | (al >> 56), xl = al << 8;
  const uint64_t ch = ((uint64_t)c[3] << 32) | c[2], cl = ((uint64_t)c[1] << 32) | c[0];
  const uint64_t yh = ch >> 8, yl = (cl >> 8) | (ch << 56);
  r[0] = a[0] ^ (uint32_t)xl ^ ((b[0] >> 11) & 0xdfffffefu) ^ (uint32_t)yl ^ (d[0] << 18);
  r[1] = a[1] ^ (uint32_t)(xl >> 32) ^ ((b[1] >> 11) & 0xddfecb7fu) ^ (uint32_t)(yl >> 32) ^ (d[1] << 18);
  r[2] = a[2] ^ (uint32_t)xh ^ ((b[2] >> 11) & 0xbffaffffu) ^ (uint32_t)yh ^ (d[2] << 18);
  r[3] = a[3] ^ (uint32_t)(xh >> 32) ^ ((b[3] >> 11) & 0xbffffff6u) ^ (uint32_t)(yh >> 32) ^ (d[3] << 18);
}

// Single-lane refill of all 156 128-bit blocks (gen_rand_all).
__host__ __device__ inline void sfmt_refill_words(uint32_t* w) {
  uint32_t r1[4], r2[4];
  for (int k = 0; k < 4; ++k) { r1[k] = w[4 * 154 + k]; r2[k] = w[4 * 155 + k]; }
  for (int i = 0; i < 156; ++i) {
    const int j = i < 34 ? i + 122 : i - 34;
    uint32_t a[4], b[4], out[4];
    for (int k = 0; k < 4; ++k) { a[k] = w[4 * i + k]; b[k] = w[4 * j + k]; }
    sfmt_recursion(out, a, b, r1, r2);
    for (int k = 0; k < 4; ++k) {
      w[4 * i + k] = out[k];
      r1[k] = r2[k];
      r2[k] = out[k];
    }
  }
}
__host__ __device__ inline void sfmt_refill(SfmtLds& s) { sfmt_refill_words(s.w); }

__host__ __device__ inline uint32_t sfmt_u32(SfmtLds& s) {
  if (s.idx >= 624) {
    sfmt_refill(s);
    s.idx = 0;
  }
  return s.w[s.idx++];
}
__host__ __device__ inline double sfmt_real2(SfmtLds& s) { return (double)sfmt_u32(s) * (1.0 / 4294967296.0); }
__host__ __device__ inline uint32_t sfmt_mod(SfmtLds& s, uint32_t n) { return sfmt_u32(s) % n; }

// Host-side seeding (same recurrence; used when creating the ensemble).
inline void sfmt_seed_host(uint32_t* w, uint32_t seed) {
  w[0] = seed;
  for (int i = 1; i < 624; ++i) w[i] = 1812433253u * (w[i - 1] ^ (w[i - 1] >> 30)) + (uint32_t)i;
  static const uint32_t parity[4] = {0x00000001u, 0x00000000u, 0x00000000u, 0x13c9e684u};
  uint32_t inner = 0;
  for (int i = 0; i < 4; ++i) inner ^= w[i] & parity[i];
  for (int i = 16; i > 0; i >>= 1) inner ^= inner >> i;
  if (inner & 1) return;
  for (int i = 0; i < 4; ++i) {
    uint32_t work = 1;
    for (int j = 0; j < 32; ++j) {
      if (work & parity[i]) { w[i] ^= work; return; }
      work <<= 1;
    }
  }
}

}  // namespace mvmc_hip
