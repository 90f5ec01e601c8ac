// ORACLE (test infrastructure only): scalar SFMT19937 generator.
// Restates crates/sfmt19937/src/{init.rs:1-40,state.rs:104-180,lib.rs:138-182,
// real.rs:20-30}: init_gen_rand (1812433253 recurrence + period
// certification), gen_rand_all (128-bit recursion, POS1=122, SL1=18, SL2=1 byte,
// SR1=11, SR2=1 byte), one 32-bit word per draw, real2 = u32 * 2^-32.
// Pinned by tests/golden/sfmt_golden.json (C-reference words for six seeds
// and around three refills, crates/sfmt19937/tests/golden_vs_c.rs:40-160).
#pragma once
#include <cstdint>
#include <cstring>

namespace oracle {

struct Sfmt {
  static const int N = 156, N32 = 624, POS1 = 122;
  uint32_t w[N32];
  int idx;
  uint64_t consumed = 0;

  void init(uint32_t seed) {
    w[0] = seed;
    for (int i = 1; i < N32; ++i) w[i] = 1812433253u * (w[i - 1] ^ (w[i - 1] >> 30)) + (uint32_t)i;
    idx = N32;
    consumed = 0;
    certify();
  }
  void certify() {
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
  static void recursion(uint32_t* r, const uint32_t* a, const uint32_t* b, const uint32_t* c, const uint32_t* d) {
    // x = a << 8 (128-bit), y = c >> 8 (128-bit), little-endian word order.
    uint64_t ah = ((uint64_t)a[3] << 32) | a[2], al = ((uint64_t)a[1] << 32) | a[0];
    uint64_t xh = (ah << 8) | (al >> 56), xl = al << 8;
    uint64_t ch = ((uint64_t)c[3] << 32) | c[2], cl = ((uint64_t)c[1] << 32) | c[0];
    uint64_t yh = ch >> 8, yl = (cl >> 8) | (ch << 56);
    const uint32_t x[4] = {(uint32_t)xl, (uint32_t)(xl >> 32), (uint32_t)xh, (uint32_t)(xh >> 32)};
    const uint32_t y[4] = {(uint32_t)yl, (uint32_t)(yl >> 32), (uint32_t)yh, (uint32_t)(yh >> 32)};
    static const uint32_t msk[4] = {0xdfffffefu, 0xddfecb7fu, 0xbffaffffu, 0xbffffff6u};
    for (int k = 0; k < 4; ++k) r[k] = a[k] ^ x[k] ^ ((b[k] >> 11) & msk[k]) ^ y[k] ^ (d[k] << 18);
  }
  void refill() {
    uint32_t r1[4], r2[4];
    std::memcpy(r1, w + 4 * (N - 2), 16);
    std::memcpy(r2, w + 4 * (N - 1), 16);
    for (int i = 0; i < N; ++i) {
      int j = i < N - POS1 ? i + POS1 : i + POS1 - N;
      uint32_t out[4];
      recursion(out, w + 4 * i, w + 4 * j, r1, r2);
      std::memcpy(w + 4 * i, out, 16);
      std::memcpy(r1, r2, 16);
      std::memcpy(r2, out, 16);
    }
  }
  uint32_t gen32() {
    if (idx >= N32) { refill(); idx = 0; }
    ++consumed;
    return w[idx++];
  }
  double real2() { return (double)gen32() * (1.0 / 4294967296.0); }
  uint32_t mod(uint32_t n) { return gen32() % n; }
};

}  // namespace oracle
