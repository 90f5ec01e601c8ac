// ORACLE — test infrastructure only.  A plain CPU FP64 restatement of the
// reference's real sampling + SR-CG optimisation path, used by tests/ and by
// bench.py's cpu_baseline leg as the checker.  The product (the HIP library)
// never links or calls this file.
//
// Every function names the reference lines whose operation order it follows.
// Pins (tests/test_oracle_*.py):
//   * SFMT words:            tests/golden/sfmt_golden.json
//   * Pfaffian / inverse:    closed forms, brute force, Pf^2=det, residuals
//   * sampler + measurement: native-C step-1 operands of the Hubbard chain
//                            (tests/golden/hubbard_cg_store0.json)
//   * multi-walker sampler:  native-C per-rank counters / EleIdx / SFMT state
//                            (tests/golden/mpi179_*.json, Heisenberg chain)
//   * SR-CG iterates:        native-C iterates (tests/golden/sr_cg_real*.txt)
//   * hopping candidate:     C retry-cap word tapes
//                            (tests/golden/candidate_hopping_retry_cap.txt)
//   * Metropolis decision:   closed forms (weight, non-finite, real2 bound)
//   * local energy:          brute force, every Green ratio against full
//                            Pfaffians of the moved configuration
//   * complex Pfaffian:      see pfapack_c.h
//   * direct SR (NSRCG 0):   full <OO>, <HO> and rank states of the native-C
//                            direct cells (tests/golden/mpi179_real_d*.json);
//                            S/g assembly + Cholesky against the C dposv
//                            solutions (tests/golden/negative_stepdt_c_sr.json)
//   * gram_complex, sr_s_g:  exact sums on integer / dyadic inputs, exact
//                            Hermitian symmetry
// Parity unpinned: the inverse against OpenBLAS dtrtri/dtrmm bits (residual
// checks only), the Cholesky solve against OpenBLAS dpotrf/dpotrs bits
// (within 1e-13 of the C solutions), glibc exp/log bits on a different glibc
// version.
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <stdexcept>
#include <string>
#include <thread>
#include <vector>

#include "../mvmc-rs_amd/csrc/ddmath.h"
#include "expert.h"
#include "pfapack.h"
#include "pfapack_c.h"
#include "sfmt.h"

using std::vector;

namespace oracle {

// exp/log used by the Metropolis weight, the overlap logarithm and the
// projection ratios.  Default: the host libm, as the reference (glibc).
// Mode 1: the correctly rounded double-double routines that the device uses,
// so that GPU-parity runs compare identical arithmetic; the two agree except
// where glibc misrounds a near-midpoint case (measured in
// tests/test_oracle_math.py).
static int g_cr_math = 0;
static inline double o_exp(double x) { return g_cr_math ? ddm::cr_exp(x) : std::exp(x); }
static inline double o_log(double x) { return g_cr_math ? ddm::cr_log(x) : std::log(x); }

// ---------------------------------------------------------------------------
// Walker state (one independent chain = one ungrouped C rank,
// crates/mvmc-core/src/multichain.rs:1-60)
// ---------------------------------------------------------------------------
struct Walker {
  Sfmt rng;
  vector<int> ele_idx, ele_cfg, ele_num, proj_cnt;
  vector<int> burn_idx, burn_cfg, burn_num, burn_proj;
  vector<double> inv, pf;  // inv: nqp planes of n*n
  long long counter[10] = {0};
  // saved samples
  vector<int> s_idx, s_cfg, s_num, s_proj;
  // accumulators of the current SR step
  double wc = 0, etot = 0, etot2 = 0, sztot = 0, sztot2 = 0;
  vector<double> ho, oo, o_store, o_last;
  int n_saved = 0;
  long long accepted = 0;
};

struct Run {
  Model m;
  vector<double> slater;  // S[qp][2Ns][2Ns]
  vector<Walker> w;
  vector<double> dproj_g, dproj_j;  // direct projection tables
  vector<int> tsite, tsgn;          // translation maps
  // reduced results of the last step
  double wc = 0, etot = 0, etot2 = 0, sztot = 0, sztot2 = 0;
  vector<double> oo, ho;
  vector<double> energies;  // per step zvo_out rows (6 values each)
  int cg_status = 0;
  vector<double> cg_solution;
  double t_sample = 0, t_measure = 0, t_sr = 0;
  int threads = 1;
};

// ---------------------------------------------------------------------------
// calc_m_all: crates/mvmc-core/src/pfaffian.rs:141-199, 871-927, 1107-1188
// ---------------------------------------------------------------------------
// Returns 0 on success, 1 zero pivot / all-zero / non-finite Pfaffian.
static int calc_m_all(const Run& R, const int* ele_idx, double* inv, double* pf, bool native_status,
                      int qp_begin = 0, int qp_end = -1) {
  const Model& m = R.m;
  const int n = m.nsize(), ne = m.ne(), ns = m.nsite(), n2 = m.nsite2(), nqp = m.nqp();
  if (qp_end < 0) qp_end = nqp;
  vector<int> piv(n), rs(n);
  for (int msi = 0; msi < n; ++msi) rs[msi] = ele_idx[msi] + (msi / ne) * ns;
  int status = 0;
  for (int qp = qp_begin; qp < qp_end; ++qp) {
    double* plane = inv + (size_t)qp * n * n;
    const double* S = R.slater.data() + (size_t)qp * n2 * n2;
    double max_abs2 = 0.0;
    for (int msi = 0; msi < n; ++msi) {
      const double* row = S + (size_t)rs[msi] * n2;
      for (int msj = 0; msj < n; ++msj) {
        double v = -row[rs[msj]];
        plane[msj + (size_t)msi * n] = v;
        max_abs2 = std::fmax(max_abs2, v * v);
      }
    }
    if (!native_status && max_abs2 < 1.0e-28) return 1;
    int info = dsktf2(plane, n, piv.data());
    if (info != 0) { status = 1; if (!native_status) return 1; continue; }
    double p = utu2pfa(plane, n, piv.data());
    if (!std::isfinite(p)) { status = 1; if (!native_status) return 1; continue; }
    pf[qp] = p;
    utu2inv(plane, n, piv.data());
    for (size_t k = 0; k < (size_t)n * n; ++k) plane[k] = -plane[k];
  }
  return status;
}

// observables.rs:128-145
static double calc_ip(const Model& m, const double* pf) {
  double ip = 0.0;
  for (int q = 0; q < m.nqp(); ++q) ip += m.qp_full_weight[q] * pf[q];
  return ip;
}

// ---------------------------------------------------------------------------
// Projection counters: crates/mvmc-core/src/sampling/projection.rs:179-403
// ---------------------------------------------------------------------------
static void make_proj_cnt(const Model& m, const int* ele_num, int* cnt) {
  const int ns = m.nsite();
  const int* n0 = ele_num;
  const int* n1 = ele_num + ns;
  for (int i = 0; i < m.n_proj(); ++i) cnt[i] = 0;
  if (m.n_gutz > 0)
    for (int ri = 0; ri < ns; ++ri) {
      int idx = m.gutz_idx[ri];
      if (idx >= 0 && idx < m.n_proj()) cnt[idx] += n0[ri] * n1[ri];
    }
  if (m.n_jast > 0)
    for (int ri = 0; ri < ns; ++ri) {
      int xi = n0[ri] + n1[ri] - 1;
      if (xi == 0) continue;
      for (int rj = ri + 1; rj < ns; ++rj) {
        int idx = m.jast_idx[(size_t)ri * ns + rj];
        if (idx < 0) continue;
        int xj = n0[rj] + n1[rj] - 1;
        cnt[m.n_gutz + idx] += xi * xj;
      }
    }
}

// projection.rs:278-398; ele_num is the post-hop occupation.
static void update_proj_cnt(const Model& m, int ri, int rj, int* cnt, const int* ele_num) {
  if (ri == rj) return;
  const int ns = m.nsite();
  const int* n0 = ele_num;
  const int* n1 = ele_num + ns;
  if (m.n_gutz > 0) {
    int a = m.gutz_idx[ri], b = m.gutz_idx[rj];
    if (a >= 0) cnt[a] -= n0[ri] + n1[ri];
    if (b >= 0) cnt[b] += n0[rj] * n1[rj];
  }
  if (m.n_jast > 0) {
    const int off = m.n_gutz;
    int xri = n0[ri] + n1[ri], xrj = n0[rj] + n1[rj];
    int idx = m.jast(ri, rj);
    if (idx >= 0) cnt[off + idx] += xri - xrj + 1;
    for (int rk = 0; rk < ns; ++rk) {
      if (rk == ri || rk == rj) continue;
      int x = n0[rk] + n1[rk] - 1;
      int a = m.jast(ri, rk);
      if (a >= 0) cnt[off + a] -= x;
      int b = m.jast(rj, rk);
      if (b >= 0) cnt[off + b] += x;
    }
  }
}

static double log_proj_ratio(const Model& m, const int* cnew, const int* cold) {
  double z = 0.0;
  for (int i = 0; i < m.n_proj(); ++i) z += m.proj[i] * (double)(cnew[i] - cold[i]);
  return z;
}

static void update_ele_config(const Model& m, int mi, int ri, int rj, int s, int* ele_idx, int* ele_cfg,
                              int* ele_num) {
  const int msa = mi + s * m.ne(), old = ri + s * m.nsite(), nw = rj + s * m.nsite();
  ele_idx[msa] = rj;
  ele_cfg[old] = -1;
  ele_cfg[nw] = mi;
  ele_num[old] = 0;
  ele_num[nw] = 1;
}

static void revert_ele_config(const Model& m, int mi, int ri, int rj, int s, int* ele_idx, int* ele_cfg,
                              int* ele_num) {
  const int msa = mi + s * m.ne(), old = ri + s * m.nsite(), nw = rj + s * m.nsite();
  ele_idx[msa] = ri;
  ele_cfg[old] = mi;
  ele_cfg[nw] = -1;
  ele_num[old] = 1;
  ele_num[nw] = 0;
}

// ---------------------------------------------------------------------------
// Pfaffian ratios and inverse updates: crates/mvmc-core/src/sampling/updates.rs
// ---------------------------------------------------------------------------
// :73-158 calculate_new_pf_m2_real_flat
static void new_pf_hop(const Run& R, const Walker& W, int ma, int s, double* pf_new) {
  const Model& m = R.m;
  const int n = m.nsize(), ne = m.ne(), ns = m.nsite(), n2 = m.nsite2();
  const int msa = ma + s * ne;
  const int rsa = W.ele_idx[msa] + s * ns;
  for (int q = 0; q < m.nqp(); ++q) {
    const double* row = R.slater.data() + ((size_t)q * n2 + rsa) * n2;
    const double* inv = W.inv.data() + (size_t)q * n * n + (size_t)msa * n;
    double ratio = 0.0;
    for (int msj = 0; msj < n; ++msj) {
      int rsj = W.ele_idx[msj] + (msj < ne ? 0 : ns);
      ratio += inv[msj] * row[rsj];
    }
    pf_new[q] = -ratio * W.pf[q];
  }
}

// :588-638 + :1184-1224 update_one_real_scalar
static void update_hop(const Run& R, Walker& W, int ma, int s) {
  const Model& m = R.m;
  const int n = m.nsize(), ne = m.ne(), ns = m.nsite(), n2 = m.nsite2();
  const int msa = ma + s * ne;
  const int rsa = W.ele_idx[msa] + s * ns;
  vector<double> slt(n), v1(n), v2(n);
  for (int q = 0; q < m.nqp(); ++q) {
    const double* row = R.slater.data() + ((size_t)q * n2 + rsa) * n2;
    for (int msj = 0; msj < n; ++msj) slt[msj] = row[W.ele_idx[msj] + (msj < ne ? 0 : ns)];
    double* inv = W.inv.data() + (size_t)q * n * n;
    for (int j = 0; j < n; ++j) v1[j] = 0.0;
    for (int i = 0; i < n; ++i) {
      const double* col = inv + (size_t)i * n;
      for (int j = 0; j < n; ++j) v1[j] += -col[j] * slt[i];
    }
    const double tmp = v1[msa];
    W.pf[q] *= -tmp;
    const double inv_a = -1.0 / tmp;
    for (int j = 0; j < n; ++j) v2[j] = inv[(size_t)msa * n + j] * inv_a;
    for (int i = 0; i < n; ++i) {
      double* r = inv + (size_t)i * n;
      for (int j = 0; j < n; ++j) r[j] += v1[i] * v2[j] - v1[j] * v2[i];
      r[msa] -= v2[i];
    }
    for (int j = 0; j < n; ++j) inv[(size_t)msa * n + j] += v2[j];
  }
}

// :1591-1622 two_hop_bilinear_c_order
static double two_hop_bilinear(const double* inv, int n, const double* a, const double* b) {
  double bma = 0.0;
  for (int i = 0; i < n; ++i) {
    double tmp = 0.0;
    for (int j = 0; j < n; ++j) tmp += inv[(size_t)i * n + j] * a[j];
    bma += b[i] * tmp;
  }
  return bma;
}

// :395-441 + :1539-1568 (two_ratio_real, C bilinear)
static void new_pf_exchange(const Run& R, const Walker& W, int ma, int s, int mb, int t, double* pf_new) {
  const Model& m = R.m;
  const int n = m.nsize(), ne = m.ne(), ns = m.nsite(), n2 = m.nsite2();
  const int msa = ma + s * ne, msb = mb + t * ne;
  if (msa == msb) { new_pf_hop(R, W, mb, t, pf_new); return; }
  const int rsa = W.ele_idx[msa] + s * ns, rsb = W.ele_idx[msb] + t * ns;
  vector<double> va(n), vb(n);
  for (int q = 0; q < m.nqp(); ++q) {
    const double* S = R.slater.data() + (size_t)q * n2 * n2;
    for (int msi = 0; msi < n; ++msi) {
      int rsi = W.ele_idx[msi] + (msi < ne ? 0 : ns);
      va[msi] = S[(size_t)rsa * n2 + rsi];
      vb[msi] = S[(size_t)rsb * n2 + rsi];
    }
    const double* inv = W.inv.data() + (size_t)q * n * n;
    const double vba = vb[msa];
    double pa = 0, pb = 0, qa = 0, qb = 0;
    for (int i = 0; i < n; ++i) {
      double ia = inv[(size_t)msa * n + i], ib = inv[(size_t)msb * n + i];
      pa += ia * va[i];
      pb += ib * va[i];
      qa += ia * vb[i];
      qb += ib * vb[i];
    }
    const double iab = inv[(size_t)msa * n + msb];
    const double bma = two_hop_bilinear(inv, n, va.data(), vb.data());
    const double ratio = iab * vba + iab * bma + pa * qb - pb * qa;
    pf_new[q] = ratio * W.pf[q];
  }
}

// :779-830 + :1429-1503 update_two_real
static void update_exchange(const Run& R, Walker& W, int ma, int s, int mb, int t, int ra_old, int rb_old) {
  const Model& m = R.m;
  const int n = m.nsize(), ne = m.ne(), ns = m.nsite(), n2 = m.nsite2();
  const int msa = ma + s * ne, msb = mb + t * ne;
  const int rsa = W.ele_idx[msa] + s * ns, rsb = W.ele_idx[msb] + t * ns;
  const int rsa_old = ra_old + s * ns, rsb_old = rb_old + t * ns;
  vector<double> P(n), Q(n), Sv(n), T(n);
  for (int q = 0; q < m.nqp(); ++q) {
    const double* S = R.slater.data() + (size_t)q * n2 * n2;
    const double m_old_ab = S[(size_t)rsa_old * n2 + rsb_old];
    for (int msi = 0; msi < n; ++msi) {
      int rsi = W.ele_idx[msi] + (msi < ne ? 0 : ns);
      Sv[msi] = S[(size_t)rsa * n2 + rsi];
      T[msi] = S[(size_t)rsb * n2 + rsi];
    }
    Sv[msb] = m_old_ab;
    double* inv = W.inv.data() + (size_t)q * n * n;
    for (int i = 0; i < n; ++i) { P[i] = 0.0; Q[i] = 0.0; }
    for (int i = 0; i < n; ++i)
      for (int j = 0; j < n; ++j) {
        double x = inv[(size_t)i * n + j];
        P[i] += x * Sv[j];
        Q[i] += x * T[j];
      }
    double bma = 0.0;
    for (int i = 0; i < n; ++i) bma += T[i] * P[i];
    const double iab = inv[(size_t)msa * n + msb];
    const double ratio = iab * T[msa] + iab * bma + P[msa] * Q[msb] - P[msb] * Q[msa];
    W.pf[q] *= ratio;
    const double a = -P[msa], b = P[msb], c = Q[msa], d = -Q[msb];
    const double e = -bma - T[msa];
    const double f = iab;
    const double det = a * d - b * c - e * f;
    const double idet = 1.0 / det;
    for (int i = 0; i < n; ++i) {
      Sv[i] = idet * inv[(size_t)msa * n + i];
      T[i] = idet * inv[(size_t)msb * n + i];
    }
    for (int i = 0; i < n; ++i) {
      const double pi = P[i], qi = Q[i], si = Sv[i], ti = T[i];
      for (int j = 0; j < n; ++j) {
        const double pj = P[j], qj = Q[j], sj = Sv[j], tj = T[j];
        inv[(size_t)i * n + j] += a * (qi * tj - qj * ti) + b * (qi * sj - qj * si) + c * (pi * tj - pj * ti) +
                                  d * (pi * sj - pj * si) + e * det * (si * tj - sj * ti) +
                                  f * idet * (pi * qj - qi * pj);
      }
      inv[(size_t)i * n + msa] += -c * ti - d * si - f * idet * qi;
      inv[(size_t)i * n + msb] += -a * ti - b * si + f * idet * pi;
    }
    for (int j = 0; j < n; ++j) {
      const double pj = P[j], qj = Q[j], sj = Sv[j], tj = T[j];
      inv[(size_t)msa * n + j] += c * tj + d * sj + f * idet * qj;
      inv[(size_t)msb * n + j] += a * tj + b * sj - f * idet * pj;
    }
    inv[(size_t)msa * n + msb] += f * idet;
    inv[(size_t)msb * n + msa] -= f * idet;
  }
}

// ---------------------------------------------------------------------------
// Sampler: crates/mvmc-core/src/sampling/{initial.rs:77-141,
// normal_initial.rs:160-330, candidate.rs:137-385, metropolis.rs:42-83,
// engine.rs:188-475, driver.rs:175-458}
// ---------------------------------------------------------------------------
static void make_initial_sample(const Run& R, Walker& W) {
  const Model& m = R.m;
  const int ns = m.nsite(), ne = m.ne();
  std::fill(W.ele_idx.begin(), W.ele_idx.end(), -1);
  std::fill(W.ele_cfg.begin(), W.ele_cfg.end(), -1);
  for (int ri = 0; ri < ns; ++ri) {
    if (m.loc_spn[ri] != 1) continue;
    for (;;) {
      int mi = (int)W.rng.mod((uint32_t)ne);
      double r = W.rng.real2();
      int si = r < 0.5 ? 0 : 1;
      if (W.ele_idx[mi + si * ne] == -1) {
        W.ele_cfg[ri + si * ns] = mi;
        W.ele_idx[mi + si * ne] = ri;
        break;
      }
    }
  }
  for (int si = 0; si < 2; ++si)
    for (int mi = 0; mi < ne; ++mi) {
      if (W.ele_idx[mi + si * ne] != -1) continue;
      for (;;) {
        int ri = (int)W.rng.mod((uint32_t)ns);
        if (W.ele_cfg[ri + si * ns] == -1 && m.loc_spn[ri] != 1) {
          W.ele_cfg[ri + si * ns] = mi;
          W.ele_idx[mi + si * ne] = ri;
          break;
        }
      }
    }
  for (int rsi = 0; rsi < 2 * ns; ++rsi) W.ele_num[rsi] = W.ele_cfg[rsi] < 0 ? 0 : 1;
  make_proj_cnt(m, W.ele_num.data(), W.proj_cnt.data());
}

static void make_initial_sample_validated(const Run& R, Walker& W) {
  for (int attempt = 1; attempt <= 101; ++attempt) {
    make_initial_sample(R, W);
    int info = calc_m_all(R, W.ele_idx.data(), W.inv.data(), W.pf.data(), true);
    if (attempt > 100) throw std::runtime_error("makeInitialSample: too many loops");
    if (info <= 0) return;
  }
}

static double log_ip_of(const Model& m, const double* pf) { return o_log(std::fabs(calc_ip(m, pf))); }

// metropolis_weight / metropolis_decision (metropolis.rs:42-83): the weight
// w = exp(2((dproj + 0) + lip_new - lip_old)), a non-finite weight becomes -1,
// accept iff w > real2 (one word).
template <class RNG>
static bool metropolis_rng(RNG& rng, double lpd, double lip_new, double lip_old, double* w_out = nullptr) {
  const double ex = 2.0 * (((lpd + 0.0) + lip_new) - lip_old);
  double w = o_exp(ex);
  if (!std::isfinite(w)) w = -1.0;
  if (w_out) *w_out = w;
  const double u = rng.real2();
  return w > u;
}
static bool metropolis(Walker& W, double lpd, double lip_new, double lip_old) {
  return metropolis_rng(W.rng, lpd, lip_new, lip_old);
}

// makeCandidate_hopping (candidate.rs:137-227; C vmcmake.c): draw-until-
// itinerant source electron (cap Nsite^2 + 1 draws, then reject if no
// itinerant electron exists), then draw-until-empty destination with the C
// order "draw, test the cap, test occupancy".  Returns true on rejection; the
// last drawn mi, s, ri, rj are reported either way.  RNG: gen32 / mod / real2.
template <class RNG>
static bool hop_candidate(int ns, int ne, const int* loc_spn, const int* ele_idx, const int* ele_cfg, RNG& rng,
                          int& mi, int& s, int& ri, int& rj) {
  auto is_local = [&](int r) { return r < 0 || r >= ns ? true : loc_spn[r] == 1; };
  const long long icnt_max = (long long)ns * ns;
  long long icnt_e = 0;
  mi = s = ri = rj = 0;
  for (;;) {
    if (icnt_e == icnt_max + 1) {
      bool any = false;
      for (int k = 0; k < 2 * ne; ++k) any = any || !is_local(ele_idx[k] < 0 ? 0 : ele_idx[k]);
      if (!any) return true;
    }
    ++icnt_e;
    mi = (int)rng.mod((uint32_t)ne);
    s = rng.real2() < 0.5 ? 0 : 1;
    ri = ele_idx[mi + s * ne];
    if (!is_local(ri)) break;
  }
  long long icnt = 0;
  for (;;) {
    rj = (int)rng.mod((uint32_t)ns);
    if (icnt > icnt_max) return true;
    ++icnt;
    if (ele_cfg[rj + s * ns] == -1 && !is_local(rj)) break;
  }
  return false;
}

// FSZ candidates (the sampler that uses them is not built; restated here so
// their draw order is pinned ahead of it).  Lookups outside the arrays read
// as "local site" and "not empty", as the reference's bounds-checked reads.
// makeCandidate_hopping_fsz / _csz (candidate.rs:387-492): source drawn over
// all n_size electrons until itinerant (no cap), spin ele_spn[mi]; then
// destination site and, with 2Sz = -1, target spin; cap tested after each
// draw, then occupancy.
template <class RNG>
static bool hop_candidate_fsz(int ns, int nsize, int two_sz, const int* loc_spn, const int* ele_idx,
                              const int* ele_cfg, const int* ele_spn, RNG& rng, int& mi, int& s, int& t, int& ri,
                              int& rj) {
  auto loc = [&](int r) { return (r < 0 || r >= ns) ? 1 : loc_spn[r]; };
  auto cfg = [&](int k) { return (k < 0 || k >= 2 * ns) ? 0 : ele_cfg[k]; };
  for (;;) {
    mi = (int)rng.mod((uint32_t)nsize);
    s = std::min(std::max(ele_spn[mi], 0), 1);
    ri = std::max(ele_idx[mi], 0);
    if (loc(ri) == 0) break;
  }
  const long long icnt_max = (long long)ns * ns;
  long long icnt = 0;
  for (;;) {
    rj = (int)rng.mod((uint32_t)ns);
    t = two_sz == -1 ? (rng.real2() < 0.5 ? 0 : 1) : s;
    if (icnt > icnt_max) return true;
    ++icnt;
    if (cfg(rj + t * ns) == -1 && loc(rj) == 0) break;
  }
  return false;
}

// makeCandidate_LocalSpinFlip_conduction (candidate.rs:494-566): draw an
// electron, flip its spin; cap tested after each draw, then "itinerant site
// with the flipped spin-site empty".  rj = ri.
template <class RNG>
static bool spin_flip_conduction(int ns, int nsize, const int* loc_spn, const int* ele_idx, const int* ele_cfg,
                                 const int* ele_spn, RNG& rng, int& mi, int& s, int& t, int& ri) {
  auto loc = [&](int r) { return (r < 0 || r >= ns) ? 1 : loc_spn[r]; };
  auto cfg = [&](int k) { return (k < 0 || k >= 2 * ns) ? 0 : ele_cfg[k]; };
  const long long icnt_max = (long long)ns * ns;
  long long icnt = 0;
  for (;;) {
    mi = (int)rng.mod((uint32_t)nsize);
    s = std::min(std::max(ele_spn[mi], 0), 1);
    t = 1 - s;
    ri = std::max(ele_idx[mi], 0);
    if (icnt > icnt_max) return true;
    ++icnt;
    if (loc(ri) == 0 && cfg(ri + t * ns) == -1) break;
  }
  return false;
}

// Raw-word tape with SFMT's draw conversions (mod: word % n, real2: word 2^-32).
struct WordTape {
  const uint32_t* w;
  int n, pos = 0;
  uint32_t gen32() {
    if (pos >= n) throw std::runtime_error("word tape exhausted");
    return w[pos++];
  }
  double real2() { return (double)gen32() * (1.0 / 4294967296.0); }
  uint32_t mod(uint32_t k) { return gen32() % k; }
};

static void run_block(const Run& R, Walker& W, int n_in, double& log_ip_old, int& n_accept) {
  const Model& m = R.m;
  const int ns = m.nsite(), ne = m.ne(), nqp = m.nqp(), np = m.n_proj();
  vector<double> pf_new(nqp);
  vector<int> cnew(np), mid(np);
  for (int step = 0; step < n_in; ++step) {
    int type;  // 0 hop, 1 exchange, 3 none
    switch (m.mp.nex_update_path) {
      case 0: type = 0; break;
      case 1: type = W.rng.real2() < 0.5 ? 1 : 0; break;
      case 2: type = 1; break;
      default: type = 3;
    }
    if (type == 0) {
      W.counter[0]++;
      int mi, s, ri, rj;
      if (hop_candidate(ns, ne, m.loc_spn.data(), W.ele_idx.data(), W.ele_cfg.data(), W.rng, mi, s, ri, rj)) continue;
      update_ele_config(m, mi, ri, rj, s, W.ele_idx.data(), W.ele_cfg.data(), W.ele_num.data());
      std::copy(W.proj_cnt.begin(), W.proj_cnt.end(), cnew.begin());
      update_proj_cnt(m, ri, rj, cnew.data(), W.ele_num.data());
      new_pf_hop(R, W, mi, s, pf_new.data());
      const double lip_new = log_ip_of(m, pf_new.data());
      const double lpd = log_proj_ratio(m, cnew.data(), W.proj_cnt.data());
      if (metropolis(W, lpd, lip_new, log_ip_old)) {
        update_hop(R, W, mi, s);
        std::copy(pf_new.begin(), pf_new.end(), W.pf.begin());
        std::copy(cnew.begin(), cnew.end(), W.proj_cnt.begin());
        log_ip_old = lip_new;
        ++n_accept;
        ++W.accepted;
        W.counter[1]++;
      } else {
        revert_ele_config(m, mi, ri, rj, s, W.ele_idx.data(), W.ele_cfg.data(), W.ele_num.data());
      }
    } else if (type == 1) {
      W.counter[2]++;
      bool single = false;
      for (int ri = 0; ri < ns; ++ri)
        if (W.ele_num[ri] + W.ele_num[ri + ns] == 1) { single = true; break; }
      if (!single) continue;
      int mi, s, ri, mj, rj;
      for (;;) {
        mi = (int)W.rng.mod((uint32_t)ne);
        s = W.rng.real2() < 0.5 ? 0 : 1;
        ri = W.ele_idx[mi + s * ne];
        if (W.ele_cfg[ri + (1 - s) * ns] == -1) break;
      }
      const int t = 1 - s;
      for (;;) {
        mj = (int)W.rng.mod((uint32_t)ne);
        rj = W.ele_idx[mj + t * ne];
        if (W.ele_cfg[rj + s * ns] == -1) break;
      }
      update_ele_config(m, mi, ri, rj, s, W.ele_idx.data(), W.ele_cfg.data(), W.ele_num.data());
      std::copy(W.proj_cnt.begin(), W.proj_cnt.end(), cnew.begin());
      update_proj_cnt(m, ri, rj, cnew.data(), W.ele_num.data());
      update_ele_config(m, mj, rj, ri, t, W.ele_idx.data(), W.ele_cfg.data(), W.ele_num.data());
      update_proj_cnt(m, rj, ri, cnew.data(), W.ele_num.data());
      new_pf_exchange(R, W, mi, s, mj, t, pf_new.data());
      const double lip_new = log_ip_of(m, pf_new.data());
      const double lpd = log_proj_ratio(m, cnew.data(), W.proj_cnt.data());
      if (metropolis(W, lpd, lip_new, log_ip_old)) {
        update_exchange(R, W, mi, s, mj, t, ri, rj);
        std::copy(cnew.begin(), cnew.end(), W.proj_cnt.begin());
        log_ip_old = lip_new;
        ++n_accept;
        ++W.accepted;
        W.counter[3]++;
      } else {
        revert_ele_config(m, mj, rj, ri, t, W.ele_idx.data(), W.ele_cfg.data(), W.ele_num.data());
        revert_ele_config(m, mi, ri, rj, s, W.ele_idx.data(), W.ele_cfg.data(), W.ele_num.data());
      }
    } else {
      // no update; the recompute guard below still runs
    }
    if (n_accept > ns) {
      int failed = calc_m_all(R, W.ele_idx.data(), W.inv.data(), W.pf.data(), false);
      if (!failed) log_ip_old = log_ip_of(m, W.pf.data());
      n_accept = 0;
    }
  }
}

static void sample_walker(const Run& R, Walker& W) {
  const Model& m = R.m;
  bool burn = W.counter[9] != 0;
  if (burn) {
    W.ele_idx = W.burn_idx;
    W.ele_cfg = W.burn_cfg;
    W.ele_num = W.burn_num;
    W.proj_cnt = W.burn_proj;
  } else {
    make_initial_sample_validated(R, W);
  }
  calc_m_all(R, W.ele_idx.data(), W.inv.data(), W.pf.data(), true);
  double log_ip = log_ip_of(m, W.pf.data());
  if (!std::isfinite(log_ip)) {
    make_initial_sample_validated(R, W);
    calc_m_all(R, W.ele_idx.data(), W.inv.data(), W.pf.data(), true);
    log_ip = log_ip_of(m, W.pf.data());
    burn = false;
  }
  const int nsample = m.mp.nvmc_sample;
  const int n_out = burn ? nsample + 1 : m.mp.nvmc_warmup + nsample;
  const int n_in = m.mp.nvmc_interval * (m.nsite() > 1 ? m.nsite() : 1);
  for (auto& c : W.counter) c = 0;
  int n_accept = 0;
  const int n = m.nsize(), n2 = m.nsite2(), np = m.n_proj();
  W.s_idx.assign((size_t)nsample * n, 0);
  W.s_cfg.assign((size_t)nsample * n2, 0);
  W.s_num.assign((size_t)nsample * n2, 0);
  W.s_proj.assign((size_t)nsample * np, 0);
  W.n_saved = 0;
  for (int out = 0; out < n_out; ++out) {
    run_block(R, W, n_in, log_ip, n_accept);
    if (out + nsample >= n_out) {
      int smp = out + nsample - n_out;
      std::copy(W.ele_idx.begin(), W.ele_idx.end(), W.s_idx.begin() + (size_t)smp * n);
      std::copy(W.ele_cfg.begin(), W.ele_cfg.end(), W.s_cfg.begin() + (size_t)smp * n2);
      std::copy(W.ele_num.begin(), W.ele_num.end(), W.s_num.begin() + (size_t)smp * n2);
      std::copy(W.proj_cnt.begin(), W.proj_cnt.end(), W.s_proj.begin() + (size_t)smp * np);
      W.n_saved++;
    }
  }
  W.burn_idx = W.ele_idx;
  W.burn_cfg = W.ele_cfg;
  W.burn_num = W.ele_num;
  W.burn_proj = W.proj_cnt;
  W.counter[9] = 1;
}

// ---------------------------------------------------------------------------
// Measurement: crates/mvmc-core/src/run.rs:2937-3616,
// crates/mvmc-core/src/observables.rs:128-366, 483-552, 1638-1768,
// crates/mvmc-core/src/slater_derivative.rs:75-99, 189-285, 490-548
// ---------------------------------------------------------------------------
// Diagonal terms: observables.rs:483-552
static double energy_diagonal(const Model& m, const int* ele_num) {
  const int ns = m.nsite();
  const int* n0 = ele_num;
  const int* n1 = ele_num + ns;
  double e = 0.0;
  for (auto& t : m.coulomb_intra) e += t.second * (double)(n0[t.first] * n1[t.first]);
  for (auto& t : m.coulomb_inter) {
    double oi = (double)(n0[t.site1] + n1[t.site1]), oj = (double)(n0[t.site2] + n1[t.site2]);
    e += t.value * oi * oj;
  }
  for (auto& t : m.hund) {
    double su = (double)(n0[t.site1] * n0[t.site2]), sd = (double)(n1[t.site1] * n1[t.site2]);
    e -= t.value * (su + sd);
  }
  return e;
}

// observables.rs:1638-1672 calh1_direct_projection_ratio_moved
static double direct_proj_ratio_moved(const Run& R, int source, int dest, int spin, const int* ele_num) {
  const int n = R.m.nsite();
  const int src_slot = source + spin * n, dst_slot = dest + spin * n;
  auto occ = [&](int slot) { return slot == src_slot ? 0 : (slot == dst_slot ? 1 : ele_num[slot]); };
  auto charge_moved = [&](int site) { return occ(site) + occ(n + site) - 1; };
  auto charge = [&](int site) { return ele_num[site] + ele_num[n + site] - 1; };
  const double* row_dest = R.dproj_j.data() + (size_t)dest * n;
  const double* row_source = R.dproj_j.data() + (size_t)source * n;
  double z = 0.0;
  z -= R.dproj_g[source] * (double)(occ(source) + occ(n + source));
  z += R.dproj_g[dest] * (double)(occ(dest) * occ(n + dest));
  z += row_source[dest] * (double)(charge_moved(source) - charge_moved(dest) + 1);
  for (int site = 0; site < n; ++site)
    if (site != source && site != dest) z += (row_dest[site] - row_source[site]) * (double)charge(site);
  return o_exp(z);
}

// updates.rs:168-262 calculate_new_pf_m2_ip_real_flat
static double new_ip_hop_moved(const Run& R, const int* ele_idx, const double* inv, const double* pf, int mj,
                               int spin, int moved_slot, int moved_site) {
  const Model& m = R.m;
  const int n = m.nsize(), ne = m.ne(), ns = m.nsite(), n2 = m.nsite2();
  const int msa = mj + spin * ne;
  auto site_of = [&](int msj) { return msj == moved_slot ? moved_site : ele_idx[msj]; };
  const int rsa = site_of(msa) + spin * ns;
  double ip = 0.0;
  for (int q = 0; q < m.nqp(); ++q) {
    const double* row = R.slater.data() + ((size_t)q * n2 + rsa) * n2;
    const double* iv = inv + (size_t)q * n * n + (size_t)msa * n;
    double ratio = 0.0;
    for (int msj = 0; msj < n; ++msj) ratio += iv[msj] * row[site_of(msj) + (msj < ne ? 0 : ns)];
    const double npf = -ratio * pf[q];
    ip += m.qp_full_weight[q] * npf;
  }
  return ip;
}

// observables.rs:1680-1755 calh1_real_fast
static double transfer_energy(const Run& R, double ip, const int* ele_idx, const int* ele_cfg, const int* ele_num,
                              const double* inv, const double* pf) {
  const Model& m = R.m;
  const int ns = m.nsite(), ne = m.ne();
  double te = 0.0;
  for (auto& t : m.transfer) {
    const int ri = t.site1, rj = t.site2, s1 = t.spin1, s2 = t.spin2;
    const int dst = ri + s1 * ns, src = rj + s2 * ns;
    double g1;
    if (s1 == s2 && ri == rj) g1 = (double)ele_num[src];
    else if (s1 != s2 || ele_num[dst] == 1 || ele_num[src] == 0 || ele_cfg[src] < 0) g1 = 0.0;
    else {
      const int mj = ele_cfg[src];
      const int msj = mj + s2 * ne;
      const double pr = direct_proj_ratio_moved(R, rj, ri, s2, ele_num);
      const double nip = new_ip_hop_moved(R, ele_idx, inv, pf, mj, s2, msj, ri);
      g1 = pr * nip / ip;
    }
    te -= t.re * g1;
  }
  return te;
}

// Green function ratio of the exchange term c+_{ri,up} c_{rj,up} c+_{rj,dn} c_{ri,dn}
// (observables.rs:558-880, general case; coincident-index branches reduce
// to zero for ri != rj).
static double green_exchange(const Run& R, int ri, int rj, int spin, int spin_other, double ip, const int* ele_idx,
                             const int* ele_cfg, const int* ele_num, const int* proj_cnt, const double* inv,
                             const double* pf) {
  const Model& m = R.m;
  const int ns = m.nsite(), ne = m.ne();
  const int rk = rj, rl = ri;
  if (ri == rj) return 0.0;
  if (ele_num[ri + spin * ns] != 0 || ele_num[rj + spin * ns] != 1 || ele_num[rk + spin_other * ns] != 0 ||
      ele_num[rl + spin_other * ns] != 1)
    return 0.0;
  const int mj = ele_cfg[rj + spin * ns], mi = ele_cfg[rl + spin_other * ns];
  if (mj < 0 || mi < 0) return 0.0;
  const int msj = mj + spin * ne, msi = mi + spin_other * ne;
  Walker tmp;
  tmp.ele_idx.assign(ele_idx, ele_idx + m.nsize());
  vector<int> num(ele_num, ele_num + 2 * ns);
  vector<int> cnt(proj_cnt, proj_cnt + m.n_proj());
  tmp.ele_idx[msi] = rk;
  num[rl + spin_other * ns] = 0;
  num[rk + spin_other * ns] = 1;
  update_proj_cnt(m, rl, rk, cnt.data(), num.data());
  tmp.ele_idx[msj] = ri;
  num[rj + spin * ns] = 0;
  num[ri + spin * ns] = 1;
  update_proj_cnt(m, rj, ri, cnt.data(), num.data());
  const double lpd = log_proj_ratio(m, cnt.data(), proj_cnt);
  const double proj_ratio = o_exp(lpd);
  tmp.inv.assign(inv, inv + (size_t)m.nqp() * m.nsize() * m.nsize());
  tmp.pf.assign(pf, pf + m.nqp());
  vector<double> pf_new(m.nqp());
  new_pf_exchange(R, tmp, mi, spin_other, mj, spin, pf_new.data());
  const double nip = calc_ip(m, pf_new.data());
  // Real operands: Julia's complex quotient (proj_ratio*nip)/ip with zero
  // imaginary parts reduces to the real quotient (Smith's algorithm, d = 0).
  return (proj_ratio * nip) / ip;
}

static double local_energy(const Run& R, double ip, const int* ele_idx, const int* ele_cfg, const int* ele_num,
                           const int* proj_cnt, const double* inv, const double* pf) {
  const Model& m = R.m;
  double e = energy_diagonal(m, ele_num);
  if (!m.transfer.empty() && std::fabs(ip) > 0.0) e += transfer_energy(R, ip, ele_idx, ele_cfg, ele_num, inv, pf);
  if (!m.exchange.empty() && std::fabs(ip) > 0.0)
    for (auto& t : m.exchange) {
      double g01 = green_exchange(R, t.site1, t.site2, 0, 1, ip, ele_idx, ele_cfg, ele_num, proj_cnt, inv, pf);
      double g10 = green_exchange(R, t.site1, t.site2, 1, 0, ip, ele_idx, ele_cfg, ele_num, proj_cnt, inv, pf);
      e += t.value * (g01 + g10);
    }
  return e;
}

// Julia's complex reciprocal for a real operand (crates/pfapack/src/julia_complex.rs:78-110).
static double julia_reciprocal_re(double c) {
  const double cd = std::fabs(c);
  if (std::sqrt(2.2250738585072014e-308 / 2.0) <= cd && cd <= std::sqrt(1.7976931348623157e308 / 2.0)) {
    const double norm = std::fma(cd, cd, 0.0 * 0.0);
    return c / norm;
  }
  return 1.0 / c;
}

// slater_derivative.rs:490-548 + :75-99 (real fast path).
static void slater_diff(const Run& R, double ip, const int* ele_idx, const double* inv, const double* pf,
                        double* out) {
  const Model& m = R.m;
  const int n = m.nsize(), ne = m.ne(), ns = m.nsite(), nleg = m.n_sp_gl(), nqp = m.nqp(), nsl = m.n_slater;
  vector<double> qo((size_t)nqp * nsl, 0.0);
  vector<int> tri(n), sgi(n);
  for (int qp = 0; qp < nqp; ++qp) {
    const int t = qp / nleg, sp = qp % nleg;
    for (int msi = 0; msi < n; ++msi) {
      tri[msi] = R.tsite[(size_t)t * ns + ele_idx[msi]];
      sgi[msi] = R.tsgn[(size_t)t * ns + ele_idx[msi]];
    }
    const double p = pf[qp];
    const double cs = p * m.spgl_cos_sin[sp], cc = p * m.spgl_cos_cos[sp], ss = p * m.spgl_sin_sin[sp];
    const double* plane = inv + (size_t)qp * n * n;
    double* o = qo.data() + (size_t)qp * nsl;
    for (int msi = 0; msi < n; ++msi) {
      const double first = msi < ne ? cs : ss;
      const double second = msi < ne ? cc : cs;
      for (int msj = 0; msj < n; ++msj) {
        const int orb = m.orb_idx[(size_t)tri[msi] * ns + tri[msj]];
        if (orb < 0 || orb >= nsl) continue;
        const int sg = sgi[msi] * sgi[msj] * m.orb_sgn[(size_t)tri[msi] * ns + tri[msj]];
        const double x = plane[msj + (size_t)msi * n];
        const double value = msj < ne ? x * first : -x * second;
        o[orb] += value * (double)sg;
      }
    }
  }
  const double inv_ip = julia_reciprocal_re(ip);
  for (int orb = 0; orb < nsl; ++orb) {
    double acc = 0.0;
    for (int qp = 0; qp < nqp; ++qp) acc += m.qp_full_weight[qp] * qo[(size_t)qp * nsl + orb];
    out[orb] = acc * inv_ip;
  }
}

static void measure_walker(const Run& R, Walker& W) {
  const Model& m = R.m;
  const int n = m.nsize(), n2 = m.nsite2(), np = m.n_proj(), size = m.sr_opt_size(), nqp = m.nqp();
  const int nsample = m.mp.nvmc_sample;
  const bool use_store = m.mp.nstore != 0 || m.mp.nsrcg == 1;
  W.wc = W.etot = W.etot2 = W.sztot = W.sztot2 = 0.0;
  W.ho.assign(size, 0.0);
  W.oo.assign(m.mp.nsrcg ? 2 * (size_t)size : (size_t)size * size, 0.0);
  if (use_store) W.o_store.assign((size_t)nsample * size, 0.0);
  W.o_last.assign(size, 0.0);
  vector<double> inv((size_t)nqp * n * n), pf(nqp), o(size);
  for (int smp = 0; smp < W.n_saved; ++smp) {
    const int* ei = W.s_idx.data() + (size_t)smp * n;
    const int* ec = W.s_cfg.data() + (size_t)smp * n2;
    const int* en = W.s_num.data() + (size_t)smp * n2;
    const int* epc = W.s_proj.data() + (size_t)smp * np;
    bool all0 = true, allneg = true;
    for (int k = 0; k < n; ++k) { all0 = all0 && ei[k] == 0; allneg = allneg && ei[k] < 0; }
    if (all0 || allneg) continue;
    if (calc_m_all(R, ei, inv.data(), pf.data(), false)) continue;
    const double ip = calc_ip(m, pf.data());
    if (std::fabs(ip) < 1.0e-100) continue;
    const double w = 1.0;
    const double e = local_energy(R, ip, ei, ec, en, epc, inv.data(), pf.data());
    if (!std::isfinite(e + 0.0)) continue;
    double sz = 0.0;
    {
      long long s = 0;
      for (int ri = 0; ri < m.nsite(); ++ri) s += en[ri] - en[ri + m.nsite()];
      sz = (double)s;
    }
    W.wc += w;
    W.etot += w * e;
    W.etot2 += (w * e) * e;
    W.sztot += w * sz;
    W.sztot2 += w * sz * sz;
    if (m.mp.vmc_calc_mode != 0) continue;
    o[0] = 1.0;
    for (int i = 0; i < np; ++i) o[1 + i] = (double)epc[i];
    if (m.n_slater > 0) slater_diff(R, ip, ei, inv.data(), pf.data(), o.data() + 1 + np);
    W.o_last = o;
    const double we = w * e;
    if (use_store) {
      const double sw = std::sqrt(w);
      for (int i = 0; i < size; ++i) {
        W.o_store[(size_t)smp * size + i] = sw * o[i];
        W.ho[i] += we * o[i];
      }
    } else {
      for (int j = 0; j < size; ++j)
        for (int i = 0; i < size; ++i) W.oo[(size_t)j * size + i] += w * o[i] * o[j];
      for (int i = 0; i < size; ++i) W.ho[i] += we * o[i];
    }
  }
  if (use_store && m.mp.vmc_calc_mode == 0 && m.mp.nsrcg == 0) {
    // C calculateOO_Store_real (vmccal.c:714-718, sr_backend.rs:305-333): the
    // full Gram of the stored rows, each entry summed over the samples in
    // order (DGEMM's own order is the BLAS build's)
    for (int j = 0; j < size; ++j)
      for (int i = 0; i < size; ++i) {
        double acc = 0.0;
        for (int s = 0; s < W.n_saved; ++s)
          acc += W.o_store[(size_t)s * size + i] * W.o_store[(size_t)s * size + j];
        W.oo[(size_t)j * size + i] = acc;
      }
  }
  if (use_store && m.mp.vmc_calc_mode == 0 && m.mp.nsrcg != 0) {
    for (int i = 0; i < size; ++i) {
      double mean = 0.0, diag = 0.0;
      for (int s = 0; s < nsample; ++s) {
        double v = W.o_store[(size_t)s * size + i];
        mean += v;
        diag += v * v;
      }
      W.oo[i] = mean;
      W.oo[size + i] = diag;
    }
  }
}

// ---------------------------------------------------------------------------
// SR-CG: crates/mvmc-core/src/sr_cg.rs:201-484 (driver), :611-1004 (solver)
// ---------------------------------------------------------------------------
struct CgOperator {
  int n = 0, samples = 0;
  vector<double> mean, diag;
  vector<const double*> parts;   // per-walker sample blocks [samples][n], walker (= rank) order
  vector<int> part_samples;
};

// Reduction order of the SR-CG solver.  The reference evaluates S x with
// BLAS dgemv (sr_backend.rs:520-597), whose summation order belongs to the
// BLAS build, and reduces over ranks with MPI; this order is therefore a
// choice, fixed here and reproduced by the device kernels:
//   bdot(a, b): 256 strided partial sums p_l = sum_{i = l mod 256} a_i b_i
//               (ascending i), then the pairwise tree p_l += p_{l+h},
//               h = 128, 64, ..., 1; the result is p_0.
//   S x:        y_r = bdot(O_r, x) per sample row r; rows are cut into chunks
//               of CG_RCH consecutive rows (walker order); per chunk
//               c_i = sum_r O_ri y_r (ascending r); chunks are folded left
//               within superchunks of CG_SCH chunks, superchunk results folded
//               left in order.
constexpr int BD_LANES = 256;
constexpr int CG_RCH = 32;
constexpr int CG_SCH = 16;

static double bdot(const double* a, const double* b, int n) {
  double p[BD_LANES];
  for (int l = 0; l < BD_LANES; ++l) {
    double acc = 0.0;
    for (int i = l; i < n; i += BD_LANES) acc += a[i] * b[i];
    p[l] = acc;
  }
  for (int h = BD_LANES / 2; h >= 1; h >>= 1)
    for (int l = 0; l < h; ++l) p[l] += p[l + h];
  return p[0];
}

// z = inv_w (O^T O x) - (mean.x) mean + shift diag x (sr_cg.rs:952-1004).
static void cg_apply(const CgOperator& op, const double* x, double* z, double inv_w, double shift) {
  const int n = op.n;
  vector<const double*> rows;
  for (size_t part = 0; part < op.parts.size(); ++part)
    for (int s = 0; s < op.part_samples[part]; ++s) rows.push_back(op.parts[part] + (size_t)s * n);
  const int nrows = (int)rows.size();
  const int nch = (nrows + CG_RCH - 1) / CG_RCH;
  vector<double> tot(n, 0.0), sup(n, 0.0), c(n);
  for (int b = 0; b < nch; ++b) {
    std::fill(c.begin(), c.end(), 0.0);
    for (int r = b * CG_RCH; r < std::min(nrows, (b + 1) * CG_RCH); ++r) {
      const double y = bdot(rows[r], x, n);
      for (int i = 0; i < n; ++i) c[i] += rows[r][i] * y;
    }
    if (b % CG_SCH == 0) sup = c;
    else
      for (int i = 0; i < n; ++i) sup[i] += c[i];
    if (b % CG_SCH == CG_SCH - 1 || b == nch - 1) {
      if (b < CG_SCH) tot = sup;
      else
        for (int i = 0; i < n; ++i) tot[i] += sup[i];
    }
  }
  const double coef = bdot(op.mean.data(), x, n);
  for (int i = 0; i < n; ++i) z[i] = inv_w * tot[i] - coef * op.mean[i] + shift * op.diag[i] * x[i];
}

static long long cg_solve(const CgOperator& op, const double* g, double inv_w, double shift, double tol,
                          int max_it, bool diag_scale, double* x, double* r_out = nullptr,
                          double* d_out = nullptr) {
  const int n = op.n;
  const double CG_TINY = 1.0e-20, EPS = 1.0e-12;
  const double threshold = tol * tol * (double)n * (double)n;
  const double pfac = 1.0 + shift;
  auto precond = [&](int i, double r) {
    double dq = pfac * op.diag[i];
    if (!std::isfinite(dq) || std::fabs(dq) < EPS) dq = dq < 0.0 ? -EPS : EPS;
    return r / dq;
  };
  vector<double> r(g, g + n), q(n, 0.0), d;
  for (int i = 0; i < n; ++i) x[i] = 0.0;
  double delta;
  if (diag_scale) {
    cg_apply(op, x, r.data(), inv_w, shift);
    for (int i = 0; i < n; ++i) r[i] = g[i] - r[i];
    d.assign(n, 0.0);
    for (int i = 0; i < n; ++i) d[i] = precond(i, r[i]);
    delta = bdot(r.data(), d.data(), n);
  } else {
    d.assign(g, g + n);
    delta = bdot(r.data(), r.data(), n);
  }
  auto finish = [&](long long st) {
    if (r_out) std::copy(r.begin(), r.end(), r_out);
    if (d_out) std::copy(d.begin(), d.end(), d_out);
    return st;
  };
  if (!std::isfinite(delta)) return finish(-1);
  int iter = 0;
  while (iter < max_it) {
    const long long failed = -(long long)(iter + 1);
    if (!std::isfinite(delta)) return finish(failed);
    if (std::fabs(delta) < threshold) break;
    cg_apply(op, d.data(), q.data(), inv_w, shift);
    const double dq = bdot(d.data(), q.data(), n);
    if (!std::isfinite(dq) || std::fabs(dq) < CG_TINY) return finish(failed);
    const double alpha = delta / dq;
    if (!std::isfinite(alpha)) return finish(failed);
    for (int i = 0; i < n; ++i) x[i] += alpha * d[i];
    if ((iter + 1) % 20 == 0) {
      cg_apply(op, x, r.data(), inv_w, shift);
      for (int i = 0; i < n; ++i) r[i] = g[i] - r[i];
    } else {
      for (int i = 0; i < n; ++i) r[i] -= alpha * q[i];
    }
    double dnew;
    if (diag_scale) {
      for (int i = 0; i < n; ++i) q[i] = precond(i, r[i]);
      dnew = bdot(r.data(), q.data(), n);
    } else {
      dnew = bdot(r.data(), r.data(), n);
    }
    if (!std::isfinite(dnew) || std::fabs(delta) < CG_TINY) return finish(failed);
    const double beta = dnew / delta;
    if (!std::isfinite(beta)) return finish(failed);
    if (diag_scale) {
      delta = dnew;
      for (int i = 0; i < n; ++i) d[i] = q[i] + beta * d[i];
    } else {
      delta *= beta;
      for (int i = 0; i < n; ++i) d[i] = r[i] + beta * d[i];
    }
    ++iter;
  }
  if (iter == max_it && std::fabs(delta) >= threshold) return finish(-(long long)(max_it + 1));
  return finish(iter);
}

// sr_cg.rs:201-484 for real parameters (offset 1, no RescaleSmat).
static int stochastic_opt_cg(Run& R) {
  Model& m = R.m;
  const int np = m.n_para(), size = m.sr_opt_size();
  const vector<double>& oo = R.oo;
  const vector<double>& ho = R.ho;
  vector<double> var(np);
  bool nan = false;
  for (int pi = 0; pi < np; ++pi) {
    const int idx = pi + 1;
    var[pi] = oo[size + idx] - oo[idx] * oo[idx];
    nan = nan || std::isnan(var[pi]);
  }
  double vmax = -INFINITY;
  if (nan) vmax = NAN;
  else
    for (double v : var) vmax = std::fmax(vmax, v);
  const double thr = vmax * m.mp.dsr_opt_red_cut;
  vector<int> map;
  for (int pi = 0; pi < np; ++pi) {
    if (m.opt_flags[2 * pi] != 1) continue;
    if (var[pi] < thr) continue;
    map.push_back(pi);
  }
  if (map.empty()) return 0;
  const int na = (int)map.size();
  CgOperator op;
  op.n = na;
  op.mean.resize(na);
  op.diag.resize(na);
  vector<double> g(na);
  const double dt = 2.0 * m.mp.dsr_opt_step_dt;
  for (int si = 0; si < na; ++si) {
    op.mean[si] = oo[map[si] + 1];
    op.diag[si] = var[map[si]];
    const int idx = map[si] + 1;
    g[si] = -dt * (ho[idx] - ho[0] * oo[idx]);
  }
  // Per-walker packed samples [samples][active], walker order (= rank order).
  vector<vector<double>> packs(R.w.size());
  const int nsample = m.mp.nvmc_sample;
  for (size_t wi = 0; wi < R.w.size(); ++wi) {
    packs[wi].assign((size_t)nsample * na, 0.0);
    for (int s = 0; s < nsample; ++s)
      for (int si = 0; si < na; ++si) packs[wi][(size_t)s * na + si] = R.w[wi].o_store[(size_t)s * size + map[si] + 1];
    op.parts.push_back(packs[wi].data());
    op.part_samples.push_back(nsample);
  }
  const int max_it = m.mp.nsr_opt_cg_max_iter > 0 ? m.mp.nsr_opt_cg_max_iter : na;
  const double inv_w = 1.0 / R.wc;
  vector<double> x(na);
  long long st = cg_solve(op, g.data(), inv_w, m.mp.dsr_opt_sta_del, m.mp.dsr_opt_cg_tol, max_it,
                          m.mp.use_diag_scale != 0, x.data());
  long long info = st;
  if (st < 0) {
    if (m.mp.nsrcg_fallback) {
      long long st2 = cg_solve(op, g.data(), inv_w, m.mp.dsr_opt_sta_del, m.mp.dsr_opt_cg_tol, max_it,
                               m.mp.use_diag_scale == 0, x.data());
      info = -((-st - 1) + 1);
      if (st2 < 0 && m.mp.nsrcg_abort_on_fail) throw std::runtime_error("SR-CG fallback failed");
    } else {
      if (m.mp.nsrcg_abort_on_fail) throw std::runtime_error("SR-CG failed (NSRCGAbortOnFail)");
      info = -((-st - 1) + 1);
    }
  }
  R.cg_status = (int)info;
  R.cg_solution = x;
  for (double v : x)
    if (!std::isfinite(v)) {
      if (m.mp.nsrcg_abort_on_fail) throw std::runtime_error("SR-CG produced non-finite updates");
      return 1;
    }
  for (int si = 0; si < na; ++si) {
    const int pi = map[si];
    if (pi < m.n_proj()) m.proj[pi] += x[si];
    else m.slater[pi - m.n_proj()] += x[si];
  }
  return 0;
}

// ---------------------------------------------------------------------------
// Direct SR (NSRCG 0) and the other C-order SrStages
// ---------------------------------------------------------------------------
// A real view of a real array or of the real parts of a complex one
// (sr_backend.rs:53-70 RealView).
struct RView {
  const double* p;
  int cplx;
  double at(size_t i) const { return cplx ? p[2 * i] : p[i]; }
};

// COrderSr::assemble_s_g (sr_backend.rs:484-516, C stcopt.c:69): column sj of S,
// S[si][sj] = oo[(pi+off) ld + (pj+off)] - oo[pi+off] oo[pj+off], the diagonal
// times 1 + sta_del; g[si] = (-2 dt) (ho[pi+off] - ho[0] oo[pi+off]).
static void sr_assemble_s_g(RView oo, RView ho, const int* map, int nm, size_t ld, int off, double sta_del,
                            double step_dt, double* s, double* g) {
  const double ratio_diag = 1.0 + sta_del;
  for (int sj = 0; sj < nm; ++sj) {
    const size_t pj = (size_t)map[sj];
    for (int si = 0; si < nm; ++si) {
      const size_t pi = (size_t)map[si];
      const double tmp = oo.at(pi + off);
      double v = oo.at((pi + off) * ld + (pj + off)) - tmp * oo.at(pj + off);
      if (si == sj) v *= ratio_diag;
      s[(size_t)sj * nm + si] = v;
    }
  }
  const double ho0 = ho.at(0);
  for (int si = 0; si < nm; ++si) {
    const size_t pi = (size_t)map[si];
    const double v = ho.at(pi + off) - ho0 * oo.at(pi + off);
    g[si] = -2.0 * step_dt * v;
  }
}

// sr.rs:535-577 cholesky_solve: LAPACK dpotrf('U') + dpotrs('U') restated in
// the reference-BLAS unblocked order (dpotf2: ddot, dgemv 'T', dscal by the
// reciprocal; dtrsm 'L','U','T','N' then 'L','U','N','N').  The reference
// links OpenBLAS, whose blocked kernels sum in another order: bit-level
// parity with it is unpinned (pinned to the C dposv solutions within
// 1e-15 + 1e-13 |x|, tests/golden/negative_stepdt_c_sr.txt).  Returns 0, or
// the factorisation's positive info, or -1 for a non-finite solution.
static int cholesky_solve(double* a, double* b, int n) {
  auto A = [&](int i, int j) -> double& { return a[i + (size_t)j * n]; };
  for (int j = 0; j < n; ++j) {
    double dot = 0.0;
    for (int k = 0; k < j; ++k) dot += A(k, j) * A(k, j);
    double ajj = A(j, j) - dot;
    if (ajj <= 0.0 || std::isnan(ajj)) {
      A(j, j) = ajj;
      return j + 1;
    }
    ajj = std::sqrt(ajj);
    A(j, j) = ajj;
    if (j + 1 < n) {
      for (int c = j + 1; c < n; ++c) {
        double t = 0.0;
        for (int k = 0; k < j; ++k) t += A(k, c) * A(k, j);
        A(j, c) = A(j, c) + (-1.0) * t;
      }
      const double r = 1.0 / ajj;
      for (int c = j + 1; c < n; ++c) A(j, c) = r * A(j, c);
    }
  }
  for (int i = 0; i < n; ++i) {  // U^T y = b
    double t = b[i];
    for (int k = 0; k < i; ++k) t -= A(k, i) * b[k];
    b[i] = t / A(i, i);
  }
  for (int k = n - 1; k >= 0; --k) {  // U x = y
    if (b[k] != 0.0) {
      b[k] = b[k] / A(k, k);
      for (int i = 0; i < k; ++i) b[i] -= b[k] * A(i, k);
    }
  }
  for (int i = 0; i < n; ++i)
    if (!std::isfinite(b[i])) return -1;
  return 0;
}

// COrderSr::gram_complex (sr_backend.rs:386-420, gram_block :349-384):
// G[i + j n] = sum_s O[i,s] conj(O[j,s]), samples in order, each product the
// num_complex one with separately rounded parts.  store: [samples][n] complex.
static void sr_gram_complex(const double* store, int n, int samples, double* out) {
  for (int j = 0; j < n; ++j)
    for (int i = 0; i < n; ++i) {
      double re = 0.0, im = 0.0;
      for (int s = 0; s < samples; ++s) {
        const double* o = store + 2 * (size_t)s * n;
        const double cr = o[2 * j], ci = -o[2 * j + 1];
        const double pr = o[2 * i] * cr - o[2 * i + 1] * ci;
        const double pim = o[2 * i] * ci + o[2 * i + 1] * cr;
        re += pr;
        im += pim;
      }
      out[2 * ((size_t)j * n + i)] = re;
      out[2 * ((size_t)j * n + i) + 1] = im;
    }
}

// COrderSr::sr_s_g (sr_backend.rs:436-460) over centered() (stage_backend.rs:167-184):
// d = O - <O>_w per column, de = E - <E>_w; S[p + q np] = sum_k (w_k d_kp) d_kq,
// g[q] = sum_k (w_k d_kq) de_k, k in order.  o: column-major [ns][np].
static void sr_s_g(const double* o, int ns, int np, const double* e, const double* w, double* s, double* g) {
  vector<double> d(o, o + (size_t)ns * np), de(ns);
  for (int p = 0; p < np; ++p) {
    double mean = 0.0;
    for (int k = 0; k < ns; ++k) mean += w[k] * o[k + (size_t)p * ns];
    for (int k = 0; k < ns; ++k) d[k + (size_t)p * ns] -= mean;
  }
  double emean = 0.0;
  for (int k = 0; k < ns; ++k) emean += w[k] * e[k];
  for (int k = 0; k < ns; ++k) de[k] = e[k] - emean;
  for (int q = 0; q < np; ++q) {
    for (int p = 0; p < np; ++p) {
      double acc = 0.0;
      for (int k = 0; k < ns; ++k) acc += w[k] * d[k + (size_t)p * ns] * d[k + (size_t)q * ns];
      s[p + (size_t)q * np] = acc;
    }
    double acc = 0.0;
    for (int k = 0; k < ns; ++k) acc += w[k] * d[k + (size_t)q * ns] * de[k];
    g[q] = acc;
  }
}

// stochastic_opt_real (sr.rs:60-160): redundancy cut on the diagonal of the
// full <OO> (collect_active_real / select_active_real, sr.rs:361-406: the
// maximum folded from 0.0), S/g assembly, Cholesky solve, parameter update.
static int stochastic_opt_direct(Run& R) {
  Model& m = R.m;
  const int np = m.n_para(), size = m.sr_opt_size();
  if (np == 0) return 0;
  const vector<double>& oo = R.oo;
  vector<double> sd(np, 0.0);
  bool nan = false;
  for (int pi = 0; pi < np; ++pi) {
    const size_t idx_diag = (size_t)(pi + 1) * size + (pi + 1), idx0 = (size_t)(pi + 1);
    if (idx_diag < oo.size() && idx0 < oo.size()) sd[pi] = oo[idx_diag] - oo[idx0] * oo[idx0];
    nan = nan || std::isnan(sd[pi]);
  }
  double vmax = 0.0;
  if (nan) vmax = NAN;
  else
    for (double v : sd) vmax = std::fmax(vmax, v);
  const double cut = vmax * m.mp.dsr_opt_red_cut;
  vector<int> map;
  for (int pi = 0; pi < np; ++pi) {
    if (m.opt_flags[2 * pi] != 1) continue;
    if (sd[pi] < cut) continue;
    map.push_back(pi);
  }
  if (map.empty()) return 0;
  const int nm = (int)map.size();
  vector<double> S((size_t)nm * nm), g(nm);
  sr_assemble_s_g(RView{oo.data(), 0}, RView{R.ho.data(), 0}, map.data(), nm, (size_t)size, 1, m.mp.dsr_opt_sta_del,
                  m.mp.dsr_opt_step_dt, S.data(), g.data());
  R.cg_status = 0;
  if (cholesky_solve(S.data(), g.data(), nm) != 0) {
    R.cg_solution.clear();
    return 1;
  }
  R.cg_solution = g;
  for (int si = 0; si < nm; ++si) {
    const int pi = map[si];
    if (pi < m.n_proj()) m.proj[pi] += g[si];
    else m.slater[pi - m.n_proj()] += g[si];
  }
  return 0;
}

// ---------------------------------------------------------------------------
// Optimisation driver: crates/mvmc-core/src/run.rs:1294-1475 with W walkers
// (multichain.rs:647-689, rank-order left-fold reductions)
// ---------------------------------------------------------------------------
// observables.rs:1551-1605 build_direct_projection_tables: per-site
// Gutzwiller values and the symmetric Jastrow value table (0 on the diagonal).
static void projection_tables(const Model& m, vector<double>& g, vector<double>& j) {
  const int n = m.nsite();
  g.assign(n, 0.0);
  j.assign((size_t)n * n, 0.0);
  for (int s = 0; s < n && m.n_gutz > 0; ++s) {
    const int k = m.gutz_idx[s];
    if (k >= 0 && k < m.n_gutz) g[s] = m.proj[k];
  }
  for (int a = 0; a < n && m.n_jast > 0; ++a)
    for (int b = 0; b < n; ++b) {
      const int k = a == b ? -1 : m.jast(a < b ? a : b, a < b ? b : a);
      if (k >= 0 && k < m.n_jast) j[(size_t)a * n + b] = m.proj[m.n_gutz + k];
    }
}

static void refresh_tables(Run& R) {
  slater_table(R.m, R.slater);
  projection_tables(R.m, R.dproj_g, R.dproj_j);
}

template <class F>
static void for_walkers(Run& R, F f) {
  const int T = R.threads > 1 ? R.threads : 1;
  if (T == 1) {
    for (size_t i = 0; i < R.w.size(); ++i) f(R.w[i]);
    return;
  }
  std::vector<std::thread> th;
  for (int t = 0; t < T; ++t)
    th.emplace_back([&, t] {
      for (size_t i = t; i < R.w.size(); i += T) f(R.w[i]);
    });
  for (auto& x : th) x.join();
}

static double now_s() {
  struct timespec ts;
  clock_gettime(CLOCK_MONOTONIC, &ts);
  return ts.tv_sec + 1e-9 * ts.tv_nsec;
}

static int opt_step(Run& R, bool do_sr) {
  Model& m = R.m;
  refresh_tables(R);
  double t0 = now_s();
  for_walkers(R, [&](Walker& W) { sample_walker(R, W); });
  double t1 = now_s();
  for_walkers(R, [&](Walker& W) { measure_walker(R, W); });
  double t2 = now_s();
  R.t_sample += t1 - t0;
  R.t_measure += t2 - t1;
  const int size = m.sr_opt_size();
  // rank-order reduction
  R.wc = R.etot = R.etot2 = R.sztot = R.sztot2 = 0.0;
  R.oo.assign(R.w[0].oo.size(), 0.0);
  R.ho.assign(size, 0.0);
  for (size_t wi = 0; wi < R.w.size(); ++wi) {
    const Walker& W = R.w[wi];
    if (wi == 0) {
      R.wc = W.wc; R.etot = W.etot; R.etot2 = W.etot2; R.sztot = W.sztot; R.sztot2 = W.sztot2;
      R.oo = W.oo;
      R.ho = W.ho;
    } else {
      R.wc += W.wc; R.etot += W.etot; R.etot2 += W.etot2; R.sztot += W.sztot; R.sztot2 += W.sztot2;
      for (size_t i = 0; i < R.oo.size(); ++i) R.oo[i] += W.oo[i];
      for (int i = 0; i < size; ++i) R.ho[i] += W.ho[i];
    }
  }
  // weight_average_we (complex (1,0)/wc) and weight_average_sr_opt_real (average.rs:18-63)
  if (std::fabs(R.wc) >= 1.0e-15) {
    const double inv_c = R.wc / (R.wc * R.wc + 0.0 * 0.0);
    R.etot *= inv_c;
    R.etot2 *= inv_c;
    R.sztot *= inv_c;
    R.sztot2 *= inv_c;
    const double inv = 1.0 / R.wc;
    const size_t active = m.mp.nsrcg ? 2 * (size_t)size : (size_t)size * size;
    for (size_t i = 0; i < active && i < R.oo.size(); ++i) R.oo[i] *= inv;
    for (int i = 0; i < size; ++i) R.ho[i] *= inv;
  }
  double var = 0.0;
  if (std::fabs(R.etot) > 1.0e-14) var = (R.etot2 - R.etot * R.etot) / (R.etot * R.etot);
  R.energies.insert(R.energies.end(), {R.etot, 0.0, R.etot2, var, R.sztot, R.sztot2});
  if (!do_sr) return 0;
  double t3 = now_s();
  int rc = 0;
  if (m.mp.nsrcg != 0) rc = stochastic_opt_cg(R);
  else rc = stochastic_opt_direct(R);
  R.t_sr += now_s() - t3;
  if (rc != 0) return rc;
  sync_parameters(m, true);
  return 0;
}

}  // namespace oracle

// ---------------------------------------------------------------------------
// C API used by tests/ (ctypes)
// ---------------------------------------------------------------------------
using namespace oracle;

extern "C" {

void oracle_set_math(int cr) { g_cr_math = cr; }

// SrStages of the C order on host arrays.
void oracle_sr_assemble_s_g(const double* oo, int oo_cplx, const double* ho, int ho_cplx, const int* map, int nm,
                            long long ld, int offset, double sta_del, double step_dt, double* s, double* g) {
  sr_assemble_s_g(RView{oo, oo_cplx}, RView{ho, ho_cplx}, map, nm, (size_t)ld, offset, sta_del, step_dt, s, g);
}
int oracle_cholesky_solve(double* s, double* rhs, int n) { return cholesky_solve(s, rhs, n); }
void oracle_sr_gram_complex(const double* store, int n, int samples, double* out) {
  sr_gram_complex(store, n, samples, out);
}
void oracle_sr_s_g(const double* o, int ns, int np, const double* e, const double* w, double* s, double* g) {
  sr_s_g(o, ns, np, e, w, s, g);
}
double oracle_cr_exp(double x) { return ddm::cr_exp(x); }
double oracle_cr_log(double x) { return ddm::cr_log(x); }
// Bulk evaluation: which = 0 exp (fast path + fallback), 1 exp accurate path
// only, 2 log, 3 log accurate path only.
void oracle_cr_eval(int which, const double* x, double* y, long long n) {
  for (long long i = 0; i < n; ++i) {
    switch (which) {
      case 0: y[i] = ddm::cr_exp(x[i]); break;
      case 1: y[i] = ddm::exp_accurate(x[i]); break;
      case 2: y[i] = ddm::cr_log(x[i]); break;
      default: y[i] = ddm::log_accurate(x[i]);
    }
  }
}

void oracle_sfmt_words(uint32_t seed, int count, uint32_t* out) {
  Sfmt r;
  r.init(seed);
  for (int i = 0; i < count; ++i) out[i] = r.gen32();
}

// BatchedEngine-style: planes [n,n,count] column-major -> pf, true inverse, status.
void oracle_pfinv(const double* planes, int n, int count, double* pf, double* inv, int* status) {
  std::vector<int> piv(n);
  for (int p = 0; p < count; ++p) {
    double* a = inv + (size_t)p * n * n;
    std::memcpy(a, planes + (size_t)p * n * n, sizeof(double) * n * n);
    int info = dsktf2(a, n, piv.data());
    if (info != 0) {
      status[p] = info;
      pf[p] = 0.0;
      std::memset(a, 0, sizeof(double) * n * n);
      continue;
    }
    double v = utu2pfa(a, n, piv.data());
    // a non-finite Pfaffian is reported with its computed value
    // (crates/mvmc-gpu/src/lib.rs:14-15, scalar.rs:84-87)
    if (!std::isfinite(v)) { status[p] = -1; pf[p] = v; std::memset(a, 0, sizeof(double) * n * n); continue; }
    pf[p] = v;
    utu2inv(a, n, piv.data());
    status[p] = 0;
  }
}

// BatchedEngine::run_c64 (crates/mvmc-gpu/src/scalar.rs:146-170): complex
// planes (re, im interleaved) -> pf, true inverse, status.  order 0: zsktf2
// (the run_c64 reference); order 1: zsktf2_c_compat (production calc_m_all).
void oracle_pfinv_c128(const double* planes, int n, int count, int order, double* pf, double* inv, int* status) {
  std::vector<int> piv(n);
  for (int p = 0; p < count; ++p) {
    zc* a = reinterpret_cast<zc*>(inv) + (size_t)p * n * n;
    std::memcpy(a, planes + (size_t)p * n * n * 2, sizeof(zc) * n * n);
    const int info = zsktf2(a, n, piv.data(), order);
    if (info != 0) {
      status[p] = info;
      pf[2 * p] = 0.0;
      pf[2 * p + 1] = 0.0;
      std::memset(a, 0, sizeof(zc) * n * n);
      continue;
    }
    const zc v = zutu2pfa(a, n, piv.data());
    pf[2 * p] = v.re;
    pf[2 * p + 1] = v.im;
    if (!(std::isfinite(v.re) && std::isfinite(v.im))) {
      status[p] = -1;
      std::memset(a, 0, sizeof(zc) * n * n);
      continue;
    }
    zutu2inv(a, n, piv.data());
    status[p] = 0;
  }
}

// Julia's complex division / reciprocal, elementwise over interleaved arrays
// (which 0: z / w, 1: 1 / w).
void oracle_julia_complex(int which, const double* z, const double* w, double* out, long long count) {
  for (long long i = 0; i < count; ++i) {
    const zc wi{w[2 * i], w[2 * i + 1]};
    const zc r = which == 0 ? julia_divide(zc{z[2 * i], z[2 * i + 1]}, wi) : julia_reciprocal(wi);
    out[2 * i] = r.re;
    out[2 * i + 1] = r.im;
  }
}

// Hopping candidate on a raw-word tape (the C retry-cap fixture of
// tests/fixtures/candidate): out8 = mi ri rj s t reject candidate_words
// words_after_metropolis (the Metropolis decision draws one more word when
// the candidate is not rejected; the tapes hold only the candidate's words).
// Returns -1 if the tape ran out.
int oracle_hop_candidate_tape(int ns, int ne, const int* loc_spn, const int* ele_idx, const int* ele_cfg,
                              const uint32_t* words, int nwords, int* out8) {
  WordTape tape{words, nwords};
  try {
    int mi, s, ri, rj;
    const bool rej = hop_candidate(ns, ne, loc_spn, ele_idx, ele_cfg, tape, mi, s, ri, rj);
    const int used = tape.pos;
    const int vals[8] = {mi, ri, rj, s, s, rej ? 1 : 0, used, used + (rej ? 0 : 1)};
    std::memcpy(out8, vals, sizeof(vals));
    return 0;
  } catch (const std::exception&) {
    return -1;
  }
}

// Overlap of one configuration from scratch: ip = sum_q w_q Pf_q (full
// calc_m_all, no ratio formula) and its projection counts.  Returns the
// calc_m_all status.
int oracle_overlap_idx(void* h, const int* ele_idx, double* ip_out, int* cnt_out) {
  Run& R = *(Run*)h;
  refresh_tables(R);
  const Model& m = R.m;
  const int ns = m.nsite(), ne = m.ne(), n = m.nsize();
  std::vector<int> cfg(2 * ns, -1), num(2 * ns, 0), cnt(m.n_proj() > 0 ? m.n_proj() : 1, 0);
  for (int mi = 0; mi < n; ++mi) cfg[ele_idx[mi] + (mi / ne) * ns] = mi % ne;
  for (int rsi = 0; rsi < 2 * ns; ++rsi) num[rsi] = cfg[rsi] < 0 ? 0 : 1;
  make_proj_cnt(m, num.data(), cnt.data());
  std::vector<double> inv((size_t)m.nqp() * n * n), pf(m.nqp());
  const int st = calc_m_all(R, ele_idx, inv.data(), pf.data(), false);
  *ip_out = st ? 0.0 : calc_ip(m, pf.data());
  for (int i = 0; i < m.n_proj(); ++i) cnt_out[i] = cnt[i];
  return st;
}

// FSZ candidates on a raw-word tape: kind 0 hopping_fsz (2Sz = -1), 1
// hopping_csz, 2 LocalSpinFlip_conduction; out8 as oracle_hop_candidate_tape.
int oracle_fsz_candidate_tape(int kind, int ns, int nsize, const int* loc_spn, const int* ele_idx,
                              const int* ele_cfg, const int* ele_spn, const uint32_t* words, int nwords,
                              int* out8) {
  WordTape tape{words, nwords};
  try {
    int mi = 0, s = 0, t = 0, ri = 0, rj = 0;
    bool rej;
    if (kind == 2) {
      rej = spin_flip_conduction(ns, nsize, loc_spn, ele_idx, ele_cfg, ele_spn, tape, mi, s, t, ri);
      rj = ri;
    } else {
      rej = hop_candidate_fsz(ns, nsize, kind == 0 ? -1 : 0, loc_spn, ele_idx, ele_cfg, ele_spn, tape, mi, s, t,
                              ri, rj);
    }
    const int used = tape.pos;
    const int vals[8] = {mi, ri, rj, s, t, rej ? 1 : 0, used, used + (rej ? 0 : 1)};
    std::memcpy(out8, vals, sizeof(vals));
    return 0;
  } catch (const std::exception&) {
    return -1;
  }
}

// Metropolis decision with the uniform draw given as a raw word.
int oracle_metropolis_word(double lpd, double lip_new, double lip_old, uint32_t word, double* w_out) {
  WordTape tape{&word, 1};
  return metropolis_rng(tape, lpd, lip_new, lip_old, w_out) ? 1 : 0;
}

// LTL only (for Pfaffian bit checks): returns info, writes Pf.
int oracle_pfaffian(const double* plane, int n, double* pf) {
  std::vector<double> a(plane, plane + (size_t)n * n);
  std::vector<int> piv(n);
  int info = dsktf2(a.data(), n, piv.data());
  *pf = info ? 0.0 : utu2pfa(a.data(), n, piv.data());
  return info;
}

double oracle_two_hop_bilinear(const double* inv, int n, const double* a, const double* b) {
  return two_hop_bilinear(inv, n, a, b);
}

// CG on a fixed operand set: O is [samples][n]; returns the C status.
long long oracle_cg(int n, int samples, const double* mean, const double* diag, const double* O, const double* g,
                    double inv_w, double shift, double tol, int max_it, int diag_scale, double* x, double* r,
                    double* d) {
  CgOperator op;
  op.n = n;
  op.mean.assign(mean, mean + n);
  op.diag.assign(diag, diag + n);
  op.parts.push_back(O);
  op.part_samples.push_back(samples);
  return cg_solve(op, g, inv_w, shift, tol, max_it, diag_scale != 0, x, r, d);
}

void oracle_cg_apply(int n, int samples, const double* mean, const double* diag, const double* O, const double* x,
                     double* z, double inv_w, double shift) {
  CgOperator op;
  op.n = n;
  op.mean.assign(mean, mean + n);
  op.diag.assign(diag, diag + n);
  op.parts.push_back(O);
  op.part_samples.push_back(samples);
  cg_apply(op, x, z, inv_w, shift);
}

// Optimisation handle.  initial_def may be NULL (then C's init_parameter
// draws from walker 0's stream; every walker consumes the same draws).
void* oracle_run_create(const char* namelist, const char* initial_def, int walkers, long long seed_override,
                        int use_seed_override, char* err, int errlen) {
  try {
    Run* R = new Run();
    R->m = read_expert(namelist);
    Model& m = R->m;
    const long long seed = use_seed_override ? seed_override : m.mp.rnd_seed;
    R->w.resize(walkers);
    // init_parameter (parameter_init.rs:63-160): projection 0, Slater 2(r-1/2)
    // for each optimised slot, one real2 draw per active slot, on every rank.
    for (int wi = 0; wi < walkers; ++wi) {
      Walker& W = R->w[wi];
      W.rng.init((uint32_t)(seed + wi));
      for (int i = 0; i < m.n_slater; ++i) {
        const bool active = m.opt_flags[2 * (m.n_proj() + i)] > 0;
        double v = 0.0;
        if (active) v = 2.0 * (W.rng.real2() - 0.5);
        if (wi == 0) m.slater[i] = v;
      }
    }
    for (double& p : m.proj) p = 0.0;
    if (initial_def && *initial_def) read_initial(m, initial_def);
    sync_parameters(m, true);
    {
      const int ns = m.nsite(), nt = m.n_mp_trans();
      R->tsite.assign((size_t)nt * ns, -1);
      R->tsgn.assign((size_t)nt * ns, 1);
      for (int t = 0; t < nt; ++t)
        for (int r = 0; r < ns; ++r) translate(m, r, t, R->tsite[(size_t)t * ns + r], R->tsgn[(size_t)t * ns + r]);
    }
    const int n = m.nsize(), n2 = m.nsite2();
    for (auto& W : R->w) {
      W.ele_idx.assign(n, -1);
      W.ele_cfg.assign(n2, -1);
      W.ele_num.assign(n2, 0);
      W.proj_cnt.assign(m.n_proj(), 0);
      W.inv.assign((size_t)m.nqp() * n * n, 0.0);
      W.pf.assign(m.nqp(), 0.0);
    }
    return R;
  } catch (std::exception& e) {
    if (err && errlen > 0) std::snprintf(err, errlen, "%s", e.what());
    return nullptr;
  }
}

void oracle_run_destroy(void* h) { delete (Run*)h; }
void oracle_run_set_threads(void* h, int t) { ((Run*)h)->threads = t; }

int oracle_run_step(void* h, int do_sr, char* err, int errlen) {
  try {
    return opt_step(*(Run*)h, do_sr != 0);
  } catch (std::exception& e) {
    if (err && errlen > 0) std::snprintf(err, errlen, "%s", e.what());
    return -1000;
  }
}

// Sampling only (one SR step's worth) without measurement.
int oracle_run_sample(void* h) {
  Run& R = *(Run*)h;
  refresh_tables(R);
  for_walkers(R, [&](Walker& W) { sample_walker(R, W); });
  return 0;
}

int oracle_dims(void* h, int* out) {
  Run& R = *(Run*)h;
  const Model& m = R.m;
  out[0] = m.nsite(); out[1] = m.ne(); out[2] = m.nsize(); out[3] = m.nqp(); out[4] = m.n_proj();
  out[5] = m.n_slater; out[6] = m.sr_opt_size(); out[7] = m.mp.nvmc_sample; out[8] = (int)R.w.size();
  out[9] = (int)R.oo.size();
  return 0;
}

void oracle_get_params(void* h, double* proj, double* slater) {
  Run& R = *(Run*)h;
  std::copy(R.m.proj.begin(), R.m.proj.end(), proj);
  std::copy(R.m.slater.begin(), R.m.slater.end(), slater);
}
void oracle_set_params(void* h, const double* proj, const double* slater) {
  Run& R = *(Run*)h;
  std::copy(proj, proj + R.m.n_proj(), R.m.proj.begin());
  std::copy(slater, slater + R.m.n_slater, R.m.slater.begin());
}

void oracle_get_energy(void* h, double* out5) {
  Run& R = *(Run*)h;
  out5[0] = R.wc; out5[1] = R.etot; out5[2] = R.etot2; out5[3] = R.sztot; out5[4] = R.sztot2;
}
int oracle_get_energies(void* h, double* out) {
  Run& R = *(Run*)h;
  std::copy(R.energies.begin(), R.energies.end(), out);
  return (int)R.energies.size() / 6;
}
void oracle_get_operands(void* h, double* oo, double* ho) {
  Run& R = *(Run*)h;
  std::copy(R.oo.begin(), R.oo.end(), oo);
  std::copy(R.ho.begin(), R.ho.end(), ho);
}
void oracle_get_o_last(void* h, int walker, double* o) {
  Run& R = *(Run*)h;
  std::copy(R.w[walker].o_last.begin(), R.w[walker].o_last.end(), o);
}
void oracle_get_o_store(void* h, int walker, double* o) {
  Run& R = *(Run*)h;
  std::copy(R.w[walker].o_store.begin(), R.w[walker].o_store.end(), o);
}
void oracle_get_walker(void* h, int walker, int* ele_idx_samples, int* counter, uint32_t* rng_words, int* rng_idx,
                       long long* accepted) {
  Run& R = *(Run*)h;
  Walker& W = R.w[walker];
  if (ele_idx_samples) std::copy(W.s_idx.begin(), W.s_idx.end(), ele_idx_samples);
  if (counter)
    for (int i = 0; i < 10; ++i) counter[i] = (int)W.counter[i];
  if (rng_words) std::memcpy(rng_words, W.rng.w, sizeof(W.rng.w));
  if (rng_idx) *rng_idx = W.rng.idx;
  if (accepted) *accepted = W.accepted;
}
void oracle_get_walker_tables(void* h, int walker, double* inv, double* pf) {
  Run& R = *(Run*)h;
  Walker& W = R.w[walker];
  std::copy(W.inv.begin(), W.inv.end(), inv);
  std::copy(W.pf.begin(), W.pf.end(), pf);
}
void oracle_get_burn(void* h, int walker, int* ele_idx, int* ele_cfg, int* ele_num, int* proj_cnt) {
  Run& R = *(Run*)h;
  Walker& W = R.w[walker];
  std::copy(W.burn_idx.begin(), W.burn_idx.end(), ele_idx);
  std::copy(W.burn_cfg.begin(), W.burn_cfg.end(), ele_cfg);
  std::copy(W.burn_num.begin(), W.burn_num.end(), ele_num);
  std::copy(W.burn_proj.begin(), W.burn_proj.end(), proj_cnt);
}
void oracle_get_slater_table(void* h, double* out) {
  Run& R = *(Run*)h;
  refresh_tables(R);
  std::copy(R.slater.begin(), R.slater.end(), out);
}
void oracle_get_qp_weights(void* h, double* out) {
  Run& R = *(Run*)h;
  std::copy(R.m.qp_full_weight.begin(), R.m.qp_full_weight.end(), out);
}
void oracle_get_times(void* h, double* out3) {
  Run& R = *(Run*)h;
  out3[0] = R.t_sample; out3[1] = R.t_measure; out3[2] = R.t_sr;
}
int oracle_get_cg(void* h, double* x) {
  Run& R = *(Run*)h;
  std::copy(R.cg_solution.begin(), R.cg_solution.end(), x);
  return R.cg_status;
}

// Measurement of one configuration with given tables (for GPU parity of (c)/(3)).
double oracle_measure_config(void* h, const int* ele_idx, const int* ele_cfg, const int* ele_num, const int* proj_cnt,
                             double* o_out, double* ip_out, int* status) {
  Run& R = *(Run*)h;
  refresh_tables(R);
  const Model& m = R.m;
  std::vector<double> inv((size_t)m.nqp() * m.nsize() * m.nsize()), pf(m.nqp());
  *status = calc_m_all(R, ele_idx, inv.data(), pf.data(), false);
  if (*status) return 0.0;
  const double ip = calc_ip(m, pf.data());
  *ip_out = ip;
  const double e = local_energy(R, ip, ele_idx, ele_cfg, ele_num, proj_cnt, inv.data(), pf.data());
  if (o_out) {
    o_out[0] = 1.0;
    for (int i = 0; i < m.n_proj(); ++i) o_out[1 + i] = proj_cnt[i];
    if (m.n_slater > 0) slater_diff(R, ip, ele_idx, inv.data(), pf.data(), o_out + 1 + m.n_proj());
  }
  return e;
}

// Measurement of one configuration given only its EleIdx: EleCfg, EleNum and
// the projection counts derived as make_initial_sample leaves them
// (projection.rs:179-277).  Returns the local energy (0 when invalid).
double oracle_measure_idx(void* h, const int* ele_idx, double* o_out, int* valid) {
  Run& R = *(Run*)h;
  const Model& m = R.m;
  const int ns = m.nsite(), ne = m.ne(), n = m.nsize();
  std::vector<int> cfg(2 * ns, -1), num(2 * ns, 0), cnt(m.n_proj() > 0 ? m.n_proj() : 1, 0);
  for (int mi = 0; mi < n; ++mi) {
    const int si = mi / ne;
    cfg[ele_idx[mi] + si * ns] = mi % ne;
  }
  for (int rsi = 0; rsi < 2 * ns; ++rsi) num[rsi] = cfg[rsi] < 0 ? 0 : 1;
  make_proj_cnt(m, num.data(), cnt.data());
  double ip = 0.0;
  int status = 0;
  std::vector<double> o(m.sr_opt_size(), 0.0);
  const double e = oracle_measure_config(h, ele_idx, cfg.data(), num.data(), cnt.data(), o.data(), &ip, &status);
  const bool ok = status == 0 && std::fabs(ip) >= 1.0e-100;
  if (o_out) std::copy(o.begin(), o.end(), o_out);
  *valid = ok ? 1 : 0;
  return ok ? e : 0.0;
}

// Gauge sync of bare parameter arrays (expert.cpp sync_parameters) and its
// Gutzwiller / Jastrow gate, for the pins on the C SetFlagShift records.
int oracle_gj_shift_gate(int n_gutz, int n_jast, const int* flags, int nflags) {
  return gj_shift_gate(n_gutz, n_jast, std::vector<int>(flags, flags + nflags)) ? 1 : 0;
}

void oracle_sync_parameters(int n_gutz, int n_jast, const int* flags, int nflags, double* proj, int n_slater,
                            double* slater) {
  Model m;
  m.n_gutz = n_gutz;
  m.n_jast = n_jast;
  m.n_slater = n_slater;
  m.opt_flags.assign(flags, flags + nflags);
  m.proj.assign(proj, proj + n_gutz + n_jast);
  m.slater.assign(slater, slater + n_slater);
  sync_parameters(m, true);
  std::copy(m.proj.begin(), m.proj.end(), proj);
  std::copy(m.slater.begin(), m.slater.end(), slater);
}

// Gauss-Legendre nodes and weights of the oracle's QP weights on [x1, x2].
void oracle_gauss_leg(int n, double x1, double x2, double* x, double* w) {
  std::vector<double> xv, wv;
  gauss_leg(x1, x2, n, xv, wv);
  std::copy(xv.begin(), xv.end(), x);
  std::copy(wv.begin(), wv.end(), w);
}

// Spin-projection Gauss-Legendre points of the oracle's QP weights:
// out[0..nl) cos^2(beta/2), [nl..2nl) sin^2, [2nl..3nl) cos sin, then the
// QP full weights of NMPTrans = 1 with translation weight 1.
void oracle_spin_gl(int nl, int nstot, double* out) {
  Model m;
  m.mp.nsp_gauss_leg = nl;
  m.mp.nsp_stot = nstot;
  m.mp.nmp_trans = 1;
  m.tr_weight.assign(1, 1.0);
  qp_weights(m);
  for (int i = 0; i < nl; ++i) {
    out[i] = m.spgl_cos_cos[i];
    out[nl + i] = m.spgl_sin_sin[i];
    out[2 * nl + i] = m.spgl_cos_sin[i];
    out[3 * nl + i] = m.qp_full_weight[i];
  }
}

}  // extern "C"
