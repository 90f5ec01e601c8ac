// ORACLE — test infrastructure only.  The oracle's own Expert-mode reader and
// parameter-dependent tables, restated from the reference's parsers and
// table builders so that no product code runs inside the checker:
//   namelist / .def readers  crates/mvmc-expert-parsers/src/lib.rs:230-320,
//                            src/parsers/{modpara,locspin,trans,coulomb,hund,
//                            exchange,gutzwiller,jastrow,orbital,qptrans}.rs
//   initial.def              crates/mvmc-core/src/initial_params.rs:9-133
//   QP weights               crates/mvmc-expert-parsers/src/utils/qp_weight.rs:24-146
//   Slater table             crates/mvmc-core/src/slater_update.rs:17-126 (real branch)
//   gauge sync               crates/mvmc-expert-parsers/src/utils/parameter_init.rs:167-260
//   translated sites         crates/mvmc-core/src/qp.rs:33-70 (no OptTrans)
// Scope: real (ComplexType 0) anti-parallel orbitals, no RBM / doublon-holon /
// OptTrans sections -- the inputs the sampling path is tested on.
#pragma once
#include <string>
#include <vector>

namespace oracle {

struct Hop {  // t c^dagger_{site1,spin1} c_{site2,spin2} (trans.def row)
  int site1, spin1, site2, spin2;
  double re;
};

struct Pair {  // CoulombInter / Hund / Exchange row
  int site1, site2;
  double value;
};

struct Params {  // ModPara keywords used by the path
  int nsite = 16, nelec = 8, nlocspin = 0, ncond = -1;
  int vmc_calc_mode = 0, nsp_gauss_leg = 1, nsp_stot = 0, nmp_trans = 0;
  int nsr_opt_itr_step = 1000, nsr_opt_itr_smp = 100;
  double dsr_opt_red_cut = 0.001, dsr_opt_sta_del = 0.02, dsr_opt_step_dt = 0.02, dsr_opt_cg_tol = 1e-10;
  int nsr_opt_cg_max_iter = 0, nvmc_warmup = 10, nvmc_interval = 1, nvmc_sample = 10;
  int nex_update_path = 0;
  long long rnd_seed = 11272;
  int nstore = 1, nsrcg = 0, use_diag_scale = 0, nsrcg_fallback = 0, nsrcg_abort_on_fail = 1;
  int rescale_smat = 0;
  int sr_flag = 0;  // diagonalisation mode (negative DSROptStepDt)
};

struct Model {
  Params mp;
  std::vector<Hop> transfer;
  std::vector<std::pair<int, double>> coulomb_intra;
  std::vector<Pair> coulomb_inter, hund, exchange;
  std::vector<int> loc_spn;            // [Nsite] 1 = local spin
  int n_gutz = 0, n_jast = 0, n_slater = 0;
  std::vector<int> gutz_idx;           // [Nsite]
  std::vector<int> jast_idx;           // [Nsite*Nsite], -1 on the diagonal
  std::vector<int> orb_idx, orb_sgn;   // [Nsite*Nsite]
  std::vector<std::vector<int>> tr_site, tr_sign;  // [NQPTrans][Nsite]
  std::vector<double> tr_weight;       // [NQPTrans]
  std::vector<double> qp_full_weight;  // [NQP]
  std::vector<double> spgl_cos_sin, spgl_cos_cos, spgl_sin_sin;
  std::vector<double> proj, slater;    // variational parameters
  std::vector<int> opt_flags;          // [2 * NPara] (re, im) component flags

  int nsite() const { return mp.nsite; }
  int ne() const { return mp.nelec; }
  int nsize() const { return 2 * mp.nelec; }
  int nsite2() const { return 2 * mp.nsite; }
  int n_proj() const { return n_gutz + n_jast; }
  int n_para() const { return n_proj() + n_slater; }
  int sr_opt_size() const { return 1 + n_para(); }
  int n_mp_trans() const { return mp.nmp_trans < 0 ? -mp.nmp_trans : mp.nmp_trans; }
  int n_sp_gl() const { return mp.nsp_gauss_leg > 0 ? mp.nsp_gauss_leg : 1; }
  int nqp() const { return n_sp_gl() * n_mp_trans(); }
  int jast(int a, int b) const { return a == b ? -1 : jast_idx[(size_t)a * mp.nsite + b]; }
};

// Reads namelist.def and the sections it names; throws std::runtime_error.
Model read_expert(const std::string& namelist);
// Last complete (6 + 3 NPara)-number record of initial.def / zqp_opt.dat.
bool read_initial(Model& m, const std::string& path);
// Gauss-Legendre nodes / weights on [x1, x2] (GaussLeg of qp.c).
void gauss_leg(double x1, double x2, int n, std::vector<double>& x, std::vector<double>& w);
// GL spin-projection weights and QPFullWeight.
void qp_weights(Model& m);
// S[qp][2Ns][2Ns] (row-major) from the orbital parameters.
void slater_table(const Model& m, std::vector<double>& S);
// Gutzwiller / Jastrow mean shift (all G and J real flags == 1) and the
// rescale of the Slater parameters to max |f| = 4.
void sync_parameters(Model& m, bool shift_correlations);
// Gate of the Gutzwiller / Jastrow shift from the component flags.
bool gj_shift_gate(int n_gutz, int n_jast, const std::vector<int>& flags);
// (site, sign) of `site` under translation t.
void translate(const Model& m, int site, int t, int& out, int& sign);

}  // namespace oracle
