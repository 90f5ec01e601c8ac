// Model description for the sampling + SR-optimisation path: an Expert-mode
// input set (namelist.def and the .def files it names) reduced to the tables
// the sampler, the measurement and the SR-CG solver read.
//
// Real-arithmetic, spin-conserving ("normal") mode only: every variational
// parameter is real (ComplexType 0), orbitals are anti-parallel, no RBM /
// doublon-holon / OptTrans terms.  This is the mode the reference selects for
// its real sampler (crates/mvmc-core/src/run.rs:1336-1340, all_complex=false).
//
// Layout conventions follow the reference so that a drop-in caller can hand
// its own buffers over unchanged:
//   ele_idx[mi + s*Ne]      site of electron mi of spin s      (state.rs:607)
//   ele_cfg[ri + s*Nsite]   electron index at (ri,s) or -1
//   ele_num[ri + s*Nsite]   occupation 0/1
//   slater[qp][rsi][rsj]    row-major, rs = r + s*Nsite         (state.rs:42)
//   invM plane qp: inv[msi*Nsize + msj], planes of Nsize^2 doubles
//   projection parameters: Gutzwiller block, then Jastrow block
//   variational parameters: projection | Slater                 (sync.rs:26)
#pragma once
#include <cstdint>
#include <string>
#include <vector>

namespace mvmc {

struct TransferTerm {
  int site1, spin1;   // creation  c^dagger_{site1,spin1}
  int site2, spin2;   // annihilation c_{site2,spin2}
  double re, im;
};

struct PairTerm {
  int site1, site2;
  double value;
};

struct ModPara {
  int nsite = 16;
  int nelec = 8;        // electrons per spin (Ne)
  int nlocspin = 0;
  int ncond = -1;
  int two_sz = -1;
  int vmc_calc_mode = 0;
  int nsp_gauss_leg = 1;
  int nsp_stot = 0;
  int nmp_trans = 0;
  int nsr_opt_itr_step = 1000;
  int nsr_opt_itr_smp = 100;
  std::string c_para_file_head = "zqp";  // CParaFileHead of the banner (parameter output files)
  std::string c_data_file_head = "zvo";  // CDataFileHead of the banner (per-step output files)
  double dsr_opt_red_cut = 0.001;
  double dsr_opt_sta_del = 0.02;
  double dsr_opt_step_dt = 0.02;
  double dsr_opt_cg_tol = 1e-10;
  int nsr_opt_cg_max_iter = 0;
  int nvmc_warmup = 10;
  int nvmc_interval = 1;
  int nvmc_sample = 10;
  int nex_update_path = 0;
  long long rnd_seed = 11272;
  int nsplit_size = 1;
  int nstore = 1;
  int nsrcg = 0;
  int use_diag_scale = 0;
  int nsrcg_fallback = 0;
  int nsrcg_abort_on_fail = 1;
  int rescale_smat = 0;
  int sr_flag = 0;  // diagonalisation mode (negative DSROptStepDt)
};

struct Model {
  ModPara mp;
  // Hamiltonian
  std::vector<TransferTerm> transfer;
  std::vector<std::pair<int, double>> coulomb_intra;
  std::vector<PairTerm> coulomb_inter, hund, exchange;
  std::vector<int> loc_spn;          // [nsite] 1 = local spin site
  // projection (Gutzwiller / Jastrow)
  int n_gutz = 0, n_jast = 0;
  std::vector<int> gutz_idx;         // [nsite] parameter index (or -1)
  std::vector<int> jast_idx;         // [nsite*nsite], -1 on the diagonal
  // orbitals (anti-parallel, real)
  int n_slater = 0;
  std::vector<int> orb_idx, orb_sgn; // [nsite*nsite]
  // quantum-number projection
  int n_qp_trans = 0;                // entries of qptransidx.def
  std::vector<std::vector<int>> qp_trans;      // [n_qp_trans][nsite]
  std::vector<std::vector<int>> qp_trans_sgn;  // [n_qp_trans][nsite]
  std::vector<double> para_qp_trans;           // weights (real part)
  std::vector<double> para_qp_trans_im;
  // derived quantum-projection weights
  std::vector<double> qp_full_weight;          // [nqp]
  std::vector<double> spgl_cos_sin, spgl_cos_cos, spgl_sin_sin; // [nsp_gauss_leg]
  // variational parameters (real mode)
  std::vector<double> proj;          // [n_proj]
  std::vector<double> slater;        // [n_slater]
  std::vector<int> opt_flags;        // [2*n_para]

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
  int jast(int a, int b) const {  // upper-triangle lookup, smaller site first
    if (a == b) return -1;
    return a < b ? jast_idx[a * mp.nsite + b] : jast_idx[b * mp.nsite + a];
  }
};

// Parse namelist.def and the files it names.  Throws std::runtime_error with
// the file and reason on malformed input (the reference treats bad input as a
// hard error, crates/mvmc-expert-parsers/src/lib.rs:230-247).
Model load_namelist(const std::string& namelist_path);

// initial.def / zqp_opt.dat layout: six leading numbers, then (re, im, x)
// per variational parameter; the last complete record wins
// (crates/mvmc-core/src/initial_params.rs:9-106).
bool load_initial_def(Model& m, const std::string& path);

// Gauss-Legendre / Legendre polynomial quantum-projection weights
// (crates/mvmc-expert-parsers/src/utils/qp_weight.rs:24-146).
void init_qp_weight(Model& m);

// Slater table S[qp][2Ns][2Ns] from the current orbital parameters
// (crates/mvmc-core/src/slater_update.rs:17-126, real branch).
void build_slater_table(const Model& m, std::vector<double>& table);

// Gauge fixing after every parameter change: Gutzwiller/Jastrow mean shift
// when every projection parameter is optimised, then Slater rescaled to
// max|f| = 4 (crates/mvmc-expert-parsers/src/utils/parameter_init.rs:168-265,
// crates/mvmc-core/src/sync.rs:96-115).
void sync_modified_parameter(Model& m, bool shift_correlations);

// Translated site / sign of `site` under translation `mpidx`
// (crates/mvmc-core/src/qp.rs:30-70 with no OptTrans).
void translated_site(const Model& m, int site, int mpidx, int& out_site, int& out_sign);

// Flattened tables of the direct transfer projection ratio
// (crates/mvmc-core/src/observables.rs:1553-1605).
void direct_projection_tables(const Model& m, std::vector<double>& gutz,
                              std::vector<double>& jast);

// Per-translation orbital index/sign of electron pairs are data-dependent;
// this returns the static translated site map [n_mp_trans][nsite] and sign.
void translation_maps(const Model& m, std::vector<int>& tsite, std::vector<int>& tsgn);

}  // namespace mvmc
