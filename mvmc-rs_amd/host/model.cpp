// Expert-mode input loader and parameter-dependent tables (host side).
// Formats follow the reference readers:
//   modpara.def    crates/mvmc-expert-parsers/src/parsers/modpara.rs:100-215
//   trans.def      crates/mvmc-expert-parsers/src/parsers/trans.rs
//   coulomb*.def   crates/mvmc-expert-parsers/src/parsers/coulomb.rs
//   hund/exchange  crates/mvmc-expert-parsers/src/parsers/{hund,exchange}.rs
//   locspn.def     crates/mvmc-expert-parsers/src/parsers/locspin.rs
//   gutzwiller/jastrow/orbital/qptransidx  parsers/{gutzwiller,jastrow,orbital,qptrans}.rs
#include "model.h"

#include <algorithm>
#include <cctype>
#include <cmath>
#include <cstdlib>
#include <fstream>
#include <map>
#include <sstream>
#include <stdexcept>

namespace mvmc {
namespace {

std::string dirname_of(const std::string& p) {
  auto pos = p.find_last_of('/');
  return pos == std::string::npos ? std::string(".") : p.substr(0, pos);
}

std::vector<std::string> read_lines(const std::string& path) {
  std::ifstream f(path);
  if (!f) throw std::runtime_error("cannot open " + path);
  std::vector<std::string> lines;
  std::string s;
  while (std::getline(f, s)) {
    if (!s.empty() && s.back() == '\r') s.pop_back();
    lines.push_back(s);
  }
  return lines;
}

std::vector<std::string> tokens(const std::string& s) {
  std::vector<std::string> out;
  std::istringstream is(s);
  std::string t;
  while (is >> t) out.push_back(t);
  return out;
}

std::string lower(std::string s) {
  for (auto& c : s) c = (char)std::tolower((unsigned char)c);
  return s;
}

// Second whitespace field of a header line as an integer ("NTransfer 24").
int header_count(const std::vector<std::string>& lines, size_t line, const std::string& path) {
  if (lines.size() <= line) throw std::runtime_error(path + ": missing header line");
  auto t = tokens(lines[line]);
  if (t.size() < 2) throw std::runtime_error(path + ": malformed header '" + lines[line] + "'");
  return std::atoi(t[1].c_str());
}

// All whitespace tokens after the five header lines.
std::vector<std::string> body_tokens(const std::vector<std::string>& lines) {
  std::vector<std::string> out;
  for (size_t i = 5; i < lines.size(); ++i) {
    auto t = tokens(lines[i]);
    out.insert(out.end(), t.begin(), t.end());
  }
  return out;
}

std::vector<std::vector<std::string>> body_rows(const std::vector<std::string>& lines) {
  std::vector<std::vector<std::string>> rows;
  for (size_t i = 5; i < lines.size(); ++i) {
    auto t = tokens(lines[i]);
    if (!t.empty()) rows.push_back(t);
  }
  return rows;
}

int to_i(const std::string& s, const std::string& path) {
  char* end = nullptr;
  long v = std::strtol(s.c_str(), &end, 10);
  if (end == s.c_str()) throw std::runtime_error(path + ": bad integer '" + s + "'");
  return (int)v;
}

double to_d(const std::string& s, const std::string& path) {
  char* end = nullptr;
  double v = std::strtod(s.c_str(), &end);
  if (end == s.c_str()) throw std::runtime_error(path + ": bad number '" + s + "'");
  return v;
}

void parse_modpara(const std::string& path, ModPara& p) {
  auto lines = read_lines(path);
  // The first eight lines are a fixed banner (file heads); keywords follow.
  for (size_t i = 0; i < 8 && i < lines.size(); ++i) {
    auto t = tokens(lines[i]);
    if (t.size() >= 2 && lower(t[0]) == "cparafilehead") p.c_para_file_head = t[1];
    if (t.size() >= 2 && lower(t[0]) == "cdatafilehead") p.c_data_file_head = t[1];
  }
  for (size_t i = 8; i < lines.size(); ++i) {
    const std::string& l = lines[i];
    if (l.empty() || l[0] == '-') continue;
    auto t = tokens(l);
    if (t.empty()) continue;
    if (t.size() < 2) throw std::runtime_error(path + ": keyword without value: " + l);
    const std::string k = lower(t[0]);
    const double v = to_d(t[1], path);
    const int iv = (int)v;
    if (k == "nvmccalmode") p.vmc_calc_mode = iv;
    else if (k == "nlanczosmode" || k == "ndataidxstart" || k == "ndataqtysmp") {}
    else if (k == "nsite") p.nsite = iv;
    else if (k == "ne" || k == "nelectron") p.nelec = iv;
    else if (k == "ncond") p.ncond = iv;
    else if (k == "2sz") p.two_sz = iv;
    else if (k == "nspgaussleg") p.nsp_gauss_leg = iv;
    else if (k == "nspstot") p.nsp_stot = iv;
    else if (k == "nmptrans") p.nmp_trans = iv;
    else if (k == "nsroptitrstep") p.nsr_opt_itr_step = iv;
    else if (k == "nsroptitrsmp") p.nsr_opt_itr_smp = iv;
    else if (k == "dsroptredcut") p.dsr_opt_red_cut = v;
    else if (k == "dsroptstadel") p.dsr_opt_sta_del = v;
    else if (k == "dsroptstepdt") p.dsr_opt_step_dt = v;
    else if (k == "nsroptcgmaxiter") p.nsr_opt_cg_max_iter = iv;
    else if (k == "dsroptcgtol") p.dsr_opt_cg_tol = v;
    else if (k == "nvmcwarmup") p.nvmc_warmup = iv;
    else if (k == "nvmcinterval") p.nvmc_interval = iv;
    else if (k == "nvmcsample") p.nvmc_sample = iv;
    else if (k == "nexupdatepath") p.nex_update_path = iv;
    else if (k == "rndseed") p.rnd_seed = (long long)v;
    else if (k == "nsplitsize") p.nsplit_size = iv;
    else if (k == "nstore") p.nstore = iv;
    else if (k == "nsrcg") p.nsrcg = iv;
    else if (k == "nsrcgfallback") p.nsrcg_fallback = iv;
    else if (k == "nsrcgabortonfail") p.nsrcg_abort_on_fail = iv;
    else if (k == "rescalesmat") p.rescale_smat = iv;
    else if (k == "usediagscale") p.use_diag_scale = iv;
    else throw std::runtime_error(path + ": unsupported keyword '" + t[0] + "'");
  }
  // readdef.c:927-957: NSRCG 2 selects the diagonally scaled CG.
  if (p.nsrcg == 2) { p.nsrcg = 1; p.use_diag_scale = 1; }
  if (p.nsrcg != 1) p.use_diag_scale = 0;
  // readdef.c:959-975 (crates/mvmc-expert-parsers/src/parsers/modpara.rs:250-266):
  // a negative DSROptStepDt selects the diagonalisation mode (SRFlag) and is
  // negated before use; SR-CG has no such mode.  The direct solve is the same
  // Cholesky solve in both modes (only the SRinfo header differs in C).
  p.sr_flag = p.dsr_opt_step_dt < 0.0 ? 1 : 0;
  if (p.sr_flag) p.dsr_opt_step_dt *= -1.0;
  if (p.sr_flag && p.nsrcg != 0)
    throw std::runtime_error(path + ": SR-CG cannot be combined with a negative DSROptStepDt");
}

int complex_type(const std::vector<std::string>& lines, const std::string& path) {
  int ct = header_count(lines, 2, path);
  if (ct != 0) throw std::runtime_error(path + ": complex parameters are not supported (ComplexType " +
                                        std::to_string(ct) + ")");
  return ct;
}

void read_flags(const std::vector<std::string>& tok, size_t first, int count, int offset,
                std::vector<int>& flags, const std::string& path) {
  if (tok.size() != first + 2 * (size_t)count)
    throw std::runtime_error(path + ": flag section must hold exactly the declared pairs");
  for (int k = 0; k < count; ++k) {
    int idx = to_i(tok[first + 2 * k], path);
    int flag = to_i(tok[first + 2 * k + 1], path);
    if (idx < 0 || idx >= count) throw std::runtime_error(path + ": flag index out of range");
    flags[2 * (offset + idx)] = flag;
    flags[2 * (offset + idx) + 1] = 0;
  }
}

}  // namespace

Model load_namelist(const std::string& namelist_path) {
  Model m;
  const std::string dir = dirname_of(namelist_path);
  std::map<std::string, std::string> files;
  for (auto& l : read_lines(namelist_path)) {
    auto t = tokens(l);
    if (t.size() < 2) continue;
    files[lower(t[0])] = dir + "/" + t[1];
  }
  auto has = [&](const char* k) { return files.count(k) != 0; };
  if (!has("modpara")) throw std::runtime_error(namelist_path + ": no ModPara entry");
  parse_modpara(files["modpara"], m.mp);
  for (const char* k : {"pairhop", "interall", "dh2", "dh4", "orbitalparallel", "orbitalgeneral", "opttrans"})
    if (has(k)) throw std::runtime_error(std::string("unsupported Expert section: ") + k);

  const int ns = m.mp.nsite;
  if (ns <= 0) throw std::runtime_error("Nsite must be positive");

  // LocSpin
  m.loc_spn.assign(ns, 0);
  if (has("locspin")) {
    const auto& p = files["locspin"];
    auto lines = read_lines(p);
    m.mp.nlocspin = header_count(lines, 1, p);
    auto rows = body_rows(lines);
    for (auto& r : rows) {
      if (r.size() < 2) continue;
      int i = to_i(r[0], p), f = to_i(r[1], p);
      if (i < 0 || i >= ns) throw std::runtime_error(p + ": site out of range");
      m.loc_spn[i] = (f == 1) ? 1 : 0;
    }
  }
  if (m.mp.ncond != -1) m.mp.nelec = (m.mp.nlocspin + m.mp.ncond) / 2;

  if (has("trans")) {
    const auto& p = files["trans"];
    auto lines = read_lines(p);
    int n = header_count(lines, 1, p);
    auto rows = body_rows(lines);
    if ((int)rows.size() < n) throw std::runtime_error(p + ": incomplete transfer section");
    for (int k = 0; k < n; ++k) {
      auto& r = rows[k];
      if (r.size() < 5) throw std::runtime_error(p + ": short transfer row");
      TransferTerm t{to_i(r[0], p), to_i(r[1], p), to_i(r[2], p), to_i(r[3], p), to_d(r[4], p),
                     r.size() > 5 ? to_d(r[5], p) : 0.0};
      if (t.im != 0.0) throw std::runtime_error(p + ": complex transfer is not supported");
      m.transfer.push_back(t);
    }
  }
  if (has("coulombintra")) {
    const auto& p = files["coulombintra"];
    auto lines = read_lines(p);
    int n = header_count(lines, 1, p);
    auto rows = body_rows(lines);
    for (int k = 0; k < n; ++k) m.coulomb_intra.push_back({to_i(rows.at(k)[0], p), to_d(rows.at(k)[1], p)});
  }
  auto pair_section = [&](const char* key, std::vector<PairTerm>& out) {
    if (!has(key)) return;
    const auto& p = files[key];
    auto lines = read_lines(p);
    int n = header_count(lines, 1, p);
    auto rows = body_rows(lines);
    for (int k = 0; k < n; ++k)
      out.push_back({to_i(rows.at(k)[0], p), to_i(rows.at(k)[1], p), to_d(rows.at(k)[2], p)});
  };
  pair_section("coulombinter", m.coulomb_inter);
  pair_section("hund", m.hund);
  pair_section("exchange", m.exchange);

  // Projection and orbital sections: indices first, flags after every
  // count is known (flags are stored per parameter as (re, im) pairs).
  std::vector<std::string> gtok, jtok, otok;
  std::string gpath, jpath, opath;
  m.gutz_idx.assign(ns, -1);
  if (has("gutzwiller")) {
    gpath = files["gutzwiller"];
    auto lines = read_lines(gpath);
    m.n_gutz = header_count(lines, 1, gpath);
    complex_type(lines, gpath);
    gtok = body_tokens(lines);
    if (gtok.size() < (size_t)2 * ns) throw std::runtime_error(gpath + ": incomplete site section");
    for (int k = 0; k < ns; ++k) {
      int i = to_i(gtok[2 * k], gpath), idx = to_i(gtok[2 * k + 1], gpath);
      if (i < 0 || i >= ns || idx >= m.n_gutz) throw std::runtime_error(gpath + ": bad Gutzwiller row");
      m.gutz_idx[i] = idx;
    }
  }
  m.jast_idx.assign((size_t)ns * ns, -1);
  if (has("jastrow")) {
    jpath = files["jastrow"];
    auto lines = read_lines(jpath);
    m.n_jast = header_count(lines, 1, jpath);
    complex_type(lines, jpath);
    jtok = body_tokens(lines);
    size_t nmap = (size_t)ns * (ns - 1);
    if (jtok.size() < 3 * nmap) throw std::runtime_error(jpath + ": incomplete mapping section");
    for (size_t k = 0; k < nmap; ++k) {
      int i = to_i(jtok[3 * k], jpath), j = to_i(jtok[3 * k + 1], jpath), idx = to_i(jtok[3 * k + 2], jpath);
      if (i == j || i < 0 || j < 0 || i >= ns || j >= ns || idx < 0 || idx >= m.n_jast)
        throw std::runtime_error(jpath + ": bad Jastrow row");
      m.jast_idx[(size_t)i * ns + j] = idx;
    }
    for (int i = 0; i < ns; ++i)
      for (int j = i + 1; j < ns; ++j)
        if (m.jast_idx[(size_t)i * ns + j] != m.jast_idx[(size_t)j * ns + i])
          throw std::runtime_error(jpath + ": Jastrow pair has inconsistent classes");
  }
  m.orb_idx.assign((size_t)ns * ns, -1);
  m.orb_sgn.assign((size_t)ns * ns, 1);
  if (has("orbital") || has("orbitalantiparallel")) {
    opath = has("orbital") ? files["orbital"] : files["orbitalantiparallel"];
    auto lines = read_lines(opath);
    m.n_slater = header_count(lines, 1, opath);
    complex_type(lines, opath);
    auto rows = body_rows(lines);
    size_t nmap = (size_t)ns * ns;
    if (rows.size() < nmap) throw std::runtime_error(opath + ": incomplete mapping section");
    int sign = 1;  // a sign column, when present, persists to later rows
    for (size_t k = 0; k < nmap; ++k) {
      auto& r = rows[k];
      if (r.size() < 3) throw std::runtime_error(opath + ": short orbital row");
      int i = to_i(r[0], opath), j = to_i(r[1], opath), idx = to_i(r[2], opath);
      if (r.size() >= 4) sign = to_i(r[3], opath);
      if (i < 0 || j < 0 || i >= ns || j >= ns || idx < 0 || idx >= m.n_slater)
        throw std::runtime_error(opath + ": bad orbital row");
      m.orb_idx[(size_t)i * ns + j] = idx;
      m.orb_sgn[(size_t)i * ns + j] = sign;
    }
    for (size_t k = nmap; k < rows.size(); ++k)
      for (auto& t : rows[k]) otok.push_back(t);
  }
  m.opt_flags.assign(2 * (size_t)m.n_para(), 0);
  if (!gpath.empty()) {
    std::vector<std::string> f(gtok.begin() + 2 * ns, gtok.end());
    read_flags(f, 0, m.n_gutz, 0, m.opt_flags, gpath);
  }
  if (!jpath.empty()) {
    std::vector<std::string> f(jtok.begin() + 3 * (size_t)ns * (ns - 1), jtok.end());
    read_flags(f, 0, m.n_jast, m.n_gutz, m.opt_flags, jpath);
  }
  if (!opath.empty()) read_flags(otok, 0, m.n_slater, m.n_proj(), m.opt_flags, opath);

  // TransSym
  if (has("transsym") || has("qptrans")) {
    const auto& p = has("transsym") ? files["transsym"] : files["qptrans"];
    auto lines = read_lines(p);
    int n = header_count(lines, 1, p);
    m.n_qp_trans = n;
    auto rows = body_rows(lines);
    if ((int)rows.size() != n + n * ns) throw std::runtime_error(p + ": record counts differ from declaration");
    m.para_qp_trans.assign(n, 0.0);
    m.para_qp_trans_im.assign(n, 0.0);
    m.qp_trans.assign(n, std::vector<int>(ns, -1));
    m.qp_trans_sgn.assign(n, std::vector<int>(ns, 1));
    for (int k = 0; k < n; ++k) {
      auto& r = rows[k];
      int idx = to_i(r.at(0), p);
      if (idx < 0 || idx >= n) throw std::runtime_error(p + ": weight index out of range");
      m.para_qp_trans[idx] = r.size() > 1 ? to_d(r[1], p) : 0.0;
      m.para_qp_trans_im[idx] = r.size() > 2 ? to_d(r[2], p) : 0.0;
      if (m.para_qp_trans_im[idx] != 0.0) throw std::runtime_error(p + ": complex QP weights are not supported");
    }
    for (int k = 0; k < n * ns; ++k) {
      auto& r = rows[n + k];
      int t = to_i(r.at(0), p), i = to_i(r.at(1), p), j = to_i(r.at(2), p);
      int s = r.size() > 3 ? to_i(r[3], p) : 1;
      if (t < 0 || t >= n || i < 0 || i >= ns || j < 0 || j >= ns) throw std::runtime_error(p + ": bad map row");
      m.qp_trans[t][i] = j;
      m.qp_trans_sgn[t][i] = s;
    }
  }
  if (m.n_mp_trans() > m.n_qp_trans) throw std::runtime_error("NMPTrans exceeds the TransSym entries");

  m.proj.assign(m.n_proj(), 0.0);
  m.slater.assign(m.n_slater, 0.0);
  init_qp_weight(m);
  return m;
}

bool load_initial_def(Model& m, const std::string& path) {
  std::ifstream f(path);
  if (!f) return false;
  std::vector<double> v;
  std::string t;
  while (f >> t) v.push_back(to_d(t, path));
  if (v.empty()) return false;
  const size_t np = (size_t)m.n_para();
  const size_t expected = 6 + 3 * np;
  if (v.size() < expected || v.size() % expected != 0)
    throw std::runtime_error(path + ": not a whole number of parameter records");
  const double* r = v.data() + (v.size() - expected);
  for (size_t i = 0; i < np; ++i) {
    double re = r[6 + 3 * i], im = r[7 + 3 * i];
    if (im != 0.0) throw std::runtime_error(path + ": complex parameter values are not supported");
    double value = re + 0.0 * im;
    if (i < (size_t)m.n_proj()) m.proj[i] = value;
    else m.slater[i - m.n_proj()] = value;
  }
  return true;
}

// Gauss-Legendre nodes/weights on [x1,x2], Newton iteration with the
// reference's stopping rule |z - z1| <= 5e-14 (qp_weight.rs:24-60).
static void gauss_legendre(double x1, double x2, int n, std::vector<double>& x, std::vector<double>& w) {
  x.assign(n, 0.0);
  w.assign(n, 0.0);
  const int mh = (n + 1) / 2;
  const double xm = 0.5 * (x2 + x1), xl = 0.5 * (x2 - x1);
  const double nf = (double)n;
  for (int i = 0; i < mh; ++i) {
    double z = std::cos(M_PI * ((double)i + 0.75) / (nf + 0.5));
    double pp;
    for (;;) {
      double p1 = 1.0, p2 = 0.0;
      for (int j = 1; j <= n; ++j) {
        double p3 = p2;
        p2 = p1;
        double jf = (double)j;
        p1 = ((2.0 * jf - 1.0) * z * p2 - (jf - 1.0) * p3) / jf;
      }
      pp = nf * (z * p1 - p2) / (z * z - 1.0);
      double z1 = z;
      z = z1 - p1 / pp;
      if (std::fabs(z - z1) <= 5.0e-14) break;
    }
    x[i] = xm - xl * z;
    x[n - 1 - i] = xm + xl * z;
    w[i] = 2.0 * xl / ((1.0 - z * z) * pp * pp);
    w[n - 1 - i] = w[i];
  }
}

static double legendre_poly(double x, int n) {
  if (n <= 0) return 1.0;
  if (n == 1) return x;
  double p01 = 1.0, p02 = x, p03 = x;
  for (int i = 2; i <= n; ++i) {
    double fi = (double)i;
    p03 = (1.0 / fi) * ((2.0 * fi - 1.0) * x * p02 - (fi - 1.0) * p01);
    p01 = p02;
    p02 = p03;
  }
  return p03;
}

void init_qp_weight(Model& m) {
  const int nleg = m.n_sp_gl();
  const int nmp = m.n_mp_trans();
  m.qp_full_weight.assign((size_t)nleg * nmp, 0.0);
  m.spgl_cos_sin.assign(nleg, 0.0);
  m.spgl_cos_cos.assign(nleg, 0.0);
  m.spgl_sin_sin.assign(nleg, 0.0);
  if (nleg == 1) {
    m.spgl_cos_cos[0] = 1.0;
    for (int j = 0; j < nmp; ++j) m.qp_full_weight[j] = j < (int)m.para_qp_trans.size() ? m.para_qp_trans[j] : 0.0;
    return;
  }
  std::vector<double> beta, wgl;
  gauss_legendre(0.0, M_PI, nleg, beta, wgl);
  for (int i = 0; i < nleg; ++i) {
    double ch = std::cos(0.5 * beta[i]);
    double sh = std::sin(0.5 * beta[i]);
    m.spgl_cos_sin[i] = ch * sh;
    m.spgl_cos_cos[i] = ch * ch;
    m.spgl_sin_sin[i] = sh * sh;
    double cb = std::cos(beta[i]);
    double w = 0.5 * std::sin(beta[i]) * wgl[i] * legendre_poly(cb, m.mp.nsp_stot);
    for (int j = 0; j < nmp; ++j)
      if (j < (int)m.para_qp_trans.size()) m.qp_full_weight[i + j * nleg] = m.para_qp_trans[j] * w;
  }
}

void translated_site(const Model& m, int site, int mpidx, int& out_site, int& out_sign) {
  if (mpidx < (int)m.qp_trans.size()) {
    out_site = m.qp_trans[mpidx][site];
    out_sign = m.mp.nmp_trans < 0 ? m.qp_trans_sgn[mpidx][site] : 1;
  } else {
    out_site = site;
    out_sign = 1;
  }
}

void translation_maps(const Model& m, std::vector<int>& tsite, std::vector<int>& tsgn) {
  const int ns = m.nsite(), nmp = m.n_mp_trans();
  tsite.assign((size_t)nmp * ns, -1);
  tsgn.assign((size_t)nmp * ns, 1);
  for (int t = 0; t < nmp; ++t)
    for (int r = 0; r < ns; ++r) translated_site(m, r, t, tsite[(size_t)t * ns + r], tsgn[(size_t)t * ns + r]);
}

void build_slater_table(const Model& m, std::vector<double>& table) {
  const int ns = m.nsite(), n2 = m.nsite2(), nleg = m.n_sp_gl(), nqp = m.nqp();
  table.assign((size_t)nqp * n2 * n2, 0.0);
  for (int qp = 0; qp < nqp; ++qp) {
    double* plane = table.data() + (size_t)qp * n2 * n2;
    const int spidx = qp % nleg;
    const int mpidx = (qp % (nleg * m.n_mp_trans())) / nleg;
    const double cs = m.spgl_cos_sin[spidx], cc = m.spgl_cos_cos[spidx], ss = m.spgl_sin_sin[spidx];
    for (int ri = 0; ri < ns; ++ri) {
      int tri, sgni;
      translated_site(m, ri, mpidx, tri, sgni);
      if (tri < 0 || tri >= ns) continue;
      for (int rj = 0; rj < ns; ++rj) {
        int trj, sgnj;
        translated_site(m, rj, mpidx, trj, sgnj);
        if (trj < 0 || trj >= ns) continue;
        const int qpsgn = sgni * sgnj;
        const int idx_ij = m.orb_idx[(size_t)tri * ns + trj];
        const int idx_ji = m.orb_idx[(size_t)trj * ns + tri];
        if (idx_ij < 0 || idx_ji < 0) continue;
        const int sgn_ij = m.orb_sgn[(size_t)tri * ns + trj] * qpsgn;
        const int sgn_ji = m.orb_sgn[(size_t)trj * ns + tri] * qpsgn;
        const double slt_ij = m.slater[idx_ij] * (double)sgn_ij;
        const double slt_ji = m.slater[idx_ji] * (double)sgn_ji;
        const int rsi0 = ri, rsi1 = ri + ns, rsj0 = rj, rsj1 = rj + ns;
        plane[(size_t)rsi0 * n2 + rsj0] = -(slt_ij - slt_ji) * cs;
        plane[(size_t)rsi0 * n2 + rsj1] = slt_ij * cc + slt_ji * ss;
        plane[(size_t)rsi1 * n2 + rsj0] = -slt_ij * ss - slt_ji * cc;
        plane[(size_t)rsi1 * n2 + rsj1] = (slt_ij - slt_ji) * cs;
      }
    }
  }
}

void sync_modified_parameter(Model& m, bool shift_correlations) {
  const int ng = m.n_gutz, nj = m.n_jast;
  bool all_active = true;
  for (int i = 0; i < ng + nj; ++i) all_active = all_active && m.opt_flags[2 * i] == 1;
  if (shift_correlations && ng > 0 && nj > 0 && all_active) {
    double shift = 0.0;
    for (int i = 0; i < ng + nj; ++i) shift += m.proj[i];
    shift /= (double)(ng + nj);
    for (int i = 0; i < ng + nj; ++i) m.proj[i] -= shift;
  }
  double xmax = 0.0;
  for (double v : m.slater) {
    double a = std::fabs(v);
    if (a > xmax) xmax = a;
  }
  if (xmax > 0.0) {
    const double ratio = 4.0 / xmax;  // D_AMP_MAX
    for (double& v : m.slater) v *= ratio;
  }
}

void direct_projection_tables(const Model& m, std::vector<double>& gutz, std::vector<double>& jast) {
  const int n = m.nsite();
  gutz.assign(n, 0.0);
  jast.assign((size_t)n * n, 0.0);
  for (int s = 0; s < n; ++s) {
    int idx = m.gutz_idx[s];
    if (m.n_gutz > 0 && idx >= 0 && idx < m.n_gutz) gutz[s] = m.proj[idx];
  }
  if (m.n_jast > 0)
    for (int a = 0; a < n; ++a)
      for (int b = 0; b < n; ++b) {
        if (a == b) continue;
        int idx = m.jast(a, b);
        if (idx >= 0 && idx < m.n_jast) jast[(size_t)a * n + b] = m.proj[m.n_gutz + idx];
      }
}

}  // namespace mvmc
