// ORACLE — test infrastructure only (see expert.h for the reference lines each
// part restates).  Written independently of the product's loader: the checker
// must not run the code it checks.
#include "expert.h"

#include <cctype>
#include <cmath>
#include <cstdlib>
#include <fstream>
#include <map>
#include <sstream>
#include <stdexcept>

namespace oracle {
namespace {

[[noreturn]] void bad(const std::string& path, const std::string& why) {
  throw std::runtime_error(path + ": " + why);
}

// A .def file: physical lines, the "%s %d" count of header line 2 and the
// whitespace fields of everything after the five header lines.
struct DefFile {
  std::string path;
  std::vector<std::string> lines;
  explicit DefFile(const std::string& p) : path(p) {
    std::ifstream f(p);
    if (!f) bad(p, "cannot open");
    std::string s;
    while (std::getline(f, s)) {
      if (!s.empty() && s.back() == '\r') s.pop_back();
      lines.push_back(s);
    }
  }
  static std::vector<std::string> fields(const std::string& s) {
    std::vector<std::string> out;
    std::istringstream is(s);
    for (std::string t; is >> t;) out.push_back(t);
    return out;
  }
  // second field of header line `k` (0-based) as an integer
  int header_int(size_t k) const {
    if (lines.size() <= k) bad(path, "missing header line");
    auto f = fields(lines[k]);
    if (f.size() < 2) bad(path, "header line without a count");
    return as_int(f[1]);
  }
  std::vector<std::string> body_fields() const {
    std::vector<std::string> out;
    for (size_t i = 5; i < lines.size(); ++i)
      for (auto& t : fields(lines[i])) out.push_back(t);
    return out;
  }
  std::vector<std::vector<std::string>> body_rows() const {
    std::vector<std::vector<std::string>> out;
    for (size_t i = 5; i < lines.size(); ++i) {
      auto f = fields(lines[i]);
      if (!f.empty()) out.push_back(f);
    }
    return out;
  }
  int as_int(const std::string& s) const {
    char* e = nullptr;
    long v = std::strtol(s.c_str(), &e, 10);
    if (e == s.c_str() || *e) bad(path, "not an integer: '" + s + "'");
    return (int)v;
  }
  double as_double(const std::string& s) const {
    char* e = nullptr;
    double v = std::strtod(s.c_str(), &e);
    if (e == s.c_str()) bad(path, "not a number: '" + s + "'");
    return v;
  }
};

std::string lowered(std::string s) {
  for (auto& c : s) c = (char)std::tolower((unsigned char)c);
  return s;
}

// GetInfoFromModPara (parsers/modpara.rs:82-260): lines 6 and 7 carry the file
// heads, line 8 is skipped, then "keyword value" lines; '-' lines are rules.
void read_modpara(const std::string& path, Params& p) {
  DefFile f(path);
  if (f.lines.size() < 8) bad(path, "fewer than 8 header lines");
  int diag_scale = 0;
  double value = 0.0;
  for (size_t i = 8; i < f.lines.size(); ++i) {
    const std::string& l = f.lines[i];
    if (l.empty() || l[0] == '-') continue;
    auto t = DefFile::fields(l);
    if (t.empty()) continue;
    if (t.size() > 1) value = f.as_double(t[1]);
    const int iv = (int)value;
    const std::string k = lowered(t[0]);
    if (k == "nvmccalmode") p.vmc_calc_mode = iv;
    else if (k == "nlanczosmode" || k == "ndataidxstart" || k == "ndataqtysmp" || k == "nsplitsize" ||
             k == "2sz" || k == "reweight") {
    } else if (k == "nsite") p.nsite = iv;
    else if (k == "ne" || k == "nelectron") p.nelec = iv;
    else if (k == "ncond") p.ncond = iv;
    else if (k == "nspgaussleg") p.nsp_gauss_leg = iv;
    else if (k == "nspstot") p.nsp_stot = iv;
    else if (k == "nmptrans") p.nmp_trans = iv;
    else if (k == "nsroptitrstep") p.nsr_opt_itr_step = iv;
    else if (k == "nsroptitrsmp") p.nsr_opt_itr_smp = iv;
    else if (k == "dsroptredcut") p.dsr_opt_red_cut = value;
    else if (k == "dsroptstadel") p.dsr_opt_sta_del = value;
    else if (k == "dsroptstepdt") p.dsr_opt_step_dt = value;
    else if (k == "nsroptcgmaxiter") p.nsr_opt_cg_max_iter = iv;
    else if (k == "dsroptcgtol") p.dsr_opt_cg_tol = value;
    else if (k == "nvmcwarmup") p.nvmc_warmup = iv;
    else if (k == "nvmcinterval") p.nvmc_interval = iv;
    else if (k == "nvmcsample") p.nvmc_sample = iv;
    else if (k == "nexupdatepath") p.nex_update_path = iv;
    else if (k == "rndseed") p.rnd_seed = iv;
    else if (k == "nstore") p.nstore = iv;
    else if (k == "nsrcg") p.nsrcg = iv;
    else if (k == "nsrcgfallback") p.nsrcg_fallback = iv != 0;
    else if (k == "nsrcgabortonfail") p.nsrcg_abort_on_fail = iv != 0;
    else if (k == "rescalesmat") p.rescale_smat = iv != 0;
    else if (k == "usediagscale") diag_scale = iv;
    else bad(path, "keyword '" + t[0] + "' is not one of the path's");
  }
  // readdef.c:927-957: NSRCG 2 is the diagonally scaled CG; useDiagScale
  // only survives with NSRCG 1
  if (p.nsrcg == 2) {
    diag_scale = 1;
    p.nsrcg = 1;
  }
  p.use_diag_scale = (diag_scale != 0 && p.nsrcg == 1) ? 1 : 0;
  // readdef.c:959-975 (modpara.rs:250-266): a negative DSROptStepDt selects
  // the diagonalisation mode (SRFlag) and is negated; SR-CG has no such mode
  if (p.dsr_opt_step_dt < 0.0) {
    p.sr_flag = 1;
    p.dsr_opt_step_dt *= -1.0;
  } else {
    p.sr_flag = 0;
  }
  if (p.sr_flag && p.nsrcg != 0) bad(path, "SR-CG cannot be combined with a negative DSROptStepDt");
}

// (index, flag) pairs in file order: the flag of parameter k is the k-th
// pair's second field, whatever index is printed (parsers/gutzwiller.rs,
// jastrow.rs, orbital.rs); real components only, imaginary flags 0.
void take_flags(const DefFile& f, const std::vector<std::string>& t, size_t first, int count, int offset,
                std::vector<int>& flags) {
  if (t.size() - first != 2 * (size_t)count) bad(f.path, "flag section must hold exactly the declared pairs");
  for (int k = 0; k < count; ++k) {
    (void)f.as_int(t[first + 2 * k]);
    flags[2 * (offset + k)] = f.as_int(t[first + 2 * k + 1]);
    flags[2 * (offset + k) + 1] = 0;
  }
}

void require_real(const DefFile& f) {
  if (f.header_int(2) > 0) bad(f.path, "complex parameters are outside the oracle's scope");
}

}  // namespace

Model read_expert(const std::string& namelist) {
  Model m;
  const auto slash = namelist.find_last_of('/');
  const std::string dir = slash == std::string::npos ? std::string(".") : namelist.substr(0, slash);
  std::map<std::string, std::string> file;
  {
    DefFile nl(namelist);
    for (auto& l : nl.lines) {
      auto t = DefFile::fields(l);
      if (t.size() >= 2) file[lowered(t[0])] = dir + "/" + t[1];
    }
  }
  auto named = [&](const char* k) { return file.count(k) > 0; };
  if (!named("modpara")) bad(namelist, "no ModPara entry");
  read_modpara(file["modpara"], m.mp);
  for (const char* k : {"pairhop", "interall", "dh2", "dh4", "orbitalparallel", "orbitalgeneral", "opttrans", "rbm"})
    if (named(k)) bad(namelist, std::string("section outside the oracle's scope: ") + k);
  const int ns = m.mp.nsite;
  if (ns <= 0) bad(namelist, "Nsite must be positive");

  // LocSpin (parsers/locspin.rs:34-105): one record per site after the header
  m.loc_spn.assign(ns, 0);
  if (named("locspin")) {
    DefFile f(file["locspin"]);
    m.mp.nlocspin = f.header_int(1);
    auto rows = f.body_rows();
    if ((int)rows.size() != ns) bad(f.path, "one record per site is required");
    for (auto& r : rows) {
      if (r.size() < 2) bad(f.path, "short record");
      const int i = f.as_int(r[0]);
      if (i < 0 || i >= ns) bad(f.path, "site out of range");
      m.loc_spn[i] = f.as_int(r[1]) == 1 ? 1 : 0;
    }
  }
  // lib.rs:313-320 (readdef.c:661): Ne = (Nlocspin + Ncond) / 2 for an even Ncond
  if (m.mp.ncond != -1 && m.mp.ncond % 2 == 0) m.mp.nelec = (m.mp.nlocspin + m.mp.ncond) / 2;

  if (named("trans")) {
    DefFile f(file["trans"]);
    const int n = f.header_int(1);
    auto rows = f.body_rows();
    if ((int)rows.size() < n) bad(f.path, "fewer transfer rows than declared");
    for (int k = 0; k < n; ++k) {
      auto& r = rows[k];
      if (r.size() < 5) bad(f.path, "short transfer row");
      const double im = r.size() > 5 ? f.as_double(r[5]) : 0.0;
      if (im != 0.0) bad(f.path, "complex hopping is outside the oracle's real path");
      m.transfer.push_back({f.as_int(r[0]), f.as_int(r[1]), f.as_int(r[2]), f.as_int(r[3]), f.as_double(r[4])});
    }
  }
  if (named("coulombintra")) {
    DefFile f(file["coulombintra"]);
    const int n = f.header_int(1);
    auto rows = f.body_rows();
    if ((int)rows.size() < n) bad(f.path, "fewer rows than declared");
    for (int k = 0; k < n; ++k) m.coulomb_intra.push_back({f.as_int(rows[k].at(0)), f.as_double(rows[k].at(1))});
  }
  for (auto [key, dst] : {std::pair<const char*, std::vector<Pair>*>{"coulombinter", &m.coulomb_inter},
                          {"hund", &m.hund},
                          {"exchange", &m.exchange}}) {
    if (!named(key)) continue;
    DefFile f(file[key]);
    const int n = f.header_int(1);
    auto rows = f.body_rows();
    if ((int)rows.size() < n) bad(f.path, "fewer rows than declared");
    for (int k = 0; k < n; ++k)
      dst->push_back({f.as_int(rows[k].at(0)), f.as_int(rows[k].at(1)), f.as_double(rows[k].at(2))});
  }

  // Gutzwiller (parsers/gutzwiller.rs:19-95): Nsite (site, idx) pairs, then the flags
  std::vector<std::string> gf, jf, of;
  std::string gp, jp, op;
  m.gutz_idx.assign(ns, -1);
  if (named("gutzwiller")) {
    DefFile f(file["gutzwiller"]);
    gp = f.path;
    m.n_gutz = f.header_int(1);
    if (m.n_gutz <= 0) bad(f.path, "declared width must be positive");
    require_real(f);
    gf = f.body_fields();
    if (gf.size() != 2 * ((size_t)ns + m.n_gutz)) bad(f.path, "Nsite pairs and the declared flags expected");
    for (int k = 0; k < ns; ++k) {
      const int site = f.as_int(gf[2 * k]), idx = f.as_int(gf[2 * k + 1]);
      if (site < 0 || site >= ns || idx < 0 || idx >= m.n_gutz) bad(f.path, "mapping out of range");
      m.gutz_idx[site] = idx;
    }
  }
  // Jastrow (parsers/jastrow.rs:22-150): Nsite(Nsite-1) directed triples,
  // each unordered pair in one class, then the flags
  m.jast_idx.assign((size_t)ns * ns, -1);
  if (named("jastrow")) {
    DefFile f(file["jastrow"]);
    jp = f.path;
    m.n_jast = f.header_int(1);
    if (m.n_jast <= 0) bad(f.path, "declared width must be positive");
    require_real(f);
    jf = f.body_fields();
    const size_t nmap = (size_t)ns * (ns - 1);
    if (jf.size() != 3 * nmap + 2 * (size_t)m.n_jast) bad(f.path, "directed triples and the declared flags expected");
    std::vector<int> seen((size_t)ns * ns, 0);
    for (size_t k = 0; k < nmap; ++k) {
      const int a = f.as_int(jf[3 * k]), b = f.as_int(jf[3 * k + 1]), idx = f.as_int(jf[3 * k + 2]);
      if (a == b || a < 0 || b < 0 || a >= ns || b >= ns || idx < 0 || idx >= m.n_jast) bad(f.path, "bad triple");
      if (seen[(size_t)a * ns + b]++) bad(f.path, "duplicated pair");
      m.jast_idx[(size_t)a * ns + b] = idx;
      if (seen[(size_t)b * ns + a] && m.jast_idx[(size_t)b * ns + a] != idx) bad(f.path, "pair in two classes");
    }
  }
  // Orbital, anti-parallel (parsers/orbital.rs:33-170): Nsite^2 rows
  // "i j idx [sign]" (a sign persists to later rows), then the flags
  m.orb_idx.assign((size_t)ns * ns, -1);
  m.orb_sgn.assign((size_t)ns * ns, 1);
  if (named("orbital") || named("orbitalantiparallel")) {
    DefFile f(named("orbital") ? file["orbital"] : file["orbitalantiparallel"]);
    op = f.path;
    m.n_slater = f.header_int(1);
    if (m.n_slater <= 0) bad(f.path, "declared width must be positive");
    require_real(f);
    const size_t nmap = (size_t)ns * ns;
    if (f.lines.size() < 5 + nmap) bad(f.path, "incomplete mapping section");
    int sign = 1;
    for (size_t k = 0; k < nmap; ++k) {
      auto r = DefFile::fields(f.lines[5 + k]);
      if (r.size() < 3) bad(f.path, "short mapping row");
      const int i = f.as_int(r[0]), j = f.as_int(r[1]), idx = f.as_int(r[2]);
      if (i < 0 || j < 0 || i >= ns || j >= ns || idx < 0 || idx >= m.n_slater) bad(f.path, "mapping out of range");
      if (r.size() >= 4) sign = f.as_int(r[3]);
      m.orb_idx[(size_t)i * ns + j] = idx;
      m.orb_sgn[(size_t)i * ns + j] = sign;
    }
    for (size_t k = 5 + nmap; k < f.lines.size(); ++k)
      for (auto& t : DefFile::fields(f.lines[k])) of.push_back(t);
  }
  m.opt_flags.assign(2 * (size_t)m.n_para(), 0);
  if (!gp.empty()) take_flags(DefFile(gp), gf, 2 * (size_t)ns, m.n_gutz, 0, m.opt_flags);
  if (!jp.empty()) take_flags(DefFile(jp), jf, 3 * (size_t)ns * (ns - 1), m.n_jast, m.n_gutz, m.opt_flags);
  if (!op.empty()) take_flags(DefFile(op), of, 0, m.n_slater, m.n_proj(), m.opt_flags);

  // TransSym (parsers/qptrans.rs:41-200): count weight rows "k w_re [w_im]",
  // then count*Nsite rows "k origin target [sign]"
  const char* tkey = named("transsym") ? "transsym" : (named("qptrans") ? "qptrans" : nullptr);
  if (tkey) {
    DefFile f(file[tkey]);
    const int n = f.header_int(1);
    auto rows = f.body_rows();
    if ((int)rows.size() != n + n * ns) bad(f.path, "record counts differ from the declaration");
    m.tr_weight.assign(n, 0.0);
    m.tr_site.assign(n, std::vector<int>(ns, -1));
    m.tr_sign.assign(n, std::vector<int>(ns, 1));
    for (int k = 0; k < n; ++k) {
      const int idx = f.as_int(rows[k].at(0));
      if (idx < 0 || idx >= n) bad(f.path, "weight index out of range");
      if (rows[k].size() > 2 && f.as_double(rows[k][2]) != 0.0) bad(f.path, "complex weights are outside scope");
      m.tr_weight[idx] = rows[k].size() > 1 ? f.as_double(rows[k][1]) : 0.0;
    }
    for (int k = 0; k < n * ns; ++k) {
      auto& r = rows[n + k];
      const int t = f.as_int(r.at(0)), a = f.as_int(r.at(1)), b = f.as_int(r.at(2));
      if (t < 0 || t >= n || a < 0 || a >= ns || b < 0 || b >= ns) bad(f.path, "map row out of range");
      m.tr_site[t][a] = b;
      m.tr_sign[t][a] = r.size() > 3 ? f.as_int(r[3]) : 1;
    }
  }
  if (m.n_mp_trans() > (int)m.tr_weight.size()) bad(namelist, "NMPTrans exceeds the TransSym entries");
  m.proj.assign(m.n_proj(), 0.0);
  m.slater.assign(m.n_slater, 0.0);
  qp_weights(m);
  return m;
}

// initial_params.rs:9-133: every complete record is applied in turn, so the
// last one holds; value = re + 0*im as C builds it.
bool read_initial(Model& m, const std::string& path) {
  std::ifstream f(path);
  if (!f) return false;
  std::vector<double> v;
  for (std::string t; f >> t;) {
    char* e = nullptr;
    const double x = std::strtod(t.c_str(), &e);
    if (e == t.c_str()) bad(path, "not a number: '" + t + "'");
    v.push_back(x);
  }
  if (v.empty()) return false;
  const size_t np = (size_t)m.n_para(), rec = 6 + 3 * np;
  if (v.size() < rec || v.size() % rec) bad(path, "not a whole number of parameter records");
  const double* r = v.data() + (v.size() - rec) + 6;
  for (size_t i = 0; i < np; ++i) {
    if (r[3 * i + 1] != 0.0) bad(path, "complex values are outside the oracle's real path");
    const double x = r[3 * i] + 0.0 * r[3 * i + 1];
    if (i < (size_t)m.n_proj()) m.proj[i] = x;
    else m.slater[i - m.n_proj()] = x;
  }
  return true;
}

// qp_weight.rs:24-60 GaussLeg on [x1, x2] (Newton until |dz| <= 5e-14) and
// qp_weight.rs:64-82 the Legendre recurrence.
void gauss_leg(double x1, double x2, int n, std::vector<double>& x, std::vector<double>& w) {
  x.assign(n, 0.0);
  w.assign(n, 0.0);
  const double mid = 0.5 * (x2 + x1), half = 0.5 * (x2 - x1), fn = (double)n;
  for (int i = 0; i < (n + 1) / 2; ++i) {
    double z = std::cos(M_PI * ((double)i + 0.75) / (fn + 0.5)), dp = 0.0;
    for (;;) {
      double pk = 1.0, pkm1 = 0.0;  // P_k(z), P_{k-1}(z)
      for (int k = 1; k <= n; ++k) {
        const double pkm2 = pkm1;
        pkm1 = pk;
        const double fk = (double)k;
        pk = ((2.0 * fk - 1.0) * z * pkm1 - (fk - 1.0) * pkm2) / fk;
      }
      dp = fn * (z * pk - pkm1) / (z * z - 1.0);
      const double zold = z;
      z = zold - pk / dp;
      if (std::fabs(z - zold) <= 5.0e-14) break;
    }
    x[i] = mid - half * z;
    x[n - 1 - i] = mid + half * z;
    w[i] = 2.0 * half / ((1.0 - z * z) * dp * dp);
    w[n - 1 - i] = w[i];
  }
}

static double legendre(double x, int n) {
  if (n <= 0) return 1.0;
  if (n == 1) return x;
  double a = 1.0, b = x, c = x;
  for (int i = 2; i <= n; ++i) {
    const double fi = (double)i;
    c = (1.0 / fi) * ((2.0 * fi - 1.0) * x * b - (fi - 1.0) * a);
    a = b;
    b = c;
  }
  return c;
}

// qp_weight.rs:86-146: one GL point -> cos^2 = 1 and the translation
// weights; otherwise w = 0.5 sin(beta) w_GL P_S(cos beta) times each
// translation weight, QP index i + j*NSPGaussLeg.
void qp_weights(Model& m) {
  const int nl = m.n_sp_gl(), nt = m.n_mp_trans();
  m.qp_full_weight.assign((size_t)nl * nt, 0.0);
  m.spgl_cos_sin.assign(nl, 0.0);
  m.spgl_cos_cos.assign(nl, 0.0);
  m.spgl_sin_sin.assign(nl, 0.0);
  auto tw = [&](int j) { return j < (int)m.tr_weight.size() ? m.tr_weight[j] : 0.0; };
  if (nl == 1) {
    m.spgl_cos_cos[0] = 1.0;
    for (int j = 0; j < nt; ++j) m.qp_full_weight[j] = tw(j);
    return;
  }
  std::vector<double> beta, wgl;
  gauss_leg(0.0, M_PI, nl, beta, wgl);
  for (int i = 0; i < nl; ++i) {
    const double c = std::cos(0.5 * beta[i]), s = std::sin(0.5 * beta[i]);
    m.spgl_cos_sin[i] = c * s;
    m.spgl_cos_cos[i] = c * c;
    m.spgl_sin_sin[i] = s * s;
    const double w = 0.5 * std::sin(beta[i]) * wgl[i] * legendre(std::cos(beta[i]), m.mp.nsp_stot);
    for (int j = 0; j < nt; ++j)
      if (j < (int)m.tr_weight.size()) m.qp_full_weight[i + j * nl] = tw(j) * w;
  }
}

// qp.rs:33-70 with no OptTrans: the TransSym map, its sign only when
// NMPTrans < 0 (anti-periodic boundary, types.rs:627-633).
void translate(const Model& m, int site, int t, int& out, int& sign) {
  out = site;
  sign = 1;
  if (t < (int)m.tr_site.size()) {
    out = m.tr_site[t][site];
    if (m.mp.nmp_trans < 0) sign = m.tr_sign[t][site];
  }
}

// slater_update.rs:17-126, real branch: for QP plane qp (spin point
// qp % NSPGaussLeg, translation (qp % (NSPGaussLeg NMPTrans)) / NSPGaussLeg)
// the 2x2 spin block of sites (ri, rj) from f_ij, f_ji at the translated sites.
void slater_table(const Model& m, std::vector<double>& S) {
  const int ns = m.nsite(), n2 = m.nsite2(), nl = m.n_sp_gl(), nq = m.nqp();
  S.assign((size_t)nq * n2 * n2, 0.0);
  for (int q = 0; q < nq; ++q) {
    const int sp = q % nl, tr = (q % (nl * m.n_mp_trans())) / nl;
    const double cs = m.spgl_cos_sin[sp], cc = m.spgl_cos_cos[sp], ss = m.spgl_sin_sin[sp];
    double* P = S.data() + (size_t)q * n2 * n2;
    for (int ri = 0; ri < ns; ++ri) {
      int ti, si;
      translate(m, ri, tr, ti, si);
      if (ti < 0 || ti >= ns) continue;
      for (int rj = 0; rj < ns; ++rj) {
        int tj, sj;
        translate(m, rj, tr, tj, sj);
        if (tj < 0 || tj >= ns) continue;
        const int a = m.orb_idx[(size_t)ti * ns + tj], b = m.orb_idx[(size_t)tj * ns + ti];
        if (a < 0 || b < 0) continue;
        const int sg = si * sj;
        const double fij = m.slater[a] * (double)(m.orb_sgn[(size_t)ti * ns + tj] * sg);
        const double fji = m.slater[b] * (double)(m.orb_sgn[(size_t)tj * ns + ti] * sg);
        P[(size_t)ri * n2 + rj] = -(fij - fji) * cs;
        P[(size_t)ri * n2 + rj + ns] = fij * cc + fji * ss;
        P[(size_t)(ri + ns) * n2 + rj] = -fij * ss - fji * cc;
        P[(size_t)(ri + ns) * n2 + rj + ns] = (fij - fji) * cs;
      }
    }
  }
}

// parameter_init.rs:222-225 / opt_flag.rs:311-322: every Gutzwiller and
// Jastrow real component flag is exactly 1 (a missing entry counts as not
// optimised), and both families exist.
bool gj_shift_gate(int n_gutz, int n_jast, const std::vector<int>& flags) {
  if (n_gutz <= 0 || n_jast <= 0) return false;
  if (flags.empty()) return true;
  for (int i = 0; i < n_gutz + n_jast; ++i) {
    const size_t c = 2 * (size_t)i;
    if (c >= flags.size() || flags[c] != 1) return false;
  }
  return true;
}

// parameter_init.rs:167-260 without doublon-holon sections: the mean of the
// Gutzwiller and Jastrow values, summed Gutzwiller first in index order, is
// subtracted from each; then every Slater value is scaled by 4 / max |f|.
void sync_parameters(Model& m, bool shift_correlations) {
  if (shift_correlations && gj_shift_gate(m.n_gutz, m.n_jast, m.opt_flags)) {
    double mean = 0.0;
    for (int i = 0; i < m.n_gutz + m.n_jast; ++i) mean += m.proj[i];
    mean /= (double)(m.n_gutz + m.n_jast);
    for (int i = 0; i < m.n_gutz + m.n_jast; ++i) m.proj[i] -= mean;
  }
  double big = 0.0;
  for (double f : m.slater) big = std::hypot(f, 0.0) > big ? std::hypot(f, 0.0) : big;
  if (big > 0.0) {
    const double scale = 4.0 / big;  // D_AMP_MAX
    for (double& f : m.slater) f *= scale;
  }
}

}  // namespace oracle
