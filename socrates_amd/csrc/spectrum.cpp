// Host-side reader of SOCRATES spectral files and McICA data files into the
// device table layout of sg_types.h.
//
// Formats: the text blocks read by read_spectrum.F90 (block dispatch at
// :212-240; readers 0_0_1 :557-621, 2_0_0 :764-789, 3_0_0 :872-900,
// 4_0_0 :951-992, 5_0_0/5_0_1 :1040-1641, 6_0_0 :1746-1777, 6_0_1 polynomial
// :1781-1877, 8_0_0 :1881-1935, 10_0_2 :2245-2331) and by read_mcica_data
// (def_mcica.F90:61-190). Features off the two-stream flux path - k-term
// look-up tables, continua, SES2, aerosol-only data - are rejected with a
// clear error, never silently dropped.
#include <cerrno>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <sstream>
#include <string>
#include <vector>

#include "sg_types.h"
#include "spectrum.h"

namespace sg {

namespace {

std::string trim(const std::string &s) {
  size_t a = s.find_first_not_of(" \t\r"), b = s.find_last_not_of(" \t\r");
  return a == std::string::npos ? std::string() : s.substr(a, b - a + 1);
}

// Fixed-column field of a Fortran formatted record (0-based start, width).
std::string col(const std::string &ln, size_t c0, size_t w) {
  if (c0 >= ln.size()) return std::string();
  return ln.substr(c0, w);
}
double fnum(const std::string &ln, size_t c0, size_t w) {
  std::string f = trim(col(ln, c0, w));
  return f.empty() ? 0.0 : std::strtod(f.c_str(), nullptr);
}
int inum(const std::string &ln, size_t c0, size_t w) {
  std::string f = trim(col(ln, c0, w));
  return f.empty() ? 0 : std::atoi(f.c_str());
}
std::vector<double> numbers(const std::string &ln) {
  std::vector<double> v;
  const char *p = ln.c_str();
  char *e;
  for (;;) {
    double x = std::strtod(p, &e);
    if (e == p) break;
    v.push_back(x);
    p = e;
  }
  return v;
}
int after_eq_int(const std::string &ln) {
  size_t e = ln.find('=');
  return e == std::string::npos ? 0 : std::atoi(ln.c_str() + e + 1);
}

struct Lines {
  std::vector<std::string> v;
  size_t i = 0;
  bool more() const { return i < v.size(); }
  const std::string &next() {
    static const std::string empty;
    return i < v.size() ? v[i++] : empty;
  }
};

int gas_slot(int type_absorb) {
  switch (type_absorb) {  // identifiers of rad_pcf.F90 (ip_h2o, ip_co2, ip_o3)
    case 1: return SG_GAS_H2O;
    case 2: return SG_GAS_CO2;
    case 3: return SG_GAS_O3;
    default: return SG_GAS_ZERO;
  }
}

}  // namespace

int read_spectrum(const char *path, SgSpec *s, std::string *err) {
  std::ifstream f(path);
  if (!f) { *err = std::string("cannot open spectral file ") + path; return -1; }
  Lines L;
  for (std::string ln; std::getline(f, ln);) {
    if (!ln.empty() && ln.back() == '\r') ln.pop_back();
    L.v.push_back(ln);
  }
  std::memset(s, 0, sizeof(*s));
  std::vector<int> type_absorb;
  std::vector<int> n_band_absorb;
  std::vector<std::vector<int>> index_absorb;
  // per (band, gas index): term list
  struct Raw { int nk = 0, scale = 0, fnc = 0; double pref = 0, tref = 0;
               std::vector<double> k, w, a, b; };
  std::vector<std::vector<Raw>> raw;
  bool have5 = false, have6 = false;
  while (L.more()) {
    const std::string ln = L.next();
    if (ln.compare(0, 6, "*BLOCK") != 0) continue;
    int type = inum(ln, 15, 4), ver = inum(ln, 47, 4);
    if (type == 0) {
      L.next();
      int nabs = 0;
      for (;;) {
        std::string h = L.next();
        std::string lab = trim(h.substr(0, h.find('=')));
        if (lab == "Number of spectral bands") s->n_band = after_eq_int(h);
        else if (lab == "Total number of gaseous absorbers") nabs = after_eq_int(h);
        else if (lab.rfind("List of indexing numbers and absorbers", 0) == 0) break;
        else if (h.compare(0, 4, "*END") == 0 || !L.more()) { *err = "bad block 0"; return -1; }
      }
      L.next();
      for (int i = 0; i < nabs; i++) type_absorb.push_back(inum(L.next(), 12, 5));
      if (nabs > SG_MAXABS) { *err = "more than 32 gaseous absorbers"; return -3; }
      s->n_absorb = nabs;
      for (int i = 0; i < nabs; i++) s->type_absorb[i] = type_absorb[i];
      if (s->n_band <= 0 || s->n_band > SG_MAXB) { *err = "band count outside 1..16"; return -1; }
    } else if (type == 2) {
      if (ver != 0) { *err = "solar block version not supported"; return -3; }
      L.next(); L.next();
      for (int b = 0; b < s->n_band; b++) s->band[b].solar_flux = fnum(L.next(), 12, 16);
    } else if (type == 3) {
      if (ver != 0) { *err = "only the total Rayleigh scheme (block 3.0.0) is supported"; return -3; }
      L.next(); L.next(); L.next();
      for (int b = 0; b < s->n_band; b++) s->band[b].rayleigh = fnum(L.next(), 12, 16);
    } else if (type == 4) {
      for (int j = 0; j < 5; j++) L.next();
      n_band_absorb.assign(s->n_band, 0);
      index_absorb.assign(s->n_band, {});
      for (int b = 0; b < s->n_band; b++) {
        const std::string h = L.next();
        n_band_absorb[b] = inum(h, 12, 5);
        bool first = true;
        while ((int)index_absorb[b].size() < n_band_absorb[b]) {
          const std::string d = L.next();
          size_t c0 = first ? 5 : 0;
          first = false;
          for (int q = 0; q < 4 && (int)index_absorb[b].size() < n_band_absorb[b]; q++)
            index_absorb[b].push_back(inum(d, c0 + 5 * q + 2, 3));
        }
      }
    } else if (type == 5) {
      if (n_band_absorb.empty()) { *err = "block 5 before block 4"; return -1; }
      raw.assign(s->n_band, std::vector<Raw>(type_absorb.size()));
      if (ver == 0) { L.next(); L.next(); L.next(); }
      else {
        while (L.more() && L.next().compare(0, 34, "Band        Gas, Number of k-terms") != 0) {}
        L.next(); L.next();
      }
      for (int b = 0; b < s->n_band; b++)
        for (int j = 0; j < n_band_absorb[b]; j++) {
          const std::string h = L.next();
          int band = inum(h, 0, 5) - 1, gas = inum(h, 12, 5) - 1;
          if (band < 0 || band >= s->n_band || gas < 0 || gas >= (int)type_absorb.size()) {
            *err = "bad k-term header in block 5"; return -1;
          }
          Raw &r = raw[band][gas];
          r.nk = inum(h, 24, 5); r.scale = inum(h, 36, 5); r.fnc = inum(h, 48, 5);
          if (r.fnc != 1 || r.scale != 2) {
            *err = "only per-term power-law scaling (ip_scale_term, ip_scale_power_law) is on this path";
            return -3;
          }
          const std::string pt = L.next();
          r.pref = fnum(pt, 6, 16); r.tref = fnum(pt, 28, 16);
          for (int k = 0; k < r.nk; k++) {
            const std::string d = L.next();
            r.k.push_back(fnum(d, 3, 16)); r.w.push_back(fnum(d, 22, 16));
            r.a.push_back(fnum(d, 41, 16)); r.b.push_back(fnum(d, 60, 16));
          }
        }
      have5 = true;
    } else if (type == 6) {
      std::string h;
      if (ver == 1) {
        h = L.next();  // "Thermal source function."
        h = L.next();
        if (trim(col(h, 14, 10)).compare(0, 5, "table") == 0) {
          *err = "tabulated Planck functions are not supported"; return -3;
        }
      } else {
        L.next();
      }
      h = L.next();
      s->n_deg = after_eq_int(h);
      s->t_ref_planck = std::strtod(h.c_str() + h.rfind('=') + 1, nullptr);
      if (s->n_deg < 0 || s->n_deg >= SG_NPL) { *err = "Planck polynomial degree too high"; return -3; }
      L.next(); L.next();
      for (int b = 0; b < s->n_band; b++) {
        std::vector<double> v;
        while ((int)v.size() < s->n_deg + 2) {
          std::vector<double> x = numbers(L.next());
          v.insert(v.end(), x.begin(), x.end());
        }
        for (int k = 0; k <= s->n_deg; k++) s->band[b].planck[k] = v[k + 1];
      }
      have6 = true;
    } else if (type == 8) {
      for (int j = 0; j < 5; j++) L.next();
      for (int b = 0; b < s->n_band; b++) {
        if (inum(L.next(), 12, 5) != 0) {
          *err = "continuum absorption is not on this path"; return -3;
        }
      }
    } else if (type == 10) {
      if (ver != 2) { *err = "droplet block version not supported"; return -3; }
      L.next();
      int idrop = inum(L.next(), 27, 5);
      const std::string h = L.next();
      int scheme = inum(h, 34, 5), npar = inum(h, 66, 5);
      L.next();
      const std::string r = L.next();
      if (idrop < 1 || idrop > SG_NDROP) { *err = "droplet type index above 8"; return -3; }
      SgDrop &d = s->drop[idrop - 1];
      d.present = 1; d.scheme = scheme; d.n_parm = npar;
      d.dim_min = fnum(r, 39, 12); d.dim_max = fnum(r, 55, 12);
      if (npar > SG_NDPARM) { *err = "too many droplet parameters"; return -3; }
      for (int b = 0; b < s->n_band; b++) {
        L.next();
        int got = 0;
        while (got < npar) {
          std::vector<double> x = numbers(L.next());
          for (double v : x) if (got < npar) d.parm[b][got++] = v;
        }
      }
    }
    // other blocks (1, 9 without continua, 11, 12, 14, ...) carry nothing the
    // flux path reads; they are skipped to their *END by the scan above
  }
  if (!have5) { *err = "no k-distribution block (5)"; return -1; }
  // assemble the term tables: major gas is the first absorber listed for the
  // band (radiance_calc.F90: index_abs(1) is the major absorber)
  int ng = 0, nm = 0;
  for (int b = 0; b < s->n_band; b++) {
    SgBand &B = s->band[b];
    if (n_band_absorb[b] < 1) { *err = "band without gaseous absorption"; return -3; }
    for (int j = 0; j < n_band_absorb[b]; j++) {
      int gi = index_absorb[b][j] - 1;
      const Raw &r = raw[b][gi];
      int slot = gas_slot(type_absorb[gi]);
      if (j == 0) { B.n_major = r.nk; B.g0 = ng; }
      else {
        if (j - 1 >= SG_MAXMINOR_GAS) { *err = "too many minor gases in a band"; return -3; }
        if (j == 1) B.minor0 = nm;
        B.minor_nk[j - 1] = r.nk;
        B.n_minor_gas = j;
      }
      for (int k = 0; k < r.nk; k++) {
        SgTerm t;
        std::memset(&t, 0, sizeof t);
        t.k = r.k[k]; t.w = r.w[k]; t.a = r.a[k]; t.b = r.b[k];
        t.lnpref = std::log(r.pref); t.lntref = std::log(r.tref);
        t.band = b; t.gas = slot;
        if (j == 0) {
          if (ng >= SG_MAXG) { *err = "too many major k-terms"; return -3; }
          s->g[ng++] = t;
        } else {
          if (nm >= SG_MAXMINOR) { *err = "too many minor k-terms"; return -3; }
          s->minor[nm++] = t;
        }
      }
    }
  }
  s->n_g = ng;
  s->n_minor = nm;
  s->n_mband = 0;
  for (int b = 0; b < s->n_band; b++)
    if (s->band[b].n_minor_gas > 0) s->mband[s->n_mband++] = b;
  (void)have6;
  return 0;
}

int read_mcica(const char *path, McicaData *m, std::string *err) {
  std::ifstream f(path);
  if (!f) { *err = std::string("cannot open McICA data file ") + path; return -1; }
  std::string ln;
  bool data = false;
  while (std::getline(f, ln)) if (ln.compare(0, 5, "*DATA") == 0) { data = true; break; }
  if (!data) { *err = "no *DATA block in McICA data file"; return -1; }
  auto read_int = [&](int &x) { f >> x; std::getline(f, ln); };
  while (std::getline(f, ln)) {
    std::string t = trim(ln);
    if (t.compare(0, 4, "*END") == 0) break;
    if (t == "tot_subcol_gen" || t == "n_subcol_gen") read_int(m->n_gen);
    else if (t == "subcol_need_single" || t == "n_subcol_req_single") read_int(m->n_req_single);
    else if (t == "subcol_need_optimal" || t == "n_subcol_req_optimal") read_int(m->n_req_optimal);
    else if (t == "ipph") read_int(m->ipph);
    else if (t == "lw_subcol_reorder_single" || t == "lw_subcol_reorder_optimal") {
      std::vector<int> &v = t == "lw_subcol_reorder_single" ? m->reorder_single : m->reorder_optimal;
      int n = t == "lw_subcol_reorder_single" ? m->n_req_single : m->n_req_optimal;
      v.resize(n);
      for (int i = 0; i < n; i++) f >> v[i];
      std::getline(f, ln);
    } else if (t == "sw_subcol_k" || t == "lw_subcol_k") {
      auto &tab = t[0] == 's' ? m->sw_subcol_k : m->lw_subcol_k;
      for (;;) {
        if (!std::getline(f, ln)) { *err = "truncated subcol_k table"; return -1; }
        int b, k, n;
        if (std::sscanf(ln.c_str(), "%d %d %d", &b, &k, &n) != 3) { *err = "bad subcol_k line"; return -1; }
        if (b == -99) break;
        tab.push_back({b, k, n});
      }
    } else if (t == "n1") read_int(m->n1);
    else if (t == "n2") read_int(m->n2);
    else if (t == "xcw") {
      m->xcw.resize((size_t)m->n1 * m->n2);
      for (auto &x : m->xcw) f >> x;
      std::getline(f, ln);
    }
  }
  if (m->n_gen <= 0 || m->n1 < 2 || m->n2 < 2 || m->xcw.size() != (size_t)m->n1 * m->n2) {
    *err = "incomplete McICA data"; return -1;
  }
  return 0;
}

// Sampling tables of set_cld_mcica (socrates_set_cld_mcica.F90:103-151):
// subcol_k per (band, k), first_subcol_k as running sums across bands, and
// the sampled generated sub-column of each sample m:
// subcol_reorder(mod(m, nd_subcol_req)).
int attach_mcica(SgSpec *s, const McicaData &m, bool solar, std::string *err) {
  if (m.n_gen > SG_MAXSUBCOL) { *err = "more than 64 generated sub-columns"; return -3; }
  s->has_mcica = 1;
  s->n_gen = m.n_gen;
  s->ipph = m.ipph;
  s->n1 = m.n1;
  s->n2 = m.n2;
  for (int mode = 0; mode < 2; mode++) {  // 0 single, 1 optimal
    int n_req = mode == 0 ? m.n_req_single : m.n_req_optimal;
    if (n_req <= 0) { s->n_req[mode] = 0; continue; }
    s->n_req[mode] = n_req;
    std::vector<int> reorder(n_req + 1);
    for (int l = 1; l <= n_req; l++) {
      int v;
      if (solar) v = l % m.n_gen;
      else {
        const std::vector<int> &src = mode == 0 ? m.reorder_single : m.reorder_optimal;
        if ((int)src.size() < n_req) { *err = "McICA reorder table too short"; return -1; }
        v = src[l - 1] % m.n_gen;
      }
      reorder[l] = v == 0 ? m.n_gen : v;
    }
    int item = 0, msamp = 1;
    for (int gi = 0; gi < s->n_g; gi++) {
      const SgTerm &t = s->g[gi];
      int kk = gi - s->band[t.band].g0 + 1;
      int nsub = 1;
      if (mode == 1)
        for (const auto &e : (solar ? m.sw_subcol_k : m.lw_subcol_k))
          if (e[0] == t.band + 1 && e[1] == kk) nsub = e[2];
      s->g_n_sub[mode][gi] = nsub;
      s->g_sub0[mode][gi] = item;
      for (int q = 0; q < nsub; q++, msamp++) {
        if (item >= SG_MAXSUB) { *err = "too many McICA samples"; return -3; }
        int idx = msamp % n_req;
        if (idx == 0) idx = n_req;
        s->sub_g[mode][item] = gi;
        s->sub_isub[mode][item] = reorder[idx];
        s->sub_inv_s[mode][item] = 1.0 / (double)nsub;
        item++;
      }
    }
    s->n_sub_items[mode] = item;
  }
  return 0;
}

}  // namespace sg
