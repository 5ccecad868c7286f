// Host-side reader of SOCRATES spectral files and McICA data files into the
// device table layout of sg_types.h.
//
// Formats: the text blocks read by read_spectrum.F90 (block dispatch at
// :212-240; readers 0_0_1/0_0_2 :557-728, 2_0_0 :764-789, 3_0_0 :872-900,
// 4_0_0/4_0_1 :951-1036, 5_0_0/5_0_1 :1040-1641 with the k-term look-up
// tables of the '_k' file :1404-1507, 6_0_0 :1746-1777, 6_0_1 polynomial
// :1781-1877, 8_0_0 :1881-1935, 9_0_0 :1939-1997, 10_0_2 :2245-2331) and by
// read_mcica_data (def_mcica.F90:61-190). Features off the two-stream flux
// path - SES2 and temperature-only look-up scaling, self-broadened look-up
// tables, generalised continua, aerosol-only data - are rejected with a clear
// error, never silently dropped.
#include <algorithm>
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

#include <algorithm>
#include "wenyi_tables.h"

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
  switch (type_absorb) {  // identifiers of gas_list_pcf.F90
    case 1: return SG_GAS_H2O;
    case 2: return SG_GAS_CO2;
    case 3: return SG_GAS_O3;
    case 4: return SG_GAS_WM0 + 0;   // N2O
    case 6: return SG_GAS_WM0 + 1;   // CH4
    case 7: return SG_GAS_WM0 + 2;   // O2
    case 9: return SG_GAS_WM0 + 3;   // SO2
    case 25: return SG_GAS_WM0 + 4;  // OCS
    case 14: return SG_GAS_WM0 + 5;  // CFC11
    case 15: return SG_GAS_WM0 + 6;  // CFC12
    case 16: return SG_GAS_WM0 + 7;  // CFC113
    case 17: return SG_GAS_WM0 + 8;  // HCFC22
    case 19: return SG_GAS_WM0 + 9;  // HFC134A
    default: return SG_GAS_ZERO;
  }
}

// n_scale_variable of a scaling function (rad_pcf.F90:505-506)
int n_scale_variable(int fnc) {
  static const int n[11] = {0, 2, 3, 4, 0, 0, 6, 8, 10, 0, 0};
  return fnc >= 0 && fnc <= 10 ? n[fnc] : 0;
}

Lines read_lines(const char *path, bool *ok) {
  Lines L;
  std::ifstream f(path);
  *ok = (bool)f;
  for (std::string ln; std::getline(f, ln);) {
    if (!ln.empty() && ln.back() == '\r') ln.pop_back();
    L.v.push_back(ln);
  }
  return L;
}

// n values of '(6(1PE13.6))' records
bool e13_values(Lines &L, int n, double *out) {
  int got = 0;
  while (got < n) {
    if (!L.more()) return false;
    const std::string &ln = L.next();
    for (int q = 0; q < 6 && got < n; q++) out[got++] = fnum(ln, 13 * q, 13);
  }
  return true;
}

}  // namespace

int read_spectrum(const char *path, SgSpec *s, SpecTables *tab, std::string *err, const std::vector<int> *retain) {
  bool ok = false;
  Lines L = read_lines(path, &ok);
  if (!ok) { *err = std::string("cannot open spectral file ") + path; return -1; }
  std::memset(s, 0, sizeof(*s));
  tab->data.clear();
  tab->wenyi_off = -1;
  std::vector<int> type_absorb;
  std::vector<int> n_band_absorb;
  std::vector<std::vector<int>> index_absorb;
  // per (band, gas index): term list
  struct Raw { int nk = 0, scale = 0, fnc = 0; double pref = 0, tref = 0;
               std::vector<double> k, w, s1, s2, s3; std::vector<int> iscat; int lut0 = -1; };
  std::vector<std::vector<Raw>> raw;
  std::vector<std::vector<int>> order;  // gases per band in block 5's order
  // continua by band: listed types, and the block 9 data stored by type
  std::vector<std::vector<int>> cont_index;
  struct Cont { int fnc = 0; double k = 0, pref = 0, tref = 0, sc[3] = {0, 0, 0}; bool set = false; };
  std::vector<std::vector<Cont>> cont;
  bool have5 = false, have6 = false, lookup = false;
  // spectral sub-bands (block 17): band and k-term of each, in file order
  std::vector<std::array<int, 2>> sub_band;
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
        if (ver == 1) {
          // '(i5, 2(7x, i5))' then '(5x, 12(2x, i3))' (:1012-1023); the
          // per-band overlap method (third field) only matters with
          // ip_overlap_hybrid, which the flux path does not use
          if (n_band_absorb[b] > 12) { *err = "more than 12 absorbers in a band"; return -3; }
          if (n_band_absorb[b] > 0) {
            const std::string d = L.next();
            for (int q = 0; q < n_band_absorb[b]; q++) index_absorb[b].push_back(inum(d, 5 + 5 * q + 2, 3));
          }
          continue;
        }
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
      order.assign(s->n_band, {});
      if (ver == 0) { L.next(); L.next(); L.next(); }
      else {
        for (;;) {
          if (!L.more()) { *err = "bad block 5"; return -1; }
          const std::string h = L.next();
          if (trim(h) == "Self-broadened indexing numbers of all absorbers.") {
            L.next();
            for (int q = 0; q < 6 && L.more(); q++)
              if (inum(L.v[L.i], 2 + 5 * q, 3) > 0) {
                *err = "self-broadened look-up tables are not on this path"; return -3;
              }
            L.next();
          }
          if (h.compare(0, 34, "Band        Gas, Number of k-terms") == 0) break;
        }
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
          order[band].push_back(gas);
          r.nk = inum(h, 24, 5); r.scale = inum(h, 36, 5); r.fnc = inum(h, 48, 5);
          if (r.scale != 2) { *err = "only per-term scaling (ip_scale_term) is on this path"; return -3; }
          if (r.fnc != SG_FNC_POWER_LAW && r.fnc != SG_FNC_POWER_QUAD && r.fnc != SG_FNC_LOOKUP &&
              !(r.fnc == SG_FNC_WENYI && band == 3)) {
            *err = "k-term scaling function not on this path (power law, power quadratic, look-up "
                   "table, and Wenyi's CO2 tables in band 4 are)";
            return -3;
          }
          const std::string pt = L.next();
          if (r.fnc == SG_FNC_LOOKUP) {
            // '(2(7x, i5))' num_ref_p, num_ref_t (:1333-1349)
            const int np = inum(pt, 7, 5), nt = inum(pt, 19, 5);
            if (s->n_pre == 0) { s->n_pre = np; s->n_tmp = nt; }
            else if (np != s->n_pre || nt != s->n_tmp) { *err = "P/T lookup table size is not consistent"; return -1; }
            if (np < 2 || nt < 2 || np > SG_MAXPRE || nt > SG_MAXTMP) {
              *err = "P/T lookup table larger than 64 pressures x 8 temperatures"; return -3;
            }
            lookup = true;
          } else {
            r.pref = fnum(pt, 6, 16); r.tref = fnum(pt, 28, 16);
          }
          const int nsv = n_scale_variable(r.fnc);
          for (int k = 0; k < r.nk; k++) {
            const std::string d = L.next();
            r.k.push_back(fnum(d, 3, 16)); r.w.push_back(fnum(d, 22, 16));
            r.iscat.push_back(inum(d, 38, 3));  // i_scat (read_spectrum.F90:1374-1381)
            r.s1.push_back(nsv >= 1 ? fnum(d, 41, 16) : 0.0);
            r.s2.push_back(nsv >= 2 ? fnum(d, 60, 16) : 0.0);
            r.s3.push_back(nsv >= 3 ? fnum(L.next(), 41, 16) : 0.0);
          }
          if (r.fnc == SG_FNC_WENYI && r.nk > 14) { *err = "Wenyi scaling has 14 k-terms"; return -3; }
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
      // '(////)', per band '(i5, 7x, i5)' and '(5x, 4(2x, i3))' (:1895-1921)
      for (int j = 0; j < 5; j++) L.next();
      cont_index.assign(s->n_band, {});
      for (int b = 0; b < s->n_band; b++) {
        const int n = inum(L.next(), 12, 5);
        if (n > SG_MAXCONT) { *err = "more than 2 continua in a band"; return -3; }
        if (n > 0) {
          const std::string d = L.next();
          for (int q = 0; q < n; q++) cont_index[b].push_back(inum(d, 5 + 5 * q + 2, 3));
        }
      }
    } else if (type == 9) {
      if (ver != 0) { *err = "only block 9.0.0 continua are on this path"; return -3; }
      if (cont_index.empty()) { *err = "block 9 before block 8"; return -1; }
      cont.assign(s->n_band, std::vector<Cont>(SG_MAXCONT + 1));
      L.next(); L.next(); L.next();
      for (int b = 0; b < s->n_band; b++)
        for (size_t j = 0; j < cont_index[b].size(); j++) {
          const std::string h = L.next();
          const int band = inum(h, 0, 5) - 1, ic = inum(h, 12, 5), fnc = inum(h, 24, 5);
          if (band < 0 || band >= s->n_band || ic < 1 || ic > SG_MAXCONT) {
            *err = "continuum outside the self- and foreign-broadened water vapour continua"; return -3;
          }
          if (fnc != SG_FNC_POWER_LAW && fnc != SG_FNC_POWER_QUAD) {
            *err = "continuum scaling function not on this path"; return -3;
          }
          Cont &c = cont[band][ic];
          c.fnc = fnc; c.set = true;
          const std::string pt = L.next();
          c.pref = fnum(pt, 6, 16); c.tref = fnum(pt, 28, 16);
          // '(6x, 1pe16.9, (t23, 2(6x, 1pe16.9)))' (:1985-1989)
          const std::string d = L.next();
          c.k = fnum(d, 6, 16);
          c.sc[0] = fnum(d, 28, 16);
          c.sc[1] = fnum(d, 50, 16);
          if (n_scale_variable(fnc) >= 3) c.sc[2] = fnum(L.next(), 28, 16);
        }
    } else if (type == 17) {
      L.next(); L.next();
      const std::string h = L.next();
      if (h.find('=') == std::string::npos) { *err = "bad block 17"; return -1; }
      const int n = after_eq_int(h);
      L.next();
      for (int q = 0; q < n; q++) {
        std::vector<double> x = numbers(L.next());
        if (x.size() < 3) { *err = "bad sub-band line in block 17"; return -1; }
        sub_band.push_back({(int)x[1], (int)x[2]});
      }
    } else if (type == 12) {
      // ice crystals, laid out as the droplet block (read_spectrum.F90:
      // 1175-1262); a fit longer than the record keeps is read past and the
      // type left absent (a call that names it is refused)
      if (ver != 2) { *err = "ice block version not supported"; return -3; }
      L.next();
      const std::string t = L.next();
      const size_t eq = t.find('=');
      if (eq == std::string::npos) { *err = "bad ice block"; return -1; }
      const int iice = std::atoi(t.c_str() + eq + 1);
      const std::string h = L.next();
      int scheme = inum(h, 34, 5), npar = inum(h, 66, 5);
      L.next();
      const std::string r = L.next();
      const bool keep = iice >= 1 && iice <= SG_NICE && npar <= SG_NDPARM;
      SgDrop *d = keep ? &s->ice[iice - 1] : nullptr;
      if (d) {
        d->present = 1; d->scheme = scheme; d->n_parm = npar;
        d->dim_min = fnum(r, 39, 12); d->dim_max = fnum(r, 55, 12);
      }
      for (int b = 0; b < s->n_band; b++) {
        L.next();
        int got = 0;
        while (got < npar) {
          std::vector<double> x = numbers(L.next());
          if (x.empty()) { *err = "truncated ice block"; return -1; }
          for (double v : x) {
            if (got < npar && d) d->parm[b][got] = v;
            got++;
          }
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
    // other blocks (1, 11, 12, 14, 17, ...) carry nothing the flux path
    // reads; they are skipped to their *END by the scan above
  }
  if (!have5) { *err = "no k-distribution block (5)"; return -1; }
  // k-term look-up tables from the '_k' file: the block headed 'k-table',
  // two header records, the P/T grid, then per band and gas in block 5's
  // order, two header records and per term and pressure n_tmp values
  if (lookup) {
    const std::string kp = std::string(path) + "_k";
    Lines K = read_lines(kp.c_str(), &ok);
    if (!ok) { *err = "k-term look-up tables: cannot open " + kp; return -1; }
    bool found = false;
    while (K.more()) {
      const std::string &h = K.next();
      if (h.compare(0, 6, "*BLOCK") == 0 && h.compare(8, 7, "k-table") == 0) { found = true; break; }
    }
    if (!found) { *err = "no k-table in " + kp; return -1; }
    K.next(); K.next();
    std::vector<double> row(s->n_tmp + 1);
    for (int ip = 0; ip < s->n_pre; ip++) {
      if (!e13_values(K, s->n_tmp + 1, row.data())) { *err = "truncated P/T grid in " + kp; return -1; }
      s->lut_lnp[ip] = std::log(row[0]);
      for (int it = 0; it < s->n_tmp; it++) s->lut_t[ip][it] = row[it + 1];
    }
    const size_t per = (size_t)s->n_pre * s->n_tmp;
    for (int b = 0; b < s->n_band; b++)
      for (int gas : order[b]) {
        Raw &r = raw[b][gas];
        if (r.fnc != SG_FNC_LOOKUP) continue;
        K.next(); K.next();
        r.lut0 = (int)(tab->data.size() / per);
        const size_t at = tab->data.size();
        tab->data.resize(at + per * r.nk);
        for (int k = 0; k < r.nk; k++)
          for (int ip = 0; ip < s->n_pre; ip++)
            if (!e13_values(K, s->n_tmp, tab->data.data() + at + k * per + (size_t)ip * s->n_tmp)) {
              *err = "truncated k-term look-up table in " + kp; return -1;
            }
      }
    s->has_lut = 1;
  }
  if (retain) {
    // gas selection (compress_spectrum, socrates_set_spectrum.F90:185-303):
    // a band keeps the retained absorbers of its list in order, and its
    // first absorber also where its first sub-band (scanned on from the
    // previous band's, as the reference does) names a k-term
    size_t i_sub = 0;
    for (int b = 0; b < s->n_band; b++) {
      bool keep_major = false;
      for (;;) {
        i_sub++;
        if (i_sub > sub_band.size()) break;
        if (sub_band[i_sub - 1][0] == b + 1) {
          keep_major = sub_band[i_sub - 1][1] > 0;
          break;
        }
      }
      std::vector<int> kept;
      for (int j = 0; j < n_band_absorb[b]; j++) {
        const int gi = index_absorb[b][j] - 1;
        const bool listed = gi >= 0 && gi < (int)type_absorb.size() &&
                            std::find(retain->begin(), retain->end(), type_absorb[gi]) != retain->end();
        if (listed || (j == 0 && keep_major)) kept.push_back(index_absorb[b][j]);
      }
      index_absorb[b] = kept;
      n_band_absorb[b] = (int)kept.size();
    }
  }
  // assemble the term tables: major gas is the first absorber listed for the
  // band (radiance_calc.F90: index_abs(1) is the major absorber)
  int ng = 0, nm = 0;
  bool wenyi = false;
  for (int b = 0; b < s->n_band; b++) {
    SgBand &B = s->band[b];
    if (n_band_absorb[b] < 1) { *err = "a band without gaseous absorption is not on this path"; return -3; }
    for (int j = 0; j < n_band_absorb[b]; j++) {
      int gi = index_absorb[b][j] - 1;
      if (gi < 0 || gi >= (int)type_absorb.size()) { *err = "bad absorber index in block 4"; return -1; }
      const Raw &r = raw[b][gi];
      int slot = gas_slot(type_absorb[gi]);
      if (j == 0) {
        if (slot >= SG_GAS_WM0) { *err = "a band's major gas other than H2O, CO2 or O3 is not on this path"; return -3; }
        B.n_major = r.nk; B.g0 = ng;
      } else {
        if (j - 1 >= SG_MAXMINOR_GAS) { *err = "too many minor gases in a band"; return -3; }
        if (j == 1) B.minor0 = nm;
        B.minor_nk[j - 1] = r.nk;
        B.n_minor_gas = j;
      }
      for (int k = 0; k < r.nk; k++) {
        SgTerm t;
        std::memset(&t, 0, sizeof t);
        t.k = r.k[k]; t.w = r.w[k]; t.a = r.s1[k]; t.b = r.s2[k]; t.c = r.s3[k];
        t.fnc = r.fnc;
        t.iscat = r.iscat[k];
        if (r.fnc != SG_FNC_LOOKUP) {
          t.lnpref = std::log(r.pref); t.lntref = std::log(r.tref);
          t.tinv = 1.0 / r.tref;
        }
        if (r.fnc == SG_FNC_LOOKUP) t.lut = r.lut0 + k;
        if (r.fnc == SG_FNC_WENYI) { t.lut = k; wenyi = true; }
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
    // continua in the band's order; radiance_calc.F90:910-922 passes the
    // data stored under the continuum's position as its index number, which
    // is the continuum's own data when the band lists them as 1, 2
    if (!cont_index.empty()) {
      B.n_cont = (int)cont_index[b].size();
      for (int j = 0; j < B.n_cont; j++) {
        if (cont_index[b][j] != j + 1) {
          *err = "continua listed out of index order in a band are not on this path"; return -3;
        }
        if (cont.empty() || !cont[b][j + 1].set) { *err = "continuum without block 9 data"; return -1; }
        const Cont &c = cont[b][j + 1];
        B.cont_type[j] = cont_index[b][j];
        B.cont_fnc[j] = c.fnc;
        B.cont_k[j] = c.k; B.cont_pref[j] = c.pref; B.cont_tref[j] = c.tref;
        for (int q = 0; q < 3; q++) B.cont_s[j][q] = c.sc[q];
        s->has_cont = 1;
      }
    }
  }
  if (wenyi) {
    // the Wenyi tables: gk4 (10 x 26 x 14), gk250b, plg, ttb
    tab->wenyi_off = (long)tab->data.size();
    const double *g4 = &kWenyiGk4[0][0][0];
    tab->data.insert(tab->data.end(), g4, g4 + 10 * 26 * 14);
    tab->data.insert(tab->data.end(), kWenyiGk250b, kWenyiGk250b + 14);
    tab->data.insert(tab->data.end(), kWenyiPlg, kWenyiPlg + 26);
    tab->data.insert(tab->data.end(), kWenyiTtb, kWenyiTtb + 10);
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
