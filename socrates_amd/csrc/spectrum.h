// Host-side readers of spectral files and McICA data (see spectrum.cpp).
#pragma once

#include <array>
#include <string>
#include <vector>

#include "sg_types.h"

namespace sg {

struct McicaData {
  int n_gen = 0, n_req_single = 0, n_req_optimal = 0, ipph = 0, n1 = 0, n2 = 0;
  std::vector<int> reorder_single, reorder_optimal;
  std::vector<std::array<int, 3>> sw_subcol_k, lw_subcol_k;  // (band, k, n)
  std::vector<double> xcw;                                   // n1 x n2, column-major
};

// Variable-size tables of a spectral file, uploaded beside the SgSpec: the
// k-term look-up tables (term t.lut at t.lut * n_pre * n_tmp) and, from
// wenyi_off (or -1), the Wenyi CO2 tables.
struct SpecTables {
  std::vector<double> data;
  long wenyi_off = -1;
};

// retain: absorber identifiers to keep (set_spectrum's gas flags), or null
// for every absorber of the file.
int read_spectrum(const char *path, SgSpec *s, SpecTables *tab, std::string *err,
                  const std::vector<int> *retain = nullptr);
int read_mcica(const char *path, McicaData *m, std::string *err);
int attach_mcica(SgSpec *s, const McicaData &m, bool solar, std::string *err);

}  // namespace sg
