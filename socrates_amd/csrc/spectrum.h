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

int read_spectrum(const char *path, SgSpec *s, std::string *err);
int read_mcica(const char *path, McicaData *m, std::string *err);
int attach_mcica(SgSpec *s, const McicaData &m, bool solar, std::string *err);

}  // namespace sg
