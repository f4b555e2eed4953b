// CPU oracle -- TEST INFRASTRUCTURE ONLY (see oracle_core.hpp).
// Essential-matrix estimator restatement (filled in below).
#pragma once
#include "oracle_core.hpp"
#include "oracle_gcransac.hpp"

namespace orc {
inline int find_essential(const double*, size_t, const double*, const double*, int, int, int, int, double, double,
                          double, size_t, size_t, int, double, size_t, uint64_t, Model&, std::vector<size_t>&,
                          Statistics&, double&) {
  return -2;
}
}  // namespace orc
