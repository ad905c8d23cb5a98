// The smoothed-aggregation multigrid preconditioner: its host-built hierarchy (multigrid.cpp) and
// the device V-cycle and conjugate gradient that run under it (multigrid.hip).
#pragma once

#include <cstdint>
#include <vector>

namespace hg {

struct HostCsr {
  int64_t rows = 0, cols = 0;
  std::vector<int32_t> ptr, idx;
  std::vector<double> val;
};

struct HostLevel {
  HostCsr a;                    // the level's operator
  HostCsr p, r;                 // prolongator to this level from the next, and its transpose
  std::vector<double> smoother; // omega / sum_j |A_ij|: the damped l1-Jacobi scaling
};

struct HostHierarchy {
  bool usable = false;  // false: the operator does not coarsen to a dense-invertible bottom
  std::vector<HostLevel> levels;       // the last one is the bottom (no p, r)
  int64_t coarse_n = 0;
  std::vector<double> coarse_inverse;  // coarse_n^2, row-major: a symmetric generalized inverse
};

// The hierarchy of a symmetric positive semi-definite operator (a zero row stays zero throughout).
HostHierarchy build_hierarchy(const HostCsr& fine, uint64_t seed);

}  // namespace hg
