// Device-side Cholesky plan, factors and solve workspace (cholesky.hip).
#pragma once

#include "kernels.h"

namespace hg {

// The analysis' maps on the device (uploaded once per mesh, shared by both systems).
struct DevicePlan {
  int64_t n = 0;
  int32_t n_super = 0;
  int32_t n_levels = 0;
  std::vector<int32_t> level_ptr;  // host copy: the launches are per level
  std::vector<int32_t> host_col_begin, host_ncol, host_nrow;
  std::vector<int64_t> host_front_off;
  int64_t arena_size = 0, w_size = 0, vec_size = 0, n_entries = 0, nnz_w = 0;
  int32_t max_front = 0;
  double flops = 0.0;
  DeviceArray<int32_t> perm, iperm, col_begin, ncol, nrow, row_ptr, rows, child_ptr, children, rel,
      level_nodes;
  DeviceArray<int64_t> front_off, w_off, vec_off;
  DeviceArray<int32_t> entry_src, entry_row, entry_col;
  DeviceArray<int64_t> entry_dst;

  void upload(const Plan& plan, hipStream_t stream);
  int64_t bytes() const;
};

// One system's solve blocks [L11^-1; L21 L11^-1] per supernode, FP64 and optionally FP32.
struct Factor {
  DeviceArray<double> w64;
  DeviceArray<float> w32;
  bool held() const { return w64.ptr != nullptr; }
  int64_t bytes() const { return w64.bytes() + w32.bytes(); }
  void reset() {
    w64.reset();
    w32.reset();
  }
};

struct SolveWorkspace {
  DeviceArray<double> r, rp, y, xp, fvec, bmax;
  DeviceArray<int32_t> pending;
  int64_t capacity = 0, fcap = 0;
  void reserve(const DevicePlan& plan, int64_t k);
};

// Numeric factorization of `values` (the operator over the mesh pattern) with the rows flagged in
// `pinned_new` (elimination order; may be null) replaced by the identity's. Throws HG_ERR_NOT_PD.
void factor_numeric(const DevicePlan& plan, const double* values, const int32_t* pinned_new,
                    const std::vector<int64_t>& pin_slots, Factor& factor, bool keep_fp32,
                    hipStream_t stream);

// x = A^-1 b refined against the operator until every row passes (componentwise or normwise) or
// max_rounds solves ran; returns the solves. k columns of n entries each.
int32_t refined_solve(const DevicePlan& plan, const Factor& factor, bool use_fp32,
                      const int32_t* offsets, const int32_t* columns, const double* values,
                      const double* b, double* x, int64_t k, double tolerance, bool componentwise,
                      int32_t max_rounds, const int32_t* pinned_new, const int32_t* dropped_old,
                      SolveWorkspace& ws, hipStream_t stream, PhaseTimer* timer = nullptr,
                      int sweep_phase = kPhaseHeatSweep);

// Each listed component's constant set to zero mean, per column.
void remove_null_space(const int32_t* comp_ptr, const int32_t* members, int32_t n_comp, int64_t n,
                       int64_t k, double* x, hipStream_t stream);

}  // namespace hg
