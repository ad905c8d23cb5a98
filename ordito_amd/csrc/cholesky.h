// Device-side Cholesky plan, factors and solve workspace (cholesky.hip, solve.hip).
#pragma once

#include "kernels.h"

namespace hg {

// The analysis' maps on the device (uploaded once per mesh, shared by both systems).
struct DevicePlan {
  int64_t n = 0;
  int64_t nnz = 0;
  int32_t n_super = 0;
  int32_t n_levels = 0;
  std::vector<int32_t> level_ptr;  // host copies: the launches are per level
  std::vector<int32_t> level_gathers;  // 1 where some front of the level has children
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
  // Sweep tasks, grouped by level. A tile is 64 rows (forward) or 64 solve-block columns
  // (backward) of one front; a product task is a tile and a range of the inner dimension. Where a
  // level has too few tiles to fill the device its products are split along the inner dimension
  // into partial slots that a reduction adds in slot order.
  struct SweepTasks {
    std::vector<int32_t> level_ptr;   // product tasks per level
    std::vector<int32_t> reduce_ptr;  // reduction groups per level (empty level: no split)
    std::vector<int32_t> slots;       // partial slots each level uses
    DeviceArray<int32_t> s, tile, k0, k1, slot;  // per product task; slot -1 writes directly
    // The same per task with its front's figures, three int4 a task read in one round trip:
    // {s, tile, k0, k1}, {slot, ncol, col_begin, ncol + nrow}, {w_off (two words), vec_off
    // (forward) or row_ptr (backward) as two words}.
    DeviceArray<int4> desc;
    // The reduction groups likewise: {s, tile, first slot, slot count}, {0, ncol, col_begin, m},
    // {0, 0, vec_off (forward) as two words}.
    DeviceArray<int4> rdesc;
    DeviceArray<int32_t> rs, rtile, rfirst, rcount;  // per reduction group
  };
  SweepTasks fwd, bwd;
  int32_t max_slots = 0;
  int32_t split_target = kDefaultSplitTarget;
  // Forward gather lists: each front position's children contributions (absolute rows of the
  // front-vector arena), in ascending child order.
  DeviceArray<int32_t> gather_ptr, gather_src, gather_pair;
  // The operator's pattern in the elimination order (P A P^T) and each entry's original slot.
  DeviceArray<int32_t> p_offsets, p_columns, p_src;
  // The numeric factorization's tile launches (host list) and tasks (device, 4 int32 each).
  std::vector<FactorLaunch> factor_launches;
  DeviceArray<int32_t> factor_tasks;

  void upload(const Plan& plan, const HostPattern& pattern, int32_t target_tasks,
              hipStream_t stream);
  // Rebuilds the sweep tasks for another split target (the factors are unaffected).
  void upload_sweeps(const Plan& plan, int32_t target_tasks, hipStream_t stream);
  int64_t bytes() const;
};

// One system's inverse factors Z = [L11^-1; -L21 L11^-1] per supernode, FP64 and optionally FP32,
// and the operator's values in the elimination order for the refinement residual.
struct Factor {
  DeviceArray<double> w64;
  DeviceArray<float> w32;
  DeviceArray<double> pvalues;
  bool held() const { return w64.ptr != nullptr; }
  int64_t bytes() const { return w64.bytes() + w32.bytes() + pvalues.bytes(); }
  void reset() {
    w64.reset();
    w32.reset();
    pvalues.reset();
  }
};

struct SolveWorkspace {
  DeviceArray<double> bp, xa, xb, r, y, d, fvec, bmax, partials;
  DeviceArray<int32_t> pending;
  int64_t capacity = 0, fcap = 0, pcap = 0;
  void reserve(const DevicePlan& plan, int64_t kk);
  int64_t bytes() const {
    return bp.bytes() + xa.bytes() + xb.bytes() + r.bytes() + y.bytes() + d.bytes() +
           fvec.bytes() + bmax.bytes() + partials.bytes() + pending.bytes();
  }
  void release() {
    for (DeviceArray<double>* a : {&bp, &xa, &xb, &r, &y, &d, &fvec, &bmax, &partials}) a->reset();
    pending.reset();
    capacity = fcap = pcap = 0;
  }
};

// Numeric factorization of `values` (the operator over the mesh pattern) with the rows flagged in
// `pinned_new` (elimination order; may be null) replaced by the identity's. Throws HG_ERR_NOT_PD.
void factor_numeric(const DevicePlan& plan, const double* values, const int32_t* pinned_new,
                    const std::vector<int64_t>& pin_slots, Factor& factor, bool keep_fp32,
                    hipStream_t stream);

// x = A^-1 b (k columns of n entries in b_layout / x_layout) refined against the operator (the
// factor's permuted values) until every row passes the test (componentwise or normwise) or
// max_rounds sweeps ran; returns the most sweeps any 64-column chunk took. Pinned rows are
// dropped. b stored by columns is only read; b stored by rows is the solve's own right-hand side,
// its pinned rows zeroed in place.
// speculate (host memory from hipHostMalloc, one int a 64-column chunk; FP64 solve blocks only):
// run the first round alone and return without waiting for its test, whose verdict (nonzero: a
// row still above tolerance) lands in speculate[chunk] with the stream's next synchronization. The
// caller checks it there and solves again without speculating when any is set.
int32_t refined_solve(const DevicePlan& plan, const Factor& factor, bool use_fp32,
                      double* b, double* x, int64_t k, double tolerance, bool componentwise,
                      int32_t max_rounds, const int32_t* pinned_new, SolveWorkspace& ws,
                      hipStream_t stream, PhaseTimer* timer = nullptr,
                      int sweep_phase = kPhaseHeatSweep, FieldLayout b_layout = FieldLayout::kColumns,
                      FieldLayout x_layout = FieldLayout::kColumns, int32_t* speculate = nullptr);

// The components' members cut into 256-member chunks, for the null-space projection's reductions.
struct NullSpace {
  int32_t n_comp = 0, n_chunks = 0;
  DeviceArray<int32_t> chunk_comp, chunk_first, comp_chunk_ptr;
  int64_t bytes() const { return chunk_comp.bytes() + chunk_first.bytes() + comp_chunk_ptr.bytes(); }
  void build(const std::vector<int32_t>& comp_ptr_host, hipStream_t stream);
};

// Each component's constant set to zero mean, per column (fixed-order chunk sums).
void remove_null_space(const NullSpace& ns, const int32_t* comp_ptr, const int32_t* members,
                       int64_t n, int64_t k, double* x, DeviceArray<double>& scratch,
                       hipStream_t stream);

// The distance from Poisson solutions stored by rows (k columns, FieldLayout::kRows): each
// component's mean removed, then each column shifted to the mean over its sources (sources
// [source_offsets[q], source_offsets[q + 1]) in the caller's numbering) and oriented, written in
// the caller's layout. Every component's rows must be contiguous in the elimination order, in
// component order (comp_ptr), which the analysis guarantees; labels: each vertex's component.
// out_slots (device): chunk c of 64 columns goes to out_slots[c], read when its kernel runs, so a
// replayed graph writes wherever the slots point that call.
// remove_means false (a nonsingular Poisson system): the components' means are left in.
void finish_rows(const DevicePlan& plan, const NullSpace& ns, const int32_t* comp_ptr,
                 const int32_t* labels, const double* x, int64_t k, const int32_t* sources,
                 const int64_t* source_offsets, double* const* out_slots,
                 DeviceArray<double>& scratch, hipStream_t stream, bool remove_means = true);

}  // namespace hg
