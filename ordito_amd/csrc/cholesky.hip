// Multifrontal sparse Cholesky on the device: numeric factorization level by level and the
// partitioned-inverse triangular solves, with FP64 refinement against the operator.
//
// Follows ordito's design (cholesky.py:1-35, kernels/cholesky.py:1-31): each supernode's front is
// a dense matrix assembled from the operator's entries and its children's Schur complements
// (extend-add), partially factored, and its solve block [L11^-1; L21 L11^-1] stored so both
// triangular sweeps are level-scheduled matrix products (the partitioned inverse of Alvarado,
// Pothen and Schreiber 1993). The forward sweep here is multifrontal: each front gathers its
// children's update vectors through the same relative index maps the factorization's extend-add
// uses, so no sum is atomic and every sum has one fixed order.
#include "cholesky.h"

namespace hg {
namespace {

constexpr int kFrontThreads = 256;

// ---- numeric factorization ---------------------------------------------------------------------

__global__ void scatter_entries_kernel(const double* __restrict__ values,
                                       const int32_t* __restrict__ src,
                                       const int64_t* __restrict__ dst,
                                       const int32_t* __restrict__ erow,
                                       const int32_t* __restrict__ ecol,
                                       const int32_t* __restrict__ pinned, int64_t n_entries,
                                       double* __restrict__ arena) {
  int64_t e = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (e >= n_entries) return;
  int32_t r = erow[e], c = ecol[e];
  double value = values[src[e]];
  if (pinned && (pinned[r] || pinned[c])) value = (r == c) ? 1.0 : 0.0;
  arena[dst[e]] = value;
}

// A pinned row whose operator stores no diagonal (an empty row) still gets the identity's.
__global__ void pin_diagonal_kernel(const int32_t* __restrict__ pin_front_slot, int32_t n_pins,
                                    double* __restrict__ arena) {
  int32_t k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k < n_pins) arena[pin_front_slot[k]] = 1.0;
}

// Children's Schur complements into this level's fronts, children in ascending order.
__global__ void extend_add_kernel(const int32_t* __restrict__ nodes, const int32_t* __restrict__ ncol,
                                  const int32_t* __restrict__ nrow,
                                  const int32_t* __restrict__ child_ptr,
                                  const int32_t* __restrict__ children,
                                  const int32_t* __restrict__ row_ptr,
                                  const int32_t* __restrict__ rel,
                                  const int64_t* __restrict__ front_off,
                                  double* __restrict__ arena) {
  const int32_t s = nodes[blockIdx.x];
  const int64_t mp = ncol[s] + nrow[s];
  double* parent = arena + front_off[s];
  for (int32_t k = child_ptr[s]; k < child_ptr[s + 1]; ++k) {
    const int32_t c = children[k];
    const int32_t nc = ncol[c], rc = nrow[c];
    const int64_t mc = nc + rc;
    const double* child = arena + front_off[c];
    const int32_t* map = rel + row_ptr[c];
    for (int32_t j = 0; j < rc; ++j) {
      const int64_t pj = map[j];
      const double* col = child + (nc + j) * mc + nc;
      for (int32_t i = j + threadIdx.x; i < rc; i += blockDim.x)
        parent[pj * mp + map[i]] += col[i];
    }
    __syncthreads();
  }
}

// Right-looking dense partial Cholesky of each front's first ncol columns, in place
// (column-major), leaving the Schur complement in the trailing block. A non-positive pivot raises
// status[0] (cholesky.py:260-266).
__global__ void factor_fronts_kernel(const int32_t* __restrict__ nodes,
                                     const int32_t* __restrict__ ncol,
                                     const int32_t* __restrict__ nrow,
                                     const int64_t* __restrict__ front_off,
                                     double* __restrict__ arena, int32_t* __restrict__ status) {
  __shared__ double pivot;
  const int32_t s = nodes[blockIdx.x];
  const int32_t nc = ncol[s];
  const int64_t m = nc + nrow[s];
  double* F = arena + front_off[s];
  for (int32_t j = 0; j < nc; ++j) {
    if (threadIdx.x == 0) {
      double d = F[j * m + j];
      if (!(d > 0.0)) {
        atomicOr(status, 1);
        d = 1.0;
      }
      d = sqrt(d);
      F[j * m + j] = d;
      pivot = d;
    }
    __syncthreads();
    const double d = pivot;
    double* colj = F + j * m;
    for (int64_t i = j + 1 + threadIdx.x; i < m; i += blockDim.x) colj[i] /= d;
    __syncthreads();
    for (int64_t k = j + 1 + threadIdx.x; k < m; k += blockDim.x) {
      const double lkj = colj[k];
      double* colk = F + k * m;
      for (int64_t i = k; i < m; ++i) colk[i] -= colj[i] * lkj;
    }
    __syncthreads();
  }
}

// Solve block of each front: W11 = L11^-1 by forward substitution a column a thread, then
// W21 = L21 W11; stored column-major (leading dimension m) at w_off.
template <typename T>
__global__ void solve_blocks_kernel(const int32_t* __restrict__ nodes,
                                    const int32_t* __restrict__ ncol,
                                    const int32_t* __restrict__ nrow,
                                    const int64_t* __restrict__ front_off,
                                    const int64_t* __restrict__ w_off,
                                    const double* __restrict__ arena, double* __restrict__ w) {
  const int32_t s = nodes[blockIdx.x];
  const int32_t nc = ncol[s], nr = nrow[s];
  const int64_t m = nc + nr;
  const double* L = arena + front_off[s];
  double* W = w + w_off[s];
  for (int32_t j = threadIdx.x; j < nc; j += blockDim.x) {
    double* wj = W + j * m;
    for (int32_t i = 0; i < j; ++i) wj[i] = 0.0;
    wj[j] = 1.0 / L[j * m + j];
    for (int32_t i = j + 1; i < nc; ++i) {
      double sum = 0.0;
      for (int32_t k = j; k < i; ++k) sum += L[k * m + i] * wj[k];
      wj[i] = -sum / L[i * m + i];
    }
  }
  __syncthreads();
  const int64_t total = static_cast<int64_t>(nr) * nc;
  for (int64_t t = threadIdx.x; t < total; t += blockDim.x) {
    const int32_t i = static_cast<int32_t>(t % nr);
    const int32_t j = static_cast<int32_t>(t / nr);
    const double* wj = W + j * m;
    double sum = 0.0;
    for (int32_t k = j; k < nc; ++k) sum += L[k * m + nc + i] * wj[k];
    W[j * m + nc + i] = sum;
  }
}

__global__ void narrow_kernel(const double* __restrict__ src, float* __restrict__ dst, int64_t n) {
  int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i < n) dst[i] = static_cast<float>(src[i]);
}

// The operator's values in the elimination order's CSR (P A P^T), for the refinement residuals.
__global__ void permute_values_kernel(const double* __restrict__ values,
                                      const int32_t* __restrict__ src, int64_t nnz,
                                      double* __restrict__ out) {
  int64_t e = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (e < nnz) out[e] = values[src[e]];
}

inline unsigned blocks_for(int64_t n) { return static_cast<unsigned>((n + 255) / 256); }

}  // namespace

// ---- host side --------------------------------------------------------------------------------

// Product tasks a level is split into at least, where its tiles allow: a few per compute unit.
constexpr int32_t kTargetTasks = 1024;

HostSweep sweep_tasks(const Plan& plan, bool forward) {
  HostSweep out;
  out.level_ptr.assign(plan.n_levels + 1, 0);
  out.reduce_ptr.assign(plan.n_levels + 1, 0);
  out.slots.assign(plan.n_levels, 0);
  for (int32_t level = 0; level < plan.n_levels; ++level) {
    int64_t tiles = 0;
    for (int32_t k = plan.level_ptr[level]; k < plan.level_ptr[level + 1]; ++k) {
      const int32_t s = plan.level_nodes[k];
      const int32_t extent = forward ? plan.ncol[s] + plan.nrow[s] : plan.ncol[s];
      tiles += (extent + 63) / 64;
    }
    const int64_t per = std::max<int64_t>(tiles, 1);
    const int32_t split = static_cast<int32_t>(std::max<int64_t>(1, (kTargetTasks + per - 1) / per));
    int32_t slots = 0;
    for (int32_t k = plan.level_ptr[level]; k < plan.level_ptr[level + 1]; ++k) {
      const int32_t s = plan.level_nodes[k];
      const int32_t nc = plan.ncol[s], m = nc + plan.nrow[s];
      const int32_t extent = forward ? m : nc;
      for (int32_t tile = 0; tile < extent; tile += 64) {
        // Forward rows below nc need j <= i (W11 is lower triangular); backward columns j need
        // i >= j.
        const int32_t begin = forward ? 0 : tile;
        const int32_t end = forward ? ((tile + 64 <= nc) ? tile + 64 : nc) : m;
        const int32_t inner = end - begin;
        const int32_t chunks = std::max(1, std::min(split, (inner + 63) / 64));
        if (chunks == 1) {
          out.s.push_back(s);
          out.tile.push_back(tile);
          out.k0.push_back(begin);
          out.k1.push_back(end);
          out.slot.push_back(-1);
          continue;
        }
        const int32_t width = ((inner + chunks - 1) / chunks + 31) / 32 * 32;
        const int32_t first = slots;
        for (int32_t k0 = begin; k0 < end; k0 += width) {
          out.s.push_back(s);
          out.tile.push_back(tile);
          out.k0.push_back(k0);
          out.k1.push_back(std::min(end, k0 + width));
          out.slot.push_back(slots++);
        }
        out.rs.push_back(s);
        out.rtile.push_back(tile);
        out.rfirst.push_back(first);
        out.rcount.push_back(slots - first);
      }
    }
    out.level_ptr[level + 1] = static_cast<int32_t>(out.s.size());
    out.reduce_ptr[level + 1] = static_cast<int32_t>(out.rs.size());
    out.slots[level] = slots;
    out.max_slots = std::max(out.max_slots, slots);
  }
  return out;
}

HostGather gather_lists(const Plan& plan) {
  HostGather g;
  g.ptr.assign(plan.vec_size + 1, 0);
  for (int32_t c = 0; c < plan.n_super; ++c) {
    const int32_t p = plan.parent[c];
    if (p < 0) continue;
    for (int32_t i = 0; i < plan.nrow[c]; ++i) g.ptr[plan.vec_off[p] + plan.rel[plan.row_ptr[c] + i] + 1]++;
  }
  for (int64_t v = 0; v < plan.vec_size; ++v) g.ptr[v + 1] += g.ptr[v];
  g.src.resize(g.ptr[plan.vec_size]);
  std::vector<int32_t> at(g.ptr.begin(), g.ptr.end() - 1);
  // Children are numbered before their parents and visited in ascending order, so each position's
  // list comes out in ascending child order.
  for (int32_t c = 0; c < plan.n_super; ++c) {
    const int32_t p = plan.parent[c];
    if (p < 0) continue;
    for (int32_t i = 0; i < plan.nrow[c]; ++i) {
      const int64_t target = plan.vec_off[p] + plan.rel[plan.row_ptr[c] + i];
      g.src[at[target]++] = static_cast<int32_t>(plan.vec_off[c] + plan.ncol[c] + i);
    }
  }
  g.task_ptr.assign(plan.n_levels + 1, 0);
  for (int32_t level = 0; level < plan.n_levels; ++level) {
    for (int32_t k = plan.level_ptr[level]; k < plan.level_ptr[level + 1]; ++k) {
      const int32_t s = plan.level_nodes[k];
      for (int32_t p0 = 0; p0 < plan.ncol[s] + plan.nrow[s]; p0 += 64) {
        g.task_s.push_back(s);
        g.task_p0.push_back(p0);
      }
    }
    g.task_ptr[level + 1] = static_cast<int32_t>(g.task_s.size());
  }
  return g;
}

void validate_sweeps(const Plan& plan) {
  auto fail = [](const std::string& what) { throw Error{HG_ERR_RUNTIME, "plan check: " + what}; };
  const int64_t S = plan.n_super;
  for (int dir = 0; dir < 2; ++dir) {
    const bool forward = dir == 0;
    const HostSweep h = sweep_tasks(plan, forward);
    std::vector<int64_t> covered(S, 0);
    for (size_t t = 0; t < h.s.size(); ++t) {
      const int32_t s = h.s[t];
      if (s < 0 || s >= S) fail("task front out of range");
      const int32_t nc = plan.ncol[s], m = nc + plan.nrow[s];
      const int32_t extent = forward ? m : nc;
      if (h.tile[t] < 0 || h.tile[t] >= extent || h.tile[t] % 64) fail("task tile out of range");
      if (h.k0[t] < 0 || h.k1[t] <= h.k0[t]) fail("empty or negative inner range");
      if (forward && h.k1[t] > nc) fail("forward inner range past the front's columns");
      if (!forward && (h.k1[t] > m || h.k0[t] < h.tile[t])) fail("backward inner range out of range");
      if (h.slot[t] >= 0) {
        int32_t level = plan.level[s];
        if (h.slot[t] >= h.slots[level]) fail("slot past the level's partials");
      }
      covered[s] += h.k1[t] - h.k0[t];
    }
    // Each tile's inner range is covered exactly once.
    for (int64_t s = 0; s < S; ++s) {
      const int64_t nc = plan.ncol[s], m = nc + plan.nrow[s];
      int64_t expect = 0;
      const int64_t extent = forward ? m : nc;
      for (int64_t tile = 0; tile < extent; tile += 64)
        expect += forward ? ((tile + 64 <= nc) ? tile + 64 : nc) : (m - tile);
      if (covered[s] != expect) fail("inner ranges do not tile the products");
    }
    for (size_t g = 0; g < h.rs.size(); ++g) {
      const int32_t s = h.rs[g];
      if (s < 0 || s >= S) fail("reduction front out of range");
      if (h.rcount[g] < 2 || h.rfirst[g] < 0 || h.rfirst[g] + h.rcount[g] > h.slots[plan.level[s]])
        fail("reduction slots out of range");
    }
  }
  const HostGather g = gather_lists(plan);
  if (static_cast<int64_t>(g.ptr.size()) != plan.vec_size + 1) fail("gather offsets size");
  for (int32_t v : g.src)
    if (v < 0 || v >= plan.vec_size) fail("gather source out of the front-vector arena");
  for (size_t t = 0; t < g.task_s.size(); ++t) {
    const int32_t s = g.task_s[t];
    if (s < 0 || s >= S || g.task_p0[t] >= plan.ncol[s] + plan.nrow[s]) fail("gather task range");
  }
  // Every structure row of every child lands in exactly one parent position.
  int64_t rows_with_parent = 0;
  for (int64_t s = 0; s < S; ++s)
    if (plan.parent[s] >= 0) rows_with_parent += plan.nrow[s];
  if (static_cast<int64_t>(g.src.size()) != rows_with_parent) fail("gather list count");
}

namespace {

void upload_sweep(const HostSweep& h, DevicePlan::SweepTasks& out, hipStream_t stream) {
  out.level_ptr = h.level_ptr;
  out.reduce_ptr = h.reduce_ptr;
  out.slots = h.slots;
  out.s.upload(h.s, stream);
  out.tile.upload(h.tile, stream);
  out.k0.upload(h.k0, stream);
  out.k1.upload(h.k1, stream);
  out.slot.upload(h.slot, stream);
  out.rs.upload(h.rs, stream);
  out.rtile.upload(h.rtile, stream);
  out.rfirst.upload(h.rfirst, stream);
  out.rcount.upload(h.rcount, stream);
}

}  // namespace

void DevicePlan::upload(const Plan& plan, const HostPattern& pattern, hipStream_t stream) {
  const HostSweep hf = sweep_tasks(plan, true);
  const HostSweep hb = sweep_tasks(plan, false);
  max_slots = std::max(hf.max_slots, hb.max_slots);
  upload_sweep(hf, fwd, stream);
  upload_sweep(hb, bwd, stream);
  {
    const HostGather g = gather_lists(plan);
    gather_ptr.upload(g.ptr, stream);
    gather_src.upload(g.src, stream);
    gather_task_ptr = g.task_ptr;
    gather_s.upload(g.task_s, stream);
    gather_p0.upload(g.task_p0, stream);
  }
  // The operator's pattern in the elimination order: row i is old row perm[i], its columns
  // renumbered and sorted; src names each entry's slot in the original values.
  {
    const int64_t nn = plan.n;
    std::vector<int32_t> offs(nn + 1, 0), cols, src;
    cols.reserve(pattern.columns.size());
    src.reserve(pattern.columns.size());
    std::vector<std::pair<int32_t, int32_t>> row;
    for (int64_t i = 0; i < nn; ++i) {
      const int32_t old = plan.perm[i];
      row.clear();
      for (int32_t e = pattern.offsets[old]; e < pattern.offsets[old + 1]; ++e)
        row.push_back({plan.iperm[pattern.columns[e]], e});
      std::sort(row.begin(), row.end());
      for (const auto& entry : row) {
        cols.push_back(entry.first);
        src.push_back(entry.second);
      }
      offs[i + 1] = static_cast<int32_t>(cols.size());
    }
    nnz = static_cast<int64_t>(cols.size());
    p_offsets.upload(offs, stream);
    p_columns.upload(cols, stream);
    p_src.upload(src, stream);
  }
  n = plan.n;
  n_super = plan.n_super;
  n_levels = plan.n_levels;
  level_ptr = plan.level_ptr;
  arena_size = plan.arena_size;
  w_size = plan.w_size;
  vec_size = plan.vec_size;
  max_front = plan.max_front;
  nnz_w = plan.nnz_w;
  flops = plan.flops;
  n_entries = static_cast<int64_t>(plan.entry_src.size());
  perm.upload(plan.perm, stream);
  iperm.upload(plan.iperm, stream);
  col_begin.upload(plan.col_begin, stream);
  ncol.upload(plan.ncol, stream);
  nrow.upload(plan.nrow, stream);
  row_ptr.upload(plan.row_ptr, stream);
  rows.upload(plan.rows, stream);
  child_ptr.upload(plan.child_ptr, stream);
  children.upload(plan.children, stream);
  rel.upload(plan.rel, stream);
  level_nodes.upload(plan.level_nodes, stream);
  front_off.upload(plan.front_off, stream);
  w_off.upload(plan.w_off, stream);
  vec_off.upload(plan.vec_off, stream);
  entry_src.upload(plan.entry_src, stream);
  entry_dst.upload(plan.entry_dst, stream);
  entry_row.upload(plan.entry_row, stream);
  entry_col.upload(plan.entry_col, stream);
  host_col_begin = plan.col_begin;
  host_ncol = plan.ncol;
  host_front_off = plan.front_off;
  host_nrow = plan.nrow;
}

int64_t DevicePlan::bytes() const {
  return perm.bytes() + iperm.bytes() + col_begin.bytes() + ncol.bytes() + nrow.bytes() +
         row_ptr.bytes() + rows.bytes() + child_ptr.bytes() + children.bytes() + rel.bytes() +
         level_nodes.bytes() + front_off.bytes() + w_off.bytes() + vec_off.bytes() +
         entry_src.bytes() + entry_dst.bytes() + entry_row.bytes() + entry_col.bytes() +
         fwd.s.bytes() * 5 + fwd.rs.bytes() * 4 + bwd.s.bytes() * 5 + bwd.rs.bytes() * 4 +
         gather_ptr.bytes() + gather_src.bytes() + gather_s.bytes() + gather_p0.bytes() +
         p_offsets.bytes() + p_columns.bytes() + p_src.bytes();
}

void factor_numeric(const DevicePlan& plan, const double* values, const int32_t* pinned_new,
                    const std::vector<int64_t>& pin_slots, Factor& factor, bool keep_fp32,
                    hipStream_t stream) {
  DeviceArray<double> arena(plan.arena_size);
  HG_HIP(hipMemsetAsync(arena.ptr, 0, sizeof(double) * plan.arena_size, stream));
  DeviceArray<int32_t> status(1);
  HG_HIP(hipMemsetAsync(status.ptr, 0, sizeof(int32_t), stream));
  if (plan.n_entries > 0) {
    scatter_entries_kernel<<<blocks_for(plan.n_entries), 256, 0, stream>>>(
        values, plan.entry_src.ptr, plan.entry_dst.ptr, plan.entry_row.ptr, plan.entry_col.ptr,
        pinned_new, plan.n_entries, arena.ptr);
    HG_HIP(hipGetLastError());
  }
  if (!pin_slots.empty()) {
    std::vector<int32_t> slots32(pin_slots.begin(), pin_slots.end());
    DeviceArray<int32_t> d_slots;
    d_slots.upload(slots32, stream);
    pin_diagonal_kernel<<<blocks_for(static_cast<int64_t>(slots32.size())), 256, 0, stream>>>(
        d_slots.ptr, static_cast<int32_t>(slots32.size()), arena.ptr);
    HG_HIP(hipGetLastError());
    HG_HIP(hipStreamSynchronize(stream));
  }
  factor.w64.allocate(plan.w_size);
  for (int32_t level = 0; level < plan.n_levels; ++level) {
    const int32_t first = plan.level_ptr[level];
    const int32_t count = plan.level_ptr[level + 1] - first;
    if (count == 0) continue;
    const int32_t* nodes = plan.level_nodes.ptr + first;
    extend_add_kernel<<<count, kFrontThreads, 0, stream>>>(
        nodes, plan.ncol.ptr, plan.nrow.ptr, plan.child_ptr.ptr, plan.children.ptr,
        plan.row_ptr.ptr, plan.rel.ptr, plan.front_off.ptr, arena.ptr);
    factor_fronts_kernel<<<count, kFrontThreads, 0, stream>>>(nodes, plan.ncol.ptr, plan.nrow.ptr,
                                                              plan.front_off.ptr, arena.ptr,
                                                              status.ptr);
    solve_blocks_kernel<double><<<count, kFrontThreads, 0, stream>>>(
        nodes, plan.ncol.ptr, plan.nrow.ptr, plan.front_off.ptr, plan.w_off.ptr, arena.ptr,
        factor.w64.ptr);
    HG_HIP(hipGetLastError());
  }
  int32_t host_status = 0;
  HG_HIP(hipMemcpyAsync(&host_status, status.ptr, sizeof(int32_t), hipMemcpyDeviceToHost, stream));
  HG_HIP(hipStreamSynchronize(stream));
  if (host_status != 0) {
    factor.w64.reset();
    throw Error{HG_ERR_NOT_PD, "the operator met a pivot that is not positive"};
  }
  factor.pvalues.allocate(plan.nnz);
  if (plan.nnz > 0) {
    permute_values_kernel<<<blocks_for(plan.nnz), 256, 0, stream>>>(values, plan.p_src.ptr,
                                                                   plan.nnz, factor.pvalues.ptr);
    HG_HIP(hipGetLastError());
  }
  if (keep_fp32) {
    factor.w32.allocate(plan.w_size);
    if (plan.w_size > 0) {
      narrow_kernel<<<blocks_for(plan.w_size), 256, 0, stream>>>(factor.w64.ptr, factor.w32.ptr,
                                                                 plan.w_size);
      HG_HIP(hipGetLastError());
    }
  } else {
    factor.w32.reset();
  }
}

}  // namespace hg
