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

// ---- solves -----------------------------------------------------------------------------------

// Forward sweep of one level: each front gathers [b(cols); 0] plus its children's update vectors,
// then y(cols) = W11 f(cols) and its own update vector f(rows) -= W21 f(cols).
template <typename T>
__global__ void forward_level_kernel(const int32_t* __restrict__ nodes,
                                     const int32_t* __restrict__ col_begin,
                                     const int32_t* __restrict__ ncol,
                                     const int32_t* __restrict__ nrow,
                                     const int32_t* __restrict__ child_ptr,
                                     const int32_t* __restrict__ children,
                                     const int32_t* __restrict__ row_ptr,
                                     const int32_t* __restrict__ rel,
                                     const int64_t* __restrict__ w_off,
                                     const int64_t* __restrict__ vec_off, const T* __restrict__ w,
                                     const double* __restrict__ b, double* __restrict__ y,
                                     double* __restrict__ fvec, int64_t n, int32_t k) {
  const int32_t s = nodes[blockIdx.x];
  const int32_t nc = ncol[s], nr = nrow[s], c0 = col_begin[s];
  const int64_t m = nc + nr;
  double* f = fvec + vec_off[s] * k;
  for (int64_t t = threadIdx.x; t < m * k; t += blockDim.x) {
    const int64_t q = t / m, i = t % m;
    f[t] = i < nc ? b[q * n + c0 + i] : 0.0;
  }
  __syncthreads();
  for (int32_t e = child_ptr[s]; e < child_ptr[s + 1]; ++e) {
    const int32_t c = children[e];
    const int64_t mc = ncol[c] + nrow[c];
    const double* fc = fvec + vec_off[c] * k + ncol[c];
    const int32_t* map = rel + row_ptr[c];
    const int32_t rc = nrow[c];
    for (int64_t t = threadIdx.x; t < static_cast<int64_t>(rc) * k; t += blockDim.x) {
      const int64_t q = t / rc, i = t % rc;
      f[q * m + map[i]] += fc[q * mc + i];
    }
    __syncthreads();
  }
  const T* W = w + w_off[s];
  for (int64_t t = threadIdx.x; t < m * k; t += blockDim.x) {
    const int64_t q = t / m, i = t % m;
    const double* fq = f + q * m;
    double sum = 0.0;
    const int32_t jmax = i < nc ? static_cast<int32_t>(i) + 1 : nc;
    for (int32_t j = 0; j < jmax; ++j) sum += static_cast<double>(W[j * m + i]) * fq[j];
    if (i < nc) y[q * n + c0 + i] = sum;
    else f[q * m + i] -= sum;
  }
}

// Backward sweep of one level: x(cols) = W11^T y(cols) - W21^T x(rows).
template <typename T>
__global__ void backward_level_kernel(const int32_t* __restrict__ nodes,
                                      const int32_t* __restrict__ col_begin,
                                      const int32_t* __restrict__ ncol,
                                      const int32_t* __restrict__ nrow,
                                      const int32_t* __restrict__ row_ptr,
                                      const int32_t* __restrict__ rows,
                                      const int64_t* __restrict__ w_off, const T* __restrict__ w,
                                      const double* __restrict__ y, double* __restrict__ x,
                                      int64_t n, int32_t k) {
  const int32_t s = nodes[blockIdx.x];
  const int32_t nc = ncol[s], nr = nrow[s], c0 = col_begin[s];
  const int64_t m = nc + nr;
  const T* W = w + w_off[s];
  const int32_t* r = rows + row_ptr[s];
  for (int64_t t = threadIdx.x; t < static_cast<int64_t>(nc) * k; t += blockDim.x) {
    const int64_t q = t / nc;
    const int32_t j = static_cast<int32_t>(t % nc);
    const T* wj = W + j * m;
    const double* yq = y + q * n + c0;
    const double* xq = x + q * n;
    double sum = 0.0;
    for (int32_t i = j; i < nc; ++i) sum += static_cast<double>(wj[i]) * yq[i];
    for (int32_t i = 0; i < nr; ++i) sum -= static_cast<double>(wj[nc + i]) * xq[r[i]];
    x[q * n + c0 + j] = sum;
  }
}

// rp[new] = r[perm[new]] (zero at pinned rows: their equations are dropped).
__global__ void gather_perm_kernel(const double* __restrict__ r, const int32_t* __restrict__ perm,
                                   const int32_t* __restrict__ pinned, int64_t n, int64_t k,
                                   double* __restrict__ rp) {
  int64_t t = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (t >= n * k) return;
  int64_t q = t / n, i = t % n;
  rp[t] = (pinned && pinned[i]) ? 0.0 : r[q * n + perm[i]];
}

// x[perm[new]] += xp[new].
__global__ void scatter_add_perm_kernel(const double* __restrict__ xp,
                                        const int32_t* __restrict__ perm, int64_t n, int64_t k,
                                        double* __restrict__ x) {
  int64_t t = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (t >= n * k) return;
  int64_t q = t / n, i = t % n;
  x[q * n + perm[i]] += xp[t];
}

// Per column, the largest |b| (the normwise test's scale, kernels/cholesky.py:1357-1390).
__global__ void column_absmax_kernel(const double* __restrict__ b, int64_t n,
                                     double* __restrict__ out) {
  __shared__ double scratch[kFrontThreads];
  const int64_t q = blockIdx.x;
  double v = 0.0;
  for (int64_t i = threadIdx.x; i < n; i += kFrontThreads) v = fmax(v, fabs(b[q * n + i]));
  scratch[threadIdx.x] = v;
  __syncthreads();
  for (int stride = kFrontThreads / 2; stride > 0; stride >>= 1) {
    if (threadIdx.x < stride) scratch[threadIdx.x] = fmax(scratch[threadIdx.x], scratch[threadIdx.x + stride]);
    __syncthreads();
  }
  if (threadIdx.x == 0) out[q] = scratch[0];
}

// r = b - A x and whether any row is still above tolerance: componentwise against
// (|A||x| + |b|)_i, normwise against the larger of that and the column's largest |b|
// (kernels/cholesky.py:1393-1440). Dropped (pinned) rows are not tested. "Not within" so NaN fails.
__global__ void residual_test_kernel(const int32_t* __restrict__ offsets,
                                     const int32_t* __restrict__ columns,
                                     const double* __restrict__ values, const double* __restrict__ x,
                                     const double* __restrict__ b, const int32_t* __restrict__ dropped,
                                     const double* __restrict__ bmax, int64_t n, int64_t k,
                                     double tolerance, int32_t componentwise,
                                     double* __restrict__ r, int32_t* __restrict__ pending) {
  int64_t t = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (t >= n * k) return;
  int64_t q = t / n, i = t % n;
  const double* xq = x + q * n;
  double bi = b[t];
  double total = bi, scale = fabs(bi);
  for (int32_t e = offsets[i]; e < offsets[i + 1]; ++e) {
    double term = values[e] * xq[columns[e]];
    total -= term;
    scale += fabs(term);
  }
  r[t] = total;
  if (!componentwise) scale = fmax(scale, bmax[q]);
  if (!(dropped && dropped[i]) && !(fabs(total) <= tolerance * scale)) atomicOr(pending, 1);
}

// Each singular component's constant set to zero mean: one block a (component, column).
__global__ void null_space_kernel(const int32_t* __restrict__ comp_ptr,
                                  const int32_t* __restrict__ members, int64_t n,
                                  double* __restrict__ x) {
  __shared__ double scratch[kFrontThreads];
  const int32_t c = blockIdx.x;
  const int64_t q = blockIdx.y;
  const int32_t first = comp_ptr[c], last = comp_ptr[c + 1];
  double sum = 0.0;
  for (int32_t e = first + threadIdx.x; e < last; e += kFrontThreads) sum += x[q * n + members[e]];
  scratch[threadIdx.x] = sum;
  __syncthreads();
  for (int stride = kFrontThreads / 2; stride > 0; stride >>= 1) {
    if (threadIdx.x < stride) scratch[threadIdx.x] += scratch[threadIdx.x + stride];
    __syncthreads();
  }
  const double mean = scratch[0] / static_cast<double>(last - first);
  for (int32_t e = first + threadIdx.x; e < last; e += kFrontThreads) x[q * n + members[e]] -= mean;
}

inline unsigned blocks_for(int64_t n) { return static_cast<unsigned>((n + 255) / 256); }

}  // namespace

// ---- host side --------------------------------------------------------------------------------

void DevicePlan::upload(const Plan& plan, hipStream_t stream) {
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
         entry_src.bytes() + entry_dst.bytes() + entry_row.bytes() + entry_col.bytes();
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

void SolveWorkspace::reserve(const DevicePlan& plan, int64_t k) {
  const int64_t entries = plan.n * k;
  if (capacity < entries) {
    r.allocate(entries);
    rp.allocate(entries);
    y.allocate(entries);
    xp.allocate(entries);
    capacity = entries;
  }
  if (fcap < plan.vec_size * k) {
    fvec.allocate(std::max<int64_t>(plan.vec_size * k, 1));
    fcap = plan.vec_size * k;
  }
  if (bmax.size < k) bmax.allocate(std::max<int64_t>(k, 64));
  if (pending.size < 1) pending.allocate(1);
}

template <typename T>
static void triangular_solve(const DevicePlan& plan, const T* w, const double* rp, double* y,
                             double* xp, double* fvec, int64_t k, hipStream_t stream) {
  const int64_t n = plan.n;
  for (int32_t level = 0; level < plan.n_levels; ++level) {
    const int32_t first = plan.level_ptr[level];
    const int32_t count = plan.level_ptr[level + 1] - first;
    if (count == 0) continue;
    forward_level_kernel<T><<<count, kFrontThreads, 0, stream>>>(
        plan.level_nodes.ptr + first, plan.col_begin.ptr, plan.ncol.ptr, plan.nrow.ptr,
        plan.child_ptr.ptr, plan.children.ptr, plan.row_ptr.ptr, plan.rel.ptr, plan.w_off.ptr,
        plan.vec_off.ptr, w, rp, y, fvec, n, static_cast<int32_t>(k));
  }
  for (int32_t level = plan.n_levels - 1; level >= 0; --level) {
    const int32_t first = plan.level_ptr[level];
    const int32_t count = plan.level_ptr[level + 1] - first;
    if (count == 0) continue;
    backward_level_kernel<T><<<count, kFrontThreads, 0, stream>>>(
        plan.level_nodes.ptr + first, plan.col_begin.ptr, plan.ncol.ptr, plan.nrow.ptr,
        plan.row_ptr.ptr, plan.rows.ptr, plan.w_off.ptr, w, y, xp, n, static_cast<int32_t>(k));
  }
  HG_HIP(hipGetLastError());
}

int32_t refined_solve(const DevicePlan& plan, const Factor& factor, bool use_fp32,
                      const int32_t* offsets, const int32_t* columns, const double* values,
                      const double* b, double* x, int64_t k, double tolerance,
                      bool componentwise, int32_t max_rounds, const int32_t* pinned_new,
                      const int32_t* dropped_old, SolveWorkspace& ws, hipStream_t stream,
                      PhaseTimer* timer, int sweep_phase) {
  const int64_t n = plan.n;
  if (n == 0 || k == 0) return 0;
  ws.reserve(plan, k);
  const unsigned grid = blocks_for(n * k);
  HG_HIP(hipMemsetAsync(x, 0, sizeof(double) * n * k, stream));
  HG_HIP(hipMemcpyAsync(ws.r.ptr, b, sizeof(double) * n * k, hipMemcpyDeviceToDevice, stream));
  if (!componentwise) {
    column_absmax_kernel<<<static_cast<unsigned>(k), kFrontThreads, 0, stream>>>(b, n, ws.bmax.ptr);
  }
  int32_t rounds = 0;
  for (; rounds < max_rounds;) {
    {
      PhaseScope scope(timer, sweep_phase, stream);
      gather_perm_kernel<<<grid, 256, 0, stream>>>(ws.r.ptr, plan.perm.ptr, pinned_new, n, k,
                                                   ws.rp.ptr);
      if (use_fp32) {
        triangular_solve<float>(plan, factor.w32.ptr, ws.rp.ptr, ws.y.ptr, ws.xp.ptr, ws.fvec.ptr,
                                k, stream);
      } else {
        triangular_solve<double>(plan, factor.w64.ptr, ws.rp.ptr, ws.y.ptr, ws.xp.ptr,
                                 ws.fvec.ptr, k, stream);
      }
      scatter_add_perm_kernel<<<grid, 256, 0, stream>>>(ws.xp.ptr, plan.perm.ptr, n, k, x);
    }
    ++rounds;
    {
      PhaseScope scope(timer, kPhaseResidual, stream);
      HG_HIP(hipMemsetAsync(ws.pending.ptr, 0, sizeof(int32_t), stream));
      residual_test_kernel<<<grid, 256, 0, stream>>>(offsets, columns, values, x, b, dropped_old,
                                                     ws.bmax.ptr, n, k, tolerance,
                                                     componentwise ? 1 : 0, ws.r.ptr,
                                                     ws.pending.ptr);
    }
    HG_HIP(hipGetLastError());
    int32_t pending = 0;
    HG_HIP(hipMemcpyAsync(&pending, ws.pending.ptr, sizeof(int32_t), hipMemcpyDeviceToHost, stream));
    HG_HIP(hipStreamSynchronize(stream));
    if (timer) timer->resolve();
    if (!pending) break;
  }
  return rounds;
}

void remove_null_space(const int32_t* comp_ptr, const int32_t* members, int32_t n_comp, int64_t n,
                       int64_t k, double* x, hipStream_t stream) {
  if (n_comp == 0 || k == 0) return;
  dim3 grid(static_cast<unsigned>(n_comp), static_cast<unsigned>(k));
  null_space_kernel<<<grid, kFrontThreads, 0, stream>>>(comp_ptr, members, n, x);
  HG_HIP(hipGetLastError());
}

}  // namespace hg
