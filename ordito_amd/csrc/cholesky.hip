// Multifrontal sparse Cholesky on the device: numeric factorization level by level and the
// partitioned-inverse triangular solves, with FP64 refinement against the operator.
//
// Follows ordito's design (cholesky.py:1-35, kernels/cholesky.py:1-31): each supernode's front is
// a dense matrix assembled from the operator's entries and its children's Schur complements
// (extend-add), partially factored, and its solve block stored so both triangular sweeps are
// level-scheduled matrix products (the partitioned inverse of Alvarado, Pothen and Schreiber 1993).
// The forward sweep here is multifrontal: each front gathers its children's update vectors through
// the same relative index maps the factorization's extend-add uses, so no sum is atomic and every
// sum has one fixed order.
//
// The partial factorization is blocked right-looking over 64-column panels, every front of a level
// in the same launches. The solve block is the inverse factor Z = [L11^-1; -L21 L11^-1], the first
// ncol columns of the inverse of the unit-extended lower factor [L11 0; L21 I]: that inverse is
// the product of the panels' elementary inverses [I 0 0; 0 D^-1 0; 0 -P D^-1 I] (D the panel's
// diagonal block factor, P the panel below it), so each panel step updates Z with the same tile
// product that updates the trailing matrix, and no separate triangular inversion is needed.
#include <utility>

#include "cholesky.h"

namespace hg {
namespace {

constexpr int kFt = kFactorTile;

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
__global__ void pin_diagonal_kernel(const int64_t* __restrict__ pin_front_slot, int32_t n_pins,
                                    double* __restrict__ arena) {
  int32_t k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k < n_pins) arena[pin_front_slot[k]] = 1.0;
}

struct FactorArgs {
  const int32_t* ncol;
  const int32_t* nrow;
  const int32_t* row_ptr;
  const int32_t* rel;
  const int64_t* front_off;
  const int64_t* w_off;
  double* arena;   // fronts, column-major m x m
  double* z;       // inverse factors, column-major m x ncol, zero on entry
  int32_t* status;
};

// One child's Schur-complement tile added into its parent through the relative map (lower part).
// Tasks of one launch are one child rank, so no two of them write the same parent entry.
__global__ __launch_bounds__(256) void extend_add_kernel(const int32_t* __restrict__ tasks,
                                                         FactorArgs a) {
  const int32_t* task = tasks + 4 * static_cast<int64_t>(blockIdx.x);
  const int32_t s = task[0], c = task[1], i0 = task[2], k0 = task[3];
  const int32_t ncc = a.ncol[c], rc = a.nrow[c];
  const int64_t mc = ncc + rc;
  const int64_t ms = a.ncol[s] + a.nrow[s];
  const double* child = a.arena + a.front_off[c];
  double* parent = a.arena + a.front_off[s];
  const int32_t* map = a.rel + a.row_ptr[c];
  // A thread's sixteen entries: the maps, then the child's and the parent's entries, all loaded
  // before any sum is stored. Entries past the child's rows read its last row instead (a task
  // only exists for rc >= 1), so every load is in bounds and none waits behind a branch.
  int32_t mi[16], mk[16];
  double add[16], cur[16];
  bool live[16];
#pragma unroll
  for (int u = 0; u < 16; ++u) {
    const int e = threadIdx.x + 256 * u;
    const int32_t i = i0 + e % kFt, k = k0 + e / kFt;
    live[u] = i < rc && k < rc && i >= k;
    const int32_t ic = min(i, rc - 1), kc = min(k, rc - 1);
    mi[u] = map[ic];
    mk[u] = map[kc];
    add[u] = child[(ncc + kc) * mc + ncc + ic];
  }
  int64_t at[16];
#pragma unroll
  for (int u = 0; u < 16; ++u) at[u] = static_cast<int64_t>(mk[u]) * ms + mi[u];
#pragma unroll
  for (int u = 0; u < 16; ++u) cur[u] = parent[at[u]];
#pragma unroll
  for (int u = 0; u < 16; ++u)
    if (live[u]) parent[at[u]] = cur[u] + add[u];
}

// The FP64 matrix cores: v_mfma_f64_16x16x4 takes a 16 x 4 block of A (lane l holds
// A[l % 16][l / 16]) and a 4 x 16 block of B (lane l holds B[l / 16][l % 16]) and accumulates the
// 16 x 16 product, lane l holding rows l / 16 + 4v (v < 4) of column l % 16. A 64 x 64 tile is four
// waves of 16 rows, each accumulating four 16-column blocks over the inner dimension four at a
// time; operands read from LDS, zero past the inner extent b (rounded up to a multiple of four).
typedef double f64x4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ f64x4 mfma16x16x4(double a, double b, f64x4 c) {
  return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
}

// C += A B for a 64 x 64 tile: A(r, k) = a_at(r, k), B(k, c) = b_at(k, c), k < kb_end.
template <typename FA, typename FB>
__device__ __forceinline__ void tile_product(FA a_at, FB b_at, int32_t kb_end, int wave, int lane,
                                             f64x4 acc[4]) {
  const int li = lane % 16, lk = lane / 16;
  for (int32_t kb = 0; kb < kb_end; kb += 4) {
    const double av = a_at(16 * wave + li, kb + lk);
#pragma unroll
    for (int n = 0; n < 4; ++n) acc[n] = mfma16x16x4(av, b_at(kb + lk, 16 * n + li), acc[n]);
  }
}

// Panel step at column j0 of front s, first launch: the task of kind 2 eliminates the b x b
// diagonal block D while applying the same eliminations to the identity, which leaves
// Zd = L_D^-1 (right-looking: at pivot k, rows i > k lose D(i,k)/D(k,k) times row k of D's
// trailing part and of Zd, then Zd's row k is divided by sqrt(D(k,k))), and stores Zd as Z's
// diagonal block; the launch's other tasks return. Once per front panel, where every task of the
// panel once repeated it. A non-positive pivot raises status[0] (cholesky.py:260-266).
//
// Thread (row i, column group cg = its wave) keeps entries (i, cg + 4v), v < 16, in registers:
// D's lower triangle (zero above) on entry. Zd's column k only becomes nonzero at pivot k, when
// D's column k is spent, so one register serves both: column j holds D's entry before pivot j and
// Zd's after it. A pivot's shared operands (Zd's row k before its scaling, 1 at the diagonal, D's
// column k below it, and D(k,k)) are published by the threads that hold them into a two-slot LDS
// buffer before the pivot's barrier; the next pivot's go to the other slot, so one barrier a pivot
// orders both. The pivots are unrolled (one instance per k), so every register index is fixed at
// compile time. Slot layout: entry j at (j % 4) * 16 + j / 4, so wave cg's sixteen operands are
// contiguous.
__device__ __forceinline__ int pivot_slot(int j) { return (j % 4) * 16 + j / 4; }

struct PanelPivots {
  double e[16];
  double (*ops)[kFt];
  double* piv;
  int i, cg, b;
  bool bad;
  double dk;  // D(i, i) at pivot i (1 if it was not positive)

  template <int K>
  __device__ __forceinline__ void pivot() {
    constexpr int p = K & 1, q = p ^ 1;
    __syncthreads();
    double d = piv[p];
    if (!(d > 0.0)) {
      bad = true;
      d = 1.0;
    }
    if (i == K) dk = d;
    if (i > K && i < b) {
      const double lik = ops[p][pivot_slot(i)] / d;
      double other[16];
#pragma unroll
      for (int v = 0; v < 16; ++v) other[v] = ops[p][cg * 16 + v];
      // Zd's entry (i, K) is zero before this pivot (the identity's) while e holds D's there: the
      // update starts from zero.
      if (cg == K % 4) e[K / 4] = 0.0;
#pragma unroll
      for (int v = 0; v < 16; ++v) e[v] = e[v] - lik * other[v];
    }
    // Pivot K + 1's operands, from this pivot's results: row K + 1 of Zd (its lane in every wave),
    // D's column K + 1 below the diagonal and D(K + 1, K + 1) (wave (K + 1) % 4).
    if (K + 1 < b) {
      if (i == K + 1) {
#pragma unroll
        for (int v = 0; v < 16; ++v)
          if (cg + 4 * v <= K) ops[q][cg * 16 + v] = e[v];
      }
      if (cg == (K + 1) % 4 && i >= K + 1) {
        const double c = e[(K + 1) / 4];
        ops[q][pivot_slot(i)] = i == K + 1 ? 1.0 : c;
        if (i == K + 1) piv[q] = c;
      }
    }
  }

  // Pivots 0, 1, ... in order while K < b (&& stops at the first that is not).
  template <int... K>
  __device__ __forceinline__ void run(std::integer_sequence<int, K...>) {
    (void)((K < b && (pivot<K>(), true)) && ...);
  }
};

__global__ __launch_bounds__(256) void panel_diag_kernel(const int32_t* __restrict__ tasks,
                                                         FactorArgs a) {
  __shared__ double ops[2][kFt];
  __shared__ double piv[2];
  const int32_t* task = tasks + 4 * static_cast<int64_t>(blockIdx.x);
  if (task[2] != 2) return;
  const int32_t s = task[0], j0 = task[1];
  const int32_t nc = a.ncol[s];
  const int64_t m = nc + a.nrow[s];
  const double* F = a.arena + a.front_off[s];
  double* Z = a.z + a.w_off[s];
  const int t = threadIdx.x;
  PanelPivots pp;
  pp.ops = ops;
  pp.piv = piv;
  pp.i = t % kFt;
  pp.cg = t / kFt;
  pp.b = min(kFt, nc - j0);
  pp.bad = false;
  pp.dk = 1.0;
  const int i = pp.i, cg = pp.cg, b = pp.b;
#pragma unroll
  for (int v = 0; v < 16; ++v) {
    const int j = cg + 4 * v;
    pp.e[v] = (i < b && j <= i) ? F[(j0 + j) * m + j0 + i] : 0.0;
  }
  // Pivot 0's operands: D's column 0 (wave 0 holds it in e[0]), 1 at the diagonal, D(0,0).
  if (cg == 0) {
    ops[0][pivot_slot(i)] = i == 0 ? 1.0 : pp.e[0];
    if (i == 0) piv[0] = pp.e[0];
  }
  pp.run(std::make_integer_sequence<int, kFt>{});
  if (pp.bad && t == 0) atomicOr(a.status, 1);
  // Zd's row i divided by sqrt(D(i,i)): no pivot after i reads it, so the scaling waits for the
  // end (its diagonal entry is the identity's 1 until then).
  const double inv = 1.0 / sqrt(pp.dk);
#pragma unroll
  for (int v = 0; v < 16; ++v) {
    const int j = cg + 4 * v;
    const double z = j == i ? 1.0 * inv : pp.e[v] * inv;
    if (i < b && j <= i) Z[static_cast<int64_t>(j0 + j) * m + j0 + i] = z;
  }
}

// Panel step, second launch: with Zd read from Z's diagonal block (lower triangle; zero above),
// kind 0 solves panel rows [t0, t0+64): X = F(rows, J) Zd^T; kind 1 applies Zd to Z's block rows
// J at columns [t0, t0+64): Zd Z(J, cols). Kind 2 tasks return. Products on the matrix cores.
__global__ __launch_bounds__(256) void panel_apply_kernel(const int32_t* __restrict__ tasks,
                                                          FactorArgs a) {
  __shared__ double Zs[kFt][kFt + 1];
  __shared__ double Ts[kFt][kFt + 1];
  const int32_t* task = tasks + 4 * static_cast<int64_t>(blockIdx.x);
  const int32_t kind = task[2];
  if (kind == 2) return;
  const int32_t s = task[0], j0 = task[1], t0 = task[3];
  const int32_t nc = a.ncol[s];
  const int64_t m = nc + a.nrow[s];
  const int32_t b = min(kFt, nc - j0);
  double* F = a.arena + a.front_off[s];
  double* Z = a.z + a.w_off[s];
  const int t = threadIdx.x;
  // Every load is issued before any LDS store (see update_kernel).
  double zv[16], tv[16];
#pragma unroll
  for (int u = 0; u < 16; ++u) {
    const int e = t + 256 * u;
    const int i = e % kFt, j = e / kFt;
    zv[u] = (i < b && j < b && i >= j) ? Z[static_cast<int64_t>(j0 + j) * m + j0 + i] : 0.0;
    if (kind == 0) {
      const int64_t r = t0 + i;
      tv[u] = (r < m && j < b) ? F[(j0 + j) * m + r] : 0.0;
    } else {
      // Ts[k][cc] = Z(j0 + k, t0 + cc), read along k (contiguous).
      tv[u] = (i < b) ? Z[static_cast<int64_t>(t0 + j) * m + j0 + i] : 0.0;
    }
  }
#pragma unroll
  for (int u = 0; u < 16; ++u) {
    const int e = t + 256 * u;
    const int i = e % kFt, j = e / kFt;
    Zs[i][j] = zv[u];
    Ts[i][j] = tv[u];
  }
  __syncthreads();
  const int wave = t / 64, lane = t % 64;
  f64x4 acc[4];
#pragma unroll
  for (int n = 0; n < 4; ++n) acc[n] = f64x4{0.0, 0.0, 0.0, 0.0};
  if (kind == 0) {
    tile_product([&](int r, int k) { return Ts[r][k]; }, [&](int k, int c) { return Zs[c][k]; },
                 b, wave, lane, acc);
  } else {
    tile_product([&](int r, int k) { return Zs[r][k]; }, [&](int k, int c) { return Ts[k][c]; },
                 b, wave, lane, acc);
  }
#pragma unroll
  for (int n = 0; n < 4; ++n) {
    const int cc = 16 * n + lane % 16;
#pragma unroll
    for (int v = 0; v < 4; ++v) {
      const int rr = 16 * wave + lane / 16 + 4 * v;
      if (kind == 0) {
        const int64_t r = t0 + rr;
        if (r < m && cc < b) F[(j0 + cc) * m + r] = acc[n][v];
      } else if (rr < b) {
        Z[static_cast<int64_t>(t0 + cc) * m + j0 + rr] = acc[n][v];
      }
    }
  }
}

// Update after the panel step at j0: rows [i0, i0+64) of the trailing matrix (c0 >= j0 + b, lower
// part) lose P(i0..) P(c0..)^T, or of Z (c0 < j0 + b) lose P(i0..) Z(J, c0..), where P is the
// solved panel (the front's columns J = [j0, j0+b)). Products on the matrix cores.
__global__ __launch_bounds__(256) void update_kernel(const int32_t* __restrict__ tasks,
                                                     FactorArgs a) {
  __shared__ double As[kFt][kFt + 1];  // [k][row]
  __shared__ double Bs[kFt][kFt + 1];  // [k][column]
  const int32_t* task = tasks + 4 * static_cast<int64_t>(blockIdx.x);
  const int32_t s = task[0], j0 = task[1], i0 = task[2], c0 = task[3];
  const int32_t nc = a.ncol[s];
  const int64_t m = nc + a.nrow[s];
  const int32_t b = min(kFt, nc - j0);
  const bool trailing = c0 >= j0 + b;
  double* F = a.arena + a.front_off[s];
  double* Z = a.z + a.w_off[s];
  const int t = threadIdx.x;
  // A thread's sixteen entries of each operand: every load is issued before any LDS store, so
  // they overlap instead of each waiting behind the store before it.
  double av[16], bv[16];
#pragma unroll
  for (int u = 0; u < 16; ++u) {
    const int e = t + 256 * u;
    const int r = e % kFt, k = e / kFt;
    av[u] = (i0 + r < m && k < b) ? F[(j0 + k) * m + i0 + r] : 0.0;
    if (trailing) {
      bv[u] = (c0 + r < m && k < b) ? F[(j0 + k) * m + c0 + r] : 0.0;
    } else {
      // Bs[k][cc] = Z(j0 + k, c0 + cc): e % 64 runs along k here, contiguous in Z.
      bv[u] = (r < b && c0 + k < j0 + b) ? Z[static_cast<int64_t>(c0 + k) * m + j0 + r] : 0.0;
    }
  }
  // The output tile's entries are loaded with the operands: no task of the launch writes them
  // before this one does, and its products read none of them.
  const int wave = t / 64, lane = t % 64;
  double* C = trailing ? F : Z;
  double cv[4][4];
  bool live[4][4];
#pragma unroll
  for (int n = 0; n < 4; ++n) {
    const int64_t c = c0 + 16 * n + lane % 16;
#pragma unroll
    for (int v = 0; v < 4; ++v) {
      const int64_t r = i0 + 16 * wave + lane / 16 + 4 * v;
      live[n][v] = r < m && (trailing ? (c < m && r >= c) : c < j0 + b);
      cv[n][v] = live[n][v] ? C[c * m + r] : 0.0;
    }
  }
#pragma unroll
  for (int u = 0; u < 16; ++u) {
    const int e = t + 256 * u;
    const int r = e % kFt, k = e / kFt;
    As[k][r] = av[u];
    if (trailing) Bs[k][r] = bv[u];
    else Bs[r][k] = bv[u];
  }
  __syncthreads();
  f64x4 acc[4];
#pragma unroll
  for (int n = 0; n < 4; ++n) acc[n] = f64x4{0.0, 0.0, 0.0, 0.0};
  tile_product([&](int r, int k) { return As[k][r]; }, [&](int k, int c) { return Bs[k][c]; }, b,
               wave, lane, acc);
#pragma unroll
  for (int n = 0; n < 4; ++n) {
    const int64_t c = c0 + 16 * n + lane % 16;
#pragma unroll
    for (int v = 0; v < 4; ++v) {
      const int64_t r = i0 + 16 * wave + lane / 16 + 4 * v;
      if (live[n][v]) C[c * m + r] = cv[n][v] - acc[n][v];
    }
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

HostSweep sweep_tasks(const Plan& plan, bool forward, int32_t target_tasks) {
  const int64_t kTargetTasks = std::max<int32_t>(target_tasks, 1);
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
  // The lists hold front-vector rows as int32.
  HG_REQUIRE(plan.vec_size < (int64_t{1} << 31), HG_ERR_VALUE,
             "system too large: its front vectors exceed the int32 range");
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
  // The first two sources of each position inline (-1: none), so a gather is one 8-byte load
  // and the update rows it names; -2 in the second marks a position with more, read on from the
  // lists.
  g.pair.assign(2 * plan.vec_size, -1);
  for (int64_t v = 0; v < plan.vec_size; ++v) {
    const int32_t count = g.ptr[v + 1] - g.ptr[v];
    if (count >= 1) g.pair[2 * v] = g.src[g.ptr[v]];
    if (count == 2) g.pair[2 * v + 1] = g.src[g.ptr[v] + 1];
    if (count > 2) g.pair[2 * v + 1] = -2;
  }
  return g;
}

void validate_sweeps(const Plan& plan, int32_t target_tasks) {
  auto fail = [](const std::string& what) { throw Error{HG_ERR_RUNTIME, "plan check: " + what}; };
  const int64_t S = plan.n_super;
  for (int dir = 0; dir < 2; ++dir) {
    const bool forward = dir == 0;
    const HostSweep h = sweep_tasks(plan, forward, target_tasks);
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
  if (static_cast<int64_t>(g.pair.size()) != 2 * plan.vec_size) fail("gather pair size");
  for (int64_t v = 0; v < plan.vec_size; ++v) {
    const int32_t count = g.ptr[v + 1] - g.ptr[v];
    const int32_t first = g.pair[2 * v], second = g.pair[2 * v + 1];
    if (first != (count >= 1 ? g.src[g.ptr[v]] : -1)) fail("gather pair first source");
    if (second != (count == 2 ? g.src[g.ptr[v] + 1] : (count > 2 ? -2 : -1)))
      fail("gather pair second source");
  }
  // Every structure row of every child lands in exactly one parent position.
  int64_t rows_with_parent = 0;
  for (int64_t s = 0; s < S; ++s)
    if (plan.parent[s] >= 0) rows_with_parent += plan.nrow[s];
  if (static_cast<int64_t>(g.src.size()) != rows_with_parent) fail("gather list count");
}

HostFactorTasks factor_tasks(const Plan& plan) {
  HostFactorTasks out;
  auto push = [&](int32_t a, int32_t b, int32_t c, int32_t d) {
    out.tasks.insert(out.tasks.end(), {a, b, c, d});
  };
  auto task_count = [&]() { return static_cast<int32_t>(out.tasks.size() / 4); };
  for (int32_t level = 0; level < plan.n_levels; ++level) {
    const int32_t lo = plan.level_ptr[level], hi = plan.level_ptr[level + 1];
    int32_t max_children = 0, max_ncol = 0;
    for (int32_t k = lo; k < hi; ++k) {
      const int32_t s = plan.level_nodes[k];
      max_children = std::max(max_children, plan.child_ptr[s + 1] - plan.child_ptr[s]);
      max_ncol = std::max(max_ncol, plan.ncol[s]);
    }
    for (int32_t rank = 0; rank < max_children; ++rank) {
      const int32_t first = task_count();
      for (int32_t k = lo; k < hi; ++k) {
        const int32_t s = plan.level_nodes[k];
        if (plan.child_ptr[s] + rank >= plan.child_ptr[s + 1]) continue;
        const int32_t c = plan.children[plan.child_ptr[s] + rank];
        for (int32_t k0 = 0; k0 < plan.nrow[c]; k0 += kFt)
          for (int32_t i0 = k0; i0 < plan.nrow[c]; i0 += kFt) push(s, c, i0, k0);
      }
      if (task_count() > first) out.launches.push_back({kFactorExtend, level, first, task_count() - first});
    }
    for (int32_t j0 = 0; j0 < max_ncol; j0 += kFt) {
      int32_t first = task_count();
      for (int32_t k = lo; k < hi; ++k) {
        const int32_t s = plan.level_nodes[k];
        const int32_t nc = plan.ncol[s], m = nc + plan.nrow[s];
        if (j0 >= nc) continue;
        const int32_t b = std::min(kFt, nc - j0);
        push(s, j0, 2, j0);
        for (int32_t r0 = j0 + b; r0 < m; r0 += kFt) push(s, j0, 0, r0);
        for (int32_t c0 = 0; c0 < j0; c0 += kFt) push(s, j0, 1, c0);
      }
      if (task_count() > first) out.launches.push_back({kFactorPanel, level, first, task_count() - first});
      first = task_count();
      for (int32_t k = lo; k < hi; ++k) {
        const int32_t s = plan.level_nodes[k];
        const int32_t nc = plan.ncol[s], m = nc + plan.nrow[s];
        if (j0 >= nc) continue;
        const int32_t b = std::min(kFt, nc - j0);
        for (int32_t i0 = j0 + b; i0 < m; i0 += kFt) {
          for (int32_t c0 = j0 + b; c0 <= i0; c0 += kFt) push(s, j0, i0, c0);
          for (int32_t c0 = 0; c0 < j0 + b; c0 += kFt) push(s, j0, i0, c0);
        }
      }
      if (task_count() > first) out.launches.push_back({kFactorUpdate, level, first, task_count() - first});
    }
  }
  return out;
}

void validate_factor_tasks(const Plan& plan, const HostFactorTasks& h) {
  auto fail = [](const std::string& what) { throw Error{HG_ERR_RUNTIME, "factor plan check: " + what}; };
  const int32_t S = plan.n_super;
  const int64_t n_tasks = static_cast<int64_t>(h.tasks.size() / 4);
  if (static_cast<int64_t>(h.tasks.size()) != 4 * n_tasks) fail("task array not a multiple of 4");
  int64_t expect_first = 0;
  int32_t last_level = 0;
  // Per front: the panel steps seen, in order, and the extend-add tiles per child.
  std::vector<int32_t> next_j0(S, 0);
  std::vector<int64_t> ext_tiles(S, 0);
  for (const FactorLaunch& L : h.launches) {
    if (L.first != expect_first || L.count <= 0) fail("launches do not tile the task list");
    expect_first += L.count;
    if (L.level < last_level || L.level >= plan.n_levels) fail("launch level out of order");
    last_level = L.level;
    for (int64_t t = L.first; t < L.first + L.count; ++t) {
      const int32_t* f = h.tasks.data() + 4 * t;
      const int32_t s = f[0];
      if (s < 0 || s >= S) fail("task front out of range");
      if (plan.level[s] != L.level) fail("task front not on the launch's level");
      const int32_t nc = plan.ncol[s], m = nc + plan.nrow[s];
      if (L.kind == kFactorExtend) {
        const int32_t c = f[1];
        if (c < 0 || c >= S || plan.parent[c] != s) fail("extend-add child is not a child of the front");
        const int32_t rc = plan.nrow[c];
        if (f[2] < 0 || f[3] < 0 || f[2] >= rc || f[3] >= rc || f[2] < f[3] || f[2] % kFt ||
            f[3] % kFt)
          fail("extend-add tile out of range");
        ext_tiles[c] += 1;
        continue;
      }
      const int32_t j0 = f[1];
      if (j0 < 0 || j0 >= nc || j0 % kFt) fail("panel column out of range");
      const int32_t b = std::min(kFt, nc - j0);
      if (L.kind == kFactorPanel) {
        const int32_t kind = f[2], t0 = f[3];
        if (kind == 2) {
          if (t0 != j0 || next_j0[s] != j0) fail("diagonal task out of step order");
          next_j0[s] = j0 + b;
        } else if (kind == 0) {
          if (t0 < j0 + b || t0 >= m || (t0 - j0 - b) % kFt) fail("panel row tile out of range");
        } else if (kind == 1) {
          if (t0 < 0 || t0 + kFt > j0 || t0 % kFt) fail("inverse-factor tile out of range");
        } else {
          fail("unknown panel task kind");
        }
      } else if (L.kind == kFactorUpdate) {
        const int32_t i0 = f[2], c0 = f[3];
        if (next_j0[s] != j0 + b) fail("update before its panel");
        if (i0 < j0 + b || i0 >= m || (i0 - j0 - b) % kFt) fail("update row tile out of range");
        if (c0 >= j0 + b) {
          if (c0 > i0 || (c0 - j0 - b) % kFt) fail("trailing tile out of range");
        } else if (c0 < 0 || c0 % kFt) {
          fail("inverse-factor update tile out of range");
        }
      } else {
        fail("unknown launch kind");
      }
    }
  }
  if (expect_first != n_tasks) fail("tasks outside every launch");
  for (int32_t s = 0; s < S; ++s) {
    if (next_j0[s] != plan.ncol[s]) fail("a front's panels do not cover its columns");
    const int64_t tiles = (plan.nrow[s] + kFt - 1) / kFt;
    if (plan.parent[s] >= 0 && ext_tiles[s] != tiles * (tiles + 1) / 2)
      fail("extend-add tiles do not cover a child's Schur complement");
  }
}

namespace {

void upload_sweep(const Plan& plan, bool forward, const HostSweep& h,
                  DevicePlan::SweepTasks& out, hipStream_t stream) {
  std::vector<int4> desc(3 * h.s.size());
  for (size_t t = 0; t < h.s.size(); ++t) {
    const int32_t s = h.s[t];
    const int64_t w_off = plan.w_off[s];
    const int64_t third = forward ? plan.vec_off[s] : static_cast<int64_t>(plan.row_ptr[s]);
    desc[3 * t] = make_int4(s, h.tile[t], h.k0[t], h.k1[t]);
    desc[3 * t + 1] = make_int4(h.slot[t], plan.ncol[s], plan.col_begin[s], plan.ncol[s] + plan.nrow[s]);
    desc[3 * t + 2] = make_int4(static_cast<int32_t>(w_off & 0xffffffff), static_cast<int32_t>(w_off >> 32),
                                static_cast<int32_t>(third & 0xffffffff), static_cast<int32_t>(third >> 32));
  }
  out.desc.upload(desc, stream);
  std::vector<int4> rdesc(3 * h.rs.size());
  for (size_t g = 0; g < h.rs.size(); ++g) {
    const int32_t s = h.rs[g];
    const int64_t third = forward ? plan.vec_off[s] : 0;
    rdesc[3 * g] = make_int4(s, h.rtile[g], h.rfirst[g], h.rcount[g]);
    rdesc[3 * g + 1] = make_int4(0, plan.ncol[s], plan.col_begin[s], plan.ncol[s] + plan.nrow[s]);
    rdesc[3 * g + 2] = make_int4(0, 0, static_cast<int32_t>(third & 0xffffffff),
                                 static_cast<int32_t>(third >> 32));
  }
  out.rdesc.upload(rdesc, stream);
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

void DevicePlan::upload_sweeps(const Plan& plan, int32_t target_tasks, hipStream_t stream) {
  const HostSweep hf = sweep_tasks(plan, true, target_tasks);
  const HostSweep hb = sweep_tasks(plan, false, target_tasks);
  max_slots = std::max(hf.max_slots, hb.max_slots);
  upload_sweep(plan, true, hf, fwd, stream);
  upload_sweep(plan, false, hb, bwd, stream);
  split_target = target_tasks;
}

void DevicePlan::upload(const Plan& plan, const HostPattern& pattern, int32_t target_tasks,
                        hipStream_t stream) {
  {
    HostFactorTasks ft = hg::factor_tasks(plan);
    factor_launches = std::move(ft.launches);
    factor_tasks.upload(ft.tasks, stream);
  }
  upload_sweeps(plan, target_tasks, stream);
  {
    const HostGather g = gather_lists(plan);
    gather_ptr.upload(g.ptr, stream);
    gather_src.upload(g.src, stream);
    gather_pair.upload(g.pair, stream);
  }
  // The operator's pattern in the elimination order: row i is old row perm[i], its columns
  // renumbered and sorted; src names each entry's slot in the original values.
  {
    const int64_t nn = plan.n;
    std::vector<int32_t> offs(nn + 1, 0);
    for (int64_t i = 0; i < nn; ++i) {
      const int32_t old = plan.perm[i];
      offs[i + 1] = offs[i] + (pattern.offsets[old + 1] - pattern.offsets[old]);
    }
    std::vector<int32_t> cols(static_cast<size_t>(offs[nn])), src(static_cast<size_t>(offs[nn]));
    parallel_chunks(nn, 1 << 14, [&](int64_t b, int64_t e_end, int64_t) {
      std::vector<std::pair<int32_t, int32_t>> row;
      for (int64_t i = b; i < e_end; ++i) {
        const int32_t old = plan.perm[i];
        row.clear();
        for (int32_t e = pattern.offsets[old]; e < pattern.offsets[old + 1]; ++e)
          row.push_back({plan.iperm[pattern.columns[e]], e});
        std::sort(row.begin(), row.end());
        int32_t at = offs[i];
        for (const auto& entry : row) {
          cols[at] = entry.first;
          src[at] = entry.second;
          ++at;
        }
      }
    });
    nnz = static_cast<int64_t>(cols.size());
    p_offsets.upload(offs, stream);
    p_columns.upload(cols, stream);
    p_src.upload(src, stream);
  }
  n = plan.n;
  n_super = plan.n_super;
  n_levels = plan.n_levels;
  level_ptr = plan.level_ptr;
  level_gathers.assign(plan.n_levels, 0);
  for (int32_t level = 0; level < plan.n_levels; ++level)
    for (int32_t k = plan.level_ptr[level]; k < plan.level_ptr[level + 1]; ++k) {
      const int32_t s = plan.level_nodes[k];
      if (plan.child_ptr[s + 1] > plan.child_ptr[s]) level_gathers[level] = 1;
    }
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
         fwd.desc.bytes() + fwd.rdesc.bytes() + bwd.desc.bytes() + bwd.rdesc.bytes() +
         gather_ptr.bytes() + gather_src.bytes() + gather_pair.bytes() +
         p_offsets.bytes() + p_columns.bytes() + p_src.bytes() + factor_tasks.bytes();
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
    // Arena offsets: a large plan's exceed the int32 range.
    for (int64_t slot : pin_slots)
      HG_REQUIRE(slot >= 0 && slot < plan.arena_size, HG_ERR_RUNTIME,
                 "pinned diagonal outside the front arena");
    DeviceArray<int64_t> d_slots;
    d_slots.upload(pin_slots, stream);
    pin_diagonal_kernel<<<blocks_for(static_cast<int64_t>(pin_slots.size())), 256, 0, stream>>>(
        d_slots.ptr, static_cast<int32_t>(pin_slots.size()), arena.ptr);
    HG_HIP(hipGetLastError());
    HG_HIP(hipStreamSynchronize(stream));
  }
  factor.w64.allocate(plan.w_size);
  if (plan.w_size > 0)
    HG_HIP(hipMemsetAsync(factor.w64.ptr, 0, sizeof(double) * plan.w_size, stream));
  FactorArgs a{};
  a.ncol = plan.ncol.ptr;
  a.nrow = plan.nrow.ptr;
  a.row_ptr = plan.row_ptr.ptr;
  a.rel = plan.rel.ptr;
  a.front_off = plan.front_off.ptr;
  a.w_off = plan.w_off.ptr;
  a.arena = arena.ptr;
  a.z = factor.w64.ptr;
  a.status = status.ptr;
  for (const FactorLaunch& L : plan.factor_launches) {
    const int32_t* tasks = plan.factor_tasks.ptr + 4 * static_cast<int64_t>(L.first);
    const unsigned grid = static_cast<unsigned>(L.count);
    if (L.kind == kFactorExtend) extend_add_kernel<<<grid, 256, 0, stream>>>(tasks, a);
    else if (L.kind == kFactorPanel) {
      panel_diag_kernel<<<grid, 256, 0, stream>>>(tasks, a);
      panel_apply_kernel<<<grid, 256, 0, stream>>>(tasks, a);
    } else {
      update_kernel<<<grid, 256, 0, stream>>>(tasks, a);
    }
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
