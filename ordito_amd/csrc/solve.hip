// Refined multi-column solves through the partitioned-inverse factor, level by level.
//
// The whole refinement (ordito kernels/cholesky.py:1357-1463, cholesky.py:1606-1902) runs in the
// elimination order with the k right-hand sides of a chunk interleaved ([row][k]): the operator is
// held permuted (P A P^T), so a right-hand side in the caller's layout is permuted in once and the
// solution out once (fields stored by rows are already in that order), and every round is forward
// sweep, backward sweep, residual -- no per-round permutation. Each level of
// a sweep is a batch of dense tile products over every front of the level:
//   forward : y(cols) = Z11 f(cols) and the front's update vector u = f(rows) + Z21 f(cols), where
//             f = [r(cols); 0] + the children's update vectors, gathered as the tile loads it
//             (children in ascending order);
//   backward: d(cols) = Z11^T y(cols) + Z21^T d(rows), and the next iterate x + d beside it;
// Z = [L11^-1; -L21 L11^-1] is the front's inverse factor (cholesky.hip).
// A product task is one 64-wide tile of one front and a range of the inner dimension; a level with
// too few tiles to fill the device splits the inner dimension into partial slots, added in slot
// order by a reduction, so every sum keeps one fixed order. Tiles stage the solve block and the
// right-hand sides through LDS and accumulate in FP64 (FP32 solve blocks widen as they load).
#include "cholesky.h"

namespace hg {
namespace {

constexpr int kTile = 64;   // rows (forward) or columns (backward) of W per tile
constexpr int kCols = 64;   // right-hand sides per solve chunk
static_assert(kCols == kFieldChunk, "a row-layout field chunk is one solve chunk");
constexpr int kReduceSplit = 4;  // blocks a reduction group is spread over (16 rows each)

template <typename T>
__device__ __forceinline__ double widen(T v) {
  return static_cast<double>(v);
}

// b (user layout, column q at q*n) -> bp[i*kk + q] in elimination order through an LDS transpose
// of 64 rows; a pinned row's equation is dropped (zero).
__global__ __launch_bounds__(256) void perm_in_kernel(const double* __restrict__ b,
                                                      const int32_t* __restrict__ perm,
                                                      const int32_t* __restrict__ pinned,
                                                      int64_t n, int32_t kk,
                                                      double* __restrict__ bp) {
  __shared__ double tile[kTile][kCols + 1];
  const int64_t i0 = static_cast<int64_t>(blockIdx.x) * kTile;
  const int ii = threadIdx.x % kTile, qq = threadIdx.x / kTile;
  const int64_t i = i0 + ii;
  const bool live = i < n;
  const int64_t src = live ? perm[i] : 0;
  const bool drop = !live || (pinned && pinned[i]);
  for (int q = qq; q < kk; q += 4) tile[ii][q] = drop ? 0.0 : b[q * n + src];
  __syncthreads();
  const int32_t rows = static_cast<int32_t>(min<int64_t>(kTile, n - i0));
  for (int e = threadIdx.x; e < rows * kk; e += 256) {
    const int r_ = e / kk, q = e % kk;
    bp[(i0 + r_) * kk + q] = tile[r_][q];
  }
}

// out (user layout) [q*n + perm[i]] = x[i*kk + q], through the same transpose.
__global__ __launch_bounds__(256) void perm_out_kernel(const double* __restrict__ x,
                                                       const int32_t* __restrict__ perm,
                                                       int64_t n, int32_t kk,
                                                       double* __restrict__ out) {
  __shared__ double tile[kTile][kCols + 1];
  const int64_t i0 = static_cast<int64_t>(blockIdx.x) * kTile;
  const int32_t rows = static_cast<int32_t>(min<int64_t>(kTile, n - i0));
  for (int e = threadIdx.x; e < rows * kk; e += 256) tile[e / kk][e % kk] = x[i0 * kk + e];
  __syncthreads();
  const int ii = threadIdx.x % kTile, qq = threadIdx.x / kTile;
  if (ii >= rows) return;
  const int64_t dst = perm[i0 + ii];
  for (int q = qq; q < kk; q += 4) out[q * n + dst] = tile[ii][q];
}

// Pinned rows of a right-hand side stored by rows: zero, the pins' values (a component's last
// row, an unreferenced vertex) held there by the identity rows of the factor.
__global__ __launch_bounds__(256) void zero_pinned_rows_kernel(const int32_t* __restrict__ pinned,
                                                               int64_t n, int32_t kk,
                                                               double* __restrict__ b) {
  const int64_t i = static_cast<int64_t>(blockIdx.x) * 256 + threadIdx.x;
  if (i >= n || !pinned[i]) return;
  for (int32_t q = 0; q < kk; ++q) b[i * kk + q] = 0.0;
}

struct SweepArgs {
  const int32_t* col_begin;
  const int32_t* ncol;
  const int32_t* nrow;
  const int32_t* row_ptr;
  const int32_t* rows;
  const int64_t* w_off;
  const int64_t* vec_off;
  const int32_t* gather_pair;
  const int32_t* gather_ptr;
  const int32_t* gather_src;
  const double* r;   // residual, the sweep's right-hand side
  double* fvec;      // update vectors (rows part of each front's vector)
  double* y;         // forward result
  double* d;         // backward result (the correction)
  const double* x;   // current iterate; null on the first round (x = 0: d is the next iterate)
  double* x_next;    // x + d; null on the first round
  double* partials;
  int32_t kk;
  int32_t gathers;   // the level's fronts have children (forward sweep): f gathers their updates
};

// One gathered entry: v plus the children's update rows adding into front-vector row `row`,
// column q, in ascending child order (the sources as in gather_pairs).
__device__ __forceinline__ double gather1(const SweepArgs& a, int64_t row, int q, double v) {
  const int2 p = reinterpret_cast<const int2*>(a.gather_pair)[row];
  if (p.x >= 0) v += a.fvec[static_cast<int64_t>(p.x) * a.kk + q];
  if (p.y >= 0) {
    v += a.fvec[static_cast<int64_t>(p.y) * a.kk + q];
  } else if (p.y == -2) {
    for (int32_t g = a.gather_ptr[row] + 1; g < a.gather_ptr[row + 1]; ++g)
      v += a.fvec[static_cast<int64_t>(a.gather_src[g]) * a.kk + q];
  }
  return v;
}

// The children's update rows that add into front-vector rows row[u] (u < 4), in two steps so a
// caller can issue other loads between them: gather_pairs loads each row's first two sources
// (inline in gather_pair; -1: none; -2 in the second: more than two, the rest in the lists after
// the first; live[u] false: none), gather_add loads every update row of column q[u] and then adds
// them to v[u] in ascending child order. Two waits on memory for four entries, not two per entry.
// Absent sources read row 0 and are not added.
__device__ __forceinline__ void gather_pairs(const SweepArgs& a, const int64_t row[4],
                                             const bool live[4], int2 p[4]) {
  const int2* pairs = reinterpret_cast<const int2*>(a.gather_pair);
#pragma unroll
  for (int u = 0; u < 4; ++u) p[u] = live[u] ? pairs[row[u]] : make_int2(-1, -1);
}

__device__ __forceinline__ void gather_add(const SweepArgs& a, const int64_t row[4],
                                           const int q[4], const int2 p[4], double v[4]) {
  double first[4], second[4];
#pragma unroll
  for (int u = 0; u < 4; ++u) {
    first[u] = a.fvec[p[u].x >= 0 ? static_cast<int64_t>(p[u].x) * a.kk + q[u] : 0];
    second[u] = a.fvec[p[u].y >= 0 ? static_cast<int64_t>(p[u].y) * a.kk + q[u] : 0];
  }
#pragma unroll
  for (int u = 0; u < 4; ++u) {
    if (p[u].x >= 0) v[u] += first[u];
    if (p[u].y >= 0) {
      v[u] += second[u];
    } else if (p[u].y == -2) {
      for (int32_t g = a.gather_ptr[row[u]] + 1; g < a.gather_ptr[row[u] + 1]; ++g)
        v[u] += a.fvec[static_cast<int64_t>(a.gather_src[g]) * a.kk + q[u]];
    }
  }
}

// The tile products run on the FP64 matrix cores: v_mfma_f64_16x16x4 takes a 16 x 4 block of A
// (lane l holds A[l % 16][l / 16]) and a 4 x 16 block of B (lane l holds B[l / 16][l % 16]) and
// accumulates the 16 x 16 product, lane l holding rows l / 16 + 4v (v < 4) of column l % 16. A
// tile's four waves each own 16 of its 64 rows and accumulate all 64 right-hand sides as four
// 16-column blocks. Both operands are staged through LDS kInner inner indices at a time; the next
// stage's global loads are issued before the current stage's products, so they overlap.
typedef double f64x4 __attribute__((ext_vector_type(4)));
constexpr int kInner = 16;
// Staged rows are kTile + kPad doubles apart: the four inner rows an MFMA operand read spans start
// 34 banks apart (mod 64), and the backward's transposed stores of sixteen inner rows at one
// column land in distinct bank pairs.
constexpr int kPad = 17;

__device__ __forceinline__ f64x4 mfma16x16x4(double a, double b, f64x4 c) {
  return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
}

// The 4-deep steps over one stage: A from As[k][16 * wave + l % 16] (the wave's 16 output rows),
// B from Bs[k][16 n + l % 16] for each of the four 16-column blocks n of right-hand sides (columns
// past the chunk's are staged as zeros, so every block is multiplied). kBatch steps' operands are
// read before their products, so the reads overlap and the products issue back to back.
template <int kDepth = kInner, int kBatch = 2>
__device__ __forceinline__ void stage_products(const double (*As)[kTile + kPad],
                                               const double (*Bs)[kCols + kPad], int wave, int lane,
                                               f64x4 acc[4]) {
  const int li = lane % 16, lk = lane / 16;
#pragma unroll
  for (int kb = 0; kb < kDepth; kb += 4 * kBatch) {
    double av[kBatch], bv[kBatch][4];
#pragma unroll
    for (int h = 0; h < kBatch; ++h) {
      av[h] = As[kb + 4 * h + lk][16 * wave + li];
#pragma unroll
      for (int n = 0; n < 4; ++n) bv[h][n] = Bs[kb + 4 * h + lk][16 * n + li];
    }
#pragma unroll
    for (int h = 0; h < kBatch; ++h)
#pragma unroll
      for (int n = 0; n < 4; ++n) acc[n] = mfma16x16x4(av[h], bv[h][n], acc[n]);
  }
}

// stage_products over the column blocks that hold right-hand sides only (blocks of them), a step
// at a time: the backward's form, whose registers the unrolled form would spill.
template <int kDepth>
__device__ __forceinline__ void stage_products_blocks(const double (*As)[kTile + kPad],
                                                      const double (*Bs)[kCols + kPad], int wave,
                                                      int lane, int blocks, f64x4 acc[4]) {
  const int li = lane % 16, lk = lane / 16;
#pragma unroll 1
  for (int kb = 0; kb < kDepth; kb += 4) {
    const double av = As[kb + lk][16 * wave + li];
#pragma unroll
    for (int n = 0; n < 4; ++n)
      if (n < blocks) acc[n] = mfma16x16x4(av, Bs[kb + lk][16 * n + li], acc[n]);
  }
}

// Forward tile product for up to 64 right-hand sides: rows [r0, r0+64) of the front's inverse
// factor times f over the inner range [k0, k1), f gathered as it is staged. slot < 0 applies the
// product; otherwise it is stored as a partial for the reduction.
// kGathers: the level's fronts have children whose update vectors f gathers (a.gathers); the
// leaves' instance carries none of that code, and fits 80 registers: six waves a SIMD instead of
// four.
// A task's descriptor: {s, tile, k0, k1}, {slot, ncol, col_begin, m}, {a 64-bit offset of the
// front's solve block, a 64-bit second offset}, read together.
struct Task {
  int32_t s, tile, k0, k1, slot, nc, c0, m;
  int64_t w_off, other;
};

__device__ __forceinline__ Task load_task(const int4* __restrict__ desc, int32_t task) {
  const int4* d = desc + 3 * static_cast<int64_t>(task);
  const int4 d0 = d[0], d1 = d[1], d2 = d[2];
  Task t;
  t.s = d0.x;
  t.tile = d0.y;
  t.k0 = d0.z;
  t.k1 = d0.w;
  t.slot = d1.x;
  t.nc = d1.y;
  t.c0 = d1.z;
  t.m = d1.w;
  t.w_off = static_cast<int64_t>(static_cast<uint32_t>(d2.x)) | (static_cast<int64_t>(d2.y) << 32);
  t.other = static_cast<int64_t>(static_cast<uint32_t>(d2.z)) | (static_cast<int64_t>(d2.w) << 32);
  return t;
}

template <typename TW, bool kGathers>
__global__ __launch_bounds__(256, kGathers ? 4 : 6) void fwd_tile_kernel(
    const int4* __restrict__ desc, const TW* __restrict__ w, SweepArgs a) {
  __shared__ double Ws[kInner][kTile + kPad];  // [inner][row]
  __shared__ double Fs[kInner][kCols + kPad];  // [inner][right-hand side]
  const Task tk = load_task(desc, xcd_block());
  const int32_t r0 = tk.tile;
  const int32_t k0 = tk.k0, k1 = tk.k1;
  const int32_t slot = tk.slot;
  const int32_t nc = tk.nc, c0 = tk.c0;
  const int32_t m = tk.m;
  const int32_t kk = a.kk;
  const TW* W = w + tk.w_off;
  const int64_t vbase = tk.other;
  const int t = threadIdx.x, wave = t / 64, lane = t % 64;
  const bool rows_live = r0 + 16 * wave < m;  // wave-uniform
  // The epilogue's gather pairs (this lane's update rows lane / 16 + 4v), loaded before the stages.
  int2 epair[4];
  {
    int64_t erow[4];
    bool elive[4];
#pragma unroll
    for (int v = 0; v < 4; ++v) {
      const int32_t i = r0 + 16 * wave + lane / 16 + 4 * v;
      erow[v] = vbase + i;
      elive[v] = kGathers && slot < 0 && i >= nc && i < m;
    }
    gather_pairs(a, erow, elive, epair);
  }
  // Staging: entry e = t + 256u of a stage is W row e % 64 / f column e % 64 at inner e / 64.
  double wreg[4], freg[4];
  // f = r(cols) plus the children's updates: a stage's gather pairs are loaded a stage ahead
  // (load_pairs), so its W, r and update-row loads issue together.
  int2 pair[4];
  auto stage_rows = [&](int32_t j0, int64_t grow[4], int gq[4], bool glive[4]) {
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      const int e = t + 256 * u;
      const int32_t j = j0 + e / kTile;
      gq[u] = e % kCols;
      glive[u] = gq[u] < kk && j < k1;
      grow[u] = vbase + j;
    }
  };
  auto load_pairs = [&](int32_t j0) {
    if (!kGathers) return;
    int64_t grow[4];
    int gq[4];
    bool glive[4];
    stage_rows(j0, grow, gq, glive);
    gather_pairs(a, grow, glive, pair);
  };
  auto load = [&](int32_t j0) {
    int64_t grow[4];
    int gq[4];
    bool glive[4];
    stage_rows(j0, grow, gq, glive);
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      const int e = t + 256 * u;
      const int ii = e % kTile, jj = e / kTile;
      const int32_t i = r0 + ii, j = j0 + jj;
      wreg[u] = (i < m && j < k1) ? widen(W[static_cast<int64_t>(j) * m + i]) : 0.0;
      freg[u] = glive[u] ? a.r[static_cast<int64_t>(c0 + j) * kk + gq[u]] : 0.0;
    }
    if (kGathers) gather_add(a, grow, gq, pair, freg);
  };
  f64x4 acc[4];
#pragma unroll
  for (int n = 0; n < 4; ++n) acc[n] = f64x4{0.0, 0.0, 0.0, 0.0};
  load_pairs(k0);
  load(k0);
  if (k0 + kInner < k1) load_pairs(k0 + kInner);
  for (int32_t j0 = k0; j0 < k1; j0 += kInner) {
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      const int e = t + 256 * u;
      Ws[e / kTile][e % kTile] = wreg[u];
      Fs[e / kCols][e % kCols] = freg[u];
    }
    __syncthreads();
    if (j0 + kInner < k1) {
      load(j0 + kInner);
      if (j0 + 2 * kInner < k1) load_pairs(j0 + 2 * kInner);
    }
    if (rows_live) stage_products(Ws, Fs, wave, lane, acc);
    __syncthreads();
  }
  if (!rows_live) return;
  // Epilogue: every load (the children's update rows that add into this tile's update rows)
  // before any store, so the loads overlap instead of each waiting behind a store it might alias.
  // A lane's outputs are rows lane / 16 + 4v of its wave's block; the sums keep the gather's order
  // (0, then each child in ascending order, then the product).
  double base[4][4] = {};
  if (kGathers && slot < 0) {
#pragma unroll
    for (int n = 0; n < 4; ++n) {
      if (16 * n + lane % 16 >= kk) continue;
      int64_t erow[4];
      int gq[4];
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        erow[v] = vbase + r0 + 16 * wave + lane / 16 + 4 * v;
        gq[v] = 16 * n + lane % 16;
      }
      gather_add(a, erow, gq, epair, base[n]);
    }
  }
#pragma unroll
  for (int n = 0; n < 4; ++n) {
    const int q = 16 * n + lane % 16;
    if (q >= kk) continue;
#pragma unroll
    for (int v = 0; v < 4; ++v) {
      const int32_t ii = 16 * wave + lane / 16 + 4 * v;
      const int32_t i = r0 + ii;
      if (i >= m) continue;
      const double value = acc[n][v];
      if (slot >= 0) {
        a.partials[(static_cast<int64_t>(slot) * kTile + ii) * kk + q] = value;
      } else if (i < nc) {
        a.y[static_cast<int64_t>(c0 + i) * kk + q] = value;
      } else {
        a.fvec[(vbase + i) * kk + q] = base[n][v] + value;
      }
    }
  }
}

// Adds the reduction partials of entries off[u] (offsets inside a slot of `stride` doubles) over
// slots [first, first + count) in slot order into sum[u]; four slots' loads are issued before
// their adds, so the loads overlap while each sum keeps its order.
__device__ __forceinline__ void add_partials(const double* __restrict__ partials, int64_t stride,
                                             int32_t first, int32_t count, const int64_t off[4],
                                             double sum[4]) {
  const double* p = partials + static_cast<int64_t>(first) * stride;
  int32_t c = 0;
  for (; c + 4 <= count; c += 4) {
    double v[4][4];
#pragma unroll
    for (int cc = 0; cc < 4; ++cc)
#pragma unroll
      for (int u = 0; u < 4; ++u) v[cc][u] = p[(c + cc) * stride + off[u]];
#pragma unroll
    for (int cc = 0; cc < 4; ++cc)
#pragma unroll
      for (int u = 0; u < 4; ++u) sum[u] += v[cc][u];
  }
  for (; c < count; ++c)
#pragma unroll
    for (int u = 0; u < 4; ++u) sum[u] += p[c * stride + off[u]];
}

// Adds a split forward tile's partials in slot order and applies the sum; blockIdx.y picks 16 of
// the tile's 64 rows.
__global__ __launch_bounds__(256) void fwd_reduce_kernel(const int4* __restrict__ rdesc,
                                                         SweepArgs a) {
  const Task g = load_task(rdesc, xcd_block());
  const int32_t r0 = g.tile;
  const int32_t first = g.k0, count = g.k1;
  const int32_t nc = g.nc, m = g.m, c0 = g.c0;
  const int32_t kk = a.kk;
  const int64_t vbase = g.other;
  constexpr int kRows = kTile / kReduceSplit;
  const int32_t lo = blockIdx.y * kRows;
  const int32_t rows = min(kRows, m - r0 - lo);
  // At most kRows * kCols = 4 x 256 entries: each thread's four, every load before any store.
  // Entries past the group's rows read slot offset 0 and store nothing.
  double sum[4] = {}, base[4] = {};
  int64_t off[4], grow[4];
  int gq[4];
  bool glive[4];
#pragma unroll
  for (int u = 0; u < 4; ++u) {
    const int32_t e = threadIdx.x + 256 * u;
    const bool live = e < rows * kk;
    const int32_t ii = live ? lo + e / kk : 0;
    off[u] = live ? static_cast<int64_t>(ii) * kk + e % kk : 0;
    grow[u] = vbase + r0 + ii;
    gq[u] = live ? e % kk : 0;
    glive[u] = live && r0 + ii >= nc;
  }
  int2 pair[4];
  if (a.gathers) gather_pairs(a, grow, glive, pair);
  add_partials(a.partials, static_cast<int64_t>(kTile) * kk, first, count, off, sum);
  if (a.gathers) gather_add(a, grow, gq, pair, base);
#pragma unroll
  for (int u = 0; u < 4; ++u) {
    const int32_t e = threadIdx.x + 256 * u;
    if (e >= rows * kk) continue;
    const int32_t ii = lo + e / kk, q = e % kk;
    const int32_t i = r0 + ii;
    if (i < nc) a.y[static_cast<int64_t>(c0 + i) * kk + q] = sum[u];
    else a.fvec[(vbase + i) * kk + q] = base[u] + sum[u];
  }
}

// Backward tile product: inverse-factor columns [j0, j0+64) transposed, times
// g = [y(cols); d(rows)] over the inner range [k0, k1); each wave owns 16 of the columns. Applied
// results also give x_next = x + d when kIterate (a later round; the first round's d is the
// iterate itself, and its instance holds no iterate registers).
// The backward stage's depth. 32 (256 contiguous bytes of an inverse-factor column a stage) needs
// 120 registers, four waves a SIMD, and measured slower than 16 at six waves.
constexpr int kBwdInner = 16;

template <typename TW, bool kIterate>
__global__ __launch_bounds__(256, 6) void bwd_tile_kernel(
    const int4* __restrict__ desc, const TW* __restrict__ w, SweepArgs a) {
  constexpr int kD = kBwdInner, kPer = kD * kTile / 256;  // entries a thread stages per operand
  __shared__ double Ws[kD][kTile + kPad];  // [inner][inverse-factor column]
  __shared__ double Gs[kD][kCols + kPad];  // [inner][right-hand side]
  const Task tk = load_task(desc, xcd_block());
  const int32_t j0 = tk.tile;
  const int32_t k0 = tk.k0, k1 = tk.k1;
  const int32_t slot = tk.slot;
  const int32_t nc = tk.nc, c0 = tk.c0;
  const int32_t m = tk.m;
  const int32_t kk = a.kk;
  const TW* W = w + tk.w_off;
  const int32_t* srows = a.rows + tk.other;
  const int t = threadIdx.x, wave = t / 64, lane = t % 64;
  const bool cols_live = j0 + 16 * wave < nc;  // wave-uniform
  // Staging: entry e = t + 256u is inverse-factor column e / kD at inner e % kD (contiguous in
  // memory), and g's right-hand side e % 64 at inner e / 64.
  double wreg[kPer], greg[kPer];
  // A stage's structure-row indices are loaded a stage ahead (load_rows), so its W and g loads
  // issue together.
  int32_t srow[kPer];
  auto load_rows = [&](int32_t i0) {
#pragma unroll
    for (int u = 0; u < kPer; ++u) {
      const int32_t ig = i0 + (t + 256 * u) / kCols;
      srow[u] = (ig >= nc && ig < k1) ? srows[ig - nc] : 0;
    }
  };
  auto load = [&](int32_t i0) {
    // 32-bit element offsets from wave-uniform bases (a front's block and a field both stay
    // under 2^31 entries), one register an address.
#pragma unroll
    for (int u = 0; u < kPer; ++u) {
      const int e = t + 256 * u;
      const int ii = e % kD, jj = e / kD;
      const int32_t i = i0 + ii, j = j0 + jj;
      wreg[u] = (i < k1 && j < nc)
                    ? widen(W[static_cast<uint32_t>(j) * static_cast<uint32_t>(m) + i])
                    : 0.0;
    }
#pragma unroll
    for (int u = 0; u < kPer; ++u) {
      const int e = t + 256 * u;
      const int q = e % kCols;
      const int32_t ig = i0 + e / kCols;
      double g = 0.0;
      if (q < kk && ig < k1) {
        if (ig < nc) g = a.y[static_cast<uint32_t>(c0 + ig) * static_cast<uint32_t>(kk) + q];
        else g = a.d[static_cast<uint32_t>(srow[u]) * static_cast<uint32_t>(kk) + q];
      }
      greg[u] = g;
    }
  };
  f64x4 acc[4];
#pragma unroll
  for (int n = 0; n < 4; ++n) acc[n] = f64x4{0.0, 0.0, 0.0, 0.0};
  load_rows(k0);
  load(k0);
  if (k0 + kD < k1) load_rows(k0 + kD);
  for (int32_t i0 = k0; i0 < k1; i0 += kD) {
#pragma unroll
    for (int u = 0; u < kPer; ++u) {
      const int e = t + 256 * u;
      Ws[e % kD][e / kD] = wreg[u];
      Gs[e / kCols][e % kCols] = greg[u];
    }
    __syncthreads();
    if (i0 + kD < k1) {
      load(i0 + kD);
      if (i0 + 2 * kD < k1) load_rows(i0 + 2 * kD);
    }
    if (cols_live) stage_products_blocks<kD>(Ws, Gs, wave, lane, (kk + 15) / 16, acc);
    __syncthreads();
  }
  if (!cols_live) return;
  // Epilogue: the iterate's entries are all loaded before any store (see the forward tile).
  double xv[4][4] = {};
  if (kIterate && slot < 0) {
#pragma unroll
    for (int n = 0; n < 4; ++n) {
      const int q = 16 * n + lane % 16;
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const int32_t j = j0 + 16 * wave + lane / 16 + 4 * v;
        if (q < kk && j < nc) xv[n][v] = a.x[static_cast<int64_t>(c0 + j) * kk + q];
      }
    }
  }
#pragma unroll
  for (int n = 0; n < 4; ++n) {
    const int q = 16 * n + lane % 16;
    if (q >= kk) continue;
#pragma unroll
    for (int v = 0; v < 4; ++v) {
      const int32_t jj = 16 * wave + lane / 16 + 4 * v;
      const int32_t j = j0 + jj;
      if (j >= nc) continue;
      const double value = acc[n][v];
      if (slot >= 0) {
        a.partials[(static_cast<int64_t>(slot) * kTile + jj) * kk + q] = value;
      } else {
        const int64_t at = static_cast<int64_t>(c0 + j) * kk + q;
        a.d[at] = value;
        if (kIterate) a.x_next[at] = xv[n][v] + value;
      }
    }
  }
}

__global__ __launch_bounds__(256) void bwd_reduce_kernel(const int4* __restrict__ rdesc,
                                                         SweepArgs a) {
  const Task g = load_task(rdesc, xcd_block());
  const int32_t j0 = g.tile;
  const int32_t first = g.k0, count = g.k1;
  const int32_t nc = g.nc, c0 = g.c0;
  const int32_t kk = a.kk;
  constexpr int kRows = kTile / kReduceSplit;
  const int32_t lo = blockIdx.y * kRows;
  const int32_t cols = min(kRows, nc - j0 - lo);
  // Each thread's (at most four) entries: every load before any store.
  const bool iterate = a.x != nullptr;
  double sum[4] = {}, xv[4] = {};
  int64_t off[4];
#pragma unroll
  for (int u = 0; u < 4; ++u) {
    const int32_t e = threadIdx.x + 256 * u;
    off[u] = e < cols * kk ? static_cast<int64_t>(lo + e / kk) * kk + e % kk : 0;
    if (iterate && e < cols * kk)
      xv[u] = a.x[static_cast<int64_t>(c0 + j0 + lo + e / kk) * kk + e % kk];
  }
  add_partials(a.partials, static_cast<int64_t>(kTile) * kk, first, count, off, sum);
#pragma unroll
  for (int u = 0; u < 4; ++u) {
    const int32_t e = threadIdx.x + 256 * u;
    if (e >= cols * kk) continue;
    const int32_t jj = lo + e / kk, q = e % kk;
    const int64_t at = static_cast<int64_t>(c0 + j0 + jj) * kk + q;
    a.d[at] = sum[u];
    if (iterate) a.x_next[at] = xv[u] + sum[u];
  }
}

// One right-hand side: f gathered into LDS 256 entries at a time, then four lanes a row split the
// inner range (j = lane mod 4), folded in lane order.
template <typename TW>
__global__ __launch_bounds__(256) void fwd_tile1_kernel(
    const int32_t* __restrict__ task_s, const int32_t* __restrict__ task_r0,
    const int32_t* __restrict__ task_k0, const int32_t* __restrict__ task_k1,
    const int32_t* __restrict__ task_slot, const TW* __restrict__ w, SweepArgs a) {
  __shared__ double Fs[256];
  __shared__ double part[4][kTile];
  const int32_t task = xcd_block();
  const int32_t s = task_s[task];
  const int32_t r0 = task_r0[task];
  const int32_t k0 = task_k0[task], k1 = task_k1[task];
  const int32_t slot = task_slot[task];
  const int32_t nc = a.ncol[s], c0 = a.col_begin[s];
  const int32_t m = nc + a.nrow[s];
  const TW* W = w + a.w_off[s];
  const int64_t vbase = a.vec_off[s];
  const int ii = threadIdx.x % kTile, lane = threadIdx.x / kTile;
  const int32_t i = r0 + ii;
  double acc = 0.0;
  for (int32_t j0 = k0; j0 < k1; j0 += 256) {
    const int32_t j = j0 + threadIdx.x;
    Fs[threadIdx.x] =
        j < k1 ? (a.gathers ? gather1(a, vbase + j, 0, a.r[c0 + j]) : a.r[c0 + j]) : 0.0;
    __syncthreads();
    const int32_t jend = min(k1 - j0, 256);
    if (i < m)
      for (int32_t jj = lane; jj < jend; jj += 4)
        acc += widen(W[static_cast<int64_t>(j0 + jj) * m + i]) * Fs[jj];
    __syncthreads();
  }
  part[lane][ii] = acc;
  __syncthreads();
  if (lane == 0 && i < m) {
    const double sum = ((part[0][ii] + part[1][ii]) + part[2][ii]) + part[3][ii];
    if (slot >= 0) {
      a.partials[static_cast<int64_t>(slot) * kTile + ii] = sum;
    } else if (i < nc) {
      a.y[c0 + i] = sum;
    } else {
      const int64_t row = vbase + i;
      a.fvec[row] =
          (a.gathers ? gather1(a, row, 0, 0.0) : 0.0) + sum;
    }
  }
}

template <typename TW>
__global__ __launch_bounds__(256) void bwd_tile1_kernel(
    const int32_t* __restrict__ task_s, const int32_t* __restrict__ task_j0,
    const int32_t* __restrict__ task_k0, const int32_t* __restrict__ task_k1,
    const int32_t* __restrict__ task_slot, const TW* __restrict__ w, SweepArgs a) {
  constexpr int kStage = 1024;  // entries of g staged at a time
  __shared__ double Gs[kStage];
  const int32_t task = xcd_block();
  const int32_t s = task_s[task];
  const int32_t j0 = task_j0[task];
  const int32_t k0 = task_k0[task], k1 = task_k1[task];
  const int32_t slot = task_slot[task];
  const int32_t nc = a.ncol[s], c0 = a.col_begin[s];
  const int32_t m = nc + a.nrow[s];
  const TW* W = w + a.w_off[s];
  const int32_t* srows = a.rows + a.row_ptr[s];
  // g = [y(cols); d(rows)] over the inner range is staged in LDS once (its structure-row gathers
  // off the products' path); then sixteen lanes a solve-block column, each group of sixteen
  // reading 16 consecutive entries of its column (contiguous in memory) a step. The 64 columns go
  // in four passes of sixteen; each column's sixteen lane sums fold in a fixed butterfly.
  const int sub = threadIdx.x % 16, grp = threadIdx.x / 16;
  double acc[4] = {0.0, 0.0, 0.0, 0.0};
  for (int32_t s0 = k0; s0 < k1; s0 += kStage) {
    const int32_t s1 = min(k1, s0 + kStage);
    for (int32_t i = s0 + static_cast<int32_t>(threadIdx.x); i < s1; i += 256)
      Gs[i - s0] = i < nc ? a.y[c0 + i] : a.d[srows[i - nc]];
    __syncthreads();
#pragma unroll
    for (int p = 0; p < 4; ++p) {
      const int32_t j = j0 + grp + 16 * p;
      if (j >= nc) continue;
      const TW* wj = W + static_cast<int64_t>(j) * m;
      // Rows above the diagonal of the inverse's leading block are zero: start at j's group.
      const int32_t first = s0 + max(0, (j - s0) / 16 * 16);
      for (int32_t i = first + sub; i < s1; i += 16) acc[p] += widen(wj[i]) * Gs[i - s0];
    }
    __syncthreads();
  }
#pragma unroll
  for (int p = 0; p < 4; ++p)
#pragma unroll
    for (int off = 8; off >= 1; off >>= 1) acc[p] += __shfl_xor(acc[p], off, 16);
  if (sub != 0) return;
#pragma unroll
  for (int p = 0; p < 4; ++p) {
    const int32_t jj = grp + 16 * p, j = j0 + jj;
    if (j >= nc) continue;
    if (slot >= 0) {
      a.partials[static_cast<int64_t>(slot) * kTile + jj] = acc[p];
    } else {
      a.d[c0 + j] = acc[p];
      if (a.x) a.x_next[c0 + j] = a.x[c0 + j] + acc[p];
    }
  }
}

// Raises *flag once per block that fails, and not at all once the flag reads set: with every row
// failing in a solve's first rounds, an atomic a wave on one word serializes the whole launch.
// Block-collective.
// Lane `lane`'s double, as a wave-uniform value.
__device__ __forceinline__ double readlane_f64(double v, int lane) {
  const uint64_t bits = __builtin_bit_cast(uint64_t, v);
  const uint32_t lo = __builtin_amdgcn_readlane(static_cast<uint32_t>(bits), lane);
  const uint32_t hi = __builtin_amdgcn_readlane(static_cast<uint32_t>(bits >> 32), lane);
  return __builtin_bit_cast(double, (static_cast<uint64_t>(hi) << 32) | lo);
}

__device__ __forceinline__ void raise_flag(bool fail, int32_t* flag) {
  const int any = __syncthreads_or(fail ? 1 : 0);
  if (any && threadIdx.x == 0 &&
      __hip_atomic_load(flag, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) == 0)
    atomicOr(flag, 1);
}

// r = b - A x in the elimination order and whether any row is still above tolerance
// (componentwise against (|A||x| + |b|)_i, normwise against the larger of that and the column's
// largest |b|; kernels/cholesky.py:1393-1440). Pinned rows are dropped. One thread a (row,
// right-hand side); the row's operator entries are shared by its kk threads.
// kWaveRow: kk is the full 64, so each wave holds one row and the row's offsets, columns and
// values are wave-uniform (read once per wave through the scalar unit).
template <bool kWaveRow>
__global__ void residual_kernel(const int32_t* __restrict__ offsets,
                                const int32_t* __restrict__ columns,
                                const double* __restrict__ values, const double* __restrict__ b,
                                const double* __restrict__ x, const int32_t* __restrict__ pinned,
                                const double* __restrict__ bmax, int64_t n, int32_t kk,
                                double tolerance, int32_t componentwise, double* __restrict__ r,
                                int32_t* __restrict__ pending) {
  int64_t t = static_cast<int64_t>(xcd_block()) * blockDim.x + threadIdx.x;
  bool fail = false;
  if (t < n * kk) {
    const int64_t i =
        kWaveRow ? __builtin_amdgcn_readfirstlane(static_cast<int32_t>(t / kCols)) : t / kk;
    const int64_t q = t % kk;
    const double bi = b[t];
    const bool dropped = pinned && pinned[i];
    const double bscale = componentwise ? 0.0 : bmax[q];
    double total = bi, scale = fabs(bi);
    // Eight entries at a time (a row of a triangle mesh has about seven): their column and value
    // loads, then their iterate loads, issue together; entries past the row's end repeat its last
    // one and are not added, and the sums stay in entry order. A wave on one row loads the eight
    // columns and values in its first lanes at once and reads them out lane by lane.
    const int32_t e0 = offsets[i], e1 = offsets[i + 1];
    for (int32_t e = e0; e < e1; e += 8) {
      int32_t c[8];
      double v[8], xv[8];
      if (kWaveRow) {
        const int32_t at = min(e + static_cast<int32_t>(threadIdx.x % 8), e1 - 1);
        const int32_t lane_c = columns[at];
        const double lane_v = values[at];
#pragma unroll
        for (int u = 0; u < 8; ++u) {
          c[u] = __builtin_amdgcn_readlane(lane_c, u);
          v[u] = readlane_f64(lane_v, u);
        }
      } else {
#pragma unroll
        for (int u = 0; u < 8; ++u) {
          const int32_t at = min(e + u, e1 - 1);
          c[u] = columns[at];
          v[u] = values[at];
        }
      }
#pragma unroll
      for (int u = 0; u < 8; ++u) xv[u] = x[static_cast<int64_t>(c[u]) * kk + q];
#pragma unroll
      for (int u = 0; u < 8; ++u) {
        const double term = e + u < e1 ? v[u] * xv[u] : 0.0;  // subtracting 0 changes nothing
        total -= term;
        scale += fabs(term);
      }
    }
    if (r) r[t] = dropped ? 0.0 : total;
    if (!componentwise) scale = fmax(scale, bscale);
    fail = !dropped && !(fabs(total) <= tolerance * scale);
  }
  raise_flag(fail, pending);
}

// The residual of residual_kernel for 64 right-hand sides, kRowsPerWave rows a wave: a lane is a
// right-hand side of each of its wave's rows. Every round of loads covers all the wave's rows
// (their offsets; then eight entries of each, columns and values loaded lane-parallel and read out
// lane by lane; then every iterate entry they name), so a wave keeps several rows' loads in flight
// where one row would wait on three dependent rounds alone. Sums keep entry order.
constexpr int kRowsPerWave = 4;

__global__ __launch_bounds__(256) void residual_rows_kernel(
    const int32_t* __restrict__ offsets, const int32_t* __restrict__ columns,
    const double* __restrict__ values, const double* __restrict__ b,
    const double* __restrict__ x, const int32_t* __restrict__ pinned,
    const double* __restrict__ bmax, int64_t n, double tolerance, int32_t componentwise,
    double* __restrict__ r, int32_t* __restrict__ pending) {
  constexpr int R = kRowsPerWave;
  const int32_t lane = static_cast<int32_t>(threadIdx.x % 64);
  const int64_t first = static_cast<int64_t>(__builtin_amdgcn_readfirstlane(
                            static_cast<int32_t>(xcd_block() * 4 + threadIdx.x / 64))) * R;
  const double bscale = componentwise ? 0.0 : bmax[lane];
  int32_t e0[R], e1[R];
  double bi[R], total[R], scale[R];
  bool dropped[R];
#pragma unroll
  for (int u = 0; u < R; ++u) {
    const int64_t i = first + u;
    const bool live = i < n;
    e0[u] = live ? offsets[i] : 0;
    e1[u] = live ? offsets[i + 1] : 0;
    bi[u] = live ? b[i * kCols + lane] : 0.0;
    dropped[u] = !live || (pinned && pinned[i]);
    total[u] = bi[u];
    scale[u] = fabs(bi[u]);
  }
  for (int32_t base = 0;; base += 8) {
    bool more = false;
#pragma unroll
    for (int u = 0; u < R; ++u) more |= e0[u] + base < e1[u];
    if (!more) break;  // wave-uniform
    // Lane 8u + j loads entry base + j of row u (an entry past the row's end, or of an empty
    // row, reads one in range and is not added).
    const int u_lane = min(lane / 8, R - 1);
    int32_t at = e0[0];
#pragma unroll
    for (int u = 0; u < R; ++u)
      if (u == u_lane) at = e1[u] > e0[u] ? min(e0[u] + base + lane % 8, e1[u] - 1) : 0;
    const int32_t lane_c = columns[at];
    const double lane_v = values[at];
    double xv[R][8];
#pragma unroll
    for (int u = 0; u < R; ++u)
#pragma unroll
      for (int j = 0; j < 8; ++j) {
        const int32_t c = __builtin_amdgcn_readlane(lane_c, 8 * u + j);
        xv[u][j] = x[static_cast<int64_t>(c) * kCols + lane];
      }
#pragma unroll
    for (int u = 0; u < R; ++u)
#pragma unroll
      for (int j = 0; j < 8; ++j) {
        const double v = readlane_f64(lane_v, 8 * u + j);
        const double term = e0[u] + base + j < e1[u] ? v * xv[u][j] : 0.0;
        total[u] -= term;
        scale[u] += fabs(term);
      }
  }
  bool fail = false;
#pragma unroll
  for (int u = 0; u < R; ++u) {
    const int64_t i = first + u;
    if (i >= n) continue;
    if (r) r[i * kCols + lane] = dropped[u] ? 0.0 : total[u];
    const double sc = componentwise ? scale[u] : fmax(scale[u], bscale);
    fail |= !dropped[u] && !(fabs(total[u]) <= tolerance * sc);
  }
  raise_flag(fail, pending);
}

// Per-column partial maxima of |b| (user layout), one block a (chunk, column).
__global__ void absmax_partial_kernel(const double* __restrict__ b, int64_t n,
                                      double* __restrict__ partials) {
  __shared__ double scratch[256];
  const int64_t q = blockIdx.y;
  const int64_t i = static_cast<int64_t>(blockIdx.x) * 256 + threadIdx.x;
  scratch[threadIdx.x] = i < n ? fabs(b[q * n + i]) : 0.0;
  __syncthreads();
  for (int stride = 128; stride > 0; stride >>= 1) {
    if (threadIdx.x < stride) scratch[threadIdx.x] = fmax(scratch[threadIdx.x], scratch[threadIdx.x + stride]);
    __syncthreads();
  }
  if (threadIdx.x == 0) partials[q * gridDim.x + blockIdx.x] = scratch[0];
}

// The same partial maxima for a chunk stored by rows: block b covers rows [256 b, 256 b + 256);
// thread (q, part) = (t % 64, t / 64) folds every fourth of them for column q. A maximum does not
// depend on the order it is taken in.
__global__ void absmax_rows_partial_kernel(const double* __restrict__ b, int64_t n, int32_t kk,
                                           double* __restrict__ partials) {
  __shared__ double scratch[4][64];
  __shared__ unsigned long long best[kCols];
  const int q = threadIdx.x % 64, part = threadIdx.x / 64;
  const int64_t i0 = static_cast<int64_t>(blockIdx.x) * 256;
  if (kk < kCols / 4) {
    // Few columns: the block's rows are read as one contiguous run, every lane busy, and each
    // column's maximum kept by an LDS atomic on the bits (for non-negative doubles their order is
    // the values' order, and a maximum does not depend on the order it is taken in).
    if (threadIdx.x < kCols) best[threadIdx.x] = 0ull;
    __syncthreads();
    const int64_t count = min<int64_t>(n - i0, 256) * kk;
    const double* run = b + i0 * kk;
    for (int64_t e = threadIdx.x; e < count; e += 256)
      atomicMax(&best[e % kk], static_cast<unsigned long long>(
                                   __builtin_bit_cast(uint64_t, fabs(run[e]))));
    __syncthreads();
    if (threadIdx.x < kk)
      partials[static_cast<int64_t>(threadIdx.x) * gridDim.x + blockIdx.x] =
          __builtin_bit_cast(double, static_cast<uint64_t>(best[threadIdx.x]));
    return;
  }
  double v = 0.0;
  if (q < kk)
    for (int64_t i = i0 + part; i < min<int64_t>(n, i0 + 256); i += 4) v = fmax(v, fabs(b[i * kk + q]));
  scratch[part][q] = v;
  __syncthreads();
  if (part == 0 && q < kk)
    partials[static_cast<int64_t>(q) * gridDim.x + blockIdx.x] =
        fmax(fmax(scratch[0][q], scratch[1][q]), fmax(scratch[2][q], scratch[3][q]));
}

__global__ void absmax_final_kernel(const double* __restrict__ partials, int32_t blocks,
                                    double* __restrict__ out) {
  __shared__ double scratch[256];
  const int64_t q = blockIdx.x;
  double v = 0.0;
  for (int32_t b = threadIdx.x; b < blocks; b += 256) v = fmax(v, partials[q * blocks + b]);
  scratch[threadIdx.x] = v;
  __syncthreads();
  for (int stride = 128; stride > 0; stride >>= 1) {
    if (threadIdx.x < stride) scratch[threadIdx.x] = fmax(scratch[threadIdx.x], scratch[threadIdx.x + stride]);
    __syncthreads();
  }
  if (threadIdx.x == 0) out[q] = scratch[0];
}

// Component sums for the null-space projection: block partials per (component chunk, column).
__global__ void component_partial_kernel(const int32_t* __restrict__ chunk_comp,
                                         const int32_t* __restrict__ chunk_first,
                                         const int32_t* __restrict__ comp_ptr,
                                         const int32_t* __restrict__ members, int64_t n,
                                         const double* __restrict__ x,
                                         double* __restrict__ partials) {
  __shared__ double scratch[256];
  const int32_t chunk = blockIdx.x;
  const int64_t q = blockIdx.y;
  const int32_t c = chunk_comp[chunk];
  const int32_t e = chunk_first[chunk] + threadIdx.x;
  scratch[threadIdx.x] = e < comp_ptr[c + 1] ? x[q * n + members[e]] : 0.0;
  __syncthreads();
  for (int stride = 128; stride > 0; stride >>= 1) {
    if (threadIdx.x < stride) scratch[threadIdx.x] += scratch[threadIdx.x + stride];
    __syncthreads();
  }
  if (threadIdx.x == 0) partials[q * gridDim.x + chunk] = scratch[0];
}

// Each (component, column)'s mean, from its chunks' partials in chunk order.
__global__ void component_mean_kernel(const int32_t* __restrict__ comp_chunk_ptr,
                                      const int32_t* __restrict__ comp_ptr, int32_t n_chunks,
                                      const double* __restrict__ partials,
                                      double* __restrict__ means) {
  __shared__ double scratch[256];
  const int32_t c = blockIdx.x;
  const int64_t q = blockIdx.y;
  double sum = 0.0;
  for (int32_t k = comp_chunk_ptr[c] + threadIdx.x; k < comp_chunk_ptr[c + 1]; k += 256)
    sum += partials[q * n_chunks + k];
  scratch[threadIdx.x] = sum;
  __syncthreads();
  for (int stride = 128; stride > 0; stride >>= 1) {
    if (threadIdx.x < stride) scratch[threadIdx.x] += scratch[threadIdx.x + stride];
    __syncthreads();
  }
  if (threadIdx.x == 0)
    means[q * gridDim.x + c] = scratch[0] / static_cast<double>(comp_ptr[c + 1] - comp_ptr[c]);
}

__global__ void component_shift_kernel(const int32_t* __restrict__ chunk_comp,
                                       const int32_t* __restrict__ chunk_first,
                                       const int32_t* __restrict__ comp_ptr,
                                       const int32_t* __restrict__ members, int32_t n_comp,
                                       int64_t n, const double* __restrict__ means,
                                       double* __restrict__ x) {
  const int32_t chunk = blockIdx.x;
  const int64_t q = blockIdx.y;
  const int32_t c = chunk_comp[chunk];
  const int32_t e = chunk_first[chunk] + threadIdx.x;
  if (e < comp_ptr[c + 1]) x[q * n + members[e]] -= means[q * n_comp + c];
}

// ---- the distance from a Poisson solution stored by rows ------------------------------------
// Every component's rows are contiguous in the elimination order (each component is dissected on
// its own and numbered in component order), so the null-space chunks (256 members of one
// component) are row ranges too: chunk k covers rows [chunk_first[k], min(+256, comp end)).

// Column sums over each chunk: thread (q, part) adds rows part, part + 4, ... (four loads issued
// before their adds), then the four parts add in order. partials[q * n_chunks + k], the layout
// component_mean_kernel reads.
__global__ __launch_bounds__(256) void rows_partial_kernel(const int32_t* __restrict__ chunk_comp,
                                                           const int32_t* __restrict__ chunk_first,
                                                           const int32_t* __restrict__ comp_ptr,
                                                           const double* __restrict__ x,
                                                           int32_t kk,
                                                           double* __restrict__ partials) {
  __shared__ double part_sum[4][kCols];
  const int32_t chunk = blockIdx.x;
  const int32_t first = chunk_first[chunk];
  const int32_t end = min(first + 256, comp_ptr[chunk_comp[chunk] + 1]);
  const int q = threadIdx.x % kCols, part = threadIdx.x / kCols;
  double sum = 0.0;
  if (q < kk) {
    int32_t r = first + part;
    for (; r + 12 < end; r += 16) {
      double v[4];
#pragma unroll
      for (int u = 0; u < 4; ++u) v[u] = x[static_cast<int64_t>(r + 4 * u) * kk + q];
#pragma unroll
      for (int u = 0; u < 4; ++u) sum += v[u];
    }
    for (; r < end; r += 4) sum += x[static_cast<int64_t>(r) * kk + q];
  }
  part_sum[part][q] = sum;
  __syncthreads();
  if (part == 0)
    partials[static_cast<int64_t>(q) * gridDim.x + chunk] =
        ((part_sum[0][q] + part_sum[1][q]) + part_sum[2][q]) + part_sum[3][q];
}

// One block per column q, after each component's mean (component_mean_kernel,
// kernels/cholesky.py:1467-1660): the shift of the mean-removed field (kernels/heat.py:3383-3424),
// offset = its mean over the column's sources (vertex v at row iperm[v]), negated when its mean
// over every vertex (the chunks' sums less their share of the means) is below the offset.
// shift[2q] = offset, shift[2q + 1] = 1 to negate.
__global__ __launch_bounds__(256) void rows_decide_kernel(
    const int32_t* __restrict__ comp_ptr, const int32_t* __restrict__ chunk_comp,
    const int32_t* __restrict__ chunk_first, int32_t n_comp, int32_t n_chunks,
    const double* __restrict__ partials, const double* __restrict__ means,
    const double* __restrict__ x, int32_t kk, int64_t n, const int32_t* __restrict__ iperm,
    const int32_t* __restrict__ labels, const int32_t* __restrict__ sources,
    const int64_t* __restrict__ source_offsets, double* __restrict__ shift) {
  __shared__ double s_total[256];
  __shared__ double s_src[256];
  const int q = blockIdx.x, t = threadIdx.x;
  const double* mean = means + static_cast<int64_t>(q) * n_comp;
  double total = 0.0;
  for (int32_t k = t; k < n_chunks; k += 256) {
    const int32_t c = chunk_comp[k];
    const int32_t len = min(256, comp_ptr[c + 1] - chunk_first[k]);
    total += partials[static_cast<int64_t>(q) * n_chunks + k] - static_cast<double>(len) * mean[c];
  }
  const int64_t first = source_offsets[q], last = source_offsets[q + 1];
  double src = 0.0;
  for (int64_t e = first + t; e < last; e += 256) {
    const int32_t v = sources[e];
    if (v < 0 || v >= n) continue;  // out of range: the call fails once the stream is synced
    src += x[static_cast<int64_t>(iperm[v]) * kk + q] - mean[labels[v]];
  }
  s_total[t] = total;
  s_src[t] = src;
  __syncthreads();
  for (int stride = 128; stride > 0; stride >>= 1) {
    if (t < stride) {
      s_total[t] += s_total[t + stride];
      s_src[t] += s_src[t + stride];
    }
    __syncthreads();
  }
  if (t == 0) {
    const int64_t count = last - first;
    const double offset = count > 0 ? s_src[0] / static_cast<double>(count) : 0.0;
    shift[2 * q] = offset;
    shift[2 * q + 1] =
        (count > 0 && s_total[0] / static_cast<double>(n) - offset < 0.0) ? 1.0 : 0.0;
  }
}

// The distance in the caller's layout (column q at q * n): vertex v's entry is the mean-removed
// value at row iperm[v], shifted and oriented. 64 vertices a block, transposed through LDS so the
// row reads and the column writes are both contiguous.
__global__ __launch_bounds__(256) void rows_out_kernel(const double* __restrict__ x, int32_t kk,
                                                       int64_t n, const int32_t* __restrict__ iperm,
                                                       const int32_t* __restrict__ labels,
                                                       int32_t n_comp,
                                                       const double* __restrict__ means,
                                                       const double* __restrict__ shift,
                                                       double* const* __restrict__ out_slot) {
  __shared__ double tile[kTile][kCols + 1];
  double* __restrict__ out = *out_slot;
  const int64_t v0 = static_cast<int64_t>(blockIdx.x) * kTile;
  const int32_t rows = static_cast<int32_t>(min<int64_t>(kTile, n - v0));
  if (kk == kCols) {
    // A lane is a column q; wave w reads rows w, w + 4, ... (each a contiguous row of x), all
    // sixteen rows' loads issued before any is used.
    const int q = threadIdx.x % kCols, w = threadIdx.x / kCols;
    const double offset = shift[2 * q];
    const bool flip = shift[2 * q + 1] != 0.0;
    double xv[kTile / 4], mv[kTile / 4];
#pragma unroll
    for (int j = 0; j < kTile / 4; ++j) {
      const int lv = w + 4 * j;
      const int64_t v = v0 + min(lv, rows - 1);
      xv[j] = x[static_cast<int64_t>(iperm[v]) * kCols + q];
      mv[j] = means[static_cast<int64_t>(q) * n_comp + labels[v]];
    }
#pragma unroll
    for (int j = 0; j < kTile / 4; ++j) {
      const double value = xv[j] - mv[j];
      tile[w + 4 * j][q] = flip ? offset - value : value - offset;
    }
  } else {
    for (int e = threadIdx.x; e < rows * kk; e += 256) {
      const int lv = e / kk, q = e % kk;
      const int64_t v = v0 + lv;
      const double value = x[static_cast<int64_t>(iperm[v]) * kk + q] -
                           means[static_cast<int64_t>(q) * n_comp + labels[v]];
      const double offset = shift[2 * q];
      tile[lv][q] = shift[2 * q + 1] != 0.0 ? offset - value : value - offset;
    }
  }
  __syncthreads();
  const int lv = threadIdx.x % kTile, qq = threadIdx.x / kTile;
  if (lv >= rows) return;
  for (int q = qq; q < kk; q += 4) out[static_cast<int64_t>(q) * n + v0 + lv] = tile[lv][q];
}

inline unsigned blocks_for(int64_t n) { return static_cast<unsigned>((n + 255) / 256); }

template <typename TW>
void sweep(const DevicePlan& plan, const TW* w, SweepArgs a, hipStream_t stream) {
  const DevicePlan::SweepTasks& F = plan.fwd;
  const DevicePlan::SweepTasks& B = plan.bwd;
  for (int32_t level = 0; level < plan.n_levels; ++level) {
    const int32_t t0 = F.level_ptr[level];
    const int32_t tasks = F.level_ptr[level + 1] - t0;
    if (tasks == 0) continue;
    a.gathers = plan.level_gathers[level];
    if (a.kk == 1) {
      fwd_tile1_kernel<TW><<<tasks, 256, 0, stream>>>(F.s.ptr + t0, F.tile.ptr + t0,
                                                      F.k0.ptr + t0, F.k1.ptr + t0,
                                                      F.slot.ptr + t0, w, a);
    } else {
      auto tile = a.gathers ? fwd_tile_kernel<TW, true> : fwd_tile_kernel<TW, false>;
      tile<<<tasks, 256, 0, stream>>>(F.desc.ptr + 3 * static_cast<int64_t>(t0), w, a);
    }
    const int32_t r0 = F.reduce_ptr[level];
    const int32_t groups = F.reduce_ptr[level + 1] - r0;
    if (groups > 0) {
      fwd_reduce_kernel<<<dim3(groups, kReduceSplit), 256, 0, stream>>>(
          F.rdesc.ptr + 3 * static_cast<int64_t>(r0), a);
    }
  }
  a.gathers = 0;
  for (int32_t level = plan.n_levels - 1; level >= 0; --level) {
    const int32_t t0 = B.level_ptr[level];
    const int32_t tasks = B.level_ptr[level + 1] - t0;
    if (tasks == 0) continue;
    if (a.kk == 1) {
      bwd_tile1_kernel<TW><<<tasks, 256, 0, stream>>>(B.s.ptr + t0, B.tile.ptr + t0,
                                                      B.k0.ptr + t0, B.k1.ptr + t0,
                                                      B.slot.ptr + t0, w, a);
    } else {
      auto tile = a.x ? bwd_tile_kernel<TW, true> : bwd_tile_kernel<TW, false>;
      tile<<<tasks, 256, 0, stream>>>(B.desc.ptr + 3 * static_cast<int64_t>(t0), w, a);
    }
    const int32_t r0 = B.reduce_ptr[level];
    const int32_t groups = B.reduce_ptr[level + 1] - r0;
    if (groups > 0) {
      bwd_reduce_kernel<<<dim3(groups, kReduceSplit), 256, 0, stream>>>(
          B.rdesc.ptr + 3 * static_cast<int64_t>(r0), a);
    }
  }
  HG_HIP(hipGetLastError());
}

}  // namespace

void SolveWorkspace::reserve(const DevicePlan& plan, int64_t kk_needed) {
  // Sized for the next power of two of columns: a caller whose batches vary in width reallocates
  // (a hipFree waits on every stream's work) a handful of times, not once per new width.
  int64_t kk = 1;
  while (kk < kk_needed) kk *= 2;
  kk = std::min<int64_t>(kk, kCols);
  const int64_t entries = plan.n * kk;
  if (capacity < entries) {
    bp.allocate(entries);
    xa.allocate(entries);
    xb.allocate(entries);
    r.allocate(entries);
    y.allocate(entries);
    d.allocate(entries);
    capacity = entries;
  }
  if (fcap < plan.vec_size * kk) {
    fvec.allocate(std::max<int64_t>(plan.vec_size * kk, 1));
    fcap = plan.vec_size * kk;
  }
  const int64_t pentries = static_cast<int64_t>(plan.max_slots) * kTile * kk;
  if (pcap < pentries) {
    partials.allocate(std::max<int64_t>(pentries, 1));
    pcap = pentries;
  }
  const int64_t absmax = static_cast<int64_t>(blocks_for(plan.n)) * kCols + kCols;
  if (bmax.size < absmax) bmax.allocate(absmax);
  if (pending.size < 1) pending.allocate(1);
}

int32_t refined_solve(const DevicePlan& plan, const Factor& factor, bool use_fp32,
                      double* b, double* x, int64_t k, double tolerance, bool componentwise,
                      int32_t max_rounds, const int32_t* pinned_new, SolveWorkspace& ws,
                      hipStream_t stream, PhaseTimer* timer, int sweep_phase,
                      FieldLayout b_layout, FieldLayout x_layout, int32_t* speculate) {
  const int64_t n = plan.n;
  if (n == 0 || k == 0) return 0;
  int32_t max_used = 0;
  if (use_fp32) speculate = nullptr;
  for (int64_t q0 = 0; q0 < k; q0 += kCols) {
    const int32_t kk = static_cast<int32_t>(std::min<int64_t>(kCols, k - q0));
    // The backward tiles address a chunk's fields with 32-bit element offsets.
    HG_REQUIRE(n * kk < (int64_t{1} << 31), HG_ERR_VALUE,
               "a solve chunk of " + std::to_string(n) + " rows x " + std::to_string(kk) +
                   " columns exceeds 2^31 entries; solve fewer columns per call");
    ws.reserve(plan, kk);
    double* bq = b + q0 * n;
    double* xq = x + q0 * n;
    const unsigned grid = blocks_for(n * kk);
    const unsigned row_tiles = static_cast<unsigned>((n + kTile - 1) / kTile);
    // The right-hand side in the elimination order with its pinned rows zero: b itself by rows
    // (zeroed in place below, after the normwise scale has read it), else permuted into ws.bp.
    double* bp = b_layout == FieldLayout::kRows ? bq : ws.bp.ptr;
    if (b_layout == FieldLayout::kColumns)
      perm_in_kernel<<<row_tiles, 256, 0, stream>>>(bq, plan.perm.ptr, pinned_new, n, kk, bp);
    double* bmax = ws.bmax.ptr + static_cast<int64_t>(blocks_for(n)) * kCols;
    if (!componentwise) {
      const unsigned blocks = blocks_for(n);
      if (b_layout == FieldLayout::kRows)
        absmax_rows_partial_kernel<<<blocks, 256, 0, stream>>>(bq, n, kk, ws.bmax.ptr);
      else
        absmax_partial_kernel<<<dim3(blocks, kk), 256, 0, stream>>>(bq, n, ws.bmax.ptr);
      absmax_final_kernel<<<kk, 256, 0, stream>>>(ws.bmax.ptr, static_cast<int32_t>(blocks), bmax);
    }
    if (b_layout == FieldLayout::kRows && pinned_new)
      zero_pinned_rows_kernel<<<blocks_for(n), 256, 0, stream>>>(pinned_new, n, kk, bp);
    SweepArgs a{};
    a.col_begin = plan.col_begin.ptr;
    a.ncol = plan.ncol.ptr;
    a.nrow = plan.nrow.ptr;
    a.row_ptr = plan.row_ptr.ptr;
    a.rows = plan.rows.ptr;
    a.w_off = plan.w_off.ptr;
    a.vec_off = plan.vec_off.ptr;
    a.gather_pair = plan.gather_pair.ptr;
    a.gather_ptr = plan.gather_ptr.ptr;
    a.gather_src = plan.gather_src.ptr;
    a.fvec = ws.fvec.ptr;
    a.y = ws.y.ptr;
    a.partials = ws.partials.ptr;
    a.kk = kk;
    // The first round solves for the iterate itself from x = 0: its right-hand side is b and its
    // correction is written straight into the iterate. Later rounds correct it: d into ws.d and
    // x + d into the other iterate buffer. By rows the iterate lives in x itself (and ws.xa).
    double* const x_home = x_layout == FieldLayout::kRows ? xq : ws.xb.ptr;
    double* x_cur = x_home;
    int32_t rounds = 0;
    while (rounds < max_rounds) {
      double* x_next = x_cur == x_home ? ws.xa.ptr : x_home;
      if (rounds == 0) {
        a.r = bp;
        a.x = nullptr;
        a.x_next = nullptr;
        a.d = x_cur;
      } else {
        a.r = ws.r.ptr;
        a.x = x_cur;
        a.x_next = x_next;
        a.d = ws.d.ptr;
      }
      {
        PhaseScope scope(timer, sweep_phase, stream);
        if (use_fp32) sweep<float>(plan, factor.w32.ptr, a, stream);
        else sweep<double>(plan, factor.w64.ptr, a, stream);
      }
      if (rounds > 0) x_cur = x_next;
      ++rounds;
      // Through FP64 factors the first round almost always passes, so its residual is formed for
      // the test alone and written (a second pass) only when another round follows.
      const bool test_only = !use_fp32 && rounds == 1;
      auto residual = [&](double* r_out) {
        if (kk == kCols)
          residual_rows_kernel<<<static_cast<unsigned>((n + 4 * kRowsPerWave - 1) /
                                                       (4 * kRowsPerWave)),
                                 256, 0, stream>>>(plan.p_offsets.ptr, plan.p_columns.ptr,
                                                   factor.pvalues.ptr, bp, x_cur, pinned_new,
                                                   bmax, n, tolerance, componentwise ? 1 : 0,
                                                   r_out, ws.pending.ptr);
        else
          residual_kernel<false><<<grid, 256, 0, stream>>>(
              plan.p_offsets.ptr, plan.p_columns.ptr, factor.pvalues.ptr, bp, x_cur, pinned_new,
              bmax, n, kk, tolerance, componentwise ? 1 : 0, r_out, ws.pending.ptr);
      };
      {
        PhaseScope scope(timer, kPhaseResidual, stream);
        launch_zero(ws.pending.ptr, sizeof(int32_t), stream);
        residual(test_only ? nullptr : ws.r.ptr);
      }
      HG_HIP(hipGetLastError());
      if (speculate) {
        HG_HIP(hipMemcpyAsync(speculate + q0 / kCols, ws.pending.ptr, sizeof(int32_t),
                              hipMemcpyDeviceToHost, stream));
        break;
      }
      int32_t pending = 0;
      HG_HIP(hipMemcpyAsync(&pending, ws.pending.ptr, sizeof(int32_t), hipMemcpyDeviceToHost,
                            stream));
      HG_HIP(hipStreamSynchronize(stream));
      if (timer) timer->resolve();
      if (!pending) break;
      if (test_only && rounds < max_rounds) {
        PhaseScope scope(timer, kPhaseResidual, stream);
        residual(ws.r.ptr);
        HG_HIP(hipGetLastError());
      }
    }
    if (x_layout == FieldLayout::kColumns)
      perm_out_kernel<<<row_tiles, 256, 0, stream>>>(x_cur, plan.perm.ptr, n, kk, xq);
    else if (x_cur != xq)
      HG_HIP(hipMemcpyAsync(xq, x_cur, sizeof(double) * n * kk, hipMemcpyDeviceToDevice, stream));
    HG_HIP(hipGetLastError());
    max_used = std::max(max_used, rounds);
  }
  return max_used;
}

void NullSpace::build(const std::vector<int32_t>& comp_ptr_host, hipStream_t stream) {
  n_comp = static_cast<int32_t>(comp_ptr_host.size()) - 1;
  std::vector<int32_t> cc, cf, ccp(n_comp + 1, 0);
  for (int32_t c = 0; c < n_comp; ++c) {
    for (int32_t e = comp_ptr_host[c]; e < comp_ptr_host[c + 1]; e += 256) {
      cc.push_back(c);
      cf.push_back(e);
    }
    ccp[c + 1] = static_cast<int32_t>(cc.size());
  }
  n_chunks = static_cast<int32_t>(cc.size());
  chunk_comp.upload(cc, stream);
  chunk_first.upload(cf, stream);
  comp_chunk_ptr.upload(ccp, stream);
}

void remove_null_space(const NullSpace& ns, const int32_t* comp_ptr, const int32_t* members,
                       int64_t n, int64_t k, double* x, DeviceArray<double>& scratch,
                       hipStream_t stream) {
  if (ns.n_chunks == 0 || k == 0) return;
  const int64_t need = static_cast<int64_t>(ns.n_chunks) * k + static_cast<int64_t>(ns.n_comp) * k;
  if (scratch.size < need) scratch.allocate(need);
  double* partials = scratch.ptr;
  double* means = scratch.ptr + static_cast<int64_t>(ns.n_chunks) * k;
  dim3 grid(static_cast<unsigned>(ns.n_chunks), static_cast<unsigned>(k));
  component_partial_kernel<<<grid, 256, 0, stream>>>(ns.chunk_comp.ptr, ns.chunk_first.ptr,
                                                     comp_ptr, members, n, x, partials);
  component_mean_kernel<<<dim3(static_cast<unsigned>(ns.n_comp), static_cast<unsigned>(k)), 256, 0,
                          stream>>>(ns.comp_chunk_ptr.ptr, comp_ptr, ns.n_chunks, partials, means);
  component_shift_kernel<<<grid, 256, 0, stream>>>(ns.chunk_comp.ptr, ns.chunk_first.ptr,
                                                   comp_ptr, members, ns.n_comp, n, means, x);
  HG_HIP(hipGetLastError());
}

void finish_rows(const DevicePlan& plan, const NullSpace& ns, const int32_t* comp_ptr,
                 const int32_t* labels, const double* x, int64_t k, const int32_t* sources,
                 const int64_t* source_offsets, double* const* out_slots,
                 DeviceArray<double>& scratch, hipStream_t stream, bool remove_means) {
  const int64_t n = plan.n;
  if (n == 0 || k == 0 || ns.n_chunks == 0) return;
  const int64_t need = static_cast<int64_t>(ns.n_chunks) * kCols +
                       static_cast<int64_t>(ns.n_comp) * kCols + 2 * kCols;
  if (scratch.size < need) scratch.allocate(need);
  double* partials = scratch.ptr;
  double* means = partials + static_cast<int64_t>(ns.n_chunks) * kCols;
  double* shift = means + static_cast<int64_t>(ns.n_comp) * kCols;
  for (int64_t q0 = 0; q0 < k; q0 += kCols) {
    const int32_t kk = static_cast<int32_t>(std::min<int64_t>(kCols, k - q0));
    const double* xq = x + q0 * n;
    rows_partial_kernel<<<static_cast<unsigned>(ns.n_chunks), 256, 0, stream>>>(
        ns.chunk_comp.ptr, ns.chunk_first.ptr, comp_ptr, xq, kk, partials);
    if (remove_means)
      component_mean_kernel<<<dim3(static_cast<unsigned>(ns.n_comp), static_cast<unsigned>(kk)),
                              256, 0, stream>>>(ns.comp_chunk_ptr.ptr, comp_ptr, ns.n_chunks,
                                                partials, means);
    else
      launch_zero(means, static_cast<int64_t>(sizeof(double)) * ns.n_comp * kk, stream);
    rows_decide_kernel<<<static_cast<unsigned>(kk), 256, 0, stream>>>(
        comp_ptr, ns.chunk_comp.ptr, ns.chunk_first.ptr, ns.n_comp, ns.n_chunks, partials, means,
        xq, kk, n, plan.iperm.ptr, labels, sources, source_offsets + q0, shift);
    rows_out_kernel<<<static_cast<unsigned>((n + kTile - 1) / kTile), 256, 0, stream>>>(
        xq, kk, n, plan.iperm.ptr, labels, ns.n_comp, means, shift, out_slots + q0 / kCols);
  }
  HG_HIP(hipGetLastError());
}

}  // namespace hg
