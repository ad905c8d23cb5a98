// The band march: the settled heat solve of an operator whose factorization is refused under the
// memory budget (ordito linalg.py:2765-2939, kernels/linalg.py:1292-1530).
//
// A band is the unfinished rows within `hops` hops of the finished ones (of the right-hand side's
// support at first). It is solved by conjugate gradient to 1e-14 with the finished values as its
// Dirichlet data and zero beyond it, scaled so its data are O(1); its rows both resolved (within
// e^14 of the band's peak) and clear of the zero it stands on (e^16 above its outer layer) are
// finished, and the march goes on from them until no unfinished row borders a finished one. Each
// band's rows are numbered in ascending row order (a compaction of the claimed rows), so the band
// system, and every sum over it, is the same run to run. The host reads three things a band: its
// row count, its stored entries and its four extreme words (scale, peak, outer layer, whether it
// has one), which decide the next band's hops and the rows this one finishes.
#include <cmath>

#include "kernels.h"

namespace hg {
namespace {

constexpr int kScanThreads = 1024;
constexpr int kScanItems = 4;  // entries a thread scans
constexpr int64_t kScanBlock = static_cast<int64_t>(kScanThreads) * kScanItems;

// The march's policy (linalg.py:185-206): a row is finished at kResolved e-folds or fewer below
// the band's peak and kMargin or more above its outer layer; a band is solved to kTolerance; a
// row far below a neighbour finishes at kDominated past the threshold; the first band reaches
// kFirstHops hops, later ones steer their reach toward kResolved + kMargin e-folds within a factor
// of kGrowth up and two down.
constexpr double kResolved = 14.0;
constexpr double kMargin = 16.0;
constexpr double kTolerance = 1e-14;
constexpr double kDominatedFactor = 54.598150033144236;  // e^4
constexpr int32_t kFirstHops = 32;
constexpr int32_t kGrowth = 4;

// Word indices of a band's extremes.
constexpr int kScale = 0, kPeak = 1, kOuter = 2, kHasOuter = 3;

// Non-negative doubles order as their bit patterns: a maximum by an integer atomic, whatever the
// order the blocks arrive in.
__device__ __forceinline__ void raise_word(double* word, double value) {
  unsigned long long* bits = reinterpret_cast<unsigned long long*>(word);
  const unsigned long long v = static_cast<unsigned long long>(__double_as_longlong(value));
  atomicMax(bits, v);
}

__device__ __forceinline__ double block_max_256(double v, double* scratch) {
  scratch[threadIdx.x] = v;
  __syncthreads();
  for (int stride = kBlock / 2; stride > 0; stride >>= 1) {
    if (threadIdx.x < stride) scratch[threadIdx.x] = fmax(scratch[threadIdx.x], scratch[threadIdx.x + stride]);
    __syncthreads();
  }
  return scratch[0];
}

// A claim: the row's band slot, appended once.
__device__ __forceinline__ void append(int32_t v, int32_t* local, int32_t* band, int32_t* count) {
  const int32_t slot = atomicAdd(count, 1);
  local[v] = slot;
  band[slot] = v;
}

// The first band's first hop: the right-hand side's support (march_seed_support).
__global__ void seed_support_kernel(const double* __restrict__ rhs, int64_t n, int32_t* local,
                                    int32_t* band, int32_t* count) {
  const int64_t v = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (v < n && rhs[v] != 0.0) append(static_cast<int32_t>(v), local, band, count);
}

// A later band's first hop: every unfinished row with a finished neighbour (march_seed_front).
__global__ void seed_front_kernel(const int32_t* __restrict__ offsets,
                                  const int32_t* __restrict__ columns,
                                  const int32_t* __restrict__ finished, int64_t n, int32_t* local,
                                  int32_t* band, int32_t* count) {
  const int64_t v = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (v >= n || finished[v]) return;
  bool front = false;
  for (int32_t e = offsets[v]; e < offsets[v + 1] && !front; ++e) front = finished[columns[e]] != 0;
  if (front) append(static_cast<int32_t>(v), local, band, count);
}

// bounds[hop] = the band's length once hop - 1 is in (march_close_hop).
__global__ void close_hop_kernel(const int32_t* __restrict__ count, int32_t hop, int32_t* bounds) {
  bounds[hop] = count[0];
}

// Hop `hop` of the band's breadth-first growth: each unfinished neighbour of the last hop's rows
// not yet in the band joins it, claimed by a compare-and-swap so it is appended once
// (march_expand_band). The hop's rows are [bounds[hop - 1], bounds[hop]), known on the device only:
// a fixed grid strides over them.
__global__ void expand_kernel(const int32_t* __restrict__ offsets,
                              const int32_t* __restrict__ columns,
                              const int32_t* __restrict__ finished, const int32_t* bounds,
                              int32_t hop, int32_t* local, int32_t* band, int32_t* count) {
  const int32_t begin = bounds[hop - 1], end = bounds[hop];
  const int32_t stride = static_cast<int32_t>(gridDim.x * blockDim.x);
  for (int32_t i = begin + static_cast<int32_t>(blockIdx.x * blockDim.x + threadIdx.x); i < end;
       i += stride) {
    const int32_t v = band[i];
    for (int32_t e = offsets[v]; e < offsets[v + 1]; ++e) {
      const int32_t j = columns[e];
      if (finished[j] == 0 && local[j] == -1 && atomicCAS(&local[j], -1, -2) == -1) {
        append(j, local, band, count);
      }
    }
  }
}

// flags[v] = 1 for a row of the band (claimed, slot >= 0).
__global__ void claimed_kernel(const int32_t* __restrict__ local, int64_t n, int32_t* flags) {
  const int64_t v = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (v < n) flags[v] = local[v] >= 0 ? 1 : 0;
}

// Exclusive scan of int32 counts, three passes: block totals, a one-block scan of the totals,
// each block's own scan plus its offset. out has n + 1 entries, out[n] the total.
__device__ __forceinline__ int32_t block_exclusive_scan(int32_t value, int32_t* scratch,
                                                        int32_t* total) {
  scratch[threadIdx.x] = value;
  __syncthreads();
  // Hillis-Steele over the block (inclusive), then shift.
  for (int stride = 1; stride < kScanThreads; stride <<= 1) {
    const int32_t add = threadIdx.x >= stride ? scratch[threadIdx.x - stride] : 0;
    __syncthreads();
    scratch[threadIdx.x] += add;
    __syncthreads();
  }
  const int32_t inclusive = scratch[threadIdx.x];
  *total = scratch[kScanThreads - 1];
  __syncthreads();
  return inclusive - value;
}

__global__ __launch_bounds__(kScanThreads) void scan_totals_kernel(const int32_t* __restrict__ in,
                                                                   int64_t n,
                                                                   int32_t* __restrict__ totals) {
  __shared__ int32_t scratch[kScanThreads];
  const int64_t base = static_cast<int64_t>(blockIdx.x) * kScanBlock;
  int32_t sum = 0;
  for (int u = 0; u < kScanItems; ++u) {
    const int64_t i = base + static_cast<int64_t>(threadIdx.x) * kScanItems + u;
    if (i < n) sum += in[i];
  }
  int32_t total = 0;
  block_exclusive_scan(sum, scratch, &total);
  if (threadIdx.x == 0) totals[blockIdx.x] = total;
}

__global__ __launch_bounds__(kScanThreads) void scan_offsets_kernel(int32_t* totals,
                                                                    int32_t n_blocks) {
  __shared__ int32_t scratch[kScanThreads];
  int32_t carry = 0;
  for (int32_t b0 = 0; b0 < n_blocks; b0 += kScanThreads) {
    const int32_t b = b0 + static_cast<int32_t>(threadIdx.x);
    const int32_t value = b < n_blocks ? totals[b] : 0;
    int32_t chunk = 0;
    const int32_t before = block_exclusive_scan(value, scratch, &chunk);
    if (b < n_blocks) totals[b] = carry + before;
    carry += chunk;
  }
  if (threadIdx.x == 0) totals[n_blocks] = carry;
}

__global__ __launch_bounds__(kScanThreads) void scan_apply_kernel(const int32_t* __restrict__ in,
                                                                  int64_t n,
                                                                  const int32_t* __restrict__ offs,
                                                                  int32_t* __restrict__ out) {
  __shared__ int32_t scratch[kScanThreads];
  const int64_t base = static_cast<int64_t>(blockIdx.x) * kScanBlock;
  int32_t items[kScanItems];
  int32_t sum = 0;
  for (int u = 0; u < kScanItems; ++u) {
    const int64_t i = base + static_cast<int64_t>(threadIdx.x) * kScanItems + u;
    items[u] = i < n ? in[i] : 0;
    sum += items[u];
  }
  int32_t total = 0;
  int32_t running = offs[blockIdx.x] + block_exclusive_scan(sum, scratch, &total);
  for (int u = 0; u < kScanItems; ++u) {
    const int64_t i = base + static_cast<int64_t>(threadIdx.x) * kScanItems + u;
    if (i < n) out[i] = running;
    running += items[u];
  }
  if (blockIdx.x == gridDim.x - 1 && threadIdx.x == 0) out[n] = offs[gridDim.x];
}

// The band in ascending row order: band[pos[v]] = v, local[v] = pos[v].
__global__ void number_kernel(const int32_t* __restrict__ flags, const int32_t* __restrict__ pos,
                              int64_t n, int32_t* local, int32_t* band) {
  const int64_t v = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (v >= n || !flags[v]) return;
  local[v] = pos[v];
  band[pos[v]] = static_cast<int32_t>(v);
}

// A band row's entries inside the band (counts[i], scanned to the band system's offsets), and the
// band's scale: the largest datum it reads, a right-hand side entry or a finished neighbour's
// value (march_band_rows).
__global__ void band_rows_kernel(const int32_t* __restrict__ offsets,
                                 const int32_t* __restrict__ columns,
                                 const int32_t* __restrict__ band, const int32_t* __restrict__ local,
                                 const int32_t* __restrict__ finished,
                                 const double* __restrict__ solution,
                                 const double* __restrict__ rhs, int32_t m, int32_t* counts,
                                 double* words) {
  __shared__ double scratch[kBlock];
  const int32_t i = static_cast<int32_t>(blockIdx.x * blockDim.x + threadIdx.x);
  double top = 0.0;
  if (i < m) {
    const int32_t v = band[i];
    int32_t count = 0;
    top = fabs(rhs[v]);
    for (int32_t e = offsets[v]; e < offsets[v + 1]; ++e) {
      const int32_t j = columns[e];
      if (local[j] >= 0) ++count;
      else if (finished[j] != 0) top = fmax(top, fabs(solution[j]));
    }
    counts[i] = count;
  }
  top = block_max_256(top, scratch);
  if (threadIdx.x == 0) raise_word(words + kScale, top);
}

// A band row of the band system: its in-band entries renumbered (ascending, as the operator's
// columns are), its finished neighbours' terms moved to the right-hand side, the row's data divided
// by the band's scale (zero for a band whose data are all zero), its guess the last band's value
// where one reached it, and the inverse diagonal for the Jacobi preconditioner
// (march_band_system).
__global__ void band_system_kernel(const int32_t* __restrict__ offsets,
                                   const int32_t* __restrict__ columns,
                                   const double* __restrict__ values,
                                   const int32_t* __restrict__ band,
                                   const int32_t* __restrict__ local,
                                   const int32_t* __restrict__ finished,
                                   const double* __restrict__ solution,
                                   const double* __restrict__ rhs,
                                   const double* __restrict__ tentative,
                                   const double* __restrict__ words,
                                   const int32_t* __restrict__ band_offsets, int32_t m,
                                   int32_t* band_columns, double* band_values, double* band_rhs,
                                   double* guess, double* inv_diag) {
  const int32_t i = static_cast<int32_t>(blockIdx.x * blockDim.x + threadIdx.x);
  if (i >= m) return;
  const int32_t v = band[i];
  int32_t slot = band_offsets[i];
  double r = rhs[v], diag = 0.0;
  for (int32_t e = offsets[v]; e < offsets[v + 1]; ++e) {
    const int32_t j = columns[e];
    const int32_t lj = local[j];
    if (lj >= 0) {
      band_columns[slot] = lj;
      band_values[slot] = values[e];
      if (j == v) diag = values[e];
      ++slot;
    } else if (finished[j] != 0) {
      r -= values[e] * solution[j];
    }
  }
  const double scale = words[kScale];
  const double inverse = scale > 0.0 ? 1.0 / scale : 0.0;
  band_rhs[i] = r * inverse;
  guess[i] = tentative[v] * inverse;
  inv_diag[i] = diag != 0.0 ? 1.0 / diag : 0.0;
}

// The band solution's peak magnitude, and its peak over the outer layer: the rows with an
// unfinished neighbour outside the band, where the zero beyond stands in for the field
// (march_band_extremes).
__global__ void band_extremes_kernel(const int32_t* __restrict__ offsets,
                                     const int32_t* __restrict__ columns,
                                     const int32_t* __restrict__ band,
                                     const int32_t* __restrict__ local,
                                     const int32_t* __restrict__ finished,
                                     const double* __restrict__ x, int32_t m, double* words) {
  __shared__ double scratch[kBlock];
  const int32_t i = static_cast<int32_t>(blockIdx.x * blockDim.x + threadIdx.x);
  double magnitude = 0.0, outer = 0.0, has_outer = 0.0;
  if (i < m) {
    const int32_t v = band[i];
    magnitude = fabs(x[i]);
    for (int32_t e = offsets[v]; e < offsets[v + 1]; ++e) {
      const int32_t j = columns[e];
      if (local[j] < 0 && finished[j] == 0) has_outer = 1.0;
    }
    if (has_outer != 0.0) outer = magnitude;
  }
  magnitude = block_max_256(magnitude, scratch);
  __syncthreads();
  outer = block_max_256(outer, scratch);
  __syncthreads();
  has_outer = block_max_256(has_outer, scratch);
  if (threadIdx.x == 0) {
    raise_word(words + kPeak, magnitude);
    raise_word(words + kOuter, outer);
    raise_word(words + kHasOuter, has_outer);
  }
}

// Finish the band's rows at `threshold` or above, and a row far smaller than a neighbour (in the
// band, or finished, at the band's scale) at `dominated` or above; every row's value seeds the
// next band's guess (march_band_commit).
__global__ void band_commit_kernel(const int32_t* __restrict__ offsets,
                                   const int32_t* __restrict__ columns,
                                   const int32_t* __restrict__ band,
                                   const int32_t* __restrict__ local,
                                   const int32_t* __restrict__ finished,
                                   const double* __restrict__ x, const double* __restrict__ words,
                                   double threshold, double dominated, int32_t m,
                                   double* solution, double* tentative, int32_t* keep_out) {
  const int32_t i = static_cast<int32_t>(blockIdx.x * blockDim.x + threadIdx.x);
  if (i >= m) return;
  const int32_t v = band[i];
  bool keep = fabs(x[i]) >= threshold;
  const double scale = words[kScale];
  if (!keep) {
    const double inverse = scale > 0.0 ? 1.0 / scale : 0.0;
    for (int32_t e = offsets[v]; e < offsets[v + 1] && !keep; ++e) {
      const int32_t j = columns[e];
      const int32_t lj = local[j];
      if (lj >= 0) keep = fabs(x[lj]) >= dominated;
      else if (finished[j] != 0) keep = fabs(solution[j]) * inverse >= dominated;
    }
  }
  const double value = x[i] * scale;
  tentative[v] = value;
  if (keep) solution[v] = value;
  keep_out[i] = keep ? 1 : 0;
}

// Close the band: its rows leave the local numbering and its kept rows are finished
// (march_band_release).
__global__ void band_release_kernel(const int32_t* __restrict__ band,
                                    const int32_t* __restrict__ keep, int32_t m, int32_t* local,
                                    int32_t* finished) {
  const int32_t i = static_cast<int32_t>(blockIdx.x * blockDim.x + threadIdx.x);
  if (i >= m) return;
  const int32_t v = band[i];
  local[v] = -1;
  if (keep[i]) finished[v] = 1;
}

__global__ void fill_kernel(int32_t* p, int64_t n, int32_t value) {
  const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i < n) p[i] = value;
}

unsigned grid_for(int64_t n) { return static_cast<unsigned>((std::max<int64_t>(n, 1) + kBlock - 1) / kBlock); }

template <typename T>
void grow(DeviceArray<T>& a, int64_t need) {
  if (a.size < need) a.allocate(std::max<int64_t>({need, 2 * a.size, 1024}));
}

// out[0..n] = exclusive scan of in[0..n) (out[n] the total); returns the total.
int32_t scan(MarchWorkspace& ws, const int32_t* in, int64_t n, int32_t* out, hipStream_t stream,
             bool read_total) {
  const int64_t blocks = std::max<int64_t>((n + kScanBlock - 1) / kScanBlock, 1);
  grow(ws.scan_totals, blocks + 1);
  scan_totals_kernel<<<static_cast<unsigned>(blocks), kScanThreads, 0, stream>>>(in, n,
                                                                                ws.scan_totals.ptr);
  scan_offsets_kernel<<<1, kScanThreads, 0, stream>>>(ws.scan_totals.ptr,
                                                      static_cast<int32_t>(blocks));
  scan_apply_kernel<<<static_cast<unsigned>(blocks), kScanThreads, 0, stream>>>(
      in, n, ws.scan_totals.ptr, out);
  HG_HIP(hipGetLastError());
  int32_t total = 0;
  if (read_total) {
    HG_HIP(hipMemcpyAsync(&total, out + n, sizeof(int32_t), hipMemcpyDeviceToHost, stream));
    HG_HIP(hipStreamSynchronize(stream));
  }
  return total;
}

// One column's march; returns the band solves' rounds.
int32_t march_column(const MarchSystem& sys, const double* rhs, double* solution,
                     MarchWorkspace& ws, hipStream_t stream) {
  const int64_t n = sys.n;
  const unsigned rows_grid = grid_for(n);
  fill_kernel<<<rows_grid, kBlock, 0, stream>>>(ws.finished.ptr, n, 0);
  fill_kernel<<<rows_grid, kBlock, 0, stream>>>(ws.local.ptr, n, -1);
  launch_zero(ws.tentative.ptr, sizeof(double) * n, stream);
  launch_zero(solution, sizeof(double) * n, stream);
  HG_HIP(hipGetLastError());
  const double resolved = std::exp(-kResolved), margin = std::exp(kMargin);
  const double span_target = kResolved + kMargin;
  int32_t hops = kFirstHops;
  bool started = false;
  int32_t rounds = 0;
  int64_t whole = -1;
  // Each band finishes a row, grows its reach or ends the march; the bound only guards a
  // non-finite solve that would otherwise repeat a band.
  for (int64_t bands = 0; bands < 4 * n + 64; ++bands) {
    grow(ws.bounds, hops + 2);
    launch_zero(ws.bounds.ptr, sizeof(int32_t) * (hops + 2), stream);
    launch_zero(ws.count.ptr, sizeof(int32_t), stream);
    if (started)
      seed_front_kernel<<<rows_grid, kBlock, 0, stream>>>(sys.offsets, sys.columns,
                                                          ws.finished.ptr, n, ws.local.ptr,
                                                          ws.band.ptr, ws.count.ptr);
    else
      seed_support_kernel<<<rows_grid, kBlock, 0, stream>>>(rhs, n, ws.local.ptr, ws.band.ptr,
                                                            ws.count.ptr);
    close_hop_kernel<<<1, 1, 0, stream>>>(ws.count.ptr, 1, ws.bounds.ptr);
    const unsigned hop_grid = std::min<unsigned>(rows_grid, 1024);
    // Every kHopCheck hops the length is read: a band that stopped growing (it holds every
    // unfinished row it can reach) skips the remaining, empty hops.
    constexpr int32_t kHopCheck = 32;
    int32_t m = -1;
    for (int32_t hop = 1; hop <= hops; ++hop) {
      expand_kernel<<<hop_grid, kBlock, 0, stream>>>(sys.offsets, sys.columns, ws.finished.ptr,
                                                     ws.bounds.ptr, hop, ws.local.ptr,
                                                     ws.band.ptr, ws.count.ptr);
      close_hop_kernel<<<1, 1, 0, stream>>>(ws.count.ptr, hop + 1, ws.bounds.ptr);
      if (hop % kHopCheck == 0 && hop < hops) {
        int32_t now = 0;
        HG_HIP(hipMemcpyAsync(&now, ws.count.ptr, sizeof(int32_t), hipMemcpyDeviceToHost,
                              stream));
        HG_HIP(hipStreamSynchronize(stream));
        if (now == m) break;
        m = now;
      }
    }
    HG_HIP(hipGetLastError());
    HG_HIP(hipMemcpyAsync(&m, ws.count.ptr, sizeof(int32_t), hipMemcpyDeviceToHost, stream));
    HG_HIP(hipStreamSynchronize(stream));
    if (m == 0 || m == whole) return rounds;
    // Number the band's rows in ascending order.
    claimed_kernel<<<rows_grid, kBlock, 0, stream>>>(ws.local.ptr, n, ws.flags.ptr);
    scan(ws, ws.flags.ptr, n, ws.positions.ptr, stream, false);
    number_kernel<<<rows_grid, kBlock, 0, stream>>>(ws.flags.ptr, ws.positions.ptr, n,
                                                    ws.local.ptr, ws.band.ptr);
    const unsigned band_grid = grid_for(m);
    launch_zero(ws.words.ptr, sizeof(double) * 4, stream);
    band_rows_kernel<<<band_grid, kBlock, 0, stream>>>(sys.offsets, sys.columns, ws.band.ptr,
                                                       ws.local.ptr, ws.finished.ptr, solution,
                                                       rhs, m, ws.counts.ptr, ws.words.ptr);
    HG_HIP(hipGetLastError());
    const int32_t nnz = scan(ws, ws.counts.ptr, m, ws.band_offsets.ptr, stream, true);
    grow(ws.band_columns, nnz);
    grow(ws.band_values, nnz);
    band_system_kernel<<<band_grid, kBlock, 0, stream>>>(
        sys.offsets, sys.columns, sys.values, ws.band.ptr, ws.local.ptr, ws.finished.ptr, solution,
        rhs, ws.tentative.ptr, ws.words.ptr, ws.band_offsets.ptr, m, ws.band_columns.ptr,
        ws.band_values.ptr, ws.band_rhs.ptr, ws.band_x.ptr, ws.band_inv.ptr);
    HG_HIP(hipGetLastError());
    CgProblem p;
    p.n = m;
    p.n_columns = 1;
    p.offsets = ws.band_offsets.ptr;
    p.columns = ws.band_columns.ptr;
    p.values64 = ws.band_values.ptr;
    p.inv_diag = ws.band_inv.ptr;
    p.rhs = ws.band_rhs.ptr;
    p.guess = ws.band_x.ptr;
    p.x = ws.band_x.ptr;
    p.settle = 0;
    p.tolerance = kTolerance;
    rounds += run_cg(p, ws.cg, stream);
    band_extremes_kernel<<<band_grid, kBlock, 0, stream>>>(sys.offsets, sys.columns, ws.band.ptr,
                                                           ws.local.ptr, ws.finished.ptr,
                                                           ws.band_x.ptr, m, ws.words.ptr);
    HG_HIP(hipGetLastError());
    double words[4] = {0.0, 0.0, 0.0, 0.0};
    HG_HIP(hipMemcpyAsync(words, ws.words.ptr, sizeof(words), hipMemcpyDeviceToHost, stream));
    HG_HIP(hipStreamSynchronize(stream));
    const double peak = words[kPeak];
    double outer = words[kOuter];
    const bool has_outer = words[kHasOuter] > 0.0;
    double threshold = resolved * peak;
    if (has_outer) {
      // An outer layer the solve left below its own tolerance of the peak is unresolved, not zero.
      outer = std::max(outer, kTolerance * peak);
      threshold = std::max(threshold, margin * outer);
      const double span = peak > outer ? std::log(peak / outer) : 0.0;
      const int64_t target = span <= 0.0
                                 ? static_cast<int64_t>(hops) * kGrowth
                                 : static_cast<int64_t>(std::ceil(hops * span_target / span));
      const int64_t next = std::min<int64_t>(
          static_cast<int64_t>(hops) * kGrowth,
          std::max<int64_t>({static_cast<int64_t>(hops / 2), target, int64_t{1}}));
      hops = static_cast<int32_t>(std::min<int64_t>(next, n + 1));
    }
    band_commit_kernel<<<band_grid, kBlock, 0, stream>>>(
        sys.offsets, sys.columns, ws.band.ptr, ws.local.ptr, ws.finished.ptr, ws.band_x.ptr,
        ws.words.ptr, threshold, threshold * kDominatedFactor, m, solution, ws.tentative.ptr,
        ws.keep.ptr);
    band_release_kernel<<<band_grid, kBlock, 0, stream>>>(ws.band.ptr, ws.keep.ptr, m,
                                                          ws.local.ptr, ws.finished.ptr);
    HG_HIP(hipGetLastError());
    // A band too thin to finish anything is grown from the same rows again.
    started = started || peak >= threshold;
    whole = has_outer ? -1 : m;
  }
  HG_REQUIRE(false, HG_ERR_RUNTIME, "the band march did not finish (a non-finite band solve?)");
  return rounds;
}

}  // namespace

int64_t MarchWorkspace::bytes() const {
  return finished.bytes() + local.bytes() + band.bytes() + flags.bytes() + positions.bytes() +
         counts.bytes() + band_offsets.bytes() + band_columns.bytes() + keep.bytes() +
         scan_totals.bytes() + bounds.bytes() + count.bytes() + tentative.bytes() +
         band_values.bytes() + band_rhs.bytes() + band_x.bytes() + band_inv.bytes() +
         words.bytes() + cg.vectors.bytes() + cg.partials.bytes() + cg.counters.bytes();
}

void MarchWorkspace::release() {
  *this = MarchWorkspace();
}

int32_t march_settled(const MarchSystem& sys, const double* rhs, double* x, int64_t n_columns,
                      MarchWorkspace& ws, hipStream_t stream) {
  const int64_t n = sys.n;
  if (n == 0 || n_columns == 0) return 0;
  HG_REQUIRE(n < (int64_t{1} << 31) - 1, HG_ERR_VALUE, "the march numbers rows in int32");
  grow(ws.finished, n);
  grow(ws.local, n);
  grow(ws.band, n);
  grow(ws.flags, n);
  grow(ws.positions, n + 1);
  grow(ws.counts, n);
  grow(ws.band_offsets, n + 1);
  grow(ws.keep, n);
  grow(ws.count, 1);
  grow(ws.tentative, n);
  grow(ws.band_rhs, n);
  grow(ws.band_x, n);
  grow(ws.band_inv, n);
  grow(ws.words, 4);
  int32_t rounds = 0;
  for (int64_t q = 0; q < n_columns; ++q)
    rounds += march_column(sys, rhs + q * n, x + q * n, ws, stream);
  return rounds;
}

}  // namespace hg
