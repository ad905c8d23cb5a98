// Internal declarations of libheatgeo: mesh state, the Cholesky plan and error plumbing.
// Everything here is private to the library; the public C ABI is include/heatgeo.h.
#pragma once

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>
#include <future>
#include <memory>
#include <string>
#include <thread>
#include <vector>

#include "../../include/heatgeo.h"

namespace hg {

// ------------------------------------------------------------------------------------------------
// Errors: every entry point returns an hg_status and leaves a message for hg_last_error().
// ------------------------------------------------------------------------------------------------

void set_error(const std::string& message);

struct Error {
  int status;
  std::string message;
};

#define HG_HIP(call)                                                                          \
  do {                                                                                        \
    hipError_t hg_err_ = (call);                                                              \
    if (hg_err_ != hipSuccess) {                                                              \
      throw ::hg::Error{HG_ERR_RUNTIME, std::string(#call) + ": " + hipGetErrorString(hg_err_)}; \
    }                                                                                         \
  } while (0)

#define HG_REQUIRE(cond, status, message)          \
  do {                                             \
    if (!(cond)) throw ::hg::Error{(status), (message)}; \
  } while (0)

// A device allocation owned by the library.
template <typename T>
struct DeviceArray {
  T* ptr = nullptr;
  int64_t size = 0;
  DeviceArray() = default;
  explicit DeviceArray(int64_t n) { allocate(n); }
  DeviceArray(const DeviceArray&) = delete;
  DeviceArray& operator=(const DeviceArray&) = delete;
  DeviceArray(DeviceArray&& o) noexcept : ptr(o.ptr), size(o.size) { o.ptr = nullptr; o.size = 0; }
  DeviceArray& operator=(DeviceArray&& o) noexcept {
    if (this != &o) { reset(); ptr = o.ptr; size = o.size; o.ptr = nullptr; o.size = 0; }
    return *this;
  }
  ~DeviceArray() { reset(); }
  void allocate(int64_t n) {
    reset();
    size = n;
    if (n > 0) HG_HIP(hipMalloc(&ptr, sizeof(T) * static_cast<size_t>(n)));
  }
  void reset() {
    if (ptr) (void)hipFree(ptr);
    ptr = nullptr;
    size = 0;
  }
  int64_t bytes() const { return size * static_cast<int64_t>(sizeof(T)); }
  void upload(const T* host, int64_t n, hipStream_t stream) {
    if (n > 0) HG_HIP(hipMemcpyAsync(ptr, host, sizeof(T) * n, hipMemcpyHostToDevice, stream));
  }
  void upload(const std::vector<T>& host, hipStream_t stream) {
    allocate(static_cast<int64_t>(host.size()));
    upload(host.data(), size, stream);
  }
  // Copies host into the array, reallocating only to grow: per-call arrays (sources, offsets)
  // keep their memory, so a solve allocates nothing once warm (hipMalloc and hipFree synchronize
  // the device, which would serialize solves on other streams).
  void assign(const std::vector<T>& host, hipStream_t stream) {
    const int64_t count = static_cast<int64_t>(host.size());
    if (size < std::max<int64_t>(count, 1)) allocate(std::max<int64_t>(count, 1));
    upload(host.data(), count, stream);
  }
};

// A pinned host allocation (hipHostMalloc): device-to-host copies into it stay asynchronous.
template <typename T>
struct PinnedArray {
  T* ptr = nullptr;
  int64_t size = 0;
  PinnedArray() = default;
  PinnedArray(const PinnedArray&) = delete;
  PinnedArray& operator=(const PinnedArray&) = delete;
  ~PinnedArray() { reset(); }
  void allocate(int64_t n) {
    reset();
    HG_HIP(hipHostMalloc(reinterpret_cast<void**>(&ptr), sizeof(T) * static_cast<size_t>(n),
                         hipHostMallocDefault));
    size = n;
  }
  void reset() {
    if (ptr) (void)hipHostFree(ptr);
    ptr = nullptr;
    size = 0;
  }
};

// ------------------------------------------------------------------------------------------------
// Host analysis (pattern.cpp, analysis.cpp)
// ------------------------------------------------------------------------------------------------

// The cotangent operator's sparsity: every face edge in both directions plus a diagonal per
// referenced vertex, rows sorted by column, and for each stored entry its run of contributing
// corner slots 3f+e (corner e opposite the edge) in ascending order.
struct HostPattern {
  int64_t n = 0;
  std::vector<int32_t> offsets;       // n+1
  std::vector<int32_t> columns;       // nnz
  std::vector<int32_t> corner_ptr;    // nnz+1
  std::vector<int32_t> corner_slots;  // <= 6F
  std::vector<int32_t> vertex_ptr;    // n+1: each vertex's corners, ascending
  std::vector<int32_t> vertex_corners;
  std::vector<int32_t> diagonal;      // n: slot of the diagonal entry, -1 for an empty row
};

// The pattern's device arrays (the HostPattern fields but the diagonal).
struct PatternArrays {
  DeviceArray<int32_t>& offsets;
  DeviceArray<int32_t>& columns;
  DeviceArray<int32_t>& corner_ptr;
  DeviceArray<int32_t>& corner_slots;
  DeviceArray<int32_t>& vertex_ptr;
  DeviceArray<int32_t>& vertex_corners;
};

// Builds the pattern on the device from the device faces (range-checked there) and, when host is
// not null, downloads it there for the host-side work.
void build_pattern_device(int64_t n_vertices, const int32_t* d_faces, int64_t n_faces,
                          hipStream_t stream, PatternArrays out, HostPattern* host);

// Connected components of the pattern's graph.
std::vector<int32_t> component_labels(const HostPattern& pattern, int32_t* n_components);

// A supernodal Cholesky plan: nested-dissection order, supernode tree, level schedule, the maps
// the numeric factorization and the multifrontal solves use.
struct Plan {
  int64_t n = 0;
  int32_t n_super = 0;
  int32_t n_levels = 0;
  std::vector<int32_t> perm;        // new -> old
  std::vector<int32_t> iperm;       // old -> new
  std::vector<int32_t> col_begin;   // per supernode, new index of its first column
  std::vector<int32_t> ncol;
  std::vector<int32_t> nrow;
  std::vector<int32_t> parent;
  std::vector<int32_t> level;
  std::vector<int32_t> row_ptr;     // CSR over supernodes: the structure rows (new indices)
  std::vector<int32_t> rows;
  std::vector<int32_t> child_ptr;   // CSR over supernodes: children, ascending
  std::vector<int32_t> children;
  std::vector<int32_t> rel;         // per supernode s, nrow[s] positions of rows(s) in parent's front
  std::vector<int32_t> rel_ptr;     // = row_ptr
  std::vector<int32_t> level_ptr;   // CSR over levels: supernodes at each level (bottom-up)
  std::vector<int32_t> level_nodes;
  std::vector<int64_t> front_off;   // arena offset (doubles) of each m x m front
  std::vector<int64_t> w_off;       // offset of each m x ncol solve block
  std::vector<int64_t> vec_off;     // offset (per column) of each front's m-vector in solve scratch
  int64_t arena_size = 0;
  int64_t w_size = 0;
  int64_t vec_size = 0;
  // Lower-triangle entries of the operator in front coordinates: value slot -> arena offset.
  std::vector<int32_t> entry_src;
  std::vector<int64_t> entry_dst;
  std::vector<int32_t> entry_row;   // new row index of the entry
  std::vector<int32_t> entry_col;   // new column index
  double flops = 0.0;               // dense partial-factorization flops
  int64_t nnz_w = 0;                // stored solve-block entries
  int32_t max_front = 0;
};

// Leaf 128: on the 1024 x 512 torus with 64 right-hand sides, leaves of 96-192 rows serve fastest
// (fewer, fuller levels), 128 by a small margin over 48 (3%) and 512 (16%).
struct AnalysisOptions {
  int32_t leaf_rows = 128;
  int32_t top_rows = 512;
};

Plan analyse(const HostPattern& pattern, const float* coordinates, const AnalysisOptions& options);

// The triangular sweeps' product tasks of one direction (cholesky.hip): per level, every 64-wide
// tile of every front and its inner range [k0, k1); slot >= 0 names a partial that the level's
// reduction groups (rs, rtile, rfirst, rcount) add in slot order.
struct HostSweep {
  std::vector<int32_t> level_ptr, reduce_ptr, slots;
  std::vector<int32_t> s, tile, k0, k1, slot;
  std::vector<int32_t> rs, rtile, rfirst, rcount;
  int32_t max_slots = 0;
};
// A level whose tiles are fewer than `target_tasks` has its products split along the inner
// dimension until it has about that many tasks (a few per compute unit by default).
constexpr int32_t kDefaultSplitTarget = 1024;
HostSweep sweep_tasks(const Plan& plan, bool forward, int32_t target_tasks = kDefaultSplitTarget);

// The forward sweep's gather lists: for each front-vector row (vec_off[s] + position) the rows of
// the children's update vectors that add into it, and the first two of them inline per position
// (pair[2v], pair[2v+1]; -1 none, -2 in the second: more than two, the rest in the lists).
struct HostGather {
  std::vector<int32_t> ptr, src, pair;
};
HostGather gather_lists(const Plan& plan);

// Checks every index the device sweeps form from the plan against its buffer; throws on the first
// violation.
void validate_sweeps(const Plan& plan, int32_t target_tasks = kDefaultSplitTarget);

// The numeric factorization as launches of 64 x 64 tile tasks (cholesky.hip), in execution order.
// A front's inverse factor Z is the first ncol columns of the inverse of its unit-extended lower
// factor [L11 0; L21 I], i.e. [L11^-1; -L21 L11^-1] (m x ncol, column-major, leading dimension
// m); it is what the sweeps multiply by. Task fields (4 int32 each):
//   kFactorExtend (parent s, child c, i0, k0): the child's Schur-complement tile rows
//       [i0, i0+64) x columns [k0, k0+64), i0 >= k0, added into the parent through the child's
//       relative map. One launch per child rank, so siblings add in ascending order.
//   kFactorPanel (s, j0, kind, t0): the diagonal block at j0 (width b = min(64, ncol - j0)) is
//       eliminated in LDS together with its inverse; kind 0 solves the panel rows [t0, t0+64)
//       below it, kind 1 applies the inverse to Z's block rows at columns [t0, t0+64) (t0 < j0),
//       kind 2 stores the inverse as Z's diagonal block.
//   kFactorUpdate (s, j0, i0, c0): rows [i0, i0+64) minus the panel's rows [i0, i0+64) times
//       the panel rows [c0, c0+64) transposed (the trailing matrix, c0 >= j0 + b, c0 <= i0) or
//       times Z's block rows at columns [c0, c0+64) (c0 < j0 + b).
enum FactorKind : int32_t { kFactorExtend = 0, kFactorPanel = 1, kFactorUpdate = 2 };
struct FactorLaunch {
  int32_t kind, level, first, count;
};
struct HostFactorTasks {
  std::vector<FactorLaunch> launches;
  std::vector<int32_t> tasks;  // 4 per task
};
constexpr int32_t kFactorTile = 64;
HostFactorTasks factor_tasks(const Plan& plan);

// Checks every factorization task against the plan (fronts, children, tile ranges, the step
// order); throws on the first violation.
void validate_factor_tasks(const Plan& plan, const HostFactorTasks& tasks);

// Runs body(begin, end, worker) over [0, count) in contiguous chunks, one a worker thread (the
// caller's thread takes the first); waits for all, rethrowing the first failure.
template <typename F>
void parallel_chunks(int64_t count, int64_t grain, F&& body) {
  const int64_t hw = std::max<int64_t>(1, std::thread::hardware_concurrency());
  const int64_t workers = std::max<int64_t>(1, std::min<int64_t>(hw, (count + grain - 1) / grain));
  const int64_t chunk = (count + workers - 1) / std::max<int64_t>(workers, 1);
  std::vector<std::future<void>> pool;
  for (int64_t w = 1; w < workers; ++w) {
    const int64_t b = w * chunk, e = std::min(count, b + chunk);
    if (b >= e) break;
    pool.push_back(std::async(std::launch::async, [&body, b, e, w] { body(b, e, w); }));
  }
  body(0, std::min(count, chunk), 0);
  for (std::future<void>& f : pool) f.get();
}

}  // namespace hg
