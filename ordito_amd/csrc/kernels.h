// Launchers of the device kernels, called by the C ABI layer (api.hip).
#pragma once

#include "hg_internal.h"

namespace hg {

constexpr int kBlock = 256;

// How a field of k columns over n vertices is stored. kColumns: the caller's layout, column q at
// q * n. kRows: in chunks of 64 columns (the solves' width), chunk c at 64 c * n holding its kk
// columns interleaved per row, rows in the factorization's elimination order (vertex v's entries
// at iperm[v] * kk): the layout the refined solves work in, so the seeds, the heat step, the
// divergence and the Poisson step hand fields on with no permutation pass between them.
enum class FieldLayout : int { kColumns = 0, kRows = 1 };
constexpr int64_t kFieldChunk = 64;

// The work item a block of a one-dimensional launch takes, numbered so that each XCD's share of the
// grid is one contiguous range: the hardware deals consecutive blocks to the eight XCDs in turn,
// and neighbouring work items (a front's tiles, neighbouring rows) then meet in one XCD's L2.
// A bijection of [0, gridDim.x).
__device__ __forceinline__ int32_t xcd_block() {
  const int32_t blocks = static_cast<int32_t>(gridDim.x), b = static_cast<int32_t>(blockIdx.x);
  const int32_t per = blocks / 8, extra = blocks % 8, xcd = b % 8;
  return xcd * per + min(xcd, extra) + b / 8;
}

// Device time of the phases of a solve, from HIP events on the stream the work is issued on.
enum Phase : int {
  kPhaseFactor = 0,
  kPhaseHeatSweep = 1,     // one forward + backward triangular sweep of the heat system
  kPhasePoissonSweep = 2,  // the same for the Poisson system
  kPhaseResidual = 3,      // a refinement residual and its test
  kPhaseDivergence = 4,
  kPhaseShift = 5,
  kPhaseCg = 6,
  kPhaseAssembly = 7,
  kPhaseCount = 8
};

struct PhaseTimer {
  bool enabled = false;
  std::vector<hipEvent_t> pool;
  size_t used = 0;
  struct Span {
    int phase;
    hipEvent_t start, stop;
  };
  std::vector<Span> pending;
  double ms[kPhaseCount] = {};
  int64_t count[kPhaseCount] = {};

  hipEvent_t mark(hipStream_t stream) {
    if (used == pool.size()) {
      hipEvent_t e;
      HG_HIP(hipEventCreate(&e));
      pool.push_back(e);
    }
    hipEvent_t e = pool[used++];
    HG_HIP(hipEventRecord(e, stream));
    return e;
  }
  void add(int phase, hipEvent_t start, hipEvent_t stop) { pending.push_back({phase, start, stop}); }
  // Call after the stream is synchronized: folds the recorded spans into the totals.
  void resolve() {
    for (const Span& s : pending) {
      float t = 0.0f;
      if (hipEventElapsedTime(&t, s.start, s.stop) == hipSuccess) {
        ms[s.phase] += t;
        count[s.phase] += 1;
      }
    }
    pending.clear();
    used = 0;
  }
  void reset() {
    pending.clear();
    used = 0;
    for (int i = 0; i < kPhaseCount; ++i) {
      ms[i] = 0.0;
      count[i] = 0;
    }
  }
  ~PhaseTimer() {
    for (hipEvent_t e : pool) (void)hipEventDestroy(e);
  }
};

// Records a phase's span when the timer is enabled.
struct PhaseScope {
  PhaseTimer* timer;
  int phase;
  hipStream_t stream;
  hipEvent_t start = nullptr;
  PhaseScope(PhaseTimer* t, int p, hipStream_t s) : timer(t), phase(p), stream(s) {
    if (timer && timer->enabled) start = timer->mark(stream);
  }
  ~PhaseScope() {
    if (start) timer->add(phase, start, timer->mark(stream));
  }
};

// assembly.hip
void launch_face_terms(const float* xyz, const int32_t* faces, int64_t n_faces, float* cot,
                       float* normals, float* areas, hipStream_t stream);
// use_robust (laplacian.py:779-874, 269-340): cot (F,3) float32 rebuilt from the side lengths
// plus the smallest constant that leaves every triangle-inequality margin at least
// kMollifyTolerance times the mean side; returns that constant (0 on a clean mesh). Synchronizes
// `stream` (the constant is decided on the host).
constexpr double kMollifyTolerance = 1e-5;  // constants.py:31
double launch_mollified_cot(const float* xyz, const int32_t* faces, int64_t n_faces, float* cot,
                            hipStream_t stream);
// Each corner's Voronoi share of its face's area (F,3) float32.
void launch_voronoi_shares(const float* xyz, const int32_t* faces, int64_t n_faces, float* shares,
                           hipStream_t stream);
// The lumped mass: barycentric (shares null) or the sum of the corners' Voronoi shares.
void launch_mass(const int32_t* vertex_ptr, const int32_t* vertex_corners, const float* areas,
                 const float* shares, int64_t n, double* mass, hipStream_t stream);
// The rim-pinned heat system's values (rim flags by vertex), float32 copy and inverse diagonal.
void launch_pinned_rows(const int32_t* offsets, const int32_t* columns, const int32_t* rim,
                        const double* heat, int64_t n, double* out, float* out32, double* inv_diag,
                        hipStream_t stream);
// out = values with `shift` added on each stored diagonal entry, and its inverse diagonal.
void launch_shifted_rows(const int32_t* offsets, const int32_t* columns, const double* values,
                         int64_t n, double shift, double* out, double* inv_diag,
                         hipStream_t stream);
// out = in with the flagged rows zeroed (in == out allowed); rim_of by vertex for kColumns, by
// row of the elimination order for kRows.
void launch_off_rim(const double* in, double* out, const int32_t* rim_of, int64_t n,
                    int64_t n_columns, FieldLayout layout, hipStream_t stream);
// a = (a + b) / 2 elementwise.
void launch_average(double* a, const double* b, int64_t count, hipStream_t stream);
int32_t edge_sum_blocks(int64_t n);
// Edge-length sums over rows [row0, row0 + n) into scalars[0..4).
void launch_edge_sums(const int32_t* offsets, const int32_t* columns, const int32_t* corner_ptr,
                      const float* xyz, int64_t row0, int64_t n, int32_t face_edges,
                      double* partials, int32_t from_edges, double t, double* scalars,
                      hipStream_t stream);
// row_segment (nullable): each row's segment s, whose heat factor is scalars[4 s + 2].
void launch_operator_rows(const int32_t* offsets, const int32_t* columns,
                          const int32_t* corner_ptr, const int32_t* corner_slots, const float* cot,
                          const double* mass, const double* scalars, const int32_t* row_segment,
                          int64_t n, double* lap, double* heat, double* poisson, float* heat32,
                          double* heat_inv_diag, double* poisson_inv_diag, hipStream_t stream);
// rows_k > 0: u0 stored by rows (FieldLayout::kRows) over rows_k columns, vertex v at row
// row_of[v]; 0: by columns.
// count_at (device, nullable): read the entry count there, at most n_entries.
void launch_seed(const int32_t* sources, const int64_t* owner, int64_t n_entries,
                 const int64_t* count_at, int64_t n, int64_t rows_k, const int32_t* row_of,
                 double* u0, int32_t* invalid, hipStream_t stream);
// Zeroes `bytes` (whole 32-bit words) with a kernel (graph-capturable, unlike a memset node).
void launch_zero(void* p, int64_t bytes, hipStream_t stream);
// Per face, kFaceGeom doubles: two of the gradient basis vectors a_k = N x e_k (the third is minus
// their sum) and the three corners' integration vectors.
constexpr int64_t kFaceGeom = 15;
void launch_face_geometry(const float* xyz, const int32_t* faces, const float* normals,
                          const float* areas, const float* cot, int64_t n_faces, double* geom,
                          hipStream_t stream);
// -b per vertex and column from the heat (both in `layout`), gathered over each vertex's corners.
// kRows: vertex v's row is row_of[v], `faces` name rows (row_of applied to every corner), and
// row_corners[row_corner_ptr[r] ..] are row r's corners in ascending corner order, each
// {corner id 3f + c, rows of face f's vertices 0, 1, 2}.
void launch_divergence(const int32_t* faces, const int32_t* row_of, const double* geom,
                       const int32_t* vertex_ptr, const int32_t* vertex_corners,
                       const int4* row_corners, const int32_t* row_corner_ptr,
                       const double* heat, int64_t n, int64_t n_columns, double* out,
                       FieldLayout layout, hipStream_t stream);
int32_t shift_partial_blocks(int64_t n);
// Columns start at col_off[q] with col_len[q] entries (device; both null: q * n and n), n the
// longest.
void launch_shift(double* phi, const int32_t* sources, const int64_t* source_offsets, int64_t n,
                  int64_t n_columns, const int64_t* col_off, const int64_t* col_len,
                  double* partials, double* shift, hipStream_t stream);
void launch_extract_fields(const double* phi, const int64_t* src_off, const int64_t* dst_off,
                           const int64_t* len, int64_t max_len, int64_t n_queries, double* out,
                           hipStream_t stream);

// cg.hip: the persistent conjugate-gradient solves.
struct CgProblem {
  int64_t n = 0;
  int64_t n_columns = 1;
  const int32_t* offsets = nullptr;
  const int32_t* columns = nullptr;
  const float* values32 = nullptr;   // narrowed values, or null to read values64
  const double* values64 = nullptr;
  const double* inv_diag = nullptr;
  const double* rhs = nullptr;
  const double* guess = nullptr;  // initial iterate (may alias x), or null for zero
  double* x = nullptr;
  // Stop rule: settle (the heat step) or relative residual tol (the Poisson step).
  int32_t settle = 1;
  int32_t check_rounds = 16;
  double change_tolerance = 1e-6;
  int32_t settle_rounds = 192;
  double tolerance = 1e-8;
  int64_t maxiter = 0;
  // Jacobi-Chebyshev polynomial preconditioner: steps > 0 applies it, with its interval.
  int32_t chebyshev_steps = 0;
  double lambda_min = 0.0;
  double lambda_max = 0.0;
};

struct CgWorkspace {
  DeviceArray<double> vectors;   // r, u, p, s, w, prev, and two polynomial vectors
  DeviceArray<double> partials;
  DeviceArray<int32_t> counters;
  int64_t capacity_entries = 0;
  int32_t capacity_blocks = 0;
  int32_t capacity_columns = 0;
  int64_t bytes() const { return vectors.bytes() + partials.bytes() + counters.bytes(); }
  void release() {
    vectors.reset();
    partials.reset();
    counters.reset();
    capacity_entries = 0;
    capacity_blocks = 0;
    capacity_columns = 0;
  }
};

// Runs the solve; returns the rounds that stepped.
int32_t run_cg(const CgProblem& problem, CgWorkspace& workspace, hipStream_t stream);

// Componentwise backward error |b - Ax|_i <= tol (|A||x| + |b|)_i everywhere, as a flag.
void launch_backward_error(const int32_t* offsets, const int32_t* columns, const double* values,
                           const double* x, const double* b, int64_t n, int64_t n_columns,
                           double tolerance, int32_t* flag, hipStream_t stream);

// Largest Gershgorin ratio sum_{j != i} |A_ij| / |A_ii| over the rows (for the polynomial's
// interval), reduced into *out (one block).
void launch_dominance_bound(const int32_t* offsets, const int32_t* columns, const double* values,
                            int64_t n, double* partials, double* out, hipStream_t stream);

// march.hip: the band march of a system refused under the memory budget.
struct MarchSystem {
  int64_t n = 0;
  const int32_t* offsets = nullptr;
  const int32_t* columns = nullptr;
  const double* values = nullptr;
};

struct MarchWorkspace {
  DeviceArray<int32_t> finished, local, band, flags, positions, counts, band_offsets, band_columns,
      keep, scan_totals, bounds, count;
  DeviceArray<double> tentative, band_values, band_rhs, band_x, band_inv, words;
  CgWorkspace cg;
  int64_t bytes() const;
  void release();
};

// Solves A x = rhs for n_columns columns (column q at q * n), each marched on its own; x is
// overwritten. Returns the band solves' conjugate-gradient rounds.
int32_t march_settled(const MarchSystem& system, const double* rhs, double* x, int64_t n_columns,
                      MarchWorkspace& workspace, hipStream_t stream);

// multigrid.hip: a smoothed-aggregation hierarchy on the device and the conjugate gradient it
// preconditions (the hierarchy is built on the host, multigrid.h).
struct HostCsr;
struct MultigridLevel {
  struct Matrix {
    int64_t rows = 0;
    DeviceArray<int32_t> ptr, idx;
    DeviceArray<double> val;
    int64_t bytes() const { return ptr.bytes() + idx.bytes() + val.bytes(); }
  };
  Matrix a, p, r;
  DeviceArray<double> smoother;
  DeviceArray<double> u, v, res, b, x;  // a cycle's vectors at this level (b, x below the top)
};

struct Multigrid {
  std::vector<MultigridLevel> levels;
  DeviceArray<double> coarse_inverse;
  DeviceArray<double> r, z, p, w, partials, scalars;
  DeviceArray<int32_t> flags;
  int32_t blocks = 0;
  int64_t bytes() const;
};

// The hierarchy of a host operator, uploaded; null when the operator does not coarsen to a
// bottom small enough to invert densely (the caller keeps its single-level preconditioner).
std::unique_ptr<Multigrid> build_multigrid(const HostCsr& fine, uint64_t seed, hipStream_t stream);

// Conjugate gradient under one V-cycle a round, from zero, to |r| <= tolerance |b| per column;
// returns the most rounds a column took.
int32_t multigrid_cg(Multigrid& mg, const double* rhs, double* x, int64_t n_columns,
                     double tolerance, int64_t maxiter, hipStream_t stream);

}  // namespace hg
