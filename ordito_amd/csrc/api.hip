// The C ABI of libheatgeo (include/heatgeo.h): mesh state, operator assembly, the heat and Poisson
// solves in both modes, the divergence, the shift and the fused geodesic path.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <limits>
#include <map>
#include <mutex>

#include "cholesky.h"
#include "multigrid.h"

namespace hg {

thread_local std::string g_last_error;

void set_error(const std::string& message) { g_last_error = message; }

// Constants of the reference's solver policy.
constexpr int32_t kHeatCheckRounds = 16;          // heat.py:147
constexpr double kHeatChangeTolerance = 1e-6;     // heat.py:151
constexpr int32_t kHeatSettleRounds = 192;        // heat.py:152
constexpr double kDiffusionDecay = 400.0;         // heat.py:158
constexpr double kSettleBackwardError = 1e-6;     // linalg.py:179
constexpr double kSettleDirectTolerance = 1e-12;  // linalg.py:181
constexpr double kPoissonTolerance = 1e-8;        // heat.py:108, 4138-4140
constexpr int32_t kRefineRounds = 12;             // cholesky.py:74
// The mixed path (FP32 solve blocks, FP64 residuals) refines to a served accuracy instead of the
// reference's: componentwise 1e-6 for the heat step (each row, however small its value, to six
// digits, so every face's gradient direction is kept) and normwise 1e-6 for the Poisson step.
// Measured on a 32k-vertex jittered torus this leaves the distance within about 3e-8 of the range
// of the FP64 direct answer, far inside a 1e-4 bar; the reference's tolerances cost about twice
// the refinement rounds for digits the served answer does not need.
constexpr double kMixedHeatTolerance = 1e-6;
constexpr double kMixedPoissonTolerance = 1e-6;
constexpr int32_t kMixedRefineRounds = 24;
constexpr double kPotpourriPoissonShift = 1e-6;   // heat.py:193: geometry-central's -L + 1e-6 I
constexpr int32_t kChebyshevSteps = 12;           // linalg.py CHEBYSHEV_DEGREE: 11 products
constexpr double kChebyshevInterval = 80.0;       // linalg.py CHEBYSHEV_INTERVAL
constexpr double kChebyshevCap = 0.5;

}  // namespace hg

using namespace hg;

struct hg_mesh {
  hipStream_t stream = nullptr;
  int64_t n = 0, f = 0;
  std::vector<float> h_xyz;
  std::vector<int32_t> h_faces;
  // Segments: meshes packed block-diagonally into this one (hg_mesh_create_packed); segment s owns
  // vertices [seg_v[s], seg_v[s+1]) and faces [seg_f[s], seg_f[s+1]). A plain mesh is one segment.
  std::vector<int64_t> seg_v{0, 0}, seg_f{0, 0};
  std::vector<double> seg_t;
  DeviceArray<int32_t> row_segment;  // packed meshes only
  int64_t segments() const { return static_cast<int64_t>(seg_v.size()) - 1; }
  HostPattern pattern;
  DeviceArray<float> xyz;
  DeviceArray<int32_t> faces, offsets, columns, corner_ptr, corner_slots, vertex_ptr,
      vertex_corners;
  // Operators.
  bool assembled = false;
  int32_t method = HG_METHOD_MESHLAB;
  double t = 0.0;
  int32_t settle_rounds = kHeatSettleRounds;
  int32_t time_from_edges = 1;
  DeviceArray<float> cot, normals, areas, heat32;
  DeviceArray<double> mass, lap, heat, poisson, heat_inv, poisson_inv, scalars, edge_partials;
  DeviceArray<double> face_geom;  // kFaceGeom doubles a face: the divergence's per-face data
  DeviceArray<float> voronoi;     // HG_METHOD_IGL: each corner's Voronoi share (F,3)
  // The rim (vertices with an edge on one face) and the heat system pinned to zero there, built on
  // its first use (HeatSolver._pinned_system, heat.py:4202-4233).
  int64_t n_rim = 0;
  std::vector<int32_t> rim;
  DeviceArray<int32_t> d_rim, d_rim_rows;  // by vertex; by row of the elimination order
  bool pinned_built = false;
  DeviceArray<double> pinned, pinned_inv;
  DeviceArray<float> pinned32;
  DeviceArray<double> pinned_rhs;
  double poisson_bound = -1.0;
  // HG_METHOD_POTPOURRI3D's Poisson step: -L + 1e-6 I over the same pattern, built on first use
  // (HeatSolver._solve_regularized_poisson, heat.py:4085-4110); no null space, its empty rows
  // (unreferenced vertices) pinned as the heat system's are.
  bool reg_built = false;
  DeviceArray<double> reg, reg_inv;
  double reg_bound = -1.0;
  // Components: the Poisson system pins each component's last-eliminated row.
  int32_t n_comp = 0;
  std::vector<int32_t> labels;
  std::vector<int32_t> comp_ptr_host;  // components' member ranges in label order
  DeviceArray<int32_t> d_labels;
  bool rows_by_component = false;      // every component's rows contiguous in the elimination order
  DeviceArray<int32_t> comp_ptr, comp_members;
  DeviceArray<int32_t> poisson_pinned_new, poisson_dropped_old, heat_pinned_new, heat_dropped_old;
  std::vector<int64_t> poisson_pin_slots, heat_pin_slots;
  bool has_heat_pins = false;
  // Cholesky.
  AnalysisOptions options;                  // nested-dissection leaf and top-merge sizes
  int32_t split_target = kDefaultSplitTarget;  // sweep tasks a level is split towards
  std::unique_ptr<Plan> host_plan;
  std::unique_ptr<DevicePlan> plan;
  DeviceArray<int32_t> faces_rows;  // the faces' corners as rows of the elimination order
  DeviceArray<int4> row_corners;    // each row's corners {corner id, its face's three rows}
  DeviceArray<int32_t> row_corner_ptr;
  Factor heat_factor, poisson_factor, pinned_factor, reg_factor;
  // The factorizations' memory budget in bytes (0: 90 % of the device memory free when a
  // factorization starts) and the systems refused under it (HG_SYSTEM_* bits).
  int64_t budget = 0;
  int32_t refused = 0;
  // Whether a system is asked of its analysis before it iterates (hg_set_budget_probe).
  bool budget_probe = true;
  // use_robust: the half-cotangents from side lengths mollified by mollify_delta.
  bool robust = false;
  double mollify_delta = 0.0;
  // The settle rule of the heat solves' iteration (hg_set_settle_rule); settle_window 0 keeps the
  // window the operators' time gives (settle_rounds).
  int32_t check_rounds = kHeatCheckRounds;
  double change_tolerance = kHeatChangeTolerance;
  int32_t settle_window = 0;
  // Speculative refinement (FP64 factors): the verdicts of the first rounds of the heat, pinned and
  // Poisson solves of a call, [system][chunk], read at the call's last synchronization.
  PinnedArray<int32_t> h_speculate;
  bool speculating = false;
  int64_t spec_chunks = 0;
  // Refinement targets of the factored solves, [precision][0 heat (componentwise), 1 Poisson
  // (normwise)].
  double refine_tolerance[2][2] = {{kSettleDirectTolerance, kPoissonTolerance},
                                   {kMixedHeatTolerance, kMixedPoissonTolerance}};
  // Workspaces.
  CgWorkspace cg;
  MarchWorkspace march;  // the band march of a system refused under the budget
  // The Poisson system's multigrid hierarchy, built on its first solve past the budget.
  std::unique_ptr<Multigrid> poisson_mg;
  bool poisson_mg_tried = false;
  SolveWorkspace solve;
  SolveWorkspace solve_side;  // the rim-pinned heat solve's, when it runs beside the Neumann one
  bool on_side = false;       // the solves being enqueued use solve_side
  DeviceArray<double> fields, shift_partials, shift_values, null_partials;
  NullSpace null_space;
  DeviceArray<int32_t> d_sources, invalid;
  PinnedArray<int32_t> h_invalid;  // the seeds' range flag, read back with the call's last sync
  DeviceArray<int64_t> d_owner, d_offsets;
  PinnedArray<int64_t> h_stage;  // the owners and offsets of a call, staged for an async upload
  bool sources_on_device = false;  // d_offsets[k] holds this call's source count
  // The distance's destination per 64-column chunk, read by the finishing kernel (so a replayed
  // graph writes to each call's buffer), staged in pinned memory.
  DeviceArray<double*> d_slots;
  PinnedArray<double*> h_slots;
  // Replay of the factored FP64 path: one instantiated graph per column count, valid while every
  // buffer it names is the one it was captured with (signature) and no state change intervened.
  struct Graph {
    hipGraphExec_t exec = nullptr;
    std::vector<uintptr_t> signature;
    hg_solve_stats stats{};
    Graph() = default;
    Graph(const Graph&) = delete;
    Graph& operator=(const Graph&) = delete;
    ~Graph() {
      if (exec) (void)hipGraphExecDestroy(exec);
    }
  };
  std::map<int64_t, std::unique_ptr<Graph>> graphs;
  bool use_graphs = true;
  bool graphs_broken = false;  // a capture failed: this mesh runs eagerly from then on
  int64_t field_capacity = 0;
  DeviceArray<double> packed_phi;  // packed meshes: the solved slot columns before extraction
  int64_t packed_capacity = 0;
  DeviceArray<int64_t> q_src_off, q_dst_off, q_len;  // packed meshes: each query's field
  DeviceArray<int32_t> q_local;                       // packed meshes: sources in mesh numbering
  hg_solve_stats stats{};
  PhaseTimer timer;
};

namespace {

template <typename F>
int guarded(F&& body) {
  try {
    body();
    return HG_OK;
  } catch (const Error& e) {
    set_error(e.message);
    return e.status;
  } catch (const std::bad_alloc&) {
    set_error("host out of memory");
    return HG_ERR_RUNTIME;
  } catch (const std::exception& e) {
    set_error(e.what());
    return HG_ERR_RUNTIME;
  }
}

void require_mesh(hg_mesh* mesh) {
  HG_REQUIRE(mesh != nullptr, HG_ERR_RUNTIME, "the mesh handle is null (released or never created)");
}

// Packed meshes solve through the factors only: the settled iteration's window and stopping rule
// are per mesh (heat.py:5234-5243), which one packed iteration cannot honour.
void require_single_or_factored(const hg_mesh& m, int32_t mode) {
  HG_REQUIRE(m.segments() == 1 || mode == HG_MODE_FACTORED, HG_ERR_VALUE,
             "packed meshes solve in HG_MODE_FACTORED only");
}

// Drops every captured graph: the state they were captured against is changing.
void invalidate_graphs(hg_mesh& m) { m.graphs.clear(); }

// Every device byte the mesh holds: its copy of the mesh and pattern, the operators, the plan and
// factors, and the per-call workspaces.
int64_t mesh_device_bytes(const hg_mesh& m) {
  int64_t total = 0;
  for (const DeviceArray<int32_t>* a :
       {&m.row_segment, &m.faces, &m.offsets, &m.columns, &m.corner_ptr, &m.corner_slots,
        &m.vertex_ptr, &m.vertex_corners, &m.d_rim, &m.d_rim_rows, &m.d_labels, &m.comp_ptr,
        &m.comp_members, &m.poisson_pinned_new, &m.poisson_dropped_old, &m.heat_pinned_new,
        &m.heat_dropped_old, &m.faces_rows, &m.row_corner_ptr, &m.d_sources, &m.invalid,
        &m.q_local})
    total += a->bytes();
  for (const DeviceArray<float>* a : {&m.xyz, &m.cot, &m.normals, &m.areas, &m.heat32, &m.voronoi,
                                      &m.pinned32})
    total += a->bytes();
  for (const DeviceArray<double>* a :
       {&m.mass, &m.lap, &m.heat, &m.poisson, &m.heat_inv, &m.poisson_inv, &m.scalars,
        &m.edge_partials, &m.face_geom, &m.pinned, &m.pinned_inv, &m.pinned_rhs, &m.reg,
        &m.reg_inv, &m.fields,
        &m.shift_partials, &m.shift_values, &m.null_partials, &m.packed_phi})
    total += a->bytes();
  for (const DeviceArray<int64_t>* a : {&m.d_owner, &m.d_offsets, &m.q_src_off, &m.q_dst_off,
                                        &m.q_len})
    total += a->bytes();
  total += m.row_corners.bytes() + m.d_slots.bytes() + m.march.bytes();
  if (m.poisson_mg) total += m.poisson_mg->bytes();
  total += m.heat_factor.bytes() + m.poisson_factor.bytes() + m.pinned_factor.bytes() +
           m.reg_factor.bytes();
  if (m.plan) total += m.plan->bytes();
  total += m.cg.bytes() + m.solve.bytes() + m.solve_side.bytes() + m.null_space.bytes();
  return total;
}

// Frees what solving built beyond the operators: the per-call workspaces (they grow to the widest
// batch a mesh has served and are rebuilt by the next solve).
void release_workspaces(hg_mesh& m) {
  invalidate_graphs(m);
  m.cg.release();
  m.march.release();
  m.poisson_mg.reset();
  m.poisson_mg_tried = false;
  m.solve.release();
  m.solve_side.release();
  m.fields.reset();
  m.field_capacity = 0;
  m.packed_phi.reset();
  m.packed_capacity = 0;
  m.pinned_rhs.reset();
  m.shift_partials.reset();
  m.shift_values.reset();
  m.null_partials.reset();
}

// ---- operators --------------------------------------------------------------------------------

// The mean unique-edge length of segment s: each edge once, over the pattern's strict upper
// triangle.
double host_mean_unique_edge(const hg_mesh& m, int64_t s) {
  double total = 0.0;
  int64_t count = 0;
  for (int64_t i = m.seg_v[s]; i < m.seg_v[s + 1]; ++i) {
    for (int32_t k = m.pattern.offsets[i]; k < m.pattern.offsets[i + 1]; ++k) {
      int32_t j = m.pattern.columns[k];
      if (j <= i) continue;
      const float* a = &m.h_xyz[3 * i];
      const float* b = &m.h_xyz[3 * j];
      float dx = b[0] - a[0], dy = b[1] - a[1], dz = b[2] - a[2];
      total += static_cast<double>(sqrtf(dx * dx + dy * dy + dz * dz));
      ++count;
    }
  }
  return count > 0 ? total / static_cast<double>(count) : 0.0;
}

double host_mean_face_edge(const hg_mesh& m, int64_t s) {
  const int64_t f0 = m.seg_f[s], f1 = m.seg_f[s + 1];
  if (f1 == f0) return 0.0;
  double total = 0.0;
  for (int64_t f = f0; f < f1; ++f) {
    for (int e = 0; e < 3; ++e) {
      const float* a = &m.h_xyz[3 * m.h_faces[3 * f + e]];
      const float* b = &m.h_xyz[3 * m.h_faces[3 * f + (e + 1) % 3]];
      float dx = b[0] - a[0], dy = b[1] - a[1], dz = b[2] - a[2];
      total += static_cast<double>(sqrtf(dx * dx + dy * dy + dz * dz));
    }
  }
  return total / static_cast<double>(3 * (f1 - f0));
}

double host_diffusion_floor(const hg_mesh& m, int64_t s) {
  if (m.seg_v[s + 1] == m.seg_v[s]) return 0.0;
  float lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
  for (int64_t v = m.seg_v[s]; v < m.seg_v[s + 1]; ++v)
    for (int d = 0; d < 3; ++d) {
      lo[d] = std::min(lo[d], m.h_xyz[3 * v + d]);
      hi[d] = std::max(hi[d], m.h_xyz[3 * v + d]);
    }
  double squares = 0.0;
  for (int d = 0; d < 3; ++d) {
    double e = static_cast<double>(hi[d]) - static_cast<double>(lo[d]);
    squares += e * e;
  }
  double diag = std::sqrt(squares);
  return (diag / kDiffusionDecay) * (diag / kDiffusionDecay);
}

void assemble(hg_mesh& m, double t_request, int32_t method, bool solver_default,
              bool robust = false) {
  const hipStream_t st = m.stream;
  const int64_t n = m.n, nnz = static_cast<int64_t>(m.pattern.columns.size());
  const int32_t face_edges = method == HG_METHOD_MESHLAB ? 1 : 0;
  // The time rule of HeatSolver._vertex_time (heat.py:3714-3736) and _settle_window (5234-5243),
  // per segment: each packed mesh keeps the time it would have alone.
  const int64_t n_seg = m.segments();
  std::vector<double> t_values(n_seg);
  std::vector<int32_t> from_edges(n_seg, 0);
  for (int64_t s = 0; s < n_seg; ++s) {
    const double mean = face_edges ? host_mean_face_edge(m, s) : host_mean_unique_edge(m, s);
    const double def = mean * mean;
    double t_value = t_request;
    if (std::isnan(t_request)) {
      const double floor = solver_default ? host_diffusion_floor(m, s) : 0.0;
      if (floor > def) t_value = floor;
      else from_edges[s] = 1;
    }
    t_values[s] = t_value;
    if (s == 0) {
      if (from_edges[s] || !(t_value > def && def > 0.0)) m.settle_rounds = kHeatSettleRounds;
      else
        m.settle_rounds =
            static_cast<int32_t>(std::ceil(kHeatSettleRounds * std::sqrt(t_value / def)));
      m.time_from_edges = from_edges[s];
    }
  }

  m.method = method;
  m.cot.allocate(3 * m.f);
  m.normals.allocate(3 * m.f);
  m.areas.allocate(m.f);
  m.face_geom.allocate(kFaceGeom * m.f);
  m.mass.allocate(n);
  m.lap.allocate(nnz);
  m.heat.allocate(nnz);
  m.poisson.allocate(nnz);
  m.heat32.allocate(nnz);
  m.heat_inv.allocate(n);
  m.poisson_inv.allocate(n);
  m.scalars.allocate(4 * n_seg);
  int64_t longest = 0;
  for (int64_t s = 0; s < n_seg; ++s) longest = std::max(longest, m.seg_v[s + 1] - m.seg_v[s]);
  m.edge_partials.allocate(2 * edge_sum_blocks(longest));
  launch_face_terms(m.xyz.ptr, m.faces.ptr, m.f, m.cot.ptr, m.normals.ptr, m.areas.ptr, st);
  // use_robust: the half-cotangents (the Laplacian's and the divergence's) from mollified side
  // lengths, the areas, normals, mass and time from the positions (heat.py:2041-2052).
  m.robust = robust;
  m.mollify_delta = robust ? launch_mollified_cot(m.xyz.ptr, m.faces.ptr, m.f, m.cot.ptr, st) : 0.0;
  launch_face_geometry(m.xyz.ptr, m.faces.ptr, m.normals.ptr, m.areas.ptr, m.cot.ptr, m.f,
                       m.face_geom.ptr, st);
  if (method == HG_METHOD_IGL) {
    m.voronoi.allocate(3 * m.f);
    launch_voronoi_shares(m.xyz.ptr, m.faces.ptr, m.f, m.voronoi.ptr, st);
  }
  launch_mass(m.vertex_ptr.ptr, m.vertex_corners.ptr, m.areas.ptr,
              method == HG_METHOD_IGL ? m.voronoi.ptr : nullptr, n, m.mass.ptr, st);
  // One edge-sum reduction per segment, its blocks counted from the segment's first row: the sums
  // (and so the times) are those the segment's mesh forms alone. The launches share the partials
  // buffer in stream order.
  for (int64_t s = 0; s < n_seg; ++s)
    launch_edge_sums(m.offsets.ptr, m.columns.ptr, m.corner_ptr.ptr, m.xyz.ptr, m.seg_v[s],
                     m.seg_v[s + 1] - m.seg_v[s], face_edges, m.edge_partials.ptr, from_edges[s],
                     from_edges[s] ? 0.0 : t_values[s], m.scalars.ptr + 4 * s, st);
  launch_operator_rows(m.offsets.ptr, m.columns.ptr, m.corner_ptr.ptr, m.corner_slots.ptr,
                       m.cot.ptr, m.mass.ptr, m.scalars.ptr, n_seg > 1 ? m.row_segment.ptr : nullptr,
                       n, m.lap.ptr, m.heat.ptr, m.poisson.ptr, m.heat32.ptr, m.heat_inv.ptr,
                       m.poisson_inv.ptr, st);
  std::vector<double> host_scalars(4 * n_seg, 0.0);
  HG_HIP(hipMemcpyAsync(host_scalars.data(), m.scalars.ptr, sizeof(double) * 4 * n_seg,
                        hipMemcpyDeviceToHost, st));
  HG_HIP(hipStreamSynchronize(st));
  m.seg_t.resize(n_seg);
  for (int64_t s = 0; s < n_seg; ++s) m.seg_t[s] = host_scalars[4 * s + 3];
  m.t = m.seg_t[0];
  m.assembled = true;
  m.poisson_bound = -1.0;
  m.heat_factor.reset();
  m.poisson_factor.reset();
  m.pinned_factor.reset();
  m.pinned_built = false;
  m.reg_factor.reset();
  m.reg_built = false;
  m.reg_bound = -1.0;
  invalidate_graphs(m);
}

void ensure_operators(hg_mesh& m) {
  if (!m.assembled) assemble(m, std::numeric_limits<double>::quiet_NaN(), HG_METHOD_MESHLAB, true);
}

// ---- pins and the factorization -----------------------------------------------------------------

void ensure_plan(hg_mesh& m) {
  if (m.plan) return;
  m.host_plan = std::make_unique<Plan>(analyse(m.pattern, m.h_xyz.data(), m.options));
  const Plan& plan = *m.host_plan;
  auto device = std::make_unique<DevicePlan>();
  device->upload(plan, m.pattern, m.split_target, m.stream);
  {
    std::vector<int32_t> rows(m.h_faces.size());
    parallel_chunks(static_cast<int64_t>(rows.size()), 1 << 16, [&](int64_t b, int64_t e, int64_t) {
      for (int64_t i = b; i < e; ++i) rows[i] = plan.iperm[m.h_faces[i]];
    });
    m.faces_rows.upload(rows, m.stream);
    // Row r's corners are vertex perm[r]'s, in its ascending corner order: the row offsets in
    // one pass, then each chunk of rows fills its own records.
    std::vector<int32_t> ptr(static_cast<size_t>(m.n) + 1, 0);
    for (int64_t r = 0; r < m.n; ++r) {
      const int32_t v = plan.perm[r];
      ptr[r + 1] = ptr[r] + (m.pattern.vertex_ptr[v + 1] - m.pattern.vertex_ptr[v]);
    }
    std::vector<int4> corners(m.pattern.vertex_corners.size());
    parallel_chunks(m.n, 1 << 14, [&](int64_t b, int64_t e, int64_t) {
      for (int64_t r = b; r < e; ++r) {
        const int32_t v = plan.perm[r];
        int32_t at = ptr[r];
        for (int32_t k = m.pattern.vertex_ptr[v]; k < m.pattern.vertex_ptr[v + 1]; ++k) {
          const int32_t c = m.pattern.vertex_corners[k], f = c / 3;
          corners[at++] = make_int4(c, rows[3 * f], rows[3 * f + 1], rows[3 * f + 2]);
        }
      }
    });
    m.row_corners.upload(corners, m.stream);
    m.row_corner_ptr.upload(ptr, m.stream);
  }
  if (m.n_rim > 0) {
    std::vector<int32_t> rim_rows(static_cast<size_t>(m.n), 0);
    for (int64_t v = 0; v < m.n; ++v) rim_rows[plan.iperm[v]] = m.rim[v];
    m.d_rim_rows.upload(rim_rows, m.stream);
  }
  // The distance's finishing pass reads components as row ranges of the elimination order.
  m.rows_by_component = true;
  for (int64_t v = 0; v < m.n && m.rows_by_component; ++v) {
    const int32_t c = m.labels[v], row = plan.iperm[v];
    m.rows_by_component = row >= m.comp_ptr_host[c] && row < m.comp_ptr_host[c + 1];
  }
  // Pins: the Poisson system's one row per component, its last-eliminated vertex
  // (kernels/cholesky.py:195-214); the heat system's empty rows (unreferenced vertices).
  const int64_t n = m.n;
  std::vector<int32_t> last(m.n_comp, -1);
  for (int64_t v = 0; v < n; ++v) {
    int32_t c = m.labels[v];
    if (last[c] < 0 || plan.iperm[v] > plan.iperm[last[c]]) last[c] = static_cast<int32_t>(v);
  }
  std::vector<int32_t> pinned_new(n, 0), dropped_old(n, 0);
  std::vector<int32_t> heat_pinned_new(n, 0), heat_dropped_old(n, 0);
  auto front_slot = [&](int32_t new_index) {
    auto it = std::upper_bound(plan.col_begin.begin(), plan.col_begin.end(), new_index);
    int32_t s = static_cast<int32_t>(it - plan.col_begin.begin()) - 1;
    int64_t m_s = plan.ncol[s] + plan.nrow[s];
    int64_t local = new_index - plan.col_begin[s];
    return plan.front_off[s] + local * m_s + local;
  };
  m.poisson_pin_slots.clear();
  m.heat_pin_slots.clear();
  for (int32_t c = 0; c < m.n_comp; ++c) {
    int32_t v = last[c];
    pinned_new[plan.iperm[v]] = 1;
    dropped_old[v] = 1;
    m.poisson_pin_slots.push_back(front_slot(plan.iperm[v]));
  }
  m.has_heat_pins = false;
  for (int64_t v = 0; v < n; ++v) {
    if (m.pattern.diagonal[v] < 0) {
      heat_pinned_new[plan.iperm[v]] = 1;
      heat_dropped_old[v] = 1;
      m.heat_pin_slots.push_back(front_slot(plan.iperm[v]));
      m.has_heat_pins = true;
    }
  }
  m.poisson_pinned_new.upload(pinned_new, m.stream);
  m.poisson_dropped_old.upload(dropped_old, m.stream);
  m.heat_pinned_new.upload(heat_pinned_new, m.stream);
  m.heat_dropped_old.upload(heat_dropped_old, m.stream);
  HG_HIP(hipStreamSynchronize(m.stream));
  m.plan = std::move(device);
}

// HG_METHOD_POTPOURRI3D's Poisson system -L + 1e-6 I over the Poisson pattern (built on first
// use, so its factorization shares the heat analysis; heat.py:4097-4108).
void ensure_regularized_system(hg_mesh& m) {
  ensure_operators(m);
  if (m.reg_built) return;
  const int64_t nnz = static_cast<int64_t>(m.pattern.columns.size());
  m.reg.allocate(std::max<int64_t>(nnz, 1));
  m.reg_inv.allocate(std::max<int64_t>(m.n, 1));
  launch_shifted_rows(m.offsets.ptr, m.columns.ptr, m.poisson.ptr, m.n, kPotpourriPoissonShift,
                      m.reg.ptr, m.reg_inv.ptr, m.stream);
  m.reg_built = true;
  m.reg_bound = -1.0;
  invalidate_graphs(m);
}

// The rim-pinned heat system over the heat system's pattern (meshes with a rim only).
void ensure_pinned_system(hg_mesh& m) {
  ensure_operators(m);
  if (m.pinned_built) return;
  const int64_t nnz = static_cast<int64_t>(m.pattern.columns.size());
  m.pinned.allocate(nnz);
  m.pinned32.allocate(nnz);
  m.pinned_inv.allocate(m.n);
  launch_pinned_rows(m.offsets.ptr, m.columns.ptr, m.d_rim.ptr, m.heat.ptr, m.n, m.pinned.ptr,
                     m.pinned32.ptr, m.pinned_inv.ptr, m.stream);
  m.pinned_built = true;
  invalidate_graphs(m);
}

// A factorization's peak device memory: the transient front arena plus the blocks it keeps.
int64_t factor_peak_bytes(const DevicePlan& plan, bool keep_fp32) {
  return 8 * plan.arena_size + (keep_fp32 ? 12 : 8) * plan.w_size + 8 * plan.nnz_w;
}

// Refuse (HG_ERR_BUDGET) a factorization whose peak exceeds the budget, as sparse_cholesky refuses
// one past CHOLESKY_MEMORY_BUDGET (cholesky.py:72, linalg.py:2393-2420); the system is then solved
// by iteration. With no budget set the limit is 90 % of the device memory free now.
void require_budget(hg_mesh& m, int32_t system, bool keep_fp32) {
  const int64_t need = factor_peak_bytes(*m.plan, keep_fp32);
  int64_t limit = m.budget;
  if (limit <= 0) {
    size_t free_bytes = 0, total_bytes = 0;
    HG_HIP(hipMemGetInfo(&free_bytes, &total_bytes));
    limit = static_cast<int64_t>(0.9 * static_cast<double>(free_bytes));
  }
  if (need > limit) m.refused |= system;
  HG_REQUIRE(need <= limit, HG_ERR_BUDGET,
             "the factorization needs " + std::to_string(need) + " bytes at its peak, over the " +
                 std::to_string(limit) + "-byte budget");
}

bool is_budget_refusal(const Error& e) { return e.status == HG_ERR_BUDGET; }

// Whether `system`'s factorization would be refused, asked of the analysis before any iteration
// (OperatorFactorization.exceeds_budget, linalg.py:2393-2420): with no budget set only a system
// already refused is; with one set, the plan's peak decides and a refusal is remembered.
bool exceeds_budget(hg_mesh& m, int32_t system) {
  if (m.refused & system) return true;
  if (m.budget <= 0 || !m.budget_probe) return false;
  ensure_plan(m);
  if (factor_peak_bytes(*m.plan, false) <= m.budget) return false;
  m.refused |= system;
  return true;
}

void factor_system(hg_mesh& m, int32_t system, bool keep_fp32) {
  ensure_operators(m);
  ensure_plan(m);
  if (system == HG_SYSTEM_PINNED) {
    if (m.n_rim == 0) return;
    ensure_pinned_system(m);
    PhaseScope scope(&m.timer, kPhaseFactor, m.stream);
    if (!m.pinned_factor.held() || (keep_fp32 && !m.pinned_factor.w32.ptr)) {
      require_budget(m, system, keep_fp32);
      invalidate_graphs(m);
      factor_numeric(*m.plan, m.pinned.ptr, m.has_heat_pins ? m.heat_pinned_new.ptr : nullptr,
                     m.heat_pin_slots, m.pinned_factor, keep_fp32, m.stream);
    }
    return;
  }
  if (system == HG_SYSTEM_REGULARIZED) {
    ensure_regularized_system(m);
    PhaseScope scope(&m.timer, kPhaseFactor, m.stream);
    if (!m.reg_factor.held() || (keep_fp32 && !m.reg_factor.w32.ptr)) {
      require_budget(m, system, keep_fp32);
      invalidate_graphs(m);
      factor_numeric(*m.plan, m.reg.ptr, m.has_heat_pins ? m.heat_pinned_new.ptr : nullptr,
                     m.heat_pin_slots, m.reg_factor, keep_fp32, m.stream);
    }
    return;
  }
  Factor& target = system == HG_SYSTEM_HEAT ? m.heat_factor : m.poisson_factor;
  if (!target.held() || (keep_fp32 && !target.w32.ptr)) {
    require_budget(m, system, keep_fp32);
    invalidate_graphs(m);
  }
  PhaseScope scope(&m.timer, kPhaseFactor, m.stream);
  if (system == HG_SYSTEM_HEAT) {
    if (!m.heat_factor.held())
      factor_numeric(*m.plan, m.heat.ptr, m.has_heat_pins ? m.heat_pinned_new.ptr : nullptr,
                     m.heat_pin_slots, m.heat_factor, keep_fp32, m.stream);
    else if (keep_fp32 && !m.heat_factor.w32.ptr)
      factor_numeric(*m.plan, m.heat.ptr, m.has_heat_pins ? m.heat_pinned_new.ptr : nullptr,
                     m.heat_pin_slots, m.heat_factor, true, m.stream);
  } else {
    if (!m.poisson_factor.held() || (keep_fp32 && !m.poisson_factor.w32.ptr))
      factor_numeric(*m.plan, m.poisson.ptr, m.poisson_pinned_new.ptr, m.poisson_pin_slots,
                     m.poisson_factor, keep_fp32, m.stream);
  }
}

// ---- solves -----------------------------------------------------------------------------------

// The heat system (Neumann on a rim) or the one pinned to zero on the rim.
struct HeatSystem {
  const double* values;
  const float* values32;
  const double* inv_diag;
  Factor* factor;
  int32_t code;  // HG_SYSTEM_HEAT or HG_SYSTEM_PINNED
};

HeatSystem heat_system(hg_mesh& m, bool pinned) {
  if (pinned) {
    ensure_pinned_system(m);
    return {m.pinned.ptr, m.pinned32.ptr, m.pinned_inv.ptr, &m.pinned_factor, HG_SYSTEM_PINNED};
  }
  return {m.heat.ptr, m.heat32.ptr, m.heat_inv.ptr, &m.heat_factor, HG_SYSTEM_HEAT};
}

int32_t diffuse_iterated(hg_mesh& m, const HeatSystem& sys, const double* rhs, double* x,
                         int64_t k) {
  PhaseScope scope(&m.timer, kPhaseCg, m.stream);
  int32_t total = 0;
  for (int64_t q0 = 0; q0 < k; q0 += 64) {
    int64_t kk = std::min<int64_t>(64, k - q0);
    CgProblem p;
    p.n = m.n;
    p.n_columns = kk;
    p.offsets = m.offsets.ptr;
    p.columns = m.columns.ptr;
    p.values32 = sys.values32;
    p.values64 = sys.values;
    p.inv_diag = sys.inv_diag;
    p.rhs = rhs + q0 * m.n;
    p.x = x + q0 * m.n;
    p.settle = 1;
    p.check_rounds = m.check_rounds;
    p.change_tolerance = m.change_tolerance;
    p.settle_rounds = m.settle_window > 0 ? m.settle_window : m.settle_rounds;
    total = std::max(total, run_cg(p, m.cg, m.stream));
  }
  return total;
}

bool backward_error_fails(hg_mesh& m, const double* values, const double* x, const double* b,
                          int64_t k) {
  if (m.invalid.size < 1) m.invalid.allocate(1);
  HG_HIP(hipMemsetAsync(m.invalid.ptr, 0, sizeof(int32_t), m.stream));
  launch_backward_error(m.offsets.ptr, m.columns.ptr, values, x, b, m.n, k, kSettleBackwardError,
                        m.invalid.ptr, m.stream);
  int32_t flag = 0;
  HG_HIP(hipMemcpyAsync(&flag, m.invalid.ptr, sizeof(int32_t), hipMemcpyDeviceToHost, m.stream));
  HG_HIP(hipStreamSynchronize(m.stream));
  return flag != 0;
}

int32_t diffuse_factored(hg_mesh& m, const HeatSystem& sys, double* rhs, double* x, int64_t k,
                         int32_t precision, FieldLayout layout = FieldLayout::kColumns) {
  const bool mixed = precision == HG_PRECISION_MIXED;
  factor_system(m, sys.code, mixed);
  int32_t* speculate =
      m.speculating ? m.h_speculate.ptr + (sys.code == HG_SYSTEM_PINNED ? m.spec_chunks : 0)
                    : nullptr;
  return refined_solve(*m.plan, *sys.factor, mixed, rhs, x, k,
                       m.refine_tolerance[mixed ? 1 : 0][0], true,
                       mixed ? kMixedRefineRounds : kRefineRounds,
                       m.has_heat_pins ? m.heat_pinned_new.ptr : nullptr,
                       m.on_side ? m.solve_side : m.solve, m.stream,
                       &m.timer, kPhaseHeatSweep, layout, layout, speculate);
}

// The settled iterate, verified, falling back to the factorization (solve_spd_settled,
// linalg.py:2659-2727), of the heat system or (pinned) the rim-pinned one. Returns the flags of
// hg_diffuse; the solve's rounds or refinement solves go to the call's statistics. Fields stored
// by rows (FieldLayout::kRows) go through the factors only.
int32_t diffuse(hg_mesh& m, double* rhs, double* x, int64_t k, int32_t mode,
                int32_t precision, FieldLayout layout = FieldLayout::kColumns,
                bool pinned = false) {
  ensure_operators(m);
  const HeatSystem sys = heat_system(m, pinned);
  int32_t flags = 0;
  int32_t rounds = 0, refinements = 0;
  const MarchSystem marched{m.n, m.offsets.ptr, m.columns.ptr, sys.values};
  auto march = [&] {
    HG_REQUIRE(layout == FieldLayout::kColumns, HG_ERR_RUNTIME, "a marched solve takes columns");
    PhaseScope scope(&m.timer, kPhaseCg, m.stream);
    rounds = march_settled(marched, rhs, x, k, m.march, m.stream);
    flags |= 4;
    m.stats.heat_marched = 1;
  };
  if (mode == HG_MODE_FACTORED || sys.factor->held()) {
    try {
      refinements = diffuse_factored(m, sys, rhs, x, k, precision, layout);
    } catch (const Error& e) {
      // Refused under the budget before anything was solved: marched instead.
      if (!is_budget_refusal(e) || layout != FieldLayout::kColumns) throw;
      march();
    }
  } else if (exceeds_budget(m, sys.code)) {
    // Past the budget the march costs less than the settled iteration, whether its iterate would
    // pass the check or not: marched at once (linalg.py:2646-2651).
    march();
  } else {
    HG_REQUIRE(layout == FieldLayout::kColumns, HG_ERR_RUNTIME, "an iterated solve takes columns");
    rounds = diffuse_iterated(m, sys, rhs, x, k);
    if (backward_error_fails(m, sys.values, x, rhs, k)) {
      flags |= 1;
      m.stats.backward_error_failed = 1;
      const bool was_held = sys.factor->held();
      try {
        refinements = diffuse_factored(m, sys, rhs, x, k, HG_PRECISION_FP64);
        if (!was_held) flags |= 2;
      } catch (const Error& e) {
        // The fallback refused under the budget before anything was solved: marched, which
        // reaches every entry with its own precision too (linalg.py:2710-2716). The call's rounds
        // stay the iteration's, as the reference returns them.
        if (!is_budget_refusal(e)) throw;
        const int32_t iterated = rounds;
        march();
        rounds = iterated;
      }
    }
  }
  if (pinned) {
    m.stats.pinned_solves = refinements > 0 ? refinements : rounds;
  } else {
    m.stats.heat_rounds = rounds;
    m.stats.heat_refinements = refinements;
  }
  return flags;
}

// The largest Gershgorin ratio of D^-1 A over a Poisson-type system's rows, reduced once.
double dominance_bound(hg_mesh& m, const double* values, double& cached) {
  if (cached < 0.0) {
    int32_t blocks = static_cast<int32_t>((std::max<int64_t>(m.n, 1) + kBlock - 1) / kBlock);
    DeviceArray<double> partials(blocks), out(1);
    launch_dominance_bound(m.offsets.ptr, m.columns.ptr, values, m.n, partials.ptr, out.ptr,
                           m.stream);
    double bound = 0.0;
    HG_HIP(hipMemcpyAsync(&bound, out.ptr, sizeof(double), hipMemcpyDeviceToHost, m.stream));
    HG_HIP(hipStreamSynchronize(m.stream));
    cached = bound;
  }
  return cached;
}

// The right-hand side and the solution may be stored by rows (factored only). A solution stored by
// columns has each component's mean removed here; one stored by rows is left for finish_rows.
// `regularized` solves HG_METHOD_POTPOURRI3D's -L + 1e-6 I instead of -L: no null space, its
// empty rows pinned as the heat system's are, the same normwise tolerance (heat.py:4127-4151).
void solve_poisson(hg_mesh& m, double* rhs, double* x, int64_t k, int32_t mode,
                   int32_t precision, FieldLayout rhs_layout = FieldLayout::kColumns,
                   FieldLayout x_layout = FieldLayout::kColumns, bool regularized = false) {
  ensure_operators(m);
  if (regularized) ensure_regularized_system(m);
  const bool mixed = precision == HG_PRECISION_MIXED;
  const int32_t system = regularized ? HG_SYSTEM_REGULARIZED : HG_SYSTEM_POISSON;
  Factor& factor = regularized ? m.reg_factor : m.poisson_factor;
  const double* values = regularized ? m.reg.ptr : m.poisson.ptr;
  const double* inv_diag = regularized ? m.reg_inv.ptr : m.poisson_inv.ptr;
  const int32_t* pins = regularized ? (m.has_heat_pins ? m.heat_pinned_new.ptr : nullptr)
                                    : m.poisson_pinned_new.ptr;
  bool factored = mode == HG_MODE_FACTORED || factor.held() || m.heat_factor.held();
  if (factored && rhs_layout == FieldLayout::kColumns && x_layout == FieldLayout::kColumns) {
    // By columns a refusal under the budget leaves the solve to the iteration.
    try {
      factor_system(m, system, mixed);
    } catch (const Error& e) {
      if (!is_budget_refusal(e)) throw;
      factored = false;
    }
  }
  if (factored) {
    factor_system(m, system, mixed);
    m.stats.poisson_refinements = refined_solve(
        *m.plan, factor, mixed, rhs, x, k, m.refine_tolerance[mixed ? 1 : 0][1], false,
        mixed ? kMixedRefineRounds : kRefineRounds, pins, m.solve, m.stream, &m.timer,
        kPhasePoissonSweep, rhs_layout, x_layout,
        m.speculating ? m.h_speculate.ptr + 2 * m.spec_chunks : nullptr);
    m.stats.poisson_rounds = 0;
  } else {
    HG_REQUIRE(rhs_layout == FieldLayout::kColumns && x_layout == FieldLayout::kColumns,
               HG_ERR_RUNTIME, "an iterated solve takes columns");
    // No factorization past the budget (the Poisson system's, or the mesh's heat system refused):
    // a multigrid hierarchy's rounds, unlike the polynomial's, do not grow with the mesh
    // (heat.py:4141-4149). An operator that does not coarsen keeps the polynomial.
    const bool unfactorable =
        !regularized &&
        (exceeds_budget(m, HG_SYSTEM_POISSON) || (m.refused & HG_SYSTEM_HEAT) != 0);
    if (unfactorable && !m.poisson_mg && !m.poisson_mg_tried) {
      m.poisson_mg_tried = true;
      HostCsr fine;
      fine.rows = fine.cols = m.n;
      fine.ptr = m.pattern.offsets;
      fine.idx = m.pattern.columns;
      fine.val.resize(m.pattern.columns.size());
      if (!fine.val.empty())
        HG_HIP(hipMemcpyAsync(fine.val.data(), m.poisson.ptr, sizeof(double) * fine.val.size(),
                              hipMemcpyDeviceToHost, m.stream));
      HG_HIP(hipStreamSynchronize(m.stream));
      m.poisson_mg = build_multigrid(fine, 0, m.stream);
    }
    if (unfactorable && m.poisson_mg) {
      PhaseScope scope(&m.timer, kPhaseCg, m.stream);
      m.stats.poisson_rounds =
          multigrid_cg(*m.poisson_mg, rhs, x, k, kPoissonTolerance, 10 * m.n, m.stream);
      m.stats.poisson_refinements = 0;
      m.stats.poisson_multigrid = 1;
    } else {
      // The Jacobi-Chebyshev interval (linalg.py:6074-6090): [max(min(80/n, 0.5), 1 - r), 1 + r],
      // r the largest Gershgorin ratio of D^-1 A, at least 1e-3.
      double r = std::max(dominance_bound(m, values, regularized ? m.reg_bound : m.poisson_bound),
                          1e-3);
      double lo = std::max(std::min(kChebyshevInterval / static_cast<double>(m.n), kChebyshevCap),
                           1.0 - r);
      double hi = 1.0 + r;
      PhaseScope scope(&m.timer, kPhaseCg, m.stream);
      int32_t rounds = 0;
      for (int64_t q0 = 0; q0 < k; q0 += 64) {
        int64_t kk = std::min<int64_t>(64, k - q0);
        CgProblem p;
        p.n = m.n;
        p.n_columns = kk;
        p.offsets = m.offsets.ptr;
        p.columns = m.columns.ptr;
        p.values64 = values;
        p.inv_diag = inv_diag;
        p.rhs = rhs + q0 * m.n;
        p.x = x + q0 * m.n;
        p.settle = 0;
        p.tolerance = kPoissonTolerance;
        p.chebyshev_steps = kChebyshevSteps;
        p.lambda_min = lo;
        p.lambda_max = hi;
        rounds = std::max(rounds, run_cg(p, m.cg, m.stream));
      }
      m.stats.poisson_rounds = rounds;
      m.stats.poisson_refinements = 0;
    }
  }
  if (x_layout == FieldLayout::kColumns && !regularized)
    remove_null_space(m.null_space, m.comp_ptr.ptr, m.comp_members.ptr, m.n, k, x, m.null_partials,
                      m.stream);
}

// Per-call buffers grow to at least twice their size and never below 1024 entries: a caller
// whose batches vary reallocates (a free waits on every stream's work) a few times, not per call.
template <typename Array>
void reserve(Array& a, int64_t need) {
  if (a.size < need) a.allocate(std::max<int64_t>({need, 2 * a.size, 1024}));
}

// Each source entry's query and the queries' offsets, staged in pinned memory so both uploads stay
// asynchronous (the stage is rewritten only by the next call, after this one's synchronization).
void upload_sources(hg_mesh& m, const int32_t* d_sources, const int64_t* offsets, int64_t k) {
  const int64_t total = offsets[k] - offsets[0];
  const int64_t owners = std::max<int64_t>(total, 1);
  reserve(m.h_stage, owners + k + 1);
  int64_t* owner = m.h_stage.ptr;
  int64_t* rel_offsets = m.h_stage.ptr + owners;
  owner[0] = 0;
  for (int64_t q = 0; q < k; ++q)
    for (int64_t e = offsets[q]; e < offsets[q + 1]; ++e) owner[e - offsets[0]] = q;
  for (int64_t q = 0; q <= k; ++q) rel_offsets[q] = offsets[q] - offsets[0];
  reserve(m.d_owner, owners);
  reserve(m.d_offsets, k + 1);
  m.d_owner.upload(owner, owners, m.stream);
  m.d_offsets.upload(rel_offsets, k + 1, m.stream);
  m.sources_on_device = true;
  reserve(m.d_sources, owners);
  if (total > 0)
    HG_HIP(hipMemcpyAsync(m.d_sources.ptr, d_sources + offsets[0], sizeof(int32_t) * total,
                          hipMemcpyDefault, m.stream));
}

void seed(hg_mesh& m, const int64_t* offsets, int64_t k, double* u0,
          FieldLayout layout = FieldLayout::kColumns) {
  const int64_t total = offsets[k] - offsets[0];
  launch_zero(u0, static_cast<int64_t>(sizeof(double)) * m.n * k, m.stream);
  if (m.invalid.size < 1) m.invalid.allocate(1);
  launch_zero(m.invalid.ptr, sizeof(int32_t), m.stream);
  const bool rows = layout == FieldLayout::kRows;
  // With the offsets on the device the kernel reads the count there (a replay's own count).
  const bool on_device = m.sources_on_device;
  launch_seed(m.d_sources.ptr, m.d_owner.ptr, on_device ? m.d_owner.size : total,
              on_device ? m.d_offsets.ptr + k : nullptr, m.n, rows ? k : 0,
              rows ? m.plan->iperm.ptr : nullptr, u0, m.invalid.ptr, m.stream);
  // The range flag comes back with the call's next synchronization (check_seeds); until then the
  // solve runs on, every kernel that reads the sources skipping the ones out of range.
  if (!m.h_invalid.ptr) m.h_invalid.allocate(1);
  HG_HIP(hipMemcpyAsync(m.h_invalid.ptr, m.invalid.ptr, sizeof(int32_t), hipMemcpyDeviceToHost,
                        m.stream));
}

// After a synchronization of the mesh's stream: fail the call if a seed was out of range
// (kernels/heat.py:599-623, heat.py:5488-5503).
void check_seeds(const hg_mesh& m) {
  HG_REQUIRE(!m.h_invalid.ptr || *m.h_invalid.ptr == 0, HG_ERR_VALUE,
             "every entry of sources must be a vertex index in [0, " + std::to_string(m.n) + ")");
}

void divergence(hg_mesh& m, const double* heat, double* out, int64_t k,
                FieldLayout layout = FieldLayout::kColumns) {
  PhaseScope scope(&m.timer, kPhaseDivergence, m.stream);
  const bool rows = layout == FieldLayout::kRows;
  launch_divergence(rows ? m.faces_rows.ptr : m.faces.ptr, rows ? m.plan->iperm.ptr : nullptr,
                    m.face_geom.ptr, m.vertex_ptr.ptr, m.vertex_corners.ptr,
                    rows ? m.row_corners.ptr : nullptr, rows ? m.row_corner_ptr.ptr : nullptr,
                    heat, m.n, k, out, layout, m.stream);
}

// The call's three fields, sized for the next power of two of columns (see SolveWorkspace).
double* reserve_fields(hg_mesh& m, int64_t k) {
  if (m.field_capacity < 3 * m.n * k) {
    int64_t columns = 1;
    while (columns < k) columns *= 2;
    m.fields.allocate(std::max<int64_t>(3 * m.n * columns, 1));
    m.field_capacity = 3 * m.n * columns;
  }
  return m.fields.ptr;
}

void shift(hg_mesh& m, double* phi, int64_t k) {
  int32_t blocks = shift_partial_blocks(m.n);
  reserve(m.shift_partials, blocks * k);
  reserve(m.shift_values, 2 * k);
  PhaseScope scope(&m.timer, kPhaseShift, m.stream);
  launch_shift(phi, m.d_sources.ptr, m.d_offsets.ptr, m.n, k, nullptr, nullptr,
               m.shift_partials.ptr, m.shift_values.ptr, m.stream);
}

// Arm (or disarm) speculative refinement for a call of k columns: with FP64 factors the first
// refinement round almost always passes its test, so the call runs on without waiting for the
// verdicts and checks them at its last synchronization (speculation_failed), solving again without
// speculating in the rare case one failed. Mixed precision always refines, and never speculates.
void arm_speculation(hg_mesh& m, int64_t k, bool on) {
  m.speculating = on;
  if (!on) return;
  m.spec_chunks = (k + kFieldChunk - 1) / kFieldChunk;
  if (m.h_speculate.size < 3 * m.spec_chunks) m.h_speculate.allocate(3 * m.spec_chunks);
  std::fill(m.h_speculate.ptr, m.h_speculate.ptr + 3 * m.spec_chunks, 0);
}

// After the stream's synchronization: whether a speculated first round failed its test.
bool speculation_failed(hg_mesh& m) {
  const bool was = m.speculating;
  m.speculating = false;
  if (!was) return false;
  return std::any_of(m.h_speculate.ptr, m.h_speculate.ptr + 3 * m.spec_chunks,
                     [](int32_t v) { return v != 0; });
}

// Every buffer and setting a captured geodesic graph names: a replay is valid while all are the
// same (buffers only grow, so one at the same address is at least as large).
std::vector<uintptr_t> graph_signature(const hg_mesh& m, int64_t k) {
  auto p = [](const void* q) { return reinterpret_cast<uintptr_t>(q); };
  const SolveWorkspace& w = m.solve;
  uint64_t tol[4];
  std::memcpy(tol, m.refine_tolerance, sizeof(tol));
  return {static_cast<uintptr_t>(k), p(m.fields.ptr), p(w.bp.ptr), p(w.xa.ptr), p(w.xb.ptr),
          p(w.r.ptr), p(w.y.ptr), p(w.d.ptr), p(w.fvec.ptr), p(w.bmax.ptr), p(w.partials.ptr),
          p(w.pending.ptr), static_cast<uintptr_t>(w.capacity), p(m.d_sources.ptr),
          p(m.d_owner.ptr), static_cast<uintptr_t>(m.d_owner.size), p(m.d_offsets.ptr),
          p(m.invalid.ptr), p(m.h_invalid.ptr), p(m.h_speculate.ptr), p(m.d_slots.ptr),
          p(m.h_slots.ptr), p(m.null_partials.ptr), p(m.heat_factor.w64.ptr),
          p(m.heat_factor.pvalues.ptr), p(m.poisson_factor.w64.ptr),
          p(m.poisson_factor.pvalues.ptr), p(m.pinned_factor.w64.ptr),
          p(m.pinned_factor.pvalues.ptr), p(m.reg_factor.w64.ptr), p(m.reg_factor.pvalues.ptr),
          p(m.plan.get()), static_cast<uintptr_t>(m.method),
          static_cast<uintptr_t>(m.n_rim), tol[0], tol[1], tol[2], tol[3],
          p(m.solve_side.bp.ptr), p(m.solve_side.xa.ptr), p(m.solve_side.xb.ptr),
          p(m.solve_side.r.ptr), p(m.solve_side.y.ptr), p(m.solve_side.d.ptr),
          p(m.solve_side.fvec.ptr), p(m.solve_side.bmax.ptr), p(m.solve_side.partials.ptr),
          p(m.solve_side.pending.ptr), static_cast<uintptr_t>(m.solve_side.capacity),
          p(m.pinned_rhs.ptr)};
}

// Whether this call can replay a graph: FP64 factors held for every system the path solves,
// the fused finish, no phase timing, a single mesh.
bool graphable(const hg_mesh& m, int32_t mode, int32_t precision) {
  return m.use_graphs && !m.graphs_broken && precision == HG_PRECISION_FP64 &&
         !m.timer.enabled && m.segments() == 1 && m.plan && m.rows_by_component &&
         (mode == HG_MODE_FACTORED || m.heat_factor.held()) && m.heat_factor.held() &&
         (m.method == HG_METHOD_POTPOURRI3D ? m.reg_factor.held() : m.poisson_factor.held()) &&
         (m.method != HG_METHOD_IGL || m.n_rim == 0 || m.pinned_factor.held());
}

bool solve_columns(hg_mesh& m, const int64_t* offsets, int64_t k, double* phi, int32_t mode,
                   int32_t precision, bool finish);

// The factored FP64 path of hg_geodesic replayed as one graph: captured on the first such call at
// this column count (capturing runs nothing), then launched; each call's inputs reach it through
// the staged sources, offsets and output slots. Returns false when the caller must solve eagerly
// (a capture failed, or the speculated first round failed its test). The sources must be uploaded.
// Graphs are captured and launched on a stream of the library's own, one per host thread (a
// caller's stream may be the null stream, which cannot capture; threads serving different meshes
// keep theirs apart), ordered after the caller's stream by an event. Both live until the process
// ends.
struct GraphStream {
  hipStream_t stream = nullptr;
  hipEvent_t ready = nullptr;
};

// Issues the mesh's work on another stream (optionally through the side workspace) for a scope,
// restoring the mesh's own stream and workspace choice however the scope ends.
struct StreamSwap {
  hg_mesh& m;
  hipStream_t stream;
  bool on_side;
  StreamSwap(hg_mesh& mesh, hipStream_t other, bool side = false)
      : m(mesh), stream(mesh.stream), on_side(mesh.on_side) {
    m.stream = other;
    m.on_side = side;
  }
  ~StreamSwap() {
    m.stream = stream;
    m.on_side = on_side;
  }
  StreamSwap(const StreamSwap&) = delete;
  StreamSwap& operator=(const StreamSwap&) = delete;
};

// Leaves speculation disarmed when a call ends, normally or by an exception: a later call must
// never stop after a first refinement round whose verdict nobody reads.
struct SpeculationDisarm {
  hg_mesh& m;
  explicit SpeculationDisarm(hg_mesh& mesh) : m(mesh) { m.speculating = false; }
  ~SpeculationDisarm() { m.speculating = false; }
  SpeculationDisarm(const SpeculationDisarm&) = delete;
  SpeculationDisarm& operator=(const SpeculationDisarm&) = delete;
};
thread_local GraphStream t_graph_stream;

// The second stream of the igl path's side-by-side heat solves, per host thread, with its fork and
// join events; both live until the process ends.
struct SideStream {
  hipStream_t stream = nullptr;
  hipEvent_t fork = nullptr, join = nullptr;
};
thread_local SideStream t_side_stream;

bool run_graph(hg_mesh& m, const int64_t* offsets, int64_t k, double* d_out, int32_t precision) {
  GraphStream& gs = t_graph_stream;
  if (!gs.stream) {
    HG_HIP(hipStreamCreateWithFlags(&gs.stream, hipStreamNonBlocking));
    HG_HIP(hipEventCreateWithFlags(&gs.ready, hipEventDisableTiming));
  }
  // The uploads on the caller's stream come first.
  HG_HIP(hipEventRecord(gs.ready, m.stream));
  HG_HIP(hipStreamWaitEvent(gs.stream, gs.ready, 0));
  std::unique_ptr<hg_mesh::Graph>& g = m.graphs[k];
  if (!g || g->signature != graph_signature(m, k)) {
    g.reset();
    auto graph = std::make_unique<hg_mesh::Graph>();
    arm_speculation(m, k, true);
    m.stats = hg_solve_stats{};
    hipGraph_t captured = nullptr;
    bool ok = false;
    {
      // Whatever the capture throws, the stream leaves capture mode and the mesh gets its own
      // stream back (and its main workspace) before the exception goes on.
      StreamSwap swap(m, gs.stream);
      ok = hipStreamBeginCapture(gs.stream, hipStreamCaptureModeRelaxed) == hipSuccess;
      if (ok) {
        try {
          solve_columns(m, offsets, k, d_out, HG_MODE_FACTORED, precision, true);
        } catch (...) {
          ok = false;
        }
        ok = hipStreamEndCapture(gs.stream, &captured) == hipSuccess && ok;
      }
    }
    if (ok) ok = hipGraphInstantiate(&graph->exec, captured, nullptr, nullptr, 0) == hipSuccess;
    if (captured) (void)hipGraphDestroy(captured);
    m.speculating = false;
    if (!ok) {
      graph->exec = nullptr;
      m.graphs_broken = true;  // from now on this mesh solves eagerly
      (void)hipGetLastError();
      return false;
    }
    graph->stats = m.stats;
    graph->signature = graph_signature(m, k);
    g = std::move(graph);
  }
  arm_speculation(m, k, true);
  const int64_t chunks = (k + kFieldChunk - 1) / kFieldChunk;
  for (int64_t c = 0; c < chunks; ++c) m.h_slots.ptr[c] = d_out + c * kFieldChunk * m.n;
  m.stats = g->stats;
  HG_HIP(hipGraphLaunch(g->exec, gs.stream));
  HG_HIP(hipStreamSynchronize(gs.stream));
  return !speculation_failed(m);
}

// Seed, heat step, divergence and Poisson step over k columns of the whole (possibly packed)
// system into phi (the caller's layout). The sources must be uploaded. With `finish` and the
// fields stored by rows, phi is the finished distance (null space removed, shifted, oriented) and
// the call returns true; otherwise phi is the unshifted distance with each component's mean
// removed.
bool solve_columns(hg_mesh& m, const int64_t* offsets, int64_t k, double* phi, int32_t mode,
                   int32_t precision, bool finish = false) {
  const int64_t n = m.n;
  double* fields = reserve_fields(m, k);
  double* u0 = fields;
  double* heat = fields + n * k;
  double* negdiv = fields + 2 * n * k;
  // Through the factors the intermediate fields stay in the solves' row layout, in the elimination
  // order: the seeds land there, the heat step's solution stays there for the divergence, which
  // writes the Poisson step's right-hand side there; only the distance is permuted out. Every
  // system the path solves is factored first; one refused under the memory budget sends the whole
  // call down the iterated path (packed meshes have none: the refusal stands).
  bool factored = mode == HG_MODE_FACTORED || m.heat_factor.held();
  // geometry-central's Poisson step at HG_METHOD_POTPOURRI3D (heat.py:3121-3128).
  const bool regularized = m.method == HG_METHOD_POTPOURRI3D;
  if (factored) {
    ensure_operators(m);
    const bool mixed = precision == HG_PRECISION_MIXED;
    try {
      factor_system(m, HG_SYSTEM_HEAT, mixed);
      factor_system(m, regularized ? HG_SYSTEM_REGULARIZED : HG_SYSTEM_POISSON, mixed);
      if (m.method == HG_METHOD_IGL && m.n_rim > 0) factor_system(m, HG_SYSTEM_PINNED, mixed);
    } catch (const Error& e) {
      if (!is_budget_refusal(e) || m.segments() > 1) throw;
      factored = false;
    }
  }
  mode = factored ? HG_MODE_FACTORED : HG_MODE_ITERATE;
  const FieldLayout layout = factored ? FieldLayout::kRows : FieldLayout::kColumns;
  seed(m, offsets, k, u0, layout);
  const bool igl = m.method == HG_METHOD_IGL && m.n_rim > 0;
  int32_t flags = 0;
  if (igl && layout == FieldLayout::kRows && m.speculating) {
    // igl::heat_geodesics' boundary rule (heat.py:594-597) with the two heat solves side by side:
    // the rim-pinned one, from the seeds zeroed on the rim, runs on a second stream with its own
    // workspace and first-round verdict while the Neumann one runs here; neither waits on the host
    // (speculation), and each one's latency-bound separator levels fill the other's idle time.
    reserve(m.pinned_rhs, n * k);
    launch_off_rim(u0, m.pinned_rhs.ptr, m.d_rim_rows.ptr, n, k, layout, m.stream);
    SideStream& side = t_side_stream;
    if (!side.stream) {
      HG_HIP(hipStreamCreateWithFlags(&side.stream, hipStreamNonBlocking));
      HG_HIP(hipEventCreateWithFlags(&side.fork, hipEventDisableTiming));
      HG_HIP(hipEventCreateWithFlags(&side.join, hipEventDisableTiming));
    }
    HG_HIP(hipEventRecord(side.fork, m.stream));
    HG_HIP(hipStreamWaitEvent(side.stream, side.fork, 0));
    {
      StreamSwap swap(m, side.stream, true);
      flags |= diffuse(m, m.pinned_rhs.ptr, negdiv, k, HG_MODE_FACTORED, precision, layout, true);
    }
    flags |= diffuse(m, u0, heat, k, mode, precision, layout);
    HG_HIP(hipEventRecord(side.join, side.stream));
    HG_HIP(hipStreamWaitEvent(m.stream, side.join, 0));
    launch_average(heat, negdiv, n * k, m.stream);
  } else {
    flags = diffuse(m, u0, heat, k, mode, precision, layout);
  }
  if (igl && !(layout == FieldLayout::kRows && m.speculating)) {
    // igl::heat_geodesics' boundary rule (heat.py:594-597): the Neumann heat averaged with the
    // heat pinned to zero on the rim, solved from the seeds zeroed there. The divergence's buffer
    // holds the pinned heat until the average has read it.
    launch_off_rim(u0, u0, layout == FieldLayout::kRows ? m.d_rim_rows.ptr : m.d_rim.ptr, n, k,
                   layout, m.stream);
    flags |= diffuse(m, u0, negdiv, k, layout == FieldLayout::kRows ? HG_MODE_FACTORED : mode,
                     precision, layout, true);
    launch_average(heat, negdiv, n * k, m.stream);
  }
  m.stats.backward_error_failed = flags & 1;
  divergence(m, heat, negdiv, k, layout);
  if (finish && layout == FieldLayout::kRows && m.rows_by_component) {
    // The heat field is spent once the divergence has read it: its buffer takes the Poisson
    // solution, which one pass turns into the distance in the caller's layout.
    solve_poisson(m, negdiv, heat, k, mode, precision, layout, FieldLayout::kRows, regularized);
    PhaseScope scope(&m.timer, kPhaseShift, m.stream);
    const int64_t chunks = (k + kFieldChunk - 1) / kFieldChunk;
    reserve(m.d_slots, chunks);
    reserve(m.h_slots, chunks);
    for (int64_t c = 0; c < chunks; ++c) m.h_slots.ptr[c] = phi + c * kFieldChunk * n;
    m.d_slots.upload(m.h_slots.ptr, chunks, m.stream);
    finish_rows(*m.plan, m.null_space, m.comp_ptr.ptr, m.d_labels.ptr, heat, k, m.d_sources.ptr,
                m.d_offsets.ptr, m.d_slots.ptr, m.null_partials, m.stream, !regularized);
    return true;
  }
  solve_poisson(m, negdiv, phi, k, mode, precision, layout, FieldLayout::kColumns, regularized);
  return false;
}

// The host and device state of a mesh from its host positions and faces (global indices).
void bind(hg_mesh& m) {
  const int64_t n_vertices = m.n;
  hipStream_t st = m.stream;
  m.faces.upload(m.h_faces, st);
  build_pattern_device(n_vertices, m.faces.ptr, m.f, st,
                       {m.offsets, m.columns, m.corner_ptr, m.corner_slots, m.vertex_ptr,
                        m.vertex_corners},
                       &m.pattern);
  m.labels = component_labels(m.pattern, &m.n_comp);
  std::vector<int32_t> comp_ptr(m.n_comp + 1, 0), members(n_vertices);
  for (int64_t v = 0; v < n_vertices; ++v) comp_ptr[m.labels[v] + 1]++;
  for (int32_t c = 0; c < m.n_comp; ++c) comp_ptr[c + 1] += comp_ptr[c];
  {
    std::vector<int32_t> at(comp_ptr.begin(), comp_ptr.end() - 1);
    for (int64_t v = 0; v < n_vertices; ++v) members[at[m.labels[v]]++] = static_cast<int32_t>(v);
  }
  m.xyz.upload(m.h_xyz, st);
  // The rim: rows with an off-diagonal entry of a single corner, i.e. an edge on one face
  // (kernels/laplacian.py:1286-1321).
  m.rim.assign(static_cast<size_t>(n_vertices), 0);
  m.n_rim = 0;
  for (int64_t v = 0; v < n_vertices; ++v) {
    for (int32_t e = m.pattern.offsets[v]; e < m.pattern.offsets[v + 1]; ++e) {
      if (m.pattern.columns[e] != v && m.pattern.corner_ptr[e + 1] - m.pattern.corner_ptr[e] == 1) {
        m.rim[v] = 1;
        ++m.n_rim;
        break;
      }
    }
  }
  if (m.n_rim > 0) m.d_rim.upload(m.rim, st);
  m.comp_ptr.upload(comp_ptr, st);
  m.comp_ptr_host = comp_ptr;
  m.d_labels.upload(m.labels, st);
  m.null_space.build(comp_ptr, st);
  m.comp_members.upload(members, st);
  if (m.segments() > 1) {
    std::vector<int32_t> seg(static_cast<size_t>(n_vertices));
    for (int64_t s = 0; s < m.segments(); ++s)
      std::fill(seg.begin() + m.seg_v[s], seg.begin() + m.seg_v[s + 1], static_cast<int32_t>(s));
    m.row_segment.upload(seg, st);
  }
  HG_HIP(hipStreamSynchronize(st));
}

}  // namespace

extern "C" {

const char* hg_last_error(void) { return g_last_error.c_str(); }

const char* hg_version(void) { return "heatgeo 0.1.0 gfx950"; }

int hg_mesh_create(const float* d_vertices, int64_t n_vertices, const int32_t* d_faces,
                   int64_t n_faces, void* stream, hg_mesh** out_mesh) {
  return guarded([&] {
    HG_REQUIRE(out_mesh != nullptr, HG_ERR_VALUE, "out_mesh is null");
    HG_REQUIRE(n_vertices >= 0 && n_faces >= 0, HG_ERR_VALUE, "negative size");
    HG_REQUIRE(n_vertices < (1ll << 31) && 3 * n_faces < (1ll << 31), HG_ERR_VALUE,
               "mesh too large for int32 indices");
    auto m = std::make_unique<hg_mesh>();
    m->stream = static_cast<hipStream_t>(stream);
    m->n = n_vertices;
    m->f = n_faces;
    m->seg_v = {0, n_vertices};
    m->seg_f = {0, n_faces};
    m->h_xyz.resize(3 * n_vertices);
    m->h_faces.resize(3 * n_faces);
    if (n_vertices > 0)
      HG_HIP(hipMemcpy(m->h_xyz.data(), d_vertices, sizeof(float) * 3 * n_vertices, hipMemcpyDefault));
    if (n_faces > 0)
      HG_HIP(hipMemcpy(m->h_faces.data(), d_faces, sizeof(int32_t) * 3 * n_faces, hipMemcpyDefault));
    bind(*m);
    *out_mesh = m.release();
  });
}

int hg_mesh_create_packed(const float* d_vertices, const int64_t* vertex_offsets,
                          const int32_t* d_faces, const int64_t* face_offsets, int64_t n_meshes,
                          void* stream, hg_mesh** out_mesh) {
  return guarded([&] {
    HG_REQUIRE(out_mesh != nullptr && vertex_offsets != nullptr && face_offsets != nullptr,
               HG_ERR_VALUE, "null argument");
    HG_REQUIRE(n_meshes >= 1, HG_ERR_VALUE, "a packed mesh needs at least one mesh");
    HG_REQUIRE(vertex_offsets[0] == 0 && face_offsets[0] == 0, HG_ERR_VALUE,
               "offsets must start at 0");
    for (int64_t s = 0; s < n_meshes; ++s)
      HG_REQUIRE(vertex_offsets[s + 1] >= vertex_offsets[s] && face_offsets[s + 1] >= face_offsets[s],
                 HG_ERR_VALUE, "offsets must not decrease");
    const int64_t n_vertices = vertex_offsets[n_meshes], n_faces = face_offsets[n_meshes];
    HG_REQUIRE(n_vertices < (1ll << 31) && 3 * n_faces < (1ll << 31), HG_ERR_VALUE,
               "packed meshes too large for int32 indices");
    auto m = std::make_unique<hg_mesh>();
    m->stream = static_cast<hipStream_t>(stream);
    m->n = n_vertices;
    m->f = n_faces;
    m->seg_v.assign(vertex_offsets, vertex_offsets + n_meshes + 1);
    m->seg_f.assign(face_offsets, face_offsets + n_meshes + 1);
    m->h_xyz.resize(3 * n_vertices);
    m->h_faces.resize(3 * n_faces);
    if (n_vertices > 0)
      HG_HIP(hipMemcpy(m->h_xyz.data(), d_vertices, sizeof(float) * 3 * n_vertices, hipMemcpyDefault));
    if (n_faces > 0)
      HG_HIP(hipMemcpy(m->h_faces.data(), d_faces, sizeof(int32_t) * 3 * n_faces, hipMemcpyDefault));
    // Each mesh's faces index its own vertices; shift them to the packed numbering.
    for (int64_t s = 0; s < n_meshes; ++s) {
      const int64_t base = vertex_offsets[s], count = vertex_offsets[s + 1] - base;
      for (int64_t e = 3 * face_offsets[s]; e < 3 * face_offsets[s + 1]; ++e) {
        const int32_t v = m->h_faces[e];
        HG_REQUIRE(v >= 0 && v < count, HG_ERR_VALUE,
                   "mesh " + std::to_string(s) + " has a face index outside [0, " +
                       std::to_string(count) + ")");
        m->h_faces[e] = static_cast<int32_t>(v + base);
      }
    }
    bind(*m);
    *out_mesh = m.release();
  });
}

int hg_mesh_destroy(hg_mesh* mesh) {
  return guarded([&] {
    if (mesh) {
      (void)hipStreamSynchronize(mesh->stream);
      delete mesh;
    }
  });
}

int hg_mesh_set_stream(hg_mesh* mesh, void* stream) {
  return guarded([&] {
    require_mesh(mesh);
    mesh->stream = static_cast<hipStream_t>(stream);
  });
}

int hg_operators(hg_mesh* mesh, double t, int32_t method, int32_t solver_default_time) {
  return guarded([&] {
    require_mesh(mesh);
    const bool robust = (method & HG_METHOD_ROBUST) != 0;
    method &= ~HG_METHOD_ROBUST;
    HG_REQUIRE(method == HG_METHOD_MESHLAB || method == HG_METHOD_IGL ||
                   method == HG_METHOD_POTPOURRI3D,
               HG_ERR_VALUE, "method must be 'meshlab', 'igl' or 'potpourri3d'");
    HG_REQUIRE(!robust || method != HG_METHOD_POTPOURRI3D, HG_ERR_VALUE,
               "use_robust at 'potpourri3d' is geometry-central's intrinsic Delaunay method, which "
               "this library does not build; use 'meshlab' or 'igl'");
    HG_REQUIRE(!robust || mesh->segments() == 1, HG_ERR_VALUE,
               "use_robust takes one mesh, not a packed set");
    HG_REQUIRE(std::isnan(t) || t > 0.0, HG_ERR_VALUE, "t must be positive");
    assemble(*mesh, t, method, solver_default_time != 0, robust);
  });
}

int hg_operator_pattern(const int32_t* d_faces, int64_t n_faces, int64_t n_vertices, void* stream,
                        int32_t* d_offsets, int32_t* d_columns, int32_t* d_corner_ptr,
                        int32_t* d_corner_slots, int32_t* d_vertex_ptr, int32_t* d_vertex_corners,
                        int64_t* out_sizes) {
  return guarded([&] {
    HG_REQUIRE(n_vertices >= 0 && n_faces >= 0, HG_ERR_VALUE, "negative size");
    HG_REQUIRE(n_vertices < (1ll << 31) && 3 * n_faces < (1ll << 31), HG_ERR_VALUE,
               "mesh too large for int32 indices");
    HG_REQUIRE(out_sizes != nullptr && (n_faces == 0 || d_faces != nullptr), HG_ERR_VALUE,
               "bad arguments");
    hipStream_t st = static_cast<hipStream_t>(stream);
    // The faces are copied into library memory first (from device or host memory alike), as a
    // mesh's are, so no kernel reads the caller's pointer directly.
    DeviceArray<int32_t> faces(std::max<int64_t>(3 * n_faces, 1));
    if (n_faces > 0)
      HG_HIP(hipMemcpyAsync(faces.ptr, d_faces, sizeof(int32_t) * 3 * n_faces, hipMemcpyDefault,
                            st));
    DeviceArray<int32_t> offsets, columns, corner_ptr, corner_slots, vertex_ptr, vertex_corners;
    build_pattern_device(n_vertices, faces.ptr, n_faces, st,
                         {offsets, columns, corner_ptr, corner_slots, vertex_ptr, vertex_corners},
                         nullptr);
    int32_t entries = 0;
    HG_HIP(hipMemcpyAsync(&entries, offsets.ptr + n_vertices, sizeof(int32_t),
                          hipMemcpyDeviceToHost, st));
    HG_HIP(hipStreamSynchronize(st));
    // Every face key carries its corner slot: two a corner.
    const int64_t nnz = entries, n_slots = 6 * n_faces;
    auto copy = [&](int32_t* dst, const DeviceArray<int32_t>& src, int64_t count) {
      if (dst != nullptr && count > 0)
        HG_HIP(hipMemcpyAsync(dst, src.ptr, sizeof(int32_t) * count, hipMemcpyDefault, st));
    };
    copy(d_offsets, offsets, n_vertices + 1);
    copy(d_columns, columns, nnz);
    copy(d_corner_ptr, corner_ptr, nnz + 1);
    copy(d_corner_slots, corner_slots, n_slots);
    copy(d_vertex_ptr, vertex_ptr, n_vertices + 1);
    copy(d_vertex_corners, vertex_corners, 3 * n_faces);
    HG_HIP(hipStreamSynchronize(st));
    out_sizes[0] = nnz;
    out_sizes[1] = n_slots;
  });
}

int hg_operator_views(hg_mesh* mesh, hg_operator_view* out) {
  return guarded([&] {
    require_mesh(mesh);
    HG_REQUIRE(out != nullptr, HG_ERR_VALUE, "out_view is null");
    ensure_operators(*mesh);
    out->n_vertices = mesh->n;
    out->n_faces = mesh->f;
    out->nnz = static_cast<int64_t>(mesh->pattern.columns.size());
    out->offsets = mesh->offsets.ptr;
    out->columns = mesh->columns.ptr;
    out->laplacian = mesh->lap.ptr;
    out->heat = mesh->heat.ptr;
    out->poisson = mesh->poisson.ptr;
    out->mass = mesh->mass.ptr;
    out->cot = mesh->cot.ptr;
    out->normals = mesh->normals.ptr;
    out->areas = mesh->areas.ptr;
    out->t = mesh->t;
    out->settle_rounds = mesh->settle_rounds;
    out->time_from_edges = mesh->time_from_edges;
    out->mollify_delta = mesh->mollify_delta;
  });
}

int hg_factor(hg_mesh* mesh, int32_t systems, int32_t precision, int32_t* out_factored) {
  return guarded([&] {
    require_mesh(mesh);
    const bool mixed = precision == HG_PRECISION_MIXED;
    // A refusal under the memory budget is an answer, not an error (OperatorFactorization.factor
    // returns False, linalg.py:2393-2420): *out_factored reports it.
    for (int32_t system :
         {HG_SYSTEM_HEAT, HG_SYSTEM_POISSON, HG_SYSTEM_PINNED, HG_SYSTEM_REGULARIZED}) {
      if (!(systems & system)) continue;
      try {
        factor_system(*mesh, system, mixed);
      } catch (const Error& e) {
        if (!is_budget_refusal(e)) throw;
      }
    }
    if (out_factored) {
      bool ok = (!(systems & HG_SYSTEM_HEAT) || mesh->heat_factor.held()) &&
                (!(systems & HG_SYSTEM_POISSON) || mesh->poisson_factor.held()) &&
                (!(systems & HG_SYSTEM_PINNED) || mesh->n_rim == 0 || mesh->pinned_factor.held()) &&
                (!(systems & HG_SYSTEM_REGULARIZED) || mesh->reg_factor.held());
      *out_factored = ok ? 1 : 0;
    }
  });
}

int hg_factored(hg_mesh* mesh, int32_t* out_systems) {
  return guarded([&] {
    require_mesh(mesh);
    *out_systems = (mesh->heat_factor.held() ? HG_SYSTEM_HEAT : 0) |
                   (mesh->poisson_factor.held() ? HG_SYSTEM_POISSON : 0) |
                   (mesh->pinned_factor.held() ? HG_SYSTEM_PINNED : 0) |
                   (mesh->reg_factor.held() ? HG_SYSTEM_REGULARIZED : 0) | (mesh->refused << 4);
  });
}

int hg_set_plan_options(hg_mesh* mesh, int32_t leaf_rows, int32_t top_rows,
                        int32_t split_target) {
  return guarded([&] {
    require_mesh(mesh);
    HG_REQUIRE(leaf_rows >= 0 && top_rows >= 0 && split_target >= 0, HG_ERR_VALUE,
               "plan options must not be negative");
    hg_mesh& m = *mesh;
    const AnalysisOptions defaults;
    const int32_t leaf = leaf_rows > 0 ? leaf_rows : defaults.leaf_rows;
    const int32_t top = top_rows > 0 ? top_rows : defaults.top_rows;
    const int32_t split = split_target > 0 ? split_target : kDefaultSplitTarget;
    if (leaf != m.options.leaf_rows || top != m.options.top_rows) {
      // A new ordering: the plan and every factor built on it go.
      HG_HIP(hipStreamSynchronize(m.stream));
      invalidate_graphs(m);
      m.heat_factor.reset();
      m.poisson_factor.reset();
      m.pinned_factor.reset();
      m.reg_factor.reset();
      m.plan.reset();
      m.host_plan.reset();
      m.options.leaf_rows = leaf;
      m.options.top_rows = top;
      m.split_target = split;
      return;
    }
    if (split != m.split_target) {
      m.split_target = split;
      invalidate_graphs(m);
      if (m.plan) {
        HG_HIP(hipStreamSynchronize(m.stream));
        m.plan->upload_sweeps(*m.host_plan, split, m.stream);
        HG_HIP(hipStreamSynchronize(m.stream));
      }
    }
  });
}

int hg_set_refinement(hg_mesh* mesh, int32_t precision, double heat_tolerance,
                      double poisson_tolerance) {
  return guarded([&] {
    require_mesh(mesh);
    HG_REQUIRE(precision == HG_PRECISION_FP64 || precision == HG_PRECISION_MIXED, HG_ERR_VALUE,
               "unknown precision");
    const bool keep_heat = std::isnan(heat_tolerance), keep_poisson = std::isnan(poisson_tolerance);
    HG_REQUIRE(keep_heat || (heat_tolerance > 0.0 && heat_tolerance < 1.0), HG_ERR_VALUE,
               "the heat refinement tolerance must lie in (0, 1)");
    HG_REQUIRE(keep_poisson || (poisson_tolerance > 0.0 && poisson_tolerance < 1.0), HG_ERR_VALUE,
               "the Poisson refinement tolerance must lie in (0, 1)");
    if (!keep_heat) mesh->refine_tolerance[precision][0] = heat_tolerance;
    if (!keep_poisson) mesh->refine_tolerance[precision][1] = poisson_tolerance;
    invalidate_graphs(*mesh);
  });
}

int hg_diffuse(hg_mesh* mesh, const double* d_rhs, double* d_solution, int64_t n_columns,
               int32_t mode, int32_t precision, int32_t* out_flags) {
  return guarded([&] {
    require_mesh(mesh);
    HG_REQUIRE(n_columns >= 0, HG_ERR_VALUE, "negative column count");
    require_single_or_factored(*mesh, mode);
    SpeculationDisarm disarm(*mesh);
    mesh->stats = hg_solve_stats{};
    // By columns (the caller's layout) the solve only reads its right-hand side.
    int32_t flags = diffuse(*mesh, const_cast<double*>(d_rhs), d_solution, n_columns, mode,
                            precision);
    HG_HIP(hipStreamSynchronize(mesh->stream));
    if (out_flags) *out_flags = flags;
  });
}

int hg_diffuse_pinned(hg_mesh* mesh, const double* d_rhs, double* d_solution, int64_t n_columns,
                      int32_t mode, int32_t precision, int32_t* out_flags,
                      int32_t* out_has_rim) {
  return guarded([&] {
    require_mesh(mesh);
    HG_REQUIRE(n_columns >= 0, HG_ERR_VALUE, "negative column count");
    require_single_or_factored(*mesh, mode);
    hg_mesh& m = *mesh;
    SpeculationDisarm disarm(m);
    m.stats = hg_solve_stats{};
    ensure_operators(m);
    int32_t flags = 0;
    if (m.n_rim > 0 && n_columns > 0) {
      // The caller's right-hand side is only read: its copy with the rim rows zeroed is solved.
      const int64_t count = m.n * n_columns;
      if (m.pinned_rhs.size < count) m.pinned_rhs.allocate(count);
      launch_off_rim(d_rhs, m.pinned_rhs.ptr, m.d_rim.ptr, m.n, n_columns, FieldLayout::kColumns,
                     m.stream);
      flags = diffuse(m, m.pinned_rhs.ptr, d_solution, n_columns, mode, precision,
                      FieldLayout::kColumns, true);
    }
    HG_HIP(hipStreamSynchronize(m.stream));
    if (out_flags) *out_flags = flags;
    if (out_has_rim) *out_has_rim = m.n_rim > 0 ? 1 : 0;
  });
}

int hg_set_graphs(hg_mesh* mesh, int32_t enable) {
  return guarded([&] {
    require_mesh(mesh);
    mesh->use_graphs = enable != 0;
    if (!mesh->use_graphs) invalidate_graphs(*mesh);
  });
}

int hg_graph_state(hg_mesh* mesh, int64_t* out_cached, int32_t* out_broken) {
  return guarded([&] {
    require_mesh(mesh);
    if (out_cached) *out_cached = static_cast<int64_t>(mesh->graphs.size());
    if (out_broken) *out_broken = mesh->graphs_broken ? 1 : 0;
  });
}

int hg_last_stats(hg_mesh* mesh, hg_solve_stats* out_stats) {
  return guarded([&] {
    require_mesh(mesh);
    HG_REQUIRE(out_stats != nullptr, HG_ERR_VALUE, "out_stats is null");
    *out_stats = mesh->stats;
  });
}

int hg_set_budget_probe(hg_mesh* mesh, int32_t probe) {
  return guarded([&] {
    require_mesh(mesh);
    mesh->budget_probe = probe != 0;
  });
}

int hg_set_settle_rule(hg_mesh* mesh, int32_t check_rounds, double change_tolerance,
                       int32_t settle_rounds) {
  return guarded([&] {
    require_mesh(mesh);
    HG_REQUIRE(check_rounds >= 1, HG_ERR_VALUE, "check_rounds must be at least 1");
    HG_REQUIRE(change_tolerance > 0.0 && std::isfinite(change_tolerance), HG_ERR_VALUE,
               "change_tolerance must be positive and finite");
    HG_REQUIRE(settle_rounds >= 0, HG_ERR_VALUE, "settle_rounds must not be negative");
    mesh->check_rounds = check_rounds;
    mesh->change_tolerance = change_tolerance;
    mesh->settle_window = settle_rounds;
  });
}

int hg_set_memory_budget(hg_mesh* mesh, int64_t bytes) {
  return guarded([&] {
    require_mesh(mesh);
    HG_REQUIRE(bytes >= 0, HG_ERR_VALUE, "the memory budget must not be negative");
    mesh->budget = bytes;
    mesh->refused = 0;
    invalidate_graphs(*mesh);
  });
}

int hg_rim_count(hg_mesh* mesh, int64_t* out_count) {
  return guarded([&] {
    require_mesh(mesh);
    HG_REQUIRE(out_count != nullptr, HG_ERR_VALUE, "out_count is null");
    *out_count = mesh->n_rim;
  });
}

int hg_solve_poisson(hg_mesh* mesh, const double* d_rhs, double* d_solution, int64_t n_columns,
                     int32_t mode, int32_t precision) {
  return guarded([&] {
    require_mesh(mesh);
    HG_REQUIRE(n_columns >= 0, HG_ERR_VALUE, "negative column count");
    require_single_or_factored(*mesh, mode);
    SpeculationDisarm disarm(*mesh);
    mesh->stats = hg_solve_stats{};
    // By columns (the caller's layout) the solve only reads its right-hand side.
    solve_poisson(*mesh, const_cast<double*>(d_rhs), d_solution, n_columns, mode, precision);
    HG_HIP(hipStreamSynchronize(mesh->stream));
  });
}

int hg_unit_gradient_divergence(hg_mesh* mesh, const double* d_heat, double* d_neg_divergence,
                                int64_t n_columns) {
  return guarded([&] {
    require_mesh(mesh);
    ensure_operators(*mesh);
    divergence(*mesh, d_heat, d_neg_divergence, n_columns);
    HG_HIP(hipStreamSynchronize(mesh->stream));
  });
}

int hg_seed_sources(hg_mesh* mesh, const int32_t* d_sources, const int64_t* source_offsets,
                    int64_t n_columns, double* d_u0) {
  return guarded([&] {
    require_mesh(mesh);
    upload_sources(*mesh, d_sources, source_offsets, n_columns);
    seed(*mesh, source_offsets, n_columns, d_u0);
    HG_HIP(hipStreamSynchronize(mesh->stream));
    check_seeds(*mesh);
  });
}

int hg_shift_to_sources(hg_mesh* mesh, double* d_phi, const int32_t* d_sources,
                        const int64_t* source_offsets, int64_t n_columns) {
  return guarded([&] {
    require_mesh(mesh);
    HG_REQUIRE(mesh->segments() == 1, HG_ERR_VALUE,
               "packed meshes shift each query in hg_geodesic_packed");
    upload_sources(*mesh, d_sources, source_offsets, n_columns);
    shift(*mesh, d_phi, n_columns);
    HG_HIP(hipStreamSynchronize(mesh->stream));
  });
}

// ------------------------------------------------------------------------------------------------
// The shift without allocation: its column partials, decisions and the queries' offsets live in
// one file-scope device array (each device holds its own copy, found by hipGetSymbolAddress once
// per device). Layout: offsets (k+1 int64) at 0, the decisions (2k doubles) at kNoAllocShiftAt,
// the column partials (blocks x k doubles) at kNoAllocPartialsAt. Every word a call reads it
// writes first, so nothing needs clearing; calls that overlap on one device share the array and
// race.
// ------------------------------------------------------------------------------------------------

namespace {

constexpr int64_t kNoAllocColumns = 4096;
constexpr int64_t kNoAllocPartials = int64_t{1} << 20;
constexpr size_t kNoAllocShiftAt = 64 * 1024;
constexpr size_t kNoAllocPartialsAt = 128 * 1024;
constexpr size_t kNoAllocBytes = kNoAllocPartialsAt + sizeof(double) * kNoAllocPartials;
static_assert(sizeof(int64_t) * (kNoAllocColumns + 1) <= kNoAllocShiftAt, "offsets overrun");
static_assert(kNoAllocShiftAt + sizeof(double) * 2 * kNoAllocColumns <= kNoAllocPartialsAt,
              "decisions overrun");

constexpr int kNoAllocDevices = 64;

}  // namespace

static __device__ __align__(256) unsigned char scratch[kNoAllocBytes];

namespace {

std::mutex g_scratch_mutex;
char* g_scratch[kNoAllocDevices] = {};

// The current device's copy of `scratch`: its address on that device, looked up once.
char* device_scratch() {
  int device = 0;
  HG_HIP(hipGetDevice(&device));
  HG_REQUIRE(device >= 0 && device < kNoAllocDevices, HG_ERR_RUNTIME,
             "device ordinal past the scratch table");
  std::lock_guard<std::mutex> lock(g_scratch_mutex);
  if (!g_scratch[device]) {
    void* ptr = nullptr;
    HG_HIP(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
    HG_REQUIRE(ptr != nullptr, HG_ERR_RUNTIME, "no device address for the shift scratch");
    g_scratch[device] = static_cast<char*>(ptr);
  }
  return g_scratch[device];
}

}  // namespace

int hg_shift_to_sources_noalloc(hg_mesh* mesh, double* d_phi, const int32_t* d_sources,
                                const int64_t* source_offsets, int64_t n_columns) {
  return guarded([&] {
    require_mesh(mesh);
    HG_REQUIRE(mesh->segments() == 1, HG_ERR_VALUE,
               "packed meshes shift each query in hg_geodesic_packed");
    HG_REQUIRE(n_columns >= 0 && n_columns <= kNoAllocColumns, HG_ERR_VALUE,
               "hg_shift_to_sources_noalloc takes at most " + std::to_string(kNoAllocColumns) +
                   " columns");
    if (n_columns == 0 || mesh->n == 0) return;
    HG_REQUIRE(source_offsets != nullptr, HG_ERR_VALUE, "source_offsets is null");
    for (int64_t q = 0; q < n_columns; ++q)
      HG_REQUIRE(source_offsets[q] >= 0 && source_offsets[q] <= source_offsets[q + 1],
                 HG_ERR_VALUE, "source_offsets must be non-negative and non-decreasing");
    const int64_t blocks = shift_partial_blocks(mesh->n);
    HG_REQUIRE(blocks * n_columns <= kNoAllocPartials, HG_ERR_VALUE,
               "hg_shift_to_sources_noalloc holds " + std::to_string(kNoAllocPartials) +
                   " column partials (one per 256 vertices per column): " +
                   std::to_string(blocks * n_columns) + " asked");
    char* base = device_scratch();
    int64_t* d_offsets = reinterpret_cast<int64_t*>(base);
    double* d_shift = reinterpret_cast<double*>(base + kNoAllocShiftAt);
    double* d_partials = reinterpret_cast<double*>(base + kNoAllocPartialsAt);
    // The caller's offsets index d_sources as they stand; the kernels read sources[e] for e in
    // [off[q], off[q+1]), the same entries the allocating path copies out relative to off[0].
    HG_HIP(hipMemcpyAsync(d_offsets, source_offsets, sizeof(int64_t) * (n_columns + 1),
                          hipMemcpyHostToDevice, mesh->stream));
    launch_shift(d_phi, d_sources, d_offsets, mesh->n, n_columns, nullptr, nullptr, d_partials,
                 d_shift, mesh->stream);
    HG_HIP(hipStreamSynchronize(mesh->stream));
  });
}

int hg_geodesic(hg_mesh* mesh, const int32_t* d_sources, const int64_t* source_offsets,
                int64_t n_queries, double* d_out, int32_t mode, int32_t precision,
                hg_solve_stats* out_stats) {
  return guarded([&] {
    require_mesh(mesh);
    HG_REQUIRE(n_queries >= 0, HG_ERR_VALUE, "negative query count");
    HG_REQUIRE(source_offsets != nullptr, HG_ERR_VALUE, "source_offsets is null");
    HG_REQUIRE(mesh->segments() == 1, HG_ERR_VALUE,
               "a packed mesh answers queries through hg_geodesic_packed");
    for (int64_t q = 0; q < n_queries; ++q)
      HG_REQUIRE(source_offsets[q] <= source_offsets[q + 1], HG_ERR_VALUE,
                 "source_offsets must not decrease");
    hg_mesh& m = *mesh;
    SpeculationDisarm disarm(m);
    m.stats = hg_solve_stats{};
    const int64_t n = m.n, k = n_queries;
    if (k == 0) return;
    if (n == 0 || m.f == 0) {
      if (n > 0) HG_HIP(hipMemsetAsync(d_out, 0, sizeof(double) * n * k, m.stream));
      HG_HIP(hipStreamSynchronize(m.stream));
      return;
    }
    ensure_operators(m);
    upload_sources(m, d_sources, source_offsets, k);
    bool done = false;
    if (graphable(m, mode, precision)) done = run_graph(m, source_offsets, k, d_out, precision);
    if (!done) {
      for (bool speculate : {precision == HG_PRECISION_FP64, false}) {
        m.stats = hg_solve_stats{};
        arm_speculation(m, k, speculate);
        if (!solve_columns(m, source_offsets, k, d_out, mode, precision, true)) shift(m, d_out, k);
        HG_HIP(hipStreamSynchronize(m.stream));
        if (!speculation_failed(m)) break;
      }
    }
    check_seeds(m);
    m.timer.resolve();
    if (out_stats) *out_stats = m.stats;
  });
}

int hg_geodesic_packed(hg_mesh* mesh, const int32_t* d_sources, const int64_t* source_offsets,
                       const int32_t* query_mesh, int64_t n_queries, double* d_out,
                       int32_t precision, hg_solve_stats* out_stats) {
  return guarded([&] {
    require_mesh(mesh);
    HG_REQUIRE(n_queries >= 0, HG_ERR_VALUE, "negative query count");
    HG_REQUIRE(source_offsets != nullptr && query_mesh != nullptr, HG_ERR_VALUE,
               "source_offsets or query_mesh is null");
    hg_mesh& m = *mesh;
    SpeculationDisarm disarm(m);
    m.stats = hg_solve_stats{};
    if (n_queries == 0) return;
    const int64_t n_seg = m.segments();
    // Slots: a mesh's queries take columns 0, 1, ... of the packed system in their order, so one
    // column carries one query of every mesh that has one; the batch is as wide as the mesh with
    // the most queries.
    std::vector<int64_t> next_slot(n_seg, 0), slot(n_queries);
    int64_t width = 0, max_len = 0;
    std::vector<int64_t> src_off(n_queries), dst_off(n_queries), len(n_queries);
    int64_t out_at = 0;
    for (int64_t q = 0; q < n_queries; ++q) {
      const int32_t s = query_mesh[q];
      HG_REQUIRE(s >= 0 && s < n_seg, HG_ERR_VALUE,
                 "query " + std::to_string(q) + " names mesh " + std::to_string(s) + " of " +
                     std::to_string(n_seg));
      slot[q] = next_slot[s]++;
      width = std::max(width, slot[q] + 1);
      len[q] = m.seg_v[s + 1] - m.seg_v[s];
      max_len = std::max(max_len, len[q]);
      dst_off[q] = out_at;
      out_at += len[q];
    }
    // Sources: each query's indices are its own mesh's; check them on the host copy, then shift
    // them into the packed numbering and give each the column of its query's slot.
    const int64_t total = source_offsets[n_queries] - source_offsets[0];
    std::vector<int32_t> local(static_cast<size_t>(std::max<int64_t>(total, 1)));
    if (total > 0) {
      HG_HIP(hipMemcpyAsync(local.data(), d_sources + source_offsets[0], sizeof(int32_t) * total,
                            hipMemcpyDefault, m.stream));
      HG_HIP(hipStreamSynchronize(m.stream));
    }
    std::vector<int32_t> packed(local.size());
    std::vector<int64_t> owner(local.size()), local_offsets(n_queries + 1);
    for (int64_t q = 0; q < n_queries; ++q) {
      const int32_t s = query_mesh[q];
      for (int64_t e = source_offsets[q]; e < source_offsets[q + 1]; ++e) {
        const int64_t at = e - source_offsets[0];
        HG_REQUIRE(local[at] >= 0 && local[at] < len[q], HG_ERR_VALUE,
                   "query " + std::to_string(q) + ": every entry of sources must be a vertex index "
                   "in [0, " + std::to_string(len[q]) + ")");
        packed[at] = static_cast<int32_t>(local[at] + m.seg_v[s]);
        owner[at] = slot[q];
      }
      local_offsets[q] = source_offsets[q] - source_offsets[0];
      src_off[q] = slot[q] * m.n + m.seg_v[s];
    }
    local_offsets[n_queries] = total;
    if (m.n == 0 || m.f == 0 || max_len == 0) {
      if (out_at > 0) HG_HIP(hipMemsetAsync(d_out, 0, sizeof(double) * out_at, m.stream));
      HG_HIP(hipStreamSynchronize(m.stream));
      return;
    }
    ensure_operators(m);
    // The seed reads the packed sources with their slot owners; the offsets are unused there.
    m.d_sources.assign(packed, m.stream);
    m.d_owner.assign(owner, m.stream);
    std::vector<int64_t> seed_offsets(width + 1, 0);
    seed_offsets[width] = total;
    if (m.packed_capacity < m.n * width) {
      m.packed_phi.allocate(m.n * width);
      m.packed_capacity = m.n * width;
    }
    arm_speculation(m, width, precision == HG_PRECISION_FP64);
    m.sources_on_device = false;  // the seeds' count is the host's
    solve_columns(m, seed_offsets.data(), width, m.packed_phi.ptr, HG_MODE_FACTORED, precision);
    // A speculated first round that failed its test: solve again without speculating.
    HG_HIP(hipStreamSynchronize(m.stream));
    if (speculation_failed(m)) {
      m.stats = hg_solve_stats{};
      solve_columns(m, seed_offsets.data(), width, m.packed_phi.ptr, HG_MODE_FACTORED, precision);
    }
    // Each query's segment of its slot's column, then its shift against its own sources.
    m.q_src_off.assign(src_off, m.stream);
    m.q_dst_off.assign(dst_off, m.stream);
    m.q_len.assign(len, m.stream);
    m.d_offsets.assign(local_offsets, m.stream);
    m.q_local.assign(local, m.stream);
    launch_extract_fields(m.packed_phi.ptr, m.q_src_off.ptr, m.q_dst_off.ptr, m.q_len.ptr, max_len,
                          n_queries, d_out, m.stream);
    const int32_t blocks = shift_partial_blocks(max_len);
    if (m.shift_partials.size < blocks * n_queries) m.shift_partials.allocate(blocks * n_queries);
    if (m.shift_values.size < 2 * n_queries) m.shift_values.allocate(2 * n_queries);
    {
      PhaseScope scope(&m.timer, kPhaseShift, m.stream);
      launch_shift(d_out, m.q_local.ptr, m.d_offsets.ptr, max_len, n_queries, m.q_dst_off.ptr,
                   m.q_len.ptr, m.shift_partials.ptr, m.shift_values.ptr, m.stream);
    }
    HG_HIP(hipStreamSynchronize(m.stream));
    check_seeds(m);
    m.timer.resolve();
    if (out_stats) *out_stats = m.stats;
  });
}

int hg_packed_times(hg_mesh* mesh, double* out_times) {
  return guarded([&] {
    require_mesh(mesh);
    HG_REQUIRE(out_times != nullptr, HG_ERR_VALUE, "out_times is null");
    ensure_operators(*mesh);
    for (int64_t s = 0; s < mesh->segments(); ++s) out_times[s] = mesh->seg_t[s];
  });
}

int hg_timing_enable(hg_mesh* mesh, int32_t enable) {
  return guarded([&] {
    require_mesh(mesh);
    HG_HIP(hipStreamSynchronize(mesh->stream));
    mesh->timer.resolve();
    mesh->timer.enabled = enable != 0;
  });
}

int hg_timing_read(hg_mesh* mesh, double* out_ms, int64_t* out_counts, int32_t reset) {
  return guarded([&] {
    require_mesh(mesh);
    HG_HIP(hipStreamSynchronize(mesh->stream));
    mesh->timer.resolve();
    for (int i = 0; i < kPhaseCount; ++i) {
      if (out_ms) out_ms[i] = mesh->timer.ms[i];
      if (out_counts) out_counts[i] = mesh->timer.count[i];
    }
    if (reset) mesh->timer.reset();
  });
}

int hg_nbytes(hg_mesh* mesh, int64_t* out_bytes) {
  return guarded([&] {
    require_mesh(mesh);
    int64_t bytes = mesh->heat_factor.bytes() + mesh->poisson_factor.bytes() +
                    mesh->pinned_factor.bytes() + mesh->reg_factor.bytes();
    if (bytes > 0 && mesh->plan) bytes += mesh->plan->bytes();
    *out_bytes = bytes;
  });
}

int hg_release(hg_mesh* mesh) {
  return guarded([&] {
    require_mesh(mesh);
    HG_HIP(hipStreamSynchronize(mesh->stream));
    release_workspaces(*mesh);
    mesh->heat_factor.reset();
    mesh->poisson_factor.reset();
    mesh->pinned_factor.reset();
    mesh->reg_factor.reset();
    mesh->plan.reset();
  });
}

int hg_device_bytes(hg_mesh* mesh, int64_t* out_bytes) {
  return guarded([&] {
    require_mesh(mesh);
    HG_REQUIRE(out_bytes != nullptr, HG_ERR_VALUE, "out_bytes is null");
    *out_bytes = mesh_device_bytes(*mesh);
  });
}

int hg_plan_stats(hg_mesh* mesh, int64_t* out) {
  return guarded([&] {
    require_mesh(mesh);
    ensure_operators(*mesh);
    ensure_plan(*mesh);
    const DevicePlan& p = *mesh->plan;
    out[0] = p.n_super;
    out[1] = p.n_levels;
    out[2] = p.nnz_w;
    out[3] = static_cast<int64_t>(p.flops);
    out[4] = p.arena_size;
    out[5] = p.max_front;
    out[6] = p.vec_size;
    out[7] = p.n;
  });
}

int hg_inverse_factor_entries(hg_mesh* mesh, int64_t* out3) {
  return guarded([&] {
    require_mesh(mesh);
    HG_REQUIRE(out3 != nullptr, HG_ERR_VALUE, "out3 is null");
    ensure_operators(*mesh);
    ensure_plan(*mesh);
    const Plan& p = *mesh->host_plan;
    int64_t nonzero = 0, read = 0, stored = 0;
    for (int32_t s = 0; s < p.n_super; ++s) {
      const int64_t nc = p.ncol[s], m = nc + p.nrow[s];
      stored += nc * m;
      nonzero += nc * (nc + 1) / 2 + (m - nc) * nc;
      // A forward tile of rows [t, t + 64) reads inner [0, min(nc, t + 64)) inside the leading
      // block and all nc below it (the backward reads the transposed set).
      for (int64_t t = 0; t < m; t += 64)
        read += std::min<int64_t>(64, m - t) * (t < nc ? std::min<int64_t>(nc, t + 64) : nc);
    }
    out3[0] = nonzero;
    out3[1] = read;
    out3[2] = stored;
  });
}

struct hg_plan {
  Plan plan;
};

int hg_plan_create(const int32_t* offsets, const int32_t* columns, int64_t n, const float* xyz,
                   int32_t leaf_rows, int32_t top_rows, hg_plan** out_plan) {
  return guarded([&] {
    HG_REQUIRE(out_plan != nullptr && n >= 0, HG_ERR_VALUE, "bad arguments");
    HostPattern pattern;
    pattern.n = n;
    pattern.offsets.assign(offsets, offsets + n + 1);
    pattern.columns.assign(columns, columns + offsets[n]);
    AnalysisOptions options;
    if (leaf_rows > 0) options.leaf_rows = leaf_rows;
    if (top_rows > 0) options.top_rows = top_rows;
    auto p = std::make_unique<hg_plan>();
    p->plan = analyse(pattern, xyz, options);
    *out_plan = p.release();
  });
}

int hg_plan_sizes(hg_plan* plan, int64_t* out) {
  return guarded([&] {
    HG_REQUIRE(plan != nullptr, HG_ERR_VALUE, "null plan");
    const Plan& p = plan->plan;
    out[0] = p.n_super;
    out[1] = p.n_levels;
    out[2] = static_cast<int64_t>(p.rows.size());
    out[3] = static_cast<int64_t>(p.children.size());
    out[4] = static_cast<int64_t>(p.entry_src.size());
    out[5] = p.arena_size;
    out[6] = p.w_size;
    out[7] = p.vec_size;
  });
}

int hg_plan_copy(hg_plan* plan, int32_t field, void* out) {
  return guarded([&] {
    HG_REQUIRE(plan != nullptr && out != nullptr, HG_ERR_VALUE, "null argument");
    const Plan& p = plan->plan;
    auto put32 = [&](const std::vector<int32_t>& v) {
      if (!v.empty()) std::memcpy(out, v.data(), v.size() * sizeof(int32_t));
    };
    auto put64 = [&](const std::vector<int64_t>& v) {
      if (!v.empty()) std::memcpy(out, v.data(), v.size() * sizeof(int64_t));
    };
    switch (field) {
      case 0: put32(p.perm); break;
      case 1: put32(p.col_begin); break;
      case 2: put32(p.ncol); break;
      case 3: put32(p.nrow); break;
      case 4: put32(p.parent); break;
      case 5: put32(p.level); break;
      case 6: put32(p.row_ptr); break;
      case 7: put32(p.rows); break;
      case 8: put32(p.rel); break;
      case 9: put32(p.child_ptr); break;
      case 10: put32(p.children); break;
      case 11: put32(p.level_ptr); break;
      case 12: put32(p.level_nodes); break;
      case 13: put64(p.front_off); break;
      case 14: put64(p.w_off); break;
      case 15: put64(p.vec_off); break;
      case 16: put64(p.entry_dst); break;
      case 17: put32(p.entry_src); break;
      case 18: put32(p.entry_row); break;
      case 19: put32(p.entry_col); break;
      default: throw Error{HG_ERR_VALUE, "unknown plan field"};
    }
  });
}

int hg_plan_validate(hg_plan* plan) {
  return guarded([&] {
    HG_REQUIRE(plan != nullptr, HG_ERR_VALUE, "null plan");
    validate_sweeps(plan->plan);
    validate_factor_tasks(plan->plan, factor_tasks(plan->plan));
  });
}

int hg_plan_factor_tasks(hg_plan* plan, int64_t* out_counts, int32_t* launches, int32_t* tasks) {
  return guarded([&] {
    HG_REQUIRE(plan != nullptr && out_counts != nullptr, HG_ERR_VALUE, "null argument");
    const HostFactorTasks h = factor_tasks(plan->plan);
    out_counts[0] = static_cast<int64_t>(h.launches.size());
    out_counts[1] = static_cast<int64_t>(h.tasks.size() / 4);
    if (launches) {
      for (size_t k = 0; k < h.launches.size(); ++k) {
        launches[4 * k + 0] = h.launches[k].kind;
        launches[4 * k + 1] = h.launches[k].level;
        launches[4 * k + 2] = h.launches[k].first;
        launches[4 * k + 3] = h.launches[k].count;
      }
    }
    if (tasks && !h.tasks.empty()) std::memcpy(tasks, h.tasks.data(), h.tasks.size() * sizeof(int32_t));
  });
}

int hg_plan_destroy(hg_plan* plan) {
  return guarded([&] { delete plan; });
}

}  // extern "C"
