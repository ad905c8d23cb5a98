/*
 * heatgeo.h -- C ABI of libheatgeo, the MI355X (gfx950) heat-method geodesic library.
 *
 * The boundary a host binding (ctypes here, see ordito_amd/_lib.py) calls to back ordito's
 * array-in/array-out heat path: ``ordito.heat.heat_geodesic`` (ordito heat.py:370-549),
 * ``ordito.heat.heat_operators`` (heat.py:200-367), ``ordito.heat.HeatSolver``
 * (heat.py:3491-4456) and the per-mesh cache ``ordito.mesh.Trimesh.heat_solver`` /
 * ``release_factorizations`` (mesh.py:2187-2265, 3457-3495).
 *
 * Conventions
 *  - Every pointer named d_* is device memory on the current HIP device; sizes are element counts.
 *  - Vertices are float32 xyz triples (V,3), faces a flat int32 buffer (3F,), sources int32,
 *    fields float64 laid out column-major over queries: column q of a (k, V) field starts at q*V.
 *  - Every function returns an hg status code; hg_last_error() holds the calling thread's last
 *    message. HG_ERR_VALUE maps to Python ValueError, HG_ERR_RUNTIME to RuntimeError.
 *  - Work is issued on the mesh's stream (hg_mesh_set_stream); functions that return results
 *    through host memory synchronize that stream.
 *  - A mesh is not safe for concurrent use from two threads (the reference's rule,
 *    heat.py:298-308: two solves against one operator run on one stream).
 */
#ifndef HEATGEO_H
#define HEATGEO_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define HG_OK 0
#define HG_ERR_VALUE 1   /* bad argument: out-of-range source, shape mismatch, unknown method */
#define HG_ERR_RUNTIME 2 /* HIP failure, missing state, dead mesh */
#define HG_ERR_NOT_PD 3  /* a non-positive pivot: the operator is not definite (cholesky.py:90-118) */
#define HG_ERR_BUDGET 4  /* the factorization would exceed the memory budget */

/* Discretizations (heat.py:2002-2012): the lumped mass and the mean edge length of t = h^2. */
#define HG_METHOD_MESHLAB 0     /* barycentric mass, h over every face's three edges */
#define HG_METHOD_IGL 1         /* Voronoi mass, h over the unique edges; a geodesic averages the
                                   Neumann heat with the heat pinned to zero on the rim */
#define HG_METHOD_POTPOURRI3D 2 /* barycentric mass, h over the unique edges; a geodesic's Poisson
                                   step solves -L + 1e-6 I (geometry-central's regularization) */
#define HG_METHOD_ROBUST 0x100  /* or'ed into HG_METHOD_MESHLAB / HG_METHOD_IGL: use_robust
                                   (heat.py:229-243, 2041-2052), the half-cotangents from side
                                   lengths mollified by the smallest constant that leaves every
                                   triangle-inequality margin 1e-5 of the mean side
                                   (mollify_intrinsic, laplacian.py:779-874) */

/* How a solve runs (HeatSolver policy, heat.py:4348-4423). */
#define HG_MODE_ITERATE 0  /* settled Jacobi-CG heat solve; Jacobi-Chebyshev CG Poisson solve */
#define HG_MODE_FACTORED 1 /* refined solves through the held sparse Cholesky factor */

/* Precision of a factored solve. */
#define HG_PRECISION_FP64 0  /* FP64 solve blocks and vectors, FP64 refinement */
#define HG_PRECISION_MIXED 1 /* FP32 solve blocks, FP64 residuals and refinement to a served
                                accuracy: componentwise 1e-6 (heat), normwise 1e-6 (Poisson) */

/* Which systems hg_factor builds. */
#define HG_SYSTEM_HEAT 1
#define HG_SYSTEM_POISSON 2
#define HG_SYSTEM_PINNED 4 /* the heat system pinned to zero on the rim (nothing on a closed mesh) */
#define HG_SYSTEM_REGULARIZED 8 /* HG_METHOD_POTPOURRI3D's Poisson step, -L + 1e-6 I
                                   (HeatSolver._solve_regularized_poisson, heat.py:4085-4110) */

typedef struct hg_mesh hg_mesh;

/* Device views of the assembled operators: heat_operators' 8-tuple (heat.py:169-178). */
typedef struct hg_operator_view {
  int64_t n_vertices;
  int64_t n_faces;
  int64_t nnz;
  const int32_t* offsets;   /* (V+1) CSR row offsets shared by the three operators */
  const int32_t* columns;   /* (nnz) */
  const double* laplacian;  /* (nnz) L, igl sign convention (negative diagonal) */
  const double* heat;       /* (nnz) M - t L */
  const double* poisson;    /* (nnz) -L */
  const double* mass;       /* (V) lumped mass: barycentric, or Voronoi at HG_METHOD_IGL */
  const float* cot;         /* (F,3) half-cotangents */
  const float* normals;     /* (F,3) unit face normals */
  const float* areas;       /* (F) face areas */
  double t;                 /* the diffusion time the heat system was built at */
  int32_t settle_rounds;    /* the settle window of the heat solve (heat.py:5234-5243) */
  int32_t time_from_edges;  /* 1: t is the device-formed squared mean edge length (the method's
                               edges) */
  double mollify_delta;     /* HG_METHOD_ROBUST: the constant added to every side (0: none) */
} hg_operator_view;

/* Statistics of the last solve on a mesh. */
typedef struct hg_solve_stats {
  int32_t heat_rounds;       /* CG rounds that stepped, or refinement solves */
  int32_t poisson_rounds;
  int32_t heat_refinements;  /* refinement solves of the factored heat step */
  int32_t poisson_refinements;
  int32_t backward_error_failed; /* 1 when a settled iterate failed its backward-error check */
  int32_t pinned_solves;     /* the rim-pinned heat step's CG rounds or refinement solves */
  int32_t heat_marched;      /* 1 when the heat step was marched band by band (refused under the
                                memory budget; heat_rounds then counts the bands' CG rounds) */
  int32_t poisson_multigrid; /* 1 when the Poisson step ran under the multigrid V-cycle */
} hg_solve_stats;

/* Thread-local message of the last failing call ("" if none). */
const char* hg_last_error(void);

/* Library version string. */
const char* hg_version(void);

/*
 * Bind a mesh: copies the caller's device buffers (V,3) float32 and (3F,) int32 into library-owned
 * memory and builds the operator pattern. Replaces what ordito.mesh.Trimesh(vertices, faces) plus
 * mesh_operator_pattern (laplacian.py:1071-1389) hold. Faces with an index outside [0, V) are a
 * ValueError. stream: a hipStream_t or NULL for the null stream.
 */
int hg_mesh_create(const float* d_vertices, int64_t n_vertices, const int32_t* d_faces,
                   int64_t n_faces, void* stream, hg_mesh** out_mesh);

/* Free a mesh and everything it holds (operators, factors). */
int hg_mesh_destroy(hg_mesh* mesh);

/* Issue the mesh's later work on this stream. */
int hg_mesh_set_stream(hg_mesh* mesh, void* stream);

/*
 * Assemble the heat operators at time t (heat_operators / _heat_operators, heat.py:200-367,
 * 2015-2206). t = NaN selects the default h^2 with the method's h (heat.py:2002-2012); with
 * solver_default_time set it is the solver's rule instead (HeatSolver._vertex_time,
 * heat.py:3714-3736): raised to (diag/400)^2 when that is longer, with the settle window grown to
 * match. Re-assembling drops any held factorization.
 */
int hg_operators(hg_mesh* mesh, double t, int32_t method, int32_t solver_default_time);

/* Device views of the operators (assembling at the default t first if needed). */
int hg_operator_views(hg_mesh* mesh, hg_operator_view* out_view);

/*
 * The cotangent operator's pattern built on the device from the faces (device or host memory,
 * copied into the library's first), with no mesh bound:
 * mesh_operator_pattern(faces, n_vertices) (laplacian.py:1071-1141, operator="cotmatrix"; the
 * sorted-key build of _mesh_operator_pattern, laplacian.py:1212-1389, and
 * kernels/laplacian.py:707-760). Every output buffer is caller-owned (device or host memory) and
 * may be NULL to skip it; capacities as in the reference, of which out_sizes reports what was
 * written:
 *   d_offsets (V+1) row bounds; d_columns (cap 6F+V) sorted columns per row;
 *   d_corner_ptr (cap 6F+V+1) each entry's run into d_corner_slots (the reference's run_start);
 *   d_corner_slots (cap 6F) the corner slots 3f+e opposite each entry's edge, ascending per run
 *     (the reference's payload order);
 *   d_vertex_ptr (V+1), d_vertex_corners (3F) each vertex's corners ascending (_corner_order,
 *     heat.py:3170-3190).
 * out_sizes[0] = entries (offsets[V]), out_sizes[1] = corner slots. A face index outside [0, V)
 * is a ValueError. Returns after the copies are complete.
 */
int hg_operator_pattern(const int32_t* d_faces, int64_t n_faces, int64_t n_vertices, void* stream,
                        int32_t* d_offsets, int32_t* d_columns, int32_t* d_corner_ptr,
                        int32_t* d_corner_slots, int32_t* d_vertex_ptr, int32_t* d_vertex_corners,
                        int64_t* out_sizes);

/*
 * Factor the systems named by `systems` (HeatSolver.factor, heat.py:3863-3888; the Poisson
 * factorization shares the heat analysis, cholesky.py:313-362). Sets *out_factored to 1 when
 * every named system holds a factor. `precision` HG_PRECISION_MIXED also keeps FP32 solve blocks.
 * A factorization whose peak (front arena plus kept blocks) exceeds the memory budget is refused,
 * not an error: the system stays unfactored, its solves iterate, and *out_factored is 0.
 */
int hg_factor(hg_mesh* mesh, int32_t systems, int32_t precision, int32_t* out_factored);

/* Whether each system holds a factor: bit 0 heat, bit 1 Poisson, bit 2 rim-pinned heat, bit 3 the
 * regularized Poisson system; bits 4-7 the same systems' refusals under the memory budget. */
int hg_factored(hg_mesh* mesh, int32_t* out_systems);

/*
 * Graph replay of hg_geodesic (no counterpart in the reference, whose kernels are launched one by
 * one): with every system factored in FP64 and timing off, the whole path (seeds, refined sweeps,
 * divergence, Poisson, finish) is captured once per column count as a HIP graph on a stream of the
 * library's own and replayed, the call's sources, offsets and output reaching it through staged
 * buffers. On by default; enable = 0 solves eagerly. A capture that fails leaves the mesh eager.
 * hg_graph_state reports the graphs held and whether a capture failed.
 */
int hg_set_graphs(hg_mesh* mesh, int32_t enable);
int hg_graph_state(hg_mesh* mesh, int64_t* out_cached, int32_t* out_broken);

/*
 * The factorizations' memory budget in bytes (sparse_cholesky's CHOLESKY_MEMORY_BUDGET,
 * cholesky.py:72, there a fixed 4 GiB): a factorization whose peak would exceed it is refused and
 * its system solved by iteration. 0 (the default) takes 90 % of the device memory free when each
 * factorization starts. Setting it forgets earlier refusals.
 */
int hg_set_memory_budget(hg_mesh* mesh, int64_t bytes);

/*
 * Whether a system is asked of its analysis, before it iterates, if its factorization would exceed
 * an explicitly set memory budget (default 1): a refused system is then marched at once. With 0
 * only a refusal already met counts, and a refusal is found when a rejected settled iterate falls
 * back to a factorization -- ordito's size test (_could_exceed_budget, linalg.py:2729-2745)
 * skipping the analysis of an operator it deems small enough.
 */
int hg_set_budget_probe(hg_mesh* mesh, int32_t probe);

/*
 * The settle rule of the heat solves' iteration (solve_spd_settled's check_rounds,
 * change_tolerance and settle_rounds, linalg.py:2490-2560): every check_rounds rounds the iterate
 * stops once each entry changed by less than change_tolerance of itself since the last check, or
 * once no new entry was reached for settle_rounds rounds. Defaults 16, 1e-6 and the window the
 * operators' time gives (heat.py:5234-5243), which settle_rounds 0 restores.
 */
int hg_set_settle_rule(hg_mesh* mesh, int32_t check_rounds, double change_tolerance,
                       int32_t settle_rounds);

/*
 * Refinement targets of the factored solves at one precision: the heat step's componentwise
 * tolerance and the Poisson step's normwise one, each in (0, 1); NaN keeps the current value. No
 * counterpart in the reference, whose targets are fixed (linalg.py:181, heat.py:108); the
 * defaults are those for HG_PRECISION_FP64 and 1e-6 / 1e-6 for HG_PRECISION_MIXED. A serving
 * layer sets them from the accuracy its queries ask for.
 */
int hg_set_refinement(hg_mesh* mesh, int32_t precision, double heat_tolerance,
                      double poisson_tolerance);

/*
 * Tunables of the sparse Cholesky (0 keeps the default): the nested dissection's leaf size
 * (default 128 rows) and top merge (512 rows; ordito's leaf 32 / top 512, cholesky.py:538-660),
 * and the task count a sweep level's products are split towards (1024). A new leaf or top size
 * drops the plan and every factor (the next solve analyses and factors again); a new split target
 * only rebuilds the sweep tasks. The autotuner (ordito_amd/tune.py) times candidates through it.
 */
int hg_set_plan_options(hg_mesh* mesh, int32_t leaf_rows, int32_t top_rows, int32_t split_target);

/*
 * Diffuse: solve (M - tL) u = rhs for n_columns columns of V entries (HeatSolver.diffuse,
 * heat.py:3890-3914). HG_MODE_ITERATE runs the settled Jacobi-CG (solve_spd_settled,
 * linalg.py:2490-2727) and its componentwise backward-error check; on failure it factors the
 * heat system and solves again refined (the reference's fallback), reporting it in *out_flags
 * (bit 0: the iterate failed its check, bit 1: a factorization was built, bit 2: the system's
 * factorization is refused under the memory budget and the system was marched band by band
 * (_march_settled, linalg.py:2765-2939): at once when the refusal was known before iterating, as
 * after a refusal or with a budget set that the analysis says the factorization exceeds, else in
 * place of the rejected iterate). HG_MODE_FACTORED solves through the held factor, refined
 * componentwise to 1e-12 (linalg.py:2659-2665), factoring first if none is held (marching when
 * that is refused). d_solution is overwritten (the iterate starts from zero).
 */
int hg_diffuse(hg_mesh* mesh, const double* d_rhs, double* d_solution, int64_t n_columns,
               int32_t mode, int32_t precision, int32_t* out_flags);

/*
 * Diffuse with the solution pinned to zero on the mesh's rim (HeatSolver.diffuse_pinned,
 * heat.py:4153-4200): (M - tL) u = rhs on the interior vertices, u = 0 on every vertex with an edge
 * on a single face, the system built on first use (heat.py:4202-4233) and settled, verified and
 * factored as hg_diffuse's; the rim entries of d_rhs are ignored. *out_has_rim is 0 for a mesh
 * without a rim, whose d_solution is then left as it was.
 */
int hg_diffuse_pinned(hg_mesh* mesh, const double* d_rhs, double* d_solution, int64_t n_columns,
                      int32_t mode, int32_t precision, int32_t* out_flags, int32_t* out_has_rim);

/* The number of rim vertices (an edge on a single face; kernels/laplacian.py:1286-1321). */
int hg_rim_count(hg_mesh* mesh, int64_t* out_count);

/*
 * Solve the Poisson system -L phi = rhs (HeatSolver.solve_poisson, heat.py:4062-4151): iterated
 * by conjugate gradient under the Jacobi-Chebyshev polynomial to relative residual 1e-8, or
 * through the held factor with normwise refinement to 1e-8 (heat.py:4138-4140). One row per
 * singular component is pinned and each component's constant set to zero mean.
 */
int hg_solve_poisson(hg_mesh* mesh, const double* d_rhs, double* d_solution, int64_t n_columns,
                     int32_t mode, int32_t precision);

/*
 * The heat method's right-hand side: -div of the unit gradient of each heat column
 * (_distance_from_heat / unit_gradient_divergence, heat.py:3052-3131, kernels/heat.py:626-687),
 * summed per vertex in ascending corner order.
 */
int hg_unit_gradient_divergence(hg_mesh* mesh, const double* d_heat, double* d_neg_divergence,
                                int64_t n_columns);

/*
 * Seed the source indicators: column q gets 1 at each of its sources d_sources[off[q]..off[q+1])
 * (kernels/heat.py:599-623). An index outside [0, V) is a ValueError (heat.py:5488-5503).
 * d_source_offsets is a host array of n_columns+1 entries.
 */
int hg_seed_sources(hg_mesh* mesh, const int32_t* d_sources, const int64_t* source_offsets,
                    int64_t n_columns, double* d_u0);

/*
 * Shift each column so its mean over its sources is zero and orient it positive
 * (_shift_to_sources, heat.py:3134-3162; kernels/heat.py:3383-3424).
 */
int hg_shift_to_sources(hg_mesh* mesh, double* d_phi, const int32_t* d_sources,
                        const int64_t* source_offsets, int64_t n_columns);

/*
 * hg_shift_to_sources without allocating: no counterpart in the reference. For callers that cannot
 * allocate device memory (code under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc): the same kernels and result as hg_shift_to_sources, with the queries' offsets, the
 * decisions and the column partials in a fixed device array of the library (8 MiB + 128 KiB per
 * device, held until the process ends) instead of per-mesh buffers. Limits: n_columns <= 4096 and
 * n_columns * ceil(V / 256) <= 2^20 (about 64 columns at a million vertices); past them
 * HG_ERR_VALUE. Every call on one device shares that array: calls that overlap on a device (two
 * streams, two threads) race. Synchronizes the mesh's stream before returning.
 */
int hg_shift_to_sources_noalloc(hg_mesh* mesh, double* d_phi, const int32_t* d_sources,
                                const int64_t* source_offsets, int64_t n_columns);

/*
 * The whole path for n_queries source sets (heat_geodesic, heat.py:370-609): seed, diffuse (at
 * HG_METHOD_IGL averaged with the rim-pinned diffusion on a mesh with a rim), divergence, Poisson,
 * shift. Query q's sources are d_sources[off[q]..off[q+1]) with
 * source_offsets a host array; its (V,) float64 distance goes to d_out + q*V. A query with no
 * sources returns zeros. Statistics of the call go to *out_stats when not NULL.
 */
int hg_geodesic(hg_mesh* mesh, const int32_t* d_sources, const int64_t* source_offsets,
                int64_t n_queries, double* d_out, int32_t mode, int32_t precision,
                hg_solve_stats* out_stats);

/*
 * Packed meshes: many meshes bound as one block-diagonal system, so every solve, level sweep and
 * factorization launch covers all of them at once. No counterpart in the reference, which solves
 * one mesh per call (heat.py:370-609); each mesh keeps what it would have alone (its own diffusion
 * time rule, components, pins and shift). Mesh s has vertices
 * d_vertices[vertex_offsets[s] .. vertex_offsets[s+1]) and faces
 * d_faces[3 face_offsets[s] .. 3 face_offsets[s+1]) indexing its own vertices from 0; the offset
 * arrays are host arrays of n_meshes + 1 entries starting at 0. Solves run in HG_MODE_FACTORED;
 * hg_geodesic, hg_shift_to_sources and HG_MODE_ITERATE refuse a packed mesh (HG_ERR_VALUE).
 */
int hg_mesh_create_packed(const float* d_vertices, const int64_t* vertex_offsets,
                          const int32_t* d_faces, const int64_t* face_offsets, int64_t n_meshes,
                          void* stream, hg_mesh** out_mesh);

/*
 * n_queries source sets on a packed mesh: query q asks for mesh query_mesh[q] (host array) from
 * the sources d_sources[off[q]..off[q+1]) (indices of that mesh's vertices; source_offsets is a
 * host array). A mesh's queries fill columns 0, 1, ... of one batched solve. Query q's distance
 * (its mesh's V_q entries, float64) goes to d_out at the sum of the earlier queries' V.
 */
int hg_geodesic_packed(hg_mesh* mesh, const int32_t* d_sources, const int64_t* source_offsets,
                       const int32_t* query_mesh, int64_t n_queries, double* d_out,
                       int32_t precision, hg_solve_stats* out_stats);

/* Each packed mesh's diffusion time (n_meshes entries; assembles at the default time if needed). */
int hg_packed_times(hg_mesh* mesh, double* out_times);

/*
 * Device-time instrumentation (no counterpart in the reference, which quotes measurements in
 * comments): with timing enabled every phase of a solve is bracketed by HIP events on the mesh's
 * stream. hg_timing_read returns, per phase, the accumulated milliseconds and the span count:
 * 0 numeric factorization, 1 heat triangular sweep (forward + backward), 2 Poisson sweep,
 * 3 refinement residual, 4 divergence, 5 shift, 6 conjugate gradient, 7 assembly.
 */
int hg_timing_enable(hg_mesh* mesh, int32_t enable);
int hg_timing_read(hg_mesh* mesh, double* out_ms8, int64_t* out_counts8, int32_t reset);

/*
 * The geodesic query server: continuous batching over resident per-mesh factors, with admission
 * by memory. No counterpart in the reference, which answers one call at a time (heat.py:370-609)
 * and keeps a mesh's factors in its HeatSolver (mesh.py:2187-2265). `workers` threads (1-16) each
 * own a HIP stream; a free worker takes the mesh whose oldest waiting request is oldest, with up
 * to `max_batch` of its requests, and solves them as one hg_geodesic batch through the mesh's
 * factors (built on its first batch, at `precision`), so batches of different meshes overlap on
 * the device and requests that arrive meanwhile join their mesh's next batch. Before a mesh is
 * factored the least recently used idle meshes are released (hg_release) until its footprint --
 * its analysis' factor and front-arena bytes and a max_batch-wide batch's workspaces -- fits
 * `budget_bytes` beside every device byte the resident meshes hold.
 *
 * hg_server_add_mesh binds a mesh (copied, as hg_mesh_create; MeshLab's discretization at the
 * solver's default time) and returns its id. hg_server_submit queues one query: `sources` is a
 * host array copied at the call; the (V,) float64 distance is written to `d_out` (device memory
 * the caller keeps until the ticket completes). hg_server_wait waits for a ticket (timeout_ms < 0:
 * without limit) and sets *out_done; a completed ticket is consumed by the wait that sees it, and
 * a query's failure is that wait's status. hg_server_stats reports the queries completed at or
 * after `since_seconds` (seconds since the server's creation). hg_server_destroy answers every
 * queued query first, then frees the meshes.
 */
typedef struct hg_server hg_server;
typedef struct hg_server_report {
  int64_t completed;      /* queries completed since `since_seconds` */
  int64_t batches;        /* batches solved since creation */
  int64_t evictions;      /* meshes released to admit others */
  int32_t resident;       /* meshes holding factors now */
  int32_t reserved;
  int64_t resident_bytes; /* device bytes the meshes hold now */
  double queries_per_s;   /* completed / (last completion - first submission) */
  double p50_ms, p99_ms, mean_ms; /* latency from submission to completion */
  double now_seconds;     /* seconds since the server's creation */
} hg_server_report;

int hg_server_create(int64_t budget_bytes, int32_t workers, int32_t max_batch, int32_t precision,
                     hg_server** out_server);
int hg_server_add_mesh(hg_server* server, const float* d_vertices, int64_t n_vertices,
                       const int32_t* d_faces, int64_t n_faces, int64_t* out_mesh_id);
int hg_server_submit(hg_server* server, int64_t mesh_id, const int32_t* sources,
                     int64_t n_sources, double* d_out, int64_t* out_ticket);
int hg_server_wait(hg_server* server, int64_t ticket, double timeout_ms, int32_t* out_done);
int hg_server_stats(hg_server* server, double since_seconds, hg_server_report* out_stats);
int hg_server_destroy(hg_server* server);

/* Statistics of the mesh's last solve (hg_diffuse, hg_diffuse_pinned, hg_solve_poisson,
 * hg_geodesic, hg_geodesic_packed). No counterpart in the reference, which logs them. */
int hg_last_stats(hg_mesh* mesh, hg_solve_stats* out_stats);

/* Device bytes the held factorizations take (HeatSolver.nbytes, heat.py:3838-3847). */
int hg_nbytes(hg_mesh* mesh, int64_t* out_bytes);

/* Drop every held factorization and the per-call solve workspaces; the operators stay
 * (HeatSolver.release, heat.py:3849-3861). */
int hg_release(hg_mesh* mesh);

/*
 * Every device byte the mesh holds: its copy of the mesh, the pattern, the operators, the plan and
 * factors, and the per-call workspaces (which grow to the widest batch served). No counterpart in
 * the reference; a serving layer sums it over its resident meshes to admit new ones by memory.
 */
int hg_device_bytes(hg_mesh* mesh, int64_t* out_bytes);

/* Analysis statistics of the mesh's plan (analysing it if needed): supernodes, levels, stored
 * solve-block entries, factorization flops, arena entries, largest front, front-vector entries
 * (n + total structure rows), rows. */
int hg_plan_stats(hg_mesh* mesh, int64_t* out_stats8);

/*
 * The inverse factors' entry counts (analysing if needed): out3[0] the structural nonzeros (each
 * front's lower-triangular leading block and its rows below), out3[1] the entries one sweep
 * direction reads (whole 64-row tiles of the leading block), out3[2] the stored entries (every
 * front's full ncol x m block). No counterpart in the reference: the algorithmic bytes of the
 * sweep's roofline are the nonzeros'.
 */
int hg_inverse_factor_entries(hg_mesh* mesh, int64_t* out3);

/*
 * Host-only access to the Cholesky analysis (sparse_cholesky's ordering and symbolic phase,
 * cholesky.py:538-1273), for inspection and tests; needs no GPU. The pattern is a symmetric CSR
 * with sorted rows, xyz float32 (n,3) host coordinates for the geometric nested dissection.
 */
typedef struct hg_plan hg_plan;
int hg_plan_create(const int32_t* offsets, const int32_t* columns, int64_t n, const float* xyz,
                   int32_t leaf_rows, int32_t top_rows, hg_plan** out_plan);
/* Sizes: [n_super, n_levels, total structure rows, total children, lower entries, arena entries,
 *         solve-block entries, vector entries]. */
int hg_plan_sizes(hg_plan* plan, int64_t* out8);
/* Copy one array out: 0 perm, 1 col_begin, 2 ncol, 3 nrow, 4 parent, 5 level, 6 row_ptr, 7 rows,
 * 8 rel, 9 child_ptr, 10 children, 11 level_ptr, 12 level_nodes (int32); 13 front_off, 14 w_off,
 * 15 vec_off, 16 entry_dst (int64); 17 entry_src, 18 entry_row, 19 entry_col (int32). */
int hg_plan_copy(hg_plan* plan, int32_t field, void* out);
/* Checks every index the device sweeps and the numeric factorization form from this plan (task
 * ranges, partial slots, gather sources, tile tasks) against its buffer; HG_ERR_RUNTIME names the
 * first violation. */
int hg_plan_validate(hg_plan* plan);
/* The numeric factorization's tile launches in execution order (the blocked form of _factor,
 * cholesky.py:1399-1540, kernels/cholesky.py:289-768): out_counts = [launches, tasks];
 * when the arrays are non-null, launches receives 4 int32 per launch (kind 0 extend-add, 1 panel,
 * 2 update; level; first task; task count) and tasks 4 int32 per task. */
int hg_plan_factor_tasks(hg_plan* plan, int64_t* out_counts, int32_t* launches, int32_t* tasks);
int hg_plan_destroy(hg_plan* plan);

#ifdef __cplusplus
}
#endif

#endif /* HEATGEO_H */
