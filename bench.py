"""
Benchmark of the heat-method geodesic path on one MI355X.

Workload (the single-GPU configuration the headline metric is quoted on): a torus of 1024 x 512
quads (524,288 vertices, 1,048,576 faces), R = 1, r = 0.3, normal jitter N(0, 0.1 h), FP32
positions, 64 source vertices drawn from a seed; the diffusion time is the solver's default rule
(h^2 raised to (bbox diagonal / 400)^2). The systems are factored once (setup, reported apart) and
stay resident in HBM; a *step* is one batch of the 64 single-source queries through the served
path: seed, heat solve, unit-gradient divergence, Poisson solve, shift -- all on the GPU.

- ``value``: served queries/s over the timed steps: the 64 queries of a step as one batch, one
  64-column refined solve per system through the FP64 inverse factors, with two batches in flight
  (``--lanes``): two handles of the mesh (each with its own factors, the serving layer's lanes) on
  two host threads and HIP streams, the K steps split between them, so one batch's latency-bound
  sweep levels overlap the other's bandwidth-bound ones. ``single_lane`` is the same with one
  batch at a time (the form the roofline and phase times are measured in).
- ``served_mixed``: the same batch through FP32 inverse factors with FP64 refinement (the
  mixed-precision form), timed beside it, with its refinement rounds and its error against the
  served fields.
- ``plain_fp64``: the reference's form, one query per call through the FP64 factors.
- ``roofline``: the triangular sweep of the served path (forward + backward, the dominant kernel),
  algorithmic bytes per sweep over its HIP-event-measured mean duration against the 8 TB/s HBM
  peak; ``traffic`` is the HBM bytes per sweep that rocprofv3's FETCH_SIZE / WRITE_SIZE passes
  measured (profiles/triangular_sweep_torus1024x512_traffic.json, scaled by a 1 GiB copy:
  ``bench.py --calibrate-copy``), null when the sweep kernels' source has changed since.
- ``cpu_baseline``: the FP64 oracle (SciPy SuperLU, one thread) on a bounded sample of the same
  queries on this host, factorization excluded as for the GPU; its results also give the error
  of the GPU fields.

    python bench.py [--gpus 1] [--steps K] [--warmup W]

Three further workloads print their own JSON line (the default run above is the headline):
``--workload sphere`` (icosphere(9) displaced by spherical harmonics, 2.6 M vertices, 16 queries,
FP64 and mixed refined to 1e-10, mixed against FP64 at full size and both against the oracle on the
same surface at icosphere(7)), ``--workload batch`` (256 meshes of
1e3-2e5 vertices packed into one system, one query each) and ``--workload serve`` (64 meshes of
1e4-1e6 vertices behind the continuous-batching server, Poisson arrivals at 2,000 queries/s for
60 s, 200 GB budget).
"""

from __future__ import annotations

import argparse
import json
import os
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

METRIC = "heat-geodesic queries/s & p50/p99 latency, served vs plain, ≤1e-4 rel err"
HBM_PEAK_GBS = 8000.0


def parse() -> argparse.Namespace:
    parser = argparse.ArgumentParser()
    parser.add_argument("--gpus", type=int, default=1)
    parser.add_argument("--steps", type=int, default=30)
    parser.add_argument("--warmup", type=int, default=3)
    parser.add_argument("--queries", type=int, default=64)
    parser.add_argument("--plain-queries", type=int, default=8)
    parser.add_argument("--lanes", type=int, default=2,
                        help="batches in flight for the served value (handles of the mesh)")
    parser.add_argument("--cpu-sample", type=int, default=4)
    parser.add_argument(
        "--cpu-sample-all", type=int, default=32, help="queries of the all-cores CPU baseline"
    )
    parser.add_argument("--no-cpu-baseline", action="store_true")
    parser.add_argument("--no-mixed", action="store_true", help="skip the served_mixed line")
    parser.add_argument(
        "--no-extra",
        action="store_true",
        help="skip the sub-objects of the other workloads (one-shot C1/C2, displaced sphere, "
        "256 packed meshes, 64 served meshes)",
    )
    parser.add_argument("--extra-steps", type=int, default=10, help="timed steps of the extras")
    parser.add_argument(
        "--serve-probes", type=str, default="6000,12000", help="capacity probe rates (serve extra)"
    )
    parser.add_argument("--major", type=int, default=1024)
    parser.add_argument("--minor", type=int, default=512)
    parser.add_argument(
        "--calibrate-copy",
        action="store_true",
        help="only run three 1 GiB device copies (the reference kernel for scaling the FETCH_SIZE "
        "and WRITE_SIZE counters) and exit",
    )
    parser.add_argument(
        "--workload",
        choices=("torus", "sphere", "batch", "serve", "igl", "oneshot", "pattern"),
        default="torus",
        help="torus: the headline configuration (default); sphere: the 2.6 M-vertex displaced "
        "sphere, mixed and FP64 against the oracle; batch: many meshes packed into one system; "
        "serve: the continuous-batching server under Poisson arrivals; igl: method='igl' on an "
        "open grid (its rim-pinned heat solve), beside method='meshlab'; oneshot: only the one-shot "
        "(iterated) calls on icosphere(4) and the torus; pattern: the operator pattern built on "
        "the device (mesh_operator_pattern) for the torus, --steps builds",
    )
    parser.add_argument(
        "--summarize-profiles",
        metavar="DIR",
        help="print the markdown summary of the rocprofv3 outputs under DIR (kt/, fetch/, write/, "
        "mfma/, copy_fetch/, copy_write/) and exit",
    )
    parser.add_argument(
        "--traffic-out",
        help="with --summarize-profiles: write the sweep's measured HBM bytes per sweep there "
        "(the record bench.py reads back as roofline.traffic)",
    )
    parser.add_argument(
        "--solve-block-entries",
        type=int,
        default=0,
        help="with --traffic-out: the profiled plan's stored solve-block entries",
    )
    many = parser.add_argument_group("--workload batch / serve")
    many.add_argument("--meshes", type=int, help="default 256 (batch), 64 (serve)")
    many.add_argument("--min-vertices", type=float, help="default 1e3 (batch), 1e4 (serve)")
    many.add_argument("--max-vertices", type=float, help="default 2e5 (batch), 1e6 (serve)")
    many.add_argument("--seed", type=int, default=0)
    many.add_argument("--tune", action="store_true", help="batch: autotune first")
    many.add_argument("--rate", type=float, default=2000.0, help="serve: Poisson arrivals per second")
    many.add_argument("--duration", type=float, default=60.0, help="serve: seconds of arrivals")
    many.add_argument("--budget-gb", type=float, default=200.0, help="serve: HBM budget")
    many.add_argument("--workers", type=int, default=4, help="serve: batches solved at once")
    many.add_argument("--pack-vertices", type=int, default=0,
                      help="serve: meshes of at most this many vertices served in packed groups")
    args = parser.parse_args()
    defaults = {"batch": (256, 1e3, 2e5), "serve": (64, 1e4, 1e6)}.get(args.workload)
    if defaults:
        args.meshes = args.meshes or defaults[0]
        args.min_vertices = args.min_vertices or defaults[1]
        args.max_vertices = args.max_vertices or defaults[2]
    return args


def sphere_result(args, oracle_check: bool = True) -> dict:
    """
    The large-mesh configuration: an icosphere subdivided 9 times (2,621,442 vertices) displaced by
    real spherical harmonics up to degree 16 with N(0, 0.05 / l) coefficients, 16 single-source
    queries as one batch, through FP64 factors and through FP32 factors refined to 1e-10
    (componentwise heat, normwise Poisson). Errors: the mixed fields against the FP64 fields at full
    size, and both against the FP64 oracle (SciPy SuperLU) on the same surface subdivided 7 times
    (163,842 vertices), since SuperLU's factorization of the full size exhausts its memory.
    """
    import torch

    from ordito_amd import creation, heat
    from oracle import heat_oracle as ho

    def solve_both(vertices, faces, sources, steps, warmup):
        sets = [torch.tensor([s], dtype=torch.int32, device="cuda") for s in sources]
        v_gpu, f_gpu = torch.from_numpy(vertices).cuda(), torch.from_numpy(faces).cuda()
        fields, lines = {}, {}
        for precision in ("fp64", "mixed"):
            mesh = heat.Trimesh(v_gpu, f_gpu)
            solver = mesh.heat_solver(method="meshlab", precision=precision)
            if precision == "mixed":
                solver.set_refinement(heat=1e-10, poisson=1e-10)
            t0 = time.perf_counter()
            assert solver.factor()
            torch.cuda.synchronize()
            setup = time.perf_counter() - t0
            elapsed, latencies, timing, out, stats = timed_steps(solver, sets, steps, warmup)
            fields[precision] = out.cpu().numpy()
            lines[precision] = {
                "value": round(len(sets) * steps / elapsed, 2),
                "ms_per_step": round(1e3 * elapsed / steps, 3),
                "p50_ms": ms(latencies, 50),
                "p99_ms": ms(latencies, 99),
                "factor_s": round(setup, 3),
                "factor_gb": round(solver.nbytes / 1e9, 3),
                "refinements": stats,
                "phase_ms_per_step": {k: round(v[0] / steps, 3) for k, v in timing.items() if v[1]},
                "t": solver.time,
            }
            log(f"   {precision}: {lines[precision]['value']} queries/s, {stats}")
            del solver, mesh
            torch.cuda.empty_cache()
        return fields, lines

    log("== generating the displaced sphere (icosphere(9), SH degree 16)")
    vertices, faces = creation.displaced_sphere(9, lmax=16, amplitude=0.05, seed=args.seed)
    n = len(vertices)
    sources = creation.random_sources(n, 16, seed=args.seed + 1)
    fields, lines = solve_both(vertices, faces, sources, args.steps, args.warmup)
    result = {
        "metric": METRIC,
        "unit": "queries/s",
        "n_gpus": 1,
        "higher_is_better": True,
        "data": "synthetic",
        "config": {
            "workload": "icosphere(9) displaced by real SH (l <= 16, N(0, 0.05/l)), 16 "
            "single-source queries as one batch, factors resident",
            "n_vertices": int(n),
            "n_faces": int(faces.size // 3),
        },
        "value": lines["fp64"]["value"],
        "dtype": "f64",
        "fp64": lines["fp64"],
        "mixed": lines["mixed"],
    }
    diffs = [ho.relative_error(fields["mixed"][q], fields["fp64"][q]) for q in range(len(sources))]
    result["mixed"]["vs_fp64_rel_err_max"] = max(d[0] for d in diffs)
    result["mixed"]["vs_fp64_rel_err_mean"] = float(np.mean([d[1] for d in diffs]))
    del fields
    if not oracle_check:
        return result
    log("== the oracle on the same surface at icosphere(7) (SciPy SuperLU under COLAMD, 1 thread)")
    v7, f7 = creation.displaced_sphere(7, lmax=16, amplitude=0.05, seed=args.seed)
    sample = creation.random_sources(len(v7), max(1, args.cpu_sample), seed=args.seed + 1)
    fields7, _ = solve_both(v7, f7, sample, 1, 1)
    t0 = time.perf_counter()
    oracle = ho.FactoredOracle(v7, f7, ordering="COLAMD")
    log(f"   oracle factored in {time.perf_counter() - t0:.1f}s")
    check = {"n_vertices": int(len(v7)), "queries": len(sample),
             "oracle_factor_s": round(time.perf_counter() - t0, 1)}
    for precision in ("fp64", "mixed"):
        errs = [ho.relative_error(fields7[precision][q], oracle.distance(np.array([s])))
                for q, s in enumerate(sample)]
        check[precision] = {"rel_err_max": max(e[0] for e in errs),
                            "rel_err_mean": float(np.mean([e[1] for e in errs]))}
        log(f"   {precision}: max {check[precision]['rel_err_max']:.3g}")
    result["oracle_check_icosphere7"] = check
    return result


def sphere_workload(args) -> int:
    print(json.dumps(sphere_result(args)))
    return 0


def igl_workload(args) -> int:
    """
    ``method="igl"`` on an open surface: a jittered 724 x 724 planar grid (524,176 vertices, about
    the torus's size, 2,892 on the rim), 64 single-source queries a step through FP64 factors of all
    three systems (the Neumann heat, the rim-pinned heat, the Poisson), beside ``method="meshlab"``
    on the same grid; the error of ``"igl"`` against the FP64 oracle on a 64 x 64 grid of the same
    kind. Prints one JSON line.
    """
    import torch

    from ordito_amd import creation, heat
    from oracle import heat_oracle as ho

    vertices, faces = creation.planar_grid(724, 724, 1.0, 0.2, args.seed)
    rng = np.random.default_rng(args.seed + 1)
    sources = rng.integers(0, len(vertices), size=args.queries).astype(np.int32)
    source_sets = [torch.tensor([s], dtype=torch.int32, device="cuda") for s in sources]
    lines = {}
    for method in ("meshlab", "igl"):
        mesh = heat.Trimesh(torch.from_numpy(vertices).cuda(), torch.from_numpy(faces).cuda())
        solver = mesh.heat_solver(method=method, precision="fp64")
        assert solver.factor()
        solver._solves = {"heat": 2, "poisson": 2, "pinned": 2}  # past rent-or-buy
        log(f"== {method}: {args.steps} steps of {len(source_sets)} queries")
        elapsed, latencies, timing, _, stats = timed_steps(solver, source_sets, args.steps, args.warmup)
        lines[method] = {
            "value": round(len(source_sets) * args.steps / elapsed, 2),
            "ms_per_step": round(1e3 * elapsed / args.steps, 3),
            "p50_ms": ms(latencies, 50),
            "p99_ms": ms(latencies, 99),
            "phase_ms_per_step": {k: round(v[0] / args.steps, 3) for k, v in timing.items() if v[1]},
            "refinements": stats,
            "factor_gb": round(solver.nbytes / 1e9, 3),
        }
        log(f"   {lines[method]['value']} queries/s, {stats}")
        del solver, mesh
        torch.cuda.empty_cache()
    small_v, small_f = creation.planar_grid(64, 64, 1.0, 0.2, args.seed)
    picks = [0, len(small_v) // 2 + 17]
    errors = [
        ho.relative_error(
            heat.heat_geodesic(small_v, small_f, np.array([p], np.int32), method="igl"),
            ho.geodesic(small_v, small_f, np.array([p]), method="igl").distance,
        )[0]
        for p in picks
    ]
    result = {
        "metric": METRIC,
        "unit": "queries/s",
        "n_gpus": 1,
        "higher_is_better": True,
        "data": "synthetic",
        "dtype": "f64",
        "config": {
            "workload": "planar grid 724x724 jittered 0.2h, 64 single-source queries a step as one "
            "batch, factors resident",
            "n_vertices": int(len(vertices)),
        },
        "value": lines["igl"]["value"],
        "igl": lines["igl"],
        "meshlab": lines["meshlab"],
        # One-shot calls (a new mesh each: the settled CG, the iterated bar 1e-5).
        "igl_one_shot_oracle_rel_err_max_64x64": max(errors),
    }
    print(json.dumps(result))
    return 0


def batch_result(args) -> dict:
    """
    The batched configuration: ``--meshes`` (default 256) meshes with log-uniform vertex counts in
    [``--min-vertices``, ``--max-vertices``] (default 1e3..2e5), one query each with 1-16 random
    sources, packed into one block-diagonal system (``MeshBatch``): one factorization and one
    batched solve for all of them, every level sweep one launch over every mesh. ``--tune`` runs
    the autotuner first. Prints one JSON line: queries/s over the timed batches, setup, and the
    error of sampled answers against the FP64 oracle.
    """
    import torch

    from ordito_amd import creation
    from ordito_amd.batch import MeshBatch

    rng = np.random.default_rng(args.seed)
    log(f"== generating {args.meshes} meshes")
    meshes = creation.mesh_collection(args.meshes, args.min_vertices, args.max_vertices, rng)
    sizes = [len(meshes[i][0]) for i in range(args.meshes)]
    queries = []
    for i in range(args.meshes):
        count = min(int(rng.integers(1, 17)), sizes[i])
        queries.append((i, rng.choice(sizes[i], size=count, replace=False).astype(np.int32)))
    t0 = time.perf_counter()
    batch = MeshBatch([(torch.from_numpy(v).cuda(), torch.from_numpy(f).cuda()) for v, f in meshes.values()])
    bind_s = time.perf_counter() - t0
    tuned = None
    if args.tune:
        from ordito_amd.tune import autotune

        log("== autotuning")
        tuned = autotune(batch, repeats=2)
        log(f"   best {tuned['best']}")
    t0 = time.perf_counter()
    assert batch.factor()
    torch.cuda.synchronize()
    factor_s = time.perf_counter() - t0
    device_queries = [(m, torch.from_numpy(s).cuda()) for m, s in queries]
    log(f"== {args.steps} batches of {len(queries)} queries on {sum(sizes)} vertices")
    for _ in range(args.warmup):
        batch.geodesic(device_queries)
    torch.cuda.synchronize()
    latencies = []
    fields = None
    start = time.perf_counter()
    for _ in range(args.steps):
        s0 = time.perf_counter()
        fields = batch.geodesic(device_queries)
        torch.cuda.synchronize()
        latencies.append(time.perf_counter() - s0)
    elapsed = time.perf_counter() - start
    from oracle import heat_oracle as ho

    errors = []
    for i in sorted(range(args.meshes), key=lambda i: sizes[i])[:3]:
        v, f = meshes[i]
        errors.append(ho.relative_error(fields[i].cpu().numpy(), ho.geodesic(v, f, queries[i][1]).distance)[0])
    result = {
        "metric": METRIC,
        "value": round(len(queries) * args.steps / elapsed, 2),
        "unit": "queries/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(1e3 * elapsed / args.steps, 3),
        "higher_is_better": True,
        "dtype": "f64",
        "data": "synthetic",
        "config": {
            "workload": f"{args.meshes} meshes packed block-diagonally, V log-uniform in "
            f"[{args.min_vertices:.0f}, {args.max_vertices:.0f}] (icosphere / torus / planar "
            f"grid), one query of 1-16 sources each, all in one batched solve",
            "total_vertices": int(sum(sizes)),
        },
        "p50_ms": ms(latencies, 50),
        "p99_ms": ms(latencies, 99),
        "setup_s": {"bind": round(bind_s, 3), "factor": round(factor_s, 3)},
        "refinements": dict(batch.last_stats),
        "autotune_best": tuned["best"] if tuned else None,
        "sampled_rel_err_max": max(errors),
        "factor_gb": round(batch.nbytes / 1e9, 3),
    }
    batch.close()
    return result


def batch_workload(args) -> int:
    print(json.dumps(batch_result(args)))
    return 0


def serve_result(args, probe_rates=()) -> dict:
    """
    The serving configuration: ``--meshes`` meshes with log-uniform vertex counts in
    [min, max] (seeded jittered icospheres at the nearest level, tori and jittered planar grids),
    queries arriving as a Poisson process at ``--rate`` per second for ``--duration`` seconds, each
    asking for 1-32 random sources on a uniformly chosen mesh, served by ``GeodesicServer`` under a
    ``--budget-gb`` HBM budget. Every mesh is warmed (factored) first; that setup is reported
    apart. Prints one JSON line: throughput, p50/p99 latency, and the error of sampled answers
    against the FP64 oracle.
    """
    from ordito_amd import creation
    from ordito_amd.serve import GeodesicServer

    rng = np.random.default_rng(args.seed)
    log(f"== generating {args.meshes} meshes")
    meshes = creation.mesh_collection(args.meshes, args.min_vertices, args.max_vertices, rng)
    sizes = {i: len(v) for i, (v, _) in meshes.items()}
    server = GeodesicServer(meshes, budget_bytes=args.budget_gb * 1e9, workers=args.workers,
                            pack_vertices=args.pack_vertices)
    log(f"== warming {args.meshes} meshes ({sum(sizes.values())} vertices)")
    t0 = time.perf_counter()
    server.warm()
    setup = time.perf_counter() - t0
    log(f"   warm in {setup:.1f}s; {args.duration:.0f}s of arrivals at {args.rate:.0f}/s")
    start = time.perf_counter()
    arrival = start
    submitted = 0
    small = set(sorted(meshes, key=lambda i: sizes[i])[:3])
    sampled = []
    next_report = start + 5.0
    while True:
        arrival += rng.exponential(1.0 / args.rate)
        if arrival - start > args.duration:
            break
        delay = arrival - time.perf_counter()
        if delay > 0:
            time.sleep(delay)
        m = int(rng.integers(args.meshes))
        count = min(int(rng.integers(1, 33)), sizes[m])
        sources = rng.choice(sizes[m], size=count, replace=False).astype(np.int32)
        future = server.submit(m, sources)
        submitted += 1
        # Answers are kept only for the sampled accuracy check: a client holds its own results.
        if m in small and len(sampled) < 3:
            sampled.append((m, sources, future))
        if time.perf_counter() > next_report:
            log(f"   {submitted} submitted, {server.stats(since=start).get('completed', 0)} done")
            next_report += 5.0
    server.drain()
    stats = server.stats(since=start)
    # Capacity: short streams at higher offered rates, each measured on its own.
    probes = []
    for rate in probe_rates:
        log(f"   probe: 10s of arrivals at {rate:.0f}/s")
        p_start = time.perf_counter()
        p_arrival = p_start
        offered = 0
        while True:
            p_arrival += rng.exponential(1.0 / rate)
            if p_arrival - p_start > 10.0:
                break
            delay = p_arrival - time.perf_counter()
            if delay > 0:
                time.sleep(delay)
            m = int(rng.integers(args.meshes))
            count = min(int(rng.integers(1, 33)), sizes[m])
            server.submit(m, rng.choice(sizes[m], size=count, replace=False).astype(np.int32))
            offered += 1
        server.drain()
        p = server.stats(since=p_start)
        probes.append({"offered_per_s": round(offered / 10.0, 1),
                       "completed_per_s": round(p["queries_per_s"], 1),
                       "p50_ms": round(p["p50_ms"], 3), "p99_ms": round(p["p99_ms"], 3),
                       "lanes": p["lanes"], "resident_gb": round(p["resident_bytes"] / 1e9, 3)})
        log(f"   {probes[-1]}")
    server.close()  # every queued request finishes first
    # Accuracy: sampled answers on the smallest meshes against the FP64 oracle.
    from oracle import heat_oracle as ho

    errors = []
    for m, sources, future in sampled:
        v, f = meshes[m]
        errors.append(ho.relative_error(future.result().cpu().numpy(), ho.geodesic(v, f, sources).distance)[0])
    result = {
        "metric": METRIC,
        "value": round(stats["queries_per_s"], 2),
        "unit": "queries/s",
        "n_gpus": 1,
        "higher_is_better": True,
        "dtype": "f64",
        "data": "synthetic",
        "config": {
            "workload": f"{args.meshes} meshes, V log-uniform in [{args.min_vertices:.0f}, "
            f"{args.max_vertices:.0f}] (icosphere / torus / planar grid), Poisson arrivals at "
            f"{args.rate:.0f}/s for {args.duration:.0f}s, 1-32 sources per query, "
            f"{args.budget_gb:.0f} GB budget, {args.workers} workers"
            + (f", meshes of at most {args.pack_vertices} vertices in packed groups"
               if args.pack_vertices else ""),
            "total_vertices": int(sum(sizes.values())),
        },
        "offered_per_s": round(submitted / args.duration, 2),
        "completed": stats["completed"],
        "p50_ms": round(stats["p50_ms"], 3),
        "p99_ms": round(stats["p99_ms"], 3),
        "mean_ms": round(stats["mean_ms"], 3),
        "batches": stats["batches"],
        "evictions": stats["evictions"],
        "resident_meshes": stats["resident"],
        "packs": stats.get("packs", 0),
        "lanes": stats["lanes"],
        "resident_gb": round(stats["resident_bytes"] / 1e9, 3),
        "setup_s": round(setup, 2),
        "sampled_rel_err_max": max(errors) if errors else None,
    }
    if probes:
        result["capacity_probes"] = probes
    return result


def serve_workload(args) -> int:
    probes = [float(r) for r in args.serve_probes.split(",") if r.strip()]
    print(json.dumps(serve_result(args, probe_rates=probes)))
    return 0


def pattern_workload(args) -> int:
    """The operator pattern of the torus built on the device from device-resident faces
    (radix sorts of the 6F + V edge keys and of the 3F corners), `steps` builds timed after
    `warmup`; checked once against the oracle's sorted keys."""
    import torch

    from oracle import heat_oracle as ho
    from ordito_amd import creation
    from ordito_amd.laplacian import mesh_operator_pattern

    vertices, faces = creation.torus(args.major, args.minor, 1.0, 0.3, jitter=0.1, seed=0)
    n = vertices.shape[0]
    f_gpu = torch.from_numpy(faces.reshape(-1)).cuda()
    for _ in range(max(args.warmup, 1)):
        pattern = mesh_operator_pattern(f_gpu, n)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        pattern = mesh_operator_pattern(f_gpu, n)
    torch.cuda.synchronize()
    seconds = (time.perf_counter() - t0) / args.steps
    want = ho.operator_pattern(n, faces.reshape(-1, 3))
    exact = bool(
        np.array_equal(pattern.offsets.cpu().numpy(), want.indptr)
        and np.array_equal(pattern.columns.cpu().numpy(), want.indices)
        and np.array_equal(pattern.corner_ptr.cpu().numpy(), want.corner_ptr)
        and np.array_equal(pattern.corner_slots.cpu().numpy(), want.corner_slots)
    )
    keys = 6 * (faces.size // 3) + n
    print(json.dumps({
        "metric": "operator pattern built on the device (mesh_operator_pattern)",
        "value": round(1e3 * seconds, 4),
        "unit": "ms per build",
        "higher_is_better": False,
        "steps": args.steps,
        "config": {"workload": f"torus {args.major}x{args.minor}", "n_vertices": int(n),
                   "n_faces": int(faces.size // 3), "keys": int(keys),
                   "entries": int(pattern.columns.numel())},
        "keys_per_s": round(keys / seconds, 1),
        "matches_oracle": exact,
    }))
    return 0


def icosphere_rows() -> dict:
    """
    ordito's own ``heat_geodesic`` benchmark rows (``benchmarks/test_heat.py:136-230``, published in
    ``docs/benchmarks/heat.md:52-57``): unit icospheres of 4, 6 and 7 subdivisions (sphere_small,
    sphere_med, sphere_large), one source at vertex 0. ``full`` times
    ``heat_geodesic(vertices, faces, sources)`` (operators built and both solves run inside the
    call); ``amortized`` times ``heat_geodesic(mesh, sources)`` on a ``Trimesh`` warmed by two
    calls, so both systems are factored and each call is two refined solves. Medians over 20 calls
    (10 for the full icosphere(7)), each checked against the FP64 oracle on icosphere(4).
    """
    import torch

    from ordito_amd import creation, heat
    from oracle import heat_oracle as ho

    published = {4: (3.47, 0.3243), 6: (6.60, 0.6897), 7: (21.63, 1.48)}
    rows = {}
    src = torch.tensor([0], dtype=torch.int32, device="cuda")
    for level in (4, 6, 7):
        v, f = creation.icosphere(level)
        vg, fg = torch.from_numpy(v).cuda(), torch.from_numpy(f).cuda()

        def timed(call, count):
            lat = []
            for _ in range(count):
                t0 = time.perf_counter()
                got = call()
                torch.cuda.synchronize()
                lat.append(time.perf_counter() - t0)
            return got, lat

        for _ in range(2):
            heat.heat_geodesic(vg, fg, src)
        torch.cuda.synchronize()
        got_full, full = timed(lambda: heat.heat_geodesic(vg, fg, src), 10 if level == 7 else 20)
        mesh = heat.Trimesh(vg, fg)
        for _ in range(2):
            heat.heat_geodesic(mesh, src)
        torch.cuda.synchronize()
        got_amortized, amortized = timed(lambda: heat.heat_geodesic(mesh, src), 20)
        row = {
            "n_vertices": int(len(v)),
            "full_p50_ms": ms(full, 50),
            "amortized_p50_ms": ms(amortized, 50),
            "ordito_rtx5090_ms": {"full": published[level][0], "amortized": published[level][1]},
            "full_vs_amortized_max_abs": float(
                (got_full - got_amortized).abs().max().item()),
        }
        if level == 4:
            expected = ho.geodesic(v, f, np.array([0])).distance
            row["rel_err_max"] = ho.relative_error(got_amortized.cpu().numpy(), expected)[0]
        rows[f"icosphere{level}"] = row
        log(f"   icosphere({level}): {row}")
        del mesh
        torch.cuda.empty_cache()
    return rows


def one_shot_result(oracle_c2, c2_vertices, c2_faces, c2_source: int) -> dict:
    """
    The reference's one-shot form (``heat_geodesic(vertices, faces, sources)``: every call builds
    the operators and iterates both solves, nothing kept): the jittered icosphere(4) from vertex 0
    (BASELINE's first configuration; ordito's published 3.47 ms is an unjittered icosphere(4) on
    an RTX 5090, ``docs/benchmarks/heat.md:52``), and one source on the torus 1024 x 512.
    """
    import torch

    from ordito_amd import creation, heat
    from oracle import heat_oracle as ho

    out = {}
    v, f = creation.jittered_icosphere(4, 0.01, 0)
    vg, fg = torch.from_numpy(v).cuda(), torch.from_numpy(f).cuda()
    src = torch.tensor([0], dtype=torch.int32, device="cuda")
    expected = ho.geodesic(v, f, np.array([0])).distance
    for _ in range(2):
        heat.heat_geodesic(vg, fg, src)
    torch.cuda.synchronize()
    lat = []
    for _ in range(20):
        t0 = time.perf_counter()
        got = heat.heat_geodesic(vg, fg, src)
        torch.cuda.synchronize()
        lat.append(time.perf_counter() - t0)
    out["icosphere4_jittered"] = {
        "n_vertices": int(len(v)),
        "mean_ms": round(1e3 * float(np.mean(lat)), 3),
        "p50_ms": ms(lat, 50),
        "p99_ms": ms(lat, 99),
        "calls": len(lat),
        "rel_err_max": ho.relative_error(got.cpu().numpy(), expected)[0],
    }
    log(f"   one-shot icosphere(4): {out['icosphere4_jittered']}")
    out["icospheres"] = icosphere_rows()
    vg, fg = torch.from_numpy(c2_vertices).cuda(), torch.from_numpy(c2_faces).cuda()
    src = torch.tensor([c2_source], dtype=torch.int32, device="cuda")
    lat = []
    stats = {}
    cg_ms = []
    plan = {}
    for _ in range(2):
        t0 = time.perf_counter()
        trimesh = heat.Trimesh(vg, fg)
        solver = trimesh.heat_solver(method="meshlab")
        solver.enable_timing(True)
        got = heat.heat_geodesic(trimesh, src)
        torch.cuda.synchronize()
        lat.append(time.perf_counter() - t0)
        stats = dict(solver.last_stats)
        cg_ms.append(solver.read_timing(reset=True)["cg"][0])
        view = heat._views(solver._native())
        plan = {"n": int(view.n_vertices), "nnz": int(view.nnz)}
        del solver, trimesh
    # The settled Jacobi-CG's algorithmic bytes a round (one column): the FP32 operator values and
    # their columns once, the row offsets, the direction gathered once, and the Chronopoulos-Gear
    # update's seven float64 vector reads (u, p, w, s, x, r, the inverse diagonal) and six writes
    # (w, p, s, x, r, u).
    rounds = max(stats.get("heat_rounds", 0), 1)
    round_bytes = plan["nnz"] * 8 + plan["n"] * (4 + 14 * 8)
    round_s = float(np.mean(cg_ms)) / 1e3 / rounds
    out["torus1024x512"] = {
        "n_vertices": int(len(c2_vertices)),
        "mean_ms": round(1e3 * float(np.mean(lat)), 3),
        "calls": len(lat),
        "rounds": stats,
        "heat_cg_ms": round(float(np.mean(cg_ms)), 3),
        "cg_round_us": round(1e6 * round_s, 2),
        "cg_round_bytes": round_bytes,
        "cg_round_gbs": round(round_bytes / round_s / 1e9, 1),
        "rel_err_max": ho.relative_error(got.cpu().numpy(), oracle_c2.distance(np.array([c2_source])))[0],
    }
    log(f"   one-shot torus: {out['torus1024x512']}")
    return out


def extra_workloads(args, oracle_c2, c2_vertices, c2_faces, c2_sources) -> dict:
    """The other configurations' figures, each a sub-object of the default line."""
    import copy

    import torch

    extras = {}
    log("== one-shot (iterated) calls")
    extras["one_shot"] = one_shot_result(oracle_c2, c2_vertices, c2_faces, int(c2_sources[0]))
    torch.cuda.empty_cache()
    log("== displaced icosphere(9), FP64 and mixed")
    sub = copy.copy(args)
    sub.steps, sub.warmup = args.extra_steps, 2
    extras["displaced_sphere"] = sphere_result(sub, oracle_check=False)
    torch.cuda.empty_cache()
    log("== 256 packed meshes")
    sub = copy.copy(args)
    sub.steps, sub.warmup = args.extra_steps, 2
    sub.meshes, sub.min_vertices, sub.max_vertices = 256, 1e3, 2e5
    extras["packed_256"] = batch_result(sub)
    torch.cuda.empty_cache()
    log("== 64 served meshes")
    sub = copy.copy(args)
    sub.meshes, sub.min_vertices, sub.max_vertices = 64, 1e4, 1e6
    probes = [float(r) for r in args.serve_probes.split(",") if r.strip()]
    extras["served_64"] = serve_result(sub, probe_rates=probes)
    torch.cuda.empty_cache()
    for value in extras.values():
        for key in ("metric", "n_gpus", "higher_is_better", "data"):
            value.pop(key, None)
    return extras


# The served window of a profiled C2 run: every dispatch from the first to the last 64-column
# sweep tile, whose sweep kernels (tiles and split-K reductions) make the triangular sweeps.
_SWEEP_TILES = ("fwd_tile_kernel<double", "bwd_tile_kernel<double")
_SWEEP_KERNELS = _SWEEP_TILES + ("fwd_reduce_kernel", "bwd_reduce_kernel")


def _short_name(name: str) -> str:
    import re

    name = name.replace("hg::(anonymous namespace)::", "").replace("hg::", "").replace("void ", "")
    return re.sub(r"\(.*", "", name)


def _rows(directory: Path, suffix: str) -> list[dict]:
    import csv

    files = sorted(directory.rglob(f"*{suffix}"))
    return list(csv.DictReader(files[0].open())) if files else []


def _dispatch_counters(directory: Path) -> dict[int, dict]:
    out: dict[int, dict] = {}
    for r in _rows(directory, "counter_collection.csv"):
        d = out.setdefault(int(r["Dispatch_Id"]), {"name": _short_name(r["Kernel_Name"])})
        d[r["Counter_Name"]] = d.get(r["Counter_Name"], 0.0) + float(r["Counter_Value"])
    return out


def _copy_scale(directory: Path, counter: str) -> float:
    """Bytes a 1 GiB device copy moves per KiB its counter reads (reads: 1 GiB, writes: 1 GiB)."""
    copies = [d for d in _dispatch_counters(directory).values() if "copy" in d["name"].lower()]
    if not copies:
        return float("nan")
    kib = max(d.get(counter, 0.0) for d in copies)
    return (1 << 30) / (kib * 1024.0) if kib > 0 else float("nan")


def _window(names: list[str]) -> tuple[int, int]:
    idx = [i for i, n in enumerate(names) if n.startswith(_SWEEP_TILES)]
    return (idx[0], idx[-1] + 1) if idx else (0, 0)


def summarize_profiles(root: Path, traffic_out: Path | None = None,
                       solve_block_entries: int = 0) -> str:
    """Markdown summary of the profiling runs listed in profiles/triangular_sweep_torus1024x512.md:
    per-level sweep times, per-kernel times and HBM traffic (FETCH_SIZE / WRITE_SIZE scaled by a
    1 GiB copy's readings), the factorization's time and matrix-core busy fraction, the one-shot
    CG."""
    import collections

    lines: list[str] = []
    trace = sorted(_rows(root / "kt", "kernel_trace.csv"), key=lambda r: int(r["Start_Timestamp"]))
    names = [_short_name(r["Kernel_Name"]) for r in trace]
    dur = [(int(r["End_Timestamp"]) - int(r["Start_Timestamp"])) / 1e3 for r in trace]
    lo, hi = _window(names)
    sweeps = sum(1 for n in names[lo:hi] if n.startswith("fwd_tile_kernel<double, false>"))
    sweep_us = sum(d for n, d in zip(names[lo:hi], dur[lo:hi]) if n.startswith(_SWEEP_KERNELS))
    lines.append(f"- sweeps in the served window: {sweeps}")
    lines.append(f"- sweep kernels' time per sweep (rocprofv3): **{sweep_us / max(sweeps, 1) / 1e3:.4f} ms**")
    # The last sweep level by level.
    starts = [i for i in range(lo, hi) if names[i].startswith("fwd_tile_kernel<double, false>")]
    if starts:
        lines += ["", "One sweep, launch by launch (the last of the window):", "",
                  "| launch | grid | µs |", "|---|---:|---:|"]
        i = starts[-1]
        while i < hi and names[i].startswith(_SWEEP_KERNELS):
            grid = f"{trace[i].get('Grid_Size_X', trace[i].get('Grid_Size', '?'))}"
            lines.append(f"| `{names[i]}` | {grid} | {dur[i]:.2f} |")
            i += 1
            if i < hi and names[i].startswith("fwd_tile_kernel<double, false>"):
                break
    agg = collections.defaultdict(lambda: [0, 0.0])
    for n, d in zip(names[lo:hi], dur[lo:hi]):
        agg[n][0] += 1
        agg[n][1] += d
    lines += ["", "Kernels in the served window:", "", "| kernel | calls | total ms | avg µs |",
              "|---|---:|---:|---:|"]
    for n, (c, t) in sorted(agg.items(), key=lambda kv: -kv[1][1]):
        lines.append(f"| `{n}` | {c} | {t / 1e3:.3f} | {t / c:.2f} |")
    whole = collections.defaultdict(lambda: [0, 0.0])
    for n, d in zip(names, dur):
        whole[n][0] += 1
        whole[n][1] += d
    avg_us = {n: t / c for n, (c, t) in whole.items()}
    # HBM traffic.
    fs = _copy_scale(root / "copy_fetch", "FETCH_SIZE")
    ws = _copy_scale(root / "copy_write", "WRITE_SIZE")
    fetch = _dispatch_counters(root / "fetch")
    write = _dispatch_counters(root / "write")
    lines += ["", f"Counter scales from a 1 GiB device copy: FETCH_SIZE x {fs:.4f}, WRITE_SIZE x {ws:.4f}"
              " (bytes per KiB read)."]
    per_sweep: dict[str, float] = {}
    for label, counters, key, scale in (("FETCH_SIZE", fetch, "FETCH_SIZE", fs),
                                        ("WRITE_SIZE", write, "WRITE_SIZE", ws)):
        ids = sorted(counters)
        nm = [counters[d]["name"] for d in ids]
        a, b = _window(nm)
        n_sw = sum(1 for x in nm[a:b] if x.startswith("fwd_tile_kernel<double, false>"))
        total = sum(counters[ids[j]].get(key, 0.0) for j in range(a, b) if nm[j].startswith(_SWEEP_KERNELS))
        per_sweep[label] = total * 1024 * scale / max(n_sw, 1)
        lines.append(f"- {label}: {per_sweep[label] / 1e9:.3f} GB per sweep "
                     f"({n_sw} sweeps in that run's window)")
    lines += ["", "Per-kernel HBM bytes per dispatch (counters x scale) and the rate over the kernel "
              "trace's mean duration:", "", "| kernel | dispatches | read MB | written MB | µs | GB/s |",
              "|---|---:|---:|---:|---:|---:|"]
    per = collections.defaultdict(lambda: [0, 0.0, 0.0])
    for d in fetch.values():
        per[d["name"]][0] += 1
        per[d["name"]][1] += d.get("FETCH_SIZE", 0.0) * 1024 * fs
    for d in write.values():
        per[d["name"]][2] += d.get("WRITE_SIZE", 0.0) * 1024 * ws
    ranked = sorted(per.items(), key=lambda kv: -(kv[1][1] + kv[1][2]))
    assembly = ("face_terms_kernel", "edge_sums_kernel", "mass_kernel", "operator_rows_kernel",
                "face_geometry_kernel")
    shown = ranked[:24] + [kv for kv in ranked[24:] if kv[0] in assembly]
    for n, (c, rd, wr) in shown:
        us = avg_us.get(n)
        rate = (rd + wr) / c / (us * 1e-6) / 1e9 if us else float("nan")
        lines.append(f"| `{n}` | {c} | {rd / c / 1e6:.2f} | {wr / c / 1e6:.2f} | "
                     f"{us if us else float('nan'):.1f} | {rate:.0f} |")
    if traffic_out is not None and per_sweep:
        record = {
            "kernel": "triangular sweep (forward + backward)",
            "workload": "torus 1024x512 R=1 r=0.3 normal jitter 0.1h, 64 single-source queries per "
            "step as one batch, factors resident",
            "columns": 64,
            "solve_block_entries": solve_block_entries,
            "factor_bytes": 8,
            "traffic_bytes_per_sweep": round(per_sweep["FETCH_SIZE"] + per_sweep["WRITE_SIZE"]),
            "fetch_bytes": round(per_sweep["FETCH_SIZE"]),
            "write_bytes": round(per_sweep["WRITE_SIZE"]),
            "fetch_scale": round(fs, 4),
            "write_scale": round(ws, 4),
            "solve_hip_sha256_16": _solve_source_hash(),
        }
        traffic_out.write_text(json.dumps(record, indent=1) + "\n")
    # Factorization: time and matrix-core busy.
    mfma = _dispatch_counters(root / "mfma")
    fact = ("panel_diag_kernel", "panel_apply_kernel", "update_kernel", "extend_add_kernel")
    lines += ["", "Numeric factorization kernels (whole run: both systems' factorizations), time and "
              "matrix-core busy (SQ_VALU_MFMA_BUSY_CYCLES over GRBM_GUI_ACTIVE x 1024 SIMDs / 8 XCDs):",
              "", "| kernel | calls | total ms | MFMA busy | FP64 MFMA GFLOP |", "|---|---:|---:|---:|---:|"]
    for n in fact:
        c, t = whole.get(n, [0, 0.0])
        busy = [d for d in mfma.values() if d["name"] == n]
        cycles = sum(d.get("GRBM_GUI_ACTIVE", 0.0) for d in busy) / 8.0
        mfma_busy = sum(d.get("SQ_VALU_MFMA_BUSY_CYCLES", 0.0) for d in busy)
        flops = sum(d.get("SQ_INSTS_VALU_MFMA_MOPS_F64", 0.0) for d in busy) * 512
        frac = mfma_busy / (cycles * 1024) if cycles else float("nan")
        lines.append(f"| `{n}` | {c} | {t / 1e3:.3f} | {100 * frac:.1f}% | {flops / 1e9:.2f} |")
    return "\n".join(lines)


def log(message: str) -> None:
    print(message, file=sys.stderr, flush=True)


def calibrate_copy() -> int:
    import torch

    x = torch.empty(256 * 1024 * 1024, dtype=torch.float32, device="cuda")  # 1 GiB
    x.fill_(1.0)
    y = torch.empty_like(x)
    for _ in range(3):
        y.copy_(x)
    torch.cuda.synchronize()
    log("copy calibration: 3 device copies of 1 GiB")
    return 0


TRAFFIC_FILE = ROOT / "profiles" / "triangular_sweep_torus1024x512_traffic.json"


def _solve_source_hash() -> str:
    import hashlib

    return hashlib.sha256((ROOT / "ordito_amd" / "csrc" / "solve.hip").read_bytes()).hexdigest()[:16]


def measured_traffic(columns: int, solve_block_entries: int, factor_bytes: int) -> float | None:
    """HBM bytes per sweep from the committed FETCH_SIZE / WRITE_SIZE passes, when they were taken
    on this workload (its column count, stored solve-block entries and factor width) with the sweep
    kernels' current source; otherwise None."""
    try:
        record = json.loads(TRAFFIC_FILE.read_text())
    except (OSError, ValueError):
        return None
    if (
        record.get("solve_hip_sha256_16") != _solve_source_hash()
        or record.get("columns") != columns
        or record.get("solve_block_entries") != solve_block_entries
        or record.get("factor_bytes") != factor_bytes
    ):
        return None
    return float(record["traffic_bytes_per_sweep"])


def sweep_bytes(stats: dict[str, int], columns: int, factor_bytes: int, first_round: bool) -> float:
    """Algorithmic bytes of one forward + backward sweep over `columns` right-hand sides: the
    inverse factors' structural nonzeros (each front's lower-triangular leading block and the rows
    below it; not the zero triangle the stored blocks carry) read twice, and per column these
    float64 vector touches -- forward: r read,
    y write (n each), update vectors written and gathered by the parents (structure rows each);
    backward: y read, d write (n each), d gathered at the structure rows, and after the first
    round x read and x + d write (n each; the first round starts from x = 0 and its d is x)."""
    n = stats["rows"]
    structure_rows = stats["front_vector_entries"] - n
    w = stats["solve_block_nonzeros"]
    vector_n = 4.0 if first_round else 6.0
    return 2.0 * w * factor_bytes + columns * (vector_n * n + 3.0 * structure_rows) * 8.0


def timed_steps(solver, source_sets, steps: int, warmup: int):
    """Warm up, then time `steps` batches with HIP-event phase timing on; returns (seconds,
    per-step latencies, phase timing, last output, last stats). The batch's sources are one flat
    device tensor with host offsets, assembled once (a server's batch arrives that way)."""
    import torch

    flat = torch.cat(source_sets).contiguous()
    offsets = [0]
    for s in source_sets:
        offsets.append(offsets[-1] + s.numel())
    for _ in range(warmup):
        solver.geodesic_flat(flat, offsets)
    solver.enable_timing(True)
    solver.read_timing(reset=True)
    torch.cuda.synchronize()
    latencies = []
    out = None
    t_start = time.perf_counter()
    for _ in range(steps):
        s0 = time.perf_counter()
        out = solver.geodesic_flat(flat, offsets)
        torch.cuda.synchronize()
        latencies.append(time.perf_counter() - s0)
    elapsed = time.perf_counter() - t_start
    timing = solver.read_timing(reset=True)
    solver.enable_timing(False)
    return elapsed, latencies, timing, out, dict(solver.last_stats)


def in_flight_steps(solvers, source_sets, steps: int, warmup: int):
    """The served batch with len(solvers) batches in flight: each solver (an independent handle
    of the mesh, its own factors) runs on its own host thread and stream, the `steps` steps split
    between them. Returns (seconds, per-step latencies, each solver's last output)."""
    import threading

    import torch

    flat = torch.cat(source_sets).contiguous()
    offsets = [0]
    for s in source_sets:
        offsets.append(offsets[-1] + s.numel())
    lanes = len(solvers)
    streams = [torch.cuda.Stream() for _ in solvers]
    # Eager launches on each lane's own stream, as the server's workers run: a replayed graph runs
    # on a stream the library keeps per host thread, and which hardware queue that lands on is not
    # the caller's to choose (two lanes on one queue run one after the other).
    for solver in solvers:
        solver.set_graphs(False)
    torch.cuda.synchronize()
    counts = [steps // lanes + (1 if i < steps % lanes else 0) for i in range(lanes)]
    latencies = [[] for _ in solvers]
    outs = [None] * lanes
    errors = []
    start = threading.Barrier(lanes + 1, timeout=600)

    def run(i: int) -> None:
        try:
            with torch.cuda.stream(streams[i]):
                # Each lane warms up on its own thread and stream, as a server's worker would, the
                # lanes together (at least five steps: the first concurrent steps of a process have
                # run one lane after the other now and then).
                for _ in range(max(warmup, 5)):
                    solvers[i].geodesic_flat(flat, offsets)
                streams[i].synchronize()
                start.wait()
                for _ in range(counts[i]):
                    s0 = time.perf_counter()
                    outs[i] = solvers[i].geodesic_flat(flat, offsets)
                    streams[i].synchronize()
                    latencies[i].append(time.perf_counter() - s0)
        except BaseException as error:  # reported by the caller
            errors.append(error)
            start.abort()

    threads = [threading.Thread(target=run, args=(i,)) for i in range(lanes)]
    for thread in threads:
        thread.start()
    torch.cuda.synchronize()
    start.wait()
    t0 = time.perf_counter()
    for thread in threads:
        thread.join()
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    for solver in solvers:
        solver.set_graphs(True)
    if errors:
        raise errors[0]
    return elapsed, [x for lane in latencies for x in lane], outs


def ms(values, q) -> float:
    return round(1e3 * float(np.percentile(values, q)), 3)


def main() -> int:
    args = parse()
    if args.summarize_profiles:
        out = Path(args.traffic_out) if args.traffic_out else None
        print(summarize_profiles(Path(args.summarize_profiles), out, args.solve_block_entries))
        return 0
    if args.gpus != 1:
        print(
            "bench.py measures one GPU: the heat-geodesic path here is built and measured on a "
            "single MI355X (multi-GPU sharding is a design note, not a measured configuration)",
            file=sys.stderr,
        )
        return 2
    import torch

    from ordito_amd import creation, heat

    if not torch.cuda.is_available():
        print("bench.py needs a GPU", file=sys.stderr)
        return 2
    torch.cuda.init()
    if args.calibrate_copy:
        return calibrate_copy()
    if args.workload == "serve":
        return serve_workload(args)
    if args.workload == "batch":
        return batch_workload(args)
    if args.workload == "sphere":
        return sphere_workload(args)
    if args.workload == "igl":
        return igl_workload(args)
    if args.workload == "pattern":
        return pattern_workload(args)
    if args.workload == "oneshot":
        log("== one-shot (iterated) calls")
        vertices, faces = creation.torus(args.major, args.minor, 1.0, 0.3, jitter=0.1, seed=0)
        from oracle import heat_oracle as ho

        source = int(creation.random_sources(vertices.shape[0], args.queries, seed=1)[0])
        line = {
            "metric": "one-shot heat_geodesic calls (operators built and both solves iterated "
            "each call)",
            "n_gpus": 1,
            "data": "synthetic",
            "workloads": one_shot_result(ho.FactoredOracle(vertices, faces), vertices, faces, source),
        }
        print(json.dumps(line))
        return 0
    log(f"== generating torus {args.major}x{args.minor}")
    vertices, faces = creation.torus(args.major, args.minor, 1.0, 0.3, jitter=0.1, seed=0)
    n = vertices.shape[0]
    sources = creation.random_sources(n, args.queries, seed=1)
    v_gpu = torch.from_numpy(vertices).cuda()
    f_gpu = torch.from_numpy(faces).cuda()
    source_sets = [torch.tensor([s], dtype=torch.int32, device="cuda") for s in sources]
    k = len(source_sets)

    log("== factoring (setup)")
    mesh = heat.Trimesh(v_gpu, f_gpu)
    solver = mesh.heat_solver(method="meshlab", precision="fp64")
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    assert solver.factor()
    torch.cuda.synchronize()
    setup_fp64 = time.perf_counter() - t0
    stats = solver.plan_stats()
    log(f"   factor fp64 {setup_fp64:.3f}s plan {stats}")
    # Factored: every later solve goes through the held factors (the rent-or-buy count is past).
    solver._solves = {"heat": 2, "poisson": 2}

    log(f"== served, one batch at a time: {args.steps} steps of {k} queries, FP64 inverse factors")
    elapsed1, latencies1, timing, out, served_stats = timed_steps(solver, source_sets, args.steps, args.warmup)
    value1 = k * args.steps / elapsed1
    log(f"   {value1:.1f} queries/s, step {1e3 * elapsed1 / args.steps:.2f} ms, stats {served_stats}")
    served_host = out.cpu().numpy()
    lanes = max(1, args.lanes)
    solvers = [solver]
    lane_meshes = []
    t0 = time.perf_counter()
    for _ in range(lanes - 1):
        lane_mesh = heat.Trimesh(v_gpu, f_gpu)
        lane = lane_mesh.heat_solver(method="meshlab", precision="fp64")
        assert lane.factor()
        lane._solves = {"heat": 2, "poisson": 2}
        lane_meshes.append(lane_mesh)
        solvers.append(lane)
    torch.cuda.synchronize()
    setup_lanes = time.perf_counter() - t0
    log(f"== served, {lanes} batches in flight: {args.steps} steps of {k} queries")
    elapsed, latencies, lane_outs = in_flight_steps(solvers, source_sets, args.steps, args.warmup)
    value = k * args.steps / elapsed
    lanes_equal = all(o is None or np.array_equal(o.cpu().numpy(), served_host) for o in lane_outs)
    log(f"   served {value:.1f} queries/s, {1e3 * elapsed / args.steps:.2f} ms a step, lanes equal {lanes_equal}")
    del solvers, lane_outs

    log(f"== plain FP64, one query per call, {args.plain_queries} queries")
    solver.geodesic_flat(source_sets[0], [0, 1])
    torch.cuda.synchronize()
    plain_lat = []
    plain_out = []
    t0 = time.perf_counter()
    for q in range(args.plain_queries):
        s0 = time.perf_counter()
        plain_out.append(solver.geodesic_flat(source_sets[q % k], [0, 1])[0])
        torch.cuda.synchronize()
        plain_lat.append(time.perf_counter() - s0)
    plain_value = args.plain_queries / (time.perf_counter() - t0)
    plain_stats = dict(solver.last_stats)

    # Roofline of the dominant kernel: the served path's triangular sweeps (heat and Poisson).
    sweep_ms = timing["heat_sweep"][0] + timing["poisson_sweep"][0]
    sweep_count = timing["heat_sweep"][1] + timing["poisson_sweep"][1]
    mean_sweep_s = sweep_ms / max(sweep_count, 1) / 1e3
    # Each step's heat and Poisson solves start with a first-round sweep per 64-column chunk.
    first_rounds = min(sweep_count, 2 * args.steps * -(-k // 64))
    bytes_per_sweep = (
        first_rounds * sweep_bytes(stats, k, 8, True)
        + (sweep_count - first_rounds) * sweep_bytes(stats, k, 8, False)
    ) / max(sweep_count, 1)
    achieved = bytes_per_sweep / mean_sweep_s / 1e9 if mean_sweep_s > 0 else 0.0
    roofline = {
        "bound": "hbm",
        "kernel": "triangular sweep (forward + backward level sweeps, FP64 inverse factors, 64 columns)",
        "achieved": round(achieved, 1),
        "peak": HBM_PEAK_GBS,
        "unit": "GB/s",
        "frac": round(achieved / HBM_PEAK_GBS, 4),
        "traffic": measured_traffic(k, stats["solve_block_entries"], 8),
        "bytes_per_launch": bytes_per_sweep,
        "mean_launch_ms": round(mean_sweep_s * 1e3, 4),
        "launches": sweep_count,
    }

    result = {
        "metric": METRIC,
        "value": round(value, 2),
        "unit": "queries/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(1e3 * elapsed / args.steps, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {
            "workload": f"torus {args.major}x{args.minor} R=1 r=0.3 normal jitter 0.1h, "
            f"{k} single-source queries per step as one batch, factors resident, "
            f"{lanes} batches in flight",
            "in_flight": lanes,
            "n_vertices": int(n),
            "n_faces": int(faces.size // 3),
            "queries_per_step": k,
            "t": solver.time,
            "supernodes": stats["supernodes"],
            "levels": stats["levels"],
            "solve_block_entries": stats["solve_block_entries"],
            "solve_block_nonzeros": stats["solve_block_nonzeros"],
        },
        "p50_ms": ms(latencies, 50),
        "p99_ms": ms(latencies, 99),
        "latency_note": "a query's latency is its step's: all 64 complete together; with "
        "batches in flight a step's latency is measured on its own lane",
        "single_lane": {
            "value": round(value1, 2),
            "ms_per_step": round(1e3 * elapsed1 / args.steps, 3),
            "p50_ms": ms(latencies1, 50),
            "p99_ms": ms(latencies1, 99),
        },
        "lanes_bit_equal": lanes_equal,
        "plain_fp64": {
            "value": round(plain_value, 2),
            "unit": "queries/s",
            "p50_ms": ms(plain_lat, 50),
            "p99_ms": ms(plain_lat, 99),
            "queries": args.plain_queries,
            "refinements": plain_stats,
        },
        "served_vs_plain": round(value / plain_value, 2) if plain_value > 0 else None,
        "setup_s": {"factor_fp64": round(setup_fp64, 3), "further_lanes": round(setup_lanes, 3)},
        "phase_ms_per_step": {key: round(v[0] / args.steps, 3) for key, v in timing.items() if v[1]},
        "refinements": served_stats,
        "roofline": roofline,
    }

    if not args.no_mixed:
        log("== served_mixed: the same batches through FP32 inverse factors")
        mesh_mixed = heat.Trimesh(v_gpu, f_gpu)
        mixed = mesh_mixed.heat_solver(method="meshlab", precision="mixed")
        t0 = time.perf_counter()
        assert mixed.factor()
        torch.cuda.synchronize()
        result["setup_s"]["factor_mixed"] = round(time.perf_counter() - t0, 3)
        mixed._solves = {"heat": 2, "poisson": 2}
        m_elapsed, m_lat, m_timing, m_out, m_stats = timed_steps(mixed, source_sets, args.steps, 1)
        mixed_host = m_out.cpu().numpy()
        from_served = [
            float(np.max(np.abs(mixed_host[q] - served_host[q])) / np.ptp(served_host[q]))
            for q in range(k)
        ]
        # The same batches with the FP64 line's batches in flight (lanes of the mixed factors).
        mixed_lanes = [mixed]
        mixed_meshes = []
        for _ in range(lanes - 1):
            lane_mesh = heat.Trimesh(v_gpu, f_gpu)
            lane = lane_mesh.heat_solver(method="meshlab", precision="mixed")
            assert lane.factor()
            lane._solves = {"heat": 2, "poisson": 2}
            mixed_meshes.append(lane_mesh)
            mixed_lanes.append(lane)
        mf_elapsed, mf_lat, mf_outs = in_flight_steps(mixed_lanes, source_sets, args.steps, 1)
        mixed_equal = all(o is None or np.array_equal(o.cpu().numpy(), mixed_host) for o in mf_outs)
        log(f"   mixed: {k * args.steps / mf_elapsed:.1f} queries/s with {lanes} in flight, "
            f"{k * args.steps / m_elapsed:.1f} one at a time")
        result["served_mixed"] = {
            "value": round(k * args.steps / mf_elapsed, 2),
            "unit": "queries/s",
            "in_flight": lanes,
            "ms_per_step": round(1e3 * mf_elapsed / args.steps, 3),
            "p50_ms": ms(mf_lat, 50),
            "p99_ms": ms(mf_lat, 99),
            "single_lane": {
                "value": round(k * args.steps / m_elapsed, 2),
                "ms_per_step": round(1e3 * m_elapsed / args.steps, 3),
                "p50_ms": ms(m_lat, 50),
                "p99_ms": ms(m_lat, 99),
            },
            "lanes_bit_equal": mixed_equal,
            "vs_plain_fp64": round(k * args.steps / mf_elapsed / plain_value, 2)
            if plain_value > 0 else None,
            "refinements": m_stats,
            "max_rel_diff_from_served": max(from_served),
            "phase_ms_per_step": {key: round(v[0] / args.steps, 3) for key, v in m_timing.items() if v[1]},
        }
        del mixed, mesh_mixed, mixed_lanes, mixed_meshes, mf_outs

    oracle = None
    oracle_setup = 0.0
    if not args.no_cpu_baseline or not args.no_extra:
        from oracle import heat_oracle as ho

        t0 = time.perf_counter()
        oracle = ho.FactoredOracle(vertices, faces)
        oracle_setup = time.perf_counter() - t0
    if not args.no_cpu_baseline:
        log("== cpu baseline (oracle, SciPy SuperLU, 1 thread)")
        sample = list(range(min(args.cpu_sample, k)))
        t0 = time.perf_counter()
        expected = [oracle.distance(np.array([sources[q]])) for q in sample]
        cpu_elapsed = time.perf_counter() - t0
        errs = [ho.relative_error(served_host[q], expected[i]) for i, q in enumerate(sample)]
        plain_errs = [
            ho.relative_error(plain_out[q].cpu().numpy(), expected[i])
            for i, q in enumerate(sample)
            if q < len(plain_out)
        ]
        all_cores = cpu_all_cores(oracle, sources, args.cpu_sample_all)
        result["cpu_baseline"] = {
            "value": round(len(sample) / cpu_elapsed, 4),
            "unit": "queries/s",
            "cores": 1,
            "kind": "port",
            "sample": f"{len(sample)} single-source queries on the same torus through the FP64 "
            f"oracle with both systems pre-factored (factorization {oracle_setup:.1f}s excluded)",
            "cpu": _cpu_model(),
            "nproc": os.cpu_count(),
            "all_cores": all_cores,
        }
        result["rel_err"] = {
            "served_max": max(e[0] for e in errs),
            "served_mean": float(np.mean([e[1] for e in errs])),
            "plain_max": max(e[0] for e in plain_errs) if plain_errs else None,
            "queries_checked": len(sample),
        }
    if not args.no_extra:
        del solver, mesh, lane_meshes
        torch.cuda.empty_cache()
        result["workloads"] = extra_workloads(args, oracle, vertices, faces, sources)
    print(json.dumps(result))
    return 0


def _oracle_query(q: int) -> float:
    # Runs in a forked worker: the factored oracle is inherited, not pickled.
    _FORK_STATE["oracle"].distance(np.array([_FORK_STATE["sources"][q]]))
    return 0.0


_FORK_STATE: dict = {}


def cpu_all_cores(oracle, sources, count: int) -> dict | None:
    """The same oracle queries on every host core: worker processes forked after the
    factorization (each solve single-threaded), ``count`` queries spread over them."""
    import multiprocessing

    cores = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else os.cpu_count()
    cores = max(1, min(int(cores or 1), 16))
    if count <= 0 or cores <= 1:
        return None
    _FORK_STATE["oracle"] = oracle
    _FORK_STATE["sources"] = sources
    context = multiprocessing.get_context("fork")
    queries = [q % len(sources) for q in range(count)]
    with context.Pool(cores) as pool:
        pool.map(_oracle_query, queries[:cores])  # workers up
        t0 = time.perf_counter()
        pool.map(_oracle_query, queries, chunksize=1)
        elapsed = time.perf_counter() - t0
    _FORK_STATE.clear()
    return {
        "value": round(count / elapsed, 4),
        "unit": "queries/s",
        "cores": cores,
        "sample": f"{count} single-source queries over {cores} forked worker processes sharing the "
        "pre-factored oracle (SciPy SuperLU, one thread each)",
    }


def _cpu_model() -> str:
    try:
        for line in Path("/proc/cpuinfo").read_text().splitlines():
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


if __name__ == "__main__":
    sys.exit(main())
