"""
Benchmark of the heat-method geodesic path on one MI355X.

Workload (the single-GPU configuration the headline metric is quoted on): a torus of 1024 x 512
quads (524,288 vertices, 1,048,576 faces), R = 1, r = 0.3, normal jitter N(0, 0.1 h), FP32
positions, 64 source vertices drawn from a seed; the diffusion time is the solver's default rule
(h^2 raised to (bbox diagonal / 400)^2). Both systems are factored once (setup, reported apart),
then a *step* is one batch of the 64 single-source queries through the served path: seed, heat
solve, unit-gradient divergence, Poisson solve, shift -- all on the GPU, factors resident in HBM.

- ``value``: served (mixed-precision) queries/s over the timed steps.
- ``plain_fp64``: the plain FP64 GPU path (FP64 factors, one query per call), timed beside it.
- ``roofline``: the triangular sweep (forward + backward, the dominant kernel), algorithmic bytes
  per sweep over its HIP-event-measured mean duration against the 8 TB/s HBM peak; ``traffic`` is
  the HBM bytes per sweep that rocprofv3's FETCH_SIZE / WRITE_SIZE passes measured
  (profiles/triangular_sweep_torus1024x512_traffic.json, scaled by a 1 GiB copy:
  ``bench.py --calibrate-copy``), null when the sweep kernels' source has changed since.
- ``cpu_baseline``: the FP64 oracle (SciPy SuperLU, one thread) on a bounded sample of the same
  queries on this host, factorization excluded as for the GPU; its results also give the error
  of the GPU fields.

    python bench.py [--gpus 1] [--steps K] [--warmup W]
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
    parser.add_argument("--steps", type=int, default=5)
    parser.add_argument("--warmup", type=int, default=1)
    parser.add_argument("--queries", type=int, default=64)
    parser.add_argument("--plain-queries", type=int, default=8)
    parser.add_argument("--cpu-sample", type=int, default=4)
    parser.add_argument("--no-cpu-baseline", action="store_true")
    parser.add_argument("--major", type=int, default=1024)
    parser.add_argument("--minor", type=int, default=512)
    parser.add_argument(
        "--calibrate-copy",
        action="store_true",
        help="only run three 1 GiB device copies (the reference kernel for scaling the FETCH_SIZE "
        "and WRITE_SIZE counters) and exit",
    )
    return parser.parse_args()


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


def measured_traffic(columns: int, solve_block_entries: int) -> float | None:
    """HBM bytes per sweep from the committed FETCH_SIZE / WRITE_SIZE passes, when they were taken
    on this workload with the sweep kernels' current source; otherwise None."""
    import hashlib

    try:
        record = json.loads(TRAFFIC_FILE.read_text())
    except (OSError, ValueError):
        return None
    source = (ROOT / "ordito_amd" / "csrc" / "solve.hip").read_bytes()
    if (
        record.get("solve_hip_sha256_16") != hashlib.sha256(source).hexdigest()[:16]
        or record.get("columns") != columns
        or record.get("solve_block_entries") != solve_block_entries
    ):
        return None
    return float(record["traffic_bytes_per_sweep"])


def log(message: str) -> None:
    print(message, file=sys.stderr, flush=True)


def sweep_bytes(stats: dict[str, int], columns: int, factor_bytes: int) -> float:
    """Algorithmic bytes of one forward + backward sweep over `columns` right-hand sides:
    the solve blocks read twice, and per column 9 n + 3 (sum of structure rows) float64 vector
    touches (permute in/out, b read, y write/read, update vectors write/read, x gather, x write,
    scatter-add read/write)."""
    n = stats["rows"]
    structure_rows = stats["front_vector_entries"] - n
    w = stats["solve_block_entries"]
    return 2.0 * w * factor_bytes + columns * (9.0 * n + 3.0 * structure_rows) * 8.0


def main() -> int:
    args = parse()
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
    log(f"== generating torus {args.major}x{args.minor}")
    vertices, faces = creation.torus(args.major, args.minor, 1.0, 0.3, jitter=0.1, seed=0)
    n = vertices.shape[0]
    sources = creation.random_sources(n, args.queries, seed=1)
    v_gpu = torch.from_numpy(vertices).cuda()
    f_gpu = torch.from_numpy(faces).cuda()
    source_sets = [torch.tensor([s], dtype=torch.int32, device="cuda") for s in sources]

    mesh_served = heat.Trimesh(v_gpu, f_gpu)
    served = mesh_served.heat_solver(method="meshlab", precision="mixed")
    mesh_plain = heat.Trimesh(v_gpu, f_gpu)
    plain = mesh_plain.heat_solver(method="meshlab", precision="fp64")

    log("== factoring (setup)")
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    assert served.factor()
    torch.cuda.synchronize()
    setup_served = time.perf_counter() - t0
    t0 = time.perf_counter()
    assert plain.factor()
    torch.cuda.synchronize()
    setup_plain = time.perf_counter() - t0
    stats = served.plan_stats()
    log(f"   factor mixed {setup_served:.2f}s fp64 {setup_plain:.2f}s plan {stats}")
    # The first solve after factoring warms the solve-side allocations; keep it off the clock.
    served._solves = {"heat": 2, "poisson": 2}
    plain._solves = {"heat": 2, "poisson": 2}

    log(f"== served warmup {args.warmup}")
    for _ in range(args.warmup):
        served.geodesic(source_sets)
    served.enable_timing(True)
    served.read_timing(reset=True)
    torch.cuda.synchronize()
    log(f"== served timed {args.steps} steps of {len(source_sets)} queries")
    latencies = []
    out = None
    t_start = time.perf_counter()
    for _ in range(args.steps):
        s0 = time.perf_counter()
        out = served.geodesic(source_sets)
        torch.cuda.synchronize()
        latencies.append(time.perf_counter() - s0)
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t_start
    timing = served.read_timing(reset=True)
    served.enable_timing(False)
    last_stats = dict(served.last_stats)
    value = len(source_sets) * args.steps / elapsed
    log(f"   served {value:.1f} queries/s, step {1e3 * elapsed / args.steps:.2f} ms, stats {last_stats}")

    log(f"== plain FP64, one query per call, {args.plain_queries} queries")
    plain.geodesic([source_sets[0]])
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    plain_lat = []
    plain_out = []
    for q in range(args.plain_queries):
        s0 = time.perf_counter()
        plain_out.append(plain.geodesic([source_sets[q % len(source_sets)]])[0])
        torch.cuda.synchronize()
        plain_lat.append(time.perf_counter() - s0)
    plain_elapsed = time.perf_counter() - t0
    plain_value = args.plain_queries / plain_elapsed

    # Roofline of the dominant kernel: the heat and Poisson triangular sweeps.
    sweep_ms = timing["heat_sweep"][0] + timing["poisson_sweep"][0]
    sweep_count = timing["heat_sweep"][1] + timing["poisson_sweep"][1]
    mean_sweep_s = sweep_ms / max(sweep_count, 1) / 1e3
    bytes_per_sweep = sweep_bytes(stats, len(source_sets), 4)
    achieved = bytes_per_sweep / mean_sweep_s / 1e9 if mean_sweep_s > 0 else 0.0
    traffic = measured_traffic(len(source_sets), stats["solve_block_entries"])
    roofline = {
        "bound": "hbm",
        "kernel": "triangular sweep (forward + backward level sweeps, mixed FP32 solve blocks)",
        "achieved": round(achieved, 1),
        "peak": HBM_PEAK_GBS,
        "unit": "GB/s",
        "frac": round(achieved / HBM_PEAK_GBS, 4),
        "traffic": traffic,
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
        "dtype": "f32 solve blocks, f64 vectors/refinement",
        "data": "synthetic",
        "config": {
            "workload": f"torus {args.major}x{args.minor} R=1 r=0.3 normal jitter 0.1h, "
            f"{len(source_sets)} single-source queries per step, factors resident",
            "n_vertices": int(n),
            "n_faces": int(faces.size // 3),
            "queries_per_step": len(source_sets),
            "t": served.time,
            "supernodes": stats["supernodes"],
            "levels": stats["levels"],
            "solve_block_entries": stats["solve_block_entries"],
        },
        "p50_ms": round(1e3 * float(np.percentile(latencies, 50)), 3),
        "p99_ms": round(1e3 * float(np.percentile(latencies, 99)), 3),
        "plain_fp64": {
            "value": round(plain_value, 2),
            "unit": "queries/s",
            "p50_ms": round(1e3 * float(np.percentile(plain_lat, 50)), 3),
            "p99_ms": round(1e3 * float(np.percentile(plain_lat, 99)), 3),
            "queries": args.plain_queries,
        },
        "served_vs_plain": round(value / plain_value, 2) if plain_value > 0 else None,
        "setup_s": {"factor_mixed": round(setup_served, 3), "factor_fp64": round(setup_plain, 3)},
        "phase_ms_per_step": {k: round(v[0] / args.steps, 3) for k, v in timing.items() if v[1]},
        "refinements": last_stats,
        "roofline": roofline,
    }

    if not args.no_cpu_baseline:
        from oracle import heat_oracle as ho

        log("== cpu baseline (oracle, SciPy SuperLU, 1 thread)")
        t0 = time.perf_counter()
        oracle = ho.FactoredOracle(vertices, faces)
        oracle_setup = time.perf_counter() - t0
        sample = list(range(min(args.cpu_sample, len(source_sets))))
        errs = []
        plain_errs = []
        t0 = time.perf_counter()
        expected = [oracle.distance(np.array([sources[q]])) for q in sample]
        cpu_elapsed = time.perf_counter() - t0
        served_host = out.cpu().numpy()
        for i, q in enumerate(sample):
            errs.append(ho.relative_error(served_host[q], expected[i]))
            if q < len(plain_out):
                plain_errs.append(ho.relative_error(plain_out[q].cpu().numpy(), expected[i]))
        result["cpu_baseline"] = {
            "value": round(len(sample) / cpu_elapsed, 4),
            "unit": "queries/s",
            "cores": 1,
            "kind": "port",
            "sample": f"{len(sample)} single-source queries on the same torus through the FP64 "
            f"oracle with both systems pre-factored (factorization {oracle_setup:.1f}s excluded)",
            "cpu": _cpu_model(),
            "nproc": os.cpu_count(),
        }
        result["rel_err"] = {
            "served_max": max(e[0] for e in errs),
            "served_mean": float(np.mean([e[1] for e in errs])),
            "plain_max": max(e[0] for e in plain_errs) if plain_errs else None,
            "queries_checked": len(sample),
        }
    print(json.dumps(result))
    return 0


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
