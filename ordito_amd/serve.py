"""
A geodesic query server with continuous batching over resident per-mesh factors.

ordito answers one call at a time (``heat.py:370-609``) and keeps a mesh's factors in its
``HeatSolver`` (``mesh.py:2187-2265``). Serving many meshes from one GPU adds what a request loop
needs around that, with no counterpart in the reference:

- **Continuous batching.** Requests queue per mesh. Whenever a worker is free, the scheduler takes
  the mesh whose oldest waiting request is oldest, and up to ``max_batch`` of its requests, and
  solves them as one batch (one multi-column refined solve per system). Requests that arrive while
  a batch runs wait for the next batch of their mesh; batches of different meshes run at once on
  the workers' own HIP streams, so their work overlaps on the device.
- **Admission by memory.** A mesh's operators and factors are built on its first batch and stay
  resident. Every resident mesh counts every device byte it holds (operators, plan, factors and
  its solve workspaces, measured by its worker after each batch); before building a new one, the
  scheduler evicts the least recently used idle meshes -- freeing all of their device state -- until
  the new mesh's estimated footprint (factors, the transient front arena and the workspaces of a
  ``max_batch``-wide batch) fits the HBM budget. A mesh whose factorization the library refuses
  under its own budget is served by iteration.
- **Latency.** Every request records its submit and completion times; ``stats()`` reports
  throughput and p50/p99.

- **Lanes.** An ``hg_mesh`` serves one stream at a time (``heat.py:298-308``), so a mesh is
  solved through *lanes*: independent handles of it, each with its own operators, factors and
  workspaces. A mesh opens a second lane (up to ``replicas``) when a full batch of its requests
  waits while every lane it has is busy and the budget holds another copy; its batches then run on
  two workers at once, and the latency-bound levels of one batch's triangular sweeps overlap the
  bandwidth-bound levels of the other's. The factorization is deterministic, so every lane returns
  the same bits for the same sources.

- **Batching across meshes.** With ``pack_vertices``, meshes of at most that many vertices are
  packed at start-up into groups (``MeshBatch``: one block-diagonal system of up to
  ``pack_group_vertices`` vertices, ``hg_mesh_create_packed``), and a group is served like one
  mesh: its requests share one queue, and a batch takes them whatever their mesh, solved in one
  batched call (``hg_geodesic_packed``), each query a column of its own mesh's rows. A packed
  batch sweeps the whole group's factors whichever meshes it asks about, so packing pays where the
  group's meshes are all asked at once (high load, many small meshes) and not at a trickle.

A batch's results are the ones each request gets alone, within the refinement's agreement (both
are refined solves through the same factors).
"""

from __future__ import annotations

import collections
import concurrent.futures
import threading
import time
from dataclasses import dataclass, field
from typing import Any

import numpy as np

from . import _arrays
from .batch import MeshBatch
from .heat import HeatSolver
from .mesh import Trimesh


@dataclass(frozen=True)
class _Pack:
    """The key of a group of packed meshes."""

    index: int


@dataclass
class _Request:
    mesh_id: Any
    sources: Any
    future: concurrent.futures.Future
    submitted: float
    consumer: Any = None  # the submitting thread's current stream, which will read the result
    started: float = 0.0
    done: float = 0.0
    member: int = 0  # the mesh's position in its packed group (0 for a mesh served alone)


@dataclass(frozen=True)
class Completion:
    """When a finished request was submitted, started and done (perf_counter seconds): the server
    keeps these, not the requests, whose results belong to the callers."""

    mesh_id: Any
    submitted: float
    started: float
    done: float


@dataclass
class _Lane:
    """One handle of a mesh: its own operators, factors and workspaces."""

    trimesh: Trimesh | None = None  # made by the lane's first batch's worker, on its stream
    solver: HeatSolver | MeshBatch | None = None  # a packed group's lane holds a MeshBatch
    factored: bool = False
    nbytes: int = 0  # every device byte the lane holds, measured after each batch
    busy: bool = False


@dataclass
class _Resident:
    n_vertices: int
    n_faces: int
    estimate: int = 0  # a lane's footprint before its first batch measures it
    last_used: float = 0.0
    lanes: list = field(default_factory=list)
    pending: collections.deque = field(default_factory=collections.deque)

    @property
    def busy(self) -> bool:
        return any(lane.busy for lane in self.lanes)

    @property
    def factored(self) -> bool:
        return any(lane.factored for lane in self.lanes)


def footprint_estimate(n_vertices: int, n_faces: int, precision: str = "fp64",
                       columns: int = 16) -> int:
    """Bytes a mesh holds once factored, plus the transient arena of a factorization and the solve
    buffers of a ``columns``-wide batch, from its size.

    Measured on the benchmark torus (524,288 vertices): 56.8 M inverse-factor entries (108 per
    vertex) and 220 M front-arena entries (420 per vertex); surfaces of other sizes scale close to
    V log V, so the per-vertex figures grow with log2(V) / 19.
    """
    scale = max(1.0, np.log2(max(n_vertices, 2)) / 19.0)
    factor_entries = 108.0 * n_vertices * scale
    arena_entries = 420.0 * n_vertices * scale
    factor_bytes = factor_entries * (8 + (4 if precision == "mixed" else 0)) * 2  # two systems
    operators = 3.67 / 0.524 * n_vertices * 8 * 5 + n_faces * (40 + 120)  # + the face geometry
    # Three fields and six refinement vectors of n x columns, the update vectors (about 2.5 n).
    buffers = (3 + 6 + 2.5 * scale) * n_vertices * columns * 8
    return int(factor_bytes + arena_entries * 8 + operators + buffers)


class GeodesicServer:
    """
    ``meshes``: mapping from a mesh id to ``(vertices, faces)``. ``budget_bytes``: the HBM the
    resident meshes may hold. ``workers``: batches solved at once (each on its own stream).
    ``replicas``: the lanes (independent handles) one mesh may hold, so that up to that many of its
    batches run at once. ``pack_vertices``: meshes of at most this many vertices are served in
    packed groups of up to ``pack_group_vertices`` vertices (0: none).
    """

    def __init__(self, meshes: dict, *, budget_bytes: float = 200e9, max_batch: int = 64,
                 workers: int = 4, precision: str = "fp64", warm_columns: int = 16,
                 replicas: int = 2, pack_vertices: int = 0,
                 pack_group_vertices: int = 2_000_000) -> None:
        _arrays.require_gpu()
        self._torch = _arrays.torch_module()
        self._meshes = dict(meshes)
        self.budget = int(budget_bytes)
        self.max_batch = int(max_batch)
        self.warm_columns = int(warm_columns)
        self.precision = precision
        if int(replicas) < 1:
            raise ValueError(f"replicas must be at least 1, got {replicas}")
        self.replicas = int(replicas)
        # Serving units: a mesh alone, or a packed group (greedily in mesh order).
        self._units: dict[Any, list] = {}
        self._unit_of: dict[Any, tuple[Any, int]] = {}
        group: list = []
        group_vertices = 0

        def close_group() -> None:
            key = _Pack(sum(isinstance(k, _Pack) for k in self._units))
            self._units[key] = list(group)
            for member, mesh_id in enumerate(group):
                self._unit_of[mesh_id] = (key, member)

        for mesh_id, (vertices, _) in self._meshes.items():
            n = int(np.prod(vertices.shape)) // 3
            if int(pack_vertices) > 0 and n <= int(pack_vertices):
                if group and (group_vertices + n > int(pack_group_vertices) or len(group) >= 256):
                    close_group()
                    group, group_vertices = [], 0
                group.append(mesh_id)
                group_vertices += n
            else:
                self._units[mesh_id] = [mesh_id]
                self._unit_of[mesh_id] = (mesh_id, 0)
        if group:
            close_group()
        self._resident: dict[Any, _Resident] = {}
        self._lock = threading.Condition()
        self._closing = False
        self._completed: list[Completion] = []
        self._evictions = 0
        self._batches = 0
        self._streams = [self._torch.cuda.Stream() for _ in range(workers)]
        self._idle = list(range(workers))
        self._pool = concurrent.futures.ThreadPoolExecutor(max_workers=workers)
        self._scheduler = threading.Thread(target=self._schedule, name="geodesic-scheduler", daemon=True)
        self._scheduler.start()

    # -- client side -----------------------------------------------------------------------------

    def warm(self, columns: int | None = None) -> None:
        """
        Factor every mesh and size its solve buffers for batches of up to ``columns`` queries
        (default ``warm_columns``, rounded up to a power of two), so that serving such batches
        allocates nothing: a device free waits on every stream's work, and would stall every worker.
        """
        columns = self.warm_columns if columns is None else int(columns)
        futures = []
        for unit, members in self._units.items():
            with self._lock:
                entry = self._resident.get(unit) or self._new_entry(unit)
                for q in range(int(columns)):
                    member = q % len(members)
                    request = _Request(members[member], np.zeros(1, dtype=np.int32),
                                       concurrent.futures.Future(), time.perf_counter(),
                                       member=member)
                    entry.pending.append(request)
                    futures.append(request.future)
                self._lock.notify_all()
        for future in futures:
            future.result()

    def submit(self, mesh_id, sources) -> concurrent.futures.Future:
        """Queue one query; the future resolves to the (V,) float64 distance on the GPU."""
        if mesh_id not in self._meshes:
            raise ValueError(f"unknown mesh {mesh_id!r}")
        unit, member = self._unit_of[mesh_id]
        request = _Request(mesh_id, np.asarray(sources, dtype=np.int32).reshape(-1),
                           concurrent.futures.Future(), time.perf_counter(),
                           consumer=self._torch.cuda.current_stream(), member=member)
        with self._lock:
            if self._closing:
                raise RuntimeError("the server is closed")
            entry = self._resident.get(unit)
            if entry is None:
                entry = self._new_entry(unit)
            entry.pending.append(request)
            self._lock.notify_all()
        return request.future

    def drain(self, timeout: float | None = None) -> bool:
        """Wait until every request queued so far has been answered (the server keeps running);
        False if ``timeout`` seconds passed first."""
        deadline = None if timeout is None else time.perf_counter() + timeout
        with self._lock:
            while any(e.pending or e.busy for e in self._resident.values()):
                left = None if deadline is None else deadline - time.perf_counter()
                if left is not None and left <= 0:
                    return False
                self._lock.wait(timeout=0.05 if left is None else min(left, 0.05))
        return True

    def close(self) -> None:
        """Finish every queued request, then stop."""
        with self._lock:
            self._closing = True
            self._lock.notify_all()
        self._scheduler.join()
        self._pool.shutdown(wait=True)

    def stats(self, since: float = 0.0) -> dict:
        """Throughput and latency of the requests completed after ``since`` (perf_counter time)."""
        with self._lock:
            done = [r for r in self._completed if r.done >= since]
        if not done:
            return {"completed": 0}
        latency = np.array([r.done - r.submitted for r in done])
        first = min(r.submitted for r in done)
        last = max(r.done for r in done)
        return {
            "completed": len(done),
            "queries_per_s": len(done) / max(last - first, 1e-9),
            "p50_ms": 1e3 * float(np.percentile(latency, 50)),
            "p99_ms": 1e3 * float(np.percentile(latency, 99)),
            "mean_ms": 1e3 * float(latency.mean()),
            "batches": self._batches,
            "evictions": self._evictions,
            "resident": sum(len(self._units[k]) for k, e in self._resident.items() if e.factored),
            "packs": sum(1 for k, e in self._resident.items() if isinstance(k, _Pack) and e.factored),
            "lanes": sum(1 for e in self._resident.values() for lane in e.lanes if lane.factored),
            "resident_bytes": self._held_bytes(),
        }

    # -- scheduling ------------------------------------------------------------------------------

    def _new_entry(self, unit) -> _Resident:
        n_vertices = n_faces = estimate = 0
        for mesh_id in self._units[unit]:
            vertices, faces = self._meshes[mesh_id]
            v = int(np.prod(vertices.shape)) // 3
            f = int(np.prod(faces.shape)) // 3
            n_vertices += v
            n_faces += f
            estimate += footprint_estimate(v, f, self.precision, self.max_batch)
        entry = _Resident(n_vertices, n_faces, estimate=estimate)
        self._resident[unit] = entry
        return entry

    def _held_bytes(self, exclude: _Lane | None = None) -> int:
        """Device bytes of every lane with state on the device but ``exclude`` (a busy lane's as of
        its last batch; a lane being built for the first time counts its mesh's estimate)."""
        return sum(max(lane.nbytes, e.estimate if lane.busy and not lane.factored else 0)
                   for e in self._resident.values() for lane in e.lanes
                   if lane is not exclude and (lane.solver is not None or lane.busy))

    def _admit(self, entry: _Resident, lane: _Lane) -> bool:
        """Evict least recently used idle lanes, freeing all their device state, until ``lane``
        fits the budget. A mesh's first lane is admitted whatever is left (it must serve); a further
        lane only if it fits: False leaves the budget as it was and the lane unbuilt."""
        need = max(entry.estimate, lane.nbytes)
        first = len(entry.lanes) == 1
        if not first and self._held_bytes(exclude=lane) + need > self.budget:
            # A second copy of a mesh is worth no other mesh's eviction.
            return False
        while self._held_bytes(exclude=lane) + need > self.budget:
            idle = [(e.last_used, n, k) for k, e in self._resident.items()
                    for n, other in enumerate(e.lanes)
                    if other.solver is not None and not other.busy and other is not lane]
            if not idle:
                break
            _, n, victim = min(idle, key=lambda item: item[:2])  # never compares unit keys
            e = self._resident[victim]
            other = e.lanes.pop(n)
            other.solver.release()
            # Its hg_mesh (operators, pattern, workspaces) goes with the solver.
            other.solver = None
            other.trimesh = None
            self._evictions += 1
        return True

    def _lane_for(self, entry: _Resident) -> _Lane | None:
        """An idle lane of ``entry``, else a new one when a full batch waits behind busy lanes (or
        the mesh has none) and fewer than ``replicas`` exist, else None."""
        idle = [lane for lane in entry.lanes if not lane.busy]
        if idle:
            return max(idle, key=lambda lane: lane.factored)
        if not entry.lanes or (len(entry.lanes) < self.replicas
                               and len(entry.pending) >= self.max_batch):
            lane = _Lane()
            entry.lanes.append(lane)
            return lane
        return None

    def _next_batch(self):
        """The serving unit (a mesh, or a packed group) whose oldest waiting request is oldest and
        which has a lane to solve it on, with up to max_batch of its requests."""
        order = sorted(((e.pending[0].submitted, k) for k, e in self._resident.items() if e.pending),
                       key=lambda head: head[0])
        for _, unit in order:
            entry = self._resident[unit]
            lane = self._lane_for(entry)
            if lane is None:
                continue
            lane.busy = True
            if not lane.factored and not self._admit(entry, lane) and lane.solver is None:
                # A further lane that does not fit is not built; the batch waits for a lane.
                lane.busy = False
                entry.lanes.remove(lane)
                continue
            batch = [entry.pending.popleft() for _ in range(min(self.max_batch, len(entry.pending)))]
            return unit, entry, lane, batch
        return None

    def _schedule(self) -> None:
        while True:
            with self._lock:
                while True:
                    pick = self._next_batch() if self._idle else None
                    if pick is not None:
                        break
                    if self._closing and not any(e.pending for e in self._resident.values()) \
                            and len(self._idle) == len(self._streams):
                        return
                    self._lock.wait(timeout=0.05)
                unit, entry, lane, batch = pick
                worker = self._idle.pop()
            self._pool.submit(self._run, worker, unit, entry, lane, batch)

    def _run(self, worker: int, unit, entry: _Resident, lane: _Lane,
             batch: list[_Request]) -> None:
        torch = self._torch
        try:
            start = time.perf_counter()
            for r in batch:
                r.started = start
            packed = isinstance(unit, _Pack)
            with torch.cuda.stream(self._streams[worker]):
                if lane.solver is None and packed:
                    lane.solver = MeshBatch([self._meshes[m] for m in self._units[unit]],
                                            precision=self.precision)
                elif lane.solver is None:
                    vertices, faces = self._meshes[unit]
                    lane.trimesh = Trimesh(vertices, faces)
                    lane.solver = lane.trimesh.heat_solver(method="meshlab", precision=self.precision)
                    # Graph replay pays per new (mesh, batch width) capture, and launches from
                    # several threads contend: measured p50 4.5 ms eager against 10.6 ms replayed.
                    lane.solver.set_graphs(False)
                if not lane.factored:
                    # A refusal under the library's own budget leaves the lane iterating (a packed
                    # group solves through its factors only, and raises instead).
                    lane.factored = lane.solver.factor()
                if packed:
                    # One batched call over the group: each query a column of its mesh's rows.
                    fields = lane.solver.geodesic([(r.member, r.sources) for r in batch])
                    storage = fields[0] if fields else None
                else:
                    # The batch's sources in one flat upload, its sets delimited by offsets.
                    offsets = [0]
                    for r in batch:
                        offsets.append(offsets[-1] + int(r.sources.size))
                    flat = torch.from_numpy(np.concatenate([r.sources for r in batch])).to("cuda")
                    fields = lane.solver.geodesic_flat(flat, offsets)
                    storage = fields
                self._streams[worker].synchronize()
                lane.nbytes = lane.solver.device_bytes
            # The results' memory comes from this worker stream's pool: each consumer's stream is
            # recorded on it (a packed batch's fields are views of one allocation), so a result
            # freed while its reads are still queued there is not reused by this worker's next
            # batch before they finish.
            for consumer in {r.consumer for r in batch if r.consumer is not None}:
                if consumer != self._streams[worker] and storage is not None:
                    storage.record_stream(consumer)
            done = time.perf_counter()
            for q, r in enumerate(batch):
                r.done = done
                r.future.set_result(fields[q])
        except BaseException as error:  # every waiting future learns of the failure
            for r in batch:
                if not r.future.done():
                    r.future.set_exception(error)
        finally:
            with self._lock:
                lane.busy = False
                entry.last_used = time.perf_counter()
                self._completed.extend(Completion(r.mesh_id, r.submitted, r.started, r.done)
                                       for r in batch if r.done)
                self._batches += 1
                self._idle.append(worker)
                self._lock.notify_all()


class NativeServer:
    """
    The same request loop in the library itself (``hg_server_*`` in ``include/heatgeo.h``): worker
    threads and streams in C++, for hosts that drive the C ABI directly. ``meshes`` maps an id to
    ``(vertices, faces)``; ``submit`` returns a ticket whose ``result()`` is the (V,) float64
    distance on the GPU.
    """

    class Ticket:
        def __init__(self, server: "NativeServer", ticket: int, out) -> None:
            self._server, self._ticket, self._out, self._done = server, ticket, out, False

        def result(self, timeout: float | None = None):
            if not self._done:
                import ctypes

                from . import _lib

                done = ctypes.c_int32()
                _lib.call("hg_server_wait", self._server._ptr, self._ticket,
                          -1.0 if timeout is None else 1e3 * timeout, ctypes.byref(done))
                if not done.value:
                    raise TimeoutError(f"ticket {self._ticket} is still running")
                self._done = True
            return self._out

    def __init__(self, meshes: dict, *, budget_bytes: float = 200e9, max_batch: int = 64,
                 workers: int = 4, precision: str = "fp64") -> None:
        import ctypes

        from . import _lib

        _arrays.require_gpu()
        torch = _arrays.torch_module()
        codes = {"fp64": _lib.HG_PRECISION_FP64, "mixed": _lib.HG_PRECISION_MIXED}
        handle = ctypes.c_void_p()
        _lib.call("hg_server_create", int(budget_bytes), int(workers), int(max_batch),
                  codes[precision], ctypes.byref(handle))
        self._ptr = handle
        self._ids: dict[Any, int] = {}
        self._sizes: dict[Any, int] = {}
        for key, (vertices, faces) in meshes.items():
            v = _arrays.vertices_tensor(vertices)
            f = _arrays.faces_tensor(faces)
            out = ctypes.c_int64()
            _lib.call("hg_server_add_mesh", self._ptr, _arrays.pointer(v), int(v.shape[0]),
                      _arrays.pointer(f), int(f.numel() // 3), ctypes.byref(out))
            self._ids[key] = int(out.value)
            self._sizes[key] = int(v.shape[0])
        torch.cuda.synchronize()

    def submit(self, mesh_id, sources) -> "NativeServer.Ticket":
        import ctypes

        from . import _lib

        if mesh_id not in self._ids:
            raise ValueError(f"unknown mesh {mesh_id!r}")
        torch = _arrays.torch_module()
        src = np.ascontiguousarray(np.asarray(sources, dtype=np.int32).reshape(-1))
        out = torch.empty(self._sizes[mesh_id], dtype=torch.float64, device="cuda")
        ticket = ctypes.c_int64()
        _lib.call("hg_server_submit", self._ptr, self._ids[mesh_id],
                  src.ctypes.data_as(ctypes.c_void_p), int(src.size), out.data_ptr(),
                  ctypes.byref(ticket))
        return NativeServer.Ticket(self, int(ticket.value), out)

    def stats(self, since: float = 0.0) -> dict:
        import ctypes

        from . import _lib

        report = _lib.ServerStats()
        _lib.call("hg_server_stats", self._ptr, float(since), ctypes.byref(report))
        return report.as_dict()

    def close(self) -> None:
        from . import _lib

        if self._ptr is not None:
            _lib.call("hg_server_destroy", self._ptr)
            self._ptr = None
