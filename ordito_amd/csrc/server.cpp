// The geodesic query server behind hg_server_* (include/heatgeo.h): continuous batching over
// resident per-mesh factors, in C++ on the library's own C ABI. No counterpart in the reference,
// which answers one call at a time (heat.py:370-609) and keeps a mesh's factors in its HeatSolver
// (mesh.py:2187-2265).
//
// - Requests queue per mesh. Each free worker thread takes the mesh whose oldest waiting request
//   is oldest, with up to max_batch of its requests; each worker owns a HIP stream (default
//   priority) and solves its batch through the mesh's factors, so batches of different meshes
//   overlap on the device while requests arriving meanwhile join the next batch of their mesh. One
//   mesh is never solved by two workers at once (an hg_mesh serves one stream at a time).
// - A mesh is factored on its first batch. Before that its worker evicts least recently used
//   idle meshes (hg_release: their factors and workspaces) until the mesh's footprint -- its
//   analysis' factor and front-arena bytes plus the workspaces of a max_batch-wide batch -- fits
//   the budget beside every device byte the resident meshes hold (hg_device_bytes).
// - Each request keeps its submit and completion times; hg_server_stats reports throughput and
//   p50/p99 latency.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <limits>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "hg_internal.h"

namespace hg {

namespace {

using Clock = std::chrono::steady_clock;

double seconds_since(Clock::time_point origin) {
  return std::chrono::duration<double>(Clock::now() - origin).count();
}

struct Request {
  int64_t ticket = 0;
  std::vector<int32_t> sources;
  double* d_out = nullptr;
  double submitted = 0.0;
};

struct Ticket {
  bool done = false;
  int32_t status = HG_OK;
  std::string message;
};

struct Completion {
  double submitted, done;
};

struct Mesh {
  hg_mesh* handle = nullptr;
  int64_t n = 0;
  bool factored = false;
  bool busy = false;
  int64_t nbytes = 0;    // every device byte its state holds, measured after each batch
  int64_t estimate = 0;  // factored footprint before its first factorization
  double last_used = 0.0;
  std::deque<Request> pending;
};

struct Worker {
  hipStream_t stream = nullptr;
  int32_t* d_sources = nullptr;
  int64_t sources_capacity = 0;
  double* d_fields = nullptr;
  int64_t fields_capacity = 0;
};

}  // namespace

}  // namespace hg

struct hg_server {
  std::mutex lock;
  std::condition_variable wake;      // the workers: a request arrived or a mesh was freed
  std::condition_variable finished;  // waiters: a ticket completed
  hg::Clock::time_point origin = hg::Clock::now();
  int64_t budget = 0;
  int32_t max_batch = 64;
  int32_t precision = HG_PRECISION_FP64;
  bool closing = false;
  int64_t next_ticket = 1;
  std::map<int64_t, hg::Mesh> meshes;
  std::map<int64_t, hg::Ticket> tickets;
  std::vector<hg::Completion> completions;
  int64_t batches = 0, evictions = 0;
  std::vector<hg::Worker> workers;
  std::vector<std::thread> threads;  // one a worker: each takes the next batch when it is free
};

namespace hg {
namespace {

template <typename F>
int guarded_server(F&& body) {
  try {
    body();
    return HG_OK;
  } catch (const Error& e) {
    set_error(e.message);
    return e.status;
  } catch (const std::exception& e) {
    set_error(e.what());
    return HG_ERR_RUNTIME;
  }
}

void check(int status) {
  if (status != HG_OK) throw Error{status, hg_last_error()};
}

// Bytes of every mesh with device state but `exclude` (a mesh being built counts its estimate).
int64_t held_bytes(const hg_server& s, int64_t exclude) {
  int64_t total = 0;
  for (const auto& [id, m] : s.meshes)
    if (id != exclude) total += std::max(m.nbytes, m.busy && !m.factored ? m.estimate : 0);
  return total;
}

// Evict least recently used idle meshes until `id` fits (the lock is held).
void admit(hg_server& s, int64_t id, Mesh& mesh) {
  const int64_t need = std::max(mesh.estimate, mesh.nbytes);
  while (held_bytes(s, id) + need > s.budget) {
    int64_t victim = -1;
    double oldest = 0.0;
    for (const auto& [k, m] : s.meshes)
      if (k != id && !m.busy && m.factored && (victim < 0 || m.last_used < oldest)) {
        victim = k;
        oldest = m.last_used;
      }
    if (victim < 0) break;
    Mesh& v = s.meshes[victim];
    (void)hg_release(v.handle);
    (void)hg_device_bytes(v.handle, &v.nbytes);
    v.factored = false;
    ++s.evictions;
  }
}

// The factored footprint of a mesh from its analysis: both systems' inverse factors and the
// transient front arena, plus a max_batch-wide batch's fields and refinement vectors.
int64_t footprint(hg_mesh* mesh, int64_t n, int32_t max_batch, int32_t precision) {
  int64_t stats[8] = {};
  check(hg_plan_stats(mesh, stats));
  const int64_t w_bytes = stats[2] * (precision == HG_PRECISION_MIXED ? 12 : 8);
  const int64_t arena_bytes = 8 * stats[4];
  const int64_t vectors = 8 * (3 + 6) * n * std::min<int64_t>(max_batch, 64) + 8 * stats[6] * 64;
  int64_t now = 0;
  check(hg_device_bytes(mesh, &now));
  return now + 2 * w_bytes + arena_bytes + vectors;
}

// Solves one batch on worker w's stream (the mesh is held busy by this worker).
void run_batch(hg_server* s, Worker& worker, Mesh* mesh, std::vector<Request>& batch) {
  int32_t status = HG_OK;
  std::string message;
  try {
    check(hg_mesh_set_stream(mesh->handle, worker.stream));
    if (!mesh->factored) {
      int32_t ok = 0;
      check(hg_factor(mesh->handle, HG_SYSTEM_HEAT | HG_SYSTEM_POISSON, s->precision, &ok));
      mesh->factored = true;  // a refusal leaves the mesh iterating, which is an answer too
    }
    std::vector<int64_t> offsets(batch.size() + 1, 0);
    for (size_t q = 0; q < batch.size(); ++q)
      offsets[q + 1] = offsets[q] + static_cast<int64_t>(batch[q].sources.size());
    const int64_t total = std::max<int64_t>(offsets.back(), 1);
    if (worker.sources_capacity < total) {
      HG_HIP(hipStreamSynchronize(worker.stream));
      if (worker.d_sources) (void)hipFree(worker.d_sources);
      worker.d_sources = nullptr;
      worker.sources_capacity = std::max<int64_t>(total, 2 * worker.sources_capacity);
      HG_HIP(hipMalloc(&worker.d_sources, sizeof(int32_t) * worker.sources_capacity));
    }
    std::vector<int32_t> flat(static_cast<size_t>(total), 0);
    for (size_t q = 0; q < batch.size(); ++q)
      std::copy(batch[q].sources.begin(), batch[q].sources.end(), flat.begin() + offsets[q]);
    HG_HIP(hipMemcpyAsync(worker.d_sources, flat.data(), sizeof(int32_t) * total,
                          hipMemcpyHostToDevice, worker.stream));
    const int64_t k = static_cast<int64_t>(batch.size());
    if (worker.fields_capacity < k * mesh->n) {
      HG_HIP(hipStreamSynchronize(worker.stream));
      if (worker.d_fields) (void)hipFree(worker.d_fields);
      worker.d_fields = nullptr;
      worker.fields_capacity = std::max<int64_t>(k * mesh->n, 2 * worker.fields_capacity);
      HG_HIP(hipMalloc(&worker.d_fields, sizeof(double) * worker.fields_capacity));
    }
    check(hg_geodesic(mesh->handle, worker.d_sources, offsets.data(), k, worker.d_fields,
                      HG_MODE_FACTORED, s->precision, nullptr));
    for (int64_t q = 0; q < k; ++q)
      HG_HIP(hipMemcpyAsync(batch[q].d_out, worker.d_fields + q * mesh->n,
                            sizeof(double) * mesh->n, hipMemcpyDeviceToDevice, worker.stream));
    HG_HIP(hipStreamSynchronize(worker.stream));
    int64_t bytes = 0;
    check(hg_device_bytes(mesh->handle, &bytes));
    mesh->nbytes = bytes;
  } catch (const Error& e) {
    status = e.status;
    message = e.message;
  } catch (const std::exception& e) {
    status = HG_ERR_RUNTIME;
    message = e.what();
  }
  std::lock_guard<std::mutex> guard(s->lock);
  const double done = seconds_since(s->origin);
  for (const Request& r : batch) {
    Ticket& t = s->tickets[r.ticket];
    t.done = true;
    t.status = status;
    t.message = message;
    if (status == HG_OK) s->completions.push_back({r.submitted, done});
  }
  mesh->busy = false;
  mesh->last_used = done;
  ++s->batches;
  s->finished.notify_all();
  s->wake.notify_all();
}

// A worker's loop: take the mesh whose oldest waiting request is oldest and that no other worker
// holds, with up to max_batch of its requests; admit it if it has no factors; solve; repeat. Ends
// when the server closes and nothing is queued.
void work(hg_server* s, int32_t w) {
  Worker& worker = s->workers[w];
  std::unique_lock<std::mutex> guard(s->lock);
  for (;;) {
    int64_t pick = -1;
    double head = 0.0;
    bool queued = false;
    for (auto& [id, m] : s->meshes) {
      queued = queued || !m.pending.empty();
      if (!m.busy && !m.pending.empty() && (pick < 0 || m.pending.front().submitted < head)) {
        pick = id;
        head = m.pending.front().submitted;
      }
    }
    if (pick < 0) {
      if (s->closing && !queued) return;
      s->wake.wait_for(guard, std::chrono::milliseconds(20));
      continue;
    }
    Mesh& mesh = s->meshes[pick];
    std::vector<Request> batch;
    while (!mesh.pending.empty() && static_cast<int32_t>(batch.size()) < s->max_batch) {
      batch.push_back(std::move(mesh.pending.front()));
      mesh.pending.pop_front();
    }
    mesh.busy = true;
    if (!mesh.factored) admit(*s, pick, mesh);
    guard.unlock();
    run_batch(s, worker, &mesh, batch);
    guard.lock();
  }
}

}  // namespace
}  // namespace hg

using namespace hg;

int hg_server_create(int64_t budget_bytes, int32_t workers, int32_t max_batch, int32_t precision,
                     hg_server** out_server) {
  return guarded_server([&] {
    HG_REQUIRE(out_server != nullptr, HG_ERR_VALUE, "out_server is null");
    HG_REQUIRE(budget_bytes > 0, HG_ERR_VALUE, "the server's memory budget must be positive");
    HG_REQUIRE(workers >= 1 && workers <= 16, HG_ERR_VALUE, "workers must be 1-16");
    HG_REQUIRE(max_batch >= 1 && max_batch <= 4096, HG_ERR_VALUE, "max_batch must be 1-4096");
    HG_REQUIRE(precision == HG_PRECISION_FP64 || precision == HG_PRECISION_MIXED, HG_ERR_VALUE,
               "precision must be HG_PRECISION_FP64 or HG_PRECISION_MIXED");
    auto s = std::make_unique<hg_server>();
    s->budget = budget_bytes;
    s->max_batch = max_batch;
    s->precision = precision;
    s->workers.resize(workers);
    for (int32_t w = 0; w < workers; ++w)
      HG_HIP(hipStreamCreateWithFlags(&s->workers[w].stream, hipStreamNonBlocking));
    for (int32_t w = 0; w < workers; ++w) s->threads.emplace_back(work, s.get(), w);
    *out_server = s.release();
  });
}

int hg_server_add_mesh(hg_server* server, const float* d_vertices, int64_t n_vertices,
                       const int32_t* d_faces, int64_t n_faces, int64_t* out_mesh_id) {
  return guarded_server([&] {
    HG_REQUIRE(server != nullptr && out_mesh_id != nullptr, HG_ERR_VALUE, "null argument");
    hg_mesh* handle = nullptr;
    check(hg_mesh_create(d_vertices, n_vertices, d_faces, n_faces, nullptr, &handle));
    try {
      check(hg_operators(handle, std::numeric_limits<double>::quiet_NaN(), HG_METHOD_MESHLAB, 1));
      // Graph replay pays a capture per new (mesh, batch width) and launches from several threads
      // contend; the server's batches launch eagerly.
      check(hg_set_graphs(handle, 0));
      std::lock_guard<std::mutex> guard(server->lock);
      const int64_t id = static_cast<int64_t>(server->meshes.size());
      Mesh& mesh = server->meshes[id];
      mesh.handle = handle;
      mesh.n = n_vertices;
      mesh.estimate = footprint(handle, n_vertices, server->max_batch, server->precision);
      check(hg_device_bytes(handle, &mesh.nbytes));
      *out_mesh_id = id;
    } catch (...) {
      (void)hg_mesh_destroy(handle);
      throw;
    }
  });
}

int hg_server_submit(hg_server* server, int64_t mesh_id, const int32_t* sources,
                     int64_t n_sources, double* d_out, int64_t* out_ticket) {
  return guarded_server([&] {
    HG_REQUIRE(server != nullptr && out_ticket != nullptr, HG_ERR_VALUE, "null argument");
    HG_REQUIRE(n_sources >= 0 && (n_sources == 0 || sources != nullptr), HG_ERR_VALUE,
               "sources is null");
    HG_REQUIRE(d_out != nullptr, HG_ERR_VALUE, "d_out is null");
    std::lock_guard<std::mutex> guard(server->lock);
    HG_REQUIRE(!server->closing, HG_ERR_RUNTIME, "the server is closing");
    auto it = server->meshes.find(mesh_id);
    HG_REQUIRE(it != server->meshes.end(), HG_ERR_VALUE,
               "unknown mesh " + std::to_string(mesh_id));
    for (int64_t e = 0; e < n_sources; ++e)
      HG_REQUIRE(sources[e] >= 0 && sources[e] < it->second.n, HG_ERR_VALUE,
                 "every entry of sources must be a vertex index in [0, " +
                     std::to_string(it->second.n) + ")");
    Request r;
    r.ticket = server->next_ticket++;
    r.sources.assign(sources, sources + n_sources);
    r.d_out = d_out;
    r.submitted = seconds_since(server->origin);
    server->tickets[r.ticket] = Ticket{};
    *out_ticket = r.ticket;
    it->second.pending.push_back(std::move(r));
    server->wake.notify_all();
  });
}

int hg_server_wait(hg_server* server, int64_t ticket, double timeout_ms, int32_t* out_done) {
  return guarded_server([&] {
    HG_REQUIRE(server != nullptr, HG_ERR_VALUE, "null server");
    std::unique_lock<std::mutex> guard(server->lock);
    auto it = server->tickets.find(ticket);
    HG_REQUIRE(it != server->tickets.end(), HG_ERR_VALUE,
               "unknown ticket " + std::to_string(ticket));
    auto ready = [&] { return server->tickets[ticket].done; };
    bool done;
    if (timeout_ms < 0.0) {
      server->finished.wait(guard, ready);
      done = true;
    } else {
      done = server->finished.wait_for(
          guard, std::chrono::duration<double, std::milli>(timeout_ms), ready);
    }
    if (out_done) *out_done = done ? 1 : 0;
    if (done) {
      const Ticket t = server->tickets[ticket];
      server->tickets.erase(ticket);
      if (t.status != HG_OK) throw Error{t.status, t.message};
    }
  });
}

int hg_server_stats(hg_server* server, double since_seconds, hg_server_report* out) {
  return guarded_server([&] {
    HG_REQUIRE(server != nullptr && out != nullptr, HG_ERR_VALUE, "null argument");
    std::lock_guard<std::mutex> guard(server->lock);
    std::vector<double> latency;
    double first = 0.0, last = 0.0;
    for (const Completion& c : server->completions) {
      if (c.done < since_seconds) continue;
      if (latency.empty() || c.submitted < first) first = c.submitted;
      last = std::max(last, c.done);
      latency.push_back(c.done - c.submitted);
    }
    *out = hg_server_report{};
    out->completed = static_cast<int64_t>(latency.size());
    out->batches = server->batches;
    out->evictions = server->evictions;
    for (const auto& [id, m] : server->meshes) {
      out->resident += m.factored ? 1 : 0;
      out->resident_bytes += m.nbytes;
    }
    out->now_seconds = seconds_since(server->origin);
    if (latency.empty()) return;
    std::sort(latency.begin(), latency.end());
    auto pct = [&](double p) {
      const double at = p / 100.0 * static_cast<double>(latency.size() - 1);
      const size_t lo = static_cast<size_t>(at);
      const size_t hi = std::min(lo + 1, latency.size() - 1);
      return 1e3 * (latency[lo] + (at - static_cast<double>(lo)) * (latency[hi] - latency[lo]));
    };
    out->p50_ms = pct(50.0);
    out->p99_ms = pct(99.0);
    double sum = 0.0;
    for (double l : latency) sum += l;
    out->mean_ms = 1e3 * sum / static_cast<double>(latency.size());
    out->queries_per_s = static_cast<double>(latency.size()) / std::max(last - first, 1e-9);
  });
}

int hg_server_destroy(hg_server* server) {
  return guarded_server([&] {
    if (!server) return;
    {
      std::lock_guard<std::mutex> guard(server->lock);
      server->closing = true;
      server->wake.notify_all();
    }
    for (std::thread& t : server->threads) t.join();
    for (auto& [id, m] : server->meshes) (void)hg_mesh_destroy(m.handle);
    for (Worker& w : server->workers) {
      if (w.d_sources) (void)hipFree(w.d_sources);
      if (w.d_fields) (void)hipFree(w.d_fields);
      if (w.stream) (void)hipStreamDestroy(w.stream);
    }
    delete server;
  });
}
