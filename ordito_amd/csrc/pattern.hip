// Device construction of the cotangent operator's sparsity pattern and its corner runs, from sorted
// keys as ordito builds it (laplacian.py:1212-1389, kernels/laplacian.py:707-760): an entry (i, j)
// for both directions of every face edge and a diagonal for every referenced vertex, rows sorted
// by column; for each entry the run of corner slots 3f+e whose opposite edge it is, in ascending
// slot order (the order its half-cotangents are summed in, kernels/laplacian.py:970-989); and each
// vertex's corners in ascending order. One stable radix sort of the keys i*n+j carrying their
// slots gives the entries and runs; one of the corners by vertex gives the vertex lists; row and
// vertex bounds are bisections of the sorted keys (no atomic counts). The temporaries live in two
// allocations. The host keeps a copy for the work that runs there (components, the rim, the
// analysis).
#include <hipcub/hipcub.hpp>

#include "kernels.h"

namespace hg {
namespace {

__global__ void check_faces_kernel(const int32_t* __restrict__ faces, int64_t count, int64_t n,
                                   int32_t* __restrict__ referenced, int32_t* __restrict__ invalid) {
  const int64_t q = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (q >= count) return;
  const int32_t v = faces[q];
  if (v < 0 || v >= n) {
    invalid[0] = 1;
    return;
  }
  referenced[v] = 1;
}

// Keys i*n + j and their slots: slot q = 3f + e weights the edge opposite corner e, both directions
// at 2q and 2q + 1, so that equal keys keep their slots ascending through a stable sort; after the
// 6F face keys, a diagonal key per referenced vertex with slot -1.
__global__ void pattern_keys_kernel(const int32_t* __restrict__ faces, int64_t n_faces,
                                    const int32_t* __restrict__ referenced,
                                    const int32_t* __restrict__ ref_pos, int64_t n,
                                    uint64_t* __restrict__ keys, int32_t* __restrict__ slots) {
  const int64_t t = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  const int64_t corners = 3 * n_faces;
  if (t < corners) {
    const int64_t f = t / 3;
    const int e = static_cast<int>(t - 3 * f);
    const uint64_t i = static_cast<uint32_t>(faces[3 * f + (e + 1) % 3]);
    const uint64_t j = static_cast<uint32_t>(faces[3 * f + (e + 2) % 3]);
    keys[2 * t] = i * n + j;
    slots[2 * t] = static_cast<int32_t>(t);
    keys[2 * t + 1] = j * n + i;
    slots[2 * t + 1] = static_cast<int32_t>(t);
  } else if (t < corners + n) {
    const int64_t v = t - corners;
    if (referenced[v]) {
      const int64_t at = 2 * corners + ref_pos[v];
      keys[at] = static_cast<uint64_t>(v) * n + v;
      slots[at] = -1;
    }
  }
}

// Per sorted key: whether it starts an entry, and whether it carries a slot.
__global__ void pattern_heads_kernel(const uint64_t* __restrict__ keys,
                                     const int32_t* __restrict__ slots, int64_t count,
                                     int32_t* __restrict__ head, int32_t* __restrict__ carries) {
  const int64_t k = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (k >= count) return;
  head[k] = (k == 0 || keys[k] != keys[k - 1]) ? 1 : 0;
  carries[k] = slots[k] >= 0 ? 1 : 0;
}

// With entry = the inclusive count of heads and position = the exclusive count of slots: each
// entry's column and the start of its run; every slot into the runs.
__global__ void pattern_entries_kernel(const uint64_t* __restrict__ keys,
                                       const int32_t* __restrict__ slots,
                                       const int32_t* __restrict__ head,
                                       const int32_t* __restrict__ entry,
                                       const int32_t* __restrict__ position, int64_t count,
                                       int64_t n, int32_t* __restrict__ columns,
                                       int32_t* __restrict__ corner_ptr,
                                       int32_t* __restrict__ corner_slots) {
  const int64_t k = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (k >= count) return;
  if (head[k]) {
    const int32_t e = entry[k] - 1;
    columns[e] = static_cast<int32_t>(keys[k] % static_cast<uint64_t>(n));
    corner_ptr[e] = position[k];
  }
  if (slots[k] >= 0) corner_slots[position[k]] = slots[k];
}

// Row bounds from the sorted keys: row v starts at the entries before the first key of row v or
// later (its lower bound among the keys, by bisection), so rows with no entries get bounds too
// and no count is accumulated.
__global__ void row_offsets_kernel(const uint64_t* __restrict__ keys,
                                   const int32_t* __restrict__ entry, int64_t count, int64_t n,
                                   int32_t* __restrict__ offsets) {
  const int64_t v = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (v > n) return;
  const uint64_t target = static_cast<uint64_t>(v) * static_cast<uint64_t>(n);
  int64_t lo = 0, hi = count;
  while (lo < hi) {
    const int64_t mid = lo + (hi - lo) / 2;
    if (keys[mid] < target) lo = mid + 1;
    else hi = mid;
  }
  offsets[v] = lo == 0 ? 0 : entry[lo - 1];
}

// Each vertex's first corner in the corners sorted by vertex (its lower bound, by bisection).
__global__ void vertex_offsets_kernel(const uint32_t* __restrict__ vertex_of, int64_t count,
                                      int64_t n, int32_t* __restrict__ vertex_ptr) {
  const int64_t v = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (v > n) return;
  int64_t lo = 0, hi = count;
  while (lo < hi) {
    const int64_t mid = lo + (hi - lo) / 2;
    if (static_cast<int64_t>(vertex_of[mid]) < v) lo = mid + 1;
    else hi = mid;
  }
  vertex_ptr[v] = static_cast<int32_t>(lo);
}

__global__ void iota_kernel(int32_t* __restrict__ out, int64_t count) {
  const int64_t q = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (q < count) out[q] = static_cast<int32_t>(q);
}

__global__ void set_word_kernel(int32_t* p, int32_t value) { p[0] = value; }

unsigned grid_of(int64_t n) {
  return static_cast<unsigned>((std::max<int64_t>(n, 1) + kBlock - 1) / kBlock);
}

// Bits a radix sort needs for keys in [0, bound].
int bits_for(uint64_t bound) {
  int bits = 1;
  while (bits < 64 && (bound >> bits) != 0) ++bits;
  return bits;
}

// hipCUB calls on caller-provided temporary storage: each first asked its size (temp null), then
// run with it.
size_t exclusive_scan(const int32_t* in, int32_t* out, int64_t items, void* temp, size_t bytes,
                      hipStream_t stream) {
  HG_HIP(hipcub::DeviceScan::ExclusiveSum(temp, bytes, in, out, static_cast<int>(items), stream));
  return bytes;
}

size_t inclusive_scan(const int32_t* in, int32_t* out, int64_t items, void* temp, size_t bytes,
                      hipStream_t stream) {
  HG_HIP(hipcub::DeviceScan::InclusiveSum(temp, bytes, in, out, static_cast<int>(items), stream));
  return bytes;
}

template <typename K, typename V>
size_t sort_pairs(const K* keys_in, K* keys_out, const V* values_in, V* values_out, int64_t items,
                  int end_bit, void* temp, size_t bytes, hipStream_t stream) {
  if (items == 0) return 0;
  HG_HIP(hipcub::DeviceRadixSort::SortPairs(temp, bytes, keys_in, keys_out, values_in, values_out,
                                            static_cast<int>(items), 0, end_bit, stream));
  return bytes;
}

// Carves 256-byte aligned pieces out of one allocation.
struct Carver {
  uint8_t* base;
  size_t at = 0;
  template <typename T>
  T* take(int64_t count) {
    T* p = base ? reinterpret_cast<T*>(base + at) : nullptr;  // a null base only sizes
    at += (sizeof(T) * static_cast<size_t>(std::max<int64_t>(count, 1)) + 255) & ~size_t{255};
    return p;
  }
};

template <typename T>
std::vector<T> download(const T* p, int64_t n, hipStream_t stream) {
  std::vector<T> out(static_cast<size_t>(std::max<int64_t>(n, 0)));
  if (n > 0) HG_HIP(hipMemcpyAsync(out.data(), p, sizeof(T) * n, hipMemcpyDeviceToHost, stream));
  return out;
}

}  // namespace

void build_pattern_device(int64_t n, const int32_t* faces, int64_t n_faces, hipStream_t stream,
                          PatternArrays out, HostPattern* host) {
  const int64_t corners = 3 * n_faces;
  HG_REQUIRE(6 * n_faces + n < (int64_t{1} << 31), HG_ERR_VALUE,
             "mesh too large for int32 pattern offsets");
  // Range check and the referenced vertices, counted by a scan (each array has room for a total),
  // in one allocation.
  size_t scan_a = exclusive_scan(nullptr, nullptr, n + 1, nullptr, 0, stream);
  Carver size_a{nullptr};
  size_a.take<int32_t>(n + 1);
  size_a.take<int32_t>(n + 1);
  size_a.take<int32_t>(1);
  size_a.take<uint8_t>(static_cast<int64_t>(scan_a));
  DeviceArray<uint8_t> arena_a(static_cast<int64_t>(size_a.at));
  Carver ca{arena_a.ptr};
  int32_t* referenced = ca.take<int32_t>(n + 1);
  int32_t* ref_pos = ca.take<int32_t>(n + 1);
  int32_t* invalid = ca.take<int32_t>(1);
  void* temp_a = ca.take<uint8_t>(static_cast<int64_t>(scan_a));
  launch_zero(referenced, sizeof(int32_t) * (n + 1), stream);
  launch_zero(invalid, sizeof(int32_t), stream);
  if (corners > 0)
    check_faces_kernel<<<grid_of(corners), kBlock, 0, stream>>>(faces, corners, n, referenced,
                                                                 invalid);
  HG_HIP(hipGetLastError());
  exclusive_scan(referenced, ref_pos, n + 1, temp_a, scan_a, stream);
  int32_t words[2] = {0, 0};
  HG_HIP(hipMemcpyAsync(&words[0], invalid, sizeof(int32_t), hipMemcpyDeviceToHost, stream));
  HG_HIP(hipMemcpyAsync(&words[1], ref_pos + n, sizeof(int32_t), hipMemcpyDeviceToHost, stream));
  HG_HIP(hipStreamSynchronize(stream));
  HG_REQUIRE(words[0] == 0, HG_ERR_VALUE,
             "every face index must be a vertex index in [0, n_vertices)");
  const int64_t count = 2 * corners + words[1];
  const int64_t n_slots = 2 * corners;  // every face key carries its slot, no diagonal does

  // Everything else in a second allocation: the keys and slots (sorted and not), the heads and
  // slot flags and their scans, the corners sorted by vertex, and the sorts' and scans' storage.
  const int key_bits = bits_for(n > 0 ? static_cast<uint64_t>(n) * n - 1 : 0);
  const int vertex_bits = bits_for(n > 0 ? static_cast<uint64_t>(n - 1) : 0);
  size_t temp_bytes = 0;
  temp_bytes = std::max(temp_bytes, sort_pairs<uint64_t, int32_t>(nullptr, nullptr, nullptr,
                                                                  nullptr, count, key_bits,
                                                                  nullptr, 0, stream));
  temp_bytes = std::max(temp_bytes, sort_pairs<uint32_t, int32_t>(nullptr, nullptr, nullptr,
                                                                  nullptr, corners, vertex_bits,
                                                                  nullptr, 0, stream));
  temp_bytes = std::max(temp_bytes, inclusive_scan(nullptr, nullptr, count + 1, nullptr, 0, stream));
  temp_bytes = std::max(temp_bytes, exclusive_scan(nullptr, nullptr, count + 1, nullptr, 0, stream));
  auto carve_b = [&](Carver& c) {
    struct {
      uint64_t *keys, *sorted_keys;
      int32_t *slots, *sorted_slots, *head, *carries, *entry, *position, *corner_ids;
      uint32_t* vertex_of;
      void* temp;
    } b;
    b.keys = c.take<uint64_t>(count);
    b.sorted_keys = c.take<uint64_t>(count);
    b.slots = c.take<int32_t>(count);
    b.sorted_slots = c.take<int32_t>(count);
    b.head = c.take<int32_t>(count + 1);
    b.carries = c.take<int32_t>(count + 1);
    b.entry = c.take<int32_t>(count + 1);
    b.position = c.take<int32_t>(count + 1);
    b.corner_ids = c.take<int32_t>(corners);
    b.vertex_of = c.take<uint32_t>(corners);
    b.temp = c.take<uint8_t>(static_cast<int64_t>(temp_bytes));
    return b;
  };
  Carver size_b{nullptr};
  carve_b(size_b);
  DeviceArray<uint8_t> arena_b(static_cast<int64_t>(size_b.at));
  Carver cb{arena_b.ptr};
  const auto b = carve_b(cb);

  // The entries: one stable sort of the keys carrying their slots.
  pattern_keys_kernel<<<grid_of(corners + n), kBlock, 0, stream>>>(faces, n_faces, referenced,
                                                                    ref_pos, n, b.keys, b.slots);
  HG_HIP(hipGetLastError());
  sort_pairs(b.keys, b.sorted_keys, b.slots, b.sorted_slots, count, key_bits, b.temp, temp_bytes,
             stream);
  launch_zero(b.head + count, sizeof(int32_t), stream);
  launch_zero(b.carries + count, sizeof(int32_t), stream);
  if (count > 0)
    pattern_heads_kernel<<<grid_of(count), kBlock, 0, stream>>>(b.sorted_keys, b.sorted_slots,
                                                                 count, b.head, b.carries);
  HG_HIP(hipGetLastError());
  inclusive_scan(b.head, b.entry, count + 1, b.temp, temp_bytes, stream);
  exclusive_scan(b.carries, b.position, count + 1, b.temp, temp_bytes, stream);
  int32_t entries = 0;
  HG_HIP(hipMemcpyAsync(&entries, b.entry + count, sizeof(int32_t), hipMemcpyDeviceToHost, stream));
  HG_HIP(hipStreamSynchronize(stream));
  const int64_t nnz = entries;
  out.columns.allocate(std::max<int64_t>(nnz, 1));
  out.corner_ptr.allocate(nnz + 1);
  out.corner_slots.allocate(std::max<int64_t>(n_slots, 1));
  out.offsets.allocate(n + 1);
  if (count > 0)
    pattern_entries_kernel<<<grid_of(count), kBlock, 0, stream>>>(
        b.sorted_keys, b.sorted_slots, b.head, b.entry, b.position, count, n, out.columns.ptr,
        out.corner_ptr.ptr, out.corner_slots.ptr);
  set_word_kernel<<<1, 1, 0, stream>>>(out.corner_ptr.ptr + nnz, static_cast<int32_t>(n_slots));
  row_offsets_kernel<<<grid_of(n + 1), kBlock, 0, stream>>>(b.sorted_keys, b.entry, count, n,
                                                             out.offsets.ptr);
  HG_HIP(hipGetLastError());

  // Each vertex's corners: a stable sort of the corner ids by vertex.
  out.vertex_corners.allocate(std::max<int64_t>(corners, 1));
  out.vertex_ptr.allocate(n + 1);
  if (corners > 0) iota_kernel<<<grid_of(corners), kBlock, 0, stream>>>(b.corner_ids, corners);
  HG_HIP(hipGetLastError());
  sort_pairs(reinterpret_cast<const uint32_t*>(faces), b.vertex_of,
             static_cast<const int32_t*>(b.corner_ids), out.vertex_corners.ptr, corners,
             vertex_bits, b.temp, temp_bytes, stream);
  vertex_offsets_kernel<<<grid_of(n + 1), kBlock, 0, stream>>>(b.vertex_of, corners, n,
                                                                out.vertex_ptr.ptr);
  HG_HIP(hipGetLastError());
  if (host == nullptr) {
    // The temporaries are freed on return: their work must be done first.
    HG_HIP(hipStreamSynchronize(stream));
    return;
  }

  // The host copy, and each row's diagonal entry.
  HostPattern& h = *host;
  h = HostPattern{};
  h.n = n;
  h.offsets = download(out.offsets.ptr, n + 1, stream);
  h.columns = download(out.columns.ptr, nnz, stream);
  h.corner_ptr = download(out.corner_ptr.ptr, nnz + 1, stream);
  h.corner_slots = download(out.corner_slots.ptr, n_slots, stream);
  h.vertex_ptr = download(out.vertex_ptr.ptr, n + 1, stream);
  h.vertex_corners = download(out.vertex_corners.ptr, corners, stream);
  HG_HIP(hipStreamSynchronize(stream));
  h.diagonal.assign(static_cast<size_t>(n), -1);
  for (int64_t v = 0; v < n; ++v)
    for (int32_t e = h.offsets[v]; e < h.offsets[v + 1]; ++e)
      if (h.columns[e] == v) h.diagonal[v] = e;
}

}  // namespace hg
