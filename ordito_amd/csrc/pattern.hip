// Device construction of the cotangent operator's sparsity pattern and its corner runs, from sorted
// keys as ordito builds it (laplacian.py:1212-1389, kernels/laplacian.py:707-760): an entry (i, j)
// for both directions of every face edge and a diagonal for every referenced vertex, rows sorted
// by column; for each entry the run of corner slots 3f+e whose opposite edge it is, in ascending
// slot order (the order its half-cotangents are summed in, kernels/laplacian.py:970-989); and each
// vertex's corners in ascending order. One stable radix sort of the keys i*n+j carrying their
// slots gives the entries and runs; one of the corners by vertex gives the vertex lists. The host
// keeps a copy for the work that runs there (components, the rim, the analysis).
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
// entry's column, the start of its run and its row's count; every slot into the runs.
__global__ void pattern_entries_kernel(const uint64_t* __restrict__ keys,
                                       const int32_t* __restrict__ slots,
                                       const int32_t* __restrict__ head,
                                       const int32_t* __restrict__ entry,
                                       const int32_t* __restrict__ position, int64_t count,
                                       int64_t n, int32_t* __restrict__ columns,
                                       int32_t* __restrict__ corner_ptr,
                                       int32_t* __restrict__ row_count,
                                       int32_t* __restrict__ corner_slots) {
  const int64_t k = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (k >= count) return;
  if (head[k]) {
    const int32_t e = entry[k] - 1;
    const uint64_t row = keys[k] / static_cast<uint64_t>(n);
    columns[e] = static_cast<int32_t>(keys[k] - row * n);
    corner_ptr[e] = position[k];
    atomicAdd(&row_count[row], 1);
  }
  if (slots[k] >= 0) corner_slots[position[k]] = slots[k];
}

__global__ void vertex_counts_kernel(const int32_t* __restrict__ faces, int64_t count,
                                     int32_t* __restrict__ corner_ids,
                                     int32_t* __restrict__ vertex_count) {
  const int64_t q = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (q >= count) return;
  corner_ids[q] = static_cast<int32_t>(q);
  atomicAdd(&vertex_count[faces[q]], 1);
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

void grow(DeviceArray<uint8_t>& temp, size_t bytes) {
  if (temp.size < static_cast<int64_t>(bytes)) temp.allocate(static_cast<int64_t>(bytes));
}

// out[0..n] = the exclusive scan of in[0..n], so out[n] is the total of in[0..n) when in[n] = 0.
void exclusive_scan(const int32_t* in, int32_t* out, int64_t n, DeviceArray<uint8_t>& temp,
                    hipStream_t stream) {
  size_t bytes = 0;
  const int items = static_cast<int>(n + 1);
  HG_HIP(hipcub::DeviceScan::ExclusiveSum(nullptr, bytes, in, out, items, stream));
  grow(temp, bytes);
  HG_HIP(hipcub::DeviceScan::ExclusiveSum(temp.ptr, bytes, in, out, items, stream));
}

void inclusive_scan(const int32_t* in, int32_t* out, int64_t n, DeviceArray<uint8_t>& temp,
                    hipStream_t stream) {
  size_t bytes = 0;
  const int items = static_cast<int>(n);
  HG_HIP(hipcub::DeviceScan::InclusiveSum(nullptr, bytes, in, out, items, stream));
  grow(temp, bytes);
  HG_HIP(hipcub::DeviceScan::InclusiveSum(temp.ptr, bytes, in, out, items, stream));
}

template <typename K, typename V>
void sort_pairs(const K* keys_in, K* keys_out, const V* values_in, V* values_out, int64_t count,
                int end_bit, DeviceArray<uint8_t>& temp, hipStream_t stream) {
  if (count == 0) return;
  size_t bytes = 0;
  const int items = static_cast<int>(count);
  HG_HIP(hipcub::DeviceRadixSort::SortPairs(nullptr, bytes, keys_in, keys_out, values_in,
                                            values_out, items, 0, end_bit, stream));
  grow(temp, bytes);
  HG_HIP(hipcub::DeviceRadixSort::SortPairs(temp.ptr, bytes, keys_in, keys_out, values_in,
                                            values_out, items, 0, end_bit, stream));
}

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
  DeviceArray<uint8_t> temp;
  // Range check and the referenced vertices, counted by a scan (each array has room for a total).
  DeviceArray<int32_t> referenced(n + 1), ref_pos(n + 1), invalid(1);
  launch_zero(referenced.ptr, sizeof(int32_t) * (n + 1), stream);
  launch_zero(invalid.ptr, sizeof(int32_t), stream);
  if (corners > 0)
    check_faces_kernel<<<grid_of(corners), kBlock, 0, stream>>>(faces, corners, n, referenced.ptr,
                                                                 invalid.ptr);
  HG_HIP(hipGetLastError());
  exclusive_scan(referenced.ptr, ref_pos.ptr, n, temp, stream);
  int32_t words[2] = {0, 0};
  HG_HIP(hipMemcpyAsync(&words[0], invalid.ptr, sizeof(int32_t), hipMemcpyDeviceToHost, stream));
  HG_HIP(hipMemcpyAsync(&words[1], ref_pos.ptr + n, sizeof(int32_t), hipMemcpyDeviceToHost,
                        stream));
  HG_HIP(hipStreamSynchronize(stream));
  HG_REQUIRE(words[0] == 0, HG_ERR_VALUE,
             "every face index must be a vertex index in [0, n_vertices)");
  const int64_t count = 2 * corners + words[1];
  const int64_t room = std::max<int64_t>(count, 1);

  // The entries: one stable sort of the keys carrying their slots.
  DeviceArray<uint64_t> keys(room), sorted_keys(room);
  DeviceArray<int32_t> slots(room), sorted_slots(room);
  pattern_keys_kernel<<<grid_of(corners + n), kBlock, 0, stream>>>(faces, n_faces, referenced.ptr,
                                                                    ref_pos.ptr, n, keys.ptr,
                                                                    slots.ptr);
  HG_HIP(hipGetLastError());
  const uint64_t largest = n > 0 ? static_cast<uint64_t>(n) * n - 1 : 0;
  sort_pairs(keys.ptr, sorted_keys.ptr, slots.ptr, sorted_slots.ptr, count, bits_for(largest),
             temp, stream);
  DeviceArray<int32_t> head(count + 1), carries(count + 1), entry(count + 1), position(count + 1);
  launch_zero(head.ptr, sizeof(int32_t) * (count + 1), stream);
  launch_zero(carries.ptr, sizeof(int32_t) * (count + 1), stream);
  if (count > 0)
    pattern_heads_kernel<<<grid_of(count), kBlock, 0, stream>>>(sorted_keys.ptr, sorted_slots.ptr,
                                                                 count, head.ptr, carries.ptr);
  HG_HIP(hipGetLastError());
  inclusive_scan(head.ptr, entry.ptr, count + 1, temp, stream);
  exclusive_scan(carries.ptr, position.ptr, count, temp, stream);
  int32_t sizes[2] = {0, 0};  // entries, slots
  HG_HIP(hipMemcpyAsync(&sizes[0], entry.ptr + count, sizeof(int32_t), hipMemcpyDeviceToHost,
                        stream));
  HG_HIP(hipMemcpyAsync(&sizes[1], position.ptr + count, sizeof(int32_t), hipMemcpyDeviceToHost,
                        stream));
  HG_HIP(hipStreamSynchronize(stream));
  const int64_t nnz = sizes[0], n_slots = sizes[1];
  DeviceArray<int32_t> row_count(n + 1);
  launch_zero(row_count.ptr, sizeof(int32_t) * (n + 1), stream);
  out.columns.allocate(std::max<int64_t>(nnz, 1));
  out.corner_ptr.allocate(nnz + 1);
  out.corner_slots.allocate(std::max<int64_t>(n_slots, 1));
  if (count > 0)
    pattern_entries_kernel<<<grid_of(count), kBlock, 0, stream>>>(
        sorted_keys.ptr, sorted_slots.ptr, head.ptr, entry.ptr, position.ptr, count, n,
        out.columns.ptr, out.corner_ptr.ptr, row_count.ptr, out.corner_slots.ptr);
  set_word_kernel<<<1, 1, 0, stream>>>(out.corner_ptr.ptr + nnz, static_cast<int32_t>(n_slots));
  HG_HIP(hipGetLastError());
  out.offsets.allocate(n + 1);
  exclusive_scan(row_count.ptr, out.offsets.ptr, n, temp, stream);

  // Each vertex's corners: a stable sort of the corner ids by vertex.
  DeviceArray<int32_t> corner_ids(std::max<int64_t>(corners, 1)), vertex_count(n + 1);
  DeviceArray<uint32_t> vertex_of(std::max<int64_t>(corners, 1));
  launch_zero(vertex_count.ptr, sizeof(int32_t) * (n + 1), stream);
  out.vertex_corners.allocate(std::max<int64_t>(corners, 1));
  if (corners > 0)
    vertex_counts_kernel<<<grid_of(corners), kBlock, 0, stream>>>(faces, corners, corner_ids.ptr,
                                                                   vertex_count.ptr);
  HG_HIP(hipGetLastError());
  sort_pairs(reinterpret_cast<const uint32_t*>(faces), vertex_of.ptr, corner_ids.ptr, out.vertex_corners.ptr, corners,
             bits_for(n > 0 ? static_cast<uint64_t>(n - 1) : 0), temp, stream);
  out.vertex_ptr.allocate(n + 1);
  exclusive_scan(vertex_count.ptr, out.vertex_ptr.ptr, n, temp, stream);

  // The host copy, and each row's diagonal entry.
  if (host == nullptr) return;
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
