// Host construction of the cotangent operator's sparsity pattern and its corner runs.
//
// The pattern ordito builds on the device from sorted keys (laplacian.py:1071-1389,
// kernels/laplacian.py:707-760): an entry (i, j) for both directions of every face edge, a diagonal
// for every referenced vertex, rows sorted by column, and for each off-diagonal entry the run of
// corner slots 3f+e whose opposite edge it is, in ascending slot order (the stable sort's order,
// the order its half-cotangents are summed in, kernels/laplacian.py:970-989). Built here by a
// counting sort over rows, once per mesh.
#include <algorithm>
#include <numeric>

#include "hg_internal.h"

namespace hg {

HostPattern build_pattern(int64_t n_vertices, const int32_t* faces, int64_t n_faces) {
  HostPattern p;
  p.n = n_vertices;
  const int64_t n = n_vertices;
  for (int64_t q = 0; q < 3 * n_faces; ++q) {
    HG_REQUIRE(faces[q] >= 0 && faces[q] < n, HG_ERR_VALUE,
               "every face index must be a vertex index in [0, n_vertices)");
  }
  std::vector<uint8_t> referenced(static_cast<size_t>(n), 0);
  for (int64_t q = 0; q < 3 * n_faces; ++q) referenced[faces[q]] = 1;

  // Directed entries per row: (column, slot); slot -1 marks the diagonal.
  std::vector<int64_t> count(static_cast<size_t>(n) + 1, 0);
  for (int64_t f = 0; f < n_faces; ++f) {
    for (int e = 0; e < 3; ++e) {
      int32_t i = faces[3 * f + (e + 1) % 3];
      int32_t j = faces[3 * f + (e + 2) % 3];
      count[i + 1]++;
      count[j + 1]++;
    }
  }
  for (int64_t v = 0; v < n; ++v) count[v + 1] += referenced[v];
  std::partial_sum(count.begin(), count.end(), count.begin());
  const int64_t total = count[n];
  std::vector<int64_t> fill(count.begin(), count.end() - 1);
  std::vector<int64_t> keys(static_cast<size_t>(total));  // (column << 32) | (slot + 1)
  for (int64_t f = 0; f < n_faces; ++f) {
    for (int e = 0; e < 3; ++e) {
      int32_t i = faces[3 * f + (e + 1) % 3];
      int32_t j = faces[3 * f + (e + 2) % 3];
      int64_t slot = 3 * f + e + 1;
      keys[fill[i]++] = (static_cast<int64_t>(j) << 32) | slot;
      keys[fill[j]++] = (static_cast<int64_t>(i) << 32) | slot;
    }
  }
  for (int64_t v = 0; v < n; ++v) {
    if (referenced[v]) keys[fill[v]++] = (static_cast<int64_t>(v) << 32);
  }

  p.offsets.assign(static_cast<size_t>(n) + 1, 0);
  p.diagonal.assign(static_cast<size_t>(n), -1);
  p.columns.reserve(static_cast<size_t>(total / 2 + n));
  p.corner_ptr.reserve(static_cast<size_t>(total / 2 + n) + 1);
  p.corner_slots.reserve(static_cast<size_t>(6 * n_faces));
  p.corner_ptr.push_back(0);
  for (int64_t v = 0; v < n; ++v) {
    auto begin = keys.begin() + count[v];
    auto end = keys.begin() + count[v + 1];
    std::sort(begin, end);
    int64_t previous = -1;
    for (auto it = begin; it != end; ++it) {
      int64_t column = *it >> 32;
      int64_t slot = (*it & 0xffffffffLL) - 1;
      if (column != previous) {
        if (previous >= 0) p.corner_ptr.push_back(static_cast<int32_t>(p.corner_slots.size()));
        if (column == v) p.diagonal[v] = static_cast<int32_t>(p.columns.size());
        p.columns.push_back(static_cast<int32_t>(column));
        previous = column;
      }
      if (slot >= 0) p.corner_slots.push_back(static_cast<int32_t>(slot));
    }
    if (previous >= 0) p.corner_ptr.push_back(static_cast<int32_t>(p.corner_slots.size()));
    p.offsets[v + 1] = static_cast<int32_t>(p.columns.size());
  }
  HG_REQUIRE(p.columns.size() < (1ull << 31), HG_ERR_VALUE, "pattern too large for int32 offsets");

  // Each vertex's corners in ascending slot order.
  p.vertex_ptr.assign(static_cast<size_t>(n) + 1, 0);
  for (int64_t q = 0; q < 3 * n_faces; ++q) p.vertex_ptr[faces[q] + 1]++;
  std::partial_sum(p.vertex_ptr.begin(), p.vertex_ptr.end(), p.vertex_ptr.begin());
  p.vertex_corners.resize(static_cast<size_t>(3 * n_faces));
  std::vector<int32_t> at(p.vertex_ptr.begin(), p.vertex_ptr.end() - 1);
  for (int64_t q = 0; q < 3 * n_faces; ++q) p.vertex_corners[at[faces[q]]++] = static_cast<int32_t>(q);
  return p;
}

std::vector<int32_t> component_labels(const HostPattern& pattern, int32_t* n_components) {
  const int64_t n = pattern.n;
  std::vector<int32_t> label(static_cast<size_t>(n), -1);
  std::vector<int32_t> stack;
  int32_t next = 0;
  for (int64_t s = 0; s < n; ++s) {
    if (label[s] >= 0) continue;
    label[s] = next;
    stack.push_back(static_cast<int32_t>(s));
    while (!stack.empty()) {
      int32_t v = stack.back();
      stack.pop_back();
      for (int32_t k = pattern.offsets[v]; k < pattern.offsets[v + 1]; ++k) {
        int32_t u = pattern.columns[k];
        if (label[u] < 0) {
          label[u] = next;
          stack.push_back(u);
        }
      }
    }
    ++next;
  }
  *n_components = next;
  return label;
}

}  // namespace hg
