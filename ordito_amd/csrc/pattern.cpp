// Connected components of the operator pattern's graph, on the host (the pattern itself is built
// on the device, pattern.hip).
#include <vector>

#include "hg_internal.h"

namespace hg {

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
