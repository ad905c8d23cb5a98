// Host analysis of a sparse Cholesky factorization: geometric nested dissection, the supernode
// tree, its level schedule, and the index maps the device factorization and solves read.
//
// Follows ordito's analysis (cholesky.py:881-1273, 1915-2450; kernels/cholesky.py:1663-2577) in
// what it decides, not how: the ordering is geometric nested dissection (split a vertex set at the
// median of its principal axis; the separator is the smaller set of endpoints of the crossing
// edges; sets of at most `leaf_rows` rows are leaves), the top of each tree is merged into one
// supernode while it holds at most `top_rows` rows (CHOLESKY_TOP_ROWS), and each supernode of the
// tree is one dense front of the multifrontal factorization (Liu 1992). The analysis runs once per
// mesh on the host, where ordito runs it on the device; it is not per query.
#include <algorithm>
#include <atomic>
#include <cmath>
#include <functional>
#include <future>
#include <mutex>
#include <numeric>
#include <thread>

#include "hg_internal.h"

namespace hg {
namespace {

struct Node {
  std::vector<int32_t> cols;  // old vertex indices
  std::vector<int32_t> kids;
};

// The two halves of a split share no vertex and no edge (the separator takes every crossing edge's
// endpoint on one side), and neither do two components: their dissections run on their own threads,
// marking disjoint entries of `side` and reading no entry the other marks. Only the node list is
// shared, behind a lock; a node's id is its position there, and the tree (each node's children, in
// order) does not depend on the order the threads append in.
struct Dissector {
  const HostPattern& pattern;
  const float* xyz;
  int32_t leaf_rows;
  std::vector<Node> nodes;
  std::mutex nodes_lock;
  std::vector<int8_t> side;  // 0: outside the current set, 1: part A, 2: part B
  // Splits of at least this many vertices hand one half to a thread of its own, down to this depth
  // (at most 2^depth threads a component).
  static constexpr size_t kThreadSet = 16384;
  static constexpr int kThreadDepth = 4;

  int32_t add(Node node) {
    std::lock_guard<std::mutex> hold(nodes_lock);
    nodes.push_back(std::move(node));
    return static_cast<int32_t>(nodes.size() - 1);
  }

  Dissector(const HostPattern& p, const float* c, int32_t leaf)
      : pattern(p), xyz(c), leaf_rows(leaf), side(static_cast<size_t>(p.n), 0) {}

  // Largest-variance direction of the points of `set` (power iteration on the covariance).
  void principal_axis(const std::vector<int32_t>& set, double axis[3], double center[3]) const {
    double mean[3] = {0, 0, 0};
    for (int32_t v : set)
      for (int d = 0; d < 3; ++d) mean[d] += xyz[3 * v + d];
    for (int d = 0; d < 3; ++d) mean[d] /= static_cast<double>(set.size());
    double cov[3][3] = {{0}};
    for (int32_t v : set) {
      double x[3];
      for (int d = 0; d < 3; ++d) x[d] = xyz[3 * v + d] - mean[d];
      for (int a = 0; a < 3; ++a)
        for (int b = 0; b < 3; ++b) cov[a][b] += x[a] * x[b];
    }
    // Start from the coordinate axis of largest spread, so a degenerate cloud still splits.
    int best = 0;
    for (int d = 1; d < 3; ++d)
      if (cov[d][d] > cov[best][best]) best = d;
    double a[3] = {0, 0, 0};
    a[best] = 1.0;
    for (int it = 0; it < 64; ++it) {
      double b[3];
      for (int r = 0; r < 3; ++r) b[r] = cov[r][0] * a[0] + cov[r][1] * a[1] + cov[r][2] * a[2];
      double norm = std::sqrt(b[0] * b[0] + b[1] * b[1] + b[2] * b[2]);
      if (!(norm > 0.0)) break;
      for (int r = 0; r < 3; ++r) a[r] = b[r] / norm;
    }
    for (int d = 0; d < 3; ++d) {
      axis[d] = a[d];
      center[d] = mean[d];
    }
  }

  // Returns the roots of the dissection forest of `set` (consumed). `depth` counts the splits
  // above that may still hand a half to a thread of its own (0: none).
  std::vector<int32_t> dissect(std::vector<int32_t> set, int depth = 0) {
    std::vector<int32_t> roots;
    if (set.empty()) return roots;
    if (static_cast<int32_t>(set.size()) <= leaf_rows) {
      std::sort(set.begin(), set.end());
      roots.push_back(add(Node{std::move(set), {}}));
      return roots;
    }
    double axis[3], center[3];
    principal_axis(set, axis, center);
    std::vector<std::pair<double, int32_t>> keyed(set.size());
    for (size_t k = 0; k < set.size(); ++k) {
      int32_t v = set[k];
      double proj = (xyz[3 * v] - center[0]) * axis[0] + (xyz[3 * v + 1] - center[1]) * axis[1] +
                    (xyz[3 * v + 2] - center[2]) * axis[2];
      keyed[k] = {proj, v};
    }
    size_t half = keyed.size() / 2;
    std::nth_element(keyed.begin(), keyed.begin() + half, keyed.end());
    for (size_t k = 0; k < keyed.size(); ++k) side[keyed[k].second] = k < half ? 1 : 2;
    // Endpoints of the crossing edges on each side.
    std::vector<int32_t> sep_a, sep_b;
    for (const auto& kv : keyed) {
      int32_t v = kv.second;
      int8_t sv = side[v];
      bool crossing = false;
      for (int32_t e = pattern.offsets[v]; e < pattern.offsets[v + 1]; ++e) {
        int32_t u = pattern.columns[e];
        if (side[u] != 0 && side[u] != sv) {
          crossing = true;
          break;
        }
      }
      if (crossing) (sv == 1 ? sep_a : sep_b).push_back(v);
    }
    bool use_a = sep_a.size() <= sep_b.size();
    std::vector<int32_t>& sep = use_a ? sep_a : sep_b;
    int8_t sep_side = use_a ? 1 : 2;
    for (int32_t v : sep) side[v] = 3;
    std::vector<int32_t> part_a, part_b;
    part_a.reserve(half);
    part_b.reserve(keyed.size() - half);
    for (const auto& kv : keyed) {
      int32_t v = kv.second;
      if (side[v] == 1) part_a.push_back(v);
      else if (side[v] == 2) part_b.push_back(v);
    }
    (void)sep_side;
    for (const auto& kv : keyed) side[kv.second] = 0;
    // Order the separator along its own largest spread, for locality in the fronts above it.
    if (sep.size() > 2) {
      double sep_axis[3], sep_center[3];
      principal_axis(sep, sep_axis, sep_center);
      std::vector<std::pair<double, int32_t>> order(sep.size());
      for (size_t k = 0; k < sep.size(); ++k) {
        int32_t v = sep[k];
        order[k] = {(xyz[3 * v] - sep_center[0]) * sep_axis[0] +
                        (xyz[3 * v + 1] - sep_center[1]) * sep_axis[1] +
                        (xyz[3 * v + 2] - sep_center[2]) * sep_axis[2],
                    v};
      }
      std::sort(order.begin(), order.end());
      for (size_t k = 0; k < sep.size(); ++k) sep[k] = order[k].second;
    }
    std::vector<int32_t> sep_copy = sep;
    std::vector<int32_t> kids_a, kids_b;
    if (depth > 0 && part_a.size() >= kThreadSet && part_b.size() >= kThreadSet) {
      auto half = std::async(std::launch::async, [this, &part_a, depth] {
        return dissect(std::move(part_a), depth - 1);
      });
      kids_b = dissect(std::move(part_b), depth - 1);
      kids_a = half.get();
    } else {
      kids_a = dissect(std::move(part_a), depth > 0 ? depth - 1 : 0);
      kids_b = dissect(std::move(part_b), depth > 0 ? depth - 1 : 0);
    }
    std::vector<int32_t> kids;
    kids.insert(kids.end(), kids_a.begin(), kids_a.end());
    kids.insert(kids.end(), kids_b.begin(), kids_b.end());
    if (sep_copy.empty()) return kids;
    roots.push_back(add(Node{std::move(sep_copy), std::move(kids)}));
    return roots;
  }
};

}  // namespace

Plan analyse(const HostPattern& pattern, const float* coordinates, const AnalysisOptions& options) {
  Plan plan;
  const int64_t n = pattern.n;
  plan.n = n;
  Dissector dissector(pattern, coordinates, std::max(options.leaf_rows, 1));
  // Each connected component is dissected on its own, so a block-diagonal system (several meshes
  // packed into one) gets one tree per block and every block's leaves sit on level 0.
  std::vector<int32_t> roots;
  {
    int32_t n_comp = 0;
    const std::vector<int32_t> label = component_labels(pattern, &n_comp);
    std::vector<std::vector<int32_t>> members(static_cast<size_t>(n_comp));
    for (int64_t v = 0; v < n; ++v) members[label[v]].push_back(static_cast<int32_t>(v));
    // One component: its splits spread over threads. Several: a pool of threads takes them in
    // turn, each dissected on one thread. Either way every component's roots are kept in
    // component order.
    std::vector<std::vector<int32_t>> comp_roots(static_cast<size_t>(n_comp));
    if (n_comp == 1) {
      comp_roots[0] = dissector.dissect(std::move(members[0]), Dissector::kThreadDepth);
    } else if (n_comp > 1) {
      const int32_t workers = std::max<int32_t>(
          1, std::min<int32_t>(n_comp, static_cast<int32_t>(std::thread::hardware_concurrency())));
      std::atomic<int32_t> next_comp{0};
      auto work = [&] {
        for (int32_t c = next_comp++; c < n_comp; c = next_comp++)
          comp_roots[c] = dissector.dissect(std::move(members[c]));
      };
      std::vector<std::future<void>> pool;
      for (int32_t w = 1; w < workers; ++w) pool.push_back(std::async(std::launch::async, work));
      work();
      for (std::future<void>& f : pool) f.get();
    }
    for (const std::vector<int32_t>& r : comp_roots) roots.insert(roots.end(), r.begin(), r.end());
  }
  std::vector<Node>& nodes = dissector.nodes;

  // Merge the top of each tree into its root while the root holds at most top_rows rows.
  for (int32_t root : roots) {
    bool merged = true;
    while (merged) {
      merged = false;
      Node& r = nodes[root];
      for (size_t k = 0; k < r.kids.size(); ++k) {
        int32_t c = r.kids[k];
        if (static_cast<int64_t>(r.cols.size() + nodes[c].cols.size()) <= options.top_rows) {
          std::vector<int32_t> cols = nodes[c].cols;
          cols.insert(cols.end(), r.cols.begin(), r.cols.end());
          std::vector<int32_t> kids(r.kids.begin(), r.kids.begin() + k);
          kids.insert(kids.end(), nodes[c].kids.begin(), nodes[c].kids.end());
          kids.insert(kids.end(), r.kids.begin() + k + 1, r.kids.end());
          r.cols = std::move(cols);
          r.kids = std::move(kids);
          nodes[c].cols.clear();
          nodes[c].kids.clear();
          merged = true;
          break;
        }
      }
    }
  }

  // Postorder numbering of the supernodes; columns numbered consecutively in that order.
  std::vector<int32_t> post;  // node ids in postorder
  post.reserve(nodes.size());
  {
    std::vector<std::pair<int32_t, size_t>> stack;
    for (int32_t root : roots) {
      stack.push_back({root, 0});
      while (!stack.empty()) {
        auto& top = stack.back();
        if (top.second < nodes[top.first].kids.size()) {
          int32_t child = nodes[top.first].kids[top.second++];
          stack.push_back({child, 0});
        } else {
          post.push_back(top.first);
          stack.pop_back();
        }
      }
    }
  }
  const int32_t S = static_cast<int32_t>(post.size());
  plan.n_super = S;
  std::vector<int32_t> super_of_node(nodes.size(), -1);
  for (int32_t s = 0; s < S; ++s) super_of_node[post[s]] = s;
  plan.perm.resize(static_cast<size_t>(n));
  plan.iperm.resize(static_cast<size_t>(n));
  plan.col_begin.resize(S);
  plan.ncol.resize(S);
  plan.parent.assign(S, -1);
  plan.level.assign(S, 0);
  std::vector<int32_t> super_of_col(static_cast<size_t>(n));
  int32_t next = 0;
  for (int32_t s = 0; s < S; ++s) {
    const Node& node = nodes[post[s]];
    plan.col_begin[s] = next;
    plan.ncol[s] = static_cast<int32_t>(node.cols.size());
    for (int32_t v : node.cols) {
      plan.perm[next] = v;
      plan.iperm[v] = next;
      super_of_col[next] = s;
      ++next;
    }
    for (int32_t kid : node.kids) plan.parent[super_of_node[kid]] = s;
  }
  HG_REQUIRE(next == n, HG_ERR_RUNTIME, "nested dissection lost vertices");
  // Children lists (ascending) and levels (height above the leaves).
  plan.child_ptr.assign(S + 1, 0);
  for (int32_t s = 0; s < S; ++s)
    if (plan.parent[s] >= 0) plan.child_ptr[plan.parent[s] + 1]++;
  std::partial_sum(plan.child_ptr.begin(), plan.child_ptr.end(), plan.child_ptr.begin());
  plan.children.resize(plan.child_ptr[S]);
  {
    std::vector<int32_t> at(plan.child_ptr.begin(), plan.child_ptr.end() - 1);
    for (int32_t s = 0; s < S; ++s)
      if (plan.parent[s] >= 0) plan.children[at[plan.parent[s]]++] = s;
  }
  int32_t max_level = 0;
  for (int32_t s = 0; s < S; ++s) {
    int32_t lv = 0;
    for (int32_t k = plan.child_ptr[s]; k < plan.child_ptr[s + 1]; ++k)
      lv = std::max(lv, plan.level[plan.children[k]] + 1);
    plan.level[s] = lv;
    max_level = std::max(max_level, lv);
  }
  plan.n_levels = S > 0 ? max_level + 1 : 0;
  plan.level_ptr.assign(plan.n_levels + 1, 0);
  for (int32_t s = 0; s < S; ++s) plan.level_ptr[plan.level[s] + 1]++;
  std::partial_sum(plan.level_ptr.begin(), plan.level_ptr.end(), plan.level_ptr.begin());
  plan.level_nodes.resize(S);
  {
    std::vector<int32_t> at(plan.level_ptr.begin(), plan.level_ptr.end() - 1);
    for (int32_t s = 0; s < S; ++s) plan.level_nodes[at[plan.level[s]]++] = s;
  }

  // Symbolic structure: rows(s) = (adjacency of cols(s) past s) U (rows(children) past s). The
  // supernodes of a level are independent (their children sit on lower levels): each level's are
  // formed on threads, each thread with its own stamps.
  plan.row_ptr.assign(S + 1, 0);
  plan.nrow.assign(S, 0);
  std::vector<std::vector<int32_t>> rows(S);
  const int64_t hw_threads = std::max<int64_t>(1, std::thread::hardware_concurrency());
  std::vector<std::vector<int32_t>> stamps(static_cast<size_t>(hw_threads));
  for (int32_t lv = 0; lv < plan.n_levels; ++lv) {
    const int32_t l0 = plan.level_ptr[lv], l1 = plan.level_ptr[lv + 1];
    parallel_chunks(l1 - l0, 16, [&](int64_t b, int64_t e, int64_t worker) {
      std::vector<int32_t>& stamp = stamps[worker];
      if (stamp.empty()) stamp.assign(static_cast<size_t>(n), -1);
      for (int64_t q = b; q < e; ++q) {
        const int32_t s = plan.level_nodes[l0 + q];
        const int32_t first = plan.col_begin[s];
        const int32_t last = first + plan.ncol[s] - 1;
        std::vector<int32_t>& list = rows[s];
        for (int32_t c = first; c <= last; ++c) {
          int32_t v = plan.perm[c];
          for (int32_t e2 = pattern.offsets[v]; e2 < pattern.offsets[v + 1]; ++e2) {
            int32_t u = plan.iperm[pattern.columns[e2]];
            if (u > last && stamp[u] != s) {
              stamp[u] = s;
              list.push_back(u);
            }
          }
        }
        for (int32_t k = plan.child_ptr[s]; k < plan.child_ptr[s + 1]; ++k) {
          int32_t child = plan.children[k];
          for (int32_t r : rows[child]) {
            HG_REQUIRE(r >= first, HG_ERR_RUNTIME, "supernode tree is not an assembly tree");
            if (r > last && stamp[r] != s) {
              stamp[r] = s;
              list.push_back(r);
            }
          }
        }
        std::sort(list.begin(), list.end());
        plan.nrow[s] = static_cast<int32_t>(list.size());
      }
    });
  }
  for (int32_t s = 0; s < S; ++s) plan.row_ptr[s + 1] = plan.row_ptr[s] + plan.nrow[s];
  plan.rows.resize(plan.row_ptr[S]);
  for (int32_t s = 0; s < S; ++s)
    std::copy(rows[s].begin(), rows[s].end(), plan.rows.begin() + plan.row_ptr[s]);

  // Each row's position in its parent's front [cols(p); rows(p)].
  plan.rel_ptr = plan.row_ptr;
  plan.rel.resize(plan.rows.size());
  for (int32_t s = 0; s < S; ++s) {
    int32_t p = plan.parent[s];
    if (p < 0) {
      HG_REQUIRE(plan.nrow[s] == 0, HG_ERR_RUNTIME, "a root supernode has structure rows");
      continue;
    }
    const int32_t pfirst = plan.col_begin[p], plast = pfirst + plan.ncol[p] - 1;
    const int32_t* prow = plan.rows.data() + plan.row_ptr[p];
    for (int32_t k = 0; k < plan.nrow[s]; ++k) {
      int32_t r = plan.rows[plan.row_ptr[s] + k];
      int32_t pos;
      if (r <= plast) {
        pos = r - pfirst;
      } else {
        const int32_t* it = std::lower_bound(prow, prow + plan.nrow[p], r);
        HG_REQUIRE(it != prow + plan.nrow[p] && *it == r, HG_ERR_RUNTIME,
                   "child structure row missing from its parent's front");
        pos = plan.ncol[p] + static_cast<int32_t>(it - prow);
      }
      plan.rel[plan.row_ptr[s] + k] = pos;
    }
  }

  // Arena, solve-block and vector offsets; flops.
  plan.front_off.resize(S);
  plan.w_off.resize(S);
  plan.vec_off.resize(S);
  int64_t arena = 0, w = 0, vec = 0;
  double flops = 0.0;
  for (int32_t s = 0; s < S; ++s) {
    int64_t m = plan.ncol[s] + plan.nrow[s];
    plan.front_off[s] = arena;
    plan.w_off[s] = w;
    plan.vec_off[s] = vec;
    arena += m * m;
    w += m * plan.ncol[s];
    vec += m;
    double c = plan.ncol[s], r = plan.nrow[s];
    flops += c * c * c / 3.0 + c * c * r + c * r * r;
    plan.max_front = std::max<int32_t>(plan.max_front, static_cast<int32_t>(m));
  }
  plan.arena_size = arena;
  plan.w_size = w;
  plan.vec_size = vec;
  plan.flops = flops;
  plan.nnz_w = w;

  // Lower-triangle operator entries in front coordinates, in operator order: each thread forms a
  // contiguous range of rows' into its own lists, appended in range order.
  {
    std::vector<std::vector<int32_t>> src(static_cast<size_t>(hw_threads)),
        erow(static_cast<size_t>(hw_threads)), ecol(static_cast<size_t>(hw_threads));
    std::vector<std::vector<int64_t>> dst(static_cast<size_t>(hw_threads));
    parallel_chunks(n, 4096, [&](int64_t b, int64_t e_end, int64_t worker) {
      // Local lists, moved out at the end: the workers' list headers sit side by side, and
      // appending to them in place would share their cache lines.
      std::vector<int32_t> my_src, my_row, my_col;
      std::vector<int64_t> my_dst;
      for (int64_t i = b; i < e_end; ++i) {
        const int32_t ni = plan.iperm[i];
        for (int32_t e = pattern.offsets[i]; e < pattern.offsets[i + 1]; ++e) {
          const int32_t nj = plan.iperm[pattern.columns[e]];
          if (ni < nj) continue;
          const int32_t s = super_of_col[nj];
          const int64_t m = plan.ncol[s] + plan.nrow[s];
          int64_t rpos;
          const int32_t first = plan.col_begin[s], last = first + plan.ncol[s] - 1;
          if (ni <= last) {
            rpos = ni - first;
          } else {
            const int32_t* srow = plan.rows.data() + plan.row_ptr[s];
            const int32_t* it = std::lower_bound(srow, srow + plan.nrow[s], ni);
            HG_REQUIRE(it != srow + plan.nrow[s] && *it == ni, HG_ERR_RUNTIME,
                       "operator entry outside its supernode's structure");
            rpos = plan.ncol[s] + (it - srow);
          }
          const int64_t cpos = nj - first;
          my_src.push_back(e);
          my_dst.push_back(plan.front_off[s] + cpos * m + rpos);
          my_row.push_back(ni);
          my_col.push_back(nj);
        }
      }
      src[worker] = std::move(my_src);
      dst[worker] = std::move(my_dst);
      erow[worker] = std::move(my_row);
      ecol[worker] = std::move(my_col);
    });
    for (int64_t w = 0; w < hw_threads; ++w) {
      plan.entry_src.insert(plan.entry_src.end(), src[w].begin(), src[w].end());
      plan.entry_dst.insert(plan.entry_dst.end(), dst[w].begin(), dst[w].end());
      plan.entry_row.insert(plan.entry_row.end(), erow[w].begin(), erow[w].end());
      plan.entry_col.insert(plan.entry_col.end(), ecol[w].begin(), ecol[w].end());
    }
  }
  return plan;
}

}  // namespace hg
