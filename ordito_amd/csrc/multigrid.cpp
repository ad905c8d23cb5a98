// Smoothed-aggregation multigrid hierarchy of a symmetric positive semi-definite operator, built on
// the host (ordito's multigrid_preconditioner, linalg.py:5020, 6284-7500): strength-of-connection
// aggregation, a piecewise-constant tentative prolongator smoothed by one damped l1-Jacobi step,
// Galerkin coarse operators, and a dense generalized inverse at the bottom. The method is Vanek,
// Mandel and Brezina's (Computing 56, 1996); ordito's policy numbers (strength threshold, damping,
// sweeps, coarse size, level cap, stall ratio) are kept. The aggregation here is the paper's serial
// three-phase greedy one in row order (ordito aggregates by a randomized parallel independent set):
// deterministic, and an aggregation is a choice of the coarse space, not of the answer.
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <numeric>
#include <vector>

#include "multigrid.h"

namespace hg {
namespace {

// ordito's policy (linalg.py:6284-6365).
constexpr int64_t kMaxCoarse = 384;      // a level this small is inverted densely
constexpr int32_t kMaxLevels = 12;
constexpr double kMinCoarsening = 0.9;   // a level must shrink to below this fraction
constexpr int64_t kMaxDense = 1024;      // past this the hierarchy has no usable bottom
constexpr double kTheta = 0.05;          // strength of connection
constexpr double kJacobiFactor = 4.0 / 3.0;
constexpr int32_t kPowerSteps = 8;

// y = A x for a host CSR.
void multiply(const HostCsr& a, const std::vector<double>& x, std::vector<double>& y) {
  y.assign(a.rows, 0.0);
  for (int64_t i = 0; i < a.rows; ++i) {
    double sum = 0.0;
    for (int32_t e = a.ptr[i]; e < a.ptr[i + 1]; ++e) sum += a.val[e] * x[a.idx[e]];
    y[i] = sum;
  }
}

// C = A B (Gustavson's row-by-row product with a dense accumulator; columns of each row sorted).
HostCsr product(const HostCsr& a, const HostCsr& b) {
  HostCsr c;
  c.rows = a.rows;
  c.cols = b.cols;
  c.ptr.assign(a.rows + 1, 0);
  std::vector<double> acc(b.cols, 0.0);
  std::vector<int32_t> mark(b.cols, -1), touched;
  for (int64_t i = 0; i < a.rows; ++i) {
    touched.clear();
    for (int32_t e = a.ptr[i]; e < a.ptr[i + 1]; ++e) {
      const int32_t k = a.idx[e];
      const double av = a.val[e];
      for (int32_t f = b.ptr[k]; f < b.ptr[k + 1]; ++f) {
        const int32_t j = b.idx[f];
        if (mark[j] != static_cast<int32_t>(i)) {
          mark[j] = static_cast<int32_t>(i);
          acc[j] = 0.0;
          touched.push_back(j);
        }
        acc[j] += av * b.val[f];
      }
    }
    std::sort(touched.begin(), touched.end());
    for (int32_t j : touched) {
      c.idx.push_back(j);
      c.val.push_back(acc[j]);
    }
    c.ptr[i + 1] = static_cast<int32_t>(c.idx.size());
  }
  return c;
}

HostCsr transpose(const HostCsr& a) {
  HostCsr t;
  t.rows = a.cols;
  t.cols = a.rows;
  t.ptr.assign(a.cols + 1, 0);
  for (int32_t j : a.idx) ++t.ptr[j + 1];
  for (int64_t j = 0; j < a.cols; ++j) t.ptr[j + 1] += t.ptr[j];
  t.idx.resize(a.idx.size());
  t.val.resize(a.val.size());
  std::vector<int32_t> next(t.ptr.begin(), t.ptr.end() - 1);
  for (int64_t i = 0; i < a.rows; ++i)
    for (int32_t e = a.ptr[i]; e < a.ptr[i + 1]; ++e) {
      const int32_t at = next[a.idx[e]]++;
      t.idx[at] = static_cast<int32_t>(i);
      t.val[at] = a.val[e];
    }
  return t;
}

std::vector<double> diagonal(const HostCsr& a) {
  std::vector<double> d(a.rows, 0.0);
  for (int64_t i = 0; i < a.rows; ++i)
    for (int32_t e = a.ptr[i]; e < a.ptr[i + 1]; ++e)
      if (a.idx[e] == i) d[i] += a.val[e];
  return d;
}

// Each row's aggregate (-1: none, an empty row): the three phases of Vanek, Mandel and Brezina
// over the strong graph |A_ij| >= theta sqrt(|A_ii A_jj|). Returns the aggregate count.
int64_t aggregate(const HostCsr& a, const std::vector<double>& d, std::vector<int64_t>& agg) {
  const int64_t n = a.rows;
  auto strong = [&](int64_t i, int32_t e) {
    const int32_t j = a.idx[e];
    return j != i && std::fabs(a.val[e]) >= kTheta * std::sqrt(std::fabs(d[i] * d[j])) &&
           a.val[e] != 0.0;
  };
  agg.assign(n, -1);
  int64_t count = 0;
  // Phase 1: a row whose strong neighbours are all free starts an aggregate with them.
  for (int64_t i = 0; i < n; ++i) {
    if (agg[i] != -1 || d[i] == 0.0) continue;
    bool free = true;
    for (int32_t e = a.ptr[i]; e < a.ptr[i + 1] && free; ++e)
      if (strong(i, e) && agg[a.idx[e]] != -1) free = false;
    if (!free) continue;
    agg[i] = count;
    for (int32_t e = a.ptr[i]; e < a.ptr[i + 1]; ++e)
      if (strong(i, e)) agg[a.idx[e]] = count;
    ++count;
  }
  // Phase 2: a free row joins the phase-1 aggregate it is most strongly connected to.
  const std::vector<int64_t> first = agg;
  for (int64_t i = 0; i < n; ++i) {
    if (agg[i] != -1 || d[i] == 0.0) continue;
    double best = -1.0;
    for (int32_t e = a.ptr[i]; e < a.ptr[i + 1]; ++e) {
      if (!strong(i, e) || first[a.idx[e]] < 0) continue;
      const double w = std::fabs(a.val[e]);
      if (w > best) {
        best = w;
        agg[i] = first[a.idx[e]];
      }
    }
  }
  // Phase 3: the rows still free aggregate with their free strong neighbours.
  for (int64_t i = 0; i < n; ++i) {
    if (agg[i] != -1 || d[i] == 0.0) continue;
    agg[i] = count;
    for (int32_t e = a.ptr[i]; e < a.ptr[i + 1]; ++e)
      if (strong(i, e) && agg[a.idx[e]] == -1) agg[a.idx[e]] = count;
    ++count;
  }
  return count;
}

// The l1 row sums sum_j |A_ij| (the smoother's scaling; ordito's l1_smoother).
std::vector<double> l1_rows(const HostCsr& a) {
  std::vector<double> s(a.rows, 0.0);
  for (int64_t i = 0; i < a.rows; ++i)
    for (int32_t e = a.ptr[i]; e < a.ptr[i + 1]; ++e) s[i] += std::fabs(a.val[e]);
  return s;
}

// The spectral radius of D^-1 A by an unnormalized power iteration from a seeded vector: the
// growth over kPowerSteps steps.
double spectral_radius(const HostCsr& a, const std::vector<double>& dinv, uint64_t seed) {
  std::vector<double> v(a.rows), w;
  uint64_t s = seed * 0x9E3779B97F4A7C15ull + 0x2545F4914F6CDD1Dull;
  for (double& x : v) {
    s ^= s << 13;
    s ^= s >> 7;
    s ^= s << 17;
    x = static_cast<double>(s >> 11) * 0x1.0p-53 - 0.5;
  }
  double n0 = 0.0;
  for (double x : v) n0 += x * x;
  if (n0 == 0.0) return 1.0;
  double scale = 0.0;  // log of the accumulated growth
  for (int32_t k = 0; k < kPowerSteps; ++k) {
    multiply(a, v, w);
    double norm = 0.0;
    for (int64_t i = 0; i < a.rows; ++i) {
      w[i] *= dinv[i];
      norm += w[i] * w[i];
    }
    double prev = 0.0;
    for (double x : v) prev += x * x;
    if (norm == 0.0 || prev == 0.0) return 1.0;
    scale += 0.5 * std::log(norm / prev);
    v.swap(w);
  }
  const double rho = std::exp(scale / kPowerSteps);
  return rho > 0.0 ? rho : 1.0;
}

// A generalized inverse G = L^-T D^+ L^-1 of the dense symmetric positive semi-definite matrix
// from its LDL^T with the pivots below 1e-12 of the largest diagonal dropped (the null directions:
// each component's constant): symmetric, and exact on every consistent right-hand side.
std::vector<double> generalized_inverse(const HostCsr& a) {
  const int64_t n = a.rows;
  std::vector<double> m(n * n, 0.0);
  for (int64_t i = 0; i < n; ++i)
    for (int32_t e = a.ptr[i]; e < a.ptr[i + 1]; ++e) m[i * n + a.idx[e]] = a.val[e];
  // Symmetrize against product round-off.
  for (int64_t i = 0; i < n; ++i)
    for (int64_t j = 0; j < i; ++j) {
      const double s = 0.5 * (m[i * n + j] + m[j * n + i]);
      m[i * n + j] = m[j * n + i] = s;
    }
  double top = 0.0;
  for (int64_t i = 0; i < n; ++i) top = std::max(top, std::fabs(m[i * n + i]));
  const double drop = 1e-12 * top;
  std::vector<double> l(n * n, 0.0), dg(n, 0.0);
  for (int64_t k = 0; k < n; ++k) {
    double dk = m[k * n + k];
    for (int64_t j = 0; j < k; ++j) dk -= l[k * n + j] * l[k * n + j] * dg[j];
    l[k * n + k] = 1.0;
    if (!(dk > drop)) {
      dg[k] = 0.0;
      continue;
    }
    dg[k] = dk;
    for (int64_t i = k + 1; i < n; ++i) {
      double v = m[i * n + k];
      for (int64_t j = 0; j < k; ++j) v -= l[i * n + j] * l[k * n + j] * dg[j];
      l[i * n + k] = v / dk;
    }
  }
  std::vector<double> g(n * n, 0.0), y(n);
  for (int64_t c = 0; c < n; ++c) {
    // L y = e_c, y /= D (dropped: 0), L^T x = y.
    std::fill(y.begin(), y.end(), 0.0);
    y[c] = 1.0;
    for (int64_t i = c + 1; i < n; ++i) {
      double v = 0.0;
      for (int64_t j = c; j < i; ++j) v += l[i * n + j] * y[j];
      y[i] = -v;
    }
    for (int64_t i = 0; i < n; ++i) y[i] = dg[i] > 0.0 ? y[i] / dg[i] : 0.0;
    for (int64_t i = n - 1; i >= 0; --i) {
      double v = y[i];
      for (int64_t j = i + 1; j < n; ++j) v -= l[j * n + i] * g[j * n + c];
      g[i * n + c] = v;
    }
  }
  return g;
}

}  // namespace

HostHierarchy build_hierarchy(const HostCsr& fine, uint64_t seed) {
  HostHierarchy h;
  HostCsr a = fine;
  for (int32_t level = 0;; ++level) {
    HostLevel lev;
    const std::vector<double> l1 = l1_rows(a);
    std::vector<double> dinv(a.rows);
    for (int64_t i = 0; i < a.rows; ++i) dinv[i] = l1[i] > 0.0 ? 1.0 / l1[i] : 0.0;
    const double omega = kJacobiFactor / spectral_radius(a, dinv, seed + level);
    lev.smoother.resize(a.rows);
    for (int64_t i = 0; i < a.rows; ++i) lev.smoother[i] = omega * dinv[i];
    const bool bottom = a.rows <= kMaxCoarse || level + 1 >= kMaxLevels;
    std::vector<int64_t> agg;
    int64_t n_agg = 0;
    if (!bottom) n_agg = aggregate(a, diagonal(a), agg);
    if (bottom || n_agg == 0 || static_cast<double>(n_agg) > kMinCoarsening * a.rows) {
      // The bottom: inverted densely, or no usable hierarchy.
      if (a.rows > kMaxDense) return HostHierarchy{};
      h.coarse_n = a.rows;
      h.coarse_inverse = generalized_inverse(a);
      lev.a = std::move(a);
      h.levels.push_back(std::move(lev));
      h.usable = true;
      return h;
    }
    // Tentative prolongator: each aggregate's indicator, unit 2-norm.
    std::vector<int64_t> size(n_agg, 0);
    for (int64_t i = 0; i < a.rows; ++i)
      if (agg[i] >= 0) ++size[agg[i]];
    HostCsr t;
    t.rows = a.rows;
    t.cols = n_agg;
    t.ptr.assign(a.rows + 1, 0);
    for (int64_t i = 0; i < a.rows; ++i) {
      if (agg[i] >= 0) {
        t.idx.push_back(static_cast<int32_t>(agg[i]));
        t.val.push_back(1.0 / std::sqrt(static_cast<double>(size[agg[i]])));
      }
      t.ptr[i + 1] = static_cast<int32_t>(t.idx.size());
    }
    // P = (I - omega D_l1^-1 A) T.
    HostCsr at = product(a, t);
    HostCsr p;
    p.rows = a.rows;
    p.cols = n_agg;
    p.ptr.assign(a.rows + 1, 0);
    for (int64_t i = 0; i < a.rows; ++i) {
      const double w = lev.smoother[i];
      // Merge row i of T (one entry at most) with -w (A T)_i, both sorted.
      int32_t e = at.ptr[i];
      const int32_t end = at.ptr[i + 1];
      const bool has_t = t.ptr[i + 1] > t.ptr[i];
      const int32_t tj = has_t ? t.idx[t.ptr[i]] : -1;
      bool placed = !has_t;
      for (; e < end; ++e) {
        const int32_t j = at.idx[e];
        if (!placed && tj < j) {
          p.idx.push_back(tj);
          p.val.push_back(t.val[t.ptr[i]]);
          placed = true;
        }
        double v = -w * at.val[e];
        if (j == tj) {
          v += t.val[t.ptr[i]];
          placed = true;
        }
        p.idx.push_back(j);
        p.val.push_back(v);
      }
      if (!placed) {
        p.idx.push_back(tj);
        p.val.push_back(t.val[t.ptr[i]]);
      }
      p.ptr[i + 1] = static_cast<int32_t>(p.idx.size());
    }
    HostCsr r = transpose(p);
    HostCsr coarse = product(r, product(a, p));
    lev.a = std::move(a);
    lev.p = std::move(p);
    lev.r = std::move(r);
    h.levels.push_back(std::move(lev));
    a = std::move(coarse);
  }
}

}  // namespace hg
