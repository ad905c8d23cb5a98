// Assembly of the heat method's operators and the per-face / per-vertex passes of a query:
// face terms, lumped mass, the time step's edge sums, the fused row kernel (L, M - tL, -L), the
// source seeding, the unit-gradient divergence and the source shift.
//
// Atomic-free: every face -> vertex sum is a gather over the vertex's corners in ascending corner
// order, and every reduction a fixed tree, so a call is bit-reproducible. Multiply-add
// contraction is off in this file: the float32 face terms and the float64 sums are formed one
// rounding per operation, exactly the arithmetic the reference's kernels spell.
#pragma clang fp contract(off)

#include <algorithm>
#include <cmath>
#include <vector>

#include "kernels.h"

namespace hg {
namespace {

struct F3 {
  float x, y, z;
};

__device__ __forceinline__ F3 load3(const float* p, int64_t v) {
  return F3{p[3 * v], p[3 * v + 1], p[3 * v + 2]};
}
__device__ __forceinline__ F3 sub(F3 a, F3 b) { return F3{a.x - b.x, a.y - b.y, a.z - b.z}; }
__device__ __forceinline__ F3 neg(F3 a) { return F3{-a.x, -a.y, -a.z}; }
__device__ __forceinline__ float dot(F3 a, F3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
__device__ __forceinline__ F3 cross(F3 u, F3 v) {
  return F3{u.y * v.z - u.z * v.y, u.z * v.x - u.x * v.z, u.x * v.y - u.y * v.x};
}

struct D3 {
  double x, y, z;
};
__device__ __forceinline__ D3 widen(F3 a) { return D3{a.x, a.y, a.z}; }
__device__ __forceinline__ D3 sub(D3 a, D3 b) { return D3{a.x - b.x, a.y - b.y, a.z - b.z}; }
__device__ __forceinline__ D3 add(D3 a, D3 b) { return D3{a.x + b.x, a.y + b.y, a.z + b.z}; }
__device__ __forceinline__ D3 scale(double s, D3 a) { return D3{s * a.x, s * a.y, s * a.z}; }
__device__ __forceinline__ double dot(D3 a, D3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
__device__ __forceinline__ D3 cross(D3 u, D3 v) {
  return D3{u.y * v.z - u.z * v.y, u.z * v.x - u.x * v.z, u.x * v.y - u.y * v.x};
}

// cross(b - a, c - a) formed at the corner opposite the longest edge (the largest angle), ties to
// corner a then b: ordito kernels/predicates.py:106-130.
__device__ __forceinline__ F3 area_vector(F3 a, F3 b, F3 c) {
  F3 ab = sub(b, a), bc = sub(c, b), ca = sub(a, c);
  float l_ab = dot(ab, ab), l_bc = dot(bc, bc), l_ca = dot(ca, ca);
  if (l_bc >= l_ca && l_bc >= l_ab) return cross(ab, neg(ca));
  if (l_ca >= l_ab) return cross(bc, neg(ab));
  return cross(ca, neg(bc));
}

}  // namespace

// Per face: float32 half-cotangents (dot / (2 |cross|), kernels/laplacian.py:93-106), unit normal
// and area (kernels/triangles.py:332-358). One thread a face, 48 bytes read, 28 written.
__global__ void face_terms_kernel(const float* __restrict__ xyz, const int32_t* __restrict__ faces,
                                  int64_t n_faces, float* __restrict__ cot,
                                  float* __restrict__ normals, float* __restrict__ areas) {
  int64_t f = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (f >= n_faces) return;
  F3 v0 = load3(xyz, faces[3 * f]), v1 = load3(xyz, faces[3 * f + 1]),
     v2 = load3(xyz, faces[3 * f + 2]);
  F3 c = area_vector(v0, v1, v2);
  float norm = sqrtf(dot(c, c));
  F3 n = c;
  if (norm > 0.0f) n = F3{c.x / norm, c.y / norm, c.z / norm};
  normals[3 * f] = n.x;
  normals[3 * f + 1] = n.y;
  normals[3 * f + 2] = n.z;
  areas[f] = 0.5f * norm;
  float denominator = 2.0f * norm;
  float c0 = 0.0f, c1 = 0.0f, c2 = 0.0f;
  if (denominator > 0.0f) {
    c0 = dot(sub(v1, v0), sub(v2, v0)) / denominator;
    c1 = dot(sub(v2, v1), sub(v0, v1)) / denominator;
    c2 = dot(sub(v0, v2), sub(v1, v2)) / denominator;
  }
  cot[3 * f] = c0;
  cot[3 * f + 1] = c1;
  cot[3 * f + 2] = c2;
}

// use_robust's first pass (laplacian.py:779-874, kernels/laplacian.py:1585-1630): per face the
// float32 side lengths, column e opposite corner e (|v1 - v2|, |v2 - v0|, |v0 - v1|), and per
// block of faces the smallest triangle-inequality margin a + b - c over its faces and sides and
// the float64 sum of its lengths, added in a fixed tree.
__global__ __launch_bounds__(256) void face_lengths_kernel(
    const float* __restrict__ xyz, const int32_t* __restrict__ faces, int64_t n_faces,
    float* __restrict__ lengths, float* __restrict__ block_margin,
    double* __restrict__ block_sum) {
  __shared__ float low[256];
  __shared__ double total[256];
  const int64_t f = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  float margin = INFINITY;
  double sum = 0.0;
  if (f < n_faces) {
    const F3 v0 = load3(xyz, faces[3 * f]), v1 = load3(xyz, faces[3 * f + 1]),
             v2 = load3(xyz, faces[3 * f + 2]);
    const F3 e0 = sub(v1, v2), e1 = sub(v2, v0), e2 = sub(v0, v1);
    const float a = sqrtf(dot(e0, e0)), b = sqrtf(dot(e1, e1)), c = sqrtf(dot(e2, e2));
    lengths[3 * f] = a;
    lengths[3 * f + 1] = b;
    lengths[3 * f + 2] = c;
    margin = fminf(fminf((a + b) - c, (b + c) - a), (c + a) - b);
    sum = (static_cast<double>(a) + static_cast<double>(b)) + static_cast<double>(c);
  }
  low[threadIdx.x] = margin;
  total[threadIdx.x] = sum;
  __syncthreads();
  for (int stride = 128; stride > 0; stride >>= 1) {
    if (static_cast<int>(threadIdx.x) < stride) {
      low[threadIdx.x] = fminf(low[threadIdx.x], low[threadIdx.x + stride]);
      total[threadIdx.x] += total[threadIdx.x + stride];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    block_margin[blockIdx.x] = low[0];
    block_sum[blockIdx.x] = total[0];
  }
}

// use_robust's half-cotangents from the side lengths plus the mollifying constant delta, in
// float32 (kernels/laplacian.py:59-90, kernels/predicates.py:194-210): twice the area by Kahan's
// ordering of Heron's formula (largest side first, a negative product clamped, NaN to zero) and
// (l_{e+1}^2 + l_{e+2}^2 - l_e^2) / (4 * 2A); a zero-area face weighs nothing.
__global__ void intrinsic_cot_kernel(const float* __restrict__ lengths, float delta,
                                     int64_t n_faces, float* __restrict__ cot) {
  const int64_t f = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (f >= n_faces) return;
  float l[3];
#pragma unroll
  for (int e = 0; e < 3; ++e) l[e] = delta > 0.0f ? lengths[3 * f + e] + delta : lengths[3 * f + e];
  // Ascending c <= b <= a: the first smallest side, the last largest, and the third.
  int lo = 0, hi = 0;
  for (int e = 1; e < 3; ++e) {
    if (l[e] < l[lo]) lo = e;
    if (l[e] >= l[hi]) hi = e;
  }
  const float c = l[lo], a = l[hi], b = l[3 - lo - hi];
  const float arg = (a + (b + c)) * (c - (a - b)) * (c + (a - b)) * (a + (b - c));
  float dbl = 0.5f * sqrtf(fmaxf(arg, 0.0f));
  if (isnan(dbl)) dbl = 0.0f;
  const float den = 4.0f * dbl;
  float c0 = 0.0f, c1 = 0.0f, c2 = 0.0f;
  if (den > 0.0f) {
    const float s0 = l[0] * l[0], s1 = l[1] * l[1], s2 = l[2] * l[2];
    c0 = ((s1 + s2) - s0) / den;
    c1 = ((s2 + s0) - s1) / den;
    c2 = ((s0 + s1) - s2) / den;
  }
  cot[3 * f] = c0;
  cot[3 * f + 1] = c1;
  cot[3 * f + 2] = c2;
}

// Per face, each corner's share of the area under Meyer et al.'s mixed Voronoi lumping, as
// igl::massmatrix forms it at MASSMATRIX_TYPE_VORONOI (kernels/triangles.py:276-330): in float64
// from the float32 corners, edge lengths l_e = |p_{e+1} - p_{e+2}|, area |(p1 - p0) x (p2 - p0)|/2,
// cosines by the law of cosines over the rounded lengths, circumcentric partials
// (cos_e l_e / sum) area and each corner the mean of the two partials beside it; an obtuse corner
// takes half the area and the others a quarter; a zero-length edge or a zero cosine sum splits
// the area evenly. Rounded once to float32. One thread a face.
__global__ void voronoi_shares_kernel(const float* __restrict__ xyz,
                                      const int32_t* __restrict__ faces, int64_t n_faces,
                                      float* __restrict__ shares) {
  int64_t f = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (f >= n_faces) return;
  const D3 p0 = widen(load3(xyz, faces[3 * f])), p1 = widen(load3(xyz, faces[3 * f + 1])),
           p2 = widen(load3(xyz, faces[3 * f + 2]));
  const D3 e0 = sub(p1, p2), e1 = sub(p2, p0), e2 = sub(p0, p1);
  const double l0 = sqrt(dot(e0, e0)), l1 = sqrt(dot(e1, e1)), l2 = sqrt(dot(e2, e2));
  const D3 c = cross(sub(p1, p0), sub(p2, p0));
  const double area = 0.5 * sqrt(dot(c, c));
  const double even = area / 3.0;
  double q0 = even, q1 = even, q2 = even;
  if (l0 > 0.0 && l1 > 0.0 && l2 > 0.0) {
    const double c0 = (l2 * l2 + l1 * l1 - l0 * l0) / (2.0 * l1 * l2);
    const double c1 = (l0 * l0 + l2 * l2 - l1 * l1) / (2.0 * l2 * l0);
    const double c2 = (l1 * l1 + l0 * l0 - l2 * l2) / (2.0 * l0 * l1);
    const double b0 = c0 * l0, b1 = c1 * l1, b2 = c2 * l2;
    const double weight = b0 + b1 + b2;
    if (weight != 0.0) {
      const double part0 = b0 / weight * area, part1 = b1 / weight * area,
                   part2 = b2 / weight * area;
      q0 = 0.5 * (part1 + part2);
      q1 = 0.5 * (part2 + part0);
      q2 = 0.5 * (part0 + part1);
    }
    const double half = 0.5 * area, quarter = 0.25 * area;
    if (c0 < 0.0) {
      q0 = half;
      q1 = quarter;
      q2 = quarter;
    } else if (c1 < 0.0) {
      q0 = quarter;
      q1 = half;
      q2 = quarter;
    } else if (c2 < 0.0) {
      q0 = quarter;
      q1 = quarter;
      q2 = half;
    }
  }
  shares[3 * f] = static_cast<float>(q0);
  shares[3 * f + 1] = static_cast<float>(q1);
  shares[3 * f + 2] = static_cast<float>(q2);
}

// The lumped mass, summed over each vertex's corners in ascending corner order: with `shares`
// null, m_i = (sum of the float64-widened areas of the faces at i) / 3 (barycentric,
// kernels/heat.py:125-131 then 423-425); otherwise m_i = the sum of the float64-widened Voronoi
// shares of i's corners, with no division (kernels/heat.py:126-128, heat.py:2169).
__global__ void mass_kernel(const int32_t* __restrict__ vertex_ptr,
                            const int32_t* __restrict__ vertex_corners,
                            const float* __restrict__ areas, const float* __restrict__ shares,
                            int64_t n, double* __restrict__ mass) {
  int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= n) return;
  double sum = 0.0;
  if (shares) {
    for (int32_t k = vertex_ptr[i]; k < vertex_ptr[i + 1]; ++k)
      sum += static_cast<double>(shares[vertex_corners[k]]);
    mass[i] = sum;
    return;
  }
  for (int32_t k = vertex_ptr[i]; k < vertex_ptr[i + 1]; ++k)
    sum += static_cast<double>(areas[vertex_corners[k] / 3]);
  mass[i] = sum / 3.0;
}

// First stage of (sum of w |v_j - v_i|, sum of w) over the strict upper triangle of rows
// [row0, row0 + n), w the faces on the edge (MeshLab's face-edge mean, kernels/heat.py:143-235):
// one thread a row, a fixed block tree counted from row0, one partial pair a block.
__global__ void edge_sums_kernel(const int32_t* __restrict__ offsets,
                                 const int32_t* __restrict__ columns,
                                 const int32_t* __restrict__ corner_ptr,
                                 const float* __restrict__ xyz, int64_t row0, int64_t n,
                                 int32_t face_edges, double* __restrict__ partials) {
  __shared__ double s_len[kBlock];
  __shared__ double s_cnt[kBlock];
  const int64_t local = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  const int64_t i = row0 + local;
  double len = 0.0, cnt = 0.0;
  if (local < n) {
    F3 vi = load3(xyz, i);
    for (int32_t k = offsets[i]; k < offsets[i + 1]; ++k) {
      int32_t j = columns[k];
      if (j <= i) continue;
      F3 d = sub(load3(xyz, j), vi);
      double w = face_edges ? static_cast<double>(corner_ptr[k + 1] - corner_ptr[k]) : 1.0;
      len += w * static_cast<double>(sqrtf(dot(d, d)));
      cnt += w;
    }
  }
  s_len[threadIdx.x] = len;
  s_cnt[threadIdx.x] = cnt;
  __syncthreads();
  for (int stride = kBlock / 2; stride > 0; stride >>= 1) {
    if (threadIdx.x < stride) {
      s_len[threadIdx.x] += s_len[threadIdx.x + stride];
      s_cnt[threadIdx.x] += s_cnt[threadIdx.x + stride];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    partials[2 * blockIdx.x] = s_len[0];
    partials[2 * blockIdx.x + 1] = s_cnt[0];
  }
}

// Fold the block partials in block order, and form the heat system's factor: -t, or -h^2 from the
// sums (kernels/linalg.py:1593-1609). scalars = [sum, count, factor, t].
__global__ void edge_sums_final_kernel(const double* __restrict__ partials, int32_t blocks,
                                       int32_t from_edges, double t, double* __restrict__ scalars) {
  __shared__ double s_len[kBlock];
  __shared__ double s_cnt[kBlock];
  double len = 0.0, cnt = 0.0;
  for (int32_t b = threadIdx.x; b < blocks; b += kBlock) {
    len += partials[2 * b];
    cnt += partials[2 * b + 1];
  }
  s_len[threadIdx.x] = len;
  s_cnt[threadIdx.x] = cnt;
  __syncthreads();
  for (int stride = kBlock / 2; stride > 0; stride >>= 1) {
    if (threadIdx.x < stride) {
      s_len[threadIdx.x] += s_len[threadIdx.x + stride];
      s_cnt[threadIdx.x] += s_cnt[threadIdx.x + stride];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    scalars[0] = s_len[0];
    scalars[1] = s_cnt[0];
    if (from_edges) {
      double h = s_cnt[0] > 0.0 ? s_len[0] / s_cnt[0] : 0.0;
      scalars[2] = -1.0 * (h * h);
      scalars[3] = h * h;
    } else {
      scalars[2] = -t;
      scalars[3] = t;
    }
  }
}

// The three operators over one pattern from one walk of each row (kernels/heat.py:298-465):
// L_ij = sum of the entry's float64-widened half-cotangents in ascending corner order
// (kernels/laplacian.py:970-989); L_ii = -(sum of the row's off-diagonal entries in column order);
// heat = factor L (+ m_i on the diagonal, kernels/linalg.py:1612-1621); Poisson = -L. Also the
// heat values narrowed to float32 (the settled CG's stream, linalg.py:2516-2521) and both
// systems' inverse diagonals (Jacobi). One thread a row; packed meshes read their own segment's
// factor (row_segment non-null).
__global__ void operator_rows_kernel(
    const int32_t* __restrict__ offsets, const int32_t* __restrict__ columns,
    const int32_t* __restrict__ corner_ptr, const int32_t* __restrict__ corner_slots,
    const float* __restrict__ cot, const double* __restrict__ mass,
    const double* __restrict__ scalars, const int32_t* __restrict__ row_segment, int64_t n,
    double* __restrict__ lap, double* __restrict__ heat, double* __restrict__ poisson,
    float* __restrict__ heat32, double* __restrict__ heat_inv_diag,
    double* __restrict__ poisson_inv_diag) {
  int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double factor = scalars[4 * (row_segment ? row_segment[i] : 0) + 2];
  double total = 0.0;
  int32_t slot = -1;
  for (int32_t k = offsets[i]; k < offsets[i + 1]; ++k) {
    if (columns[k] == i) {
      slot = k;
      continue;
    }
    double value = 0.0;
    for (int32_t p = corner_ptr[k]; p < corner_ptr[k + 1]; ++p)
      value += static_cast<double>(cot[corner_slots[p]]);
    lap[k] = value;
    total += value;
    double h = factor * value;
    heat[k] = h;
    heat32[k] = static_cast<float>(h);
    poisson[k] = -value;
  }
  double hd = 0.0, pd = 0.0;
  if (slot >= 0) {
    double diagonal = -total;
    lap[slot] = diagonal;
    hd = factor * diagonal + mass[i];
    heat[slot] = hd;
    heat32[slot] = static_cast<float>(hd);
    pd = -diagonal;
    poisson[slot] = pd;
  }
  heat_inv_diag[i] = hd != 0.0 ? 1.0 / hd : 0.0;
  poisson_inv_diag[i] = pd != 0.0 ? 1.0 / pd : 0.0;
}

// The heat system with its solution pinned to zero on the rim (kernels/heat.py:487-508): a rim row
// becomes the identity's and an interior row drops its rim columns, over the heat system's own
// pattern. Symmetric positive definite; against a right-hand side zero on the rim its solution is
// the interior system's padded with zeros. Also the values narrowed to float32 and the inverse
// diagonal, which the settled CG reads. One thread a row.
__global__ void pinned_rows_kernel(const int32_t* __restrict__ offsets,
                                   const int32_t* __restrict__ columns,
                                   const int32_t* __restrict__ rim, const double* __restrict__ heat,
                                   int64_t n, double* __restrict__ out, float* __restrict__ out32,
                                   double* __restrict__ inv_diag) {
  int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const bool on_rim = rim[i] != 0;
  double diagonal = 0.0;
  for (int32_t k = offsets[i]; k < offsets[i + 1]; ++k) {
    const int32_t c = columns[k];
    double value = heat[k];
    if (on_rim) value = c == i ? 1.0 : 0.0;
    else if (rim[c] != 0) value = 0.0;
    out[k] = value;
    out32[k] = static_cast<float>(value);
    if (c == i) diagonal = value;
  }
  inv_diag[i] = diagonal != 0.0 ? 1.0 / diagonal : 0.0;
}

// The regularized Poisson system -L + shift I (kernels/heat.py add_to_diagonal, heat.py:4097-4108):
// the Poisson values with shift added to each stored diagonal entry, and its inverse diagonal.
__global__ void shifted_rows_kernel(const int32_t* __restrict__ offsets,
                                    const int32_t* __restrict__ columns,
                                    const double* __restrict__ values, int64_t n, double shift,
                                    double* __restrict__ out, double* __restrict__ inv_diag) {
  int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= n) return;
  double diagonal = 0.0;
  for (int32_t k = offsets[i]; k < offsets[i + 1]; ++k) {
    double value = values[k];
    if (columns[k] == i) {
      value += shift;
      diagonal = value;
    }
    out[k] = value;
  }
  inv_diag[i] = diagonal != 0.0 ? 1.0 / diagonal : 0.0;
}

// out = in with row r zeroed wherever rim_of[r] is set (kernels/heat.py:511-514, the pinned
// solve's right-hand side), over n_columns columns in `layout` (rim_of indexed by vertex for
// columns, by row of the elimination order for rows). in == out is allowed.
__global__ void off_rim_kernel(const double* in, double* out, const int32_t* __restrict__ rim_of,
                               int64_t n, int64_t n_columns, int32_t rows_layout) {
  const int64_t e = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (e >= n * n_columns) return;
  int64_t r;
  if (rows_layout) {
    const int64_t chunk = e / (kFieldChunk * n);
    const int64_t kk = min(kFieldChunk, n_columns - chunk * kFieldChunk);
    r = (e - chunk * kFieldChunk * n) / kk;
  } else {
    r = e % n;
  }
  out[e] = rim_of[r] != 0 ? 0.0 : in[e];
}

// a = (a + b) / 2: the Neumann heat averaged with the rim-pinned one (kernels/heat.py:517-520).
__global__ void average_kernel(double* __restrict__ a, const double* __restrict__ b, int64_t count) {
  const int64_t e = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (e >= count) return;
  a[e] = 0.5 * (a[e] + b[e]);
}

// u0[q, s] = 1 at each source of column q (duplicates allowed); an index outside [0, n) writes
// nothing and raises the flag (kernels/heat.py:599-623). One thread a source entry.
// rows_k > 0: u0 is stored by rows (FieldLayout::kRows) over rows_k columns, vertex v at row
// row_of[v].
// count_at (device, nullable): the entry count is min(*count_at, n_entries) (a replayed launch
// reads each call's count); the grid strides over the entries.
__global__ void seed_kernel(const int32_t* __restrict__ sources, const int64_t* __restrict__ owner,
                            int64_t n_entries, const int64_t* __restrict__ count_at, int64_t n,
                            int64_t rows_k, const int32_t* __restrict__ row_of,
                            double* __restrict__ u0, int32_t* __restrict__ invalid) {
  const int64_t count = count_at ? min(*count_at, n_entries) : n_entries;
  for (int64_t e = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; e < count;
       e += static_cast<int64_t>(gridDim.x) * blockDim.x) {
    const int32_t v = sources[e];
    if (v < 0 || v >= n) {
      atomicOr(invalid, 1);
      continue;
    }
    const int64_t q = owner[e];
    if (rows_k > 0) {
      const int64_t chunk = q / kFieldChunk;
      const int64_t kk = min(kFieldChunk, rows_k - chunk * kFieldChunk);
      u0[chunk * kFieldChunk * n + static_cast<int64_t>(row_of[v]) * kk + q % kFieldChunk] = 1.0;
    } else {
      u0[q * n + v] = 1.0;
    }
  }
}

// Zeroes `words` 32-bit words: the clears of a solve, as kernels so a captured graph holds no
// memset node.
__global__ void zero_words_kernel(uint32_t* __restrict__ p, int64_t words) {
  for (int64_t e = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; e < words;
       e += static_cast<int64_t>(gridDim.x) * blockDim.x)
    p[e] = 0u;
}

// Per face, the float64 quantities every query's divergence reuses (kernels/triangles.py:834-894,
// kernels/heat.py:626-660), from the float32 positions, unit normal N, area and half-cotangents,
// kFaceGeom doubles a face:
//   [0, 6): a_0 = N x (v2 - v1) and a_1 = N x (v0 - v2). With a_2 = N x (v1 - v0) = -a_0 - a_1,
//       sum_k u_k a_k = (u0 - u2) a_0 + (u1 - u2) a_1 is the heat gradient times 2A (a zero-area
//       face stores zeros: no gradient);
//   [6 + 3c, 9 + 3c), c < 3: corner c's integration vector, corner 0
//       g_0 = c_2 (v1 - v0) + c_1 (v2 - v0), corner 1 g_1 = c_0 (v2 - v1) + c_2 (v0 - v1),
//       corner 2 g_2 = c_1 (v0 - v2) + c_0 (v1 - v2): the corner's term is g_c . x for the unit
//       gradient direction x.
__global__ void face_geometry_kernel(const float* __restrict__ xyz,
                                     const int32_t* __restrict__ faces,
                                     const float* __restrict__ normals,
                                     const float* __restrict__ areas,
                                     const float* __restrict__ cot, int64_t n_faces,
                                     double* __restrict__ geom) {
  const int64_t f = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (f >= n_faces) return;
  const D3 v0 = widen(load3(xyz, faces[3 * f])), v1 = widen(load3(xyz, faces[3 * f + 1])),
           v2 = widen(load3(xyz, faces[3 * f + 2]));
  const D3 nrm = widen(load3(normals, f));
  const bool live = static_cast<double>(areas[f]) > 0.0;
  const D3 zero{0.0, 0.0, 0.0};
  const D3 a[2] = {live ? cross(nrm, sub(v2, v1)) : zero, live ? cross(nrm, sub(v0, v2)) : zero};
  const double c0 = cot[3 * f], c1 = cot[3 * f + 1], c2 = cot[3 * f + 2];
  const D3 g[3] = {add(scale(c2, sub(v1, v0)), scale(c1, sub(v2, v0))),
                   add(scale(c0, sub(v2, v1)), scale(c2, sub(v0, v1))),
                   add(scale(c1, sub(v0, v2)), scale(c0, sub(v1, v2)))};
  double* out = geom + kFaceGeom * f;
#pragma unroll
  for (int k = 0; k < 2; ++k) {
    out[3 * k] = a[k].x;
    out[3 * k + 1] = a[k].y;
    out[3 * k + 2] = a[k].z;
  }
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    out[6 + 3 * k] = g[k].x;
    out[6 + 3 * k + 1] = g[k].y;
    out[6 + 3 * k + 2] = g[k].z;
  }
}

// One corner's divergence term for heat values u0, u1, u2 at its face's vertices: the gradient
// direction s = (u0 - u2) a_0 + (u1 - u2) a_1 is scaled by the power of two that brings its largest
// component into [0.5, 1) (exact, so tiny heat values neither underflow nor lose bits; the
// reference divides by the largest component, kernels/predicates.py:718-724), normalized, and
// dotted with g_c. A zero gradient contributes zero.
__device__ __forceinline__ double corner_term(const double* __restrict__ geom, int32_t corner,
                                              double u0, double u1, double u2) {
#pragma clang fp contract(fast)
  const double* a = geom;
  const double d0 = u0 - u2, d1 = u1 - u2;
  const double sx = d0 * a[0] + d1 * a[3];
  const double sy = d0 * a[1] + d1 * a[4];
  const double sz = d0 * a[2] + d1 * a[5];
  const double largest = fmax(fmax(fabs(sx), fabs(sy)), fabs(sz));
  if (largest == 0.0) return 0.0;
  int e;
  (void)frexp(largest, &e);
  const double wx = ldexp(sx, -e), wy = ldexp(sy, -e), wz = ldexp(sz, -e);
  const double* g = geom + 6 + 3 * corner;
  return (g[0] * wx + g[1] * wy + g[2] * wz) * rsqrt(wx * wx + wy * wy + wz * wz);
}

// -b at vertex i and column q: the terms of i's corners for that column summed in ascending corner
// order (kernels/heat.py:626-687, heat.py:3218-3232), each term recomputed from the face's heat
// values (no per-corner intermediate in memory). heat and out share a layout: entry (row r,
// column q) at r * vstride + q * qstride, vertex i at row row_of[i] (row i when row_of is null),
// and `faces` name rows. Corners go four at a time: their corner ids, then their faces' vertices,
// then the other two heat values and the face data, each step's loads issued together (a vertex
// has about six corners).
//
// kWaveVertex: one wave a vertex, its lanes the columns (kk <= 64 columns stored by rows, so the
// heat rows it reads and the row it writes are contiguous); the vertex, its corners and their face
// data are wave-uniform. Blocks are numbered so that each XCD's share of the grid is one
// contiguous range of vertices (xcd_block), whose neighbours' heat rows then stay in that XCD's
// L2. Otherwise one thread per (vertex, column), the contiguous index fastest.
template <bool kWaveVertex>
__global__ __launch_bounds__(256) void divergence_kernel(
    const int32_t* __restrict__ faces, const double* __restrict__ geom,
    const int32_t* __restrict__ vertex_ptr, const int32_t* __restrict__ vertex_corners,
    const int32_t* __restrict__ row_of, const double* __restrict__ heat, int64_t n, int64_t kk,
    int64_t vstride, int64_t qstride, double* __restrict__ out) {
  int64_t i, q;
  if (kWaveVertex) {
    i = __builtin_amdgcn_readfirstlane(static_cast<int32_t>(xcd_block() * 4 + threadIdx.x / 64));
    q = threadIdx.x % 64;
    if (i >= n || q >= kk) return;
  } else {
    // Whichever index is contiguous in memory runs fastest across the threads.
    const int64_t e = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (e >= n * kk) return;
    if (vstride == 1) {
      q = e / n;
      i = e - q * n;
    } else {
      i = e / kk;
      q = e - i * kk;
    }
  }
  const double* u = heat + q * qstride;
  const int64_t row = row_of ? row_of[i] : i;
  const double ui = u[row * vstride];
  double sum = 0.0;
  const int32_t kbeg = vertex_ptr[i], kend = vertex_ptr[i + 1];
  for (int32_t kb = kbeg; kb < kend; kb += 4) {
    int32_t f[4], corner[4];
#pragma unroll
    for (int w = 0; w < 4; ++w) {
      const int32_t c = vertex_corners[min(kb + w, kend - 1)];  // past the end: repeats, unused
      f[w] = c / 3;
      corner[w] = c - 3 * f[w];
    }
    int32_t v[4][3];
#pragma unroll
    for (int w = 0; w < 4; ++w)
#pragma unroll
      for (int k = 0; k < 3; ++k) v[w][k] = faces[3 * static_cast<int64_t>(f[w]) + k];
    // The corner's own vertex is i: two heat values a corner are read.
    double h[4][3];
#pragma unroll
    for (int w = 0; w < 4; ++w)
#pragma unroll
      for (int k = 0; k < 3; ++k) h[w][k] = k == corner[w] ? ui : u[v[w][k] * vstride];
#pragma unroll
    for (int w = 0; w < 4; ++w) {
      const double term = corner_term(geom + kFaceGeom * static_cast<int64_t>(f[w]), corner[w],
                                      h[w][0], h[w][1], h[w][2]);
      if (kb + w < kend) sum += term;
    }
  }
  out[row * vstride + q * qstride] = sum;
}

// The same -b for a chunk of kk columns stored by rows, one wave a row r of the elimination order
// (its lanes the columns): row r's corners are corners[corner_ptr[r] ..], each {corner id 3f + c,
// rows of face f's vertices 0, 1, 2}, in the vertex's ascending corner order. A wave reads eight
// corner records, then every heat value and face datum they name, then adds the terms of the
// corners that exist (a uniform branch: a vertex's last batch computes no padding); the result row
// is written contiguously. Same terms, same order as divergence_kernel.
__global__ __launch_bounds__(256) void divergence_rows_kernel(
    const int4* __restrict__ corners, const int32_t* __restrict__ corner_ptr,
    const double* __restrict__ geom, const double* __restrict__ heat, int64_t n, int64_t kk,
    double* __restrict__ out) {
  const int64_t r =
      __builtin_amdgcn_readfirstlane(static_cast<int32_t>(xcd_block() * 4 + threadIdx.x / 64));
  const int64_t q = threadIdx.x % 64;
  if (r >= n || q >= kk) return;
  const int32_t kbeg = corner_ptr[r], kend = corner_ptr[r + 1];
  double sum = 0.0;
  for (int32_t kb = kbeg; kb < kend; kb += 8) {
    int4 c[8];
#pragma unroll
    for (int w = 0; w < 8; ++w) c[w] = corners[min(kb + w, kend - 1)];  // past the end: unused
    // A row's base is wave-uniform and the column the lane's offset.
    const uint32_t lane = static_cast<uint32_t>(q);
    double h[8][3];
#pragma unroll
    for (int w = 0; w < 8; ++w) {
      h[w][0] = (heat + static_cast<int64_t>(c[w].y) * kk)[lane];
      h[w][1] = (heat + static_cast<int64_t>(c[w].z) * kk)[lane];
      h[w][2] = (heat + static_cast<int64_t>(c[w].w) * kk)[lane];
    }
#pragma unroll
    for (int w = 0; w < 8; ++w) {
      if (kb + w >= kend) break;
      const int32_t f = c[w].x / 3;
      sum += corner_term(geom + kFaceGeom * static_cast<int64_t>(f), c[w].x - 3 * f, h[w][0],
                         h[w][1], h[w][2]);
    }
  }
  out[r * kk + q] = sum;
}

// divergence_rows_kernel for kk <= 32 columns: 64 / kk rows a wave, lane l holding column l % kk
// of row l / kk (lanes past the last whole row idle), so narrow batches keep their lanes busy.
// Each lane walks its own row's corners four at a time (loads first, then the terms in ascending
// corner order): the same terms in the same order as divergence_rows_kernel.
__global__ __launch_bounds__(256) void divergence_rows_packed_kernel(
    const int4* __restrict__ corners, const int32_t* __restrict__ corner_ptr,
    const double* __restrict__ geom, const double* __restrict__ heat, int64_t n, int32_t kk,
    double* __restrict__ out) {
  const int32_t per_wave = 64 / kk;
  const int32_t lane = static_cast<int32_t>(threadIdx.x % 64);
  const int32_t sub = lane / kk;
  const int64_t r =
      static_cast<int64_t>(xcd_block() * 4 + static_cast<int32_t>(threadIdx.x / 64)) * per_wave +
      sub;
  const int32_t q = lane % kk;
  if (sub >= per_wave || r >= n) return;
  const int32_t kbeg = corner_ptr[r], kend = corner_ptr[r + 1];
  double sum = 0.0;
  for (int32_t kb = kbeg; kb < kend; kb += 4) {
    int4 c[4];
#pragma unroll
    for (int w = 0; w < 4; ++w) c[w] = corners[min(kb + w, kend - 1)];  // past the end: unused
    double h[4][3];
#pragma unroll
    for (int w = 0; w < 4; ++w) {
      h[w][0] = heat[static_cast<int64_t>(c[w].y) * kk + q];
      h[w][1] = heat[static_cast<int64_t>(c[w].z) * kk + q];
      h[w][2] = heat[static_cast<int64_t>(c[w].w) * kk + q];
    }
#pragma unroll
    for (int w = 0; w < 4; ++w) {
      if (kb + w >= kend) break;
      const int32_t f = c[w].x / 3;
      sum += corner_term(geom + kFaceGeom * static_cast<int64_t>(f), c[w].x - 3 * f, h[w][0],
                         h[w][1], h[w][2]);
    }
  }
  out[r * kk + q] = sum;
}

// Columns of the shift: column q starts at col_off[q] and holds col_len[q] entries (packed meshes'
// fields, each its own length), or at q * n with n entries when col_off is null.
struct Columns {
  const int64_t* off;
  const int64_t* len;
  int64_t n;
  __device__ __forceinline__ int64_t start(int64_t q) const { return off ? off[q] : q * n; }
  __device__ __forceinline__ int64_t size(int64_t q) const { return len ? len[q] : n; }
};

// First stage of each column's (sum over its sources, sum over every vertex) of phi
// (kernels/heat.py:3383-3404): block partials of the global sum; the source sum is folded by the
// final stage, sources being few. A column shorter than the grid adds zero partials, so its sum
// is the one a grid of its own length forms.
__global__ void column_sum_kernel(const double* __restrict__ phi, Columns cols, int64_t n_columns,
                                  double* __restrict__ partials) {
  __shared__ double s[kBlock];
  int64_t q = blockIdx.y;
  int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  s[threadIdx.x] = (i < cols.size(q)) ? phi[cols.start(q) + i] : 0.0;
  __syncthreads();
  for (int stride = kBlock / 2; stride > 0; stride >>= 1) {
    if (threadIdx.x < stride) s[threadIdx.x] += s[threadIdx.x + stride];
    __syncthreads();
  }
  if (threadIdx.x == 0) partials[q * gridDim.x + blockIdx.x] = s[0];
}

// One block a column: fold its partials in block order and its sources in order, and decide the
// shift: offset = mean over the sources, orient when the shifted mean is negative
// (kernels/heat.py:3407-3424). shift[2q] = offset, shift[2q+1] = 1 to negate.
__global__ void shift_decide_kernel(const double* __restrict__ phi,
                                    const double* __restrict__ partials, int32_t blocks,
                                    const int32_t* __restrict__ sources,
                                    const int64_t* __restrict__ source_offsets, Columns cols,
                                    double* __restrict__ shift) {
  __shared__ double s[kBlock];
  __shared__ double s_src[kBlock];
  int64_t q = blockIdx.x;
  const int64_t base = cols.start(q), n = cols.size(q);
  double total = 0.0;
  for (int32_t b = threadIdx.x; b < blocks; b += kBlock) total += partials[q * blocks + b];
  int64_t first = source_offsets[q], last = source_offsets[q + 1];
  double src = 0.0;
  // A source out of range (the call then fails) is skipped.
  for (int64_t e = first + threadIdx.x; e < last; e += kBlock) {
    const int32_t v = sources[e];
    if (v >= 0 && v < n) src += phi[base + v];
  }
  s[threadIdx.x] = total;
  s_src[threadIdx.x] = src;
  __syncthreads();
  for (int stride = kBlock / 2; stride > 0; stride >>= 1) {
    if (threadIdx.x < stride) {
      s[threadIdx.x] += s[threadIdx.x + stride];
      s_src[threadIdx.x] += s_src[threadIdx.x + stride];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    int64_t count = last - first;
    double offset = count > 0 ? s_src[0] / static_cast<double>(count) : 0.0;
    shift[2 * q] = offset;
    shift[2 * q + 1] = (count > 0 && s[0] / static_cast<double>(n) - offset < 0.0) ? 1.0 : 0.0;
  }
}

__global__ void shift_apply_kernel(double* __restrict__ phi, const double* __restrict__ shift,
                                   Columns cols, int64_t n_columns) {
  int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  int64_t q = blockIdx.y;
  if (q >= n_columns || i >= cols.size(q)) return;
  const int64_t at = cols.start(q) + i;
  double offset = shift[2 * q];
  double value = phi[at];
  phi[at] = shift[2 * q + 1] != 0.0 ? offset - value : value - offset;
}

// Packed meshes: query q's field is rows [src_off[q], src_off[q] + len[q]) of the solved columns
// (its mesh's segment of its slot's column), copied to out[dst_off[q] ...].
__global__ void extract_fields_kernel(const double* __restrict__ phi,
                                      const int64_t* __restrict__ src_off,
                                      const int64_t* __restrict__ dst_off,
                                      const int64_t* __restrict__ len, double* __restrict__ out) {
  int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  int64_t q = blockIdx.y;
  if (i >= len[q]) return;
  out[dst_off[q] + i] = phi[src_off[q] + i];
}

// ------------------------------------------------------------------------------------------------
// Launchers
// ------------------------------------------------------------------------------------------------

static inline unsigned blocks_for(int64_t n) {
  return static_cast<unsigned>((n + kBlock - 1) / kBlock);
}

void launch_face_terms(const float* xyz, const int32_t* faces, int64_t n_faces, float* cot,
                       float* normals, float* areas, hipStream_t stream) {
  if (n_faces == 0) return;
  face_terms_kernel<<<blocks_for(n_faces), kBlock, 0, stream>>>(xyz, faces, n_faces, cot, normals,
                                                                 areas);
  HG_HIP(hipGetLastError());
}

double launch_mollified_cot(const float* xyz, const int32_t* faces, int64_t n_faces, float* cot,
                            hipStream_t stream) {
  if (n_faces == 0) return 0.0;
  const unsigned blocks = blocks_for(n_faces);
  DeviceArray<float> lengths(3 * n_faces), margins(blocks);
  DeviceArray<double> sums(blocks);
  face_lengths_kernel<<<blocks, kBlock, 0, stream>>>(xyz, faces, n_faces, lengths.ptr, margins.ptr,
                                                     sums.ptr);
  HG_HIP(hipGetLastError());
  std::vector<float> h_margin(blocks);
  std::vector<double> h_sum(blocks);
  HG_HIP(hipMemcpyAsync(h_margin.data(), margins.ptr, sizeof(float) * blocks,
                        hipMemcpyDeviceToHost, stream));
  HG_HIP(hipMemcpyAsync(h_sum.data(), sums.ptr, sizeof(double) * blocks, hipMemcpyDeviceToHost,
                        stream));
  HG_HIP(hipStreamSynchronize(stream));
  float margin = INFINITY;
  double total = 0.0;
  for (unsigned b = 0; b < blocks; ++b) {
    margin = std::min(margin, h_margin[b]);
    total += h_sum[b];
  }
  // delta = max(epsilon * mean - margin, 0) in float32 (laplacian.py:866-870).
  const float eps_scaled = static_cast<float>(kMollifyTolerance * (total / (3.0 * n_faces)));
  const float shortfall = eps_scaled - margin;
  const float delta = shortfall > 0.0f ? shortfall : 0.0f;
  intrinsic_cot_kernel<<<blocks, kBlock, 0, stream>>>(lengths.ptr, delta, n_faces, cot);
  HG_HIP(hipGetLastError());
  HG_HIP(hipStreamSynchronize(stream));
  return delta;
}

void launch_voronoi_shares(const float* xyz, const int32_t* faces, int64_t n_faces, float* shares,
                           hipStream_t stream) {
  if (n_faces == 0) return;
  voronoi_shares_kernel<<<blocks_for(n_faces), kBlock, 0, stream>>>(xyz, faces, n_faces, shares);
  HG_HIP(hipGetLastError());
}

void launch_mass(const int32_t* vertex_ptr, const int32_t* vertex_corners, const float* areas,
                 const float* shares, int64_t n, double* mass, hipStream_t stream) {
  if (n == 0) return;
  mass_kernel<<<blocks_for(n), kBlock, 0, stream>>>(vertex_ptr, vertex_corners, areas, shares, n,
                                                    mass);
  HG_HIP(hipGetLastError());
}

void launch_pinned_rows(const int32_t* offsets, const int32_t* columns, const int32_t* rim,
                        const double* heat, int64_t n, double* out, float* out32, double* inv_diag,
                        hipStream_t stream) {
  if (n == 0) return;
  pinned_rows_kernel<<<blocks_for(n), kBlock, 0, stream>>>(offsets, columns, rim, heat, n, out,
                                                           out32, inv_diag);
  HG_HIP(hipGetLastError());
}

void launch_shifted_rows(const int32_t* offsets, const int32_t* columns, const double* values,
                         int64_t n, double shift, double* out, double* inv_diag,
                         hipStream_t stream) {
  if (n == 0) return;
  shifted_rows_kernel<<<blocks_for(n), kBlock, 0, stream>>>(offsets, columns, values, n, shift,
                                                            out, inv_diag);
  HG_HIP(hipGetLastError());
}

void launch_off_rim(const double* in, double* out, const int32_t* rim_of, int64_t n,
                    int64_t n_columns, FieldLayout layout, hipStream_t stream) {
  if (n * n_columns == 0) return;
  off_rim_kernel<<<blocks_for(n * n_columns), kBlock, 0, stream>>>(
      in, out, rim_of, n, n_columns, layout == FieldLayout::kRows ? 1 : 0);
  HG_HIP(hipGetLastError());
}

void launch_average(double* a, const double* b, int64_t count, hipStream_t stream) {
  if (count == 0) return;
  average_kernel<<<blocks_for(count), kBlock, 0, stream>>>(a, b, count);
  HG_HIP(hipGetLastError());
}

int32_t edge_sum_blocks(int64_t n) { return static_cast<int32_t>(blocks_for(std::max<int64_t>(n, 1))); }

void launch_edge_sums(const int32_t* offsets, const int32_t* columns, const int32_t* corner_ptr,
                      const float* xyz, int64_t row0, int64_t n, int32_t face_edges,
                      double* partials, int32_t from_edges, double t, double* scalars,
                      hipStream_t stream) {
  int32_t blocks = edge_sum_blocks(n);
  if (n > 0)
    edge_sums_kernel<<<blocks, kBlock, 0, stream>>>(offsets, columns, corner_ptr, xyz, row0, n,
                                                    face_edges, partials);
  else
    HG_HIP(hipMemsetAsync(partials, 0, sizeof(double) * 2 * blocks, stream));
  edge_sums_final_kernel<<<1, kBlock, 0, stream>>>(partials, blocks, from_edges, t, scalars);
  HG_HIP(hipGetLastError());
}

void launch_operator_rows(const int32_t* offsets, const int32_t* columns,
                          const int32_t* corner_ptr, const int32_t* corner_slots, const float* cot,
                          const double* mass, const double* scalars, const int32_t* row_segment,
                          int64_t n, double* lap, double* heat, double* poisson, float* heat32,
                          double* heat_inv_diag, double* poisson_inv_diag, hipStream_t stream) {
  if (n == 0) return;
  operator_rows_kernel<<<blocks_for(n), kBlock, 0, stream>>>(
      offsets, columns, corner_ptr, corner_slots, cot, mass, scalars, row_segment, n, lap, heat,
      poisson, heat32, heat_inv_diag, poisson_inv_diag);
  HG_HIP(hipGetLastError());
}

void launch_seed(const int32_t* sources, const int64_t* owner, int64_t n_entries,
                 const int64_t* count_at, int64_t n, int64_t rows_k, const int32_t* row_of,
                 double* u0, int32_t* invalid, hipStream_t stream) {
  if (n_entries == 0) return;
  const unsigned blocks = std::min<unsigned>(blocks_for(n_entries), 1024u);
  seed_kernel<<<blocks, kBlock, 0, stream>>>(sources, owner, n_entries, count_at, n, rows_k,
                                             row_of, u0, invalid);
  HG_HIP(hipGetLastError());
}

void launch_zero(void* p, int64_t bytes, hipStream_t stream) {
  if (bytes <= 0) return;
  HG_REQUIRE(bytes % 4 == 0, HG_ERR_RUNTIME, "a cleared range is whole 32-bit words");
  const int64_t words = bytes / 4;
  const unsigned blocks = std::min<unsigned>(blocks_for(words), 2048u);
  zero_words_kernel<<<blocks, kBlock, 0, stream>>>(static_cast<uint32_t*>(p), words);
  HG_HIP(hipGetLastError());
}

void launch_face_geometry(const float* xyz, const int32_t* faces, const float* normals,
                          const float* areas, const float* cot, int64_t n_faces, double* geom,
                          hipStream_t stream) {
  if (n_faces == 0) return;
  face_geometry_kernel<<<blocks_for(n_faces), kBlock, 0, stream>>>(xyz, faces, normals, areas, cot,
                                                                    n_faces, geom);
  HG_HIP(hipGetLastError());
}

void launch_divergence(const int32_t* faces, const int32_t* row_of, const double* geom,
                       const int32_t* vertex_ptr, const int32_t* vertex_corners,
                       const int4* row_corners, const int32_t* row_corner_ptr,
                       const double* heat, int64_t n, int64_t n_columns, double* out,
                       FieldLayout layout, hipStream_t stream) {
  if (n == 0 || n_columns == 0) return;
  if (layout == FieldLayout::kRows) {
    for (int64_t c0 = 0; c0 < n_columns; c0 += kFieldChunk) {
      const int64_t kk = std::min<int64_t>(kFieldChunk, n_columns - c0);
      // A wave a row while at least half its lanes hold columns; 64 / kk rows a wave below.
      if (kk > kFieldChunk / 2) {
        divergence_rows_kernel<<<static_cast<unsigned>((n + 3) / 4), kBlock, 0, stream>>>(
            row_corners, row_corner_ptr, geom, heat + c0 * n, n, kk, out + c0 * n);
      } else {
        const int64_t rows_per_block = 4 * (kFieldChunk / kk);
        divergence_rows_packed_kernel<<<static_cast<unsigned>((n + rows_per_block - 1) /
                                                              rows_per_block),
                                        kBlock, 0, stream>>>(row_corners, row_corner_ptr, geom,
                                                             heat + c0 * n, n,
                                                             static_cast<int32_t>(kk),
                                                             out + c0 * n);
      }
    }
  } else {
    divergence_kernel<false><<<blocks_for(n * n_columns), kBlock, 0, stream>>>(
        faces, geom, vertex_ptr, vertex_corners, nullptr, heat, n, n_columns, 1, n, out);
  }
  HG_HIP(hipGetLastError());
}

int32_t shift_partial_blocks(int64_t n) { return static_cast<int32_t>(blocks_for(n)); }

void launch_shift(double* phi, const int32_t* sources, const int64_t* source_offsets, int64_t n,
                  int64_t n_columns, const int64_t* col_off, const int64_t* col_len,
                  double* partials, double* shift, hipStream_t stream) {
  if (n == 0 || n_columns == 0) return;
  int32_t blocks = shift_partial_blocks(n);
  dim3 grid(blocks, static_cast<unsigned>(n_columns));
  const Columns cols{col_off, col_len, n};
  column_sum_kernel<<<grid, kBlock, 0, stream>>>(phi, cols, n_columns, partials);
  shift_decide_kernel<<<static_cast<unsigned>(n_columns), kBlock, 0, stream>>>(
      phi, partials, blocks, sources, source_offsets, cols, shift);
  shift_apply_kernel<<<grid, kBlock, 0, stream>>>(phi, shift, cols, n_columns);
  HG_HIP(hipGetLastError());
}

void launch_extract_fields(const double* phi, const int64_t* src_off, const int64_t* dst_off,
                           const int64_t* len, int64_t max_len, int64_t n_queries, double* out,
                           hipStream_t stream) {
  if (max_len == 0 || n_queries == 0) return;
  dim3 grid(blocks_for(max_len), static_cast<unsigned>(n_queries));
  extract_fields_kernel<<<grid, kBlock, 0, stream>>>(phi, src_off, dst_off, len, out);
  HG_HIP(hipGetLastError());
}

}  // namespace hg
