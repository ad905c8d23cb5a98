// The multigrid-preconditioned conjugate gradient of a refused Poisson system (ordito
// HeatSolver._solve_poisson_system, heat.py:4127-4151, under multigrid_preconditioner,
// linalg.py:5020): the hierarchy is built on the host (multigrid.cpp) and uploaded once; each
// conjugate-gradient round applies one V-cycle -- two damped l1-Jacobi sweeps down, the restricted
// residual, the coarser cycle, the prolonged correction, two sweeps up, the dense generalized
// inverse at the bottom -- over up to kChunk columns at once (column q of a level's vectors at
// q * n). Pre- and post-smoothing are the same symmetric sweeps, so the cycle is a symmetric
// preconditioner. Every reduction is a fixed-order block tree; the host reads one count a round.
#include <algorithm>

#include "kernels.h"
#include "multigrid.h"

namespace hg {
namespace {

constexpr int kChunk = 16;  // columns a cycle carries at once

struct Csr {
  int64_t rows;
  const int32_t* ptr;
  const int32_t* idx;
  const double* val;
};

__device__ __forceinline__ double row_product(const Csr& m, int64_t i, const double* x) {
  double sum = 0.0;
  for (int32_t e = m.ptr[i]; e < m.ptr[i + 1]; ++e) sum += m.val[e] * x[m.idx[e]];
  return sum;
}

// y = M x (add: y += M x) for each column; x has `cols` entries a column, y m.rows.
__global__ void apply_kernel(Csr m, const double* __restrict__ x, int64_t cols, double* y,
                             int32_t add) {
  const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  const int64_t q = blockIdx.y;
  if (i >= m.rows) return;
  const double v = row_product(m, i, x + q * cols);
  if (add) y[q * m.rows + i] += v;
  else y[q * m.rows + i] = v;
}

// One damped Jacobi sweep out = in + s (b - A in); from zero (in null) out = s b.
__global__ void jacobi_kernel(Csr a, const double* __restrict__ s, const double* __restrict__ b,
                              const double* __restrict__ in, double* __restrict__ out) {
  const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  const int64_t q = blockIdx.y;
  if (i >= a.rows) return;
  const int64_t o = q * a.rows;
  if (in == nullptr) {
    out[o + i] = s[i] * b[o + i];
  } else {
    out[o + i] = in[o + i] + s[i] * (b[o + i] - row_product(a, i, in + o));
  }
}

__global__ void residual_kernel(Csr a, const double* __restrict__ b, const double* __restrict__ x,
                                double* __restrict__ r) {
  const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  const int64_t q = blockIdx.y;
  if (i >= a.rows) return;
  const int64_t o = q * a.rows;
  r[o + i] = b[o + i] - row_product(a, i, x + o);
}

// x = G b at the bottom, G dense row-major n x n.
__global__ void dense_kernel(const double* __restrict__ g, int64_t n, const double* __restrict__ b,
                             double* __restrict__ x) {
  const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  const int64_t q = blockIdx.y;
  if (i >= n) return;
  double sum = 0.0;
  for (int64_t j = 0; j < n; ++j) sum += g[i * n + j] * b[q * n + j];
  x[q * n + i] = sum;
}

__device__ __forceinline__ double block_sum(double v, double* scratch) {
  scratch[threadIdx.x] = v;
  __syncthreads();
  for (int stride = kBlock / 2; stride > 0; stride >>= 1) {
    if (threadIdx.x < stride) scratch[threadIdx.x] += scratch[threadIdx.x + stride];
    __syncthreads();
  }
  return scratch[0];
}

// partials[q * G + block] = the block's share of a_q . b_q.
__global__ void dot_partials_kernel(const double* __restrict__ a, const double* __restrict__ b,
                                    int64_t n, double* __restrict__ partials) {
  __shared__ double scratch[kBlock];
  const int64_t q = blockIdx.y;
  double v = 0.0;
  for (int64_t i = static_cast<int64_t>(blockIdx.x) * kBlock + threadIdx.x; i < n;
       i += static_cast<int64_t>(gridDim.x) * kBlock)
    v += a[q * n + i] * b[q * n + i];
  v = block_sum(v, scratch);
  if (threadIdx.x == 0) partials[q * gridDim.x + blockIdx.x] = v;
}

// out[q] = the column's partials folded in block order (one block a column).
__global__ void fold_kernel(const double* __restrict__ partials, int32_t blocks,
                            double* __restrict__ out) {
  __shared__ double scratch[kBlock];
  const int32_t q = blockIdx.x;
  double v = 0.0;
  for (int32_t b = threadIdx.x; b < blocks; b += kBlock) v += partials[q * blocks + b];
  v = block_sum(v, scratch);
  if (threadIdx.x == 0) out[q] = v;
}

// Per column: alpha = rz / pw (zero for a finished column or a vanishing denominator).
__global__ void alpha_kernel(const double* rz, const double* pw, const int32_t* done, int32_t k,
                             double* alpha) {
  const int32_t q = threadIdx.x;
  if (q >= k) return;
  alpha[q] = (!done[q] && pw[q] != 0.0) ? rz[q] / pw[q] : 0.0;
}

// x += alpha p, r -= alpha w.
__global__ void update_kernel(double* x, double* r, const double* __restrict__ p,
                              const double* __restrict__ w, const double* __restrict__ alpha,
                              int64_t n) {
  const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  const int64_t q = blockIdx.y;
  if (i >= n) return;
  const double a = alpha[q];
  x[q * n + i] += a * p[q * n + i];
  r[q * n + i] -= a * w[q * n + i];
}

// done[q] once |r|^2 <= tol^2 |b|^2; *open = the columns not done.
__global__ void converge_kernel(const double* rr, const double* bb, double tol_sq, int32_t k,
                                int32_t* done, int32_t* open) {
  __shared__ int32_t count;
  if (threadIdx.x == 0) count = 0;
  __syncthreads();
  const int32_t q = threadIdx.x;
  if (q < k) {
    if (!(rr[q] > tol_sq * bb[q])) done[q] = 1;
    if (!done[q]) atomicAdd(&count, 1);
  }
  __syncthreads();
  if (threadIdx.x == 0) *open = count;
}

// beta = rz_new / rz for a live column (then rz = rz_new); p = z + beta p.
__global__ void beta_kernel(const double* rz_new, double* rz, const int32_t* done, int32_t k,
                            int32_t first, double* beta) {
  const int32_t q = threadIdx.x;
  if (q >= k) return;
  beta[q] = (first || done[q] || rz[q] == 0.0) ? 0.0 : rz_new[q] / rz[q];
  rz[q] = rz_new[q];
}

__global__ void direction_kernel(double* p, const double* __restrict__ z,
                                 const double* __restrict__ beta, int64_t n) {
  const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  const int64_t q = blockIdx.y;
  if (i >= n) return;
  p[q * n + i] = z[q * n + i] + beta[q] * p[q * n + i];
}

dim3 grid2(int64_t n, int64_t k) {
  return dim3(static_cast<unsigned>((std::max<int64_t>(n, 1) + kBlock - 1) / kBlock),
              static_cast<unsigned>(k));
}

template <typename T>
void upload(DeviceArray<T>& dst, const std::vector<T>& src, hipStream_t stream) {
  dst.allocate(std::max<int64_t>(static_cast<int64_t>(src.size()), 1));
  if (!src.empty())
    HG_HIP(hipMemcpyAsync(dst.ptr, src.data(), sizeof(T) * src.size(), hipMemcpyHostToDevice,
                          stream));
}

Csr view(const MultigridLevel::Matrix& m) { return Csr{m.rows, m.ptr.ptr, m.idx.ptr, m.val.ptr}; }

// out = V-cycle(b) at `level`, for k columns.
void vcycle(Multigrid& mg, size_t level, const double* b, double* out, int64_t k,
            hipStream_t stream) {
  MultigridLevel& L = mg.levels[level];
  const int64_t n = L.a.rows;
  const dim3 grid = grid2(n, k);
  if (level + 1 == mg.levels.size()) {
    dense_kernel<<<grid, kBlock, 0, stream>>>(mg.coarse_inverse.ptr, n, b, out);
    HG_HIP(hipGetLastError());
    return;
  }
  const Csr a = view(L.a);
  MultigridLevel& next = mg.levels[level + 1];
  jacobi_kernel<<<grid, kBlock, 0, stream>>>(a, L.smoother.ptr, b, nullptr, L.u.ptr);
  jacobi_kernel<<<grid, kBlock, 0, stream>>>(a, L.smoother.ptr, b, L.u.ptr, L.v.ptr);
  residual_kernel<<<grid, kBlock, 0, stream>>>(a, b, L.v.ptr, L.res.ptr);
  apply_kernel<<<grid2(next.a.rows, k), kBlock, 0, stream>>>(view(L.r), L.res.ptr, n, next.b.ptr,
                                                             0);
  HG_HIP(hipGetLastError());
  vcycle(mg, level + 1, next.b.ptr, next.x.ptr, k, stream);
  apply_kernel<<<grid, kBlock, 0, stream>>>(view(L.p), next.x.ptr, next.a.rows, L.v.ptr, 1);
  jacobi_kernel<<<grid, kBlock, 0, stream>>>(a, L.smoother.ptr, b, L.v.ptr, L.u.ptr);
  jacobi_kernel<<<grid, kBlock, 0, stream>>>(a, L.smoother.ptr, b, L.u.ptr, out);
  HG_HIP(hipGetLastError());
}

}  // namespace

int64_t Multigrid::bytes() const {
  int64_t total = coarse_inverse.bytes() + r.bytes() + z.bytes() + p.bytes() + w.bytes() +
                  partials.bytes() + scalars.bytes() + flags.bytes();
  for (const MultigridLevel& L : levels)
    total += L.a.bytes() + L.p.bytes() + L.r.bytes() + L.smoother.bytes() + L.u.bytes() +
             L.v.bytes() + L.res.bytes() + L.b.bytes() + L.x.bytes();
  return total;
}

std::unique_ptr<Multigrid> build_multigrid(const HostCsr& fine, uint64_t seed,
                                           hipStream_t stream) {
  HostHierarchy h = build_hierarchy(fine, seed);
  if (!h.usable) return nullptr;
  auto mg = std::make_unique<Multigrid>();
  auto put = [&](MultigridLevel::Matrix& dst, const HostCsr& src) {
    dst.rows = src.rows;
    upload(dst.ptr, src.ptr.empty() ? std::vector<int32_t>{0} : src.ptr, stream);
    upload(dst.idx, src.idx, stream);
    upload(dst.val, src.val, stream);
  };
  mg->levels.resize(h.levels.size());
  for (size_t l = 0; l < h.levels.size(); ++l) {
    const HostLevel& hl = h.levels[l];
    MultigridLevel& L = mg->levels[l];
    put(L.a, hl.a);
    if (l + 1 < h.levels.size()) {
      put(L.p, hl.p);
      put(L.r, hl.r);
    }
    upload(L.smoother, hl.smoother, stream);
    const int64_t n = hl.a.rows * kChunk;
    L.u.allocate(n);
    L.v.allocate(n);
    L.res.allocate(n);
    if (l > 0) {
      L.b.allocate(n);
      L.x.allocate(n);
    }
  }
  upload(mg->coarse_inverse, h.coarse_inverse, stream);
  const int64_t n = fine.rows * kChunk;
  mg->r.allocate(n);
  mg->z.allocate(n);
  mg->p.allocate(n);
  mg->w.allocate(n);
  mg->blocks = static_cast<int32_t>(std::min<int64_t>((fine.rows + kBlock - 1) / kBlock, 1024));
  mg->partials.allocate(static_cast<int64_t>(mg->blocks) * kChunk);
  mg->scalars.allocate(8 * kChunk);
  mg->flags.allocate(kChunk + 1);
  HG_HIP(hipStreamSynchronize(stream));
  return mg;
}

int32_t multigrid_cg(Multigrid& mg, const double* rhs, double* x, int64_t n_columns,
                     double tolerance, int64_t maxiter, hipStream_t stream) {
  const MultigridLevel& top = mg.levels[0];
  const int64_t n = top.a.rows;
  const Csr a = view(top.a);
  int32_t most = 0;
  for (int64_t c0 = 0; c0 < n_columns; c0 += kChunk) {
    const int64_t k = std::min<int64_t>(kChunk, n_columns - c0);
    const double* b = rhs + c0 * n;
    double* xc = x + c0 * n;
    const dim3 grid = grid2(n, k);
    double* bb = mg.scalars.ptr;
    double* rz = bb + kChunk;
    double* rz_new = rz + kChunk;
    double* pw = rz_new + kChunk;
    double* rr = pw + kChunk;
    double* alpha = rr + kChunk;
    double* beta = alpha + kChunk;
    int32_t* done = mg.flags.ptr;
    int32_t* open = mg.flags.ptr + kChunk;
    const dim3 dots(static_cast<unsigned>(mg.blocks), static_cast<unsigned>(k));
    auto dot = [&](const double* u, const double* v, double* out) {
      dot_partials_kernel<<<dots, kBlock, 0, stream>>>(u, v, n, mg.partials.ptr);
      fold_kernel<<<static_cast<unsigned>(k), kBlock, 0, stream>>>(mg.partials.ptr, mg.blocks,
                                                                    out);
    };
    // x = 0, r = b, z = M r, p = z.
    launch_zero(xc, sizeof(double) * n * k, stream);
    HG_HIP(hipMemcpyAsync(mg.r.ptr, b, sizeof(double) * n * k, hipMemcpyDeviceToDevice, stream));
    launch_zero(mg.p.ptr, sizeof(double) * n * k, stream);
    launch_zero(done, sizeof(int32_t) * kChunk, stream);
    dot(b, b, bb);
    converge_kernel<<<1, 64, 0, stream>>>(bb, bb, 0.0, static_cast<int32_t>(k), done, open);
    int32_t rounds = 0;
    for (int64_t it = 0;; ++it) {
      int32_t live = 0;
      HG_HIP(hipMemcpyAsync(&live, open, sizeof(int32_t), hipMemcpyDeviceToHost, stream));
      HG_HIP(hipStreamSynchronize(stream));
      if (live == 0 || it >= maxiter) break;
      vcycle(mg, 0, mg.r.ptr, mg.z.ptr, k, stream);
      dot(mg.r.ptr, mg.z.ptr, rz_new);
      beta_kernel<<<1, 64, 0, stream>>>(rz_new, rz, done, static_cast<int32_t>(k), it == 0 ? 1 : 0,
                                        beta);
      direction_kernel<<<grid, kBlock, 0, stream>>>(mg.p.ptr, mg.z.ptr, beta, n);
      apply_kernel<<<grid, kBlock, 0, stream>>>(a, mg.p.ptr, n, mg.w.ptr, 0);
      dot(mg.p.ptr, mg.w.ptr, pw);
      alpha_kernel<<<1, 64, 0, stream>>>(rz, pw, done, static_cast<int32_t>(k), alpha);
      update_kernel<<<grid, kBlock, 0, stream>>>(xc, mg.r.ptr, mg.p.ptr, mg.w.ptr, alpha, n);
      dot(mg.r.ptr, mg.r.ptr, rr);
      converge_kernel<<<1, 64, 0, stream>>>(rr, bb, tolerance * tolerance,
                                            static_cast<int32_t>(k), done, open);
      HG_HIP(hipGetLastError());
      ++rounds;
    }
    most = std::max(most, rounds);
  }
  return most;
}

}  // namespace hg
