// Persistent conjugate-gradient solves: the settled Jacobi-CG of the heat step and the
// Jacobi-Chebyshev preconditioned CG of the Poisson step, each one cooperative launch.
//
// The iteration is ordito's Chronopoulos-Gear form (kernels/algorithms/conjugate_gradient.py:
// 381-866): a round is w = A u with the three dots r.u, w.u, r.r in one pass, then
// beta = gamma/gamma_old, alpha = gamma / (delta - beta gamma / alpha_old) and the update
// p = u + beta p, s = w + beta s, x += alpha p, r -= alpha s, u = M^-1 r. Where ordito records
// two launches a round into a conditional CUDA graph, this runs the whole loop -- rounds, the
// settle monitor (conjugate_gradient.py:1491-1580) and the stop decisions -- inside one kernel
// with grid-wide barriers, so nothing returns to the host until the solve ends. Every block folds
// the per-block partials in the same fixed order, so all blocks agree on alpha, beta and the stop
// decision bit for bit (the redundant fold ordito uses instead of a barrier, here beside one).
#include "kernels.h"

namespace hg {
namespace {

constexpr int kMaxColumns = 64;

struct CgArgs {
  int64_t n;
  int32_t k;
  const int32_t* offsets;
  const int32_t* columns;
  const float* values32;
  const double* values64;
  const double* inv_diag;
  const double* rhs;
  const double* guess;  // the initial iterate, or null for zero (may alias x)
  double* x;
  double* r;
  double* u;
  double* p;
  double* s;
  double* w;
  double* prev;
  double* res;
  double* da;
  double* db;
  double* partials;  // [k][G][3], then [G][2] for the settle check, then [k][G] for |b|^2
  // [0] stepped rounds, [1] set when a grid barrier gave up waiting, [2] barrier groups arrived
  // at the current barrier, [3] barrier generation, [4, 12) each group's blocks arrived; [1..12)
  // zero at launch.
  int32_t* counters;
  int64_t spin_limit;  // polls of a grid barrier before it gives up (every block then exits)
  int32_t settle;
  int32_t check_rounds;
  double change_tolerance;
  int32_t settle_rounds;
  double tol_sq;
  int64_t maxiter;
  int32_t cheb_steps;
  double theta;
  double delta;
};

// A block's view of the grid barrier: the generation it waits past and whether a barrier gave up.
struct SyncState {
  int32_t generation;
  int32_t aborted;
};

// Grid-wide barrier over co-resident blocks (the launch is cooperative), counted in two stages so
// no counter takes more than about a sixteenth of the grid's arrivals: a block's thread 0 releases
// the block's writes and counts itself into its group's counter (blocks b with the same b mod 8,
// the XCD the hardware deals b to); the last of a group counts the group into the grid's counter,
// and the last group resets the counts and advances the generation the others poll. The polls are
// relaxed loads, one acquire after the wait. The wait is bounded: past a.spin_limit polls, or once
// another block has given up, the block records the abort and stops waiting, and every later
// barrier returns at once, so each block leaves its loop within a round and the kernel ends; the
// host reads the abort and fails the call.
constexpr int kBarrierGroups = 8;

template <int kThreads, bool kSingle>
__device__ __forceinline__ void barrier(const CgArgs& a, SyncState* sync) {
  if constexpr (kSingle) {
    __syncthreads();
  } else {
    __syncthreads();
    if (threadIdx.x == 0 && !sync->aborted) {
      int32_t* abort = a.counters + 1;
      int32_t* grid_count = a.counters + 2;
      int32_t* generation = a.counters + 3;
      int32_t* group_count = a.counters + 4;
      const int32_t blocks = static_cast<int32_t>(gridDim.x);
      const int32_t group = static_cast<int32_t>(blockIdx.x) % kBarrierGroups;
      const int32_t groups = min(blocks, kBarrierGroups);
      const int32_t members = blocks / kBarrierGroups + (group < blocks % kBarrierGroups ? 1 : 0);
      const int32_t gen = sync->generation;
      const int32_t in_group = __hip_atomic_fetch_add(group_count + group, 1, __ATOMIC_ACQ_REL,
                                                      __HIP_MEMORY_SCOPE_AGENT);
      bool last = false;
      if (in_group == members - 1) {
        __hip_atomic_store(group_count + group, 0, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        const int32_t in_grid = __hip_atomic_fetch_add(grid_count, 1, __ATOMIC_ACQ_REL,
                                                       __HIP_MEMORY_SCOPE_AGENT);
        last = in_grid == groups - 1;
      }
      if (last) {
        __hip_atomic_store(grid_count, 0, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        __hip_atomic_store(generation, gen + 1, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
      } else {
        int64_t polls = 0;
        while (__hip_atomic_load(generation, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) == gen) {
          if (__hip_atomic_load(abort, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0 ||
              ++polls > a.spin_limit) {
            __hip_atomic_store(abort, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            sync->aborted = 1;
            break;
          }
          __builtin_amdgcn_s_sleep(2);
        }
        __atomic_thread_fence(__ATOMIC_ACQUIRE);
      }
      sync->generation = gen + 1;
    }
    __syncthreads();
  }
}

// Fixed-order block sum of up to three values, every thread getting the same bits: a butterfly
// within each wave (each step adds two lanes' values, which commute, so every lane holds the
// same sum), then the waves' sums in wave order. Two block barriers.
template <int kThreads, int kWidth>
__device__ __forceinline__ void block_sum(double (&v)[kWidth], double* scratch) {
  constexpr int kWaves = kThreads / 64;
  const int wave = threadIdx.x / 64, lane = threadIdx.x % 64;
#pragma unroll
  for (int c = 0; c < kWidth; ++c)
#pragma unroll
    for (int offset = 32; offset > 0; offset >>= 1) v[c] += __shfl_xor(v[c], offset, 64);
  if (lane == 0)
#pragma unroll
    for (int c = 0; c < kWidth; ++c) scratch[c * kWaves + wave] = v[c];
  __syncthreads();
#pragma unroll
  for (int c = 0; c < kWidth; ++c) {
    double total = 0.0;
    for (int w = 0; w < kWaves; ++w) total += scratch[c * kWaves + w];
    v[c] = total;
  }
  __syncthreads();
}

template <int kThreads>
__device__ __forceinline__ double block_max(double v, double* scratch) {
  constexpr int kWaves = kThreads / 64;
  const int wave = threadIdx.x / 64, lane = threadIdx.x % 64;
#pragma unroll
  for (int offset = 32; offset > 0; offset >>= 1) v = fmax(v, __shfl_xor(v, offset, 64));
  if (lane == 0) scratch[wave] = v;
  __syncthreads();
  double out = scratch[0];
  for (int w = 1; w < kWaves; ++w) out = fmax(out, scratch[w]);
  __syncthreads();
  return out;
}

__device__ __forceinline__ double row_dot(const CgArgs& a, int64_t i, const double* vec) {
  double sum = 0.0;
  const int32_t begin = a.offsets[i], end = a.offsets[i + 1];
  if (a.values32) {
    for (int32_t e = begin; e < end; ++e)
      sum += static_cast<double>(a.values32[e]) * vec[a.columns[e]];
  } else {
    for (int32_t e = begin; e < end; ++e) sum += a.values64[e] * vec[a.columns[e]];
  }
  return sum;
}

// u = p(D^-1 A) D^-1 r for every column: the Chebyshev semi-iteration on D^-1 A z = D^-1 r from
// z = 0 (Saad, Iterative Methods for Sparse Linear Systems, 2nd ed., Algorithm 12.1), cheb_steps
// steps, cheb_steps - 1 products; the search direction is double-buffered so each step is one
// barrier. Leaves every thread past a barrier.
template <int kThreads, bool kSingle>
__device__ void chebyshev_apply(const CgArgs& a, SyncState* sync) {
  const int64_t stride = static_cast<int64_t>(gridDim.x) * kThreads;
  const int64_t first = static_cast<int64_t>(blockIdx.x) * kThreads + threadIdx.x;
  const double sigma = a.theta / a.delta;
  double rho = 1.0 / sigma;
  for (int32_t q = 0; q < a.k; ++q) {
    const int64_t o = static_cast<int64_t>(q) * a.n;
    for (int64_t i = first; i < a.n; i += stride) {
      double c = a.inv_diag[i] * a.r[o + i];
      double d = c / a.theta;
      a.res[o + i] = c;
      a.da[o + i] = d;
      a.u[o + i] = d;
    }
  }
  barrier<kThreads, kSingle>(a, sync);
  double* d_old = a.da;
  double* d_new = a.db;
  for (int32_t step = 1; step < a.cheb_steps; ++step) {
    const double rho_new = 1.0 / (2.0 * sigma - rho);
    const double keep = rho_new * rho;
    const double push = 2.0 * rho_new / a.delta;
    for (int32_t q = 0; q < a.k; ++q) {
      const int64_t o = static_cast<int64_t>(q) * a.n;
      for (int64_t i = first; i < a.n; i += stride) {
        double ad = row_dot(a, i, d_old + o);
        double res = a.res[o + i] - a.inv_diag[i] * ad;
        a.res[o + i] = res;
        double d = keep * d_old[o + i] + push * res;
        d_new[o + i] = d;
        a.u[o + i] += d;
      }
    }
    rho = rho_new;
    barrier<kThreads, kSingle>(a, sync);
    double* t = d_old;
    d_old = d_new;
    d_new = t;
  }
}

template <int kThreads, bool kSingle>
__global__ __launch_bounds__(kThreads) void cg_kernel(CgArgs a) {
  __shared__ double scratch[3 * kThreads];
  __shared__ double gamma_old[kMaxColumns];
  __shared__ double alpha_old[kMaxColumns];
  __shared__ double atol_sq[kMaxColumns];
  __shared__ double coeff[2 * kMaxColumns];
  __shared__ int stepping[kMaxColumns];
  __shared__ SyncState sync_state;
  SyncState* sync = &sync_state;
  if (threadIdx.x == 0) {
    sync_state.generation = 0;
    sync_state.aborted = 0;
  }
  __syncthreads();

  const int32_t G = gridDim.x;
  const int64_t stride = static_cast<int64_t>(G) * kThreads;
  const int64_t first = static_cast<int64_t>(blockIdx.x) * kThreads + threadIdx.x;
  double* round_partials = a.partials;
  double* settle_partials = a.partials + static_cast<int64_t>(a.k) * G * 3;
  double* norm_partials = settle_partials + 2 * G;

  // Initial state from x = 0 (or the guess x0): r = b - A x0, p = s = 0, |b|^2 for the
  // tolerance. Each thread writes only its own rows of x, with the guess's own value.
  for (int32_t q = 0; q < a.k; ++q) {
    const int64_t o = static_cast<int64_t>(q) * a.n;
    double acc[1] = {0.0};
    for (int64_t i = first; i < a.n; i += stride) {
      const double b = a.rhs[o + i];
      double ri = b, xi = 0.0;
      if (a.guess) {
        xi = a.guess[o + i];
        ri = b - row_dot(a, i, a.guess + o);
      }
      a.x[o + i] = xi;
      a.r[o + i] = ri;
      a.p[o + i] = 0.0;
      a.s[o + i] = 0.0;
      a.prev[o + i] = 0.0;
      if (a.cheb_steps == 0) a.u[o + i] = a.inv_diag[i] * ri;
      acc[0] += b * b;
    }
    block_sum<kThreads, 1>(acc, scratch);
    if (threadIdx.x == 0) norm_partials[static_cast<int64_t>(q) * G + blockIdx.x] = acc[0];
  }
  barrier<kThreads, kSingle>(a, sync);
  if (a.cheb_steps > 0) chebyshev_apply<kThreads, kSingle>(a, sync);
  for (int32_t q = 0; q < a.k; ++q) {
    double acc[1] = {0.0};
    for (int32_t b = threadIdx.x; b < G; b += kThreads)
      acc[0] += norm_partials[static_cast<int64_t>(q) * G + b];
    block_sum<kThreads, 1>(acc, scratch);
    if (threadIdx.x == 0) {
      atol_sq[q] = a.tol_sq * acc[0];
      gamma_old[q] = INFINITY;
      alpha_old[q] = 1.0;
    }
  }
  __syncthreads();

  int32_t stepped_rounds = 0;
  int64_t round = 0;
  double settle_previous = 0.0, reach_check = 0.0, checks = 0.0;
  bool running = true;
  while (running) {
    if (sync->aborted) break;
    ++round;
    // w = A u and the round's three dots.
    for (int32_t q = 0; q < a.k; ++q) {
      const int64_t o = static_cast<int64_t>(q) * a.n;
      double acc[3] = {0.0, 0.0, 0.0};
      for (int64_t i = first; i < a.n; i += stride) {
        double wi = row_dot(a, i, a.u + o);
        a.w[o + i] = wi;
        double ri = a.r[o + i], ui = a.u[o + i];
        acc[0] += ri * ui;
        acc[1] += wi * ui;
        acc[2] += ri * ri;
      }
      block_sum<kThreads, 3>(acc, scratch);
      if (threadIdx.x == 0) {
        double* dst = round_partials + (static_cast<int64_t>(q) * G + blockIdx.x) * 3;
        dst[0] = acc[0];
        dst[1] = acc[1];
        dst[2] = acc[2];
      }
    }
    barrier<kThreads, kSingle>(a, sync);
    // Every block folds every column's partials in block order: alpha, beta, whether it steps.
    for (int32_t q = 0; q < a.k; ++q) {
      double acc[3] = {0.0, 0.0, 0.0};
      for (int32_t b = threadIdx.x; b < G; b += kThreads) {
        const double* src = round_partials + (static_cast<int64_t>(q) * G + b) * 3;
        acc[0] += src[0];
        acc[1] += src[1];
        acc[2] += src[2];
      }
      block_sum<kThreads, 3>(acc, scratch);
      if (threadIdx.x == 0) {
        int step = 0;
        double alpha = 0.0, beta = 0.0;
        if (acc[2] > atol_sq[q] && round <= a.maxiter && acc[0] != 0.0) {
          beta = acc[0] / gamma_old[q];
          double denominator = acc[1] - beta * acc[0] / alpha_old[q];
          if (denominator != 0.0) {
            alpha = acc[0] / denominator;
            step = 1;
          }
        }
        stepping[q] = step;
        coeff[2 * q] = alpha;
        coeff[2 * q + 1] = beta;
        if (step) {
          gamma_old[q] = acc[0];
          alpha_old[q] = alpha;
        }
      }
    }
    __syncthreads();
    bool any = false;
    for (int32_t q = 0; q < a.k; ++q) any = any || stepping[q];
    if (!any) break;
    ++stepped_rounds;
    for (int32_t q = 0; q < a.k; ++q) {
      if (!stepping[q]) continue;
      const double alpha = coeff[2 * q], beta = coeff[2 * q + 1];
      const int64_t o = static_cast<int64_t>(q) * a.n;
      for (int64_t i = first; i < a.n; i += stride) {
        double pk = a.u[o + i] + beta * a.p[o + i];
        double sk = a.w[o + i] + beta * a.s[o + i];
        a.p[o + i] = pk;
        a.s[o + i] = sk;
        a.x[o + i] += alpha * pk;
        double rk = a.r[o + i] - alpha * sk;
        a.r[o + i] = rk;
        if (a.cheb_steps == 0) a.u[o + i] = a.inv_diag[i] * rk;
      }
    }
    if (round >= a.maxiter) running = false;
    if (a.cheb_steps > 0) {
      barrier<kThreads, kSingle>(a, sync);
      chebyshev_apply<kThreads, kSingle>(a, sync);
    }
    if (a.settle && round % a.check_rounds == 0) {
      // Per-entry relative change since the last check and the count of reached entries.
      double change = 0.0;
      double reached[1] = {0.0};
      for (int32_t q = 0; q < a.k; ++q) {
        const int64_t o = static_cast<int64_t>(q) * a.n;
        for (int64_t i = first; i < a.n; i += stride) {
          double xv = a.x[o + i];
          if (xv != 0.0) {
            change = fmax(change, fabs(xv - a.prev[o + i]) / fabs(xv));
            reached[0] += 1.0;
          }
          a.prev[o + i] = xv;
        }
      }
      change = block_max<kThreads>(change, scratch);
      block_sum<kThreads, 1>(reached, scratch);
      if (threadIdx.x == 0) {
        settle_partials[2 * blockIdx.x] = change;
        settle_partials[2 * blockIdx.x + 1] = reached[0];
      }
      barrier<kThreads, kSingle>(a, sync);
      double c = 0.0;
      double count[1] = {0.0};
      for (int32_t b = threadIdx.x; b < G; b += kThreads) {
        c = fmax(c, settle_partials[2 * b]);
        count[0] += settle_partials[2 * b + 1];
      }
      c = block_max<kThreads>(c, scratch);
      block_sum<kThreads, 1>(count, scratch);
      checks += 1.0;
      double total = count[0] + 1.0;
      if (total != settle_previous) {
        settle_previous = total;
        reach_check = checks;
      } else {
        double idle = (checks - reach_check) * static_cast<double>(a.check_rounds);
        if (c < a.change_tolerance || idle >= static_cast<double>(a.settle_rounds)) running = false;
      }
    } else if (a.cheb_steps == 0) {
      barrier<kThreads, kSingle>(a, sync);
    }
  }
  if (blockIdx.x == 0 && threadIdx.x == 0) a.counters[0] = stepped_rounds;
}

// One column of at most kRows * 1024 rows in one workgroup, the iteration of cg_kernel with its
// state where one compute unit reaches it fastest: each thread's rows (tid + 1024 j, j < kRows)
// keep x, r, u, p, s, prev and the Chebyshev residual in registers, and the vectors the products
// read from every row (u, and the Chebyshev directions) live in LDS, so a product is one round
// trip for the operator's entries (issued a row at a time, four or eight together) and LDS
// gathers.
// Same recurrences, stop rules, sums in fixed order (a thread's rows in order, then block_sum).
constexpr int kSmallThreads = 1024;
constexpr int kSmallRows = 3;  // rows a thread: systems of at most 3,072 rows

// The products of a thread's rows, kBatch entries of every row a step: each step's columns and
// values of all the rows are loaded together (one round trip for the rows, where a row at a time
// waited once per row), then gathered from LDS and summed, each row in entry order. kWide: the
// operator's float64 values only (the Chebyshev instance's: the Poisson systems have no narrowed
// copy), else the narrowed values when there are.
template <int kRows, bool kWide>
__device__ __forceinline__ void small_row_products(const CgArgs& a, const int32_t (&begin)[kRows],
                                                   const int32_t (&end)[kRows],
                                                   const double* __restrict__ vec,
                                                   double (&out)[kRows]) {
  constexpr int kBatch = kRows == 1 ? 8 : (kRows == 2 ? 4 : (kWide ? 2 : 3));  // no spills
  int32_t longest = 0;
#pragma unroll
  for (int j = 0; j < kRows; ++j) {
    out[j] = 0.0;
    longest = max(longest, end[j] - begin[j]);
  }
  for (int32_t base = 0; base < longest; base += kBatch) {
    int32_t col[kRows][kBatch];
    double val[kRows][kBatch];
#pragma unroll
    for (int j = 0; j < kRows; ++j)
#pragma unroll
      for (int u = 0; u < kBatch; ++u) {
        const bool has = begin[j] + base + u < end[j];
        const int32_t e = has ? begin[j] + base + u : 0;
        col[j][u] = has ? a.columns[e] : 0;
        val[j][u] = !has ? 0.0
                    : kWide || !a.values32 ? a.values64[e]
                                           : static_cast<double>(a.values32[e]);
      }
#pragma unroll
    for (int j = 0; j < kRows; ++j)
#pragma unroll
      for (int u = 0; u < kBatch; ++u)
        if (begin[j] + base + u < end[j]) out[j] += val[j][u] * vec[col[j][u]];
  }
}

template <int kRows, bool kCheb>
__global__ __launch_bounds__(kSmallThreads) void cg_small_kernel(CgArgs a) {
  constexpr int kWaves = kSmallThreads / 64;
  constexpr int kN = kRows * kSmallThreads;
  __shared__ double u_lds[kN];
  __shared__ double d_lds[2][kN];
  __shared__ double scratch[3 * kWaves];
  const int32_t n = static_cast<int32_t>(a.n);
  const int t = threadIdx.x;
  bool live[kRows];
  int32_t begin[kRows], end[kRows];
  // The Chebyshev residual and direction (the Poisson step) or the settle check's last iterate
  // (the heat step): an instance holds one or the other.
  constexpr int kCheck = kCheb ? 1 : kRows, kPoly = kCheb ? kRows : 1;
  double x[kRows], r[kRows], u[kRows], p[kRows], sv[kRows], dinv[kRows];
  double prev[kCheck], res[kPoly], dcur[kPoly];
  double bb[1] = {0.0};
#pragma unroll
  for (int j = 0; j < kRows; ++j) {
    const int32_t i = t + kSmallThreads * j;
    live[j] = i < n;
    begin[j] = live[j] ? a.offsets[i] : 0;
    end[j] = live[j] ? a.offsets[i + 1] : 0;
    dinv[j] = live[j] ? a.inv_diag[i] : 0.0;
    x[j] = live[j] && a.guess ? a.guess[i] : 0.0;
    u_lds[i] = x[j];
  }
  __syncthreads();
  // r = b - A x0 (x0 the guess, else zero).
  double ax[kRows] = {};
  if (a.guess) small_row_products<kRows, kCheb>(a, begin, end, u_lds, ax);
#pragma unroll
  for (int j = 0; j < kRows; ++j) {
    const int32_t i = t + kSmallThreads * j;
    const double b = live[j] ? a.rhs[i] : 0.0;
    r[j] = live[j] && a.guess ? b - ax[j] : b;
    if (live[j]) a.x[i] = x[j];
    p[j] = 0.0;
    sv[j] = 0.0;
    if constexpr (!kCheb) prev[j] = 0.0;
    u[j] = dinv[j] * r[j];
    bb[0] += b * b;
  }
  block_sum<kSmallThreads, 1>(bb, scratch);
  const double atol_sq = a.tol_sq * bb[0];
  double gamma_old = INFINITY, alpha_old = 1.0;
  const double sigma = a.cheb_steps > 0 ? a.theta / a.delta : 0.0;
  // u = p(D^-1 A) D^-1 r for the Poisson step's rounds (chebyshev_apply's recurrence).
  auto chebyshev = [&]() {
    if constexpr (kCheb) {
      double rho = 1.0 / sigma;
#pragma unroll
      for (int j = 0; j < kRows; ++j) {
        const int32_t i = t + kSmallThreads * j;
        const double c = dinv[j] * r[j];
        const double d = c / a.theta;
        res[j] = c;
        dcur[j] = d;
        u[j] = d;
        if (live[j]) d_lds[0][i] = d;
      }
      __syncthreads();
      int cur = 0;
      for (int32_t step = 1; step < a.cheb_steps; ++step) {
        const double rho_new = 1.0 / (2.0 * sigma - rho);
        const double keep = rho_new * rho;
        const double push = 2.0 * rho_new / a.delta;
        double ad[kRows];
        small_row_products<kRows, kCheb>(a, begin, end, d_lds[cur], ad);
#pragma unroll
        for (int j = 0; j < kRows; ++j) {
          const int32_t i = t + kSmallThreads * j;
          if (!live[j]) continue;
          res[j] = res[j] - dinv[j] * ad[j];
          const double d = keep * dcur[j] + push * res[j];
          dcur[j] = d;
          u[j] += d;
          d_lds[cur ^ 1][i] = d;
        }
        rho = rho_new;
        cur ^= 1;
        __syncthreads();
      }
    }
  };
  if (kCheb) chebyshev();
#pragma unroll
  for (int j = 0; j < kRows; ++j)
    if (live[j]) u_lds[t + kSmallThreads * j] = u[j];
  __syncthreads();

  int32_t stepped_rounds = 0;
  int64_t round = 0;
  double settle_previous = 0.0, reach_check = 0.0, checks = 0.0;
  bool running = true;
  while (running) {
    ++round;
    double w[kRows];
    double acc[3] = {0.0, 0.0, 0.0};
    small_row_products<kRows, kCheb>(a, begin, end, u_lds, w);
#pragma unroll
    for (int j = 0; j < kRows; ++j) {
      acc[0] += r[j] * u[j];
      acc[1] += w[j] * u[j];
      acc[2] += r[j] * r[j];
    }
    block_sum<kSmallThreads, 3>(acc, scratch);
    double alpha = 0.0, beta = 0.0;
    bool step = false;
    if (acc[2] > atol_sq && round <= a.maxiter && acc[0] != 0.0) {
      beta = acc[0] / gamma_old;
      const double denominator = acc[1] - beta * acc[0] / alpha_old;
      if (denominator != 0.0) {
        alpha = acc[0] / denominator;
        step = true;
      }
    }
    if (!step) break;
    gamma_old = acc[0];
    alpha_old = alpha;
    ++stepped_rounds;
#pragma unroll
    for (int j = 0; j < kRows; ++j) {
      p[j] = u[j] + beta * p[j];
      sv[j] = w[j] + beta * sv[j];
      x[j] += alpha * p[j];
      r[j] = r[j] - alpha * sv[j];
      if (!kCheb) u[j] = dinv[j] * r[j];
    }
    if (round >= a.maxiter) running = false;
    if (kCheb) chebyshev();
    if (!kCheb && a.settle && round % a.check_rounds == 0) {
      double change = 0.0;
      double reached[1] = {0.0};
#pragma unroll
      for (int j = 0; j < kCheck; ++j) {
        if (live[j] && x[j] != 0.0) {
          change = fmax(change, fabs(x[j] - prev[j]) / fabs(x[j]));
          reached[0] += 1.0;
        }
        prev[j] = x[j];
      }
      const double c = block_max<kSmallThreads>(change, scratch);
      block_sum<kSmallThreads, 1>(reached, scratch);
      checks += 1.0;
      const double total = reached[0] + 1.0;
      if (total != settle_previous) {
        settle_previous = total;
        reach_check = checks;
      } else {
        const double idle = (checks - reach_check) * static_cast<double>(a.check_rounds);
        if (c < a.change_tolerance || idle >= static_cast<double>(a.settle_rounds)) running = false;
      }
    }
    // The next round's product reads u from LDS: every thread is past this round's (block_sum's
    // barriers), so the rows are rewritten, then published.
#pragma unroll
    for (int j = 0; j < kRows; ++j)
      if (live[j]) u_lds[t + kSmallThreads * j] = u[j];
    __syncthreads();
  }
#pragma unroll
  for (int j = 0; j < kRows; ++j)
    if (live[j]) a.x[t + kSmallThreads * j] = x[j];
  if (t == 0) a.counters[0] = stepped_rounds;
}

// Componentwise backward error flag (kernels/linalg.py:1128-1230): "not within" so NaN flags.
__global__ void backward_error_kernel(const int32_t* __restrict__ offsets,
                                      const int32_t* __restrict__ columns,
                                      const double* __restrict__ values,
                                      const double* __restrict__ x, const double* __restrict__ b,
                                      int64_t n, double tolerance, int32_t* __restrict__ flag) {
  int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  int64_t q = blockIdx.y;
  bool fail = false;
  if (i < n) {
    const double* xq = x + q * n;
    double bi = b[q * n + i];
    double total = bi, scale = fabs(bi);
    for (int32_t e = offsets[i]; e < offsets[i + 1]; ++e) {
      double term = values[e] * xq[columns[e]];
      total -= term;
      scale += fabs(term);
    }
    fail = !(fabs(total) <= tolerance * scale);
  }
  // One atomic a failing block, none once the flag reads set.
  const int any = __syncthreads_or(fail ? 1 : 0);
  if (any && threadIdx.x == 0 &&
      __hip_atomic_load(flag, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) == 0)
    atomicOr(flag, 1);
}

__global__ void dominance_partial_kernel(const int32_t* __restrict__ offsets,
                                         const int32_t* __restrict__ columns,
                                         const double* __restrict__ values, int64_t n,
                                         double* __restrict__ partials) {
  __shared__ double scratch[kBlock];
  int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  double ratio = 0.0;
  if (i < n) {
    double diag = 0.0, off = 0.0;
    for (int32_t e = offsets[i]; e < offsets[i + 1]; ++e) {
      if (columns[e] == i) diag = fabs(values[e]);
      else off += fabs(values[e]);
    }
    if (diag > 0.0) ratio = off / diag;
  }
  ratio = block_max<kBlock>(ratio, scratch);
  if (threadIdx.x == 0) partials[blockIdx.x] = ratio;
}

__global__ void dominance_final_kernel(const double* __restrict__ partials, int32_t blocks,
                                       double* __restrict__ out) {
  __shared__ double scratch[kBlock];
  double v = 0.0;
  for (int32_t b = threadIdx.x; b < blocks; b += kBlock) v = fmax(v, partials[b]);
  v = block_max<kBlock>(v, scratch);
  if (threadIdx.x == 0) out[0] = v;
}

// The multi-block solve's grid: one 1024-lane block per compute unit at most (every block folds
// every block's partials each round, and each grid barrier counts every block in, so fewer, wider
// blocks keep a round short), and never more blocks than can be resident at once.
int32_t cooperative_grid(int64_t n) {
  int device = 0;
  HG_HIP(hipGetDevice(&device));
  int cus = 0;
  HG_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, device));
  int per_cu = 0;
  HG_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, cg_kernel<1024, false>, 1024, 0));
  HG_REQUIRE(per_cu >= 1, HG_ERR_RUNTIME, "the conjugate-gradient kernel fits no compute unit");
  const int64_t needed = (n + 1023) / 1024;
  return static_cast<int32_t>(std::max<int64_t>(1, std::min<int64_t>(cus, needed)));
}

// Polls of a grid barrier before it gives up: each poll is an agent-scope load plus a short sleep
// (about a microsecond), so about two seconds, thousands of times a round's longest legitimate
// wait.
constexpr int64_t kSpinLimit = int64_t{1} << 21;

}  // namespace

int32_t run_cg(const CgProblem& problem, CgWorkspace& ws, hipStream_t stream) {
  const int64_t n = problem.n;
  const int64_t k = problem.n_columns;
  HG_REQUIRE(k >= 1 && k <= kMaxColumns, HG_ERR_VALUE, "conjugate gradient takes 1-64 columns");
  if (n == 0) return 0;
  // Small systems run in one block of 1024 lanes with block barriers only.
  const bool single = n <= 8192;
  const int32_t G = single ? 1 : cooperative_grid(n);
  const int64_t entries = n * k;
  if (ws.capacity_entries < entries) {
    ws.vectors.allocate(9 * entries);
    ws.capacity_entries = entries;
  }
  const int64_t partial_size = k * G * 3 + 2 * G + k * G;
  if (ws.partials.size < partial_size) ws.partials.allocate(partial_size);
  if (ws.counters.size < 4 + kBarrierGroups) ws.counters.allocate(4 + kBarrierGroups);
  CgArgs a{};
  a.n = n;
  a.k = static_cast<int32_t>(k);
  a.offsets = problem.offsets;
  a.columns = problem.columns;
  a.values32 = problem.values32;
  a.values64 = problem.values64;
  a.inv_diag = problem.inv_diag;
  a.rhs = problem.rhs;
  a.guess = problem.guess;
  a.x = problem.x;
  double* base = ws.vectors.ptr;
  a.r = base;
  a.u = base + entries;
  a.p = base + 2 * entries;
  a.s = base + 3 * entries;
  a.w = base + 4 * entries;
  a.prev = base + 5 * entries;
  a.res = base + 6 * entries;
  a.da = base + 7 * entries;
  a.db = base + 8 * entries;
  a.partials = ws.partials.ptr;
  a.counters = ws.counters.ptr;
  a.spin_limit = kSpinLimit;
  a.settle = problem.settle;
  a.check_rounds = std::max(problem.check_rounds, 1);
  a.change_tolerance = problem.change_tolerance;
  a.settle_rounds = problem.settle_rounds;
  a.tol_sq = problem.settle ? 0.0 : problem.tolerance * problem.tolerance;
  a.maxiter = problem.maxiter > 0 ? problem.maxiter : 10 * n;
  a.cheb_steps = problem.chebyshev_steps;
  a.theta = 0.5 * (problem.lambda_max + problem.lambda_min);
  a.delta = 0.5 * (problem.lambda_max - problem.lambda_min);
  void* args[] = {&a};
  launch_zero(ws.counters.ptr, sizeof(int32_t) * (4 + kBarrierGroups), stream);
  if (k == 1 && n <= static_cast<int64_t>(kSmallRows) * kSmallThreads) {
    const int64_t rows = (n + kSmallThreads - 1) / kSmallThreads;
    const bool cheb = a.cheb_steps > 0;
    // The Chebyshev instance reads the float64 values only.
    HG_REQUIRE(!cheb || a.values32 == nullptr, HG_ERR_RUNTIME,
               "the Chebyshev-preconditioned solve takes float64 operator values");
    auto kernel = rows == 1 ? (cheb ? cg_small_kernel<1, true> : cg_small_kernel<1, false>)
                  : rows == 2 ? (cheb ? cg_small_kernel<2, true> : cg_small_kernel<2, false>)
                              : (cheb ? cg_small_kernel<3, true> : cg_small_kernel<3, false>);
    hipLaunchKernelGGL(kernel, dim3(1), dim3(kSmallThreads), 0, stream, a);
    HG_HIP(hipGetLastError());
  } else if (single) {
    hipLaunchKernelGGL((cg_kernel<1024, true>), dim3(1), dim3(1024), 0, stream, a);
    HG_HIP(hipGetLastError());
  } else {
    HG_HIP(hipLaunchCooperativeKernel(reinterpret_cast<const void*>(&cg_kernel<1024, false>),
                                      dim3(G), dim3(1024), args, 0, stream));
  }
  int32_t words[2] = {0, 0};
  HG_HIP(hipMemcpyAsync(words, ws.counters.ptr, sizeof(words), hipMemcpyDeviceToHost, stream));
  HG_HIP(hipStreamSynchronize(stream));
  HG_REQUIRE(words[1] == 0, HG_ERR_RUNTIME,
             "conjugate gradient: a grid barrier waited past its limit; the solve was abandoned");
  return words[0];
}

void launch_backward_error(const int32_t* offsets, const int32_t* columns, const double* values,
                           const double* x, const double* b, int64_t n, int64_t n_columns,
                           double tolerance, int32_t* flag, hipStream_t stream) {
  if (n == 0) return;
  dim3 grid(static_cast<unsigned>((n + kBlock - 1) / kBlock), static_cast<unsigned>(n_columns));
  backward_error_kernel<<<grid, kBlock, 0, stream>>>(offsets, columns, values, x, b, n, tolerance,
                                                     flag);
  HG_HIP(hipGetLastError());
}

void launch_dominance_bound(const int32_t* offsets, const int32_t* columns, const double* values,
                            int64_t n, double* partials, double* out, hipStream_t stream) {
  int32_t blocks = static_cast<int32_t>((std::max<int64_t>(n, 1) + kBlock - 1) / kBlock);
  if (n > 0)
    dominance_partial_kernel<<<blocks, kBlock, 0, stream>>>(offsets, columns, values, n, partials);
  else
    HG_HIP(hipMemsetAsync(partials, 0, sizeof(double) * blocks, stream));
  dominance_final_kernel<<<1, kBlock, 0, stream>>>(partials, blocks, out);
  HG_HIP(hipGetLastError());
}

}  // namespace hg
