// krylov.hip -- GMRES and the pseudo-transient Newton driver.
//
// GmresSolver::solve (KspSolver.C:353-492): right-preconditioned GMRES(m)
// with Givens rotations and back-substitution (KspSolver.C:919-982).  The
// whole cycle runs on the GPU as a fixed sequence of kernels with no host
// synchronisation: the Hessenberg matrix, rotations, residual estimate and
// the stop flag live in device memory, a one-lane kernel applies each
// rotation, and every vector kernel of a later Arnoldi step returns at once
// when the flag is set.  The host synchronises once per cycle.
//
// Orthogonalisation is classical Gram-Schmidt with one reorthogonalisation
// pass (CGS2) in FP64, fused into three sweeps over the basis per step:
//   sweep 1  h1 = V^T w
//   sweep 2  w -= V h1, then h2 = V^T w   (one read of V serves both)
//   sweep 3  w -= V h2, then ||w||^2
// instead of the reference's j+1 dependent dot/axpy pairs (modified
// Gram-Schmidt, KspSolver.C:404-411).  Reductions are deterministic: a fixed
// grid of kDotBlocks blocks sums contiguous slices (wave butterflies in a
// fixed pattern), and one block sums the partials in order.
//
// The Newton driver follows TsSolver/NewtonSolver for a steady problem with
// one Newton iteration per pseudo-time step (NewtonSolver.h:70-231,
// ImplicitTsDesc.C:225-241): F = R(U) (+ (U - Un)/dt = 0 at the first
// iteration), H1 + diag(1/dt), block-Jacobi, GMRES, U += dQ; the CFL number
// follows CflLaw "Old" (TsParameters.C:198-200) and the run stops when
// ||R||_2 <= eps ||R_0||_2 (TsDesc.C:1261-1278).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <vector>

#include "context.h"
#include "host_internal.h"
#include "host_mesh.h"

namespace aerofx {

constexpr int kDotBlocks = 1024;  // fixed: the reduction order never depends on the device
constexpr int kSweepBlocks = 256;  // blocks of the basis sweeps: one per CU
constexpr int kMaxBasis = 32;     // basis vectors handled by one sweep (m <= 31)
constexpr int kWave = 64;

// Device-resident GMRES state (one per right-hand side).
struct GmresState {
  double H[(kMaxBasis + 1) * kMaxBasis];  // row-major (i, j) -> H[i * kMaxBasis + j]
  double cs[2 * kMaxBasis];               // (sin, cos) per rotation
  double g[kMaxBasis + 1];
  double y[kMaxBasis];
  double dots[3 * kMaxBasis];             // h1 | h2 | ||w||^2
  double beta, target, l2res, invhn;
  int iter, jlast, done, maxits, m;
};

// ---------------------------------------------------------------------------
// deterministic block reductions
// ---------------------------------------------------------------------------
// K = kMaxBasis values per lane: a transposing butterfly halves the values a
// lane carries at each of the first five exchange steps (a lane keeps the
// half selected by its lane bit and adds the partner's copy of that half),
// so after steps 32..2 lane l holds one index, summed over the lanes that
// share bit 0, and the step 1 exchange completes the wave sum: 32 shuffles
// instead of 6 per value.  The fixed pattern keeps the sums deterministic.
template <int HALF, int OFF>
__device__ __forceinline__ void halve_step(double (&v)[kMaxBasis], int lane) {
  const bool up = (lane & OFF) != 0;
#pragma unroll
  for (int i = 0; i < HALF; ++i) {
    const double keep = up ? v[HALF + i] : v[i];
    const double give = up ? v[i] : v[HALF + i];
    v[i] = keep + __shfl_xor(give, OFF, kWave);
  }
}

__device__ __forceinline__ void block_reduce_store_k(double (&v)[kMaxBasis], int nk, int nparts,
                                                     double *__restrict__ partial) {
  static_assert(kMaxBasis == 32 && kWave == 64, "the butterfly below is written for 32 values, 64 lanes");
  __shared__ double sh[kBlock / kWave][kMaxBasis];
  const int lane = threadIdx.x % kWave, wave = threadIdx.x / kWave;
  halve_step<16, 32>(v, lane);
  halve_step<8, 16>(v, lane);
  halve_step<4, 8>(v, lane);
  halve_step<2, 4>(v, lane);
  halve_step<1, 2>(v, lane);
  const double tot = v[0] + __shfl_xor(v[0], 1, kWave);
  // index held by this lane: bits 5..1 of the lane select halves 16, 8, 4, 2, 1
  const int idx = ((lane >> 5) & 1) * 16 + ((lane >> 4) & 1) * 8 + ((lane >> 3) & 1) * 4 + ((lane >> 2) & 1) * 2 +
                  ((lane >> 1) & 1);
  if ((lane & 1) == 0) sh[wave][idx] = tot;
  __syncthreads();
  if ((int)threadIdx.x < nk) {
    double s = 0.0;
#pragma unroll
    for (int w = 0; w < kBlock / kWave; ++w) s += sh[w][threadIdx.x];
    partial[(size_t)threadIdx.x * kDotBlocks + blockIdx.x] = s;
  }
  (void)nparts;
}

__device__ __forceinline__ void block_reduce_store_1(double v, double *__restrict__ partial) {
  __shared__ double sh1[kBlock / kWave];
  const int lane = threadIdx.x % kWave, wave = threadIdx.x / kWave;
#pragma unroll
  for (int off = kWave / 2; off > 0; off >>= 1) v += __shfl_xor(v, off, kWave);
  if (lane == 0) sh1[wave] = v;
  __syncthreads();
  if (threadIdx.x == 0) {
    double s = 0.0;
#pragma unroll
    for (int w = 0; w < kBlock / kWave; ++w) s += sh1[w];
    partial[blockIdx.x] = s;
  }
}

// Arnoldi step j: Hessenberg column, rotations, residual estimate, stop flag
// (KspSolver.C:404-447, 919-962).
// Givens update of Arnoldi step j.  The block's threads stage what it reads
// (the column h1 + h2 of H, ||w||, the stored rotations, g[j]) in LDS in one
// parallel round; thread 0 then runs the rotation recurrence from LDS with
// unconditional loads, so the update costs one global latency.  Every thread
// of the block must call it (it synchronises).
struct GivensStage {
  double h[kMaxBasis + 1];
  double rot[2 * kMaxBasis];
  double gj;
  int done;
};

__device__ void givens_step(GmresState *s, int j, const double *h2, double wsq, GivensStage &g) {
  const int t = threadIdx.x;
  const double hn = sqrt(wsq);
  if (t <= kMaxBasis) {
    double v = 0.0;
    if (t <= j) v = s->dots[t] + h2[t];
    if (t == j + 1) v = hn;
    g.h[t] = v;
  }
  if (t < 2 * kMaxBasis) g.rot[t] = s->cs[t];
  if (t == 0) {
    g.gj = s->g[j];
    g.done = s->done;
  }
  __syncthreads();
  if (t != 0 || g.done) return;
  const int M = kMaxBasis;
  double h[kMaxBasis + 1];
#pragma unroll
  for (int i = 0; i <= kMaxBasis; ++i) h[i] = g.h[i];
#pragma unroll
  for (int i = 0; i < kMaxBasis - 1; ++i) {
    const double sn = g.rot[2 * i], cs = g.rot[2 * i + 1];
    const double hj = cs * h[i] + sn * h[i + 1];
    const double hjp1 = -sn * h[i] + cs * h[i + 1];
    if (i < j) {
      h[i] = hj;
      h[i + 1] = hjp1;
    }
  }
  double hjj = 0.0, hj1 = 0.0;
#pragma unroll
  for (int i = 0; i < kMaxBasis; ++i)
    if (i == j) {
      hjj = h[i];
      hj1 = h[i + 1];
    }
  double sn, cs;
  if (hjj * hjj == 0.0) {
    cs = 0.0;
    sn = 1.0;
  } else {
    double coef = 1.0 / sqrt(hjj * hjj + hj1 * hj1);
    cs = sqrt(hjj * hjj) * coef;
    sn = cs * hj1 / hjj;
  }
  s->cs[2 * j + 1] = cs;
  s->cs[2 * j] = sn;
#pragma unroll
  for (int i = 0; i < kMaxBasis; ++i)
    if (i < j) s->H[i * M + j] = h[i];
  s->H[j * M + j] = cs * hjj + sn * hj1;
  s->H[(j + 1) * M + j] = 0.0;
  const double gj = g.gj;
  const double gj1 = -sn * gj;
  s->g[j + 1] = gj1;
  s->g[j] = cs * gj;
  s->iter += 1;
  s->jlast = j;
  if (hn == 0.0) {
    s->done = 1;
    return;
  }
  s->invhn = 1.0 / hn;
  s->l2res = fabs(gj1);
  if (s->l2res <= s->target || s->iter >= s->maxits) s->done = 1;
}

// mode 0: partial_k = sum w V_k
// mode 1: w -= sum_k h_k V_k, then partial_k = sum w V_k
// mode 2: w -= sum_k h_k V_k, then partial_0 = sum w w
// mode 3: w -= sum_k h_k V_k, then partial_k = sum w V_k and partial_ny = sum w w
// The grid has nparts blocks (fixed by n); block b sums a contiguous slice
// into partial[.][b].  A lane takes two consecutive rows per step (16-B
// loads: half the load instructions, twice the bytes in flight) and adds
// them in row order.
template <int MODE>
__global__ __launch_bounds__(kBlock) void k_sweep(size_t n, double *__restrict__ w, const double *__restrict__ V,
                                                  size_t ldv, int ny, const double *__restrict__ h,
                                                  double *__restrict__ partial, const int *__restrict__ done) {
  if (done && *done) return;
  // slices hold an even number of rows, so that row pairs never straddle two blocks
  const size_t slice = ((n + gridDim.x - 1) / gridDim.x + 1) & ~(size_t)1;
  // rounding the slice up can leave the last blocks without rows (b0 >= n)
  const size_t b0 = min(n, blockIdx.x * slice), b1 = min(n, b0 + slice);
  double acc[kMaxBasis];
#pragma unroll
  for (int k = 0; k < kMaxBasis; ++k) acc[k] = 0.0;
  double hk[kMaxBasis];
  if (MODE != 0) {
#pragma unroll
    for (int k = 0; k < kMaxBasis; ++k) hk[k] = k < ny ? h[k] : 0.0;
  }
  // one row of the sweep: w_i update and accumulation, in row order per lane
  auto row = [&](size_t i, double wi, const double *vk) {
    if (MODE != 0) {
#pragma unroll
      for (int k = 0; k < kMaxBasis; ++k)
        if (k < ny) wi -= hk[k] * vk[k];
      if (ny > 0) w[i] = wi;
    }
    if (MODE == 2) {
      acc[0] += wi * wi;
    } else if (MODE == 3) {
#pragma unroll
      for (int k = 0; k < kMaxBasis; ++k) acc[k] += k == ny ? wi * wi : wi * vk[k];
    } else {
#pragma unroll
      for (int k = 0; k < kMaxBasis; ++k) acc[k] += wi * vk[k];
    }
  };
  const size_t even_end = b0 + ((b1 - b0) & ~(size_t)1);
  for (size_t i = b0 + 2 * threadIdx.x; i < even_end; i += 2 * kBlock) {
    double2 v[kMaxBasis];
#pragma unroll
    for (int k = 0; k < kMaxBasis; ++k)
      v[k] = k < ny ? *reinterpret_cast<const double2 *>(V + (size_t)k * ldv + i) : make_double2(0.0, 0.0);
    const double2 w2 = *reinterpret_cast<const double2 *>(w + i);
    double vk[kMaxBasis];
#pragma unroll
    for (int k = 0; k < kMaxBasis; ++k) vk[k] = v[k].x;
    row(i, w2.x, vk);
#pragma unroll
    for (int k = 0; k < kMaxBasis; ++k) vk[k] = v[k].y;
    row(i + 1, w2.y, vk);
  }
  // an odd row count (last block only): the lane whose next pair would start
  // at the last row takes it, after its other rows
  if (even_end < b1 && threadIdx.x == ((b1 - 1 - b0) / 2) % kBlock) {
    double vk[kMaxBasis];
#pragma unroll
    for (int k = 0; k < kMaxBasis; ++k) vk[k] = k < ny ? V[(size_t)k * ldv + b1 - 1] : 0.0;
    row(b1 - 1, w[b1 - 1], vk);
  }
  if (MODE == 2)
    block_reduce_store_1(acc[0], partial);
  else
    block_reduce_store_k(acc, MODE == 3 ? ny + 1 : ny, gridDim.x, partial);
}

// out[k] = sum_b partial[k][b] in a fixed tree, one block per k
__global__ __launch_bounds__(kBlock) void k_reduce_partials(const double *__restrict__ partial, int nparts,
                                                            double *__restrict__ out, const int *__restrict__ done,
                                                            GmresState *gs, int j) {
  if (done && *done) return;
  __shared__ double sh[kBlock];
  const int k = blockIdx.x;
  double s = 0.0;
  for (int b = threadIdx.x; b < nparts; b += kBlock) s += partial[(size_t)k * kDotBlocks + b];
  sh[threadIdx.x] = s;
  __syncthreads();
  for (int st = kBlock / 2; st > 0; st >>= 1) {
    if ((int)threadIdx.x < st) sh[threadIdx.x] += sh[threadIdx.x + st];
    __syncthreads();
  }
  if (threadIdx.x == 0) out[k] = sh[0];
  if (gs) {
    // after the norm of step j: the Givens update (single block launch)
    __shared__ GivensStage stage;
    givens_step(gs, j, gs->dots + kMaxBasis, sh[0], stage);
  }
}

// After the mode-3 sweep of Arnoldi step j: out[0..j] = h2 = V^T w1 and
// ||w1||^2, summed over the partials (32 lanes per value, each lane's slice
// loaded in independent batches, then a fixed butterfly); then
// ||w1 - V h2||^2 = ||w1||^2 - ||h2||^2 (V orthonormal) goes to the norm slot
// and the Givens update runs.  One block of kReduceThreads.
constexpr int kReduceThreads = 1024;
__global__ __launch_bounds__(kReduceThreads) void k_reduce_cgs2(const double *__restrict__ partial, int nparts,
                                                                double *__restrict__ out, GmresState *gs, int j) {
  if (gs->done) return;
  const int ny = j + 1;
  __shared__ double sh[kMaxBasis + 1];
  const int k = threadIdx.x / 32, sub = threadIdx.x % 32;
  double t = 0.0;
  if (k <= ny) {
    const double *row = partial + (size_t)k * kDotBlocks;
    for (int b0 = sub; b0 < nparts; b0 += 32 * 8) {
      double v[8];
#pragma unroll
      for (int u = 0; u < 8; ++u) v[u] = b0 + 32 * u < nparts ? row[b0 + 32 * u] : 0.0;
#pragma unroll
      for (int u = 0; u < 8; ++u) t += v[u];
    }
  }
  t += __shfl_xor(t, 16, 32);
  t += __shfl_xor(t, 8, 32);
  t += __shfl_xor(t, 4, 32);
  t += __shfl_xor(t, 2, 32);
  t += __shfl_xor(t, 1, 32);
  if (sub == 0 && k <= ny) {
    sh[k] = t;
    if (k < ny) out[k] = t;
  }
  __syncthreads();
  __shared__ double wsq;
  if (threadIdx.x == 0) {
    double h2sq = 0.0;
    for (int i = 0; i < ny; ++i) h2sq += sh[i] * sh[i];
    const double r = sh[ny] - h2sq;
    wsq = r > 0.0 ? r : 0.0;
    gs->dots[2 * kMaxBasis] = wsq;
  }
  __syncthreads();
  __shared__ GivensStage stage;
  givens_step(gs, j, sh, wsq, stage);
}

// y = a x + b y, a possibly scaled by a device scalar (*as)
__global__ __launch_bounds__(kBlock) void k_axpby(size_t n, double a, const double *__restrict__ as,
                                                  const double *__restrict__ x, double b, double *__restrict__ y,
                                                  const int *__restrict__ done) {
  if (done && *done) return;
  const size_t i = blockIdx.x * (size_t)kBlock + threadIdx.x;
  if (i >= n) return;
  const double aa = as ? a * (*as) : a;
  y[i] = aa * x[i] + b * y[i];
}

__global__ __launch_bounds__(kBlock) void k_fill(size_t n, double v, double *__restrict__ y) {
  const size_t i = blockIdx.x * (size_t)kBlock + threadIdx.x;
  if (i < n) y[i] = v;
}

// w = sum_{k <= jlast} y_k V_k
__global__ __launch_bounds__(kBlock) void k_combine(size_t n, double *__restrict__ w, const double *__restrict__ V,
                                                    size_t ldv, const GmresState *__restrict__ s) {
  const size_t i = blockIdx.x * (size_t)kBlock + threadIdx.x;
  if (i >= n) return;
  const int nk = s->jlast + 1;
  double acc = 0.0;
  for (int k = 0; k < nk; ++k) acc += s->y[k] * V[(size_t)k * ldv + i];
  w[i] = acc;
}

// start of a cycle: beta = ||r|| (r = b - A x), target from ||b|| on the first
__global__ void k_gmres_start(GmresState *s, const double *__restrict__ nrm2, int first, double eps) {
  if (threadIdx.x != 0) return;
  const double beta = sqrt(nrm2[0]);
  s->beta = beta;
  if (first) {
    s->target = eps * beta;  // x0 = 0: r = b
    s->iter = 0;
  }
  for (int k = 0; k <= s->m; ++k) s->g[k] = 0.0;
  s->g[0] = beta;
  s->done = (beta == 0.0) ? 2 : 0;
  s->invhn = beta == 0.0 ? 0.0 : 1.0 / beta;
  s->jlast = -1;
}

// back-substitution (KspSolver.C:968-982)
__global__ void k_gmres_backsolve(GmresState *s) {
  if (threadIdx.x != 0) return;
  const int M = kMaxBasis, j = s->jlast;
  for (int i = j; i >= 0; --i) {
    double yi = s->g[i];
    for (int k = j; k >= i + 1; --k) yi -= s->H[i * M + k] * s->y[k];
    s->y[i] = yi / s->H[i * M + i];
  }
}

// count nodes whose density or pressure is not positive
__global__ __launch_bounds__(kBlock) void k_check_states(int N, const double *__restrict__ U, Gas g,
                                                         int *__restrict__ bad) {
  const int n = blockIdx.x * kBlock + threadIdx.x;
  if (n >= N) return;
  const double *u = U + 5 * (size_t)n;
  double rho = u[0];
  double vel2 = (u[1] * u[1] + u[2] * u[2] + u[3] * u[3]) / (rho * rho);
  double p = (g.gam - 1.0) * (u[4] - 0.5 * rho * vel2) - g.gam * g.pstiff;
  if (!(rho > 0.0) || !(p + g.pstiff > 0.0)) atomicAdd(bad, 1);
}

// ---------------------------------------------------------------------------
// MatVecProdFD (MatVecProd.C:262-372, 879-929)
// ---------------------------------------------------------------------------
// partial sums of eps0 |U| + eps0 over all entries (computeEpsilon)
__global__ __launch_bounds__(kBlock) void k_eps_sum(size_t n, const double *__restrict__ U,
                                                    double *__restrict__ partial) {
  const size_t slice = (n + kDotBlocks - 1) / kDotBlocks;
  const size_t b0 = blockIdx.x * slice, b1 = min(n, b0 + slice);
  const double eps0 = 1.e-6;
  double acc = 0.0;
  for (size_t i = b0 + threadIdx.x; i < b1; i += kBlock) acc += eps0 * fabs(U[i]) + eps0;
  block_reduce_store_1(acc, partial);
}

// scal[0] = p.p, scal[1] = sum(eps0 |U| + eps0) -> scal[2] = eps, [3] = 1/eps, [4] = 0.5/eps
__global__ void k_fd_eps(double *scal, double size) {
  if (threadIdx.x != 0) return;
  const double norm = sqrt(scal[0]);
  const double eps = norm < 1.0e-14 ? 1.e-6 : scal[1] / (size * norm);
  scal[2] = eps;
  scal[3] = 1.0 / eps;
  scal[4] = 0.5 / eps;
}

// Q = U + s eps p (s = +1 or -1)
__global__ __launch_bounds__(kBlock) void k_fd_shift(size_t n, const double *__restrict__ U,
                                                     const double *__restrict__ p, const double *__restrict__ scal,
                                                     double s, double *__restrict__ Q) {
  const size_t i = blockIdx.x * (size_t)kBlock + threadIdx.x;
  if (i >= n) return;
  const double ep = scal[2] * p[i];
  Q[i] = s > 0.0 ? U[i] + ep : U[i] - ep;
}

// TimeState::add_dAW_dt, backward Euler, Un = U (TimeState.C:40-73):
// F += (1/dt) (1 Q + (-1) Un)
__global__ __launch_bounds__(kBlock) void k_fd_time(size_t n, const double *__restrict__ Q,
                                                    const double *__restrict__ U, const double *__restrict__ dt,
                                                    double *__restrict__ F) {
  const size_t i = blockIdx.x * (size_t)kBlock + threadIdx.x;
  if (i >= n) return;
  const double invDt = 1.0 / dt[i / 5];
  F[i] += invDt * (Q[i] - U[i]);
}

// prod = scal[idx] (Fa - Fb)
__global__ __launch_bounds__(kBlock) void k_fd_diff(size_t n, const double *__restrict__ Fa,
                                                    const double *__restrict__ Fb, const double *__restrict__ scal,
                                                    int idx, double *__restrict__ prod) {
  const size_t i = blockIdx.x * (size_t)kBlock + threadIdx.x;
  if (i >= n) return;
  prod[i] = scal[idx] * (Fa[i] - Fb[i]);
}

struct Krylov {
  int m = 0;
  size_t n = 0;
  size_t ldv = 0;             // basis stride: n rounded up to 256 B, so every V_k is 16-B aligned
  int nparts = 0;             // blocks of the dot-product sweeps (fixed by n)
  double *V = nullptr;        // [m+1][n] basis
  double *w = nullptr, *z = nullptr, *r = nullptr, *x = nullptr, *b = nullptr, *R = nullptr;
  double *partial = nullptr;  // [kMaxBasis][kDotBlocks]
  double *scal = nullptr;     // device scalars
  GmresState *st = nullptr;   // device
  GmresState *h_st = nullptr; // pinned host mirror
  // the Arnoldi sequence of one cycle, captured once per (m, precision, transpose)
  hipStream_t cap_stream = nullptr;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  hipEvent_t order_ev = nullptr;
  int graph_key = -1;
  double *h_scal = nullptr;   // pinned
};

void free_krylov(aerofx_ctx *c) {
  if (!c->ksp) return;
  if (c->ksp->exec) (void)hipGraphExecDestroy(c->ksp->exec);
  if (c->ksp->graph) (void)hipGraphDestroy(c->ksp->graph);
  if (c->ksp->cap_stream) (void)hipStreamDestroy(c->ksp->cap_stream);
  if (c->ksp->order_ev) (void)hipEventDestroy(c->ksp->order_ev);
  if (c->ksp->h_st) (void)hipHostFree(c->ksp->h_st);
  if (c->ksp->h_scal) (void)hipHostFree(c->ksp->h_scal);
  delete c->ksp;
  c->ksp = nullptr;  // device buffers belong to c->allocs
}

static int ensure_krylov(aerofx_ctx *c, int m) {
  if (m < 1 || m > kMaxBasis - 1) return set_error(AEROFX_ERR_ARG, "gmres: Krylov dimension must be in [1, 31]");
  if (c->ksp && c->ksp->m >= m) return 0;
  free_krylov(c);
  Krylov *k = new Krylov();
  k->m = m;
  k->n = 5 * (size_t)c->N;
  // about 8 entries per lane per sweep, between 64 blocks and one per CU (the
  // sweeps hold up to 32 basis rows per lane: one wave per SIMD)
  k->nparts = (int)std::min<size_t>(kSweepBlocks, std::max<size_t>(64, (k->n + 8 * kBlock - 1) / (8 * kBlock)));
  void *p = nullptr;
  int rc;
  k->ldv = (k->n + 31) & ~(size_t)31;
  if ((rc = ctx_alloc(c, sizeof(double) * (size_t)(m + 1) * k->ldv, &p))) { delete k; return rc; }
  k->V = (double *)p;
  double **vecs[6] = {&k->w, &k->z, &k->r, &k->x, &k->b, &k->R};
  for (double **v : vecs) {
    if ((rc = ctx_alloc(c, sizeof(double) * k->n, &p))) { delete k; return rc; }
    *v = (double *)p;
  }
  if ((rc = ctx_alloc(c, sizeof(double) * kMaxBasis * kDotBlocks, &p))) { delete k; return rc; }
  k->partial = (double *)p;
  if ((rc = ctx_alloc(c, sizeof(double) * 16, &p))) { delete k; return rc; }
  k->scal = (double *)p;

  if ((rc = ctx_alloc(c, sizeof(GmresState), &p))) { delete k; return rc; }
  k->st = (GmresState *)p;
  hipError_t e = hipHostMalloc((void **)&k->h_st, sizeof(GmresState), hipHostMallocDefault);
  if (e == hipSuccess) e = hipHostMalloc((void **)&k->h_scal, sizeof(double) * 16, hipHostMallocDefault);
  if (e != hipSuccess) { delete k; return set_hip_error(e, "hipHostMalloc"); }
  c->ksp = k;
  return 0;
}

static inline dim3 grid_for(size_t n) { return dim3((unsigned)((n + kBlock - 1) / kBlock)); }

template <int MODE>
static void sweep(aerofx_ctx *c, double *w, const double *V, size_t ldv, int ny, const double *h, double *out,
                  const int *done, GmresState *gs = nullptr, int j = -1) {
  Krylov *k = c->ksp;
  hipLaunchKernelGGL((k_sweep<MODE>), dim3(k->nparts), dim3(kBlock), 0, c->stream, k->n, w, V, ldv, ny, h,
                     k->partial, done);
  hipLaunchKernelGGL(k_reduce_partials, dim3(MODE == 2 ? 1 : ny), dim3(kBlock), 0, c->stream, k->partial,
                     k->nparts, out, done, MODE == 2 ? gs : nullptr, j);
}

static int fetch_state(aerofx_ctx *c) {
  hipError_t e = hipMemcpyAsync(c->ksp->h_st, c->ksp->st, sizeof(GmresState), hipMemcpyDeviceToHost, c->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
  return e == hipSuccess ? 0 : set_hip_error(e, "gmres state copy");
}

// The m Arnoldi steps of one cycle (no host decisions: a converged cycle
// leaves the remaining kernels to return at once on the stop flag).
static int enqueue_arnoldi(aerofx_ctx *c, int m, bool trans) {
  Krylov *k = c->ksp;
  const size_t n = k->n;
  int *done = &k->st->done;
  int rc;
  for (int j = 0; j < m; ++j) {
    double *Vj = k->V + (size_t)j * k->ldv;
    {
      // V_j = (r, or w1 - V h2 of the previous step) * invhn, z = P V_j
      ProfScope ps(c, AEROFX_TIMER_PREC_APPLY);
      if ((rc = launch_prec_norm_flag(c, j == 0 ? k->r : k->w, &k->st->invhn, k->V, k->ldv, j,
                                      k->st->dots + kMaxBasis, Vj, k->z, done, trans)))
        return rc;
    }
    {
      ProfScope ps(c, AEROFX_TIMER_JAC_APPLY);
      if ((rc = launch_h1_apply_flag(c, k->z, k->w, done, trans))) return rc;
    }
    {
      ProfScope ps(c, AEROFX_TIMER_GRAM_SCHMIDT);
      // CGS2 in two passes: h1 = V^T w; w1 = w - V h1, h2 = V^T w1, ||w1||;
      // the second update w1 - V h2 is applied by the next step's first kernel
      sweep<0>(c, k->w, k->V, k->ldv, j + 1, nullptr, k->st->dots, done);
      hipLaunchKernelGGL((k_sweep<3>), dim3(k->nparts), dim3(kBlock), 0, c->stream, k->n, k->w, k->V, k->ldv, j + 1,
                         k->st->dots, k->partial, done);
      hipLaunchKernelGGL(k_reduce_cgs2, dim3(1), dim3(kReduceThreads), 0, c->stream, k->partial, k->nparts,
                         k->st->dots + kMaxBasis, k->st, j);
    }
  }
  hipError_t e = hipGetLastError();
  return e == hipSuccess ? 0 : set_hip_error(e, "arnoldi launch");
}

// Runs the cycle's Arnoldi steps: eagerly when profiling (per-kernel events),
// otherwise as one hipGraph captured on a private stream, ordered after the
// context stream by events.  The graph holds kernel nodes only.
static int arnoldi_cycle(aerofx_ctx *c, int m, bool trans) {
  Krylov *k = c->ksp;
  if (c->prof) return enqueue_arnoldi(c, m, trans);
  const int key = (m * 2 + (trans ? 1 : 0)) * 2 + c->jac_precision;
  hipError_t e = hipSuccess;
  if (!k->cap_stream) e = hipStreamCreateWithFlags(&k->cap_stream, hipStreamNonBlocking);
  if (e == hipSuccess && !k->order_ev) e = hipEventCreateWithFlags(&k->order_ev, hipEventDisableTiming);
  if (e != hipSuccess) return set_hip_error(e, "arnoldi stream");
  if (!k->exec || k->graph_key != key) {
    if (k->exec) (void)hipGraphExecDestroy(k->exec);
    if (k->graph) (void)hipGraphDestroy(k->graph);
    k->exec = nullptr;
    k->graph = nullptr;
    hipStream_t saved = c->stream;
    c->stream = k->cap_stream;
    e = hipStreamBeginCapture(k->cap_stream, hipStreamCaptureModeThreadLocal);
    int rc = e == hipSuccess ? enqueue_arnoldi(c, m, trans) : 0;
    hipError_t e2 = hipStreamEndCapture(k->cap_stream, &k->graph);
    c->stream = saved;
    if (rc) return rc;
    if (e == hipSuccess) e = e2;
    if (e == hipSuccess) e = hipGraphInstantiate(&k->exec, k->graph, nullptr, nullptr, 0);
    if (e != hipSuccess) return set_hip_error(e, "arnoldi graph");
    k->graph_key = key;
  }
  hipEvent_t ev = k->order_ev;
  e = hipEventRecord(ev, c->stream);
  if (e == hipSuccess) e = hipStreamWaitEvent(k->cap_stream, ev, 0);
  if (e == hipSuccess) e = hipGraphLaunch(k->exec, k->cap_stream);
  if (e == hipSuccess) e = hipEventRecord(ev, k->cap_stream);
  if (e == hipSuccess) e = hipStreamWaitEvent(c->stream, ev, 0);
  return e == hipSuccess ? 0 : set_hip_error(e, "arnoldi graph launch");
}

struct GmresOut {
  int its = 0;
  double res = 0, target = 0;
};

// Right-preconditioned GMRES(m) on H1 with block-Jacobi: A x = b, x0 = 0
// (ImplicitCoupledTsDesc::solveLinearSystem sets dQ = 0).  Enqueues one
// cycle at a time and synchronises once per cycle.
static int gmres(aerofx_ctx *c, const double *b, double *x, int m, double eps, int maxits, GmresOut *out,
                 bool trans = false) {
  int rc = ensure_krylov(c, m);
  if (rc) return rc;
  Krylov *k = c->ksp;
  const size_t n = k->n;
  *k->h_st = GmresState{};
  k->h_st->maxits = maxits;
  k->h_st->m = m;
  hipError_t e = hipMemcpyAsync(k->st, k->h_st, sizeof(GmresState), hipMemcpyHostToDevice, c->stream);
  if (e != hipSuccess) return set_hip_error(e, "gmres init copy");
  int *done = &k->st->done;
  hipLaunchKernelGGL(k_fill, grid_for(n), dim3(kBlock), 0, c->stream, n, 0.0, x);
  for (int cycle = 0;; ++cycle) {
    // r = b - A x
    if (cycle == 0) {
      hipLaunchKernelGGL(k_axpby, grid_for(n), dim3(kBlock), 0, c->stream, n, 1.0, nullptr, b, 0.0, k->r, nullptr);
    } else {
      if ((rc = launch_h1_apply(c, x, k->w, trans))) return rc;
      hipLaunchKernelGGL(k_axpby, grid_for(n), dim3(kBlock), 0, c->stream, n, 1.0, nullptr, b, 0.0, k->r, nullptr);
      hipLaunchKernelGGL(k_axpby, grid_for(n), dim3(kBlock), 0, c->stream, n, -1.0, nullptr, k->w, 1.0, k->r,
                         nullptr);
    }
    sweep<2>(c, k->r, nullptr, 0, 0, nullptr, k->scal, nullptr);
    hipLaunchKernelGGL(k_gmres_start, dim3(1), dim3(64), 0, c->stream, k->st, k->scal, cycle == 0 ? 1 : 0, eps);
    if ((rc = arnoldi_cycle(c, m, trans))) return rc;
    if ((rc = fetch_state(c))) return rc;
    const GmresState &hs = *k->h_st;
    if (hs.jlast >= 0) {
      hipLaunchKernelGGL(k_gmres_backsolve, dim3(1), dim3(64), 0, c->stream, k->st);
      hipLaunchKernelGGL(k_combine, grid_for(n), dim3(kBlock), 0, c->stream, n, k->w, k->V, k->ldv, k->st);
      if ((rc = launch_prec_apply(c, k->w, k->r, trans))) return rc;
      hipLaunchKernelGGL(k_axpby, grid_for(n), dim3(kBlock), 0, c->stream, n, 1.0, nullptr, k->r, 1.0, x, nullptr);
    }
    if (hs.done) break;  // converged, maxits, zero residual or lucky breakdown
  }
  e = hipGetLastError();
  if (e != hipSuccess) return set_hip_error(e, "gmres launch");
  out->its = k->h_st->iter;
  out->res = k->h_st->l2res;
  out->target = k->h_st->target;
  return 0;
}

// MatVecProdFD::evaluate + apply: prod = (F(U + eps p) - F(U)) / eps (order 1)
// or (F(U + eps p) - F(U - eps p)) / (2 eps) (order 2), F = R + time term.
static int jvp_fd(aerofx_ctx *c, const double *U, const double *p, double *prod, int order, const double *dt) {
  int rc = ensure_krylov(c, c->ksp ? c->ksp->m : 1);
  if (rc) return rc;
  Krylov *k = c->ksp;
  const size_t n = k->n;
  // computeEpsilon
  sweep<2>(c, const_cast<double *>(p), nullptr, 0, 0, nullptr, k->scal, nullptr);
  hipLaunchKernelGGL(k_eps_sum, dim3(kDotBlocks), dim3(kBlock), 0, c->stream, n, U, k->partial);
  hipLaunchKernelGGL(k_reduce_partials, dim3(1), dim3(kBlock), 0, c->stream, k->partial, kDotBlocks, k->scal + 1,
                     nullptr, nullptr, -1);
  hipLaunchKernelGGL(k_fd_eps, dim3(1), dim3(64), 0, c->stream, k->scal, (double)n);
  double *Fa = k->w, *Fb = k->z, *Q = k->r;
  // F(U + eps p)
  hipLaunchKernelGGL(k_fd_shift, grid_for(n), dim3(kBlock), 0, c->stream, n, U, p, k->scal, 1.0, Q);
  if ((rc = launch_residual(c, Q, Fa, true))) return rc;
  if (dt) hipLaunchKernelGGL(k_fd_time, grid_for(n), dim3(kBlock), 0, c->stream, n, Q, U, dt, Fa);
  if (order == 1) {
    // F(U): the time term vanishes at Q = Un
    if ((rc = launch_residual(c, U, Fb, true))) return rc;
    if (dt) hipLaunchKernelGGL(k_fd_time, grid_for(n), dim3(kBlock), 0, c->stream, n, U, U, dt, Fb);
    hipLaunchKernelGGL(k_fd_diff, grid_for(n), dim3(kBlock), 0, c->stream, n, Fa, Fb, k->scal, 3, prod);
  } else {
    hipLaunchKernelGGL(k_fd_shift, grid_for(n), dim3(kBlock), 0, c->stream, n, U, p, k->scal, -1.0, Q);
    if ((rc = launch_residual(c, Q, Fb, true))) return rc;
    if (dt) hipLaunchKernelGGL(k_fd_time, grid_for(n), dim3(kBlock), 0, c->stream, n, Q, U, dt, Fb);
    hipLaunchKernelGGL(k_fd_diff, grid_for(n), dim3(kBlock), 0, c->stream, n, Fa, Fb, k->scal, 4, prod);
  }
  hipError_t e = hipGetLastError();
  return e == hipSuccess ? 0 : set_hip_error(e, "jvp_fd launch");
}

}  // namespace aerofx

using namespace aerofx;

extern "C" {

int aerofx_jvp_fd(aerofx_ctx *c, const double *U, const double *p, double *Jp, int order, const double *dt) {
  if (!c || !U || !p || !Jp) return set_error(AEROFX_ERR_ARG, "jvp_fd: null argument");
  if (order != 1 && order != 2) return set_error(AEROFX_ERR_ARG, "jvp_fd: order 1 or 2");
  int rc = jvp_fd(c, U, p, Jp, order, dt);
  if (rc) return rc;
  hipError_t e = hipMemcpyAsync(c->h_err, c->d_err, sizeof(int), hipMemcpyDeviceToHost, c->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
  if (e != hipSuccess) return set_hip_error(e, "jvp_fd sync");
  return c->h_err[0];
}

int aerofx_adjoint_solve(aerofx_ctx *c, const double *U, const double *rhs, double *psi,
                         const aerofx_adjoint_opts *o, aerofx_adjoint_stats *st) {
  if (!c || !U || !rhs || !psi || !o || !st) return set_error(AEROFX_ERR_ARG, "adjoint_solve: null argument");
  if (o->precision != 0 && o->precision != 1) return set_error(AEROFX_ERR_ARG, "adjoint_solve: precision 0 or 1");
  int rc = ensure_krylov(c, o->gmres_m);
  if (rc) return rc;
  Krylov *k = c->ksp;
  const size_t n = k->n;
  auto t0 = std::chrono::steady_clock::now();
  *st = aerofx_adjoint_stats{};
  // A = H1 + diag(1/dt) at U (pseudo-time step of the given CFL), its
  // block-Jacobi inverse in the requested precision
  if ((rc = launch_primitive(c, U, c->d_V))) return rc;
  if (o->cfl > 0.0) {
    if ((rc = launch_timestep(c, o->cfl, c->d_dt))) return rc;
  }
  if ((rc = launch_jacobian(c, o->cfl > 0.0 ? c->d_dt : nullptr, o->precision))) return rc;
  // ||rhs||
  sweep<2>(c, const_cast<double *>(rhs), nullptr, 0, 0, nullptr, k->scal, nullptr);
  hipError_t e = hipMemcpyAsync(k->h_scal, k->scal, sizeof(double), hipMemcpyDeviceToHost, c->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
  if (e != hipSuccess) return set_hip_error(e, "adjoint_solve sync");
  const double rnorm0 = std::sqrt(k->h_scal[0]);
  hipLaunchKernelGGL(k_fill, grid_for(n), dim3(kBlock), 0, c->stream, n, 0.0, psi);
  double res = rnorm0;
  int it;
  for (it = 0;; ++it) {
    // r = rhs - (dR/dU)^T psi (FP64, exact second-order adjoint)
    if (it == 0) {
      hipLaunchKernelGGL(k_axpby, grid_for(n), dim3(kBlock), 0, c->stream, n, 1.0, nullptr, rhs, 0.0, k->b,
                         nullptr);
    } else {
      if ((rc = launch_adjoint(c, U, psi, k->R, nullptr))) return rc;
      hipLaunchKernelGGL(k_axpby, grid_for(n), dim3(kBlock), 0, c->stream, n, 1.0, nullptr, rhs, 0.0, k->b,
                         nullptr);
      hipLaunchKernelGGL(k_axpby, grid_for(n), dim3(kBlock), 0, c->stream, n, -1.0, nullptr, k->R, 1.0, k->b,
                         nullptr);
    }
    sweep<2>(c, k->b, nullptr, 0, 0, nullptr, k->scal, nullptr);
    e = hipMemcpyAsync(k->h_scal, k->scal, sizeof(double), hipMemcpyDeviceToHost, c->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    if (e != hipSuccess) return set_hip_error(e, "adjoint_solve sync");
    res = std::sqrt(k->h_scal[0]);
    if (res <= o->eps * rnorm0 || rnorm0 == 0.0) {
      st->converged = 1;
      break;
    }
    if (it >= o->max_its) break;
    // A^T d = r with the transposed H1 and block-Jacobi
    GmresOut go;
    if ((rc = gmres(c, k->b, k->x, o->gmres_m, o->gmres_eps, o->gmres_maxits, &go, true))) return rc;
    st->krylov_its += go.its;
    hipLaunchKernelGGL(k_axpby, grid_for(n), dim3(kBlock), 0, c->stream, n, 1.0, nullptr, k->x, 1.0, psi, nullptr);
  }
  e = hipStreamSynchronize(c->stream);
  if (e != hipSuccess) return set_hip_error(e, "adjoint_solve sync");
  st->its = it;
  st->res0 = rnorm0;
  st->res = res;
  st->seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  return AEROFX_OK;
}

int aerofx_jvp_transpose(aerofx_ctx *c, const double *p, double *JTp) {
  if (!c || !p || !JTp) return set_error(AEROFX_ERR_ARG, "jvp_transpose: null argument");
  return launch_h1_apply(c, p, JTp, true);
}

int aerofx_prec_apply_transpose(aerofx_ctx *c, const double *y, double *x) {
  if (!c || !y || !x) return set_error(AEROFX_ERR_ARG, "prec_apply_transpose: null argument");
  return launch_prec_apply(c, y, x, true);
}

int aerofx_gmres_transpose(aerofx_ctx *c, const double *b, double *x, int m, double eps, int maxits, int *its,
                           double *final_res) {
  if (!c || !b || !x) return set_error(AEROFX_ERR_ARG, "gmres_transpose: null argument");
  if (!c->d_jacD) return set_error(AEROFX_ERR_STATE, "gmres_transpose: operators not set up");
  GmresOut o;
  int rc = gmres(c, b, x, m, eps, maxits, &o, true);
  if (rc) return rc;
  if (its) *its = o.its;
  if (final_res) *final_res = o.res;
  hipError_t e = hipStreamSynchronize(c->stream);
  return e == hipSuccess ? AEROFX_OK : set_hip_error(e, "gmres_transpose sync");
}

int aerofx_timestep(aerofx_ctx *c, const double *U, double cfl, double *dt) {
  if (!c || !U || !dt) return set_error(AEROFX_ERR_ARG, "timestep: null argument");
  int rc = launch_primitive(c, U, c->d_V);
  return rc ? rc : launch_timestep(c, cfl, dt);
}

int aerofx_jacobian_setup(aerofx_ctx *c, const double *U, const double *dt, int precision) {
  if (!c || !U) return set_error(AEROFX_ERR_ARG, "jacobian_setup: null argument");
  if (precision != 0 && precision != 1) return set_error(AEROFX_ERR_ARG, "jacobian_setup: precision 0 or 1");
  int rc = launch_primitive(c, U, c->d_V);
  if (rc) return rc;
  if ((rc = launch_jacobian(c, dt, precision))) return rc;
  hipError_t e = hipMemcpyAsync(c->h_err + 1, c->d_err + 1, sizeof(int), hipMemcpyDeviceToHost, c->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
  if (e != hipSuccess) return set_hip_error(e, "jacobian_setup sync");
  if (c->h_err[1]) return set_error(AEROFX_ERR_STATE, "jacobian_setup: zero pivot in a diagonal block");
  return AEROFX_OK;
}

int aerofx_jacobian_blocks(aerofx_ctx *c, double *D, double *Aedge) {
  if (!c || !c->d_jacD) return set_error(AEROFX_ERR_STATE, "jacobian_blocks: no Jacobian");
  const size_t esz = c->jac_precision ? sizeof(float) : sizeof(double);
  std::vector<char> hd(esz * blk64_count(c->N)), hb(esz * blk64_count(2 * (size_t)c->E));
  hipError_t e = hipStreamSynchronize(c->stream);
  if (e == hipSuccess) e = hipMemcpy(hd.data(), c->d_jacD, hd.size(), hipMemcpyDeviceToHost);
  if (e == hipSuccess) e = hipMemcpy(hb.data(), c->d_jacB, hb.size(), hipMemcpyDeviceToHost);
  if (e != hipSuccess) return set_hip_error(e, "jacobian_blocks copy");
  auto at = [&](const std::vector<char> &v, size_t idx) -> double {
    return c->jac_precision ? (double)((const float *)v.data())[idx] : ((const double *)v.data())[idx];
  };
  if (D)
    for (size_t n = 0; n < (size_t)c->N; ++n)
      for (int kk = 0; kk < 25; ++kk) D[25 * n + kk] = at(hd, blk64_base(n) + kk * kBlkStride);
  if (Aedge) {
    // adjacency entry q of node n -> edge layout [E][2][25] = (A_ij, A_ji)
    std::vector<int> rowptr, adj, eid;
    node_csr(c->N, c->E, c->ptr.data(), rowptr, adj, eid);
    for (int n = 0; n < c->N; ++n)
      for (int q = rowptr[n]; q < rowptr[n + 1]; ++q) {
        int l = eid[q];
        int side = n < adj[q] ? 0 : 1;
        for (int kk = 0; kk < 25; ++kk)
          Aedge[50 * (size_t)l + 25 * side + kk] = at(hb, blk64_base(q) + kk * kBlkStride);
      }
  }
  return AEROFX_OK;
}

int aerofx_jvp(aerofx_ctx *c, const double *p, double *Jp) {
  if (!c || !p || !Jp) return set_error(AEROFX_ERR_ARG, "jvp: null argument");
  return launch_h1_apply(c, p, Jp);
}

int aerofx_prec_apply(aerofx_ctx *c, const double *y, double *x) {
  if (!c || !y || !x) return set_error(AEROFX_ERR_ARG, "prec_apply: null argument");
  return launch_prec_apply(c, y, x);
}

int aerofx_gmres(aerofx_ctx *c, const double *b, double *x, int m, double eps, int maxits, int *its,
                 double *final_res) {
  if (!c || !b || !x) return set_error(AEROFX_ERR_ARG, "gmres: null argument");
  if (!c->d_jacD) return set_error(AEROFX_ERR_STATE, "gmres: operators not set up");
  GmresOut o;
  int rc = gmres(c, b, x, m, eps, maxits, &o);
  if (rc) return rc;
  if (its) *its = o.its;
  if (final_res) *final_res = o.res;
  hipError_t e = hipStreamSynchronize(c->stream);
  return e == hipSuccess ? AEROFX_OK : set_hip_error(e, "gmres sync");
}

int aerofx_steady_solve(aerofx_ctx *c, double *U, const aerofx_newton_opts *o, aerofx_newton_stats *st,
                        double *history, int hist_cap) {
  if (!c || !U || !o || !st) return set_error(AEROFX_ERR_ARG, "steady_solve: null argument");
  if (o->precision != 0 && o->precision != 1) return set_error(AEROFX_ERR_ARG, "steady_solve: precision 0 or 1");
  int rc = ensure_krylov(c, o->gmres_m);
  if (rc) return rc;
  Krylov *k = c->ksp;
  const size_t n = k->n;
  auto t0 = std::chrono::steady_clock::now();
  *st = aerofx_newton_stats{};
  double res0 = 0.0, res = 0.0, cfl = o->cfl0;
  int step;
  for (step = 0; step <= o->max_steps; ++step) {
    // spatial residual and its norm (TsDesc::computeResidualNorm); one
    // synchronisation per step reads the norm and the error counters
    if ((rc = launch_residual(c, U, k->R, true))) return rc;
    sweep<2>(c, k->R, nullptr, 0, 0, nullptr, k->scal, nullptr);
    hipError_t e = hipMemcpyAsync(k->h_scal, k->scal, sizeof(double), hipMemcpyDeviceToHost, c->stream);
    if (e == hipSuccess) e = hipMemcpyAsync(c->h_err, c->d_err, 3 * sizeof(int), hipMemcpyDeviceToHost, c->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    if (e != hipSuccess) return set_hip_error(e, "steady_solve sync");
    if (step > 0 && c->h_err[2]) {
      st->steps = step;
      return set_error(AEROFX_ERR_STATE, "steady_solve: update produced non-physical states");
    }
    if (c->h_err[0]) {
      st->steps = step;
      return set_error(AEROFX_ERR_STATE, "steady_solve: non-physical reconstructed state");
    }
    res = std::sqrt(k->h_scal[0]);
    if (step == 0) res0 = res;
    if (history && step < hist_cap) history[step] = res;
    if (res == 0.0 || res <= o->eps * res0) {
      st->converged = 1;
      break;
    }
    if (step == o->max_steps) break;
    // CflLaw Old (TsParameters.C:198-200) with its = step
    double rr = res / res0;
    cfl = std::min(std::max(std::max(o->cfl_coef1, std::min(step * o->cfl_coef2, cfl + o->cfl_coef2)),
                            o->cfl0 / std::pow(rr, o->ser)),
                   o->cfl_max);
    if ((rc = launch_timestep(c, cfl, c->d_dt))) return rc;
    if ((rc = launch_jacobian(c, c->d_dt, o->precision))) return rc;
    // b = -F
    hipLaunchKernelGGL(k_axpby, grid_for(n), dim3(kBlock), 0, c->stream, n, -1.0, nullptr, k->R, 0.0, k->b,
                       nullptr);
    GmresOut go;
    if ((rc = gmres(c, k->b, k->x, o->gmres_m, o->gmres_eps, o->gmres_maxits, &go))) return rc;
    st->krylov_its += go.its;
    // Q += dQ; the state check is read at the next step's synchronisation
    hipLaunchKernelGGL(k_axpby, grid_for(n), dim3(kBlock), 0, c->stream, n, 1.0, nullptr, k->x, 1.0, U, nullptr);
    hipLaunchKernelGGL(k_zero_int, dim3(1), dim3(64), 0, c->stream, c->d_err + 2, 1);
    hipLaunchKernelGGL(k_check_states, dim3((c->N + kBlock - 1) / kBlock), dim3(kBlock), 0, c->stream, c->N, U,
                       c->gas, c->d_err + 2);
    st->final_cfl = cfl;
  }
  hipError_t e = hipStreamSynchronize(c->stream);
  if (e != hipSuccess) return set_hip_error(e, "steady_solve sync");
  st->steps = step;
  st->res0 = res0;
  st->res = res;
  st->seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  return AEROFX_OK;
}

}  // extern "C"
