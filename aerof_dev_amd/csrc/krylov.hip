// krylov.hip -- GMRES and the pseudo-transient Newton driver.
//
// GmresSolver::solve (KspSolver.C:353-492) with right preconditioning,
// Givens rotations and back-substitution (KspSolver.C:919-982) on the host;
// every vector operation runs on the GPU in FP64.  Orthogonalisation is
// classical Gram-Schmidt with one reorthogonalisation pass (CGS2): each pass
// is one fused multi-dot (all basis vectors against w in one sweep) and one
// fused update, instead of the reference's j+1 sequential dot/axpy pairs
// (modified Gram-Schmidt, KspSolver.C:404-411).  Dot products reduce
// deterministically: a fixed grid of kDotBlocks blocks, each summing its
// contiguous slice with a fixed tree, then one block summing the partials in
// order.
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
constexpr int kDotGroup = 8;      // basis vectors per multi-dot pass

// partial[g * kDotBlocks + b] = sum over block b's slice of x * Y_g
__global__ __launch_bounds__(kBlock) void k_multidot_partial(size_t n, const double *__restrict__ x,
                                                             const double *__restrict__ Y, size_t ldy, int ny,
                                                             double *__restrict__ partial) {
  __shared__ double sh[kDotGroup][kBlock];
  const int g0 = blockIdx.y * kDotGroup;
  const int ng = min(kDotGroup, ny - g0);
  const size_t slice = (n + kDotBlocks - 1) / kDotBlocks;
  const size_t b0 = blockIdx.x * slice, b1 = min(n, b0 + slice);
  double acc[kDotGroup];
#pragma unroll
  for (int g = 0; g < kDotGroup; ++g) acc[g] = 0.0;
  for (size_t i = b0 + threadIdx.x; i < b1; i += kBlock) {
    const double xi = x[i];
#pragma unroll
    for (int g = 0; g < kDotGroup; ++g)
      if (g < ng) acc[g] += xi * Y[(size_t)(g0 + g) * ldy + i];
  }
#pragma unroll
  for (int g = 0; g < kDotGroup; ++g) sh[g][threadIdx.x] = acc[g];
  __syncthreads();
  for (int s = kBlock / 2; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s)
#pragma unroll
      for (int g = 0; g < kDotGroup; ++g) sh[g][threadIdx.x] += sh[g][threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x < (unsigned)ng) partial[(size_t)(g0 + threadIdx.x) * kDotBlocks + blockIdx.x] = sh[threadIdx.x][0];
}

// out[g] = sum_b partial[g][b], fixed tree; one block per g
__global__ __launch_bounds__(kBlock) void k_multidot_final(const double *__restrict__ partial,
                                                           double *__restrict__ out) {
  __shared__ double sh[kBlock];
  const int g = blockIdx.x;
  double s = 0.0;
  for (int b = threadIdx.x; b < kDotBlocks; b += kBlock) s += partial[(size_t)g * kDotBlocks + b];
  sh[threadIdx.x] = s;
  __syncthreads();
  for (int st = kBlock / 2; st > 0; st >>= 1) {
    if ((int)threadIdx.x < st) sh[threadIdx.x] += sh[threadIdx.x + st];
    __syncthreads();
  }
  if (threadIdx.x == 0) out[g] = sh[0];
}

// w -= sum_k h[k] Y_k  (h on the device)
__global__ __launch_bounds__(kBlock) void k_update(size_t n, double *__restrict__ w, const double *__restrict__ Y,
                                                   size_t ldy, int ny, const double *__restrict__ h) {
  const size_t i = blockIdx.x * (size_t)kBlock + threadIdx.x;
  if (i >= n) return;
  double s = w[i];
  for (int k = 0; k < ny; ++k) s -= h[k] * Y[(size_t)k * ldy + i];
  w[i] = s;
}

// y = a x + b y
__global__ __launch_bounds__(kBlock) void k_axpby(size_t n, double a, const double *__restrict__ x, double b,
                                                  double *__restrict__ y) {
  const size_t i = blockIdx.x * (size_t)kBlock + threadIdx.x;
  if (i >= n) return;
  y[i] = a * x[i] + b * y[i];
}

// w = sum_k c[k] Y_k (c passed by value through a small struct)
struct Coefs {
  double c[64];
};
__global__ __launch_bounds__(kBlock) void k_combine(size_t n, double *__restrict__ w, const double *__restrict__ Y,
                                                    size_t ldy, int ny, Coefs c) {
  const size_t i = blockIdx.x * (size_t)kBlock + threadIdx.x;
  if (i >= n) return;
  double s = 0.0;
  for (int k = 0; k < ny; ++k) s += c.c[k] * Y[(size_t)k * ldy + i];
  w[i] = s;
}

__global__ __launch_bounds__(kBlock) void k_fill(size_t n, double v, double *__restrict__ y) {
  const size_t i = blockIdx.x * (size_t)kBlock + threadIdx.x;
  if (i < n) y[i] = v;
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

struct Krylov {
  int m = 0;
  size_t n = 0;
  double *V = nullptr;        // [m+1][n] basis
  double *w = nullptr, *z = nullptr, *r = nullptr, *x = nullptr, *b = nullptr, *R = nullptr;
  double *partial = nullptr;  // [64][kDotBlocks]
  double *dots = nullptr;     // [128] device results
  double *h_dots = nullptr;   // pinned host copy
};

void free_krylov(aerofx_ctx *c) {
  if (!c->ksp) return;
  if (c->ksp->h_dots) (void)hipHostFree(c->ksp->h_dots);
  delete c->ksp;
  c->ksp = nullptr;  // device buffers belong to c->allocs
}

static int ensure_krylov(aerofx_ctx *c, int m) {
  if (m < 1 || m > 62) return set_error(AEROFX_ERR_ARG, "gmres: Krylov dimension must be in [1, 62]");
  if (c->ksp && c->ksp->m >= m) return 0;
  free_krylov(c);
  Krylov *k = new Krylov();
  k->m = m;
  k->n = 5 * (size_t)c->N;
  void *p = nullptr;
  int rc;
  if ((rc = ctx_alloc(c, sizeof(double) * (size_t)(m + 1) * k->n, &p))) { delete k; return rc; }
  k->V = (double *)p;
  double **vecs[6] = {&k->w, &k->z, &k->r, &k->x, &k->b, &k->R};
  for (double **v : vecs) {
    if ((rc = ctx_alloc(c, sizeof(double) * k->n, &p))) { delete k; return rc; }
    *v = (double *)p;
  }
  if ((rc = ctx_alloc(c, sizeof(double) * 64 * kDotBlocks, &p))) { delete k; return rc; }
  k->partial = (double *)p;
  if ((rc = ctx_alloc(c, sizeof(double) * 128, &p))) { delete k; return rc; }
  k->dots = (double *)p;
  hipError_t e = hipHostMalloc((void **)&k->h_dots, sizeof(double) * 128, hipHostMallocDefault);
  if (e != hipSuccess) { delete k; return set_hip_error(e, "hipHostMalloc"); }
  c->ksp = k;
  return 0;
}

static inline dim3 grid_for(size_t n) { return dim3((unsigned)((n + kBlock - 1) / kBlock)); }

// dots[0..ny) = x . Y_k on the device (no synchronisation)
static void multidot(aerofx_ctx *c, const double *x, const double *Y, size_t ldy, int ny, double *out) {
  Krylov *k = c->ksp;
  dim3 grid(kDotBlocks, (ny + kDotGroup - 1) / kDotGroup);
  hipLaunchKernelGGL(k_multidot_partial, grid, dim3(kBlock), 0, c->stream, k->n, x, Y, ldy, ny, k->partial);
  hipLaunchKernelGGL(k_multidot_final, dim3(ny), dim3(kBlock), 0, c->stream, k->partial, out);
}

static int fetch(aerofx_ctx *c, const double *dev, int count) {
  hipError_t e = hipMemcpyAsync(c->ksp->h_dots, dev, sizeof(double) * count, hipMemcpyDeviceToHost, c->stream);
  if (e != hipSuccess) return set_hip_error(e, "hipMemcpyAsync");
  e = hipStreamSynchronize(c->stream);
  if (e != hipSuccess) return set_hip_error(e, "hipStreamSynchronize");
  return 0;
}

// ||x||_2 with the deterministic reduction (synchronises)
static int norm2(aerofx_ctx *c, const double *x, double *out) {
  multidot(c, x, x, 0, 1, c->ksp->dots);
  int rc = fetch(c, c->ksp->dots, 1);
  if (rc) return rc;
  *out = std::sqrt(c->ksp->h_dots[0]);
  return 0;
}

struct GmresOut {
  int its = 0;
  double res0 = 0, res = 0, target = 0;
};

// Right-preconditioned GMRES(m) on H1 with block-Jacobi: solves A x = b,
// x starts at 0 (ImplicitCoupledTsDesc::solveLinearSystem sets dQ = 0).
static int gmres(aerofx_ctx *c, const double *b, double *x, int m, double eps, int maxits, GmresOut *out) {
  int rc = ensure_krylov(c, m);
  if (rc) return rc;
  Krylov *k = c->ksp;
  const size_t n = k->n;
  std::vector<double> H((size_t)(m + 1) * m, 0.0), cs(2 * (size_t)m), g(m + 1), y(m);
  auto Hij = [&](int i, int j) -> double & { return H[(size_t)i * m + j]; };
  hipLaunchKernelGGL(k_fill, grid_for(n), dim3(kBlock), 0, c->stream, n, 0.0, x);
  double target = 0.0, l2res = 0.0;
  int iter = 0;
  bool exitLoop = false;
  do {
    // r = b - A x
    if ((rc = launch_h1_apply(c, x, k->w))) return rc;
    hipLaunchKernelGGL(k_axpby, grid_for(n), dim3(kBlock), 0, c->stream, n, 1.0, b, 0.0, k->r);
    hipLaunchKernelGGL(k_axpby, grid_for(n), dim3(kBlock), 0, c->stream, n, -1.0, k->w, 1.0, k->r);
    double beta;
    if ((rc = norm2(c, k->r, &beta))) return rc;
    if (iter == 0) {
      double bn;
      if ((rc = norm2(c, b, &bn))) return rc;
      target = eps * bn;
      out->res0 = bn;
    }
    if (beta == 0.0) break;
    // V0 = r / beta
    hipLaunchKernelGGL(k_axpby, grid_for(n), dim3(kBlock), 0, c->stream, n, 1.0 / beta, k->r, 0.0, k->V);
    std::fill(g.begin(), g.end(), 0.0);
    g[0] = beta;
    int j;
    for (j = 0; j < m; ++j) {
      double *Vj = k->V + (size_t)j * n;
      if ((rc = launch_prec_apply(c, Vj, k->z))) return rc;
      if ((rc = launch_h1_apply(c, k->z, k->w))) return rc;
      {
        ProfScope ps(c, AEROFX_TIMER_GRAM_SCHMIDT);
        // CGS2: two passes of h = V^T w, w -= V h; then ||w||
        multidot(c, k->w, k->V, n, j + 1, k->dots);
        hipLaunchKernelGGL(k_update, grid_for(n), dim3(kBlock), 0, c->stream, n, k->w, k->V, n, j + 1, k->dots);
        multidot(c, k->w, k->V, n, j + 1, k->dots + 64);
        hipLaunchKernelGGL(k_update, grid_for(n), dim3(kBlock), 0, c->stream, n, k->w, k->V, n, j + 1,
                           k->dots + 64);
        multidot(c, k->w, k->w, 0, 1, k->dots + 127);
      }
      if ((rc = fetch(c, k->dots, 128))) return rc;
      for (int i = 0; i <= j; ++i) Hij(i, j) = k->h_dots[i] + k->h_dots[64 + i];
      double hn = std::sqrt(k->h_dots[127]);
      Hij(j + 1, j) = hn;
      // rotations (KspSolver.C:919-962)
      for (int i = 0; i < j; ++i) {
        double hj = cs[2 * i + 1] * Hij(i, j) + cs[2 * i] * Hij(i + 1, j);
        double hjp1 = -cs[2 * i] * Hij(i, j) + cs[2 * i + 1] * Hij(i + 1, j);
        Hij(i, j) = hj;
        Hij(i + 1, j) = hjp1;
      }
      if (Hij(j, j) * Hij(j, j) == 0.0) {
        cs[2 * j + 1] = 0.0;
        cs[2 * j] = 1.0;
      } else {
        double coef = 1.0 / std::sqrt(Hij(j, j) * Hij(j, j) + Hij(j + 1, j) * Hij(j + 1, j));
        cs[2 * j + 1] = std::sqrt(Hij(j, j) * Hij(j, j)) * coef;
        cs[2 * j] = cs[2 * j + 1] * Hij(j + 1, j) / Hij(j, j);
      }
      Hij(j, j) = cs[2 * j + 1] * Hij(j, j) + cs[2 * j] * Hij(j + 1, j);
      Hij(j + 1, j) = 0.0;
      g[j + 1] = -cs[2 * j] * g[j];
      g[j] = cs[2 * j + 1] * g[j];
      if (hn == 0.0) {
        ++iter;
        exitLoop = true;
        break;
      }
      hipLaunchKernelGGL(k_axpby, grid_for(n), dim3(kBlock), 0, c->stream, n, 1.0 / hn, k->w, 0.0,
                         k->V + (size_t)(j + 1) * n);
      l2res = std::fabs(g[j + 1]);
      ++iter;
      if (l2res <= target || iter >= maxits) {
        exitLoop = true;
        break;
      }
    }
    if (j == m) --j;
    // back-substitution (KspSolver.C:968-982)
    for (int i = j; i >= 0; --i) {
      y[i] = g[i];
      for (int kk = j; kk >= i + 1; --kk) y[i] -= Hij(i, kk) * y[kk];
      y[i] /= Hij(i, i);
    }
    Coefs cf;
    for (int i = 0; i <= j; ++i) cf.c[i] = y[i];
    hipLaunchKernelGGL(k_combine, grid_for(n), dim3(kBlock), 0, c->stream, n, k->w, k->V, n, j + 1, cf);
    if ((rc = launch_prec_apply(c, k->w, k->r))) return rc;
    hipLaunchKernelGGL(k_axpby, grid_for(n), dim3(kBlock), 0, c->stream, n, 1.0, k->r, 1.0, x);
  } while (!exitLoop);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return set_hip_error(e, "gmres launch");
  out->its = iter;
  out->res = l2res;
  out->target = target;
  return 0;
}

}  // namespace aerofx

using namespace aerofx;

extern "C" {

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
  std::vector<char> hd(25 * esz * (size_t)c->N), hb(25 * esz * 2 * (size_t)c->E);
  hipError_t e = hipStreamSynchronize(c->stream);
  if (e == hipSuccess) e = hipMemcpy(hd.data(), c->d_jacD, hd.size(), hipMemcpyDeviceToHost);
  if (e == hipSuccess) e = hipMemcpy(hb.data(), c->d_jacB, hb.size(), hipMemcpyDeviceToHost);
  if (e != hipSuccess) return set_hip_error(e, "jacobian_blocks copy");
  auto at = [&](const std::vector<char> &v, size_t idx) -> double {
    return c->jac_precision ? (double)((const float *)v.data())[idx] : ((const double *)v.data())[idx];
  };
  if (D)
    for (size_t k = 0; k < 25 * (size_t)c->N; ++k) D[k] = at(hd, k);
  if (Aedge) {
    // adjacency entry e of node n -> edge layout [E][2][25] = (A_ij, A_ji)
    std::vector<int> rowptr, adj, eid;
    node_csr(c->N, c->E, c->ptr.data(), rowptr, adj, eid);
    for (int n = 0; n < c->N; ++n)
      for (int q = rowptr[n]; q < rowptr[n + 1]; ++q) {
        int l = eid[q];
        int side = n < adj[q] ? 0 : 1;
        for (int k = 0; k < 25; ++k) Aedge[50 * (size_t)l + 25 * side + k] = at(hb, 25 * (size_t)q + k);
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
    // spatial residual and its norm (TsDesc::computeResidualNorm)
    int ierr = launch_residual(c, U, k->R, true);
    if (ierr) return ierr;
    int cnt = 0;
    if ((rc = aerofx_error_count(c, &cnt))) return rc;
    if (cnt) {
      st->steps = step;
      return set_error(AEROFX_ERR_STATE, "steady_solve: non-physical reconstructed state");
    }
    if ((rc = norm2(c, k->R, &res))) return rc;
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
    hipLaunchKernelGGL(k_axpby, grid_for(n), dim3(kBlock), 0, c->stream, n, -1.0, k->R, 0.0, k->b);
    GmresOut go;
    if ((rc = gmres(c, k->b, k->x, o->gmres_m, o->gmres_eps, o->gmres_maxits, &go)))
      return rc;
    st->krylov_its += go.its;
    // Q += dQ
    hipLaunchKernelGGL(k_axpby, grid_for(n), dim3(kBlock), 0, c->stream, n, 1.0, k->x, 1.0, U);
    hipLaunchKernelGGL(k_zero_int, dim3(1), dim3(64), 0, c->stream, c->d_err + 2, 1);
    hipLaunchKernelGGL(k_check_states, dim3((c->N + kBlock - 1) / kBlock), dim3(kBlock), 0, c->stream, c->N, U,
                       c->gas, c->d_err + 2);
    hipError_t e = hipMemcpyAsync(c->h_err + 2, c->d_err + 2, sizeof(int), hipMemcpyDeviceToHost, c->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    if (e != hipSuccess) return set_hip_error(e, "steady_solve sync");
    if (c->h_err[2]) {
      st->steps = step + 1;
      return set_error(AEROFX_ERR_STATE, "steady_solve: update produced non-physical states");
    }
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
