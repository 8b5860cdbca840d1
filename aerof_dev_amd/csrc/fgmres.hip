// fgmres.hip -- flexible GMRES over up to 16 right-hand sides at once, for the
// derivative solves of the sensitivity handler: the exact second-order
// derivative as the operator (forward-mode products, or the hand-written
// adjoint for the transposed system) and, as a variable right
// preconditioner, an inexact solve with the first-order Jacobian H1 + 1/dt
// (GMRES for one right-hand side, block GMRES for several: every outer step
// preconditions all right-hand sides in one block solve).
//
// Each right-hand side keeps its own Krylov space (V_q, Z_q = M^-1 V_q); the
// Gram-Schmidt is classical with one reorthogonalisation (two passes, the
// second folded with the norm), reductions in a fixed order, so results are
// deterministic.  The small Hessenberg problems (Givens rotations, back
// substitution) run on the host between steps.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <vector>

#include "context.h"
#include "host_internal.h"

namespace aerofx {

namespace {

constexpr int kFgBlock = 256;
constexpr int kFgMaxK = 32;    // outer basis vectors per right-hand side
constexpr int kFgParts = 256;  // blocks of the dot-product sweeps

inline dim3 fg_grid(size_t n, int nq) { return dim3((unsigned)((n + kFgBlock - 1) / kFgBlock), (unsigned)nq); }

// Slice [b0, b1) of block b in a fixed partition of n into parts.
__device__ __forceinline__ void fg_slice(size_t n, int parts, int b, size_t &b0, size_t &b1) {
  const size_t per = (n + parts - 1) / parts;
  b0 = std::min(n, (size_t)b * per);
  b1 = std::min(n, b0 + per);
}

// Fixed-order block sum of v over kFgBlock threads (LDS tree); result in thread 0.
__device__ __forceinline__ double fg_block_sum(double v, double *sh) {
  const int t = threadIdx.x;
  sh[t] = v;
  __syncthreads();
  for (int s = kFgBlock / 2; s > 0; s >>= 1) {
    if (t < s) sh[t] += sh[t + s];
    __syncthreads();
  }
  const double r = sh[0];
  __syncthreads();
  return r;
}

// partial[(q * kFgMaxK + k) * parts + b] = V_k,q . W_q over block b's slice,
// k < nk; V_k,q at V + k * ldk + q * n.  With nk == 0: the squared norm of
// W_q into slot 0.
__global__ __launch_bounds__(kFgBlock) void k_fg_dots(size_t n, int nk, const double *__restrict__ V, size_t ldk,
                                                      const double *__restrict__ W, int parts,
                                                      double *__restrict__ partial) {
  __shared__ double sh[kFgBlock];
  const int b = blockIdx.x, q = blockIdx.y;
  size_t b0, b1;
  fg_slice(n, parts, b, b0, b1);
  const double *w = W + (size_t)q * n;
  const int kk = nk > 0 ? nk : 1;
  for (int k = 0; k < kk; ++k) {
    const double *v = nk > 0 ? V + (size_t)k * ldk + (size_t)q * n : w;
    double s = 0.0;
    for (size_t i = b0 + threadIdx.x; i < b1; i += kFgBlock) s += v[i] * w[i];
    const double r = fg_block_sum(s, sh);
    if (threadIdx.x == 0) partial[((size_t)q * kFgMaxK + k) * parts + b] = r;
  }
}

// out[q * kFgMaxK + k] = sum_b partial[...] in block order (one thread each)
__global__ void k_fg_reduce(int nq, int nk, int parts, const double *__restrict__ partial, double *__restrict__ out) {
  const int idx = blockIdx.x * blockDim.x + threadIdx.x;
  const int kk = nk > 0 ? nk : 1;
  if (idx >= nq * kk) return;
  const int q = idx / kk, k = idx - q * kk;
  const double *p = partial + ((size_t)q * kFgMaxK + k) * parts;
  double s = 0.0;
  for (int b = 0; b < parts; ++b) s += p[b];
  out[q * kFgMaxK + k] = s;
}

// W_q -= sum_{k < nk} h[q][k] V_k,q  (k ascending)
__global__ __launch_bounds__(kFgBlock) void k_fg_update(size_t n, int nk, const double *__restrict__ V, size_t ldk,
                                                        const double *__restrict__ h, double *__restrict__ W) {
  const size_t i = blockIdx.x * (size_t)kFgBlock + threadIdx.x;
  const int q = blockIdx.y;
  if (i >= n) return;
  double w = W[(size_t)q * n + i];
  for (int k = 0; k < nk; ++k) w -= h[q * kFgMaxK + k] * V[(size_t)k * ldk + (size_t)q * n + i];
  W[(size_t)q * n + i] = w;
}

// dst_q = src_{sel[q]} * scale[q]
__global__ __launch_bounds__(kFgBlock) void k_fg_scale(size_t n, const double *__restrict__ src,
                                                       const int *__restrict__ sel, const double *__restrict__ scale,
                                                       double *__restrict__ dst) {
  const size_t i = blockIdx.x * (size_t)kFgBlock + threadIdx.x;
  const int q = blockIdx.y;
  if (i >= n) return;
  dst[(size_t)q * n + i] = src[(size_t)sel[q] * n + i] * scale[q];
}

// X_{sel[q]} += sum_{k < nk} y[q][k] Z_k,q  (k ascending)
__global__ __launch_bounds__(kFgBlock) void k_fg_combine(size_t n, int nk, const double *__restrict__ Z, size_t ldk,
                                                         const double *__restrict__ y, const int *__restrict__ sel,
                                                         double *__restrict__ X) {
  const size_t i = blockIdx.x * (size_t)kFgBlock + threadIdx.x;
  const int q = blockIdx.y;
  if (i >= n) return;
  double acc = 0.0;
  for (int k = 0; k < nk; ++k) acc += y[q * kFgMaxK + k] * Z[(size_t)k * ldk + (size_t)q * n + i];
  X[(size_t)sel[q] * n + i] += acc;
}

// R_q = B_q - A X_q
__global__ __launch_bounds__(kFgBlock) void k_fg_residual(size_t n, const double *__restrict__ B,
                                                          const double *__restrict__ AX, double *__restrict__ R) {
  const size_t i = blockIdx.x * (size_t)kFgBlock + threadIdx.x;
  const int q = blockIdx.y;
  if (i >= n) return;
  R[(size_t)q * n + i] = B[(size_t)q * n + i] - AX[(size_t)q * n + i];
}

__global__ __launch_bounds__(kFgBlock) void k_fg_fill0(size_t n, double *__restrict__ X) {
  const size_t i = blockIdx.x * (size_t)kFgBlock + threadIdx.x;
  if (i < n) X[(size_t)blockIdx.y * n + i] = 0.0;
}

}  // namespace

struct FgWork {
  int nq = 0, mo = 0;
  size_t n = 0;
  double *V = nullptr, *Z = nullptr;  // [mo+1][nq][n], [mo][nq][n]
  double *W = nullptr, *R = nullptr;  // [nq][n] each
  double *partial = nullptr;          // [nq][kFgMaxK][kFgParts]
  double *h1 = nullptr, *h2 = nullptr, *nrm = nullptr;  // [nq][kFgMaxK]
  double *y = nullptr, *scale = nullptr;                // [nq][kFgMaxK], [nq]
  int *sel = nullptr;                                   // [nq]
  double *hbuf = nullptr;                               // pinned host [3][nq][kFgMaxK]
};

void free_fgmres(aerofx_ctx *c) {
  FgWork *f = c->fgm;
  if (!f) return;
  if (f->hbuf) (void)hipHostFree(f->hbuf);
  for (void *p : {(void *)f->V, (void *)f->Z, (void *)f->W, (void *)f->partial, (void *)f->h1})
    if (p) ctx_release(c, p);
  delete f;
  c->fgm = nullptr;
}

namespace {

int ensure_fg(aerofx_ctx *c, int nq, int mo, size_t n) {
  if (c->fgm && c->fgm->nq >= nq && c->fgm->mo >= mo && c->fgm->n == n) return 0;
  free_fgmres(c);
  FgWork *f = new FgWork();
  f->nq = nq;
  f->mo = mo;
  f->n = n;
  c->fgm = f;  // free_fgmres releases whatever of it exists on a failure below
  void *p = nullptr;
  int rc;
  auto fail = [&](int r) {
    free_fgmres(c);
    return r;
  };
  if ((rc = ctx_alloc(c, sizeof(double) * (size_t)(mo + 1) * nq * n, &p))) return fail(rc);
  f->V = (double *)p;
  if ((rc = ctx_alloc(c, sizeof(double) * (size_t)mo * nq * n, &p))) return fail(rc);
  f->Z = (double *)p;
  if ((rc = ctx_alloc(c, sizeof(double) * 2 * (size_t)nq * n, &p))) return fail(rc);
  f->W = (double *)p;
  f->R = f->W + (size_t)nq * n;
  if ((rc = ctx_alloc(c, sizeof(double) * (size_t)nq * kFgMaxK * kFgParts, &p))) return fail(rc);
  f->partial = (double *)p;
  if ((rc = ctx_alloc(c, sizeof(double) * (size_t)nq * kFgMaxK * 5 + sizeof(int) * 16, &p))) return fail(rc);
  f->h1 = (double *)p;
  f->h2 = f->h1 + (size_t)nq * kFgMaxK;
  f->nrm = f->h2 + (size_t)nq * kFgMaxK;
  f->y = f->nrm + (size_t)nq * kFgMaxK;
  f->scale = f->y + (size_t)nq * kFgMaxK;
  f->sel = (int *)(f->scale + (size_t)nq * kFgMaxK);
  hipError_t e = hipHostMalloc((void **)&f->hbuf, sizeof(double) * 3 * (size_t)nq * kFgMaxK, hipHostMallocDefault);
  if (e != hipSuccess) {
    f->hbuf = nullptr;
    return fail(set_hip_error(e, "fgmres host buffer"));
  }
  return 0;
}

}  // namespace

int fgmres_solve(aerofx_ctx *c, int nrhs, const double *rhs, double *x, int mo, double eps, int max_outer,
                 const FgOps &ops, FgResult *res) {
  if (nrhs < 1 || nrhs > 16) return set_error(AEROFX_ERR_ARG, "fgmres: 1 <= nrhs <= 16");
  if (mo < 1 || mo > kFgMaxK - 1) return set_error(AEROFX_ERR_ARG, "fgmres: outer restart in [1, 31]");
  const size_t n = 5 * (size_t)c->N;
  int rc = ensure_fg(c, nrhs, mo, n);
  if (rc) return rc;
  FgWork &f = *c->fgm;
  const size_t ldk = (size_t)nrhs * n;  // stride between basis slots: all right-hand sides of slot k
  hipStream_t s = c->stream;
  *res = FgResult{};
  hipError_t e;
  auto sync_copy = [&](void *dst, const void *src, size_t bytes) -> int {
    hipError_t ee = hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, s);
    if (ee == hipSuccess) ee = hipStreamSynchronize(s);
    return ee == hipSuccess ? 0 : set_hip_error(ee, "fgmres sync");
  };
  // squared norms of the nq vectors at base ([nq][n]) -> host out[q]
  auto norms2 = [&](const double *base, int nq, std::vector<double> &out) -> int {
    hipLaunchKernelGGL(k_fg_dots, dim3(kFgParts, nq), dim3(kFgBlock), 0, s, n, 0, nullptr, (size_t)0, base, kFgParts,
                       f.partial);
    hipLaunchKernelGGL(k_fg_reduce, dim3((nq + 63) / 64), dim3(64), 0, s, nq, 0, kFgParts, f.partial, f.nrm);
    int r = sync_copy(f.hbuf, f.nrm, sizeof(double) * (size_t)nq * kFgMaxK);
    if (r) return r;
    out.resize(nq);
    for (int q = 0; q < nq; ++q) out[q] = f.hbuf[q * kFgMaxK];
    return 0;
  };
  std::vector<double> bnorm, rnorm;
  if ((rc = norms2(rhs, nrhs, bnorm))) return rc;
  for (double &v : bnorm) v = std::sqrt(v);
  hipLaunchKernelGGL(k_fg_fill0, fg_grid(n, nrhs), dim3(kFgBlock), 0, s, n, x);
  int outer = 0;
  std::vector<int> sel;
  std::vector<double> H, g, cs, sn;
  for (int cycle = 0;; ++cycle) {
    // R = rhs - A x (x = 0 on the first cycle)
    if (cycle == 0) {
      e = hipMemcpyAsync(f.R, rhs, sizeof(double) * nrhs * n, hipMemcpyDeviceToDevice, s);
      if (e != hipSuccess) return set_hip_error(e, "fgmres copy");
    } else {
      if ((rc = ops.apply(x, f.W, nrhs))) return rc;
      hipLaunchKernelGGL(k_fg_residual, fg_grid(n, nrhs), dim3(kFgBlock), 0, s, n, rhs, f.W, f.R);
    }
    if ((rc = norms2(f.R, nrhs, rnorm))) return rc;
    double worst = 0.0;
    sel.clear();
    for (int q = 0; q < nrhs; ++q) {
      const double r = std::sqrt(rnorm[q]);
      const double rel = bnorm[q] > 0.0 ? r / bnorm[q] : 0.0;
      if (rel >= worst) {
        worst = rel;
        res->res = r;
        res->res0 = bnorm[q];
      }
      if (bnorm[q] > 0.0 && r > eps * bnorm[q]) sel.push_back(q);
    }
    if (sel.empty()) {
      res->converged = 1;
      break;
    }
    if (outer >= max_outer) break;
    const int na = (int)sel.size();
    // V_0,a = R_sel[a] / ||R_sel[a]||
    std::vector<double> beta(na);
    for (int a = 0; a < na; ++a) {
      beta[a] = std::sqrt(rnorm[sel[a]]);
      f.hbuf[a] = 1.0 / beta[a];
    }
    e = hipMemcpyAsync(f.scale, f.hbuf, sizeof(double) * na, hipMemcpyHostToDevice, s);
    if (e == hipSuccess) e = hipMemcpyAsync(f.sel, sel.data(), sizeof(int) * na, hipMemcpyHostToDevice, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);  // the host buffer is reused below
    if (e != hipSuccess) return set_hip_error(e, "fgmres start copy");
    const size_t lda = (size_t)na * n;  // basis slot stride for this cycle's active set
    hipLaunchKernelGGL(k_fg_scale, fg_grid(n, na), dim3(kFgBlock), 0, s, n, f.R, f.sel, f.scale, f.V);
    H.assign((size_t)na * (mo + 1) * mo, 0.0);
    g.assign((size_t)na * (mo + 1), 0.0);
    cs.assign((size_t)na * mo, 0.0);
    sn.assign((size_t)na * mo, 0.0);
    for (int a = 0; a < na; ++a) g[(size_t)a * (mo + 1)] = beta[a];
    int kused = 0;
    for (int k = 0; k < mo; ++k) {
      double *Vk = f.V + (size_t)k * lda, *Zk = f.Z + (size_t)k * lda;
      // Z_k = M^-1 V_k (all active right-hand sides in one inexact solve)
      int inner = 0;
      if ((rc = ops.prec(Vk, Zk, na, &inner))) return rc;
      res->inner_its += inner;
      // W = A Z_k, exact products
      if ((rc = ops.apply(Zk, f.W, na))) return rc;
      // two Gram-Schmidt passes against V_0..V_k, the second with ||W||^2
      hipLaunchKernelGGL(k_fg_dots, dim3(kFgParts, na), dim3(kFgBlock), 0, s, n, k + 1, f.V, lda, f.W, kFgParts,
                         f.partial);
      hipLaunchKernelGGL(k_fg_reduce, dim3((na * (k + 1) + 63) / 64), dim3(64), 0, s, na, k + 1, kFgParts, f.partial,
                         f.h1);
      hipLaunchKernelGGL(k_fg_update, fg_grid(n, na), dim3(kFgBlock), 0, s, n, k + 1, f.V, lda, f.h1, f.W);
      hipLaunchKernelGGL(k_fg_dots, dim3(kFgParts, na), dim3(kFgBlock), 0, s, n, k + 1, f.V, lda, f.W, kFgParts,
                         f.partial);
      hipLaunchKernelGGL(k_fg_reduce, dim3((na * (k + 1) + 63) / 64), dim3(64), 0, s, na, k + 1, kFgParts, f.partial,
                         f.h2);
      hipLaunchKernelGGL(k_fg_update, fg_grid(n, na), dim3(kFgBlock), 0, s, n, k + 1, f.V, lda, f.h2, f.W);
      hipLaunchKernelGGL(k_fg_dots, dim3(kFgParts, na), dim3(kFgBlock), 0, s, n, 0, nullptr, (size_t)0, f.W,
                         kFgParts, f.partial);
      hipLaunchKernelGGL(k_fg_reduce, dim3((na + 63) / 64), dim3(64), 0, s, na, 0, kFgParts, f.partial, f.nrm);
      e = hipGetLastError();
      if (e != hipSuccess) return set_hip_error(e, "fgmres launch");
      // h1, h2, ||W||^2 to the host (contiguous on the device, capacity f.nq)
      if ((rc = sync_copy(f.hbuf, f.h1, sizeof(double) * 3 * (size_t)f.nq * kFgMaxK))) return rc;
      ++outer;
      kused = k + 1;
      bool all_converged = true, breakdown = false;
      std::vector<double> inv_norm(na, 0.0);
      for (int a = 0; a < na; ++a) {
        const double *p1 = f.hbuf + (size_t)a * kFgMaxK, *p2 = f.hbuf + (size_t)(f.nq + a) * kFgMaxK;
        const double wn2 = f.hbuf[(size_t)(2 * f.nq + a) * kFgMaxK];
        double *Hc = H.data() + (size_t)a * (mo + 1) * mo;  // column-major: Hc[i + (mo+1) j]
        double hsq = 0.0;
        for (int i = 0; i <= k; ++i) {
          Hc[i + (size_t)(mo + 1) * k] = p1[i] + p2[i];
          hsq += (p1[i] + p2[i]) * (p1[i] + p2[i]);
        }
        const double hn = wn2 > 0.0 ? std::sqrt(wn2) : 0.0;
        Hc[k + 1 + (size_t)(mo + 1) * k] = hn;
        // W in the span of V_0..V_k (to rounding): the space holds the solution
        if (hn <= 1e-14 * std::sqrt(hsq + wn2)) breakdown = true;
        inv_norm[a] = hn > 0.0 ? 1.0 / hn : 0.0;
        // Givens rotations (KspSolver.C:919-966): the previous ones, then a new one
        for (int i = 0; i < k; ++i) {
          const double c0 = cs[(size_t)a * mo + i], s0 = sn[(size_t)a * mo + i];
          const double h0 = Hc[i + (size_t)(mo + 1) * k], h1v = Hc[i + 1 + (size_t)(mo + 1) * k];
          Hc[i + (size_t)(mo + 1) * k] = c0 * h0 + s0 * h1v;
          Hc[i + 1 + (size_t)(mo + 1) * k] = -s0 * h0 + c0 * h1v;
        }
        const double hkk = Hc[k + (size_t)(mo + 1) * k], hk1 = Hc[k + 1 + (size_t)(mo + 1) * k];
        const double den = std::sqrt(hkk * hkk + hk1 * hk1);
        const double c1 = den > 0.0 ? hkk / den : 1.0, s1 = den > 0.0 ? hk1 / den : 0.0;
        cs[(size_t)a * mo + k] = c1;
        sn[(size_t)a * mo + k] = s1;
        Hc[k + (size_t)(mo + 1) * k] = c1 * hkk + s1 * hk1;
        Hc[k + 1 + (size_t)(mo + 1) * k] = 0.0;
        double *ga = g.data() + (size_t)a * (mo + 1);
        ga[k + 1] = -s1 * ga[k];
        ga[k] = c1 * ga[k];
        if (std::fabs(ga[k + 1]) > eps * bnorm[sel[a]]) all_converged = false;
      }
      if (all_converged || breakdown || k + 1 == mo || outer >= max_outer) break;
      // V_{k+1} = W / ||W||
      std::vector<int> ident(na);
      for (int a = 0; a < na; ++a) ident[a] = a;
      e = hipMemcpyAsync(f.scale, inv_norm.data(), sizeof(double) * na, hipMemcpyHostToDevice, s);
      if (e == hipSuccess) e = hipMemcpyAsync(f.sel, ident.data(), sizeof(int) * na, hipMemcpyHostToDevice, s);
      if (e == hipSuccess) e = hipStreamSynchronize(s);
      if (e != hipSuccess) return set_hip_error(e, "fgmres scale copy");
      hipLaunchKernelGGL(k_fg_scale, fg_grid(n, na), dim3(kFgBlock), 0, s, n, f.W, f.sel, f.scale,
                         f.V + (size_t)(k + 1) * lda);
    }
    // back substitution per right-hand side and x_sel += Z y
    std::vector<double> yall((size_t)na * kFgMaxK, 0.0);
    for (int a = 0; a < na; ++a) {
      const double *Hc = H.data() + (size_t)a * (mo + 1) * mo;
      const double *ga = g.data() + (size_t)a * (mo + 1);
      double *ya = yall.data() + (size_t)a * kFgMaxK;
      for (int i = kused - 1; i >= 0; --i) {
        double v = ga[i];
        for (int j = i + 1; j < kused; ++j) v -= Hc[i + (size_t)(mo + 1) * j] * ya[j];
        const double d = Hc[i + (size_t)(mo + 1) * i];
        ya[i] = d != 0.0 ? v / d : 0.0;
      }
    }
    e = hipMemcpyAsync(f.y, yall.data(), sizeof(double) * yall.size(), hipMemcpyHostToDevice, s);
    if (e == hipSuccess) e = hipMemcpyAsync(f.sel, sel.data(), sizeof(int) * na, hipMemcpyHostToDevice, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    if (e != hipSuccess) return set_hip_error(e, "fgmres combine copy");
    hipLaunchKernelGGL(k_fg_combine, fg_grid(n, na), dim3(kFgBlock), 0, s, n, kused, f.Z, lda, f.y, f.sel, x);
  }
  res->outer_its = outer;
  e = hipStreamSynchronize(s);
  return e == hipSuccess ? 0 : set_hip_error(e, "fgmres sync");
}

}  // namespace aerofx
