// tangent.hip -- forward-mode derivative of the residual on gfx950:
//   dR = (dR/dU) dU + (dR/dX) dX
// for the second-order residual R_i = (1/V_i)(sum_edges +-F(V_ij, V_ji, n_ij)
// + sum_faces F_b(V_i, n_f/3)).  This is the directional derivative
// SpaceOperator::computeDerivativeOfResidual (SpaceOperator.C:917-1070) and
// EdgeSet::computeDerivativeOfFiniteVolumeTerm (Edge.C:1494-1573) form for a
// shape perturbation dX (with dU the state perturbation: the exact product of
// the second-order Jacobian that MatVecProdH2 applies, MatVecProd.C:1293+).
// Every quantity the residual computes from X is differentiated: dual-face
// normals and control volumes (per tet, ElemTetCore.C:31-44, 281-330), face
// normals (FaceTriaCore.C:31-37), LSQ moments, Cholesky factors and weights,
// the gradients, the reconstruction distances and the fluxes.
//
// The numbers are Dual (value, derivative along the one seeded direction)
// pushed through the same templated routines the residual uses.  Stages:
//   1. forward residual: V, LSQ gradients, raw flux sums S (the MUSCL
//      gradients are recomputed, nothing from an earlier call is kept)
//   2. geometry: dn per edge (its tets in tet order), dV_i per node (its tets
//      in tet order), dn_f per boundary face
//   3. per node: dV(U), the LSQ moments and factors and the gradients in Dual
//   4. per edge: reconstruction and Roe flux in Dual -> dF
//   5. per node: dS = sum over upper edges of dF - sum over lower edges of dF
//      + face fluxes in Dual; dR = dS / V - S dV / V^2
// Sums are in fixed orders: deterministic from run to run.
#include <hip/hip_runtime.h>

#include "context.h"
#include "device_dual.h"
#include "device_sens.h"
#include "host_internal.h"

namespace aerofx {

struct Tangent {
  double *S = nullptr;    // [N][5] raw flux sums of the forward pass
  double *dV = nullptr;   // [N][5]
  double *dG = nullptr;   // [N][16]
  double *dn = nullptr;   // [E][3]
  double *dcv = nullptr;  // [N]
  double *dfn = nullptr;  // [F][3] full face normal
  double *dF = nullptr;   // [E][5]
  // the state part of the exact product operator (MatVecProdH2), kept from
  // its selection: primitive state [N][5] and LSQ gradients [N][16]
  double *V0 = nullptr, *G0 = nullptr;
};

void free_tangent(aerofx_ctx *c) {
  delete c->tan;  // device memory is owned by the context allocation list
  c->tan = nullptr;
}

namespace {

constexpr int kTanBlock = 256;
// ElemTet edge -> (end, end) and edge -> (face, face) tables (ElemTetCore.C:19-21)
constexpr int kEdgeEnd[6][2] = {{0, 1}, {0, 2}, {0, 3}, {1, 2}, {1, 3}, {2, 3}};
constexpr int kEdgeFace[6][2] = {{0, 1}, {2, 0}, {1, 2}, {0, 3}, {3, 1}, {2, 3}};

__device__ __forceinline__ Dual coord(const double *__restrict__ X4, const double *__restrict__ dX, int n, int a) {
  return Dual(X4[4 * (size_t)n + a], dX ? dX[3 * (size_t)n + a] : 0.0);
}

__device__ __forceinline__ void load_tet(const int *__restrict__ tets, size_t t, const double *__restrict__ X4,
                                         const double *__restrict__ dX, int nd[4], Dual x[4][3]) {
#pragma unroll
  for (int v = 0; v < 4; ++v) {
    nd[v] = tets[4 * t + v];
#pragma unroll
    for (int a = 0; a < 3; ++a) x[v][a] = coord(X4, dX, nd[v], a);
  }
}

// ---------------------------------------------------------------------------
// 2. geometry
// ---------------------------------------------------------------------------
// dn_l: the tets containing edge l in tet order; each adds its dual-face
// normal of the edge (ElemTet::computeEdgeNormalsGCL1 with Xn == Xnp1,
// ElemTetCore.C:281-330), oriented from the lower to the higher node id
__global__ __launch_bounds__(kTanBlock) void k_tan_edge_normals(int E, const int *__restrict__ etrow,
                                                                const int *__restrict__ etslot,
                                                                const int *__restrict__ tets,
                                                                const double *__restrict__ X4,
                                                                const double *__restrict__ dX,
                                                                double *__restrict__ dn) {
  const int l = blockIdx.x * kTanBlock + threadIdx.x;
  if (l >= E) return;
  double acc[3] = {0.0, 0.0, 0.0};
  const double third = 1.0 / 3.0;
  for (int q = etrow[l]; q < etrow[l + 1]; ++q) {
    const size_t s = etslot[q];
    const size_t t = s / 6;
    const int e = (int)(s - 6 * t);
    int nd[4];
    Dual x[4][3];
    load_tet(tets, t, X4, dX, nd, x);
    Dual f[4][3], gc[3];
#pragma unroll
    for (int a = 0; a < 3; ++a) {
      f[0][a] = third * ((x[0][a] + x[1][a]) + x[2][a]);
      f[1][a] = third * ((x[0][a] + x[1][a]) + x[3][a]);
      f[2][a] = third * ((x[0][a] + x[2][a]) + x[3][a]);
      f[3][a] = third * ((x[1][a] + x[2][a]) + x[3][a]);
      gc[a] = 0.25 * (((x[0][a] + x[1][a]) + x[2][a]) + x[3][a]);
    }
    int fa = 0, fb = 0, ea = 0, eb = 0;
#pragma unroll
    for (int k = 0; k < 6; ++k)
      if (k == e) {
        fa = kEdgeFace[k][1];
        fb = kEdgeFace[k][0];
        ea = kEdgeEnd[k][0];
        eb = kEdgeEnd[k][1];
      }
    Dual p[3], r[3];
#pragma unroll
    for (int a = 0; a < 3; ++a) {
      p[a] = f[fa][a] - gc[a];
      r[a] = f[fb][a] - gc[a];
    }
    Dual A[3] = {p[1] * r[2] - p[2] * r[1], p[2] * r[0] - p[0] * r[2], p[0] * r[1] - p[1] * r[0]};
    const double sg = nd[ea] < nd[eb] ? 1.0 : -1.0;
#pragma unroll
    for (int a = 0; a < 3; ++a) acc[a] += sg * (0.5 * (A[a] + A[a]) + 0.25 * (A[a] + A[a])).d;
  }
#pragma unroll
  for (int a = 0; a < 3; ++a) dn[3 * (size_t)l + a] = acc[a];
}

// dV_i: the tets around node i in tet order, a quarter of each tet's volume
// derivative (ElemTet::computeVolume, ElemTetCore.C:31-44)
__global__ __launch_bounds__(kTanBlock) void k_tan_cv(int N, const int *__restrict__ ntrow,
                                                      const int *__restrict__ ntslot, const int *__restrict__ tets,
                                                      const double *__restrict__ X4, const double *__restrict__ dX,
                                                      double *__restrict__ dcv) {
  const int n = blockIdx.x * kTanBlock + threadIdx.x;
  if (n >= N) return;
  const double sixth = 1.0 / 6.0;
  double acc = 0.0;
  for (int q = ntrow[n]; q < ntrow[n + 1]; ++q) {
    const size_t t = (size_t)ntslot[q] >> 2;
    int nd[4];
    Dual x[4][3];
    load_tet(tets, t, X4, dX, nd, x);
    Dual v1[3], v2[3], v3[3];
#pragma unroll
    for (int a = 0; a < 3; ++a) {
      v1[a] = x[1][a] - x[0][a];
      v2[a] = x[2][a] - x[0][a];
      v3[a] = x[3][a] - x[0][a];
    }
    Dual c0 = v1[1] * v2[2] - v1[2] * v2[1];
    Dual c1 = v1[2] * v2[0] - v1[0] * v2[2];
    Dual c2 = v1[0] * v2[1] - v1[1] * v2[0];
    Dual vol = sixth * (v3[0] * c0 + v3[1] * c1 + v3[2] * c2);
    acc += 0.25 * vol.d;
  }
  dcv[n] = acc;
}

// dn_f = d(0.5 (x2 - x0) x (x1 - x0)) (FaceTria::computeNormal, FaceTriaCore.C:31-37)
__global__ __launch_bounds__(kTanBlock) void k_tan_faces(int F, const int *__restrict__ faces,
                                                         const double *__restrict__ X4,
                                                         const double *__restrict__ dX, double *__restrict__ dfn) {
  const int f = blockIdx.x * kTanBlock + threadIdx.x;
  if (f >= F) return;
  Dual x[3][3];
#pragma unroll
  for (int v = 0; v < 3; ++v) {
    const int nd = faces[3 * (size_t)f + v];
#pragma unroll
    for (int a = 0; a < 3; ++a) x[v][a] = coord(X4, dX, nd, a);
  }
  Dual p[3], q[3];
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    p[a] = x[2][a] - x[0][a];
    q[a] = x[1][a] - x[0][a];
  }
  dfn[3 * (size_t)f] = (0.5 * (p[1] * q[2] - p[2] * q[1])).d;
  dfn[3 * (size_t)f + 1] = (0.5 * (p[2] * q[0] - p[0] * q[2])).d;
  dfn[3 * (size_t)f + 2] = (0.5 * (p[0] * q[1] - p[1] * q[0])).d;
}

__global__ __launch_bounds__(kTanBlock) void k_fill0(size_t n, double *__restrict__ y) {
  const size_t i = blockIdx.x * (size_t)kTanBlock + threadIdx.x;
  if (i < n) y[i] = 0.0;
}

// ---------------------------------------------------------------------------
// 3. per node: dV and, at second order, the gradients in Dual
// ---------------------------------------------------------------------------
__device__ __forceinline__ void load_v(const double *__restrict__ U, const double *__restrict__ dU, int n,
                                       const Gas &gas, Dual v[5]) {
  Dual u[5];
#pragma unroll
  for (int k = 0; k < 5; ++k) u[k] = Dual(U[5 * (size_t)n + k], dU ? dU[5 * (size_t)n + k] : 0.0);
  prim_from_cons(u, gas, v);
}

template <bool SECOND>
__global__ __launch_bounds__(kTanBlock) void k_tan_grad(DevTables t, const double *__restrict__ U,
                                                        const double *__restrict__ dU,
                                                        const double *__restrict__ dX, Gas gas,
                                                        double *__restrict__ dV, double *__restrict__ dG,
                                                        const int *__restrict__ done) {
  if (done && *done) return;
  const int n = blockIdx.x * kTanBlock + threadIdx.x;
  if (n >= t.N) return;
  Dual Vn[5];
  load_v(U, dU, n, gas, Vn);
#pragma unroll
  for (int k = 0; k < 5; ++k) dV[5 * (size_t)n + k] = Vn[k].d;
  if (!SECOND) return;
  const Dual xn[3] = {coord(t.X4, dX, n, 0), coord(t.X4, dX, n, 1), coord(t.X4, dX, n, 2)};
  const int e0 = t.rowptr[n], e1 = t.rowptr[n + 1];
  // LSQ moments in edge order (SubDomainCore.C:1334-1381) and factors
  Dual M[6];
  for (int e = e0; e < e1; ++e) {
    const int m = t.adj[e];
    const Dual d0 = coord(t.X4, dX, m, 0) - xn[0], d1 = coord(t.X4, dX, m, 1) - xn[1],
               d2 = coord(t.X4, dX, m, 2) - xn[2];
    M[0] += d0 * d0;
    M[1] += d0 * d1;
    M[2] += d0 * d2;
    M[3] += d1 * d1;
    M[4] += d1 * d2;
    M[5] += d2 * d2;
  }
  Dual L[6];
  lsq_params(M, L);
  Dual acc[15];
  for (int e = e0; e < e1; ++e) {
    const int m = t.adj[e];
    Dual Vm[5], w[3];
    load_v(U, dU, m, gas, Vm);
    lsq_weights_t(L, coord(t.X4, dX, m, 0) - xn[0], coord(t.X4, dX, m, 1) - xn[1], coord(t.X4, dX, m, 2) - xn[2], w);
#pragma unroll
    for (int q = 0; q < 5; ++q) {
      const Dual d = Vm[q] - Vn[q];
      acc[q] += w[0] * d;
      acc[5 + q] += w[1] * d;
      acc[10 + q] += w[2] * d;
    }
  }
#pragma unroll
  for (int q = 0; q < 15; ++q) dG[16 * (size_t)n + q] = acc[q].d;
  dG[16 * (size_t)n + 15] = 0.0;
}

// ---------------------------------------------------------------------------
// 4. per edge: reconstruction and Roe flux in Dual
// ---------------------------------------------------------------------------
template <bool SECOND>
__global__ __launch_bounds__(kTanBlock) void k_tan_flux(DevTables t, const double *__restrict__ V,
                                                        const double *__restrict__ grad,
                                                        const double *__restrict__ dV,
                                                        const double *__restrict__ dG,
                                                        const double *__restrict__ dX,
                                                        const double *__restrict__ dn, FluxConsts fc,
                                                        double *__restrict__ dF, const int *__restrict__ done) {
  if (done && *done) return;
  const int l = blockIdx.x * kTanBlock + threadIdx.x;
  if (l >= t.E) return;
  const int2 ij = t.eij[l];
  const int i = ij.x, j = ij.y;
  Dual Vi[5], Vj[5], Vij[5], Vji[5];
#pragma unroll
  for (int k = 0; k < 5; ++k) {
    Vi[k] = Dual(V[5 * (size_t)i + k], dV[5 * (size_t)i + k]);
    Vj[k] = Dual(V[5 * (size_t)j + k], dV[5 * (size_t)j + k]);
  }
  if (SECOND) {
    const Dual dx0 = coord(t.X4, dX, j, 0) - coord(t.X4, dX, i, 0);
    const Dual dx1 = coord(t.X4, dX, j, 1) - coord(t.X4, dX, i, 1);
    const Dual dx2 = coord(t.X4, dX, j, 2) - coord(t.X4, dX, i, 2);
    const double *gi = grad + 16 * (size_t)i, *gj = grad + 16 * (size_t)j;
    const double *hi = dG + 16 * (size_t)i, *hj = dG + 16 * (size_t)j;
#pragma unroll
    for (int k = 0; k < 5; ++k) {
      const Dual ddVij = dx0 * Dual(gi[k], hi[k]) + dx1 * Dual(gi[5 + k], hi[5 + k]) +
                         dx2 * Dual(gi[10 + k], hi[10 + k]);
      const Dual ddVji = dx0 * Dual(gj[k], hj[k]) + dx1 * Dual(gj[5 + k], hj[5 + k]) +
                         dx2 * Dual(gj[10 + k], hj[10 + k]);
      vanalbada_t(fc.beta, Vi[k], ddVij, Vj[k], ddVji, Vij[k], Vji[k]);
    }
  } else {
#pragma unroll
    for (int k = 0; k < 5; ++k) {
      Vij[k] = Vi[k];
      Vji[k] = Vj[k];
    }
  }
  const Dual nv[3] = {Dual(t.normal[3 * (size_t)l], dn[3 * (size_t)l]),
                      Dual(t.normal[3 * (size_t)l + 1], dn[3 * (size_t)l + 1]),
                      Dual(t.normal[3 * (size_t)l + 2], dn[3 * (size_t)l + 2])};
  Dual f[5];
  if (sqrt(nv[0].v * nv[0].v + nv[1].v * nv[1].v + nv[2].v * nv[2].v) >= 1e-18) roeflux7(fc.gas, nv, Dual(0.0), Vij, Vji, f);
#pragma unroll
  for (int k = 0; k < 5; ++k) dF[5 * (size_t)l + k] = f[k].d;
}

// ---------------------------------------------------------------------------
// 5. per node: dS and dR
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(kTanBlock) void k_tan_nodes(DevTables t, const double *__restrict__ V,
                                                         const double *__restrict__ dV,
                                                         const double *__restrict__ S,
                                                         const double *__restrict__ dcv,
                                                         const double *__restrict__ dfn,
                                                         const double *__restrict__ dF, FluxConsts fc,
                                                         double *__restrict__ dR, const int *__restrict__ done) {
  if (done && *done) return;
  const int n = blockIdx.x * kTanBlock + threadIdx.x;
  if (n >= t.N) return;
  double s[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
  for (int l = t.eup[n]; l < t.eup[n + 1]; ++l)
#pragma unroll
    for (int k = 0; k < 5; ++k) s[k] += dF[5 * (size_t)l + k];
  for (int q = t.lrow[n]; q < t.lrow[n + 1]; ++q) {
    const size_t l = t.ledge[q];
#pragma unroll
    for (int k = 0; k < 5; ++k) s[k] -= dF[5 * l + k];
  }
  const int f0 = t.frow[n], f1 = t.frow[n + 1];
  if (f0 < f1) {
    Dual Vn[5];
#pragma unroll
    for (int k = 0; k < 5; ++k) Vn[k] = Dual(V[5 * (size_t)n + k], dV[5 * (size_t)n + k]);
    const double third = 1.0 / 3.0;
    for (int q = f0; q < f1; ++q) {
      const int face = t.fidx[q];
      Dual nv[3], fl[5];
#pragma unroll
      for (int a = 0; a < 3; ++a) nv[a] = Dual(t.fnorm3[3 * (size_t)face + a], third * dfn[3 * (size_t)face + a]);
      face_flux(fc, t.fcode[face], nv, Vn, fl);
#pragma unroll
      for (int k = 0; k < 5; ++k) s[k] += fl[k].d;
    }
  }
  // R = S * (1/V): dR = dS / V + S d(1/V) = dS / V - S dV / V^2
  const double ic = t.invcv[n], dic = -dcv[n] * ic * ic;
#pragma unroll
  for (int k = 0; k < 5; ++k) dR[5 * (size_t)n + k] = s[k] * ic + S[5 * (size_t)n + k] * dic;
}

// y += p / dt (the backward-Euler time term of the exact product, dt fixed)
__global__ __launch_bounds__(kTanBlock) void k_tan_time(size_t n, const double *__restrict__ p,
                                                        const double *__restrict__ dt, double *__restrict__ y,
                                                        const int *__restrict__ done) {
  if (done && *done) return;
  const size_t i = blockIdx.x * (size_t)kTanBlock + threadIdx.x;
  if (i >= n) return;
  y[i] += p[i] * (1.0 / dt[i / 5]);
}

template <class T>
int alloc_n(aerofx_ctx *c, size_t n, T **out) {
  void *p = nullptr;
  int rc = ctx_alloc(c, n * sizeof(T), &p);
  *out = static_cast<T *>(p);
  return rc;
}

int ensure_tangent(aerofx_ctx *c) {
  if (c->tan) return 0;
  int rc = ensure_elem_tables(c);
  if (rc) return rc;
  Tangent *g = new Tangent();
  const size_t N = c->N, E = c->E, F = std::max(c->F, 1);
  if ((rc = alloc_n(c, 5 * N, &g->S)) || (rc = alloc_n(c, 5 * N, &g->dV)) || (rc = alloc_n(c, 16 * N, &g->dG)) ||
      (rc = alloc_n(c, 3 * E, &g->dn)) || (rc = alloc_n(c, N, &g->dcv)) || (rc = alloc_n(c, 3 * F, &g->dfn)) ||
      (rc = alloc_n(c, 5 * E, &g->dF))) {
    delete g;
    return rc;
  }
  c->tan = g;
  return 0;
}

inline dim3 grid_for(long long n) { return dim3((unsigned)((n + kTanBlock - 1) / kTanBlock)); }

}  // namespace

int launch_tangent(aerofx_ctx *c, const double *U, const double *dU, const double *dX, double *dR) {
  int rc = ensure_tangent(c);
  if (rc) return rc;
  Tangent *g = c->tan;
  const ElemTables &el = *c->elem;
  const DevTables &t = c->tab;
  // 1. forward pass: V, gradients, raw flux sums
  if ((rc = launch_residual(c, U, g->S, false))) return rc;
  // 2. geometry
  if (dX) {
    if (t.E > 0)
      hipLaunchKernelGGL(k_tan_edge_normals, grid_for(t.E), dim3(kTanBlock), 0, c->stream, t.E, el.etrow, el.etslot,
                         el.tets, t.X4, dX, g->dn);
    hipLaunchKernelGGL(k_tan_cv, grid_for(t.N), dim3(kTanBlock), 0, c->stream, t.N, el.ntrow, el.ntslot, el.tets,
                       t.X4, dX, g->dcv);
    if (t.F > 0)
      hipLaunchKernelGGL(k_tan_faces, grid_for(t.F), dim3(kTanBlock), 0, c->stream, t.F, el.faces, t.X4, dX, g->dfn);
  } else {
    hipLaunchKernelGGL(k_fill0, grid_for(3LL * t.E), dim3(kTanBlock), 0, c->stream, 3 * (size_t)t.E, g->dn);
    hipLaunchKernelGGL(k_fill0, grid_for(t.N), dim3(kTanBlock), 0, c->stream, (size_t)t.N, g->dcv);
    hipLaunchKernelGGL(k_fill0, grid_for(3LL * t.F), dim3(kTanBlock), 0, c->stream, 3 * (size_t)t.F, g->dfn);
  }
  // 3.-5.
  const bool second = c->p.second_order != 0;
  if (second)
    hipLaunchKernelGGL(k_tan_grad<true>, grid_for(t.N), dim3(kTanBlock), 0, c->stream, t, U, dU, dX, c->gas, g->dV,
                       g->dG, nullptr);
  else
    hipLaunchKernelGGL(k_tan_grad<false>, grid_for(t.N), dim3(kTanBlock), 0, c->stream, t, U, dU, dX, c->gas, g->dV,
                       g->dG, nullptr);
  if (t.E > 0) {
    if (second)
      hipLaunchKernelGGL(k_tan_flux<true>, grid_for(t.E), dim3(kTanBlock), 0, c->stream, t, c->d_V, c->d_grad, g->dV,
                         g->dG, dX, g->dn, c->fc, g->dF, nullptr);
    else
      hipLaunchKernelGGL(k_tan_flux<false>, grid_for(t.E), dim3(kTanBlock), 0, c->stream, t, c->d_V, c->d_grad,
                         g->dV, g->dG, dX, g->dn, c->fc, g->dF, nullptr);
  }
  hipLaunchKernelGGL(k_tan_nodes, grid_for(t.N), dim3(kTanBlock), 0, c->stream, t, c->d_V, g->dV, g->S, g->dcv, g->dfn,
                     g->dF, c->fc, dR, nullptr);
  // the Navier-Stokes Galerkin term (state derivative)
  if ((rc = launch_visc_tangent(c, c->d_V, dU, dX, dR, nullptr))) return rc;
  hipError_t e = hipGetLastError();
  return e == hipSuccess ? 0 : set_hip_error(e, "tangent launch");
}

// MatVecProdH2::evaluate (MatVecProd.C:1560-1609) as a matrix-free exact
// product: the forward pass at Q (primitive state, LSQ gradients, raw flux
// sums) is kept in the tangent workspace, and the geometry derivatives are
// zero (the mesh does not move).
int launch_h2_select(aerofx_ctx *c, const double *Q) {
  int rc = ensure_tangent(c);
  if (rc) return rc;
  Tangent *g = c->tan;
  const DevTables &t = c->tab;
  if (!g->V0 && ((rc = alloc_n(c, 5 * (size_t)t.N, &g->V0)) || (rc = alloc_n(c, 16 * (size_t)t.N, &g->G0))))
    return rc;
  if ((rc = launch_residual(c, Q, g->S, false))) return rc;
  hipError_t e = hipMemcpyAsync(g->V0, c->d_V, 5 * sizeof(double) * (size_t)t.N, hipMemcpyDeviceToDevice, c->stream);
  if (e == hipSuccess)
    e = hipMemcpyAsync(g->G0, c->d_grad, 16 * sizeof(double) * (size_t)t.N, hipMemcpyDeviceToDevice, c->stream);
  if (e != hipSuccess) return set_hip_error(e, "h2 select copy");
  hipLaunchKernelGGL(k_fill0, grid_for(3LL * t.E), dim3(kTanBlock), 0, c->stream, 3 * (size_t)t.E, g->dn);
  hipLaunchKernelGGL(k_fill0, grid_for(t.N), dim3(kTanBlock), 0, c->stream, (size_t)t.N, g->dcv);
  hipLaunchKernelGGL(k_fill0, grid_for(3LL * t.F), dim3(kTanBlock), 0, c->stream, 3 * (size_t)t.F, g->dfn);
  e = hipGetLastError();
  return e == hipSuccess ? 0 : set_hip_error(e, "h2 select launch");
}

// MatVecProdH2::apply (MatVecProd.C:1647-1670): y = (dR/dU) p + p / dt at the
// selected Q (dt NULL: no time term); every kernel returns at once when
// *done != 0 (device-resident Krylov loops).
int launch_h2_apply(aerofx_ctx *c, const double *Q, const double *p, const double *dt, double *y,
                    const int *done) {
  Tangent *g = c->tan;
  if (!g || !g->V0) return set_error(AEROFX_ERR_STATE, "h2 apply: operator not selected");
  const DevTables &t = c->tab;
  if (c->p.second_order)
    hipLaunchKernelGGL(k_tan_grad<true>, grid_for(t.N), dim3(kTanBlock), 0, c->stream, t, Q, p, nullptr, c->gas,
                       g->dV, g->dG, done);
  else
    hipLaunchKernelGGL(k_tan_grad<false>, grid_for(t.N), dim3(kTanBlock), 0, c->stream, t, Q, p, nullptr, c->gas,
                       g->dV, g->dG, done);
  if (t.E > 0) {
    if (c->p.second_order)
      hipLaunchKernelGGL(k_tan_flux<true>, grid_for(t.E), dim3(kTanBlock), 0, c->stream, t, g->V0, g->G0, g->dV,
                         g->dG, nullptr, g->dn, c->fc, g->dF, done);
    else
      hipLaunchKernelGGL(k_tan_flux<false>, grid_for(t.E), dim3(kTanBlock), 0, c->stream, t, g->V0, g->G0, g->dV,
                         g->dG, nullptr, g->dn, c->fc, g->dF, done);
  }
  hipLaunchKernelGGL(k_tan_nodes, grid_for(t.N), dim3(kTanBlock), 0, c->stream, t, g->V0, g->dV, g->S, g->dcv,
                     g->dfn, g->dF, c->fc, y, done);
  int rc;
  if ((rc = launch_visc_tangent(c, g->V0, p, nullptr, y, done))) return rc;
  if (dt)
    hipLaunchKernelGGL(k_tan_time, grid_for(5LL * t.N), dim3(kTanBlock), 0, c->stream, 5 * (size_t)t.N, p, dt, y,
                       done);
  // the wall rows of the Navier-Stokes conditions (after the time term)
  if ((rc = launch_bcs_product(c, Q, p, y, done))) return rc;
  hipError_t e = hipGetLastError();
  return e == hipSuccess ? 0 : set_hip_error(e, "h2 apply launch");
}

}  // namespace aerofx

using namespace aerofx;

extern "C" int aerofx_residual_derivative(aerofx_ctx *c, const double *U, const double *dU, const double *dX,
                                          double *dR) {
  if (!c || !U || !dR) return set_error(AEROFX_ERR_ARG, "residual_derivative: null argument");
  if (c->fc.low_mach)
    return set_error(AEROFX_ERR_STATE, "residual_derivative: not available with low-Mach preconditioning");
  if (int rf = refuse_failsafe_nodes(c, "residual_derivative")) return rf;
  int rc = launch_tangent(c, U, dU, dX, dR);
  if (rc) return rc;
  hipError_t e = hipMemcpyAsync(c->h_err, c->d_err, sizeof(int), hipMemcpyDeviceToHost, c->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
  if (e != hipSuccess) return set_hip_error(e, "residual_derivative sync");
  return c->h_err[0];
}
