// viscous.hip -- laminar Navier-Stokes terms on the GPU: the Galerkin term
// of the residual (FemEquationTermNS, ElemTet::computeGalerkinTerm), its
// Jacobian blocks in H1 (ElemTet::computeJacobianGalerkinTerm), the viscous
// time-step term (ElemSet::computeTimeStep) and BcFcnNS's strong wall
// conditions (SubDomain::applyBCsToResidual / applyBCsToJacobian /
// applyBCsToSolutionVector).
//
// The element terms are gathered, not scattered: a node's residual sums its
// elements' contributions in element order, an H1 entry (n, m) sums those of
// the elements sharing edge (n, m), a diagonal block those of the node's
// elements -- the reference's per-node (per-block) order, with no atomics.
// Each element is evaluated once per vertex (residual) or per entry (H1):
// the gathers stay in L2 and the sums deterministic.
#include <algorithm>
#include <cstring>
#include <string>
#include <vector>

#include "context.h"
#include "device_dual.h"
#include "device_viscous.h"
#include "host_internal.h"

namespace aerofx {

namespace {

constexpr int kLocalEdge[6][2] = {{0, 1}, {0, 2}, {0, 3}, {1, 2}, {1, 3}, {2, 3}};

__device__ __forceinline__ void load_element(const int4 *__restrict__ tets, const double *__restrict__ X4,
                                             const double *__restrict__ V, int t, double x[4][3], double v[4][5]) {
  const int4 nd = tets[t];
  const int id[4] = {nd.x, nd.y, nd.z, nd.w};
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    const double4 p = *reinterpret_cast<const double4 *>(X4 + 4 * (size_t)id[k]);
    x[k][0] = p.x;
    x[k][1] = p.y;
    x[k][2] = p.z;
#pragma unroll
    for (int q = 0; q < 5; ++q) v[k][q] = V[5 * (size_t)id[k] + q];
  }
}

__device__ __forceinline__ bool wall_isothermal(int type) { return type == 1 || type == -1; }

// row k of a [4][W] register array without dynamic indexing (which would
// send the array to scratch)
template <int W>
__device__ __forceinline__ void pick_row(const double a[4][W], int k, double out[W]) {
#pragma unroll
  for (int q = 0; q < W; ++q) out[q] = a[0][q];
#pragma unroll
  for (int r = 1; r < 4; ++r)
#pragma unroll
    for (int q = 0; q < W; ++q) out[q] = k == r ? a[r][q] : out[q];
}

__device__ __forceinline__ double pick4(const double a[4], int k) {
  return k == 0 ? a[0] : (k == 1 ? a[1] : (k == 2 ? a[2] : a[3]));
}

// R[n] += (sum over the node's elements of vol (r . g_n)) (/ V_n if scale)
__global__ __launch_bounds__(kBlock) void k_visc_residual(int N, const int *__restrict__ ntrow,
                                                          const int *__restrict__ ntslot,
                                                          const int4 *__restrict__ tets, const double *__restrict__ X4,
                                                          const double *__restrict__ V,
                                                          const double *__restrict__ invcv, NsConsts ns, int scale,
                                                          double *__restrict__ R, const int *__restrict__ done) {
  if (done && *done) return;
  const int n = blockIdx.x * kBlock + threadIdx.x;
  if (n >= N) return;
  double acc[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
  for (int s = ntrow[n]; s < ntrow[n + 1]; ++s) {
    const int slot = ntslot[s], t = slot >> 2, loc = slot & 3;
    double x[4][3], v[4][5], g[4][3];
    load_element(tets, X4, V, t, x, v);
    const double vol = p1_gradients(x, g);
    NsElem e;
    ns_element(ns, g, v, e);
    double c[5], gl[3];
    pick_row<3>(g, loc, gl);
    ns_node_residual(e, gl, vol, c);
#pragma unroll
    for (int k = 0; k < 5; ++k) acc[k] += c[k];
  }
  const double ic = invcv[n];
#pragma unroll
  for (int k = 0; k < 5; ++k) R[5 * (size_t)n + k] += scale ? acc[k] * ic : acc[k];
}

// Element pass of the residual: each element once, its four vertex
// contributions vol (r . g_j) to C[pos[4 t + j]][5], pos the place of slot
// (t, j) in the node->element slot list, so that the node pass below reads
// each node's contributions as one contiguous run (in element order: the
// sums of k_visc_residual, term for term) and consecutive nodes read
// consecutive runs.
__global__ void k_slot_positions(int n, const int *__restrict__ ntslot, int *__restrict__ pos) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s < n) pos[ntslot[s]] = s;
}

__global__ __launch_bounds__(kBlock) void k_visc_elements(int T, const int4 *__restrict__ tets,
                                                          const double *__restrict__ X4, const double *__restrict__ V,
                                                          NsConsts ns, const int4 *__restrict__ pos,
                                                          double *__restrict__ C, const int *__restrict__ done) {
  if (done && *done) return;
  const int t = blockIdx.x * kBlock + threadIdx.x;
  if (t >= T) return;
  double x[4][3], v[4][5], g[4][3];
  load_element(tets, X4, V, t, x, v);
  const double vol = p1_gradients(x, g);
  NsElem e;
  ns_element(ns, g, v, e);
  const int4 p4 = pos[t];
  const int pj[4] = {p4.x, p4.y, p4.z, p4.w};
#pragma unroll
  for (int j = 0; j < 4; ++j) {
    double c[5];
    ns_node_residual(e, g[j], vol, c);
    double *out = C + 5 * (size_t)pj[j];
#pragma unroll
    for (int k = 0; k < 5; ++k) out[k] = c[k];
  }
}

// Transposed element pass: each element once, C[pos[4 t + k]] = vol (d r /
// d U_k)^T w with w[j][m] = sum_i g_i[j] lambda_i[m] / V_i (the residual of
// vertex i carries 1/V_i); k_visc_gather then adds the node sums to the
// transposed product.
__global__ __launch_bounds__(kBlock) void k_visc_adj_elements(int T, const int4 *__restrict__ tets,
                                                              const double *__restrict__ X4,
                                                              const double *__restrict__ V,
                                                              const double *__restrict__ lam,
                                                              const double *__restrict__ invcv, NsConsts ns,
                                                              const int4 *__restrict__ pos, double *__restrict__ C) {
  const int t = blockIdx.x * kBlock + threadIdx.x;
  if (t >= T) return;
  double x[4][3], v[4][5], g[4][3];
  load_element(tets, X4, V, t, x, v);
  const double vol = p1_gradients(x, g);
  NsElem e;
  ns_element(ns, g, v, e);
  const int4 nd = tets[t];
  const int id[4] = {nd.x, nd.y, nd.z, nd.w};
  double w[3][5];
#pragma unroll
  for (int j = 0; j < 3; ++j)
#pragma unroll
    for (int m = 0; m < 5; ++m) w[j][m] = 0.0;
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    const double ic = invcv[id[i]];
#pragma unroll
    for (int m = 0; m < 5; ++m) {
      const double l = lam[5 * (size_t)id[i] + m] * ic;
#pragma unroll
      for (int j = 0; j < 3; ++j) w[j][m] += g[i][j] * l;
    }
  }
  const int4 p4 = pos[t];
  const int pk[4] = {p4.x, p4.y, p4.z, p4.w};
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    double out[5];
    ns_node_adjoint(ns, e, g[k], v[k], e.T[k], w, out);
    double *o = C + 5 * (size_t)pk[k];
#pragma unroll
    for (int q = 0; q < 5; ++q) o[q] = vol * out[q];
  }
}

// The shape part of the element's vertex contributions: with the
// coordinates moving, only the P1 gradients g and the volume change (the
// states, mean temperature and transport coefficients stay), so
//   d c_j = dvol (r . g_j) + vol (dr . g_j + r . dg_j),
// dr from the velocity and temperature gradients along dg.
template <class T>
__device__ __forceinline__ T p1_gradients_t(const T x[4][3], T g[4][3]) {
  T J[3][3];
#pragma unroll
  for (int col = 0; col < 3; ++col)
#pragma unroll
    for (int r = 0; r < 3; ++r) J[r][col] = x[col + 1][r] - x[0][r];
  const T det = J[0][0] * (J[1][1] * J[2][2] - J[1][2] * J[2][1]) + J[1][0] * (J[0][2] * J[2][1] - J[0][1] * J[2][2]) +
                J[2][0] * (J[0][1] * J[1][2] - J[0][2] * J[1][1]);
  const T s = -1.0 / det;
  g[1][0] = (-J[1][1] * J[2][2] + J[1][2] * J[2][1]) * s;
  g[1][1] = (J[0][1] * J[2][2] - J[0][2] * J[2][1]) * s;
  g[1][2] = -(J[0][1] * J[1][2] - J[0][2] * J[1][1]) * s;
  g[2][0] = -(-J[1][0] * J[2][2] + J[1][2] * J[2][0]) * s;
  g[2][1] = -(J[0][0] * J[2][2] - J[0][2] * J[2][0]) * s;
  g[2][2] = (J[0][0] * J[1][2] - J[0][2] * J[1][0]) * s;
  g[3][0] = -(J[1][0] * J[2][1] - J[1][1] * J[2][0]) * s;
  g[3][1] = (J[0][0] * J[2][1] - J[0][1] * J[2][0]) * s;
  g[3][2] = -(J[0][0] * J[1][1] - J[0][1] * J[1][0]) * s;
#pragma unroll
  for (int d = 0; d < 3; ++d) g[0][d] = -(g[1][d] + g[2][d] + g[3][d]);
  return (1.0 / 6.0) * det;
}

// dc[j][m] for the coordinate change dx[4][3] of the element's vertices
__device__ __forceinline__ void ns_shape_contrib(const NsElem &e, const double v[4][5], const double x[4][3],
                                                 const double dx[4][3], double dc[4][5]) {
  Dual xd[4][3], gd[4][3];
#pragma unroll
  for (int k = 0; k < 4; ++k)
#pragma unroll
    for (int d = 0; d < 3; ++d) xd[k][d] = Dual(x[k][d], dx[k][d]);
  const Dual vd = p1_gradients_t<Dual>(xd, gd);
  double g[4][3], dg[4][3];
#pragma unroll
  for (int k = 0; k < 4; ++k)
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      g[k][d] = gd[k][d].v;
      dg[k][d] = gd[k][d].d;
    }
  const double vol = vd.v, dvol = vd.d;
  double dgu[3][3], dgT[3];
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j)
      dgu[i][j] = dg[0][j] * v[0][i + 1] + dg[1][j] * v[1][i + 1] + dg[2][j] * v[2][i + 1] + dg[3][j] * v[3][i + 1];
#pragma unroll
  for (int j = 0; j < 3; ++j) dgT[j] = dg[0][j] * e.T[0] + dg[1][j] * e.T[1] + dg[2][j] * e.T[2] + dg[3][j] * e.T[3];
  const double ddiv = dgu[0][0] + dgu[1][1] + dgu[2][2];
  double dt[3][3];
#pragma unroll
  for (int i = 0; i < 3; ++i) dt[i][i] = e.lambda * ddiv + 2.0 * e.mu * dgu[i][i];
  dt[0][1] = dt[1][0] = e.mu * (dgu[1][0] + dgu[0][1]);
  dt[0][2] = dt[2][0] = e.mu * (dgu[2][0] + dgu[0][2]);
  dt[1][2] = dt[2][1] = e.mu * (dgu[2][1] + dgu[1][2]);
  double r[3][5], dr[3][5];
#pragma unroll
  for (int j = 0; j < 3; ++j) {
    r[j][0] = dr[j][0] = 0.0;
#pragma unroll
    for (int i = 0; i < 3; ++i) {
      r[j][i + 1] = e.tij[i][j];
      dr[j][i + 1] = dt[i][j];
    }
    r[j][4] = e.ucg[0] * e.tij[0][j] + e.ucg[1] * e.tij[1][j] + e.ucg[2] * e.tij[2][j] + e.kappa * e.dTdxj[j];
    dr[j][4] = e.ucg[0] * dt[0][j] + e.ucg[1] * dt[1][j] + e.ucg[2] * dt[2][j] + e.kappa * dgT[j];
  }
#pragma unroll
  for (int j = 0; j < 4; ++j)
#pragma unroll
    for (int m = 0; m < 5; ++m)
      dc[j][m] = dvol * (r[0][m] * g[j][0] + r[1][m] * g[j][1] + r[2][m] * g[j][2]) +
                 vol * (dr[0][m] * g[j][0] + dr[1][m] * g[j][1] + dr[2][m] * g[j][2] + r[0][m] * dg[j][0] +
                        r[1][m] * dg[j][1] + r[2][m] * dg[j][2]);
}

// Adjoint of the shape part: TX[t][3 a + d] += sum_j lam_j / V_j . d c_j /
// d x_a[d] (added to the element's other geometry terms before the node
// gather of the adjoint)
__global__ __launch_bounds__(kBlock) void k_visc_adj_x(int T, const int4 *__restrict__ tets,
                                                       const double *__restrict__ X4, const double *__restrict__ V,
                                                       const double *__restrict__ lam,
                                                       const double *__restrict__ invcv, NsConsts ns,
                                                       double *__restrict__ TX) {
  const int t = blockIdx.x * kBlock + threadIdx.x;
  if (t >= T) return;
  double x[4][3], v[4][5], g[4][3];
  load_element(tets, X4, V, t, x, v);
  p1_gradients(x, g);
  NsElem e;
  ns_element(ns, g, v, e);
  const int4 nd = tets[t];
  const int id[4] = {nd.x, nd.y, nd.z, nd.w};
  double l[4][5];
#pragma unroll
  for (int j = 0; j < 4; ++j) {
    const double ic = invcv[id[j]];
#pragma unroll
    for (int m = 0; m < 5; ++m) l[j][m] = lam[5 * (size_t)id[j] + m] * ic;
  }
#pragma unroll 1
  for (int dir = 0; dir < 12; ++dir) {
    double dx[4][3];
#pragma unroll
    for (int k = 0; k < 4; ++k)
#pragma unroll
      for (int d = 0; d < 3; ++d) dx[k][d] = (3 * k + d == dir) ? 1.0 : 0.0;
    double dc[4][5];
    ns_shape_contrib(e, v, x, dx, dc);
    double s = 0.0;
#pragma unroll
    for (int j = 0; j < 4; ++j)
#pragma unroll
      for (int m = 0; m < 5; ++m) s += dc[j][m] * l[j][m];
    TX[12 * (size_t)t + dir] += s;
  }
}

// Forward element pass of the state derivative: each element once,
// dr = sum_k (d r / d U_k) dU_k and its vertex contributions vol (dr . g_j)
// to C[pos[4 t + j]]; k_visc_gather then adds the node sums (/ V).
__global__ __launch_bounds__(kBlock) void k_visc_tan_elements(int T, const int4 *__restrict__ tets,
                                                              const double *__restrict__ X4,
                                                              const double *__restrict__ V,
                                                              const double *__restrict__ dU,
                                                              const double *__restrict__ dX, NsConsts ns,
                                                              const int4 *__restrict__ pos, double *__restrict__ C,
                                                              const int *__restrict__ done) {
  if (done && *done) return;
  const int t = blockIdx.x * kBlock + threadIdx.x;
  if (t >= T) return;
  double x[4][3], v[4][5], g[4][3];
  load_element(tets, X4, V, t, x, v);
  const double vol = p1_gradients(x, g);
  NsElem e;
  ns_element(ns, g, v, e);
  const int4 nd = tets[t];
  const int id[4] = {nd.x, nd.y, nd.z, nd.w};
  double dc[4][5];
  if (dX) {
    double dx[4][3];
#pragma unroll
    for (int k = 0; k < 4; ++k)
#pragma unroll
      for (int d = 0; d < 3; ++d) dx[k][d] = dX[3 * (size_t)id[k] + d];
    ns_shape_contrib(e, v, x, dx, dc);
  } else {
#pragma unroll
    for (int j = 0; j < 4; ++j)
#pragma unroll
      for (int m = 0; m < 5; ++m) dc[j][m] = 0.0;
  }
  double dr[3][5];
#pragma unroll
  for (int j = 0; j < 3; ++j)
#pragma unroll
    for (int m = 0; m < 5; ++m) dr[j][m] = 0.0;
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    if (!dU) break;
#pragma unroll
    for (int cc = 0; cc < 5; ++cc) {
      const double du = dU[5 * (size_t)id[k] + cc];
      double col[3][5];
      ns_node_column(ns, e, g[k], v[k], e.T[k], cc, col);
#pragma unroll
      for (int j = 0; j < 3; ++j)
#pragma unroll
        for (int m = 0; m < 5; ++m) dr[j][m] += col[j][m] * du;
    }
  }
  const int4 p4 = pos[t];
  const int pj[4] = {p4.x, p4.y, p4.z, p4.w};
#pragma unroll
  for (int j = 0; j < 4; ++j) {
    double *o = C + 5 * (size_t)pj[j];
#pragma unroll
    for (int m = 0; m < 5; ++m) o[m] = vol * (dr[0][m] * g[j][0] + dr[1][m] * g[j][1] + dr[2][m] * g[j][2]) + dc[j][m];
  }
}

// Viscous wall force (PostFcnNS::computeViscousForce, PostFcn.C:1578-1612;
// FaceTria::computeForce, FaceTria.C:28-60): on every wall face f with its
// element e, F_v = -tau_e n_f (n_f the full outward face normal
// 0.5 (x2 - x0) x (x1 - x0)), J += F_v . d.  With FU / FX given, the
// derivatives of J_f with respect to the states [4][5] and coordinates
// [4][3] of the element's vertices go to the face's slots 4 f + k.
__global__ __launch_bounds__(kBlock) void k_visc_force(int F, const int *__restrict__ faces,
                                                       const int *__restrict__ fcode,
                                                       const double *__restrict__ fnorm,
                                                       const int *__restrict__ face_elem,
                                                       const int4 *__restrict__ tets, const double *__restrict__ X4,
                                                       const double *__restrict__ V, NsConsts ns, double d0, double d1,
                                                       double d2, double *__restrict__ Jf, double *__restrict__ FU,
                                                       double *__restrict__ FX) {
  const int f = blockIdx.x * kBlock + threadIdx.x;
  if (f >= F) return;
  const int t = face_elem[f];
  if (t < 0 || !is_wall_code(fcode[f])) return;
  double x[4][3], v[4][5], g[4][3];
  load_element(tets, X4, V, t, x, v);
  p1_gradients(x, g);
  NsElem e;
  ns_element(ns, g, v, e);
  const double n[3] = {fnorm[3 * (size_t)f], fnorm[3 * (size_t)f + 1], fnorm[3 * (size_t)f + 2]};
  const double dd[3] = {d0, d1, d2};
  // J_f = -sum_ij d_i tau_ij n_j
  double jf = 0.0;
#pragma unroll
  for (int i = 0; i < 3; ++i) jf += -(e.tij[i][0] * n[0] + e.tij[i][1] * n[1] + e.tij[i][2] * n[2]) * dd[i];
  Jf[f] += jf;
  if (!FU) return;
#pragma unroll 1
  for (int k = 0; k < 4; ++k) {
    double vk[5], gk[3];
    pick_row<5>(v, k, vk);
    pick_row<3>(g, k, gk);
#pragma unroll
    for (int cc = 0; cc < 5; ++cc) {
      double col[3][5];
      ns_node_column(ns, e, gk, vk, pick4(e.T, k), cc, col);
      // col[j][i + 1] = d tau_ij / d U_k[cc]
      double s = 0.0;
#pragma unroll
      for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j) s += -col[j][i + 1] * n[j] * dd[i];
      FU[5 * (4 * (size_t)f + k) + cc] = s;
    }
  }
  // coordinates: through the P1 gradients (dual numbers on the element map)
  // and through the face normal (the face's vertices)
  const int4 nd = tets[t];
  const int id[4] = {nd.x, nd.y, nd.z, nd.w};
  const int fv[3] = {faces[3 * (size_t)f], faces[3 * (size_t)f + 1], faces[3 * (size_t)f + 2]};
#pragma unroll 1
  for (int dir = 0; dir < 12; ++dir) {
    const int k = dir / 3, dc = dir - 3 * k;
    Dual xd[4][3], gd[4][3];
#pragma unroll
    for (int a = 0; a < 4; ++a)
#pragma unroll
      for (int b = 0; b < 3; ++b) xd[a][b] = Dual(x[a][b], (a == k && b == dc) ? 1.0 : 0.0);
    p1_gradients_t<Dual>(xd, gd);
    double dgu[3][3];
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
      for (int j = 0; j < 3; ++j)
        dgu[i][j] = gd[0][j].d * v[0][i + 1] + gd[1][j].d * v[1][i + 1] + gd[2][j].d * v[2][i + 1] +
                    gd[3][j].d * v[3][i + 1];
    const double ddiv = dgu[0][0] + dgu[1][1] + dgu[2][2];
    double dt[3][3];
#pragma unroll
    for (int i = 0; i < 3; ++i) dt[i][i] = e.lambda * ddiv + 2.0 * e.mu * dgu[i][i];
    dt[0][1] = dt[1][0] = e.mu * (dgu[1][0] + dgu[0][1]);
    dt[0][2] = dt[2][0] = e.mu * (dgu[2][0] + dgu[0][2]);
    dt[1][2] = dt[2][1] = e.mu * (dgu[2][1] + dgu[1][2]);
    // the face normal's derivative when the vertex is on the face
    double dn[3] = {0.0, 0.0, 0.0};
#pragma unroll
    for (int q = 0; q < 3; ++q) {
      if (fv[q] != id[k]) continue;
      Dual p0[3], p1[3], p2[3];
#pragma unroll
      for (int b = 0; b < 3; ++b) {
        const double4 a0 = *reinterpret_cast<const double4 *>(X4 + 4 * (size_t)fv[0]);
        const double4 a1 = *reinterpret_cast<const double4 *>(X4 + 4 * (size_t)fv[1]);
        const double4 a2 = *reinterpret_cast<const double4 *>(X4 + 4 * (size_t)fv[2]);
        const double c0 = b == 0 ? a0.x : (b == 1 ? a0.y : a0.z);
        const double c1 = b == 0 ? a1.x : (b == 1 ? a1.y : a1.z);
        const double c2 = b == 0 ? a2.x : (b == 1 ? a2.y : a2.z);
        p0[b] = Dual(c0, (q == 0 && b == dc) ? 1.0 : 0.0);
        p1[b] = Dual(c1, (q == 1 && b == dc) ? 1.0 : 0.0);
        p2[b] = Dual(c2, (q == 2 && b == dc) ? 1.0 : 0.0);
      }
      Dual u[3], w[3];
#pragma unroll
      for (int b = 0; b < 3; ++b) {
        u[b] = p2[b] - p0[b];
        w[b] = p1[b] - p0[b];
      }
      const Dual c0 = u[1] * w[2] - u[2] * w[1], c1 = u[2] * w[0] - u[0] * w[2], c2 = u[0] * w[1] - u[1] * w[0];
      dn[0] = 0.5 * c0.d;
      dn[1] = 0.5 * c1.d;
      dn[2] = 0.5 * c2.d;
    }
    double s = 0.0;
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
      for (int j = 0; j < 3; ++j) s += -(dt[i][j] * n[j] + e.tij[i][j] * dn[j]) * dd[i];
    FX[3 * (4 * (size_t)f + k) + dc] = s;
  }
}

// dJdU[n] += sum over the node's wall-face element slots of FU, and dJdX
// likewise, in face order
__global__ __launch_bounds__(kBlock) void k_visc_force_gather(int N, const int *__restrict__ row,
                                                              const int *__restrict__ slot,
                                                              const double *__restrict__ FU,
                                                              const double *__restrict__ FX, double *__restrict__ dJdU,
                                                              double *__restrict__ dJdX) {
  const int n = blockIdx.x * kBlock + threadIdx.x;
  if (n >= N) return;
  double u[5] = {0.0, 0.0, 0.0, 0.0, 0.0}, xs[3] = {0.0, 0.0, 0.0};
  for (int s = row[n]; s < row[n + 1]; ++s) {
    const size_t q = slot[s];
#pragma unroll
    for (int k = 0; k < 5; ++k) u[k] += FU[5 * q + k];
#pragma unroll
    for (int k = 0; k < 3; ++k) xs[k] += FX[3 * q + k];
  }
  if (dJdU)
#pragma unroll
    for (int k = 0; k < 5; ++k) dJdU[5 * (size_t)n + k] += u[k];
  if (dJdX)
#pragma unroll
    for (int k = 0; k < 3; ++k) dJdX[3 * (size_t)n + k] += xs[k];
}

// The wall rows of a product y = A p (H2 products): y_r = p_r on the
// momentum rows (the isothermal energy row too)
__global__ void k_bcs_product(int nwall, const int2 *__restrict__ wall, double wall_T, const double *__restrict__ U,
                              const double *__restrict__ p, double *__restrict__ y, const int *__restrict__ done) {
  if (done && *done) return;
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nwall) return;
  const int n = wall[i].x;
  const double *q = p + 5 * (size_t)n;
  double *o = y + 5 * (size_t)n;
  o[1] = q[1];
  o[2] = q[2];
  o[3] = q[3];
  if (wall_isothermal(wall[i].y)) {
    const double *u = U + 5 * (size_t)n;
    o[4] = (u[4] - 2.0 * wall_T * u[0]) * q[0] - u[1] * q[1] - u[2] * q[2] - u[3] * q[3] + u[0] * q[4];
  }
}

// The wall rows in a transposed product: P zeroes the rows that the wall
// conditions replace (momentum; energy on isothermal walls) ...
__global__ void k_bcs_adj_mask(int nwall, const int2 *__restrict__ wall, double *__restrict__ y) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nwall) return;
  const int n = wall[i].x;
  const int r1 = wall_isothermal(wall[i].y) ? 5 : 4;
  for (int r = 1; r < r1; ++r) y[5 * (size_t)n + r] = 0.0;
}

// ... and y += B^T psi adds the transposed rows of the conditions
// (d F_r / d U = e_r on the momentum rows; the isothermal energy row
// (U4 - 2 Tw U0, -U1, -U2, -U3, U0))
__global__ void k_bcs_adj_rows(int nwall, const int2 *__restrict__ wall, double wall_T, const double *__restrict__ U,
                               const double *__restrict__ psi, double *__restrict__ y) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nwall) return;
  const int n = wall[i].x;
  const double *p = psi + 5 * (size_t)n;
  double *o = y + 5 * (size_t)n;
  o[1] += p[1];
  o[2] += p[2];
  o[3] += p[3];
  if (wall_isothermal(wall[i].y)) {
    const double *u = U + 5 * (size_t)n;
    o[0] += (u[4] - 2.0 * wall_T * u[0]) * p[4];
    o[1] += -u[1] * p[4];
    o[2] += -u[2] * p[4];
    o[3] += -u[3] * p[4];
    o[4] += u[0] * p[4];
  }
}

// Node pass: R[n] += sum over the node's slots (element order) (/ V_n)
__global__ __launch_bounds__(kBlock) void k_visc_gather(int N, const int *__restrict__ ntrow,
                                                        const double *__restrict__ C,
                                                        const double *__restrict__ invcv, int scale,
                                                        double *__restrict__ R, const int *__restrict__ done) {
  if (done && *done) return;
  const int n = blockIdx.x * kBlock + threadIdx.x;
  if (n >= N) return;
  double acc[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
  for (int s = ntrow[n]; s < ntrow[n + 1]; ++s) {
    const double *c = C + 5 * (size_t)s;
#pragma unroll
    for (int k = 0; k < 5; ++k) acc[k] += c[k];
  }
  const double ic = invcv[n];
#pragma unroll
  for (int k = 0; k < 5; ++k) R[5 * (size_t)n + k] += scale ? acc[k] * ic : acc[k];
}

// H1 entries: index q < 2E is adjacency entry q = (n, m), summed over the
// elements of edge (n, m) and added (times 1/V_n per contribution) to its
// off-diagonal block; q = 2E + n is the diagonal block of n, summed over the
// node's elements and added times 1/V_n.
template <class T>
__global__ __launch_bounds__(kBlock) void k_visc_jacobian(DevTables tb, const int *__restrict__ ntrow,
                                                          const int *__restrict__ ntslot,
                                                          const int *__restrict__ etrow,
                                                          const int *__restrict__ etslot,
                                                          const int4 *__restrict__ tets, const double *__restrict__ V,
                                                          NsConsts ns, T *__restrict__ D, T *__restrict__ B) {
  const long long q = (long long)blockIdx.x * kBlock + threadIdx.x;
  const int E2 = 2 * tb.E;
  if (q >= (long long)E2 + tb.N) return;
  double acc[25];
#pragma unroll
  for (int k = 0; k < 25; ++k) acc[k] = 0.0;
  if (q < E2) {
    const int e = (int)q;
    // row of entry e
    int lo = 0, hi = tb.N;
    while (hi - lo > 1) {
      const int mid = (lo + hi) >> 1;
      if (tb.rowptr[mid] <= e) lo = mid; else hi = mid;
    }
    const int n = lo, l = tb.eidx[e];
    const double cn = tb.invcv[n];
    for (int s = etrow[l]; s < etrow[l + 1]; ++s) {
      const int slot = etslot[s], t = slot / 6, le = slot - 6 * t;
      const int4 nd = tets[t];
      const int id[4] = {nd.x, nd.y, nd.z, nd.w};
      const int a = kLocalEdge[le][0], b = kLocalEdge[le][1];
      const int ln = id[a] == n ? a : b, lm = id[a] == n ? b : a;
      double x[4][3], v[4][5], g[4][3];
      load_element(tets, tb.X4, V, t, x, v);
      const double vol = p1_gradients(x, g);
      NsElem el;
      ns_element(ns, g, v, el);
      double blk[25], gm[3], gn[3], vm[5];
      pick_row<3>(g, lm, gm);
      pick_row<3>(g, ln, gn);
      pick_row<5>(v, lm, vm);
      ns_node_jacobian(ns, el, gm, vm, pick4(el.T, lm), gn, blk);
#pragma unroll
      for (int k = 0; k < 25; ++k) acc[k] += cn * blk[k] * vol;
    }
    T *Be = B + blk64_base(e);
#pragma unroll
    for (int k = 0; k < 25; ++k) Be[k * kBlkStride] = (T)((double)Be[k * kBlkStride] + acc[k]);
  } else {
    const int n = (int)(q - E2);
    for (int s = ntrow[n]; s < ntrow[n + 1]; ++s) {
      const int slot = ntslot[s], t = slot >> 2, loc = slot & 3;
      double x[4][3], v[4][5], g[4][3];
      load_element(tets, tb.X4, V, t, x, v);
      const double vol = p1_gradients(x, g);
      NsElem el;
      ns_element(ns, g, v, el);
      double blk[25], gl[3], vl[5];
      pick_row<3>(g, loc, gl);
      pick_row<5>(v, loc, vl);
      ns_node_jacobian(ns, el, gl, vl, pick4(el.T, loc), gl, blk);
#pragma unroll
      for (int k = 0; k < 25; ++k) acc[k] += blk[k] * vol;
    }
    const double ic = tb.invcv[n];
    T *Dn = D + blk64_base(n);
#pragma unroll
    for (int k = 0; k < 25; ++k) Dn[k * kBlkStride] = (T)((double)Dn[k * kBlkStride] + acc[k] * ic);
  }
}

// idtv[n] = sum over the node's elements of 9 (mu/Re)/rho_mid |T| |g_n|^2,
// the element's mean state (Vmid summed in vertex order, times 1/4)
__global__ __launch_bounds__(kBlock) void k_visc_idtv(int N, const int *__restrict__ ntrow,
                                                      const int *__restrict__ ntslot, const int4 *__restrict__ tets,
                                                      const double *__restrict__ X4, const double *__restrict__ V,
                                                      NsConsts ns, double *__restrict__ idtv) {
  const int n = blockIdx.x * kBlock + threadIdx.x;
  if (n >= N) return;
  double acc = 0.0;
  for (int s = ntrow[n]; s < ntrow[n + 1]; ++s) {
    const int slot = ntslot[s], t = slot >> 2, loc = slot & 3;
    double x[4][3], v[4][5], g[4][3];
    load_element(tets, X4, V, t, x, v);
    double Vmid[5];
#pragma unroll
    for (int k = 0; k < 5; ++k) Vmid[k] = (((0.0 + v[0][k]) + v[1][k]) + v[2][k]) + v[3][k];
#pragma unroll
    for (int k = 0; k < 5; ++k) Vmid[k] *= 1.0 / 4.0;
    const double mul = ns.ooRe * ns_mu(ns, ns_temperature(ns, Vmid));
    double a[3], b[3], w[3];
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      a[d] = x[1][d] - x[0][d];
      b[d] = x[2][d] - x[0][d];
      w[d] = x[3][d] - x[0][d];
    }
    const double cr[3] = {a[1] * b[2] - a[2] * b[1], a[2] * b[0] - a[0] * b[2], a[0] * b[1] - a[1] * b[0]};
    const double volume = (1.0 / 6.0) * (w[0] * cr[0] + w[1] * cr[1] + w[2] * cr[2]);
    const double localDt = 9.0 * (mul / Vmid[0]) * volume;
    p1_gradients(x, g);
    double gl[3];
    pick_row<3>(g, loc, gl);
    acc += localDt * (gl[0] * gl[0] + gl[1] * gl[1] + gl[2] * gl[2]);
  }
  idtv[n] = acc;
}


// BcFcnNS::applyToResidualTerm (BcFcnCore.C:230-245), fixed walls
__global__ void k_bcs_residual(int nwall, const int2 *__restrict__ wall, double wall_T, const double *__restrict__ U,
                               double *__restrict__ F, const int *__restrict__ done) {
  if (done && *done) return;
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nwall) return;
  const int n = wall[i].x;
  const double Vw[5] = {0.0, 0.0, 0.0, 0.0, wall_T};
  const double *u = U + 5 * (size_t)n;
  double *f = F + 5 * (size_t)n;
  f[1] = -Vw[1] * u[0] + u[1];
  f[2] = -Vw[2] * u[0] + u[2];
  f[3] = -Vw[3] * u[0] + u[3];
  if (wall_isothermal(wall[i].y)) f[4] = (u[4] - u[0] * Vw[4]) * u[0] - 0.5 * (u[1] * u[1] + u[2] * u[2] + u[3] * u[3]);
}

// BcFcnNS::applyToSolutionVectorTerm (BcFcnCore.C:212-226)
__global__ void k_bcs_solution(int nwall, const int2 *__restrict__ wall, double wall_T, double *__restrict__ U) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nwall) return;
  const int n = wall[i].x;
  const double Vw[5] = {0.0, 0.0, 0.0, 0.0, wall_T};
  double *u = U + 5 * (size_t)n;
  u[1] = u[0] * Vw[1];
  u[2] = u[0] * Vw[2];
  u[3] = u[0] * Vw[3];
  if (wall_isothermal(wall[i].y)) u[4] = u[0] * (Vw[4] + 0.5 * (Vw[1] * Vw[1] + Vw[2] * Vw[2] + Vw[3] * Vw[3]));
}

// BcFcnNS::applyToOffDiagonalTerm / applyToDiagonalTerm (BcFcn.C:8-122) on
// every block of the wall rows
template <class T>
__global__ void k_bcs_jacobian(int nwall, const int2 *__restrict__ wall, const int *__restrict__ rowptr, double wall_T,
                               const double *__restrict__ U, T *__restrict__ D, T *__restrict__ B) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nwall) return;
  const int n = wall[i].x;
  const bool iso = wall_isothermal(wall[i].y);
  const int r1 = iso ? 5 : 4;
  for (int e = rowptr[n]; e < rowptr[n + 1]; ++e) {
    T *Be = B + blk64_base(e);
    for (int r = 1; r < r1; ++r)
      for (int k = 0; k < 5; ++k) Be[(5 * r + k) * kBlkStride] = (T)0.0;
  }
  const double Vw[5] = {0.0, 0.0, 0.0, 0.0, wall_T};
  const double *u = U + 5 * (size_t)n;
  T *Dn = D + blk64_base(n);
  for (int r = 1; r < r1; ++r)
    for (int k = 0; k < 5; ++k) Dn[(5 * r + k) * kBlkStride] = (T)0.0;
  Dn[5 * kBlkStride] = (T)(-Vw[1]);
  Dn[6 * kBlkStride] = (T)1.0;
  Dn[10 * kBlkStride] = (T)(-Vw[2]);
  Dn[12 * kBlkStride] = (T)1.0;
  Dn[15 * kBlkStride] = (T)(-Vw[3]);
  Dn[18 * kBlkStride] = (T)1.0;
  if (iso) {
    Dn[20 * kBlkStride] = (T)(u[4] - 2.0 * Vw[4] * u[0]);
    Dn[21 * kBlkStride] = (T)(-u[1]);
    Dn[22 * kBlkStride] = (T)(-u[2]);
    Dn[23 * kBlkStride] = (T)(-u[3]);
    Dn[24 * kBlkStride] = (T)u[0];
  }
}

int check(const char *what) {
  hipError_t e = hipGetLastError();
  return e == hipSuccess ? 0 : set_hip_error(e, what);
}

// Domain::setNodeType priorities (DomainCore.C:618-640)
int bc_priority(int code) {
  static const int pr[22] = {15, 16, 13, 14, 17, 11, 12, 20, 18, 19, -1, 9, 8, 10, 2, 1, 7, 4, 3, 6, 5, 0};
  return (code < -10 || code > 11) ? -1 : pr[code + 10];
}

}  // namespace

int launch_visc_elements(aerofx_ctx *c, const double *V, const int *done) {
  const ElemTables *el = c->elem;
  ProfScope ps(c, AEROFX_TIMER_FE_TERM);
  hipLaunchKernelGGL(k_visc_elements, dim3((c->T + kBlock - 1) / kBlock), dim3(kBlock), 0, c->stream, c->T,
                     (const int4 *)el->tets, c->tab.X4, V, c->ns, (const int4 *)c->d_visc_pos, c->d_visc_c, done);
  return check("viscous element launch");
}

int launch_visc_residual(aerofx_ctx *c, const double *V, double *R, bool scale, const int *done) {
  if (!c->viscous) return 0;
  const ElemTables *el = c->elem;
  ProfScope ps(c, AEROFX_TIMER_FE_TERM);
  if (c->d_visc_c) {
    hipLaunchKernelGGL(k_visc_elements, dim3((c->T + kBlock - 1) / kBlock), dim3(kBlock), 0, c->stream, c->T,
                       (const int4 *)el->tets, c->tab.X4, V, c->ns, (const int4 *)c->d_visc_pos, c->d_visc_c, done);
    hipLaunchKernelGGL(k_visc_gather, dim3((c->N + kBlock - 1) / kBlock), dim3(kBlock), 0, c->stream, c->N,
                       el->ntrow, c->d_visc_c, c->tab.invcv, scale ? 1 : 0, R, done);
  } else {
    hipLaunchKernelGGL(k_visc_residual, dim3((c->N + kBlock - 1) / kBlock), dim3(kBlock), 0, c->stream, c->N,
                       el->ntrow, el->ntslot, (const int4 *)el->tets, c->tab.X4, V, c->tab.invcv, c->ns,
                       scale ? 1 : 0, R, done);
  }
  return check("viscous residual launch");
}

int launch_visc_jacobian(aerofx_ctx *c, const double *U, int precision) {
  if (!c->viscous) return 0;
  const ElemTables *el = c->elem;
  const long long nq = 2LL * c->E + c->N;
  const dim3 grid((unsigned)((nq + kBlock - 1) / kBlock));
  {
    ProfScope ps(c, AEROFX_TIMER_FE_JAC);
    if (precision)
      hipLaunchKernelGGL(k_visc_jacobian<float>, grid, dim3(kBlock), 0, c->stream, c->tab, el->ntrow, el->ntslot,
                         el->etrow, el->etslot, (const int4 *)el->tets, c->d_V, c->ns, (float *)c->d_jacD,
                         (float *)c->d_jacB);
    else
      hipLaunchKernelGGL(k_visc_jacobian<double>, grid, dim3(kBlock), 0, c->stream, c->tab, el->ntrow, el->ntslot,
                         el->etrow, el->etslot, (const int4 *)el->tets, c->d_V, c->ns, (double *)c->d_jacD,
                         (double *)c->d_jacB);
  }
  int rc = check("viscous jacobian launch");
  if (rc || !c->nwall) return rc;
  if (!U) return set_error(AEROFX_ERR_ARG, "jacobian: the wall conditions need the state");
  const dim3 g((c->nwall + 63) / 64), b(64);
  if (precision)
    hipLaunchKernelGGL(k_bcs_jacobian<float>, g, b, 0, c->stream, c->nwall, c->d_wall, c->tab.rowptr, c->ns.wall_T,
                       U, (float *)c->d_jacD, (float *)c->d_jacB);
  else
    hipLaunchKernelGGL(k_bcs_jacobian<double>, g, b, 0, c->stream, c->nwall, c->d_wall, c->tab.rowptr, c->ns.wall_T,
                       U, (double *)c->d_jacD, (double *)c->d_jacB);
  return check("wall jacobian launch");
}

int launch_visc_adjoint(aerofx_ctx *c, const double *lam, double *Ub) {
  if (!c->viscous) return 0;
  if (!c->d_visc_c) return set_error(AEROFX_ERR_STATE, "adjoint: the viscous contribution buffer is missing");
  const ElemTables *el = c->elem;
  hipLaunchKernelGGL(k_visc_adj_elements, dim3((c->T + kBlock - 1) / kBlock), dim3(kBlock), 0, c->stream, c->T,
                     (const int4 *)el->tets, c->tab.X4, c->d_V, lam, c->tab.invcv, c->ns,
                     (const int4 *)c->d_visc_pos, c->d_visc_c);
  hipLaunchKernelGGL(k_visc_gather, dim3((c->N + kBlock - 1) / kBlock), dim3(kBlock), 0, c->stream, c->N,
                     el->ntrow, c->d_visc_c, c->tab.invcv, 0, Ub, nullptr);
  return check("viscous adjoint launch");
}

int launch_visc_tangent(aerofx_ctx *c, const double *V, const double *dU, const double *dX, double *dR,
                        const int *done) {
  if (!c->viscous || (!dU && !dX)) return 0;
  if (!c->d_visc_c) return set_error(AEROFX_ERR_STATE, "tangent: the viscous contribution buffer is missing");
  const ElemTables *el = c->elem;
  hipLaunchKernelGGL(k_visc_tan_elements, dim3((c->T + kBlock - 1) / kBlock), dim3(kBlock), 0, c->stream, c->T,
                     (const int4 *)el->tets, c->tab.X4, V, dU, dX, c->ns, (const int4 *)c->d_visc_pos, c->d_visc_c,
                     done);
  hipLaunchKernelGGL(k_visc_gather, dim3((c->N + kBlock - 1) / kBlock), dim3(kBlock), 0, c->stream, c->N,
                     el->ntrow, c->d_visc_c, c->tab.invcv, 1, dR, done);
  return check("viscous tangent launch");
}

int launch_visc_adjoint_x(aerofx_ctx *c, const double *lam, double *TX) {
  if (!c->viscous) return 0;
  const ElemTables *el = c->elem;
  hipLaunchKernelGGL(k_visc_adj_x, dim3((c->T + kBlock - 1) / kBlock), dim3(kBlock), 0, c->stream, c->T,
                     (const int4 *)el->tets, c->tab.X4, c->d_V, lam, c->tab.invcv, c->ns, TX);
  return check("viscous shape adjoint launch");
}

int launch_visc_force(aerofx_ctx *c, const double dir[3], double *Jf, bool derivs) {
  if (!c->viscous || !c->F) return 0;
  const ElemTables *el = c->elem;
  hipLaunchKernelGGL(k_visc_force, dim3((c->F + kBlock - 1) / kBlock), dim3(kBlock), 0, c->stream, c->F, el->faces,
                     c->tab.fcode, c->tab.fnorm, c->d_face_elem, (const int4 *)el->tets, c->tab.X4, c->d_V, c->ns,
                     dir[0], dir[1], dir[2], Jf, derivs ? c->d_force_u : nullptr, derivs ? c->d_force_x : nullptr);
  return check("viscous force launch");
}

int launch_visc_force_gather(aerofx_ctx *c, double *dJdU, double *dJdX) {
  if (!c->viscous || (!dJdU && !dJdX)) return 0;
  hipLaunchKernelGGL(k_visc_force_gather, dim3((c->N + kBlock - 1) / kBlock), dim3(kBlock), 0, c->stream, c->N,
                     c->d_nfe_row, c->d_nfe_slot, c->d_force_u, c->d_force_x, dJdU, dJdX);
  return check("viscous force gather launch");
}

int launch_bcs_product(aerofx_ctx *c, const double *U, const double *p, double *y, const int *done) {
  if (!c->viscous || !c->nwall) return 0;
  hipLaunchKernelGGL(k_bcs_product, dim3((c->nwall + 63) / 64), dim3(64), 0, c->stream, c->nwall, c->d_wall,
                     c->ns.wall_T, U, p, y, done);
  return check("wall product launch");
}

int launch_bcs_adjoint(aerofx_ctx *c, const double *U, const double *psi, double *Ppsi, double *y, bool pre) {
  if (!c->viscous || !c->nwall) return 0;
  const dim3 g((c->nwall + 63) / 64), b(64);
  if (pre) {
    hipError_t e = hipMemcpyAsync(Ppsi, psi, sizeof(double) * 5 * (size_t)c->N, hipMemcpyDeviceToDevice, c->stream);
    if (e != hipSuccess) return set_hip_error(e, "adjoint wall rows copy");
    hipLaunchKernelGGL(k_bcs_adj_mask, g, b, 0, c->stream, c->nwall, c->d_wall, Ppsi);
  } else {
    hipLaunchKernelGGL(k_bcs_adj_rows, g, b, 0, c->stream, c->nwall, c->d_wall, c->ns.wall_T, U, psi, y);
  }
  return check("adjoint wall rows launch");
}

int launch_bcs_residual(aerofx_ctx *c, const double *U, double *F, const int *done) {
  if (!c->viscous || !c->nwall) return 0;
  hipLaunchKernelGGL(k_bcs_residual, dim3((c->nwall + 63) / 64), dim3(64), 0, c->stream, c->nwall, c->d_wall,
                     c->ns.wall_T, U, F, done);
  return check("wall residual launch");
}

int launch_visc_idtv(aerofx_ctx *c, double **idtv) {
  *idtv = nullptr;
  if (!c->viscous || c->ns.viscous_cst == 0.0) return 0;
  if (!c->d_idtv) {
    void *p = nullptr;
    int rc = ctx_alloc(c, sizeof(double) * (size_t)c->N, &p);
    if (rc) return rc;
    c->d_idtv = (double *)p;
  }
  const ElemTables *el = c->elem;
  hipLaunchKernelGGL(k_visc_idtv, dim3((c->N + kBlock - 1) / kBlock), dim3(kBlock), 0, c->stream, c->N, el->ntrow,
                     el->ntslot, (const int4 *)el->tets, c->tab.X4, c->d_V, c->ns, c->d_idtv);
  *idtv = c->d_idtv;
  return check("viscous time-step launch");
}

int refuse_viscous(aerofx_ctx *c, const char *who) {
  if (!c->viscous) return 0;
  return set_error(AEROFX_ERR_STATE, std::string(who) + ": not available with the Navier-Stokes terms");
}

}  // namespace aerofx

using namespace aerofx;

extern "C" {

int aerofx_ctx_set_viscous(aerofx_ctx *c, const aerofx_viscous *v) {
  if (!c) return set_error(AEROFX_ERR_ARG, "set_viscous: null context");
  if (!v) {
    c->viscous = 0;
    c->nwall = 0;
    return AEROFX_OK;
  }
  if (c->fc.low_mach) return set_error(AEROFX_ERR_STATE, "set_viscous: not with the low-Mach preconditioner");
  if (!(v->reynolds > 0.0)) return set_error(AEROFX_ERR_ARG, "set_viscous: Reynolds number must be > 0");
  if (v->viscosity_model < 0 || v->viscosity_model > 2 || v->conductivity_model < 0 || v->conductivity_model > 1)
    return set_error(AEROFX_ERR_ARG, "set_viscous: unknown viscosity or conductivity model");
  if (v->conductivity_model == AEROFX_COND_CONSTANT_PRANDTL && !(v->prandtl > 0.0))
    return set_error(AEROFX_ERR_ARG, "set_viscous: Prandtl number must be > 0");
  if (v->viscosity_model == AEROFX_VISC_SUTHERLAND && !(v->sutherland_ts > 0.0))
    return set_error(AEROFX_ERR_ARG, "set_viscous: Sutherland temperature ratio must be > 0");
  int rc = ensure_elem_tables(c);
  if (rc) return rc;
  // the element contributions of the residual [4T][5] (allocated here, at
  // set-up, so that no residual call allocates); without the memory the
  // per-node kernel evaluates each element once per vertex instead
  if (!c->d_visc_c) {
    void *p = nullptr, *q = nullptr;
    if (hipMalloc(&p, sizeof(double) * 20 * (size_t)c->T) == hipSuccess &&
        hipMalloc(&q, sizeof(int) * 4 * (size_t)c->T) == hipSuccess) {
      c->allocs.push_back(p);
      c->allocs.push_back(q);
      c->d_visc_c = (double *)p;
      c->d_visc_pos = (int *)q;
      const int ns4 = 4 * c->T;
      hipLaunchKernelGGL(k_slot_positions, dim3((ns4 + 255) / 256), dim3(256), 0, c->stream, ns4, c->elem->ntslot,
                         c->d_visc_pos);
      if ((rc = check("slot positions launch"))) return rc;
    } else {
      (void)hipGetLastError();
      if (p) (void)hipFree(p);
    }
  }
  NsConsts ns{};
  ns.visc = v->viscosity_model;
  ns.cond = v->conductivity_model;
  ns.ooRe = 1.0 / v->reynolds;
  ns.gam = c->p.gamma;
  ns.invgam1 = 1.0 / (c->p.gamma - 1.0);
  ns.pstiff = c->p.pstiff;
  ns.alpha = c->p.gamma * (c->p.gamma - 1.0) * c->p.mach * c->p.mach;
  ns.dalpha = 2.0 * c->p.gamma * (c->p.gamma - 1.0) * c->p.mach;
  ns.Ts = v->sutherland_ts;
  ns.bulk = v->bulk_viscosity;
  ns.ooPr = v->conductivity_model == AEROFX_COND_CONSTANT_PRANDTL ? 1.0 / v->prandtl : 0.0;
  ns.kcond = v->conductivity;
  ns.wall_T = v->wall_temperature;
  ns.viscous_cst = v->viscous_cst;
  // the wall nodes: node type = code of the highest-priority face touching
  // the node (Domain::setNodeType, Face::setNodeType)
  const int N = c->N, F = c->F;
  std::vector<int> type(N, 0);
  for (int f = 0; f < F; ++f) {
    const int code = c->fcodes[f];
    if (code < -10 || code > 11) continue;
    for (int a = 0; a < 3; ++a) {
      const int n = c->faces[3 * (size_t)f + a];
      if (bc_priority(code) > bc_priority(type[n])) type[n] = code;
    }
  }
  std::vector<int2> wall;
  for (int n = 0; n < N; ++n)
    if (type[n] == 1 || type[n] == -1 || type[n] == 3 || type[n] == -3) wall.push_back(make_int2(n, type[n]));
  if (!wall.empty() && (int)wall.size() > c->wall_cap) {
    void *p = nullptr;
    if ((rc = ctx_alloc(c, sizeof(int2) * wall.size(), &p))) return rc;
    if (c->d_wall) ctx_release(c, c->d_wall);
    c->d_wall = (int2 *)p;
    c->wall_cap = (int)wall.size();
  }
  if (!wall.empty()) {
    hipError_t e = hipMemcpyAsync(c->d_wall, wall.data(), sizeof(int2) * wall.size(), hipMemcpyHostToDevice, c->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    if (e != hipSuccess) return set_hip_error(e, "set_viscous upload");
  }
  // the element of every face (FaceTria::elemNum) and, for the wall faces,
  // the slots (4 f + k) of their elements' vertices per node, in face order
  if (!c->d_face_elem && F > 0) {
    std::vector<int> nrow(N + 1, 0), nlist(4 * (size_t)c->T);
    for (size_t k = 0; k < 4 * (size_t)c->T; ++k) nrow[c->tets[k] + 1]++;
    for (int n = 0; n < N; ++n) nrow[n + 1] += nrow[n];
    {
      std::vector<int> fill(nrow.begin(), nrow.end() - 1);
      for (size_t k = 0; k < 4 * (size_t)c->T; ++k) nlist[fill[c->tets[k]]++] = (int)(k / 4);
    }
    std::vector<int> felem(F, -1), frow(N + 1, 0), fslot;
    for (int f = 0; f < F; ++f) {
      const int a = c->faces[3 * (size_t)f], b = c->faces[3 * (size_t)f + 1], d = c->faces[3 * (size_t)f + 2];
      for (int s = nrow[a]; s < nrow[a + 1] && felem[f] < 0; ++s) {
        const int t = nlist[s];
        const int *tt = &c->tets[4 * (size_t)t];
        bool hb = false, hd = false;
        for (int q = 0; q < 4; ++q) {
          hb = hb || tt[q] == b;
          hd = hd || tt[q] == d;
        }
        if (hb && hd) felem[f] = t;
      }
    }
    std::vector<std::vector<int>> per(N);
    for (int f = 0; f < F; ++f) {
      const int code = c->fcodes[f];
      const int ac = code < 0 ? -code : code;
      if (felem[f] < 0 || !(ac == 1 || ac == 2 || ac == 3)) continue;
      for (int k = 0; k < 4; ++k) per[c->tets[4 * (size_t)felem[f] + k]].push_back(4 * f + k);
    }
    for (int n = 0; n < N; ++n) frow[n + 1] = frow[n] + (int)per[n].size();
    for (int n = 0; n < N; ++n) fslot.insert(fslot.end(), per[n].begin(), per[n].end());
    if (fslot.empty()) fslot.push_back(0);
    void *p1 = nullptr, *p2 = nullptr, *p3 = nullptr, *p4 = nullptr, *p5 = nullptr;
    if ((rc = ctx_alloc(c, sizeof(int) * F, &p1)) || (rc = ctx_alloc(c, sizeof(int) * (N + 1), &p2)) ||
        (rc = ctx_alloc(c, sizeof(int) * fslot.size(), &p3)) ||
        (rc = ctx_alloc(c, sizeof(double) * 20 * (size_t)F, &p4)) ||
        (rc = ctx_alloc(c, sizeof(double) * 12 * (size_t)F, &p5)))
      return rc;
    hipError_t e = hipMemcpyAsync(p1, felem.data(), sizeof(int) * F, hipMemcpyHostToDevice, c->stream);
    if (e == hipSuccess) e = hipMemcpyAsync(p2, frow.data(), sizeof(int) * (N + 1), hipMemcpyHostToDevice, c->stream);
    if (e == hipSuccess)
      e = hipMemcpyAsync(p3, fslot.data(), sizeof(int) * fslot.size(), hipMemcpyHostToDevice, c->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    if (e != hipSuccess) return set_hip_error(e, "set_viscous face tables");
    c->d_face_elem = (int *)p1;
    c->d_nfe_row = (int *)p2;
    c->d_nfe_slot = (int *)p3;
    c->d_force_u = (double *)p4;
    c->d_force_x = (double *)p5;
  }
  c->node_type = type;
  c->nwall = (int)wall.size();
  c->ns = ns;
  c->viscous = 1;
  return AEROFX_OK;
}

int aerofx_ctx_node_types(const aerofx_ctx *c, int *types) {
  if (!c || !types) return set_error(AEROFX_ERR_ARG, "node_types: null argument");
  if (!c->viscous) return set_error(AEROFX_ERR_STATE, "node_types: the Navier-Stokes terms are off");
  std::memcpy(types, c->node_type.data(), sizeof(int) * (size_t)c->N);
  return AEROFX_OK;
}

int aerofx_apply_bcs_residual(aerofx_ctx *c, const double *U, double *F) {
  if (!c || !U || !F) return set_error(AEROFX_ERR_ARG, "apply_bcs_residual: null argument");
  int rc = launch_bcs_residual(c, U, F, nullptr);
  if (rc) return rc;
  hipError_t e = hipStreamSynchronize(c->stream);
  return e == hipSuccess ? AEROFX_OK : set_hip_error(e, "apply_bcs_residual sync");
}

int aerofx_apply_bcs_solution(aerofx_ctx *c, double *U) {
  if (!c || !U) return set_error(AEROFX_ERR_ARG, "apply_bcs_solution: null argument");
  if (c->viscous && c->nwall) {
    hipLaunchKernelGGL(k_bcs_solution, dim3((c->nwall + 63) / 64), dim3(64), 0, c->stream, c->nwall, c->d_wall,
                       c->ns.wall_T, U);
    int rc = check("wall solution launch");
    if (rc) return rc;
  }
  hipError_t e = hipStreamSynchronize(c->stream);
  return e == hipSuccess ? AEROFX_OK : set_hip_error(e, "apply_bcs_solution sync");
}

}  // extern "C"
