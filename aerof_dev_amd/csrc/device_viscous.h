// device_viscous.h -- laminar Navier-Stokes element terms (host and device).
//
// One P1 tetrahedron of FemEquationTermNS with the reference's operation
// order (built with -ffp-contract=off, so host and device round alike):
//   * shape-function gradients and signed volume from the inverse of the
//     element map (ElemTet::computeGradientP1Function, ElemTetCore.C:1961-2012);
//   * velocities and temperatures at the nodes and their element means,
//     P1 velocity / temperature gradients (NavierStokesTerm.h:205-225,
//     283-292, 357-388, 485-498);
//   * mu, lambda, kappa at the mean temperature (ViscoFcn.h, ThermalCondFcn.h,
//     FemEquationTermDesc.C:46-54), times 1/Re;
//   * the stress tensor and heat flux (NavierStokesTerm.h:561-576, 647-654)
//     and the volume term r[3][5] (NavierStokesTerm.h:690-725);
//   * the column blocks of d r / d U_k contracted with one node's gradient,
//     sum_j dRdU[k][j] g[j] (FemEquationTermDesc.C:338-449,
//     NavierStokesTerm.h:825-1146, completeJac for the perfect gas).
#pragma once
#include <hip/hip_runtime.h>

namespace aerofx {

// Constants of the viscous terms, fixed at aerofx_ctx_set_viscous.
struct NsConsts {
  int visc;      // 0 constant, 1 Sutherland, 2 Prandtl
  int cond;      // 0 constant Prandtl number, 1 constant conductivity
  double ooRe;   // 1 / Re
  double gam, invgam1, pstiff;
  double alpha, dalpha;  // gam (gam - 1) M^2 and its Mach derivative
  double Ts, bulk, ooPr, kcond;
  double wall_T;         // isothermal wall temperature
  double viscous_cst;    // weight of the viscous time-step term
};

__host__ __device__ inline double ns_temperature(const NsConsts &c, const double *V) {
  return c.invgam1 * (V[4] + c.pstiff) / V[0];
}

__host__ __device__ inline double ns_mu(const NsConsts &c, double Tadim) {
  if (c.visc == 1) {
    const double T = c.alpha * Tadim;
    return T * sqrt(T) * (1.0 + c.Ts) / (T + c.Ts);
  }
  if (c.visc == 2) return c.alpha * Tadim;
  return 1.0;
}

// compute_muDerivative with dMach = 0
__host__ __device__ inline double ns_dmu(const NsConsts &c, double Tadim, double dTadim) {
  if (c.visc == 1) {
    const double dAlpha = c.dalpha * 0.0;
    const double T = c.alpha * Tadim;
    const double dT = dAlpha * Tadim + c.alpha * dTadim;
    return (1.5 * (1.0 + c.Ts) * sqrt(T) * dT * (T + c.Ts) - T * sqrt(T) * (1.0 + c.Ts) * dT) /
           ((T + c.Ts) * (T + c.Ts));
  }
  if (c.visc == 2) return c.dalpha * 0.0 * Tadim + c.alpha * dTadim;
  return 0.0;
}

__host__ __device__ inline double ns_kappa(const NsConsts &c, double Tadim) {
  return c.cond == 1 ? c.kcond : c.ooPr * c.gam * ns_mu(c, Tadim);
}

__host__ __device__ inline double ns_dkappa(const NsConsts &c, double Tadim, double dTadim) {
  return c.cond == 1 ? 0.0 : c.ooPr * c.gam * ns_dmu(c, Tadim, dTadim);
}

// Gradients g[4][3] of the P1 shape functions of the element with vertex
// coordinates x[4][3]; returns its signed volume.
__host__ __device__ inline double p1_gradients(const double x[4][3], double g[4][3]) {
  double J[3][3];
#pragma unroll
  for (int col = 0; col < 3; ++col)
#pragma unroll
    for (int r = 0; r < 3; ++r) J[r][col] = x[col + 1][r] - x[0][r];
  const double det = J[0][0] * (J[1][1] * J[2][2] - J[1][2] * J[2][1]) +
                     J[1][0] * (J[0][2] * J[2][1] - J[0][1] * J[2][2]) +
                     J[2][0] * (J[0][1] * J[1][2] - J[0][2] * J[1][1]);
  const double s = -1.0 / det;
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

// Element state shared by the volume term and its Jacobian.
struct NsElem {
  double ucg[3], T[4], Tcg;
  double dudxj[3][3], dTdxj[3];
  double mu, lambda, kappa;
  double tij[3][3];
};

// v[k][5]: primitive states of the four vertices
__host__ __device__ inline void ns_element(const NsConsts &c, const double g[4][3], const double v[4][5],
                                           NsElem &e) {
  const double q = 1.0 / 4.0;
#pragma unroll
  for (int i = 0; i < 3; ++i) e.ucg[i] = q * (v[0][i + 1] + v[1][i + 1] + v[2][i + 1] + v[3][i + 1]);
#pragma unroll
  for (int k = 0; k < 4; ++k) e.T[k] = ns_temperature(c, v[k]);
  e.Tcg = q * (e.T[0] + e.T[1] + e.T[2] + e.T[3]);
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j)
      e.dudxj[i][j] = g[0][j] * v[0][i + 1] + g[1][j] * v[1][i + 1] + g[2][j] * v[2][i + 1] + g[3][j] * v[3][i + 1];
#pragma unroll
  for (int j = 0; j < 3; ++j) e.dTdxj[j] = g[0][j] * e.T[0] + g[1][j] * e.T[1] + g[2][j] * e.T[2] + g[3][j] * e.T[3];
  const double mu = ns_mu(c, e.Tcg);
  const double lambda = c.visc == 0 ? c.bulk - (2.0 / 3.0) * mu : -(2.0 / 3.0) * mu;
  const double kappa = ns_kappa(c, e.Tcg);
  e.mu = mu * c.ooRe;
  e.lambda = lambda * c.ooRe;
  e.kappa = kappa * c.ooRe;
  const double div = e.dudxj[0][0] + e.dudxj[1][1] + e.dudxj[2][2];
#pragma unroll
  for (int i = 0; i < 3; ++i) e.tij[i][i] = e.lambda * div + 2.0 * e.mu * e.dudxj[i][i];
  e.tij[0][1] = e.mu * (e.dudxj[1][0] + e.dudxj[0][1]);
  e.tij[0][2] = e.mu * (e.dudxj[2][0] + e.dudxj[0][2]);
  e.tij[1][2] = e.mu * (e.dudxj[2][1] + e.dudxj[1][2]);
  e.tij[1][0] = e.tij[0][1];
  e.tij[2][0] = e.tij[0][2];
  e.tij[2][1] = e.tij[1][2];
}

// Contribution vol (r . g_j) of the element to its vertex with gradient gj.
__host__ __device__ inline void ns_node_residual(const NsElem &e, const double gj[3], double vol, double out[5]) {
  double r[3][5];
#pragma unroll
  for (int j = 0; j < 3; ++j) {
    const double qj = -e.kappa * e.dTdxj[j];
    r[j][0] = 0.0;
    r[j][1] = e.tij[0][j];
    r[j][2] = e.tij[1][j];
    r[j][3] = e.tij[2][j];
    r[j][4] = e.ucg[0] * e.tij[0][j] + e.ucg[1] * e.tij[1][j] + e.ucg[2] * e.tij[2][j] - qj;
  }
#pragma unroll
  for (int k = 0; k < 5; ++k) out[k] = vol * ((r[0][k] * gj[0] + r[1][k] * gj[1] + r[2][k] * gj[2]) - 0.25 * 0.0);
}

// Column cc of d r / d U_k for the vertex k with primitive state V (the
// other vertices enter through e): col[j][m] = d r[j][m] / d U_k[cc].
__host__ __device__ inline void ns_node_column(const NsConsts &c, const NsElem &e, const double gk[3],
                                               const double V[5], const double Tk, int cc, double col[3][5]) {
  const double rho = V[0];
  const double usq = V[1] * V[1] + V[2] * V[2] + V[3] * V[3];
  double dTcg, dTk;
  if (cc == 0) {
    dTcg = -0.25 / rho * (Tk - 0.5 * usq);
    dTk = -1.0 / rho * (Tk - 0.5 * usq);
  } else if (cc < 4) {
    dTcg = -0.25 / rho * V[cc];
    dTk = -1.0 / rho * V[cc];
  } else {
    dTcg = 0.25 / rho;
    dTk = 1.0 / rho;
  }
  const double div = e.dudxj[0][0] + e.dudxj[1][1] + e.dudxj[2][2];
  const double dkap = c.ooRe * ns_dkappa(c, e.Tcg, dTcg);
  double dmu = ns_dmu(c, e.Tcg, dTcg);
  const double dlam = c.ooRe * (-(2.0 / 3.0) * dmu);
  dmu *= c.ooRe;
  // d u_k[i] / d U_k[cc] and d u_cg[i] / d U_k[cc]
  double du[3], ducg[3];
#pragma unroll
  for (int i = 0; i < 3; ++i) {
    du[i] = cc == 0 ? -V[i + 1] / rho : (cc == i + 1 ? 1.0 / rho : 0.0);
    ducg[i] = cc == 0 ? -0.25 * V[i + 1] / rho : (cc == i + 1 ? 0.25 / rho : 0.0);
  }
  double dg[3][3];
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j) dg[i][j] = gk[j] * du[i];
  const double ddiv = dg[0][0] + dg[1][1] + dg[2][2];
  double dt[3][3];
#pragma unroll
  for (int i = 0; i < 3; ++i)
    dt[i][i] = 2.0 * dmu * e.dudxj[i][i] + 2.0 * e.mu * dg[i][i] + dlam * div + e.lambda * ddiv;
  dt[0][1] = e.mu * (dg[1][0] + dg[0][1]) + dmu * (e.dudxj[1][0] + e.dudxj[0][1]);
  dt[0][2] = e.mu * (dg[2][0] + dg[0][2]) + dmu * (e.dudxj[2][0] + e.dudxj[0][2]);
  dt[1][2] = e.mu * (dg[2][1] + dg[1][2]) + dmu * (e.dudxj[2][1] + e.dudxj[1][2]);
  dt[1][0] = dt[0][1];
  dt[2][0] = dt[0][2];
  dt[2][1] = dt[1][2];
#pragma unroll
  for (int j = 0; j < 3; ++j) {
    const double dq = -dkap * e.dTdxj[j] - e.kappa * (gk[j] * dTk);
    col[j][0] = 0.0;
    col[j][1] = dt[0][j];
    col[j][2] = dt[1][j];
    col[j][3] = dt[2][j];
    col[j][4] = ducg[0] * e.tij[0][j] + e.ucg[0] * dt[0][j] + ducg[1] * e.tij[1][j] + e.ucg[1] * dt[1][j] +
                ducg[2] * e.tij[2][j] + e.ucg[2] * dt[2][j] - dq;
  }
}

// out[5 m + c] = sum_j dRdU[k][j][m][c] gi[j] for the vertex k
__host__ __device__ inline void ns_node_jacobian(const NsConsts &c, const NsElem &e, const double gk[3],
                                                 const double V[5], const double Tk, const double gi[3],
                                                 double out[25]) {
#pragma unroll
  for (int cc = 0; cc < 5; ++cc) {
    double col[3][5];
    ns_node_column(c, e, gk, V, Tk, cc, col);
#pragma unroll
    for (int m = 0; m < 5; ++m) out[5 * m + cc] = col[0][m] * gi[0] + col[1][m] * gi[1] + col[2][m] * gi[2];
  }
}

// out[c] = sum_j sum_m dRdU[k][j][m][c] w[j][m]: the transposed element
// derivative applied to w[j][m] = sum_i g_i[j] lambda_i[m]
__host__ __device__ inline void ns_node_adjoint(const NsConsts &c, const NsElem &e, const double gk[3],
                                                const double V[5], const double Tk, const double w[3][5],
                                                double out[5]) {
#pragma unroll
  for (int cc = 0; cc < 5; ++cc) {
    double col[3][5];
    ns_node_column(c, e, gk, V, Tk, cc, col);
    double s = 0.0;
#pragma unroll
    for (int j = 0; j < 3; ++j)
#pragma unroll
      for (int m = 0; m < 5; ++m) s += col[j][m] * w[j][m];
    out[cc] = s;
  }
}

}  // namespace aerofx
