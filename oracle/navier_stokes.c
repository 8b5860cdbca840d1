/* Laminar Navier-Stokes terms of the residual and of H1, restated on the
 * CPU.  TEST INFRASTRUCTURE: the oracle the GPU kernels are checked against;
 * the product never links it.
 *
 * What is restated (perfect gas, dim 5, no turbulence model, full wall
 * integration, fixed walls):
 *   - the P1 shape-function gradients and signed volume of a tetrahedron
 *     (ElemTet::computeGradientP1Function, ElemTetCore.C:1961-2012);
 *   - the transport coefficients (FemEquationTermNS::computeTransportCoefficients,
 *     FemEquationTermDesc.C:46-54; ViscoFcn.h constant / Sutherland /
 *     Prandtl viscosity; ThermalCondFcn.h constant-Prandtl and constant
 *     conductivity), all scaled by 1/Re (FemEquationTermDesc.C:126-129);
 *   - the volume term r[3][5] (FemEquationTermNS::computeVolumeTerm,
 *     FemEquationTermDesc.C:98-158, NavierStokesTerm::computeVolumeTermNS,
 *     NavierStokesTerm.h:690-725, with the stress tensor and heat flux of
 *     NavierStokesTerm.h:561-576, 647-654) and its assembly
 *     R_j += vol (r . dphi_j) (ElemTet::computeGalerkinTerm,
 *     ElemTet.C:24-120, non-embedded branch) in element order;
 *   - the exact derivative of r with respect to the conservative states of
 *     the four nodes (FemEquationTermNS::computeJacobianVolumeTerm,
 *     FemEquationTermDesc.C:338-449, completeJac for a perfect gas;
 *     NavierStokesTerm::computeJacobianVolumeTermNS, NavierStokesTerm.h:
 *     825-1146) and its assembly into H1 (ElemTet::computeJacobianGalerkinTerm,
 *     ElemTet.C:1912-1957: diagonal blocks unscaled, off-diagonal blocks
 *     times 1/V of their row; SpaceOperator::computeJacobian runs it before
 *     the finite-volume term, SpaceOperator.C:2131-2136);
 *   - the elemental viscous time-step term (ElemSet::computeTimeStep,
 *     Elem.C:25-83) weighted by ViscosityParameter (Domain.C:91);
 *   - the node types (Domain::setNodeType priorities, DomainCore.C:604-670;
 *     Face::setNodeType, FaceCore.C:83-95) and BcFcnNS's strong wall
 *     conditions on the solution, the residual and the Jacobian
 *     (BcFcnCore.C:212-265, BcFcn.C:8-122; SubDomain::applyBCsToResidual,
 *     SubDomain.C:2852-2885; SubDomain::applyBCsToJacobian,
 *     SubDomain.C:3060-3091).
 * The adiabatic-wall face term (FaceTria::computeGalerkinTerm with
 * NavierStokesTerm::computeSurfaceTermNS, NavierStokesTerm.h) is the wall
 * velocity times the wall stress: zero for fixed walls, so it adds nothing.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "aerof_oracle.h"

static const double kTwoThird = 2.0 / 3.0;
static const double kFourth = 1.0 / 4.0;
static const double kSixth = 1.0 / 6.0;

/* ElemTet::computeGradientP1Function (ElemTetCore.C:1961-2012): the
 * Jacobian of the map from the unit tetrahedron, its determinant, and the
 * gradients of the four P1 shape functions from its inverse. */
double or_p1_gradients(const double *X, const int *tet, double g[4][3]) {
  const double *x0 = X + 3 * (size_t)tet[0];
  double J[3][3];
  for (int c = 0; c < 3; ++c) {
    const double *xc = X + 3 * (size_t)tet[c + 1];
    for (int r = 0; r < 3; ++r) J[r][c] = xc[r] - x0[r];
  }
  double det = J[0][0] * (J[1][1] * J[2][2] - J[1][2] * J[2][1]) + J[1][0] * (J[0][2] * J[2][1] - J[0][1] * J[2][2]) +
               J[2][0] * (J[0][1] * J[1][2] - J[0][2] * J[1][1]);
  double s = -1.0 / det;
  g[1][0] = (-J[1][1] * J[2][2] + J[1][2] * J[2][1]) * s;
  g[1][1] = (J[0][1] * J[2][2] - J[0][2] * J[2][1]) * s;
  g[1][2] = -(J[0][1] * J[1][2] - J[0][2] * J[1][1]) * s;
  g[2][0] = -(-J[1][0] * J[2][2] + J[1][2] * J[2][0]) * s;
  g[2][1] = -(J[0][0] * J[2][2] - J[0][2] * J[2][0]) * s;
  g[2][2] = (J[0][0] * J[1][2] - J[0][2] * J[1][0]) * s;
  g[3][0] = -(J[1][0] * J[2][1] - J[1][1] * J[2][0]) * s;
  g[3][1] = (J[0][0] * J[2][1] - J[0][1] * J[2][0]) * s;
  g[3][2] = -(J[0][0] * J[1][1] - J[0][1] * J[1][0]) * s;
  for (int d = 0; d < 3; ++d) g[0][d] = -(g[1][d] + g[2][d] + g[3][d]);
  return kSixth * det;
}

/* VarFcnSGEuler::computeTemperature (VarFcnSGEuler.h:120-126). */
static double temperature(const or_params *p, const double *V) {
  double invgam1 = 1.0 / (p->gamma - 1.0);
  return invgam1 * (V[4] + p->pstiff) / V[0];
}

/* ViscoFcn::compute_mu and the Sutherland / Prandtl scale
 * alpha = gam (gam - 1) M^2 (ViscoFcn.h). */
static double visc_alpha(const or_params *p) { return p->gamma * (p->gamma - 1.0) * p->mach * p->mach; }

static double visc_mu(const or_params *p, double Tadim) {
  if (p->visc_model == 1) {
    double T = visc_alpha(p) * Tadim, Ts = p->sutherland_ts;
    return T * sqrt(T) * (1.0 + Ts) / (T + Ts);
  }
  if (p->visc_model == 2) return visc_alpha(p) * Tadim;
  return 1.0;
}

/* compute_muDerivative with dMach = 0 (ViscoFcn.h). */
static double visc_dmu(const or_params *p, double Tadim, double dTadim) {
  if (p->visc_model == 1) {
    double alpha = visc_alpha(p), Ts = p->sutherland_ts;
    double dalpha = 2.0 * p->gamma * (p->gamma - 1.0) * p->mach;
    double dAlpha = dalpha * 0.0;
    double T = alpha * Tadim;
    double dT = dAlpha * Tadim + alpha * dTadim;
    return (1.5 * (1.0 + Ts) * sqrt(T) * dT * (T + Ts) - T * sqrt(T) * (1.0 + Ts) * dT) / ((T + Ts) * (T + Ts));
  }
  if (p->visc_model == 2) {
    double dalpha = 2.0 * p->gamma * (p->gamma - 1.0) * p->mach;
    return dalpha * 0.0 * Tadim + visc_alpha(p) * dTadim;
  }
  return 0.0;
}

static double visc_lambda(const or_params *p, double mu) {
  if (p->visc_model == 0) return p->bulk_viscosity - kTwoThird * mu;
  return -kTwoThird * mu;
}

/* ThermalCondFcn::compute; cp = gam for the perfect gas
 * (VarFcnSGEuler::specificHeatCstPressure). */
static double cond_kappa(const or_params *p, double Tadim) {
  if (p->cond_model == 1) return p->conductivity;
  return 1.0 / p->prandtl * p->gamma * visc_mu(p, Tadim);
}

static double cond_dkappa(const or_params *p, double Tadim, double dTadim) {
  if (p->cond_model == 1) return 0.0;
  return 1.0 / p->prandtl * p->gamma * visc_dmu(p, Tadim, dTadim);
}

void or_ns_transport(const or_params *p, double T, double *mu, double *lambda, double *kappa) {
  double ooRe = 1.0 / p->reynolds;
  double m = visc_mu(p, T);
  double l = visc_lambda(p, m);
  double k = cond_kappa(p, T);
  *mu = m * ooRe;
  *lambda = l * ooRe;
  *kappa = k * ooRe;
}

/* The element quantities shared by the residual and the Jacobian:
 * velocities, temperatures, their P1 gradients, transport coefficients. */
typedef struct {
  double u[4][3], ucg[3], T[4], Tcg, dudxj[3][3], dTdxj[3], mu, lambda, kappa, tij[3][3];
} ns_elem;

static void ns_element(const or_params *p, double g[4][3], const double *v[4], ns_elem *e) {
  for (int k = 0; k < 4; ++k)
    for (int i = 0; i < 3; ++i) e->u[k][i] = v[k][i + 1];
  for (int i = 0; i < 3; ++i) e->ucg[i] = kFourth * (e->u[0][i] + e->u[1][i] + e->u[2][i] + e->u[3][i]);
  for (int k = 0; k < 4; ++k) e->T[k] = temperature(p, v[k]);
  e->Tcg = kFourth * (e->T[0] + e->T[1] + e->T[2] + e->T[3]);
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j)
      e->dudxj[i][j] = g[0][j] * e->u[0][i] + g[1][j] * e->u[1][i] + g[2][j] * e->u[2][i] + g[3][j] * e->u[3][i];
  for (int j = 0; j < 3; ++j)
    e->dTdxj[j] = g[0][j] * e->T[0] + g[1][j] * e->T[1] + g[2][j] * e->T[2] + g[3][j] * e->T[3];
  or_ns_transport(p, e->Tcg, &e->mu, &e->lambda, &e->kappa);
  /* NavierStokesTerm::computeStressTensor */
  double div = e->dudxj[0][0] + e->dudxj[1][1] + e->dudxj[2][2];
  for (int i = 0; i < 3; ++i) e->tij[i][i] = e->lambda * div + 2.0 * e->mu * e->dudxj[i][i];
  e->tij[0][1] = e->mu * (e->dudxj[1][0] + e->dudxj[0][1]);
  e->tij[0][2] = e->mu * (e->dudxj[2][0] + e->dudxj[0][2]);
  e->tij[1][2] = e->mu * (e->dudxj[2][1] + e->dudxj[1][2]);
  e->tij[1][0] = e->tij[0][1];
  e->tij[2][0] = e->tij[0][2];
  e->tij[2][1] = e->tij[1][2];
}

/* r[j][.] = (0, tau_0j, tau_1j, tau_2j, u.tau_.j - q_j), q = -kappa grad T. */
void or_ns_volume_term(const or_params *p, double g[4][3], const double *v[4], double r[3][5]) {
  ns_elem e;
  ns_element(p, g, v, &e);
  double qj[3];
  for (int j = 0; j < 3; ++j) qj[j] = -e.kappa * e.dTdxj[j];
  for (int j = 0; j < 3; ++j) {
    r[j][0] = 0.0;
    r[j][1] = e.tij[0][j];
    r[j][2] = e.tij[1][j];
    r[j][3] = e.tij[2][j];
    r[j][4] = e.ucg[0] * e.tij[0][j] + e.ucg[1] * e.tij[1][j] + e.ucg[2] * e.tij[2][j] - qj[j];
  }
}

/* d r / d U_k for the four nodes k: dRdU[k][j][m][c] = d r[j][m] / d U_k[c]. */
void or_ns_volume_jacobian(const or_params *p, double g[4][3], const double *v[4], double dRdU[4][3][5][5]) {
  ns_elem e;
  ns_element(p, g, v, &e);
  double ooRe = 1.0 / p->reynolds;
  for (int k = 0; k < 4; ++k) {
    const double *V = v[k];
    double rho = V[0];
    /* dTcg/dU_k (completeJac) and the coefficient derivatives */
    double dTcg[5];
    dTcg[0] = -0.25 / rho * (e.T[k] - 0.5 * (V[1] * V[1] + V[2] * V[2] + V[3] * V[3]));
    dTcg[1] = -0.25 / rho * V[1];
    dTcg[2] = -0.25 / rho * V[2];
    dTcg[3] = -0.25 / rho * V[3];
    dTcg[4] = 0.25 / rho;
    double dmu[5], dlam[5], dkap[5];
    for (int c = 0; c < 5; ++c) {
      dkap[c] = ooRe * cond_dkappa(p, e.Tcg, dTcg[c]);
      dmu[c] = visc_dmu(p, e.Tcg, dTcg[c]);
      dlam[c] = ooRe * (-kTwoThird * dmu[c]);
      dmu[c] *= ooRe;
    }
    /* d u_cg / d U_k */
    double ducg[5][3];
    for (int c = 0; c < 5; ++c)
      for (int i = 0; i < 3; ++i) ducg[c][i] = 0.0;
    for (int i = 0; i < 3; ++i) {
      ducg[0][i] = -kFourth * V[i + 1] / rho;
      ducg[i + 1][i] = kFourth / rho;
    }
    /* d tau / d U_k: the velocity gradients depend on U_k through
     * u_k = m_k / rho_k (d u_k / d rho = -u_k / rho, d u_k / d m = 1 / rho) */
    const double *gk = g[k];
    double div = e.dudxj[0][0] + e.dudxj[1][1] + e.dudxj[2][2];
    double dt[5][3][3];
    for (int c = 0; c < 5; ++c) {
      /* d(du_i/dx_j)/dU_k[c] = gk[j] * d u_k[i] / dU_k[c] */
      double dg[3][3];
      for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) {
          double du = 0.0;
          if (c == 0) du = -V[i + 1] / rho;
          else if (c == i + 1) du = 1.0 / rho;
          dg[i][j] = gk[j] * du;
        }
      double ddiv = dg[0][0] + dg[1][1] + dg[2][2];
      for (int i = 0; i < 3; ++i)
        dt[c][i][i] = 2.0 * dmu[c] * e.dudxj[i][i] + 2.0 * e.mu * dg[i][i] + dlam[c] * div + e.lambda * ddiv;
      dt[c][0][1] = e.mu * (dg[1][0] + dg[0][1]) + dmu[c] * (e.dudxj[1][0] + e.dudxj[0][1]);
      dt[c][0][2] = e.mu * (dg[2][0] + dg[0][2]) + dmu[c] * (e.dudxj[2][0] + e.dudxj[0][2]);
      dt[c][1][2] = e.mu * (dg[2][1] + dg[1][2]) + dmu[c] * (e.dudxj[2][1] + e.dudxj[1][2]);
      dt[c][1][0] = dt[c][0][1];
      dt[c][2][0] = dt[c][0][2];
      dt[c][2][1] = dt[c][1][2];
    }
    /* d q / d U_k: q_j = -kappa dT/dx_j, dT_k/dU_k from T = (E/rho - |u|^2/2) */
    double dTk[5];
    dTk[0] = -1.0 / rho * (e.T[k] - 0.5 * (V[1] * V[1] + V[2] * V[2] + V[3] * V[3]));
    dTk[1] = -1.0 / rho * V[1];
    dTk[2] = -1.0 / rho * V[2];
    dTk[3] = -1.0 / rho * V[3];
    dTk[4] = 1.0 / rho;
    for (int j = 0; j < 3; ++j)
      for (int c = 0; c < 5; ++c) {
        double dq = -dkap[c] * e.dTdxj[j] - e.kappa * (gk[j] * dTk[c]);
        dRdU[k][j][0][c] = 0.0;
        dRdU[k][j][1][c] = dt[c][0][j];
        dRdU[k][j][2][c] = dt[c][1][j];
        dRdU[k][j][3][c] = dt[c][2][j];
        dRdU[k][j][4][c] = ducg[c][0] * e.tij[0][j] + e.ucg[0] * dt[c][0][j] + ducg[c][1] * e.tij[1][j] +
                           e.ucg[1] * dt[c][1][j] + ducg[c][2] * e.tij[2][j] + e.ucg[2] * dt[c][2][j] - dq;
      }
  }
}

/* ElemSet::computeGalerkinTerm -> ElemTet::computeGalerkinTerm, element
 * order; Rabs (optional) receives the sums of |contribution|. */
void or_galerkin_term(const or_mesh *m, const or_params *p, const double *V, double *R, double *Rabs) {
  for (int t = 0; t < m->T; ++t) {
    const int *tet = m->tets + 4 * (size_t)t;
    double g[4][3], r[3][5];
    double vol = or_p1_gradients(m->X, tet, g);
    const double *v[4] = {V + 5 * (size_t)tet[0], V + 5 * (size_t)tet[1], V + 5 * (size_t)tet[2],
                          V + 5 * (size_t)tet[3]};
    or_ns_volume_term(p, g, v, r);
    for (int j = 0; j < 4; ++j)
      for (int k = 0; k < 5; ++k) {
        double c = vol * ((r[0][k] * g[j][0] + r[1][k] * g[j][1] + r[2][k] * g[j][2]) - kFourth * 0.0);
        R[5 * (size_t)tet[j] + k] += c;
        if (Rabs) Rabs[5 * (size_t)tet[j] + k] += fabs(c);
      }
  }
}

/* ElemTet::computeJacobianGalerkinTerm (non-embedded branch): Adiag gets the
 * unscaled diagonal sums, Aedge [E][2][25] (A_ij, A_ji), i < j, the
 * off-diagonal blocks times 1/V of their row. */
void or_galerkin_jacobian(const or_mesh *m, const or_params *p, const double *V, double *Adiag, double *Aedge) {
  static const int ends[6][2] = {{0, 1}, {0, 2}, {0, 3}, {1, 2}, {1, 3}, {2, 3}};
  for (int t = 0; t < m->T; ++t) {
    const int *tet = m->tets + 4 * (size_t)t;
    double g[4][3];
    double vol = or_p1_gradients(m->X, tet, g);
    const double *v[4] = {V + 5 * (size_t)tet[0], V + 5 * (size_t)tet[1], V + 5 * (size_t)tet[2],
                          V + 5 * (size_t)tet[3]};
    double dRdU[4][3][5][5];
    or_ns_volume_jacobian(p, g, v, dRdU);
    for (int i = 0; i < 4; ++i) {
      double *Aii = Adiag + 25 * (size_t)tet[i];
      for (int q = 0; q < 25; ++q) {
        int mm = q / 5, c = q % 5;
        Aii[q] += (dRdU[i][0][mm][c] * g[i][0] + dRdU[i][1][mm][c] * g[i][1] + dRdU[i][2][mm][c] * g[i][2]) * vol;
      }
    }
    for (int e = 0; e < 6; ++e) {
      int a = ends[e][0], b = ends[e][1];
      int i = tet[a] < tet[b] ? a : b, j = tet[a] < tet[b] ? b : a;
      int l = m->tet_edges[6 * (size_t)t + e];
      double *Aij = Aedge + 50 * (size_t)l, *Aji = Aij + 25;
      double cij = 1.0 / m->cv[tet[i]], cji = 1.0 / m->cv[tet[j]];
      for (int q = 0; q < 25; ++q) {
        int mm = q / 5, c = q % 5;
        Aij[q] += cij * (dRdU[j][0][mm][c] * g[i][0] + dRdU[j][1][mm][c] * g[i][1] + dRdU[j][2][mm][c] * g[i][2]) *
                  vol;
        Aji[q] += cji * (dRdU[i][0][mm][c] * g[j][0] + dRdU[i][1][mm][c] * g[j][1] + dRdU[i][2][mm][c] * g[j][2]) *
                  vol;
      }
    }
  }
}

/* ElemSet::computeTimeStep (Elem.C:25-83): idtv_j += 9 (mu/Re)/rho_mid |T|
 * |dphi_j|^2 with the element's mean state; ElemTet::computeVolume is the
 * signed volume (ElemTetCore.C:31-44). */
void or_viscous_idtv(const or_mesh *m, const or_params *p, const double *V, double *idtv) {
  double ooRe = 1.0 / p->reynolds;
  for (int t = 0; t < m->T; ++t) {
    const int *tet = m->tets + 4 * (size_t)t;
    double Vmid[5];
    for (int k = 0; k < 5; ++k) {
      Vmid[k] = 0.0;
      for (int j = 0; j < 4; ++j) Vmid[k] += V[5 * (size_t)tet[j] + k];
      Vmid[k] *= 1.0 / 4.0;
    }
    double Tm = temperature(p, Vmid);
    double mul = ooRe * visc_mu(p, Tm);
    const double *x[4] = {m->X + 3 * (size_t)tet[0], m->X + 3 * (size_t)tet[1], m->X + 3 * (size_t)tet[2],
                          m->X + 3 * (size_t)tet[3]};
    double v1[3], v2[3], v3[3];
    for (int d = 0; d < 3; ++d) {
      v1[d] = x[1][d] - x[0][d];
      v2[d] = x[2][d] - x[0][d];
      v3[d] = x[3][d] - x[0][d];
    }
    double cr[3] = {v1[1] * v2[2] - v1[2] * v2[1], v1[2] * v2[0] - v1[0] * v2[2], v1[0] * v2[1] - v1[1] * v2[0]};
    double volume = kSixth * (v3[0] * cr[0] + v3[1] * cr[1] + v3[2] * cr[2]);
    double localDt = 9.0 * (mul / Vmid[0]) * volume;
    double g[4][3];
    or_p1_gradients(m->X, tet, g);
    for (int j = 0; j < 4; ++j)
      idtv[tet[j]] += localDt * (g[j][0] * g[j][0] + g[j][1] * g[j][1] + g[j][2] * g[j][2]);
  }
}

/* Domain::setNodeType with its priority table (DomainCore.C:604-670):
 * each node takes the code of the highest-priority face touching it. */
static int bc_priority(int code) {
  switch (code) {
    case -3: return 20;
    case -1: return 19;
    case -2: return 18;
    case -6: return 17;
    case -9: return 16;
    case -10: return 15;
    case -7: return 14;
    case -8: return 13;
    case -4: return 12;
    case -5: return 11;
    case 3: return 10;
    case 1: return 9;
    case 2: return 8;
    case 6: return 7;
    case 9: return 6;
    case 10: return 5;
    case 7: return 4;
    case 8: return 3;
    case 4: return 2;
    case 5: return 1;
    case 11: return 0;
    default: return -1;
  }
}

void or_node_types(const or_mesh *m, int *type) {
  for (int i = 0; i < m->N; ++i) type[i] = 0;
  for (int f = 0; f < m->F; ++f) {
    int code = m->codes[f];
    if (code < -10 || code > 11) continue;
    for (int a = 0; a < 3; ++a) {
      int n = m->faces[3 * f + a];
      if (bc_priority(code) > bc_priority(type[n])) type[n] = code;
    }
  }
}

static int is_wall(int t) { return t == 1 || t == -1 || t == 3 || t == -3; }
static int is_isothermal(int t) { return t == 1 || t == -1; }

/* BcFcnNS::applyToSolutionVector on the wall nodes (fixed walls: the wall
 * velocity is zero; the wall temperature is p->wall_temperature). */
void or_apply_bcs_solution(const or_mesh *m, const or_params *p, const int *type, double *U) {
  const double Vw[5] = {0.0, 0.0, 0.0, 0.0, p->wall_temperature};
  for (int i = 0; i < m->N; ++i) {
    if (!is_wall(type[i])) continue;
    double *u = U + 5 * (size_t)i;
    u[1] = u[0] * Vw[1];
    u[2] = u[0] * Vw[2];
    u[3] = u[0] * Vw[3];
    if (is_isothermal(type[i])) u[4] = u[0] * (Vw[4] + 0.5 * (Vw[1] * Vw[1] + Vw[2] * Vw[2] + Vw[3] * Vw[3]));
  }
}

/* BcFcnNS::applyToResidualTerm (BcFcnCore.C:230-245). */
void or_apply_bcs_residual(const or_mesh *m, const or_params *p, const int *type, const double *U, double *F) {
  const double Vw[5] = {0.0, 0.0, 0.0, 0.0, p->wall_temperature};
  for (int i = 0; i < m->N; ++i) {
    if (!is_wall(type[i])) continue;
    const double *u = U + 5 * (size_t)i;
    double *f = F + 5 * (size_t)i;
    f[1] = -Vw[1] * u[0] + u[1];
    f[2] = -Vw[2] * u[0] + u[2];
    f[3] = -Vw[3] * u[0] + u[3];
    if (is_isothermal(type[i])) f[4] = (u[4] - u[0] * Vw[4]) * u[0] - 0.5 * (u[1] * u[1] + u[2] * u[2] + u[3] * u[3]);
  }
}

/* BcFcnNS::applyToOffDiagonalTerm / applyToDiagonalTerm (BcFcn.C:8-122)
 * over the edge blocks of the wall rows and the wall diagonals. */
void or_apply_bcs_jacobian(const or_mesh *m, const or_params *p, const int *type, const double *U, double *Adiag,
                           double *Aedge) {
  const double Vw[5] = {0.0, 0.0, 0.0, 0.0, p->wall_temperature};
  for (int l = 0; l < m->E; ++l) {
    int i = m->ptr[2 * l], j = m->ptr[2 * l + 1];
    double *blk[2] = {Aedge + 50 * (size_t)l, Aedge + 50 * (size_t)l + 25};
    int row[2] = {i, j};
    for (int s = 0; s < 2; ++s) {
      if (!is_wall(type[row[s]])) continue;
      for (int k = 0; k < 5; ++k) {
        blk[s][5 + k] = 0.0;
        blk[s][10 + k] = 0.0;
        blk[s][15 + k] = 0.0;
        if (is_isothermal(type[row[s]])) blk[s][20 + k] = 0.0;
      }
    }
  }
  for (int i = 0; i < m->N; ++i) {
    if (!is_wall(type[i])) continue;
    double *A = Adiag + 25 * (size_t)i;
    const double *u = U + 5 * (size_t)i;
    for (int k = 0; k < 5; ++k) {
      A[5 + k] = 0.0;
      A[10 + k] = 0.0;
      A[15 + k] = 0.0;
      if (is_isothermal(type[i])) A[20 + k] = 0.0;
    }
    A[5] = -Vw[1];
    A[6] = 1.0;
    A[10] = -Vw[2];
    A[12] = 1.0;
    A[15] = -Vw[3];
    A[18] = 1.0;
    if (is_isothermal(type[i])) {
      A[20] = u[4] - 2.0 * Vw[4] * u[0];
      A[21] = -u[1];
      A[22] = -u[2];
      A[23] = -u[3];
      A[24] = u[0];
    }
  }
}
