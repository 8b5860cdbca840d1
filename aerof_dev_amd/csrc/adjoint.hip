// adjoint.hip -- transposed residual derivatives on gfx950:
//   dRdU^T lambda and dRdX^T lambda for the second-order residual
//   R_i = (1/V_i) (sum_edges +-F(V_ij, V_ji, n_ij) + sum_faces F_b(V_i, n_f/3))
// (SpaceOperator::computeTransposeDerivativeOfResidual, SpaceOperator.C:1235-1304,
//  and the sensitivity operators of EdgeSet, Edge.C:1452-1728, on the same
//  residual; the reference builds per-edge sparse operator blocks, here each
//  edge contracts its local derivatives with the adjoint weights at once).
//
// The MUSCL gradients are recomputed by a forward residual pass at the start
// of each call; nothing from an earlier forward evaluation is kept.
//
// Stages (all deterministic: per-edge, per-tet and per-face results go to
// their own slots and each node sums its slots in a fixed order):
//   1. forward residual: V, LSQ gradients, R (scaled)
//   2. k_adj_edges_u / _x  one thread per edge: mu = lambda/V, g = mu_i - mu_j;
//      g^T dF/d(V_ij, V_ji, n) by the reverse mode of the Roe flux,
//      then reverse through the Van Albada reconstruction by hand
//      -> per edge: Vbar_i, Vbar_j, Gbar_i, Gbar_j, dxbar, nbar
//   3. k_adj_nodes  one thread per node: sums the edge slots, adds the
//      boundary-face terms -> Vbar, Gbar, the dx part of Xbar
//   4. k_adj_lsq    one thread per node: reverse of the LSQ gradients
//      (grad_i = sum_j W_ij (V_j - V_i)), then Ubar = (dV/dU)^T Vbar; for X
//      also the weight and Cholesky-factor terms -> S_i = d/dM_i, symmetric
//   5. k_adj_tets / k_adj_faces: control-volume, dual-normal and face-normal
//      geometry, reverse over each element's coordinates
//   6. k_adj_xgather: Xbar_i = slots of 3-5 and -sum_j (S_i + S_j) dx_ij
#include <hip/hip_runtime.h>

#include "context.h"
#include "device_dual.h"
#include "device_sens.h"
#include "host_internal.h"

namespace aerofx {

struct Adjoint {
  double *EV = nullptr;   // [E][20] Vbar_i, c_i, Vbar_j, c_j (Gbar_i[d*5+k] = c_i[k] dx[d])
  double *EX = nullptr;   // [E][6]  dxbar, nbar (coordinate derivatives only)
  double *Vb = nullptr;   // [N][5]
  double *Gb = nullptr;   // [N][16]
  double *Xb = nullptr;   // [N][4]
  double *S = nullptr;    // [N][6]
  double *R = nullptr;    // [N][5] forward residual (scaled)
  double *TX = nullptr;   // [T][12]
  double *FX = nullptr;   // [F][9]
  double *Jf = nullptr;   // [F] per-face objective values, then the sum at [F]
  double *h_J = nullptr;  // pinned
  int *tets = nullptr;    // [T][4]
  int *tedge = nullptr;   // [T][6]
  int *faces = nullptr;   // [F][3]
  int *ntrow = nullptr;   // [N+1]
  int *ntslot = nullptr;  // [4T] 4*t + local vertex
  int *nfslot = nullptr;  // [3F] 3*f + local vertex, in the order of tab.fidx
};

namespace {

constexpr int kAdjBlock = 256;
// ElemTet edge -> (end, end) and edge -> (face, face) tables (ElemTetCore.C:19-21)
__constant__ int cEdgeEnd[6][2] = {{0, 1}, {0, 2}, {0, 3}, {1, 2}, {1, 3}, {2, 3}};
__constant__ int cEdgeFace[6][2] = {{0, 1}, {2, 0}, {1, 2}, {0, 3}, {3, 1}, {2, 3}};

// LSQ weights W(dx; L) as the gradient pass evaluates them.
__device__ __forceinline__ void lsq_weights(const double L[6], double dx0, double dx1, double dx2, double w[3]) {
  const double or11 = L[0], or22 = L[1], or33 = L[2], r12or11 = L[3], r23or22 = L[4], psi = L[5];
  const double alpha1 = dx0 * or11 * or11;
  const double alpha2 = (dx1 - r12or11 * dx0) * or22 * or22;
  const double alpha3 = (dx2 - r23or22 * dx1 + psi * dx0) * or33 * or33;
  w[0] = alpha1 - r12or11 * alpha2 + psi * alpha3;
  w[1] = alpha2 - r23or22 * alpha3;
  w[2] = alpha3;
}

// Reverse of lsq_weights: given Wbar, accumulate dxbar and (optionally) Lbar.
__device__ __forceinline__ void lsq_weights_rev(const double L[6], double dx0, double dx1, double dx2,
                                                const double wb[3], double dxb[3], double *Lb) {
  const double or11 = L[0], or22 = L[1], or33 = L[2], r12or11 = L[3], r23or22 = L[4], psi = L[5];
  const double a3b = wb[2] + psi * wb[0] - r23or22 * wb[1];
  const double a2b = wb[1] - r12or11 * wb[0];
  const double a1b = wb[0];
  const double s11 = or11 * or11, s22 = or22 * or22, s33 = or33 * or33;
  const double c3 = a3b * s33, c2 = a2b * s22;
  dxb[0] += a1b * s11 - c2 * r12or11 + c3 * psi;
  dxb[1] += c2 - c3 * r23or22;
  dxb[2] += c3;
  if (Lb) {
    const double alpha2 = (dx1 - r12or11 * dx0) * s22;
    const double alpha3 = (dx2 - r23or22 * dx1 + psi * dx0) * s33;
    const double t2 = dx1 - r12or11 * dx0;
    const double t3 = dx2 - r23or22 * dx1 + psi * dx0;
    Lb[0] += a1b * dx0 * 2.0 * or11;
    Lb[1] += a2b * t2 * 2.0 * or22;
    Lb[2] += a3b * t3 * 2.0 * or33;
    Lb[3] += -alpha2 * wb[0] - c2 * dx0;
    Lb[4] += -alpha3 * wb[1] - c3 * dx1;
    Lb[5] += alpha3 * wb[0] + c3 * dx0;
  }
}

__device__ __forceinline__ double seed(int dir, int k) { return dir == k ? 1.0 : 0.0; }

// Reverse mode of roeflux7 (device_physics.h) with evitno = 0, the edge
// call: given the weights g on the five flux components, returns
// bL = g^T dphi/dUg, bR = g^T dphi/dUd and be = g^T dphi/den.  The forward
// statements are repeated as written (one rounding each, the same upwind
// branch), then each is reversed in the opposite order; one forward and one
// reverse sweep replace the 13 directional passes of the Dual form.
__device__ __forceinline__ void roeflux7_vjp(const Gas &gs, const double e[3], const double L[5], const double R[5],
                                             const double g[5], double bL[5], double bR[5], double be[3]) {
  const double gam = gs.gam, gam1 = gs.gam1, invgam1 = gs.invgam1, ps = gs.pstiff;
  // ---- forward
  const double rn2 = e[0] * e[0] + e[1] * e[1] + e[2] * e[2];
  const double rnorm = sqrt(rn2);
  const double invnorm = 1.0 / rnorm;
  const double n0 = e[0] * invnorm, n1 = e[1] * invnorm, n2 = e[2] * invnorm;
  const double vitg2 = L[1] * L[1] + L[2] * L[2] + L[3] * L[3];
  const double vitd2 = R[1] * R[1] + R[2] * R[2] + R[3] * R[3];
  const double s1 = sqrt(L[0]), s2 = sqrt(R[0]);
  const double ener1 = (L[4] + gam * ps) * invgam1 + 0.5 * L[0] * vitg2;
  const double ener2 = (R[4] + gam * ps) * invgam1 + 0.5 * R[0] * vitd2;
  const double usro = 1.0 / (s1 + s2);
  const double m2 = s1 * L[1] + s2 * R[1], m3 = s1 * L[2] + s2 * R[2], m4n = s1 * L[3] + s2 * R[3];
  const double u2 = m2 * usro, u3 = m3 * usro, u4 = m4n * usro;
  const double A1 = (ener1 + L[4]) / s1, A2 = (ener2 + R[4]) / s2;
  const double u5 = (A1 + A2) * usro;
  const double VdotNt = n0 * u2 + n1 * u3 + n2 * u4;
  const bool left = VdotNt > 0.0;
  const double *S = left ? L : R;
  const double vit2 = left ? vitg2 : vitd2;
  const double VdotN = S[1] * n0 + S[2] * n1 + S[3] * n2;
  const double H0 = gam * (S[4] + ps) + 0.5 * gam1 * S[0] * vit2;
  const double H = H0 / (gam1 * S[0]);
  const double p0 = S[0] * VdotN;
  const double p1 = p0 * S[1] + S[4] * n0, p2 = p0 * S[2] + S[4] * n1, p3 = p0 * S[3] + S[4] * n2;
  const double p4 = p0 * H;
  const double qir = 0.5 * (u2 * u2 + u3 * u3 + u4 * u4);
  const double cr2 = gam1 * (u5 - qir);
  const double cr = sqrt(cr2);
  const double icr2 = 1.0 / cr2;
  const double vp4 = VdotNt + cr, vp5 = VdotNt - cr;
  const double r = cr, sm = -cr, cb = -cr, itb = -1.0 / cr;
  const double d1 = -L[0] + R[0];
  const double d2 = -L[0] * L[1] + R[0] * R[1];
  const double d3 = -L[0] * L[2] + R[0] * R[2];
  const double d4 = -L[0] * L[3] + R[0] * R[3];
  const double a1 = VdotNt * d1 - n0 * d2 - n1 * d3 - n2 * d4;
  const double a2 = qir * d1 - u2 * d2 - u3 * d3 - u4 * d4 + (ener2 - ener1);
  double flur4 = 0.0, flur5 = 0.0, mm = 0.0, tt = 0.0;
  if (!left) {
    mm = vp4 > 0.0 ? vp4 : 0.0;
    tt = cb * a1 + gam1 * sm * itb * a2;
    flur4 = mm * tt;
  } else {
    mm = vp5 < 0.0 ? vp5 : 0.0;
    tt = -cb * a1 + (-gam1 * r * itb * a2);
    flur5 = -mm * tt;
  }
  const double fs = flur4 + flur5;
  const double rs = r * flur4 + sm * flur5;
  const double X0 = fs, X1 = n0 * rs + u2 * fs, X2 = n1 * rs + u3 * fs, X3 = n2 * rs + u4 * fs,
               X4 = VdotNt * rs + u5 * fs;
  const double q0 = p0 - 0.5 * X0 * icr2, q1 = p1 - X1 * 0.5 * icr2, q2 = p2 - X2 * 0.5 * icr2,
               q3 = p3 - X3 * 0.5 * icr2, q4 = p4 - X4 * 0.5 * icr2;
  // ---- reverse
  double bLl[5] = {0.0, 0.0, 0.0, 0.0, 0.0}, bRr[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
  // phi_k = q_k rnorm
  double b_rnorm = g[0] * q0 + g[1] * q1 + g[2] * q2 + g[3] * q3 + g[4] * q4;
  const double bq0 = g[0] * rnorm, bq1 = g[1] * rnorm, bq2 = g[2] * rnorm, bq3 = g[3] * rnorm, bq4 = g[4] * rnorm;
  // q_k = p_k - X_k icr2 / 2
  const double c = 0.5 * icr2;
  const double b_icr2 = -0.5 * (bq0 * X0 + bq1 * X1 + bq2 * X2 + bq3 * X3 + bq4 * X4);
  const double bX0 = -c * bq0, bX1 = -c * bq1, bX2 = -c * bq2, bX3 = -c * bq3, bX4 = -c * bq4;
  double b_p0 = bq0;
  const double b_p1 = bq1, b_p2 = bq2, b_p3 = bq3, b_p4 = bq4;
  // X_k
  double b_fs = bX0 + u2 * bX1 + u3 * bX2 + u4 * bX3 + u5 * bX4;
  const double b_rs = n0 * bX1 + n1 * bX2 + n2 * bX3 + VdotNt * bX4;
  double b_n0 = rs * bX1, b_n1 = rs * bX2, b_n2 = rs * bX3;
  double b_u2 = fs * bX1, b_u3 = fs * bX2, b_u4 = fs * bX3, b_u5 = fs * bX4;
  double b_VdotNt = rs * bX4;
  // rs = r flur4 + sm flur5, fs = flur4 + flur5
  const double b_flur4 = b_fs + r * b_rs, b_flur5 = b_fs + sm * b_rs;
  double b_r = flur4 * b_rs, b_sm = flur5 * b_rs, b_cb = 0.0, b_itb = 0.0, b_a1 = 0.0, b_a2 = 0.0;
  double b_vp4 = 0.0, b_vp5 = 0.0;
  if (!left) {
    // flur4 = mm tt, tt = cb a1 + ((gam1 sm) itb) a2
    const double b_mm = tt * b_flur4, b_tt = mm * b_flur4;
    b_vp4 = vp4 > 0.0 ? b_mm : 0.0;
    b_cb += a1 * b_tt;
    b_a1 += cb * b_tt;
    const double k = gam1 * sm * itb;
    b_a2 += k * b_tt;
    const double b_k = a2 * b_tt;
    b_sm += gam1 * itb * b_k;
    b_itb += gam1 * sm * b_k;
  } else {
    // flur5 = -mm tt, tt = -cb a1 + (((-gam1) r) itb) a2
    const double b_mm = -tt * b_flur5, b_tt = -mm * b_flur5;
    b_vp5 = vp5 < 0.0 ? b_mm : 0.0;
    b_cb += -a1 * b_tt;
    b_a1 += -cb * b_tt;
    const double k = -gam1 * r * itb;
    b_a2 += k * b_tt;
    const double b_k = a2 * b_tt;
    b_r += -gam1 * itb * b_k;
    b_itb += -gam1 * r * b_k;
  }
  // a2 = qir d1 - u2 d2 - u3 d3 - u4 d4 + (ener2 - ener1)
  double b_qir = d1 * b_a2;
  double b_d1 = qir * b_a2, b_d2 = -u2 * b_a2, b_d3 = -u3 * b_a2, b_d4 = -u4 * b_a2;
  b_u2 += -d2 * b_a2;
  b_u3 += -d3 * b_a2;
  b_u4 += -d4 * b_a2;
  double b_ener1 = -b_a2, b_ener2 = b_a2;
  // a1 = VdotNt d1 - n0 d2 - n1 d3 - n2 d4
  b_VdotNt += d1 * b_a1;
  b_d1 += VdotNt * b_a1;
  b_n0 += -d2 * b_a1;
  b_n1 += -d3 * b_a1;
  b_n2 += -d4 * b_a1;
  b_d2 += -n0 * b_a1;
  b_d3 += -n1 * b_a1;
  b_d4 += -n2 * b_a1;
  // d1..d4
  bLl[0] += -b_d1 - L[1] * b_d2 - L[2] * b_d3 - L[3] * b_d4;
  bRr[0] += b_d1 + R[1] * b_d2 + R[2] * b_d3 + R[3] * b_d4;
  bLl[1] += -L[0] * b_d2;
  bLl[2] += -L[0] * b_d3;
  bLl[3] += -L[0] * b_d4;
  bRr[1] += R[0] * b_d2;
  bRr[2] += R[0] * b_d3;
  bRr[3] += R[0] * b_d4;
  // r = cr, sm = -cr, cb = -cr, itb = -1/cr; vp4 = VdotNt + cr, vp5 = VdotNt - cr
  double b_cr = b_r - b_sm - b_cb + b_itb / (cr * cr) + b_vp4 - b_vp5;
  b_VdotNt += b_vp4 + b_vp5;
  // icr2 = 1/cr2, cr = sqrt(cr2), cr2 = gam1 (u5 - qir)
  const double b_cr2 = -icr2 * icr2 * b_icr2 + 0.5 * b_cr / cr;
  b_u5 += gam1 * b_cr2;
  b_qir += -gam1 * b_cr2;
  // qir = (u2^2 + u3^2 + u4^2) / 2
  b_u2 += u2 * b_qir;
  b_u3 += u3 * b_qir;
  b_u4 += u4 * b_qir;
  // p1..p4 = p0 S_k + S4 n_k, p0 H
  b_p0 += S[1] * b_p1 + S[2] * b_p2 + S[3] * b_p3 + H * b_p4;
  double bS[5];
  bS[1] = p0 * b_p1;
  bS[2] = p0 * b_p2;
  bS[3] = p0 * b_p3;
  bS[4] = n0 * b_p1 + n1 * b_p2 + n2 * b_p3;
  b_n0 += S[4] * b_p1;
  b_n1 += S[4] * b_p2;
  b_n2 += S[4] * b_p3;
  const double b_H = p0 * b_p4;
  // p0 = S0 VdotN
  bS[0] = VdotN * b_p0;
  const double b_VdotN = S[0] * b_p0;
  // H = H0 / (gam1 S0), H0 = gam (S4 + ps) + gam1 S0 vit2 / 2
  const double b_H0 = b_H / (gam1 * S[0]);
  bS[0] += -H * b_H / S[0] + 0.5 * gam1 * vit2 * b_H0;
  bS[4] += gam * b_H0;
  const double b_vit2 = 0.5 * gam1 * S[0] * b_H0;
  // VdotN = S1 n0 + S2 n1 + S3 n2
  bS[1] += n0 * b_VdotN;
  bS[2] += n1 * b_VdotN;
  bS[3] += n2 * b_VdotN;
  b_n0 += S[1] * b_VdotN;
  b_n1 += S[2] * b_VdotN;
  b_n2 += S[3] * b_VdotN;
  double b_vitg2 = 0.0, b_vitd2 = 0.0;
  if (left) {
#pragma unroll
    for (int k = 0; k < 5; ++k) bLl[k] += bS[k];
    b_vitg2 += b_vit2;
  } else {
#pragma unroll
    for (int k = 0; k < 5; ++k) bRr[k] += bS[k];
    b_vitd2 += b_vit2;
  }
  // VdotNt = n0 u2 + n1 u3 + n2 u4
  b_n0 += u2 * b_VdotNt;
  b_n1 += u3 * b_VdotNt;
  b_n2 += u4 * b_VdotNt;
  b_u2 += n0 * b_VdotNt;
  b_u3 += n1 * b_VdotNt;
  b_u4 += n2 * b_VdotNt;
  // u5 = (A1 + A2) usro, A1 = (ener1 + L4) / s1, A2 = (ener2 + R4) / s2
  double b_usro = (A1 + A2) * b_u5;
  const double b_A1 = usro * b_u5, b_A2 = usro * b_u5;
  b_ener1 += b_A1 / s1;
  bLl[4] += b_A1 / s1;
  double b_s1 = -A1 / s1 * b_A1;
  b_ener2 += b_A2 / s2;
  bRr[4] += b_A2 / s2;
  double b_s2 = -A2 / s2 * b_A2;
  // u_c = m_c usro, m_c = s1 L_c + s2 R_c
  b_usro += m2 * b_u2 + m3 * b_u3 + m4n * b_u4;
  const double b_m2 = usro * b_u2, b_m3 = usro * b_u3, b_m4 = usro * b_u4;
  b_s1 += L[1] * b_m2 + L[2] * b_m3 + L[3] * b_m4;
  b_s2 += R[1] * b_m2 + R[2] * b_m3 + R[3] * b_m4;
  bLl[1] += s1 * b_m2;
  bLl[2] += s1 * b_m3;
  bLl[3] += s1 * b_m4;
  bRr[1] += s2 * b_m2;
  bRr[2] += s2 * b_m3;
  bRr[3] += s2 * b_m4;
  // usro = 1 / (s1 + s2)
  b_s1 += -usro * usro * b_usro;
  b_s2 += -usro * usro * b_usro;
  // ener = (p + gam ps) / gam1 + rho vit / 2
  bLl[4] += invgam1 * b_ener1;
  bLl[0] += 0.5 * vitg2 * b_ener1;
  b_vitg2 += 0.5 * L[0] * b_ener1;
  bRr[4] += invgam1 * b_ener2;
  bRr[0] += 0.5 * vitd2 * b_ener2;
  b_vitd2 += 0.5 * R[0] * b_ener2;
  // s = sqrt(rho)
  bLl[0] += 0.5 * b_s1 / s1;
  bRr[0] += 0.5 * b_s2 / s2;
  // vit2 = |u|^2
#pragma unroll
  for (int k = 1; k < 4; ++k) {
    bLl[k] += 2.0 * L[k] * b_vitg2;
    bRr[k] += 2.0 * R[k] * b_vitd2;
  }
  // n = e invnorm, invnorm = 1/rnorm, rnorm = sqrt(rn2)
  const double b_invnorm = e[0] * b_n0 + e[1] * b_n1 + e[2] * b_n2;
  b_rnorm += -invnorm * invnorm * b_invnorm;
  const double b_rn2 = 0.5 * b_rnorm / rnorm;
  be[0] = invnorm * b_n0 + 2.0 * e[0] * b_rn2;
  be[1] = invnorm * b_n1 + 2.0 * e[1] * b_rn2;
  be[2] = invnorm * b_n2 + 2.0 * e[2] * b_rn2;
#pragma unroll
  for (int k = 0; k < 5; ++k) {
    bL[k] = bLl[k];
    bR[k] = bRr[k];
  }
}

// d VA(a, b) / da and / db of the Van Albada function (RecFcn.h:269-281)
__device__ __forceinline__ void va_partials(double a, double b, double &da, double &db) {
  const double eps0 = 1.e-16;
  da = va_fcn(eps0, Dual(a, 1.0), Dual(b, 0.0)).d;
  db = va_fcn(eps0, Dual(a, 0.0), Dual(b, 1.0)).d;
}

}  // namespace

// ---------------------------------------------------------------------------
// 2. per edge
// ---------------------------------------------------------------------------
template <bool SECOND, bool WANTX>
__device__ __forceinline__ void adj_edge(const DevTables &t, const double *__restrict__ V,
                                         const double *__restrict__ grad, const double *__restrict__ lam,
                                         const FluxConsts &fc, double *__restrict__ EV, double *__restrict__ EX) {
  const long long l = (long long)blockIdx.x * kAdjBlock + threadIdx.x;
  if (l >= t.E) return;
  const int2 ij = t.eij[l];
  const int i = ij.x, j = ij.y;
  double Vi[5], Vj[5], g[5];
  const double ici = t.invcv[i], icj = t.invcv[j];
#pragma unroll
  for (int k = 0; k < 5; ++k) {
    Vi[k] = V[5 * (size_t)i + k];
    Vj[k] = V[5 * (size_t)j + k];
    g[k] = lam[5 * (size_t)i + k] * ici - lam[5 * (size_t)j + k] * icj;
  }
  double dx[3] = {0.0, 0.0, 0.0}, gi[15], gj[15];
  double VL[5], VR[5];
  if (SECOND) {
    const double4 xi = *reinterpret_cast<const double4 *>(t.X4 + 4 * (size_t)i);
    const double4 xj = *reinterpret_cast<const double4 *>(t.X4 + 4 * (size_t)j);
    dx[0] = xj.x - xi.x;
    dx[1] = xj.y - xi.y;
    dx[2] = xj.z - xi.z;
#pragma unroll
    for (int q = 0; q < 15; ++q) {
      gi[q] = grad[16 * (size_t)i + q];
      gj[q] = grad[16 * (size_t)j + q];
    }
#pragma unroll
    for (int k = 0; k < 5; ++k) {
      const double ddVij = dx[0] * gi[k] + dx[1] * gi[5 + k] + dx[2] * gi[10 + k];
      const double ddVji = dx[0] * gj[k] + dx[1] * gj[5 + k] + dx[2] * gj[10 + k];
      vanalbada(fc.beta, Vi[k], ddVij, Vj[k], ddVji, VL[k], VR[k]);
    }
  } else {
#pragma unroll
    for (int k = 0; k < 5; ++k) {
      VL[k] = Vi[k];
      VR[k] = Vj[k];
    }
  }
  const double nv[3] = {t.normal[3 * (size_t)l], t.normal[3 * (size_t)l + 1], t.normal[3 * (size_t)l + 2]};
  double fb[13];
#pragma unroll
  for (int d = 0; d < 13; ++d) fb[d] = 0.0;
  if (!(sqrt(nv[0] * nv[0] + nv[1] * nv[1] + nv[2] * nv[2]) < 1e-18))
    roeflux7_vjp(fc.gas, nv, VL, VR, g, fb, fb + 5, fb + 10);
  double vbi[5], vbj[5], ci[5] = {0.0, 0.0, 0.0, 0.0, 0.0}, cj[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
  double dxb[3] = {0.0, 0.0, 0.0};
  if (SECOND) {
    const double beta = fc.beta, beta2 = 1.0 - 2.0 * beta;
#pragma unroll
    for (int k = 0; k < 5; ++k) {
      const double ddVij = dx[0] * gi[k] + dx[1] * gi[5 + k] + dx[2] * gi[10 + k];
      const double ddVji = dx[0] * gj[k] + dx[1] * gj[5 + k] + dx[2] * gj[10 + k];
      const double dVji = Vj[k] - Vi[k];
      const double dV = beta2 * dVji;
      const double dVi = dV + 2.0 * beta * ddVij;
      const double dVj = dV + 2.0 * beta * ddVji;
      double da_i, db_i, da_j, db_j;
      va_partials(dVi, dVji, da_i, db_i);
      va_partials(dVj, dVji, da_j, db_j);
      const double Lb = fb[k], Rb = fb[5 + k];
      const double A = 0.5 * Lb * da_i, B = 0.5 * Lb * db_i;
      const double A2 = -0.5 * Rb * da_j, B2 = -0.5 * Rb * db_j;
      const double dVjib = B + B2 + beta2 * (A + A2);
      const double ddVijb = 2.0 * beta * A, ddVjib = 2.0 * beta * A2;
      vbi[k] = Lb - dVjib;
      vbj[k] = Rb + dVjib;
      // Gbar_i[5a+k] = ddVijb dx[a]: the node pass forms the products
      ci[k] = ddVijb;
      cj[k] = ddVjib;
      if (WANTX) {
#pragma unroll
        for (int a = 0; a < 3; ++a) dxb[a] += ddVijb * gi[5 * a + k] + ddVjib * gj[5 * a + k];
      }
    }
  } else {
#pragma unroll
    for (int k = 0; k < 5; ++k) {
      vbi[k] = fb[k];
      vbj[k] = fb[5 + k];
    }
  }
  double2 *ev = reinterpret_cast<double2 *>(EV + 20 * (size_t)l);
  ev[0] = make_double2(vbi[0], vbi[1]);
  ev[1] = make_double2(vbi[2], vbi[3]);
  ev[2] = make_double2(vbi[4], ci[0]);
  ev[3] = make_double2(ci[1], ci[2]);
  ev[4] = make_double2(ci[3], ci[4]);
  ev[5] = make_double2(vbj[0], vbj[1]);
  ev[6] = make_double2(vbj[2], vbj[3]);
  ev[7] = make_double2(vbj[4], cj[0]);
  ev[8] = make_double2(cj[1], cj[2]);
  ev[9] = make_double2(cj[3], cj[4]);
  if (WANTX) {
    double *ex = EX + 6 * (size_t)l;
#pragma unroll
    for (int a = 0; a < 3; ++a) {
      ex[a] = dxb[a];
      ex[3 + a] = fb[10 + a];
    }
  }
}

// The (dR/dU)^T-only edge pass fits two waves per SIMD without spilling
// (256 VGPRs; 2.10 -> 1.83 ms at 815k nodes); with the coordinate slots it
// would spill 68 registers and is left at the compiler's choice.
template <bool SECOND>
__global__ __launch_bounds__(kAdjBlock) __attribute__((amdgpu_waves_per_eu(2, 2))) void k_adj_edges_u(
    DevTables t, const double *__restrict__ V, const double *__restrict__ grad, const double *__restrict__ lam,
    FluxConsts fc, double *__restrict__ EV, double *__restrict__ EX) {
  adj_edge<SECOND, false>(t, V, grad, lam, fc, EV, EX);
}
template <bool SECOND>
__global__ __launch_bounds__(kAdjBlock) void k_adj_edges_x(DevTables t, const double *__restrict__ V,
                                                           const double *__restrict__ grad,
                                                           const double *__restrict__ lam, FluxConsts fc,
                                                           double *__restrict__ EV, double *__restrict__ EX) {
  adj_edge<SECOND, true>(t, V, grad, lam, fc, EV, EX);
}

// ---------------------------------------------------------------------------
// 3. per node: edge slots in edge order, then faces
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(kAdjBlock) void k_adj_nodes(DevTables t, const double *__restrict__ V,
                                                         const double *__restrict__ lam, FluxConsts fc,
                                                         const double *__restrict__ EV,
                                                         const double *__restrict__ EX, int second, int want_x,
                                                         double *__restrict__ Vb, double *__restrict__ Gb,
                                                         double *__restrict__ Xb) {
  const int n = blockIdx.x * kAdjBlock + threadIdx.x;
  if (n >= t.N) return;
  double vb[5] = {0.0, 0.0, 0.0, 0.0, 0.0}, gb[15], xb[3] = {0.0, 0.0, 0.0};
#pragma unroll
  for (int q = 0; q < 15; ++q) gb[q] = 0.0;
  const double4 xn = *reinterpret_cast<const double4 *>(t.X4 + 4 * (size_t)n);
  // one edge slot: Vbar, and Gbar = c dx^T with dx = X_j - X_i as the edge
  // pass formed it (the same products, added in the same order)
  auto add_slot = [&](const double *ev, int m, bool upper) {
    double c[5];
#pragma unroll
    for (int k = 0; k < 5; ++k) {
      vb[k] += ev[k];
      c[k] = ev[5 + k];
    }
    if (second) {
      const double4 xm = *reinterpret_cast<const double4 *>(t.X4 + 4 * (size_t)m);
      const double dx[3] = {upper ? xm.x - xn.x : xn.x - xm.x, upper ? xm.y - xn.y : xn.y - xm.y,
                            upper ? xm.z - xn.z : xn.z - xm.z};
#pragma unroll
      for (int a = 0; a < 3; ++a)
#pragma unroll
        for (int k = 0; k < 5; ++k) gb[5 * a + k] += c[k] * dx[a];
    }
  };
  // edges (m, n), m < n: the j-side slots
  for (int q = t.lrow[n]; q < t.lrow[n + 1]; ++q) {
    const size_t l = t.ledge[q];
    add_slot(EV + 20 * l + 10, t.eij[l].x, false);
    if (want_x) {
#pragma unroll
      for (int a = 0; a < 3; ++a) xb[a] += EX[6 * l + a];
    }
  }
  // edges (n, m), n < m: the i-side slots
  for (int l = t.eup[n]; l < t.eup[n + 1]; ++l) {
    add_slot(EV + 20 * (size_t)l, t.eij[l].y, true);
    if (want_x) {
#pragma unroll
      for (int a = 0; a < 3; ++a) xb[a] -= EX[6 * (size_t)l + a];
    }
  }
  const int f0 = t.frow[n], f1 = t.frow[n + 1];
  if (f0 < f1) {
    const double ic = t.invcv[n];
    double mu[5], Vn[5];
#pragma unroll
    for (int k = 0; k < 5; ++k) {
      mu[k] = lam[5 * (size_t)n + k] * ic;
      Vn[k] = V[5 * (size_t)n + k];
    }
    for (int q = f0; q < f1; ++q) {
      const int face = t.fidx[q];
      const int code = t.fcode[face];
#pragma unroll 1
      for (int dir = 0; dir < 5; ++dir) {
        Dual nv[3], vv[5], fl[5];
#pragma unroll
        for (int a = 0; a < 3; ++a) nv[a] = Dual(t.fnorm3[3 * (size_t)face + a]);
#pragma unroll
        for (int k = 0; k < 5; ++k) vv[k] = Dual(Vn[k], seed(dir, k));
        face_flux(fc, code, nv, vv, fl);
        double s = 0.0;
#pragma unroll
        for (int k = 0; k < 5; ++k) s += mu[k] * fl[k].d;
        vb[dir] += s;
      }
    }
  }
#pragma unroll
  for (int k = 0; k < 5; ++k) Vb[5 * (size_t)n + k] = vb[k];
#pragma unroll
  for (int r = 0; r < 15; ++r) Gb[16 * (size_t)n + r] = gb[r];
#pragma unroll
  for (int a = 0; a < 3; ++a) Xb[4 * (size_t)n + a] = xb[a];
}

// ---------------------------------------------------------------------------
// 4. per node: reverse of the LSQ gradients, then U
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(kAdjBlock) void k_adj_lsq(DevTables t, const double *__restrict__ V,
                                                       const double *__restrict__ U, Gas gas,
                                                       const double *__restrict__ Vb,
                                                       const double *__restrict__ Gb, int second, int want_x,
                                                       double *__restrict__ Ub, double *__restrict__ Xb,
                                                       double *__restrict__ S) {
  const int n = blockIdx.x * kAdjBlock + threadIdx.x;
  if (n >= t.N) return;
  double vb[5], Vn[5];
#pragma unroll
  for (int k = 0; k < 5; ++k) {
    vb[k] = Vb[5 * (size_t)n + k];
    Vn[k] = V[5 * (size_t)n + k];
  }
  if (second) {
    double Ln[6], gn[15];
#pragma unroll
    for (int q = 0; q < 6; ++q) Ln[q] = t.lsq6[6 * (size_t)n + q];
#pragma unroll
    for (int q = 0; q < 15; ++q) gn[q] = Gb[16 * (size_t)n + q];
    const double4 xn = *reinterpret_cast<const double4 *>(t.X4 + 4 * (size_t)n);
    double xacc[3] = {0.0, 0.0, 0.0}, Lb[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
    double M[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
    for (int e = t.rowptr[n]; e < t.rowptr[n + 1]; ++e) {
      const int m = t.adj[e];
      const double4 xm = *reinterpret_cast<const double4 *>(t.X4 + 4 * (size_t)m);
      const double dx0 = xm.x - xn.x, dx1 = xm.y - xn.y, dx2 = xm.z - xn.z;
      double Lm[6], gm[15], Vm[5];
#pragma unroll
      for (int q = 0; q < 6; ++q) Lm[q] = t.lsq6[6 * (size_t)m + q];
#pragma unroll
      for (int q = 0; q < 15; ++q) gm[q] = Gb[16 * (size_t)m + q];
#pragma unroll
      for (int k = 0; k < 5; ++k) Vm[k] = V[5 * (size_t)m + k];
      double wn[3], wm[3];
      lsq_weights(Ln, dx0, dx1, dx2, wn);      // row n, neighbour m
      lsq_weights(Lm, -dx0, -dx1, -dx2, wm);   // row m, neighbour n (dx_mn = X_n - X_m)
#pragma unroll
      for (int k = 0; k < 5; ++k)
        vb[k] += (wm[0] * gm[k] + wm[1] * gm[5 + k] + wm[2] * gm[10 + k]) -
                 (wn[0] * gn[k] + wn[1] * gn[5 + k] + wn[2] * gn[10 + k]);
      if (want_x) {
        double wbn[3], wbm[3];
#pragma unroll
        for (int a = 0; a < 3; ++a) {
          double s1 = 0.0, s2 = 0.0;
#pragma unroll
          for (int k = 0; k < 5; ++k) {
            s1 += gn[5 * a + k] * (Vm[k] - Vn[k]);
            s2 += gm[5 * a + k] * (Vn[k] - Vm[k]);
          }
          wbn[a] = s1;
          wbm[a] = s2;
        }
        double dxn[3] = {0.0, 0.0, 0.0}, dxm[3] = {0.0, 0.0, 0.0};
        lsq_weights_rev(Ln, dx0, dx1, dx2, wbn, dxn, Lb);
        lsq_weights_rev(Lm, -dx0, -dx1, -dx2, wbm, dxm, nullptr);
#pragma unroll
        for (int a = 0; a < 3; ++a) xacc[a] += dxm[a] - dxn[a];
        // LSQ moments in edge order (rows are in ascending edge id)
        M[0] += dx0 * dx0;
        M[1] += dx0 * dx1;
        M[2] += dx0 * dx2;
        M[3] += dx1 * dx1;
        M[4] += dx1 * dx2;
        M[5] += dx2 * dx2;
      }
    }
    if (want_x) {
      double Mb[6];
#pragma unroll 1
      for (int dir = 0; dir < 6; ++dir) {
        Dual Md[6], Ld[6];
#pragma unroll
        for (int q = 0; q < 6; ++q) Md[q] = Dual(M[q], seed(dir, q));
        lsq_params(Md, Ld);
        double s = 0.0;
#pragma unroll
        for (int q = 0; q < 6; ++q) s += Lb[q] * Ld[q].d;
        Mb[dir] = s;
      }
      double *sn = S + 6 * (size_t)n;
      sn[0] = 2.0 * Mb[0];
      sn[1] = Mb[1];
      sn[2] = Mb[2];
      sn[3] = 2.0 * Mb[3];
      sn[4] = Mb[4];
      sn[5] = 2.0 * Mb[5];
#pragma unroll
      for (int a = 0; a < 3; ++a) Xb[4 * (size_t)n + a] += xacc[a];
    }
  }
  if (Ub) {
    double u[5], ub[5];
#pragma unroll
    for (int k = 0; k < 5; ++k) u[k] = U[5 * (size_t)n + k];
#pragma unroll 1
    for (int dir = 0; dir < 5; ++dir) {
      Dual ud[5], vd[5];
#pragma unroll
      for (int k = 0; k < 5; ++k) ud[k] = Dual(u[k], seed(dir, k));
      prim_from_cons(ud, gas, vd);
      double s = 0.0;
#pragma unroll
      for (int k = 0; k < 5; ++k) s += vb[k] * vd[k].d;
      ub[dir] = s;
    }
#pragma unroll
    for (int k = 0; k < 5; ++k) Ub[5 * (size_t)n + k] = ub[k];
  }
}

// ---------------------------------------------------------------------------
// 5. geometry: per tet (control volumes, dual normals), per boundary face
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(kAdjBlock) void k_adj_tets(int T, const int *__restrict__ tets,
                                                        const int *__restrict__ tedge,
                                                        const double *__restrict__ X4,
                                                        const double *__restrict__ invcv,
                                                        const double *__restrict__ lam,
                                                        const double *__restrict__ R,
                                                        const double *__restrict__ EX, double *__restrict__ TX) {
  const long long t = (long long)blockIdx.x * kAdjBlock + threadIdx.x;
  if (t >= T) return;
  int nd[4];
  double x[4][3], cvb = 0.0;
#pragma unroll
  for (int v = 0; v < 4; ++v) {
    nd[v] = tets[4 * t + v];
    const double4 xv = *reinterpret_cast<const double4 *>(X4 + 4 * (size_t)nd[v]);
    x[v][0] = xv.x;
    x[v][1] = xv.y;
    x[v][2] = xv.z;
    // R_i = F_i / cv_i: d/dcv_i (lambda_i . R_i) = -lambda_i . R_i / cv_i
    double s = 0.0;
#pragma unroll
    for (int k = 0; k < 5; ++k) s += lam[5 * (size_t)nd[v] + k] * R[5 * (size_t)nd[v] + k];
    cvb += -s * invcv[nd[v]];
  }
  double nb[6][3], sg[6];
#pragma unroll
  for (int e = 0; e < 6; ++e) {
    const size_t l = tedge[6 * t + e];
    sg[e] = nd[cEdgeEnd[e][0]] < nd[cEdgeEnd[e][1]] ? 1.0 : -1.0;
#pragma unroll
    for (int a = 0; a < 3; ++a) nb[e][a] = EX[6 * l + 3 + a];
  }
  const double third = 1.0 / 3.0, sixth = 1.0 / 6.0;
#pragma unroll 1
  for (int dir = 0; dir < 12; ++dir) {
    Dual xd[4][3];
#pragma unroll
    for (int v = 0; v < 4; ++v)
#pragma unroll
      for (int a = 0; a < 3; ++a) xd[v][a] = Dual(x[v][a], seed(dir, 3 * v + a));
    // ElemTet::computeVolume (ElemTetCore.C:31-44)
    Dual v1[3], v2[3], v3[3];
#pragma unroll
    for (int a = 0; a < 3; ++a) {
      v1[a] = xd[1][a] - xd[0][a];
      v2[a] = xd[2][a] - xd[0][a];
      v3[a] = xd[3][a] - xd[0][a];
    }
    Dual c0 = v1[1] * v2[2] - v1[2] * v2[1];
    Dual c1 = v1[2] * v2[0] - v1[0] * v2[2];
    Dual c2 = v1[0] * v2[1] - v1[1] * v2[0];
    Dual vol = sixth * (v3[0] * c0 + v3[1] * c1 + v3[2] * c2);
    double acc = 0.25 * vol.d * cvb;
    // dual-face normals (ElemTetCore.C:281-330)
    Dual f[4][3], gc[3];
#pragma unroll
    for (int a = 0; a < 3; ++a) {
      f[0][a] = third * ((xd[0][a] + xd[1][a]) + xd[2][a]);
      f[1][a] = third * ((xd[0][a] + xd[1][a]) + xd[3][a]);
      f[2][a] = third * ((xd[0][a] + xd[2][a]) + xd[3][a]);
      f[3][a] = third * ((xd[1][a] + xd[2][a]) + xd[3][a]);
      gc[a] = 0.25 * (((xd[0][a] + xd[1][a]) + xd[2][a]) + xd[3][a]);
    }
#pragma unroll
    for (int e = 0; e < 6; ++e) {
      Dual p[3], q[3];
#pragma unroll
      for (int a = 0; a < 3; ++a) {
        p[a] = f[cEdgeFace[e][1]][a] - gc[a];
        q[a] = f[cEdgeFace[e][0]][a] - gc[a];
      }
      Dual A[3] = {p[1] * q[2] - p[2] * q[1], p[2] * q[0] - p[0] * q[2], p[0] * q[1] - p[1] * q[0]};
      double s = 0.0;
#pragma unroll
      for (int a = 0; a < 3; ++a) {
        Dual nn = 0.5 * (A[a] + A[a]) + 0.25 * (A[a] + A[a]);
        s += nb[e][a] * nn.d;
      }
      acc += sg[e] * s;
    }
    TX[12 * t + dir] = acc;
  }
}

__global__ __launch_bounds__(kAdjBlock) void k_adj_faces(int F, const int *__restrict__ faces,
                                                         const int *__restrict__ fcode,
                                                         const double *__restrict__ fnorm3,
                                                         const double *__restrict__ X4,
                                                         const double *__restrict__ invcv,
                                                         const double *__restrict__ V,
                                                         const double *__restrict__ lam, FluxConsts fc,
                                                         double *__restrict__ FX) {
  const int f = blockIdx.x * kAdjBlock + threadIdx.x;
  if (f >= F) return;
  const int code = fcode[f];
  double nb[3] = {0.0, 0.0, 0.0};
  int nd[3];
#pragma unroll
  for (int v = 0; v < 3; ++v) nd[v] = faces[3 * (size_t)f + v];
#pragma unroll 1
  for (int v = 0; v < 3; ++v) {
    const int n = nd[v];
    const double ic = invcv[n];
    double mu[5];
    Dual vv[5];
#pragma unroll
    for (int k = 0; k < 5; ++k) {
      mu[k] = lam[5 * (size_t)n + k] * ic;
      vv[k] = Dual(V[5 * (size_t)n + k]);
    }
#pragma unroll 1
    for (int dir = 0; dir < 3; ++dir) {
      Dual nv[3], fl[5];
#pragma unroll
      for (int a = 0; a < 3; ++a) nv[a] = Dual(fnorm3[3 * (size_t)f + a], seed(dir, a));
      face_flux(fc, code, nv, vv, fl);
      double s = 0.0;
#pragma unroll
      for (int k = 0; k < 5; ++k) s += mu[k] * fl[k].d;
      nb[dir] += s;
    }
  }
  // nv = (1/3) n_f (FaceTriaCore.C:288-292)
  const double third = 1.0 / 3.0;
#pragma unroll
  for (int a = 0; a < 3; ++a) nb[a] *= third;
  double x[3][3];
#pragma unroll
  for (int v = 0; v < 3; ++v) {
    const double4 xv = *reinterpret_cast<const double4 *>(X4 + 4 * (size_t)nd[v]);
    x[v][0] = xv.x;
    x[v][1] = xv.y;
    x[v][2] = xv.z;
  }
#pragma unroll 1
  for (int dir = 0; dir < 9; ++dir) {
    Dual xd[3][3];
#pragma unroll
    for (int v = 0; v < 3; ++v)
#pragma unroll
      for (int a = 0; a < 3; ++a) xd[v][a] = Dual(x[v][a], seed(dir, 3 * v + a));
    // FaceTria::computeNormal (FaceTriaCore.C:31-37): 0.5 (x2 - x0) x (x1 - x0)
    Dual p[3], q[3];
#pragma unroll
    for (int a = 0; a < 3; ++a) {
      p[a] = xd[2][a] - xd[0][a];
      q[a] = xd[1][a] - xd[0][a];
    }
    Dual c0 = p[1] * q[2] - p[2] * q[1];
    Dual c1 = p[2] * q[0] - p[0] * q[2];
    Dual c2 = p[0] * q[1] - p[1] * q[0];
    FX[9 * (size_t)f + dir] = nb[0] * (0.5 * c0).d + nb[1] * (0.5 * c1).d + nb[2] * (0.5 * c2).d;
  }
}

// ---------------------------------------------------------------------------
// 6. per node: Xbar
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(kAdjBlock) void k_adj_xgather(DevTables t, const double *__restrict__ Xb,
                                                           const double *__restrict__ S, int second,
                                                           const int *__restrict__ ntrow,
                                                           const int *__restrict__ ntslot,
                                                           const double *__restrict__ TX,
                                                           const int *__restrict__ nfslot,
                                                           const double *__restrict__ FX,
                                                           double *__restrict__ out) {
  // XCD-aware block order (each XCD a contiguous node range): 0.1 ms less at
  // 815k nodes; the same order made the node-sum and LSQ reverse passes slower
  const int nb = (t.N + kAdjBlock - 1) / kAdjBlock;
  const int n = xcd_chunk(blockIdx.x, nb) * kAdjBlock + (int)threadIdx.x;
  if (n >= t.N) return;
  double xb[3] = {Xb[4 * (size_t)n], Xb[4 * (size_t)n + 1], Xb[4 * (size_t)n + 2]};
  if (second) {
    const double *sn = S + 6 * (size_t)n;
    const double4 xn = *reinterpret_cast<const double4 *>(t.X4 + 4 * (size_t)n);
    for (int e = t.rowptr[n]; e < t.rowptr[n + 1]; ++e) {
      const int m = t.adj[e];
      const double *sm = S + 6 * (size_t)m;
      const double4 xm = *reinterpret_cast<const double4 *>(t.X4 + 4 * (size_t)m);
      const double d0 = xm.x - xn.x, d1 = xm.y - xn.y, d2 = xm.z - xn.z;
      const double a0 = sn[0] + sm[0], a1 = sn[1] + sm[1], a2 = sn[2] + sm[2];
      const double a3 = sn[3] + sm[3], a4 = sn[4] + sm[4], a5 = sn[5] + sm[5];
      xb[0] -= a0 * d0 + a1 * d1 + a2 * d2;
      xb[1] -= a1 * d0 + a3 * d1 + a4 * d2;
      xb[2] -= a2 * d0 + a4 * d1 + a5 * d2;
    }
  }
  for (int q = ntrow[n]; q < ntrow[n + 1]; ++q) {
    const size_t s = ntslot[q];  // 4 t + v
    const size_t tt = s >> 2, v = s & 3;
#pragma unroll
    for (int a = 0; a < 3; ++a) xb[a] += TX[12 * tt + 3 * v + a];
  }
  for (int q = t.frow[n]; q < t.frow[n + 1]; ++q) {
    const size_t s = nfslot[q];  // 3 f + v
    const size_t f = s / 3, v = s - 3 * f;
#pragma unroll
    for (int a = 0; a < 3; ++a) xb[a] += FX[9 * f + 3 * v + a];
  }
#pragma unroll
  for (int a = 0; a < 3; ++a) out[3 * (size_t)n + a] = xb[a];
}

// ---------------------------------------------------------------------------
// host side
// ---------------------------------------------------------------------------
void free_adjoint(aerofx_ctx *c) {
  if (c->adj && c->adj->h_J) (void)hipHostFree(c->adj->h_J);
  delete c->adj;  // device memory is owned by the context allocation list
  c->adj = nullptr;
}

namespace {

template <class T>
int alloc_n(aerofx_ctx *c, size_t n, T **out) {
  void *p = nullptr;
  int rc = ctx_alloc(c, n * sizeof(T), &p);
  *out = static_cast<T *>(p);
  return rc;
}

template <class T>
int upload_vec(aerofx_ctx *c, const std::vector<T> &v, T **out) {
  int rc = alloc_n(c, v.size() ? v.size() : 1, out);
  if (rc) return rc;
  if (v.empty()) return 0;
  hipError_t e = hipMemcpy(*out, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice);
  return e == hipSuccess ? 0 : set_hip_error(e, "adjoint table upload");
}

int ensure_adjoint(aerofx_ctx *c) {
  if (c->adj) return 0;
  const int N = c->N, E = c->E, T = c->T, F = c->F;
  int rc = ensure_elem_tables(c);
  if (rc) return rc;
  Adjoint *a = new Adjoint();
  const ElemTables &el = *c->elem;
  a->tets = el.tets;
  a->tedge = el.tedge;
  a->faces = el.faces;
  a->ntrow = el.ntrow;
  a->ntslot = el.ntslot;
  a->nfslot = el.nfslot;
  if ((rc = alloc_n(c, 20 * (size_t)E, &a->EV)) ||
      (rc = alloc_n(c, 6 * (size_t)E, &a->EX)) || (rc = alloc_n(c, 5 * (size_t)N, &a->Vb)) ||
      (rc = alloc_n(c, 16 * (size_t)N, &a->Gb)) || (rc = alloc_n(c, 4 * (size_t)N, &a->Xb)) ||
      (rc = alloc_n(c, 6 * (size_t)N, &a->S)) || (rc = alloc_n(c, 5 * (size_t)N, &a->R)) ||
      (rc = alloc_n(c, 12 * (size_t)std::max(T, 1), &a->TX)) ||
      (rc = alloc_n(c, 9 * (size_t)std::max(F, 1), &a->FX)) || (rc = alloc_n(c, (size_t)F + 1, &a->Jf))) {
    delete a;
    return rc;
  }
  hipError_t e = hipHostMalloc((void **)&a->h_J, sizeof(double), hipHostMallocDefault);
  if (e != hipSuccess) {
    delete a;
    return set_hip_error(e, "adjoint set-up");
  }
  c->adj = a;
  return 0;
}

inline dim3 grid_for(long long n) { return dim3((unsigned)((n + kAdjBlock - 1) / kAdjBlock)); }

}  // namespace

int ensure_elem_tables(aerofx_ctx *c) {
  if (c->elem) return 0;
  const int N = c->N, E = c->E, T = c->T, F = c->F;
  if ((int)c->tets.size() != 4 * T || (int)c->tet_edges.size() != 6 * T || (int)c->faces.size() != 3 * F)
    return set_error(AEROFX_ERR_STATE, "element tables missing from the context");
  // node -> (tet, vertex) slots in tet order; node -> (face, vertex) in the
  // order of tab.fidx; edge -> (tet, local edge) slots in tet order
  std::vector<int> ntrow(N + 1, 0), ntslot(4 * (size_t)T), nfslot(3 * (size_t)F);
  std::vector<int> etrow(E + 1, 0), etslot(6 * (size_t)T);
  for (size_t k = 0; k < 4 * (size_t)T; ++k) ntrow[c->tets[k] + 1]++;
  for (int n = 0; n < N; ++n) ntrow[n + 1] += ntrow[n];
  {
    std::vector<int> fill(ntrow.begin(), ntrow.end() - 1);
    for (size_t k = 0; k < 4 * (size_t)T; ++k) ntslot[fill[c->tets[k]]++] = (int)k;
  }
  {
    std::vector<int> frow(N + 1, 0);
    for (size_t k = 0; k < 3 * (size_t)F; ++k) frow[c->faces[k] + 1]++;
    for (int n = 0; n < N; ++n) frow[n + 1] += frow[n];
    std::vector<int> fill(frow.begin(), frow.end() - 1);
    for (size_t k = 0; k < 3 * (size_t)F; ++k) nfslot[fill[c->faces[k]]++] = (int)k;
  }
  for (size_t k = 0; k < 6 * (size_t)T; ++k) etrow[c->tet_edges[k] + 1]++;
  for (int l = 0; l < E; ++l) etrow[l + 1] += etrow[l];
  {
    std::vector<int> fill(etrow.begin(), etrow.end() - 1);
    for (size_t k = 0; k < 6 * (size_t)T; ++k) etslot[fill[c->tet_edges[k]]++] = (int)k;
  }
  ElemTables *el = new ElemTables();
  int rc;
  if ((rc = upload_vec(c, c->tets, &el->tets)) || (rc = upload_vec(c, c->tet_edges, &el->tedge)) ||
      (rc = upload_vec(c, c->faces, &el->faces)) || (rc = upload_vec(c, ntrow, &el->ntrow)) ||
      (rc = upload_vec(c, ntslot, &el->ntslot)) || (rc = upload_vec(c, nfslot, &el->nfslot)) ||
      (rc = upload_vec(c, etrow, &el->etrow)) || (rc = upload_vec(c, etslot, &el->etslot))) {
    delete el;
    return rc;
  }
  c->elem = el;
  return 0;
}

int launch_adjoint(aerofx_ctx *c, const double *U, const double *lam, double *Ub, double *Xout) {
  int rc = ensure_adjoint(c);
  if (rc) return rc;
  Adjoint *a = c->adj;
  const DevTables &t = c->tab;
  const int second = c->p.second_order ? 1 : 0;
  const int want_x = Xout ? 1 : 0;
  // 1. forward pass: V, gradients (recomputed), R
  if ((rc = launch_residual(c, U, a->R, true))) return rc;
  // 2.
  if (t.E > 0) {
    const dim3 g = grid_for(t.E), b(kAdjBlock);
    if (second && want_x)
      hipLaunchKernelGGL((k_adj_edges_x<true>), g, b, 0, c->stream, t, c->d_V, c->d_grad, lam, c->fc, a->EV, a->EX);
    else if (second)
      hipLaunchKernelGGL((k_adj_edges_u<true>), g, b, 0, c->stream, t, c->d_V, c->d_grad, lam, c->fc, a->EV, a->EX);
    else if (want_x)
      hipLaunchKernelGGL((k_adj_edges_x<false>), g, b, 0, c->stream, t, c->d_V, c->d_grad, lam, c->fc, a->EV, a->EX);
    else
      hipLaunchKernelGGL((k_adj_edges_u<false>), g, b, 0, c->stream, t, c->d_V, c->d_grad, lam, c->fc, a->EV, a->EX);
  }
  // 3.
  hipLaunchKernelGGL(k_adj_nodes, grid_for(t.N), dim3(kAdjBlock), 0, c->stream, t, c->d_V, lam, c->fc, a->EV, a->EX,
                     second, want_x, a->Vb, a->Gb, a->Xb);
  // 4.
  hipLaunchKernelGGL(k_adj_lsq, grid_for(t.N), dim3(kAdjBlock), 0, c->stream, t, c->d_V, U, c->gas, a->Vb, a->Gb,
                     second, want_x, Ub, a->Xb, a->S);
  // the Navier-Stokes Galerkin term (its element derivative, transposed)
  if (c->viscous && Ub && (rc = launch_visc_adjoint(c, lam, Ub))) return rc;
  if (want_x) {
    // 5.
    if (c->T > 0)
      hipLaunchKernelGGL(k_adj_tets, grid_for(c->T), dim3(kAdjBlock), 0, c->stream, c->T, a->tets, a->tedge, t.X4,
                         t.invcv, lam, a->R, a->EX, a->TX);
    // the Navier-Stokes Galerkin term's coordinate derivative, per element
    if (c->T > 0 && (rc = launch_visc_adjoint_x(c, lam, a->TX))) return rc;
    if (c->F > 0)
      hipLaunchKernelGGL(k_adj_faces, grid_for(c->F), dim3(kAdjBlock), 0, c->stream, c->F, a->faces, t.fcode,
                         t.fnorm3, t.X4, t.invcv, c->d_V, lam, c->fc, a->FX);
    // 6.
    hipLaunchKernelGGL(k_adj_xgather, grid_for(t.N), dim3(kAdjBlock), 0, c->stream, t, a->Xb, a->S, second,
                       a->ntrow, a->ntslot, a->TX, a->nfslot, a->FX, Xout);
  }
  hipError_t e = hipGetLastError();
  return e == hipSuccess ? 0 : set_hip_error(e, "adjoint launch");
}

// ---------------------------------------------------------------------------
// force objective on slip-wall faces: J = sum_f sum_{v in f} p_v (n_f / 3) . d
// (the wall flux carries p n, FluxFcnDescSG.C:779-789, so this is the
// pressure force on the wall projected on d)
// ---------------------------------------------------------------------------
// per face: J_f, and the coordinate derivative of J_f (9 slots)
__global__ __launch_bounds__(kAdjBlock) void k_lift_faces(int F, const int *__restrict__ faces,
                                                          const int *__restrict__ fcode,
                                                          const double *__restrict__ fnorm3,
                                                          const double *__restrict__ X4,
                                                          const double *__restrict__ V, double d0, double d1,
                                                          double d2, double *__restrict__ Jf,
                                                          double *__restrict__ FX) {
  const int f = blockIdx.x * kAdjBlock + threadIdx.x;
  if (f >= F) return;
  const int code = fcode[f];
  const bool wall = is_wall_code(code);
  int nd[3];
  double psum = 0.0;
#pragma unroll
  for (int v = 0; v < 3; ++v) {
    nd[v] = faces[3 * (size_t)f + v];
    psum += V[5 * (size_t)nd[v] + 4];
  }
  if (!wall) psum = 0.0;
  const double *nv = fnorm3 + 3 * (size_t)f;
  Jf[f] = psum * (nv[0] * d0 + nv[1] * d1 + nv[2] * d2);
  if (!FX) return;
  // d J_f / d x: J_f = psum (1/3) (0.5 (x2 - x0) x (x1 - x0)) . d
  double x[3][3];
#pragma unroll
  for (int v = 0; v < 3; ++v) {
    const double4 xv = *reinterpret_cast<const double4 *>(X4 + 4 * (size_t)nd[v]);
    x[v][0] = xv.x;
    x[v][1] = xv.y;
    x[v][2] = xv.z;
  }
  const double third = 1.0 / 3.0;
#pragma unroll 1
  for (int dir = 0; dir < 9; ++dir) {
    Dual p[3], q[3];
#pragma unroll
    for (int a = 0; a < 3; ++a) {
      p[a] = Dual(x[2][a], seed(dir, 6 + a)) - Dual(x[0][a], seed(dir, a));
      q[a] = Dual(x[1][a], seed(dir, 3 + a)) - Dual(x[0][a], seed(dir, a));
    }
    Dual c0 = p[1] * q[2] - p[2] * q[1];
    Dual c1 = p[2] * q[0] - p[0] * q[2];
    Dual c2 = p[0] * q[1] - p[1] * q[0];
    FX[9 * (size_t)f + dir] = psum * third * 0.5 * (c0.d * d0 + c1.d * d1 + c2.d * d2);
  }
}

// per node: dJ/dU (through p(U)) and dJ/dX (face slots)
__global__ __launch_bounds__(kAdjBlock) void k_lift_nodes(DevTables t, const double *__restrict__ U, Gas gas,
                                                          const int *__restrict__ nfslot,
                                                          const double *__restrict__ FX, double d0, double d1,
                                                          double d2, double *__restrict__ dJdU,
                                                          double *__restrict__ dJdX) {
  const int n = blockIdx.x * kAdjBlock + threadIdx.x;
  if (n >= t.N) return;
  double w = 0.0, xb[3] = {0.0, 0.0, 0.0};
  for (int q = t.frow[n]; q < t.frow[n + 1]; ++q) {
    const int f = t.fidx[q];
    const int code = t.fcode[f];
    if (!is_wall_code(code)) continue;
    const double *nv = t.fnorm3 + 3 * (size_t)f;
    w += nv[0] * d0 + nv[1] * d1 + nv[2] * d2;
    if (dJdX) {
      const size_t sl = nfslot[q];
      const size_t ff = sl / 3, v = sl - 3 * ff;
#pragma unroll
      for (int a = 0; a < 3; ++a) xb[a] += FX[9 * ff + 3 * v + a];
    }
  }
  if (dJdU) {
    double u[5];
#pragma unroll
    for (int k = 0; k < 5; ++k) u[k] = U[5 * (size_t)n + k];
#pragma unroll 1
    for (int dir = 0; dir < 5; ++dir) {
      Dual ud[5], vd[5];
#pragma unroll
      for (int k = 0; k < 5; ++k) ud[k] = Dual(u[k], seed(dir, k));
      prim_from_cons(ud, gas, vd);
      dJdU[5 * (size_t)n + dir] = w * vd[4].d;
    }
  }
  if (dJdX)
#pragma unroll
    for (int a = 0; a < 3; ++a) dJdX[3 * (size_t)n + a] = xb[a];
}

// J = sum of the per-face values in face order (one block)
__global__ void k_sum_ordered(int F, const double *__restrict__ Jf, double *__restrict__ out) {
  __shared__ double sh[256];
  const int per = (F + 255) / 256;
  const int b0 = threadIdx.x * per, b1 = min(F, b0 + per);
  double s = 0.0;
  for (int f = b0; f < b1; ++f) s += Jf[f];
  sh[threadIdx.x] = s;
  __syncthreads();
  if (threadIdx.x == 0) {
    double tot = 0.0;
    for (int k = 0; k < 256; ++k) tot += sh[k];
    *out = tot;
  }
}

}  // namespace aerofx

using namespace aerofx;

extern "C" int aerofx_adjoint_residual(aerofx_ctx *ctx, const double *U, const double *lambda,
                                       double *dRdU_T_lambda, double *dRdX_T_lambda) {
  if (!ctx || !U || !lambda) return set_error(AEROFX_ERR_ARG, "adjoint_residual: null argument");
  if (ctx->fc.low_mach)
    return set_error(AEROFX_ERR_STATE, "adjoint_residual: not available with low-Mach preconditioning");
  if (int rf = refuse_failsafe_nodes(ctx, "adjoint_residual")) return rf;
  if (!dRdU_T_lambda && !dRdX_T_lambda) return AEROFX_OK;
  int rc = launch_adjoint(ctx, U, lambda, dRdU_T_lambda, dRdX_T_lambda);
  if (rc) return rc;
  hipError_t e = hipStreamSynchronize(ctx->stream);
  return e == hipSuccess ? AEROFX_OK : set_hip_error(e, "adjoint_residual");
}

extern "C" int aerofx_wall_force(aerofx_ctx *ctx, const double *U, const double dir[3], double *J,
                                 double *dJdU, double *dJdX) {
  if (!ctx || !U || !dir || !J) return set_error(AEROFX_ERR_ARG, "wall_force: null argument");
  int rc = ensure_adjoint(ctx);
  if (rc) return rc;
  Adjoint *a = ctx->adj;
  const DevTables &t = ctx->tab;
  if ((rc = launch_primitive(ctx, U, ctx->d_V))) return rc;
  if (ctx->F > 0) {
    hipLaunchKernelGGL(k_lift_faces, grid_for(ctx->F), dim3(kAdjBlock), 0, ctx->stream, ctx->F, a->faces, t.fcode,
                       t.fnorm3, t.X4, ctx->d_V, dir[0], dir[1], dir[2], a->Jf, dJdX ? a->FX : nullptr);
    // the viscous stress on the wall faces (PostFcnNS::computeForce)
    if ((rc = launch_visc_force(ctx, dir, a->Jf, dJdU || dJdX))) return rc;
    hipLaunchKernelGGL(k_sum_ordered, dim3(1), dim3(256), 0, ctx->stream, ctx->F, a->Jf, a->Jf + ctx->F);
  } else {
    hipLaunchKernelGGL(k_sum_ordered, dim3(1), dim3(256), 0, ctx->stream, 0, a->Jf, a->Jf);  // no faces: J = 0
  }
  if (dJdU || dJdX)
    hipLaunchKernelGGL(k_lift_nodes, grid_for(t.N), dim3(kAdjBlock), 0, ctx->stream, t, U, ctx->gas, a->nfslot,
                       a->FX, dir[0], dir[1], dir[2], dJdU, dJdX);
  if ((dJdU || dJdX) && (rc = launch_visc_force_gather(ctx, dJdU, dJdX))) return rc;
  hipError_t e = hipGetLastError();
  if (e == hipSuccess) e = hipMemcpyAsync(a->h_J, a->Jf + ctx->F, sizeof(double), hipMemcpyDeviceToHost, ctx->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
  if (e != hipSuccess) return set_hip_error(e, "wall_force");
  *J = *a->h_J;
  return AEROFX_OK;
}
