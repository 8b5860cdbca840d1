// device_jacobian.h -- flux Jacobians as gfx950 device code.
//
//   roejac6      ROEJAC6, type 0, prec 0 or 1 (f77src/roejac6.f:1-316):
//                0.5 |n| (A(Ug) +/- |A_roe(Ug,Ud)|), row-major 5x5
//   post_dVdU    VarFcnSGEuler::computedVdU + VarFcnBase::postMultiplyBydVdU
//                (VarFcnSGEuler.h:588-603, VarFcnBase.h:383-411)
//   wall_jac     FluxFcnSGWallEuler3D::computeJacobian (FluxFcnDescSG.C:838-886)
//   farfield_jac FluxFcnFD<5>::computeJacobian over BOUNDFLUX5
//                (FluxFcnBase.h:112-159)
//   invert5      LINPACKdgefa + LINPACKdgedi on the transposed block
//                (DenseMatrixOps.h:1006-1181, DiagMatrix.C:44-82)
#pragma once
#include "device_physics.h"

namespace aerofx {

// With prec, Turkel's beta of the edge (turkel_beta, the law of
// roejac6.f:135-143) replaces 1: the Jacobian of the low-Mach flux ROEFLUX5.
__device__ __forceinline__ void roejac6(double gamma_diss, const Gas &g, const double en[3], double evitno,
                                        const double Ug[5], const double Ud[5], int sw, double jac[25],
                                        bool prec = false, const LowMach &lm = LowMach{}, double length = 1.0) {
  const double gam = g.gam, gam1 = g.gam1, pstiff = g.pstiff;
  double H[25];
  double rnorm = sqrt(en[0] * en[0] + en[1] * en[1] + en[2] * en[2]);
  double invnorm = 1.0 / rnorm;
  double n1 = en[0] * invnorm, n2 = en[1] * invnorm, n3 = en[2] * invnorm;
  double vitno = evitno * invnorm;
  double VdotN = Ug[1] * n1 + Ug[2] * n2 + Ug[3] * n3;
  double qir = 0.5 * (Ug[1] * Ug[1] + Ug[2] * Ug[2] + Ug[3] * Ug[3]);
  double enth = gam * (Ug[4] + pstiff) + gam1 * Ug[0] * qir;
  enth = enth / (gam1 * Ug[0]);
  H[0] = -vitno; H[1] = n1; H[2] = n2; H[3] = n3; H[4] = 0.0;
  H[5] = gam1 * n1 * qir - Ug[1] * VdotN;
  H[6] = (1.0 - gam1) * n1 * Ug[1] + VdotN - vitno;
  H[7] = -gam1 * n1 * Ug[2] + Ug[1] * n2;
  H[8] = -gam1 * n1 * Ug[3] + Ug[1] * n3;
  H[9] = gam1 * n1;
  H[10] = gam1 * n2 * qir - Ug[2] * VdotN;
  H[11] = -gam1 * n2 * Ug[1] + Ug[2] * n1;
  H[12] = (1.0 - gam1) * n2 * Ug[2] + VdotN - vitno;
  H[13] = -gam1 * n2 * Ug[3] + Ug[2] * n3;
  H[14] = gam1 * n2;
  H[15] = gam1 * n3 * qir - Ug[3] * VdotN;
  H[16] = -gam1 * n3 * Ug[1] + Ug[3] * n1;
  H[17] = -gam1 * n3 * Ug[2] + Ug[3] * n2;
  H[18] = (1.0 - gam1) * n3 * Ug[3] + VdotN - vitno;
  H[19] = gam1 * n3;
  H[20] = (gam1 * qir - enth) * VdotN;
  H[21] = enth * n1 - gam1 * Ug[1] * VdotN;
  H[22] = enth * n2 - gam1 * Ug[2] * VdotN;
  H[23] = enth * n3 - gam1 * Ug[3] * VdotN;
  H[24] = (1.0 + gam1) * VdotN - vitno;
  double squsr1 = sqrt(Ug[0]);
  double squsr2 = sqrt(Ud[0]);
  double ener1 = (Ug[4] + gam * pstiff) / gam1 + 0.5 * Ug[0] * (Ug[1] * Ug[1] + Ug[2] * Ug[2] + Ug[3] * Ug[3]);
  double ener2 = (Ud[4] + gam * pstiff) / gam1 + 0.5 * Ud[0] * (Ud[1] * Ud[1] + Ud[2] * Ud[2] + Ud[3] * Ud[3]);
  double usro = 1.0 / (squsr1 + squsr2);
  double uar2 = (squsr1 * Ug[1] + squsr2 * Ud[1]) * usro;
  double uar3 = (squsr1 * Ug[2] + squsr2 * Ud[2]) * usro;
  double uar4 = (squsr1 * Ug[3] + squsr2 * Ud[3]) * usro;
  double uar5 = ((ener1 + Ug[4]) / squsr1 + (ener2 + Ud[4]) / squsr2) * usro;
  enth = uar5;
  VdotN = n1 * uar2 + n2 * uar3 + n3 * uar4;
  qir = 0.5 * (uar2 * uar2 + uar3 * uar3 + uar4 * uar4);
  double cr2 = gam1 * (uar5 - qir);
  double cr = sqrt(cr2);
  double gc = gam1 / (cr * cr);
  double usc2 = 0.5 / (cr * cr);
  double betaRT2 = 1.0;  // beta = 1 without low-Mach preconditioning
  if (prec) {
    const double beta = turkel_beta(lm, Ug[4], Ud[4], length, qir, 1.0 / cr2);
    betaRT2 = beta * beta;
  }
  double vp1 = VdotN;
  double tt = (1.0 - betaRT2) * VdotN;
  double root = sqrt(tt * tt + 4.0 * betaRT2 * (cr * cr));
  double vp4 = 0.5 * ((1.0 + betaRT2) * VdotN + root);
  double vp5 = 0.5 * ((1.0 + betaRT2) * VdotN - root);
  double rRT = vp4 - vp1 * betaRT2;
  double sRT = vp5 - vp1 * betaRT2;
  double tRT = 0.5 * (vp5 - vp4);
  vp1 = fabs(vp1 - vitno) * gamma_diss;
  vp4 = fabs(vp4 - vitno) * gamma_diss;
  vp5 = fabs(vp5 - vitno) * gamma_diss;
  double bt = betaRT2 * tRT;
  double cc = cr * cr;
  double tmw[25];
  tmw[0] = n1 + n2 * uar4 - uar3 * n3 - gc * n1 * qir;
  tmw[1] = gc * n1 * uar2;
  tmw[2] = gc * n1 * uar3 + n3;
  tmw[3] = gc * n1 * uar4 - n2;
  tmw[4] = -gc * n1;
  tmw[5] = n2 + n3 * uar2 - uar4 * n1 - gc * n2 * qir;
  tmw[6] = gc * n2 * uar2 - n3;
  tmw[7] = gc * n2 * uar3;
  tmw[8] = gc * n2 * uar4 + n1;
  tmw[9] = -gc * n2;
  tmw[10] = n3 + n1 * uar3 - uar2 * n2 - gc * n3 * qir;
  tmw[11] = gc * n3 * uar2 + n2;
  tmw[12] = gc * n3 * uar3 - n1;
  tmw[13] = gc * n3 * uar4;
  tmw[14] = -gc * n3;
  tmw[15] = gam1 * qir * sRT / bt + VdotN * cc / tRT;
  tmw[16] = -gam1 * uar2 * sRT / bt - cc * n1 / tRT;
  tmw[17] = -gam1 * uar3 * sRT / bt - cc * n2 / tRT;
  tmw[18] = -gam1 * uar4 * sRT / bt - cc * n3 / tRT;
  tmw[19] = gam1 * sRT / bt;
  tmw[20] = -gam1 * qir * rRT / bt - VdotN * cc / tRT;
  tmw[21] = gam1 * uar2 * rRT / bt + cc * n1 / tRT;
  tmw[22] = gam1 * uar3 * rRT / bt + cc * n2 / tRT;
  tmw[23] = gam1 * uar4 * rRT / bt + cc * n3 / tRT;
  tmw[24] = -gam1 * rRT / bt;
  const double c2a = n2 * uar2 - n3, c3a = n3 * uar2 + n2;
  const double c1b = n1 * uar3 + n3, c3b = n3 * uar3 - n1;
  const double c1c = n1 * uar4 - n2, c2c = n2 * uar4 + n1;
  const double c1d = n1 * qir + (n3 * uar3 - n2 * uar4);
  const double c2d = n2 * qir + (n1 * uar4 - n3 * uar2);
  const double c3d = n3 * qir + (n2 * uar2 - n1 * uar3);
  const double sgn = sw == 1 ? 1.0 : -1.0;
#pragma unroll
  for (int j = 0; j < 5; ++j) {
    const double t1 = tmw[j], t2 = tmw[5 + j], t3 = tmw[10 + j], t4 = tmw[15 + j], t5 = tmw[20 + j];
    double d0 = vp1 * (n1 * t1 + n2 * t2 + n3 * t3) + vp4 * usc2 * t4 + vp5 * usc2 * t5;
    double d1 = vp1 * (n1 * uar2 * t1 + c2a * t2 + c3a * t3) + vp4 * (usc2 * (uar2 + rRT * n1)) * t4 +
                vp5 * (usc2 * (uar2 + sRT * n1)) * t5;
    double d2 = vp1 * (n2 * uar3 * t2 + c1b * t1 + c3b * t3) + vp4 * (usc2 * (uar3 + rRT * n2)) * t4 +
                vp5 * (usc2 * (uar3 + sRT * n2)) * t5;
    double d3 = vp1 * (n3 * uar4 * t3 + c1c * t1 + c2c * t2) + vp4 * (usc2 * (uar4 + rRT * n3)) * t4 +
                vp5 * (usc2 * (uar4 + sRT * n3)) * t5;
    double d4 = vp1 * (c1d * t1 + c2d * t2 + c3d * t3) + vp4 * (usc2 * (enth + VdotN * rRT)) * t4 +
                vp5 * (usc2 * (enth + VdotN * sRT)) * t5;
    if (sw == 1 || sw == 2) {
      H[j] = H[j] + sgn * d0;
      H[5 + j] = H[5 + j] + sgn * d1;
      H[10 + j] = H[10 + j] + sgn * d2;
      H[15 + j] = H[15 + j] + sgn * d3;
      H[20 + j] = H[20 + j] + sgn * d4;
    }
  }
#pragma unroll
  for (int k = 0; k < 25; ++k) jac[k] = 0.5 * H[k] * rnorm;
}

__device__ __forceinline__ void post_dVdU(const Gas &g, const double V[5], const double m[25], double res[25]) {
  const double gam1 = g.gam1;
  double invrho = 1.0 / V[0];
  double d0 = 1.0, d5 = -invrho * V[1], d6 = invrho, d10 = -invrho * V[2], d12 = invrho, d15 = -invrho * V[3],
         d18 = invrho;
  double d20 = gam1 * 0.5 * (V[1] * V[1] + V[2] * V[2] + V[3] * V[3]);
  double d21 = -gam1 * V[1], d22 = -gam1 * V[2], d23 = -gam1 * V[3], d24 = gam1;
#pragma unroll
  for (int r = 0; r < 5; ++r) {
    const double *a = m + 5 * r;
    double *o = res + 5 * r;
    o[0] = a[0] * d0 + a[1] * d5 + a[2] * d10 + a[3] * d15 + a[4] * d20;
    o[1] = a[1] * d6 + a[4] * d21;
    o[2] = a[2] * d12 + a[4] * d22;
    o[3] = a[3] * d18 + a[4] * d23;
    o[4] = a[4] * d24;
  }
}

__device__ __forceinline__ void wall_jac(const Gas &g, const double n[3], double nvel, const double V[5],
                                         double jac[25]) {
  double dfdV[25];
#pragma unroll
  for (int k = 0; k < 25; ++k) dfdV[k] = 0.0;
  dfdV[9] = n[0];
  dfdV[14] = n[1];
  dfdV[19] = n[2];
  dfdV[24] = nvel;
  post_dVdU(g, V, dfdV, jac);
}

__device__ __forceinline__ void farfield_jac(const Gas &g, const double n[3], double sig, const double VL[5],
                                             const double Ub[5], double jac[25]) {
  const double eps0 = 1.e-6;
  double Veps[5], flux[5], fluxeps[5], dfdVL[25];
  boundflux5(g, n, sig, VL, Ub, flux);
#pragma unroll
  for (int k = 0; k < 5; ++k) Veps[k] = VL[k];
#pragma unroll
  for (int k = 0; k < 5; ++k) {
    double eps = fabs(Veps[k]) < 1.e-10 ? eps0 : eps0 * Veps[k];
    double inveps = 1.0 / eps;
    Veps[k] += eps;
    if (k != 0) Veps[k - 1] = VL[k - 1];
    boundflux5(g, n, sig, Veps, Ub, fluxeps);
#pragma unroll
    for (int j = 0; j < 5; ++j) dfdVL[5 * j + k] = (fluxeps[j] - flux[j]) * inveps;
  }
  post_dVdU(g, VL, dfdVL, jac);
}

// In-place inverse of a row-major 5x5 block by LINPACK dgefa/dgedi in the
// storage precision T (float for the mixed mode, as the reference's
// JacobiPrec<float,5>).  The reference transposes the block, runs the
// column-major LINPACK routines and transposes back; that is the same
// arithmetic as running them on M(i,j) = blk[5i+j] directly, which is what
// A_(i,j) addresses.  Returns 0 or the 1-based zero-pivot column.
template <class T>
__device__ __forceinline__ int invert5(T blk[25]) {
  T *a = blk;
  int ipvt[5];
  T work[5];
#define A_(i, j) a[5 * (i) + (j)]
  for (int k = 0; k < 4; ++k) {
    // pivot: largest |a| in column k, rows k..4, first on ties
    int l = k;
    double mx = fabs((double)A_(k, k));
#pragma unroll
    for (int i = k + 1; i < 5; ++i) {
      double t2 = fabs((double)A_(i, k));
      if (t2 > mx) { mx = t2; l = i; }
    }
    ipvt[k] = l;
    if (A_(l, k) == (T)0) return k + 1;
    if (l != k) { T t = A_(l, k); A_(l, k) = A_(k, k); A_(k, k) = t; }
    T t = (T)(-1. / A_(k, k));
    for (int i = k + 1; i < 5; ++i) A_(i, k) *= t;
    for (int j = k + 1; j < 5; ++j) {
      t = A_(l, j);
      if (l != k) { A_(l, j) = A_(k, j); A_(k, j) = t; }
      for (int i = k + 1; i < 5; ++i) A_(i, j) += t * A_(i, k);
    }
  }
  ipvt[4] = 4;
  if (A_(4, 4) == (T)0) return 5;
  // inverse(U)
  for (int k = 0; k < 5; ++k) {
    A_(k, k) = (T)(1.0 / A_(k, k));
    T t = -A_(k, k);
    for (int i = 0; i < k; ++i) A_(i, k) *= t;
    for (int j = k + 1; j < 5; ++j) {
      t = A_(k, j);
      A_(k, j) = (T)0;
      for (int i = 0; i <= k; ++i) A_(i, j) += t * A_(i, k);
    }
  }
  // inverse(U) * inverse(L)
  for (int k = 3; k >= 0; --k) {
    for (int i = k + 1; i < 5; ++i) { work[i] = A_(i, k); A_(i, k) = (T)0; }
    for (int j = k + 1; j < 5; ++j) {
      T t = work[j];
      for (int i = 0; i < 5; ++i) A_(i, k) += t * A_(i, j);
    }
    int l = ipvt[k];
    if (l != k)
      for (int i = 0; i < 5; ++i) { T t3 = A_(i, k); A_(i, k) = A_(i, l); A_(i, l) = t3; }
  }
#undef A_
  return 0;
}

}  // namespace aerofx
