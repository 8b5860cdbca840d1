// device_physics.h -- per-edge and per-face physics as gfx950 device code.
//
// Operation order follows the reference kernels so that, built without FMA
// contraction, each routine reproduces the reference bit for bit:
//   prim_from_cons  VarFcnSGEuler::conservativeToPrimitive (VarFcnSGEuler.h:307-336)
//   va_fcn / vanalbada  RecFcn::computeVanAlbadaFcn, RecFcn::vanalbada
//                   (RecFcn.h:269-281, 433-446; eps0 = 1e-16, RecFcn.h:47)
//   roeflux7        ROEFLUX7, type 0 (f77src/roeflux7.f:26-187)
//   boundflux5      BOUNDFLUX5, type 0 (f77src/boundflux5.f:23-98)
//   wallflux        FluxFcnSGWallEuler3D::compute (FluxFcnDescSG.C:779-789)
//   roe_turkel_flux ROEFLUX5, type 0, gamma = 1, Ugr = Ug, Udr = Ud
//                   (f77src/roeflux5.f; called by FluxFcnSGApprJacRoeEuler3D::
//                   compute, FluxFcnDescSG.C:125-139, when the spatial
//                   low-Mach preconditioner is on)
#pragma once
#include <hip/hip_runtime.h>

#include "device_dual.h"

namespace aerofx {

struct Gas {
  double gam, gam1, invgam1, pstiff;
};

// Division and square root of the flux routines: IEEE-rounded, as the
// reference's; the Dual overloads (device_dual.h) carry derivatives.
__device__ __forceinline__ double ddiv(double a, double b) { return a / b; }
__device__ __forceinline__ double dsqrt(double x) { return sqrt(x); }

template <class T>
__device__ __forceinline__ void prim_from_cons(const T U[5], const Gas &g, T V[5]) {
  V[0] = U[0];
  T invRho = 1.0 / U[0];
  V[1] = U[1] * invRho;
  V[2] = U[2] * invRho;
  V[3] = U[3] * invRho;
  T vel2 = V[1] * V[1] + V[2] * V[2] + V[3] * V[3];
  V[4] = (g.gam - 1.0) * (U[4] - 0.5 * U[0] * vel2) - g.gam * g.pstiff;
}

// Branch-free: the numerator is selected, then one division (0/den = 0 for
// ab < 0), so lanes of a wave never diverge.
template <class T>
__device__ __forceinline__ T va_fcn(double eps, T a, T b) {
  T a2 = a * a;
  T b2 = b * b;
  T ab = a * b;
  T den = a2 + b2 + 2.0 * eps;
  T num = ab > 0.0 ? (a * (b2 + eps) + b * (a2 + eps)) : (ab == 0.0 ? 0.5 * eps * (a + b) : T(0.0));
  return ddiv(num, den);
}

// V_ij = V_i + 1/2 VA(d_i, dV), V_ji = V_j - 1/2 VA(d_j, dV) for one variable.
__device__ __forceinline__ void vanalbada(double beta, double Vi, double ddVij, double Vj,
                                          double ddVji, double &Vij, double &Vji) {
  const double beta2 = 1.0 - 2.0 * beta;
  const double eps0 = 1.e-16;
  double dVji = Vj - Vi;
  double dV = beta2 * dVji;
  double dVi = dV + 2.0 * beta * ddVij;
  double dVj = dV + 2.0 * beta * ddVji;
  Vij = Vi + 0.5 * va_fcn(eps0, dVi, dVji);
  Vji = Vj - 0.5 * va_fcn(eps0, dVj, dVji);
}

// Roe flux through the dual face with area vector en (not normalised) and
// normal face velocity evitno; states Ug (left = i side), Ud (right) are
// primitive (rho, u, v, w, p).
template <class T>
__device__ __forceinline__ void roeflux7(const Gas &g, const T en[3], T evitno, const T Ug[5], const T Ud[5],
                                         T phi[5]) {
  const double gam = g.gam, gam1 = g.gam1, invgam1 = g.invgam1, pstiff = g.pstiff;
  T rnorm = dsqrt(en[0] * en[0] + en[1] * en[1] + en[2] * en[2]);
  T invnorm = ddiv(1.0, rnorm);
  T n0 = en[0] * invnorm, n1 = en[1] * invnorm, n2 = en[2] * invnorm;
  T vitno = evitno * invnorm;
  T vitg2 = Ug[1] * Ug[1] + Ug[2] * Ug[2] + Ug[3] * Ug[3];
  T vitd2 = Ud[1] * Ud[1] + Ud[2] * Ud[2] + Ud[3] * Ud[3];
  T squsr1 = dsqrt(Ug[0]);
  T squsr2 = dsqrt(Ud[0]);
  T ener1 = (Ug[4] + gam * pstiff) * invgam1 + 0.5 * Ug[0] * vitg2;
  T ener2 = (Ud[4] + gam * pstiff) * invgam1 + 0.5 * Ud[0] * vitd2;
  T usro = ddiv(1.0, squsr1 + squsr2);
  T uar2 = (squsr1 * Ug[1] + squsr2 * Ud[1]) * usro;
  T uar3 = (squsr1 * Ug[2] + squsr2 * Ud[2]) * usro;
  T uar4 = (squsr1 * Ug[3] + squsr2 * Ud[3]) * usro;
  T uar5 = (ddiv(ener1 + Ug[4], squsr1) + ddiv(ener2 + Ud[4], squsr2)) * usro;
  T VdotNt = n0 * uar2 + n1 * uar3 + n2 * uar4;
  const bool left = (VdotNt - vitno > 0);
  // centred term: one-sided, upwinded by the Roe normal velocity
  const T *S = left ? Ug : Ud;
  T vit2 = left ? vitg2 : vitd2;
  T VdotN = S[1] * n0 + S[2] * n1 + S[3] * n2;
  T H = gam * (S[4] + pstiff) + 0.5 * gam1 * S[0] * vit2;
  H = ddiv(H, gam1 * S[0]);
  phi[0] = S[0] * (VdotN - vitno);
  phi[1] = phi[0] * S[1] + S[4] * n0;
  phi[2] = phi[0] * S[2] + S[4] * n1;
  phi[3] = phi[0] * S[3] + S[4] * n2;
  phi[4] = phi[0] * H + S[4] * vitno;
  // single acoustic wave of the Roe dissipation
  VdotN = VdotNt;
  T qir = 0.5 * (uar2 * uar2 + uar3 * uar3 + uar4 * uar4);
  T cr2 = gam1 * (uar5 - qir);
  T cr = dsqrt(cr2);
  cr2 = ddiv(1.0, cr2);
  T vp4 = VdotN + cr;
  T vp5 = VdotN - cr;
  T r = cr;
  T s = -cr;
  T cr2byt = -cr;
  T invtbeta2 = ddiv(-1.0, cr);
  T dif1 = -Ug[0] + Ud[0];
  T dif2 = -Ug[0] * Ug[1] + Ud[0] * Ud[1];
  T dif3 = -Ug[0] * Ug[2] + Ud[0] * Ud[2];
  T dif4 = -Ug[0] * Ug[3] + Ud[0] * Ud[3];
  T dif5 = -ener1 + ener2;
  vp4 = (vp4 - vitno);
  vp5 = (vp5 - vitno);
  T a1 = VdotN * dif1 - n0 * dif2 - n1 * dif3 - n2 * dif4;
  T a2 = qir * dif1 - uar2 * dif2 - uar3 * dif3 - uar4 * dif4 + dif5;
  T flur4, flur5;
  if (!left) {
    T m4 = vp4 > 0.0 ? vp4 : T(0.0);
    flur4 = m4 * (cr2byt * a1 + gam1 * s * invtbeta2 * a2);
    flur5 = T(0.0);
  } else {
    flur4 = T(0.0);
    T m5 = vp5 < 0.0 ? vp5 : T(0.0);
    flur5 = -m5 * (-cr2byt * a1 + (-gam1 * r * invtbeta2 * a2));
  }
  T fs = flur4 + flur5;
  T rs = r * flur4 + s * flur5;
  phi[0] = phi[0] - 0.5 * fs * cr2;
  phi[1] = phi[1] - (n0 * rs + uar2 * fs) * 0.5 * cr2;
  phi[2] = phi[2] - (n1 * rs + uar3 * fs) * 0.5 * cr2;
  phi[3] = phi[3] - (n2 * rs + uar4 * fs) * 0.5 * cr2;
  phi[4] = phi[4] - (VdotN * rs + uar5 * fs) * 0.5 * cr2;
  phi[0] = phi[0] * rnorm;
  phi[1] = phi[1] * rnorm;
  phi[2] = phi[2] * rnorm;
  phi[3] = phi[3] * rnorm;
  phi[4] = phi[4] * rnorm;
}

// Steger-Warming farfield flux A+(V) W(V) + A-(V) Ub through area vector nv.
template <class T>
__device__ __forceinline__ void boundflux5(const Gas &g, const T nv[3], double sig, const T ua[5],
                                           const double ub[5], T flux[5]) {
  const double gam = g.gam, gam1 = g.gam1;
  T cap = gsqrt(nv[0] * nv[0] + nv[1] * nv[1] + nv[2] * nv[2]);
  T capd1 = nv[0] / cap, capd2 = nv[1] / cap, capd3 = nv[2] / cap;
  T ro = ua[0];
  T usro = 1.0 / ro;
  T u = ua[1], v = ua[2], w = ua[3];
  T vit = u * u + v * v + w * w;
  T pm = ua[4];
  T rhoE = pm / gam1 + 0.5 * ro * vit;
  T c2 = gam * pm * usro;
  T c = gsqrt(c2);
  T usc = 1.0 / c;
  T cc = 0.5 * vit + c2 / gam1;
  T unx = capd1 * u + capd2 * v + capd3 * w;
  T t1 = cap * unx - sig, t4 = cap * (unx + c) - sig, t5 = cap * (unx - c) - sig;
  T vp1 = t1 > 0.0 ? t1 : T(0.0);
  T vp4 = t4 > 0.0 ? t4 : T(0.0);
  T vp5 = t5 > 0.0 ? t5 : T(0.0);
  T vpp = vp4 + vp5;
  T vpm = vp4 - vp5;
  T fgp0 = (gam1 * ro * vp1 + 0.5 * ro * vpp) / gam;
  T fgp1 = fgp0 * u + 0.5 * pm * usc * capd1 * vpm;
  T fgp2 = fgp0 * v + 0.5 * pm * usc * capd2 * vpm;
  T fgp3 = fgp0 * w + 0.5 * pm * usc * capd3 * vpm;
  T fgp4 = 0.5 * ((gam1 / gam) * ro * vit * (vp1 - 0.5 * vpp) + rhoE * vpp + ro * unx * c * vpm / gam);
  vp1 = t1 < 0.0 ? t1 : T(0.0);
  vp4 = t4 < 0.0 ? t4 : T(0.0);
  vp5 = t5 < 0.0 ? t5 : T(0.0);
  vpp = vp1 - 0.5 * (vp4 + vp5);
  vpm = 0.5 * (vp5 - vp4);
  T xn1 = gam1 * usc * (-0.5 * vit * ub[0] + u * ub[1] + v * ub[2] + w * ub[3] - ub[4]);
  T xn2 = unx * ub[0] - capd1 * ub[1] - capd2 * ub[2] - capd3 * ub[3];
  T ff = usc * (vpp * xn1 + vpm * xn2);
  T gg = vpm * xn1 + vpp * xn2;
  flux[0] = fgp0 + (vp1 * ub[0] + ff);
  flux[1] = fgp1 + (vp1 * ub[1] + u * ff + capd1 * gg);
  flux[2] = fgp2 + (vp1 * ub[2] + v * ff + capd2 * gg);
  flux[3] = fgp3 + (vp1 * ub[3] + w * ff + capd3 * gg);
  flux[4] = fgp4 + (vp1 * ub[4] + cc * ff + unx * gg);
}

// Spatial low-Mach preconditioner constants (SpatialLowMachPrec,
// LowMachPrec.h): minimum Mach, slope k1 on the local Mach number, cut-off
// Mach, shock reducer, and the 1/Re of the edge (0 for inviscid flow).
struct LowMach {
  double min_mach, slope, cutoff, shock, irey;
};

// Turkel's preconditioning parameter for the edge: beta = min((1 + sqrt(1/Re))
// max(k1 M_roe, M_min) + c_shock |p_g - p_d| / ((p_g + p_d) L), M_cut).
__host__ __device__ __forceinline__ double turkel_beta(const LowMach &lm, double pg, double pd, double length,
                                                       double q2, double inv_c2) {
  const double shock = fabs(pg - pd) / (pg + pd) / length;
  const double mroe = sqrt(2.0 * q2 * inv_c2);
  double b = fmax(lm.slope * mroe, lm.min_mach);
  b = (1.0 + sqrt(lm.irey)) * b + lm.shock * shock;
  return fmin(b, lm.cutoff);
}

// Roe flux with Turkel-preconditioned dissipation through area vector en
// (not normalised), static face, for primitive states Ug (i side), Ud (j
// side):  F = 1/2 |n| [F(Ug).n + F(Ud).n - P^-1 |P A_roe| (Ud - Ug)].
// The preconditioned system has one convective eigenvalue u_n (entropy and
// two shear waves) and two acoustic ones
//   lambda_+- = 1/2 ((1 + beta^2) u_n +- sqrt(((1 - beta^2) u_n)^2 + 4 beta^2 c^2)),
// with r, s = lambda_+- - beta^2 u_n and t = 1/2 (lambda_- - lambda_+); with
// beta = 1 it is Roe's flux.  Each intermediate is evaluated in the order
// of roeflux5.f so that, built without FMA contraction, the result matches
// the reference's bit for bit.
__host__ __device__ __forceinline__ void roe_turkel_flux(const Gas &g, const LowMach &lm, bool prec,
                                                         const double en[3], double length, const double Ug[5],
                                                         const double Ud[5], double phi[5]) {
  const double gam = g.gam, gam1 = g.gam1, pstiff = g.pstiff;
  const double rnorm = sqrt(en[0] * en[0] + en[1] * en[1] + en[2] * en[2]);
  const double invnorm = 1.0 / rnorm;
  const double n[3] = {en[0] * invnorm, en[1] * invnorm, en[2] * invnorm};
  // skew[d][k] = (n x e_k)_d style rotation terms of the shear waves
  const double skew[3][3] = {{0.0, n[2], -n[1]}, {-n[2], 0.0, n[0]}, {n[1], -n[0], 0.0}};
  // centred part: F(Ug).n + F(Ud).n
  double q2g = Ug[1] * Ug[1] + Ug[2] * Ug[2] + Ug[3] * Ug[3];
  double q2d = Ud[1] * Ud[1] + Ud[2] * Ud[2] + Ud[3] * Ud[3];
  {
    const double ung = Ug[1] * n[0] + Ug[2] * n[1] + Ug[3] * n[2];
    const double Hg = (gam * (Ug[4] + pstiff) + 0.5 * gam1 * Ug[0] * q2g) / (gam1 * Ug[0]);
    const double und = Ud[1] * n[0] + Ud[2] * n[1] + Ud[3] * n[2];
    const double Hd = (gam * (Ud[4] + pstiff) + 0.5 * gam1 * Ud[0] * q2d) / (gam1 * Ud[0]);
    const double mg = Ug[0] * ung;
    phi[0] = mg + Ud[0] * und;
    for (int d = 0; d < 3; ++d) phi[1 + d] = (mg * Ug[1 + d] + Ug[4] * n[d]) + Ud[0] * Ud[1 + d] * und + Ud[4] * n[d];
    phi[4] = mg * Hg + Ud[0] * und * Hd;
  }
  // Roe average
  const double sg = sqrt(Ug[0]), sd = sqrt(Ud[0]);
  const double Eg = (Ug[4] + gam * pstiff) / gam1 + 0.5 * Ug[0] * q2g;
  const double Ed = (Ud[4] + gam * pstiff) / gam1 + 0.5 * Ud[0] * q2d;
  const double inv_s = 1.0 / (sg + sd);
  double ua[3];
  for (int d = 0; d < 3; ++d) ua[d] = (sg * Ug[1 + d] + sd * Ud[1 + d]) * inv_s;
  const double Ha = ((Eg + Ug[4]) / sg + (Ed + Ud[4]) / sd) * inv_s;
  const double un = n[0] * ua[0] + n[1] * ua[1] + n[2] * ua[2];
  const double q2 = 0.5 * (ua[0] * ua[0] + ua[1] * ua[1] + ua[2] * ua[2]);
  const double tet[3] = {n[2] * ua[1] - n[1] * ua[2], n[0] * ua[2] - n[2] * ua[0], n[1] * ua[0] - n[0] * ua[1]};
  const double c2 = gam1 * (Ha - q2);
  const double c = sqrt(c2);
  const double ic2 = 1.0 / c2;
  const double beta = prec ? turkel_beta(lm, Ug[4], Ud[4], length, q2, ic2) : 1.0;
  const double b2 = beta * beta;
  // jump in the conservative variables
  double dU[5];
  dU[0] = -Ug[0] + Ud[0];
  for (int d = 0; d < 3; ++d) dU[1 + d] = -Ug[0] * Ug[1 + d] + Ud[0] * Ud[1 + d];
  {
    const double eg = (Ug[4] + gam * pstiff) / gam1 + 0.5 * Ug[0] * (Ug[1] * Ug[1] + Ug[2] * Ug[2] + Ug[3] * Ug[3]);
    const double ed = (Ud[4] + gam * pstiff) / gam1 + 0.5 * Ud[0] * (Ud[1] * Ud[1] + Ud[2] * Ud[2] + Ud[3] * Ud[3]);
    dU[4] = -eg + ed;
  }
  // eigenvalues of the preconditioned system
  const double sq = ((1.0 - b2) * un) * ((1.0 - b2) * un);
  const double lp = 0.5 * ((1.0 + b2) * un + sqrt(sq + 4.0 * b2 * (c * c)));
  const double lm_ = 0.5 * ((1.0 + b2) * un - sqrt(sq + 4.0 * b2 * (c * c)));
  const double r = lp - un * b2, s = lm_ - un * b2, t = 0.5 * (lm_ - lp);
  const double a0 = fabs(un), ap = fabs(lp), am = fabs(lm_);
  // wave strengths times |eigenvalue|: entropy + shear (3), acoustic (2)
  double w[3];
  for (int d = 0; d < 3; ++d) {
    double acc = (n[d] * (1.0 - gam1 * q2 * ic2) - tet[d]) * dU[0];
    for (int k = 0; k < 3; ++k) {
      const double base = n[d] * gam1 * ua[k] * ic2;
      acc = acc + (k == d ? base : skew[d][k] + base) * dU[1 + k];
    }
    w[d] = a0 * (acc - (n[d] * gam1 * ic2) * dU[4]);
  }
  const double cc2 = c * c;
  double wp = (cc2 * un / t + gam1 * q2 * s / (t * b2)) * dU[0];
  for (int k = 0; k < 3; ++k) wp = wp - (cc2 * n[k] / t + gam1 * ua[k] * s / (t * b2)) * dU[1 + k];
  wp = ap * (wp + gam1 * s / (t * b2) * dU[4]);
  double wm = -(cc2 * un / t + gam1 * q2 * r / (b2 * t)) * dU[0];
  for (int k = 0; k < 3; ++k) wm = wm + (cc2 * n[k] / t + gam1 * ua[k] * r / (b2 * t)) * dU[1 + k];
  wm = am * (wm - gam1 * r / (b2 * t) * dU[4]);
  // dissipation: right eigenvectors times the wave strengths
  const double sum = wp + wm, rs = r * wp + s * wm;
  phi[0] = phi[0] - (n[0] * w[0] + n[1] * w[1] + n[2] * w[2] + 0.5 * sum * ic2);
  for (int m = 0; m < 3; ++m) {
    double acc = (m == 0 ? ua[m] * n[0] : ua[m] * n[0] + skew[0][m]) * w[0];
    acc = acc + (m == 1 ? ua[m] * n[1] : ua[m] * n[1] + skew[1][m]) * w[1];
    acc = acc + (m == 2 ? ua[m] * n[2] : ua[m] * n[2] + skew[2][m]) * w[2];
    phi[1 + m] = phi[1 + m] - (acc + (n[m] * rs + ua[m] * sum) * 0.5 * ic2);
  }
  phi[4] = phi[4] - ((q2 * n[0] + tet[0]) * w[0] + (q2 * n[1] + tet[1]) * w[1] + (q2 * n[2] + tet[2]) * w[2] +
                     (un * rs + Ha * sum) * 0.5 * ic2);
  for (int k = 0; k < 5; ++k) phi[k] = phi[k] * 0.5 * rnorm;
}

template <class T>
__device__ __forceinline__ void wallflux(const T nv[3], double nvel, const T V[5], T flux[5]) {
  flux[0] = 0.0;
  flux[1] = V[4] * nv[0];
  flux[2] = V[4] * nv[1];
  flux[3] = V[4] * nv[2];
  flux[4] = V[4] * nvel;
}

}  // namespace aerofx
