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

template <class T>
__device__ __forceinline__ void wallflux(const T nv[3], double nvel, const T V[5], T flux[5]) {
  flux[0] = 0.0;
  flux[1] = V[4] * nv[0];
  flux[2] = V[4] * nv[1];
  flux[3] = V[4] * nv[2];
  flux[4] = V[4] * nvel;
}

}  // namespace aerofx
