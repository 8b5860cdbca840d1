// device_sens.h -- templated pieces of the residual shared by its derivative
// passes (adjoint.hip: reverse mode, tangent.hip: forward mode).  With
// T = double they are the primal arithmetic; with T = Dual they carry one
// directional derivative.
#pragma once
#include <hip/hip_runtime.h>

#include "context.h"
#include "device_dual.h"

namespace aerofx {

// (1/r11, 1/r22, 1/r33, r12/r11, r23/r22, psi) from the LSQ moments
// M = sum dx dx^T (xx, xy, xz, yy, yz, zz), in the order of the set-up
// (SubDomainCore.C:1755-1777, SubDomain.C:137-159).
template <class T>
__device__ __forceinline__ void lsq_params(const T M[6], T L[6]) {
  T r11 = gsqrt(M[0]);
  T or11 = 1.0 / r11;
  T r12 = M[1] * or11;
  T r13 = M[2] * or11;
  T r22 = gsqrt(M[3] - r12 * r12);
  T r23 = (M[4] - r12 * r13) / r22;
  T r33 = gsqrt(M[5] - (r13 * r13 + r23 * r23));
  T o11 = 1.0 / r11, o22 = 1.0 / r22, o33 = 1.0 / r33;
  L[0] = o11;
  L[1] = o22;
  L[2] = o33;
  L[3] = r12 * o11;
  L[4] = r23 * o22;
  L[5] = (r12 * r23 - r13 * r22) * o11 * o22;
}

// LSQ weights W(dx; L) (computeLocalWeightsLeastSquares, SubDomain.C:137-159)
// in the gradient pass's operation order.
template <class T>
__device__ __forceinline__ void lsq_weights_t(const T L[6], T dx0, T dx1, T dx2, T w[3]) {
  const T or11 = L[0], or22 = L[1], or33 = L[2], r12or11 = L[3], r23or22 = L[4], psi = L[5];
  const T alpha1 = dx0 * or11 * or11;
  const T alpha2 = (dx1 - r12or11 * dx0) * or22 * or22;
  const T alpha3 = (dx2 - r23or22 * dx1 + psi * dx0) * or33 * or33;
  w[0] = alpha1 - r12or11 * alpha2 + psi * alpha3;
  w[1] = alpha2 - r23or22 * alpha3;
  w[2] = alpha3;
}

// boundary flux of one face node; nv = (1/3) face normal
template <class T>
__device__ __forceinline__ void face_flux(const FluxConsts &fc, int code, const T nv[3], const T V[5], T fl[5]) {
  if (is_wall_code(code)) {
    wallflux(nv, 0.0, V, fl);
  } else {
    const double *ub = (code == AEROFX_BC_INLET || code == -AEROFX_BC_INLET) ? fc.ub_in : fc.ub_out;
    boundflux5(fc.gas, nv, 0.0, V, ub, fl);
  }
}

// Van Albada reconstruction of one variable (RecFcn.h:433-446) for either
// number type: V_ij = V_i + VA(d_i, dV)/2, V_ji = V_j - VA(d_j, dV)/2.
template <class T>
__device__ __forceinline__ void vanalbada_t(double beta, T Vi, T ddVij, T Vj, T ddVji, T &Vij, T &Vji) {
  const double beta2 = 1.0 - 2.0 * beta;
  const double eps0 = 1.e-16;
  T dVji = Vj - Vi;
  T dV = beta2 * dVji;
  T dVi = dV + 2.0 * beta * ddVij;
  T dVj = dV + 2.0 * beta * ddVji;
  Vij = Vi + 0.5 * va_fcn(eps0, dVi, dVji);
  Vji = Vj - 0.5 * va_fcn(eps0, dVj, dVji);
}

}  // namespace aerofx
