// physics_eval.hip -- the per-edge, per-face and per-block routines that the
// residual, Jacobian and preconditioner kernels inline (device_physics.h,
// device_jacobian.h), evaluated one case per thread on caller arrays.  It lets
// a test hold the device code itself against stored outputs of the
// reference's own routines, case by case and bit for bit.
#include <hip/hip_runtime.h>

#include "context.h"
#include "device_jacobian.h"
#include "host_internal.h"

namespace aerofx {
namespace {

constexpr int kEvalBlock = 128;

__global__ __launch_bounds__(kEvalBlock) void k_physics_eval(int kind, int n, Gas g, double scalar,
                                                             const double *__restrict__ a,
                                                             const double *__restrict__ b,
                                                             const double *__restrict__ c,
                                                             const double *__restrict__ d, double *__restrict__ out,
                                                             int *__restrict__ bad) {
  const int q = blockIdx.x * kEvalBlock + threadIdx.x;
  if (q >= n) return;
  const size_t s = (size_t)q;
  switch (kind) {
    case AEROFX_PHYS_ROEFLUX7: {
      double nv[3], vl[5], vr[5], f[5];
      for (int k = 0; k < 3; ++k) nv[k] = a[3 * s + k];
      for (int k = 0; k < 5; ++k) {
        vl[k] = b[5 * s + k];
        vr[k] = c[5 * s + k];
      }
      roeflux7(g, nv, 0.0, vl, vr, f);
      for (int k = 0; k < 5; ++k) out[5 * s + k] = f[k];
      break;
    }
    case AEROFX_PHYS_BOUNDFLUX5: {
      double nv[3], v[5], ub[5], f[5];
      for (int k = 0; k < 3; ++k) nv[k] = a[3 * s + k];
      for (int k = 0; k < 5; ++k) {
        v[k] = b[5 * s + k];
        ub[k] = c[5 * s + k];
      }
      boundflux5(g, nv, 0.0, v, ub, f);
      for (int k = 0; k < 5; ++k) out[5 * s + k] = f[k];
      break;
    }
    case AEROFX_PHYS_VANALBADA: {
      for (int k = 0; k < 5; ++k) {
        double vij, vji;
        vanalbada(scalar, a[5 * s + k], b[5 * s + k], c[5 * s + k], d[5 * s + k], vij, vji);
        out[10 * s + k] = vij;
        out[10 * s + 5 + k] = vji;
      }
      break;
    }
    case AEROFX_PHYS_ROEJAC6_SW1:
    case AEROFX_PHYS_ROEJAC6_SW2: {
      double nv[3], ug[5], ud[5], jac[25];
      for (int k = 0; k < 3; ++k) nv[k] = a[3 * s + k];
      for (int k = 0; k < 5; ++k) {
        ug[k] = b[5 * s + k];
        ud[k] = c[5 * s + k];
      }
      roejac6(1.0, g, nv, 0.0, ug, ud, kind == AEROFX_PHYS_ROEJAC6_SW1 ? 1 : 2, jac);
      for (int k = 0; k < 25; ++k) out[25 * s + k] = jac[k];
      break;
    }
    case AEROFX_PHYS_ROEJAC6_LM_SW1:
    case AEROFX_PHYS_ROEJAC6_LM_SW2: {
      double nv[3], ug[5], ud[5], jac[25];
      for (int k = 0; k < 3; ++k) nv[k] = a[3 * s + k];
      for (int k = 0; k < 5; ++k) {
        ug[k] = b[5 * s + k];
        ud[k] = c[5 * s + k];
      }
      const double *pr = d + 7 * s;
      const LowMach lm{pr[0], pr[1], pr[2], pr[3], pr[4]};
      roejac6(1.0, g, nv, 0.0, ug, ud, kind == AEROFX_PHYS_ROEJAC6_LM_SW1 ? 1 : 2, jac, pr[6] != 0.0, lm, pr[5]);
      for (int k = 0; k < 25; ++k) out[25 * s + k] = jac[k];
      break;
    }
    case AEROFX_PHYS_INVERT5_F64: {
      double blk[25];
      for (int k = 0; k < 25; ++k) blk[k] = a[25 * s + k];
      if (invert5<double>(blk) != 0) atomicAdd(bad, 1);
      for (int k = 0; k < 25; ++k) out[25 * s + k] = blk[k];
      break;
    }
    case AEROFX_PHYS_INVERT5_F32: {
      float blk[25];
      for (int k = 0; k < 25; ++k) blk[k] = (float)a[25 * s + k];
      if (invert5<float>(blk) != 0) atomicAdd(bad, 1);
      for (int k = 0; k < 25; ++k) out[25 * s + k] = (double)blk[k];
      break;
    }
    case AEROFX_PHYS_ROEFLUX5: {
      double nv[3], vl[5], vr[5], f[5];
      for (int k = 0; k < 3; ++k) nv[k] = a[3 * s + k];
      for (int k = 0; k < 5; ++k) {
        vl[k] = b[5 * s + k];
        vr[k] = c[5 * s + k];
      }
      const double *pr = d + 7 * s;
      const LowMach lm{pr[0], pr[1], pr[2], pr[3], pr[4]};
      roe_turkel_flux(g, lm, pr[6] != 0.0, nv, pr[5], vl, vr, f);
      for (int k = 0; k < 5; ++k) out[5 * s + k] = f[k];
      break;
    }
    default:
      break;
  }
}

}  // namespace
}  // namespace aerofx

using namespace aerofx;

extern "C" int aerofx_physics_eval(int kind, int n, double scalar, const double *a, const double *b, const double *c,
                                   const double *d, double *out) {
  if (kind < AEROFX_PHYS_ROEFLUX7 || kind > AEROFX_PHYS_ROEJAC6_LM_SW2)
    return set_error(AEROFX_ERR_ARG, "physics_eval: unknown kind");
  if (n < 0) return set_error(AEROFX_ERR_ARG, "physics_eval: n < 0");
  const bool need_bc = kind != AEROFX_PHYS_INVERT5_F64 && kind != AEROFX_PHYS_INVERT5_F32;
  if (!a || !out || (need_bc && (!b || !c)) || ((kind == AEROFX_PHYS_VANALBADA || kind >= AEROFX_PHYS_ROEFLUX5) && !d))
    return set_error(AEROFX_ERR_ARG, "physics_eval: null argument");
  if (n == 0) return AEROFX_OK;
  Gas g;
  g.gam = kind == AEROFX_PHYS_VANALBADA ? 1.4 : scalar;
  g.gam1 = g.gam - 1.0;
  g.invgam1 = 1.0 / g.gam1;
  g.pstiff = 0.0;
  int *bad = nullptr;
  hipError_t e = hipMalloc(&bad, sizeof(int));
  if (e != hipSuccess) return set_hip_error(e, "physics_eval: hipMalloc");
  int h_bad = 0;
  e = hipMemset(bad, 0, sizeof(int));
  if (e == hipSuccess) {
    hipLaunchKernelGGL(k_physics_eval, dim3((n + kEvalBlock - 1) / kEvalBlock), dim3(kEvalBlock), 0, 0, kind, n, g,
                       scalar, a, b, c, d, out, bad);
    e = hipGetLastError();
  }
  if (e == hipSuccess) e = hipMemcpy(&h_bad, bad, sizeof(int), hipMemcpyDeviceToHost);
  (void)hipFree(bad);
  if (e != hipSuccess) return set_hip_error(e, "physics_eval");
  return h_bad;
}
