// device_dual.h -- first-order forward-mode numbers for the adjoint kernels.
//
// A Dual carries a value and one directional derivative.  The templated
// physics routines in device_physics.h, instantiated with Dual, give the
// derivative of a flux along one seeded input direction; the adjoint kernels
// contract those with the adjoint weights.  Comparisons and selections look
// at the value only, so a piecewise-defined routine is differentiated on the
// branch its value takes.
#pragma once
#include <hip/hip_runtime.h>

namespace aerofx {

struct Dual {
  double v, d;
  __device__ __forceinline__ Dual() : v(0.0), d(0.0) {}
  __device__ __forceinline__ Dual(double x) : v(x), d(0.0) {}  // NOLINT: constants promote
  __device__ __forceinline__ Dual(double x, double dx) : v(x), d(dx) {}
};

__device__ __forceinline__ Dual operator+(Dual a, Dual b) { return Dual(a.v + b.v, a.d + b.d); }
__device__ __forceinline__ Dual operator-(Dual a, Dual b) { return Dual(a.v - b.v, a.d - b.d); }
__device__ __forceinline__ Dual operator-(Dual a) { return Dual(-a.v, -a.d); }
__device__ __forceinline__ Dual operator*(Dual a, Dual b) { return Dual(a.v * b.v, a.d * b.v + a.v * b.d); }
__device__ __forceinline__ Dual operator/(Dual a, Dual b) {
  const double q = a.v / b.v;
  return Dual(q, (a.d - q * b.d) / b.v);
}
__device__ __forceinline__ Dual operator+(double a, Dual b) { return Dual(a + b.v, b.d); }
__device__ __forceinline__ Dual operator+(Dual a, double b) { return Dual(a.v + b, a.d); }
__device__ __forceinline__ Dual operator-(double a, Dual b) { return Dual(a - b.v, -b.d); }
__device__ __forceinline__ Dual operator-(Dual a, double b) { return Dual(a.v - b, a.d); }
__device__ __forceinline__ Dual operator*(double a, Dual b) { return Dual(a * b.v, a * b.d); }
__device__ __forceinline__ Dual operator*(Dual a, double b) { return Dual(a.v * b, a.d * b); }
__device__ __forceinline__ Dual operator/(Dual a, double b) { return Dual(a.v / b, a.d / b); }
__device__ __forceinline__ Dual operator/(double a, Dual b) {
  const double q = a / b.v;
  return Dual(q, -q * b.d / b.v);
}
__device__ __forceinline__ Dual &operator+=(Dual &a, Dual b) { return a = a + b; }
__device__ __forceinline__ Dual &operator-=(Dual &a, Dual b) { return a = a - b; }

__device__ __forceinline__ bool operator>(Dual a, double b) { return a.v > b; }
__device__ __forceinline__ bool operator<(Dual a, double b) { return a.v < b; }
__device__ __forceinline__ bool operator>=(Dual a, double b) { return a.v >= b; }
__device__ __forceinline__ bool operator<=(Dual a, double b) { return a.v <= b; }
__device__ __forceinline__ bool operator==(Dual a, double b) { return a.v == b; }
__device__ __forceinline__ bool operator>(Dual a, Dual b) { return a.v > b.v; }
__device__ __forceinline__ bool operator<(Dual a, Dual b) { return a.v < b.v; }

// value of a double or a Dual
__device__ __forceinline__ double val(double x) { return x; }
__device__ __forceinline__ double val(Dual x) { return x.v; }

// square root for the templated routines (IEEE sqrt on the value)
__device__ __forceinline__ double gsqrt(double x) { return sqrt(x); }
__device__ __forceinline__ Dual gsqrt(Dual x) {
  const double s = sqrt(x.v);
  return Dual(s, 0.5 * x.d / s);
}

// the quotient and square root the Roe flux uses (ddiv / dsqrt for double)
__device__ __forceinline__ Dual ddiv(Dual a, Dual b) { return a / b; }
__device__ __forceinline__ Dual dsqrt(Dual x) { return gsqrt(x); }

}  // namespace aerofx
