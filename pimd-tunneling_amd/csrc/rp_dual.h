// rp_dual.h -- forward-mode dual numbers for the Hessian-vector products.
//
// The potentials in rp_pes.h are written once as reverse-mode (adjoint) energy +
// gradient code templated on the scalar type.  Instantiated with Dual<double>
// and seeded with x + eps*v, the gradient code returns grad(x) + eps*H(x) v:
// forward-over-reverse, exact to rounding, no finite differences.
#pragma once
#include <hip/hip_runtime.h>
#include <math.h>

namespace rp {

// keep the double overloads visible next to the Dual ones below
using ::acos;
using ::exp;
using ::sqrt;

struct Dual {
  double v, d;
  __host__ __device__ Dual() : v(0.0), d(0.0) {}
  __host__ __device__ Dual(double a) : v(a), d(0.0) {}
  __host__ __device__ Dual(double a, double b) : v(a), d(b) {}
};

__host__ __device__ inline Dual operator+(Dual a, Dual b) { return Dual(a.v + b.v, a.d + b.d); }
__host__ __device__ inline Dual operator-(Dual a, Dual b) { return Dual(a.v - b.v, a.d - b.d); }
__host__ __device__ inline Dual operator-(Dual a) { return Dual(-a.v, -a.d); }
__host__ __device__ inline Dual operator*(Dual a, Dual b) { return Dual(a.v * b.v, a.d * b.v + a.v * b.d); }
__host__ __device__ inline Dual operator/(Dual a, Dual b) {
  double q = a.v / b.v;
  return Dual(q, (a.d - q * b.d) / b.v);
}
__host__ __device__ inline Dual operator+(Dual a, double b) { return Dual(a.v + b, a.d); }
__host__ __device__ inline Dual operator+(double a, Dual b) { return Dual(a + b.v, b.d); }
__host__ __device__ inline Dual operator-(Dual a, double b) { return Dual(a.v - b, a.d); }
__host__ __device__ inline Dual operator-(double a, Dual b) { return Dual(a - b.v, -b.d); }
__host__ __device__ inline Dual operator*(Dual a, double b) { return Dual(a.v * b, a.d * b); }
__host__ __device__ inline Dual operator*(double a, Dual b) { return Dual(a * b.v, a * b.d); }
__host__ __device__ inline Dual operator/(Dual a, double b) { return Dual(a.v / b, a.d / b); }
__host__ __device__ inline Dual operator/(double a, Dual b) {
  double q = a / b.v;
  return Dual(q, -q * b.d / b.v);
}
__host__ __device__ inline Dual& operator+=(Dual& a, Dual b) { a = a + b; return a; }
__host__ __device__ inline Dual& operator-=(Dual& a, Dual b) { a = a - b; return a; }
__host__ __device__ inline Dual& operator*=(Dual& a, Dual b) { a = a * b; return a; }

__host__ __device__ inline Dual sqrt(Dual a) {
  double s = ::sqrt(a.v);
  return Dual(s, 0.5 * a.d / s);
}
__host__ __device__ inline Dual exp(Dual a) {
  double e = ::exp(a.v);
  return Dual(e, e * a.d);
}
__host__ __device__ inline Dual acos(Dual a) {
  return Dual(::acos(a.v), -a.d / ::sqrt(1.0 - a.v * a.v));
}

__host__ __device__ inline double value(double a) { return a; }
__host__ __device__ inline double value(Dual a) { return a.v; }
__host__ __device__ inline double tangent(double) { return 0.0; }
__host__ __device__ inline double tangent(Dual a) { return a.d; }

}  // namespace rp
