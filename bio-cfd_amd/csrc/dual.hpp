// A value and one directional derivative (forward-mode differentiation), for
// the tangent-linear and adjoint passes: the stencil and reconstruction code
// is evaluated on these where it is generic in its scalar type.  |x| takes
// the derivative sign(x), and 0 at x = 0 (the value central differences give
// there).
#pragma once

#include <hip/hip_runtime.h>

namespace bcfd {

struct Dual {
  double v, t;
};
__device__ inline Dual operator+(Dual a, Dual b) { return {a.v + b.v, a.t + b.t}; }
__device__ inline Dual operator+(Dual a, double b) { return {a.v + b, a.t}; }
__device__ inline Dual operator+(double a, Dual b) { return {a + b.v, b.t}; }
__device__ inline Dual operator-(Dual a, Dual b) { return {a.v - b.v, a.t - b.t}; }
__device__ inline Dual operator-(Dual a, double b) { return {a.v - b, a.t}; }
__device__ inline Dual operator-(double a, Dual b) { return {a - b.v, -b.t}; }
__device__ inline Dual operator-(Dual a) { return {-a.v, -a.t}; }
__device__ inline Dual operator*(Dual a, Dual b) { return {a.v * b.v, a.v * b.t + a.t * b.v}; }
__device__ inline Dual operator*(double s, Dual a) { return {s * a.v, s * a.t}; }
__device__ inline Dual operator*(Dual a, double s) { return {a.v * s, a.t * s}; }
__device__ inline Dual operator/(Dual a, double s) { return {a.v / s, a.t / s}; }
__device__ inline Dual operator/(Dual a, Dual b) {
  const double q = a.v / b.v;
  return {q, (a.t - q * b.t) / b.v};
}
__device__ inline Dual operator/(double a, Dual b) {
  const double q = a / b.v;
  return {q, -q * b.t / b.v};
}
__device__ inline Dual fabs(Dual a) {
  return a.v < 0.0 ? Dual{-a.v, -a.t} : (a.v > 0.0 ? a : Dual{0.0, 0.0});
}
__device__ inline Dual sqrt(Dual a) {
  const double r = ::sqrt(a.v);
  return {r, a.t / (2.0 * r)};
}
using ::fabs;  // the double overloads stay visible beside the Dual ones
using ::sqrt;

__device__ inline double val_of(double x) { return x; }
__device__ inline double val_of(Dual x) { return x.v; }

}  // namespace bcfd
