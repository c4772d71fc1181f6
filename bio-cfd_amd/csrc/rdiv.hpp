// Correctly rounded division by a fixed divisor from its reciprocal.
//
// x / b from b and y = 1/b rounded once: q0 = x y is within two ulps of
// x / b; one residual correction q1 = q0 - (b q0 - x) y (the residual exact
// in one FMA) makes it faithful, and a second correction from a faithful
// quotient with the correctly rounded reciprocal gives x / b rounded to
// nearest (Markstein's theorem, the correction step of IEEE division by
// Newton-Raphson iteration with an FMA), bit for bit the division's result
// while x, x y and the residuals are normal numbers.  Zero keeps its sign:
// the residual is formed as b q - x and subtracted.
//
// Host builds of this header run it in tests against the division.
#pragma once

#include <cmath>

namespace bcfd {

__host__ __device__ inline double rdiv_core(double x, double b, double y) {
  const double q0 = x * y;
  const double q1 = std::fma(-std::fma(b, q0, -x), y, q0);
  return std::fma(-std::fma(b, q1, -x), y, q1);
}

}  // namespace bcfd
