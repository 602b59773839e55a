// btpm_device.h -- device functions shared by the conduction paths.
#pragma once

#include <hip/hip_runtime.h>

namespace btpm_dev {

// Newton root of the surface energy balance in FP64,
//   g(T) = src + m1 (t1 - T) - es T^4 = 0,   T <- T + g / (m1 + 4 es T^3),
// stopped when |dT| < tol.  g' is inverted with v_rcp_f64 plus one
// Newton-Raphson refinement of the reciprocal (relative error ~1e-16, so the
// iterates are the exact-division iterates to rounding); g is concave and
// decreasing, so the iteration converges from any positive start.
__device__ __forceinline__ double sbc_newton(double src, double m1, double t1, double es,
                                             double x, double tol, int cap) {
  const double es4 = 4.0 * es;
  for (int it = 0; it < cap; ++it) {
    const double x2 = x * x;
    const double x3 = x2 * x;
    const double g = fma(-es * x3, x, fma(m1, t1 - x, src));
    const double gp = fma(es4, x3, m1);
    double rg = __builtin_amdgcn_rcp(gp);
    rg = fma(rg, fma(-gp, rg, 1.0), rg);
    const double d = g * rg;
    x += d;
    if (fabs(d) < tol) break;
  }
  return x;
}

}  // namespace btpm_dev
