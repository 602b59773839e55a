// btpm_device.h -- device functions shared by the conduction paths.
#pragma once

#include <hip/hip_runtime.h>

namespace btpm_dev {

// Newton root of the surface energy balance in FP64,
//   g(T) = src + m1 (t1 - T) - es T^4 = K - m1 T - es T^4,  K = src + m1 t1,
//   T <- T + g / (m1 + 4 es T^3),
// stopped when |dT| < tol.  g' is inverted with v_rcp_f64 alone: an
// approximate reciprocal only scales the step by (1 + e), which leaves the
// root (fixed by g, evaluated in FP64) unchanged and moves the returned
// iterate by e |last step| (< 1e-10 K); g is concave and decreasing, so the
// iteration converges from any positive start.  sbc_newton below is the
// whole solve; its two pieces serve callers that interleave the first
// iteration with other work.

// One Newton iteration from x (updated in place); returns the step.
__device__ __forceinline__ double newton_step(double K, double m1, double es, double& x) {
  const double x3 = (x * x) * x;
  const double ex3 = es * x3;
  const double g = fma(-ex3, x, fma(-m1, x, K));
  const double gp = fma(4.0, ex3, m1);
  const double d = g * __builtin_amdgcn_rcp(gp);
  x += d;
  return d;
}

// Iterations after a first step d0 (callers that peel the first iteration
// to schedule it among other work): continues while |d| >= tol, at most
// cap iterations in all.
__device__ __forceinline__ double newton_finish(double K, double m1, double es, double x,
                                                double d0, double tol, int cap) {
  if (!(fabs(d0) < tol))
    for (int it = 1; it < cap; ++it)
      if (fabs(newton_step(K, m1, es, x)) < tol) break;
  return x;
}

__device__ __forceinline__ double sbc_newton(double src, double m1, double t1, double es,
                                             double x, double tol, int cap) {
  const double K = fma(m1, t1, src);
  const double d0 = newton_step(K, m1, es, x);
  return newton_finish(K, m1, es, x, d0, tol, cap);
}

}  // namespace btpm_dev
