// observe.hip -- the Observe stage's flux synthesis on the GPU: the thermal
// infrared flux of the smooth and the rough model at the observer, per time
// step and band, from the Thermal stage's temperatures and the Visibility
// stage's masks (175706/Run-04/observe_generator.cpp, asteroidgeometry
// :530-698, surfacegeometry :465-529, observe :699-836).
//
//   btpm_obs_geom  one (step, facet) per lane: the vertices turned by their
//                  body's rotation and placed at the heliocentric position in
//                  metres, the observer vector, unit normal, facet frame, the
//                  view vector in that frame, its length and the view angle's
//                  cosine; facets seen beyond 90 degrees are invisible.  The
//                  step ntau-1 pass also stores each facet's area (the
//                  reference's flux uses the area of its last geometry step).
//   btpm_obs_flux  one wave per (facet, 16 steps): per step the crater
//                  sub-facets across the lanes (Planck intensity x emissivity
//                  x area x visibility x cos(view angle) / d^2 in three bands),
//                  summed over the wave, scaled by A_facet / S_crater; the
//                  smooth facet term; both written per facet, band and step.
//   btpm_obs_sum   64 steps per workgroup, 16 facet ranges summed in order
//                  and combined in order; the colour temperature fit (100..566 K, band ratios,
//                  the table split over the 16 waves, first minimum kept; the
//                  smooth fit keeps the reference's comma-operator terms,
//                  :782, i.e. fits band 1 alone) and the colour corrections.
//
// FP64 throughout, contraction off and the reference's operation order; the
// rotation trigonometry, pole matrix, positions, Planck prefactors and the
// colour-fit table come from the host (the same libm as the restatement), so
// what differs from it is the device's exp / acos / cos (ulps) and the
// order of the crater and facet sums: fluxes agree to ~1e-15 relative.
#pragma clang fp contract(off)

#include <hip/hip_runtime.h>

#include "btpm_internal.h"
#include "btpm_kernels.h"

namespace {
constexpr int kObsRec = 6;   // per (step, facet): view vector[3], |view|, cos(view angle), visible
constexpr int kObsSteps = 16;
}

extern "C" __global__ __launch_bounds__(256) void btpm_obs_geom(
    const double* __restrict__ equpos, const int* __restrict__ fv, int nf, int nv_primary,
    const double* __restrict__ trig, const double* __restrict__ frame, int s0, int nsteps,
    int ntau, const int* __restrict__ visible, double* __restrict__ rec,
    double* __restrict__ area) {
  const int f = blockIdx.x * blockDim.x + threadIdx.x;
  const int sl = blockIdx.y;
  if (f >= nf || sl >= nsteps) return;
  const int s = s0 + sl;
  double p[3][3];
#pragma unroll
  for (int i = 0; i < 3; ++i) {
    const int v = fv[3 * f + i];
    const double* t = trig + 4 * (size_t)s + (v < nv_primary ? 0 : 2);
    const double c = t[0], sn = t[1];
    const double* e = equpos + 3 * (size_t)v;
    const double q0 = (e[0] * c) - (e[1] * sn);
    const double q1 = (e[0] * sn) + (e[1] * c);
    const double q2 = e[2];
#pragma unroll
    for (int k = 0; k < 3; ++k)
      p[i][k] = frame[9 + k] +
                ((frame[3 * k] * q0) + (frame[3 * k + 1] * q1) + (frame[3 * k + 2] * q2));
  }
  double ov[3], a[3], b[3], n[3], y[3];
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    ov[k] = ((p[0][k] + p[1][k] + p[2][k]) / 3.0) - frame[12 + k];
    a[k] = p[1][k] - p[0][k];
    b[k] = p[2][k] - p[0][k];
  }
  const double o = sqrt((ov[0] * ov[0]) + (ov[1] * ov[1]) + (ov[2] * ov[2]));
  n[0] = (a[1] * b[2]) - (a[2] * b[1]);
  n[1] = (a[2] * b[0]) - (a[0] * b[2]);
  n[2] = (a[0] * b[1]) - (a[1] * b[0]);
  const double nl = sqrt((n[0] * n[0]) + (n[1] * n[1]) + (n[2] * n[2]));
  if (s == ntau - 1) area[f] = nl / 2.0;
#pragma unroll
  for (int k = 0; k < 3; ++k) n[k] = n[k] / nl;
  const double x = sqrt((a[0] * a[0]) + (a[1] * a[1]) + (a[2] * a[2]));
#pragma unroll
  for (int k = 0; k < 3; ++k) a[k] = a[k] / x;
  y[0] = (n[1] * a[2]) - (n[2] * a[1]);
  y[1] = (n[2] * a[0]) - (n[0] * a[2]);
  y[2] = (n[0] * a[1]) - (n[1] * a[0]);
  const double v0 = -1.0 * ((a[0] * ov[0]) + (a[1] * ov[1]) + (a[2] * ov[2]));
  const double v1 = -1.0 * ((y[0] * ov[0]) + (y[1] * ov[1]) + (y[2] * ov[2]));
  const double v2 = -1.0 * ((n[0] * ov[0]) + (n[1] * ov[1]) + (n[2] * ov[2]));
  const double c = v2 / o;
  // acos(c) > pi/2 exactly for -1 <= c <= BTPM_OBS_COS_BEHIND (glibc's
  // acos; NaN beyond -1 fails the comparison in the reference)
  const bool behind = c <= BTPM_OBS_COS_BEHIND && c >= -1.0;
  double* r = rec + ((size_t)sl * nf + f) * kObsRec;
  r[0] = v0;
  r[1] = v1;
  r[2] = v2;
  r[3] = sqrt((v0 * v0) + (v1 * v1) + (v2 * v2));
  r[4] = cos(acos(c));
  r[5] = (visible[(size_t)f * ntau + s] != 0 && !behind) ? (double)visible[(size_t)f * ntau + s]
                                                          : 0.0;
}

// pl[3*i] = 2hc^2/lambda^5, pl[3*i+1] = h c, pl[3*i+2] = lambda k (band i)
__device__ __forceinline__ double planck(const double* pl, int i, double T) {
  return pl[3 * i] * (1.0 / (exp(pl[3 * i + 1] / (pl[3 * i + 2] * T)) - 1.0));
}

// csub[j] = (unit normal[3], area) of crater sub-facet j; out[f][k][s]:
// k = 0..2 smooth bands, 3..5 rough bands (before summing over facets)
extern "C" __global__ __launch_bounds__(64) void btpm_obs_flux(
    const double* __restrict__ rec, const double* __restrict__ area,
    const float* __restrict__ stemp, const float* __restrict__ cvis,
    const float* __restrict__ ctemp, const double4* __restrict__ csub, int ncf, double stot,
    int rough, const double* __restrict__ pl, double eps, int nf, int ntau, int s0, int nsteps,
    double* __restrict__ out) {
  const int f = blockIdx.x;
  const int lane = threadIdx.x;
  const double af = area[f];
  const int sl0 = blockIdx.y * kObsSteps;
  for (int sl = sl0; sl < min(sl0 + kObsSteps, nsteps); ++sl) {
    const int s = s0 + sl;
    const double* r = rec + ((size_t)sl * nf + f) * kObsRec;
    const double vis = r[5];
    double sm[3] = {0.0, 0.0, 0.0}, ro[3] = {0.0, 0.0, 0.0};
    if (vis != 0.0) {
      const double w0 = r[0], w1 = r[1], w2 = r[2], d = r[3];
      if (rough) {
        double sf[3] = {0.0, 0.0, 0.0};
        const size_t base = ((size_t)f * ntau + s) * ncf;
        for (int j = lane; j < ncf; j += 64) {
          const double cv = (double)cvis[base + j];
          if (!(cv > 0.0)) continue;
          const double4 cn = csub[j];
          const double va = acos(((w0 * cn.x) + (w1 * cn.y) + (w2 * cn.z)) / d);
          const double T = (double)ctemp[base + j];
#pragma unroll
          for (int i = 0; i < 3; ++i) {
            double it = planck(pl, i, T);
            double mu = (eps * cn.w * cv * cos(va)) / (d * d);
            if (it != it) it = 0.0;
            if (mu != mu) mu = 0.0;
            sf[i] = sf[i] + (it * mu);
          }
        }
#pragma unroll
        for (int i = 0; i < 3; ++i) {
          double v = sf[i];
          for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
          ro[i] = v * af / stot;
        }
      }
      const double T = (double)stemp[(size_t)f * ntau + s];
#pragma unroll
      for (int i = 0; i < 3; ++i) {
        double it = planck(pl, i, T);
        double mu = (eps * af * vis * r[4]) / (d * d);
        if (it != it) it = 0.0;
        if (mu != mu) mu = 0.0;
        sm[i] = it * mu;
      }
    }
    if (lane == 0) {
      double* o = out + (size_t)f * 6 * nsteps + sl;
#pragma unroll
      for (int i = 0; i < 3; ++i) {
        o[(size_t)i * nsteps] = sm[i];
        o[(size_t)(3 + i) * nsteps] = ro[i];
      }
    }
  }
}

// bn[t][3] = the band intensities of a blackbody at tmin + t, divided by
// their mean; colour[3][tcount].  A workgroup per 64 steps (lane = step);
// its 16 waves sum 16 consecutive facet ranges in facet order, the ranges'
// sums are then combined in order.
extern "C" __global__ __launch_bounds__(1024) void btpm_obs_sum(
    const double* __restrict__ terms, int nf, int nsteps, const double* __restrict__ bn,
    const double* __restrict__ colour, int tmin, int tcount, double* __restrict__ smooth,
    double* __restrict__ rough) {
  __shared__ double part[16][6][64];
  __shared__ double fchi[2][16][64];
  __shared__ int fk[2][16][64];
  const int lane = threadIdx.x & 63, seg = threadIdx.x >> 6;
  const int sl = blockIdx.x * 64 + lane;
  const int per = (nf + 15) / 16;
  const int f0 = seg * per, f1 = min(nf, f0 + per);
  double acc[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
  if (sl < nsteps) {
    for (int f = f0; f < f1; ++f) {
      const double* t = terms + (size_t)f * 6 * nsteps + sl;
#pragma unroll
      for (int k = 0; k < 6; ++k) acc[k] = acc[k] + t[(size_t)k * nsteps];
    }
  }
#pragma unroll
  for (int k = 0; k < 6; ++k) part[seg][k][lane] = acc[k];
  __syncthreads();
  // every wave forms the step's band fluxes (the 16 ranges in order) ...
  double sm[3] = {0.0, 0.0, 0.0}, ro[3] = {0.0, 0.0, 0.0};
  for (int g = 0; g < 16; ++g)
#pragma unroll
    for (int i = 0; i < 3; ++i) {
      sm[i] = sm[i] + part[g][i][lane];
      ro[i] = ro[i] + part[g][3 + i][lane];
    }
  // ... and scans its own range of colour temperatures: the first minimum of
  // the range (strict <, from 100.0 like the full scan); squares as products,
  // the correctly rounded value of pow(x, 2.0)
  const int kper = (tcount + 15) / 16;
  const int k0 = seg * kper, k1 = min(tcount, k0 + kper);
  double cs_s = 100.0, cs_r = 100.0;
  int bt_s = -1, bt_r = -1;
  {
    double am = (sm[0] + sm[1] + sm[2]) / 3.0;
    const double s2 = sm[0] / am;
    am = (ro[0] + ro[1] + ro[2]) / 3.0;
    const double r2 = ro[0] / am, r3 = ro[1] / am, r4 = ro[2] / am;
    for (int k = k0; k < k1; ++k) {
      const double b0 = bn[3 * k], b1 = bn[3 * k + 1], b2 = bn[3 * k + 2];
      const double e0 = (b0 - s2) / s2;
      double chi = e0 * e0 + 2.0 + 2.0;
      if (chi < cs_s) {
        cs_s = chi;
        bt_s = k;
      }
      const double d0 = (b0 - r2) / r2, d1 = (b1 - r3) / r3, d2 = (b2 - r4) / r4;
      chi = d0 * d0 + d1 * d1 + d2 * d2;
      if (chi < cs_r) {
        cs_r = chi;
        bt_r = k;
      }
    }
  }
  fchi[0][seg][lane] = cs_s;
  fchi[1][seg][lane] = cs_r;
  fk[0][seg][lane] = bt_s;
  fk[1][seg][lane] = bt_r;
  __syncthreads();
  if (seg != 0 || sl >= nsteps) return;
  // ranges in order with strict <: the first minimum of the whole table
  cs_s = 100.0;
  cs_r = 100.0;
  bt_s = 0;
  bt_r = 0;
  for (int g = 0; g < 16; ++g) {
    if (fchi[0][g][lane] < cs_s) {
      cs_s = fchi[0][g][lane];
      bt_s = fk[0][g][lane];
    }
    if (fchi[1][g][lane] < cs_r) {
      cs_r = fchi[1][g][lane];
      bt_r = fk[1][g][lane];
    }
  }
#pragma unroll
  for (int i = 0; i < 3; ++i) {
    smooth[3 * (size_t)sl + i] = sm[i] * colour[(size_t)i * tcount + bt_s];
    rough[3 * (size_t)sl + i] = ro[i] * colour[(size_t)i * tcount + bt_r];
  }
}

// ===========================================================================
#include <algorithm>
#include <cmath>
#include <vector>

#include "runtime.h"

using namespace btpm;

namespace {
// Where btpm_observe_solve takes the crater visibility from: with obs_facet
// (the observer's direction in each facet frame, [nf][ntau][3]) it is
// computed on the device, the crater sunlight kernel with the observer as the
// source and in->visible as the facet flags, with m x m rim samples per
// sub-facet and the crater radius.
struct ObsVis {
  const float* obs_facet = nullptr;
  int m = 0;
  float radius = 0.0f;
};

// btpm_observe and btpm_observe_solve: the temperatures come from the host
// arrays of `in`, or with `solve` from that finished solve's device memory
int observe_run(btpm_ctx* ctx, const btpm_observe_in* in, btpm_solve* solve, const ObsVis& ov,
                double* smooth, double* rough) {
  if (!ctx || !in || !smooth || !rough || !in->equpos || !in->facets || !in->visible ||
      (!solve && !in->stemp) || !in->colour || !in->wavelength_um)
    return fail(BTPM_ERR_ARG, "null argument");
  if (in->nv < 3 || in->nf < 1 || in->ntau < 1 || in->nv_primary < 0 || in->nv_primary > in->nv)
    return fail(BTPM_ERR_ARG, "bad sizes: nv %d nf %d ntau %d nv_primary %d", in->nv, in->nf,
                in->ntau, in->nv_primary);
  if (in->tcount < 1 || in->tmin < 1) return fail(BTPM_ERR_ARG, "bad colour table");
  if (!(in->hel[0] > 0.0) || !(in->obs[0] > 0.0) || !(in->au_m > 0.0))
    return fail(BTPM_ERR_ARG, "bad positions");
  for (int i = 0; i < 3 * in->nf; ++i)
    if (in->facets[i] < 0 || in->facets[i] >= in->nv)
      return fail(BTPM_ERR_ARG, "facet %d: vertex %d out of range", i / 3, in->facets[i]);
  const int rough_on = in->rough != 0;
  if (rough_on) {
    if (!in->crater_vtx || !in->crater_facets || (!ov.obs_facet && !in->crater_visible) ||
        (!solve && !in->crater_temp))
      return fail(BTPM_ERR_ARG, "null crater arrays");
    if (ov.obs_facet && (in->ncf > 128 || ov.m < 1 || ov.m > 64 || !(ov.radius > 0.0f)))
      return fail(BTPM_ERR_ARG, "crater visibility: ncf %d, m %d, radius %g", in->ncf, ov.m,
                  (double)ov.radius);
    if (in->ncv < 3 || in->ncf < 1) return fail(BTPM_ERR_ARG, "bad crater sizes");
    if (solve && in->ncf != BTPM_NS)
      return fail(BTPM_ERR_ARG, "the solve's crater model has %d facets, not %d", BTPM_NS,
                  in->ncf);
    for (int i = 0; i < 3 * in->ncf; ++i)
      if (in->crater_facets[i] < 0 || in->crater_facets[i] >= in->ncv)
        return fail(BTPM_ERR_ARG, "crater facet %d: vertex out of range", i / 3);
  }
  CtxLock lk(ctx);
  API_BEGIN
  HIPCHK(hipSetDevice(ctx->device));
  hipStream_t st = ctx->stream;
  const int nv = in->nv, nf = in->nf, ntau = in->ntau, ncf = rough_on ? in->ncf : 0;
  // host: the reference's trigonometry and constants (observe_generator.cpp:
  // 8-14, 531-549, 557-575), with the same libm as the restatement
  const double PI = 4.0 * std::atan(1.0);
  const double hcon = 6.626068963e-34, cvel = 299792458.0, kcon = 1.380650424e-23;
  double frame[15];
  {
    const double r0 = in->rot[0], r1 = in->rot[1];
    const double gamma = std::acos(std::cos(r0) * std::sin(std::fabs(r1)));
    double* tm = frame;
    tm[0] = std::sin(r0) / std::sin(gamma);
    tm[1] = -1.0 * std::cos(r0) * std::cos(r0) * std::sin(r1) * std::cos(r1) / std::sin(gamma);
    tm[2] = std::cos(r0) * std::cos(r1);
    tm[3] = 0.0;
    tm[4] = ((std::sin(r0) * std::sin(r0)) +
             (std::cos(r0) * std::cos(r0) * std::cos(r1) * std::cos(r1))) / std::sin(gamma);
    tm[5] = std::cos(r0) * std::sin(r1);
    tm[6] = -1.0 * std::cos(r0) * std::cos(r1) / std::sin(gamma);
    tm[7] = -1.0 * std::sin(r0) * std::cos(r0) * std::sin(r1) / std::sin(gamma);
    tm[8] = std::sin(r0);
    const double* h = in->hel;
    const double* ob = in->obs;
    frame[9] = in->au_m * h[0] * std::cos(h[1]) * std::cos(h[2]);
    frame[10] = in->au_m * h[0] * std::cos(h[1]) * std::sin(h[2]);
    frame[11] = in->au_m * h[0] * std::sin(h[1]);
    frame[12] = frame[9] - (in->au_m * ob[0] * std::cos(ob[1]) * std::cos(ob[2]));
    frame[13] = frame[10] - (in->au_m * ob[0] * std::cos(ob[1]) * std::sin(ob[2]));
    frame[14] = frame[11] - (in->au_m * ob[0] * std::sin(ob[1]));
  }
  std::vector<double> trig(4 * (size_t)ntau);
  for (int s = 0; s < ntau; ++s) {
    const double ap = 2.0 * PI * s * in->rev_primary / ntau;
    const double as = 2.0 * PI * s * in->rev_secondary / ntau;
    trig[4 * (size_t)s] = std::cos(ap);
    trig[4 * (size_t)s + 1] = std::sin(ap);
    trig[4 * (size_t)s + 2] = std::cos(as);
    trig[4 * (size_t)s + 3] = std::sin(as);
  }
  double pl[9];
  for (int i = 0; i < 3; ++i) {
    const double lam = in->wavelength_um[i] * 1.0e-6;
    pl[3 * i] = (2.0 * hcon * cvel * cvel) / std::pow(lam, 5.0);
    pl[3 * i + 1] = hcon * cvel;
    pl[3 * i + 2] = lam * kcon;
  }
  std::vector<double> bn(3 * (size_t)in->tcount);
  for (int k = 0; k < in->tcount; ++k) {
    double b[3];
    for (int i = 0; i < 3; ++i)
      b[i] = pl[3 * i] * (1.0 / (std::exp(pl[3 * i + 1] / (pl[3 * i + 2] * (double)(in->tmin + k))) -
                                 1.0));
    const double ab = (b[0] + b[1] + b[2]) / 3.0;
    for (int i = 0; i < 3; ++i) bn[3 * (size_t)k + i] = b[i] / ab;
  }
  // crater sub-facets (surfacegeometry :480-492)
  std::vector<double4> csub(std::max(1, ncf));
  double stot = 0.0;
  for (int j = 0; j < ncf; ++j) {
    double p[3][3], a[3], b[3], n[3];
    for (int v = 0; v < 3; ++v)
      for (int k = 0; k < 3; ++k)
        p[v][k] = (double)in->crater_vtx[3 * (size_t)in->crater_facets[3 * j + v] + k];
    for (int k = 0; k < 3; ++k) {
      a[k] = p[1][k] - p[0][k];
      b[k] = p[2][k] - p[0][k];
    }
    n[0] = (a[1] * b[2]) - (a[2] * b[1]);
    n[1] = (a[2] * b[0]) - (a[0] * b[2]);
    n[2] = (a[0] * b[1]) - (a[1] * b[0]);
    const double l = std::sqrt((n[0] * n[0]) + (n[1] * n[1]) + (n[2] * n[2]));
    const double ar = l / 2.0;
    csub[j] = make_double4(n[0] / l, n[1] / l, n[2] / l, ar);
    stot = stot + (ar * (n[2] / l));
  }
  const size_t nser = (size_t)nf * ntau;
  DBuf d_eq, d_fv, d_trig, d_frame, d_vis, d_st, d_cvis, d_ct, d_cs, d_pl, d_bn, d_col, d_area;
  upload(d_eq, in->equpos, 3 * (size_t)nv, st);
  upload(d_fv, in->facets, 3 * (size_t)nf, st);
  upload(d_trig, trig.data(), trig.size(), st);
  upload(d_frame, frame, 15, st);
  upload(d_vis, in->visible, nser, st);
  if (rough_on) {  // the large inputs (4 B per sub-facet and step each) through pinned slices
    d_cvis.alloc(nser * ncf * sizeof(float));
    if (ov.obs_facet)
      crater_fractions_on_device(ctx, in->visible, ov.obs_facet, (long long)nser, in->crater_vtx,
                                 in->crater_facets, ncf, ov.m, ov.radius, d_cvis.as<float>());
    else
      upload_staged(ctx, d_cvis.p, in->crater_visible, nser * ncf * sizeof(float));
    d_ct.alloc(nser * ncf * sizeof(float));
  }
  if (solve) {  // the temperatures never leave the device
    d_st.alloc(nser * sizeof(float));
    const int rc = solve_temps_on_device(solve, ctx, nf, ntau, d_st.as<float>(),
                                         rough_on ? d_ct.as<float>() : nullptr);
    if (rc != BTPM_OK) return rc;
  } else {
    upload(d_st, in->stemp, nser, st);
    if (rough_on) upload_staged(ctx, d_ct.p, in->crater_temp, nser * ncf * sizeof(float));
  }
  upload(d_cs, csub.data(), csub.size(), st);
  upload(d_pl, pl, 9, st);
  upload(d_bn, bn.data(), bn.size(), st);
  upload(d_col, in->colour, 3 * (size_t)in->tcount, st);
  d_area.alloc((size_t)nf * sizeof(double));
  // the last step's areas first
  DBuf d_rec, d_terms, d_sm, d_ro;
  // steps per chunk: geometry records and facet terms of at most ~512 MB
  const size_t per_step = (size_t)nf * (kObsRec + 6) * sizeof(double);
  const int sc = (int)std::max<size_t>(1, std::min<size_t>({(size_t)ntau, (512u << 20) / per_step,
                                                             (size_t)65535}));
  d_rec.alloc((size_t)sc * nf * kObsRec * sizeof(double));
  d_terms.alloc((size_t)sc * nf * 6 * sizeof(double));
  d_sm.alloc(3 * (size_t)ntau * sizeof(double));
  d_ro.alloc(3 * (size_t)ntau * sizeof(double));
  hipLaunchKernelGGL(btpm_obs_geom, dim3(grid1(nf, 256), 1), dim3(256), 0, st, d_eq.as<double>(),
                     d_fv.as<int>(), nf, in->nv_primary, d_trig.as<double>(),
                     d_frame.as<double>(), ntau - 1, 1, ntau, d_vis.as<int>(), d_rec.as<double>(),
                     d_area.as<double>());
  for (int s0 = 0; s0 < ntau; s0 += sc) {
    const int ns = std::min(sc, ntau - s0);
    hipLaunchKernelGGL(btpm_obs_geom, dim3(grid1(nf, 256), ns), dim3(256), 0, st,
                       d_eq.as<double>(), d_fv.as<int>(), nf, in->nv_primary, d_trig.as<double>(),
                       d_frame.as<double>(), s0, ns, ntau, d_vis.as<int>(), d_rec.as<double>(),
                       d_area.as<double>());
    hipLaunchKernelGGL(btpm_obs_flux, dim3(nf, grid1(ns, kObsSteps)), dim3(64), 0, st,
                       d_rec.as<double>(), d_area.as<double>(), d_st.as<float>(),
                       rough_on ? d_cvis.as<float>() : (const float*)nullptr,
                       rough_on ? d_ct.as<float>() : (const float*)nullptr, d_cs.as<double4>(),
                       ncf, stot, rough_on, d_pl.as<double>(), in->eps, nf, ntau, s0, ns,
                       d_terms.as<double>());
    hipLaunchKernelGGL(btpm_obs_sum, dim3(grid1(ns, 64)), dim3(1024), 0, st,
                       d_terms.as<double>(), nf, ns, d_bn.as<double>(), d_col.as<double>(),
                       in->tmin, in->tcount, d_sm.as<double>() + 3 * (size_t)s0,
                       d_ro.as<double>() + 3 * (size_t)s0);
    HIPCHK(hipGetLastError());
  }
  HIPCHK(hipMemcpyAsync(smooth, d_sm.p, 3 * (size_t)ntau * sizeof(double), hipMemcpyDeviceToHost,
                        st));
  HIPCHK(hipMemcpyAsync(rough, d_ro.p, 3 * (size_t)ntau * sizeof(double), hipMemcpyDeviceToHost,
                        st));
  HIPCHK(hipStreamSynchronize(st));
  API_END
}
}  // namespace

extern "C" int btpm_observe(btpm_ctx* ctx, const btpm_observe_in* in, double* smooth,
                            double* rough) {
  return observe_run(ctx, in, nullptr, ObsVis{}, smooth, rough);
}

extern "C" int btpm_observe_solve(btpm_ctx* ctx, const btpm_observe_in* in, btpm_solve* solve,
                                  const float* obs_facet, int m, float crater_radius,
                                  double* smooth, double* rough) {
  if (!solve) return fail(BTPM_ERR_ARG, "null argument");
  ObsVis ov;
  ov.obs_facet = obs_facet;
  ov.m = m;
  ov.radius = crater_radius;
  return observe_run(ctx, in, solve, ov, smooth, rough);
}
