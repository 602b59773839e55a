// geometry.hip -- illumination geometry and shadows of a (binary) shape on
// the GPU: the producer of the Thermal stage's direct facet illumination.
//
// Per time step (T175706/thermal.cpp:1272-1471, asteroidgeometry): every
// vertex turned by its body's rotation angle, taken to the ecliptic frame and
// placed at the system's heliocentric position; per facet the centroid, unit
// normal, local frame and the angle of the Sun; back-facing facets unlit.
// Then (astshadowtest / astshadow, :1612-1707) the ray from the Sun through
// every lit facet's centroid against the plane and edges of every facet
// nearer the Sun; and the direct term solar cos(angle) shadow (:1757).
//
// The shadow flags are integers decided by signs of FP64 expressions, so the
// kernels evaluate every expression in the reference's operation order with
// no fused multiply-adds (the pragma below), and the trigonometry (rotation
// angles, the pole matrix, the position) comes from the host, evaluated by
// the same libm as the reference: the remaining operations (+ - * / sqrt)
// are IEEE correctly rounded, so the flags are those of the FP64 restatement
// bit for bit.
#pragma clang fp contract(off)

#include <hip/hip_runtime.h>

#include "btpm_internal.h"
#include "btpm_kernels.h"

namespace {
constexpr int kGR = BTPM_GEOM_REC;  // doubles per facet record
// record layout: centroid[3], |centroid|, circumradius, centroid . normal,
// normal[3], vertices[3][3], index (as double), pad
constexpr int kMid = 0, kDist = 3, kRad = 4, kTc1 = 5, kNrm = 6, kVtx = 9, kIdx = 18;
constexpr int kShChunk = 64;        // nearer-facet records staged per LDS chunk
}  // namespace

// One (step, facet) per lane: the facet record of the step's chunk slot, the
// back-face flag, cos(angle) and the unit Sun direction in the facet frame.
//   trig[s] = (cos, sin) of the primary's and the secondary's rotation angle
//   frame[0..8] = ecliptic <- equatorial matrix (row-major), frame[9..11] =
//   heliocentric position of the system (km)
extern "C" __global__ __launch_bounds__(256) void btpm_geom_facets(
    const double* __restrict__ equpos, const int* __restrict__ fv, int nf, int nv_primary,
    const double* __restrict__ trig, const double* __restrict__ frame, int s0, int nsteps,
    double* __restrict__ rec, int* __restrict__ lit, float* __restrict__ cosang,
    float* __restrict__ sun, double sun_scale) {
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
  double mid[3], a[3], b[3], n[3], y[3];
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    mid[k] = (p[0][k] + p[1][k] + p[2][k]) / 3.0;
    a[k] = p[1][k] - p[0][k];
    b[k] = p[2][k] - p[0][k];
  }
  const double t = sqrt((mid[0] * mid[0]) + (mid[1] * mid[1]) + (mid[2] * mid[2]));
  n[0] = (a[1] * b[2]) - (a[2] * b[1]);
  n[1] = (a[2] * b[0]) - (a[0] * b[2]);
  n[2] = (a[0] * b[1]) - (a[1] * b[0]);
  const double nn = sqrt((n[0] * n[0]) + (n[1] * n[1]) + (n[2] * n[2]));
#pragma unroll
  for (int k = 0; k < 3; ++k) n[k] = n[k] / nn;
  const double xl = sqrt((a[0] * a[0]) + (a[1] * a[1]) + (a[2] * a[2]));
#pragma unroll
  for (int k = 0; k < 3; ++k) a[k] = a[k] / xl;
  y[0] = (n[1] * a[2]) - (n[2] * a[1]);
  y[1] = (n[2] * a[0]) - (n[0] * a[2]);
  y[2] = (n[0] * a[1]) - (n[1] * a[0]);
  const double iv0 = -1.0 * ((a[0] * mid[0]) + (a[1] * mid[1]) + (a[2] * mid[2]));
  const double iv1 = -1.0 * ((y[0] * mid[0]) + (y[1] * mid[1]) + (y[2] * mid[2]));
  const double iv2 = -1.0 * ((n[0] * mid[0]) + (n[1] * mid[1]) + (n[2] * mid[2]));
  const double ang = acos(iv2 / t);
  const double kPi = 4.0 * 0.78539816339744830962;  // = 4 atan(1) exactly in double
  const size_t o = (size_t)sl * nf + f;
  lit[o] = ang > (kPi / 2.0) ? 0 : 1;
  cosang[o] = (float)cos(ang);
  if (sun_scale > 0.0) {  // thermal.cpp:1453-1455
    sun[3 * o] = (float)(sun_scale * iv0 / t);
    sun[3 * o + 1] = (float)(sun_scale * iv1 / t);
    sun[3 * o + 2] = (float)(sun_scale * iv2 / t);
  } else {
    sun[3 * o] = (float)(iv0 / t);
    sun[3 * o + 1] = (float)(iv1 / t);
    sun[3 * o + 2] = (float)(iv2 / t);
  }
  // record for the shadow tests
  double r2 = 0.0;
#pragma unroll
  for (int i = 0; i < 3; ++i) {
    const double d0 = p[i][0] - mid[0], d1 = p[i][1] - mid[1], d2 = p[i][2] - mid[2];
    r2 = fmax(r2, (d0 * d0) + (d1 * d1) + (d2 * d2));
  }
  double* r = rec + o * kGR;
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    r[kMid + k] = mid[k];
    r[kNrm + k] = n[k];
    r[kVtx + k] = p[0][k];
    r[kVtx + 3 + k] = p[1][k];
    r[kVtx + 6 + k] = p[2][k];
  }
  r[kDist] = t;
  r[kRad] = sqrt(r2);
  r[kTc1] = (mid[0] * n[0]) + (mid[1] * n[1]) + (mid[2] * n[2]);
  r[kIdx] = (double)f;
  r[kIdx + 1] = 0.0;
}

namespace {
// thermal.cpp:1612-1667 with H's record: true if the ray from the Sun
// through g meets H inside its triangle
__device__ __forceinline__ bool ray_hits(const double (&g)[3], const double* __restrict__ H) {
  const double tcalc2 = g[0] * H[kNrm] + g[1] * H[kNrm + 1] + g[2] * H[kNrm + 2];
  const double t = H[kTc1] / tcalc2;
  const double* v0 = H + kVtx;
  const double* v1 = H + kVtx + 3;
  const double* v2 = H + kVtx + 6;
  double x[3], e1[3], e2[3], e3[3], x1[3], x2[3], x3[3];
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    x[k] = t * g[k];
    e1[k] = v1[k] - v0[k];
    e2[k] = v0[k] - v2[k];
    e3[k] = v2[k] - v1[k];
    x1[k] = x[k] - v0[k];
    x3[k] = x[k] - v1[k];
    x2[k] = x[k] - v2[k];
  }
  const double* nH = H + kNrm;
  auto side = [&](const double (&e)[3], const double (&xv)[3]) {
    const double c0 = e[1] * xv[2] - e[2] * xv[1];
    const double c1 = e[2] * xv[0] - e[0] * xv[2];
    const double c2 = e[0] * xv[1] - e[1] * xv[0];
    return c0 * nH[0] + c1 * nH[1] + c2 * nH[2];
  };
  return side(e1, x1) >= 0.0 && side(e2, x2) >= 0.0 && side(e3, x3) >= 0.0;
}
}  // namespace

// Shadow flags of one step: a workgroup owns 256 facets G of a step (lane =
// G) and walks every facet H of the step in LDS chunks of 64 records.  H is
// tested against G if it is nearer the Sun (|H| < |G|, ties to the lower
// index: the pair order of thermal.cpp:1686-1701), and only if the Sun->G
// ray passes within H's circumradius + margin of H's centroid, measured in
// the body's frame (w = H - G, |w|^2 - (w.u)^2 with u the unit ray): the
// rounding of that distance and of the plane intersection is ~1e-16 |G|
// (1e-8 km at 1 AU), the margin 1e-6 km + 1e-14 |G|, so a skipped H is one
// whose triangle the ray misses by far more than the edge tests can resolve,
// and the flags equal testing every pair.
extern "C" __global__ __launch_bounds__(256) void btpm_geom_shadow(const double* __restrict__ rec,
                                                                   int nf, int nsteps,
                                                                   int* __restrict__ lit) {
  __shared__ double H[kShChunk][kGR];
  const int tid = (int)threadIdx.x;
  const int f = blockIdx.x * blockDim.x + tid;
  const int sl = blockIdx.y;
  if (sl >= nsteps) return;  // uniform
  const double* R = rec + (size_t)sl * nf * kGR;
  const bool in = f < nf;
  int on = 0;
  double g[3] = {0.0, 0.0, 0.0}, u[3] = {0.0, 0.0, 0.0}, gd = 0.0;
  if (in) {
    on = lit[(size_t)sl * nf + f];
    const double* G = R + (size_t)f * kGR;
    g[0] = G[kMid];
    g[1] = G[kMid + 1];
    g[2] = G[kMid + 2];
    gd = G[kDist];
    u[0] = g[0] / gd;
    u[1] = g[1] / gd;
    u[2] = g[2] / gd;
  }
  if (__syncthreads_count(on) == 0) return;
  // skip margin: well above the ~1e-16 |G| rounding of w and of the plane
  // intersection, well below any facet
  const double margin = 1e-6 + 1e-14 * gd;
  for (int c0 = 0; c0 < nf; c0 += kShChunk) {
    const int nc = min(kShChunk, nf - c0);
    __syncthreads();
    for (int e = tid; e < nc * kGR; e += (int)blockDim.x)
      (&H[0][0])[e] = R[(size_t)c0 * kGR + e];
    __syncthreads();
    if (!on) continue;
    for (int h = 0; h < nc; ++h) {
      const double hd = H[h][kDist];
      const int hi = c0 + h;
      if (!(hd < gd || (hd == gd && hi < f))) continue;
      const double w0 = H[h][kMid] - g[0], w1 = H[h][kMid + 1] - g[1], w2 = H[h][kMid + 2] - g[2];
      const double wu = w0 * u[0] + w1 * u[1] + w2 * u[2];
      const double d2 = (w0 * w0 + w1 * w1 + w2 * w2) - wu * wu;
      const double reach = H[h][kRad] + margin;
      if (d2 > reach * reach) continue;
      if (ray_hits(g, H[h])) {
        on = 0;
        break;
      }
    }
  }
  if (in) lit[(size_t)sl * nf + f] = on;
}

// [step][facet] staging of a chunk -> [facet][step] outputs, through a
// 64 x 64 LDS tile: shadow flags, illum = solar cos shadow, sun directions.
extern "C" __global__ __launch_bounds__(256) void btpm_geom_finish(
    const int* __restrict__ lit, const float* __restrict__ cosang, const float* __restrict__ sun,
    int nf, int nsteps, int s0, int ntau, float solar, int* __restrict__ shadow,
    float* __restrict__ illum, float* __restrict__ sun_out) {
  __shared__ int tl[64][65];
  __shared__ float tc[64][65];
  const int f0 = blockIdx.x * 64, q0 = blockIdx.y * 64;
  const int tx = (int)threadIdx.x & 63, ty = (int)threadIdx.x >> 6;  // 64 x 4
  for (int r = ty; r < 64; r += 4) {  // rows = steps, columns = facets
    const int q = q0 + r, f = f0 + tx;
    int l = 0;
    float c = 0.0f;
    if (q < nsteps && f < nf) {
      l = lit[(size_t)q * nf + f];
      c = cosang[(size_t)q * nf + f];
    }
    tl[r][tx] = l;
    tc[r][tx] = c;
  }
  __syncthreads();
  for (int r = ty; r < 64; r += 4) {  // rows = facets, columns = steps
    const int f = f0 + r, q = q0 + tx;
    if (f >= nf || q >= nsteps) continue;
    const int l = tl[tx][r];
    const size_t o = (size_t)f * ntau + s0 + q;
    shadow[o] = l;
    illum[o] = solar * tc[tx][r] * (float)l;
    if (sun_out) {
      const size_t i = (size_t)q * nf + f;
      sun_out[3 * o] = sun[3 * i];
      sun_out[3 * o + 1] = sun[3 * i + 1];
      sun_out[3 * o + 2] = sun[3 * i + 2];
    }
  }
}

// ===========================================================================
// host entry point
// ===========================================================================
#include <cmath>
#include <vector>

#include "runtime.h"

using namespace btpm;

extern "C" int btpm_geometry_illumination(btpm_ctx* ctx, const btpm_geometry* g, int* shadow,
                                          float* illum, float* sun_facet, float* solar) {
  if (!ctx || !g || !shadow || !g->equpos || !g->facets) return fail(BTPM_ERR_ARG, "null argument");
  if (g->nv < 3 || g->nf < 1 || g->ntau < 1 || g->nv_primary < 0 || g->nv_primary > g->nv)
    return fail(BTPM_ERR_ARG, "bad sizes: nv %d nf %d ntau %d nv_primary %d", g->nv, g->nf,
                g->ntau, g->nv_primary);
  if (!(g->pol[0] > 0.0) || !(g->au_km > 0.0)) return fail(BTPM_ERR_ARG, "bad position or AU");
  for (int i = 0; i < 3 * g->nf; ++i)
    if (g->facets[i] < 0 || g->facets[i] >= g->nv)
      return fail(BTPM_ERR_ARG, "facet %d: vertex %d out of range", i / 3, g->facets[i]);
  CtxLock lk(ctx);
  API_BEGIN
  HIPCHK(hipSetDevice(ctx->device));
  hipStream_t st = ctx->stream;
  const int nv = g->nv, nf = g->nf, ntau = g->ntau;
  // the reference's trigonometry, on the host (thermal.cpp:12, 1274-1296, 1349, 1371)
  const double PI = 4.0 * std::atan(1.0);
  double frame[12];
  {
    const double r0 = g->rot[0], r1 = g->rot[1];
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
    frame[9] = g->au_km * g->pol[0] * std::cos(g->pol[1]) * std::cos(g->pol[2]);
    frame[10] = g->au_km * g->pol[0] * std::cos(g->pol[1]) * std::sin(g->pol[2]);
    frame[11] = g->au_km * g->pol[0] * std::sin(g->pol[1]);
  }
  const float sol = (float)(1367.0 / (g->pol[0] * g->pol[0]));
  if (solar) *solar = sol;
  std::vector<double> trig(4 * (size_t)ntau);
  for (int s = 0; s < ntau; ++s) {
    const double ap = (2.0 * PI * s * g->rev_primary / ntau) + 0.0;
    const double as = (2.0 * PI * s * g->rev_secondary / ntau) + 0.0;
    trig[4 * (size_t)s] = std::cos(ap);
    trig[4 * (size_t)s + 1] = std::sin(ap);
    trig[4 * (size_t)s + 2] = std::cos(as);
    trig[4 * (size_t)s + 3] = std::sin(as);
  }
  DBuf d_eq, d_fv, d_trig, d_frame;
  upload(d_eq, g->equpos, 3 * (size_t)nv, st);
  upload(d_fv, g->facets, 3 * (size_t)nf, st);
  upload(d_trig, trig.data(), trig.size(), st);
  upload(d_frame, frame, 12, st);
  const size_t nser = (size_t)nf * ntau;
  DBuf d_shadow, d_illum, d_sun;
  d_shadow.alloc(nser * sizeof(int));
  d_illum.alloc(nser * sizeof(float));
  if (sun_facet) d_sun.alloc(3 * nser * sizeof(float));
  // steps per chunk: records of at most ~256 MB
  const size_t rec_step = (size_t)nf * BTPM_GEOM_REC * sizeof(double);
  const int sc = (int)std::max<size_t>(1, std::min<size_t>({(size_t)ntau, (256u << 20) / rec_step,
                                                             (size_t)65535}));
  DBuf d_rec, d_lit, d_cos, d_sunc;
  d_rec.alloc(rec_step * sc);
  d_lit.alloc((size_t)sc * nf * sizeof(int));
  d_cos.alloc((size_t)sc * nf * sizeof(float));
  d_sunc.alloc(3 * (size_t)sc * nf * sizeof(float));
  for (int s0 = 0; s0 < ntau; s0 += sc) {
    const int ns = std::min(sc, ntau - s0);
    hipLaunchKernelGGL(btpm_geom_facets, dim3(grid1(nf, 256), ns), dim3(256), 0, st,
                       d_eq.as<double>(), d_fv.as<int>(), nf, g->nv_primary, d_trig.as<double>(),
                       d_frame.as<double>(), s0, ns, d_rec.as<double>(), d_lit.as<int>(),
                       d_cos.as<float>(), d_sunc.as<float>(), g->sun_scale);
    hipLaunchKernelGGL(btpm_geom_shadow, dim3(grid1(nf, 256), ns), dim3(256), 0, st,
                       d_rec.as<double>(), nf, ns, d_lit.as<int>());
    hipLaunchKernelGGL(btpm_geom_finish, dim3(grid1(nf, 64), grid1(ns, 64)), dim3(256), 0, st,
                       d_lit.as<int>(), d_cos.as<float>(), d_sunc.as<float>(), nf, ns, s0, ntau, sol,
                       d_shadow.as<int>(), d_illum.as<float>(),
                       sun_facet ? d_sun.as<float>() : nullptr);
    HIPCHK(hipGetLastError());
  }
  HIPCHK(hipMemcpyAsync(shadow, d_shadow.p, nser * sizeof(int), hipMemcpyDeviceToHost, st));
  if (illum)
    HIPCHK(hipMemcpyAsync(illum, d_illum.p, nser * sizeof(float), hipMemcpyDeviceToHost, st));
  if (sun_facet)
    HIPCHK(hipMemcpyAsync(sun_facet, d_sun.p, 3 * nser * sizeof(float), hipMemcpyDeviceToHost,
                          st));
  HIPCHK(hipStreamSynchronize(st));
  API_END
}
