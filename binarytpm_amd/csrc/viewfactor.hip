// viewfactor.hip -- the view factors of the Thermal stage's self-heating
// terms, on the GPU:
//
//  * facet view factors of a shape (the Shape_Selfheating producer,
//    SH/main.cpp = 3905/Shape_Selfheating/main.cpp): every ordered facet pair
//    that faces both ways (:628-735), the occlusion cull (hidden, :187-368),
//    the n x n sub-facet refinement of pairs with VF >= 0.05
//    (betterviewfactors, :374-622) and the per-facet totals (:817-835);
//  * shape->crater view factors (calculate_surface_global_viewfactors,
//    T3905/main.cpp:1054-1181): per facet-pair entry and crater sub-facet, the
//    view factor times the crater sub-facet's cosine and its m*m-sample rim
//    visibility.
//
// The producer computes in float and its list membership is decided by signs
// and comparisons of float expressions, so every expression is evaluated in
// the reference's operation order with no fused multiply-adds (the pragma
// below); + - * / and sqrt are IEEE correctly rounded, so the lists, the
// refined sums and the shape->crater factors equal the float restatement bit
// for bit.  The one transcendental decision, acosf(c) <= float(pi)/2, is the
// exact interval BTPM_VF_COS_MIN <= c <= 1; the centroid view factor itself
// keeps the reference's cosf(acosf(c)), which the device libm evaluates
// within an ulp or two of the host's.
#pragma clang fp contract(off)

#include <hip/hip_runtime.h>

#include "btpm_internal.h"
#include "btpm_kernels.h"

namespace {
constexpr int kVR = BTPM_VF_REC;  // floats per facet record
// record: centroid[3], unnormalised normal[3], |normal|, area, vertices
// v0 v1 v2 [3][3], edge vectors v1-v0, v0-v2, v2-v1 [3][3], pad[2]
constexpr int kMid = 0, kNrm = 3, kNlen = 6, kArea = 7, kV0 = 8, kV1 = 11, kV2 = 14, kE1 = 17,
              kE2 = 20, kE3 = 23;
constexpr int kHidChunk = 256;    // partner records staged per LDS chunk
constexpr float kGap = 1.0f - 0x1p-20f;

__device__ __forceinline__ bool facing_cos(float c) {
  return c >= BTPM_VF_COS_MIN && c <= 1.0f;
}

// SH/main.cpp:689-706: separation, its length and the two cosines
__device__ __forceinline__ void pair_cos(const float* __restrict__ ri, const float* __restrict__ rj,
                                         float sep[3], float& sd, float& ci, float& cj) {
  sep[0] = rj[kMid] - ri[kMid];
  sep[1] = rj[kMid + 1] - ri[kMid + 1];
  sep[2] = rj[kMid + 2] - ri[kMid + 2];
  sd = sqrtf((sep[0] * sep[0]) + (sep[1] * sep[1]) + (sep[2] * sep[2]));
  ci = ((ri[kNrm] * sep[0]) + (ri[kNrm + 1] * sep[1]) + (ri[kNrm + 2] * sep[2])) /
       (ri[kNlen] * sd);
  cj = (((-1.0f * rj[kNrm]) * sep[0]) + ((-1.0f * rj[kNrm + 1]) * sep[1]) +
        ((-1.0f * rj[kNrm + 2]) * sep[2])) / (rj[kNlen] * sd);
}

// vertex (row, col) of a triangle split n x n (SH/main.cpp:466-479, rows
// from vertex 2 towards vertex 0, 1-based)
__device__ __forceinline__ void sub_vertex(const float* __restrict__ r, const float t1[3],
                                           const float t2[3], int row, int col, float p[3]) {
#pragma unroll
  for (int q = 0; q < 3; ++q)
    p[q] = r[kV2 + q] + ((float)(col - 1) * t1[q]) - ((float)(row - 1) * t2[q]);
}

// centroid of sub-facet l (SH/main.cpp:481-539: row j holds 2j-1 facets from
// index (j-1)^2; pairs of facets per column k, then the row's last facet)
__device__ void sub_centroid(const float* __restrict__ r, int n, int l, float m[3]) {
  float t1[3], t2[3];
#pragma unroll
  for (int q = 0; q < 3; ++q) {
    t1[q] = (r[kV1 + q] - r[kV0 + q]) / (float)n;
    t2[q] = (r[kV2 + q] - r[kV0 + q]) / (float)n;
  }
  int j = (int)sqrtf((float)l);
  while (j * j > l) --j;
  while ((j + 1) * (j + 1) <= l) ++j;
  const int row = j + 1, rr = l - j * j;
  int ar, ac, br, bc, cr, cc;
  if (rr == 2 * (row - 1)) {
    ar = row + 1, ac = row, br = row + 1, bc = row + 1, cr = row, cc = row;
  } else {
    const int k = rr / 2 + 1;
    if ((rr & 1) == 0)
      ar = row, ac = k + 1, br = row, bc = k, cr = row + 1, cc = k;
    else
      ar = row + 1, ac = k, br = row + 1, bc = k + 1, cr = row, cc = k + 1;
  }
  float pa[3], pb[3], pc[3];
  sub_vertex(r, t1, t2, ar, ac, pa);
  sub_vertex(r, t1, t2, br, bc, pb);
  sub_vertex(r, t1, t2, cr, cc, pc);
#pragma unroll
  for (int q = 0; q < 3; ++q) m[q] = (pa[q] + pb[q] + pc[q]) / 3.0f;
}
}  // namespace

// SH/main.cpp:750-808: one facet per lane
extern "C" __global__ __launch_bounds__(256) void btpm_vf_prep(const float* __restrict__ vtx,
                                                               const int* __restrict__ fv, int nf,
                                                               float* __restrict__ rec) {
  const int f = blockIdx.x * blockDim.x + threadIdx.x;
  if (f >= nf) return;
  float p[3][3];
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    const float* v = vtx + 3 * (size_t)fv[3 * f + a];
    p[a][0] = v[0], p[a][1] = v[1], p[a][2] = v[2];
  }
  float* r = rec + (size_t)f * kVR;
  float e0[3], e1[3], n[3];
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    r[kMid + k] = (p[0][k] + p[1][k] + p[2][k]) / 3.0f;
    e0[k] = p[1][k] - p[0][k];
    e1[k] = p[2][k] - p[0][k];
  }
  n[0] = (e0[1] * e1[2]) - (e0[2] * e1[1]);
  n[1] = (e0[2] * e1[0]) - (e0[0] * e1[2]);
  n[2] = (e0[0] * e1[1]) - (e0[1] * e1[0]);
  const float nl = sqrtf((n[0] * n[0]) + (n[1] * n[1]) + (n[2] * n[2]));
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    r[kNrm + k] = n[k];
    r[kV0 + k] = p[0][k];
    r[kV1 + k] = p[1][k];
    r[kV2 + k] = p[2][k];
    r[kE1 + k] = e0[k];
    r[kE2 + k] = -1.0f * e1[k];
    r[kE3 + k] = p[2][k] - p[1][k];
  }
  r[kNlen] = nl;
  r[kArea] = nl / 2.0f;
  r[26] = 0.0f;
  r[27] = 0.0f;
}

// SH/main.cpp:641-735: row I per workgroup, partners J ascending in chunks of
// 256 lanes.  Count mode (rowptr == nullptr) stores the row's number of
// facing partners; fill mode stores them in order from rowptr[I] with the
// centroid view factor (:709, clamped at 0) and the vector A_I VF d/|d|.
extern "C" __global__ __launch_bounds__(256) void btpm_vf_facing(
    const float* __restrict__ rec, int nf, const int* __restrict__ rowptr, int* __restrict__ count,
    int* __restrict__ col, float* __restrict__ vf, float* __restrict__ vec) {
  __shared__ int wcount[4];
  const int I = blockIdx.x;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const float* ri = rec + (size_t)I * kVR;
  const float PI = 4.0f * atanf(1.0f);
  const unsigned long long lt = (1ull << lane) - 1ull;
  int base = rowptr ? rowptr[I] : 0;
  int total = 0;
  for (int J0 = 0; J0 < nf; J0 += 256) {
    const int J = J0 + (int)threadIdx.x;
    float sep[3], sd = 0.0f, ci = 0.0f, cj = 0.0f;
    bool f = false;
    if (J < nf && J != I) {
      pair_cos(ri, rec + (size_t)J * kVR, sep, sd, ci, cj);
      f = facing_cos(ci) && facing_cos(cj);
    }
    const unsigned long long b = __ballot(f);
    if (lane == 0) wcount[wave] = __popcll(b);
    __syncthreads();
    int off = base;
    for (int w = 0; w < wave; ++w) off += wcount[w];
    const int chunk = wcount[0] + wcount[1] + wcount[2] + wcount[3];
    __syncthreads();
    if (rowptr && f) {
      const int e = off + __popcll(b & lt);
      if (e >= rowptr[I + 1]) continue;  // unreachable: the count pass decided the same pairs
      const float* rj = rec + (size_t)J * kVR;
      float v = (cosf(acosf(ci)) * cosf(acosf(cj)) * rj[kArea]) / (PI * sd * sd);
      if (v < 0.0f) v = 0.0f;
      col[e] = J;
      vf[e] = v;
#pragma unroll
      for (int k = 0; k < 3; ++k) vec[3 * (size_t)e + k] = ri[kArea] * v * (sep[k] / sd);
    }
    base += chunk;
    total += chunk;
  }
  if (!rowptr && threadIdx.x == 0) count[I] = total;
}

// SH/main.cpp:208-360: work item (G, z0) per workgroup, partner z0 + lane of
// G's facing list per lane; every other partner S of G is staged through LDS
// and tested against the line from Z's centroid towards G's.  keep = 0 when
// some S hides Z from G.
extern "C" __global__ __launch_bounds__(256) void btpm_vf_hidden(const float* __restrict__ rec,
                                                                 const int* __restrict__ rowptr1,
                                                                 const int* __restrict__ col1,
                                                                 const int2* __restrict__ work,
                                                                 unsigned char* __restrict__ keep) {
  __shared__ float4 srec[kHidChunk * (kVR / 4)];
  __shared__ int sidx[kHidChunk];
  const int2 wk = work[blockIdx.x];
  const int G = wk.x;
  const int beg = rowptr1[G], nd = rowptr1[G + 1] - beg;
  const int zi = wk.y + (int)threadIdx.x;
  const bool act = zi < nd;
  const int Z = act ? col1[beg + zi] : G;
  const float* gr = rec + (size_t)G * kVR;
  const float* zr = rec + (size_t)Z * kVR;
  const float zm0 = zr[kMid], zm1 = zr[kMid + 1], zm2 = zr[kMid + 2];
  const float gv0 = gr[kMid] - zm0, gv1 = gr[kMid + 1] - zm1, gv2 = gr[kMid + 2] - zm2;
  const float g2 = (gv0 * gv0) + (gv1 * gv1) + (gv2 * gv2);
  const float g = sqrtf(g2);
  bool vis = act;
  for (int s0 = 0; s0 < nd; s0 += kHidChunk) {
    const int cnt = min(kHidChunk, nd - s0);
    __syncthreads();
    if ((int)threadIdx.x < cnt) {
      const int S = col1[beg + s0 + threadIdx.x];
      sidx[threadIdx.x] = S;
      const float4* src = reinterpret_cast<const float4*>(rec + (size_t)S * kVR);
#pragma unroll
      for (int c = 0; c < kVR / 4; ++c) srec[threadIdx.x * (kVR / 4) + c] = src[c];
    }
    __syncthreads();
    for (int r = 0; r < cnt && vis; ++r) {
      const int Sx = sidx[r];
      if (Sx == G || Sx == Z) continue;
      const float* s = reinterpret_cast<const float*>(&srec[r * (kVR / 4)]);
      const float sv0 = s[kMid] - zm0, sv1 = s[kMid + 1] - zm1, sv2 = s[kMid + 2] - zm2;
      // the reference's g > s with s = sqrtf(s2): sqrtf is monotone, so
      // s2 >= g2 rules it out, and s2 below g2 (1 - 2^-20) (normal range)
      // decides it without the root (two correctly rounded roots cannot meet
      // across that gap); in between, the root itself
      const float s2 = (sv0 * sv0) + (sv1 * sv1) + (sv2 * sv2);
      if (!(s2 < g2)) continue;
      if (!(s2 < g2 * kGap && g2 > 1e-30f) && !(g > sqrtf(s2))) continue;
      const float n0 = s[kNrm], n1 = s[kNrm + 1], n2 = s[kNrm + 2];
      const float tc1 = (sv0 * n0) + (sv1 * n1) + (sv2 * n2);
      const float tc2 = (gv0 * n0) + (gv1 * n1) + (gv2 * n2);
      const float t = tc1 / tc2;
      if (!(t > 0.0f)) continue;
      const float x0 = (t * gv0) + zm0, x1 = (t * gv1) + zm1, x2 = (t * gv2) + zm2;
      // edges from v0 (e1), v2 (e2), v1 (e3), each crossed with x - its
      // origin and dotted with the normal: inside when all three are >= 0
      // (tested one at a time: most points fail the first)
      const float* e1 = s + kE1;
      const float a0 = x0 - s[kV0], a1 = x1 - s[kV0 + 1], a2 = x2 - s[kV0 + 2];
      const float k1 = (((e1[1] * a2) - (e1[2] * a1)) * n0) + (((e1[2] * a0) - (e1[0] * a2)) * n1) +
                       (((e1[0] * a1) - (e1[1] * a0)) * n2);
      if (!(k1 >= 0.0f)) continue;
      const float* e2 = s + kE2;
      const float b0 = x0 - s[kV2], b1 = x1 - s[kV2 + 1], b2 = x2 - s[kV2 + 2];
      const float k2 = (((e2[1] * b2) - (e2[2] * b1)) * n0) + (((e2[2] * b0) - (e2[0] * b2)) * n1) +
                       (((e2[0] * b1) - (e2[1] * b0)) * n2);
      if (!(k2 >= 0.0f)) continue;
      const float* e3 = s + kE3;
      const float c0 = x0 - s[kV1], c1 = x1 - s[kV1 + 1], c2 = x2 - s[kV1 + 2];
      const float k3 = (((e3[1] * c2) - (e3[2] * c1)) * n0) + (((e3[2] * c0) - (e3[0] * c2)) * n1) +
                       (((e3[0] * c1) - (e3[1] * c0)) * n2);
      if (k3 >= 0.0f) vis = false;
    }
  }
  if (act) keep[beg + zi] = vis ? 1 : 0;
}

// kept partners per row (one wave per row)
extern "C" __global__ __launch_bounds__(64) void btpm_vf_kept(const int* __restrict__ rowptr1,
                                                              const unsigned char* __restrict__ keep,
                                                              int* __restrict__ count) {
  const int G = blockIdx.x;
  const int end = rowptr1[G + 1];
  int c = 0;
  for (int e0 = rowptr1[G]; e0 < end; e0 += 64) {
    const int e = e0 + (int)threadIdx.x;
    c += __popcll(__ballot(e < end && keep[e] != 0));
  }
  if (threadIdx.x == 0) count[G] = c;
}

// SH/main.cpp:337-347: the kept partners of each row in order (one wave per
// row); pairs to refine (VF >= refine_min, :443) are appended to reflist as
// (entry, row), in no particular order.
extern "C" __global__ __launch_bounds__(64) void btpm_vf_compact(
    const int* __restrict__ rowptr1, const int* __restrict__ col1, const float* __restrict__ vf1,
    const float* __restrict__ vec1, const unsigned char* __restrict__ keep,
    const int* __restrict__ rowptr, int* __restrict__ col, float* __restrict__ vf,
    float* __restrict__ vec, int refine, float refine_min, int* __restrict__ nref,
    int2* __restrict__ reflist) {
  const int G = blockIdx.x;
  const int lane = threadIdx.x;
  const unsigned long long lt = (1ull << lane) - 1ull;
  const int end = rowptr1[G + 1];
  int o = rowptr[G];
  for (int e0 = rowptr1[G]; e0 < end; e0 += 64) {
    const int e = e0 + lane;
    const bool k = e < end && (keep == nullptr || keep[e] != 0);
    const unsigned long long b = __ballot(k);
    const int dst = o + __popcll(b & lt);
    const float v = k ? vf1[e] : 0.0f;
    if (k) {
      col[dst] = col1[e];
      vf[dst] = v;
#pragma unroll
      for (int q = 0; q < 3; ++q) vec[3 * (size_t)dst + q] = vec1[3 * (size_t)e + q];
    }
    const bool rf = refine && k && v >= refine_min;
    const unsigned long long rb = __ballot(rf);
    if (rb) {
      int rbase = 0;
      if (lane == 0) rbase = atomicAdd(nref, __popcll(rb));
      rbase = __shfl(rbase, 0);
      if (rf) reflist[rbase + __popcll(rb & lt)] = make_int2(dst, G);
    }
    o += __popcll(b);
  }
}

// SH/main.cpp:415-615: one refined pair per workgroup.  Lane j owns
// sub-facet j of facet I and sums over the sub-facets k of J (centroids in
// LDS) in order; the row sums are then combined in order j = 0.. by lane 0.
extern "C" __global__ __launch_bounds__(128) void btpm_vf_refine(const float* __restrict__ rec,
                                                                 const int* __restrict__ col,
                                                                 const int2* __restrict__ reflist,
                                                                 int nsub, float* __restrict__ vf,
                                                                 float* __restrict__ vec) {
  extern __shared__ float sm[];
  const int nsf = nsub * nsub;
  float* m2 = sm;                  // [nsf][3]
  float* tot = sm + 3 * nsf;       // [nsf]
  float* totv = sm + 4 * nsf;      // [nsf][3]
  const int2 pr = reflist[blockIdx.x];
  const int e = pr.x;
  const float* ri = rec + (size_t)pr.y * kVR;
  const float* rj = rec + (size_t)col[e] * kVR;
  const float PI = 4.0f * atanf(1.0f);
  for (int l = threadIdx.x; l < nsf; l += blockDim.x) sub_centroid(rj, nsub, l, m2 + 3 * l);
  __syncthreads();
  const float a1 = ri[kArea] / (float)nsf, a2 = rj[kArea] / (float)nsf;
  const float ni0 = ri[kNrm], ni1 = ri[kNrm + 1], ni2 = ri[kNrm + 2], nil = ri[kNlen];
  const float nj0 = rj[kNrm], nj1 = rj[kNrm + 1], nj2 = rj[kNrm + 2], njl = rj[kNlen];
  for (int j = threadIdx.x; j < nsf; j += blockDim.x) {
    float m1[3];
    sub_centroid(ri, nsub, j, m1);
    float t = 0.0f, tv0 = 0.0f, tv1 = 0.0f, tv2 = 0.0f;
    for (int k = 0; k < nsf; ++k) {
      const float s0 = m2[3 * k] - m1[0], s1 = m2[3 * k + 1] - m1[1], s2 = m2[3 * k + 2] - m1[2];
      const float sd = sqrtf((s0 * s0) + (s1 * s1) + (s2 * s2));
      const float ci = ((ni0 * s0) + (ni1 * s1) + (ni2 * s2)) / (nil * sd);
      const float cj = (((-1.0f * nj0) * s0) + ((-1.0f * nj1) * s1) + ((-1.0f * nj2) * s2)) /
                       (njl * sd);
      float v = (ci * cj * a2) / (PI * sd * sd);
      if (v < 0.0f) v = 0.0f;
      t = t + v;
      tv0 = tv0 + (v * (s0 / sd));
      tv1 = tv1 + (v * (s1 / sd));
      tv2 = tv2 + (v * (s2 / sd));
    }
    tot[j] = t;
    totv[3 * j] = tv0;
    totv[3 * j + 1] = tv1;
    totv[3 * j + 2] = tv2;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    float eff = 0.0f, v0 = 0.0f, v1 = 0.0f, v2 = 0.0f;
    for (int j = 0; j < nsf; ++j) {
      eff = eff + ((a1 / ri[kArea]) * tot[j]);
      v0 = v0 + (a1 * totv[3 * j]);
      v1 = v1 + (a1 * totv[3 * j + 1]);
      v2 = v2 + (a1 * totv[3 * j + 2]);
    }
    vf[e] = eff;
    vec[3 * (size_t)e] = v0;
    vec[3 * (size_t)e + 1] = v1;
    vec[3 * (size_t)e + 2] = v2;
  }
}

// SH/main.cpp:822-835 and :874-878: per row (one wave) the total view
// factor and photon momentum vector, summed in list order, and each entry's
// vector-list value vec / (VF A_I)
extern "C" __global__ __launch_bounds__(64) void btpm_vf_finish(
    const int* __restrict__ rowptr, const float* __restrict__ vf, const float* __restrict__ vec,
    const float* __restrict__ rec, float* __restrict__ dir, float* __restrict__ total,
    float* __restrict__ photon) {
  const int G = blockIdx.x;
  const int lane = threadIdx.x;
  const float area = rec[(size_t)G * kVR + kArea];
  const int beg = rowptr[G], end = rowptr[G + 1];
  float t = 0.0f, p0 = 0.0f, p1 = 0.0f, p2 = 0.0f;
  for (int e0 = beg; e0 < end; e0 += 64) {
    const int e = e0 + lane;
    float v = 0.0f, q0 = 0.0f, q1 = 0.0f, q2 = 0.0f;
    if (e < end) {
      v = vf[e];
      const float x0 = vec[3 * (size_t)e], x1 = vec[3 * (size_t)e + 1], x2 = vec[3 * (size_t)e + 2];
      q0 = x0 / area, q1 = x1 / area, q2 = x2 / area;
      dir[3 * (size_t)e] = x0 / (v * area);
      dir[3 * (size_t)e + 1] = x1 / (v * area);
      dir[3 * (size_t)e + 2] = x2 / (v * area);
    }
    const int cnt = min(64, end - e0);
    for (int i = 0; i < cnt; ++i) {
      t = t + __shfl(v, i);
      p0 = p0 + __shfl(q0, i);
      p1 = p1 + __shfl(q1, i);
      p2 = p2 + __shfl(q2, i);
    }
  }
  if (lane == 0) {
    total[G] = end > beg ? t : 0.0f;
    photon[3 * (size_t)G] = p0;
    photon[3 * (size_t)G + 1] = p1;
    photon[3 * (size_t)G + 2] = p2;
  }
}

// T3905/main.cpp:1059-1097: the facet frame (rows: unit first edge, normal x
// edge, unit normal) in double, one facet per lane
extern "C" __global__ __launch_bounds__(256) void btpm_drf_frames(const double* __restrict__ vtx,
                                                                  const int* __restrict__ fv, int nf,
                                                                  double* __restrict__ frame) {
  const int f = blockIdx.x * blockDim.x + threadIdx.x;
  if (f >= nf) return;
  const double* p0 = vtx + 3 * (size_t)fv[3 * f];
  const double* p1 = vtx + 3 * (size_t)fv[3 * f + 1];
  const double* p2 = vtx + 3 * (size_t)fv[3 * f + 2];
  double e[3], b[3], n[3], y[3];
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    e[k] = p1[k] - p0[k];
    b[k] = p2[k] - p0[k];
  }
  n[0] = (e[1] * b[2]) - (e[2] * b[1]);
  n[1] = (e[2] * b[0]) - (e[0] * b[2]);
  n[2] = (e[0] * b[1]) - (e[1] * b[0]);
  const double nl = sqrt((n[0] * n[0]) + (n[1] * n[1]) + (n[2] * n[2]));
#pragma unroll
  for (int k = 0; k < 3; ++k) n[k] = n[k] / nl;
  const double x = sqrt((e[0] * e[0]) + (e[1] * e[1]) + (e[2] * e[2]));
#pragma unroll
  for (int k = 0; k < 3; ++k) e[k] = e[k] / x;
  y[0] = (n[1] * e[2]) - (n[2] * e[1]);
  y[1] = (n[2] * e[0]) - (n[0] * e[2]);
  y[2] = (n[0] * e[1]) - (n[1] * e[0]);
  double* m = frame + 9 * (size_t)f;
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    m[k] = e[k];
    m[3 + k] = y[k];
    m[6 + k] = n[k];
  }
}

// row index of every CSR entry (one wave per row)
extern "C" __global__ __launch_bounds__(64) void btpm_drf_rows(const int* __restrict__ rowptr,
                                                               int* __restrict__ row) {
  const int f = blockIdx.x;
  for (int e = rowptr[f] + (int)threadIdx.x; e < rowptr[f + 1]; e += 64) row[e] = f;
}

// T3905/main.cpp:1127-1176: a workgroup per tile of 64 list entries (lane =
// entry), wave w taking crater sub-facets j = w, w+4, ...; the per-entry
// direction in the facet frame is computed once per lane.
//   csub[j]       = unit normal of crater sub-facet j (xyz, pad)
//   samp[j][p]    = rim-visibility sample p of sub-facet j (xyz, pad): the
//                   reference's vertex 2 + mul_p0 trivector0 + mul_p1 trivector1
//                   (:1159-1161), the same float operations, done on the host
//   vtab[c]       = the visibility after c shadowed samples (the reference's
//                   running 1 - c/(m*m) with its clamp at 0, :1165-1166)
// A sample is shadowed when its height is at most |w_z|/|w_xy| times its
// distance to the rim along w (:1162-1165); the count of shadowed samples
// decides the visibility, so the loop counts them.  Results go through an LDS
// tile so that each entry's row of ncf factors is stored contiguously.
extern "C" __global__ __launch_bounds__(256) void btpm_drf(
    const double* __restrict__ frame, const int* __restrict__ row, const float* __restrict__ vfv,
    const float* __restrict__ dir, const float4* __restrict__ csub,
    const float4* __restrict__ samp, const float* __restrict__ vtab, int ncf, int nsamp,
    float radius, long long nnz, float* __restrict__ drf) {
  extern __shared__ float sm[];
  float* tile = sm;                        // [64][ncf + 1]
  float* svt = sm + 64 * (ncf + 1);        // [nsamp + 1]
  for (int i = threadIdx.x; i <= nsamp; i += blockDim.x) svt[i] = vtab[i];
  const long long q0 = (long long)blockIdx.x * 64;
  const int lane = threadIdx.x & 63;
  // the wave's crater sub-facets are uniform across its lanes: their normals
  // and samples come through scalar loads
  const int wave = __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
  const long long q = q0 + lane;
  const bool act = q < nnz;
  float w0 = 0.0f, w1 = 0.0f, w2 = 1.0f, a = 0.0f, vf = 0.0f;
  if (act) {
    const double* M = frame + 9 * (size_t)row[q];
    const double d0 = dir[3 * q], d1 = dir[3 * q + 1], d2 = dir[3 * q + 2];
    w0 = (float)((M[0] * d0) + (M[1] * d1) + (M[2] * d2));
    w1 = (float)((M[3] * d0) + (M[4] * d1) + (M[5] * d2));
    w2 = (float)((M[6] * d0) + (M[7] * d1) + (M[8] * d2));
    a = fabsf(w2) / sqrtf((w0 * w0) + (w1 * w1));
    vf = vfv[q];
  }
  const float nw1 = -1.0f * w1;
  const float rr = radius * radius;
  __syncthreads();
  for (int j = wave; j < ncf; j += 4) {
    const float4 nj = csub[j];
    const float ang = (w0 * nj.x) + (w1 * nj.y) + (w2 * nj.z);
    float vis = 0.0f;
    if (act && !(ang < 0.0f)) {
      const float4* sp = samp + (size_t)__builtin_amdgcn_readfirstlane(j) * nsamp;
      auto shadowed = [&](const float4 m) -> int {
        const float tc1 = (nw1 * m.x) + (w0 * m.y);
        const float sr = sqrtf(rr - (tc1 * tc1));
        const float tc0 = (w0 * m.x) + (w1 * m.y) - sr;
        return (m.z <= (a * tc0)) ? 1 : 0;
      };
      int hits = 0, p = 0;
      for (; p + 4 <= nsamp; p += 4) {
        const int pu = __builtin_amdgcn_readfirstlane(p);
        const float4 m0 = sp[pu], m1 = sp[pu + 1], m2 = sp[pu + 2], m3 = sp[pu + 3];
        hits += shadowed(m0) + shadowed(m1) + shadowed(m2) + shadowed(m3);
      }
      for (; p < nsamp; ++p) hits += shadowed(sp[__builtin_amdgcn_readfirstlane(p)]);
      vis = svt[hits];
    }
    float r = (vf * vis * ang) / w2;
    if (r > 1.0f) r = 1.0f;
    tile[lane * (ncf + 1) + j] = r;
  }
  __syncthreads();
  const long long rows = min(64ll, nnz - q0);
  float* out = drf + q0 * ncf;
  for (int i = threadIdx.x; i < rows * ncf; i += blockDim.x) {
    const int e = i / ncf, j = i - e * ncf;
    out[i] = tile[e * (ncf + 1) + j];
  }
}

// Crater sunlight (T3905/main.cpp:863-886, surfacegeometry; T3905/
// GPU_code.cu:152-210, sfcpartialshadow): a workgroup per 64 (facet, step)
// entries, lane = entry, wave w taking crater sub-facets j = w, w+4, ...
// with the rim samples through scalar loads (as btpm_drf).  Per entry the
// Sun vector sv in the facet frame; cos_j = (sv . n_j)/|sv|; the sunlit
// fraction starts at 1 (0 when the facet is dark or cos_j <= 0) and loses
// 1/(m*m) per sample below the rim along sv (normalised by its horizontal
// length) while above 0 -- a host table gives it after c hits.
extern "C" __global__ __launch_bounds__(256) void btpm_crater_sun(
    const int* __restrict__ lit_in, const float* __restrict__ sun,
    const float4* __restrict__ csub, const float4* __restrict__ samp,
    const float* __restrict__ vtab, int ncf, int nsamp, float radius, long long n,
    float* __restrict__ shadow, float* __restrict__ cosang) {
  extern __shared__ float sm[];
  float* tsh = sm;                         // [64][ncf + 1]
  float* tco = sm + 64 * (ncf + 1);        // [64][ncf + 1]
  float* svt = sm + 128 * (ncf + 1);       // [nsamp + 1]
  for (int i = threadIdx.x; i <= nsamp; i += blockDim.x) svt[i] = vtab[i];
  const long long q0 = (long long)blockIdx.x * 64;
  const int lane = threadIdx.x & 63;
  const int wave = __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
  const long long q = q0 + lane;
  const bool act = q < n;
  float s0 = 0.0f, s1 = 0.0f, s2 = 1.0f, t = 1.0f, w0 = 0.0f, w1 = 0.0f, a = 0.0f;
  int lit = 0;
  if (act) {
    s0 = sun[3 * q], s1 = sun[3 * q + 1], s2 = sun[3 * q + 2];
    lit = lit_in[q];
    t = sqrtf((s0 * s0) + (s1 * s1) + (s2 * s2));
    if (lit == 1) {
      const float h = sqrtf((s0 * s0) + (s1 * s1));
      w0 = s0 / h;
      w1 = s1 / h;
      const float w2 = s2 / h;
      a = fabsf(w2) / sqrtf((w0 * w0) + (w1 * w1));
    }
  }
  const float nw1 = -1.0f * w1;
  const float rr = radius * radius;
  __syncthreads();
  for (int j = wave; j < ncf; j += 4) {
    const float4 nj = csub[j];
    const float ill = ((s0 * nj.x) + (s1 * nj.y) + (s2 * nj.z)) / t;
    float sh = 0.0f;
    if (act && lit == 1 && !(ill <= 0.0f)) {
      const float4* sp = samp + (size_t)__builtin_amdgcn_readfirstlane(j) * nsamp;
      auto below = [&](const float4 m) -> int {
        const float tc1 = (nw1 * m.x) + (w0 * m.y);
        const float sr = sqrtf(rr - (tc1 * tc1));
        const float tc0 = (w0 * m.x) + (w1 * m.y) - sr;
        return (m.z <= (a * tc0)) ? 1 : 0;
      };
      int hits = 0, p = 0;
      for (; p + 4 <= nsamp; p += 4) {
        const int pu = __builtin_amdgcn_readfirstlane(p);
        const float4 m0 = sp[pu], m1 = sp[pu + 1], m2 = sp[pu + 2], m3 = sp[pu + 3];
        hits += below(m0) + below(m1) + below(m2) + below(m3);
      }
      for (; p < nsamp; ++p) hits += below(sp[__builtin_amdgcn_readfirstlane(p)]);
      sh = svt[hits];
    } else if (act && lit != 0 && lit != 1 && !(ill <= 0.0f)) {
      sh = 1.0f;   // a flag other than 0 or 1: lit, no partial shadow (as the reference)
    }
    tsh[lane * (ncf + 1) + j] = sh;
    tco[lane * (ncf + 1) + j] = ill;
  }
  __syncthreads();
  const long long rows = min(64ll, n - q0);
  for (int i = threadIdx.x; i < rows * ncf; i += blockDim.x) {
    const int e = i / ncf, j = i - e * ncf;
    shadow[q0 * ncf + i] = tsh[e * (ncf + 1) + j];
    if (cosang) cosang[q0 * ncf + i] = tco[e * (ncf + 1) + j];  // null: fractions only
  }
}

// ===========================================================================
#include <algorithm>
#include <climits>
#include <cmath>
#include <memory>
#include <vector>

#include "runtime.h"

using namespace btpm;

// facet view factors of one mesh, resident on its device until freed
struct btpm_vf {
  int device = 0;
  int nf = 0;
  long long n_facing = 0, nnz = 0, n_refined = 0;
  std::vector<int> rowptr;
  DBuf col, vf, vec, dir, total, photon;
};

namespace {
// the crater model's sub-facet unit normals, its m*m rim-visibility samples
// per sub-facet and the visibility after c shadowed samples
void crater_samples(const float* cvtx, const int* cfv, int ncf, int m, std::vector<float4>& csub,
                    std::vector<float4>& samp, std::vector<float>& vtab) {
  const int nsamp = m * m;
  // crater sub-facets (T3905/main.cpp:819-860): unit normal, and the
  // rim-visibility samples vertex 2 + mul0 trivector0 + mul1 trivector1
  // with the multipliers of :1101-1114
  std::vector<float> mul(2 * (size_t)nsamp);
  for (int k = 1; k != m + 1; ++k)
    for (int l = 1; l != k + 1; ++l) {
      const int p = ((k - 1) * (k - 1)) + (2 * (l - 1));
      mul[2 * p] = (float)k - (1.0f / 3.0f);
      mul[2 * p + 1] = (float)l - (2.0f / 3.0f);
    }
  for (int k = 2; k != m + 1; ++k)
    for (int l = 1; l != k; ++l) {
      const int p = ((k - 1) * (k - 1)) + (2 * (l - 1)) + 1;
      mul[2 * p] = (float)k - (2.0f / 3.0f);
      mul[2 * p + 1] = (float)l - (1.0f / 3.0f);
    }
  csub.assign(ncf, make_float4(0.0f, 0.0f, 0.0f, 0.0f));
  samp.assign((size_t)ncf * nsamp, make_float4(0.0f, 0.0f, 0.0f, 0.0f));
  for (int j = 0; j < ncf; ++j) {
    const float* p0 = cvtx + 3 * (size_t)cfv[3 * j];
    const float* p1 = cvtx + 3 * (size_t)cfv[3 * j + 1];
    const float* p2 = cvtx + 3 * (size_t)cfv[3 * j + 2];
    float a[3], b[3], n[3], t0[3], t1[3];
    for (int k = 0; k < 3; ++k) {
      a[k] = p1[k] - p0[k];
      b[k] = p2[k] - p0[k];
    }
    n[0] = (a[1] * b[2]) - (a[2] * b[1]);
    n[1] = (a[2] * b[0]) - (a[0] * b[2]);
    n[2] = (a[0] * b[1]) - (a[1] * b[0]);
    const float l = std::sqrt((n[0] * n[0]) + (n[1] * n[1]) + (n[2] * n[2]));
    csub[j] = make_float4(n[0] / l, n[1] / l, n[2] / l, 0.0f);
    for (int k = 0; k < 3; ++k) {
      t0[k] = -1.0f * b[k] / (float)m;
      t1[k] = a[k] / (float)m;
    }
    for (int p = 0; p < nsamp; ++p) {
      float mp[3];
      for (int k = 0; k < 3; ++k)
        mp[k] = p2[k] + (mul[2 * p] * t0[k]) + (mul[2 * p + 1] * t1[k]);
      samp[(size_t)j * nsamp + p] = make_float4(mp[0], mp[1], mp[2], 0.0f);
    }
  }
  // visibility after c shadowed samples (T3905/main.cpp:1165-1166)
  vtab.assign(nsamp + 1, 0.0f);
  {
    const float step = 1.0f / (float)(m * m);
    float vis = 1.0f;
    vtab[0] = vis;
    for (int c = 1; c <= nsamp; ++c) {
      if (vis != 0.0f) vis = vis - step;
      if (vis < 0.0f) vis = 0.0f;
      vtab[c] = vis;
    }
  }
}

// exclusive scan of per-row counts into a CSR row pointer; false on overflow
bool scan_rows(const std::vector<int>& cnt, std::vector<int>& rp) {
  rp.assign(cnt.size() + 1, 0);
  long long s = 0;
  for (size_t i = 0; i < cnt.size(); ++i) {
    s += cnt[i];
    if (s > INT_MAX) return false;
    rp[i + 1] = (int)s;
  }
  return true;
}
}  // namespace

extern "C" int btpm_vf_build(btpm_ctx* ctx, const btpm_vf_mesh* m, btpm_vf** out) {
  if (!ctx || !m || !out || !m->vtx || !m->facets) return fail(BTPM_ERR_ARG, "null argument");
  *out = nullptr;
  if (m->nv < 3 || m->nf < 1) return fail(BTPM_ERR_ARG, "bad sizes: nv %d nf %d", m->nv, m->nf);
  if (m->nsub < 0 || m->nsub > BTPM_VF_NSUB_MAX)
    return fail(BTPM_ERR_ARG, "nsub %d outside 0..%d", m->nsub, BTPM_VF_NSUB_MAX);
  if (std::isnan(m->refine_min)) return fail(BTPM_ERR_ARG, "refine_min is NaN");
  for (int i = 0; i < 3 * m->nf; ++i)
    if (m->facets[i] < 0 || m->facets[i] >= m->nv)
      return fail(BTPM_ERR_ARG, "facet %d: vertex %d out of range", i / 3, m->facets[i]);
  CtxLock lk(ctx);
  API_BEGIN
  HIPCHK(hipSetDevice(ctx->device));
  hipStream_t st = ctx->stream;
  const int nf = m->nf;
  btpm_vf* r = new btpm_vf;
  std::unique_ptr<btpm_vf> guard(r);
  r->device = ctx->device;
  r->nf = nf;
  DBuf d_vtx, d_fv, d_rec;
  upload(d_vtx, m->vtx, 3 * (size_t)m->nv, st);
  upload(d_fv, m->facets, 3 * (size_t)nf, st);
  d_rec.alloc((size_t)nf * BTPM_VF_REC * sizeof(float));
  hipLaunchKernelGGL(btpm_vf_prep, dim3(grid1(nf, 256)), dim3(256), 0, st, d_vtx.as<float>(),
                     d_fv.as<int>(), nf, d_rec.as<float>());
  // facing pairs: counts, row pointers, then the ordered lists
  DBuf d_cnt;
  d_cnt.alloc((size_t)nf * sizeof(int));
  hipLaunchKernelGGL(btpm_vf_facing, dim3(nf), dim3(256), 0, st, d_rec.as<float>(), nf,
                     (const int*)nullptr, d_cnt.as<int>(), (int*)nullptr, (float*)nullptr,
                     (float*)nullptr);
  HIPCHK(hipGetLastError());
  std::vector<int> cnt(nf), rp1;
  HIPCHK(hipMemcpyAsync(cnt.data(), d_cnt.p, nf * sizeof(int), hipMemcpyDeviceToHost, st));
  HIPCHK(hipStreamSynchronize(st));
  if (!scan_rows(cnt, rp1)) return fail(BTPM_ERR_ARG, "more than 2^31 facing pairs");
  const long long n1 = rp1[nf];
  r->n_facing = n1;
  DBuf d_rp1, d_col1, d_vf1, d_vec1;
  upload(d_rp1, rp1.data(), rp1.size(), st);
  d_col1.alloc(std::max<size_t>(1, n1) * sizeof(int));
  d_vf1.alloc(std::max<size_t>(1, n1) * sizeof(float));
  d_vec1.alloc(std::max<size_t>(1, n1) * 3 * sizeof(float));
  hipLaunchKernelGGL(btpm_vf_facing, dim3(nf), dim3(256), 0, st, d_rec.as<float>(), nf,
                     d_rp1.as<int>(), (int*)nullptr, d_col1.as<int>(), d_vf1.as<float>(),
                     d_vec1.as<float>());
  HIPCHK(hipGetLastError());
  // occlusion cull: (row, first partner) work items, longest rows first
  DBuf d_keep;
  std::vector<int> rp;
  if (m->cull && n1 > 0) {
    std::vector<int> rows;
    for (int G = 0; G < nf; ++G)
      if (cnt[G] > 0) rows.push_back(G);
    std::stable_sort(rows.begin(), rows.end(), [&](int a, int b) { return cnt[a] > cnt[b]; });
    std::vector<int2> work;
    for (int G : rows)
      for (int z0 = 0; z0 < cnt[G]; z0 += 256) work.push_back(make_int2(G, z0));
    DBuf d_work;
    upload(d_work, work.data(), work.size(), st);
    d_keep.alloc((size_t)n1);
    for (size_t w0 = 0; w0 < work.size(); w0 += 1u << 30) {
      const size_t nw = std::min<size_t>(work.size() - w0, 1u << 30);
      hipLaunchKernelGGL(btpm_vf_hidden, dim3((unsigned)nw), dim3(256), 0, st, d_rec.as<float>(),
                         d_rp1.as<int>(), d_col1.as<int>(), d_work.as<int2>() + w0,
                         d_keep.as<unsigned char>());
    }
    hipLaunchKernelGGL(btpm_vf_kept, dim3(nf), dim3(64), 0, st, d_rp1.as<int>(),
                       d_keep.as<unsigned char>(), d_cnt.as<int>());
    HIPCHK(hipGetLastError());
    HIPCHK(hipMemcpyAsync(cnt.data(), d_cnt.p, nf * sizeof(int), hipMemcpyDeviceToHost, st));
    HIPCHK(hipStreamSynchronize(st));
    scan_rows(cnt, rp);
  } else {
    rp = rp1;
  }
  const long long nnz = rp[nf];
  r->nnz = nnz;
  r->rowptr = rp;
  DBuf d_rp, d_ref, d_nref;
  upload(d_rp, rp.data(), rp.size(), st);
  r->col.alloc(std::max<size_t>(1, nnz) * sizeof(int));
  r->vf.alloc(std::max<size_t>(1, nnz) * sizeof(float));
  r->vec.alloc(std::max<size_t>(1, nnz) * 3 * sizeof(float));
  r->dir.alloc(std::max<size_t>(1, nnz) * 3 * sizeof(float));
  r->total.alloc((size_t)nf * sizeof(float));
  r->photon.alloc((size_t)nf * 3 * sizeof(float));
  d_ref.alloc(std::max<size_t>(1, nnz) * sizeof(int2));
  const int zero = 0;
  upload(d_nref, &zero, 1, st);
  hipLaunchKernelGGL(btpm_vf_compact, dim3(nf), dim3(64), 0, st, d_rp1.as<int>(), d_col1.as<int>(),
                     d_vf1.as<float>(), d_vec1.as<float>(),
                     d_keep.p ? d_keep.as<unsigned char>() : (const unsigned char*)nullptr,
                     d_rp.as<int>(), r->col.as<int>(), r->vf.as<float>(), r->vec.as<float>(),
                     m->nsub > 0 ? 1 : 0, m->refine_min, d_nref.as<int>(), d_ref.as<int2>());
  HIPCHK(hipGetLastError());
  int nref = 0;
  HIPCHK(hipMemcpyAsync(&nref, d_nref.p, sizeof(int), hipMemcpyDeviceToHost, st));
  HIPCHK(hipStreamSynchronize(st));
  r->n_refined = nref;
  if (nref > 0) {
    const int nsf = m->nsub * m->nsub;
    hipLaunchKernelGGL(btpm_vf_refine, dim3(nref), dim3(128), 7 * nsf * sizeof(float), st,
                       d_rec.as<float>(), r->col.as<int>(), d_ref.as<int2>(), m->nsub,
                       r->vf.as<float>(), r->vec.as<float>());
  }
  hipLaunchKernelGGL(btpm_vf_finish, dim3(nf), dim3(64), 0, st, d_rp.as<int>(), r->vf.as<float>(),
                     r->vec.as<float>(), d_rec.as<float>(), r->dir.as<float>(),
                     r->total.as<float>(), r->photon.as<float>());
  HIPCHK(hipGetLastError());
  HIPCHK(hipStreamSynchronize(st));
  *out = guard.release();
  API_END
}

extern "C" int btpm_vf_counts(const btpm_vf* v, long long* n_facing, long long* nnz,
                              long long* n_refined) {
  if (!v) return fail(BTPM_ERR_ARG, "null view-factor handle");
  if (n_facing) *n_facing = v->n_facing;
  if (nnz) *nnz = v->nnz;
  if (n_refined) *n_refined = v->n_refined;
  return BTPM_OK;
}

extern "C" int btpm_vf_fetch(const btpm_vf* v, int* rowptr, int* col, float* vf, float* dir,
                             float* vec, float* total, float* photon) {
  if (!v) return fail(BTPM_ERR_ARG, "null view-factor handle");
  API_BEGIN
  HIPCHK(hipSetDevice(v->device));
  const size_t n = (size_t)v->nnz, nf = (size_t)v->nf;
  if (rowptr) std::copy(v->rowptr.begin(), v->rowptr.end(), rowptr);
  if (col && n) HIPCHK(hipMemcpy(col, v->col.p, n * sizeof(int), hipMemcpyDeviceToHost));
  if (vf && n) HIPCHK(hipMemcpy(vf, v->vf.p, n * sizeof(float), hipMemcpyDeviceToHost));
  if (dir && n) HIPCHK(hipMemcpy(dir, v->dir.p, 3 * n * sizeof(float), hipMemcpyDeviceToHost));
  if (vec && n) HIPCHK(hipMemcpy(vec, v->vec.p, 3 * n * sizeof(float), hipMemcpyDeviceToHost));
  if (total) HIPCHK(hipMemcpy(total, v->total.p, nf * sizeof(float), hipMemcpyDeviceToHost));
  if (photon) HIPCHK(hipMemcpy(photon, v->photon.p, 3 * nf * sizeof(float), hipMemcpyDeviceToHost));
  API_END
}

extern "C" void btpm_vf_free(btpm_vf* v) {
  if (!v) return;
  (void)hipSetDevice(v->device);
  delete v;
}

extern "C" int btpm_crater_viewfactors(btpm_ctx* ctx, const btpm_drf_in* in, float* drf) {
  if (!ctx || !in || !drf || !in->vtx || !in->facets || !in->rowptr || !in->crater_vtx ||
      !in->crater_facets)
    return fail(BTPM_ERR_ARG, "null argument");
  if (in->nv < 3 || in->nf < 1 || in->ncv < 3 || in->ncf < 1 || in->ncf > 128 || in->m < 1 ||
      in->m > 64)
    return fail(BTPM_ERR_ARG, "bad sizes: nv %d nf %d ncv %d ncf %d m %d", in->nv, in->nf, in->ncv,
                in->ncf, in->m);
  if (in->rowptr[0] != 0) return fail(BTPM_ERR_ARG, "rowptr[0] must be 0");
  for (int f = 0; f < in->nf; ++f)
    if (in->rowptr[f + 1] < in->rowptr[f]) return fail(BTPM_ERR_ARG, "rowptr decreases at %d", f);
  const long long nnz = in->rowptr[in->nf];
  if (nnz > 0 && (!in->vf || !in->dir)) return fail(BTPM_ERR_ARG, "null view factors");
  for (int i = 0; i < 3 * in->nf; ++i)
    if (in->facets[i] < 0 || in->facets[i] >= in->nv)
      return fail(BTPM_ERR_ARG, "facet %d: vertex %d out of range", i / 3, in->facets[i]);
  for (int i = 0; i < 3 * in->ncf; ++i)
    if (in->crater_facets[i] < 0 || in->crater_facets[i] >= in->ncv)
      return fail(BTPM_ERR_ARG, "crater facet %d: vertex %d out of range", i / 3,
                  in->crater_facets[i]);
  if (nnz == 0) return BTPM_OK;
  CtxLock lk(ctx);
  API_BEGIN
  HIPCHK(hipSetDevice(ctx->device));
  hipStream_t st = ctx->stream;
  const int ncf = in->ncf, m = in->m, nsamp = m * m;
  std::vector<float4> csub, samp;
  std::vector<float> vtab;
  crater_samples(in->crater_vtx, in->crater_facets, ncf, m, csub, samp, vtab);
  DBuf d_vtx, d_fv, d_rp, d_vf, d_dir, d_cs, d_samp, d_vt, d_frame, d_row, d_out;
  upload(d_vtx, in->vtx, 3 * (size_t)in->nv, st);
  upload(d_fv, in->facets, 3 * (size_t)in->nf, st);
  upload(d_rp, in->rowptr, (size_t)in->nf + 1, st);
  upload(d_vf, in->vf, (size_t)nnz, st);
  upload(d_dir, in->dir, 3 * (size_t)nnz, st);
  upload(d_cs, csub.data(), csub.size(), st);
  upload(d_samp, samp.data(), samp.size(), st);
  upload(d_vt, vtab.data(), vtab.size(), st);
  d_frame.alloc(9 * (size_t)in->nf * sizeof(double));
  d_row.alloc((size_t)nnz * sizeof(int));
  const long long n = nnz * ncf;
  d_out.alloc((size_t)n * sizeof(float));
  hipLaunchKernelGGL(btpm_drf_frames, dim3(grid1(in->nf, 256)), dim3(256), 0, st,
                     d_vtx.as<double>(), d_fv.as<int>(), in->nf, d_frame.as<double>());
  hipLaunchKernelGGL(btpm_drf_rows, dim3(in->nf), dim3(64), 0, st, d_rp.as<int>(), d_row.as<int>());
  const size_t shm = (64 * ((size_t)ncf + 1) + (size_t)nsamp + 1) * sizeof(float);
  hipLaunchKernelGGL(btpm_drf, dim3(grid1((size_t)nnz, 64)), dim3(256), shm, st,
                     d_frame.as<double>(), d_row.as<int>(), d_vf.as<float>(), d_dir.as<float>(),
                     d_cs.as<float4>(), d_samp.as<float4>(), d_vt.as<float>(), ncf, nsamp,
                     in->crater_radius, nnz, d_out.as<float>());
  HIPCHK(hipGetLastError());
  download_staged(ctx, drf, d_out.p, (size_t)n * sizeof(float));
  API_END
}

extern "C++" {
namespace btpm {
// The fractions of btpm_crater_sunlight (lit, or visible with the observer as
// the source) of n facet-steps straight into device memory frac[n][ncf], from
// host flags lit[n] and source vectors src[n][3] in the facet frames (no
// cosines, no readback).  The caller holds the context's lock inside its
// API_BEGIN (HIP failures throw); the work goes on the context's stream.
void crater_fractions_on_device(btpm_ctx* ctx, const int* lit, const float* src, long long n,
                                const float* cvtx, const int* cfv, int ncf, int m, float radius,
                                float* frac) {
  hipStream_t st = ctx->stream;
  const int nsamp = m * m;
  std::vector<float4> csub, samp;
  std::vector<float> vtab;
  crater_samples(cvtx, cfv, ncf, m, csub, samp, vtab);
  DBuf d_cs, d_samp, d_vt, d_lit, d_src;
  upload(d_cs, csub.data(), csub.size(), st);
  upload(d_samp, samp.data(), samp.size(), st);
  upload(d_vt, vtab.data(), vtab.size(), st);
  const size_t shm = (128 * ((size_t)ncf + 1) + (size_t)nsamp + 1) * sizeof(float);
  const long long chunk = 1ll << 22;
  const long long first = std::min(n, chunk);
  d_lit.alloc((size_t)first * sizeof(int));
  d_src.alloc((size_t)first * 3 * sizeof(float));
  for (long long q0 = 0; q0 < n; q0 += chunk) {
    const long long k = std::min(chunk, n - q0);
    HIPCHK(hipMemcpyAsync(d_lit.p, lit + q0, (size_t)k * sizeof(int), hipMemcpyHostToDevice, st));
    HIPCHK(hipMemcpyAsync(d_src.p, src + 3 * q0, (size_t)k * 3 * sizeof(float),
                          hipMemcpyHostToDevice, st));
    hipLaunchKernelGGL(btpm_crater_sun, dim3((unsigned)((k + 63) / 64)), dim3(256), shm, st,
                       d_lit.as<int>(), d_src.as<float>(), d_cs.as<float4>(), d_samp.as<float4>(),
                       d_vt.as<float>(), ncf, nsamp, radius, k, frac + q0 * ncf,
                       (float*)nullptr);
    HIPCHK(hipGetLastError());
  }
  HIPCHK(hipStreamSynchronize(st));  // the chunk buffers are freed on return
}
}  // namespace btpm
}  // extern "C++"

extern "C" int btpm_crater_sunlight(btpm_ctx* ctx, const btpm_crater_sun_in* in, float* shadow,
                                    float* cosang) {
  if (!ctx || !in || !shadow || !cosang || !in->shadow || !in->sun || !in->crater_vtx ||
      !in->crater_facets)
    return fail(BTPM_ERR_ARG, "null argument");
  if (in->nf < 1 || in->ntau < 1 || in->ncv < 3 || in->ncf < 1 || in->ncf > 128 || in->m < 1 ||
      in->m > 64)
    return fail(BTPM_ERR_ARG, "bad sizes: nf %d ntau %d ncv %d ncf %d m %d", in->nf, in->ntau,
                in->ncv, in->ncf, in->m);
  for (int i = 0; i < 3 * in->ncf; ++i)
    if (in->crater_facets[i] < 0 || in->crater_facets[i] >= in->ncv)
      return fail(BTPM_ERR_ARG, "crater facet %d: vertex %d out of range", i / 3,
                  in->crater_facets[i]);
  CtxLock lk(ctx);
  API_BEGIN
  HIPCHK(hipSetDevice(ctx->device));
  hipStream_t st = ctx->stream;
  const int ncf = in->ncf, m = in->m, nsamp = m * m;
  std::vector<float4> csub, samp;
  std::vector<float> vtab;
  crater_samples(in->crater_vtx, in->crater_facets, ncf, m, csub, samp, vtab);
  DBuf d_cs, d_samp, d_vt;
  upload(d_cs, csub.data(), csub.size(), st);
  upload(d_samp, samp.data(), samp.size(), st);
  upload(d_vt, vtab.data(), vtab.size(), st);
  const size_t shm = (128 * ((size_t)ncf + 1) + (size_t)nsamp + 1) * sizeof(float);
  // facet-step chunks of at most 2^22 entries (the [entries][ncf] outputs of
  // a chunk: 1.6 GB for 100 sub-facets)
  const long long total = (long long)in->nf * in->ntau;
  const long long chunk = 1ll << 22;
  DBuf d_lit, d_sun, d_sh, d_co;
  const long long first = std::min(total, chunk);
  d_lit.alloc((size_t)first * sizeof(int));
  d_sun.alloc((size_t)first * 3 * sizeof(float));
  d_sh.alloc((size_t)first * ncf * sizeof(float));
  d_co.alloc((size_t)first * ncf * sizeof(float));
  for (long long q0 = 0; q0 < total; q0 += chunk) {
    const long long n = std::min(chunk, total - q0);
    HIPCHK(hipMemcpyAsync(d_lit.p, in->shadow + q0, (size_t)n * sizeof(int),
                          hipMemcpyHostToDevice, st));
    HIPCHK(hipMemcpyAsync(d_sun.p, in->sun + 3 * q0, (size_t)n * 3 * sizeof(float),
                          hipMemcpyHostToDevice, st));
    hipLaunchKernelGGL(btpm_crater_sun, dim3((unsigned)((n + 63) / 64)), dim3(256), shm, st,
                       d_lit.as<int>(), d_sun.as<float>(), d_cs.as<float4>(), d_samp.as<float4>(),
                       d_vt.as<float>(), ncf, nsamp, in->crater_radius, n, d_sh.as<float>(),
                       d_co.as<float>());
    HIPCHK(hipGetLastError());
    download_staged(ctx, shadow + q0 * ncf, d_sh.p, (size_t)n * ncf * sizeof(float));
    download_staged(ctx, cosang + q0 * ncf, d_co.p, (size_t)n * ncf * sizeof(float));
  }
  API_END
}
