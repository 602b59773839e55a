// aux.hip -- once-per-solve and once-per-rotation kernels around the
// conduction kernels: layout packing, initial temperatures, shape->crater
// coupling (RGI/RGH), crater and facet sunlight scattering, the facet
// self-heating SpMM and the per-set status reduction.
#include <hip/hip_runtime.h>

#include "btpm_internal.h"
#include "btpm_kernels.h"

// ---------------------------------------------------------------------------
// layout: facet-major x[c][ntau]  <->  blocked x[b][c][BLK]
// ---------------------------------------------------------------------------

// out[b][c][k] = scale[c] * in[c][8b+k]  (scale may be null = 1); steps past
// ntau are padded with the last step's value.
extern "C" __global__ void btpm_pack_blocked(const float* __restrict__ in, float* __restrict__ out,
                                             const float* __restrict__ scale, int ncols, int ntau,
                                             int nb) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= (size_t)ncols * nb) return;
  const int b = (int)(i / ncols);
  const int c = (int)(i - (size_t)b * ncols);
  const float sc = scale ? scale[c] : 1.0f;
  float v[BTPM_BLK];
#pragma unroll
  for (int k = 0; k < BTPM_BLK; ++k) {
    const int s = min(b * BTPM_BLK + k, ntau - 1);
    v[k] = sc * in[(size_t)c * ntau + s];
  }
  float4* o = reinterpret_cast<float4*>(out + ((size_t)b * ncols + c) * BTPM_BLK);
  o[0] = make_float4(v[0], v[1], v[2], v[3]);
  o[1] = make_float4(v[4], v[5], v[6], v[7]);
}

// Every quantity the unpack kernels move out of the device (surface and
// crater temperatures, absorbed fluxes) is non-negative; values are floored
// at +0.  A column that never receives energy cools towards 0 K, where the
// solver's rounding can leave -1e-15 K or -0: a Planck function downstream
// (the Observe stage's exp(hc / (lambda k T)) - 1) turns any T <= -0 into a
// large negative intensity instead of the zero it gives at +0.  NaN passes.
__device__ __forceinline__ float floor0(float v) { return v <= 0.0f ? 0.0f : v; }

extern "C" __global__ void btpm_unpack_blocked(const float* __restrict__ in, float* __restrict__ out,
                                               int ncols, int ntau, int nb) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= (size_t)ncols * nb) return;
  const int b = (int)(i / ncols);
  const int c = (int)(i - (size_t)b * ncols);
  const float* p = in + ((size_t)b * ncols + c) * BTPM_BLK;
#pragma unroll
  for (int k = 0; k < BTPM_BLK; ++k) {
    const int s = b * BTPM_BLK + k;
    if (s < ntau) out[(size_t)c * ntau + s] = floor0(p[k]);
  }
}

// crater reference layout x[f][s][j] <-> blocked x[b][f*NS+j][k]
extern "C" __global__ void btpm_pack_crater(const float* __restrict__ in, float* __restrict__ out,
                                            int nfac, int ntau, int nb) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int ncols = nfac * BTPM_NS;
  if (i >= (size_t)ncols * nb) return;
  const int b = (int)(i / ncols);
  const int c = (int)(i - (size_t)b * ncols);
  const int f = c / BTPM_NS, j = c - f * BTPM_NS;
  float v[BTPM_BLK];
#pragma unroll
  for (int k = 0; k < BTPM_BLK; ++k) {
    const int s = min(b * BTPM_BLK + k, ntau - 1);
    v[k] = in[((size_t)f * ntau + s) * BTPM_NS + j];
  }
  float4* o = reinterpret_cast<float4*>(out + ((size_t)b * ncols + c) * BTPM_BLK);
  o[0] = make_float4(v[0], v[1], v[2], v[3]);
  o[1] = make_float4(v[4], v[5], v[6], v[7]);
}

extern "C" __global__ void btpm_unpack_crater(const float* __restrict__ in, float* __restrict__ out,
                                              int nfac, int ntau, int nb) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int ncols = nfac * BTPM_NS;
  if (i >= (size_t)ncols * nb) return;
  const int b = (int)(i / ncols);
  const int c = (int)(i - (size_t)b * ncols);
  const int f = c / BTPM_NS, j = c - f * BTPM_NS;
  const float* p = in + ((size_t)b * ncols + c) * BTPM_BLK;
#pragma unroll
  for (int k = 0; k < BTPM_BLK; ++k) {
    const int s = b * BTPM_BLK + k;
    if (s < ntau) out[((size_t)f * ntau + s) * BTPM_NS + j] = floor0(p[k]);
  }
}

// facets [c0/100, (c0+nc)/100) of the device layout (step blocks x columns x
// 8 steps) in the crater temperature file's order [column][step]
extern "C" __global__ void btpm_unpack_crater_rows(const float* __restrict__ in,
                                                   float* __restrict__ out, int ncols, int c0,
                                                   int nc, int ntau, int nb) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= (size_t)nc * nb) return;
  const int b = (int)(i / nc);
  const int c = (int)(i - (size_t)b * nc);
  const float* p = in + ((size_t)b * ncols + c0 + c) * BTPM_BLK;
  float* o = out + (size_t)c * ntau + (size_t)b * BTPM_BLK;
#pragma unroll
  for (int k = 0; k < BTPM_BLK; ++k)
    if (b * BTPM_BLK + k < ntau) o[k] = floor0(p[k]);
}

// ---------------------------------------------------------------------------
// initial temperatures
// ---------------------------------------------------------------------------

// initialisetemperatures, T175706/thermal.cpp:1861-1889, from the absorbed
// flux src = (1-A)F (blocked):  T = (mean_s src / es)^(1/4) everywhere.
extern "C" __global__ void btpm_init_smooth(btpm_set_dev S) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= S.ncols) return;
  double sum = 0.0;
  for (int b = 0; b < S.nb; ++b) {
    const float* p = S.src + ((size_t)b * S.ncols + c) * BTPM_BLK;
    const int kend = min(BTPM_BLK, S.ntau - b * BTPM_BLK);
    for (int k = 0; k < kend; ++k) sum += (double)p[k];
  }
  const float t = (float)pow(sum / S.ntau / (double)S.es, 0.25);
  for (int i = 0; i < BTPM_PSTRIDE; ++i) S.prof[(size_t)c * BTPM_PSTRIDE + i] = t;
  const float e = S.esia * t * t * t * t;
  for (int b = 0; b < S.nb; ++b) {
    float* p = S.T + ((size_t)b * S.ncols + c) * BTPM_BLK;
    for (int k = 0; k < BTPM_BLK; ++k) p[k] = t;
  }
  if (S.emis)  // facet-major [c][nb*BLK]
    for (int k = 0; k < S.nb * BTPM_BLK; ++k) S.emis[(size_t)c * S.nb * BTPM_BLK + k] = e;
  S.frozen[c] = 0;
  S.dmax[c] = 0.0f;
}

// Warm start of a smooth set whose T and profile were copied from a finished
// solve (the TI sweep's carried state, T175706/thermal.cpp:3676-3700): the
// self-heating emission of the carried rotation, nothing frozen.
extern "C" __global__ void btpm_warm_smooth(btpm_set_dev S) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= S.ncols) return;
  if (S.emis)  // facet-major [c][nb*BLK] from the blocked T
    for (int b = 0; b < S.nb; ++b) {
      const float* p = S.T + ((size_t)b * S.ncols + c) * BTPM_BLK;
      for (int k = 0; k < BTPM_BLK; ++k) {
        const float t = p[k];
        S.emis[(size_t)c * S.nb * BTPM_BLK + (size_t)b * BTPM_BLK + k] = S.esia * (t * t) * (t * t);
      }
    }
  S.frozen[c] = 0;
  S.dmax[c] = 0.0f;
}

// roughness_initialise_temperatures, T175706/thermal_GPU.cu:374-423: two
// passes per crater; src = (1-A) illum + RGH (blocked).  vf32[i*NS+j] =
// VF(j->i).  One workgroup (128 lanes) per parent facet.
// Memory traffic is coalesced: a lane reads and writes its column's 8-step
// blocks as float4 pairs, and the facet's 100 profile records (contiguous,
// 64 floats each) are written by the whole workgroup as one float4 stream.
extern "C" __global__ __launch_bounds__(128) void btpm_init_crater(btpm_set_dev S,
                                                                   const float* __restrict__ vf32) {
  __shared__ float e1[BTPM_NS];
  __shared__ float tcol[BTPM_NS];
  const int f = blockIdx.x;
  const int j = threadIdx.x;
  const bool ok = j < BTPM_NS;
  const int c = f * BTPM_NS + (ok ? j : 0);
  double sum = 0.0;
  if (ok) {
    // the steps in order, block by block; the full blocks' loads are issued
    // eight blocks at a time (sixteen loads in flight per lane) ahead of the
    // additions, the partial last block is added step by step
    const float* srcc = S.src + (size_t)c * BTPM_BLK;
    const size_t bstr = (size_t)S.ncols * BTPM_BLK;
    const int nfull = S.ntau / BTPM_BLK;
    for (int b0 = 0; b0 < nfull; b0 += 8) {
      float4 x[8][2];
#pragma unroll
      for (int j = 0; j < 8; ++j)
        if (b0 + j < nfull) {
          const float4* p = reinterpret_cast<const float4*>(srcc + (size_t)(b0 + j) * bstr);
          x[j][0] = p[0];
          x[j][1] = p[1];
        }
#pragma unroll
      for (int j = 0; j < 8; ++j)
        if (b0 + j < nfull) {
          sum += (double)x[j][0].x; sum += (double)x[j][0].y;
          sum += (double)x[j][0].z; sum += (double)x[j][0].w;
          sum += (double)x[j][1].x; sum += (double)x[j][1].y;
          sum += (double)x[j][1].z; sum += (double)x[j][1].w;
        }
    }
    if (nfull < S.nb) {
      const float* p = srcc + (size_t)nfull * bstr;
      for (int k = 0; k < S.ntau - nfull * BTPM_BLK; ++k) sum += (double)p[k];
    }
  }
  const double mean = sum / S.ntau;
  const double t1 = pow(mean / (double)S.es, 0.25);
  if (ok) e1[j] = (float)((double)S.esia * t1 * t1 * t1 * t1);
  __syncthreads();
  if (ok) {
    double self = 0.0;
    for (int i = 0; i < BTPM_NS; ++i) self += (double)e1[i] * (double)vf32[i * BTPM_NS + j];
    const float t = (float)pow((mean + self) / (double)S.es, 0.25);
    tcol[j] = t;
    const float4 t4 = make_float4(t, t, t, t);
    for (int b = 0; b < S.nb; ++b) {
      float4* p = reinterpret_cast<float4*>(S.T + ((size_t)b * S.ncols + c) * BTPM_BLK);
      p[0] = t4;
      p[1] = t4;
    }
  }
  __syncthreads();
  static_assert(BTPM_PSTRIDE % 4 == 0, "float4 profile records");
  float4* rec = reinterpret_cast<float4*>(S.prof + (size_t)f * BTPM_NS * BTPM_PSTRIDE);
  for (int e = j; e < BTPM_NS * BTPM_PSTRIDE / 4; e += (int)blockDim.x) {
    const float t = tcol[e / (BTPM_PSTRIDE / 4)];
    rec[e] = make_float4(t, t, t, t);
  }
  if (j == 0) {
    S.frozen[f] = 0;
    S.dmax[f] = 0.0f;
  }
}

// ---------------------------------------------------------------------------
// shape -> crater coupling (load_GPU_roughness_globalillumination /
// _globalheating, T175706/thermal.cpp:2693-2803) as a per-facet GEMM
//   out[j][s] = sum_k drf[k][j] w[g_k][s]        (j < 100, s in a 64-step tile)
// w = weights in facet-major rows w[g][s] (s < nb*BLK, btpm_weights_fm):
//   RGI: w = A_g F_g(s);  RGH: w = es T_g(s)^4.
// A wave's gather of one entry's weights is then one 256-byte run (64
// consecutive steps of one row) instead of eight 32-byte pieces of the
// blocked series a quarter of a cache line each.
// The result is written blocked over crater columns c = f*NS+j as
//   dst[b][c][k] = base_gain * base_scale_f * base[b][c][k] + out   (base may be null).
// Workgroup: 256 lanes = 4 waves per (facet, 64-step tile), on a 1-D grid
// mapped so that all tiles of a facet run on one XCD (workgroups are dealt
// to the 8 XCDs round-robin by id): the facet's drf rows, re-read by each of
// its tiles, then come from that XCD's L2 instead of HBM.
//
// The product runs on v_mfma_f32_16x16x32_f16 with the three-product split
//   drf w ~ drf_hi w_hi + drf_hi w_lo + drf_lo w_hi     (FP32 accumulation),
// x_hi = fp16(x), x_lo = fp16(x - x_hi); drf is multiplied by the facet's
// power-of-two drf_scale (largest |drf| of the facet in [0.5, 1)) so its low
// halves stay normal fp16 numbers, and the accumulator by 1 / drf_scale,
// which is exact.  The facet's entries go through LDS in chunks of 32 as
// drf^T [m = sub-facet][k] and w^T [n = step][k] hi/lo halves: the drf^T
// tiles are packed once per shape (btpm_pack_drf) and copied in 16-byte
// pieces, the w^T halves are split here, a thread writing 8 consecutive k of
// one step as one 16-byte store (the gathers stay coalesced: lanes run over
// steps); the next chunk is loaded into registers while the current one is
// on the matrix cores.  Wave
// w owns the 16 steps of n tile w and all 7 m tiles (M = 100 padded to 112),
// so each w fragment is read once per chunk.  Fragment layout: lane l
// supplies A[m = l%16][k = 8(l/16) .. +7] and B[k = 8(l/16) .. +7][n = l%16]
// and receives D[m = 4(l/16) + i][n = l%16], i = 0..3.
// ---------------------------------------------------------------------------
namespace {
constexpr int kRK = 32;        // entries per K chunk (one 16x16x32 product)
constexpr int kRS = kRK + 8;   // LDS row stride in halves: 20 dwords, so the 8
                               // rows of one 16-byte access hit distinct bank quads
constexpr int kRM = 112;       // padded sub-facets (7 m tiles)
constexpr int kRN = 64;        // steps per workgroup (4 n tiles, one per wave)
constexpr int kRDrfTile = 2 * kRM * kRK;  // halves per packed drf^T chunk tile [2][112][32]
static_assert(kRDrfTile == BTPM_DRF_TILE, "packed drf tile size");
typedef _Float16 rhalf8 __attribute__((ext_vector_type(8)));
typedef float rfloat4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ void split8(const float (&v)[8], float sc, rhalf8& hi, rhalf8& lo) {
#pragma unroll
  for (int i = 0; i < 8; ++i) {
    const float x = v[i] * sc;
    const _Float16 h = (_Float16)x;
    hi[i] = h;
    lo[i] = (_Float16)(x - (float)h);
  }
}
}  // namespace

extern "C" __global__ __launch_bounds__(256) void btpm_shape_to_crater(
    const int* __restrict__ rowptr, const int* __restrict__ colidx,
    const _Float16* __restrict__ drfA, const int* __restrict__ cbase,
    const float* __restrict__ drf_scale, const float* __restrict__ w, int nfac, int ntau,
    int nb, const float* __restrict__ base,
    const float* __restrict__ base_scale, float base_gain, float* __restrict__ dst) {
  constexpr int kMT = kRM / 16;
  constexpr int kAPieces = kRDrfTile / 8;  // 16-byte pieces of a packed drf^T tile: 896
  constexpr int kAPer = (kAPieces + 255) / 256;
  static_assert(kRN * (kRK / 8) == 256, "staging shape");
  __shared__ __align__(16) _Float16 la[2][kRM * kRS];  // drf^T hi / lo
  __shared__ __align__(16) _Float16 lb[2][kRN * kRS];  // w^T hi / lo
  const int ntile = (nb * BTPM_BLK + kRN - 1) / kRN;
  const int xcd = (int)(blockIdx.x & 7), local = (int)(blockIdx.x >> 3);
  const int f = 8 * (local / ntile) + xcd;
  const int s0 = kRN * (local % ntile);
  if (f >= nfac) return;
  const int tid = (int)threadIdx.x;
  const int wave = tid >> 6, lane = tid & 63;
  const int r = lane & 15, g = lane >> 4;
  const int ncols = nfac * BTPM_NS;
  const float unsc = 1.0f / drf_scale[f];  // exact: a power of two
  const rhalf8* __restrict__ tiles = reinterpret_cast<const rhalf8*>(drfA) +
                                     (size_t)cbase[f] * (kRDrfTile / 8);
  const float bsc = base_gain * (base_scale ? base_scale[f] : 1.0f);
  const int k0 = rowptr[f], k1 = rowptr[f + 1];
  // staging roles: drf^T tile pieces p = tid + 256 q (a straight copy of the
  // packed tile); w^T row tid % 64 (a step), group tid / 64 = wave
  const int sb = min(s0 + (tid & (kRN - 1)), ntau - 1);
  const int npad = nb * BTPM_BLK;
  rhalf8 ra[kAPer];
  float rb[8];
  // the wave's 8 entry indices of a chunk are wave-uniform; read through an
  // opaque zero VGPR they come by vector loads, on the vector memory
  // counter: as scalar loads they would share the LDS counter, and every
  // wait for the products' LDS reads would wait for them too (scalar loads
  // complete out of order).  Indices run one chunk ahead of the gathers that
  // use them; every index is clamped into the facet's row, and gathers past
  // its end read a valid weight that is zeroed where it is used
  int vz;
  asm volatile("v_mov_b32 %0, 0" : "=v"(vz));
  int ci[8], cn[8];
  auto load_idx = [&](int kb, int (&dst)[8]) {
#pragma unroll
    for (int i = 0; i < 8; ++i) dst[i] = colidx[min(kb + 8 * wave + i, k1 - 1) + vz];
  };
  auto fetch = [&](int kb) {
    const rhalf8* __restrict__ tile = tiles + (size_t)((kb - k0) / kRK) * (kRDrfTile / 8);
#pragma unroll
    for (int q = 0; q < kAPer; ++q) {
      const int p = tid + 256 * q;
      if (p < kAPieces) ra[q] = tile[p];
    }
#pragma unroll
    for (int i = 0; i < 8; ++i) rb[i] = w[(size_t)ci[i] * npad + sb];
    load_idx(kb + kRK, cn);
  };
  rfloat4 acc[kMT];
#pragma unroll
  for (int t = 0; t < kMT; ++t) acc[t] = rfloat4{0.0f, 0.0f, 0.0f, 0.0f};
  if (k0 < k1) {
    load_idx(k0, ci);
    fetch(k0);
  }
  for (int kb = k0; kb < k1; kb += kRK) {
    __syncthreads();  // the previous chunk's products have read LDS
#pragma unroll
    for (int q = 0; q < kAPer; ++q) {
      const int p = tid + 256 * q;  // piece (h, m, 8-entry group) of the [2][112][32] tile
      if (p < kAPieces) {
        const int h = p / (kAPieces / 2), rem = p - h * (kAPieces / 2);
        *reinterpret_cast<rhalf8*>(&la[h][(rem / (kRK / 8)) * kRS + 8 * (rem % (kRK / 8))]) = ra[q];
      }
    }
    {
      const int kn = min(kRK, k1 - kb);
      float v[8];
#pragma unroll
      for (int i = 0; i < 8; ++i) v[i] = 8 * wave + i < kn ? rb[i] : 0.0f;
      rhalf8 hi, lo;
      split8(v, 1.0f, hi, lo);
      const int n = tid & (kRN - 1);
      *reinterpret_cast<rhalf8*>(&lb[0][n * kRS + 8 * wave]) = hi;
      *reinterpret_cast<rhalf8*>(&lb[1][n * kRS + 8 * wave]) = lo;
    }
    __syncthreads();
    if (kb + kRK < k1) {
#pragma unroll
      for (int i = 0; i < 8; ++i) ci[i] = cn[i];
      fetch(kb + kRK);
    }
    const int bn = 16 * wave + r;
    const rhalf8 bhi = *reinterpret_cast<const rhalf8*>(&lb[0][bn * kRS + 8 * g]);
    const rhalf8 blo = *reinterpret_cast<const rhalf8*>(&lb[1][bn * kRS + 8 * g]);
#pragma unroll
    for (int t = 0; t < kMT; ++t) {
      const int am = 16 * t + r;
      const rhalf8 ahi = *reinterpret_cast<const rhalf8*>(&la[0][am * kRS + 8 * g]);
      const rhalf8 alo = *reinterpret_cast<const rhalf8*>(&la[1][am * kRS + 8 * g]);
      acc[t] = __builtin_amdgcn_mfma_f32_16x16x32_f16(alo, bhi, acc[t], 0, 0, 0);
      acc[t] = __builtin_amdgcn_mfma_f32_16x16x32_f16(ahi, blo, acc[t], 0, 0, 0);
      acc[t] = __builtin_amdgcn_mfma_f32_16x16x32_f16(ahi, bhi, acc[t], 0, 0, 0);
    }
  }
  const int s = s0 + 16 * wave + r;
  if (s >= nb * BTPM_BLK) return;
#pragma unroll
  for (int t = 0; t < kMT; ++t) {
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      const int j = 16 * t + 4 * g + i;
      if (j >= BTPM_NS) continue;
      const int c = f * BTPM_NS + j;
      const size_t o = ((size_t)(s / BTPM_BLK) * ncols + c) * BTPM_BLK + (s % BTPM_BLK);
      float v = acc[t][i] * unsc;
      if (base) v += bsc * base[o];
      dst[o] = v;
    }
  }
}

// Weights of btpm_shape_to_crater as facet-major rows out[c][s] (s < nb*BLK)
// from a blocked [nb][ncols][BLK] series x: mode 0 gives (gain * scale[c]) x
// (RGI: A_c F_c), mode 1 gives es x^4 (RGH: the emission of the smooth
// surface temperatures), each value computed once here instead of once per
// view-factor entry that gathers it.  One thread per (block, column).
extern "C" __global__ __launch_bounds__(256) void btpm_weights_fm(const float* __restrict__ x,
                                                                 const float* __restrict__ scale,
                                                                 float gain, int mode, float es,
                                                                 int ncols, int nb,
                                                                 float* __restrict__ out) {
  const size_t i = (size_t)blockIdx.x * 256 + threadIdx.x;  // b * ncols + c
  if (i >= (size_t)nb * ncols) return;
  const int b = (int)(i / ncols), c = (int)(i % ncols);
  const float4 a0 = *reinterpret_cast<const float4*>(x + i * BTPM_BLK);
  const float4 a1 = *reinterpret_cast<const float4*>(x + i * BTPM_BLK + 4);
  float v[BTPM_BLK] = {a0.x, a0.y, a0.z, a0.w, a1.x, a1.y, a1.z, a1.w};
  const float sc = mode == 1 ? 0.0f : gain * scale[c];
#pragma unroll
  for (int k = 0; k < BTPM_BLK; ++k) v[k] = mode == 1 ? es * (v[k] * v[k]) * (v[k] * v[k]) : sc * v[k];
  float* o = out + (size_t)c * nb * BTPM_BLK + (size_t)b * BTPM_BLK;
  *reinterpret_cast<float4*>(o) = make_float4(v[0], v[1], v[2], v[3]);
  *reinterpret_cast<float4*>(o + 4) = make_float4(v[4], v[5], v[6], v[7]);
}

// Packs the shape->crater view factors once per shape for btpm_shape_to_crater:
// facet f's entries rowptr[f] .. rowptr[f+1]-1 in chunks of 32, chunk c as the
// tile drfA[cbase[f] + c] = [hi, lo][m = sub-facet, 112][k = entry in chunk, 32]
// of drf * drf_scale[f], where drf_scale[f] is the power of two that puts the
// facet's largest |drf| in [0.5, 1) (zero padding for m >= 100 and for entries
// past the facet's last).  One workgroup per facet; each chunk goes through
// LDS so the fp32 rows are read and the fp16 tiles written contiguously.
extern "C" __global__ __launch_bounds__(256) void btpm_pack_drf(const int* __restrict__ rowptr,
                                                               const float* __restrict__ drf,
                                                               const int* __restrict__ cbase,
                                                               float* __restrict__ drf_scale,
                                                               _Float16* __restrict__ drfA) {
  __shared__ float red[256];
  __shared__ float chunk[kRK][BTPM_NS + 1];
  const int f = blockIdx.x;
  const int tid = (int)threadIdx.x;
  const int k0 = rowptr[f], k1 = rowptr[f + 1];
  float mx = 0.0f;
  for (size_t e = (size_t)k0 * BTPM_NS + tid; e < (size_t)k1 * BTPM_NS; e += 256)
    mx = fmaxf(mx, fabsf(drf[e]));
  red[tid] = mx;
  __syncthreads();
  for (int h = 128; h > 0; h >>= 1) {
    if (tid < h) red[tid] = fmaxf(red[tid], red[tid + h]);
    __syncthreads();
  }
  mx = red[0];
  float sc = 1.0f;
  if (mx > 0.0f && mx <= 3.0e38f) {
    int ex = 0;
    (void)frexpf(mx, &ex);  // mx = m 2^ex, m in [0.5, 1)
    sc = ldexpf(1.0f, max(-120, min(120, -ex)));
  }
  if (tid == 0) drf_scale[f] = sc;
  for (int kb = k0, c = 0; kb < k1; kb += kRK, ++c) {
    const int kn = min(kRK, k1 - kb);
    __syncthreads();
    for (int e = tid; e < kRK * BTPM_NS; e += 256) {
      const int k = e / BTPM_NS, m = e - (e / BTPM_NS) * BTPM_NS;
      chunk[k][m] = k < kn ? drf[(size_t)(kb + k) * BTPM_NS + m] * sc : 0.0f;
    }
    __syncthreads();
    _Float16* tile = drfA + (size_t)(cbase[f] + c) * kRDrfTile;
    for (int e = tid; e < kRM * kRK; e += 256) {
      const int m = e / kRK, k = e - (e / kRK) * kRK;
      const float x = m < BTPM_NS ? chunk[k][m] : 0.0f;
      const _Float16 hi = (_Float16)x;
      tile[e] = hi;
      tile[kRM * kRK + e] = (_Float16)(x - (float)hi);
    }
  }
}

// ---------------------------------------------------------------------------
// crater sunlight scattering (roughness_single_scattering /
// roughness_multiple_scattering, T175706/thermal_GPU.cu:241-341).
// Direct sunlight comes either from shadow, cosang [f][s][j] or, when sun is
// given, from the Sun direction in the facet's crater frame sun[f][s][3]
// (its length is the facet's own sunlit fraction, 0 when eclipsed): the
// sub-facet's sunlit fraction is interpolated bilinearly in vis_table
// [n_el][n_az][100] (elevation i*90/(n_el-1) deg, azimuth i*360/n_az deg)
// and its cosine is max(0, n_j . u) with the unit direction u.
// rgi [f][s][j] may be null.  Output illum blocked [b][f*NS+j][k].  One
// workgroup (128 lanes) per (facet, block): lane j holds the 8 steps of
// sub-facet j.
// ---------------------------------------------------------------------------
namespace {
__device__ __forceinline__ float direct_from_sun(const float* __restrict__ d3,
                                                 const float* __restrict__ vis, int n_el,
                                                 int n_az, float nx, float ny, float nz, int j) {
  const float dx = d3[0], dy = d3[1], dz = d3[2];
  const float len = sqrtf(dx * dx + dy * dy + dz * dz);
  if (!(len > 0.0f) || !(dz > 0.0f)) return 0.0f;
  const float il = 1.0f / len;
  const double ux = (double)(dx * il), uy = (double)(dy * il), uz = (double)(dz * il);
  const double kPi = 3.14159265358979323846;
  const double el = asin(fmin(uz, 1.0));
  double az = atan2(uy, ux);
  if (az < 0.0) az += 2.0 * kPi;
  const double fe = fmin(el / (0.5 * kPi / (n_el - 1)), (double)(n_el - 1) - 1e-9);
  const double fa = az / (2.0 * kPi / n_az);
  const int ie = (int)floor(fe);
  const int ia = ((int)floor(fa)) % n_az;
  const float we = (float)(fe - ie), wa = (float)(fa - floor(fa));
  const int ie1 = min(ie + 1, n_el - 1), ia1 = (ia + 1) % n_az;
  const float v00 = vis[((size_t)ie * n_az + ia) * BTPM_NS + j];
  const float v01 = vis[((size_t)ie * n_az + ia1) * BTPM_NS + j];
  const float v10 = vis[((size_t)ie1 * n_az + ia) * BTPM_NS + j];
  const float v11 = vis[((size_t)ie1 * n_az + ia1) * BTPM_NS + j];
  const float v = (1.0f - we) * ((1.0f - wa) * v00 + wa * v01) +
                  we * ((1.0f - wa) * v10 + wa * v11);
  const float c = (float)ux * nx + (float)uy * ny + (float)uz * nz;
  return (v * len) * (c > 0.0f ? c : 0.0f);
}
}  // namespace

extern "C" __global__ __launch_bounds__(128) void btpm_crater_scatter(
    const float* __restrict__ vf32, const float* __restrict__ albedo,
    const float* __restrict__ shadow, const float* __restrict__ cosang,
    const float* __restrict__ sun, const float* __restrict__ vis_table, int n_el, int n_az,
    const float* __restrict__ normals, const float* __restrict__ rgi, float solar, int multiple,
    float acc, int max_iter, int nfac, int ntau, int nb, float* __restrict__ illum_out) {
  __shared__ float vfs[BTPM_NS * BTPM_NS];
  __shared__ float il[BTPM_BLK][BTPM_NS];
  const int f = blockIdx.y;
  const int b = blockIdx.x;
  const int j = threadIdx.x;
  const bool ok = j < BTPM_NS;
  for (int e = j; e < BTPM_NS * BTPM_NS; e += 128) vfs[e] = vf32[e];
  const float a = albedo[f];
  float nx = 0.0f, ny = 0.0f, nz = 0.0f;
  if (sun && ok) {
    nx = normals[3 * j];
    ny = normals[3 * j + 1];
    nz = normals[3 * j + 2];
  }
  float sp[BTPM_BLK], res[BTPM_BLK];
#pragma unroll
  for (int k = 0; k < BTPM_BLK; ++k) {
    const int s = min(b * BTPM_BLK + k, ntau - 1);
    sp[k] = 0.0f;
    if (ok) {
      const size_t o = ((size_t)f * ntau + s) * BTPM_NS + j;
      const float direct =
          sun ? direct_from_sun(sun + ((size_t)f * ntau + s) * 3, vis_table, n_el, n_az, nx, ny,
                                nz, j)
              : shadow[o] * cosang[o];
      sp[k] = solar * direct + (rgi ? rgi[o] : 0.0f);
      il[k][j] = sp[k];
    }
  }
  __syncthreads();
  // single scattering: tot_j = il_j + A sum_i VF(j->i) il_i, clamped >= 0
#pragma unroll
  for (int k = 0; k < BTPM_BLK; ++k) {
    float t = sp[k];
    if (ok)
      for (int i = 0; i < BTPM_NS; ++i) t = fmaf(vfs[i * BTPM_NS + j] * a, il[k][i], t);
    res[k] = t > 0.0f ? t : 0.0f;
  }
  if (multiple && a > 0.0f) {
    // Jacobi on p = A (sp + VF p) from p = A * single; stop when every
    // |p_new - p_old| <= acc over the block's 8 steps
    __syncthreads();
#pragma unroll
    for (int k = 0; k < BTPM_BLK; ++k)
      if (ok) il[k][j] = a * res[k];
    for (int it = 0; it < max_iter; ++it) {
      __syncthreads();  // the a*res / previous sweep writes are visible
      float pn[BTPM_BLK];
      int ch = 0;
#pragma unroll
      for (int k = 0; k < BTPM_BLK; ++k) {
        float sum = 0.0f;
        if (ok)
          for (int i = 0; i < BTPM_NS; ++i) sum = fmaf(vfs[i * BTPM_NS + j], il[k][i], sum);
        pn[k] = a * (sp[k] + sum);
        if (ok && fabsf(pn[k] - il[k][j]) > acc) ch = 1;
      }
      __syncthreads();  // every lane has read the old sweep
#pragma unroll
      for (int k = 0; k < BTPM_BLK; ++k)
        if (ok) il[k][j] = pn[k];
      if (!__syncthreads_or(ch)) break;  // block-uniform exit
    }
#pragma unroll
    for (int k = 0; k < BTPM_BLK; ++k) res[k] = ok ? il[k][j] / a : 0.0f;
  } else if (multiple) {
#pragma unroll
    for (int k = 0; k < BTPM_BLK; ++k) res[k] = sp[k];
  }
  if (!ok) return;
  const int ncols = nfac * BTPM_NS;
  float4* o = reinterpret_cast<float4*>(illum_out + ((size_t)b * ncols + f * BTPM_NS + j) * BTPM_BLK);
  o[0] = make_float4(res[0], res[1], res[2], res[3]);
  o[1] = make_float4(res[4], res[5], res[6], res[7]);
}

// ---------------------------------------------------------------------------
// sparse facet coupling over blocked time series (CSR rows = receivers):
//   y[b][f][k] = alpha_f * (x0[b][f][k] + sum_{(f,g)} vf * x[b][g][k])
// x0, alpha may be null (0 and 1).  Lane = receiving facet, each lane covers
// BPT consecutive time blocks.  Used for the self-heating flux of the next
// rotation (x = emis) and for the facet scattering sweeps.
// ---------------------------------------------------------------------------
extern "C" __global__ __launch_bounds__(256) void btpm_spmm_blocked(
    const int* __restrict__ rowptr, const int* __restrict__ colidx, const float* __restrict__ vals,
    const float* __restrict__ x, const float* __restrict__ x0, const float* __restrict__ alpha,
    float* __restrict__ y, int nrows, int nb, int bpt, const btpm_status_dev* __restrict__ st) {
  if (st && st->done) return;
  const int f = blockIdx.x * blockDim.x + threadIdx.x;
  const int b0 = blockIdx.y * bpt;
  if (f >= nrows) return;
  const int k0 = rowptr[f], k1 = rowptr[f + 1];
  const float al = alpha ? alpha[f] : 1.0f;
  for (int bb = b0; bb < min(b0 + bpt, nb); bb += 4) {
    const int nbb = min(4, min(b0 + bpt, nb) - bb);
    float acc[4][BTPM_BLK];
#pragma unroll
    for (int q = 0; q < 4; ++q)
#pragma unroll
      for (int k = 0; k < BTPM_BLK; ++k) acc[q][k] = 0.0f;
    for (int kk = k0; kk < k1; ++kk) {
      const int g = colidx[kk];
      const float v = vals[kk];
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        if (q < nbb) {
          const float4* p = reinterpret_cast<const float4*>(x + ((size_t)(bb + q) * nrows + g) * BTPM_BLK);
          const float4 lo = p[0], hi = p[1];
          acc[q][0] = fmaf(v, lo.x, acc[q][0]);
          acc[q][1] = fmaf(v, lo.y, acc[q][1]);
          acc[q][2] = fmaf(v, lo.z, acc[q][2]);
          acc[q][3] = fmaf(v, lo.w, acc[q][3]);
          acc[q][4] = fmaf(v, hi.x, acc[q][4]);
          acc[q][5] = fmaf(v, hi.y, acc[q][5]);
          acc[q][6] = fmaf(v, hi.z, acc[q][6]);
          acc[q][7] = fmaf(v, hi.w, acc[q][7]);
        }
      }
    }
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      if (q < nbb) {
        const size_t o = ((size_t)(bb + q) * nrows + f) * BTPM_BLK;
        float r[BTPM_BLK];
#pragma unroll
        for (int k = 0; k < BTPM_BLK; ++k) r[k] = al * ((x0 ? x0[o + k] : 0.0f) + acc[q][k]);
        float4* yp = reinterpret_cast<float4*>(y + o);
        yp[0] = make_float4(r[0], r[1], r[2], r[3]);
        yp[1] = make_float4(r[4], r[5], r[6], r[7]);
      }
    }
  }
}

// Facet self-heating of the next rotation for every smooth set of one shape
// in one launch (grid z = set):
//   qext[b][f][k] = src[b][f][k] + sum_g VF_fg emis[g][8b + k]
// A workgroup owns 256 receiving facets (lane = facet) x 64 steps of one set
// and walks the emitters in chunks of 256: each chunk's emission rows are
// staged in LDS once for all 256 receivers (row stride 68 floats keeps the
// lanes' 16-byte reads of different rows on different banks), and each lane
// accumulates its 64 steps in registers from its CSR entries that fall in
// the chunk (entries sorted by emitter, packed as (emitter, view factor)
// pairs; cptr[c][f] = first entry of row f in chunk c).  The staged emission is reused by every receiver of the block,
// instead of being gathered from L2 once per view factor.
namespace {
// Global-memory access to a set's arrays.  Through a generic pointer (the
// set table holds generic ones) loads compile to FLAT instructions, which
// count on the LDS / scalar-load wait counter as well: every wait for an LDS
// read or a scalar load would then also wait for the next chunk's emission
// rows from HBM, the prefetch meant to overlap it.
typedef float gvec4 __attribute__((ext_vector_type(4)));
__device__ __forceinline__ float4 gload4(const float* p) {
  const gvec4 v = *reinterpret_cast<const __attribute__((address_space(1))) gvec4*>(
      (const __attribute__((address_space(1))) float*)p);
  return make_float4(v.x, v.y, v.z, v.w);
}
__device__ __forceinline__ void gstore4(float* p, float4 v) {
  *reinterpret_cast<__attribute__((address_space(1))) gvec4*>(
      (__attribute__((address_space(1))) float*)p) = gvec4{v.x, v.y, v.z, v.w};
}
__device__ __forceinline__ float gload1(const float* p) {
  return *(const __attribute__((address_space(1))) float*)p;
}
__device__ __forceinline__ void gstore1(float* p, float v) {
  *(__attribute__((address_space(1))) float*)p = v;
}
constexpr int kSHRows = BTPM_SH_ROWS;    // receivers per workgroup (one per lane)
constexpr int kSHChunk = BTPM_SH_CHUNK;  // emitters per staged chunk
constexpr int kSHSteps = BTPM_SH_STEPS;  // steps per workgroup
constexpr int kSHStride = BTPM_SH_STRIDE;  // LDS row stride in floats
}  // namespace

extern "C" __global__ __launch_bounds__(kSHRows) void btpm_selfheat_tiled(
    const int* __restrict__ cptr, const int2* __restrict__ ent, int nf, int nchunk,
    const int* __restrict__ slots, const btpm_set_dev* __restrict__ sets,
    const btpm_status_dev* __restrict__ status) {
  __shared__ __align__(16) float E[kSHChunk * kSHStride];
  const int slot = slots[blockIdx.z];
  if (status[slot].done) return;
  const btpm_set_dev& S = sets[slot];
  const int npad = S.nb * BTPM_BLK;
  const int s0 = blockIdx.y * kSHSteps;
  if (s0 >= npad) return;
  const int tid = (int)threadIdx.x;
  const int f = blockIdx.x * kSHRows + tid;
  const bool live = f < nf && (S.inst || !S.frozen[f]);
  if (__syncthreads_count(live) == 0) return;  // every receiver here is frozen
  const float* __restrict__ emis = S.emis;
  float acc[kSHSteps];
#pragma unroll
  for (int q = 0; q < kSHSteps; ++q) acc[q] = 0.0f;
  for (int c = 0; c < nchunk; ++c) {
    const int g0 = c * kSHChunk;
    __syncthreads();  // the previous chunk has been consumed
    for (int e = tid; e < kSHChunk * (kSHSteps / 4); e += kSHRows) {
      const int gl = e / (kSHSteps / 4), q4 = 4 * (e % (kSHSteps / 4));
      float4 v = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
      if (g0 + gl < nf && s0 + q4 < npad)
        v = gload4(emis + (size_t)(g0 + gl) * npad + s0 + q4);
      *reinterpret_cast<float4*>(&E[gl * kSHStride + q4]) = v;
    }
    __syncthreads();
    if (live) {
      const int k0 = cptr[(size_t)c * nf + f], k1 = cptr[(size_t)(c + 1) * nf + f];
      for (int k = k0; k < k1; ++k) {
        const int2 gv = ent[k];
        const float v = __int_as_float(gv.y);
        const float* er = &E[(gv.x - g0) * kSHStride];
#pragma unroll
        for (int q = 0; q < kSHSteps; q += 4) {
          const float4 e4 = *reinterpret_cast<const float4*>(er + q);
          acc[q] = fmaf(v, e4.x, acc[q]);
          acc[q + 1] = fmaf(v, e4.y, acc[q + 1]);
          acc[q + 2] = fmaf(v, e4.z, acc[q + 2]);
          acc[q + 3] = fmaf(v, e4.w, acc[q + 3]);
        }
      }
    }
  }
  if (!live) return;
  const float* __restrict__ src = S.src;
  float* __restrict__ y = S.qext;
#pragma unroll
  for (int bq = 0; bq < kSHSteps / BTPM_BLK; ++bq) {
    if (s0 + bq * BTPM_BLK >= npad) break;
    const size_t o = ((size_t)(s0 / BTPM_BLK + bq) * nf + f) * BTPM_BLK;
    const float4 a = gload4(src + o);
    const float4 b = gload4(src + o + 4);
    const int q = bq * BTPM_BLK;
    gstore4(y + o,
            make_float4(a.x + acc[q], a.y + acc[q + 1], a.z + acc[q + 2], a.w + acc[q + 3]));
    gstore4(y + o + 4,
            make_float4(b.x + acc[q + 4], b.y + acc[q + 5], b.z + acc[q + 6], b.w + acc[q + 7]));
  }
}

// The same product in the lane-interleaved layout: for receiver block rb,
// emitter chunk c and wave w, wseg[(rb * nchunk + c) * 4 + w] = (first, n)
// and entry k of lane l is ent[first + 64 k + l] = (its emitter's index in
// the chunk, view-factor bits), lanes with fewer than n entries padded with
// (0, +0.0f).  H steps per workgroup (64 or 32: one body, two entry points;
// 32 halves the staged chunk's LDS and doubles the workgroups per CU).  Each entry load of a wave is one 512-byte
// contiguous read whose address is known ahead, so the next two are
// prefetched; each lane still adds its entries in emitter order, so the sums
// equal btpm_selfheat_tiled's bit for bit.  The next chunk's emission rows
// are fetched into registers while the current chunk is consumed.
template <int H>
__device__ __forceinline__ void selfheat_ell_body(const int2* __restrict__ wseg,
                                                  const int2* __restrict__ ent, int nf,
                                                  int nchunk, const int* __restrict__ slots,
                                                  const btpm_set_dev* __restrict__ sets,
                                                  const btpm_status_dev* __restrict__ status) {
  constexpr int kStride = H + 4;                     // LDS row stride in floats
  constexpr int kQ4 = H / 4;                         // float4 per emission row
  constexpr int kPer = kSHChunk * kQ4 / kSHRows;     // float4 staged per lane
  static_assert(kSHChunk * kQ4 % kSHRows == 0 && kSHRows % kQ4 == 0, "staging shape");
  __shared__ __align__(16) float E[kSHChunk * kStride];
  const int slot = slots[blockIdx.z];
  if (status[slot].done) return;
  const btpm_set_dev& S = sets[slot];
  const int npad = S.nb * BTPM_BLK;
  const int s0 = blockIdx.y * H;
  if (s0 >= npad) return;
  const int tid = (int)threadIdx.x;
  const int f = blockIdx.x * kSHRows + tid;
  const bool live = f < nf && (S.inst || !S.frozen[f]);
  if (__syncthreads_count(live) == 0) return;  // every receiver here is frozen
  const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
  const int lane = tid & 63;
  const bool wave_live = __ballot(live) != 0;
  const float* __restrict__ emis = S.emis;
  // staging: lane tid holds rows tid / kQ4 + i * (kSHRows / kQ4), float4 column tid % kQ4
  const int srow = tid / kQ4, sq4 = 4 * (tid % kQ4);
  const bool scol = s0 + sq4 < npad;
  float4 pre[kPer];
  auto fetch = [&](int c) {
    const int g0 = c * kSHChunk;
#pragma unroll
    for (int i = 0; i < kPer; ++i) {
      const int g = g0 + srow + i * (kSHRows / kQ4);
      pre[i] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
      if (g < nf && scol) pre[i] = gload4(emis + (size_t)g * npad + s0 + sq4);
    }
  };
  float acc[H];
#pragma unroll
  for (int q = 0; q < H; ++q) acc[q] = 0.0f;
  // the wave's segment (first, n) of a chunk is wave-uniform; read through
  // an opaque zero VGPR it comes by a vector load, one chunk ahead (a scalar
  // load would share the LDS wait counter, and the chunk's first LDS wait
  // would drain it)
  int vz;
  asm volatile("v_mov_b32 %0, 0" : "=v"(vz));
  const int2* __restrict__ wsb = wseg + (size_t)blockIdx.x * nchunk * BTPM_SH_WAVES + wave;
  int2 dn = wsb[vz];
  fetch(0);
  for (int c = 0; c < nchunk; ++c) {
    __syncthreads();  // the previous chunk has been consumed
#pragma unroll
    for (int i = 0; i < kPer; ++i)
      *reinterpret_cast<float4*>(&E[(srow + i * (kSHRows / kQ4)) * kStride + sq4]) = pre[i];
    if (c + 1 < nchunk) fetch(c + 1);
    const int2 dc = dn;
    dn = wsb[(size_t)min(c + 1, nchunk - 1) * BTPM_SH_WAVES + vz];
    __syncthreads();
    const int n = wave_live ? __builtin_amdgcn_readfirstlane(dc.y) : 0;
    if (n == 0) continue;
    const int2* __restrict__ e = ent + (size_t)__builtin_amdgcn_readfirstlane(dc.x) + lane;
    // entries are prefetched three ahead by unconditional loads (the index
    // clamped to the segment's last entry) into the registers of the entry
    // just consumed (the loop is unrolled by three, so no copy waits for a
    // load); the empty asm with a memory clobber keeps each load where it is
    // written instead of next to its use, and the one on the emitter index
    // keeps that entry's address arithmetic (and its wait) at its own step
    const int last = n - 1;
    auto step = [&](const int2& gv) {
      int gx = gv.x;
      asm volatile("" : "+v"(gx));
      const float v = __int_as_float(gv.y);
      const float* er = &E[gx * kStride];
#pragma unroll
      for (int q = 0; q < H; q += 4) {
        const float4 e4 = *reinterpret_cast<const float4*>(er + q);
        acc[q] = fmaf(v, e4.x, acc[q]);
        acc[q + 1] = fmaf(v, e4.y, acc[q + 1]);
        acc[q + 2] = fmaf(v, e4.z, acc[q + 2]);
        acc[q + 3] = fmaf(v, e4.w, acc[q + 3]);
      }
    };
    int2 qa = e[0], qb = e[(size_t)min(1, last) * 64], qc = e[(size_t)min(2, last) * 64];
    asm volatile("" ::: "memory");
    int k = 0;
    for (; k + 3 <= n; k += 3) {
      step(qa);
      qa = e[(size_t)min(k + 3, last) * 64];
      asm volatile("" ::: "memory");
      step(qb);
      qb = e[(size_t)min(k + 4, last) * 64];
      asm volatile("" ::: "memory");
      step(qc);
      qc = e[(size_t)min(k + 5, last) * 64];
      asm volatile("" ::: "memory");
    }
    if (k < n) step(qa);
    if (k + 1 < n) step(qb);
  }
  if (!live) return;
  const float* __restrict__ src = S.src;
  float* __restrict__ y = S.qext;
#pragma unroll
  for (int bq = 0; bq < H / BTPM_BLK; ++bq) {
    if (s0 + bq * BTPM_BLK >= npad) break;
    const size_t o = ((size_t)(s0 / BTPM_BLK + bq) * nf + f) * BTPM_BLK;
    const float4 a = gload4(src + o);
    const float4 b = gload4(src + o + 4);
    const int q = bq * BTPM_BLK;
    gstore4(y + o,
            make_float4(a.x + acc[q], a.y + acc[q + 1], a.z + acc[q + 2], a.w + acc[q + 3]));
    gstore4(y + o + 4,
            make_float4(b.x + acc[q + 4], b.y + acc[q + 5], b.z + acc[q + 6], b.w + acc[q + 7]));
  }
}

extern "C" __global__ __launch_bounds__(kSHRows) void btpm_selfheat_ell(
    const int2* wseg, const int2* ent, int nf, int nchunk, const int* slots,
    const btpm_set_dev* sets, const btpm_status_dev* status) {
  selfheat_ell_body<64>(wseg, ent, nf, nchunk, slots, sets, status);
}
extern "C" __global__ __launch_bounds__(kSHRows) void btpm_selfheat_ell_h32(
    const int2* wseg, const int2* ent, int nf, int nchunk, const int* slots,
    const btpm_set_dev* sets, const btpm_status_dev* status) {
  selfheat_ell_body<32>(wseg, ent, nf, nchunk, slots, sets, status);
}

// The same product with a wave's 64 receivers taken 8 at a time.  In pass p
// (0..7) of wave w the lanes 8g .. 8g+7 work for the receiver in slot
// 64w + 8p + g of the block, lane s = l % 8 on steps 4s + 32j .. 4s + 32j + 3
// (j < H/32) of the workgroup's H steps.  Each LDS read of a group is one
// 128-byte run of a staged row (the one-receiver-per-lane kernels read 64
// random rows at once and collide on banks), and a pass walks the longest of
// 8 receivers' entries in the chunk instead of the longest of 64.  A block's receivers sit in slots by decreasing entry count
// (gperm[rb * 256 + slot] = receiver, nf for an empty slot), so long rows
// share passes.  For receiver block rb, chunk c, wave w and pass p,
// gseg[((rb * nchunk + c) * 4 + w) * 8 + p] = (first, n): n is the longest
// group's entry count in the chunk rounded up to even, and entries k, k + 1
// of group g are the int4 gent[first + 8 (k / 2) + g] = (emitter within the
// chunk, view-factor bits) twice, shorter groups padded with (0, +0.0f).
// Each receiver still adds its entries in emitter order (a padded entry adds
// +0 x E >= 0, which leaves the sum unchanged), so the sums equal
// btpm_selfheat_tiled's bit for bit.
template <int H>
__device__ __forceinline__ void selfheat_grp_body(const int2* __restrict__ gseg,
                                                  const int4* __restrict__ gent,
                                                  const int* __restrict__ gperm, int nf,
                                                  int nchunk, const int* __restrict__ slots,
                                                  const btpm_set_dev* __restrict__ sets,
                                                  const btpm_status_dev* __restrict__ status) {
  constexpr int kJ = H / 32;                      // 128-byte runs of a staged row
  constexpr int kQ4 = H / 4;                      // float4 per staged row
  constexpr int kPer = kSHChunk * kQ4 / kSHRows;  // float4 staged per lane
  constexpr int kPass = 8;                        // receivers per lane: 64 per wave / 8 groups
  static_assert(H % 32 == 0 && kSHChunk * kQ4 % kSHRows == 0 && kSHRows % kQ4 == 0,
                "staging shape");
  __shared__ __align__(16) float E[kSHChunk * H];
  const int slot = slots[blockIdx.z];
  if (status[slot].done) return;
  const btpm_set_dev& S = sets[slot];
  const int npad = S.nb * BTPM_BLK;
  const int s0 = blockIdx.y * H;
  if (s0 >= npad) return;
  const int tid = (int)threadIdx.x;
  const int rb = (int)blockIdx.x;
  const uint8_t* __restrict__ frozen = S.frozen;
  {
    const int f = gperm[rb * kSHRows + tid];
    const bool live = f < nf && (S.inst || !frozen[f]);
    if (__syncthreads_count(live) == 0) return;  // every receiver here is frozen
  }
  const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
  const int lane = tid & 63, g = lane >> 3, sl = lane & 7;
  int off[kJ];  // this lane's steps in the staged rows
#pragma unroll
  for (int j = 0; j < kJ; ++j) off[j] = 32 * j + 4 * sl;
  int fr[kPass];
  unsigned mine = 0, pass_live = 0;  // this lane's live passes; the wave's
#pragma unroll
  for (int p = 0; p < kPass; ++p) {
    fr[p] = gperm[rb * kSHRows + 64 * wave + 8 * p + g];
    const bool live = fr[p] < nf && (S.inst || !frozen[fr[p]]);
    mine |= (unsigned)live << p;
    if (__ballot(live) != 0) pass_live |= 1u << p;
  }
  const float* __restrict__ emis = S.emis;
  // staging: lane tid holds rows tid / kQ4 + i * (kSHRows / kQ4), float4 column tid % kQ4
  const int srow = tid / kQ4, sq4 = 4 * (tid % kQ4);
  const bool scol = s0 + sq4 < npad;
  float4 pre[kPer];
  auto fetch = [&](int c) {
    const int g0 = c * kSHChunk;
#pragma unroll
    for (int i = 0; i < kPer; ++i) {
      const int e = g0 + srow + i * (kSHRows / kQ4);
      pre[i] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
      if (e < nf && scol) pre[i] = gload4(emis + (size_t)e * npad + s0 + sq4);
    }
  };
  float acc[kPass][kJ][4];
#pragma unroll
  for (int p = 0; p < kPass; ++p)
#pragma unroll
    for (int j = 0; j < kJ; ++j)
#pragma unroll
      for (int q = 0; q < 4; ++q) acc[p][j][q] = 0.0f;
  // the wave's 8 segment descriptors of a chunk, lane l holding pass l % 8,
  // loaded one chunk ahead (read with readlane where they are used)
  const int2* __restrict__ sgb =
      gseg + (size_t)rb * nchunk * (BTPM_SH_WAVES * kPass) + wave * kPass + (lane & 7);
  int2 dn = sgb[0];
  fetch(0);
  for (int c = 0; c < nchunk; ++c) {
    __syncthreads();  // the previous chunk has been consumed
#pragma unroll
    for (int i = 0; i < kPer; ++i)
      *reinterpret_cast<float4*>(&E[(srow + i * (kSHRows / kQ4)) * H + sq4]) = pre[i];
    if (c + 1 < nchunk) fetch(c + 1);
    const int2 dc = dn;
    dn = sgb[(size_t)min(c + 1, nchunk - 1) * (BTPM_SH_WAVES * kPass)];
    __syncthreads();
#pragma unroll
    for (int p = 0; p < kPass; ++p) {
      const int n = (pass_live >> p) & 1 ? __builtin_amdgcn_readlane(dc.y, p) : 0;
      if (n == 0) continue;
      const int4* __restrict__ e = gent + (size_t)__builtin_amdgcn_readlane(dc.x, p) + g;
      // pairs are prefetched three ahead by unconditional loads (the index
      // clamped to the segment's last pair) into the registers of the pair
      // just consumed (the loop is unrolled by three, so no copy waits for a
      // load); the empty asm with a memory clobber keeps each load where it
      // is written (otherwise the compiler sinks it to its use and waits for
      // it there) without waiting for it
      const int last = n / 2 - 1;
      auto step = [&](const int4& q) {
        // the empty asm makes the pair's emitter indices opaque until here,
        // so their address arithmetic (and the wait for the pair's load)
        // stays at this step instead of being hoisted to the loop's top
        int xa = q.x, xb = q.z;
        asm volatile("" : "+v"(xa), "+v"(xb));
        float4 a[kJ], b[kJ];
#pragma unroll
        for (int j = 0; j < kJ; ++j) {
          a[j] = *reinterpret_cast<const float4*>(&E[xa * H + off[j]]);
          b[j] = *reinterpret_cast<const float4*>(&E[xb * H + off[j]]);
        }
        const float va = __int_as_float(q.y), vb = __int_as_float(q.w);
#pragma unroll
        for (int j = 0; j < kJ; ++j) {
          acc[p][j][0] = fmaf(vb, b[j].x, fmaf(va, a[j].x, acc[p][j][0]));
          acc[p][j][1] = fmaf(vb, b[j].y, fmaf(va, a[j].y, acc[p][j][1]));
          acc[p][j][2] = fmaf(vb, b[j].z, fmaf(va, a[j].z, acc[p][j][2]));
          acc[p][j][3] = fmaf(vb, b[j].w, fmaf(va, a[j].w, acc[p][j][3]));
        }
      };
      int4 qa = e[0], qb = e[min(1, last) * 8], qc = e[min(2, last) * 8];
      asm volatile("" ::: "memory");
      int kp = 0;
      for (; kp + 3 <= last + 1; kp += 3) {
        step(qa);
        qa = e[min(kp + 3, last) * 8];
        asm volatile("" ::: "memory");
        step(qb);
        qb = e[min(kp + 4, last) * 8];
        asm volatile("" ::: "memory");
        step(qc);
        qc = e[min(kp + 5, last) * 8];
        asm volatile("" ::: "memory");
      }
      if (kp <= last) step(qa);
      if (kp + 1 <= last) step(qb);
    }
  }
  const float* __restrict__ src = S.src;
  float* __restrict__ y = S.qext;
#pragma unroll
  for (int p = 0; p < kPass; ++p) {
    if (!((mine >> p) & 1)) continue;
    const int f = fr[p];
#pragma unroll
    for (int j = 0; j < kJ; ++j) {
      const int st = s0 + off[j];  // a multiple of 4: st < npad covers st + 3
      if (st >= npad) continue;
      const size_t o = ((size_t)(st / BTPM_BLK) * nf + f) * BTPM_BLK + (st % BTPM_BLK);
      const float4 a = gload4(src + o);
      gstore4(y + o,
            make_float4(a.x + acc[p][j][0], a.y + acc[p][j][1],
                                                      a.z + acc[p][j][2], a.w + acc[p][j][3]));
    }
  }
}

extern "C" __global__ __launch_bounds__(kSHRows) void btpm_selfheat_grp(
    const int2* gseg, const int4* gent, const int* gperm, int nf, int nchunk, const int* slots,
    const btpm_set_dev* sets, const btpm_status_dev* status) {
  selfheat_grp_body<32>(gseg, gent, gperm, nf, nchunk, slots, sets, status);
}

// The same product with the lanes along time: lane l of a wave holds step
// s0 + l, and the wave's 32 receivers are accumulators acc[r] in every lane.
// A receiver's entries are the same for all lanes of the wave, so they come
// through the scalar cache and the wave walks exactly its receivers' entries
// (no padding), and every entry reads one contiguous emission row segment
// E[g][s0 .. s0+63] from LDS (consecutive lanes on consecutive banks).  The
// workgroup (8 waves, 256 receivers x 64 steps) stages chunks of 256
// emitters x 64 steps (64 KB), two workgroups per CU.  Each receiver adds its
// entries in emitter order (chunks in order, the per-chunk CSR of
// btpm_selfheat_tiled within a chunk), so the sums equal that kernel's bit
// for bit.
namespace {
constexpr int kRowsR = 32;                   // receivers per wave
constexpr int kRowsWaves = 8;                // waves per workgroup
constexpr int kRowsF = kRowsR * kRowsWaves;  // receivers per workgroup
constexpr int kRowsH = 64;                   // steps per workgroup (one per lane)
static_assert(kRowsF == BTPM_SH_ROWS_T, "receiver block of the time-lane kernel");
}  // namespace

extern "C" __global__ __launch_bounds__(kRowsWaves * 64) void btpm_selfheat_rows(
    const int* __restrict__ cptr, const int2* __restrict__ ent, int nf, int nchunk,
    const int* __restrict__ slots, const btpm_set_dev* __restrict__ sets,
    const btpm_status_dev* __restrict__ status) {
  constexpr int kPer = kSHChunk * kRowsH / 4 / (kRowsWaves * 64);  // float4 staged per lane
  static_assert(kPer * kRowsWaves * 64 * 4 == kSHChunk * kRowsH, "staging shape");
  __shared__ __align__(16) float E[kSHChunk * kRowsH];
  const int slot = slots[blockIdx.z];
  if (status[slot].done) return;
  const btpm_set_dev& S = sets[slot];
  const int npad = S.nb * BTPM_BLK;
  const int s0 = blockIdx.y * kRowsH;
  if (s0 >= npad) return;
  const int tid = (int)threadIdx.x;
  const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
  const int lane = tid & 63;
  const int f0 = blockIdx.x * kRowsF + wave * kRowsR;  // the wave's first receiver
  // live receivers of the wave: bit r for receiver f0 + r (uniform)
  const int fl = f0 + lane;
  const bool lv = lane < kRowsR && fl < nf && (S.inst || !S.frozen[fl]);
  const unsigned long long live = __ballot(lv);
  if (__syncthreads_count(lane == 0 && live != 0ull) == 0) return;  // all frozen
  const float* __restrict__ emis = S.emis;
  float4 pre[kPer];
  auto fetch = [&](int c) {
    const int g0 = c * kSHChunk;
#pragma unroll
    for (int i = 0; i < kPer; ++i) {
      const int idx = tid + i * kRowsWaves * 64;      // float4 index in the chunk
      const int row = idx / (kRowsH / 4), q4 = 4 * (idx % (kRowsH / 4));
      pre[i] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
      if (g0 + row < nf && s0 + q4 < npad)
        pre[i] = gload4(emis + (size_t)(g0 + row) * npad + s0 + q4);
    }
  };
  float acc[kRowsR];
#pragma unroll
  for (int r = 0; r < kRowsR; ++r) acc[r] = 0.0f;
  fetch(0);
  for (int c = 0; c < nchunk; ++c) {
    const int g0 = c * kSHChunk;
    __syncthreads();  // the previous chunk has been consumed
#pragma unroll
    for (int i = 0; i < kPer; ++i)
      *reinterpret_cast<float4*>(&E[4 * (tid + i * kRowsWaves * 64)]) = pre[i];
    if (c + 1 < nchunk) fetch(c + 1);
    __syncthreads();
    // E[(g - g0) * kRowsH + lane]: emitter g of the chunk at this lane's step
    auto row = [&](int g) { return E[(g - g0) * kRowsH + lane]; };
#pragma unroll
    for (int r = 0; r < kRowsR; ++r) {
      if (!((live >> r) & 1ull)) continue;
      const int f = f0 + r;
      const int k0 = cptr[(size_t)c * nf + f], k1 = cptr[(size_t)(c + 1) * nf + f];
      float a = acc[r];
      int k = k0;
      for (; k + 4 <= k1; k += 4) {
        const int2 e0 = ent[k], e1 = ent[k + 1], e2 = ent[k + 2], e3 = ent[k + 3];
        const float x0 = row(e0.x), x1 = row(e1.x), x2 = row(e2.x), x3 = row(e3.x);
        a = fmaf(__int_as_float(e0.y), x0, a);
        a = fmaf(__int_as_float(e1.y), x1, a);
        a = fmaf(__int_as_float(e2.y), x2, a);
        a = fmaf(__int_as_float(e3.y), x3, a);
      }
      for (; k < k1; ++k) {
        const int2 e0 = ent[k];
        a = fmaf(__int_as_float(e0.y), row(e0.x), a);
      }
      acc[r] = a;
    }
  }
  const int s = s0 + lane;
  if (s >= npad) return;
  const float* __restrict__ src = S.src;
  float* __restrict__ y = S.qext;
#pragma unroll
  for (int r = 0; r < kRowsR; ++r) {
    if (!((live >> r) & 1ull)) continue;
    const size_t o = ((size_t)(s / BTPM_BLK) * nf + (f0 + r)) * BTPM_BLK + (s % BTPM_BLK);
    gstore1(y + o, gload1(src + o) + acc[r]);
  }
}

// elementwise helpers over blocked arrays of n floats
//   y = gain * scale_c * x
extern "C" __global__ void btpm_scale_cols(float* __restrict__ y, const float* __restrict__ x,
                                           const float* __restrict__ scale, float gain, int ncols,
                                           size_t n) {
  // grid-stride: a crater series can hold more than 2^32 elements
  const size_t stride = (size_t)gridDim.x * blockDim.x;
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
    const int c = (int)((i / BTPM_BLK) % ncols);
    y[i] = (gain * scale[c]) * x[i];
  }
}

// p_new = alpha*(d + q) has been written to y; record max |y - p| per
// workgroup into part[] and copy y into p.
extern "C" __global__ void btpm_delta_max(float* __restrict__ p, const float* __restrict__ y,
                                          float* __restrict__ part, size_t n) {
  __shared__ float red[256];
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  float d = 0.0f;
  if (i < n) {
    d = fabsf(y[i] - p[i]);
    p[i] = y[i];
  }
  red[threadIdx.x] = d;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) red[threadIdx.x] = fmaxf(red[threadIdx.x], red[threadIdx.x + s]);
    __syncthreads();
  }
  if (threadIdx.x == 0) part[blockIdx.x] = red[0];
}

// F = p / A  (A > 0)  else direct d
extern "C" __global__ void btpm_finish_scatter(float* __restrict__ F, const float* __restrict__ p,
                                               const float* __restrict__ d,
                                               const float* __restrict__ albedo, int ncols,
                                               size_t n) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int c = (int)((i / BTPM_BLK) % ncols);
  const float a = albedo[c];
  F[i] = a > 0.0f ? p[i] / a : d[i];
}

// ---------------------------------------------------------------------------
// per-set status after a rotation (one workgroup per set)
// ---------------------------------------------------------------------------
extern "C" __global__ __launch_bounds__(256) void btpm_status_kernel(
    const int* __restrict__ set_ids, const btpm_set_dev* __restrict__ sets,
    btpm_status_dev* __restrict__ status) {
  __shared__ long long red_w[256];
  __shared__ int red_a[256];
  __shared__ int scan[256];
  const int sid = set_ids[blockIdx.x];
  const btpm_set_dev& S = sets[sid];
  btpm_status_dev* st = status + sid;
  if (st->done) return;
  // the tiles this rotation could fill: every launch list covers them and
  // each of their workgroups wrote its counts (later tiles may be stale)
  const int ntl = min(S.ntiles, (st->norder + S.tile_cols - 1) / S.tile_cols);
  long long w = 0;
  int a = 0;
  for (int t = threadIdx.x; t < ntl; t += 256) {
    a += S.tile_active[t];
    w += S.tile_work[t];
  }
  red_w[threadIdx.x] = w;
  red_a[threadIdx.x] = a;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) {
      red_w[threadIdx.x] += red_w[threadIdx.x + s];
      red_a[threadIdx.x] += red_a[threadIdx.x + s];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    const int rev = st->rev;          // index of the rotation just executed
    st->active = red_a[0];
    st->colsteps += red_w[0] * (long long)S.ntau;
    st->rev = rev + 1;
    if (rev > S.check_after && red_a[0] == 0) {
      st->done = 1;
      st->converged = 1;
    } else if (rev + 1 >= S.max_rev) {
      st->done = 1;
    }
  }
  // stable compaction of the groups still integrating into order[] for the
  // next rotation (chunks of 256 with a block-wide inclusive scan)
  const int tid = (int)threadIdx.x;
  int base = 0;
  for (int g0 = 0; g0 < S.ngroups; g0 += 256) {
    const int g = g0 + tid;
    const int live = (g < S.ngroups && !S.frozen[g]) ? 1 : 0;
    __syncthreads();  // the previous chunk's scan[255] has been read
    scan[tid] = live;
    __syncthreads();
    for (int off = 1; off < 256; off <<= 1) {
      const int v = tid >= off ? scan[tid - off] : 0;
      __syncthreads();
      scan[tid] += v;
      __syncthreads();
    }
    if (live) S.order[base + scan[tid] - 1] = g;
    base += scan[255];
  }
  if (tid == 0) st->norder = base;
}
