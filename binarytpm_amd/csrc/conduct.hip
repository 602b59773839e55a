// conduct.hip -- persistent per-rotation conduction kernels for gfx950.
//
// One launch advances every listed tile by one full rotation (ntau steps):
// the 57-node depth profile of each column lives in VGPRs for the whole
// rotation, inputs/outputs stream through HBM in 8-step blocks (blocked SoA,
// see btpm_internal.h) with the next block prefetched while the current one
// is integrated, and the per-column rotation-to-rotation change is reduced
// in-kernel for the freeze test.
//
// Algorithm per column and step (the reference's inertiadifference,
// T175706/thermal.cpp:1895-1979, and the crater kernels finite_difference +
// boundary_condition, T175706/thermal_GPU.cu:430-574):
//   FTCS   T_i += r (T_{i+1} - 2 T_i + T_{i-1}),  i = 1..NX-1 (old values)
//   bottom T_NX = T_{NX-1}
//   SBC    Newton on (1-A)F + Q + m1 (T_1 - T_0) - es T_0^4 = 0 from the old
//          T_0 until |dT| < tol, in FP64
// where Q (self-heating) always uses the PREVIOUS rotation's surface
// temperatures at the same step.
//
// Precision: the profile is held as offsets u_i = T_i - T_ref with a
// per-column reference T_ref (the bottom node at rotation start).  The FTCS
// operator is exact on constants, so the offsets integrate the same
// recurrence while their FP32 rounding is set by the size of the offsets
// (K) rather than of the absolute temperature (hundreds of K).
#include <hip/hip_runtime.h>

#include <type_traits>

#include "btpm_internal.h"
#include "btpm_kernels.h"
#include "btpm_device.h"

typedef _Float16 half4_t __attribute__((ext_vector_type(4)));
typedef _Float16 half8_t __attribute__((ext_vector_type(8)));
// Pointers read from the set descriptor are generic; casting them to the
// global address space lets the compiler emit global (not flat) memory ops,
// whose completion is tracked by vmcnt alone so prefetches stay in flight.
#define GLOBAL_AS __attribute__((address_space(1)))
typedef GLOBAL_AS float gfloat;
typedef float float4_t __attribute__((ext_vector_type(4)));

namespace {

// FTCS in scaled form.  With c = 1 - 2r the step is u_i' = c u_i + r (u_{i-1}
// + u_{i+1}); holding U = u / s with s = c^k after k steps turns it into
//     U_i' = U_i + kr (U_{i-1} + U_{i+1}),   kr = r / c,
// one add and one FMA accumulated in place per node instead of three
// operations.  The per-step factor is made consistent with the rounded kr as
// sig = 1 / (1 + 2 kr) (so constants stay exactly constant), tracked in FP64
// with its inverse, and folded back into U once per 8-step block (U grows by
// at most c^-8 < 1e32 in between for r <= BTPM_R_MAX); the crater kernel
// folds every 1 to 3 blocks, as rarely as the FP32 range allows for its r.
// The zero-flux bottom is the reference's copy node u[NX] = u[NX-1]: node
// NX-1 sees its own old value in its place, and u[NX] itself is not carried
// (a profile enters with u[NX] = u[NX-1], which every FD step leaves so, and
// leaves through store_profile with the copy).
template <int NX>
__device__ __forceinline__ void ftcs_scaled(float (&u)[NX + 1], float kr) {
  float before = u[0];
#pragma unroll
  for (int i = 1; i < NX - 1; ++i) {
    const float cur = u[i];
    u[i] = fmaf(kr, u[i + 1] + before, cur);
    before = cur;
  }
  const float cur = u[NX - 1];
  u[NX - 1] = fmaf(kr, cur + before, cur);
}

// Per-column scale state of the scaled recurrence: real offset = sd * U.
struct Scale {
  float kr;         // r / (1 - 2r), FP32
  double sig, isig; // per-step factor 1 / (1 + 2 kr) and its inverse (exact)
  double sd, isd;   // current scale and its inverse
};

__device__ __forceinline__ float scaled_kr(float r) { return r / (1.0f - 2.0f * r); }

__device__ __forceinline__ Scale make_scale(float r) {
  Scale sc;
  sc.kr = scaled_kr(r);
  sc.isig = fma(2.0, (double)sc.kr, 1.0);
  sc.sig = 1.0 / sc.isig;
  sc.sd = 1.0;
  sc.isd = 1.0;
  return sc;
}

// Scale of the n steps between two rescales for one r, shared by all columns
// with that r: sk[k] = (sig^(k+1), sig^-(k+1)) in FP32 for the per-step
// conversions and pk[k] = sig^(k+1) in FP64 for the rescale.
__device__ __forceinline__ void scale_table(float r, float2* sk, double* pk, int n) {
  const double isig = fma(2.0, (double)scaled_kr(r), 1.0);
  const double sig = 1.0 / isig;
  double p = 1.0, ip = 1.0;
  for (int k = 0; k < n; ++k) {
    p *= sig;
    ip *= isig;
    sk[k] = make_float2((float)p, (float)ip);
    pk[k] = p;
  }
}

// The crater kernel folds the scale back into the profile every 1 to
// kRescaleMax blocks: as rarely as U <= 1e4 K x sig^-(8 blocks) stays inside
// the FP32 range for every crater of the workgroup (sig = 1 - 2r).
constexpr int kRescaleMax = 3;
__device__ __forceinline__ int rescale_blocks(float r) {
  const double c = 1.0 - 2.0 * (double)r;
  const double c2 = c * c, c4 = c2 * c2, c8 = c4 * c4;
  const double lim = 1e4 / 3.4e38;  // 1e4 K / FLT_MAX
  return c8 * c8 * c8 > lim ? 3 : c8 * c8 > lim ? 2 : 1;
}

// Per-crater constants of the crater kernel, shared by the crater's 100 lanes.
struct CraterConst {
  float2 sk[kRescaleMax * BTPM_BLK];  // scale_table
  double pk[kRescaleMax * BTPM_BLK];
  double r0[2];         // residual of the scale carried from the last rescale
  double m1;            // Gamma sqrt(2 pi n / P) / dz
  double es, tol;       // emissivity x sigma and the Newton tolerance, read from
                        // LDS at each step (held in VGPRs they would cost 4)
};

// Fold the accumulated scale into the profile: U *= f with f = fp32(sd); the
// residual sd / f (within 6e-8 of 1) is carried, so the block-to-block gain
// has no systematic rounding bias.  sd = 1 / isd to second order.
template <int NX>
__device__ __forceinline__ void rescale(float (&u)[NX + 1], Scale& sc) {
  const float f = (float)sc.sd;
#pragma unroll
  for (int i = 0; i < NX; ++i) u[i] *= f;
  sc.isd *= (double)f;
  sc.sd = 2.0 - sc.isd;
}

__device__ __forceinline__ void load8(const gfloat* p, float (&v)[BTPM_BLK]) {
  const GLOBAL_AS float4_t* q = reinterpret_cast<const GLOBAL_AS float4_t*>(p);
  const float4_t a = q[0], b = q[1];
  v[0] = a.x; v[1] = a.y; v[2] = a.z; v[3] = a.w;
  v[4] = b.x; v[5] = b.y; v[6] = b.z; v[7] = b.w;
}

__device__ __forceinline__ void load4(const gfloat* p, float (&v)[BTPM_BLK], int k0) {
  const float4_t a = *reinterpret_cast<const GLOBAL_AS float4_t*>(p + k0);
  v[k0] = a.x; v[k0 + 1] = a.y; v[k0 + 2] = a.z; v[k0 + 3] = a.w;
}

// FTCS + bottom boundary + FP64 Newton SBC for one step; returns T_0.
// (s, is) is the scale after this step and its inverse (FP32 or FP64): the
// Newton solve sees the real T_1 = Tref + s U_1 and the root goes back into
// the scaled profile as U_0 = (T_0 - Tref) / s.
// The Newton start x0 is a predictor: the previous rotation's surface
// temperature at this step shifted by the difference between the two
// rotations, extrapolated linearly from the two previous steps (the caller's
// dl = t - t_prev at step k-1, dl2 at step k-2: x0 = t_prev + 2 dl - dl2).
template <int NX, typename SC>
__device__ __forceinline__ double conduct_step(float (&u)[NX + 1], float kr, SC s, SC is,
                                               double src, double m1, double es, double Tref,
                                               double tol, int cap, double x0) {
  ftcs_scaled<NX>(u, kr);
  double t1;
  if constexpr (std::is_same<SC, double>::value)
    t1 = fma(s, (double)u[1], Tref);
  else
    t1 = Tref + (double)(s * u[1]);
  // the first Newton iteration is straight-line code next to the FTCS of
  // nodes 2..NX, which do not depend on it: the empty asm statements keep
  // those nodes' updates ahead of the (rarely taken) iteration loop, so the
  // scheduler fills the FP64 chain's latency with them
  const double K = fma(m1, t1, src);
  double x = fmax(x0, 1.0);
  const double d0 = btpm_dev::newton_step(K, m1, es, x);
#pragma unroll
  for (int i = 2; i < NX; ++i) asm volatile("" : "+v"(u[i]));
  x = btpm_dev::newton_finish(K, m1, es, x, d0, tol, cap);
  if constexpr (std::is_same<SC, double>::value)
    u[0] = (float)((x - Tref) * is);
  else
    u[0] = (float)(x - Tref) * is;
  return x;
}

// profile of one column: 57 nodes in a 64-float record (one base address,
// immediate offsets), held as offsets from its bottom node
template <int NX>
__device__ __forceinline__ float load_profile(const gfloat* rec, float (&u)[NX + 1]) {
#pragma unroll
  for (int i = 0; i + 3 <= NX; i += 4) {
    const float4_t a = *reinterpret_cast<const GLOBAL_AS float4_t*>(rec + i);
    u[i] = a.x; u[i + 1] = a.y; u[i + 2] = a.z; u[i + 3] = a.w;
  }
#pragma unroll
  for (int i = (NX + 1) & ~3; i <= NX; ++i) u[i] = rec[i];
  const float tref = u[NX];
#pragma unroll
  for (int i = 0; i <= NX; ++i) u[i] -= tref;
  return tref;
}

// (node NX is written as the copy of node NX-1)
template <int NX>
__device__ __forceinline__ void store_profile(gfloat* rec, const float (&u)[NX + 1], float tref) {
  auto node = [&](int i) { return tref + u[i < NX ? i : NX - 1]; };
#pragma unroll
  for (int i = 0; i + 3 <= NX; i += 4)
    *reinterpret_cast<GLOBAL_AS float4_t*>(rec + i) =
        float4_t{node(i), node(i + 1), node(i + 2), node(i + 3)};
#pragma unroll
  for (int i = (NX + 1) & ~3; i <= NX; ++i) rec[i] = node(i);
}

__device__ __forceinline__ void store8(gfloat* p, const float (&v)[BTPM_BLK]) {
  GLOBAL_AS float4_t* q = reinterpret_cast<GLOBAL_AS float4_t*>(p);
  q[0] = float4_t{v[0], v[1], v[2], v[3]};
  q[1] = float4_t{v[4], v[5], v[6], v[7]};
}

}  // namespace

// ---------------------------------------------------------------------------
// Smooth facets: one lane per column, W columns per workgroup (W = 256, 128
// or 64, instantiated below; the set's tile_cols is W).  The width changes
// only which lanes share a workgroup, never a column's arithmetic.
// ---------------------------------------------------------------------------
template <int W>
__global__ __launch_bounds__(W, 4) void btpm_conduct_smooth(const btpm_tile* __restrict__ tiles,
                                                           const btpm_set_dev* __restrict__ sets,
                                                           const btpm_status_dev* __restrict__ status) {
  // previous-rotation T of the block, parked in LDS (lane-private slots)
  __shared__ float lds_p[BTPM_BLK][W];
  constexpr int NX = BTPM_NX;
  const btpm_tile tl = tiles[blockIdx.x];
  const btpm_set_dev& S = sets[tl.set];
  const btpm_status_dev st = status[tl.set];
  if (st.done || st.rev >= S.max_rev) return;
  const bool check = st.rev > S.check_after;

  const int ncols = S.ncols;
  // lane -> column through the compacted order of unfrozen columns; lanes
  // past its end take a valid column for their (unused) loads
  const int idx = tl.tile * W + (int)threadIdx.x;
  const int norder = st.norder;
  const bool valid = idx < norder;
  const int c = S.order[valid ? idx : (norder > 0 ? norder - 1 : 0)];
  const int cc = c;
  const bool active = valid && (S.inst || !S.frozen[cc]);
  const int nwork = __syncthreads_count(active);
  if (nwork == 0) {
    if (threadIdx.x == 0) {
      S.tile_active[tl.tile] = 0;
      S.tile_work[tl.tile] = 0;
    }
    return;
  }

  Scale sc = make_scale(S.r[cc]);
  const double m1 = (double)S.m1[cc];
  const double es = (double)S.es;
  const double tol = (double)S.newton_tol;
  const int cap = S.newton_cap;
  // with facet self-heating the SpMM has written src + Q into qext
  const gfloat* __restrict__ srcp = (const gfloat*)(S.qext != nullptr ? S.qext : S.src);
  const gfloat* __restrict__ Tin = (const gfloat*)S.T;
  gfloat* __restrict__ Tout = (gfloat*)S.T;
  gfloat* __restrict__ Eout = (gfloat*)S.emis;
  const bool has_e = S.emis != nullptr;
  const float esia = S.esia;

  float u[NX + 1];
  const float Tref = load_profile<NX>((const gfloat*)S.prof + (size_t)cc * BTPM_PSTRIDE, u);
  const double Trefd = (double)Tref;

  const int nb = S.nb;
  const int ntau = S.ntau;
  const size_t bstride = (size_t)ncols * BTPM_BLK;
  const size_t coff = (size_t)cc * BTPM_BLK;

  const int tid = (int)threadIdx.x;
  float nsrc[BTPM_BLK], nprev[BTPM_BLK];
  load8(srcp + coff, nsrc);
  load8(Tin + coff, nprev);

  float dm = 0.0f;
  float dl = 0.0f;   // T - T_prev at the previous step (predictor shift)
  float dl2 = 0.0f;  // and at the step before
  // the TI = 0 loop is a separate instantiation so the main loop stays lean
  auto rotation = [&](auto inst_tag) {
    constexpr bool INST = decltype(inst_tag)::value;
    for (int b = 0; b < nb; ++b) {
#pragma unroll
      for (int k = 0; k < BTPM_BLK; ++k) lds_p[k][tid] = nprev[k];
      // compiler barrier: keep the parked values in LDS, not in VGPRs
      asm volatile("" ::: "memory");
      const bool more = b + 1 < nb;
      const size_t o = (size_t)b * bstride + coff;
      const size_t on = (size_t)(b + 1) * bstride + coff;
      if (more) load8(Tin + on, nprev);
      const int kend = min(BTPM_BLK, ntau - b * BTPM_BLK);
#pragma unroll
      for (int k = 0; k < BTPM_BLK; ++k) {
        const float tp = lds_p[k][tid];
        float t = tp;
        if (k < kend) {
          if constexpr (INST) {  // instantaneous equilibrium (TI = 0), thermal.cpp:1985-2033
            t = (float)sqrt(sqrt((double)nsrc[k] / es));
          } else {
            sc.sd *= sc.sig;
            sc.isd *= sc.isig;
            t = (float)conduct_step<NX>(u, sc.kr, sc.sd, sc.isd, (double)nsrc[k], m1, es, Trefd,
                                        tol, cap, (double)(tp + (2.0f * dl - dl2)));
          }
          dl2 = dl;
          dl = t - tp;
          dm = fmaxf(dm, fabsf(dl));
        }
        lds_p[k][tid] = t;  // the slot now holds this rotation's T
        if (k == BTPM_BLK / 2 - 1 && more) load4(srcp + on, nsrc, 0);
        if (k == BTPM_BLK - 1 && more) load4(srcp + on, nsrc, 4);
      }
      asm volatile("" ::: "memory");
      if constexpr (!INST) rescale<NX>(u, sc);
      if (active) {
        float out[BTPM_BLK];
#pragma unroll
        for (int k = 0; k < BTPM_BLK; ++k) out[k] = lds_p[k][tid];
        store8(Tout + o, out);
        if (has_e) {
#pragma unroll
          for (int k = 0; k < BTPM_BLK; ++k) {
            const float t = out[k];
            out[k] = esia * (t * t) * (t * t);
          }
          store8(Eout + (size_t)cc * nb * BTPM_BLK + (size_t)b * BTPM_BLK, out);  // facet-major
        }
      }
    }
  };
  if (S.inst) rotation(std::true_type{});
  else rotation(std::false_type{});

  bool still = active;
  if (active) {
    if (!S.inst) {
      int cw = c;
      asm volatile("" : "+v"(cw));  // keep the record address out of the step loop
      store_profile<NX>((gfloat*)S.prof + (size_t)cw * BTPM_PSTRIDE, u, Tref);
    }
    S.dmax[c] = dm;
    if (check && dm < S.tacc) {
      still = false;
      if (!S.inst) S.frozen[c] = 1;
    }
  }
  const int nact = __syncthreads_count(still);
  if (threadIdx.x == 0) {
    S.tile_active[tl.tile] = nact;
    S.tile_work[tl.tile] = S.inst ? 0 : nwork;
  }
}

// every width with the same register budget (4 waves per SIMD, 128 VGPRs)
template __global__ void btpm_conduct_smooth<256>(const btpm_tile*, const btpm_set_dev*,
                                                  const btpm_status_dev*);
template __global__ void btpm_conduct_smooth<128>(const btpm_tile*, const btpm_set_dev*,
                                                  const btpm_status_dev*);
template __global__ void btpm_conduct_smooth<64>(const btpm_tile*, const btpm_set_dev*,
                                                 const btpm_status_dev*);

// ---------------------------------------------------------------------------
// Crater columns: BTPM_CRATERS_WG craters x 100 sub-facets per workgroup of
// 1024 lanes (lane -> sub-facet column, 24 idle lanes).  Every 8-step block
// the workgroup evaluates the crater self-heating of all its craters for the
// 8 steps at once as one MFMA product
//     Q[j][n] = sum_i VF(j->i) E[i][n],   n = (crater, step in block)
// with M = 100 (padded 112), K = 100 (padded 112), N = 80, using
// v_mfma_f32_16x16x16_f16 and the three-product split
//     VF E ~ VFhi Ehi + VFhi Elo + VFlo Ehi   (FP32 accumulation),
// where x_hi = fp16(x) and x_lo = fp16(x - x_hi).  VF is stored pre-scaled by
// 2^s (max entry in [0.5, 1)) so its low halves stay normal fp16 numbers;
// the FP32 accumulator is multiplied by 2^-s, which is exact.
// ---------------------------------------------------------------------------
namespace {
constexpr int kMP = 112;           // padded M (receiving sub-facet j)
constexpr int kKP = 112;           // padded K (emitting sub-facet i)
// LDS row stride in halves: 60 dwords, so the 8 rows of one 16-byte fragment
// read (8 lanes) start 4 banks apart and cover all 32 banks
constexpr int kLdsK = 120;
constexpr int kNP = BTPM_CRATERS_WG * BTPM_BLK;  // 80 columns of the product
// Q row stride: 1000 columns + 24 idle lanes + 4 floats of padding, so the
// 8 step rows start in different LDS banks and the product's float4 stores
// for steps 0..7 of one column (8 lanes) do not collide on one bank
constexpr int kQCols = 1028;
constexpr int kMT = kMP / 16;      // 7 M tiles
constexpr int kNT = kNP / 16;      // 5 N tiles

// Q[n][m] = sum_k A[m][k] B[n][k] over the padded 112 x 112 x 80 tile held in
// LDS (A = VF^T hi/lo, B = E hi/lo).  Q is written column-major for the
// conduction lanes: product column n = crater * 8 + step goes to
// lds_q[step][crater * 100 + m], m < 100.  Output layout of the 16 x 16 MFMAs:
// lane l receives D[m = 4(l/16) + i][n = l%16], i = 0..3.
//
// The phase is bound by the latency of its LDS fragment reads, so a wave
// computes a strip of one m tile and 3 (or 2) n tiles, reading each A fragment
// once per K chunk for the whole strip; K runs as 3 chunks of 32
// (v_mfma_f32_16x16x32_f16, 16-byte fragment reads) plus the tail k = 96..99
// as one 16x16x16 chunk in which only the 16 lanes holding sub-facets read.
// The 14 strips (7 m tiles x {n 0-2, n 3-4}) sit on the waves so that each
// SIMD (waves s, s+4, s+8, s+12) carries at most 9 of the 35 output tiles;
// waves 12 and 13 have none.  Within a lane, fragment slot j of A and of B
// hold the same k, which is all the product needs of the K layout.
//
// One K chunk of 32 (v_mfma_f32_16x16x32_f16): lane l supplies
// A[m = l%16][k = 8(l/16) .. +7] and B[k = 8(l/16) .. +7][n = l%16].
template <int NNT>
__device__ __forceinline__ void strip_chunk32(const _Float16 (*lds_a)[kMP * kLdsK],
                                              const _Float16 (*lds_b)[kNP * kLdsK], int am, int k0,
                                              int nt0, int r, float4_t (&acc)[NNT]) {
  const half8_t ahi = *reinterpret_cast<const half8_t*>(&lds_a[0][am * kLdsK + k0]);
  const half8_t alo = *reinterpret_cast<const half8_t*>(&lds_a[1][am * kLdsK + k0]);
#pragma unroll
  for (int t = 0; t < NNT; ++t) {
    const int bn = (nt0 + t) * 16 + r;
    const half8_t bhi = *reinterpret_cast<const half8_t*>(&lds_b[0][bn * kLdsK + k0]);
    const half8_t blo = *reinterpret_cast<const half8_t*>(&lds_b[1][bn * kLdsK + k0]);
    acc[t] = __builtin_amdgcn_mfma_f32_16x16x32_f16(alo, bhi, acc[t], 0, 0, 0);
    acc[t] = __builtin_amdgcn_mfma_f32_16x16x32_f16(ahi, blo, acc[t], 0, 0, 0);
    acc[t] = __builtin_amdgcn_mfma_f32_16x16x32_f16(ahi, bhi, acc[t], 0, 0, 0);
  }
}

template <int NNT>
__device__ __forceinline__ void product_strip(const _Float16 (*lds_a)[kMP * kLdsK],
                                              const _Float16 (*lds_b)[kNP * kLdsK],
                                              float* lds_q, int mt, int nt0, int lane,
                                              float unscale) {
  constexpr int kFull = BTPM_NS / 32 * 32;  // 96: sub-facets in full chunks of 32
  static_assert(BTPM_NS - kFull == 4 && kKP >= kFull + 16, "one partial K chunk of 4");
  float4_t acc[NNT];
#pragma unroll
  for (int t = 0; t < NNT; ++t) acc[t] = float4_t{0.0f, 0.0f, 0.0f, 0.0f};
  const int r = lane & 15;
  const int g = lane >> 4;
  const int am = mt * 16 + r;
#pragma unroll 1
  for (int kt = 0; kt < kFull / 32; ++kt)
    strip_chunk32<NNT>(lds_a, lds_b, am, kt * 32 + 8 * g, nt0, r, acc);
  // the tail k = 96..99 as one v_mfma_f32_16x16x16f16 chunk (k = 4(l/16)..+3):
  // only lanes 0-15 hold sub-facets, the others contribute zero operands
  // instead of reading the zero padding
  {
    const half4_t zero = {0, 0, 0, 0};
    half4_t ahi = zero, alo = zero, bhi[NNT], blo[NNT];
#pragma unroll
    for (int t = 0; t < NNT; ++t) bhi[t] = blo[t] = zero;
    if (g == 0) {
      ahi = *reinterpret_cast<const half4_t*>(&lds_a[0][am * kLdsK + kFull]);
      alo = *reinterpret_cast<const half4_t*>(&lds_a[1][am * kLdsK + kFull]);
#pragma unroll
      for (int t = 0; t < NNT; ++t) {
        const int bn = (nt0 + t) * 16 + r;
        bhi[t] = *reinterpret_cast<const half4_t*>(&lds_b[0][bn * kLdsK + kFull]);
        blo[t] = *reinterpret_cast<const half4_t*>(&lds_b[1][bn * kLdsK + kFull]);
      }
    }
#pragma unroll
    for (int t = 0; t < NNT; ++t) {
      acc[t] = __builtin_amdgcn_mfma_f32_16x16x16f16(alo, bhi[t], acc[t], 0, 0, 0);
      acc[t] = __builtin_amdgcn_mfma_f32_16x16x16f16(ahi, blo[t], acc[t], 0, 0, 0);
      acc[t] = __builtin_amdgcn_mfma_f32_16x16x16f16(ahi, bhi[t], acc[t], 0, 0, 0);
    }
  }
  const int m0 = mt * 16 + 4 * g;  // 4 receiving sub-facets m0 .. m0 + 3
  if (m0 < BTPM_NS) {
#pragma unroll
    for (int t = 0; t < NNT; ++t) {
      const int bn = (nt0 + t) * 16 + r;
      float* q = &lds_q[(bn & (BTPM_BLK - 1)) * kQCols + (bn / BTPM_BLK) * BTPM_NS + m0];
      *reinterpret_cast<float4*>(q) = make_float4(acc[t][0] * unscale, acc[t][1] * unscale,
                                                  acc[t][2] * unscale, acc[t][3] * unscale);
    }
  }
}

// wave -> strip, one nibble per wave: 0..6 = m tile with n tiles 0-2,
// 7..13 = m tile (value - 7) with n tiles 3-4, 15 = none
constexpr unsigned long long kStripOfWave = 0xD9FFC852B741A630ull;

__device__ __forceinline__ void crater_product(const _Float16 (*lds_a)[kMP * kLdsK],
                                               const _Float16 (*lds_b)[kNP * kLdsK],
                                               float* lds_q, int wave, int lane, float unscale) {
  static_assert(kMT == 7 && kNT == 5 && BTPM_CRATER_THREADS == 1024, "strip table shape");
  const int s = (int)((kStripOfWave >> (4 * wave)) & 15ull);
  if (s < kMT)
    product_strip<3>(lds_a, lds_b, lds_q, s, 0, lane, unscale);
  else if (s < 2 * kMT)
    product_strip<2>(lds_a, lds_b, lds_q, s - kMT, 3, lane, unscale);
}

// Stage VF^T hi/lo ([2][112][112] fp16 in global memory) into LDS.
__device__ __forceinline__ void stage_vf(const _Float16* src, _Float16 (*lds_a)[kMP * kLdsK],
                                         int tid) {
  for (int e = tid; e < 2 * kMP * kKP; e += BTPM_CRATER_THREADS) {
    const int h = e / (kMP * kKP);
    const int rem = e - h * kMP * kKP;
    const int m = rem / kKP, k = rem - m * kKP;
    lds_a[h][m * kLdsK + k] = src[e];
  }
}
}  // namespace

extern "C" __global__ __launch_bounds__(BTPM_CRATER_THREADS) void btpm_conduct_crater(
    const btpm_tile* __restrict__ tiles, const btpm_set_dev* __restrict__ sets,
    const btpm_status_dev* __restrict__ status) {
  constexpr int NX = BTPM_NX;
  constexpr int NS = BTPM_NS;
  constexpr int P = BTPM_CRATERS_WG;
  __shared__ _Float16 lds_a[2][kMP * kLdsK];   // VF^T hi / lo  (52 KB)
  __shared__ _Float16 lds_b[2][kNP * kLdsK];   // E hi / lo     (37 KB)
  __shared__ float lds_q[BTPM_BLK * kQCols];   // product [step][column] (32 KB)
  __shared__ float lds_p[BTPM_BLK][BTPM_CRATER_THREADS];  // T_prev, then T (32 KB)
  __shared__ CraterConst lds_cc[P];            // per-crater constants
  __shared__ unsigned lds_cmax[P];
  __shared__ int lds_cnt;
  __shared__ int lds_nres[P];                  // blocks per rescale per crater

  const btpm_tile tl = tiles[blockIdx.x];
  const btpm_set_dev& S = sets[tl.set];
  const btpm_status_dev st = status[tl.set];
  if (st.done || st.rev >= S.max_rev) return;
  const bool check = st.rev > S.check_after;

  const int tid = (int)threadIdx.x;
  const int lc = tid < P * NS ? tid : P * NS - 1;  // local column (clamped)
  const int lcr = lc / NS;                         // local crater
  const int j = lc - lcr * NS;                     // sub-facet
  // local crater -> crater through the compacted order of unfrozen craters;
  // slots past its end and the idle lanes (tid >= 1000) mirror a valid
  // crater's column, so they stay well-behaved
  const int norder = st.norder;
  const int slot = tl.tile * P + lcr;
  const bool valid = tid < P * NS && slot < norder;
  const int last = norder > 0 ? norder - 1 : 0;
  const int g = S.order[slot < norder ? slot : last];
  const int gg = g;
  const int c = gg * NS + j;                       // global column
  const bool active = valid && !S.frozen[gg];
  const int nwork = __syncthreads_count(active);
  if (nwork == 0) {
    if (tid == 0) {
      S.tile_active[tl.tile] = 0;
      S.tile_work[tl.tile] = 0;
    }
    return;
  }

  // stage VF^T (hi, lo) once per launch; zero the padded E rows
  {
    stage_vf(reinterpret_cast<const _Float16*>(S.vfA), lds_a, tid);
    for (int e = tid; e < 2 * kNP * kLdsK; e += BTPM_CRATER_THREADS)
      (&lds_b[0][0])[e] = (_Float16)0.0f;
    if (tid < P) {
      lds_cmax[tid] = 0u;
      const int gt = S.order[tl.tile * P + tid < norder ? tl.tile * P + tid : last];
      scale_table(S.r[gt], lds_cc[tid].sk, lds_cc[tid].pk, kRescaleMax * BTPM_BLK);
      lds_nres[tid] = rescale_blocks(S.r[gt]);
      lds_cc[tid].r0[0] = 1.0;
      lds_cc[tid].m1 = (double)S.m1[gt];
      lds_cc[tid].es = (double)S.es;
      lds_cc[tid].tol = (double)S.newton_tol;
    }
  }
  __syncthreads();
  // the idle lanes' Q slots are never written by the product
  if (tid >= P * NS)
    for (int k = 0; k < BTPM_BLK; ++k) lds_q[k * kQCols + tid] = 0.0f;

  const float kr = scaled_kr(S.r[gg]);
  CraterConst& cc = lds_cc[lcr];
  const int cap = S.newton_cap;
  int nres = kRescaleMax;  // blocks per rescale: the workgroup's smallest
  for (int i = 0; i < P; ++i) nres = min(nres, lds_nres[i]);
  nres = __builtin_amdgcn_readfirstlane(nres);
  int bph = 0;   // blocks since the last rescale
  int rpar = 0;  // rescales so far, mod 2
  const float esia = S.esia;
  const int ncols = S.ncols;
  // held in a register, so the block loop has no scalar load (read through S
  // there, it is reloaded after every barrier, and a pending scalar load
  // makes the product's first LDS wait a full drain of the shared counter)
  const float vf_unscale = S.vf_unscale;

  float u[NX + 1];
  const float Tref = load_profile<NX>((const gfloat*)S.prof + (size_t)c * BTPM_PSTRIDE, u);
  const double Trefd = (double)Tref;

  const int nb = S.nb;
  const int ntau = S.ntau;
  const size_t bstride = (size_t)ncols * BTPM_BLK;
  const unsigned coff = (unsigned)c * BTPM_BLK;   // < 2^32 / 4: lane part of an address
  const int wave = tid >> 6;
  const int lane = tid & 63;

  const gfloat* __restrict__ srcp = (const gfloat*)S.src;
  const gfloat* __restrict__ Tin = (const gfloat*)S.T;
  gfloat* __restrict__ Tout = (gfloat*)S.T;
  float nsrc[BTPM_BLK], nprev[BTPM_BLK];
  load8(srcp + coff, nsrc);
  load8(Tin + coff, nprev);

  float dm = 0.0f;
  float dl = 0.0f;   // T - T_prev at the previous step (predictor shift)
  float dl2 = 0.0f;  // and at the step before
  // the lane's part of a block address, 32 bits, added to the uniform block
  // base at each use (the asm statement keeps the compiler from folding it
  // into 64-bit per-lane pointers held across the loop)
  auto lane_off = [&]() {
    unsigned o = coff;
    asm volatile("" : "+v"(o));
    return o;
  };

  // previous-rotation emission of block b -> LDS (fp16 hi/lo split), T_prev
  // parked in lane-private LDS slots, the next block's T_prev prefetched,
  // then the block's self-heating product into lds_q
  auto block_head = [&](int b, bool more) {
    if (tid < P * NS) {
#pragma unroll
      for (int k = 0; k < BTPM_BLK; ++k) {
        const float t = nprev[k];
        const float e = esia * (t * t) * (t * t);
        const _Float16 hi = (_Float16)e;
        const _Float16 lo = (_Float16)(e - (float)hi);
        const int n = lcr * BTPM_BLK + k;
        lds_b[0][n * kLdsK + j] = hi;
        lds_b[1][n * kLdsK + j] = lo;
      }
    }
#pragma unroll
    for (int k = 0; k < BTPM_BLK; ++k) lds_p[k][tid] = nprev[k];
    if (more) load8(Tin + (size_t)(b + 1) * bstride + lane_off(), nprev);
    __syncthreads();
    crater_product(lds_a, lds_b, lds_q, wave, lane, vf_unscale);
    __syncthreads();
  };
  auto block_tail = [&](int b) {
    if (active) {
      float out[BTPM_BLK];
#pragma unroll
      for (int k = 0; k < BTPM_BLK; ++k) out[k] = lds_p[k][tid];
      store8(Tout + (size_t)b * bstride + lane_off(), out);
    }
  };
  // U *= f = fp32(r0 sig^n) at a rescale, n the steps since the last one; the
  // residual r0 sig^n / f is carried (cc.r0 is double buffered by rescale
  // parity: every lane reads one half, sub-facet 0 writes the other, and the
  // barriers of the blocks in between order them)
  auto block_scale = [&](int n) {
    const double pb = cc.r0[rpar] * cc.pk[n - 1];
    const float f = (float)pb;
    if (j == 0) cc.r0[rpar ^ 1] = fma(pb - (double)f, (double)__builtin_amdgcn_rcpf(f), 1.0);
    rpar ^= 1;
    return f;
  };
  auto after_step = [&](int k, float tp, float t) {
    dl2 = dl;
    dl = t - tp;
    dm = fmaxf(dm, fabsf(dl));
    lds_p[k][tid] = t;  // the slot now holds this rotation's T
  };
  auto next_src = [&](int b, int k) {
    const gfloat* next = srcp + (size_t)(b + 1) * bstride;  // uniform block base
    if (k == BTPM_BLK / 2 - 1) load4(next + lane_off(), nsrc, 0);
    if (k == BTPM_BLK - 1) load4(next + lane_off(), nsrc, 4);
  };

  // the steps of block b, all 57 nodes on the vector ALUs.  FULL: the block
  // has all 8 steps and a successor whose inputs are prefetched (every block
  // but the last), so the unrolled steps carry no step-count tests
  auto valu_steps = [&](int b, auto inst_tag, auto full_tag) {
    constexpr bool INST = decltype(inst_tag)::value;
    constexpr bool FULL = decltype(full_tag)::value;
    const int kend = FULL ? BTPM_BLK : min(BTPM_BLK, ntau - b * BTPM_BLK);
    // this block's row of the scale table (steps since the last rescale)
    const float2* sk8 = &cc.sk[bph * BTPM_BLK];
#pragma unroll
    for (int k = 0; k < BTPM_BLK; ++k) {
      const float tp = lds_p[k][tid];
      if (FULL || k < kend) {
        const double tot = (double)(nsrc[k] + lds_q[k * kQCols + tid]);
        float t;
        if constexpr (INST) {  // roughness_inst_equilibrium, thermal_GPU.cu:581-641
          t = (float)sqrt(sqrt(tot / cc.es));
        } else {
          // the per-step conversions leave out the residual (within 6e-8 of
          // 1): a bounded perturbation of T_1 and T_0, not a gain on the profile
          const float2 sk = sk8[k];
          t = (float)conduct_step<NX>(u, kr, sk.x, sk.y, tot, cc.m1, cc.es, Trefd, cc.tol, cap,
                                      (double)(tp + (2.0f * dl - dl2)));
        }
        after_step(k, tp, t);
      }
      if constexpr (FULL) next_src(b, k);
    }
    asm volatile("" ::: "memory");
    if constexpr (!INST) {
      // every nres blocks, and at the end of the rotation (the profile goes
      // out unscaled)
      if (!FULL || bph + 1 == nres) {
        const float f = block_scale(bph * BTPM_BLK + kend);
#pragma unroll
        for (int i = 0; i < NX; ++i) u[i] *= f;
        bph = 0;
      } else {
        ++bph;
      }
    }
  };

  // the TI = 0 loop is a separate instantiation so the main loop stays lean
  auto rotation = [&](auto inst_tag) {
    for (int b = 0; b + 1 < nb; ++b) {
      block_head(b, true);
      valu_steps(b, inst_tag, std::true_type{});
      block_tail(b);
    }
    block_head(nb - 1, false);
    valu_steps(nb - 1, inst_tag, std::false_type{});
    block_tail(nb - 1);
  };
  if (S.inst) rotation(std::true_type{});
  else rotation(std::false_type{});

  // per-crater freeze test (converged kernel, T175706/thermal_GPU.cu:648-692)
  if (active) {
    if (!S.inst) {
      int cw = c;
      asm volatile("" : "+v"(cw));  // keep the record address out of the step loop
      store_profile<NX>((gfloat*)S.prof + (size_t)cw * BTPM_PSTRIDE, u, Tref);
    }
    atomicMax(&lds_cmax[lcr], __float_as_uint(dm));
  }
  if (tid == 0) lds_cnt = 0;
  __syncthreads();
  if (active && j == 0) {
    const float cm = __uint_as_float(lds_cmax[lcr]);
    S.dmax[g] = cm;
    bool still = true;
    if (check && cm < S.tacc) {
      S.frozen[g] = 1;
      still = false;
    }
    if (still) atomicAdd(&lds_cnt, 1);
  }
  __syncthreads();
  if (tid == 0) {
    S.tile_active[tl.tile] = lds_cnt;
    S.tile_work[tl.tile] = S.inst ? 0 : nwork;
  }
}

// Standalone crater self-heating product (tests, tools and the
// allocation-free entry point):
//   q[n][j] = sum_i VF(j->i) e[n][i],  n < ncol, with the kernel's own split.
// The power-of-two unscale comes by value, or from device memory when
// unscale_p is set (tiles packed on the device).
extern "C" __global__ __launch_bounds__(BTPM_CRATER_THREADS) void btpm_contract_kernel(
    const _Float16* __restrict__ vfA, const float* __restrict__ e, float* __restrict__ q,
    int ncol, float unscale_v, const float* __restrict__ unscale_p) {
  __shared__ _Float16 lds_a[2][kMP * kLdsK];
  __shared__ _Float16 lds_b[2][kNP * kLdsK];
  __shared__ float lds_q[BTPM_BLK * kQCols];
  const float unscale = unscale_p != nullptr ? *unscale_p : unscale_v;
  const int tid = (int)threadIdx.x;
  const int n0 = blockIdx.x * kNP;
  stage_vf(vfA, lds_a, tid);
  for (int x = tid; x < kNP * kLdsK; x += BTPM_CRATER_THREADS) {
    const int n = x / kLdsK, i = x - n * kLdsK;
    float v = 0.0f;
    if (i < BTPM_NS && n0 + n < ncol) v = e[(size_t)(n0 + n) * BTPM_NS + i];
    const _Float16 hi = (_Float16)v;
    lds_b[0][x] = hi;
    lds_b[1][x] = (_Float16)(v - (float)hi);
  }
  __syncthreads();
  crater_product(lds_a, lds_b, lds_q, tid >> 6, tid & 63, unscale);
  __syncthreads();
  for (int x = tid; x < kNP * BTPM_NS; x += BTPM_CRATER_THREADS) {
    const int n = x / BTPM_NS, jj = x - n * BTPM_NS;
    if (n0 + n < ncol)
      q[(size_t)(n0 + n) * BTPM_NS + jj] =
          lds_q[(n % BTPM_BLK) * kQCols + (n / BTPM_BLK) * BTPM_NS + jj];
  }
}
