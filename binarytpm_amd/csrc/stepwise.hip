// stepwise.hip -- the reference's GPU execution structure, as a baseline.
//
// One finite-difference launch and one boundary-condition launch per time
// step with a host synchronisation after each, the depth profile kept in
// global memory between the two, the crater self-heating as a per-lane loop
// over the 100 emitting sub-facets, and a separate convergence kernel whose
// flags the host reads back after every rotation (the control flow of
// GPU_inertia_difference / GPU_converged, T175706/thermal_GPU.cu:430-574,
// 648-692, 1021-1026; the smooth facets follow inertiadifference,
// T175706/thermal.cpp:1895-1979, in the same shape).  The arithmetic is the
// reference's: absolute FP32 temperatures with
//   T_i += r (T_{i+1} - 2 T_i + T_{i-1}),   T_NX = T_{NX-1},
// and the FP64 Newton of the surface energy balance started from the old
// T_0.  It shares the set layouts (blocked series, profile records) with the
// persistent kernels, so its results are read through the same accessors;
// it exists to measure what the persistent, batched path gains.
#include <hip/hip_runtime.h>

#include "btpm_device.h"
#include "btpm_internal.h"
#include "btpm_kernels.h"

namespace {
__device__ __forceinline__ size_t blocked(int s, int c, int ncols) {
  return ((size_t)(s / BTPM_BLK) * ncols + c) * BTPM_BLK + (s % BTPM_BLK);
}
__device__ __forceinline__ int group_of(const btpm_set_dev& S, int c) {
  return S.kind == BTPM_SET_CRATER ? c / BTPM_NS : c;
}
}  // namespace

// profile records [c][BTPM_PSTRIDE] <-> node-major temp [NX+1][ncols]
extern "C" __global__ void btpm_step_profile(float* __restrict__ temp, float* __restrict__ prof,
                                             int ncols, int to_records) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= ncols) return;
  for (int i = 0; i <= BTPM_NX; ++i) {
    if (to_records)
      prof[(size_t)c * BTPM_PSTRIDE + i] = temp[(size_t)i * ncols + c];
    else
      temp[(size_t)i * ncols + c] = prof[(size_t)c * BTPM_PSTRIDE + i];
  }
}

// finite_difference: interior nodes from the old values, then the bottom.
extern "C" __global__ void btpm_step_fd(btpm_set_dev S, float* __restrict__ temp) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= S.ncols || S.inst) return;
  const int g = group_of(S, c);
  if (S.frozen[g]) return;
  const int n = S.ncols;
  const float r = S.r[g];
  float t[BTPM_NX + 1];
  for (int i = 0; i <= BTPM_NX; ++i) t[i] = temp[(size_t)i * n + c];
  float before = t[0];
  for (int i = 1; i < BTPM_NX; ++i) {
    const float cur = t[i];
    t[i] = cur + r * (t[i + 1] - 2.0f * cur + before);
    before = cur;
  }
  t[BTPM_NX] = t[BTPM_NX - 1];
  for (int i = 1; i <= BTPM_NX; ++i) temp[(size_t)i * n + c] = t[i];
}

// boundary_condition for smooth facets: Q = sum_g VF_fg es_ia T_g,prev(s)^4
// gathered from the previous rotation's series, then Newton (or the TI = 0
// equilibrium).
extern "C" __global__ void btpm_step_bc_smooth(btpm_set_dev S, float* __restrict__ temp,
                                               const float* __restrict__ Tp, int s,
                                               const int* __restrict__ rowptr,
                                               const int* __restrict__ colidx,
                                               const float* __restrict__ vf,
                                               float* __restrict__ dm) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= S.ncols || S.frozen[c]) return;
  const int n = S.ncols;
  const size_t o = blocked(s, c, n);
  float src = S.src[o];
  if (rowptr) {
    float q = 0.0f;
    for (int k = rowptr[c]; k < rowptr[c + 1]; ++k) {
      const float t = Tp[blocked(s, colidx[k], n)];
      q += vf[k] * (S.esia * (t * t) * (t * t));
    }
    src += q;
  }
  double x;
  if (S.inst)
    x = sqrt(sqrt((double)src / S.es));
  else
    x = btpm_dev::sbc_newton((double)src, (double)S.m1[c], (double)temp[n + c], (double)S.es,
                             fmax((double)temp[c], 1.0), (double)S.newton_tol, S.newton_cap);
  const float t0 = (float)x;
  temp[c] = t0;
  S.T[o] = t0;
  dm[c] = fmaxf(dm[c], fabsf(t0 - Tp[o]));
}

// boundary_condition for crater columns: one workgroup per crater,
// E_i = es_ia T_i,prev(s)^4 in LDS, Q_j = src_j + sum_i VF(j->i) E_i as a
// per-lane loop (vf32[i*100+j] = VF(j->i)), then Newton per sub-facet.
extern "C" __global__ __launch_bounds__(128) void btpm_step_bc_crater(
    btpm_set_dev S, float* __restrict__ temp, const float* __restrict__ Tp, int s,
    const float* __restrict__ vf32, float* __restrict__ dm) {
  __shared__ float E[BTPM_NS];
  const int f = blockIdx.x;
  const int j = threadIdx.x;
  if (f >= S.ngroups || S.frozen[f]) return;  // uniform over the workgroup
  const int n = S.ncols;
  const int c = f * BTPM_NS + j;
  const bool ok = j < BTPM_NS;
  size_t o = 0;
  float tp = 0.0f;
  if (ok) {
    o = blocked(s, c, n);
    tp = Tp[o];
    E[j] = S.esia * (tp * tp) * (tp * tp);
  }
  __syncthreads();
  if (!ok) return;
  float q = S.src[o];
  for (int i = 0; i < BTPM_NS; ++i) q += vf32[i * BTPM_NS + j] * E[i];
  double x;
  if (S.inst)
    x = sqrt(sqrt((double)q / S.es));
  else
    x = btpm_dev::sbc_newton((double)q, (double)S.m1[f], (double)temp[n + c], (double)S.es,
                             fmax((double)temp[c], 1.0), (double)S.newton_tol, S.newton_cap);
  const float t0 = (float)x;
  temp[c] = t0;
  S.T[o] = t0;
  dm[c] = fmaxf(dm[c], fabsf(t0 - tp));
}

// converged: per group max |T - T_prev| over the rotation; freeze below TACC
// once `check` is set (rev > check_after).
extern "C" __global__ void btpm_step_converged(btpm_set_dev S, const float* __restrict__ dm,
                                               int check) {
  const int g = blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= S.ngroups || S.frozen[g]) return;
  const int w = S.kind == BTPM_SET_CRATER ? BTPM_NS : 1;
  float m = 0.0f;
  for (int j = 0; j < w; ++j) m = fmaxf(m, dm[(size_t)g * w + j]);
  S.dmax[g] = m;
  if (check && m < S.tacc) S.frozen[g] = 1;
}
