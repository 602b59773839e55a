// noalloc.hip -- the allocation-free variant of the crater self-heating
// product (btpm_crater_contract_noalloc): inputs and outputs in the caller's
// device memory, the packed fp16 hi/lo view-factor tiles in a file-scope
// device array instead of a buffer the library allocates, so the call makes
// no hipMalloc / hipFree (usable under HIP graph capture and in loops that
// must not allocate).
//
// The array is one per device, held until the process ends, and shared by
// every call on that device: calls that overlap on one device race (the
// second call's packing may overwrite the tiles the first call's product is
// reading).  Calls on one stream are ordered and safe.  The packing kernel
// writes every byte the product reads (the zero padding included), so the
// array needs no clearing.
#include <hip/hip_runtime.h>

#include <cmath>
#include <mutex>

#include "../../include/btpm.h"
#include "btpm_internal.h"
#include "btpm_kernels.h"
#include "runtime.h"

using namespace btpm;

namespace {
constexpr int kTile = 112;                                 // padded M and K of the product
constexpr size_t kTileBytes = 2 * kTile * kTile * sizeof(_Float16);   // VF^T hi, lo
constexpr size_t kScratchBytes = (kTileBytes + sizeof(float) + 255) & ~(size_t)255;  // 50,432
constexpr int kMaxDevices = 64;

// scratch = [VF^T hi tile][VF^T lo tile][unscale float][pad]
__device__ __align__(256) unsigned char scratch[kScratchBytes];

std::mutex g_mu;
void* g_scratch[kMaxDevices] = {};

// One workgroup: the largest |VF| (block reduction), then the tiles of
// vf_tiles() (host.hip) on the device: A[h][m = j][k = i] = hi/lo of
// VF(j->i) * 2^s, zero padded, with 2^s putting the largest entry in
// [0.5, 1) (clamped to 2^-14 .. 2^24), and the unscale 2^-s.
__global__ __launch_bounds__(1024) void noalloc_pack_vf(const float* __restrict__ vf,
                                                      unsigned char* __restrict__ scr) {
  __shared__ float red[1024];
  const int tid = (int)threadIdx.x;
  float m = 0.0f;
  for (int e = tid; e < BTPM_NS * BTPM_NS; e += 1024) m = fmaxf(m, fabsf(vf[e]));
  red[tid] = m;
  __syncthreads();
  for (int w = 512; w > 0; w >>= 1) {
    if (tid < w) red[tid] = fmaxf(red[tid], red[tid + w]);
    __syncthreads();
  }
  const float vmax = red[0];
  int s = 0;
  if (vmax > 0.0f) {
    int ex = 0;
    (void)frexpf(vmax, &ex);  // vmax = f 2^ex, f in [0.5, 1)
    s = max(-14, min(24, -ex));
  }
  const float sc = ldexpf(1.0f, s);
  _Float16* a = reinterpret_cast<_Float16*>(scr);
  for (int e = tid; e < kTile * kTile; e += 1024) {
    const int jj = e / kTile, ii = e - jj * kTile;
    _Float16 hi = (_Float16)0.0f, lo = (_Float16)0.0f;
    if (jj < BTPM_NS && ii < BTPM_NS) {
      const float v = vf[ii * BTPM_NS + jj] * sc;
      hi = (_Float16)v;
      lo = (_Float16)(v - (float)hi);
    }
    a[e] = hi;
    a[kTile * kTile + e] = lo;
  }
  if (tid == 0) *reinterpret_cast<float*>(scr + kTileBytes) = ldexpf(1.0f, -s);
}
}  // namespace

namespace btpm {
// Called by btpm_ctx_create with `device` current: the device address of this
// device's copy of `scratch` (the array's name is a host-side stand-in).
void noalloc_setup(int device) {
  if (device < 0 || device >= kMaxDevices) return;
  std::lock_guard<std::mutex> lk(g_mu);
  if (g_scratch[device]) return;
  void* ptr = nullptr;
  HIPCHK(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
  g_scratch[device] = ptr;
}
}  // namespace btpm

extern "C" int btpm_crater_contract_noalloc(btpm_ctx* ctx, const float* vf, const float* e,
                                            int ncol, float* q, void* stream) {
  if (!ctx || !vf || !e || !q) return fail(BTPM_ERR_ARG, "null argument");
  if (ncol < 1 || ncol > BTPM_NOALLOC_MAX_COLS)
    return fail(BTPM_ERR_ARG, "ncol = %d outside [1, %d] (the allocation-free product's limit)",
                ncol, BTPM_NOALLOC_MAX_COLS);
  if (ctx->device < 0 || ctx->device >= kMaxDevices)
    return fail(BTPM_ERR_ARG, "device %d beyond the scratch table", ctx->device);
  void* scr = nullptr;
  {
    std::lock_guard<std::mutex> lk(g_mu);
    scr = g_scratch[ctx->device];
  }
  if (!scr) return fail(BTPM_ERR_STATE, "scratch of device %d was not set up", ctx->device);
  auto run = [&](hipStream_t st) {
    HIPCHK(hipSetDevice(ctx->device));
    unsigned char* s8 = static_cast<unsigned char*>(scr);
    hipLaunchKernelGGL(noalloc_pack_vf, dim3(1), dim3(1024), 0, st, vf, s8);
    const int per = BTPM_CRATERS_WG * BTPM_BLK;  // product columns per workgroup
    hipLaunchKernelGGL(btpm_contract_kernel, dim3((ncol + per - 1) / per),
                       dim3(BTPM_CRATER_THREADS), 0, st, reinterpret_cast<const _Float16*>(s8), e,
                       q, ncol, 1.0f, reinterpret_cast<const float*>(s8 + kTileBytes));
    HIPCHK(hipGetLastError());
  };
  if (stream) {
    // the caller's stream (it may be capturing a graph): enqueue only
    API_BEGIN
    run(static_cast<hipStream_t>(stream));
    API_END
  }
  CtxLock lk(ctx);
  API_BEGIN
  run(ctx->stream);
  HIPCHK(hipStreamSynchronize(ctx->stream));
  API_END
}
