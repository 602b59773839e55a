// probe.hip -- calibration kernel for the benchmark (built into its own
// libbtpm_probe.so, not part of libbtpm): the VALU issue ceiling of the
// device for plain (unpacked) FP32 instructions, the unit in which the
// compute-bound kernels of the library are measured.  Every lane runs 8 independent
// dependency chains of one instruction kind, so issue, not latency, bounds
// the loop; the kernel does no memory traffic besides one store per lane.
#pragma clang fp contract(off)

#include <hip/hip_runtime.h>


namespace {
constexpr int kChains = 8;
}

// op 0: v_fma_f32, op 1: v_add_f32, op 2: v_mul_f32, op 3: v_sqrt_f32
// (the hardware square root), op 4: v_add_u32 (32-bit integer add);
// kChains instructions of that kind per iteration and lane
template <int OP>
__global__ __launch_bounds__(256) void btpm_probe_valu_kernel(float* __restrict__ out, int iters,
                                                              float a, float b) {
  float x[kChains];
#pragma unroll
  for (int k = 0; k < kChains; ++k) x[k] = (float)(threadIdx.x + k) * 1e-3f;
  for (int i = 0; i < iters; ++i) {
#pragma unroll
    for (int k = 0; k < kChains; ++k) {
      if (OP == 0) x[k] = __builtin_fmaf(x[k], a, b);
      if (OP == 1) x[k] = x[k] + a;
      if (OP == 2) x[k] = x[k] * a;
      if (OP == 3) x[k] = __builtin_amdgcn_sqrtf(x[k]);
      if (OP == 4) x[k] = __uint_as_float(__float_as_uint(x[k]) + __float_as_uint(a));
    }
  }
  float s = 0.0f;
#pragma unroll
  for (int k = 0; k < kChains; ++k) s += x[k];
  out[blockIdx.x * blockDim.x + threadIdx.x] = s;
}

// Wave instructions per second of one instruction kind: op 0 fma, 1 add,
// 2 mul, 3 hardware sqrt, 4 integer add; iters iterations, 8 waves per SIMD.
// Returns 0, or -1 on a bad argument, -2 on a HIP failure.
extern "C" int btpm_probe_valu(int device, int op, int iters, double* wave_instr_per_s) {
  if (!wave_instr_per_s || op < 0 || op > 4 || iters < 1 || iters > (1 << 20)) return -1;
  if (hipSetDevice(device) != hipSuccess) return -2;
  int cus = 0;
  if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, device) != hipSuccess)
    return -2;
  const int blocks = cus * 32;   // 8 waves per SIMD
  float* d_out = nullptr;
  hipStream_t st = nullptr;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  int rc = -2;
  float ms = 0.0f;
  if (hipMalloc(&d_out, (size_t)blocks * 256 * sizeof(float)) != hipSuccess) return -2;
  if (hipStreamCreateWithFlags(&st, hipStreamNonBlocking) != hipSuccess) goto out;
  if (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess) goto out;
  {
    auto launch = [&](int n) {
      switch (op) {
        case 0: hipLaunchKernelGGL(btpm_probe_valu_kernel<0>, dim3(blocks), dim3(256), 0, st,
                                   d_out, n, 1.0001f, 1e-7f); break;
        case 1: hipLaunchKernelGGL(btpm_probe_valu_kernel<1>, dim3(blocks), dim3(256), 0, st,
                                   d_out, n, 1e-7f, 0.0f); break;
        case 2: hipLaunchKernelGGL(btpm_probe_valu_kernel<2>, dim3(blocks), dim3(256), 0, st,
                                   d_out, n, 0.99999f, 0.0f); break;
        case 3: hipLaunchKernelGGL(btpm_probe_valu_kernel<3>, dim3(blocks), dim3(256), 0, st,
                                   d_out, n, 0.0f, 0.0f); break;
        default: hipLaunchKernelGGL(btpm_probe_valu_kernel<4>, dim3(blocks), dim3(256), 0, st,
                                    d_out, n, 1.0f, 0.0f); break;
      }
    };
    launch(16);  // warm-up
    if (hipEventRecord(e0, st) != hipSuccess) goto out;
    launch(iters);
    if (hipEventRecord(e1, st) != hipSuccess || hipEventSynchronize(e1) != hipSuccess) goto out;
    if (hipGetLastError() != hipSuccess) goto out;
    if (hipEventElapsedTime(&ms, e0, e1) != hipSuccess) goto out;
    const double waves = (double)blocks * 4.0;
    *wave_instr_per_s = waves * (double)iters * kChains / (ms * 1e-3);
    rc = 0;
  }
out:
  if (e0) (void)hipEventDestroy(e0);
  if (e1) (void)hipEventDestroy(e1);
  if (st) (void)hipStreamDestroy(st);
  (void)hipFree(d_out);
  return rc;
}
