// Measured rates of the device the library runs on: the HBM copy bandwidth,
// the FP32 VALU issue rate of the non-fused mul/add stream the exact-order
// kernels execute, and the FP16 MFMA rate of the instruction the mixed
// kernels use.  The benchmark reports them next to the published peaks.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "peaks.hpp"

namespace imops {
namespace {

typedef _Float16 half8 __attribute__((ext_vector_type(8)));
typedef float floatx4 __attribute__((ext_vector_type(4)));

__global__ void copy_rate_kernel(const float4* __restrict__ src, float4* __restrict__ dst, size_t n) {
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) dst[i] = src[i];
}

// kValuPerIter VALU instructions per iteration: 8 independent chains, each a
// v_mul_f32 then a v_add_f32 (no contraction, no packed FP32), unrolled 4x.
constexpr int kValuPerIter = 4 * 16;

__global__ void __launch_bounds__(256) valu_rate_kernel(float* out, int iters, float seed) {
  float a[8];
#pragma unroll
  for (int k = 0; k < 8; ++k) a[k] = seed + (float)threadIdx.x * 1e-7f + (float)k;
  const float m = 0.999999f, c = 1e-7f;
  for (int i = 0; i < iters; ++i) {
#pragma unroll
    for (int u = 0; u < 4; ++u) {
#pragma unroll
      for (int k = 0; k < 8; ++k) a[k] = a[k] * m;
#pragma unroll
      for (int k = 0; k < 8; ++k) a[k] = a[k] + c;
    }
  }
  float s = 0.0f;
#pragma unroll
  for (int k = 0; k < 8; ++k) s += a[k];
  if (s == 1234.5f) out[blockIdx.x] = s;  // never true; keeps the chains live
}

// 4 independent accumulator chains of v_mfma_f32_16x16x32_f16 per wave,
// 16 products per loop iteration.
__global__ void __launch_bounds__(256) mfma_rate_kernel(float* out, int iters) {
  const int lane = threadIdx.x & 63;
  half8 a, b;
#pragma unroll
  for (int e = 0; e < 8; ++e) {
    a[e] = (_Float16)(0.001f * (float)(lane + e));
    b[e] = (_Float16)(0.002f * (float)(lane - e));
  }
  floatx4 c0 = {0.0f, 0.0f, 0.0f, 0.0f}, c1 = c0, c2 = c0, c3 = c0;
  for (int i = 0; i < iters; i += 4) {
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      c0 = __builtin_amdgcn_mfma_f32_16x16x32_f16(a, b, c0, 0, 0, 0);
      c1 = __builtin_amdgcn_mfma_f32_16x16x32_f16(a, b, c1, 0, 0, 0);
      c2 = __builtin_amdgcn_mfma_f32_16x16x32_f16(a, b, c2, 0, 0, 0);
      c3 = __builtin_amdgcn_mfma_f32_16x16x32_f16(a, b, c3, 0, 0, 0);
    }
  }
  const floatx4 t = c0 + c1 + c2 + c3;
  if (t[0] + t[1] + t[2] + t[3] == 1234.5f) out[blockIdx.x] = t[0];
}

}  // namespace

hipError_t measure_peaks(hipStream_t s, void* buf_a, void* buf_b, size_t bytes, int device, MeasuredPeaks* out) {
  hipDeviceProp_t prop;
  hipError_t e = hipGetDeviceProperties(&prop, device);
  if (e != hipSuccess) return e;
  const int cus = prop.multiProcessorCount;
  const double clock_hz = (double)prop.clockRate * 1e3;
  hipEvent_t e0, e1;
  if ((e = hipEventCreate(&e0)) != hipSuccess) return e;
  if ((e = hipEventCreate(&e1)) != hipSuccess) {
    (void)hipEventDestroy(e0);
    return e;
  }
  auto timed = [&](auto&& launch, int reps, float* ms) -> hipError_t {
    launch();  // warm-up
    hipError_t r = hipEventRecord(e0, s);
    for (int i = 0; r == hipSuccess && i < reps; ++i) launch();
    if (r == hipSuccess) r = hipGetLastError();
    if (r == hipSuccess) r = hipEventRecord(e1, s);
    if (r == hipSuccess) r = hipEventSynchronize(e1);
    if (r == hipSuccess) r = hipEventElapsedTime(ms, e0, e1);
    *ms /= reps;
    return r;
  };
  float ms = 0.0f;
  // HBM: copy `bytes` (reads and writes that many bytes)
  const size_t n4 = bytes / 16;
  e = timed([&] { hipLaunchKernelGGL(copy_rate_kernel, dim3(cus * 16), dim3(256), 0, s, (const float4*)buf_a, (float4*)buf_b, n4); },
            10, &ms);
  if (e == hipSuccess) out->copy_GBps = 2.0 * (double)(n4 * 16) / (ms * 1e-3) / 1e9;
  // VALU: 8 waves per SIMD
  const int vblocks = cus * 8, viters = 4096;
  if (e == hipSuccess)
    e = timed([&] { hipLaunchKernelGGL(valu_rate_kernel, dim3(vblocks), dim3(256), 0, s, (float*)buf_b, viters, 1.0f); }, 5,
              &ms);
  if (e == hipSuccess) {
    const double instr = (double)vblocks * 4 * viters * kValuPerIter;  // wave64 instructions
    out->valu_per_clk_per_simd = instr / (ms * 1e-3) / clock_hz / (4.0 * cus);
    out->valu_nonfused_TFLOPs = instr * 64 / (ms * 1e-3) / 1e12;  // one flop per lane per mul or add
  }
  // MFMA: 4 waves per SIMD
  const int mblocks = cus * 4, miters = 4096;
  if (e == hipSuccess)
    e = timed([&] { hipLaunchKernelGGL(mfma_rate_kernel, dim3(mblocks), dim3(256), 0, s, (float*)buf_b, miters); }, 5, &ms);
  if (e == hipSuccess) out->mfma_f16_TFLOPs = (double)mblocks * 4 * miters * 4 * (16.0 * 16 * 32 * 2) / (ms * 1e-3) / 1e12;
  out->clock_GHz = clock_hz / 1e9;
  out->compute_units = cus;
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  return e;
}

}  // namespace imops
