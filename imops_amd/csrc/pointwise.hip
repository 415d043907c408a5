// Per-pixel kernels: raw normalisation, the 3x3 "Markesteijn" demosaic and a
// fused chain of the elementwise colour stages.
//
// All of these are HBM-bound streaming kernels: each thread handles whole
// pixels with unit-stride loads/stores, and consecutive elementwise stages of
// the stage list (CFACoeffs, Exp, CST, Gamma) are applied in registers in one
// pass so an RGB image is read and written once per run of such stages.
// Each stage keeps the reference's f32 operation order, so fusing them does
// not change a single rounding.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "pointwise.hpp"

namespace imops {
namespace {

// u16 crop window -> (DN - black) / (white - black)   (demosaic.rs:61-91 +
// the black/white scaling done at raw ingest, extern_pipeline.rs:263).
__global__ void normalize_kernel(const uint16_t* __restrict__ raw, size_t full_w, size_t cx, size_t cy, size_t cw,
                                 size_t ch, float black, float range, float* __restrict__ out) {
  const size_t n = cw * ch;
  for (size_t k = (size_t)blockIdx.x * blockDim.x + threadIdx.x; k < n; k += (size_t)gridDim.x * blockDim.x) {
    const size_t y = k / cw, x = k - y * cw;
    out[k] = ((float)raw[(cy + y) * full_w + cx + x] - black) / range;
  }
}

// demosaic.rs:108-186: mean of the same-colour pixels in the clipped 3x3
// window (row-major order, centre skipped); native channel copied.
__global__ void markesteijn_kernel(const float* __restrict__ in, int width, int height, CfaPattern cfa,
                                   float* __restrict__ out) {
  const size_t n = (size_t)width * height;
  for (size_t k = (size_t)blockIdx.x * blockDim.x + threadIdx.x; k < n; k += (size_t)gridDim.x * blockDim.x) {
    const int row = (int)(k / width), col = (int)(k - (size_t)row * width);
    const int y0 = row > 0 ? row - 1 : 0, y1 = min(row + 1, height - 1);
    const int x0 = col > 0 ? col - 1 : 0, x1 = min(col + 1, width - 1);
    const int centre = cfa.at(row, col);
    float o[3];
    for (int ch = 0; ch < 3; ++ch) {
      if (ch == centre) { o[ch] = in[k]; continue; }
      float sum = 0.0f, count = 0.0f;
      for (int ny = y0; ny <= y1; ++ny)
        for (int nx = x0; nx <= x1; ++nx) {
          if (ny == row && nx == col) continue;
          if (cfa.at(ny, nx) == ch) { sum += in[(size_t)ny * width + nx]; count += 1.0f; }
        }
      o[ch] = count > 0.0f ? sum / count : 0.0f;
    }
    out[k * 3] = o[0];
    out[k * 3 + 1] = o[1];
    out[k * 3 + 2] = o[2];
  }
}

__device__ __forceinline__ float oetf(float x) {
  const float ax = fabsf(x);
  if (ax > 0.0031308f) {
    const float y = 1.055f * powf(ax, 1.0f / 2.4f) - 0.055f;
    return x < 0.0f ? -y : y;
  }
  return 12.92f * x;
}

__device__ __forceinline__ float eotf(float x) {
  const float ax = fabsf(x);
  if (ax > 0.04045f) {
    const float y = powf((ax + 0.055f) / 1.055f, 2.4f);
    return x < 0.0f ? -y : y;
  }
  return x / 12.92f;
}

__global__ void pointwise_kernel(float* __restrict__ rgb, size_t npix, PointwiseProgram prog) {
  for (size_t p = (size_t)blockIdx.x * blockDim.x + threadIdx.x; p < npix; p += (size_t)gridDim.x * blockDim.x) {
    float v0 = rgb[p * 3], v1 = rgb[p * 3 + 1], v2 = rgb[p * 3 + 2];
    for (int k = 0; k < prog.n; ++k) {
      const PointwiseOp& op = prog.ops[k];
      switch (op.kind) {
        case PW_MAT3: {
          const float* m = op.m;
          const float a = m[0] * v0 + m[1] * v1 + m[2] * v2;
          const float b = m[3] * v0 + m[4] * v1 + m[5] * v2;
          const float c = m[6] * v0 + m[7] * v1 + m[8] * v2;
          v0 = a; v1 = b; v2 = c;
          break;
        }
        case PW_SCALE3:
          v0 = v0 * op.m[0]; v1 = v1 * op.m[1]; v2 = v2 * op.m[2];
          break;
        case PW_OETF:
          v0 = oetf(v0); v1 = oetf(v1); v2 = oetf(v2);
          break;
        case PW_EOTF:
          v0 = eotf(v0); v1 = eotf(v1); v2 = eotf(v2);
          break;
        case PW_GAMMA: {
          const float inv = op.m[0];
          v0 = v0 <= 0.0f ? 0.0f : powf(v0, inv);
          v1 = v1 <= 0.0f ? 0.0f : powf(v1, inv);
          v2 = v2 <= 0.0f ? 0.0f : powf(v2, inv);
          break;
        }
        default: break;
      }
    }
    rgb[p * 3] = v0;
    rgb[p * 3 + 1] = v1;
    rgb[p * 3 + 2] = v2;
  }
}

int grid_for(size_t n, int block) {
  size_t g = (n + block - 1) / block;
  if (g > 65536) g = 65536;
  return (int)(g == 0 ? 1 : g);
}

}  // namespace

hipError_t normalize_launch(const uint16_t* raw, size_t full_w, size_t cx, size_t cy, size_t cw, size_t ch,
                            float black, float white, float* out, hipStream_t s) {
  const float range = white - black;
  hipLaunchKernelGGL(normalize_kernel, dim3(grid_for(cw * ch, 256)), dim3(256), 0, s, raw, full_w, cx, cy, cw, ch,
                     black, range, out);
  return hipGetLastError();
}

hipError_t markesteijn_launch(const float* in, int width, int height, const CfaPattern& cfa, float* out,
                              hipStream_t s) {
  hipLaunchKernelGGL(markesteijn_kernel, dim3(grid_for((size_t)width * height, 256)), dim3(256), 0, s, in, width,
                     height, cfa, out);
  return hipGetLastError();
}

hipError_t pointwise_launch(float* rgb, size_t npix, const PointwiseProgram& prog, hipStream_t s) {
  if (prog.n == 0) return hipSuccess;
  hipLaunchKernelGGL(pointwise_kernel, dim3(grid_for(npix, 256)), dim3(256), 0, s, rgb, npix, prog);
  return hipGetLastError();
}

}  // namespace imops
