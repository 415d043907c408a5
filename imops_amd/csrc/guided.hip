// Luma-guided chroma denoise (pichromatic/src/chroma_denoise.rs:14-152) for
// gfx950: the guided filter of He et al. with the luma as guide, smoothing the
// chroma differences C_R = R - Y and C_B = B - Y.
//
// The box filter (chroma_denoise.rs:100-152) is a running sum along each row,
// then along each column, normalised by the in-bounds window size.  Its f32
// value depends on the order of the adds and subtracts, so each line keeps the
// reference's single running sum; lines run in parallel, one lane per line.
// A wave owns 64 lines and walks them in chunks of 64 positions: the chunk of
// the leading stream (the samples entering the window) and of the trailing
// stream (the samples leaving it) are loaded with coalesced 256-byte accesses
// into registers one chunk ahead, parked in LDS, and read back one position per
// step.  Row passes write their outputs through LDS as well, so every global
// access is a 256-byte line.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <algorithm>

#include "runtime.hpp"
#include "stages.hpp"

namespace imops {
namespace {

// Relative luminance weights of pixel.rs:33-35.
constexpr float kWr = 0.2722287168f, kWg = 0.6740817658f, kWb = 0.0536895174f;

__device__ __forceinline__ float luma(float r, float g, float b) { return kWr * r + kWg * g + kWb * b; }

// Planes 0..5: Y, C_R, C_B, Y*Y, Y*C_R, Y*C_B (chroma_denoise.rs:21-27, 71-72).
__global__ void guided_prep_kernel(const float* __restrict__ rgb, size_t n, float* __restrict__ planes) {
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
    const float r = rgb[i * 3], g = rgb[i * 3 + 1], b = rgb[i * 3 + 2];
    const float y = luma(r, g, b);
    const float cr = r - y, cb = b - y;
    planes[i] = y;
    planes[n + i] = cr;
    planes[2 * n + i] = cb;
    planes[3 * n + i] = y * y;
    planes[4 * n + i] = y * cr;
    planes[5 * n + i] = y * cb;
  }
}

// Box means of the prep planes -> a, b per chroma channel (chroma_denoise.rs:77-97):
// out planes a_R, b_R, a_B, b_B.
__global__ void guided_coef_kernel(const float* __restrict__ means, size_t n, float eps, float* __restrict__ ab) {
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
    const float mg = means[i], mgg = means[3 * n + i];
    const float var = mgg - mg * mg;
#pragma unroll
    for (int c = 0; c < 2; ++c) {
      const float mp = means[(1 + c) * n + i], mgp = means[(4 + c) * n + i];
      const float cov = mgp - mg * mp;
      const float a = cov / (var + eps);
      ab[(2 * c) * n + i] = a;
      ab[(2 * c + 1) * n + i] = mp - a * mg;
    }
  }
}

// q = mean_a Y + mean_b per chroma channel, then the RGB recombination (chroma_denoise.rs:29-37, 99-103).
__global__ void guided_apply_kernel(float* __restrict__ rgb, const float* __restrict__ mab, size_t n) {
  const float kr = kWr / kWg, kb = kWb / kWg;
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
    const float r = rgb[i * 3], g = rgb[i * 3 + 1], b = rgb[i * 3 + 2];
    const float y = luma(r, g, b);
    const float qr = mab[i] * y + mab[n + i];
    const float qb = mab[2 * n + i] * y + mab[3 * n + i];
    rgb[i * 3] = y + qr;
    rgb[i * 3 + 1] = y - kr * qr - kb * qb;
    rgb[i * 3 + 2] = y + qb;
  }
}

constexpr int kLines = 64;  // lines per wave
constexpr int kChunk = 64;  // positions per chunk

// One pass of box_filter over `planes` planes of w x h (plane stride w*h).
// kRows: lines are image rows (the horizontal pass), else columns (the vertical
// pass, which also divides by the in-bounds window area).  r is the radius
// clamped to the line length - 1 (the same sums and window sizes).
template <bool kRows>
__global__ void __launch_bounds__(kLines) box_pass_kernel(const float* __restrict__ in, float* __restrict__ out, int w,
                                                          int h, int r, int rx) {
  __shared__ float lead_s[kChunk][kLines + 1], trail_s[kChunk][kLines + 1], out_s[kChunk][kLines + 1];
  const size_t plane = (size_t)w * h;
  in += blockIdx.y * plane;
  out += blockIdx.y * plane;
  const int nl = kRows ? h : w, len = kRows ? w : h;
  const int lane = threadIdx.x, base = blockIdx.x * kLines, line = base + lane;
  // element (line l, position p)
  auto at = [&](int l, int p) -> size_t { return kRows ? (size_t)l * w + p : (size_t)p * w + l; };
  // one chunk block of a stream: positions p0 .. p0 + 63 of the wave's 64 lines, lane-coalesced
  auto load_block = [&](int p0, float (&v)[kChunk]) {
#pragma unroll
    for (int j = 0; j < kChunk; ++j) {
      // rows: load j reads line base + j at position p0 + lane; columns: line base + lane at p0 + j
      const int l = kRows ? base + j : line, p = kRows ? p0 + lane : p0 + j;
      v[j] = (l < nl && p >= 0 && p < len) ? in[at(l, p)] : 0.0f;
    }
  };
  auto park = [&](const float (&v)[kChunk], float (*s)[kLines + 1]) {
#pragma unroll
    for (int j = 0; j < kChunk; ++j) {
      if (kRows) s[lane][j] = v[j];  // load j held position lane of line j
      else s[j][lane] = v[j];
    }
  };
  // initial window sum of position 0 (positions 0 .. r in order)
  float sum = 0.0f;
  if (line < nl)
    for (int p = 0; p <= r; ++p) sum += in[at(line, p)];
  float lead[kChunk], trail[kChunk];
  load_block(r, lead);          // chunk 0 leading positions y + r
  load_block(-r - 1, trail);    // chunk 0 trailing positions y - r - 1
  for (int p0 = 0; p0 < len; p0 += kChunk) {
    __syncthreads();  // the previous chunk's LDS reads are done
    park(lead, lead_s);
    park(trail, trail_s);
    __syncthreads();
    if (p0 + kChunk < len) {  // next chunk's loads in flight during this chunk's steps
      load_block(p0 + kChunk + r, lead);
      load_block(p0 + kChunk - r - 1, trail);
    }
    const int steps = std::min(kChunk, len - p0);
    for (int j = 0; j < steps; ++j) {
      const int p = p0 + j;
      if (p > 0) {
        if (p > r) sum -= trail_s[j][lane];
        if (p + r < len) sum += lead_s[j][lane];
      }
      float v = sum;
      if (!kRows) {
        const float cy = (float)(std::min(r, p) + std::min(r, h - 1 - p) + 1);
        const float cx = (float)(std::min(rx, line) + std::min(rx, w - 1 - line) + 1);
        v = v / (cx * cy);
        if (line < nl) out[at(line, p)] = v;
      } else {
        out_s[j][lane] = v;
      }
    }
    if (kRows) {
      __syncthreads();
#pragma unroll 4
      for (int j = 0; j < kChunk; ++j) {  // line base + j, positions p0 + lane
        const int l = base + j, p = p0 + lane;
        if (l < nl && p < len) out[at(l, p)] = out_s[lane][j];
      }
    }
  }
}

int grid1d(size_t n) {
  size_t g = (n + 255) / 256;
  return (int)(g > 65536 ? 65536 : (g ? g : 1));
}

// box_filter (chroma_denoise.rs:100-152) of `planes` planes: in -> tmp (rows) -> out (columns).
void box_filter(const float* in, float* tmp, float* out, int w, int h, int radius, int planes, hipStream_t s) {
  const int rx = std::min(radius, w - 1), ry = std::min(radius, h - 1);
  hipLaunchKernelGGL(box_pass_kernel<true>, dim3((h + kLines - 1) / kLines, planes), dim3(kLines), 0, s, in, tmp, w, h,
                     rx, rx);
  hipLaunchKernelGGL(box_pass_kernel<false>, dim3((w + kLines - 1) / kLines, planes), dim3(kLines), 0, s, tmp, out, w,
                     h, ry, rx);
}

}  // namespace

// chroma_denoise (chroma_denoise.rs:14-38) on the image's RGB, in place.
void stage_luma_chroma_denoise(imops_ctx* ctx, imops_image* img, int radius, float epsilon) {
  require_rgb(img, "LumaGuidedChromaDenoise");
  if (radius < 0) throw Error{IMOPS_E_INVALID, "LumaGuidedChromaDenoise: radius must be >= 0"};
  TimedRegion stage_timer(ctx, "luma_chroma_denoise");
  const int w = (int)img->width, h = (int)img->height;
  const size_t n = (size_t)w * h;
  if (n == 0) return;
  hipStream_t s = ctx->stream;
  ctx->scr[SCR_BM3D_0].reserve(18 * n * sizeof(float));
  float* f = ctx->scr[SCR_BM3D_0].as<float>();
  float *prep = f, *tmp = f + 6 * n, *means = f + 12 * n;
  float* rgb = img->rgb.as<float>();
  hipLaunchKernelGGL(guided_prep_kernel, dim3(grid1d(n)), dim3(256), 0, s, rgb, n, prep);
  box_filter(prep, tmp, means, w, h, radius, 6, s);
  float* ab = prep;  // the prep planes are free again
  hipLaunchKernelGGL(guided_coef_kernel, dim3(grid1d(n)), dim3(256), 0, s, means, n, epsilon, ab);
  float* mab = means;
  box_filter(ab, tmp, mab, w, h, radius, 4, s);
  hipLaunchKernelGGL(guided_apply_kernel, dim3(grid1d(n)), dim3(256), 0, s, rgb, mab, n);
  IMOPS_HIP_CHECK(hipGetLastError());
}

}  // namespace imops

using namespace imops;

extern "C" int imops_luma_chroma_denoise(imops_ctx* ctx, imops_image* img, int radius, float epsilon) {
  return guarded([&] {
    if (!ctx || !img) throw Error{IMOPS_E_INVALID, "luma_chroma_denoise: null argument"};
    stage_luma_chroma_denoise(ctx, img, radius, epsilon);
  });
}
