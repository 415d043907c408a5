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

// crop_bayer_center (extern_pipeline.rs:113-179): destination (x, y) takes
// source 2x2 block (x/2, y/2) * factor, site (x%2, y%2): every factor-th CFA
// cell, so the Bayer phase is kept; samples past the plane read 0.
__global__ void crop_bayer_center_kernel(const float* __restrict__ in, size_t width, size_t n_in, size_t factor,
                                         size_t nw, size_t nh, float* __restrict__ out) {
  const size_t n = nw * nh;
  for (size_t k = (size_t)blockIdx.x * blockDim.x + threadIdx.x; k < n; k += (size_t)gridDim.x * blockDim.x) {
    const size_t y = k / nw, x = k - y * nw;
    const size_t sy = (y / 2) * factor * 2 + y % 2, sx = (x / 2) * factor * 2 + x % 2;
    const size_t idx = sy * width + sx;
    out[k] = idx < n_in ? in[idx] : 0.0f;
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

__global__ void preview_demosaic_kernel(const float* __restrict__ in, int width, int height, int f, int c00, int c01,
                                        int c10, int c11, float* __restrict__ out) {
  const int nw = width / f, nh = height / f;
  const size_t n = (size_t)nw * nh;
  for (size_t k = (size_t)blockIdx.x * blockDim.x + threadIdx.x; k < n; k += (size_t)gridDim.x * blockDim.x) {
    const size_t r = (k / nw) * f, c = (k % nw) * f;
    const size_t tl = r * width + c;
    const float s00 = in[tl], s01 = in[tl + 1], s10 = in[tl + width], s11 = in[tl + width + 1];
#pragma unroll
    for (int ch = 0; ch < 3; ++ch) {  // the sites are written in order: the last one of a colour wins
      float v = 0.0f;
      if (c00 == ch) v = s00;
      if (c01 == ch) v = s01;
      if (c10 == ch) v = s10;
      if (c11 == ch) v = s11;
      out[k * 3 + ch] = v;
    }
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

// Oklab with Ottosson's published matrices (linear sRGB -> LMS, cube root,
// LMS' -> Lab) in the polar form the colour crate calls Oklch: C = |(a, b)|,
// hue = atan2(b, a) in degrees, wrapped to [0, 360).
__device__ __forceinline__ void lin_to_oklch(float r, float g, float b, float& L, float& C, float& H) {
  const float l = 0.4122214708f * r + 0.5363325363f * g + 0.0514459929f * b;
  const float m = 0.2119034982f * r + 0.6806995451f * g + 0.1073969566f * b;
  const float s = 0.0883024619f * r + 0.2817188376f * g + 0.6299787005f * b;
  const float l_ = cbrtf(l), m_ = cbrtf(m), s_ = cbrtf(s);
  L = 0.2104542553f * l_ + 0.7936177850f * m_ - 0.0040720468f * s_;
  const float A = 1.9779984951f * l_ - 2.4285922050f * m_ + 0.4505937099f * s_;
  const float B = 0.0259040371f * l_ + 0.7827717662f * m_ - 0.8086757660f * s_;
  C = sqrtf(A * A + B * B);
  float h = atan2f(B, A) * (180.0f / 3.14159265358979f);
  if (h < 0.0f) h += 360.0f;
  H = h;
}

__device__ __forceinline__ void oklch_to_lin(float L, float C, float H, float& r, float& g, float& b) {
  const float hr = H * (3.14159265358979f / 180.0f);
  const float A = C * cosf(hr), B = C * sinf(hr);
  const float l_ = L + 0.3963377774f * A + 0.2158037573f * B;
  const float m_ = L - 0.1055613458f * A - 0.0638541728f * B;
  const float s_ = L - 0.0894841775f * A - 1.2914855480f * B;
  const float l = l_ * l_ * l_, m = m_ * m_ * m_, s = s_ * s_ * s_;
  r = 4.0767416621f * l - 3.3077115913f * m + 0.2309699292f * s;
  g = -1.2684380046f * l + 2.6097574011f * m - 0.3413193965f * s;
  b = -0.0041960863f * l - 0.7034186147f * m + 1.7076147010f * s;
}

// ACES reference gamut compression, forward (tone_map.rs:75-170, RgcParams::default)
__device__ __forceinline__ float rgc_compress(float dist, float lim, float thr, float power) {
  if (dist < thr) return dist;
  const float base_inner = (1.0f - thr) / (lim - thr);
  if (base_inner <= 0.0f || fabsf(lim - thr) < 1e-6f) return dist;
  const float denom_inner = powf(base_inner, -power) - 1.0f;
  if (denom_inner <= 0.0f) return dist;
  const float s = (lim - thr) / powf(denom_inner, 1.0f / power);
  const float dn = (dist - thr) / s;
  const float den = powf(1.0f + powf(dn, power), 1.0f / power);
  return thr + s * (dn / den);
}

__device__ __forceinline__ float square_sigmoid(float x) {  // tone_map.rs:7-15
  const float c = 1.219512f;
  if (x <= 0.0f) return 0.0f;
  return 1.0f / (1.0f + (1.0f / (c * x)));
}

__device__ __forceinline__ void mat3(const float* m, float& v0, float& v1, float& v2) {
  const float a = m[0] * v0 + m[1] * v1 + m[2] * v2;
  const float b = m[3] * v0 + m[4] * v1 + m[5] * v2;
  const float c = m[6] * v0 + m[7] * v1 + m[8] * v2;
  v0 = a; v1 = b; v2 = c;
}

// sigmoid (tone_map.rs:16-31) on one ACEScg pixel; a2l / l2a convert to and from linear sRGB
__device__ void sigmoid_tone_map(const float* a2l, const float* l2a, float& v0, float& v1, float& v2) {
  float g0, g1, g2;
  const float ach = fmaxf(fmaxf(v0, v1), v2);
  if (ach == 0.0f) {
    g0 = g1 = g2 = 0.0f;
  } else {
    const float aa = fabsf(ach);
    g0 = ach - rgc_compress((ach - v0) / aa, 1.147f, 0.815f, 1.2f) * aa;
    g1 = ach - rgc_compress((ach - v1) / aa, 1.264f, 0.803f, 1.2f) * aa;
    g2 = ach - rgc_compress((ach - v2) / aa, 1.312f, 0.88f, 1.2f) * aa;
  }
  float x0 = g0, x1 = g1, x2 = g2, L, C, H;
  mat3(a2l, x0, x1, x2);
  lin_to_oklch(x0, x1, x2, L, C, H);
  const float p0 = square_sigmoid(g0), p1 = square_sigmoid(g1), p2 = square_sigmoid(g2);
  const float sl = 0.2722287168f * p0 + 0.6740817658f * p1 + 0.0536895174f * p2;  // luminance (pixel.rs:43-47)
  const float mult = 1.0f - powf(sl, 5.0f);
  float y0 = p0, y1 = p1, y2 = p2, L2, C2, H2;
  mat3(a2l, y0, y1, y2);
  lin_to_oklch(y0, y1, y2, L2, C2, H2);
  oklch_to_lin(L2, C2 * mult, H, v0, v1, v2);
  mat3(l2a, v0, v1, v2);
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
        case PW_CONTRAST: {  // contrast.rs:4-8, MIDDLE_GRAY 0.185 (pixel.rs:36)
          const float c = op.m[0];
          v0 = 0.185f * powf(v0 / 0.185f, c);
          v1 = 0.185f * powf(v1 / 0.185f, c);
          v2 = 0.185f * powf(v2 / 0.185f, c);
          break;
        }
        case PW_HIGHLIGHT: {  // highlight_reconstruction.rs:5-14
          const float f = v1 / op.m[0];
          v1 = ((1.0f - f) * v1) + (f * (v0 + v2) * 0.5f);
          break;
        }
        case PW_LIN_TO_OKLCH: {
          float L, C, H;
          lin_to_oklch(v0, v1, v2, L, C, H);
          v0 = L; v1 = C; v2 = H;
          break;
        }
        case PW_OKLCH_TO_LIN: {
          float r, g, b;
          oklch_to_lin(v0, v1, v2, r, g, b);
          v0 = r; v1 = g; v2 = b;
          break;
        }
        case PW_SIGMOID: sigmoid_tone_map(op.m, op.m + 9, v0, v1, v2); break;
        default: break;
      }
    }
    rgb[p * 3] = v0;
    rgb[p * 3 + 1] = v1;
    rgb[p * 3 + 2] = v2;
  }
}

// apply_vignette_radial_correction (vignette.rs:71-96) for one FixVignetteRadial opcode
__global__ void vignette_kernel(float* __restrict__ rgb, int width, int height, VignetteParams v) {
  const size_t n = (size_t)width * height;
  const float wf = (float)(width - 1), hf = (float)(height - 1);
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
    const float py = (float)(i / width), px = (float)(i % width);
    const float u = px / wf, vv = py / hf;
    const float du = u - v.cx, dv = vv - v.cy;
    const float d = sqrtf(du * du + dv * dv);
    const float r = d / v.d_max;
    const float r2 = r * r, r4 = r2 * r2, r6 = r4 * r2, r8 = r4 * r4, r10 = r8 * r2;
    const float corr = v.k[0] * r2 + v.k[1] * r4 + v.k[2] * r6 + v.k[3] * r8 + v.k[4] * r10;
    const float gain = fmaxf(1.0f + v.strength * corr, 0.0f);
    rgb[i * 3] *= gain;
    rgb[i * 3 + 1] *= gain;
    rgb[i * 3 + 2] *= gain;
  }
}

// output pixel of input (x, y) under the orientation; the output is w x h, or h x w when rotated by 90/270
__device__ __forceinline__ size_t oriented_index(int x, int y, int w, int h, int orientation) {
  switch (orientation) {
    case 1: return (size_t)x * h + (h - 1 - y);              // Rotate90: (x, y) -> (h-1-y, x)
    case 2: return (size_t)(h - 1 - y) * w + (w - 1 - x);    // Rotate180
    case 3: return (size_t)(w - 1 - x) * h + y;              // Rotate270: (x, y) -> (y, w-1-x)
    default: return (size_t)y * w + x;
  }
}

__global__ void encode_u8_kernel(const float* __restrict__ rgb, int w, int h, int orientation,
                                 uint8_t* __restrict__ out) {
  const size_t n = (size_t)w * h;
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
    const size_t o = oriented_index((int)(i % w), (int)(i / w), w, h, orientation);
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      const float v = rgb[i * 3 + c];
      const float q = (v < 0.0f ? 0.0f : (v > 1.0f ? 1.0f : v)) * 255.0f;  // NaN stays NaN
      out[o * 3 + c] = q == q ? (uint8_t)(int)q : (uint8_t)0;              // `as u8`: NaN -> 0, truncation
    }
  }
}

__global__ void encode_f16_kernel(const float* __restrict__ rgb, int w, int h, int orientation,
                                  uint16_t* __restrict__ out) {
  const size_t n = (size_t)w * h;
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
    const size_t o = oriented_index((int)(i % w), (int)(i / w), w, h, orientation);
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      const _Float16 hv = (_Float16)rgb[i * 3 + c];  // v_cvt_f16_f32: round to nearest even
      out[o * 3 + c] = __builtin_bit_cast(uint16_t, hv);
    }
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

hipError_t crop_bayer_center_launch(const float* in, size_t width, size_t height, size_t factor, size_t nw, size_t nh,
                                   float* out, hipStream_t s) {
  hipLaunchKernelGGL(crop_bayer_center_kernel, dim3(grid_for(nw * nh, 256)), dim3(256), 0, s, in, width,
                     width * height, factor, nw, nh, out);
  return hipGetLastError();
}

hipError_t preview_demosaic_launch(const float* in, int width, int height, int factor, const int cell[4], float* out,
                                   hipStream_t s) {
  const size_t n = (size_t)(width / factor) * (height / factor);
  hipLaunchKernelGGL(preview_demosaic_kernel, dim3(grid_for(n, 256)), dim3(256), 0, s, in, width, height, factor,
                     cell[0], cell[1], cell[2], cell[3], out);
  return hipGetLastError();
}

hipError_t markesteijn_launch(const float* in, int width, int height, const CfaPattern& cfa, float* out,
                              hipStream_t s) {
  hipLaunchKernelGGL(markesteijn_kernel, dim3(grid_for((size_t)width * height, 256)), dim3(256), 0, s, in, width,
                     height, cfa, out);
  return hipGetLastError();
}

hipError_t vignette_launch(float* rgb, int width, int height, const VignetteParams& v, hipStream_t s) {
  hipLaunchKernelGGL(vignette_kernel, dim3(grid_for((size_t)width * height, 256)), dim3(256), 0, s, rgb, width, height,
                     v);
  return hipGetLastError();
}

hipError_t encode_u8_launch(const float* rgb, int width, int height, int orientation, uint8_t* out, hipStream_t s) {
  hipLaunchKernelGGL(encode_u8_kernel, dim3(grid_for((size_t)width * height, 256)), dim3(256), 0, s, rgb, width, height,
                     orientation, out);
  return hipGetLastError();
}

hipError_t encode_f16_launch(const float* rgb, int width, int height, int orientation, uint16_t* out, hipStream_t s) {
  hipLaunchKernelGGL(encode_f16_kernel, dim3(grid_for((size_t)width * height, 256)), dim3(256), 0, s, rgb, width, height,
                     orientation, out);
  return hipGetLastError();
}

hipError_t pointwise_launch(float* rgb, size_t npix, const PointwiseProgram& prog, hipStream_t s) {
  if (prog.n == 0) return hipSuccess;
  hipLaunchKernelGGL(pointwise_kernel, dim3(grid_for(npix, 256)), dim3(256), 0, s, rgb, npix, prog);
  return hipGetLastError();
}

}  // namespace imops
