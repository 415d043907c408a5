// Three-pass Markesteijn X-Trans demosaic (the markesteijn3 extension of the
// Demosaic stage) for gfx950.  The method (green along four directions clamped
// to the local green range; with three passes four more green candidates
// recomputed from the adjacent pixels' interpolated greens; red / blue of every
// candidate by colour differences along its direction; a (Y, Pb, Pr)
// derivative per candidate; 5x5 homogeneity counts; the mean of the candidates
// within 7/8 of the best count) and its FP32 operation order are those of the
// oracle's definition; this file reproduces them bit for bit (no contraction,
// correctly rounded division, min/max written as the comparisons they are).
//
// Layout: candidate planes in scratch (8 green, 8 red, 8 blue f32 planes,
// W x H each) and one byte plane of per-candidate homogeneity flags; every
// kernel is one thread per pixel over the frame.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "pointwise.hpp"

namespace imops {
namespace {

__constant__ int kDy[4] = {0, 1, 1, 1};
__constant__ int kDx[4] = {1, 0, 1, -1};

struct M3Args {
  const float* in;
  int W, H;
  CfaPattern cfa;
  int nc;      // candidates: 4 (one pass) or 8 (three passes)
  float* G;    // nc planes
  float* R;
  float* B;
  uint8_t* flags;
  float* out;  // interleaved RGB
};

__device__ __forceinline__ float fmax_(float a, float b) { return a < b ? b : a; }  // std::max
__device__ __forceinline__ float fmin_(float a, float b) { return b < a ? b : a; }  // std::min

__device__ __forceinline__ bool inb(const M3Args& a, int y, int x) { return y >= 0 && y < a.H && x >= 0 && x < a.W; }
__device__ __forceinline__ int col(const M3Args& a, int y, int x) { return a.cfa.at(y, x); }
__device__ __forceinline__ float val(const M3Args& a, int y, int x) { return a.in[(size_t)y * a.W + x]; }

__device__ __forceinline__ int near(const M3Args& a, int y, int x, int d, int s, int ch) {
  for (int k = 1; k <= 3; ++k) {
    const int yy = y + s * k * kDy[d], xx = x + s * k * kDx[d];
    if (!inb(a, yy, xx)) return 0;
    if (col(a, yy, xx) == ch) return k;
  }
  return 0;
}

// min / max of the green CFA values of the 3x3 window, else the 5x5 window
__device__ bool green_bounds(const M3Args& a, int y, int x, float& lo, float& hi) {
  bool any = false;
  for (int r = 1; r <= 2 && !any; ++r)
    for (int yy = y - r; yy <= y + r; ++yy)
      for (int xx = x - r; xx <= x + r; ++xx) {
        if (!inb(a, yy, xx) || col(a, yy, xx) != 1) continue;
        const float g = val(a, yy, xx);
        if (!any) {
          lo = hi = g;
          any = true;
        } else {
          lo = fmin_(lo, g);
          hi = fmax_(hi, g);
        }
      }
  return any;
}

__global__ void m3_green_kernel(M3Args a) {
  const size_t n = (size_t)a.W * a.H;
  for (size_t p = (size_t)blockIdx.x * blockDim.x + threadIdx.x; p < n; p += (size_t)gridDim.x * blockDim.x) {
    const int y = (int)(p / a.W), x = (int)(p - (size_t)y * a.W);
    const int c = col(a, y, x);
    float lo = 0.0f, hi = 0.0f;
    const bool hb = c != 1 && green_bounds(a, y, x, lo, hi);
    for (int d = 0; d < 4; ++d) {
      float g;
      if (c == 1) {
        g = val(a, y, x);
      } else {
        const int ka = near(a, y, x, d, -1, 1), kb = near(a, y, x, d, 1, 1);
        if (ka && kb)
          g = ((float)kb * val(a, y - ka * kDy[d], x - ka * kDx[d]) + (float)ka * val(a, y + kb * kDy[d], x + kb * kDx[d])) /
              (float)(ka + kb);
        else if (ka) g = val(a, y - ka * kDy[d], x - ka * kDx[d]);
        else if (kb) g = val(a, y + kb * kDy[d], x + kb * kDx[d]);
        else g = hb ? (lo + hi) * 0.5f : val(a, y, x);
        if (hb) g = fmin_(fmax_(g, lo), hi);
      }
      a.G[(size_t)d * n + p] = g;
    }
  }
}

// candidates 4..7: where the nearest green along -d or +d is not adjacent, the
// mean of candidate d at the two adjacent pixels
__global__ void m3_green2_kernel(M3Args a) {
  const size_t n = (size_t)a.W * a.H;
  for (size_t p = (size_t)blockIdx.x * blockDim.x + threadIdx.x; p < n; p += (size_t)gridDim.x * blockDim.x) {
    const int y = (int)(p / a.W), x = (int)(p - (size_t)y * a.W);
    const int c = col(a, y, x);
    float lo = 0.0f, hi = 0.0f;
    const bool hb = c != 1 && green_bounds(a, y, x, lo, hi);
    for (int d = 0; d < 4; ++d) {
      float g = a.G[(size_t)d * n + p];
      if (c != 1) {
        const int ka = near(a, y, x, d, -1, 1), kb = near(a, y, x, d, 1, 1);
        const int ya = y - kDy[d], xa = x - kDx[d], yb = y + kDy[d], xb = x + kDx[d];
        if (!(ka == 1 && kb == 1) && inb(a, ya, xa) && inb(a, yb, xb))
          g = (a.G[(size_t)d * n + (size_t)ya * a.W + xa] + a.G[(size_t)d * n + (size_t)yb * a.W + xb]) * 0.5f;
        if (hb) g = fmin_(fmax_(g, lo), hi);
      }
      a.G[(size_t)(4 + d) * n + p] = g;
    }
  }
}

__global__ void m3_chroma_kernel(M3Args a) {
  const size_t n = (size_t)a.W * a.H;
  for (size_t p = (size_t)blockIdx.x * blockDim.x + threadIdx.x; p < n; p += (size_t)gridDim.x * blockDim.x) {
    const int y = (int)(p / a.W), x = (int)(p - (size_t)y * a.W);
    const int c = col(a, y, x);
    for (int k = 0; k < a.nc; ++k) {
      const int d = k & 3;
      const float* Gk = a.G + (size_t)k * n;
      float res[2];
      for (int j = 0; j < 2; ++j) {
        const int ch = j == 0 ? 0 : 2;
        if (c == ch) {
          res[j] = val(a, y, x);
          continue;
        }
        const int ka = near(a, y, x, d, -1, ch), kb = near(a, y, x, d, 1, ch);
        float cd = 0.0f;
        if (ka || kb) {
          const int ya = y - ka * kDy[d], xa = x - ka * kDx[d], yb = y + kb * kDy[d], xb = x + kb * kDx[d];
          if (ka && kb)
            cd = ((float)kb * (val(a, ya, xa) - Gk[(size_t)ya * a.W + xa]) +
                  (float)ka * (val(a, yb, xb) - Gk[(size_t)yb * a.W + xb])) /
                 (float)(ka + kb);
          else if (ka) cd = val(a, ya, xa) - Gk[(size_t)ya * a.W + xa];
          else cd = val(a, yb, xb) - Gk[(size_t)yb * a.W + xb];
        } else {
          for (int r = 1; r <= 2; ++r) {
            float acc = 0.0f;
            int cnt = 0;
            for (int yy = y - r; yy <= y + r; ++yy)
              for (int xx = x - r; xx <= x + r; ++xx) {
                if (!inb(a, yy, xx) || col(a, yy, xx) != ch) continue;
                acc += val(a, yy, xx) - Gk[(size_t)yy * a.W + xx];
                ++cnt;
              }
            if (cnt) {
              cd = acc / (float)cnt;
              break;
            }
          }
        }
        res[j] = Gk[p] + cd;
      }
      a.R[(size_t)k * n + p] = res[0];
      a.B[(size_t)k * n + p] = res[1];
    }
  }
}

__device__ __forceinline__ void ypp(const M3Args& a, size_t n, int k, size_t q, float o[3]) {
  const float r = a.R[(size_t)k * n + q], g = a.G[(size_t)k * n + q], b = a.B[(size_t)k * n + q];
  const float Y = 0.2627f * r + 0.678f * g + 0.0593f * b;
  o[0] = Y;
  o[1] = (b - Y) * 0.5315f;
  o[2] = (r - Y) * 0.6781f;
}

__global__ void m3_flags_kernel(M3Args a) {
  const size_t n = (size_t)a.W * a.H;
  for (size_t p = (size_t)blockIdx.x * blockDim.x + threadIdx.x; p < n; p += (size_t)gridDim.x * blockDim.x) {
    const int y = (int)(p / a.W), x = (int)(p - (size_t)y * a.W);
    float dd[8];
    float mn = 0.0f;
    for (int k = 0; k < a.nc; ++k) {
      const int d = k & 3;
      float c0[3];
      ypp(a, n, k, p, c0);
      float part[2];
      for (int j = 0; j < 2; ++j) {
        const int s = j == 0 ? -1 : 1;
        const int yy = min(max(y + s * kDy[d], 0), a.H - 1), xx = min(max(x + s * kDx[d], 0), a.W - 1);
        float c1[3];
        ypp(a, n, k, (size_t)yy * a.W + xx, c1);
        const float e0 = c0[0] - c1[0], e1 = c0[1] - c1[1], e2 = c0[2] - c1[2];
        part[j] = e0 * e0 + e1 * e1 + e2 * e2;
      }
      dd[k] = part[0] + part[1];
      mn = k == 0 ? dd[k] : fmin_(mn, dd[k]);
    }
    const float tr = 8.0f * mn;
    unsigned m = 0;
    for (int k = 0; k < a.nc; ++k) m |= (dd[k] <= tr ? 1u : 0u) << k;
    a.flags[p] = (uint8_t)m;
  }
}

__global__ void m3_select_kernel(M3Args a) {
  const size_t n = (size_t)a.W * a.H;
  for (size_t p = (size_t)blockIdx.x * blockDim.x + threadIdx.x; p < n; p += (size_t)gridDim.x * blockDim.x) {
    const int y = (int)(p / a.W), x = (int)(p - (size_t)y * a.W);
    int sc[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    for (int yy = max(y - 2, 0); yy <= min(y + 2, a.H - 1); ++yy)
      for (int xx = max(x - 2, 0); xx <= min(x + 2, a.W - 1); ++xx) {
        const unsigned m = a.flags[(size_t)yy * a.W + xx];
#pragma unroll
        for (int k = 0; k < 8; ++k) sc[k] += (m >> k) & 1u;
      }
    int mx = 0;
    for (int k = 0; k < a.nc; ++k) mx = max(mx, sc[k]);
    float acc[3] = {0.0f, 0.0f, 0.0f};
    int cnt = 0;
    for (int k = 0; k < a.nc; ++k) {
      if (8 * sc[k] < 7 * mx) continue;
      acc[0] += a.R[(size_t)k * n + p];
      acc[1] += a.G[(size_t)k * n + p];
      acc[2] += a.B[(size_t)k * n + p];
      ++cnt;
    }
    for (int j = 0; j < 3; ++j) a.out[p * 3 + j] = acc[j] / (float)cnt;
  }
}

int grid_of(size_t n) {
  const size_t g = (n + 255) / 256;
  return (int)(g > 65536 ? 65536 : (g ? g : 1));
}

}  // namespace

size_t markesteijn3_scratch_bytes(int width, int height, int passes) {
  const size_t n = (size_t)width * height;
  const int nc = passes > 1 ? 8 : 4;
  return 3 * (size_t)nc * n * sizeof(float) + n;
}

hipError_t markesteijn3_launch(const float* in, int width, int height, const CfaPattern& cfa, int passes,
                               void* scratch, float* out, hipStream_t s) {
  if (passes != 1 && passes != 3) return hipErrorInvalidValue;
  M3Args a;
  a.in = in;
  a.W = width;
  a.H = height;
  a.cfa = cfa;
  a.nc = passes > 1 ? 8 : 4;
  const size_t n = (size_t)width * height;
  a.G = static_cast<float*>(scratch);
  a.R = a.G + (size_t)a.nc * n;
  a.B = a.R + (size_t)a.nc * n;
  a.flags = reinterpret_cast<uint8_t*>(a.B + (size_t)a.nc * n);
  a.out = out;
  const dim3 g(grid_of(n)), b(256);
  hipLaunchKernelGGL(m3_green_kernel, g, b, 0, s, a);
  if (a.nc == 8) hipLaunchKernelGGL(m3_green2_kernel, g, b, 0, s, a);
  hipLaunchKernelGGL(m3_chroma_kernel, g, b, 0, s, a);
  hipLaunchKernelGGL(m3_flags_kernel, g, b, 0, s, a);
  hipLaunchKernelGGL(m3_select_kernel, g, b, 0, s, a);
  return hipGetLastError();
}

}  // namespace imops
