// Spectral film simulation kernels for gfx950 (pichromatic/src/film).
//
// Plane layout: every emulsion quantity is a planar f32 W x H array
// (emulsion-major), so each stage streams unit-stride rows.
//
//  * expose: the reference's spectral walk (upsample -> photon weighting ->
//    Beer-Lambert through every absorbing layer -> trapezoid band mean,
//    exposure/mod.rs:68-81, absorption.rs:24-108, upsample.rs:145-170) is linear
//    in the three basis weights, so the host collapses it over the grid into a
//    3 x E matrix and the device evaluates relu(M^-1 rgb) . C per pixel in f64
//    (film_fit_expose in film_fit.hip);
//  * blur: reflect-101 separable Gaussian, one output per thread, taps summed
//    in the reference's order (film/blur.rs:10-89), with fused epilogues for
//    local scatter, adjacency and plain stores;
//  * capture: halation bounce add, reciprocity, log-spaced LUT sample (f64),
//    H&D reduction to image/mask dye in one pass (exposure/mod.rs:97-115,
//    reduction.rs:16-68);
//  * DIR inhibition, grain (SplitMix64 in counter form: value k of the row
//    stream is mix(key + k * golden), bit-identical to the reference's
//    sequential generator), and the densitometric scan + optional positive
//    invert (f64 per bin in the reference's order; the mixed form in f32 with
//    log2(10)-scaled absorptivities and folded trapezoid weights).
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "film.hpp"

namespace imops {
namespace {

__device__ __forceinline__ int reflect_index(long long i, long long n) {
  long long x = i;
  for (;;) {
    if (x < 0) x = -x;
    else if (x >= n) x = 2 * n - 2 - x;
    else return (int)x;
    if (n == 1) return 0;
  }
}

__global__ void blur_h_kernel(const float* __restrict__ in, float* __restrict__ out, int w, int h,
                              const float* __restrict__ taps, int radius, BlurPlanes pl) {
  const size_t n = (size_t)w * h;
  in += blockIdx.y * pl.in;
  out += blockIdx.y * n;
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
    const int y = (int)(i / w), x = (int)(i - (size_t)y * w);
    const float* row = in + (size_t)y * w;
    float acc = 0.0f;
    if (x >= radius && x + radius < w) {
      for (int k = 0; k <= 2 * radius; ++k) acc += row[x + k - radius] * taps[k];
    } else {
      for (int k = 0; k <= 2 * radius; ++k) acc += row[reflect_index((long long)x + k - radius, w)] * taps[k];
    }
    out[i] = acc;
  }
}

// Vertical pass + epilogue on the unblurred plane `orig` (may alias `out`).
__global__ void blur_v_kernel(const float* __restrict__ tmp, const float* orig, float* out, int w, int h,
                              const float* __restrict__ taps, int radius, int mode, float a, BlurPlanes pl) {
  const size_t n = (size_t)w * h;
  tmp += blockIdx.y * n;
  if (orig) orig += blockIdx.y * pl.orig;
  out += blockIdx.y * pl.out;
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
    const int y = (int)(i / w), x = (int)(i - (size_t)y * w);
    float acc = 0.0f;
    if (y >= radius && y + radius < h) {
      for (int k = 0; k <= 2 * radius; ++k) acc += tmp[(size_t)(y + k - radius) * w + x] * taps[k];
    } else {
      for (int k = 0; k <= 2 * radius; ++k) acc += tmp[(size_t)reflect_index((long long)y + k - radius, h) * w + x] * taps[k];
    }
    float v;
    switch (mode) {
      case BLUR_MIX: {  // (1-f) p + f blur(p)      halation.rs:164-181
        const float keep = 1.0f - a;
        v = keep * orig[i] + a * acc;
        break;
      }
      case BLUR_ADJ: {  // d + beta (d - blur(d))   diffusion.rs:23-26
        const float d = orig[i];
        v = d + a * (d - acc);
        break;
      }
      default: v = acc;
    }
    out[i] = v;
  }
}

// Halation bounce + reciprocity + capture LUT + reduction (image and mask dye).
template <typename Real>
__global__ void capture_reduce_kernel(const float* __restrict__ absorbed, const float* __restrict__ bounce, size_t n,
                                      const FilmDeviceTables* __restrict__ T, float* __restrict__ image,
                                      float* __restrict__ mask) {
  const size_t b = blockIdx.y, planes = (size_t)T->n_emulsions * n;  // image b of a batch
  absorbed += b * planes;
  if (bounce) bounce += b * n;
  image += b * planes;
  mask += b * planes;
  for (size_t p = (size_t)blockIdx.x * blockDim.x + threadIdx.x; p < n; p += (size_t)gridDim.x * blockDim.x) {
    for (int e = 0; e < T->n_emulsions; ++e) {
      const FilmEmulsionTables& E = T->em[e];
      float phi = absorbed[(size_t)e * n + p];
      if (bounce && E.gain > 0.0f) phi += E.gain * bounce[p];
      const Real x = (Real)(phi * E.eta);
      Real f;
      if (!(x > (Real)0)) {
        f = (Real)E.lut_frac[0];
      } else {
        const Real lp = log10(x);
        const int last = E.lut_n - 1;
        if (lp <= (Real)E.lut_log[0]) {
          f = (Real)E.lut_frac[0];
        } else if (lp >= (Real)E.lut_log[last]) {
          f = (Real)E.lut_frac[last];
        } else {
          int lo = 0, hi = last;
          while (hi - lo > 1) {
            const int mid = (lo + hi) / 2;
            if ((Real)E.lut_log[mid] <= lp) lo = mid;
            else hi = mid;
          }
          const Real t = (lp - (Real)E.lut_log[lo]) / ((Real)E.lut_log[hi] - (Real)E.lut_log[lo]);
          f = (Real)E.lut_frac[lo] * ((Real)1 - t) + (Real)E.lut_frac[hi] * t;
        }
      }
      const float fr = (float)f;
      const float c = fr < 0.0f ? 0.0f : (fr > 1.0f ? 1.0f : fr);
      const float eff = E.reversal ? 1.0f - c : c;
      const float di = E.d_max * powf(eff, E.inv_gamma);
      image[(size_t)e * n + p] = di;
      float m = 0.0f;
      if (E.has_mask) {
        const float q = di / E.d_max;
        m = E.mask_scale * (1.0f - (q < 0.0f ? 0.0f : (q > 1.0f ? 1.0f : q)));
      }
      mask[(size_t)e * n + p] = m;
    }
  }
}

// DIR inhibition (diffusion.rs:62-82): dye_j *= exp(-sum_i M[i][j] blur(dye_i)).
__global__ void dir_apply_kernel(float* __restrict__ image, const float* __restrict__ inh, size_t n,
                                 const FilmDeviceTables* __restrict__ T) {
  const int E = T->n_emulsions;
  for (size_t p = (size_t)blockIdx.x * blockDim.x + threadIdx.x; p < n; p += (size_t)gridDim.x * blockDim.x) {
    // the pixel's blurred dyes, loaded together (fixed trip counts)
    float hv[kFilmMaxEmulsions];
#pragma unroll
    for (int i = 0; i < kFilmMaxEmulsions; ++i) hv[i] = i < E ? inh[(size_t)i * n + p] : 0.0f;
#pragma unroll
    for (int j = 0; j < kFilmMaxEmulsions; ++j) {
      if (j >= E) break;
      float total = 0.0f;
      bool active = false;
#pragma unroll
      for (int i = 0; i < kFilmMaxEmulsions; ++i) {
        if (i >= E) break;
        const float wgt = T->dir[i][j];
        if (T->dir_valid[i][j] && fabsf(wgt) > 1e-6f) {
          active = true;
          total += wgt * hv[i];
        }
      }
      if (active) image[(size_t)j * n + p] *= expf(-total);
    }
  }
}

__device__ __forceinline__ uint64_t mix64(uint64_t z) {
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

// Irwin-Hall noise of grain.rs:38-45/72-82: pixel x of row y uses outputs
// 12x+1 .. 12x+12 of the stream keyed (seed, layer, y).
__global__ void grain_noise_kernel(float* __restrict__ noise, int w, int h, uint64_t key0) {
  const size_t n = (size_t)w * h;
  const float umax = (float)UINT64_MAX;
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
    const uint64_t y = i / w, x = i - y * w;
    const uint64_t key = key0 + y;
    uint64_t state = key + (12ull * x) * 0x9E3779B97F4A7C15ull;
    float acc = 0.0f;
    for (int k = 0; k < 12; ++k) {
      state += 0x9E3779B97F4A7C15ull;
      acc += __ull2float_rn(mix64(state)) / umax;
    }
    noise[i] = acc - 6.0f;
  }
}

// Philox4x32-10 (Salmon, Moraes, Dror, Shaw: "Parallel random numbers: as easy
// as 1, 2, 3", SC 2011): ten rounds of two 32x32->64 multiplies with the
// Weyl key schedule.
__device__ __forceinline__ void philox4x32_10(uint32_t c[4], uint32_t k0, uint32_t k1) {
  for (int r = 0; r < 10; ++r) {
    const uint64_t p0 = (uint64_t)0xD2511F53u * c[0], p1 = (uint64_t)0xCD9E8D57u * c[2];
    const uint32_t h0 = (uint32_t)(p0 >> 32), l0 = (uint32_t)p0, h1 = (uint32_t)(p1 >> 32), l1 = (uint32_t)p1;
    const uint32_t n0 = h1 ^ c[1] ^ k0, n2 = h0 ^ c[3] ^ k1;
    c[0] = n0;
    c[1] = l1;
    c[2] = n2;
    c[3] = l0;
    k0 += 0x9E3779B9u;
    k1 += 0xBB67AE85u;
  }
}

// Gaussian grain noise from Philox: counter (pixel index low, high, layer, 0),
// key (seed low, high); Box-Muller on the first two words, in f64.
__global__ void grain_philox_kernel(float* __restrict__ noise, size_t n, uint64_t seed, uint32_t layer) {
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
    uint32_t c[4] = {(uint32_t)i, (uint32_t)((uint64_t)i >> 32), layer, 0u};
    philox4x32_10(c, (uint32_t)seed, (uint32_t)(seed >> 32));
    const double u1 = ((double)c[0] + 0.5) * 2.3283064365386963e-10;  // (x + 1/2) / 2^32, in (0, 1)
    const double u2 = ((double)c[1] + 0.5) * 2.3283064365386963e-10;
    noise[i] = (float)(sqrt(-2.0 * log(u1)) * cos(6.283185307179586477 * u2));
  }
}

// Sum of squares in f64: per-block partials, then one ordered pass.
__global__ void sumsq_partial_kernel(const float* __restrict__ x, size_t n, double* __restrict__ partial) {
  __shared__ double s[256];
  double acc = 0.0;
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
    const double v = (double)x[i];
    acc += v * v;
  }
  s[threadIdx.x] = acc;
  __syncthreads();
  for (int off = 128; off > 0; off >>= 1) {
    if ((int)threadIdx.x < off) s[threadIdx.x] += s[threadIdx.x + off];
    __syncthreads();
  }
  if (threadIdx.x == 0) partial[blockIdx.x] = s[0];
}

__global__ void grain_norm_kernel(const double* __restrict__ partial, int nparts, size_t n, float* __restrict__ norm) {
  if (threadIdx.x == 0 && blockIdx.x == 0) {
    double sum = 0.0;
    for (int k = 0; k < nparts; ++k) sum += partial[k];
    const double var = sum / (double)n;
    *norm = var > 1e-12 ? (float)(1.0 / sqrt(var)) : 1.0f;
  }
}

// grain.rs:91-98
// absolute: the density sigma is `kappa` itself (no taper, no sqrt(D (Dmax - D)))
__global__ void grain_apply_kernel(float* __restrict__ dye, const float* __restrict__ noise, size_t n,
                                   const float* __restrict__ norm_p, float d_max, float kappa, int absolute) {
  const float norm = *norm_p;
  const float eps_toe = 0.02f * d_max;
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
    const float v0 = dye[i];
    const float dens = v0 < 0.0f ? 0.0f : (v0 > d_max ? d_max : v0);
    const float taper = fminf(dens / (dens + eps_toe), 1.0f);
    const float sigma_d = absolute ? 1.0f : taper * sqrtf(fmaxf(dens * (d_max - dens), 0.0f));
    float v = dens + kappa * sigma_d * noise[i] * norm;
    const float hi = d_max * 1.05f;
    dye[i] = v < 0.0f ? 0.0f : (v > hi ? hi : v);
  }
}

// Densitometric scan (densitometry.rs:47-73) and the positive invert (invert.rs:36-108),
// streamed bin by bin: the density of bin i (emulsions in order), its 10^-d, and
// each CMF trapezoid's term in bin order (no per-bin arrays).
template <typename Real, int NB>
__global__ void scan_kernel(const float* __restrict__ image, const float* __restrict__ mask, size_t n,
                            const FilmDeviceTables* __restrict__ T, float* __restrict__ rgb) {
  const int E = T->n_emulsions;
  const size_t b = blockIdx.y;  // image b of a batch
  image += b * E * n;
  mask += b * E * n;
  rgb += b * 3 * n;
  const Real dl = (Real)T->dl;
  for (size_t p = (size_t)blockIdx.x * blockDim.x + threadIdx.x; p < n; p += (size_t)gridDim.x * blockDim.x) {
    Real di[kFilmMaxEmulsions], dm[kFilmMaxEmulsions];
#pragma unroll
    for (int e = 0; e < kFilmMaxEmulsions; ++e) {
      di[e] = e < E ? (Real)image[(size_t)e * n + p] : (Real)0;
      dm[e] = e < E ? (Real)mask[(size_t)e * n + p] : (Real)0;
    }
    Real xyz[3] = {0, 0, 0}, prev[3] = {0, 0, 0};
    for (int i = 0; i < NB; ++i) {
      Real dens = 0;
#pragma unroll
      for (int e = 0; e < kFilmMaxEmulsions; ++e) {
        if (e >= E) break;
        const FilmEmulsionTables& M = T->em[e];
        dens += di[e] * (Real)M.eps[i];
        if (M.has_mask) dens += dm[e] * (Real)M.mask_eps[i];
      }
      const Real t = exp(-dens * (Real)2.302585092994045684) * (Real)T->scanner[i];  // 10^-d
#pragma unroll
      for (int c = 0; c < 3; ++c) {
        const Real a = t * (Real)T->cmf[c][i];
        if (i > 0) xyz[c] += (Real)0.5 * (prev[c] + a) * dl;
        prev[c] = a;
      }
    }
    float out[3];
    for (int c = 0; c < 3; ++c) {
      const Real v = (Real)T->xyz2ap1[c][0] * xyz[0] + (Real)T->xyz2ap1[c][1] * xyz[1] + (Real)T->xyz2ap1[c][2] * xyz[2];
      out[c] = (float)((Real)(float)v * (Real)T->dmin_scale);
    }
    if (T->invert) {
      const float eps = 1e-6f;
      float raw[3];
      for (int c = 0; c < 3; ++c) {
        float tt = out[c] / T->inv_dmin[c];
        tt = tt < eps ? eps : (tt > 1.0f ? 1.0f : tt);
        const float inv = fmaxf(1.0f / tt - 1.0f, 0.0f);
        const float iv = fmaxf(inv, 0.0f);
        const float sh = fmaxf(T->inv_shoulder[c], 1e-6f);
        raw[c] = fmaxf(T->inv_gain[c] * (iv / (1.0f + iv / sh)), 0.0f);
      }
      const float y = fmaxf(0.2722287168f * raw[0] + 0.6740817658f * raw[1] + 0.0536895174f * raw[2], 0.0f);
      float k;
      if (y <= 0.0f) k = 0.0f;
      else if (y >= 0.02f) k = 0.55f;
      else k = 0.55f * (y / 0.02f);
      for (int c = 0; c < 3; ++c) out[c] = fmaxf(y + k * (raw[c] - y), 0.0f);
    }
    rgb[p * 3] = out[0];
    rgb[p * 3 + 1] = out[1];
    rgb[p * 3 + 2] = out[2];
  }
}

int grid1d(size_t n) {
  size_t g = (n + 255) / 256;
  return (int)(g > 65536 ? 65536 : (g ? g : 1));
}

// LDS-tiled separable passes (gaussian_blur_separable, blur.rs:10-37), with the
// same k-ordered sums as blur_h_kernel / blur_v_kernel.  Rows: one workgroup
// per segment of kRowSeg outputs; the segment and its reflected halo are staged
// with coalesced loads.  Columns: one workgroup per 64-column x th-row tile,
// one lane per column, each tile row a 256-byte load; th is kColH on large
// images and halves (down to kColHMin) until the grid has kColTiles tiles, so
// small images (the adjoint's crops) still fill the chip.  Each thread makes 4
// neighbouring outputs at once and slides a register window along the staged
// data, so one LDS read feeds 4 (columns) or 16 (rows: 16-byte quads) products;
// every output still adds its taps one at a time in tap order.
constexpr int BT = 256;
constexpr int kRowSeg = 2048;
constexpr int kColH = 128;
constexpr int kColHMin = 16;
constexpr int kColTiles = 1024;

// Row LDS: taps (padded to a quad), then the segment + halo at a 16-byte aligned
// offset, with 8 slack floats for the last quads.  Column LDS: taps, then
// th + 2r rows of 64 with 4 slack rows for the last window.
__host__ __device__ inline int row_buf_off(int nt) { return (nt + 3) & ~3; }
size_t blur_row_smem(int radius) {
  return (size_t)(row_buf_off(2 * radius + 1) + kRowSeg + 2 * radius + 8) * sizeof(float);
}
size_t blur_col_smem(int radius, int th) { return (size_t)(2 * radius + 1 + (th + 2 * radius + 4) * 64) * sizeof(float); }

int blur_col_tiles(int w, int h, int th) { return ((w + 63) / 64) * ((h + th - 1) / th); }

int blur_col_height(int w, int h) {
  int th = kColH;
  while (th > kColHMin && blur_col_tiles(w, h, th) < kColTiles) th /= 2;
  return th;
}

// a[j] = sum_k tp[k] bp[j + k], j < 4, over a 16-byte-aligned LDS row (tp zero-padded
// to a multiple of 4; reads up to bp[nt + 6]).  q holds bp[k .. k+3], nq the next quad.
__device__ __forceinline__ void row_conv4(const float* bp, const float* tp, int nt, float (&a)[4]) {
  float a0 = 0.0f, a1 = 0.0f, a2 = 0.0f, a3 = 0.0f;
  float4 q = *reinterpret_cast<const float4*>(bp);
  int k = 0;
  for (; k + 4 <= nt; k += 4) {
    const float4 nq = *reinterpret_cast<const float4*>(bp + k + 4);
    const float4 tk = *reinterpret_cast<const float4*>(tp + k);
    a0 += q.x * tk.x; a1 += q.y * tk.x; a2 += q.z * tk.x; a3 += q.w * tk.x;
    a0 += q.y * tk.y; a1 += q.z * tk.y; a2 += q.w * tk.y; a3 += nq.x * tk.y;
    a0 += q.z * tk.z; a1 += q.w * tk.z; a2 += nq.x * tk.z; a3 += nq.y * tk.z;
    a0 += q.w * tk.w; a1 += nq.x * tk.w; a2 += nq.y * tk.w; a3 += nq.z * tk.w;
    q = nq;
  }
  if (k < nt) {  // 1 or 3 taps left (nt is odd)
    const float4 nq = *reinterpret_cast<const float4*>(bp + k + 4);
    const float4 tk = *reinterpret_cast<const float4*>(tp + k);
    a0 += q.x * tk.x; a1 += q.y * tk.x; a2 += q.z * tk.x; a3 += q.w * tk.x;
    if (k + 1 < nt) {
      a0 += q.y * tk.y; a1 += q.z * tk.y; a2 += q.w * tk.y; a3 += nq.x * tk.y;
      a0 += q.z * tk.z; a1 += q.w * tk.z; a2 += nq.x * tk.z; a3 += nq.y * tk.z;
    }
  }
  a[0] = a0;
  a[1] = a1;
  a[2] = a2;
  a[3] = a3;
}

// a[j] = sum_k tp[k] bp[(j + k) 64], j < 4, down one lane's column of a 64-wide
// LDS tile; the window w0..w3 slides down one row per tap (reads up to row nt + 2).
__device__ __forceinline__ void col_conv4(const float* bp, const float* tp, int nt, float (&a)[4]) {
  float a0 = 0.0f, a1 = 0.0f, a2 = 0.0f, a3 = 0.0f;
  float w0 = bp[0], w1 = bp[64], w2 = bp[128];
#pragma unroll 4
  for (int k = 0; k < nt; ++k) {
    const float w3 = bp[(k + 3) * 64];
    const float tk = tp[k];
    a0 += w0 * tk; a1 += w1 * tk; a2 += w2 * tk; a3 += w3 * tk;
    w0 = w1; w1 = w2; w2 = w3;
  }
  a[0] = a0;
  a[1] = a1;
  a[2] = a2;
  a[3] = a3;
}

__global__ void __launch_bounds__(BT) blur_row_kernel(const float* __restrict__ in, float* __restrict__ out, int w,
                                                      int h, const float* __restrict__ taps, int radius,
                                                      BlurPlanes pl) {
  extern __shared__ float sh[];
  in += blockIdx.y * pl.in;
  out += blockIdx.y * (size_t)w * h;
  const int tid = threadIdx.x, nt = 2 * radius + 1;
  const int segs = (w + kRowSeg - 1) / kRowSeg;
  const int y = blockIdx.x / segs, x0 = (blockIdx.x % segs) * kRowSeg;
  const int nx = min(kRowSeg, w - x0);
  float* tp = sh;
  float* buf = sh + row_buf_off(nt);
  for (int k = tid; k < row_buf_off(nt); k += BT) tp[k] = k < nt ? taps[k] : 0.0f;
  const float* row = in + (size_t)y * w;
  for (int j = tid; j < nx + 2 * radius; j += BT) buf[j] = row[reflect_index((long long)x0 - radius + j, w)];
  __syncthreads();
  float* orow = out + (size_t)y * w + x0;
  for (int u = tid; 4 * u < nx; u += BT) {
    float a[4];
    row_conv4(buf + 4 * u, tp, nt, a);
    const int o = 4 * u;
    orow[o] = a[0];
    if (o + 1 < nx) orow[o + 1] = a[1];
    if (o + 2 < nx) orow[o + 2] = a[2];
    if (o + 3 < nx) orow[o + 3] = a[3];
  }
}

// Column epilogue: plain store, scatter mix or adjacency on the unblurred plane.
__device__ __forceinline__ float blur_epilogue(int mode, float a, const float* orig, size_t i, float acc) {
  switch (mode) {
    case BLUR_MIX: {  // (1-f) p + f blur(p)      halation.rs:164-181
      const float keep = 1.0f - a;
      return keep * orig[i] + a * acc;
    }
    case BLUR_ADJ: {  // d + beta (d - blur(d))   diffusion.rs:23-26
      const float d = orig[i];
      return d + a * (d - acc);
    }
    default: return acc;
  }
}

__global__ void __launch_bounds__(BT) blur_col_kernel(const float* __restrict__ tmp, const float* orig, float* out,
                                                      int w, int h, const float* __restrict__ taps, int radius,
                                                      int mode, float a, int th, BlurPlanes pl) {
  extern __shared__ float sh[];
  tmp += blockIdx.y * (size_t)w * h;
  if (orig) orig += blockIdx.y * pl.orig;
  out += blockIdx.y * pl.out;
  const int tid = threadIdx.x, nt = 2 * radius + 1;
  const int tiles_x = (w + 63) / 64;
  const int xs = (blockIdx.x % tiles_x) * 64, ys = (blockIdx.x / tiles_x) * th;
  const int nyo = min(th, h - ys);
  const int c = tid & 63, g = tid >> 6, x = xs + c;
  float* tp = sh;
  float* buf = sh + nt;
  for (int k = tid; k < nt; k += BT) tp[k] = taps[k];
  const int rows = nyo + 2 * radius;
  if (x < w)
    for (int j = g; j < rows; j += BT / 64) buf[j * 64 + c] = tmp[(size_t)reflect_index((long long)ys - radius + j, h) * w + x];
  __syncthreads();
  if (x >= w) return;
  // rows t0..t0+3 read buf rows t0 + j + k
  for (int t0 = 4 * g; t0 < nyo; t0 += 4 * (BT / 64)) {
    float acc[4];
    col_conv4(buf + t0 * 64 + c, tp, nt, acc);
    const size_t i = (size_t)(ys + t0) * w + x;
    out[i] = blur_epilogue(mode, a, orig, i, acc[0]);
    if (t0 + 1 < nyo) out[i + w] = blur_epilogue(mode, a, orig, i + w, acc[1]);
    if (t0 + 2 < nyo) out[i + 2 * (size_t)w] = blur_epilogue(mode, a, orig, i + 2 * (size_t)w, acc[2]);
    if (t0 + 3 < nyo) out[i + 3 * (size_t)w] = blur_epilogue(mode, a, orig, i + 3 * (size_t)w, acc[3]);
  }
}

constexpr size_t kBlurSmemMax = 120 * 1024;

// Horizontal then vertical pass over pl.count planes (tiled kernels while their
// LDS fits, else the one-pixel-per-thread ones).  tmp holds pl.count planes.
void blur_passes(const float* in, float* tmp, const float* orig, float* out, int w, int h, const float* taps_h,
                 const float* taps_v, int radius, int mode, float a, const BlurPlanes& pl, hipStream_t s) {
  const size_t n = (size_t)w * h;
  if (blur_row_smem(radius) <= kBlurSmemMax) {
    const int segs = (w + kRowSeg - 1) / kRowSeg;
    hipLaunchKernelGGL(blur_row_kernel, dim3(h * segs, pl.count), dim3(BT), blur_row_smem(radius), s, in, tmp, w, h,
                       taps_h, radius, pl);
  } else {
    hipLaunchKernelGGL(blur_h_kernel, dim3(grid1d(n), pl.count), dim3(256), 0, s, in, tmp, w, h, taps_h, radius, pl);
  }
  // the column tile height counts every plane's tiles towards filling the chip
  const int th = blur_col_height(w, h * pl.count);
  if (blur_col_smem(radius, th) <= kBlurSmemMax) {
    hipLaunchKernelGGL(blur_col_kernel, dim3(blur_col_tiles(w, h, th), pl.count), dim3(BT), blur_col_smem(radius, th),
                       s, tmp, orig, out, w, h, taps_v, radius, mode, a, th, pl);
  } else {
    hipLaunchKernelGGL(blur_v_kernel, dim3(grid1d(n), pl.count), dim3(256), 0, s, tmp, orig, out, w, h, taps_v, radius,
                       mode, a, pl);
  }
}

// Transposed blur (the adjoint of blur_passes, for the film-model gradients).
// Along one axis of length n the forward reads source reflect(p + k - r) with
// tap k; its transpose at q sums tap j times v[q - j + r] over the p inside
// [0, n), plus the sources that reflect onto q: -q (q <= r) and 2n - 2 - q
// (q >= n - 1 - r); one reflection at most, as radius <= n - 1.
__device__ __forceinline__ float adj_edges(const float* __restrict__ v, size_t stride, int n, int q,
                                           const float* __restrict__ taps, int r, float acc) {
  if (q >= 1 && q <= r) {  // sources -q reflect to q
    for (int j = 0; j <= 2 * r; ++j) {
      const int p = -q - j + r;
      if (p >= 0 && p < n) acc += taps[j] * v[(size_t)p * stride];
    }
  }
  if (q <= n - 2 && q >= n - 1 - r) {  // sources 2n-2-q reflect to q
    for (int j = 0; j <= 2 * r; ++j) {
      const int p = 2 * n - 2 - q - j + r;
      if (p >= 0 && p < n) acc += taps[j] * v[(size_t)p * stride];
    }
  }
  return acc;
}

__device__ __forceinline__ float adj_sum(const float* __restrict__ v, size_t stride, int n, int q,
                                         const float* __restrict__ taps, int r) {
  float acc = 0.0f;
  for (int j = 0; j <= 2 * r; ++j) {
    const int p = q - j + r;
    if (p >= 0 && p < n) acc += taps[j] * v[(size_t)p * stride];
  }
  return adj_edges(v, stride, n, q, taps, r, acc);
}

// The reflected sources of a transposed blur, read from an LDS window: v[(p -
// base) stride] holds source p (zero outside [0, n)); tr holds the taps
// reversed (tr[m] = tap 2r - m).  Sources -q (q <= r) and 2n - 2 - q (q >= n -
// 1 - r) reflect onto q; the window of the tile holding q always covers the
// sources they read (p in [0, r - q] and [n - r, n - 1]).
__device__ __forceinline__ float fold_edges(const float* v, int stride, int base, int n, int q, const float* tr, int r,
                                            float acc) {
  if (q >= 1 && q <= r)  // p = r - q - j >= 0
    for (int j = 0; j <= r - q; ++j) acc += tr[2 * r - j] * v[(r - q - j - base) * stride];
  if (q <= n - 2 && q >= n - 1 - r)  // p = 2n - 2 - q + r - j < n
    for (int j = n - 1 - q + r; j <= 2 * r; ++j) acc += tr[2 * r - j] * v[(2 * n - 2 - q + r - j - base) * stride];
  return acc;
}

// Column transpose, LDS-tiled like blur_col_kernel: the tile's rows and a halo
// of r above and below, zero outside the plane (those sources do not exist).
// The transpose at row q sums tap k times source q + r - k, i.e. the forward
// column loop with the taps reversed, plus the folded reflections.
__global__ void __launch_bounds__(BT) blur_colT_kernel(const float* __restrict__ in, float* __restrict__ out, int w,
                                                       int h, const float* __restrict__ taps, int radius, int th,
                                                       BlurPlanes pl) {
  extern __shared__ float sh[];
  in += blockIdx.y * pl.in;
  out += blockIdx.y * (size_t)w * h;
  const int tid = threadIdx.x, nt = 2 * radius + 1;
  const int tiles_x = (w + 63) / 64;
  const int xs = (blockIdx.x % tiles_x) * 64, ys = (blockIdx.x / tiles_x) * th;
  const int nyo = min(th, h - ys);
  const int c = tid & 63, g = tid >> 6, x = xs + c;
  float* tr = sh;
  float* buf = sh + nt;
  for (int k = tid; k < nt; k += BT) tr[k] = taps[nt - 1 - k];
  const int rows = nyo + 2 * radius;
  if (x < w)
    for (int j = g; j < rows; j += BT / 64) {
      const int yy = ys - radius + j;
      buf[j * 64 + c] = (yy >= 0 && yy < h) ? in[(size_t)yy * w + x] : 0.0f;
    }
  __syncthreads();
  if (x >= w) return;
  for (int t0 = 4 * g; t0 < nyo; t0 += 4 * (BT / 64)) {
    float a[4];
    col_conv4(buf + t0 * 64 + c, tr, nt, a);
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const int q = ys + t0 + j;
      if (t0 + j >= nyo) break;
      float acc = a[j];
      if (q <= radius || q >= h - 1 - radius) acc = fold_edges(buf + c, 64, ys - radius, h, q, tr, radius, acc);
      out[(size_t)q * w + x] = acc;
    }
  }
}

// Row transpose, LDS-tiled like blur_row_kernel (taps reversed, zero halo, the
// reflections folded from the same row buffer), then out = c0 orig + c1 acc
// (orig may alias out).
__global__ void __launch_bounds__(BT) blur_rowT_kernel(const float* __restrict__ in, const float* orig, float* out,
                                                       int w, int h, const float* __restrict__ taps, int radius,
                                                       float c0, float c1, BlurPlanes pl) {
  extern __shared__ float sh[];
  in += blockIdx.y * (size_t)w * h;
  if (orig) orig += blockIdx.y * pl.orig;
  out += blockIdx.y * pl.out;
  const int tid = threadIdx.x, nt = 2 * radius + 1;
  const int segs = (w + kRowSeg - 1) / kRowSeg;
  const int y = blockIdx.x / segs, x0 = (blockIdx.x % segs) * kRowSeg;
  const int nx = min(kRowSeg, w - x0);
  float* tr = sh;
  float* buf = sh + row_buf_off(nt);
  for (int k = tid; k < row_buf_off(nt); k += BT) tr[k] = k < nt ? taps[nt - 1 - k] : 0.0f;
  const float* row = in + (size_t)y * w;
  for (int j = tid; j < nx + 2 * radius; j += BT) {
    const int xx = x0 - radius + j;
    buf[j] = (xx >= 0 && xx < w) ? row[xx] : 0.0f;
  }
  __syncthreads();
  for (int u = tid; 4 * u < nx; u += BT) {
    float a[4];
    row_conv4(buf + 4 * u, tr, nt, a);
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const int t = 4 * u + j;
      if (t >= nx) break;
      const int q = x0 + t;
      float acc = a[j];
      if (q <= radius || q >= w - 1 - radius) acc = fold_edges(buf, 1, x0 - radius, w, q, tr, radius, acc);
      const size_t i = (size_t)y * w + q;
      out[i] = (c0 != 0.0f ? c0 * orig[i] : 0.0f) + c1 * acc;
    }
  }
}

// One-pixel-per-thread transposes, for supports whose tiles exceed the LDS budget.
__global__ void blur_vT_kernel(const float* __restrict__ in, float* __restrict__ out, int w, int h,
                               const float* __restrict__ taps, int r, BlurPlanes pl) {
  const size_t n = (size_t)w * h;
  in += blockIdx.y * pl.in;
  out += blockIdx.y * n;
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
    const int y = (int)(i / w), x = (int)(i - (size_t)y * w);
    out[i] = adj_sum(in + x, (size_t)w, h, y, taps, r);
  }
}

__global__ void blur_hT_kernel(const float* __restrict__ in, const float* orig, float* out, int w, int h,
                               const float* __restrict__ taps, int r, float c0, float c1, BlurPlanes pl) {
  const size_t n = (size_t)w * h;
  in += blockIdx.y * n;
  if (orig) orig += blockIdx.y * pl.orig;
  out += blockIdx.y * pl.out;
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
    const int y = (int)(i / w), x = (int)(i - (size_t)y * w);
    const float acc = adj_sum(in + (size_t)y * w, 1, w, x, taps, r);
    out[i] = (c0 != 0.0f ? c0 * orig[i] : 0.0f) + c1 * acc;
  }
}

}  // namespace

hipError_t film_blur(const float* in, float* tmp, const float* orig, float* out, int w, int h, const float* taps,
                     int radius, int mode, float a, hipStream_t s, const BlurPlanes& pl) {
  blur_passes(in, tmp, orig, out, w, h, taps, taps, radius, mode, a, pl, s);
  return hipGetLastError();
}

hipError_t film_blur2(const float* in, float* tmp, float* out, int w, int h, const float* taps_h, const float* taps_v,
                      int radius, hipStream_t s, const BlurPlanes& pl) {
  blur_passes(in, tmp, nullptr, out, w, h, taps_h, taps_v, radius, (int)BLUR_STORE, 0.0f, pl, s);
  return hipGetLastError();
}

hipError_t film_blur_adjoint(const float* in, float* tmp, const float* orig, float* out, int w, int h,
                             const float* taps, int radius, float c0, float c1, hipStream_t s, const BlurPlanes& pl) {
  const size_t n = (size_t)w * h;
  const int th = blur_col_height(w, h * pl.count);
  if (blur_col_smem(radius, th) <= kBlurSmemMax) {
    hipLaunchKernelGGL(blur_colT_kernel, dim3(blur_col_tiles(w, h, th), pl.count), dim3(BT), blur_col_smem(radius, th),
                       s, in, tmp, w, h, taps, radius, th, pl);
  } else {
    hipLaunchKernelGGL(blur_vT_kernel, dim3(grid1d(n), pl.count), dim3(256), 0, s, in, tmp, w, h, taps, radius, pl);
  }
  if (blur_row_smem(radius) <= kBlurSmemMax) {
    const int segs = (w + kRowSeg - 1) / kRowSeg;
    hipLaunchKernelGGL(blur_rowT_kernel, dim3(h * segs, pl.count), dim3(BT), blur_row_smem(radius), s, tmp, orig, out,
                       w, h, taps, radius, c0, c1, pl);
  } else {
    hipLaunchKernelGGL(blur_hT_kernel, dim3(grid1d(n), pl.count), dim3(256), 0, s, tmp, orig, out, w, h, taps, radius,
                       c0, c1, pl);
  }
  return hipGetLastError();
}

hipError_t film_capture_reduce(const float* absorbed, const float* bounce, size_t n, const FilmDeviceTables* T,
                               float* image, float* mask, hipStream_t s, int images, bool fp32) {
  if (fp32)
    hipLaunchKernelGGL(capture_reduce_kernel<float>, dim3(grid1d(n), images), dim3(256), 0, s, absorbed, bounce, n, T,
                       image, mask);
  else
    hipLaunchKernelGGL(capture_reduce_kernel<double>, dim3(grid1d(n), images), dim3(256), 0, s, absorbed, bounce, n, T,
                       image, mask);
  return hipGetLastError();
}

hipError_t film_dir_apply(float* image, const float* inh, size_t n, const FilmDeviceTables* T, hipStream_t s) {
  hipLaunchKernelGGL(dir_apply_kernel, dim3(grid1d(n)), dim3(256), 0, s, image, inh, n, T);
  return hipGetLastError();
}

hipError_t film_grain(float* dye, float* noise, float* tmp, int w, int h, uint64_t key0, const float* taps, int radius,
                      double* partial, float d_max, float kappa, hipStream_t s, bool philox, uint64_t seed, int layer,
                      bool absolute) {
  const size_t n = (size_t)w * h;
  if (philox)
    hipLaunchKernelGGL(grain_philox_kernel, dim3(grid1d(n)), dim3(256), 0, s, noise, n, seed, (uint32_t)layer);
  else
    hipLaunchKernelGGL(grain_noise_kernel, dim3(grid1d(n)), dim3(256), 0, s, noise, w, h, key0);
  blur_passes(noise, tmp, noise, noise, w, h, taps, taps, radius, (int)BLUR_STORE, 0.0f, BlurPlanes{}, s);
  const int parts = kFilmReduceBlocks;
  hipLaunchKernelGGL(sumsq_partial_kernel, dim3(parts), dim3(256), 0, s, noise, n, partial);
  float* norm = reinterpret_cast<float*>(partial + parts);
  hipLaunchKernelGGL(grain_norm_kernel, dim3(1), dim3(64), 0, s, partial, parts, n, norm);
  hipLaunchKernelGGL(grain_apply_kernel, dim3(grid1d(n)), dim3(256), 0, s, dye, noise, n, norm, d_max, kappa,
                     absolute ? 1 : 0);
  return hipGetLastError();
}

namespace {
// The mixed form's scan in FP32: per bin one density sum over the emulsions'
// dye and mask terms (pre-scaled by log2 10), one v_exp_f32 and three
// products with scanner x CMF x trapezoid weight; the XYZ -> AP1 step, the
// Dmin normalisation and the inversion as scan_kernel.
template <int NB>
__global__ void __launch_bounds__(256) scan_f32_kernel(const float* __restrict__ image, const float* __restrict__ mask,
                                                       size_t n, const FilmDeviceTables* __restrict__ T,
                                                       float* __restrict__ rgb) {
  const int E = T->n_emulsions;
  const size_t b = blockIdx.y;
  image += b * E * n;
  mask += b * E * n;
  rgb += b * 3 * n;
  for (size_t p = (size_t)blockIdx.x * blockDim.x + threadIdx.x; p < n; p += (size_t)gridDim.x * blockDim.x) {
    float di[kFilmMaxEmulsions], dm[kFilmMaxEmulsions];
#pragma unroll
    for (int e = 0; e < kFilmMaxEmulsions; ++e) {
      di[e] = e < E ? image[(size_t)e * n + p] : 0.0f;
      dm[e] = e < E ? mask[(size_t)e * n + p] : 0.0f;
    }
    float x0 = 0.0f, x1 = 0.0f, x2 = 0.0f;
    for (int i = 0; i < NB; ++i) {
      float d = 0.0f;
#pragma unroll
      for (int e = 0; e < kFilmMaxEmulsions; ++e) {
        if (e >= E) break;
        d = fmaf(di[e], T->scan_eps2[e][i], d);
        d = fmaf(dm[e], T->scan_mask2[e][i], d);
      }
      const float t = __builtin_amdgcn_exp2f(-d);
      x0 = fmaf(t, T->scan_w[0][i], x0);
      x1 = fmaf(t, T->scan_w[1][i], x1);
      x2 = fmaf(t, T->scan_w[2][i], x2);
    }
    const float ds = (float)T->dmin_scale;
    float out[3];
    for (int c = 0; c < 3; ++c)
      out[c] = (T->scan_xyz2ap1[c][0] * x0 + T->scan_xyz2ap1[c][1] * x1 + T->scan_xyz2ap1[c][2] * x2) * ds;
    if (T->invert) {
      const float eps = 1e-6f;
      float raw[3];
      for (int c = 0; c < 3; ++c) {
        float tt = out[c] / T->inv_dmin[c];
        tt = tt < eps ? eps : (tt > 1.0f ? 1.0f : tt);
        const float inv = fmaxf(1.0f / tt - 1.0f, 0.0f);
        const float iv = fmaxf(inv, 0.0f);
        const float sh = fmaxf(T->inv_shoulder[c], 1e-6f);
        raw[c] = fmaxf(T->inv_gain[c] * (iv / (1.0f + iv / sh)), 0.0f);
      }
      const float y = fmaxf(0.2722287168f * raw[0] + 0.6740817658f * raw[1] + 0.0536895174f * raw[2], 0.0f);
      float k;
      if (y <= 0.0f) k = 0.0f;
      else if (y >= 0.02f) k = 0.55f;
      else k = 0.55f * (y / 0.02f);
      for (int c = 0; c < 3; ++c) out[c] = fmaxf(y + k * (raw[c] - y), 0.0f);
    }
    rgb[p * 3] = out[0];
    rgb[p * 3 + 1] = out[1];
    rgb[p * 3 + 2] = out[2];
  }
}
}  // namespace

hipError_t film_scan(const float* image, const float* mask, size_t n, const FilmDeviceTables* T, float* rgb,
                     hipStream_t s, int images, bool fp32, int bins) {
  if (bins != 16 && bins != 81) return hipErrorInvalidValue;
  const dim3 g(grid1d(n), images), blk(256);
  if (bins == 16) {
    if (fp32) hipLaunchKernelGGL((scan_f32_kernel<16>), g, blk, 0, s, image, mask, n, T, rgb);
    else hipLaunchKernelGGL((scan_kernel<double, 16>), g, blk, 0, s, image, mask, n, T, rgb);
  } else {
    if (fp32) hipLaunchKernelGGL((scan_f32_kernel<81>), g, blk, 0, s, image, mask, n, T, rgb);
    else hipLaunchKernelGGL((scan_kernel<double, 81>), g, blk, 0, s, image, mask, n, T, rgb);
  }
  return hipGetLastError();
}

}  // namespace imops
