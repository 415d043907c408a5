// BM3D (plain FP32) for gfx950: block matching, collaborative filtering with
// fixed-point aggregation, and the luma detail blend.
//
// Behaviour follows pichromatic/src/bm3d.rs bit for bit:
//  * matching (bm3d.rs:281-391): every candidate's 3-channel SSD is summed in
//    the reference's order (row-major, one running f32 sum per channel) with
//    the same per-row early exits and per-channel stop thresholds, accepted
//    candidates are capped at max_candidates in scan order, and the group is
//    the first prev_pow2(min(16, n)) entries of the (dist, scan order) sort;
//  * filtering (bm3d.rs:128-238, 490-614): 8x8 DCT as the reference's
//    row-dot-products, WHT across the group with 1/sqrt(n), hard threshold /
//    Wiener, inverse; every output sample is one thread's sequential sum in
//    the reference's order, and the Wiener weight's running sum is taken in
//    stack order;
//  * aggregation (bm3d.rs:463-488): each contribution is converted to i64
//    fixed point (x1e5, truncated) and summed with integer adds, so any
//    summation order gives the reference's exact totals.  A workgroup owns a
//    tile of reference patches and accumulates its footprint in LDS before
//    flushing non-zero sums to the global i64 planes with atomics.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "bm3d.hpp"

namespace imops {
namespace {

__device__ __forceinline__ long long to_fixed(float v) {
  const float s = v * 100000.0f;
  if (isnan(s)) return 0;
  if (s >= 9.2233720368547758e18f) return INT64_MAX;
  if (s <= -9.2233720368547758e18f) return INT64_MIN;
  return (long long)s;  // truncation toward zero
}

// ------------------------------------------------------------------ colour
__global__ void rgb_to_ycbcr_kernel(const float* __restrict__ rgb, size_t n, float* __restrict__ y,
                                    float* __restrict__ cb, float* __restrict__ cr) {
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
    const float r = rgb[i * 3] * 255.0f, g = rgb[i * 3 + 1] * 255.0f, b = rgb[i * 3 + 2] * 255.0f;
    y[i] = 0.299f * r + 0.587f * g + 0.114f * b;
    cb[i] = -0.168736f * r - 0.331264f * g + 0.5f * b + 128.0f;
    cr[i] = 0.5f * r - 0.418688f * g - 0.081312f * b + 128.0f;
  }
}

// ---------------------------------------------------------------- matching
// One workgroup per reference patch; the clamped search window of the three
// guide planes is staged in LDS.
constexpr int MT = 256;               // matching threads
constexpr int kMaxWin = kBm3dMaxWindow; // largest supported search window
constexpr int kMaxSpan = kMaxWin + 7;  // staged rows/cols

struct MatchShared {
  float win[3][kMaxSpan * kMaxSpan];
  float ref[3][64];
  float dist[kMaxWin * kMaxWin];      // +inf when rejected
  int scan[MT + 1];
  unsigned long long best[MT / 64];
};

__global__ void __launch_bounds__(MT) bm3d_match_kernel(Bm3dMatchArgs a) {
  extern __shared__ unsigned char smem_raw[];
  MatchShared& S = *reinterpret_cast<MatchShared*>(smem_raw);
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int ref_id = blockIdx.x;
  const int rx = (ref_id % a.refs_x) * a.stride, ry = (ref_id / a.refs_x) * a.stride;
  const int half = a.window / 2;
  const int wmax = a.width - 8, hmax = a.height - 8;
  const int sx0 = rx >= half ? rx - half : 0, sx1 = min(rx + half, wmax);
  const int sy0 = ry >= half ? ry - half : 0, sy1 = min(ry + half, hmax);
  const int nx = sx1 - sx0 + 1, ny = sy1 - sy0 + 1;
  const int span_x = nx + 7, span_y = ny + 7;

  for (int ch = 0; ch < 3; ++ch) {
    const float* plane = a.guide[ch];
    for (int k = tid; k < span_x * span_y; k += MT) {
      const int yy = k / span_x, xx = k - yy * span_x;
      S.win[ch][yy * span_x + xx] = plane[(size_t)(sy0 + yy) * a.width + sx0 + xx];
    }
    if (tid < 64) S.ref[ch][tid] = plane[(size_t)(ry + (tid >> 3)) * a.width + rx + (tid & 7)];
  }
  __syncthreads();

  const float thr = a.threshold * 64.0f;
  const int ncand = nx * ny;
  // Candidate SSDs in scan order; the reference position is excluded here and
  // prepended below with distance 0.
  for (int base = 0; base < ncand; base += MT) {
    const int c = base + tid;
    const bool valid = c < ncand;
    const int cy = valid ? c / nx : 0, cx = valid ? c - cy * nx : 0;
    const bool is_ref = valid && (sx0 + cx == rx) && (sy0 + cy == ry);
    bool live = valid && !is_ref;
    float d[3] = {0.0f, 0.0f, 0.0f};
    float stop = thr;
    for (int ch = 0; ch < 3; ++ch) {
      // exit when no lane of this wave still needs this channel
      if (!__any(live)) break;
      const float* win = &S.win[ch][cy * span_x + cx];
      const float* rp = S.ref[ch];
      float acc = 0.0f;
      bool run = live;
      for (int dy = 0; dy < 8; ++dy) {
        if (!__any(run)) break;
        if (run) {
          const float* row = win + dy * span_x;
          for (int dx = 0; dx < 8; ++dx) {
            const float diff = row[dx] - rp[dy * 8 + dx];
            acc += diff * diff;
          }
          if (acc > stop) run = false;  // compute_ssd_flat returns the partial sum here
        }
      }
      d[ch] = acc;
      if (live) {
        if (ch == 0) {
          if (d[0] > thr) live = false;
          else stop = thr - d[0];
        } else if (ch == 1) {
          if (d[0] + d[1] > thr) live = false;
          else stop = thr - (d[0] + d[1]);
        }
      }
    }
    float dist = __builtin_inff();
    if (live) {
      const float total = d[0] + d[1] + d[2];
      if (total < thr) dist = total / 64.0f;
    }
    if (valid) S.dist[c] = dist;
  }
  __syncthreads();

  // Cap: only the first max_candidates-1 accepted candidates (scan order)
  // join the reference patch.  Block-wide exclusive scan of accept counts over
  // contiguous chunks of the scan order.
  const int chunk = (ncand + MT - 1) / MT;
  const int c0 = tid * chunk, c1 = min(c0 + chunk, ncand);
  int cnt = 0;
  for (int c = c0; c < c1; ++c) cnt += S.dist[c] != __builtin_inff();
  S.scan[tid + 1] = cnt;
  __syncthreads();
  if (tid == 0) {
    S.scan[0] = 0;
    for (int t = 1; t <= MT; ++t) S.scan[t] += S.scan[t - 1];
  }
  __syncthreads();
  const int total_acc = S.scan[MT];
  int rank = S.scan[tid];
  for (int c = c0; c < c1; ++c)
    if (S.dist[c] != __builtin_inff()) {
      if (rank >= a.max_candidates - 1) S.dist[c] = __builtin_inff();
      ++rank;
    }
  __syncthreads();
  const int n_list = 1 + min(total_acc, a.max_candidates - 1);
  const int limit = min(a.max_group, n_list);
  int g = 1;
  while (g * 2 <= limit) g *= 2;

  // Selection: the reference patch first, then g-1 rounds of block argmin on
  // (dist bits, scan index); dist >= 0 so its bit pattern orders like the value.
  uint32_t* out = a.groups + (size_t)ref_id * 17;
  if (tid == 0) {
    out[0] = (uint32_t)g;
    out[1] = (uint32_t)rx | ((uint32_t)ry << 16);
  }
  for (int k = 1; k < g; ++k) {
    unsigned long long best = ~0ull;
    for (int c = tid; c < ncand; c += MT) {
      const float dv = S.dist[c];
      if (dv != __builtin_inff()) {
        const unsigned long long key = ((unsigned long long)__float_as_uint(dv) << 32) | (unsigned)c;
        best = key < best ? key : best;
      }
    }
    for (int off = 32; off > 0; off >>= 1) {
      const unsigned long long o = __shfl_xor(best, off, 64);
      best = o < best ? o : best;
    }
    if (lane == 0) S.best[wave] = best;
    __syncthreads();
    unsigned long long b = S.best[0];
    for (int w = 1; w < MT / 64; ++w) b = S.best[w] < b ? S.best[w] : b;
    const int c = (int)(b & 0xffffffffu);
    if (tid == 0 && b != ~0ull) {
      const int cy = c / nx, cx = c - cy * nx;
      out[1 + k] = (uint32_t)(sx0 + cx) | ((uint32_t)(sy0 + cy) << 16);
      S.dist[c] = __builtin_inff();
    }
    __syncthreads();
  }
}

// --------------------------------------------------------------- filtering
constexpr int FT = 256;  // 4 waves; each wave filters one (reference patch, channel) at a time

// 8x8 2D DCT of the value each lane holds (lane = position r*8+c) in the
// reference's order; `pad` is a per-wave 64-float LDS scratch.
__device__ __forceinline__ float dct2_lane(float v, float* pad, const float* C, int lane) {
  const int r = lane >> 3, k = lane & 7;
  pad[lane] = v;
  __builtin_amdgcn_wave_barrier();
  // y[r][k] = sum_m x[r][m] * C[k][m]
  const float* xr = pad + r * 8;
  const float* ck = C + k * 8;
  float y = xr[0] * ck[0] + xr[1] * ck[1] + xr[2] * ck[2] + xr[3] * ck[3] + xr[4] * ck[4] + xr[5] * ck[5] +
            xr[6] * ck[6] + xr[7] * ck[7];
  __builtin_amdgcn_wave_barrier();
  pad[lane] = y;
  __builtin_amdgcn_wave_barrier();
  // out[j][k] = sum_n y[n][k] * C[j][n]   (lane = j*8+k)
  const int j = r;
  const float* cj = C + j * 8;
  float o = pad[0 * 8 + k] * cj[0] + pad[1 * 8 + k] * cj[1] + pad[2 * 8 + k] * cj[2] + pad[3 * 8 + k] * cj[3] +
            pad[4 * 8 + k] * cj[4] + pad[5 * 8 + k] * cj[5] + pad[6 * 8 + k] * cj[6] + pad[7 * 8 + k] * cj[7];
  __builtin_amdgcn_wave_barrier();
  return o;
}

// Inverse: u[k][n] = sum_j x[j][k] * D[n][j]; out[n][m] = sum_k u[k][n] * D[m][k].
__device__ __forceinline__ float idct2_lane(float v, float* pad, const float* D, int lane) {
  pad[lane] = v;
  __builtin_amdgcn_wave_barrier();
  {
    const int k = lane >> 3, n = lane & 7;  // lane computes u[k][n]
    const float* dn = D + n * 8;
    float u = pad[0 * 8 + k] * dn[0] + pad[1 * 8 + k] * dn[1] + pad[2 * 8 + k] * dn[2] + pad[3 * 8 + k] * dn[3] +
              pad[4 * 8 + k] * dn[4] + pad[5 * 8 + k] * dn[5] + pad[6 * 8 + k] * dn[6] + pad[7 * 8 + k] * dn[7];
    __builtin_amdgcn_wave_barrier();
    pad[lane] = u;
    __builtin_amdgcn_wave_barrier();
  }
  const int n = lane >> 3, m = lane & 7;  // lane computes out[n][m]
  const float* dm = D + m * 8;
  float o = pad[0 * 8 + n] * dm[0] + pad[1 * 8 + n] * dm[1] + pad[2 * 8 + n] * dm[2] + pad[3 * 8 + n] * dm[3] +
            pad[4 * 8 + n] * dm[4] + pad[5 * 8 + n] * dm[5] + pad[6 * 8 + n] * dm[6] + pad[7 * 8 + n] * dm[7];
  __builtin_amdgcn_wave_barrier();
  return o;
}

// In-place WHT across the group (registers v[0..g-1]) with the reference's
// butterfly order and 1/sqrt(g) scale.
__device__ __forceinline__ void wht_regs(float* v, int g, float scale) {
  for (int h = 1; h < g; h *= 2)
    for (int i = 0; i < g; i += 2 * h)
      for (int j = i; j < i + h; ++j) {
        const float x = v[j], y = v[j + h];
        v[j] = x + y;
        v[j + h] = x - y;
      }
  for (int i = 0; i < g; ++i) v[i] *= scale;
}

struct FilterShared {
  float pad[FT / 64][64];
  float ssum[FT / 64][16 * 64];  // Wiener coefficients in stack order
  float tab[3][64];              // dct, idct, kaiser
};

template <bool kStep1>
__global__ void __launch_bounds__(FT) bm3d_filter_kernel(Bm3dFilterArgs a) {
  extern __shared__ unsigned char smem_raw[];
  FilterShared& S = *reinterpret_cast<FilterShared*>(smem_raw);
  unsigned long long* acc = reinterpret_cast<unsigned long long*>(smem_raw + sizeof(FilterShared));
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  // tile of reference patches owned by this workgroup
  const int tiles_x = (a.refs_x + a.tile - 1) / a.tile;
  const int tx = blockIdx.x % tiles_x, ty = blockIdx.x / tiles_x;
  const int r0x = tx * a.tile, r0y = ty * a.tile;
  const int r1x = min(r0x + a.tile, a.refs_x), r1y = min(r0y + a.tile, a.refs_y);
  const int nrx = r1x - r0x, nref = nrx * (r1y - r0y);
  // footprint: every patch a group of this tile can contain
  const int half = a.window / 2;
  const int fx0 = max(r0x * a.stride - half, 0), fy0 = max(r0y * a.stride - half, 0);
  const int fx1 = min((r1x - 1) * a.stride + half, a.width - 8) + 8;   // exclusive
  const int fy1 = min((r1y - 1) * a.stride + half, a.height - 8) + 8;
  const int fw = fx1 - fx0, fh = fy1 - fy0;

  if (tid < 64) {
    S.tab[0][tid] = a.dct[tid];
    S.tab[1][tid] = a.idct[tid];
    S.tab[2][tid] = a.kaiser[tid];
  }
  float* pad = S.pad[wave];

  for (int ch = 0; ch < 3; ++ch) {
    for (int k = tid; k < 2 * fw * fh; k += FT) acc[k] = 0ull;
    __syncthreads();
    const float ch_sigma = ch == 0 ? a.sigma : a.sigma * a.chroma_scale;
    const float* gplane = a.guide[ch];
    const float* nplane = a.noisy[ch];
    for (int r = wave; r < nref; r += FT / 64) {
      const int ref_id = (r0y + r / nrx) * a.refs_x + r0x + r % nrx;
      const uint32_t* grp = a.groups + (size_t)ref_id * 17;
      const int g = (int)grp[0];
      const int py = lane >> 3, px = lane & 7;
      float gv[16], nv[16];
      int gx[16], gyy[16];
      for (int k = 0; k < 16; ++k) {
        if (k < g) {
          const uint32_t p = grp[1 + k];
          gx[k] = (int)(p & 0xffffu);
          gyy[k] = (int)(p >> 16);
          const size_t off = (size_t)(gyy[k] + py) * a.width + gx[k] + px;
          gv[k] = gplane[off];
          nv[k] = kStep1 ? 0.0f : nplane[off];
        }
      }
      // forward 3D transform (bm3d.rs:490-506)
      for (int k = 0; k < g; ++k) gv[k] = dct2_lane(gv[k], pad, S.tab[0], lane);
      const float wscale = 1.0f / sqrtf((float)g);
      wht_regs(gv, g, wscale);
      if (!kStep1) {
        for (int k = 0; k < g; ++k) nv[k] = dct2_lane(nv[k], pad, S.tab[0], lane);
        wht_regs(nv, g, wscale);
      }
      float weight;
      if (kStep1) {
        // hard_threshold (bm3d.rs:240-255): element 0 of the stack is kept
        const float th = a.lambda * ch_sigma;
        int nz = 0;
        for (int k = 0; k < g; ++k) {
          if (k == 0 && lane == 0) { ++nz; continue; }
          if (fabsf(gv[k]) < th) gv[k] = 0.0f;
          else ++nz;
        }
        for (int off = 32; off > 0; off >>= 1) nz += __shfl_xor(nz, off, 64);
        weight = nz > 0 ? 1.0f / (float)nz : 1.0f;
        for (int k = 0; k < g; ++k) nv[k] = gv[k];
      } else {
        // wiener_filter (bm3d.rs:257-272): coefficient per element, running sum in stack order
        const float s2 = ch_sigma * ch_sigma;
        float* ss = S.ssum[wave];
        for (int k = 0; k < g; ++k) {
          if (k == 0 && lane == 0) { ss[0] = 0.0f; continue; }
          const float e = gv[k] * gv[k];
          const float coef = e / (e + s2 + 1e-5f);
          nv[k] *= coef;
          ss[k * 64 + lane] = coef * coef;
        }
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        float sum = 0.0f;
        if (lane == 0) {
          sum = 1.0f;
          const int n = g * 64;
          for (int i = 1; i < n; ++i) sum += ss[i];
        }
        sum = __shfl(sum, 0, 64);
        weight = sum > 0.0f ? 1.0f / sum : 1.0f;
        __builtin_amdgcn_wave_barrier();
      }
      // inverse (bm3d.rs:508-524)
      wht_regs(nv, g, wscale);
      const float kw = S.tab[2][lane] * weight;
      const long long dfix = to_fixed(kw);
      for (int k = 0; k < g; ++k) {
        const float v = idct2_lane(nv[k], pad, S.tab[1], lane);
        const int ax = gx[k] + px - fx0, ay = gyy[k] + py - fy0;
        const int fi = ay * fw + ax;
        atomicAdd(&acc[2 * fi], (unsigned long long)to_fixed(v * kw));
        atomicAdd(&acc[2 * fi + 1], (unsigned long long)dfix);
      }
    }
    __syncthreads();
    // flush the footprint to the global i64 planes
    unsigned long long* gnum = a.num + (size_t)ch * a.width * a.height;
    unsigned long long* gden = a.den + (size_t)ch * a.width * a.height;
    for (int k = tid; k < fw * fh; k += FT) {
      const unsigned long long n = acc[2 * k], d = acc[2 * k + 1];
      if (n | d) {
        const size_t gi = (size_t)(fy0 + k / fw) * a.width + fx0 + k % fw;
        if (n) atomicAdd(&gnum[gi], n);
        if (d) atomicAdd(&gden[gi], d);
      }
    }
    __syncthreads();
  }
}

// num/den -> estimate (bm3d.rs:225-236); optionally clears the accumulators.
__global__ void bm3d_finalize_kernel(const unsigned long long* __restrict__ num, const unsigned long long* __restrict__ den,
                                     const float* __restrict__ noisy0, const float* __restrict__ noisy1,
                                     const float* __restrict__ noisy2, float* __restrict__ out0,
                                     float* __restrict__ out1, float* __restrict__ out2, size_t n) {
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < 3 * n; i += (size_t)gridDim.x * blockDim.x) {
    const int ch = (int)(i / n);
    const size_t p = i - (size_t)ch * n;
    const float nn = (float)(long long)num[i] / 100000.0f;
    const float dd = (float)(long long)den[i] / 100000.0f;
    const float* noisy = ch == 0 ? noisy0 : (ch == 1 ? noisy1 : noisy2);
    float* out = ch == 0 ? out0 : (ch == 1 ? out1 : out2);
    out[p] = dd > 1e-6f ? nn / dd : noisy[p];
  }
}

__global__ void fill_u64_kernel(unsigned long long* p, size_t n) {
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) p[i] = 0ull;
}

// Separable in-bounds-normalised Gaussian of gaussian_blur_1ch (bm3d.rs:648-699).
template <bool kVertical>
__global__ void blur1_kernel(const float* __restrict__ in, float* __restrict__ out, int w, int h,
                             Bm3dBlurTaps taps) {
  const size_t n = (size_t)w * h;
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
    const int y = (int)(i / w), x = (int)(i - (size_t)y * w);
    float val = 0.0f, ws = 0.0f;
    for (int k = 0; k < taps.len; ++k) {
      const int q = (kVertical ? y : x) - taps.radius + k;
      if (q >= 0 && q < (kVertical ? h : w)) {
        val += in[kVertical ? (size_t)q * w + x : (size_t)y * w + q] * taps.k[k];
        ws += taps.k[k];
      }
    }
    out[i] = val / ws;
  }
}

// Detail blend of the denoised luma + YCbCr -> RGB /255 (bm3d.rs:51-60, 79-92, 629-635).
__global__ void bm3d_merge_kernel(const float* __restrict__ yd, const float* __restrict__ cb,
                                  const float* __restrict__ cr, const float* __restrict__ yorig,
                                  const float* __restrict__ yblur, float detail, float* __restrict__ rgb, size_t n) {
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
    const float hf = yorig[i] - yblur[i];
    float yv = yd[i] + detail * hf;
    yv = yv < 0.0f ? 0.0f : (yv > 255.0f ? 255.0f : yv);
    const float cbv = cb[i] - 128.0f, crv = cr[i] - 128.0f;
    const float r = yv + 1.402f * crv;
    const float g = yv - 0.344136f * cbv - 0.714136f * crv;
    const float b = yv + 1.772f * cbv;
    rgb[i * 3] = fmaxf(r, 0.0f) / 255.0f;
    rgb[i * 3 + 1] = fmaxf(g, 0.0f) / 255.0f;
    rgb[i * 3 + 2] = fmaxf(b, 0.0f) / 255.0f;
  }
}

int grid1d(size_t n) {
  size_t g = (n + 255) / 256;
  return (int)(g > 65536 ? 65536 : (g ? g : 1));
}

}  // namespace

size_t bm3d_match_smem() { return sizeof(MatchShared); }
size_t bm3d_filter_smem(int footprint_pixels) {
  return sizeof(FilterShared) + (size_t)footprint_pixels * 2 * sizeof(unsigned long long);
}

hipError_t bm3d_rgb_to_ycbcr(const float* rgb, size_t n, float* y, float* cb, float* cr, hipStream_t s) {
  hipLaunchKernelGGL(rgb_to_ycbcr_kernel, dim3(grid1d(n)), dim3(256), 0, s, rgb, n, y, cb, cr);
  return hipGetLastError();
}

hipError_t bm3d_match_launch(const Bm3dMatchArgs& a, int nrefs, hipStream_t s) {
  hipLaunchKernelGGL(bm3d_match_kernel, dim3(nrefs), dim3(MT), bm3d_match_smem(), s, a);
  return hipGetLastError();
}

hipError_t bm3d_filter_launch(const Bm3dFilterArgs& a, bool step1, int nblocks, size_t smem, hipStream_t s) {
  if (step1) hipLaunchKernelGGL(bm3d_filter_kernel<true>, dim3(nblocks), dim3(FT), smem, s, a);
  else hipLaunchKernelGGL(bm3d_filter_kernel<false>, dim3(nblocks), dim3(FT), smem, s, a);
  return hipGetLastError();
}

hipError_t bm3d_finalize_launch(const unsigned long long* num, const unsigned long long* den, const float* const noisy[3],
                                float* const out[3], size_t n, hipStream_t s) {
  hipLaunchKernelGGL(bm3d_finalize_kernel, dim3(grid1d(3 * n)), dim3(256), 0, s, num, den, noisy[0], noisy[1],
                     noisy[2], out[0], out[1], out[2], n);
  return hipGetLastError();
}

hipError_t fill_zero_u64(unsigned long long* p, size_t n, hipStream_t s) {
  hipLaunchKernelGGL(fill_u64_kernel, dim3(grid1d(n)), dim3(256), 0, s, p, n);
  return hipGetLastError();
}

hipError_t bm3d_blur_launch(const float* in, float* tmp, float* out, int w, int h, const Bm3dBlurTaps& taps,
                            hipStream_t s) {
  const size_t n = (size_t)w * h;
  hipLaunchKernelGGL(blur1_kernel<false>, dim3(grid1d(n)), dim3(256), 0, s, in, tmp, w, h, taps);
  hipLaunchKernelGGL(blur1_kernel<true>, dim3(grid1d(n)), dim3(256), 0, s, tmp, out, w, h, taps);
  return hipGetLastError();
}

hipError_t bm3d_merge_launch(const float* yd, const float* cb, const float* cr, const float* yorig, const float* yblur,
                             float detail, float* rgb, size_t n, hipStream_t s) {
  hipLaunchKernelGGL(bm3d_merge_kernel, dim3(grid1d(n)), dim3(256), 0, s, yd, cb, cr, yorig, yblur, detail, rgb, n);
  return hipGetLastError();
}

}  // namespace imops
