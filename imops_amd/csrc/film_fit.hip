// Differentiable film model on gfx950: the forward pieces that differ from the
// film stage (spectral exposure as a per-stock 3 x E contraction) and the
// reverse-mode (adjoint) kernels of
//
//   A_e = sum_k w_k C[k][e]                       (w = relu(Minv rgb); C from the sensitivities)
//   S_e = 0.82 A_e + 0.18 G_local(A_e)            (local scatter, reflect-101 Gaussian)
//   B   = G_halation(S_deepest)                   (halation bounce, width scaled by the radius parameter)
//   P_e = S_e + gain_e B                          (gain = strength * r_e * bleed_e)
//   f_e = LUT_e(P_e),  D_e = d_max f^(1/gamma),  M_e = 0.4 d_max (1 - D/d_max)
//   out = scale * XYZ->AP1 * CMF . (scanner 10^-(sum D eps + M eps_mask))
//
// (pichromatic/src/film: exposure/mod.rs:45-137, halation.rs:24-181,
// capture.rs:37-106, reduction.rs:16-68, densitometry.rs:13-73) with grain,
// DIR inhibition and adjacency off.  Parameter gradients are reduced in f64.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "film.hpp"

namespace imops {
namespace {

int grid_of(size_t n) {
  size_t g = (n + 255) / 256;
  return (int)(g > 16384 ? 16384 : (g ? g : 1));
}

__global__ void fit_expose_kernel(const float* __restrict__ rgb, size_t n, FitExposeConsts k, float* __restrict__ planes) {
  for (size_t p = (size_t)blockIdx.x * blockDim.x + threadIdx.x; p < n; p += (size_t)gridDim.x * blockDim.x) {
    const double r = rgb[p * 3], g = rgb[p * 3 + 1], b = rgb[p * 3 + 2];
    double w[3];
    for (int i = 0; i < 3; ++i) {
      const double v = k.minv[i][0] * r + k.minv[i][1] * g + k.minv[i][2] * b;
      w[i] = v < 0.0 ? 0.0 : v;
    }
    for (int e = 0; e < k.n_emulsions; ++e)
      planes[(size_t)e * n + p] = (float)(w[0] * k.c[0][e] + w[1] * k.c[1][e] + w[2] * k.c[2][e]);
  }
}

// sum over the taps whose reflect-101 source index is q (one reflection at most:
// radius <= n - 1), i.e. the transpose of the forward blur along one axis
__device__ __forceinline__ float adj_sum(const float* __restrict__ v, size_t stride, int n, int q,
                                         const float* __restrict__ taps, int r) {
  float acc = 0.0f;
  for (int j = 0; j <= 2 * r; ++j) {
    const int p = q - j + r;
    if (p >= 0 && p < n) acc += taps[j] * v[(size_t)p * stride];
  }
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

__global__ void blur_vT_kernel(const float* __restrict__ in, float* __restrict__ out, int w, int h,
                               const float* __restrict__ taps, int r) {
  const size_t n = (size_t)w * h;
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
    const int y = (int)(i / w), x = (int)(i - (size_t)y * w);
    out[i] = adj_sum(in + x, (size_t)w, h, y, taps, r);
  }
}

// out = c0 * orig + c1 * H^T(in)   (orig may alias out)
__global__ void blur_hT_kernel(const float* __restrict__ in, const float* orig, float* out, int w, int h,
                               const float* __restrict__ taps, int r, float c0, float c1) {
  const size_t n = (size_t)w * h;
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
    const int y = (int)(i / w), x = (int)(i - (size_t)y * w);
    const float acc = adj_sum(in + (size_t)y * w, 1, w, x, taps, r);
    out[i] = (c0 != 0.0f ? c0 * orig[i] : 0.0f) + c1 * acc;
  }
}

// Block-wide f64 sums of `cnt` per-thread values, added atomically to out[0..cnt).
template <int kMax>
__device__ void block_accumulate(double (&v)[kMax], int cnt, double* out) {
  __shared__ double part[4][kMax];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  for (int k = 0; k < cnt; ++k) {
    double x = v[k];
    for (int off = 32; off > 0; off >>= 1) x += __shfl_xor(x, off, 64);
    if (lane == 0) part[wave][k] = x;
  }
  __syncthreads();
  if (threadIdx.x < cnt) {
    double s = 0.0;
    for (int wv = 0; wv < (int)(blockDim.x >> 6); ++wv) s += part[wv][threadIdx.x];
    atomicAdd(out + threadIdx.x, s);
  }
}

// Reverse of capture + reduction + scan for one image.  In: S planes, bounce B,
// dL/dout.  Out: dL/dS planes (into sbar), dL/dB, and f64 sums of dL/dd_max,
// dL/dgamma, dL/dgain (3 x E).
__global__ void __launch_bounds__(256) fit_back_scan_kernel(const float* __restrict__ S, const float* __restrict__ B,
                                                            const FilmDeviceTables* __restrict__ T,
                                                            const float* __restrict__ gout, size_t n,
                                                            float* __restrict__ sbar, float* __restrict__ bbar,
                                                            double* __restrict__ acc) {
  const int E = T->n_emulsions;
  double part[3 * kFilmMaxEmulsions];
  for (int k = 0; k < 3 * kFilmMaxEmulsions; ++k) part[k] = 0.0;
  const double ln10 = 2.302585092994045684;
  for (size_t p = (size_t)blockIdx.x * blockDim.x + threadIdx.x; p < n; p += (size_t)gridDim.x * blockDim.x) {
    float D[kFilmMaxEmulsions], M[kFilmMaxEmulsions], dDdf[kFilmMaxEmulsions], P[kFilmMaxEmulsions];
    float eff[kFilmMaxEmulsions];
    double slope[kFilmMaxEmulsions];
    const float bv = B ? B[p] : 0.0f;
    for (int e = 0; e < E; ++e) {
      const FilmEmulsionTables& Em = T->em[e];
      const float phi = S[(size_t)e * n + p] + (B && Em.gain > 0.0f ? Em.gain * bv : 0.0f);
      P[e] = phi;
      const double x = (double)(phi * Em.eta);
      double f, s = 0.0;
      if (!(x > 0.0)) {
        f = Em.lut_frac[0];
      } else {
        const double lp = log10(x);
        if (lp <= Em.lut_log[0]) {
          f = Em.lut_frac[0];
        } else if (lp >= Em.lut_log[63]) {
          f = Em.lut_frac[63];
        } else {
          int lo = 0, hi = 63;
          while (hi - lo > 1) {
            const int mid = (lo + hi) / 2;
            if (Em.lut_log[mid] <= lp) lo = mid;
            else hi = mid;
          }
          const double dl = Em.lut_log[hi] - Em.lut_log[lo];
          const double t = (lp - Em.lut_log[lo]) / dl;
          f = Em.lut_frac[lo] * (1.0 - t) + Em.lut_frac[hi] * t;
          s = (Em.lut_frac[hi] - Em.lut_frac[lo]) / dl / (x * ln10) * (double)Em.eta;  // df/dP
        }
      }
      slope[e] = s;
      const float fr = (float)f;
      const bool inside = fr > 0.0f && fr < 1.0f;
      const float c = fr < 0.0f ? 0.0f : (fr > 1.0f ? 1.0f : fr);
      const float ef = Em.reversal ? 1.0f - c : c;
      eff[e] = ef;
      D[e] = Em.d_max * powf(ef, Em.inv_gamma);
      const float deff = ef > 0.0f ? Em.d_max * Em.inv_gamma * powf(ef, Em.inv_gamma - 1.0f) : 0.0f;
      dDdf[e] = inside ? (Em.reversal ? -deff : deff) : 0.0f;
      float m = 0.0f;
      if (Em.has_mask) {
        const float q = D[e] / Em.d_max;
        m = Em.mask_scale * (1.0f - (q < 0.0f ? 0.0f : (q > 1.0f ? 1.0f : q)));
      }
      M[e] = m;
    }
    // forward scan to the transmittances
    double dens[16], t[16];
    for (int i = 0; i < 16; ++i) dens[i] = 0.0;
    for (int e = 0; e < E; ++e)
      for (int i = 0; i < 16; ++i) {
        dens[i] += (double)D[e] * T->em[e].eps[i];
        if (T->em[e].has_mask) dens[i] += (double)M[e] * T->em[e].mask_eps[i];
      }
    for (int i = 0; i < 16; ++i) t[i] = pow(10.0, -dens[i]) * T->scanner[i];
    // reverse: out = scale * X xyz,  xyz_c = sum_i trap_i t_i cmf_c(i)
    double gx[3];
    for (int c = 0; c < 3; ++c) {
      gx[c] = 0.0;
      for (int r = 0; r < 3; ++r) gx[c] += T->xyz2ap1[r][c] * (double)gout[p * 3 + r];
      gx[c] *= T->dmin_scale;
    }
    double gdens[16];
    for (int i = 0; i < 16; ++i) {
      const double tw = (i == 0 || i == 15) ? 10.0 : 20.0;
      const double gt = tw * (gx[0] * T->cmf[0][i] + gx[1] * T->cmf[1][i] + gx[2] * T->cmf[2][i]);
      gdens[i] = -ln10 * t[i] * gt;
    }
    float gB = 0.0f;
    for (int e = 0; e < E; ++e) {
      const FilmEmulsionTables& Em = T->em[e];
      double gD = 0.0, gM = 0.0;
      for (int i = 0; i < 16; ++i) {
        gD += gdens[i] * Em.eps[i];
        if (Em.has_mask) gM += gdens[i] * Em.mask_eps[i];
      }
      double gdmax = 0.0;
      if (Em.has_mask) {
        const float q = D[e] / Em.d_max;
        if (q > 0.0f && q < 1.0f) {
          gD += -0.4 * gM;  // dM/dD
          gdmax += 0.4 * gM;
        } else if (q <= 0.0f) {
          gdmax += 0.4 * gM;
        }
      }
      const double ef = eff[e];
      gdmax += gD * (ef > 0.0 ? pow(ef, (double)Em.inv_gamma) : 0.0);
      const double ggamma = ef > 0.0 ? gD * (-(double)D[e] * log(ef) * (double)Em.inv_gamma * (double)Em.inv_gamma) : 0.0;
      const double gP = gD * (double)dDdf[e] * slope[e];
      sbar[(size_t)e * n + p] = (float)gP;
      part[e] += gdmax;
      part[kFilmMaxEmulsions + e] += ggamma;
      if (B && Em.gain > 0.0f) {
        part[2 * kFilmMaxEmulsions + e] += gP * (double)bv;
        gB += (float)(gP * (double)Em.gain);
      }
    }
    if (bbar) bbar[p] = gB;
  }
  block_accumulate<3 * kFilmMaxEmulsions>(part, 3 * kFilmMaxEmulsions, acc);
}

// dL/dC[k][e] = sum_p w_k(p) Abar_e(p)
__global__ void __launch_bounds__(256) fit_back_expose_kernel(const float* __restrict__ rgb,
                                                              const float* __restrict__ abar, size_t n,
                                                              FitExposeConsts k, double* __restrict__ acc) {
  double part[3 * kFilmMaxEmulsions];
  for (int i = 0; i < 3 * kFilmMaxEmulsions; ++i) part[i] = 0.0;
  for (size_t p = (size_t)blockIdx.x * blockDim.x + threadIdx.x; p < n; p += (size_t)gridDim.x * blockDim.x) {
    const double r = rgb[p * 3], g = rgb[p * 3 + 1], b = rgb[p * 3 + 2];
    double w[3];
    for (int i = 0; i < 3; ++i) {
      const double v = k.minv[i][0] * r + k.minv[i][1] * g + k.minv[i][2] * b;
      w[i] = v < 0.0 ? 0.0 : v;
    }
    for (int e = 0; e < k.n_emulsions; ++e) {
      const double a = abar[(size_t)e * n + p];
      for (int i = 0; i < 3; ++i) part[i * kFilmMaxEmulsions + e] += w[i] * a;
    }
  }
  block_accumulate<3 * kFilmMaxEmulsions>(part, 3 * kFilmMaxEmulsions, acc);
}

// acc[0] += sum_p a(p) (b(p) + c(p))
__global__ void __launch_bounds__(256) fit_dot_kernel(const float* __restrict__ a, const float* __restrict__ b,
                                                      const float* __restrict__ c, size_t n, double* __restrict__ acc) {
  double part[1] = {0.0};
  for (size_t p = (size_t)blockIdx.x * blockDim.x + threadIdx.x; p < n; p += (size_t)gridDim.x * blockDim.x)
    part[0] += (double)a[p] * ((double)b[p] + (double)c[p]);
  block_accumulate<1>(part, 1, acc);
}

__global__ void fill_f64_kernel(double* p, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = 0.0;
}

}  // namespace

hipError_t film_fit_expose(const float* rgb, size_t n, const FitExposeConsts& k, float* planes, hipStream_t s) {
  hipLaunchKernelGGL(fit_expose_kernel, dim3(grid_of(n)), dim3(256), 0, s, rgb, n, k, planes);
  return hipGetLastError();
}

hipError_t film_blur_adjoint(const float* in, float* tmp, const float* orig, float* out, int w, int h,
                             const float* taps, int radius, float c0, float c1, hipStream_t s) {
  const size_t n = (size_t)w * h;
  hipLaunchKernelGGL(blur_vT_kernel, dim3(grid_of(n)), dim3(256), 0, s, in, tmp, w, h, taps, radius);
  hipLaunchKernelGGL(blur_hT_kernel, dim3(grid_of(n)), dim3(256), 0, s, tmp, orig, out, w, h, taps, radius, c0, c1);
  return hipGetLastError();
}

hipError_t film_fit_back_scan(const float* S, const float* B, const FilmDeviceTables* T, const float* gout, size_t n,
                              float* sbar, float* bbar, double* acc, hipStream_t s) {
  hipLaunchKernelGGL(fit_back_scan_kernel, dim3(grid_of(n)), dim3(256), 0, s, S, B, T, gout, n, sbar, bbar, acc);
  return hipGetLastError();
}

hipError_t film_fit_back_expose(const float* rgb, const float* abar, size_t n, const FitExposeConsts& k, double* acc,
                                hipStream_t s) {
  hipLaunchKernelGGL(fit_back_expose_kernel, dim3(grid_of(n)), dim3(256), 0, s, rgb, abar, n, k, acc);
  return hipGetLastError();
}

hipError_t film_fit_dot(const float* a, const float* b, const float* c, size_t n, double* acc, hipStream_t s) {
  hipLaunchKernelGGL(fit_dot_kernel, dim3(grid_of(n)), dim3(256), 0, s, a, b, c, n, acc);
  return hipGetLastError();
}

hipError_t film_fit_zero(double* p, int n, hipStream_t s) {
  hipLaunchKernelGGL(fill_f64_kernel, dim3((n + 255) / 256), dim3(256), 0, s, p, n);
  return hipGetLastError();
}

}  // namespace imops
