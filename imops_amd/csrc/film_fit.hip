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
// DIR inhibition and adjacency off.  The reverse pass is f32 arithmetic (the
// capture LUT lookup aside); parameter gradients are summed in f64 in a fixed
// order, so they are deterministic.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <algorithm>
#include <type_traits>

#include "film.hpp"

namespace imops {
namespace {

int grid_of(size_t n) {
  size_t g = (n + 255) / 256;
  return (int)(g > 16384 ? 16384 : (g ? g : 1));
}

__global__ void fit_expose_kernel(const float* __restrict__ rgb, size_t n, FitExposeConsts k, float* __restrict__ planes) {
  rgb += blockIdx.y * 3 * n;
  planes += blockIdx.y * k.n_emulsions * n;
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

// Block-wide f64 sums of kGroups x kGroup per-thread values, stored to this
// block's row of the partial-sum table (kFitSlots doubles per block, block
// (x, y) at row y * gridDim.x + x; group g's entry e at slot0 + 8 g + e);
// fit_reduce_kernel adds the rows in a fixed order, so the gradients do not
// depend on the order in which blocks finish.  Sizes are compile-time so the
// per-thread sums stay in registers.
template <int kGroup, int kGroups>
__device__ void block_partials(double (&v)[kGroup * kGroups], double* part_rows, int slot0) {
  constexpr int kCnt = kGroup * kGroups;
  __shared__ double part[4][kCnt];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
  for (int k = 0; k < kCnt; ++k) {
    double x = v[k];
    for (int off = 32; off > 0; off >>= 1) x += __shfl_xor(x, off, 64);
    if (lane == 0) part[wave][k] = x;
  }
  __syncthreads();
  if (threadIdx.x < kCnt) {
    double s = 0.0;
    for (int wv = 0; wv < (int)(blockDim.x >> 6); ++wv) s += part[wv][threadIdx.x];
    const int slot = slot0 + (threadIdx.x / kGroup) * kFilmMaxEmulsions + threadIdx.x % kGroup;
    part_rows[((size_t)blockIdx.y * gridDim.x + blockIdx.x) * kFitSlots + slot] = s;
  }
}

// One batch image's pixels [0, n) split over gridDim.x blocks of 256 threads.
#define FIT_PIXELS(p, n) \
  for (size_t p = (size_t)blockIdx.x * blockDim.x + threadIdx.x; p < (n); p += (size_t)gridDim.x * blockDim.x)

// Reverse of capture + reduction + scan for one image of a batch.  In: S planes,
// bounce B, dL/dout.  Out: dL/dS planes (into sbar), dL/dB, and f64 sums of
// dL/dd_max, dL/dgamma, dL/dgain (3 x E).  f32 arithmetic on f32 copies of
// the capture curves and spectral tables (staged in LDS), with the per-block
// parameter sums in f64.
// kNB: bins of the spectral grid (16 or 81).
// kK: 0 for the 64-entry crystal LUT, else the knots of a knot curve; then the
// f64 sums of dL/d(developable fraction at knot k) go to kFitSlotKnot + 8 k + e
// (linear interpolation in log10 fluence: weights 1 - t and t at the bracketing
// knots, weight 1 at the clamped end).
// kL2: `gout` holds a target image instead; the kernel evaluates the model
// output from the same per-pixel quantities (f32), takes dL/dout = 2 (out -
// target) and adds the f64 sum of (out - target)^2 at kFitSlotLoss.
template <int kE, bool kL2, int kNB, int kK>
__global__ void __launch_bounds__(256, 2) fit_back_scan_kernel(const float* __restrict__ S, const float* __restrict__ B,
                                                            const FilmDeviceTables* __restrict__ T,
                                                            const float* __restrict__ gout, size_t n,
                                                            float* __restrict__ sbar, float* __restrict__ bbar,
                                                            double* __restrict__ part_rows) {
  constexpr int kN = kK > 0 ? kK : 64;  // curve samples
  __shared__ float eps_s[kE][kNB], meps_s[kE][kNB], cmf_s[3][kNB], scan_s[kNB], x2a_s[3][3];
  __shared__ float llog_s[kE][kN], lfrac_s[kE][kN];
  constexpr int E = kE;  // == T->n_emulsions
  for (int i = threadIdx.x; i < E * kN; i += blockDim.x) {
    llog_s[i / kN][i % kN] = (float)T->em[i / kN].lut_log[i % kN];
    lfrac_s[i / kN][i % kN] = (float)T->em[i / kN].lut_frac[i % kN];
  }
  for (int i = threadIdx.x; i < kE * kNB; i += blockDim.x) {
    const int e = i / kNB, l = i % kNB;
    eps_s[e][l] = (float)T->em[e].eps[l];
    meps_s[e][l] = T->em[e].has_mask ? (float)T->em[e].mask_eps[l] : 0.0f;
  }
  for (int i = threadIdx.x; i < 3 * kNB; i += blockDim.x) {
    const int c = i / kNB, l = i % kNB;
    const double tw = (l == 0 || l == kNB - 1) ? 0.5 * T->dl : T->dl;  // trapezoid weights of the scan integral
    cmf_s[c][l] = (float)(tw * T->cmf[c][l]);
  }
  for (int i = threadIdx.x; i < kNB; i += blockDim.x) scan_s[i] = (float)T->scanner[i];
  if (threadIdx.x < 9) x2a_s[threadIdx.x / 3][threadIdx.x % 3] = (float)(T->xyz2ap1[threadIdx.x / 3][threadIdx.x % 3] * T->dmin_scale);
  __syncthreads();
  const size_t b = blockIdx.y;  // image b of a batch
  S += b * E * n;
  sbar += b * E * n;
  if (B) B += b * n;
  if (bbar) bbar += b * n;
  gout += b * 3 * n;
  double part[3 * kE];
  for (int k = 0; k < 3 * kE; ++k) part[k] = 0.0;
  double kpart[kK > 0 ? kK * kE : 1];
  for (int k = 0; k < (kK > 0 ? kK * kE : 1); ++k) kpart[k] = 0.0;
  double loss[1] = {0.0};
  const float ln10f = 2.30258512f, log2_10f = 3.32192802f;
  FIT_PIXELS(p, n) {
    float D[kE], M[kE], dDdf[kE], eff[kE], slope[kE];
    int klo[kE];
    float kt[kE];
    const float bv = B ? B[p] : 0.0f;
    for (int e = 0; e < E; ++e) {
      const FilmEmulsionTables& Em = T->em[e];
      const float phi = S[(size_t)e * n + p] + (B && Em.gain > 0.0f ? Em.gain * bv : 0.0f);
      const float x = phi * Em.eta;
      const float* ll = llog_s[e];
      const float* lf = lfrac_s[e];
      float fr, sl = 0.0f;
      int lo = 0;
      float t = 0.0f;
      if (!(x > 0.0f)) {
        fr = lf[0];
      } else {
        const float lp = log10f(x);
        if (lp <= ll[0]) {
          fr = lf[0];
        } else if (lp >= ll[kN - 1]) {
          fr = lf[kN - 1];
          lo = kN - 2;
          t = 1.0f;
        } else {
          int hi = kN - 1;
          while (hi - lo > 1) {
            const int mid = (lo + hi) / 2;
            if (ll[mid] <= lp) lo = mid;
            else hi = mid;
          }
          const float dl = ll[hi] - ll[lo];
          t = (lp - ll[lo]) / dl;
          fr = lf[lo] * (1.0f - t) + lf[hi] * t;
          sl = (lf[hi] - lf[lo]) / dl / (x * ln10f) * Em.eta;  // df/dP
        }
      }
      klo[e] = lo;
      kt[e] = t;
      slope[e] = sl;
      const bool inside = fr > 0.0f && fr < 1.0f;
      const float c = fr < 0.0f ? 0.0f : (fr > 1.0f ? 1.0f : fr);
      const float ef = Em.reversal ? 1.0f - c : c;
      eff[e] = ef;
      D[e] = Em.d_max * powf(ef, Em.inv_gamma);
      const float deff = ef > 0.0f ? Em.inv_gamma * D[e] / ef : 0.0f;  // d_max g^-1 ef^(1/g - 1)
      dDdf[e] = inside ? (Em.reversal ? -deff : deff) : 0.0f;
      float m = 0.0f;
      if (Em.has_mask) {
        const float q = D[e] / Em.d_max;
        m = Em.mask_scale * (1.0f - (q < 0.0f ? 0.0f : (q > 1.0f ? 1.0f : q)));
      }
      M[e] = m;
    }
    // reverse: out_r = sum_c X[r][c] xyz_c,  xyz_c = sum_i trap_i t_i cmf_c(i)
    float gx[3];
    float gDs[kE], gMs[kE];
    for (int e = 0; e < kE; ++e) gDs[e] = gMs[e] = 0.0f;
    if constexpr (kL2) {
      // one pass over the grid: xyz_c = sum_i u_c(i) with u_c(i) = trap_i t_i cmf_c(i),
      // and the per-channel parts of the reverse, K_c,e = sum_i u_c(i) eps_e(i) (and the
      // mask's), so that once the residual gives gx after the loop,
      // dL/dD_e = -ln10 sum_c gx_c K_c,e (one 10^-d per bin, every table read once)
      float xyz[3] = {0.0f, 0.0f, 0.0f}, Kd[3][kE], Km[3][kE];
      for (int c = 0; c < 3; ++c)
        for (int e = 0; e < kE; ++e) Kd[c][e] = Km[c][e] = 0.0f;
#pragma unroll 2
      for (int i = 0; i < kNB; ++i) {
        float d = 0.0f;
        for (int e = 0; e < kE; ++e) d += D[e] * eps_s[e][i] + M[e] * meps_s[e][i];
        const float t = exp2f(-d * log2_10f) * scan_s[i];
        for (int c = 0; c < 3; ++c) {
          const float u = t * cmf_s[c][i];
          xyz[c] += u;
          for (int e = 0; e < kE; ++e) {
            Kd[c][e] += u * eps_s[e][i];
            Km[c][e] += u * meps_s[e][i];
          }
        }
      }
      float g[3];
      for (int r = 0; r < 3; ++r) {
        const float o = x2a_s[r][0] * xyz[0] + x2a_s[r][1] * xyz[1] + x2a_s[r][2] * xyz[2];
        const float diff = o - gout[p * 3 + r];
        loss[0] += (double)diff * (double)diff;
        g[r] = 2.0f * diff;
      }
      for (int c = 0; c < 3; ++c) gx[c] = x2a_s[0][c] * g[0] + x2a_s[1][c] * g[1] + x2a_s[2][c] * g[2];
      for (int e = 0; e < kE; ++e) {
        gDs[e] = -ln10f * (gx[0] * Kd[0][e] + gx[1] * Kd[1][e] + gx[2] * Kd[2][e]);
        gMs[e] = -ln10f * (gx[0] * Km[0][e] + gx[1] * Km[1][e] + gx[2] * Km[2][e]);
      }
    } else {
      for (int c = 0; c < 3; ++c)
        gx[c] = x2a_s[0][c] * gout[p * 3] + x2a_s[1][c] * gout[p * 3 + 1] + x2a_s[2][c] * gout[p * 3 + 2];
      // per wavelength: the forward transmittance t = scanner 10^-dens, then
      // dL/ddens = -ln10 t (gx . cmf) into the per-emulsion dye and mask sums
#pragma unroll 2
      for (int i = 0; i < kNB; ++i) {
        float d = 0.0f;
        for (int e = 0; e < kE; ++e) d += D[e] * eps_s[e][i] + M[e] * meps_s[e][i];
        const float t = exp2f(-d * log2_10f) * scan_s[i];
        const float gd = -ln10f * t * (gx[0] * cmf_s[0][i] + gx[1] * cmf_s[1][i] + gx[2] * cmf_s[2][i]);
        for (int e = 0; e < kE; ++e) {
          gDs[e] += gd * eps_s[e][i];
          gMs[e] += gd * meps_s[e][i];
        }
      }
    }
    float gB = 0.0f;
    for (int e = 0; e < E; ++e) {
      const FilmEmulsionTables& Em = T->em[e];
      float gD = gDs[e];
      const float gM = gMs[e];
      float gdmax = 0.0f;
      if (Em.has_mask) {
        const float q = D[e] / Em.d_max;
        if (q > 0.0f && q < 1.0f) {
          gD += -0.4f * gM;  // dM/dD
          gdmax += 0.4f * gM;
        } else if (q <= 0.0f) {
          gdmax += 0.4f * gM;
        }
      }
      const float ef = eff[e];
      gdmax += gD * (D[e] / Em.d_max);  // dD/dd_max = ef^(1/g)
      const float ggamma = ef > 0.0f ? -gD * D[e] * logf(ef) * Em.inv_gamma * Em.inv_gamma : 0.0f;
      const float gF = gD * dDdf[e];  // dL/d(developable fraction)
      const float gP = gF * slope[e];
      sbar[(size_t)e * n + p] = gP;
      part[e] += (double)gdmax;
      part[kE + e] += (double)ggamma;
      if (B && Em.gain > 0.0f) {
        part[2 * kE + e] += (double)gP * (double)bv;
        gB += gP * Em.gain;
      }
      if constexpr (kK > 0) {
        const double w1 = (double)gF * (double)kt[e], w0 = (double)gF - w1;
#pragma unroll
        for (int k = 0; k < kK; ++k)
          kpart[k * kE + e] += k == klo[e] ? w0 : (k == klo[e] + 1 ? w1 : 0.0);
      }
    }
    if (bbar) bbar[p] = gB;
  }
  block_partials<kE, 3>(part, part_rows, kFitSlotDmax);
  if constexpr (kK > 0) block_partials<kE, kK>(kpart, part_rows, kFitSlotKnot);
  if constexpr (kL2) block_partials<1, 1>(loss, part_rows, kFitSlotLoss);
}

// dL/dC[k][e] = sum_p w_k(p) Abar_e(p)
template <int kE>
__global__ void __launch_bounds__(256) fit_back_expose_kernel(const float* __restrict__ rgb,
                                                              const float* __restrict__ abar, size_t n,
                                                              FitExposeConsts k, double* __restrict__ part_rows) {
  rgb += blockIdx.y * 3 * n;
  abar += blockIdx.y * kE * n;
  double part[3 * kE];
  for (int i = 0; i < 3 * kE; ++i) part[i] = 0.0;
  FIT_PIXELS(p, n) {
    const double r = rgb[p * 3], g = rgb[p * 3 + 1], b = rgb[p * 3 + 2];
    double w[3];
    for (int i = 0; i < 3; ++i) {
      const double v = k.minv[i][0] * r + k.minv[i][1] * g + k.minv[i][2] * b;
      w[i] = v < 0.0 ? 0.0 : v;
    }
    for (int e = 0; e < kE; ++e) {
      const double a = abar[(size_t)e * n + p];
      for (int i = 0; i < 3; ++i) part[i * kE + e] += w[i] * a;
    }
  }
  block_partials<kE, 3>(part, part_rows, kFitSlotC);
}

// sum_p a(p) (b(p) + c(p)) over one plane per image
__global__ void __launch_bounds__(256) fit_dot_kernel(const float* __restrict__ a, const float* __restrict__ b,
                                                      const float* __restrict__ c, size_t n,
                                                      double* __restrict__ part_rows) {
  const size_t off = blockIdx.y * n;
  double part[1] = {0.0};
  FIT_PIXELS(p, n) part[0] += (double)a[off + p] * ((double)b[off + p] + (double)c[off + p]);
  block_partials<1, 1>(part, part_rows, kFitSlotSigma);
}

__global__ void fill_f64_kernel(double* p, size_t n) {
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) p[i] = 0.0;
}

// acc[slot] = sum over the partial rows, one block per slot: thread t adds rows
// t, t + 256, ... in order, then a fixed butterfly within each wave and the four
// wave sums in wave order (a fixed order, whatever the timing).
constexpr int kReduceThreads = 256;
__global__ void __launch_bounds__(kReduceThreads) fit_reduce_kernel(const double* __restrict__ part_rows, int rows,
                                                                    double* __restrict__ acc) {
  __shared__ double wsum[kReduceThreads / 64];
  const int slot = blockIdx.x, lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  double s = 0.0;
  for (int r = threadIdx.x; r < rows; r += kReduceThreads) s += part_rows[(size_t)r * kFitSlots + slot];
  for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off, 64);
  if (lane == 0) wsum[wave] = s;
  __syncthreads();
  if (threadIdx.x == 0) {
    double t = 0.0;
    for (int k = 0; k < kReduceThreads / 64; ++k) t += wsum[k];
    acc[slot] = t;
  }
}

}  // namespace

hipError_t film_fit_expose(const float* rgb, size_t n, const FitExposeConsts& k, float* planes, hipStream_t s,
                           int images) {
  hipLaunchKernelGGL(fit_expose_kernel, dim3(grid_of(n), images), dim3(256), 0, s, rgb, n, k, planes);
  return hipGetLastError();
}

int fit_blocks_per_image(size_t n, int images) {
  const int want = std::max(1, kFitReduceRows / std::max(images, 1));
  return std::min(want, grid_of(n));
}

int fit_partial_rows(size_t n, int images) { return fit_blocks_per_image(n, images) * images; }

// Calls f(std::integral_constant<int, E>{}) for the emulsion count E in 1..kFilmMaxEmulsions.
template <class F>
hipError_t with_emulsions(int emulsions, F&& f) {
  switch (emulsions) {
    case 1: f(std::integral_constant<int, 1>{}); break;
    case 2: f(std::integral_constant<int, 2>{}); break;
    case 3: f(std::integral_constant<int, 3>{}); break;
    case 4: f(std::integral_constant<int, 4>{}); break;
    case 5: f(std::integral_constant<int, 5>{}); break;
    case 6: f(std::integral_constant<int, 6>{}); break;
    case 7: f(std::integral_constant<int, 7>{}); break;
    case 8: f(std::integral_constant<int, 8>{}); break;
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

template <int kE, int kNB, int kK>
void launch_back_scan(dim3 grid, hipStream_t s, const float* S, const float* B, const FilmDeviceTables* T,
                      const float* gout, size_t n, float* sbar, float* bbar, double* part_rows, bool l2_target) {
  if (l2_target)
    hipLaunchKernelGGL((fit_back_scan_kernel<kE, true, kNB, kK>), grid, dim3(256), 0, s, S, B, T, gout, n, sbar, bbar,
                       part_rows);
  else
    hipLaunchKernelGGL((fit_back_scan_kernel<kE, false, kNB, kK>), grid, dim3(256), 0, s, S, B, T, gout, n, sbar, bbar,
                       part_rows);
}

hipError_t film_fit_back_scan(const float* S, const float* B, const FilmDeviceTables* T, int emulsions, int bins,
                              int knots, const float* gout, size_t n, float* sbar, float* bbar, double* part_rows,
                              hipStream_t s, int images, bool l2_target) {
  const dim3 grid(fit_blocks_per_image(n, images), images);
  if (bins == 16 && knots == 0)
    return with_emulsions(emulsions, [&](auto e) {
      launch_back_scan<decltype(e)::value, 16, 0>(grid, s, S, B, T, gout, n, sbar, bbar, part_rows, l2_target);
    });
  // the 81-bin grid and the 5-knot curve: up to kFitMaxEmulsionsExt emulsions
  if (emulsions < 1 || emulsions > kFitMaxEmulsionsExt || (bins != 16 && bins != 81) || (knots != 0 && knots != 5))
    return hipErrorInvalidValue;
  return with_emulsions(emulsions, [&](auto e) {
    constexpr int kE = decltype(e)::value <= kFitMaxEmulsionsExt ? decltype(e)::value : 1;
    if (bins == 81 && knots == 5) launch_back_scan<kE, 81, 5>(grid, s, S, B, T, gout, n, sbar, bbar, part_rows, l2_target);
    else if (bins == 81) launch_back_scan<kE, 81, 0>(grid, s, S, B, T, gout, n, sbar, bbar, part_rows, l2_target);
    else launch_back_scan<kE, 16, 5>(grid, s, S, B, T, gout, n, sbar, bbar, part_rows, l2_target);
  });
}

hipError_t film_fit_back_expose(const float* rgb, const float* abar, size_t n, const FitExposeConsts& k,
                                double* part_rows, hipStream_t s, int images) {
  const dim3 grid(fit_blocks_per_image(n, images), images);
  return with_emulsions(k.n_emulsions, [&](auto e) {
    hipLaunchKernelGGL(fit_back_expose_kernel<decltype(e)::value>, grid, dim3(256), 0, s, rgb, abar, n, k, part_rows);
  });
}

hipError_t film_fit_dot(const float* a, const float* b, const float* c, size_t n, double* part_rows, hipStream_t s,
                        int images) {
  hipLaunchKernelGGL(fit_dot_kernel, dim3(fit_blocks_per_image(n, images), images), dim3(256), 0, s, a, b, c, n,
                     part_rows);
  return hipGetLastError();
}

hipError_t film_fit_zero(double* p, size_t n, hipStream_t s) {
  hipLaunchKernelGGL(fill_f64_kernel, dim3((unsigned)std::min<size_t>((n + 255) / 256, 4096)), dim3(256), 0, s, p, n);
  return hipGetLastError();
}

hipError_t film_fit_reduce(const double* part_rows, int rows, double* acc, hipStream_t s) {
  hipLaunchKernelGGL(fit_reduce_kernel, dim3(kFitSlots), dim3(kReduceThreads), 0, s, part_rows, rows, acc);
  return hipGetLastError();
}

}  // namespace imops
