// Film simulation kernel interfaces (film.hip).
#pragma once
#include <hip/hip_runtime.h>
#include <stddef.h>
#include <stdint.h>

namespace imops {

constexpr int kFilmMaxLayers = 12;
constexpr int kFilmMaxEmulsions = 8;
// Spectral grids: the reference's 16 samples 400..700 nm (step 20) and the
// 81-sample extension 380..780 nm (step 5).
constexpr int kFilmMaxBins = 81;
constexpr int kFilmLutMax = 64;  // capture curve knots (the reference's LUT has 64)
constexpr int kFilmReduceBlocks = 1024;  // f64 partial sums of the grain variance (+1 slot for the norm)

enum FilmBlurMode : int { BLUR_STORE = 0, BLUR_MIX = 1, BLUR_ADJ = 2 };

// A batch of W x H planes blurred by one launch: plane i reads in + i * in,
// its epilogue source orig + i * orig and writes out + i * out (element
// strides); the row pass keeps plane i at tmp + i * W * H.
struct BlurPlanes {
  int count = 1;
  size_t in = 0, orig = 0, out = 0;
};

struct FilmEmulsionTables {
  double lut_log[kFilmLutMax], lut_frac[kFilmLutMax];  // capture curve knots (log10 fluence -> developable fraction)
  int lut_n;                                           // knots in use: 64 (crystal LUT) or a knot curve's count
  double eps[kFilmMaxBins], mask_eps[kFilmMaxBins];    // dye / coloured-mask absorptivity
  float gain;                        // halation bounce gain r_e * bleed_e (<= 0: none)
  float eta;                         // reciprocity efficiency at the shutter time
  float d_max, inv_gamma, mask_scale;
  int has_mask, reversal;
};

struct FilmDeviceTables {
  int n_bins;              // 16 or 81 samples of the spectral grid
  double dl;               // grid step (nm): 20 or 5
  double band_inv;         // 1 / (last - first wavelength): the band mean's divisor (1/300 on the reference grid)
  double minv[3][3];       // ACEScg -> basis weights
  double basis[3][kFilmMaxBins];   // Gaussian basis spectra
  double e_rel[kFilmMaxBins];      // 550 / lambda
  double inv_e_rel[kFilmMaxBins];  // 1 / e_rel: the per-pixel division becomes a product (<= 1 ulp in f64)
  int n_absorbing;         // absorbing layers, top to bottom
  int is_emulsion[kFilmMaxLayers];
  double trans[kFilmMaxLayers][kFilmMaxBins];
  int n_emulsions;
  FilmEmulsionTables em[kFilmMaxEmulsions];
  float dir[kFilmMaxEmulsions][kFilmMaxEmulsions];  // [source][target]
  int dir_valid[kFilmMaxEmulsions][kFilmMaxEmulsions];
  double scanner[kFilmMaxBins];
  double cmf[3][kFilmMaxBins];
  double xyz2ap1[3][3];
  double dmin_scale;       // 1 / peak of the Dmin scan
  int invert;              // PositiveLinear
  float inv_dmin[3], inv_shoulder[3], inv_gain[3];
  // The mixed form's FP32 scan: absorptivities pre-scaled by log2(10), so that
  // 10^-d = exp2(-d'), and scanner x CMF x trapezoid weight per bin (0 where unused)
  float scan_eps2[kFilmMaxEmulsions][kFilmMaxBins];
  float scan_mask2[kFilmMaxEmulsions][kFilmMaxBins];
  float scan_w[3][kFilmMaxBins];
  float scan_xyz2ap1[3][3];
};

// Differentiable film model (film_fit.hip).  Parameter gradients are f64 sums:
// each reducing kernel stores one row of kFitSlots partial sums per block
// (d_max, gamma, gain per emulsion at kFitSlotDmax + {0, 8, 16}; the exposure
// contraction C[k][e] at kFitSlotC + 8 k + e; the halation width at
// kFitSlotSigma; the L2 loss at kFitSlotLoss; the developable fraction at knot
// k of a knot curve at kFitSlotKnot + 8 k + e), and film_fit_reduce adds the
// rows in order into acc.
constexpr int kFitSlots = 128;
constexpr int kFitSlotDmax = 0, kFitSlotC = 24, kFitSlotSigma = 48, kFitSlotLoss = 49, kFitSlotKnot = 64;
constexpr int kFitMaxKnots = 8;
constexpr int kFitMaxEmulsionsExt = 4;  // emulsions of the 81-bin / knot-curve reverse kernels
constexpr int kFitReduceRows = 1024;  // blocks of a reducing launch (whole batch), when the pixels allow
struct FitExposeConsts {
  double minv[3][3];                 // ACEScg -> basis weights
  double c[3][kFilmMaxEmulsions];    // band-absorbed fluence of emulsion e per unit basis weight k
  int n_emulsions;
};

// fp32: the mixed form's FP32 per-pixel arithmetic instead of the reference's FP64 (capture LUT, scan)
// bins: the tables' spectral grid (16 or 81; the scan kernels are specialised on it)
// The exposure, collapsed over the spectral grid to relu(M^-1 rgb) . C in f64 (the stage and the
// differentiable model), batched over `images` images: rgb image b at rgb + 3 b n, its E planes at planes + b E n.
hipError_t film_fit_expose(const float* rgb, size_t n, const FitExposeConsts& k, float* planes, hipStream_t s,
                           int images = 1);
// out = c0 * orig + c1 * (V o H)^T in   (the transpose of film_blur's reflect-101 Gaussian; radius <= w-1, h-1);
// orig, when used, is out (same strides)
hipError_t film_blur_adjoint(const float* in, float* tmp, const float* orig, float* out, int w, int h,
                             const float* taps, int radius, float c0, float c1, hipStream_t s,
                             const BlurPlanes& pl = BlurPlanes{});
// Rows of partial sums a reducing launch over `images` images of n pixels writes.
int fit_partial_rows(size_t n, int images);
// S, sbar: E planes per image; B, bbar: one plane per image; gout: 3n per image
// (emulsions = T->n_emulsions, 1..kFilmMaxEmulsions; bins = T->n_bins, 16 or
// 81; knots: 0 for the crystal LUT, 5 for a 5-knot curve; the 81-bin grid and
// the knot curve take up to kFitMaxEmulsionsExt emulsions).  l2_target: gout
// is a target image and dL/dout = 2 (model output - target), the loss summed at kFitSlotLoss.
hipError_t film_fit_back_scan(const float* S, const float* B, const FilmDeviceTables* T, int emulsions, int bins,
                              int knots, const float* gout, size_t n, float* sbar, float* bbar, double* part_rows,
                              hipStream_t s, int images = 1, bool l2_target = false);
hipError_t film_fit_back_expose(const float* rgb, const float* abar, size_t n, const FitExposeConsts& k,
                                double* part_rows, hipStream_t s, int images = 1);
// sum_p a (b + c) over one plane per image (planes of consecutive images n apart)
hipError_t film_fit_dot(const float* a, const float* b, const float* c, size_t n, double* part_rows, hipStream_t s,
                        int images = 1);
hipError_t film_fit_zero(double* p, size_t n, hipStream_t s);
hipError_t film_fit_reduce(const double* part_rows, int rows, double* acc, hipStream_t s);
hipError_t film_blur(const float* in, float* tmp, const float* orig, float* out, int w, int h, const float* taps,
                     int radius, int mode, float a, hipStream_t s, const BlurPlanes& pl = BlurPlanes{});
// separable blur with different horizontal and vertical taps (same radius), plain store
hipError_t film_blur2(const float* in, float* tmp, float* out, int w, int h, const float* taps_h, const float* taps_v,
                      int radius, hipStream_t s, const BlurPlanes& pl = BlurPlanes{});
// Batched over `images`: absorbed, image, mask E planes per image, bounce one plane per image.
hipError_t film_capture_reduce(const float* absorbed, const float* bounce, size_t n, const FilmDeviceTables* T,
                               float* image, float* mask, hipStream_t s, int images = 1, bool fp32 = false);
hipError_t film_dir_apply(float* image, const float* inh, size_t n, const FilmDeviceTables* T, hipStream_t s);
// Grain (grain.rs:53-100).  Noise: the reference's SplitMix64 Irwin-Hall rows
// (philox = false, key0 = the row key of y = 0), or Philox4x32-10 normals
// (philox = true; key = the film seed, counter = (pixel, layer)).  Amplitude:
// the reference's kappa * taper * sqrt(D (Dmax - D)), or a constant density
// sigma `kappa` when `absolute`.
hipError_t film_grain(float* dye, float* noise, float* tmp, int w, int h, uint64_t key0, const float* taps, int radius,
                      double* partial, float d_max, float kappa, hipStream_t s, bool philox = false,
                      uint64_t seed = 0, int layer = 0, bool absolute = false);
hipError_t film_scan(const float* image, const float* mask, size_t n, const FilmDeviceTables* T, float* rgb,
                     hipStream_t s, int images = 1, bool fp32 = false, int bins = 16);

}  // namespace imops
