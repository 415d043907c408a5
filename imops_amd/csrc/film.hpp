// Film simulation kernel interfaces (film.hip).
#pragma once
#include <hip/hip_runtime.h>
#include <stddef.h>
#include <stdint.h>

namespace imops {

constexpr int kFilmMaxLayers = 12;
constexpr int kFilmMaxEmulsions = 8;
constexpr int kFilmReduceBlocks = 1024;  // f64 partial sums of the grain variance (+1 slot for the norm)

enum FilmBlurMode : int { BLUR_STORE = 0, BLUR_MIX = 1, BLUR_ADJ = 2 };

struct FilmEmulsionTables {
  double lut_log[64], lut_frac[64];  // developable-fraction LUT (log10 fluence -> fraction)
  double eps[16], mask_eps[16];      // dye / coloured-mask absorptivity
  float gain;                        // halation bounce gain r_e * bleed_e (<= 0: none)
  float eta;                         // reciprocity efficiency at the shutter time
  float d_max, inv_gamma, mask_scale;
  int has_mask, reversal;
};

struct FilmDeviceTables {
  double minv[3][3];       // ACEScg -> basis weights
  double basis[3][16];     // Gaussian basis spectra
  double e_rel[16];        // 550 / lambda
  int n_absorbing;         // absorbing layers, top to bottom
  int is_emulsion[kFilmMaxLayers];
  double trans[kFilmMaxLayers][16];
  int n_emulsions;
  FilmEmulsionTables em[kFilmMaxEmulsions];
  float dir[kFilmMaxEmulsions][kFilmMaxEmulsions];  // [source][target]
  int dir_valid[kFilmMaxEmulsions][kFilmMaxEmulsions];
  double scanner[16];
  double cmf[3][16];
  double xyz2ap1[3][3];
  double dmin_scale;       // 1 / peak of the Dmin scan
  int invert;              // PositiveLinear
  float inv_dmin[3], inv_shoulder[3], inv_gain[3];
};

// Differentiable film model (film_fit.hip).
struct FitExposeConsts {
  double minv[3][3];                 // ACEScg -> basis weights
  double c[3][kFilmMaxEmulsions];    // band-absorbed fluence of emulsion e per unit basis weight k
  int n_emulsions;
};

hipError_t film_expose(const float* rgb, size_t n, const FilmDeviceTables* T, float* planes, hipStream_t s);
hipError_t film_fit_expose(const float* rgb, size_t n, const FitExposeConsts& k, float* planes, hipStream_t s);
// out = c0 * orig + c1 * (V o H)^T in   (the transpose of film_blur's reflect-101 Gaussian; radius <= w-1, h-1)
hipError_t film_blur_adjoint(const float* in, float* tmp, const float* orig, float* out, int w, int h,
                             const float* taps, int radius, float c0, float c1, hipStream_t s);
hipError_t film_fit_back_scan(const float* S, const float* B, const FilmDeviceTables* T, const float* gout, size_t n,
                              float* sbar, float* bbar, double* acc, hipStream_t s);
hipError_t film_fit_back_expose(const float* rgb, const float* abar, size_t n, const FitExposeConsts& k, double* acc,
                                hipStream_t s);
hipError_t film_fit_dot(const float* a, const float* b, const float* c, size_t n, double* acc, hipStream_t s);
hipError_t film_fit_zero(double* p, int n, hipStream_t s);
hipError_t film_blur(const float* in, float* tmp, const float* orig, float* out, int w, int h, const float* taps,
                     int radius, int mode, float a, hipStream_t s);
// separable blur with different horizontal and vertical taps (same radius), plain store
hipError_t film_blur2(const float* in, float* tmp, float* out, int w, int h, const float* taps_h, const float* taps_v,
                      int radius, hipStream_t s);
hipError_t film_capture_reduce(const float* absorbed, const float* bounce, size_t n, const FilmDeviceTables* T,
                               float* image, float* mask, hipStream_t s);
hipError_t film_dir_apply(float* image, const float* inh, size_t n, const FilmDeviceTables* T, hipStream_t s);
hipError_t film_grain(float* dye, float* noise, float* tmp, int w, int h, uint64_t key0, const float* taps, int radius,
                      double* partial, float d_max, float kappa, hipStream_t s);
hipError_t film_scan(const float* image, const float* mask, size_t n, const FilmDeviceTables* T, float* rgb,
                     hipStream_t s);

}  // namespace imops
