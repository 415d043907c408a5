// Oracle (test infrastructure only): film simulation declarations.
#pragma once
#include <cstddef>
#include <cstdint>
#include <vector>

namespace oracle {

// Intermediate planes of one run (emulsion-major, width*height each).
struct FilmTrace {
  std::vector<double> latent, image_dye, mask_dye;
  float mid[3] = {0, 0, 0}, dmin[3] = {0, 0, 0};
};

// Extension parameters of the film stage (defaults: the reference's behaviour).
struct FilmExt {
  int bins = 16;               // spectral grid: 16 (400..700 / 20 nm) or 81 (380..780 / 5 nm)
  uint64_t stock_seed = 0;     // the synthetic stock (stock index 7)
  bool philox = false;         // Philox4x32-10 normal grain noise
  float grain_sigma = -1.0f;   // >= 0: constant grain density sigma
  float hal_radius = -1.0f;    // >= 0: halation radius in pixels (sigma = radius / 3)
  float hal_strength = -1.0f;  // >= 0: halation gain strength * bleed_e
};

// The differentiable film model: stock (7: the seeded synthetic stock), spectral
// grid (16 or 81 bins) and pixel pitch.
struct FilmFitModel {
  int stock = 1;
  int bins = 16;
  uint64_t stock_seed = 0;
  float pitch_um = 36000.0f / 6016.0f;
};

// Parameters of the differentiable film model (emulsions top to bottom).
struct FilmFitParams {
  int n_emulsions = 0;
  int n_bins = 16;
  int n_knots = 0;             // knot-curve stock: the developable fraction at each knot
  double sens[8][81] = {};     // sensitivity samples on the grid
  double knot_frac[8][8] = {};
  double d_max[8] = {}, gamma[8] = {};  // characteristic curve D = d_max f^(1/gamma)
  double hal_strength = 1.0;   // halation gain multiplier
  double hal_radius = 1.0;     // halation PSF width multiplier
};

template <class P>
int film_fit_process(float* rgb, size_t w, size_t h, const FilmFitModel& m, const FilmFitParams& fit);
int film_fit_nominal(const FilmFitModel& m, FilmFitParams* out);
// dL/dparams of film_fit_process<double> on one image for dL/dout = gout (an f64
// reverse pass written from the forward's equations); out_rgb, when given,
// receives the forward's output (w x h x 3) in f32.
int film_fit_gradient(const float* rgb, const float* gout, size_t w, size_t h, const FilmFitModel& m,
                      const FilmFitParams& fit, FilmFitParams* grad, float* out_rgb = nullptr);

// film::process on an interleaved ACEScg buffer (absolute luminance), in place.
// output: 0 NegativeLinear, 1 PositiveLinear.  shutter_seconds <= 0: 1/box ISO.
template <class P>
int film_process(float* rgb, size_t w, size_t h, int stock_idx, int format, uint64_t seed, int output,
                 float shutter_seconds, FilmTrace* trace, const FilmExt& ext = FilmExt{});

void film_blur(float* buf, size_t w, size_t h, float sigma);
void film_grain_noise(float* out, size_t w, size_t h, uint64_t seed, size_t layer);
void film_philox(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]);
void film_grain_noise_philox(float* out, size_t n, uint64_t seed, size_t layer);
int film_stock_knots(int stock_idx, int bins, uint64_t stock_seed, int emulsion, double* logs, double* fracs);
int film_stock_sensitivity(int stock_idx, int bins, uint64_t stock_seed, int emulsion, double* out);
void film_cmf(int bins, double* out);
int film_capture_lut(int stock_idx, int emulsion, double* logs, double* fracs, double params[3]);
double film_lut_sample(int stock_idx, int emulsion, double phi);
void film_upsample(const float rgb[3], double out[16]);
void film_spectrum_to_acescg(const double s[16], double out[3]);
int film_normalized_dmin(int stock_idx, float out[3]);

}  // namespace oracle
