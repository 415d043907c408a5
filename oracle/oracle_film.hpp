// Oracle (test infrastructure only): film simulation declarations.
#pragma once
#include <cstddef>
#include <cstdint>
#include <vector>

namespace film_data { struct Stock; }

namespace oracle {

// Intermediate planes of one run (emulsion-major, width*height each).
struct FilmTrace {
  std::vector<double> latent, image_dye, mask_dye;
  float mid[3] = {0, 0, 0}, dmin[3] = {0, 0, 0};
};

const film_data::Stock* film_stock(int idx);

// Parameters of the differentiable film model (emulsions top to bottom).
struct FilmFitParams {
  int n_emulsions;
  double sens[8][16];          // sensitivity samples 400..700 nm
  double d_max[8], gamma[8];   // characteristic curve D = d_max f^(1/gamma)
  double hal_strength;         // halation gain multiplier
  double hal_radius;           // halation PSF width multiplier
};

template <class P>
int film_fit_process(float* rgb, size_t w, size_t h, int stock_idx, float pitch_um, const FilmFitParams& fit);
int film_fit_nominal(int stock_idx, FilmFitParams* out);

// film::process on an interleaved ACEScg buffer (absolute luminance), in place.
// output: 0 NegativeLinear, 1 PositiveLinear.  shutter_seconds <= 0: 1/box ISO.
template <class P>
int film_process(float* rgb, size_t w, size_t h, int stock_idx, int format, uint64_t seed, int output,
                 float shutter_seconds, FilmTrace* trace);

void film_blur(float* buf, size_t w, size_t h, float sigma);
void film_grain_noise(float* out, size_t w, size_t h, uint64_t seed, size_t layer);
int film_capture_lut(int stock_idx, int emulsion, double* logs, double* fracs, double params[3]);
double film_lut_sample(int stock_idx, int emulsion, double phi);
void film_upsample(const float rgb[3], double out[16]);
void film_spectrum_to_acescg(const double s[16], double out[3]);
int film_normalized_dmin(int stock_idx, float out[3]);

}  // namespace oracle
