// Oracle (test infrastructure only): C entry points loaded by the tests, the
// smoke check and the benchmark's CPU-baseline leg through ctypes.
#include <cstdint>
#include <algorithm>
#include <vector>

#include "oracle_bm3d.hpp"
#include "oracle_colour.hpp"
#include "oracle_film.hpp"

using namespace oracle;

extern "C" {

int oracle_amaze(const float* in, size_t w, size_t h, const char* cfa_pattern, size_t crop_x, size_t crop_y,
                 float* out, int fp64) {
  Cfa base;
  if (!base.parse(cfa_pattern)) return -1;
  if (!amaze_size_supported(w, h)) return -2;
  Cfa cfa = base.shift(crop_x, crop_y);  // demosaic.rs:93-98
  if (fp64) amaze_demosaic<double>(in, w, h, cfa, out);
  else amaze_demosaic<float>(in, w, h, cfa, out);
  return 0;
}

// Test hooks of the AMaZE restatement: a planted mistake (0 = none) and the
// per-pass trace of one tile (16 planes of 128 x 128 floats, or NULL).
void oracle_amaze_hooks(int plant, float* trace, int tile_top, int tile_left) {
  g_amaze_hooks.plant = plant;
  g_amaze_hooks.trace = trace;
  g_amaze_hooks.trace_top = tile_top;
  g_amaze_hooks.trace_left = tile_left;
}

int oracle_markesteijn(const float* in, size_t w, size_t h, const char* cfa_pattern, size_t crop_x, size_t crop_y,
                       float* out) {
  Cfa base;
  if (!base.parse(cfa_pattern)) return -1;
  Cfa cfa = base.shift(crop_x, crop_y);
  markesteijn(in, w, h, cfa, out);
  return 0;
}

void oracle_markesteijn3_plant(int plant) { oracle::g_m3_plant = plant; }

int oracle_markesteijn3(const float* in, size_t w, size_t h, const char* cfa_pattern, size_t crop_x, size_t crop_y,
                        int passes, int fp64, float* out) {
  Cfa base;
  if (!base.parse(cfa_pattern) || (passes != 1 && passes != 3) || w == 0 || h == 0) return -1;
  Cfa cfa = base.shift(crop_x, crop_y);
  if (fp64) markesteijn3<double>(in, w, h, cfa, passes, out);
  else markesteijn3<float>(in, w, h, cfa, passes, out);
  return 0;
}

// Fast / SuperFast / SuperSuperFast (demosaic.rs:189-354): factor 2, 4 or 8; out (w/f) x (h/f) x 3.
int oracle_preview_demosaic(const float* in, size_t w, size_t h, const char* cfa_pattern, size_t crop_x,
                            size_t crop_y, int factor, float* out) {
  Cfa base;
  if (!base.parse(cfa_pattern)) return -1;
  Cfa cfa = base.shift(crop_x, crop_y);
  const int c00 = cfa.color_at(0, 0), c01 = cfa.color_at(0, 1), c10 = cfa.color_at(1, 0), c11 = cfa.color_at(1, 1);
  if (c00 > 2 || c01 > 2 || c10 > 2 || c11 > 2) return -1;  // the reference indexes a 3-channel pixel
  const size_t nw = w / factor, nh = h / factor;
  for (size_t k = 0; k < nw * nh; ++k) {
    const size_t r = (k / nw) * factor, c = (k % nw) * factor, tl = r * w + c;
    float v[3] = {0.0f, 0.0f, 0.0f};
    v[c00] = in[tl];
    v[c01] = in[tl + 1];
    v[c10] = in[tl + w];
    v[c11] = in[tl + w + 1];
    for (int ch = 0; ch < 3; ++ch) out[k * 3 + ch] = v[ch];
  }
  return 0;
}

// u16 raw -> normalised f32 CFA plane of the crop window ((DN-black)/(white-black)).
int oracle_normalize_raw(const uint16_t* raw, size_t full_w, size_t full_h, size_t cx, size_t cy, size_t cw,
                         size_t ch, float black, float white, float* out) {
  if (cx + cw > full_w || cy + ch > full_h) return -1;
  const float range = white - black;
  for (size_t y = 0; y < ch; ++y)
    for (size_t x = 0; x < cw; ++x) out[y * cw + x] = ((float)raw[(cy + y) * full_w + cx + x] - black) / range;
  return 0;
}

int oracle_camera_forward_matrix(const float* cal, size_t n, float* fwd9) {
  float f[3][3];
  if (!camera_forward_matrix(cal, n, f)) return -1;
  for (int i = 0; i < 9; ++i) fwd9[i] = f[i / 3][i % 3];
  return 0;
}

int oracle_camera_cst(float* rgb, size_t n, const float* cal, size_t cal_len, int target) {
  float f[3][3];
  if (!camera_forward_matrix(cal, cal_len, f)) return -1;
  camera_cst(rgb, n, f, target);
  return 0;
}

int oracle_cst(float* rgb, size_t n, int src, int dst) {
  cst(rgb, n, src, dst);
  return 0;
}

int oracle_exp(float* rgb, size_t n, float ev) { exp_ev(rgb, n, ev); return 0; }
int oracle_gamma(float* rgb, size_t n, float gamma) { gamma_encode(rgb, n, gamma); return 0; }
int oracle_cfa_coeffs(float* rgb, size_t n, const float* wb) { cfa_coeffs(rgb, n, wb); return 0; }
float oracle_srgb_oetf(float x) { return srgb_oetf(x); }

// BM3D with explicit parameters; params[11] = {intensity, sigma, lambda,
// max_dist, chroma_scale, detail, stride, window, max_group, max_candidates,
// group_transform (0 Walsh-Hadamard, 1 Haar)}.
int oracle_bm3d(float* rgb, size_t w, size_t h, const float* params, int fp64, long long* stats2) {
  Bm3dParams p;
  p.intensity = params[0];
  p.sigma = params[1];
  p.lambda = params[2];
  p.max_dist = params[3];
  p.chroma_scale = params[4];
  p.detail = params[5];
  p.stride = (int)params[6];
  p.window = (int)params[7];
  p.max_group = (int)params[8];
  p.max_candidates = (int)params[9];
  p.group_transform = (int)params[10];
  if (p.group_transform != 0 && p.group_transform != 1) return -1;
  if (p.stride < 1 || p.window < 1 || (p.window & 1) == 0 || p.max_group < 1 || p.max_group > 32 ||
      p.max_candidates < 1)
    return -1;
  Bm3dStats st;
  if (fp64) bm3d<double>(rgb, w, h, p, &st);
  else bm3d<float>(rgb, w, h, p, &st);
  if (stats2) { stats2[0] = st.refs; stats2[1] = st.groups; }
  return 0;
}

int oracle_group_transform_1d(double* d, int n, int kind, int inverse, int fp64) {
  if (n < 1 || n > 32 || (n & (n - 1)) || (kind != 0 && kind != 1)) return -1;
  group_transform_1d(d, n, kind, inverse != 0, fp64 != 0);
  return 0;
}

int oracle_bm3d_intensity_params(float intensity, float* params) {
  Bm3dParams p = Bm3dParams::from_intensity(intensity);
  const float v[10] = {p.intensity, p.sigma, p.lambda, p.max_dist, p.chroma_scale, p.detail,
                       (float)p.stride, (float)p.window, (float)p.max_group, (float)p.max_candidates};
  for (int i = 0; i < 10; ++i) params[i] = v[i];
  return 0;
}

int oracle_chroma_bm3d(float* rgb, size_t w, size_t h, float intensity) {
  chroma_bm3d(rgb, w, h, intensity);
  return 0;
}

int oracle_bm3d_tables(float* dct, float* idct, float* kaiser) {
  bm3d_tables(dct, idct, kaiser);
  return 0;
}

// bm3d.rs:765-776 test image: channel sinusoids in [0,1].
int oracle_bm3d_test_image(float* rgb, size_t w, size_t h) {
  for (size_t y = 0; y < h; ++y)
    for (size_t x = 0; x < w; ++x) {
      float r = (sinf((float)x * 0.12f) * 0.5f + 0.5f) * 255.0f;
      float g = (cosf((float)y * 0.17f) * 0.5f + 0.5f) * 255.0f;
      float b = (sinf((float)(x + y) * 0.09f) * 0.5f + 0.5f) * 255.0f;
      float* o = rgb + (y * w + x) * 3;
      o[0] = r / 255.0f; o[1] = g / 255.0f; o[2] = b / 255.0f;
    }
  return 0;
}

// Kahan-compensated f64 channel sum (bm3d.rs:778-789).
double oracle_refined_sum(const float* rgb, size_t n, int channel) {
  double sum = 0.0, c = 0.0;
  for (size_t i = 0; i < n; ++i) {
    double x = (double)rgb[i * 3 + channel];
    double y = x - c;
    double t = sum + y;
    c = (t - sum) - y;
    sum = t;
  }
  return sum;
}

// Film simulation; trace planes (latent, image dye, mask dye: E*w*h each as f64)
// are written when the pointers are non-null.
// ext (may be null: the reference's behaviour) = {bins, stock_seed, philox, grain_sigma,
// halation_radius, halation_strength} as doubles.
int oracle_film(float* rgb, size_t w, size_t h, int stock, int format, unsigned long long seed, int output,
                float shutter, int fp64, double* latent, double* image_dye, double* mask_dye, float* mid_dmin6,
                const double* ext) {
  FilmTrace tr;
  FilmExt x;
  if (ext) {
    x.bins = (int)ext[0];
    x.stock_seed = (uint64_t)ext[1];
    x.philox = ext[2] != 0.0;
    x.grain_sigma = (float)ext[3];
    x.hal_radius = (float)ext[4];
    x.hal_strength = (float)ext[5];
  }
  const bool want = latent || image_dye || mask_dye || mid_dmin6;
  int rc = fp64 ? film_process<double>(rgb, w, h, stock, format, seed, output, shutter, want ? &tr : nullptr, x)
                : film_process<float>(rgb, w, h, stock, format, seed, output, shutter, want ? &tr : nullptr, x);
  if (rc != 0) return rc;
  if (latent) std::copy(tr.latent.begin(), tr.latent.end(), latent);
  if (image_dye) std::copy(tr.image_dye.begin(), tr.image_dye.end(), image_dye);
  if (mask_dye) std::copy(tr.mask_dye.begin(), tr.mask_dye.end(), mask_dye);
  if (mid_dmin6)
    for (int c = 0; c < 3; ++c) { mid_dmin6[c] = tr.mid[c]; mid_dmin6[3 + c] = tr.dmin[c]; }
  return 0;
}

void oracle_film_blur(float* buf, size_t w, size_t h, float sigma) { film_blur(buf, w, h, sigma); }
void oracle_film_grain_noise(float* out, size_t w, size_t h, unsigned long long seed, size_t layer) {
  film_grain_noise(out, w, h, seed, layer);
}
void oracle_film_philox(const uint32_t* ctr, const uint32_t* key, uint32_t* out) { film_philox(ctr, key, out); }
void oracle_film_grain_noise_philox(float* out, size_t n, unsigned long long seed, size_t layer) {
  film_grain_noise_philox(out, n, seed, layer);
}
int oracle_film_stock_knots(int stock, int bins, unsigned long long stock_seed, int emulsion, double* logs,
                            double* fracs) {
  return film_stock_knots(stock, bins, stock_seed, emulsion, logs, fracs);
}
int oracle_film_stock_sensitivity(int stock, int bins, unsigned long long stock_seed, int emulsion, double* out) {
  return film_stock_sensitivity(stock, bins, stock_seed, emulsion, out);
}
void oracle_film_cmf(int bins, double* out) { film_cmf(bins, out); }
int oracle_film_capture_lut(int stock, int emulsion, double* logs, double* fracs, double* params) {
  return film_capture_lut(stock, emulsion, logs, fracs, params);
}
double oracle_film_lut_sample(int stock, int emulsion, double phi) { return film_lut_sample(stock, emulsion, phi); }
void oracle_film_upsample(const float* rgb, double* out) { film_upsample(rgb, out); }
void oracle_film_spectrum_to_acescg(const double* s, double* out) { film_spectrum_to_acescg(s, out); }
int oracle_film_normalized_dmin(int stock, float* out) { return film_normalized_dmin(stock, out); }

// Differentiable film model: params = [n_emulsions, n_bins, n_knots, sens (8x81),
// knot fractions (8x8), d_max (8), gamma (8), hal_strength, hal_radius] as doubles
// (733 values); the model = (stock, bins, stock_seed, pitch).
static constexpr int kFitSens = 3, kFitKnots = kFitSens + 8 * 81, kFitDmax = kFitKnots + 64, kFitGamma = kFitDmax + 8,
                     kFitHal = kFitGamma + 8, kFitN = kFitHal + 2;
static FilmFitParams unpack_fit(const double* p) {
  FilmFitParams f{};
  f.n_emulsions = (int)p[0];
  f.n_bins = (int)p[1];
  f.n_knots = (int)p[2];
  for (int e = 0; e < 8; ++e) {
    for (int i = 0; i < 81; ++i) f.sens[e][i] = p[kFitSens + e * 81 + i];
    for (int k = 0; k < 8; ++k) f.knot_frac[e][k] = p[kFitKnots + e * 8 + k];
    f.d_max[e] = p[kFitDmax + e];
    f.gamma[e] = p[kFitGamma + e];
  }
  f.hal_strength = p[kFitHal];
  f.hal_radius = p[kFitHal + 1];
  return f;
}
static void pack_fit(const FilmFitParams& f, double* p) {
  for (int i = 0; i < kFitN; ++i) p[i] = 0.0;
  p[0] = f.n_emulsions;
  p[1] = f.n_bins;
  p[2] = f.n_knots;
  for (int e = 0; e < 8; ++e) {
    for (int i = 0; i < 81; ++i) p[kFitSens + e * 81 + i] = f.sens[e][i];
    for (int k = 0; k < 8; ++k) p[kFitKnots + e * 8 + k] = f.knot_frac[e][k];
    p[kFitDmax + e] = f.d_max[e];
    p[kFitGamma + e] = f.gamma[e];
  }
  p[kFitHal] = f.hal_strength;
  p[kFitHal + 1] = f.hal_radius;
}
static FilmFitModel fit_model(int stock, int bins, uint64_t seed, float pitch) {
  FilmFitModel m;
  m.stock = stock;
  m.bins = bins;
  m.stock_seed = seed;
  m.pitch_um = pitch;
  return m;
}
int oracle_film_fit_nparams() { return kFitN; }
int oracle_film_fit(float* rgb, size_t w, size_t h, int stock, int bins, uint64_t seed, float pitch_um,
                    const double* params, int fp64) {
  const FilmFitParams f = unpack_fit(params);
  const FilmFitModel m = fit_model(stock, bins, seed, pitch_um);
  return fp64 ? film_fit_process<double>(rgb, w, h, m, f) : film_fit_process<float>(rgb, w, h, m, f);
}
int oracle_film_fit_nominal(int stock, int bins, uint64_t seed, double* params) {
  FilmFitParams f;
  const int rc = film_fit_nominal(fit_model(stock, bins, seed, 1.0f), &f);
  if (rc) return rc;
  pack_fit(f, params);
  return 0;
}
int oracle_film_fit_gradient(const float* rgb, const float* gout, size_t w, size_t h, int stock, int bins,
                             uint64_t seed, float pitch_um, const double* params, double* grad, float* out_rgb) {
  FilmFitParams g;
  const int rc = film_fit_gradient(rgb, gout, w, h, fit_model(stock, bins, seed, pitch_um), unpack_fit(params), &g,
                                   out_rgb);
  if (rc) return rc;
  pack_fit(g, grad);
  return 0;
}

}  // extern "C"
