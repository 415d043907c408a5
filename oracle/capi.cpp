// Oracle (test infrastructure only): C entry points loaded by the tests, the
// smoke check and the benchmark's CPU-baseline leg through ctypes.
#include <cstdint>
#include <vector>

#include "oracle_colour.hpp"

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

int oracle_markesteijn(const float* in, size_t w, size_t h, const char* cfa_pattern, size_t crop_x, size_t crop_y,
                       float* out) {
  Cfa base;
  if (!base.parse(cfa_pattern)) return -1;
  Cfa cfa = base.shift(crop_x, crop_y);
  markesteijn(in, w, h, cfa, out);
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

}  // extern "C"
