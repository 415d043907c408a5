// Oracle (test infrastructure only): the display-tail stages restated in f32
// from pichromatic/src:
//   contrast.rs:4-8 (Contrast), highlight_reconstruction.rs:5-14,
//   lch.rs:5-21 (LCH), tone_map.rs:7-170 (SigmoidToneMap with the ACES
//   reference gamut compression), vignette.rs:3-99 (FixVignetteRadial), and
//   BaselineExposureCompensation (pichromatic-pipeline/src/modules.rs:335-371,
//   radiance.rs:19-31).
// Oklab/Oklch come from the third-party `color` crate (0.3.x, absent here):
// restated from Ottosson's published Oklab matrices, converted through linear
// sRGB as that crate converts, hue in degrees wrapped to [0, 360); parity
// unpinned beyond the published Oklab values the tests check.
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>

#include "oracle_colour.hpp"

namespace oracle {
namespace {

constexpr float kPi = 3.14159265358979f;

void lin_to_oklch(const float in[3], float out[3]) {
  const float r = in[0], g = in[1], b = in[2];
  const float l = 0.4122214708f * r + 0.5363325363f * g + 0.0514459929f * b;
  const float m = 0.2119034982f * r + 0.6806995451f * g + 0.1073969566f * b;
  const float s = 0.0883024619f * r + 0.2817188376f * g + 0.6299787005f * b;
  const float l_ = std::cbrt(l), m_ = std::cbrt(m), s_ = std::cbrt(s);
  const float L = 0.2104542553f * l_ + 0.7936177850f * m_ - 0.0040720468f * s_;
  const float A = 1.9779984951f * l_ - 2.4285922050f * m_ + 0.4505937099f * s_;
  const float B = 0.0259040371f * l_ + 0.7827717662f * m_ - 0.8086757660f * s_;
  float h = std::atan2(B, A) * (180.0f / kPi);
  if (h < 0.0f) h += 360.0f;
  out[0] = L;
  out[1] = std::sqrt(A * A + B * B);
  out[2] = h;
}

void oklch_to_lin(const float in[3], float out[3]) {
  const float hr = in[2] * (kPi / 180.0f);
  const float A = in[1] * std::cos(hr), B = in[1] * std::sin(hr), L = in[0];
  const float l_ = L + 0.3963377774f * A + 0.2158037573f * B;
  const float m_ = L - 0.1055613458f * A - 0.0638541728f * B;
  const float s_ = L - 0.0894841775f * A - 1.2914855480f * B;
  const float l = l_ * l_ * l_, m = m_ * m_ * m_, s = s_ * s_ * s_;
  out[0] = 4.0767416621f * l - 3.3077115913f * m + 0.2309699292f * s;
  out[1] = -1.2684380046f * l + 2.6097574011f * m - 0.3413193965f * s;
  out[2] = -0.0041960863f * l - 0.7034186147f * m + 1.7076147010f * s;
}

// compress_value (tone_map.rs:75-130), forward direction
float rgc_compress(float dist, float lim, float thr, float power) {
  if (dist < thr) return dist;
  const float base_inner = (1.0f - thr) / (lim - thr);
  if (base_inner <= 0.0f || std::fabs(lim - thr) < 1e-6f) return dist;
  const float denom_inner = std::pow(base_inner, -power) - 1.0f;
  if (denom_inner <= 0.0f) return dist;
  const float s = (lim - thr) / std::pow(denom_inner, 1.0f / power);
  const float dn = (dist - thr) / s;
  const float den = std::pow(1.0f + std::pow(dn, power), 1.0f / power);
  return thr + s * (dn / den);
}

float square_sigmoid(float x) {
  const float c = 1.219512f;
  if (x <= 0.0f) return 0.0f;
  return 1.0f / (1.0f + (1.0f / (c * x)));
}

}  // namespace

void oklch_of(const float lin[3], float lch_out[3]) { lin_to_oklch(lin, lch_out); }

void rgc(const float in[3], float out[3]) {  // gamut_compress_pixel with RgcParams::default
  const float ach = std::fmax(std::fmax(in[0], in[1]), in[2]);
  if (ach == 0.0f) {
    out[0] = out[1] = out[2] = 0.0f;
    return;
  }
  const float aa = std::fabs(ach);
  const float thr[3] = {0.815f, 0.803f, 0.88f}, lim[3] = {1.147f, 1.264f, 1.312f};
  for (int c = 0; c < 3; ++c) out[c] = ach - rgc_compress((ach - in[c]) / aa, lim[c], thr[c], 1.2f) * aa;
}

float rgc_curve(float dist, int channel) {
  const float thr[3] = {0.815f, 0.803f, 0.88f}, lim[3] = {1.147f, 1.264f, 1.312f};
  return rgc_compress(dist, lim[channel], thr[channel], 1.2f);
}

void display_contrast(float* rgb, size_t n, float c) {
  for (size_t i = 0; i < 3 * n; ++i) rgb[i] = 0.185f * std::pow(rgb[i] / 0.185f, c);
}

void display_highlight(float* rgb, size_t n, float clip_g) {
  for (size_t i = 0; i < n; ++i) {
    float* p = rgb + 3 * i;
    const float f = p[1] / clip_g;
    p[1] = ((1.0f - f) * p[1]) + (f * (p[0] + p[2]) * (1.0f / 2.0f));
  }
}

void display_lch(float* rgb, size_t n, int cs, float lc, float cc, float hc) {
  for (size_t i = 0; i < n; ++i) {
    float lin[3], lch[3];
    convert(cs, CS_LINEAR_SRGB, rgb + 3 * i, lin);
    lin_to_oklch(lin, lch);
    const float scaled[3] = {lch[0] * lc, lch[1] * cc, lch[2] * hc};
    oklch_to_lin(scaled, lin);
    convert(CS_LINEAR_SRGB, cs, lin, rgb + 3 * i);
  }
}

void display_sigmoid(float* rgb, size_t n) {
  for (size_t i = 0; i < n; ++i) {
    float* px = rgb + 3 * i;
    float gc[3], lin[3], lch_gc[3], p[3], lch_p[3];
    rgc(px, gc);
    convert(CS_ACESCG, CS_LINEAR_SRGB, gc, lin);
    lin_to_oklch(lin, lch_gc);
    for (int c = 0; c < 3; ++c) p[c] = square_sigmoid(gc[c]);
    const float s = 0.2722287168f * p[0] + 0.6740817658f * p[1] + 0.0536895174f * p[2];
    const float m = 1.0f - std::pow(s, 5.0f);
    convert(CS_ACESCG, CS_LINEAR_SRGB, p, lin);
    lin_to_oklch(lin, lch_p);
    const float out_lch[3] = {lch_p[0], lch_p[1] * m, lch_gc[2]};
    oklch_to_lin(out_lch, lin);
    convert(CS_LINEAR_SRGB, CS_ACESCG, lin, px);
  }
}

void display_baseline(float* rgb, size_t n, float ev, float t, float fn, float iso) {
  if (std::fabs(ev) > 1e-6f) exp_ev(rgb, n, ev);
  if (t > 0.0f && fn > 0.0f && iso > 0.0f) {
    const float g = (float)(12.5 * 1.0 * ((double)fn * (double)fn) / ((double)t * (double)iso));
    for (size_t i = 0; i < 3 * n; ++i) rgb[i] = rgb[i] * g;
  }
}

void display_vignette(float* rgb, size_t width, size_t height, const uint8_t* op, size_t len, float strength) {
  if (len < 4) return;
  auto u32 = [&](size_t o) {
    return ((uint32_t)op[o] << 24) | ((uint32_t)op[o + 1] << 16) | ((uint32_t)op[o + 2] << 8) | (uint32_t)op[o + 3];
  };
  auto f64 = [&](size_t o) {
    uint64_t b = 0;
    for (int i = 0; i < 8; ++i) b = (b << 8) | op[o + i];
    double d;
    std::memcpy(&d, &b, sizeof d);
    return d;
  };
  const size_t count = u32(0);
  size_t off = 4;
  for (size_t c = 0; c < count; ++c) {
    if (off + 16 > len) break;
    const uint32_t id = u32(off);
    const size_t ps = u32(off + 12);
    off += 16;
    if (off + ps > len) break;
    const size_t p = off;
    off += ps;
    if (id != 3 || ps < 56) continue;
    float k[5];
    for (int j = 0; j < 5; ++j) k[j] = (float)f64(p + 8 * j);
    const float cx = (float)f64(p + 40), cy = (float)f64(p + 48);
    const float wf = (float)(width - 1), hf = (float)(height - 1);
    const float d00 = std::sqrt(cx * cx + cy * cy);
    const float d10 = std::sqrt((1.0f - cx) * (1.0f - cx) + cy * cy);
    const float d01 = std::sqrt(cx * cx + (1.0f - cy) * (1.0f - cy));
    const float d11 = std::sqrt((1.0f - cx) * (1.0f - cx) + (1.0f - cy) * (1.0f - cy));
    const float dmax = std::fmax(std::fmax(std::fmax(d00, d10), d01), d11);
    if (!(dmax > 0.0f)) continue;
    for (size_t i = 0; i < width * height; ++i) {
      const float py = (float)(i / width), px = (float)(i % width);
      const float u = px / wf, v = py / hf;
      const float du = u - cx, dv = v - cy;
      const float d = std::sqrt(du * du + dv * dv);
      const float r = d / dmax;
      const float r2 = r * r, r4 = r2 * r2, r6 = r4 * r2, r8 = r4 * r4, r10 = r8 * r2;
      const float corr = k[0] * r2 + k[1] * r4 + k[2] * r6 + k[3] * r8 + k[4] * r10;
      const float gain = std::fmax(1.0f + strength * corr, 0.0f);
      for (int ch = 0; ch < 3; ++ch) rgb[3 * i + ch] *= gain;
    }
  }
}

}  // namespace oracle

extern "C" {
int oracle_contrast(float* rgb, size_t n, float c) { oracle::display_contrast(rgb, n, c); return 0; }
int oracle_highlight(float* rgb, size_t n, float clip_g) { oracle::display_highlight(rgb, n, clip_g); return 0; }
int oracle_lch(float* rgb, size_t n, int cs, float lc, float cc, float hc) {
  oracle::display_lch(rgb, n, cs, lc, cc, hc);
  return 0;
}
int oracle_sigmoid(float* rgb, size_t n) { oracle::display_sigmoid(rgb, n); return 0; }
int oracle_baseline(float* rgb, size_t n, float ev, float t, float fn, float iso) {
  oracle::display_baseline(rgb, n, ev, t, fn, iso);
  return 0;
}
int oracle_vignette(float* rgb, size_t w, size_t h, const uint8_t* op, size_t len, float strength) {
  oracle::display_vignette(rgb, w, h, op, len, strength);
  return 0;
}
int oracle_oklch(const float* lin, float* lch) {
  oracle::oklch_of(lin, lch);
  return 0;
}
int oracle_rgc(const float* in, float* out) { oracle::rgc(in, out); return 0; }
float oracle_rgc_curve(float dist, int channel) { return oracle::rgc_curve(dist, channel); }
}
