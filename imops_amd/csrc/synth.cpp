// Deterministic synthetic raw generator.
//
// Produces the seeded test scenes the benchmarks and parity tests run on:
// smooth per-channel gradients, 40 flat rectangles painted in order and three
// oriented sinusoids per channel, sampled through a CFA and pushed through a
// Poisson shot-noise + Gaussian read-noise sensor model, quantised to u16.
//
// Every random draw comes from SplitMix64 (the generator the film grain uses,
// pichromatic/src/film/development/grain.rs:20-45).  Scene parameters are drawn
// sequentially from the stream keyed 0x5EED0000 + frame; per-pixel noise uses
// the counter form of the same generator (output n of a stream is a pure
// function of key + n * golden), so rows can be generated in parallel and the
// bytes do not depend on the thread count.
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>

#include "imops_gpu.h"
#include "cfa_table.hpp"

namespace {

constexpr uint64_t kGolden = 0x9E3779B97F4A7C15ull;

inline uint64_t mix64(uint64_t z) {
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

struct Stream {
  uint64_t state;
  uint64_t next() { state += kGolden; return mix64(state); }
  // uniform in [0,1) with 53 random bits
  double uniform() { return (double)(next() >> 11) * (1.0 / 9007199254740992.0); }
  double uniform(double a, double b) { return a + (b - a) * uniform(); }
};

// n-th output (n >= 0) of the stream whose initial state is key.
inline uint64_t counter_u64(uint64_t key, uint64_t n) { return mix64(key + (n + 1) * kGolden); }
inline double counter_uniform(uint64_t key, uint64_t n) {
  return (double)(counter_u64(key, n) >> 11) * (1.0 / 9007199254740992.0);
}

struct Rect { double x0, y0, x1, y1; double level[3]; };
struct Sine { double kx, ky, phase, amp; };

struct Scene {
  double grad[3][3];  // a, b, c per channel
  Rect rects[40];
  Sine sines[3][3];
  size_t w, h;

  Scene(size_t w_, size_t h_, uint64_t frame) : w(w_), h(h_) {
    Stream s{0x5EED0000ull + frame};
    for (int c = 0; c < 3; ++c)
      for (int k = 0; k < 3; ++k) grad[c][k] = s.uniform(0.05, 0.5);
    for (auto& r : rects) {
      double rw = s.uniform((double)w / 50.0, (double)w / 5.0);
      double rh = s.uniform((double)h / 50.0, (double)h / 5.0);
      r.x0 = s.uniform(0.0, (double)w - rw);
      r.y0 = s.uniform(0.0, (double)h - rh);
      r.x1 = r.x0 + rw;
      r.y1 = r.y0 + rh;
      for (int c = 0; c < 3; ++c) r.level[c] = s.uniform(0.02, 0.9);
    }
    const double two_pi = 6.283185307179586;
    for (int c = 0; c < 3; ++c)
      for (int k = 0; k < 3; ++k) {
        double period = s.uniform(2.0, 32.0);
        double angle = s.uniform(0.0, 3.141592653589793);
        double phase = s.uniform(0.0, two_pi);
        double amp = s.uniform(0.02, 0.08);
        sines[c][k] = {two_pi * std::cos(angle) / period, two_pi * std::sin(angle) / period, phase, amp};
      }
  }

  double value(size_t x, size_t y, int c) const {
    double v = -1.0;
    for (int i = 39; i >= 0; --i) {  // the last painted rectangle wins
      const Rect& r = rects[i];
      if ((double)x >= r.x0 && (double)x < r.x1 && (double)y >= r.y0 && (double)y < r.y1) {
        v = r.level[c];
        break;
      }
    }
    if (v < 0.0) v = grad[c][0] + grad[c][1] * (double)x / (double)w + grad[c][2] * (double)y / (double)h;
    for (int k = 0; k < 3; ++k) {
      const Sine& sn = sines[c][k];
      v += sn.amp * std::sin(sn.kx * (double)x + sn.ky * (double)y + sn.phase);
    }
    return v < 0.0 ? 0.0 : (v > 1.0 ? 1.0 : v);
  }
};

// Poisson(lambda) from counter uniforms: Knuth's product method below 20,
// a rounded normal approximation above.
double poisson(double lambda, uint64_t key, uint64_t base) {
  if (lambda <= 0.0) return 0.0;
  if (lambda < 20.0) {
    double limit = std::exp(-lambda), p = 1.0;
    int k = 0;
    for (;;) {
      p *= counter_uniform(key, base + 4 + (uint64_t)k);
      if (p <= limit || k > 200) break;
      ++k;
    }
    return (double)k;
  }
  double u1 = counter_uniform(key, base), u2 = counter_uniform(key, base + 1);
  if (u1 < 1e-300) u1 = 1e-300;
  double z = std::sqrt(-2.0 * std::log(u1)) * std::cos(6.283185307179586 * u2);
  double e = std::floor(lambda + std::sqrt(lambda) * z + 0.5);
  return e < 0.0 ? 0.0 : e;
}

}  // namespace

extern "C" int imops_synth_scene_rgb(float* out, size_t width, size_t height, uint64_t frame) {
  if (!out || width == 0 || height == 0) return IMOPS_E_INVALID;
  Scene sc(width, height, frame);
#pragma omp parallel for schedule(static)
  for (long long y = 0; y < (long long)height; ++y)
    for (size_t x = 0; x < width; ++x)
      for (int c = 0; c < 3; ++c) out[((size_t)y * width + x) * 3 + c] = (float)sc.value(x, (size_t)y, c);
  return IMOPS_OK;
}

extern "C" int imops_synth_raw_u16(uint16_t* out, size_t width, size_t height, const char* cfa_pattern,
                                   uint64_t frame, float full_well, float black, float white,
                                   float read_sigma) {
  if (!out || width == 0 || height == 0 || !cfa_pattern) return IMOPS_E_INVALID;
  imops::CfaTable cfa;
  if (!cfa.parse(cfa_pattern)) return IMOPS_E_INVALID;
  if (!(white > black) || !(full_well > 0.0f)) return IMOPS_E_INVALID;
  Scene sc(width, height, frame);
  const uint64_t noise_key = mix64(0xA11CE5EEDull ^ (0x5EED0000ull + frame));
  const double gain = ((double)white - (double)black) / (double)full_well;
#pragma omp parallel for schedule(static)
  for (long long yy = 0; yy < (long long)height; ++yy) {
    size_t y = (size_t)yy;
    for (size_t x = 0; x < width; ++x) {
      int c = cfa.color_at(y, x);
      if (c > 2) c = 1;
      double s = sc.value(x, y, c);
      uint64_t base = ((uint64_t)y * width + x) * 256ull;
      double e = poisson(s * (double)full_well, noise_key, base);
      double u1 = counter_uniform(noise_key, base + 2), u2 = counter_uniform(noise_key, base + 3);
      if (u1 < 1e-300) u1 = 1e-300;
      double z = std::sqrt(-2.0 * std::log(u1)) * std::cos(6.283185307179586 * u2);
      double dn = (double)black + e * gain + (double)read_sigma * z;
      dn = std::floor(dn + 0.5);
      if (dn < 0.0) dn = 0.0;
      if (dn > (double)white) dn = (double)white;
      out[y * width + x] = (uint16_t)dn;
    }
  }
  return IMOPS_OK;
}
