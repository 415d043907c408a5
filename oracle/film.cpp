// Oracle (test infrastructure only): the spectral film simulation restated
// from pichromatic/src/film.
//
//   process                      film/mod.rs:77-129, mid_negative_acescg 136-163
//   FilmFormat::pixel_pitch_um   film/types.rs:45-48
//   expose_with_pitch_and_shutter film/exposure/mod.rs:45-137 (upsample_acescg
//                                upsample.rs:114-170, absorb_stack absorption.rs:24-108)
//   apply_spatial_exposure_effects film/exposure/halation.rs:24-181
//   DevelopableFractionLut       film/exposure/capture.rs:37-106; reciprocity radiance.rs:43-49
//   develop                      film/development/mod.rs:18-48, reduction.rs:16-68,
//                                diffusion.rs:11-84, grain.rs:20-100
//   scan                         film/scan/densitometry.rs:13-117, invert.rs:36-125
//   gaussian_blur_separable      film/blur.rs:10-89
//   FilmStock::finalize          film/stock/mod.rs:152-214
//
// `P` is the type of the planes the reference keeps in f32 (float: the
// reference's rounding; double: the FP64 oracle).  The reference's f64 islands
// (spectral exposure, LUT, scan) are f64 in both.
//
// Extensions (the film stage's optional fields): an 81-bin spectral grid
// 380..780 nm (CIE 1931 CMFs from the multi-lobe Gaussian fit of Wyman, Sloan
// and Shirley, JCGT 2(2) 2013), the seeded synthetic stock (ColorNeg200's stack
// with two-Gaussian sensitivities and 5-knot characteristic curves), Philox4x32-10
// grain noise (Salmon et al., SC 2011) with a constant density sigma, and the
// halation radius / strength mapping.
//
// Stock data: oracle_stock_tables.hpp, generated from the reference's
// film/stock/*.rs by tests/golden/make_stock_tables.py (not the GPU library's
// records).
//
// Pins: grain noise = SplitMix64 row streams (Python-integer brute force plus
// the published test vector); Philox = the Random123 known-answer vectors;
// blur = reflect-101 brute force; capture LUT = adaptive quadrature;
// upsample/scan round trip; the CMF fit against the reference's 20 nm table;
// the reference's end-to-end film test gates.  The per-stock absolute outputs
// are otherwise pinned only by those gates (the reference's own tests gate them
// loosely).
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <vector>

#include "oracle_common.hpp"
#include "oracle_film.hpp"
#include "oracle_stock_tables.hpp"

namespace oracle {

namespace {

constexpr int NB = 16;       // the reference's grid (the differentiable model's)
constexpr int NBMAX = 81;

struct Grid {
  int n;
  double l0, dl;
  double lam(int i) const { return l0 + dl * i; }
};
const Grid kG16{16, 400.0, 20.0};  // WavelengthGrid::mvp (spectrum.rs:17-26)
const Grid kG81{81, 380.0, 5.0};
const Grid& grid_for(int bins) { return bins == 81 ? kG81 : kG16; }

// The reference rounds its f64 islands to f32 where it stores planes; the FP64
// oracle keeps the f64 value.
template <class P> inline P store(double v) { return (P)(float)v; }
template <> inline double store<double>(double v) { return v; }

double lambda_nm(int i) { return kG16.lam(i); }

double integrate(const double* s, const Grid& g) {  // SpectralCurve::integrate (spectrum.rs:106-118)
  double acc = 0.0;
  for (int i = 0; i < g.n - 1; ++i) acc += 0.5 * (s[i] + s[i + 1]) * (g.lam(i + 1) - g.lam(i));
  return acc;
}

double evaluate(const double* s, const Grid& g, double l) {  // SpectralCurve::evaluate (spectrum.rs:60-78)
  if (l <= g.lam(0)) return s[0];
  if (l >= g.lam(g.n - 1)) return s[g.n - 1];
  for (int i = 0; i < g.n - 1; ++i)
    if (l >= g.lam(i) && l <= g.lam(i + 1)) {
      const double t = (l - g.lam(i)) / (g.lam(i + 1) - g.lam(i));
      return s[i] * (1.0 - t) + s[i + 1] * t;
    }
  return s[g.n - 1];
}

const double XBAR[16] = {0.01431, 0.13438, 0.34828, 0.29080, 0.09564, 0.00490, 0.06327, 0.29040,
                         0.59450, 0.91630, 1.06220, 0.85445, 0.44790, 0.16490, 0.04677, 0.01136};
const double YBAR[16] = {0.000396, 0.004000, 0.023000, 0.060000, 0.139020, 0.323000, 0.710000, 0.954000,
                         0.995000, 0.870000, 0.631000, 0.381000, 0.175000, 0.061000, 0.017000, 0.004100};
const double ZBAR[16] = {0.06785, 0.64560, 1.74706, 1.66920, 0.81295, 0.27200, 0.07825, 0.02030,
                         0.00390, 0.00165, 0.00080, 0.00019, 0.00002, 0.00000, 0.00000, 0.00000};
const double XYZ2AP1[3][3] = {{1.6410233797, -0.3248032942, -0.2364246952},
                              {-0.6636628587, 1.6153315917, 0.0167563477},
                              {0.0117218943, -0.0082844165, 0.9883948585}};

// CIE 1931 CMFs on a grid: the reference's table (upsample.rs:17-30) on its grid,
// else the Wyman-Sloan-Shirley multi-lobe fit: sums of Gaussian lobes whose
// width is s1 below the centre and s2 above it.
struct Cmf {
  double c[3][NBMAX];
  explicit Cmf(const Grid& g) {
    auto lobe = [](double l, double m, double s1, double s2) {
      const double t = (l - m) / (l < m ? s1 : s2);
      return std::exp(-0.5 * t * t);
    };
    for (int i = 0; i < g.n; ++i) {
      if (g.n == 16) {
        c[0][i] = XBAR[i];
        c[1][i] = YBAR[i];
        c[2][i] = ZBAR[i];
        continue;
      }
      const double l = g.lam(i);
      c[0][i] = 1.056 * lobe(l, 599.8, 37.9, 31.0) + 0.362 * lobe(l, 442.0, 16.0, 26.7) -
                0.065 * lobe(l, 501.1, 20.4, 26.2);
      c[1][i] = 0.821 * lobe(l, 568.8, 46.9, 40.5) + 0.286 * lobe(l, 530.9, 16.3, 31.1);
      c[2][i] = 1.217 * lobe(l, 437.0, 11.8, 36.0) + 0.681 * lobe(l, 459.0, 26.0, 13.8);
    }
  }
};
const Cmf& cmf_for(const Grid& g) {
  static const Cmf c16(kG16), c81(kG81);
  return g.n == 81 ? c81 : c16;
}

double integrate_cmf(const double* s, const double* cmf, const Grid& g) {  // upsample.rs:85-95
  double acc = 0.0;
  for (int i = 0; i < g.n - 1; ++i) {
    const double a = s[i] * cmf[i], b = s[i + 1] * cmf[i + 1];
    acc += 0.5 * (a + b) * g.dl;
  }
  return acc;
}

void spectrum_to_acescg(const double* s, const Grid& g, double out[3]) {
  const Cmf& C = cmf_for(g);
  const double x = integrate_cmf(s, C.c[0], g), y = integrate_cmf(s, C.c[1], g), z = integrate_cmf(s, C.c[2], g);
  for (int i = 0; i < 3; ++i) out[i] = XYZ2AP1[i][0] * x + XYZ2AP1[i][1] * y + XYZ2AP1[i][2] * z;
}

void inv3(const double m[3][3], double o[3][3]) {  // upsample.rs:56-79
  const double det = m[0][0] * (m[1][1] * m[2][2] - m[1][2] * m[2][1]) - m[0][1] * (m[1][0] * m[2][2] - m[1][2] * m[2][0]) +
                     m[0][2] * (m[1][0] * m[2][1] - m[1][1] * m[2][0]);
  const double id = 1.0 / det;
  o[0][0] = (m[1][1] * m[2][2] - m[1][2] * m[2][1]) * id;
  o[0][1] = (m[0][2] * m[2][1] - m[0][1] * m[2][2]) * id;
  o[0][2] = (m[0][1] * m[1][2] - m[0][2] * m[1][1]) * id;
  o[1][0] = (m[1][2] * m[2][0] - m[1][0] * m[2][2]) * id;
  o[1][1] = (m[0][0] * m[2][2] - m[0][2] * m[2][0]) * id;
  o[1][2] = (m[0][2] * m[1][0] - m[0][0] * m[1][2]) * id;
  o[2][0] = (m[1][0] * m[2][1] - m[1][1] * m[2][0]) * id;
  o[2][1] = (m[0][1] * m[2][0] - m[0][0] * m[2][1]) * id;
  o[2][2] = (m[0][0] * m[1][1] - m[0][1] * m[1][0]) * id;
}

struct Basis {  // UpsampleBasis::new (upsample.rs:123-143) on a grid
  double b[3][NBMAX];
  double minv[3][3];
  explicit Basis(const Grid& g) {
    const double peaks[3] = {450.0, 550.0, 650.0};
    double cols[3][3];
    for (int k = 0; k < 3; ++k) {
      for (int i = 0; i < g.n; ++i) {
        const double d = (g.lam(i) - peaks[k]) / 40.0;
        b[k][i] = std::exp(-0.5 * d * d);
      }
      spectrum_to_acescg(b[k], g, cols[k]);
    }
    double m[3][3];
    for (int r = 0; r < 3; ++r)
      for (int k = 0; k < 3; ++k) m[r][k] = cols[k][r];
    inv3(m, minv);
  }
};

const Basis& basis(const Grid& g) {
  static const Basis b16(kG16), b81(kG81);
  return g.n == 81 ? b81 : b16;
}

// Capture curve: the reference's DevelopableFractionLut (capture.rs:37-88) with
// 64 knots, or a characteristic curve given by knots (the synthetic stock); both
// sampled linearly in log10 fluence, clamped at the ends.
struct Lut {
  int n = 64;
  double logs[64], fracs[64];
  void build(double mu_ln, double sigma_ln, double k) {
    n = 64;
    for (int i = 0; i < 64; ++i) {
      const double t = (double)i / 63.0;
      const double log_phi = -4.0 + t * (8.0 - (-4.0));
      const double phi = std::pow(10.0, log_phi);
      // expected_survival (capture.rs:91-106)
      const double dz = (2.0 * 8.0) / 64.0;
      const double inv_sqrt_2pi = 1.0 / std::sqrt(2.0 * 3.14159265358979323846);
      double acc = 0.0;
      for (int j = 0; j <= 64; ++j) {
        const double z = -8.0 + (double)j * dz;
        const double w = (j == 0 || j == 64) ? 0.5 : 1.0;
        const double pdf = inv_sqrt_2pi * std::exp(-0.5 * z * z);
        const double s = std::exp(mu_ln + sigma_ln * z);
        const double g = std::exp(-k * s * s * phi);
        acc += w * pdf * g * dz;
      }
      logs[i] = log_phi;
      fracs[i] = rclamp(1.0 - acc, 0.0, 1.0);
    }
  }
  double sample(double phi) const {
    if (!(phi > 0.0)) return fracs[0];
    const double lp = std::log10(phi);
    if (lp <= logs[0]) return fracs[0];
    if (lp >= logs[n - 1]) return fracs[n - 1];
    int lo = 0, hi = n - 1;
    while (hi - lo > 1) {
      const int mid = (lo + hi) / 2;
      if (logs[mid] <= lp) lo = mid;
      else hi = mid;
    }
    const double t = (lp - logs[lo]) / (logs[hi] - logs[lo]);
    return fracs[lo] * (1.0 - t) + fracs[hi] * t;
  }
};

// ---- stock source: a stock's parameters with every curve sampled on a grid
struct SrcLayer {
  int kind;  // 0 emulsion, 1 filter, 2 overcoat, 3 antihalation, 4 support
  float thickness_um;
  bool has_sens;
  double sens[NBMAX];
  double mu_ln, sigma_ln;
  float agx;
  double eps[NBMAX], mask[NBMAX];
  bool has_mask;
  float d_max, gamma;
  double capture_k;
  float recip_p;
  bool reversal;
  int knots = 0;  // > 0: characteristic curve knots instead of the crystal LUT
  double knot_x[8], knot_f[8];
};

struct Src {
  Grid g;
  float box_iso;
  std::vector<SrcLayer> layers;
  double reflect[NBMAX], scanner[NBMAX];
  float psf_local_um, psf_halation_um, developer_diffusion_um, adjacency_beta, dir_diffusion_um;
  int n_dir;
  float dir[8][8];
};

void take(const oracle_stocks::OCurve& c, const Grid& g, double* out) {
  for (int i = 0; i < g.n; ++i) out[i] = c.present ? (g.n == 81 ? c.s81[i] : c.s16[i]) : 0.0;
}

Src src_from_tables(int idx, const Grid& g) {
  const oracle_stocks::OStock& S = oracle_stocks::kStocks[idx];
  Src r;
  r.g = g;
  r.box_iso = S.box_iso;
  for (int l = 0; l < S.n_layers; ++l) {
    const oracle_stocks::OLayer& L = S.layers[l];
    SrcLayer o{};
    o.kind = L.kind;
    o.thickness_um = L.thickness_um;
    o.has_sens = L.sensitivity.present;
    take(L.sensitivity, g, o.sens);
    o.mu_ln = L.mu_ln;
    o.sigma_ln = L.sigma_ln;
    o.agx = L.silver_halide_fraction;
    take(L.epsilon, g, o.eps);
    o.has_mask = L.has_coupler && L.mask_epsilon.present;
    take(L.mask_epsilon, g, o.mask);
    o.d_max = L.d_max;
    o.gamma = L.gamma_contrast;
    o.capture_k = L.capture_k;
    o.recip_p = L.reciprocity_p;
    o.reversal = L.is_reversal;
    r.layers.push_back(o);
  }
  take(S.ah_reflectance, g, r.reflect);
  take(S.scanner_light, g, r.scanner);
  r.psf_local_um = S.psf_local_um;
  r.psf_halation_um = S.psf_halation_um;
  r.developer_diffusion_um = S.developer_diffusion_um;
  r.adjacency_beta = S.adjacency_beta;
  r.dir_diffusion_um = S.dir_diffusion_um;
  r.n_dir = S.n_dir;
  for (int i = 0; i < 8; ++i)
    for (int j = 0; j < 8; ++j) r.dir[i][j] = S.dir[i][j];
  return r;
}

// The synthetic stock: ColorNeg200's stack; emulsion k (blue, green, red;
// base 450, 550, 650 nm) gets the sensitivity
//   exp(-(l - mu1)^2 / (2 s1^2)) + a2 exp(-(l - mu2)^2 / (2 s2^2))
// and a 5-knot characteristic curve, with no reciprocity failure.  Uniforms
// from a SplitMix64 stream started at `seed`
// (u = top 53 bits / 2^53): mu1 = base + 30 (u - 1/2), s1 = 25 + 20 u,
// mu2 = mu1 + 100 (u - 1/2), s2 = 20 + 30 u, a2 = 0.15 + 0.35 u, five knot
// positions -2.5 + j + 0.3 (u - 1/2), then four increments 0.2 + u (knots 1..4)
// whose running share c_j gives the fraction 0.005 + 0.965 c_j (c_0 = 0: the toe).
// Fourteen uniforms per emulsion.
Src src_synthetic(uint64_t seed, const Grid& g) {
  Src r = src_from_tables(1, g);  // ColorNeg200
  uint64_t state = seed;
  auto uniform = [&]() {
    state += 0x9E3779B97F4A7C15ull;
    uint64_t z = state;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    z ^= z >> 31;
    return (double)(z >> 11) / 9007199254740992.0;
  };
  const double base[3] = {450.0, 550.0, 650.0};
  int k = 0;
  for (SrcLayer& L : r.layers) {
    if (L.kind != 0) continue;
    const double mu1 = base[k] + 30.0 * (uniform() - 0.5);
    const double s1 = 25.0 + 20.0 * uniform();
    const double mu2 = mu1 + 100.0 * (uniform() - 0.5);
    const double s2 = 20.0 + 30.0 * uniform();
    const double a2 = 0.15 + 0.35 * uniform();
    for (int i = 0; i < g.n; ++i) {
      const double d1 = (g.lam(i) - mu1) / s1, d2 = (g.lam(i) - mu2) / s2;
      L.sens[i] = std::exp(-0.5 * d1 * d1) + a2 * std::exp(-0.5 * d2 * d2);
    }
    L.recip_p = 1.0f;
    L.knots = 5;
    for (int j = 0; j < 5; ++j) L.knot_x[j] = -2.5 + (double)j + 0.3 * (uniform() - 0.5);
    double inc[5] = {0.0, 0.0, 0.0, 0.0, 0.0}, total = 0.0;
    for (int j = 1; j < 5; ++j) {
      inc[j] = 0.2 + uniform();
      total += inc[j];
    }
    double run = 0.0;
    for (int j = 0; j < 5; ++j) {
      run += inc[j];
      L.knot_f[j] = 0.005 + 0.965 * (run / total);
    }
    ++k;
  }
  return r;
}

Src load_src(int idx, int bins, uint64_t stock_seed) {
  const Grid& g = grid_for(bins);
  return idx == 7 ? src_synthetic(stock_seed, g) : src_from_tables(idx, g);
}

struct Emulsion {
  int layer;
  double sens[NBMAX], eps[NBMAX], mask[NBMAX];
  bool has_mask, reversal;
  float d_max, gamma, recip_p;
  Lut lut;
  float kappa_ref;
};

struct Derived {
  Src src;
  const Src* s;
  Grid g;
  std::vector<Emulsion> em;
  double trans[12][NBMAX];  // per-layer Beer-Lambert transmittance (absorbing layers)
  bool absorbs[12], is_emulsion[12];
  double reflect[NBMAX], scanner[NBMAX];
  bool any_reversal = false;
  // differentiable-model switches (FilmFitParams); defaults reproduce film::process
  double hal_strength = 1.0, hal_radius = 1.0, scan_scale = 0.0;  // scan_scale <= 0: from this stock's Dmin
  bool grain = true, dir = true, adjacency = true;
  FilmExt ext;  // extension parameters (defaults: the reference's behaviour)
  Derived(const Derived&) = delete;  // `s` points into this object
  Derived& operator=(const Derived&) = delete;
  explicit Derived(const Src& source, const FilmFitParams* fit = nullptr) : src(source), s(&src), g(source.g) {
    const Src* st = &src;
    for (int i = 0; i < g.n; ++i) {
      reflect[i] = st->reflect[i];
      scanner[i] = st->scanner[i];
    }
    int ei = 0;
    for (size_t l = 0; l < st->layers.size(); ++l) {
      const SrcLayer& L = st->layers[l];
      if (L.reversal) any_reversal = true;
      double curve[NBMAX];
      for (int i = 0; i < g.n; ++i) curve[i] = L.sens[i];
      if (fit && L.kind == 0 && L.has_sens)
        for (int i = 0; i < g.n; ++i) curve[i] = fit->sens[ei][i];
      is_emulsion[l] = L.kind == 0;
      absorbs[l] = L.kind != 4 && L.has_sens;
      const double thickness = (double)L.thickness_um;
      for (int i = 0; i < g.n; ++i) {
        double od;
        if (L.kind == 0) {
          const double sigma = curve[i] * 2.0;  // ABSORPTION_SIGMA_SCALE_PER_UM (constants.rs:51)
          od = sigma * (double)L.agx * thickness;
        } else {
          od = curve[i] * thickness;
        }
        trans[l][i] = std::exp(-od);
      }
      if (L.kind != 0) continue;
      Emulsion e;
      e.layer = (int)l;
      for (int i = 0; i < g.n; ++i) {
        e.sens[i] = curve[i];
        e.eps[i] = L.eps[i];
        e.mask[i] = L.mask[i];
      }
      e.has_mask = L.has_mask;
      e.reversal = L.reversal;
      e.d_max = fit ? (float)fit->d_max[ei] : L.d_max;
      e.gamma = fit ? (float)fit->gamma[ei] : L.gamma;
      const int fe = ei++;
      e.recip_p = L.recip_p;
      const double mu = L.mu_ln, sg = L.sigma_ln;
      if (L.knots > 0) {
        e.lut.n = L.knots;
        for (int j = 0; j < L.knots; ++j) {
          e.lut.logs[j] = L.knot_x[j];
          e.lut.fracs[j] = fit && fit->n_knots == L.knots ? fit->knot_frac[fe][j] : L.knot_f[j];
        }
      } else {
        e.lut.build(mu, sg, L.capture_k);
      }
      // grain kappa at 1 um pitch (stock/mod.rs:193-211)
      const double mean_s = std::exp(mu + 0.5 * sg * sg);
      const double r = std::max(mean_s * 0.5, 1e-6);
      const double volume = 1.04719755119659774615 * 4.0 * r * r * r;  // FRAC_PI_3 * 4 r^3
      const double rho = (double)L.agx * thickness / std::max(volume, 1e-18);
      e.kappa_ref = (float)(1.0 / std::sqrt(rho));
      em.push_back(e);
    }
  }
};

// film/blur.rs:10-89
template <class P>
void blur(std::vector<P>& buf, size_t w, size_t h, float sigma) {
  if (sigma < 1e-3f || w == 0 || h == 0) return;
  const int radius = (int)std::max(std::ceil(3.0f * sigma), 1.0f);
  std::vector<P> k(2 * radius + 1);
  const P inv_2s2 = (P)1 / ((P)2 * (P)sigma * (P)sigma);
  P sum = 0;
  for (int i = 0; i <= 2 * radius; ++i) {
    const P x = (P)i - (P)radius;
    const P v = std::exp(-x * x * inv_2s2);
    k[i] = v;
    sum += v;
  }
  for (auto& v : k) v /= sum;
  auto reflect = [](long long i, long long n) -> size_t {
    long long x = i;
    for (;;) {
      if (x < 0) x = -x;
      else if (x >= n) x = 2 * n - 2 - x;
      else return (size_t)x;
      if (n == 1) return 0;
    }
  };
  std::vector<P> tmp(buf.size());
#pragma omp parallel for schedule(static)
  for (long long yy = 0; yy < (long long)h; ++yy)
    for (size_t x = 0; x < w; ++x) {
      P acc = 0;
      for (int j = 0; j <= 2 * radius; ++j) acc += buf[yy * w + reflect((long long)x + j - radius, (long long)w)] * k[j];
      tmp[yy * w + x] = acc;
    }
#pragma omp parallel for schedule(static)
  for (long long yy = 0; yy < (long long)h; ++yy)
    for (size_t x = 0; x < w; ++x) {
      P acc = 0;
      for (int j = 0; j <= 2 * radius; ++j) acc += tmp[reflect(yy + j - radius, (long long)h) * w + x] * k[j];
      buf[yy * w + x] = acc;
    }
}

template <class P>
using Planes = std::vector<std::vector<P>>;

// exposure/mod.rs:45-122 (band mean over the grid's span: 300 nm on the
// reference grid).  Halation extension: hal_radius >= 0 px gives sigma =
// radius / 3; hal_strength >= 0 gives the gain strength * bleed_e whatever the
// backing reflectance.
template <class P>
Planes<P> expose(const float* rgb, size_t w, size_t h, const Derived& d, float pitch, float shutter) {
  const size_t n = w * h;
  const size_t E = d.em.size();
  const Src& st = *d.s;
  const Grid& g = d.g;
  const Basis& B = basis(g);
  const double span = g.lam(g.n - 1) - g.lam(0);
  Planes<P> abs(E, std::vector<P>(n));
#pragma omp parallel for schedule(static)
  for (long long p = 0; p < (long long)n; ++p) {
    const double px[3] = {(double)rgb[p * 3], (double)rgb[p * 3 + 1], (double)rgb[p * 3 + 2]};
    double wgt[3];
    for (int r = 0; r < 3; ++r) {
      wgt[r] = B.minv[r][0] * px[0] + B.minv[r][1] * px[1] + B.minv[r][2] * px[2];
      if (wgt[r] < 0.0) wgt[r] = 0.0;
    }
    double phi[NBMAX];
    for (int i = 0; i < g.n; ++i) {
      const double s = wgt[0] * B.b[0][i] + wgt[1] * B.b[1][i] + wgt[2] * B.b[2][i];
      const double e_rel = 550.0 / g.lam(i);
      phi[i] = s / e_rel;
    }
    size_t e = 0;
    for (size_t l = 0; l < st.layers.size(); ++l) {
      if (!d.absorbs[l]) continue;
      double absorbed[NBMAX];
      for (int i = 0; i < g.n; ++i) {
        const double phi_t = phi[i] * d.trans[l][i];
        absorbed[i] = phi[i] - phi_t;
        phi[i] = phi_t;
      }
      if (d.is_emulsion[l]) {
        double acc = 0.0;
        for (int i = 0; i < g.n - 1; ++i) acc += 0.5 * (absorbed[i] + absorbed[i + 1]) * g.dl;
        abs[e++][p] = store<P>(acc / span);
      }
    }
  }
  // halation.rs:24-93
  const float sigma_local = std::max(st.psf_local_um / std::max(pitch, 1e-6f), 0.0f);
  if (sigma_local >= 1e-3f) {
    for (auto& plane : abs) {
      std::vector<P> sc = plane;
      blur(sc, w, h, sigma_local);
      const P f = (P)0.18f, keep = (P)1 - f;
      for (size_t i = 0; i < n; ++i) plane[i] = keep * plane[i] + f * sc[i];
    }
  }
  const FilmExt& x = d.ext;
  const float sigma_wide =
      x.hal_radius >= 0.0f ? x.hal_radius / 3.0f
                           : (float)((double)std::max(st.psf_halation_um / std::max(pitch, 1e-6f), 0.0f) * d.hal_radius);
  bool do_wide = sigma_wide >= 1e-3f;
  if (do_wide) {
    if (x.hal_strength >= 0.0f) {
      do_wide = x.hal_strength > 0.0f;
    } else {
      float max_r = 0.0f;
      const double bands[4] = {450.0, 550.0, 650.0, 700.0};
      for (double bl : bands) max_r = std::max(max_r, (float)evaluate(d.reflect, g, bl));
      do_wide = max_r > 0.0f;
    }
  }
  if (do_wide && E > 0) {
    std::vector<P> bounce = abs[E - 1];
    blur(bounce, w, h, sigma_wide);
    for (size_t e = 0; e < E; ++e) {
      const Emulsion& em = d.em[e];
      // bleed weight from the sensitivity peak (halation.rs:102-132)
      double peak = 550.0, mx = -INFINITY;
      for (int i = 0; i < g.n; ++i)
        if (em.sens[i] > mx) { mx = em.sens[i]; peak = g.lam(i); }
      const float wb = 0.20f, wg = 0.45f, wr = 1.0f;
      float bleed;
      if (peak <= 450.0) bleed = wb;
      else if (peak <= 550.0) bleed = wb + (wg - wb) * (float)((peak - 450.0) / 100.0);
      else if (peak <= 650.0) bleed = wg + (wr - wg) * (float)((peak - 550.0) / 100.0);
      else bleed = wr;
      float gain;
      if (x.hal_strength >= 0.0f) {
        gain = x.hal_strength * bleed;
      } else {
        // effective_reflectance (halation.rs:155-161)
        float r_e;
        const double norm = integrate(em.sens, g);
        if (norm <= 1e-9) {
          r_e = (float)evaluate(d.reflect, g, 650.0);
        } else {
          double prod[NBMAX];
          for (int i = 0; i < g.n; ++i) prod[i] = d.reflect[i] * em.sens[i];
          r_e = (float)(integrate(prod, g) / norm);
        }
        gain = (float)((double)(r_e * bleed) * d.hal_strength);
      }
      if (gain <= 0.0f) continue;
      for (size_t i = 0; i < n; ++i) abs[e][i] += (P)gain * bounce[i];
    }
  }
  // capture LUT with reciprocity (exposure/mod.rs:97-115)
  for (size_t e = 0; e < E; ++e) {
    const Emulsion& em = d.em[e];
    float eta = 1.0f;
    if (!((double)shutter <= 1.0 || (double)em.recip_p >= 0.9999))
      eta = (float)std::pow(std::max((double)shutter, 1.0), (double)em.recip_p - 1.0);
    auto& plane = abs[e];
#pragma omp parallel for schedule(static)
    for (long long i = 0; i < (long long)n; ++i)
      plane[i] = store<P>(em.lut.sample((double)(P)(plane[i] * (P)eta)));
  }
  return abs;
}

// Philox4x32-10 (Salmon, Moraes, Dror, Shaw, SC 2011): per round the products
// M0 c0 and M1 c2 (32x32 -> 64 bits) are folded with the key, and the key
// advances by the Weyl increments.
void philox4x32_10(uint32_t c[4], uint32_t k0, uint32_t k1) {
  for (int round = 0; round < 10; ++round) {
    const uint64_t p0 = (uint64_t)0xD2511F53u * c[0];
    const uint64_t p1 = (uint64_t)0xCD9E8D57u * c[2];
    const uint32_t out0 = (uint32_t)(p1 >> 32) ^ c[1] ^ k0;
    const uint32_t out2 = (uint32_t)(p0 >> 32) ^ c[3] ^ k1;
    c[1] = (uint32_t)p1;
    c[3] = (uint32_t)p0;
    c[0] = out0;
    c[2] = out2;
    k0 += 0x9E3779B9u;
    k1 += 0xBB67AE85u;
  }
}

// Grain noise of the Philox extension: counter (pixel index low / high 32 bits,
// layer, 0), key (seed low / high 32 bits), a standard normal from the first
// two words by Box-Muller in f64 ((x + 1/2) / 2^32 keeps both uniforms in (0, 1)).
template <class P>
void grain_noise_philox(P* noise, size_t n, uint64_t seed, size_t layer) {
#pragma omp parallel for schedule(static)
  for (long long i = 0; i < (long long)n; ++i) {
    uint32_t c[4] = {(uint32_t)(uint64_t)i, (uint32_t)((uint64_t)i >> 32), (uint32_t)layer, 0u};
    philox4x32_10(c, (uint32_t)seed, (uint32_t)(seed >> 32));
    const double u1 = ((double)c[0] + 0.5) / 4294967296.0, u2 = ((double)c[1] + 0.5) / 4294967296.0;
    noise[i] = (P)(float)(std::sqrt(-2.0 * std::log(u1)) * std::cos(2.0 * 3.14159265358979323846 * u2));
  }
}

// Unit grain noise n(x) = sum of 12 uniforms - 6 from one SplitMix64 stream
// per row (grain.rs:20-45, 64-82).
template <class P>
void grain_noise(P* noise, size_t w, size_t h, uint64_t seed, size_t layer) {
#pragma omp parallel for schedule(static)
  for (long long y = 0; y < (long long)h; ++y) {
    uint64_t state = seed * 0xD1B54A32D192ED03ull + (uint64_t)layer * 0x9E3779B97F4A7C15ull + (uint64_t)y;
    for (size_t x = 0; x < w; ++x) {
      float acc = 0.0f;
      for (int k = 0; k < 12; ++k) {
        state += 0x9E3779B97F4A7C15ull;
        uint64_t z = state;
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        z ^= z >> 31;
        acc += (float)z / (float)UINT64_MAX;
      }
      noise[y * w + x] = (P)(acc - 6.0f);
    }
  }
}

template <class P>
struct Dyes {
  Planes<P> image, mask;
};

// development/mod.rs:18-48
template <class P>
Dyes<P> develop(const Derived& d, const Planes<P>& latent, size_t w, size_t h, uint64_t seed, float pitch) {
  const size_t n = w * h, E = d.em.size();
  const Src& st = *d.s;
  Dyes<P> dy;
  // reduce (reduction.rs:16-68)
  for (size_t e = 0; e < E; ++e) {
    const Emulsion& em = d.em[e];
    const P d_max = (P)em.d_max;
    const P inv_gamma = (P)1 / (P)std::max(em.gamma, 1e-6f);
    std::vector<P> img(n), msk(n, (P)0);
    for (size_t i = 0; i < n; ++i) {
      const P f = rclamp(latent[e][i], (P)0, (P)1);
      const P eff = em.reversal ? (P)1 - f : f;
      img[i] = d_max * std::pow(eff, inv_gamma);
    }
    if (em.has_mask) {
      const P mask_scale = d_max * (P)0.4f;
      for (size_t i = 0; i < n; ++i) msk[i] = mask_scale * ((P)1 - rclamp(img[i] / d_max, (P)0, (P)1));
    }
    dy.image.push_back(img);
    dy.mask.push_back(msk);
  }
  // DIR inhibition (diffusion.rs:39-84)
  const float sigma_dir = st.dir_diffusion_um / std::max(pitch, 1e-6f);
  if (d.dir && E > 0 && st.n_dir > 0 && sigma_dir >= 1e-3f) {
    Planes<P> inh = dy.image;
    for (auto& p : inh) blur(p, w, h, sigma_dir);
    for (size_t j = 0; j < E; ++j) {
      std::vector<P> total(n, (P)0);
      bool active = false;
      for (size_t i = 0; i < E; ++i) {
        if ((int)i < st.n_dir && (int)j < st.n_dir) {
          const float wgt = st.dir[i][j];
          if (std::fabs(wgt) > 1e-6f) {
            active = true;
            for (size_t p = 0; p < n; ++p) total[p] += (P)wgt * inh[i][p];
          }
        }
      }
      if (active)
        for (size_t p = 0; p < n; ++p) dy.image[j][p] *= std::exp(-total[p]);
    }
  }
  // adjacency (diffusion.rs:11-28)
  const float sigma_adj = st.developer_diffusion_um / std::max(pitch, 1e-6f);
  if (d.adjacency && !(std::fabs(st.adjacency_beta) < 1e-8f || sigma_adj < 1e-3f)) {
    for (auto& plane : dy.image) {
      std::vector<P> b = plane;
      blur(b, w, h, sigma_adj);
      for (size_t i = 0; i < n; ++i) plane[i] += (P)st.adjacency_beta * (plane[i] - b[i]);
    }
  }
  // grain (grain.rs:53-100)
  const float sigma_g = 3.0f / std::max(pitch, 1e-6f);
  // extension: grain_sigma >= 0 is a constant density sigma; philox: Philox normals
  const bool absolute = d.ext.grain_sigma >= 0.0f;
  for (size_t e = 0; e < (d.grain ? E : 0); ++e) {
    const P d_max = (P)d.em[e].d_max;
    const P kappa = absolute ? (P)d.ext.grain_sigma : (P)((d.em[e].kappa_ref / std::max(pitch, 1e-6f)) * 0.15f);
    if (kappa <= (P)0 || d_max <= (P)0) continue;
    std::vector<P> noise(n);
    if (d.ext.philox) grain_noise_philox(noise.data(), n, seed, e);
    else grain_noise(noise.data(), w, h, seed, e);
    blur(noise, w, h, std::max(sigma_g, 1.0f));
    double var = 0.0;
    for (size_t i = 0; i < n; ++i) var += (double)noise[i] * (double)noise[i];
    var /= (double)n;
    const P norm = var > 1e-12 ? store<P>(1.0 / std::sqrt(var)) : (P)1;
    auto& plane = dy.image[e];
    for (size_t i = 0; i < n; ++i) {
      const P dens = rclamp(plane[i], (P)0, d_max);
      const P eps_toe = (P)0.02f * d_max;
      const P taper = std::min(dens / (dens + eps_toe), (P)1);
      const P sigma_d = absolute ? (P)1 : taper * std::sqrt(std::max(dens * (d_max - dens), (P)0));
      P v = dens + kappa * sigma_d * noise[i] * norm;
      plane[i] = rclamp(v, (P)0, d_max * (P)1.05f);
    }
  }
  return dy;
}

// densitometry.rs:13-73 (before the Dmin normalisation)
void scan_pixel(const Derived& d, const double* di, const double* dm, double out[3]) {
  const Grid& g = d.g;
  double dens[NBMAX] = {0};
  for (size_t e = 0; e < d.em.size(); ++e)
    for (int l = 0; l < g.n; ++l) {
      dens[l] += di[e] * d.em[e].eps[l];
      if (d.em[e].has_mask) dens[l] += dm[e] * d.em[e].mask[l];
    }
  double t[NBMAX];
  for (int i = 0; i < g.n; ++i) t[i] = std::pow(10.0, -dens[i]) * d.scanner[i];
  spectrum_to_acescg(t, g, out);
}

void dmin_reference(const Derived& d, double out[3]) {
  double di[8] = {0}, dm[8] = {0};
  for (size_t e = 0; e < d.em.size(); ++e) dm[e] = d.em[e].has_mask ? (double)(d.em[e].d_max * 0.4f) : 0.0;
  scan_pixel(d, di, dm, out);
}

template <class P>
void scan(const Derived& d, const Dyes<P>& dy, size_t n, float* rgb) {
  double dmin[3];
  dmin_reference(d, dmin);
  const double peak = std::max(std::max(std::max(dmin[0], dmin[1]), dmin[2]), 1e-12);
  const double scale = d.scan_scale > 0.0 ? d.scan_scale : 1.0 / peak;
  const size_t E = d.em.size();
#pragma omp parallel for schedule(static)
  for (long long p = 0; p < (long long)n; ++p) {
    double di[8], dm[8], c[3];
    for (size_t e = 0; e < E; ++e) {
      di[e] = (double)dy.image[e][p];
      dm[e] = (double)dy.mask[e][p];
    }
    scan_pixel(d, di, dm, c);
    for (int k = 0; k < 3; ++k) rgb[p * 3 + k] = (float)((double)store<P>(c[k]) * scale);
  }
}

void normalized_dmin(const Derived& d, float out[3]) {
  double dm[3];
  dmin_reference(d, dm);
  const double peak = std::max(std::max(std::max(dm[0], dm[1]), dm[2]), 1e-12);
  for (int k = 0; k < 3; ++k) out[k] = (float)(dm[k] / peak);
}

// invert.rs:36-108
void invert(float* rgb, size_t n, const float mid[3], const float dmin_in[3]) {
  const float eps = 1e-6f;
  auto soft = [](float inv, float sh) {
    inv = std::fmax(inv, 0.0f);
    const float s = std::fmax(sh, 1e-6f);
    return inv / (1.0f + inv / s);
  };
  float dmin[3], mid_t[3], inv_mid[3], shoulder[3], soft_mid[3], g[3];
  for (int c = 0; c < 3; ++c) {
    dmin[c] = std::fmax(dmin_in[c], eps);
    mid_t[c] = rclamp(mid[c] / dmin[c], eps, 1.0f - eps);
    inv_mid[c] = std::fmax(1.0f / mid_t[c] - 1.0f, 0.0f);
    shoulder[c] = std::fmax(inv_mid[c] * 8.0f, eps);
    soft_mid[c] = std::fmax(soft(inv_mid[c], shoulder[c]), eps);
    g[c] = 0.185f / soft_mid[c];
  }
  for (size_t p = 0; p < n; ++p) {
    float* px = rgb + p * 3;
    float raw[3];
    for (int c = 0; c < 3; ++c) {
      const float t = rclamp(px[c] / dmin[c], eps, 1.0f);
      const float inv = std::fmax(1.0f / t - 1.0f, 0.0f);
      raw[c] = std::fmax(g[c] * soft(inv, shoulder[c]), 0.0f);
    }
    const float y = std::fmax(0.2722287168f * raw[0] + 0.6740817658f * raw[1] + 0.0536895174f * raw[2], 0.0f);
    float k;
    if (y <= 0.0f) k = 0.0f;
    else if (y >= 0.02f) k = 0.55f;
    else k = 0.55f * (y / 0.02f);
    for (int c = 0; c < 3; ++c) px[c] = std::fmax(y + k * (raw[c] - y), 0.0f);
  }
}

float format_width_mm(int format) { return format == 1 ? 56.0f : (format == 2 ? 101.6f : 36.0f); }

template <class P>
void run_negative(const Derived& d, const float* rgb_in, size_t w, size_t h, float pitch, float shutter, uint64_t seed,
                  float* rgb_out, FilmTrace* trace) {
  Planes<P> latent = expose<P>(rgb_in, w, h, d, pitch, shutter);
  if (trace) {
    trace->latent.clear();
    for (auto& p : latent) trace->latent.insert(trace->latent.end(), p.begin(), p.end());
  }
  Dyes<P> dy = develop<P>(d, latent, w, h, seed, pitch);
  if (trace) {
    trace->image_dye.clear();
    trace->mask_dye.clear();
    for (auto& p : dy.image) trace->image_dye.insert(trace->image_dye.end(), p.begin(), p.end());
    for (auto& p : dy.mask) trace->mask_dye.insert(trace->mask_dye.end(), p.begin(), p.end());
  }
  scan<P>(d, dy, w * h, rgb_out);
}

}  // namespace

bool stock_ok(int idx) { return idx >= 0 && idx <= 7; }

// film/mod.rs:77-129 (and the extension parameters of `ext`)
template <class P>
int film_process(float* rgb, size_t w, size_t h, int stock_idx, int format, uint64_t seed, int output,
                 float shutter_seconds, FilmTrace* trace, const FilmExt& ext) {
  if (!stock_ok(stock_idx) || w == 0 || h == 0 || (ext.bins != 16 && ext.bins != 81)) return -1;
  Derived d(load_src(stock_idx, ext.bins, ext.stock_seed));
  d.ext = ext;
  const float pitch = format_width_mm(format) * 1000.0f / (float)w;
  const float shutter = shutter_seconds > 0.0f ? shutter_seconds : 1.0f / d.s->box_iso;
  std::vector<float> out(w * h * 3);
  run_negative<P>(d, rgb, w, h, pitch, shutter, seed, out.data(), trace);
  if (output == 1 && !d.any_reversal) {
    float dmin[3], mid[3];
    normalized_dmin(d, dmin);
    // mid-gray negative: 32x32 flat absolute-luminance gray, seed 0 (film/mod.rs:136-163)
    const int N = 32;
    // relative_to_absolute_y with sunny-16 metering (film/mod.rs:154-163, radiance.rs:19-25, 52-59)
    const float iso = std::max(d.s->box_iso, 1.0f);
    const double t = (double)(1.0f / iso), fn = 8.0, v = (double)0.185f;
    const float gv = (float)(12.5 * v * (fn * fn) / (t * (double)iso));
    std::vector<float> gray(N * N * 3, gv), gout(N * N * 3);
    run_negative<P>(d, gray.data(), N, N, pitch, shutter, 0, gout.data(), nullptr);
    double acc[3] = {0, 0, 0};
    for (int p = 0; p < N * N; ++p)
      for (int c = 0; c < 3; ++c) acc[c] += (double)gout[p * 3 + c];
    for (int c = 0; c < 3; ++c) mid[c] = (float)(acc[c] / (double)(N * N));
    if (trace) {
      for (int c = 0; c < 3; ++c) { trace->mid[c] = mid[c]; trace->dmin[c] = dmin[c]; }
    }
    invert(out.data(), w * h, mid, dmin);
  }
  std::copy(out.begin(), out.end(), rgb);
  return 0;
}

template int film_process<float>(float*, size_t, size_t, int, int, uint64_t, int, float, FilmTrace*, const FilmExt&);
template int film_process<double>(float*, size_t, size_t, int, int, uint64_t, int, float, FilmTrace*, const FilmExt&);

// The differentiable film model: film::process (NegativeLinear, shutter 1/box
// ISO, a given pixel pitch) on the model's grid with grain, DIR inhibition and
// adjacency off, the emulsion sensitivities, D_max, gamma and (knot-curve
// stocks) the developable fraction at each knot taken from `fit`, the halation
// gain and PSF width scaled by fit.hal_strength / fit.hal_radius, and the scan
// normalised by the nominal stock's Dmin (so the normalisation is not a
// function of the parameters).
namespace {

int fit_model_check(const FilmFitModel& m, const Src& src, const FilmFitParams* fit) {
  if (!stock_ok(m.stock) || (m.bins != 16 && m.bins != 81) || !(m.pitch_um > 0.0f)) return -1;
  if (!fit) return 0;
  int E = 0, knots = 0;
  for (const SrcLayer& L : src.layers)
    if (L.kind == 0) {
      ++E;
      knots = L.knots;
    }
  if (fit->n_emulsions != E || fit->n_bins != m.bins || fit->n_knots != knots || knots > 8) return -1;
  return 0;
}

double fit_scan_scale(const Src& src) {
  const Derived nom(src);
  double dmin[3];
  dmin_reference(nom, dmin);
  return 1.0 / std::max(std::max(std::max(dmin[0], dmin[1]), dmin[2]), 1e-12);
}

}  // namespace

template <class P>
int film_fit_process(float* rgb, size_t w, size_t h, const FilmFitModel& m, const FilmFitParams& fit) {
  if (w == 0 || h == 0 || !stock_ok(m.stock)) return -1;
  const Src src = load_src(m.stock, m.bins, m.stock_seed);
  if (fit_model_check(m, src, &fit)) return -1;
  Derived d(src, &fit);
  d.hal_strength = fit.hal_strength;
  d.hal_radius = fit.hal_radius;
  d.scan_scale = fit_scan_scale(src);
  d.grain = d.dir = d.adjacency = false;
  std::vector<float> out(w * h * 3);
  run_negative<P>(d, rgb, w, h, m.pitch_um, 1.0f / src.box_iso, 0, out.data(), nullptr);
  std::copy(out.begin(), out.end(), rgb);
  return 0;
}
template int film_fit_process<float>(float*, size_t, size_t, const FilmFitModel&, const FilmFitParams&);
template int film_fit_process<double>(float*, size_t, size_t, const FilmFitModel&, const FilmFitParams&);

int film_fit_nominal(const FilmFitModel& m, FilmFitParams* out) {
  if (!stock_ok(m.stock) || (m.bins != 16 && m.bins != 81)) return -1;
  const Derived d(load_src(m.stock, m.bins, m.stock_seed));
  *out = FilmFitParams{};
  out->n_emulsions = (int)d.em.size();
  out->n_bins = d.g.n;
  for (size_t e = 0; e < d.em.size(); ++e) {
    for (int i = 0; i < d.g.n; ++i) out->sens[e][i] = d.em[e].sens[i];
    out->d_max[e] = d.em[e].d_max;
    out->gamma[e] = d.em[e].gamma;
    if (d.em[e].lut.n < 64) {
      out->n_knots = d.em[e].lut.n;
      for (int j = 0; j < d.em[e].lut.n; ++j) out->knot_frac[e][j] = d.em[e].lut.fracs[j];
    }
  }
  out->hal_strength = 1.0;
  out->hal_radius = 1.0;
  return 0;
}

// ---- the model's reverse pass in f64 -------------------------------------
// The forward is film_fit_process<double> step by step (exposure walk, local
// scatter, halation bounce, capture curve, H&D reduction with the coloured
// mask, densitometric scan); the reverse applies the transpose of each step:
//   scan:       dL/dt_i = w_i sum_j xyzbar_j cmf_j(i),  dL/ddens_i = -ln10 t_i dL/dt_i
//   reduction:  D = d_max c^(1/g), M = 0.4 d_max (1 - clamp(D/d_max)) (mask stocks)
//   capture:    f = (1 - t) f_lo + t f_hi at the bracketing knots, t linear in log10 P
//   halation:   P_e = S_e + gain_e B, B = G_sigma(S_deepest), dB/dsigma = V'H + VH'
//   scatter:    S = 0.82 A + 0.18 G_local(A)   (the blurs' transposes by scattering
//               every output's taps back onto the sources it read, reflect-101)
//   exposure:   A_e = (1/span) sum_i w_i phi_i (prod of T_l above e)(1 - T_e),
//               phi_i = (M^-1 rgb)_+ . basis(i) / e_rel(i), T_e = exp(-2 rho t sens_e)
// The f32 roundings of d_max, gamma, the halation gain and sigma in the forward
// are taken as the identity; the bleed factor (the sensitivity's argmax, piecewise
// constant) is held fixed.
namespace {

struct GradTaps {
  int r = 0;
  std::vector<double> k, dk;  // normalised taps of blur() and their derivative in sigma
};

GradTaps grad_taps(float sigma) {
  GradTaps t;
  t.r = (int)std::max(std::ceil(3.0f * sigma), 1.0f);
  const double s = (double)sigma, inv_2s2 = 1.0 / (2.0 * s * s);
  const size_t nt = (size_t)(2 * t.r + 1);
  std::vector<double> g(nt), gd(nt);
  double sg = 0.0, sgd = 0.0;
  for (size_t i = 0; i < nt; ++i) {
    const double x = (double)i - (double)t.r;
    g[i] = std::exp(-x * x * inv_2s2);
    gd[i] = g[i] * x * x / (s * s * s);
    sg += g[i];
    sgd += gd[i];
  }
  t.k.resize(nt);
  t.dk.resize(nt);
  for (size_t i = 0; i < nt; ++i) {
    t.k[i] = g[i] / sg;
    t.dk[i] = (gd[i] * sg - g[i] * sgd) / (sg * sg);
  }
  return t;
}

size_t refl101(long long x, long long n) {
  for (;;) {
    if (x < 0) x = -x;
    else if (x >= n) x = 2 * n - 2 - x;
    else return (size_t)x;
    if (n == 1) return 0;
  }
}

// out = V(H in): rows with taps kh, then columns with kv (reflect-101)
void sep_blur(const std::vector<double>& in, std::vector<double>& out, size_t w, size_t h,
              const std::vector<double>& kh, const std::vector<double>& kv, int r) {
  std::vector<double> tmp(w * h);
#pragma omp parallel for schedule(static)
  for (long long y = 0; y < (long long)h; ++y)
    for (size_t x = 0; x < w; ++x) {
      double acc = 0.0;
      for (int j = 0; j <= 2 * r; ++j) acc += in[y * w + refl101((long long)x + j - r, (long long)w)] * kh[j];
      tmp[y * w + x] = acc;
    }
  out.assign(w * h, 0.0);
#pragma omp parallel for schedule(static)
  for (long long y = 0; y < (long long)h; ++y)
    for (size_t x = 0; x < w; ++x) {
      double acc = 0.0;
      for (int j = 0; j <= 2 * r; ++j) acc += tmp[refl101(y + j - r, (long long)h) * w + x] * kv[j];
      out[y * w + x] = acc;
    }
}

// out = (V(H .))^T in = H^T (V^T in), each forward output's taps scattered onto its sources
void sep_blur_transpose(const std::vector<double>& in, std::vector<double>& out, size_t w, size_t h,
                        const std::vector<double>& k, int r) {
  std::vector<double> tmp(w * h, 0.0);
#pragma omp parallel for schedule(static)
  for (long long x = 0; x < (long long)w; ++x)  // one column per thread: its scatter stays in the column
    for (size_t y = 0; y < h; ++y)
      for (int j = 0; j <= 2 * r; ++j) tmp[refl101((long long)y + j - r, (long long)h) * w + x] += k[j] * in[y * w + x];
  out.assign(w * h, 0.0);
#pragma omp parallel for schedule(static)
  for (long long y = 0; y < (long long)h; ++y)  // one row per thread
    for (size_t x = 0; x < w; ++x)
      for (int j = 0; j <= 2 * r; ++j) out[y * w + refl101((long long)x + j - r, (long long)w)] += k[j] * tmp[y * w + x];
}

}  // namespace

int film_fit_gradient(const float* rgb, const float* gout, size_t w, size_t h, const FilmFitModel& m,
                      const FilmFitParams& fit, FilmFitParams* grad, float* out_rgb) {
  if (w == 0 || h == 0 || !grad || !stock_ok(m.stock)) return -1;
  const Src src = load_src(m.stock, m.bins, m.stock_seed);
  if (fit_model_check(m, src, &fit)) return -1;
  Derived d(src, &fit);
  if (d.any_reversal) return -1;
  const double scale = fit_scan_scale(src);
  const size_t n = w * h, E = d.em.size();
  const Grid& g = d.g;
  const int nb = g.n;
  const Basis& Bs = basis(g);
  const Cmf& Cm = cmf_for(g);
  const double span = g.lam(nb - 1) - g.lam(0);
  const float pitch = m.pitch_um;
  double tw[NBMAX];
  for (int i = 0; i < nb; ++i) tw[i] = (i == 0 || i == nb - 1) ? 0.5 * g.dl : g.dl;
  // absorbing layers above each emulsion: transmittance product tab_e(i), emulsion layer index
  std::vector<int> em_layer(E);
  for (size_t e = 0; e < E; ++e) em_layer[e] = d.em[e].layer;
  std::vector<std::vector<double>> tab(E, std::vector<double>(nb, 1.0));
  for (size_t e = 0; e < E; ++e)
    for (int l = 0; l < em_layer[e]; ++l)
      if (d.absorbs[l])
        for (int i = 0; i < nb; ++i) tab[e][i] *= d.trans[l][i];

  // ---- forward: exposure planes A_e and the basis weights per pixel
  std::vector<std::vector<double>> A(E, std::vector<double>(n));
  std::vector<double> wgt(3 * n);
#pragma omp parallel for schedule(static)
  for (long long p = 0; p < (long long)n; ++p) {
    const double px[3] = {(double)rgb[p * 3], (double)rgb[p * 3 + 1], (double)rgb[p * 3 + 2]};
    for (int r = 0; r < 3; ++r) {
      const double v = Bs.minv[r][0] * px[0] + Bs.minv[r][1] * px[1] + Bs.minv[r][2] * px[2];
      wgt[p * 3 + r] = v < 0.0 ? 0.0 : v;
    }
    for (size_t e = 0; e < E; ++e) {
      double acc = 0.0;
      const size_t l = (size_t)em_layer[e];
      for (int i = 0; i < nb; ++i) {
        const double phi = (wgt[p * 3] * Bs.b[0][i] + wgt[p * 3 + 1] * Bs.b[1][i] + wgt[p * 3 + 2] * Bs.b[2][i]) /
                           (550.0 / g.lam(i));
        acc += tw[i] * phi * tab[e][i] * (1.0 - d.trans[l][i]);
      }
      A[e][p] = acc / span;
    }
  }
  // local scatter
  const float sigma_local = std::max(src.psf_local_um / std::max(pitch, 1e-6f), 0.0f);
  const bool do_local = sigma_local >= 1e-3f;
  const double f_loc = (double)0.18f, keep = 1.0 - f_loc;
  GradTaps tl;
  std::vector<std::vector<double>> S(E);
  for (size_t e = 0; e < E; ++e) {
    if (!do_local) {
      S[e] = A[e];
      continue;
    }
    if (tl.k.empty()) tl = grad_taps(sigma_local);
    std::vector<double> bl;
    sep_blur(A[e], bl, w, h, tl.k, tl.k, tl.r);
    S[e].resize(n);
    for (size_t i = 0; i < n; ++i) S[e][i] = keep * A[e][i] + f_loc * bl[i];
  }
  // halation
  const double sigma_unit = (double)std::max(src.psf_halation_um / std::max(pitch, 1e-6f), 0.0f);
  const float sigma_wide = (float)(sigma_unit * fit.hal_radius);
  bool do_wide = sigma_wide >= 1e-3f && E > 0;
  if (do_wide) {
    float max_r = 0.0f;
    const double bands[4] = {450.0, 550.0, 650.0, 700.0};
    for (double bl : bands) max_r = std::max(max_r, (float)evaluate(d.reflect, g, bl));
    do_wide = max_r > 0.0f;
  }
  std::vector<double> gain(E, 0.0), rb(E, 0.0), bleed(E, 0.0), rnum(E, 0.0), rden(E, 0.0);
  std::vector<double> Bn;
  GradTaps th;
  if (do_wide) {
    th = grad_taps(sigma_wide);
    sep_blur(S[E - 1], Bn, w, h, th.k, th.k, th.r);
    for (size_t e = 0; e < E; ++e) {
      const Emulsion& em = d.em[e];
      double peak = 550.0, mx = -INFINITY;
      for (int i = 0; i < nb; ++i)
        if (em.sens[i] > mx) { mx = em.sens[i]; peak = g.lam(i); }
      const float wb = 0.20f, wg = 0.45f, wr = 1.0f;
      float bl;
      if (peak <= 450.0) bl = wb;
      else if (peak <= 550.0) bl = wb + (wg - wb) * (float)((peak - 450.0) / 100.0);
      else if (peak <= 650.0) bl = wg + (wr - wg) * (float)((peak - 550.0) / 100.0);
      else bl = wr;
      const double norm = integrate(em.sens, g);
      double prod[NBMAX];
      for (int i = 0; i < nb; ++i) prod[i] = d.reflect[i] * em.sens[i];
      rnum[e] = integrate(prod, g);
      rden[e] = norm;
      const float r_e = norm <= 1e-9 ? (float)evaluate(d.reflect, g, 650.0) : (float)(rnum[e] / norm);
      bleed[e] = bl;
      rb[e] = (double)(r_e * bl);
      const float gn = (float)(rb[e] * fit.hal_strength);
      gain[e] = gn > 0.0f ? (double)gn : 0.0;
    }
  }

  // ---- capture, reduction, scan forward and their reverse, per pixel
  std::vector<std::vector<double>> Sbar(E, std::vector<double>(n, 0.0));
  std::vector<double> Bbar(do_wide ? n : 0, 0.0);
  // per-row sums (added in row order): d_max, gamma, gain, knot fractions
  const size_t K = (size_t)fit.n_knots, nsum = 3 * E + E * K;
  std::vector<double> rows(h * nsum, 0.0);
  const double ln10 = 2.30258509299404568402;
#pragma omp parallel for schedule(static)
  for (long long y = 0; y < (long long)h; ++y) {
    double* acc = rows.data() + (size_t)y * nsum;
    for (size_t x = 0; x < w; ++x) {
      const size_t p = (size_t)y * w + x;
      double D[8], M[8], eff[8], fr[8], tt[8], dfdx[8], xv[8];
      int lo[8];
      bool clamp_lo[8], clamp_hi[8];
      for (size_t e = 0; e < E; ++e) {
        const Emulsion& em = d.em[e];
        const Lut& L = em.lut;
        const double P = S[e][p] + (do_wide && gain[e] > 0.0 ? gain[e] * Bn[p] : 0.0);
        const double xx = P;  // reciprocity efficiency 1 at the shutter 1/box ISO
        xv[e] = xx;
        lo[e] = 0;
        tt[e] = 0.0;
        dfdx[e] = 0.0;
        clamp_lo[e] = clamp_hi[e] = false;
        double f;
        if (!(xx > 0.0)) {
          f = L.fracs[0];
          clamp_lo[e] = true;
        } else {
          const double lp = std::log10(xx);
          if (lp <= L.logs[0]) {
            f = L.fracs[0];
            clamp_lo[e] = true;
          } else if (lp >= L.logs[L.n - 1]) {
            f = L.fracs[L.n - 1];
            clamp_hi[e] = true;
          } else {
            int a = 0, b = L.n - 1;
            while (b - a > 1) {
              const int mid = (a + b) / 2;
              if (L.logs[mid] <= lp) a = mid;
              else b = mid;
            }
            const double t = (lp - L.logs[a]) / (L.logs[b] - L.logs[a]);
            f = L.fracs[a] * (1.0 - t) + L.fracs[b] * t;
            lo[e] = a;
            tt[e] = t;
            dfdx[e] = (L.fracs[b] - L.fracs[a]) / (L.logs[b] - L.logs[a]) / (xx * ln10);
          }
        }
        fr[e] = f;
        const double c = rclamp(f, 0.0, 1.0);
        eff[e] = c;
        const double dm = (double)em.d_max, ig = 1.0 / (double)std::max(em.gamma, 1e-6f);
        D[e] = dm * std::pow(c, ig);
        M[e] = em.has_mask ? dm * (double)0.4f * (1.0 - rclamp(D[e] / dm, 0.0, 1.0)) : 0.0;
      }
      double dens[NBMAX], t[NBMAX];
      for (int i = 0; i < nb; ++i) dens[i] = 0.0;
      for (size_t e = 0; e < E; ++e)
        for (int i = 0; i < nb; ++i) {
          dens[i] += D[e] * d.em[e].eps[i];
          if (d.em[e].has_mask) dens[i] += M[e] * d.em[e].mask[i];
        }
      for (int i = 0; i < nb; ++i) t[i] = std::pow(10.0, -dens[i]) * d.scanner[i];
      if (out_rgb) {
        double cc[3];
        spectrum_to_acescg(t, g, cc);
        for (int k = 0; k < 3; ++k) out_rgb[p * 3 + k] = (float)(cc[k] * scale);
      }
      // reverse
      double xyzbar[3];
      for (int j = 0; j < 3; ++j) {
        xyzbar[j] = 0.0;
        for (int k = 0; k < 3; ++k) xyzbar[j] += XYZ2AP1[k][j] * scale * (double)gout[p * 3 + k];
      }
      double Dbar[8] = {0}, Mbar[8] = {0};
      for (int i = 0; i < nb; ++i) {
        const double tbar = tw[i] * (xyzbar[0] * Cm.c[0][i] + xyzbar[1] * Cm.c[1][i] + xyzbar[2] * Cm.c[2][i]);
        const double densbar = -ln10 * t[i] * tbar;
        for (size_t e = 0; e < E; ++e) {
          Dbar[e] += densbar * d.em[e].eps[i];
          if (d.em[e].has_mask) Mbar[e] += densbar * d.em[e].mask[i];
        }
      }
      double bsum = 0.0;
      for (size_t e = 0; e < E; ++e) {
        const Emulsion& em = d.em[e];
        const double dm = (double)em.d_max, ig = 1.0 / (double)std::max(em.gamma, 1e-6f);
        double Db = Dbar[e], dmb = 0.0;
        if (em.has_mask) {
          const double q = D[e] / dm;
          if (q > 0.0 && q < 1.0) {
            Db += -(double)0.4f * Mbar[e];
            dmb += (double)0.4f * Mbar[e];
          } else if (q <= 0.0) {
            dmb += (double)0.4f * Mbar[e];
          }
        }
        const double c = eff[e];
        dmb += Db * std::pow(c, ig);
        const double gb = c > 0.0 ? Db * D[e] * std::log(c) * (-ig * ig) : 0.0;
        const bool inside = fr[e] > 0.0 && fr[e] < 1.0;
        const double fbar = inside && c > 0.0 ? Db * dm * ig * std::pow(c, ig - 1.0) : 0.0;
        acc[e] += dmb;
        acc[E + e] += gb;
        for (size_t k = 0; k < K; ++k) {
          double wk = 0.0;
          if (clamp_lo[e]) wk = k == 0 ? 1.0 : 0.0;
          else if (clamp_hi[e]) wk = k == K - 1 ? 1.0 : 0.0;
          else if ((int)k == lo[e]) wk = 1.0 - tt[e];
          else if ((int)k == lo[e] + 1) wk = tt[e];
          acc[3 * E + e * K + k] += fbar * wk;
        }
        const double Pb = fbar * dfdx[e];
        Sbar[e][p] = Pb;
        if (do_wide && gain[e] > 0.0) {
          bsum += gain[e] * Pb;
          acc[2 * E + e] += Pb * Bn[p];
        }
      }
      if (do_wide) Bbar[p] = bsum;
    }
  }
  std::vector<double> sums(nsum, 0.0);
  for (size_t y = 0; y < h; ++y)
    for (size_t k = 0; k < nsum; ++k) sums[k] += rows[y * nsum + k];

  *grad = FilmFitParams{};
  grad->n_emulsions = (int)E;
  grad->n_bins = nb;
  grad->n_knots = (int)K;
  for (size_t e = 0; e < E; ++e) {
    grad->d_max[e] = sums[e];
    grad->gamma[e] = sums[E + e];
    for (size_t k = 0; k < K; ++k) grad->knot_frac[e][k] = sums[3 * E + e * K + k];
  }
  // ---- halation reverse
  if (do_wide) {
    std::vector<double> back, d1, d2;
    sep_blur_transpose(Bbar, back, w, h, th.k, th.r);
    for (size_t i = 0; i < n; ++i) Sbar[E - 1][i] += back[i];
    sep_blur(S[E - 1], d1, w, h, th.k, th.dk, th.r);
    sep_blur(S[E - 1], d2, w, h, th.dk, th.k, th.r);
    double sb = 0.0;
    for (size_t i = 0; i < n; ++i) sb += Bbar[i] * (d1[i] + d2[i]);
    grad->hal_radius = sb * sigma_unit;
    double hs = 0.0;
    for (size_t e = 0; e < E; ++e)
      if (gain[e] > 0.0) hs += sums[2 * E + e] * rb[e];
    grad->hal_strength = hs;
  }
  // ---- local scatter reverse: Abar = 0.82 Sbar + 0.18 G^T Sbar
  std::vector<std::vector<double>> Abar(E);
  for (size_t e = 0; e < E; ++e) {
    if (!do_local) {
      Abar[e] = Sbar[e];
      continue;
    }
    std::vector<double> back;
    sep_blur_transpose(Sbar[e], back, w, h, tl.k, tl.r);
    Abar[e].resize(n);
    for (size_t i = 0; i < n; ++i) Abar[e][i] = keep * Sbar[e][i] + f_loc * back[i];
  }
  // ---- exposure reverse: Q[e][k] = sum_p Abar_e(p) w_k(p), then the transmittance chain
  std::vector<double> Q(E * 3, 0.0);
  for (size_t e = 0; e < E; ++e)
    for (int k = 0; k < 3; ++k) {
      double acc = 0.0;
      for (size_t p = 0; p < n; ++p) acc += Abar[e][p] * wgt[p * 3 + k];
      Q[e * 3 + k] = acc;
    }
  for (size_t e2 = 0; e2 < E; ++e2) {
    const SrcLayer& L2 = src.layers[(size_t)em_layer[e2]];
    const double alpha2 = 2.0 * (double)L2.agx * (double)L2.thickness_um;
    for (int i = 0; i < nb; ++i) {
      double gsum = 0.0;
      for (size_t e = 0; e < E; ++e) {
        // Phi_e(i) = sum_p Abar_e(p) w_i phi_i(p) / span
        const double phi_sum = tw[i] / span / (550.0 / g.lam(i)) *
                               (Q[e * 3] * Bs.b[0][i] + Q[e * 3 + 1] * Bs.b[1][i] + Q[e * 3 + 2] * Bs.b[2][i]);
        const double Te = d.trans[(size_t)em_layer[e]][i];
        double dtau = 0.0;  // d(tab_e (1 - T_e)) / dsens_e2(i)
        if (e == e2) dtau = alpha2 * tab[e][i] * Te;
        else if (em_layer[e2] < em_layer[e]) dtau = -alpha2 * tab[e][i] * (1.0 - Te);
        gsum += phi_sum * dtau;
      }
      if (do_wide && gain[e2] > 0.0 && rden[e2] > 1e-9) {
        const double dr = tw[i] * (d.reflect[i] * rden[e2] - rnum[e2]) / (rden[e2] * rden[e2]);
        gsum += sums[2 * E + e2] * fit.hal_strength * bleed[e2] * dr;
      }
      grad->sens[e2][i] = gsum;
    }
  }
  return 0;
}

// Pieces of the restatement exposed on their own so tests can check them
// against independent computations.
void film_blur(float* buf, size_t w, size_t h, float sigma) {
  std::vector<float> v(buf, buf + w * h);
  blur(v, w, h, sigma);
  std::copy(v.begin(), v.end(), buf);
}

void film_grain_noise(float* out, size_t w, size_t h, uint64_t seed, size_t layer) {
  grain_noise(out, w, h, seed, layer);
}

void film_philox(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]) {
  for (int i = 0; i < 4; ++i) out[i] = ctr[i];
  philox4x32_10(out, key[0], key[1]);
}

void film_grain_noise_philox(float* out, size_t n, uint64_t seed, size_t layer) {
  grain_noise_philox(out, n, seed, layer);
}

int film_capture_lut(int stock_idx, int emulsion, double* logs, double* fracs, double params[3]) {
  if (!stock_ok(stock_idx) || stock_idx == 7) return -1;
  const Derived d(load_src(stock_idx, 16, 0));
  if (emulsion < 0 || emulsion >= (int)d.em.size()) return -1;
  const Emulsion& e = d.em[emulsion];
  std::copy(e.lut.logs, e.lut.logs + 64, logs);
  std::copy(e.lut.fracs, e.lut.fracs + 64, fracs);
  const SrcLayer& L = d.s->layers[e.layer];
  params[0] = L.mu_ln;
  params[1] = L.sigma_ln;
  params[2] = L.capture_k;
  return 0;
}

// The characteristic curve knots of an emulsion of a stock (64 LUT knots, or a knot curve's).
int film_stock_knots(int stock_idx, int bins, uint64_t stock_seed, int emulsion, double* logs, double* fracs) {
  if (!stock_ok(stock_idx) || (bins != 16 && bins != 81)) return -1;
  const Derived d(load_src(stock_idx, bins, stock_seed));
  if (emulsion < 0 || emulsion >= (int)d.em.size()) return -1;
  const Lut& L = d.em[emulsion].lut;
  std::copy(L.logs, L.logs + L.n, logs);
  std::copy(L.fracs, L.fracs + L.n, fracs);
  return L.n;
}

// An emulsion's sensitivity samples on the grid.
int film_stock_sensitivity(int stock_idx, int bins, uint64_t stock_seed, int emulsion, double* out) {
  if (!stock_ok(stock_idx) || (bins != 16 && bins != 81)) return -1;
  const Derived d(load_src(stock_idx, bins, stock_seed));
  if (emulsion < 0 || emulsion >= (int)d.em.size()) return -1;
  std::copy(d.em[emulsion].sens, d.em[emulsion].sens + d.g.n, out);
  return d.g.n;
}

void film_cmf(int bins, double* out) {
  const Grid& g = grid_for(bins);
  const Cmf& c = cmf_for(g);
  for (int k = 0; k < 3; ++k)
    for (int i = 0; i < g.n; ++i) out[k * g.n + i] = c.c[k][i];
}

double film_lut_sample(int stock_idx, int emulsion, double phi) {
  const Derived d(load_src(stock_idx, 16, 0));
  return d.em[emulsion].lut.sample(phi);
}

void film_upsample(const float rgb[3], double out[16]) {
  const Basis& B = basis(kG16);
  double wgt[3];
  for (int r = 0; r < 3; ++r) {
    wgt[r] = B.minv[r][0] * rgb[0] + B.minv[r][1] * rgb[1] + B.minv[r][2] * rgb[2];
    if (wgt[r] < 0.0) wgt[r] = 0.0;
  }
  for (int i = 0; i < NB; ++i) out[i] = wgt[0] * B.b[0][i] + wgt[1] * B.b[1][i] + wgt[2] * B.b[2][i];
}

void film_spectrum_to_acescg(const double s[16], double out[3]) { spectrum_to_acescg(s, kG16, out); }

int film_normalized_dmin(int stock_idx, float out[3]) {
  if (!stock_ok(stock_idx)) return -1;
  const Derived d(load_src(stock_idx, 16, 0));
  normalized_dmin(d, out);
  return 0;
}

}  // namespace oracle
