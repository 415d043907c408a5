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
// Pins: grain noise = SplitMix64 row streams (Python-integer brute force plus
// the published test vector); blur = reflect-101 brute force; capture LUT =
// adaptive quadrature; upsample/scan round trip; the reference's end-to-end
// film test gates.  The per-stock absolute outputs are otherwise parity
// unpinned (the reference's own tests only gate them loosely).
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <vector>

#include "../imops_amd/csrc/film_stock_data.hpp"
#include "oracle_common.hpp"
#include "oracle_film.hpp"

namespace oracle {

using namespace film_data;

namespace {

constexpr int NB = 16;

// The reference rounds its f64 islands to f32 where it stores planes; the FP64
// oracle keeps the f64 value.
template <class P> inline P store(double v) { return (P)(float)v; }
template <> inline double store<double>(double v) { return v; }

double lambda_nm(int i) { return 400.0 + 20.0 * i; }

void curve_samples(const Curve& c, double* out) {
  for (int i = 0; i < NB; ++i) {
    const double l = lambda_nm(i);
    switch (c.kind) {
      case CURVE_CONST: out[i] = c.a; break;
      case CURVE_GAUSS: {
        const double d = (l - c.b) / c.c;
        out[i] = c.a * std::exp(-0.5 * d * d);
        break;
      }
      case CURVE_ORANGE: {  // orange_mask_epsilon (color_neg_200.rs:51-65 and the other colour negatives)
        const double u = (l - 450.0) / 50.0, v = (l - 520.0) / 60.0;
        const double blue = std::exp(-0.5 * (u * u));
        const double green = c.c * std::exp(-0.5 * (v * v));
        out[i] = c.a * blue + c.b * green + c.d;
        break;
      }
      case CURVE_PANCHRO: {  // panchromatic_sensitivity (bw_stub.rs:19-32, trix_400.rs:16-28)
        const double t = (l - 400.0) / 300.0;
        const double envelope = std::fmax(std::sin(M_PI * t), 0.0);
        out[i] = c.a + c.b * envelope;
        break;
      }
      default: out[i] = 0.0;
    }
  }
}

double integrate(const double* s) {  // SpectralCurve::integrate (spectrum.rs:106-118)
  double acc = 0.0;
  for (int i = 0; i < NB - 1; ++i) acc += 0.5 * (s[i] + s[i + 1]) * (lambda_nm(i + 1) - lambda_nm(i));
  return acc;
}

double evaluate(const double* s, double l) {  // SpectralCurve::evaluate (spectrum.rs:60-78)
  if (l <= lambda_nm(0)) return s[0];
  if (l >= lambda_nm(NB - 1)) return s[NB - 1];
  for (int i = 0; i < NB - 1; ++i)
    if (l >= lambda_nm(i) && l <= lambda_nm(i + 1)) {
      const double t = (l - lambda_nm(i)) / (lambda_nm(i + 1) - lambda_nm(i));
      return s[i] * (1.0 - t) + s[i + 1] * t;
    }
  return s[NB - 1];
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

double integrate_cmf(const double* s, const double* cmf) {  // upsample.rs:85-95
  double acc = 0.0;
  for (int i = 0; i < 15; ++i) {
    const double a = s[i] * cmf[i], b = s[i + 1] * cmf[i + 1];
    acc += 0.5 * (a + b) * 20.0;
  }
  return acc;
}

void spectrum_to_acescg(const double* s, double out[3]) {
  const double x = integrate_cmf(s, XBAR), y = integrate_cmf(s, YBAR), z = integrate_cmf(s, ZBAR);
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

struct Basis {  // UpsampleBasis::new (upsample.rs:123-143)
  double b[3][NB];
  double minv[3][3];
  Basis() {
    const double peaks[3] = {450.0, 550.0, 650.0};
    double cols[3][3];
    for (int k = 0; k < 3; ++k) {
      for (int i = 0; i < NB; ++i) {
        const double d = (lambda_nm(i) - peaks[k]) / 40.0;
        b[k][i] = std::exp(-0.5 * d * d);
      }
      spectrum_to_acescg(b[k], cols[k]);
    }
    double m[3][3];
    for (int r = 0; r < 3; ++r)
      for (int k = 0; k < 3; ++k) m[r][k] = cols[k][r];
    inv3(m, minv);
  }
};

const Basis& basis() {
  static Basis b;
  return b;
}

struct Lut {  // DevelopableFractionLut (capture.rs:37-88)
  double logs[64], fracs[64];
  void build(double mu_ln, double sigma_ln, double k) {
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
    if (lp >= logs[63]) return fracs[63];
    int lo = 0, hi = 63;
    while (hi - lo > 1) {
      const int mid = (lo + hi) / 2;
      if (logs[mid] <= lp) lo = mid;
      else hi = mid;
    }
    const double t = (lp - logs[lo]) / (logs[hi] - logs[lo]);
    return fracs[lo] * (1.0 - t) + fracs[hi] * t;
  }
};

struct Emulsion {
  int layer;
  double sens[NB], eps[NB], mask[NB];
  bool has_mask, reversal;
  float d_max, gamma, recip_p;
  Lut lut;
  float kappa_ref;
};

struct Derived {
  const Stock* s;
  std::vector<Emulsion> em;
  double trans[12][NB];  // per-layer Beer-Lambert transmittance (absorbing layers)
  bool absorbs[12], is_emulsion[12];
  double reflect[NB], scanner[NB];
  bool any_reversal = false;
  // differentiable-model switches (FilmFitParams); defaults reproduce film::process
  double hal_strength = 1.0, hal_radius = 1.0, scan_scale = 0.0;  // scan_scale <= 0: from this stock's Dmin
  bool grain = true, dir = true, adjacency = true;
  explicit Derived(const Stock* st, const FilmFitParams* fit = nullptr) : s(st) {
    curve_samples(st->ah_reflectance, reflect);
    curve_samples(st->scanner_light, scanner);
    int ei = 0;
    for (int l = 0; l < st->n_layers; ++l) {
      const Layer& L = st->layers[l];
      if (L.is_reversal) any_reversal = true;
      double curve[NB];
      curve_samples(L.sensitivity, curve);
      if (fit && L.kind == L_EMULSION && L.sensitivity.kind != CURVE_NONE)
        for (int i = 0; i < NB; ++i) curve[i] = fit->sens[ei][i];
      is_emulsion[l] = L.kind == L_EMULSION;
      absorbs[l] = L.kind != L_SUPPORT && L.sensitivity.kind != CURVE_NONE;
      const double thickness = (double)L.thickness_um;
      for (int i = 0; i < NB; ++i) {
        double od;
        if (L.kind == L_EMULSION) {
          const double sigma = curve[i] * 2.0;  // ABSORPTION_SIGMA_SCALE_PER_UM (constants.rs:51)
          od = sigma * (double)L.silver_halide_fraction * thickness;
        } else {
          od = curve[i] * thickness;
        }
        trans[l][i] = std::exp(-od);
      }
      if (L.kind != L_EMULSION) continue;
      Emulsion e;
      e.layer = l;
      for (int i = 0; i < NB; ++i) e.sens[i] = curve[i];
      curve_samples(L.dye_epsilon, e.eps);
      e.has_mask = L.mask_epsilon.kind != CURVE_NONE;
      curve_samples(L.mask_epsilon, e.mask);
      e.reversal = L.is_reversal;
      e.d_max = fit ? (float)fit->d_max[ei] : L.d_max;
      e.gamma = fit ? (float)fit->gamma[ei] : L.gamma_contrast;
      ++ei;
      e.recip_p = L.reciprocity_p;
      const double mu = std::log(L.crystal_mean_um), sg = L.crystal_sigma_ln;
      e.lut.build(mu, sg, L.capture_k);
      // grain kappa at 1 um pitch (stock/mod.rs:193-211)
      const double mean_s = std::exp(mu + 0.5 * sg * sg);
      const double r = std::max(mean_s * 0.5, 1e-6);
      const double volume = 1.04719755119659774615 * 4.0 * r * r * r;  // FRAC_PI_3 * 4 r^3
      const double rho = (double)L.silver_halide_fraction * thickness / std::max(volume, 1e-18);
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

// exposure/mod.rs:45-122
template <class P>
Planes<P> expose(const float* rgb, size_t w, size_t h, const Derived& d, float pitch, float shutter) {
  const size_t n = w * h;
  const size_t E = d.em.size();
  const Stock& st = *d.s;
  const Basis& B = basis();
  Planes<P> abs(E, std::vector<P>(n));
#pragma omp parallel for schedule(static)
  for (long long p = 0; p < (long long)n; ++p) {
    const double px[3] = {(double)rgb[p * 3], (double)rgb[p * 3 + 1], (double)rgb[p * 3 + 2]};
    double wgt[3];
    for (int r = 0; r < 3; ++r) {
      wgt[r] = B.minv[r][0] * px[0] + B.minv[r][1] * px[1] + B.minv[r][2] * px[2];
      if (wgt[r] < 0.0) wgt[r] = 0.0;
    }
    double phi[NB];
    for (int i = 0; i < NB; ++i) {
      const double s = wgt[0] * B.b[0][i] + wgt[1] * B.b[1][i] + wgt[2] * B.b[2][i];
      const double e_rel = 550.0 / lambda_nm(i);
      phi[i] = s / e_rel;
    }
    size_t e = 0;
    for (int l = 0; l < st.n_layers; ++l) {
      if (!d.absorbs[l]) continue;
      double absorbed[NB];
      for (int i = 0; i < NB; ++i) {
        const double phi_t = phi[i] * d.trans[l][i];
        absorbed[i] = phi[i] - phi_t;
        phi[i] = phi_t;
      }
      if (d.is_emulsion[l]) {
        double acc = 0.0;
        for (int i = 0; i < 15; ++i) acc += 0.5 * (absorbed[i] + absorbed[i + 1]) * 20.0;
        abs[e++][p] = store<P>(acc / 300.0);
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
  const float sigma_wide = (float)((double)std::max(st.psf_halation_um / std::max(pitch, 1e-6f), 0.0f) * d.hal_radius);
  bool do_wide = sigma_wide >= 1e-3f;
  if (do_wide) {
    float max_r = 0.0f;
    const double bands[4] = {450.0, 550.0, 650.0, 700.0};
    for (double bl : bands) max_r = std::max(max_r, (float)evaluate(d.reflect, bl));
    do_wide = max_r > 0.0f;
  }
  if (do_wide && E > 0) {
    std::vector<P> bounce = abs[E - 1];
    blur(bounce, w, h, sigma_wide);
    for (size_t e = 0; e < E; ++e) {
      const Emulsion& em = d.em[e];
      // bleed weight from the sensitivity peak (halation.rs:102-132)
      double peak = 550.0, mx = -INFINITY;
      for (int i = 0; i < NB; ++i)
        if (em.sens[i] > mx) { mx = em.sens[i]; peak = lambda_nm(i); }
      const float wb = 0.20f, wg = 0.45f, wr = 1.0f;
      float bleed;
      if (peak <= 450.0) bleed = wb;
      else if (peak <= 550.0) bleed = wb + (wg - wb) * (float)((peak - 450.0) / 100.0);
      else if (peak <= 650.0) bleed = wg + (wr - wg) * (float)((peak - 550.0) / 100.0);
      else bleed = wr;
      // effective_reflectance (halation.rs:155-161)
      float r_e;
      const double norm = integrate(em.sens);
      if (norm <= 1e-9) {
        r_e = (float)evaluate(d.reflect, 650.0);
      } else {
        double prod[NB];
        for (int i = 0; i < NB; ++i) prod[i] = d.reflect[i] * em.sens[i];
        r_e = (float)(integrate(prod) / norm);
      }
      const float gain = (float)((double)(r_e * bleed) * d.hal_strength);
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
  const Stock& st = *d.s;
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
  for (size_t e = 0; e < (d.grain ? E : 0); ++e) {
    const P d_max = (P)d.em[e].d_max;
    const P kappa = (P)((d.em[e].kappa_ref / std::max(pitch, 1e-6f)) * 0.15f);
    if (kappa <= (P)0 || d_max <= (P)0) continue;
    std::vector<P> noise(n);
    grain_noise(noise.data(), w, h, seed, e);
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
      const P sigma_d = taper * std::sqrt(std::max(dens * (d_max - dens), (P)0));
      P v = dens + kappa * sigma_d * noise[i] * norm;
      plane[i] = rclamp(v, (P)0, d_max * (P)1.05f);
    }
  }
  return dy;
}

// densitometry.rs:13-73 (before the Dmin normalisation)
void scan_pixel(const Derived& d, const double* di, const double* dm, double out[3]) {
  double dens[NB] = {0};
  for (size_t e = 0; e < d.em.size(); ++e)
    for (int l = 0; l < NB; ++l) {
      dens[l] += di[e] * d.em[e].eps[l];
      if (d.em[e].has_mask) dens[l] += dm[e] * d.em[e].mask[l];
    }
  double t[NB];
  for (int i = 0; i < NB; ++i) t[i] = std::pow(10.0, -dens[i]) * d.scanner[i];
  spectrum_to_acescg(t, out);
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

const film_data::Stock* film_stock(int idx) { return stock_by_index(idx); }

// film/mod.rs:77-129
template <class P>
int film_process(float* rgb, size_t w, size_t h, int stock_idx, int format, uint64_t seed, int output,
                 float shutter_seconds, FilmTrace* trace) {
  const Stock* st = stock_by_index(stock_idx);
  if (!st || w == 0 || h == 0) return -1;
  const Derived d(st);
  const float pitch = format_width_mm(format) * 1000.0f / (float)w;
  const float shutter = shutter_seconds > 0.0f ? shutter_seconds : 1.0f / st->box_iso;
  std::vector<float> out(w * h * 3);
  run_negative<P>(d, rgb, w, h, pitch, shutter, seed, out.data(), trace);
  if (output == 1 && !d.any_reversal) {
    float dmin[3], mid[3];
    normalized_dmin(d, dmin);
    // mid-gray negative: 32x32 flat absolute-luminance gray, seed 0 (film/mod.rs:136-163)
    const int N = 32;
    // relative_to_absolute_y with sunny-16 metering (film/mod.rs:154-163, radiance.rs:19-25, 52-59)
    const float iso = std::max(st->box_iso, 1.0f);
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

template int film_process<float>(float*, size_t, size_t, int, int, uint64_t, int, float, FilmTrace*);
template int film_process<double>(float*, size_t, size_t, int, int, uint64_t, int, float, FilmTrace*);

// The differentiable film model: film::process (NegativeLinear, shutter 1/box
// ISO, a given pixel pitch) with grain, DIR inhibition and adjacency off, the emulsion sensitivities,
// D_max and gamma taken from `fit`, the halation gain and PSF width scaled by
// fit->hal_strength / fit->hal_radius, and the scan normalised by the nominal
// stock's Dmin (so the normalisation is not a function of the parameters).
template <class P>
int film_fit_process(float* rgb, size_t w, size_t h, int stock_idx, float pitch_um, const FilmFitParams& fit) {
  const Stock* st = stock_by_index(stock_idx);
  if (!st || w == 0 || h == 0) return -1;
  double nominal_scale;
  {
    const Derived nom(st);
    double dmin[3];
    dmin_reference(nom, dmin);
    nominal_scale = 1.0 / std::max(std::max(std::max(dmin[0], dmin[1]), dmin[2]), 1e-12);
  }
  Derived d(st, &fit);
  d.hal_strength = fit.hal_strength;
  d.hal_radius = fit.hal_radius;
  d.scan_scale = nominal_scale;
  d.grain = d.dir = d.adjacency = false;
  std::vector<float> out(w * h * 3);
  run_negative<P>(d, rgb, w, h, pitch_um, 1.0f / st->box_iso, 0, out.data(), nullptr);
  std::copy(out.begin(), out.end(), rgb);
  return 0;
}
template int film_fit_process<float>(float*, size_t, size_t, int, float, const FilmFitParams&);
template int film_fit_process<double>(float*, size_t, size_t, int, float, const FilmFitParams&);

int film_fit_nominal(int stock_idx, FilmFitParams* out) {
  const Stock* st = stock_by_index(stock_idx);
  if (!st) return -1;
  const Derived d(st);
  *out = FilmFitParams{};
  out->n_emulsions = (int)d.em.size();
  for (size_t e = 0; e < d.em.size(); ++e) {
    for (int i = 0; i < NB; ++i) out->sens[e][i] = d.em[e].sens[i];
    out->d_max[e] = d.em[e].d_max;
    out->gamma[e] = d.em[e].gamma;
  }
  out->hal_strength = 1.0;
  out->hal_radius = 1.0;
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

int film_capture_lut(int stock_idx, int emulsion, double* logs, double* fracs, double params[3]) {
  const Stock* st = stock_by_index(stock_idx);
  if (!st) return -1;
  const Derived d(st);
  if (emulsion < 0 || emulsion >= (int)d.em.size()) return -1;
  const Emulsion& e = d.em[emulsion];
  std::copy(e.lut.logs, e.lut.logs + 64, logs);
  std::copy(e.lut.fracs, e.lut.fracs + 64, fracs);
  const Layer& L = st->layers[e.layer];
  params[0] = std::log(L.crystal_mean_um);
  params[1] = L.crystal_sigma_ln;
  params[2] = L.capture_k;
  return 0;
}

double film_lut_sample(int stock_idx, int emulsion, double phi) {
  const Derived d(stock_by_index(stock_idx));
  return d.em[emulsion].lut.sample(phi);
}

void film_upsample(const float rgb[3], double out[16]) {
  const Basis& B = basis();
  double wgt[3];
  for (int r = 0; r < 3; ++r) {
    wgt[r] = B.minv[r][0] * rgb[0] + B.minv[r][1] * rgb[1] + B.minv[r][2] * rgb[2];
    if (wgt[r] < 0.0) wgt[r] = 0.0;
  }
  for (int i = 0; i < NB; ++i) out[i] = wgt[0] * B.b[0][i] + wgt[1] * B.b[1][i] + wgt[2] * B.b[2][i];
}

void film_spectrum_to_acescg(const double s[16], double out[3]) { spectrum_to_acescg(s, out); }

int film_normalized_dmin(int stock_idx, float out[3]) {
  const Stock* st = stock_by_index(stock_idx);
  if (!st) return -1;
  const Derived d(st);
  normalized_dmin(d, out);
  return 0;
}

}  // namespace oracle
