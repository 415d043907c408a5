// Host side of the film stage: stock tables -> device tables (the derived
// quantities of FilmStock::finalize, pichromatic/src/film/stock/mod.rs:152-214,
// and of the exposure/scan set-up), and the launch sequence of film::process
// (pichromatic/src/film/mod.rs:77-163).  Everything computed here is a
// per-stock or per-call constant evaluated with the host libm exactly as the
// reference evaluates it, so the device kernels only stream pixels.
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <vector>

#include "film.hpp"
#include "film_stock_data.hpp"
#include "stages.hpp"

namespace imops {

using namespace film_data;

namespace {

// The reference's spectral grid (16 samples, 400..700 nm); the differentiable
// model below works on it.
constexpr int NB = 16;
double lam(int i) { return 400.0 + 20.0 * i; }

// A spectral grid: the reference's (16 bins) or the 81-bin extension.
struct Grid {
  int n;
  double l0, dl;
  double lam(int i) const { return l0 + dl * i; }
};
const Grid kGridRef{16, 400.0, 20.0};
const Grid kGrid81{81, 380.0, 5.0};
const Grid& grid_of(int bins) { return bins == 81 ? kGrid81 : kGridRef; }

double gauss_at(double l, double peak, double sigma) {
  const double d = (l - peak) / sigma;
  return std::exp(-0.5 * d * d);
}

void sample_curve(const Curve& c, const Grid& g, double* out) {
  for (int i = 0; i < g.n; ++i) {
    const double l = g.lam(i);
    if (c.kind == CURVE_CONST) {
      out[i] = c.a;
    } else if (c.kind == CURVE_GAUSS) {  // gaussian_curve (color_neg_200.rs:38-50)
      out[i] = c.a * gauss_at(l, c.b, c.c);
    } else if (c.kind == CURVE_ORANGE) {  // orange_mask_epsilon (color_neg_200.rs:51-65, portra_400.rs:28-40)
      const double u = (l - 450.0) / 50.0, v = (l - 520.0) / 60.0;
      out[i] = c.a * std::exp(-0.5 * (u * u)) + c.b * (c.c * std::exp(-0.5 * (v * v))) + c.d;
    } else if (c.kind == CURVE_PANCHRO) {  // panchromatic_sensitivity (bw_stub.rs:19-32, trix_400.rs:16-28)
      const double t = (l - 400.0) / 300.0;
      out[i] = c.a + c.b * std::fmax(std::sin(3.14159265358979323846 * t), 0.0);
    } else if (c.kind == CURVE_GAUSS2) {  // two-component Gaussian mixture (synthetic stock)
      out[i] = c.a * gauss_at(l, c.b, c.c) + c.d * gauss_at(l, c.e, c.f);
    } else {
      out[i] = 0.0;
    }
  }
}
void sample_curve(const Curve& c, double* out) { sample_curve(c, kGridRef, out); }

double trapz(const double* s, const Grid& g) {  // SpectralCurve::integrate (spectrum.rs:106-118)
  double acc = 0.0;
  for (int i = 0; i < g.n - 1; ++i) acc += 0.5 * (s[i] + s[i + 1]) * (g.lam(i + 1) - g.lam(i));
  return acc;
}
double trapz(const double* s) { return trapz(s, kGridRef); }

double eval_curve(const double* s, const Grid& g, double l) {  // SpectralCurve::evaluate (spectrum.rs:60-78)
  if (l <= g.lam(0)) return s[0];
  if (l >= g.lam(g.n - 1)) return s[g.n - 1];
  for (int i = 0; i < g.n - 1; ++i)
    if (l >= g.lam(i) && l <= g.lam(i + 1)) {
      const double t = (l - g.lam(i)) / (g.lam(i + 1) - g.lam(i));
      return s[i] * (1.0 - t) + s[i + 1] * t;
    }
  return s[g.n - 1];
}
double eval_curve(const double* s, double l) { return eval_curve(s, kGridRef, l); }

// CIE 1931 2-degree CMFs on the 20 nm grid and XYZ(D65) -> AP1 (upsample.rs:17-45).
const double kCmf[3][16] = {
    {0.01431, 0.13438, 0.34828, 0.29080, 0.09564, 0.00490, 0.06327, 0.29040, 0.59450, 0.91630, 1.06220, 0.85445,
     0.44790, 0.16490, 0.04677, 0.01136},
    {0.000396, 0.004000, 0.023000, 0.060000, 0.139020, 0.323000, 0.710000, 0.954000, 0.995000, 0.870000, 0.631000,
     0.381000, 0.175000, 0.061000, 0.017000, 0.004100},
    {0.06785, 0.64560, 1.74706, 1.66920, 0.81295, 0.27200, 0.07825, 0.02030, 0.00390, 0.00165, 0.00080, 0.00019,
     0.00002, 0.00000, 0.00000, 0.00000}};
const double kXyz2Ap1[3][3] = {{1.6410233797, -0.3248032942, -0.2364246952},
                               {-0.6636628587, 1.6153315917, 0.0167563477},
                               {0.0117218943, -0.0082844165, 0.9883948585}};

// CMFs on a grid: the reference's table on its own grid; on the 81-bin grid the
// multi-lobe piecewise-Gaussian fit of the CIE 1931 2-degree observer (Wyman,
// Sloan, Shirley, "Simple analytic approximations to the CIE XYZ colour matching
// functions", JCGT 2(2), 2013), each lobe with separate widths below and above
// its centre.
double lobe(double l, double mu, double s_lo, double s_hi) {
  const double t = (l - mu) / (l < mu ? s_lo : s_hi);
  return std::exp(-0.5 * t * t);
}
void cmf_on(const Grid& g, double out[3][kFilmMaxBins]) {
  for (int i = 0; i < g.n; ++i) {
    if (g.n == 16) {
      for (int c = 0; c < 3; ++c) out[c][i] = kCmf[c][i];
      continue;
    }
    const double l = g.lam(i);
    out[0][i] = 1.056 * lobe(l, 599.8, 37.9, 31.0) + 0.362 * lobe(l, 442.0, 16.0, 26.7) -
                0.065 * lobe(l, 501.1, 20.4, 26.2);
    out[1][i] = 0.821 * lobe(l, 568.8, 46.9, 40.5) + 0.286 * lobe(l, 530.9, 16.3, 31.1);
    out[2][i] = 1.217 * lobe(l, 437.0, 11.8, 36.0) + 0.681 * lobe(l, 459.0, 26.0, 13.8);
  }
}

// spectrum_to_acescg_rgb (upsample.rs:85-95, 179-184): CMF trapezoid -> XYZ -> AP1.
void spectrum_ap1(const double* s, const Grid& g, const double cmf[3][kFilmMaxBins], double out[3]) {
  double xyz[3];
  for (int c = 0; c < 3; ++c) {
    double acc = 0.0;
    for (int i = 0; i < g.n - 1; ++i) {
      const double a = s[i] * cmf[c][i], b = s[i + 1] * cmf[c][i + 1];
      acc += 0.5 * (a + b) * g.dl;
    }
    xyz[c] = acc;
  }
  for (int r = 0; r < 3; ++r) out[r] = kXyz2Ap1[r][0] * xyz[0] + kXyz2Ap1[r][1] * xyz[1] + kXyz2Ap1[r][2] * xyz[2];
}

// Developable-fraction LUT (capture.rs:37-106): 64 log-spaced fluences,
// trapezoid over 65 nodes of the standardised log crystal size.
void build_lut(double mu_ln, double sigma_ln, double k, double* logs, double* fracs) {
  const double dz = (2.0 * 8.0) / 64.0;
  const double inv_sqrt_2pi = 1.0 / std::sqrt(2.0 * 3.14159265358979323846);
  for (int i = 0; i < 64; ++i) {
    const double t = (double)i / 63.0;
    const double lp = -4.0 + t * (8.0 - (-4.0));
    const double phi = std::pow(10.0, lp);
    double acc = 0.0;
    for (int j = 0; j <= 64; ++j) {
      const double z = -8.0 + (double)j * dz;
      const double wt = (j == 0 || j == 64) ? 0.5 : 1.0;
      const double pdf = inv_sqrt_2pi * std::exp(-0.5 * z * z);
      const double s = std::exp(mu_ln + sigma_ln * z);
      acc += wt * pdf * std::exp(-k * s * s * phi) * dz;
    }
    const double f = 1.0 - acc;
    logs[i] = lp;
    fracs[i] = f < 0.0 ? 0.0 : (f > 1.0 ? 1.0 : f);
  }
}

// A characteristic curve given by knots (log10 fluence, developable fraction),
// sampled like the reference's LUT: linear in log10 fluence, clamped at the ends.
struct KnotCurve {
  int n = 0;
  double log10_phi[kFilmLutMax], frac[kFilmLutMax];
};

// The seeded synthetic colour negative: the ColorNeg200 layer stack
// (color_neg_200.rs:67-209) with each emulsion's sensitivity a two-Gaussian
// mixture, its capture a 5-knot monotone characteristic curve, and no
// reciprocity failure.  Draws, in order, from SplitMix64 (grain.rs:20-45
// generator, state = seed), u = (next >> 11) * 2^-53, for emulsions blue, green,
// red (base 450, 550, 650 nm):
//   mu1 = base + 30 (u - 1/2), s1 = 25 + 20 u, mu2 = mu1 + 100 (u - 1/2),
//   s2 = 20 + 30 u, a2 = 0.15 + 0.35 u;
//   knot j = 0..4: log10 fluence -2.5 + j + 0.3 (u - 1/2);
//   knot j = 1..4: increment 0.2 + u; fraction_0 = 0.005 (the toe: unexposed film is
//   nearly the base), fraction_j = 0.005 + 0.965 (increments 1..j) / (increments 1..4).
struct SplitMix {
  uint64_t state;
  uint64_t next() {
    state += 0x9E3779B97F4A7C15ull;
    uint64_t z = state;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
  }
  double uniform() { return (double)(next() >> 11) * (1.0 / 9007199254740992.0); }
};

struct SyntheticStock {
  Stock stock;
  KnotCurve knots[3];
};

SyntheticStock make_synthetic_stock(uint64_t seed) {
  SyntheticStock out{color_neg_200(), {}};
  Stock& st = out.stock;
  st.name = "Synthetic";
  SplitMix rng{seed};
  const double base[3] = {450.0, 550.0, 650.0};
  int e = 0;
  for (int l = 0; l < st.n_layers; ++l) {
    Layer& L = st.layers[l];
    if (L.kind != L_EMULSION) continue;
    const double mu1 = base[e] + 30.0 * (rng.uniform() - 0.5);
    const double s1 = 25.0 + 20.0 * rng.uniform();
    const double mu2 = mu1 + 100.0 * (rng.uniform() - 0.5);
    const double s2 = 20.0 + 30.0 * rng.uniform();
    const double a2 = 0.15 + 0.35 * rng.uniform();
    L.sensitivity = Curve{CURVE_GAUSS2, 1.0, mu1, s1, a2, mu2, s2};
    L.reciprocity_p = 1.0f;
    KnotCurve& K = out.knots[e];
    K.n = 5;
    for (int j = 0; j < 5; ++j) K.log10_phi[j] = -2.5 + (double)j + 0.3 * (rng.uniform() - 0.5);
    double inc[5] = {0.0, 0.0, 0.0, 0.0, 0.0}, total = 0.0;
    for (int j = 1; j < 5; ++j) {
      inc[j] = 0.2 + rng.uniform();
      total += inc[j];
    }
    double run = 0.0;
    for (int j = 0; j < 5; ++j) {
      run += inc[j];
      K.frac[j] = 0.005 + 0.965 * (run / total);
    }
    ++e;
  }
  return out;
}

struct HostStock {
  const Stock* s = nullptr;
  Stock own{};                 // the record itself, for a synthetic stock
  Grid grid = kGridRef;
  FilmDeviceTables t{};        // per-stock part (no per-call fields yet)
  float kappa_ref[kFilmMaxEmulsions] = {};
  float recip_p[kFilmMaxEmulsions] = {};
  float bleed[kFilmMaxEmulsions] = {};       // bleed_e of the sensitivity peak (halation.rs:102-132)
  float bleed_gain[kFilmMaxEmulsions] = {};  // r_e * bleed_e
  bool wide_reflect = false;                 // any backing reflectance > 0 at the four bands
  bool any_reversal = false;
  double dmin_raw[3] = {0, 0, 0};
  double alpha[kFilmMaxLayers] = {};  // emulsion layers: optical density per unit sensitivity (2 rho t)
  int layer_of[kFilmMaxLayers] = {};   // stock layer of each absorbing layer
  double reflect[kFilmMaxBins] = {};  // antihalation backing reflectance samples
  int knots = 0;                      // knot-curve stock: knots per emulsion (0: crystal LUTs)
};

// knots: per emulsion, a characteristic curve replacing the crystal LUT (or null)
HostStock build_stock(const Stock* st, const Grid& g, const KnotCurve* knots = nullptr) {
  HostStock h;
  h.s = st;
  h.grid = g;
  FilmDeviceTables& T = h.t;
  T.n_bins = g.n;
  T.dl = g.dl;
  T.band_inv = 1.0 / (g.lam(g.n - 1) - g.lam(0));
  double cmf[3][kFilmMaxBins];
  cmf_on(g, cmf);
  // upsample basis (upsample.rs:104-143)
  const double peaks[3] = {450.0, 550.0, 650.0};
  double cols[3][3];
  for (int k = 0; k < 3; ++k) {
    for (int i = 0; i < g.n; ++i) {
      const double d = (g.lam(i) - peaks[k]) / 40.0;
      T.basis[k][i] = std::exp(-0.5 * d * d);
    }
    spectrum_ap1(T.basis[k], g, cmf, cols[k]);
  }
  double m[3][3];
  for (int r = 0; r < 3; ++r)
    for (int k = 0; k < 3; ++k) m[r][k] = cols[k][r];
  const double det = m[0][0] * (m[1][1] * m[2][2] - m[1][2] * m[2][1]) - m[0][1] * (m[1][0] * m[2][2] - m[1][2] * m[2][0]) +
                     m[0][2] * (m[1][0] * m[2][1] - m[1][1] * m[2][0]);
  const double id = 1.0 / det;
  T.minv[0][0] = (m[1][1] * m[2][2] - m[1][2] * m[2][1]) * id;
  T.minv[0][1] = (m[0][2] * m[2][1] - m[0][1] * m[2][2]) * id;
  T.minv[0][2] = (m[0][1] * m[1][2] - m[0][2] * m[1][1]) * id;
  T.minv[1][0] = (m[1][2] * m[2][0] - m[1][0] * m[2][2]) * id;
  T.minv[1][1] = (m[0][0] * m[2][2] - m[0][2] * m[2][0]) * id;
  T.minv[1][2] = (m[0][2] * m[1][0] - m[0][0] * m[1][2]) * id;
  T.minv[2][0] = (m[1][0] * m[2][1] - m[1][1] * m[2][0]) * id;
  T.minv[2][1] = (m[0][1] * m[2][0] - m[0][0] * m[2][1]) * id;
  T.minv[2][2] = (m[0][0] * m[1][1] - m[0][1] * m[1][0]) * id;
  for (int i = 0; i < g.n; ++i) {
    T.e_rel[i] = 550.0 / g.lam(i);
    T.inv_e_rel[i] = 1.0 / T.e_rel[i];
  }
  for (int c = 0; c < 3; ++c)
    for (int i = 0; i < g.n; ++i) T.cmf[c][i] = cmf[c][i];
  for (int r = 0; r < 3; ++r)
    for (int c = 0; c < 3; ++c) T.xyz2ap1[r][c] = kXyz2Ap1[r][c];
  sample_curve(st->scanner_light, g, T.scanner);
  double* reflect = h.reflect;
  sample_curve(st->ah_reflectance, g, reflect);
  const double bands[4] = {450.0, 550.0, 650.0, 700.0};
  float max_r = 0.0f;
  for (double b : bands) max_r = std::max(max_r, (float)eval_curve(reflect, g, b));
  h.wide_reflect = max_r > 0.0f;

  int e = 0;
  for (int l = 0; l < st->n_layers; ++l) {
    const Layer& L = st->layers[l];
    if (L.is_reversal) h.any_reversal = true;
    if (L.kind == L_SUPPORT || L.sensitivity.kind == CURVE_NONE) continue;  // no absorption (absorption.rs:68-70)
    double curve[kFilmMaxBins];
    sample_curve(L.sensitivity, g, curve);
    const int a = T.n_absorbing++;
    h.layer_of[a] = l;
    const double thickness = (double)L.thickness_um;
    for (int i = 0; i < g.n; ++i) {
      const double od = L.kind == L_EMULSION ? (curve[i] * 2.0) * (double)L.silver_halide_fraction * thickness
                                             : curve[i] * thickness;
      T.trans[a][i] = std::exp(-od);
    }
    T.is_emulsion[a] = L.kind == L_EMULSION;
    if (L.kind != L_EMULSION) continue;
    h.alpha[a] = 2.0 * (double)L.silver_halide_fraction * thickness;
    FilmEmulsionTables& E = T.em[e];
    const double mu = std::log(L.crystal_mean_um), sg = L.crystal_sigma_ln;
    if (knots && knots[e].n > 0) {
      h.knots = knots[e].n;
      E.lut_n = knots[e].n;
      for (int j = 0; j < E.lut_n; ++j) {
        E.lut_log[j] = knots[e].log10_phi[j];
        E.lut_frac[j] = knots[e].frac[j];
      }
    } else {
      E.lut_n = 64;
      build_lut(mu, sg, L.capture_k, E.lut_log, E.lut_frac);
    }
    sample_curve(L.dye_epsilon, g, E.eps);
    E.has_mask = L.mask_epsilon.kind != CURVE_NONE;
    sample_curve(L.mask_epsilon, g, E.mask_eps);
    E.reversal = L.is_reversal;
    E.d_max = L.d_max;
    E.inv_gamma = 1.0f / std::max(L.gamma_contrast, 1e-6f);
    E.mask_scale = L.d_max * 0.4f;  // MASK_DENSITY_FRACTION_OF_DMAX (constants.rs:67)
    h.recip_p[e] = L.reciprocity_p;
    // grain kappa at 1 um pitch (stock/mod.rs:193-211)
    const double mean_s = std::exp(mu + 0.5 * sg * sg);
    const double r = std::max(mean_s * 0.5, 1e-6);
    const double vol = 1.04719755119659774615 * 4.0 * r * r * r;
    const double rho = (double)L.silver_halide_fraction * thickness / std::max(vol, 1e-18);
    h.kappa_ref[e] = (float)(1.0 / std::sqrt(rho));
    // halation bleed (halation.rs:102-132) and effective reflectance (155-161)
    double peak = 550.0, mx = -INFINITY;
    for (int i = 0; i < g.n; ++i)
      if (curve[i] > mx) { mx = curve[i]; peak = g.lam(i); }
    const float wb = 0.20f, wg = 0.45f, wr = 1.0f;
    float bleed;
    if (peak <= 450.0) bleed = wb;
    else if (peak <= 550.0) bleed = wb + (wg - wb) * (float)((peak - 450.0) / 100.0);
    else if (peak <= 650.0) bleed = wg + (wr - wg) * (float)((peak - 550.0) / 100.0);
    else bleed = wr;
    const double norm = trapz(curve, g);
    float r_e;
    if (norm <= 1e-9) {
      r_e = (float)eval_curve(reflect, g, 650.0);
    } else {
      double prod[kFilmMaxBins];
      for (int i = 0; i < g.n; ++i) prod[i] = reflect[i] * curve[i];
      r_e = (float)(trapz(prod, g) / norm);
    }
    h.bleed[e] = bleed;
    h.bleed_gain[e] = r_e * bleed;
    ++e;
  }
  T.n_emulsions = e;
  for (int i = 0; i < kFilmMaxEmulsions; ++i)
    for (int j = 0; j < kFilmMaxEmulsions; ++j) {
      const bool v = i < st->n_dir && j < st->n_dir && i < e && j < e;
      T.dir_valid[i][j] = v;
      T.dir[i][j] = v ? st->dir[i][j] : 0.0f;
    }
  // Dmin reference scan: image dye 0, mask at d_max * 0.4 (densitometry.rs:76-105)
  double dens[kFilmMaxBins] = {0};
  for (int k = 0; k < e; ++k) {
    const double dm = T.em[k].has_mask ? (double)(T.em[k].d_max * 0.4f) : 0.0;
    for (int l = 0; l < g.n; ++l) {
      dens[l] += 0.0 * T.em[k].eps[l];
      if (T.em[k].has_mask) dens[l] += dm * T.em[k].mask_eps[l];
    }
  }
  double tt[kFilmMaxBins];
  for (int i = 0; i < g.n; ++i) tt[i] = std::pow(10.0, -dens[i]) * T.scanner[i];
  spectrum_ap1(tt, g, cmf, h.dmin_raw);
  const double pk = std::max(std::max(std::max(h.dmin_raw[0], h.dmin_raw[1]), h.dmin_raw[2]), 1e-12);
  T.dmin_scale = 1.0 / pk;
  return h;
}

// Host tables of a stock on a grid (built once per (stock, bins, synthetic seed)).
const HostStock& host_stock(int idx, int bins = 16, uint64_t stock_seed = 0) {
  static std::mutex mu;
  static std::map<std::string, std::unique_ptr<HostStock>> cache;
  if (bins != 16 && bins != 81) throw Error{IMOPS_E_INVALID, "film: spectral_bins must be 16 or 81"};
  const std::string key = std::to_string(idx) + "/" + std::to_string(bins) + "/" +
                          (idx == SYNTHETIC ? std::to_string(stock_seed) : std::string("-"));
  std::lock_guard<std::mutex> lock(mu);
  auto it = cache.find(key);
  if (it != cache.end()) return *it->second;
  std::unique_ptr<HostStock> h;
  if (idx == SYNTHETIC) {
    const SyntheticStock syn = make_synthetic_stock(stock_seed);
    h = std::make_unique<HostStock>(build_stock(&syn.stock, grid_of(bins), syn.knots));
    h->own = syn.stock;
    h->s = &h->own;
  } else {
    const Stock* st = stock_by_index(idx);
    if (!st) throw Error{IMOPS_E_UNSUPPORTED, "film stock tables not available on the GPU path"};
    h = std::make_unique<HostStock>(build_stock(st, grid_of(bins)));
  }
  return *(cache[key] = std::move(h));
}

// Reflect-101 Gaussian taps of film/blur.rs:39-55.
std::vector<float> gauss_taps(float sigma, int* radius) {
  const int r = (int)std::max(std::ceil(3.0f * sigma), 1.0f);
  std::vector<float> k(2 * r + 1);
  const float inv_2s2 = 1.0f / (2.0f * sigma * sigma);
  float sum = 0.0f;
  for (int i = 0; i <= 2 * r; ++i) {
    const float x = (float)i - (float)r;
    const float v = expf(-x * x * inv_2s2);
    k[i] = v;
    sum += v;
  }
  for (auto& v : k) v /= sum;
  *radius = r;
  return k;
}

float format_width_mm(int f) { return f == 1 ? 56.0f : (f == 2 ? 101.6f : 36.0f); }

// Device buffers of one film run.
struct FilmBuffers {
  float* absorbed;   // E planes (absorbed fluence, then unused)
  float* image;      // E planes image dye
  float* mask;       // E planes mask dye
  float* inh;        // E planes (DIR inhibitors)
  float* tmp;        // blur scratch
  float* bounce;     // halation bounce / grain noise
  double* partial;   // grain variance partials + norm
};

unsigned bits_of(float f) {
  unsigned u;
  std::memcpy(&u, &f, sizeof u);
  return u;
}

// Per-shape constants of a film run: blur taps and the device tables with the
// per-call fields (reciprocity, bounce gains, PositiveLinear inversion) filled in.
struct FilmPrepared {
  imops::DevBuf taps, tables;
  FitExposeConsts ex{};  // the exposure collapsed over the spectral grid (collapse_exposure)
  int radius[5] = {}, off[5] = {};
  float sigmas[5] = {};
  bool do_wide = false;
  FilmDeviceTables T{};  // host copy of what `tables` holds
  ~FilmPrepared() {
    taps.release();
    tables.release();
  }
};

// Taps and per-call table fields at this pitch and shutter (radiance.rs:43-49,
// halation.rs:24-93, diffusion.rs, grain.rs:53-100); uploads both and waits
// (the host sources are locals).
// Halation extension: hal_radius >= 0 sets the wide blur's radius in pixels
// (sigma = radius / 3, so the kernel's ceil(3 sigma) radius is that radius) and
// hal_strength >= 0 its gain strength * bleed_e, applied whatever the backing's
// reflectance; negative values keep the reference's psf_halation / pitch and
// r_e * bleed_e.
// The exposure (exposure/mod.rs:45-137) is linear in the three basis weights:
// emulsion e absorbs band_inv * sum_i w_i s(i)/e_rel(i) * (transmittance of the
// layers above it)(i) * (1 - its own transmittance)(i), with s(i) = sum_k
// w_k basis_k(i) and trapezoid weights w_i, so it is a 3 x E matrix per stock
// and grid, formed here in f64; the device evaluates relu(M^-1 rgb) . C per
// pixel in f64 (film_fit_expose), which differs from the per-bin walk only by
// the order of f64 roundings.
FitExposeConsts collapse_exposure(const FilmDeviceTables& T) {
  FitExposeConsts k{};
  for (int m = 0; m < 3; ++m)
    for (int c = 0; c < 3; ++c) k.minv[m][c] = T.minv[m][c];
  k.n_emulsions = T.n_emulsions;
  const int nb = T.n_bins;
  std::vector<double> above(nb, 1.0);
  int e = 0;
  for (int l = 0; l < T.n_absorbing; ++l) {
    if (T.is_emulsion[l]) {
      for (int kk = 0; kk < 3; ++kk) {
        double acc = 0.0;
        for (int i = 0; i < nb; ++i) {
          const double wi = (i == 0 || i == nb - 1) ? 0.5 * T.dl : T.dl;
          acc += wi * T.basis[kk][i] * T.inv_e_rel[i] * above[i] * (1.0 - T.trans[l][i]);
        }
        k.c[kk][e] = acc * T.band_inv;
      }
      ++e;
    }
    for (int i = 0; i < nb; ++i) above[i] *= T.trans[l][i];
  }
  return k;
}

// The FP32 scan's tables (scan_f32_kernel): log2(10)-scaled absorptivities and
// scanner x CMF x trapezoid weight per bin.
void scan_f32_tables(FilmDeviceTables& T) {
  const int nb = T.n_bins;
  const double l2_10 = 3.321928094887362347870;
  for (int e = 0; e < kFilmMaxEmulsions; ++e)
    for (int i = 0; i < kFilmMaxBins; ++i) {
      const bool on = e < T.n_emulsions && i < nb;
      T.scan_eps2[e][i] = on ? (float)(T.em[e].eps[i] * l2_10) : 0.0f;
      T.scan_mask2[e][i] = on && T.em[e].has_mask ? (float)(T.em[e].mask_eps[i] * l2_10) : 0.0f;
    }
  for (int c = 0; c < 3; ++c)
    for (int i = 0; i < kFilmMaxBins; ++i) {
      const double wi = i >= nb ? 0.0 : ((i == 0 || i == nb - 1) ? 0.5 * T.dl : T.dl);
      T.scan_w[c][i] = (float)(T.scanner[i] * T.cmf[c][i] * wi);
    }
  for (int r = 0; r < 3; ++r)
    for (int c = 0; c < 3; ++c) T.scan_xyz2ap1[r][c] = (float)T.xyz2ap1[r][c];
}

void prepare_film(hipStream_t s, const HostStock& hs, const FilmDeviceTables& base, float pitch, float shutter,
                  const FilmStageParams& fp, FilmPrepared& P) {
  const Stock& st = *hs.s;
  P.T = base;
  P.ex = collapse_exposure(P.T);
  scan_f32_tables(P.T);
  const int E = P.T.n_emulsions;
  const float sigma_local = std::max(st.psf_local_um / std::max(pitch, 1e-6f), 0.0f);
  const float sigma_wide = fp.hal_radius >= 0.0f ? fp.hal_radius / 3.0f
                                                 : std::max(st.psf_halation_um / std::max(pitch, 1e-6f), 0.0f);
  const bool reflects = fp.hal_strength >= 0.0f ? fp.hal_strength > 0.0f : hs.wide_reflect;
  P.do_wide = sigma_wide >= 1e-3f && reflects && E > 0;
  for (int e = 0; e < E; ++e) {
    const double p = (double)hs.recip_p[e];
    P.T.em[e].eta = ((double)shutter <= 1.0 || p >= 0.9999) ? 1.0f
                                                            : (float)std::pow(std::max((double)shutter, 1.0), p - 1.0);
    const float gain = fp.hal_strength >= 0.0f ? fp.hal_strength * hs.bleed[e] : hs.bleed_gain[e];
    P.T.em[e].gain = P.do_wide ? gain : 0.0f;
  }
  // blur tap sets: local scatter, halation, DIR, adjacency, grain
  P.sigmas[0] = sigma_local;
  P.sigmas[1] = sigma_wide;
  P.sigmas[2] = st.dir_diffusion_um / std::max(pitch, 1e-6f);
  P.sigmas[3] = st.developer_diffusion_um / std::max(pitch, 1e-6f);
  P.sigmas[4] = std::max(3.0f / std::max(pitch, 1e-6f), 1.0f);
  std::vector<float> all;
  for (int k = 0; k < 5; ++k) {
    P.off[k] = (int)all.size();
    if (P.sigmas[k] >= 1e-3f) {
      auto t = gauss_taps(P.sigmas[k], &P.radius[k]);
      all.insert(all.end(), t.begin(), t.end());
    } else {
      P.radius[k] = 0;
    }
  }
  P.taps.reserve(std::max<size_t>(all.size(), 1) * sizeof(float));
  P.tables.reserve(sizeof(FilmDeviceTables));
  IMOPS_HIP_CHECK(hipMemcpyAsync(P.taps.ptr, all.data(), all.size() * sizeof(float), hipMemcpyHostToDevice, s));
  IMOPS_HIP_CHECK(hipMemcpyAsync(P.tables.ptr, &P.T, sizeof(P.T), hipMemcpyHostToDevice, s));
  IMOPS_HIP_CHECK(hipStreamSynchronize(s));  // `all` is a local
}

// Expose -> develop -> scan (NegativeLinear), or with the inversion when the
// prepared tables carry it.  Kernels only: nothing is uploaded or awaited.
// fp32: the mixed form's FP32 per-pixel spectral arithmetic (expose, capture LUT, scan)
void run_film(hipStream_t s, const HostStock& hs, const FilmPrepared& P, const float* rgb_in, float* rgb_out, int w,
              int h, float pitch, uint64_t seed, FilmBuffers& B, const FilmStageParams& fp, bool fp32 = false) {
  const size_t n = (size_t)w * h;
  const Stock& st = *hs.s;
  const int E = P.T.n_emulsions;
  const float* taps = P.taps.as<float>();
  const FilmDeviceTables* T = P.tables.as<FilmDeviceTables>();
  const float sigma_local = P.sigmas[0], sig_dir = P.sigmas[2], sig_adj = P.sigmas[3];

  const int bins = P.T.n_bins;
  (void)bins;
  IMOPS_HIP_CHECK(film_fit_expose(rgb_in, n, P.ex, B.absorbed, s, 1));
  if (sigma_local >= 1e-3f)  // local gelatin scatter, mix 0.18 (halation.rs:38-44)
    for (int e = 0; e < E; ++e) {
      float* p = B.absorbed + (size_t)e * n;
      IMOPS_HIP_CHECK(film_blur(p, B.tmp, p, p, w, h, taps + P.off[0], P.radius[0], BLUR_MIX, 0.18f, s));
    }
  if (P.do_wide) {
    const float* deep = B.absorbed + (size_t)(E - 1) * n;
    IMOPS_HIP_CHECK(film_blur(deep, B.tmp, deep, B.bounce, w, h, taps + P.off[1], P.radius[1], BLUR_STORE, 0.0f, s));
  }
  IMOPS_HIP_CHECK(film_capture_reduce(B.absorbed, P.do_wide ? B.bounce : nullptr, n, T, B.image, B.mask, s, 1, fp32));
  // DIR inhibition (diffusion.rs:39-84)
  if (E > 0 && st.n_dir > 0 && sig_dir >= 1e-3f) {
    for (int e = 0; e < E; ++e) {
      const float* p = B.image + (size_t)e * n;
      float* q = B.inh + (size_t)e * n;
      IMOPS_HIP_CHECK(film_blur(p, B.tmp, p, q, w, h, taps + P.off[2], P.radius[2], BLUR_STORE, 0.0f, s));
    }
    IMOPS_HIP_CHECK(film_dir_apply(B.image, B.inh, n, T, s));
  }
  // adjacency (diffusion.rs:11-28)
  if (!(std::fabs(st.adjacency_beta) < 1e-8f || sig_adj < 1e-3f))
    for (int e = 0; e < E; ++e) {
      float* p = B.image + (size_t)e * n;
      IMOPS_HIP_CHECK(film_blur(p, B.tmp, p, p, w, h, taps + P.off[3], P.radius[3], BLUR_ADJ, st.adjacency_beta, s));
    }
  // grain (grain.rs:53-100)
  // (grain_sigma >= 0: that constant density sigma; philox: Philox4x32-10 normals)
  const bool absolute = fp.grain_sigma >= 0.0f;
  for (int e = 0; e < E; ++e) {
    const float d_max = P.T.em[e].d_max;
    const float kappa = absolute ? fp.grain_sigma : (hs.kappa_ref[e] / std::max(pitch, 1e-6f)) * 0.15f;
    if (kappa <= 0.0f || d_max <= 0.0f) continue;
    const uint64_t key0 = seed * 0xD1B54A32D192ED03ull + (uint64_t)e * 0x9E3779B97F4A7C15ull;
    IMOPS_HIP_CHECK(film_grain(B.image + (size_t)e * n, B.bounce, B.tmp, w, h, key0, taps + P.off[4], P.radius[4],
                               B.partial, d_max, kappa, s, fp.philox, seed, e, absolute));
  }
  IMOPS_HIP_CHECK(film_scan(B.image, B.mask, n, T, rgb_out, s, 1, fp32, bins));
}

// Prepared runs of a context, by (stock, format, width, shutter, output).
struct FilmCache : imops::StageCache {
  std::vector<std::pair<std::string, std::unique_ptr<FilmPrepared>>> runs;
};

}  // namespace

FilmStageParams film_stage_params_from_table(const std::string& stock, const std::string& format, uint64_t seed,
                                             const std::string& output) {
  static const char* stocks[] = {"BwStub", "ColorNeg200", "Portra400", "Ektar100", "FujiPro400H", "EktachromeE100",
                                 "TriX400", "Synthetic"};
  FilmStageParams p;
  p.stock = -1;
  for (int i = 0; i < 8; ++i)
    if (stock == stocks[i]) p.stock = i;
  if (p.stock < 0) throw Error{IMOPS_E_INVALID, "Unknown film stock: " + stock};
  if (format == "Film35mm") p.format = 0;
  else if (format == "Film6x6") p.format = 1;
  else if (format == "Film4x5") p.format = 2;
  else throw Error{IMOPS_E_INVALID, "Unknown film format: " + format};
  if (output == "NegativeLinear") p.output = 0;
  else if (output == "PositiveLinear") p.output = 1;
  else throw Error{IMOPS_E_INVALID, "Unknown film output: " + output};
  p.seed = seed;
  return p;
}

void stage_film(imops_ctx* ctx, imops_image* img, const FilmStageParams& p) {
  require_rgb(img, "Film");
  TimedRegion stage_timer(ctx, "film");
  if (img->color_space != IMOPS_CS_ACESCG)
    throw Error{IMOPS_E_STATE, "film: wrong color space: expected ACEScg"};  // film/mod.rs:78-87
  const int w = (int)img->width, h = (int)img->height;
  if (w <= 0 || h <= 0) throw Error{IMOPS_E_INVALID, "film: invalid dimensions"};
  if (p.stock < 0 || p.stock > SYNTHETIC) throw Error{IMOPS_E_INVALID, "film: unknown stock"};
  const HostStock& hs = host_stock(p.stock, p.bins, p.stock_seed);
  if (hs.t.n_emulsions > kFilmMaxEmulsions) throw Error{IMOPS_E_UNSUPPORTED, "film: too many emulsion layers"};
  hipStream_t s = ctx->stream;
  const size_t n = (size_t)w * h;
  const int E = hs.t.n_emulsions;
  const float pitch = format_width_mm(p.format) * 1000.0f / (float)w;  // types.rs:45-48
  const float shutter = img->has_shutter ? img->shutter : 1.0f / hs.s->box_iso;

  // scratch: 4E planes + tmp + bounce (+ the 32x32 mid-gray run), tables, partials
  const size_t mid_n = 32 * 32;
  const size_t planes = (size_t)(4 * E + 2);
  ctx->scr[SCR_FILM_0].reserve(planes * n * sizeof(float));
  ctx->scr[SCR_FILM_1].reserve(planes * mid_n * sizeof(float) + 2 * mid_n * 3 * sizeof(float));
  ctx->scr[SCR_FILM_2].reserve((kFilmReduceBlocks + 1) * sizeof(double));
  auto layout = [&](float* base, size_t np, FilmBuffers& B) {
    B.absorbed = base;
    B.image = base + (size_t)E * np;
    B.mask = base + (size_t)2 * E * np;
    B.inh = base + (size_t)3 * E * np;
    B.tmp = base + (size_t)4 * E * np;
    B.bounce = base + (size_t)(4 * E + 1) * np;
    B.partial = reinterpret_cast<double*>(ctx->scr[SCR_FILM_2].ptr);
  };
  FilmBuffers B;
  layout(ctx->scr[SCR_FILM_0].as<float>(), n, B);

  // prepared constants of this shape (computed once per context)
  char keybuf[256];
  snprintf(keybuf, sizeof keybuf, "%d/%d/%d/%08x/%d/%d/%llu/%08x/%08x", p.stock, p.format, w, bits_of(shutter),
           p.output, p.bins, (unsigned long long)p.stock_seed, bits_of(p.hal_radius), bits_of(p.hal_strength));
  const std::string key(keybuf);
  if (!ctx->film_cache) ctx->film_cache.reset(new FilmCache());
  FilmCache& cache = static_cast<FilmCache&>(*ctx->film_cache);
  FilmPrepared* P = nullptr;
  for (auto& r : cache.runs)
    if (r.first == key) P = r.second.get();
  if (!P) {
    if (ctx->capturing) throw Error{IMOPS_E_STATE, "film: shape not prepared before stream capture"};
    if (cache.runs.size() >= 64 && !ctx->in_burst) cache.runs.clear();  // a burst keeps every shape it prepared
    // prepared in full before it enters the cache: a failure part-way (an
    // allocation, a copy) must not leave an entry with missing device tables
    std::unique_ptr<FilmPrepared> fresh(new FilmPrepared());
    P = fresh.get();
    FilmDeviceTables T = hs.t;
    T.invert = 0;
    prepare_film(s, hs, T, pitch, shutter, p, *P);
    if (p.output == 1 && !hs.any_reversal) {
      // PositiveLinear: Dmin of the film base and the mean of a 32x32 flat
      // mid-gray run with seed 0 at this image's pitch and shutter (film/mod.rs:110-121, 136-151).
      FilmBuffers M;
      float* mb = ctx->scr[SCR_FILM_1].as<float>();
      layout(mb, mid_n, M);
      float* gray = mb + planes * mid_n;
      float* gout = gray + mid_n * 3;
      const float iso = std::max(hs.s->box_iso, 1.0f);
      const float gv = (float)(12.5 * (double)0.185f * (8.0 * 8.0) / ((double)(1.0f / iso) * (double)iso));
      std::vector<float> hgray(mid_n * 3, gv), hout(mid_n * 3);
      IMOPS_HIP_CHECK(hipMemcpyAsync(gray, hgray.data(), hgray.size() * sizeof(float), hipMemcpyHostToDevice, s));
      run_film(s, hs, *P, gray, gout, 32, 32, pitch, 0, M, p);
      IMOPS_HIP_CHECK(hipMemcpyAsync(hout.data(), gout, hout.size() * sizeof(float), hipMemcpyDeviceToHost, s));
      IMOPS_HIP_CHECK(hipStreamSynchronize(s));
      double acc[3] = {0, 0, 0};
      for (size_t q = 0; q < mid_n; ++q)
        for (int c = 0; c < 3; ++c) acc[c] += (double)hout[q * 3 + c];
      float mid[3], dmin[3];
      const double pk = std::max(std::max(std::max(hs.dmin_raw[0], hs.dmin_raw[1]), hs.dmin_raw[2]), 1e-12);
      for (int c = 0; c < 3; ++c) {
        mid[c] = (float)(acc[c] / (double)mid_n);
        dmin[c] = (float)(hs.dmin_raw[c] / pk);
      }
      // invert constants (invert.rs:41-77)
      const float eps = 1e-6f;
      for (int c = 0; c < 3; ++c) {
        const float dm = std::fmax(dmin[c], eps);
        float mt = mid[c] / dm;
        mt = mt < eps ? eps : (mt > 1.0f - eps ? 1.0f - eps : mt);
        const float inv_mid = std::fmax(1.0f / mt - 1.0f, 0.0f);
        const float sh = std::fmax(inv_mid * 8.0f, eps);
        const float iv = std::fmax(inv_mid, 0.0f);
        const float soft = std::fmax(iv / (1.0f + iv / std::fmax(sh, 1e-6f)), eps);
        P->T.inv_dmin[c] = dm;
        P->T.inv_shoulder[c] = sh;
        P->T.inv_gain[c] = 0.185f / soft;
      }
      P->T.invert = 1;
      IMOPS_HIP_CHECK(hipMemcpyAsync(P->tables.ptr, &P->T, sizeof(P->T), hipMemcpyHostToDevice, s));
      IMOPS_HIP_CHECK(hipStreamSynchronize(s));
    }
    cache.runs.emplace_back(key, std::move(fresh));
  }
  // the scan writes straight into the image: every pixel's input was consumed by expose
  // the 32x32 mid-gray calibration above stays FP64 in both forms
  run_film(s, hs, *P, img->rgb.as<float>(), img->rgb.as<float>(), w, h, pitch, p.seed, B, p, p.fp32);
  img->color_space = IMOPS_CS_ACESCG;
}

// ---------------------------------------------------------------- film fit
// The differentiable film model: the film stage's expose -> local scatter ->
// halation -> capture -> reduction -> scan (NegativeLinear, shutter 1/box ISO)
// with grain, DIR inhibition and adjacency off, the emulsion sensitivities,
// D_max and gamma as parameters, the halation gain and PSF width scaled by two
// more, and the scan normalised by the nominal stock's Dmin.
namespace {

struct FitSetup {
  FilmDeviceTables T;                 // stock tables with the parameter-dependent fields
  FitExposeConsts k;
  int nb = 0;                         // bins of the spectral grid
  double tw[kFilmMaxBins];            // trapezoid weights of the grid
  double W[3][kFilmMaxBins];          // trapezoid weight * basis_k / e_rel / band
  double tab[kFilmMaxEmulsions][kFilmMaxBins];  // transmittance of every absorbing layer above emulsion e
  double te[kFilmMaxEmulsions][kFilmMaxBins];   // transmittance of emulsion e
  double alpha[kFilmMaxEmulsions];
  int above[kFilmMaxEmulsions][kFilmMaxEmulsions];  // [e'][e]: emulsion e' lies above emulsion e
  float rb[kFilmMaxEmulsions];        // r_e * bleed_e
  double bleed[kFilmMaxEmulsions], rnum[kFilmMaxEmulsions], rden[kFilmMaxEmulsions];
  float sigma_l, sigma_h, sigma_h_unit;
  bool do_local, do_wide;
};

FitSetup fit_setup(const HostStock& hs, const imops_film_fit_params& p, float pitch) {
  FitSetup f{};
  const Stock& st = *hs.s;
  const Grid& g = hs.grid;
  const int E = hs.t.n_emulsions, nb = g.n;
  if (p.n_emulsions != E) throw Error{IMOPS_E_INVALID, "film_fit: n_emulsions does not match the stock"};
  if (p.n_bins != nb) throw Error{IMOPS_E_INVALID, "film_fit: n_bins does not match the model's spectral grid"};
  if (p.n_knots != hs.knots)
    throw Error{IMOPS_E_INVALID, "film_fit: n_knots must be the stock's knot count (0 for crystal LUTs)"};
  f.nb = nb;
  f.T = hs.t;
  f.T.invert = 0;
  for (int i = 0; i < nb; ++i) f.tw[i] = (i == 0 || i == nb - 1) ? 0.5 * g.dl : g.dl;
  // sensitivities -> per-layer transmittances, C matrix (exposure/mod.rs:68-81 collapsed over the basis)
  double trans[kFilmMaxLayers][kFilmMaxBins];
  int emul_of_layer[kFilmMaxLayers];
  int e = 0;
  for (int a = 0; a < hs.t.n_absorbing; ++a) {
    emul_of_layer[a] = -1;
    if (!hs.t.is_emulsion[a]) {
      for (int i = 0; i < nb; ++i) trans[a][i] = hs.t.trans[a][i];
      continue;
    }
    const Layer& L = st.layers[hs.layer_of[a]];
    for (int i = 0; i < nb; ++i)
      trans[a][i] = std::exp(-((double)p.sens[e][i] * 2.0) * (double)L.silver_halide_fraction * (double)L.thickness_um);
    f.alpha[e] = hs.alpha[a];
    emul_of_layer[a] = e++;
  }
  for (int k = 0; k < 3; ++k)
    for (int i = 0; i < nb; ++i) f.W[k][i] = f.tw[i] * hs.t.basis[k][i] / hs.t.e_rel[i] * hs.t.band_inv;
  double above_t[kFilmMaxBins];
  for (int i = 0; i < nb; ++i) above_t[i] = 1.0;
  for (int a = 0; a < hs.t.n_absorbing; ++a) {
    const int em = emul_of_layer[a];
    if (em >= 0) {
      for (int i = 0; i < nb; ++i) {
        f.tab[em][i] = above_t[i];
        f.te[em][i] = trans[a][i];
      }
      for (int e2 = 0; e2 < em; ++e2) f.above[e2][em] = 1;
    }
    for (int i = 0; i < nb; ++i) above_t[i] *= trans[a][i];
  }
  for (int m = 0; m < 3; ++m)
    for (int c = 0; c < 3; ++c) f.k.minv[m][c] = hs.t.minv[m][c];
  f.k.n_emulsions = E;
  for (int k = 0; k < 3; ++k)
    for (int em = 0; em < E; ++em) {
      double acc = 0.0;
      for (int i = 0; i < nb; ++i) acc += f.W[k][i] * f.tab[em][i] * (1.0 - f.te[em][i]);
      f.k.c[k][em] = acc;
    }
  // halation (halation.rs:24-132): bleed from the sensitivity peak, reflectance weighted by the sensitivity
  f.sigma_l = std::max(st.psf_local_um / std::max(pitch, 1e-6f), 0.0f);
  f.sigma_h_unit = std::max(st.psf_halation_um / std::max(pitch, 1e-6f), 0.0f);
  f.sigma_h = (float)((double)f.sigma_h_unit * (double)p.hal_radius);
  f.do_local = f.sigma_l >= 1e-3f;
  f.do_wide = f.sigma_h >= 1e-3f && hs.wide_reflect && E > 0;
  for (int em = 0; em < E; ++em) {
    double curve[kFilmMaxBins];
    for (int i = 0; i < nb; ++i) curve[i] = p.sens[em][i];
    double peak = 550.0, mx = -INFINITY;
    for (int i = 0; i < nb; ++i)
      if (curve[i] > mx) { mx = curve[i]; peak = g.lam(i); }
    const float wb = 0.20f, wg = 0.45f, wr = 1.0f;
    float bl;
    if (peak <= 450.0) bl = wb;
    else if (peak <= 550.0) bl = wb + (wg - wb) * (float)((peak - 450.0) / 100.0);
    else if (peak <= 650.0) bl = wg + (wr - wg) * (float)((peak - 550.0) / 100.0);
    else bl = wr;
    const double norm = trapz(curve, g);
    float r;
    double prod[kFilmMaxBins];
    for (int i = 0; i < nb; ++i) prod[i] = hs.reflect[i] * curve[i];
    f.rnum[em] = trapz(prod, g);
    f.rden[em] = norm;
    r = norm <= 1e-9 ? (float)eval_curve(hs.reflect, g, 650.0) : (float)(f.rnum[em] / norm);
    f.bleed[em] = bl;
    f.rb[em] = r * bl;
    FilmEmulsionTables& Em = f.T.em[em];
    Em.gain = f.do_wide ? (float)((double)f.rb[em] * (double)p.hal_strength) : 0.0f;
    Em.eta = 1.0f;  // shutter 1/box ISO <= 1 s: no reciprocity failure
    Em.d_max = p.d_max[em];
    Em.inv_gamma = 1.0f / std::max(p.gamma[em], 1e-6f);
    Em.mask_scale = p.d_max[em] * 0.4f;
    for (int j = 0; j < p.n_knots; ++j) Em.lut_frac[j] = p.knot_frac[em][j];  // knot positions stay the stock's
  }
  return f;
}

// d/dsigma of the normalised Gaussian taps of gauss_taps (fixed radius)
std::vector<float> gauss_taps_dsigma(float sigma, int radius) {
  std::vector<double> g(2 * radius + 1), gd(2 * radius + 1);
  double sg = 0.0, sgd = 0.0;
  const double s = sigma;
  for (int i = 0; i <= 2 * radius; ++i) {
    const double x = (double)i - (double)radius;
    g[i] = std::exp(-x * x / (2.0 * s * s));
    gd[i] = g[i] * x * x / (s * s * s);
    sg += g[i];
    sgd += gd[i];
  }
  std::vector<float> out(2 * radius + 1);
  for (int i = 0; i <= 2 * radius; ++i) out[i] = (float)((gd[i] * sg - g[i] * sgd) / (sg * sg));
  return out;
}

// Device planes of one batched forward/backward over N images (image-major:
// image b's E planes of a quantity at base + b E n, its single planes at base + b n).
struct FitRun {
  imops_ctx* ctx;
  const HostStock* hs;
  FitSetup f;
  int w, h, E, N;
  size_t n;
  float* S;      // E planes per image (scratch, or the caller's saved state)
  float* Sbar;   // E planes per image
  float* Bn;     // one plane per image (follows S in the saved state)
  float* Bbar;
  float* tmp;    // E planes per image
  float* u1;
  float* u2;
  float* img;    // E planes per image: image dye
  float* msk;    // E planes per image: mask dye
  float* taps;   // local, halation, halation d/dsigma
  int rl, rh, off_l, off_h, off_hd;
  FilmDeviceTables* T;
  double* acc;   // kFitSlots sums: 3E (d_max, gamma, gain) at [0, 8, 16], C at [24 + k*8 + e], sigma at [48]
  double* part;  // rows x kFitSlots partial sums of the reducing kernels
  int rows;
};

size_t fit_state_floats(int emulsions, size_t images, size_t n) { return images * n * (size_t)(emulsions + 1); }

// The film model's per-call tables on the device (its own buffers, no other
// stage writes them), uploaded from pinned staging so the upload does not wait
// for the stream; an upload identical to the last one is skipped (a backward
// pass after its forward).  `done` marks the end of the last copy out of the
// staging, which must pass before the host rewrites it.
struct FitUpload : StageCache {
  DevBuf tables;  // FilmDeviceTables, then kFitSlots sums, then the partial-sum rows
  DevBuf taps;
  void* pinned = nullptr;
  size_t pinned_bytes = 0;
  hipEvent_t done = nullptr;
  std::vector<unsigned char> key;  // what the device tables were built from
  ~FitUpload() override {
    if (done) (void)hipEventDestroy(done);
    if (pinned) (void)hipHostFree(pinned);
    tables.release();
    taps.release();
  }
};

template <class T>
void append_bytes(std::vector<unsigned char>& v, const T& x) {
  const unsigned char* p = reinterpret_cast<const unsigned char*>(&x);
  v.insert(v.end(), p, p + sizeof(T));
}

// The model's stock tables (validated).
const HostStock& fit_stock(const imops_film_fit_model* m) {
  if (!m) throw Error{IMOPS_E_INVALID, "film_fit: null model"};
  if (m->stock < 0 || m->stock > SYNTHETIC) throw Error{IMOPS_E_INVALID, "film_fit: unknown stock"};
  if (m->spectral_bins != 16 && m->spectral_bins != 81)
    throw Error{IMOPS_E_INVALID, "film_fit: spectral_bins must be 16 or 81"};
  const HostStock& hs = host_stock(m->stock, m->spectral_bins, m->stock_seed);
  if (hs.knots > kFitMaxKnots) throw Error{IMOPS_E_UNSUPPORTED, "film_fit: more than 8 curve knots"};
  return hs;
}

FitRun fit_prepare(imops_ctx* ctx, const imops_film_fit_model* m, const imops_film_fit_params& p, size_t w, size_t h,
                   size_t images, const float* state) {
  if (!ctx) throw Error{IMOPS_E_INVALID, "film_fit: null context"};
  const HostStock& hs = fit_stock(m);
  const float pitch = m->pitch_um;
  if (w < 2 || h < 2 || !(pitch > 0.0f)) throw Error{IMOPS_E_INVALID, "film_fit: invalid dimensions or pitch"};
  if (images < 1 || images > 65535 || w > (size_t)INT32_MAX / 2 || h > (size_t)INT32_MAX / 2)
    throw Error{IMOPS_E_INVALID, "film_fit: image count must be 1..65535"};
  FitRun R{};
  R.ctx = ctx;
  R.hs = &hs;
  if (R.hs->any_reversal) throw Error{IMOPS_E_UNSUPPORTED, "film_fit: reversal stocks are not supported"};
  const bool ext = m->spectral_bins != 16 || hs.knots != 0;
  if (ext && (hs.t.n_emulsions > kFitMaxEmulsionsExt || (hs.knots != 0 && hs.knots != 5)))
    throw Error{IMOPS_E_UNSUPPORTED, "film_fit: the 81-bin grid and knot curves take at most 4 emulsions and 5 knots"};
  R.f = fit_setup(*R.hs, p, pitch);
  R.w = (int)w;
  R.h = (int)h;
  R.E = R.hs->t.n_emulsions;
  R.n = w * h;
  R.N = (int)images;
  if (images * (size_t)std::max(R.E, 1) > 65535)
    throw Error{IMOPS_E_INVALID, "film_fit: at most 65535 emulsion planes (images x emulsions) per call"};
  std::vector<float> all;
  R.off_l = R.off_h = R.off_hd = 0;
  R.rl = R.rh = 0;
  if (R.f.do_local) {
    auto t = gauss_taps(R.f.sigma_l, &R.rl);
    R.off_l = (int)all.size();
    all.insert(all.end(), t.begin(), t.end());
  }
  if (R.f.do_wide) {
    auto t = gauss_taps(R.f.sigma_h, &R.rh);
    R.off_h = (int)all.size();
    all.insert(all.end(), t.begin(), t.end());
    auto d = gauss_taps_dsigma(R.f.sigma_h, R.rh);
    R.off_hd = (int)all.size();
    all.insert(all.end(), d.begin(), d.end());
  }
  const int rmax = std::max(R.rl, R.rh);
  if (rmax > (int)std::min(w, h) - 1)
    throw Error{IMOPS_E_UNSUPPORTED, "film_fit: image smaller than the blur support"};
  // planes per image: Sbar, img, msk, tmp (E each), Bbar, u1, u2, plus S and Bn (E + 1) without a state
  const size_t plane = R.n * images;  // one plane for every image
  const size_t np = (size_t)(4 * R.E + 3) + (state ? 0 : (size_t)(R.E + 1));
  ctx->scr[SCR_FILM_0].reserve(np * plane * sizeof(float));
  R.rows = fit_partial_rows(R.n, R.N);
  if (!ctx->fit_cache) ctx->fit_cache.reset(new FitUpload());
  FitUpload& U = static_cast<FitUpload&>(*ctx->fit_cache);
  const size_t tab_bytes = sizeof(FilmDeviceTables) + (size_t)(1 + R.rows) * kFitSlots * sizeof(double);
  const size_t taps_bytes = std::max<size_t>(all.size(), 1) * sizeof(float);
  std::vector<unsigned char> key;
  append_bytes(key, m->stock);
  append_bytes(key, m->spectral_bins);
  append_bytes(key, m->stock_seed);
  append_bytes(key, m->pitch_um);
  append_bytes(key, p);
  const void* old_tables = U.tables.ptr;
  const void* old_taps = U.taps.ptr;
  U.tables.reserve(tab_bytes);
  U.taps.reserve(taps_bytes);
  const bool fresh = key == U.key && U.tables.ptr == old_tables && U.taps.ptr == old_taps;
  float* b = ctx->scr[SCR_FILM_0].as<float>();
  R.Sbar = b;
  R.img = b + (size_t)R.E * plane;
  R.msk = b + (size_t)2 * R.E * plane;
  R.tmp = b + (size_t)3 * R.E * plane;
  R.Bbar = b + (size_t)4 * R.E * plane;
  R.u1 = R.Bbar + plane;
  R.u2 = R.u1 + plane;
  float* sv = state ? const_cast<float*>(state) : R.u2 + plane;  // S then Bn (fit_state_floats)
  R.S = sv;
  R.Bn = sv + (size_t)R.E * plane;
  R.T = U.tables.as<FilmDeviceTables>();
  R.acc = reinterpret_cast<double*>(reinterpret_cast<char*>(U.tables.ptr) + sizeof(FilmDeviceTables));
  R.part = R.acc + kFitSlots;
  R.taps = U.taps.as<float>();
  if (!fresh) {
    hipStream_t s = ctx->stream;
    const size_t need = sizeof(FilmDeviceTables) + all.size() * sizeof(float);
    if (U.done) IMOPS_HIP_CHECK(hipEventSynchronize(U.done));  // the staging is free again
    if (need > U.pinned_bytes) {
      if (U.pinned) IMOPS_HIP_CHECK(hipHostFree(U.pinned));
      U.pinned = nullptr;
      U.pinned_bytes = 0;
      IMOPS_HIP_CHECK(hipHostMalloc(&U.pinned, need, hipHostMallocDefault));
      U.pinned_bytes = need;
    }
    if (!U.done) IMOPS_HIP_CHECK(hipEventCreateWithFlags(&U.done, hipEventDisableTiming));
    char* st = static_cast<char*>(U.pinned);
    std::memcpy(st, &R.f.T, sizeof(FilmDeviceTables));
    if (!all.empty()) std::memcpy(st + sizeof(FilmDeviceTables), all.data(), all.size() * sizeof(float));
    IMOPS_HIP_CHECK(hipMemcpyAsync(R.T, st, sizeof(FilmDeviceTables), hipMemcpyHostToDevice, s));
    if (!all.empty())
      IMOPS_HIP_CHECK(hipMemcpyAsync(R.taps, st + sizeof(FilmDeviceTables), all.size() * sizeof(float),
                                     hipMemcpyHostToDevice, s));
    IMOPS_HIP_CHECK(hipEventRecord(U.done, s));
    U.key = std::move(key);
  }
  return R;
}

// expose + local scatter + halation bounce of all N images into R.S / R.Bn
void fit_front(FitRun& R, const float* rgb) {
  hipStream_t s = R.ctx->stream;
  const size_t n = R.n, En = (size_t)R.E * n;
  IMOPS_HIP_CHECK(film_fit_expose(rgb, n, R.f.k, R.S, s, R.N));
  if (R.f.do_local)  // every emulsion plane of every image in one launch pair
    IMOPS_HIP_CHECK(film_blur(R.S, R.tmp, R.S, R.S, R.w, R.h, R.taps + R.off_l, R.rl, BLUR_MIX, 0.18f, s,
                              BlurPlanes{R.N * R.E, n, n, n}));
  if (R.f.do_wide) {
    const float* deep = R.S + (size_t)(R.E - 1) * n;
    IMOPS_HIP_CHECK(film_blur(deep, R.tmp, deep, R.Bn, R.w, R.h, R.taps + R.off_h, R.rh, BLUR_STORE, 0.0f, s,
                              BlurPlanes{R.N, En, En, n}));
  }
}

}  // namespace

}  // namespace imops

using namespace imops;

extern "C" int imops_film(imops_ctx* ctx, imops_image* img, int stock, int film_format, uint64_t seed, int output) {
  return guarded([&] {
    if (!ctx || !img) throw Error{IMOPS_E_INVALID, "film: null argument"};
    FilmStageParams p;
    p.stock = stock;
    p.format = film_format;
    p.seed = seed;
    p.output = output;
    if (stock < 0 || stock >= STOCK_COUNT || film_format < 0 || film_format > 2 || output < 0 || output > 1)
      throw Error{IMOPS_E_INVALID, "film: stock / format / output out of range"};
    stage_film(ctx, img, p);
  });
}

extern "C" int imops_film_default_params(imops_film_params* out) {
  return guarded([&] {
    if (!out) throw Error{IMOPS_E_INVALID, "film_default_params: null argument"};
    const FilmStageParams d;
    *out = imops_film_params{};
    out->stock = d.stock;
    out->film_format = d.format;
    out->output = d.output;
    out->precision = 0;
    out->spectral_bins = d.bins;
    out->grain = 0;
    out->seed = d.seed;
    out->stock_seed = d.stock_seed;
    out->grain_sigma = d.grain_sigma;
    out->halation_radius = d.hal_radius;
    out->halation_strength = d.hal_strength;
  });
}

extern "C" int imops_film_ex(imops_ctx* ctx, imops_image* img, const imops_film_params* q) {
  return guarded([&] {
    if (!ctx || !img || !q) throw Error{IMOPS_E_INVALID, "film_ex: null argument"};
    if (q->stock < 0 || q->stock > SYNTHETIC || q->film_format < 0 || q->film_format > 2 || q->output < 0 ||
        q->output > 1 || q->precision < 0 || q->precision > 1 || (q->spectral_bins != 16 && q->spectral_bins != 81) ||
        q->grain < 0 || q->grain > 1)
      throw Error{IMOPS_E_INVALID, "film_ex: parameter out of range"};
    FilmStageParams p;
    p.stock = q->stock;
    p.format = q->film_format;
    p.seed = q->seed;
    p.output = q->output;
    p.fp32 = q->precision == 1;
    p.bins = q->spectral_bins;
    p.stock_seed = q->stock_seed;
    p.philox = q->grain == 1;
    p.grain_sigma = q->grain_sigma;
    p.hal_radius = q->halation_radius;
    p.hal_strength = q->halation_strength;
    stage_film(ctx, img, p);
  });
}

extern "C" int imops_film_fit_nominal(const imops_film_fit_model* m, imops_film_fit_params* out) {
  return guarded([&] {
    if (!out) throw Error{IMOPS_E_INVALID, "film_fit_nominal: null argument"};
    const HostStock& hs = fit_stock(m);
    *out = imops_film_fit_params{};
    out->n_emulsions = hs.t.n_emulsions;
    out->n_bins = hs.grid.n;
    out->n_knots = hs.knots;
    for (int a = 0, e = 0; a < hs.t.n_absorbing; ++a) {
      if (!hs.t.is_emulsion[a]) continue;
      const Layer& L = hs.s->layers[hs.layer_of[a]];
      double curve[kFilmMaxBins];
      sample_curve(L.sensitivity, hs.grid, curve);
      for (int i = 0; i < hs.grid.n; ++i) out->sens[e][i] = (float)curve[i];
      for (int j = 0; j < hs.knots; ++j) out->knot_frac[e][j] = (float)hs.t.em[e].lut_frac[j];
      out->d_max[e] = L.d_max;
      out->gamma[e] = L.gamma_contrast;
      ++e;
    }
    out->hal_strength = 1.0f;
    out->hal_radius = 1.0f;
  });
}

extern "C" int imops_film_fit_knots(const imops_film_fit_model* m, int emulsion, double* log10_fluence) {
  return guarded([&] {
    if (!log10_fluence) throw Error{IMOPS_E_INVALID, "film_fit_knots: null argument"};
    const HostStock& hs = fit_stock(m);
    if (emulsion < 0 || emulsion >= hs.t.n_emulsions) throw Error{IMOPS_E_INVALID, "film_fit_knots: no such emulsion"};
    for (int j = 0; j < hs.knots; ++j) log10_fluence[j] = hs.t.em[emulsion].lut_log[j];
  });
}

extern "C" int imops_film_fit_state_floats(const imops_film_fit_model* m, size_t n_images, size_t width, size_t height,
                                           size_t* floats) {
  return guarded([&] {
    if (!floats) throw Error{IMOPS_E_INVALID, "film_fit_state_floats: null argument"};
    *floats = fit_state_floats(fit_stock(m).t.n_emulsions, n_images, width * height);
  });
}

extern "C" int imops_film_fit_forward(imops_ctx* ctx, const imops_film_fit_model* m, const imops_film_fit_params* p,
                                      const float* rgb_in, float* rgb_out, size_t n_images, size_t width,
                                      size_t height, float* state) {
  return guarded([&] {
    if (!p || !rgb_in || !rgb_out) throw Error{IMOPS_E_INVALID, "film_fit_forward: null argument"};
    FitRun R = fit_prepare(ctx, m, *p, width, height, n_images, state);
    hipStream_t s = ctx->stream;
    TimedRegion tr(ctx, "film_fit_forward");
    fit_front(R, rgb_in);
    IMOPS_HIP_CHECK(film_capture_reduce(R.S, R.f.do_wide ? R.Bn : nullptr, R.n, R.T, R.img, R.msk, s, R.N));
    IMOPS_HIP_CHECK(film_scan(R.img, R.msk, R.n, R.T, rgb_out, s, R.N, false, R.f.nb));
  });
}

namespace imops {
namespace {

// The reverse pass of one batched call: dL/dparams into *grad for dL/dout =
// gout, or, with l2_target, for L = sum (out - gout)^2 (its value into *loss).
// front: run the exposure and blurs first (no saved state).
void fit_reverse(FitRun& R, const imops_film_fit_params& p, const float* rgb_in, const float* gout, bool l2_target,
                 bool front, imops_film_fit_params* grad, double* loss) {
  imops_ctx* ctx = R.ctx;
  hipStream_t s = ctx->stream;
  const int E = R.E;
  const size_t n = R.n, En = (size_t)E * n;
  {
    TimedRegion tr(ctx, l2_target ? "film_fit_l2_step" : "film_fit_backward");
    IMOPS_HIP_CHECK(film_fit_zero(R.part, (size_t)R.rows * kFitSlots, s));
    if (front) fit_front(R, rgb_in);
    IMOPS_HIP_CHECK(film_fit_back_scan(R.S, R.f.do_wide ? R.Bn : nullptr, R.T, E, R.f.nb, R.hs->knots, gout, n,
                                       R.Sbar, R.f.do_wide ? R.Bbar : nullptr, R.part, s, R.N, l2_target));
    if (R.f.do_wide) {
      const float* deep = R.S + (size_t)(E - 1) * n;
      float* deep_bar = R.Sbar + (size_t)(E - 1) * n;
      const float* k = R.taps + R.off_h;
      const float* kd = R.taps + R.off_hd;
      // d/dsigma of V(H(x)) = V'(H x) + V(H' x)
      IMOPS_HIP_CHECK(film_blur2(deep, R.tmp, R.u1, R.w, R.h, k, kd, R.rh, s, BlurPlanes{R.N, En, 0, n}));
      IMOPS_HIP_CHECK(film_blur2(deep, R.tmp, R.u2, R.w, R.h, kd, k, R.rh, s, BlurPlanes{R.N, En, 0, n}));
      IMOPS_HIP_CHECK(film_fit_dot(R.Bbar, R.u1, R.u2, n, R.part, s, R.N));
      IMOPS_HIP_CHECK(film_blur_adjoint(R.Bbar, R.tmp, deep_bar, deep_bar, R.w, R.h, k, R.rh, 1.0f, 1.0f, s,
                                        BlurPlanes{R.N, n, En, En}));
    }
    if (R.f.do_local)
      IMOPS_HIP_CHECK(film_blur_adjoint(R.Sbar, R.tmp, R.Sbar, R.Sbar, R.w, R.h, R.taps + R.off_l, R.rl,
                                        1.0f - 0.18f, 0.18f, s, BlurPlanes{R.N * E, n, n, n}));
    IMOPS_HIP_CHECK(film_fit_back_expose(rgb_in, R.Sbar, n, R.f.k, R.part, s, R.N));
    IMOPS_HIP_CHECK(film_fit_reduce(R.part, R.rows, R.acc, s));
  }
  double acc[kFitSlots];
  IMOPS_HIP_CHECK(hipMemcpyAsync(acc, R.acc, sizeof acc, hipMemcpyDeviceToHost, s));
  IMOPS_HIP_CHECK(hipStreamSynchronize(s));
  const FitSetup& f = R.f;
  *grad = imops_film_fit_params{};
  grad->n_emulsions = E;
  grad->n_bins = f.nb;
  grad->n_knots = p.n_knots;
  for (int k = 0; k < p.n_knots; ++k)
    for (int e = 0; e < E; ++e) grad->knot_frac[e][k] = (float)acc[kFitSlotKnot + k * kFilmMaxEmulsions + e];
  double g_hs = 0.0;
  for (int e = 0; e < E; ++e) {
    grad->d_max[e] = (float)acc[e];
    grad->gamma[e] = (float)acc[8 + e];
    g_hs += acc[16 + e] * (double)f.rb[e];
  }
  grad->hal_strength = f.do_wide ? (float)g_hs : 0.0f;
  grad->hal_radius = f.do_wide ? (float)(acc[48] * (double)f.sigma_h_unit) : 0.0f;
  // sensitivities: through the exposure contraction C and the halation reflectance r_e
  for (int e2 = 0; e2 < E; ++e2)
    for (int i = 0; i < f.nb; ++i) {
      double g = 0.0;
      for (int k = 0; k < 3; ++k)
        for (int e = 0; e < E; ++e) {
          const double cbar = acc[24 + k * 8 + e];
          double d = 0.0;
          if (e2 == e) d += f.tab[e][i] * f.alpha[e] * f.te[e][i];
          if (f.above[e2][e]) d -= f.alpha[e2] * f.tab[e][i] * (1.0 - f.te[e][i]);
          g += cbar * f.W[k][i] * d;
        }
      if (f.do_wide && f.rden[e2] > 1e-9) {
        const double dr = f.tw[i] * (R.hs->reflect[i] * f.rden[e2] - f.rnum[e2]) / (f.rden[e2] * f.rden[e2]);
        g += acc[16 + e2] * (double)p.hal_strength * f.bleed[e2] * dr;
      }
      grad->sens[e2][i] = (float)g;
    }
  if (loss) *loss = acc[kFitSlotLoss];
}

}  // namespace
}  // namespace imops

extern "C" int imops_film_fit_backward(imops_ctx* ctx, const imops_film_fit_model* m, const imops_film_fit_params* p,
                                       const float* rgb_in, const float* grad_out, size_t n_images, size_t width,
                                       size_t height, const float* state, imops_film_fit_params* grad) {
  return guarded([&] {
    if (!p || !rgb_in || !grad_out || !grad) throw Error{IMOPS_E_INVALID, "film_fit_backward: null argument"};
    FitRun R = fit_prepare(ctx, m, *p, width, height, n_images, state);
    fit_reverse(R, *p, rgb_in, grad_out, false, !state, grad, nullptr);
  });
}

extern "C" int imops_film_fit_l2_step(imops_ctx* ctx, const imops_film_fit_model* m, const imops_film_fit_params* p,
                                      const float* rgb_in, const float* target, size_t n_images, size_t width,
                                      size_t height, imops_film_fit_params* grad, double* loss) {
  return guarded([&] {
    if (!p || !rgb_in || !target || !grad || !loss) throw Error{IMOPS_E_INVALID, "film_fit_l2_step: null argument"};
    FitRun R = fit_prepare(ctx, m, *p, width, height, n_images, nullptr);
    fit_reverse(R, *p, rgb_in, target, true, true, grad, loss);
  });
}
