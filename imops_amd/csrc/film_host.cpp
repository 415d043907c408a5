// Host side of the film stage: stock tables -> device tables (the derived
// quantities of FilmStock::finalize, pichromatic/src/film/stock/mod.rs:152-214,
// and of the exposure/scan set-up), and the launch sequence of film::process
// (pichromatic/src/film/mod.rs:77-163).  Everything computed here is a
// per-stock or per-call constant evaluated with the host libm exactly as the
// reference evaluates it, so the device kernels only stream pixels.
#include <algorithm>
#include <cmath>
#include <cstring>
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

constexpr int NB = 16;
double lam(int i) { return 400.0 + 20.0 * i; }

void sample_curve(const Curve& c, double* out) {
  for (int i = 0; i < NB; ++i) {
    const double l = lam(i);
    if (c.kind == CURVE_CONST) {
      out[i] = c.a;
    } else if (c.kind == CURVE_GAUSS) {  // gaussian_curve (color_neg_200.rs:38-50)
      const double d = (l - c.b) / c.c;
      out[i] = c.a * std::exp(-0.5 * d * d);
    } else if (c.kind == CURVE_ORANGE) {  // orange_mask_epsilon (color_neg_200.rs:52-66)
      const double u = (l - 450.0) / 50.0, v = (l - 520.0) / 60.0;
      out[i] = 0.35 * std::exp(-0.5 * (u * u)) + 0.25 * (c.a * std::exp(-0.5 * (v * v))) + 0.05;
    } else {
      out[i] = 0.0;
    }
  }
}

double trapz(const double* s) {  // SpectralCurve::integrate (spectrum.rs:106-118)
  double acc = 0.0;
  for (int i = 0; i < NB - 1; ++i) acc += 0.5 * (s[i] + s[i + 1]) * (lam(i + 1) - lam(i));
  return acc;
}

double eval_curve(const double* s, double l) {  // SpectralCurve::evaluate (spectrum.rs:60-78)
  if (l <= lam(0)) return s[0];
  if (l >= lam(NB - 1)) return s[NB - 1];
  for (int i = 0; i < NB - 1; ++i)
    if (l >= lam(i) && l <= lam(i + 1)) {
      const double t = (l - lam(i)) / (lam(i + 1) - lam(i));
      return s[i] * (1.0 - t) + s[i + 1] * t;
    }
  return s[NB - 1];
}

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

void spectrum_ap1(const double* s, double out[3]) {
  double xyz[3];
  for (int c = 0; c < 3; ++c) {
    double acc = 0.0;
    for (int i = 0; i < 15; ++i) {
      const double a = s[i] * kCmf[c][i], b = s[i + 1] * kCmf[c][i + 1];
      acc += 0.5 * (a + b) * 20.0;
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

struct HostStock {
  const Stock* s = nullptr;
  FilmDeviceTables t{};        // per-stock part (no per-call fields yet)
  float kappa_ref[kFilmMaxEmulsions] = {};
  float recip_p[kFilmMaxEmulsions] = {};
  float bleed_gain[kFilmMaxEmulsions] = {};  // r_e * bleed_e
  bool wide_reflect = false;                 // any backing reflectance > 0 at the four bands
  bool any_reversal = false;
  double dmin_raw[3] = {0, 0, 0};
};

HostStock build_stock(const Stock* st) {
  HostStock h;
  h.s = st;
  FilmDeviceTables& T = h.t;
  // upsample basis (upsample.rs:104-143)
  const double peaks[3] = {450.0, 550.0, 650.0};
  double cols[3][3];
  for (int k = 0; k < 3; ++k) {
    for (int i = 0; i < NB; ++i) {
      const double d = (lam(i) - peaks[k]) / 40.0;
      T.basis[k][i] = std::exp(-0.5 * d * d);
    }
    spectrum_ap1(T.basis[k], cols[k]);
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
  for (int i = 0; i < NB; ++i) T.e_rel[i] = 550.0 / lam(i);
  for (int c = 0; c < 3; ++c)
    for (int i = 0; i < NB; ++i) T.cmf[c][i] = kCmf[c][i];
  for (int r = 0; r < 3; ++r)
    for (int c = 0; c < 3; ++c) T.xyz2ap1[r][c] = kXyz2Ap1[r][c];
  sample_curve(st->scanner_light, T.scanner);
  double reflect[NB];
  sample_curve(st->ah_reflectance, reflect);
  const double bands[4] = {450.0, 550.0, 650.0, 700.0};
  float max_r = 0.0f;
  for (double b : bands) max_r = std::max(max_r, (float)eval_curve(reflect, b));
  h.wide_reflect = max_r > 0.0f;

  int e = 0;
  for (int l = 0; l < st->n_layers; ++l) {
    const Layer& L = st->layers[l];
    if (L.is_reversal) h.any_reversal = true;
    if (L.kind == L_SUPPORT || L.sensitivity.kind == CURVE_NONE) continue;  // no absorption (absorption.rs:68-70)
    double curve[NB];
    sample_curve(L.sensitivity, curve);
    const int a = T.n_absorbing++;
    const double thickness = (double)L.thickness_um;
    for (int i = 0; i < NB; ++i) {
      const double od = L.kind == L_EMULSION ? (curve[i] * 2.0) * (double)L.silver_halide_fraction * thickness
                                             : curve[i] * thickness;
      T.trans[a][i] = std::exp(-od);
    }
    T.is_emulsion[a] = L.kind == L_EMULSION;
    if (L.kind != L_EMULSION) continue;
    FilmEmulsionTables& E = T.em[e];
    const double mu = std::log(L.crystal_mean_um), sg = L.crystal_sigma_ln;
    build_lut(mu, sg, L.capture_k, E.lut_log, E.lut_frac);
    sample_curve(L.dye_epsilon, E.eps);
    E.has_mask = L.mask_epsilon.kind != CURVE_NONE;
    sample_curve(L.mask_epsilon, E.mask_eps);
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
    for (int i = 0; i < NB; ++i)
      if (curve[i] > mx) { mx = curve[i]; peak = lam(i); }
    const float wb = 0.20f, wg = 0.45f, wr = 1.0f;
    float bleed;
    if (peak <= 450.0) bleed = wb;
    else if (peak <= 550.0) bleed = wb + (wg - wb) * (float)((peak - 450.0) / 100.0);
    else if (peak <= 650.0) bleed = wg + (wr - wg) * (float)((peak - 550.0) / 100.0);
    else bleed = wr;
    const double norm = trapz(curve);
    float r_e;
    if (norm <= 1e-9) {
      r_e = (float)eval_curve(reflect, 650.0);
    } else {
      double prod[NB];
      for (int i = 0; i < NB; ++i) prod[i] = reflect[i] * curve[i];
      r_e = (float)(trapz(prod) / norm);
    }
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
  double dens[NB] = {0};
  for (int k = 0; k < e; ++k) {
    const double dm = T.em[k].has_mask ? (double)(T.em[k].d_max * 0.4f) : 0.0;
    for (int l = 0; l < NB; ++l) {
      dens[l] += 0.0 * T.em[k].eps[l];
      if (T.em[k].has_mask) dens[l] += dm * T.em[k].mask_eps[l];
    }
  }
  double tt[NB];
  for (int i = 0; i < NB; ++i) tt[i] = std::pow(10.0, -dens[i]) * T.scanner[i];
  spectrum_ap1(tt, h.dmin_raw);
  const double pk = std::max(std::max(std::max(h.dmin_raw[0], h.dmin_raw[1]), h.dmin_raw[2]), 1e-12);
  T.dmin_scale = 1.0 / pk;
  return h;
}

const HostStock& host_stock(int idx) {
  static std::unique_ptr<HostStock> cache[STOCK_COUNT];
  static std::once_flag once[STOCK_COUNT];
  const Stock* st = stock_by_index(idx);
  if (!st) throw Error{IMOPS_E_UNSUPPORTED, "film stock tables not available on the GPU path"};
  std::call_once(once[idx], [&] { cache[idx] = std::make_unique<HostStock>(build_stock(st)); });
  return *cache[idx];
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
  float* taps;       // blur taps (up to 5 kernels)
  double* partial;   // grain variance partials + norm
  FilmDeviceTables* tables;
};

// Expose -> develop -> scan (NegativeLinear) or with invert when T.invert.
void run_film(hipStream_t s, const HostStock& hs, FilmDeviceTables& T, const float* rgb_in, float* rgb_out, int w,
              int h, float pitch, float shutter, uint64_t seed, FilmBuffers& B, imops::DevBuf& tapbuf) {
  const size_t n = (size_t)w * h;
  const Stock& st = *hs.s;
  const int E = T.n_emulsions;
  // per-call constants: reciprocity (radiance.rs:43-49), bounce gains
  const float sigma_local = std::max(st.psf_local_um / std::max(pitch, 1e-6f), 0.0f);
  const float sigma_wide = std::max(st.psf_halation_um / std::max(pitch, 1e-6f), 0.0f);
  const bool do_wide = sigma_wide >= 1e-3f && hs.wide_reflect && E > 0;
  for (int e = 0; e < E; ++e) {
    const double p = (double)hs.recip_p[e];
    T.em[e].eta = ((double)shutter <= 1.0 || p >= 0.9999) ? 1.0f
                                                          : (float)std::pow(std::max((double)shutter, 1.0), p - 1.0);
    T.em[e].gain = do_wide ? hs.bleed_gain[e] : 0.0f;
  }
  // blur tap sets: local scatter, halation, DIR, adjacency, grain
  const float sig_dir = st.dir_diffusion_um / std::max(pitch, 1e-6f);
  const float sig_adj = st.developer_diffusion_um / std::max(pitch, 1e-6f);
  const float sig_grain = std::max(3.0f / std::max(pitch, 1e-6f), 1.0f);
  const float sigmas[5] = {sigma_local, sigma_wide, sig_dir, sig_adj, sig_grain};
  int radius[5], off[5];
  std::vector<float> all;
  for (int k = 0; k < 5; ++k) {
    off[k] = (int)all.size();
    if (sigmas[k] >= 1e-3f) {
      auto t = gauss_taps(sigmas[k], &radius[k]);
      all.insert(all.end(), t.begin(), t.end());
    } else {
      radius[k] = 0;
    }
  }
  tapbuf.reserve(std::max<size_t>(all.size(), 1) * sizeof(float));
  IMOPS_HIP_CHECK(hipMemcpyAsync(tapbuf.ptr, all.data(), all.size() * sizeof(float), hipMemcpyHostToDevice, s));
  IMOPS_HIP_CHECK(hipMemcpyAsync(B.tables, &T, sizeof(T), hipMemcpyHostToDevice, s));
  IMOPS_HIP_CHECK(hipStreamSynchronize(s));  // the host sources of both uploads are locals
  const float* taps = tapbuf.as<float>();

  IMOPS_HIP_CHECK(film_expose(rgb_in, n, B.tables, B.absorbed, s));
  if (sigma_local >= 1e-3f)  // local gelatin scatter, mix 0.18 (halation.rs:38-44)
    for (int e = 0; e < E; ++e) {
      float* p = B.absorbed + (size_t)e * n;
      IMOPS_HIP_CHECK(film_blur(p, B.tmp, p, p, w, h, taps + off[0], radius[0], BLUR_MIX, 0.18f, s));
    }
  if (do_wide) {
    const float* deep = B.absorbed + (size_t)(E - 1) * n;
    IMOPS_HIP_CHECK(film_blur(deep, B.tmp, deep, B.bounce, w, h, taps + off[1], radius[1], BLUR_STORE, 0.0f, s));
  }
  IMOPS_HIP_CHECK(film_capture_reduce(B.absorbed, do_wide ? B.bounce : nullptr, n, B.tables, B.image, B.mask, s));
  // DIR inhibition (diffusion.rs:39-84)
  if (E > 0 && st.n_dir > 0 && sig_dir >= 1e-3f) {
    for (int e = 0; e < E; ++e) {
      const float* p = B.image + (size_t)e * n;
      float* q = B.inh + (size_t)e * n;
      IMOPS_HIP_CHECK(film_blur(p, B.tmp, p, q, w, h, taps + off[2], radius[2], BLUR_STORE, 0.0f, s));
    }
    IMOPS_HIP_CHECK(film_dir_apply(B.image, B.inh, n, B.tables, s));
  }
  // adjacency (diffusion.rs:11-28)
  if (!(std::fabs(st.adjacency_beta) < 1e-8f || sig_adj < 1e-3f))
    for (int e = 0; e < E; ++e) {
      float* p = B.image + (size_t)e * n;
      IMOPS_HIP_CHECK(film_blur(p, B.tmp, p, p, w, h, taps + off[3], radius[3], BLUR_ADJ, st.adjacency_beta, s));
    }
  // grain (grain.rs:53-100)
  for (int e = 0; e < E; ++e) {
    const float d_max = T.em[e].d_max;
    const float kappa = (hs.kappa_ref[e] / std::max(pitch, 1e-6f)) * 0.15f;
    if (kappa <= 0.0f || d_max <= 0.0f) continue;
    const uint64_t key0 = seed * 0xD1B54A32D192ED03ull + (uint64_t)e * 0x9E3779B97F4A7C15ull;
    IMOPS_HIP_CHECK(film_grain(B.image + (size_t)e * n, B.bounce, B.tmp, w, h, key0, taps + off[4], radius[4],
                               B.partial, d_max, kappa, s));
  }
  IMOPS_HIP_CHECK(film_scan(B.image, B.mask, n, B.tables, rgb_out, s));
}

}  // namespace

FilmStageParams film_stage_params_from_table(const std::string& stock, const std::string& format, uint64_t seed,
                                             const std::string& output) {
  static const char* stocks[] = {"BwStub", "ColorNeg200", "Portra400", "Ektar100", "FujiPro400H", "EktachromeE100",
                                 "TriX400"};
  FilmStageParams p;
  p.stock = -1;
  for (int i = 0; i < 7; ++i)
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
  if (p.stock < 0 || p.stock >= STOCK_COUNT) throw Error{IMOPS_E_INVALID, "film: unknown stock"};
  const HostStock& hs = host_stock(p.stock);
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
  ctx->scr[SCR_FILM_2].reserve(sizeof(FilmDeviceTables) + (kFilmReduceBlocks + 1) * sizeof(double));
  auto layout = [&](float* base, size_t np, FilmBuffers& B) {
    B.absorbed = base;
    B.image = base + (size_t)E * np;
    B.mask = base + (size_t)2 * E * np;
    B.inh = base + (size_t)3 * E * np;
    B.tmp = base + (size_t)4 * E * np;
    B.bounce = base + (size_t)(4 * E + 1) * np;
    B.tables = ctx->scr[SCR_FILM_2].as<FilmDeviceTables>();
    B.partial = reinterpret_cast<double*>(reinterpret_cast<char*>(ctx->scr[SCR_FILM_2].ptr) + sizeof(FilmDeviceTables));
  };
  FilmDeviceTables T = hs.t;
  T.invert = 0;

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
    run_film(s, hs, T, gray, gout, 32, 32, pitch, shutter, 0, M, ctx->scr[SCR_FILM_3]);
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
      T.inv_dmin[c] = dm;
      T.inv_shoulder[c] = sh;
      T.inv_gain[c] = 0.185f / soft;
    }
    T.invert = 1;
  }
  FilmBuffers B;
  layout(ctx->scr[SCR_FILM_0].as<float>(), n, B);
  // the scan writes straight into the image: every pixel's input was consumed by expose
  run_film(s, hs, T, img->rgb.as<float>(), img->rgb.as<float>(), w, h, pitch, shutter, p.seed, B, ctx->scr[SCR_FILM_3]);
  img->color_space = IMOPS_CS_ACESCG;
}

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
