// Film stock parameter tables (data only).
//
// The values are the stock definitions of pichromatic/src/film/stock/*.rs
// (e.g. ColorNeg200: color_neg_200.rs:67-209), written as plain records from
// which the library's host code builds its derived tables (curve samples on
// the spectral grid, capture LUTs, grain kappa).  Spectral curves are given by
// their analytic form, so they can be sampled on any grid (the reference's 16
// samples 400..700 nm, or 81 samples 380..780 nm).
#pragma once

namespace film_data {

enum CurveKind : int {
  CURVE_NONE = 0,
  CURVE_CONST = 1,     // a
  CURVE_GAUSS = 2,     // a * exp(-0.5 ((l - b) / c)^2): a = amplitude, b = peak, c = sigma
  CURVE_ORANGE = 3,    // a * exp(-0.5 ((l-450)/50)^2) + b * (c * exp(-0.5 ((l-520)/60)^2)) + d  (coloured mask)
  CURVE_PANCHRO = 4,   // a + b * max(sin(pi (l - 400) / 300), 0)  (panchromatic B&W sensitivity)
  CURVE_GAUSS2 = 5,    // a exp(-0.5 ((l - b) / c)^2) + d exp(-0.5 ((l - e) / f)^2)  (synthetic stock sensitivities)
};

struct Curve {
  int kind;
  double a, b, c, d, e, f;
};

enum LayerKindId : int { L_EMULSION = 0, L_FILTER = 1, L_OVERCOAT = 2, L_ANTIHALATION = 3, L_SUPPORT = 4 };

struct Layer {
  int kind;
  float depth_um, thickness_um;
  Curve sensitivity;              // emulsion sensitivity, or absorption curve of a filter / overcoat / AH layer
  double crystal_mean_um;         // ln(s) ~ N(ln(mean), sigma_ln^2); 0 for non-emulsion layers
  double crystal_sigma_ln;
  float silver_halide_fraction;
  Curve dye_epsilon;              // CURVE_NONE for non-emulsion layers
  Curve mask_epsilon;             // CURVE_NONE when the coupler has no coloured mask
  float d_max;
  float gamma_contrast;
  double capture_k;
  float reciprocity_p;
  bool is_reversal;
};

struct Stock {
  const char* name;
  float box_iso;
  int n_layers;
  Layer layers[12];
  Curve ah_reflectance;
  float psf_local_um, psf_halation_um;
  float developer_diffusion_um;
  float adjacency_beta;
  float dir_diffusion_um;
  int n_dir;                      // DIR matrix is n_dir x n_dir, matrix[source][target]
  float dir[7][7];
  Curve scanner_light;
};

constexpr Curve C_NONE{CURVE_NONE, 0, 0, 0, 0};
constexpr Curve cconst(double a) { return Curve{CURVE_CONST, a, 0, 0, 0}; }
constexpr Curve gauss(double peak, double sigma, double amp) { return Curve{CURVE_GAUSS, amp, peak, sigma, 0}; }
// orange_mask_epsilon of each colour-negative stock: blue_w * blue + green_w * (green_amp * green) + offset
constexpr Curve orange(double blue_w, double green_w, double green_amp, double offset) {
  return Curve{CURVE_ORANGE, blue_w, green_w, green_amp, offset};
}
// panchromatic_sensitivity of the B&W stocks: base + amp * max(sin(pi t), 0), t = (l - 400) / 300
constexpr Curve panchro(double base, double amp) { return Curve{CURVE_PANCHRO, base, amp, 0, 0}; }

constexpr Layer emulsion(float depth, float thick, Curve sens, double crystal_mean, double sigma_ln, float agx,
                         Curve dye, Curve mask, float d_max, float gamma, double capture_k, float recip_p,
                         bool reversal = false) {
  return Layer{L_EMULSION, depth, thick, sens, crystal_mean, sigma_ln, agx, dye, mask, d_max, gamma, capture_k,
               recip_p, reversal};
}

constexpr Layer passive(int kind, float depth, float thick, Curve absorption) {
  return Layer{kind, depth, thick, absorption, 0.0, 0.0, 0.0f, C_NONE, C_NONE, 0.0f, 1.0f, 1.0, 1.0f, false};
}

// Stock order of StockId (film/stock/mod.rs:124-133).
enum StockIdx : int { BW_STUB = 0, COLOR_NEG_200 = 1, PORTRA_400 = 2, EKTAR_100 = 3, FUJI_PRO_400H = 4,
                      EKTACHROME_E100 = 5, TRIX_400 = 6, STOCK_COUNT = 7,
                      SYNTHETIC = 7 };  // seeded extension stock (not one of the reference's StockIds)

// ColorNeg200: overcoat, blue emulsion, yellow filter, green emulsion, red emulsion, antihalation.
inline const Stock& color_neg_200() {
  constexpr Curve mask = orange(0.35, 0.25, 0.6, 0.05);
  static const Stock s{
      "ColorNeg200",
      200.0f,
      6,
      {passive(L_OVERCOAT, 0.0f, 1.0f, cconst(0.01)),
       emulsion(1.0f, 4.0f, gauss(450, 35, 1.0), 0.55, 0.30, 0.18f, gauss(450, 40, 1.0), mask, 2.2f, 0.65f, 4.2, 0.88f),
       passive(L_FILTER, 5.0f, 2.0f, gauss(430, 40, 1.2)),
       emulsion(7.0f, 5.0f, gauss(550, 40, 1.0), 0.60, 0.32, 0.18f, gauss(550, 40, 1.0), mask, 2.2f, 0.65f, 2.1, 0.90f),
       emulsion(12.0f, 6.0f, gauss(650, 45, 1.0), 0.70, 0.35, 0.18f, gauss(650, 45, 1.0), mask, 2.2f, 0.65f, 1.35, 0.92f),
       passive(L_ANTIHALATION, 18.0f, 2.0f, gauss(650, 80, 0.3))},
      gauss(680, 60, 0.08),
      3.0f,
      80.0f,
      8.0f,
      0.35f,
      15.0f,
      3,
      {{0.02f, 0.05f, 0.03f}, {0.05f, 0.02f, 0.06f}, {0.03f, 0.06f, 0.02f}},
      cconst(1.0)};
  return s;
}

// Portra400 (portra_400.rs:42-245): fast/slow sub-layers per colour, DIR between all six.
inline const Stock& portra_400() {
  constexpr Curve mask = orange(0.32, 0.22, 0.6, 0.04);
  static const Stock s{
      "Portra400",
      400.0f,
      9,
      {passive(L_OVERCOAT, 0.0f, 1.0f, cconst(0.01)),
       emulsion(1.0f, 2.5f, gauss(450, 35, 1.0), 0.85, 0.35, 0.16f, gauss(450, 40, 1.0), mask, 1.1f, 0.54f, 4.8, 0.85f),
       emulsion(3.5f, 2.5f, gauss(450, 35, 0.8), 0.42, 0.28, 0.20f, gauss(450, 40, 1.0), mask, 1.1f, 0.54f, 2.2, 0.89f),
       passive(L_FILTER, 6.0f, 2.0f, gauss(430, 40, 1.3)),
       emulsion(8.0f, 3.0f, gauss(550, 40, 1.0), 0.90, 0.35, 0.16f, gauss(550, 40, 1.0), mask, 1.1f, 0.54f, 2.4, 0.87f),
       emulsion(11.0f, 3.0f, gauss(550, 40, 0.8), 0.45, 0.28, 0.20f, gauss(550, 40, 1.0), mask, 1.1f, 0.54f, 1.1, 0.91f),
       emulsion(14.0f, 3.5f, gauss(650, 45, 1.0), 0.95, 0.35, 0.16f, gauss(650, 45, 1.0), mask, 1.1f, 0.54f, 1.5, 0.89f),
       emulsion(17.5f, 3.5f, gauss(650, 45, 0.8), 0.48, 0.28, 0.20f, gauss(650, 45, 1.0), mask, 1.1f, 0.54f, 0.7, 0.93f),
       passive(L_ANTIHALATION, 21.0f, 2.0f, gauss(650, 80, 0.3))},
      gauss(680, 60, 0.06),
      3.0f,
      70.0f,
      7.0f,
      0.30f,
      16.0f,
      6,
      {{0.02f, 0.01f, 0.04f, 0.02f, 0.03f, 0.01f},
       {0.01f, 0.01f, 0.02f, 0.01f, 0.02f, 0.01f},
       {0.04f, 0.02f, 0.02f, 0.01f, 0.05f, 0.02f},
       {0.02f, 0.01f, 0.01f, 0.01f, 0.02f, 0.01f},
       {0.03f, 0.01f, 0.05f, 0.02f, 0.02f, 0.01f},
       {0.01f, 0.01f, 0.02f, 0.01f, 0.01f, 0.01f}},
      cconst(1.0)};
  return s;
}

// Ektar100 (ektar_100.rs:46-268): hard-gamma slow layers, stronger DIR and adjacency.
inline const Stock& ektar_100() {
  constexpr Curve mask = orange(0.38, 0.26, 0.6, 0.05);
  static const Stock s{
      "Ektar100",
      100.0f,
      9,
      {passive(L_OVERCOAT, 0.0f, 1.0f, cconst(0.01)),
       emulsion(1.0f, 2.8f, gauss(450, 32, 1.0), 0.82, 0.32, 0.17f, gauss(450, 30, 1.25), mask, 1.25f, 0.58f, 4.8, 0.86f),
       emulsion(3.8f, 2.5f, gauss(450, 32, 0.85), 0.38, 0.24, 0.21f, gauss(450, 30, 1.25), mask, 1.90f, 0.34f, 6.5, 0.90f),
       passive(L_FILTER, 6.0f, 2.0f, gauss(430, 38, 1.35)),
       emulsion(8.0f, 3.2f, gauss(545, 36, 1.0), 0.88, 0.32, 0.17f, gauss(545, 30, 1.25), mask, 1.25f, 0.58f, 2.6, 0.88f),
       emulsion(11.2f, 3.0f, gauss(545, 36, 0.85), 0.40, 0.24, 0.21f, gauss(545, 30, 1.25), mask, 1.90f, 0.34f, 3.6, 0.92f),
       emulsion(14.2f, 3.6f, gauss(650, 40, 1.0), 0.92, 0.32, 0.17f, gauss(650, 32, 1.25), mask, 1.25f, 0.58f, 1.6, 0.90f),
       emulsion(17.8f, 3.5f, gauss(650, 40, 0.85), 0.42, 0.24, 0.21f, gauss(650, 32, 1.25), mask, 1.90f, 0.34f, 2.4, 0.94f),
       passive(L_ANTIHALATION, 21.5f, 2.0f, gauss(650, 80, 0.25))},
      gauss(680, 60, 0.05),
      2.5f,
      55.0f,
      6.0f,
      0.55f,
      14.0f,
      6,
      {{0.03f, 0.02f, 0.09f, 0.04f, 0.07f, 0.03f},
       {0.02f, 0.02f, 0.04f, 0.02f, 0.04f, 0.02f},
       {0.09f, 0.04f, 0.03f, 0.02f, 0.10f, 0.04f},
       {0.04f, 0.02f, 0.02f, 0.02f, 0.04f, 0.02f},
       {0.07f, 0.03f, 0.10f, 0.04f, 0.03f, 0.02f},
       {0.03f, 0.02f, 0.04f, 0.02f, 0.02f, 0.02f}},
      cconst(1.0)};
  return s;
}

// FujiPro400H (fuji_pro_400h.rs:43-271): a fourth (cyan-sensitive 490 nm) layer between the
// filter and the green pair; non-symmetric DIR over seven emulsions.
inline const Stock& fuji_pro_400h() {
  constexpr Curve mask = orange(0.28, 0.20, 0.55, 0.03);
  static const Stock s{
      "FujiPro400H",
      400.0f,
      10,
      {passive(L_OVERCOAT, 0.0f, 1.0f, cconst(0.01)),
       emulsion(1.0f, 2.5f, gauss(445, 35, 1.0), 0.82, 0.33, 0.17f, gauss(445, 40, 1.0), mask, 1.1f, 0.56f, 4.5, 0.86f),
       emulsion(3.5f, 2.5f, gauss(445, 35, 0.8), 0.40, 0.28, 0.19f, gauss(445, 40, 1.0), mask, 1.1f, 0.56f, 2.1, 0.90f),
       passive(L_FILTER, 6.0f, 2.0f, gauss(430, 40, 1.3)),
       emulsion(8.0f, 2.0f, gauss(490, 30, 0.9), 0.50, 0.30, 0.15f, gauss(650, 45, 0.6), mask, 0.5f, 0.56f, 2.0, 0.88f),
       emulsion(10.0f, 3.0f, gauss(545, 40, 1.0), 0.88, 0.34, 0.17f, gauss(545, 40, 1.0), mask, 1.1f, 0.56f, 2.3, 0.87f),
       emulsion(13.0f, 3.0f, gauss(545, 40, 0.8), 0.44, 0.28, 0.19f, gauss(545, 40, 1.0), mask, 1.1f, 0.56f, 1.1, 0.91f),
       emulsion(16.0f, 3.5f, gauss(650, 45, 1.0), 0.92, 0.34, 0.17f, gauss(650, 45, 1.0), mask, 1.1f, 0.56f, 1.4, 0.89f),
       emulsion(19.5f, 3.5f, gauss(650, 45, 0.8), 0.46, 0.28, 0.19f, gauss(650, 45, 1.0), mask, 1.1f, 0.56f, 0.7, 0.93f),
       passive(L_ANTIHALATION, 23.0f, 2.0f, gauss(650, 80, 0.3))},
      gauss(670, 50, 0.05),
      2.5f,
      65.0f,
      6.5f,
      0.32f,
      14.0f,
      7,
      {{0.02f, 0.01f, 0.03f, 0.04f, 0.02f, 0.03f, 0.01f},
       {0.01f, 0.01f, 0.02f, 0.02f, 0.01f, 0.02f, 0.01f},
       {0.03f, 0.02f, 0.02f, 0.06f, 0.03f, 0.04f, 0.02f},
       {0.04f, 0.02f, 0.05f, 0.02f, 0.01f, 0.05f, 0.02f},
       {0.02f, 0.01f, 0.02f, 0.01f, 0.01f, 0.02f, 0.01f},
       {0.03f, 0.01f, 0.04f, 0.05f, 0.02f, 0.02f, 0.01f},
       {0.01f, 0.01f, 0.02f, 0.02f, 0.01f, 0.01f, 0.01f}},
      cconst(1.0)};
  return s;
}

// EktachromeE100 (ektachrome_e100.rs:28-156): reversal (E-6) slide film, no coloured mask.
// Every layer carries is_reversal, as in the reference.
inline const Stock& ektachrome_e100() {
  static const Stock s{
      "EktachromeE100",
      100.0f,
      6,
      {Layer{L_OVERCOAT, 0.0f, 1.0f, cconst(0.01), 0.0, 0.0, 0.0f, C_NONE, C_NONE, 0.0f, 1.0f, 1.0, 1.0f, true},
       emulsion(1.0f, 4.0f, gauss(440, 30, 1.0), 0.42, 0.25, 0.22f, gauss(440, 35, 1.0), C_NONE, 3.2f, 1.65f, 3.5, 0.93f,
                true),
       Layer{L_FILTER, 5.0f, 2.0f, gauss(430, 35, 1.4), 0.0, 0.0, 0.0f, C_NONE, C_NONE, 0.0f, 1.0f, 1.0, 1.0f, true},
       emulsion(7.0f, 4.0f, gauss(545, 35, 1.0), 0.45, 0.25, 0.22f, gauss(545, 35, 1.0), C_NONE, 3.2f, 1.65f, 1.8, 0.94f,
                true),
       emulsion(11.0f, 5.0f, gauss(650, 40, 1.0), 0.50, 0.26, 0.22f, gauss(650, 40, 1.0), C_NONE, 3.2f, 1.65f, 1.1, 0.95f,
                true),
       Layer{L_ANTIHALATION, 16.0f, 2.0f, gauss(650, 80, 0.3), 0.0, 0.0, 0.0f, C_NONE, C_NONE, 0.0f, 1.0f, 1.0, 1.0f,
             true}},
      gauss(660, 40, 0.02),
      1.5f,
      40.0f,
      4.0f,
      0.40f,
      8.0f,
      3,
      {{0.01f, 0.02f, 0.01f}, {0.02f, 0.01f, 0.02f}, {0.01f, 0.02f, 0.01f}},
      cconst(1.0)};
  return s;
}

// BwStub (bw_stub.rs:34-82): one panchromatic emulsion, neutral silver density.
inline const Stock& bw_stub() {
  static const Stock s{
      "BwStub",
      100.0f,
      1,
      {emulsion(0.0f, 10.0f, panchro(0.3, 0.7), 0.7, 0.35, 0.2f, cconst(1.0), C_NONE, 2.5f, 0.65f, 1.0, 1.0f)},
      cconst(0.0),
      2.0f,
      40.0f,
      5.0f,
      0.3f,
      10.0f,
      1,
      {{0.0f}},
      cconst(1.0)};
  return s;
}

// TriX400 (trix_400.rs:30-91): fast and slow panchromatic silver layers.
inline const Stock& trix_400() {
  static const Stock s{
      "TriX400",
      400.0f,
      2,
      {emulsion(0.0f, 6.0f, panchro(0.35, 0.65), 0.95, 0.38, 0.22f, cconst(1.0), C_NONE, 1.5f, 0.70f, 2.2, 0.85f),
       emulsion(6.0f, 6.0f, panchro(0.35, 0.65), 0.45, 0.28, 0.24f, cconst(1.0), C_NONE, 1.5f, 0.70f, 0.9, 0.89f)},
      cconst(0.01),
      3.5f,
      50.0f,
      6.0f,
      0.35f,
      12.0f,
      2,
      {{0.03f, 0.02f}, {0.02f, 0.02f}},
      cconst(1.0)};
  return s;
}

// Stock by StockIdx (StockId::load, stock/mod.rs:136-148).
inline const Stock* stock_by_index(int idx) {
  switch (idx) {
    case BW_STUB: return &bw_stub();
    case COLOR_NEG_200: return &color_neg_200();
    case PORTRA_400: return &portra_400();
    case EKTAR_100: return &ektar_100();
    case FUJI_PRO_400H: return &fuji_pro_400h();
    case EKTACHROME_E100: return &ektachrome_e100();
    case TRIX_400: return &trix_400();
    default: return nullptr;
  }
}

}  // namespace film_data
