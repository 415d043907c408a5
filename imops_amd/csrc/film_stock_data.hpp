// Film stock parameter tables (data only).
//
// The values are the stock definitions of pichromatic/src/film/stock/*.rs
// (e.g. ColorNeg200: color_neg_200.rs:62-218), written as plain records so
// both the GPU library's host code and the oracle can build their own derived
// tables (curve samples, capture LUTs, grain kappa) from the same inputs.
// Spectral curves are given by their analytic form on the 16-sample grid
// 400..700 nm (step 20 nm).
#pragma once

namespace film_data {

enum CurveKind : int {
  CURVE_NONE = 0,
  CURVE_CONST = 1,     // a
  CURVE_GAUSS = 2,     // a * exp(-0.5 ((l - b) / c)^2): a = amplitude, b = peak, c = sigma
  CURVE_ORANGE = 3,    // 0.35 exp(-0.5 ((l-450)/50)^2) + 0.25 * a * exp(-0.5 ((l-520)/60)^2) + 0.05
  CURVE_PANCHRO = 4,   // panchromatic B&W sensitivity (bw_stub.rs:19-32 / trix_400.rs:16-28), a = variant
};

struct Curve {
  int kind;
  double a, b, c;
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

constexpr Curve C_NONE{CURVE_NONE, 0, 0, 0};
constexpr Curve cconst(double a) { return Curve{CURVE_CONST, a, 0, 0}; }
constexpr Curve gauss(double peak, double sigma, double amp) { return Curve{CURVE_GAUSS, amp, peak, sigma}; }
constexpr Curve orange(double green_amp) { return Curve{CURVE_ORANGE, green_amp, 0, 0}; }

constexpr Layer passive(int kind, float depth, float thick, Curve absorption) {
  return Layer{kind, depth, thick, absorption, 0.0, 0.0, 0.0f, C_NONE, C_NONE, 0.0f, 1.0f, 1.0, 1.0f, false};
}

// Stock order of StockId (film/stock/mod.rs:124-133).
enum StockIdx : int { BW_STUB = 0, COLOR_NEG_200 = 1, PORTRA_400 = 2, EKTAR_100 = 3, FUJI_PRO_400H = 4,
                      EKTACHROME_E100 = 5, TRIX_400 = 6, STOCK_COUNT = 7 };

// ColorNeg200: overcoat, blue emulsion, yellow filter, green emulsion, red emulsion, antihalation.
inline const Stock& color_neg_200() {
  static const Stock s{
      "ColorNeg200",
      200.0f,
      6,
      {passive(L_OVERCOAT, 0.0f, 1.0f, cconst(0.01)),
       Layer{L_EMULSION, 1.0f, 4.0f, gauss(450, 35, 1.0), 0.55, 0.30, 0.18f, gauss(450, 40, 1.0), orange(0.6), 2.2f,
             0.65f, 4.2, 0.88f, false},
       passive(L_FILTER, 5.0f, 2.0f, gauss(430, 40, 1.2)),
       Layer{L_EMULSION, 7.0f, 5.0f, gauss(550, 40, 1.0), 0.60, 0.32, 0.18f, gauss(550, 40, 1.0), orange(0.6), 2.2f,
             0.65f, 2.1, 0.90f, false},
       Layer{L_EMULSION, 12.0f, 6.0f, gauss(650, 45, 1.0), 0.70, 0.35, 0.18f, gauss(650, 45, 1.0), orange(0.6), 2.2f,
             0.65f, 1.35, 0.92f, false},
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

// Stock by StockIdx; nullptr for stocks whose tables are not provided.
inline const Stock* stock_by_index(int idx) {
  switch (idx) {
    case COLOR_NEG_200: return &color_neg_200();
    default: return nullptr;
  }
}

}  // namespace film_data
