// Parameters and entry points of the non-elementwise stages (BM3D, film).
#pragma once
#include <cstdint>
#include <string>

#include "runtime.hpp"

namespace imops {

// BM3D parameters.  from_intensity reproduces Bm3dParams::from_intensity
// (pichromatic/src/bm3d.rs:22-32) and the fixed geometry of bm3d.rs:7-12;
// the explicit fields let a stage list set the geometry and noise model
// directly (defaults leave the reference behaviour unchanged).
struct Bm3dStageParams {
  float intensity = 0.0f;
  float sigma = 0.08f;          // noise sigma in 0..255 units
  float lambda = 2.0025f;       // hard-threshold multiplier
  float max_dist = 3020.0f;     // step-1 match threshold (per-pixel mean SSD)
  float chroma_scale = 1.8f;    // sigma multiplier for Cb/Cr
  float detail = 0.0f;          // luma detail blend strength
  int stride = 6;               // reference-patch grid step
  int window = 19;              // search window (odd)
  int max_group = 16;           // group size cap (power of two)
  int max_candidates = 1024;    // accepted-candidate cap per reference patch
  int precision = IMOPS_BM3D_FP32;  // IMOPS_BM3D_FP32, _MIXED or _MIXED_MATCH
  int group_transform = IMOPS_BM3D_HADAMARD;  // IMOPS_BM3D_HADAMARD (bm3d.rs:595-614) or IMOPS_BM3D_HAAR
  static Bm3dStageParams from_intensity(float i);
};

Bm3dStageParams bm3d_stage_params_from_table(double intensity, double sigma_norm, double lambda, int stride,
                                             int window, int max_group);
void stage_bm3d(imops_ctx* ctx, imops_image* img, const Bm3dStageParams& p);
// chroma_bm3d (pichromatic/src/bm3d.rs:705-759)
void stage_chroma_bm3d(imops_ctx* ctx, imops_image* img, float intensity);
// chroma_denoise (pichromatic/src/chroma_denoise.rs:14-152)
void stage_luma_chroma_denoise(imops_ctx* ctx, imops_image* img, int radius, float epsilon);

struct FilmStageParams {
  int stock = 1;    // StockId order of film/stock/mod.rs (0 BwStub, 1 ColorNeg200, ...; 7 the synthetic stock)
  int format = 0;   // 0 Film35mm, 1 Film6x6, 2 Film4x5
  uint64_t seed = 1;
  int output = 0;   // 0 NegativeLinear, 1 PositiveLinear
  bool fp32 = false;  // precision = "mixed": FP32 per-pixel spectral arithmetic (the reference's is FP64)
  // Extensions (the defaults are the reference's behaviour):
  int bins = 16;              // spectral grid: 16 (400..700 nm / 20) or 81 (380..780 nm / 5)
  uint64_t stock_seed = 0;    // parameters of the synthetic stock
  bool philox = false;        // grain noise from Philox4x32-10 normals instead of SplitMix64 Irwin-Hall rows
  float grain_sigma = -1.0f;  // >= 0: constant grain density sigma instead of kappa * taper * sqrt(D (Dmax - D))
  float hal_radius = -1.0f;   // >= 0: halation blur radius in pixels (sigma = radius / 3) instead of psf_halation / pitch
  float hal_strength = -1.0f; // >= 0: halation gain strength * bleed_e instead of r_e * bleed_e
};

FilmStageParams film_stage_params_from_table(const std::string& stock, const std::string& format, uint64_t seed,
                                             const std::string& output);
void stage_film(imops_ctx* ctx, imops_image* img, const FilmStageParams& p);

}  // namespace imops
