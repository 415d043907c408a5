// BM3D and film stages not yet on the GPU: they fail loudly.
#include "stages.hpp"

namespace imops {

Bm3dStageParams Bm3dStageParams::from_intensity(float i) {
  Bm3dStageParams p;
  const float val = i < 0.001f ? 0.001f : (i > 1.0f ? 1.0f : i);
  p.intensity = i;
  p.sigma = val * 80.0f;
  p.lambda = 2.0f + (val * 2.5f);
  p.max_dist = 3000.0f + (val * 20000.0f);
  p.chroma_scale = 1.8f;
  const float d = i * 0.5f;
  p.detail = d < 0.0f ? 0.0f : (d > 0.5f ? 0.5f : d);
  return p;
}

Bm3dStageParams bm3d_stage_params_from_table(double intensity, double sigma_norm, double lambda, int stride,
                                             int window, int max_group) {
  Bm3dStageParams p = Bm3dStageParams::from_intensity((float)intensity);
  (void)sigma_norm; (void)lambda; (void)stride; (void)window; (void)max_group;
  return p;
}

void stage_bm3d(imops_ctx*, imops_image*, const Bm3dStageParams&) {
  throw Error{IMOPS_E_UNSUPPORTED, "BM3D is not implemented yet"};
}

FilmStageParams film_stage_params_from_table(const std::string&, const std::string&, uint64_t seed,
                                             const std::string&) {
  FilmStageParams p;
  p.seed = seed;
  return p;
}

void stage_film(imops_ctx*, imops_image*, const FilmStageParams&) {
  throw Error{IMOPS_E_UNSUPPORTED, "Film is not implemented yet"};
}

}  // namespace imops
