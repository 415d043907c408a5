// Film stage not yet on the GPU: it fails loudly.
#include "stages.hpp"

namespace imops {

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
