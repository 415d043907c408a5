// Oracle (test infrastructure only): BM3D declarations.
#pragma once
#include <cstddef>

namespace oracle {

struct Bm3dParams {
  float intensity = 0.0f;
  float sigma = 0.08f;
  float lambda = 2.0025f;
  float max_dist = 3020.0f;
  float chroma_scale = 1.8f;
  float detail = 0.0f;
  int stride = 6;
  int window = 19;
  int max_group = 16;
  int max_candidates = 1024;
  int group_transform = 0;  // 0: Walsh-Hadamard (bm3d.rs:595-614), 1: Haar (extension)
  // Bm3dParams::from_intensity (bm3d.rs:22-32) + detail strength (bm3d.rs:53)
  static Bm3dParams from_intensity(float i);
};

struct Bm3dStats {
  long long refs = 0, groups = 0;
};

template <class T>
void bm3d(float* rgb, size_t w, size_t h, const Bm3dParams& p, Bm3dStats* stats);
void chroma_bm3d(float* rgb, size_t w, size_t h, float intensity);
void bm3d_tables(float* dct, float* idct, float* kaiser);
void group_transform_1d(double* d, int n, int kind, bool inverse, bool fp64);

}  // namespace oracle
