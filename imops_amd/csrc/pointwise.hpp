// Per-pixel kernel interfaces (pointwise.hip).
#pragma once
#include <hip/hip_runtime.h>
#include <stddef.h>
#include <stdint.h>

namespace imops {

// Repeating CFA pattern (period up to 12x12); colour at (r, c) for r, c >= 0.
struct CfaPattern {
  int w, h;
  uint8_t base[12][12];
  __host__ __device__ int at(int r, int c) const { return base[r % h][c % w]; }
};

enum PointwiseKind : int {
  PW_MAT3 = 1, PW_SCALE3 = 2, PW_OETF = 3, PW_EOTF = 4, PW_GAMMA = 5,
  PW_CONTRAST = 6,      // m[0] = contrast exponent
  PW_HIGHLIGHT = 7,     // m[0] = green clip (white-balance coefficient 1)
  PW_LIN_TO_OKLCH = 8,  // linear sRGB -> Oklch (L, C, hue in degrees)
  PW_OKLCH_TO_LIN = 9,  // Oklch -> linear sRGB
  PW_SIGMOID = 10,      // SigmoidToneMap on ACEScg; m[0..8] ACEScg -> linear sRGB, m[9..17] back
};

struct PointwiseOp {
  int kind;
  float m[18];
};

// A short program of elementwise ops applied in order to every pixel.
struct PointwiseProgram {
  static constexpr int kMax = 12;
  int n = 0;
  PointwiseOp ops[kMax];
  bool push(int kind, const float* m, int nm) {
    if (n >= kMax) return false;
    ops[n].kind = kind;
    for (int i = 0; i < 18; ++i) ops[n].m[i] = i < nm ? m[i] : 0.0f;
    ++n;
    return true;
  }
};

hipError_t normalize_launch(const uint16_t* raw, size_t full_w, size_t cx, size_t cy, size_t cw, size_t ch,
                            float black, float white, float* out, hipStream_t s);
// crop_bayer_center (extern_pipeline.rs:113-179): out is nw x nh.
hipError_t crop_bayer_center_launch(const float* in, size_t width, size_t height, size_t factor, size_t nw, size_t nh,
                                   float* out, hipStream_t s);
// Fast / SuperFast / SuperSuperFast (demosaic.rs:189-354): output (w/f) x (h/f); cell[k] is the
// colour of CFA site (k/2, k%2), written in that order (a repeated colour keeps the later site).
hipError_t preview_demosaic_launch(const float* in, int width, int height, int factor, const int cell[4], float* out,
                                   hipStream_t s);
// Three-pass Markesteijn (markesteijn3.hip): passes 1 or 3; scratch of
// markesteijn3_scratch_bytes.
size_t markesteijn3_scratch_bytes(int width, int height, int passes);
hipError_t markesteijn3_launch(const float* in, int width, int height, const CfaPattern& cfa, int passes,
                               void* scratch, float* out, hipStream_t s);
hipError_t markesteijn_launch(const float* in, int width, int height, const CfaPattern& cfa, float* out,
                              hipStream_t s);
hipError_t pointwise_launch(float* rgb, size_t npix, const PointwiseProgram& prog, hipStream_t s);

// FixVignetteRadial (vignette.rs:56-96): k[0..4] polynomial, (cx, cy) centre, d_max corner distance.
struct VignetteParams {
  float k[5], cx, cy, d_max, strength;
};
hipError_t vignette_launch(float* rgb, int width, int height, const VignetteParams& v, hipStream_t s);

// Output encodings of interleaved RGB, with the EXIF/DNG orientation applied
// (0 none, 1 Rotate90, 2 Rotate180, 3 Rotate270; apply_orientation, output.rs:31-70):
// u8 = (clamp(c, 0, 1) * 255) truncated (get_image_rgb_data_c, extern_pipeline.rs:362-395),
// f16 = IEEE binary16 round-to-nearest-even (save_exr's f16::from_f32, output.rs:201-209).
hipError_t encode_u8_launch(const float* rgb, int width, int height, int orientation, uint8_t* out, hipStream_t s);
hipError_t encode_f16_launch(const float* rgb, int width, int height, int orientation, uint16_t* out, hipStream_t s);

}  // namespace imops
