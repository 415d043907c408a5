// AMaZE demosaic kernel interface (device-side entry in amaze.hip).
#pragma once
#include <hip/hip_runtime.h>
#include <stddef.h>
#include <stdint.h>

namespace imops {

// Floats of global scratch one workgroup uses for one 128x128 tile:
// 11 full-resolution and 11 half-resolution f32 planes.
constexpr size_t kAmazeSlotFloats = 11 * 16384 + 11 * 8192;

struct AmazeArgs {
  const float* in;   // normalised CFA plane, width*height
  float* out;        // interleaved RGB f32, width*height*3
  int width, height;
  int tiles_x, tiles_y;
  float* scratch;    // gridDim.x * kAmazeSlotFloats floats
  uint8_t cfa[2][2]; // colour of the (crop-shifted) Bayer pattern at (r&1, c&1)
};

size_t amaze_tiles(int width, int height, int* tiles_x, int* tiles_y);
hipError_t amaze_launch(const AmazeArgs& args, int grid, hipStream_t stream);

}  // namespace imops
