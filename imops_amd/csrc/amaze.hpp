// AMaZE demosaic kernel interface (device-side entry in amaze.hip).
#pragma once
#include <hip/hip_runtime.h>
#include <stddef.h>
#include <stdint.h>

namespace imops {

// Floats of global scratch one workgroup uses for one 128x128 tile:
// 11 full-resolution and 11 half-resolution f32 planes.
constexpr size_t kAmazeSlotFloats = 8 * 16384 + 11 * 8192;

struct AmazeArgs {
  const float* in;   // normalised CFA plane, width*height (or null: read in16)
  // u16 frame read directly by the tile fill: sample (r, c) of the crop window is
  // (in16[(in16_y + r) * in16_pitch + in16_x + c] - black) / range
  const uint16_t* in16;
  size_t in16_pitch, in16_x, in16_y;
  float black, range;
  float* out;        // interleaved RGB f32, width*height*3
  int width, height;
  int tiles_x, tiles_y;
  float* scratch;    // gridDim.x * kAmazeSlotFloats floats
  uint8_t cfa[2][2]; // colour of the (crop-shifted) Bayer pattern at (r&1, c&1)
};

// Tile slots of the allocation-free entry point, whose scratch is a static
// device array of kAmazeNoallocSlots * kAmazeSlotFloats floats (69 MB) per GPU.
constexpr int kAmazeNoallocSlots = 64;

size_t amaze_tiles(int width, int height, int* tiles_x, int* tiles_y);
// Device address of the static scratch array on `device`, which must be the
// current device.  Looked up once per device (later calls return the cached
// address); safe to call from several threads.
hipError_t amaze_noalloc_scratch(int device, float** out);
hipError_t amaze_launch(const AmazeArgs& args, int grid, hipStream_t stream);

}  // namespace imops
