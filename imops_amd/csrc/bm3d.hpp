// BM3D kernel interfaces (bm3d.hip).
#pragma once
#include <hip/hip_runtime.h>
#include <stddef.h>
#include <stdint.h>

namespace imops {

constexpr int kBm3dMaxWindow = 41;  // search windows up to 41x41 (odd)
constexpr int kBm3dMaxGroup = 16;

struct Bm3dMatchArgs {
  const float* guide[3];
  int width, height;
  int refs_x, stride, window;
  float threshold;        // per-pixel mean SSD limit of this step
  int max_candidates, max_group;
  uint32_t* groups;       // per reference patch: [g, (x | y<<16) x 16]
  const uint32_t* ref_list = nullptr;   // list launch: the patches to match
  const uint32_t* ref_count = nullptr;  // device count of ref_list entries
  int block_width = 0;                  // candidates per lane and pass (2, 3, 4, 5, 7); 0: chosen per window
  // staging layout (bm3d_match_layout): adjacent patches per workgroup, LDS row
  // pitch and floats per staged plane
  int match_group = 1, match_pitch = 0, match_plane = 0;
};

struct Bm3dFilterArgs {
  const float* guide[3];
  const float* noisy[3];
  // plain filter: the row pass of the forward DCT of every 8-pixel row segment
  // of guide and noisy (bm3d_row_dct_launch); noisy rows are read in step 2 only
  const float* guide_rows[3];
  const float* noisy_rows[3];
  int width, height;
  int refs_x, refs_y, stride, window, tile;
  float sigma, chroma_scale, lambda;
  int haar;  // group transform: 0 Walsh-Hadamard (bm3d.rs:595-614), 1 Haar
  const uint32_t* groups;
  unsigned long long* num;  // 3 planes of i64 (two's complement in u64)
  unsigned long long* den;
  float dct[64], idct[64], kaiser[64];  // DctTables (bm3d.rs:417-461), computed on the host
};

// Matrix-core matcher (bm3d_mfma.hip): one workgroup per 8x8 block of
// reference patches; LDS offsets come from bm3d_mfma_layout().
constexpr int kBm3dMfmaListCap = 48;
struct Bm3dMfmaArgs {
  const float* guide[3];
  int width, height;
  int refs_x, refs_y, stride, window;
  float threshold;
  int max_candidates, max_group;
  uint32_t* groups;
  uint32_t* overflow_list;   // patches whose candidate list overflowed (matched by the plain kernel)
  uint32_t* overflow_count;  // [0]: patches for the plain matcher, [1]: patches selected approximately
  int blocks_x, n_blocks;
  int spx, spy;              // largest staged region (pixels)
  int pitch;                 // LDS row pitch of the FP16 copies (halves)
  int off_w16, off_w16b, off_nb, off_list, off_cnt, off_meta, off_red, smem;
  bool max_tiles_ok;         // the geometry fits the kernel's tile bookkeeping
};

struct Bm3dBlurTaps {
  int radius, len;
  float k[33];
};

size_t bm3d_filter_smem(int footprint_pixels);
hipError_t bm3d_rgb_to_ycbcr(const float* rgb, size_t n, float* y, float* cb, float* cr, hipStream_t s);
hipError_t bm3d_match_launch(const Bm3dMatchArgs& a, int nrefs, hipStream_t s, bool trial = false);
// the matcher's staging layout for this window and stride (match_group,
// match_pitch, match_plane)
void bm3d_match_layout(Bm3dMatchArgs& a);
hipError_t bm3d_match_list_launch(const Bm3dMatchArgs& a, int grid, hipStream_t s);
Bm3dMfmaArgs bm3d_mfma_layout(int width, int height, int stride, int window);
hipError_t bm3d_match_mfma_launch(const Bm3dMfmaArgs& a, hipStream_t s);
// Mixed-precision collaborative filter (bm3d_mfma.hip): DCT / Hadamard products on FP16 MFMA.
constexpr size_t kBm3dMfmaFragBytes = 2 * 4 * 2 * 64 * 16;
void bm3d_filter_mfma_fragments(const float* dct, uint16_t* out);
size_t bm3d_filter_mfma_smem(int footprint_pixels, int tile);
hipError_t bm3d_filter_mfma_launch(const Bm3dFilterArgs& a, bool step1, int nblocks, size_t smem, const void* frags,
                                   hipStream_t s, bool trial = false);
// trial: an autotuner trial launch (a separate kernel instantiation, same code)
hipError_t bm3d_filter_launch(const Bm3dFilterArgs& a, bool step1, int nblocks, size_t smem, hipStream_t s,
                              bool trial = false);
// Row pass of dct_2d_8x8 (bm3d.rs:526-530) for every 8-pixel row segment of
// three w x h planes (w >= 8): rows[c][(y * (w - 7) + x) * 8 + k] in the
// reference's order.  Needs the tables of bm3d_upload_tables.
hipError_t bm3d_row_dct_launch(const float* const in[3], int w, int h, float* const rows[3], hipStream_t s);
inline size_t bm3d_row_dct_floats(int w, int h) { return w >= 8 ? (size_t)(w - 7) * h * 8 : 0; }
// copies the DCT / IDCT / Kaiser tables the plain filter reads (64 floats each)
hipError_t bm3d_upload_tables(const float* dct, const float* idct, const float* kaiser, hipStream_t s);
hipError_t bm3d_finalize_launch(const unsigned long long* num, const unsigned long long* den, const float* const noisy[3],
                                float* const out[3], size_t n, hipStream_t s);
hipError_t fill_zero_u64(unsigned long long* p, size_t n, hipStream_t s);
hipError_t bm3d_blur_launch(const float* in, float* tmp, float* out, int w, int h, const Bm3dBlurTaps& taps,
                            hipStream_t s);
// chroma-only denoise of the Cb and Cr planes (dct_chroma_denoise, bm3d.rs:719-759); th is the hard threshold
hipError_t chroma_dct_launch(const float* cb, const float* cr, float* out_cb, float* out_cr, int w, int h, float th,
                            hipStream_t s);
// YCbCr (0..255, +128 offsets) -> RGB / 255 clamped at 0 (bm3d.rs:83-96, 629-635)
hipError_t ycbcr_merge_launch(const float* y, const float* cb, const float* cr, float* rgb, size_t n, hipStream_t s);
hipError_t bm3d_merge_launch(const float* yd, const float* cb, const float* cr, const float* yorig, const float* yblur,
                             float detail, float* rgb, size_t n, hipStream_t s);

}  // namespace imops
