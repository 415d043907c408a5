/*
 * imops_gpu.h - C ABI of the MI355X (gfx950) image-stage library.
 *
 * The library replaces the CPU bodies of the pichromatic stages on the
 * raw -> demosaic -> BM3D -> colour -> film path.  Its calls mirror the
 * reference's stage interface:
 *   - per-stage calls  <-> Module<T>::process (pichromatic-pipeline/src/modules.rs:175-181,
 *                          stage impls 492-894) and the Image methods they call
 *                          (pichromatic/src/pixel.rs:55-171);
 *   - pipeline calls   <-> get_pixel_pipeline_c / run_pixel_pipeline_c / free_pipeline_c
 *                          (pichromatic-pipeline/src/extern_pipeline.rs:326-359, 434-443)
 *                          with parse_config (config.rs:13-34) semantics;
 *   - imops_last_error <-> get_last_error_c (extern_pipeline.rs:27-38).
 *
 * Conventions: every call returns IMOPS_OK (0) or a negative IMOPS_E_* code
 * and never unwinds across the boundary; the message of the last failure on
 * the calling thread is returned by imops_last_error().  Host buffers belong
 * to the caller.  Device buffers belong to the library and sit behind opaque
 * handles with explicit free calls.  A context owns one HIP stream; calls on
 * one context are serialised on that stream and are not thread-safe against
 * each other, separate contexts are independent.
 */
#ifndef IMOPS_GPU_H
#define IMOPS_GPU_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define IMOPS_OK 0
#define IMOPS_E_INVALID (-1)     /* bad argument (null pointer, size, unknown name) */
#define IMOPS_E_HIP (-2)         /* HIP runtime failure */
#define IMOPS_E_STATE (-3)       /* image not in the state the stage needs */
#define IMOPS_E_UNSUPPORTED (-4) /* input the reference itself cannot process */

/* Colour spaces (the reference's ColorSpaceTag subset: modules.rs:107-112). */
#define IMOPS_CS_NONE (-1) /* camera RGB straight out of demosaic (color_space = None) */
#define IMOPS_CS_SRGB 0
#define IMOPS_CS_LINEAR_SRGB 1
#define IMOPS_CS_ACESCG 2
#define IMOPS_CS_XYZD65 3

/* BM3D group transform across the patches of a group
 * (imops_bm3d_params.group_transform): the reference's Walsh-Hadamard
 * (walsh_hadamard_1d, bm3d.rs:595-614) or the orthonormal multi-level Haar
 * transform (an extension; identical for groups of two). */
#define IMOPS_BM3D_HADAMARD 0
#define IMOPS_BM3D_HAAR 1

/* BM3D arithmetic (imops_bm3d_params.precision). */
#define IMOPS_BM3D_FP32 0
#define IMOPS_BM3D_MIXED 1
#define IMOPS_BM3D_MIXED_MATCH 2

/* Demosaic algorithms (DemosaicAlgorithmType, modules.rs:73-100). */
#define IMOPS_DEMOSAIC_AMAZE 0
#define IMOPS_DEMOSAIC_MARKESTEIJN 1
/* Preview modes (demosaic.rs:189-354): the top-left 2x2 CFA cell of every 2x2,
 * 4x4 or 8x8 block gives one RGB pixel, so the frame shrinks by 2, 4 or 8 per
 * side; the raw plane no longer matches the new size and is dropped. */
#define IMOPS_DEMOSAIC_FAST 2
#define IMOPS_DEMOSAIC_SUPERFAST 3
#define IMOPS_DEMOSAIC_SUPERSUPERFAST 4
/* The three-pass Markesteijn X-Trans demosaic (an extension: the reference's
 * "markesteijn" is the 3x3 same-colour mean): directional green clamped to the
 * local green range, green recomputed from the closer pixels, colour-difference
 * red / blue, Y Pb Pr homogeneity selection; any CFA.  MARKESTEIJN3 runs the
 * three passes (eight candidates), MARKESTEIJN3_1PASS one pass (four). */
#define IMOPS_DEMOSAIC_MARKESTEIJN3 5
#define IMOPS_DEMOSAIC_MARKESTEIJN3_1PASS 6

typedef struct imops_ctx imops_ctx;
typedef struct imops_image imops_image;
typedef struct imops_pipeline imops_pipeline;

/* Message of the last failed call on this thread ("" if none).  Valid until the
 * next failing call on the same thread.  Replaces get_last_error_c. */
const char* imops_last_error(void);

/* Context: device ordinal, one stream, cached tables and scratch. */
int imops_ctx_create(int device, imops_ctx** out);
int imops_ctx_destroy(imops_ctx* ctx);
int imops_ctx_synchronize(imops_ctx* ctx);
/* The hipStream_t every call of this context is enqueued on. */
void* imops_ctx_stream(imops_ctx* ctx);
/* GPU timing of named regions (HIP events on the context stream around the
 * launches: "amaze", "markesteijn", "normalize", "pointwise", "bm3d",
 * "bm3d_match", "bm3d_filter_ht", "bm3d_filter_wiener", "film").  Enabling or
 * disabling clears the recorded events; imops_ctx_timing synchronises the
 * stream and returns the summed milliseconds and the number of recorded
 * launches of one region.  The name "bm3d_fallback_refs" instead sums the
 * reference patches the mixed-precision matcher handed to the FP32 matcher
 * (recording it synchronises the stream once per BM3D step while timing is on).
 * No counterpart in the reference (its CLI prints wall-clock pipeline time
 * only, src/bin/imops-cli.rs:59-63). */
int imops_ctx_set_timing(imops_ctx* ctx, int enabled);
/* Launch-shape autotuning (no counterpart in the reference).  When enabled, the
 * first call of a tunable kernel at a new shape (BM3D filter: references per
 * workgroup tile; FP32 BM3D matcher: candidates per thread) times every
 * candidate on the call's own data and keeps the fastest for the context.
 * Every candidate computes identical results.  imops_ctx_tuning writes the
 * choices as "key=value" lines (NUL-terminated, truncated to len) and the size
 * the whole report needs to *needed. */
int imops_ctx_set_autotune(imops_ctx* ctx, int enabled);
int imops_ctx_tuning(imops_ctx* ctx, char* buf, size_t len, size_t* needed);
int imops_ctx_timing(imops_ctx* ctx, const char* region, double* total_ms, int* count);

/* Measured rates of the context's device (no counterpart in the reference: the
 * benchmark reports them beside the published peaks).  Allocates two 1 GiB
 * buffers for the call and writes n >= 6 values: out[0] HBM copy GB/s (read +
 * written bytes), out[1] FP32 non-fused VALU wave64 instructions per clock per
 * SIMD at the rated clock, out[2] the same stream in TFLOP/s (one flop per lane
 * per mul or add), out[3] dense v_mfma_f32_16x16x32_f16 TFLOP/s, out[4] the rated
 * clock in GHz, out[5] the number of compute units. */
int imops_measure_peaks(imops_ctx* ctx, double* out, int n);

/* Test hook (no counterpart in the reference): the next `count` device
 * allocations the library makes fail as hipMalloc would (the call returns
 * IMOPS_E_HIP and imops_last_error names the allocation); 0 clears it.  Lets
 * tests reach the device-allocation error path from the host. */
int imops_debug_fail_allocations(int count);

/* Images: device-resident replacement of pichromatic::pixel::Image
 * (pixel.rs:22-27, image.rs:7-43). */
int imops_image_create(imops_ctx* ctx, imops_image** out);
int imops_image_free(imops_image* img);

/* Raw ingest fused into the demosaic load: binds the caller's DEVICE u16 frame
 * (full_w x full_h, the crop window cx, cy, cw, ch, black / white levels as
 * imops_image_set_raw_u16) without copying or normalising it.  AMaZE then
 * reads and normalises the u16 samples in its tile fill (crop_and_normalize,
 * demosaic.rs:61-91, and the CFA shift of the crop origin, demosaic.rs:93-98,
 * folded into the load); any other consumer of the raw plane first converts
 * it as imops_image_set_raw_u16 would.  The frame must stay valid and
 * unchanged until the demosaic has run (synchronise the context). */
int imops_image_bind_raw_u16(imops_ctx* ctx, imops_image* img, const uint16_t* dev_src, size_t full_w, size_t full_h,
                             size_t cx, size_t cy, size_t cw, size_t ch, float black, float white,
                             const char* cfa_pattern);

/* crop_bayer_center (extern_pipeline.rs:113-179): keeps every crop_factor-th
 * 2x2 CFA cell of the raw plane, so the frame becomes
 * ((w / f) & ~1) x ((h / f) & ~1) with the same CFA phase; metadata kept.
 * crop_factor <= 1 leaves the image unchanged. */
int imops_crop_bayer_center(imops_ctx* ctx, imops_image* img, size_t crop_factor);

/* The calibration matrix parse_raw_image chooses (extern_pipeline.rs:254-270,
 * interpolate_matrices 457-495): with both the illuminant-A and D65 colour
 * matrices, their mix weighted by where the white balance's B/R ratio falls
 * between the two illuminants' (in log space, clamped to [0, 1]); else
 * whichever exists (A, then D65, then any other), else the identity.  Any
 * matrix may be NULL; wb_coeffs (4 values, NaN read as 0) may be NULL. */
int imops_dng_calibration_matrix(const float* color_matrix_a, const float* color_matrix_d65,
                                 const float* color_matrix_other, const float* wb_coeffs, float out[9]);

/* Raw ingest: u16 sensor frame -> normalised f32 CFA plane of the crop window,
 * value (DN - black) / (white - black).  `src` is a host pointer, or a device
 * pointer when src_on_device != 0.  `cfa_pattern` is the sensor pattern string
 * (cfa.rs:98-131); the crop origin shifts it as get_cfa does (demosaic.rs:93-98).
 * Replaces the crop_and_normalize step of parse_raw_image (extern_pipeline.rs:239-324,
 * demosaic.rs:61-91). */
int imops_image_set_raw_u16(imops_ctx* ctx, imops_image* img, const uint16_t* src, int src_on_device,
                            size_t full_width, size_t full_height, size_t crop_x, size_t crop_y,
                            size_t crop_width, size_t crop_height, float black, float white,
                            const char* cfa_pattern);
/* An already normalised CFA plane (Image.raw_data). */
int imops_image_set_raw_f32(imops_ctx* ctx, imops_image* img, const float* src, int src_on_device, size_t width,
                            size_t height, const char* cfa_pattern, size_t crop_x, size_t crop_y);
/* Interleaved RGB f32 (Image.rgb_data) in colour space `color_space`. */
int imops_image_set_rgb_f32(imops_ctx* ctx, imops_image* img, const float* src, int src_on_device, size_t width,
                            size_t height, int color_space);
/* Copy the interleaved RGB f32 out (count = width*height*3 floats). */
int imops_image_get_rgb_f32(imops_ctx* ctx, const imops_image* img, float* dst, int dst_on_device, size_t count);
int imops_image_get_info(const imops_image* img, size_t* width, size_t* height, int* color_space, int* has_rgb);
/* Encoded readback (count = width*height*3 values), encoded on the device so only the
 * encoded bytes cross the link; orientation 0 none, 1 Rotate90, 2 Rotate180,
 * 3 Rotate270 (apply_orientation, src/output.rs:31-70: rotated frames are height x width).
 * u8: (clamp(c, 0, 1) * 255) truncated, NaN -> 0 (get_image_rgb_data_c,
 * pichromatic-pipeline/src/extern_pipeline.rs:362-395; save_jpeg, src/output.rs:221-233).
 * f16: IEEE binary16 bits, round to nearest even (save_exr's f16::from_f32, src/output.rs:201-209). */
int imops_image_get_rgb_u8(imops_ctx* ctx, const imops_image* img, uint8_t* dst, int dst_on_device, int orientation,
                           size_t count);
int imops_image_get_rgb_f16(imops_ctx* ctx, const imops_image* img, uint16_t* dst, int dst_on_device,
                            int orientation, size_t count);
/* Device address of the interleaved RGB buffer (valid until the next stage call). */
int imops_image_rgb_device_ptr(const imops_image* img, float** out);

/* Metadata used by the stages (ImageMetadata fields wb_coeffs,
 * calibration_matrix_d65, shutter_seconds). */
int imops_image_set_wb_coeffs(imops_image* img, const float wb[4]);
int imops_image_set_calibration_matrix(imops_image* img, const float* m, size_t len);
int imops_image_set_shutter_seconds(imops_image* img, float seconds);
/* ImageMetadata baseline_exposure, f_number and iso (BaselineExposureCompensation). */
int imops_image_set_exposure_metadata(imops_image* img, float baseline_ev, float f_number, float iso);
/* ImageMetadata opcode_list3: the DNG OpcodeList3 bytes, big-endian as stored (Vignette). */
int imops_image_set_opcode_list3(imops_image* img, const uint8_t* bytes, size_t len);

/* Stages ---------------------------------------------------------------- */

/* Module<Demosaic>::process (modules.rs:629-668): raw plane -> RGB, colour
 * space reset to None.  AMaZE (demosaic.rs:358-1379) needs a Bayer pattern and
 * rejects sizes where a dimension mod 96 is 1..15 (the reference indexes past
 * its tile there); markesteijn (demosaic.rs:108-186) takes any pattern. */
int imops_demosaic(imops_ctx* ctx, imops_image* img, int algorithm);

/* AMaZE demosaic that allocates nothing.  No counterpart in the reference: it
 * is for callers that cannot allocate device memory (code under HIP graph
 * capture, serving loops, harnesses that forbid hipMalloc).  Reads the
 * caller's device f32 CFA plane (width*height, normalised, of the crop window
 * whose origin is crop_x, crop_y in the pattern's frame) and writes the
 * caller's device RGB buffer (width*height*3 f32), both on ctx's stream; the
 * call only enqueues (synchronise the context, or order with events, before
 * reading the output).  Same results and the same size rules as
 * imops_demosaic(IMOPS_DEMOSAIC_AMAZE).  Its tile scratch is a static device
 * array of 64 tile slots (69 MB) that each GPU holds until the process ends;
 * the slot count bounds the tiles in flight, not the image size, and every
 * pass rewrites the slot planes it reads, so nothing has to be cleared between
 * calls.  All calls on one GPU share that array: calls that overlap on one GPU
 * (two contexts' streams, or two threads) race and must be serialised by the
 * caller. */
int imops_demosaic_amaze_noalloc(imops_ctx* ctx, const float* cfa_plane, float* rgb_out, size_t width,
                                 size_t height, const char* cfa_pattern, size_t crop_x, size_t crop_y);

/* Module<CFACoeffs>::process (modules.rs:533-540, cfa_coeffs.rs:4-12);
 * wb = NULL uses the image's wb_coeffs. */
int imops_cfa_coeffs(imops_ctx* ctx, imops_image* img, const float* wb);

/* Module<Exp>::process (modules.rs:308-311, exp.rs:4-11). */
int imops_exp(imops_ctx* ctx, imops_image* img, float ev);

/* Module<Gamma>::process (modules.rs:431-433, tone_map.rs:38-49). */
int imops_gamma(imops_ctx* ctx, imops_image* img, float gamma);

/* Module<CST>::process (modules.rs:576-592): camera_cst with the image's
 * calibration matrix when the colour space is None (cst.rs:12-48), else
 * cst (cst.rs:6-10). */
int imops_cst(imops_ctx* ctx, imops_image* img, int target_color_space);

/* Image::camera_cst with an explicit XYZ->camera matrix (3 rows, cst.rs:12-48). */
int imops_camera_cst(imops_ctx* ctx, imops_image* img, int target_color_space, const float* calibration_matrix,
                     size_t len);

/* BM3D (pichromatic/src/bm3d.rs:34-61) in plain FP32, bit-exact with the
 * reference's arithmetic.  imops_bm3d takes the reference's single knob
 * (Module<BM3D>, modules.rs:506-510; Bm3dParams::from_intensity, bm3d.rs:22-32);
 * imops_bm3d_ex takes every parameter explicitly. */
typedef struct imops_bm3d_params {
  float sigma;          /* noise sigma in 0..255 units (reference: 80*clamp(intensity)) */
  float lambda;         /* hard-threshold multiplier (2 + 2.5*clamp(intensity)) */
  float max_dist;       /* step-1 mean-SSD match limit (3000 + 20000*clamp(intensity)); step 2 uses half */
  float chroma_scale;   /* sigma multiplier for Cb/Cr (1.8) */
  float detail;         /* luma detail blend strength (clamp(0.5*intensity, 0, 0.5)) */
  int stride;           /* reference-patch grid step (6) */
  int search_window;    /* odd, <= 41 (19) */
  int max_group;        /* 1..16, groups are the largest power of two <= min(this, matches) (16) */
  int max_candidates;   /* accepted-candidate cap in scan order (1024) */
  int precision;        /* IMOPS_BM3D_FP32 (0): plain FP32 throughout;
                           IMOPS_BM3D_MIXED_MATCH (2): FP16 matrix-core block matching with
                           FP32 re-ranking (same groups as FP32), FP32 filter;
                           IMOPS_BM3D_MIXED (1): that matcher plus the collaborative filter's
                           DCT / Hadamard transforms on FP16 matrix cores (FP32 accumulation) */
  int group_transform;  /* IMOPS_BM3D_HADAMARD (0, the reference's) or IMOPS_BM3D_HAAR (1) */
} imops_bm3d_params;

/* Fill `out` with the parameters intensity maps to (from_intensity). */
int imops_bm3d_default_params(float intensity, imops_bm3d_params* out);
int imops_bm3d(imops_ctx* ctx, imops_image* img, float intensity);
int imops_bm3d_ex(imops_ctx* ctx, imops_image* img, const imops_bm3d_params* params);

/* Chroma-only denoise (Image::chroma_bm3d, pichromatic/src/pixel.rs:122-126;
 * chroma_bm3d + dct_chroma_denoise, bm3d.rs:705-759; stage ChromaDenoise,
 * modules.rs:688-722): Y kept, every 8x8 Cb/Cr patch at a 6-pixel step DCT'd,
 * hard-thresholded at (2 + 2.5 i) 80 i 1.8 (i = intensity clamped to
 * [0.001, 1]), inverse-DCT'd and aggregated with the Kaiser window in i64
 * fixed point; bit-exact with the reference. */
int imops_chroma_bm3d(imops_ctx* ctx, imops_image* img, float intensity);

/* Luma-guided chroma denoise (Image::chroma_denoise, pixel.rs:128-137;
 * chroma_denoise.rs:14-152; stage LumaGuidedChromaDenoise, modules.rs:724-760):
 * guided filter (radius, epsilon) of C_R = R - Y and C_B = B - Y with the luma
 * as guide, box means as the reference's running sums; bit-exact with the
 * reference's f32 arithmetic. */
int imops_luma_chroma_denoise(imops_ctx* ctx, imops_image* img, int radius, float epsilon);

/* Display tail: elementwise stages, fused with neighbouring elementwise stages
 * by the stage runner.  The Oklch conversions restate Oklab from Ottosson's
 * published matrices through linear sRGB, as the colour crate converts. */
/* Image::contrast (pixel.rs:107-110; contrast.rs:4-8): 0.185 (x / 0.185)^c. */
int imops_contrast(imops_ctx* ctx, imops_image* img, float c);
/* Image::highlight_reconstruction (pixel.rs:111-114; highlight_reconstruction.rs:5-14);
 * needs the image's white balance (green clip = wb[1]). */
int imops_highlight_reconstruction(imops_ctx* ctx, imops_image* img);
/* Image::lch (pixel.rs:60-68; lch.rs:5-21): Oklch L, C, hue scaled in the image's colour space. */
int imops_lch(imops_ctx* ctx, imops_image* img, float lc, float cc, float hc);
/* Image::sigmoid_tone_map (pixel.rs:91-94; tone_map.rs:7-170): ACES reference gamut
 * compression, square sigmoid, chroma roll-off by luminance^5, on ACEScg. */
int imops_sigmoid_tone_map(imops_ctx* ctx, imops_image* img);
/* BaselineExposureCompensation (modules.rs:332-371): 2^baseline_ev, then the
 * absolute-luminance gain 12.5 N^2 / (t S) when shutter, f-number and ISO are set. */
int imops_baseline_exposure(imops_ctx* ctx, imops_image* img);
/* Image::vignette (pixel.rs:149-160; vignette.rs:3-99): FixVignetteRadial opcodes of
 * the image's OpcodeList3 (none set: no change). */
int imops_vignette(imops_ctx* ctx, imops_image* img, float strength);

/* Film simulation (film::process, pichromatic/src/film/mod.rs:77-129) on an
 * absolute-luminance ACEScg image; output ACEScg.  Module<Film>
 * (modules.rs:841-876).  stock: 0 BwStub, 1 ColorNeg200, 2 Portra400,
 * 3 Ektar100, 4 FujiPro400H, 5 EktachromeE100, 6 TriX400 (StockId order);
 * film_format: 0 Film35mm, 1 Film6x6, 2 Film4x5; output: 0 NegativeLinear,
 * 1 PositiveLinear.  The shutter time is the image's shutter_seconds, else
 * 1/box ISO.  A stock whose tables the library does not carry returns
 * IMOPS_E_UNSUPPORTED. */
int imops_film(imops_ctx* ctx, imops_image* img, int stock, int film_format, uint64_t seed, int output);

/* Film stage with the extension parameters (the Film stage fields of a stage
 * list: stock, film_format, seed, output, precision, spectral_bins,
 * stock_seed, grain, grain_sigma, halation_radius, halation_strength).
 * imops_film_default_params fills the reference's behaviour; the extensions:
 *   stock 7 = the seeded synthetic colour negative (stock_seed);
 *   spectral_bins 16 (the reference's 400..700 nm / 20 nm grid) or 81
 *     (380..780 nm / 5 nm);
 *   grain 0 = the reference's SplitMix64 Irwin-Hall rows, 1 = Philox4x32-10
 *     normals; grain_sigma >= 0 replaces the density-dependent amplitude by a
 *     constant density sigma;
 *   halation_radius >= 0 (pixels) and halation_strength >= 0 replace the
 *     stock's halation width (psf_halation / pitch) and gains (r_e bleed_e);
 *   precision 0 = the reference's FP64 spectral arithmetic, 1 = FP32. */
typedef struct imops_film_params {
  int stock, film_format, output, precision, spectral_bins, grain;
  uint64_t seed, stock_seed;
  float grain_sigma, halation_radius, halation_strength;
} imops_film_params;
int imops_film_default_params(imops_film_params* out);
int imops_film_ex(imops_ctx* ctx, imops_image* img, const imops_film_params* params);

/* Differentiable film model --------------------------------------------------
 * No counterpart in the reference (its film simulation has no gradients); it is
 * film::process (film/mod.rs:77-129) in the NegativeLinear form with shutter
 * 1/box ISO and a given pixel pitch, grain, DIR inhibition and adjacency off
 * (development/mod.rs:18-48), on the stock's spectral grid (16 samples 400..700
 * nm as the reference, or 81 samples 380..780 nm), parameterised by the emulsion
 * sensitivities on that grid (stock/*.rs layer curves), the characteristic curve
 * D = d_max f^(1/gamma) (reduction.rs:16-68) with, for a knot-curve stock (the
 * seeded synthetic stock: 5 knots), the developable fraction f at each knot, a
 * halation gain multiplier and a halation PSF width multiplier
 * (halation.rs:24-132); the scan is normalised by the nominal stock's Dmin.
 * Images are N x H x W x 3 ACEScg (absolute luminance) f32 on the device.
 * imops_film_fit_backward returns dL/dparams for dL/dout = grad_out, summed over
 * the images.  Every stage runs once over all N images (one launch per stage,
 * not per image).  The 81-bin grid and knot curves take stocks of up to 4
 * emulsions.
 *
 * `state` (optional, device, imops_film_fit_state_floats floats): the forward
 * pass saves the blurred exposure planes and the halation bounce there, and a
 * backward pass given the same state, parameters and images reads them instead
 * of running the exposure and blurs again.  NULL: nothing saved / recomputed. */
typedef struct imops_film_fit_model {
  int stock;             /* as imops_film_ex: 0..6 the reference's stocks, 7 the seeded synthetic stock */
  int spectral_bins;     /* 16 or 81 */
  uint64_t stock_seed;   /* stock 7 */
  float pitch_um;        /* pixel pitch */
} imops_film_fit_model;

typedef struct imops_film_fit_params {
  int n_emulsions;       /* emulsion layers top to bottom (the stock's count) */
  int n_bins;            /* the model's spectral_bins */
  int n_knots;           /* the stock's curve knots (0: crystal LUTs) */
  float sens[8][81];     /* sensitivity samples on the grid (n_bins used) */
  float knot_frac[8][8]; /* developable fraction at each knot (knot positions are the stock's) */
  float d_max[8];
  float gamma[8];
  float hal_strength;    /* 1 = the stock's halation gain */
  float hal_radius;      /* 1 = the stock's halation PSF width */
} imops_film_fit_params;

int imops_film_fit_nominal(const imops_film_fit_model* model, imops_film_fit_params* out);
/* log10 fluence of each knot of one emulsion's curve (n_knots values) */
int imops_film_fit_knots(const imops_film_fit_model* model, int emulsion, double* log10_fluence);
int imops_film_fit_state_floats(const imops_film_fit_model* model, size_t n_images, size_t width, size_t height,
                                size_t* floats);
int imops_film_fit_forward(imops_ctx* ctx, const imops_film_fit_model* model, const imops_film_fit_params* params,
                           const float* rgb_in, float* rgb_out, size_t n_images, size_t width, size_t height,
                           float* state);
int imops_film_fit_backward(imops_ctx* ctx, const imops_film_fit_model* model, const imops_film_fit_params* params,
                            const float* rgb_in, const float* grad_out, size_t n_images, size_t width, size_t height,
                            const float* state, imops_film_fit_params* grad);
/* One fitting step of the L2 loss L = sum over images, pixels and channels of
 * (model(rgb_in) - target)^2: the exposure and blurs, then the reverse pass,
 * whose per-pixel kernel evaluates the model output (f32) from the quantities
 * it already forms, so the output image is never written.  Returns L in *loss
 * and dL/dparams in *grad (the same gradient as imops_film_fit_backward with
 * grad_out = 2 (out - target), up to the f32 evaluation of out); one
 * synchronisation, for the readback.  No counterpart in the reference. */
int imops_film_fit_l2_step(imops_ctx* ctx, const imops_film_fit_model* model, const imops_film_fit_params* params,
                           const float* rgb_in, const float* target, size_t n_images, size_t width, size_t height,
                           imops_film_fit_params* grad, double* loss);

/* Pipeline --------------------------------------------------------------- */

/* Parse a TOML or JSON stage list ([[pipeline_modules]] tables, names matched
 * case-insensitively; config.rs:13-34).  Replaces get_pixel_pipeline_c. */
int imops_pipeline_parse(const char* text, size_t len, imops_pipeline** out);
int imops_pipeline_free(imops_pipeline* p);
size_t imops_pipeline_num_stages(const imops_pipeline* p);
/* Canonical stage name ("Demosaic", "CST", ...) of stage i, NULL if out of range. */
const char* imops_pipeline_stage_name(const imops_pipeline* p, size_t i);
/* Run the stages in order on the image (run_pixel_pipeline, pipeline.rs:4-13). */
int imops_pipeline_run(imops_ctx* ctx, imops_pipeline* p, imops_image* img);
/* Run the stages on each of n frames (any mix of sizes and CFA patterns the stage
 * list accepts); no counterpart in the reference, which processes one image per
 * call.  use_graphs = 0: imops_pipeline_run per frame.  use_graphs != 0: frames
 * are grouped by shape (size, CFA and crop, white balance, calibration, shutter,
 * colour space); the first frame of each group runs directly, largest group
 * first, then each group with more frames is captured once into a HIP graph on
 * a staging image of its shape and replayed per frame (input copied in, output
 * copied out).  Results are identical either way.  Timing regions are not
 * recorded for replayed frames. */
int imops_pipeline_run_burst(imops_ctx* ctx, imops_pipeline* p, imops_image* const* frames, size_t n, int use_graphs);

/* Synthetic inputs --------------------------------------------------------- */

/* Seeded scene in linear RGB [0,1] (gradients + 40 rectangles + sinusoids). */
int imops_synth_scene_rgb(float* out, size_t width, size_t height, uint64_t frame);
/* The same scene through a CFA with Poisson shot noise (full_well e-) and
 * Gaussian read noise (read_sigma DN), quantised to u16 in [0, white]. */
int imops_synth_raw_u16(uint16_t* out, size_t width, size_t height, const char* cfa_pattern, uint64_t frame,
                        float full_well, float black, float white, float read_sigma);

#ifdef __cplusplus
}
#endif

#endif /* IMOPS_GPU_H */
