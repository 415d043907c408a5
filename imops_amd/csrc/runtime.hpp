// Internal host-side state behind the C ABI: contexts, device images, error
// reporting and the shared launch helpers.
#pragma once
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>
#include <functional>
#include <map>
#include <memory>
#include <set>
#include <string>
#include <vector>

#include "cfa_table.hpp"
#include "imops_gpu.h"
#include "pointwise.hpp"

namespace imops {

// Failure carried to the C boundary: status code + message.
struct Error {
  int code;
  std::string msg;
};

void set_error(int code, const std::string& msg);
void clear_error();

#define IMOPS_HIP_CHECK(expr)                                                                  \
  do {                                                                                         \
    hipError_t e__ = (expr);                                                                   \
    if (e__ != hipSuccess)                                                                     \
      throw ::imops::Error{IMOPS_E_HIP, std::string(#expr) + ": " + hipGetErrorString(e__)};   \
  } while (0)

// State a stage keeps in its context across calls (prepared tables etc.).
struct StageCache {
  virtual ~StageCache() = default;
};

// Device buffer that only grows.
struct DevBuf {
  void* ptr = nullptr;
  size_t bytes = 0;
  void reserve(size_t n);
  void release();
  template <class T> T* as() const { return static_cast<T*>(ptr); }
};

}  // namespace imops

// Scratch buffers a context owns, by use.
enum ScratchId : int { SCR_AMAZE = 0, SCR_STAGING, SCR_BM3D_0, SCR_BM3D_1, SCR_BM3D_2, SCR_BM3D_3, SCR_FILM_0,
                       SCR_FILM_1, SCR_FILM_2, SCR_FILM_3, SCR_DEMOSAIC, SCR_COUNT };

namespace imops {
// Optional per-region GPU timing: HIP events recorded around launches on the
// context stream while enabled.
struct KernelTimer {
  bool enabled = false;
  struct Region {
    std::string name;
    std::vector<std::pair<hipEvent_t, hipEvent_t>> events;
  };
  std::vector<Region> regions;
  // Named event counts recorded while timing is on (e.g. patches a stage
  // handed to a slower path); reported through imops_ctx_timing as the sum.
  // Each is a device word copied asynchronously into a pinned host slot, read
  // once the stream has drained, so counting never stalls the stream.
  struct Count {
    std::string name;
    const uint32_t* host;
  };
  std::vector<Count> counts;
  std::vector<uint32_t*> pinned;  // blocks of kPinnedWords slots, kept until release()
  size_t pinned_used = 0;
  static constexpr size_t kPinnedWords = 1024;
  Region& region(const std::string& name);
  // queues the copy of n consecutive device words, counted under names[0..n)
  void count_async(const char* const* names, const uint32_t* dev, int n, hipStream_t s);
  void clear();
  void release();  // clear() and free the pinned blocks (stream drained)
};
}  // namespace imops

struct imops_ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  int num_cus = 0;
  float* amaze_noalloc = nullptr;  // this device's static AMaZE scratch (allocation-free entry point)
  imops::DevBuf scr[SCR_COUNT];
  imops::KernelTimer timer;
  std::unique_ptr<imops::StageCache> film_cache;  // prepared film runs by shape
  std::unique_ptr<imops::StageCache> fit_cache;   // film-model tables on the device and their pinned staging
  imops::DevBuf bm3d_frags;                       // FP16 DCT fragments of the mixed BM3D filter
  bool bm3d_tables_ready = false;                 // constant tables of the FP32 BM3D filter uploaded
  // Set while the stage code runs under stream capture: stages then only
  // enqueue work (no host uploads, synchronisation or allocation).
  bool capturing = false;
  // Set while a burst runs: per-shape prepared state (film runs) is then kept
  // rather than evicted, since the burst's captures rely on it.
  bool in_burst = false;
  // Launch-shape autotuning: when on, the first call of a tunable kernel at a
  // new shape times every candidate and keeps the fastest (by key).
  bool autotune = false;
  std::map<std::string, int> tuned;
  // Images created on this context and not yet freed: destroying the context
  // releases their device buffers and detaches them, so a later
  // imops_image_free only deletes the handle.
  std::set<imops_image*> images;
};

namespace imops {
// Records start/stop events on the context stream around its scope when timing is on.
// A roctx range on the calling host thread for the scope (what a profiler's
// marker trace shows: `rocprofv3 --marker-trace`); the library's stages and
// timed regions open one each.
class MarkerRange {
 public:
  explicit MarkerRange(const char* name);
  ~MarkerRange();
  MarkerRange(const MarkerRange&) = delete;
  MarkerRange& operator=(const MarkerRange&) = delete;
};

class TimedRegion {
 public:
  TimedRegion(imops_ctx* ctx, const char* name);
  ~TimedRegion();
  TimedRegion(const TimedRegion&) = delete;
  TimedRegion& operator=(const TimedRegion&) = delete;

 private:
  MarkerRange range_;
  imops_ctx* ctx_;
  hipEvent_t stop_ = nullptr;
};
}  // namespace imops

struct imops_image {
  imops_ctx* ctx = nullptr;
  size_t width = 0, height = 0;
  bool has_raw = false, has_rgb = false;
  imops::DevBuf raw;   // f32 CFA plane (width*height)
  // A caller's device u16 frame bound without a copy (imops_image_bind_raw_u16):
  // the raw plane is pending until a stage needs it as f32 (AMaZE reads the
  // u16 samples directly).  Not part of copy_state_from: it is buffer state.
  const uint16_t* raw16 = nullptr;
  size_t raw16_pitch = 0, raw16_x = 0, raw16_y = 0;
  float raw16_black = 0.0f, raw16_white = 1.0f;
  imops::DevBuf rgb;   // f32 interleaved RGB (width*height*3)
  imops::DevBuf tmp;   // second RGB buffer for out-of-place stages
  imops::CfaTable cfa;
  bool has_cfa = false;
  size_t crop_x = 0, crop_y = 0;
  float wb[4] = {1, 1, 1, 1};
  bool has_wb = false;
  std::vector<float> calib;
  int color_space = IMOPS_CS_NONE;
  float shutter = 0.0f;
  bool has_shutter = false;
  // exposure metadata of BaselineExposureCompensation (ImageMetadata baseline_exposure, f_number, iso)
  float baseline_ev = 0.0f, f_number = 0.0f, iso = 0.0f;
  bool has_baseline = false, has_f_number = false, has_iso = false;
  std::vector<uint8_t> opcode_list3;  // DNG OpcodeList3 bytes (Vignette)
  bool has_opcodes = false;
  void swap_rgb_tmp() { std::swap(rgb, tmp); }
  // Every field but the device buffers and the context: the per-frame state a
  // stage list reads or updates (shape, plane flags, CFA and crop, colour
  // space and all metadata).
  void copy_state_from(const imops_image& o) {
    width = o.width;
    height = o.height;
    has_raw = o.has_raw;
    has_rgb = o.has_rgb;
    cfa = o.cfa;
    has_cfa = o.has_cfa;
    crop_x = o.crop_x;
    crop_y = o.crop_y;
    std::copy(o.wb, o.wb + 4, wb);
    has_wb = o.has_wb;
    calib = o.calib;
    color_space = o.color_space;
    shutter = o.shutter;
    has_shutter = o.has_shutter;
    baseline_ev = o.baseline_ev;
    f_number = o.f_number;
    iso = o.iso;
    has_baseline = o.has_baseline;
    has_f_number = o.has_f_number;
    has_iso = o.has_iso;
    opcode_list3 = o.opcode_list3;
    has_opcodes = o.has_opcodes;
  }
};

namespace imops {

// Candidate `fallback` unless the context autotunes: then, once per key, each
// candidate is run (after `reset`) twice on the context stream and timed with
// HIP events, and the fastest is remembered.  Every candidate must compute the
// same result; the caller runs the chosen one afterwards.
int tuned_choice(imops_ctx* ctx, const std::string& key, const std::vector<int>& candidates, int fallback,
                 const std::function<void(int)>& run, const std::function<void()>& reset);

// Turns a bound u16 frame into the f32 raw plane (no-op when none is pending).
void materialize_raw(imops_ctx* ctx, imops_image* img);

// Stage implementations shared by the per-stage C calls and the pipeline runner.
void stage_demosaic(imops_ctx* ctx, imops_image* img, int algorithm);
void stage_pointwise(imops_ctx* ctx, imops_image* img, const PointwiseProgram& prog);
// Display-tail program builders (each adds at most kMaxOpsPerStage ops).
constexpr int kMaxOpsPerStage = 6;
void push_contrast(PointwiseProgram& prog, float c);
void push_highlight(PointwiseProgram& prog, const float wb[4]);
void push_lch(PointwiseProgram& prog, int color_space, float lc, float cc, float hc);
void push_sigmoid(PointwiseProgram& prog);
// BaselineExposureCompensation: ev scale then the absolute-luminance gain, from the image's metadata.
void push_baseline_exposure(PointwiseProgram& prog, const imops_image* img);
// Vignette (vignette.rs:3-99) from the image's OpcodeList3; no opcodes: nothing to do.
void stage_vignette(imops_ctx* ctx, imops_image* img, float strength);
// Append the ops of one colour conversion / stage to a pointwise program.
void push_convert(PointwiseProgram& prog, int src, int dst);
void push_camera_cst(PointwiseProgram& prog, const float* calib, size_t len, int target);
void push_exp(PointwiseProgram& prog, float ev);
void push_gamma(PointwiseProgram& prog, float gamma);
void push_wb(PointwiseProgram& prog, const float wb[4]);
void require_rgb(const imops_image* img, const char* stage);
int parse_color_space(const std::string& name);

// Run f, converting exceptions into a status code + last error.
template <class F>
int guarded(F&& f) {
  try {
    f();
    clear_error();
    return IMOPS_OK;
  } catch (const Error& e) {
    set_error(e.code, e.msg);
    return e.code;
  } catch (const std::bad_alloc&) {
    set_error(IMOPS_E_HIP, "host allocation failed");
    return IMOPS_E_HIP;
  } catch (const std::exception& e) {
    set_error(IMOPS_E_INVALID, e.what());
    return IMOPS_E_INVALID;
  } catch (...) {
    set_error(IMOPS_E_INVALID, "unknown failure");
    return IMOPS_E_INVALID;
  }
}

}  // namespace imops
