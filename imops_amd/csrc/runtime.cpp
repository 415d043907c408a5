// Host runtime behind the C ABI: contexts, device-resident images and the
// per-stage entry points.  Errors never cross the boundary as exceptions:
// each entry point returns a status and records the message for
// imops_last_error() (the reference's catch_panic + LAST_ERROR,
// pichromatic-pipeline/src/extern_pipeline.rs:18-58).
#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstring>
#include <string>

#include "amaze.hpp"
#include "peaks.hpp"
#include "runtime.hpp"

#include <rocprofiler-sdk-roctx/roctx.h>

namespace imops {

namespace {
thread_local std::string t_last_error;
}

void set_error(int code, const std::string& msg) {
  (void)code;
  t_last_error = msg;
}
void clear_error() {}

// Allocation failures injected by imops_debug_fail_allocations (a test hook):
// that many of the next device allocations fail as hipMalloc would.
static std::atomic<int> g_injected_alloc_failures{0};

void DevBuf::reserve(size_t n) {
  if (n <= bytes) return;
  release();
  void* p = nullptr;
  int pending = g_injected_alloc_failures.load();
  while (pending > 0 && !g_injected_alloc_failures.compare_exchange_weak(pending, pending - 1)) {
  }
  hipError_t e = pending > 0 ? hipErrorOutOfMemory : hipMalloc(&p, n);
  if (e != hipSuccess)
    throw Error{IMOPS_E_HIP, "hipMalloc(" + std::to_string(n) + " bytes): " + hipGetErrorString(e)};
  ptr = p;
  bytes = n;
}

void DevBuf::release() {
  if (ptr) (void)hipFree(ptr);
  ptr = nullptr;
  bytes = 0;
}

int tuned_choice(imops_ctx* ctx, const std::string& key, const std::vector<int>& candidates, int fallback,
                 const std::function<void(int)>& run, const std::function<void()>& reset) {
  const auto it = ctx->tuned.find(key);
  if (it != ctx->tuned.end()) return it->second;
  if (!ctx->autotune || ctx->capturing || candidates.size() < 2) return fallback;
  hipEvent_t e0, e1;
  IMOPS_HIP_CHECK(hipEventCreate(&e0));
  IMOPS_HIP_CHECK(hipEventCreate(&e1));
  int best = fallback;
  float best_ms = 0.0f;
  try {
    for (int c : candidates) {
      float ms = 0.0f;
      for (int rep = 0; rep < 2; ++rep) {
        reset();
        IMOPS_HIP_CHECK(hipEventRecord(e0, ctx->stream));
        run(c);
        IMOPS_HIP_CHECK(hipEventRecord(e1, ctx->stream));
        IMOPS_HIP_CHECK(hipEventSynchronize(e1));
        float t = 0.0f;
        IMOPS_HIP_CHECK(hipEventElapsedTime(&t, e0, e1));
        ms = rep == 0 ? t : std::min(ms, t);
      }
      if (best_ms == 0.0f || ms < best_ms) {
        best_ms = ms;
        best = c;
      }
    }
  } catch (...) {
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    throw;
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  ctx->tuned[key] = best;
  return best;
}

KernelTimer::Region& KernelTimer::region(const std::string& name) {
  for (auto& r : regions)
    if (r.name == name) return r;
  regions.push_back(Region{name, {}});
  return regions.back();
}

void KernelTimer::clear() {
  for (auto& r : regions)
    for (auto& e : r.events) {
      (void)hipEventDestroy(e.first);
      (void)hipEventDestroy(e.second);
    }
  regions.clear();
  counts.clear();
  pinned_used = 0;
}

void KernelTimer::count_async(const char* const* names, const uint32_t* dev, int n, hipStream_t s) {
  if (n <= 0 || (size_t)n > kPinnedWords) return;
  size_t blk = pinned_used / kPinnedWords, off = pinned_used % kPinnedWords;
  if (off + n > kPinnedWords) {  // the words stay in one block
    ++blk;
    off = 0;
  }
  if (blk == pinned.size()) {
    void* b = nullptr;
    IMOPS_HIP_CHECK(hipHostMalloc(&b, kPinnedWords * sizeof(uint32_t), hipHostMallocDefault));
    pinned.push_back(static_cast<uint32_t*>(b));
  }
  uint32_t* h = pinned[blk] + off;
  pinned_used = blk * kPinnedWords + off + n;
  IMOPS_HIP_CHECK(hipMemcpyAsync(h, dev, n * sizeof(uint32_t), hipMemcpyDeviceToHost, s));
  for (int i = 0; i < n; ++i) counts.push_back({names[i], h + i});
}

void KernelTimer::release() {
  clear();
  for (uint32_t* b : pinned) (void)hipHostFree(b);
  pinned.clear();
}

MarkerRange::MarkerRange(const char* name) { roctxRangePushA(name); }
MarkerRange::~MarkerRange() { roctxRangePop(); }

TimedRegion::TimedRegion(imops_ctx* ctx, const char* name) : range_(name), ctx_(ctx) {
  if (!ctx_ || !ctx_->timer.enabled) return;
  hipEvent_t start = nullptr;
  if (hipEventCreate(&start) != hipSuccess || hipEventCreate(&stop_) != hipSuccess) {
    if (start) (void)hipEventDestroy(start);
    stop_ = nullptr;
    return;
  }
  (void)hipEventRecord(start, ctx_->stream);
  ctx_->timer.region(name).events.push_back({start, stop_});
}

TimedRegion::~TimedRegion() {
  if (stop_) (void)hipEventRecord(stop_, ctx_->stream);
}

void require_rgb(const imops_image* img, const char* stage) {
  if (!img->has_rgb) throw Error{IMOPS_E_STATE, std::string(stage) + ": image has no RGB data (demosaic first)"};
}

int parse_color_space(const std::string& name) {
  // serde_plain names of color::ColorSpaceTag (modules.rs:578-582)
  if (name == "Srgb") return IMOPS_CS_SRGB;
  if (name == "LinearSrgb") return IMOPS_CS_LINEAR_SRGB;
  if (name == "AcesCg") return IMOPS_CS_ACESCG;
  if (name == "XyzD65") return IMOPS_CS_XYZD65;
  return -100;
}

// ---------------------------------------------------------------- colour ops
namespace {


// 3x3 helpers in double for deriving colour matrices on the host.
static void m3mul(const double a[9], const double b[9], double o[9]) {
  double t[9];
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) t[i * 3 + j] = a[i * 3] * b[j] + a[i * 3 + 1] * b[3 + j] + a[i * 3 + 2] * b[6 + j];
  for (int k = 0; k < 9; ++k) o[k] = t[k];
}
static void m3inv(const double m[9], double o[9]) {
  const double c00 = m[4] * m[8] - m[5] * m[7], c01 = m[5] * m[6] - m[3] * m[8], c02 = m[3] * m[7] - m[4] * m[6];
  const double id = 1.0 / (m[0] * c00 + m[1] * c01 + m[2] * c02);
  o[0] = c00 * id; o[1] = (m[2] * m[7] - m[1] * m[8]) * id; o[2] = (m[1] * m[5] - m[2] * m[4]) * id;
  o[3] = c01 * id; o[4] = (m[0] * m[8] - m[2] * m[6]) * id; o[5] = (m[2] * m[3] - m[0] * m[5]) * id;
  o[6] = c02 * id; o[7] = (m[1] * m[6] - m[0] * m[7]) * id; o[8] = (m[0] * m[4] - m[1] * m[3]) * id;
}
// RGB->XYZ of primaries (xy chromaticities) normalised to white (wx, wy).
static void rgb_to_xyz(const double xy[6], double wx, double wy, double o[9]) {
  double p[9];
  for (int k = 0; k < 3; ++k) {
    const double x = xy[2 * k], y = xy[2 * k + 1];
    p[k] = x / y; p[3 + k] = 1.0; p[6 + k] = (1.0 - x - y) / y;
  }
  double pi[9], s[3];
  m3inv(p, pi);
  const double w[3] = {wx / wy, 1.0, (1.0 - wx - wy) / wy};
  for (int i = 0; i < 3; ++i) s[i] = pi[i * 3] * w[0] + pi[i * 3 + 1] * w[1] + pi[i * 3 + 2] * w[2];
  for (int i = 0; i < 3; ++i)
    for (int k = 0; k < 3; ++k) o[i * 3 + k] = p[i * 3 + k] * s[k];
}
// Bradford adaptation from white (sx, sy) to white (dx, dy).
static void bradford(double sx, double sy, double dx, double dy, double o[9]) {
  const double b[9] = {0.8951, 0.2664, -0.1614, -0.7502, 1.7135, 0.0367, 0.0389, -0.0685, 1.0296};
  double bi[9];
  m3inv(b, bi);
  const double ws[3] = {sx / sy, 1.0, (1.0 - sx - sy) / sy}, wd[3] = {dx / dy, 1.0, (1.0 - dx - dy) / dy};
  double d[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
  for (int i = 0; i < 3; ++i) {
    const double cs = b[i * 3] * ws[0] + b[i * 3 + 1] * ws[1] + b[i * 3 + 2] * ws[2];
    const double cd = b[i * 3] * wd[0] + b[i * 3 + 1] * wd[1] + b[i * 3 + 2] * wd[2];
    d[i * 4] = cd / cs;
  }
  m3mul(bi, d, o);
  m3mul(o, b, o);
}
// ACEScg (AP1 primaries, ACES white) -> linear sRGB through XYZ(D65), with the
// Bradford transform between the two whites.
static void acescg_to_linear_srgb(const double xyz_to_lin[9], double o[9]) {
  const double ap1[6] = {0.713, 0.293, 0.165, 0.830, 0.128, 0.044};
  double a2x[9], cat[9];
  rgb_to_xyz(ap1, 0.32168, 0.33767, a2x);
  bradford(0.32168, 0.33767, 0.3127, 0.3290, cat);
  m3mul(cat, a2x, o);
  m3mul(xyz_to_lin, o, o);
}

struct ColourMats {
  float xyz_to_lin[9], lin_to_xyz[9], aces_to_lin[9], lin_to_aces[9];
  ColourMats() {
    // CSS Color 4 XYZ(D65) <-> linear sRGB in rational form; ACEScg from the
    // AP1 primaries and ACES white, Bradford-adapted to D65.  Rounded to f32
    // once, on the host.
    const double x2s[9] = {12831.0 / 3959.0, -329.0 / 214.0, -1974.0 / 3959.0,
                           -851781.0 / 878810.0, 1648619.0 / 878810.0, 36519.0 / 878810.0,
                           705.0 / 12673.0, -2585.0 / 12673.0, 705.0 / 667.0};
    const double s2x[9] = {506752.0 / 1228815.0, 87881.0 / 245763.0, 12673.0 / 70218.0,
                           87098.0 / 409605.0, 175762.0 / 245763.0, 12673.0 / 175545.0,
                           7918.0 / 409605.0, 87881.0 / 737289.0, 1001167.0 / 1053270.0};
    double a2s[9], s2a[9];
    acescg_to_linear_srgb(x2s, a2s);
    m3inv(a2s, s2a);
    for (int k = 0; k < 9; ++k) {
      aces_to_lin[k] = (float)a2s[k];
      lin_to_aces[k] = (float)s2a[k];
      xyz_to_lin[k] = (float)x2s[k];
      lin_to_xyz[k] = (float)s2x[k];
    }
  }
};

const ColourMats& mats() {
  static ColourMats m;
  return m;
}

// cst.rs:51-95 in f32 with the reference's loop order (no pivoting).
void pseudo_inverse3(const float m[3][3], float out[3][3]) {
  float work[3][6];
  for (int i = 0; i < 3; ++i) {
    for (int j = 0; j < 6; ++j) work[i][j] = (j == i + 3) ? 1.0f : 0.0f;
    for (int j = 0; j < 3; ++j)
      for (int k = 0; k < 3; ++k) work[i][j] += m[k][i] * m[k][j];
  }
  for (int i = 0; i < 3; ++i) {
    float num = work[i][i];
    for (int j = 0; j < 6; ++j) work[i][j] /= num;
    for (int k = 0; k < 3; ++k) {
      if (k == i) continue;
      num = work[k][i];
      for (int j = 0; j < 6; ++j) work[k][j] -= work[i][j] * num;
    }
  }
  float tmp[3][3];
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) {
      tmp[i][j] = 0.0f;
      for (int k = 0; k < 3; ++k) tmp[i][j] += work[j][k + 3] * m[i][k];
    }
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) out[i][j] = tmp[j][i];
}

}  // namespace

void push_convert(PointwiseProgram& prog, int src, int dst) {
  if (src == dst) return;
  const ColourMats& m = mats();
  bool ok = true;
  switch (src) {  // src -> linear sRGB
    case IMOPS_CS_SRGB: ok = prog.push(PW_EOTF, nullptr, 0); break;
    case IMOPS_CS_LINEAR_SRGB: break;
    case IMOPS_CS_ACESCG: ok = prog.push(PW_MAT3, m.aces_to_lin, 9); break;
    case IMOPS_CS_XYZD65: ok = prog.push(PW_MAT3, m.xyz_to_lin, 9); break;
    default: throw Error{IMOPS_E_INVALID, "unsupported source colour space"};
  }
  switch (dst) {  // linear sRGB -> dst
    case IMOPS_CS_SRGB: ok = ok && prog.push(PW_OETF, nullptr, 0); break;
    case IMOPS_CS_LINEAR_SRGB: break;
    case IMOPS_CS_ACESCG: ok = ok && prog.push(PW_MAT3, m.lin_to_aces, 9); break;
    case IMOPS_CS_XYZD65: ok = ok && prog.push(PW_MAT3, m.lin_to_xyz, 9); break;
    default: throw Error{IMOPS_E_INVALID, "unsupported target colour space"};
  }
  if (!ok) throw Error{IMOPS_E_INVALID, "too many fused pointwise ops"};
}

// cst.rs:12-48
void push_camera_cst(PointwiseProgram& prog, const float* cal, size_t len, int target) {
  if (len != 9) throw Error{IMOPS_E_UNSUPPORTED, "camera_cst: calibration matrix must have 3 rows of 3"};
  const float xyz_d65[3] = {0.95042f, 1.0f, 1.08890f};
  float rgb_d65[3], mult[3] = {1.0f, 1.0f, 1.0f}, xyz2cam[3][3], fwd[3][3];
  for (int i = 0; i < 3; ++i)
    rgb_d65[i] = cal[i * 3] * xyz_d65[0] + cal[i * 3 + 1] * xyz_d65[1] + cal[i * 3 + 2] * xyz_d65[2];
  for (int i = 0; i < 3; ++i)
    if (rgb_d65[i] > 0.0f) mult[i] = 1.0f / rgb_d65[i];
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) xyz2cam[i][j] = cal[i * 3 + j] * mult[i];
  pseudo_inverse3(xyz2cam, fwd);
  if (!prog.push(PW_MAT3, &fwd[0][0], 9)) throw Error{IMOPS_E_INVALID, "too many fused pointwise ops"};
  push_convert(prog, IMOPS_CS_XYZD65, target);
}

void push_exp(PointwiseProgram& prog, float ev) {
  const float v = powf(2.0f, ev);  // exp.rs:5, evaluated on the host like the reference
  const float s[3] = {v, v, v};
  if (!prog.push(PW_SCALE3, s, 3)) throw Error{IMOPS_E_INVALID, "too many fused pointwise ops"};
}

void push_gamma(PointwiseProgram& prog, float gamma) {
  const float inv = 1.0f / std::fmax(gamma, 1e-6f);  // tone_map.rs:39
  if (!prog.push(PW_GAMMA, &inv, 1)) throw Error{IMOPS_E_INVALID, "too many fused pointwise ops"};
}

void push_wb(PointwiseProgram& prog, const float wb[4]) {
  if (!prog.push(PW_SCALE3, wb, 3)) throw Error{IMOPS_E_INVALID, "too many fused pointwise ops"};
}

namespace {
void push_or_throw(PointwiseProgram& prog, int kind, const float* m, int nm) {
  if (!prog.push(kind, m, nm)) throw Error{IMOPS_E_INVALID, "too many fused pointwise ops"};
}
}  // namespace

void push_contrast(PointwiseProgram& prog, float c) { push_or_throw(prog, PW_CONTRAST, &c, 1); }

// highlight_reconstruction.rs:5-14: clip_g is wb_coeffs[1]
void push_highlight(PointwiseProgram& prog, const float wb[4]) { push_or_throw(prog, PW_HIGHLIGHT, &wb[1], 1); }

// lch.rs:5-21: source space -> Oklch (through linear sRGB, as the colour crate
// converts), scale L, C and hue, and back
void push_lch(PointwiseProgram& prog, int color_space, float lc, float cc, float hc) {
  if (color_space == IMOPS_CS_NONE) throw Error{IMOPS_E_STATE, "LCH needs the image's colour space"};
  push_convert(prog, color_space, IMOPS_CS_LINEAR_SRGB);
  push_or_throw(prog, PW_LIN_TO_OKLCH, nullptr, 0);
  const float s[3] = {lc, cc, hc};
  push_or_throw(prog, PW_SCALE3, s, 3);
  push_or_throw(prog, PW_OKLCH_TO_LIN, nullptr, 0);
  push_convert(prog, IMOPS_CS_LINEAR_SRGB, color_space);
}

// tone_map.rs:16-31: the buffer is taken as ACEScg whatever its tag
void push_sigmoid(PointwiseProgram& prog) {
  const ColourMats& m = mats();
  float both[18];
  for (int k = 0; k < 9; ++k) {
    both[k] = m.aces_to_lin[k];
    both[9 + k] = m.lin_to_aces[k];
  }
  push_or_throw(prog, PW_SIGMOID, both, 18);
}

// modules.rs:335-371: 2^ev when |ev| > 1e-6, then K N^2 / (t S) (radiance.rs:19-31,
// K = 12.5) when shutter, f-number and ISO are all known and positive
void push_baseline_exposure(PointwiseProgram& prog, const imops_image* img) {
  const float ev = img->has_baseline ? img->baseline_ev : 0.0f;
  if (std::fabs(ev) > 1e-6f) push_exp(prog, ev);
  if (img->has_shutter && img->has_f_number && img->has_iso && img->shutter > 0.0f && img->f_number > 0.0f &&
      img->iso > 0.0f) {
    const double t = img->shutter, n = img->f_number, iso = img->iso;
    const float g = (float)(12.5 * 1.0 * (n * n) / (t * iso));
    const float s[3] = {g, g, g};
    push_or_throw(prog, PW_SCALE3, s, 3);
  }
}

// apply_vignette_radial_correction (vignette.rs:3-99): big-endian opcode list,
// one pass per FixVignetteRadial (id 3) opcode with >= 56 parameter bytes
void stage_vignette(imops_ctx* ctx, imops_image* img, float strength) {
  require_rgb(img, "Vignette");
  if (!img->has_opcodes) return;
  const std::vector<uint8_t>& op = img->opcode_list3;
  if (op.size() < 4) return;
  auto u32 = [&](size_t o) {
    return ((uint32_t)op[o] << 24) | ((uint32_t)op[o + 1] << 16) | ((uint32_t)op[o + 2] << 8) | (uint32_t)op[o + 3];
  };
  auto f64 = [&](size_t o) {
    uint64_t b = 0;
    for (int i = 0; i < 8; ++i) b = (b << 8) | op[o + i];
    double d;
    std::memcpy(&d, &b, sizeof d);
    return d;
  };
  const size_t count = u32(0);
  size_t off = 4;
  TimedRegion tr(ctx, "vignette");
  for (size_t c = 0; c < count; ++c) {
    if (off + 16 > op.size()) break;
    const uint32_t id = u32(off);
    const size_t psize = u32(off + 12);
    off += 16;
    if (off + psize > op.size()) break;
    const size_t p = off;
    off += psize;
    if (id != 3 || psize < 56) continue;
    VignetteParams v;
    for (int k = 0; k < 5; ++k) v.k[k] = (float)f64(p + 8 * k);
    v.cx = (float)f64(p + 40);
    v.cy = (float)f64(p + 48);
    const float cx = v.cx, cy = v.cy;
    const float d00 = std::sqrt(cx * cx + cy * cy);
    const float d10 = std::sqrt((1.0f - cx) * (1.0f - cx) + cy * cy);
    const float d01 = std::sqrt(cx * cx + (1.0f - cy) * (1.0f - cy));
    const float d11 = std::sqrt((1.0f - cx) * (1.0f - cx) + (1.0f - cy) * (1.0f - cy));
    v.d_max = std::fmax(std::fmax(std::fmax(d00, d10), d01), d11);
    v.strength = strength;
    if (v.d_max > 0.0f)
      IMOPS_HIP_CHECK(vignette_launch(img->rgb.as<float>(), (int)img->width, (int)img->height, v, ctx->stream));
  }
}

void stage_pointwise(imops_ctx* ctx, imops_image* img, const PointwiseProgram& prog) {
  TimedRegion tr(ctx, "pointwise");
  IMOPS_HIP_CHECK(pointwise_launch(img->rgb.as<float>(), img->width * img->height, prog, ctx->stream));
}

// AMaZE launch arguments for a w x h CFA plane with the crop-shifted pattern `cfa`;
// rejects the shapes the reference cannot process.
AmazeArgs amaze_args(const CfaTable& cfa, size_t w, size_t h, const float* in, float* out) {
  if (!cfa.is_bayer()) throw Error{IMOPS_E_UNSUPPORTED, "AMaZE needs a 2x2 Bayer pattern, got " + cfa.name};
  const size_t rw = w % 96, rh = h % 96;
  if (w < 2 || h < 2 || (rw >= 1 && rw <= 15) || (rh >= 1 && rh <= 15))
    throw Error{IMOPS_E_UNSUPPORTED,
                "AMaZE: width and height mod 96 must not be 1..15 (the reference indexes past its tile buffer)"};
  if (w > (1u << 30) / 3 || h > (1u << 20)) throw Error{IMOPS_E_INVALID, "AMaZE: image too large"};
  AmazeArgs a{};
  a.in = in;
  a.out = out;
  a.width = (int)w;
  a.height = (int)h;
  amaze_tiles(a.width, a.height, &a.tiles_x, &a.tiles_y);
  for (int r = 0; r < 2; ++r)
    for (int c = 0; c < 2; ++c) a.cfa[r][c] = cfa.pat[r][c];
  return a;
}

void materialize_raw(imops_ctx* ctx, imops_image* img) {
  if (!img->raw16) return;
  const size_t w = img->width, h = img->height;
  img->raw.reserve(w * h * sizeof(float));
  TimedRegion tr(ctx, "normalize");
  IMOPS_HIP_CHECK(normalize_launch(img->raw16, img->raw16_pitch, img->raw16_x, img->raw16_y, w, h, img->raw16_black,
                                   img->raw16_white, img->raw.as<float>(), ctx->stream));
  img->raw16 = nullptr;
}

void stage_demosaic(imops_ctx* ctx, imops_image* img, int algorithm) {
  if (!img->has_raw) throw Error{IMOPS_E_STATE, "Demosaic: image has no raw CFA data"};
  if (algorithm != IMOPS_DEMOSAIC_AMAZE) materialize_raw(ctx, img);
  if (!img->has_cfa) throw Error{IMOPS_E_STATE, "A CFA must be set to demosaic raw images"};
  const size_t w = img->width, h = img->height;
  const CfaTable cfa = img->cfa.shift(img->crop_x, img->crop_y);  // demosaic.rs:93-98
  img->rgb.reserve(w * h * 3 * sizeof(float));
  if (algorithm == IMOPS_DEMOSAIC_AMAZE) {
    AmazeArgs a = amaze_args(cfa, w, h, img->raw16 ? nullptr : img->raw.as<float>(), img->rgb.as<float>());
    if (img->raw16) {  // fused ingest: the tile fill reads and normalises the bound u16 frame
      a.in16 = img->raw16;
      a.in16_pitch = img->raw16_pitch;
      a.in16_x = img->raw16_x;
      a.in16_y = img->raw16_y;
      a.black = img->raw16_black;
      a.range = img->raw16_white - img->raw16_black;  // as normalize_launch forms it
    }
    const size_t ntiles = (size_t)a.tiles_x * a.tiles_y;
    int grid = ctx->num_cus > 0 ? ctx->num_cus : 256;  // one 144 KiB-LDS workgroup per CU
    if ((size_t)grid > ntiles) grid = (int)ntiles;
    ctx->scr[SCR_AMAZE].reserve((size_t)grid * kAmazeSlotFloats * sizeof(float));
    a.scratch = ctx->scr[SCR_AMAZE].as<float>();
    TimedRegion tr(ctx, "amaze");
    IMOPS_HIP_CHECK(amaze_launch(a, grid, ctx->stream));
  } else if (algorithm == IMOPS_DEMOSAIC_MARKESTEIJN) {
    CfaPattern p{};
    p.w = cfa.width;
    p.h = cfa.height;
    for (int r = 0; r < 12; ++r)
      for (int c = 0; c < 12; ++c) p.base[r][c] = cfa.pat[r][c];
    TimedRegion tr(ctx, "markesteijn");
    IMOPS_HIP_CHECK(markesteijn_launch(img->raw.as<float>(), (int)w, (int)h, p, img->rgb.as<float>(), ctx->stream));
  } else if (algorithm == IMOPS_DEMOSAIC_MARKESTEIJN3 || algorithm == IMOPS_DEMOSAIC_MARKESTEIJN3_1PASS) {
    for (int r = 0; r < 2; ++r)
      for (int c = 0; c < 2; ++c)
        if (cfa.pat[r][c] > 2) throw Error{IMOPS_E_UNSUPPORTED, "markesteijn3 needs an RGB pattern, got " + cfa.name};
    if (w > (1u << 30) / 3 || h > (1u << 20)) throw Error{IMOPS_E_INVALID, "markesteijn3: image too large"};
    CfaPattern p{};
    p.w = cfa.width;
    p.h = cfa.height;
    for (int r = 0; r < 12; ++r)
      for (int c = 0; c < 12; ++c) p.base[r][c] = cfa.pat[r][c];
    const int passes = algorithm == IMOPS_DEMOSAIC_MARKESTEIJN3 ? 3 : 1;
    ctx->scr[SCR_DEMOSAIC].reserve(markesteijn3_scratch_bytes((int)w, (int)h, passes));
    TimedRegion tr(ctx, "markesteijn3");
    IMOPS_HIP_CHECK(markesteijn3_launch(img->raw.as<float>(), (int)w, (int)h, p, passes, ctx->scr[SCR_DEMOSAIC].ptr,
                                        img->rgb.as<float>(), ctx->stream));
  } else if (algorithm >= IMOPS_DEMOSAIC_FAST && algorithm <= IMOPS_DEMOSAIC_SUPERSUPERFAST) {
    const int f = algorithm == IMOPS_DEMOSAIC_FAST ? 2 : (algorithm == IMOPS_DEMOSAIC_SUPERFAST ? 4 : 8);
    const size_t nw = w / f, nh = h / f;
    if (nw == 0 || nh == 0) throw Error{IMOPS_E_UNSUPPORTED, "Demosaic preview: frame smaller than its block"};
    int cell[4];
    for (int k = 0; k < 4; ++k) {
      cell[k] = cfa.pat[k / 2][k % 2];
      if (cell[k] > 2) throw Error{IMOPS_E_UNSUPPORTED, "Demosaic preview needs an RGB pattern, got " + cfa.name};
    }
    TimedRegion tr(ctx, "preview_demosaic");
    IMOPS_HIP_CHECK(preview_demosaic_launch(img->raw.as<float>(), (int)w, (int)h, f, cell, img->rgb.as<float>(),
                                            ctx->stream));
    img->width = nw;
    img->height = nh;
    img->has_raw = false;  // the raw plane keeps the old size
    img->raw16 = nullptr;
  } else {
    throw Error{IMOPS_E_INVALID, "unknown demosaic algorithm"};
  }
  img->has_rgb = true;
  img->color_space = IMOPS_CS_NONE;  // modules.rs:666
}

}  // namespace imops

using namespace imops;

// ---------------------------------------------------------------- C ABI
namespace imops {
template <class F>
int run_program(imops_ctx* ctx, imops_image* img, const char* name, F&& build) {
  return guarded([&] {
    if (!ctx || !img) throw Error{IMOPS_E_INVALID, std::string(name) + ": null argument"};
    require_rgb(img, name);
    PointwiseProgram p;
    build(p);
    stage_pointwise(ctx, img, p);
  });
}
}  // namespace imops

namespace imops {
// Encoded readback: the kernel writes dst directly when it is device memory, else a
// device staging buffer that is copied out (the encoded bytes only cross the link).
template <class T, class Launch>
void get_encoded(imops_ctx* ctx, const imops_image* img, T* dst, int dst_on_device, int orientation, size_t count,
                 const char* name, Launch launch) {
  if (!ctx || !img || !dst) throw Error{IMOPS_E_INVALID, std::string(name) + ": null argument"};
  require_rgb(img, name);
  if (orientation < 0 || orientation > 3)
    throw Error{IMOPS_E_INVALID, std::string(name) + ": orientation must be 0 (none), 1 (90), 2 (180) or 3 (270)"};
  if (count != img->width * img->height * 3)
    throw Error{IMOPS_E_INVALID, std::string(name) + ": count must be width*height*3"};
  T* out = dst;
  if (!dst_on_device) {
    ctx->scr[SCR_STAGING].reserve(count * sizeof(T));
    out = ctx->scr[SCR_STAGING].as<T>();
  }
  IMOPS_HIP_CHECK(launch(img->rgb.as<float>(), (int)img->width, (int)img->height, orientation, out, ctx->stream));
  if (!dst_on_device)
    IMOPS_HIP_CHECK(hipMemcpyAsync(dst, out, count * sizeof(T), hipMemcpyDeviceToHost, ctx->stream));
  IMOPS_HIP_CHECK(hipStreamSynchronize(ctx->stream));
}
}  // namespace imops

extern "C" {

const char* imops_last_error(void) { return t_last_error.c_str(); }

int imops_ctx_create(int device, imops_ctx** out) {
  return guarded([&] {
    if (!out) throw Error{IMOPS_E_INVALID, "imops_ctx_create: out is null"};
    *out = nullptr;
    int n = 0;
    IMOPS_HIP_CHECK(hipGetDeviceCount(&n));
    if (device < 0 || device >= n) throw Error{IMOPS_E_INVALID, "imops_ctx_create: no such device"};
    IMOPS_HIP_CHECK(hipSetDevice(device));
    auto* c = new imops_ctx();
    c->device = device;
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, device) == hipSuccess) c->num_cus = prop.multiProcessorCount;
    // the static scratch of the allocation-free entry point, looked up once per device
    hipError_t e = amaze_noalloc_scratch(device, &c->amaze_noalloc);
    if (e != hipSuccess) {
      delete c;
      throw Error{IMOPS_E_HIP, std::string("hipGetSymbolAddress(AMaZE scratch): ") + hipGetErrorString(e)};
    }
    e = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking);
    if (e != hipSuccess) {
      delete c;
      throw Error{IMOPS_E_HIP, std::string("hipStreamCreate: ") + hipGetErrorString(e)};
    }
    *out = c;
  });
}

int imops_ctx_set_autotune(imops_ctx* ctx, int enabled) {
  return guarded([&] {
    if (!ctx) throw Error{IMOPS_E_INVALID, "null context"};
    ctx->autotune = enabled != 0;
  });
}

int imops_ctx_tuning(imops_ctx* ctx, char* buf, size_t len, size_t* needed) {
  return guarded([&] {
    if (!ctx) throw Error{IMOPS_E_INVALID, "null context"};
    std::string out;
    for (const auto& kv : ctx->tuned) out += kv.first + "=" + std::to_string(kv.second) + "\n";
    if (needed) *needed = out.size() + 1;
    if (buf && len) {
      const size_t m = std::min(len - 1, out.size());
      std::memcpy(buf, out.data(), m);
      buf[m] = '\0';
    }
  });
}

int imops_ctx_set_timing(imops_ctx* ctx, int enabled) {
  return guarded([&] {
    if (!ctx) throw Error{IMOPS_E_INVALID, "null context"};
    IMOPS_HIP_CHECK(hipStreamSynchronize(ctx->stream));
    ctx->timer.clear();
    ctx->timer.enabled = enabled != 0;
  });
}

int imops_ctx_timing(imops_ctx* ctx, const char* region, double* total_ms, int* count) {
  return guarded([&] {
    if (!ctx || !region || !total_ms || !count) throw Error{IMOPS_E_INVALID, "ctx_timing: null argument"};
    IMOPS_HIP_CHECK(hipStreamSynchronize(ctx->stream));
    double total = 0.0;
    int n = 0;
    for (auto& r : ctx->timer.regions)
      if (r.name == region)
        for (auto& e : r.events) {
          float ms = 0.0f;
          IMOPS_HIP_CHECK(hipEventElapsedTime(&ms, e.first, e.second));
          total += ms;
          ++n;
        }
    for (auto& c : ctx->timer.counts)
      if (c.name == region) {
        total += (double)*c.host;
        ++n;
      }
    *total_ms = total;
    *count = n;
  });
}

int imops_ctx_destroy(imops_ctx* ctx) {
  return guarded([&] {
    if (!ctx) return;
    (void)hipSetDevice(ctx->device);
    (void)hipStreamSynchronize(ctx->stream);
    for (imops_image* img : ctx->images) {
      img->raw.release();
      img->rgb.release();
      img->tmp.release();
      img->raw16 = nullptr;
      img->has_raw = img->has_rgb = false;
      img->ctx = nullptr;
    }
    ctx->images.clear();
    ctx->timer.release();
    for (auto& b : ctx->scr) b.release();
    ctx->bm3d_frags.release();
    ctx->film_cache.reset();
    ctx->fit_cache.reset();
    (void)hipStreamDestroy(ctx->stream);
    delete ctx;
  });
}

int imops_ctx_synchronize(imops_ctx* ctx) {
  return guarded([&] {
    if (!ctx) throw Error{IMOPS_E_INVALID, "null context"};
    IMOPS_HIP_CHECK(hipStreamSynchronize(ctx->stream));
  });
}

int imops_debug_fail_allocations(int count) {
  return guarded([&] {
    if (count < 0) throw Error{IMOPS_E_INVALID, "imops_debug_fail_allocations: count < 0"};
    g_injected_alloc_failures.store(count);
  });
}

int imops_measure_peaks(imops_ctx* ctx, double* out, int n) {
  return guarded([&] {
    if (!ctx || !out || n < 6) throw Error{IMOPS_E_INVALID, "imops_measure_peaks: null argument or n < 6"};
    IMOPS_HIP_CHECK(hipSetDevice(ctx->device));
    const size_t bytes = (size_t)1 << 30;
    struct Held {  // released on every exit
      imops::DevBuf b;
      ~Held() { b.release(); }
    } a, b;
    a.b.reserve(bytes);
    b.b.reserve(bytes);
    IMOPS_HIP_CHECK(hipMemsetAsync(a.b.ptr, 0, bytes, ctx->stream));
    IMOPS_HIP_CHECK(hipMemsetAsync(b.b.ptr, 0, bytes, ctx->stream));
    imops::MeasuredPeaks p;
    const hipError_t e = imops::measure_peaks(ctx->stream, a.b.ptr, b.b.ptr, bytes, ctx->device, &p);
    (void)hipStreamSynchronize(ctx->stream);
    IMOPS_HIP_CHECK(e);
    out[0] = p.copy_GBps;
    out[1] = p.valu_per_clk_per_simd;
    out[2] = p.valu_nonfused_TFLOPs;
    out[3] = p.mfma_f16_TFLOPs;
    out[4] = p.clock_GHz;
    out[5] = p.compute_units;
  });
}

void* imops_ctx_stream(imops_ctx* ctx) { return ctx ? (void*)ctx->stream : nullptr; }

int imops_image_create(imops_ctx* ctx, imops_image** out) {
  return guarded([&] {
    if (!ctx || !out) throw Error{IMOPS_E_INVALID, "imops_image_create: null argument"};
    auto* img = new imops_image();
    img->ctx = ctx;
    ctx->images.insert(img);
    *out = img;
  });
}

int imops_image_free(imops_image* img) {
  return guarded([&] {
    if (!img) return;
    if (img->ctx) {  // else its context was destroyed first and released the buffers
      (void)hipSetDevice(img->ctx->device);
      (void)hipStreamSynchronize(img->ctx->stream);
      img->raw.release();
      img->rgb.release();
      img->tmp.release();
      img->ctx->images.erase(img);
    }
    delete img;
  });
}

int imops_image_set_raw_u16(imops_ctx* ctx, imops_image* img, const uint16_t* src, int src_on_device,
                            size_t full_w, size_t full_h, size_t cx, size_t cy, size_t cw, size_t ch, float black,
                            float white, const char* cfa_pattern) {
  return guarded([&] {
    if (!ctx || !img || !src || !cfa_pattern) throw Error{IMOPS_E_INVALID, "set_raw_u16: null argument"};
    if (!img->ctx) throw Error{IMOPS_E_STATE, "set_raw_u16: the image's context was destroyed"};
    if (cw == 0 || ch == 0 || cx + cw > full_w || cy + ch > full_h)
      throw Error{IMOPS_E_INVALID, "set_raw_u16: crop window outside the frame"};
    if (!(white > black)) throw Error{IMOPS_E_INVALID, "set_raw_u16: white level must exceed black level"};
    CfaTable cfa;
    if (!cfa.parse(cfa_pattern)) throw Error{IMOPS_E_INVALID, std::string("Unknown CFA pattern \"") + cfa_pattern + "\""};
    const uint16_t* dsrc = src;
    if (!src_on_device) {
      ctx->scr[SCR_STAGING].reserve(full_w * full_h * sizeof(uint16_t));
      IMOPS_HIP_CHECK(hipMemcpyAsync(ctx->scr[SCR_STAGING].ptr, src, full_w * full_h * sizeof(uint16_t),
                                     hipMemcpyHostToDevice, ctx->stream));
      dsrc = ctx->scr[SCR_STAGING].as<uint16_t>();
    }
    img->raw.reserve(cw * ch * sizeof(float));
    img->raw16 = nullptr;
    TimedRegion tr(ctx, "normalize");
    IMOPS_HIP_CHECK(normalize_launch(dsrc, full_w, cx, cy, cw, ch, black, white, img->raw.as<float>(), ctx->stream));
    img->width = cw;
    img->height = ch;
    img->cfa = cfa;
    img->has_cfa = true;
    img->crop_x = cx;
    img->crop_y = cy;
    img->has_raw = true;
    img->has_rgb = false;
    img->color_space = IMOPS_CS_NONE;
    if (!src_on_device) IMOPS_HIP_CHECK(hipStreamSynchronize(ctx->stream));  // staging buffer reuse
  });
}

int imops_image_bind_raw_u16(imops_ctx* ctx, imops_image* img, const uint16_t* dev_src, size_t full_w, size_t full_h,
                             size_t cx, size_t cy, size_t cw, size_t ch, float black, float white,
                             const char* cfa_pattern) {
  return guarded([&] {
    if (!ctx || !img || !dev_src || !cfa_pattern) throw Error{IMOPS_E_INVALID, "bind_raw_u16: null argument"};
    if (!img->ctx) throw Error{IMOPS_E_STATE, "bind_raw_u16: the image's context was destroyed"};
    if (cw == 0 || ch == 0 || cx + cw > full_w || cy + ch > full_h)
      throw Error{IMOPS_E_INVALID, "bind_raw_u16: crop window outside the frame"};
    if (!(white > black)) throw Error{IMOPS_E_INVALID, "bind_raw_u16: white level must exceed black level"};
    CfaTable cfa;
    if (!cfa.parse(cfa_pattern)) throw Error{IMOPS_E_INVALID, std::string("Unknown CFA pattern \"") + cfa_pattern + "\""};
    hipPointerAttribute_t attr;
    if (hipPointerGetAttributes(&attr, dev_src) != hipSuccess || attr.type != hipMemoryTypeDevice) {
      (void)hipGetLastError();
      throw Error{IMOPS_E_INVALID, "bind_raw_u16: the frame must be device memory"};
    }
    img->raw16 = dev_src;
    img->raw16_pitch = full_w;
    img->raw16_x = cx;
    img->raw16_y = cy;
    img->raw16_black = black;
    img->raw16_white = white;
    img->width = cw;
    img->height = ch;
    img->cfa = cfa;
    img->has_cfa = true;
    img->crop_x = cx;
    img->crop_y = cy;
    img->has_raw = true;
    img->has_rgb = false;
    img->color_space = IMOPS_CS_NONE;
  });
}

int imops_crop_bayer_center(imops_ctx* ctx, imops_image* img, size_t crop_factor) {
  return guarded([&] {
    if (!ctx || !img) throw Error{IMOPS_E_INVALID, "crop_bayer_center: null argument"};
    if (!img->has_raw) throw Error{IMOPS_E_STATE, "crop_bayer_center: image has no raw CFA data"};
    if (crop_factor <= 1) return;  // the image itself (extern_pipeline.rs:118-120)
    materialize_raw(ctx, img);
    const size_t w = img->width, h = img->height;
    const size_t nw = (w / crop_factor) & ~(size_t)1, nh = (h / crop_factor) & ~(size_t)1;
    DevBuf out;
    out.reserve(std::max<size_t>(nw * nh, 1) * sizeof(float));
    if (nw * nh > 0)
      IMOPS_HIP_CHECK(crop_bayer_center_launch(img->raw.as<float>(), w, h, crop_factor, nw, nh, out.as<float>(),
                                               ctx->stream));
    IMOPS_HIP_CHECK(hipStreamSynchronize(ctx->stream));
    std::swap(img->raw, out);
    out.release();
    img->width = nw;  // the metadata (CFA, crop origin, white balance, ...) is kept
    img->height = nh;
    img->has_rgb = false;
  });
}

// interpolate_matrices (extern_pipeline.rs:457-495) in the reference's f32 order.
static void interpolate_matrices(const float* m1, const float* m2, const float wb[4], float out[9]) {
  const float xyz_a[3] = {1.09850f, 1.0f, 0.35585f};
  const float xyz_d65[3] = {0.95047f, 1.0f, 1.08883f};
  float s1[3], s2[3];
  for (int i = 0; i < 3; ++i) {
    s1[i] = m1[i * 3] * xyz_a[0] + m1[i * 3 + 1] * xyz_a[1] + m1[i * 3 + 2] * xyz_a[2];
    s2[i] = m2[i * 3] * xyz_d65[0] + m2[i * 3 + 1] * xyz_d65[1] + m2[i * 3 + 2] * xyz_d65[2];
  }
  const float ratio1 = s1[2] != 0.0f ? s1[0] / s1[2] : 1.0f;
  const float ratio2 = s2[2] != 0.0f ? s2[0] / s2[2] : 1.0f;
  const float ratio_s = (wb[0] > 0.0f && wb[2] > 0.0f) ? wb[2] / wb[0] : ratio2;
  float w = 0.0f;
  if (ratio1 != ratio2 && ratio1 > 0.0f && ratio2 > 0.0f && ratio_s > 0.0f) {
    const float v = (std::log(ratio_s) - std::log(ratio2)) / (std::log(ratio1) - std::log(ratio2));
    w = v < 0.0f ? 0.0f : (v > 1.0f ? 1.0f : v);  // f32::clamp (NaN passes through)
  }
  for (int i = 0; i < 9; ++i) out[i] = w * m1[i] + (1.0f - w) * m2[i];
}

int imops_dng_calibration_matrix(const float* color_matrix_a, const float* color_matrix_d65,
                                 const float* color_matrix_other, const float* wb_coeffs, float out[9]) {
  return guarded([&] {
    if (!out) throw Error{IMOPS_E_INVALID, "dng_calibration_matrix: null output"};
    float wb[4] = {0.0f, 0.0f, 0.0f, 0.0f};
    for (int i = 0; i < 4 && wb_coeffs; ++i) wb[i] = std::isnan(wb_coeffs[i]) ? 0.0f : wb_coeffs[i];
    const float* pick = nullptr;
    if (color_matrix_a && color_matrix_d65) {
      interpolate_matrices(color_matrix_a, color_matrix_d65, wb, out);
      return;
    }
    if (color_matrix_a) pick = color_matrix_a;
    else if (color_matrix_d65) pick = color_matrix_d65;
    else if (color_matrix_other) pick = color_matrix_other;
    if (pick) {
      for (int i = 0; i < 9; ++i) out[i] = pick[i];
    } else {
      for (int i = 0; i < 9; ++i) out[i] = (i % 4 == 0) ? 1.0f : 0.0f;
    }
  });
}

int imops_image_set_raw_f32(imops_ctx* ctx, imops_image* img, const float* src, int src_on_device, size_t w,
                            size_t h, const char* cfa_pattern, size_t crop_x, size_t crop_y) {
  return guarded([&] {
    if (!ctx || !img || !src || !cfa_pattern) throw Error{IMOPS_E_INVALID, "set_raw_f32: null argument"};
    if (!img->ctx) throw Error{IMOPS_E_STATE, "set_raw_f32: the image's context was destroyed"};
    if (w == 0 || h == 0) throw Error{IMOPS_E_INVALID, "set_raw_f32: empty image"};
    CfaTable cfa;
    if (!cfa.parse(cfa_pattern)) throw Error{IMOPS_E_INVALID, std::string("Unknown CFA pattern \"") + cfa_pattern + "\""};
    img->raw.reserve(w * h * sizeof(float));
    img->raw16 = nullptr;
    IMOPS_HIP_CHECK(hipMemcpyAsync(img->raw.ptr, src, w * h * sizeof(float),
                                   src_on_device ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice, ctx->stream));
    IMOPS_HIP_CHECK(hipStreamSynchronize(ctx->stream));
    img->width = w;
    img->height = h;
    img->cfa = cfa;
    img->has_cfa = true;
    img->crop_x = crop_x;
    img->crop_y = crop_y;
    img->has_raw = true;
    img->has_rgb = false;
    img->color_space = IMOPS_CS_NONE;
  });
}

int imops_image_set_rgb_f32(imops_ctx* ctx, imops_image* img, const float* src, int src_on_device, size_t w,
                            size_t h, int cs) {
  return guarded([&] {
    if (!ctx || !img || !src) throw Error{IMOPS_E_INVALID, "set_rgb_f32: null argument"};
    if (!img->ctx) throw Error{IMOPS_E_STATE, "set_rgb_f32: the image's context was destroyed"};
    if (w == 0 || h == 0) throw Error{IMOPS_E_INVALID, "set_rgb_f32: empty image"};
    if (cs < IMOPS_CS_NONE || cs > IMOPS_CS_XYZD65) throw Error{IMOPS_E_INVALID, "set_rgb_f32: unknown colour space"};
    img->rgb.reserve(w * h * 3 * sizeof(float));
    IMOPS_HIP_CHECK(hipMemcpyAsync(img->rgb.ptr, src, w * h * 3 * sizeof(float),
                                   src_on_device ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice, ctx->stream));
    IMOPS_HIP_CHECK(hipStreamSynchronize(ctx->stream));
    img->width = w;
    img->height = h;
    img->has_rgb = true;
    img->has_raw = false;
    img->raw16 = nullptr;
    img->color_space = cs;
  });
}

int imops_image_get_rgb_f32(imops_ctx* ctx, const imops_image* img, float* dst, int dst_on_device, size_t count) {
  return guarded([&] {
    if (!ctx || !img || !dst) throw Error{IMOPS_E_INVALID, "get_rgb_f32: null argument"};
    require_rgb(img, "get_rgb_f32");
    if (count != img->width * img->height * 3)
      throw Error{IMOPS_E_INVALID, "get_rgb_f32: count must be width*height*3"};
    IMOPS_HIP_CHECK(hipMemcpyAsync(dst, img->rgb.ptr, count * sizeof(float),
                                   dst_on_device ? hipMemcpyDeviceToDevice : hipMemcpyDeviceToHost, ctx->stream));
    IMOPS_HIP_CHECK(hipStreamSynchronize(ctx->stream));
  });
}


int imops_image_get_rgb_u8(imops_ctx* ctx, const imops_image* img, uint8_t* dst, int dst_on_device, int orientation,
                           size_t count) {
  return guarded(
      [&] { get_encoded(ctx, img, dst, dst_on_device, orientation, count, "get_rgb_u8", encode_u8_launch); });
}

int imops_image_get_rgb_f16(imops_ctx* ctx, const imops_image* img, uint16_t* dst, int dst_on_device,
                            int orientation, size_t count) {
  return guarded(
      [&] { get_encoded(ctx, img, dst, dst_on_device, orientation, count, "get_rgb_f16", encode_f16_launch); });
}

int imops_image_get_info(const imops_image* img, size_t* w, size_t* h, int* cs, int* has_rgb) {
  return guarded([&] {
    if (!img) throw Error{IMOPS_E_INVALID, "get_info: null image"};
    if (w) *w = img->width;
    if (h) *h = img->height;
    if (cs) *cs = img->color_space;
    if (has_rgb) *has_rgb = img->has_rgb ? 1 : 0;
  });
}

int imops_image_rgb_device_ptr(const imops_image* img, float** out) {
  return guarded([&] {
    if (!img || !out) throw Error{IMOPS_E_INVALID, "rgb_device_ptr: null argument"};
    require_rgb(img, "rgb_device_ptr");
    *out = img->rgb.as<float>();
  });
}

int imops_image_set_wb_coeffs(imops_image* img, const float wb[4]) {
  return guarded([&] {
    if (!img || !wb) throw Error{IMOPS_E_INVALID, "set_wb_coeffs: null argument"};
    for (int i = 0; i < 4; ++i) img->wb[i] = wb[i];
    img->has_wb = true;
  });
}

int imops_image_set_calibration_matrix(imops_image* img, const float* m, size_t len) {
  return guarded([&] {
    if (!img || !m || len == 0 || len % 3 != 0) throw Error{IMOPS_E_INVALID, "set_calibration_matrix: bad matrix"};
    img->calib.assign(m, m + len);
  });
}

int imops_image_set_shutter_seconds(imops_image* img, float s) {
  return guarded([&] {
    if (!img) throw Error{IMOPS_E_INVALID, "set_shutter_seconds: null image"};
    img->shutter = s;
    img->has_shutter = true;
  });
}

int imops_image_set_exposure_metadata(imops_image* img, float baseline_ev, float f_number, float iso) {
  return guarded([&] {
    if (!img) throw Error{IMOPS_E_INVALID, "set_exposure_metadata: null image"};
    img->baseline_ev = baseline_ev;
    img->f_number = f_number;
    img->iso = iso;
    img->has_baseline = img->has_f_number = img->has_iso = true;
  });
}

int imops_image_set_opcode_list3(imops_image* img, const uint8_t* bytes, size_t len) {
  return guarded([&] {
    if (!img || (!bytes && len)) throw Error{IMOPS_E_INVALID, "set_opcode_list3: null argument"};
    img->opcode_list3.assign(bytes, bytes + len);
    img->has_opcodes = true;
  });
}

int imops_contrast(imops_ctx* ctx, imops_image* img, float c) {
  return run_program(ctx, img, "Contrast", [&](PointwiseProgram& p) { push_contrast(p, c); });
}

int imops_highlight_reconstruction(imops_ctx* ctx, imops_image* img) {
  return run_program(ctx, img, "HighlightReconstruction", [&](PointwiseProgram& p) {
    if (!img->has_wb) throw Error{IMOPS_E_STATE, "wb_coeffs needs to be set to perform highlight_reconstruction"};
    push_highlight(p, img->wb);
  });
}

int imops_lch(imops_ctx* ctx, imops_image* img, float lc, float cc, float hc) {
  return run_program(ctx, img, "LCH", [&](PointwiseProgram& p) { push_lch(p, img->color_space, lc, cc, hc); });
}

int imops_sigmoid_tone_map(imops_ctx* ctx, imops_image* img) {
  return run_program(ctx, img, "SigmoidToneMap", [&](PointwiseProgram& p) { push_sigmoid(p); });
}

int imops_baseline_exposure(imops_ctx* ctx, imops_image* img) {
  return run_program(ctx, img, "BaselineExposureCompensation",
                     [&](PointwiseProgram& p) { push_baseline_exposure(p, img); });
}

int imops_vignette(imops_ctx* ctx, imops_image* img, float strength) {
  return guarded([&] {
    if (!ctx || !img) throw Error{IMOPS_E_INVALID, "Vignette: null argument"};
    stage_vignette(ctx, img, strength);
  });
}

int imops_demosaic(imops_ctx* ctx, imops_image* img, int algorithm) {
  return guarded([&] {
    if (!ctx || !img) throw Error{IMOPS_E_INVALID, "demosaic: null argument"};
    stage_demosaic(ctx, img, algorithm);
  });
}

int imops_demosaic_amaze_noalloc(imops_ctx* ctx, const float* cfa_plane, float* rgb_out, size_t width,
                                 size_t height, const char* cfa_pattern, size_t crop_x, size_t crop_y) {
  return guarded([&] {
    if (!ctx || !cfa_plane || !rgb_out || !cfa_pattern)
      throw Error{IMOPS_E_INVALID, "demosaic_amaze_noalloc: null argument"};
    if (!ctx->amaze_noalloc) throw Error{IMOPS_E_STATE, "demosaic_amaze_noalloc: no static scratch on this device"};
    CfaTable cfa;
    if (!cfa.parse(cfa_pattern)) throw Error{IMOPS_E_INVALID, std::string("Unknown CFA pattern \"") + cfa_pattern + "\""};
    AmazeArgs a = amaze_args(cfa.shift(crop_x, crop_y), width, height, cfa_plane, rgb_out);
    // the scratch holds kAmazeNoallocSlots tiles in flight whatever the image size, so no
    // image is too large for it; the grid never exceeds the slots it has
    const size_t ntiles = (size_t)a.tiles_x * a.tiles_y;
    int grid = kAmazeNoallocSlots;
    if (ctx->num_cus > 0 && ctx->num_cus < grid) grid = ctx->num_cus;
    if ((size_t)grid > ntiles) grid = (int)ntiles;
    if ((size_t)grid * kAmazeSlotFloats > (size_t)kAmazeNoallocSlots * kAmazeSlotFloats)
      throw Error{IMOPS_E_INVALID, "demosaic_amaze_noalloc: grid exceeds the static scratch"};
    a.scratch = ctx->amaze_noalloc;
    IMOPS_HIP_CHECK(amaze_launch(a, grid, ctx->stream));
  });
}

int imops_cfa_coeffs(imops_ctx* ctx, imops_image* img, const float* wb) {
  return guarded([&] {
    if (!ctx || !img) throw Error{IMOPS_E_INVALID, "cfa_coeffs: null argument"};
    require_rgb(img, "CFACoeffs");
    if (!wb && !img->has_wb) throw Error{IMOPS_E_STATE, "wb_coeffs needs to be set to perform highlight_reconstruction"};
    PointwiseProgram p;
    push_wb(p, wb ? wb : img->wb);
    stage_pointwise(ctx, img, p);
  });
}

int imops_exp(imops_ctx* ctx, imops_image* img, float ev) {
  return guarded([&] {
    if (!ctx || !img) throw Error{IMOPS_E_INVALID, "exp: null argument"};
    require_rgb(img, "Exp");
    PointwiseProgram p;
    push_exp(p, ev);
    stage_pointwise(ctx, img, p);
  });
}

int imops_gamma(imops_ctx* ctx, imops_image* img, float gamma) {
  return guarded([&] {
    if (!ctx || !img) throw Error{IMOPS_E_INVALID, "gamma: null argument"};
    require_rgb(img, "Gamma");
    PointwiseProgram p;
    push_gamma(p, gamma);
    stage_pointwise(ctx, img, p);
  });
}

int imops_cst(imops_ctx* ctx, imops_image* img, int target) {
  return guarded([&] {
    if (!ctx || !img) throw Error{IMOPS_E_INVALID, "cst: null argument"};
    require_rgb(img, "CST");
    if (target < IMOPS_CS_SRGB || target > IMOPS_CS_XYZD65) throw Error{IMOPS_E_INVALID, "cst: unknown target"};
    PointwiseProgram p;
    if (img->color_space == IMOPS_CS_NONE) {
      if (img->calib.empty())
        throw Error{IMOPS_E_STATE, "A calibration matrix must be set to perform a camera cst transform"};
      push_camera_cst(p, img->calib.data(), img->calib.size(), target);
    } else {
      push_convert(p, img->color_space, target);
    }
    stage_pointwise(ctx, img, p);
    img->color_space = target;
  });
}

int imops_camera_cst(imops_ctx* ctx, imops_image* img, int target, const float* m, size_t len) {
  return guarded([&] {
    if (!ctx || !img || !m) throw Error{IMOPS_E_INVALID, "camera_cst: null argument"};
    require_rgb(img, "camera_cst");
    if (target < IMOPS_CS_SRGB || target > IMOPS_CS_XYZD65) throw Error{IMOPS_E_INVALID, "camera_cst: unknown target"};
    PointwiseProgram p;
    push_camera_cst(p, m, len, target);
    stage_pointwise(ctx, img, p);
    img->color_space = target;
  });
}

}  // extern "C"
