// Host side of the BM3D stage: parameters (Bm3dParams::from_intensity,
// pichromatic/src/bm3d.rs:22-32), the f32 DCT/IDCT/Kaiser tables
// (DctTables::new, bm3d.rs:417-461) and the Gaussian taps of the detail blend
// (bm3d.rs:648-660) computed with the host libm exactly as the reference
// computes them, and the launch sequence of bm3d() (bm3d.rs:34-61).
#include <algorithm>
#include <cmath>
#include <string>
#include <vector>

#include "bm3d.hpp"
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

// Stage-list fields: `intensity` (the reference's only knob) plus optional
// explicit `sigma` (normalised to 0..1), `lambda`, `stride`, `search_window`
// and `max_group`.  With sigma set and no intensity, the intensity that maps
// to that sigma (sigma*255/80) drives the match threshold and detail blend.
Bm3dStageParams bm3d_stage_params_from_table(double intensity, double sigma_norm, double lambda, int stride,
                                             int window, int max_group) {
  float i = (float)intensity;
  if (sigma_norm >= 0.0 && intensity == 0.0) i = (float)(sigma_norm * 255.0 / 80.0);
  Bm3dStageParams p = Bm3dStageParams::from_intensity(i);
  if (sigma_norm >= 0.0) p.sigma = (float)(sigma_norm * 255.0);
  if (lambda >= 0.0) p.lambda = (float)lambda;
  p.stride = stride;
  p.window = window;
  p.max_group = max_group;
  if (stride < 1 || window < 1 || (window & 1) == 0 || window > kBm3dMaxWindow || max_group < 1 ||
      max_group > kBm3dMaxGroup)
    throw Error{IMOPS_E_INVALID, "Invalid BM3D config: stride >= 1, odd search_window <= 41, max_group 1..16"};
  return p;
}

namespace {

struct HostTables {
  float dct[64], idct[64], kaiser[64];
  HostTables() {
    const float pi = 3.14159265358979323846f;  // std::f32::consts::PI
    for (int k = 0; k < 8; ++k)
      for (int n = 0; n < 8; ++n) {
        const float c = (float)k * pi / 8.0f;
        const float val = cosf(((float)n + 0.5f) * c);
        dct[k * 8 + n] = val * (k == 0 ? 0.35355339f : 0.5f);
      }
    for (int n = 0; n < 8; ++n)
      for (int k = 0; k < 8; ++k) {
        const float theta = (pi / 8.0f) * ((float)n + 0.5f) * (float)k;
        idct[n * 8 + k] = (k == 0 ? 0.35355339f : 0.5f) * cosf(theta);
      }
    auto i0 = [](float x) {
      float sum = 1.0f, term = 1.0f;
      for (int k = 1; k <= 10; ++k) {
        const float t = x / 2.0f / (float)k;
        term *= t * t;
        sum += term;
      }
      return sum;
    };
    const float beta = 2.0f, i0b = i0(beta);
    for (int y = 0; y < 8; ++y)
      for (int x = 0; x < 8; ++x) {
        const float ux = (2.0f * (float)x / 7.0f) - 1.0f, uy = (2.0f * (float)y / 7.0f) - 1.0f;
        kaiser[y * 8 + x] = (i0(beta * sqrtf(1.0f - ux * ux)) / i0b) * (i0(beta * sqrtf(1.0f - uy * uy)) / i0b);
        // the filters' 32-bit fixed-point path for the weights relies on this
        if (!(kaiser[y * 8 + x] >= 0.0f && kaiser[y * 8 + x] <= 1.0f)) throw Error{IMOPS_E_STATE, "internal: Kaiser table outside [0, 1]"};
      }
  }
};

Bm3dBlurTaps blur_taps(float sigma) {
  Bm3dBlurTaps t{};
  t.radius = (int)ceilf(3.0f * sigma);
  t.len = 2 * t.radius + 1;
  const float two_s2 = 2.0f * sigma * sigma;
  float sum = 0.0f;
  for (int i = 0; i < t.len; ++i) {
    const float k = (float)i - (float)t.radius;
    t.k[i] = expf(-k * k / two_s2);
  }
  for (int i = 0; i < t.len; ++i) sum += t.k[i];
  for (int i = 0; i < t.len; ++i) t.k[i] /= sum;
  return t;
}

}  // namespace

namespace {

const HostTables& host_tables() {
  static const HostTables t;
  return t;
}

// The DCT/IDCT/Kaiser constant tables and the FP16 DCT fragments, uploaded once
// per context (the host sources are static).
void ensure_bm3d_tables(imops_ctx* ctx) {
  if (ctx->bm3d_tables_ready) return;
  if (ctx->capturing) throw Error{IMOPS_E_STATE, "BM3D: tables not uploaded before stream capture"};
  const HostTables& tables = host_tables();
  static const std::vector<uint16_t> frags = [&] {
    std::vector<uint16_t> v(kBm3dMfmaFragBytes / 2);
    bm3d_filter_mfma_fragments(tables.dct, v.data());
    return v;
  }();
  hipStream_t s = ctx->stream;
  ctx->bm3d_frags.reserve(kBm3dMfmaFragBytes);
  IMOPS_HIP_CHECK(bm3d_upload_tables(tables.dct, tables.idct, tables.kaiser, s));
  IMOPS_HIP_CHECK(hipMemcpyAsync(ctx->bm3d_frags.ptr, frags.data(), kBm3dMfmaFragBytes, hipMemcpyHostToDevice, s));
  IMOPS_HIP_CHECK(hipStreamSynchronize(s));
  ctx->bm3d_tables_ready = true;
}

}  // namespace

// chroma_bm3d (bm3d.rs:705-717): Y is kept, Cb and Cr are denoised patch-wise.
void stage_chroma_bm3d(imops_ctx* ctx, imops_image* img, float intensity) {
  require_rgb(img, "ChromaDenoise");
  TimedRegion stage_timer(ctx, "chroma_bm3d");
  const int w = (int)img->width, h = (int)img->height;
  const size_t n = (size_t)w * h;
  hipStream_t s = ctx->stream;
  const float ic = std::min(std::max(intensity, 0.001f), 1.0f);
  const float sigma = ic * 80.0f;
  const float threshold = (2.0f + (ic * 2.5f)) * sigma * 1.8f;
  ensure_bm3d_tables(ctx);
  ctx->scr[SCR_BM3D_0].reserve(5 * n * sizeof(float));
  float* f = ctx->scr[SCR_BM3D_0].as<float>();
  float *y = f, *cb = f + n, *cr = f + 2 * n, *cb2 = f + 3 * n, *cr2 = f + 4 * n;
  IMOPS_HIP_CHECK(bm3d_rgb_to_ycbcr(img->rgb.as<float>(), n, y, cb, cr, s));
  IMOPS_HIP_CHECK(chroma_dct_launch(cb, cr, cb2, cr2, w, h, threshold, s));
  IMOPS_HIP_CHECK(ycbcr_merge_launch(y, cb2, cr2, img->rgb.as<float>(), n, s));
}

void stage_bm3d(imops_ctx* ctx, imops_image* img, const Bm3dStageParams& p) {
  require_rgb(img, "BM3D");
  // group tables pack patch positions as x | y << 16
  if (img->width > 65535 || img->height > 65535)
    throw Error{IMOPS_E_INVALID, "BM3D: width and height must be at most 65535"};
  TimedRegion stage_timer(ctx, "bm3d");
  const int w = (int)img->width, h = (int)img->height;
  const size_t n = (size_t)w * h;
  hipStream_t s = ctx->stream;

  // scratch: 3 noisy planes, original Y, 3 basic planes, 3 final planes, blur tmp/out,
  // i64 num/den (3 planes each), group table, device tables
  const int refs_x = w > 8 ? (w - 8 + p.stride - 1) / p.stride : 0;
  const int refs_y = h > 8 ? (h - 8 + p.stride - 1) / p.stride : 0;
  const size_t nrefs = (size_t)refs_x * refs_y;
  ctx->scr[SCR_BM3D_0].reserve(12 * n * sizeof(float));
  ctx->scr[SCR_BM3D_1].reserve(6 * n * sizeof(unsigned long long));
  const size_t gwords = (nrefs ? nrefs : 1) * 17;
  ctx->scr[SCR_BM3D_2].reserve((gwords + nrefs + 64) * sizeof(uint32_t));

  float* f = ctx->scr[SCR_BM3D_0].as<float>();
  float* noisy[3] = {f, f + n, f + 2 * n};
  float* yorig = f + 3 * n;
  float* basic[3] = {f + 4 * n, f + 5 * n, f + 6 * n};
  float* fin[3] = {f + 7 * n, f + 8 * n, f + 9 * n};
  float* btmp = f + 10 * n;
  float* bout = f + 11 * n;
  unsigned long long* num = ctx->scr[SCR_BM3D_1].as<unsigned long long>();
  unsigned long long* den = num + 3 * n;
  uint32_t* groups = ctx->scr[SCR_BM3D_2].as<uint32_t>();
  uint32_t* overflow_count = groups + gwords;           // 64-word aligned slot
  uint32_t* overflow_list = overflow_count + 64;
  // Matrix-core matching needs one 8x8 block of reference patches' search
  // region in LDS; larger geometries use the plain matcher.
  Bm3dMfmaArgs mf = bm3d_mfma_layout(w, h, p.stride, p.window);
  const bool use_mfma = p.precision != IMOPS_BM3D_FP32 && nrefs > 0 && mf.smem <= 160 * 1024 && mf.max_tiles_ok;
  static const HostTables& tables = host_tables();
  const bool mixed_filter = p.precision == IMOPS_BM3D_MIXED;
  ensure_bm3d_tables(ctx);
  const void* filter_frags = ctx->bm3d_frags.ptr;

  IMOPS_HIP_CHECK(bm3d_rgb_to_ycbcr(img->rgb.as<float>(), n, noisy[0], noisy[1], noisy[2], s));
  IMOPS_HIP_CHECK(hipMemcpyAsync(yorig, noisy[0], n * sizeof(float), hipMemcpyDeviceToDevice, s));

  // plain filter: row-pass planes of the noisy image (step 1's guide and step
  // 2's noisy stack) and of the basic estimate (step 2's guide)
  const size_t rfl = mixed_filter || nrefs == 0 ? 0 : bm3d_row_dct_floats(w, h);
  float* noisy_rows[3] = {nullptr, nullptr, nullptr};
  float* basic_rows[3] = {nullptr, nullptr, nullptr};
  if (rfl) {
    ctx->scr[SCR_BM3D_3].reserve(6 * rfl * sizeof(float));
    float* r = ctx->scr[SCR_BM3D_3].as<float>();
    for (int c = 0; c < 3; ++c) {
      noisy_rows[c] = r + c * rfl;
      basic_rows[c] = r + (3 + c) * rfl;
    }
  }

  // workgroup tile of reference patches: the largest whose LDS footprint lets
  // four plain-filter workgroups (16 waves) share a CU, or one mixed-filter
  // workgroup (16 waves for an 8x8 tile) fit
  const int half = p.window / 2;
  auto filter_lds = [&](int t) {
    const int sp = (t - 1) * p.stride + 2 * half + 8;
    return mixed_filter ? bm3d_filter_mfma_smem(sp * sp, t) : bm3d_filter_smem(sp * sp);
  };
  const size_t lds_budget = mixed_filter ? 128 * 1024 : 40 * 1024;
  int default_tile = 8;
  while (default_tile > 1 && filter_lds(default_tile) > lds_budget) default_tile /= 2;
  // candidates for the autotuner: every tile whose footprint fits one CU's LDS
  std::vector<int> tile_cands;
  for (int t = 2; t <= 8; t *= 2)
    if (filter_lds(t) <= 160 * 1024) tile_cands.push_back(t);
  const std::string shape_key = std::to_string(w) + "x" + std::to_string(h) + "/s" + std::to_string(p.stride) + "/w" +
                                std::to_string(p.window) + "/g" + std::to_string(p.max_group);

  auto run_step = [&](const float* const guide[3], bool step1, float* const out[3]) {
    IMOPS_HIP_CHECK(fill_zero_u64(num, 6 * n, s));
    if (nrefs > 0) {
      Bm3dMatchArgs m{};
      for (int c = 0; c < 3; ++c) m.guide[c] = guide[c];
      m.width = w;
      m.height = h;
      m.refs_x = refs_x;
      m.stride = p.stride;
      m.window = p.window;
      m.threshold = step1 ? p.max_dist : p.max_dist * 0.5f;
      m.max_candidates = p.max_candidates;
      m.max_group = p.max_group;
      m.groups = groups;
      bm3d_match_layout(m);
      {
        TimedRegion tr(ctx, "bm3d_match");
        if (use_mfma) {
          for (int c = 0; c < 3; ++c) mf.guide[c] = guide[c];
          mf.width = w;
          mf.height = h;
          mf.refs_x = refs_x;
          mf.refs_y = refs_y;
          mf.stride = p.stride;
          mf.window = p.window;
          mf.threshold = m.threshold;
          mf.max_candidates = p.max_candidates;
          mf.max_group = p.max_group;
          mf.groups = groups;
          mf.overflow_list = overflow_list;
          mf.overflow_count = overflow_count;
          mf.blocks_x = (refs_x + 7) / 8;
          mf.n_blocks = mf.blocks_x * ((refs_y + 7) / 8);
          IMOPS_HIP_CHECK(bm3d_match_mfma_launch(mf, s));
          if (ctx->timer.enabled && !ctx->capturing) {  // diagnostics: patches sent to the plain matcher / selected approximately
            static const char* const names[2] = {"bm3d_fallback_refs", "bm3d_approx_refs"};
            ctx->timer.count_async(names, overflow_count, 2, s);
          }
          // patches whose contender list overflowed: the plain matcher, grid-stride over the list
          Bm3dMatchArgs fb = m;
          fb.ref_list = overflow_list;
          fb.ref_count = overflow_count;
          IMOPS_HIP_CHECK(bm3d_match_list_launch(fb, (int)std::min<size_t>(nrefs, 512), s));
        } else {
          // candidate block width per thread: every choice selects the same groups
          m.block_width = tuned_choice(
              ctx, "bm3d_match/" + shape_key + (step1 ? "/1" : "/2"), {0, 2, 4, 5, 7}, 0,
              [&](int q) {
                Bm3dMatchArgs t = m;
                t.block_width = q;
                IMOPS_HIP_CHECK(bm3d_match_launch(t, (int)nrefs, s, true));
              },
              [] {});
          IMOPS_HIP_CHECK(bm3d_match_launch(m, (int)nrefs, s));
        }
      }
      Bm3dFilterArgs fa{};
      for (int c = 0; c < 3; ++c) {
        fa.guide[c] = guide[c];
        fa.noisy[c] = noisy[c];
        fa.guide_rows[c] = step1 ? noisy_rows[c] : basic_rows[c];
        fa.noisy_rows[c] = noisy_rows[c];
      }
      if (rfl) {
        TimedRegion tr(ctx, "bm3d_row_dct");
        if (step1) IMOPS_HIP_CHECK(bm3d_row_dct_launch(noisy, w, h, noisy_rows, s));
        else IMOPS_HIP_CHECK(bm3d_row_dct_launch(guide, w, h, basic_rows, s));
      }
      fa.width = w;
      fa.height = h;
      fa.refs_x = refs_x;
      fa.refs_y = refs_y;
      fa.stride = p.stride;
      fa.window = p.window;
      fa.tile = default_tile;
      fa.sigma = p.sigma;
      fa.chroma_scale = p.chroma_scale;
      fa.lambda = p.lambda;
      fa.haar = p.group_transform == IMOPS_BM3D_HAAR;
      fa.groups = groups;
      fa.num = num;
      fa.den = den;
      for (int i = 0; i < 64; ++i) {
        fa.dct[i] = tables.dct[i];
        fa.idct[i] = tables.idct[i];
        fa.kaiser[i] = tables.kaiser[i];
      }
      // reference tile per workgroup: the aggregation is exact integer
      // arithmetic, so every tile gives the same sums
      auto launch_filter = [&](int t, bool trial) {
        Bm3dFilterArgs ft = fa;
        ft.tile = t;
        const int nblk = ((refs_x + t - 1) / t) * ((refs_y + t - 1) / t);
        if (mixed_filter)
          IMOPS_HIP_CHECK(bm3d_filter_mfma_launch(ft, step1, nblk, filter_lds(t), filter_frags, s, trial));
        else IMOPS_HIP_CHECK(bm3d_filter_launch(ft, step1, nblk, filter_lds(t), s, trial));
      };
      const size_t tuned_before = ctx->tuned.size();
      const int tile = tuned_choice(
          ctx, std::string(mixed_filter ? "bm3d_filter_mfma/" : "bm3d_filter/") + shape_key + (step1 ? "/ht" : "/wiener"),
          tile_cands, default_tile, [&](int t) { launch_filter(t, true); },
          [&] { IMOPS_HIP_CHECK(fill_zero_u64(num, 6 * n, s)); });
      if (ctx->tuned.size() != tuned_before) IMOPS_HIP_CHECK(fill_zero_u64(num, 6 * n, s));  // trials ran
      TimedRegion tr(ctx, step1 ? "bm3d_filter_ht" : "bm3d_filter_wiener");
      launch_filter(tile, false);
    }
    const float* nz[3] = {noisy[0], noisy[1], noisy[2]};
    IMOPS_HIP_CHECK(bm3d_finalize_launch(num, den, nz, out, n, s));
  };
  const float* g1[3] = {noisy[0], noisy[1], noisy[2]};
  run_step(g1, true, basic);
  const float* g2[3] = {basic[0], basic[1], basic[2]};
  run_step(g2, false, fin);

  const Bm3dBlurTaps taps = blur_taps(3.0f);
  IMOPS_HIP_CHECK(bm3d_blur_launch(yorig, btmp, bout, w, h, taps, s));
  IMOPS_HIP_CHECK(bm3d_merge_launch(fin[0], fin[1], fin[2], yorig, bout, p.detail, img->rgb.as<float>(), n, s));
}

}  // namespace imops

using namespace imops;

extern "C" {

int imops_bm3d_default_params(float intensity, imops_bm3d_params* out) {
  return guarded([&] {
    if (!out) throw Error{IMOPS_E_INVALID, "bm3d_default_params: null argument"};
    const Bm3dStageParams p = Bm3dStageParams::from_intensity(intensity);
    out->sigma = p.sigma;
    out->lambda = p.lambda;
    out->max_dist = p.max_dist;
    out->chroma_scale = p.chroma_scale;
    out->detail = p.detail;
    out->stride = p.stride;
    out->search_window = p.window;
    out->max_group = p.max_group;
    out->max_candidates = p.max_candidates;
    out->precision = p.precision;
    out->group_transform = p.group_transform;
  });
}

int imops_bm3d(imops_ctx* ctx, imops_image* img, float intensity) {
  return guarded([&] {
    if (!ctx || !img) throw Error{IMOPS_E_INVALID, "bm3d: null argument"};
    stage_bm3d(ctx, img, Bm3dStageParams::from_intensity(intensity));
  });
}

int imops_chroma_bm3d(imops_ctx* ctx, imops_image* img, float intensity) {
  return guarded([&] {
    if (!ctx || !img) throw Error{IMOPS_E_INVALID, "chroma_bm3d: null argument"};
    stage_chroma_bm3d(ctx, img, intensity);
  });
}

int imops_bm3d_ex(imops_ctx* ctx, imops_image* img, const imops_bm3d_params* q) {
  return guarded([&] {
    if (!ctx || !img || !q) throw Error{IMOPS_E_INVALID, "bm3d_ex: null argument"};
    Bm3dStageParams p;
    p.sigma = q->sigma;
    p.lambda = q->lambda;
    p.max_dist = q->max_dist;
    p.chroma_scale = q->chroma_scale;
    p.detail = q->detail;
    p.stride = q->stride;
    p.window = q->search_window;
    p.max_group = q->max_group;
    p.max_candidates = q->max_candidates;
    p.precision = q->precision;
    p.group_transform = q->group_transform;
    if (p.group_transform != IMOPS_BM3D_HADAMARD && p.group_transform != IMOPS_BM3D_HAAR)
      throw Error{IMOPS_E_INVALID, "bm3d_ex: group_transform must be IMOPS_BM3D_HADAMARD or IMOPS_BM3D_HAAR"};
    if (p.precision != IMOPS_BM3D_FP32 && p.precision != IMOPS_BM3D_MIXED && p.precision != IMOPS_BM3D_MIXED_MATCH)
      throw Error{IMOPS_E_INVALID, "bm3d_ex: precision must be IMOPS_BM3D_FP32, _MIXED or _MIXED_MATCH"};
    if (p.stride < 1 || p.window < 1 || (p.window & 1) == 0 || p.window > kBm3dMaxWindow || p.max_group < 1 ||
        p.max_group > kBm3dMaxGroup || p.max_candidates < 1)
      throw Error{IMOPS_E_INVALID, "bm3d_ex: stride >= 1, odd search_window <= 41, max_group 1..16, max_candidates >= 1"};
    stage_bm3d(ctx, img, p);
  });
}

}  // extern "C"
