// BM3D (plain FP32) for gfx950: block matching, collaborative filtering with
// fixed-point aggregation, and the luma detail blend.
//
// Behaviour follows pichromatic/src/bm3d.rs bit for bit:
//  * matching (bm3d.rs:281-391): every candidate's 3-channel SSD is summed in
//    the reference's order (row-major, one running f32 sum per channel); the
//    reference's per-row early exits change no decision and no accepted
//    distance (ssd_channel), so the full sums are taken; accepted candidates
//    are capped at max_candidates in scan order, and the group is the first
//    prev_pow2(min(16, n)) entries of the (dist, scan order) sort;
//  * filtering (bm3d.rs:128-238, 490-614): 8x8 DCT as the reference's
//    row-dot-products, WHT across the group with 1/sqrt(n), hard threshold /
//    Wiener, inverse; every output sample is one thread's sequential sum in
//    the reference's order, and the Wiener weight's running sum is taken in
//    stack order;
//  * aggregation (bm3d.rs:463-488): each contribution is converted to i64
//    fixed point (x1e5, truncated) and summed with integer adds, so any
//    summation order gives the reference's exact totals.  A workgroup owns a
//    tile of reference patches and accumulates its footprint in LDS before
//    flushing non-zero sums to the global i64 planes with atomics.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <type_traits>

#include "bm3d.hpp"

namespace imops {
namespace {

__device__ __forceinline__ long long to_fixed(float v) {
  const float s = v * 100000.0f;
  if (fabsf(s) < 2147483648.0f) return (long long)(int)s;  // truncation; the common case
  if (isnan(s)) return 0;
  if (s >= 9.2233720368547758e18f) return INT64_MAX;
  if (s <= -9.2233720368547758e18f) return INT64_MIN;
  return (long long)s;  // truncation toward zero
}

// to_fixed for a weight in [0, 1] (Kaiser sample x group weight, both <= 1):
// the product is below 2^31, so the truncation is the 32-bit conversion.
__device__ __forceinline__ long long to_fixed_unit(float w) { return (long long)(int)(w * 100000.0f); }

// ------------------------------------------------------------------ colour
__global__ void rgb_to_ycbcr_kernel(const float* __restrict__ rgb, size_t n, float* __restrict__ y,
                                    float* __restrict__ cb, float* __restrict__ cr) {
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
    const float r = rgb[i * 3] * 255.0f, g = rgb[i * 3 + 1] * 255.0f, b = rgb[i * 3 + 2] * 255.0f;
    y[i] = 0.299f * r + 0.587f * g + 0.114f * b;
    cb[i] = -0.168736f * r - 0.331264f * g + 0.5f * b + 128.0f;
    cr[i] = 0.5f * r - 0.418688f * g - 0.081312f * b + 128.0f;
  }
}

// ---------------------------------------------------------------- matching
// One wave per reference patch.  A workgroup of up to four waves takes that
// many horizontally adjacent reference patches of one patch row: their clamped
// search windows share rows and overlap in columns, so the union is staged once
// in LDS per guide plane (row pitch a.match_pitch <= 64, one lane per column),
// and each wave then matches its own patch without further barriers.
constexpr int kMatchWaves = 4;   // reference patches (waves) per workgroup at most
constexpr int kMatchPad = 16;    // floats after each staged plane (reads past a row end)
constexpr int kTopK = 8;         // smallest keys a lane keeps; refilled when a lane runs out
constexpr uint32_t kNoKey = 0xffffffffu;

// Wave-wide minimum of a u32: butterflies inside each 16-lane DPP row (xor 1
// and 2 as quad permutes, xor 4 as a half-row mirror then a quad reversal,
// xor 8 as a row rotation), then the four row results through scalar registers.
__device__ __forceinline__ uint32_t wave_min_u32(uint32_t v) {
  v = min(v, (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0xB1, 0xF, 0xF, true));
  v = min(v, (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x4E, 0xF, 0xF, true));
  int t = __builtin_amdgcn_update_dpp(0, (int)v, 0x141, 0xF, 0xF, true);
  t = __builtin_amdgcn_update_dpp(0, t, 0x1B, 0xF, 0xF, true);
  v = min(v, (uint32_t)t);
  v = min(v, (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x128, 0xF, 0xF, true));
  const uint32_t r0 = (uint32_t)__builtin_amdgcn_readlane((int)v, 0), r1 = (uint32_t)__builtin_amdgcn_readlane((int)v, 16);
  const uint32_t r2 = (uint32_t)__builtin_amdgcn_readlane((int)v, 32), r3 = (uint32_t)__builtin_amdgcn_readlane((int)v, 48);
  return min(min(r0, r1), min(r2, r3));
}

// compute_ssd_flat (bm3d.rs:367-391) for Q horizontally adjacent candidates of
// one channel: the running sum over the rows in the reference's order, every
// row.  The reference stops at the first row end whose sum exceeds a stop
// value (thr for channel 0, thr minus the distance so far for channels 1 and
// 2); the full sum decides the same:
//  * the sums never decrease (non-negative terms, monotone rounding), so a
//    stop is hit exactly when the full sum exceeds it;
//  * a candidate that hits a stop is rejected by the reference: in channel 0
//    by d0 > thr; in channel k > 0 the partial sum d > fl(thr - s) (s the
//    rounded distance so far, s <= thr) forces fl(s + d) >= thr, since
//    fl(s + d) < thr would need s + d < thr, i.e. d < thr - s, and then
//    d <= fl(thr - s) for the float d; the total can then not be below thr,
//    and a larger (full) sum keeps every test failing the same way;
//  * a candidate that hits no stop has the full sums as its distances.
// (NaN sums fail every threshold test on both sides.)  `ref` is the
// reference patch, 8 floats per row.
template <int Q>
__device__ __forceinline__ void ssd_channel(const float* win, const float* ref, int pitch, float* out) {
  float acc[Q];
#pragma unroll
  for (int q = 0; q < Q; ++q) acc[q] = 0.0f;
#pragma unroll 1
  for (int dy = 0; dy < 8; ++dy) {
    const float* row = win + dy * pitch;
    float wv[Q + 7], rv[8];
    // values past the row end (at most Q - 1) only reach candidates with
    // cx >= nx, which are dead
#pragma unroll
    for (int t = 0; t < Q + 7; ++t) wv[t] = row[t];
#pragma unroll
    for (int t = 0; t < 8; ++t) rv[t] = ref[dy * 8 + t];
#pragma unroll
    for (int q = 0; q < Q; ++q) {
      float a_ = acc[q];
#pragma unroll
      for (int dx = 0; dx < 8; ++dx) {
        const float diff = wv[q + dx] - rv[dx];
        a_ += diff * diff;
      }
      acc[q] = a_;
    }
  }
#pragma unroll
  for (int q = 0; q < Q; ++q) out[q] = acc[q];
}

// Joint 3-channel distances of one lane's block of Q candidates (window offset
// `off` of its first candidate; bit q of `valid` set for real candidates) --
// block_matching_joint's per-candidate logic (bm3d.rs:318-350) on the full
// channel sums.  Its tests d0 > thr and d0 + d1 > thr are implied by the last
// one, total = (d0 + d1) + d2 < thr (the sums are non-negative and rounding
// is monotone), so they only serve here to skip a channel for the whole wave
// when no candidate of any lane can still pass.  dist = total / 64 for
// accepted candidates, +inf otherwise.
template <int Q>
__device__ __forceinline__ void block_dists(const float* win, int plane, int pitch, const float* ref, int off,
                                            uint32_t valid, float thr, float (&dist)[Q]) {
  float d0[Q], d1[Q], d2[Q];
#pragma unroll
  for (int q = 0; q < Q; ++q) d0[q] = d1[q] = d2[q] = 0.0f;
  if (__any(valid != 0)) ssd_channel<Q>(win + off, ref, pitch, d0);
  float m = __builtin_inff();
#pragma unroll
  for (int q = 0; q < Q; ++q) m = (valid >> q) & 1 ? fminf(m, d0[q]) : m;
  if (__any(m <= thr)) ssd_channel<Q>(win + plane + off, ref + 64, pitch, d1);
  m = __builtin_inff();
#pragma unroll
  for (int q = 0; q < Q; ++q) {
    d0[q] = d0[q] + d1[q];
    m = (valid >> q) & 1 ? fminf(m, d0[q]) : m;
  }
  if (__any(m <= thr)) ssd_channel<Q>(win + 2 * plane + off, ref + 128, pitch, d2);
#pragma unroll
  for (int q = 0; q < Q; ++q) {
    const float total = d0[q] + d2[q];
    dist[q] = ((valid >> q) & 1) && total < thr ? total / 64.0f : __builtin_inff();
  }
}

// Candidate block width for a wave over ny rows of nx candidates: the width
// with the least per-lane work (cost ~ passes x per-row work).
__device__ __forceinline__ int ssd_block_width(int nx, int ny) {
  int best = 2, best_cost = 1 << 30;
  const int widths[5] = {2, 3, 4, 5, 7};
  for (int i = 0; i < 5; ++i) {
    const int q = widths[i];
    const int passes = ((nx + q - 1) / q * ny + 63) / 64;
    const int cost = passes * (26 * q + (q + 7) / 2 + 4);
    if (cost < best_cost) {
      best_cost = cost;
      best = q;
    }
  }
  return best;
}

// Search window of one reference patch (clamped to the image, bm3d.rs:296-305).
struct RefGeom {
  int rx, ry, sx0, sy0, nx, ny;
};
__device__ __forceinline__ RefGeom ref_geom(const Bm3dMatchArgs& a, int ref_id) {
  RefGeom G;
  G.rx = (ref_id % a.refs_x) * a.stride;
  G.ry = (ref_id / a.refs_x) * a.stride;
  const int half = a.window / 2;
  const int wmax = a.width - 8, hmax = a.height - 8;
  G.sx0 = G.rx >= half ? G.rx - half : 0;
  G.sy0 = G.ry >= half ? G.ry - half : 0;
  G.nx = min(G.rx + half, wmax) - G.sx0 + 1;
  G.ny = min(G.ry + half, hmax) - G.sy0 + 1;
  return G;
}

// Match one reference patch with one wave against the staged window (columns
// for this patch from col0) and write its group (bm3d.rs:281-365).
//
// The lanes walk the window in scan order, 64 blocks of Q candidates per pass;
// after each pass a ballot prefix over the lanes' accepted counts gives every
// candidate its rank among the accepted ones, so only the first max_candidates
// - 1 acceptances are eligible, and the walk stops once that many are found
// (later candidates cannot join).  Each lane keeps its kTopK smallest eligible
// keys (distance bits, packed position: the position orders like the scan
// index) sorted in registers; the group is the reference patch followed by the
// g - 1 smallest keys, taken one per round as the wave minimum of the lanes'
// heads.  A lane that had more than kTopK eligible keys and has given up all
// it kept triggers a refill: the walk runs again and every lane keeps the
// kTopK smallest keys above the last key it gave up (the sums are
// deterministic, so the ranks and keys come out the same).
template <int Q>
__device__ __forceinline__ void match_wave(const Bm3dMatchArgs& a, const float* win, const float* ref, int col0,
                                           const RefGeom& G, int ref_id) {
  const int lane = threadIdx.x & 63;
  const int nx = G.nx, ny = G.ny, pitch = a.match_pitch, plane = a.match_plane;
  const int ref_cx = G.rx - G.sx0, ref_cy = G.ry - G.sy0;
  const float thr = a.threshold * 64.0f;
  const int cap = a.max_candidates - 1;
  const int nbx = (nx + Q - 1) / Q, nblk = nbx * ny;
  const uint32_t pos0 = (uint32_t)G.sx0 | ((uint32_t)G.sy0 << 16);
  uint32_t kd[kTopK], kc[kTopK];
  bool over = false;
  int total = 0;  // accepted candidates so far in scan order (wave-uniform)
  long long fk = -1;  // the lane's last selected key (distance bits << 32 | position), -1: none yet

  // One lane-block of QB candidates starting at (cx0, cy), all lanes in scan
  // order: distances, ranks among the accepted, kept keys.  kFloor: keep only
  // keys above the lane's last selected one (the refill walk).
  auto step = [&](auto qb, auto floor_tag, bool vblk, int cy, int cx0) {
    constexpr int QB = decltype(qb)::value;
    constexpr bool kFloor = decltype(floor_tag)::value;
    uint32_t valid = 0;
#pragma unroll
    for (int q = 0; q < QB; ++q) valid |= (uint32_t)(vblk && cx0 + q < nx) << q;
    if (cy == ref_cy && (unsigned)(ref_cx - cx0) < (unsigned)QB) valid &= ~(1u << (ref_cx - cx0));
    float dist[QB];
    block_dists<QB>(win, plane, pitch, ref, cy * pitch + col0 + cx0, valid, thr, dist);
    int cnt = 0;
#pragma unroll
    for (int q = 0; q < QB; ++q) cnt += dist[q] != __builtin_inff() ? 1 : 0;
    // exclusive prefix of the lanes' counts (< 8) and their sum, bit by bit
    int rank = total, wsum = 0;
#pragma unroll
    for (int bit = 0; bit < 3; ++bit) {
      if (bit > 0 && QB < (1 << bit)) break;
      const unsigned long long m = __ballot((cnt >> bit) & 1);
      rank += (int)__builtin_amdgcn_mbcnt_hi((uint32_t)(m >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)m, 0u)) << bit;
      wsum += __popcll(m) << bit;
    }
    const uint32_t crow = pos0 + ((uint32_t)cy << 16) + (uint32_t)cx0;
#pragma unroll
    for (int q = 0; q < QB; ++q) {
      const uint32_t x = __float_as_uint(dist[q]);
      const bool acc = dist[q] != __builtin_inff();
      const uint32_t c = crow + q;
      bool keep = acc && rank < cap;
      if (kFloor) keep = keep && (((long long)x << 32) | c) > fk;
      const uint32_t key = keep ? x : kNoKey;
      rank += acc ? 1 : 0;
      // a key falls off (or is not kept) when the list is full
      over |= keep && kd[kTopK - 1] != kNoKey;
      if (__any(key < kd[kTopK - 1])) {
        // sorted insert; an equal key stays behind the ones already kept
        // (their positions are smaller)
        bool lt[kTopK];
#pragma unroll
        for (int i = 0; i < kTopK; ++i) lt[i] = key < kd[i];
#pragma unroll
        for (int i = kTopK - 1; i >= 1; --i) {
          kc[i] = lt[i - 1] ? kc[i - 1] : (lt[i] ? c : kc[i]);
          kd[i] = lt[i - 1] ? kd[i - 1] : (lt[i] ? key : kd[i]);
        }
        kc[0] = lt[0] ? c : kc[0];
        kd[0] = lt[0] ? key : kd[0];
      }
    }
    total += wsum;
  };

  // The window in scan order: passes of Q-wide blocks (window rows split into
  // nbx blocks; a lane's block moves on by 64 blocks a pass) while a whole
  // pass cannot reach the cap, then passes of single candidates from the first
  // position not walked, so that at most 63 candidates past the cap are
  // evaluated.
  auto walk = [&](auto floor_tag) {
#pragma unroll
    for (int i = 0; i < kTopK; ++i) kd[i] = kc[i] = kNoKey;
    over = false;
    total = 0;
    if (!(thr >= 0.0f)) return;  // every first row already exceeds a negative (or NaN) threshold
    const int drow = 64 / nbx, dbx = 64 - drow * nbx;
    int cy = lane / nbx, bx = lane - cy * nbx;  // this lane's block in the pass
    int base = 0;
    for (; base < nblk && total + 64 * Q <= cap; base += 64) {
      step(std::integral_constant<int, Q>{}, floor_tag, cy < ny, cy < ny ? cy : 0, cy < ny ? bx * Q : 0);
      cy += drow;
      bx += dbx;
      if (bx >= nbx) {
        bx -= nbx;
        ++cy;
      }
    }
    if (base >= nblk) return;
    const int brow = base / nbx;
    const int ncand = nx * ny;
    for (int p = brow * nx + (base - brow * nbx) * Q; p < ncand && total < cap; p += 64) {
      const int c = p + lane;
      const bool v = c < ncand;
      const int y = v ? c / nx : 0;
      step(std::integral_constant<int, 1>{}, floor_tag, v, y, v ? c - y * nx : 0);
    }
  };

  walk(std::false_type{});
  const int n_list = 1 + min(total, cap);
  const int limit = min(a.max_group, n_list);
  int g = 1;
  while (g * 2 <= limit) g *= 2;
  const int need = g - 1;
  uint32_t* out = a.groups + (size_t)ref_id * 17;
  for (int r = 0; r < need; ++r) {
    if (__any(kd[0] == kNoKey && over)) walk(std::true_type{});  // rare: refill above each lane's last selected key
    const uint32_t wb = wave_min_u32(kd[0]);
    const unsigned long long m = __ballot(kd[0] == wb);
    // equal distances: the smallest position (scan index) wins
    const uint32_t wc = __popcll(m) == 1 ? (uint32_t)__builtin_amdgcn_readlane((int)kc[0], __ffsll((long long)m) - 1)
                                         : wave_min_u32(kd[0] == wb ? kc[0] : kNoKey);
    if (kd[0] == wb && kc[0] == wc) {
      out[2 + r] = wc;
      fk = ((long long)wb << 32) | wc;
#pragma unroll
      for (int i = 0; i < kTopK - 1; ++i) {
        kd[i] = kd[i + 1];
        kc[i] = kc[i + 1];
      }
      kd[kTopK - 1] = kc[kTopK - 1] = kNoKey;
    }
  }
  if (lane == 0) {
    out[0] = (uint32_t)g;
    out[1] = (uint32_t)G.rx | ((uint32_t)G.ry << 16);
  }
}

// match_wave at the block width the launch asks for (or the one for the rows
// the cap usually needs)
__device__ __forceinline__ void match_wave_q(const Bm3dMatchArgs& a, const float* win, const float* ref, int col0,
                                             const RefGeom& G, int ref_id) {
  const int q = a.block_width ? a.block_width
                              : ssd_block_width(G.nx, min(G.ny, (a.max_candidates - 1) / G.nx + 4));
  switch (q) {
    case 7: match_wave<7>(a, win, ref, col0, G, ref_id); break;
    case 5: match_wave<5>(a, win, ref, col0, G, ref_id); break;
    case 4: match_wave<4>(a, win, ref, col0, G, ref_id); break;
    case 3: match_wave<3>(a, win, ref, col0, G, ref_id); break;
    default: match_wave<2>(a, win, ref, col0, G, ref_id); break;
  }
}

// Stage rows sy0 .. sy0 + rows - 1 and columns sx0 .. sx0 + cols - 1 (cols <= 64)
// of the three guide planes at row pitch a.match_pitch, one lane per column,
// rows spread over the workgroup's nw waves; twelve rows of the three planes
// are loaded before they are stored (one batch covers a 4-wave workgroup's
// window).
__device__ __forceinline__ void stage_window(const Bm3dMatchArgs& a, float* win, int sx0, int sy0, int cols, int rows,
                                             int wave, int nw) {
  constexpr int kRows = 12;
  const int lane = threadIdx.x & 63;
  const int pitch = a.match_pitch, plane = a.match_plane;
  const bool col_ok = lane < cols;
  for (int y = wave; y < rows; y += kRows * nw) {
    float v[kRows][3];
#pragma unroll
    for (int u = 0; u < kRows; ++u) {
      const int yy = y + u * nw;
      const bool ok = col_ok && yy < rows;
      const size_t g = (size_t)(sy0 + (ok ? yy : 0)) * a.width + sx0 + (ok ? lane : 0);
#pragma unroll
      for (int ch = 0; ch < 3; ++ch) v[u][ch] = ok ? a.guide[ch][g] : 0.0f;
    }
#pragma unroll
    for (int u = 0; u < kRows; ++u) {
      const int yy = y + u * nw;
      if (col_ok && yy < rows) {
#pragma unroll
        for (int ch = 0; ch < 3; ++ch) win[ch * plane + yy * pitch + lane] = v[u][ch];
      }
    }
  }
}

// The reference patch (3 x 8 x 8 floats) of one wave, from the guide planes.
__device__ __forceinline__ void stage_ref(const Bm3dMatchArgs& a, float* ref, const RefGeom& G) {
  const int lane = threadIdx.x & 63;
  const size_t g = (size_t)(G.ry + (lane >> 3)) * a.width + G.rx + (lane & 7);
#pragma unroll
  for (int ch = 0; ch < 3; ++ch) ref[ch * 64 + lane] = a.guide[ch][g];
}

// a.match_group reference patches of one patch row per workgroup, one per
// wave.  kTrial only separates autotuner trial launches from real ones in
// kernel traces.
template <bool kTrial>
__global__ void __launch_bounds__(kMatchWaves * 64) bm3d_match_kernel(Bm3dMatchArgs a) {
  extern __shared__ float match_smem[];
  const int R = a.match_group;
  const int groups_x = (a.refs_x + R - 1) / R;
  const int row = blockIdx.x / groups_x, px = R * (blockIdx.x - row * groups_x);
  const int r0 = row * a.refs_x + px;
  const int nref = min(R, a.refs_x - px);
  const int wave = threadIdx.x >> 6, nw = blockDim.x >> 6;
  const RefGeom G0 = ref_geom(a, r0);
  const RefGeom GL = ref_geom(a, r0 + nref - 1);
  const RefGeom G = wave < nref ? ref_geom(a, r0 + wave) : G0;
  float* ref = match_smem + 3 * a.match_plane + wave * 192;
  stage_window(a, match_smem, G0.sx0, G0.sy0, GL.sx0 + GL.nx - G0.sx0 + 7, G0.ny + 7, wave, nw);
  if (wave < nref) stage_ref(a, ref, G);
  __syncthreads();
  if (wave >= nref) return;  // wave-uniform; no barrier follows
  match_wave_q(a, match_smem, ref, G.sx0 - G0.sx0, G, r0 + wave);
}

// The patches named in ref_list[0 .. *ref_count) (the patches the matrix-core
// matcher hands back), one wave per workgroup, grid-stride.
__global__ void __launch_bounds__(64) bm3d_match_list_kernel(Bm3dMatchArgs a) {
  extern __shared__ float match_smem[];
  float* ref = match_smem + 3 * a.match_plane;
  const uint32_t n = *a.ref_count;
  for (uint32_t it = blockIdx.x; it < n; it += gridDim.x) {
    const int ref_id = (int)a.ref_list[it];
    const RefGeom G = ref_geom(a, ref_id);
    stage_window(a, match_smem, G.sx0, G.sy0, G.nx + 7, G.ny + 7, 0, 1);
    stage_ref(a, ref, G);
    __syncthreads();
    match_wave_q(a, match_smem, ref, 0, G, ref_id);
    __syncthreads();
  }
}

// --------------------------------------------------------------- filtering
// DCT, IDCT and Kaiser tables (DctTables, bm3d.rs:417-461) in constant memory,
// copied in by bm3d_upload_tables before each filter launch.  The kernels read
// them through a pointer the compiler cannot prove loop-invariant, so each
// table is fetched into scalar registers next to its use instead of 192 values
// being held (and spilled) for the whole kernel.
__constant__ float g_bm3d_tables[3 * 64];
typedef const __attribute__((address_space(4))) float* ctab_ptr;
__device__ __forceinline__ ctab_ptr bm3d_table(int which) {
  ctab_ptr p = (ctab_ptr)g_bm3d_tables + which * 64;
  asm volatile("" : "+s"(p));
  return p;
}

// A wave filters four (reference patch, channel) groups at once: a 16-lane
// slot per group, lane k of a slot holding patch k of the group (all 64
// samples in registers).  The 8x8 DCTs are register dot products in the
// reference's order, the WHT across the group is a butterfly between slot
// lanes (lane k pairs with k^h), and the Wiener weight's running sum walks
// the stack in order by passing the partial sum from lane k to lane k+1.
constexpr int FT = 256;

// coefficient k of the row pass of one 8-sample row: sum_m x[m] C[k][m]
// (dct_1d_8, bm3d.rs:552-566)
__device__ __forceinline__ float dct_row_coef(const float* xr, ctab_ptr C, int k) {
  const auto ck = C + k * 8;
  return xr[0] * ck[0] + xr[1] * ck[1] + xr[2] * ck[2] + xr[3] * ck[3] + xr[4] * ck[4] + xr[5] * ck[5] +
         xr[6] * ck[6] + xr[7] * ck[7];
}

// column pass of dct_2d_8x8 on the row-pass output y: x[j][k] = sum_n y[n][k] C[j][n].
// Column by column, so each column's inputs die as its outputs appear (x may be y).
__device__ __forceinline__ void dct2_cols(const float* y, float* x, ctab_ptr C) {
#pragma unroll
  for (int k = 0; k < 8; ++k) {
    float o[8];
#pragma unroll
    for (int j = 0; j < 8; ++j) {
      const auto cj = C + j * 8;
      o[j] = y[0 * 8 + k] * cj[0] + y[1 * 8 + k] * cj[1] + y[2 * 8 + k] * cj[2] + y[3 * 8 + k] * cj[3] +
             y[4 * 8 + k] * cj[4] + y[5 * 8 + k] * cj[5] + y[6 * 8 + k] * cj[6] + y[7 * 8 + k] * cj[7];
    }
#pragma unroll
    for (int j = 0; j < 8; ++j) x[j * 8 + k] = o[j];
  }
}

// out[j][k] = sum_n (sum_m x[n][m] C[k][m]) C[j][n]   (dct_2d_8x8, bm3d.rs:526-536)
__device__ __forceinline__ void dct2_regs(float* x, ctab_ptr C) {
  float y[64];
#pragma unroll
  for (int r = 0; r < 8; ++r)
#pragma unroll
    for (int k = 0; k < 8; ++k) y[r * 8 + k] = dct_row_coef(x + r * 8, C, k);
  dct2_cols(y, x, C);
}

// The row pass of a patch's forward DCT depends only on the 8-pixel row
// segment, which every patch and group containing it shares: it is computed
// once per segment here, and the filter runs only the column pass.
struct RowDctPlanes {
  const float* in[3];
  float* rows[3];
};
__global__ void row_dct_kernel(RowDctPlanes pl, int w, int h) {
  const float* __restrict__ in = pl.in[blockIdx.y];
  float* __restrict__ rows = pl.rows[blockIdx.y];
  const int rw = w - 7;
  const size_t n = (size_t)rw * h;
  const ctab_ptr C = bm3d_table(0);
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
    const int y = (int)(i / rw), x = (int)(i - (size_t)y * rw);
    const float* src = in + (size_t)y * w + x;
    float xr[8], o[8];
#pragma unroll
    for (int m = 0; m < 8; ++m) xr[m] = src[m];
#pragma unroll
    for (int k = 0; k < 8; ++k) o[k] = dct_row_coef(xr, C, k);
    float4* dst = reinterpret_cast<float4*>(rows + i * 8);
    dst[0] = make_float4(o[0], o[1], o[2], o[3]);
    dst[1] = make_float4(o[4], o[5], o[6], o[7]);
  }
}

// u[k][n] = sum_j x[j][k] D[n][j]; out[n][m] = sum_k u[k][n] D[m][k]   (idct_2d_8x8, bm3d.rs:538-548)
__device__ __forceinline__ void idct2_regs(float* x, ctab_ptr D) {
  float u[64];
#pragma unroll
  for (int k = 0; k < 8; ++k)
#pragma unroll
    for (int n = 0; n < 8; ++n) {
      const auto dn = D + n * 8;
      u[k * 8 + n] = x[0 * 8 + k] * dn[0] + x[1 * 8 + k] * dn[1] + x[2 * 8 + k] * dn[2] + x[3 * 8 + k] * dn[3] +
                     x[4 * 8 + k] * dn[4] + x[5 * 8 + k] * dn[5] + x[6 * 8 + k] * dn[6] + x[7 * 8 + k] * dn[7];
    }
#pragma unroll
  for (int n = 0; n < 8; ++n)
#pragma unroll
    for (int m = 0; m < 8; ++m) {
      const auto dm = D + m * 8;
      x[n * 8 + m] = u[0 * 8 + n] * dm[0] + u[1 * 8 + n] * dm[1] + u[2 * 8 + n] * dm[2] + u[3 * 8 + n] * dm[3] +
                     u[4 * 8 + n] * dm[4] + u[5 * 8 + n] * dm[5] + u[6 * 8 + n] * dm[6] + u[7 * 8 + n] * dm[7];
    }
}

// Walsh-Hadamard across the group (walsh_hadamard_1d, bm3d.rs:595-614):
// stage h pairs lane k with k^h, the lower lane keeping x + y and the upper
// x - y.  Each stage is one fused multiply-add per value with the partner read
// through DPP inside the slot's 16-lane row (h = 1, 2: quad permutes; h = 4: a
// half-row mirror, then the quad reversal, 7 - (k^3) = k^4; h = 8: a row
// rotation by 8): w_k <- w_k + m_k w_(k^h), m_k = +1 on the lower lane and -1
// on the upper one.  The multiply by +-1 is exact and the add rounds once, so
// the lower lane gets x + y and the upper y - x = -(x - y) (round to nearest is
// symmetric): the upper lane holds the reference's value negated.  Partners
// have been negated alike (they differ only in bits not yet processed), so the
// per-lane sign is (-1)^popcount(k & (g - 1)) after the stages, and it is folded
// into the final scale (exact).  Only zeros can come out with the other sign,
// which no later step distinguishes (thresholds compare |x|, the fixed-point
// aggregation truncates both to 0).
//
// The DPP operand must not read a VGPR written by the previous two VALU
// instructions, nor follow an EXEC write by fewer than five wait states; each
// block starts with s_nop 4 (the stages run under per-slot EXEC masks).
#define IMOPS_FMAC_DPP(i, ctrl) "v_fmac_f32_dpp %" #i ", %" #i ", %16 " ctrl " row_mask:0xf bank_mask:0xf bound_ctrl:1\n"
#define IMOPS_FMAC_DPP16(ctrl)                                                                                   \
  "s_nop 4\n" IMOPS_FMAC_DPP(0, ctrl) IMOPS_FMAC_DPP(1, ctrl) IMOPS_FMAC_DPP(2, ctrl) IMOPS_FMAC_DPP(3, ctrl)   \
      IMOPS_FMAC_DPP(4, ctrl) IMOPS_FMAC_DPP(5, ctrl) IMOPS_FMAC_DPP(6, ctrl) IMOPS_FMAC_DPP(7, ctrl)           \
          IMOPS_FMAC_DPP(8, ctrl) IMOPS_FMAC_DPP(9, ctrl) IMOPS_FMAC_DPP(10, ctrl) IMOPS_FMAC_DPP(11, ctrl)     \
              IMOPS_FMAC_DPP(12, ctrl) IMOPS_FMAC_DPP(13, ctrl) IMOPS_FMAC_DPP(14, ctrl) IMOPS_FMAC_DPP(15, ctrl)
#define IMOPS_V16(b)                                                                                          \
  "+v"(v[b + 0]), "+v"(v[b + 1]), "+v"(v[b + 2]), "+v"(v[b + 3]), "+v"(v[b + 4]), "+v"(v[b + 5]),          \
      "+v"(v[b + 6]), "+v"(v[b + 7]), "+v"(v[b + 8]), "+v"(v[b + 9]), "+v"(v[b + 10]), "+v"(v[b + 11]),    \
      "+v"(v[b + 12]), "+v"(v[b + 13]), "+v"(v[b + 14]), "+v"(v[b + 15])
// h = 4: the mirror into temporaries, then the quad reversal fused into the FMA

template <int H>
__device__ __forceinline__ void wht_stage(float* v, float m) {
  if (H == 1) {
#pragma unroll
    for (int b = 0; b < 64; b += 16) asm volatile(IMOPS_FMAC_DPP16("quad_perm:[1,0,3,2]") : IMOPS_V16(b) : "v"(m));
  } else if (H == 2) {
#pragma unroll
    for (int b = 0; b < 64; b += 16) asm volatile(IMOPS_FMAC_DPP16("quad_perm:[2,3,0,1]") : IMOPS_V16(b) : "v"(m));
  } else if (H == 8) {
#pragma unroll
    for (int b = 0; b < 64; b += 16) asm volatile(IMOPS_FMAC_DPP16("row_ror:8") : IMOPS_V16(b) : "v"(m));
  } else {
#pragma unroll
    for (int b = 0; b < 64; b += 8) {
      float t0, t1, t2, t3, t4, t5, t6, t7;
      asm volatile(
          "s_nop 4\n"
          "v_mov_b32_dpp %8, %0 row_half_mirror row_mask:0xf bank_mask:0xf\n"
          "v_mov_b32_dpp %9, %1 row_half_mirror row_mask:0xf bank_mask:0xf\n"
          "v_mov_b32_dpp %10, %2 row_half_mirror row_mask:0xf bank_mask:0xf\n"
          "v_mov_b32_dpp %11, %3 row_half_mirror row_mask:0xf bank_mask:0xf\n"
          "v_mov_b32_dpp %12, %4 row_half_mirror row_mask:0xf bank_mask:0xf\n"
          "v_mov_b32_dpp %13, %5 row_half_mirror row_mask:0xf bank_mask:0xf\n"
          "v_mov_b32_dpp %14, %6 row_half_mirror row_mask:0xf bank_mask:0xf\n"
          "v_mov_b32_dpp %15, %7 row_half_mirror row_mask:0xf bank_mask:0xf\n"
          "s_nop 1\n"
          "v_fmac_f32_dpp %0, %8, %16 quad_perm:[3,2,1,0] row_mask:0xf bank_mask:0xf bound_ctrl:1\n"
          "v_fmac_f32_dpp %1, %9, %16 quad_perm:[3,2,1,0] row_mask:0xf bank_mask:0xf bound_ctrl:1\n"
          "v_fmac_f32_dpp %2, %10, %16 quad_perm:[3,2,1,0] row_mask:0xf bank_mask:0xf bound_ctrl:1\n"
          "v_fmac_f32_dpp %3, %11, %16 quad_perm:[3,2,1,0] row_mask:0xf bank_mask:0xf bound_ctrl:1\n"
          "v_fmac_f32_dpp %4, %12, %16 quad_perm:[3,2,1,0] row_mask:0xf bank_mask:0xf bound_ctrl:1\n"
          "v_fmac_f32_dpp %5, %13, %16 quad_perm:[3,2,1,0] row_mask:0xf bank_mask:0xf bound_ctrl:1\n"
          "v_fmac_f32_dpp %6, %14, %16 quad_perm:[3,2,1,0] row_mask:0xf bank_mask:0xf bound_ctrl:1\n"
          "v_fmac_f32_dpp %7, %15, %16 quad_perm:[3,2,1,0] row_mask:0xf bank_mask:0xf bound_ctrl:1\n"
          : "+v"(v[b + 0]), "+v"(v[b + 1]), "+v"(v[b + 2]), "+v"(v[b + 3]), "+v"(v[b + 4]), "+v"(v[b + 5]),
            "+v"(v[b + 6]), "+v"(v[b + 7]), "=&v"(t0), "=&v"(t1), "=&v"(t2), "=&v"(t3), "=&v"(t4), "=&v"(t5),
            "=&v"(t6), "=&v"(t7)
          : "v"(m));
    }
  }
}

__device__ __forceinline__ void wht_lanes(float* v, int k, int g, float scale) {
  if (g > 1) wht_stage<1>(v, (k & 1) ? -1.0f : 1.0f);
  if (g > 2) wht_stage<2>(v, (k & 2) ? -1.0f : 1.0f);
  if (g > 4) wht_stage<4>(v, (k & 4) ? -1.0f : 1.0f);
  if (g > 8) wht_stage<8>(v, (k & 8) ? -1.0f : 1.0f);
  const float s = (__popc(k & (g - 1)) & 1) ? -scale : scale;
#pragma unroll
  for (int i = 0; i < 64; ++i) v[i] *= s;
}

// Haar group transform on the lanes (the in-place lifting layout: lane k of
// the group ends as the coefficient whose lowest set bit is h = k & -k, the
// block sum over [k - h, k) minus the one over [k, k + h), lane 0 as the sum of
// all g; scales 1/sqrt(2h) and 1/sqrt(g)).  Each stage is the Walsh-Hadamard
// stage's DPP fused multiply-add with a multiplier of 0 on the lanes that do
// not take part (a lane that is not a multiple of h keeps its value).  The
// upper lane of a pair computes y - x, the exact negation of x - y.  Forward:
// every lane that takes part at stage h has only been a lower lane before, so
// the negation shows only in the lanes' final values and the scale carries a
// sign of -1 for k > 0.  Inverse (the transpose: scales first, then stages
// from g/2 down to 1): before stage h the lower lane k of a pair holds its
// value negated unless k = 0 (it was the upper lane at the stage of its
// lowest bit) and the upper lane is untouched, so the multipliers are chosen
// to keep each pair's relation, and every lane but 0 ends negated; the caller
// folds that sign into the aggregation weight.  Groups of two (where both
// transforms are the same matrix) take the Walsh-Hadamard path in the inverse,
// whose order (butterfly, then scale) they then share bit for bit.
__device__ __forceinline__ float haar_scale_lane(int k, int g) {
  const int span = k == 0 ? g : 2 * (k & -k);
  return span > 0 ? 1.0f / sqrtf((float)span) : 0.0f;  // g = 0: no group, lanes hold zeros
}
__device__ __forceinline__ float haar_mult(int k, int h, bool inverse) {
  if (k & (h - 1)) return 0.0f;
  if ((k & h) == 0) return (inverse && k != 0) ? -1.0f : 1.0f;  // lower lane
  return (inverse && k != h) ? 1.0f : -1.0f;                    // upper lane
}
__device__ __forceinline__ void haar_lanes_fwd(float* v, int k, int g) {
  if (g > 1) wht_stage<1>(v, haar_mult(k, 1, false));
  if (g > 2) wht_stage<2>(v, haar_mult(k, 2, false));
  if (g > 4) wht_stage<4>(v, haar_mult(k, 4, false));
  if (g > 8) wht_stage<8>(v, haar_mult(k, 8, false));
  const float s = k == 0 ? haar_scale_lane(0, g) : -haar_scale_lane(k, g);
#pragma unroll
  for (int i = 0; i < 64; ++i) v[i] *= s;
}
__device__ __forceinline__ void haar_lanes_inv(float* v, int k, int g) {
  const float s = haar_scale_lane(k, g);
#pragma unroll
  for (int i = 0; i < 64; ++i) v[i] *= s;
  if (g > 8) wht_stage<8>(v, haar_mult(k, 8, true));
  if (g > 4) wht_stage<4>(v, haar_mult(k, 4, true));
  if (g > 2) wht_stage<2>(v, haar_mult(k, 2, true));
  if (g > 1) wht_stage<1>(v, haar_mult(k, 1, true));
}

// LDS: the footprint's i64 numerators, then its u32 denominators.  A
// workgroup's denominator sum at one pixel is at most 64 references x 16
// patches x 1e5 < 2^32, so the 32-bit partial sums are exact.
// kHaar: the Haar group transform instead of the reference's Walsh-Hadamard
// (a separate instantiation, so the reference's path compiles unchanged).
template <bool kStep1, bool kTrial, bool kHaar>
__global__ void __launch_bounds__(FT, 3) bm3d_filter_kernel(Bm3dFilterArgs a) {
  extern __shared__ unsigned long long acc[];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int slot = lane >> 4, k = lane & 15;
  const int tiles_x = (a.refs_x + a.tile - 1) / a.tile;
  const int tx = blockIdx.x % tiles_x, ty = blockIdx.x / tiles_x;
  const int r0x = tx * a.tile, r0y = ty * a.tile;
  const int r1x = min(r0x + a.tile, a.refs_x), r1y = min(r0y + a.tile, a.refs_y);
  const int nrx = r1x - r0x, nref = nrx * (r1y - r0y);
  const int half = a.window / 2;
  const int fx0 = max(r0x * a.stride - half, 0), fy0 = max(r0y * a.stride - half, 0);
  const int fx1 = min((r1x - 1) * a.stride + half, a.width - 8) + 8;  // exclusive
  const int fy1 = min((r1y - 1) * a.stride + half, a.height - 8) + 8;
  const int fw = fx1 - fx0, fh = fy1 - fy0, fp = fw * fh;
  uint32_t* den = reinterpret_cast<uint32_t*>(acc + fp);

  for (int ch = 0; ch < 3; ++ch) {
    for (int q = tid; q < fp; q += FT) {
      acc[q] = 0ull;
      den[q] = 0u;
    }
    __syncthreads();
    const float ch_sigma = ch == 0 ? a.sigma : a.sigma * a.chroma_scale;
    const int rw = a.width - 7;
    const float* __restrict__ grows = a.guide_rows[ch];
    const float* __restrict__ nrows = a.noisy_rows[ch];
    // 16 groups in flight per workgroup (4 waves x 4 slots); every lane runs the
    // same number of iterations so the slot shuffles stay converged.
    for (int base = 0; base < nref; base += 16) {
      const int r = base + wave * 4 + slot;
      const bool have = r < nref;
      int g = 0, px = 0, py = 0;
      if (have) {
        const int ref_id = (r0y + r / nrx) * a.refs_x + r0x + r % nrx;
        const uint32_t* grp = a.groups + (size_t)ref_id * 17;
        g = (int)grp[0];
        if (k < g) {
          const uint32_t pp = grp[1 + k];
          px = (int)(pp & 0xffffu);
          py = (int)(pp >> 16);
        }
      }
      const bool act = k < g;
      // the patch's row segments, already through the DCT's row pass.  Lanes
      // without a patch hold position (0, 0): they load in bounds and keep zeros.
      float gv[64], nv[64], yr[64];
      auto load_rows = [&](const float* __restrict__ rows, float* y) {
#pragma unroll
        for (int dy = 0; dy < 8; ++dy) {
          const size_t off = ((size_t)(py + dy) * rw + px) * 8;
          const float4 r0 = *reinterpret_cast<const float4*>(rows + off);
          const float4 r1 = *reinterpret_cast<const float4*>(rows + off + 4);
          const float rr[8] = {r0.x, r0.y, r0.z, r0.w, r1.x, r1.y, r1.z, r1.w};
#pragma unroll
          for (int i = 0; i < 8; ++i) y[dy * 8 + i] = act ? rr[i] : 0.0f;
        }
      };
      load_rows(grows, yr);
      const float wscale = g > 0 ? 1.0f / sqrtf((float)g) : 0.0f;
      dct2_cols(yr, gv, bm3d_table(0));
      if (kHaar) haar_lanes_fwd(gv, k, g);
      else wht_lanes(gv, k, g, wscale);
      float weight;
      if (kStep1) {
        // hard_threshold (bm3d.rs:240-255): stack element 0 (patch 0, coefficient 0) is kept
        const float th = a.lambda * ch_sigma;
        int nz = 0;
#pragma unroll
        for (int i = 0; i < 64; ++i) {
          if (i == 0 && k == 0) { nz += act ? 1 : 0; continue; }
          if (fabsf(gv[i]) < th) gv[i] = 0.0f;
          else nz += act ? 1 : 0;
        }
        for (int off = 8; off > 0; off >>= 1) nz += __shfl_xor(nz, off, 16);
        weight = nz > 0 ? 1.0f / (float)nz : 1.0f;
#pragma unroll
        for (int i = 0; i < 64; ++i) nv[i] = gv[i];
      } else {
        // wiener_filter (bm3d.rs:257-272): the coefficients depend on the guide
        // stack only, so they (and the weight) are formed before the noisy stack
        // is transformed, which keeps one 64-value stack fewer live
        const float s2 = ch_sigma * ch_sigma;
#pragma unroll
        for (int i = 0; i < 64; ++i) {
          if (i == 0 && k == 0) { gv[0] = 1.0f; continue; }  // element 0: weight term 1, sample untouched
          const float e = gv[i] * gv[i];
          gv[i] = e / (e + s2 + 1e-5f);
        }
        // running sum of coef * coef in stack order: 1 + patch 0 (from its element 1), then patch 1, ...
        float run = 1.0f;
        int gmax = g;
        for (int off = 16; off < 64; off <<= 1) gmax = max(gmax, __shfl_xor(gmax, off, 64));
        for (int kk = 0; kk < gmax; ++kk) {
          if (k == kk && kk < g) {
#pragma unroll
            for (int i = 0; i < 64; ++i)
              if (!(kk == 0 && i == 0)) run += gv[i] * gv[i];
          }
          run = __shfl(run, (slot << 4) + kk, 64);
        }
        weight = run > 0.0f ? 1.0f / run : 1.0f;
        load_rows(nrows, yr);
        dct2_cols(yr, nv, bm3d_table(0));
        if (kHaar) haar_lanes_fwd(nv, k, g);
        else wht_lanes(nv, k, g, wscale);
#pragma unroll
        for (int i = 0; i < 64; ++i) nv[i] *= gv[i];  // element 0 of patch 0: coefficient 1 (exact)
      }
      // inverse (bm3d.rs:508-524) and aggregation into the footprint
      // the inverse Haar leaves lanes k > 0 negated (see haar_lanes_inv); the
      // inverse DCT is odd, so the numerator's weight carries the sign back
      // (an exact product of two negated operands)
      const bool haar_inv = kHaar && g > 2;
      if (haar_inv) haar_lanes_inv(nv, k, g);
      else wht_lanes(nv, k, g, wscale);
      idct2_regs(nv, bm3d_table(1));
      const ctab_ptr kaiser = bm3d_table(2);
      const float nsign = (haar_inv && k != 0) ? -1.0f : 1.0f;
      if (act) {
#pragma unroll
        for (int dy = 0; dy < 8; ++dy)
#pragma unroll
          for (int dx = 0; dx < 8; ++dx) {
            const float kw = kaiser[dy * 8 + dx] * weight;
            const int fi = (py + dy - fy0) * fw + (px + dx - fx0);
            atomicAdd(&acc[fi], (unsigned long long)to_fixed(nv[dy * 8 + dx] * (kw * nsign)));
            atomicAdd(&den[fi], (uint32_t)to_fixed_unit(kw));
          }
      }
    }
    __syncthreads();
    unsigned long long* gnum = a.num + (size_t)ch * a.width * a.height;
    unsigned long long* gden = a.den + (size_t)ch * a.width * a.height;
    for (int q = tid; q < fp; q += FT) {
      const unsigned long long n = acc[q], d = den[q];
      if (n | d) {
        const size_t gi = (size_t)(fy0 + q / fw) * a.width + fx0 + q % fw;
        if (n) atomicAdd(&gnum[gi], n);
        if (d) atomicAdd(&gden[gi], d);
      }
    }
    __syncthreads();
  }
}

// num/den -> estimate (bm3d.rs:225-236); optionally clears the accumulators.
__global__ void bm3d_finalize_kernel(const unsigned long long* __restrict__ num, const unsigned long long* __restrict__ den,
                                     const float* __restrict__ noisy0, const float* __restrict__ noisy1,
                                     const float* __restrict__ noisy2, float* __restrict__ out0,
                                     float* __restrict__ out1, float* __restrict__ out2, size_t n) {
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < 3 * n; i += (size_t)gridDim.x * blockDim.x) {
    const int ch = (int)(i / n);
    const size_t p = i - (size_t)ch * n;
    const float nn = (float)(long long)num[i] / 100000.0f;
    const float dd = (float)(long long)den[i] / 100000.0f;
    const float* noisy = ch == 0 ? noisy0 : (ch == 1 ? noisy1 : noisy2);
    float* out = ch == 0 ? out0 : (ch == 1 ? out1 : out2);
    out[p] = dd > 1e-6f ? nn / dd : noisy[p];
  }
}

__global__ void fill_u64_kernel(unsigned long long* p, size_t n) {
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) p[i] = 0ull;
}

// Separable in-bounds-normalised Gaussian of gaussian_blur_1ch (bm3d.rs:648-699).
template <bool kVertical>
__global__ void blur1_kernel(const float* __restrict__ in, float* __restrict__ out, int w, int h,
                             Bm3dBlurTaps taps) {
  const size_t n = (size_t)w * h;
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
    const int y = (int)(i / w), x = (int)(i - (size_t)y * w);
    float val = 0.0f, ws = 0.0f;
    for (int k = 0; k < taps.len; ++k) {
      const int q = (kVertical ? y : x) - taps.radius + k;
      if (q >= 0 && q < (kVertical ? h : w)) {
        val += in[kVertical ? (size_t)q * w + x : (size_t)y * w + q] * taps.k[k];
        ws += taps.k[k];
      }
    }
    out[i] = val / ws;
  }
}

// The same blur with the radius fixed at compile time: each thread makes Q
// neighbouring outputs of a line from one register window of Q + 2R samples
// (every tap still added in order, out-of-bounds taps skipped as above).
template <bool kVertical, int R>
__global__ void blur1_win_kernel(const float* __restrict__ in, float* __restrict__ out, int w, int h,
                                 Bm3dBlurTaps taps) {
  constexpr int Q = 4, L = 2 * R + 1;
  const int len = kVertical ? h : w, nseg = (len + Q - 1) / Q;
  const size_t total = (size_t)(kVertical ? w : h) * nseg;
  auto at = [&](int line, int q) -> size_t { return kVertical ? (size_t)q * w + line : (size_t)line * w + q; };
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < total; i += (size_t)gridDim.x * blockDim.x) {
    // consecutive threads: neighbouring columns (vertical) or segments of one row
    const int line = kVertical ? (int)(i % w) : (int)(i / nseg);
    const int p0 = (kVertical ? (int)(i / w) : (int)(i % nseg)) * Q;
    float win[Q + 2 * R];
#pragma unroll
    for (int t = 0; t < Q + 2 * R; ++t) {
      const int q = p0 - R + t;
      win[t] = (q >= 0 && q < len) ? in[at(line, q)] : 0.0f;
    }
#pragma unroll
    for (int j = 0; j < Q; ++j) {
      const int p = p0 + j;
      if (p >= len) break;
      float val = 0.0f, ws = 0.0f;
#pragma unroll
      for (int k = 0; k < L; ++k) {
        const int q = p - R + k;
        if (q >= 0 && q < len) {
          val += win[j + k] * taps.k[k];
          ws += taps.k[k];
        }
      }
      out[at(line, p)] = val / ws;
    }
  }
}

// Detail blend of the denoised luma + YCbCr -> RGB /255 (bm3d.rs:51-60, 79-92, 629-635).
__global__ void bm3d_merge_kernel(const float* __restrict__ yd, const float* __restrict__ cb,
                                  const float* __restrict__ cr, const float* __restrict__ yorig,
                                  const float* __restrict__ yblur, float detail, float* __restrict__ rgb, size_t n) {
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
    const float hf = yorig[i] - yblur[i];
    float yv = yd[i] + detail * hf;
    yv = yv < 0.0f ? 0.0f : (yv > 255.0f ? 255.0f : yv);
    const float cbv = cb[i] - 128.0f, crv = cr[i] - 128.0f;
    const float r = yv + 1.402f * crv;
    const float g = yv - 0.344136f * cbv - 0.714136f * crv;
    const float b = yv + 1.772f * cbv;
    rgb[i * 3] = fmaxf(r, 0.0f) / 255.0f;
    rgb[i * 3 + 1] = fmaxf(g, 0.0f) / 255.0f;
    rgb[i * 3 + 2] = fmaxf(b, 0.0f) / 255.0f;
  }
}

// ----------------------------------------------------- chroma-only denoise
// dct_chroma_denoise (bm3d.rs:719-759): every 8x8 patch at a 6-pixel step
// (x < W-8, y < H-8) is DCT'd, hard-thresholded (all but coefficient 0),
// inverse-DCT'd and aggregated with the Kaiser window in i64 fixed point.
// Gather form, no atomics: a workgroup owns a kChromaTile-pixel square and
// its 256 threads each filter one of the 16 x 16 patches touching it (those
// of the first row and column also touch the previous tiles, which filter
// them again); the Kaiser-weighted samples wait in LDS, then every pixel of
// the tile adds the fixed-point terms of the one to four patches over it.
// Integer sums do not depend on order, so the totals are the reference's.
constexpr int kChromaSteps = 15;                   // patch steps per tile side
constexpr int kChromaTile = kChromaSteps * 6;      // 90 pixels
constexpr int kChromaPos = kChromaSteps + 1;       // patch positions per side (one before the tile)

__global__ void __launch_bounds__(kChromaPos * kChromaPos) chroma_dct_kernel(
    const float* __restrict__ in_cb, const float* __restrict__ in_cr, float* __restrict__ out_cb,
    float* __restrict__ out_cr, int w, int h, float th) {
  __shared__ float prod[kChromaPos * kChromaPos][65];  // odd pitch: lane t's row starts in bank t
  __shared__ int valid[kChromaPos * kChromaPos];
  const float* __restrict__ in = blockIdx.y == 0 ? in_cb : in_cr;
  float* __restrict__ out = blockIdx.y == 0 ? out_cb : out_cr;
  const int tiles_x = (w + kChromaTile - 1) / kChromaTile;
  const int X0 = (blockIdx.x % tiles_x) * kChromaTile, Y0 = (blockIdx.x / tiles_x) * kChromaTile;
  const int t = threadIdx.x;
  {
    const int px = X0 - 6 + 6 * (t % kChromaPos), py = Y0 - 6 + 6 * (t / kChromaPos);
    const bool ok = px >= 0 && py >= 0 && px < w - 8 && py < h - 8;
    valid[t] = ok;
    if (ok) {
      float v[64];
#pragma unroll
      for (int dy = 0; dy < 8; ++dy)
#pragma unroll
        for (int dx = 0; dx < 8; ++dx) v[dy * 8 + dx] = in[(size_t)(py + dy) * w + px + dx];
      dct2_regs(v, bm3d_table(0));
#pragma unroll
      for (int i = 1; i < 64; ++i)  // hard_threshold (bm3d.rs:240-255) keeps coefficient 0
        if (fabsf(v[i]) < th) v[i] = 0.0f;
      idct2_regs(v, bm3d_table(1));
      const ctab_ptr kaiser = bm3d_table(2);
#pragma unroll
      for (int i = 0; i < 64; ++i) prod[t][i] = v[i] * kaiser[i];
    }
  }
  __syncthreads();
  const ctab_ptr kaiser = bm3d_table(2);
  for (int q = t; q < kChromaTile * kChromaTile; q += kChromaPos * kChromaPos) {
    const int lx = q % kChromaTile, ly = q / kChromaTile;
    const int x = X0 + lx, y = Y0 + ly;
    if (x >= w || y >= h) continue;
    long long num = 0, den = 0;
    // patch positions i with X0 - 6 + 6 i <= x <= X0 + 1 + 6 i (likewise j for rows)
    for (int j = (ly + 4) / 6; j <= (ly + 6) / 6; ++j)
      for (int i = (lx + 4) / 6; i <= (lx + 6) / 6; ++i) {
        const int pid = j * kChromaPos + i;
        if (!valid[pid]) continue;
        const int e = (ly + 6 - 6 * j) * 8 + (lx + 6 - 6 * i);
        num += to_fixed(prod[pid][e]);
        den += to_fixed_unit(kaiser[e]);
      }
    const size_t o = (size_t)y * w + x;
    const float nn = (float)num / 100000.0f, dd = (float)den / 100000.0f;
    out[o] = dd > 1e-6f ? nn / dd : in[o];
  }
}

// ycbcr_to_rgb_inplace + merge_channels_inplace (bm3d.rs:83-96, 629-635)
__global__ void ycbcr_merge_kernel(const float* __restrict__ y, const float* __restrict__ cb,
                                   const float* __restrict__ cr, float* __restrict__ rgb, size_t n) {
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
    const float yv = y[i], cbv = cb[i] - 128.0f, crv = cr[i] - 128.0f;
    const float r = yv + 1.402f * crv;
    const float g = yv - 0.344136f * cbv - 0.714136f * crv;
    const float b = yv + 1.772f * cbv;
    rgb[i * 3] = fmaxf(r, 0.0f) / 255.0f;
    rgb[i * 3 + 1] = fmaxf(g, 0.0f) / 255.0f;
    rgb[i * 3 + 2] = fmaxf(b, 0.0f) / 255.0f;
  }
}

int grid1d(size_t n) {
  size_t g = (n + 255) / 256;
  return (int)(g > 65536 ? 65536 : (g ? g : 1));
}

}  // namespace


hipError_t bm3d_upload_tables(const float* dct, const float* idct, const float* kaiser, hipStream_t s) {
  hipError_t e = hipMemcpyToSymbolAsync(HIP_SYMBOL(g_bm3d_tables), dct, 64 * sizeof(float), 0, hipMemcpyHostToDevice, s);
  if (e == hipSuccess)
    e = hipMemcpyToSymbolAsync(HIP_SYMBOL(g_bm3d_tables), idct, 64 * sizeof(float), 64 * sizeof(float),
                               hipMemcpyHostToDevice, s);
  if (e == hipSuccess)
    e = hipMemcpyToSymbolAsync(HIP_SYMBOL(g_bm3d_tables), kaiser, 64 * sizeof(float), 128 * sizeof(float),
                               hipMemcpyHostToDevice, s);
  return e;
}
size_t bm3d_filter_smem(int footprint_pixels) {
  return (size_t)footprint_pixels * (sizeof(unsigned long long) + sizeof(uint32_t));
}

hipError_t bm3d_row_dct_launch(const float* const in[3], int w, int h, float* const rows[3], hipStream_t s) {
  if (w < 8 || h <= 0) return hipSuccess;
  const RowDctPlanes pl{{in[0], in[1], in[2]}, {rows[0], rows[1], rows[2]}};
  hipLaunchKernelGGL(row_dct_kernel, dim3(grid1d((size_t)(w - 7) * h), 3), dim3(256), 0, s, pl, w, h);
  return hipGetLastError();
}

hipError_t bm3d_rgb_to_ycbcr(const float* rgb, size_t n, float* y, float* cb, float* cr, hipStream_t s) {
  hipLaunchKernelGGL(rgb_to_ycbcr_kernel, dim3(grid1d(n)), dim3(256), 0, s, rgb, n, y, cb, cr);
  return hipGetLastError();
}

void bm3d_match_layout(Bm3dMatchArgs& a) {
  // the widest union of R adjacent windows is that of interior patches,
  // (R - 1) * stride + window + 7 columns; it must fit one lane per column
  int R = kMatchWaves;
  while (R > 1 && (R - 1) * a.stride + a.window + 7 > 64) --R;
  a.match_group = R;
  a.match_pitch = (R - 1) * a.stride + a.window + 7;
  a.match_plane = (a.window + 7) * a.match_pitch + kMatchPad;
}

static size_t match_smem(const Bm3dMatchArgs& a, int waves) {
  return (size_t)(3 * a.match_plane + waves * 192) * sizeof(float);
}

hipError_t bm3d_match_launch(const Bm3dMatchArgs& a, int nrefs, hipStream_t s, bool trial) {
  if (a.match_group < 1 || a.match_group > kMatchWaves || a.match_pitch < a.window + 7 || a.match_pitch > 64 ||
      a.match_plane < (a.window + 7) * a.match_pitch + kMatchPad)
    return hipErrorInvalidValue;
  const int R = a.match_group;
  const int grid = (nrefs / a.refs_x) * ((a.refs_x + R - 1) / R);
  const size_t smem = match_smem(a, R);
  if (trial) hipLaunchKernelGGL(bm3d_match_kernel<true>, dim3(grid), dim3(64 * R), smem, s, a);
  else hipLaunchKernelGGL(bm3d_match_kernel<false>, dim3(grid), dim3(64 * R), smem, s, a);
  return hipGetLastError();
}

hipError_t bm3d_match_list_launch(const Bm3dMatchArgs& a, int grid, hipStream_t s) {
  if (a.match_pitch < a.window + 7 || a.match_pitch > 64 || a.match_plane < (a.window + 7) * a.match_pitch + kMatchPad)
    return hipErrorInvalidValue;
  hipLaunchKernelGGL(bm3d_match_list_kernel, dim3(grid), dim3(64), match_smem(a, 1), s, a);
  return hipGetLastError();
}

hipError_t bm3d_filter_launch(const Bm3dFilterArgs& a, bool step1, int nblocks, size_t smem, hipStream_t s, bool trial) {
  if (a.haar) {  // (trial launches share the real kernels)
    if (step1) hipLaunchKernelGGL((bm3d_filter_kernel<true, false, true>), dim3(nblocks), dim3(FT), smem, s, a);
    else hipLaunchKernelGGL((bm3d_filter_kernel<false, false, true>), dim3(nblocks), dim3(FT), smem, s, a);
  } else if (trial) {
    if (step1) hipLaunchKernelGGL((bm3d_filter_kernel<true, true, false>), dim3(nblocks), dim3(FT), smem, s, a);
    else hipLaunchKernelGGL((bm3d_filter_kernel<false, true, false>), dim3(nblocks), dim3(FT), smem, s, a);
  } else {
    if (step1) hipLaunchKernelGGL((bm3d_filter_kernel<true, false, false>), dim3(nblocks), dim3(FT), smem, s, a);
    else hipLaunchKernelGGL((bm3d_filter_kernel<false, false, false>), dim3(nblocks), dim3(FT), smem, s, a);
  }
  return hipGetLastError();
}

hipError_t bm3d_finalize_launch(const unsigned long long* num, const unsigned long long* den, const float* const noisy[3],
                                float* const out[3], size_t n, hipStream_t s) {
  hipLaunchKernelGGL(bm3d_finalize_kernel, dim3(grid1d(3 * n)), dim3(256), 0, s, num, den, noisy[0], noisy[1],
                     noisy[2], out[0], out[1], out[2], n);
  return hipGetLastError();
}

hipError_t fill_zero_u64(unsigned long long* p, size_t n, hipStream_t s) {
  hipLaunchKernelGGL(fill_u64_kernel, dim3(grid1d(n)), dim3(256), 0, s, p, n);
  return hipGetLastError();
}

hipError_t bm3d_blur_launch(const float* in, float* tmp, float* out, int w, int h, const Bm3dBlurTaps& taps,
                            hipStream_t s) {
  const size_t n = (size_t)w * h;
  if (taps.radius == 9 && taps.len == 19) {  // sigma 3, the detail blend's blur
    const size_t segs = (size_t)h * ((w + 3) / 4), segs_v = (size_t)w * ((h + 3) / 4);
    hipLaunchKernelGGL((blur1_win_kernel<false, 9>), dim3(grid1d(segs)), dim3(256), 0, s, in, tmp, w, h, taps);
    hipLaunchKernelGGL((blur1_win_kernel<true, 9>), dim3(grid1d(segs_v)), dim3(256), 0, s, tmp, out, w, h, taps);
    return hipGetLastError();
  }
  hipLaunchKernelGGL(blur1_kernel<false>, dim3(grid1d(n)), dim3(256), 0, s, in, tmp, w, h, taps);
  hipLaunchKernelGGL(blur1_kernel<true>, dim3(grid1d(n)), dim3(256), 0, s, tmp, out, w, h, taps);
  return hipGetLastError();
}

hipError_t chroma_dct_launch(const float* cb, const float* cr, float* out_cb, float* out_cr, int w, int h, float th,
                            hipStream_t s) {
  const int tiles = ((w + kChromaTile - 1) / kChromaTile) * ((h + kChromaTile - 1) / kChromaTile);
  hipLaunchKernelGGL(chroma_dct_kernel, dim3(tiles, 2), dim3(kChromaPos * kChromaPos), 0, s, cb, cr, out_cb, out_cr, w,
                     h, th);
  return hipGetLastError();
}

hipError_t ycbcr_merge_launch(const float* y, const float* cb, const float* cr, float* rgb, size_t n, hipStream_t s) {
  hipLaunchKernelGGL(ycbcr_merge_kernel, dim3(grid1d(n)), dim3(256), 0, s, y, cb, cr, rgb, n);
  return hipGetLastError();
}

hipError_t bm3d_merge_launch(const float* yd, const float* cb, const float* cr, const float* yorig, const float* yblur,
                             float detail, float* rgb, size_t n, hipStream_t s) {
  hipLaunchKernelGGL(bm3d_merge_kernel, dim3(grid1d(n)), dim3(256), 0, s, yd, cb, cr, yorig, yblur, detail, rgb, n);
  return hipGetLastError();
}

}  // namespace imops
