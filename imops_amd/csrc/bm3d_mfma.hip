// BM3D block matching on the gfx950 matrix cores, selecting exactly the groups
// the plain FP32 matcher (and pichromatic/src/bm3d.rs:281-391) selects.
//
// Distances are evaluated as ||a||^2 + ||b||^2 - 2 a.b: the three guide planes
// of a workgroup's search region are shifted by their mean and rounded to FP16
// in LDS, the cross terms of 16 reference patches x 16 candidates run on
// v_mfma_f32_16x16x32_f16 (FP32 accumulation, K = 192 = 3 channels x 8 x 8),
// and the norms are 8x8 box sums of the same FP16 values.  Each FP16 distance d
// comes with a rigorous interval for the FP32 sum the reference computes:
//
//   |sqrt(SSD) - sqrt(d_fp16_exact)| <= r = u' (||a'|| + ||b'||)  (rounding of the shifted values, u' = 2^-11 + 2^-23)
//   |d - d_fp16_exact| <= eps = 2e-5 (na + nb) + 1e-3           (FP32 norms and MFMA accumulation)
//   |fl32_reference_sum - SSD| <= 5e-5 SSD                       (192-term sequential FP32 sum)
//
// Acceptance (threshold), the 1024-candidate cap in scan order and the group
// size are decided from those intervals, falling back to the reference's own
// early-exit FP32 evaluation when an interval straddles the threshold.  The
// group itself is chosen in two passes over the candidates: pass 1 tracks the
// two smallest d per lane and reference, which bounds the 15th smallest
// distance from above (T); pass 2 keeps every capped, accepted candidate whose
// lower bound is <= T.  That list contains the true top 15, and its members are
// re-evaluated in the reference's FP32 order (from the FP32 planes in global
// memory) and ranked by (distance, scan order).  When more than 48 candidates
// are indistinguishable at FP16 precision, the group is taken from the
// approximate top 15 (FP16 distance <= the 15th smallest tracked one),
// re-ranked in FP32; if even that list exceeds 32 entries (ties in flat areas),
// the reference patch goes to the plain matcher.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <algorithm>
#include <type_traits>

#include "bm3d.hpp"

namespace imops {
namespace {

typedef _Float16 half8 __attribute__((ext_vector_type(8)));
typedef float floatx4 __attribute__((ext_vector_type(4)));

constexpr int XW = 4;  // waves per workgroup; wave w owns a 4x4 sub-block of the 8x8 reference block
constexpr int XT = 64 * XW;
constexpr int kListCap = kBm3dMfmaListCap;     // exact-band contenders per reference
constexpr int kListCapB = 32;                  // approximate top-K contenders per reference
constexpr int kRowWords = (kBm3dMaxWindow + 1) / 2;  // 16-bit rejection counters per window row, packed in pairs
static_assert(kRowWords * 4 <= kListCap * 2, "rejection counters alias the candidate lists");
constexpr float kU = 4.8852e-4f;    // (2^-11 + 2^-23) * 1.001
constexpr float kEpsRel = 2e-5f;
constexpr float kEpsAbs = 1e-3f;
constexpr float kSumRel = 5e-5f;
constexpr float kInf = __builtin_inff();
constexpr float kNaN = __builtin_nanf("");

struct RefMeta {
  float d15;  // K-th smallest tracked upper bound
  float dK;   // K-th smallest tracked FP16 distance (approximate-list criterion)
  int limit;  // scan index of the last candidate inside the cap
  int g;      // group size
  int pad;    // 1: the cap needs the row walk (rejections before the geometric cap)
};

__device__ __forceinline__ half8 load8(const _Float16* p) {
  half8 v;
  __builtin_memcpy(&v, p, 16);
  return v;
}

typedef float floatx4v __attribute__((ext_vector_type(4)));
__device__ __forceinline__ void load8g(const float* p, float* v) {
  floatx4v a, b;
  __builtin_memcpy(&a, p, 16);
  __builtin_memcpy(&b, p + 4, 16);
  v[0] = a[0]; v[1] = a[1]; v[2] = a[2]; v[3] = a[3];
  v[4] = b[0]; v[5] = b[1]; v[6] = b[2]; v[7] = b[3];
}

// One channel's 8x8 patches of the candidate and the reference, every row
// loaded before any is used (16 independent loads per patch in flight).
__device__ __forceinline__ void load_pair(const float* plane, int W, int rx, int ry, int cx, int cy, float (&c)[8][8],
                                          float (&r)[8][8]) {
#pragma unroll
  for (int dy = 0; dy < 8; ++dy) {
    load8g(plane + (size_t)(cy + dy) * W + cx, c[dy]);
    load8g(plane + (size_t)(ry + dy) * W + rx, r[dy]);
  }
}

// The plain matcher's acceptance of one candidate (early exit per row, per-channel
// stop thresholds) on the FP32 guide planes.  kPreload: every row of a channel is
// loaded before the sum starts (for callers outside the candidate sweeps, which
// have the registers to spare); otherwise row by row.
template <bool kPreload>
__device__ bool exact_accept(const float* const g[3], int W, int rx, int ry, int cx, int cy, float thr) {
  float d[3] = {0.0f, 0.0f, 0.0f};
  float stop = thr;
  for (int ch = 0; ch < 3; ++ch) {
    float acc = 0.0f;
    if (kPreload) {
      float c[8][8], r[8][8];
      load_pair(g[ch], W, rx, ry, cx, cy, c, r);
#pragma unroll
      for (int dy = 0; dy < 8; ++dy) {
#pragma unroll
        for (int dx = 0; dx < 8; ++dx) {
          const float diff = c[dy][dx] - r[dy][dx];
          acc += diff * diff;
        }
        if (acc > stop) break;
      }
    } else {
      for (int dy = 0; dy < 8; ++dy) {
        float c[8], r[8];
        load8g(g[ch] + (size_t)(cy + dy) * W + cx, c);
        load8g(g[ch] + (size_t)(ry + dy) * W + rx, r);
#pragma unroll
        for (int dx = 0; dx < 8; ++dx) {
          const float diff = c[dx] - r[dx];
          acc += diff * diff;
        }
        if (acc > stop) break;
      }
    }
    d[ch] = acc;
    if (ch == 0) {
      if (d[0] > thr) return false;
      stop = thr - d[0];
    } else if (ch == 1) {
      if (d[0] + d[1] > thr) return false;
      stop = thr - (d[0] + d[1]);
    }
  }
  return d[0] + d[1] + d[2] < thr;
}

// Full SSD of an accepted candidate in the reference's order (no early exit can
// have fired for it).
__device__ float exact_total(const float* const g[3], int W, int rx, int ry, int cx, int cy) {
  float d[3];
#pragma unroll
  for (int ch = 0; ch < 3; ++ch) {
    float c[8][8], r[8][8];
    load_pair(g[ch], W, rx, ry, cx, cy, c, r);
    float acc = 0.0f;
#pragma unroll
    for (int dy = 0; dy < 8; ++dy)
#pragma unroll
      for (int dx = 0; dx < 8; ++dx) {
        const float diff = c[dy][dx] - r[dy][dx];
        acc += diff * diff;
      }
    d[ch] = acc;
  }
  return d[0] + d[1] + d[2];
}

// Interval arithmetic below uses the hardware square root (<= 1 ulp); the
// (1 +- 1e-6) factors absorb it.
__device__ __forceinline__ float hsqrt(float x) { return __builtin_amdgcn_sqrtf(x); }

// median of three (v_med3_i32)
__device__ __forceinline__ int med3i(int a, int b, int c) { return max(min(a, b), min(max(a, b), c)); }

__device__ __forceinline__ float wave_sum(float v) {
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}
__device__ __forceinline__ float wave_max(float v) {
  for (int off = 32; off > 0; off >>= 1) v = fmaxf(v, __shfl_xor(v, off, 64));
  return v;
}
__device__ __forceinline__ int wave_max_i(int v) {
  for (int off = 32; off > 0; off >>= 1) v = max(v, __shfl_xor(v, off, 64));
  return __builtin_amdgcn_readfirstlane(v);
}
// Sum over the 16 lanes of this lane's group.
__device__ __forceinline__ int group_sum(int v) {
  for (int off = 8; off > 0; off >>= 1) v += __shfl_xor(v, off, 16);
  return v;
}

// B operand chunk: 8 halves starting at a 4-byte aligned LDS address.
__device__ __forceinline__ half8 load8a(const _Float16* p) {
  const uint32_t* q = reinterpret_cast<const uint32_t*>(__builtin_assume_aligned(p, 4));
  typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));
  const u32x4 v = {q[0], q[1], q[2], q[3]};
  return __builtin_bit_cast(half8, v);
}

// kPD: the staged rows' pitch in dwords when known at compile time (the row
// offsets inside a B fragment then become immediate LDS offsets), 0 otherwise.
template <int kPD>
__global__ void __launch_bounds__(XT) bm3d_match_mfma_kernel(Bm3dMfmaArgs a) {
  extern __shared__ __attribute__((aligned(16))) unsigned char smem[];
  // c0[ch][y][x] = fp16(v - shift); c1[ch][y][x] = c0[ch][y][x + 1], so every
  // 8-half candidate row starts on a 4-byte boundary in one of the two copies.
  _Float16* c0 = reinterpret_cast<_Float16*>(smem + a.off_w16);
  _Float16* c1 = reinterpret_cast<_Float16*>(smem + a.off_w16b);
  float* nb = reinterpret_cast<float*>(smem + a.off_nb);
  float* hs = reinterpret_cast<float*>(smem + a.off_w16b);         // setup scratch, before c1 is staged
  uint64_t* keys = reinterpret_cast<uint64_t*>(smem + a.off_w16b);  // exact phase, after pass 2: rank keys
  uint16_t* list = reinterpret_cast<uint16_t*>(smem + a.off_list);
  uint32_t* rejrow = reinterpret_cast<uint32_t*>(smem + a.off_list);  // pass 1: rejections per (reference, window row)
  int* lcnt = reinterpret_cast<int*>(smem + a.off_cnt);
  int* lcntb = lcnt + XW * 16;
  uint16_t* listb = reinterpret_cast<uint16_t*>(smem + a.off_list) + XW * 16 * kListCap;
  RefMeta* meta = reinterpret_cast<RefMeta*>(smem + a.off_meta);
  float* red = reinterpret_cast<float*>(smem + a.off_red);

  const int tid = threadIdx.x, lane = tid & 63, wave = __builtin_amdgcn_readfirstlane(tid >> 6);
  const int q = lane >> 4, n = lane & 15;

  // Workgroups are dealt round-robin to the 8 XCDs: give each XCD a contiguous
  // run of reference blocks so neighbouring search regions share its L2.
  const int nwg = gridDim.x, wg = blockIdx.x;
  const int xcd = wg & 7;
  const int b = xcd * (nwg >> 3) + min(xcd, nwg & 7) + (wg >> 3);
  const int bx = b % a.blocks_x, by = b / a.blocks_x;
  const int r0x = bx * 8, r0y = by * 8;
  const int nvx = min(8, a.refs_x - r0x), nvy = min(8, a.refs_y - r0y);
  const int st = a.stride, half = a.window / 2;
  const int W = a.width, wmax = a.width - 8, hmax = a.height - 8;
  const int X0 = max(r0x * st - half, 0), X1 = min((r0x + nvx - 1) * st + half, wmax);
  const int Y0 = max(r0y * st - half, 0), Y1 = min((r0y + nvy - 1) * st + half, hmax);
  const int SX = X1 - X0 + 8, SY = Y1 - Y0 + 8;
  const int NX = X1 - X0 + 1, NY = Y1 - Y0 + 1;
  const int P = a.pitch, PN = a.spx - 7, plane = a.spy * a.pitch;
  const float thr = a.threshold * 64.0f;
  const int cap = a.max_candidates - 1;  // accepted candidates kept besides the reference

  // ---- shift = mean of the middle reference patch (per channel); norms = 8x8
  // box sums of the squared FP16 values; then the two FP16 copies
  for (int k = tid; k < XW * 16 * kRowWords; k += XT) rejrow[k] = 0u;
  if (tid < XW * 16) {
    lcnt[tid] = 0;
    lcntb[tid] = 0;
    meta[tid].d15 = kInf;
    meta[tid].dK = kInf;
    meta[tid].limit = 0x7fffffff;
    meta[tid].g = 1;
    meta[tid].pad = 0;
  }
  // every wave computes the shift itself (the same loads and sums, so the same values)
  float shift[3];
  {
    const int mx = (r0x + (nvx - 1) / 2) * st, my = (r0y + (nvy - 1) / 2) * st;
    const size_t at = (size_t)(my + (lane >> 3)) * W + mx + (lane & 7);
    const float g0 = a.guide[0][at], g1 = a.guide[1][at], g2 = a.guide[2][at];
    shift[0] = wave_sum(g0) * (1.0f / 64.0f);
    shift[1] = wave_sum(g1) * (1.0f / 64.0f);
    shift[2] = wave_sum(g2) * (1.0f / 64.0f);
  }
  // the FP16 copy c0, straight from the guide planes: the workgroup's threads
  // take consecutive elements of the 3 x SY x SX region, 16 loads in flight each.
  // A shifted value that is not finite or does not fit FP16 (|v| > 65504) would
  // make distances NaN, which the sweeps read as "outside the window": such a
  // block is matched by the plain matcher instead.
  bool unfit = false;
  {
    const int tot_el = 3 * SY * SX;
    const float* const g0 = a.guide[0];
    const float* const g1 = a.guide[1];
    const float* const g2 = a.guide[2];
    const int drow = XT / SX, dcol = XT - drow * SX;  // a step of XT elements
    for (int base = 0; base < tot_el; base += 16 * XT) {
      float v[16];
      int dst[16], chs[16];
      int e = base + tid, r = e / SX, xx = e - r * SX;
#pragma unroll
      for (int u = 0; u < 16; ++u) {
        const int ch = r >= 2 * SY ? 2 : (r >= SY ? 1 : 0), yy = r - ch * SY;
        chs[u] = ch;
        dst[u] = e < tot_el ? ch * plane + yy * P + xx : -1;
        // plane pointer by select: indexing the argument array with a per-lane
        // channel would load the pointer from memory before every element
        const float* gp = ch == 0 ? g0 : (ch == 1 ? g1 : g2);
        v[u] = e < tot_el ? gp[(size_t)(Y0 + yy) * W + X0 + xx] : 0.0f;
        e += XT;
        xx += dcol;
        r += drow;
        if (xx >= SX) {
          xx -= SX;
          ++r;
        }
      }
#pragma unroll
      for (int u = 0; u < 16; ++u)
        if (dst[u] >= 0) {
          const float hv = v[u] - (chs[u] == 0 ? shift[0] : (chs[u] == 1 ? shift[1] : shift[2]));
          unfit |= !(fabsf(hv) <= 65504.0f);
          c0[dst[u]] = (_Float16)hv;
        }
    }
  }
  if (__syncthreads_or(unfit)) {  // workgroup-uniform: every reference of the block to the plain matcher
    if (tid < 64 && (tid & 7) < nvx && (tid >> 3) < nvy) {
      const uint32_t at = atomicAdd(a.overflow_count, 1u);
      a.overflow_list[at] = (uint32_t)((r0y + (tid >> 3)) * a.refs_x + r0x + (tid & 7));
    }
    return;
  }
  // norms: per pixel row, the sum over channels, then over 8 columns, of the squared
  // FP16 values (read back from the copy) into the row-sum scratch hs (c1's area);
  // a lane takes four adjacent columns from twelve loaded values per channel
  const int ngx = (NX + 3) / 4;
  for (int it = tid; it < SY * ngx; it += XT) {
    const int row = it / ngx, x4 = 4 * (it - row * ngx);
    float acc[4] = {0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
    for (int ch = 0; ch < 3; ++ch) {
      const _Float16* src = c0 + ch * plane + row * P + x4;  // x4 even: 4-byte aligned
      float f[12];
#pragma unroll
      for (int k = 0; k < 6; ++k) {
        const uint32_t w2 = *reinterpret_cast<const uint32_t*>(src + 2 * k);
        f[2 * k] = (float)__builtin_bit_cast(_Float16, (uint16_t)(w2 & 0xffffu));
        f[2 * k + 1] = (float)__builtin_bit_cast(_Float16, (uint16_t)(w2 >> 16));
      }
#pragma unroll
      for (int o = 0; o < 4; ++o)
#pragma unroll
        for (int dx = 0; dx < 8; ++dx) acc[o] += f[o + dx] * f[o + dx];
    }
#pragma unroll
    for (int o = 0; o < 4; ++o)
      if (x4 + o < NX) hs[row * PN + x4 + o] = acc[o];
  }
  __syncthreads();
  // 8-row sums: a thread walks down one column over a quarter of the rows
  float nbm = 0.0f;
  const int seg = (NY + 3) / 4;
  for (int it = tid; it < 4 * NX; it += XT) {
    const int xx = it % NX, r0 = (it / NX) * seg, r1 = min(r0 + seg, NY);
    if (r0 >= r1) continue;
    float win8[8];
#pragma unroll
    for (int dy = 0; dy < 7; ++dy) win8[dy] = hs[(r0 + dy) * PN + xx];
    for (int row = r0; row < r1; ++row) {
      win8[7] = hs[(row + 7) * PN + xx];
      float acc = 0.0f;
#pragma unroll
      for (int dy = 0; dy < 8; ++dy) acc += win8[dy];
      nb[row * PN + xx] = acc;
      nbm = fmaxf(nbm, acc);
#pragma unroll
      for (int dy = 0; dy < 7; ++dy) win8[dy] = win8[dy + 1];
    }
  }
  nbm = wave_max(nbm);
  if (lane == 0) red[3 * XW + wave] = nbm;
  __syncthreads();  // hs is dead: c1[x] = c0[x + 1], a dword (two halves) per lane
  {
    const int pw = P / 2;  // dwords per row (P and plane are even)
    const int tot_dw = 3 * SY * pw;
    for (int base = 0; base < tot_dw; base += 8 * XT) {
      uint32_t lo[8], hi[8];
      int at[8];
#pragma unroll
      for (int u = 0; u < 8; ++u) {
        const int k = base + u * XT + tid;
        const int r = k / pw, c = k - r * pw;
        const int ch = r >= 2 * SY ? 2 : (r >= SY ? 1 : 0), yy = r - ch * SY;
        at[u] = k < tot_dw ? (ch * plane + yy * P) / 2 + c : -1;
        const uint32_t* s0 = reinterpret_cast<const uint32_t*>(c0);
        lo[u] = at[u] >= 0 ? s0[at[u]] : 0u;
        hi[u] = at[u] >= 0 && c + 1 < pw ? s0[at[u] + 1] : 0u;
      }
#pragma unroll
      for (int u = 0; u < 8; ++u)
        if (at[u] >= 0) reinterpret_cast<uint32_t*>(c1)[at[u]] = __builtin_amdgcn_alignbit(hi[u], lo[u], 16);
    }
  }
  __syncthreads();
  float nbmax = red[3 * XW];
  for (int w = 1; w < XW; ++w) nbmax = fmaxf(nbmax, red[3 * XW + w]);
  const float rbmax = kU * hsqrt(nbmax) * (1.0f + 1e-6f) + 1e-5f;

  // ---- this wave's 4x4 reference sub-block.  D layout of the MFMA: lane
  // (q, n), register j holds reference m = 4q + j (sub-block row q, column j)
  // against candidate n of the tile.
  const int wrx0 = r0x + 4 * (wave & 1), wry0 = r0y + 4 * (wave >> 1);
  const int wnx = min(4, a.refs_x - wrx0), wny = min(4, a.refs_y - wry0);
  const bool wave_active = 4 * (wave & 1) < nvx && 4 * (wave >> 1) < nvy;
  const int K = a.max_group - 1;

  const int ry = (wry0 + q) * st;
  const bool yvalid = wave_active && q < wny;
  const int wy0 = max(ry - half, 0), wy1 = min(ry + half, hmax);
  int rxa[4], wx0[4], ew[4], ewv[4], limgeo[4], npos[4];
  bool vj[4];
  float na[4], ra[4], asure[4], arej[4], epsmax[4];
#pragma unroll
  for (int j = 0; j < 4; ++j) {
    rxa[j] = (wrx0 + j) * st;
    vj[j] = yvalid && j < wnx;
    wx0[j] = max(rxa[j] - half, 0);
    ew[j] = min(rxa[j] + half, wmax) - wx0[j] + 1;
    ewv[j] = vj[j] ? ew[j] : 0;  // window width for the unsigned column test (0: no such reference)
    // candidates of the window besides the reference, and the scan index of
    // the last of the first `cap` of them
    npos[j] = vj[j] ? ew[j] * (wy1 - wy0 + 1) - 1 : 0;
    const int self = (ry - wy0) * ew[j] + (rxa[j] - wx0[j]);
    limgeo[j] = npos[j] <= cap ? 0x7fffffff : (cap - 1 < self ? cap - 1 : cap);
    na[j] = vj[j] ? nb[(ry - Y0) * PN + (rxa[j] - X0)] : 0.0f;
    ra[j] = kU * hsqrt(na[j]) * (1.0f + 1e-6f);
    epsmax[j] = kEpsRel * (na[j] + nbmax) + kEpsAbs;
    const float root = hsqrt(thr / (1.0f + kSumRel)) * (1.0f - 1e-6f) - ra[j] - rbmax;
    asure[j] = root > 0.0f ? root * root * (1.0f - 1e-6f) - epsmax[j] : -kInf;
    // d > arej: every candidate's lower bound exceeds the threshold (surely rejected;
    // the per-pair classification would reject it too)
    const float rfar = hsqrt(thr / ((1.0f - kSumRel) * (1.0f - 1e-6f))) * (1.0f + 1e-6f) + ra[j] + rbmax;
    arej[j] = rfar * rfar * ((1.0f + 1e-5f) / ((1.0f - 1e-6f) * (1.0f - 1e-6f))) + epsmax[j];
  }

  int ux0 = 0, ux1 = 0, uy0 = 0, uy1 = -1;
  half8 A[6];
  if (wave_active) {
    ux0 = max(wrx0 * st - half, 0);
    ux1 = min((wrx0 + wnx - 1) * st + half, wmax);
    uy0 = max(wry0 * st - half, 0);
    uy1 = min((wry0 + wny - 1) * st + half, hmax);
    // A operand: lane (q, n) holds reference n (clamped into the valid part of
    // the sub-block), K-group q: see load_b for the K order.
    const int aix = min(n & 3, wnx - 1), aiy = min(n >> 2, wny - 1);
    const int arx = (wrx0 + aix) * st - X0, ary = (wry0 + aiy) * st - Y0;
#pragma unroll
    for (int s = 0; s < 6; ++s) {
      const _Float16* r = c0 + (s >> 1) * plane + (ary + 4 * (s & 1)) * P + arx + 2 * q;
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        A[s][2 * i] = r[i * P];
        A[s][2 * i + 1] = r[i * P + 1];
      }
    }
  }
  const int ntc = (ux1 - ux0) / 16 + 1;

  // Precise classification of one pair from its own bounds; rb = the
  // candidate's share of r.
  auto classify = [&](int j, float d, float nbv, float rb, int cx, int cy) -> bool {
    const float eps = kEpsRel * (na[j] + nbv) + kEpsAbs;
    const float r = ra[j] + rb;
    const float hi = hsqrt(fmaxf(d + eps, 0.0f)) * (1.0f + 1e-6f) + r;
    if (hi * hi * (1.0f + kSumRel) * (1.0f + 1e-6f) < thr) return true;
    const float lo = fmaxf(hsqrt(fmaxf(d - eps, 0.0f)) * (1.0f - 1e-6f) - r, 0.0f);
    if (lo * lo * (1.0f - kSumRel) * (1.0f - 1e-6f) > thr) return false;
    return exact_accept<false>(a.guide, W, rxa[j], ry, cx, cy, thr);
  };

  // ---- tile walk.  Lane n evaluates candidate column cx = ux0 + 16 tc + n in
  // every row: tile columns tc outermost, rows innermost.  Window membership is
  // folded into the distance: na[j] is replaced by NaN in columns outside
  // reference j's window, the candidate norm by NaN in rows outside the
  // window of the lane's reference row, and the reference's own pair by NaN in
  // its row; every comparison with a NaN is false, so such a pair is neither
  // accepted, rejected, tracked nor collected.  The cap is a scan-index test
  // per pair.  Columns past the union are clamped to one of the same parity
  // for the loads and lie outside every window.
  const int nrows = yvalid ? wy1 - wy0 + 1 : 0;
  const int sr = ry - wy0;  // window row of the references themselves
  int cxc = 0;              // this lane's candidate column in the current tile column
  int oxc[4];               // its window column, per reference
  float naC[4];             // na[j] inside reference j's window columns, NaN outside
  auto column = [&](int cx) {
    cxc = cx;
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      oxc[j] = cx - wx0[j];
      naC[j] = (unsigned)oxc[j] < (unsigned)ewv[j] ? na[j] : kNaN;
    }
  };
  // window rows wholly before the cap position `lim` (scan index of the last
  // capped candidate)
  auto rows_before = [&](const int* lim, int* full) {
#pragma unroll
    for (int j = 0; j < 4; ++j) full[j] = !vj[j] ? 0 : (lim[j] >= nrows * ew[j] ? nrows : (lim[j] + 1) / ew[j]);
  };
  // union rows holding a reference of the wave (wave-uniform 128-bit map)
  uint64_t self_lo = 0, self_hi = 0;
  {
    uint32_t m[4] = {0u, 0u, 0u, 0u};
    if (yvalid) {
      const int u = ry - uy0;
      m[u >> 5] |= 1u << (u & 31);
    }
#pragma unroll
    for (int k = 0; k < 4; ++k)
      for (int off = 1; off < 64; off <<= 1) m[k] |= (uint32_t)__shfl_xor((int)m[k], off, 64);
    self_lo = ((uint64_t)__builtin_amdgcn_readfirstlane(m[1]) << 32) | (uint32_t)__builtin_amdgcn_readfirstlane(m[0]);
    self_hi = ((uint64_t)__builtin_amdgcn_readfirstlane(m[3]) << 32) | (uint32_t)__builtin_amdgcn_readfirstlane(m[2]);
  }
  auto self_row = [&](int cy) {
    const int u = cy - uy0;
    return (((u < 64 ? self_lo >> (u & 63) : self_hi >> (u & 63))) & 1ull) != 0;
  };
  // distances of the lane's four pairs in one row; out-of-window pairs are NaN
  auto dists = [&](int cy, const floatx4& acc, float nbv, float* d) {
    const int rr = cy - wy0;
    const float nbq = (unsigned)rr < (unsigned)nrows ? nbv : kNaN;
#pragma unroll
    for (int j = 0; j < 4; ++j) d[j] = fmaf(-2.0f, acc[j], naC[j] + nbq);
    if (self_row(cy)) {
#pragma unroll
      for (int j = 0; j < 4; ++j)
        if (rr == sr && cxc == rxa[j]) d[j] = kNaN;
    }
  };
  // MFMA cross terms of the 16 references against the 16 candidates of one row
  // of a tile column.  The row's B fragments start at `row` (this lane's
  // column, copy c0 or c1 by parity).  K order: chunk s covers rows
  // 4 (s & 1) .. 4 (s & 1) + 3 of channel s >> 1, and K-group q of it the patch
  // columns 2q and 2q + 1 of those four rows (element 2i + j: row i, column
  // 2q + j).  The four K-groups then read neighbouring dwords of one row in
  // each LDS pass, shared between groups, instead of four rows whose dwords
  // fall on the same banks.
  const int PD = kPD ? kPD : P / 2;
  auto load_b = [&](const _Float16* row, half8* B) {
    const uint32_t* p = reinterpret_cast<const uint32_t*>(__builtin_assume_aligned(row + 2 * q, 4));
#pragma unroll
    for (int s = 0; s < 6; ++s) {
      const uint32_t* r = p + (s >> 1) * (plane / 2) + 4 * (s & 1) * PD;
      typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));
      const u32x4 v = {r[0], r[PD], r[2 * PD], r[3 * PD]};
      B[s] = __builtin_bit_cast(half8, v);
    }
  };
  auto mfma_row = [&](const half8* B) -> floatx4 {
    floatx4 acc = {0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
    for (int s = 0; s < 6; ++s) acc = __builtin_amdgcn_mfma_f32_16x16x32_f16(A[s], B[s], acc, 0, 0, 0);
    return acc;
  };
  // Visits the tiles of union rows [rb, re): on_column(cx) once per tile column,
  // then on_tile(cy, acc, nbv) per row.  The products of row r + 1 are issued
  // before row r is evaluated, and the loads of row r + 2 after them.
  auto sweep = [&](int rb, int re, auto&& on_column, auto&& on_tile) {
    const int nr = re - rb;
    if (nr <= 0) return;
    for (int tc = 0; tc < ntc; ++tc) {
      const int cx = ux0 + 16 * tc + n;
      int cxl = cx;
      if (cx > ux1) {
        cxl = ux1 - ((cx - ux1) & 1);
        if (cxl < X0) cxl += 2;  // one-column union: read past it (inside the pitch)
      }
      const int odd = (cxl - X0) & 1;
      const _Float16* bp = (odd ? c1 : c0) + (uy0 + rb - Y0) * P + (cxl - X0 - odd);
      const float* np = nb + (uy0 + rb - Y0) * PN + (min(cxl, X1) - X0);
      on_column(cx);
      half8 B[6];
      load_b(bp, B);
      float nb_next = np[0];
      floatx4 acc_next = mfma_row(B);
      if (nr > 1) load_b(bp + P, B);
      for (int r = 0; r < nr; ++r) {
        const floatx4 acc = acc_next;
        const float nbv = nb_next;
        if (r + 1 < nr) {
          nb_next = np[(r + 1) * PN];
          acc_next = mfma_row(B);
          if (r + 2 < nr) load_b(bp + (r + 2) * P, B);
        }
        on_tile(uy0 + rb + r, acc, nbv);
      }
    }
  };
  // distance of pair j from the cross term: (na + nb) - 2 a.b, one rounding after the sum
  auto dist = [&](int j, float acc, float nbv) { return fmaf(-2.0f, acc, na[j] + nbv); };

  // ---- pass 1: rejections, and per lane the two smallest d among accepted
  // candidates inside the geometric cap.  Pairs that are surely accepted take
  // no further work; one branch per tile handles the rest.
  // The two smallest are tracked on the distances' bit patterns in integer
  // order (= float order for d >= 0; among negative d, whose tracked order is
  // immaterial, reversed), b1 <= b2.  Every fourth row the trackers rotate
  // within the 16-lane group, so candidates along a vertical structure spread
  // over several lanes' pairs.
  int b1[4], b2[4];
  constexpr int kInfBits = 0x7f800000;
#pragma unroll
  for (int j = 0; j < 4; ++j) {
    b1[j] = b2[j] = kInfBits;
  }
  auto track = [&](int j, bool on, float d) {
    const int dv = on ? __float_as_int(d) : kInfBits;
    b2[j] = med3i(b1[j], b2[j], dv);
    b1[j] = min(b1[j], dv);
  };
  int full1[4] = {0, 0, 0, 0};  // window rows wholly before the geometric cap
  const int nru = uy1 - uy0 + 1;   // union rows
  int rows1 = nru;                 // union rows pass 1 swept
  // rejections of reference (q, j) counted so far: its row counters, lane n
  // reads words n and n + 16
  auto rej_total = [&](int j) {
    const uint32_t* rc = rejrow + (wave * 16 + 4 * q + j) * kRowWords;
    const uint32_t w0 = rc[n], w1 = n + 16 < kRowWords ? rc[n + 16] : 0u;
    return group_sum((int)((w0 & 0xffffu) + (w0 >> 16) + (w1 & 0xffffu) + (w1 >> 16)));
  };
  if (wave_active) {
    rows_before(limgeo, full1);
    auto on_tile = [&](int cy, const floatx4& acc, float nbv) {
      const int rr = cy - wy0;
      float d[4];
      dists(cy, acc, nbv, d);
      bool rj[4], am[4], trk[4], ok[4];
      bool anyr = false, anya = false;
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        const bool sure = d[j] <= asure[j];
        rj[j] = d[j] > arej[j];
        am[j] = (d[j] <= arej[j]) & !sure;
        trk[j] = rr * ew[j] + oxc[j] <= limgeo[j];  // inside the geometric cap
        ok[j] = sure & trk[j];
        anyr |= rj[j];
        anya |= am[j];
      }
      // surely rejected pairs: lane n < 4 of each quadrant adds the quadrant's
      // count for reference (q, n) to its row counter (two 16-bit row counters per word)
      if (__ballot(anyr)) {
        uint32_t mine = 0u;
#pragma unroll
        for (int j = 0; j < 4; ++j) {
          const uint32_t c = (uint32_t)__popc((uint32_t)(__ballot(rj[j]) >> (16 * q)) & 0xffffu);
          mine = n == j ? c : mine;
        }
        if (n < 4 && mine)
          atomicAdd(&rejrow[(wave * 16 + 4 * q + n) * kRowWords + (rr >> 1)], mine << (16 * (rr & 1)));
      }
      if (__ballot(anya)) {  // the threshold is within reach of some pair
        const float rb = kU * hsqrt(nbv) * (1.0f + 1e-6f) + 1e-5f;
#pragma unroll
        for (int j = 0; j < 4; ++j) {
          if (am[j]) {
            const bool acc_j = classify(j, d[j], nbv, rb, cxc, cy);
            ok[j] = acc_j && trk[j];
            if (!acc_j) atomicAdd(&rejrow[(wave * 16 + 4 * q + j) * kRowWords + (rr >> 1)], (rr & 1) ? 65536u : 1u);
          }
        }
      }
#pragma unroll
      for (int j = 0; j < 4; ++j) track(j, ok[j], d[j]);
      if ((cy - uy0) % 4 == 3) {
#pragma unroll
        for (int j = 0; j < 4; ++j) {
          b1[j] = __builtin_amdgcn_update_dpp(0, b1[j], 0x127, 0xf, 0xf, false);  // row_ror:7
          b2[j] = __builtin_amdgcn_update_dpp(0, b2[j], 0x127, 0xf, 0xf, false);
        }
      }
    };
    // Rows past every reference's cap cannot change a decision, so the sweep
    // covers the union rows through the last geometric-cap row first (the
    // tracked candidates all lie there) and then extends only while some
    // reference has fewer than `cap` acceptances in the rows swept so far.
    int rlast = -1;
#pragma unroll
    for (int j = 0; j < 4; ++j)
      if (vj[j]) rlast = max(rlast, wy0 - uy0 + (limgeo[j] == 0x7fffffff ? nrows - 1 : limgeo[j] / ew[j]));
    int R = wave_max_i(rlast) + 1;
    sweep(0, R, column, on_tile);
    while (R < nru) {
      int need = 0;
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        const int rs = min(max(R - (wy0 - uy0), 0), nrows);
        const int acc_so_far = rs * ew[j] - (sr < rs ? 1 : 0) - rej_total(j);
        if (vj[j] && rs < nrows && acc_so_far < cap) need = max(need, (cap - acc_so_far + ew[j] - 1) / ew[j]);
      }
      need = wave_max_i(need);
      if (need == 0) break;
      const int R2 = min(nru, R + need);
      sweep(R, R2, column, on_tile);
      R = R2;
    }
    rows1 = R;
  }
  // per reference: accepted count and group size.  A reference whose window
  // was not swept to its end has at least `cap` acceptances in the swept rows:
  // any count above the cap takes the same decisions below.
  int nacc[4];
  bool anyrej[4];
#pragma unroll
  for (int j = 0; j < 4; ++j) {
    // rejections of reference (q, j): its row counters, lane n reads words n and n + 16
    const uint32_t* rc = rejrow + (wave * 16 + 4 * q + j) * kRowWords;
    const uint32_t w0 = rc[n], w1 = n + 16 < kRowWords ? rc[n + 16] : 0u;
    const int complete = rows1 - (wy0 - uy0) >= nrows;
    nacc[j] = complete ? npos[j] - rej_total(j) : cap + 1;
    // rejections up to the row holding the geometric cap (all of that row: a
    // reference flagged without need gets the geometric cap from the walk)
    const int cr = min(full1[j], (yvalid ? wy1 - wy0 + 1 : 0) - 1);
    auto upto = [&](uint32_t w, int row) {
      return (row <= cr ? (int)(w & 0xffffu) : 0) + (row + 1 <= cr ? (int)(w >> 16) : 0);
    };
    anyrej[j] = group_sum(upto(w0, 2 * n) + (n + 16 < kRowWords ? upto(w1, 2 * n + 32) : 0)) > 0;
  }
  if (wave_active && n == 0) {
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const int lim = min(a.max_group, 1 + min(nacc[j], cap));
      int g = 1;
      while (g * 2 <= lim) g *= 2;
      meta[wave * 16 + 4 * q + j].g = g;
      meta[wave * 16 + 4 * q + j].limit = nacc[j] > cap ? limgeo[j] : 0x7fffffff;
      meta[wave * 16 + 4 * q + j].pad = (nacc[j] > cap && anyrej[j]) ? 1 : 0;
    }
  }
  __syncthreads();
  // ---- the cap when candidates before the geometric cap were rejected.  Each
  // wave settles its own sub-block's references: a prefix sum over the window
  // rows (lane = row) of the accepted counts finds the row holding the cap-th
  // acceptance; that row's candidates are re-evaluated from the same products
  // as pass 1 (same operands, so the same distances and classifications), the
  // rare undecided pairs by the reference's own test.
  if (wave_active) {
    const bool padl = lane < 16 && meta[wave * 16 + lane].pad;
    uint64_t padm = __ballot(padl);
    while (padm) {
      const int m = __builtin_ctzll(padm);  // reference m = (row m >> 2, column m & 3) of the sub-block
      padm &= padm - 1;
      const int mq = m >> 2, mj = m & 3, slot = wave * 16 + m;
      const int mrx = (wrx0 + mj) * st, mry = (wry0 + mq) * st;
      const int mx0 = max(mrx - half, 0), mw = min(mrx + half, wmax) - mx0 + 1;
      const int my0 = max(mry - half, 0), mh = min(mry + half, hmax) - my0 + 1;
      auto rejs = [&](int row) { return (int)((rejrow[slot * kRowWords + (row >> 1)] >> (16 * (row & 1))) & 0xffffu); };
      int inrow = 0;
      if (lane < mh) inrow = mw - (my0 + lane == mry ? 1 : 0) - rejs(lane);
      int pref = inrow;  // inclusive prefix over rows
      for (int off = 1; off < 64; off <<= 1) {
        const int v = __shfl_up(pref, off, 64);
        if (lane >= off) pref += v;
      }
      const uint64_t reach = __ballot(lane < mh && pref >= cap);
      if (!reach) continue;  // fewer than cap acceptances (not the case for a flagged reference)
      const int row = __builtin_ctzll(reach);
      const int seen = __shfl(pref - inrow, row, 64);  // acceptances before the row
      const int cy = my0 + row;
      // accepted candidates of the row as a column mask (bit c = column mx0 + c)
      uint64_t acc_cols = 0;
      if (rejs(row) == 0) {
        acc_cols = mw >= 64 ? ~0ull : ((1ull << mw) - 1ull);
      } else {
        auto eval = [&](auto jtag) {
          constexpr int J = decltype(jtag)::value;
          for (int tt = 0; tt * 16 < mw; ++tt) {
            const int cx = mx0 + 16 * tt + n;
            int cxl = cx;
            if (cx > ux1) {
              cxl = ux1 - ((cx - ux1) & 1);
              if (cxl < X0) cxl += 2;
            }
            const int odd = (cxl - X0) & 1;
            half8 B[6];
            load_b((odd ? c1 : c0) + (cy - Y0) * P + (cxl - X0 - odd), B);
            const float nbv = nb[(cy - Y0) * PN + (min(cxl, X1) - X0)];
            const floatx4 acc = mfma_row(B);
            bool ok = false;
            if (q == mq && 16 * tt + n < mw) {
              const float d = dist(J, acc[J], nbv);
              if (d <= asure[J]) ok = true;
              else if (d <= arej[J]) ok = classify(J, d, nbv, kU * hsqrt(nbv) * (1.0f + 1e-6f) + 1e-5f, cx, cy);
            }
            acc_cols |= ((__ballot(ok) >> (16 * mq)) & 0xffffull) << (16 * tt);
          }
        };
        switch (mj) {
          case 0: eval(std::integral_constant<int, 0>{}); break;
          case 1: eval(std::integral_constant<int, 1>{}); break;
          case 2: eval(std::integral_constant<int, 2>{}); break;
          default: eval(std::integral_constant<int, 3>{}); break;
        }
      }
      if (cy == mry) acc_cols &= ~(1ull << (mrx - mx0));  // the reference itself
      // lane c: column c holds the (cap - seen)-th acceptance of the row
      const bool mine = lane < mw && ((acc_cols >> lane) & 1ull);
      const int before = seen + __popcll(acc_cols & ((1ull << lane) - 1ull));
      if (mine && before == cap - 1) meta[slot].limit = row * mw + lane;
    }
  }
  // ---- T: K-th smallest upper bound among the tracked candidates.  The bound
  // grows with the distance (same reference, block-wide norm terms), so it is
  // the bound of the K-th smallest tracked distance (ranked with ties by lane,
  // slot); that K-th distance is kept as well.
  if (wave_active && K > 0) {
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      float u1 = kInf, u2 = kInf;
      // upper bounds with the block's largest candidate norm (each term grows with the norm)
      const float e1 = __int_as_float(b1[j]), e2 = __int_as_float(b2[j]);
      if (b1[j] != kInfBits) {
        const float f = hsqrt(fmaxf(e1 + epsmax[j], 0.0f)) * (1.0f + 1e-6f) + ra[j] + rbmax;
        u1 = f * f * ((1.0f + kSumRel) * (1.0f + 1e-6f));
      }
      if (b2[j] != kInfBits) {
        const float f = hsqrt(fmaxf(e2 + epsmax[j], 0.0f)) * (1.0f + 1e-6f) + ra[j] + rbmax;
        u2 = f * f * ((1.0f + kSumRel) * (1.0f + 1e-6f));
      }
      int s1 = 0, s2 = 0;
      for (int k = 0; k < 16; ++k) {
        const float p1 = __shfl(e1, 16 * q + k, 64), p2 = __shfl(e2, 16 * q + k, 64);
        s1 += (p1 < e1 || (p1 == e1 && k < n)) + (p2 < e1 || (p2 == e1 && k < n));
        s2 += (p1 < e2 || (p1 == e2 && k <= n)) + (p2 < e2 || (p2 == e2 && k < n));
      }
      if (s1 == K - 1) {
        meta[wave * 16 + 4 * q + j].dK = e1;
        meta[wave * 16 + 4 * q + j].d15 = u1;
      }
      if (s2 == K - 1) {
        meta[wave * 16 + 4 * q + j].dK = e2;
        meta[wave * 16 + 4 * q + j].d15 = u2;
      }
    }
  }
  __syncthreads();

  // ---- pass 2: collect every capped, accepted candidate whose lower bound reaches T.
  // A candidate's lower bound is <= T only if d <= (sT + r)^2 + eps with
  // sT = sqrt(T / ((1 - kSumRel)(1 - 1e-6))); r and eps are taken at the block's
  // largest candidate norm, so the test keeps a superset of those candidates.
  const int UW = ux1 - ux0 + 1;
  if (wave_active) {
    float sT[4], dK[4], dA[4];
    int limit[4];
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const RefMeta& mt = meta[wave * 16 + 4 * q + j];
      limit[j] = mt.limit;
      if (K <= 0) sT[j] = -1.0f;                                        // group of one: nothing to rank
      else if (min(nacc[j], cap) <= K || mt.d15 == kInf) sT[j] = kInf;  // every accepted candidate
      else sT[j] = hsqrt(mt.d15 / ((1.0f - kSumRel) * (1.0f - 1e-6f))) * (1.0f + 1e-6f);
      dK[j] = K <= 0 ? -kInf : (sT[j] == kInf ? kInf : mt.dK);
      // d bound of the exact band with the block's largest candidate norm (a superset
      // of every candidate's own band)
      if (K <= 0) {
        dA[j] = -kInf;
      } else {
        const float f = sT[j] + ra[j] + rbmax;
        dA[j] = f * f * (1.0f + 1e-5f) + epsmax[j];
      }
    }
    // list lengths of reference (q, j), the same in the quadrant's 16 lanes;
    // a selected lane's slot is the count plus the selected lanes below it in the quadrant
    int cntA[4] = {0, 0, 0, 0}, cntB[4] = {0, 0, 0, 0};
    const uint64_t qmask = 0xffffull << (16 * q);
    auto append = [&](bool sel, int& cnt, uint16_t* lst, int capn, uint16_t li) {
      const uint64_t m = __ballot(sel);
      if (m) {
        const uint64_t mq = m & qmask;
        const int below = (int)__builtin_amdgcn_mbcnt_hi((uint32_t)(mq >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)mq, 0u));
        const int pos = cnt + below;
        if (sel && pos < capn) lst[pos] = li;
        cnt += __popcll(mq);
      }
    };
    // The approximate top-K criterion (d <= dK) selects a subset of the exact
    // band (dK <= dA), and its list is only used where the band's list
    // overflowed: the sweep collects the band (kB false), and a second sweep
    // collects the top-K list of such references (kB true, dsel = dK there).
    // Capped candidates: scan index <= limit.
    float dsel[4];
    auto tile = [&](auto b_tag, int cy, const floatx4& acc, float nbv) {
      constexpr bool kB = decltype(b_tag)::value;
      const int rr = cy - wy0;
      float d[4];
      dists(cy, acc, nbv, d);
      bool cand[4];
      bool any = false;
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        cand[j] = (d[j] <= dsel[j]) & (rr * ew[j] + oxc[j] <= limit[j]);
        any |= cand[j];
      }
      // wave-uniform: every lane of a quadrant takes part in the appends, so the
      // quadrant's list lengths stay equal in its 16 lanes
      if (__ballot(any)) {
        const float rb = kU * hsqrt(nbv) * (1.0f + 1e-6f) + 1e-5f;
        const uint16_t li = (uint16_t)((cy - uy0) * UW + (cxc - ux0));
#pragma unroll
        for (int j = 0; j < 4; ++j) {
          bool ok = cand[j] && d[j] <= asure[j];
          if (cand[j] && !ok && d[j] <= arej[j]) ok = classify(j, d[j], nbv, rb, cxc, cy);
          const int slot = wave * 16 + 4 * q + j;
          if (kB) append(ok, cntB[j], listb + slot * kListCapB, kListCapB, li);
          else append(ok, cntA[j], list + slot * kListCap, kListCap, li);
        }
      }
    };
#pragma unroll
    for (int j = 0; j < 4; ++j) dsel[j] = dA[j];
    // union rows through the last row holding a capped candidate
    int rlast = -1;
#pragma unroll
    for (int j = 0; j < 4; ++j)
      if (vj[j]) rlast = max(rlast, wy0 - uy0 + (limit[j] >= nrows * ew[j] ? nrows - 1 : limit[j] / ew[j]));
    const int R2 = wave_max_i(rlast) + 1;
    sweep(0, R2, column, [&](int cy, const floatx4& acc, float nbv) { tile(std::false_type{}, cy, acc, nbv); });
    bool needb = false;
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      dsel[j] = cntA[j] > kListCap ? dK[j] : -kInf;
      needb |= cntA[j] > kListCap;
    }
    if (__ballot(needb)) {
      sweep(0, R2, column, [&](int cy, const floatx4& acc, float nbv) { tile(std::true_type{}, cy, acc, nbv); });
    }
    if (n == 0) {
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        lcnt[wave * 16 + 4 * q + j] = cntA[j];
        lcntb[wave * 16 + 4 * q + j] = cntB[j];
      }
    }
  }
  __syncthreads();  // c1 is dead: its space holds the rank keys

  // ---- exact re-rank of the lists; entries of the wave's 16 references are
  // spread over the 64 lanes.  A reference uses its exact-band list when that
  // fit; otherwise the approximate top-K list (mixed-precision selection);
  // otherwise it goes to the plain matcher.
  int nms = 0, offs = 0, total = 0, usea = 1;
  if (wave_active) {
    const int m = lane & 15;
    const bool valid = (m & 3) < wnx && (m >> 2) < wny;
    const int ca = valid ? lcnt[wave * 16 + m] : 0, cb = valid ? lcntb[wave * 16 + m] : 0;
    usea = ca <= kListCap;
    nms = usea ? ca : (cb <= kListCapB ? cb : 0);
    for (int k = 0; k < 16; ++k) {
      const int v = __builtin_amdgcn_readlane(nms, k);
      offs += k < m ? v : 0;
      total += v;
    }
  }
  // entry e of the flattened lists -> (reference rm, index idx, list length nm,
  // exact-band list used); lane k < 16 holds reference k's offset and length
  auto entry = [&](int e, int& rm, int& idx, int& nm) -> int {
    rm = 0;
    int o = 0, ua = 1;
    nm = 0;
    for (int k = 0; k < 16; ++k) {
      const int ok = __builtin_amdgcn_readlane(offs, k), nk = __builtin_amdgcn_readlane(nms, k);
      if (e >= ok && e < ok + nk) {
        rm = k;
        o = ok;
        nm = nk;
        ua = __builtin_amdgcn_readlane(usea, k);
      }
    }
    idx = e - o;
    return ua;
  };
  if (wave_active) {
    for (int base = 0; base < total; base += 64) {
      const int e = base + lane;
      int rm, idx, nm;
      const int ua = entry(e, rm, idx, nm);
      if (e < total) {
        const int slot = wave * 16 + rm;
        const int li = ua ? list[slot * kListCap + idx] : listb[slot * kListCapB + idx];
        const int py = min(uy0 + li / UW, hmax), px = min(ux0 + li % UW, wmax);  // inside the image whatever the entry
        const int rrx = (wrx0 + (rm & 3)) * st, rry = (wry0 + (rm >> 2)) * st;
        const float t = exact_total(a.guide, W, rrx, rry, px, py);
        // rank key (distance, scan index): a sum of squares is >= +0, whose bit
        // patterns order as the values
        const int ex0 = max(rrx - half, 0), ex1 = min(rrx + half, wmax), ey0 = max(rry - half, 0);
        const uint32_t scan = (uint32_t)((py - ey0) * (ex1 - ex0 + 1) + (px - ex0));
        keys[slot * kListCap + idx] = ((uint64_t)__float_as_uint(t) << 32) | scan;
      }
    }
  }
  __syncthreads();
  if (wave_active) {
    for (int base = 0; base < total; base += 64) {
      const int e = base + lane;
      int rm, idx, nm;
      const int ua = entry(e, rm, idx, nm);
      if (e < total) {
        const int slot = wave * 16 + rm;
        const uint16_t* lst = ua ? list + slot * kListCap : listb + slot * kListCapB;
        const int li = lst[idx];
        const int py = min(uy0 + li / UW, hmax), px = min(ux0 + li % UW, wmax);  // inside the image whatever the entry
        // rank by (distance, scan order)
        const uint64_t* kk = keys + slot * kListCap;
        const uint64_t mine = kk[idx];
        int rank = 0;
        for (int k = 0; k < nm; ++k) rank += kk[k] < mine ? 1 : 0;
        const int g = meta[slot].g;
        const int ref_id = (wry0 + (rm >> 2)) * a.refs_x + wrx0 + (rm & 3);
        if (rank < g - 1) a.groups[(size_t)ref_id * 17 + 2 + rank] = (uint32_t)px | ((uint32_t)py << 16);
      }
    }
    if (lane < 16) {
      const int m = lane;
      if ((m & 3) < wnx && (m >> 2) < wny) {
        const int slot = wave * 16 + m;
        const int rrx = (wrx0 + (m & 3)) * st, rry = (wry0 + (m >> 2)) * st;
        const int ref_id = (wry0 + (m >> 2)) * a.refs_x + wrx0 + (m & 3);
        if (lcnt[slot] > kListCap && lcntb[slot] > kListCapB) {  // both lists overflowed: plain matcher
          const uint32_t at = atomicAdd(a.overflow_count, 1u);
          a.overflow_list[at] = (uint32_t)ref_id;
        } else {
          if (lcnt[slot] > kListCap) atomicAdd(a.overflow_count + 1, 1u);  // approximate selection
          a.groups[(size_t)ref_id * 17] = (uint32_t)meta[slot].g;
          a.groups[(size_t)ref_id * 17 + 1] = (uint32_t)rrx | ((uint32_t)rry << 16);
        }
      }
    }
  }
}

// ------------------------------------------------------------ collaborative filter
// Mixed-precision filter: one wave per (group, channel).  With the group's
// patches as rows, the 2-D DCT is one product with Kd = D (x) D (64 x 64, FP16),
// the Walsh-Hadamard transform across the group a product with the g x g
// Sylvester matrix / sqrt(g) (16 x 16, rows past g zero), then the hard
// threshold or Wiener shrinkage, the inverse Hadamard product, a transpose
// through LDS and the product with Kd^T; FP32 accumulation throughout.
// Aggregation is the FP32 form's fixed-point footprint (bm3d.rs:463-488).
//
// Before the FP16 conversion every patch of a group has the same level c (the
// mean of the reference patch's four centre pixels) subtracted.  A constant
// moves only stack element 0 (its DC coefficient, through the first Hadamard
// row), which the hard threshold keeps and the Wiener filter passes untouched,
// so c is added back after the inverse transforms; meanwhile the FP16 operands
// hold residuals of the local contrast instead of absolute levels.
typedef _Float16 half4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ long long to_fixed_m(float v) {
  const float s = v * 100000.0f;
  if (fabsf(s) < 2147483648.0f) return (long long)(int)s;  // truncation; the common case
  if (isnan(s)) return 0;
  if (s >= 9.2233720368547758e18f) return INT64_MAX;
  if (s <= -9.2233720368547758e18f) return INT64_MIN;
  return (long long)s;
}

__device__ __forceinline__ half4 to_h4(const floatx4& v) {
  half4 h = {(_Float16)v[0], (_Float16)v[1], (_Float16)v[2], (_Float16)v[3]};
  return h;
}

// Group level: mean of the reference patch's centre 2x2 pixels (wave-uniform).
__device__ __forceinline__ float group_level(const float* __restrict__ plane, int W, uint32_t p0) {
  const float* c = plane + (size_t)((p0 >> 16) + 3) * W + (p0 & 0xffffu) + 3;
  const float v = 0.25f * ((c[0] + c[1]) + (c[W] + c[W + 1]));
  return __int_as_float(__builtin_amdgcn_readfirstlane(__float_as_int(v)));
}

// A fragments of the group's patches: lane (patch n, q) holds rows q and 4 + q, minus the level c.
__device__ __forceinline__ void load_patch_rows(const float* __restrict__ plane, int W, bool act, int px, int py, int q,
                                                float c, half8& a0, half8& a1) {
  float v0[8], v1[8];
  if (act) {
    load8g(plane + (size_t)(py + q) * W + px, v0);
    load8g(plane + (size_t)(py + 4 + q) * W + px, v1);
  } else {
#pragma unroll
    for (int e = 0; e < 8; ++e) v0[e] = v1[e] = 0.0f;
  }
#pragma unroll
  for (int e = 0; e < 8; ++e) {
    a0[e] = (_Float16)(act ? v0[e] - c : 0.0f);
    a1[e] = (_Float16)(act ? v1[e] - c : 0.0f);
  }
}

// LDS: DCT fragments (16 KB), one 16 x 64 FP16 transpose per wave (2 KB each), the
// footprint's i64 numerators and u32 denominators, then the tile's group tables
// (read once, used by the three channels).  A workgroup's denominator
// sum at one pixel is at most 64 references x 16 patches x 1e5 < 2^32, so the
// 32-bit partial sums are exact.
// Workgroups of 8 waves for tiles of up to 4x4 reference patches, 16 waves for
// 8x8 tiles (one such workgroup per CU still holds 16 waves).
constexpr int filter_threads(int tile) { return tile >= 8 ? 1024 : 512; }
constexpr int filter_fixed_lds(int threads) { return 16384 + (threads / 64) * 2048; }

// kTrial only separates autotuner trial launches from real ones in kernel traces.
// kHaar: the Haar group transform instead of the reference's Walsh-Hadamard.
template <bool kStep1, bool kTrial, int FM, bool kHaar>
__global__ void __launch_bounds__(FM) bm3d_filter_mfma_kernel(Bm3dFilterArgs a, const half8* __restrict__ frags) {
  constexpr int kFilterFixedLds = filter_fixed_lds(FM);
  extern __shared__ __align__(16) unsigned char fsh[];
  const int tid = threadIdx.x, lane = tid & 63;
  const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
  const int q = lane >> 4, n = lane & 15;
  const int tiles_x = (a.refs_x + a.tile - 1) / a.tile;
  const int tx = blockIdx.x % tiles_x, ty = blockIdx.x / tiles_x;
  const int r0x = tx * a.tile, r0y = ty * a.tile;
  const int r1x = min(r0x + a.tile, a.refs_x), r1y = min(r0y + a.tile, a.refs_y);
  const int nrx = r1x - r0x, nref = nrx * (r1y - r0y);
  const int half_w = a.window / 2;
  const int fx0 = max(r0x * a.stride - half_w, 0), fy0 = max(r0y * a.stride - half_w, 0);
  const int fx1 = min((r1x - 1) * a.stride + half_w, a.width - 8) + 8;  // exclusive
  const int fy1 = min((r1y - 1) * a.stride + half_w, a.height - 8) + 8;
  const int fw = fx1 - fx0, fh = fy1 - fy0, fp = fw * fh;
  half8* lfrag = reinterpret_cast<half8*>(fsh);
  _Float16* tbuf = reinterpret_cast<_Float16*>(fsh + 16384) + wave * 16 * 64;
  unsigned long long* fnum = reinterpret_cast<unsigned long long*>(fsh + kFilterFixedLds);
  uint32_t* fden = reinterpret_cast<uint32_t*>(fnum + fp);
  uint32_t* lgrp = fden + fp;  // the tile's group tables, shared by the three channels
  // constant B fragments: forward (Kd^T) at [t][s], inverse (Kd) at [4 + t][s]
  for (int i = tid; i < 1024; i += FM) lfrag[i] = frags[i];
  for (int i = tid; i < nref * 17; i += FM) {
    const int r = i / 17;
    lgrp[i] = a.groups[(size_t)((r0y + r / nrx) * a.refs_x + r0x + r % nrx) * 17 + (i - r * 17)];
  }
#define KF(t, s) lfrag[((t) * 2 + (s)) * 64 + lane]
#define KI(t, s) lfrag[((4 + (t)) * 2 + (s)) * 64 + lane]

  for (int ch = 0; ch < 3; ++ch) {
    for (int k = tid; k < fp; k += FM) {
      fnum[k] = 0ull;
      fden[k] = 0u;
    }
    __syncthreads();
    const float ch_sigma = ch == 0 ? a.sigma : a.sigma * a.chroma_scale;
    const float* __restrict__ gplane = a.guide[ch];
    const float* __restrict__ nplane = a.noisy[ch];
    for (int r = wave; r < nref; r += FM / 64) {  // wave-uniform
      const uint32_t* grp = lgrp + r * 17;
      const int g = (int)grp[0];
      if (g == 0) continue;  // nothing to aggregate
      // patch n (A rows) and patches 4q + j (C rows) positions
      const bool actn = n < g;
      const uint32_t pn = actn ? grp[1 + n] : 0u;
      const int pxn = (int)(pn & 0xffffu), pyn = (int)(pn >> 16);
      int cbase[4];  // footprint index of patch 4q + j's top-left pixel
      bool cact[4];
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        cact[j] = 4 * q + j < g;
        const uint32_t pp = cact[j] ? grp[1 + 4 * q + j] : 0u;
        cbase[j] = ((int)(pp >> 16) - fy0) * fw + ((int)(pp & 0xffffu) - fx0);
      }
      // group-transform fragments (A of the 16x16x16 products): lane (row n, q)
      // holds T[n][4q .. 4q+3] (forward) and T[4q .. 4q+3][n] (inverse, the
      // transpose).  Walsh-Hadamard: the Sylvester matrix / sqrt(g), symmetric.
      // Haar (in-place lifting order): row 0 = 1/sqrt(g) everywhere; row k > 0
      // with lowest set bit h = +1/sqrt(2h) on [k - h, k), -1/sqrt(2h) on [k, k + h).
      half4 Hf, Hi;
      if (!kHaar) {
        const float hs = 1.0f / sqrtf((float)g);
#pragma unroll
        for (int e = 0; e < 4; ++e) {
          const int kk = 4 * q + e;
          const float v = (n < g && kk < g) ? ((__popc(n & kk) & 1) ? -hs : hs) : 0.0f;
          Hf[e] = (_Float16)v;
        }
        Hi = Hf;
      } else {
        auto tcoef = [&](int r, int c) -> float {
          if (r >= g || c >= g) return 0.0f;
          if (r == 0) return 1.0f / sqrtf((float)g);
          const int h = r & -r;
          const float v = 1.0f / sqrtf((float)(2 * h));
          return (c >= r - h && c < r) ? v : ((c >= r && c < r + h) ? -v : 0.0f);
        };
#pragma unroll
        for (int e = 0; e < 4; ++e) {
          Hf[e] = (_Float16)tcoef(n, 4 * q + e);
          Hi[e] = (_Float16)tcoef(4 * q + e, n);
        }
      }
      // forward: 2-D DCT (rows = patches) then Hadamard across the group
      const uint32_t p0 = grp[1];
      const float cg = group_level(gplane, a.width, p0);
      half8 A0, A1;
      load_patch_rows(gplane, a.width, actn, pxn, pyn, q, cg, A0, A1);
      floatx4 Z[4];
#pragma unroll
      for (int t = 0; t < 4; ++t) {
        floatx4 y = {0.0f, 0.0f, 0.0f, 0.0f};
        y = __builtin_amdgcn_mfma_f32_16x16x32_f16(A0, KF(t, 0), y, 0, 0, 0);
        y = __builtin_amdgcn_mfma_f32_16x16x32_f16(A1, KF(t, 1), y, 0, 0, 0);
        const floatx4 zero = {0.0f, 0.0f, 0.0f, 0.0f};
        Z[t] = __builtin_amdgcn_mfma_f32_16x16x16f16(Hf, to_h4(y), zero, 0, 0, 0);
      }
      float weight, level;
      if (kStep1) {
        level = cg;
        const float th = a.lambda * ch_sigma;
        int nz = 0;
#pragma unroll
        for (int t = 0; t < 4; ++t)
#pragma unroll
          for (int j = 0; j < 4; ++j) {
            const bool dc = t == 0 && n == 0 && q == 0 && j == 0;  // stack element 0 is kept
            const bool keep = dc || !(fabsf(Z[t][j]) < th);
            if (!keep) Z[t][j] = 0.0f;
            nz += (keep && (4 * q + j < g)) ? 1 : 0;
          }
        for (int off = 32; off > 0; off >>= 1) nz += __shfl_xor(nz, off, 64);
        weight = nz > 0 ? 1.0f / (float)nz : 1.0f;
      } else {
        level = group_level(nplane, a.width, p0);
        half8 N0, N1;
        load_patch_rows(nplane, a.width, actn, pxn, pyn, q, level, N0, N1);
        const float s2 = ch_sigma * ch_sigma;
        float wsum = 0.0f;
#pragma unroll
        for (int t = 0; t < 4; ++t) {
          floatx4 y = {0.0f, 0.0f, 0.0f, 0.0f};
          y = __builtin_amdgcn_mfma_f32_16x16x32_f16(N0, KF(t, 0), y, 0, 0, 0);
          y = __builtin_amdgcn_mfma_f32_16x16x32_f16(N1, KF(t, 1), y, 0, 0, 0);
          const floatx4 zero = {0.0f, 0.0f, 0.0f, 0.0f};
          floatx4 zn = __builtin_amdgcn_mfma_f32_16x16x16f16(Hf, to_h4(y), zero, 0, 0, 0);
#pragma unroll
          for (int j = 0; j < 4; ++j) {
            const bool dc = t == 0 && n == 0 && q == 0 && j == 0;
            const float e = Z[t][j] * Z[t][j];
            // hardware reciprocal (1 ulp): far below the FP16 operands' rounding
            const float coef = e * __builtin_amdgcn_rcpf(e + s2 + 1e-5f);
            if (!dc) {
              zn[j] *= coef;
              wsum += coef * coef;
            }
          }
          Z[t] = zn;
        }
        for (int off = 32; off > 0; off >>= 1) wsum += __shfl_xor(wsum, off, 64);
        const float run = 1.0f + wsum;
        weight = run > 0.0f ? 1.0f / run : 1.0f;
      }
      // inverse Hadamard, transpose to patch rows through LDS, inverse DCT.  The
      // 16 x 64 transpose buffer stores each row's 8-half blocks XOR-swizzled by
      // 2 (row / 4), so the four rows 4q + j one store writes fall in different
      // banks; the reads undo the swizzle (a block stays 8 contiguous halves).
#pragma unroll
      for (int t = 0; t < 4; ++t) {
        const floatx4 zero = {0.0f, 0.0f, 0.0f, 0.0f};
        const floatx4 y = __builtin_amdgcn_mfma_f32_16x16x16f16(Hi, to_h4(Z[t]), zero, 0, 0, 0);
        const int blk = ((2 * t + (n >> 3)) ^ (2 * q)) * 8 + (n & 7);
#pragma unroll
        for (int j = 0; j < 4; ++j) tbuf[(4 * q + j) * 64 + blk] = (_Float16)y[j];
      }
      __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
      __builtin_amdgcn_wave_barrier();
      __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
      const int rsw = 2 * (n >> 2);  // row n's swizzle
      const half8 B0 = load8(tbuf + n * 64 + 8 * (q ^ rsw)), B1 = load8(tbuf + n * 64 + 8 * ((4 + q) ^ rsw));
      __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
      __builtin_amdgcn_wave_barrier();
      __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
#pragma unroll
      for (int t = 0; t < 4; ++t) {
        floatx4 pv = {0.0f, 0.0f, 0.0f, 0.0f};
        pv = __builtin_amdgcn_mfma_f32_16x16x32_f16(B0, KI(t, 0), pv, 0, 0, 0);
        pv = __builtin_amdgcn_mfma_f32_16x16x32_f16(B1, KI(t, 1), pv, 0, 0, 0);
        // lane (pixel 16t + n, q), register j -> patch 4q + j
        const int pix = 16 * t + n, poff = (pix >> 3) * fw + (pix & 7);
        const float kw = a.kaiser[pix] * weight;
        const long long kwf = (long long)(int)(kw * 100000.0f);  // kw in [0, 1]: exact 32-bit truncation
#pragma unroll
        for (int j = 0; j < 4; ++j) {
          if (!cact[j]) continue;
          const int fi = cbase[j] + poff;
          atomicAdd(&fnum[fi], (unsigned long long)to_fixed_m((pv[j] + level) * kw));
          atomicAdd(&fden[fi], (uint32_t)kwf);
        }
      }
    }
    __syncthreads();
    unsigned long long* gnum = a.num + (size_t)ch * a.width * a.height;
    unsigned long long* gden = a.den + (size_t)ch * a.width * a.height;
    for (int k = tid; k < fp; k += FM) {
      const unsigned long long nv = fnum[k], dv = fden[k];
      if (nv | dv) {
        const size_t gi = (size_t)(fy0 + k / fw) * a.width + fx0 + k % fw;
        if (nv) atomicAdd(&gnum[gi], nv);
        if (dv) atomicAdd(&gden[gi], dv);
      }
    }
    __syncthreads();
  }
#undef KF
#undef KI
}

__global__ void fill_u32_kernel(uint32_t* p, size_t n, uint32_t v) {
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) p[i] = v;
}

}  // namespace

Bm3dMfmaArgs bm3d_mfma_layout(int width, int height, int stride, int window) {
  Bm3dMfmaArgs a{};
  const int half = window / 2;
  a.spx = std::min(7 * stride + 2 * half + 8, width);
  a.spy = std::min(7 * stride + 2 * half + 8, height);
  a.pitch = (a.spx + 1 + 1) & ~1;  // even, with room for reads one column past the staged span
  auto up16 = [](size_t v) { return (v + 15) & ~(size_t)15; };
  const size_t copy = (size_t)3 * a.spy * a.pitch * 2;
  const size_t tot = (size_t)XW * 16 * kListCap * 8;  // rank keys of the exact phase
  size_t off = 0;
  a.off_w16 = 0;
  off = up16(copy);
  a.off_w16b = (int)off;  // second copy; also the setup row sums and the exact-phase rank keys
  off = up16(off + std::max(copy, tot));
  a.off_nb = (int)off;
  off = up16(off + (size_t)(a.spy - 7) * (a.spx - 7) * 4);
  a.off_list = (int)off;
  off = up16(off + (size_t)XW * 16 * (kListCap + kListCapB) * 2);
  a.off_cnt = (int)off;
  off = up16(off + XW * 16 * 4 * 2);
  a.off_meta = (int)off;
  off = up16(off + XW * 16 * sizeof(RefMeta));
  a.off_red = (int)off;
  off = up16(off + 4 * XW * 4);
  a.smem = (int)off;
  // the union's rows are tracked in a 128-bit map; list entries index the
  // wave's candidate union in 16 bits
  const int uw = 3 * stride + 2 * half + 1;
  a.max_tiles_ok = uw <= 128 && (size_t)uw * uw < 65536;
  return a;
}

// B-operand fragments of the DCT products in lane order: [fwd|inv][N-tile t][K-step s][lane][8 halves],
// fwd: Kd^T (pixel x coefficient), inv: Kd (coefficient x pixel), Kd[(j,k)][(n,m)] = dct[j][n] dct[k][m].
void bm3d_filter_mfma_fragments(const float* dct, uint16_t* out) {
  for (int which = 0; which < 2; ++which)
    for (int t = 0; t < 4; ++t)
      for (int s = 0; s < 2; ++s)
        for (int l = 0; l < 64; ++l)
          for (int e = 0; e < 8; ++e) {
            const int nn = l & 15, qq = l >> 4;
            const int kidx = 32 * s + 8 * qq + e, col = 16 * t + nn;
            const int coef = which == 0 ? col : kidx, pix = which == 0 ? kidx : col;
            const double v = (double)dct[(coef >> 3) * 8 + (pix >> 3)] * (double)dct[(coef & 7) * 8 + (pix & 7)];
            const _Float16 h = (_Float16)(float)v;
            out[((((which * 4 + t) * 2 + s) * 64 + l) * 8) + e] = __builtin_bit_cast(uint16_t, h);
          }
}

size_t bm3d_filter_mfma_smem(int footprint_pixels, int tile) {
  return filter_fixed_lds(filter_threads(tile)) +
         (size_t)footprint_pixels * (sizeof(unsigned long long) + sizeof(uint32_t)) +
         (size_t)tile * tile * 17 * sizeof(uint32_t);
}

hipError_t bm3d_filter_mfma_launch(const Bm3dFilterArgs& a, bool step1, int nblocks, size_t smem, const void* frags,
                                   hipStream_t s, bool trial) {
  const half8* f = reinterpret_cast<const half8*>(frags);
  const int threads = filter_threads(a.tile);
#define LAUNCH_FILTER(S1, TR, TH) \
  hipLaunchKernelGGL((bm3d_filter_mfma_kernel<S1, TR, TH, false>), dim3(nblocks), dim3(TH), smem, s, a, f)
#define LAUNCH_HAAR(S1, TH) \
  hipLaunchKernelGGL((bm3d_filter_mfma_kernel<S1, false, TH, true>), dim3(nblocks), dim3(TH), smem, s, a, f)
  if (a.haar) {  // (trial launches share the real kernels)
    if (threads == 1024) {
      if (step1) LAUNCH_HAAR(true, 1024);
      else LAUNCH_HAAR(false, 1024);
    } else {
      if (step1) LAUNCH_HAAR(true, 512);
      else LAUNCH_HAAR(false, 512);
    }
  } else if (threads == 1024) {
    if (trial) {
      if (step1) LAUNCH_FILTER(true, true, 1024);
      else LAUNCH_FILTER(false, true, 1024);
    } else {
      if (step1) LAUNCH_FILTER(true, false, 1024);
      else LAUNCH_FILTER(false, false, 1024);
    }
  } else {
    if (trial) {
      if (step1) LAUNCH_FILTER(true, true, 512);
      else LAUNCH_FILTER(false, true, 512);
    } else {
      if (step1) LAUNCH_FILTER(true, false, 512);
      else LAUNCH_FILTER(false, false, 512);
    }
  }
#undef LAUNCH_FILTER
#undef LAUNCH_HAAR
  return hipGetLastError();
}

hipError_t bm3d_match_mfma_launch(const Bm3dMfmaArgs& a, hipStream_t s) {
  hipLaunchKernelGGL(fill_u32_kernel, dim3(1), dim3(64), 0, s, a.overflow_count, (size_t)2, 0u);
  if (a.pitch == 68)  // 39x39 windows at stride 3
    hipLaunchKernelGGL(bm3d_match_mfma_kernel<34>, dim3(a.n_blocks), dim3(XT), (size_t)a.smem, s, a);
  else
    hipLaunchKernelGGL(bm3d_match_mfma_kernel<0>, dim3(a.n_blocks), dim3(XT), (size_t)a.smem, s, a);
  return hipGetLastError();
}

}  // namespace imops
