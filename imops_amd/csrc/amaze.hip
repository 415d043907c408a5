// AMaZE Bayer demosaic for gfx950.
//
// Behaviour follows pichromatic/src/demosaic.rs:358-1379 exactly: the image
// is processed in 128x128 tiles on the grid top,left = -16 + 96k with the
// reference's border fill, and every tile runs the same sequence of passes.
// Three passes of the reference update a plane in place while reading
// neighbours that the same pass has already rewritten (hcd/vcd variance
// selection, the hvwt diagonal smoothing and the pmwt diagonal smoothing), so
// the result depends on the tile a pixel is computed in; the kernel keeps that
// tile structure instead of computing the image as one plane.
//
// Mapping: one 1024-thread workgroup per tile (persistent over tiles; 16 waves
// per CU hide the scratch-slot and barrier latency better than 8).  The
// CFA tile, hvwt, pmwt and the two Nyquist masks live in LDS (144 KiB); the
// other planes of the tile live in a per-workgroup global scratch slot that
// stays L2/MALL resident while the tile is processed.  Recurrences run as
// register-carried chains (hcd along rows, vcd along columns) or row by row
// in a single wave (hvwt, pmwt) while the other waves do independent passes.
//
// All arithmetic is IEEE f32 in the reference's operation order: the file is
// compiled with -ffp-contract=off and correctly rounded division.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <mutex>

#include "amaze.hpp"

namespace imops {
namespace {

constexpr int TS = 128, TSH = 64;
constexpr int V1 = 128, V2 = 256, V3 = 384;
constexpr int P1 = 127, P2 = 254, P3 = 381;
constexpr int M1 = 129, M2 = 258, M3 = 387;
constexpr int NT = 1024;  // threads per workgroup

constexpr float eps = 1e-5f, epssq = 1e-10f, arthresh = 0.75f;
constexpr float clip_pt = 1.0f, clip_pt8 = 1.0f * 0.8f;

// Offsets (in floats) of the planes inside one scratch slot.
enum : int {
  S_RGBGREEN = 0,
  S_VCD = 1 * 16384,
  S_HCD = 2 * 16384,
  S_VCDALT = 3 * 16384,
  S_HCDALT = 4 * 16384,
  S_CDDIFFSQ = 5 * 16384,
  S_DGINTV = 6 * 16384,
  S_DGINTH = 7 * 16384,
  S_HALF = 8 * 16384,
  S_DGRB0 = S_HALF + 0 * 8192,
  S_DGRB1 = S_HALF + 1 * 8192,
  S_DELP = S_HALF + 2 * 8192,
  S_DELM = S_HALF + 3 * 8192,
  S_RBINT = S_HALF + 4 * 8192,
  S_DGRB2H = S_HALF + 5 * 8192,
  S_DGRB2V = S_HALF + 6 * 8192,
  S_DGRBSQ1M = S_HALF + 7 * 8192,
  S_DGRBSQ1P = S_HALF + 8 * 8192,
  S_RBM = S_HALF + 9 * 8192,
  S_RBP = S_HALF + 10 * 8192,
  S_SLOT_FLOATS = S_HALF + 11 * 8192,
};
static_assert(S_SLOT_FLOATS == kAmazeSlotFloats, "slot layout");

struct Tile {
  int top, left, rr1, cc1, rrmin, ccmin, rrmax, ccmax;
};

__device__ __forceinline__ float sqr(float x) { return x * x; }
__device__ __forceinline__ float lim(float x, float lo, float hi) { return fmaxf(lo, fminf(x, hi)); }
__device__ __forceinline__ float ulim(float x, float y, float z) { return y < z ? lim(x, y, z) : lim(x, z, y); }
__device__ __forceinline__ float clampnan(float x, float m, float mx) {
  return isnan(x) ? (m + mx) * 0.5f : (x < m ? m : (x > mx ? mx : x));
}

struct Ctx {
  const AmazeArgs& a;
  Tile t;
  uint32_t cfa4;     // the 2x2 pattern, one byte per site ((r & 1) * 2 + (c & 1)), in a register
  float* s;          // scratch slot
  float* cfa;        // LDS 128x128
  float* hvwt;       // LDS 128x64
  float* pmwt;       // LDS 128x64
  uint8_t* nyq;      // LDS 128x64
  uint8_t* nyq2;     // LDS 128x64
  // (the kernel-argument bytes indexed per lane would be a memory load per call)
  __device__ int color_at(int r, int c) const { return (int)((cfa4 >> ((((r & 1) << 1) | (c & 1)) << 3)) & 0xffu); }
  __device__ int par(int rr) const { return color_at(t.top + rr, t.left + 2) & 1; }
  __device__ float get_in(int r, int c) const {
    r = r < 0 ? 0 : (r > a.height - 1 ? a.height - 1 : r);
    c = c < 0 ? 0 : (c > a.width - 1 ? a.width - 1 : c);
    if (a.in16)  // raw ingest fused into the tile fill: the normalisation of crop_and_normalize
      return ((float)a.in16[(a.in16_y + (size_t)r) * a.in16_pitch + a.in16_x + (size_t)c] - a.black) / a.range;
    return a.in[(size_t)r * a.width + c];
  }
  __device__ float* P(int off) const { return s + off; }
};

// Source row of tile row rr (border fill demosaic.rs:500-607); -1 = zero.
__device__ __forceinline__ int src_row(const Ctx& c, int rr, bool* upper, bool* lower) {
  const Tile& t = c.t;
  *upper = *lower = false;
  if (rr >= t.rr1) return INT32_MIN;
  if (rr < t.rrmin) { *upper = true; return rr; }
  if (rr >= t.rrmax) { *lower = true; return rr - t.rrmax; }
  return rr + t.top;
}

__device__ void pass_fill(Ctx& c, int tid) {
  const Tile& t = c.t;
  const int H = c.a.height, W = c.a.width;
  float* rgbgreen = c.P(S_RGBGREEN);
  for (int idx = tid; idx < 16384; idx += NT) {
    const int rr = idx >> 7, cc = idx & 127;
    float v = 0.0f;
    if (rr < t.rr1 && cc < t.cc1) {
      const bool up = rr < t.rrmin, lo = rr >= t.rrmax;
      const bool le = cc < t.ccmin, ri = cc >= t.ccmax;
      const int br = lo ? rr - t.rrmax : rr;  // index inside the 16-wide border band
      const int bc = ri ? cc - t.ccmax : cc;
      int row, col;
      if ((up || lo) && (le || ri)) {  // corners
        row = up ? 32 - br : H - br - 2;
        col = le ? 32 - bc : W - bc - 2;
      } else {
        row = up ? 32 - br + t.top : (lo ? H - br - 2 : rr + t.top);
        col = le ? 32 - bc + t.left : (ri ? W - bc - 2 : cc + t.left);
      }
      v = c.get_in(row, col);
    }
    c.cfa[idx] = v;
    rgbgreen[idx] = v;
  }
}

// dirwts[0] and dirwts[1] of demosaic.rs:612-622 at site i, formed from the
// LDS CFA tile where a pass needs them instead of being kept as two planes of
// the scratch slot (every read is at rows and columns 2 .. rr1-3 / cc1-3, the
// reference's computed range).
__device__ __forceinline__ float dirw0(const float* cfa, int i) {
  const float delv = fabsf(cfa[i + V1] - cfa[i - V1]);
  return eps + fabsf(cfa[i + V2] - cfa[i]) + fabsf(cfa[i] - cfa[i - V2]) + delv;
}
__device__ __forceinline__ float dirw1(const float* cfa, int i) {
  const float delh = fabsf(cfa[i + 1] - cfa[i - 1]);
  return eps + fabsf(cfa[i + 2] - cfa[i]) + fabsf(cfa[i] - cfa[i - 2]) + delh;
}

// demosaic.rs:626-698
__device__ void pass_colour_diff(Ctx& c, int tid) {
  const Tile& t = c.t;
  const float* cfa = c.cfa;
  float* vcd = c.P(S_VCD);
  float* hcd = c.P(S_HCD);
  float* vcdalt = c.P(S_VCDALT);
  float* hcdalt = c.P(S_HCDALT);
  float* dgintv = c.P(S_DGINTV);
  float* dginth = c.P(S_DGINTH);
  for (int i = tid; i < 16384; i += NT) {
    const int rr = i >> 7, cc = i & 127;
    if (!(rr >= 4 && rr < t.rr1 - 4 && cc >= 4 && cc < t.cc1 - 4)) {
      vcd[i] = 0.0f; hcd[i] = 0.0f; vcdalt[i] = 0.0f; hcdalt[i] = 0.0f; dgintv[i] = 0.0f; dginth[i] = 0.0f;
      continue;
    }
    const bool fcswitch = (((c.color_at(t.top + rr, t.left + 4) & 1) != 0) != (((cc - 4) & 1) != 0));
    const float ci = cfa[i];
    const float dw0m2 = dirw0(cfa, i - V2), dw0 = dirw0(cfa, i), dw0p2 = dirw0(cfa, i + V2);
    const float dw1m2 = dirw1(cfa, i - 2), dw1 = dirw1(cfa, i), dw1p2 = dirw1(cfa, i + 2);
    float cru = cfa[i - V1] * (dw0m2 + dw0) / fmaxf(dw0m2 * (eps + ci) + dw0 * (eps + cfa[i - V2]), 1e-10f);
    float crd = cfa[i + V1] * (dw0p2 + dw0) / fmaxf(dw0p2 * (eps + ci) + dw0 * (eps + cfa[i + V2]), 1e-10f);
    float crl = cfa[i - 1] * (dw1m2 + dw1) / fmaxf(dw1m2 * (eps + ci) + dw1 * (eps + cfa[i - 2]), 1e-10f);
    float crr = cfa[i + 1] * (dw1p2 + dw1) / fmaxf(dw1p2 * (eps + ci) + dw1 * (eps + cfa[i + 2]), 1e-10f);
    float guha = cfa[i - V1] + (ci - cfa[i - V2]) * 0.5f;
    float gdha = cfa[i + V1] + (ci - cfa[i + V2]) * 0.5f;
    float glha = cfa[i - 1] + (ci - cfa[i - 2]) * 0.5f;
    float grha = cfa[i + 1] + (ci - cfa[i + 2]) * 0.5f;
    float guar = fabsf(1.0f - cru) < arthresh ? ci * cru : guha;
    float gdar = fabsf(1.0f - crd) < arthresh ? ci * crd : gdha;
    float glar = fabsf(1.0f - crl) < arthresh ? ci * crl : glha;
    float grar = fabsf(1.0f - crr) < arthresh ? ci * crr : grha;
    const float dw1l = dirw1(cfa, i - 1), dw1r = dirw1(cfa, i + 1);
    const float dw0u = dirw0(cfa, i - V1), dw0d = dirw0(cfa, i + V1);
    float hwt = dw1l / fmaxf(dw1l + dw1r, 1e-10f);
    float vwt = dw0u / fmaxf(dw0d + dw0u, 1e-10f);
    float gintvha = vwt * gdha + (1.0f - vwt) * guha;
    float ginthha = hwt * grha + (1.0f - hwt) * glha;
    float vcd_v, hcd_v, vcdalt_v, hcdalt_v;
    if (fcswitch) {
      vcd_v = ci - (vwt * gdar + (1.0f - vwt) * guar);
      hcd_v = ci - (hwt * grar + (1.0f - hwt) * glar);
      vcdalt_v = ci - gintvha;
      hcdalt_v = ci - ginthha;
    } else {
      vcd_v = (vwt * gdar + (1.0f - vwt) * guar) - ci;
      hcd_v = (hwt * grar + (1.0f - hwt) * glar) - ci;
      vcdalt_v = gintvha - ci;
      hcdalt_v = ginthha - ci;
    }
    float gu = guar, gd = gdar, gl = glar, gr = grar;
    if (ci > clip_pt8 || gintvha > clip_pt8 || ginthha > clip_pt8) {
      vcd_v = vcdalt_v;
      hcd_v = hcdalt_v;
      gu = guha; gd = gdha; gl = glha; gr = grha;
    }
    vcd[i] = vcd_v;
    hcd[i] = hcd_v;
    vcdalt[i] = vcdalt_v;
    hcdalt[i] = hcdalt_v;
    dgintv[i] = fminf(sqr(guha - gdha), sqr(gu - gd));
    dginth[i] = fminf(sqr(glha - grha), sqr(gl - gr));
  }
}

// One step of the variance selection + clamp of demosaic.rs:700-786 for one
// direction.  `prev` is the already-updated value two pixels back along the
// chain, `cur`/`next` are the values written by pass_colour_diff.
// Straight-line form of the reference's nested branches: at a G site
// (gsite) the reference works with gint = -v + ci and the clamp target
// -ulim(gint) + ci, elsewhere with gint = v + ci and ulim(gint) - ci; both
// sides' expressions are formed on the selected operands exactly as the
// reference writes them (one division, whichever side the lane is on), and
// the reference's branch conditions pick the result.  A chain step is a
// serial dependence, so the branch-free form (no divergent paths, no
// per-branch issue gaps) is what sets the pass time.
__device__ __forceinline__ float cd_update(float prev, float cur, float next, float aprev, float acur, float anext,
                                           bool gsite, float ci, float nm, float np) {
  float var = 3.0f * (sqr(prev) + sqr(cur) + sqr(next)) - sqr(prev + cur + next);
  float avar = 3.0f * (sqr(aprev) + sqr(acur) + sqr(anext)) - sqr(aprev + acur + anext);
  const float v = avar < var ? acur : cur;
  const float gint = gsite ? -v + ci : v + ci;
  const float u = ulim(gint, nm, np);
  const float target = gsite ? -u + ci : u - ci;  // the clamp value
  const float q = 3.0f * v / (eps + gint + ci);
  const float w = gsite ? 1.0f - q : 1.0f + q;
  const float blend = w * v + (1.0f - w) * target;
  const bool side = gsite ? v > 0.0f : v < 0.0f;                       // the reference's first test
  const bool hard = gsite ? 3.0f * v > (gint + ci) : 3.0f * v < -(gint + ci);  // its second
  float r = side ? (hard ? target : blend) : v;
  if (gint > clip_pt) r = target;
  return r;
}

// hcd chains along rows (threads 0..255) and vcd chains along columns
// (threads 256..511); each thread walks both parities of its line.  A step
// reads the not-yet-updated values two pixels ahead (`next`), which no step
// of the chain has written yet (a step writes only its own pixel), so they
// are loaded a block of kAhead steps early: the loads of block k + 1 are in
// flight while block k runs, instead of one slot round trip per step.
constexpr int kAhead = 4;

// delhvsqsum of demosaic.rs:612-622 at site p (inside rows and columns
// 2 .. rr1-3 / cc1-3, where the reference computes it), from the LDS CFA tile.
__device__ __forceinline__ float delhvsq(const float* cfa, int p) {
  const float delh = fabsf(cfa[p + 1] - cfa[p - 1]);
  const float delv = fabsf(cfa[p + V1] - cfa[p - V1]);
  return sqr(delh) + sqr(delv);
}

// The Nyquist test's gradient term (demosaic.rs:830-848, the second sum) at
// site i (its 5x5 neighbourhood lies inside delhvsqsum's computed range at
// every site pass_hvwt_nyq evaluates); it reads only the CFA tile, so it is
// formed while the chains run and delhvsqsum is never stored.
__device__ __forceinline__ float nyq_grad_term(const float* cfa, int i) {
  auto dq = [cfa](int p) { return delhvsq(cfa, p); };
  const float nyqthresh = 0.5f;
  const float gg0 = nyqthresh * 0.07384411893421103f, gg1 = nyqthresh * 0.06207511968171489f,
              gg2 = nyqthresh * 0.0521818194747806f, gg3 = nyqthresh * 0.03687419286733595f,
              gg4 = nyqthresh * 0.03099732204057846f, gg5 = nyqthresh * 0.018413194161458882f;
  return gg0 * dq(i) + gg1 * (dq(i - V1) + dq(i + 1) + dq(i - 1) + dq(i + V1)) +
         gg2 * (dq(i - M1) + dq(i - P1) + dq(i + P1) + dq(i + M1)) +
         gg3 * (dq(i - V2) + dq(i - 2) + dq(i + 2) + dq(i + V2)) +
         gg4 * (dq(i - V2 - 1) + dq(i - V2 + 1) + dq(i - TS - 2) + dq(i - TS + 2) + dq(i + TS - 2) + dq(i + TS + 2) +
                dq(i + V2 - 1) + dq(i + V2 + 1)) +
         gg5 * (dq(i - M2) + dq(i - P2) + dq(i + P2) + dq(i + M2));
}

// Threads 512..1023 while the chains run: nyq_grad_term of every site
// pass_hvwt_nyq evaluates, into the pmwt LDS plane (unused until pass_diag).
__device__ void pass_nyq_grad(Ctx& c, int tid) {
  const Tile& t = c.t;
  for (int h = tid - 512; h < 8192; h += 512) {
    const int rr = h >> 6, j = h & 63;
    if (rr < 6 || rr >= t.rr1 - 6) continue;
    const int p = c.par(rr);
    const int cc = 2 * j + p;
    if (cc < 6 + p || cc >= t.cc1 - 6) continue;
    c.pmwt[h] = nyq_grad_term(c.cfa, rr * TS + cc);
  }
}

__device__ void pass_cd_chains(Ctx& c, int tid) {
  const Tile& t = c.t;
  const float* cfa = c.cfa;
  if (tid >= 512) {  // 128 lines x 2 parities per direction; the rest form the Nyquist gradient term
    pass_nyq_grad(c, tid);
    return;
  }
  const bool horiz = tid < 256;
  const int line = (tid & 255) >> 1, parity = tid & 1;
  float* cd = c.P(horiz ? S_HCD : S_VCD);
  const float* alt = c.P(horiz ? S_HCDALT : S_VCDALT);
  const int lo = 4, hi = horiz ? t.cc1 - 4 : t.rr1 - 4;          // along the chain
  const int lline = 4 + line, hline = horiz ? t.rr1 - 4 : t.cc1 - 4;  // the line index
  if (lline >= hline) return;
  const int step = horiz ? 1 : V1;   // index distance of one pixel along the chain
  const int base = horiz ? lline * TS : lline;
  const int pos0 = lo + parity;
  if (pos0 >= hi) return;
  const int i0 = base + pos0 * step;
  // the colour is the same at every pixel of a chain (it steps by two pixels)
  const bool gsite = (c.color_at(t.top + (horiz ? lline : pos0), t.left + (horiz ? pos0 : lline)) & 1) != 0;
  float prev = cd[i0 - 2 * step];
  float aprev = alt[i0 - 2 * step];
  float cur = cd[i0], acur = alt[i0];
  // nx[k], anx[k]: `next` of step s0 + k (the values at pixel i0 + 2 step (s0 + k + 1))
  float nx[kAhead], anx[kAhead];
#pragma unroll
  for (int k = 0; k < kAhead; ++k) {
    nx[k] = anx[k] = 0.0f;
    if (pos0 + 2 * k < hi) {
      nx[k] = cd[i0 + 2 * step * (k + 1)];
      anx[k] = alt[i0 + 2 * step * (k + 1)];
    }
  }
  for (int s0 = 0; pos0 + 2 * s0 < hi; s0 += kAhead) {
    float ny[kAhead], any[kAhead];
#pragma unroll
    for (int k = 0; k < kAhead; ++k) {
      ny[k] = any[k] = 0.0f;
      const int sk = s0 + kAhead + k;
      if (pos0 + 2 * sk < hi) {
        ny[k] = cd[i0 + 2 * step * (sk + 1)];
        any[k] = alt[i0 + 2 * step * (sk + 1)];
      }
    }
    // the block's CFA values: cf[m] = cfa at pixel i0 + (2 s0 - 1 + m) step (read-only here);
    // positions past the chain's end are clamped into the tile and never used
    float cf[2 * kAhead + 1];
#pragma unroll
    for (int m = 0; m < 2 * kAhead + 1; ++m) cf[m] = cfa[min(i0 + (2 * s0 - 1 + m) * step, 16383)];
#pragma unroll
    for (int k = 0; k < kAhead; ++k) {
      const int pos = pos0 + 2 * (s0 + k);
      if (pos >= hi) break;
      const int i = i0 + 2 * step * (s0 + k);
      const float next = nx[k], anext = anx[k];
      const float v = cd_update(prev, cur, next, aprev, acur, anext, gsite, cf[2 * k + 1], cf[2 * k], cf[2 * k + 2]);
      cd[i] = v;
      prev = v; aprev = acur;
      cur = next; acur = anext;
    }
#pragma unroll
    for (int k = 0; k < kAhead; ++k) {
      nx[k] = ny[k];
      anx[k] = any[k];
    }
  }
}

__device__ void pass_cddiffsq(Ctx& c, int tid) {
  const Tile& t = c.t;
  const float* vcd = c.P(S_VCD);
  const float* hcd = c.P(S_HCD);
  float* cddiffsq = c.P(S_CDDIFFSQ);
  for (int i = tid; i < 16384; i += NT) {
    const int rr = i >> 7, cc = i & 127;
    float v = 0.0f;
    if (rr >= 4 && rr < t.rr1 - 4 && cc >= 4 && cc < t.cc1 - 4 && (c.color_at(t.top + rr, t.left + cc) & 1) == 0)
      v = sqr(vcd[i] - hcd[i]);
    cddiffsq[i] = v;
  }
}

// demosaic.rs:788-828 (hvwt) fused with the Nyquist test of 830-868
// (nyqutest only feeds the mask and its bounding box).
__device__ void pass_hvwt_nyq(Ctx& c, int tid, int* bbox) {
  const Tile& t = c.t;
  const float* vcd = c.P(S_VCD);
  const float* hcd = c.P(S_HCD);
  const float* cfa = c.cfa;
  const float* dgintv = c.P(S_DGINTV);
  const float* dginth = c.P(S_DGINTH);
  const float* cddiffsq = c.P(S_CDDIFFSQ);
  const float gaussodd0 = 0.14659727707323927f, gaussodd1 = 0.103592713382435f, gaussodd2 = 0.0732036125103057f,
              gaussodd3 = 0.0365543548389495f;
  int minr = INT32_MAX, maxr = -1, minc = INT32_MAX, maxc = -1;
  for (int h = tid; h < 8192; h += NT) {
    const int rr = h >> 6, j = h & 63;
    float hv = 0.0f;
    uint8_t ny = 0;
    if (rr >= 6 && rr < t.rr1 - 6) {
      const int p = c.par(rr);
      const int cc = 2 * j + p;
      if (cc >= 6 + p && cc < t.cc1 - 6) {
        const int i = rr * TS + cc;
        float uave = vcd[i] + vcd[i - V1] + vcd[i - V2] + vcd[i - V3];
        float dave = vcd[i] + vcd[i + V1] + vcd[i + V2] + vcd[i + V3];
        float lave = hcd[i] + hcd[i - 1] + hcd[i - 2] + hcd[i - 3];
        float rave = hcd[i] + hcd[i + 1] + hcd[i + 2] + hcd[i + 3];
        float dgrbvvaru = sqr(vcd[i] - uave) + sqr(vcd[i - V1] - uave) + sqr(vcd[i - V2] - uave) + sqr(vcd[i - V3] - uave);
        float dgrbvvard = sqr(vcd[i] - dave) + sqr(vcd[i + V1] - dave) + sqr(vcd[i + V2] - dave) + sqr(vcd[i + V3] - dave);
        float dgrbhvarl = sqr(hcd[i] - lave) + sqr(hcd[i - 1] - lave) + sqr(hcd[i - 2] - lave) + sqr(hcd[i - 3] - lave);
        float dgrbhvarr = sqr(hcd[i] - rave) + sqr(hcd[i + 1] - rave) + sqr(hcd[i + 2] - rave) + sqr(hcd[i + 3] - rave);
        const float dw1l = dirw1(cfa, i - 1), dw1r = dirw1(cfa, i + 1);
        const float dw0u = dirw0(cfa, i - V1), dw0d = dirw0(cfa, i + V1);
        float hwt = dw1l / fmaxf(dw1l + dw1r, 1e-10f);
        float vwt = dw0u / fmaxf(dw0d + dw0u, 1e-10f);
        float vcdvar = epssq + vwt * dgrbvvard + (1.0f - vwt) * dgrbvvaru;
        float hcdvar = epssq + hwt * dgrbhvarr + (1.0f - hwt) * dgrbhvarl;
        float dgrbvvaru1 = dgintv[i] + dgintv[i - V1] + dgintv[i - V2];
        float dgrbvvard1 = dgintv[i] + dgintv[i + V1] + dgintv[i + V2];
        float dgrbhvarl1 = dginth[i] + dginth[i - 1] + dginth[i - 2];
        float dgrbhvarr1 = dginth[i] + dginth[i + 1] + dginth[i + 2];
        float vcdvar1 = epssq + vwt * dgrbvvard1 + (1.0f - vwt) * dgrbvvaru1;
        float hcdvar1 = epssq + hwt * dgrbhvarr1 + (1.0f - hwt) * dgrbhvarl1;
        float varwt = hcdvar / fmaxf(vcdvar + hcdvar, 1e-10f);
        float diffwt = hcdvar1 / fmaxf(vcdvar1 + hcdvar1, 1e-10f);
        hv = ((0.5f - varwt) * (0.5f - diffwt) > 0.0f && fabsf(0.5f - diffwt) < fabsf(0.5f - varwt)) ? varwt : diffwt;
        float nt = (gaussodd0 * cddiffsq[i] +
                    gaussodd1 * (cddiffsq[i - M1] + cddiffsq[i - P1] + cddiffsq[i + P1] + cddiffsq[i + M1]) +
                    gaussodd2 * (cddiffsq[i - V2] + cddiffsq[i - 2] + cddiffsq[i + 2] + cddiffsq[i + V2]) +
                    gaussodd3 * (cddiffsq[i - M2] + cddiffsq[i - P2] + cddiffsq[i + P2] + cddiffsq[i + M2])) -
                   c.pmwt[h];  // nyq_grad_term, formed while the chains ran
        if (nt > 0.0f) {
          ny = 1;
          minr = min(minr, rr); maxr = max(maxr, rr);
          minc = min(minc, cc); maxc = max(maxc, cc);
        }
      }
    }
    c.hvwt[h] = hv;
    c.nyq[h] = ny;
  }
  if (maxr >= 0) {
    atomicMin(&bbox[0], minr);
    atomicMax(&bbox[1], maxr);
    atomicMin(&bbox[2], minc);
    atomicMax(&bbox[3], maxc);
  }
}

// nyquist2 (demosaic.rs:888-910); bb holds the adjusted bounding box.
__device__ void pass_nyq2(Ctx& c, int tid, bool do_nyq, const int* bb) {
  for (int h = tid; h < 8192; h += NT) {
    uint8_t v = 0;
    if (do_nyq) {
      const int rr = h >> 6, j = h & 63;
      if (rr >= bb[0] && rr < bb[1]) {
        const int cc = 2 * j + c.par(rr);
        if (cc >= bb[2] + c.par(rr) && cc < bb[3]) {
          const int i = rr * TS + cc;
          const uint8_t* n = c.nyq;
          unsigned s = (unsigned)n[(i - V2) >> 1] + n[(i - M1) >> 1] + n[(i - P1) >> 1] + n[(i - 2) >> 1] +
                       n[(i + 2) >> 1] + n[(i + P1) >> 1] + n[(i + M1) >> 1] + n[(i + V2) >> 1];
          v = s > 4 ? 1 : (s < 4 ? 0 : n[i >> 1]);
        }
      }
    }
    c.nyq2[h] = v;
  }
}

// Area statistics on Nyquist sites (demosaic.rs:912-945).
__device__ void pass_nyq_area(Ctx& c, int tid, const int* bb) {
  const float* cfa = c.cfa;
  for (int h = tid; h < 8192; h += NT) {
    const int rr = h >> 6, j = h & 63;
    if (rr < bb[0] || rr >= bb[1]) continue;
    const int cc = 2 * j + c.par(rr);
    if (cc < bb[2] + c.par(rr) || cc >= bb[3]) continue;
    const int i = rr * TS + cc;
    if (!c.nyq2[i >> 1]) continue;
    float sumcfa = 0.0f, sumh = 0.0f, sumv = 0.0f, sumsqh = 0.0f, sumsqv = 0.0f, areawt = 0.0f;
    for (int a = -6; a < 7; a += 2) {
      int i1 = i + a * TS - 6;
      for (int b = -6; b < 7; b += 2, i1 += 2) {
        if (c.nyq2[i1 >> 1]) {
          float ct = cfa[i1];
          sumcfa += ct;
          sumh += cfa[i1 - 1] + cfa[i1 + 1];
          sumv += cfa[i1 - V1] + cfa[i1 + V1];
          sumsqh += sqr(ct - cfa[i1 - 1]) + sqr(ct - cfa[i1 + 1]);
          sumsqv += sqr(ct - cfa[i1 - V1]) + sqr(ct - cfa[i1 + V1]);
          areawt += 1.0f;
        }
      }
    }
    sumh = sumcfa - sumh * 0.5f;
    sumv = sumcfa - sumv * 0.5f;
    areawt = areawt * 0.5f;
    float hcdvar = epssq + fabsf(areawt * sumsqh - sumh * sumh);
    float vcdvar = epssq + fabsf(areawt * sumsqv - sumv * sumv);
    c.hvwt[i >> 1] = hcdvar / (vcdvar + hcdvar);
  }
}

// Diagonal smoothing of a half-resolution weight plane, in place, row by row
// in one wave (demosaic.rs:951-956 for hvwt, 1113-1117 for pmwt): a row reads
// the already-smoothed row above and the untouched row below.  Lane l owns
// half-plane column j = margin/2 + l of every row (site column 2j + parity);
// on a Bayer pattern the row parity alternates, so the diagonal neighbours of
// lane l are lanes l - 1 + p and l + p of the rows above and below (p = this
// row's parity).  The smoothed row above stays in registers and the row below
// is this wave's next read, so a row costs one LDS read and four lane
// exchanges instead of a round trip through LDS per row; the one neighbour
// outside the wave (lane -1, a column left of the margin that the smoothing
// never writes) is read from LDS.
__device__ void pass_diag_smooth(Ctx& c, int lane, float* w, int r0, int r1, int margin) {
  if (r0 >= r1) return;
  const int jl = margin / 2 + lane;  // this lane's half-plane column
  const bool inrow = jl < 64;
  auto own = [&](int rr) { return inrow ? w[rr * 64 + jl] : 0.0f; };
  float above = own(r0 - 1);  // row r0 - 1 is outside the smoothed range: unchanged
  float cur = own(r0);
  for (int rr = r0; rr < r1; ++rr) {
    const int p = c.par(rr);
    const float below = own(rr + 1);  // rows below are not smoothed yet
    const int sa = lane - 1 + p, sb = lane + p;
    float a = __shfl(above, sa < 0 ? 0 : sa, 64), b = __shfl(above, sb > 63 ? 63 : sb, 64);
    float cv = __shfl(below, sa < 0 ? 0 : sa, 64), d = __shfl(below, sb > 63 ? 63 : sb, 64);
    if (sa < 0) {  // column margin/2 - 1: left of the smoothed range in both rows
      a = w[(rr - 1) * 64 + margin / 2 - 1];
      cv = w[(rr + 1) * 64 + margin / 2 - 1];
    }
    const int cc = margin + p + 2 * lane;
    const bool act = cc < c.t.cc1 - margin;
    float nv = cur;
    if (act) {
      const float alt = (a + b + cv + d) / 4.0f;
      nv = fabsf(0.5f - cur) < fabsf(0.5f - alt) ? alt : cur;
      w[rr * 64 + jl] = nv;
    }
    above = nv;
    cur = below;
  }
}

// G at R/B sites after the hvwt smoothing (demosaic.rs:957-971).
__device__ void pass_green(Ctx& c, int tid, int nthreads) {
  const Tile& t = c.t;
  const float* cfa = c.cfa;
  const float* vcd = c.P(S_VCD);
  const float* hcd = c.P(S_HCD);
  float* rgbgreen = c.P(S_RGBGREEN);
  float* dgrb0 = c.P(S_DGRB0);
  float* dgrb2h = c.P(S_DGRB2H);
  float* dgrb2v = c.P(S_DGRB2V);
  for (int h = tid; h < 8192; h += nthreads) {
    const int rr = h >> 6, j = h & 63;
    float d0 = 0.0f, d2h = 0.0f, d2v = 0.0f;
    if (rr >= 8 && rr < t.rr1 - 8) {
      const int p = c.par(rr);
      const int cc = 2 * j + p;
      if (cc >= 8 + p && cc < t.cc1 - 8) {
        const int i = rr * TS + cc;
        const float hw = c.hvwt[h];
        d0 = hw * vcd[i] + (1.0f - hw) * hcd[i];
        const float g = cfa[i] + d0;
        rgbgreen[i] = g;
        if (c.nyq2[h]) {
          d2h = sqr(g - (rgbgreen[i - 1] + rgbgreen[i + 1]) * 0.5f);
          d2v = sqr(g - (rgbgreen[i - V1] + rgbgreen[i + V1]) * 0.5f);
        }
      }
    }
    dgrb0[h] = d0;
    dgrb2h[h] = d2h;
    dgrb2v[h] = d2v;
  }
}

// Nyquist-region refinement of G (demosaic.rs:975-998).
__device__ void pass_green_nyq(Ctx& c, int tid, const int* bb, int nthreads) {
  const float gq0 = 0.169917f, gq1 = 0.108947f, gq2 = 0.069855f, gq3 = 0.0287182f;
  const float* cfa = c.cfa;
  const float* vcd = c.P(S_VCD);
  const float* hcd = c.P(S_HCD);
  const float* d2h = c.P(S_DGRB2H);
  const float* d2v = c.P(S_DGRB2V);
  float* rgbgreen = c.P(S_RGBGREEN);
  float* dgrb0 = c.P(S_DGRB0);
  for (int h = tid; h < 8192; h += nthreads) {
    const int rr = h >> 6, j = h & 63;
    if (rr < bb[0] || rr >= bb[1]) continue;
    const int cc = 2 * j + c.par(rr);
    if (cc < bb[2] + c.par(rr) || cc >= bb[3]) continue;
    const int i = rr * TS + cc;
    if (!c.nyq2[i >> 1]) continue;
    float gvarh = epssq + (gq0 * d2h[i >> 1] +
                           gq1 * (d2h[(i - M1) >> 1] + d2h[(i - P1) >> 1] + d2h[(i + P1) >> 1] + d2h[(i + M1) >> 1]) +
                           gq2 * (d2h[(i - V2) >> 1] + d2h[(i - 2) >> 1] + d2h[(i + 2) >> 1] + d2h[(i + V2) >> 1]) +
                           gq3 * (d2h[(i - M2) >> 1] + d2h[(i - P2) >> 1] + d2h[(i + P2) >> 1] + d2h[(i + M2) >> 1]));
    float gvarv = epssq + (gq0 * d2v[i >> 1] +
                           gq1 * (d2v[(i - M1) >> 1] + d2v[(i - P1) >> 1] + d2v[(i + P1) >> 1] + d2v[(i + M1) >> 1]) +
                           gq2 * (d2v[(i - V2) >> 1] + d2v[(i - 2) >> 1] + d2v[(i + 2) >> 1] + d2v[(i + V2) >> 1]) +
                           gq3 * (d2v[(i - M2) >> 1] + d2v[(i - P2) >> 1] + d2v[(i + P2) >> 1] + d2v[(i + M2) >> 1]));
    float d0 = (hcd[i] * gvarv + vcd[i] * gvarh) / (gvarv + gvarh);
    dgrb0[i >> 1] = d0;
    rgbgreen[i] = cfa[i] + d0;
  }
}

// Diagonal gradient setup (demosaic.rs:1001-1019).
__device__ void pass_diag_setup(Ctx& c, int tid, int nthreads) {
  const Tile& t = c.t;
  const float* cfa = c.cfa;
  float* delp = c.P(S_DELP);
  float* delm = c.P(S_DELM);
  float* sq1p = c.P(S_DGRBSQ1P);
  float* sq1m = c.P(S_DGRBSQ1M);
  for (int h = tid; h < 8192; h += nthreads) {
    const int rr = h >> 6, j = h & 63;
    const int cc = 2 * j;
    float dp = 0.0f, dm = 0.0f, sp = 0.0f, sm = 0.0f;
    if (rr >= 6 && rr < t.rr1 - 6 && cc >= 6 && cc < t.cc1 - 6) {
      const int i = rr * TS + cc;
      if (c.par(rr) == 0) {
        dp = fabsf(cfa[i - P1] - cfa[i + P1]);
        dm = fabsf(cfa[i + M1] - cfa[i - M1]);
        sp = sqr(cfa[i + 1] - cfa[i + 1 - P1]) + sqr(cfa[i + 1] - cfa[i + 1 + P1]);
        sm = sqr(cfa[i + 1] - cfa[i + 1 + M1]) + sqr(cfa[i + 1] - cfa[i + 1 - M1]);
      } else {
        sp = sqr(cfa[i] - cfa[i - P1]) + sqr(cfa[i] - cfa[i + P1]);
        sm = sqr(cfa[i] - cfa[i - M1]) + sqr(cfa[i] - cfa[i + M1]);
        dp = fabsf(cfa[i + 1 - P1] - cfa[i + 1 + P1]);
        dm = fabsf(cfa[i + 1 - M1] - cfa[i + 1 + M1]);
      }
    }
    delp[h] = dp;
    delm[h] = dm;
    sq1p[h] = sp;
    sq1m[h] = sm;
  }
}

// Diagonal R/B estimates and their weight (demosaic.rs:1023-1107).
__device__ void pass_diag(Ctx& c, int tid) {
  const Tile& t = c.t;
  const float* cfa = c.cfa;
  const float* delp = c.P(S_DELP);
  const float* delm = c.P(S_DELM);
  const float* sq1p = c.P(S_DGRBSQ1P);
  const float* sq1m = c.P(S_DGRBSQ1M);
  float* rbm = c.P(S_RBM);
  float* rbp = c.P(S_RBP);
  const float ge0 = 0.13719494435797422f, ge1 = 0.05640252782101291f;
  for (int h = tid; h < 8192; h += NT) {
    const int rr = h >> 6, j = h & 63;
    float vm = 0.0f, vp = 0.0f, pw = 0.0f;
    if (rr >= 8 && rr < t.rr1 - 8) {
      const int p = c.par(rr);
      const int cc = 2 * j + p;
      if (cc >= 8 + p && cc < t.cc1 - 8) {
        const int i = rr * TS + cc, i1 = i >> 1;
        const float ci = cfa[i];
        float crse = (cfa[i + M1] * 2.0f) / (eps + ci + cfa[i + M2]);
        float crnw = (cfa[i - M1] * 2.0f) / (eps + ci + cfa[i - M2]);
        float crne = (cfa[i - P1] * 2.0f) / (eps + ci + cfa[i - P2]);
        float crsw = (cfa[i + P1] * 2.0f) / (eps + ci + cfa[i + P2]);
        float rbse = fabsf(1.0f - crse) < arthresh ? ci * crse : cfa[i + M1] + (ci - cfa[i + M2]) * 0.5f;
        float rbnw = fabsf(1.0f - crnw) < arthresh ? ci * crnw : cfa[i - M1] + (ci - cfa[i - M2]) * 0.5f;
        float rbne = fabsf(1.0f - crne) < arthresh ? ci * crne : cfa[i - P1] + (ci - cfa[i - P2]) * 0.5f;
        float rbsw = fabsf(1.0f - crsw) < arthresh ? ci * crsw : cfa[i + P1] + (ci - cfa[i + P2]) * 0.5f;
        float wtse = eps + delm[i1] + delm[(i + M1) >> 1] + delm[(i + M2) >> 1];
        float wtnw = eps + delm[i1] + delm[(i - M1) >> 1] + delm[(i - M2) >> 1];
        float wtne = eps + delp[i1] + delp[(i - P1) >> 1] + delp[(i - P2) >> 1];
        float wtsw = eps + delp[i1] + delp[(i + P1) >> 1] + delp[(i + P2) >> 1];
        vm = (wtse * rbnw + wtnw * rbse) / (wtse + wtnw);
        vp = (wtne * rbsw + wtsw * rbne) / (wtne + wtsw);
        float rbvarm = epssq + ge0 * (sq1m[(i - V1) >> 1] + sq1m[(i - 1) >> 1] + sq1m[(i + 1) >> 1] + sq1m[(i + V1) >> 1]) +
                       ge1 * (sq1m[(i - V2 - 1) >> 1] + sq1m[(i - V2 + 1) >> 1] + sq1m[(i - 2 - V1) >> 1] +
                              sq1m[(i + 2 - V1) >> 1] + sq1m[(i - 2 + V1) >> 1] + sq1m[(i + 2 + V1) >> 1] +
                              sq1m[(i + V2 - 1) >> 1] + sq1m[(i + V2 + 1) >> 1]);
        float rbvarp = epssq + ge0 * (sq1p[(i - V1) >> 1] + sq1p[(i - 1) >> 1] + sq1p[(i + 1) >> 1] + sq1p[(i + V1) >> 1]) +
                       ge1 * (sq1p[(i - V2 - 1) >> 1] + sq1p[(i - V2 + 1) >> 1] + sq1p[(i - 2 - V1) >> 1] +
                              sq1p[(i + 2 - V1) >> 1] + sq1p[(i - 2 + V1) >> 1] + sq1p[(i + 2 + V1) >> 1] +
                              sq1p[(i + V2 - 1) >> 1] + sq1p[(i + V2 + 1) >> 1]);
        pw = rbvarm / (rbvarp + rbvarm);
        if (vp < ci) {
          if (vp * 2.0f < ci) {
            vp = ulim(vp, cfa[i - P1], cfa[i + P1]);
          } else {
            float pwt = (ci - vp) * 2.0f / (eps + vp + ci);
            vp = pwt * vp + (1.0f - pwt) * ulim(vp, cfa[i - P1], cfa[i + P1]);
          }
        }
        if (vm < ci) {
          if (vm * 2.0f < ci) {
            vm = ulim(vm, cfa[i - M1], cfa[i + M1]);
          } else {
            float mwt = (ci - vm) * 2.0f / (eps + vm + ci);
            vm = mwt * vm + (1.0f - mwt) * ulim(vm, cfa[i - M1], cfa[i + M1]);
          }
        }
        if (vp > clip_pt) vp = ulim(vp, cfa[i - P1], cfa[i + P1]);
        if (vm > clip_pt) vm = ulim(vm, cfa[i - M1], cfa[i + M1]);
      }
    }
    rbm[h] = vm;
    rbp[h] = vp;
    c.pmwt[h] = pw;
  }
}

// rbint after the pmwt smoothing (demosaic.rs:1118-1121).
__device__ void pass_rbint(Ctx& c, int tid) {
  const Tile& t = c.t;
  const float* cfa = c.cfa;
  const float* rbm = c.P(S_RBM);
  const float* rbp = c.P(S_RBP);
  float* rbint = c.P(S_RBINT);
  for (int h = tid; h < 8192; h += NT) {
    const int rr = h >> 6, j = h & 63;
    float v = 0.0f;
    if (rr >= 10 && rr < t.rr1 - 10) {
      const int p = c.par(rr);
      const int cc = 2 * j + p;
      if (cc >= 10 + p && cc < t.cc1 - 10) {
        const int i = rr * TS + cc;
        const float pw = c.pmwt[h];
        v = (cfa[i] + rbm[h] * (1.0f - pw) + rbp[h] * pw) * 0.5f;
      }
    }
    rbint[h] = v;
  }
}

// G refinement at R/B sites from the diagonal estimates (demosaic.rs:1125-1194).
__device__ void pass_green_diag(Ctx& c, int tid) {
  const Tile& t = c.t;
  const float* cfa = c.cfa;
  const float* rbint = c.P(S_RBINT);
  float* rgbgreen = c.P(S_RGBGREEN);
  float* dgrb0 = c.P(S_DGRB0);
  for (int h = tid; h < 8192; h += NT) {
    const int rr = h >> 6, j = h & 63;
    if (rr < 12 || rr >= t.rr1 - 12) continue;
    const int p = c.par(rr);
    const int cc = 2 * j + p;
    if (cc < 12 + p || cc >= t.cc1 - 12) continue;
    const int i = rr * TS + cc, i1 = i >> 1;
    if (!(fabsf(0.5f - c.pmwt[i1]) >= fabsf(0.5f - c.hvwt[i1]))) continue;
    const float rb = rbint[i1];
    float cru = cfa[i - V1] * 2.0f / (eps + rb + rbint[i1 - V1]);
    float crd = cfa[i + V1] * 2.0f / (eps + rb + rbint[i1 + V1]);
    float crl = cfa[i - 1] * 2.0f / (eps + rb + rbint[i1 - 1]);
    float crr = cfa[i + 1] * 2.0f / (eps + rb + rbint[i1 + 1]);
    float gu = fabsf(1.0f - cru) < arthresh ? rb * cru : cfa[i - V1] + (rb - rbint[i1 - V1]) * 0.5f;
    float gd = fabsf(1.0f - crd) < arthresh ? rb * crd : cfa[i + V1] + (rb - rbint[i1 + V1]) * 0.5f;
    float gl = fabsf(1.0f - crl) < arthresh ? rb * crl : cfa[i - 1] + (rb - rbint[i1 - 1]) * 0.5f;
    float gr = fabsf(1.0f - crr) < arthresh ? rb * crr : cfa[i + 1] + (rb - rbint[i1 + 1]) * 0.5f;
    const float dw0u = dirw0(cfa, i - V1), dw0d = dirw0(cfa, i + V1);
    const float dw1l = dirw1(cfa, i - 1), dw1r = dirw1(cfa, i + 1);
    float gintv = (dw0u * gd + dw0d * gu) / fmaxf(dw0d + dw0u, 1e-10f);
    float ginth = (dw1l * gr + dw1r * gl) / fmaxf(dw1l + dw1r, 1e-10f);
    if (gintv < rb) {
      if (2.0f * gintv < rb) {
        gintv = ulim(gintv, cfa[i - V1], cfa[i + V1]);
      } else {
        float vwt = 2.0f * (rb - gintv) / (eps + gintv + rb);
        gintv = vwt * gintv + (1.0f - vwt) * ulim(gintv, cfa[i - V1], cfa[i + V1]);
      }
    }
    if (ginth < rb) {
      if (2.0f * ginth < rb) {
        ginth = ulim(ginth, cfa[i - 1], cfa[i + 1]);
      } else {
        float hwt = 2.0f * (rb - ginth) / (eps + ginth + rb);
        ginth = hwt * ginth + (1.0f - hwt) * ulim(ginth, cfa[i - 1], cfa[i + 1]);
      }
    }
    if (ginth > clip_pt) ginth = ulim(ginth, cfa[i - 1], cfa[i + 1]);
    if (gintv > clip_pt) gintv = ulim(gintv, cfa[i - V1], cfa[i + V1]);
    const float hw = c.hvwt[i1];
    const float g = ginth * (1.0f - hw) + gintv * hw;
    rgbgreen[i] = g;
    dgrb0[i1] = g - cfa[i];
  }
}

// Split G-R / G-B differences by colour (demosaic.rs:1204-1213).
__device__ void pass_split(Ctx& c, int tid, int ex, int ey) {
  const Tile& t = c.t;
  float* dgrb0 = c.P(S_DGRB0);
  float* dgrb1 = c.P(S_DGRB1);
  for (int h = tid; h < 8192; h += NT) {
    const int rr = h >> 6;
    float v1 = 0.0f;
    if (rr >= 13 - ey && rr < t.rr1 - 12 && ((rr - (13 - ey)) & 1) == 0 && h >= ((rr * TS + 13 - ex) >> 1) &&
        h < ((rr * TS + t.cc1 - 12) >> 1)) {
      v1 = dgrb0[h];
      dgrb0[h] = 0.0f;
    }
    dgrb1[h] = v1;
  }
}

// Chroma interpolation at R/B sites (demosaic.rs:1215-1252).
__device__ void pass_chroma(Ctx& c, int tid) {
  const Tile& t = c.t;
  for (int h = tid; h < 8192; h += NT) {
    const int rr = h >> 6, j = h & 63;
    if (rr < 14 || rr >= t.rr1 - 14) continue;
    const int p = c.par(rr);
    const int cc = 2 * j + p;
    if (cc < 14 + p || cc >= t.cc1 - 14) continue;
    const int i = rr * TS + cc;
    const int ch = 1 - c.color_at(t.top + rr, t.left + cc) / 2;
    float* d = c.P(ch == 0 ? S_DGRB0 : S_DGRB1);
    float wtnw = 1.0f / (eps + fabsf(d[(i - M1) >> 1] - d[(i + M1) >> 1]) + fabsf(d[(i - M1) >> 1] - d[(i - M3) >> 1]) +
                         fabsf(d[(i + M1) >> 1] - d[(i - M3) >> 1]));
    float wtne = 1.0f / (eps + fabsf(d[(i - P1) >> 1] - d[(i + P1) >> 1]) + fabsf(d[(i - P1) >> 1] - d[(i - P3) >> 1]) +
                         fabsf(d[(i + P1) >> 1] - d[(i - P3) >> 1]));
    float wtsw = 1.0f / (eps + fabsf(d[(i + P1) >> 1] - d[(i - P1) >> 1]) + fabsf(d[(i + P1) >> 1] - d[(i + M3) >> 1]) +
                         fabsf(d[(i - P1) >> 1] - d[(i + P3) >> 1]));
    float wtse = 1.0f / (eps + fabsf(d[(i + M1) >> 1] - d[(i - M1) >> 1]) + fabsf(d[(i + M1) >> 1] - d[(i + P3) >> 1]) +
                         fabsf(d[(i - M1) >> 1] - d[(i + M3) >> 1]));
    float val = (wtnw * (1.325f * d[(i - M1) >> 1] - 0.175f * d[(i - M3) >> 1] - 0.075f * d[(i - M1 - 2) >> 1] -
                         0.075f * d[(i - M1 - V2) >> 1]) +
                 wtne * (1.325f * d[(i - P1) >> 1] - 0.175f * d[(i - P3) >> 1] - 0.075f * d[(i - P1 + 2) >> 1] -
                         0.075f * d[(i - P1 + V2) >> 1]) +
                 wtsw * (1.325f * d[(i + P1) >> 1] - 0.175f * d[(i + P3) >> 1] - 0.075f * d[(i + P1 - 2) >> 1] -
                         0.075f * d[(i - M1) >> 1]) +
                 wtse * (1.325f * d[(i + M1) >> 1] - 0.175f * d[(i + M3) >> 1] - 0.075f * d[(i + M1 + 2) >> 1] -
                         0.075f * d[(i + M1 + V2) >> 1])) /
                (wtnw + wtne + wtsw + wtse);
    d[i >> 1] = val;
  }
}

// Output of the tile's 96x96 core (demosaic.rs:1254-1366).
__device__ void pass_output(Ctx& c, int tid) {
  const Tile& t = c.t;
  const float* rgbgreen = c.P(S_RGBGREEN);
  const float* dgrb0 = c.P(S_DGRB0);
  const float* dgrb1 = c.P(S_DGRB1);
  const float* hvwt = c.hvwt;
  const int ncol = t.cc1 - 32 > 0 ? t.cc1 - 32 : 0;  // output columns 16 .. cc1-16
  const int nrow = t.rr1 - 32 > 0 ? t.rr1 - 32 : 0;
  for (int k = tid; k < nrow * ncol; k += NT) {
    const int rr = 16 + k / ncol, cc = 16 + k % ncol;
    const int row = rr + t.top, col = cc + t.left;
    if (!(col >= 0 && col < c.a.width && row >= 0 && row < c.a.height)) continue;
    const int i = rr * TS + cc;
    // G-first rows compute the full form at even offsets from 16, the direct
    // form at odd ones (and vice versa); an odd trailing column takes the
    // first form of its row.
    const bool g_first = c.par(rr) == 1;
    const bool even_off = ((cc - 16) & 1) == 0;
    const bool full = g_first == even_off;
    float r, g, b;
    const float gg = rgbgreen[i];
    if (full) {
      const float hu = hvwt[(i - V1) >> 1], hr = hvwt[(i + 1) >> 1], hl = hvwt[(i - 1) >> 1], hd = hvwt[(i + V1) >> 1];
      const float temp = 1.0f / fmaxf(hu + 2.0f - hr - hl + hd, 1e-5f);
      r = clampnan(gg - (hu * dgrb0[(i - V1) >> 1] + (1.0f - hr) * dgrb0[(i + 1) >> 1] + (1.0f - hl) * dgrb0[(i - 1) >> 1] +
                         hd * dgrb0[(i + V1) >> 1]) * temp,
                   0.0f, 1.0f);
      b = clampnan(gg - (hu * dgrb1[(i - V1) >> 1] + (1.0f - hr) * dgrb1[(i + 1) >> 1] + (1.0f - hl) * dgrb1[(i - 1) >> 1] +
                         hd * dgrb1[(i + V1) >> 1]) * temp,
                   0.0f, 1.0f);
    } else {
      r = clampnan(gg - dgrb0[i >> 1], 0.0f, 1.0f);
      b = clampnan(gg - dgrb1[i >> 1], 0.0f, 1.0f);
    }
    g = clampnan(gg, 0.0f, 1.0f);
    float* o = c.a.out + ((size_t)row * c.a.width + col) * 3;
    o[0] = r; o[1] = g; o[2] = b;
  }
}

__global__ void __launch_bounds__(NT) amaze_kernel(AmazeArgs a) {
  __shared__ float s_cfa[16384];
  __shared__ float s_hvwt[8192];
  __shared__ float s_pmwt[8192];
  __shared__ uint8_t s_nyq[8192];
  __shared__ uint8_t s_nyq2[8192];
  __shared__ int s_bbox[4];

  const int tid = threadIdx.x;
  const int wave = tid >> 6, lane = tid & 63;
  float* slot = a.scratch + (size_t)blockIdx.x * kAmazeSlotFloats;

  // (ey, ex): location of the R site in the pattern (demosaic.rs:1196-1202)
  int ex, ey;
  if (a.cfa[0][0] == 1) {
    if (a.cfa[0][1] == 0) { ex = 1; ey = 0; } else { ex = 0; ey = 1; }
  } else {
    if (a.cfa[0][0] == 0) { ex = 0; ey = 0; } else { ex = 1; ey = 1; }
  }

  const uint32_t cfa4 = (uint32_t)a.cfa[0][0] | ((uint32_t)a.cfa[0][1] << 8) | ((uint32_t)a.cfa[1][0] << 16) |
                        ((uint32_t)a.cfa[1][1] << 24);
  for (int tile = blockIdx.x; tile < a.tiles_x * a.tiles_y; tile += gridDim.x) {
    Ctx c{a, {}, cfa4, slot, s_cfa, s_hvwt, s_pmwt, s_nyq, s_nyq2};
    const int ty = tile / a.tiles_x, tx = tile % a.tiles_x;
    Tile& t = c.t;
    t.top = -16 + 96 * ty;
    t.left = -16 + 96 * tx;
    const int bottom = min(t.top + 128, a.height + 16), right = min(t.left + 128, a.width + 16);
    t.rr1 = bottom - t.top;
    t.cc1 = right - t.left;
    t.rrmin = t.top < 0 ? 16 : 0;
    t.ccmin = t.left < 0 ? 16 : 0;
    t.rrmax = bottom > a.height ? a.height - t.top : t.rr1;
    t.ccmax = right > a.width ? a.width - t.left : t.cc1;

    if (tid == 0) { s_bbox[0] = INT32_MAX; s_bbox[1] = -1; s_bbox[2] = INT32_MAX; s_bbox[3] = -1; }
    pass_fill(c, tid);
    __syncthreads();
    pass_colour_diff(c, tid);
    __syncthreads();
    pass_cd_chains(c, tid);
    __syncthreads();
    pass_cddiffsq(c, tid);
    __syncthreads();
    pass_hvwt_nyq(c, tid, s_bbox);
    __syncthreads();

    // Nyquist bounding box (demosaic.rs:849-886)
    int bb[4] = {0, 0, 0, 0};
    bool do_nyq = false;
    if (s_bbox[1] >= 0) {
      int sr = s_bbox[0], er = s_bbox[1], sc = s_bbox[2], ec = s_bbox[3];
      do_nyq = sr != er && sc != ec;
      if (do_nyq) {
        er += 1; ec += 1;
        sc -= sc & 1;
        sr = max(sr, 8); er = min(er, t.rr1 - 8);
        sc = max(sc, 8); ec = min(ec, t.cc1 - 8);
        bb[0] = sr; bb[1] = er; bb[2] = sc; bb[3] = ec;
      }
    }
    pass_nyq2(c, tid, do_nyq, bb);
    __syncthreads();
    if (do_nyq) pass_nyq_area(c, tid, bb);
    __syncthreads();

    // hvwt smoothing (wave 0) overlaps the diagonal setup (waves 1..15)
    if (wave == 0) pass_diag_smooth(c, lane, s_hvwt, 8, t.rr1 - 8, 8);
    else pass_diag_setup(c, tid - 64, NT - 64);
    __syncthreads();
    pass_diag(c, tid);  // reads neither hvwt nor anything pass_green writes
    __syncthreads();
    // pmwt smoothing (wave 0) overlaps G at R/B sites (waves 1..15; it reads the
    // smoothed hvwt and no pmwt)
    if (wave == 0) pass_diag_smooth(c, lane, s_pmwt, 10, t.rr1 - 10, 10);
    else pass_green(c, tid - 64, NT - 64);
    __syncthreads();
    if (do_nyq) {  // the Nyquist G refinement reads pass_green's planes at neighbouring sites
      pass_green_nyq(c, tid, bb, NT);
      __syncthreads();
    }
    pass_rbint(c, tid);
    __syncthreads();
    pass_green_diag(c, tid);
    __syncthreads();
    pass_split(c, tid, ex, ey);
    __syncthreads();
    pass_chroma(c, tid);
    __syncthreads();
    pass_output(c, tid);
    __syncthreads();
  }
}

// Scratch of the allocation-free entry point: kAmazeNoallocSlots tile slots.
// Every pass writes each element of the plane it defines before any pass
// reads it (zeros outside the pass's valid region included), so nothing in a
// slot has to start at zero, and a slot's contents from an earlier tile or
// call are never read.  Host code must use the address from
// amaze_noalloc_scratch, never this name.
__device__ __align__(256) float scratch[(size_t)kAmazeNoallocSlots * kAmazeSlotFloats];

}  // namespace

hipError_t amaze_noalloc_scratch(int device, float** out) {
  constexpr int kMaxDevices = 64;
  static std::mutex mu;
  static float* addr[kMaxDevices] = {};
  if (device < 0 || device >= kMaxDevices || !out) return hipErrorInvalidValue;
  std::lock_guard<std::mutex> lock(mu);
  if (!addr[device]) {
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return e;
    addr[device] = static_cast<float*>(p);
  }
  *out = addr[device];
  return hipSuccess;
}

size_t amaze_tiles(int width, int height, int* tiles_x, int* tiles_y) {
  int tx = 0, ty = 0;
  for (int left = -16; left < width; left += 96) ++tx;
  for (int top = -16; top < height; top += 96) ++ty;
  if (tiles_x) *tiles_x = tx;
  if (tiles_y) *tiles_y = ty;
  return (size_t)tx * ty;
}

hipError_t amaze_launch(const AmazeArgs& args, int grid, hipStream_t stream) {
  hipLaunchKernelGGL(amaze_kernel, dim3(grid), dim3(NT), 0, stream, args);
  return hipGetLastError();
}

}  // namespace imops
