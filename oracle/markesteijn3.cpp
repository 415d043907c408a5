// Oracle (test infrastructure only): the three-pass Markesteijn X-Trans
// demosaic, the `markesteijn3` extension of Demosaic (the reference's
// "markesteijn", demosaic.rs:108-186, is a 3x3 same-colour mean; it has no
// multi-pass method).  Restated from the structure of F. Markesteijn's
// published X-Trans algorithm: green interpolated along four directions and
// clamped to the local green range; with more than one pass, green recomputed
// along each direction from the interpolated greens of the adjacent (closer)
// pixels, giving four more candidates; red and blue per candidate by colour
// differences; a perceptual (Y, Pb, Pr) derivative per candidate; homogeneity
// maps; the average of the most homogeneous candidates.  The formulas are this
// project's definition of each step:
//
//   directions d: (0,1), (1,0), (1,1), (1,-1) (dy, dx)
//   near(p, s d, ch): the first k in 1..3 with colour ch at p + s k d (in bounds), else none
//   green bounds at p: min/max of the green CFA values of the 3x3 window (5x5 if it has none)
//   candidate d (d = 0..3), non-green p: greens at p - a d and p + b d ->
//       (b g(p - a d) + a g(p + b d)) / (a + b), one side's value if only one exists,
//       the bounds' midpoint if neither; clamped to the bounds
//   candidate 4 + d (passes > 1), non-green p: candidate d, except where the nearest
//       green along -d or +d is not adjacent and both adjacent pixels exist: the mean
//       of candidate d at p - d and p + d; clamped to the bounds
//   red / blue of a candidate along its direction d, non-native p: its green at p plus
//       the same weighted interpolation of (v - G) at the nearest native pixels
//       along -+d; the raster-order mean of (v - G) over the native pixels of the 3x3
//       (else 5x5) window if neither exists
//   Y = 0.2627 R + 0.678 G + 0.0593 B, Pb = (B - Y) 0.5315, Pr = (R - Y) 0.6781;
//   diff_c(p) = D(p, p - d) + D(p, p + d) (neighbours clamped into the image, d the
//       candidate's direction), D = dY^2 + dPb^2 + dPr^2
//   flag_c(p) = diff_c(p) <= 8 min over candidates of diff(p); score_c(p) = sum of
//       flag_c over the in-bounds 5x5 window; the candidates with 8 score >= 7 max
//       score are averaged (sum in candidate order, divided by their count).
//
// passes = 1 uses candidates 0..3, passes = 3 all eight.  T = float is the FP32
// form the GPU reproduces bit for bit (same operation order, no contraction);
// T = double is the FP64 oracle.
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <vector>

#include "oracle_colour.hpp"

namespace oracle {

// Test hook (tests/test_oracle_markesteijn3_cpu.py only): a deliberate mistake
// (0 = none) that the tests must catch: 1 the first diagonal's search steps one
// column short on its + side, 2 the two-sided colour-difference weights swapped,
// 3 the green clamp's lower bound replaced by the upper one.
int g_m3_plant = 0;

namespace {

constexpr int DY[4] = {0, 1, 1, 1}, DX[4] = {1, 0, 1, -1};

struct M3 {
  const float* in;
  int W, H;
  const Cfa& cf;
  M3(const float* i, int w, int h, const Cfa& c) : in(i), W(w), H(h), cf(c) {}
  int col(int y, int x) const { return cf.color_at((size_t)y, (size_t)x); }
  bool inb(int y, int x) const { return y >= 0 && y < H && x >= 0 && x < W; }
  int near(int y, int x, int d, int s, int ch) const {
    for (int k = 1; k <= 3; ++k) {
      const int yy = y + s * k * DY[d], xx = x + s * k * DX[d] - (g_m3_plant == 1 && d == 2 && s > 0 ? 1 : 0);
      if (!inb(yy, xx)) return 0;
      if (col(yy, xx) == ch) return k;
    }
    return 0;
  }
};

}  // namespace

template <class T>
void markesteijn3(const float* in, size_t width, size_t height, const Cfa& cfa, int passes, float* out_rgb) {
  const int W = (int)width, H = (int)height;
  const size_t n = width * height;
  const M3 m(in, W, H, cfa);
  const int nc = passes > 1 ? 8 : 4;
  auto v = [&](int y, int x) { return (T)in[(size_t)y * W + x]; };
  std::vector<T> G(8 * n), R(8 * n), B(8 * n);
  std::vector<uint8_t> flag(8 * n);
  std::vector<T> lo(n), hi(n);
  std::vector<uint8_t> hb(n);
  // green bounds
#pragma omp parallel for schedule(static)
  for (long long y = 0; y < H; ++y)
    for (int x = 0; x < W; ++x) {
      const size_t p = (size_t)y * W + x;
      hb[p] = 0;
      if (m.col((int)y, x) == 1) continue;
      for (int r = 1; r <= 2 && !hb[p]; ++r)
        for (int yy = (int)y - r; yy <= (int)y + r; ++yy)
          for (int xx = x - r; xx <= x + r; ++xx) {
            if (!m.inb(yy, xx) || m.col(yy, xx) != 1) continue;
            const T g = v(yy, xx);
            if (!hb[p]) { lo[p] = hi[p] = g; hb[p] = 1; }
            else { lo[p] = std::min(lo[p], g); hi[p] = std::max(hi[p], g); }
          }
    }
  // candidates 0..3: directional green
#pragma omp parallel for schedule(static)
  for (long long y = 0; y < H; ++y)
    for (int x = 0; x < W; ++x) {
      const size_t p = (size_t)y * W + x;
      const int c = m.col((int)y, x);
      for (int d = 0; d < 4; ++d) {
        T g;
        if (c == 1) {
          g = v((int)y, x);
        } else {
          const int a = m.near((int)y, x, d, -1, 1), b = m.near((int)y, x, d, 1, 1);
          if (a && b)
            g = ((T)b * v((int)y - a * DY[d], x - a * DX[d]) + (T)a * v((int)y + b * DY[d], x + b * DX[d])) / (T)(a + b);
          else if (a) g = v((int)y - a * DY[d], x - a * DX[d]);
          else if (b) g = v((int)y + b * DY[d], x + b * DX[d]);
          else g = hb[p] ? (lo[p] + hi[p]) * (T)0.5 : v((int)y, x);
          if (hb[p]) g = std::min(std::max(g, g_m3_plant == 3 ? hi[p] : lo[p]), hi[p]);
        }
        G[d * n + p] = g;
      }
    }
  // candidates 4..7: green recomputed from the closer pixels
  if (nc == 8) {
#pragma omp parallel for schedule(static)
    for (long long y = 0; y < H; ++y)
      for (int x = 0; x < W; ++x) {
        const size_t p = (size_t)y * W + x;
        const int c = m.col((int)y, x);
        for (int d = 0; d < 4; ++d) {
          T g = G[d * n + p];
          if (c != 1) {
            const int a = m.near((int)y, x, d, -1, 1), b = m.near((int)y, x, d, 1, 1);
            const int ya = (int)y - DY[d], xa = x - DX[d], yb = (int)y + DY[d], xb = x + DX[d];
            if (!(a == 1 && b == 1) && m.inb(ya, xa) && m.inb(yb, xb))
              g = (G[d * n + (size_t)ya * W + xa] + G[d * n + (size_t)yb * W + xb]) * (T)0.5;
            if (hb[p]) g = std::min(std::max(g, lo[p]), hi[p]);
          }
          G[(4 + d) * n + p] = g;
        }
      }
  }
  // red and blue per candidate
#pragma omp parallel for schedule(static)
  for (long long y = 0; y < H; ++y)
    for (int x = 0; x < W; ++x) {
      const size_t p = (size_t)y * W + x;
      const int c = m.col((int)y, x);
      for (int k = 0; k < nc; ++k) {
        const int d = k & 3;
        const T* Gk = G.data() + k * n;
        T res[2];
        for (int j = 0; j < 2; ++j) {
          const int ch = j == 0 ? 0 : 2;
          if (c == ch) {
            res[j] = v((int)y, x);
            continue;
          }
          const int a = m.near((int)y, x, d, -1, ch), b = m.near((int)y, x, d, 1, ch);
          T cd = 0;
          if (a || b) {
            const int ya = (int)y - a * DY[d], xa = x - a * DX[d], yb = (int)y + b * DY[d], xb = x + b * DX[d];
            const int wa = g_m3_plant == 2 ? a : b, wb = g_m3_plant == 2 ? b : a;
            if (a && b)
              cd = ((T)wa * (v(ya, xa) - Gk[(size_t)ya * W + xa]) + (T)wb * (v(yb, xb) - Gk[(size_t)yb * W + xb])) /
                   (T)(a + b);
            else if (a) cd = v(ya, xa) - Gk[(size_t)ya * W + xa];
            else cd = v(yb, xb) - Gk[(size_t)yb * W + xb];
          } else {
            for (int r = 1; r <= 2; ++r) {
              T acc = 0;
              int cnt = 0;
              for (int yy = (int)y - r; yy <= (int)y + r; ++yy)
                for (int xx = x - r; xx <= x + r; ++xx) {
                  if (!m.inb(yy, xx) || m.col(yy, xx) != ch) continue;
                  acc += v(yy, xx) - Gk[(size_t)yy * W + xx];
                  ++cnt;
                }
              if (cnt) {
                cd = acc / (T)cnt;
                break;
              }
            }
          }
          res[j] = Gk[p] + cd;
        }
        R[k * n + p] = res[0];
        B[k * n + p] = res[1];
      }
    }
  // perceptual derivatives and flags
  auto ypp = [&](int k, size_t q, T o[3]) {
    const T r = R[k * n + q], g = G[k * n + q], b = B[k * n + q];
    const T Y = (T)0.2627 * r + (T)0.678 * g + (T)0.0593 * b;
    o[0] = Y;
    o[1] = (b - Y) * (T)0.5315;
    o[2] = (r - Y) * (T)0.6781;
  };
#pragma omp parallel for schedule(static)
  for (long long y = 0; y < H; ++y)
    for (int x = 0; x < W; ++x) {
      const size_t p = (size_t)y * W + x;
      T dd[8];
      T mn = 0;
      for (int k = 0; k < nc; ++k) {
        const int d = k & 3;
        T c0[3];
        ypp(k, p, c0);
        T part[2];
        for (int j = 0; j < 2; ++j) {
          const int s = j == 0 ? -1 : 1;
          const int yy = std::min(std::max((int)y + s * DY[d], 0), H - 1), xx = std::min(std::max(x + s * DX[d], 0), W - 1);
          T c1[3];
          ypp(k, (size_t)yy * W + xx, c1);
          const T e0 = c0[0] - c1[0], e1 = c0[1] - c1[1], e2 = c0[2] - c1[2];
          part[j] = e0 * e0 + e1 * e1 + e2 * e2;
        }
        dd[k] = part[0] + part[1];
        mn = k == 0 ? dd[k] : std::min(mn, dd[k]);
      }
      const T tr = (T)8 * mn;
      for (int k = 0; k < nc; ++k) flag[k * n + p] = dd[k] <= tr ? 1 : 0;
    }
  // homogeneity scores and the average of the selected candidates
#pragma omp parallel for schedule(static)
  for (long long y = 0; y < H; ++y)
    for (int x = 0; x < W; ++x) {
      const size_t p = (size_t)y * W + x;
      int sc[8] = {0, 0, 0, 0, 0, 0, 0, 0};
      for (int yy = std::max((int)y - 2, 0); yy <= std::min((int)y + 2, H - 1); ++yy)
        for (int xx = std::max(x - 2, 0); xx <= std::min(x + 2, W - 1); ++xx)
          for (int k = 0; k < nc; ++k) sc[k] += flag[k * n + (size_t)yy * W + xx];
      int mx = 0;
      for (int k = 0; k < nc; ++k) mx = std::max(mx, sc[k]);
      T acc[3] = {0, 0, 0};
      int cnt = 0;
      for (int k = 0; k < nc; ++k) {
        if (8 * sc[k] < 7 * mx) continue;
        acc[0] += R[k * n + p];
        acc[1] += G[k * n + p];
        acc[2] += B[k * n + p];
        ++cnt;
      }
      for (int j = 0; j < 3; ++j) out_rgb[p * 3 + j] = (float)(acc[j] / (T)cnt);
    }
}

template void markesteijn3<float>(const float*, size_t, size_t, const Cfa&, int, float*);
template void markesteijn3<double>(const float*, size_t, size_t, const Cfa&, int, float*);

}  // namespace oracle
