// Oracle (test infrastructure only): AMaZE Bayer demosaic restated from
// pichromatic/src/demosaic.rs:358-1379, one 128x128 tile at a time on the
// tile grid top,left = -16 + 96k (demosaic.rs:397-407).  Instantiated for
// float (exact reference op order) and double (FP64 error oracle).
// Pins: uniform input (the reference's own demosaic tests), closed-form linear
// ramps, native-site passthrough, and on textured input the exact transpose
// equivariance of the passes demosaic.rs:612-786, the rotation equivariance of
// the stencil passes and the closed-form Nyquist boundary
// (tests/test_oracle_amaze_cpu.py, which also plants mistakes through the
// hooks below to show each pin catches one).
#include <vector>

#include "oracle_colour.hpp"

namespace oracle {

// Test hooks, used only by tests/test_oracle_amaze_cpu.py: `plant` selects a
// deliberate restatement mistake (0: none) so the tests can show that their
// checks catch it; `trace`, when set, receives per-pass planes of the tile
// whose top-left corner is (trace_top, trace_left), each as a 128 x 128 float
// plane in tile coordinates (half-resolution planes expanded to the sites
// they belong to, NaN elsewhere).
AmazeHooks g_amaze_hooks;

namespace {

template <class T>
struct AmazeTile {
  static constexpr int TS = 128, TSH = 64;
  static constexpr int V1 = 128, V2 = 256, V3 = 384;
  static constexpr int P1 = 127, P2 = 254, P3 = 381;
  static constexpr int M1 = 129, M2 = 258, M3 = 387;

  std::vector<T> rgbgreen, delhvsqsum, dirwts0, dirwts1, vcd, hcd, vcdalt, hcdalt, cddiffsq, hvwt, dgrb0,
      dgrb1, delp, delm, rbint, dgrb2_h, dgrb2_v, dgintv, dginth, dgrbsq1m, dgrbsq1p, cfa, pmwt, rbm, rbp,
      nyqutest;
  std::vector<uint8_t> nyquist, nyquist2;

  AmazeTile()
      : rgbgreen(16384), delhvsqsum(16384), dirwts0(16384), dirwts1(16384), vcd(16384), hcd(16384),
        vcdalt(16384), hcdalt(16384), cddiffsq(16384), hvwt(8192), dgrb0(8192), dgrb1(8192), delp(8192),
        delm(8192), rbint(8192), dgrb2_h(8192), dgrb2_v(8192), dgintv(16384), dginth(16384), dgrbsq1m(8192),
        dgrbsq1p(8192), cfa(16384), pmwt(8192), rbm(8192), rbp(8192), nyqutest(8192), nyquist(8192),
        nyquist2(8192) {}

  static T sqr(T x) { return x * x; }
  static T lim(T x, T lo, T hi) { return rmax(lo, rmin(x, hi)); }
  static T ulim(T x, T y, T z) { return y < z ? lim(x, y, z) : lim(x, z, y); }
  static T clampnan(T x, T m, T mx) { return std::isnan(x) ? (m + mx) * (T)0.5 : rclamp(x, m, mx); }
  static T interpolatef(T w, T a, T b) { return w * a + ((T)1 - w) * b; }

  float* trace_plane(int top, int left, int k) const {
    const AmazeHooks& hk = g_amaze_hooks;
    if (!hk.trace || top != hk.trace_top || left != hk.trace_left) return nullptr;
    return hk.trace + (size_t)k * 16384;
  }
  void trace_full(int top, int left, int k, const std::vector<T>& v) const {
    if (float* d = trace_plane(top, left, k))
      for (int i = 0; i < 16384; ++i) d[i] = (float)v[i];
  }
  template <class V>
  void trace_half(int top, int left, int k, const std::vector<V>& v, const Cfa& cf) const {
    float* d = trace_plane(top, left, k);
    if (!d) return;
    for (int rr = 0; rr < 128; ++rr)
      for (int cc = 0; cc < 128; ++cc) {
        const int par = cf.color_at((size_t)((top + rr) % 48 + 48) % 48, (size_t)((left + 2) % 48 + 48) % 48) & 1;
        d[rr * 128 + cc] = ((cc & 1) == par) ? (float)v[(rr * 128 + cc) >> 1] : NAN;
      }
  }

  void run(const float* input, size_t width, size_t height, const Cfa& cf, int top, int left, float* out,
           T clip_pt) {
    const T eps = (T)1e-5f, epssq = (T)1e-10f, arthresh = (T)0.75f;
    const T gaussodd[4] = {(T)0.14659727707323927f, (T)0.103592713382435f, (T)0.0732036125103057f,
                           (T)0.0365543548389495f};
    const T nyqthresh = g_amaze_hooks.plant == 5 ? (T)5.0f : (g_amaze_hooks.plant == 6 ? (T)0.05f : (T)0.5f);
    const T gaussgrad[6] = {nyqthresh * (T)0.07384411893421103f, nyqthresh * (T)0.06207511968171489f,
                            nyqthresh * (T)0.0521818194747806f,  nyqthresh * (T)0.03687419286733595f,
                            nyqthresh * (T)0.03099732204057846f, nyqthresh * (T)0.018413194161458882f};
    const T gausseven[2] = {(T)0.13719494435797422f, (T)0.05640252782101291f};
    const T gquinc[4] = {(T)0.169917f, (T)0.108947f, (T)0.069855f, (T)0.0287182f};
    const int H = (int)height, W = (int)width;

    auto get_in = [&](int r, int c) -> T {
      int rc = r < 0 ? 0 : (r > H - 1 ? H - 1 : r);
      int cc = c < 0 ? 0 : (c > W - 1 ? W - 1 : c);
      return (T)input[(size_t)rc * width + (size_t)cc];
    };
    auto color_at = [&](int r, int c) -> int { return cf.color_at((size_t)((r % 48 + 48) % 48), (size_t)((c % 48 + 48) % 48)); };

    const int bottom = std::min(top + 128, H + 16);
    const int right_edge = std::min(left + 128, W + 16);
    const int rr1 = bottom - top, cc1 = right_edge - left;
    const int rrmin = top < 0 ? 16 : 0, ccmin = left < 0 ? 16 : 0;
    const int rrmax = bottom > H ? H - top : rr1;
    const int ccmax = right_edge > W ? W - left : cc1;

    auto put = [&](int idx, T v) { cfa[idx] = v; rgbgreen[idx] = v; };
    if (rrmin > 0)
      for (int rr = 0; rr < 16; ++rr)
        for (int cc = ccmin; cc < ccmax; ++cc) put(rr * TS + cc, get_in(32 - rr + top, cc + left));
    for (int rr = rrmin; rr < rrmax; ++rr)
      for (int cc = ccmin; cc < ccmax; ++cc) put(rr * TS + cc, get_in(rr + top, cc + left));
    if (rrmax < rr1)
      for (int rr = 0; rr < 16; ++rr)
        for (int cc = ccmin; cc < ccmax; ++cc) put((rrmax + rr) * TS + cc, get_in(H - rr - 2, left + cc));
    if (ccmin > 0)
      for (int rr = rrmin; rr < rrmax; ++rr)
        for (int cc = 0; cc < 16; ++cc) put(rr * TS + cc, get_in(rr + top, 32 - cc + left));
    if (ccmax < cc1)
      for (int rr = rrmin; rr < rrmax; ++rr)
        for (int cc = 0; cc < 16; ++cc) put(rr * TS + ccmax + cc, get_in(top + rr, W - cc - 2));
    if (rrmin > 0 && ccmin > 0)
      for (int rr = 0; rr < 16; ++rr)
        for (int cc = 0; cc < 16; ++cc) put(rr * TS + cc, get_in(32 - rr, 32 - cc));
    if (rrmax < rr1 && ccmax < cc1)
      for (int rr = 0; rr < 16; ++rr)
        for (int cc = 0; cc < 16; ++cc) put((rrmax + rr) * TS + ccmax + cc, get_in(H - rr - 2, W - cc - 2));
    if (rrmin > 0 && ccmax < cc1)
      for (int rr = 0; rr < 16; ++rr)
        for (int cc = 0; cc < 16; ++cc) put(rr * TS + ccmax + cc, get_in(32 - rr, W - cc - 2));
    if (rrmax < rr1 && ccmin > 0)
      for (int rr = 0; rr < 16; ++rr)
        for (int cc = 0; cc < 16; ++cc) put((rrmax + rr) * TS + cc, get_in(H - rr - 2, 32 - cc));

    // gradients (demosaic.rs:612-622)
    for (int rr = 2; rr < rr1 - 2; ++rr)
      for (int cc = 2; cc < cc1 - 2; ++cc) {
        int i = rr * TS + cc;
        T delh = std::fabs(cfa[i + 1] - cfa[i - 1]);
        T delv = std::fabs(cfa[i + V1] - cfa[i - V1]);
        dirwts0[i] = eps + std::fabs(cfa[i + V2] - cfa[i]) + std::fabs(cfa[i] - cfa[i - V2]) + delv;
        dirwts1[i] = eps + std::fabs(cfa[i + 2] - cfa[i]) +
                     std::fabs(cfa[i] - cfa[g_amaze_hooks.plant == 1 ? i + 2 : i - 2]) + delh;
        delhvsqsum[i] = sqr(delh) + sqr(delv);
      }
    trace_full(top, left, 0, dirwts0);
    trace_full(top, left, 1, dirwts1);
    trace_full(top, left, 2, delhvsqsum);

    const T clip_pt8 = clip_pt * (T)0.8f;

    // colour differences (demosaic.rs:626-698)
    for (int rr = 4; rr < rr1 - 4; ++rr) {
      bool fcswitch = (color_at(top + rr, left + 4) & 1) != 0;
      for (int cc = 4; cc < cc1 - 4; ++cc) {
        int i = rr * TS + cc;
        T cru = cfa[i - V1] * (dirwts0[i - V2] + dirwts0[i]) /
                rmax(dirwts0[i - V2] * (eps + cfa[i]) + dirwts0[i] * (eps + cfa[i - V2]), (T)1e-10f);
        T crd = cfa[i + V1] * (dirwts0[i + V2] + dirwts0[i]) /
                rmax(dirwts0[i + V2] * (eps + cfa[i]) + dirwts0[i] * (eps + cfa[i + V2]), (T)1e-10f);
        T crl = cfa[i - 1] * (dirwts1[i - 2] + dirwts1[i]) /
                rmax(dirwts1[g_amaze_hooks.plant == 2 ? i + 2 : i - 2] * (eps + cfa[i]) +
                         dirwts1[i] * (eps + cfa[i - 2]),
                     (T)1e-10f);
        T crr = cfa[i + 1] * (dirwts1[i + 2] + dirwts1[i]) /
                rmax(dirwts1[i + 2] * (eps + cfa[i]) + dirwts1[i] * (eps + cfa[i + 2]), (T)1e-10f);
        T guha = cfa[i - V1] + (cfa[i] - cfa[i - V2]) * (T)0.5f;
        T gdha = cfa[i + V1] + (cfa[i] - cfa[i + V2]) * (T)0.5f;
        T glha = cfa[i - 1] + (cfa[i] - cfa[i - 2]) * (T)0.5f;
        T grha = cfa[i + 1] + (cfa[i] - cfa[i + 2]) * (T)0.5f;
        T guar = std::fabs((T)1 - cru) < arthresh ? cfa[i] * cru : guha;
        T gdar = std::fabs((T)1 - crd) < arthresh ? cfa[i] * crd : gdha;
        T glar = std::fabs((T)1 - crl) < arthresh ? cfa[i] * crl : glha;
        T grar = std::fabs((T)1 - crr) < arthresh ? cfa[i] * crr : grha;
        T hwt = dirwts1[i - 1] / rmax(dirwts1[i - 1] + dirwts1[i + 1], (T)1e-10f);
        T vwt = dirwts0[i - V1] / rmax(dirwts0[i + V1] + dirwts0[i - V1], (T)1e-10f);
        T gintvha = vwt * gdha + ((T)1 - vwt) * guha;
        T ginthha = hwt * grha + ((T)1 - hwt) * glha;
        T vcd_v, hcd_v, vcdalt_v, hcdalt_v;
        if (fcswitch) {
          vcd_v = cfa[i] - (vwt * gdar + ((T)1 - vwt) * guar);
          hcd_v = cfa[i] - (hwt * grar + ((T)1 - hwt) * glar);
          vcdalt_v = cfa[i] - gintvha;
          hcdalt_v = cfa[i] - ginthha;
        } else {
          vcd_v = (vwt * gdar + ((T)1 - vwt) * guar) - cfa[i];
          hcd_v = (hwt * grar + ((T)1 - hwt) * glar) - cfa[i];
          vcdalt_v = gintvha - cfa[i];
          hcdalt_v = ginthha - cfa[i];
        }
        fcswitch = !fcswitch;
        T gu = guar, gd = gdar, gl = glar, gr = grar;
        if (cfa[i] > clip_pt8 || gintvha > clip_pt8 || ginthha > clip_pt8) {
          vcd_v = vcdalt_v;
          hcd_v = hcdalt_v;
          gu = guha; gd = gdha; gl = glha; gr = grha;
        }
        vcd[i] = vcd_v;
        hcd[i] = hcd_v;
        vcdalt[i] = vcdalt_v;
        hcdalt[i] = hcdalt_v;
        dgintv[i] = rmin(sqr(guha - gdha), sqr(gu - gd));
        dginth[i] = rmin(sqr(glha - grha), sqr(gl - gr));
      }
    }
    trace_full(top, left, 3, vcd);
    trace_full(top, left, 4, hcd);
    trace_full(top, left, 5, vcdalt);
    trace_full(top, left, 6, hcdalt);
    trace_full(top, left, 7, dgintv);
    trace_full(top, left, 8, dginth);

    // variance-based choice and clamp, in place (demosaic.rs:700-786)
    for (int rr = 4; rr < rr1 - 4; ++rr)
      for (int cc = 4; cc < cc1 - 4; ++cc) {
        int i = rr * TS + cc;
        bool c = (color_at(top + rr, left + cc) & 1) != 0;
        T hcdvar = (T)3 * (sqr(hcd[i - 2]) + sqr(hcd[i]) + sqr(hcd[i + 2])) - sqr(hcd[i - 2] + hcd[i] + hcd[i + 2]);
        T hcdaltvar = (T)3 * (sqr(hcdalt[i - 2]) + sqr(hcdalt[i]) + sqr(hcdalt[i + 2])) -
                      sqr(hcdalt[i - 2] + hcdalt[i] + hcdalt[i + 2]);
        T vcdvar = (T)3 * (sqr(vcd[i - V2]) + sqr(vcd[i]) + sqr(vcd[i + V2])) - sqr(vcd[i - V2] + vcd[i] + vcd[i + V2]);
        T vcdaltvar = (T)3 * (sqr(vcdalt[i - V2]) + sqr(vcdalt[i]) + sqr(vcdalt[i + V2])) -
                      sqr(vcdalt[i - V2] + vcdalt[i] + vcdalt[i + V2]);
        if (hcdaltvar < hcdvar) hcd[i] = hcdalt[i];
        if (vcdaltvar < vcdvar) vcd[i] = vcdalt[i];
        if (c) {
          T ginth = -hcd[i] + cfa[i];
          T gintv = -vcd[i] + cfa[i];
          if (hcd[i] > 0) {
            if ((T)3 * hcd[i] > (ginth + cfa[i])) {
              hcd[i] = -ulim(ginth, cfa[i - 1], cfa[i + 1]) + cfa[i];
            } else {
              T hwt = (T)1 - (T)3 * hcd[i] / (eps + ginth + cfa[i]);
              hcd[i] = hwt * hcd[i] + ((T)1 - hwt) * (-ulim(ginth, cfa[i - 1], cfa[i + 1]) + cfa[i]);
            }
          }
          if (vcd[i] > 0) {
            if ((T)3 * vcd[i] > (gintv + cfa[i])) {
              vcd[i] = -ulim(gintv, cfa[i - V1], cfa[i + V1]) + cfa[i];
            } else {
              T vwt = (T)1 - (T)3 * vcd[i] / (eps + gintv + cfa[i]);
              vcd[i] = vwt * vcd[i] + ((T)1 - vwt) * (-ulim(gintv, cfa[i - V1], cfa[i + V1]) + cfa[i]);
            }
          }
          if (ginth > clip_pt) hcd[i] = -ulim(ginth, cfa[i - 1], cfa[i + 1]) + cfa[i];
          if (gintv > clip_pt) vcd[i] = -ulim(gintv, cfa[i - V1], cfa[i + V1]) + cfa[i];
        } else {
          T ginth = hcd[i] + cfa[i];
          T gintv = vcd[i] + cfa[i];
          if (hcd[i] < 0) {
            if ((T)3 * hcd[i] < -(ginth + cfa[i])) {
              hcd[i] = ulim(ginth, cfa[g_amaze_hooks.plant == 3 ? i + 1 : i - 1], cfa[i + 1]) - cfa[i];
            } else {
              T hwt = (T)1 + (T)3 * hcd[i] / (eps + ginth + cfa[i]);
              hcd[i] = hwt * hcd[i] + ((T)1 - hwt) * (ulim(ginth, cfa[i - 1], cfa[i + 1]) - cfa[i]);
            }
          }
          if (vcd[i] < 0) {
            if ((T)3 * vcd[i] < -(gintv + cfa[i])) {
              vcd[i] = ulim(gintv, cfa[i - V1], cfa[i + V1]) - cfa[i];
            } else {
              T vwt = (T)1 + (T)3 * vcd[i] / (eps + gintv + cfa[i]);
              vcd[i] = vwt * vcd[i] + ((T)1 - vwt) * (ulim(gintv, cfa[i - V1], cfa[i + V1]) - cfa[i]);
            }
          }
          if (ginth > clip_pt) hcd[i] = ulim(ginth, cfa[i - 1], cfa[i + 1]) - cfa[i];
          if (gintv > clip_pt) vcd[i] = ulim(gintv, cfa[i - V1], cfa[i + V1]) - cfa[i];
          cddiffsq[i] = sqr(vcd[i] - hcd[i]);
        }
      }

    trace_full(top, left, 9, vcd);
    trace_full(top, left, 10, hcd);
    trace_full(top, left, 11, cddiffsq);

    // horizontal/vertical weights (demosaic.rs:788-828)
    for (int rr = 6; rr < rr1 - 6; ++rr)
      for (int cc = 6 + (color_at(top + rr, left + 2) & 1); cc < cc1 - 6; cc += 2) {
        int i = rr * TS + cc;
        T uave = vcd[i] + vcd[i - V1] + vcd[i - V2] + vcd[i - V3];
        T dave = vcd[i] + vcd[i + V1] + vcd[i + V2] + vcd[i + V3];
        T lave = hcd[i] + hcd[i - 1] + hcd[i - 2] + hcd[i - 3];
        T rave = hcd[i] + hcd[i + 1] + hcd[i + 2] + hcd[i + 3];
        T dgrbvvaru = sqr(vcd[i] - uave) + sqr(vcd[i - V1] - uave) + sqr(vcd[i - V2] - uave) + sqr(vcd[i - V3] - uave);
        T dgrbvvard = sqr(vcd[i] - dave) + sqr(vcd[i + V1] - dave) + sqr(vcd[i + V2] - dave) + sqr(vcd[i + V3] - dave);
        T dgrbhvarl = sqr(hcd[i] - lave) + sqr(hcd[i - 1] - lave) + sqr(hcd[i - 2] - lave) + sqr(hcd[i - 3] - lave);
        T dgrbhvarr = sqr(hcd[i] - rave) + sqr(hcd[i + 1] - rave) + sqr(hcd[i + 2] - rave) + sqr(hcd[i + 3] - rave);
        T hwt = dirwts1[i - 1] / rmax(dirwts1[i - 1] + dirwts1[i + 1], (T)1e-10f);
        T vwt = dirwts0[i - V1] / rmax(dirwts0[i + V1] + dirwts0[i - V1], (T)1e-10f);
        T vcdvar = epssq + vwt * dgrbvvard + ((T)1 - vwt) * dgrbvvaru;
        T hcdvar = epssq + hwt * dgrbhvarr + ((T)1 - hwt) * dgrbhvarl;
        T dgrbvvaru1 = dgintv[i] + dgintv[i - V1] + dgintv[i - V2];
        T dgrbvvard1 = dgintv[i] + dgintv[i + V1] + dgintv[i + V2];
        T dgrbhvarl1 = dginth[i] + dginth[i - 1] + dginth[i - 2];
        T dgrbhvarr1 = dginth[i] + dginth[i + 1] + dginth[i + 2];
        T vcdvar1 = epssq + vwt * dgrbvvard1 + ((T)1 - vwt) * dgrbvvaru1;
        T hcdvar1 = epssq + hwt * dgrbhvarr1 + ((T)1 - hwt) * dgrbhvarl1;
        T varwt = hcdvar / rmax(vcdvar + hcdvar, (T)1e-10f);
        T diffwt = hcdvar1 / rmax(vcdvar1 + hcdvar1, (T)1e-10f);
        if (g_amaze_hooks.plant == 4) diffwt = hcdvar1 / rmax(vcdvar1 + hcdvar1 + dginth[i + 1] * (T)1e-3f, (T)1e-10f);
        if (((T)0.5f - varwt) * ((T)0.5f - diffwt) > 0 && std::fabs((T)0.5f - diffwt) < std::fabs((T)0.5f - varwt))
          hvwt[i >> 1] = varwt;
        else
          hvwt[i >> 1] = diffwt;
      }
    trace_half(top, left, 12, hvwt, cf);

    // Nyquist texture test (demosaic.rs:830-847)
    for (int rr = 6; rr < rr1 - 6; ++rr)
      for (int cc = 6 + (color_at(top + rr, left + 2) & 1); cc < cc1 - 6; cc += 2) {
        int i = rr * TS + cc;
        nyqutest[i >> 1] =
            (gaussodd[0] * cddiffsq[i] +
             gaussodd[1] * (cddiffsq[i - M1] + cddiffsq[i - P1] + cddiffsq[i + P1] + cddiffsq[i + M1]) +
             gaussodd[2] * (cddiffsq[i - V2] + cddiffsq[i - 2] + cddiffsq[i + 2] + cddiffsq[i + V2]) +
             gaussodd[3] * (cddiffsq[i - M2] + cddiffsq[i - P2] + cddiffsq[i + P2] + cddiffsq[i + M2])) -
            (gaussgrad[0] * delhvsqsum[i] +
             gaussgrad[1] * (delhvsqsum[i - V1] + delhvsqsum[i + 1] + delhvsqsum[i - 1] + delhvsqsum[i + V1]) +
             gaussgrad[2] * (delhvsqsum[i - M1] + delhvsqsum[i - P1] + delhvsqsum[i + P1] + delhvsqsum[i + M1]) +
             gaussgrad[3] * (delhvsqsum[i - V2] + delhvsqsum[i - 2] + delhvsqsum[i + 2] + delhvsqsum[i + V2]) +
             gaussgrad[4] * (delhvsqsum[i - V2 - 1] + delhvsqsum[i - V2 + 1] + delhvsqsum[i - TS - 2] +
                             delhvsqsum[i - TS + 2] + delhvsqsum[i + TS - 2] + delhvsqsum[i + TS + 2] +
                             delhvsqsum[i + V2 - 1] + delhvsqsum[i + V2 + 1]) +
             gaussgrad[5] * (delhvsqsum[i - M2] + delhvsqsum[i - P2] + delhvsqsum[i + P2] + delhvsqsum[i + M2]));
      }

    trace_half(top, left, 13, nyqutest, cf);
    int nystartrow = 0, nyendrow = 0, nystartcol = TS + 1, nyendcol = 0;
    for (int rr = 6; rr < rr1 - 6; ++rr)
      for (int cc = 6 + (color_at(top + rr, left + 2) & 1); cc < cc1 - 6; cc += 2) {
        int i = rr * TS + cc;
        if (nyqutest[i >> 1] > 0) {
          nyquist[i >> 1] = 1;
          if (nystartrow == 0) nystartrow = rr;
          nyendrow = rr;
          nystartcol = std::min(nystartcol, cc);
          nyendcol = std::max(nyendcol, cc);
        }
      }
    trace_half(top, left, 14, nyquist, cf);
    const bool do_nyquist = nystartrow != nyendrow && nystartcol != nyendcol;
    if (float* d = trace_plane(top, left, 15)) d[0] = do_nyquist ? 1.0f : 0.0f;
    if (do_nyquist) {
      nyendrow += 1;
      nyendcol += 1;
      nystartcol -= nystartcol & 1;
      nystartrow = std::max(nystartrow, 8);
      nyendrow = std::min(nyendrow, rr1 - 8);
      nystartcol = std::max(nystartcol, 8);
      nyendcol = std::min(nyendcol, cc1 - 8);
      for (int k = 4 * TSH; k < (TS - 8) * TSH; ++k) nyquist2[k] = 0;
      for (int rr = nystartrow; rr < nyendrow; ++rr)
        for (int i = rr * TS + nystartcol + (color_at(top + rr, left + 2) & 1); i < rr * TS + nyendcol; i += 2) {
          unsigned t = (unsigned)nyquist[(i - V2) >> 1] + nyquist[(i - M1) >> 1] + nyquist[(i - P1) >> 1] +
                       nyquist[(i - 2) >> 1] + nyquist[(i + 2) >> 1] + nyquist[(i + P1) >> 1] +
                       nyquist[(i + M1) >> 1] + nyquist[(i + V2) >> 1];
          nyquist2[i >> 1] = t > 4 ? 1 : (t < 4 ? 0 : nyquist[i >> 1]);
        }
      for (int rr = nystartrow; rr < nyendrow; ++rr)
        for (int i = rr * TS + nystartcol + (color_at(top + rr, left + 2) & 1); i < rr * TS + nyendcol; i += 2) {
          if (!nyquist2[i >> 1]) continue;
          T sumcfa = 0, sumh = 0, sumv = 0, sumsqh = 0, sumsqv = 0, areawt = 0;
          for (int a = -6; a < 7; a += 2) {
            int i1 = i + a * TS - 6;
            for (int b = -6; b < 7; b += 2, i1 += 2) {
              if (nyquist2[i1 >> 1]) {
                T ct = cfa[i1];
                sumcfa += ct;
                sumh += cfa[i1 - 1] + cfa[i1 + 1];
                sumv += cfa[i1 - V1] + cfa[i1 + V1];
                sumsqh += sqr(ct - cfa[i1 - 1]) + sqr(ct - cfa[i1 + 1]);
                sumsqv += sqr(ct - cfa[i1 - V1]) + sqr(ct - cfa[i1 + V1]);
                areawt += 1;
              }
            }
          }
          sumh = sumcfa - sumh * (T)0.5f;
          sumv = sumcfa - sumv * (T)0.5f;
          areawt = areawt * (T)0.5f;
          T hcdvar = epssq + std::fabs(areawt * sumsqh - sumh * sumh);
          T vcdvar = epssq + std::fabs(areawt * sumsqv - sumv * sumv);
          hvwt[i >> 1] = hcdvar / (vcdvar + hcdvar);
        }
    }

    // G at R/B sites (demosaic.rs:949-971)
    for (int rr = 8; rr < rr1 - 8; ++rr)
      for (int i = rr * TS + 8 + (color_at(top + rr, left + 2) & 1); i < rr * TS + cc1 - 8; i += 2) {
        T hvwtalt = (hvwt[(i - M1) >> 1] + hvwt[(i - P1) >> 1] + hvwt[(i + P1) >> 1] + hvwt[(i + M1) >> 1]) / (T)4;
        hvwt[i >> 1] = std::fabs((T)0.5f - hvwt[i >> 1]) < std::fabs((T)0.5f - hvwtalt) ? hvwtalt : hvwt[i >> 1];
        dgrb0[i >> 1] = interpolatef(hvwt[i >> 1], vcd[i], hcd[i]);
        rgbgreen[i] = cfa[i] + dgrb0[i >> 1];
        dgrb2_h[i >> 1] = nyquist2[i >> 1] ? sqr(rgbgreen[i] - (rgbgreen[i - 1] + rgbgreen[i + 1]) * (T)0.5f) : (T)0;
        dgrb2_v[i >> 1] = nyquist2[i >> 1] ? sqr(rgbgreen[i] - (rgbgreen[i - V1] + rgbgreen[i + V1]) * (T)0.5f) : (T)0;
      }

    if (do_nyquist)  // demosaic.rs:975-998
      for (int rr = nystartrow; rr < nyendrow; ++rr)
        for (int i = rr * TS + nystartcol + (color_at(top + rr, left + 2) & 1); i < rr * TS + nyendcol; i += 2) {
          if (!nyquist2[i >> 1]) continue;
          T gvarh = epssq + (gquinc[0] * dgrb2_h[i >> 1] +
                             gquinc[1] * (dgrb2_h[(i - M1) >> 1] + dgrb2_h[(i - P1) >> 1] + dgrb2_h[(i + P1) >> 1] + dgrb2_h[(i + M1) >> 1]) +
                             gquinc[2] * (dgrb2_h[(i - V2) >> 1] + dgrb2_h[(i - 2) >> 1] + dgrb2_h[(i + 2) >> 1] + dgrb2_h[(i + V2) >> 1]) +
                             gquinc[3] * (dgrb2_h[(i - M2) >> 1] + dgrb2_h[(i - P2) >> 1] + dgrb2_h[(i + P2) >> 1] + dgrb2_h[(i + M2) >> 1]));
          T gvarv = epssq + (gquinc[0] * dgrb2_v[i >> 1] +
                             gquinc[1] * (dgrb2_v[(i - M1) >> 1] + dgrb2_v[(i - P1) >> 1] + dgrb2_v[(i + P1) >> 1] + dgrb2_v[(i + M1) >> 1]) +
                             gquinc[2] * (dgrb2_v[(i - V2) >> 1] + dgrb2_v[(i - 2) >> 1] + dgrb2_v[(i + 2) >> 1] + dgrb2_v[(i + V2) >> 1]) +
                             gquinc[3] * (dgrb2_v[(i - M2) >> 1] + dgrb2_v[(i - P2) >> 1] + dgrb2_v[(i + P2) >> 1] + dgrb2_v[(i + M2) >> 1]));
          dgrb0[i >> 1] = (hcd[i] * gvarv + vcd[i] * gvarh) / (gvarv + gvarh);
          rgbgreen[i] = cfa[i] + dgrb0[i >> 1];
        }

    // diagonal setup (demosaic.rs:1001-1019)
    for (int rr = 6; rr < rr1 - 6; ++rr) {
      const bool even_row = (color_at(top + rr, left + 2) & 1) == 0;
      for (int cc = 6; cc < cc1 - 6; cc += 2) {
        int i = rr * TS + cc;
        if (even_row) {
          delp[i >> 1] = std::fabs(cfa[i - P1] - cfa[i + P1]);
          delm[i >> 1] = std::fabs(cfa[i + M1] - cfa[i - M1]);
          dgrbsq1p[i >> 1] = sqr(cfa[i + 1] - cfa[i + 1 - P1]) + sqr(cfa[i + 1] - cfa[i + 1 + P1]);
          dgrbsq1m[i >> 1] = sqr(cfa[i + 1] - cfa[i + 1 + M1]) + sqr(cfa[i + 1] - cfa[i + 1 - M1]);
        } else {
          dgrbsq1p[i >> 1] = sqr(cfa[i] - cfa[i - P1]) + sqr(cfa[i] - cfa[i + P1]);
          dgrbsq1m[i >> 1] = sqr(cfa[i] - cfa[i - M1]) + sqr(cfa[i] - cfa[i + M1]);
          delp[i >> 1] = std::fabs(cfa[i + 1 - P1] - cfa[i + 1 + P1]);
          delm[i >> 1] = std::fabs(cfa[i + 1 - M1] - cfa[i + 1 + M1]);
        }
      }
    }

    // diagonal interpolation (demosaic.rs:1023-1107)
    for (int rr = 8; rr < rr1 - 8; ++rr)
      for (int cc = 8 + (color_at(top + rr, left + 2) & 1); cc < cc1 - 8; cc += 2) {
        int i = rr * TS + cc, i1 = i >> 1;
        T crse = (cfa[i + M1] * (T)2) / (eps + cfa[i] + cfa[i + M2]);
        T crnw = (cfa[i - M1] * (T)2) / (eps + cfa[i] + cfa[i - M2]);
        T crne = (cfa[i - P1] * (T)2) / (eps + cfa[i] + cfa[i - P2]);
        T crsw = (cfa[i + P1] * (T)2) / (eps + cfa[i] + cfa[i + P2]);
        T rbse = std::fabs((T)1 - crse) < arthresh ? cfa[i] * crse : cfa[i + M1] + (cfa[i] - cfa[i + M2]) * (T)0.5f;
        T rbnw = std::fabs((T)1 - crnw) < arthresh ? cfa[i] * crnw : cfa[i - M1] + (cfa[i] - cfa[i - M2]) * (T)0.5f;
        T rbne = std::fabs((T)1 - crne) < arthresh ? cfa[i] * crne : cfa[i - P1] + (cfa[i] - cfa[i - P2]) * (T)0.5f;
        T rbsw = std::fabs((T)1 - crsw) < arthresh ? cfa[i] * crsw : cfa[i + P1] + (cfa[i] - cfa[i + P2]) * (T)0.5f;
        T wtse = eps + delm[i1] + delm[(i + M1) >> 1] + delm[(i + M2) >> 1];
        T wtnw = eps + delm[i1] + delm[(i - M1) >> 1] + delm[(i - M2) >> 1];
        T wtne = eps + delp[i1] + delp[(i - P1) >> 1] + delp[(i - P2) >> 1];
        T wtsw = eps + delp[i1] + delp[(i + P1) >> 1] + delp[(i + P2) >> 1];
        rbm[i1] = (wtse * rbnw + wtnw * rbse) / (wtse + wtnw);
        rbp[i1] = (wtne * rbsw + wtsw * rbne) / (wtne + wtsw);
        T rbvarm = epssq + gausseven[0] * (dgrbsq1m[(i - V1) >> 1] + dgrbsq1m[(i - 1) >> 1] + dgrbsq1m[(i + 1) >> 1] + dgrbsq1m[(i + V1) >> 1]) +
                   gausseven[1] * (dgrbsq1m[(i - V2 - 1) >> 1] + dgrbsq1m[(i - V2 + 1) >> 1] + dgrbsq1m[(i - 2 - V1) >> 1] +
                                   dgrbsq1m[(i + 2 - V1) >> 1] + dgrbsq1m[(i - 2 + V1) >> 1] + dgrbsq1m[(i + 2 + V1) >> 1] +
                                   dgrbsq1m[(i + V2 - 1) >> 1] + dgrbsq1m[(i + V2 + 1) >> 1]);
        T rbvarp = epssq + gausseven[0] * (dgrbsq1p[(i - V1) >> 1] + dgrbsq1p[(i - 1) >> 1] + dgrbsq1p[(i + 1) >> 1] + dgrbsq1p[(i + V1) >> 1]) +
                   gausseven[1] * (dgrbsq1p[(i - V2 - 1) >> 1] + dgrbsq1p[(i - V2 + 1) >> 1] + dgrbsq1p[(i - 2 - V1) >> 1] +
                                   dgrbsq1p[(i + 2 - V1) >> 1] + dgrbsq1p[(i - 2 + V1) >> 1] + dgrbsq1p[(i + 2 + V1) >> 1] +
                                   dgrbsq1p[(i + V2 - 1) >> 1] + dgrbsq1p[(i + V2 + 1) >> 1]);
        pmwt[i1] = rbvarm / (rbvarp + rbvarm);
        if (rbp[i1] < cfa[i]) {
          if (rbp[i1] * (T)2 < cfa[i]) {
            rbp[i1] = ulim(rbp[i1], cfa[i - P1], cfa[i + P1]);
          } else {
            T pwt = (cfa[i] - rbp[i1]) * (T)2 / (eps + rbp[i1] + cfa[i]);
            rbp[i1] = pwt * rbp[i1] + ((T)1 - pwt) * ulim(rbp[i1], cfa[i - P1], cfa[i + P1]);
          }
        }
        if (rbm[i1] < cfa[i]) {
          if (rbm[i1] * (T)2 < cfa[i]) {
            rbm[i1] = ulim(rbm[i1], cfa[i - M1], cfa[i + M1]);
          } else {
            T mwt = (cfa[i] - rbm[i1]) * (T)2 / (eps + rbm[i1] + cfa[i]);
            rbm[i1] = mwt * rbm[i1] + ((T)1 - mwt) * ulim(rbm[i1], cfa[i - M1], cfa[i + M1]);
          }
        }
        if (rbp[i1] > clip_pt) rbp[i1] = ulim(rbp[i1], cfa[i - P1], cfa[i + P1]);
        if (rbm[i1] > clip_pt) rbm[i1] = ulim(rbm[i1], cfa[i - M1], cfa[i + M1]);
      }

    for (int rr = 10; rr < rr1 - 10; ++rr)  // demosaic.rs:1109-1123
      for (int cc = 10 + (color_at(top + rr, left + 2) & 1); cc < cc1 - 10; cc += 2) {
        int i = rr * TS + cc, i1 = i >> 1;
        T pmwtalt = (pmwt[(i - M1) >> 1] + pmwt[(i - P1) >> 1] + pmwt[(i + P1) >> 1] + pmwt[(i + M1) >> 1]) / (T)4;
        if (std::fabs((T)0.5f - pmwt[i1]) < std::fabs((T)0.5f - pmwtalt)) pmwt[i1] = pmwtalt;
        rbint[i1] = (cfa[i] + rbm[i1] * ((T)1 - pmwt[i1]) + rbp[i1] * pmwt[i1]) * (T)0.5f;
      }

    for (int rr = 12; rr < rr1 - 12; ++rr)  // demosaic.rs:1125-1194
      for (int cc = 12 + (color_at(top + rr, left + 2) & 1); cc < cc1 - 12; cc += 2) {
        int i = rr * TS + cc, i1 = i >> 1;
        if (!(std::fabs((T)0.5f - pmwt[i >> 1]) >= std::fabs((T)0.5f - hvwt[i >> 1]))) continue;
        T cru = cfa[i - V1] * (T)2 / (eps + rbint[i1] + rbint[i1 - V1]);
        T crd = cfa[i + V1] * (T)2 / (eps + rbint[i1] + rbint[i1 + V1]);
        T crl = cfa[i - 1] * (T)2 / (eps + rbint[i1] + rbint[i1 - 1]);
        T crr = cfa[i + 1] * (T)2 / (eps + rbint[i1] + rbint[i1 + 1]);
        T gu = std::fabs((T)1 - cru) < arthresh ? rbint[i1] * cru : cfa[i - V1] + (rbint[i1] - rbint[i1 - V1]) * (T)0.5f;
        T gd = std::fabs((T)1 - crd) < arthresh ? rbint[i1] * crd : cfa[i + V1] + (rbint[i1] - rbint[i1 + V1]) * (T)0.5f;
        T gl = std::fabs((T)1 - crl) < arthresh ? rbint[i1] * crl : cfa[i - 1] + (rbint[i1] - rbint[i1 - 1]) * (T)0.5f;
        T gr = std::fabs((T)1 - crr) < arthresh ? rbint[i1] * crr : cfa[i + 1] + (rbint[i1] - rbint[i1 + 1]) * (T)0.5f;
        T gintv = (dirwts0[i - V1] * gd + dirwts0[i + V1] * gu) / rmax(dirwts0[i + V1] + dirwts0[i - V1], (T)1e-10f);
        T ginth = (dirwts1[i - 1] * gr + dirwts1[i + 1] * gl) / rmax(dirwts1[i - 1] + dirwts1[i + 1], (T)1e-10f);
        if (gintv < rbint[i1]) {
          if ((T)2 * gintv < rbint[i1]) {
            gintv = ulim(gintv, cfa[i - V1], cfa[i + V1]);
          } else {
            T vwt = (T)2 * (rbint[i1] - gintv) / (eps + gintv + rbint[i1]);
            gintv = vwt * gintv + ((T)1 - vwt) * ulim(gintv, cfa[i - V1], cfa[i + V1]);
          }
        }
        if (ginth < rbint[i1]) {
          if ((T)2 * ginth < rbint[i1]) {
            ginth = ulim(ginth, cfa[i - 1], cfa[i + 1]);
          } else {
            T hwt = (T)2 * (rbint[i1] - ginth) / (eps + ginth + rbint[i1]);
            ginth = hwt * ginth + ((T)1 - hwt) * ulim(ginth, cfa[i - 1], cfa[i + 1]);
          }
        }
        if (ginth > clip_pt) ginth = ulim(ginth, cfa[i - 1], cfa[i + 1]);
        if (gintv > clip_pt) gintv = ulim(gintv, cfa[i - V1], cfa[i + V1]);
        rgbgreen[i] = ginth * ((T)1 - hvwt[i1]) + gintv * hvwt[i1];
        dgrb0[i1] = rgbgreen[i] - cfa[i];
      }

    // chroma (demosaic.rs:1196-1252); (ey, ex) is the R site of the unshifted pattern
    int ex, ey;
    if (color_at(0, 0) == 1) {
      if (color_at(0, 1) == 0) { ex = 1; ey = 0; } else { ex = 0; ey = 1; }
    } else {
      if (color_at(0, 0) == 0) { ex = 0; ey = 0; } else { ex = 1; ey = 1; }
    }
    for (int rr = 13 - ey; rr < rr1 - 12; rr += 2)
      for (int i1 = (rr * TS + 13 - ex) >> 1; i1 < ((rr * TS + cc1 - 12) >> 1); ++i1) {
        dgrb1[i1] = dgrb0[i1];
        dgrb0[i1] = 0;
      }
    for (int rr = 14; rr < rr1 - 14; ++rr)
      for (int cc = 14 + (color_at(top + rr, left + 2) & 1); cc < cc1 - 14; cc += 2) {
        int i = rr * TS + cc;
        int c = 1 - color_at(top + rr, left + cc) / 2;
        std::vector<T>& d = c == 0 ? dgrb0 : dgrb1;
        T wtnw = (T)1 / (eps + std::fabs(d[(i - M1) >> 1] - d[(i + M1) >> 1]) + std::fabs(d[(i - M1) >> 1] - d[(i - M3) >> 1]) +
                         std::fabs(d[(i + M1) >> 1] - d[(i - M3) >> 1]));
        T wtne = (T)1 / (eps + std::fabs(d[(i - P1) >> 1] - d[(i + P1) >> 1]) + std::fabs(d[(i - P1) >> 1] - d[(i - P3) >> 1]) +
                         std::fabs(d[(i + P1) >> 1] - d[(i - P3) >> 1]));
        T wtsw = (T)1 / (eps + std::fabs(d[(i + P1) >> 1] - d[(i - P1) >> 1]) + std::fabs(d[(i + P1) >> 1] - d[(i + M3) >> 1]) +
                         std::fabs(d[(i - P1) >> 1] - d[(i + P3) >> 1]));
        T wtse = (T)1 / (eps + std::fabs(d[(i + M1) >> 1] - d[(i - M1) >> 1]) + std::fabs(d[(i + M1) >> 1] - d[(i + P3) >> 1]) +
                         std::fabs(d[(i - M1) >> 1] - d[(i + M3) >> 1]));
        T val = (wtnw * ((T)1.325f * d[(i - M1) >> 1] - (T)0.175f * d[(i - M3) >> 1] - (T)0.075f * d[(i - M1 - 2) >> 1] -
                         (T)0.075f * d[(i - M1 - V2) >> 1]) +
                 wtne * ((T)1.325f * d[(i - P1) >> 1] - (T)0.175f * d[(i - P3) >> 1] - (T)0.075f * d[(i - P1 + 2) >> 1] -
                         (T)0.075f * d[(i - P1 + V2) >> 1]) +
                 wtsw * ((T)1.325f * d[(i + P1) >> 1] - (T)0.175f * d[(i + P3) >> 1] - (T)0.075f * d[(i + P1 - 2) >> 1] -
                         (T)0.075f * d[(i - M1) >> 1]) +
                 wtse * ((T)1.325f * d[(i + M1) >> 1] - (T)0.175f * d[(i + M3) >> 1] - (T)0.075f * d[(i + M1 + 2) >> 1] -
                         (T)0.075f * d[(i + M1 + V2) >> 1])) /
                (wtnw + wtne + wtsw + wtse);
        d[i >> 1] = val;
      }

    // output (demosaic.rs:1254-1366)
    auto full = [&](int i, int row, int col) {
      T temp = (T)1 / rmax(hvwt[(i - V1) >> 1] + (T)2 - hvwt[(i + 1) >> 1] - hvwt[(i - 1) >> 1] + hvwt[(i + V1) >> 1], (T)1e-5f);
      T r = clampnan(rgbgreen[i] - (hvwt[(i - V1) >> 1] * dgrb0[(i - V1) >> 1] + ((T)1 - hvwt[(i + 1) >> 1]) * dgrb0[(i + 1) >> 1] +
                                    ((T)1 - hvwt[(i - 1) >> 1]) * dgrb0[(i - 1) >> 1] + hvwt[(i + V1) >> 1] * dgrb0[(i + V1) >> 1]) * temp,
                     (T)0, (T)1);
      T b = clampnan(rgbgreen[i] - (hvwt[(i - V1) >> 1] * dgrb1[(i - V1) >> 1] + ((T)1 - hvwt[(i + 1) >> 1]) * dgrb1[(i + 1) >> 1] +
                                    ((T)1 - hvwt[(i - 1) >> 1]) * dgrb1[(i - 1) >> 1] + hvwt[(i + V1) >> 1] * dgrb1[(i + V1) >> 1]) * temp,
                     (T)0, (T)1);
      T g = clampnan(rgbgreen[i], (T)0, (T)1);
      float* o = out + ((size_t)row * width + (size_t)col) * 3;
      o[0] = (float)r; o[1] = (float)g; o[2] = (float)b;
    };
    auto direct = [&](int i, int row, int col) {
      T r = clampnan(rgbgreen[i] - dgrb0[i >> 1], (T)0, (T)1);
      T b = clampnan(rgbgreen[i] - dgrb1[i >> 1], (T)0, (T)1);
      T g = clampnan(rgbgreen[i], (T)0, (T)1);
      float* o = out + ((size_t)row * width + (size_t)col) * 3;
      o[0] = (float)r; o[1] = (float)g; o[2] = (float)b;
    };
    auto inside = [&](int row, int col) { return col >= 0 && col < W && row >= 0 && row < H; };
    for (int rr = 16; rr < rr1 - 16; ++rr) {
      const int row = rr + top;
      int col = left + 16;
      int i = rr * TS + 16;
      const bool g_first = (color_at(top + rr, left + 2) & 1) == 1;
      while (i < rr * TS + cc1 - 16 - (cc1 & 1)) {
        if (inside(row, col)) { if (g_first) full(i, row, col); else direct(i, row, col); }
        ++i; ++col;
        if (inside(row, col)) { if (g_first) direct(i, row, col); else full(i, row, col); }
        ++i; ++col;
      }
      if ((cc1 & 1) != 0 && inside(row, col)) { if (g_first) full(i, row, col); else direct(i, row, col); }
    }
  }
};

}  // namespace

// The reference indexes past its 128x128 tile buffer when a dimension leaves
// 1..15 pixels in the last tile row/column (demosaic.rs:527-535, 549-556), so
// such sizes are rejected here as they are by the GPU path.
bool amaze_size_supported(size_t width, size_t height) {
  size_t rw = width % 96, rh = height % 96;
  return width >= 2 && height >= 2 && !(rw >= 1 && rw <= 15) && !(rh >= 1 && rh <= 15);
}

template <class T>
void amaze_demosaic(const float* input, size_t width, size_t height, const Cfa& cfa, float* out) {
  std::vector<std::pair<int, int>> tiles;
  for (int top = -16; top < (int)height; top += 96)
    for (int left = -16; left < (int)width; left += 96) tiles.push_back({top, left});
#pragma omp parallel for schedule(dynamic, 4)
  for (long long k = 0; k < (long long)tiles.size(); ++k) {
    AmazeTile<T> t;  // fresh zeroed scratch per tile, as demosaic.rs:470-497 allocates it
    t.run(input, width, height, cfa, tiles[k].first, tiles[k].second, out, (T)1.0f);
  }
}

template void amaze_demosaic<float>(const float*, size_t, size_t, const Cfa&, float*);
template void amaze_demosaic<double>(const float*, size_t, size_t, const Cfa&, float*);

}  // namespace oracle
