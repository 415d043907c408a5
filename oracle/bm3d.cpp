// Oracle (test infrastructure only): BM3D restated from pichromatic/src/bm3d.rs.
//
//   bm3d                  bm3d.rs:34-61      (YCbCr x255, two steps, detail blend, back to RGB)
//   rgb_to_ycbcr / back   bm3d.rs:63-92
//   run_bm3d_step_joint   bm3d.rs:128-238    (per reference patch on the stride grid)
//   hard_threshold        bm3d.rs:240-255
//   wiener_filter         bm3d.rs:257-272
//   block_matching_joint  bm3d.rs:281-365    (early-exit SSD, 1024-candidate cap, sort, pow2 group)
//   DctTables             bm3d.rs:417-461    (f32 cosf tables, Kaiser beta=2 via 10-term I0)
//   AtomicAccumulator     bm3d.rs:463-488    (value*1e5 truncated to i64; order-independent sums)
//   transforms            bm3d.rs:490-614    (2D DCT as matrix products, WHT with 1/sqrt(n))
//   gaussian_blur_1ch     bm3d.rs:648-699
//   chroma_bm3d           bm3d.rs:705-759
//
// The geometry (stride, search window, group cap, candidate cap) and the noise
// model are parameters; the reference's values are the defaults the callers
// pass.  Ties in the match sort are broken by scan order (the reference's
// sort_unstable leaves them unspecified; it is stable below 21 candidates).
#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdint>
#include <vector>

#include "oracle_bm3d.hpp"
#include "oracle_common.hpp"

namespace oracle {

namespace {

template <class T>
struct Tables {
  T dct[64], idct[64], kaiser[64];
  Tables() {
    const T pi = (T)3.14159265358979323846;  // std::f32::consts::PI for T = float
    for (int k = 0; k < 8; ++k)
      for (int n = 0; n < 8; ++n) {
        T c = (T)k * pi / (T)8;
        T val = std::cos(((T)n + (T)0.5) * c);
        T scale = k == 0 ? (T)0.35355339f : (T)0.5;
        dct[k * 8 + n] = val * scale;
      }
    for (int n = 0; n < 8; ++n)
      for (int k = 0; k < 8; ++k) {
        T theta = (pi / (T)8) * ((T)n + (T)0.5) * (T)k;
        T scale = k == 0 ? (T)0.35355339f : (T)0.5;
        idct[n * 8 + k] = scale * std::cos(theta);
      }
    auto i0 = [](T x) {
      T sum = 1, term = 1;
      for (int k = 1; k <= 10; ++k) {
        T t = x / (T)2 / (T)k;
        term *= t * t;
        sum += term;
      }
      return sum;
    };
    const T beta = 2;
    const T i0b = i0(beta);
    for (int y = 0; y < 8; ++y)
      for (int x = 0; x < 8; ++x) {
        T ux = ((T)2 * (T)x / (T)7) - (T)1;
        T uy = ((T)2 * (T)y / (T)7) - (T)1;
        T wx = i0(beta * std::sqrt((T)1 - ux * ux)) / i0b;
        T wy = i0(beta * std::sqrt((T)1 - uy * uy)) / i0b;
        kaiser[y * 8 + x] = wx * wy;
      }
  }
};

template <class T>
const Tables<T>& tables() {
  static Tables<T> t;
  return t;
}

// (value * 1e5) as i64: truncation toward zero, saturating, NaN -> 0.
template <class T>
inline int64_t to_fixed(T value) {
  T v = value * (T)100000.0f;
  if (std::isnan(v)) return 0;
  if (v >= (T)9.2233720368547758e18) return INT64_MAX;
  if (v <= (T)-9.2233720368547758e18) return INT64_MIN;
  return (int64_t)v;
}

template <class T>
void dct_1d(T* x, const T* c) {
  T t[8];
  for (int i = 0; i < 8; ++i) t[i] = x[i];
  for (int k = 0; k < 8; ++k) {
    const T* r = c + k * 8;
    x[k] = t[0] * r[0] + t[1] * r[1] + t[2] * r[2] + t[3] * r[3] + t[4] * r[4] + t[5] * r[5] + t[6] * r[6] + t[7] * r[7];
  }
}

template <class T>
void transpose8(T* b) {
  for (int y = 0; y < 8; ++y)
    for (int x = y + 1; x < 8; ++x) std::swap(b[y * 8 + x], b[x * 8 + y]);
}

template <class T>
void dct2(T* b, const T* c) {
  for (int i = 0; i < 8; ++i) dct_1d(b + i * 8, c);
  transpose8(b);
  for (int i = 0; i < 8; ++i) dct_1d(b + i * 8, c);
  transpose8(b);
}

template <class T>
void idct2(T* b, const T* c) {
  transpose8(b);
  for (int i = 0; i < 8; ++i) dct_1d(b + i * 8, c);
  transpose8(b);
  for (int i = 0; i < 8; ++i) dct_1d(b + i * 8, c);
}

template <class T>
void wht(T* d, int n) {
  for (int h = 1; h < n; h *= 2)
    for (int i = 0; i < n; i += 2 * h)
      for (int j = i; j < i + h; ++j) {
        T x = d[j], y = d[j + h];
        d[j] = x + y;
        d[j + h] = x - y;
      }
  T scale = (T)1 / std::sqrt((T)n);
  for (int i = 0; i < n; ++i) d[i] *= scale;
}

// Haar group transform (an option beside the reference's Walsh-Hadamard,
// which bm3d.rs:595-614 fixes; no reference counterpart).  Orthonormal
// multi-level Haar in the in-place lifting layout: stage h = 1, 2, 4, .. < n
// takes the butterfly (x, y) -> (x + y, x - y) of positions k and k + h for
// every k that is a multiple of 2h, the same butterfly and operand order as
// the Walsh-Hadamard stages but on those pairs only.  Position 0 ends as the
// sum of all n values and position k > 0, whose lowest set bit is h, as the
// sum over [k - h, k) minus the sum over [k, k + h); the scale 1/sqrt(n) at
// position 0 and 1/sqrt(2h) elsewhere makes the rows orthonormal.  The
// inverse is the transpose: the same scales first, then the stages from
// h = n/2 down to 1.  For n = 2 the two transforms are the same matrix; groups
// of two are evaluated in the Walsh-Hadamard order (butterfly, then scale) in
// both directions, so they give the Walsh-Hadamard bits.
template <class T>
T haar_scale(int k, int n) {
  const int span = k == 0 ? n : 2 * (k & -k);
  return (T)1 / std::sqrt((T)span);
}

template <class T>
void haar_stage(T* d, int n, int h) {
  for (int k = 0; k + h < n; k += 2 * h) {
    T x = d[k], y = d[k + h];
    d[k] = x + y;
    d[k + h] = x - y;
  }
}

template <class T>
void haar_fwd(T* d, int n) {
  for (int h = 1; h < n; h *= 2) haar_stage(d, n, h);
  for (int k = 0; k < n; ++k) d[k] *= haar_scale<T>(k, n);
}

template <class T>
void haar_inv(T* d, int n) {
  for (int k = 0; k < n; ++k) d[k] *= haar_scale<T>(k, n);
  for (int h = n / 2; h >= 1; h /= 2) haar_stage(d, n, h);
}

template <class T>
void group_fwd(T* col, int g, int kind) {
  if (kind == 1) haar_fwd(col, g);
  else wht(col, g);
}
template <class T>
void group_inv(T* col, int g, int kind) {
  if (kind == 1 && g > 2) haar_inv(col, g);
  else wht(col, g);  // the Walsh-Hadamard transform with 1/sqrt(n) is its own inverse
}

template <class T>
void transform3d(T* s, int g, const Tables<T>& tb, int kind) {
  for (int i = 0; i < g; ++i) dct2(s + i * 64, tb.dct);
  for (int i = 0; i < 64; ++i) {
    T col[32];
    for (int k = 0; k < g; ++k) col[k] = s[k * 64 + i];
    group_fwd(col, g, kind);
    for (int k = 0; k < g; ++k) s[k * 64 + i] = col[k];
  }
}

template <class T>
void inverse3d(T* s, int g, const Tables<T>& tb, int kind) {
  for (int i = 0; i < 64; ++i) {
    T col[32];
    for (int k = 0; k < g; ++k) col[k] = s[k * 64 + i];
    group_inv(col, g, kind);
    for (int k = 0; k < g; ++k) s[k * 64 + i] = col[k];
  }
  for (int i = 0; i < g; ++i) idct2(s + i * 64, tb.idct);
}

template <class T>
int hard_threshold(T* s, int n, T th) {
  int c = 0;
  for (int i = 0; i < n; ++i) {
    if (i == 0) { ++c; continue; }
    if (std::fabs(s[i]) < th) s[i] = 0;
    else ++c;
  }
  return c;
}

template <class T>
T wiener(T* noisy, const T* guide, int n, T sigma) {
  T sum = 0;
  const T s2 = sigma * sigma;
  for (int i = 0; i < n; ++i) {
    if (i == 0) { sum += 1; continue; }
    T e = guide[i] * guide[i];
    T coef = e / (e + s2 + (T)1e-5f);
    noisy[i] *= coef;
    sum += coef * coef;
  }
  return sum > 0 ? (T)1 / sum : (T)1;
}

template <class T>
void extract(const std::vector<T>& img, size_t w, size_t x, size_t y, T* out) {
  for (int dy = 0; dy < 8; ++dy)
    for (int dx = 0; dx < 8; ++dx) out[dy * 8 + dx] = img[(y + dy) * w + x + dx];
}

// compute_ssd_flat (bm3d.rs:367-391): running sum, exit after a row once above stop.
template <class T>
T ssd(const std::vector<T>& img, size_t w, size_t x, size_t y, const T* ref, T stop) {
  T d = 0;
  for (int dy = 0; dy < 8; ++dy) {
    const T* row = &img[(y + dy) * w + x];
    for (int dx = 0; dx < 8; ++dx) {
      T diff = row[dx] - ref[dy * 8 + dx];
      d += diff * diff;
    }
    if (d > stop) return d;
  }
  return d;
}

template <class T>
struct MatchT {
  T dist;
  int x, y;
};

template <class T>
int block_match(const std::vector<T>* ch, size_t w, size_t h, size_t rx, size_t ry, bool step1,
                const Bm3dParams& p, int* gx, int* gy) {
  const int max_cand = p.max_candidates;
  std::vector<MatchT<T>> cand;
  cand.reserve(std::min(max_cand, p.window * p.window));
  const T threshold = step1 ? (T)p.max_dist : (T)p.max_dist * (T)0.5f;
  const T thr = threshold * (T)64;
  T rr[64], rg[64], rb[64];
  extract(ch[0], w, rx, ry, rr);
  extract(ch[1], w, rx, ry, rg);
  extract(ch[2], w, rx, ry, rb);
  const size_t half = (size_t)(p.window / 2);
  const size_t wmax = w >= 8 ? w - 8 : 0, hmax = h >= 8 ? h - 8 : 0;
  const size_t sx0 = rx >= half ? rx - half : 0, sx1 = std::min(rx + half, wmax);
  const size_t sy0 = ry >= half ? ry - half : 0, sy1 = std::min(ry + half, hmax);
  cand.push_back({(T)0, (int)rx, (int)ry});
  for (size_t y = sy0; y <= sy1; ++y)
    for (size_t x = sx0; x <= sx1; ++x) {
      if (x == rx && y == ry) continue;
      T dr = ssd(ch[0], w, x, y, rr, thr);
      if (dr > thr) continue;
      T dg = ssd(ch[1], w, x, y, rg, thr - dr);
      if (dr + dg > thr) continue;
      T db = ssd(ch[2], w, x, y, rb, thr - (dr + dg));
      T total = dr + dg + db;
      if (total < thr && (int)cand.size() < max_cand) cand.push_back({total / (T)64, (int)x, (int)y});
    }
  std::stable_sort(cand.begin(), cand.end(), [](const MatchT<T>& a, const MatchT<T>& b) { return a.dist < b.dist; });
  int limit = std::min(p.max_group, (int)cand.size());
  int g = 1;
  while (g * 2 <= limit) g *= 2;
  for (int i = 0; i < g; ++i) { gx[i] = cand[i].x; gy[i] = cand[i].y; }
  return g;
}

template <class T>
std::vector<std::vector<T>> run_step(const std::vector<std::vector<T>>& noisy, const std::vector<std::vector<T>>& guide,
                                     size_t w, size_t h, const Bm3dParams& p, bool step1, Bm3dStats* stats) {
  const Tables<T>& tb = tables<T>();
  const size_t count = w * h;
  std::vector<std::atomic<int64_t>> num(3 * count), den(3 * count);
  for (auto& a : num) a.store(0, std::memory_order_relaxed);
  for (auto& a : den) a.store(0, std::memory_order_relaxed);
  std::vector<std::pair<size_t, size_t>> refs;
  const size_t ylim = h >= 8 ? h - 8 : 0, xlim = w >= 8 ? w - 8 : 0;
  for (size_t y = 0; y < ylim; y += (size_t)p.stride)
    for (size_t x = 0; x < xlim; x += (size_t)p.stride) refs.push_back({x, y});
  long long groups_total = 0;
#pragma omp parallel for schedule(dynamic, 16) reduction(+ : groups_total)
  for (long long r = 0; r < (long long)refs.size(); ++r) {
    const size_t rx = refs[r].first, ry = refs[r].second;
    int gx[32], gy[32];
    const int g = block_match(guide.data(), w, h, rx, ry, step1, p, gx, gy);
    groups_total += g;
    std::vector<T> gs(g * 64), ns(g * 64);
    for (int ch = 0; ch < 3; ++ch) {
      const T chs = ch == 0 ? (T)p.sigma : (T)p.sigma * (T)p.chroma_scale;
      for (int k = 0; k < g; ++k) extract(guide[ch], w, gx[k], gy[k], &gs[k * 64]);
      if (step1) ns = gs;
      else for (int k = 0; k < g; ++k) extract(noisy[ch], w, gx[k], gy[k], &ns[k * 64]);
      transform3d(gs.data(), g, tb, p.group_transform);
      if (!step1) transform3d(ns.data(), g, tb, p.group_transform);
      T weight;
      if (step1) {
        const T th = (T)p.lambda * chs;
        int nz = hard_threshold(gs.data(), g * 64, th);
        weight = nz > 0 ? (T)1 / (T)nz : (T)1;
        ns = gs;
      } else {
        weight = wiener(ns.data(), gs.data(), g * 64, chs);
      }
      inverse3d(ns.data(), g, tb, p.group_transform);
      std::atomic<int64_t>* nacc = &num[ch * count];
      std::atomic<int64_t>* dacc = &den[ch * count];
      for (int k = 0; k < g; ++k)
        for (int dy = 0; dy < 8; ++dy)
          for (int dx = 0; dx < 8; ++dx) {
            const size_t idx = (gy[k] + dy) * w + gx[k] + dx;
            const T val = ns[k * 64 + dy * 8 + dx];
            const T wv = tb.kaiser[dy * 8 + dx] * weight;
            nacc[idx].fetch_add(to_fixed(val * wv), std::memory_order_relaxed);
            dacc[idx].fetch_add(to_fixed(wv), std::memory_order_relaxed);
          }
    }
  }
  if (stats) {
    stats->refs += (long long)refs.size();
    stats->groups += groups_total;
  }
  std::vector<std::vector<T>> out(3, std::vector<T>(count));
  for (int ch = 0; ch < 3; ++ch)
    for (size_t i = 0; i < count; ++i) {
      T n = (T)num[ch * count + i].load() / (T)100000.0f;
      T d = (T)den[ch * count + i].load() / (T)100000.0f;
      out[ch][i] = d > (T)1e-6f ? n / d : noisy[ch][i];
    }
  return out;
}

template <class T>
std::vector<T> blur1(const std::vector<T>& data, size_t w, size_t h, T sigma) {
  const int radius = (int)std::ceil((T)3 * sigma);
  const int klen = 2 * radius + 1;
  std::vector<T> k(klen);
  const T two_s2 = (T)2 * sigma * sigma;
  for (int i = 0; i < klen; ++i) {
    T kk = (T)i - (T)radius;
    k[i] = std::exp(-kk * kk / two_s2);
  }
  T ksum = 0;
  for (T v : k) ksum += v;
  for (T& v : k) v /= ksum;
  std::vector<T> tmp(w * h), out(w * h);
#pragma omp parallel for schedule(static)
  for (long long yy = 0; yy < (long long)h; ++yy) {
    const size_t y = (size_t)yy;
    for (size_t x = 0; x < w; ++x) {
      T val = 0, ws = 0;
      for (int i = 0; i < klen; ++i) {
        long long kx = (long long)x - radius + i;
        if (kx >= 0 && kx < (long long)w) { val += data[y * w + kx] * k[i]; ws += k[i]; }
      }
      tmp[y * w + x] = val / ws;
    }
  }
#pragma omp parallel for schedule(static)
  for (long long yy = 0; yy < (long long)h; ++yy) {
    const size_t y = (size_t)yy;
    for (size_t x = 0; x < w; ++x) {
      T val = 0, ws = 0;
      for (int i = 0; i < klen; ++i) {
        long long ky = (long long)y - radius + i;
        if (ky >= 0 && ky < (long long)h) { val += tmp[ky * w + x] * k[i]; ws += k[i]; }
      }
      out[y * w + x] = val / ws;
    }
  }
  return out;
}

template <class T>
void to_ycbcr(const float* rgb, size_t n, std::vector<std::vector<T>>& ch) {
  ch.assign(3, std::vector<T>(n));
  for (size_t i = 0; i < n; ++i) {
    const T r = (T)(rgb[i * 3] * 255.0f), g = (T)(rgb[i * 3 + 1] * 255.0f), b = (T)(rgb[i * 3 + 2] * 255.0f);
    ch[0][i] = (T)0.299f * r + (T)0.587f * g + (T)0.114f * b;
    ch[1][i] = (T)-0.168736f * r - (T)0.331264f * g + (T)0.5f * b + (T)128;
    ch[2][i] = (T)0.5f * r - (T)0.418688f * g - (T)0.081312f * b + (T)128;
  }
}

template <class T>
void from_ycbcr(const std::vector<T>& y, const std::vector<T>& cb, const std::vector<T>& cr, float* rgb, size_t n) {
  for (size_t i = 0; i < n; ++i) {
    const T yv = y[i], cbv = cb[i] - (T)128, crv = cr[i] - (T)128;
    const T r = yv + (T)1.402f * crv;
    const T g = yv - (T)0.344136f * cbv - (T)0.714136f * crv;
    const T b = yv + (T)1.772f * cbv;
    rgb[i * 3] = (float)(std::fmax(r, (T)0) / (T)255);
    rgb[i * 3 + 1] = (float)(std::fmax(g, (T)0) / (T)255);
    rgb[i * 3 + 2] = (float)(std::fmax(b, (T)0) / (T)255);
  }
}

}  // namespace

Bm3dParams Bm3dParams::from_intensity(float i) {
  Bm3dParams p;
  const float val = rclamp(i, 0.001f, 1.0f);
  p.intensity = i;
  p.sigma = val * 80.0f;
  p.lambda = 2.0f + (val * 2.5f);
  p.max_dist = 3000.0f + (val * 20000.0f);
  p.chroma_scale = 1.8f;
  p.detail = rclamp(i * 0.5f, 0.0f, 0.5f);
  return p;
}

template <class T>
void bm3d(float* rgb, size_t w, size_t h, const Bm3dParams& p, Bm3dStats* stats) {
  const size_t n = w * h;
  std::vector<std::vector<T>> ch;
  to_ycbcr<T>(rgb, n, ch);
  const std::vector<T> orig_y = ch[0];
  auto basic = run_step<T>(ch, ch, w, h, p, true, stats);
  auto den = run_step<T>(ch, basic, w, h, p, false, stats);
  const std::vector<T> blurred = blur1<T>(orig_y, w, h, (T)3);
  const T ds = (T)p.detail;
  for (size_t i = 0; i < n; ++i) {
    const T hf = orig_y[i] - blurred[i];
    den[0][i] = rclamp(den[0][i] + ds * hf, (T)0, (T)255);
  }
  from_ycbcr<T>(den[0], den[1], den[2], rgb, n);
}

template void bm3d<float>(float*, size_t, size_t, const Bm3dParams&, Bm3dStats*);
template void bm3d<double>(float*, size_t, size_t, const Bm3dParams&, Bm3dStats*);

// chroma_bm3d (bm3d.rs:705-759)
void chroma_bm3d(float* rgb, size_t w, size_t h, float intensity) {
  const float sigma = rclamp(intensity, 0.001f, 1.0f) * 80.0f;
  const float threshold = (2.0f + (rclamp(intensity, 0.001f, 1.0f) * 2.5f)) * sigma * 1.8f;
  const Tables<float>& tb = tables<float>();
  const size_t n = w * h;
  std::vector<std::vector<float>> ch;
  to_ycbcr<float>(rgb, n, ch);
  for (int c = 1; c < 3; ++c) {
    std::vector<std::atomic<int64_t>> num(n), den(n);
    for (size_t i = 0; i < n; ++i) { num[i].store(0); den[i].store(0); }
    std::vector<std::pair<size_t, size_t>> refs;
    for (size_t y = 0; y < (h >= 8 ? h - 8 : 0); y += 6)
      for (size_t x = 0; x < (w >= 8 ? w - 8 : 0); x += 6) refs.push_back({x, y});
#pragma omp parallel for schedule(static)
    for (long long r = 0; r < (long long)refs.size(); ++r) {
      float b[64];
      extract(ch[c], w, refs[r].first, refs[r].second, b);
      dct2(b, tb.dct);
      hard_threshold(b, 64, threshold);
      idct2(b, tb.idct);
      for (int dy = 0; dy < 8; ++dy)
        for (int dx = 0; dx < 8; ++dx) {
          const size_t idx = (refs[r].second + dy) * w + refs[r].first + dx;
          const float wk = tb.kaiser[dy * 8 + dx];
          num[idx].fetch_add(to_fixed(b[dy * 8 + dx] * wk), std::memory_order_relaxed);
          den[idx].fetch_add(to_fixed(wk), std::memory_order_relaxed);
        }
    }
    for (size_t i = 0; i < n; ++i) {
      float nn = (float)num[i].load() / 100000.0f, dd = (float)den[i].load() / 100000.0f;
      if (dd > 1e-6f) ch[c][i] = nn / dd;
    }
  }
  from_ycbcr<float>(ch[0], ch[1], ch[2], rgb, n);
}

void bm3d_tables(float* dct, float* idct, float* kaiser) {
  const Tables<float>& t = tables<float>();
  for (int i = 0; i < 64; ++i) { dct[i] = t.dct[i]; idct[i] = t.idct[i]; kaiser[i] = t.kaiser[i]; }
}

}  // namespace oracle

namespace oracle {
// The 1-D group transforms on one column of n values (kind 0: Walsh-Hadamard,
// 1: Haar), forward or inverse, in f32 or f64.
void group_transform_1d(double* d, int n, int kind, bool inverse, bool fp64) {
  if (fp64) {
    if (inverse) group_inv(d, n, kind);
    else group_fwd(d, n, kind);
  } else {
    float f[32];
    for (int i = 0; i < n; ++i) f[i] = (float)d[i];
    if (inverse) group_inv(f, n, kind);
    else group_fwd(f, n, kind);
    for (int i = 0; i < n; ++i) d[i] = f[i];
  }
}
}  // namespace oracle
