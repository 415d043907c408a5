// Oracle (test infrastructure only): the luma-guided chroma denoise restated
// from pichromatic/src/chroma_denoise.rs:14-152 in f32, in the reference's
// operation order: the box filter's running sums walk each row, then each
// column, one add or subtract at a time, and the window-size normalisation
// divides by the product of the two in-bounds counts.
#include <algorithm>
#include <vector>

namespace oracle {
namespace {

// pixel.rs:33-35
constexpr float kWr = 0.2722287168f, kWg = 0.6740817658f, kWb = 0.0536895174f;

float luma(const float* p) { return kWr * p[0] + kWg * p[1] + kWb * p[2]; }

}  // namespace

// box_filter (chroma_denoise.rs:100-152)
std::vector<float> box_filter(const std::vector<float>& data, size_t w, size_t h, size_t r) {
  std::vector<float> tmp(data.size(), 0.0f), out(data.size(), 0.0f);
  for (size_t y = 0; y < h; ++y) {
    const size_t row = y * w;
    float sum = 0.0f;
    for (size_t dx = 0; dx <= std::min(r, w - 1); ++dx) sum += data[row + dx];
    tmp[row] = sum;
    for (size_t x = 1; x < w; ++x) {
      if (x > r) sum -= data[row + x - r - 1];
      if (x + r < w) sum += data[row + x + r];
      tmp[row + x] = sum;
    }
  }
  for (size_t x = 0; x < w; ++x) {
    float sum = 0.0f;
    for (size_t dy = 0; dy <= std::min(r, h - 1); ++dy) sum += tmp[dy * w + x];
    out[x] = sum;
    for (size_t y = 1; y < h; ++y) {
      if (y > r) sum -= tmp[(y - r - 1) * w + x];
      if (y + r < h) sum += tmp[(y + r) * w + x];
      out[y * w + x] = sum;
    }
  }
  for (size_t y = 0; y < h; ++y) {
    const float cy = (float)(std::min(r, y) + std::min(r, h - 1 - y) + 1);
    for (size_t x = 0; x < w; ++x) {
      const float cx = (float)(std::min(r, x) + std::min(r, w - 1 - x) + 1);
      out[y * w + x] /= cx * cy;
    }
  }
  return out;
}

// guided_filter (chroma_denoise.rs:41-99)
std::vector<float> guided_filter(const std::vector<float>& g, const std::vector<float>& p, size_t w, size_t h,
                                 size_t r, float eps) {
  const size_t n = g.size();
  const std::vector<float> mg = box_filter(g, w, h, r), mp = box_filter(p, w, h, r);
  std::vector<float> gg(n), gp(n);
  for (size_t i = 0; i < n; ++i) {
    gg[i] = g[i] * g[i];
    gp[i] = g[i] * p[i];
  }
  const std::vector<float> mgg = box_filter(gg, w, h, r), mgp = box_filter(gp, w, h, r);
  std::vector<float> a(n), b(n);
  for (size_t i = 0; i < n; ++i) {
    const float var = mgg[i] - mg[i] * mg[i];
    const float cov = mgp[i] - mg[i] * mp[i];
    a[i] = cov / (var + eps);
    b[i] = mp[i] - a[i] * mg[i];
  }
  const std::vector<float> ma = box_filter(a, w, h, r), mb = box_filter(b, w, h, r);
  std::vector<float> q(n);
  for (size_t i = 0; i < n; ++i) q[i] = ma[i] * g[i] + mb[i];
  return q;
}

// chroma_denoise (chroma_denoise.rs:14-38), rgb interleaved in place
void luma_chroma_denoise(float* rgb, size_t w, size_t h, size_t r, float eps) {
  const size_t n = w * h;
  std::vector<float> y(n), cr(n), cb(n);
  for (size_t i = 0; i < n; ++i) {
    y[i] = luma(rgb + 3 * i);
    cr[i] = rgb[3 * i] - y[i];
    cb[i] = rgb[3 * i + 2] - y[i];
  }
  const std::vector<float> crs = guided_filter(y, cr, w, h, r, eps), cbs = guided_filter(y, cb, w, h, r, eps);
  const float kr = kWr / kWg, kb = kWb / kWg;
  for (size_t i = 0; i < n; ++i) {
    float* px = rgb + 3 * i;
    const float yv = luma(px);
    px[0] = yv + crs[i];
    px[1] = yv - kr * crs[i] - kb * cbs[i];
    px[2] = yv + cbs[i];
  }
}

}  // namespace oracle

extern "C" int oracle_box_filter(const float* in, float* out, size_t w, size_t h, size_t r) {
  if (w == 0 || h == 0) return 0;
  std::vector<float> v(in, in + w * h);
  const std::vector<float> o = oracle::box_filter(v, w, h, r);
  std::copy(o.begin(), o.end(), out);
  return 0;
}

extern "C" int oracle_luma_chroma_denoise(float* rgb, size_t w, size_t h, size_t r, float eps) {
  if (w == 0 || h == 0) return 0;
  oracle::luma_chroma_denoise(rgb, w, h, r, eps);
  return 0;
}
