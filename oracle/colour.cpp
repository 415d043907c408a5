// Oracle (test infrastructure only): the per-pixel colour stages and the
// 3x3 "Markesteijn" demosaic, restated from
//   pichromatic/src/cst.rs:6-95 (camera_cst, pseudo_inverse_matrix, cst),
//   pichromatic/src/exp.rs:4-11, cfa_coeffs.rs:4-12, tone_map.rs:38-49,
//   pichromatic/src/demosaic.rs:108-186 (Markesteijn).
// Colour-space matrices/transfer functions of the third-party `color` crate
// (0.3.x, absent here) are restated from their published definitions (CSS
// Color 4 sRGB/XYZ-D65; ACEScg = AP1 primaries with the ACES white, Bradford-
// adapted to D65) and converted through linear sRGB as that crate does:
// parity unpinned for the absolute matrix entries.
#include <vector>

#include "oracle_common.hpp"
#include "oracle_colour.hpp"

namespace oracle {

// cst.rs:51-95, f32, no pivoting, same loop order.
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

// cst.rs:12-39: XYZ->camera rows normalised by the camera response to D65,
// then inverted.  Only 3-row matrices (components == 3) are accepted.
bool camera_forward_matrix(const float* cal, size_t n, float fwd[3][3]) {
  if (n != 9) return false;
  const float xyz_d65[3] = {0.95042f, 1.0f, 1.08890f};
  float rgb_d65[3], mult[3] = {1.0f, 1.0f, 1.0f};
  for (int i = 0; i < 3; ++i)
    rgb_d65[i] = cal[i * 3] * xyz_d65[0] + cal[i * 3 + 1] * xyz_d65[1] + cal[i * 3 + 2] * xyz_d65[2];
  for (int i = 0; i < 3; ++i)
    if (rgb_d65[i] > 0.0f) mult[i] = 1.0f / rgb_d65[i];
  float xyz2cam[3][3];
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) xyz2cam[i][j] = cal[i * 3 + j] * mult[i];
  pseudo_inverse3(xyz2cam, fwd);
  return true;
}


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

static void mat_rows(const double src[3][3], float dst[3][3]) {
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) dst[i][j] = (float)src[i][j];
}

ColourTables::ColourTables() {
  // CSS Color 4 XYZ(D65) <-> linear sRGB (rational forms).
  const double x2s[3][3] = {{12831.0 / 3959.0, -329.0 / 214.0, -1974.0 / 3959.0},
                            {-851781.0 / 878810.0, 1648619.0 / 878810.0, 36519.0 / 878810.0},
                            {705.0 / 12673.0, -2585.0 / 12673.0, 705.0 / 667.0}};
  const double s2x[3][3] = {{506752.0 / 1228815.0, 87881.0 / 245763.0, 12673.0 / 70218.0},
                            {87098.0 / 409605.0, 175762.0 / 245763.0, 12673.0 / 175545.0},
                            {7918.0 / 409605.0, 87881.0 / 737289.0, 1001167.0 / 1053270.0}};
  double a2s_flat[9], s2a_flat[9];
  acescg_to_linear_srgb(&x2s[0][0], a2s_flat);
  m3inv(a2s_flat, s2a_flat);
  double a2s[3][3], s2a[3][3];
  for (int k = 0; k < 9; ++k) { a2s[k / 3][k % 3] = a2s_flat[k]; s2a[k / 3][k % 3] = s2a_flat[k]; }
  mat_rows(x2s, xyz_to_lin);
  mat_rows(s2x, lin_to_xyz);
  mat_rows(a2s, aces_to_lin);
  mat_rows(s2a, lin_to_aces);
}

const ColourTables& colour_tables() {
  static ColourTables t;
  return t;
}

static inline void mat3(const float m[3][3], const float in[3], float out[3]) {
  float a = m[0][0] * in[0] + m[0][1] * in[1] + m[0][2] * in[2];
  float b = m[1][0] * in[0] + m[1][1] * in[1] + m[1][2] * in[2];
  float c = m[2][0] * in[0] + m[2][1] * in[1] + m[2][2] * in[2];
  out[0] = a; out[1] = b; out[2] = c;
}

float srgb_oetf(float x) {
  float ax = std::fabs(x);
  if (ax > 0.0031308f) {
    float y = 1.055f * powf(ax, 1.0f / 2.4f) - 0.055f;
    return x < 0.0f ? -y : y;
  }
  return 12.92f * x;
}

float srgb_eotf(float x) {
  float ax = std::fabs(x);
  if (ax > 0.04045f) {
    float y = powf((ax + 0.055f) / 1.055f, 2.4f);
    return x < 0.0f ? -y : y;
  }
  return x / 12.92f;
}

void to_linear_srgb(int cs, const float in[3], float out[3]) {
  const ColourTables& t = colour_tables();
  switch (cs) {
    case CS_SRGB: for (int c = 0; c < 3; ++c) out[c] = srgb_eotf(in[c]); break;
    case CS_LINEAR_SRGB: for (int c = 0; c < 3; ++c) out[c] = in[c]; break;
    case CS_ACESCG: mat3(t.aces_to_lin, in, out); break;
    case CS_XYZD65: mat3(t.xyz_to_lin, in, out); break;
    default: out[0] = out[1] = out[2] = NAN;
  }
}

void from_linear_srgb(int cs, const float in[3], float out[3]) {
  const ColourTables& t = colour_tables();
  switch (cs) {
    case CS_SRGB: for (int c = 0; c < 3; ++c) out[c] = srgb_oetf(in[c]); break;
    case CS_LINEAR_SRGB: for (int c = 0; c < 3; ++c) out[c] = in[c]; break;
    case CS_ACESCG: mat3(t.lin_to_aces, in, out); break;
    case CS_XYZD65: mat3(t.lin_to_xyz, in, out); break;
    default: out[0] = out[1] = out[2] = NAN;
  }
}

void convert(int src, int dst, const float in[3], float out[3]) {
  if (src == dst) { out[0] = in[0]; out[1] = in[1]; out[2] = in[2]; return; }
  float lin[3];
  to_linear_srgb(src, in, lin);
  from_linear_srgb(dst, lin, out);
}

void camera_cst(float* rgb, size_t n, const float fwd[3][3], int target) {
#pragma omp parallel for schedule(static)
  for (long long p = 0; p < (long long)n; ++p) {
    float* px = rgb + (size_t)p * 3;
    float r = px[0], g = px[1], b = px[2];
    float xyz[3] = {fwd[0][0] * r + fwd[0][1] * g + fwd[0][2] * b, fwd[1][0] * r + fwd[1][1] * g + fwd[1][2] * b,
                    fwd[2][0] * r + fwd[2][1] * g + fwd[2][2] * b};
    convert(CS_XYZD65, target, xyz, px);
  }
}

void cst(float* rgb, size_t n, int src, int dst) {
#pragma omp parallel for schedule(static)
  for (long long p = 0; p < (long long)n; ++p) {
    float in[3] = {rgb[p * 3], rgb[p * 3 + 1], rgb[p * 3 + 2]};
    convert(src, dst, in, rgb + (size_t)p * 3);
  }
}

void exp_ev(float* rgb, size_t n, float ev) {
  const float v = powf(2.0f, ev);  // exp.rs:5
  for (size_t i = 0; i < n * 3; ++i) rgb[i] = rgb[i] * v;
}

void gamma_encode(float* rgb, size_t n, float gamma) {
  const float inv = 1.0f / rmax(gamma, 1e-6f);  // tone_map.rs:39
  for (size_t i = 0; i < n * 3; ++i) rgb[i] = rgb[i] <= 0.0f ? 0.0f : powf(rgb[i], inv);
}

void cfa_coeffs(float* rgb, size_t n, const float wb[4]) {
  for (size_t p = 0; p < n; ++p) {
    rgb[p * 3] = rgb[p * 3] * wb[0];
    rgb[p * 3 + 1] = rgb[p * 3 + 1] * wb[1];
    rgb[p * 3 + 2] = rgb[p * 3 + 2] * wb[2];
  }
}

// demosaic.rs:108-186
void markesteijn(const float* in, size_t width, size_t height, const Cfa& cfa, float* out) {
#pragma omp parallel for schedule(static)
  for (long long rowl = 0; rowl < (long long)height; ++rowl) {
    size_t row = (size_t)rowl;
    size_t y0 = row == 0 ? 0 : row - 1;
    size_t y1 = std::min(row + 1, height - 1);
    for (size_t col = 0; col < width; ++col) {
      int centre = cfa.color_at(row, col);
      size_t x0 = col == 0 ? 0 : col - 1;
      size_t x1 = std::min(col + 1, width - 1);
      float* o = out + (row * width + col) * 3;
      for (int ch = 0; ch < 3; ++ch) {
        if (ch == centre) { o[ch] = in[row * width + col]; continue; }
        float sum = 0.0f, count = 0.0f;
        for (size_t ny = y0; ny <= y1; ++ny)
          for (size_t nx = x0; nx <= x1; ++nx) {
            if (ny == row && nx == col) continue;
            if (cfa.color_at(ny, nx) == ch) { sum += in[ny * width + nx]; count += 1.0f; }
          }
        o[ch] = count > 0.0f ? sum / count : 0.0f;
      }
    }
  }
}

}  // namespace oracle
