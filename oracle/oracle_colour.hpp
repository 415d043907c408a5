// Oracle (test infrastructure only): colour-stage declarations.
#pragma once
#include <cstddef>

#include "oracle_common.hpp"

namespace oracle {

enum { CS_SRGB = 0, CS_LINEAR_SRGB = 1, CS_ACESCG = 2, CS_XYZD65 = 3 };

struct ColourTables {
  float xyz_to_lin[3][3], lin_to_xyz[3][3], aces_to_lin[3][3], lin_to_aces[3][3];
  ColourTables();
};
const ColourTables& colour_tables();

void pseudo_inverse3(const float m[3][3], float out[3][3]);
bool camera_forward_matrix(const float* cal, size_t n, float fwd[3][3]);
float srgb_oetf(float x);
float srgb_eotf(float x);
void convert(int src, int dst, const float in[3], float out[3]);
void camera_cst(float* rgb, size_t n, const float fwd[3][3], int target);
void cst(float* rgb, size_t n, int src, int dst);
void exp_ev(float* rgb, size_t n, float ev);
void gamma_encode(float* rgb, size_t n, float gamma);
void cfa_coeffs(float* rgb, size_t n, const float wb[4]);
void markesteijn(const float* in, size_t width, size_t height, const Cfa& cfa, float* out);
// The three-pass Markesteijn extension (markesteijn3.cpp); passes 1 or 3.
extern int g_m3_plant;  // test hook: a planted mistake (0 = none)
template <class T>
void markesteijn3(const float* in, size_t width, size_t height, const Cfa& cfa, int passes, float* out_rgb);

// Test hooks of the AMaZE restatement (see amaze.cpp): trace receives 16
// planes of 128 x 128 floats for the tile at (trace_top, trace_left).
struct AmazeHooks {
  int plant = 0;
  float* trace = nullptr;
  int trace_top = 0, trace_left = 0;
};
extern AmazeHooks g_amaze_hooks;

bool amaze_size_supported(size_t width, size_t height);
template <class T>
void amaze_demosaic(const float* input, size_t width, size_t height, const Cfa& cfa, float* out);

}  // namespace oracle
