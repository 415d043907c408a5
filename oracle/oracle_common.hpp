// Oracle (test infrastructure only): CPU restatement of the reference
// algorithms, used by the tests, the smoke check and the CPU-baseline leg of
// the benchmark as the checker.  Nothing in the product links or calls it.
//
// Build flags (oracle/Makefile): -O2 -ffp-contract=off -fno-fast-math, so
// every f32 expression rounds exactly as the reference's Rust does (Rust never
// contracts a*b+c into an FMA and evaluates left to right).
#pragma once
#include <cmath>
#include <cstddef>
#include <cstdint>
#include <cstring>

namespace oracle {

// Colour-filter table: pichromatic/src/cfa.rs:98-150 (pattern letters, 48x48
// extension) and cfa.rs:204-233 (crop shift).
struct Cfa {
  int w = 0, h = 0;
  uint8_t pat[48][48] = {};
  bool parse(const char* s) {
    size_t n = std::strlen(s);
    if (n == 4) { w = 2; h = 2; }
    else if (n == 36) { w = 6; h = 6; }
    else if (n == 16) { w = 2; h = 8; }
    else if (n == 144) { w = 12; h = 12; }
    else return false;
    uint8_t base[12][12] = {};
    for (size_t i = 0; i < n; ++i) {
      int c;
      switch (s[i]) {
        case 'R': c = 0; break;
        case 'G': c = 1; break;
        case 'B': c = 2; break;
        case 'E': case 'C': c = 3; break;
        case 'M': c = 4; break;
        case 'Y': c = 5; break;
        default: return false;
      }
      base[i / w][i % w] = (uint8_t)c;
    }
    for (int r = 0; r < 48; ++r)
      for (int c = 0; c < 48; ++c) pat[r][c] = base[r % h][c % w];
    return true;
  }
  // cfa.rs:140-142
  int color_at(size_t r, size_t c) const { return pat[r % 48][c % 48]; }
  Cfa shift(size_t x, size_t y) const {
    Cfa o;
    o.w = w; o.h = h;
    for (int r = 0; r < 48; ++r)
      for (int c = 0; c < 48; ++c) o.pat[r][c] = (uint8_t)color_at((size_t)r + y, (size_t)c + x);
    return o;
  }
};

// Rust f32/f64 min/max/clamp semantics (max/min ignore a NaN operand; clamp
// passes NaN through).
template <class T> inline T rmax(T a, T b) { return std::fmax(a, b); }
template <class T> inline T rmin(T a, T b) { return std::fmin(a, b); }
template <class T> inline T rclamp(T x, T lo, T hi) { return x < lo ? lo : (x > hi ? hi : x); }

}  // namespace oracle
