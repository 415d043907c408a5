// CFA descriptor shared by host code: the colour index at (row, col) of a
// repeating colour-filter pattern, as pichromatic/src/cfa.rs:98-150 defines it
// (pattern string of 4, 16, 36 or 144 letters, extended to a 48x48 table) and
// the crop shift of cfa.rs:204-233.
#pragma once
#include <cstdint>
#include <cstring>
#include <string>

namespace imops {

struct CfaTable {
  int width = 0, height = 0;   // repeat period
  uint8_t pat[48][48] = {};    // colour index, 0=R 1=G 2=B 3=C 4=M 5=Y
  std::string name;

  static int letter(char ch) {
    switch (ch) {
      case 'R': return 0;
      case 'G': return 1;
      case 'B': return 2;
      case 'E': return 3;
      case 'C': return 3;
      case 'M': return 4;
      case 'Y': return 5;
      default: return -1;
    }
  }

  bool parse(const char* s) {
    size_t n = std::strlen(s);
    int w, h;
    switch (n) {
      case 4: w = 2; h = 2; break;
      case 36: w = 6; h = 6; break;
      case 16: w = 2; h = 8; break;
      case 144: w = 12; h = 12; break;
      default: return false;
    }
    uint8_t base[12][12] = {};
    for (size_t i = 0; i < n; ++i) {
      int c = letter(s[i]);
      if (c < 0) return false;
      base[i / w][i % w] = (uint8_t)c;
    }
    for (int r = 0; r < 48; ++r)
      for (int c = 0; c < 48; ++c) pat[r][c] = base[r % h][c % w];
    width = w;
    height = h;
    name = s;
    return true;
  }

  int color_at(size_t row, size_t col) const { return pat[row % 48][col % 48]; }

  // The pattern seen by an image cropped at (x, y).
  CfaTable shift(size_t x, size_t y) const {
    CfaTable out;
    out.width = width;
    out.height = height;
    for (int r = 0; r < 48; ++r)
      for (int c = 0; c < 48; ++c) out.pat[r][c] = (uint8_t)color_at((size_t)r + y, (size_t)c + x);
    static const char kNames[] = "RGBCMY";
    out.name.clear();
    for (int r = 0; r < height; ++r)
      for (int c = 0; c < width; ++c) out.name.push_back(kNames[out.pat[r][c] % 6]);
    return out;
  }

  bool is_bayer() const {
    if (width != 2 || height != 2) return false;
    int cnt[6] = {};
    for (int r = 0; r < 2; ++r)
      for (int c = 0; c < 2; ++c) cnt[pat[r][c] % 6]++;
    if (cnt[0] != 1 || cnt[1] != 2 || cnt[2] != 1) return false;
    // the two greens sit on one diagonal
    return (pat[0][1] == 1 && pat[1][0] == 1) || (pat[0][0] == 1 && pat[1][1] == 1);
  }
};

}  // namespace imops
