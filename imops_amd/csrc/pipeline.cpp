// Stage-list configuration and runner: the C++ counterpart of
// pichromatic-pipeline/src/config.rs:13-34 (parse TOML, falling back to JSON;
// `[[pipeline_modules]]` tables; module names matched case-insensitively
// against the registry of modules.rs:897-915) and pipeline.rs:4-13 (run the
// stages in order).
//
// Runs of consecutive elementwise stages (CFACoeffs, Exp, CST, Gamma) are
// fused into one pointwise kernel launch; each stage keeps its own f32
// arithmetic, so the output is the same as running them one by one.
#include <cctype>
#include <cerrno>
#include <climits>
#include <cstdlib>
#include <map>
#include <memory>
#include <algorithm>
#include <cstring>
#include <string>
#include <vector>

#include "runtime.hpp"
#include "stages.hpp"

namespace imops {
namespace {

// ------------------------------------------------------------------ values
struct Value {
  enum Kind { STR, NUM, INT, BOOL, ARR, OBJ } kind = STR;
  std::string s;
  double num = 0;
  long long i = 0;
  unsigned long long u = 0;  // INT: the value when it is non-negative (covers 2^63 .. 2^64-1)
  bool neg = false;          // INT: written with a minus sign
  bool b = false;
  std::vector<Value> arr;
  std::map<std::string, Value> obj;
};
using Table = std::map<std::string, Value>;

[[noreturn]] void bad(const std::string& m) { throw Error{IMOPS_E_INVALID, "cant decode config from string: " + m}; }

struct Cursor {
  const std::string& t;
  size_t p = 0;
  explicit Cursor(const std::string& s) : t(s) {}
  bool eof() const { return p >= t.size(); }
  char peek() const { return eof() ? '\0' : t[p]; }
  void skip_ws(bool newlines) {
    while (!eof()) {
      char c = t[p];
      if (c == ' ' || c == '\t' || c == '\r' || (newlines && c == '\n')) ++p;
      else if (c == '#') { while (!eof() && t[p] != '\n') ++p; }
      else break;
    }
  }
  bool take(char c) {
    if (peek() == c) { ++p; return true; }
    return false;
  }
};

std::string parse_string(Cursor& c, char q) {
  std::string out;
  while (!c.eof() && c.peek() != q) {
    char ch = c.t[c.p++];
    if (ch == '\\' && q == '"' && !c.eof()) {
      char e = c.t[c.p++];
      switch (e) {
        case 'n': out += '\n'; break;
        case 't': out += '\t'; break;
        case '"': out += '"'; break;
        case '\\': out += '\\'; break;
        case '/': out += '/'; break;
        default: out += e;
      }
    } else {
      out += ch;
    }
  }
  if (!c.take(q)) bad("unterminated string");
  return out;
}

Value parse_number(Cursor& c) {
  size_t s = c.p;
  while (!c.eof() && (std::isdigit((unsigned char)c.peek()) || std::strchr("+-.eE_", c.peek()))) ++c.p;
  std::string tok;
  for (size_t k = s; k < c.p; ++k)
    if (c.t[k] != '_') tok += c.t[k];
  if (tok.empty()) bad("expected a value");
  Value v;
  char* end = nullptr;
  if (tok.find_first_of(".eE") == std::string::npos) {
    v.kind = Value::INT;
    v.neg = tok[0] == '-';
    errno = 0;
    if (v.neg) {
      v.i = std::strtoll(tok.c_str(), &end, 10);
      v.num = (double)v.i;
    } else {
      v.u = std::strtoull(tok.c_str(), &end, 10);
      v.i = v.u > (unsigned long long)LLONG_MAX ? LLONG_MAX : (long long)v.u;
      v.num = (double)v.u;
    }
    if (errno == ERANGE) bad("integer out of range '" + tok + "'");
    if (v.neg && v.i == 0) v.neg = false;  // -0
  } else {
    v.kind = Value::NUM;
    v.num = std::strtod(tok.c_str(), &end);
  }
  if (!end || *end) bad("bad number '" + tok + "'");
  return v;
}

bool parse_word(Cursor& c, const char* w) {
  size_t n = std::strlen(w);
  if (c.t.compare(c.p, n, w) == 0) { c.p += n; return true; }
  return false;
}

Value parse_toml_value(Cursor& c);

Value parse_toml_inline_table(Cursor& c) {
  Value v;
  v.kind = Value::OBJ;
  c.skip_ws(false);
  if (c.take('}')) return v;
  for (;;) {
    c.skip_ws(false);
    std::string key;
    if (c.take('"')) key = parse_string(c, '"');
    else {
      while (!c.eof() && (std::isalnum((unsigned char)c.peek()) || c.peek() == '_' || c.peek() == '-')) key += c.t[c.p++];
    }
    if (key.empty()) bad("bad inline-table key");
    c.skip_ws(false);
    if (!c.take('=')) bad("expected '=' in inline table");
    c.skip_ws(false);
    v.obj[key] = parse_toml_value(c);
    c.skip_ws(false);
    if (c.take('}')) return v;
    if (!c.take(',')) bad("expected ',' in inline table");
  }
}

Value parse_toml_value(Cursor& c) {
  char ch = c.peek();
  if (ch == '"' || ch == '\'') {
    ++c.p;
    Value v;
    v.kind = Value::STR;
    v.s = parse_string(c, ch);
    return v;
  }
  if (ch == '[') {
    ++c.p;
    Value v;
    v.kind = Value::ARR;
    for (;;) {
      c.skip_ws(true);
      if (c.take(']')) return v;
      v.arr.push_back(parse_toml_value(c));
      c.skip_ws(true);
      if (c.take(']')) return v;
      if (!c.take(',')) bad("expected ',' in array");
    }
  }
  if (ch == '{') {
    ++c.p;
    return parse_toml_inline_table(c);
  }
  if (parse_word(c, "true")) { Value v; v.kind = Value::BOOL; v.b = true; return v; }
  if (parse_word(c, "false")) { Value v; v.kind = Value::BOOL; v.b = false; return v; }
  return parse_number(c);
}

std::vector<Table> parse_toml(const std::string& text) {
  std::vector<Table> mods;
  Cursor c(text);
  bool in_table = false;
  for (;;) {
    c.skip_ws(true);
    if (c.eof()) break;
    if (c.peek() == '[') {
      if (c.t.compare(c.p, 20, "[[pipeline_modules]]") != 0) bad("only [[pipeline_modules]] tables are allowed");
      c.p += 20;
      mods.emplace_back();
      in_table = true;
      continue;
    }
    std::string key;
    if (c.take('"')) key = parse_string(c, '"');
    else
      while (!c.eof() && (std::isalnum((unsigned char)c.peek()) || c.peek() == '_' || c.peek() == '-')) key += c.t[c.p++];
    if (key.empty()) bad("expected a key");
    if (!in_table) bad("key outside [[pipeline_modules]]");
    c.skip_ws(false);
    if (!c.take('=')) bad("expected '=' after key " + key);
    c.skip_ws(false);
    mods.back()[key] = parse_toml_value(c);
    c.skip_ws(false);
    if (!c.eof() && c.peek() != '\n') bad("trailing characters after value of " + key);
  }
  return mods;
}

Value parse_json_value(Cursor& c) {
  c.skip_ws(true);
  char ch = c.peek();
  if (ch == '"') { ++c.p; Value v; v.kind = Value::STR; v.s = parse_string(c, '"'); return v; }
  if (ch == '[') {
    ++c.p;
    Value v; v.kind = Value::ARR;
    c.skip_ws(true);
    if (c.take(']')) return v;
    for (;;) {
      v.arr.push_back(parse_json_value(c));
      c.skip_ws(true);
      if (c.take(']')) return v;
      if (!c.take(',')) bad("expected ',' in JSON array");
    }
  }
  if (ch == '{') {
    ++c.p;
    Value v; v.kind = Value::OBJ;
    c.skip_ws(true);
    if (c.take('}')) return v;
    for (;;) {
      c.skip_ws(true);
      if (!c.take('"')) bad("expected JSON key");
      std::string k = parse_string(c, '"');
      c.skip_ws(true);
      if (!c.take(':')) bad("expected ':'");
      v.obj[k] = parse_json_value(c);
      c.skip_ws(true);
      if (c.take('}')) return v;
      if (!c.take(',')) bad("expected ',' in JSON object");
    }
  }
  if (parse_word(c, "true")) { Value v; v.kind = Value::BOOL; v.b = true; return v; }
  if (parse_word(c, "false")) { Value v; v.kind = Value::BOOL; v.b = false; return v; }
  if (parse_word(c, "null")) bad("null value");
  return parse_number(c);
}

std::vector<Table> parse_json(const std::string& text) {
  Cursor c(text);
  Value root = parse_json_value(c);
  c.skip_ws(true);
  if (!c.eof()) bad("trailing JSON");
  if (root.kind != Value::OBJ || !root.obj.count("pipeline_modules")) bad("missing pipeline_modules");
  const Value& arr = root.obj.at("pipeline_modules");
  if (arr.kind != Value::ARR) bad("pipeline_modules must be an array");
  std::vector<Table> mods;
  for (const Value& m : arr.arr) {
    if (m.kind != Value::OBJ) bad("module must be an object");
    mods.push_back(m.obj);
  }
  return mods;
}

// Parameter<T> accepts a bare value or {value = ..., description = ...}
// (modules.rs:40-67).
const Value* param(const Table& t, const char* key) {
  auto it = t.find(key);
  if (it == t.end()) return nullptr;
  const Value* v = &it->second;
  if (v->kind == Value::OBJ) {
    auto jt = v->obj.find("value");
    if (jt == v->obj.end()) throw Error{IMOPS_E_INVALID, std::string("missing field `value` in ") + key};
    v = &jt->second;
  }
  return v;
}

double num_param(const Table& t, const char* key, double def, const char* mod) {
  const Value* v = param(t, key);
  if (!v) return def;
  if (v->kind != Value::NUM && v->kind != Value::INT) throw Error{IMOPS_E_INVALID, std::string("Invalid ") + mod + " config"};
  return v->num;
}

// A u64 field (serde rejects floats, negative and out-of-range values for it).
uint64_t u64_param(const Table& t, const char* key, uint64_t def, const char* mod) {
  const Value* v = param(t, key);
  if (!v) return def;
  if (v->kind != Value::INT || v->neg) throw Error{IMOPS_E_INVALID, std::string("Invalid ") + mod + " config"};
  return (uint64_t)v->u;
}

std::string str_param(const Table& t, const char* key, const std::string& def, const char* mod) {
  const Value* v = param(t, key);
  if (!v) return def;
  if (v->kind != Value::STR) throw Error{IMOPS_E_INVALID, std::string("Invalid ") + mod + " config"};
  return v->s;
}

std::string lower(std::string s) {
  for (auto& ch : s) ch = (char)std::tolower((unsigned char)ch);
  return s;
}

std::string trim(const std::string& s) {
  size_t a = s.find_first_not_of(" \t\r\n"), b = s.find_last_not_of(" \t\r\n");
  return a == std::string::npos ? std::string() : s.substr(a, b - a + 1);
}

}  // namespace

enum StageKind {
  K_DEMOSAIC, K_CFACOEFFS, K_EXP, K_GAMMA, K_CST, K_CONTRAST, K_HIGHLIGHT, K_LCH, K_SIGMOID, K_BASELINE,
  K_BM3D, K_CHROMA_BM3D, K_LUMA_CHROMA, K_VIGNETTE, K_FILM, K_OTHER
};

struct Stage {
  std::string name;  // canonical schema name
  StageKind kind = K_OTHER;
  int algorithm = IMOPS_DEMOSAIC_AMAZE;
  double ev = 0.0, gamma = 2.2, intensity = 0.0, epsilon = 0.01, contrast = 1.0, strength = 1.0;
  double lch[3] = {1.0, 1.0, 1.0};
  int radius = 4;
  std::string target;
  Bm3dStageParams bm3d;
  FilmStageParams film;
};

}  // namespace imops

struct imops_pipeline {
  std::vector<imops::Stage> stages;
};

namespace imops {
namespace {

// Registry names in modules.rs:897-915 order.
const char* kModuleNames[] = {"Demosaic", "ChromaDenoise", "LumaGuidedChromaDenoise", "CFACoeffs", "Vignette",
                              "HighlightReconstruction", "BaselineExposureCompensation", "Exp", "Contrast", "CST",
                              "Film", "LCH", "SigmoidToneMap", "Gamma", "BM3D"};

Stage make_stage(const Table& t) {
  auto it = t.find("name");
  if (it == t.end() || it->second.kind != Value::STR) throw Error{IMOPS_E_INVALID, "Module must have a name field"};
  const std::string want = lower(it->second.s);
  Stage s;
  for (const char* n : kModuleNames)
    if (lower(n) == want) s.name = n;
  if (s.name.empty()) throw Error{IMOPS_E_INVALID, "wrong pipeline module name " + it->second.s};
  if (s.name == "Demosaic") {
    s.kind = K_DEMOSAIC;
    std::string a = str_param(t, "algorithm", "amaze", "Demosaic");  // rename_all = lowercase
    if (a == "amaze") s.algorithm = IMOPS_DEMOSAIC_AMAZE;
    else if (a == "markesteijn") s.algorithm = IMOPS_DEMOSAIC_MARKESTEIJN;
    else if (a == "markesteijn3") {
      const uint64_t passes = u64_param(t, "passes", 3, "Demosaic");
      if (passes != 1 && passes != 3) throw Error{IMOPS_E_INVALID, "Invalid Demosaic config: passes must be 1 or 3"};
      s.algorithm = passes == 3 ? IMOPS_DEMOSAIC_MARKESTEIJN3 : IMOPS_DEMOSAIC_MARKESTEIJN3_1PASS;
    }
    else if (a == "fast") s.algorithm = IMOPS_DEMOSAIC_FAST;
    else if (a == "superfast") s.algorithm = IMOPS_DEMOSAIC_SUPERFAST;
    else if (a == "supersuperfast") s.algorithm = IMOPS_DEMOSAIC_SUPERSUPERFAST;
    else throw Error{IMOPS_E_INVALID, "Invalid Demosaic config"};
  } else if (s.name == "CFACoeffs") {
    s.kind = K_CFACOEFFS;
  } else if (s.name == "Exp") {
    s.kind = K_EXP;
    s.ev = num_param(t, "ev", 0.0, "Exp");
  } else if (s.name == "Gamma") {
    s.kind = K_GAMMA;
    s.gamma = num_param(t, "gamma", 2.2, "Gamma");
  } else if (s.name == "CST") {
    s.kind = K_CST;
    s.target = str_param(t, "target_color_space", "", "CST");
  } else if (s.name == "BM3D") {
    s.kind = K_BM3D;
    s.bm3d = bm3d_stage_params_from_table(
        num_param(t, "intensity", 0.0, "BM3D"), param(t, "sigma") ? num_param(t, "sigma", 0, "BM3D") : -1.0,
        param(t, "lambda") ? num_param(t, "lambda", 0, "BM3D") : -1.0,
        (int)num_param(t, "stride", 6, "BM3D"), (int)num_param(t, "search_window", 19, "BM3D"),
        (int)num_param(t, "max_group", 16, "BM3D"));
    const std::string prec = str_param(t, "precision", "fp32", "BM3D");
    if (prec == "fp32") s.bm3d.precision = IMOPS_BM3D_FP32;
    else if (prec == "mixed") s.bm3d.precision = IMOPS_BM3D_MIXED;
    else if (prec == "mixed-match") s.bm3d.precision = IMOPS_BM3D_MIXED_MATCH;
    else throw Error{IMOPS_E_INVALID, "Invalid BM3D config"};
    // extension key: the group transform across a group's patches
    const std::string gt = str_param(t, "group_transform", "hadamard", "BM3D");
    if (gt == "hadamard") s.bm3d.group_transform = IMOPS_BM3D_HADAMARD;
    else if (gt == "haar") s.bm3d.group_transform = IMOPS_BM3D_HAAR;
    else throw Error{IMOPS_E_INVALID, "Invalid BM3D config"};
  } else if (s.name == "ChromaDenoise") {  // modules.rs:688-722
    s.kind = K_CHROMA_BM3D;
    s.intensity = num_param(t, "intensity", 0.0, "ChromaDenoise");
  } else if (s.name == "Contrast") {  // modules.rs:454-490
    s.kind = K_CONTRAST;
    s.contrast = num_param(t, "c", 1.0, "Contrast");
  } else if (s.name == "HighlightReconstruction") {  // modules.rs:265-296
    s.kind = K_HIGHLIGHT;
  } else if (s.name == "LCH") {  // modules.rs:221-262
    s.kind = K_LCH;
    s.lch[0] = num_param(t, "lc", 1.0, "LCH");
    s.lch[1] = num_param(t, "cc", 1.0, "LCH");
    s.lch[2] = num_param(t, "hc", 1.0, "LCH");
  } else if (s.name == "SigmoidToneMap") {  // modules.rs:392-416
    s.kind = K_SIGMOID;
  } else if (s.name == "BaselineExposureCompensation") {  // modules.rs:331-384
    s.kind = K_BASELINE;
  } else if (s.name == "Vignette") {  // modules.rs:762-798
    s.kind = K_VIGNETTE;
    s.strength = num_param(t, "strength", 1.0, "Vignette");
  } else if (s.name == "LumaGuidedChromaDenoise") {  // modules.rs:724-760
    s.kind = K_LUMA_CHROMA;
    const double r = num_param(t, "radius", 4, "LumaGuidedChromaDenoise");
    if (!(r >= 0 && r <= 1e9) || r != (double)(long long)r)
      throw Error{IMOPS_E_INVALID, "Invalid LumaGuidedChromaDenoise config"};
    s.radius = (int)r;
    s.epsilon = num_param(t, "epsilon", 0.01, "LumaGuidedChromaDenoise");
  } else if (s.name == "Film") {
    s.kind = K_FILM;
    s.film = film_stage_params_from_table(str_param(t, "stock", "ColorNeg200", "Film"),
                                          str_param(t, "film_format", "Film35mm", "Film"),
                                          u64_param(t, "seed", 1, "Film"),
                                          str_param(t, "output", "NegativeLinear", "Film"));
    const std::string prec = str_param(t, "precision", "fp64", "Film");
    if (prec == "mixed") s.film.fp32 = true;
    else if (prec != "fp64") throw Error{IMOPS_E_INVALID, "Invalid Film config"};
    // extensions (absent: the reference's behaviour)
    s.film.stock_seed = u64_param(t, "stock_seed", 0, "Film");
    const uint64_t bins = u64_param(t, "spectral_bins", 16, "Film");
    if (bins != 16 && bins != 81) throw Error{IMOPS_E_INVALID, "Invalid Film config: spectral_bins must be 16 or 81"};
    s.film.bins = (int)bins;
    const std::string grain = str_param(t, "grain", "splitmix", "Film");
    if (grain == "philox") s.film.philox = true;
    else if (grain != "splitmix") throw Error{IMOPS_E_INVALID, "Invalid Film config"};
    auto nonneg = [&](const char* key) -> float {
      if (!param(t, key)) return -1.0f;
      const double v = num_param(t, key, 0.0, "Film");
      if (!(v >= 0.0 && v < 1e6)) throw Error{IMOPS_E_INVALID, std::string("Invalid Film config: ") + key};
      return (float)v;
    };
    s.film.grain_sigma = nonneg("grain_sigma");
    s.film.hal_radius = nonneg("halation_radius");
    s.film.hal_strength = nonneg("halation_strength");
  } else {
    s.kind = K_OTHER;
  }
  return s;
}

bool is_pointwise(const Stage& s) {
  return s.kind == K_CFACOEFFS || s.kind == K_EXP || s.kind == K_GAMMA || s.kind == K_CST || s.kind == K_CONTRAST ||
         s.kind == K_HIGHLIGHT || s.kind == K_LCH || s.kind == K_SIGMOID || s.kind == K_BASELINE;
}

}  // namespace

void run_stages(imops_ctx* ctx, const std::vector<Stage>& stages, imops_image* img) {
  size_t k = 0;
  while (k < stages.size()) {
    const Stage& s = stages[k];
    // a roctx range per stage (one for a fused run of elementwise stages)
    const MarkerRange range(is_pointwise(s) ? "elementwise stages" : s.name.c_str());
    if (is_pointwise(s)) {
      // Fuse the run of elementwise stages, tracking the colour space as the
      // stages would update it.
      require_rgb(img, s.name.c_str());
      PointwiseProgram prog;
      int cs = img->color_space;
      for (; k < stages.size() && is_pointwise(stages[k]); ++k) {
        const Stage& e = stages[k];
        if (prog.n + kMaxOpsPerStage > PointwiseProgram::kMax) {
          stage_pointwise(ctx, img, prog);
          prog = PointwiseProgram();
        }
        switch (e.kind) {
          case K_CFACOEFFS:
            if (!img->has_wb) throw Error{IMOPS_E_STATE, "wb_coeffs needs to be set to perform highlight_reconstruction"};
            push_wb(prog, img->wb);
            break;
          case K_EXP: push_exp(prog, (float)e.ev); break;
          case K_CONTRAST: push_contrast(prog, (float)e.contrast); break;
          case K_HIGHLIGHT:
            if (!img->has_wb) throw Error{IMOPS_E_STATE, "wb_coeffs needs to be set to perform highlight_reconstruction"};
            push_highlight(prog, img->wb);
            break;
          case K_LCH: push_lch(prog, cs, (float)e.lch[0], (float)e.lch[1], (float)e.lch[2]); break;
          case K_SIGMOID: push_sigmoid(prog); break;
          case K_BASELINE: push_baseline_exposure(prog, img); break;
          case K_GAMMA: push_gamma(prog, (float)e.gamma); break;
          case K_CST: {
            const int target = parse_color_space(trim(e.target));
            if (target < 0) throw Error{IMOPS_E_INVALID, "Color space not recognized: " + e.target};
            if (cs == IMOPS_CS_NONE) {
              if (img->calib.empty())
                throw Error{IMOPS_E_STATE, "A calibration matrix must be set to perform a camera cst transform"};
              push_camera_cst(prog, img->calib.data(), img->calib.size(), target);
            } else {
              push_convert(prog, cs, target);
            }
            cs = target;
            break;
          }
          default: break;
        }
      }
      stage_pointwise(ctx, img, prog);
      img->color_space = cs;
      continue;
    }
    switch (s.kind) {
      case K_DEMOSAIC: stage_demosaic(ctx, img, s.algorithm); break;
      case K_BM3D: stage_bm3d(ctx, img, s.bm3d); break;
      case K_CHROMA_BM3D: stage_chroma_bm3d(ctx, img, (float)s.intensity); break;
      case K_LUMA_CHROMA: stage_luma_chroma_denoise(ctx, img, s.radius, (float)s.epsilon); break;
      case K_VIGNETTE: stage_vignette(ctx, img, (float)s.strength); break;
      case K_FILM: stage_film(ctx, img, s.film); break;
      default:
        throw Error{IMOPS_E_UNSUPPORTED, "stage " + s.name + " is not implemented by the GPU library"};
    }
    ++k;
  }
}

// ------------------------------------------------------------------ bursts
// A burst of frames of uneven sizes and CFA patterns through one stage list.
// Frames are grouped by everything the stage list reads besides pixel data
// (size, CFA and crop, white balance, calibration matrix, shutter, colour
// space).  The first frame of every group runs eagerly, largest group first, so
// every scratch buffer and per-shape table reaches its final size and content.
// Each group with more frames is then captured once into a HIP graph on a
// staging image of its shape and replayed per frame: the frame's input planes
// are copied into the staging image, the graph runs, the output is copied back.
namespace {

std::string shape_key(const imops_image* f) {
  std::string k = std::to_string(f->width) + "x" + std::to_string(f->height) + "|" + (f->has_raw ? "r" : "-") +
                  (f->has_rgb ? "c" : "-") + "|" + std::to_string(f->color_space) + "|";
  if (f->has_cfa) k += f->cfa.name + "@" + std::to_string(f->crop_x) + "," + std::to_string(f->crop_y);
  auto add_floats = [&](const float* v, size_t m) {
    k += "|";
    for (size_t i = 0; i < m; ++i) {
      uint32_t u;
      std::memcpy(&u, &v[i], sizeof u);
      k += std::to_string(u) + ",";
    }
  };
  if (f->has_wb) add_floats(f->wb, 4);
  add_floats(f->calib.data(), f->calib.size());
  if (f->has_shutter) add_floats(&f->shutter, 1);
  k += std::string("|e") + (f->has_baseline ? "b" : "-") + (f->has_f_number ? "f" : "-") + (f->has_iso ? "i" : "-");
  const float e[3] = {f->baseline_ev, f->f_number, f->iso};
  add_floats(e, 3);
  if (f->has_opcodes) {
    k += "|o";
    for (uint8_t b : f->opcode_list3) k += std::to_string(b) + ",";
  }
  return k;
}

struct BurstShape {
  std::vector<size_t> frames;
  imops_image* staging = nullptr;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
};

void run_burst_graphs(imops_ctx* ctx, const std::vector<Stage>& stages, imops_image* const* frames, size_t n) {
  // bound u16 frames become f32 raw planes: the captured graphs read each frame's copy
  for (size_t i = 0; i < n; ++i) materialize_raw(ctx, frames[i]);
  // every frame's state before the stage list runs (the eager first frames change theirs)
  std::vector<imops_image> pre(n);
  std::vector<BurstShape> shapes;
  std::vector<std::string> keys;
  std::vector<size_t> shape_of(n);
  for (size_t i = 0; i < n; ++i) {
    pre[i].copy_state_from(*frames[i]);
    const std::string k = shape_key(frames[i]);
    size_t j = 0;
    while (j < keys.size() && keys[j] != k) ++j;
    if (j == keys.size()) {
      keys.push_back(k);
      shapes.emplace_back();
    }
    shapes[j].frames.push_back(i);
    shape_of[i] = j;
  }
  std::vector<size_t> order(shapes.size());
  for (size_t j = 0; j < order.size(); ++j) order[j] = j;
  std::stable_sort(order.begin(), order.end(), [&](size_t a, size_t b) {
    const imops_image* fa = frames[shapes[a].frames[0]];
    const imops_image* fb = frames[shapes[b].frames[0]];
    return fa->width * fa->height > fb->width * fb->height;
  });
  for (size_t j : order) run_stages(ctx, stages, frames[shapes[j].frames[0]]);

  const bool timing = ctx->timer.enabled;
  ctx->timer.enabled = false;
  ctx->in_burst = true;
  auto cleanup = [&] {
    ctx->timer.enabled = timing;
    ctx->capturing = false;
    ctx->in_burst = false;
    (void)hipStreamSynchronize(ctx->stream);
    for (BurstShape& sh : shapes) {
      if (sh.exec) (void)hipGraphExecDestroy(sh.exec);
      if (sh.graph) (void)hipGraphDestroy(sh.graph);
      if (sh.staging) {
        sh.staging->raw.release();
        sh.staging->rgb.release();
        sh.staging->tmp.release();
        delete sh.staging;
      }
    }
  };
  try {
    for (size_t j : order) {
      BurstShape& sh = shapes[j];
      if (sh.frames.size() < 2) continue;
      imops_image* S = new imops_image();
      sh.staging = S;
      S->copy_state_from(pre[sh.frames[0]]);  // all metadata the stages read, as the frames have it
      S->ctx = ctx;
      const size_t px = S->width * S->height;
      if (S->has_raw) S->raw.reserve(px * sizeof(float));
      S->rgb.reserve(px * 3 * sizeof(float));
      void* before[SCR_COUNT];
      for (int b = 0; b < SCR_COUNT; ++b) before[b] = ctx->scr[b].ptr;
      ctx->capturing = true;
      IMOPS_HIP_CHECK(hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeGlobal));
      try {
        run_stages(ctx, stages, S);
      } catch (...) {
        hipGraph_t g = nullptr;
        (void)hipStreamEndCapture(ctx->stream, &g);
        if (g) (void)hipGraphDestroy(g);
        throw;
      }
      IMOPS_HIP_CHECK(hipStreamEndCapture(ctx->stream, &sh.graph));
      ctx->capturing = false;
      for (int b = 0; b < SCR_COUNT; ++b)
        if (ctx->scr[b].ptr != before[b]) throw Error{IMOPS_E_STATE, "burst: scratch moved during graph capture"};
      IMOPS_HIP_CHECK(hipGraphInstantiate(&sh.exec, sh.graph, nullptr, nullptr, 0));
    }
    for (size_t i = 0; i < n; ++i) {
      BurstShape& sh = shapes[shape_of[i]];
      if (i == sh.frames[0]) continue;
      imops_image* F = frames[i];
      imops_image* S = sh.staging;  // holds the state the captured stage list left (shape, flags, colour space)
      const size_t px = pre[i].width * pre[i].height;
      if (pre[i].has_raw)
        IMOPS_HIP_CHECK(hipMemcpyAsync(S->raw.ptr, F->raw.ptr, px * sizeof(float), hipMemcpyDeviceToDevice, ctx->stream));
      if (pre[i].has_rgb)
        IMOPS_HIP_CHECK(hipMemcpyAsync(S->rgb.ptr, F->rgb.ptr, px * 3 * sizeof(float), hipMemcpyDeviceToDevice, ctx->stream));
      IMOPS_HIP_CHECK(hipGraphLaunch(sh.exec, ctx->stream));
      const size_t px_out = S->width * S->height;  // a preview demosaic bins the frame
      if (S->has_rgb) {
        F->rgb.reserve(px_out * 3 * sizeof(float));
        IMOPS_HIP_CHECK(hipMemcpyAsync(F->rgb.ptr, S->rgb.ptr, px_out * 3 * sizeof(float), hipMemcpyDeviceToDevice,
                                       ctx->stream));
      }
      F->copy_state_from(*S);
    }
  } catch (...) {
    cleanup();
    throw;
  }
  cleanup();
}

}  // namespace

}  // namespace imops

using namespace imops;

extern "C" {

int imops_pipeline_parse(const char* text, size_t len, imops_pipeline** out) {
  return guarded([&] {
    if (!text || !out) throw Error{IMOPS_E_INVALID, "pipeline_parse: null argument"};
    *out = nullptr;
    const std::string s(text, len);
    std::vector<Table> mods;
    try {
      mods = parse_toml(s);
    } catch (const Error& toml_err) {
      try {
        mods = parse_json(s);
      } catch (const Error&) {
        throw toml_err;
      }
    }
    auto p = std::make_unique<imops_pipeline>();
    for (const Table& t : mods) p->stages.push_back(make_stage(t));
    *out = p.release();
  });
}

int imops_pipeline_free(imops_pipeline* p) {
  delete p;
  return IMOPS_OK;
}

size_t imops_pipeline_num_stages(const imops_pipeline* p) { return p ? p->stages.size() : 0; }

const char* imops_pipeline_stage_name(const imops_pipeline* p, size_t i) {
  if (!p || i >= p->stages.size()) return nullptr;
  return p->stages[i].name.c_str();
}

int imops_pipeline_run(imops_ctx* ctx, imops_pipeline* p, imops_image* img) {
  return guarded([&] {
    if (!ctx || !p || !img) throw Error{IMOPS_E_INVALID, "pipeline_run: null argument"};
    run_stages(ctx, p->stages, img);
  });
}

int imops_pipeline_run_burst(imops_ctx* ctx, imops_pipeline* p, imops_image* const* frames, size_t n, int use_graphs) {
  return guarded([&] {
    if (!ctx || !p || (!frames && n)) throw Error{IMOPS_E_INVALID, "pipeline_run_burst: null argument"};
    for (size_t i = 0; i < n; ++i)
      if (!frames[i] || frames[i]->ctx != ctx)
        throw Error{IMOPS_E_INVALID, "pipeline_run_burst: frame " + std::to_string(i) + " is null or of another context"};
    if (!use_graphs) {
      for (size_t i = 0; i < n; ++i) run_stages(ctx, p->stages, frames[i]);
      return;
    }
    run_burst_graphs(ctx, p->stages, frames, n);
  });
}

}  // extern "C"
