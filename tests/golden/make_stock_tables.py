"""Generate the oracle's film stock tables from the reference's stock sources.

Reads pichromatic/src/film/stock/{bw_stub,color_neg_200,portra_400,ektar_100,
fuji_pro_400h,ektachrome_e100,trix_400}.rs under the reference checkout as text,
evaluates each stock's `load()` (struct literals, numeric fields, the
per-file curve helpers such as gaussian_curve / orange_mask_epsilon /
panchromatic_sensitivity evaluated in f64 at every wavelength of the grid) and
writes oracle/oracle_stock_tables.hpp: plain data (numbers only), sampled on the
reference's 16-bin grid (400..700 nm, 20 nm) and on the 81-bin grid
(380..780 nm, 5 nm).  f32 fields of the reference (Microns, IsoSpeed, the
silver-halide fraction, d_max, gamma, reciprocity, the point-spread widths,
adjacency and DIR weights) are written as the source's decimal literal with an
f suffix, so C++ rounds them to f32 exactly as rustc does.

The oracle builds its film tables from this file only; the GPU library has its
own hand-written records (imops_amd/csrc/film_stock_data.hpp), so a
transcription error in either shows up as a film parity failure.

    python tests/golden/make_stock_tables.py [REFERENCE_ROOT]
"""
from __future__ import annotations

import math
import re
import sys
from pathlib import Path

ROOT = Path(__file__).resolve().parents[2]
REF = Path(sys.argv[1]) if len(sys.argv) > 1 and __name__ == "__main__" else Path("/root/reference")
STOCK_DIR = REF / "pichromatic" / "src" / "film" / "stock"
OUT = ROOT / "oracle" / "oracle_stock_tables.hpp"

# StockId order (film/stock/mod.rs:124-133)
STOCKS = [("BwStub", "bw_stub.rs"), ("ColorNeg200", "color_neg_200.rs"), ("Portra400", "portra_400.rs"),
          ("Ektar100", "ektar_100.rs"), ("FujiPro400H", "fuji_pro_400h.rs"),
          ("EktachromeE100", "ektachrome_e100.rs"), ("TriX400", "trix_400.rs")]
GRID16 = [400.0 + 20.0 * i for i in range(16)]
GRID81 = [380.0 + 5.0 * i for i in range(81)]
KINDS = {"Emulsion": 0, "Filter": 1, "Overcoat": 2, "Antihalation": 3, "Support": 4}


# ---------------------------------------------------------------- tokens
TOKEN = re.compile(r"""
    (?P<num>\d[\d_]*(?:\.\d[\d_]*)?(?:[eE][+-]?\d+)?(?:_?f(?:32|64))?)
  | (?P<str>"(?:[^"\\]|\\.)*")
  | (?P<id>[A-Za-z_][A-Za-z0-9_]*(?:::[A-Za-z_][A-Za-z0-9_]*)*!?)
  | (?P<op>\|&|::|->|=>|[-+*/(){}\[\],;:.=|&<>!])
  | (?P<ws>\s+)
""", re.X)


def strip_comments(src: str) -> str:
    src = src.split("#[cfg(test)]")[0]  # the unit tests are not stock data
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return re.sub(r"//[^\n]*", "", src)


def tokenize(src: str):
    out, pos = [], 0
    while pos < len(src):
        m = TOKEN.match(src, pos)
        if not m:
            raise SyntaxError(f"unexpected {src[pos:pos + 20]!r}")
        pos = m.end()
        kind = m.lastgroup
        if kind != "ws":
            out.append((kind, m.group()))
    return out


class Lit:
    """A numeric literal, keeping its source text (for f32 fields)."""

    def __init__(self, text: str):
        self.text = text.replace("_f64", "").replace("_f32", "").replace("f64", "").replace("f32", "").replace("_", "")
        self.value = float(self.text)


def num(v):
    return v.value if isinstance(v, Lit) else float(v)


class Curve:
    def __init__(self, fn):
        self.fn = fn

    def samples(self, grid):
        return [self.fn(l) for l in grid]


# ---------------------------------------------------------------- parser/evaluator
class Parser:
    def __init__(self, tokens, env, helpers):
        self.t, self.i, self.env, self.helpers = tokens, 0, env, helpers

    def peek(self, k=0):
        return self.t[self.i + k] if self.i + k < len(self.t) else ("eof", "")

    def take(self, want=None):
        tok = self.t[self.i]
        if want is not None and tok[1] != want:
            raise SyntaxError(f"expected {want!r}, got {tok[1]!r}")
        self.i += 1
        return tok

    def expr(self):
        v = self.term()
        while self.peek()[1] in ("+", "-"):
            op = self.take()[1]
            r = self.term()
            v = num(v) + num(r) if op == "+" else num(v) - num(r)
        return v

    def term(self):
        v = self.unary()
        while self.peek()[1] in ("*", "/"):
            op = self.take()[1]
            r = self.unary()
            v = num(v) * num(r) if op == "*" else num(v) / num(r)
        return v

    def unary(self):
        if self.peek()[1] == "-":
            self.take()
            v = self.unary()
            return Lit("-" + v.text) if isinstance(v, Lit) else -num(v)
        return self.postfix(self.primary())

    def args(self):
        self.take("(")
        out = []
        while self.peek()[1] != ")":
            out.append(self.expr())
            if self.peek()[1] == ",":
                self.take()
        self.take(")")
        return out

    def postfix(self, v):
        while self.peek()[1] == ".":
            self.take()
            name = self.take()[1]
            a = self.args()
            x = num(v)
            if name == "exp": v = math.exp(x)
            elif name == "ln": v = math.log(x)
            elif name == "sin": v = math.sin(x)
            elif name == "cos": v = math.cos(x)
            elif name == "sqrt": v = math.sqrt(x)
            elif name == "abs": v = abs(x)
            elif name == "powi":
                n = int(num(a[0]))
                v = 1.0
                for _ in range(n):
                    v = v * x
            elif name == "powf": v = math.pow(x, num(a[0]))
            elif name == "max": v = max(x, num(a[0]))
            elif name == "min": v = min(x, num(a[0]))
            else: raise SyntaxError(f"method {name}")
        return v

    def primary(self):
        kind, text = self.peek()
        if kind == "num":
            self.take()
            return Lit(text)
        if kind == "str":
            self.take()
            return text[1:-1]
        if text == "(":
            self.take()
            v = self.expr()
            self.take(")")
            return v
        if text == "vec!":
            self.take()
            self.take("[")
            out = []
            while self.peek()[1] != "]":
                out.append(self.expr())
                if self.peek()[1] == ",":
                    self.take()
            self.take("]")
            return out
        if kind == "id":
            self.take()
            if text == "std::f64::consts::PI":
                return math.pi
            if text == "None":
                return None
            if text in ("true", "false"):
                return text == "true"
            if text.startswith("LayerKind::"):
                return KINDS[text.split("::")[1]]
            if self.peek()[1] == "{":  # struct literal
                self.take("{")
                fields = {"__type__": text}
                while self.peek()[1] != "}":
                    key = self.take()[1]
                    self.take(":")
                    fields[key] = self.expr()
                    if self.peek()[1] == ",":
                        self.take()
                self.take("}")
                return fields
            if self.peek()[1] == "(":
                a = self.args()
                if text in ("Some", "Microns", "IsoSpeed"):
                    return a[0]
                if text == "SpectralCurve::constant":
                    c = num(a[0])
                    return Curve(lambda l, c=c: c)
                if text in self.helpers:
                    return self.helpers[text](*[num(x) for x in a])
                raise SyntaxError(f"call {text}")
            if text in self.env:
                return self.env[text]
            raise SyntaxError(f"name {text}")
        raise SyntaxError(f"unexpected {text!r}")


def block_body(tokens, start):
    """Tokens of the {...} block opening at tokens[start] (exclusive of the braces)."""
    depth, j = 0, start
    while True:
        if tokens[j][1] == "{": depth += 1
        elif tokens[j][1] == "}":
            depth -= 1
            if depth == 0:
                return tokens[start + 1:j], j
        j += 1


def run_block(tokens, env, helpers):
    """`let x = expr;` statements then a final expression; returns its value."""
    p = Parser(tokens, dict(env), helpers)
    while p.i < len(p.t):
        if p.peek()[1] == "let":
            p.take()
            name = p.take()[1]
            if p.peek()[1] == ":":  # type annotation
                while p.peek()[1] != "=":
                    p.take()
            p.take("=")
            p.env[name] = p.expr()
            p.take(";")
        else:
            v = p.expr()
            return v, p.env
    return None, p.env


def helpers_of(tokens):
    """Curve helper functions: fn name(params) -> SpectralCurve { ... .map(|&l| { body }) ... }."""
    helpers = {}
    for k in range(len(tokens) - 1):
        if tokens[k][1] == "fn" and tokens[k + 1][1] != "load":
            name = tokens[k + 1][1]
            j = k + 2
            params = []
            depth = 0
            while True:
                t = tokens[j][1]
                if t == "(": depth += 1
                elif t == ")":
                    depth -= 1
                    if depth == 0:
                        break
                elif depth == 1 and tokens[j + 1][1] == ":" and tokens[j][0] == "id":
                    params.append(t)
                j += 1
            while tokens[j][1] != "{":
                j += 1
            body, _ = block_body(tokens, j)
            if not any(t[1] == "SpectralCurve::new" for t in body):
                continue
            m = next(i for i in range(len(body)) if body[i][1] == "|&")
            var = body[m + 1][1]
            inner_start = m + 3  # after `|&l|`
            if body[inner_start][1] == "{":
                inner, _ = block_body(body, inner_start)
            else:
                end = inner_start
                depth = 0
                while not (depth == 0 and body[end][1] == ")"):
                    depth += body[end][1] == "("
                    depth -= body[end][1] == ")"
                    end += 1
                inner = body[inner_start:end]

            def make(inner=inner, params=params, var=var, helpers=helpers):
                def curve_fn(*vals):
                    env0 = dict(zip(params, vals))
                    return Curve(lambda l: num(run_block(inner, {**env0, var: l}, helpers)[0]))
                return curve_fn
            helpers[name] = make()
    return helpers


def load_stock(path: Path) -> dict:
    toks = tokenize(strip_comments(path.read_text()))
    helpers = helpers_of(toks)
    k = next(i for i in range(len(toks) - 1) if toks[i][1] == "fn" and toks[i + 1][1] == "load")
    while toks[k][1] != "{":
        k += 1
    body, _ = block_body(toks, k)
    cut = next(i for i in range(len(body)) if body[i][1] == "stock" and body[i + 1][1] == "." )
    _, env = run_block(body[:cut], {}, helpers)
    return env["stock"]


# ---------------------------------------------------------------- C++ emission
def f32(v) -> str:
    if isinstance(v, Lit):
        t = v.text
        if "." not in t and "e" not in t.lower():
            t += ".0"
        return t + "f"
    return repr(float(v)) + "f"


def f64(v) -> str:
    return repr(float(num(v)))


def curve(c) -> str:
    if c is None:
        return "{false, {}, {}}"
    s16 = ", ".join(repr(x) for x in c.samples(GRID16))
    s81 = ", ".join(repr(x) for x in c.samples(GRID81))
    return "{true, {" + s16 + "}, {" + s81 + "}}"


def layer(L) -> str:
    cs = L.get("crystal_size")
    cp = L.get("coupler")
    fields = [
        str(L["kind"]), f32(L["depth_from_surface"]), f32(L["thickness"]), curve(L.get("spectral_sensitivity")),
        "true" if cs else "false", f64(cs["mu_ln"]) if cs else "0.0", f64(cs["sigma_ln"]) if cs else "0.0",
        f32(L["silver_halide_fraction"]), "true" if cp else "false",
        curve(cp["epsilon"]) if cp else curve(None), curve(cp.get("mask_epsilon")) if cp else curve(None),
        f32(cp["d_max"]) if cp else "0.0f", f32(L["gamma_contrast"]), f64(L["capture_k"]), f32(L["reciprocity_p"]),
        "true" if L["is_reversal"] else "false"]
    return "{" + ", ".join(fields) + "}"


def stock_cpp(s) -> str:
    ah = s["antihalation"]
    dirm = s["dir_inhibition_matrix"]
    rows = []
    for i in range(8):
        r = dirm[i] if i < len(dirm) else []
        rows.append("{" + ", ".join(f32(r[j]) if j < len(r) else "0.0f" for j in range(8)) + "}")
    parts = [
        '"' + s["name"] + '"', f32(s["box_iso"]), str(len(s["layers"])),
        "{" + ",\n      ".join(layer(L) for L in s["layers"]) + "}",
        curve(ah["reflectance"]), f32(ah["psf_local_um"]), f32(ah["psf_halation_um"]),
        f32(s["developer_diffusion_length"]), f32(s["adjacency_beta"]), f32(s["dir_diffusion_length"]),
        str(len(dirm)), "{" + ", ".join(rows) + "}", curve(s["scanner_light"])]
    return "  {" + ",\n    ".join(parts) + "}"


def render() -> str:
    """The header text generated from the reference sources."""
    stocks = []
    for name, fname in STOCKS:
        s = load_stock(STOCK_DIR / fname)
        assert s["name"] == name, (s["name"], name)
        stocks.append(s)
    text = f"""// GENERATED by tests/golden/make_stock_tables.py from the reference's
// pichromatic/src/film/stock/*.rs (each stock's load()); do not edit.
// Test infrastructure: the oracle's film stock data, independent of the GPU
// library's hand-written records.  Curves are sampled in f64 on the 16-bin
// grid 400..700 nm (step 20) and the 81-bin grid 380..780 nm (step 5).
#pragma once

namespace oracle_stocks {{

struct OCurve {{
  bool present;
  double s16[16];
  double s81[81];
}};

// LayerKind: 0 emulsion, 1 filter, 2 overcoat, 3 antihalation, 4 support
struct OLayer {{
  int kind;
  float depth_um, thickness_um;
  OCurve sensitivity;
  bool has_crystal;
  double mu_ln, sigma_ln;
  float silver_halide_fraction;
  bool has_coupler;
  OCurve epsilon, mask_epsilon;
  float d_max, gamma_contrast;
  double capture_k;
  float reciprocity_p;
  bool is_reversal;
}};

struct OStock {{
  const char* name;
  float box_iso;
  int n_layers;
  OLayer layers[12];
  OCurve ah_reflectance;
  float psf_local_um, psf_halation_um;
  float developer_diffusion_um, adjacency_beta, dir_diffusion_um;
  int n_dir;
  float dir[8][8];
  OCurve scanner_light;
}};

// StockId order (film/stock/mod.rs:124-133)
static const OStock kStocks[{len(stocks)}] = {{
""" + ",\n".join(stock_cpp(s) for s in stocks) + """
};

}  // namespace oracle_stocks
"""
    return text


def main():
    OUT.write_text(render())
    print(f"wrote {OUT}")


if __name__ == "__main__":
    main()
