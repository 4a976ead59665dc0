/*
 * scenefile.cpp -- the reference's scene files, read on the host into the
 * flat scene descriptor of the C ABI (clv_scene_file_parse).
 *
 * Restates the scene assembly of clovers-cli/src/scenefile.rs:21-117 and
 * clovers/src/objects.rs:59-188: internally tagged objects, materials and
 * textures, untagged MaterialInit (a shared name or an inline material) and
 * ColorInit (a legacy encoded-sRGB array or one typed colour), the priority
 * list, the default material appended last, one top-level object per element
 * of a top-level ObjectList / STL / PLY, and the background colour.  Mesh
 * files follow clovers/src/objects/stl.rs:38-96 and ply.rs:37-103 (readers:
 * stl_io 0.10.0 and ply-rs 0.1.3 formats; rotation: nalgebra 0.34.1
 * Rotation3::from_euler_angles = Rz(yaw) Ry(pitch) Rx(roll)), evaluated in
 * f32 in the reference's operation order (compiled with -ffp-contract=off).
 * Colour conversions (palette 0.7.6) are evaluated in f64 and rounded once
 * to f32.  What the reference's schema allows but this build does not render
 * (GLTF, ConstantMedium boundaries other than Sphere/Boxy, other illuminants
 * and colour spaces) is CLV_EUNSUPPORTED; what makes the reference panic is
 * CLV_EINVAL.
 */
#include "../../include/clovers_hip.h"

#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <strings.h>

#include <algorithm>
#include <array>
#include <memory>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

namespace {

/* ---------------------------------------------------------------- JSON --- */
struct J {
  enum T { NUL, BOOL, NUM, STR, ARR, OBJ } t = NUL;
  bool b = false;
  double n = 0.0;
  std::string s;
  std::vector<J> a;
  std::vector<std::pair<std::string, J>> o;
  const J *get(const char *k) const {
    if (t != OBJ) return nullptr;
    for (const auto &kv : o)
      if (kv.first == k) return &kv.second;
    return nullptr;
  }
};

struct Fail : std::runtime_error {
  int status;
  Fail(int st, const std::string &m) : std::runtime_error(m), status(st) {}
};
[[noreturn]] void invalid(const std::string &m) { throw Fail(CLV_EINVAL, m); }
[[noreturn]] void unsupported(const std::string &m) { throw Fail(CLV_EUNSUPPORTED, m); }

struct Parser {
  const char *p, *end;
  void ws() {
    while (p < end && (*p == ' ' || *p == '\t' || *p == '\n' || *p == '\r')) p++;
  }
  [[noreturn]] void bad(const char *what) { invalid(std::string("JSON: ") + what); }
  void lit(const char *w) {
    size_t k = strlen(w);
    if ((size_t)(end - p) < k || memcmp(p, w, k) != 0) bad("unexpected token");
    p += k;
  }
  static void utf8(std::string &s, unsigned c) {
    if (c < 0x80) {
      s += (char)c;
    } else if (c < 0x800) {
      s += (char)(0xC0 | (c >> 6));
      s += (char)(0x80 | (c & 0x3F));
    } else if (c < 0x10000) {
      s += (char)(0xE0 | (c >> 12));
      s += (char)(0x80 | ((c >> 6) & 0x3F));
      s += (char)(0x80 | (c & 0x3F));
    } else {
      s += (char)(0xF0 | (c >> 18));
      s += (char)(0x80 | ((c >> 12) & 0x3F));
      s += (char)(0x80 | ((c >> 6) & 0x3F));
      s += (char)(0x80 | (c & 0x3F));
    }
  }
  unsigned hex4() {
    if (end - p < 4) bad("short \\u escape");
    unsigned v = 0;
    for (int i = 0; i < 4; i++) {
      char c = *p++;
      v <<= 4;
      if (c >= '0' && c <= '9') v |= (unsigned)(c - '0');
      else if (c >= 'a' && c <= 'f') v |= (unsigned)(c - 'a' + 10);
      else if (c >= 'A' && c <= 'F') v |= (unsigned)(c - 'A' + 10);
      else bad("bad \\u escape");
    }
    return v;
  }
  std::string str() {
    if (p >= end || *p != '"') bad("expected a string");
    p++;
    std::string s;
    while (p < end && *p != '"') {
      char c = *p++;
      if (c != '\\') {
        s += c;
        continue;
      }
      if (p >= end) bad("unterminated escape");
      char e = *p++;
      switch (e) {
        case '"': s += '"'; break;
        case '\\': s += '\\'; break;
        case '/': s += '/'; break;
        case 'b': s += '\b'; break;
        case 'f': s += '\f'; break;
        case 'n': s += '\n'; break;
        case 'r': s += '\r'; break;
        case 't': s += '\t'; break;
        case 'u': {
          unsigned c1 = hex4();
          if (c1 >= 0xD800 && c1 < 0xDC00 && end - p >= 6 && p[0] == '\\' && p[1] == 'u') {
            p += 2;
            unsigned c2 = hex4();
            c1 = 0x10000 + ((c1 - 0xD800) << 10) + (c2 - 0xDC00);
          }
          utf8(s, c1);
          break;
        }
        default: bad("bad escape");
      }
    }
    if (p >= end) bad("unterminated string");
    p++;
    return s;
  }
  J value(int depth) {
    if (depth > 256) bad("nesting too deep");
    ws();
    if (p >= end) bad("unexpected end");
    J v;
    char c = *p;
    if (c == '{') {
      v.t = J::OBJ;
      p++;
      ws();
      if (p < end && *p == '}') { p++; return v; }
      for (;;) {
        ws();
        std::string k = str();
        ws();
        if (p >= end || *p != ':') bad("expected ':'");
        p++;
        v.o.emplace_back(std::move(k), value(depth + 1));
        ws();
        if (p < end && *p == ',') { p++; continue; }
        if (p < end && *p == '}') { p++; return v; }
        bad("expected ',' or '}'");
      }
    }
    if (c == '[') {
      v.t = J::ARR;
      p++;
      ws();
      if (p < end && *p == ']') { p++; return v; }
      for (;;) {
        v.a.push_back(value(depth + 1));
        ws();
        if (p < end && *p == ',') { p++; continue; }
        if (p < end && *p == ']') { p++; return v; }
        bad("expected ',' or ']'");
      }
    }
    if (c == '"') { v.t = J::STR; v.s = str(); return v; }
    if (c == 't') { lit("true"); v.t = J::BOOL; v.b = true; return v; }
    if (c == 'f') { lit("false"); v.t = J::BOOL; return v; }
    if (c == 'n') { lit("null"); return v; }
    /* number: strtod rounds correctly, as every JSON reader should */
    std::string num;
    while (p < end && (strchr("+-0123456789.eE", *p) != nullptr)) num += *p++;
    if (num.empty()) bad("unexpected character");
    char *q = nullptr;
    v.t = J::NUM;
    v.n = strtod(num.c_str(), &q);
    if (!q || *q) bad("bad number");
    return v;
  }
};

J parse_json(const char *text) {
  Parser ps{text, text + strlen(text)};
  J v = ps.value(0);
  ps.ws();
  if (ps.p != ps.end) invalid("JSON: trailing characters");
  return v;
}

/* --------------------------------------------------------- JSON helpers --- */
double num(const J *v, const char *what) {
  if (!v || v->t != J::NUM) invalid(std::string("expected a number for ") + what);
  return v->n;
}
double num_or(const J &o, const char *k, double dflt) {
  const J *v = o.get(k);
  return v ? num(v, k) : dflt;
}
const J &need(const J &o, const char *k) {
  const J *v = o.get(k);
  if (!v) invalid(std::string("missing field '") + k + "'");
  return *v;
}
std::string kind_of(const J &o) {
  const J *k = o.get("kind");
  if (!k || k->t != J::STR) invalid("missing 'kind'");
  return k->s;
}
/* a 3-vector: an array, or an object with the given field names */
void vec3(const J &v, double out[3], const char *n0 = "x", const char *n1 = "y", const char *n2 = "z") {
  if (v.t == J::ARR) {
    if (v.a.size() != 3) invalid("expected 3 components");
    for (int i = 0; i < 3; i++) out[i] = num(&v.a[i], "vector component");
    return;
  }
  if (v.t == J::OBJ) {
    out[0] = num(v.get(n0), n0);
    out[1] = num(v.get(n1), n1);
    out[2] = num(v.get(n2), n2);
    return;
  }
  invalid("expected a 3-vector");
}
void vec3f(const J &v, float out[3]) {
  double d[3];
  vec3(v, d);
  for (int i = 0; i < 3; i++) out[i] = (float)d[i];
}

/* -------------------------------------------------------------- colour --- */
/* sRGB primaries, D65 white, Bradford cone matrix (colorinit.rs:45-60 via palette) */
struct Colour {
  double m_srgb_to_d65[3][3], m_d65_to_e[3][3];
  static void inv3(const double a[3][3], double r[3][3]) {
    double det = a[0][0] * (a[1][1] * a[2][2] - a[1][2] * a[2][1]) - a[0][1] * (a[1][0] * a[2][2] - a[1][2] * a[2][0]) +
                 a[0][2] * (a[1][0] * a[2][1] - a[1][1] * a[2][0]);
    r[0][0] = (a[1][1] * a[2][2] - a[1][2] * a[2][1]) / det;
    r[0][1] = (a[0][2] * a[2][1] - a[0][1] * a[2][2]) / det;
    r[0][2] = (a[0][1] * a[1][2] - a[0][2] * a[1][1]) / det;
    r[1][0] = (a[1][2] * a[2][0] - a[1][0] * a[2][2]) / det;
    r[1][1] = (a[0][0] * a[2][2] - a[0][2] * a[2][0]) / det;
    r[1][2] = (a[0][2] * a[1][0] - a[0][0] * a[1][2]) / det;
    r[2][0] = (a[1][0] * a[2][1] - a[1][1] * a[2][0]) / det;
    r[2][1] = (a[0][1] * a[2][0] - a[0][0] * a[2][1]) / det;
    r[2][2] = (a[0][0] * a[1][1] - a[0][1] * a[1][0]) / det;
  }
  static void mul(const double a[3][3], const double b[3][3], double r[3][3]) {
    for (int i = 0; i < 3; i++)
      for (int j = 0; j < 3; j++) r[i][j] = a[i][0] * b[0][j] + a[i][1] * b[1][j] + a[i][2] * b[2][j];
  }
  Colour() {
    const double xy[3][2] = {{0.64, 0.33}, {0.30, 0.60}, {0.15, 0.06}};
    const double white[3] = {0.95047, 1.0, 1.08883};
    double P[3][3], Pi[3][3];
    for (int c = 0; c < 3; c++) {
      P[0][c] = xy[c][0] / xy[c][1];
      P[1][c] = 1.0;
      P[2][c] = (1.0 - xy[c][0] - xy[c][1]) / xy[c][1];
    }
    inv3(P, Pi);
    double S[3];
    for (int i = 0; i < 3; i++) S[i] = Pi[i][0] * white[0] + Pi[i][1] * white[1] + Pi[i][2] * white[2];
    for (int i = 0; i < 3; i++)
      for (int j = 0; j < 3; j++) m_srgb_to_d65[i][j] = P[i][j] * S[j];
    const double B[3][3] = {{0.8951, 0.2664, -0.1614}, {-0.7502, 1.7135, 0.0367}, {0.0389, -0.0685, 1.0296}};
    double Bi[3][3], D[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}}, T[3][3];
    inv3(B, Bi);
    for (int i = 0; i < 3; i++) {
      double rs = B[i][0] * white[0] + B[i][1] * white[1] + B[i][2] * white[2];
      double rd = B[i][0] + B[i][1] + B[i][2];  /* white E = (1, 1, 1) */
      D[i][i] = rd / rs;
    }
    mul(D, B, T);
    mul(Bi, T, m_d65_to_e);
  }
  static double eotf(double c) { return c <= 0.04045 ? c / 12.92 : pow((c + 0.055) / 1.055, 2.4); }
  static void apply(const double m[3][3], const double v[3], double r[3]) {
    for (int i = 0; i < 3; i++) r[i] = m[i][0] * v[0] + m[i][1] * v[1] + m[i][2] * v[2];
  }
  void lin_srgb_to_e(const double rgb[3], float out[3]) const {
    double x[3], e[3];
    apply(m_srgb_to_d65, rgb, x);
    apply(m_d65_to_e, x, e);
    for (int i = 0; i < 3; i++) out[i] = (float)e[i];
  }
  void srgb_to_e(const double rgb[3], float out[3]) const {
    double l[3] = {eotf(rgb[0]), eotf(rgb[1]), eotf(rgb[2])};
    lin_srgb_to_e(l, out);
  }
  /* ColorInit (colorinit.rs:17-60): legacy [r, g, b] / {red, green, blue} are encoded sRGB */
  void init(const J &c, float out[3]) const {
    double v[3];
    if (c.t == J::ARR) {
      vec3(c, v);
      srgb_to_e(v, out);
      return;
    }
    if (c.t != J::OBJ || c.o.size() != 1) invalid("unrecognised colour");
    const std::string &tag = c.o[0].first;
    const J &val = c.o[0].second;
    if (tag == "hex") {
      if (val.t != J::STR) invalid("hex colour must be a string");
      std::string h = val.s;
      if (!h.empty() && h[0] == '#') h = h.substr(1);
      if (h.size() == 3) h = std::string{h[0], h[0], h[1], h[1], h[2], h[2]};
      if (h.size() != 6) invalid("bad hex colour");
      for (int i = 0; i < 3; i++) v[i] = (double)strtol(h.substr(2 * i, 2).c_str(), nullptr, 16) / 255.0;
      srgb_to_e(v, out);
    } else if (tag == "lin_srgb") {
      vec3(val, v, "red", "green", "blue");
      lin_srgb_to_e(v, out);
    } else if (tag == "srgb") {
      vec3(val, v, "red", "green", "blue");
      srgb_to_e(v, out);
    } else if (tag == "xyz_e") {
      vec3(val, v);
      for (int i = 0; i < 3; i++) out[i] = (float)v[i];
    } else if (tag == "xyz_d65") {
      double e[3];
      vec3(val, v);
      apply(m_d65_to_e, v, e);
      for (int i = 0; i < 3; i++) out[i] = (float)e[i];
    } else {
      unsupported("colour space '" + tag + "' is not supported");
    }
  }
};

const Colour &colour() {
  static const Colour c;
  return c;
}

/* ----------------------------------------------------------- materials --- */
/* Texture (textures.rs:22-109) into the material's texture fields */
void texture(const J *t, const float dflt[3], clv_material &r) {
  if (!t) {  /* Texture::default() = SolidColor::default() (solid_color.rs:50-56) */
    r.texture = CLV_TEX_SOLID;
    memcpy(r.color, dflt, sizeof(r.color));
    return;
  }
  const std::string k = kind_of(*t);
  if (k == "SolidColor") {
    r.texture = CLV_TEX_SOLID;
    colour().init(need(*t, "color"), r.color);
  } else if (k == "IlluminantD65" || k == "IlluminantD50") {
    r.texture = k == "IlluminantD65" ? CLV_TEX_ILLUMINANT_D65 : CLV_TEX_ILLUMINANT_D50;
    r.color[0] = r.color[1] = r.color[2] = 0.0f;
    r.intensity = (float)num_or(*t, "intensity", 1.0);
  } else if (k == "SpatialChecker" || k == "SurfaceChecker") {
    /* defaults: even 0.8, odd 0.5 Xyz<E>; density 1 / 10 (spatial_checker.rs:46-61, surface_checker.rs:44-59) */
    const bool spatial = k == "SpatialChecker";
    r.texture = spatial ? CLV_TEX_SPATIAL_CHECKER : CLV_TEX_SURFACE_CHECKER;
    const J *e = t->get("even"), *o = t->get("odd");
    if (e) colour().init(*e, r.color);
    else r.color[0] = r.color[1] = r.color[2] = 0.8f;
    if (o) colour().init(*o, r.odd_color);
    else r.odd_color[0] = r.odd_color[1] = r.odd_color[2] = 0.5f;
    r.density = (float)num_or(*t, "density", spatial ? 1.0 : 10.0);
  } else {
    unsupported("texture kind '" + k + "' is not supported");
  }
}

/* Material (materials.rs:63-103) -> clv_material; m == nullptr is Material::default() */
clv_material material(const J *m) {
  clv_material r;
  memset(&r, 0, sizeof(r));
  r.intensity = 1.0f;
  r.refractive_index = 1.5f;  /* dielectric.rs:27-29 */
  r.cauchy_a = 1.5046f;       /* dispersive.rs:39-45 */
  r.cauchy_b = 0.00420f;
  r.film_index = 1.5f;        /* thin_film.rs:57-63 */
  r.film_thickness = 500.0f;
  r.color[0] = r.color[1] = r.color[2] = 1.0f;
  const float grey[3] = {0.5f, 0.5f, 0.5f}, bright[3] = {100.0f, 100.0f, 100.0f};
  std::string k = "Lambertian";
  if (m) {
    const J *kk = m->get("kind");
    if (kk && kk->t == J::STR) k = kk->s;
  }
  if (k == "Lambertian") {
    r.kind = CLV_MAT_LAMBERTIAN;
    texture(m ? m->get("albedo") : nullptr, grey, r);
  } else if (k == "Metal") {
    r.kind = CLV_MAT_METAL;
    texture(m->get("albedo"), grey, r);
    r.fuzz = (float)num_or(*m, "fuzz", 0.0);  /* not clamped on the serde path (metal.rs:60) */
  } else if (k == "Dielectric") {
    r.kind = CLV_MAT_DIELECTRIC;
    r.refractive_index = (float)num_or(*m, "refractive_index", 1.5);
    if (const J *c = m->get("color")) vec3f(*c, r.color);  /* Xyz<E> directly (dielectric.rs:19-23) */
  } else if (k == "Dispersive") {
    r.kind = CLV_MAT_DISPERSIVE;
    r.cauchy_a = (float)num_or(*m, "cauchy_a", 1.5046);
    r.cauchy_b = (float)num_or(*m, "cauchy_b", 0.00420);
  } else if (k == "DiffuseLight") {
    r.kind = CLV_MAT_DIFFUSE_LIGHT;  /* default emit SolidColor Xyz(100, 100, 100) (diffuse_light.rs:23-30) */
    texture(m->get("emit"), bright, r);
  } else if (k == "ConeLight") {
    /* both fields are required by the schema (cone_light.rs:13-18 has no serde defaults) */
    if (!m->get("spread") || !m->get("emit")) invalid("ConeLight needs 'spread' and 'emit'");
    r.kind = CLV_MAT_CONE_LIGHT;
    r.spread = (float)num(m->get("spread"), "spread");
    texture(m->get("emit"), bright, r);
  } else if (k == "Isotropic") {  /* isotropic.rs:13-19: albedo texture, default grey */
    r.kind = CLV_MAT_ISOTROPIC;
    texture(m->get("albedo"), grey, r);
    if (r.texture == CLV_TEX_SURFACE_CHECKER) unsupported("a SurfaceChecker texture on an Isotropic material");
  } else {
    unsupported("material kind '" + k + "' is not supported");
  }
  if (m) {
    const J *tf = m->get("thin_film");
    if (tf && tf->t != J::NUL) {
      r.thin_film = 1u;
      r.film_index = (float)num_or(*tf, "refractive_index", 1.5);
      r.film_thickness = (float)num_or(*tf, "thickness", 500.0);
    }
  }
  return r;
}

/* --------------------------------------------------------------- meshes --- */
struct Tri {
  float v[3][3];
};

std::vector<char> read_file(const std::string &path) {
  FILE *f = fopen(path.c_str(), "rb");
  if (!f) invalid("cannot open '" + path + "'");
  std::vector<char> buf;
  char tmp[1 << 16];
  size_t k;
  while ((k = fread(tmp, 1, sizeof(tmp), f)) > 0) buf.insert(buf.end(), tmp, tmp + k);
  fclose(f);
  return buf;
}

template <typename T>
T load_le(const char *p) {
  T v;
  memcpy(&v, p, sizeof(T));
  return v;
}

/* STL, binary (80-byte header, count, 50-byte facets) or ASCII */
std::vector<Tri> read_stl(const std::string &path) {
  std::vector<char> d = read_file(path);
  std::vector<Tri> out;
  if (d.size() >= 84) {
    uint32_t n = load_le<uint32_t>(d.data() + 80);
    if (84ull + 50ull * n == d.size()) {
      out.resize(n);
      for (uint32_t i = 0; i < n; i++) memcpy(out[i].v, d.data() + 84 + 50ull * i + 12, 36);
      return out;
    }
  }
  std::string s(d.begin(), d.end());
  size_t a = s.find_first_not_of(" \t\r\n");
  if (a == std::string::npos || strncasecmp(s.c_str() + a, "solid", 5) != 0) invalid("'" + path + "' is not an STL file");
  std::vector<std::array<float, 3>> verts;
  size_t pos = 0;
  while (pos < s.size()) {
    size_t e = s.find('\n', pos);
    if (e == std::string::npos) e = s.size();
    std::string line = s.substr(pos, e - pos);
    pos = e + 1;
    char key[32] = {0};
    if (sscanf(line.c_str(), "%31s", key) != 1) continue;
    if (!strcasecmp(key, "vertex")) {
      double x, y, z;
      char w[32];
      if (sscanf(line.c_str(), "%31s %lf %lf %lf", w, &x, &y, &z) != 4) invalid("bad STL vertex");
      verts.push_back({(float)x, (float)y, (float)z});
    } else if (!strcasecmp(key, "endloop")) {
      if (verts.size() != 3) invalid("STL facet without 3 vertices");
      Tri t;
      for (int i = 0; i < 3; i++)
        for (int c = 0; c < 3; c++) t.v[i][c] = verts[i][c];
      out.push_back(t);
      verts.clear();
    }
  }
  return out;
}

/* PLY: ascii / binary little or big endian; vertex x, y, z and face vertex_indices */
struct PlyProp {
  std::string name;
  bool list = false;
  std::string type, item;  /* scalar type, or list count type + item type */
};
struct PlyElem {
  std::string name;
  uint64_t count = 0;
  std::vector<PlyProp> props;
};

int type_size(const std::string &t) {
  if (t == "char" || t == "int8" || t == "uchar" || t == "uint8") return 1;
  if (t == "short" || t == "int16" || t == "ushort" || t == "uint16") return 2;
  if (t == "int" || t == "int32" || t == "uint" || t == "uint32" || t == "float" || t == "float32") return 4;
  if (t == "double" || t == "float64") return 8;
  invalid("unknown PLY type '" + t + "'");
}
/* canonical kind: i (signed int), u (unsigned int), f (float) */
char type_kind(const std::string &t) {
  if (t == "float" || t == "float32" || t == "double" || t == "float64") return 'f';
  if (t[0] == 'u') return 'u';
  return 'i';
}

struct PlyReader {
  const std::vector<char> &d;
  size_t pos;
  bool big;
  double read(const std::string &t) {
    int n = type_size(t);
    if (pos + (size_t)n > d.size()) invalid("PLY data ends early");
    unsigned char b[8];
    memcpy(b, d.data() + pos, (size_t)n);
    pos += (size_t)n;
    if (big) std::reverse(b, b + n);
    char k = type_kind(t);
    if (k == 'f') {
      if (n == 4) { float f; memcpy(&f, b, 4); return f; }
      double x; memcpy(&x, b, 8); return x;
    }
    if (n == 1) return k == 'u' ? (double)b[0] : (double)(int8_t)b[0];
    if (n == 2) { uint16_t u; memcpy(&u, b, 2); return k == 'u' ? (double)u : (double)(int16_t)u; }
    uint32_t u; memcpy(&u, b, 4); return k == 'u' ? (double)u : (double)(int32_t)u;
  }
};

std::vector<Tri> read_ply(const std::string &path) {
  std::vector<char> d = read_file(path);
  size_t pos = 0;
  auto line = [&]() {
    size_t e = pos;
    while (e < d.size() && d[e] != '\n') e++;
    std::string l(d.begin() + (long)pos, d.begin() + (long)e);
    pos = e < d.size() ? e + 1 : e;
    if (!l.empty() && l.back() == '\r') l.pop_back();
    return l;
  };
  if (line() != "ply") invalid("'" + path + "' is not a PLY file");
  std::string fmt;
  std::vector<PlyElem> els;
  for (;;) {
    if (pos >= d.size()) invalid("PLY header without end_header");
    std::string l = line();
    char w[5][64] = {{0}};
    int k = sscanf(l.c_str(), "%63s %63s %63s %63s %63s", w[0], w[1], w[2], w[3], w[4]);
    if (k < 1 || !strcmp(w[0], "comment") || !strcmp(w[0], "obj_info")) continue;
    if (!strcmp(w[0], "format")) fmt = w[1];
    else if (!strcmp(w[0], "element")) {
      PlyElem e;
      e.name = w[1];
      e.count = strtoull(w[2], nullptr, 10);
      els.push_back(e);
    } else if (!strcmp(w[0], "property")) {
      if (els.empty()) invalid("PLY property before element");
      PlyProp p;
      if (!strcmp(w[1], "list")) { p.list = true; p.type = w[2]; p.item = w[3]; p.name = w[4]; type_size(p.item); }
      else { p.type = w[1]; p.name = w[2]; }
      type_size(p.type);
      els.back().props.push_back(p);
    } else if (!strcmp(w[0], "end_header")) break;
  }
  std::vector<std::array<float, 3>> verts;
  std::vector<std::array<int64_t, 3>> faces;
  const bool ascii = fmt == "ascii";
  if (!ascii && fmt != "binary_little_endian" && fmt != "binary_big_endian") invalid("unknown PLY format '" + fmt + "'");
  PlyReader rd{d, pos, fmt == "binary_big_endian"};
  /* ascii tokens */
  const char *tp = d.data() + pos, *tend = d.data() + d.size();
  auto token = [&]() {
    while (tp < tend && (*tp == ' ' || *tp == '\n' || *tp == '\r' || *tp == '\t')) tp++;
    if (tp >= tend) invalid("PLY data ends early");
    char *q;
    double v = strtod(tp, &q);
    if (q == tp) invalid("bad PLY number");
    tp = q;
    return v;
  };
  for (const PlyElem &e : els) {
    const bool is_v = e.name == "vertex", is_f = e.name == "face";
    int ix = -1, iy = -1, iz = -1, ii = -1;
    for (size_t k = 0; k < e.props.size(); k++) {
      const PlyProp &p = e.props[k];
      if (is_v && (p.name == "x" || p.name == "y" || p.name == "z")) {
        /* ply.rs:110-129: int, uint, float and double coordinates convert; the rest is unimplemented!() */
        if (p.list || type_size(p.type) < 4) unsupported("PLY: unsupported vertex property format");
        (p.name == "x" ? ix : p.name == "y" ? iy : iz) = (int)k;
      }
      if (is_f && p.name == "vertex_indices") {
        /* ply.rs:131-153: lists of unsigned or signed (|i|) integers except char */
        if (!p.list || type_kind(p.item) == 'f' || p.item == "char" || p.item == "int8")
          unsupported("PLY: unsupported face index format");
        ii = (int)k;
      }
    }
    if (is_v && (ix < 0 || iy < 0 || iz < 0)) invalid("PLY vertex element without x, y, z");
    if (is_f && ii < 0) invalid("PLY face element without vertex_indices");
    for (uint64_t r = 0; r < e.count; r++) {
      std::array<float, 3> v = {0, 0, 0};
      std::array<int64_t, 3> f = {0, 0, 0};
      for (size_t k = 0; k < e.props.size(); k++) {
        const PlyProp &p = e.props[k];
        if (!p.list) {
          double x = ascii ? token() : rd.read(p.type);
          if ((int)k == ix) v[0] = (float)x;
          if ((int)k == iy) v[1] = (float)x;
          if ((int)k == iz) v[2] = (float)x;
          continue;
        }
        double cnt = ascii ? token() : rd.read(p.type);
        if (!(cnt >= 0)) invalid("bad PLY list length");
        for (uint64_t j = 0; j < (uint64_t)cnt; j++) {
          double x = ascii ? token() : rd.read(p.item);
          if ((int)k == ii && j < 3) f[j] = (int64_t)fabs(x);
        }
        if ((int)k == ii && cnt < 3) invalid("PLY face with fewer than 3 vertices");
      }
      if (is_v) verts.push_back(v);
      if (is_f) faces.push_back(f);
    }
  }
  std::vector<Tri> out(faces.size());
  for (size_t i = 0; i < faces.size(); i++)
    for (int c = 0; c < 3; c++) {
      if ((uint64_t)faces[i][c] >= verts.size()) invalid("PLY face index out of range");
      for (int a = 0; a < 3; a++) out[i].v[c][a] = verts[(size_t)faces[i][c]][a];
    }
  return out;
}

/* (R p) * scale + center in f32, R = Rotation3::from_euler_angles (nalgebra
 * 0.34.1): entries as its ZYX formula, rows summed left to right */
void transform(std::vector<Tri> &tris, const float rot_deg[3], float scale, const float center[3]) {
  const float deg = 3.14159265358979323846f / 180.0f;  /* f32::to_radians */
  float s[3], c[3];
  for (int i = 0; i < 3; i++) {
    float a = rot_deg[i] * deg;
    s[i] = sinf(a);
    c[i] = cosf(a);
  }
  const float sr = s[0], cr = c[0], sp = s[1], cp = c[1], sy = s[2], cy = c[2];
  const float R[3][3] = {{cy * cp, cy * sp * sr - sy * cr, cy * sp * cr + sy * sr},
                         {sy * cp, sy * sp * sr + cy * cr, sy * sp * cr - cy * sr},
                         {-sp, cp * sr, cp * cr}};
  for (Tri &t : tris)
    for (int v = 0; v < 3; v++) {
      float p[3] = {t.v[v][0], t.v[v][1], t.v[v][2]};
      for (int r = 0; r < 3; r++) {
        float rp = (R[r][0] * p[0] + R[r][1] * p[1]) + R[r][2] * p[2];
        t.v[v][r] = rp * scale + center[r];
      }
    }
}

/* ---------------------------------------------------------------- scene --- */
struct Loader {
  std::vector<clv_object> objects;
  std::vector<clv_material> materials;
  std::vector<clv_xform> xforms;
  std::vector<std::pair<std::string, uint32_t>> shared;  /* name -> material; the default ("") last */
  std::string mesh_root;

  uint32_t add_material(const clv_material &m) {
    materials.push_back(m);
    return (uint32_t)materials.size() - 1;
  }
  /* MaterialInit (objects.rs:159-188): a name, falling back to the default, or an inline material */
  uint32_t material_of(const J *init, bool mesh) {
    if (init && init->t == J::OBJ) return add_material(material(init));
    std::string name;
    if (init && init->t == J::STR) name = init->s;
    else if (init && init->t != J::NUL) invalid("bad material reference");
    for (auto &kv : shared)
      if (kv.first == name) return kv.second;
    /* STL/PLY unwrap the lookup (stl.rs:50, ply.rs:43): no fallback */
    if (mesh) invalid("mesh material '" + name + "' not found (the reference panics)");
    for (auto &kv : shared)
      if (kv.first.empty()) return kv.second;
    invalid("no default material");
  }
  uint32_t add_xform(const std::vector<clv_xform_op> &ops) {
    if (ops.empty()) return CLV_NO_XFORM;
    if (ops.size() > CLV_MAX_XFORM_OPS) unsupported("more than 4 nested RotateY/Translate wrappers");
    clv_xform x;
    memset(&x, 0, sizeof(x));
    x.n_ops = (uint32_t)ops.size();
    for (size_t i = 0; i < ops.size(); i++) x.ops[i] = ops[i];
    xforms.push_back(x);
    return (uint32_t)xforms.size() - 1;
  }
  void push(uint32_t kind, const float *p, int np, uint32_t mat, bool priority, uint32_t xf) {
    clv_object o;
    memset(&o, 0, sizeof(o));
    o.kind = kind;
    o.material = mat;
    o.priority = priority ? 1u : 0u;
    o.xform = xf;
    for (int i = 0; i < np; i++) o.p[i] = p[i];
    objects.push_back(o);
  }
  void primitive(const J &o, bool priority, const std::vector<clv_xform_op> &ops) {
    const std::string k = kind_of(o);
    float p[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
    uint32_t kind;
    int np;
    if (k == "Sphere") {
      vec3f(need(o, "center"), p);
      p[3] = (float)num(o.get("radius"), "radius");
      kind = CLV_OBJ_SPHERE, np = 4;
    } else if (k == "Quad" || k == "Triangle") {
      vec3f(need(o, "q"), p);
      vec3f(need(o, "u"), p + 3);
      vec3f(need(o, "v"), p + 6);
      kind = k == "Quad" ? CLV_OBJ_QUAD : CLV_OBJ_TRIANGLE_QUV, np = 9;
    } else if (k == "Boxy") {
      vec3f(need(o, "corner_0"), p);
      vec3f(need(o, "corner_1"), p + 3);
      kind = CLV_OBJ_BOXY, np = 6;
    } else if (k == "MovingSphere") {  /* the file's times are not read (objects.rs:104-109) */
      vec3f(need(o, "center_0"), p);
      vec3f(need(o, "center_1"), p + 3);
      p[6] = (float)num(o.get("radius"), "radius");
      kind = CLV_OBJ_MOVING_SPHERE, np = 7;
    } else {
      unsupported("object kind '" + k + "' is not supported");
    }
    uint32_t mat = material_of(o.get("material"), false);
    push(kind, p, np, mat, priority, add_xform(ops));
  }
  /* ConstantMedium (objects.rs:99-103, constant_medium.rs:19-46): a Sphere or
   * Boxy boundary (its own wrappers move onto the medium: the same hit along
   * the same ray), density (default 0.1), texture (default SolidColor grey)
   * as the Isotropic phase function; the boundary's material is never read */
  void medium(const J &o, bool priority, std::vector<clv_xform_op> ops) {
    if (priority) invalid("a ConstantMedium cannot be a priority object (its random() is unimplemented: the reference panics)");
    const J *b = &need(o, "boundary");
    for (;;) {
      const std::string bk = kind_of(*b);
      if (bk != "Translate" && bk != "RotateY") break;
      clv_xform_op op;
      memset(&op, 0, sizeof(op));
      if (bk == "Translate") {
        op.kind = CLV_XF_TRANSLATE;
        vec3f(need(*b, "offset"), op.v);
      } else {
        op.kind = CLV_XF_ROTATE_Y;
        op.v[0] = (float)num(b->get("angle"), "angle");
      }
      ops.push_back(op);
      b = &need(*b, "object");
    }
    const std::string bk = kind_of(*b);
    const float density = (float)num_or(o, "density", 0.1);
    float p[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
    uint32_t kind;
    int np;
    if (bk == "Sphere") {
      vec3f(need(*b, "center"), p);
      p[3] = (float)num(b->get("radius"), "radius");
      p[4] = density;
      kind = CLV_OBJ_CONSTANT_MEDIUM_SPHERE, np = 5;
    } else if (bk == "Boxy") {
      vec3f(need(*b, "corner_0"), p);
      vec3f(need(*b, "corner_1"), p + 3);
      p[6] = density;
      kind = CLV_OBJ_CONSTANT_MEDIUM_BOXY, np = 7;
    } else {
      unsupported("a ConstantMedium boundary of kind '" + bk + "' is not supported");
    }
    clv_material iso = material(nullptr);  /* the defaults of every material field */
    iso.kind = CLV_MAT_ISOTROPIC;
    const float grey[3] = {0.5f, 0.5f, 0.5f};
    texture(o.get("texture"), grey, iso);
    if (iso.texture == CLV_TEX_SURFACE_CHECKER) unsupported("a SurfaceChecker texture on a ConstantMedium");
    push(kind, p, np, add_material(iso), false, add_xform(ops));
  }
  void mesh(const J &o, bool priority, bool ply) {
    std::string path = need(o, "path").t == J::STR ? need(o, "path").s : "";
    if (path.empty()) invalid("mesh without a path");
    if (!mesh_root.empty() && path[0] != '/') path = mesh_root + "/" + path;
    std::vector<Tri> tris = ply ? read_ply(path) : read_stl(path);
    float rot[3], center[3];
    vec3f(need(o, "rotation"), rot);
    vec3f(need(o, "center"), center);
    transform(tris, rot, (float)num(o.get("scale"), "scale"), center);
    uint32_t mat = material_of(o.get("material"), true);
    for (const Tri &t : tris) push(CLV_OBJ_TRIANGLE, &t.v[0][0], 9, mat, priority, CLV_NO_XFORM);
  }
  /* one top-level object (scenefile.rs:66-100) */
  void add(const J &o, bool priority, std::vector<clv_xform_op> ops, bool top) {
    const std::string k = kind_of(o);
    if (k == "Translate") {
      clv_xform_op op;
      op.kind = CLV_XF_TRANSLATE;
      vec3f(need(o, "offset"), op.v);
      ops.push_back(op);
      add(need(o, "object"), priority, ops, false);
      return;
    }
    if (k == "RotateY") {
      clv_xform_op op;
      op.kind = CLV_XF_ROTATE_Y;
      op.v[0] = (float)num(o.get("angle"), "angle");
      op.v[1] = op.v[2] = 0.0f;
      ops.push_back(op);
      add(need(o, "object"), priority, ops, false);
      return;
    }
    if (k == "ObjectList" || k == "STL" || k == "PLY") {
      if (!ops.empty()) unsupported("a list inside RotateY/Translate (hitable.rs:158 returns None on any miss)");
      if (!top) invalid("Recursing too hard in scene initialization");
      if (k == "STL" || k == "PLY") {
        mesh(o, priority, k == "PLY");
        return;
      }
      const J &l = need(o, "objects");
      if (l.t != J::ARR) invalid("ObjectList objects must be a list");
      for (const J &c : l.a) add(c, priority, {}, false);  /* flattened one level (scenefile.rs:84-94) */
      return;
    }
    if (k == "ConstantMedium") {
      medium(o, priority, ops);
      return;
    }
    primitive(o, priority, ops);
  }
};

bool flag(const J &o, const char *k) {
  const J *v = o.get(k);
  return v && v->t == J::BOOL && v->b;
}

}  // namespace

struct clv_scene_file {
  std::vector<clv_object> objects;
  std::vector<clv_material> materials;
  std::vector<clv_xform> xforms;
  clv_scene_desc desc;
};

extern "C" {

clv_status clv_scene_file_parse(const char *json, const char *mesh_root, uint32_t bvh_algorithm,
                                clv_scene_file **out, char *err, uint32_t err_len) {
  if (err && err_len) err[0] = 0;
  if (!out) return CLV_EINVAL;
  *out = nullptr;
  auto report = [&](int st, const std::string &m) {
    if (err && err_len) snprintf(err, err_len, "%s", m.c_str());
    return (clv_status)st;
  };
  if (!json) return report(CLV_EINVAL, "json is NULL");
  if (bvh_algorithm > CLV_BVH_SAH) return report(CLV_EINVAL, "unknown BVH algorithm");
  try {
    J sf = parse_json(json);
    if (sf.t != J::OBJ) invalid("a scene file is a JSON object");
    std::unique_ptr<clv_scene_file> f(new clv_scene_file());
    clv_scene_desc &d = f->desc;
    memset(&d, 0, sizeof(d));
    d.abi_version = CLV_ABI_VERSION;
    d.mis_bvh_algorithm = bvh_algorithm;
    /* background: legacy encoded sRGB (scenefile.rs:43-44) */
    double bg[3] = {0, 0, 0};
    if (const J *b = sf.get("background_color")) vec3(*b, bg, "red", "green", "blue");
    colour().srgb_to_e(bg, d.background);
    const J &cam = need(sf, "camera");
    vec3f(need(cam, "look_from"), d.camera.look_from);
    vec3f(need(cam, "look_at"), d.camera.look_at);
    vec3f(need(cam, "up"), d.camera.up);
    d.camera.vertical_fov = (float)num(cam.get("vertical_fov"), "vertical_fov");
    d.camera.aperture = (float)num(cam.get("aperture"), "aperture");
    d.camera.focus_distance = (float)num(cam.get("focus_distance"), "focus_distance");
    d.camera.time_0 = (float)num_or(sf, "time_0", 0.0);
    d.camera.time_1 = (float)num_or(sf, "time_1", 1.0);
    Loader ld;
    if (mesh_root) ld.mesh_root = mesh_root;
    if (const J *ms = sf.get("materials")) {
      if (ms->t != J::ARR) invalid("materials must be a list");
      for (const J &m : ms->a) {
        const J *nm = m.get("name");
        ld.shared.emplace_back(nm && nm->t == J::STR ? nm->s : "", ld.add_material(material(&m)));
      }
    }
    ld.shared.emplace_back("", ld.add_material(material(nullptr)));  /* SharedMaterial::default() (scenefile.rs:58-59) */
    if (const J *os = sf.get("objects")) {
      if (os->t != J::ARR) invalid("objects must be a list");
      for (const J &o : os->a) ld.add(o, flag(o, "priority"), {}, true);
    }
    f->objects.swap(ld.objects);
    f->materials.swap(ld.materials);
    f->xforms.swap(ld.xforms);
    d.n_objects = (uint32_t)f->objects.size();
    d.objects = f->objects.data();
    d.n_materials = (uint32_t)f->materials.size();
    d.materials = f->materials.data();
    d.n_xforms = (uint32_t)f->xforms.size();
    d.xforms = f->xforms.empty() ? nullptr : f->xforms.data();
    *out = f.release();
    return CLV_OK;
  } catch (const Fail &e) {
    return report(e.status, e.what());
  } catch (const std::bad_alloc &) {
    return report(CLV_ENOMEM, "out of host memory");
  } catch (const std::exception &e) {
    return report(CLV_EINTERNAL, e.what());
  }
}

const clv_scene_desc *clv_scene_file_desc(const clv_scene_file *f) { return f ? &f->desc : nullptr; }

void clv_scene_file_free(clv_scene_file *f) { delete f; }

}  // extern "C"
