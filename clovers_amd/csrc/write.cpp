/*
 * write.cpp -- image files of a rendered pixel buffer (clovers-cli/src/write.rs).
 *
 * clv_write_png: Xyz<E> -> sRGB 8-bit PNG, as write.rs:18-84 does with palette
 * 0.7.6: Bradford adaptation E -> D65, XYZ(D65) -> linear sRGB, the sRGB
 * transfer function (IEC 61966-2-1), clamp to [0, 1] and round to 8 bits;
 * tEXt chunks "Comment" and "Software".  The colour arithmetic is f64 here
 * (palette's is f32), so a channel that lands within ~1e-6 of a rounding
 * boundary may differ by one code value.
 *
 * clv_write_exr: the linear buffer as an OpenEXR scanline image with float
 * channels R = X, G = Y, B = Z and the chromaticities attribute of the XYZ
 * convention (red (1,0), green (0,1), blue (0,0), white (1/3,1/3)), as
 * write.rs:86-125; stored uncompressed (the reference's exr crate picks a
 * lossless compression, which changes the bytes, not the pixels).
 */
#include "../../include/clovers_hip.h"

#include <math.h>
#include <stdio.h>
#include <string.h>
#include <zlib.h>

#include <string>
#include <vector>

namespace {

/* XYZ(E) -> linear sRGB(D65), built from the same primaries, white and
 * Bradford matrix as the scene-file colour path, inverted */
struct ToSrgb {
  double m[3][3];
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
  ToSrgb() {
    const double xy[3][2] = {{0.64, 0.33}, {0.30, 0.60}, {0.15, 0.06}};
    const double white[3] = {0.95047, 1.0, 1.08883};
    double P[3][3], Pi[3][3], M[3][3], Mi[3][3];
    for (int c = 0; c < 3; c++) {
      P[0][c] = xy[c][0] / xy[c][1];
      P[1][c] = 1.0;
      P[2][c] = (1.0 - xy[c][0] - xy[c][1]) / xy[c][1];
    }
    inv3(P, Pi);
    double S[3];
    for (int i = 0; i < 3; i++) S[i] = Pi[i][0] * white[0] + Pi[i][1] * white[1] + Pi[i][2] * white[2];
    for (int i = 0; i < 3; i++)
      for (int j = 0; j < 3; j++) M[i][j] = P[i][j] * S[j];  /* linear sRGB -> XYZ(D65) */
    inv3(M, Mi);
    const double B[3][3] = {{0.8951, 0.2664, -0.1614}, {-0.7502, 1.7135, 0.0367}, {0.0389, -0.0685, 1.0296}};
    double Bi[3][3], D[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}}, T[3][3], A[3][3];
    inv3(B, Bi);
    for (int i = 0; i < 3; i++) {
      double rd = B[i][0] * white[0] + B[i][1] * white[1] + B[i][2] * white[2];  /* destination D65 */
      double rs = B[i][0] + B[i][1] + B[i][2];                                     /* source E */
      D[i][i] = rd / rs;
    }
    mul(D, B, T);
    mul(Bi, T, A);  /* XYZ(E) -> XYZ(D65) */
    mul(Mi, A, m);
  }
  static double oetf(double c) { return c <= 0.0031308 ? 12.92 * c : 1.055 * pow(c, 1.0 / 2.4) - 0.055; }
  static uint8_t to8(double c) {
    if (!(c > 0.0)) return 0;  /* negative and NaN */
    if (c >= 1.0) return 255;
    return (uint8_t)floor(c * 255.0 + 0.5);
  }
  void px(const float xyz[3], uint8_t out[3]) const {
    for (int i = 0; i < 3; i++) {
      double l = m[i][0] * xyz[0] + m[i][1] * xyz[1] + m[i][2] * xyz[2];
      out[i] = to8(oetf(l > 0.0 ? l : 0.0));
    }
  }
};

void be32(std::vector<uint8_t> &b, uint32_t v) {
  b.push_back((uint8_t)(v >> 24));
  b.push_back((uint8_t)(v >> 16));
  b.push_back((uint8_t)(v >> 8));
  b.push_back((uint8_t)v);
}

void chunk(std::vector<uint8_t> &out, const char type[4], const uint8_t *data, size_t n) {
  be32(out, (uint32_t)n);
  size_t start = out.size();
  out.insert(out.end(), type, type + 4);
  out.insert(out.end(), data, data + n);
  uLong crc = crc32(0L, Z_NULL, 0);
  crc = crc32(crc, out.data() + start, (uInt)(n + 4));
  be32(out, (uint32_t)crc);
}

bool write_all(const char *path, const void *data, size_t n) {
  FILE *f = fopen(path, "wb");
  if (!f) return false;
  bool ok = fwrite(data, 1, n, f) == n;
  ok = (fclose(f) == 0) && ok;
  return ok;
}

template <typename T>
void le(std::vector<uint8_t> &b, T v) {
  const uint8_t *p = reinterpret_cast<const uint8_t *>(&v);
  b.insert(b.end(), p, p + sizeof(T));
}
void attr(std::vector<uint8_t> &h, const char *name, const char *type, const std::vector<uint8_t> &val) {
  h.insert(h.end(), name, name + strlen(name) + 1);
  h.insert(h.end(), type, type + strlen(type) + 1);
  le<int32_t>(h, (int32_t)val.size());
  h.insert(h.end(), val.begin(), val.end());
}

}  // namespace

extern "C" {

clv_status clv_write_png(const float *xyz, uint32_t width, uint32_t height, const char *path, const char *comment) {
  if (!xyz || !path || width == 0 || height == 0 || width > (1u << 16) || height > (1u << 16)) return CLV_EINVAL;
  static const ToSrgb conv;
  const size_t stride = 1 + 3 * (size_t)width;  /* filter byte 0 (None) + RGB */
  std::vector<uint8_t> raw(stride * height);
  for (uint32_t y = 0; y < height; y++) {
    uint8_t *row = raw.data() + stride * y;
    row[0] = 0;
    for (uint32_t x = 0; x < width; x++) conv.px(xyz + 3 * ((size_t)y * width + x), row + 1 + 3 * x);
  }
  uLongf zn = compressBound((uLong)raw.size());
  std::vector<uint8_t> z(zn);
  if (compress2(z.data(), &zn, raw.data(), (uLong)raw.size(), 6) != Z_OK) return CLV_EINTERNAL;
  std::vector<uint8_t> out = {0x89, 'P', 'N', 'G', '\r', '\n', 0x1A, '\n'};
  std::vector<uint8_t> ihdr;
  be32(ihdr, width);
  be32(ihdr, height);
  ihdr.insert(ihdr.end(), {8, 2, 0, 0, 0});  /* 8-bit RGB, deflate, adaptive filters, no interlace */
  chunk(out, "IHDR", ihdr.data(), ihdr.size());
  const std::string texts[2] = {std::string("Comment") + '\0' + (comment ? comment : ""),
                                std::string("Software") + '\0' + "clovers_amd"};
  for (const std::string &t : texts) chunk(out, "tEXt", reinterpret_cast<const uint8_t *>(t.data()), t.size());
  chunk(out, "IDAT", z.data(), zn);
  chunk(out, "IEND", nullptr, 0);
  return write_all(path, out.data(), out.size()) ? CLV_OK : CLV_EINVAL;
}

clv_status clv_write_exr(const float *xyz, uint32_t width, uint32_t height, const char *path) {
  if (!xyz || !path || width == 0 || height == 0 || width > (1u << 20) || height > (1u << 20)) return CLV_EINVAL;
  std::vector<uint8_t> f = {0x76, 0x2f, 0x31, 0x01, 2, 0, 0, 0};  /* magic, version 2, single-part scanline */
  std::vector<uint8_t> v;
  /* channels in the file's (alphabetical) order: B = Z, G = Y, R = X; 2 = FLOAT */
  for (const char *c : {"B", "G", "R"}) {
    v.push_back((uint8_t)c[0]);
    v.push_back(0);
    le<int32_t>(v, 2);
    v.insert(v.end(), {0, 0, 0, 0});  /* pLinear + reserved */
    le<int32_t>(v, 1);
    le<int32_t>(v, 1);
  }
  v.push_back(0);
  attr(f, "channels", "chlist", v);
  v.clear();
  for (float c : {1.0f, 0.0f, 0.0f, 1.0f, 0.0f, 0.0f, 1.0f / 3.0f, 1.0f / 3.0f}) le<float>(v, c);
  attr(f, "chromaticities", "chromaticities", v);
  attr(f, "compression", "compression", {0});
  v.clear();
  for (int32_t c : {0, 0, (int32_t)width - 1, (int32_t)height - 1}) le<int32_t>(v, c);
  attr(f, "dataWindow", "box2i", v);
  attr(f, "displayWindow", "box2i", v);
  attr(f, "lineOrder", "lineOrder", {0});
  v.clear();
  le<float>(v, 1.0f);
  attr(f, "pixelAspectRatio", "float", v);
  v.clear();
  le<float>(v, 0.0f);
  le<float>(v, 0.0f);
  attr(f, "screenWindowCenter", "v2f", v);
  v.clear();
  le<float>(v, 1.0f);
  attr(f, "screenWindowWidth", "float", v);
  v.clear();
  const char *sw = "clovers_amd";
  v.insert(v.end(), sw, sw + strlen(sw));
  attr(f, "software", "string", v);
  f.push_back(0);  /* end of header */
  const size_t line_bytes = 3ull * 4ull * width;
  const size_t table = f.size();
  f.resize(table + 8ull * height);
  for (uint32_t y = 0; y < height; y++) {
    const uint64_t off = f.size();
    memcpy(f.data() + table + 8ull * y, &off, 8);
    le<int32_t>(f, (int32_t)y);
    le<int32_t>(f, (int32_t)line_bytes);
    const float *row = xyz + 3ull * width * y;
    for (int c = 2; c >= 0; c--)  /* B, G, R = Z, Y, X */
      for (uint32_t x = 0; x < width; x++) le<float>(f, row[3 * x + c]);
  }
  return write_all(path, f.data(), f.size()) ? CLV_OK : CLV_EINVAL;
}

}  // extern "C"
