// PNG files: output of RGBA8 images, the file Images::saveBufferPNG writes through
// lodepng::encode (Images.cu:22-36): 8-bit RGBA, non-interlaced, here with filter 0 on
// every scanline and zlib deflate; and input of textures (below).  Host only.
#include <algorithm>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>
#include <zlib.h>

#include "../../include/dumbray.h"

namespace dr {
int fail(int code, const std::string &msg);
}

namespace {
void put32(std::vector<unsigned char> &v, uint32_t x) {
  v.push_back((unsigned char)(x >> 24));
  v.push_back((unsigned char)(x >> 16));
  v.push_back((unsigned char)(x >> 8));
  v.push_back((unsigned char)x);
}
void chunk(std::vector<unsigned char> &out, const char *type, const std::vector<unsigned char> &data) {
  put32(out, (uint32_t)data.size());
  const size_t at = out.size();
  out.insert(out.end(), type, type + 4);
  out.insert(out.end(), data.begin(), data.end());
  const uLong crc = crc32(0L, out.data() + at, (uInt)(out.size() - at));
  put32(out, (uint32_t)crc);
}
}  // namespace

extern "C" int dr_png_write(const char *path, const uint8_t *rgba, int width, int height) {
  if (!path || !rgba || width <= 0 || height <= 0) return dr::fail(DR_ERR_ARG, "dr_png_write: bad arguments");
  const size_t row = (size_t)width * 4;
  std::vector<unsigned char> raw((row + 1) * (size_t)height);
  for (int y = 0; y < height; y++) {
    raw[(row + 1) * y] = 0;   // filter: none
    std::copy(rgba + row * y, rgba + row * (y + 1), raw.begin() + (row + 1) * y + 1);
  }
  uLongf zlen = compressBound((uLong)raw.size());
  std::vector<unsigned char> z(zlen);
  if (compress2(z.data(), &zlen, raw.data(), (uLong)raw.size(), 6) != Z_OK)
    return dr::fail(DR_ERR_ALLOC, "dr_png_write: deflate failed");
  z.resize(zlen);
  std::vector<unsigned char> out = {0x89, 'P', 'N', 'G', '\r', '\n', 0x1a, '\n'};
  std::vector<unsigned char> ihdr;
  put32(ihdr, (uint32_t)width);
  put32(ihdr, (uint32_t)height);
  ihdr.insert(ihdr.end(), {8, 6, 0, 0, 0});   // bit depth 8, colour type RGBA, deflate, filter 0, no interlace
  chunk(out, "IHDR", ihdr);
  chunk(out, "IDAT", z);
  chunk(out, "IEND", {});
  FILE *f = std::fopen(path, "wb");
  if (!f) return dr::fail(DR_ERR_ARG, std::string("dr_png_write: cannot open ") + path);
  const bool ok = std::fwrite(out.data(), 1, out.size(), f) == out.size();
  if (std::fclose(f) != 0 || !ok) return dr::fail(DR_ERR_ARG, std::string("dr_png_write: cannot write ") + path);
  return DR_OK;
}

// ---------------------------------------------------------------- reading
// PNG decoding to 8-bit RGBA, as lodepng::decode gives it to Images::getTexturePNG
// (Images.cu:38-56): every colour type (grey, RGB, palette, grey+alpha, RGBA) and bit depth
// (1-16), tRNS transparency, both interlace methods; 16-bit samples keep their high byte,
// sub-byte grey is scaled to 0-255.
namespace {
uint32_t rd32(const unsigned char *p) { return ((uint32_t)p[0] << 24) | ((uint32_t)p[1] << 16) | ((uint32_t)p[2] << 8) | p[3]; }

int paeth(int a, int b, int c) {
  const int p = a + b - c;
  const int pa = std::abs(p - a), pb = std::abs(p - b), pc = std::abs(p - c);
  if (pa <= pb && pa <= pc) return a;
  if (pb <= pc) return b;
  return c;
}

constexpr uint64_t kMaxPixels = 1ull << 28;   // 1 GiB of RGBA8

struct Png {
  uint32_t w = 0, h = 0;
  int depth = 0, ctype = 0, interlace = 0;
  std::vector<unsigned char> idat, plte, trns;
};

bool parse(const std::vector<unsigned char> &f, Png &p, std::string &err) {
  static const unsigned char sig[8] = {0x89, 'P', 'N', 'G', '\r', '\n', 0x1a, '\n'};
  if (f.size() < 8 || std::memcmp(f.data(), sig, 8) != 0) { err = "not a PNG file"; return false; }
  size_t at = 8;
  bool ihdr = false;
  while (at + 12 <= f.size()) {
    const uint32_t len = rd32(f.data() + at);
    if (at + 12 + (size_t)len > f.size()) { err = "truncated chunk"; return false; }
    const unsigned char *type = f.data() + at + 4, *data = f.data() + at + 8;
    if (!std::memcmp(type, "IHDR", 4) && len >= 13) {
      p.w = rd32(data);
      p.h = rd32(data + 4);
      p.depth = data[8];
      p.ctype = data[9];
      p.interlace = data[12];
      ihdr = true;
    } else if (!std::memcmp(type, "PLTE", 4)) {
      p.plte.assign(data, data + len);
    } else if (!std::memcmp(type, "tRNS", 4)) {
      p.trns.assign(data, data + len);
    } else if (!std::memcmp(type, "IDAT", 4)) {
      p.idat.insert(p.idat.end(), data, data + len);
    } else if (!std::memcmp(type, "IEND", 4)) {
      break;
    }
    at += 12 + (size_t)len;
  }
  if (!ihdr || p.w == 0 || p.h == 0 || p.w > (1u << 16) || p.h > (1u << 16)) { err = "bad or missing IHDR"; return false; }
  if ((uint64_t)p.w * p.h > kMaxPixels) { err = "image larger than 2^28 pixels"; return false; }
  const int d = p.depth, c = p.ctype;
  const bool ok = (c == 0 && (d == 1 || d == 2 || d == 4 || d == 8 || d == 16)) ||
                  (c == 3 && (d == 1 || d == 2 || d == 4 || d == 8)) || ((c == 2 || c == 4 || c == 6) && (d == 8 || d == 16));
  if (!ok || p.interlace > 1) { err = "unsupported colour type / bit depth / interlace"; return false; }
  if (c == 3 && p.plte.empty()) { err = "palette image without PLTE"; return false; }
  return true;
}

int channels(int ctype) { return ctype == 0 ? 1 : ctype == 2 ? 3 : ctype == 3 ? 1 : ctype == 4 ? 2 : 4; }

// Unfilters one (sub)image of w x h pixels from `in` (filter byte + row bytes per row) into
// `rows`, returning the bytes consumed, or 0 on a bad filter.
size_t unfilter(const unsigned char *in, size_t avail, uint32_t w, uint32_t h, int bpp_bits, std::vector<unsigned char> &rows) {
  const size_t stride = ((size_t)w * bpp_bits + 7) / 8, bpp = std::max(1, bpp_bits / 8);
  if (w == 0 || h == 0) { rows.clear(); return 0; }
  if (avail < (stride + 1) * h) return 0;
  rows.assign(stride * h, 0);
  for (uint32_t y = 0; y < h; y++) {
    const unsigned char ft = in[(stride + 1) * y];
    const unsigned char *src = in + (stride + 1) * y + 1;
    unsigned char *dst = &rows[stride * y];
    const unsigned char *up = y ? &rows[stride * (y - 1)] : nullptr;
    for (size_t i = 0; i < stride; i++) {
      const int a = i >= bpp ? dst[i - bpp] : 0, b = up ? up[i] : 0, c = (up && i >= bpp) ? up[i - bpp] : 0;
      int v;
      switch (ft) {
        case 0: v = src[i]; break;
        case 1: v = src[i] + a; break;
        case 2: v = src[i] + b; break;
        case 3: v = src[i] + ((a + b) >> 1); break;
        case 4: v = src[i] + paeth(a, b, c); break;
        default: return 0;
      }
      dst[i] = (unsigned char)v;
    }
  }
  return (stride + 1) * h;
}

// RGBA8 of pixel x of an unfiltered row.
void pixel_rgba(const Png &p, const unsigned char *row, uint32_t x, unsigned char out[4]) {
  const int d = p.depth, c = p.ctype;
  auto sample = [&](int k) -> int {   // channel k, 8 bits (16-bit: high byte)
    if (d == 16) return row[(x * channels(c) + k) * 2];
    return row[x * channels(c) + k];
  };
  auto sample16 = [&](int k) -> int {  // full-precision channel k for tRNS comparison
    if (d == 16) return (row[(x * channels(c) + k) * 2] << 8) | row[(x * channels(c) + k) * 2 + 1];
    return row[x * channels(c) + k];
  };
  if (c == 0 || c == 3) {
    int v;
    if (d < 8) {
      const int per = 8 / d, shift = 8 - d * (int)(x % per + 1);
      v = (row[x / per] >> shift) & ((1 << d) - 1);
    } else {
      v = (d == 16) ? sample16(0) : row[x];
    }
    if (c == 3) {
      const size_t n = p.plte.size() / 3;
      if ((size_t)v < n) { out[0] = p.plte[3 * v]; out[1] = p.plte[3 * v + 1]; out[2] = p.plte[3 * v + 2]; }
      else { out[0] = out[1] = out[2] = 0; }
      out[3] = (size_t)v < p.trns.size() ? p.trns[v] : 255;
      return;
    }
    const int g = (d == 16) ? (v >> 8) : (d == 8 ? v : v * 255 / ((1 << d) - 1));
    out[0] = out[1] = out[2] = (unsigned char)g;
    out[3] = 255;
    if (p.trns.size() >= 2 && v == ((p.trns[0] << 8) | p.trns[1])) out[3] = 0;
    return;
  }
  if (c == 2) {
    out[0] = (unsigned char)sample(0); out[1] = (unsigned char)sample(1); out[2] = (unsigned char)sample(2); out[3] = 255;
    if (p.trns.size() >= 6 && sample16(0) == ((p.trns[0] << 8) | p.trns[1]) && sample16(1) == ((p.trns[2] << 8) | p.trns[3]) &&
        sample16(2) == ((p.trns[4] << 8) | p.trns[5]))
      out[3] = 0;
    return;
  }
  if (c == 4) { out[0] = out[1] = out[2] = (unsigned char)sample(0); out[3] = (unsigned char)sample(1); return; }
  out[0] = (unsigned char)sample(0); out[1] = (unsigned char)sample(1); out[2] = (unsigned char)sample(2);
  out[3] = (unsigned char)sample(3);
}

bool decode(const std::vector<unsigned char> &file, std::vector<unsigned char> &rgba, uint32_t &w, uint32_t &h, std::string &err) {
  Png p;
  if (!parse(file, p, err)) return false;
  w = p.w;
  h = p.h;
  const int bits = p.depth * channels(p.ctype);
  // the scanline bytes the image needs (filter byte + row per row, per Adam7 pass)
  static const int X0[7] = {0, 4, 0, 2, 0, 1, 0}, Y0[7] = {0, 0, 4, 0, 2, 0, 1};
  static const int DX[7] = {8, 8, 4, 4, 2, 2, 1}, DY[7] = {8, 8, 8, 4, 4, 2, 2};
  auto pass_size = [&](int pass, uint32_t &pw, uint32_t &ph) {
    pw = (w > (uint32_t)X0[pass]) ? (w - X0[pass] + DX[pass] - 1) / DX[pass] : 0;
    ph = (h > (uint32_t)Y0[pass]) ? (h - Y0[pass] + DY[pass] - 1) / DY[pass] : 0;
  };
  size_t need = 0;
  if (p.interlace == 0) {
    need = (((size_t)w * bits + 7) / 8 + 1) * h;
  } else {
    for (int pass = 0; pass < 7; pass++) {
      uint32_t pw, ph;
      pass_size(pass, pw, ph);
      if (pw && ph) need += (((size_t)pw * bits + 7) / 8 + 1) * ph;
    }
  }
  // inflate (no further than the image needs)
  z_stream zs{};
  if (inflateInit(&zs) != Z_OK) { err = "zlib init"; return false; }
  std::vector<unsigned char> raw;
  zs.next_in = p.idat.data();
  zs.avail_in = (uInt)p.idat.size();
  unsigned char buf[1 << 16];
  int zr;
  do {
    zs.next_out = buf;
    zs.avail_out = sizeof(buf);
    zr = inflate(&zs, Z_NO_FLUSH);
    if (zr != Z_OK && zr != Z_STREAM_END && zr != Z_BUF_ERROR) { inflateEnd(&zs); err = "corrupt image data"; return false; }
    raw.insert(raw.end(), buf, buf + (sizeof(buf) - zs.avail_out));
    if (zr == Z_BUF_ERROR) break;                 // no progress possible: the input ended
  } while (zr != Z_STREAM_END && raw.size() < need);
  inflateEnd(&zs);
  rgba.assign((size_t)w * h * 4, 0);
  std::vector<unsigned char> rows;
  const size_t stride_bits = (size_t)bits;
  if (p.interlace == 0) {
    if (!unfilter(raw.data(), raw.size(), w, h, bits, rows)) { err = "bad scanlines"; return false; }
    const size_t stride = ((size_t)w * stride_bits + 7) / 8;
    for (uint32_t y = 0; y < h; y++)
      for (uint32_t x = 0; x < w; x++) pixel_rgba(p, &rows[stride * y], x, &rgba[4 * ((size_t)y * w + x)]);
    return true;
  }
  // Adam7
  size_t at = 0;
  for (int pass = 0; pass < 7; pass++) {
    uint32_t pw, ph;
    pass_size(pass, pw, ph);
    if (pw == 0 || ph == 0) continue;
    const size_t used = unfilter(raw.data() + at, raw.size() - at, pw, ph, bits, rows);
    if (!used) { err = "bad interlaced scanlines"; return false; }
    at += used;
    const size_t stride = ((size_t)pw * stride_bits + 7) / 8;
    for (uint32_t y = 0; y < ph; y++)
      for (uint32_t x = 0; x < pw; x++) {
        const size_t X = X0[pass] + (size_t)x * DX[pass], Y = Y0[pass] + (size_t)y * DY[pass];
        pixel_rgba(p, &rows[stride * y], x, &rgba[4 * (Y * w + X)]);
      }
  }
  return true;
}

bool read_file(const char *path, std::vector<unsigned char> &out) {
  FILE *f = std::fopen(path, "rb");
  if (!f) return false;
  unsigned char buf[1 << 16];
  size_t n;
  while ((n = std::fread(buf, 1, sizeof(buf), f)) > 0) out.insert(out.end(), buf, buf + n);
  std::fclose(f);
  return true;
}
}  // namespace

namespace dr {
// Decodes a PNG file to RGBA8 (row-major, row 0 at the top).
bool png_decode_file(const std::string &path, std::vector<unsigned char> &rgba, uint32_t &w, uint32_t &h, std::string &err) {
  std::vector<unsigned char> file;
  if (!read_file(path.c_str(), file)) { err = "cannot open " + path; return false; }
  return decode(file, rgba, w, h, err);
}
}  // namespace dr

extern "C" int dr_png_info(const char *path, int *width, int *height) {
  if (!path) return dr::fail(DR_ERR_ARG, "dr_png_info: bad arguments");
  try {
    std::vector<unsigned char> rgba;
    uint32_t w = 0, h = 0;
    std::string err;
    if (!dr::png_decode_file(path, rgba, w, h, err)) return dr::fail(DR_ERR_ARG, "dr_png_info: " + err);
    if (width) *width = (int)w;
    if (height) *height = (int)h;
    return DR_OK;
  } catch (const std::bad_alloc &) {
    return dr::fail(DR_ERR_ALLOC, "dr_png_info: out of host memory");
  }
}

extern "C" int dr_png_read(const char *path, uint8_t *rgba, int width, int height) {
  if (!path || !rgba) return dr::fail(DR_ERR_ARG, "dr_png_read: bad arguments");
  try {
    std::vector<unsigned char> px;
    uint32_t w = 0, h = 0;
    std::string err;
    if (!dr::png_decode_file(path, px, w, h, err)) return dr::fail(DR_ERR_ARG, "dr_png_read: " + err);
    if ((int)w != width || (int)h != height) return dr::fail(DR_ERR_ARG, "dr_png_read: size differs from dr_png_info's");
    std::memcpy(rgba, px.data(), px.size());
    return DR_OK;
  } catch (const std::bad_alloc &) {
    return dr::fail(DR_ERR_ALLOC, "dr_png_read: out of host memory");
  }
}
