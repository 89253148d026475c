// PNG output of RGBA8 images: the file Images::saveBufferPNG writes through
// lodepng::encode (Images.cu:22-36): 8-bit RGBA, non-interlaced, here with filter 0 on
// every scanline and zlib deflate.  Host only.
#include <cstdint>
#include <cstdio>
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
