// Mutation fuzzing of the library's host-side parsers: the scene-file loader (Dson text,
// OBJ meshes, the PNG textures they name) and the PNG reader.  Built with AddressSanitizer
// and UndefinedBehaviorSanitizer by tests/test_fuzz_host.py (host code only); any memory
// error, leak or undefined behaviour ends the run with a sanitizer report.
//
//   fuzz_host scene N DIR SEED...   mutated scene texts parsed with base directory DIR
//   fuzz_host obj N DIR SCENE OBJ   SCENE parsed with DIR/mesh.obj replaced by a mutated OBJ
//   fuzz_host png N DIR SEED...     mutated PNG files (chunk data with valid CRCs, or the
//                                   inflated scanlines re-deflated) read through dr_png_*
//
// Accepted inputs are also read back through every dr_scene_file_* accessor.
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <random>
#include <sstream>
#include <string>
#include <vector>
#include <zlib.h>
#include "../../include/dumbray.h"
namespace dr {   // the library's error hook (dr_api.hip), without the HIP parts
std::string g_err;
int fail(int code, const std::string &msg) { g_err = msg; return code; }
}
static std::string slurp(const char *p) { std::ifstream f(p, std::ios::binary); std::stringstream s; s << f.rdbuf(); return s.str(); }
static void spit(const std::string &p, const std::string &d) { std::ofstream f(p, std::ios::binary); f << d; }
static std::mt19937_64 rng(12345);   // fixed: runs are reproducible
static std::string mutate(std::string s) {
  int n = 1 + (int)(rng() % 3);
  for (int k = 0; k < n; k++) {
    if (s.empty()) { s.push_back((char)rng()); continue; }
    size_t i = rng() % s.size();
    switch (rng() % 7) {
      case 0: s[i] = (char)rng(); break;
      case 1: s.erase(i, 1 + rng() % 16); break;
      case 2: s.insert(i, 1, "{}[]\":,-.0123456789eE \n/"[rng() % 25]); break;
      case 3: { size_t j = rng() % s.size(); size_t l = 1 + rng() % 64; s.insert(i, s.substr(j, l)); } break;
      case 4: s[i] ^= (char)(1 << (rng() % 8)); break;
      case 5: s.insert(i, std::to_string((long long)rng() % 100000 - 50000)); break;
      case 6: s.resize(i); break;
    }
  }
  return s;
}

// PNG-aware mutation: chunk data mutated with CRCs fixed, or the inflated IDAT stream
// mutated and deflated again.
static uint32_t be32(const std::string &s, size_t i) {
  return ((uint32_t)(uint8_t)s[i] << 24) | ((uint32_t)(uint8_t)s[i + 1] << 16) | ((uint32_t)(uint8_t)s[i + 2] << 8) | (uint8_t)s[i + 3];
}
static void put32(std::string &o, uint32_t v) { for (int k = 3; k >= 0; k--) o.push_back((char)(v >> (8 * k))); }
static std::string chunk(const std::string &type, const std::string &data) {
  std::string o;
  put32(o, (uint32_t)data.size());
  std::string td = type + data;
  o += td;
  put32(o, (uint32_t)crc32(0, (const Bytef *)td.data(), (uInt)td.size()));
  return o;
}
static std::string png_mutate(const std::string &s) {
  std::vector<std::pair<std::string, std::string>> ch;
  size_t i = 8;
  while (i + 12 <= s.size()) {
    uint32_t n = be32(s, i);
    if (i + 12 + n > s.size()) break;
    ch.push_back({s.substr(i + 4, 4), s.substr(i + 8, n)});
    i += 12 + n;
  }
  if (ch.empty()) return mutate(s);
  std::string idat;
  for (auto &c : ch) if (c.first == "IDAT") idat += c.second;
  const int what = (int)(rng() % 3);
  if (what == 0 && !idat.empty()) {
    std::vector<unsigned char> raw(1 << 22);
    uLongf rl = raw.size();
    if (uncompress(raw.data(), &rl, (const Bytef *)idat.data(), idat.size()) == Z_OK) {
      std::string r((const char *)raw.data(), rl);
      r = mutate(r);
      std::vector<unsigned char> z(compressBound(r.size()) + 16);
      uLongf zl = z.size();
      compress(z.data(), &zl, (const Bytef *)r.data(), r.size());
      std::string out = s.substr(0, 8);
      bool done = false;
      for (auto &c : ch) {
        if (c.first == "IDAT") { if (!done) out += chunk("IDAT", std::string((const char *)z.data(), zl)); done = true; }
        else out += chunk(c.first, c.second);
      }
      return out;
    }
  }
  size_t k = rng() % ch.size();
  if (what == 2) ch[k].first[rng() % 4] = "IHDRPLTEtRNSIDATIENDabcz"[rng() % 24];   // another chunk type
  else ch[k].second = mutate(ch[k].second);
  std::string out = s.substr(0, 8);
  for (auto &c : ch) out += chunk(c.first, c.second);
  return out;
}
static void exercise(dr_scene_file f) {
  dr_scene_file_settings st;
  int n, nm, nl;
  if (dr_scene_file_info(f, &st, &n, &nm, &nl) != DR_OK) return;
  std::vector<float> t((size_t)n * 9), nr((size_t)n * 9), uv((size_t)n * 6);
  std::vector<int32_t> mid(n);
  dr_scene_file_geometry(f, t.data(), nr.data(), mid.data());
  if (n) dr_scene_file_texcoords(f, uv.data());
  std::vector<dr_material> m(nm);
  std::vector<dr_light> l(nl);
  dr_scene_file_tables(f, m.data(), l.data());
  int nt, nmp; int64_t ntx;
  if (dr_scene_file_texture_info(f, &nt, &ntx, &nmp) != DR_OK) return;
  std::vector<dr_texture> tx(nt);
  std::vector<float> texels((size_t)ntx * 4);
  std::vector<dr_material_maps> maps(nmp);
  dr_scene_file_textures(f, tx.data(), texels.data(), maps.data());
}
int main(int argc, char **argv) {
  const std::string mode = argv[1];
  const int iters = atoi(argv[2]);
  const std::string dir = argv[3];     // scratch folder
  int ok = 0;
  if (mode == "scene") {
    std::vector<std::string> seeds;
    for (int i = 4; i < argc; i++) seeds.push_back(slurp(argv[i]));
    for (int it = 0; it < iters; it++) {
      std::string s = mutate(seeds[rng() % seeds.size()]);
      dr_scene_file f = nullptr;
      if (dr_scene_file_parse(s.c_str(), dir.c_str(), &f) == DR_OK) { ok++; exercise(f); dr_scene_file_destroy(f); }
    }
  } else if (mode == "obj") {     // argv[4] scene text naming mesh.obj in dir, argv[5] the obj seed
    const std::string scene = slurp(argv[4]), obj = slurp(argv[5]);
    for (int it = 0; it < iters; it++) {
      spit(dir + "/mesh.obj", mutate(obj));
      dr_scene_file f = nullptr;
      if (dr_scene_file_parse(scene.c_str(), dir.c_str(), &f) == DR_OK) { ok++; exercise(f); dr_scene_file_destroy(f); }
    }
  } else if (mode == "png") {     // argv[4..] png seeds
    std::vector<std::string> seeds;
    for (int i = 4; i < argc; i++) seeds.push_back(slurp(argv[i]));
    const std::string p = dir + "/f.png";
    for (int it = 0; it < iters; it++) {
      std::string s = seeds[rng() % seeds.size()];
      // keep the signature mostly, mutate the chunks
      std::string m = (rng() % 8 == 0) ? mutate(s) : png_mutate(s);
      spit(p, m);
      int w = 0, h = 0;
      if (dr_png_info(p.c_str(), &w, &h) == DR_OK && (long long)w * h <= (1 << 24)) {
        std::vector<uint8_t> px((size_t)w * h * 4);
        if (dr_png_read(p.c_str(), px.data(), w, h) == DR_OK) ok++;
      }
    }
  }
  printf("%s: %d iterations, %d accepted\n", mode.c_str(), iters, ok);
  return 0;
}
