// Host build of the render kernels' camera and framebuffer helpers (dr_shade.h: frame_k,
// camera_ray, block_pixel, udiv), checked against the oracle and integer division.
// Built and run by tests/test_shade_host.py (host code only, no GPU).
//
//   shade_probe udiv     n / d through udiv against the / operator: every n < 2^17 for
//                        d <= 4096, n around the multiples of d up to 2^31, random 32-bit pairs
//   shade_probe camera   camera_ray(P, frame_k(P), ...) against the oracle's ora_camera_ray, bit for
//                        bit, for every lens and boxing mode on several image and subpixel sizes
//   shade_probe blocks   block_pixel against the framebuffer layout of 8 x 16 blocks
//
// Prints "ok" and exits 0, or prints the first mismatch and exits 1.
#include <cstdio>
#include <cstring>
#include <random>
#include "../../dumbray_amd/csrc/dr_shade.h"
#include "../../oracle/dr_oracle_render.h"

static uint32_t bits(float f) { uint32_t u; std::memcpy(&u, &f, 4); return u; }

static int check_udiv() {
  for (uint32_t d = 1; d <= 4096; d++) {
    const dr::UDiv D = dr::make_udiv(d);
    for (uint32_t n = 0; n < (1u << 17); n++)
      if (dr::udiv(n, D) != n / d) { std::printf("udiv(%u, %u) = %u\n", n, d, dr::udiv(n, D)); return 1; }
    for (uint64_t k = 1; k * d < (1ull << 31); k += 1 + k / 64) {
      for (int e = -1; e <= 1; e++) {
        const uint32_t n = (uint32_t)((int64_t)(k * d) + e);
        if (dr::udiv(n, D) != n / d) { std::printf("udiv(%u, %u) = %u\n", n, d, dr::udiv(n, D)); return 1; }
      }
    }
  }
  std::mt19937_64 rng(7);
  for (int t = 0; t < 20000000; t++) {
    const uint32_t d = 1 + (uint32_t)(rng() % 0xffffffffull), n = (uint32_t)rng();
    if (dr::udiv(n, dr::make_udiv(d)) != n / d) { std::printf("udiv(%u, %u)\n", n, d); return 1; }
  }
  std::printf("ok\n");
  return 0;
}

static int check_camera() {
  std::mt19937 rng(11);
  std::uniform_real_distribution<float> U01(0.0f, 1.0f), Um(-1.0f, 1.0f);
  const int sizes[][4] = {{1920, 1080, 4, 4}, {256, 256, 2, 2}, {33, 20, 3, 5}, {1, 1, 1, 1}, {640, 961, 1, 7}};
  long checked = 0;
  for (int lens = 0; lens <= 2; lens++)
    for (int boxing = 0; boxing <= 3; boxing++)
      for (const auto &sz : sizes) {
        dr_render_params P;
        std::memset(&P, 0, sizeof(P));
        ora_render_cfg C;
        std::memset(&C, 0, sizeof(C));
        P.width = C.width = sz[0];
        P.height = C.height = sz[1];
        P.fsaa_x = C.fsaa_x = sz[2];
        P.fsaa_y = C.fsaa_y = sz[3];
        P.iteration = C.iteration = 1;
        P.lens = C.lens = lens;
        P.boxing = C.boxing = boxing;
        for (int k = 0; k < 13; k++) P.cam[k] = C.cam[k] = Um(rng) * 3.0f;
        P.cam[12] = C.cam[12] = 0.5f + U01(rng) * 2.0f;
        const float lp[8] = {70.0f, 40.0f, 25.0f, 0.6f, 1.0f, 0.5f, 0.25f, 0.0f};
        for (int k = 0; k < 8; k++) P.lens_p[k] = C.lens_p[k] = lp[k];
        for (int k = 0; k < 9; k++) P.cam_rows[k] = C.cam_rows[k] = Um(rng);
        const dr::FrameK K = dr::frame_k(P);
        for (int t = 0; t < 4000; t++) {
          const int px = (int)(rng() % (uint32_t)sz[0]), py = (int)(rng() % (uint32_t)sz[1]);
          const int s = (int)(rng() % (uint32_t)(sz[2] * sz[3]));
          // the kernels' subpixel split of sample s, through udiv as in start_path
          const int i = (int)dr::udiv((uint32_t)s, K.fy), j = s - i * P.fsaa_y;
          if (i != s / sz[3] || j != s % sz[3]) { std::printf("subpixel split of %d\n", s); return 1; }
          const float u[4] = {U01(rng), U01(rng), U01(rng), U01(rng)};
          dr::V3 o, d;
          dr::camera_ray(P, K, px, py, i, j, u[0], u[1], u[2], u[3], o, d);
          float oo[3], od[3];
          ora_camera_ray(&C, px, py, i, j, u, oo, od);
          const float got[6] = {o.x, o.y, o.z, d.x, d.y, d.z}, want[6] = {oo[0], oo[1], oo[2], od[0], od[1], od[2]};
          for (int k = 0; k < 6; k++)
            if (bits(got[k]) != bits(want[k])) {
              std::printf("lens %d boxing %d size %dx%d/%dx%d pixel (%d,%d) sub (%d,%d): component %d %.9g vs %.9g\n",
                          lens, boxing, sz[0], sz[1], sz[2], sz[3], px, py, i, j, k, got[k], want[k]);
              return 1;
            }
          checked++;
        }
      }
  std::printf("ok %ld rays\n", checked);
  return 0;
}

static int check_blocks() {
  const int sizes[][2] = {{1920, 1080}, {33, 20}, {8, 16}, {9, 17}, {1, 1}, {1024, 1}};
  for (const auto &sz : sizes) {
    dr_render_params P;
    std::memset(&P, 0, sizeof(P));
    P.width = sz[0];
    P.height = sz[1];
    P.fsaa_x = P.fsaa_y = 1;
    const dr::FrameK K = dr::frame_k(P);
    const int nbw = (sz[0] + 7) / 8, nbh = (sz[1] + 15) / 16;
    // every pixel of the image is reached by exactly one (block, local pixel), row-major in the block
    std::vector<int> seen((size_t)sz[0] * sz[1], 0);
    for (int b = 0; b < nbw * nbh; b++)
      for (int lp = 0; lp < 128; lp++) {
        int x, y;
        const bool in = dr::block_pixel(K, P.width, P.height, b, lp, x, y);
        if (x != 8 * (b % nbw) + lp % 8 || y != 16 * (b / nbw) + lp / 8 || in != (x < sz[0] && y < sz[1])) {
          std::printf("block %d pixel %d of %dx%d -> (%d, %d)\n", b, lp, sz[0], sz[1], x, y);
          return 1;
        }
        if (in) seen[(size_t)y * sz[0] + x]++;
      }
    for (int v : seen)
      if (v != 1) { std::printf("pixel coverage of %dx%d\n", sz[0], sz[1]); return 1; }
  }
  std::printf("ok\n");
  return 0;
}

int main(int argc, char **argv) {
  if (argc < 2) return 2;
  if (!std::strcmp(argv[1], "udiv")) return check_udiv();
  if (!std::strcmp(argv[1], "camera")) return check_camera();
  if (!std::strcmp(argv[1], "blocks")) return check_blocks();
  return 2;
}
