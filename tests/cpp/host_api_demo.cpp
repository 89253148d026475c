// Drives the C++ host mirror (include/dumbray.hpp) end to end: reads a scene and
// render settings written by tests/test_cpp_host_api.py, builds the Octree,
// renders one iteration with DumbRenderer::renderBlocksGPU and writes the
// row-major RGBA image and the node count back for comparison with the Python
// path over the same C ABI.
#include <cstdio>
#include <cstring>
#include <vector>

#include "dumbray.hpp"

using namespace dumbray;

template <typename T>
static bool rd(FILE *f, T *p, size_t n) { return fread(p, sizeof(T), n, f) == n; }

int main(int argc, char **argv) {
  if (argc != 3) { fprintf(stderr, "usage: %s scene.bin out.bin\n", argv[0]); return 2; }
  FILE *f = fopen(argv[1], "rb");
  if (!f) { perror("scene"); return 2; }
  int32_t hdr[10];
  if (!rd(f, hdr, 10)) return 3;
  const int ntris = hdr[0], nmat = hdr[1], nlights = hdr[2], W = hdr[3], H = hdr[4], fx = hdr[5], fy = hdr[6],
            depth = hdr[7], cull = hdr[8], thresh_depth = hdr[9];
  std::vector<float> tris(9 * (size_t)ntris);
  std::vector<int32_t> mid(ntris);
  std::vector<dr_material> mats(nmat);
  std::vector<dr_light> lights(nlights);
  float cam[13];
  uint64_t seed;
  if (!rd(f, tris.data(), tris.size()) || !rd(f, mid.data(), mid.size()) || !rd(f, mats.data(), mats.size()) ||
      !rd(f, lights.data(), lights.size()) || !rd(f, cam, 13) || !rd(f, &seed, 1))
    return 3;
  fclose(f);
  try {
    check(dr_device_open(0), "dr_device_open");
    Octree<> tree(thresh_depth >> 8, thresh_depth & 0xff);
    std::vector<Renderable<BakedTriFace>> faces(ntris);
    for (int i = 0; i < ntris; i++) {
      const float *t = &tris[9 * (size_t)i];
      faces[i].object.vert = Triangle{{t[0], t[1], t[2]}, {t[3], t[4], t[5]}, {t[6], t[7], t[8]}};
      faces[i].materialId = mid[i];
    }
    tree.push(faces);
    tree.build();
    Camera c{{cam[0], cam[1], cam[2]}, {cam[3], cam[4], cam[5]}, {cam[6], cam[7], cam[8]}, {cam[9], cam[10], cam[11]},
             cam[12]};
    DumbRenderer r(tree, mats, lights, c, depth, fx, fy, cull != 0);
    r.setSeed(seed);
    BlockBuffer frame(W, H);
    if (!r.renderBlocksGPU(frame, 0, frame.getBlockCount(), 1)) throw Error("renderBlocksGPU");
    hip_check(hipDeviceSynchronize(), "sync");
    const std::vector<float> img = frame.toHostImage();
    // the same request through RenderBatch (one persistent graph launch) gives the same frame
    BlockBuffer bframe(W, H);
    RenderBatch batch({{&r, &bframe, 1, 0}});
    batch.run();
    hip_check(hipDeviceSynchronize(), "sync");
    const std::vector<float> bimg = bframe.toHostImage();
    if (bimg.size() != img.size() || std::memcmp(bimg.data(), img.data(), img.size() * sizeof(float)) != 0)
      throw Error("RenderBatch frame differs from renderBlocksGPU");
    FILE *o = fopen(argv[2], "wb");
    if (!o) { perror("out"); return 2; }
    const int32_t nodes = tree.getNodeCount();
    fwrite(&nodes, sizeof(nodes), 1, o);
    fwrite(img.data(), sizeof(float), img.size(), o);
    fclose(o);
    printf("host_api_demo: %d tris, %d nodes, %dx%d rendered, batch equal\n", ntris, nodes, W, H);
  } catch (const std::exception &e) {
    fprintf(stderr, "host_api_demo: %s\n", e.what());
    return 1;
  }
  return 0;
}
