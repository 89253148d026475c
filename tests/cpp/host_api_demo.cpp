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

// The failure path at the reference's call site: renderBlocksGPU returns false (the
// reference then renders the blocks with renderBlocksCPU, DumbRenderer.cu:109-111); here the
// library reports why, leaves the frame as it was and keeps working for the next call.
static int failure_demo() {
  try {
    check(dr_device_open(0), "dr_device_open");
    Octree<> tree(16, 12);
    std::vector<Renderable<BakedTriFace>> faces(2);
    faces[0].object.vert = Triangle{{-4, -1, -4}, {-4, -1, 4}, {4, -1, 4}};
    faces[1].object.vert = Triangle{{-4, -1, -4}, {4, -1, 4}, {4, -1, -4}};
    faces[0].materialId = 0;
    faces[1].materialId = 1;                     // the renderer below has one material only
    tree.push(faces);
    tree.build();
    dr_light l{};
    l.kind = DR_AREA_LIGHT;
    const float corner[3] = {-0.5f, 3.0f, -0.5f}, eu[3] = {1, 0, 0}, ev[3] = {0, 0, 1}, em[3] = {10, 10, 10};
    for (int k = 0; k < 3; k++) { l.corner[k] = corner[k]; l.edge_u[k] = eu[k]; l.edge_v[k] = ev[k]; l.emission[k] = em[k]; }
    const Camera c{{0, 2, -6}, {1, 0, 0}, {0, 0.9486833f, 0.3162278f}, {0, -0.3162278f, 0.9486833f}, 0.8660254f};
    std::vector<dr_material> one = {Plugins::lambert({0.5f, 0.5f, 0.5f})};
    DumbRenderer bad(tree, one, {l}, c, 2, 1, 1, false);
    BlockBuffer frame(24, 20);
    const size_t bytes = sizeof(float) * 4 * (size_t)frame.getBlockSize() * frame.getBlockCount();
    hip_check(hipMemset(frame.deviceData(), 0x7f, bytes), "hipMemset");
    if (bad.renderBlocksGPU(frame, 0, frame.getBlockCount(), 1)) throw Error("renderBlocksGPU accepted a short material table");
    printf("renderBlocksGPU failed: %s\n", dr_last_error());
    std::vector<unsigned char> raw(bytes);
    hip_check(hipMemcpy(raw.data(), frame.deviceData(), bytes, hipMemcpyDeviceToHost), "hipMemcpy");
    for (unsigned char b : raw)
      if (b != 0x7f) throw Error("the failed call wrote the frame");
    printf("frame untouched\n");
    // the caller's fallback: here, a renderer with a complete material table
    std::vector<dr_material> two = {one[0], Plugins::lambert({0.2f, 0.7f, 0.3f})};
    DumbRenderer good(tree, two, {l}, c, 2, 1, 1, false);
    if (!good.renderBlocksGPU(frame, 0, frame.getBlockCount(), 1)) throw Error(dr_last_error());
    hip_check(hipDeviceSynchronize(), "sync");
    const std::vector<float> img = frame.toHostImage();
    float mx = 0.0f;
    for (size_t i = 0; i < img.size(); i += 4) mx = img[i] > mx ? img[i] : mx;
    if (!(mx > 0.0f)) throw Error("retry rendered nothing");
    printf("retry rendered (max red %g)\n", mx);
  } catch (const std::exception &e) {
    fprintf(stderr, "host_api_demo --failure: %s\n", e.what());
    return 1;
  }
  return 0;
}

int main(int argc, char **argv) {
  if (argc == 2 && strcmp(argv[1], "--failure") == 0) return failure_demo();
  if (argc != 3) { fprintf(stderr, "usage: %s scene.bin out.bin | --failure\n", argv[0]); return 2; }
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
