// The reference's scene-file workflow through the C++ host mirror (include/dumbray.hpp):
// DumbRenderContext::buildFromFile, a DumbRenderer with the file's settings, a
// BufferedRenderProcess running a fixed number of iterations into a
// BlockBasedFrameBuffer<8,16>, then Images::saveBufferPNG.
// usage: scene_file_demo scene.dumb width height iterations out.png
#include <cstdio>
#include <cstdlib>
#include <iostream>
#include <memory>

#include "dumbray.hpp"

namespace {
struct Done {
  std::mutex m;
  std::condition_variable cv;
  bool done = false;
  int iterations = 0;
};
void on_iteration(void *p) {
  Done *d = (Done *)p;
  std::lock_guard<std::mutex> g(d->m);
  d->iterations++;
}
void on_done(void *p) {
  Done *d = (Done *)p;
  std::lock_guard<std::mutex> g(d->m);
  d->done = true;
  d->cv.notify_all();
}
}  // namespace

int main(int argc, char **argv) {
  if (argc != 6) {
    std::fprintf(stderr, "usage: %s scene.dumb width height iterations out.png\n", argv[0]);
    return 2;
  }
  const int W = std::atoi(argv[2]), H = std::atoi(argv[3]), N = std::atoi(argv[4]);
  try {
    dumbray::check(dr_device_open(0), "dr_device_open");
    dumbray::DumbRenderContext ctx;
    if (!ctx.buildFromFile(argv[1], &std::cerr)) return 3;
    std::unique_ptr<dumbray::DumbRenderer> renderer(ctx.newRenderer());
    dumbray::BlockBuffer frame(8, 8);
    Done done;
    dumbray::BufferedRenderProcess process;
    process.setBuffer(&frame);
    process.setRenderer(renderer.get());
    process.setTargetResolution(W, H);
    process.setTargetIterations(N);
    process.setIterationCompletionCallback(on_iteration, &done);
    process.setRenderCompletionCallback(on_done, &done);
    process.start();
    {
      std::unique_lock<std::mutex> lock(done.m);
      if (!done.cv.wait_for(lock, std::chrono::seconds(60), [&] { return done.done; })) {
        std::fprintf(stderr, "render did not complete\n");
        return 4;
      }
    }
    process.end();
    if (renderer->iteration() != N || done.iterations != N) {
      std::fprintf(stderr, "iterations %d / callbacks %d, expected %d\n", renderer->iteration(), done.iterations, N);
      return 5;
    }
    if (dumbray::Images::saveBufferPNG(frame, argv[5]) != dumbray::Images::IMAGES_NO_ERROR) return 6;
    std::printf("triangles %zu nodes %d iterations %d\n", ctx.geometry().getData().size(),
                ctx.geometry().getNodeCount(), renderer->iteration());
  } catch (const std::exception &e) {
    std::fprintf(stderr, "%s\n", e.what());
    return 1;
  }
  return 0;
}
