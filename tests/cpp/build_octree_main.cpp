// Builds the octree of a raw float32 triangle file (ntris x 9) with the library's host
// builder and prints the node count; tests/test_build_sanitizers.py compiles it with
// ThreadSanitizer and with AddressSanitizer + UndefinedBehaviorSanitizer.
//   build_octree_main FILE THREADS SPLIT DEPTH
#include <cstdio>
#include <cstdlib>
#include <vector>
#include "dr_octree_build.h"

int main(int argc, char **argv) {
  if (argc != 5) return 2;
  FILE *f = std::fopen(argv[1], "rb");
  if (!f) return 2;
  std::fseek(f, 0, SEEK_END);
  const long bytes = std::ftell(f);
  std::fseek(f, 0, SEEK_SET);
  std::vector<float> t((size_t)bytes / 4);
  const size_t got = std::fread(t.data(), 1, (size_t)bytes, f);
  std::fclose(f);
  if (got != (size_t)bytes) return 2;
  dr::OctreeHost h;
  const int rc = dr::build_octree(t.data(), (int)(t.size() / 9), std::atoi(argv[3]), std::atoi(argv[4]),
                                  std::atoi(argv[2]), h);
  std::printf("rc %d nodes %zu\n", rc, h.bounds.size());
  return rc;
}
