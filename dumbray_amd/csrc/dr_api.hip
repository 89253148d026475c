// C ABI: errors, device selection, octree build/export, scene upload.
#include <cstring>
#include <new>
#include "dr_common.h"
#include "dr_plugins.h"

namespace dr {
static thread_local std::string g_err = "";
void set_error(const std::string &msg) { g_err = msg; }
int fail(int code, const std::string &msg) { g_err = msg; return code; }
int dev_upload(DevBuf &b, const void *host, size_t bytes) {
  b.bytes = bytes;
  if (bytes == 0) { b.p = nullptr; return DR_OK; }
  hipError_t e = hipMalloc(&b.p, bytes);
  if (e != hipSuccess) { b.p = nullptr; return fail(DR_ERR_ALLOC, std::string("hipMalloc: ") + hipGetErrorString(e)); }
  e = hipMemcpy(b.p, host, bytes, hipMemcpyHostToDevice);
  if (e != hipSuccess) return fail(DR_ERR_HIP, std::string("hipMemcpy: ") + hipGetErrorString(e));
  return DR_OK;
}
void dev_free(DevBuf &b) {
  if (b.p) (void)hipFree(b.p);
  b.p = nullptr;
  b.bytes = 0;
}
}  // namespace dr

extern "C" {

const char *dr_last_error(void) { return dr::g_err.c_str(); }
const char *dr_version(void) { return "dumbray_amd 0.1 gfx950"; }

int dr_device_open(int device) {
  int n = 0;
  DR_HIP_CHECK(hipGetDeviceCount(&n));
  if (device < 0 || device >= n) return dr::fail(DR_ERR_ARG, "dr_device_open: no such device");
  DR_HIP_CHECK(hipSetDevice(device));
  return DR_OK;
}

int dr_octree_build(const float *tris, int ntris, int split_threshold, int max_depth, int nthreads,
                    dr_octree *out) {
  if (!out || ntris < 0 || (ntris > 0 && !tris)) return dr::fail(DR_ERR_ARG, "dr_octree_build: bad arguments");
  *out = nullptr;
  dr_octree t = new (std::nothrow) dr_octree_s();
  if (!t) return dr::fail(DR_ERR_ALLOC, "dr_octree_build: out of host memory");
  try {
    const int rc = dr::build_octree(tris, ntris, split_threshold, max_depth, nthreads, t->host);
    if (rc != 0) {
      delete t;
      return dr::fail(rc == -1 ? DR_ERR_ARG : DR_ERR_BUILD,
                      rc == -1 ? "dr_octree_build: split_threshold >= 1 and 0 <= max_depth <= 24 required"
                               : "dr_octree_build: node pool exhausted");
    }
  } catch (const std::bad_alloc &) {
    delete t;
    return dr::fail(DR_ERR_ALLOC, "dr_octree_build: out of host memory");
  }
  *out = t;
  return DR_OK;
}

int dr_octree_info(dr_octree t, int *node_count, int *list_total, int *ngroups, int *nrec) {
  if (!t) return dr::fail(DR_ERR_ARG, "dr_octree_info: null octree");
  const dr::OctreeHost &h = t->host;
  if (node_count) *node_count = (int)h.bounds.size();
  if (list_total) *list_total = (int)h.list.size();
  int nleafrec = 0;
  for (size_t k = 0; k < h.list.size(); k++) {
    const dr::Tri &tr = h.tris[h.list[k]];
    if (!(dr::eq(tr.a, tr.b) || dr::eq(tr.b, tr.c) || dr::eq(tr.c, tr.a))) nleafrec++;
  }
  if (ngroups) *ngroups = (h.children.empty() || h.children[0] < 0) ? 0 : (int)(h.bounds.size() - 1) / 8;
  if (nrec) *nrec = nleafrec;
  return DR_OK;
}

int dr_octree_export(dr_octree t, float *bounds, int32_t *children, int32_t *list_off, int32_t *list) {
  if (!t) return dr::fail(DR_ERR_ARG, "dr_octree_export: null octree");
  const dr::OctreeHost &h = t->host;
  const size_t n = h.bounds.size();
  for (size_t i = 0; i < n; i++) {
    const dr::Box &b = h.bounds[i];
    if (bounds) {
      bounds[6 * i + 0] = b.s.x; bounds[6 * i + 1] = b.s.y; bounds[6 * i + 2] = b.s.z;
      bounds[6 * i + 3] = b.e.x; bounds[6 * i + 4] = b.e.y; bounds[6 * i + 5] = b.e.z;
    }
    if (children) children[i] = h.children[i];
  }
  if (list_off) memcpy(list_off, h.list_off.data(), sizeof(int32_t) * h.list_off.size());
  if (list && !h.list.empty()) memcpy(list, h.list.data(), sizeof(int32_t) * h.list.size());
  return DR_OK;
}

void dr_octree_destroy(dr_octree t) { delete t; }

int dr_scene_create(dr_octree t, dr_scene *out) {
  if (!t || !out) return dr::fail(DR_ERR_ARG, "dr_scene_create: bad arguments");
  *out = nullptr;
  dr::OctreeFlat flat;
  if (dr::flatten_octree(t->host, flat) != 0) return dr::fail(DR_ERR_ARG, "dr_scene_create: octree cannot be flattened");
  dr_scene s = new (std::nothrow) dr_scene_s();
  if (!s) return dr::fail(DR_ERR_ALLOC, "dr_scene_create: out of host memory");
  (void)hipGetDevice(&s->device);
  int rc = dr::dev_upload(s->groups, flat.groups.data(), flat.groups.size() * sizeof(float));
  if (rc == DR_OK) rc = dr::dev_upload(s->recs, flat.recs.data(), flat.recs.size() * sizeof(float));
  if (rc == DR_OK) rc = dr::dev_upload(s->urecs, flat.urecs.data(), flat.urecs.size() * sizeof(float));
  if (rc == DR_OK) rc = dr::dev_upload(s->lref, flat.lref.data(), flat.lref.size() * sizeof(int32_t));
  if (rc == DR_OK) {
    const std::vector<uint32_t> zeros(2 * dr::kWorkSlots, 0u);
    rc = dr::dev_upload(s->work, zeros.data(), zeros.size() * sizeof(uint32_t));
  }
  if (rc == DR_OK && hipDeviceGetAttribute(&s->num_cu, hipDeviceAttributeMultiprocessorCount, s->device) != hipSuccess)
    rc = dr::fail(DR_ERR_HIP, "dr_scene_create: cannot query the compute unit count");
  if (rc != DR_OK) { dr_scene_destroy(s); return rc; }
  dr::DevTree &T = s->tree;
  T.groups = (const float4 *)s->groups.p;
  T.recs = (const float4 *)s->recs.p;
  T.urecs = (const float4 *)s->urecs.p;
  T.lref = (const int *)s->lref.p;
  T.rsx = flat.root.s.x; T.rsy = flat.root.s.y; T.rsz = flat.root.s.z;
  T.rex = flat.root.e.x; T.rey = flat.root.e.y; T.rez = flat.root.e.z;
  T.root_is_leaf = flat.root_is_leaf;
  T.root_start = flat.root_start;
  T.root_count = flat.root_count;
  T.margin = flat.prune_margin;
  s->max_depth = flat.max_depth;
  s->ntris = (int)t->host.tris.size();
  s->shade_host.resize(s->ntris);
  s->faces_host.assign((size_t)s->ntris * dr::kFaceWords, 0.0f);   // texture coordinates start at 0
  for (int i = 0; i < s->ntris; i++) {
    const dr::Tri &tr = t->host.tris[i];
    const dr::V3 n = dr::cross(tr.b - tr.a, tr.c - tr.a);
    s->shade_host[i] = make_float4(n.x, n.y, n.z, 0.0f);
    float *F = &s->faces_host[(size_t)i * dr::kFaceWords];
    const float v[9] = {tr.a.x, tr.a.y, tr.a.z, tr.b.x, tr.b.y, tr.b.z, tr.c.x, tr.c.y, tr.c.z};
    memcpy(F, v, sizeof(v));
  }
  rc = dr_scene_set_normals(s, nullptr, s->ntris);
  if (rc != DR_OK) { dr_scene_destroy(s); return rc; }
  *out = s;
  return DR_OK;
}

int dr_scene_set_normals(dr_scene s, const float *normals, int ntris) {
  if (!s || ntris != s->ntris) return dr::fail(DR_ERR_ARG, "dr_scene_set_normals: one normal triple per triangle required");
  for (int i = 0; i < ntris; i++) {
    float *F = &s->faces_host[(size_t)i * dr::kFaceWords];
    if (normals) {
      memcpy(F + 9, normals + 9 * (size_t)i, 9 * sizeof(float));
    } else {
      // flat: the unit Triangle::normal (c - a) x (b - a) at every vertex
      const dr::V3 a = dr::v3(F[0], F[1], F[2]), b = dr::v3(F[3], F[4], F[5]), c = dr::v3(F[6], F[7], F[8]);
      const dr::V3 n = dr::normalized(dr::cross(c - a, b - a));
      for (int k = 0; k < 3; k++) { F[9 + 3 * k] = n.x; F[10 + 3 * k] = n.y; F[11 + 3 * k] = n.z; }
    }
  }
  dr::dev_free(s->faces);
  return dr::dev_upload(s->faces, s->faces_host.data(), s->faces_host.size() * sizeof(float));
}

int dr_scene_set_texcoords(dr_scene s, const float *uv, int ntris) {
  if (!s || ntris != s->ntris) return dr::fail(DR_ERR_ARG, "dr_scene_set_texcoords: one coordinate triple per triangle required");
  for (int i = 0; i < ntris; i++) {
    float *F = &s->faces_host[(size_t)i * dr::kFaceWords];
    for (int k = 0; k < 6; k++) F[18 + k] = uv ? uv[6 * (size_t)i + k] : 0.0f;
  }
  dr::dev_free(s->faces);
  return dr::dev_upload(s->faces, s->faces_host.data(), s->faces_host.size() * sizeof(float));
}

int dr_scene_set_materials(dr_scene s, const int32_t *mat_id, int ntris) {
  if (!s || ntris != s->ntris || (ntris > 0 && !mat_id)) return dr::fail(DR_ERR_ARG, "dr_scene_set_materials: one id per triangle required");
  int mx = -1;
  for (int i = 0; i < ntris; i++) {
    if (mat_id[i] < 0) return dr::fail(DR_ERR_ARG, "dr_scene_set_materials: negative material id");
    memcpy(&s->shade_host[i].w, &mat_id[i], 4);
    mx = mat_id[i] > mx ? mat_id[i] : mx;
  }
  s->max_mat = mx;
  dr::dev_free(s->shade);
  const int rc = dr::dev_upload(s->shade, s->shade_host.data(), sizeof(float4) * (size_t)ntris);
  if (rc != DR_OK) return rc;
  s->has_materials = true;
  return DR_OK;
}

void dr_scene_destroy(dr_scene s) {
  if (!s) return;
  dr::dev_free(s->groups);
  dr::dev_free(s->recs);
  dr::dev_free(s->urecs);
  dr::dev_free(s->lref);
  dr::dev_free(s->shade);
  dr::dev_free(s->work);
  dr::dev_free(s->faces);
  delete s;
}

int64_t dr_scene_device_bytes(dr_scene s) { return s ? s->bytes() : 0; }

}  // extern "C"
