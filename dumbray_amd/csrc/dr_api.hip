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

// Copies host data over a device buffer of the same size once the device is idle (the
// buffer's address stays, so launches and batches holding it stay valid).
static int dev_overwrite(DevBuf &b, const void *host, size_t bytes) {
  if (bytes != b.bytes) return fail(DR_ERR_ARG, "internal: table size changed");
  if (bytes == 0) return DR_OK;
  DR_HIP_CHECK(hipDeviceSynchronize());
  DR_HIP_CHECK(hipMemcpy(b.p, host, bytes, hipMemcpyHostToDevice));
  return DR_OK;
}

// Per-triangle shading record words kFaceWords: the vertices (from the scene), the vertex
// normals (given, or the flat unit Triangle::normal (c - a) x (b - a), Triangle.impl.h:233-235)
// and the texture coordinates (given, or zeros).
static void fill_faces(const dr_scene_s &s, const float *normals, const float *uv, std::vector<float> &F) {
  F.assign((size_t)s.ntris * kFaceWords, 0.0f);
  for (int i = 0; i < s.ntris; i++) {
    float *f = &F[(size_t)i * kFaceWords];
    memcpy(f, &s.faces_host[(size_t)i * kFaceWords], 9 * sizeof(float));
    if (normals) {
      memcpy(f + 9, normals + 9 * (size_t)i, 9 * sizeof(float));
    } else {
      const V3 a = v3(f[0], f[1], f[2]), b = v3(f[3], f[4], f[5]), c = v3(f[6], f[7], f[8]);
      const V3 n = normalized(cross(c - a, b - a));
      for (int k = 0; k < 3; k++) { f[9 + 3 * k] = n.x; f[10 + 3 * k] = n.y; f[11 + 3 * k] = n.z; }
    }
    if (uv)
      for (int k = 0; k < 6; k++) f[18 + k] = uv[6 * (size_t)i + k];
  }
}

// Material ids over the scene's per-triangle normals: returns the largest id, or -2 for a negative id.
static int fill_shade(const dr_scene_s &s, const int32_t *mat_id, std::vector<float4> &out) {
  out = s.shade_host;
  int mx = -1;
  for (int i = 0; i < s.ntris; i++) {
    if (mat_id[i] < 0) return -2;
    memcpy(&out[i].w, &mat_id[i], 4);
    mx = mat_id[i] > mx ? mat_id[i] : mx;
  }
  return mx;
}

void scene_retain(dr_scene s) { s->refs.fetch_add(1); }
void scene_release(dr_scene s) {
  if (s->refs.fetch_sub(1) != 1) return;
  dev_free(s->groups);
  dev_free(s->recs);
  dev_free(s->urecs);
  dev_free(s->lref);
  dev_free(s->shade);
  dev_free(s->work);
  dev_free(s->faces);
  for (auto &e : s->scratch) dev_free(e.buf);
  delete s;
}
void shading_retain(dr_shading h) { h->refs.fetch_add(1); }
void shading_release(dr_shading h) {
  if (h->refs.fetch_sub(1) != 1) return;
  dev_free(h->shade);
  dev_free(h->faces);
  if (h->scene) scene_release(h->scene);
  delete h;
}

int scene_scratch(dr_scene s, hipStream_t stream, size_t bytes, void **out) {
  std::lock_guard<std::mutex> lock(s->scratch_mu);
  dr_scene_s::StreamScratch *e = nullptr;
  for (auto &x : s->scratch)
    if (x.stream == stream) e = &x;
  if (!e) {
    s->scratch.push_back({stream, DevBuf{}});
    e = &s->scratch.back();
  }
  if (e->buf.bytes < bytes) {
    // earlier calls on this stream may still read the old buffer
    DR_HIP_CHECK(hipStreamSynchronize(stream));
    dev_free(e->buf);
    const hipError_t err = hipMalloc(&e->buf.p, bytes);
    if (err != hipSuccess) {
      e->buf.p = nullptr;
      return fail(DR_ERR_ALLOC, std::string("dr_backward: vertex records: ") + hipGetErrorString(err));
    }
    e->buf.bytes = bytes;
  }
  *out = e->buf.p;
  return DR_OK;
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
  return dr::noalloc_scratch_init(device, nullptr);
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
  // the scene's own tables: material id 0 (not yet usable: dr_scene_set_materials) and flat normals
  std::vector<float> faces;
  dr::fill_faces(*s, nullptr, nullptr, faces);
  s->faces_host = faces;
  rc = dr::dev_upload(s->shade, s->shade_host.data(), s->shade_host.size() * sizeof(float4));
  if (rc == DR_OK) rc = dr::dev_upload(s->faces, faces.data(), faces.size() * sizeof(float));
  if (rc != DR_OK) { dr_scene_destroy(s); return rc; }
  *out = s;
  return DR_OK;
}

int dr_scene_set_normals(dr_scene s, const float *normals, int ntris) {
  if (!s || ntris != s->ntris) return dr::fail(DR_ERR_ARG, "dr_scene_set_normals: one normal triple per triangle required");
  std::vector<float> F;
  std::vector<float> uv(6 * (size_t)ntris);
  for (int i = 0; i < ntris; i++)
    memcpy(&uv[6 * (size_t)i], &s->faces_host[(size_t)i * dr::kFaceWords + 18], 6 * sizeof(float));
  dr::fill_faces(*s, normals, uv.data(), F);
  const int rc = dr::dev_overwrite(s->faces, F.data(), F.size() * sizeof(float));
  if (rc == DR_OK) s->faces_host = std::move(F);
  return rc;
}

int dr_scene_set_texcoords(dr_scene s, const float *uv, int ntris) {
  if (!s || ntris != s->ntris) return dr::fail(DR_ERR_ARG, "dr_scene_set_texcoords: one coordinate triple per triangle required");
  std::vector<float> F = s->faces_host;
  for (int i = 0; i < ntris; i++)
    for (int k = 0; k < 6; k++) F[(size_t)i * dr::kFaceWords + 18 + k] = uv ? uv[6 * (size_t)i + k] : 0.0f;
  const int rc = dr::dev_overwrite(s->faces, F.data(), F.size() * sizeof(float));
  if (rc == DR_OK) s->faces_host = std::move(F);
  return rc;
}

int dr_scene_set_materials(dr_scene s, const int32_t *mat_id, int ntris) {
  if (!s || ntris != s->ntris || (ntris > 0 && !mat_id)) return dr::fail(DR_ERR_ARG, "dr_scene_set_materials: one id per triangle required");
  std::vector<float4> sh;
  const int mx = dr::fill_shade(*s, mat_id, sh);
  if (mx == -2) return dr::fail(DR_ERR_ARG, "dr_scene_set_materials: negative material id");
  const int rc = dr::dev_overwrite(s->shade, sh.data(), sizeof(float4) * (size_t)ntris);
  if (rc != DR_OK) return rc;
  s->shade_host = std::move(sh);
  s->max_mat = mx;
  s->has_materials = true;
  return DR_OK;
}

int dr_shading_create(dr_scene s, const int32_t *mat_id, const float *normals, const float *uv, dr_shading *out) {
  if (!s || !out || (s->ntris > 0 && !mat_id)) return dr::fail(DR_ERR_ARG, "dr_shading_create: bad arguments");
  *out = nullptr;
  dr_shading h = new (std::nothrow) dr_shading_s();
  if (!h) return dr::fail(DR_ERR_ALLOC, "dr_shading_create: out of host memory");
  std::vector<float4> sh;
  std::vector<float> F;
  int rc = DR_OK;
  try {
    h->max_mat = s->ntris > 0 ? dr::fill_shade(*s, mat_id, sh) : -1;
    if (h->max_mat == -2) rc = dr::fail(DR_ERR_ARG, "dr_shading_create: negative material id");
    if (rc == DR_OK) dr::fill_faces(*s, normals, uv, F);
  } catch (const std::bad_alloc &) {
    rc = dr::fail(DR_ERR_ALLOC, "dr_shading_create: out of host memory");
  }
  if (rc == DR_OK) rc = dr::dev_upload(h->shade, sh.data(), sh.size() * sizeof(float4));
  if (rc == DR_OK) rc = dr::dev_upload(h->faces, F.data(), F.size() * sizeof(float));
  if (rc != DR_OK) {
    dr::dev_free(h->shade);
    dr::dev_free(h->faces);
    delete h;
    return rc;
  }
  dr::scene_retain(s);
  h->scene = s;
  *out = h;
  return DR_OK;
}

void dr_shading_destroy(dr_shading h) {
  if (h) dr::shading_release(h);
}

void dr_scene_destroy(dr_scene s) {
  if (s) dr::scene_release(s);
}

int64_t dr_scene_device_bytes(dr_scene s) { return s ? s->bytes() : 0; }

}  // extern "C"
