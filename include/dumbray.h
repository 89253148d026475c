/*
 * dumbray_amd C ABI: the drop-in boundary of the MI355X path-tracing hot path.
 *
 * Plain C types only (pointers and sizes, no torch or HIP types in signatures).
 * Every entry point returns an int status (DR_OK = 0) unless stated otherwise;
 * dr_last_error() returns a thread-local message for the last failure.  No C++
 * exception crosses this boundary.  The library owns device memory behind opaque
 * handles; callers own host buffers and no host pointer is retained after a call
 * returns.  Pointers documented as "device" are HBM addresses on the current
 * device; `stream` is a hipStream_t passed as void* (NULL = null stream).
 * Handles are immutable after creation and safe to use from several host
 * threads on distinct streams.
 */
#ifndef DUMBRAY_H
#define DUMBRAY_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
  DR_OK = 0,
  DR_ERR_ARG = 1,      /* invalid argument / shape */
  DR_ERR_HIP = 2,      /* HIP runtime error */
  DR_ERR_ALLOC = 3,    /* out of memory */
  DR_ERR_BUILD = 4     /* octree build failed */
};

/* Flags for dr_cast. */
#define DR_CAST_CLIP_BACKFACES 1u   /* Octree::cast's clipBackfaces argument (Octree.cuh:130) */

typedef struct dr_octree_s *dr_octree;   /* host-built octree (Octree<Renderable<BakedTriFace>>) */
typedef struct dr_scene_s *dr_scene;     /* device-resident scene (geometry + shading tables) */
typedef struct dr_shading_s *dr_shading; /* one renderer's shading tables of a scene (dr_shading_create) */

/* Thread-local text of the last error (never NULL). */
const char *dr_last_error(void);
/* Library version string (never NULL). */
const char *dr_version(void);

/* Make `device` current for the calling thread.
 * Replaces the per-GPU setup of BlockRenderer (BlockRenderer.cu:42-55). */
int dr_device_open(int device);

/* Octree<Renderable<BakedTriFace>>::push(all) + build()
 * (Octree.impl.cuh:97-129 push/build, :339-533 split/reduce).
 * tris: host, ntris x 9 floats (a, b, c); triangle i gets hit id i (push order,
 * = index into Octree::getData()).  split_threshold = OCTREE_POLYCOUNT_TO_SPLIT_NODE,
 * max_depth = OCTREE_MAX_DEPTH (Octree.cuh:11-13), both runtime here (max_depth <= 24).
 * nthreads <= 0 uses every host core. */
int dr_octree_build(const float *tris, int ntris, int split_threshold, int max_depth, int nthreads,
                    dr_octree *out);
/* Sizes: node count (Octree::getNodeCount, Octree.impl.cuh:219-221), total leaf-list
 * entries, device groups and triangle records. Any pointer may be NULL. */
int dr_octree_info(dr_octree t, int *node_count, int *list_total, int *ngroups, int *nrec);
/* Nodes in breadth-first order (children of a node contiguous, reference child
 * numbering): bounds[n*6] (start.xyz, end.xyz), children[n] (first child or -1),
 * list_off[n+1], list[list_total] (triangle ids in stored order). Host buffers. */
int dr_octree_export(dr_octree t, float *bounds, int32_t *children, int32_t *list_off, int32_t *list);
void dr_octree_destroy(dr_octree t);

/* Upload an octree to the current device (replaces TypeTools<Octree>::prepareForCpyLoad
 * deep copy + pointer relocation, Octree.impl.cuh:763-796).  The scene keeps its own
 * copy; the octree may be destroyed afterwards. */
int dr_scene_create(dr_octree t, dr_scene *out);
void dr_scene_destroy(dr_scene s);
/* Device bytes held by the scene. */
int64_t dr_scene_device_bytes(dr_scene s);

/* Batch Octree::cast (Octree.impl.cuh:174-212) + validateNotSameAsObject (:163).
 * rays: device, n x 6 floats (origin, direction; direction need not be unit).
 * exclude: device int32[n] triangle id the ray must not hit, or NULL.
 * out_id: device int32[n] hit id or -1; out_t: device float[n] hit distance
 * (FLT_MAX on miss, Octree.impl.cuh:170); out_point: device float[n*3] or NULL.
 * Asynchronous on `stream`. */
int dr_cast(dr_scene s, const float *rays, int n, unsigned flags, const int32_t *exclude, int32_t *out_id,
            float *out_t, float *out_point, void *stream);

/* ---------------------------------------------------------------- path tracing */

/* Material kinds (plugins of Material<BakedTriFace>, Material.cuh:115-123).  Kinds 0-2
 * are the plugins the benchmark scenes use (the reference has no such shaders); kinds
 * 3-6 are the reference's own shaders, without textures (ColoredTexture colours only). */
enum {
  DR_LAMBERT = 0,
  DR_MIRROR = 1,
  DR_METAL = 2,
  DR_DUMB_BASED = 3,          /* DumbBasedShader, DumbBasedShader.impl.cuh:5-122 */
  DR_SIMPLE_STOCHASTIC = 4,   /* SimpleStochasticShader, SimpleStochasticShader.impl.cuh:4-37 */
  DR_FRESNEL = 5,             /* Shaders::FresnelShader, FresnelShader.impl.cuh:6-19 */
  DR_DEFAULT_SHADER = 6       /* DefaultShaderGeneric, DefaultShader.impl.cuh:5-52 */
};

/* A texture (Texture, Texture.cuh:13-60): width x height RGBA texels, row 0 at the top,
 * starting at texel `offset` of the texel array; filter 0 = FILTER_NONE, 1 = FILTER_BILINEAR. */
typedef struct {
  int32_t offset, width, height, filter;
} dr_texture;

/* A ColoredTexture's texture part (DumbStructs.cuh:70-92): the texture (-1: the colour
 * alone), sampled at texture coordinate * tiling + offset. */
typedef struct {
  int32_t texture;
  float tiling[2], offset[2];
} dr_texture_map;

/* Textures of one material (96 B), referenced by dr_material.reserved[0] = 1 + index:
 *   map[0] the colour's texture (DumbBasedShader fresnelColor, FresnelShader color),
 *   map[1] DumbBasedShader diffuseColor, map[2] normalColor, map[3] alphaCutout;
 * normal_color = normalColor's colour (the normal map's channel weights), alpha_threshold =
 * alphaCutoutThreshold (DumbBasedShader.cuh:25-31). */
typedef struct {
  dr_texture_map map[4];
  float normal_color[3];
  float alpha_threshold;
} dr_material_maps;

/* One material (48 B).  Fields by kind:
 *   LAMBERT / MIRROR / METAL: albedo = Lambert albedo / mirror reflectance / Schlick F0.
 *   DUMB_BASED: albedo = fresnel colour, diffuse = diffuse colour already scaled by
 *     (1 - metal) as the constructor does, p[0] = specular exponent, p[1] = metal in [0, 1].
 *   SIMPLE_STOCHASTIC / DEFAULT_SHADER: albedo = colour, p[0] = diffuse, p[1] = gloss
 *     (both divided by their sum when it exceeds 1, as the constructors do), p[2] = shininess.
 *   FRESNEL: albedo = fresnel colour.
 * reserved[0]: 0, or 1 + the index of the material's dr_material_maps (DUMB_BASED, FRESNEL). */
typedef struct {
  int32_t kind;
  float albedo[3];
  float diffuse[3];
  float p[3];
  int32_t reserved[2];
} dr_material;

/* Light kinds (Light::getVertexPhotons, Light.impl.cuh:34-36).  AREA is the benchmark
 * plugin; the others are the reference's own lights. */
enum {
  DR_AREA_LIGHT = 0,              /* one-sided quad emitting on the (edge_u x edge_v) side */
  DR_DIRECTIONAL_LIGHT = 1,       /* SimpleDirectionalLight, SimpleDirectionalLight.impl.cuh:7-16 */
  DR_SOFT_DIRECTIONAL_LIGHT = 2,  /* SimpleSoftDirectionalLight, SimpleSoftDirectionalLight.impl.cuh:5-19 */
  DR_SPHERICAL_LIGHT = 3,         /* SphericalLightEmitter, SphericalLightEmitter.impl.cuh:5-18 */
  DR_SPOT_LIGHT = 4               /* Lights::Spotlight, Spotlight.impl.cuh:7-90 */
};

/* One light (64 B).  Fields by kind:
 *   AREA: corner, edge_u, edge_v, emission (radiance scale).
 *   DIRECTIONAL: emission = colour, edge_u = unit direction, dist = distance of the photon origin.
 *   SOFT_DIRECTIONAL: as DIRECTIONAL, plus radius = distance * softness.
 *   SPHERICAL: emission = colour, corner = centre, radius.
 *   SPOT: emission = colour * luminosity, corner = position, edge_u = unit direction,
 *     edge_v = (cos inner angle, cos outer angle, falloff power), radius = emitter size,
 *     reserved = sample count (0-16) | 32 (casts shadows) | 64 (emits from a sphere, else a disc);
 *     the values the Spotlight constructor stores (Spotlight.impl.cuh:7-38). */
typedef struct {
  float corner[3], edge_u[3], edge_v[3], emission[3];
  int32_t kind;
  float dist, radius;
  int32_t reserved;
} dr_light;

/* Lenses (Lense plugins, Lense.cuh:46-82). */
enum {
  DR_STOCHASTIC_LENS = 0,  /* SimpleStochasticLense: jitter inside the subpixel, SimpleStochasticLense.impl.cuh:11-25 */
  DR_PERSPECTIVE_LENS = 1, /* DefaultPerspectiveLense: subpixel centre, unit direction, DefaultPerspectiveLense.impl.cuh:12-19 */
  DR_SPHERICAL_SEGMENT_LENS = 2  /* Lenses::SphericalSegmentLense: angular pixel grid, disc aperture and a colour filter,
                                    SphericalSegmentLense.impl.cuh:6-34; needs lens_p and cam_rows; dr_render_blocks only */
};

/* Plugin sets. */
enum {
  DR_PLUGINS_BENCHMARK = 0,  /* kinds 0-2 and area lights; blockers are opaque */
  DR_PLUGINS_REFERENCE = 1   /* every kind above and the reference's photon transparency
                                (DumbRenderer.cu:396-413); dr_render_blocks only */
};

/* Pixel boxing (DumbRenderer::BoxingMode, DumbRenderer.cuh:12-18): the pixel size is
 * 1 / height, 1 / width, 1 / min(width, height) or 1 / max(width, height)
 * (DumbRenderer.cu:190-200). */
enum {
  DR_BOXING_HEIGHT = 0,
  DR_BOXING_WIDTH = 1,
  DR_BOXING_MIN = 2,
  DR_BOXING_MAX = 3
};

/* Per-call render parameters (DumbRenderer::PixelRenderProcess::SceneConfiguration,
 * DumbRenderer.cuh:80-101, plus the estimator's seeds). */
typedef struct {
  int32_t width, height;          /* image size in pixels */
  int32_t fsaa_x, fsaa_y;         /* subpixel grid = samples per pixel per iteration */
  int32_t max_bounces;            /* DumbRenderer maxBounces (<= 64, DumbRenderer.cuh:7) */
  int32_t rr_start;               /* Russian roulette at layers >= rr_start */
  float rr_survive;               /* survival probability (1 disables) */
  int32_t cull_backfaces;         /* DumbRenderer ignoreBackfaces */
  uint32_t seed_lo, seed_hi;      /* Philox key */
  uint32_t request;               /* Philox counter word 3 (batched requests) */
  int32_t iteration;              /* 1-based; 1 sets the frame, k > 1 blends with 1/k (DumbRenderer.cu:64) */
  float cam[13];                  /* camera: position, right, up, forward, focal = 0.5/tan(fov/2) */
  const dr_material *materials;   /* device table [nmat] */
  const dr_light *lights;         /* device table [nlights] */
  int32_t nmat, nlights;          /* nlights <= 16 */
  int32_t lens;                   /* DR_STOCHASTIC_LENS or DR_PERSPECTIVE_LENS */
  int32_t plugins;                /* DR_PLUGINS_BENCHMARK or DR_PLUGINS_REFERENCE */
  int32_t boxing;                 /* DR_BOXING_* (0 = HEIGHT, the reference's default) */
  float lens_p[8];                /* DR_SPHERICAL_SEGMENT_LENS: angle x, angle y (degrees), focal distance,
                                     sensor size, filter r, g, b, unused */
  float cam_rows[9];              /* the camera Transform's rows trans.x, trans.y, trans.z (scale applied): the
                                     point map of lens origins (Transform.impl.h:85-87); used by that lens */
  const dr_material_maps *maps;   /* device table [nmaps] (materials with textures), or NULL */
  const dr_texture *textures;     /* device table [ntextures], or NULL */
  const float *texels;            /* device RGBA fp32 texels of every texture */
  int32_t nmaps, ntextures;
  dr_shading shading;             /* material ids, vertex normals and texture coordinates of this render
                                     (dr_shading_create for the same scene), or NULL for the scene's own
                                     (dr_scene_set_materials / _normals / _texcoords) */
} dr_render_params;

/* Attach per-triangle material ids (host int32[ntris], ntris = the octree's triangle
 * count; the scene keeps a device copy).  Needed before rendering with the scene's own
 * tables (dr_render_params.shading = NULL).  This and the two calls below wait for the
 * device to finish its work, then overwrite the scene's tables in place: renders and
 * batch replays issued afterwards read the new contents. */
int dr_scene_set_materials(dr_scene s, const int32_t *mat_id, int ntris);

/* Per-vertex normals of every triangle (BakedTriFace::norm, BakedTriMesh.h:8): host
 * float[ntris * 9] (a, b, c normals), or NULL for flat normals, the unit
 * (c - a) x (b - a) of each triangle (Triangle::normal, Triangle.impl.h:233-235), which
 * is also the default.  Used by the reference shaders (DR_PLUGINS_REFERENCE). */
int dr_scene_set_normals(dr_scene s, const float *normals, int ntris);

/* Texture coordinates of every triangle (BakedTriFace::tex, BakedTriMesh.h:8): host
 * float[ntris * 6] (u, v of a, b, c), or NULL for zeros (the default).  Used by textured
 * materials of the reference shaders. */
int dr_scene_set_texcoords(dr_scene s, const float *uv, int ntris);

/* Shading tables of one renderer of scene `s`, in device memory owned by the handle:
 * material ids (host int32[ntris], required), vertex normals (host float[ntris * 9] or NULL for
 * flat normals) and texture coordinates (host float[ntris * 6] or NULL for zeros), with the
 * layouts and defaults of dr_scene_set_materials / _normals / _texcoords.  A render passes the
 * handle in dr_render_params.shading, so renderers of one scene may shade it differently at
 * the same time (the reference's TriScene carries one material per Renderable element,
 * Scene.cuh:68; DumbRenderer reads it at Material.impl.cuh:122-170).  The handle keeps its
 * scene alive; a batch keeps the handles of its requests alive until dr_batch_destroy. */
int dr_shading_create(dr_scene s, const int32_t *mat_id, const float *normals, const float *uv, dr_shading *out);
void dr_shading_destroy(dr_shading h);

/* DumbRenderer::renderBlocksGPU (DumbRenderer.cu:87-113, override of
 * BlockRenderer::renderBlocksGPU, BlockRenderer.cuh:43-44): renders framebuffer
 * blocks [start_block, end_block) of one iteration into `frame`, a device
 * BlockBasedFrameBuffer<8,16> (RGBA fp32, pixel (x, y) of block b at
 * frame[4 * (128 * b + 8 * (y % 16) + x % 8)], blocks numbered row by row,
 * BlockBasedFrameBuffer.impl.cuh:108-156).  Iteration 1 sets pixels, iteration k
 * blends colour * (1/k) + old * (1 - 1/k).  Asynchronous on `stream`. */
int dr_render_blocks(dr_scene s, const dr_render_params *p, float *frame, int start_block, int end_block,
                     void *stream);

/* Renderer::backward: adjoint of one dr_render_blocks iteration over the whole image
 * by path replay (each path is re-traced from its Philox key, nothing is stored
 * between passes).  dl_dframe: device, same block layout as `frame` (alpha ignored).
 * Adds dL/dalbedo (Lambert albedo and mirror reflectance; device double[nmat*3]) and
 * dL/demission (device double[nlights*3]); the caller zeroes them.  Per-block
 * reductions in LDS precede one FP64 atomic per block and parameter.  The benchmark
 * plugin set only; every lens (the spherical segment lens's filter scales dL/dI). */
int dr_backward(dr_scene s, const dr_render_params *p, const float *dl_dframe, double *g_albedo, double *g_emission,
                void *stream);

/* dr_backward for callers that cannot allocate device memory (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc).  No counterpart in the
 * reference.  Same arguments and results as dr_backward; the block counter and the vertex
 * records of the path replay live in a 40 MiB array that the library holds on every device it runs on until
 * the process ends (its address is resolved by dr_device_open, or by the first call on a
 * device), and the persistent grid is sized to what that array holds (all resident
 * workgroups up to max_bounces 4, fewer beyond).  Calls no hipMalloc/hipFree.  Every call
 * on one device shares the array: calls that overlap on a device (several streams or host
 * threads) race and must be serialised by the caller.  Returns DR_ERR_ARG when one
 * workgroup's records (256 lanes x (max_bounces + 1) x 36 B) and the 256-B counter exceed
 * the array. */
int dr_backward_noalloc(dr_scene s, const dr_render_params *p, const float *dl_dframe, double *g_albedo,
                        double *g_emission, void *stream);

/* ---------------------------------------------------------------- batched requests */

/* One render request of a batch: a scene, its parameters (device material/light
 * tables) and its device BlockBasedFrameBuffer<8,16> (params.width x params.height). */
typedef struct {
  dr_scene scene;
  dr_render_params params;
  float *frame;
} dr_request;

/* Launch shape of the batched kernel. */
typedef struct {
  int32_t tile_blocks;     /* framebuffer blocks per work item of the device queue */
  int32_t groups_per_cu;   /* persistent workgroups per compute unit draining the queue */
  int32_t refill_min;      /* idle lanes a wave collects before it takes new path jobs */
} dr_batch_tuning;

typedef struct dr_batch_s *dr_batch;

/* Many renders in one persistent launch (replaces BlockRenderer's per-device block
 * batches and host round trips, BlockRenderer.cu:74-97, FrameBuffer.impl.cuh:261-268,
 * 355-379).  Work items (request, tile of blocks) are queued longest first and
 * drained by persistent workgroups through one device atomic; the launch (queue
 * reset + kernel) is captured once in a hipGraph.  tuning == NULL runs the autotuner
 * (timed proxy renders into scratch frames; request frames are not touched).
 * Requests of both plugin sets and every lens may be mixed: each plugin set has its own
 * queue and persistent kernel, launched one after the other in the same graph.
 * Every request's frame equals what dr_render_blocks would produce for it. */
int dr_batch_create(const dr_request *reqs, int n, const dr_batch_tuning *tuning, dr_batch *out);
/* Replays the captured graph on `stream`. */
int dr_batch_run(dr_batch b, void *stream);
int dr_batch_get_tuning(dr_batch b, dr_batch_tuning *out);
void dr_batch_destroy(dr_batch b);

/* ---------------------------------------------------------------- scene files */

/* A scene loaded from the reference's scene description (.dumb, Dson text) and the
 * Wavefront OBJ files it names: DumbRenderContext::fromFile / fromDson
 * (DumbRenderContext.cu:310-371) with its include, materials, lights, groups, objects,
 * camera and renderer sections (:641-787, :106-271, :476-640), MeshReader::readObj
 * (MeshReader.impl.h) and PolyMesh::bake (PolyMesh.impl.h:140-152).  Host memory only.
 * Supported plugin types: materials dumb_based, simple_stochastic, fresnel_color;
 * lights simple_soft_directional, simple_spherical, spotlight; lenses simple_stochastic,
 * spherical_segment; PNG textures.  Primitive meshes are rejected with an error.  Thread, device and block settings are checked and ignored. */
typedef struct dr_scene_file_s *dr_scene_file;

/* Renderer settings and camera of a scene file (DumbRenderContextData::RendererSettings,
 * DumbRenderContext.cu:43-50, defaults :282-289), in dr_render_params terms. */
typedef struct {
  int32_t boxing;            /* "pixel"/"boxing"; default DR_BOXING_HEIGHT */
  int32_t max_bounces;       /* "pixel"/"bounces"; default 2 */
  int32_t fsaa_x, fsaa_y;    /* "pixel"/"samples_per_pixel"; default 1 x 1 */
  int32_t ignore_backfaces;  /* "pixel"/"ignore_backfaces"; default 1 */
  int32_t lens;              /* DR_STOCHASTIC_LENS or DR_SPHERICAL_SEGMENT_LENS */
  float cam[13];             /* dr_render_params.cam: camera transform and lens (default angle 64) */
  float lens_p[8];           /* dr_render_params.lens_p */
  float cam_rows[9];         /* dr_render_params.cam_rows */
} dr_scene_file_settings;

/* DumbRenderContext::fromFile (DumbRenderContext.cu:310-335): parse `path` and every
 * file it includes or names; relative names resolve against the including file's
 * directory first.  On error returns DR_ERR_ARG and dr_last_error() holds the message. */
int dr_scene_file_load(const char *path, dr_scene_file *out);
/* DumbRenderContext::fromDson on scene text; relative file names resolve against
 * base_dir (may be NULL or "" for the working directory). */
int dr_scene_file_parse(const char *text, const char *base_dir, dr_scene_file *out);
/* Counts (triangles in push order, materials, lights) and settings; any pointer may be NULL. */
int dr_scene_file_info(dr_scene_file f, dr_scene_file_settings *settings, int *ntris, int *nmat, int *nlights);
/* Host arrays for dr_octree_build (tris[ntris*9]), dr_scene_set_normals
 * (normals[ntris*9], BakedTriFace::norm) and dr_scene_set_materials (mat_id[ntris]);
 * any pointer may be NULL. */
int dr_scene_file_geometry(dr_scene_file f, float *tris, float *normals, int32_t *mat_id);
/* Material and light tables (materials[nmat], lights[nlights]); either may be NULL. */
int dr_scene_file_tables(dr_scene_file f, dr_material *materials, dr_light *lights);
/* Textures loaded from the scene's PNG files (DumbRenderContext::getImageId,
 * DumbRenderContext.cu:374-452) and the materials' maps: counts, then the tables
 * (textures[ntextures], texels[ntexels * 4] RGBA = byte / 255, maps[nmaps]); any pointer may be NULL. */
int dr_scene_file_texture_info(dr_scene_file f, int *ntextures, int64_t *ntexels, int *nmaps);
int dr_scene_file_textures(dr_scene_file f, dr_texture *textures, float *texels, dr_material_maps *maps);
/* Texture coordinates of every triangle for dr_scene_set_texcoords (uv[ntris * 6]). */
int dr_scene_file_texcoords(dr_scene_file f, float *uv);
void dr_scene_file_destroy(dr_scene_file f);

/* ---------------------------------------------------------------- image output */

/* Images::saveBufferPNG's colour conversion (Images.cu:7-19, :28-32): frame (device
 * BlockBasedFrameBuffer<8,16>, width x height) to out (device, row-major RGBA8,
 * width*height*4 bytes), each channel 255*c clamped to [0, 255] and truncated (NaN -> 0).
 * Asynchronous on `stream`. */
int dr_frame_to_rgba8(const float *frame, int width, int height, uint8_t *out, void *stream);
/* lodepng::encode as Images::saveBufferPNG calls it (Images.cu:33): an 8-bit RGBA PNG
 * file of a host row-major RGBA8 image. */
int dr_png_write(const char *path, const uint8_t *rgba, int width, int height);
/* lodepng::decode as Images::getTexturePNG calls it (Images.cu:38-42): a PNG file (any colour
 * type and bit depth, tRNS, interlaced or not) as 8-bit RGBA, row 0 at the top.
 * dr_png_info gives the size; dr_png_read fills rgba[width*height*4] (host). */
int dr_png_info(const char *path, int *width, int *height);
int dr_png_read(const char *path, uint8_t *rgba, int width, int height);

/* dr_cast plus per-ray work counters of the traversal that ran (measurement only;
 * no counterpart in the reference, which prints timings only): out_counts is a
 * device int32[n*6] receiving, per ray, the groups expanded (196 B read each), leaves
 * tested, triangle plane tests (32 B record reads), inside tests (a further 32 B each),
 * children of the expanded groups with a non-empty leaf below (octant tests) and children
 * whose slab test ran.  Same results as dr_cast. */
int dr_cast_stats(dr_scene s, const float *rays, int n, unsigned flags, const int32_t *exclude, int32_t *out_id,
                  float *out_t, int32_t *out_counts, void *stream);

/* dr_render_blocks plus work counters of the render (measurement only; no counterpart in the
 * reference, which prints timings only).  counts: device uint64[9], added to (the caller zeroes
 * it): casts (camera, bounce and photon rays), of which light photons, group expansions (196 B
 * read each), leaves tested, triangle plane tests and inside tests (32 B of a record each), path
 * jobs started, children of expanded groups with a non-empty leaf below (octant tests) and
 * children whose slab test ran.  Same frame as dr_render_blocks. */
int dr_render_stats(dr_scene s, const dr_render_params *p, float *frame, int start_block, int end_block,
                    uint64_t *counts, void *stream);

#ifdef __cplusplus
}
#endif
#endif
