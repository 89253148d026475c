/*
 * CPU oracle of the path estimator (TEST INFRASTRUCTURE ONLY, see dr_oracle.h).
 *
 * Restates DumbRenderer::PixelRenderProcess (DumbRenderer.cu:174-443): per
 * subpixel a camera sample, per layer one Octree::cast with the previous
 * layer's triangle excluded (:285, :365-366), photons cast from each light
 * toward the hit point and counted as lit when they first hit the same
 * triangle or land within sqrt(8 eps) of it (:388-395), indirect samples
 * requested from the material (:369-385), and the parent layer accumulating
 * getReflectedColor(child photon) * sampleWeight (:330-345).  Layer colours
 * are accumulated bottom-up, as the reference does.
 *
 * The materials and the light here are the plugins the benchmark scenes need
 * (none exists in the reference): Lambert (albedo/pi, cosine-weighted
 * sampling), perfect mirror, Schlick-Fresnel metal (DumbTools::fresnel,
 * DumbStructs.cuh:136-140) and a one-sided area quad light, plus Russian
 * roulette with a fixed survival probability.  Random numbers come from
 * Philox4x32-10 (Salmon et al., SC'11) keyed by (seed) with counter
 * (pixel, sample, layer<<8 | purpose, request), replacing the reference's
 * per-thread XORWOW, so every sample is reproducible on any schedule.
 */
#ifndef DR_ORACLE_RENDER_H
#define DR_ORACLE_RENDER_H
#include <stdint.h>
#include "dr_oracle.h"
#ifdef __cplusplus
extern "C" {
#endif

enum { ORA_LAMBERT = 0, ORA_MIRROR = 1, ORA_METAL = 2, ORA_DUMB_BASED = 3, ORA_SIMPLE_STOCHASTIC = 4,
       ORA_FRESNEL = 5, ORA_DEFAULT_SHADER = 6 };

/* Same layouts as the library's dr_texture, dr_texture_map and dr_material_maps. */
typedef struct { int32_t offset, width, height, filter; } ora_texture;
typedef struct { int32_t texture; float tiling[2], offset[2]; } ora_texture_map;
typedef struct {
	ora_texture_map map[4];        /* colour (DumbBased fresnelColor / Fresnel color), diffuse, normal, alpha cutout */
	float normal_color[3];
	float alpha_threshold;
} ora_material_maps;
/* Textures of a scene: per-triangle texture coordinates (6 floats: u, v of a, b, c), the
 * materials' maps (ora_material.reserved[0] = 1 + index), texture descriptors and RGBA texels. */
typedef struct {
	const float *uv;
	const ora_material_maps *maps; int nmaps;
	const ora_texture *tex; int ntex;
	const float *texels;
} ora_textures;

/* Same layout as the library's dr_material (48 B). */
typedef struct {
	int32_t kind;
	float albedo[3];    /* Lambert albedo, mirror reflectance, metal F0; reference shaders: colour / fresnel colour */
	float diffuse[3];   /* DumbBasedShader diffuse colour (already scaled by 1 - metal) */
	float p[3];         /* DumbBased: specular, metal; SimpleStochastic / Default: diffuse, gloss, shininess */
	int32_t reserved[2];
} ora_material;

enum { ORA_AREA_LIGHT = 0, ORA_DIRECTIONAL_LIGHT = 1, ORA_SOFT_DIRECTIONAL_LIGHT = 2, ORA_SPHERICAL_LIGHT = 3,
       ORA_SPOT_LIGHT = 4 };

/* Same layout as the library's dr_light (64 B). */
typedef struct {
	float corner[3], edge_u[3], edge_v[3], emission[3];   /* area: emits on the (edge_u x edge_v) side */
	int32_t kind;
	float dist, radius;     /* directional: edge_u = unit direction, dist; soft: radius = dist * softness;
	                           spherical: corner = centre, radius; spot: corner = position, edge_u = unit
	                           direction, edge_v = (cos inner, cos outer, falloff), radius = emitter size */
	int32_t reserved;       /* spot: sample count | 32 casts shadows | 64 emits from a sphere */
} ora_light;

typedef struct {
	int32_t width, height, fsaa_x, fsaa_y, max_bounces, rr_start;
	float rr_survive;
	int32_t cull_backfaces;
	uint32_t seed_lo, seed_hi, request;
	int32_t iteration;              /* 1-based; samples of iteration k are numbered from (k-1)*fsaa_x*fsaa_y */
	float cam[13];                  /* position, right, up, forward, focal (0.5 / tan(fov/2)) */
	int32_t lens;                   /* 0: SimpleStochasticLense jitter, 1: DefaultPerspectiveLense */
	int32_t boxing;                 /* pixel size 1/H, 1/W, 1/min(W,H), 1/max(W,H) for 0..3 */
	float lens_p[8];                /* lens 2 (SphericalSegmentLense): angle x, y, distance, sensor, filter rgb */
	float cam_rows[9];              /* camera Transform rows trans.x/y/z: point map of lens-2 origins */
} ora_render_cfg;

/* Pixel colours (row-major, rows [row0, row1) x W x 3) of one iteration, into out_rgb
 * (fp32 shading) or out_rgb64 (shading accumulated in double; geometry and sampling
 * stay fp32, so the paths are identical). */
int ora_render(const ora_octree *t, const int32_t *mat_id, const ora_material *mats, int nmat, const ora_light *lights,
               int nlights, const ora_render_cfg *cfg, float *out_rgb, double *out_rgb64, int row0, int row1,
               int nthreads);
/* FP64 adjoint of the same estimator: given dL/dimage (H x W x 3), accumulates
 * dL/dalbedo (nmat x 3, Lambert albedo and mirror reflectance) and
 * dL/demission (nlights x 3).  Paths are replayed from their Philox keys. */
int ora_backward(const ora_octree *t, const int32_t *mat_id, const ora_material *mats, int nmat,
                 const ora_light *lights, int nlights, const ora_render_cfg *cfg, const float *dl_dimage,
                 double *g_albedo, double *g_emission, int nthreads);

/* The reference estimator with the reference's own plugins (DumbRenderer.cu:256-443 with
 * one sample per layer): shaders DumbBasedShader, SimpleStochasticShader, FresnelShader,
 * DefaultShaderGeneric (no textures), lights SimpleDirectionalLight,
 * SimpleSoftDirectionalLight, SphericalLightEmitter (and the area quad), and a light
 * photon that reaches another surface continuing through it with the colour that
 * surface's getReflectedColor returns when that colour is not black (:396-413).
 * normals: float[ntris*9] per-vertex normals or NULL for flat unit (c-a)x(b-a).
 * Layer colours are accumulated bottom-up in fp32. */
int ora_render_ref(const ora_octree *t, const int32_t *mat_id, const ora_material *mats, int nmat,
                   const ora_light *lights, int nlights, const float *normals, const ora_render_cfg *cfg,
                   float *out_rgb, int row0, int row1, int nthreads);
/* The same with textures (tx may be NULL). */
int ora_render_ref_tex(const ora_octree *t, const int32_t *mat_id, const ora_material *mats, int nmat,
                       const ora_light *lights, int nlights, const float *normals, const ora_textures *tx,
                       const ora_render_cfg *cfg, float *out_rgb, int row0, int row1, int nthreads);
/* Texture::operator()(Vector2) of one texture (probe for tests): RGBA into out. */
void ora_texture_sample(const ora_texture *t, const float *texels, float x, float y, float out[4]);

/* Probes of the reference plugins for tests.  photon_type: 0 direct, 1 indirect. */
float ora_pow(float x, float y);
float ora_asin(float x);          /* arcsin on [0, 1] as the SphericalSegmentLense aperture uses it */
/* camera ray of subpixel (i, j) of pixel (px, py) for the lens uniforms u (probe for tests) */
void ora_camera_ray(const ora_render_cfg *cfg, int px, int py, int i, int j, const float u[4], float o[3], float d[3]);
void ora_point_on_sphere(float ua, float ub, float radius, float out[3]);
void ora_reflected_color(const ora_material *m, const float tri[9], const float nrm[9], const float p[3],
                         const float photon_dir[3], const float photon_color[3], const float observer[3],
                         int photon_type, int sample_type, float out[3]);
/* requestIndirectSamples of one material at (tri, p) for the layer ray direction d, with the
 * uniforms of (pix, samp, layer); returns 1 with the sample direction and type, or 0. */
int ora_indirect_sample(const ora_material *m, const float tri[9], const float nrm[9], const float p[3],
                        const float d[3], const ora_render_cfg *cfg, uint32_t pix, uint32_t samp, int layer,
                        float dir[3], int *type);
/* getVertexPhotons of light l for the point p; returns 1 with the photon ray and colour, or 0. */
int ora_light_photon(const ora_light *L, int l, int k, const float p[3], const ora_render_cfg *cfg, uint32_t pix,
                     uint32_t samp, int layer, float origin[3], float dir[3], float color[3]);

/* Philox4x32-10 block and the sin/cos of 2*pi*u used by the samplers (probes for tests). */
void ora_philox(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]);
void ora_sincos_2pi(float u, float *s, float *c);

#ifdef __cplusplus
}
#endif
#endif
