/*
 * CPU oracle of the path estimator (see dr_oracle_render.h).
 * TEST INFRASTRUCTURE ONLY.  Geometry, rays and sampling are fp32 with one
 * IEEE operation per source operation (built with -ffp-contract=off); shading
 * is evaluated in fp32 (ora_render, fp64 = 0) or fp64 (fp64 = 1, ora_backward).
 */
#include "dr_oracle_render.h"
#include <math.h>
#include <pthread.h>
#include <stdlib.h>
#include <string.h>
#include <unistd.h>

#define VEPS 0.000005f                          /* VECTOR_EPSILON (Vector3.h:10) */
#define INV_PI_F 0.31830988618379067154f
#define MAX_LIGHTS 16

int ora_cast1(const ora_octree *t, const float *o, const float *d, int clip, int32_t exclude, int32_t *id, float *dist,
              float *point);
const float *ora_octree_tri(const ora_octree *t, int32_t id);

typedef struct { float x, y, z; } v3;
static inline v3 V(float x, float y, float z) { v3 r; r.x = x; r.y = y; r.z = z; return r; }
static inline v3 vadd(v3 a, v3 b) { return V(a.x + b.x, a.y + b.y, a.z + b.z); }
static inline v3 vsub(v3 a, v3 b) { return V(a.x - b.x, a.y - b.y, a.z - b.z); }
static inline v3 vmul(v3 a, float f) { return V(a.x * f, a.y * f, a.z * f); }
static inline float vdot(v3 a, v3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
static inline v3 vcross(v3 a, v3 b) { return V(a.y * b.z - b.y * a.z, b.x * a.z - a.x * b.z, a.x * b.y - b.x * a.y); }
static inline v3 vnormalized(v3 a) { float il = 1.0f / sqrtf(a.x * a.x + a.y * a.y + a.z * a.z); return vmul(a, il); }
static inline v3 ld3(const float *p) { return V(p[0], p[1], p[2]); }

/* ------------------------------------------------------------------ */
/* Philox4x32-10 (Salmon, Moraes, Dror, Shaw: "Parallel random numbers: */
/* as easy as 1, 2, 3", SC'11)                                          */
/* ------------------------------------------------------------------ */
void ora_philox(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]) {
	uint32_t c0 = ctr[0], c1 = ctr[1], c2 = ctr[2], c3 = ctr[3], k0 = key[0], k1 = key[1];
	for (int r = 0; r < 10; r++) {
		const uint64_t p0 = (uint64_t)0xD2511F53u * c0, p1 = (uint64_t)0xCD9E8D57u * c2;
		const uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0, hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
		c0 = hi1 ^ c1 ^ k0; c1 = lo1; c2 = hi0 ^ c3 ^ k1; c3 = lo0;
		k0 += 0x9E3779B9u; k1 += 0xBB67AE85u;
	}
	out[0] = c0; out[1] = c1; out[2] = c2; out[3] = c3;
}
/* sin and cos of 2*pi*u, u in [0,1): quadrant from 4u, Taylor series on [0, pi/2) */
void ora_sincos_2pi(float u, float *s, float *c) {
	const float x = u * 4.0f;
	const int q = (int)x;
	const float f = x - (float)q;
	const float a = f * 1.57079637f;
	const float a2 = a * a;
	const float sa = a * (1.0f + a2 * (-0.166666667f + a2 * (0.00833333377f + a2 * (-0.000198412701f
	                 + a2 * (2.75573188e-06f + a2 * (-2.50521079e-08f))))));
	const float ca = 1.0f + a2 * (-0.5f + a2 * (0.0416666679f + a2 * (-0.00138888892f + a2 * (2.48015876e-05f
	                 + a2 * (-2.75573188e-07f + a2 * 2.08767570e-09f)))));
	switch (q & 3) {
	case 0: *s = sa; *c = ca; break;
	case 1: *s = ca; *c = -sa; break;
	case 2: *s = -sa; *c = -ca; break;
	default: *s = -ca; *c = sa; break;
	}
}

typedef struct {
	const ora_octree *tree; const int32_t *mat_id; const ora_material *mats; int nmat;
	const ora_light *lights; int nlights; const ora_render_cfg *cfg;
	const float *normals;           /* reference estimator: per-vertex normals or NULL */
	const ora_textures *tx;         /* reference estimator: textures or NULL */
} rctx;

/* 4 uniforms in [0,1) for (pixel, sample, layer, purpose): purpose 0 = lens,
 * 1 = BSDF direction (u0, u1) + Russian roulette (u2), 16 + l = light l */
static void rng4(const rctx *C, uint32_t pix, uint32_t samp, uint32_t layer, uint32_t purpose, float u[4]) {
	const uint32_t ctr[4] = {pix, samp, (layer << 8) | purpose, C->cfg->request};
	const uint32_t key[2] = {C->cfg->seed_lo, C->cfg->seed_hi};
	uint32_t o[4];
	ora_philox(ctr, key, o);
	for (int k = 0; k < 4; k++) u[k] = (float)(o[k] >> 8) * (1.0f / 16777216.0f);
}

/* DumbRenderer::PixelRenderProcess::countPixelSize (DumbRenderer.cu:190-200) */
static float pixel_size(float W, float H, int boxing) {
	if (boxing == 1) return 1.0f / W;
	if (boxing == 2) return 1.0f / ((H <= W) ? H : W);
	if (boxing == 3) return 1.0f / ((H >= W) ? H : W);
	return 1.0f / H;
}

/* Camera sample of subpixel (i, j) of pixel (px, py): DumbRenderer.cu:190-232 with
 * the boxing's pixel size, SimpleStochasticLense jitter (SimpleStochasticLense.impl.cuh:11-21,
 * DumbRand::range = u*(max-min)+min) or DefaultPerspectiveLense, Camera transform
 * (Camera.impl.cuh:7-10, Transform::direction Transform.impl.h:166-168). */
/* sin/cos of an angle in radians through ora_sincos_2pi of the angle in turns reduced to [0, 1) */
static void sincos_rad(float a, float *s, float *c) {
	float t = a * 0.159154943f;
	t = t - floorf(t);
	ora_sincos_2pi(t, s, c);
}

/* arcsin on [0, 1]: Taylor series of arcsin to x^21 for x <= 1/2, asin(x) = pi/2 - 2 asin(sqrt((1-x)/2)) above */
static float asin_series(float x) {
	const float z = x * x;
	float p = 7.312525874e-03f;
	p = p * z + 8.390335810e-03f;
	p = p * z + 9.761609529e-03f;
	p = p * z + 1.155180090e-02f;
	p = p * z + 1.396484375e-02f;
	p = p * z + 1.735276442e-02f;
	p = p * z + 2.237215909e-02f;
	p = p * z + 3.038194444e-02f;
	p = p * z + 4.464285714e-02f;
	p = p * z + 7.500000000e-02f;
	p = p * z + 1.666666667e-01f;
	return x + x * (z * p);
}
float ora_asin(float x) {
	if (x <= 0.5f) return asin_series(x);
	return 1.57079637f - 2.0f * asin_series(sqrtf((1.0f - x) * 0.5f));
}

static void camera_ray(const ora_render_cfg *c, int px, int py, int i, int j, const float u[4], v3 *o, v3 *d) {
	const float W = (float)c->width, H = (float)c->height;
	const float ps = pixel_size(W, H, c->boxing);
	const float pdx = ps / ((float)c->fsaa_x), pdy = ps / ((float)c->fsaa_y);
	const float ppx = ((float)px - (W / 2.0f)) * ps;
	const float ppy = ((H / 2.0f) - (float)py) * ps;
	const float sx = ppx + (((float)i) + 0.5f) * pdx;
	const float sy = ppy + (((float)j) + 0.5f) * pdy;
	float lx, ly, lz;
	if (c->lens == 2) {
		/* Lenses::SphericalSegmentLense::getPixelSamples (SphericalSegmentLense.impl.cuh:13-30),
		 * then Ray >>= Transform (Transform.impl.h:85-87, 166-168) */
		const float offx = pdx * 0.5f, offy = pdy * 0.5f;
		const float ha = ((sx + ((u[0] * (offx - (-offx))) + (-offx))) * c->lens_p[0]) / 57.29577951308233f;
		const float va = ((sy + ((u[1] * (offy - (-offy))) + (-offy))) * c->lens_p[1]) / 57.29577951308233f;
		float hs, hc, vs, vc, rs, rc;
		sincos_rad(ha, &hs, &hc);
		sincos_rad(va, &vs, &vc);
		const v3 target = vmul(V(hs * vc, vs, hc * vc), c->lens_p[2]);
		ora_sincos_2pi(u[2], &rs, &rc);
		const v3 org = vmul(vmul(V(rc, rs, 0.0f), ora_asin(u[3])), c->lens_p[3]);
		const v3 dl = vnormalized(vsub(target, org));
		const float *R = c->cam_rows, *C = c->cam;
		*o = V(((org.x * R[0] + org.y * R[1]) + org.z * R[2]) + C[0], ((org.x * R[3] + org.y * R[4]) + org.z * R[5]) + C[1],
		       ((org.x * R[6] + org.y * R[7]) + org.z * R[8]) + C[2]);
		*d = V((dl.x * C[3] + dl.y * C[6]) + dl.z * C[9], (dl.x * C[4] + dl.y * C[7]) + dl.z * C[10],
		       (dl.x * C[5] + dl.y * C[8]) + dl.z * C[11]);
		return;
	}
	if (c->lens == 1) {
		/* DefaultPerspectiveLense (DefaultPerspectiveLense.impl.cuh:12-15): the subpixel centre,
		 * direction normalized in lens space before the camera transform */
		const v3 n = vnormalized(V(sx, sy, c->cam[12]));
		lx = n.x; ly = n.y; lz = n.z;
	} else {
		const float offx = pdx * 0.5f, offy = pdy * 0.5f;
		lx = sx + ((u[0] * (offx - (-offx))) + (-offx));
		ly = sy + ((u[1] * (offy - (-offy))) + (-offy));
		lz = c->cam[12];
	}
	const float *P = c->cam, *R = c->cam + 3, *U = c->cam + 6, *F = c->cam + 9;
	*o = V(P[0], P[1], P[2]);
	*d = V((lx * R[0] + ly * U[0]) + lz * F[0], (lx * R[1] + ly * U[1]) + lz * F[1], (lx * R[2] + ly * U[2]) + lz * F[2]);
}

void ora_camera_ray(const ora_render_cfg *c, int px, int py, int i, int j, const float u[4], float o[3], float d[3]) {
	v3 ov, dv;
	camera_ray(c, px, py, i, j, u, &ov, &dv);
	o[0] = ov.x; o[1] = ov.y; o[2] = ov.z;
	d[0] = dv.x; d[1] = dv.y; d[2] = dv.z;
}

/* geometric normal of triangle h, unit, facing the incoming ray; also the unit ray direction */
static void hit_frame(const rctx *C, int32_t h, v3 d, v3 *nu, v3 *du) {
	const float *T = ora_octree_tri(C->tree, h);
	const v3 a = ld3(T), b = ld3(T + 3), c = ld3(T + 6);
	v3 n = vnormalized(vcross(vsub(b, a), vsub(c, a)));
	*du = vnormalized(d);
	if (vdot(n, *du) > 0) n = V(-n.x, -n.y, -n.z);
	*nu = n;
}

/* cosine-weighted direction about unit n (orthonormal basis of Duff et al., JCGT 2017) */
static v3 cosine_dir(v3 n, float u1, float u2) {
	const float sign = copysignf(1.0f, n.z);
	const float a = -1.0f / (sign + n.z);
	const float b = n.x * n.y * a;
	const v3 b1 = V(1.0f + ((sign * n.x) * n.x) * a, sign * b, -sign * n.x);
	const v3 b2 = V(b, sign + (n.y * n.y) * a, -n.y);
	const float r = sqrtf(u1);
	float s, c;
	ora_sincos_2pi(u2, &s, &c);
	const float x = r * c, y = r * s, z = sqrtf(1.0f - u1);
	return vadd(vadd(vmul(b1, x), vmul(b2, y)), vmul(n, z));
}

/* Vector3::reflectionOnUnitVector (Vector3.impl.h:393-396) */
static v3 reflect_unit(v3 v, v3 n) {
	const float mul = (2.0f * vdot(v, n));
	return V(v.x - (n.x * mul), v.y - (n.y * mul), v.z - (n.z * mul));
}

/* Sampling decision of the layer at `layer` (requestIndirectSamples + Russian roulette).
 * Returns 1 and the direction / weight when a sample is requested. */
static int next_sample(const rctx *C, const ora_material *M, uint32_t pix, uint32_t samp, int layer, v3 nu, v3 du,
                       v3 *dir, float *weight) {
	if (layer >= C->cfg->max_bounces) return 0;
	float u[4];
	rng4(C, pix, samp, (uint32_t)layer, 1, u);
	*weight = 1.0f;
	if (layer >= C->cfg->rr_start) {
		if (u[2] >= C->cfg->rr_survive) return 0;
		*weight = 1.0f / C->cfg->rr_survive;
	}
	if (M->kind == ORA_LAMBERT) *dir = cosine_dir(nu, u[0], u[1]);
	else *dir = reflect_unit(du, nu);
	return 1;
}

/* One light photon for the point p (AreaQuadLight::getVertexPhotons): sample q on the
 * quad, photon ray q -> p with colour E * cos_l * A / d^2.  Returns 0 if the photon
 * cannot light p (outside the emitting side, or p facing away); else the shadow ray
 * (q, w), the incidence cosine at p and the scalar geometry factor cos_l * A / d^2. */
static int light_photon(const rctx *C, int l, uint32_t pix, uint32_t samp, int layer, v3 p, v3 nu, v3 *q, v3 *w,
                        float *cos_i, float *gl) {
	const ora_light *L = &C->lights[l];
	float u[4];
	rng4(C, pix, samp, (uint32_t)layer, 16u + (uint32_t)l, u);
	const v3 cor = ld3(L->corner), eu = ld3(L->edge_u), ev = ld3(L->edge_v);
	*q = vadd(vadd(cor, vmul(eu, u[0])), vmul(ev, u[1]));
	const v3 delta = vsub(p, *q);
	const float d2 = vdot(delta, delta);
	if (!(d2 > 0.0f)) return 0;
	*w = vmul(delta, 1.0f / sqrtf(d2));
	const v3 nl = vcross(eu, ev);
	const float area = sqrtf(vdot(nl, nl));
	const v3 nlu = vmul(nl, 1.0f / area);
	const float cos_l = vdot(nlu, *w);
	if (!(cos_l > 0.0f)) return 0;
	*cos_i = -vdot(nu, *w);
	if (!(*cos_i > 0.0f)) return 0;
	*gl = (cos_l * area) / d2;
	return 1;
}

/* Shadow test (DumbRenderer.cu:386-395): the photon ray's first hit is the same
 * triangle or lands within sqrt(8 eps) of the shaded point. */
static int photon_lit(const rctx *C, v3 q, v3 w, int32_t h, v3 p) {
	const float o[3] = {q.x, q.y, q.z}, d[3] = {w.x, w.y, w.z};
	int32_t id; float t; float hp[3];
	if (!ora_cast1(C->tree, o, d, C->cfg->cull_backfaces, -1, &id, &t, hp)) return 0;
	if (id == h) return 1;
	const float dx = hp[0] - p.x, dy = hp[1] - p.y, dz = hp[2] - p.z;
	return (dx * dx + dy * dy + dz * dz) <= (8.0f * VEPS);
}

#define REAL float
#define FN(x) x##_f
#include "dr_render_body.inc"
#undef REAL
#undef FN
#define REAL double
#define FN(x) x##_d
#include "dr_render_body.inc"
#undef REAL
#undef FN

#include "dr_render_ref.inc"

/* ------------------------------------------------------------------ */
/* threading                                                           */
/* ------------------------------------------------------------------ */
typedef struct {
	const rctx *C; int y0, y1, ybase; float *out; double *out64; const float *dl; double *ga, *ge; int mode;
} rjob;
static void *render_worker(void *arg) {
	rjob *J = (rjob *)arg;
	const rctx *C = J->C;
	const ora_render_cfg *cfg = C->cfg;
	for (int py = J->y0; py < J->y1; py++)
		for (int px = 0; px < cfg->width; px++) {
			const size_t k = (size_t)(py - J->ybase) * cfg->width + px;
			if (J->mode == 0) pixel_color_f(C, px, py, J->out + 3 * k);
			else if (J->mode == 1) pixel_color_d(C, px, py, J->out64 + 3 * k);
			else if (J->mode == 3) pixel_color_ref(C, px, py, J->out + 3 * k);
			else if (J->dl[3 * k] != 0.0f || J->dl[3 * k + 1] != 0.0f || J->dl[3 * k + 2] != 0.0f)
				pixel_backward_d(C, px, py, J->dl + 3 * k, J->ga, J->ge);   /* dL/dI = 0 adds nothing */
		}
	return NULL;
}
static int run(const rctx *C, int mode, float *out, double *out64, const float *dl, double *ga, double *ge, int nth,
               int row0, int row1) {
	const int H = row1 - row0;
	if (nth <= 0) { long c = sysconf(_SC_NPROCESSORS_ONLN); nth = c > 0 ? (int)c : 1; }
	if (nth > H) nth = H > 0 ? H : 1;
	pthread_t *th = (pthread_t *)malloc(sizeof(pthread_t) * nth);
	rjob *jobs = (rjob *)calloc(nth, sizeof(rjob));
	double *ga_part = NULL, *ge_part = NULL;
	if (mode == 2) {
		ga_part = (double *)calloc((size_t)nth * C->nmat * 3 + 1, sizeof(double));
		ge_part = (double *)calloc((size_t)nth * C->nlights * 3 + 1, sizeof(double));
	}
	for (int i = 0; i < nth; i++) {
		jobs[i].C = C; jobs[i].mode = mode; jobs[i].out = out; jobs[i].out64 = out64; jobs[i].dl = dl;
		jobs[i].y0 = row0 + (int)(((long long)H * i) / nth); jobs[i].y1 = row0 + (int)(((long long)H * (i + 1)) / nth);
		jobs[i].ybase = row0;
		if (mode == 2) { jobs[i].ga = ga_part + (size_t)i * C->nmat * 3; jobs[i].ge = ge_part + (size_t)i * C->nlights * 3; }
		pthread_create(&th[i], NULL, render_worker, &jobs[i]);
	}
	for (int i = 0; i < nth; i++) pthread_join(th[i], NULL);
	if (mode == 2) {
		for (int i = 0; i < nth; i++) {
			for (int k = 0; k < C->nmat * 3; k++) ga[k] += ga_part[(size_t)i * C->nmat * 3 + k];
			for (int k = 0; k < C->nlights * 3; k++) ge[k] += ge_part[(size_t)i * C->nlights * 3 + k];
		}
		free(ga_part); free(ge_part);
	}
	free(th); free(jobs);
	return 0;
}

int ora_render(const ora_octree *t, const int32_t *mat_id, const ora_material *mats, int nmat, const ora_light *lights,
               int nlights, const ora_render_cfg *cfg, float *out_rgb, double *out_rgb64, int row0, int row1,
               int nthreads) {
	if (nlights > MAX_LIGHTS || cfg->fsaa_x < 1 || cfg->fsaa_y < 1) return 1;
	if (row0 < 0 || row1 > cfg->height || row0 > row1) return 1;
	rctx C = {t, mat_id, mats, nmat, lights, nlights, cfg, NULL, NULL};
	return run(&C, out_rgb64 ? 1 : 0, out_rgb, out_rgb64, NULL, NULL, NULL, nthreads, row0, row1);
}

int ora_backward(const ora_octree *t, const int32_t *mat_id, const ora_material *mats, int nmat,
                 const ora_light *lights, int nlights, const ora_render_cfg *cfg, const float *dl_dimage,
                 double *g_albedo, double *g_emission, int nthreads) {
	if (nlights > MAX_LIGHTS || cfg->fsaa_x < 1 || cfg->fsaa_y < 1) return 1;
	rctx C = {t, mat_id, mats, nmat, lights, nlights, cfg, NULL, NULL};
	memset(g_albedo, 0, sizeof(double) * nmat * 3);
	memset(g_emission, 0, sizeof(double) * nlights * 3);
	return run(&C, 2, NULL, NULL, dl_dimage, g_albedo, g_emission, nthreads, 0, cfg->height);
}

int ora_render_ref(const ora_octree *t, const int32_t *mat_id, const ora_material *mats, int nmat,
                   const ora_light *lights, int nlights, const float *normals, const ora_render_cfg *cfg,
                   float *out_rgb, int row0, int row1, int nthreads) {
	if (nlights > MAX_LIGHTS || cfg->fsaa_x < 1 || cfg->fsaa_y < 1) return 1;
	if (row0 < 0 || row1 > cfg->height || row0 > row1) return 1;
	rctx C = {t, mat_id, mats, nmat, lights, nlights, cfg, normals, NULL};
	return run(&C, 3, out_rgb, NULL, NULL, NULL, NULL, nthreads, row0, row1);
}

int ora_render_ref_tex(const ora_octree *t, const int32_t *mat_id, const ora_material *mats, int nmat,
                       const ora_light *lights, int nlights, const float *normals, const ora_textures *tx,
                       const ora_render_cfg *cfg, float *out_rgb, int row0, int row1, int nthreads) {
	if (nlights > MAX_LIGHTS || cfg->fsaa_x < 1 || cfg->fsaa_y < 1) return 1;
	if (row0 < 0 || row1 > cfg->height || row0 > row1) return 1;
	rctx C = {t, mat_id, mats, nmat, lights, nlights, cfg, normals, tx};
	return run(&C, 3, out_rgb, NULL, NULL, NULL, NULL, nthreads, row0, row1);
}
