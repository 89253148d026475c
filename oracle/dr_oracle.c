/*
 * CPU oracle: geometry, octree build and closest-hit cast.
 * TEST INFRASTRUCTURE ONLY (see dr_oracle.h).  Built with -ffp-contract=off
 * and without -ffast-math so every fp32 operation is a single IEEE-754
 * correctly rounded operation evaluated in source order, as in the reference
 * CPU build (x86-64 SSE, no FMA).
 *
 * min/max: the reference calls unqualified min/max on the host
 * (AABB.impl.h:119-127, Octree.impl.cuh:104-105, 509-528).  This oracle pins
 * std::min/std::max semantics: min(a,b) = (b < a) ? b : a, max(a,b) = (a < b) ? b : a.
 */
#include "dr_oracle.h"
#include <float.h>
#include <math.h>
#include <pthread.h>
#include <stdlib.h>
#include <string.h>
#include <unistd.h>

#define VEPS 0.000005f                      /* VECTOR_EPSILON, Vector3.h:10 */

typedef struct { float x, y, z; } v3;
typedef struct { v3 a, b, c; } tri_t;
typedef struct { v3 s, e; } aabb_t;

/* ------------------------------------------------------------------ */
/* Vector3 (Vector3.impl.h:9-33 macros, 84-200 operators)              */
/* ------------------------------------------------------------------ */
static inline v3 V(float x, float y, float z) { v3 r; r.x = x; r.y = y; r.z = z; return r; }
static inline v3 vadd(v3 a, v3 b) { return V(a.x + b.x, a.y + b.y, a.z + b.z); }
static inline v3 vsub(v3 a, v3 b) { return V(a.x - b.x, a.y - b.y, a.z - b.z); }
static inline v3 vmul(v3 a, float f) { return V(a.x * f, a.y * f, a.z * f); }
/* operator/(float): fi = 1/f; multiply (Vector3.impl.h:170-173) */
static inline v3 vdivf(v3 a, float f) { float fi = 1.0f / f; return V(a.x * fi, a.y * fi, a.z * fi); }
static inline float vdot(v3 a, v3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
/* operator& (CrossProductWith, Vector3.impl.h:19-22) */
static inline v3 vcross(v3 a, v3 b) {
	return V(a.y * b.z - b.y * a.z, b.x * a.z - a.x * b.z, a.x * b.y - b.x * a.y);
}
static inline float vsqr(v3 a) { return a.x * a.x + a.y * a.y + a.z * a.z; }
static inline float vmag(v3 a) { return sqrtf(a.x * a.x + a.y * a.y + a.z * a.z); }
static inline int veq(v3 a, v3 b) { return (a.x == b.x) && (a.y == b.y) && (a.z == b.z); }
/* NormalOn(v) = v*(dot(this,v)/|v|^2)  (Vector3.impl.h:31) */
static inline v3 vnormal_on(v3 t, v3 v) { return vmul(v, vdot(t, v) / vsqr(v)); }
static inline float fmin_std(float a, float b) { return (b < a) ? b : a; }
static inline float fmax_std(float a, float b) { return (a < b) ? b : a; }
#define SMIN(a, b) (((a) < (b)) ? (a) : (b))   /* SHAPES_MIN, Shapes.impl.cuh:200 */
#define SMAX(a, b) (((a) > (b)) ? (a) : (b))   /* SHAPES_MAX, Shapes.impl.cuh:201 */

static inline v3 ld3(const float *p) { return V(p[0], p[1], p[2]); }
static inline tri_t ldtri(const float *p) { tri_t t; t.a = ld3(p); t.b = ld3(p + 3); t.c = ld3(p + 6); return t; }
static inline tri_t T(v3 a, v3 b, v3 c) { tri_t t; t.a = a; t.b = b; t.c = c; return t; }

/* ------------------------------------------------------------------ */
/* AABB (AABB.impl.h)                                                  */
/* ------------------------------------------------------------------ */
/* AABB(s,e) + fix() (AABB.impl.h:12-16, 141-146) */
static inline aabb_t box_make(v3 s, v3 e) {
	aabb_t b; float tmp;
	if (s.x > e.x) { tmp = s.x; s.x = e.x; e.x = tmp; }
	if (s.y > e.y) { tmp = s.y; s.y = e.y; e.y = tmp; }
	if (s.z > e.z) { tmp = s.z; s.z = e.z; e.z = tmp; }
	b.s = s; b.e = e; return b;
}
/* contains (AABB.impl.h:93-95) */
static inline int box_contains(aabb_t b, v3 v) {
	return v.x >= b.s.x && v.y >= b.s.y && v.z >= b.s.z && v.x <= b.e.x && v.y <= b.e.y && v.z <= b.e.z;
}
/* castPreInversed (AABB.impl.h:116-130) */
static inline float box_cast_preinv(aabb_t b, v3 o, v3 inv) {
	float ds = (b.s.x - o.x) * inv.x;
	float de = (b.e.x - o.x) * inv.x;
	float mn = fmin_std(ds, de), mx = fmax_std(ds, de);
	ds = (b.s.y - o.y) * inv.y;
	de = (b.e.y - o.y) * inv.y;
	mn = fmax_std(mn, fmin_std(ds, de));
	mx = fmin_std(mx, fmax_std(ds, de));
	ds = (b.s.z - o.z) * inv.z;
	de = (b.e.z - o.z) * inv.z;
	mn = fmax_std(mn, fmin_std(ds, de));
	mx = fmin_std(mx, fmax_std(ds, de));
	if (mn > mx + VEPS) return FLT_MAX;
	return mn;
}
/* Shapes::castPreInversed<AABB>(ray, box, clip=false) (Shapes.impl.cuh:144-147) */
static inline int box_hit(aabb_t b, v3 o, v3 inv) { return box_cast_preinv(b, o, inv) != FLT_MAX; }

/* Triangle::sortByMases (Triangle.impl.h:185-199) */
static inline void tri_sort(tri_t *t, float am, float bm, float cm) {
	v3 tmp;
	if (am > bm) {
		if (cm > am) { tmp = t->a; t->a = t->b; t->b = tmp; }
		else {
			if (bm > cm) { tmp = t->a; t->a = t->c; t->c = tmp; }
			else { tmp = t->a; t->a = t->b; t->b = t->c; t->c = tmp; }
		}
	} else {
		if (am > cm) { tmp = t->a; t->a = t->c; t->c = t->b; t->b = tmp; }
		else if (bm > cm) { tmp = t->b; t->b = t->c; t->c = tmp; }
	}
}
/* CROSS_POINT (AABB.impl.h:168, Shapes.impl.cuh:202) */
static inline v3 cross_point(v3 from, v3 to, float fromV, float toV, float barrier) {
	return vadd(from, vmul(vsub(to, from), (barrier - fromV) / (toV - fromV)));
}
static int box_isect_tri_dim(aabb_t b, tri_t t, int dimm);
/* intersectsTri<dimm>(Triangle) (AABB.impl.h:150-166): dimm 0 = z, 1 = x, 2 = y, 3 = true */
static int box_isect_tri_dim(aabb_t b, tri_t t, int dimm) {
	float av, bv, cv, s, e;
	if (dimm >= 3) return 1;
	if (dimm == 0) { tri_sort(&t, t.a.z, t.b.z, t.c.z); av = t.a.z; bv = t.b.z; cv = t.c.z; s = b.s.z; e = b.e.z; }
	else if (dimm == 1) { tri_sort(&t, t.a.x, t.b.x, t.c.x); av = t.a.x; bv = t.b.x; cv = t.c.x; s = b.s.x; e = b.e.x; }
	else { tri_sort(&t, t.a.y, t.b.y, t.c.y); av = t.a.y; bv = t.b.y; cv = t.c.y; s = b.s.y; e = b.e.y; }
	/* AABB.impl.h:169-221 */
	if (cv < s) return 0;
	if (av > e) return 0;
	const int nd = dimm + 1;
	if (av <= s) {
		v3 asc = cross_point(t.a, t.c, av, cv, s);
		if (bv <= s) {
			v3 bsc = cross_point(t.b, t.c, bv, cv, s);
			if (cv <= e) return box_isect_tri_dim(b, T(asc, bsc, t.c), nd);
			else {
				v3 bec = cross_point(t.b, t.c, bv, cv, e);
				if (box_isect_tri_dim(b, T(bsc, bec, asc), nd)) return 1;
				v3 aec = cross_point(t.a, t.c, av, cv, e);
				return box_isect_tri_dim(b, T(asc, bec, aec), nd);
			}
		} else if (bv <= e) {
			if (cv <= e) {
				if (box_isect_tri_dim(b, T(asc, t.b, t.c), nd)) return 1;
				v3 asb = cross_point(t.a, t.b, av, bv, s);
				return box_isect_tri_dim(b, T(asc, asb, t.b), nd);
			} else {
				v3 asb = cross_point(t.a, t.b, av, bv, s);
				v3 bec = cross_point(t.b, t.c, bv, cv, e);
				if (box_isect_tri_dim(b, T(asb, t.b, bec), nd)) return 1;
				if (box_isect_tri_dim(b, T(asc, asb, bec), nd)) return 1;
				v3 aec = cross_point(t.a, t.c, av, cv, e);
				return box_isect_tri_dim(b, T(asc, bec, aec), nd);
			}
		} else {
			v3 asb = cross_point(t.a, t.b, av, bv, s);
			v3 aeb = cross_point(t.a, t.b, av, bv, e);
			if (box_isect_tri_dim(b, T(asc, asb, aeb), nd)) return 1;
			v3 aec = cross_point(t.a, t.c, av, cv, e);
			return box_isect_tri_dim(b, T(asc, aeb, aec), nd);
		}
	} else {
		if (cv <= e) return box_isect_tri_dim(b, t, nd);
		else {
			v3 aec = cross_point(t.a, t.c, av, cv, e);
			if (bv <= e) {
				v3 bec = cross_point(t.b, t.c, bv, cv, e);
				if (box_isect_tri_dim(b, T(t.a, t.b, bec), nd)) return 1;
				return box_isect_tri_dim(b, T(t.a, aec, bec), nd);
			} else {
				v3 aeb = cross_point(t.a, t.b, av, bv, e);
				return box_isect_tri_dim(b, T(t.a, aeb, aec), nd);
			}
		}
	}
}

/* ------------------------------------------------------------------ */
/* Triangle (Triangle.impl.h)                                          */
/* ------------------------------------------------------------------ */
/* containsVertex, active branch (Triangle.impl.h:292-302) */
static inline int tri_contains(const tri_t *t, v3 p) {
	if (veq(t->a, t->b) || veq(t->b, t->c) || veq(t->c, t->a)) return 0;
	if (veq(p, t->a) || veq(p, t->b) || veq(p, t->c)) return 1;
	v3 ab = vsub(t->b, t->a);
	v3 bc = vsub(t->c, t->b);
	v3 ca = vsub(t->a, t->c);
	v3 ax = vsub(p, t->a);
	v3 bx = vsub(p, t->b);
	v3 cx = vsub(p, t->c);
	const float tol = 32 * VEPS;
	return (vdot(ab, ax) / vmag(ax) + tol >= -vdot(ab, ca) / vmag(ca))
		&& (vdot(bc, bx) / vmag(bx) + tol >= -vdot(bc, ab) / vmag(ab))
		&& (vdot(ca, cx) / vmag(cx) + tol >= -vdot(ca, bc) / vmag(bc));
}
/* Triangle::cast (Triangle.impl.h:334-348) */
static inline int tri_cast(const tri_t *t, v3 o, v3 d, int clip, float *dist, v3 *hp) {
	v3 n = vcross(vsub(t->b, t->a), vsub(t->c, t->a));
	float dp = vdot(vsub(t->a, o), n);
	if (clip && (dp < -VEPS)) return 0;
	float dd = vdot(d, n);
	if (dp * dd <= 0) return 0;
	float distance = dp / dd;
	v3 h = vadd(o, vmul(d, distance));
	if (tri_contains(t, h)) { *dist = distance; *hp = h; return 1; }
	return 0;
}
/* surface (Triangle.impl.h:206-208) */
static inline float tri_surface(tri_t t) { return vmag(vcross(vsub(t.b, t.a), vsub(t.c, t.a))) * 0.5f; }
/* massCenter() = (a+b+c)/3 (Triangle.impl.h:326-328) */
static inline v3 tri_center(tri_t t) { return vdivf(vadd(vadd(t.a, t.b), t.c), 3.0f); }

/* ------------------------------------------------------------------ */
/* Shapes (Shapes.impl.cuh:202-312)                                    */
/* ------------------------------------------------------------------ */
typedef struct { tri_t *v; int n, cap; } tlist;
static void tl_push(tlist *l, tri_t t) {
	if (l->n == l->cap) { l->cap = l->cap ? l->cap * 2 : 32; l->v = (tri_t *)realloc(l->v, sizeof(tri_t) * l->cap); }
	l->v[l->n++] = t;
}
/* SHAPES_PUSH_INTERSECTIONS for one axis (Shapes.impl.cuh:203-253) */
static void clip_axis(tlist *front, tlist *back, int axis, float s, float e) {
	for (int i = 0; i < front->n; i++) {
		tri_t *t = &front->v[i];
		float av, bv, cv;
#define AXV(p) (axis == 0 ? (p).x : (axis == 1 ? (p).y : (p).z))
		tri_sort(t, AXV(t->a), AXV(t->b), AXV(t->c));
		av = AXV(t->a); bv = AXV(t->b); cv = AXV(t->c);
#undef AXV
		if (cv < s) continue;
		else if (av > e) continue;
		else if (av <= s) {
			v3 asc = cross_point(t->a, t->c, av, cv, s);
			if (bv <= s) {
				v3 bsc = cross_point(t->b, t->c, bv, cv, s);
				if (cv <= e) tl_push(back, T(asc, bsc, t->c));
				else {
					v3 bec = cross_point(t->b, t->c, bv, cv, e);
					tl_push(back, T(bsc, bec, asc));
					v3 aec = cross_point(t->a, t->c, av, cv, e);
					tl_push(back, T(asc, bec, aec));
				}
			} else if (bv <= e) {
				if (cv <= e) {
					tl_push(back, T(asc, t->b, t->c));
					v3 asb = cross_point(t->a, t->b, av, bv, s);
					tl_push(back, T(asc, asb, t->b));
				} else {
					v3 asb = cross_point(t->a, t->b, av, bv, s);
					v3 bec = cross_point(t->b, t->c, bv, cv, e);
					tl_push(back, T(asb, t->b, bec));
					tl_push(back, T(asc, asb, bec));
					v3 aec = cross_point(t->a, t->c, av, cv, e);
					tl_push(back, T(asc, bec, aec));
				}
			} else {
				v3 asb = cross_point(t->a, t->b, av, bv, s);
				v3 aeb = cross_point(t->a, t->b, av, bv, e);
				tl_push(back, T(asc, asb, aeb));
				v3 aec = cross_point(t->a, t->c, av, cv, e);
				tl_push(back, T(asc, aeb, aec));
			}
		} else {
			if (cv <= e) tl_push(back, *t);
			else {
				v3 aec = cross_point(t->a, t->c, av, cv, e);
				if (bv <= e) {
					v3 bec = cross_point(t->b, t->c, bv, cv, e);
					tl_push(back, T(t->a, t->b, bec));
					tl_push(back, T(t->a, aec, bec));
				} else {
					v3 aeb = cross_point(t->a, t->b, av, bv, e);
					tl_push(back, T(t->a, aeb, aec));
				}
			}
		}
	}
	front->n = 0;
}
/* SHAPES_GET_INTERSECTING_TRIANGLES (Shapes.impl.cuh:255-266): x, then y, then z */
static void clip_to_box(aabb_t b, tri_t t, tlist *out, tlist *tmp) {
	out->n = 0; tmp->n = 0;
	tl_push(tmp, t);
	clip_axis(tmp, out, 0, b.s.x, b.e.x);
	clip_axis(out, tmp, 1, b.s.y, b.e.y);
	clip_axis(tmp, out, 2, b.s.z, b.e.z);
}
/* intersectionBounds<AABB,Triangle> (Shapes.impl.cuh:272-291) */
static aabb_t isect_bounds(aabb_t b, tri_t t, tlist *l, tlist *tmp) {
	clip_to_box(b, t, l, tmp);
	if (l->n > 0) {
		v3 mn = b.e, mx = b.s;
		for (int i = 0; i < l->n; i++) {
			const tri_t *q = &l->v[i];
			v3 nmx = V(SMAX(SMAX(SMAX(mx.x, q->a.x), q->b.x), q->c.x),
				SMAX(SMAX(SMAX(mx.y, q->a.y), q->b.y), q->c.y),
				SMAX(SMAX(SMAX(mx.z, q->a.z), q->b.z), q->c.z));
			v3 nmn = V(SMIN(SMIN(SMIN(mn.x, q->a.x), q->b.x), q->c.x),
				SMIN(SMIN(SMIN(mn.y, q->a.y), q->b.y), q->c.y),
				SMIN(SMIN(SMIN(mn.z, q->a.z), q->b.z), q->c.z));
			mx = nmx; mn = nmn;
		}
		return box_make(mn, mx);
	}
	return b;
}
/* intersectionCenter<AABB,Triangle> (Shapes.impl.cuh:293-312) */
static v3 isect_center(aabb_t b, tri_t t, tlist *l, tlist *tmp) {
	clip_to_box(b, t, l, tmp);
	if (l->n > 0) {
		float size = 0.0f;
		v3 sum = V(0.0f, 0.0f, 0.0f);
		for (int i = 0; i < l->n; i++) {
			float surface = (tri_surface(l->v[i]) + 0.0001f);
			sum = vadd(sum, vmul(tri_center(l->v[i]), surface));
			size += surface;
		}
		return vdivf(sum, size);
	}
	return vdivf(vadd(b.s, b.e), 2.0f);
}
/* boundingBox<Triangle> (Shapes.impl.cuh:371-382) */
static aabb_t tri_bbox(tri_t t) {
	v3 s = V(SMIN(SMIN(t.a.x, t.b.x), t.c.x), SMIN(SMIN(t.a.y, t.b.y), t.c.y), SMIN(SMIN(t.a.z, t.b.z), t.c.z));
	v3 e = V(SMAX(SMAX(t.a.x, t.b.x), t.c.x), SMAX(SMAX(t.a.y, t.b.y), t.c.y), SMAX(SMAX(t.a.z, t.b.z), t.c.z));
	return box_make(s, e);
}

/* ------------------------------------------------------------------ */
/* Octree (Octree.impl.cuh)                                            */
/* ------------------------------------------------------------------ */
typedef struct { int32_t *v; int n, cap; } ilist;
static void il_push(ilist *l, int32_t x) {
	if (l->n == l->cap) { l->cap = l->cap ? l->cap * 2 : 16; l->v = (int32_t *)realloc(l->v, sizeof(int32_t) * l->cap); }
	l->v[l->n++] = x;
}
struct ora_octree {
	aabb_t *bounds; int32_t *children; ilist *lists;
	int n, cap;
	tri_t *tris; int ntris;
	int thresh, maxdepth;
	float E;       /* OCTREE_AABB_EPSILON = 256 * VECTOR_EPSILON (Octree.impl.cuh:3-5) */
	v3 Ev;         /* OCTREE_AABB_EPSILON_VECTOR = EPSILON_VECTOR * 256 */
	tlist l, tmp;  /* scratch for clipping */
};
static int oct_add_node(ora_octree *t) {
	if (t->n == t->cap) {
		t->cap = t->cap ? t->cap * 2 : 64;
		t->bounds = (aabb_t *)realloc(t->bounds, sizeof(aabb_t) * t->cap);
		t->children = (int32_t *)realloc(t->children, sizeof(int32_t) * t->cap);
		t->lists = (ilist *)realloc(t->lists, sizeof(ilist) * t->cap);
	}
	int i = t->n++;
	t->children[i] = -1;
	memset(&t->lists[i], 0, sizeof(ilist));
	return i;
}
/* splitAABB (Octree.impl.cuh:423-434); bit0 = z-high, bit1 = y-high, bit2 = x-high */
static void oct_split_aabb(const ora_octree *t, aabb_t b, v3 c, aabb_t *r) {
	const v3 s = b.s, e = b.e; const float E = t->E;
	r[0] = box_make(s, vadd(c, t->Ev));
	r[1] = box_make(V(s.x, s.y, c.z - E), V(c.x + E, c.y + E, e.z));
	r[2] = box_make(V(s.x, c.y - E, s.z), V(c.x + E, e.y, c.z + E));
	r[3] = box_make(V(s.x, c.y - E, c.z - E), V(c.x + E, e.y, e.z));
	r[4] = box_make(V(c.x - E, s.y, s.z), V(e.x, c.y + E, c.z + E));
	r[5] = box_make(V(c.x - E, s.y, c.z - E), V(e.x, c.y + E, e.z));
	r[6] = box_make(V(c.x - E, c.y - E, s.z), V(e.x, e.y, c.z + E));
	r[7] = box_make(vsub(c, t->Ev), e);
}
/* splittingMakesSence with the intersection cache (Octree.impl.cuh:437-475) */
static int oct_split_makes_sense(ora_octree *t, int idx, const aabb_t *sub, uint8_t *cache) {
	const int nodeCount = t->lists[idx].n;
	for (int i = 0; i < nodeCount; i++) cache[i] = 0;
	int full = 0, empty = 0;
	float load = 0.0f;
	for (int i = 0; i < 8; i++) {
		int insiders = 0;
		for (int j = 0; j < nodeCount; j++)
			if (box_isect_tri_dim(sub[i], t->tris[t->lists[idx].v[j]], 0)) { insiders++; cache[j] |= (uint8_t)(1 << i); }
		if (insiders >= nodeCount) full++;
		else if (insiders == 0) empty++;
		load += ((float)insiders) / ((float)nodeCount);
	}
	(void)empty;
	if (full >= 1) return 0;
	return load < fmin_std(((float)nodeCount) / ((float)t->thresh), 6.0f);
}
/* reduceNode (Octree.impl.cuh:489-533) */
static void oct_reduce(ora_octree *t, int idx, int reduceLeaves) {
	v3 start = t->bounds[idx].e, end = t->bounds[idx].s;
	if (t->children[idx] >= 0) {
		const int c0 = t->children[idx];
		for (int i = 0; i < 8; i++) oct_reduce(t, c0 + i, reduceLeaves);
		for (int i = 0; i < 8; i++) {
			aabb_t b = t->bounds[c0 + i];
			start = V(fmin_std(start.x, b.s.x), fmin_std(start.y, b.s.y), fmin_std(start.z, b.s.z));
			end = V(fmax_std(end.x, b.e.x), fmax_std(end.y, b.e.y), fmax_std(end.z, b.e.z));
		}
	} else {
		if (!reduceLeaves) return;
		for (int i = 0; i < t->lists[idx].n; i++) {
			aabb_t ob = isect_bounds(t->bounds[idx], t->tris[t->lists[idx].v[i]], &t->l, &t->tmp);
			start = V(fmin_std(start.x, ob.s.x), fmin_std(start.y, ob.s.y), fmin_std(start.z, ob.s.z));
			end = V(fmax_std(end.x, ob.e.x), fmax_std(end.y, ob.e.y), fmax_std(end.z, ob.e.z));
		}
	}
	start = vsub(start, t->Ev);
	end = vadd(end, t->Ev);
	const v3 bs = t->bounds[idx].s, be = t->bounds[idx].e;
	start = V(fmax_std(start.x, bs.x), fmax_std(start.y, bs.y), fmax_std(start.z, bs.z));
	end = V(fmin_std(end.x, be.x), fmin_std(end.y, be.y), fmin_std(end.z, be.z));
	t->bounds[idx] = box_make(start, end);
}
/* split (Octree.impl.cuh:339-421), depth-first, single-threaded */
static void oct_split(ora_octree *t, int idx, int depth) {
	if (depth >= t->maxdepth) { oct_reduce(t, idx, 1); return; }
	const int n = t->lists[idx].n;
	if (n < t->thresh) { oct_reduce(t, idx, 1); return; }
	v3 center = V(0, 0, 0);
	const aabb_t bounds = t->bounds[idx];
	for (int i = 0; i < n; i++)
		center = vadd(center, isect_center(bounds, t->tris[t->lists[idx].v[i]], &t->l, &t->tmp));
	center = vdivf(center, (float)n);
	aabb_t sub[8];
	oct_split_aabb(t, bounds, center, sub);
	uint8_t *cache = (uint8_t *)malloc((size_t)n);
	if (!oct_split_makes_sense(t, idx, sub, cache)) { free(cache); oct_reduce(t, idx, 1); return; }
	/* splitNode (Octree.impl.cuh:477-486) */
	int c0 = -1;
	for (int i = 0; i < 8; i++) { int k = oct_add_node(t); if (i == 0) c0 = k; t->bounds[k] = sub[i]; }
	t->children[idx] = c0;
	for (int i = 0; i < n; i++) {
		const int32_t e = t->lists[idx].v[i];
		for (int j = 0; j < 8; j++) if (cache[i] & (1 << j)) il_push(&t->lists[c0 + j], e);
	}
	free(cache);
	free(t->lists[idx].v); memset(&t->lists[idx], 0, sizeof(ilist));
	for (int i = 0; i < 8; i++) oct_split(t, c0 + i, depth + 1);
}

ora_octree *ora_octree_build(const float *tris, int ntris, int split_threshold, int max_depth) {
	ora_octree *t = (ora_octree *)calloc(1, sizeof(ora_octree));
	t->thresh = split_threshold; t->maxdepth = max_depth;
	t->E = 256 * VEPS;
	t->Ev = vmul(V(VEPS, VEPS, VEPS), 256);
	t->ntris = ntris;
	t->tris = (tri_t *)malloc(sizeof(tri_t) * (size_t)(ntris > 0 ? ntris : 1));
	for (int i = 0; i < ntris; i++) t->tris[i] = ldtri(tris + 9 * (size_t)i);
	/* reset(): one root node (Octree.impl.cuh:80-87) */
	oct_add_node(t);
	t->bounds[0] = box_make(V(0, 0, 0), V(0, 0, 0));
	/* push (Octree.impl.cuh:97-111) */
	for (int i = 0; i < ntris; i++) {
		aabb_t box = tri_bbox(t->tris[i]);
		if (i == 0) t->bounds[0] = box;
		else {
			const v3 os = t->bounds[0].s, oe = t->bounds[0].e;
			const v3 bs = vsub(box.s, t->Ev), be = vadd(box.e, t->Ev);
			v3 ns = V(fmin_std(os.x, bs.x), fmin_std(os.y, bs.y), fmin_std(os.z, bs.z));
			v3 ne = V(fmax_std(oe.x, be.x), fmax_std(oe.y, be.y), fmax_std(oe.z, be.z));
			t->bounds[0] = box_make(ns, ne);
		}
		il_push(&t->lists[0], i);
	}
	/* build (Octree.impl.cuh:114-129) */
	oct_split(t, 0, 0);
	oct_reduce(t, 0, 0);
	return t;
}
void ora_octree_free(ora_octree *t) {
	if (!t) return;
	for (int i = 0; i < t->n; i++) free(t->lists[i].v);
	free(t->lists); free(t->bounds); free(t->children); free(t->tris);
	free(t->l.v); free(t->tmp.v); free(t);
}
int ora_octree_node_count(const ora_octree *t) { return t->n; }
int ora_octree_list_total(const ora_octree *t) {
	int s = 0; for (int i = 0; i < t->n; i++) s += t->lists[i].n; return s;
}
void ora_octree_export(const ora_octree *t, float *bounds, int32_t *children, int32_t *list_off, int32_t *list) {
	int off = 0;
	for (int i = 0; i < t->n; i++) {
		const aabb_t b = t->bounds[i];
		bounds[6 * i + 0] = b.s.x; bounds[6 * i + 1] = b.s.y; bounds[6 * i + 2] = b.s.z;
		bounds[6 * i + 3] = b.e.x; bounds[6 * i + 4] = b.e.y; bounds[6 * i + 5] = b.e.z;
		children[i] = t->children[i];
		list_off[i] = off;
		for (int j = 0; j < t->lists[i].n; j++) list[off++] = t->lists[i].v[j];
	}
	list_off[t->n] = off;
}

/* castInLeaf (Octree.impl.cuh:584-612) */
static int oct_cast_leaf(const ora_octree *t, v3 o, v3 d, int idx, int clip, int32_t exclude,
                         int32_t *id, float *dist, v3 *hp, int32_t *nref, int32_t *ntri) {
	const ilist *l = &t->lists[idx];
	if (l->n <= 0) return 0;
	float best = FLT_MAX; v3 bestp = V(0, 0, 0); int32_t besti = -1;
	for (int i = 0; i < l->n; i++) {
		const int32_t e = l->v[i];
		v3 p; float dd;
		(*nref)++;
		const tri_t *tr = &t->tris[e];
		if (!(veq(tr->a, tr->b) || veq(tr->b, tr->c) || veq(tr->c, tr->a))) (*ntri)++;
		int casted = tri_cast(tr, o, d, clip, &dd, &p);
		if (casted && exclude >= 0) casted = (e != exclude);   /* validateNotSameAsObject (:163) */
		if (casted && dd < best) { best = dd; bestp = p; besti = e; }
	}
	if (best != FLT_MAX && box_contains(t->bounds[idx], bestp)) { *id = besti; *dist = best; *hp = bestp; return 1; }
	return 0;
}
/* Octree::cast (Octree.impl.cuh:174-212) */
static int oct_cast(const ora_octree *t, v3 o, v3 d, int clip, int32_t exclude,
                    int32_t *id, float *dist, v3 *hp, int32_t *nbox, int32_t *nref, int32_t *ntri) {
	const v3 inv = V(1.0f / d.x, 1.0f / d.y, 1.0f / d.z);
	if (t->children[0] < 0) return oct_cast_leaf(t, o, d, 0, clip, exclude, id, dist, hp, nref, ntri);
	(*nbox)++;
	if (!box_hit(t->bounds[0], o, inv)) return 0;
	struct { int32_t node; int cur; } stack[64];
	int canonical = (d.z < 0) ? 1 : 0;
	if (d.y < 0) canonical += 2;
	if (d.x < 0) canonical += 4;
	stack[0].node = t->children[0]; stack[0].cur = 0;
	int i = 0; int32_t child = -1;
	for (;;) {
		while (stack[i].cur < 8) {
			child = stack[i].node + (canonical ^ stack[i].cur);
			(*nbox)++;
			if (box_hit(t->bounds[child], o, inv)) break;
			stack[i].cur++;
		}
		if (stack[i].cur >= 8) {
			if (i == 0) return 0;
			i--;
		} else {
			const int32_t ch = t->children[child];
			if (ch < 0) {
				if (oct_cast_leaf(t, o, d, child, clip, exclude, id, dist, hp, nref, ntri)) return 1;
				stack[i].cur++;
			} else {
				stack[i].cur++;   /* the reference increments the parent frame after pushing */
				i++;
				stack[i].node = ch; stack[i].cur = 0;
			}
		}
	}
}

/* ------------------------------------------------------------------ */
/* batch drivers                                                       */
/* ------------------------------------------------------------------ */
typedef struct {
	const ora_octree *t; const float *tris; int ntris;
	const float *rays; int lo, hi; int clip; const int32_t *exclude;
	int32_t *id; float *dist; float *point; float *second; int32_t *nb, *nr, *nt;
} job_t;
static void *cast_worker(void *arg) {
	job_t *j = (job_t *)arg;
	for (int r = j->lo; r < j->hi; r++) {
		const float *R = j->rays + 6 * (size_t)r;
		int32_t id = -1, nb = 0, nr = 0, nt = 0; float dist = FLT_MAX; v3 hp = V(0, 0, 0);
		int32_t ex = j->exclude ? j->exclude[r] : -1;
		if (!oct_cast(j->t, ld3(R), ld3(R + 3), j->clip, ex, &id, &dist, &hp, &nb, &nr, &nt)) { id = -1; dist = FLT_MAX; }
		j->id[r] = id; j->dist[r] = dist;
		if (j->point) { j->point[3 * r] = hp.x; j->point[3 * r + 1] = hp.y; j->point[3 * r + 2] = hp.z; }
		if (j->nb) j->nb[r] = nb;
		if (j->nr) j->nr[r] = nr;
		if (j->nt) j->nt[r] = nt;
	}
	return NULL;
}
static void *brute_worker(void *arg) {
	job_t *j = (job_t *)arg;
	for (int r = j->lo; r < j->hi; r++) {
		const float *R = j->rays + 6 * (size_t)r;
		const v3 o = ld3(R), d = ld3(R + 3);
		float best = FLT_MAX, second = FLT_MAX; int32_t bi = -1;
		const int32_t ex = j->exclude ? j->exclude[r] : -1;
		for (int k = 0; k < j->ntris; k++) {
			tri_t tr = ldtri(j->tris + 9 * (size_t)k);
			float dd; v3 p;
			if (k != ex && tri_cast(&tr, o, d, j->clip, &dd, &p)) {
				if (dd < best) { second = best; best = dd; bi = k; }
				else if (dd < second) second = dd;
			}
		}
		j->id[r] = bi; j->dist[r] = best;
		if (j->second) j->second[r] = second;
	}
	return NULL;
}
static void run_jobs(job_t *proto, int n, int nthreads, void *(*fn)(void *)) {
	if (nthreads <= 0) { long c = sysconf(_SC_NPROCESSORS_ONLN); nthreads = c > 0 ? (int)c : 1; }
	if (nthreads > n) nthreads = n > 0 ? n : 1;
	pthread_t *th = (pthread_t *)malloc(sizeof(pthread_t) * nthreads);
	job_t *jobs = (job_t *)malloc(sizeof(job_t) * nthreads);
	for (int i = 0; i < nthreads; i++) {
		jobs[i] = *proto;
		jobs[i].lo = (int)(((long long)n * i) / nthreads);
		jobs[i].hi = (int)(((long long)n * (i + 1)) / nthreads);
		pthread_create(&th[i], NULL, fn, &jobs[i]);
	}
	for (int i = 0; i < nthreads; i++) pthread_join(th[i], NULL);
	free(th); free(jobs);
}
void ora_octree_cast(const ora_octree *t, const float *rays, int n, int clip, const int32_t *exclude,
                     int32_t *id, float *dist, float *point, int32_t *n_box, int32_t *n_leafref,
                     int32_t *n_tri, int nthreads) {
	job_t p; memset(&p, 0, sizeof(p));
	p.t = t; p.rays = rays; p.clip = clip; p.exclude = exclude; p.id = id; p.dist = dist; p.point = point;
	p.nb = n_box; p.nr = n_leafref; p.nt = n_tri;
	run_jobs(&p, n, nthreads, cast_worker);
}
void ora_brute_cast(const float *tris, int ntris, const float *rays, int n, int clip, const int32_t *exclude,
                    int32_t *id, float *dist, float *second_dist, int nthreads) {
	job_t p; memset(&p, 0, sizeof(p));
	p.tris = tris; p.ntris = ntris; p.rays = rays; p.clip = clip; p.exclude = exclude;
	p.id = id; p.dist = dist; p.second = second_dist;
	run_jobs(&p, n, nthreads, brute_worker);
}

/* single-ray cast for the path estimator (dr_oracle_render.c) */
int ora_cast1(const ora_octree *t, const float *o, const float *d, int clip, int32_t exclude, int32_t *id, float *dist,
              float *point) {
	int32_t nb = 0, nr = 0, nt = 0; v3 hp = V(0, 0, 0);
	int r = oct_cast(t, ld3(o), ld3(d), clip, exclude, id, dist, &hp, &nb, &nr, &nt);
	if (r) { point[0] = hp.x; point[1] = hp.y; point[2] = hp.z; }
	return r;
}
const float *ora_octree_tri(const ora_octree *t, int32_t id) { return &t->tris[id].a.x; }

/* ---- primitive probes ---- */
int ora_tri_cast(const float *tri, const float *o, const float *d, int clip, float *t, float *p) {
	tri_t tr = ldtri(tri); v3 h; float dd;
	int r = tri_cast(&tr, ld3(o), ld3(d), clip, &dd, &h);
	if (r) { *t = dd; p[0] = h.x; p[1] = h.y; p[2] = h.z; }
	return r;
}
int ora_tri_contains(const float *tri, const float *p) { tri_t tr = ldtri(tri); return tri_contains(&tr, ld3(p)); }
float ora_aabb_cast_preinv(const float *box, const float *o, const float *inv) {
	aabb_t b; b.s = ld3(box); b.e = ld3(box + 3);
	return box_cast_preinv(b, ld3(o), ld3(inv));
}
int ora_aabb_intersects_tri(const float *box, const float *tri) {
	aabb_t b; b.s = ld3(box); b.e = ld3(box + 3);
	return box_isect_tri_dim(b, ldtri(tri), 0);
}
void ora_intersection_bounds(const float *box, const float *tri, float *out) {
	aabb_t b; b.s = ld3(box); b.e = ld3(box + 3);
	tlist l = {0}, tmp = {0};
	aabb_t r = isect_bounds(b, ldtri(tri), &l, &tmp);
	out[0] = r.s.x; out[1] = r.s.y; out[2] = r.s.z; out[3] = r.e.x; out[4] = r.e.y; out[5] = r.e.z;
	free(l.v); free(tmp.v);
}
void ora_intersection_center(const float *box, const float *tri, float *out) {
	aabb_t b; b.s = ld3(box); b.e = ld3(box + 3);
	tlist l = {0}, tmp = {0};
	v3 c = isect_center(b, ldtri(tri), &l, &tmp);
	out[0] = c.x; out[1] = c.y; out[2] = c.z;
	free(l.v); free(tmp.v);
}
int ora_clip_tri(const float *box, const float *tri, float *out) {
	aabb_t b; b.s = ld3(box); b.e = ld3(box + 3);
	tlist l = {0}, tmp = {0};
	clip_to_box(b, ldtri(tri), &l, &tmp);
	int n = l.n;
	for (int i = 0; i < n && i < 27; i++) {
		const tri_t *q = &l.v[i];
		float *o = out + 9 * i;
		o[0] = q->a.x; o[1] = q->a.y; o[2] = q->a.z; o[3] = q->b.x; o[4] = q->b.y; o[5] = q->b.z;
		o[6] = q->c.x; o[7] = q->c.y; o[8] = q->c.z;
	}
	free(l.v); free(tmp.v);
	return n;
}
/* Triangle::getMasses (Triangle.impl.h:248-277) */
void ora_tri_get_masses(const float *tri, const float *pp, float *out) {
	tri_t t = ldtri(tri); v3 center = ld3(pp);
	v3 ab = vsub(t.b, t.a);
	v3 bc = vsub(t.c, t.b);
	v3 ae = vsub(ab, vnormal_on(ab, bc));
	v3 ax = vsub(center, t.a);
	v3 ad = vnormal_on(ax, ae);
	if (ae.x < 0) { ad.x = -ad.x; ae.x = -ae.x; }
	if (ae.y < 0) { ad.y = -ad.y; ae.y = -ae.y; }
	if (ae.z < 0) { ad.z = -ad.z; ae.z = -ae.z; }
	float div = ad.x + ad.y + ad.z;
	if (div == 0) { out[0] = 1; out[1] = 0; out[2] = 0; return; }
	float g = (ae.x + ae.y + ae.z) / div;
	float tt;
	v3 by = vsub(vadd(t.a, vmul(ax, g)), t.b);
	if (bc.x < 0) { bc.x = -bc.x; by.x = -by.x; }
	if (bc.y < 0) { bc.y = -bc.y; by.y = -by.y; }
	if (bc.z < 0) { bc.z = -bc.z; by.z = -by.z; }
	div = bc.x + bc.y + bc.z;
	if (div == 0) tt = 0;
	else tt = (by.x + by.y + by.z) / div;
	float cc = tt, bb = (1 - tt), aa = (g - 1);
	v3 r = vdivf(V(aa, bb, cc), aa + bb + cc);
	out[0] = r.x; out[1] = r.y; out[2] = r.z;
}
/* Triangle::massCenter(masses) (Triangle.impl.h:318-320) */
void ora_tri_mass_center_w(const float *tri, const float *m, float *out) {
	tri_t t = ldtri(tri);
	v3 r = vadd(vadd(vmul(t.a, m[0]), vmul(t.b, m[1])), vmul(t.c, m[2]));
	out[0] = r.x; out[1] = r.y; out[2] = r.z;
}
