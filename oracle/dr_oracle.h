/*
 * CPU oracle for the DumbRay closest-hit / path-tracing hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker: only tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg load it.  The
 * product (dumbray_amd) never links, imports or calls anything here.
 *
 * It restates, in plain C with fp32 arithmetic evaluated left to right and
 * no FMA contraction (built with -ffp-contract=off), the reference CPU path:
 *   Vector3 ops            __source__/DataStructures/Primitives/Pure/Vector3/Vector3.impl.h:9-33
 *   AABB                   __source__/DataStructures/Primitives/Compound/AABB/AABB.impl.h:93-222
 *   Triangle               __source__/DataStructures/Primitives/Compound/Triangle/Triangle.impl.h:185-348
 *   Shapes clip/bounds     __source__/Namespaces/Shapes/Shapes.impl.cuh:202-382
 *   Octree build / cast    __source__/DataStructures/Objects/Scene/Raycasters/Octree/Octree.impl.cuh
 *   Path estimator         __source__/DataStructures/Renderers/DumbRenderer/DumbRenderer.cu:174-443
 * Each function below cites the reference lines it follows.
 *
 * Pinning (what checks each function against something other than itself)
 * is listed per function in tests/test_oracle_*.py.
 */
#ifndef DR_ORACLE_H
#define DR_ORACLE_H
#include <stdint.h>
#ifdef __cplusplus
extern "C" {
#endif

typedef struct ora_octree ora_octree;

/* ---- primitive probes (for closed-form tests) ---- */
/* Triangle::cast (Triangle.impl.h:334-348); tri = 9 floats a,b,c. Returns 1 on hit. */
int ora_tri_cast(const float *tri, const float *o, const float *d, int clip, float *t, float *p);
/* Triangle::containsVertex (Triangle.impl.h:283-316, active branch 292-302) */
int ora_tri_contains(const float *tri, const float *p);
/* AABB::castPreInversed (AABB.impl.h:116-130); box = 6 floats start,end; inv = 1/d */
float ora_aabb_cast_preinv(const float *box, const float *o, const float *inv);
/* AABB::intersectsTri<0> (AABB.impl.h:158-222) */
int ora_aabb_intersects_tri(const float *box, const float *tri);
/* Shapes::intersectionBounds<AABB,Triangle> (Shapes.impl.cuh:272-291) -> out 6 floats */
void ora_intersection_bounds(const float *box, const float *tri, float *out);
/* Shapes::intersectionCenter<AABB,Triangle> (Shapes.impl.cuh:293-312) -> out 3 floats */
void ora_intersection_center(const float *box, const float *tri, float *out);
/* clipped pieces of SHAPES_GET_INTERSECTING_TRIANGLES (Shapes.impl.cuh:255-266); returns count, out 27*9 floats */
int ora_clip_tri(const float *box, const float *tri, float *out);
/* Triangle::getMasses / massCenter (Triangle.impl.h:248-277, 318-330) */
void ora_tri_get_masses(const float *tri, const float *p, float *out);
void ora_tri_mass_center_w(const float *tri, const float *m, float *out);

/* ---- octree (Octree.impl.cuh) ---- */
ora_octree *ora_octree_build(const float *tris, int ntris, int split_threshold, int max_depth);
void ora_octree_free(ora_octree *t);
int ora_octree_node_count(const ora_octree *t);
/* export: bounds[n*6], children[n] (first child index or -1), list_off[n+1], list (tri ids) */
int ora_octree_list_total(const ora_octree *t);
void ora_octree_export(const ora_octree *t, float *bounds, int32_t *children, int32_t *list_off, int32_t *list);
/* Octree::cast (Octree.impl.cuh:174-212) over a batch; rays = n*6 (o,d); exclude = per ray tri id
 * rejected by validateNotSameAsObject (Octree.impl.cuh:163) or NULL.  Multi-threaded over rays
 * with nthreads (<=0: all cores).  Outputs id (-1 miss), t, hit point (may be NULL).  Counters
 * (may be NULL): per-ray boxes tested, leaf-list entries visited, triangles tested. */
void ora_octree_cast(const ora_octree *t, const float *rays, int n, int clip, const int32_t *exclude,
                     int32_t *id, float *dist, float *point, int32_t *n_box, int32_t *n_leafref,
                     int32_t *n_tri, int nthreads);
/* brute force closest hit over all triangles in push order with strict '<' (first wins ties) */
void ora_brute_cast(const float *tris, int ntris, const float *rays, int n, int clip, const int32_t *exclude,
                    int32_t *id, float *dist, float *second_dist, int nthreads);

#ifdef __cplusplus
}
#endif
#endif
