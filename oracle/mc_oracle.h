/*
 * mc_oracle.h -- CPU restatement of model3d's MarchingCubesSearch path.
 * TEST INFRASTRUCTURE ONLY (the checker); see mc_oracle.c.
 */
#ifndef MC_ORACLE_H
#define MC_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
    OR_SPHERE = 1,
    OR_RECT = 2,
    OR_CAPSULE = 3,
    OR_CYLINDER = 4,
    OR_CONE = 5,
    OR_CONESLICE = 6,
    OR_TORUS = 7,
    OR_JOINED = 10,
    OR_SUBTRACTED = 11,
    OR_INTERSECTED = 12,
    OR_TRANSFORM = 13,
    OR_FORCEBOUNDS = 14,
    OR_COLLIDER = 15,
    OR_HOST = 16,
};

enum { OR_XF_TRANSLATE = 1, OR_XF_SCALE = 2, OR_XF_VECSCALE = 3, OR_XF_MATRIX = 4 };

enum {
    OR_OK = 0,
    OR_E_INVALID_BOUNDS = 1,
    OR_E_INSIDE_BOUNDARY = 2,
    OR_E_NOT_ON_EDGE = 3,
    OR_E_OOM = 4,
};

typedef struct or_node or_node;
typedef struct or_mesh or_mesh;
typedef int (*or_host_fn)(void *user, const double *p);

or_node *or_sphere(const double *c, double r);
or_node *or_rect(const double *mn, const double *mx);
or_node *or_capsule(const double *p1, const double *p2, double r);
or_node *or_cylinder(const double *p1, const double *p2, double r);
or_node *or_cone(const double *tip, const double *base, double r);
or_node *or_cone_slice(const double *p1, const double *p2, double r1, double r2);
or_node *or_torus(const double *center, const double *axis, double outer, double inner);
or_node *or_joined(or_node **ch, int count);
or_node *or_intersected(or_node **ch, int count);
or_node *or_subtracted(or_node *pos, or_node *neg);
or_node *or_transform(or_node *child, const int *kinds, const double *params, int nops);
or_node *or_force_bounds(or_node *child, const double *mn, const double *mx);
or_node *or_collider(const double *tris, long ntris);
or_node *or_collider_inset(const double *tris, long ntris, double inset);
or_node *or_collider_hollow(const double *tris, long ntris, double r);
/* Triangle.SphereCollision (primitives.go:251-271) of one triangle (9 doubles) */
int or_tri_sphere_collision(const double *tri, const double *c, double r);
or_node *or_host(const double *mn, const double *mx, or_host_fn fn, void *user);
void or_free(or_node *n);

void or_bounds(const or_node *n, double *mn, double *mx);
int or_contains(const or_node *n, const double *pt);
void or_contains_batch(const or_node *n, const double *pts, long count, uint8_t *out);

int or_case_table(int32_t *out); /* 256 rows x 31: ntri, then 30 corner ids (-1 pad) */
int or_rotations(int32_t *out);  /* up to 24 rows x 8 */
long or_spacer(const or_node *s, double delta, int axis, double *out);

int or_marching_cubes_search(const or_node *s, double delta, int iters, int threads, or_mesh **out);
/* MarchingCubesInterior: the mesh plus, per unique vertex, 6 doubles
 * {searched vertex, interior point} (or_mesh_interior). */
int or_marching_cubes_interior(const or_node *s, double delta, int iters, int threads, or_mesh **out);
const double *or_mesh_interior(const or_mesh *m, long *n);
/* JoinedTransform(ops).Inverse().Apply at n points in place (3 doubles each);
 * ops as or_transform takes them. */
void or_inverse_apply(const int *kinds, const double *params, int nops, double *pts, long n);
/* MarchingCubesSearchFilter (MarchingCubesFilter when iters == 0): the
 * reference's block recursion, corners evaluated block by block. */
typedef int (*or_rect_fn)(void *user, const double *mn, const double *mx);
int or_marching_cubes_filter(const or_node *s, or_rect_fn f, void *user, double delta, int iters, int threads,
                             or_mesh **out);
/* 1 if any of nt triangles (9 doubles each) collides with the rect. */
int or_tris_rect_collision(const double *tris, long nt, const double *mn, const double *mx);
/* EncodeSTL of nt triangles (9 doubles each) into out (84 + 50 nt bytes). */
void or_encode_stl(const double *tris, long nt, uint8_t *out);
/* GroupTriangles of nt triangles (9 doubles each): perm[k] = index of the
 * k-th triangle in the grouped order (ties in the sorts by index).  0 or 4 (out of memory). */
int or_group_triangles(const double *tris, long nt, uint32_t *perm);
long or_mesh_ntris(const or_mesh *m);
const double *or_mesh_tris(const or_mesh *m);
void or_mesh_info(const or_mesh *m, long *corners, long *vertices, long *search_evals, double *ms_scan,
                  double *ms_search);
void or_mesh_free(or_mesh *m);

#ifdef __cplusplus
}
#endif
#endif
