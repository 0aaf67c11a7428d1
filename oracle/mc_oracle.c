/*
 * mc_oracle.c -- CPU restatement, in plain C and IEEE binary64, of model3d's
 * MarchingCubesSearch path (model3d/mc.go) and of every Solid.Contains the
 * device library lowers.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the checker: tests/, the smoke
 * entry point and the CPU-baseline leg of bench.py load it to produce the
 * expected outputs.  The product library never links, loads or calls it.
 *
 * Build: see oracle/Makefile (gcc -O2 -ffp-contract=off, no fast-math), so
 * every expression rounds exactly as the Go gc compiler's unfused amd64 code
 * does.  Each function cites the reference source it restates.
 *
 * Parity pins (what checks each part; see tests/test_oracle.py):
 *   - case table: determinism, one base case per key, every triangle edge is a
 *     sign-changing edge, outward orientation, closed manifold output on random
 *     occupancy grids (mc_test.go:9-42); committed fixture.
 *   - spacer: Python float loop reproduces it exactly; committed fixture.
 *   - Sphere mesh: closed forms T = 2V - 4 and |volume - 4/3 pi| < 1e-2
 *     (measurements_test.go:9-22); Contains equals the textbook predicate away
 *     from ties.
 *   - Rect/Capsule/Cylinder/Cone/ConeSlice/Torus Contains: textbook predicates
 *     away from ties, bounds property (shapes_test.go:87-155).
 *   - CSG: boolean identities (solid_test.go:8-95).
 *   - Bisection: vertex stays on its edge, |v - x*| <= delta/2^(iters+1).
 *   - Collider ray parity: brute force over all triangles, compared with an
 *     independent numpy Moller-Trumbore on small meshes.
 */
#define _GNU_SOURCE
#include <math.h>
#include <pthread.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <unistd.h>

#include "mc_oracle.h"

/* ------------------------------------------------------------------------ */
/* Coord3D arithmetic, model3d/coords.go                                      */
/* ------------------------------------------------------------------------ */

typedef struct { double x, y, z; } c3;

static inline c3 v3(double x, double y, double z) { c3 r = {x, y, z}; return r; }
static inline c3 c3_of(const double *a) { return v3(a[0], a[1], a[2]); }

/* coords.go:258-264 */
static inline c3 c3_add(c3 a, c3 b) { return v3(a.x + b.x, a.y + b.y, a.z + b.z); }
/* coords.go:241-246 */
static inline c3 c3_scale(c3 a, double s) { return v3(a.x * s, a.y * s, a.z * s); }
/* coords.go:267-269: c.Add(c1.Scale(-1)) */
static inline c3 c3_sub(c3 a, c3 b) { return c3_add(a, c3_scale(b, -1)); }
/* coords.go:225-227 */
static inline c3 c3_mul(c3 a, c3 b) { return v3(a.x * b.x, a.y * b.y, a.z * b.z); }
/* coords.go:235-237 */
static inline c3 c3_recip(c3 a) { return v3(1 / a.x, 1 / a.y, 1 / a.z); }
/* coords.go:250-255 */
static inline c3 c3_add_scalar(c3 a, double s) { return v3(a.x + s, a.y + s, a.z + s); }
/* coords.go:211-213 */
static inline double c3_dot(c3 a, c3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
/* coords.go:201-203 */
static inline double c3_norm(c3 a) { return sqrt(a.x * a.x + a.y * a.y + a.z * a.z); }
/* coords.go:272-277 */
static inline double c3_dist(c3 a, c3 b) {
    double d1 = a.x - b.x, d2 = a.y - b.y, d3 = a.z - b.z;
    return sqrt(d1 * d1 + d2 * d2 + d3 * d3);
}
/* coords.go:216-222 */
static inline c3 c3_cross(c3 a, c3 b) {
    return v3(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}
/* coords.go:379-381 */
static inline c3 c3_normalize(c3 a) { return c3_scale(a, 1 / c3_norm(a)); }
/* coords.go:374-376 */
static inline double c3_sum(c3 a) { return a.x + a.y + a.z; }
/* coords.go:196-198 */
static inline c3 c3_mid(c3 a, c3 b) { return c3_scale(c3_add(a, b), 0.5); }

/* Go's math.Min / math.Max special cases (Go stdlib math/dim.go): -Inf/+Inf
 * win over NaN, NaN propagates, and signed zeros are ordered. */
static inline double go_min(double x, double y) {
    if (isinf(x) && x < 0) return x;
    if (isinf(y) && y < 0) return y;
    if (isnan(x) || isnan(y)) return NAN;
    if (x == 0 && x == y) return signbit(x) ? x : y;
    return x < y ? x : y;
}
static inline double go_max(double x, double y) {
    if (isinf(x) && x > 0) return x;
    if (isinf(y) && y > 0) return y;
    if (isnan(x) || isnan(y)) return NAN;
    if (x == 0 && x == y) return signbit(x) ? y : x;
    return x > y ? x : y;
}
/* coords.go:342-349 */
static inline c3 c3_min(c3 a, c3 b) { return v3(go_min(a.x, b.x), go_min(a.y, b.y), go_min(a.z, b.z)); }
static inline c3 c3_max(c3 a, c3 b) { return v3(go_max(a.x, b.x), go_max(a.y, b.y), go_max(a.z, b.z)); }
/* Go struct equality on Coord3D: field-wise float == */
static inline int c3_eq(c3 a, c3 b) { return a.x == b.x && a.y == b.y && a.z == b.z; }

/* coords.go:388-421 */
static void c3_ortho_basis(c3 c, c3 *o1, c3 *o2) {
    double ax = fabs(c.x), ay = fabs(c.y), az = fabs(c.z);
    c3 b1 = v3(0, 0, 0);
    if (ax > ay && ax > az) {
        b1.x = c.y / ax;
        b1.y = -c.x / ax;
    } else {
        b1.y = c.z;
        b1.z = -c.y;
        if (ay > az) {
            b1.y /= ay;
            b1.z /= ay;
        } else {
            b1.y /= az;
            b1.z /= az;
        }
    }
    c3 b2 = v3(b1.y * c.z - b1.z * c.y, b1.z * c.x - b1.x * c.z, b1.x * c.y - b1.y * c.x);
    *o1 = c3_normalize(b1);
    *o2 = c3_normalize(b2);
}

/* ------------------------------------------------------------------------ */
/* Transforms, model3d/transform.go and matrix.go                             */
/* ------------------------------------------------------------------------ */

typedef struct {
    int kind;      /* OR_XF_* */
    double v[9];   /* translate/vecscale: v[0..2]; scale: v[0]; matrix: 9 row-major */
} xf_op;

/* matrix.go:131-137 */
static inline c3 mat_mul_col(const double *m, c3 c) {
    return v3(m[0] * c.x + m[1] * c.y + m[2] * c.z,
              m[3] * c.x + m[4] * c.y + m[5] * c.z,
              m[6] * c.x + m[7] * c.y + m[8] * c.z);
}

/* matrix.go:59-61 */
static double mat_det(const double *m) {
    return m[0] * (m[4] * m[8] - m[5] * m[7]) - m[1] * (m[3] * m[8] - m[5] * m[6]) +
           m[2] * (m[3] * m[7] - m[4] * m[6]);
}

/* matrix.go:69-90: cofactor matrix, then Scale(1/det) */
static void mat_inverse(const double *m, double *out) {
    double det = mat_det(m);
    double r[9] = {
        m[4] * m[8] - m[5] * m[7], m[2] * m[7] - m[1] * m[8], m[1] * m[5] - m[2] * m[4],
        m[5] * m[6] - m[3] * m[8], m[0] * m[8] - m[2] * m[6], m[2] * m[3] - m[0] * m[5],
        m[3] * m[7] - m[4] * m[6], m[1] * m[6] - m[0] * m[7], m[0] * m[4] - m[1] * m[3],
    };
    double s = 1 / det;
    for (int i = 0; i < 9; i++) out[i] = r[i] * s;
}

/* Inverse of one elementary transform: transform.go:52-54, 83-85, 114-116, 153-155 */
static xf_op xf_inverse(xf_op t) {
    xf_op r;
    memset(&r, 0, sizeof r);
    r.kind = t.kind;
    switch (t.kind) {
    case OR_XF_TRANSLATE: {
        c3 o = c3_scale(c3_of(t.v), -1);
        r.v[0] = o.x; r.v[1] = o.y; r.v[2] = o.z;
        break;
    }
    case OR_XF_SCALE: r.v[0] = 1 / t.v[0]; break;
    case OR_XF_VECSCALE: {
        c3 o = c3_recip(c3_of(t.v));
        r.v[0] = o.x; r.v[1] = o.y; r.v[2] = o.z;
        break;
    }
    case OR_XF_MATRIX: mat_inverse(t.v, r.v); break;
    }
    return r;
}

/* transform.go:44-46, 71-73, 103-105, 129-131 */
static c3 xf_apply(const xf_op *t, c3 c) {
    switch (t->kind) {
    case OR_XF_TRANSLATE: return c3_add(c, c3_of(t->v));
    case OR_XF_SCALE: return c3_scale(c, t->v[0]);
    case OR_XF_VECSCALE: return c3_mul(c, c3_of(t->v));
    case OR_XF_MATRIX: return mat_mul_col(t->v, c);
    }
    return c;
}

/* transform.go:48-50, 75-81, 107-111, 133-150 */
static void xf_apply_bounds(const xf_op *t, c3 *mn, c3 *mx) {
    c3 a = *mn, b = *mx;
    switch (t->kind) {
    case OR_XF_TRANSLATE:
        *mn = c3_add(a, c3_of(t->v));
        *mx = c3_add(b, c3_of(t->v));
        break;
    case OR_XF_SCALE:
        if (t->v[0] < 0) {
            *mn = c3_scale(b, t->v[0]);
            *mx = c3_scale(a, t->v[0]);
        } else {
            *mn = c3_scale(a, t->v[0]);
            *mx = c3_scale(b, t->v[0]);
        }
        break;
    case OR_XF_VECSCALE: {
        c3 s = c3_of(t->v);
        c3 p = c3_mul(a, s), q = c3_mul(b, s);
        *mn = c3_min(p, q);
        *mx = c3_max(q, p);
        break;
    }
    case OR_XF_MATRIX: {
        double xs[2] = {a.x, b.x}, ys[2] = {a.y, b.y}, zs[2] = {a.z, b.z};
        c3 nmin = v3(0, 0, 0), nmax = v3(0, 0, 0);
        for (int i = 0; i < 2; i++)
            for (int j = 0; j < 2; j++)
                for (int k = 0; k < 2; k++) {
                    c3 c = mat_mul_col(t->v, v3(xs[i], ys[j], zs[k]));
                    if (i == 0 && j == 0 && k == 0) {
                        nmin = c;
                        nmax = c;
                    } else {
                        nmin = c3_min(nmin, c);
                        nmax = c3_max(nmax, c);
                    }
                }
        *mn = nmin;
        *mx = nmax;
        break;
    }
    }
}

/* ------------------------------------------------------------------------ */
/* Solid tree                                                                 */
/* ------------------------------------------------------------------------ */

struct or_node {
    int kind;
    double p[16];
    int nchild;
    or_node **child;
    c3 mn, mx;          /* funcSolid/ColliderSolid stored bounds */
    int nxf;
    xf_op *inv;         /* inverse transform chain, in application order */
    long ntris;
    double *tris;       /* collider triangles, 9 doubles each */
    or_host_fn host_fn;
    void *host_user;
};

static or_node *node_new(int kind) {
    or_node *n = (or_node *)calloc(1, sizeof(or_node));
    n->kind = kind;
    return n;
}

static void node_min_max(const or_node *n, c3 *mn, c3 *mx);

/* circleAxisBound, shapes.go:577-585 (ProjectOut coords.go:424-427) */
static double circle_axis_bound(int axis, c3 normal, double sign) {
    double arr[3] = {0, 0, 0};
    arr[axis] = sign;
    c3 c = c3_of(arr);
    c3 normed = c3_normalize(normal);
    c3 proj = c3_sub(c, c3_scale(normed, c3_dot(normed, c)));
    proj = c3_scale(proj, 1 / (c3_norm(proj) + 1e-8));
    double pa[3] = {proj.x, proj.y, proj.z};
    return sign * (fabs(pa[axis]) + 1e-8);
}

static c3 circle_bounds(c3 axis, double sign) {
    return v3(circle_axis_bound(0, axis, sign), circle_axis_bound(1, axis, sign),
              circle_axis_bound(2, axis, sign));
}

static void node_min_max(const or_node *n, c3 *mn, c3 *mx) {
    const double *p = n->p;
    switch (n->kind) {
    case OR_SPHERE: /* shapes.go:19-26 */
        *mn = c3_add_scalar(c3_of(p), -p[3]);
        *mx = c3_add_scalar(c3_of(p), p[3]);
        return;
    case OR_RECT: /* shapes.go:161-168 */
        *mn = c3_of(p);
        *mx = c3_of(p + 3);
        return;
    case OR_CAPSULE: /* shapes.go:355-362 */
        *mn = c3_add_scalar(c3_min(c3_of(p), c3_of(p + 3)), -p[6]);
        *mx = c3_add_scalar(c3_max(c3_of(p), c3_of(p + 3)), p[6]);
        return;
    case OR_CYLINDER: { /* shapes.go:554-575 */
        c3 p1 = c3_of(p), p2 = c3_of(p + 3);
        c3 axis = c3_sub(p2, p1);
        *mn = c3_add(c3_min(p1, p2), c3_scale(circle_bounds(axis, -1), p[6]));
        *mx = c3_add(c3_max(p1, p2), c3_scale(circle_bounds(axis, 1), p[6]));
        return;
    }
    case OR_CONE: { /* shapes.go:867-886 */
        c3 tip = c3_of(p), base = c3_of(p + 3);
        c3 axis = c3_sub(tip, base);
        *mn = c3_min(c3_add(c3_scale(circle_bounds(axis, -1), p[6]), base), tip);
        *mx = c3_max(c3_add(c3_scale(circle_bounds(axis, 1), p[6]), base), tip);
        return;
    }
    case OR_CONESLICE: { /* shapes.go:1048-1072 */
        c3 p1 = c3_of(p), p2 = c3_of(p + 3);
        c3 axis = c3_sub(p2, p1);
        c3 d = circle_bounds(axis, -1);
        *mn = c3_min(c3_add(c3_scale(d, p[6]), p1), c3_add(c3_scale(d, p[7]), p2));
        d = circle_bounds(axis, 1);
        *mx = c3_max(c3_add(c3_scale(d, p[6]), p1), c3_add(c3_scale(d, p[7]), p2));
        return;
    }
    case OR_TORUS: { /* shapes.go:1342-1361 */
        c3 center = c3_of(p), axis = c3_of(p + 3);
        double outer = p[6], inner = p[7];
        c3 extra = v3(inner, inner, inner);
        *mn = c3_sub(c3_add(c3_scale(circle_bounds(axis, -1), outer), center), extra);
        *mx = c3_add(c3_add(c3_scale(circle_bounds(axis, 1), outer), center), extra);
        return;
    }
    case OR_JOINED: { /* solid.go:73-87 */
        c3 a, b;
        node_min_max(n->child[0], mn, mx);
        for (int i = 1; i < n->nchild; i++) {
            node_min_max(n->child[i], &a, &b);
            *mn = c3_min(*mn, a);
            *mx = c3_max(*mx, b);
        }
        return;
    }
    case OR_SUBTRACTED: /* solid.go:130-136 */
        node_min_max(n->child[0], mn, mx);
        return;
    case OR_INTERSECTED: { /* solid.go:146-161 */
        c3 a, b, lo, hi;
        node_min_max(n->child[0], &lo, &hi);
        for (int i = 1; i < n->nchild; i++) {
            node_min_max(n->child[i], &a, &b);
            lo = c3_max(lo, a);
            hi = c3_min(hi, b);
        }
        *mn = lo;
        *mx = c3_max(hi, lo);
        return;
    }
    default: /* funcSolid (solid.go:52-58), ColliderSolid (solid.go:282-289) */
        *mn = n->mn;
        *mx = n->mx;
        return;
    }
}

or_node *or_sphere(const double *c, double r) {
    or_node *n = node_new(OR_SPHERE);
    memcpy(n->p, c, 3 * sizeof(double));
    n->p[3] = r;
    return n;
}
or_node *or_rect(const double *mn, const double *mx) {
    or_node *n = node_new(OR_RECT);
    memcpy(n->p, mn, 3 * sizeof(double));
    memcpy(n->p + 3, mx, 3 * sizeof(double));
    return n;
}
static or_node *seg_node(int kind, const double *a, const double *b, double r) {
    or_node *n = node_new(kind);
    memcpy(n->p, a, 3 * sizeof(double));
    memcpy(n->p + 3, b, 3 * sizeof(double));
    n->p[6] = r;
    return n;
}
or_node *or_capsule(const double *p1, const double *p2, double r) { return seg_node(OR_CAPSULE, p1, p2, r); }
or_node *or_cylinder(const double *p1, const double *p2, double r) { return seg_node(OR_CYLINDER, p1, p2, r); }
or_node *or_cone(const double *tip, const double *base, double r) { return seg_node(OR_CONE, tip, base, r); }
or_node *or_cone_slice(const double *p1, const double *p2, double r1, double r2) {
    or_node *n = seg_node(OR_CONESLICE, p1, p2, r1);
    n->p[7] = r2;
    return n;
}
or_node *or_torus(const double *center, const double *axis, double outer, double inner) {
    or_node *n = seg_node(OR_TORUS, center, axis, outer);
    n->p[7] = inner;
    return n;
}
static or_node *multi_node(int kind, or_node **ch, int count) {
    or_node *n = node_new(kind);
    n->nchild = count;
    n->child = (or_node **)calloc(count > 0 ? count : 1, sizeof(or_node *));
    for (int i = 0; i < count; i++) n->child[i] = ch[i];
    return n;
}
or_node *or_joined(or_node **ch, int count) { return multi_node(OR_JOINED, ch, count); }
or_node *or_intersected(or_node **ch, int count) { return multi_node(OR_INTERSECTED, ch, count); }
or_node *or_subtracted(or_node *pos, or_node *neg) {
    or_node *ch[2] = {pos, neg};
    return multi_node(OR_SUBTRACTED, ch, 2);
}

/* TransformSolid, transform.go:240-246, with t a JoinedTransform of the given
 * elementary transforms (transform.go:191-219): bounds are pushed through each
 * op in order; the inverse chain is the reversed list of inverses. */
or_node *or_transform(or_node *child, const int *kinds, const double *params, int nops) {
    or_node *n = multi_node(OR_TRANSFORM, &child, 1);
    xf_op *fwd = (xf_op *)calloc(nops > 0 ? nops : 1, sizeof(xf_op));
    n->inv = (xf_op *)calloc(nops > 0 ? nops : 1, sizeof(xf_op));
    n->nxf = nops;
    for (int i = 0; i < nops; i++) {
        fwd[i].kind = kinds[i];
        memcpy(fwd[i].v, params + 9 * i, 9 * sizeof(double));
    }
    for (int i = 0; i < nops; i++) n->inv[i] = xf_inverse(fwd[nops - 1 - i]);
    c3 mn, mx;
    node_min_max(child, &mn, &mx);
    for (int i = 0; i < nops; i++) xf_apply_bounds(&fwd[i], &mn, &mx);
    n->mn = mn;
    n->mx = mx;
    free(fwd);
    return n;
}

/* ForceSolidBounds, solid.go:307-309 (CacheSolidBounds 316-318 passes the
 * child's own Min/Max). */
or_node *or_force_bounds(or_node *child, const double *mn, const double *mx) {
    or_node *n = multi_node(OR_FORCEBOUNDS, &child, 1);
    n->mn = c3_of(mn);
    n->mx = c3_of(mx);
    return n;
}

/* NewColliderSolid(MeshToCollider(mesh)), solid.go:256-258; the collider's
 * bounds are the union of triangle bounds (bvh/JoinedCollider min/max,
 * collisions.go:225-234, primitives.go Triangle.Min/Max). */
or_node *or_collider(const double *tris, long ntris) {
    or_node *n = node_new(OR_COLLIDER);
    n->ntris = ntris;
    n->tris = (double *)malloc((ntris > 0 ? ntris : 1) * 9 * sizeof(double));
    memcpy(n->tris, tris, ntris * 9 * sizeof(double));
    for (long i = 0; i < ntris; i++) {
        const double *t = tris + 9 * i;
        c3 a = c3_of(t), b = c3_of(t + 3), c = c3_of(t + 6);
        c3 tmn = c3_min(c3_min(a, b), c), tmx = c3_max(c3_max(a, b), c);
        if (i == 0) {
            n->mn = tmn;
            n->mx = tmx;
        } else {
            n->mn = c3_min(n->mn, tmn);
            n->mx = c3_max(n->mx, tmx);
        }
    }
    return n;
}

/* NewColliderSolidInset(MeshToCollider(mesh), inset), solid.go:260-270:
 * min = c.Min() + inset, max = min.Max(c.Max() - inset); and
 * NewColliderSolidHollow(MeshToCollider(mesh), r), solid.go:272-279:
 * min = c.Min() - r, max = c.Max() + r.  p[0] = inset, p[1] = radius. */
or_node *or_collider_inset(const double *tris, long ntris, double inset) {
    or_node *n = or_collider(tris, ntris);
    c3 iv = v3(inset, inset, inset);
    c3 mn = c3_add(n->mn, iv);
    n->mx = c3_max(mn, c3_sub(n->mx, iv));
    n->mn = mn;
    n->p[0] = inset;
    return n;
}

or_node *or_collider_hollow(const double *tris, long ntris, double r) {
    or_node *n = or_collider(tris, ntris);
    c3 rv = v3(r, r, r);
    n->mn = c3_sub(n->mn, rv);
    n->mx = c3_add(n->mx, rv);
    n->p[1] = r;
    return n;
}

or_node *or_host(const double *mn, const double *mx, or_host_fn fn, void *user) {
    or_node *n = node_new(OR_HOST);
    n->mn = c3_of(mn);
    n->mx = c3_of(mx);
    n->host_fn = fn;
    n->host_user = user;
    return n;
}

void or_free(or_node *n) {
    if (!n) return;
    for (int i = 0; i < n->nchild; i++) or_free(n->child[i]);
    free(n->child);
    free(n->inv);
    free(n->tris);
    free(n);
}

void or_bounds(const or_node *n, double *mn, double *mx) {
    c3 a, b;
    node_min_max(n, &a, &b);
    mn[0] = a.x; mn[1] = a.y; mn[2] = a.z;
    mx[0] = b.x; mx[1] = b.y; mx[2] = b.z;
}

/* ------------------------------------------------------------------------ */
/* Contains                                                                   */
/* ------------------------------------------------------------------------ */

/* Triangle.Normal, primitives.go:27-33 */
static c3 tri_normal(const double *t) {
    c3 a = c3_of(t), b = c3_of(t + 3), c = c3_of(t + 6);
    return c3_normalize(c3_cross(c3_sub(b, a), c3_sub(c, a)));
}

/* Triangle.rayCollision (primitives.go:207-249) as used by RayCollisions
 * (primitives.go:194-205): 1 if the ray hits with scale >= 0. */
static int tri_ray_hit(const double *t, c3 origin, c3 dir) {
    double d = c3_dot(tri_normal(t), c3_normalize(dir));
    if (d < 1e-8 && d > -1e-8) return 0;
    c3 t0 = c3_of(t), t1 = c3_of(t + 3), t2 = c3_of(t + 6);
    c3 v1 = c3_sub(t1, t0);
    c3 v2 = c3_sub(t2, t0);
    c3 cross1 = c3_cross(dir, v2);
    double det = c3_dot(cross1, v1);
    if (det == 0) return 0;
    double inv_det = 1 / det;
    c3 o = c3_sub(origin, t0);
    double bary1 = inv_det * c3_dot(o, cross1);
    if (bary1 < 0 || bary1 > 1) return 0;
    c3 cross2 = c3_cross(o, v1);
    double bary2 = inv_det * c3_dot(dir, cross2);
    if (bary2 < 0 || bary1 + bary2 > 1) return 0;
    double scale = inv_det * c3_dot(v2, cross2);
    /* RayCollisions: `if info == nil || scale < 0 { return 0 }` */
    return !(scale < 0);
}

/* Triangle.rayCollision (primitives.go:207-249) for Triangle.SphereCollision:
 * 1 when it reports a collision (info != nil), its scale in *scale. */
static int tri_ray_scale(const double *t, c3 origin, c3 dir, double *scale) {
    double d = c3_dot(tri_normal(t), c3_normalize(dir));
    if (d < 1e-8 && d > -1e-8) return 0;
    c3 t0 = c3_of(t), t1 = c3_of(t + 3), t2 = c3_of(t + 6);
    c3 v1 = c3_sub(t1, t0);
    c3 v2 = c3_sub(t2, t0);
    c3 cross1 = c3_cross(dir, v2);
    double det = c3_dot(cross1, v1);
    if (det == 0) return 0;
    double inv_det = 1 / det;
    c3 o = c3_sub(origin, t0);
    double bary1 = inv_det * c3_dot(o, cross1);
    if (bary1 < 0 || bary1 > 1) return 0;
    c3 cross2 = c3_cross(o, v1);
    double bary2 = inv_det * c3_dot(dir, cross2);
    if (bary2 < 0 || bary1 + bary2 > 1) return 0;
    *scale = inv_det * c3_dot(v2, cross2);
    return 1;
}

/* segmentEntersSphere, primitives.go:273-279 */
static int segment_enters_sphere(c3 p1, c3 p2, c3 c, double r) {
    c3 v = c3_sub(p2, p1);
    double frac = (c3_dot(c, v) - c3_dot(p1, v)) / c3_dot(v, v);
    c3 closest = c3_add(p1, c3_scale(v, frac));
    return frac >= 0 && frac <= 1 && c3_dist(closest, c) < r;
}

/* Triangle.SphereCollision, primitives.go:251-271 */
static int tri_sphere(const double *t, c3 c, double r) {
    for (int k = 0; k < 3; k++)
        if (c3_dist(c3_of(t + 3 * k), c) < r) return 1;
    for (int i = 0; i < 3; i++)
        if (segment_enters_sphere(c3_of(t + 3 * i), c3_of(t + 3 * ((i + 1) % 3)), c, r)) return 1;
    double scale;
    return tri_ray_scale(t, c, tri_normal(t), &scale) && fabs(scale) < r;
}

/* The mesh collider's SphereCollision (JoinedCollider, collisions.go:292-302):
 * any triangle's SphereCollision, here by brute force over all triangles. */
static int mesh_sphere(const or_node *n, c3 c, double r) {
    for (long i = 0; i < n->ntris; i++)
        if (tri_sphere(n->tris + 9 * i, c, r)) return 1;
    return 0;
}

int or_tri_sphere_collision(const double *tri, const double *c, double r) { return tri_sphere(tri, c3_of(c), r); }

/* Fixed ray direction of ColliderContains, collisions.go:120-125. */
static const double k_ray_dir[3] = {0.5224892708603626, 0.10494477243214506, 0.43558938446126527};

static int node_contains(const or_node *n, c3 c);

static int in_bounds(c3 c, c3 mn, c3 mx) {
    /* bounder.go:23-27 / solid.go:46-50 */
    return c3_eq(c3_min(c, mn), mn) && c3_eq(c3_max(c, mx), mx);
}

static int node_contains(const or_node *n, c3 c) {
    const double *p = n->p;
    switch (n->kind) {
    case OR_SPHERE: /* shapes.go:29-31 */
        return c3_dist(c, c3_of(p)) <= p[3];
    case OR_RECT: /* shapes.go:171-173 */
        return in_bounds(c, c3_of(p), c3_of(p + 3));
    case OR_CAPSULE: { /* shapes.go:365-368, primitives.go:547-593 */
        c3 p1 = c3_of(p), p2 = c3_of(p + 3);
        c3 s0, s1;
        if (p1.x < p2.x || (p1.x == p2.x && p1.y < p2.y) ||
            (p1.x == p2.x && p1.y == p2.y && p1.z < p2.z)) {
            s0 = p1; s1 = p2;
        } else {
            s0 = p2; s1 = p1;
        }
        c3 v1 = c3_sub(s1, s0);
        double norm = c3_norm(v1);
        c3 v = c3_scale(v1, 1 / norm);
        c3 v2 = c3_sub(c, s0);
        double mag = c3_dot(v, v2);
        c3 closest;
        if (mag > norm) closest = s1;
        else if (mag < 0) closest = s0;
        else closest = c3_add(c3_scale(v, mag), s0);
        return c3_dist(c, closest) <= p[6];
    }
    case OR_CYLINDER: { /* shapes.go:588-597 */
        c3 p1 = c3_of(p), p2 = c3_of(p + 3);
        c3 diff = c3_sub(p1, p2);
        c3 direction = c3_normalize(diff);
        double frac = c3_dot(c3_sub(c, p2), direction);
        if (frac < 0 || frac > c3_norm(diff)) return 0;
        c3 projection = c3_add(p2, c3_scale(direction, frac));
        return c3_dist(projection, c) <= p[6];
    }
    case OR_CONE: { /* shapes.go:888-898 */
        c3 tip = c3_of(p), base = c3_of(p + 3);
        c3 diff = c3_sub(tip, base);
        c3 direction = c3_normalize(diff);
        double frac = c3_dot(c3_sub(c, base), direction);
        double radius_frac = 1 - frac / c3_norm(diff);
        if (radius_frac < 0 || radius_frac > 1) return 0;
        c3 projection = c3_add(base, c3_scale(direction, frac));
        return c3_dist(projection, c) <= p[6] * radius_frac;
    }
    case OR_CONESLICE: { /* shapes.go:1073-1084 */
        c3 p1 = c3_of(p), p2 = c3_of(p + 3);
        c3 diff = c3_sub(p2, p1);
        c3 direction = c3_normalize(diff);
        double frac = c3_dot(c3_sub(c, p1), direction);
        double radius_frac = 1 - frac / c3_norm(diff);
        if (radius_frac < 0 || radius_frac > 1) return 0;
        c3 projection = c3_add(p1, c3_scale(direction, frac));
        double r = p[6] * radius_frac + p[7] * (1 - radius_frac);
        return c3_dist(projection, c) <= r;
    }
    case OR_TORUS: { /* shapes.go:1363-1365 -> genericSDF 1451-1484 */
        c3 center = c3_of(p), axis = c3_of(p + 3);
        c3 b1, b2;
        c3_ortho_basis(axis, &b1, &b2);
        c3 centered = c3_sub(c, center);
        double x = c3_dot(b1, centered);
        double y = c3_dot(b2, centered);
        double outer_norm = sqrt(x * x + y * y);
        if (outer_norm == 0) {
            x = 1;
            y = 0;
            outer_norm = 1;
        }
        double scale = p[6] / outer_norm;
        x *= scale;
        y *= scale;
        c3 ring = c3_add(c3_scale(b1, x), c3_scale(b2, y));
        return p[7] - c3_dist(ring, centered) >= 0;
    }
    case OR_JOINED: /* solid.go:89-96 */
        for (int i = 0; i < n->nchild; i++)
            if (node_contains(n->child[i], c)) return 1;
        return 0;
    case OR_SUBTRACTED: /* solid.go:138-140 */
        return node_contains(n->child[0], c) && !node_contains(n->child[1], c);
    case OR_INTERSECTED: /* solid.go:163-170 */
        for (int i = 0; i < n->nchild; i++)
            if (!node_contains(n->child[i], c)) return 0;
        return 1;
    case OR_TRANSFORM: { /* transform.go:240-246 inside CheckedFuncSolid solid.go:46-50 */
        if (!in_bounds(c, n->mn, n->mx)) return 0;
        c3 q = c;
        for (int i = 0; i < n->nxf; i++) q = xf_apply(&n->inv[i], q);
        return node_contains(n->child[0], q);
    }
    case OR_FORCEBOUNDS: /* solid.go:307-309 */
        if (!in_bounds(c, n->mn, n->mx)) return 0;
        return node_contains(n->child[0], c);
    case OR_COLLIDER: { /* ColliderSolid.Contains, solid.go:292-300 */
        if (!in_bounds(c, n->mn, n->mx)) return 0;
        const double inset = p[0], radius = p[1];
        if (radius != 0) return mesh_sphere(n, c, radius);
        /* ColliderContains(c, coord, inset), collisions.go:119-134 */
        c3 dir = c3_of(k_ray_dir);
        long count = 0;
        for (long i = 0; i < n->ntris; i++) count += tri_ray_hit(n->tris + 9 * i, c, dir);
        if (count % 2 == 0) {
            if (inset < 0) return mesh_sphere(n, c, -inset);
            return 0;
        }
        return inset <= 0 || !mesh_sphere(n, c, inset);
    }
    case OR_HOST: {
        double q[3] = {c.x, c.y, c.z};
        return n->host_fn(n->host_user, q) != 0;
    }
    }
    return 0;
}

int or_contains(const or_node *n, const double *pt) { return node_contains(n, c3_of(pt)); }

void or_contains_batch(const or_node *n, const double *pts, long count, uint8_t *out) {
    for (long i = 0; i < count; i++) out[i] = (uint8_t)node_contains(n, c3_of(pts + 3 * i));
}

/* ------------------------------------------------------------------------ */
/* Marching cubes case table, mc.go:285-586                                  */
/* ------------------------------------------------------------------------ */

typedef struct { uint8_t r[8]; } mc_rot;

/* mcRotation.Compose, mc.go:356-362 */
static mc_rot rot_compose(mc_rot m, mc_rot m1) {
    mc_rot res;
    for (int i = 0; i < 8; i++) res.r[i] = m.r[m1.r[i]];
    return res;
}

static int rot_cmp(const void *a, const void *b) {
    return memcmp(((const mc_rot *)a)->r, ((const mc_rot *)b)->r, 8);
}

/* allMcRotations, mc.go:312-353: closure of {zRotation, xRotation} sorted
 * lexicographically. */
static int all_rotations(mc_rot *out) {
    mc_rot zr = {{2, 0, 3, 1, 6, 4, 7, 5}};
    mc_rot xr = {{2, 3, 6, 7, 0, 1, 4, 5}};
    mc_rot queue[64];
    int qh = 0, qt = 0, count = 0;
    mc_rot id = {{0, 1, 2, 3, 4, 5, 6, 7}};
    queue[qt++] = id;
    out[count++] = id;
    while (qh < qt) {
        mc_rot next = queue[qh++];
        mc_rot ops[2] = {zr, xr};
        for (int k = 0; k < 2; k++) {
            mc_rot r = rot_compose(ops[k], next);
            int seen = 0;
            for (int i = 0; i < count; i++)
                if (!memcmp(out[i].r, r.r, 8)) seen = 1;
            if (!seen) {
                out[count++] = r;
                queue[qt++] = r;
            }
        }
    }
    qsort(out, count, sizeof(mc_rot), rot_cmp);
    return count;
}

/* baseTriangleTable, mc.go:462-586: the 23 rotation classes of the
 * "A survey of the marching cubes algorithm" (2006) table.  Each row: inside
 * corner mask, triangle count, then 6 corner indices per triangle (three
 * edges given as corner pairs, counter-clockwise seen from outside). */
typedef struct { uint8_t mask; uint8_t ntri; uint8_t t[5][6]; } base_case;

#define M1(a) (1u << (a))
static const base_case k_base[] = {
    {0, 0, {{0}}},
    {M1(0), 1, {{0, 1, 0, 2, 0, 4}}},
    {M1(0) | M1(1), 2, {{0, 4, 1, 5, 0, 2}, {1, 5, 1, 3, 0, 2}}},
    {M1(0) | M1(5), 2, {{0, 1, 0, 2, 0, 4}, {5, 7, 1, 5, 4, 5}}},
    {M1(0) | M1(7), 2, {{0, 1, 0, 2, 0, 4}, {6, 7, 3, 7, 5, 7}}},
    {M1(1) | M1(2) | M1(3), 3, {{0, 1, 1, 5, 0, 2}, {0, 2, 1, 5, 2, 6}, {2, 6, 1, 5, 3, 7}}},
    {M1(0) | M1(1) | M1(7), 3, {{0, 4, 1, 5, 0, 2}, {1, 5, 1, 3, 0, 2}, {6, 7, 3, 7, 5, 7}}},
    {M1(1) | M1(4) | M1(7), 3, {{4, 6, 4, 5, 0, 4}, {1, 5, 1, 3, 0, 1}, {6, 7, 3, 7, 5, 7}}},
    {M1(0) | M1(1) | M1(2) | M1(3), 2, {{0, 4, 1, 5, 3, 7}, {0, 4, 3, 7, 2, 6}}},
    {M1(0) | M1(2) | M1(3) | M1(6), 4,
     {{0, 1, 4, 6, 0, 4}, {0, 1, 6, 7, 4, 6}, {0, 1, 1, 3, 6, 7}, {1, 3, 3, 7, 6, 7}}},
    {M1(1) | M1(2) | M1(5) | M1(6), 4,
     {{0, 2, 2, 3, 6, 7}, {0, 2, 6, 7, 4, 6}, {0, 1, 4, 5, 5, 7}, {5, 7, 1, 3, 0, 1}}},
    {M1(0) | M1(2) | M1(3) | M1(7), 4,
     {{0, 4, 0, 1, 2, 6}, {0, 1, 5, 7, 2, 6}, {2, 6, 5, 7, 6, 7}, {0, 1, 1, 3, 5, 7}}},
    {M1(1) | M1(2) | M1(3) | M1(4), 4,
     {{0, 1, 1, 5, 0, 2}, {0, 2, 1, 5, 2, 6}, {2, 6, 1, 5, 3, 7}, {4, 5, 0, 4, 4, 6}}},
    {M1(1) | M1(2) | M1(4) | M1(7), 4,
     {{0, 1, 1, 5, 1, 3}, {0, 2, 2, 3, 2, 6}, {4, 5, 0, 4, 4, 6}, {5, 7, 6, 7, 3, 7}}},
    {M1(1) | M1(2) | M1(3) | M1(6), 4,
     {{0, 2, 0, 1, 4, 6}, {0, 1, 3, 7, 4, 6}, {0, 1, 1, 5, 3, 7}, {4, 6, 3, 7, 6, 7}}},
    {M1(0) | M1(2) | M1(3) | M1(5) | M1(6), 5,
     {{0, 1, 4, 6, 0, 4}, {0, 1, 6, 7, 4, 6}, {0, 1, 1, 3, 6, 7}, {1, 3, 3, 7, 6, 7}, {5, 7, 1, 5, 4, 5}}},
    {M1(2) | M1(3) | M1(4) | M1(5) | M1(6), 5,
     {{5, 7, 1, 5, 0, 4}, {0, 4, 6, 7, 5, 7}, {0, 2, 6, 7, 0, 4}, {0, 2, 3, 7, 6, 7}, {0, 2, 1, 3, 3, 7}}},
    {M1(0) | M1(4) | M1(5) | M1(6) | M1(7), 3, {{1, 5, 0, 1, 0, 2}, {0, 2, 2, 6, 1, 5}, {1, 5, 2, 6, 3, 7}}},
    {M1(1) | M1(2) | M1(3) | M1(4) | M1(5) | M1(6), 2, {{0, 2, 0, 1, 0, 4}, {3, 7, 6, 7, 5, 7}}},
    {M1(1) | M1(2) | M1(3) | M1(4) | M1(6) | M1(7), 4,
     {{0, 2, 4, 5, 0, 4}, {0, 2, 5, 7, 4, 5}, {0, 2, 1, 5, 5, 7}, {0, 1, 1, 5, 0, 2}}},
    {M1(2) | M1(3) | M1(4) | M1(5) | M1(6) | M1(7), 2, {{1, 5, 0, 4, 0, 2}, {1, 3, 1, 5, 0, 2}}},
    {M1(1) | M1(2) | M1(3) | M1(4) | M1(5) | M1(6) | M1(7), 1, {{0, 2, 0, 1, 0, 4}}},
    {0xff, 0, {{0}}},
};
#define N_BASE ((int)(sizeof(k_base) / sizeof(k_base[0])))

typedef struct { int ntri; uint8_t t[5][6]; int claimed_by; } mc_case;

/* mcLookupTable, mc.go:434-456: for every base case and every rotation in
 * sorted order, the first rotation reaching a key defines that key. */
static void build_table(mc_case *table, int *n_conflicts) {
    mc_rot rots[24];
    int nrot = all_rotations(rots);
    for (int k = 0; k < 256; k++) {
        table[k].ntri = -1;
        table[k].claimed_by = -1;
    }
    int conflicts = 0;
    for (int b = 0; b < N_BASE; b++) {
        for (int ri = 0; ri < nrot; ri++) {
            mc_rot rot = rots[ri];
            int key = 0;
            for (int c = 0; c < 8; c++)
                if (k_base[b].mask & (1u << c)) key |= 1 << rot.r[c];
            if (table[key].ntri >= 0) {
                if (table[key].claimed_by != b) conflicts++;
                continue;
            }
            table[key].ntri = k_base[b].ntri;
            table[key].claimed_by = b;
            for (int t = 0; t < k_base[b].ntri; t++)
                for (int i = 0; i < 6; i++) table[key].t[t][i] = rot.r[k_base[b].t[t][i]];
        }
    }
    if (n_conflicts) *n_conflicts = conflicts;
}

static mc_case g_table[256];
static pthread_once_t g_table_once = PTHREAD_ONCE_INIT;
static int g_table_conflicts;
static void init_table(void) { build_table(g_table, &g_table_conflicts); }

int or_case_table(int32_t *out) {
    pthread_once(&g_table_once, init_table);
    for (int k = 0; k < 256; k++) {
        int32_t *row = out + k * 31;
        row[0] = g_table[k].ntri;
        for (int i = 0; i < 30; i++) row[1 + i] = i < 6 * g_table[k].ntri ? g_table[k].t[i / 6][i % 6] : -1;
    }
    return g_table_conflicts;
}

int or_rotations(int32_t *out) {
    mc_rot rots[24];
    int n = all_rotations(rots);
    for (int i = 0; i < n; i++)
        for (int j = 0; j < 8; j++) out[8 * i + j] = rots[i].r[j];
    return n;
}

/* ------------------------------------------------------------------------ */
/* Spacer, mc.go:588-612                                                      */
/* ------------------------------------------------------------------------ */

typedef struct { double *v[3]; long n[3]; } spacer;

static long axis_values(double lo, double hi, double delta, double *out) {
    long n = 0;
    for (double x = lo - delta; x <= hi + delta; x += delta) {
        if (out) out[n] = x;
        n++;
    }
    return n;
}

static int spacer_make(const or_node *s, double delta, spacer *sp) {
    c3 mn, mx;
    node_min_max(s, &mn, &mx);
    double lo[3] = {mn.x, mn.y, mn.z}, hi[3] = {mx.x, mx.y, mx.z};
    for (int a = 0; a < 3; a++) {
        sp->n[a] = axis_values(lo[a], hi[a], delta, NULL);
        sp->v[a] = (double *)malloc((sp->n[a] > 0 ? sp->n[a] : 1) * sizeof(double));
        if (!sp->v[a]) return OR_E_OOM;
        axis_values(lo[a], hi[a], delta, sp->v[a]);
    }
    return 0;
}

static void spacer_free(spacer *sp) {
    for (int a = 0; a < 3; a++) free(sp->v[a]);
}

long or_spacer(const or_node *s, double delta, int axis, double *out) {
    c3 mn, mx;
    node_min_max(s, &mn, &mx);
    double lo[3] = {mn.x, mn.y, mn.z}, hi[3] = {mx.x, mx.y, mx.z};
    return axis_values(lo[axis], hi[axis], delta, out);
}

/* BoundsValid, bounder.go:30-46 */
static int bounds_valid(const or_node *s) {
    c3 mn, mx;
    node_min_max(s, &mn, &mx);
    double a = c3_sum(mn), b = c3_sum(mx);
    if (isnan(a) || isnan(b) || isinf(a) || isinf(b)) return 0;
    if (mx.x < mn.x || mx.y < mn.y || mx.z < mn.z) return 0;
    return 1;
}

/* ------------------------------------------------------------------------ */
/* Corner scan: GOMAXPROCS+1 plane ring, mc.go:614-645, 676-692               */
/* ------------------------------------------------------------------------ */

typedef struct {
    const or_node *solid;
    const spacer *sp;
    int nslots;
    uint8_t **planes;      /* nslots plane buffers */
    int *plane_z;          /* z held by each slot, -1 if free */
    int *ready;            /* slot filled */
    int next_z;            /* next z to hand out */
    int consumed_z;        /* highest z released by the consumer */
    int error;
    pthread_mutex_t mu;
    pthread_cond_t cv;
} scan_state;

/* solidCache.FetchZ, mc.go:676-692 */
static int fetch_plane(const or_node *s, const spacer *sp, long z, uint8_t *vals) {
    long nx = sp->n[0], ny = sp->n[1], nz = sp->n[2];
    int on_edge = z == 0 || z == nz - 1;
    long idx = 0;
    int err = 0;
    for (long i = 0; i < ny; i++) {
        for (long j = 0; j < nx; j++) {
            int b = node_contains(s, v3(sp->v[0][j], sp->v[1][i], sp->v[2][z]));
            vals[idx++] = (uint8_t)b;
            if (b && (on_edge || i == 0 || j == 0 || i == ny - 1 || j == nx - 1)) err = OR_E_INSIDE_BOUNDARY;
        }
    }
    return err;
}

static void *scan_worker(void *arg) {
    scan_state *st = (scan_state *)arg;
    long nz = st->sp->n[2];
    for (;;) {
        pthread_mutex_lock(&st->mu);
        /* A plane z may be produced once the consumer has released z - nslots. */
        while (st->next_z < nz && st->next_z - st->nslots > st->consumed_z - 1 && !st->error)
            pthread_cond_wait(&st->cv, &st->mu);
        if (st->next_z >= nz || st->error) {
            pthread_mutex_unlock(&st->mu);
            return NULL;
        }
        int z = st->next_z++;
        int slot = z % st->nslots;
        pthread_mutex_unlock(&st->mu);

        int err = fetch_plane(st->solid, st->sp, z, st->planes[slot]);

        pthread_mutex_lock(&st->mu);
        st->plane_z[slot] = z;
        st->ready[slot] = 1;
        if (err && !st->error) st->error = err;
        pthread_cond_broadcast(&st->cv);
        pthread_mutex_unlock(&st->mu);
    }
}

/* ------------------------------------------------------------------------ */
/* Triangle output and vertex dedup                                           */
/* ------------------------------------------------------------------------ */

struct or_mesh {
    long ntri, cap;
    double *tris;          /* 9 doubles per triangle */
    long n_corners, n_vertices, n_search_evals;
    double ms_scan, ms_search;
    double *interior;      /* MarchingCubesInterior: per unique vertex {searched vertex, interior point} */
    long n_interior;
};

static int mesh_push(or_mesh *m, const c3 *v) {
    if (m->ntri == m->cap) {
        long nc = m->cap ? m->cap * 2 : 1024;
        double *nt = (double *)realloc(m->tris, nc * 9 * sizeof(double));
        if (!nt) return OR_E_OOM;
        m->tris = nt;
        m->cap = nc;
    }
    double *t = m->tris + 9 * m->ntri++;
    for (int i = 0; i < 3; i++) {
        t[3 * i] = v[i].x;
        t[3 * i + 1] = v[i].y;
        t[3 * i + 2] = v[i].z;
    }
    return 0;
}

/* hash3 of the coordinate bits, coords.go:451-467 (same mixing constants, so
 * the dedup behaves as the reference's CoordToSlice). */
static uint64_t coord_hash(const double *p) {
    uint64_t x, y, z;
    memcpy(&x, p, 8);
    memcpy(&y, p + 1, 8);
    memcpy(&z, p + 2, 8);
    uint64_t h = x * 0x9e3779b185ebca87ULL ^ y * 0xc2b2ae3d27d4eb4fULL ^ z * 0x165667b19e3779f9ULL;
    h ^= h >> 33;
    h *= 0xff51afd7ed558ccdULL;
    h ^= h >> 33;
    h *= 0xc4ceb9fe1a85ec53ULL;
    h ^= h >> 33;
    return h;
}

/* ------------------------------------------------------------------------ */
/* mcSearchPoint, mc.go:223-258; LookupEdgePoint mc.go:647-660                */
/* ------------------------------------------------------------------------ */

static int lookup_edge_point(const spacer *sp, const double *arr, int *axis, double *lo, double *hi) {
    double origin[3] = {sp->v[0][0], sp->v[1][0], sp->v[2][0]};
    double delta = sp->v[0][1] - sp->v[0][0];
    for (int i = 0; i < 3; i++) {
        double modulus = fabs(fmod(arr[i] - origin[i], delta));
        if (modulus > delta / 4 && modulus < 3 * delta / 4) {
            long idx = (long)((arr[i] - origin[i]) / delta);
            if (idx < 0 || idx + 1 >= sp->n[i]) return OR_E_NOT_ON_EDGE;
            *axis = i;
            *lo = sp->v[i][idx];
            *hi = sp->v[i][idx + 1];
            return 0;
        }
    }
    return OR_E_NOT_ON_EDGE;
}

static int search_point(const or_node *s, int iters, const spacer *sp, const double *c, double *out,
                        double *interior) {
    int axis;
    double false_pt, true_pt;
    int err = lookup_edge_point(sp, c, &axis, &false_pt, &true_pt);
    if (err) return err;
    double arr[3] = {c[0], c[1], c[2]};
    double tp[3] = {c[0], c[1], c[2]};
    tp[axis] = true_pt;
    if (!node_contains(s, c3_of(tp))) {
        double t = false_pt;
        false_pt = true_pt;
        true_pt = t;
    }
    for (int i = 0; i < iters; i++) {
        double mid = (false_pt + true_pt) / 2;
        arr[axis] = mid;
        if (node_contains(s, c3_of(arr))) true_pt = mid;
        else false_pt = mid;
    }
    arr[axis] = (false_pt + true_pt) / 2;
    memcpy(out, arr, 3 * sizeof(double));
    if (interior) { /* mc.go:252-255 */
        arr[axis] = true_pt;
        memcpy(interior, arr, 3 * sizeof(double));
    }
    return 0;
}

typedef struct {
    const or_node *s;
    const spacer *sp;
    int iters;
    const double *in;
    double *out;
    double *interior; /* or NULL */
    long n;
    long next;
    int error;
    pthread_mutex_t mu;
} search_state;

static void *search_worker(void *arg) {
    search_state *st = (search_state *)arg;
    for (;;) {
        pthread_mutex_lock(&st->mu);
        long start = st->next;
        st->next += 256;
        pthread_mutex_unlock(&st->mu);
        if (start >= st->n) return NULL;
        long end = start + 256 < st->n ? start + 256 : st->n;
        for (long i = start; i < end; i++) {
            int err = search_point(st->s, st->iters, st->sp, st->in + 3 * i, st->out + 3 * i,
                                   st->interior ? st->interior + 3 * i : NULL);
            if (err) {
                pthread_mutex_lock(&st->mu);
                st->error = err;
                pthread_mutex_unlock(&st->mu);
            }
        }
    }
}

static double now_ms(void) {
    struct timespec ts;
    clock_gettime(CLOCK_MONOTONIC, &ts);
    return ts.tv_sec * 1e3 + ts.tv_nsec * 1e-6;
}

/* mcSearch, mc.go:182-221: unique vertices, parallel search, in-place
 * rewrite of every triangle corner that equals the original vertex.  With
 * want_interior the interior point of each unique vertex is kept beside its
 * searched position (the CoordMap of MarchingCubesInterior, mc.go:191-198). */
static int mc_search(const or_node *s, double delta, int iters, int threads, or_mesh *m, int want_interior) {
    if (iters == 0 && !want_interior) return 0;
    spacer sp;
    int err = spacer_make(s, delta, &sp);
    if (err) {
        spacer_free(&sp);
        return err;
    }
    long nslot = 3 * m->ntri;
    long cap = 1;
    while (cap < 2 * nslot + 16) cap <<= 1;
    long *table = (long *)malloc(cap * sizeof(long));
    long *slot_vertex = (long *)malloc((nslot > 0 ? nslot : 1) * sizeof(long));
    double *uniq = (double *)malloc((nslot > 0 ? nslot : 1) * 3 * sizeof(double));
    if (!table || !slot_vertex || !uniq) {
        free(table); free(slot_vertex); free(uniq);
        spacer_free(&sp);
        return OR_E_OOM;
    }
    for (long i = 0; i < cap; i++) table[i] = -1;
    long nv = 0;
    for (long k = 0; k < nslot; k++) {
        const double *p = m->tris + 3 * k;
        uint64_t h = coord_hash(p) & (uint64_t)(cap - 1);
        for (;;) {
            long v = table[h];
            if (v < 0) {
                table[h] = nv;
                memcpy(uniq + 3 * nv, p, 3 * sizeof(double));
                slot_vertex[k] = nv++;
                break;
            }
            if (!memcmp(uniq + 3 * v, p, 3 * sizeof(double))) {
                slot_vertex[k] = v;
                break;
            }
            h = (h + 1) & (uint64_t)(cap - 1);
        }
    }
    free(table);
    double *outv = (double *)malloc((nv > 0 ? nv : 1) * 3 * sizeof(double));
    double *inter = want_interior ? (double *)malloc((nv > 0 ? nv : 1) * 3 * sizeof(double)) : NULL;
    search_state st;
    memset(&st, 0, sizeof st);
    st.s = s;
    st.sp = &sp;
    st.iters = iters;
    st.in = uniq;
    st.out = outv;
    st.interior = inter;
    st.n = nv;
    pthread_mutex_init(&st.mu, NULL);
    if (threads < 1) threads = 1;
    pthread_t *th = (pthread_t *)malloc(threads * sizeof(pthread_t));
    for (int i = 0; i < threads; i++) pthread_create(&th[i], NULL, search_worker, &st);
    for (int i = 0; i < threads; i++) pthread_join(th[i], NULL);
    free(th);
    pthread_mutex_destroy(&st.mu);
    err = st.error;
    if (!err) {
        for (long k = 0; k < nslot; k++) memcpy(m->tris + 3 * k, outv + 3 * slot_vertex[k], 3 * sizeof(double));
        if (inter) {
            m->interior = (double *)malloc((nv > 0 ? nv : 1) * 6 * sizeof(double));
            for (long v = 0; v < nv; v++) {
                memcpy(m->interior + 6 * v, outv + 3 * v, 3 * sizeof(double));
                memcpy(m->interior + 6 * v + 3, inter + 3 * v, 3 * sizeof(double));
            }
            m->n_interior = nv;
        }
    }
    free(inter);
    m->n_vertices = nv;
    m->n_search_evals = nv * (long)(1 + iters);
    free(outv);
    free(uniq);
    free(slot_vertex);
    spacer_free(&sp);
    return err;
}

/* ------------------------------------------------------------------------ */
/* MarchingCubesFilter, mc.go:73-135, with mcBlock / mcBlockCache            */
/* (mc.go:724-858): single-threaded, block by block.                          */
/* ------------------------------------------------------------------------ */

typedef struct { long mn[3], mx[3]; } mc_block;

static long block_volume(const mc_block *b) {
    return (b->mx[0] - b->mn[0]) * (b->mx[1] - b->mn[1]) * (b->mx[2] - b->mn[2]);
}

typedef struct {
    const or_node *s;
    const spacer *sp;
    or_rect_fn f;
    void *user;
    double eps;
    or_mesh *m;
    uint8_t *vals;
    int err;
} filter_state;

/* blockFilter: f(m.Bounds(epsilon)), mc.go:95-101, 736-740 */
static int block_accept(filter_state *st, const mc_block *b) {
    double mn[3], mx[3];
    for (int a = 0; a < 3; a++) {
        mn[a] = st->sp->v[a][b->mn[a]] + -st->eps;
        mx[a] = st->sp->v[a][b->mx[a]] + st->eps;
    }
    return st->f(st->user, mn, mx) != 0;
}

/* mcBlockCache.Populate + the cube loop of MarchingCubesFilter (mc.go:104-121,
 * 808-858) */
static void process_block(filter_state *st, const mc_block *b) {
    const spacer *sp = st->sp;
    long sx = b->mx[0] - b->mn[0] + 1, sy = b->mx[1] - b->mn[1] + 1, sz = b->mx[2] - b->mn[2] + 1;
    uint8_t *v = (uint8_t *)realloc(st->vals, (size_t)(sx * sy * sz));
    if (!v) { st->err = OR_E_OOM; return; }
    st->vals = v;
    long k = 0;
    for (long z = b->mn[2]; z <= b->mx[2]; z++) {
        int ez = z == 0 || z == sp->n[2] - 1;
        for (long y = b->mn[1]; y <= b->mx[1]; y++) {
            int ey = y == 0 || y == sp->n[1] - 1;
            for (long x = b->mn[0]; x <= b->mx[0]; x++) {
                int ex = x == 0 || x == sp->n[0] - 1;
                int value = node_contains(st->s, v3(sp->v[0][x], sp->v[1][y], sp->v[2][z]));
                if (value && (ex || ey || ez)) { st->err = OR_E_INSIDE_BOUNDARY; return; }
                v[k++] = (uint8_t)value;
            }
        }
    }
    for (long z = b->mn[2]; z < b->mx[2]; z++)
        for (long y = b->mn[1]; y < b->mx[1]; y++)
            for (long x = b->mn[0]; x < b->mx[0]; x++) {
                int bits = 0, mask = 1;
                for (int z1 = 0; z1 < 2; z1++) /* getSquare(z) | getSquare(z+1) << 4 */
                    for (int y1 = 0; y1 < 2; y1++)
                        for (int x1 = 0; x1 < 2; x1++, mask <<= 1) {
                            long i = ((z + z1 - b->mn[2]) * sy + (y + y1 - b->mn[1])) * sx + (x + x1 - b->mn[0]);
                            if (v[i]) bits |= mask;
                        }
                const mc_case *cs = &g_table[bits];
                if (cs->ntri <= 0) continue;
                c3 mn = v3(sp->v[0][x], sp->v[1][y], sp->v[2][z]);
                c3 mx = v3(sp->v[0][x + 1], sp->v[1][y + 1], sp->v[2][z + 1]);
                c3 corners[8];
                for (int c = 0; c < 8; c++)
                    corners[c] = v3(c & 1 ? mx.x : mn.x, c & 2 ? mx.y : mn.y, c & 4 ? mx.z : mn.z);
                for (int t = 0; t < cs->ntri; t++) {
                    c3 tri[3];
                    for (int e = 0; e < 3; e++) tri[e] = c3_mid(corners[cs->t[t][2 * e]], corners[cs->t[t][2 * e + 1]]);
                    if (mesh_push(st->m, tri)) { st->err = OR_E_OOM; return; }
                }
            }
}

/* mcBlock.Pieces with Split, mc.go:760-797; leaf != 0: process, else recurse
 * again with the sub-division volume (the second level, mc.go:105-107). */
static void block_pieces(filter_state *st, const mc_block *b, long min_volume, int second) {
    if (st->err || !block_accept(st, b)) return;
    if (block_volume(b) / 2 < min_volume) {
        if (second) process_block(st, b);
        else block_pieces(st, b, 64, 1);
        return;
    }
    long lx = b->mx[0] - b->mn[0], ly = b->mx[1] - b->mn[1], lz = b->mx[2] - b->mn[2];
    int axis = 0;
    if (ly >= lx && ly >= lz) axis = 1;
    else if (lz >= lx && lz >= ly) axis = 2;
    mc_block b1 = *b, b2 = *b;
    b1.mx[axis] = (b->mx[axis] + b->mn[axis]) / 2;
    b2.mn[axis] = b1.mx[axis];
    block_pieces(st, &b1, min_volume, second);
    block_pieces(st, &b2, min_volume, second);
}

int or_marching_cubes_filter(const or_node *s, or_rect_fn f, void *user, double delta, int iters, int threads,
                             or_mesh **out) {
    *out = NULL;
    if (!bounds_valid(s)) return OR_E_INVALID_BOUNDS;
    pthread_once(&g_table_once, init_table);
    spacer sp;
    int err = spacer_make(s, delta, &sp);
    if (err) { spacer_free(&sp); return err; }
    or_mesh *m = (or_mesh *)calloc(1, sizeof(or_mesh));
    m->n_corners = sp.n[0] * sp.n[1] * sp.n[2];
    filter_state st = {s, &sp, f, user, delta * 1e-3, m, NULL, 0};
    mc_block root = {{0, 0, 0}, {sp.n[0] - 1, sp.n[1] - 1, sp.n[2] - 1}};
    long divide = block_volume(&root) / 4096;
    if (divide < 64) divide = 64; /* essentials.MaxInt(rootBlock.Volume()/4096, subDivideVolume) */
    block_pieces(&st, &root, divide, 0);
    free(st.vals);
    spacer_free(&sp);
    err = st.err;
    if (!err) err = mc_search(s, delta, iters, threads, m, 0);
    if (err) { or_mesh_free(m); return err; }
    *out = m;
    return 0;
}

/* RectCollision of a triangle list (MeshToCollider(...).RectCollision, the
 * OR over its triangles of Triangle.RectCollision, primitives.go:448-495),
 * brute force. */
static int ray_scale(const double *t, c3 origin, c3 dir, double *scale) {
    double d = c3_dot(tri_normal(t), c3_normalize(dir));
    if (d < 1e-8 && d > -1e-8) return 0;
    c3 t0 = c3_of(t), t1 = c3_of(t + 3), t2 = c3_of(t + 6);
    c3 v1 = c3_sub(t1, t0), v2 = c3_sub(t2, t0);
    c3 cross1 = c3_cross(dir, v2);
    double det = c3_dot(cross1, v1);
    if (det == 0) return 0;
    double inv_det = 1 / det;
    c3 o = c3_sub(origin, t0);
    double bary1 = inv_det * c3_dot(o, cross1);
    if (bary1 < 0 || bary1 > 1) return 0;
    c3 cross2 = c3_cross(o, v1);
    double bary2 = inv_det * c3_dot(dir, cross2);
    if (bary2 < 0 || bary1 + bary2 > 1) return 0;
    *scale = inv_det * c3_dot(v2, cross2);
    return *scale >= 0; /* FirstRayCollision, primitives.go:181-188 */
}

static void new_seg(c3 p1, c3 p2, c3 *a, c3 *b) { /* NewSegment, primitives.go:547-554 */
    if (p1.x < p2.x || (p1.x == p2.x && p1.y < p2.y) || (p1.x == p2.x && p1.y == p2.y && p1.z < p2.z)) {
        *a = p1, *b = p2;
    } else {
        *a = p2, *b = p1;
    }
}

static int seg_rect(c3 a, c3 b, const double *mn, const double *mx) { /* primitives.go:632-638, bvh.go:322-351 */
    c3 dir = c3_sub(b, a);
    double o[3] = {a.x, a.y, a.z}, r[3] = {dir.x, dir.y, dir.z};
    double min_frac = -INFINITY, max_frac = INFINITY;
    for (int axis = 0; axis < 3; axis++) {
        if (r[axis] == 0) {
            if (o[axis] < mn[axis] || o[axis] > mx[axis]) { min_frac = 0, max_frac = -1; break; }
            continue;
        }
        double t1 = (mn[axis] - o[axis]) / r[axis], t2 = (mx[axis] - o[axis]) / r[axis];
        if (t1 > t2) { double tt = t1; t1 = t2; t2 = tt; }
        if (t2 < 0) { min_frac = 0, max_frac = -1; break; }
        if (t1 > min_frac) min_frac = t1;
        if (t2 < max_frac) max_frac = t2;
    }
    return max_frac >= min_frac && max_frac >= 0 && min_frac <= 1;
}

static int tri_seg(const double *t, c3 p1, c3 p2) { /* Triangle.SegmentCollision, primitives.go:440-445 */
    c3 a, b;
    new_seg(p1, p2, &a, &b);
    double sc;
    return ray_scale(t, a, c3_sub(b, a), &sc) && sc <= 1;
}

static int tri_rect(const double *t, const double *mn, const double *mx) {
    for (int i = 0; i < 3; i++) {
        c3 a, b;
        new_seg(c3_of(t + 3 * i), c3_of(t + 3 * ((i + 1) % 3)), &a, &b);
        if (seg_rect(a, b, mn, mx)) return 1;
    }
    for (int k = 0; k < 2; k++) {
        double x = k ? mx[0] : mn[0];
        if (tri_seg(t, v3(x, mn[1], mn[2]), v3(x, mn[1], mx[2])) || tri_seg(t, v3(x, mn[1], mn[2]), v3(x, mx[1], mn[2])) ||
            tri_seg(t, v3(x, mx[1], mn[2]), v3(x, mx[1], mx[2])) || tri_seg(t, v3(x, mn[1], mx[2]), v3(x, mx[1], mx[2])))
            return 1;
    }
    return tri_seg(t, v3(mn[0], mn[1], mn[2]), v3(mx[0], mn[1], mn[2])) ||
           tri_seg(t, v3(mn[0], mx[1], mn[2]), v3(mx[0], mx[1], mn[2])) ||
           tri_seg(t, v3(mn[0], mn[1], mx[2]), v3(mx[0], mn[1], mx[2])) ||
           tri_seg(t, v3(mn[0], mx[1], mx[2]), v3(mx[0], mx[1], mx[2]));
}

int or_tris_rect_collision(const double *tris, long nt, const double *mn, const double *mx) {
    for (long i = 0; i < nt; i++)
        if (tri_rect(tris + 9 * i, mn, mx)) return 1;
    return 0;
}

/* MarchingCubes (mc.go:14-37) followed by mcSearch (mc.go:182-221); with
 * want_interior, MarchingCubesInterior (mc.go:55-60). */
static int marching_cubes(const or_node *s, double delta, int iters, int threads, int want_interior,
                          or_mesh **out) {
    *out = NULL;
    if (!bounds_valid(s)) return OR_E_INVALID_BOUNDS;
    pthread_once(&g_table_once, init_table);
    if (threads < 1) threads = 1;

    double t0 = now_ms();
    spacer sp;
    int err = spacer_make(s, delta, &sp);
    if (err) {
        spacer_free(&sp);
        return err;
    }
    or_mesh *m = (or_mesh *)calloc(1, sizeof(or_mesh));
    long nx = sp.n[0], ny = sp.n[1], nz = sp.n[2];
    m->n_corners = nx * ny * nz;

    /* Scan.numGos is capped at len(Zs)-1, mc.go:615-621. */
    int workers = threads;
    if (workers > nz - 1) workers = (int)(nz - 1);
    if (workers < 1) workers = 1;
    scan_state st;
    memset(&st, 0, sizeof st);
    st.solid = s;
    st.sp = &sp;
    st.nslots = workers + 1;
    st.planes = (uint8_t **)calloc(st.nslots, sizeof(uint8_t *));
    st.plane_z = (int *)malloc(st.nslots * sizeof(int));
    st.ready = (int *)calloc(st.nslots, sizeof(int));
    for (int i = 0; i < st.nslots; i++) {
        st.planes[i] = (uint8_t *)malloc(nx * ny);
        st.plane_z[i] = -1;
    }
    st.consumed_z = 0;
    pthread_mutex_init(&st.mu, NULL);
    pthread_cond_init(&st.cv, NULL);
    pthread_t *th = (pthread_t *)malloc(workers * sizeof(pthread_t));
    for (int i = 0; i < workers; i++) pthread_create(&th[i], NULL, scan_worker, &st);

    for (long z = 0; z < nz && !err; z++) {
        int slot = (int)(z % st.nslots);
        pthread_mutex_lock(&st.mu);
        while (!(st.ready[slot] && st.plane_z[slot] == z) && !st.error) pthread_cond_wait(&st.cv, &st.mu);
        if (st.error) err = st.error;
        pthread_mutex_unlock(&st.mu);
        if (err || z == 0) continue;
        const uint8_t *bot = st.planes[(z - 1) % st.nslots];
        const uint8_t *top = st.planes[slot];
        /* Emission loop, mc.go:20-35 */
        for (long y = 0; y < ny - 1 && !err; y++) {
            for (long x = 0; x < nx - 1; x++) {
                int bits = 0, mask = 1;
                for (int y1 = 0; y1 < 2; y1++)
                    for (int x1 = 0; x1 < 2; x1++, mask <<= 1)
                        if (bot[(x + x1) + (y + y1) * nx]) bits |= mask;
                for (int y1 = 0; y1 < 2; y1++)
                    for (int x1 = 0; x1 < 2; x1++, mask <<= 1)
                        if (top[(x + x1) + (y + y1) * nx]) bits |= mask;
                const mc_case *cs = &g_table[bits];
                if (cs->ntri <= 0) continue;
                c3 mn = v3(sp.v[0][x], sp.v[1][y], sp.v[2][z - 1]);
                c3 mx = v3(sp.v[0][x + 1], sp.v[1][y + 1], sp.v[2][z]);
                c3 corners[8];
                for (int c = 0; c < 8; c++)
                    corners[c] = v3(c & 1 ? mx.x : mn.x, c & 2 ? mx.y : mn.y, c & 4 ? mx.z : mn.z);
                for (int t = 0; t < cs->ntri; t++) {
                    c3 tri[3];
                    for (int e = 0; e < 3; e++) tri[e] = c3_mid(corners[cs->t[t][2 * e]], corners[cs->t[t][2 * e + 1]]);
                    if (mesh_push(m, tri)) {
                        err = OR_E_OOM;
                        break;
                    }
                }
            }
        }
        /* Release plane z-1 so a worker can reuse its slot. */
        pthread_mutex_lock(&st.mu);
        st.ready[(z - 1) % st.nslots] = 0;
        st.consumed_z = (int)z;
        pthread_cond_broadcast(&st.cv);
        pthread_mutex_unlock(&st.mu);
    }
    pthread_mutex_lock(&st.mu);
    if (err && !st.error) st.error = err;
    st.next_z = (int)nz;
    pthread_cond_broadcast(&st.cv);
    pthread_mutex_unlock(&st.mu);
    for (int i = 0; i < workers; i++) pthread_join(th[i], NULL);
    free(th);
    for (int i = 0; i < st.nslots; i++) free(st.planes[i]);
    free(st.planes);
    free(st.plane_z);
    free(st.ready);
    pthread_mutex_destroy(&st.mu);
    pthread_cond_destroy(&st.cv);
    spacer_free(&sp);
    double t1 = now_ms();
    m->ms_scan = t1 - t0;
    if (!err) err = mc_search(s, delta, iters, threads, m, want_interior);
    m->ms_search = now_ms() - t1;
    if (err) {
        or_mesh_free(m);
        return err;
    }
    *out = m;
    return 0;
}

int or_marching_cubes_search(const or_node *s, double delta, int iters, int threads, or_mesh **out) {
    return marching_cubes(s, delta, iters, threads, 0, out);
}

int or_marching_cubes_interior(const or_node *s, double delta, int iters, int threads, or_mesh **out) {
    return marching_cubes(s, delta, iters, threads, 1, out);
}

const double *or_mesh_interior(const or_mesh *m, long *n) {
    *n = m->n_interior;
    return m->interior;
}

/* JoinedTransform(xforms).Inverse().Apply at every point (transform.go:
 * 199-204, 213-219, with each op's Inverse and Apply): the vertex map that
 * MarchingCubesConj (mc.go:66-71) applies through Mesh.Transform
 * (mesh.go:613-642). */
void or_inverse_apply(const int *kinds, const double *params, int nops, double *pts, long n) {
    xf_op *inv = (xf_op *)malloc((nops > 0 ? nops : 1) * sizeof(xf_op));
    for (int i = 0; i < nops; i++) {
        xf_op t;
        memset(&t, 0, sizeof t);
        t.kind = kinds[nops - 1 - i];
        memcpy(t.v, params + 9 * (nops - 1 - i), 9 * sizeof(double));
        inv[i] = xf_inverse(t);
    }
    for (long k = 0; k < n; k++) {
        c3 c = c3_of(pts + 3 * k);
        for (int i = 0; i < nops; i++) c = xf_apply(&inv[i], c);
        pts[3 * k] = c.x, pts[3 * k + 1] = c.y, pts[3 * k + 2] = c.z;
    }
    free(inv);
}

/* EncodeSTL (export.go:20-66) with STLWriter (fileformats/stl.go:17-58):
 * 80 zero bytes, uint32 count, then per triangle float32(Normal()) with
 * Normal = Cross(t1 - t0, t2 - t0).Normalize() in float64
 * (primitives.go:27-33, coords.go:201-222, 267-269, 379-381), the float32
 * vertices and two zero bytes, little-endian.  tris: 9 doubles each. */
static void put_f32(uint8_t *o, float f) {
    uint32_t u;
    memcpy(&u, &f, 4);
    o[0] = (uint8_t)u, o[1] = (uint8_t)(u >> 8), o[2] = (uint8_t)(u >> 16), o[3] = (uint8_t)(u >> 24);
}

void or_encode_stl(const double *tris, long nt, uint8_t *out) {
    memset(out, 0, 80);
    const uint32_t n = (uint32_t)nt;
    out[80] = (uint8_t)n, out[81] = (uint8_t)(n >> 8), out[82] = (uint8_t)(n >> 16), out[83] = (uint8_t)(n >> 24);
    for (long t = 0; t < nt; t++) {
        const double *p = tris + 9 * t;
        c3 a = v3(p[3] + -p[0], p[4] + -p[1], p[5] + -p[2]); /* Sub = Add(Scale(-1)) */
        c3 b = v3(p[6] + -p[0], p[7] + -p[1], p[8] + -p[2]);
        c3 c = v3(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
        double norm = sqrt(c.x * c.x + c.y * c.y + c.z * c.z);
        double s = 1 / norm;
        uint8_t *o = out + 84 + 50 * t;
        put_f32(o, (float)(c.x * s));
        put_f32(o + 4, (float)(c.y * s));
        put_f32(o + 8, (float)(c.z * s));
        for (int k = 0; k < 9; k++) put_f32(o + 12 + 4 * k, (float)p[k]);
        o[48] = o[49] = 0;
    }
}

/* GroupTriangles(faces) = GroupBounders (model3d/bvh.go:118-156, 219-255):
 * three orders of the triangles by the midpoint of their boxes (Triangle
 * Min/Max, primitives.go:35-43; Coord3D.Mid, coords.go:196-198), then the
 * recursive median split along the axis whose two halves' box areas
 * (multipleBoundsArea, boundsArea, bvh.go:257-292) sum least (the first
 * such axis), the other two orders stably partitioned by the halves.  The
 * reference sorts with sort.Slice, which orders equal keys arbitrarily; this
 * restatement breaks ties by the triangle's index (a stable sort), and puts
 * a NaN key last.  perm[k] = the index of the k-th triangle of the grouped
 * order. */
typedef struct {
    double mn[3], mx[3], mid[3];
} grp_box;
static const grp_box *grp_boxes;
static int grp_axis;
static int grp_cmp(const void *a, const void *b) {
    const uint32_t i = *(const uint32_t *)a, j = *(const uint32_t *)b;
    const double x = grp_boxes[i].mid[grp_axis], y = grp_boxes[j].mid[grp_axis];
    const int xn = isnan(x), yn = isnan(y);
    if (xn != yn) return xn - yn;
    if (!xn && x < y) return -1;
    if (!xn && y < x) return 1;
    return i < j ? -1 : (i > j ? 1 : 0);
}
static double grp_area(const grp_box *bx, const uint32_t *ids, long n) {
    double mn[3], mx[3];
    for (int a = 0; a < 3; a++) mn[a] = bx[ids[0]].mn[a], mx[a] = bx[ids[0]].mx[a];
    for (long i = 1; i < n; i++)
        for (int a = 0; a < 3; a++) {
            if (bx[ids[i]].mn[a] < mn[a]) mn[a] = bx[ids[i]].mn[a];
            if (bx[ids[i]].mx[a] > mx[a]) mx[a] = bx[ids[i]].mx[a];
        }
    const double dx = mx[0] + -mn[0], dy = mx[1] + -mn[1], dz = mx[2] + -mn[2];
    return 2 * (dx * (dy + dz) + dy * dz);
}
static void grp_rec(const grp_box *bx, uint32_t *sorted[3], long n, uint32_t *out, uint8_t *flag, uint32_t *tmp) {
    if (n <= 2) {
        for (long i = 0; i < n; i++) out[i] = sorted[0][i];
        return;
    }
    const long mid = n / 2;
    int axis = 0;
    double best = grp_area(bx, sorted[0], mid) + grp_area(bx, sorted[0] + mid, n - mid);
    for (int a = 1; a < 3; a++) {
        const double v = grp_area(bx, sorted[a], mid) + grp_area(bx, sorted[a] + mid, n - mid);
        if (v < best) best = v, axis = a;
    }
    for (long i = 0; i < n; i++) flag[sorted[axis][i]] = i < mid;
    for (int a = 0; a < 3; a++) {
        if (a == axis) continue;
        long i0 = 0, i1 = mid;
        for (long i = 0; i < n; i++) {
            const uint32_t e = sorted[a][i];
            if (flag[e]) tmp[i0++] = e;
            else tmp[i1++] = e;
        }
        memcpy(sorted[a], tmp, n * sizeof(uint32_t));
    }
    uint32_t *lo[3] = {sorted[0], sorted[1], sorted[2]}, *hi[3] = {sorted[0] + mid, sorted[1] + mid, sorted[2] + mid};
    grp_rec(bx, lo, mid, out, flag, tmp);
    grp_rec(bx, hi, n - mid, out + mid, flag, tmp);
}
static double go_min_(double a, double b) {
    if (isinf(a) && a < 0) return a;
    if (isinf(b) && b < 0) return b;
    if (isnan(a) || isnan(b)) return NAN;
    if (a == 0 && a == b) return signbit(a) ? a : b;
    return a < b ? a : b;
}
static double go_max_(double a, double b) {
    if (isinf(a) && a > 0) return a;
    if (isinf(b) && b > 0) return b;
    if (isnan(a) || isnan(b)) return NAN;
    if (a == 0 && a == b) return signbit(a) ? b : a;
    return a > b ? a : b;
}
int or_group_triangles(const double *tris, long nt, uint32_t *perm) {
    if (nt <= 0) return 0;
    grp_box *bx = malloc(nt * sizeof(grp_box));
    uint32_t *s[3], *tmp = malloc(nt * sizeof(uint32_t));
    uint8_t *flag = malloc(nt);
    for (int a = 0; a < 3; a++) s[a] = malloc(nt * sizeof(uint32_t));
    if (!bx || !tmp || !flag || !s[0] || !s[1] || !s[2]) return 4;
    for (long t = 0; t < nt; t++) {
        const double *p = tris + 9 * t;
        for (int a = 0; a < 3; a++) {
            bx[t].mn[a] = go_min_(go_min_(p[a], p[3 + a]), p[6 + a]);
            bx[t].mx[a] = go_max_(go_max_(p[a], p[3 + a]), p[6 + a]);
            bx[t].mid[a] = (bx[t].mn[a] + bx[t].mx[a]) * 0.5;
        }
    }
    grp_boxes = bx;
    for (int a = 0; a < 3; a++) {
        for (long t = 0; t < nt; t++) s[a][t] = (uint32_t)t;
        grp_axis = a;
        qsort(s[a], nt, sizeof(uint32_t), grp_cmp);
    }
    grp_rec(bx, s, nt, perm, flag, tmp);
    free(bx), free(tmp), free(flag);
    for (int a = 0; a < 3; a++) free(s[a]);
    return 0;
}

long or_mesh_ntris(const or_mesh *m) { return m->ntri; }
const double *or_mesh_tris(const or_mesh *m) { return m->tris; }
void or_mesh_info(const or_mesh *m, long *corners, long *vertices, long *search_evals, double *ms_scan,
                  double *ms_search) {
    *corners = m->n_corners;
    *vertices = m->n_vertices;
    *search_evals = m->n_search_evals;
    *ms_scan = m->ms_scan;
    *ms_search = m->ms_search;
}
void or_mesh_free(or_mesh *m) {
    if (!m) return;
    free(m->interior);
    free(m->tris);
    free(m);
}
