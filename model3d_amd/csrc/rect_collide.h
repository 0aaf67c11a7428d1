// rect_collide.h -- MeshToCollider(mesh).RectCollision's float64 tests,
// compiled for the host (filter.cpp: a caller's rect collider, asked from
// the host) and the device (k_filter.hip: the block recursion of
// MarchingCubesFilter / C2F), in the reference's operation order.
//
//   Triangle.RectCollision / SegmentCollision  model3d/primitives.go:181-249, 440-495
//   Segment.RectCollision, NewSegment          model3d/primitives.go:547-554, 632-638
//   rayCollisionWithBounds                     model3d/bvh.go:322-351
//   joinedMultiCollider.RectCollision's box    model3d/collisions.go:344-350
#pragma once

#include <math.h>

#include "mc3d_internal.h"

namespace mc3d {
namespace rc {

struct V {
    double x, y, z;
};
MC3D_HD inline V mk(double x, double y, double z) { return V{x, y, z}; }
MC3D_HD inline V sub(V a, V b) { return mk(a.x + -b.x, a.y + -b.y, a.z + -b.z); }  // Sub = Add(Scale(-1)), coords.go:267-269
MC3D_HD inline double dot(V a, V b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
MC3D_HD inline V cross(V a, V b) { return mk(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x); }
MC3D_HD inline double norm(V a) { return sqrt(a.x * a.x + a.y * a.y + a.z * a.z); }
MC3D_HD inline V scale(V a, double s) { return mk(a.x * s, a.y * s, a.z * s); }
MC3D_HD inline V normalize(V a) { return scale(a, 1 / norm(a)); }
MC3D_HD inline double comp(V a, int i) { return i == 0 ? a.x : (i == 1 ? a.y : a.z); }
// math.Min / math.Max with Go's NaN and signed-zero rules
MC3D_HD inline double go_min(double a, double b) {
    if (isinf(a) && a < 0) return a;
    if (isinf(b) && b < 0) return b;
    if (isnan(a) || isnan(b)) return NAN;
    if (a == 0 && a == b) return signbit(a) ? a : b;
    return a < b ? a : b;
}
MC3D_HD inline double go_max(double a, double b) {
    if (isinf(a) && a > 0) return a;
    if (isinf(b) && b > 0) return b;
    if (isnan(a) || isnan(b)) return NAN;
    if (a == 0 && a == b) return signbit(a) ? b : a;
    return a > b ? a : b;
}

struct Seg {
    V a, b;
};
// NewSegment, primitives.go:547-554
MC3D_HD inline Seg new_segment(V p1, V p2) {
    if (p1.x < p2.x || (p1.x == p2.x && p1.y < p2.y) || (p1.x == p2.x && p1.y == p2.y && p1.z < p2.z)) return {p1, p2};
    return {p2, p1};
}

// rayCollisionWithBounds, bvh.go:322-351
MC3D_HD inline void ray_bounds(V origin, V dir, const double *mn, const double *mx, double &min_frac,
                               double &max_frac) {
    min_frac = -INFINITY, max_frac = INFINITY;
    for (int axis = 0; axis < 3; axis++) {
        const double o = comp(origin, axis), rate = comp(dir, axis);
        if (rate == 0) {
            if (o < mn[axis] || o > mx[axis]) {
                min_frac = 0, max_frac = -1;
                return;
            }
            continue;
        }
        double t1 = (mn[axis] - o) / rate, t2 = (mx[axis] - o) / rate;
        if (t1 > t2) {
            const double t = t1;
            t1 = t2, t2 = t;
        }
        if (t2 < 0) {
            min_frac = 0, max_frac = -1;
            return;
        }
        if (t1 > min_frac) min_frac = t1;
        if (t2 < max_frac) max_frac = t2;
    }
}

// Segment.RectCollision, primitives.go:632-638
MC3D_HD inline bool segment_rect(const Seg &s, const double *mn, const double *mx) {
    double lo, hi;
    ray_bounds(s.a, sub(s.b, s.a), mn, mx, lo, hi);
    return hi >= lo && hi >= 0 && lo <= 1;
}

// Triangle.rayCollision's scale (primitives.go:207-249) for FirstRayCollision
// (181-188): false when there is no collision at scale >= 0.
MC3D_HD inline bool tri_ray(const V *t, V origin, V dir, double &scale_out) {
    const V n = normalize(cross(sub(t[1], t[0]), sub(t[2], t[0])));  // Triangle.Normal
    const double d = dot(n, normalize(dir));
    if (d < 1e-8 && d > -1e-8) return false;
    const V v1 = sub(t[1], t[0]), v2 = sub(t[2], t[0]);
    const V cross1 = cross(dir, v2);
    const double det = dot(cross1, v1);
    if (det == 0) return false;
    const double inv = 1 / det;
    const V o = sub(origin, t[0]);
    const double bary1 = inv * dot(o, cross1);
    if (bary1 < 0 || bary1 > 1) return false;
    const V cross2 = cross(o, v1);
    const double bary2 = inv * dot(dir, cross2);
    if (bary2 < 0 || bary1 + bary2 > 1) return false;
    scale_out = inv * dot(v2, cross2);
    return scale_out >= 0;
}

// Triangle.SegmentCollision, primitives.go:440-445
MC3D_HD inline bool tri_segment(const V *t, const Seg &s) {
    double sc;
    return tri_ray(t, s.a, sub(s.b, s.a), sc) && sc <= 1;
}

// Triangle.RectCollision, primitives.go:448-495
MC3D_HD inline bool tri_rect(const V *t, const double *mn, const double *mx) {
    for (int i = 0; i < 3; i++)
        if (segment_rect(new_segment(t[i], t[(i + 1) % 3]), mn, mx)) return true;
    for (int side = 0; side < 2; side++) {
        const double x = side ? mx[0] : mn[0];
        const Seg segs[4] = {
            new_segment(mk(x, mn[1], mn[2]), mk(x, mn[1], mx[2])),
            new_segment(mk(x, mn[1], mn[2]), mk(x, mx[1], mn[2])),
            new_segment(mk(x, mx[1], mn[2]), mk(x, mx[1], mx[2])),
            new_segment(mk(x, mn[1], mx[2]), mk(x, mx[1], mx[2])),
        };
        for (int k = 0; k < 4; k++)
            if (tri_segment(t, segs[k])) return true;
    }
    const Seg segs[4] = {
        new_segment(mk(mn[0], mn[1], mn[2]), mk(mx[0], mn[1], mn[2])),
        new_segment(mk(mn[0], mx[1], mn[2]), mk(mx[0], mx[1], mn[2])),
        new_segment(mk(mn[0], mn[1], mx[2]), mk(mx[0], mn[1], mx[2])),
        new_segment(mk(mn[0], mx[1], mx[2]), mk(mx[0], mx[1], mx[2])),
    };
    for (int k = 0; k < 4; k++)
        if (tri_segment(t, segs[k])) return true;
    return false;
}

// The bounds test of a joined collider, collisions.go:344-350: the rect and
// the box overlap unless max(rmin, bmin) > min(rmax, bmax) on some axis.
MC3D_HD inline bool box_overlap(const double *rmn, const double *rmx, const double *bmn, const double *bmx) {
    for (int a = 0; a < 3; a++) {
        const double lo = go_max(rmn[a], bmn[a]), hi = go_min(rmx[a], bmx[a]);
        if (go_min(lo, hi) != lo) return false;
    }
    return true;
}

// Triangle.Min / Max (primitives.go:35-43) of 9 doubles.
MC3D_HD inline void tri_box(const double *p, double *tmn, double *tmx) {
    for (int a = 0; a < 3; a++)
        tmn[a] = go_min(go_min(p[a], p[3 + a]), p[6 + a]), tmx[a] = go_max(go_max(p[a], p[3 + a]), p[6 + a]);
}

// One triangle of the mesh against the rect, as the collider's tree asks
// it: its box must overlap the rect (every ancestor box contains it), then
// Triangle.RectCollision.
MC3D_HD inline bool triangle_hits_rect(const double *p, const double *mn, const double *mx) {
    double tmn[3], tmx[3];
    tri_box(p, tmn, tmx);
    if (!box_overlap(mn, mx, tmn, tmx)) return false;
    const V t[3] = {mk(p[0], p[1], p[2]), mk(p[3], p[4], p[5]), mk(p[6], p[7], p[8])};
    return tri_rect(t, mn, mx);
}

// The vertices of a triangle that are the origin of one of its segments
// (NewSegment puts the lexicographically smaller end first): all but the
// largest one (the last of equals -- an equal vertex has the same
// coordinates as the origin it stands for).  None for a triangle with a NaN
// coordinate (NewSegment's order needs ordered coordinates).
MC3D_HD inline int origin_vertices(const double *p, bool *origin) {
    origin[0] = origin[1] = origin[2] = false;
    for (int k = 0; k < 9; k++)
        if (isnan(p[k])) return 0;
    auto less = [&](int j, int k) {
        const double *a = p + 3 * j, *b = p + 3 * k;
        return a[0] < b[0] || (a[0] == b[0] && (a[1] < b[1] || (a[1] == b[1] && a[2] < b[2])));
    };
    int top = 0;
    if (!less(1, top)) top = 1;
    if (!less(2, top)) top = 2;
    for (int k = 0; k < 3; k++) origin[k] = k != top;
    return 2;
}

// A sure hit without Triangle.RectCollision's segment tests: a segment
// origin (origin_vertices) lies in the closed rect.  Segment.RectCollision
// from an origin o inside the rect is true: on every axis (mn - o) <= 0 <=
// (mx - o) exactly, so after the division the entry fraction is <= 0 and the
// exit one >= 0.  (A vertex inside also puts the triangle's box over the
// rect.)
MC3D_HD inline bool vertex_inside(const double *p, const double *mn, const double *mx) {
    bool origin[3];
    origin_vertices(p, origin);
    for (int k = 0; k < 3; k++)
        if (origin[k] && p[3 * k] >= mn[0] && p[3 * k] <= mx[0] && p[3 * k + 1] >= mn[1] && p[3 * k + 1] <= mx[1] &&
            p[3 * k + 2] >= mn[2] && p[3 * k + 2] <= mx[2])
            return true;
    return false;
}

}  // namespace rc
}  // namespace mc3d
