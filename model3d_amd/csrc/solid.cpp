// solid.cpp -- host side of the library: the Solid tree behind the C-ABI
// constructors, its Min/Max in the reference's float64 arithmetic, the
// marching-cubes case table and spacer, and the lowering of a tree to the
// flat device program the kernels evaluate.
//
// Bit parity rests on doing every float64 operation of the reference in the
// same order (compiled with -ffp-contract=off), either here (constants that
// the reference recomputes per Contains call but that do not depend on the
// query point) or on the device (everything that does).

#include <math.h>
#include <string.h>

#include <algorithm>
#include <array>
#include <atomic>
#include <initializer_list>
#include <mutex>

#include "mc3d_internal.h"

namespace mc3d {

namespace {

struct V3 {
    double x, y, z;
};
inline V3 mk(double x, double y, double z) { return V3{x, y, z}; }
inline V3 of(const double *a) { return V3{a[0], a[1], a[2]}; }
inline V3 add(V3 a, V3 b) { return mk(a.x + b.x, a.y + b.y, a.z + b.z); }
inline V3 scale(V3 a, double s) { return mk(a.x * s, a.y * s, a.z * s); }
// Coord3D.Sub is Add(Scale(-1)) (coords.go:267-269); IEEE gives the same bits
// as a plain subtraction, but keep the reference's form.
inline V3 sub(V3 a, V3 b) { return add(a, scale(b, -1)); }
inline V3 mul(V3 a, V3 b) { return mk(a.x * b.x, a.y * b.y, a.z * b.z); }
inline V3 adds(V3 a, double s) { return mk(a.x + s, a.y + s, a.z + s); }
inline double dot(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
inline double norm(V3 a) { return sqrt(a.x * a.x + a.y * a.y + a.z * a.z); }
inline V3 normalize(V3 a) { return scale(a, 1 / norm(a)); }

// Go math.Min / math.Max (special cases for infinities, NaN, signed zero).
inline double gmin(double x, double y) {
    if (isinf(x) && x < 0) return x;
    if (isinf(y) && y < 0) return y;
    if (isnan(x) || isnan(y)) return NAN;
    if (x == 0 && x == y) return signbit(x) ? x : y;
    return x < y ? x : y;
}
inline double gmax(double x, double y) {
    if (isinf(x) && x > 0) return x;
    if (isinf(y) && y > 0) return y;
    if (isnan(x) || isnan(y)) return NAN;
    if (x == 0 && x == y) return signbit(x) ? y : x;
    return x > y ? x : y;
}
inline V3 vmin(V3 a, V3 b) { return mk(gmin(a.x, b.x), gmin(a.y, b.y), gmin(a.z, b.z)); }
inline V3 vmax(V3 a, V3 b) { return mk(gmax(a.x, b.x), gmax(a.y, b.y), gmax(a.z, b.z)); }

// Coord3D.OrthoBasis, coords.go:388-421.
void ortho_basis(V3 c, V3 &o1, V3 &o2) {
    double ax = fabs(c.x), ay = fabs(c.y), az = fabs(c.z);
    V3 b1 = mk(0, 0, 0);
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
    V3 b2 = mk(b1.y * c.z - b1.z * c.y, b1.z * c.x - b1.x * c.z, b1.x * c.y - b1.y * c.x);
    o1 = normalize(b1);
    o2 = normalize(b2);
}

// circleAxisBound, shapes.go:577-585 (ProjectOut, coords.go:424-427).
double circle_axis_bound(int axis, V3 normal, double sign) {
    double arr[3] = {0, 0, 0};
    arr[axis] = sign;
    V3 c = of(arr);
    V3 normed = normalize(normal);
    V3 proj = sub(c, scale(normed, dot(normed, c)));
    proj = scale(proj, 1 / (norm(proj) + 1e-8));
    double pa[3] = {proj.x, proj.y, proj.z};
    return sign * (fabs(pa[axis]) + 1e-8);
}
V3 circle_bounds(V3 axis, double sign) {
    return mk(circle_axis_bound(0, axis, sign), circle_axis_bound(1, axis, sign),
              circle_axis_bound(2, axis, sign));
}

// Matrix3 helpers, matrix.go:59-137.
V3 mat_col(const double *m, V3 c) {
    return mk(m[0] * c.x + m[1] * c.y + m[2] * c.z, m[3] * c.x + m[4] * c.y + m[5] * c.z,
              m[6] * c.x + m[7] * c.y + m[8] * c.z);
}
void mat_inverse(const double *m, double *out) {
    double det = m[0] * (m[4] * m[8] - m[5] * m[7]) - m[1] * (m[3] * m[8] - m[5] * m[6]) +
                 m[2] * (m[3] * m[7] - m[4] * m[6]);
    double r[9] = {
        m[4] * m[8] - m[5] * m[7], m[2] * m[7] - m[1] * m[8], m[1] * m[5] - m[2] * m[4],
        m[5] * m[6] - m[3] * m[8], m[0] * m[8] - m[2] * m[6], m[2] * m[3] - m[0] * m[5],
        m[3] * m[7] - m[4] * m[6], m[1] * m[6] - m[0] * m[7], m[0] * m[4] - m[1] * m[3],
    };
    double s = 1 / det;
    for (int i = 0; i < 9; i++) out[i] = r[i] * s;
}

struct Xf {
    int kind;
    double v[9];
};

// Inverse of one elementary transform, transform.go:52-54, 83-85, 114-116, 153-155.
Xf xf_inverse(const Xf &t) {
    Xf r;
    memset(&r, 0, sizeof r);
    r.kind = t.kind;
    switch (t.kind) {
    case MC3D_XF_TRANSLATE: {
        V3 o = scale(of(t.v), -1);
        r.v[0] = o.x, r.v[1] = o.y, r.v[2] = o.z;
        break;
    }
    case MC3D_XF_SCALE: r.v[0] = 1 / t.v[0]; break;
    case MC3D_XF_VECSCALE: r.v[0] = 1 / t.v[0], r.v[1] = 1 / t.v[1], r.v[2] = 1 / t.v[2]; break;
    case MC3D_XF_MATRIX: mat_inverse(t.v, r.v); break;
    }
    return r;
}

// Transform.ApplyBounds, transform.go:48-50, 75-81, 107-111, 133-150.
void xf_apply_bounds(const Xf &t, V3 &mn, V3 &mx) {
    V3 a = mn, b = mx;
    switch (t.kind) {
    case MC3D_XF_TRANSLATE: mn = add(a, of(t.v)), mx = add(b, of(t.v)); break;
    case MC3D_XF_SCALE:
        if (t.v[0] < 0) mn = scale(b, t.v[0]), mx = scale(a, t.v[0]);
        else mn = scale(a, t.v[0]), mx = scale(b, t.v[0]);
        break;
    case MC3D_XF_VECSCALE: {
        V3 p = mul(a, of(t.v)), q = mul(b, of(t.v));
        mn = vmin(p, q);
        mx = vmax(q, p);
        break;
    }
    case MC3D_XF_MATRIX: {
        double xs[2] = {a.x, b.x}, ys[2] = {a.y, b.y}, zs[2] = {a.z, b.z};
        V3 lo = mk(0, 0, 0), hi = mk(0, 0, 0);
        for (int i = 0; i < 2; i++)
            for (int j = 0; j < 2; j++)
                for (int k = 0; k < 2; k++) {
                    V3 c = mat_col(t.v, mk(xs[i], ys[j], zs[k]));
                    if (i == 0 && j == 0 && k == 0) lo = hi = c;
                    else lo = vmin(lo, c), hi = vmax(hi, c);
                }
        mn = lo, mx = hi;
        break;
    }
    }
}

// Largest double s with fl(sqrt(s)) <= r, so that "sqrt(s) <= r" (the
// reference's Dist(...) <= r) is decided exactly by "s <= T" without a device
// sqrt.  fl o sqrt is monotone, so the accepted set is a prefix of [0, inf].
// Sums of squares are never negative; NaN fails both forms.
double sqrt_le_threshold(double r) {
    if (isnan(r) || r < 0) return -1.0;
    if (isinf(r)) return INFINITY;
    double s = r * r;
    while (s < INFINITY && sqrt(nextafter(s, INFINITY)) <= r) s = nextafter(s, INFINITY);
    while (sqrt(s) > r) s = nextafter(s, -INFINITY);
    return s;
}

// ---------------------------------------------------------------------------
// Float32 guard bands (eval.h, atom32).
//
// Each guarded atom decides "s <= T" where s is a sum of three squares d_i^2
// and T the float64 threshold.  In float32 every input is rounded (relative
// error u = 2^-24) and every operation rounds once.  For a point within 2r of
// the primitive ("near regime") the absolute error of each d_i is at most
// e_i (derived per primitive below), so
//   |q32 - s| <= E = sum_i [e_i (4r + e_i) + u (2r + e_i)^2] + 2.2 u * 12 r^2
// (squares of the perturbed d_i, their rounding, and the two additions over a
// sum of at most 12 r^2).  "Inside" is decided only when q32 <= T - E, which
// forces the float32 d_i below ~r, hence the near regime, hence s <= T; an
// inside point is in the near regime, so q32 <= T + E and it is never
// decided "outside".  A relative 1e-12 T absorbs the float64 rounding of the
// reference's own s.  Branch tests (capsule end caps, cylinder slab) get a
// margin e_br on the projection; inside the margin the atom is re-decided in
// float64.  Thresholds are rounded outward when stored as float.
// ---------------------------------------------------------------------------
const double kU = 1.0 / 16777216.0;  // 2^-24

double band_E(const double e[3], double r, double T) {
    double E = 0;
    for (int i = 0; i < 3; i++) E += e[i] * (4 * r + e[i]) + kU * (2 * r + e[i]) * (2 * r + e[i]);
    return E + 2.2 * kU * 12 * r * r + 1e-12 * fabs(T) + 1e-37;
}

// True when float32 evaluation is safe from overflow/underflow trouble.
bool guard_ok(double mag, double r) { return isfinite(mag) && isfinite(r) && r > 1e-15 && mag < 1e15 && r < 1e15; }

}  // namespace

// ---------------------------------------------------------------------------
// Solid tree
// ---------------------------------------------------------------------------
enum Kind {
    K_SPHERE = 1, K_RECT, K_CAPSULE, K_CYLINDER, K_CONE, K_CONESLICE, K_TORUS,
    K_JOINED, K_SUBTRACTED, K_INTERSECTED, K_TRANSFORM, K_FORCEBOUNDS, K_COLLIDER, K_HOST,
};

struct Solid {
    Kind kind;
    double p[8] = {0};
    std::vector<SolidPtr> children;
    std::vector<Xf> fwd;              // K_TRANSFORM: the forward JoinedTransform
    double mn[3] = {0}, mx[3] = {0};  // stored bounds (funcSolid / ColliderSolid)
    std::shared_ptr<ColliderData> collider;
    HostAtom host{nullptr, nullptr};
    // Stack entries the node's code needs, emitting the needier operand
    // first (set when the node is made; its children are made before it).
    int need = 1;
};

// Solid::need from the children's.
static int node_need(const Solid &s) {
    switch (s.kind) {
    case K_JOINED:
    case K_INTERSECTED: {
        // operands in decreasing need; each after the first sits on one
        // accumulated entry: max(n0, n1 + 1) over the two largest
        int n0 = 0, n1 = -1;
        for (auto &c : s.children) {
            if (c->need > n0) n1 = n0, n0 = c->need;
            else if (c->need > n1) n1 = c->need;
        }
        return s.children.size() > 1 ? std::max(n0, n1 + 1) : n0;
    }
    case K_SUBTRACTED: {
        const int a = s.children[0]->need, b = s.children[1]->need;
        return a == b ? a + 1 : std::max(a, b);
    }
    case K_TRANSFORM:
    case K_FORCEBOUNDS: return std::max(s.children[0]->need, 2);
    default: return 1;
    }
}

static void min_max(const Solid &s, V3 &mn, V3 &mx) {
    const double *p = s.p;
    switch (s.kind) {
    case K_SPHERE:  // shapes.go:19-26
        mn = adds(of(p), -p[3]), mx = adds(of(p), p[3]);
        return;
    case K_RECT:  // shapes.go:161-168
        mn = of(p), mx = of(p + 3);
        return;
    case K_CAPSULE:  // shapes.go:355-362
        mn = adds(vmin(of(p), of(p + 3)), -p[6]);
        mx = adds(vmax(of(p), of(p + 3)), p[6]);
        return;
    case K_CYLINDER: {  // shapes.go:554-575
        V3 p1 = of(p), p2 = of(p + 3), axis = sub(p2, p1);
        mn = add(vmin(p1, p2), scale(circle_bounds(axis, -1), p[6]));
        mx = add(vmax(p1, p2), scale(circle_bounds(axis, 1), p[6]));
        return;
    }
    case K_CONE: {  // shapes.go:867-886
        V3 tip = of(p), base = of(p + 3), axis = sub(tip, base);
        mn = vmin(add(scale(circle_bounds(axis, -1), p[6]), base), tip);
        mx = vmax(add(scale(circle_bounds(axis, 1), p[6]), base), tip);
        return;
    }
    case K_CONESLICE: {  // shapes.go:1048-1072
        V3 p1 = of(p), p2 = of(p + 3), axis = sub(p2, p1);
        V3 d = circle_bounds(axis, -1);
        mn = vmin(add(scale(d, p[6]), p1), add(scale(d, p[7]), p2));
        d = circle_bounds(axis, 1);
        mx = vmax(add(scale(d, p[6]), p1), add(scale(d, p[7]), p2));
        return;
    }
    case K_TORUS: {  // shapes.go:1342-1361
        V3 center = of(p), axis = of(p + 3);
        V3 extra = mk(p[7], p[7], p[7]);
        mn = sub(add(scale(circle_bounds(axis, -1), p[6]), center), extra);
        mx = add(add(scale(circle_bounds(axis, 1), p[6]), center), extra);
        return;
    }
    case K_JOINED: {  // solid.go:73-87
        V3 a, b;
        min_max(*s.children[0], mn, mx);
        for (size_t i = 1; i < s.children.size(); i++) {
            min_max(*s.children[i], a, b);
            mn = vmin(mn, a), mx = vmax(mx, b);
        }
        return;
    }
    case K_SUBTRACTED:  // solid.go:130-136
        min_max(*s.children[0], mn, mx);
        return;
    case K_INTERSECTED: {  // solid.go:146-161
        V3 a, b, lo, hi;
        min_max(*s.children[0], lo, hi);
        for (size_t i = 1; i < s.children.size(); i++) {
            min_max(*s.children[i], a, b);
            lo = vmax(lo, a), hi = vmin(hi, b);
        }
        mn = lo, mx = vmax(hi, lo);
        return;
    }
    default:  // funcSolid (solid.go:52-58), ColliderSolid (solid.go:282-289), host solid
        mn = of(s.mn), mx = of(s.mx);
        return;
    }
}

void solid_bounds(const Solid &s, double mn[3], double mx[3]) {
    V3 a, b;
    min_max(s, a, b);
    mn[0] = a.x, mn[1] = a.y, mn[2] = a.z;
    mx[0] = b.x, mx[1] = b.y, mx[2] = b.z;
}

// BoundsValid, bounder.go:30-46.
bool bounds_valid(const Solid &s) {
    V3 a, b;
    min_max(s, a, b);
    double sa = a.x + a.y + a.z, sb = b.x + b.y + b.z;
    if (isnan(sa) || isnan(sb) || isinf(sa) || isinf(sb)) return false;
    return !(b.x < a.x || b.y < a.y || b.z < a.z);
}

// ---------------------------------------------------------------------------
// Case table, mc.go:285-586.  The 23 rotation classes below are the
// reference's baseTriangleTable (from "A survey of the marching cubes
// algorithm", 2006): inside-corner mask, triangle count, 6 corner ids per
// triangle.  Every key is reached from exactly one class (rotations keep the
// inside-corner count and the classes are distinct orbits), so the Go map
// iteration order of the classes cannot matter; within a class the first
// rotation in lexicographic order wins.
// ---------------------------------------------------------------------------
namespace {
struct BaseCase {
    uint8_t mask, ntri;
    uint8_t t[5][6];
};
#define B(a) (1u << (a))
const BaseCase kBase[] = {
    {0, 0, {{0}}},
    {B(0), 1, {{0, 1, 0, 2, 0, 4}}},
    {B(0) | B(1), 2, {{0, 4, 1, 5, 0, 2}, {1, 5, 1, 3, 0, 2}}},
    {B(0) | B(5), 2, {{0, 1, 0, 2, 0, 4}, {5, 7, 1, 5, 4, 5}}},
    {B(0) | B(7), 2, {{0, 1, 0, 2, 0, 4}, {6, 7, 3, 7, 5, 7}}},
    {B(1) | B(2) | B(3), 3, {{0, 1, 1, 5, 0, 2}, {0, 2, 1, 5, 2, 6}, {2, 6, 1, 5, 3, 7}}},
    {B(0) | B(1) | B(7), 3, {{0, 4, 1, 5, 0, 2}, {1, 5, 1, 3, 0, 2}, {6, 7, 3, 7, 5, 7}}},
    {B(1) | B(4) | B(7), 3, {{4, 6, 4, 5, 0, 4}, {1, 5, 1, 3, 0, 1}, {6, 7, 3, 7, 5, 7}}},
    {B(0) | B(1) | B(2) | B(3), 2, {{0, 4, 1, 5, 3, 7}, {0, 4, 3, 7, 2, 6}}},
    {B(0) | B(2) | B(3) | B(6), 4, {{0, 1, 4, 6, 0, 4}, {0, 1, 6, 7, 4, 6}, {0, 1, 1, 3, 6, 7}, {1, 3, 3, 7, 6, 7}}},
    {B(1) | B(2) | B(5) | B(6), 4, {{0, 2, 2, 3, 6, 7}, {0, 2, 6, 7, 4, 6}, {0, 1, 4, 5, 5, 7}, {5, 7, 1, 3, 0, 1}}},
    {B(0) | B(2) | B(3) | B(7), 4, {{0, 4, 0, 1, 2, 6}, {0, 1, 5, 7, 2, 6}, {2, 6, 5, 7, 6, 7}, {0, 1, 1, 3, 5, 7}}},
    {B(1) | B(2) | B(3) | B(4), 4, {{0, 1, 1, 5, 0, 2}, {0, 2, 1, 5, 2, 6}, {2, 6, 1, 5, 3, 7}, {4, 5, 0, 4, 4, 6}}},
    {B(1) | B(2) | B(4) | B(7), 4, {{0, 1, 1, 5, 1, 3}, {0, 2, 2, 3, 2, 6}, {4, 5, 0, 4, 4, 6}, {5, 7, 6, 7, 3, 7}}},
    {B(1) | B(2) | B(3) | B(6), 4, {{0, 2, 0, 1, 4, 6}, {0, 1, 3, 7, 4, 6}, {0, 1, 1, 5, 3, 7}, {4, 6, 3, 7, 6, 7}}},
    {B(0) | B(2) | B(3) | B(5) | B(6), 5,
     {{0, 1, 4, 6, 0, 4}, {0, 1, 6, 7, 4, 6}, {0, 1, 1, 3, 6, 7}, {1, 3, 3, 7, 6, 7}, {5, 7, 1, 5, 4, 5}}},
    {B(2) | B(3) | B(4) | B(5) | B(6), 5,
     {{5, 7, 1, 5, 0, 4}, {0, 4, 6, 7, 5, 7}, {0, 2, 6, 7, 0, 4}, {0, 2, 3, 7, 6, 7}, {0, 2, 1, 3, 3, 7}}},
    {B(0) | B(4) | B(5) | B(6) | B(7), 3, {{1, 5, 0, 1, 0, 2}, {0, 2, 2, 6, 1, 5}, {1, 5, 2, 6, 3, 7}}},
    {B(1) | B(2) | B(3) | B(4) | B(5) | B(6), 2, {{0, 2, 0, 1, 0, 4}, {3, 7, 6, 7, 5, 7}}},
    {B(1) | B(2) | B(3) | B(4) | B(6) | B(7), 4, {{0, 2, 4, 5, 0, 4}, {0, 2, 5, 7, 4, 5}, {0, 2, 1, 5, 5, 7}, {0, 1, 1, 5, 0, 2}}},
    {B(2) | B(3) | B(4) | B(5) | B(6) | B(7), 2, {{1, 5, 0, 4, 0, 2}, {1, 3, 1, 5, 0, 2}}},
    {B(1) | B(2) | B(3) | B(4) | B(5) | B(6) | B(7), 1, {{0, 2, 0, 1, 0, 4}}},
    {0xff, 0, {{0}}},
};
#undef B

// Cube edge id of a corner pair: axis = the differing bit, then the two
// remaining corner bits (low axis first).
int edge_id(int a, int b) {
    int d = a ^ b;
    int lo = a & b;
    if (d == 1) return 0 + ((lo >> 1) & 3);
    if (d == 2) return 4 + ((lo & 1) | (((lo >> 2) & 1) << 1));
    if (d == 4) return 8 + (lo & 3);
    return -1;
}

uint8_t g_table[256 * 16];
std::once_flag g_table_once;

void build_case_table() {
    // 24 rotations: closure of the two generators under composition
    // (mc.go:312-330), sorted lexicographically (mc.go:339-350).
    typedef std::array<uint8_t, 8> Rot;
    const Rot zr = {2, 0, 3, 1, 6, 4, 7, 5};
    const Rot xr = {2, 3, 6, 7, 0, 1, 4, 5};
    std::vector<Rot> rots = {{0, 1, 2, 3, 4, 5, 6, 7}};
    for (size_t qi = 0; qi < rots.size(); qi++) {
        for (const Rot &op : {zr, xr}) {
            Rot r;
            for (int i = 0; i < 8; i++) r[i] = op[rots[qi][i]];  // Compose, mc.go:356-362
            if (std::find(rots.begin(), rots.end(), r) == rots.end()) rots.push_back(r);
        }
    }
    std::sort(rots.begin(), rots.end());
    bool set[256] = {false};
    memset(g_table, 0xff, sizeof g_table);
    for (const BaseCase &bc : kBase) {
        for (const Rot &rot : rots) {
            int key = 0;
            for (int c = 0; c < 8; c++)
                if (bc.mask & (1u << c)) key |= 1 << rot[c];  // ApplyIntersections, mc.go:380-388
            if (set[key]) continue;
            set[key] = true;
            uint8_t *row = g_table + 16 * key;
            row[0] = bc.ntri;
            for (int t = 0; t < bc.ntri; t++)
                for (int e = 0; e < 3; e++)
                    row[1 + 3 * t + e] = (uint8_t)edge_id(rot[bc.t[t][2 * e]], rot[bc.t[t][2 * e + 1]]);
        }
    }
}
}  // namespace

const uint8_t *case_table() {
    std::call_once(g_table_once, build_case_table);
    return g_table;
}

std::vector<double> spacer_axis(double lo, double hi, double delta) {
    std::vector<double> v;
    for (double x = lo - delta; x <= hi + delta; x += delta) {
        v.push_back(x);
        if (v.size() > (1u << 22)) break;  // caller rejects grids this large
    }
    return v;
}

// ---------------------------------------------------------------------------
// Ray grid of a mesh collider
// ---------------------------------------------------------------------------
namespace {
// Fixed direction of ColliderContains, collisions.go:120-125.
const double kRayDir[3] = {0.5224892708603626, 0.10494477243214506, 0.43558938446126527};

inline V3 cross(V3 a, V3 b) { return mk(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x); }
}  // namespace

const RayGrid &ColliderData::ray_grid() const {
    std::call_once(grid_once, [this] {
        RayGrid &G = grid;
        const V3 D = of(kRayDir);
        const V3 dn = normalize(D);
        // Any orthonormal basis perpendicular to the ray serves the binning.
        V3 u, v;
        ortho_basis(dn, u, v);
        G.u[0] = u.x, G.u[1] = u.y, G.u[2] = u.z;
        G.v[0] = v.x, G.v[1] = v.y, G.v[2] = v.z;
        const size_t nt = tris.size() / 9;
        struct Box2 {
            double lo[2], hi[2];
        };
        std::vector<Box2> boxes;
        boxes.reserve(nt);
        double ulo = INFINITY, uhi = -INFINITY, vlo = INFINITY, vhi = -INFINITY;
        for (size_t i = 0; i < nt; i++) {
            const double *t = &tris[9 * i];
            const V3 t0 = of(t), t1 = of(t + 3), t2 = of(t + 6);
            // Triangle.rayCollision's point-independent part (primitives.go:207-230):
            // the parallel test, v1, v2, cross1, det; such triangles never hit.
            const V3 nrm = normalize(cross(sub(t1, t0), sub(t2, t0)));
            const double dd = dot(nrm, dn);
            if (dd < 1e-8 && dd > -1e-8) continue;
            const V3 v1 = sub(t1, t0), v2 = sub(t2, t0);
            const V3 c1 = cross(D, v2);
            const double det = dot(c1, v1);
            if (det == 0) continue;
            const double inv_det = 1 / det;
            const double r[RAY_REC] = {t0.x, t0.y, t0.z, c1.x, c1.y, c1.z, inv_det, v1.x, v1.y, v1.z, v2.x, v2.y, v2.z};
            G.rec.insert(G.rec.end(), r, r + RAY_REC);
            // Projected box, widened far beyond the float64 error of a hit
            // (barycentrics lose ~1/|n.d| digits for grazing triangles).
            Box2 b{{INFINITY, INFINITY}, {-INFINITY, -INFINITY}};
            double mag = 0;
            for (const V3 &p : {t0, t1, t2}) {
                const double pu = dot(p, u), pv = dot(p, v);
                b.lo[0] = std::min(b.lo[0], pu), b.hi[0] = std::max(b.hi[0], pu);
                b.lo[1] = std::min(b.lo[1], pv), b.hi[1] = std::max(b.hi[1], pv);
                mag = std::max(mag, std::max(fabs(p.x), std::max(fabs(p.y), fabs(p.z))));
            }
            const double ext = std::max(b.hi[0] - b.lo[0], b.hi[1] - b.lo[1]);
            const double w = ext * std::max(1e-6, 1e-13 / fabs(dd)) + 1e-9 * mag + 1e-300;
            b.lo[0] -= w, b.lo[1] -= w, b.hi[0] += w, b.hi[1] += w;
            boxes.push_back(b);
            ulo = std::min(ulo, b.lo[0]), uhi = std::max(uhi, b.hi[0]);
            vlo = std::min(vlo, b.lo[1]), vhi = std::max(vhi, b.hi[1]);
        }
        const size_t nk = boxes.size();
        if (nk == 0) {
            G.u0 = G.v0 = 0, G.inv_cell = 1, G.nu = G.nv = 1;
            G.cell_start.assign(2, 0);
            return;
        }
        // About two cells per kept triangle, square cells, at most 4096 a side.
        const double su = std::max(uhi - ulo, 1e-300), sv = std::max(vhi - vlo, 1e-300);
        double cell = sqrt(su * sv / (2.0 * (double)nk));
        cell = std::max(cell, std::max(su, sv) / 4095.0);
        G.u0 = ulo, G.v0 = vlo, G.inv_cell = 1 / cell;
        // Every widened box maps below nu (same expression as the lookup).
        G.nu = (int)((uhi - ulo) * G.inv_cell) + 1;
        G.nv = (int)((vhi - vlo) * G.inv_cell) + 1;
        auto cell_range = [&](double lo, double hi, double o, int n, int &a, int &b) {
            // the device's lookup: floor((p - o) * inv_cell), monotone in p
            double fa = (lo - o) * G.inv_cell, fb = (hi - o) * G.inv_cell;
            a = (int)std::max(0.0, std::min((double)(n - 1), floor(fa)));
            b = (int)std::max(0.0, std::min((double)(n - 1), floor(fb)));
        };
        std::vector<uint32_t> count((size_t)G.nu * G.nv + 1, 0);
        for (size_t i = 0; i < nk; i++) {
            int a0, a1, b0, b1;
            cell_range(boxes[i].lo[0], boxes[i].hi[0], G.u0, G.nu, a0, a1);
            cell_range(boxes[i].lo[1], boxes[i].hi[1], G.v0, G.nv, b0, b1);
            for (int cv = b0; cv <= b1; cv++)
                for (int cu = a0; cu <= a1; cu++) count[(size_t)cv * G.nu + cu]++;
        }
        G.cell_start.assign(count.size(), 0);
        for (size_t c = 0; c + 1 < count.size(); c++) G.cell_start[c + 1] = G.cell_start[c] + count[c];
        G.cell_tris.resize(G.cell_start.back());
        std::vector<uint32_t> fill(G.cell_start.begin(), G.cell_start.end() - 1);
        for (size_t i = 0; i < nk; i++) {
            int a0, a1, b0, b1;
            cell_range(boxes[i].lo[0], boxes[i].hi[0], G.u0, G.nu, a0, a1);
            cell_range(boxes[i].lo[1], boxes[i].hi[1], G.v0, G.nv, b0, b1);
            for (int cv = b0; cv <= b1; cv++)
                for (int cu = a0; cu <= a1; cu++) G.cell_tris[fill[(size_t)cv * G.nu + cu]++] = (uint32_t)i;
        }
    });
    return grid;
}

// ---------------------------------------------------------------------------
// Sphere grid of a mesh collider
// ---------------------------------------------------------------------------
const SphereGrid &ColliderData::sphere_grid(double R) const {
    std::lock_guard<std::mutex> lk(sphere_mu);
    for (const auto &g : spheres)
        if (g->R == R) return *g;
    std::unique_ptr<SphereGrid> G(new SphereGrid);
    G->R = R;
    const size_t nt = tris.size() / 9;
    G->rec.resize(nt * SPH_REC, 0.0);
    struct B3 {
        double lo[3], hi[3];
    };
    std::vector<B3> boxes(nt);
    double glo[3] = {INFINITY, INFINITY, INFINITY}, ghi[3] = {-INFINITY, -INFINITY, -INFINITY};
    double ext_sum = 0;
    for (size_t i = 0; i < nt; i++) {
        const double *t = &tris[9 * i];
        const V3 t0 = of(t), t1 = of(t + 3), t2 = of(t + 6);
        double *r = &G->rec[SPH_REC * i];
        memcpy(r, t, 9 * sizeof(double));
        // Triangle.rayCollision along Triangle.Normal() (primitives.go:207-230,
        // 263-269): the parallel test, v1, v2, cross1, det do not depend on
        // the point.
        const V3 nrm = normalize(cross(sub(t1, t0), sub(t2, t0)));
        const double d = dot(nrm, normalize(nrm));
        const V3 v1 = sub(t1, t0), v2 = sub(t2, t0);
        const V3 c1 = cross(nrm, v2);
        const double det = dot(c1, v1);
        const bool hit = !(d < 1e-8 && d > -1e-8) && det != 0;
        const double vals[15] = {nrm.x, nrm.y, nrm.z, hit ? 1.0 : 0.0, hit ? 1 / det : 0.0, c1.x, c1.y, c1.z,
                                 v1.x,  v1.y,  v1.z,  v2.x,           v2.y,               v2.z, 0.0};
        memcpy(r + 9, vals, sizeof vals);
        // the box grown by R, and by far more than the rounding of a test
        B3 b;
        double mag = 0;
        for (int a = 0; a < 3; a++) {
            const double lo = std::min(t[a], std::min(t[3 + a], t[6 + a]));
            const double hi = std::max(t[a], std::max(t[3 + a], t[6 + a]));
            mag = std::max(mag, std::max(fabs(lo), fabs(hi)));
            b.lo[a] = lo, b.hi[a] = hi;
            ext_sum += hi - lo;
        }
        const double w = R * (1 + 1e-6) + 1e-9 * (mag + R) + 1e-300;
        for (int a = 0; a < 3; a++) {
            b.lo[a] -= w, b.hi[a] += w;
            glo[a] = std::min(glo[a], b.lo[a]), ghi[a] = std::max(ghi[a], b.hi[a]);
        }
        boxes[i] = b;
    }
    if (nt) {
        // Cubic cells about the size of a grown triangle box, at most 2^21
        // cells (a point outside the grid is farther than R from every
        // triangle).
        const double mean_ext = ext_sum / (3.0 * (double)nt);
        double cell = std::max(mean_ext + 2 * R, 1e-300);
        const double span = std::max(ghi[0] - glo[0], std::max(ghi[1] - glo[1], ghi[2] - glo[2]));
        if (std::isfinite(span) && std::isfinite(cell)) {
            for (;;) {
                double cells = 1;
                for (int a = 0; a < 3; a++) cells *= std::floor((ghi[a] - glo[a]) / cell) + 1;
                if (cells <= (double)(1u << 21)) break;
                cell *= 1.25;
            }
            cell = std::max(cell, span / 1e6);
            G->inv_cell = 1 / cell;
            for (int a = 0; a < 3; a++) {
                G->o[a] = glo[a];
                G->n[a] = (int)((ghi[a] - glo[a]) * G->inv_cell) + 1;  // every grown box maps below n
            }
        } else {
            // An infinite radius (or grown boxes beyond the float64 range):
            // one cell with every triangle, which every finite point maps to
            // ((p - 0) * 0 == 0); a non-finite point maps to NaN and is
            // outside, as it is farther than any radius from every triangle.
            G->inv_cell = 0;
            for (int a = 0; a < 3; a++) G->o[a] = 0, G->n[a] = 1;
        }
        auto range = [&](double lo, double hi, int a, int &i0, int &i1) {
            // the device's lookup: floor((p - o) * inv_cell), monotone in p
            const double fa = (lo - G->o[a]) * G->inv_cell, fb = (hi - G->o[a]) * G->inv_cell;
            i0 = (int)std::max(0.0, std::min((double)(G->n[a] - 1), floor(fa)));
            i1 = (int)std::max(0.0, std::min((double)(G->n[a] - 1), floor(fb)));
        };
        const size_t ncell = (size_t)G->n[0] * G->n[1] * G->n[2];
        std::vector<uint32_t> count(ncell + 1, 0);
        for (int pass = 0; pass < 2; pass++) {
            if (pass == 1) {
                G->cell_start.assign(ncell + 1, 0);
                for (size_t c = 0; c < ncell; c++) G->cell_start[c + 1] = G->cell_start[c] + count[c];
                G->cell_tris.resize(G->cell_start.back());
                std::copy(G->cell_start.begin(), G->cell_start.end() - 1, count.begin());
            }
            for (size_t i = 0; i < nt; i++) {
                int r0[3], r1[3];
                for (int a = 0; a < 3; a++) range(boxes[i].lo[a], boxes[i].hi[a], a, r0[a], r1[a]);
                for (int z = r0[2]; z <= r1[2]; z++)
                    for (int y = r0[1]; y <= r1[1]; y++)
                        for (int x = r0[0]; x <= r1[0]; x++) {
                            const size_t c = ((size_t)z * G->n[1] + y) * G->n[0] + x;
                            if (pass == 0) count[c]++;
                            else G->cell_tris[count[c]++] = (uint32_t)i;
                        }
            }
        }
    } else {
        G->cell_start.assign(2, 0);
    }
    spheres.push_back(std::move(G));
    return *spheres.back();
}

// ---------------------------------------------------------------------------
// Lowering
// ---------------------------------------------------------------------------
namespace {

struct Box {
    double lo[3], hi[3];
};
const double kInf = INFINITY;
Box box_all() { return Box{{-kInf, -kInf, -kInf}, {kInf, kInf, kInf}}; }

// Widen a box so that rounding in the reference's float64 Contains (relative
// errors ~1e-15) can never make an atom true outside it.
Box widen(Box b) {
    for (int a = 0; a < 3; a++) {
        double m = std::max(fabs(b.lo[a]), fabs(b.hi[a]));
        double pad = 1e-7 * (b.hi[a] - b.lo[a]) + 1e-9 * m + 1e-300;
        b.lo[a] -= pad, b.hi[a] += pad;
        if (!(b.lo[a] <= b.hi[a]) || isnan(pad)) b.lo[a] = -kInf, b.hi[a] = kInf;
    }
    return b;
}

// World-space image of a local box under the forward transforms of the frame
// chain (innermost level first).
Box to_world(Box b, const std::vector<std::vector<Xf>> &levels) {
    for (size_t li = levels.size(); li-- > 0;) {
        for (const Xf &t : levels[li]) {
            V3 mn = of(b.lo), mx = of(b.hi);
            if (t.kind == MC3D_XF_MATRIX) {
                // Reject ill-conditioned matrices: cull nothing below them.
                double inv[9];
                mat_inverse(t.v, inv);
                double n1 = 0, n2 = 0;
                for (int i = 0; i < 9; i++) n1 = std::max(n1, fabs(t.v[i])), n2 = std::max(n2, fabs(inv[i]));
                if (!(n1 * n2 < 1e6)) return box_all();
            }
            xf_apply_bounds(t, mn, mx);
            b = Box{{mn.x, mn.y, mn.z}, {mx.x, mx.y, mx.z}};
            b = widen(b);
        }
    }
    return widen(b);
}

struct Lowerer {
    Program &P;
    std::string &msg;
    std::vector<std::vector<Xf>> levels;  // forward transforms, outermost first
    std::vector<int> frame_of_level;      // frame id at each depth

    static int need(const Solid &s) { return s.need; }

    uint32_t cur_frame() const { return frame_of_level.back(); }

    void atom(uint32_t kind, std::initializer_list<double> params, Box local) {
        Insn in;
        in.x = OP_ATOM | (kind << 8) | (cur_frame() << 16);
        in.y = (uint32_t)P.f64.size();
        P.f64.insert(P.f64.end(), params.begin(), params.end());
        P.atom_pc.push_back((uint32_t)P.code.size());
        P.code.push_back(in);
        AtomBox ab;
        memcpy(ab.lo, local.lo, sizeof ab.lo);
        memcpy(ab.hi, local.hi, sizeof ab.hi);
        P.atom_box.push_back(ab);
    }
    void op(uint32_t o) { P.code.push_back(Insn{o, 0}); }
    // The float copy of param k of the last atom rounds down (-1) or up (+1).
    void round_f32(uint32_t k, int dir) { P.f32_dir.push_back({P.code.back().y + k, dir}); }

    Box local_box(const Solid &s) {
        V3 mn, mx;
        min_max(s, mn, mx);
        return to_world(widen(Box{{mn.x, mn.y, mn.z}, {mx.x, mx.y, mx.z}}), levels);
    }

    bool emit(const Solid &s) {
        const double *p = s.p;
        switch (s.kind) {
        case K_SPHERE: {
            const double r = p[3], T = sqrt_le_threshold(r);
            double e[3], mag = 0;
            for (int i = 0; i < 3; i++) {
                e[i] = kU * (2 * fabs(p[i]) + 4 * r) * (1 + 4 * kU);
                mag = std::max(mag, fabs(p[i]));
            }
            double E = band_E(e, r, T);
            bool ok = guard_ok(mag, r);
            atom(A_SPHERE, {p[0], p[1], p[2], r, T, ok ? T - E : -INFINITY, ok ? T + E : INFINITY}, local_box(s));
            round_f32(5, -1), round_f32(6, +1);
            return true;
        }
        case K_RECT:
            atom(A_RECT, {p[0], p[1], p[2], p[3], p[4], p[5]}, local_box(s));
            return true;
        case K_CAPSULE: {
            // NewSegment canonical order, primitives.go:547-554; the
            // point-independent part of Segment.Closest, primitives.go:568-582.
            V3 p1 = of(p), p2 = of(p + 3), s0, s1;
            if (p1.x < p2.x || (p1.x == p2.x && p1.y < p2.y) || (p1.x == p2.x && p1.y == p2.y && p1.z < p2.z))
                s0 = p1, s1 = p2;
            else
                s0 = p2, s1 = p1;
            V3 v1 = sub(s1, s0);
            double nrm = norm(v1);
            V3 v = scale(v1, 1 / nrm);
            const double r = p[6], T = sqrt_le_threshold(r);
            // Near regime: |p - s0| <= R = L + 2r.
            const double L = nrm, R = L + 2 * r;
            const double a0[3] = {s0.x, s0.y, s0.z}, a1[3] = {s1.x, s1.y, s1.z};
            double ev2 = 0, mag = R;
            for (int i = 0; i < 3; i++) {
                ev2 += kU * (2 * fabs(a0[i]) + 2 * R) * (1 + 4 * kU);
                mag = std::max(mag, std::max(fabs(a0[i]), fabs(a1[i])) + R);
            }
            const double emag = (ev2 + 9 * kU * R) * 1.01 + 1e-12 * R;
            const double ebr = (emag + 1.01 * kU * L + 2 * kU * R) * 1.001 + 1e-12 * R;
            double e[3];
            for (int i = 0; i < 3; i++) {
                double ecl = emag + 4 * kU * R + 2 * kU * fabs(a0[i]) + kU * fabs(a1[i]);
                e[i] = (kU * (fabs(a0[i]) + R + 2 * r) + ecl) * (1 + 4 * kU);
            }
            double E = band_E(e, r, T);
            bool ok = guard_ok(mag, r) && isfinite(nrm) && nrm > 0;
            atom(A_CAPSULE,
                 {s0.x, s0.y, s0.z, s1.x, s1.y, s1.z, v.x, v.y, v.z, nrm, r, T, ok ? ebr : INFINITY,
                  ok ? T - E : -INFINITY, ok ? T + E : INFINITY},
                 local_box(s));
            round_f32(12, +1), round_f32(13, -1), round_f32(14, +1);
            return true;
        }
        case K_CYLINDER: {  // shapes.go:588-597, point-independent part
            V3 p1 = of(p), p2 = of(p + 3);
            V3 diff = sub(p1, p2);
            V3 dir = normalize(diff);
            const double r = p[6], T = sqrt_le_threshold(r), L = norm(diff), R = L + 2 * r;
            const double a0[3] = {p2.x, p2.y, p2.z};
            double ev = 0, mag = R;
            for (int i = 0; i < 3; i++) {
                ev += kU * (2 * fabs(a0[i]) + 2 * R) * (1 + 4 * kU);
                mag = std::max(mag, fabs(a0[i]) + R);
            }
            const double efrac = (ev + 9 * kU * R) * 1.01 + 1e-12 * R;
            const double ebr = (efrac + 1.01 * kU * L + 2 * kU * R) * 1.001 + 1e-12 * R;
            double e[3];
            for (int i = 0; i < 3; i++) {
                double eproj = efrac + 2 * kU * fabs(a0[i]) + 3 * kU * R;
                e[i] = (eproj + kU * (fabs(a0[i]) + R + 2 * r)) * (1 + 4 * kU);
            }
            double E = band_E(e, r, T);
            bool ok = guard_ok(mag, r) && isfinite(L) && L > 0;
            atom(A_CYLINDER,
                 {p2.x, p2.y, p2.z, dir.x, dir.y, dir.z, L, r, T, ok ? ebr : INFINITY, ok ? T - E : -INFINITY,
                  ok ? T + E : INFINITY},
                 local_box(s));
            round_f32(9, +1), round_f32(10, -1), round_f32(11, +1);
            return true;
        }
        case K_CONE: {  // shapes.go:888-898
            V3 tip = of(p), base = of(p + 3);
            V3 diff = sub(tip, base);
            V3 dir = normalize(diff);
            atom(A_CONE, {base.x, base.y, base.z, dir.x, dir.y, dir.z, norm(diff), p[6]}, local_box(s));
            return true;
        }
        case K_CONESLICE: {  // shapes.go:1073-1084
            V3 p1 = of(p), p2 = of(p + 3);
            V3 diff = sub(p2, p1);
            V3 dir = normalize(diff);
            atom(A_CONESLICE, {p1.x, p1.y, p1.z, dir.x, dir.y, dir.z, norm(diff), p[6], p[7]}, local_box(s));
            return true;
        }
        case K_TORUS: {  // shapes.go:1451-1484, OrthoBasis is point-independent
            V3 b1, b2;
            ortho_basis(of(p + 3), b1, b2);
            atom(A_TORUS, {p[0], p[1], p[2], b1.x, b1.y, b1.z, b2.x, b2.y, b2.z, p[6], p[7]}, local_box(s));
            return true;
        }
        case K_JOINED:
        case K_INTERSECTED: {
            // operands in decreasing need, ties in their order
            const size_t nc = s.children.size();
            std::pair<int, size_t> small[8];
            std::vector<std::pair<int, size_t>> big;
            std::pair<int, size_t> *order = small;
            if (nc > 8) big.resize(nc), order = big.data();
            for (size_t i = 0; i < nc; i++) order[i] = {-need(*s.children[i]), i};
            std::stable_sort(order, order + nc);
            for (size_t k = 0; k < nc; k++) {
                if (!emit(*s.children[order[k].second])) return false;
                if (k > 0) op(s.kind == K_JOINED ? OP_OR : OP_AND);
            }
            return true;
        }
        case K_SUBTRACTED: {
            const Solid &a = *s.children[0], &b = *s.children[1];
            if (need(a) >= need(b)) {
                if (!emit(a) || !emit(b)) return false;
                op(OP_ANDNOT);
            } else {
                if (!emit(b) || !emit(a)) return false;
                op(OP_RANDNOT);
            }
            return true;
        }
        case K_FORCEBOUNDS:
        case K_TRANSFORM: {
            // CheckedFuncSolid (solid.go:46-50): bounds check in the outer
            // frame, AND the child (TransformSolid: at inv.Apply(c) in a new frame).
            if (s.kind == K_TRANSFORM) {
                std::vector<Xf> inv;
                for (size_t i = s.fwd.size(); i-- > 0;) inv.push_back(xf_inverse(s.fwd[i]));
                // New frame = parent chain followed by this level's inverse ops.
                const Frame &parent = P.frames[cur_frame()];
                Frame f;
                f.offset = (uint32_t)P.f64.size();
                f.nops = parent.nops + (uint32_t)inv.size();
                std::vector<double> chain(P.f64.begin() + parent.offset,
                                          P.f64.begin() + parent.offset + 10 * parent.nops);
                for (const Xf &t : inv) {
                    chain.push_back((double)t.kind);
                    chain.insert(chain.end(), t.v, t.v + 9);
                }
                P.f64.insert(P.f64.end(), chain.begin(), chain.end());
                P.frames.push_back(f);
                if (P.frames.size() >= 65536) {
                    msg = "too many transform frames";
                    return false;
                }
                levels.push_back(s.fwd);
                frame_of_level.push_back((uint32_t)P.frames.size() - 1);
                bool ok = emit(*s.children[0]);
                levels.pop_back();
                frame_of_level.pop_back();
                if (!ok) return false;
            } else {
                if (!emit(*s.children[0])) return false;
            }
            atom(A_RECT, {s.mn[0], s.mn[1], s.mn[2], s.mx[0], s.mx[1], s.mx[2]},
                 to_world(widen(Box{{s.mn[0], s.mn[1], s.mn[2]}, {s.mx[0], s.mx[1], s.mx[2]}}), levels));
            op(OP_AND);
            return true;
        }
        case K_COLLIDER: {  // ColliderSolid.Contains, solid.go:292-300
            // params: collider, the solid's bounds, inset, radius
            const double inset = s.p[0], radius = s.p[1];
            P.colliders.push_back(s.collider);
            P.collider_R.push_back(radius != 0 ? radius : fabs(inset));
            atom(A_COLLIDER, {(double)(P.colliders.size() - 1), s.mn[0], s.mn[1], s.mn[2], s.mx[0], s.mx[1], s.mx[2],
                              inset, radius},
                 local_box(s));
            return true;
        }
        case K_HOST: {
            P.host_atoms.push_back(s.host);
            atom(A_HOST, {(double)(P.host_atoms.size() - 1)}, local_box(s));
            return true;
        }
        }
        msg = "unknown solid kind";
        return false;
    }
};

int max_stack_depth(const Program &P) {
    int d = 0, m = 0;
    for (const Insn &in : P.code) {
        uint32_t o = insn_op(in.x);
        if (o == OP_ATOM || o == OP_TRUE || o == OP_FALSE) d++;
        else if (o != OP_NOT) d--;
        m = std::max(m, d);
    }
    return m;
}

}  // namespace

bool lower_solid(const Solid &s, Program &P, std::string &msg) {
    P = Program();
    P.frames.push_back(Frame{0, 0});  // identity frame
    Lowerer L{P, msg, {}, {0}};
    P.code.reserve(256), P.f64.reserve(2048), P.atom_pc.reserve(128), P.atom_box.reserve(128), P.f32_dir.reserve(256);
    if (!L.emit(s)) return false;
    P.max_stack = max_stack_depth(P);
    if (P.max_stack > 32) {
        msg = "solid tree too deep for the device boolean stack";
        return false;
    }
    if (P.code.size() >= (1u << 24)) {
        msg = "solid program too large";
        return false;
    }
    P.f32.resize(P.f64.size());
    for (size_t i = 0; i < P.f64.size(); i++) P.f32[i] = (float)P.f64[i];
    for (const auto &d : P.f32_dir) {
        const double v = P.f64[d.first];
        float f = (float)v;
        if (d.second < 0 && (double)f > v) f = nextafterf(f, -INFINITY);
        if (d.second > 0 && (double)f < v) f = nextafterf(f, INFINITY);
        P.f32[d.first] = f;
    }
    return true;
}

// ---------------------------------------------------------------------------
// Error string (per thread)
// ---------------------------------------------------------------------------
static thread_local std::string g_err;
void set_error(const std::string &m) { g_err = m; }

}  // namespace mc3d

// ---------------------------------------------------------------------------
// C-ABI: solids
// ---------------------------------------------------------------------------
using namespace mc3d;

namespace {
mc3d_solid *wrap(Solid *s) {
    s->need = node_need(*s);
    mc3d_solid *h = new (std::nothrow) mc3d_solid;
    if (!h) {
        delete s;
        return nullptr;
    }
    h->ptr = SolidPtr(s);
    h->refs = 1;
    return h;
}
std::mutex g_ref_mu;
}  // namespace

namespace mc3d {
// JoinedTransform(xforms).Inverse() (transform.go:213-219) as frame records
// {kind, v[9]}: the inverses of the ops, last op first.
std::vector<double> inverse_records(const mc3d_xform *xforms, size_t n) {
    std::vector<double> r;
    for (size_t i = n; i-- > 0;) {
        Xf t;
        t.kind = xforms[i].kind;
        memcpy(t.v, xforms[i].v, sizeof t.v);
        const Xf inv = xf_inverse(t);
        r.push_back((double)inv.kind);
        r.insert(r.end(), inv.v, inv.v + 9);
    }
    return r;
}
}  // namespace mc3d

extern "C" {

mc3d_solid *mc3d_sphere(const double center[3], double radius) {
    if (!center) return set_error("mc3d_sphere: null center"), nullptr;
    Solid *s = new Solid{K_SPHERE};
    memcpy(s->p, center, 3 * sizeof(double));
    s->p[3] = radius;
    return wrap(s);
}

mc3d_solid *mc3d_rect(const double min[3], const double max[3]) {
    if (!min || !max) return set_error("mc3d_rect: null bound"), nullptr;
    Solid *s = new Solid{K_RECT};
    memcpy(s->p, min, 3 * sizeof(double));
    memcpy(s->p + 3, max, 3 * sizeof(double));
    return wrap(s);
}

static mc3d_solid *seg_solid(Kind k, const double *a, const double *b, double r1, double r2) {
    if (!a || !b) return set_error("null point"), nullptr;
    Solid *s = new Solid{k};
    memcpy(s->p, a, 3 * sizeof(double));
    memcpy(s->p + 3, b, 3 * sizeof(double));
    s->p[6] = r1;
    s->p[7] = r2;
    return wrap(s);
}
mc3d_solid *mc3d_capsule(const double p1[3], const double p2[3], double r) { return seg_solid(K_CAPSULE, p1, p2, r, 0); }
mc3d_solid *mc3d_cylinder(const double p1[3], const double p2[3], double r) { return seg_solid(K_CYLINDER, p1, p2, r, 0); }
mc3d_solid *mc3d_cone(const double tip[3], const double base[3], double r) { return seg_solid(K_CONE, tip, base, r, 0); }
mc3d_solid *mc3d_cone_slice(const double p1[3], const double p2[3], double r1, double r2) {
    return seg_solid(K_CONESLICE, p1, p2, r1, r2);
}
mc3d_solid *mc3d_torus(const double center[3], const double axis[3], double outer, double inner) {
    return seg_solid(K_TORUS, center, axis, outer, inner);
}

static mc3d_solid *multi(Kind k, mc3d_solid *const *ch, size_t n) {
    if (!ch || n == 0) return set_error("composite solid needs at least one child"), nullptr;
    Solid *s = new Solid{k};
    for (size_t i = 0; i < n; i++) {
        if (!ch[i]) {
            delete s;
            return set_error("null child solid"), nullptr;
        }
        s->children.push_back(ch[i]->ptr);
    }
    return wrap(s);
}
mc3d_solid *mc3d_joined(mc3d_solid *const *children, size_t n) { return multi(K_JOINED, children, n); }
mc3d_solid *mc3d_intersected(mc3d_solid *const *children, size_t n) { return multi(K_INTERSECTED, children, n); }
mc3d_solid *mc3d_subtracted(mc3d_solid *positive, mc3d_solid *negative) {
    mc3d_solid *ch[2] = {positive, negative};
    return multi(K_SUBTRACTED, ch, 2);
}

mc3d_solid *mc3d_transform(mc3d_solid *child, const mc3d_xform *xforms, size_t n) {
    if (!child || (n && !xforms)) return set_error("mc3d_transform: null argument"), nullptr;
    Solid *s = new Solid{K_TRANSFORM};
    s->children.push_back(child->ptr);
    for (size_t i = 0; i < n; i++) {
        if (xforms[i].kind < MC3D_XF_TRANSLATE || xforms[i].kind > MC3D_XF_MATRIX) {
            delete s;
            return set_error("mc3d_transform: unknown transform kind"), nullptr;
        }
        Xf t;
        t.kind = xforms[i].kind;
        memcpy(t.v, xforms[i].v, sizeof t.v);
        s->fwd.push_back(t);
    }
    // TransformSolid, transform.go:240-246: bounds = t.ApplyBounds(s.Min(), s.Max()).
    V3 mn, mx;
    min_max(*child->ptr, mn, mx);
    for (const Xf &t : s->fwd) xf_apply_bounds(t, mn, mx);
    s->mn[0] = mn.x, s->mn[1] = mn.y, s->mn[2] = mn.z;
    s->mx[0] = mx.x, s->mx[1] = mx.y, s->mx[2] = mx.z;
    return wrap(s);
}

mc3d_solid *mc3d_force_bounds(mc3d_solid *child, const double min[3], const double max[3]) {
    if (!child || !min || !max) return set_error("mc3d_force_bounds: null argument"), nullptr;
    Solid *s = new Solid{K_FORCEBOUNDS};
    s->children.push_back(child->ptr);
    memcpy(s->mn, min, 3 * sizeof(double));
    memcpy(s->mx, max, 3 * sizeof(double));
    return wrap(s);
}

mc3d_solid *mc3d_collider_solid(const double *tris, size_t n) {
    if (!tris || n == 0) return set_error("mc3d_collider_solid: empty mesh"), nullptr;
    if (n >= (1u << 30)) return set_error("mc3d_collider_solid: mesh too large"), nullptr;
    for (size_t i = 0; i < 9 * n; i++)
        if (!isfinite(tris[i])) return set_error("mc3d_collider_solid: non-finite vertex"), nullptr;
    static std::atomic<uint64_t> next_id{1};
    auto cd = std::make_shared<ColliderData>();
    cd->id = next_id++;
    cd->tris.assign(tris, tris + 9 * n);
    // Union of triangle bounds (JoinedCollider min/max, collisions.go:225-234).
    for (size_t i = 0; i < n; i++) {
        const double *t = tris + 9 * i;
        V3 a = of(t), b = of(t + 3), c = of(t + 6);
        V3 lo = vmin(vmin(a, b), c), hi = vmax(vmax(a, b), c);
        if (i == 0) {
            cd->mn[0] = lo.x, cd->mn[1] = lo.y, cd->mn[2] = lo.z;
            cd->mx[0] = hi.x, cd->mx[1] = hi.y, cd->mx[2] = hi.z;
        } else {
            V3 m0 = vmin(of(cd->mn), lo), m1 = vmax(of(cd->mx), hi);
            cd->mn[0] = m0.x, cd->mn[1] = m0.y, cd->mn[2] = m0.z;
            cd->mx[0] = m1.x, cd->mx[1] = m1.y, cd->mx[2] = m1.z;
        }
    }
    Solid *s = new Solid{K_COLLIDER};
    s->collider = cd;
    memcpy(s->mn, cd->mn, sizeof s->mn);
    memcpy(s->mx, cd->mx, sizeof s->mx);
    return wrap(s);
}

// NewColliderSolidInset(c, inset), solid.go:260-270: the collider's bounds
// grown inward by inset (min + inset; max = Max(min, c.Max() - inset)).
mc3d_solid *mc3d_collider_solid_inset(const double *tris, size_t n, double inset) {
    if (!isfinite(inset)) return set_error("mc3d_collider_solid_inset: non-finite inset"), nullptr;
    mc3d_solid *h = mc3d_collider_solid(tris, n);
    if (!h) return nullptr;
    Solid *s = const_cast<Solid *>(h->ptr.get());
    const V3 iv = mk(inset, inset, inset);
    const V3 mn = add(of(s->mn), iv), mx = vmax(mn, sub(of(s->mx), iv));
    s->mn[0] = mn.x, s->mn[1] = mn.y, s->mn[2] = mn.z;
    s->mx[0] = mx.x, s->mx[1] = mx.y, s->mx[2] = mx.z;
    s->p[0] = inset;
    return h;
}

// NewColliderSolidHollow(c, r), solid.go:272-279: the collider's bounds grown
// by r; Contains is SphereCollision(p, r).
mc3d_solid *mc3d_collider_solid_hollow(const double *tris, size_t n, double r) {
    if (!isfinite(r)) return set_error("mc3d_collider_solid_hollow: non-finite radius"), nullptr;
    mc3d_solid *h = mc3d_collider_solid(tris, n);
    if (!h) return nullptr;
    Solid *s = const_cast<Solid *>(h->ptr.get());
    const V3 rv = mk(r, r, r);
    const V3 mn = sub(of(s->mn), rv), mx = add(of(s->mx), rv);
    s->mn[0] = mn.x, s->mn[1] = mn.y, s->mn[2] = mn.z;
    s->mx[0] = mx.x, s->mx[1] = mx.y, s->mx[2] = mx.z;
    s->p[1] = r;
    return h;
}

mc3d_solid *mc3d_host_solid(const double min[3], const double max[3], mc3d_contains_fn fn, void *user) {
    if (!min || !max || !fn) return set_error("mc3d_host_solid: null argument"), nullptr;
    Solid *s = new Solid{K_HOST};
    memcpy(s->mn, min, 3 * sizeof(double));
    memcpy(s->mx, max, 3 * sizeof(double));
    s->host = HostAtom{fn, user};
    return wrap(s);
}

mc3d_solid *mc3d_solid_retain(mc3d_solid *s) {
    if (!s) return nullptr;
    std::lock_guard<std::mutex> g(g_ref_mu);
    s->refs++;
    return s;
}

void mc3d_solid_release(mc3d_solid *s) {
    if (!s) return;
    bool del;
    {
        std::lock_guard<std::mutex> g(g_ref_mu);
        del = --s->refs == 0;
    }
    if (del) delete s;
}

int mc3d_solid_bounds(const mc3d_solid *s, double min[3], double max[3]) {
    if (!s || !min || !max) return set_error("mc3d_solid_bounds: null argument"), MC3D_E_INVALID_ARG;
    solid_bounds(*s->ptr, min, max);
    return MC3D_OK;
}

int mc3d_case_table(uint8_t out[256 * 16]) {
    if (!out) return MC3D_E_INVALID_ARG;
    memcpy(out, case_table(), 256 * 16);
    return MC3D_OK;
}

const char *mc3d_last_error(void) { return g_err.c_str(); }
const char *mc3d_version(void) { return "mc3d 0.1 (gfx950)"; }

}  // extern "C"
