// eval.h -- device-side Solid.Contains: the postfix boolean program over
// atoms, each atom one primitive Contains in float64 (bit-exact with the
// reference) or in float32 with an error band (guard mode).
#pragma once

#include <hip/hip_runtime.h>
#include <math.h>
#include <stdint.h>

#include <type_traits>

#include "mc3d_internal.h"

namespace mc3d {

struct DevProg {
    const Insn *__restrict__ code;
    const double *__restrict__ f64;
    const float *__restrict__ f32;
    const Frame *__restrict__ frames;
    const PlanNode *__restrict__ plan;  // per pc, the node record of the brick plan
    const DevCollider *__restrict__ colliders;
    // Answers of host-evaluated atoms: host_vals[atom * host_stride + point id].
    const uint8_t *__restrict__ host_vals;
    uint64_t host_stride;
    uint32_t ncode;
    uint32_t nf64;  // float64 params (without the padding after them)
};

// The same pointer, known to address global memory: pointers loaded from
// memory are generic, and generic (flat) accesses go through the vector
// memory pipeline even when they would hit LDS; a cast from the global
// address space lets the compiler emit global loads instead.
template <class T>
__device__ __forceinline__ T *global_ptr(T *p) {
    return (T *)(__attribute__((address_space(1))) T *)p;
}

// The same pointer in the constant address space (memory that nothing
// writes while the kernel runs): a load through it at a wave-uniform
// address becomes a scalar load, served by the scalar data cache into
// scalar registers that vector instructions read as operands.  The walks
// take such pointers as their own types (a cast back to a generic pointer
// would be folded away together with the address space).
#define MC3D_CONST __attribute__((address_space(4)))
template <class T>
__device__ __forceinline__ const MC3D_CONST T *scalar_ptr(const T *p) {
    return (const MC3D_CONST T *)p;
}

__device__ __forceinline__ DevProg as_global(DevProg P) {
    P.code = global_ptr(P.code);
    P.f64 = global_ptr(P.f64);
    P.f32 = global_ptr(P.f32);
    P.frames = global_ptr(P.frames);
    P.plan = global_ptr(P.plan);
    P.colliders = global_ptr(P.colliders);
    P.host_vals = global_ptr(P.host_vals);
    return P;
}

// Go math.Min(c, m) == m  <=>  m == -Inf || c >= m   (and the Max twin); this
// is Rect.Contains / InBounds (shapes.go:171-173, bounder.go:23-27) including
// the NaN and signed-zero cases.
MC3D_HD inline bool in_lo(double c, double m) { return m == -INFINITY || c >= m; }
MC3D_HD inline bool in_hi(double c, double m) { return m == INFINITY || c <= m; }

// Mesh collider SphereCollision (JoinedCollider, collisions.go:292-302): does
// any triangle of the point's cell in the sphere grid (binned for this
// radius) touch the sphere -- Triangle.SphereCollision (primitives.go:251-279)
// in the reference's float64 operation order: a vertex closer than r
// (Coord3D.Dist), an edge entering the sphere (segmentEntersSphere), or the
// ray along the triangle's normal hitting it at |scale| < r (rayCollision,
// its point-independent part precomputed per triangle).  A point outside
// the grid is farther than r from every triangle.
__device__ __forceinline__ bool sphere_collision(const DevCollider &C, double x, double y, double z, double r,
                                                 uint32_t *candidates) {
    const double fx = (x - C.so[0]) * C.sinv, fy = (y - C.so[1]) * C.sinv, fz = (z - C.so[2]) * C.sinv;
    if (!(fx >= 0 && fx < C.sn[0] && fy >= 0 && fy < C.sn[1] && fz >= 0 && fz < C.sn[2])) return false;
    const uint32_t cell = ((uint32_t)fz * (uint32_t)C.sn[1] + (uint32_t)fy) * (uint32_t)C.sn[0] + (uint32_t)fx;
    const uint32_t *__restrict__ cs = global_ptr(C.scell_start);
    const uint32_t *__restrict__ ct = global_ptr(C.scell_tris);
    const double *__restrict__ rec = global_ptr(C.srec);
    const uint32_t e = cs[cell + 1];
    if (candidates) *candidates += e - cs[cell];
    for (uint32_t i = cs[cell]; i < e; i++) {
        const double *t = rec + (size_t)SPH_REC * ct[i];
        for (int k = 0; k < 3; k++) {  // p.Dist(c) < r
            const double d1 = t[3 * k] - x, d2 = t[3 * k + 1] - y, d3 = t[3 * k + 2] - z;
            if (sqrt(d1 * d1 + d2 * d2 + d3 * d3) < r) return true;
        }
        for (int k = 0; k < 3; k++) {  // segmentEntersSphere(t[k], t[k+1], c, r)
            const double *p1 = t + 3 * k, *p2 = t + 3 * ((k + 1) % 3);
            const double vx = p2[0] + -p1[0], vy = p2[1] + -p1[1], vz = p2[2] + -p1[2];
            const double frac =
                ((x * vx + y * vy + z * vz) - (p1[0] * vx + p1[1] * vy + p1[2] * vz)) / (vx * vx + vy * vy + vz * vz);
            const double cx = p1[0] + vx * frac, cy = p1[1] + vy * frac, cz = p1[2] + vz * frac;
            const double d1 = cx - x, d2 = cy - y, d3 = cz - z;
            if (frac >= 0 && frac <= 1 && sqrt(d1 * d1 + d2 * d2 + d3 * d3) < r) return true;
        }
        if (t[12] != 0) {  // rayCollision from c along the normal t[9..11]
            const double inv = t[13];
            const double ox = x + -t[0], oy = y + -t[1], oz = z + -t[2];
            const double bary1 = inv * (ox * t[14] + oy * t[15] + oz * t[16]);
            if (bary1 < 0 || bary1 > 1) continue;
            const double c2x = oy * t[19] - oz * t[18], c2y = oz * t[17] - ox * t[19], c2z = ox * t[18] - oy * t[17];
            const double bary2 = inv * (t[9] * c2x + t[10] * c2y + t[11] * c2z);
            if (bary2 < 0 || bary1 + bary2 > 1) continue;
            const double scale = inv * (t[20] * c2x + t[21] * c2y + t[22] * c2z);
            if (fabs(scale) < r) return true;
        }
    }
    return false;
}

// ColliderSolid.Contains (solid.go:292-300): InBounds, then with a radius
// the hollow shell (SphereCollision), otherwise ColliderContains(c, p, inset)
// (collisions.go:119-134): the parity of ray hits along the fixed direction,
// and with an inset the sphere test.  Only the triangles binned in the
// point's projected cell are tested for the ray; each test is
// Triangle.rayCollision (primitives.go:207-249) from its point-independent
// constants, counted by RayCollisions when `scale < 0` is false.
// Params a: collider, min(3), max(3), inset, radius.
__device__ __forceinline__ bool collider_contains(const DevCollider &C, const double *__restrict__ a, double x,
                                                  double y, double z, uint32_t *candidates = nullptr) {
    if (!(in_lo(x, a[1]) && in_lo(y, a[2]) && in_lo(z, a[3]) && in_hi(x, a[4]) && in_hi(y, a[5]) && in_hi(z, a[6])))
        return false;
    const double inset = a[7], radius = a[8];
    if (radius != 0) return sphere_collision(C, x, y, z, radius, candidates);
    uint32_t hits = 0;
    const double fu = (x * C.u[0] + y * C.u[1] + z * C.u[2] - C.u0) * C.inv_cell;
    const double fv = (x * C.v[0] + y * C.v[1] + z * C.v[2] - C.v0) * C.inv_cell;
    if (fu >= 0 && fu < C.nu && fv >= 0 && fv < C.nv) {  // else no triangle can be hit
        const uint32_t cell = (uint32_t)fv * (uint32_t)C.nu + (uint32_t)fu;
        const uint32_t *__restrict__ cs = global_ptr(C.cell_start);
        const uint32_t *__restrict__ ct = global_ptr(C.cell_tris);
        const double *__restrict__ rec = global_ptr(C.rec);
        const uint32_t e = cs[cell + 1];
        if (candidates) *candidates += e - cs[cell];
        const double Dx = 0.5224892708603626, Dy = 0.10494477243214506, Dz = 0.43558938446126527;
        for (uint32_t i = cs[cell]; i < e; i++) {
            const double *r = rec + (size_t)RAY_REC * ct[i];
            const double ox = x - r[0], oy = y - r[1], oz = z - r[2];
            const double bary1 = r[6] * (ox * r[3] + oy * r[4] + oz * r[5]);
            if (bary1 < 0 || bary1 > 1) continue;
            const double c2x = oy * r[9] - oz * r[8], c2y = oz * r[7] - ox * r[9], c2z = ox * r[8] - oy * r[7];
            const double bary2 = r[6] * (Dx * c2x + Dy * c2y + Dz * c2z);
            if (bary2 < 0 || bary1 + bary2 > 1) continue;
            const double scale = r[6] * (r[10] * c2x + r[11] * c2y + r[12] * c2z);
            hits += !(scale < 0);
        }
    }
    if (!(hits & 1u)) return inset < 0 && sphere_collision(C, x, y, z, -inset, candidates);
    return inset <= 0 || !sphere_collision(C, x, y, z, inset, candidates);
}

// Guard statistics and algorithmic flop counts (statistics runs only), kept
// in registers and reduced once per block.
struct GStats {
    uint32_t evals = 0;       // float32 atom evaluations
    uint32_t rechecks = 0;    // of those, in the band -> float64
    uint32_t mismatches = 0;  // plain float32 decision != float64 decision
    uint32_t violations = 0;  // band said sure but float64 disagrees (must stay 0)
    uint32_t flops64 = 0;     // float64 operations of the evaluated atoms and frames
    uint32_t flops32 = 0;     // float32 operations of the guarded atoms
};

// Algorithmic operation counts (add, sub, mul, div, sqrt and compare each
// count one) of one atom evaluation, from the sequences in atom64 / atom32
// below.  A collider costs its bounds test and projection plus a fixed
// Moller-Trumbore sequence per candidate triangle of the point's cell.
__device__ __forceinline__ uint32_t atom_flops64(uint32_t kind) {
    switch (kind) {
    case A_SPHERE: return 9;
    case A_RECT: return 6;
    case A_CAPSULE: return 25;
    case A_CYLINDER: return 25;
    case A_CONE: return 29;
    case A_CONESLICE: return 32;
    case A_TORUS: return 41;
    case A_COLLIDER: return 20;
    default: return 0;
    }
}
constexpr uint32_t COLLIDER_FLOPS_PER_CANDIDATE = 37;
__device__ __forceinline__ uint32_t atom_flops32(uint32_t kind) {
    switch (kind) {
    case A_SPHERE: return 10;
    case A_CAPSULE: return 29;
    case A_CYLINDER: return 29;
    default: return 0;
    }
}

template <class F64P>
__device__ __forceinline__ void apply_frame(F64P f64, Frame fr, double &x, double &y, double &z) {
    F64P r = f64 + fr.offset;
    for (uint32_t i = 0; i < fr.nops; ++i, r += 10) {
        int k = (int)r[0];
        if (k == MC3D_XF_TRANSLATE) {  // Translate.Apply, transform.go:44-46
            x = x + r[1];
            y = y + r[2];
            z = z + r[3];
        } else if (k == MC3D_XF_SCALE) {  // Scale.Apply, transform.go:71-73
            x = x * r[1];
            y = y * r[1];
            z = z * r[1];
        } else if (k == MC3D_XF_VECSCALE) {  // VecScale.Apply, transform.go:103-105
            x = x * r[1];
            y = y * r[2];
            z = z * r[3];
        } else {  // Matrix3.MulColumn, matrix.go:131-137
            double nx = r[1] * x + r[2] * y + r[3] * z;
            double ny = r[4] * x + r[5] * y + r[6] * z;
            double nz = r[7] * x + r[8] * y + r[9] * z;
            x = nx, y = ny, z = nz;
        }
    }
}

// One primitive Contains in float64 (point already in the atom's frame).
// Constants that do not depend on the point were computed on the host with
// the reference's operation order (solid.cpp).
template <class F64P>
__device__ __forceinline__ bool atom64(uint32_t kind, F64P a, double x, double y, double z) {
    switch (kind) {
    case A_SPHERE: {  // Sphere.Contains: coord.Dist(center) <= r, shapes.go:29-31
        double d1 = x - a[0], d2 = y - a[1], d3 = z - a[2];
        return d1 * d1 + d2 * d2 + d3 * d3 <= a[4];  // sqrt(s) <= r  <=>  s <= T
    }
    case A_RECT:  // Rect.Contains, shapes.go:171-173
        // all six tests (no short circuit): the params load together
        return ((int)in_lo(x, a[0]) & (int)in_lo(y, a[1]) & (int)in_lo(z, a[2]) & (int)in_hi(x, a[3]) &
                (int)in_hi(y, a[4]) & (int)in_hi(z, a[5])) != 0;
    case A_CAPSULE: {  // Segment.Dist(c) <= r, primitives.go:568-593
        double v2x = x - a[0], v2y = y - a[1], v2z = z - a[2];
        double mag = a[6] * v2x + a[7] * v2y + a[8] * v2z;
        double cx, cy, cz;
        if (mag > a[9]) {
            cx = a[3], cy = a[4], cz = a[5];
        } else if (mag < 0) {
            cx = a[0], cy = a[1], cz = a[2];
        } else {
            cx = a[6] * mag + a[0], cy = a[7] * mag + a[1], cz = a[8] * mag + a[2];
        }
        double d1 = x - cx, d2 = y - cy, d3 = z - cz;
        return d1 * d1 + d2 * d2 + d3 * d3 <= a[11];
    }
    case A_CYLINDER: {  // Cylinder.Contains, shapes.go:588-597
        double frac = (x - a[0]) * a[3] + (y - a[1]) * a[4] + (z - a[2]) * a[5];
        if (frac < 0 || frac > a[6]) return false;
        double px = a[0] + a[3] * frac, py = a[1] + a[4] * frac, pz = a[2] + a[5] * frac;
        double d1 = px - x, d2 = py - y, d3 = pz - z;
        return d1 * d1 + d2 * d2 + d3 * d3 <= a[8];
    }
    case A_CONE: {  // Cone.Contains, shapes.go:888-898
        double frac = (x - a[0]) * a[3] + (y - a[1]) * a[4] + (z - a[2]) * a[5];
        double rf = 1 - frac / a[6];
        if (rf < 0 || rf > 1) return false;
        double px = a[0] + a[3] * frac, py = a[1] + a[4] * frac, pz = a[2] + a[5] * frac;
        double d1 = px - x, d2 = py - y, d3 = pz - z;
        return sqrt(d1 * d1 + d2 * d2 + d3 * d3) <= a[7] * rf;
    }
    case A_CONESLICE: {  // ConeSlice.Contains, shapes.go:1073-1084
        double frac = (x - a[0]) * a[3] + (y - a[1]) * a[4] + (z - a[2]) * a[5];
        double rf = 1 - frac / a[6];
        if (rf < 0 || rf > 1) return false;
        double px = a[0] + a[3] * frac, py = a[1] + a[4] * frac, pz = a[2] + a[5] * frac;
        double r = a[7] * rf + a[8] * (1 - rf);
        double d1 = px - x, d2 = py - y, d3 = pz - z;
        return sqrt(d1 * d1 + d2 * d2 + d3 * d3) <= r;
    }
    case A_TORUS: {  // Torus.Contains -> genericSDF, shapes.go:1363-1365, 1451-1484
        double cx = x - a[0], cy = y - a[1], cz = z - a[2];
        double u = a[3] * cx + a[4] * cy + a[5] * cz;
        double v = a[6] * cx + a[7] * cy + a[8] * cz;
        double on = sqrt(u * u + v * v);
        if (on == 0) {
            u = 1, v = 0, on = 1;
        }
        double s = a[9] / on;
        u *= s;
        v *= s;
        double rx = a[3] * u + a[6] * v, ry = a[4] * u + a[7] * v, rz = a[5] * u + a[8] * v;
        double d1 = rx - cx, d2 = ry - cy, d3 = rz - cz;
        return a[10] - sqrt(d1 * d1 + d2 * d2 + d3 * d3) >= 0;
    }
    }
    return false;
}

// Float32 decision with a band: 1 inside, 0 outside, -1 in the band (the
// caller re-decides in float64).  The band thresholds (Tlo, Thi) and branch
// margins were computed on the host from a forward error bound of this exact
// float32 operation sequence, valid for every point within 2r of the
// primitive; points farther away can only be decided "outside", which is
// then correct (solid.cpp, guard_*).  *plain gets the unguarded float32
// decision for the mismatch statistic.
template <class F32P>
__device__ __forceinline__ int atom32(uint32_t kind, F32P a, float x, float y, float z, bool *plain) {
    switch (kind) {
    case A_SPHERE: {  // [c0 c1 c2 r T Tlo Thi]
        float d1 = x - a[0], d2 = y - a[1], d3 = z - a[2];
        float q = d1 * d1 + d2 * d2 + d3 * d3;
        *plain = q <= a[4];
        return q <= a[5] ? 1 : (q > a[6] ? 0 : -1);
    }
    case A_CAPSULE: {  // [s0(3) s1(3) v(3) norm r T ebr Tlo Thi]
        float v2x = x - a[0], v2y = y - a[1], v2z = z - a[2];
        float mag = a[6] * v2x + a[7] * v2y + a[8] * v2z;
        float cx, cy, cz;
        bool amb = fabsf(mag - a[9]) <= a[12] || fabsf(mag) <= a[12];
        if (mag > a[9]) {
            cx = a[3], cy = a[4], cz = a[5];
        } else if (mag < 0) {
            cx = a[0], cy = a[1], cz = a[2];
        } else {
            cx = a[6] * mag + a[0], cy = a[7] * mag + a[1], cz = a[8] * mag + a[2];
        }
        float d1 = x - cx, d2 = y - cy, d3 = z - cz;
        float q = d1 * d1 + d2 * d2 + d3 * d3;
        *plain = q <= a[11];
        if (amb) return -1;
        return q <= a[13] ? 1 : (q > a[14] ? 0 : -1);
    }
    case A_CYLINDER: {  // [p2(3) dir(3) norm r T ebr Tlo Thi]
        float frac = (x - a[0]) * a[3] + (y - a[1]) * a[4] + (z - a[2]) * a[5];
        // Slab test against the margin: frac - norm is compared as one rounded
        // difference (a rounded norm + margin could swallow a point just past
        // the end cap).
        float dl = frac - a[6];
        *plain = false;
        if (frac < -a[9] || dl > a[9]) return 0;
        bool amb = fabsf(frac) <= a[9] || fabsf(dl) <= a[9];
        float px = a[0] + a[3] * frac, py = a[1] + a[4] * frac, pz = a[2] + a[5] * frac;
        float d1 = px - x, d2 = py - y, d3 = pz - z;
        float q = d1 * d1 + d2 * d2 + d3 * d3;
        *plain = frac >= 0 && frac <= a[6] && q <= a[8];
        if (amb) return -1;
        return q <= a[10] ? 1 : (q > a[11] ? 0 : -1);
    }
    }
    return -1;  // no float32 form: decide in float64
}

__device__ __forceinline__ bool has_atom32(uint32_t kind) {
    return kind == A_SPHERE || kind == A_CAPSULE || kind == A_CYLINDER;
}

// Number of params of each atom kind (the layouts above / in solid.cpp).
__device__ __forceinline__ uint32_t atom_nparams(uint32_t kind) {
    switch (kind) {
    case A_SPHERE: return 7;
    case A_RECT: return 6;
    case A_CAPSULE: return 15;
    case A_CYLINDER: return 12;
    case A_CONE: return 8;
    case A_CONESLICE: return 9;
    case A_TORUS: return 11;
    case A_COLLIDER: return 9;
    default: return 1;
    }
}
constexpr uint32_t MAX_ATOM_PARAMS = 16;

// The N params of one atom copied into vector registers (uniform values): a
// loop over a brick's rows reads them from there rather than reloading them
// (scalar registers are too few to hold them across the loop).
template <class T, int N>
struct RegParams {
    T v[N];
    __device__ __forceinline__ T operator[](int i) const { return v[i]; }
};
template <int N, class T, class SrcP>
__device__ __forceinline__ RegParams<T, N> reg_params(SrcP src) {
    RegParams<T, N> r;
    for (int i = 0; i < N; i++) {
        r.v[i] = src[i];
        asm volatile("" : "+v"(r.v[i]));  // kept in a vector register
    }
    return r;
}

// One atom at one point, with the guard statistics.  f64/f32 point at the
// atom's params.
// HOST: the program may hold host-evaluated atoms, whose answers for point
// `pid` were filled in by the host beforehand (kernels without point ids are
// never launched on such programs).
template <int MODE, bool STATS, bool HOST, class F64P, class F32P>
__device__ __forceinline__ bool eval_atom(const DevProg &P, uint32_t kind, F64P f64, F32P f32, double qx, double qy,
                                          double qz, float fx, float fy, float fz, bool live, uint64_t pid,
                                          GStats &gs) {
    if constexpr (std::is_pointer<F64P>::value) {  // params in memory (not RegParams)
        if (kind == A_COLLIDER) {
            uint32_t cand = 0;
            const bool b = collider_contains(P.colliders[(int)f64[0]], (const double *)f64, qx, qy, qz,
                                             STATS ? &cand : nullptr);
            if (STATS && live) gs.flops64 += atom_flops64(kind) + COLLIDER_FLOPS_PER_CANDIDATE * cand;
            return b;
        }
        if (kind == A_HOST) return HOST && P.host_vals[(uint64_t)f64[0] * P.host_stride + pid] != 0;
    }
    if constexpr (MODE == 1) {
        if (has_atom32(kind)) {
            bool plain;
            const int r = atom32(kind, f32, fx, fy, fz, &plain);
            if (STATS && live) gs.evals++, gs.flops32 += atom_flops32(kind);
            if (r < 0) {
                const bool b = atom64(kind, f64, qx, qy, qz);
                if (STATS && live) gs.rechecks++, gs.mismatches += plain != b, gs.flops64 += atom_flops64(kind);
                return b;
            }
            if (STATS && live) {
                const bool t = atom64(kind, f64, qx, qy, qz);
                gs.mismatches += plain != t;
                gs.violations += (r != 0) != t;
            }
            return r != 0;
        }
    }
    if (STATS && live) gs.flops64 += atom_flops64(kind);
    return atom64(kind, f64, qx, qy, qz);
}

// Operation count of one frame (3 per translate / scale, 15 per matrix).
template <class F64P>
__device__ __forceinline__ uint32_t frame_flops(F64P f64, Frame fr) {
    uint32_t n = 0;
    for (uint32_t i = 0; i < fr.nops; ++i) {
        const int k = (int)f64[fr.offset + 10 * i];
        n += (k == MC3D_XF_TRANSLATE || k == MC3D_XF_SCALE || k == MC3D_XF_VECSCALE) ? 3 : 15;
    }
    return n;
}

// Postfix boolean program over atoms for NP points at once (one walk of the
// program); each point's stack is a bit stack in a register.  The program
// and the params it refers to (insn.y offsets into f64 / f32; frames carry
// their params there too) are read through the given pointer types: with
// UNI the whole wave walks one program, read with scalar loads through
// constant-space pointers (scalar_ptr), so the walk branches on scalars;
// otherwise each lane reads its own.  MODE 0: every atom in float64.  MODE 1:
// float32 with band, float64 in the band.  `live` only gates the statistics.
// One non-atom instruction on a bit stack (bit 0 the top; see eval_code_n).
__device__ __forceinline__ uint32_t stack_op(uint32_t op, uint32_t s) {
    const uint32_t t = s & 1u;
    if (op == OP_OR) return (s >> 1) | t;
    if (op == OP_AND) return (s >> 1) & (~1u | t);
    if (op == OP_ANDNOT) return (s >> 1) & ~t;  // second & !top
    if (op == OP_RANDNOT) {                     // !second & top
        s >>= 1;
        return (s & ~1u) | (t & ~s & 1u);
    }
    if (op == OP_NOT) return s ^ 1u;
    if (op == OP_TRUE) return (s << 1) | 1u;
    return s << 1;  // OP_FALSE
}

template <int MODE, bool STATS, int NP, bool UNI, bool HOST = false, class CodeP, class F64P, class F32P,
          class FrameP>
__device__ __forceinline__ void eval_code_n(CodeP code, uint32_t n, const DevProg &P, F64P f64, F32P f32,
                                            FrameP frames, const double *px, const double *py, const double *pz,
                                            const bool *live, bool *res, GStats &gs, const uint64_t *pid = nullptr) {
    uint32_t st[NP];
    double qx[NP], qy[NP], qz[NP];
    float fx[NP], fy[NP], fz[NP];
    for (int k = 0; k < NP; k++) {
        st[k] = 0;
        qx[k] = px[k], qy[k] = py[k], qz[k] = pz[k];
        fx[k] = (float)px[k], fy[k] = (float)py[k], fz[k] = (float)pz[k];
    }
    uint32_t cur = 0;
    for (uint32_t pc = 0; pc < n; ++pc) {
        Insn in{code[pc].x, code[pc].y};  // field-wise: the pointer may be in any address space
        if (UNI) in.x = (uint32_t)__builtin_amdgcn_readfirstlane((int)in.x);
        if (UNI) in.y = (uint32_t)__builtin_amdgcn_readfirstlane((int)in.y);
        const uint32_t op = insn_op(in.x);
        if (op == OP_ATOM) {
            const uint32_t fr = insn_frame(in.x);
            if (fr != cur) {
                const Frame f{frames[fr].offset, frames[fr].nops};
                for (int k = 0; k < NP; k++) {
                    qx[k] = px[k], qy[k] = py[k], qz[k] = pz[k];
                    apply_frame(f64, f, qx[k], qy[k], qz[k]);
                    if (MODE == 1) fx[k] = (float)qx[k], fy[k] = (float)qy[k], fz[k] = (float)qz[k];
                }
                if (STATS) {
                    const uint32_t ff = frame_flops(f64, f);
                    for (int k = 0; k < NP; k++) gs.flops64 += live[k] ? ff : 0;
                }
                cur = fr;
            }
            const uint32_t kind = insn_kind(in.x);
            for (int k = 0; k < NP; k++)
                st[k] = (st[k] << 1) |
                        (uint32_t)eval_atom<MODE, STATS, HOST>(P, kind, f64 + in.y, f32 + in.y, qx[k], qy[k], qz[k],
                                                               fx[k], fy[k], fz[k], live[k], HOST ? pid[k] : 0, gs);
        } else {
            for (int k = 0; k < NP; k++) {
                uint32_t s = st[k];
                if (op == OP_OR) {
                    const uint32_t t = s & 1u;
                    s = (s >> 1) | t;
                } else if (op == OP_AND) {
                    const uint32_t t = s & 1u;
                    s = (s >> 1) & (~1u | t);
                } else if (op == OP_ANDNOT) {  // second & !top
                    const uint32_t t = s & 1u;
                    s = (s >> 1) & ~t;
                } else if (op == OP_RANDNOT) {  // !second & top
                    const uint32_t t = s & 1u;
                    s >>= 1;
                    s = (s & ~1u) | (t & ~s & 1u);
                } else if (op == OP_NOT) {
                    s ^= 1u;
                } else if (op == OP_TRUE) {
                    s = (s << 1) | 1u;
                } else {
                    s <<= 1;
                }
                st[k] = s;
            }
        }
    }
    for (int k = 0; k < NP; k++) res[k] = st[k] & 1u;
}

template <int MODE, bool STATS, bool UNI, bool HOST = false>
__device__ __forceinline__ bool eval_code(const Insn *__restrict__ code, uint32_t n, const DevProg &P, double px,
                                          double py, double pz, GStats &gs, uint64_t pid = 0) {
    const bool live = true;
    bool r;
    if constexpr (UNI)  // one program for the whole wave: scalar loads
        eval_code_n<MODE, STATS, 1, UNI, HOST>(scalar_ptr(code), n, P, scalar_ptr(P.f64), scalar_ptr(P.f32),
                                               scalar_ptr(P.frames), &px, &py, &pz, &live, &r, gs, &pid);
    else
        eval_code_n<MODE, STATS, 1, UNI, HOST>(code, n, P, P.f64, P.f32, P.frames, &px, &py, &pz, &live, &r, gs,
                                               &pid);
    return r;
}

}  // namespace mc3d
