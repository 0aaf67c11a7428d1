// mc3d.hip -- CDNA4 (gfx950) kernels and the pipeline behind
// mc3d_marching_cubes_search (model3d/mc.go:45-49).
//
// Stages, one solid per call:
//   corners  : Solid.Contains at every grid corner -> bit-packed occupancy,
//              one 64-bit word per 64 consecutive x corners of a row; one
//              wave per word, packed with a wave ballot (mc.go:614-692).
//   classify : per word, crossing masks of the x/y/z grid edges (one vertex
//              per crossed edge, which is the reference's vertex dedup,
//              mesh.go:822-849) and the triangle count of the 64 cubes
//              (mcLookupTable, mc.go:434-456); exclusive scans give every
//              word its vertex and triangle offsets.
//   emit     : per word, vertex keys (axis, x, y, z of the owning edge) and
//              triangles as uint32 vertex-index triples in table order
//              (mc.go:20-35, 401-407).
//   bisect   : per vertex, mcSearchPoint (mc.go:223-258) in float64: the
//              first probe is the edge's high corner, read from the
//              occupancy grid (same point, same answer), then `iters`
//              midpoint evaluations.
//
// Every float64 expression keeps the reference's operation order and is
// compiled with -ffp-contract=off, so Contains decisions, topology and
// vertex coordinates are bit-identical to the reference's amd64 build.

#include <hip/hip_runtime.h>
#include <math.h>
#include <stdint.h>
#include <string.h>

#include <chrono>
#include <memory>
#include <new>
#include <string>
#include <vector>

#include "mc3d_internal.h"

using namespace mc3d;

namespace {

// ---------------------------------------------------------------------------
// Device program evaluation
// ---------------------------------------------------------------------------
struct DevProg {
    const Insn *__restrict__ code;
    const double *__restrict__ f64;
    const float *__restrict__ f32;
    const Frame *__restrict__ frames;
    uint32_t ncode;
    uint32_t pad;
};

// Go math.Min(c, m) == m  <=>  m == -Inf || c >= m   (and the Max twin); this
// is Rect.Contains / InBounds (shapes.go:171-173, bounder.go:23-27) including
// the NaN and signed-zero cases.
__device__ __forceinline__ bool in_lo(double c, double m) { return m == -INFINITY || c >= m; }
__device__ __forceinline__ bool in_hi(double c, double m) { return m == INFINITY || c <= m; }

__device__ __forceinline__ void apply_frame(const double *__restrict__ f64, Frame fr, double &x, double &y,
                                            double &z) {
    const double *r = f64 + fr.offset;
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
__device__ __forceinline__ bool atom64(uint32_t kind, const double *__restrict__ a, double x, double y, double z) {
    switch (kind) {
    case A_SPHERE: {  // Sphere.Contains: coord.Dist(center) <= r, shapes.go:29-31
        double d1 = x - a[0], d2 = y - a[1], d3 = z - a[2];
        return d1 * d1 + d2 * d2 + d3 * d3 <= a[4];
    }
    case A_RECT:  // Rect.Contains, shapes.go:171-173
        return in_lo(x, a[0]) && in_lo(y, a[1]) && in_lo(z, a[2]) && in_hi(x, a[3]) && in_hi(y, a[4]) &&
               in_hi(z, a[5]);
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

// Postfix boolean program over atoms; the stack is a bit stack in a register.
__device__ __forceinline__ bool eval_point(const DevProg &P, double px, double py, double pz) {
    uint32_t st = 0;
    uint32_t cur = 0;
    double qx = px, qy = py, qz = pz;
    for (uint32_t pc = 0; pc < P.ncode; ++pc) {
        Insn in = P.code[pc];
        uint32_t op = insn_op(in.x);
        if (op == OP_ATOM) {
            uint32_t fr = insn_frame(in.x);
            if (fr != cur) {
                qx = px, qy = py, qz = pz;
                apply_frame(P.f64, P.frames[fr], qx, qy, qz);
                cur = fr;
            }
            st = (st << 1) | (uint32_t)atom64(insn_kind(in.x), P.f64 + in.y, qx, qy, qz);
        } else if (op == OP_OR) {
            uint32_t t = st & 1u;
            st = (st >> 1) | t;
        } else if (op == OP_AND) {
            uint32_t t = st & 1u;
            st = (st >> 1) & (~1u | t);
        } else if (op == OP_ANDNOT) {  // second & !top
            uint32_t t = st & 1u;
            st = (st >> 1) & ~t;
        } else if (op == OP_RANDNOT) {  // !second & top
            uint32_t t = st & 1u;
            st >>= 1;
            st = (st & ~1u) | (t & ~st & 1u);
        } else if (op == OP_NOT) {
            st ^= 1u;
        } else if (op == OP_TRUE) {
            st = (st << 1) | 1u;
        } else {
            st <<= 1;
        }
    }
    return st & 1u;
}

// ---------------------------------------------------------------------------
// Kernels
// ---------------------------------------------------------------------------
struct Grid {
    const double *__restrict__ xs;
    const double *__restrict__ ys;
    const double *__restrict__ zs;
    int nx, ny, nz, W;  // W = words per row = ceil(nx / 64)
};

__device__ __forceinline__ size_t word_index(const Grid &g, int wx, int y, int z) {
    return ((size_t)z * g.ny + y) * g.W + wx;
}

// corners: one wave per occupancy word (64 x corners of one row).
__global__ __launch_bounds__(256) void k_corners(DevProg P, Grid g, uint64_t *__restrict__ bits,
                                                 uint32_t *__restrict__ flags, uint64_t nwords) {
    const int lane = threadIdx.x & 63;
    for (uint64_t w = (uint64_t)blockIdx.x * 4 + (threadIdx.x >> 6); w < nwords; w += (uint64_t)gridDim.x * 4) {
        const int wx = (int)(w % g.W);
        const uint64_t row = w / g.W;
        const int y = (int)(row % g.ny), z = (int)(row / g.ny);
        const int x = wx * 64 + lane;
        bool inside = false;
        if (x < g.nx) inside = eval_point(P, g.xs[x], g.ys[y], g.zs[z]);
        uint64_t m = __ballot(inside);
        if (lane == 0) bits[w] = m;
        // panic("solid is true outside of bounds"), mc.go:687-689
        if (inside && (x == 0 || x == g.nx - 1 || y == 0 || y == g.ny - 1 || z == 0 || z == g.nz - 1))
            atomicOr(flags, 1u);
    }
}

struct Masks {
    uint64_t cx, cy, cz;
};

// Crossed-edge masks of one occupancy word: bit i of cx is the edge from
// corner (64wx+i, y, z) to its +x neighbour, and likewise for y and z.
// Padding bits (x >= nx) and boundary rows are zero, so no bounds checks
// beyond the neighbour words' existence are needed.
__device__ __forceinline__ Masks word_masks(const Grid &g, const uint64_t *__restrict__ bits, int wx, int y, int z) {
    size_t w = word_index(g, wx, y, z);
    uint64_t b = bits[w];
    uint64_t bn = wx + 1 < g.W ? bits[w + 1] : 0;
    uint64_t by = y + 1 < g.ny ? bits[w + g.W] : 0;
    uint64_t bz = z + 1 < g.nz ? bits[w + (size_t)g.ny * g.W] : 0;
    Masks m;
    m.cx = b ^ ((b >> 1) | (bn << 63));
    m.cy = b ^ by;
    m.cz = b ^ bz;
    return m;
}

__device__ __forceinline__ uint64_t row_word(const Grid &g, const uint64_t *__restrict__ bits, int wx, int y,
                                              int z) {
    if (wx >= g.W || y >= g.ny || z >= g.nz) return 0;
    return bits[word_index(g, wx, y, z)];
}

// Per-lane corner values of the 64 cubes of a word: c[k] bit i is corner k
// (k = dx + 2 dy + 4 dz, mc.go:260-301) of cube 64wx+i.
struct CubeCorners {
    uint64_t c[8];
};
__device__ __forceinline__ CubeCorners cube_corners(const Grid &g, const uint64_t *__restrict__ bits, int wx, int y,
                                                    int z) {
    CubeCorners cc;
    for (int dz = 0; dz < 2; dz++)
        for (int dy = 0; dy < 2; dy++) {
            uint64_t b = row_word(g, bits, wx, y + dy, z + dz);
            uint64_t bn = row_word(g, bits, wx + 1, y + dy, z + dz);
            cc.c[2 * dy + 4 * dz] = b;
            cc.c[1 + 2 * dy + 4 * dz] = (b >> 1) | (bn << 63);
        }
    return cc;
}

__device__ __forceinline__ uint64_t mixed_cubes(const CubeCorners &cc) {
    uint64_t any = 0, all = ~0ull;
    for (int k = 0; k < 8; k++) any |= cc.c[k], all &= cc.c[k];
    return any & ~all;
}

__device__ __forceinline__ int cube_case(const CubeCorners &cc, int i) {
    int c = 0;
    for (int k = 0; k < 8; k++) c |= (int)((cc.c[k] >> i) & 1) << k;
    return c;
}

// classify: per word, number of crossed edges (vertices) and triangles.
__global__ __launch_bounds__(256) void k_classify(Grid g, const uint64_t *__restrict__ bits,
                                                  const uint8_t *__restrict__ table, uint2 *__restrict__ counts,
                                                  uint64_t nwords) {
    __shared__ uint8_t ntri[256];
    ntri[threadIdx.x] = table[16 * threadIdx.x];
    __syncthreads();
    for (uint64_t w = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; w < nwords;
         w += (uint64_t)gridDim.x * blockDim.x) {
        const int wx = (int)(w % g.W);
        const uint64_t row = w / g.W;
        const int y = (int)(row % g.ny), z = (int)(row / g.ny);
        Masks m = word_masks(g, bits, wx, y, z);
        uint32_t nv = __popcll(m.cx) + __popcll(m.cy) + __popcll(m.cz);
        CubeCorners cc = cube_corners(g, bits, wx, y, z);
        uint64_t mixed = mixed_cubes(cc);
        uint32_t nt = 0;
        while (mixed) {
            int i = __ffsll((unsigned long long)mixed) - 1;
            mixed &= mixed - 1;
            nt += ntri[cube_case(cc, i)];
        }
        counts[w] = make_uint2(nv, nt);
    }
}

// ---- exclusive scan of uint2 counts (vertices, triangles) ----------------
constexpr int SCAN_BLOCK = 256;
constexpr int SCAN_ITEMS = 8;
constexpr int SCAN_TILE = SCAN_BLOCK * SCAN_ITEMS;

__device__ __forceinline__ uint2 add2(uint2 a, uint2 b) { return make_uint2(a.x + b.x, a.y + b.y); }

// Block-wide exclusive scan of one value per thread; returns the exclusive
// prefix and writes the block total.
__device__ __forceinline__ uint2 block_exclusive(uint2 v, uint2 *sh, uint2 &total) {
    const int t = threadIdx.x;
    sh[t] = v;
    __syncthreads();
    for (int off = 1; off < SCAN_BLOCK; off <<= 1) {
        uint2 o = t >= off ? sh[t - off] : make_uint2(0, 0);
        __syncthreads();
        sh[t] = add2(sh[t], o);
        __syncthreads();
    }
    total = sh[SCAN_BLOCK - 1];
    uint2 r = t > 0 ? sh[t - 1] : make_uint2(0, 0);
    __syncthreads();
    return r;
}

__global__ __launch_bounds__(SCAN_BLOCK) void k_scan_reduce(const uint2 *__restrict__ in, uint64_t n,
                                                            uint2 *__restrict__ tile_sums) {
    __shared__ uint2 sh[SCAN_BLOCK];
    uint64_t base = (uint64_t)blockIdx.x * SCAN_TILE + (uint64_t)threadIdx.x * SCAN_ITEMS;
    uint2 s = make_uint2(0, 0);
    for (int i = 0; i < SCAN_ITEMS; i++)
        if (base + i < n) s = add2(s, in[base + i]);
    uint2 total;
    block_exclusive(s, sh, total);
    if (threadIdx.x == 0) tile_sums[blockIdx.x] = total;
}

// Single block: exclusive scan of the tile sums in place; totals in 64 bits.
__global__ __launch_bounds__(SCAN_BLOCK) void k_scan_tiles(uint2 *__restrict__ tile_sums, uint64_t ntiles,
                                                           unsigned long long *__restrict__ totals) {
    __shared__ uint2 sh[SCAN_BLOCK];
    uint2 carry = make_uint2(0, 0);
    unsigned long long tv = 0, tt = 0;
    for (uint64_t start = 0; start < ntiles; start += SCAN_BLOCK) {
        uint64_t i = start + threadIdx.x;
        uint2 v = i < ntiles ? tile_sums[i] : make_uint2(0, 0);
        uint2 total;
        uint2 ex = block_exclusive(v, sh, total);
        if (i < ntiles) tile_sums[i] = add2(ex, carry);
        carry = add2(carry, total);
        tv += total.x;
        tt += total.y;
    }
    if (threadIdx.x == 0) {
        totals[0] = tv;
        totals[1] = tt;
    }
}

__global__ __launch_bounds__(SCAN_BLOCK) void k_scan_down(const uint2 *__restrict__ in, uint64_t n,
                                                          const uint2 *__restrict__ tile_sums,
                                                          uint2 *__restrict__ out) {
    __shared__ uint2 sh[SCAN_BLOCK];
    uint64_t base = (uint64_t)blockIdx.x * SCAN_TILE + (uint64_t)threadIdx.x * SCAN_ITEMS;
    uint2 vals[SCAN_ITEMS];
    uint2 s = make_uint2(0, 0);
    for (int i = 0; i < SCAN_ITEMS; i++) {
        vals[i] = base + i < n ? in[base + i] : make_uint2(0, 0);
        s = add2(s, vals[i]);
    }
    uint2 total;
    uint2 ex = add2(block_exclusive(s, sh, total), tile_sums[blockIdx.x]);
    for (int i = 0; i < SCAN_ITEMS; i++) {
        if (base + i < n) out[base + i] = ex;
        ex = add2(ex, vals[i]);
    }
}

// ---- emit -----------------------------------------------------------------
__device__ __forceinline__ uint64_t vkey(int axis, int x, int y, int z) {
    return (uint64_t)axis | ((uint64_t)x << 2) | ((uint64_t)y << 22) | ((uint64_t)z << 42);
}

// Vertex index of the grid edge (axis, x, y, z): the word's base, then the
// word's x, y, z crossing masks in that order.
__device__ __forceinline__ uint32_t vertex_of_edge(const Grid &g, const uint64_t *__restrict__ bits,
                                                   const uint2 *__restrict__ offs, int axis, int x, int y, int z) {
    int wx = x >> 6, b = x & 63;
    Masks m = word_masks(g, bits, wx, y, z);
    uint32_t idx = offs[word_index(g, wx, y, z)].x;
    uint64_t below = b ? (~0ull >> (64 - b)) : 0ull;
    if (axis == 0) return idx + __popcll(m.cx & below);
    idx += __popcll(m.cx);
    if (axis == 1) return idx + __popcll(m.cy & below);
    idx += __popcll(m.cy);
    return idx + __popcll(m.cz & below);
}

__global__ __launch_bounds__(256) void k_emit(Grid g, const uint64_t *__restrict__ bits,
                                              const uint8_t *__restrict__ table, const uint2 *__restrict__ offs,
                                              uint64_t *__restrict__ vkeys, uint32_t *__restrict__ tris,
                                              uint64_t nwords) {
    __shared__ uint8_t tab[256 * 16];
    for (int i = threadIdx.x; i < 256 * 16; i += blockDim.x) tab[i] = table[i];
    __syncthreads();
    for (uint64_t w = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; w < nwords;
         w += (uint64_t)gridDim.x * blockDim.x) {
        const int wx = (int)(w % g.W);
        const uint64_t row = w / g.W;
        const int y = (int)(row % g.ny), z = (int)(row / g.ny);
        uint2 o = offs[w];
        // vertices: x, then y, then z crossings, each in bit order
        Masks m = word_masks(g, bits, wx, y, z);
        uint32_t v = o.x;
        uint64_t ms[3] = {m.cx, m.cy, m.cz};
        for (int a = 0; a < 3; a++) {
            uint64_t mm = ms[a];
            while (mm) {
                int i = __ffsll((unsigned long long)mm) - 1;
                mm &= mm - 1;
                vkeys[v++] = vkey(a, wx * 64 + i, y, z);
            }
        }
        // triangles, cube by cube in x order, table order within a cube
        CubeCorners cc = cube_corners(g, bits, wx, y, z);
        uint64_t mixed = mixed_cubes(cc);
        uint32_t t = o.y;
        while (mixed) {
            int i = __ffsll((unsigned long long)mixed) - 1;
            mixed &= mixed - 1;
            const uint8_t *row8 = tab + 16 * cube_case(cc, i);
            int nt = row8[0];
            int x = wx * 64 + i;
            for (int k = 0; k < 3 * nt; k++) {
                int e = row8[1 + k];
                int axis = e >> 2, r = e & 3;
                int ex = x, ey = y, ez = z;
                if (axis == 0) ey += r & 1, ez += r >> 1;
                else if (axis == 1) ex += r & 1, ez += r >> 1;
                else ex += r & 1, ey += r >> 1;
                tris[3 * (size_t)t + k] = vertex_of_edge(g, bits, offs, axis, ex, ey, ez);
            }
            t += nt;
        }
    }
}

__device__ __forceinline__ bool corner_bit(const Grid &g, const uint64_t *__restrict__ bits, int x, int y, int z) {
    return (bits[word_index(g, x >> 6, y, z)] >> (x & 63)) & 1;
}

// bisect: mcSearchPoint (mc.go:223-258) per vertex; iters == 0 leaves the
// edge midpoints of MarchingCubes (Coord3D.Mid, coords.go:196-198).
__global__ __launch_bounds__(256) void k_bisect(DevProg P, Grid g, const uint64_t *__restrict__ bits,
                                                const uint64_t *__restrict__ vkeys, double *__restrict__ verts,
                                                uint64_t nverts, int iters) {
    for (uint64_t v = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; v < nverts;
         v += (uint64_t)gridDim.x * blockDim.x) {
        uint64_t key = vkeys[v];
        int axis = (int)(key & 3);
        int x = (int)((key >> 2) & 0xfffff), y = (int)((key >> 22) & 0xfffff), z = (int)((key >> 42) & 0xfffff);
        double p[3] = {g.xs[x], g.ys[y], g.zs[z]};
        double lo, hi;
        int hx = x, hy = y, hz = z;
        if (axis == 0) lo = g.xs[x], hi = g.xs[x + 1], hx++;
        else if (axis == 1) lo = g.ys[y], hi = g.ys[y + 1], hy++;
        else lo = g.zs[z], hi = g.zs[z + 1], hz++;
        double res;
        if (iters == 0) {
            res = (lo + hi) * 0.5;
        } else {
            double fp = lo, tp = hi;
            if (!corner_bit(g, bits, hx, hy, hz)) fp = hi, tp = lo;
            for (int i = 0; i < iters; i++) {
                double mid = (fp + tp) / 2;
                p[axis] = mid;
                if (eval_point(P, p[0], p[1], p[2])) tp = mid;
                else fp = mid;
            }
            res = (fp + tp) / 2;
        }
        p[axis] = res;
        verts[3 * v] = p[0];
        verts[3 * v + 1] = p[1];
        verts[3 * v + 2] = p[2];
    }
}

__global__ __launch_bounds__(256) void k_contains(DevProg P, const double *__restrict__ pts, uint8_t *__restrict__ out,
                                                  uint64_t n) {
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x)
        out[i] = eval_point(P, pts[3 * i], pts[3 * i + 1], pts[3 * i + 2]) ? 1 : 0;
}

// ---------------------------------------------------------------------------
// Host side
// ---------------------------------------------------------------------------
#define HIP_TRY(expr)                                                                        \
    do {                                                                                     \
        hipError_t e_ = (expr);                                                              \
        if (e_ != hipSuccess) {                                                              \
            set_error(std::string(#expr) + ": " + hipGetErrorString(e_));                    \
            return MC3D_E_HIP;                                                               \
        }                                                                                    \
    } while (0)

struct DevBuf {
    void *p = nullptr;
    size_t cap = 0;
    ~DevBuf() {
        if (p) (void)hipFree(p);
    }
    int ensure(size_t bytes) {
        if (bytes <= cap) return MC3D_OK;
        if (p) (void)hipFree(p);
        p = nullptr;
        cap = 0;
        size_t want = bytes + bytes / 4 + 4096;
        hipError_t e = hipMalloc(&p, want);
        if (e != hipSuccess) {
            p = nullptr;
            set_error(std::string("hipMalloc: ") + hipGetErrorString(e));
            return MC3D_E_OOM;
        }
        cap = want;
        return MC3D_OK;
    }
    template <class T>
    T *as() const {
        return (T *)p;
    }
};

unsigned grid_for(uint64_t n, unsigned block, unsigned cap = 1u << 20) {
    uint64_t b = (n + block - 1) / block;
    if (b == 0) b = 1;
    return (unsigned)(b < cap ? b : cap);
}

}  // namespace

struct mc3d_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    DevBuf prog_code, prog_f64, prog_f32, prog_frames, table;
    DevBuf axes, bits, counts, offs, tiles, totals, flags, vkeys, verts, tris, pts, out8;
    hipEvent_t ev[6] = {};
    bool table_loaded = false;
};

struct mc3d_mesh {
    mc3d_ctx *ctx = nullptr;
    uint64_t nv = 0, nt = 0;
    std::vector<double> hverts;
    std::vector<uint32_t> htris;
    bool on_host = false;
    DevBuf dverts, dtris;  // device results owned by the mesh
    mc3d_stats stats{};
};

namespace {

int upload_program(mc3d_ctx *ctx, const Program &P, DevProg &dp) {
    int rc;
    if ((rc = ctx->prog_code.ensure(P.code.size() * sizeof(Insn) + 16))) return rc;
    if ((rc = ctx->prog_f64.ensure(P.f64.size() * sizeof(double) + 16))) return rc;
    if ((rc = ctx->prog_f32.ensure(P.f32.size() * sizeof(float) + 16))) return rc;
    if ((rc = ctx->prog_frames.ensure(P.frames.size() * sizeof(Frame) + 16))) return rc;
    HIP_TRY(hipMemcpyAsync(ctx->prog_code.p, P.code.data(), P.code.size() * sizeof(Insn), hipMemcpyHostToDevice,
                           ctx->stream));
    if (!P.f64.empty()) {
        HIP_TRY(hipMemcpyAsync(ctx->prog_f64.p, P.f64.data(), P.f64.size() * sizeof(double),
                               hipMemcpyHostToDevice, ctx->stream));
        HIP_TRY(hipMemcpyAsync(ctx->prog_f32.p, P.f32.data(), P.f32.size() * sizeof(float),
                               hipMemcpyHostToDevice, ctx->stream));
    }
    HIP_TRY(hipMemcpyAsync(ctx->prog_frames.p, P.frames.data(), P.frames.size() * sizeof(Frame),
                           hipMemcpyHostToDevice, ctx->stream));
    dp.code = ctx->prog_code.as<Insn>();
    dp.f64 = ctx->prog_f64.as<double>();
    dp.f32 = ctx->prog_f32.as<float>();
    dp.frames = ctx->prog_frames.as<Frame>();
    dp.ncode = (uint32_t)P.code.size();
    dp.pad = 0;
    return MC3D_OK;
}

int check_supported(const Program &P) {
    if (!P.colliders.empty()) {
        set_error("ColliderSolid lowering is not available in this build");
        return MC3D_E_UNSUPPORTED;
    }
    if (!P.host_atoms.empty()) {
        set_error("host-evaluated solids are not available in this build");
        return MC3D_E_UNSUPPORTED;
    }
    return MC3D_OK;
}

std::string fmt_bounds(const double *mn, const double *mx) {
    char buf[512];
    snprintf(buf, sizeof buf,
             "invalid bounds: min=model3d.Coord3D{X:%.17g, Y:%.17g, Z:%.17g} max=model3d.Coord3D{X:%.17g, Y:%.17g, "
             "Z:%.17g}",
             mn[0], mn[1], mn[2], mx[0], mx[1], mx[2]);
    return buf;
}

float elapsed(hipEvent_t a, hipEvent_t b) {
    float ms = 0;
    if (hipEventElapsedTime(&ms, a, b) != hipSuccess) return 0;
    return ms;
}

}  // namespace

extern "C" {

int mc3d_ctx_create(int device, mc3d_ctx **out) {
    if (!out) return MC3D_E_INVALID_ARG;
    *out = nullptr;
    int n = 0;
    HIP_TRY(hipGetDeviceCount(&n));
    if (device < 0 || device >= n) {
        set_error("mc3d_ctx_create: no such HIP device");
        return MC3D_E_INVALID_ARG;
    }
    HIP_TRY(hipSetDevice(device));
    mc3d_ctx *c = new (std::nothrow) mc3d_ctx;
    if (!c) return MC3D_E_OOM;
    c->device = device;
    hipError_t e = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking);
    if (e != hipSuccess) {
        set_error(std::string("hipStreamCreate: ") + hipGetErrorString(e));
        delete c;
        return MC3D_E_HIP;
    }
    for (auto &ev : c->ev) {
        if (hipEventCreate(&ev) != hipSuccess) {
            set_error("hipEventCreate failed");
            mc3d_ctx_destroy(c);
            return MC3D_E_HIP;
        }
    }
    *out = c;
    return MC3D_OK;
}

void mc3d_ctx_destroy(mc3d_ctx *ctx) {
    if (!ctx) return;
    (void)hipSetDevice(ctx->device);
    if (ctx->stream) (void)hipStreamSynchronize(ctx->stream);
    for (auto &ev : ctx->ev)
        if (ev) (void)hipEventDestroy(ev);
    if (ctx->stream) (void)hipStreamDestroy(ctx->stream);
    delete ctx;
}

int mc3d_marching_cubes_search(mc3d_ctx *ctx, const mc3d_solid *s, double delta, int iters,
                               const mc3d_options *opts, mc3d_mesh **out) {
    if (!ctx || !s || !out) {
        set_error("mc3d_marching_cubes_search: null argument");
        return MC3D_E_INVALID_ARG;
    }
    *out = nullptr;
    if (iters < 0) iters = 0;
    mc3d_options o{};
    o.precision = MC3D_PREC_FP64;
    if (opts) o = *opts;
    HIP_TRY(hipSetDevice(ctx->device));
    auto t_start = std::chrono::steady_clock::now();

    // validateBounds, mc.go:15 / bounder.go:48-52
    double mn[3], mx[3];
    solid_bounds(*s->ptr, mn, mx);
    if (!bounds_valid(*s->ptr)) {
        set_error(fmt_bounds(mn, mx));
        return MC3D_E_INVALID_BOUNDS;
    }
    if (!(delta > 0)) {
        set_error("delta must be positive");
        return MC3D_E_INVALID_ARG;
    }
    // newSquareSpacer, mc.go:594-608
    std::vector<double> ax[3];
    for (int a = 0; a < 3; a++) {
        ax[a] = spacer_axis(mn[a], mx[a], delta);
        if (ax[a].size() >= (1u << 20)) {
            set_error("grid too large: more than 2^20 samples on an axis");
            return MC3D_E_UNSUPPORTED;
        }
        if (ax[a].size() < 2) {
            set_error("grid has fewer than two samples on an axis");
            return MC3D_E_INVALID_ARG;
        }
    }
    Program P;
    std::string msg;
    if (!lower_solid(*s->ptr, P, msg)) {
        set_error(msg);
        return MC3D_E_UNSUPPORTED;
    }
    int rc = check_supported(P);
    if (rc) return rc;

    Grid g;
    g.nx = (int)ax[0].size(), g.ny = (int)ax[1].size(), g.nz = (int)ax[2].size();
    g.W = (g.nx + 63) / 64;
    const uint64_t nwords = (uint64_t)g.W * g.ny * g.nz;

    DevProg dp;
    if ((rc = upload_program(ctx, P, dp))) return rc;
    if ((rc = ctx->axes.ensure((ax[0].size() + ax[1].size() + ax[2].size()) * sizeof(double)))) return rc;
    double *dax = ctx->axes.as<double>();
    HIP_TRY(hipMemcpyAsync(dax, ax[0].data(), ax[0].size() * 8, hipMemcpyHostToDevice, ctx->stream));
    HIP_TRY(hipMemcpyAsync(dax + g.nx, ax[1].data(), ax[1].size() * 8, hipMemcpyHostToDevice, ctx->stream));
    HIP_TRY(hipMemcpyAsync(dax + g.nx + g.ny, ax[2].data(), ax[2].size() * 8, hipMemcpyHostToDevice, ctx->stream));
    g.xs = dax, g.ys = dax + g.nx, g.zs = dax + g.nx + g.ny;
    if (!ctx->table_loaded) {
        if ((rc = ctx->table.ensure(256 * 16))) return rc;
        HIP_TRY(hipMemcpy(ctx->table.p, case_table(), 256 * 16, hipMemcpyHostToDevice));
        ctx->table_loaded = true;
    }
    const uint64_t ntiles = (nwords + SCAN_TILE - 1) / SCAN_TILE;
    if ((rc = ctx->bits.ensure(nwords * 8))) return rc;
    if ((rc = ctx->counts.ensure(nwords * sizeof(uint2)))) return rc;
    if ((rc = ctx->offs.ensure(nwords * sizeof(uint2)))) return rc;
    if ((rc = ctx->tiles.ensure(ntiles * sizeof(uint2)))) return rc;
    if ((rc = ctx->totals.ensure(2 * sizeof(unsigned long long)))) return rc;
    if ((rc = ctx->flags.ensure(64))) return rc;
    uint64_t *bits = ctx->bits.as<uint64_t>();
    HIP_TRY(hipMemsetAsync(ctx->flags.p, 0, 64, ctx->stream));

    HIP_TRY(hipEventRecord(ctx->ev[0], ctx->stream));
    k_corners<<<grid_for(nwords, 4), 256, 0, ctx->stream>>>(dp, g, bits, ctx->flags.as<uint32_t>(),
                                                                       nwords);
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipEventRecord(ctx->ev[1], ctx->stream));
    k_classify<<<grid_for(nwords, 256), 256, 0, ctx->stream>>>(g, bits, ctx->table.as<uint8_t>(),
                                                                 ctx->counts.as<uint2>(), nwords);
    HIP_TRY(hipGetLastError());
    k_scan_reduce<<<(unsigned)ntiles, SCAN_BLOCK, 0, ctx->stream>>>(ctx->counts.as<uint2>(), nwords,
                                                                      ctx->tiles.as<uint2>());
    k_scan_tiles<<<1, SCAN_BLOCK, 0, ctx->stream>>>(ctx->tiles.as<uint2>(), ntiles,
                                                     ctx->totals.as<unsigned long long>());
    k_scan_down<<<(unsigned)ntiles, SCAN_BLOCK, 0, ctx->stream>>>(ctx->counts.as<uint2>(), nwords,
                                                                    ctx->tiles.as<uint2>(), ctx->offs.as<uint2>());
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipEventRecord(ctx->ev[2], ctx->stream));
    unsigned long long totals[2];
    uint32_t flag = 0;
    HIP_TRY(hipMemcpyAsync(totals, ctx->totals.p, sizeof totals, hipMemcpyDeviceToHost, ctx->stream));
    HIP_TRY(hipMemcpyAsync(&flag, ctx->flags.p, sizeof flag, hipMemcpyDeviceToHost, ctx->stream));
    HIP_TRY(hipStreamSynchronize(ctx->stream));
    if (flag) {
        set_error("solid is true outside of bounds");
        return MC3D_E_INSIDE_BOUNDARY;
    }
    if (totals[0] >= (1ull << 32) || totals[1] >= (1ull << 32)) {
        set_error("mesh too large for 32-bit vertex indices");
        return MC3D_E_UNSUPPORTED;
    }
    const uint64_t nv = totals[0], nt = totals[1];

    std::unique_ptr<mc3d_mesh> mesh(new (std::nothrow) mc3d_mesh);
    mc3d_mesh *m = mesh.get();
    if (!m) return MC3D_E_OOM;
    m->ctx = ctx;
    m->nv = nv;
    m->nt = nt;
    if ((rc = ctx->vkeys.ensure(nv * 8 + 8)) || (rc = m->dverts.ensure(nv * 24 + 8)) ||
        (rc = m->dtris.ensure(nt * 12 + 8)))
        return rc;
    HIP_TRY(hipEventRecord(ctx->ev[3], ctx->stream));
    k_emit<<<grid_for(nwords, 256), 256, 0, ctx->stream>>>(g, bits, ctx->table.as<uint8_t>(), ctx->offs.as<uint2>(),
                                                             ctx->vkeys.as<uint64_t>(), m->dtris.as<uint32_t>(),
                                                             nwords);
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipEventRecord(ctx->ev[4], ctx->stream));
    if (nv) {
        k_bisect<<<grid_for(nv, 256), 256, 0, ctx->stream>>>(dp, g, bits, ctx->vkeys.as<uint64_t>(),
                                                              m->dverts.as<double>(), nv, iters);
        HIP_TRY(hipGetLastError());
    }
    HIP_TRY(hipEventRecord(ctx->ev[5], ctx->stream));
    HIP_TRY(hipStreamSynchronize(ctx->stream));

    mc3d_stats &st = m->stats;
    st.corners = (uint64_t)g.nx * g.ny * g.nz;
    st.vertices = nv;
    st.triangles = nt;
    st.bisection_evals = iters > 0 ? nv * (uint64_t)(iters + 1) : 0;
    st.ms_corners = elapsed(ctx->ev[0], ctx->ev[1]);
    st.ms_classify = elapsed(ctx->ev[1], ctx->ev[2]);
    st.ms_emit = elapsed(ctx->ev[3], ctx->ev[4]);
    st.ms_bisect = elapsed(ctx->ev[4], ctx->ev[5]);
    st.ms_device = st.ms_corners + st.ms_classify + st.ms_emit + st.ms_bisect;
    if (!o.device_only) {
        auto t0 = std::chrono::steady_clock::now();
        m->hverts.resize(nv * 3);
        m->htris.resize(nt * 3);
        if (nv) HIP_TRY(hipMemcpy(m->hverts.data(), m->dverts.p, nv * 24, hipMemcpyDeviceToHost));
        if (nt) HIP_TRY(hipMemcpy(m->htris.data(), m->dtris.p, nt * 12, hipMemcpyDeviceToHost));
        m->on_host = true;
        st.ms_d2h = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    }
    (void)t_start;
    *out = mesh.release();
    return MC3D_OK;
}

int mc3d_mesh_counts(const mc3d_mesh *m, size_t *nv, size_t *nt) {
    if (!m) return MC3D_E_INVALID_ARG;
    if (nv) *nv = m->nv;
    if (nt) *nt = m->nt;
    return MC3D_OK;
}

int mc3d_mesh_copy(mc3d_mesh *m, double *verts, uint32_t *tris) {
    if (!m) return MC3D_E_INVALID_ARG;
    if (m->on_host) {
        if (verts && m->nv) memcpy(verts, m->hverts.data(), m->nv * 24);
        if (tris && m->nt) memcpy(tris, m->htris.data(), m->nt * 12);
        return MC3D_OK;
    }
    HIP_TRY(hipSetDevice(m->ctx->device));
    if (verts && m->nv) HIP_TRY(hipMemcpy(verts, m->dverts.p, m->nv * 24, hipMemcpyDeviceToHost));
    if (tris && m->nt) HIP_TRY(hipMemcpy(tris, m->dtris.p, m->nt * 12, hipMemcpyDeviceToHost));
    return MC3D_OK;
}

int mc3d_mesh_device_ptrs(const mc3d_mesh *m, const double **verts, const uint32_t **tris) {
    if (!m) return MC3D_E_INVALID_ARG;
    if (verts) *verts = m->dverts.as<double>();
    if (tris) *tris = m->dtris.as<uint32_t>();
    return MC3D_OK;
}

int mc3d_mesh_stats(const mc3d_mesh *m, mc3d_stats *out) {
    if (!m || !out) return MC3D_E_INVALID_ARG;
    *out = m->stats;
    return MC3D_OK;
}

void mc3d_mesh_free(mc3d_mesh *m) {
    if (!m) return;
    if (m->ctx) (void)hipSetDevice(m->ctx->device);
    delete m;
}

int mc3d_contains(mc3d_ctx *ctx, const mc3d_solid *s, const double *pts, size_t n, int32_t precision,
                  uint8_t *out) {
    if (!ctx || !s || (n && (!pts || !out))) {
        set_error("mc3d_contains: null argument");
        return MC3D_E_INVALID_ARG;
    }
    (void)precision;
    if (n == 0) return MC3D_OK;
    HIP_TRY(hipSetDevice(ctx->device));
    Program P;
    std::string msg;
    if (!lower_solid(*s->ptr, P, msg)) {
        set_error(msg);
        return MC3D_E_UNSUPPORTED;
    }
    int rc = check_supported(P);
    if (rc) return rc;
    DevProg dp;
    if ((rc = upload_program(ctx, P, dp))) return rc;
    if ((rc = ctx->pts.ensure(n * 24)) || (rc = ctx->out8.ensure(n))) return rc;
    HIP_TRY(hipMemcpyAsync(ctx->pts.p, pts, n * 24, hipMemcpyHostToDevice, ctx->stream));
    k_contains<<<grid_for(n, 256), 256, 0, ctx->stream>>>(dp, ctx->pts.as<double>(), ctx->out8.as<uint8_t>(), n);
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipMemcpyAsync(out, ctx->out8.p, n, hipMemcpyDeviceToHost, ctx->stream));
    HIP_TRY(hipStreamSynchronize(ctx->stream));
    return MC3D_OK;
}

}  // extern "C"
