// mc3d.hip -- CDNA4 (gfx950) kernels and the pipeline behind
// mc3d_marching_cubes_search (model3d/mc.go:45-49).
//
// Stages, one solid per call:
//   plan     : (culling on) per brick of 64 x 4 x 4 corners, the solid's
//              program simplified against the brick's box: atoms whose
//              conservative box misses the brick are false, bounds checks and
//              spheres that contain it are true, and the boolean ops fold
//              around them.  Empty and full bricks need no evaluation.
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
//   bisect   : per vertex, mcSearchPoint (mc.go:223-258): the first probe is
//              the edge's high corner, read from the occupancy grid (same
//              point, same answer), then `iters` float64 midpoints.
//
// Every float64 expression keeps the reference's operation order and is
// compiled with -ffp-contract=off, so Contains decisions, topology and
// vertex coordinates are bit-identical to the reference's amd64 build.  In
// guard mode an atom is first decided in float32 and re-decided in float64
// when the float32 value is inside its error band (eval.h).

#include <hip/hip_runtime.h>
#include <math.h>
#include <stdint.h>
#include <string.h>

#include <chrono>
#include <memory>
#include <new>
#include <string>
#include <vector>

#include "eval.h"
#include "mc3d_internal.h"

using namespace mc3d;

namespace {

constexpr uint32_t BRICK_TRUE = 0xffffffffu;

// Per-brick programs: info[b] = {offset into code, length}; length 0 is a
// brick where the solid is false everywhere, BRICK_TRUE where it is true.
struct Bricks {
    const uint2 *__restrict__ info;  // null: no culling, use the full program
    const Insn *__restrict__ code;
    int BY, BZ, nby, nbz;
};

template <bool STATS>
__device__ __forceinline__ void flush_stats(const GStats &gs, unsigned long long *__restrict__ out) {
    if (!STATS) return;
    __shared__ unsigned int red[4];
    if (threadIdx.x < 4) red[threadIdx.x] = 0;
    __syncthreads();
    if (gs.evals) atomicAdd(&red[0], gs.evals);
    if (gs.rechecks) atomicAdd(&red[1], gs.rechecks);
    if (gs.mismatches) atomicAdd(&red[2], gs.mismatches);
    if (gs.violations) atomicAdd(&red[3], gs.violations);
    __syncthreads();
    if (threadIdx.x < 4 && red[threadIdx.x]) atomicAdd(&out[threadIdx.x], (unsigned long long)red[threadIdx.x]);
}
struct Grid {
    const double *__restrict__ xs;
    const double *__restrict__ ys;
    const double *__restrict__ zs;
    int nx, ny, nz, W;  // W = words per row = ceil(nx / 64)
};

__device__ __forceinline__ size_t word_index(const Grid &g, int wx, int y, int z) {
    return ((size_t)z * g.ny + y) * g.W + wx;
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

// Appends `item` for the lanes with `pred` set: one atomic per wave.  The
// list order is arbitrary; consumers place their output by scanned offsets.
__device__ __forceinline__ void wave_append(bool pred, uint32_t item, uint32_t *__restrict__ list,
                                            uint32_t *__restrict__ count) {
    const uint64_t bal = __ballot(pred);
    if (!bal) return;
    const int lane = threadIdx.x & 63;
    const int leader = __ffsll((unsigned long long)bal) - 1;
    uint32_t base = 0;
    if (lane == leader) base = atomicAdd(count, (uint32_t)__popcll(bal));
    base = __shfl(base, leader);
    if (pred) list[base + __popcll(bal & ((1ull << lane) - 1))] = item;
}

// classify: per word, number of crossed edges (vertices) and triangles; the
// words with any are listed for emit.
__global__ __launch_bounds__(256) void k_classify(Grid g, const uint64_t *__restrict__ bits,
                                                  const uint8_t *__restrict__ table, uint2 *__restrict__ counts,
                                                  uint32_t *__restrict__ active, uint32_t *__restrict__ nactive,
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
        wave_append((nv | nt) != 0, (uint32_t)w, active, nactive);
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

// Occupancy word, crossing masks and vertex base of one (row, word); vertex
// numbering inside a word is: its x crossings, then y, then z, each in bit
// order, starting at the word's scanned base.
struct WordInfo {
    uint64_t bits, cx, cy, cz;
    uint32_t base, pad;
};

__device__ __forceinline__ void wave_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// emit: one wave per listed word, one lane per x position.  The 2 x 2 rows
// by 2 words a cube of this word can touch are staged in LDS, so every lane
// can number the vertex on any edge of its cube.
__global__ __launch_bounds__(256) void k_emit(Grid g, const uint64_t *__restrict__ bits,
                                              const uint8_t *__restrict__ table, const uint2 *__restrict__ offs,
                                              const uint32_t *__restrict__ active,
                                              const uint32_t *__restrict__ nactive, uint64_t *__restrict__ vkeys,
                                              uint32_t *__restrict__ tris) {
    __shared__ uint8_t tab[256 * 16];
    __shared__ WordInfo wi_all[4][8];
    for (int i = threadIdx.x; i < 256 * 16; i += blockDim.x) tab[i] = table[i];
    __syncthreads();
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    WordInfo *wi = wi_all[wave];
    const uint32_t nact = *nactive;
    const uint64_t below = lane ? (~0ull >> (64 - lane)) : 0ull;
    for (uint32_t k = blockIdx.x * 4 + wave; k < nact; k += gridDim.x * 4) {
        const uint64_t w = active[k];
        const int wx = (int)(w % g.W);
        const uint64_t row = w / g.W;
        const int y = (int)(row % g.ny), z = (int)(row / g.ny);
        wave_sync();
        if (lane < 8) {  // entry: bit 0 word +1, bit 1 row +1 in y, bit 2 row +1 in z
            const int ex = wx + (lane & 1), ey = y + ((lane >> 1) & 1), ez = z + (lane >> 2);
            WordInfo r{0, 0, 0, 0, 0, 0};
            if (ex < g.W && ey < g.ny && ez < g.nz) {
                const size_t wi_idx = word_index(g, ex, ey, ez);
                Masks m = word_masks(g, bits, ex, ey, ez);
                r.bits = bits[wi_idx];
                r.cx = m.cx, r.cy = m.cy, r.cz = m.cz;
                r.base = offs[wi_idx].x;
            }
            wi[lane] = r;
        }
        wave_sync();
        const int x = wx * 64 + lane;
        // vertices of this word
        {
            const WordInfo me = wi[0];
            uint32_t v = me.base;
            if ((me.cx >> lane) & 1) vkeys[v + __popcll(me.cx & below)] = vkey(0, x, y, z);
            v += __popcll(me.cx);
            if ((me.cy >> lane) & 1) vkeys[v + __popcll(me.cy & below)] = vkey(1, x, y, z);
            v += __popcll(me.cy);
            if ((me.cz >> lane) & 1) vkeys[v + __popcll(me.cz & below)] = vkey(2, x, y, z);
        }
        // triangles of this lane's cube, placed by a ballot prefix sum
        int cs = 0;
        for (int dz = 0; dz < 2; dz++)
            for (int dy = 0; dy < 2; dy++) {
                const int e = 2 * dy + 4 * dz;
                const uint64_t b = wi[(dy << 1) | (dz << 2)].bits, bn = wi[1 | (dy << 1) | (dz << 2)].bits;
                cs |= (int)((b >> lane) & 1) << e;
                cs |= (int)((((b >> 1) | (bn << 63)) >> lane) & 1) << (e + 1);
            }
        const int nt = (cs == 0 || cs == 255) ? 0 : tab[16 * cs];
        const uint32_t pre = __popcll(__ballot(nt & 1) & below) + 2 * __popcll(__ballot(nt & 2) & below) +
                             4 * __popcll(__ballot(nt & 4) & below);
        const uint32_t t = offs[w].y + pre;
        for (int kk = 0; kk < 3 * nt; kk++) {
            const int e = tab[16 * cs + 1 + kk];
            const int axis = e >> 2, r = e & 3;
            int ddx = 0, ddy = 0, ddz = 0;
            if (axis == 0) ddy = r & 1, ddz = r >> 1;
            else if (axis == 1) ddx = r & 1, ddz = r >> 1;
            else ddx = r & 1, ddy = r >> 1;
            const int ex = x + ddx;
            const WordInfo &o = wi[((ex >> 6) - wx) | (ddy << 1) | (ddz << 2)];
            const int bb = ex & 63;
            const uint64_t bl = bb ? (~0ull >> (64 - bb)) : 0ull;
            uint32_t idx = o.base;
            if (axis == 0) {
                idx += __popcll(o.cx & bl);
            } else if (axis == 1) {
                idx += __popcll(o.cx) + __popcll(o.cy & bl);
            } else {
                idx += __popcll(o.cx) + __popcll(o.cy) + __popcll(o.cz & bl);
            }
            tris[3 * (size_t)t + kk] = idx;
        }
    }
}


// ---- plan: per-brick program simplification --------------------------------
// Atom against a brick box: 0 certainly false, 1 certainly true, 2 evaluate.
// "False" needs only the atom's conservative world box (solid.cpp widens it
// far beyond float64 rounding); "true" is proven for untransformed bounds
// checks / boxes (exact comparisons) and spheres (farthest corner inside the
// threshold with a relative margin far above the rounding of the sum).
__device__ __forceinline__ int classify_atom(const DevProg &P, uint32_t pc, Insn in, const double *lo,
                                             const double *hi) {
    const AtomBox bx = P.boxes[pc];
    if (hi[0] < bx.lo[0] || lo[0] > bx.hi[0] || hi[1] < bx.lo[1] || lo[1] > bx.hi[1] || hi[2] < bx.lo[2] ||
        lo[2] > bx.hi[2])
        return 0;
    if (insn_frame(in.x) == 0) {
        const double *a = P.f64 + in.y;
        const uint32_t kind = insn_kind(in.x);
        if (kind == A_RECT) {
            if (in_lo(lo[0], a[0]) && in_lo(lo[1], a[1]) && in_lo(lo[2], a[2]) && in_hi(hi[0], a[3]) &&
                in_hi(hi[1], a[4]) && in_hi(hi[2], a[5]))
                return 1;
        } else if (kind == A_SPHERE) {
            double s = 0;
            for (int i = 0; i < 3; i++) {
                double m = fmax(fabs(lo[i] - a[i]), fabs(hi[i] - a[i]));
                s += m * m;
            }
            if (s * (1 + 1e-12) <= a[4]) return 1;
        }
    }
    return 2;
}

constexpr uint32_t TAG_EXPR = 0u, TAG_FALSE = 1u << 30, TAG_TRUE = 2u << 30, TAG_MASK = 3u << 30;
constexpr int PLAN_BLOCK = 128;
constexpr int PLAN_STACK = 33;

// PASS 0 writes counts[b] = {code slots to reserve, 1 if the brick needs
// evaluation}; the reservation is the peak length during the walk, since
// folding a constant can discard code already emitted.  PASS 1 rewrites the
// code of active bricks at offs[b].x and sets info[b] for every brick.
template <int PASS>
__global__ __launch_bounds__(PLAN_BLOCK) void k_plan(DevProg P, Grid g, Bricks B, uint64_t nbricks,
                                                     uint2 *__restrict__ counts, const uint2 *__restrict__ offs,
                                                     uint2 *__restrict__ info, Insn *__restrict__ out,
                                                     uint32_t *__restrict__ act_list) {
    __shared__ uint32_t stk_all[PLAN_BLOCK * PLAN_STACK];
    uint32_t *S = stk_all + threadIdx.x * PLAN_STACK;
    for (uint64_t b = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; b < nbricks;
         b += (uint64_t)gridDim.x * blockDim.x) {
        const int bx = (int)(b % g.W);
        const uint64_t r = b / g.W;
        const int by = (int)(r % B.nby), bz = (int)(r / B.nby);
        const int x0 = bx * 64, y0 = by * B.BY, z0 = bz * B.BZ;
        const int x1 = min(x0 + 64, g.nx - 1), y1 = min(y0 + B.BY, g.ny - 1), z1 = min(z0 + B.BZ, g.nz - 1);
        const double lo[3] = {g.xs[x0], g.ys[y0], g.zs[z0]};
        const double hi[3] = {g.xs[x1], g.ys[y1], g.zs[z1]};
        Insn *o = (PASS == 1 && counts[b].y) ? out + offs[b].x : nullptr;
        uint32_t n = 0, peak = 0;
        int sp = 0;
        for (uint32_t pc = 0; pc < P.ncode; ++pc) {
            const Insn in = P.code[pc];
            const uint32_t op = insn_op(in.x);
            if (op == OP_ATOM) {
                int c = classify_atom(P, pc, in, lo, hi);
                if (c == 2) {
                    if (o) o[n] = in;
                    S[sp++] = TAG_EXPR | n;
                    n++;
                } else {
                    S[sp++] = c ? TAG_TRUE : TAG_FALSE;
                }
            } else if (op == OP_TRUE || op == OP_FALSE) {
                S[sp++] = op == OP_TRUE ? TAG_TRUE : TAG_FALSE;
            } else if (op == OP_NOT) {
                uint32_t a = S[sp - 1];
                if (a == TAG_TRUE) S[sp - 1] = TAG_FALSE;
                else if (a == TAG_FALSE) S[sp - 1] = TAG_TRUE;
                else {
                    if (o) o[n] = in;
                    n++;
                }
            } else {
                const uint32_t bb = S[--sp], aa = S[--sp];
                const bool aE = (aa & TAG_MASK) == TAG_EXPR, bE = (bb & TAG_MASK) == TAG_EXPR;
                const bool aT = aa == TAG_TRUE, aF = aa == TAG_FALSE, bT = bb == TAG_TRUE, bF = bb == TAG_FALSE;
                // start of the operands' code: constants emitted nothing
                const uint32_t first = aE ? (aa & ~TAG_MASK) : (bE ? (bb & ~TAG_MASK) : n);
                uint32_t res;
                bool emit = false, emit_not = false;
                if (op == OP_OR) {
                    if (aT || bT) res = TAG_TRUE, n = first;
                    else if (aF) res = bb;
                    else if (bF) res = aa;
                    else res = aa, emit = true;
                } else if (op == OP_AND) {
                    if (aF || bF) res = TAG_FALSE, n = first;
                    else if (aT) res = bb;
                    else if (bT) res = aa;
                    else res = aa, emit = true;
                } else if (op == OP_ANDNOT) {  // a & !b
                    if (aF || bT) res = TAG_FALSE, n = first;
                    else if (bF) res = aa;
                    else if (aT) res = bb, emit_not = true;  // b's code is last
                    else res = aa, emit = true;
                } else {  // OP_RANDNOT: !a & b
                    if (aT || bF) res = TAG_FALSE, n = first;
                    else if (aF) res = bb;
                    else if (bT) res = aa, emit_not = true;  // a's code is last
                    else res = aa, emit = true;
                }
                if (emit || emit_not) {
                    if (o) o[n] = emit ? in : Insn{OP_NOT, 0};
                    n++;
                }
                S[sp++] = res;
            }
            peak = max(peak, n);
        }
        const uint32_t root = S[0];
        const uint32_t len = root == TAG_FALSE ? 0u : (root == TAG_TRUE ? BRICK_TRUE : n);
        const bool active = len != 0 && len != BRICK_TRUE;
        if (PASS == 0) {
            counts[b] = make_uint2(active ? peak : 0u, active ? 1u : 0u);
        } else {
            info[b] = make_uint2(offs[b].x, len);
            if (active) act_list[offs[b].y] = (uint32_t)b;  // y: exclusive count of active bricks
        }
    }
}

template <int MODE, bool STATS>
__device__ __forceinline__ bool eval_brick(const DevProg &P, const Bricks &B, const Grid &g, int bx, int y, int z,
                                           double px, double py, double pz, GStats &gs) {
    if (!B.info) return eval_code<MODE, STATS>(P.code, P.ncode, P, px, py, pz, gs);
    const uint2 bi = B.info[((size_t)(z / B.BZ) * B.nby + (y / B.BY)) * g.W + bx];
    if (bi.y == 0) return false;
    if (bi.y == BRICK_TRUE) return true;
    return eval_code<MODE, STATS>(B.code + bi.x, bi.y, P, px, py, pz, gs);
}

// corners: one wave per occupancy word (64 x corners of one row).
template <int MODE, bool STATS>
__global__ __launch_bounds__(256) void k_corners(DevProg P, Bricks B, Grid g, uint64_t *__restrict__ bits,
                                                 uint32_t *__restrict__ flags, unsigned long long *__restrict__ gst,
                                                 uint64_t nwords) {
    const int lane = threadIdx.x & 63;
    GStats gs;
    for (uint64_t w = (uint64_t)blockIdx.x * 4 + (threadIdx.x >> 6); w < nwords; w += (uint64_t)gridDim.x * 4) {
        const int wx = (int)(w % g.W);
        const uint64_t row = w / g.W;
        const int y = (int)(row % g.ny), z = (int)(row / g.ny);
        const int x = wx * 64 + lane;
        bool inside = false;
        if (x < g.nx) inside = eval_brick<MODE, STATS>(P, B, g, wx, y, z, g.xs[x], g.ys[y], g.zs[z], gs);
        uint64_t m = __ballot(inside);
        if (lane == 0) bits[w] = m;
        // panic("solid is true outside of bounds"), mc.go:687-689
        if (inside && (x == 0 || x == g.nx - 1 || y == 0 || y == g.ny - 1 || z == 0 || z == g.nz - 1))
            atomicOr(flags, 1u);
    }
    flush_stats<STATS>(gs, gst);
}

// Culled corners, part 1: words of constant bricks are written directly (all
// false, or every valid corner true).
__global__ __launch_bounds__(256) void k_fill_const(Bricks B, Grid g, uint64_t *__restrict__ bits,
                                                    uint32_t *__restrict__ flags, uint64_t nwords) {
    for (uint64_t w = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; w < nwords;
         w += (uint64_t)gridDim.x * blockDim.x) {
        const int wx = (int)(w % g.W);
        const uint64_t row = w / g.W;
        const int y = (int)(row % g.ny), z = (int)(row / g.ny);
        const uint32_t len = B.info[((size_t)(z / B.BZ) * B.nby + (y / B.BY)) * g.W + wx].y;
        if (len == 0) {
            bits[w] = 0;
        } else if (len == BRICK_TRUE) {
            const int nvalid = min(64, g.nx - wx * 64);
            bits[w] = nvalid == 64 ? ~0ull : ((1ull << nvalid) - 1);
            // panic("solid is true outside of bounds"), mc.go:687-689
            if (y == 0 || y == g.ny - 1 || z == 0 || z == g.nz - 1 || wx == 0 || wx * 64 + nvalid == g.nx)
                atomicOr(flags, 1u);
        }
    }
}

// Culled corners, part 2: one workgroup per active brick; its 4 waves take
// the brick's rows, all sharing the brick program (scalar loads).
template <int MODE, bool STATS>
__global__ __launch_bounds__(256) void k_corners_act(DevProg P, Bricks B, Grid g, const uint32_t *__restrict__ act,
                                                     uint32_t nact, uint64_t *__restrict__ bits,
                                                     uint32_t *__restrict__ flags,
                                                     unsigned long long *__restrict__ gst) {
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    GStats gs;
    for (uint32_t k = blockIdx.x; k < nact; k += gridDim.x) {
        const uint32_t b = act[k];
        const int bx = (int)(b % g.W);
        const uint32_t r = b / g.W;
        const int by = (int)(r % B.nby), bz = (int)(r / B.nby);
        const uint2 bi = B.info[b];
        const int x = bx * 64 + lane;
        for (int j = wave; j < B.BY * B.BZ; j += 4) {
            const int y = by * B.BY + j % B.BY, z = bz * B.BZ + j / B.BY;
            if (y >= g.ny || z >= g.nz) continue;  // wave-uniform
            bool inside = false;
            if (x < g.nx) inside = eval_code<MODE, STATS>(B.code + bi.x, bi.y, P, g.xs[x], g.ys[y], g.zs[z], gs);
            const uint64_t m = __ballot(inside);
            if (lane == 0) bits[word_index(g, bx, y, z)] = m;
            if (inside && (x == 0 || x == g.nx - 1 || y == 0 || y == g.ny - 1 || z == 0 || z == g.nz - 1))
                atomicOr(flags, 1u);
        }
    }
    flush_stats<STATS>(gs, gst);
}

__device__ __forceinline__ bool corner_bit(const Grid &g, const uint64_t *__restrict__ bits, int x, int y, int z) {
    return (bits[word_index(g, x >> 6, y, z)] >> (x & 63)) & 1;
}

// bisect: mcSearchPoint (mc.go:223-258) per vertex; iters == 0 leaves the
// edge midpoints of MarchingCubes (Coord3D.Mid, coords.go:196-198).  The
// edge lies inside the brick of its low corner (bricks include their +1
// corners), so that brick's program decides every midpoint.
template <int MODE, bool STATS>
__global__ __launch_bounds__(256) void k_bisect(DevProg P, Bricks B, Grid g, const uint64_t *__restrict__ bits,
                                                const uint64_t *__restrict__ vkeys, double *__restrict__ verts,
                                                unsigned long long *__restrict__ gst, uint64_t nverts, int iters) {
    GStats gs;
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
                if (eval_brick<MODE, STATS>(P, B, g, x >> 6, y, z, p[0], p[1], p[2], gs)) tp = mid;
                else fp = mid;
            }
            res = (fp + tp) / 2;
        }
        p[axis] = res;
        verts[3 * v] = p[0];
        verts[3 * v + 1] = p[1];
        verts[3 * v + 2] = p[2];
    }
    flush_stats<STATS>(gs, gst);
}

template <int MODE>
__global__ __launch_bounds__(256) void k_contains(DevProg P, const double *__restrict__ pts, uint8_t *__restrict__ out,
                                                  uint64_t n) {
    GStats gs;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x)
        out[i] = eval_code<MODE, false>(P.code, P.ncode, P, pts[3 * i], pts[3 * i + 1], pts[3 * i + 2], gs) ? 1 : 0;
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
    DevBuf prog_code, prog_f64, prog_f32, prog_frames, prog_boxes, table;
    DevBuf axes, bits, counts, offs, tiles, totals, flags, gstats, vkeys, pts, out8;
    DevBuf bcounts, boffs, binfo, bcode, actb, actw;
    hipEvent_t ev[8] = {};
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
    if ((rc = ctx->prog_boxes.ensure(P.code.size() * sizeof(AtomBox) + 16))) return rc;
    std::vector<AtomBox> boxes(P.code.size());
    for (size_t i = 0; i < P.atom_pc.size(); i++) boxes[P.atom_pc[i]] = P.atom_box[i];
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
    HIP_TRY(hipMemcpyAsync(ctx->prog_boxes.p, boxes.data(), boxes.size() * sizeof(AtomBox), hipMemcpyHostToDevice,
                           ctx->stream));
    // The host vectors above are pageable: the copies complete before return.
    HIP_TRY(hipStreamSynchronize(ctx->stream));
    dp.code = ctx->prog_code.as<Insn>();
    dp.f64 = ctx->prog_f64.as<double>();
    dp.f32 = ctx->prog_f32.as<float>();
    dp.frames = ctx->prog_frames.as<Frame>();
    dp.boxes = ctx->prog_boxes.as<AtomBox>();
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

// Exclusive scan of n uint2 values from `in` into `out`; totals (64-bit) land
// in ctx->totals.
int scan2(mc3d_ctx *ctx, const uint2 *in, uint2 *out, uint64_t n) {
    const uint64_t ntiles = (n + SCAN_TILE - 1) / SCAN_TILE;
    int rc;
    if ((rc = ctx->tiles.ensure(ntiles * sizeof(uint2)))) return rc;
    k_scan_reduce<<<(unsigned)ntiles, SCAN_BLOCK, 0, ctx->stream>>>(in, n, ctx->tiles.as<uint2>());
    k_scan_tiles<<<1, SCAN_BLOCK, 0, ctx->stream>>>(ctx->tiles.as<uint2>(), ntiles,
                                                     ctx->totals.as<unsigned long long>());
    k_scan_down<<<(unsigned)ntiles, SCAN_BLOCK, 0, ctx->stream>>>(in, n, ctx->tiles.as<uint2>(), out);
    HIP_TRY(hipGetLastError());
    return MC3D_OK;
}

template <int MODE, bool STATS>
void launch_corners(mc3d_ctx *ctx, const DevProg &dp, const Bricks &B, const Grid &g, uint64_t nwords,
                    uint32_t nact) {
    if (!B.info) {
        k_corners<MODE, STATS><<<grid_for(nwords, 4), 256, 0, ctx->stream>>>(
            dp, B, g, ctx->bits.as<uint64_t>(), ctx->flags.as<uint32_t>(), ctx->gstats.as<unsigned long long>(),
            nwords);
        return;
    }
    k_fill_const<<<grid_for(nwords, 256), 256, 0, ctx->stream>>>(B, g, ctx->bits.as<uint64_t>(),
                                                                   ctx->flags.as<uint32_t>(), nwords);
    if (nact)
        k_corners_act<MODE, STATS><<<grid_for(nact, 1), 256, 0, ctx->stream>>>(
            dp, B, g, ctx->actb.as<uint32_t>(), nact, ctx->bits.as<uint64_t>(), ctx->flags.as<uint32_t>(),
            ctx->gstats.as<unsigned long long>());
}

template <int MODE, bool STATS>
void launch_bisect(mc3d_ctx *ctx, const DevProg &dp, const Bricks &B, const Grid &g, mc3d_mesh *m, int iters) {
    k_bisect<MODE, STATS><<<grid_for(m->nv, 256), 256, 0, ctx->stream>>>(
        dp, B, g, ctx->bits.as<uint64_t>(), ctx->vkeys.as<uint64_t>(), m->dverts.as<double>(),
        ctx->gstats.as<unsigned long long>(), m->nv, iters);
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
    o.precision = MC3D_PREC_FP32_GUARD;
    o.cull = 1;
    if (opts) o = *opts;
    const int mode = o.precision == MC3D_PREC_FP32_GUARD ? 1 : 0;
    const bool stats = o.count_mismatch != 0;
    HIP_TRY(hipSetDevice(ctx->device));

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
    if (nwords >= (1ull << 32)) {
        set_error("grid too large: more than 2^32 occupancy words");
        return MC3D_E_UNSUPPORTED;
    }
    if ((rc = ctx->bits.ensure(nwords * 8))) return rc;
    if ((rc = ctx->actw.ensure(nwords * sizeof(uint32_t)))) return rc;
    if ((rc = ctx->counts.ensure(nwords * sizeof(uint2)))) return rc;
    if ((rc = ctx->offs.ensure(nwords * sizeof(uint2)))) return rc;
    if ((rc = ctx->totals.ensure(2 * sizeof(unsigned long long)))) return rc;
    if ((rc = ctx->flags.ensure(64))) return rc;
    if ((rc = ctx->gstats.ensure(64))) return rc;
    HIP_TRY(hipMemsetAsync(ctx->flags.p, 0, 64, ctx->stream));
    HIP_TRY(hipMemsetAsync(ctx->gstats.p, 0, 64, ctx->stream));
    // The spacer vectors are pageable host memory: finish their copies while
    // they are alive.
    HIP_TRY(hipStreamSynchronize(ctx->stream));

    std::unique_ptr<mc3d_mesh> mesh(new (std::nothrow) mc3d_mesh);
    mc3d_mesh *m = mesh.get();
    if (!m) return MC3D_E_OOM;
    m->ctx = ctx;
    mc3d_stats &st = m->stats;

    // ---- plan ----
    Bricks B{nullptr, nullptr, 1, 1, 1, 1};
    HIP_TRY(hipEventRecord(ctx->ev[0], ctx->stream));
    if (o.cull) {
        int bsz = o.brick > 0 ? o.brick : 4;
        B.BY = B.BZ = bsz;
        B.nby = (g.ny + bsz - 1) / bsz;
        B.nbz = (g.nz + bsz - 1) / bsz;
        const uint64_t nb = (uint64_t)g.W * B.nby * B.nbz;
        if (nb >= (1ull << 32)) {
            set_error("too many culling bricks");
            return MC3D_E_UNSUPPORTED;
        }
        if ((rc = ctx->bcounts.ensure(nb * sizeof(uint2))) || (rc = ctx->boffs.ensure(nb * sizeof(uint2))) ||
            (rc = ctx->binfo.ensure(nb * sizeof(uint2))) || (rc = ctx->actb.ensure(nb * sizeof(uint32_t))))
            return rc;
        k_plan<0><<<grid_for(nb, PLAN_BLOCK), PLAN_BLOCK, 0, ctx->stream>>>(dp, g, B, nb, ctx->bcounts.as<uint2>(),
                                                                          nullptr, nullptr, nullptr, nullptr);
        HIP_TRY(hipGetLastError());
        if ((rc = scan2(ctx, ctx->bcounts.as<uint2>(), ctx->boffs.as<uint2>(), nb))) return rc;
        unsigned long long bt[2];
        HIP_TRY(hipMemcpyAsync(bt, ctx->totals.p, sizeof bt, hipMemcpyDeviceToHost, ctx->stream));
        HIP_TRY(hipStreamSynchronize(ctx->stream));
        if (bt[0] >= (1ull << 32)) {
            set_error("brick programs exceed 2^32 instructions");
            return MC3D_E_UNSUPPORTED;
        }
        if ((rc = ctx->bcode.ensure(bt[0] * sizeof(Insn) + 16))) return rc;
        k_plan<1><<<grid_for(nb, PLAN_BLOCK), PLAN_BLOCK, 0, ctx->stream>>>(
            dp, g, B, nb, ctx->bcounts.as<uint2>(), ctx->boffs.as<uint2>(), ctx->binfo.as<uint2>(),
            ctx->bcode.as<Insn>(), ctx->actb.as<uint32_t>());
        HIP_TRY(hipGetLastError());
        B.info = ctx->binfo.as<uint2>();
        B.code = ctx->bcode.as<Insn>();
        st.bricks_total = nb;
        st.bricks_active = bt[1];
    }
    HIP_TRY(hipEventRecord(ctx->ev[1], ctx->stream));

    // ---- corners ----
    const uint32_t nact = (uint32_t)st.bricks_active;
    if (mode == 0)
        stats ? launch_corners<0, true>(ctx, dp, B, g, nwords, nact) : launch_corners<0, false>(ctx, dp, B, g, nwords, nact);
    else
        stats ? launch_corners<1, true>(ctx, dp, B, g, nwords, nact) : launch_corners<1, false>(ctx, dp, B, g, nwords, nact);
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipEventRecord(ctx->ev[2], ctx->stream));

    // ---- classify + scan ----
    uint32_t *nactw_dev = ctx->flags.as<uint32_t>() + 1;  // words with vertices or triangles
    k_classify<<<grid_for(nwords, 256), 256, 0, ctx->stream>>>(g, ctx->bits.as<uint64_t>(), ctx->table.as<uint8_t>(),
                                                                 ctx->counts.as<uint2>(), ctx->actw.as<uint32_t>(),
                                                                 nactw_dev, nwords);
    HIP_TRY(hipGetLastError());
    if ((rc = scan2(ctx, ctx->counts.as<uint2>(), ctx->offs.as<uint2>(), nwords))) return rc;
    HIP_TRY(hipEventRecord(ctx->ev[3], ctx->stream));
    unsigned long long totals[2];
    uint32_t flag[2] = {0, 0};
    HIP_TRY(hipMemcpyAsync(totals, ctx->totals.p, sizeof totals, hipMemcpyDeviceToHost, ctx->stream));
    HIP_TRY(hipMemcpyAsync(flag, ctx->flags.p, sizeof flag, hipMemcpyDeviceToHost, ctx->stream));
    HIP_TRY(hipStreamSynchronize(ctx->stream));
    const uint32_t nactw = flag[1];
    if (flag[0]) {
        set_error("solid is true outside of bounds");
        return MC3D_E_INSIDE_BOUNDARY;
    }
    if (totals[0] >= (1ull << 32) || totals[1] >= (1ull << 32)) {
        set_error("mesh too large for 32-bit vertex indices");
        return MC3D_E_UNSUPPORTED;
    }
    const uint64_t nv = totals[0], nt = totals[1];
    m->nv = nv;
    m->nt = nt;
    if ((rc = ctx->vkeys.ensure(nv * 8 + 8)) || (rc = m->dverts.ensure(nv * 24 + 8)) ||
        (rc = m->dtris.ensure(nt * 12 + 8)))
        return rc;

    // ---- emit ----
    HIP_TRY(hipEventRecord(ctx->ev[4], ctx->stream));
    if (nactw) {
        k_emit<<<grid_for(nactw, 4), 256, 0, ctx->stream>>>(g, ctx->bits.as<uint64_t>(), ctx->table.as<uint8_t>(),
                                                              ctx->offs.as<uint2>(), ctx->actw.as<uint32_t>(),
                                                              nactw_dev, ctx->vkeys.as<uint64_t>(),
                                                              m->dtris.as<uint32_t>());
        HIP_TRY(hipGetLastError());
    }
    HIP_TRY(hipEventRecord(ctx->ev[5], ctx->stream));

    // ---- bisect ----
    if (nv) {
        if (mode == 0) stats ? launch_bisect<0, true>(ctx, dp, B, g, m, iters) : launch_bisect<0, false>(ctx, dp, B, g, m, iters);
        else stats ? launch_bisect<1, true>(ctx, dp, B, g, m, iters) : launch_bisect<1, false>(ctx, dp, B, g, m, iters);
        HIP_TRY(hipGetLastError());
    }
    HIP_TRY(hipEventRecord(ctx->ev[6], ctx->stream));
    unsigned long long gs[4] = {0, 0, 0, 0};
    if (stats) HIP_TRY(hipMemcpyAsync(gs, ctx->gstats.p, sizeof gs, hipMemcpyDeviceToHost, ctx->stream));
    HIP_TRY(hipStreamSynchronize(ctx->stream));

    st.corners = (uint64_t)g.nx * g.ny * g.nz;
    st.vertices = nv;
    st.triangles = nt;
    st.bisection_evals = iters > 0 ? nv * (uint64_t)(iters + 1) : 0;
    st.guard_evals = gs[0];
    st.guard_rechecks = gs[1];
    st.guard_mismatches = gs[2];
    st.guard_violations = gs[3];
    st.ms_plan = elapsed(ctx->ev[0], ctx->ev[1]);
    st.ms_corners = elapsed(ctx->ev[1], ctx->ev[2]);
    st.ms_classify = elapsed(ctx->ev[2], ctx->ev[3]);
    st.ms_emit = elapsed(ctx->ev[4], ctx->ev[5]);
    st.ms_bisect = elapsed(ctx->ev[5], ctx->ev[6]);
    st.ms_device = st.ms_plan + st.ms_corners + st.ms_classify + st.ms_emit + st.ms_bisect;
    if (!o.device_only) {
        auto t0 = std::chrono::steady_clock::now();
        m->hverts.resize(nv * 3);
        m->htris.resize(nt * 3);
        if (nv) HIP_TRY(hipMemcpy(m->hverts.data(), m->dverts.p, nv * 24, hipMemcpyDeviceToHost));
        if (nt) HIP_TRY(hipMemcpy(m->htris.data(), m->dtris.p, nt * 12, hipMemcpyDeviceToHost));
        m->on_host = true;
        st.ms_d2h = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    }
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
    if (precision == MC3D_PREC_FP32_GUARD)
        k_contains<1><<<grid_for(n, 256), 256, 0, ctx->stream>>>(dp, ctx->pts.as<double>(), ctx->out8.as<uint8_t>(), n);
    else
        k_contains<0><<<grid_for(n, 256), 256, 0, ctx->stream>>>(dp, ctx->pts.as<double>(), ctx->out8.as<uint8_t>(), n);
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipMemcpyAsync(out, ctx->out8.p, n, hipMemcpyDeviceToHost, ctx->stream));
    HIP_TRY(hipStreamSynchronize(ctx->stream));
    return MC3D_OK;
}

}  // extern "C"
