// kernels.h -- the device stages of mc3d_marching_cubes_search
// (model3d/mc.go:14-49), CDNA4 / gfx950, wave64.
//
//   plan      per brick (64 x BY x BZ corners), the solid's program folded
//             against the brick's box: atoms whose conservative box misses
//             the brick are false, bounds checks and spheres that contain it
//             are true.  Constant bricks need no evaluation at all.
//   corners   Solid.Contains at every grid corner into a bit-packed occupancy
//             grid, one 64-bit word per 64 consecutive x corners of a row:
//             constant bricks are filled directly, each active brick gets a
//             workgroup whose waves share the brick program (scalar loads)
//             and pack their row with a ballot (mc.go:614-692).
//   classify  per word, crossing masks of the x/y/z grid edges (one vertex
//             per crossed edge: the reference's vertex dedup, mesh.go:822-849)
//             and the triangle count of its 64 cubes (mcLookupTable,
//             mc.go:434-456); one exclusive scan gives vertex and triangle
//             offsets and the list of words that emit anything.
//   emit      one wave per listed word, a lane per cube: vertex keys (the
//             owning edge) and triangles as uint32 vertex-index triples in
//             table order (mc.go:20-35, 401-407).
//   bisect    per vertex, mcSearchPoint (mc.go:223-258): the first probe is
//             the edge's high corner, read from the occupancy grid (same
//             point, same answer), then `iters` float64 midpoints.
//
// One launch of each stage serves every solid of a call: a call is a list of
// segments (one per solid, a single MarchingCubesSearch is a list of one),
// laid end to end in the occupancy words (ranges aligned to 64 words), the
// bricks (aligned to 128) and the outputs (by the scan).  A wave finds its
// segment from a table per 64 words / 128 bricks, so it is wave-uniform and
// the segment's program, grid and brick shape are scalar loads.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <type_traits>

#include "eval.h"

namespace mc3d {

constexpr uint32_t BRICK_TRUE = 0xffffffffu;  // info.y of a brick where the solid is true
constexpr uint32_t BRICK_FULL = 0xffffffffu;  // info.x of a brick that uses the unfolded program

// Per-call control block in device memory (zeroed by the upload).
struct Ctl {
    uint32_t inside;       // panic("solid is true outside of bounds") seen
    uint32_t overflow;     // an output index exceeded the buffers' capacity
    uint32_t nact_bricks;  // bricks whose folded program is not constant
    uint32_t pad;
    unsigned long long totals[4];  // scan totals: vertices, triangles, listed words
    unsigned long long gstats[4];  // guard statistics (GStats order)
    unsigned long long flops[4];   // statistics runs: float64, float32 ops of corners; the same of bisect
    unsigned long long code_total; // instructions over all active bricks' folded programs
    uint32_t code_max;             // longest folded program
    uint32_t full_bricks;          // bricks whose folded program overflowed its slot
};

struct Grid {
    const double *__restrict__ xs;
    const double *__restrict__ ys;
    const double *__restrict__ zs;
    int nx, ny, nz, W;  // W = words per row = ceil(nx / 64)
};

// Per-brick programs: info[b] = {offset into code (or BRICK_FULL), length};
// length 0: the solid is false everywhere in the brick; BRICK_TRUE: true.
struct Bricks {
    const uint2 *__restrict__ info;  // null: no culling, use the full program
    const Insn *__restrict__ code;
    int BY, BZ, nby, nbz;
};

// One solid of a call.  B.info / B.code are the call's arrays, indexed by the
// global brick number; word and brick numbers are local to the segment
// unless named global.  Kernels also take the first segment by value (a
// call of one solid reads it from the kernel arguments, no load).
struct Seg {
    DevProg P;
    Grid g;
    Bricks B;
    uint64_t word0, nwords;    // global occupancy words [word0, word0 + nwords); word0 % 64 == 0
    uint64_t brick0, nbricks;  // global bricks; brick0 % 128 == 0
    int32_t iters;
    int32_t stage;  // the plan may stage this segment's program in LDS
};

// Per-solid results, copied back with the control block.
struct SegOut {
    unsigned long long vbase, tbase;  // first vertex / triangle of the solid in the call's outputs
    uint32_t inside;                  // panic("solid is true outside of bounds") for this solid
    uint32_t pad[3];
};

constexpr int SEG_WORD_SHIFT = 6;   // segment table entry per 64 words
constexpr int SEG_BRICK_SHIFT = 7;  // and per 128 bricks

// The segment of a wave's items (all its items share one table entry).
// MULTI: the call has several segments; otherwise segment 0, whose copy in
// the kernel arguments is used (no load at all).
template <bool MULTI>
__device__ __forceinline__ uint32_t wave_seg(const uint32_t *__restrict__ table, uint64_t item, int shift) {
    if (!MULTI) return 0;
    return (uint32_t)__builtin_amdgcn_readfirstlane((int)table[item >> shift]);
}

// Segment si by value (scalar registers): loaded for a call of several,
// the kernel-argument copy for a call of one; its pointers marked global.
template <bool MULTI>
__device__ __forceinline__ Seg pick_seg(const Seg *__restrict__ segs, const Seg &seg0, uint32_t si) {
    Seg s = MULTI ? segs[si] : seg0;
    s.P = as_global(s.P);
    s.g.xs = global_ptr(s.g.xs), s.g.ys = global_ptr(s.g.ys), s.g.zs = global_ptr(s.g.zs);
    s.B.info = global_ptr(s.B.info), s.B.code = global_ptr(s.B.code);
    return s;
}

__device__ __forceinline__ size_t word_index(const Grid &g, int wx, int y, int z) {
    return ((size_t)z * g.ny + y) * g.W + wx;
}

// Word index -> (word column, row y, row z).  Word indices fit in 32 bits
// (the host rejects larger grids), so the divisions stay 32-bit.
__device__ __forceinline__ void word_coords(const Grid &g, uint64_t w, int &wx, int &y, int &z) {
    const uint32_t w32 = (uint32_t)w;
    const uint32_t row = w32 / (uint32_t)g.W;
    wx = (int)(w32 - row * (uint32_t)g.W);
    const uint32_t zz = row / (uint32_t)g.ny;
    y = (int)(row - zz * (uint32_t)g.ny);
    z = (int)zz;
}

template <bool STATS>
__device__ __forceinline__ void flush_stats(const GStats &gs, unsigned long long *__restrict__ out,
                                            unsigned long long *__restrict__ flops) {
    if (!STATS) return;
    __shared__ unsigned long long red[6];
    if (threadIdx.x < 6) red[threadIdx.x] = 0;
    __syncthreads();
    if (gs.evals) atomicAdd(&red[0], (unsigned long long)gs.evals);
    if (gs.rechecks) atomicAdd(&red[1], (unsigned long long)gs.rechecks);
    if (gs.mismatches) atomicAdd(&red[2], (unsigned long long)gs.mismatches);
    if (gs.violations) atomicAdd(&red[3], (unsigned long long)gs.violations);
    if (gs.flops64) atomicAdd(&red[4], (unsigned long long)gs.flops64);
    if (gs.flops32) atomicAdd(&red[5], (unsigned long long)gs.flops32);
    __syncthreads();
    if (threadIdx.x < 4 && red[threadIdx.x]) atomicAdd(&out[threadIdx.x], red[threadIdx.x]);
    if (threadIdx.x >= 4 && threadIdx.x < 6 && red[threadIdx.x]) atomicAdd(&flops[threadIdx.x - 4], red[threadIdx.x]);
}

// Appends `item` for the lanes with `pred` set, one atomic per wave.  The
// list order is arbitrary: every consumer writes to places fixed by scans.
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

// s_waitcnt operand for vmcnt(0) alone (expcnt and lgkmcnt at their maxima).
constexpr int VMCNT0 = 0x0f70;

// Index of this wave in its block as a scalar: loops and row arithmetic
// derived from it stay in scalar registers.
__device__ __forceinline__ int wave_id() { return __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6)); }

__device__ __forceinline__ void wave_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// ---------------------------------------------------------------------------
// plan
// ---------------------------------------------------------------------------
// Atom against a brick box: 0 certainly false, 1 certainly true, 2 evaluate.
// "False" needs only the atom's conservative world box (solid.cpp widens it
// far beyond float64 rounding); "true" is proven for untransformed boxes and
// bounds checks (exact comparisons) and spheres (farthest brick corner inside
// the threshold with a relative margin far above the rounding of the sum).
// The program is read through the given pointer types (LDS, or constant
// memory with scalar loads).
template <class F64P, class BoxP>
__device__ __forceinline__ int classify_atom(F64P f64, BoxP boxes, uint32_t pc, Insn in, const double *lo,
                                             const double *hi) {
    AtomBox bx;
    for (int i = 0; i < 3; i++) bx.lo[i] = boxes[pc].lo[i], bx.hi[i] = boxes[pc].hi[i];
    if (hi[0] < bx.lo[0] || lo[0] > bx.hi[0] || hi[1] < bx.lo[1] || lo[1] > bx.hi[1] || hi[2] < bx.lo[2] ||
        lo[2] > bx.hi[2])
        return 0;
    if (insn_frame(in.x) == 0) {
        const auto a = f64 + in.y;
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

constexpr size_t PLAN_STACK_BYTES = PLAN_BLOCK * PLAN_STACK * sizeof(uint32_t);

// Folds the postfix program against one brick's box (lo, hi): a stack of
// {false, true, expression starting at output position n}; the folded code
// goes to o[0, cap) (`over` if it needs more).  S is the thread's stack.
template <class CodeP, class BoxP, class F64P>
__device__ __forceinline__ void plan_fold(CodeP code, BoxP boxes, F64P f64, uint32_t ncode, uint32_t *S,
                                          const double *lo, const double *hi, Insn *o, uint32_t cap, uint32_t &n,
                                          bool &over) {
    int sp = 0;
    for (uint32_t pc = 0; pc < ncode; ++pc) {
        const Insn in{code[pc].x, code[pc].y};  // field-wise: the pointer may be in any address space
        const uint32_t op = insn_op(in.x);
        if (op == OP_ATOM) {
            int c = classify_atom(f64, boxes, pc, in, lo, hi);
            if (c == 2) {
                if (n < cap) o[n] = in;
                else over = true;
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
                if (n < cap) o[n] = in;
                else over = true;
                n++;
            }
        } else {
            const uint32_t bb = S[--sp], aa = S[--sp];
            const bool aE = (aa & TAG_MASK) == TAG_EXPR, bE = (bb & TAG_MASK) == TAG_EXPR;
            const bool aT = aa == TAG_TRUE, aF = aa == TAG_FALSE, bT = bb == TAG_TRUE, bF = bb == TAG_FALSE;
            // start of the operands' code (constants emitted nothing)
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
                if (n < cap) o[n] = emit ? in : Insn{OP_NOT, 0};
                else over = true;
                n++;
            }
            S[sp++] = res;
        }
    }
}



// One thread per brick walks the postfix program with a stack of
// {false, true, expression starting at output position n} and writes the
// folded program into the brick's slot of `cap` instructions.  A brick whose
// folded code ever needs more than `cap` slots uses the unfolded program.
// A brick where the solid is constant gets its occupancy words here (all
// false, or every valid corner true, with the check that no boundary corner
// is inside: panic("solid is true outside of bounds"), mc.go:687-689); the
// stores cost little in a kernel whose waves mostly wait on the fold.
// With `stage`, the program, its atom boxes and its params are first copied
// to LDS.  The per-brick statistics are reduced per wave.
template <bool MULTI>
__global__ __launch_bounds__(PLAN_BLOCK) void k_plan(const Seg *__restrict__ segs, uint32_t nsegs, const Seg seg0,
                                                     const uint32_t *__restrict__ brick_seg, uint64_t nbricks_all,
                                                     uint32_t cap, uint2 *__restrict__ info, Insn *__restrict__ out,
                                                     uint32_t *__restrict__ act_list, uint64_t *__restrict__ bits_all,
                                                     Ctl *__restrict__ ctl, SegOut *__restrict__ sout) {
    extern __shared__ __align__(16) char plan_smem[];
    uint32_t *S = (uint32_t *)plan_smem + threadIdx.x * PLAN_STACK;
    Insn *sc = (Insn *)(plan_smem + PLAN_STACK_BYTES);
    uint32_t staged = ~0u;  // segment whose program is in LDS
    const int lane = threadIdx.x & 63;
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    const uint64_t b_end = (nbricks_all + blockDim.x - 1) / blockDim.x * blockDim.x;  // whole waves take part
    for (uint64_t gb = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; gb < b_end; gb += stride) {
        // A block's 128 bricks share one segment (ranges aligned to 128).
        const uint32_t si = wave_seg<MULTI>(brick_seg, gb - threadIdx.x, SEG_BRICK_SHIFT);
        const Seg sg = pick_seg<MULTI>(segs, seg0, si);
        const DevProg &P = sg.P;
        const Grid &g = sg.g;
        const Bricks &B = sg.B;
        if (sg.stage) {  // block-uniform
            AtomBox *sb = (AtomBox *)(sc + P.ncode);
            double *sf = (double *)(sb + P.ncode);
            if (staged != si) {
                __syncthreads();  // the previous segment's walks are done with the LDS copy
                for (uint32_t i = threadIdx.x; i < P.ncode; i += blockDim.x) sc[i] = P.code[i], sb[i] = P.boxes[i];
                for (uint32_t i = threadIdx.x; i < P.nf64; i += blockDim.x) sf[i] = P.f64[i];
                __syncthreads();
                staged = si;
            }
        }
        const uint64_t b = gb - sg.brick0;  // local brick
        bool active = false, full = false;
        uint32_t len = 0;
        if (gb < nbricks_all && b < sg.nbricks) {
            const uint32_t r = (uint32_t)b / (uint32_t)g.W;
            const int bx = (int)((uint32_t)b - r * (uint32_t)g.W);
            const int bz = (int)(r / (uint32_t)B.nby), by = (int)(r - (uint32_t)bz * (uint32_t)B.nby);
            const int x0 = bx * 64, y0 = by * B.BY, z0 = bz * B.BZ;
            const int x1 = min(x0 + 64, g.nx - 1), y1 = min(y0 + B.BY, g.ny - 1), z1 = min(z0 + B.BZ, g.nz - 1);
            const double lo[3] = {g.xs[x0], g.ys[y0], g.zs[z0]};
            const double hi[3] = {g.xs[x1], g.ys[y1], g.zs[z1]};
            Insn *o = out + gb * cap;
            uint32_t n = 0;
            bool over = false;
            // two copies of the fold, so that each reads with typed
            // instructions: LDS, or scalar loads for programs too large to
            // stage (vector loads would also wait for the fold's stores)
            if (sg.stage)
                plan_fold(sc, (const AtomBox *)(sc + P.ncode), (const double *)((const AtomBox *)(sc + P.ncode) + P.ncode),
                          P.ncode, S, lo, hi, o, cap, n, over);
            else
                plan_fold(scalar_ptr(P.code), scalar_ptr(P.boxes), scalar_ptr(P.f64), P.ncode, S, lo, hi, o, cap, n,
                          over);
            const uint32_t root = S[0];
            uint2 bi;
            if (root == TAG_FALSE) bi = make_uint2(0, 0);
            else if (root == TAG_TRUE) bi = make_uint2(0, BRICK_TRUE);
            else if (over) bi = make_uint2(BRICK_FULL, P.ncode), active = true;
            else bi = make_uint2((uint32_t)(gb * cap), n), active = true;
            info[gb] = bi;
            if (!active) {  // the constant brick's words
                const bool t = root == TAG_TRUE;
                const int nvalid = min(64, g.nx - x0);
                const uint64_t word = t ? (nvalid == 64 ? ~0ull : ((1ull << nvalid) - 1)) : 0ull;
                const int jz_end = min(B.BZ, g.nz - z0), jy_end = min(B.BY, g.ny - y0);
                const size_t W = (size_t)g.W, plane = (size_t)g.ny * W;
                uint64_t *zrow = bits_all + sg.word0 + word_index(g, bx, y0, z0);
                for (int jz = 0; jz < jz_end; jz++, zrow += plane)
                    for (int jy = 0; jy < jy_end; jy++) zrow[(size_t)jy * W] = word;
                if (t && (bx == 0 || x0 + nvalid == g.nx || y0 == 0 || y0 + jy_end == g.ny || z0 == 0 ||
                          z0 + jz_end == g.nz))
                    atomicOr(&ctl->inside, 1u), atomicOr(&sout[si].inside, 1u);
            }
            len = active ? bi.y : 0;
            full = over;
        }
        // per-wave statistics: one atomic each
        uint32_t sum = len, mx = len;
        for (int d = 32; d; d >>= 1) sum += __shfl_xor(sum, d), mx = max(mx, (uint32_t)__shfl_xor(mx, d));
        const uint32_t nfull = (uint32_t)__popcll(__ballot(full));
        if (lane == 0 && sum) {
            atomicAdd(&ctl->code_total, (unsigned long long)sum);
            atomicMax(&ctl->code_max, mx);
            if (nfull) atomicAdd(&ctl->full_bricks, nfull);
        }
        wave_append(active, (uint32_t)gb, act_list, &ctl->nact_bricks);
    }
}

// ---------------------------------------------------------------------------
// corners
// ---------------------------------------------------------------------------
// Without culling: one wave per occupancy word.
template <int MODE, bool STATS, bool MULTI>
__global__ __launch_bounds__(256) void k_corners(const Seg *__restrict__ segs, uint32_t nsegs, const Seg seg0,
                                                 const uint32_t *__restrict__ word_seg, uint64_t *__restrict__ bits,
                                                 Ctl *__restrict__ ctl, SegOut *__restrict__ sout,
                                                 uint64_t nwords_all) {
    const int lane = threadIdx.x & 63;
    GStats gs;
    for (uint64_t gw = (uint64_t)blockIdx.x * 4 + wave_id(); gw < nwords_all; gw += (uint64_t)gridDim.x * 4) {
        const uint32_t si = wave_seg<MULTI>(word_seg, gw, SEG_WORD_SHIFT);
        const Seg sg = pick_seg<MULTI>(segs, seg0, si);
        const DevProg &P = sg.P;
        const Grid &g = sg.g;
        const uint64_t w = gw - sg.word0;
        if (w >= sg.nwords) {  // alignment padding
            if (lane == 0) bits[gw] = 0;
            continue;
        }
        int wx, y, z;
        word_coords(g, w, wx, y, z);
        const int x = wx * 64 + lane;
        bool inside = false;
        // Point id of a corner (host-evaluated atoms): x + nx * (y + ny * z).
        const uint64_t pid = (uint64_t)x + (uint64_t)g.nx * ((uint64_t)y + (uint64_t)g.ny * (uint64_t)z);
        if (x < g.nx) inside = eval_code<MODE, STATS, true, true>(P.code, P.ncode, P, g.xs[x], g.ys[y], g.zs[z], gs, pid);
        const uint64_t m = __ballot(inside);
        if (lane == 0) bits[gw] = m;
        // panic("solid is true outside of bounds"), mc.go:687-689
        if (inside && (x == 0 || x == g.nx - 1 || y == 0 || y == g.ny - 1 || z == 0 || z == g.nz - 1))
            atomicOr(&ctl->inside, 1u), atomicOr(&sout[si].inside, 1u);
    }
    flush_stats<STATS>(gs, ctl->gstats, ctl->flops);
}

__device__ __forceinline__ const Insn *brick_code(const DevProg &P, const Bricks &B, uint2 bi) {
    return bi.x == BRICK_FULL ? P.code : B.code + bi.x;
}

// One row of 64 corners decided: the occupancy word (at `word`) and, for
// bricks on the grid's boundary, the check that no boundary corner is
// inside (panic("solid is true outside of bounds"), mc.go:687-689); xedge:
// the lanes on the x faces, redge: the row lies on a y or z face.
__device__ __forceinline__ void row_out(uint64_t *__restrict__ word, bool inside, bool edge, uint64_t xedge,
                                        bool redge, Ctl *__restrict__ ctl, SegOut *__restrict__ so) {
    const uint64_t m = __ballot(inside);
    if ((threadIdx.x & 63) == 0) {
        *word = m;
        if (edge && (m & (redge ? ~0ull : xedge))) atomicOr(&ctl->inside, 1u), atomicOr(&so->inside, 1u);
    }
}

// A row of a brick as a wave walks it: its coordinates, its word (index in
// the segment's occupancy grid) and whether it lies on a y or z face.
struct RowInfo {
    double y, z;
    uint32_t word, redge;
};

// Lane i describes the wave's row i (row j = wave + WPB * i of the brick,
// j = jz * BY + jy) in rows[i], once per brick, so the walks read a row with
// one broadcast instead of stepping and loading; returns the rows inside the
// grid as a mask.  Needs BY * BZ <= 64 * WPB.
template <int WPB>
__device__ __forceinline__ uint64_t wave_rows(const Grid &g, const Bricks &B, int bx, int by, int bz, int wave,
                                              RowInfo *rows) {
    const int lane = threadIdx.x & 63;
    const int j = wave + WPB * lane, jy = j % B.BY, jz = j / B.BY;
    const int y = by * B.BY + jy, z = bz * B.BZ + jz;
    const bool ok = j < B.BY * B.BZ && y < g.ny && z < g.nz;
    if (ok)
        rows[lane] = RowInfo{g.ys[y], g.zs[z], (uint32_t)word_index(g, bx, y, z),
                             (y == 0 || y == g.ny - 1 || z == 0 || z == g.nz - 1) ? 1u : 0u};
    wave_sync();
    return __ballot(ok);
}

// The rows of a brick that wave `wave` of WPB takes (row j = jz * BY + jy
// with j % WPB == wave), each decided by eval(py, pz) for the lane's x.
template <int WPB, class F>
__device__ __forceinline__ void brick_rows(const Grid &g, const Bricks &B, uint64_t *__restrict__ bits, int bx,
                                           int by, int bz, int wave, int x, bool xin, bool edge,
                                           Ctl *__restrict__ ctl, SegOut *__restrict__ so, RowInfo *rows, F &&eval) {
    if (B.BY * B.BZ <= 64 * WPB) {  // the rows from LDS
        const uint64_t xedge = __ballot(x == 0 || x == g.nx - 1);
        for (uint64_t m = wave_rows<WPB>(g, B, bx, by, bz, wave, rows); m; m &= m - 1) {
            const RowInfo r = rows[__ffsll((unsigned long long)m) - 1];
            const bool inside = eval(r.y, r.z) && xin;
            row_out(bits + r.word, inside, edge, xedge, r.redge != 0, ctl, so);
        }
        return;
    }
    const int y0 = by * B.BY, z0 = bz * B.BZ;
    const int jz_end = min(B.BZ, g.nz - z0), jy_end = min(B.BY, g.ny - y0);
    const auto ys = scalar_ptr(g.ys) + y0, zs = scalar_ptr(g.zs) + z0;
    const uint64_t xedge = __ballot(x == 0 || x == g.nx - 1);
    const size_t W = (size_t)g.W, plane = (size_t)g.ny * W;
    uint64_t *zrow = bits + word_index(g, bx, y0, z0);  // word of row (0, jz)
    for (int jz = 0; jz < jz_end; jz++, zrow += plane) {
        const int z = z0 + jz;
        const double pz = zs[jz];
        const bool zedge = z == 0 || z == g.nz - 1;
        for (int jy = ((wave - jz * B.BY) % WPB + WPB) % WPB; jy < jy_end; jy += WPB) {
            const int y = y0 + jy;
            const bool inside = eval(ys[jy], pz) && xin;
            row_out(zrow + (size_t)jy * W, inside, edge, xedge, zedge || y == 0 || y == g.ny - 1, ctl, so);
        }
    }
}

// A folded program whose atoms are all untransformed spheres, boxes,
// capsules or cylinders (scalar loads: the walk is wave-uniform).
template <class CodeP>
__device__ __forceinline__ bool fast_atoms(CodeP code, uint32_t n) {
    for (uint32_t pc = 0; pc < n; ++pc) {
        const uint32_t x = (uint32_t)__builtin_amdgcn_readfirstlane((int)code[pc].x);
        if (insn_op(x) != OP_ATOM) continue;
        const uint32_t kind = insn_kind(x);
        if (insn_frame(x) != 0 || !(kind == A_SPHERE || kind == A_RECT || kind == A_CAPSULE || kind == A_CYLINDER))
            return false;
    }
    return true;
}

// Rows one wave of a brick can take in the instruction-major walk.
constexpr int WAVE_ROWS = 16;

// Culled, part 2: WPB waves per active brick.  A brick whose folded program
// is one untransformed atom has the atom kind dispatched once for all its
// rows; other programs are walked per NP rows.  Either way the walk is
// wave-uniform: instructions, params and row coordinates come in through
// scalar loads.
template <int MODE, bool STATS, int NP, int WPB, bool MULTI>
__global__ __launch_bounds__(64 * WPB) void k_corners_act(const Seg *__restrict__ segs, uint32_t nsegs, const Seg seg0,
                                                          const uint32_t *__restrict__ brick_seg,
                                                          const uint32_t *__restrict__ act,
                                                          uint64_t *__restrict__ bits_all, Ctl *__restrict__ ctl,
                                                          SegOut *__restrict__ sout) {
    const int lane = threadIdx.x & 63, wave = wave_id();
    const uint32_t nact = ctl->nact_bricks;
    __shared__ RowInfo row_info[WPB][64];  // each wave's rows of the current brick
    GStats gs;
    for (uint32_t k = blockIdx.x; k < nact; k += gridDim.x) {
        const uint32_t gb = scalar_ptr(act)[k];
        const uint32_t si = wave_seg<MULTI>(brick_seg, gb, SEG_BRICK_SHIFT);
        const Seg sg = pick_seg<MULTI>(segs, seg0, si);
        const DevProg &P = sg.P;
        const Grid &g = sg.g;
        const Bricks &B = sg.B;
        SegOut *so = &sout[si];
        const int nrows = B.BY * B.BZ;
        uint64_t *bits = bits_all + sg.word0;
        const uint32_t b = gb - (uint32_t)sg.brick0;
        const uint32_t r = b / (uint32_t)g.W;
        const int bx = (int)(b - r * (uint32_t)g.W);
        const int bz = (int)(r / (uint32_t)B.nby), by = (int)(r - (uint32_t)bz * (uint32_t)B.nby);
        // the brick touches the grid's boundary faces
        const bool edge = bx == 0 || bx * 64 + 64 >= g.nx || by == 0 || (by + 1) * B.BY >= g.ny || bz == 0 ||
                          (bz + 1) * B.BZ >= g.nz;
        const uint64_t biw = scalar_ptr((const uint64_t *)B.info)[gb];
        const uint2 bi = make_uint2((uint32_t)biw, (uint32_t)(biw >> 32));
        const auto code = scalar_ptr(brick_code(P, B, bi));
        const auto f64 = scalar_ptr(P.f64);
        const auto f32 = scalar_ptr(P.f32);
        const int x = bx * 64 + lane;
        const bool xin = x < g.nx;
        const double px0 = xin ? g.xs[x] : 0.0;
        const float fx0 = (float)px0;
        // Wait for that load here: one counter covers vector loads and
        // stores, so a wait for it inside the row loop would also wait for
        // the previous rows' stores to complete.
        __builtin_amdgcn_s_waitcnt(VMCNT0);
        const Insn in0{code[0].x, code[0].y};
        const uint32_t kind0 = insn_kind(in0.x);
        if (bi.y == 1 && insn_op(in0.x) == OP_ATOM && insn_frame(in0.x) == 0 &&
            (kind0 == A_SPHERE || kind0 == A_RECT || kind0 == A_CAPSULE || kind0 == A_CYLINDER)) {
            // one atom: dispatched here, once for the brick's rows
            auto rows = [&](auto KIND) {
                constexpr uint32_t K = decltype(KIND)::value;
                constexpr int N = K == A_SPHERE ? 7 : K == A_RECT ? 6 : K == A_CAPSULE ? 15 : 12;
                const auto p64 = reg_params<N, double>(f64 + in0.y);
                const auto p32 = reg_params<(MODE == 1 ? N : 1), float>(f32 + in0.y);
                brick_rows<WPB>(g, B, bits, bx, by, bz, wave, x, xin, edge, ctl, so, row_info[wave],
                                [&](double py, double pz) {
                    return eval_atom<MODE, STATS, false>(P, K, p64, p32, px0, py, pz, fx0, (float)py, (float)pz,
                                                         xin, 0, gs);
                });
            };
            if (kind0 == A_SPHERE) rows(std::integral_constant<uint32_t, A_SPHERE>());
            else if (kind0 == A_RECT) rows(std::integral_constant<uint32_t, A_RECT>());
            else if (kind0 == A_CAPSULE) rows(std::integral_constant<uint32_t, A_CAPSULE>());
            else rows(std::integral_constant<uint32_t, A_CYLINDER>());
            continue;
        }
        const uint64_t xedge = __ballot(x == 0 || x == g.nx - 1);
        const int wrows = (nrows - wave + WPB - 1) / WPB;  // rows of this wave: j = wave + WPB * i
        if (wrows <= WAVE_ROWS && fast_atoms(code, bi.y)) {
            // Several untransformed atoms of the four common kinds: each
            // instruction is decoded once for all the wave's rows (a bit
            // stack per row), each atom with its params in vector registers.
            // Row i of the wave from LDS (wave_rows; rows past the grid:
            // bit i of rmask clear).
            RowInfo *rows = row_info[wave];
            const uint32_t rmask = (uint32_t)wave_rows<WPB>(g, B, bx, by, bz, wave, rows);
            uint32_t st[WAVE_ROWS];
            for (int i = 0; i < WAVE_ROWS; i++) st[i] = 0;
            for (uint32_t pc = 0; pc < bi.y; ++pc) {
                Insn in{code[pc].x, code[pc].y};
                in.x = (uint32_t)__builtin_amdgcn_readfirstlane((int)in.x);
                in.y = (uint32_t)__builtin_amdgcn_readfirstlane((int)in.y);
                const uint32_t op = insn_op(in.x);
                if (op == OP_ATOM) {
                    auto atom_rows = [&](auto KIND) {
                        constexpr uint32_t K = decltype(KIND)::value;
                        constexpr int N = K == A_SPHERE ? 7 : K == A_RECT ? 6 : K == A_CAPSULE ? 15 : 12;
                        const auto p64 = reg_params<N, double>(f64 + in.y);
                        const auto p32 = reg_params<(MODE == 1 ? N : 1), float>(f32 + in.y);
                        for (int i = 0; i < WAVE_ROWS; i++)
                            if ((rmask >> i) & 1u) {  // wave-uniform
                                const double py = rows[i].y, pz = rows[i].z;
                                const bool v = eval_atom<MODE, STATS, false>(P, K, p64, p32, px0, py, pz, fx0,
                                                                             (float)py, (float)pz, xin, 0, gs);
                                st[i] = (st[i] << 1) | (uint32_t)v;
                            }
                    };
                    const uint32_t kind = insn_kind(in.x);
                    if (kind == A_SPHERE) atom_rows(std::integral_constant<uint32_t, A_SPHERE>());
                    else if (kind == A_RECT) atom_rows(std::integral_constant<uint32_t, A_RECT>());
                    else if (kind == A_CAPSULE) atom_rows(std::integral_constant<uint32_t, A_CAPSULE>());
                    else atom_rows(std::integral_constant<uint32_t, A_CYLINDER>());
                } else {
                    for (int i = 0; i < WAVE_ROWS; i++) st[i] = stack_op(op, st[i]);
                }
            }
            for (int i = 0; i < WAVE_ROWS; i++)
                if ((rmask >> i) & 1u) row_out(bits + rows[i].word, (st[i] & 1u) && xin, edge, xedge, rows[i].redge != 0, ctl, so);
            continue;
        }
        for (int j0 = wave; j0 < nrows; j0 += WPB * NP) {
            double px[NP], py[NP], pz[NP];
            bool ok[NP], live[NP], res[NP];
            int yy[NP], zz[NP];
            for (int q = 0; q < NP; q++) {
                const int j = j0 + WPB * q;
                yy[q] = by * B.BY + j % B.BY;
                zz[q] = bz * B.BZ + j / B.BY;
                ok[q] = j < nrows && yy[q] < g.ny && zz[q] < g.nz;  // wave-uniform
                live[q] = ok[q] && xin;
                px[q] = px0;
                py[q] = ok[q] ? scalar_ptr(g.ys)[yy[q]] : 0.0;
                pz[q] = ok[q] ? scalar_ptr(g.zs)[zz[q]] : 0.0;
            }
            eval_code_n<MODE, STATS, NP, true>(code, bi.y, P, f64, f32, scalar_ptr(P.frames), px, py, pz, live, res,
                                               gs);
            for (int q = 0; q < NP; q++)
                if (ok[q])  // wave-uniform
                    row_out(bits + word_index(g, bx, yy[q], zz[q]), res[q] && xin, edge, xedge,
                            yy[q] == 0 || yy[q] == g.ny - 1 || zz[q] == 0 || zz[q] == g.nz - 1, ctl, so);
        }
    }
    flush_stats<STATS>(gs, ctl->gstats, ctl->flops);
}

// ---------------------------------------------------------------------------
// exclusive scan of uint4 counts (block sums written by classify, a scan of
// those, downsweep)
// ---------------------------------------------------------------------------
constexpr int SCAN_BLOCK = 256;
constexpr int SCAN_ITEMS = 8;
constexpr int SCAN_TILE = SCAN_BLOCK * SCAN_ITEMS;
constexpr int CLASSIFY_BLOCK = 1024;  // words per classify block and per block sum
constexpr int SUMS_BLOCK = 1024;      // threads of the one block scanning the block sums
static_assert(SCAN_TILE % CLASSIFY_BLOCK == 0, "a scan tile starts at a classify block");

__device__ __forceinline__ uint4 add4(uint4 a, uint4 b) { return make_uint4(a.x + b.x, a.y + b.y, a.z + b.z, a.w + b.w); }

__device__ __forceinline__ uint4 sub4(uint4 a, uint4 b) { return make_uint4(a.x - b.x, a.y - b.y, a.z - b.z, a.w - b.w); }

// Exclusive scan over the block: an inclusive scan in each wave (shuffles),
// then the wave totals through LDS (sh: NT / 64 entries).
template <int NT = SCAN_BLOCK>
__device__ __forceinline__ uint4 block_exclusive(uint4 v, uint4 *sh, uint4 &total) {
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    uint4 inc = v;
    for (int d = 1; d < 64; d <<= 1) {
        const uint4 o = make_uint4(__shfl_up(inc.x, d), __shfl_up(inc.y, d), __shfl_up(inc.z, d), __shfl_up(inc.w, d));
        if (lane >= d) inc = add4(inc, o);
    }
    if (lane == 63) sh[wave] = inc;
    __syncthreads();
    uint4 pre = make_uint4(0, 0, 0, 0);
    total = make_uint4(0, 0, 0, 0);
    for (int w = 0; w < NT / 64; w++) {
        const uint4 t = sh[w];
        if (w < wave) pre = add4(pre, t);
        total = add4(total, t);
    }
    __syncthreads();  // sh is reused by the caller's next scan
    return add4(pre, sub4(inc, v));
}

// Sum over the block (sh: NT / 64 entries); every thread gets it.
template <int NT>
__device__ __forceinline__ uint4 block_sum(uint4 v, uint4 *sh) {
    for (int d = 32; d; d >>= 1)
        v = add4(v, make_uint4(__shfl_xor(v.x, d), __shfl_xor(v.y, d), __shfl_xor(v.z, d), __shfl_xor(v.w, d)));
    if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = v;
    __syncthreads();
    uint4 t = make_uint4(0, 0, 0, 0);
    for (int w = 0; w < NT / 64; w++) t = add4(t, sh[w]);
    return t;
}

// ---------------------------------------------------------------------------
// classify
// ---------------------------------------------------------------------------
struct Masks {
    uint64_t cx, cy, cz;
};

// Crossed-edge masks of one occupancy word: bit i of cx is the edge from
// corner (64wx+i, y, z) to its +x neighbour, and likewise for y and z.
// Padding bits (x >= nx) and boundary corners are zero, so no checks beyond
// the neighbour words' existence are needed.
// The 2 x 2 x 2 words from (wx, y, z): n[k] is word (wx + (k & 1),
// y + ((k >> 1) & 1), z + (k >> 2)), zero beyond the grid.  All eight are
// loaded from clamped addresses (no branches), so the loads issue together.
struct Nbhd {
    uint64_t n[8];
};
__device__ __forceinline__ Nbhd load_nbhd(const Grid &g, const uint64_t *__restrict__ bits, int wx, int y, int z) {
    const bool hx = wx + 1 < g.W, hy = y + 1 < g.ny, hz = z + 1 < g.nz;
    const size_t w = word_index(g, wx, y, z);
    const size_t dx = hx ? 1 : 0, dy = hy ? (size_t)g.W : 0, dz = hz ? (size_t)g.ny * g.W : 0;
    Nbhd r;
    for (int k = 0; k < 8; k++) r.n[k] = bits[w + ((k & 1) ? dx : 0) + ((k & 2) ? dy : 0) + ((k & 4) ? dz : 0)];
    for (int k = 0; k < 8; k++)
        if (((k & 1) && !hx) || ((k & 2) && !hy) || ((k & 4) && !hz)) r.n[k] = 0;
    return r;
}

// Crossed-edge masks of word n[0]: bit i of cx is the edge from corner
// (64wx+i, y, z) to its +x neighbour, and likewise for y and z.  Padding
// bits (x >= nx) and boundary corners are zero, so no checks beyond the
// neighbour words' existence are needed.
__device__ __forceinline__ Masks nbhd_masks(const Nbhd &h) {
    Masks m;
    m.cx = h.n[0] ^ ((h.n[0] >> 1) | (h.n[1] << 63));
    m.cy = h.n[0] ^ h.n[2];
    m.cz = h.n[0] ^ h.n[4];
    return m;
}

// With a block filter (MarchingCubesFilter): keep has a bit per cube of an
// accepted block, in the occupancy layout.  An edge gives a vertex only when
// a kept cube uses it: an x edge at corner (x, y, z) belongs to cubes
// (x, y-1..y, z-1..z), a y edge to (x-1..x, y, z-1..z), a z edge to
// (x-1..x, y-1..y, z).
__device__ __forceinline__ uint64_t keep_word(const Grid &g, const uint64_t *__restrict__ keep, int wx, int y, int z) {
    if (wx < 0 || y < 0 || z < 0) return 0;
    return keep[word_index(g, wx, y, z)];
}
__device__ __forceinline__ void keep_masks(const Grid &g, const uint64_t *__restrict__ keep, int wx, int y, int z,
                                           Masks &m) {
    const uint64_t k00 = keep_word(g, keep, wx, y, z), k10 = keep_word(g, keep, wx, y - 1, z);
    const uint64_t k01 = keep_word(g, keep, wx, y, z - 1), k11 = keep_word(g, keep, wx, y - 1, z - 1);
    const uint64_t p00 = keep_word(g, keep, wx - 1, y, z), p01 = keep_word(g, keep, wx - 1, y, z - 1);
    const uint64_t p10 = keep_word(g, keep, wx - 1, y - 1, z);
    m.cx &= k00 | k10 | k01 | k11;
    const uint64_t ky = k00 | k01, kz = k00 | k10;
    m.cy &= ky | (ky << 1) | ((p00 | p01) >> 63);
    m.cz &= kz | (kz << 1) | ((p00 | p10) >> 63);
}

// c[k] bit i is corner k (k = dx + 2 dy + 4 dz, mc.go:260-301) of cube 64wx+i.
struct CubeCorners {
    uint64_t c[8];
};
__device__ __forceinline__ CubeCorners nbhd_corners(const Nbhd &h) {
    CubeCorners cc;
    for (int r = 0; r < 4; r++) {  // r = dy + 2 dz
        const uint64_t b = h.n[2 * r], bn = h.n[2 * r + 1];
        cc.c[2 * r] = b;
        cc.c[2 * r + 1] = (b >> 1) | (bn << 63);
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

// Brick-major order of a one-solid culled call's words (all words of brick
// 0 row by row, then brick 1, ...): the scan numbers vertices in this order,
// so the vertices of a brick come out together and the bisection's waves
// mostly walk one brick program.  Position i holds word (wx, y, z), or none
// for the rows of bricks past the grid.  BY and BZ are powers of two (the
// host uses this order only then).
__device__ __forceinline__ uint32_t bm_index(const Grid &g, const Bricks &B, int wx, int y, int z) {
    const int sy = __ffs(B.BY) - 1, sz = __ffs(B.BZ) - 1;
    const int by = y >> sy, jy = y & (B.BY - 1), bz = z >> sz, jz = z & (B.BZ - 1);
    return ((((uint32_t)(bz * B.nby + by) * (uint32_t)g.W + (uint32_t)wx) << (sy + sz)) |
            (uint32_t)((jz << sy) | jy));
}
__device__ __forceinline__ bool bm_word(const Grid &g, const Bricks &B, uint32_t i, int &wx, int &y, int &z) {
    const int sy = __ffs(B.BY) - 1, sz = __ffs(B.BZ) - 1;
    const uint32_t b = i >> (sy + sz), r = i & ((1u << (sy + sz)) - 1);
    const int jz = (int)(r >> sy), jy = (int)(r & (uint32_t)(B.BY - 1));
    const uint32_t t = b / (uint32_t)g.W;
    wx = (int)(b - t * (uint32_t)g.W);
    const int bz = (int)(t / (uint32_t)B.nby), by = (int)t - bz * B.nby;
    y = by * B.BY + jy;
    z = bz * B.BZ + jz;
    return y < g.ny && z < g.nz;
}

// counts[w] = crossed edges | triangles << 8 | (1 if the word emits anything)
// << 17 (at most 192 edges and 320 triangles per word); a word per thread,
// and per block (CLASSIFY_BLOCK words) the block's sums {edges, triangles,
// emitting words, 0} for the scan: part_sums[blockIdx.x].  With `bm` (one
// solid, culled) thread i takes the word at brick-major position i and
// counts[i] is that word's (nwords_all: the positions).
template <bool MULTI>
__global__ __launch_bounds__(CLASSIFY_BLOCK) void k_classify(const Seg *__restrict__ segs, uint32_t nsegs, const Seg seg0,
                                                         const uint32_t *__restrict__ word_seg,
                                                         const uint64_t *__restrict__ bits_all,
                                                         const uint8_t *__restrict__ table, uint32_t *__restrict__ counts,
                                                         uint4 *__restrict__ part_sums,
                                                         const uint64_t *__restrict__ keep_all, uint64_t nwords_all,
                                                         bool bm) {
    __shared__ uint8_t ntri[256];
    __shared__ uint4 sh[CLASSIFY_BLOCK / 64];
    if (threadIdx.x < 256) ntri[threadIdx.x] = table[16 * threadIdx.x];
    __syncthreads();
    uint4 c = make_uint4(0, 0, 0, 0);
    // consecutive words per wave (segment ranges are aligned to 64)
    const uint64_t gw = (uint64_t)blockIdx.x * CLASSIFY_BLOCK + threadIdx.x;
    if (gw < nwords_all) {  // wave-uniform: nwords_all is a multiple of 64
        const uint32_t si = wave_seg<MULTI>(word_seg, gw, SEG_WORD_SHIFT);
        const Seg sg = pick_seg<MULTI>(segs, seg0, si);
        const Grid &g = sg.g;
        const uint64_t *bits = bits_all + sg.word0;
        uint64_t w = gw - sg.word0;
        int wx = 0, y = 0, z = 0;
        bool has;
        if (MULTI || !bm) {
            has = w < sg.nwords;  // else alignment padding
            if (has) word_coords(g, w, wx, y, z);
        } else {
            has = bm_word(g, sg.B, (uint32_t)gw, wx, y, z);
            w = word_index(g, wx, y, z);
        }
        if (has) {
            const Nbhd h = load_nbhd(g, bits, wx, y, z);
            Masks m = nbhd_masks(h);
            const CubeCorners cc = nbhd_corners(h);
            uint64_t mixed = mixed_cubes(cc);
            if (keep_all) {  // a block filter: only the cubes of accepted blocks
                const uint64_t *keep = keep_all + sg.word0;
                keep_masks(g, keep, wx, y, z, m);
                mixed &= keep[w];
            }
            const uint32_t nv = __popcll(m.cx) + __popcll(m.cy) + __popcll(m.cz);
            uint32_t nt = 0;
            while (mixed) {
                const int i = __ffsll((unsigned long long)mixed) - 1;
                mixed &= mixed - 1;
                nt += ntri[cube_case(cc, i)];
            }
            c = make_uint4(nv, nt, (nv | nt) != 0 ? 1u : 0u, 0u);
        }
        counts[gw] = c.x | (c.y << 8) | (c.z << 17);
    }
    const uint4 t = block_sum<CLASSIFY_BLOCK>(c, sh);
    if (threadIdx.x == 0) part_sums[blockIdx.x] = t;
}

// One block: exclusive scan of the classify blocks' sums in place, SCAN_ITEMS
// per thread; totals in 64 bits.
__global__ __launch_bounds__(SUMS_BLOCK) void k_scan_tiles(uint4 *__restrict__ part_sums, uint64_t nparts,
                                                           unsigned long long *__restrict__ totals) {
    __shared__ uint4 sh[SUMS_BLOCK / 64];
    uint4 carry = make_uint4(0, 0, 0, 0);
    unsigned long long t0 = 0, t1 = 0, t2 = 0;
    for (uint64_t start = 0; start < nparts; start += (uint64_t)SUMS_BLOCK * SCAN_ITEMS) {
        const uint64_t base = start + (uint64_t)threadIdx.x * SCAN_ITEMS;
        uint4 vals[SCAN_ITEMS];
        uint4 s = make_uint4(0, 0, 0, 0);
        for (int i = 0; i < SCAN_ITEMS; i++) {
            vals[i] = base + i < nparts ? part_sums[base + i] : make_uint4(0, 0, 0, 0);
            s = add4(s, vals[i]);
        }
        uint4 total;
        uint4 ex = add4(block_exclusive<SUMS_BLOCK>(s, sh, total), carry);
        for (int i = 0; i < SCAN_ITEMS; i++) {
            if (base + i < nparts) part_sums[base + i] = ex;
            ex = add4(ex, vals[i]);
        }
        carry = add4(carry, total);
        t0 += total.x, t1 += total.y, t2 += total.z;
    }
    if (threadIdx.x == 0) totals[0] = t0, totals[1] = t1, totals[2] = t2;
}

// Downsweep over tiles of SCAN_TILE words (the exclusive scan of the
// classify blocks' sums gives each tile's offset); words with .z set are also
// listed at their scanned .z.
__global__ __launch_bounds__(SCAN_BLOCK) void k_scan_down(const uint32_t *__restrict__ in, uint64_t n,
                                                          const uint4 *__restrict__ part_sums,
                                                          uint2 *__restrict__ out, uint32_t *__restrict__ list,
                                                          const Seg seg0, bool bm) {
    __shared__ uint4 sh[SCAN_BLOCK / 64];
    const uint64_t base = (uint64_t)blockIdx.x * SCAN_TILE + (uint64_t)threadIdx.x * SCAN_ITEMS;
    uint4 vals[SCAN_ITEMS];
    uint4 s = make_uint4(0, 0, 0, 0);
    for (int i = 0; i < SCAN_ITEMS; i++) {
        const uint32_t c = base + i < n ? in[base + i] : 0u;
        vals[i] = make_uint4(c & 0xffu, (c >> 8) & 0x1ffu, c >> 17, 0u);
        s = add4(s, vals[i]);
    }
    uint4 total;
    uint4 ex = add4(block_exclusive(s, sh, total), part_sums[(uint64_t)blockIdx.x * (SCAN_TILE / CLASSIFY_BLOCK)]);
    for (int i = 0; i < SCAN_ITEMS; i++) {
        if (base + i < n) {
            out[base + i] = make_uint2(ex.x, ex.y);
            if (list && vals[i].z) {
                uint32_t w = (uint32_t)(base + i);
                if (bm) {  // brick-major position -> word
                    int wx, y, z;
                    bm_word(seg0.g, seg0.B, w, wx, y, z);
                    w = (uint32_t)word_index(seg0.g, wx, y, z);
                }
                list[ex.z] = w;
            }
        }
        ex = add4(ex, vals[i]);
    }
}

// ---------------------------------------------------------------------------
// emit
// ---------------------------------------------------------------------------
__device__ __forceinline__ uint64_t vkey(int axis, int x, int y, int z) {
    return (uint64_t)axis | ((uint64_t)x << 2) | ((uint64_t)y << 22) | ((uint64_t)z << 42);
}

// Occupancy word, crossing masks and vertex base of one (row, word).  The
// vertices of a word are numbered from its scanned base: its x crossings,
// then y, then z, each in bit order.
struct WordInfo {
    uint64_t bits, cx, cy, cz;
    uint32_t base, tbase;  // scanned first vertex and first triangle of the word
};

// One wave per listed word, one lane per x position.  The 2 x 2 rows by 2
// words a cube of this word can touch are staged in LDS, so every lane can
// number the vertex on any edge of its cube.  Writes at or beyond the
// capacities are dropped and flagged (the host grows the buffers and runs
// this stage again).
template <bool MULTI>
__global__ __launch_bounds__(256) void k_emit(const Seg *__restrict__ segs, uint32_t nsegs, const Seg seg0,
                                              const uint32_t *__restrict__ word_seg,
                                              const uint64_t *__restrict__ bits_all,
                                              const uint8_t *__restrict__ table, const uint2 *__restrict__ offs_all,
                                              const uint32_t *__restrict__ active, Ctl *__restrict__ ctl,
                                              uint64_t *__restrict__ vkeys, uint32_t *__restrict__ vseg,
                                              uint32_t *__restrict__ tris, const uint64_t *__restrict__ keep_all,
                                              uint64_t vcap, uint64_t tcap, bool bm) {
    __shared__ __align__(16) uint8_t tab[256 * 16];
    __shared__ WordInfo wi_all[4][8];
    __shared__ uint32_t s_pre[4][64];
    __shared__ uint8_t s_cs[4][64];
    // (table: the case table with each edge decoded to axis | dx << 2 |
    // dy << 3 | dz << 4, its offset from the cube's low corner)
    for (int i = threadIdx.x; i < 256; i += blockDim.x)  // 16 bytes a thread
        reinterpret_cast<uint4 *>(tab)[i] = reinterpret_cast<const uint4 *>(table)[i];
    __syncthreads();
    const int lane = threadIdx.x & 63, wave = wave_id();
    WordInfo *wi = wi_all[wave];
    const uint32_t nact = (uint32_t)ctl->totals[2];
    const uint64_t below = lane ? (~0ull >> (64 - lane)) : 0ull;
    bool over = false;
    const uint32_t kstep = gridDim.x * 4;
    uint32_t k = blockIdx.x * 4 + wave;
    uint64_t gw_next = k < nact ? active[k] : 0;
    for (; k < nact; k += kstep) {
        const uint64_t gw = gw_next;
        if (k + kstep < nact) gw_next = active[k + kstep];  // the next word, while this one is processed
        const uint32_t si = wave_seg<MULTI>(word_seg, gw, SEG_WORD_SHIFT);
        const Seg sg = pick_seg<MULTI>(segs, seg0, si);
        const Grid &g = sg.g;
        const uint64_t *bits = bits_all + sg.word0;
        const uint2 *offs = offs_all + sg.word0;
        const uint32_t vbase = MULTI ? offs[0].x : 0u;  // the solid's first vertex (triangles index from it)
        const uint64_t w = gw - sg.word0;
        int wx, y, z;
        word_coords(g, w, wx, y, z);
        wave_sync();
        if (lane < 8) {  // entry bits: 0 next word, 1 row y+1, 2 row z+1
            const int ex = wx + (lane & 1), ey = y + ((lane >> 1) & 1), ez = z + (lane >> 2);
            // loads from clamped addresses (no branches), zeroed where the
            // word or its neighbour lies beyond the grid
            const bool in = ex < g.W && ey < g.ny && ez < g.nz;
            const int cx = min(ex, g.W - 1), cy = min(ey, g.ny - 1), cz = min(ez, g.nz - 1);
            // a segment's word indices fit in 32 bits (checked on the host)
            const uint32_t W = (uint32_t)g.W, plane = (uint32_t)g.ny * W;
            const uint32_t idx = (uint32_t)cz * plane + (uint32_t)cy * W + (uint32_t)cx;
            const bool hx = cx + 1 < g.W, hy = cy + 1 < g.ny, hz = cz + 1 < g.nz;
            const uint64_t b = bits[idx];
            const uint64_t bn = bits[idx + (hx ? 1u : 0u)], by = bits[idx + (hy ? W : 0u)];
            const uint64_t bz = bits[idx + (hz ? plane : 0u)];
            const uint2 o = offs[bm && !MULTI ? bm_index(g, sg.B, cx, cy, cz) : idx];
            WordInfo r{0, 0, 0, 0, 0, 0};
            if (in) {
                r.bits = b;
                Masks m{b ^ ((b >> 1) | ((hx ? bn : 0) << 63)), b ^ (hy ? by : 0), b ^ (hz ? bz : 0)};
                if (keep_all) keep_masks(g, keep_all + sg.word0, ex, ey, ez, m);
                r.cx = m.cx, r.cy = m.cy, r.cz = m.cz;
                r.base = o.x;
                r.tbase = o.y;
            }
            wi[lane] = r;
        }
        wave_sync();
        const int x = wx * 64 + lane;
        {  // vertices of this word (vertex numbers are 32-bit, as the scan)
            const WordInfo me = wi[0];
            const uint64_t yz = vkey(0, 0, y, z);  // the row's part of the keys
            const uint32_t c0 = __popcll(me.cx), c1 = __popcll(me.cy);
            const uint32_t k0 = me.base + __popcll(me.cx & below);
            const uint32_t k1 = me.base + c0 + __popcll(me.cy & below);
            const uint32_t k2 = me.base + c0 + c1 + __popcll(me.cz & below);
            const uint32_t xk = (uint32_t)x << 2;
            if ((me.cx >> lane) & 1) {
                if (k0 < vcap) {
                    vkeys[k0] = yz | xk;
                    if (vseg) vseg[k0] = si;
                } else {
                    over = true;
                }
            }
            if ((me.cy >> lane) & 1) {
                if (k1 < vcap) {
                    vkeys[k1] = yz | (xk | 1u);
                    if (vseg) vseg[k1] = si;
                } else {
                    over = true;
                }
            }
            if ((me.cz >> lane) & 1) {
                if (k2 < vcap) {
                    vkeys[k2] = yz | (xk | 2u);
                    if (vseg) vseg[k2] = si;
                } else {
                    over = true;
                }
            }
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
        if (keep_all && !((keep_all[sg.word0 + w] >> lane) & 1)) cs = 0;  // not in an accepted block
        const int nt = tab[16 * cs];
        const uint32_t pre = __popcll(__ballot(nt & 1) & below) + 2 * __popcll(__ballot(nt & 2) & below) +
                             4 * __popcll(__ballot(nt & 4) & below);
        const uint32_t tw = __shfl(pre + nt, 63);  // triangles of the word
        const uint64_t t0 = wi[0].tbase;
        if (t0 + tw > tcap) {
            over = true;
            continue;  // wave-uniform
        }
        // Spread the word's 3*tw triangle corners over the lanes: corner j
        // belongs to triangle j/3 of the lane whose prefix covers it.
        s_pre[wave][lane] = pre;
        s_cs[wave][lane] = (uint8_t)cs;
        wave_sync();
        for (uint32_t j = lane; j < 3 * tw; j += 64) {
            const uint32_t ti = j / 3, ec = j - 3 * ti;
            int L = 0;
            for (int step = 32; step; step >>= 1)
                if (L + step < 64 && s_pre[wave][L + step] <= ti) L += step;
            const int c = s_cs[wave][L];
            const int kk = 3 * (int)(ti - s_pre[wave][L]) + (int)ec;
            const int e = tab[16 * c + 1 + kk];
            const int axis = e & 3, ddx = (e >> 2) & 1, ddy = (e >> 3) & 1, ddz = (e >> 4) & 1;
            const int ex = L + ddx;  // x within this word (64: the next word's first)
            const WordInfo &o = wi[(ex >> 6) | (ddy << 1) | (ddz << 2)];
            const int bb = ex & 63;
            const uint64_t bl = bb ? (~0ull >> (64 - bb)) : 0ull;
            // vertices of the word before this edge: its x crossings, then y, then z
            const uint64_t m = axis == 0 ? o.cx : (axis == 1 ? o.cy : o.cz);
            const uint32_t idx = o.base + (axis > 0 ? (uint32_t)__popcll(o.cx) : 0u) +
                                 (axis > 1 ? (uint32_t)__popcll(o.cy) : 0u) + (uint32_t)__popcll(m & bl);
            tris[3 * t0 + j] = idx - vbase;
        }
    }
    if (over) atomicOr(&ctl->overflow, 1u);
}

// Each solid's first vertex and triangle in the call's outputs.
__global__ __launch_bounds__(256) void k_seg_ranges(const Seg *__restrict__ segs, uint32_t nsegs,
                                                    const uint2 *__restrict__ offs, SegOut *__restrict__ sout) {
    for (uint32_t s = blockIdx.x * blockDim.x + threadIdx.x; s < nsegs; s += gridDim.x * blockDim.x) {
        const uint2 o = offs[segs[s].word0];
        sout[s].vbase = o.x;
        sout[s].tbase = o.y;
    }
}

// ---------------------------------------------------------------------------
// bisect
// ---------------------------------------------------------------------------
__device__ __forceinline__ bool corner_bit(const Grid &g, const uint64_t *__restrict__ bits, int x, int y, int z) {
    return (bits[word_index(g, x >> 6, y, z)] >> (x & 63)) & 1;
}

// mcSearchPoint (mc.go:223-258) per vertex; iters == 0 leaves the edge
// midpoints of MarchingCubes (Coord3D.Mid, coords.go:196-198).  An edge lies
// inside the brick of its low corner (bricks include their +1 corners), so
// that brick's program decides every midpoint.
template <int MODE, bool STATS>
__device__ __forceinline__ void bisect_vertex(const Seg &sg, const uint64_t *__restrict__ bits_all,
                                              const uint64_t *__restrict__ vkeys, double *__restrict__ verts,
                                              double *__restrict__ interior, uint64_t v, GStats &gs) {
    const DevProg &P = sg.P;
    const Grid &g = sg.g;
    const Bricks &B = sg.B;
    const uint64_t *bits = bits_all + sg.word0;
    const int iters = sg.iters;
    {
        const uint64_t key = vkeys[v];
        const int axis = (int)(key & 3);
        const int x = (int)((key >> 2) & 0xfffff), y = (int)((key >> 22) & 0xfffff), z = (int)((key >> 42) & 0xfffff);
        double p[3] = {g.xs[x], g.ys[y], g.zs[z]};
        double lo, hi;
        int hx = x, hy = y, hz = z;
        if (axis == 0) lo = g.xs[x], hi = g.xs[x + 1], hx++;
        else if (axis == 1) lo = g.ys[y], hi = g.ys[y + 1], hy++;
        else lo = g.zs[z], hi = g.zs[z + 1], hz++;
        double res, inside;  // the vertex; the bracket's inside end (mc.go:252-255)
        if (iters == 0) {
            res = (lo + hi) * 0.5;
            inside = interior && !corner_bit(g, bits, hx, hy, hz) ? lo : hi;
        } else {
            double fp = lo, tp = hi;
            if (!corner_bit(g, bits, hx, hy, hz)) fp = hi, tp = lo;
            // The edge's brick program, looked up once for all midpoints.
            const Insn *code = P.code;
            uint32_t ncode = P.ncode;
            int konst = -1;
            if (B.info) {
                const uint2 bi = B.info[sg.brick0 + ((size_t)(z / B.BZ) * B.nby + (y / B.BY)) * g.W + (x >> 6)];
                if (bi.y == 0) konst = 0;
                else if (bi.y == BRICK_TRUE) konst = 1;
                else code = brick_code(P, B, bi), ncode = bi.y;
            }
            // One program for every lane of the wave (no culling, or lanes
            // of one brick: vertices are numbered brick by brick): walked
            // with scalar loads.
            const uint64_t cp = (uint64_t)code;
            const uint64_t cp0 = (uint64_t)(uint32_t)__builtin_amdgcn_readfirstlane((int)(uint32_t)cp) |
                                 ((uint64_t)(uint32_t)__builtin_amdgcn_readfirstlane((int)(uint32_t)(cp >> 32)) << 32);
            const uint32_t n0 = (uint32_t)__builtin_amdgcn_readfirstlane((int)ncode);
            const int k0 = __builtin_amdgcn_readfirstlane(konst);
            if (__ballot(cp != cp0 || ncode != n0 || konst != k0) == 0 && k0 < 0) {
                const Insn *code0 = (const Insn *)cp0;
                for (int i = 0; i < iters; i++) {
                    const double mid = (fp + tp) / 2;
                    p[axis] = mid;
                    if (eval_code<MODE, STATS, true>(code0, n0, P, p[0], p[1], p[2], gs)) tp = mid;
                    else fp = mid;
                }
            } else {
                for (int i = 0; i < iters; i++) {
                    const double mid = (fp + tp) / 2;
                    p[axis] = mid;
                    // Lanes of one wave may sit in different bricks: no readfirstlane here.
                    const bool in = konst >= 0 ? konst != 0
                                               : eval_code<MODE, STATS, false>(code, ncode, P, p[0], p[1], p[2], gs);
                    if (in) tp = mid;
                    else fp = mid;
                }
            }
            res = (fp + tp) / 2;
            inside = tp;
            if (STATS) gs.flops64 += 2 * (iters + 1);
        }
        p[axis] = res;
        verts[3 * v] = p[0];
        verts[3 * v + 1] = p[1];
        verts[3 * v + 2] = p[2];
        if (interior) {
            p[axis] = inside;
            interior[3 * v] = p[0];
            interior[3 * v + 1] = p[1];
            interior[3 * v + 2] = p[2];
        }
    }
}

template <int MODE, bool STATS, bool MULTI>
__global__ __launch_bounds__(256) void k_bisect(const Seg *__restrict__ segs, uint32_t nsegs, const Seg seg0,
                                                const uint32_t *__restrict__ vseg,
                                                const uint64_t *__restrict__ bits_all,
                                                const uint64_t *__restrict__ vkeys, double *__restrict__ verts,
                                                double *__restrict__ interior, Ctl *__restrict__ ctl, uint64_t vcap) {
    GStats gs;
    const uint64_t nv = min((uint64_t)ctl->totals[0], vcap);
    for (uint64_t v = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; v < nv;
         v += (uint64_t)gridDim.x * blockDim.x) {
        // Lanes may belong to different solids: take them one solid at a time.
        const uint32_t mine = MULTI ? vseg[v] : 0u;
        for (;;) {
            const uint32_t si = (uint32_t)__builtin_amdgcn_readfirstlane((int)mine);
            if (mine != si) continue;
            bisect_vertex<MODE, STATS>(pick_seg<MULTI>(segs, seg0, si), bits_all, vkeys, verts, interior, v, gs);
            break;
        }
    }
    flush_stats<STATS>(gs, ctl->gstats, ctl->flops + 2);
}

// Solid.Contains at arbitrary points (the full program, no bricks).
template <int MODE>
__global__ __launch_bounds__(256) void k_contains(DevProg P, const double *__restrict__ pts, uint8_t *__restrict__ out,
                                                  uint64_t n) {
    GStats gs;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x)
        out[i] = eval_code<MODE, false, true, true>(P.code, P.ncode, P, pts[3 * i], pts[3 * i + 1], pts[3 * i + 2],
                                                     gs, i)
                     ? 1
                     : 0;
}

// ---------------------------------------------------------------------------
// host-evaluated atoms
// ---------------------------------------------------------------------------
// The points a host-evaluated atom is asked about, in the atom's own frame
// (the inverse transforms above it, applied exactly as for native atoms):
// points [start, start + n) of a grid's corners (pts == nullptr; corner id
// x + nx * (y + ny * z)) or of a point list.
__global__ __launch_bounds__(256) void k_host_points(DevProg P, Grid g, const double *__restrict__ pts,
                                                     uint32_t frame, uint64_t start, uint64_t n,
                                                     double *__restrict__ out) {
    const Frame fr = P.frames[frame];
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x) {
        const uint64_t id = start + i;
        double x, y, z;
        if (pts) {
            x = pts[3 * id], y = pts[3 * id + 1], z = pts[3 * id + 2];
        } else {
            const uint64_t row = id / (uint64_t)g.nx;
            x = g.xs[id - row * (uint64_t)g.nx];
            y = g.ys[row % (uint64_t)g.ny];
            z = g.zs[row / (uint64_t)g.ny];
        }
        apply_frame(P.f64, fr, x, y, z);
        out[3 * i] = x, out[3 * i + 1] = y, out[3 * i + 2] = z;
    }
}

// mcSearchPoint (mc.go:223-258) one probe at a time, for programs with
// host-evaluated atoms: the host answers those atoms for every vertex's
// probe between rounds.  state[2v..2v+1] = (false end, true end) on the
// edge's axis; probe[3v..] = the next midpoint, which is also written to
// verts (after the last round it is the result).  phase 0 sets up the state
// from the corner bits; phase 1 decides the current probes and moves on.
template <int MODE>
__global__ __launch_bounds__(256) void k_bisect_host(DevProg P, Grid g, const uint64_t *__restrict__ bits,
                                                     const uint64_t *__restrict__ vkeys, double *__restrict__ state,
                                                     double *__restrict__ probe, double *__restrict__ verts,
                                                     double *__restrict__ interior, uint64_t nv, int phase) {
    GStats gs;
    for (uint64_t v = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; v < nv;
         v += (uint64_t)gridDim.x * blockDim.x) {
        const uint64_t key = vkeys[v];
        const int axis = (int)(key & 3);
        const int x = (int)((key >> 2) & 0xfffff), y = (int)((key >> 22) & 0xfffff), z = (int)((key >> 42) & 0xfffff);
        double p[3] = {g.xs[x], g.ys[y], g.zs[z]};
        double fp, tp;
        if (phase == 0) {
            int hx = x, hy = y, hz = z;
            double lo, hi;
            if (axis == 0) lo = g.xs[x], hi = g.xs[x + 1], hx++;
            else if (axis == 1) lo = g.ys[y], hi = g.ys[y + 1], hy++;
            else lo = g.zs[z], hi = g.zs[z + 1], hz++;
            fp = lo, tp = hi;
            if (!corner_bit(g, bits, hx, hy, hz)) fp = hi, tp = lo;
        } else {
            fp = state[2 * v], tp = state[2 * v + 1];
            const double mid = (fp + tp) / 2;
            p[axis] = mid;
            if (eval_code<MODE, false, false, true>(P.code, P.ncode, P, p[0], p[1], p[2], gs, v)) tp = mid;
            else fp = mid;
        }
        state[2 * v] = fp, state[2 * v + 1] = tp;
        p[axis] = (fp + tp) / 2;
        probe[3 * v] = p[0], probe[3 * v + 1] = p[1], probe[3 * v + 2] = p[2];
        verts[3 * v] = p[0], verts[3 * v + 1] = p[1], verts[3 * v + 2] = p[2];
        if (interior) {  // the bracket's inside end after this round (mc.go:252-255)
            p[axis] = tp;
            interior[3 * v] = p[0], interior[3 * v + 1] = p[1], interior[3 * v + 2] = p[2];
        }
    }
}

// Mesh.Transform(t.Inverse()) of MarchingCubesConj (mc.go:66-71,
// mesh.go:613-642): every vertex mapped by the inverse chain, given as frame
// records {kind, v[9]} (the Apply of each op, transform.go:44-46, 71-73,
// 103-105, 129-131).
__global__ __launch_bounds__(256) void k_inverse_apply(const double *__restrict__ recs, uint32_t nops,
                                                       double *__restrict__ verts, const Ctl *__restrict__ ctl,
                                                       uint64_t vcap) {
    const uint64_t nv = min((uint64_t)ctl->totals[0], vcap);
    const Frame f{0, nops};
    for (uint64_t v = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; v < nv; v += (uint64_t)gridDim.x * blockDim.x) {
        double x = verts[3 * v], y = verts[3 * v + 1], z = verts[3 * v + 2];
        apply_frame(recs, f, x, y, z);
        verts[3 * v] = x, verts[3 * v + 1] = y, verts[3 * v + 2] = z;
    }
}

// Binary STL body (export.go:37-66, fileformats/stl.go:41-58): per triangle
// the float32 cast of Triangle.Normal() -- the cross product of (t1 - t0) and
// (t2 - t0), scaled by 1 / Norm() in float64 (primitives.go:27-33,
// coords.go:201-222, 379-381) -- then the float32 casts of the three
// vertices and two zero bytes: 50 bytes, written as 25 little-endian
// halfwords (a record starts at an even offset).
__global__ __launch_bounds__(256) void k_encode_stl(const double *__restrict__ verts, const uint32_t *__restrict__ tris,
                                                    uint64_t nt, uint16_t *__restrict__ body) {
    for (uint64_t t = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; t < nt; t += (uint64_t)gridDim.x * blockDim.x) {
        double p[3][3];
        for (int k = 0; k < 3; k++) {
            const uint64_t v = tris[3 * t + k];
            p[k][0] = verts[3 * v], p[k][1] = verts[3 * v + 1], p[k][2] = verts[3 * v + 2];
        }
        const double ax = p[1][0] - p[0][0], ay = p[1][1] - p[0][1], az = p[1][2] - p[0][2];
        const double bx = p[2][0] - p[0][0], by = p[2][1] - p[0][1], bz = p[2][2] - p[0][2];
        const double cx = ay * bz - az * by, cy = az * bx - ax * bz, cz = ax * by - ay * bx;
        const double s = 1 / sqrt(cx * cx + cy * cy + cz * cz);
        float rec[12] = {(float)(cx * s), (float)(cy * s), (float)(cz * s)};
        for (int k = 0; k < 3; k++)
            for (int d = 0; d < 3; d++) rec[3 + 3 * k + d] = (float)p[k][d];
        uint16_t *o = body + 25 * t;
        for (int w = 0; w < 12; w++) {
            const uint32_t u = __float_as_uint(rec[w]);
            o[2 * w] = (uint16_t)(u & 0xffffu);
            o[2 * w + 1] = (uint16_t)(u >> 16);
        }
        o[24] = 0;
    }
}

}  // namespace mc3d
