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

#include "pipeline.h"
#include "plan.h"

namespace mc3d {


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
// One thread per brick folds the postfix program (plan_fold_skip: a subtree
// whose atoms miss every brick of the wave is replaced by its constant
// without a walk, so the walk stays the same for the whole wave; a wave's
// bricks are consecutive, so its constant bricks' words are written row by
// row) and writes
// the folded program into the brick's slot of `cap` instructions.  A brick
// whose folded code ever needs more than `cap` slots uses the unfolded
// program.  A brick where the solid is constant gets its occupancy words
// here (all false, or every valid corner true, with the check that no
// boundary corner is inside: panic("solid is true outside of bounds"),
// mc.go:687-689).  With `stage`, the program's node records are first copied
// to LDS; otherwise the uniform walk reads them with scalar loads.  The per-brick statistics are reduced per wave.
// Bricks a plan wave takes at once: its first `gsz` lanes (a power of two
// <= 64) hold `gsz` consecutive bricks, or with `tile` (a call of one
// solid) a tile of one word column by ty brick rows by gsz / ty brick
// planes, ty the power of two nearest sqrt(gsz).  The fewer and closer a
// wave's bricks, the more subtrees miss all of them; with fewer bricks per
// wave more waves are in flight, which is what a latency-bound walk needs.
template <bool MULTI>
__device__ __forceinline__ uint64_t plan_units(const Seg &s0, uint64_t nbricks_all, uint32_t gsz, bool tile) {
    if (MULTI || !tile) return (nbricks_all + gsz - 1) / gsz;
    const uint32_t ty = 1u << ((31 - __clz(gsz)) / 2), tz = gsz / ty;
    return (uint64_t)s0.g.W * ((s0.B.nby + ty - 1) / ty) * ((s0.B.nbz + tz - 1) / tz);
}

template <bool MULTI>
__global__ __launch_bounds__(PLAN_BLOCK) void k_plan(const Seg *__restrict__ segs, uint32_t nsegs, const Seg seg0,
                                                     const uint32_t *__restrict__ brick_seg, uint64_t nbricks_all,
                                                     uint32_t cap, uint2 *__restrict__ info, Insn *__restrict__ out,
                                                     uint32_t *__restrict__ act_list, uint64_t *__restrict__ bits_all,
                                                     Ctl *__restrict__ ctl, SegOut *__restrict__ sout, uint32_t gsz,
                                                     bool tile) {
    extern __shared__ __align__(16) char plan_smem[];
    uint32_t *S = (uint32_t *)plan_smem + threadIdx.x * PLAN_STACK;
    uint32_t staged = ~0u;  // segment whose program is in LDS
    const int lane = threadIdx.x & 63, wave = wave_id();
    constexpr int WPB = PLAN_BLOCK / 64;
    tile = tile && !MULTI;
    const uint64_t nunits = plan_units<MULTI>(seg0, nbricks_all, gsz, tile);
    const uint64_t uend = (nunits + WPB - 1) / WPB * WPB;  // every wave of a block runs the same iterations
    for (uint64_t u0 = (uint64_t)blockIdx.x * WPB; u0 < uend; u0 += (uint64_t)gridDim.x * WPB) {
        const uint64_t u = u0 + wave;
        // this lane's brick
        uint64_t gb;
        bool valid;
        if (tile) {
            const uint32_t ty = 1u << ((31 - __clz(gsz)) / 2), tz = gsz / ty;
            const uint32_t W = (uint32_t)seg0.g.W, nty = (uint32_t)(seg0.B.nby + ty - 1) / ty;
            const uint32_t bx = (uint32_t)(u % W), t = (uint32_t)(u / W);
            const uint32_t by = (t % nty) * ty + (uint32_t)lane % ty, bz = (t / nty) * tz + (uint32_t)lane / ty;
            valid = (uint32_t)lane < gsz && u < nunits && by < (uint32_t)seg0.B.nby && bz < (uint32_t)seg0.B.nbz;
            gb = valid ? ((uint64_t)bz * seg0.B.nby + by) * W + bx : 0;
        } else {
            gb = u * gsz + lane;
            valid = (uint32_t)lane < gsz && gb < nbricks_all;
        }
        // A block's bricks share one segment (ranges aligned to 128; a block
        // takes 2 gsz <= 128 consecutive bricks).
        const uint32_t si = wave_seg<MULTI>(brick_seg, u0 * gsz, SEG_BRICK_SHIFT);
        const Seg sg = pick_seg<MULTI>(segs, seg0, si);
        const DevProg &P = sg.P;
        const Grid &g = sg.g;
        const Bricks &B = sg.B;
        const uint64_t b = gb - sg.brick0;  // local brick
        valid = valid && b < sg.nbricks;
        bool active = false, full = false;
        uint32_t len = 0;
        // the brick's box (a lane without a brick takes the first brick's
        // and keeps its output slot empty), loaded before the program is
        // staged so that the two loads overlap
        const uint64_t bl = valid ? b : 0;
        const uint32_t r = (uint32_t)bl / (uint32_t)g.W;
        const int bx = (int)((uint32_t)bl - r * (uint32_t)g.W);
        const int bz = (int)(r / (uint32_t)B.nby), by = (int)(r - (uint32_t)bz * (uint32_t)B.nby);
        const int x0 = bx * 64, y0 = by * B.BY, z0 = bz * B.BZ;
        const int x1 = min(x0 + 64, g.nx - 1), y1 = min(y0 + B.BY, g.ny - 1), z1 = min(z0 + B.BZ, g.nz - 1);
        const double lo[3] = {g.xs[x0], g.ys[y0], g.zs[z0]};
        const double hi[3] = {g.xs[x1], g.ys[y1], g.zs[z1]};
        // a staged program's node records in LDS
        PlanNode *sn = (PlanNode *)(plan_smem + PLAN_STACK_BYTES);
        if (sg.stage) {  // block-uniform
            if (staged != si) {
                __syncthreads();  // the previous segment's walks are done with the LDS copy
                const uint4 *src = (const uint4 *)P.plan;
                uint4 *dst = (uint4 *)sn;
                // copied by the LDS DMA (16 bytes a lane, 1 KiB a wave per
                // load, no registers), all loads in flight before the wait
                const uint32_t nq = P.ncode * (uint32_t)(sizeof(PlanNode) / 16);
                for (uint32_t c = (uint32_t)wave; c * 64 < nq; c += WPB)
                    if (c * 64 + (uint32_t)lane < nq)
                        __builtin_amdgcn_global_load_lds(
                            (const void __attribute__((address_space(1))) *)(src + c * 64 + lane),
                            (void __attribute__((address_space(3))) *)(dst + c * 64), 16, 0, 0);
                __builtin_amdgcn_s_waitcnt(VMCNT0);
                __syncthreads();
                staged = si;
            }
        }
        Insn *o = out + (valid ? gb * cap : 0);
        const uint32_t ocap = valid ? cap : 0u;
        uint32_t n = 0;
        bool over = false;
        // two copies of the fold, so that each reads with typed
        // instructions: LDS, or scalar loads for programs too large to stage
        if (sg.stage) {
            auto miss = [&](const PlanNode &t) { return __ballot(valid && !node_misses(t, lo, hi)) == 0; };
            plan_fold_nodes((const PlanNode *)sn, P.ncode, S, lo, hi, o, ocap, n, over, miss);
        } else {
            auto miss = [&](const PlanNode &t) { return __ballot(valid && !node_misses(t, lo, hi)) == 0; };
            plan_fold_nodes(scalar_ptr(P.plan), P.ncode, S, lo, hi, o, ocap, n, over, miss);
        }
        if (valid) {
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
                if (t && (bx == 0 || x0 + nvalid == g.nx || y0 == 0 || y0 + jy_end == g.ny || z_face(g, z0) ||
                          z_face(g, z0 + jz_end - 1)))
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
        if (inside && (x == 0 || x == g.nx - 1 || y == 0 || y == g.ny - 1 || z_face(g, z)))
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
                             (y == 0 || y == g.ny - 1 || z_face(g, z)) ? 1u : 0u};
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
        const bool zedge = z_face(g, z);
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
    // the block's first list entry loaded beside the count (the list holds
    // one entry per brick and the grid has at most that many blocks)
    const uint32_t gb0 = scalar_ptr(act)[blockIdx.x];
    const uint32_t nact = ctl->nact_bricks;
    __shared__ RowInfo row_info[WPB][64];  // each wave's rows of the current brick
    GStats gs;
    for (uint32_t k = blockIdx.x; k < nact; k += gridDim.x) {
        const uint32_t gb = k == blockIdx.x ? gb0 : scalar_ptr(act)[k];
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
                            yy[q] == 0 || yy[q] == g.ny - 1 || z_face(g, zz[q]), ctl, so);
        }
    }
    flush_stats<STATS>(gs, ctl->gstats, ctl->flops);
}

// ---------------------------------------------------------------------------
// exclusive scan of uint4 counts (block sums written by classify, a scan of
// those, downsweep)
// ---------------------------------------------------------------------------

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
    bool zin;  // a plane above: the word's z edges and cubes exist
};
__device__ __forceinline__ Nbhd load_nbhd(const Grid &g, const uint64_t *__restrict__ bits, int wx, int y, int z) {
    const bool hx = wx + 1 < g.W, hy = y + 1 < g.ny, hz = z + 1 < g.nz;
    const size_t w = word_index(g, wx, y, z);
    const size_t dx = hx ? 1 : 0, dy = hy ? (size_t)g.W : 0, dz = hz ? (size_t)g.ny * g.W : 0;
    Nbhd r;
    for (int k = 0; k < 8; k++) r.n[k] = bits[w + ((k & 1) ? dx : 0) + ((k & 2) ? dy : 0) + ((k & 4) ? dz : 0)];
    for (int k = 0; k < 8; k++)
        if (((k & 1) && !hx) || ((k & 2) && !hy) || ((k & 4) && !hz)) r.n[k] = 0;
    r.zin = hz;
    return r;
}

// load_nbhd for a one-solid call in brick-major order: a wave's 64 lanes
// hold consecutive positions, so the rows y + 1 and z + 1 of a brick column
// are lanes + 1 and + BY of the same wave.  Each lane loads its word and
// its +x neighbour; the +y / +z words come from those lanes by shuffles,
// and only lanes on a brick's last row or plane (or the wave's end) load
// them.  Every lane of the wave takes part (has: the position holds a word).
__device__ __forceinline__ uint64_t shfl64(uint64_t v, int src) {
    const uint32_t lo = (uint32_t)__shfl((int)(uint32_t)v, src), hi = (uint32_t)__shfl((int)(uint32_t)(v >> 32), src);
    return ((uint64_t)hi << 32) | lo;
}
__device__ __forceinline__ Nbhd load_nbhd_bm(const Grid &g, const Bricks &B, const uint64_t *__restrict__ bits,
                                             int wx, int y, int z, bool has, uint64_t pos) {  // pos: within the solid
    const int lane = threadIdx.x & 63;
    const bool hx = wx + 1 < g.W, hy = y + 1 < g.ny, hz = z + 1 < g.nz;
    const size_t W = (size_t)g.W, plane = (size_t)g.ny * g.W;
    const size_t w = has ? word_index(g, wx, y, z) : 0;
    // lanes + 1 / + BY hold rows y + 1 / z + 1 of this brick column
    const uint32_t jy = (uint32_t)pos & (uint32_t)(B.BY - 1);
    const uint32_t jz = (uint32_t)(pos / (uint32_t)B.BY) & (uint32_t)(B.BZ - 1);
    const bool sy = jy + 1 < (uint32_t)B.BY && lane + 1 < 64, sz = jz + 1 < (uint32_t)B.BZ && lane + B.BY < 64;
    Nbhd r;
    for (int k = 0; k < 8; k++) r.n[k] = 0;
    if (has) {
        r.n[0] = bits[w];
        if (hx) r.n[1] = bits[w + 1];
        if (hy && !sy) r.n[2] = bits[w + W];
        if (hy && hx && !sy) r.n[3] = bits[w + W + 1];
        if (hz && !sz) r.n[4] = bits[w + plane];
        if (hz && hx && !sz) r.n[5] = bits[w + plane + 1];
        if (hy && hz && !(sy && sz)) r.n[6] = bits[w + W + plane];
        if (hx && hy && hz && !(sy && sz)) r.n[7] = bits[w + W + plane + 1];
    }
    const int ly = min(lane + 1, 63), lz = min(lane + B.BY, 63), lyz = min(lane + B.BY + 1, 63);
    const uint64_t a2 = shfl64(r.n[0], ly), a3 = shfl64(r.n[1], ly), a4 = shfl64(r.n[0], lz);
    const uint64_t a5 = shfl64(r.n[1], lz), a6 = shfl64(r.n[0], lyz), a7 = shfl64(r.n[1], lyz);
    if (hy && sy) r.n[2] = a2;
    if (hy && hx && sy) r.n[3] = a3;
    if (hz && sz) r.n[4] = a4;
    if (hz && hx && sz) r.n[5] = a5;
    if (hy && hz && sy && sz) r.n[6] = a6;
    if (hx && hy && hz && sy && sz) r.n[7] = a7;
    r.zin = hz;
    return r;
}

// Crossed-edge masks of word n[0]: bit i of cx is the edge from corner
// (64wx+i, y, z) to its +x neighbour, and likewise for y and z.  Padding
// bits (x >= nx) and corners on the grid's faces are zero, so no checks
// beyond the neighbour words' existence are needed -- except on the top
// plane of a z slab, whose corners may be inside: it has no z edges.
__device__ __forceinline__ Masks nbhd_masks(const Nbhd &h) {
    Masks m;
    m.cx = h.n[0] ^ ((h.n[0] >> 1) | (h.n[1] << 63));
    m.cy = h.n[0] ^ h.n[2];
    m.cz = h.zin ? h.n[0] ^ h.n[4] : 0;
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
// The top plane of the grid (of a z slab) has no cubes: no corners.
__device__ __forceinline__ CubeCorners nbhd_corners(const Nbhd &h) {
    CubeCorners cc;
    for (int r = 0; r < 4; r++) {  // r = dy + 2 dz
        const uint64_t b = h.zin ? h.n[2 * r] : 0, bn = h.zin ? h.n[2 * r + 1] : 0;
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

// One scan position classified: {crossed edges (its vertices), triangles
// of its mixed cubes, 1 if it emits anything, 0} (at most 192 edges and 320
// triangles per word), and in `word` the position's word as the list of
// emitting words names it.  Position gw is an occupancy word, or with `bm`
// (one solid, culled) the word at brick-major position gw.
template <bool MULTI, bool SHFL = false>
__device__ __forceinline__ uint4 classify_position(const Seg *__restrict__ segs, const Seg &seg0,
                                                   const uint32_t *__restrict__ word_seg,
                                                   const uint64_t *__restrict__ bits_all, const uint8_t *ntri,
                                                   const uint64_t *__restrict__ keep_all, bool bm, uint64_t gw,
                                                   uint32_t &word) {
    // (with bm, word_seg is the table of scan positions)
    const uint32_t si = wave_seg<MULTI>(word_seg, gw, SEG_WORD_SHIFT);
    const Seg sg = pick_seg<MULTI>(segs, seg0, si);
    const Grid &g = sg.g;
    const uint64_t *bits = bits_all + sg.word0;
    uint64_t w = gw - sg.word0;
    int wx = 0, y = 0, z = 0;
    bool has;
    Nbhd h;
    if (!bm) {
        has = w < sg.nwords;  // else alignment padding
        if (!has) return make_uint4(0, 0, 0, 0);
        word_coords(g, w, wx, y, z);
        word = (uint32_t)gw;
        h = load_nbhd(g, bits, wx, y, z);
    } else {  // the solid's brick-major position gw - pos0
        has = bm_word(g, sg.B, (uint32_t)(gw - sg.pos0), wx, y, z);
        w = word_index(g, wx, y, z);
        word = (uint32_t)(sg.word0 + w);
        if (SHFL) {  // (positions are padded to whole waves)
            h = load_nbhd_bm(g, sg.B, bits, wx, y, z, has, gw - sg.pos0);
            if (!has) return make_uint4(0, 0, 0, 0);
        } else {
            if (!has) return make_uint4(0, 0, 0, 0);
            h = load_nbhd(g, bits, wx, y, z);
        }
    }
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
    return make_uint4(nv, nt, (nv | nt) != 0 ? 1u : 0u, 0u);
}

// counts[pos] = crossed edges | triangles << 8 | (1 if the word emits
// anything) << 17, a position per thread, and per block (CLASSIFY_BLOCK
// positions) the block's sums {edges, triangles, emitting words, 0} for the
// scan: part_sums[blockIdx.x] (npos: the positions).
template <bool MULTI, bool SHFL = false>
__global__ __launch_bounds__(CLASSIFY_BLOCK) void k_classify(const Seg *__restrict__ segs, uint32_t nsegs, const Seg seg0,
                                                         const uint32_t *__restrict__ word_seg,
                                                         const uint64_t *__restrict__ bits_all,
                                                         const uint8_t *__restrict__ table, uint32_t *__restrict__ counts,
                                                         uint4 *__restrict__ part_sums,
                                                         const uint64_t *__restrict__ keep_all, uint64_t npos,
                                                         bool bm) {
    __shared__ uint8_t ntri[256];
    __shared__ uint4 sh[CLASSIFY_BLOCK / 64];
    if (threadIdx.x < 256) ntri[threadIdx.x] = table[16 * threadIdx.x];
    __syncthreads();
    uint4 c = make_uint4(0, 0, 0, 0);
    // consecutive positions per wave (segment ranges are aligned to 64)
    const uint64_t gw = (uint64_t)blockIdx.x * CLASSIFY_BLOCK + threadIdx.x;
    if (gw < npos) {  // wave-uniform: npos is a multiple of 64
        uint32_t word;
        c = classify_position<MULTI, SHFL>(segs, seg0, word_seg, bits_all, ntri, keep_all, bm, gw, word);
        counts[gw] = c.x | (c.y << 8) | (c.z << 17);
    }
    const uint4 t = block_sum<CLASSIFY_BLOCK>(c, sh);
    if (threadIdx.x == 0) part_sums[blockIdx.x] = t;
}

// Classify and the scan of its counts in one pass, for at most
// SCAN_ONE_PASS blocks: a block takes tile t in the order blocks start (a
// ticket), classifies its CLASSIFY_BLOCK positions, scans them, publishes
// its sums in agg[t], waits until every tile before it has published (one
// lane per tile), adds theirs up and writes offs[pos] = {first vertex,
// first triangle} and the list of emitting words itself; the last tile
// writes the totals.  A block that has waited SCAN_SPIN_LIMIT rounds gives
// up (scan_error) rather than spin on; every tile before it has a running
// or finished block (tickets follow the start order), so the wait ends.
template <bool MULTI>
__global__ __launch_bounds__(CLASSIFY_BLOCK) void k_classify_scan(const Seg *__restrict__ segs, uint32_t nsegs,
                                                              const Seg seg0, const uint32_t *__restrict__ word_seg,
                                                              const uint64_t *__restrict__ bits_all,
                                                              const uint8_t *__restrict__ table,
                                                              const uint64_t *__restrict__ keep_all, uint64_t npos,
                                                              bool bm, ScanAgg *__restrict__ agg,
                                                              uint2 *__restrict__ offs, uint32_t *__restrict__ list,
                                                              Ctl *__restrict__ ctl) {
    __shared__ uint8_t ntri[256];
    __shared__ uint4 sh[CLASSIFY_BLOCK / 64];
    __shared__ uint32_t s_tile;
    __shared__ uint4 s_pre;
    if (threadIdx.x < 256) ntri[threadIdx.x] = table[16 * threadIdx.x];
    if (threadIdx.x == 0) s_tile = atomicAdd(&ctl->scan_ticket, 1u);
    __syncthreads();
    const uint32_t t = s_tile;
    const unsigned long long epoch = ctl->epoch;
    uint4 c = make_uint4(0, 0, 0, 0);
    uint32_t word = 0;
    const uint64_t gw = (uint64_t)t * CLASSIFY_BLOCK + threadIdx.x;
    if (gw < npos) c = classify_position<MULTI>(segs, seg0, word_seg, bits_all, ntri, keep_all, bm, gw, word);
    uint4 total;
    const uint4 ex = block_exclusive<CLASSIFY_BLOCK>(c, sh, total);
    if (threadIdx.x < 64) {
        const int lane = threadIdx.x;
        if (lane == 0) {  // this tile's sums (atomics: seen by every block at once)
            __hip_atomic_exchange(&agg[t].w[0], ((unsigned long long)total.x << 32) | epoch, __ATOMIC_RELAXED,
                                  __HIP_MEMORY_SCOPE_AGENT);
            __hip_atomic_exchange(&agg[t].w[1], ((unsigned long long)((total.y << 12) | total.z) << 32) | epoch,
                                  __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        }
        uint4 pre = make_uint4(0, 0, 0, 0);
        for (uint32_t spins = 0;; spins++) {
            uint4 v = make_uint4(0, 0, 0, 0);
            bool ready = true;
            if ((uint32_t)lane < t) {
                const unsigned long long a =
                    __hip_atomic_fetch_or(&agg[lane].w[0], 0ull, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                const unsigned long long b =
                    __hip_atomic_fetch_or(&agg[lane].w[1], 0ull, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                ready = (a & 0xffffffffull) == epoch && (b & 0xffffffffull) == epoch;
                const uint32_t tl = (uint32_t)(b >> 32);
                v = make_uint4((uint32_t)(a >> 32), tl >> 12, tl & 0xfffu, 0u);
            }
            if (!__ballot(!ready)) {
                for (int d = 32; d; d >>= 1)
                    v = add4(v, make_uint4(__shfl_xor(v.x, d), __shfl_xor(v.y, d), __shfl_xor(v.z, d), 0u));
                pre = v;
                break;
            }
            if (spins >= SCAN_SPIN_LIMIT) {
                if (lane == 0) atomicOr(&ctl->scan_error, 1u);
                break;
            }
            __builtin_amdgcn_s_sleep(2);
        }
        if (lane == 0) {
            s_pre = pre;
            if (t == gridDim.x - 1)
                ctl->totals[0] = pre.x + total.x, ctl->totals[1] = pre.y + total.y, ctl->totals[2] = pre.z + total.z;
        }
    }
    __syncthreads();
    if (gw < npos) {
        offs[gw] = make_uint2(s_pre.x + ex.x, s_pre.y + ex.y);
        if (c.z) list[s_pre.z + ex.z] = word;
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

// Listed words a wave takes per iteration of emit.
constexpr int EMIT_G = 8;
static_assert(3 * EMIT_G <= 64 && EMIT_G <= 32, "emit's per-group prefixes fit one wave");

// A word of a wave's group, as emit's later steps need it (LDS).
struct GroupWord {
    uint32_t y, z, x0;  // row, and the x of the word's first corner
    uint32_t si;        // segment
    uint32_t vsolid;    // the solid's first vertex (its triangles index from there)
    uint32_t zin;       // a plane above the word's: it has cubes
    uint64_t keep;      // block filter: the word's cubes of accepted blocks
};

// Position of the r-th set bit of m (r < popcount(m)).
__device__ __forceinline__ int select_bit(uint64_t m, int r) {
    int pos = 0;
    uint32_t h = (uint32_t)m;
    const int c = __popc(h);
    if (r >= c) r -= c, h = (uint32_t)(m >> 32), pos = 32;
    for (int w = 16; w; w >>= 1) {
        const int cw = __popc(h & ((1u << w) - 1u));
        if (r >= cw) r -= cw, h >>= w, pos += w;
    }
    return pos;
}

// Largest i < n with pre[i] <= j (pre ascending from pre[0] == 0; n <= 64).
__device__ __forceinline__ int owner(const uint32_t *pre, int n, uint32_t j) {
    int L = 0;
    for (int step = 32; step; step >>= 1)
        if (L + step < n && pre[L + step] <= j) L += step;
    return L;
}

// Emit: a wave takes EMIT_G consecutive listed words at a time.  Their
// vertices and triangles are consecutive in the outputs (the scan numbers
// them in list order), so the group's work is flattened: the 2 x 2 rows by
// 2 words around each word are staged in LDS (a lane per word and
// neighbour), then the group's vertices are spread over the lanes (vertex j
// of the group: its word, axis and crossing found from per-word prefix
// counts), then its mixed cubes (a lane per cube, its case from the staged
// bits), then the corners of its triangles (corner j belongs to triangle
// j / 3 of the cube whose prefix covers it).  No lane works on a cube
// without triangles.  Writes at or beyond the capacities are dropped and
// flagged (the host grows the buffers and runs this stage again).
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
    __shared__ WordInfo wi_all[4][EMIT_G][8];
    __shared__ GroupWord gw_all[4][EMIT_G];
    __shared__ uint64_t mixed_all[4][EMIT_G];
    __shared__ uint32_t vpre_all[4][3 * EMIT_G];
    __shared__ uint32_t mpre_all[4][EMIT_G];
    __shared__ uint32_t s_pre[4][64];
    __shared__ uint16_t s_cube[4][64];
    __shared__ uint8_t s_cs[4][64];
    // the listed words' count, loaded beside the table
    const uint32_t nact = (uint32_t)ctl->totals[2];
    // (table: the case table with each edge decoded to axis | dx << 2 |
    // dy << 3 | dz << 4, its offset from the cube's low corner)
    for (int i = threadIdx.x; i < 256; i += blockDim.x)  // 16 bytes a thread
        reinterpret_cast<uint4 *>(tab)[i] = reinterpret_cast<const uint4 *>(table)[i];
    __syncthreads();
    const int lane = threadIdx.x & 63, wave = wave_id();
    WordInfo(*wi)[8] = wi_all[wave];
    GroupWord *gwd = gw_all[wave];
    uint64_t *mixed = mixed_all[wave];
    uint32_t *vpre = vpre_all[wave], *mpre = mpre_all[wave];
    const uint64_t below = lane ? (~0ull >> (64 - lane)) : 0ull;
    bool over = false;
    const uint32_t kstep = gridDim.x * 4 * EMIT_G;
    for (uint32_t k0 = (blockIdx.x * 4 + wave) * EMIT_G; k0 < nact; k0 += kstep) {
        const int ng = (int)min((uint32_t)EMIT_G, nact - k0);  // words in the group
        wave_sync();  // the previous group's LDS reads are done
        // 1. Item 8 g + e stages neighbour e of word g: entry bits 0 next
        //    word, 1 row y+1, 2 row z+1.
        for (int q = lane; q < 8 * EMIT_G; q += 64) {
            const int g = q >> 3, e = q & 7;
            WordInfo r{0, 0, 0, 0, 0, 0};
            if (g < ng) {
                const uint32_t gw = active[k0 + g];
                const uint32_t si = MULTI ? word_seg[gw >> SEG_WORD_SHIFT] : 0u;
                Grid g_;
                Bricks B_{};
                uint64_t word0, pos0;
                if (MULTI) {
                    const Seg *sp = &segs[si];
                    g_.nx = sp->g.nx, g_.ny = sp->g.ny, g_.nz = sp->g.nz, g_.W = sp->g.W;
                    if (bm) B_.BY = sp->B.BY, B_.BZ = sp->B.BZ, B_.nby = sp->B.nby;
                    word0 = sp->word0, pos0 = sp->pos0;
                } else {
                    g_ = seg0.g;
                    B_ = seg0.B;
                    word0 = seg0.word0, pos0 = seg0.pos0;
                }
                const uint64_t *bits = bits_all + word0;
                const uint2 *offs = offs_all + pos0;  // the solid's scan positions
                int wx, y, z;
                word_coords(g_, gw - word0, wx, y, z);
                const int ex = wx + (e & 1), ey = y + ((e >> 1) & 1), ez = z + (e >> 2);
                // loads from clamped addresses (no branches), zeroed where
                // the word or its neighbour lies beyond the grid
                const bool in = ex < g_.W && ey < g_.ny && ez < g_.nz;
                const int cx = min(ex, g_.W - 1), cy = min(ey, g_.ny - 1), cz = min(ez, g_.nz - 1);
                // a segment's word indices fit in 32 bits (checked on the host)
                const uint32_t W = (uint32_t)g_.W, plane = (uint32_t)g_.ny * W;
                const uint32_t idx = (uint32_t)cz * plane + (uint32_t)cy * W + (uint32_t)cx;
                const bool hx = cx + 1 < g_.W, hy = cy + 1 < g_.ny, hz = cz + 1 < g_.nz;
                const uint64_t b = bits[idx];
                const uint64_t bn = bits[idx + (hx ? 1u : 0u)], by = bits[idx + (hy ? W : 0u)];
                const uint64_t bz = bits[idx + (hz ? plane : 0u)];
                const uint2 o = offs[bm ? bm_index(g_, B_, cx, cy, cz) : idx];
                if (in) {
                    r.bits = b;
                    // (the top plane, also of a z slab: no z edges)
                    Masks m{b ^ ((b >> 1) | ((hx ? bn : 0) << 63)), b ^ (hy ? by : 0), hz ? b ^ bz : 0};
                    if (!MULTI && keep_all) keep_masks(g_, keep_all + word0, ex, ey, ez, m);
                    r.cx = m.cx, r.cy = m.cy, r.cz = m.cz;
                    r.base = o.x;
                    r.tbase = o.y;
                }
                if (e == 0) {
                    const uint32_t vsolid = MULTI ? offs[0].x : 0u;
                    const uint64_t kw = (!MULTI && keep_all) ? keep_all[word0 + idx] : ~0ull;
                    gwd[g] = GroupWord{(uint32_t)y, (uint32_t)z, (uint32_t)wx * 64u, si, vsolid,
                                       z + 1 < g_.nz ? 1u : 0u, kw};
                }
            }
            wi[g][e] = r;
        }
        wave_sync();
        const uint32_t vbase = wi[0][0].base, tbase = wi[0][0].tbase;  // the group's first vertex / triangle
        // 2. Vertex keys.  Lane 3 g + a counts word g's crossings on axis a;
        //    their exclusive prefix numbers the group's vertices.
        uint32_t nvg;
        {
            uint32_t c = 0;
            if (lane < 3 * EMIT_G && lane / 3 < ng) {
                const WordInfo &me = wi[lane / 3][0];
                const int a = lane % 3;
                c = (uint32_t)__popcll(a == 0 ? me.cx : (a == 1 ? me.cy : me.cz));
            }
            uint32_t inc = c;
            for (int d = 1; d < 3 * EMIT_G; d <<= 1) {
                const uint32_t t = __shfl_up(inc, d);
                if (lane >= d) inc += t;
            }
            if (lane < 3 * EMIT_G) vpre[lane] = inc - c;
            nvg = __shfl(inc, 3 * EMIT_G - 1);
        }
        wave_sync();
        for (uint32_t j = lane; j < nvg; j += 64) {
            const int q = owner(vpre, 3 * EMIT_G, j), g = q / 3, a = q - 3 * g;
            const WordInfo &me = wi[g][0];
            const uint64_t m = a == 0 ? me.cx : (a == 1 ? me.cy : me.cz);
            const int bit = select_bit(m, (int)(j - vpre[q]));
            const GroupWord gw = gwd[g];
            const uint64_t v = (uint64_t)vbase + j;
            if (v < vcap) {
                vkeys[v] = vkey(a, (int)gw.x0 + bit, (int)gw.y, (int)gw.z);
                if (MULTI) vseg[v] = gw.si;
            } else {
                over = true;
            }
        }
        // 3. Each word's mixed cubes (lane g); their prefix numbers the
        //    group's mixed cubes.
        uint32_t nmg;
        {
            uint32_t c = 0;
            if (lane < ng) {
                Nbhd h;
                for (int k = 0; k < 8; k++) h.n[k] = wi[lane][k].bits;
                h.zin = gwd[lane].zin != 0;
                const uint64_t m = mixed_cubes(nbhd_corners(h)) & gwd[lane].keep;
                mixed[lane] = m;
                c = (uint32_t)__popcll(m);
            }
            uint32_t inc = c;
            for (int d = 1; d < EMIT_G; d <<= 1) {
                const uint32_t t = __shfl_up(inc, d);
                if (lane >= d) inc += t;
            }
            if (lane < EMIT_G) mpre[lane] = inc - c;
            nmg = __shfl(inc, EMIT_G - 1);
        }
        wave_sync();
        // 4. Triangles, up to 64 mixed cubes at a time.
        uint32_t tdone = 0;
        for (uint32_t c0 = 0; c0 < nmg; c0 += 64) {
            const uint32_t j = c0 + lane;
            int cs = 0, g = 0, i = 0;
            if (j < nmg) {
                g = owner(mpre, ng, j);
                i = select_bit(mixed[g], (int)(j - mpre[g]));
                for (int dz = 0; dz < 2; dz++)
                    for (int dy = 0; dy < 2; dy++) {
                        const int e = 2 * dy + 4 * dz;
                        const uint64_t b = wi[g][(dy << 1) | (dz << 2)].bits, bn = wi[g][1 | (dy << 1) | (dz << 2)].bits;
                        cs |= (int)((b >> i) & 1) << e;
                        cs |= (int)((((b >> 1) | (bn << 63)) >> i) & 1) << (e + 1);
                    }
            }
            const int nt = tab[16 * cs];
            const uint32_t pre = __popcll(__ballot(nt & 1) & below) + 2 * __popcll(__ballot(nt & 2) & below) +
                                 4 * __popcll(__ballot(nt & 4) & below);
            const uint32_t tw = __shfl(pre + nt, 63);  // triangles of these cubes
            s_pre[wave][lane] = pre;
            s_cs[wave][lane] = (uint8_t)cs;
            s_cube[wave][lane] = (uint16_t)((g << 6) | i);
            wave_sync();
            const uint64_t t0 = (uint64_t)tbase + tdone;
            if (t0 + tw > tcap) {
                over = true;
            } else {
                // triangle ti belongs to the cube whose prefix covers it (a
                // lane per triangle: one search for its three corners)
                for (uint32_t ti = lane; ti < tw; ti += 64) {
                    const int L = owner(s_pre[wave], 64, ti);
                    const int c = s_cs[wave][L], gi = s_cube[wave][L], gg = gi >> 6, ii = gi & 63;
                    const int k0 = 3 * (int)(ti - s_pre[wave][L]);
                    const uint32_t vsolid = gwd[gg].vsolid;
                    uint32_t corner[3];
                    for (int ec = 0; ec < 3; ec++) {
                        const int e = tab[16 * c + 1 + k0 + ec];
                        const int axis = e & 3, ddx = (e >> 2) & 1, ddy = (e >> 3) & 1, ddz = (e >> 4) & 1;
                        const int ex = ii + ddx;  // x within the word (64: the next word's first)
                        const WordInfo &o = wi[gg][(ex >> 6) | (ddy << 1) | (ddz << 2)];
                        const int bb = ex & 63;
                        const uint64_t bl = bb ? (~0ull >> (64 - bb)) : 0ull;
                        // vertices of the word before this edge: its x crossings, then y, then z
                        const uint64_t m = axis == 0 ? o.cx : (axis == 1 ? o.cy : o.cz);
                        const uint32_t idx = o.base + (axis > 0 ? (uint32_t)__popcll(o.cx) : 0u) +
                                             (axis > 1 ? (uint32_t)__popcll(o.cy) : 0u) + (uint32_t)__popcll(m & bl);
                        corner[ec] = idx - vsolid;
                    }
                    uint32_t *dst = tris + 3 * (t0 + ti);
                    dst[0] = corner[0], dst[1] = corner[1], dst[2] = corner[2];
                }
            }
            tdone += tw;
            wave_sync();  // s_pre / s_cs / s_cube are rewritten by the next chunk
        }
    }
    if (over) atomicOr(&ctl->overflow, 1u);
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
                                              const uint64_t key, double *__restrict__ verts,
                                              double *__restrict__ interior, uint64_t v, GStats &gs) {
    const DevProg &P = sg.P;
    const Grid &g = sg.g;
    const Bricks &B = sg.B;
    const uint64_t *bits = bits_all + sg.word0;
    const int iters = sg.iters;
    {
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
    // the first vertex's key (and solid) loaded beside the vertex count: the
    // arrays hold vcap entries, so the loads stay in bounds
    const uint64_t v0 = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const uint64_t key0 = v0 < vcap ? vkeys[v0] : 0;
    const uint32_t seg0v = MULTI && v0 < vcap ? vseg[v0] : 0u;
    const uint64_t nv = min((uint64_t)ctl->totals[0], vcap);
    for (uint64_t v = v0; v < nv; v += (uint64_t)gridDim.x * blockDim.x) {
        const uint64_t key = v == v0 ? key0 : vkeys[v];
        // Lanes may belong to different solids: take them one solid at a time.
        const uint32_t mine = MULTI ? (v == v0 ? seg0v : vseg[v]) : 0u;
        for (;;) {
            const uint32_t si = (uint32_t)__builtin_amdgcn_readfirstlane((int)mine);
            if (mine != si) continue;
            bisect_vertex<MODE, STATS>(pick_seg<MULTI>(segs, seg0, si), bits_all, key, verts, interior, v, gs);
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



}  // namespace mc3d
