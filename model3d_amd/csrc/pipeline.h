// pipeline.h -- the data structures the host and the device stages share:
// the control block, a solid's grid, bricks and segment, and the block
// sizes the host needs to size its buffers (kernels.h holds the kernels).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "eval.h"

namespace mc3d {

constexpr uint32_t BRICK_TRUE = 0xffffffffu;  // info.y of a brick where the solid is true
constexpr uint32_t BRICK_FULL = 0xffffffffu;  // info.x of a brick that uses the unfolded program

// Per-call control block in device memory (zeroed by the upload).
struct Ctl {
    uint32_t inside;       // panic("solid is true outside of bounds") seen
    uint32_t overflow;     // an output index exceeded the buffers' capacity
    uint32_t nact_bricks;  // bricks whose folded program is not constant
    uint32_t filter_overflow;  // the block filter's recursion outgrew a level's block list
    unsigned long long totals[4];  // scan totals: vertices, triangles, listed words
    unsigned long long gstats[4];  // guard statistics (GStats order)
    unsigned long long flops[4];   // statistics runs: float64, float32 ops of corners; the same of bisect
    unsigned long long code_total; // instructions over all active bricks' folded programs
    uint32_t code_max;             // longest folded program
    uint32_t full_bricks;          // bricks whose folded program overflowed its slot
    uint32_t epoch;                // this call's tag for the classify blocks' sums (set by the host)
    uint32_t scan_ticket;          // classify blocks in the order they started
    uint32_t scan_error;           // a classify block gave up waiting for the blocks before it
    uint32_t pad;
};

// A classify block's sums for the one-pass scan of small grids: two words,
// w[0] = vertices << 32 | epoch, w[1] = (triangles << 12 | emitting words)
// << 32 | epoch (a block of CLASSIFY_BLOCK words has fewer than 2^20
// triangles and at most 2^10 emitting words).  Each word is written and read
// whole by an atomic, so a reader takes the two only when both carry this
// call's epoch.  A word of another epoch is an earlier call's (the array
// starts zeroed and epochs are never reused).
struct ScanAgg {
    unsigned long long w[2];
};
constexpr uint32_t SCAN_EPOCH_MAX = 1u << 30;

struct Grid {
    const double *__restrict__ xs;
    const double *__restrict__ ys;
    const double *__restrict__ zs;
    int nx, ny, nz, W;  // W = words per row = ceil(nx / 64)
    // a z slab of a larger grid: its first / last plane is shared with the
    // slab below / above, not a face of the grid
    int zlo_open, zhi_open;
};

// Plane z is a face of the whole grid (where panic("solid is true outside
// of bounds"), mc.go:687-689, is checked).
MC3D_HD inline bool z_face(const Grid &g, int z) {
    return (z == 0 && !g.zlo_open) || (z == g.nz - 1 && !g.zhi_open);
}

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
    uint64_t pos0;             // first scan position (brick-major calls), else word0; pos0 % 64 == 0
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

// The brick plan: threads per block and the per-thread fold stack.
constexpr int PLAN_BLOCK = 128;
constexpr int PLAN_STACK = 33;
constexpr size_t PLAN_STACK_BYTES = PLAN_BLOCK * PLAN_STACK * sizeof(uint32_t);

// classify and the scan of its counts
constexpr int SCAN_BLOCK = 256;
constexpr int SCAN_ITEMS = 8;
constexpr int SCAN_TILE = SCAN_BLOCK * SCAN_ITEMS;
constexpr int CLASSIFY_BLOCK = 1024;  // words per classify block and per block sum
constexpr int SUMS_BLOCK = 1024;      // threads of the one block scanning the block sums
constexpr int SCAN_FUSE_PARTS = 4096; // up to this many block sums, each scan tile adds its prefix itself
// From this many brick-major positions classify takes the +y / +z words
// from neighbouring lanes (the 1G-corner tree 298 -> 223 us; below it the
// shuffles cost more than the loads they save: csg64 21.7 -> 22.6 us).
constexpr uint64_t CLASSIFY_SHFL_MIN = 1ull << 22;
constexpr int SCAN_ONE_PASS = 64;     // up to this many classify blocks, classify scans its counts itself
constexpr uint32_t SCAN_SPIN_LIMIT = 1u << 22;  // rounds of waiting after which a block gives up
static_assert(SCAN_TILE % CLASSIFY_BLOCK == 0, "a scan tile starts at a classify block");

}  // namespace mc3d
