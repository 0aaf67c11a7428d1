// launch.h -- host-side launchers of the device stages (kernels.h).  Each
// group of kernels is compiled in a translation unit of its own (the corner
// kernel once per precision and statistics mode), so the library builds in
// parallel; mc3d.hip only sees these declarations.
#pragma once

#include <hip/hip_runtime.h>

#include <algorithm>

#include "pipeline.h"

namespace mc3d {

// Device buffers of one call: the context's grow-only buffers, or pieces of
// the fixed scratch array of the allocation-free entry point.
struct Bufs {
    uint64_t *bits = nullptr;    // occupancy, one word per 64 x corners of a row
    uint32_t *counts = nullptr;  // packed per scan position
    uint2 *offs = nullptr;       // scanned {first vertex, first triangle}
    uint32_t *actw = nullptr;    // listed (emitting) words
    uint4 *tiles = nullptr;      // classify block sums
    ScanAgg *agg = nullptr;      // classify block sums of the one-pass scan (SCAN_ONE_PASS entries)
    uint2 *binfo = nullptr;      // per brick {code offset, length}
    uint32_t *actb = nullptr;    // active bricks
    Insn *bcode = nullptr;       // folded brick programs
    uint64_t *vkeys = nullptr;   // owning edge of each vertex
    uint32_t *vseg = nullptr;    // segment of each vertex (calls of several solids)
    const uint8_t *table = nullptr;   // 256-case table
    const uint8_t *etable = nullptr;  // the same with edges decoded for emit
    const uint64_t *keep = nullptr;   // block filter cube mask (one filtered solid)
};

// One call as the stage kernels see it.
struct Pipe {
    const Seg *segs = nullptr;
    uint32_t nsegs = 0;
    Seg seg0{};
    const uint32_t *word_seg = nullptr, *brick_seg = nullptr;  // segment tables (calls of several)
    const uint32_t *pos_seg = nullptr;  // segment per 64 scan positions (word_seg unless brick-major)
    Ctl *ctl = nullptr;
    SegOut *sout = nullptr;
    uint64_t nwords = 0, nbricks = 0, nscan = 0;
    uint32_t cap = 0;
    bool cull = false, bm = false, stats = false;
    int mode = 0, rows_per_walk = 1, wpb = 4;
    size_t plan_lds = 0;  // dynamic LDS of the plan (its stacks, the staged program)
    uint32_t plan_gsz = 64;  // bricks per plan wave
    bool plan_tile = false;  // plan waves take compact tiles of bricks (one solid)
    bool scan_one_pass = false;  // classify scans its counts itself (at most SCAN_ONE_PASS blocks)
    uint32_t emit_div = 1024;  // emit blocks: occupancy words / emit_div, in [512, 4096]
};

hipError_t launch_plan(const Pipe &p, const Bufs &b, hipStream_t st);
// Instantiated once per (MODE, STATS) in k_corners.hip.
template <int MODE, bool STATS>
hipError_t launch_corners_ms(const Pipe &p, const Bufs &b, hipStream_t st);
template <>
hipError_t launch_corners_ms<0, false>(const Pipe &p, const Bufs &b, hipStream_t st);
template <>
hipError_t launch_corners_ms<0, true>(const Pipe &p, const Bufs &b, hipStream_t st);
template <>
hipError_t launch_corners_ms<1, false>(const Pipe &p, const Bufs &b, hipStream_t st);
template <>
hipError_t launch_corners_ms<1, true>(const Pipe &p, const Bufs &b, hipStream_t st);
inline hipError_t launch_corners(const Pipe &p, const Bufs &b, hipStream_t st) {
    if (p.mode == 0) return p.stats ? launch_corners_ms<0, true>(p, b, st) : launch_corners_ms<0, false>(p, b, st);
    return p.stats ? launch_corners_ms<1, true>(p, b, st) : launch_corners_ms<1, false>(p, b, st);
}
// zeroes bytes (a multiple of 16) of device memory with a kernel
hipError_t launch_zero(void *p, size_t bytes, hipStream_t st);
hipError_t launch_classify(const Pipe &p, const Bufs &b, hipStream_t st);
// the scan of classify's counts, the list of emitting words and (calls of
// several solids) each solid's output ranges
hipError_t launch_scan(const Pipe &p, const Bufs &b, hipStream_t st);
hipError_t launch_emit(const Pipe &p, const Bufs &b, uint32_t *tris, uint64_t vcap, uint64_t tcap, hipStream_t st);
hipError_t launch_bisect(const Pipe &p, const Bufs &b, double *verts, double *interior, uint64_t vcap,
                         hipStream_t st);
hipError_t launch_bisect_host(int mode, const DevProg &P, const Grid &g, const uint64_t *bits,
                              const uint64_t *vkeys, double *state, double *probe, double *verts, double *interior,
                              uint64_t nv, int phase, hipStream_t st);
hipError_t launch_host_points(const DevProg &P, const Grid &g, const double *pts, uint32_t frame, uint64_t start,
                              uint64_t n, double *out, hipStream_t st);
hipError_t launch_contains(int mode, const DevProg &P, const double *pts, uint8_t *out, uint64_t n, hipStream_t st);
hipError_t launch_inverse_apply(const double *recs, uint32_t nops, double *verts, const Ctl *ctl, uint64_t vcap,
                                hipStream_t st);
// reads (write false) or writes `items` 16-byte items of buf; the read form
// writes one word per block (8192 blocks) to out
hipError_t launch_stream(bool write, void *buf, uint64_t items, uint32_t *out, hipStream_t st);
hipError_t launch_encode_stl(const double *verts, const uint32_t *tris, uint64_t nt, uint16_t *body,
                             hipStream_t st);

// A mesh collider's triangles (9 doubles each) binned by their boxes in a
// uniform grid of n[0] x n[1] x n[2] cells: cell (x, y, z) holds
// cell_tris[cell_start[c], cell_start[c + 1]), c = (z * n[1] + y) * n[0] + x;
// point p is in cell floor((p - o) * inv), clamped to the grid.
// Beside it, a finer grid of the vertices that are a segment's origin (all
// but a triangle's largest, NewSegment's order) as a summed-volume table:
// vsat[(z * vn[1] + y) * vn[0] + x] counts those in cells <= (x, y, z).
struct RectGridDev {
    const double *tris = nullptr;
    uint32_t ntris = 0;
    const uint32_t *cell_start = nullptr, *cell_tris = nullptr;
    double o[3] = {0, 0, 0}, inv = 1;
    int n[3] = {1, 1, 1};
    const uint32_t *vsat = nullptr;
    double vo[3] = {0, 0, 0}, vinv = 1;
    int vn[3] = {1, 1, 1};
};
// A block of cubes [mn, mx) of MarchingCubesFilter's recursion (corner indices).
struct FBlock {
    int mn[3], mx[3];
};
// The recursion's filter: RectCollision of R with a block's bounds grown by
// eps, then by extra; accepted leaves set their cubes in keep (occupancy
// word layout, W words per row, ny rows per plane).
struct BlockFilter {
    RectGridDev R;
    const double *xs, *ys, *zs;
    double eps, extra;
    uint64_t *keep;
    int W, ny;
    Ctl *ctl;  // filter_overflow: a level outgrew its block list
};
hipError_t launch_rgrid_gather(const double *verts, const uint32_t *tris, uint64_t nt, double *out, hipStream_t st);
// counts must start at zero
hipError_t launch_rgrid_count(const RectGridDev &R, uint32_t *counts, hipStream_t st);
// start (ncell + 1 entries) and cursor from the counts; tiles: ceil(ncell / 1024) <= 1024 entries
hipError_t launch_rgrid_scan(const uint32_t *counts, uint32_t ncell, uint32_t *tiles, uint32_t *start,
                             uint32_t *cursor, hipStream_t st);
hipError_t launch_rgrid_fill(const RectGridDev &R, uint32_t *cursor, uint32_t *cell_tris, hipStream_t st);
// R.vsat's table from the triangles: vsat must start at zero
hipError_t launch_vgrid(const RectGridDev &R, uint32_t *vsat, hipStream_t st);
// level 0's list = {root}; level_counts[0] = 1, level_counts[1..nlevels] = 0
hipError_t launch_filter_root(const FBlock &root, FBlock *front, uint32_t *level_counts, uint32_t nlevels,
                              hipStream_t st);
hipError_t launch_filter_level(const BlockFilter &F, const FBlock *in, const uint32_t *nin, FBlock *out,
                               uint32_t *nout, uint32_t cap, uint64_t max_in, hipStream_t st);

// GroupTriangles (k_group.hip): a triangle's box, and one level of the
// recursion: the three orders, the axis of each of the level's 2^L nodes,
// the first-half flag of each triangle.
struct GroupBox {
    double mn[3], mx[3];
};
struct GroupLevel {
    uint32_t n = 0;
    int L = 0;
    const GroupBox *box = nullptr;
    const uint32_t *perm[3] = {nullptr, nullptr, nullptr};
    uint8_t *axis = nullptr;
    uint8_t *flag = nullptr;
};
// boxes and the three midpoint key arrays of triangles (verts + tris, or 9
// doubles each when tris is null); iota[t] = t
hipError_t launch_grp_boxes(const double *verts, const uint32_t *tris, uint32_t n, GroupBox *box, uint64_t *keys[3],
                            uint32_t *iota, hipStream_t st);
// stable radix sort of (key, value) pairs / exclusive sum (hipCUB; a null
// tmp asks for the temporary bytes)
hipError_t grp_sort(void *tmp, size_t &bytes, const uint64_t *keys_in, uint64_t *keys_out, const uint32_t *vals_in,
                    uint32_t *vals_out, uint32_t n, hipStream_t st);
hipError_t grp_scan(void *tmp, size_t &bytes, const uint32_t *in, uint32_t *out, uint32_t n, hipStream_t st);
// the axis of every node (slots: 36 keys per node of a level whose nodes
// hold more than 64 triangles)
hipError_t launch_grp_axis(const GroupLevel &G, unsigned long long *slots, hipStream_t st);
hipError_t launch_grp_flags(const GroupLevel &G, hipStream_t st);
hipError_t launch_grp_gather_flags(const GroupLevel &G, int a, uint32_t *f, hipStream_t st);
hipError_t launch_grp_scatter(const GroupLevel &G, int a, const uint32_t *scan, uint32_t *out, hipStream_t st);
// out[k] = the three indices of triangle perm[k]
hipError_t launch_grp_apply(const uint32_t *tris, const uint32_t *perm, uint32_t n, uint32_t *out, hipStream_t st);

// z-slab stitching (k_slab.hip).  The x and y edges of plane z among a
// slab's vertices (keys vkeys): keys without z and vertex indices, in
// vertex order, *count of them (flags, scan: nv entries of scratch; tmp:
// grp_scan's temporary storage).
hipError_t launch_plane_list(const uint64_t *vkeys, uint32_t nv, uint32_t z, uint32_t *flags, uint32_t *scan,
                             void *tmp, size_t tmp_bytes, uint64_t *keys, uint32_t *idx, uint32_t *count,
                             hipStream_t st);
// keep[i] = 1, then 0 at the ndrop indices of drop
hipError_t launch_stitch_keep(uint32_t *keep, uint32_t nv, const uint32_t *drop, uint32_t ndrop, hipStream_t st);
// kept vertex i (rank: exclusive sums of keep) to fv[vbase + rank[i]] (and
// its interior point to fin); newidx[i] = its index in the mesh
hipError_t launch_stitch_verts(const double *v, const double *in, uint32_t nv, const uint32_t *keep,
                               const uint32_t *rank, uint32_t vbase, double *fv, double *fin, uint32_t *newidx,
                               hipStream_t st);
// newidx[up_idx[j]] = lo_global[j] (both lists sorted by key; a key
// mismatch sets *bad)
hipError_t launch_stitch_pairs(const uint64_t *lo_keys, const uint32_t *lo_global, const uint64_t *up_keys,
                               const uint32_t *up_idx, uint32_t m, uint32_t *newidx, uint32_t *bad, hipStream_t st);
hipError_t launch_stitch_tris(const uint32_t *t, uint64_t nt, const uint32_t *newidx, uint32_t *ft, hipStream_t st);
// global[j] = newidx[idx[j]]
hipError_t launch_stitch_globals(const uint32_t *idx, uint32_t m, const uint32_t *newidx, uint32_t *global,
                                 hipStream_t st);

inline unsigned grid_for(uint64_t n, unsigned block, unsigned cap = 1u << 20) {
    uint64_t b = (n + block - 1) / block;
    if (b == 0) b = 1;
    return (unsigned)(b < cap ? b : cap);
}

}  // namespace mc3d
