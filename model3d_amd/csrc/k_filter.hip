// k_filter.hip -- MarchingCubesFilter's block recursion with a mesh
// collider's RectCollision as the filter (MarchingCubesC2F, and
// MarchingCubesSearchFilter with the library's rect collider), on the
// device.
//
//   MarchingCubesFilter / C2F                  model3d/mc.go:83-180
//   mcBlock (Bounds, Volume, Pieces, Split)    model3d/mc.go:724-797
//   MeshToCollider(...).RectCollision          model3d/collisions.go:138-142, 169-179, 344-356
//
// The collider's triangles are binned in a uniform 3D grid by their boxes;
// a rect's answer is the OR over the triangles of the cells it overlaps of
// "the triangle's box overlaps the rect and Triangle.RectCollision" -- the
// answer the collider's tree gives, whose every node box contains the boxes
// of the triangles below it.  The recursion runs level by level: one wave
// per block asks the filter; an accepted block whose volume / 2 is below 64
// sets its cubes in the cube mask, a larger one pushes its two halves to the
// next level.  The reference's two-stage recursion (first down to
// max(volume / 4096, 64), then each queued block down to 64) asks the same
// deterministic filter of the same blocks, so its accepted leaves are these.
#include "launch.h"
#include "rect_collide.h"

// Blocks per wave in the recursion's level kernel.
#ifndef MC3D_FILTER_GROUP
#define MC3D_FILTER_GROUP 4
#endif

namespace mc3d {
namespace {

constexpr int FGROUP = MC3D_FILTER_GROUP;
static_assert(FGROUP >= 1 && FGROUP <= 64, "blocks per wave: 1 to 64");

__device__ __forceinline__ void cell_range(double lo, double hi, double o, double inv, int n, int &c0, int &c1) {
    // clamped to the grid (a NaN bound: cell 0; such a rect overlaps no box)
    const double a = floor((lo - o) * inv), b = floor((hi - o) * inv);
    c0 = a >= 0 ? (a <= n - 1 ? (int)a : n - 1) : 0;
    c1 = b >= 0 ? (b <= n - 1 ? (int)b : n - 1) : 0;
}

// The cells of a triangle's box; false for a box with NaN bounds (such a
// triangle's box overlaps no rect).
__device__ __forceinline__ bool tri_cells(const RectGridDev &R, const double *p, int *c0, int *c1) {
    double tmn[3], tmx[3];
    rc::tri_box(p, tmn, tmx);
    for (int a = 0; a < 3; a++) {
        if (isnan(tmn[a]) || isnan(tmx[a])) return false;
        cell_range(tmn[a], tmx[a], R.o[a], R.inv, R.n[a], c0[a], c1[a]);
    }
    return true;
}

__global__ __launch_bounds__(256) void k_rgrid_gather(const double *__restrict__ verts,
                                                      const uint32_t *__restrict__ tris, uint64_t nt,
                                                      double *__restrict__ out) {
    for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < 9 * nt;
         i += (uint64_t)gridDim.x * blockDim.x) {
        const uint64_t t = i / 9, r = i - 9 * t;
        out[i] = verts[3 * (uint64_t)tris[3 * t + r / 3] + r % 3];
    }
}

__global__ __launch_bounds__(256) void k_rgrid_count(RectGridDev R, uint32_t *__restrict__ counts) {
    for (uint32_t t = blockIdx.x * blockDim.x + threadIdx.x; t < R.ntris; t += gridDim.x * blockDim.x) {
        int c0[3], c1[3];
        if (!tri_cells(R, R.tris + 9 * (size_t)t, c0, c1)) continue;
        for (int z = c0[2]; z <= c1[2]; z++)
            for (int y = c0[1]; y <= c1[1]; y++)
                for (int x = c0[0]; x <= c1[0]; x++) atomicAdd(&counts[((size_t)z * R.n[1] + y) * R.n[0] + x], 1u);
    }
}

__global__ __launch_bounds__(256) void k_rgrid_fill(RectGridDev R, uint32_t *__restrict__ cursor,
                                                    uint32_t *__restrict__ cell_tris) {
    for (uint32_t t = blockIdx.x * blockDim.x + threadIdx.x; t < R.ntris; t += gridDim.x * blockDim.x) {
        int c0[3], c1[3];
        if (!tri_cells(R, R.tris + 9 * (size_t)t, c0, c1)) continue;
        for (int z = c0[2]; z <= c1[2]; z++)
            for (int y = c0[1]; y <= c1[1]; y++)
                for (int x = c0[0]; x <= c1[0]; x++)
                    cell_tris[atomicAdd(&cursor[((size_t)z * R.n[1] + y) * R.n[0] + x], 1u)] = t;
    }
}

// Counts each segment origin (rc::origin_vertices) in its cell of the
// vertex grid; one outside the grid is not counted.
__global__ __launch_bounds__(256) void k_vgrid_count(RectGridDev R, uint32_t *__restrict__ vsat) {
    for (uint32_t t = blockIdx.x * blockDim.x + threadIdx.x; t < R.ntris; t += gridDim.x * blockDim.x) {
        const double *p = R.tris + 9 * (size_t)t;
        bool origin[3];
        if (!rc::origin_vertices(p, origin)) continue;
        for (int k = 0; k < 3; k++) {
            if (!origin[k]) continue;
            int c[3];
            bool in = true;
            for (int a = 0; a < 3; a++) {
                const double f = floor((p[3 * k + a] - R.vo[a]) * R.vinv);
                in = in && f >= 0 && f < R.vn[a];
                c[a] = in ? (int)f : 0;
            }
            if (in) atomicAdd(&vsat[((size_t)c[2] * R.vn[1] + c[1]) * R.vn[0] + c[0]], 1u);
        }
    }
}

// Inclusive prefix sums along x (a wave per row), then y and z (a thread
// per column): the summed-volume table in place.
__global__ __launch_bounds__(256) void k_vsat_x(uint32_t *__restrict__ v, int n0, uint32_t rows) {
    const int lane = threadIdx.x & 63;
    const uint32_t wave = (blockIdx.x * blockDim.x + threadIdx.x) >> 6, nw = (gridDim.x * blockDim.x) >> 6;
    for (uint32_t r = wave; r < rows; r += nw) {
        uint32_t *row = v + (size_t)r * n0, carry = 0;
        for (int x0 = 0; x0 < n0; x0 += 64) {
            const int x = x0 + lane;
            uint32_t s = x < n0 ? row[x] : 0u;
            for (int d = 1; d < 64; d <<= 1) {
                const uint32_t u = __shfl_up(s, d, 64);
                if (lane >= d) s += u;
            }
            if (x < n0) row[x] = s + carry;
            carry += __shfl(s, 63, 64);
        }
    }
}

__global__ __launch_bounds__(256) void k_vsat_yz(uint32_t *__restrict__ v, int n0, int n1, int n2, int axis) {
    // axis 1: columns (x, z) along y; axis 2: columns (x, y) along z
    const uint32_t ncol = (uint32_t)n0 * (axis == 1 ? n2 : n1);
    for (uint32_t c = blockIdx.x * blockDim.x + threadIdx.x; c < ncol; c += gridDim.x * blockDim.x) {
        const uint32_t x = c % n0, w = c / n0;
        size_t i, step;
        int len;
        if (axis == 1) i = (size_t)w * n1 * n0 + x, step = n0, len = n1;
        else i = (size_t)w * n0 + x, step = (size_t)n0 * n1, len = n2;
        uint32_t s = 0;
        for (int k = 0; k < len; k++, i += step) v[i] = s += v[i];
    }
}

constexpr int RSCAN = 1024;  // cells per scan tile (one per thread of a 1024-thread block)

// Inclusive scan of one value per thread over a 1024-thread block.
__device__ __forceinline__ uint32_t block_scan(uint32_t v, uint32_t *sh) {
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    for (int d = 1; d < 64; d <<= 1) {
        const uint32_t u = __shfl_up(v, d, 64);
        if (lane >= d) v += u;
    }
    if (lane == 63) sh[wave] = v;
    __syncthreads();
    if (wave == 0) {
        uint32_t w = lane < 16 ? sh[lane] : 0;
        for (int d = 1; d < 16; d <<= 1) {
            const uint32_t u = __shfl_up(w, d, 64);
            if (lane >= d) w += u;
        }
        if (lane < 16) sh[lane] = w;
    }
    __syncthreads();
    return v + (wave ? sh[wave - 1] : 0);
}

// Tile sums of the counts.
__global__ __launch_bounds__(RSCAN) void k_rscan_tiles(const uint32_t *__restrict__ counts, uint32_t n,
                                                       uint32_t *__restrict__ tiles) {
    __shared__ uint32_t sh[16];
    const uint32_t i = blockIdx.x * RSCAN + threadIdx.x;
    const uint32_t s = block_scan(i < n ? counts[i] : 0u, sh);
    if (threadIdx.x == RSCAN - 1) tiles[blockIdx.x] = s;
}

// Exclusive scan of the tile sums (at most RSCAN tiles, one block).
__global__ __launch_bounds__(RSCAN) void k_rscan_top(uint32_t *__restrict__ tiles, uint32_t ntiles) {
    __shared__ uint32_t sh[16];
    const uint32_t v = threadIdx.x < ntiles ? tiles[threadIdx.x] : 0u;
    const uint32_t s = block_scan(v, sh);
    if (threadIdx.x < ntiles) tiles[threadIdx.x] = s - v;
}

// start[i] = cursor[i] = exclusive prefix of counts; start[n] = the total.
__global__ __launch_bounds__(RSCAN) void k_rscan_down(const uint32_t *__restrict__ counts, uint32_t n,
                                                      const uint32_t *__restrict__ tiles, uint32_t *__restrict__ start,
                                                      uint32_t *__restrict__ cursor) {
    __shared__ uint32_t sh[16];
    const uint32_t i = blockIdx.x * RSCAN + threadIdx.x;
    const uint32_t v = i < n ? counts[i] : 0u;
    const uint32_t s = block_scan(v, sh) + tiles[blockIdx.x];
    if (i < n) start[i] = cursor[i] = s - v;
    if (i == n - 1) start[n] = s;
}

__global__ void k_filter_root(FBlock root, FBlock *__restrict__ front, uint32_t *__restrict__ level_counts,
                              uint32_t nlevels) {
    for (uint32_t i = threadIdx.x; i <= nlevels; i += blockDim.x) level_counts[i] = i == 0 ? 1u : 0u;
    if (threadIdx.x == 0) front[0] = root;
}

// Candidate k of n in a scattered order (P prime and above n: a bijection),
// so that the first 64 lanes sample the whole rect, not its first rows.
__device__ __forceinline__ uint32_t scatter(uint32_t k, uint32_t n, uint64_t P) { return (uint32_t)(k * P % n); }

// A sure hit from the vertex table: a segment origin in a cell strictly
// inside the rect's cell range.  For such a vertex v in cell k on an axis,
// floor((v - o) * inv) = k > floor((mn - o) * inv) gives v > mn (the
// rounded subtraction and product are monotone), and k < floor((mx - o) *
// inv) gives v < mx; a vertex inside makes the collision (rc::vertex_inside).
__device__ __forceinline__ bool sure_hit(const RectGridDev &R, const double *mn, const double *mx) {
    if (!R.vsat) return false;
    int lo[3], hi[3];
    for (int a = 0; a < 3; a++) {
        if (isnan(mn[a]) || isnan(mx[a])) return false;
        int c0, c1;
        cell_range(mn[a], mx[a], R.vo[a], R.vinv, R.vn[a], c0, c1);
        lo[a] = c0 + 1, hi[a] = c1 - 1;  // (clamping only narrows the range)
        if (hi[a] < lo[a]) return false;
    }
    auto S = [&](int x, int y, int z) -> uint32_t {
        if (x < 0 || y < 0 || z < 0) return 0u;
        return R.vsat[((size_t)z * R.vn[1] + y) * R.vn[0] + x];
    };
    const int x0 = lo[0] - 1, y0 = lo[1] - 1, z0 = lo[2] - 1, x1 = hi[0], y1 = hi[1], z1 = hi[2];
    const uint32_t n = S(x1, y1, z1) - S(x0, y1, z1) - S(x1, y0, z1) - S(x1, y1, z0) + S(x0, y0, z1) + S(x0, y1, z0) +
                       S(x1, y0, z0) - S(x0, y0, z0);
    return n != 0;
}

// RectCollision of the mesh for [mn, mx] where the vertex table gave no sure
// hit, asked by a whole wave (the answer is wave-uniform): every candidate
// triangle tested once -- the triangles binned in the cells the rect
// overlaps, taken 64 at a time (each (y, z) row of those cells is one run
// of the cell lists; a triangle listed in several of the rect's cells is
// tested at the first of them), or, for a rect over more cells than half
// the triangles, all triangles in a scattered order.
__device__ bool rect_hits(const RectGridDev &R, const double *mn, const double *mx) {
    const int lane = threadIdx.x & 63;
    int c0[3], c1[3];
    for (int a = 0; a < 3; a++) {
        cell_range(mn[a], mx[a], R.o[a], R.inv, R.n[a], c0[a], c1[a]);
        if (c1[a] < c0[a]) return false;  // min above max: the rect overlaps no box
    }
    const uint32_t lx = c1[0] - c0[0] + 1, ly = c1[1] - c0[1] + 1, lz = c1[2] - c0[2] + 1;
    const uint64_t ncell = (uint64_t)lx * ly * lz;
    if (2 * ncell > R.ntris) {
        for (uint32_t t0 = 0; t0 < R.ntris; t0 += 64) {
            bool h = false;
            if (t0 + lane < R.ntris) {
                const double *p = R.tris + 9 * (size_t)scatter(t0 + lane, R.ntris, 4294967291ull);
                h = rc::vertex_inside(p, mn, mx) || rc::triangle_hits_rect(p, mn, mx);
            }
            if (__ballot(h)) return true;
        }
        return false;
    }
    const uint32_t nrows = ly * lz;
    for (uint32_t r0 = 0; r0 < nrows; r0 += 64) {
        // lane i: row r0 + i, its run [a, a + len) of cell_tris
        const uint32_t r = r0 + lane;
        uint32_t a = 0, len = 0;
        int ry = 0, rz = 0;
        if (r < nrows) {
            ry = c0[1] + (int)(r % ly), rz = c0[2] + (int)(r / ly);
            const size_t row = ((size_t)rz * R.n[1] + ry) * R.n[0];
            a = R.cell_start[row + c0[0]];
            len = R.cell_start[row + c1[0] + 1] - a;
        }
        uint32_t incl = len;
        for (int d = 1; d < 64; d <<= 1) {
            const uint32_t u = __shfl_up(incl, d, 64);
            if (lane >= d) incl += u;
        }
        const uint32_t excl = incl - len, total = __shfl(incl, 63, 64);
        for (uint32_t e0 = 0; e0 < total; e0 += 64) {
            const uint32_t e = e0 + lane;
            // the row holding entry e: the first lane j with incl_j > e
            int lo = 0, hi = 63;
            for (int k = 0; k < 6; k++) {
                const int mid = (lo + hi) >> 1;
                if (__shfl(incl, mid, 64) > e) hi = mid;
                else lo = mid + 1;
            }
            const uint32_t aj = __shfl(a, lo, 64), xj = __shfl(excl, lo, 64);
            const int yj = __shfl(ry, lo, 64), zj = __shfl(rz, lo, 64);
            bool h = false;
            if (e < total) {
                const uint32_t idx = aj + (e - xj);
                const double *p = R.tris + 9 * (size_t)R.cell_tris[idx];
                double tmn[3], tmx[3];
                rc::tri_box(p, tmn, tmx);
                if (rc::box_overlap(mn, mx, tmn, tmx)) {
                    // the first of the rect's cells that list it: row (y, z)
                    // at max(its first cell, the rect's), then the cell's run
                    int t0[3], t1[3];
                    for (int ax = 0; ax < 3; ax++) cell_range(tmn[ax], tmx[ax], R.o[ax], R.inv, R.n[ax], t0[ax], t1[ax]);
                    if (yj == max(t0[1], c0[1]) && zj == max(t0[2], c0[2])) {
                        const size_t cf = ((size_t)zj * R.n[1] + yj) * R.n[0] + max(t0[0], c0[0]);
                        if (idx >= R.cell_start[cf] && idx < R.cell_start[cf + 1]) {
                            const rc::V t[3] = {rc::mk(p[0], p[1], p[2]), rc::mk(p[3], p[4], p[5]),
                                                rc::mk(p[6], p[7], p[8])};
                            h = rc::vertex_inside(p, mn, mx) || rc::tri_rect(t, mn, mx);
                        }
                    }
                }
            }
            if (__ballot(h)) return true;
        }
    }
    return false;
}

template <class T>
__device__ __forceinline__ T lane_value(T v, int j) {
    return __shfl(v, j, 64);
}

// One level of the recursion: wave w takes blocks FGROUP w .. FGROUP w +
// FGROUP - 1 of `in` (then strides), one per lane.  Each lane asks the vertex
// table; the blocks it leaves open get the wave-cooperative full test one
// at a time; the accepted ones split into `out` (one counter update per
// wave) or set their cubes (a leaf: the wave takes its rows).
__global__ __launch_bounds__(256) void k_filter_level(BlockFilter F, const FBlock *__restrict__ in,
                                                      const uint32_t *__restrict__ nin, FBlock *__restrict__ out,
                                                      uint32_t *__restrict__ nout, uint32_t cap) {
    const int lane = threadIdx.x & 63;
    const uint32_t wave = (blockIdx.x * blockDim.x + threadIdx.x) >> 6, nw = (gridDim.x * blockDim.x) >> 6;
    const uint32_t n = min(*nin, cap);  // (never above: the lists hold every block of a level)
    const double *ax[3] = {F.xs, F.ys, F.zs};
    for (uint32_t base = wave * FGROUP; base < n; base += nw * FGROUP) {
        const uint32_t i = base + lane;
        const bool valid = lane < FGROUP && i < n;
        FBlock b = valid ? in[i] : FBlock{{0, 0, 0}, {0, 0, 0}};
        // mcBlock.Bounds(epsilon) (mc.go:736-740), then the filter's
        // Rect.Expand(extra) (shapes.go:339-344; extra 0: the rect itself)
        double mn[3], mx[3];
        for (int a = 0; a < 3; a++) {
            mn[a] = (ax[a][b.mn[a]] + -F.eps) + -F.extra;
            mx[a] = (ax[a][b.mx[a]] + F.eps) + F.extra;
        }
        bool acc = valid && sure_hit(F.R, mn, mx);
        for (uint64_t open = __ballot(valid && !acc); open; open &= open - 1) {
            const int j = __ffsll((unsigned long long)open) - 1;
            double jmn[3], jmx[3];
            for (int a = 0; a < 3; a++) jmn[a] = lane_value(mn[a], j), jmx[a] = lane_value(mx[a], j);
            const bool hit = rect_hits(F.R, jmn, jmx);
            if (lane == j) acc = hit;
        }
        const int lx = b.mx[0] - b.mn[0], ly = b.mx[1] - b.mn[1], lz = b.mx[2] - b.mn[2];
        const bool leaf = acc && (int64_t)lx * ly * lz / 2 < 64;  // mcBlock.Pieces(64, ...), mc.go:760-775
        const bool split = acc && !leaf;
        // mcBlock.Split (mc.go:777-797): halve the longest axis (y, then z
        // on ties with x); the halves go to out[k + 2 r], r: the lane's rank
        const uint64_t sm = __ballot(split);
        if (sm) {
            uint32_t k = 0;
            if (lane == 0) k = atomicAdd(nout, 2u * (uint32_t)__popcll(sm));
            k = lane_value(k, 0);
            if (split) {
                const uint32_t r = (uint32_t)__popcll(sm & ((1ull << lane) - 1));
                int axis = 0;
                if (ly >= lx && ly >= lz) axis = 1;
                else if (lz >= lx && lz >= ly) axis = 2;
                FBlock b1 = b, b2 = b;
                b1.mx[axis] = (b.mx[axis] + b.mn[axis]) / 2;
                b2.mn[axis] = b1.mx[axis];
                if (k + 2 * r + 2 <= cap) out[k + 2 * r] = b1, out[k + 2 * r + 1] = b2;
                else atomicOr(&F.ctl->filter_overflow, 1u);
            }
        }
        // the leaves' cubes, one leaf at a time, its rows across the lanes
        for (uint64_t lm = __ballot(leaf); lm; lm &= lm - 1) {
            const int j = __ffsll((unsigned long long)lm) - 1;
            int m0[3], m1[3];
            for (int a = 0; a < 3; a++) m0[a] = lane_value(b.mn[a], j), m1[a] = lane_value(b.mx[a], j);
            const int jy = m1[1] - m0[1], rows = jy * (m1[2] - m0[2]);
            for (int r = lane; r < rows; r += 64) {
                const int y = m0[1] + r % jy, z = m0[2] + r / jy;
                unsigned long long *row = (unsigned long long *)F.keep + ((size_t)z * F.ny + y) * F.W;
                for (int x = m0[0]; x < m1[0];) {
                    const int w = x >> 6, lo = x & 63, hi = min(64, m1[0] - (w << 6));
                    const uint64_t m = (hi >= 64 ? ~0ull : ((1ull << hi) - 1)) & (~0ull << lo);
                    atomicOr(&row[w], (unsigned long long)m);
                    x = (w << 6) + hi;
                }
            }
        }
    }
}

}  // namespace

hipError_t launch_rgrid_gather(const double *verts, const uint32_t *tris, uint64_t nt, double *out, hipStream_t st) {
    if (!nt) return hipSuccess;
    k_rgrid_gather<<<grid_for(9 * nt, 256, 4096), 256, 0, st>>>(verts, tris, nt, out);
    return hipGetLastError();
}

hipError_t launch_rgrid_count(const RectGridDev &R, uint32_t *counts, hipStream_t st) {
    if (!R.ntris) return hipSuccess;
    k_rgrid_count<<<grid_for(R.ntris, 256, 4096), 256, 0, st>>>(R, counts);
    return hipGetLastError();
}

hipError_t launch_rgrid_scan(const uint32_t *counts, uint32_t ncell, uint32_t *tiles, uint32_t *start,
                             uint32_t *cursor, hipStream_t st) {
    const uint32_t ntiles = (ncell + RSCAN - 1) / RSCAN;
    if (ncell == 0 || ntiles > RSCAN) return hipErrorInvalidValue;
    k_rscan_tiles<<<ntiles, RSCAN, 0, st>>>(counts, ncell, tiles);
    k_rscan_top<<<1, RSCAN, 0, st>>>(tiles, ntiles);
    k_rscan_down<<<ntiles, RSCAN, 0, st>>>(counts, ncell, tiles, start, cursor);
    return hipGetLastError();
}

hipError_t launch_rgrid_fill(const RectGridDev &R, uint32_t *cursor, uint32_t *cell_tris, hipStream_t st) {
    if (!R.ntris) return hipSuccess;
    k_rgrid_fill<<<grid_for(R.ntris, 256, 4096), 256, 0, st>>>(R, cursor, cell_tris);
    return hipGetLastError();
}

hipError_t launch_vgrid(const RectGridDev &R, uint32_t *vsat, hipStream_t st) {
    if (R.ntris) k_vgrid_count<<<grid_for(R.ntris, 256, 4096), 256, 0, st>>>(R, vsat);
    const uint32_t rows = (uint32_t)R.vn[1] * R.vn[2];
    k_vsat_x<<<grid_for(rows, 4, 4096), 256, 0, st>>>(vsat, R.vn[0], rows);
    k_vsat_yz<<<grid_for((uint64_t)R.vn[0] * R.vn[2], 256, 4096), 256, 0, st>>>(vsat, R.vn[0], R.vn[1], R.vn[2], 1);
    k_vsat_yz<<<grid_for((uint64_t)R.vn[0] * R.vn[1], 256, 4096), 256, 0, st>>>(vsat, R.vn[0], R.vn[1], R.vn[2], 2);
    return hipGetLastError();
}

hipError_t launch_filter_root(const FBlock &root, FBlock *front, uint32_t *level_counts, uint32_t nlevels,
                              hipStream_t st) {
    k_filter_root<<<1, 256, 0, st>>>(root, front, level_counts, nlevels);
    return hipGetLastError();
}

hipError_t launch_filter_level(const BlockFilter &F, const FBlock *in, const uint32_t *nin, FBlock *out,
                               uint32_t *nout, uint32_t cap, uint64_t max_in, hipStream_t st) {
    // one wave per FGROUP blocks of the level (at most max_in of them), 4
    // waves per workgroup, at most 16384 waves (they stride over the rest)
    const uint32_t waves =
        (uint32_t)std::min<uint64_t>(std::max<uint64_t>((max_in + FGROUP - 1) / FGROUP, 1), 16384);
    k_filter_level<<<(waves + 3) / 4, 256, 0, st>>>(F, in, nin, out, nout, cap);
    return hipGetLastError();
}

}  // namespace mc3d
