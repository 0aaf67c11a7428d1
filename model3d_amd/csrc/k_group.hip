// k_group.hip -- GroupTriangles (GroupBounders, model3d/bvh.go:118-156,
// 219-292) on the device: the triangle order that SaveGroupedSTL writes
// (mesh.go:721-743) and that GroupedTrianglesToCollider expects.
//
// The reference sorts the triangles three times by their boxes' midpoints
// (one order per axis), then recursively cuts each node's range at n / 2
// along the axis whose two halves' box areas sum least, stably partitioning
// the other two orders by the halves; a node of one or two triangles is
// left in x order.  Here the recursion runs level by level over all nodes
// at once: the node of position p at level L is found by halving [0, n) L
// times, so no node table is kept.  Per level: the axis of every node (box
// reductions over its two halves in each order), a flag per triangle (in
// the first half along its node's axis), and a stable partition of each
// order inside its nodes (an exclusive scan of the flags).  The three
// initial sorts are stable radix sorts, so equal midpoints stay in index
// order (the reference's sort.Slice leaves their order unspecified).
#include <hipcub/hipcub.hpp>

#include "launch.h"
#include "rect_collide.h"

namespace mc3d {
namespace {

// A double as a key whose unsigned order is the numeric order (-0 as +0,
// NaN last).
__device__ __forceinline__ uint64_t order_key(double x) {
    if (isnan(x)) return ~0ull;
    if (x == 0) x = 0;
    const uint64_t u = (uint64_t)__double_as_longlong(x);
    return (u >> 63) ? ~u : (u | (1ull << 63));
}
__device__ __forceinline__ double key_double(uint64_t k) {
    return __longlong_as_double((long long)((k >> 63) ? (k & ~(1ull << 63)) : ~k));
}

// The node of position p at level L: [lo, hi) after L halvings of [0, n)
// (the first half is the n / 2 smallest positions).
__device__ __forceinline__ uint32_t node_of(uint32_t p, uint32_t n, int L, uint32_t &lo, uint32_t &hi) {
    lo = 0, hi = n;
    uint32_t k = 0;
    for (int l = 0; l < L; l++) {
        const uint32_t mid = lo + (hi - lo) / 2;
        if (p < mid) hi = mid, k = 2 * k;
        else lo = mid, k = 2 * k + 1;
    }
    return k;
}
// The range of node k at level L.
__device__ __forceinline__ void node_range(uint32_t k, uint32_t n, int L, uint32_t &lo, uint32_t &hi) {
    lo = 0, hi = n;
    for (int l = L - 1; l >= 0; l--) {
        const uint32_t mid = lo + (hi - lo) / 2;
        if ((k >> l) & 1u) lo = mid;
        else hi = mid;
    }
}

// Triangle boxes (Triangle.Min / Max, primitives.go:35-43) and the midpoint
// keys of one axis.
__global__ __launch_bounds__(256) void k_grp_boxes(const double *__restrict__ verts, const uint32_t *__restrict__ tris,
                                                   uint32_t n, GroupBox *__restrict__ box, uint64_t *__restrict__ key0,
                                                   uint64_t *__restrict__ key1, uint64_t *__restrict__ key2,
                                                   uint32_t *__restrict__ iota) {
    for (uint32_t t = blockIdx.x * blockDim.x + threadIdx.x; t < n; t += gridDim.x * blockDim.x) {
        const double *p[3];
        for (int k = 0; k < 3; k++) p[k] = verts + 3 * (size_t)(tris ? tris[3 * (size_t)t + k] : 3 * t + k);
        GroupBox b;
        uint64_t *keys[3] = {key0, key1, key2};
        for (int a = 0; a < 3; a++) {
            b.mn[a] = rc::go_min(rc::go_min(p[0][a], p[1][a]), p[2][a]);
            b.mx[a] = rc::go_max(rc::go_max(p[0][a], p[1][a]), p[2][a]);
            keys[a][t] = order_key((b.mn[a] + b.mx[a]) * 0.5);  // Coord3D.Mid, coords.go:196-198
        }
        box[t] = b;
        iota[t] = t;
    }
}

// boundsArea (bvh.go:289-292) of a box
__device__ __forceinline__ double bounds_area(const double *mn, const double *mx) {
    const double dx = mx[0] + -mn[0], dy = mx[1] + -mn[1], dz = mx[2] + -mn[2];
    return 2 * (dx * (dy + dz) + dy * dz);
}

// The node's axis (bestSplitAxis, bvh.go:219-236) from the six half boxes.
__device__ __forceinline__ uint8_t best_axis(const double hmn[3][2][3], const double hmx[3][2][3]) {
    int axis = 0;
    double best = bounds_area(hmn[0][0], hmx[0][0]) + bounds_area(hmn[0][1], hmx[0][1]);
    for (int a = 1; a < 3; a++) {
        const double v = bounds_area(hmn[a][0], hmx[a][0]) + bounds_area(hmn[a][1], hmx[a][1]);
        if (v < best) best = v, axis = a;
    }
    return (uint8_t)axis;
}

// Small nodes (at most 64 triangles): a thread per node reduces its halves
// in order, as multipleBoundsArea does (bvh.go:257-287).
__global__ __launch_bounds__(256) void k_grp_axis_small(GroupLevel G) {
    const uint32_t nodes = 1u << G.L;
    for (uint32_t k = blockIdx.x * blockDim.x + threadIdx.x; k < nodes; k += gridDim.x * blockDim.x) {
        uint32_t lo, hi;
        node_range(k, G.n, G.L, lo, hi);
        if (hi - lo <= 2) {
            G.axis[k] = 0;
            continue;
        }
        const uint32_t mid = lo + (hi - lo) / 2;
        double hmn[3][2][3], hmx[3][2][3];
        for (int a = 0; a < 3; a++)
            for (int h = 0; h < 2; h++) {
                const uint32_t s = h ? mid : lo, e = h ? hi : mid;
                const GroupBox &b0 = G.box[G.perm[a][s]];
                for (int c = 0; c < 3; c++) hmn[a][h][c] = b0.mn[c], hmx[a][h][c] = b0.mx[c];
                for (uint32_t p = s + 1; p < e; p++) {
                    const GroupBox &b = G.box[G.perm[a][p]];
                    for (int c = 0; c < 3; c++) {
                        if (b.mn[c] < hmn[a][h][c]) hmn[a][h][c] = b.mn[c];
                        if (b.mx[c] > hmx[a][h][c]) hmx[a][h][c] = b.mx[c];
                    }
                }
            }
        G.axis[k] = best_axis(hmn, hmx);
    }
}

// Large nodes: every position of the three orders folds its box into its
// node half's slot (36 ordered keys per node: axis, half, min xyz, max xyz)
// with atomic min / max; a wave whose 64 x GRP_RUN positions all lie in one
// node half folds them first.  Equal to the in-order reduction for boxes
// without NaN (the min and max do not depend on the order; only the sign
// of a zero might, and it cannot change an area comparison).
constexpr int GRP_RUN = 16;

__device__ __forceinline__ void fold_box(double *mn, double *mx, const GroupBox &b) {
    for (int c = 0; c < 3; c++) mn[c] = fmin(mn[c], b.mn[c]), mx[c] = fmax(mx[c], b.mx[c]);
}

__device__ __forceinline__ void flush_box(unsigned long long *slot, const double *mn, const double *mx) {
    for (int c = 0; c < 3; c++) {
        atomicMin(&slot[c], (unsigned long long)order_key(mn[c]));
        atomicMax(&slot[3 + c], (unsigned long long)order_key(mx[c]));
    }
}

__global__ __launch_bounds__(256) void k_grp_fold(GroupLevel G, unsigned long long *__restrict__ slots) {
    const int lane = threadIdx.x & 63;
    const uint32_t per_axis = (G.n + GRP_RUN - 1) / GRP_RUN;
    const uint32_t wave = (blockIdx.x * blockDim.x + threadIdx.x) >> 6, nw = (gridDim.x * blockDim.x) >> 6;
    const uint32_t waves_per_axis = (per_axis + 63) / 64;
    for (uint32_t w = wave; w < 3 * waves_per_axis; w += nw) {
        const int a = (int)(w / waves_per_axis);
        const uint32_t chunk = (w - a * waves_per_axis) * 64 + lane;
        const uint32_t p0 = chunk * GRP_RUN, p1 = min(p0 + GRP_RUN, G.n);
        // the node half of the chunk's first and last positions
        uint32_t lo, hi, key0 = ~0u, key1 = ~0u;
        if (p0 < p1) {
            uint32_t k = node_of(p0, G.n, G.L, lo, hi);
            key0 = 2 * k + (p0 >= lo + (hi - lo) / 2);
            k = node_of(p1 - 1, G.n, G.L, lo, hi);
            key1 = 2 * k + (p1 - 1 >= lo + (hi - lo) / 2);
        }
        const uint32_t first = __shfl(key0, 0, 64);
        const bool uniform = __all(p0 >= p1 || (key0 == first && key1 == first));
        if (uniform) {
            double mn[3] = {INFINITY, INFINITY, INFINITY}, mx[3] = {-INFINITY, -INFINITY, -INFINITY};
            for (uint32_t p = p0; p < p1; p++) fold_box(mn, mx, G.box[G.perm[a][p]]);
            for (int d = 32; d >= 1; d >>= 1)
                for (int c = 0; c < 3; c++) {
                    mn[c] = fmin(mn[c], __shfl_xor(mn[c], d, 64));
                    mx[c] = fmax(mx[c], __shfl_xor(mx[c], d, 64));
                }
            if (lane == 0 && first != ~0u) flush_box(slots + (size_t)(first >> 1) * 36 + a * 12 + (first & 1) * 6, mn, mx);
            continue;
        }
        // runs of positions in one node half, flushed as they end
        uint32_t cur = ~0u;
        double mn[3], mx[3];
        for (uint32_t p = p0; p < p1; p++) {
            const uint32_t k = node_of(p, G.n, G.L, lo, hi);
            const uint32_t key = 2 * k + (p >= lo + (hi - lo) / 2);
            if (key != cur) {
                if (cur != ~0u) flush_box(slots + (size_t)(cur >> 1) * 36 + a * 12 + (cur & 1) * 6, mn, mx);
                cur = key;
                for (int c = 0; c < 3; c++) mn[c] = INFINITY, mx[c] = -INFINITY;
            }
            fold_box(mn, mx, G.box[G.perm[a][p]]);
        }
        if (cur != ~0u) flush_box(slots + (size_t)(cur >> 1) * 36 + a * 12 + (cur & 1) * 6, mn, mx);
    }
}

__global__ __launch_bounds__(256) void k_grp_slots_init(unsigned long long *__restrict__ slots, uint32_t nodes) {
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < nodes * 36; i += gridDim.x * blockDim.x)
        slots[i] = (i % 6) < 3 ? ~0ull : 0ull;
}

__global__ __launch_bounds__(256) void k_grp_axis_slots(GroupLevel G, const unsigned long long *__restrict__ slots) {
    const uint32_t nodes = 1u << G.L;
    for (uint32_t k = blockIdx.x * blockDim.x + threadIdx.x; k < nodes; k += gridDim.x * blockDim.x) {
        uint32_t lo, hi;
        node_range(k, G.n, G.L, lo, hi);
        if (hi - lo <= 2) {
            G.axis[k] = 0;
            continue;
        }
        double hmn[3][2][3], hmx[3][2][3];
        const unsigned long long *s = slots + (size_t)k * 36;
        for (int a = 0; a < 3; a++)
            for (int h = 0; h < 2; h++)
                for (int c = 0; c < 3; c++) {
                    hmn[a][h][c] = key_double(s[a * 12 + h * 6 + c]);
                    hmx[a][h][c] = key_double(s[a * 12 + h * 6 + 3 + c]);
                }
        G.axis[k] = best_axis(hmn, hmx);
    }
}

// flag[t] = 1 for the triangles in the first half of their node along the
// node's axis (splitBounders, bvh.go:159-217).
__global__ __launch_bounds__(256) void k_grp_flags(GroupLevel G) {
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < 3 * G.n; i += gridDim.x * blockDim.x) {
        const int a = (int)(i / G.n);
        const uint32_t p = i - a * G.n;
        uint32_t lo, hi;
        const uint32_t k = node_of(p, G.n, G.L, lo, hi);
        if (G.axis[k] == a) G.flag[G.perm[a][p]] = p < lo + (hi - lo) / 2;
    }
}

__global__ __launch_bounds__(256) void k_grp_gather_flags(GroupLevel G, int a, uint32_t *__restrict__ f) {
    for (uint32_t p = blockIdx.x * blockDim.x + threadIdx.x; p < G.n; p += gridDim.x * blockDim.x)
        f[p] = G.flag[G.perm[a][p]];
}

// The stable partition of order a inside each node: flagged triangles
// first, each side in its previous order (scan: exclusive sums of f).
__global__ __launch_bounds__(256) void k_grp_scatter(GroupLevel G, int a, const uint32_t *__restrict__ scan,
                                                     uint32_t *__restrict__ out) {
    for (uint32_t p = blockIdx.x * blockDim.x + threadIdx.x; p < G.n; p += gridDim.x * blockDim.x) {
        uint32_t lo, hi;
        node_of(p, G.n, G.L, lo, hi);
        const uint32_t mid = lo + (hi - lo) / 2, t = G.perm[a][p];
        const uint32_t r = scan[p] - scan[lo];
        out[G.flag[t] ? lo + r : mid + (p - lo - r)] = t;
    }
}

__global__ __launch_bounds__(256) void k_grp_apply(const uint32_t *__restrict__ tris, const uint32_t *__restrict__ perm,
                                                   uint32_t n, uint32_t *__restrict__ out) {
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < 3 * n; i += gridDim.x * blockDim.x)
        out[i] = tris[3 * (size_t)perm[i / 3] + i % 3];
}

}  // namespace

hipError_t launch_grp_apply(const uint32_t *tris, const uint32_t *perm, uint32_t n, uint32_t *out, hipStream_t st) {
    k_grp_apply<<<grid_for(3ull * n, 256, 8192), 256, 0, st>>>(tris, perm, n, out);
    return hipGetLastError();
}

hipError_t launch_grp_boxes(const double *verts, const uint32_t *tris, uint32_t n, GroupBox *box, uint64_t *keys[3],
                            uint32_t *iota, hipStream_t st) {
    k_grp_boxes<<<grid_for(n, 256, 4096), 256, 0, st>>>(verts, tris, n, box, keys[0], keys[1], keys[2], iota);
    return hipGetLastError();
}

hipError_t grp_sort(void *tmp, size_t &bytes, const uint64_t *keys_in, uint64_t *keys_out, const uint32_t *vals_in,
                    uint32_t *vals_out, uint32_t n, hipStream_t st) {
    return hipcub::DeviceRadixSort::SortPairs(tmp, bytes, keys_in, keys_out, vals_in, vals_out, (int)n, 0, 64, st);
}

hipError_t grp_scan(void *tmp, size_t &bytes, const uint32_t *in, uint32_t *out, uint32_t n, hipStream_t st) {
    return hipcub::DeviceScan::ExclusiveSum(tmp, bytes, in, out, (int)n, st);
}

hipError_t launch_grp_axis(const GroupLevel &G, unsigned long long *slots, hipStream_t st) {
    const uint32_t nodes = 1u << G.L;
    const uint32_t size = (G.n + nodes - 1) / nodes;  // the largest node of the level
    if (size <= 64) {
        k_grp_axis_small<<<grid_for(nodes, 256, 4096), 256, 0, st>>>(G);
        return hipGetLastError();
    }
    k_grp_slots_init<<<grid_for((uint64_t)nodes * 36, 256, 4096), 256, 0, st>>>(slots, nodes);
    const uint64_t waves = 3 * (((G.n + GRP_RUN - 1) / GRP_RUN + 63) / 64);
    k_grp_fold<<<grid_for(waves, 4, 8192), 256, 0, st>>>(G, slots);
    k_grp_axis_slots<<<grid_for(nodes, 256, 4096), 256, 0, st>>>(G, slots);
    return hipGetLastError();
}

hipError_t launch_grp_flags(const GroupLevel &G, hipStream_t st) {
    k_grp_flags<<<grid_for(3ull * G.n, 256, 8192), 256, 0, st>>>(G);
    return hipGetLastError();
}

hipError_t launch_grp_gather_flags(const GroupLevel &G, int a, uint32_t *f, hipStream_t st) {
    k_grp_gather_flags<<<grid_for(G.n, 256, 8192), 256, 0, st>>>(G, a, f);
    return hipGetLastError();
}

hipError_t launch_grp_scatter(const GroupLevel &G, int a, const uint32_t *scan, uint32_t *out, hipStream_t st) {
    k_grp_scatter<<<grid_for(G.n, 256, 8192), 256, 0, st>>>(G, a, scan, out);
    return hipGetLastError();
}

}  // namespace mc3d
