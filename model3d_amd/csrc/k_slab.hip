// k_slab.hip -- stitching the meshes of a grid searched in z slabs.
//
// Slab s evaluates corner planes [z0, z1] and owns the cubes between; the
// plane z1 is also slab s + 1's first plane.  Each slab's search emits a
// vertex for every crossed edge its cubes use, so the x and y edges of a
// shared plane come out of both slabs, bit-identical (the same two corners
// and the same bisection).  The stitch keeps the lower slab's copy: the
// upper slab's copies are dropped and its triangles pointed at the lower
// slab's vertices, matched by edge (axis, x, y) after sorting both planes'
// lists by it.
#include "launch.h"

namespace mc3d {
namespace {

constexpr uint64_t KEY_XYA = (1ull << 42) - 1;  // a vertex key without its z (axis | x << 2 | y << 22)

__device__ __forceinline__ bool on_plane(uint64_t key, uint32_t z) {
    return (key & 3) < 2 && (uint32_t)(key >> 42) == z;
}

__global__ __launch_bounds__(256) void k_plane_flags(const uint64_t *__restrict__ vkeys, uint32_t nv, uint32_t z,
                                                     uint32_t *__restrict__ f) {
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < nv; i += gridDim.x * blockDim.x)
        f[i] = on_plane(vkeys[i], z) ? 1u : 0u;
}

__global__ __launch_bounds__(256) void k_plane_compact(const uint64_t *__restrict__ vkeys, uint32_t nv, uint32_t z,
                                                       const uint32_t *__restrict__ scan, uint64_t *__restrict__ keys,
                                                       uint32_t *__restrict__ idx, uint32_t *__restrict__ count) {
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < nv; i += gridDim.x * blockDim.x) {
        const uint64_t k = vkeys[i];
        const bool f = on_plane(k, z);
        if (f) keys[scan[i]] = k & KEY_XYA, idx[scan[i]] = i;
        if (i == nv - 1) *count = scan[i] + (f ? 1u : 0u);
    }
}

__global__ __launch_bounds__(256) void k_stitch_keep(uint32_t *__restrict__ keep, uint32_t nv) {
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < nv; i += gridDim.x * blockDim.x) keep[i] = 1u;
}
__global__ __launch_bounds__(256) void k_stitch_drop(uint32_t *__restrict__ keep, const uint32_t *__restrict__ drop,
                                                     uint32_t ndrop) {
    for (uint32_t j = blockIdx.x * blockDim.x + threadIdx.x; j < ndrop; j += gridDim.x * blockDim.x)
        keep[drop[j]] = 0u;
}

// The kept vertices (and interior points) to their place in the mesh; every
// kept vertex's index in the mesh.
__global__ __launch_bounds__(256) void k_stitch_verts(const double *__restrict__ v, const double *__restrict__ in,
                                                      uint32_t nv, const uint32_t *__restrict__ keep,
                                                      const uint32_t *__restrict__ rank, uint32_t vbase,
                                                      double *__restrict__ fv, double *__restrict__ fin,
                                                      uint32_t *__restrict__ newidx) {
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < nv; i += gridDim.x * blockDim.x) {
        if (!keep[i]) continue;
        const size_t o = (size_t)(vbase + rank[i]);
        for (int c = 0; c < 3; c++) fv[3 * o + c] = v[3 * (size_t)i + c];
        if (in)
            for (int c = 0; c < 3; c++) fin[3 * o + c] = in[3 * (size_t)i + c];
        newidx[i] = vbase + rank[i];
    }
}

// The dropped vertices to the lower slab's copies (both lists sorted by
// edge); a pair of different edges flags the mismatch.
__global__ __launch_bounds__(256) void k_stitch_pairs(const uint64_t *__restrict__ lo_keys,
                                                      const uint32_t *__restrict__ lo_global,
                                                      const uint64_t *__restrict__ up_keys,
                                                      const uint32_t *__restrict__ up_idx, uint32_t m,
                                                      uint32_t *__restrict__ newidx, uint32_t *__restrict__ bad) {
    for (uint32_t j = blockIdx.x * blockDim.x + threadIdx.x; j < m; j += gridDim.x * blockDim.x) {
        if (lo_keys[j] != up_keys[j]) atomicOr(bad, 1u);
        newidx[up_idx[j]] = lo_global[j];
    }
}

__global__ __launch_bounds__(256) void k_stitch_tris(const uint32_t *__restrict__ t, uint64_t n3,
                                                     const uint32_t *__restrict__ newidx, uint32_t *__restrict__ ft) {
    for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n3; i += (uint64_t)gridDim.x * blockDim.x)
        ft[i] = newidx[t[i]];
}

__global__ __launch_bounds__(256) void k_stitch_globals(const uint32_t *__restrict__ idx, uint32_t m,
                                                        const uint32_t *__restrict__ newidx,
                                                        uint32_t *__restrict__ global) {
    for (uint32_t j = blockIdx.x * blockDim.x + threadIdx.x; j < m; j += gridDim.x * blockDim.x)
        global[j] = newidx[idx[j]];
}

}  // namespace

hipError_t launch_plane_list(const uint64_t *vkeys, uint32_t nv, uint32_t z, uint32_t *flags, uint32_t *scan,
                             void *tmp, size_t tmp_bytes, uint64_t *keys, uint32_t *idx, uint32_t *count,
                             hipStream_t st) {
    if (!nv) return hipMemsetAsync(count, 0, 4, st);
    k_plane_flags<<<grid_for(nv, 256, 8192), 256, 0, st>>>(vkeys, nv, z, flags);
    hipError_t e = grp_scan(tmp, tmp_bytes, flags, scan, nv, st);
    if (e != hipSuccess) return e;
    k_plane_compact<<<grid_for(nv, 256, 8192), 256, 0, st>>>(vkeys, nv, z, scan, keys, idx, count);
    return hipGetLastError();
}

hipError_t launch_stitch_keep(uint32_t *keep, uint32_t nv, const uint32_t *drop, uint32_t ndrop, hipStream_t st) {
    if (nv) k_stitch_keep<<<grid_for(nv, 256, 8192), 256, 0, st>>>(keep, nv);
    if (ndrop) k_stitch_drop<<<grid_for(ndrop, 256, 8192), 256, 0, st>>>(keep, drop, ndrop);
    return hipGetLastError();
}

hipError_t launch_stitch_verts(const double *v, const double *in, uint32_t nv, const uint32_t *keep,
                               const uint32_t *rank, uint32_t vbase, double *fv, double *fin, uint32_t *newidx,
                               hipStream_t st) {
    if (nv) k_stitch_verts<<<grid_for(nv, 256, 8192), 256, 0, st>>>(v, in, nv, keep, rank, vbase, fv, fin, newidx);
    return hipGetLastError();
}

hipError_t launch_stitch_pairs(const uint64_t *lo_keys, const uint32_t *lo_global, const uint64_t *up_keys,
                               const uint32_t *up_idx, uint32_t m, uint32_t *newidx, uint32_t *bad, hipStream_t st) {
    if (m) k_stitch_pairs<<<grid_for(m, 256, 8192), 256, 0, st>>>(lo_keys, lo_global, up_keys, up_idx, m, newidx, bad);
    return hipGetLastError();
}

hipError_t launch_stitch_tris(const uint32_t *t, uint64_t nt, const uint32_t *newidx, uint32_t *ft, hipStream_t st) {
    if (nt) k_stitch_tris<<<grid_for(3 * nt, 256, 8192), 256, 0, st>>>(t, 3 * nt, newidx, ft);
    return hipGetLastError();
}

hipError_t launch_stitch_globals(const uint32_t *idx, uint32_t m, const uint32_t *newidx, uint32_t *global,
                                 hipStream_t st) {
    if (m) k_stitch_globals<<<grid_for(m, 256, 8192), 256, 0, st>>>(idx, m, newidx, global);
    return hipGetLastError();
}

}  // namespace mc3d
