// k_bisect.hip -- launchers of the per-vertex kernels: the bisection
// (mcSearchPoint), its host-answered form, Contains at arbitrary points and
// the points host-evaluated atoms are asked about.
#include "kernels.h"
#include "launch.h"

namespace mc3d {

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

namespace {

template <int MODE, bool STATS>
void bisect(const Pipe &p, const Bufs &b, double *verts, double *interior, uint64_t vcap, hipStream_t st) {
    const unsigned grid = grid_for(vcap, 256, 1u << 14);
    if (p.nsegs > 1)
        k_bisect<MODE, STATS, true><<<grid, 256, 0, st>>>(p.segs, p.nsegs, p.seg0, b.vseg, b.bits, b.vkeys, verts,
                                                          interior, p.ctl, vcap);
    else
        k_bisect<MODE, STATS, false><<<grid, 256, 0, st>>>(p.segs, p.nsegs, p.seg0, nullptr, b.bits, b.vkeys, verts,
                                                           interior, p.ctl, vcap);
}

}  // namespace

hipError_t launch_bisect(const Pipe &p, const Bufs &b, double *verts, double *interior, uint64_t vcap,
                         hipStream_t st) {
    if (p.mode == 0)
        p.stats ? bisect<0, true>(p, b, verts, interior, vcap, st) : bisect<0, false>(p, b, verts, interior, vcap, st);
    else
        p.stats ? bisect<1, true>(p, b, verts, interior, vcap, st) : bisect<1, false>(p, b, verts, interior, vcap, st);
    return hipGetLastError();
}

hipError_t launch_bisect_host(int mode, const DevProg &P, const Grid &g, const uint64_t *bits,
                              const uint64_t *vkeys, double *state, double *probe, double *verts, double *interior,
                              uint64_t nv, int phase, hipStream_t st) {
    const unsigned grid = grid_for(nv, 256, 1u << 14);
    if (mode == 1)
        k_bisect_host<1><<<grid, 256, 0, st>>>(P, g, bits, vkeys, state, probe, verts, interior, nv, phase);
    else
        k_bisect_host<0><<<grid, 256, 0, st>>>(P, g, bits, vkeys, state, probe, verts, interior, nv, phase);
    return hipGetLastError();
}

hipError_t launch_host_points(const DevProg &P, const Grid &g, const double *pts, uint32_t frame, uint64_t start,
                              uint64_t n, double *out, hipStream_t st) {
    k_host_points<<<grid_for(n, 256), 256, 0, st>>>(P, g, pts, frame, start, n, out);
    return hipGetLastError();
}

hipError_t launch_contains(int mode, const DevProg &P, const double *pts, uint8_t *out, uint64_t n, hipStream_t st) {
    if (mode == 1)
        k_contains<1><<<grid_for(n, 256), 256, 0, st>>>(P, pts, out, n);
    else
        k_contains<0><<<grid_for(n, 256), 256, 0, st>>>(P, pts, out, n);
    return hipGetLastError();
}

}  // namespace mc3d
