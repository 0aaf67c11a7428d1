// k_corners.hip -- launcher of the corner kernels (Solid.Contains at every
// grid corner into the occupancy grid).  Built once per precision and
// statistics mode: the Makefile defines CORNERS_MODE (0 float64, 1 float32 +
// guard) and CORNERS_STATS (0 or 1).
#include "kernels.h"
#include "launch.h"

#ifndef CORNERS_MODE
#error "CORNERS_MODE must be defined"
#endif
#ifndef CORNERS_STATS
#error "CORNERS_STATS must be defined"
#endif

namespace mc3d {
namespace {

template <int MODE, bool STATS, int NP, bool MULTI>
void corners_act(const Pipe &p, const Bufs &b, unsigned grid, hipStream_t st) {
    if (p.wpb == 1)
        k_corners_act<MODE, STATS, NP, 1, MULTI>
            <<<grid, 64, 0, st>>>(p.segs, p.nsegs, p.seg0, p.brick_seg, b.actb, b.bits, p.ctl, p.sout);
    else if (p.wpb == 2)
        k_corners_act<MODE, STATS, NP, 2, MULTI>
            <<<grid, 128, 0, st>>>(p.segs, p.nsegs, p.seg0, p.brick_seg, b.actb, b.bits, p.ctl, p.sout);
    else
        k_corners_act<MODE, STATS, NP, 4, MULTI>
            <<<grid, 256, 0, st>>>(p.segs, p.nsegs, p.seg0, p.brick_seg, b.actb, b.bits, p.ctl, p.sout);
}

template <int MODE, bool STATS, bool MULTI>
void corners_m(const Pipe &p, const Bufs &b, hipStream_t st) {
    if (!p.cull) {  // about four words per wave (fewer wave launches for a cheap program)
        k_corners<MODE, STATS, MULTI><<<grid_for(p.nwords, 16), 256, 0, st>>>(p.segs, p.nsegs, p.seg0, p.word_seg,
                                                                              b.bits, p.ctl, p.sout, p.nwords);
        return;
    }
    // (the words of constant bricks were written by the plan)
    const unsigned grid = grid_for(p.nbricks, 1, 1u << 15);
    if (p.rows_per_walk >= 4)
        corners_act<MODE, STATS, 4, MULTI>(p, b, grid, st);
    else if (p.rows_per_walk == 2)
        corners_act<MODE, STATS, 2, MULTI>(p, b, grid, st);
    else
        corners_act<MODE, STATS, 1, MULTI>(p, b, grid, st);
}

}  // namespace

template <>
hipError_t launch_corners_ms<CORNERS_MODE, (CORNERS_STATS != 0)>(const Pipe &p, const Bufs &b, hipStream_t st) {
    if (p.nsegs > 1)
        corners_m<CORNERS_MODE, (CORNERS_STATS != 0), true>(p, b, st);
    else
        corners_m<CORNERS_MODE, (CORNERS_STATS != 0), false>(p, b, st);
    return hipGetLastError();
}

}  // namespace mc3d
