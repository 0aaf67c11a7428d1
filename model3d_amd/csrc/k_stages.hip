// k_stages.hip -- launchers of the grid and compaction stages: the brick
// plan, classify and its scan, emit, the per-solid output ranges, the
// MarchingCubesConj vertex map and the STL encoder.
#include "kernels.h"
#include "launch.h"

namespace mc3d {

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
// listed at their scanned .z.  FUSED (few classify blocks): the sums are not
// scanned beforehand -- each tile adds up the sums before it itself, and the
// last tile writes the totals.
template <bool FUSED>
__global__ __launch_bounds__(SCAN_BLOCK) void k_scan_down(const uint32_t *__restrict__ in, uint64_t n,
                                                          const uint4 *__restrict__ part_sums,
                                                          uint2 *__restrict__ out, uint32_t *__restrict__ list,
                                                          const Seg seg0, const Seg *__restrict__ segs,
                                                          const uint32_t *__restrict__ pos_seg, bool bm,
                                                          unsigned long long *__restrict__ totals) {
    static_assert(SCAN_ITEMS == 8, "a thread loads its counts as two 16-byte vectors");
    __shared__ uint4 sh[SCAN_BLOCK / 64];
    const uint64_t base = (uint64_t)blockIdx.x * SCAN_TILE + (uint64_t)threadIdx.x * SCAN_ITEMS;
    // a thread's counts: two 16-byte loads (the buffers are 256-byte aligned)
    // unless its items run past n
    const bool whole = base + SCAN_ITEMS <= n;
    uint32_t c[SCAN_ITEMS];
    if (whole) {
        const uint4 c0 = ((const uint4 *)(in + base))[0], c1 = ((const uint4 *)(in + base))[1];
        c[0] = c0.x, c[1] = c0.y, c[2] = c0.z, c[3] = c0.w, c[4] = c1.x, c[5] = c1.y, c[6] = c1.z, c[7] = c1.w;
    } else {
        for (int i = 0; i < SCAN_ITEMS; i++) c[i] = base + i < n ? in[base + i] : 0u;
    }
    uint4 vals[SCAN_ITEMS];
    uint4 s = make_uint4(0, 0, 0, 0);
    for (int i = 0; i < SCAN_ITEMS; i++) {
        vals[i] = make_uint4(c[i] & 0xffu, (c[i] >> 8) & 0x1ffu, c[i] >> 17, 0u);
        s = add4(s, vals[i]);
    }
    uint4 off;
    const uint32_t first_part = blockIdx.x * (SCAN_TILE / CLASSIFY_BLOCK);
    if (FUSED) {
        // the block sums before this tile, four loads in flight per thread
        uint4 a = make_uint4(0, 0, 0, 0);
        for (uint32_t i0 = threadIdx.x; i0 < first_part; i0 += 4 * SCAN_BLOCK) {
            uint4 v[4];
            for (int k = 0; k < 4; k++) {
                const uint32_t i = i0 + (uint32_t)k * SCAN_BLOCK;
                v[k] = i < first_part ? part_sums[i] : make_uint4(0, 0, 0, 0);
            }
            for (int k = 0; k < 4; k++) a = add4(a, v[k]);
        }
        off = block_sum<SCAN_BLOCK>(a, sh);
        __syncthreads();  // sh is reused by the scan below
    } else {
        off = part_sums[first_part];
    }
    uint4 total;
    uint4 ex = add4(block_exclusive(s, sh, total), off);
    if (FUSED && blockIdx.x == gridDim.x - 1 && threadIdx.x == 0) {
        const uint4 all = add4(off, total);
        totals[0] = all.x, totals[1] = all.y, totals[2] = all.z;
    }
    if (whole) {  // the offsets as four 16-byte stores
        uint4 *q = (uint4 *)(out + base);
        uint4 e = ex;
        for (int i = 0; i < SCAN_ITEMS; i += 2) {
            const uint4 e1 = add4(e, vals[i]);
            q[i / 2] = make_uint4(e.x, e.y, e1.x, e1.y);
            e = add4(e1, vals[i + 1]);
        }
    }
    for (int i = 0; i < SCAN_ITEMS; i++) {
        if (base + i < n) {
            if (!whole) out[base + i] = make_uint2(ex.x, ex.y);
            if (list && vals[i].z) {
                uint32_t w = (uint32_t)(base + i);
                if (bm && pos_seg) {  // brick-major position -> word of its solid
                    const Seg &sg = segs[pos_seg[w >> SEG_WORD_SHIFT]];
                    int wx, y, z;
                    bm_word(sg.g, sg.B, (uint32_t)(w - sg.pos0), wx, y, z);
                    w = (uint32_t)(sg.word0 + word_index(sg.g, wx, y, z));
                } else if (bm) {  // (one solid: position 0 is its first)
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

// Each solid's first vertex and triangle in the call's outputs.
__global__ __launch_bounds__(256) void k_seg_ranges(const Seg *__restrict__ segs, uint32_t nsegs,
                                                    const uint2 *__restrict__ offs, SegOut *__restrict__ sout) {
    for (uint32_t s = blockIdx.x * blockDim.x + threadIdx.x; s < nsegs; s += gridDim.x * blockDim.x) {
        const uint2 o = offs[segs[s].pos0];
        sout[s].vbase = o.x;
        sout[s].tbase = o.y;
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


// Streams of known size for the calibration of the HBM traffic counters
// (and a measure of the achievable bandwidth): k_stream_read reads n 16-byte
// items once and writes one word per block; k_stream_write writes n items.
__global__ __launch_bounds__(256) void k_stream_read(const uint4 *__restrict__ in, uint64_t n,
                                                     uint32_t *__restrict__ out) {
    uint32_t acc = 0;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x) {
        const uint4 v = in[i];
        acc ^= v.x ^ v.y ^ v.z ^ v.w;
    }
    for (int d = 32; d; d >>= 1) acc ^= __shfl_xor(acc, d);
    __shared__ uint32_t red[4];
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = acc;
    __syncthreads();
    if (threadIdx.x == 0) out[blockIdx.x] = red[0] ^ red[1] ^ red[2] ^ red[3];
}

__global__ __launch_bounds__(256) void k_stream_write(uint4 *__restrict__ out, uint64_t n, uint32_t v) {
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x)
        out[i] = make_uint4(v, v + 1, v + 2, (uint32_t)i);
}

// Zeroes `bytes` (a multiple of 16) with plain stores: a fill kernel, not a
// memset, so that it may sit in a captured graph.
__global__ __launch_bounds__(256) void k_zero(uint4 *__restrict__ p, uint64_t n) {
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x)
        p[i] = make_uint4(0, 0, 0, 0);
}

hipError_t launch_zero(void *p, size_t bytes, hipStream_t st) {
    const uint64_t n = bytes / 16;
    if (n) k_zero<<<grid_for(n, 256, 4096), 256, 0, st>>>((uint4 *)p, n);
    return hipGetLastError();
}

hipError_t launch_stream(bool write, void *buf, uint64_t items, uint32_t *out, hipStream_t st) {
    const unsigned grid = 8192;
    if (write)
        k_stream_write<<<grid, 256, 0, st>>>((uint4 *)buf, items, 7u);
    else
        k_stream_read<<<grid, 256, 0, st>>>((const uint4 *)buf, items, out);
    return hipGetLastError();
}

hipError_t launch_plan(const Pipe &p, const Bufs &b, hipStream_t st) {
    // one wave per unit of bricks (plan_units; past 2^17 units a grid-stride
    // loop takes the rest): tiles at the grid's edges are partial, so a
    // one-solid call has more units than nbricks / gsz
    uint64_t units = (p.nbricks + p.plan_gsz - 1) / p.plan_gsz;
    if (p.nsegs == 1 && p.plan_tile) {
        const uint32_t ty = 1u << ((31 - __builtin_clz(p.plan_gsz)) / 2), tz = p.plan_gsz / ty;
        units = (uint64_t)p.seg0.g.W * ((p.seg0.B.nby + ty - 1) / ty) * ((p.seg0.B.nbz + tz - 1) / tz);
    }
    const unsigned pgrid = grid_for(units, PLAN_BLOCK / 64, 1u << 16);
    if (p.nsegs > 1)
        k_plan<true><<<pgrid, PLAN_BLOCK, p.plan_lds, st>>>(p.segs, p.nsegs, p.seg0, p.brick_seg, p.nbricks, p.cap,
                                                            b.binfo, b.bcode, b.actb, b.bits, p.ctl, p.sout,
                                                            p.plan_gsz, false);
    else
        k_plan<false><<<pgrid, PLAN_BLOCK, p.plan_lds, st>>>(p.segs, p.nsegs, p.seg0, p.brick_seg, p.nbricks, p.cap,
                                                             b.binfo, b.bcode, b.actb, b.bits, p.ctl, p.sout,
                                                             p.plan_gsz, p.plan_tile);
    return hipGetLastError();
}

hipError_t launch_classify(const Pipe &p, const Bufs &b, hipStream_t st) {
    const unsigned ctiles = (unsigned)((p.nscan + CLASSIFY_BLOCK - 1) / CLASSIFY_BLOCK);
    if (p.scan_one_pass) {  // small grids: the scan too
        if (p.nsegs > 1)
            k_classify_scan<true><<<ctiles, CLASSIFY_BLOCK, 0, st>>>(p.segs, p.nsegs, p.seg0, p.pos_seg, b.bits,
                                                                     b.table, nullptr, p.nscan, p.bm, b.agg, b.offs,
                                                                     b.actw, p.ctl);
        else
            k_classify_scan<false><<<ctiles, CLASSIFY_BLOCK, 0, st>>>(p.segs, p.nsegs, p.seg0, p.word_seg, b.bits,
                                                                      b.table, b.keep, p.nscan, p.bm, b.agg, b.offs,
                                                                      b.actw, p.ctl);
    } else if (p.nsegs > 1) {
        if (p.bm && p.nscan >= CLASSIFY_SHFL_MIN)
            k_classify<true, true><<<ctiles, CLASSIFY_BLOCK, 0, st>>>(p.segs, p.nsegs, p.seg0, p.pos_seg, b.bits,
                                                                      b.table, b.counts, b.tiles, nullptr, p.nscan,
                                                                      p.bm);
        else
            k_classify<true><<<ctiles, CLASSIFY_BLOCK, 0, st>>>(p.segs, p.nsegs, p.seg0, p.pos_seg, b.bits,
                                                                b.table, b.counts, b.tiles, nullptr, p.nscan, p.bm);
    } else if (p.bm && p.nscan >= CLASSIFY_SHFL_MIN) {  // large brick-major grids: +y / +z words by shuffles
        k_classify<false, true><<<ctiles, CLASSIFY_BLOCK, 0, st>>>(p.segs, p.nsegs, p.seg0, p.word_seg, b.bits,
                                                                   b.table, b.counts, b.tiles, b.keep, p.nscan, p.bm);
    } else {
        k_classify<false><<<ctiles, CLASSIFY_BLOCK, 0, st>>>(p.segs, p.nsegs, p.seg0, p.word_seg, b.bits, b.table,
                                                             b.counts, b.tiles, b.keep, p.nscan, p.bm);
    }
    return hipGetLastError();
}

hipError_t launch_scan(const Pipe &p, const Bufs &b, hipStream_t st) {
    // the scan of the counts whose per-block sums classify wrote (done by
    // classify itself on small grids)
    const uint64_t ntiles = (p.nscan + SCAN_TILE - 1) / SCAN_TILE;
    const uint64_t nparts = (p.nscan + CLASSIFY_BLOCK - 1) / CLASSIFY_BLOCK;
    const uint32_t *multi_pos = p.nsegs > 1 ? p.pos_seg : nullptr;  // positions -> solids
    if (p.scan_one_pass) {
    } else if (nparts <= SCAN_FUSE_PARTS) {  // each tile sums the block sums before it
        k_scan_down<true><<<(unsigned)ntiles, SCAN_BLOCK, 0, st>>>(b.counts, p.nscan, b.tiles, b.offs, b.actw, p.seg0,
                                                                  p.segs, multi_pos, p.bm, p.ctl->totals);
    } else {
        k_scan_tiles<<<1, SUMS_BLOCK, 0, st>>>(b.tiles, nparts, p.ctl->totals);
        k_scan_down<false><<<(unsigned)ntiles, SCAN_BLOCK, 0, st>>>(b.counts, p.nscan, b.tiles, b.offs, b.actw,
                                                                   p.seg0, p.segs, multi_pos, p.bm, nullptr);
    }
    if (p.nsegs > 1)  // a single solid starts at vertex 0 and triangle 0 (zeroed by the upload)
        k_seg_ranges<<<grid_for(p.nsegs, 256), 256, 0, st>>>(p.segs, p.nsegs, b.offs, p.sout);
    return hipGetLastError();
}

hipError_t launch_emit(const Pipe &p, const Bufs &b, uint32_t *tris, uint64_t vcap, uint64_t tcap, hipStream_t st) {
    // a wave takes EMIT_G listed words at a time: 512 to 4096 blocks of four
    // waves, from the grid's words (each block also stages the case table)
    const unsigned grid = std::max(512u, grid_for(p.nwords, p.emit_div, 4096));
    if (p.nsegs > 1)
        k_emit<true><<<grid, 256, 0, st>>>(p.segs, p.nsegs, p.seg0, p.word_seg, b.bits, b.etable, b.offs, b.actw,
                                           p.ctl, b.vkeys, b.vseg, tris, nullptr, vcap, tcap, p.bm);
    else
        k_emit<false><<<grid, 256, 0, st>>>(p.segs, p.nsegs, p.seg0, p.word_seg, b.bits, b.etable, b.offs, b.actw,
                                            p.ctl, b.vkeys, nullptr, tris, b.keep, vcap, tcap, p.bm);
    return hipGetLastError();
}

hipError_t launch_inverse_apply(const double *recs, uint32_t nops, double *verts, const Ctl *ctl, uint64_t vcap,
                                hipStream_t st) {
    k_inverse_apply<<<grid_for(vcap, 256, 1u << 14), 256, 0, st>>>(recs, nops, verts, ctl, vcap);
    return hipGetLastError();
}

hipError_t launch_encode_stl(const double *verts, const uint32_t *tris, uint64_t nt, uint16_t *body,
                             hipStream_t st) {
    k_encode_stl<<<grid_for(nt, 256, 1u << 16), 256, 0, st>>>(verts, tris, nt, body);
    return hipGetLastError();
}

}  // namespace mc3d
