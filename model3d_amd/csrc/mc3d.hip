// mc3d.hip -- host side of the MI355X marching-cubes pipeline and the C-ABI
// entry points for meshing (include/mc3d.h).  The device stages live in
// kernels.h, Contains in eval.h, the solid tree and its lowering in solid.cpp.
//
// A call meshes a list of solids (one for mc3d_marching_cubes_search, many
// for mc3d_marching_cubes_search_batch): one pinned upload of every lowered
// program, the grid axes and a zeroed control block, then one launch per
// stage for all solids on the context's stream and one synchronisation at
// the end; every count a stage needs is read on the device.  Output buffers
// are pooled per context and sized from earlier calls; a call whose meshes
// do not fit grows them and re-runs the two output stages.
//
// Every float64 expression keeps the reference's operation order and is
// compiled with -ffp-contract=off, so Contains decisions, topology and vertex
// coordinates are bit-identical to the reference's amd64 build.

#include <hip/hip_runtime.h>
#include <math.h>
#include <stdint.h>
#include <string.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <memory>
#include <new>
#include <string>
#include <thread>
#include <vector>

#include "kernels.h"
#include "mc3d_internal.h"

using namespace mc3d;

namespace {

#define HIP_TRY(expr)                                                                        \
    do {                                                                                     \
        hipError_t e_ = (expr);                                                              \
        if (e_ != hipSuccess) {                                                              \
            set_error(std::string(#expr) + ": " + hipGetErrorString(e_));                    \
            return MC3D_E_HIP;                                                               \
        }                                                                                    \
    } while (0)

// Grow-only device buffer.
struct DevBuf {
    void *p = nullptr;
    size_t cap = 0;
    DevBuf() = default;
    DevBuf(const DevBuf &) = delete;
    DevBuf &operator=(const DevBuf &) = delete;
    ~DevBuf() {
        if (p) (void)hipFree(p);
    }
    int ensure(size_t bytes) {
        if (bytes <= cap) return MC3D_OK;
        if (p) (void)hipFree(p);
        p = nullptr;
        cap = 0;
        const size_t want = bytes + bytes / 4 + 4096;
        hipError_t e = hipMalloc(&p, want);
        if (e != hipSuccess) {
            p = nullptr;
            set_error(std::string("hipMalloc: ") + hipGetErrorString(e));
            return MC3D_E_OOM;
        }
        cap = want;
        return MC3D_OK;
    }
    template <class T>
    T *as() const {
        return (T *)p;
    }
};

// Pinned host staging buffer (grow-only).
struct HostBuf {
    void *p = nullptr;
    size_t cap = 0;
    ~HostBuf() {
        if (p) (void)hipHostFree(p);
    }
    int ensure(size_t bytes) {
        if (bytes <= cap) return MC3D_OK;
        if (p) (void)hipHostFree(p);
        p = nullptr;
        cap = 0;
        const size_t want = bytes + bytes / 4 + 4096;
        hipError_t e = hipHostMalloc(&p, want, hipHostMallocDefault);
        if (e != hipSuccess) {
            p = nullptr;
            set_error(std::string("hipHostMalloc: ") + hipGetErrorString(e));
            return MC3D_E_OOM;
        }
        cap = want;
        return MC3D_OK;
    }
};

// Device buffers released by freed meshes, reused by later calls.
struct Pool {
    std::vector<std::unique_ptr<DevBuf>> free;
    std::unique_ptr<DevBuf> take(size_t bytes) {
        size_t best = free.size();
        for (size_t i = 0; i < free.size(); i++)
            if (free[i]->cap >= bytes && (best == free.size() || free[i]->cap < free[best]->cap)) best = i;
        if (best < free.size()) {
            std::unique_ptr<DevBuf> b = std::move(free[best]);
            free.erase(free.begin() + best);
            return b;
        }
        return std::unique_ptr<DevBuf>(new DevBuf);
    }
    void give(std::unique_ptr<DevBuf> b) {
        if (b && b->p) free.push_back(std::move(b));
        while (free.size() > 8) free.erase(free.begin());
    }
};

// memcpy of large host buffers split over up to 16 threads (the first touch
// of freshly allocated pages dominates a single-threaded copy).
void parallel_copy(void *dst, const void *src, size_t bytes) {
    constexpr size_t CHUNK = 8u << 20;
    const size_t nt = std::min<size_t>(16, (bytes + CHUNK - 1) / CHUNK);
    if (nt <= 1) {
        if (bytes) memcpy(dst, src, bytes);
        return;
    }
    const size_t per = (bytes + nt - 1) / nt;
    std::vector<std::thread> th;
    for (size_t i = 0; i < nt; i++) {
        const size_t a = i * per, b = std::min(bytes, a + per);
        if (a < b) th.emplace_back([=] { memcpy((char *)dst + a, (const char *)src + a, b - a); });
    }
    for (std::thread &t : th) t.join();
}

unsigned grid_for(uint64_t n, unsigned block, unsigned cap = 1u << 20) {
    uint64_t b = (n + block - 1) / block;
    if (b == 0) b = 1;
    return (unsigned)(b < cap ? b : cap);
}

size_t align_up(size_t v, size_t a) { return (v + a - 1) / a * a; }

// The device outputs of one call (vertices, triangles), shared by the meshes
// it returned and given back to the context's pool with the last of them.
struct MeshBuffers {
    std::shared_ptr<Pool> pool;
    std::unique_ptr<DevBuf> verts, tris, interior;  // interior: MarchingCubesInterior only
    ~MeshBuffers() {
        if (pool) {
            pool->give(std::move(verts));
            pool->give(std::move(tris));
            pool->give(std::move(interior));
        }
    }
};

}  // namespace

// A collider's ray grid on the device, kept per context (keyed by the
// collider's id; the shared_ptr keeps the id from being reused).
struct ColliderCache {
    std::shared_ptr<const ColliderData> keep;
    DevBuf rec, start, tris;
    DevCollider dev{};
};

struct mc3d_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    std::vector<std::unique_ptr<ColliderCache>> colliders;  // most recent last
    HostBuf staging, ctl_host, segstage, outstage;
    DevBuf blob, table, etable, bits, counts, offs, tiles, actw, actb, binfo, bcode, vkeys, vseg, pts, out8;
    DevBuf hvals, hpts, hstate, hprobe;  // host-evaluated atoms: answers, points, bisection state
    HostBuf hstage;                      // pinned points / answers of one callback batch
    DevBuf stl, stl_in;                  // STL encoding: body bytes; uploaded host triangles
    DevBuf keep;                         // block filter: cube mask
    std::shared_ptr<Pool> pool = std::make_shared<Pool>();
    uint64_t v_hint = 1u << 18, t_hint = 1u << 19;  // output capacities to start from
    hipEvent_t ev[10] = {};
};

// One solid's mesh: a slice [voff, voff + nv) / [toff, toff + nt) of its
// call's outputs (triangle indices count from the slice's first vertex).
struct mc3d_mesh {
    int device = 0;
    uint64_t nv = 0, nt = 0, voff = 0, toff = 0;
    std::unique_ptr<double[]> hverts;  // on_host: 3 nv (allocated without initialisation)
    std::unique_ptr<uint32_t[]> htris;  // on_host: 3 nt
    std::unique_ptr<double[]> hinterior;  // on_host with interior points: 3 nv
    bool on_host = false, has_interior = false;
    std::shared_ptr<MeshBuffers> buf;
    mc3d_stats stats{};
};

namespace {

// One solid of a call, prepared on the host.
struct Job {
    const mc3d_solid *s = nullptr;
    double delta = 0;
    int iters = 0;
    Program P;
    std::vector<double> ax[3];
    Grid g{};    // sizes; axis pointers set by the upload
    Bricks B{};  // shape; array pointers set by the upload
    uint64_t nwords = 0, nbricks = 0, word0 = 0, brick0 = 0;
    uint32_t cap = 0;  // folded-program slot of its culling bricks (0: no culling)
    bool host = false;
    std::vector<uint32_t> frames;  // frame of each host-evaluated atom
    bool interior = false;         // MarchingCubesInterior: keep each vertex's inside point
    std::vector<double> conj;      // MarchingCubesConj: inverse chain applied to the vertices
    std::vector<uint64_t> keep;    // MarchingCubesFilter: cubes of accepted blocks (occupancy layout)
};

// Frame of each host-evaluated atom, by host atom index.
std::vector<uint32_t> host_frames(const Program &P) {
    std::vector<uint32_t> fr(P.host_atoms.size(), 0);
    for (uint32_t pc : P.atom_pc) {
        const Insn in = P.code[pc];
        if (insn_kind(in.x) == A_HOST) fr[(size_t)P.f64[in.y]] = insn_frame(in.x);
    }
    return fr;
}

std::string fmt_bounds(const double *mn, const double *mx) {
    char buf[512];
    snprintf(buf, sizeof buf,
             "invalid bounds: min=model3d.Coord3D{X:%.17g, Y:%.17g, Z:%.17g} max=model3d.Coord3D{X:%.17g, Y:%.17g, "
             "Z:%.17g}",
             mn[0], mn[1], mn[2], mx[0], mx[1], mx[2]);
    return buf;
}

// validateBounds (mc.go:15, bounder.go:48-52), the spacer (mc.go:594-608),
// the lowering, and the culling bricks: 64 x B x B corners, B growing until
// the fixed program slots fit.
int prepare_job(Job &j, const mc3d_options &o) {
    double mn[3], mx[3];
    solid_bounds(*j.s->ptr, mn, mx);
    if (!bounds_valid(*j.s->ptr)) {
        set_error(fmt_bounds(mn, mx));
        return MC3D_E_INVALID_BOUNDS;
    }
    if (!(j.delta > 0)) {
        set_error("delta must be positive");
        return MC3D_E_INVALID_ARG;
    }
    if (j.iters < 0) j.iters = 0;
    for (int a = 0; a < 3; a++) {
        j.ax[a] = spacer_axis(mn[a], mx[a], j.delta);
        if (j.ax[a].size() >= (1u << 20)) {
            set_error("grid too large: more than 2^20 samples on an axis");
            return MC3D_E_UNSUPPORTED;
        }
        if (j.ax[a].size() < 2) {
            set_error("grid has fewer than two samples on an axis");
            return MC3D_E_INVALID_ARG;
        }
    }
    std::string msg;
    if (!lower_solid(*j.s->ptr, j.P, msg)) {
        set_error(msg);
        return MC3D_E_UNSUPPORTED;
    }
    j.host = !j.P.host_atoms.empty();
    j.frames = host_frames(j.P);
    Grid &g = j.g;
    g.nx = (int)j.ax[0].size(), g.ny = (int)j.ax[1].size(), g.nz = (int)j.ax[2].size();
    g.W = (g.nx + 63) / 64;
    j.nwords = (uint64_t)g.W * g.ny * g.nz;
    if (j.nwords >= (1ull << 32)) {
        set_error("grid too large: more than 2^32 occupancy words");
        return MC3D_E_UNSUPPORTED;
    }
    j.B = Bricks{nullptr, nullptr, 1, 1, 1, 1};
    // A host-evaluated solid may be true outside its bounds, so nothing is
    // folded by bounds: no culling.
    if (o.cull && !j.host) {
        int bsz = o.brick > 0 ? o.brick : 4;
        j.cap = std::min<uint32_t>((uint32_t)j.P.code.size(), 64u);
        for (;;) {
            j.B.BY = j.B.BZ = bsz;
            j.B.nby = (g.ny + bsz - 1) / bsz;
            j.B.nbz = (g.nz + bsz - 1) / bsz;
            j.nbricks = (uint64_t)g.W * j.B.nby * j.B.nbz;
            if (j.nbricks * j.cap < (1ull << 27) || bsz >= 64) break;
            bsz *= 2;
        }
    }
    return MC3D_OK;
}

// The device copy of a collider's ray grid (uploaded once per context).
int collider_on_device(mc3d_ctx *ctx, const std::shared_ptr<const ColliderData> &cd, DevCollider &out,
                       size_t keep_at_least) {
    for (size_t i = 0; i < ctx->colliders.size(); i++) {
        if (ctx->colliders[i]->keep->id == cd->id) {
            out = ctx->colliders[i]->dev;
            std::unique_ptr<ColliderCache> hit = std::move(ctx->colliders[i]);  // now the most recent
            ctx->colliders.erase(ctx->colliders.begin() + i);
            ctx->colliders.push_back(std::move(hit));
            return MC3D_OK;
        }
    }
    const RayGrid &G = cd->ray_grid();
    std::unique_ptr<ColliderCache> c(new ColliderCache);
    c->keep = cd;
    int rc;
    if ((rc = c->rec.ensure(G.rec.size() * sizeof(double) + 8)) ||
        (rc = c->start.ensure(G.cell_start.size() * sizeof(uint32_t) + 8)) ||
        (rc = c->tris.ensure(G.cell_tris.size() * sizeof(uint32_t) + 8)))
        return rc;
    if (!G.rec.empty()) HIP_TRY(hipMemcpy(c->rec.p, G.rec.data(), G.rec.size() * sizeof(double), hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(c->start.p, G.cell_start.data(), G.cell_start.size() * sizeof(uint32_t), hipMemcpyHostToDevice));
    if (!G.cell_tris.empty())
        HIP_TRY(hipMemcpy(c->tris.p, G.cell_tris.data(), G.cell_tris.size() * sizeof(uint32_t), hipMemcpyHostToDevice));
    DevCollider &d = c->dev;
    d.rec = c->rec.as<double>();
    d.cell_start = c->start.as<uint32_t>();
    d.cell_tris = c->tris.as<uint32_t>();
    memcpy(d.u, G.u, sizeof d.u);
    memcpy(d.v, G.v, sizeof d.v);
    d.u0 = G.u0, d.v0 = G.v0, d.inv_cell = G.inv_cell, d.nu = G.nu, d.nv = G.nv;
    out = d;
    ctx->colliders.push_back(std::move(c));
    // Evict the oldest entries, never one the current call uses (those are
    // the newest `keep_at_least`, looked up or added in this upload).
    while (ctx->colliders.size() > std::max<size_t>(4, keep_at_least)) ctx->colliders.erase(ctx->colliders.begin());
    return MC3D_OK;
}

// Bytes of LDS the plan needs to stage a program: instructions, their atom
// boxes and the float64 params.
size_t plan_stage_bytes(const Program &P) {
    return PLAN_STACK_BYTES + P.code.size() * (sizeof(Insn) + sizeof(AtomBox)) + P.f64.size() * sizeof(double);
}
constexpr size_t PLAN_STAGE_LIMIT = 64 * 1024;

// Where a call's pieces sit in device memory.
struct Upload {
    Ctl *ctl = nullptr;
    SegOut *sout = nullptr;
    Seg *segs = nullptr;
    uint32_t *word_seg = nullptr, *brick_seg = nullptr;
    std::vector<Seg> hsegs;  // host copies of the segments
    size_t back_bytes = 0;   // control block + per-solid results, copied back at the end
    const double *conj = nullptr;  // job 0's MarchingCubesConj inverse chain (records of 10 doubles)
    uint32_t conj_nops = 0;
};

// Packs the control block, the per-solid results, the segments, the segment
// tables and every job's program and grid axes into the pinned staging
// buffer and copies it to the device in one transfer.
int upload(mc3d_ctx *ctx, std::vector<Job> &jobs, uint64_t nwords_all, uint64_t nbricks_all, Upload &U) {
    const size_t n = jobs.size();
    size_t ncol = 0;
    for (const Job &j : jobs) ncol += j.P.colliders.size();
    std::vector<std::vector<DevCollider>> cols(n);
    for (size_t s = 0; s < n; s++) {
        cols[s].resize(jobs[s].P.colliders.size());
        for (size_t i = 0; i < cols[s].size(); i++) {
            int rc = collider_on_device(ctx, jobs[s].P.colliders[i], cols[s][i], ncol);
            if (rc) return rc;
        }
    }
    size_t off = 0;
    auto place = [&](size_t bytes) {
        size_t o = off;
        off = align_up(off + bytes, 256);
        return o;
    };
    // the per-solid results follow the control block directly (one copy back)
    const size_t o_sout = align_up(sizeof(Ctl), 16);
    U.back_bytes = o_sout + n * sizeof(SegOut);
    off = align_up(U.back_bytes, 256);
    const size_t o_segs = place(n * sizeof(Seg));
    const bool tables = n > 1;
    const size_t o_wseg = place(tables ? nwords_all / 64 * sizeof(uint32_t) : 0);
    const size_t o_bseg = place(tables ? nbricks_all / 128 * sizeof(uint32_t) : 0);
    const size_t o_conj = place(jobs[0].conj.size() * sizeof(double));
    struct Off {
        size_t code, f64, f32, fr, box, col, ax;
    };
    std::vector<Off> po(n);
    std::vector<std::vector<AtomBox>> boxes(n);
    for (size_t s = 0; s < n; s++) {
        const Program &P = jobs[s].P;
        boxes[s].resize(P.code.size());
        for (size_t i = 0; i < P.atom_pc.size(); i++) boxes[s][P.atom_pc[i]] = P.atom_box[i];
        po[s].code = place(P.code.size() * sizeof(Insn));
        // params padded so that staging a brick's atoms never reads past the array
        po[s].f64 = place((P.f64.size() + MAX_ATOM_PARAMS) * sizeof(double));
        po[s].f32 = place((P.f32.size() + MAX_ATOM_PARAMS) * sizeof(float));
        po[s].fr = place(P.frames.size() * sizeof(Frame));
        po[s].box = place(boxes[s].size() * sizeof(AtomBox));
        po[s].col = place(cols[s].size() * sizeof(DevCollider));
        po[s].ax = place((jobs[s].ax[0].size() + jobs[s].ax[1].size() + jobs[s].ax[2].size()) * sizeof(double));
    }
    const size_t total = off;
    int rc;
    if ((rc = ctx->staging.ensure(total)) || (rc = ctx->blob.ensure(total))) return rc;
    char *h = (char *)ctx->staging.p;
    char *dv = (char *)ctx->blob.p;
    memset(h, 0, o_segs);  // control block and per-solid results start at zero
    U.hsegs.assign(n, Seg{});
    for (size_t s = 0; s < n; s++) {
        Job &j = jobs[s];
        const Program &P = j.P;
        memcpy(h + po[s].code, P.code.data(), P.code.size() * sizeof(Insn));
        memset(h + po[s].f64, 0, (P.f64.size() + MAX_ATOM_PARAMS) * sizeof(double));
        memset(h + po[s].f32, 0, (P.f32.size() + MAX_ATOM_PARAMS) * sizeof(float));
        if (!P.f64.empty()) {
            memcpy(h + po[s].f64, P.f64.data(), P.f64.size() * sizeof(double));
            memcpy(h + po[s].f32, P.f32.data(), P.f32.size() * sizeof(float));
        }
        memcpy(h + po[s].fr, P.frames.data(), P.frames.size() * sizeof(Frame));
        memcpy(h + po[s].box, boxes[s].data(), boxes[s].size() * sizeof(AtomBox));
        if (!cols[s].empty()) memcpy(h + po[s].col, cols[s].data(), cols[s].size() * sizeof(DevCollider));
        double *d = (double *)(h + po[s].ax);
        for (int a = 0; a < 3; a++) d = std::copy(j.ax[a].begin(), j.ax[a].end(), d);
        Seg &sg = U.hsegs[s];
        sg.P.code = (const Insn *)(dv + po[s].code);
        sg.P.f64 = (const double *)(dv + po[s].f64);
        sg.P.f32 = (const float *)(dv + po[s].f32);
        sg.P.frames = (const Frame *)(dv + po[s].fr);
        sg.P.boxes = (const AtomBox *)(dv + po[s].box);
        sg.P.colliders = (const DevCollider *)(dv + po[s].col);
        sg.P.host_vals = nullptr;
        sg.P.host_stride = 0;
        sg.P.ncode = (uint32_t)P.code.size();
        sg.P.nf64 = (uint32_t)P.f64.size();
        if (!j.ax[0].empty()) {
            const double *ax = (const double *)(dv + po[s].ax);
            j.g.xs = ax, j.g.ys = ax + j.g.nx, j.g.zs = ax + j.g.nx + j.g.ny;
        }
        sg.g = j.g;
        sg.B = j.B;
        if (j.cap) sg.B.info = ctx->binfo.as<uint2>(), sg.B.code = ctx->bcode.as<Insn>();
        sg.word0 = j.word0, sg.nwords = j.nwords;
        sg.brick0 = j.brick0, sg.nbricks = j.nbricks;
        sg.iters = j.iters;
        sg.stage = j.cap && plan_stage_bytes(P) <= PLAN_STAGE_LIMIT ? 1 : 0;
    }
    memcpy(h + o_segs, U.hsegs.data(), n * sizeof(Seg));
    if (!jobs[0].conj.empty()) {
        memcpy(h + o_conj, jobs[0].conj.data(), jobs[0].conj.size() * sizeof(double));
        U.conj = (const double *)(dv + o_conj);
        U.conj_nops = (uint32_t)(jobs[0].conj.size() / 10);
    }
    if (tables) {
        uint32_t *ws = (uint32_t *)(h + o_wseg), *bs = (uint32_t *)(h + o_bseg);
        for (size_t s = 0; s < n; s++) {
            const uint64_t w1 = s + 1 < n ? jobs[s + 1].word0 : nwords_all;
            for (uint64_t w = jobs[s].word0; w < w1; w += 64) ws[w / 64] = (uint32_t)s;
            const uint64_t b1 = s + 1 < n ? jobs[s + 1].brick0 : nbricks_all;
            for (uint64_t b = jobs[s].brick0; b < b1; b += 128) bs[b / 128] = (uint32_t)s;
        }
    }
    U.ctl = (Ctl *)dv;
    U.sout = (SegOut *)(dv + o_sout);
    U.segs = (Seg *)(dv + o_segs);
    U.word_seg = tables ? (uint32_t *)(dv + o_wseg) : nullptr;
    U.brick_seg = tables ? (uint32_t *)(dv + o_bseg) : nullptr;
    HIP_TRY(hipMemcpyAsync(ctx->blob.p, h, total, hipMemcpyHostToDevice, ctx->stream));
    return MC3D_OK;
}

// Rewrites segment s on the device (after its host-evaluated answers moved).
int update_seg(mc3d_ctx *ctx, const Upload &U, size_t s) {
    int rc;
    if ((rc = ctx->segstage.ensure(sizeof(Seg)))) return rc;
    // Every caller synchronised the stream after the slot's previous copy.
    memcpy(ctx->segstage.p, &U.hsegs[s], sizeof(Seg));
    HIP_TRY(hipMemcpyAsync(U.segs + s, ctx->segstage.p, sizeof(Seg), hipMemcpyHostToDevice, ctx->stream));
    return MC3D_OK;
}

constexpr uint64_t HOST_BATCH = 1u << 20;  // points per callback

// Answers every host-evaluated atom of segment s at n points -- its grid
// corners (pts == nullptr) or a device point list -- through its callback,
// in batches, into hvals[h * n + i], and points the segment's program at
// them.  The GPU computes each atom's local points; the host only runs the
// callbacks.
int host_round(mc3d_ctx *ctx, const Job &j, Upload &U, size_t s, const double *pts, uint64_t n, double &ms_cb) {
    const Program &P = j.P;
    DevProg &dp = U.hsegs[s].P;
    const size_t nh = P.host_atoms.size();
    int rc;
    if ((rc = ctx->hvals.ensure(nh * n + 1)) || (rc = ctx->hpts.ensure(HOST_BATCH * 24)) ||
        (rc = ctx->hstage.ensure(HOST_BATCH * 25)))
        return rc;
    double *hp = (double *)ctx->hstage.p;
    uint8_t *ho = (uint8_t *)(hp + 3 * HOST_BATCH);
    for (size_t h = 0; h < nh; h++) {
        for (uint64_t s0 = 0; s0 < n; s0 += HOST_BATCH) {
            const uint64_t c = std::min(HOST_BATCH, n - s0);
            k_host_points<<<grid_for(c, 256), 256, 0, ctx->stream>>>(dp, U.hsegs[s].g, pts, j.frames[h], s0, c,
                                                                    ctx->hpts.as<double>());
            HIP_TRY(hipGetLastError());
            HIP_TRY(hipMemcpyAsync(hp, ctx->hpts.p, c * 24, hipMemcpyDeviceToHost, ctx->stream));
            // Also orders the previous batch's upload of `ho` before it is rewritten.
            HIP_TRY(hipStreamSynchronize(ctx->stream));
            const auto t0 = std::chrono::steady_clock::now();
            const int r = P.host_atoms[h].fn(P.host_atoms[h].user, hp, (size_t)c, ho);
            ms_cb += std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
            if (r) {
                set_error("host solid callback failed with status " + std::to_string(r));
                return MC3D_E_CALLBACK;
            }
            HIP_TRY(hipMemcpyAsync(ctx->hvals.as<uint8_t>() + h * n + s0, ho, c, hipMemcpyHostToDevice, ctx->stream));
        }
    }
    dp.host_vals = ctx->hvals.as<uint8_t>();
    dp.host_stride = n;
    return update_seg(ctx, U, s);
}

float elapsed(hipEvent_t a, hipEvent_t b) {
    float ms = 0;
    if (hipEventElapsedTime(&ms, a, b) != hipSuccess) return 0;
    return ms;
}

// The scan of the counts whose per-block sums k_classify wrote.
int scan4(mc3d_ctx *ctx, const uint32_t *in, uint2 *out, uint64_t n, unsigned long long *totals, uint32_t *list,
          const Seg &seg0, bool bm) {
    const uint64_t ntiles = (n + SCAN_TILE - 1) / SCAN_TILE;
    k_scan_tiles<<<1, SUMS_BLOCK, 0, ctx->stream>>>(ctx->tiles.as<uint4>(), (n + CLASSIFY_BLOCK - 1) / CLASSIFY_BLOCK,
                                                    totals);
    k_scan_down<<<(unsigned)ntiles, SCAN_BLOCK, 0, ctx->stream>>>(in, n, ctx->tiles.as<uint4>(), out, list, seg0, bm);
    HIP_TRY(hipGetLastError());
    return MC3D_OK;
}

// Sizes and launch shapes of one call.
struct Call {
    uint32_t nsegs = 0;
    uint64_t nwords = 0, nbricks = 0;  // over all segments, with alignment padding
    uint32_t cap = 0;                  // folded-program slot (culling) of every brick
    bool cull = false;
    int mode = 0, rows_per_walk = 1, wpb = 4;
    bool stats = false;
    bool interior = false;  // a call of one MarchingCubesInterior solid
    const uint64_t *keep = nullptr;  // a call of one filtered solid: its cube mask on the device
    bool bm = false;     // the scan takes the words in brick-major order (one culled solid)
    uint64_t nscan = 0;  // scan positions: words, or brick-major positions
};

template <int MODE, bool STATS, int NP, bool MULTI>
void launch_corners_act(mc3d_ctx *ctx, const Call &c, const Upload &U, unsigned grid) {
    uint32_t *act = ctx->actb.as<uint32_t>();
    uint64_t *bits = ctx->bits.as<uint64_t>();
    const Seg &s0 = U.hsegs[0];
    if (c.wpb == 1)
        k_corners_act<MODE, STATS, NP, 1, MULTI>
            <<<grid, 64, 0, ctx->stream>>>(U.segs, c.nsegs, s0, U.brick_seg, act, bits, U.ctl, U.sout);
    else if (c.wpb == 2)
        k_corners_act<MODE, STATS, NP, 2, MULTI>
            <<<grid, 128, 0, ctx->stream>>>(U.segs, c.nsegs, s0, U.brick_seg, act, bits, U.ctl, U.sout);
    else
        k_corners_act<MODE, STATS, NP, 4, MULTI>
            <<<grid, 256, 0, ctx->stream>>>(U.segs, c.nsegs, s0, U.brick_seg, act, bits, U.ctl, U.sout);
}

template <int MODE, bool STATS, bool MULTI>
void launch_corners_m(mc3d_ctx *ctx, const Call &c, const Upload &U) {
    const Seg &s0 = U.hsegs[0];
    if (!c.cull) {
        k_corners<MODE, STATS, MULTI><<<grid_for(c.nwords, 4), 256, 0, ctx->stream>>>(
            U.segs, c.nsegs, s0, U.word_seg, ctx->bits.as<uint64_t>(), U.ctl, U.sout, c.nwords);
        return;
    }
    // (the words of constant bricks were written by the plan)
    const unsigned grid = grid_for(c.nbricks, 1, 1u << 15);
    if (c.rows_per_walk >= 4)
        launch_corners_act<MODE, STATS, 4, MULTI>(ctx, c, U, grid);
    else if (c.rows_per_walk == 2)
        launch_corners_act<MODE, STATS, 2, MULTI>(ctx, c, U, grid);
    else
        launch_corners_act<MODE, STATS, 1, MULTI>(ctx, c, U, grid);
}

template <int MODE, bool STATS>
void launch_corners(mc3d_ctx *ctx, const Call &c, const Upload &U) {
    c.nsegs > 1 ? launch_corners_m<MODE, STATS, true>(ctx, c, U) : launch_corners_m<MODE, STATS, false>(ctx, c, U);
}

template <int MODE, bool STATS>
void launch_bisect(mc3d_ctx *ctx, const Call &c, const Upload &U, double *verts, double *interior,
                   uint64_t vcap) {
    const unsigned grid = grid_for(vcap, 256, 1u << 14);
    if (c.nsegs > 1)
        k_bisect<MODE, STATS, true><<<grid, 256, 0, ctx->stream>>>(
            U.segs, c.nsegs, U.hsegs[0], ctx->vseg.as<uint32_t>(), ctx->bits.as<uint64_t>(),
            ctx->vkeys.as<uint64_t>(), verts, interior, U.ctl, vcap);
    else
        k_bisect<MODE, STATS, false><<<grid, 256, 0, ctx->stream>>>(U.segs, c.nsegs, U.hsegs[0], nullptr,
                                                                    ctx->bits.as<uint64_t>(),
                                                                    ctx->vkeys.as<uint64_t>(), verts, interior,
                                                                    U.ctl, vcap);
}

int take_outputs(mc3d_ctx *ctx, MeshBuffers &mb, const Call &c, uint64_t vcap, uint64_t tcap) {
    int rc;
    if (!mb.verts || mb.verts->cap < vcap * 24) {
        ctx->pool->give(std::move(mb.verts));
        mb.verts = ctx->pool->take(vcap * 24);
    }
    if (!mb.tris || mb.tris->cap < tcap * 12) {
        ctx->pool->give(std::move(mb.tris));
        mb.tris = ctx->pool->take(tcap * 12);
    }
    if ((rc = mb.verts->ensure(vcap * 24)) || (rc = mb.tris->ensure(tcap * 12)) ||
        (rc = ctx->vkeys.ensure(vcap * 8)))
        return rc;
    if (c.nsegs > 1 && (rc = ctx->vseg.ensure(vcap * 4))) return rc;
    if (c.interior) {
        if (!mb.interior || mb.interior->cap < vcap * 24) {
            ctx->pool->give(std::move(mb.interior));
            mb.interior = ctx->pool->take(vcap * 24);
        }
        if ((rc = mb.interior->ensure(vcap * 24))) return rc;
    }
    return MC3D_OK;
}

// Meshes keep their outputs until freed: a buffer sized for earlier, larger
// calls is traded for one that fits (the copy is ordered on the stream), and
// goes back to the pool for the next call.
int fit_outputs(mc3d_ctx *ctx, MeshBuffers &mb, uint64_t nv, uint64_t nt) {
    auto fit = [&](std::unique_ptr<DevBuf> &b, size_t need) -> int {
        if (!b || b->cap <= 2 * need + (4u << 20)) return MC3D_OK;
        std::unique_ptr<DevBuf> f = ctx->pool->take(need);
        int rc = f->ensure(need);
        if (rc) return rc;
        if (need) HIP_TRY(hipMemcpyAsync(f->p, b->p, need, hipMemcpyDeviceToDevice, ctx->stream));
        ctx->pool->give(std::move(b));
        b = std::move(f);
        return MC3D_OK;
    };
    int rc;
    if ((rc = fit(mb.verts, nv * 24)) || (rc = fit(mb.tris, nt * 12)) || (rc = fit(mb.interior, nv * 24)))
        return rc;
    return MC3D_OK;
}

void launch_emit(mc3d_ctx *ctx, const Call &c, const Upload &U, MeshBuffers &mb, uint64_t vcap, uint64_t tcap) {
    // 2048 blocks, up to 8192 for large grids (a few listed words per wave:
    // even tails; each block also stages the case table)
    const unsigned grid = std::max(2048u, grid_for(c.nwords, 128, 8192));
    if (c.nsegs > 1)
        k_emit<true><<<grid, 256, 0, ctx->stream>>>(U.segs, c.nsegs, U.hsegs[0], U.word_seg, ctx->bits.as<uint64_t>(),
                                                    ctx->etable.as<uint8_t>(), ctx->offs.as<uint2>(),
                                                    ctx->actw.as<uint32_t>(), U.ctl, ctx->vkeys.as<uint64_t>(),
                                                    ctx->vseg.as<uint32_t>(), mb.tris->as<uint32_t>(), nullptr, vcap,
                                                    tcap, false);
    else
        k_emit<false><<<grid, 256, 0, ctx->stream>>>(U.segs, c.nsegs, U.hsegs[0], U.word_seg,
                                                     ctx->bits.as<uint64_t>(), ctx->etable.as<uint8_t>(),
                                                     ctx->offs.as<uint2>(), ctx->actw.as<uint32_t>(), U.ctl,
                                                     ctx->vkeys.as<uint64_t>(), nullptr, mb.tris->as<uint32_t>(), c.keep,
                                                     vcap, tcap, c.bm);
}

// The output stages: emit (vertex keys, triangles) and bisect (vertices).
int run_output(mc3d_ctx *ctx, const Call &c, const Upload &U, MeshBuffers &mb, uint64_t vcap, uint64_t tcap) {
    HIP_TRY(hipEventRecord(ctx->ev[4], ctx->stream));
    launch_emit(ctx, c, U, mb, vcap, tcap);
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipEventRecord(ctx->ev[5], ctx->stream));
    double *verts = mb.verts->as<double>();
    double *inter = c.interior ? mb.interior->as<double>() : nullptr;
    if (c.mode == 0)
        c.stats ? launch_bisect<0, true>(ctx, c, U, verts, inter, vcap)
                : launch_bisect<0, false>(ctx, c, U, verts, inter, vcap);
    else
        c.stats ? launch_bisect<1, true>(ctx, c, U, verts, inter, vcap)
                : launch_bisect<1, false>(ctx, c, U, verts, inter, vcap);
    HIP_TRY(hipGetLastError());
    if (U.conj) {
        k_inverse_apply<<<grid_for(vcap, 256, 1u << 14), 256, 0, ctx->stream>>>(U.conj, U.conj_nops, verts, U.ctl,
                                                                                 vcap);
        HIP_TRY(hipGetLastError());
    }
    HIP_TRY(hipEventRecord(ctx->ev[6], ctx->stream));
    return MC3D_OK;
}

// The output stages of a solid with host-evaluated atoms (a call of one):
// the mesh size is read first (one more synchronisation), then emit, then
// the bisection in rounds with the host answering its atoms between them.
// Leaves the final control block in ctx->ctl_host.
int run_output_host(mc3d_ctx *ctx, const Call &c, const Job &j, Upload &U, MeshBuffers &mb, double &ms_cb) {
    Ctl *hc = (Ctl *)ctx->ctl_host.p;
    HIP_TRY(hipMemcpyAsync(hc, U.ctl, U.back_bytes, hipMemcpyDeviceToHost, ctx->stream));
    HIP_TRY(hipStreamSynchronize(ctx->stream));
    if (hc->inside) {
        set_error("solid is true outside of bounds");
        return MC3D_E_INSIDE_BOUNDARY;
    }
    const uint64_t nv = hc->totals[0], nt = hc->totals[1];
    if (nv >= (1ull << 32) || nt >= (1ull << 32)) {
        set_error("mesh too large for 32-bit vertex indices");
        return MC3D_E_UNSUPPORTED;
    }
    const uint64_t vcap = std::max<uint64_t>(nv, 1), tcap = std::max<uint64_t>(nt, 1);
    int rc;
    if ((rc = take_outputs(ctx, mb, c, vcap, tcap)) || (rc = ctx->hstate.ensure(vcap * 16)) ||
        (rc = ctx->hprobe.ensure(vcap * 24)))
        return rc;
    HIP_TRY(hipEventRecord(ctx->ev[4], ctx->stream));
    launch_emit(ctx, c, U, mb, vcap, tcap);
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipEventRecord(ctx->ev[5], ctx->stream));
    double *verts = mb.verts->as<double>(), *state = ctx->hstate.as<double>(), *probe = ctx->hprobe.as<double>();
    double *inter = c.interior ? mb.interior->as<double>() : nullptr;
    const unsigned grid = grid_for(nv, 256, 1u << 14);
    for (int round = 0; round <= j.iters; round++) {
        if (round > 0 && (rc = host_round(ctx, j, U, 0, probe, nv, ms_cb))) return rc;
        const DevProg &dp = U.hsegs[0].P;
        if (c.mode == 1)
            k_bisect_host<1><<<grid, 256, 0, ctx->stream>>>(dp, U.hsegs[0].g, ctx->bits.as<uint64_t>(),
                                                            ctx->vkeys.as<uint64_t>(), state, probe, verts, inter,
                                                            nv, round > 0);
        else
            k_bisect_host<0><<<grid, 256, 0, ctx->stream>>>(dp, U.hsegs[0].g, ctx->bits.as<uint64_t>(),
                                                            ctx->vkeys.as<uint64_t>(), state, probe, verts, inter,
                                                            nv, round > 0);
        HIP_TRY(hipGetLastError());
    }
    if (U.conj) {
        k_inverse_apply<<<grid_for(nv, 256, 1u << 14), 256, 0, ctx->stream>>>(U.conj, U.conj_nops, verts, U.ctl, nv);
        HIP_TRY(hipGetLastError());
    }
    HIP_TRY(hipEventRecord(ctx->ev[6], ctx->stream));
    HIP_TRY(hipMemcpyAsync(hc, U.ctl, U.back_bytes, hipMemcpyDeviceToHost, ctx->stream));
    HIP_TRY(hipStreamSynchronize(ctx->stream));
    return MC3D_OK;
}

// Meshes `jobs` (prepared) in one pass of the pipeline; out[i] gets job i's
// mesh.  A job with host-evaluated atoms is only ever run alone.
double ms_since(std::chrono::steady_clock::time_point t0) {
    return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
}

int run_call(mc3d_ctx *ctx, std::vector<Job> &jobs, const mc3d_options &o, std::vector<mc3d_mesh *> &out,
             std::chrono::steady_clock::time_point t_call, double ms_prepare) {
    const auto t_enqueue = std::chrono::steady_clock::now();
    double ms_enqueue = 0, ms_wait = 0;
    Call c;
    c.nsegs = (uint32_t)jobs.size();
    c.mode = o.precision == MC3D_PREC_FP32_GUARD ? 1 : 0;
    c.stats = o.count_mismatch != 0;
    c.rows_per_walk = o.rows_per_walk > 0 ? o.rows_per_walk : 1;
    c.wpb = o.waves_per_brick > 0 ? o.waves_per_brick : 4;
    c.cull = jobs[0].cap != 0;
    c.interior = c.nsegs == 1 && jobs[0].interior;
    const bool host = c.nsegs == 1 && jobs[0].host;
    for (Job &j : jobs) {
        j.word0 = c.nwords;
        c.nwords += align_up(j.nwords, 64);
        j.brick0 = c.nbricks;
        c.nbricks += align_up(j.nbricks, 128);
        c.cap = std::max(c.cap, j.cap);
    }
    if (c.nwords >= (1ull << 32)) {
        set_error("call too large: more than 2^32 occupancy words");
        return MC3D_E_UNSUPPORTED;
    }
    if (c.cull && c.nbricks * c.cap >= (1ull << 32) - 1) {
        set_error("too many culling bricks");
        return MC3D_E_UNSUPPORTED;
    }
    // One culled solid: the scan takes its words brick by brick (the
    // positions include the rows of bricks past the grid).
    c.nscan = c.nwords;
    const int BY = jobs[0].B.BY, BZ = jobs[0].B.BZ;
    if (c.cull && c.nsegs == 1 && !host && (BY & (BY - 1)) == 0 && (BZ & (BZ - 1)) == 0) {
        const uint64_t n = c.nbricks * (uint64_t)BY * (uint64_t)BZ;
        if (n < (1ull << 32)) c.bm = true, c.nscan = n;
    }
    const unsigned ctiles = (unsigned)((c.nscan + CLASSIFY_BLOCK - 1) / CLASSIFY_BLOCK);  // classify blocks
    int rc;
    if ((rc = ctx->bits.ensure(c.nwords * 8)) || (rc = ctx->counts.ensure(c.nscan * sizeof(uint32_t))) ||
        (rc = ctx->offs.ensure(c.nscan * sizeof(uint2))) || (rc = ctx->actw.ensure(c.nwords * sizeof(uint32_t))) ||
        (rc = ctx->tiles.ensure((uint64_t)ctiles * sizeof(uint4))))
        return rc;
    if (c.cull && ((rc = ctx->binfo.ensure(c.nbricks * sizeof(uint2))) ||
                   (rc = ctx->actb.ensure(c.nbricks * sizeof(uint32_t))) ||
                   (rc = ctx->bcode.ensure(c.nbricks * c.cap * sizeof(Insn) + 16))))
        return rc;
    if (c.nsegs == 1 && !jobs[0].keep.empty()) {
        if ((rc = ctx->keep.ensure(jobs[0].keep.size() * 8))) return rc;
        HIP_TRY(hipMemcpyAsync(ctx->keep.p, jobs[0].keep.data(), jobs[0].keep.size() * 8, hipMemcpyHostToDevice,
                               ctx->stream));
        c.keep = ctx->keep.as<uint64_t>();
    }
    Upload U;
    if ((rc = upload(ctx, jobs, c.nwords, c.nbricks, U))) return rc;
    if ((rc = ctx->ctl_host.ensure(U.back_bytes))) return rc;

    std::shared_ptr<MeshBuffers> mb = std::make_shared<MeshBuffers>();
    mb->pool = ctx->pool;
    mc3d_stats st{};
    if (host && (rc = host_round(ctx, jobs[0], U, 0, nullptr, (uint64_t)jobs[0].g.nx * jobs[0].g.ny * jobs[0].g.nz,
                                 st.ms_host_callbacks)))
        return rc;

    // ---- plan ----
    HIP_TRY(hipEventRecord(ctx->ev[0], ctx->stream));
    if (c.cull) {
        // Programs that fit are staged in LDS beside the per-thread stacks.
        size_t lds = PLAN_STACK_BYTES;
        for (size_t s = 0; s < jobs.size(); s++)
            if (U.hsegs[s].stage) lds = std::max(lds, plan_stage_bytes(jobs[s].P));
        const unsigned pgrid = grid_for(c.nbricks, PLAN_BLOCK);
        if (c.nsegs > 1)
            k_plan<true><<<pgrid, PLAN_BLOCK, lds, ctx->stream>>>(
                U.segs, c.nsegs, U.hsegs[0], U.brick_seg, c.nbricks, c.cap, ctx->binfo.as<uint2>(),
                ctx->bcode.as<Insn>(), ctx->actb.as<uint32_t>(), ctx->bits.as<uint64_t>(), U.ctl, U.sout);
        else
            k_plan<false><<<pgrid, PLAN_BLOCK, lds, ctx->stream>>>(
                U.segs, c.nsegs, U.hsegs[0], U.brick_seg, c.nbricks, c.cap, ctx->binfo.as<uint2>(),
                ctx->bcode.as<Insn>(), ctx->actb.as<uint32_t>(), ctx->bits.as<uint64_t>(), U.ctl, U.sout);
        HIP_TRY(hipGetLastError());
    }
    HIP_TRY(hipEventRecord(ctx->ev[1], ctx->stream));

    // ---- corners ----
    if (c.mode == 0)
        c.stats ? launch_corners<0, true>(ctx, c, U) : launch_corners<0, false>(ctx, c, U);
    else
        c.stats ? launch_corners<1, true>(ctx, c, U) : launch_corners<1, false>(ctx, c, U);
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipEventRecord(ctx->ev[2], ctx->stream));

    // ---- classify + scan (+ list of words that emit) ----
    if (c.nsegs > 1)
        k_classify<true><<<ctiles, CLASSIFY_BLOCK, 0, ctx->stream>>>(
            U.segs, c.nsegs, U.hsegs[0], U.word_seg, ctx->bits.as<uint64_t>(), ctx->table.as<uint8_t>(),
            ctx->counts.as<uint32_t>(), ctx->tiles.as<uint4>(), nullptr, c.nwords, false);
    else
        k_classify<false><<<ctiles, CLASSIFY_BLOCK, 0, ctx->stream>>>(
            U.segs, c.nsegs, U.hsegs[0], U.word_seg, ctx->bits.as<uint64_t>(), ctx->table.as<uint8_t>(),
            ctx->counts.as<uint32_t>(), ctx->tiles.as<uint4>(), c.keep, c.nscan, c.bm);
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipEventRecord(ctx->ev[8], ctx->stream));
    if ((rc = scan4(ctx, ctx->counts.as<uint32_t>(), ctx->offs.as<uint2>(), c.nscan, U.ctl->totals,
                    ctx->actw.as<uint32_t>(), U.hsegs[0], c.bm)))
        return rc;
    if (c.nsegs > 1) {  // a single solid starts at vertex 0 and triangle 0 (zeroed by the upload)
        k_seg_ranges<<<grid_for(c.nsegs, 256), 256, 0, ctx->stream>>>(U.segs, c.nsegs, ctx->offs.as<uint2>(), U.sout);
        HIP_TRY(hipGetLastError());
    }
    HIP_TRY(hipEventRecord(ctx->ev[3], ctx->stream));

    // ---- emit + bisect, sized from earlier calls ----
    Ctl *hc = (Ctl *)ctx->ctl_host.p;
    const SegOut *so = (const SegOut *)((const char *)hc + ((const char *)U.sout - (const char *)U.ctl));
    if (host) {
        if ((rc = run_output_host(ctx, c, jobs[0], U, *mb, st.ms_host_callbacks))) return rc;
    } else {
        uint64_t vcap = ctx->v_hint, tcap = ctx->t_hint;
        if ((rc = take_outputs(ctx, *mb, c, vcap, tcap))) return rc;
        if ((rc = run_output(ctx, c, U, *mb, vcap, tcap))) return rc;
        HIP_TRY(hipMemcpyAsync(hc, U.ctl, U.back_bytes, hipMemcpyDeviceToHost, ctx->stream));
        ms_enqueue = ms_since(t_enqueue);
        const auto t_wait = std::chrono::steady_clock::now();
        HIP_TRY(hipStreamSynchronize(ctx->stream));
        ms_wait = ms_since(t_wait);
        if (hc->inside) {
            size_t bad = 0;
            while (bad + 1 < jobs.size() && !so[bad].inside) bad++;
            set_error(c.nsegs == 1 ? std::string("solid is true outside of bounds")
                                   : "solid " + std::to_string(bad) + " of the batch is true outside of bounds");
            return MC3D_E_INSIDE_BOUNDARY;
        }
        const uint64_t nv = hc->totals[0], nt = hc->totals[1];
        if (nv >= (1ull << 32) || nt >= (1ull << 32)) {
            set_error("mesh too large for 32-bit vertex indices");
            return MC3D_E_UNSUPPORTED;
        }
        if (nv > vcap || nt > tcap) {  // grow and run the output stages again
            ctx->v_hint = std::max(ctx->v_hint, nv + nv / 4 + 1024);
            ctx->t_hint = std::max(ctx->t_hint, nt + nt / 4 + 1024);
            vcap = ctx->v_hint, tcap = ctx->t_hint;
            if ((rc = take_outputs(ctx, *mb, c, vcap, tcap))) return rc;
            // the bisection's operation counts start again with the second run
            HIP_TRY(hipMemsetAsync(U.ctl->flops + 2, 0, 2 * sizeof(unsigned long long), ctx->stream));
            if ((rc = run_output(ctx, c, U, *mb, vcap, tcap))) return rc;
            HIP_TRY(hipMemcpyAsync(hc, U.ctl, U.back_bytes, hipMemcpyDeviceToHost, ctx->stream));
            HIP_TRY(hipStreamSynchronize(ctx->stream));
        }
    }
    const uint64_t nv = hc->totals[0], nt = hc->totals[1];
    if ((rc = fit_outputs(ctx, *mb, nv, nt))) return rc;

    st.bricks_active = hc->nact_bricks;
    st.guard_evals = hc->gstats[0];
    st.guard_rechecks = hc->gstats[1];
    st.guard_mismatches = hc->gstats[2];
    st.guard_violations = hc->gstats[3];
    st.corner_flops64 = hc->flops[0];
    st.corner_flops32 = hc->flops[1];
    st.bisect_flops64 = hc->flops[2];
    st.bisect_flops32 = hc->flops[3];
    st.brick_code_total = hc->code_total;
    st.brick_code_max = hc->code_max;
    st.bricks_full = hc->full_bricks;
    st.ms_plan = elapsed(ctx->ev[0], ctx->ev[1]);
    st.ms_fill = 0;  // the plan writes the constant bricks' words
    st.ms_corners = elapsed(ctx->ev[1], ctx->ev[2]);
    st.ms_classify = elapsed(ctx->ev[2], ctx->ev[8]);
    st.ms_scan = elapsed(ctx->ev[8], ctx->ev[3]);
    st.ms_emit = elapsed(ctx->ev[4], ctx->ev[5]);
    st.ms_bisect = elapsed(ctx->ev[5], ctx->ev[6]);
    st.ms_device = elapsed(ctx->ev[0], ctx->ev[6]);

    // One mesh per job: its slice of the outputs.
    std::vector<std::unique_ptr<mc3d_mesh>> meshes(jobs.size());
    for (size_t s = 0; s < jobs.size(); s++) {
        meshes[s].reset(new (std::nothrow) mc3d_mesh);
        if (!meshes[s]) return MC3D_E_OOM;
        mc3d_mesh *m = meshes[s].get();
        const Job &j = jobs[s];
        m->device = ctx->device;
        m->buf = mb;
        m->has_interior = c.interior;
        m->voff = so[s].vbase;
        m->toff = so[s].tbase;
        m->nv = (s + 1 < jobs.size() ? so[s + 1].vbase : nv) - m->voff;
        m->nt = (s + 1 < jobs.size() ? so[s + 1].tbase : nt) - m->toff;
        m->stats = st;
        mc3d_stats &ms = m->stats;
        ms.corners = (uint64_t)j.g.nx * j.g.ny * j.g.nz;
        ms.grid_nx = j.g.nx, ms.grid_ny = j.g.ny, ms.grid_nz = j.g.nz;
        ms.vertices = m->nv;
        ms.triangles = m->nt;
        ms.bisection_evals = j.iters > 0 ? m->nv * (uint64_t)(j.iters + 1) : 0;
        ms.bricks_total = j.nbricks;
        ms.program_length = j.P.code.size();
    }
    if (!o.device_only) {
        auto t0 = std::chrono::steady_clock::now();
        // one copy of all outputs through pinned memory, then split per mesh
        const size_t ibytes = c.interior ? nv * 24 : 0;
        if ((rc = ctx->outstage.ensure(nv * 24 + nt * 12 + ibytes + 16))) return rc;
        char *hs = (char *)ctx->outstage.p;
        if (nv) HIP_TRY(hipMemcpyAsync(hs, mb->verts->p, nv * 24, hipMemcpyDeviceToHost, ctx->stream));
        if (nt) HIP_TRY(hipMemcpyAsync(hs + nv * 24, mb->tris->p, nt * 12, hipMemcpyDeviceToHost, ctx->stream));
        if (ibytes)
            HIP_TRY(hipMemcpyAsync(hs + nv * 24 + nt * 12, mb->interior->p, ibytes, hipMemcpyDeviceToHost,
                                   ctx->stream));
        HIP_TRY(hipStreamSynchronize(ctx->stream));
        const double *hv = (const double *)hs;
        const uint32_t *ht = (const uint32_t *)(hs + nv * 24);
        for (auto &m : meshes) {
            m->hverts.reset(new double[3 * m->nv]);
            m->htris.reset(new uint32_t[3 * m->nt]);
            parallel_copy(m->hverts.get(), hv + 3 * m->voff, m->nv * 24);
            parallel_copy(m->htris.get(), ht + 3 * m->toff, m->nt * 12);
            if (ibytes) {
                const double *hi = (const double *)(hs + nv * 24 + nt * 12);
                m->hinterior.reset(new double[3 * m->nv]);
                parallel_copy(m->hinterior.get(), hi + 3 * m->voff, m->nv * 24);
            }
            m->on_host = true;
        }
        const double ms_d2h = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
        for (auto &m : meshes) m->stats.ms_d2h = ms_d2h;
    }
    const double ms_call = ms_since(t_call);
    out.resize(jobs.size());
    for (size_t s = 0; s < jobs.size(); s++) {
        meshes[s]->stats.ms_call = ms_call;
        meshes[s]->stats.ms_prepare = ms_prepare;
        meshes[s]->stats.ms_enqueue = ms_enqueue;
        meshes[s]->stats.ms_wait = ms_wait;
        out[s] = meshes[s].release();
    }
    return MC3D_OK;
}

mc3d_options default_options(const mc3d_options *opts) {
    mc3d_options o{};
    o.precision = MC3D_PREC_FP32_GUARD;
    o.cull = 1;
    if (opts) o = *opts;
    return o;
}

}  // namespace

extern "C" {

int mc3d_ctx_create(int device, mc3d_ctx **out) {
    if (!out) return MC3D_E_INVALID_ARG;
    *out = nullptr;
    int n = 0;
    HIP_TRY(hipGetDeviceCount(&n));
    if (device < 0 || device >= n) {
        set_error("mc3d_ctx_create: no such HIP device");
        return MC3D_E_INVALID_ARG;
    }
    HIP_TRY(hipSetDevice(device));
    mc3d_ctx *c = new (std::nothrow) mc3d_ctx;
    if (!c) return MC3D_E_OOM;
    c->device = device;
    hipError_t e = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking);
    if (e != hipSuccess) {
        set_error(std::string("hipStreamCreate: ") + hipGetErrorString(e));
        delete c;
        return MC3D_E_HIP;
    }
    for (auto &ev : c->ev) {
        if (hipEventCreate(&ev) != hipSuccess) {
            set_error("hipEventCreate failed");
            mc3d_ctx_destroy(c);
            return MC3D_E_HIP;
        }
    }
    // The emitter's copy of the case table: each edge e = 4 * axis + r
    // decoded to axis | dx << 2 | dy << 3 | dz << 4 (its offset from the
    // cube's low corner); entry 0 of a case stays its triangle count.
    uint8_t etab[256 * 16];
    const uint8_t *tab = case_table();
    for (int i = 0; i < 256 * 16; i++) {
        const int t = tab[i];
        if ((i & 15) == 0) {
            etab[i] = (uint8_t)t;
            continue;
        }
        const int axis = (t >> 2) & 3, d0 = t & 1, d1 = (t >> 1) & 1;  // d0, d1: the other two axes in x, y, z order
        const int dx = axis == 0 ? 0 : d0, dy = axis == 0 ? d0 : (axis == 1 ? 0 : d1), dz = axis == 2 ? 0 : d1;
        etab[i] = (uint8_t)(axis | (dx << 2) | (dy << 3) | (dz << 4));
    }
    if (c->ctl_host.ensure(sizeof(Ctl) + 4096) || c->table.ensure(256 * 16) || c->etable.ensure(256 * 16) ||
        hipMemcpy(c->table.p, tab, 256 * 16, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(c->etable.p, etab, 256 * 16, hipMemcpyHostToDevice) != hipSuccess) {
        set_error("mc3d_ctx_create: setup failed");
        mc3d_ctx_destroy(c);
        return MC3D_E_HIP;
    }
    *out = c;
    return MC3D_OK;
}

void mc3d_ctx_destroy(mc3d_ctx *ctx) {
    if (!ctx) return;
    (void)hipSetDevice(ctx->device);
    if (ctx->stream) (void)hipStreamSynchronize(ctx->stream);
    for (auto &ev : ctx->ev)
        if (ev) (void)hipEventDestroy(ev);
    if (ctx->stream) (void)hipStreamDestroy(ctx->stream);
    delete ctx;
}

namespace {
// One solid, one call: MarchingCubesSearch, with interior points
// (MarchingCubesInterior) or with an inverse chain for the vertices
// (MarchingCubesConj).
using RectFilter = std::function<bool(const double *, const double *)>;

int search_one(mc3d_ctx *ctx, const mc3d_solid *s, double delta, int iters, const mc3d_options *opts,
               mc3d_mesh **out, bool interior, std::vector<double> conj, const RectFilter *filter = nullptr,
               int filter_threads = 1) {
    *out = nullptr;
    const auto t_call = std::chrono::steady_clock::now();
    const mc3d_options o = default_options(opts);
    HIP_TRY(hipSetDevice(ctx->device));
    std::vector<Job> jobs(1);
    jobs[0].s = s, jobs[0].delta = delta, jobs[0].iters = iters;
    jobs[0].interior = interior;
    jobs[0].conj = std::move(conj);
    int rc = prepare_job(jobs[0], o);
    if (rc) return rc;
    if (filter && (rc = filter_keep_mask(jobs[0].ax, delta, *filter, filter_threads, jobs[0].keep))) return rc;
    const double ms_prepare = ms_since(t_call);
    std::vector<mc3d_mesh *> meshes;
    if ((rc = run_call(ctx, jobs, o, meshes, t_call, ms_prepare))) return rc;
    *out = meshes[0];
    return MC3D_OK;
}
}  // namespace

int mc3d_marching_cubes_search_filter(mc3d_ctx *ctx, const mc3d_solid *s, mc3d_rect_filter_fn f, void *user,
                                      double delta, int iters, const mc3d_options *opts, mc3d_mesh **out) {
    if (!ctx || !s || !f || !out) {
        set_error("mc3d_marching_cubes_search_filter: null argument");
        return MC3D_E_INVALID_ARG;
    }
    const RectFilter filter = [f, user](const double *mn, const double *mx) { return f(user, mn, mx) != 0; };
    // a caller's filter is asked on this thread; the library's own rect
    // collider is thread-safe and runs on 16
    const int threads = f == mc3d_rect_collider_collides ? 16 : 1;
    return search_one(ctx, s, delta, iters, opts, out, false, {}, &filter, threads);
}

struct mc3d_rect_collider {
    RectMesh mesh;
};

mc3d_rect_collider *mc3d_rect_collider_create(const double *triangles, size_t n) {
    if (n && !triangles) {
        set_error("mc3d_rect_collider_create: null triangles");
        return nullptr;
    }
    mc3d_rect_collider *c = new (std::nothrow) mc3d_rect_collider;
    if (!c) {
        set_error("out of memory");
        return nullptr;
    }
    c->mesh.build(triangles, n);
    return c;
}

int mc3d_rect_collider_collides(void *collider, const double min[3], const double max[3]) {
    return collider && ((const mc3d_rect_collider *)collider)->mesh.collides(min, max) ? 1 : 0;
}

void mc3d_rect_collider_free(mc3d_rect_collider *c) { delete c; }

int mc3d_marching_cubes_filter(mc3d_ctx *ctx, const mc3d_solid *s, mc3d_rect_filter_fn f, void *user, double delta,
                               const mc3d_options *opts, mc3d_mesh **out) {
    return mc3d_marching_cubes_search_filter(ctx, s, f, user, delta, 0, opts, out);
}

int mc3d_marching_cubes_c2f(mc3d_ctx *ctx, const mc3d_solid *s, double big_delta, double small_delta,
                            double extra_space, int iters, const mc3d_options *opts, mc3d_mesh **out) {
    if (!ctx || !s || !out) {
        set_error("mc3d_marching_cubes_c2f: null argument");
        return MC3D_E_INVALID_ARG;
    }
    *out = nullptr;
    if (big_delta < small_delta) {
        set_error("bigDelta should be >= smallDelta for efficiency");  // mc.go:162-164
        return MC3D_E_INVALID_ARG;
    }
    extra_space += 2 * big_delta * sqrt(3.0);  // mc.go:167
    // the coarse mesh (mc.go:169), its triangles on the host
    mc3d_options co = default_options(opts);
    co.device_only = 1;
    mc3d_mesh *coarse = nullptr;
    int rc = search_one(ctx, s, big_delta, iters, &co, &coarse, false, {});
    if (rc) return rc;
    std::vector<double> v(coarse->nv * 3);
    std::vector<uint32_t> t(coarse->nt * 3);
    rc = mc3d_mesh_copy(coarse, v.data(), t.data());
    const size_t nt = coarse->nt;
    mc3d_mesh_free(coarse);
    if (rc) return rc;
    std::vector<double> tris(nt * 9);
    for (size_t i = 0; i < 3 * nt; i++)
        for (int d = 0; d < 3; d++) tris[3 * i + d] = v[3 * (size_t)t[i] + d];
    RectMesh collider;  // MeshToCollider(coarseMesh), mc.go:170
    collider.build(tris.data(), nt);
    // filter = collider.RectCollision(r.Expand(extraSpace)), mc.go:171-174
    const RectFilter filter = [&collider, extra_space](const double *mn, const double *mx) {
        const double emn[3] = {mn[0] + -extra_space, mn[1] + -extra_space, mn[2] + -extra_space};
        const double emx[3] = {mx[0] + extra_space, mx[1] + extra_space, mx[2] + extra_space};
        return collider.collides(emn, emx);
    };
    return search_one(ctx, s, small_delta, iters, opts, out, false, {}, &filter, 16);
}

int mc3d_marching_cubes_search(mc3d_ctx *ctx, const mc3d_solid *s, double delta, int iters,
                               const mc3d_options *opts, mc3d_mesh **out) {
    if (!ctx || !s || !out) {
        set_error("mc3d_marching_cubes_search: null argument");
        return MC3D_E_INVALID_ARG;
    }
    return search_one(ctx, s, delta, iters, opts, out, false, {});
}

int mc3d_marching_cubes_interior(mc3d_ctx *ctx, const mc3d_solid *s, double delta, int iters,
                                 const mc3d_options *opts, mc3d_mesh **out) {
    if (!ctx || !s || !out) {
        set_error("mc3d_marching_cubes_interior: null argument");
        return MC3D_E_INVALID_ARG;
    }
    return search_one(ctx, s, delta, iters, opts, out, true, {});
}

int mc3d_marching_cubes_conj(mc3d_ctx *ctx, const mc3d_solid *s, double delta, int iters,
                             const mc3d_xform *xforms, size_t n, const mc3d_options *opts, mc3d_mesh **out) {
    if (!ctx || !s || !out || (n && !xforms)) {
        set_error("mc3d_marching_cubes_conj: null argument");
        return MC3D_E_INVALID_ARG;
    }
    *out = nullptr;
    // TransformSolid(JoinedTransform(xforms), s), mc.go:67-68
    mc3d_solid *t = mc3d_transform(const_cast<mc3d_solid *>(s), xforms, n);
    if (!t) return MC3D_E_INVALID_ARG;  // the error is set
    const int rc = search_one(ctx, t, delta, iters, opts, out, false, inverse_records(xforms, n));
    mc3d_solid_release(t);
    return rc;
}

int mc3d_marching_cubes_search_batch(mc3d_ctx *ctx, const mc3d_solid *const *solids, const double *deltas,
                                     const int32_t *iters, size_t n, const mc3d_options *opts, mc3d_mesh **out) {
    if (!ctx || (n && (!solids || !deltas || !iters || !out))) {
        set_error("mc3d_marching_cubes_search_batch: null argument");
        return MC3D_E_INVALID_ARG;
    }
    for (size_t i = 0; i < n; i++) out[i] = nullptr;
    for (size_t i = 0; i < n; i++) {
        if (!solids[i]) {
            set_error("mc3d_marching_cubes_search_batch: solid " + std::to_string(i) + " is null");
            return MC3D_E_INVALID_ARG;
        }
    }
    if (n == 0) return MC3D_OK;
    const auto t_call = std::chrono::steady_clock::now();
    const mc3d_options o = default_options(opts);
    HIP_TRY(hipSetDevice(ctx->device));
    // Bounds, spacers and lowering of every solid, on up to 16 host threads
    // (solid trees are immutable; errors are kept per solid).
    std::vector<Job> all(n);
    std::vector<int> rcs(n, MC3D_OK);
    std::vector<std::string> errs(n);
    std::atomic<size_t> next{0};
    auto work = [&]() {
        for (size_t i; (i = next.fetch_add(1)) < n;) {
            all[i].s = solids[i], all[i].delta = deltas[i], all[i].iters = iters[i];
            rcs[i] = prepare_job(all[i], o);
            if (rcs[i]) errs[i] = mc3d_last_error();
        }
    };
    const size_t nthreads = std::min<size_t>({16, std::max(1u, std::thread::hardware_concurrency()), (n + 7) / 8});
    std::vector<std::thread> pool;
    for (size_t t = 1; t < nthreads; t++) pool.emplace_back(work);
    work();
    for (std::thread &t : pool) t.join();
    // Solids with host-evaluated atoms need their own callback rounds and run
    // alone; all others share one pass of the pipeline.
    std::vector<Job> batched, alone;
    std::vector<size_t> batched_at, alone_at;
    for (size_t i = 0; i < n; i++) {
        if (rcs[i]) {
            set_error("solid " + std::to_string(i) + " of the batch: " + errs[i]);
            return rcs[i];
        }
        Job &j = all[i];
        if (j.host) {
            alone.push_back(std::move(j));
            alone_at.push_back(i);
        } else {
            batched.push_back(std::move(j));
            batched_at.push_back(i);
        }
    }
    const double ms_prepare = ms_since(t_call);
    std::vector<mc3d_mesh *> got(n, nullptr);
    auto fail = [&](int rc) {
        for (mc3d_mesh *m : got) delete m;
        return rc;
    };
    if (!batched.empty()) {
        std::vector<mc3d_mesh *> ms;
        int rc = run_call(ctx, batched, o, ms, t_call, ms_prepare);
        if (rc) return fail(rc);
        for (size_t k = 0; k < ms.size(); k++) got[batched_at[k]] = ms[k];
    }
    for (size_t k = 0; k < alone.size(); k++) {
        std::vector<Job> one;
        one.push_back(std::move(alone[k]));
        std::vector<mc3d_mesh *> ms;
        int rc = run_call(ctx, one, o, ms, std::chrono::steady_clock::now(), 0.0);
        if (rc) {
            set_error("solid " + std::to_string(alone_at[k]) + " of the batch: " + mc3d_last_error());
            return fail(rc);
        }
        got[alone_at[k]] = ms[0];
    }
    for (size_t i = 0; i < n; i++) out[i] = got[i];
    return MC3D_OK;
}

int mc3d_mesh_counts(const mc3d_mesh *m, size_t *nv, size_t *nt) {
    if (!m) return MC3D_E_INVALID_ARG;
    if (nv) *nv = m->nv;
    if (nt) *nt = m->nt;
    return MC3D_OK;
}

int mc3d_mesh_copy(mc3d_mesh *m, double *verts, uint32_t *tris) {
    if (!m) return MC3D_E_INVALID_ARG;
    if (m->on_host) {
        if (verts && m->nv) parallel_copy(verts, m->hverts.get(), m->nv * 24);
        if (tris && m->nt) parallel_copy(tris, m->htris.get(), m->nt * 12);
        return MC3D_OK;
    }
    HIP_TRY(hipSetDevice(m->device));
    if (verts && m->nv)
        HIP_TRY(hipMemcpy(verts, m->buf->verts->as<double>() + 3 * m->voff, m->nv * 24, hipMemcpyDeviceToHost));
    if (tris && m->nt)
        HIP_TRY(hipMemcpy(tris, m->buf->tris->as<uint32_t>() + 3 * m->toff, m->nt * 12, hipMemcpyDeviceToHost));
    return MC3D_OK;
}

size_t mc3d_stl_bytes(size_t n_triangles) { return 84 + 50 * n_triangles; }

namespace {
// The binary STL of nt device triangles into out (host): the body by
// k_encode_stl into device scratch, one copy through pinned memory, the
// 84-byte header written here.
int encode_stl_device(mc3d_ctx *ctx, const double *verts, const uint32_t *tris, uint64_t nt, uint8_t *out,
                      size_t out_bytes) {
    if (nt >= (1ull << 32)) {
        set_error("too many triangles for STL format");  // export.go:38-40
        return MC3D_E_UNSUPPORTED;
    }
    if (out_bytes < mc3d_stl_bytes(nt)) {
        set_error("STL buffer too small: need " + std::to_string(mc3d_stl_bytes(nt)) + " bytes");
        return MC3D_E_INVALID_ARG;
    }
    memset(out, 0, 80);
    const uint32_t n32 = (uint32_t)nt;
    memcpy(out + 80, &n32, 4);  // little-endian host
    if (!nt) return MC3D_OK;
    const size_t body = 50 * nt;
    int rc;
    if ((rc = ctx->stl.ensure(body)) || (rc = ctx->outstage.ensure(body))) return rc;
    k_encode_stl<<<grid_for(nt, 256, 1u << 16), 256, 0, ctx->stream>>>(verts, tris, nt, ctx->stl.as<uint16_t>());
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipMemcpyAsync(ctx->outstage.p, ctx->stl.p, body, hipMemcpyDeviceToHost, ctx->stream));
    HIP_TRY(hipStreamSynchronize(ctx->stream));
    parallel_copy(out + 84, ctx->outstage.p, body);
    return MC3D_OK;
}
}  // namespace

int mc3d_mesh_encode_stl(mc3d_ctx *ctx, const mc3d_mesh *m, uint8_t *out, size_t out_bytes) {
    if (!ctx || !m || !out) {
        set_error("mc3d_mesh_encode_stl: null argument");
        return MC3D_E_INVALID_ARG;
    }
    if (m->device != ctx->device || !m->buf || !m->buf->verts || !m->buf->tris) {
        set_error("mc3d_mesh_encode_stl: the mesh is not on this context's device");
        return MC3D_E_INVALID_ARG;
    }
    HIP_TRY(hipSetDevice(ctx->device));
    // triangle indices count from the mesh's first vertex
    return encode_stl_device(ctx, m->buf->verts->as<double>() + 3 * m->voff,
                             m->buf->tris->as<uint32_t>() + 3 * m->toff, m->nt, out, out_bytes);
}

int mc3d_encode_stl(mc3d_ctx *ctx, const double *verts, size_t nv, const uint32_t *tris, size_t nt, uint8_t *out,
                    size_t out_bytes) {
    if (!ctx || !out || (nv && !verts) || (nt && !tris)) {
        set_error("mc3d_encode_stl: null argument");
        return MC3D_E_INVALID_ARG;
    }
    for (size_t i = 0; i < 3 * nt; i++)
        if (tris[i] >= nv) {
            set_error("mc3d_encode_stl: triangle " + std::to_string(i / 3) + " refers to vertex " +
                      std::to_string(tris[i]) + " of " + std::to_string(nv));
            return MC3D_E_INVALID_ARG;
        }
    HIP_TRY(hipSetDevice(ctx->device));
    int rc;
    if ((rc = ctx->stl_in.ensure(nv * 24 + nt * 12 + 16))) return rc;
    double *dv = ctx->stl_in.as<double>();
    uint32_t *dt = (uint32_t *)(ctx->stl_in.as<char>() + align_up(nv * 24, 16));
    if (nv) HIP_TRY(hipMemcpyAsync(dv, verts, nv * 24, hipMemcpyHostToDevice, ctx->stream));
    if (nt) HIP_TRY(hipMemcpyAsync(dt, tris, nt * 12, hipMemcpyHostToDevice, ctx->stream));
    return encode_stl_device(ctx, dv, dt, nt, out, out_bytes);
}

int mc3d_mesh_copy_interior(mc3d_mesh *m, double *interior) {
    if (!m || !m->has_interior) {
        set_error("mc3d_mesh_copy_interior: not a mesh of mc3d_marching_cubes_interior");
        return MC3D_E_INVALID_ARG;
    }
    if (!interior || !m->nv) return MC3D_OK;
    if (m->on_host) {
        parallel_copy(interior, m->hinterior.get(), m->nv * 24);
        return MC3D_OK;
    }
    HIP_TRY(hipSetDevice(m->device));
    HIP_TRY(hipMemcpy(interior, m->buf->interior->as<double>() + 3 * m->voff, m->nv * 24, hipMemcpyDeviceToHost));
    return MC3D_OK;
}

int mc3d_mesh_device_ptrs(const mc3d_mesh *m, const double **verts, const uint32_t **tris) {
    if (!m) return MC3D_E_INVALID_ARG;
    const bool have = m->buf && m->buf->verts && m->buf->tris;
    if (verts) *verts = have ? m->buf->verts->as<double>() + 3 * m->voff : nullptr;
    if (tris) *tris = have ? m->buf->tris->as<uint32_t>() + 3 * m->toff : nullptr;
    return MC3D_OK;
}

int mc3d_mesh_stats(const mc3d_mesh *m, mc3d_stats *out) {
    if (!m || !out) return MC3D_E_INVALID_ARG;
    *out = m->stats;
    return MC3D_OK;
}

void mc3d_mesh_free(mc3d_mesh *m) {
    if (!m) return;
    (void)hipSetDevice(m->device);
    delete m;
}

int mc3d_contains(mc3d_ctx *ctx, const mc3d_solid *s, const double *pts, size_t n, int32_t precision,
                  uint8_t *out) {
    if (!ctx || !s || (n && (!pts || !out))) {
        set_error("mc3d_contains: null argument");
        return MC3D_E_INVALID_ARG;
    }
    if (n == 0) return MC3D_OK;
    HIP_TRY(hipSetDevice(ctx->device));
    std::vector<Job> jobs(1);
    Job &j = jobs[0];
    j.s = s;
    std::string msg;
    if (!lower_solid(*s->ptr, j.P, msg)) {
        set_error(msg);
        return MC3D_E_UNSUPPORTED;
    }
    j.frames = host_frames(j.P);
    j.g = Grid{nullptr, nullptr, nullptr, 1, 1, 1, 1};
    int rc;
    Upload U;
    if ((rc = upload(ctx, jobs, 0, 0, U))) return rc;
    if ((rc = ctx->pts.ensure(n * 24)) || (rc = ctx->out8.ensure(n))) return rc;
    HIP_TRY(hipMemcpyAsync(ctx->pts.p, pts, n * 24, hipMemcpyHostToDevice, ctx->stream));
    if (!j.P.host_atoms.empty()) {
        double ms_cb = 0;
        if ((rc = host_round(ctx, j, U, 0, ctx->pts.as<double>(), n, ms_cb))) return rc;
    }
    const DevProg &dp = U.hsegs[0].P;
    if (precision == MC3D_PREC_FP32_GUARD)
        k_contains<1><<<grid_for(n, 256), 256, 0, ctx->stream>>>(dp, ctx->pts.as<double>(), ctx->out8.as<uint8_t>(), n);
    else
        k_contains<0><<<grid_for(n, 256), 256, 0, ctx->stream>>>(dp, ctx->pts.as<double>(), ctx->out8.as<uint8_t>(), n);
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipMemcpyAsync(out, ctx->out8.p, n, hipMemcpyDeviceToHost, ctx->stream));
    HIP_TRY(hipStreamSynchronize(ctx->stream));
    return MC3D_OK;
}

}  // extern "C"
