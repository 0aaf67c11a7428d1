// mc3d.hip -- host side of the MI355X marching-cubes pipeline and the C-ABI
// entry points for meshing (include/mc3d.h).  The device stages live in
// kernels.h, Contains in eval.h, the solid tree and its lowering in solid.cpp.
//
// One call = one pinned upload of the lowered program and grid axes, the
// stage kernels on the context's stream, and one synchronisation at the end;
// every count a stage needs is read on the device.  Output buffers are
// pooled per context and sized from earlier calls; a call whose mesh does
// not fit grows them and re-runs the two output stages.
//
// Every float64 expression keeps the reference's operation order and is
// compiled with -ffp-contract=off, so Contains decisions, topology and vertex
// coordinates are bit-identical to the reference's amd64 build.

#include <hip/hip_runtime.h>
#include <math.h>
#include <stdint.h>
#include <string.h>

#include <algorithm>
#include <chrono>
#include <memory>
#include <new>
#include <string>
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

unsigned grid_for(uint64_t n, unsigned block, unsigned cap = 1u << 20) {
    uint64_t b = (n + block - 1) / block;
    if (b == 0) b = 1;
    return (unsigned)(b < cap ? b : cap);
}

size_t align_up(size_t v, size_t a) { return (v + a - 1) / a * a; }

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
    HostBuf staging, ctl_host;
    DevBuf blob, table, bits, counts, offs, tiles, actw, actb, binfo, bcode, vkeys, pts, out8;
    DevBuf hvals, hpts, hstate, hprobe;  // host-evaluated atoms: answers, points, bisection state
    HostBuf hstage;                      // pinned points / answers of one callback batch
    std::shared_ptr<Pool> pool = std::make_shared<Pool>();
    uint64_t v_hint = 1u << 18, t_hint = 1u << 19;  // output capacities to start from
    hipEvent_t ev[8] = {};
    bool table_loaded = false;
};

struct mc3d_mesh {
    int device = 0;
    uint64_t nv = 0, nt = 0;
    std::vector<double> hverts;
    std::vector<uint32_t> htris;
    bool on_host = false;
    std::shared_ptr<Pool> pool;
    std::unique_ptr<DevBuf> dverts, dtris;  // device results owned by the mesh
    mc3d_stats stats{};
    ~mc3d_mesh() {
        if (pool) {
            pool->give(std::move(dverts));
            pool->give(std::move(dtris));
        }
    }
};

namespace {

struct Upload {
    DevProg dp;
    Ctl *ctl;
    double *axes;
};

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
    // Evict the oldest entries, never one the current program uses (those
    // are the newest `keep_at_least`, looked up or added in this upload).
    while (ctx->colliders.size() > std::max<size_t>(4, keep_at_least)) ctx->colliders.erase(ctx->colliders.begin());
    return MC3D_OK;
}

// Packs the lowered program, the grid axes and a zeroed control block into
// the pinned staging buffer and copies it to the device in one transfer.
int upload(mc3d_ctx *ctx, const Program &P, const std::vector<double> *ax, Upload &U) {
    std::vector<DevCollider> cols(P.colliders.size());
    for (size_t i = 0; i < cols.size(); i++) {
        int rc = collider_on_device(ctx, P.colliders[i], cols[i], cols.size());
        if (rc) return rc;
    }
    std::vector<AtomBox> boxes(P.code.size());
    for (size_t i = 0; i < P.atom_pc.size(); i++) boxes[P.atom_pc[i]] = P.atom_box[i];
    size_t off = 0;
    auto place = [&](size_t bytes) {
        size_t o = off;
        off = align_up(off + bytes, 256);
        return o;
    };
    const size_t o_ctl = place(sizeof(Ctl));
    const size_t o_code = place(P.code.size() * sizeof(Insn));
    const size_t o_f64 = place(P.f64.size() * sizeof(double));
    const size_t o_f32 = place(P.f32.size() * sizeof(float));
    const size_t o_fr = place(P.frames.size() * sizeof(Frame));
    const size_t o_box = place(boxes.size() * sizeof(AtomBox));
    const size_t o_col = place(cols.size() * sizeof(DevCollider));
    size_t n_ax = 0;
    if (ax) n_ax = ax[0].size() + ax[1].size() + ax[2].size();
    const size_t o_ax = place(n_ax * sizeof(double));
    const size_t total = off;
    int rc;
    if ((rc = ctx->staging.ensure(total)) || (rc = ctx->blob.ensure(total))) return rc;
    char *h = (char *)ctx->staging.p;
    memset(h + o_ctl, 0, sizeof(Ctl));
    memcpy(h + o_code, P.code.data(), P.code.size() * sizeof(Insn));
    if (!P.f64.empty()) {
        memcpy(h + o_f64, P.f64.data(), P.f64.size() * sizeof(double));
        memcpy(h + o_f32, P.f32.data(), P.f32.size() * sizeof(float));
    }
    memcpy(h + o_fr, P.frames.data(), P.frames.size() * sizeof(Frame));
    memcpy(h + o_box, boxes.data(), boxes.size() * sizeof(AtomBox));
    if (!cols.empty()) memcpy(h + o_col, cols.data(), cols.size() * sizeof(DevCollider));
    if (ax) {
        double *d = (double *)(h + o_ax);
        for (int a = 0; a < 3; a++) d = std::copy(ax[a].begin(), ax[a].end(), d);
    }
    HIP_TRY(hipMemcpyAsync(ctx->blob.p, h, total, hipMemcpyHostToDevice, ctx->stream));
    char *dv = (char *)ctx->blob.p;
    U.dp.code = (const Insn *)(dv + o_code);
    U.dp.f64 = (const double *)(dv + o_f64);
    U.dp.f32 = (const float *)(dv + o_f32);
    U.dp.frames = (const Frame *)(dv + o_fr);
    U.dp.boxes = (const AtomBox *)(dv + o_box);
    U.dp.colliders = (const DevCollider *)(dv + o_col);
    U.dp.host_vals = nullptr;
    U.dp.host_stride = 0;
    U.dp.ncode = (uint32_t)P.code.size();
    U.dp.pad = 0;
    U.ctl = (Ctl *)(dv + o_ctl);
    U.axes = (double *)(dv + o_ax);
    return MC3D_OK;
}

// Frame of each host-evaluated atom, by host atom index.
std::vector<uint32_t> host_frames(const Program &P) {
    std::vector<uint32_t> fr(P.host_atoms.size(), 0);
    for (uint32_t pc : P.atom_pc) {
        const Insn in = P.code[pc];
        if (insn_kind(in.x) == A_HOST) fr[(size_t)P.f64[in.y]] = insn_frame(in.x);
    }
    return fr;
}

constexpr uint64_t HOST_BATCH = 1u << 20;  // points per callback

// Answers every host-evaluated atom at n points -- grid corners (pts ==
// nullptr) or a device point list -- through its callback, in batches, into
// hvals[h * n + i], and points the program at them.  The GPU computes each
// atom's local points; the host only runs the callbacks.
int host_round(mc3d_ctx *ctx, const Program &P, const std::vector<uint32_t> &frames, Upload &U, const Grid &g,
               const double *pts, uint64_t n, double &ms_cb) {
    const size_t nh = P.host_atoms.size();
    int rc;
    if ((rc = ctx->hvals.ensure(nh * n + 1)) || (rc = ctx->hpts.ensure(HOST_BATCH * 24)) ||
        (rc = ctx->hstage.ensure(HOST_BATCH * 25)))
        return rc;
    double *hp = (double *)ctx->hstage.p;
    uint8_t *ho = (uint8_t *)(hp + 3 * HOST_BATCH);
    for (size_t h = 0; h < nh; h++) {
        for (uint64_t s = 0; s < n; s += HOST_BATCH) {
            const uint64_t c = std::min(HOST_BATCH, n - s);
            k_host_points<<<grid_for(c, 256), 256, 0, ctx->stream>>>(U.dp, g, pts, frames[h], s, c,
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
            HIP_TRY(hipMemcpyAsync(ctx->hvals.as<uint8_t>() + h * n + s, ho, c, hipMemcpyHostToDevice, ctx->stream));
        }
    }
    U.dp.host_vals = ctx->hvals.as<uint8_t>();
    U.dp.host_stride = n;
    return MC3D_OK;
}

std::string fmt_bounds(const double *mn, const double *mx) {
    char buf[512];
    snprintf(buf, sizeof buf,
             "invalid bounds: min=model3d.Coord3D{X:%.17g, Y:%.17g, Z:%.17g} max=model3d.Coord3D{X:%.17g, Y:%.17g, "
             "Z:%.17g}",
             mn[0], mn[1], mn[2], mx[0], mx[1], mx[2]);
    return buf;
}

float elapsed(hipEvent_t a, hipEvent_t b) {
    float ms = 0;
    if (hipEventElapsedTime(&ms, a, b) != hipSuccess) return 0;
    return ms;
}

int scan4(mc3d_ctx *ctx, const uint4 *in, uint4 *out, uint64_t n, unsigned long long *totals, uint32_t *list) {
    const uint64_t ntiles = (n + SCAN_TILE - 1) / SCAN_TILE;
    int rc;
    if ((rc = ctx->tiles.ensure(ntiles * sizeof(uint4)))) return rc;
    k_scan_reduce<<<(unsigned)ntiles, SCAN_BLOCK, 0, ctx->stream>>>(in, n, ctx->tiles.as<uint4>());
    k_scan_tiles<<<1, SCAN_BLOCK, 0, ctx->stream>>>(ctx->tiles.as<uint4>(), ntiles, totals);
    k_scan_down<<<(unsigned)ntiles, SCAN_BLOCK, 0, ctx->stream>>>(in, n, ctx->tiles.as<uint4>(), out, list);
    HIP_TRY(hipGetLastError());
    return MC3D_OK;
}

template <int MODE, bool STATS>
void launch_corners(mc3d_ctx *ctx, const DevProg &dp, const Bricks &B, const Grid &g, Ctl *ctl, uint64_t nwords,
                    uint64_t nbricks, int rows_per_walk) {
    if (!B.info) {
        k_corners<MODE, STATS><<<grid_for(nwords, 4), 256, 0, ctx->stream>>>(dp, g, ctx->bits.as<uint64_t>(), ctl,
                                                                             nwords);
        return;
    }
    k_fill_const<<<grid_for(nwords, 256), 256, 0, ctx->stream>>>(B, g, ctx->bits.as<uint64_t>(), ctl, nwords);
    const unsigned grid = grid_for(nbricks, 1, 1u << 15);
    if (rows_per_walk >= 4)
        k_corners_act<MODE, STATS, 4><<<grid, 256, 0, ctx->stream>>>(dp, B, g, ctx->actb.as<uint32_t>(),
                                                                     ctx->bits.as<uint64_t>(), ctl);
    else if (rows_per_walk == 2)
        k_corners_act<MODE, STATS, 2><<<grid, 256, 0, ctx->stream>>>(dp, B, g, ctx->actb.as<uint32_t>(),
                                                                     ctx->bits.as<uint64_t>(), ctl);
    else
        k_corners_act<MODE, STATS, 1><<<grid, 256, 0, ctx->stream>>>(dp, B, g, ctx->actb.as<uint32_t>(),
                                                                     ctx->bits.as<uint64_t>(), ctl);
}

template <int MODE, bool STATS>
void launch_bisect(mc3d_ctx *ctx, const DevProg &dp, const Bricks &B, const Grid &g, Ctl *ctl, double *verts,
                   uint64_t vcap, int iters) {
    k_bisect<MODE, STATS><<<grid_for(vcap, 256, 1u << 14), 256, 0, ctx->stream>>>(
        dp, B, g, ctx->bits.as<uint64_t>(), ctx->vkeys.as<uint64_t>(), verts, ctl, vcap, iters);
}

// The output stages: emit (vertex keys, triangles) and bisect (vertices).
int run_output(mc3d_ctx *ctx, const Upload &U, const Bricks &B, const Grid &g, mc3d_mesh *m, int mode, bool stats,
               int iters, uint64_t vcap, uint64_t tcap, uint64_t nwords) {
    HIP_TRY(hipEventRecord(ctx->ev[4], ctx->stream));
    k_emit<<<grid_for(nwords, 4, 2048), 256, 0, ctx->stream>>>(
        g, ctx->bits.as<uint64_t>(), ctx->table.as<uint8_t>(), ctx->offs.as<uint4>(), ctx->actw.as<uint32_t>(),
        U.ctl, ctx->vkeys.as<uint64_t>(), m->dtris->as<uint32_t>(), vcap, tcap);
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipEventRecord(ctx->ev[5], ctx->stream));
    double *verts = m->dverts->as<double>();
    if (mode == 0)
        stats ? launch_bisect<0, true>(ctx, U.dp, B, g, U.ctl, verts, vcap, iters)
              : launch_bisect<0, false>(ctx, U.dp, B, g, U.ctl, verts, vcap, iters);
    else
        stats ? launch_bisect<1, true>(ctx, U.dp, B, g, U.ctl, verts, vcap, iters)
              : launch_bisect<1, false>(ctx, U.dp, B, g, U.ctl, verts, vcap, iters);
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipEventRecord(ctx->ev[6], ctx->stream));
    return MC3D_OK;
}

int take_outputs(mc3d_ctx *ctx, mc3d_mesh *m, uint64_t vcap, uint64_t tcap);

// The output stages of a program with host-evaluated atoms: the mesh size is
// read first (one more synchronisation), then emit, then the bisection in
// rounds with the host answering its atoms between them.  Leaves the final
// control block in ctx->ctl_host.
int run_output_host(mc3d_ctx *ctx, const Program &P, const std::vector<uint32_t> &frames, Upload &U, const Grid &g,
                    mc3d_mesh *m, int mode, int iters, uint64_t nwords, double &ms_cb) {
    Ctl *hc = (Ctl *)ctx->ctl_host.p;
    HIP_TRY(hipMemcpyAsync(hc, U.ctl, sizeof(Ctl), hipMemcpyDeviceToHost, ctx->stream));
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
    if ((rc = take_outputs(ctx, m, vcap, tcap)) || (rc = ctx->hstate.ensure(vcap * 16)) ||
        (rc = ctx->hprobe.ensure(vcap * 24)))
        return rc;
    HIP_TRY(hipEventRecord(ctx->ev[4], ctx->stream));
    k_emit<<<grid_for(nwords, 4, 2048), 256, 0, ctx->stream>>>(
        g, ctx->bits.as<uint64_t>(), ctx->table.as<uint8_t>(), ctx->offs.as<uint4>(), ctx->actw.as<uint32_t>(),
        U.ctl, ctx->vkeys.as<uint64_t>(), m->dtris->as<uint32_t>(), vcap, tcap);
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipEventRecord(ctx->ev[5], ctx->stream));
    double *verts = m->dverts->as<double>(), *state = ctx->hstate.as<double>(), *probe = ctx->hprobe.as<double>();
    const unsigned grid = grid_for(nv, 256, 1u << 14);
    for (int round = 0; round <= iters; round++) {
        if (round > 0 && (rc = host_round(ctx, P, frames, U, g, probe, nv, ms_cb))) return rc;
        if (mode == 1)
            k_bisect_host<1><<<grid, 256, 0, ctx->stream>>>(U.dp, g, ctx->bits.as<uint64_t>(),
                                                            ctx->vkeys.as<uint64_t>(), state, probe, verts, nv,
                                                            round > 0);
        else
            k_bisect_host<0><<<grid, 256, 0, ctx->stream>>>(U.dp, g, ctx->bits.as<uint64_t>(),
                                                            ctx->vkeys.as<uint64_t>(), state, probe, verts, nv,
                                                            round > 0);
        HIP_TRY(hipGetLastError());
    }
    HIP_TRY(hipEventRecord(ctx->ev[6], ctx->stream));
    HIP_TRY(hipMemcpyAsync(hc, U.ctl, sizeof(Ctl), hipMemcpyDeviceToHost, ctx->stream));
    HIP_TRY(hipStreamSynchronize(ctx->stream));
    return MC3D_OK;
}

int take_outputs(mc3d_ctx *ctx, mc3d_mesh *m, uint64_t vcap, uint64_t tcap) {
    int rc;
    if (!m->dverts || m->dverts->cap < vcap * 24) {
        ctx->pool->give(std::move(m->dverts));
        m->dverts = ctx->pool->take(vcap * 24);
    }
    if (!m->dtris || m->dtris->cap < tcap * 12) {
        ctx->pool->give(std::move(m->dtris));
        m->dtris = ctx->pool->take(tcap * 12);
    }
    if ((rc = m->dverts->ensure(vcap * 24)) || (rc = m->dtris->ensure(tcap * 12)) ||
        (rc = ctx->vkeys.ensure(vcap * 8)))
        return rc;
    return MC3D_OK;
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
    if (c->ctl_host.ensure(sizeof(Ctl)) || c->table.ensure(256 * 16) ||
        hipMemcpy(c->table.p, case_table(), 256 * 16, hipMemcpyHostToDevice) != hipSuccess) {
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

int mc3d_marching_cubes_search(mc3d_ctx *ctx, const mc3d_solid *s, double delta, int iters,
                               const mc3d_options *opts, mc3d_mesh **out) {
    if (!ctx || !s || !out) {
        set_error("mc3d_marching_cubes_search: null argument");
        return MC3D_E_INVALID_ARG;
    }
    *out = nullptr;
    const auto t_call = std::chrono::steady_clock::now();
    if (iters < 0) iters = 0;
    mc3d_options o{};
    o.precision = MC3D_PREC_FP32_GUARD;
    o.cull = 1;
    if (opts) o = *opts;
    const int mode = o.precision == MC3D_PREC_FP32_GUARD ? 1 : 0;
    const bool stats = o.count_mismatch != 0;
    HIP_TRY(hipSetDevice(ctx->device));

    // validateBounds, mc.go:15 / bounder.go:48-52
    double mn[3], mx[3];
    solid_bounds(*s->ptr, mn, mx);
    if (!bounds_valid(*s->ptr)) {
        set_error(fmt_bounds(mn, mx));
        return MC3D_E_INVALID_BOUNDS;
    }
    if (!(delta > 0)) {
        set_error("delta must be positive");
        return MC3D_E_INVALID_ARG;
    }
    // newSquareSpacer, mc.go:594-608
    std::vector<double> ax[3];
    for (int a = 0; a < 3; a++) {
        ax[a] = spacer_axis(mn[a], mx[a], delta);
        if (ax[a].size() >= (1u << 20)) {
            set_error("grid too large: more than 2^20 samples on an axis");
            return MC3D_E_UNSUPPORTED;
        }
        if (ax[a].size() < 2) {
            set_error("grid has fewer than two samples on an axis");
            return MC3D_E_INVALID_ARG;
        }
    }
    Program P;
    std::string msg;
    if (!lower_solid(*s->ptr, P, msg)) {
        set_error(msg);
        return MC3D_E_UNSUPPORTED;
    }
    int rc;
    // A host-evaluated solid may be true outside its bounds, so nothing is
    // folded by bounds: no culling.
    const bool host = !P.host_atoms.empty();
    const std::vector<uint32_t> frames = host_frames(P);
    if (host) o.cull = 0;

    Grid g;
    g.nx = (int)ax[0].size(), g.ny = (int)ax[1].size(), g.nz = (int)ax[2].size();
    g.W = (g.nx + 63) / 64;
    const uint64_t nwords = (uint64_t)g.W * g.ny * g.nz;
    if (nwords >= (1ull << 32)) {
        set_error("grid too large: more than 2^32 occupancy words");
        return MC3D_E_UNSUPPORTED;
    }

    // Bricks: 64 x B x B corners; B grows until the fixed program slots fit.
    Bricks B{nullptr, nullptr, 1, 1, 1, 1};
    uint64_t nb = 0;
    uint32_t cap = 0;
    if (o.cull) {
        int bsz = o.brick > 0 ? o.brick : 4;
        cap = std::min<uint32_t>((uint32_t)P.code.size(), 64u);
        for (;;) {
            B.BY = B.BZ = bsz;
            B.nby = (g.ny + bsz - 1) / bsz;
            B.nbz = (g.nz + bsz - 1) / bsz;
            nb = (uint64_t)g.W * B.nby * B.nbz;
            if (nb * cap < (1ull << 27) || bsz >= 64) break;
            bsz *= 2;
        }
        if (nb * cap >= (1ull << 32)) {
            set_error("too many culling bricks");
            return MC3D_E_UNSUPPORTED;
        }
    }

    Upload U;
    if ((rc = upload(ctx, P, ax, U))) return rc;
    g.xs = U.axes, g.ys = U.axes + g.nx, g.zs = U.axes + g.nx + g.ny;
    if ((rc = ctx->bits.ensure(nwords * 8)) || (rc = ctx->counts.ensure(nwords * sizeof(uint4))) ||
        (rc = ctx->offs.ensure(nwords * sizeof(uint4))) || (rc = ctx->actw.ensure(nwords * sizeof(uint32_t))))
        return rc;

    std::unique_ptr<mc3d_mesh> mesh(new (std::nothrow) mc3d_mesh);
    mc3d_mesh *m = mesh.get();
    if (!m) return MC3D_E_OOM;
    m->device = ctx->device;
    m->pool = ctx->pool;
    mc3d_stats &st = m->stats;
    if (host && (rc = host_round(ctx, P, frames, U, g, nullptr, (uint64_t)g.nx * g.ny * g.nz, st.ms_host_callbacks)))
        return rc;

    // ---- plan ----
    HIP_TRY(hipEventRecord(ctx->ev[0], ctx->stream));
    if (o.cull) {
        if ((rc = ctx->binfo.ensure(nb * sizeof(uint2))) || (rc = ctx->actb.ensure(nb * sizeof(uint32_t))) ||
            (rc = ctx->bcode.ensure(nb * cap * sizeof(Insn) + 16)))
            return rc;
        // Stage the program in LDS when it fits beside the per-thread stacks.
        const size_t staged = PLAN_STACK_BYTES + P.code.size() * (sizeof(Insn) + sizeof(AtomBox));
        const int stage = staged <= 64 * 1024 ? 1 : 0;
        k_plan<<<grid_for(nb, PLAN_BLOCK), PLAN_BLOCK, stage ? staged : PLAN_STACK_BYTES, ctx->stream>>>(
            U.dp, g, B, nb, cap, ctx->binfo.as<uint2>(), ctx->bcode.as<Insn>(), ctx->actb.as<uint32_t>(), U.ctl,
            stage);
        HIP_TRY(hipGetLastError());
        B.info = ctx->binfo.as<uint2>();
        B.code = ctx->bcode.as<Insn>();
        st.bricks_total = nb;
    }
    HIP_TRY(hipEventRecord(ctx->ev[1], ctx->stream));

    // ---- corners ----
    const int rpw = o.rows_per_walk > 0 ? o.rows_per_walk : 1;
    if (mode == 0)
        stats ? launch_corners<0, true>(ctx, U.dp, B, g, U.ctl, nwords, nb, rpw)
              : launch_corners<0, false>(ctx, U.dp, B, g, U.ctl, nwords, nb, rpw);
    else
        stats ? launch_corners<1, true>(ctx, U.dp, B, g, U.ctl, nwords, nb, rpw)
              : launch_corners<1, false>(ctx, U.dp, B, g, U.ctl, nwords, nb, rpw);
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipEventRecord(ctx->ev[2], ctx->stream));

    // ---- classify + scan (+ list of words that emit) ----
    k_classify<<<grid_for(nwords, 256), 256, 0, ctx->stream>>>(g, ctx->bits.as<uint64_t>(), ctx->table.as<uint8_t>(),
                                                                 ctx->counts.as<uint4>(), nwords);
    HIP_TRY(hipGetLastError());
    if ((rc = scan4(ctx, ctx->counts.as<uint4>(), ctx->offs.as<uint4>(), nwords, U.ctl->totals,
                    ctx->actw.as<uint32_t>())))
        return rc;
    HIP_TRY(hipEventRecord(ctx->ev[3], ctx->stream));

    // ---- emit + bisect, sized from earlier calls ----
    Ctl *hc = (Ctl *)ctx->ctl_host.p;
    if (host) {
        if ((rc = run_output_host(ctx, P, frames, U, g, m, mode, iters, nwords, st.ms_host_callbacks))) return rc;
    } else {
        uint64_t vcap = ctx->v_hint, tcap = ctx->t_hint;
        if ((rc = take_outputs(ctx, m, vcap, tcap))) return rc;
        if ((rc = run_output(ctx, U, B, g, m, mode, stats, iters, vcap, tcap, nwords))) return rc;
        HIP_TRY(hipMemcpyAsync(hc, U.ctl, sizeof(Ctl), hipMemcpyDeviceToHost, ctx->stream));
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
        if (nv > vcap || nt > tcap) {  // grow and run the output stages again
            ctx->v_hint = std::max(ctx->v_hint, nv + nv / 4 + 1024);
            ctx->t_hint = std::max(ctx->t_hint, nt + nt / 4 + 1024);
            vcap = ctx->v_hint, tcap = ctx->t_hint;
            if ((rc = take_outputs(ctx, m, vcap, tcap))) return rc;
            if ((rc = run_output(ctx, U, B, g, m, mode, stats, iters, vcap, tcap, nwords))) return rc;
            HIP_TRY(hipMemcpyAsync(hc, U.ctl, sizeof(Ctl), hipMemcpyDeviceToHost, ctx->stream));
            HIP_TRY(hipStreamSynchronize(ctx->stream));
        }
    }
    const uint64_t nv = hc->totals[0], nt = hc->totals[1];
    m->nv = nv;
    m->nt = nt;

    st.corners = (uint64_t)g.nx * g.ny * g.nz;
    st.vertices = nv;
    st.triangles = nt;
    st.bisection_evals = iters > 0 ? nv * (uint64_t)(iters + 1) : 0;
    st.bricks_active = hc->nact_bricks;
    st.guard_evals = hc->gstats[0];
    st.guard_rechecks = hc->gstats[1];
    st.guard_mismatches = hc->gstats[2];
    st.guard_violations = hc->gstats[3];
    st.ms_plan = elapsed(ctx->ev[0], ctx->ev[1]);
    st.ms_corners = elapsed(ctx->ev[1], ctx->ev[2]);
    st.ms_classify = elapsed(ctx->ev[2], ctx->ev[3]);
    st.ms_emit = elapsed(ctx->ev[4], ctx->ev[5]);
    st.ms_bisect = elapsed(ctx->ev[5], ctx->ev[6]);
    st.ms_device = elapsed(ctx->ev[0], ctx->ev[6]);
    if (!o.device_only) {
        auto t0 = std::chrono::steady_clock::now();
        m->hverts.resize(nv * 3);
        m->htris.resize(nt * 3);
        if (nv) HIP_TRY(hipMemcpy(m->hverts.data(), m->dverts->p, nv * 24, hipMemcpyDeviceToHost));
        if (nt) HIP_TRY(hipMemcpy(m->htris.data(), m->dtris->p, nt * 12, hipMemcpyDeviceToHost));
        m->on_host = true;
        st.ms_d2h = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    }
    st.ms_call = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t_call).count();
    *out = mesh.release();
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
        if (verts && m->nv) memcpy(verts, m->hverts.data(), m->nv * 24);
        if (tris && m->nt) memcpy(tris, m->htris.data(), m->nt * 12);
        return MC3D_OK;
    }
    HIP_TRY(hipSetDevice(m->device));
    if (verts && m->nv) HIP_TRY(hipMemcpy(verts, m->dverts->p, m->nv * 24, hipMemcpyDeviceToHost));
    if (tris && m->nt) HIP_TRY(hipMemcpy(tris, m->dtris->p, m->nt * 12, hipMemcpyDeviceToHost));
    return MC3D_OK;
}

int mc3d_mesh_device_ptrs(const mc3d_mesh *m, const double **verts, const uint32_t **tris) {
    if (!m) return MC3D_E_INVALID_ARG;
    if (verts) *verts = m->dverts ? m->dverts->as<double>() : nullptr;
    if (tris) *tris = m->dtris ? m->dtris->as<uint32_t>() : nullptr;
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
    Program P;
    std::string msg;
    if (!lower_solid(*s->ptr, P, msg)) {
        set_error(msg);
        return MC3D_E_UNSUPPORTED;
    }
    int rc;
    Upload U;
    if ((rc = upload(ctx, P, nullptr, U))) return rc;
    if ((rc = ctx->pts.ensure(n * 24)) || (rc = ctx->out8.ensure(n))) return rc;
    HIP_TRY(hipMemcpyAsync(ctx->pts.p, pts, n * 24, hipMemcpyHostToDevice, ctx->stream));
    if (!P.host_atoms.empty()) {
        double ms_cb = 0;
        const Grid none{nullptr, nullptr, nullptr, 1, 1, 1, 1};
        if ((rc = host_round(ctx, P, host_frames(P), U, none, ctx->pts.as<double>(), n, ms_cb))) return rc;
    }
    if (precision == MC3D_PREC_FP32_GUARD)
        k_contains<1><<<grid_for(n, 256), 256, 0, ctx->stream>>>(U.dp, ctx->pts.as<double>(), ctx->out8.as<uint8_t>(), n);
    else
        k_contains<0><<<grid_for(n, 256), 256, 0, ctx->stream>>>(U.dp, ctx->pts.as<double>(), ctx->out8.as<uint8_t>(), n);
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipMemcpyAsync(out, ctx->out8.p, n, hipMemcpyDeviceToHost, ctx->stream));
    HIP_TRY(hipStreamSynchronize(ctx->stream));
    return MC3D_OK;
}

}  // extern "C"
