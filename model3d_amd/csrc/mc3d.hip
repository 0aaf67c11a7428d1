// mc3d.hip -- host side of the MI355X marching-cubes pipeline and the C-ABI
// entry points for meshing (include/mc3d.h).  The device stages live in
// kernels.h (launched from k_*.hip through launch.h), Contains in eval.h, the
// solid tree and its lowering in solid.cpp.
//
// A call meshes a list of solids (one for mc3d_marching_cubes_search, many
// for mc3d_marching_cubes_search_batch): one upload of every lowered program,
// the grid axes and a zeroed control block, then one launch per stage for all
// solids on the context's stream and one synchronisation at the end; every
// count a stage needs is read on the device.  Output buffers are pooled per
// context and sized from earlier calls; a call whose meshes do not fit grows
// them and re-runs the two output stages.  Meshes wanted on the host land in
// pooled pinned memory: the triangles are copied while the bisection runs.
//
// Every float64 expression keeps the reference's operation order and is
// compiled with -ffp-contract=off, so Contains decisions, topology and vertex
// coordinates are bit-identical to the reference's amd64 build.

#include <hip/hip_runtime.h>
#include <math.h>
#include <stdint.h>
#include <string.h>

#include <algorithm>
#include <array>
#include <atomic>
#include <chrono>
#include <condition_variable>
#include <functional>
#include <exception>
#include <map>
#include <memory>
#include <mutex>
#include <new>
#include <string>
#include <thread>
#include <vector>

#include "launch.h"
#include "mc3d_internal.h"
#include "plan.h"

using namespace mc3d;

// ---------------------------------------------------------------------------
// Scratch of the allocation-free entry point (mc3d_marching_cubes_search_noalloc):
// every device buffer that call needs beyond its inputs and outputs -- the
// program blob, occupancy, counts, offsets, the word and brick lists, the
// brick programs and the vertex keys -- is carved from this array, except its
// last SCRATCH_AGG_BYTES, which hold classify's one-pass scan sums (zeroed
// once).  Each GPU
// holds its own copy until the process ends.  In host code `scratch` is only
// a stand-in for the symbol: its device address on the current device comes
// from hipGetSymbolAddress (scratch_address), and only that address is used.
// ---------------------------------------------------------------------------
constexpr size_t SCRATCH_BYTES = 64u << 20;
constexpr size_t SCRATCH_AGG_BYTES = 4096;
static __device__ __align__(256) char scratch[SCRATCH_BYTES];

namespace {

constexpr int MAX_DEVICES = 64;
std::mutex g_scratch_mu;
void *g_scratch_ptr[MAX_DEVICES];              // device address of `scratch` per device (null: not looked up)
uint64_t g_scratch_agg_gen[MAX_DEVICES];       // epoch generation its scan sums were zeroed for
std::mutex g_noalloc_mu[MAX_DEVICES];          // the allocation-free calls of one device run one at a time

// The device address of `scratch` on `dev`, which must be the current
// device; looked up once per device.
int scratch_address(int dev, void **out) {
    if (dev < 0 || dev >= MAX_DEVICES) {
        set_error("device index beyond the scratch table");
        return MC3D_E_UNSUPPORTED;
    }
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        void *p = nullptr;
        hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p) {
            set_error(std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
            return MC3D_E_HIP;
        }
        // the one-pass scan sums start zeroed (epoch generation 0)
        if ((e = hipMemset((char *)p + SCRATCH_BYTES - SCRATCH_AGG_BYTES, 0, SCRATCH_AGG_BYTES)) != hipSuccess ||
            (e = hipDeviceSynchronize()) != hipSuccess) {
            set_error(std::string("hipMemset(scratch): ") + hipGetErrorString(e));
            return MC3D_E_HIP;
        }
        g_scratch_ptr[dev] = p;
    }
    *out = g_scratch_ptr[dev];
    return MC3D_OK;
}

// Scan epochs: every call tags its classify blocks' one-pass sums with an
// epoch no other call of the process uses (1 .. SCAN_EPOCH_MAX - 1).  When
// the sequence starts over its generation grows, and a sums array is zeroed
// again before its first use in the new generation.
std::atomic<uint64_t> g_epoch_count{0};
void next_epoch(uint32_t &epoch, uint64_t &gen) {
    const uint64_t n = g_epoch_count.fetch_add(1);
    epoch = 1 + (uint32_t)(n % (SCAN_EPOCH_MAX - 1));
    gen = n / (SCAN_EPOCH_MAX - 1);
}

#define HIP_TRY(expr)                                                                        \
    do {                                                                                     \
        hipError_t e_ = (expr);                                                              \
        if (e_ != hipSuccess) {                                                              \
            set_error(std::string(#expr) + ": " + hipGetErrorString(e_));                    \
            return MC3D_E_HIP;                                                               \
        }                                                                                    \
    } while (0)

// Runs an entry point's body; a C++ exception (an allocation that failed
// inside a container, a thread that could not start) becomes a status
// instead of unwinding through the C ABI.
template <class F>
int guarded(F &&f) {
    try {
        return f();
    } catch (const std::bad_alloc &) {
        set_error("out of host memory");
        return MC3D_E_OOM;
    } catch (const std::exception &e) {
        set_error(std::string("internal error: ") + e.what());
        return MC3D_E_INTERNAL;
    } catch (...) {
        set_error("internal error");
        return MC3D_E_INTERNAL;
    }
}

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

// Pinned host buffer (grow-only).
struct HostBuf {
    void *p = nullptr;
    size_t cap = 0;
    HostBuf() = default;
    HostBuf(const HostBuf &) = delete;
    HostBuf &operator=(const HostBuf &) = delete;
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
    template <class T>
    T *as() const {
        return (T *)p;
    }
};

// Buffers released by freed meshes, reused by later calls (device buffers,
// or pinned host buffers): the most recent 32, up to 8 GiB together (a
// z-slab search holds a mesh per slab, and freeing device memory
// synchronises the device).
template <class B>
struct Pool {
    std::mutex mu;  // meshes may be freed on any thread
    std::vector<std::unique_ptr<B>> free;
    std::unique_ptr<B> take(size_t bytes) {
        std::lock_guard<std::mutex> lk(mu);
        size_t best = free.size();
        for (size_t i = 0; i < free.size(); i++)
            if (free[i]->cap >= bytes && (best == free.size() || free[i]->cap < free[best]->cap)) best = i;
        if (best < free.size()) {
            std::unique_ptr<B> b = std::move(free[best]);
            free.erase(free.begin() + best);
            return b;
        }
        return std::unique_ptr<B>(new B);
    }
    void give(std::unique_ptr<B> b) {
        if (!b || !b->p) return;
        std::lock_guard<std::mutex> lk(mu);
        free.push_back(std::move(b));
        size_t bytes = 0;
        for (const auto &f : free) bytes += f->cap;
        while (free.size() > 32 || (free.size() > 1 && bytes > (8ull << 30))) {
            bytes -= free.front()->cap;
            free.erase(free.begin());
        }
    }
};

// memcpy of large host buffers split over up to 16 threads (the first touch
// of freshly allocated pages dominates a single-threaded copy).
// Host threads kept for the library's parallel host loops (lowering a
// batch's solids, filling its upload): run(n, fn) calls fn(i) for every
// i < n on up to 16 threads, the caller's included, and returns when all
// are done.  Runs from several threads take turns.
class HostPool {
  public:
    static HostPool &get() {
        static HostPool pool;
        return pool;
    }
    void run(size_t n, const std::function<void(size_t)> &fn) {
        if (n == 0) return;
        if (n == 1 || workers_.empty()) {
            for (size_t i = 0; i < n; i++) fn(i);
            return;
        }
        std::lock_guard<std::mutex> one(run_mu_);
        {
            std::lock_guard<std::mutex> lk(mu_);
            fn_ = &fn, n_ = n, next_ = 0, busy_ = workers_.size(), gen_++;
        }
        cv_.notify_all();
        drain();
        std::unique_lock<std::mutex> lk(mu_);
        done_cv_.wait(lk, [&] { return busy_ == 0; });
        fn_ = nullptr;
    }

  private:
    HostPool() {
        const unsigned hw = std::max(1u, std::thread::hardware_concurrency());
        for (unsigned t = 1; t < std::min(16u, hw); t++) workers_.emplace_back([this] { loop(); });
    }
    ~HostPool() {
        {
            std::lock_guard<std::mutex> lk(mu_);
            stop_ = true;
        }
        cv_.notify_all();
        for (std::thread &t : workers_) t.join();
    }
    void drain() {
        for (size_t i; (i = next_.fetch_add(1)) < n_;) (*fn_)(i);
    }
    void loop() {
        uint64_t seen = 0;
        for (;;) {
            {
                std::unique_lock<std::mutex> lk(mu_);
                cv_.wait(lk, [&] { return stop_ || gen_ != seen; });
                if (stop_) return;
                seen = gen_;
            }
            drain();
            std::lock_guard<std::mutex> lk(mu_);
            if (--busy_ == 0) done_cv_.notify_all();
        }
    }
    std::mutex run_mu_, mu_;
    std::condition_variable cv_, done_cv_;
    const std::function<void(size_t)> *fn_ = nullptr;
    size_t n_ = 0, busy_ = 0;
    std::atomic<size_t> next_{0};
    uint64_t gen_ = 0;
    bool stop_ = false;
    std::vector<std::thread> workers_;
};

void parallel_copy(void *dst, const void *src, size_t bytes) {
    constexpr size_t CHUNK = 1u << 20;
    const size_t nt = std::min<size_t>(16, (bytes + CHUNK - 1) / CHUNK);
    if (nt <= 1) {
        if (bytes) memcpy(dst, src, bytes);
        return;
    }
    const size_t per = (bytes + nt - 1) / nt;
    HostPool::get().run(nt, [=](size_t i) {
        const size_t a = i * per, b = std::min(bytes, a + per);
        if (a < b) memcpy((char *)dst + a, (const char *)src + a, b - a);
    });
}

size_t align_up(size_t v, size_t a) { return (v + a - 1) / a * a; }

// Device arrays into caller (pageable) host memory: 4 MiB chunks through a
// ring of four pinned buffers on a stream of their own, each chunk copied
// on by eight host threads while the next chunks' DMA runs, so the whole
// costs about the DMA alone (a direct copy to pageable memory goes through
// the runtime's own staging at about a third of the rate, and one host
// thread copying beside the DMA halves it).  One ring per device, made on
// first use and kept until the process ends; calls on one device take
// turns.
struct CopySpan {
    void *dst;
    const void *src;
    size_t bytes;
};

int copy_to_caller(int device, const std::vector<CopySpan> &spans) {
    constexpr size_t CH = 4u << 20, PART = CH / 8;
    constexpr int R = 4;
    struct Ring {
        std::mutex mu;
        hipStream_t st = nullptr;
        hipEvent_t ev[R] = {};
        void *buf[R] = {};
    };
    static std::mutex rings_mu;
    static std::vector<Ring *> rings;  // never freed: the runtime may be gone at exit
    Ring *r;
    {
        std::lock_guard<std::mutex> lk(rings_mu);
        if ((size_t)device >= rings.size()) rings.resize(device + 1, nullptr);
        if (!rings[device]) rings[device] = new Ring;
        r = rings[device];
    }
    std::lock_guard<std::mutex> lk(r->mu);
    HIP_TRY(hipSetDevice(device));
    if (!r->st) {
        HIP_TRY(hipStreamCreateWithFlags(&r->st, hipStreamNonBlocking));
        for (int i = 0; i < R; i++) {
            HIP_TRY(hipEventCreateWithFlags(&r->ev[i], hipEventDisableTiming));
            HIP_TRY(hipHostMalloc(&r->buf[i], CH, hipHostMallocDefault));
        }
    }
    std::vector<CopySpan> ch;
    for (const CopySpan &s : spans)
        for (size_t o = 0; o < s.bytes; o += CH)
            ch.push_back({(char *)s.dst + o, (const char *)s.src + o, std::min(CH, s.bytes - o)});
    auto issue = [&](size_t k) -> int {
        HIP_TRY(hipMemcpyAsync(r->buf[k % R], ch[k].src, ch[k].bytes, hipMemcpyDeviceToHost, r->st));
        HIP_TRY(hipEventRecord(r->ev[k % R], r->st));
        return MC3D_OK;
    };
    int rc;
    size_t next = 0;
    for (; next < std::min<size_t>(R, ch.size()); next++)
        if ((rc = issue(next))) return rc;
    for (size_t k = 0; k < ch.size(); k++) {
        HIP_TRY(hipEventSynchronize(r->ev[k % R]));
        const CopySpan c = ch[k];
        const void *src = r->buf[k % R];
        HostPool::get().run((c.bytes + PART - 1) / PART, [=](size_t i) {
            const size_t a = i * PART, n = std::min(PART, c.bytes - a);
            memcpy((char *)c.dst + a, (const char *)src + a, n);
        });
        if (next < ch.size() && (rc = issue(next++))) {  // into the buffer just emptied
            (void)hipStreamSynchronize(r->st);
            return rc;
        }
    }
    return MC3D_OK;
}

// The outputs of one call (vertices, triangles), shared by the meshes it
// returned and given back to the context's pools with the last of them.
// `host` holds the meshes' copies in pinned memory: all vertices, then all
// triangles, then all interior points.
struct MeshBuffers {
    std::shared_ptr<Pool<DevBuf>> pool;
    std::shared_ptr<Pool<HostBuf>> hpool;
    std::unique_ptr<DevBuf> verts, tris, interior;  // interior: MarchingCubesInterior only
    std::unique_ptr<HostBuf> host;
    uint64_t nv = 0, nt = 0;  // of the whole call (the layout of `host`)
    ~MeshBuffers() {
        if (pool) {
            pool->give(std::move(verts));
            pool->give(std::move(tris));
            pool->give(std::move(interior));
        }
        if (hpool) hpool->give(std::move(host));
    }
};

}  // namespace

// A collider's ray grid on the device, kept per context (keyed by the
// collider's id; the shared_ptr keeps the id from being reused).
struct ColliderCache {
    std::shared_ptr<const ColliderData> keep;
    double R = 0;  // radius of its sphere queries (0: none)
    DevBuf rec, start, tris, srec, sstart, stris;
    DevCollider dev{};
};

enum { EV_PLAN0, EV_PLAN1, EV_CORNERS, EV_CLASSIFY, EV_SCAN, EV_EMIT0, EV_EMIT1, EV_OUT, EV_CTL, EV_N };

struct mc3d_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    hipStream_t copy = nullptr;  // device-to-host copies of results (overlapping the bisection)
    void *scratch = nullptr;     // this device's `scratch` (allocation-free entry point)
    std::vector<std::unique_ptr<ColliderCache>> colliders;  // most recent last
    HostBuf staging, ctl_host, ctl_early, segstage, outstage;
    DevBuf blob, table, etable, bits, counts, offs, tiles, agg, actw, actb, binfo, bcode, vkeys, vseg, pts, out8;
    uint64_t agg_gen = ~0ull;  // epoch generation `agg` was zeroed for
    DevBuf hvals, hpts, hstate, hprobe;  // host-evaluated atoms: answers, points, bisection state
    HostBuf hstage;                      // pinned points / answers of one callback batch
    DevBuf stl, stl_in;                  // STL encoding: body bytes; uploaded host triangles
    // GroupTriangles: boxes, keys, the three orders (and a spare), flags,
    // scans, node axes and slots, hipCUB's temporary storage
    DevBuf g_box, g_keys[4], g_perm[4], g_flag, g_f, g_scan, g_axis, g_slots, g_tmp, g_tris;
    DevBuf keep;                         // block filter: cube mask
    // z slabs: a plane's vertex list and the stitch's per-vertex arrays
    DevBuf s_flags, s_scan, s_keys, s_idx, s_count, s_keep, s_rank, s_newidx, s_bad;
    // device block filter: triangles, their grid, the recursion's block lists
    DevBuf rf_tris, rf_counts, rf_start, rf_cursor, rf_cells, rf_tiles, rf_front[2], rf_levels, rf_vsat;
    HostBuf rf_total;
    std::shared_ptr<Pool<DevBuf>> pool = std::make_shared<Pool<DevBuf>>();
    std::shared_ptr<Pool<HostBuf>> hpool = std::make_shared<Pool<HostBuf>>();
    uint64_t v_hint = 1u << 18, t_hint = 1u << 19;  // output capacities to start from
    hipEvent_t ev[EV_N] = {};
    // Captured launch chains of recent call shapes (most recent last).
    struct Graph {
        std::vector<uint8_t> key;
        hipGraphExec_t exec = nullptr;
    };
    std::vector<Graph> graphs;
    bool graphs_off = false;  // a capture failed on this device: enqueue stage by stage
};

// One solid's mesh: a slice [voff, voff + nv) / [toff, toff + nt) of its
// call's outputs (triangle indices count from the slice's first vertex).
struct mc3d_mesh {
    int device = 0;
    uint64_t nv = 0, nt = 0, voff = 0, toff = 0;
    bool on_host = false, has_interior = false;
    std::shared_ptr<MeshBuffers> buf;
    mc3d_stats stats{};
    const double *hverts() const { return buf->host->as<double>() + 3 * voff; }
    const uint32_t *htris() const { return (const uint32_t *)(buf->host->as<double>() + 3 * buf->nv) + 3 * toff; }
    const double *hinterior() const {
        return (const double *)((const char *)buf->host->p + align_up(buf->nv * 24 + buf->nt * 12, 256)) + 3 * voff;
    }
};

namespace {

// MarchingCubesFilter with a mesh collider's RectCollision as its filter,
// run on the device (k_filter.hip): the collider's triangles, on the host
// or as a mesh of this context's device, and the filter's Rect.Expand.
struct DeviceRectFilter {
    const double *host_tris = nullptr;   // 9 doubles per triangle, or
    const double *dev_verts = nullptr;   // a device mesh: its vertices
    const uint32_t *dev_tris = nullptr;  // and triangle indices
    uint64_t ntris = 0;
    double extra = 0;  // Rect.Expand of every block's bounds (0: none)
    double cell = 0;   // grid cell edge: about twice a triangle's extent
};

// One solid of a call, prepared on the host.
struct Job {
    const mc3d_solid *s = nullptr;
    double delta = 0;
    int iters = 0;
    std::shared_ptr<const Lowered> low;  // the solid's (shared with its handle)
    std::vector<double> ax[3];
    Grid g{};    // sizes; axis pointers set by the upload
    Bricks B{};  // shape; array pointers set by the upload
    uint64_t nwords = 0, nbricks = 0, word0 = 0, brick0 = 0;
    uint64_t pos0 = 0, npos = 0;  // scan positions (npos 0: the scan takes the words, pos0 = word0)
    uint32_t cap = 0;  // folded-program slot of its culling bricks (0: no culling)
    bool interior = false;         // MarchingCubesInterior: keep each vertex's inside point
    std::vector<double> conj;      // MarchingCubesConj: inverse chain applied to the vertices
    std::vector<uint64_t> keep;    // MarchingCubesFilter: cubes of accepted blocks (occupancy layout)
    const DeviceRectFilter *dfilter = nullptr;  // or the filter that makes that mask on the device
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
void job_nodes(const Program &P, std::vector<PlanNode> &nodes);

// The solid's lowered program: made by the first call on the handle (under
// its lock), shared by every later call.  A failed lowering is not kept.
int lowered(const mc3d_solid *s, std::shared_ptr<const Lowered> &out) {
    std::lock_guard<std::mutex> g(s->low_mu);
    if (!s->low) {
        auto L = std::make_shared<Lowered>();
        std::string msg;
        if (!lower_solid(*s->ptr, L->P, msg)) {
            set_error(msg);
            return MC3D_E_UNSUPPORTED;
        }
        L->host = !L->P.host_atoms.empty();
        L->frames = host_frames(L->P);
        job_nodes(L->P, L->nodes);
        s->low = std::move(L);
    }
    out = s->low;
    return MC3D_OK;
}

// The occupancy grid of job j's axes (j.ax; a z slab keeps its open faces
// in j.g) and its culling bricks.  With check, a grid past 2^32 occupancy
// words is MC3D_E_UNSUPPORTED (single calls stream such grids in slabs).
int layout_job(Job &j, const mc3d_options &o, bool check) {
    Grid &g = j.g;
    g.nx = (int)j.ax[0].size(), g.ny = (int)j.ax[1].size(), g.nz = (int)j.ax[2].size();
    g.W = (g.nx + 63) / 64;
    j.nwords = (uint64_t)g.W * g.ny * g.nz;
    if (check && j.nwords >= (1ull << 32)) {
        set_error("grid too large: more than 2^32 occupancy words");
        return MC3D_E_UNSUPPORTED;
    }
    j.B = Bricks{nullptr, nullptr, 1, 1, 1, 1};
    j.cap = 0, j.nbricks = 0;
    // A host-evaluated solid may be true outside its bounds, so nothing is
    // folded by bounds: no culling.
    if (o.cull && !j.low->host) {
        int bsz = o.brick > 0 ? o.brick : 4;
        j.cap = std::min<uint32_t>((uint32_t)j.low->P.code.size(), 64u);
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

int prepare_job(Job &j, const mc3d_options &o, bool check = true) {
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
    int rc = lowered(j.s, j.low);
    if (rc) return rc;
    return layout_job(j, o, check);
}

// The device copy of a collider's ray grid and, for sphere queries of
// radius R, its sphere grid (uploaded once per context).
int collider_on_device(mc3d_ctx *ctx, const std::shared_ptr<const ColliderData> &cd, double R, DevCollider &out,
                       size_t keep_at_least) {
    for (size_t i = 0; i < ctx->colliders.size(); i++) {
        if (ctx->colliders[i]->keep->id == cd->id && ctx->colliders[i]->R == R) {
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
    c->R = R;
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
    d.srec = nullptr, d.scell_start = d.scell_tris = nullptr;
    d.so[0] = d.so[1] = d.so[2] = 0, d.sinv = 1, d.sn[0] = d.sn[1] = d.sn[2] = 0, d.pad = 0;
    if (R > 0) {
        const SphereGrid &S = cd->sphere_grid(R);
        if ((rc = c->srec.ensure(S.rec.size() * sizeof(double) + 8)) ||
            (rc = c->sstart.ensure(S.cell_start.size() * sizeof(uint32_t) + 8)) ||
            (rc = c->stris.ensure(S.cell_tris.size() * sizeof(uint32_t) + 8)))
            return rc;
        if (!S.rec.empty())
            HIP_TRY(hipMemcpy(c->srec.p, S.rec.data(), S.rec.size() * sizeof(double), hipMemcpyHostToDevice));
        HIP_TRY(hipMemcpy(c->sstart.p, S.cell_start.data(), S.cell_start.size() * sizeof(uint32_t),
                          hipMemcpyHostToDevice));
        if (!S.cell_tris.empty())
            HIP_TRY(hipMemcpy(c->stris.p, S.cell_tris.data(), S.cell_tris.size() * sizeof(uint32_t),
                              hipMemcpyHostToDevice));
        d.srec = c->srec.as<double>();
        d.scell_start = c->sstart.as<uint32_t>();
        d.scell_tris = c->stris.as<uint32_t>();
        memcpy(d.so, S.o, sizeof d.so);
        d.sinv = S.inv_cell;
        memcpy(d.sn, S.n, sizeof d.sn);
    }
    out = d;
    ctx->colliders.push_back(std::move(c));
    // Evict the oldest entries, never one the current call uses (those are
    // the newest `keep_at_least`, looked up or added in this upload).
    while (ctx->colliders.size() > std::max<size_t>(4, keep_at_least)) ctx->colliders.erase(ctx->colliders.begin());
    return MC3D_OK;
}

// Bytes of LDS the plan needs to stage a program: instructions, their atom
// boxes and the float64 params.
size_t plan_stage_bytes(const Program &P) { return PLAN_STACK_BYTES + P.code.size() * sizeof(PlanNode); }
constexpr size_t PLAN_STAGE_LIMIT = 64 * 1024;

// The node records of a postfix program for the brick plan (PlanNode): per
// node its instruction, link, all-false value, box (atoms: their own;
// operators: the union of their atoms'; constants: empty; a NaN bound makes
// the union unbounded on that axis) and the params of the "certainly true"
// tests.  boxes: each atom's box, by pc.
void plan_nodes(const Program &P, const std::vector<AtomBox> &boxes, std::vector<PlanNode> &nodes) {
    const size_t n = P.code.size();
    nodes.assign(n, PlanNode{});
    std::vector<uint32_t> st, start(n);
    st.reserve(64);
    auto unite = [](PlanNode &u, const PlanNode &a, const PlanNode &b) {
        for (int k = 0; k < 3; k++) {
            u.lo[k] = (isnan(a.lo[k]) || isnan(b.lo[k])) ? -INFINITY : std::min(a.lo[k], b.lo[k]);
            u.hi[k] = (isnan(a.hi[k]) || isnan(b.hi[k])) ? INFINITY : std::max(a.hi[k], b.hi[k]);
        }
    };
    for (size_t i = 0; i < n; i++) {
        PlanNode &nd = nodes[i];
        nd.x = P.code[i].x, nd.y = P.code[i].y;
        const uint32_t op = insn_op(nd.x);
        if (op == OP_ATOM || op == OP_TRUE || op == OP_FALSE) {
            start[i] = (uint32_t)i;
            nd.link = (uint32_t)i, nd.cfalse = op == OP_TRUE ? 1u : 0u;
            if (op == OP_ATOM) {
                memcpy(nd.lo, boxes[i].lo, sizeof nd.lo);
                memcpy(nd.hi, boxes[i].hi, sizeof nd.hi);
                const uint32_t kind = insn_kind(nd.x);
                const double *a = P.f64.data() + nd.y;
                if (kind == A_RECT) memcpy(nd.prm, a, 6 * sizeof(double));
                if (kind == A_SPHERE) nd.prm[0] = a[0], nd.prm[1] = a[1], nd.prm[2] = a[2], nd.prm[3] = a[4];
            } else {
                for (int k = 0; k < 3; k++) nd.lo[k] = INFINITY, nd.hi[k] = -INFINITY;
            }
        } else if (op == OP_NOT) {
            const uint32_t c = st.back();
            st.pop_back();
            start[i] = start[c];
            memcpy(nd.lo, nodes[c].lo, sizeof nd.lo);
            memcpy(nd.hi, nodes[c].hi, sizeof nd.hi);
            nd.link = c, nd.cfalse = nodes[c].cfalse ^ 1u;
        } else {
            const uint32_t b = st.back();
            st.pop_back();
            const uint32_t a = st.back();
            st.pop_back();
            const uint32_t fa = nodes[a].cfalse, fb = nodes[b].cfalse;
            start[i] = start[a];
            unite(nd, nodes[a], nodes[b]);
            nd.link = a;
            nd.cfalse = op == OP_OR ? (fa | fb) : op == OP_AND ? (fa & fb) : op == OP_ANDNOT ? (fa & ~fb & 1u)
                                                                                             : (~fa & fb & 1u);
        }
        st.push_back((uint32_t)i);
    }
    // each leaf links to the outermost node starting at it (the last in postfix order)
    for (size_t i = 0; i < n; i++)
        if (start[i] != i) nodes[start[i]].link = (uint32_t)i;
}

// Where a call's pieces sit in the program blob (offsets), and the blob's
// device address once placed.
struct Upload {
    // layout (bytes from the blob's start)
    size_t total = 0, o_sout = 0, o_segs = 0, o_wseg = 0, o_bseg = 0, o_pseg = 0, o_conj = 0;
    uint64_t npos_all = 0;  // brick-major scan positions of a call of several (0: none)
    struct Off {
        size_t code, f64, f32, fr, nodes, col, ax;
    };
    std::vector<Off> po;
    std::vector<std::vector<PlanNode>> nodes;
    std::vector<std::vector<DevCollider>> cols;
    bool tables = false;
    // placed
    Ctl *ctl = nullptr;
    SegOut *sout = nullptr;
    Seg *segs = nullptr;
    uint32_t *word_seg = nullptr, *brick_seg = nullptr, *pos_seg = nullptr;
    std::vector<Seg> hsegs;  // host copies of the segments
    size_t back_bytes = 0;   // control block + per-solid results, copied back at the end
    const double *conj = nullptr;  // job 0's MarchingCubesConj inverse chain (records of 10 doubles)
    uint32_t conj_nops = 0;
    uint32_t epoch = 0;  // this call's scan epoch (written into the control block)
};

// The blob's layout: the control block, the per-solid results, the
// segments, the segment tables and every job's program and grid axes.
// `cols` holds each job's colliders as the device sees them.
// plan_nodes of a program, its atoms' boxes placed by pc.
void job_nodes(const Program &P, std::vector<PlanNode> &nodes) {
    std::vector<AtomBox> boxes(P.code.size());
    for (size_t i = 0; i < P.atom_pc.size(); i++) boxes[P.atom_pc[i]] = P.atom_box[i];
    plan_nodes(P, boxes, nodes);
}

void layout_upload(const std::vector<Job> &jobs, uint64_t nwords_all, uint64_t nbricks_all, Upload &U) {
    const size_t n = jobs.size();
    size_t off = 0;
    auto place = [&](size_t bytes) {
        size_t o = off;
        off = align_up(off + bytes, 256);
        return o;
    };
    // the per-solid results follow the control block directly (one copy back)
    U.o_sout = align_up(sizeof(Ctl), 16);
    U.back_bytes = U.o_sout + n * sizeof(SegOut);
    off = align_up(U.back_bytes, 256);
    U.o_segs = place(n * sizeof(Seg));
    U.tables = n > 1;
    U.o_wseg = place(U.tables ? nwords_all / 64 * sizeof(uint32_t) : 0);
    U.o_bseg = place(U.tables ? nbricks_all / 128 * sizeof(uint32_t) : 0);
    // brick-major positions of several solids: a table of their own
    U.npos_all = U.tables && jobs[0].npos ? jobs.back().pos0 + jobs.back().npos : 0;
    U.o_pseg = place(U.npos_all / 64 * sizeof(uint32_t));
    U.o_conj = place(jobs[0].conj.size() * sizeof(double));
    U.po.resize(n);
    U.nodes.resize(n);
    for (size_t s = 0; s < n; s++) {
        const Program &P = jobs[s].low->P;
        if (jobs[s].low->nodes.size() != P.code.size()) job_nodes(P, U.nodes[s]);
        const std::vector<PlanNode> &nodes = U.nodes[s].empty() ? jobs[s].low->nodes : U.nodes[s];
        Upload::Off &po = U.po[s];
        po.code = place(P.code.size() * sizeof(Insn));
        // params padded so that staging a brick's atoms never reads past the array
        po.f64 = place((P.f64.size() + MAX_ATOM_PARAMS) * sizeof(double));
        po.f32 = place((P.f32.size() + MAX_ATOM_PARAMS) * sizeof(float));
        po.fr = place(P.frames.size() * sizeof(Frame));
        po.nodes = place(nodes.size() * sizeof(PlanNode));
        po.col = place(P.colliders.size() * sizeof(DevCollider));
        po.ax = place((jobs[s].ax[0].size() + jobs[s].ax[1].size() + jobs[s].ax[2].size()) * sizeof(double));
    }
    U.total = off;
}

// Writes the blob laid out by layout_upload into host memory h, for the
// device address dv (the caller copies it there); fills the segments'
// device pointers (culling bricks from b).
void fill_upload(std::vector<Job> &jobs, uint64_t nwords_all, uint64_t nbricks_all, const Bufs &b, char *h, char *dv,
                 Upload &U) {
    const size_t n = jobs.size();
    memset(h, 0, U.o_segs);  // control block and per-solid results start at zero
    ((Ctl *)h)->epoch = U.epoch;
    U.hsegs.assign(n, Seg{});
    // the jobs' regions are disjoint: many jobs are filled in parallel
    auto fill = [&](size_t s) {
        Job &j = jobs[s];
        const Program &P = j.low->P;
        const Upload::Off &po = U.po[s];
        if (U.tables) {  // this solid's entries of the segment tables (disjoint per solid)
            uint32_t *ws = (uint32_t *)(h + U.o_wseg), *bs = (uint32_t *)(h + U.o_bseg);
            const uint64_t w1 = s + 1 < n ? jobs[s + 1].word0 : nwords_all;
            for (uint64_t w = j.word0; w < w1; w += 64) ws[w / 64] = (uint32_t)s;
            const uint64_t b1 = s + 1 < n ? jobs[s + 1].brick0 : nbricks_all;
            for (uint64_t bb = j.brick0; bb < b1; bb += 128) bs[bb / 128] = (uint32_t)s;
            uint32_t *ps = (uint32_t *)(h + U.o_pseg);
            if (U.npos_all)
                for (uint64_t q = j.pos0; q < j.pos0 + j.npos; q += 64) ps[q / 64] = (uint32_t)s;
        }
        memcpy(h + po.code, P.code.data(), P.code.size() * sizeof(Insn));
        memset(h + po.f64, 0, (P.f64.size() + MAX_ATOM_PARAMS) * sizeof(double));
        memset(h + po.f32, 0, (P.f32.size() + MAX_ATOM_PARAMS) * sizeof(float));
        if (!P.f64.empty()) {
            memcpy(h + po.f64, P.f64.data(), P.f64.size() * sizeof(double));
            memcpy(h + po.f32, P.f32.data(), P.f32.size() * sizeof(float));
        }
        memcpy(h + po.fr, P.frames.data(), P.frames.size() * sizeof(Frame));
        const std::vector<PlanNode> &nodes = U.nodes[s].empty() ? j.low->nodes : U.nodes[s];
        memcpy(h + po.nodes, nodes.data(), nodes.size() * sizeof(PlanNode));
        if (!U.cols[s].empty()) memcpy(h + po.col, U.cols[s].data(), U.cols[s].size() * sizeof(DevCollider));
        double *d = (double *)(h + po.ax);
        for (int a = 0; a < 3; a++) d = std::copy(j.ax[a].begin(), j.ax[a].end(), d);
        Seg &sg = U.hsegs[s];
        sg.P.code = (const Insn *)(dv + po.code);
        sg.P.f64 = (const double *)(dv + po.f64);
        sg.P.f32 = (const float *)(dv + po.f32);
        sg.P.frames = (const Frame *)(dv + po.fr);
        sg.P.plan = (const PlanNode *)(dv + po.nodes);
        sg.P.colliders = (const DevCollider *)(dv + po.col);
        sg.P.host_vals = nullptr;
        sg.P.host_stride = 0;
        sg.P.ncode = (uint32_t)P.code.size();
        sg.P.nf64 = (uint32_t)P.f64.size();
        if (!j.ax[0].empty()) {
            const double *ax = (const double *)(dv + po.ax);
            j.g.xs = ax, j.g.ys = ax + j.g.nx, j.g.zs = ax + j.g.nx + j.g.ny;
        }
        sg.g = j.g;
        sg.B = j.B;
        if (j.cap) sg.B.info = b.binfo, sg.B.code = b.bcode;
        sg.word0 = j.word0, sg.nwords = j.nwords;
        sg.brick0 = j.brick0, sg.nbricks = j.nbricks;
        sg.pos0 = j.pos0;
        sg.iters = j.iters;
        sg.stage = j.cap && plan_stage_bytes(P) <= PLAN_STAGE_LIMIT ? 1 : 0;
    };
    if (n >= 16) HostPool::get().run(n, fill);
    else
        for (size_t s = 0; s < n; s++) fill(s);
    memcpy(h + U.o_segs, U.hsegs.data(), n * sizeof(Seg));
    if (!jobs[0].conj.empty()) {
        memcpy(h + U.o_conj, jobs[0].conj.data(), jobs[0].conj.size() * sizeof(double));
        U.conj = (const double *)(dv + U.o_conj);
        U.conj_nops = (uint32_t)(jobs[0].conj.size() / 10);
    }
    U.ctl = (Ctl *)dv;
    U.sout = (SegOut *)(dv + U.o_sout);
    U.segs = (Seg *)(dv + U.o_segs);
    U.word_seg = U.tables ? (uint32_t *)(dv + U.o_wseg) : nullptr;
    U.brick_seg = U.tables ? (uint32_t *)(dv + U.o_bseg) : nullptr;
    U.pos_seg = U.npos_all ? (uint32_t *)(dv + U.o_pseg) : U.word_seg;
}

// The colliders of every job on this context's device.
int upload_colliders(mc3d_ctx *ctx, const std::vector<Job> &jobs, Upload &U) {
    const size_t n = jobs.size();
    size_t ncol = 0;
    for (const Job &j : jobs) ncol += j.low->P.colliders.size();
    U.cols.assign(n, {});
    for (size_t s = 0; s < n; s++) {
        const Program &P = jobs[s].low->P;
        U.cols[s].resize(P.colliders.size());
        for (size_t i = 0; i < U.cols[s].size(); i++) {
            int rc = collider_on_device(ctx, P.colliders[i], P.collider_R[i], U.cols[s][i], ncol);
            if (rc) return rc;
        }
    }
    return MC3D_OK;
}

// The whole blob in the context's pinned staging buffer, for one transfer
// (enqueue_upload).
int prepare_upload(mc3d_ctx *ctx, std::vector<Job> &jobs, uint64_t nwords_all, uint64_t nbricks_all, const Bufs &b,
                   Upload &U) {
    int rc;
    if ((rc = upload_colliders(ctx, jobs, U))) return rc;
    layout_upload(jobs, nwords_all, nbricks_all, U);
    if ((rc = ctx->staging.ensure(U.total)) || (rc = ctx->blob.ensure(U.total))) return rc;
    fill_upload(jobs, nwords_all, nbricks_all, b, (char *)ctx->staging.p, (char *)ctx->blob.p, U);
    return MC3D_OK;
}

int enqueue_upload(mc3d_ctx *ctx, const Upload &U) {
    HIP_TRY(hipMemcpyAsync(ctx->blob.p, ctx->staging.p, U.total, hipMemcpyHostToDevice, ctx->stream));
    return MC3D_OK;
}

int upload(mc3d_ctx *ctx, std::vector<Job> &jobs, uint64_t nwords_all, uint64_t nbricks_all, const Bufs &b,
           Upload &U) {
    int rc;
    if ((rc = prepare_upload(ctx, jobs, nwords_all, nbricks_all, b, U))) return rc;
    return enqueue_upload(ctx, U);
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
    const Program &P = j.low->P;
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
            HIP_TRY(launch_host_points(dp, U.hsegs[s].g, pts, j.low->frames[h], s0, c, ctx->hpts.as<double>(),
                                       ctx->stream));
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

double ms_since(std::chrono::steady_clock::time_point t0) {
    return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
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
    bool host = false;      // a call of one solid with host-evaluated atoms
    bool bm = false;        // the scan takes the words in brick-major order (one culled solid)
    uint64_t nscan = 0;     // scan positions: words, or brick-major positions
    size_t plan_lds = 0;
};

// The call's sizes; places every job's words and bricks.
int plan_call(std::vector<Job> &jobs, const mc3d_options &o, Call &c) {
    c.nsegs = (uint32_t)jobs.size();
    c.mode = o.precision == MC3D_PREC_FP32_GUARD ? 1 : 0;
    c.stats = o.count_mismatch != 0;
    c.rows_per_walk = o.rows_per_walk > 0 ? o.rows_per_walk : 1;
    c.wpb = o.waves_per_brick > 0 ? o.waves_per_brick : 4;
    c.cull = jobs[0].cap != 0;
    c.interior = c.nsegs == 1 && jobs[0].interior;
    c.host = c.nsegs == 1 && jobs[0].low->host;
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
    // Culled solids: the scan takes each solid's words brick by brick (the
    // positions include the rows of bricks past the grid), each solid's
    // positions padded to whole waves (classify reads a wave's rows from its
    // own lanes, and a wave stays in one solid).
    c.nscan = c.nwords;
    bool bm = c.cull && !c.host;
    uint64_t np = 0;
    for (const Job &j : jobs) {
        const int BY = j.B.BY, BZ = j.B.BZ;
        bm = bm && j.cap && (BY & (BY - 1)) == 0 && (BZ & (BZ - 1)) == 0;
        if (bm) np += align_up(j.nbricks * (uint64_t)BY * (uint64_t)BZ, 64);
    }
    if (bm && np < (1ull << 32) - 64) {
        c.bm = true, c.nscan = np;
        np = 0;
        for (Job &j : jobs) {
            j.pos0 = np, j.npos = align_up(j.nbricks * (uint64_t)j.B.BY * (uint64_t)j.B.BZ, 64);
            np += j.npos;
        }
    } else {
        for (Job &j : jobs) j.pos0 = j.word0, j.npos = 0;
    }
    c.plan_lds = PLAN_STACK_BYTES;
    if (c.cull)
        for (const Job &j : jobs) {
            const size_t lds = plan_stage_bytes(j.low->P);
            if (lds <= PLAN_STAGE_LIMIT) c.plan_lds = std::max(c.plan_lds, lds);
        }
    return MC3D_OK;
}

uint64_t classify_blocks(const Call &c) { return (c.nscan + CLASSIFY_BLOCK - 1) / CLASSIFY_BLOCK; }

// The context's buffers for a call (grown as needed).
int ctx_buffers(mc3d_ctx *ctx, const Call &c, Bufs &b, uint32_t &epoch) {
    int rc;
    const void *agg_was = ctx->agg.p;
    if ((rc = ctx->agg.ensure(SCAN_ONE_PASS * sizeof(ScanAgg)))) return rc;
    uint64_t gen;
    next_epoch(epoch, gen);
    if (ctx->agg.p != agg_was || ctx->agg_gen != gen) {  // new memory, or the epochs started over
        HIP_TRY(launch_zero(ctx->agg.p, ctx->agg.cap, ctx->stream));
        ctx->agg_gen = gen;
    }
    b.agg = ctx->agg.as<ScanAgg>();
    if ((rc = ctx->bits.ensure(c.nwords * 8)) || (rc = ctx->counts.ensure(c.nscan * sizeof(uint32_t))) ||
        (rc = ctx->offs.ensure(c.nscan * sizeof(uint2))) || (rc = ctx->actw.ensure(c.nwords * sizeof(uint32_t))) ||
        (rc = ctx->tiles.ensure(classify_blocks(c) * sizeof(uint4))))
        return rc;
    if (c.cull && ((rc = ctx->binfo.ensure(c.nbricks * sizeof(uint2))) ||
                   (rc = ctx->actb.ensure(c.nbricks * sizeof(uint32_t))) ||
                   (rc = ctx->bcode.ensure(c.nbricks * c.cap * sizeof(Insn) + 16))))
        return rc;
    b.bits = ctx->bits.as<uint64_t>();
    b.counts = ctx->counts.as<uint32_t>();
    b.offs = ctx->offs.as<uint2>();
    b.actw = ctx->actw.as<uint32_t>();
    b.tiles = ctx->tiles.as<uint4>();
    b.binfo = c.cull ? ctx->binfo.as<uint2>() : nullptr;
    b.actb = c.cull ? ctx->actb.as<uint32_t>() : nullptr;
    b.bcode = c.cull ? ctx->bcode.as<Insn>() : nullptr;
    b.table = ctx->table.as<uint8_t>();
    b.etable = ctx->etable.as<uint8_t>();
    return MC3D_OK;
}

Pipe make_pipe(const Call &c, const Upload &U) {
    Pipe p;
    memset((void *)&p, 0, sizeof p);  // padding included: graph keys compare the bytes
    p.segs = U.segs;
    p.nsegs = c.nsegs;
    p.seg0 = U.hsegs[0];
    p.word_seg = U.word_seg;
    p.brick_seg = U.brick_seg;
    p.pos_seg = U.pos_seg;
    p.ctl = U.ctl;
    p.sout = U.sout;
    p.nwords = c.nwords;
    p.nbricks = c.nbricks;
    p.nscan = c.nscan;
    p.cap = c.cap;
    p.cull = c.cull;
    p.bm = c.bm;
    p.stats = c.stats;
    p.mode = c.mode;
    p.rows_per_walk = c.rows_per_walk;
    p.wpb = c.wpb;
    p.plan_lds = c.plan_lds;
    // Launch shapes measured on MI355X: a single solid's plan waves take
    // compact tiles, 16 bricks for programs that stage in LDS and 64 for the
    // large ones walked through scalar loads (csg64 39.6 -> 37.4 us, the
    // 512-primitive tree 183 -> 160 us); emit gets a block per 512 words.
    p.plan_tile = true;
    p.plan_gsz = c.plan_lds > PLAN_STACK_BYTES ? 16 : 64;
    p.emit_div = 512;
    p.scan_one_pass = classify_blocks(c) <= SCAN_ONE_PASS;
    return p;
}

// plan, corners, classify and scan, with events between them (ev: the
// context's events, or null).
int enqueue_front(const Pipe &p, const Bufs &b, hipStream_t st, const hipEvent_t *ev) {
    if (ev) HIP_TRY(hipEventRecord(ev[EV_PLAN0], st));
    if (p.cull) HIP_TRY(launch_plan(p, b, st));
    if (ev) HIP_TRY(hipEventRecord(ev[EV_PLAN1], st));
    HIP_TRY(launch_corners(p, b, st));
    if (ev) HIP_TRY(hipEventRecord(ev[EV_CORNERS], st));
    HIP_TRY(launch_classify(p, b, st));
    if (ev) HIP_TRY(hipEventRecord(ev[EV_CLASSIFY], st));
    HIP_TRY(launch_scan(p, b, st));
    if (ev) HIP_TRY(hipEventRecord(ev[EV_SCAN], st));
    return MC3D_OK;
}

// emit (vertex keys, triangles) and bisect (vertices), then the conj map.
int enqueue_output(const Pipe &p, const Bufs &b, const Upload &U, double *verts, uint32_t *tris, double *interior,
                   uint64_t vcap, uint64_t tcap, hipStream_t st, const hipEvent_t *ev) {
    if (ev) HIP_TRY(hipEventRecord(ev[EV_EMIT0], st));
    HIP_TRY(launch_emit(p, b, tris, vcap, tcap, st));
    if (ev) HIP_TRY(hipEventRecord(ev[EV_EMIT1], st));
    HIP_TRY(launch_bisect(p, b, verts, interior, vcap, st));
    if (U.conj) HIP_TRY(launch_inverse_apply(U.conj, U.conj_nops, verts, p.ctl, vcap, st));
    if (ev) HIP_TRY(hipEventRecord(ev[EV_OUT], st));
    return MC3D_OK;
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

// Meshes keep their outputs until freed: a buffer sized for much larger
// earlier calls (over 4x the mesh and 64 MiB more) is traded for one that
// fits (the copy is ordered on the stream) and goes back to the pool for the
// next call; smaller slack is kept, as trading costs a copy per call.
int fit_outputs(mc3d_ctx *ctx, MeshBuffers &mb, uint64_t nv, uint64_t nt) {
    auto fit = [&](std::unique_ptr<DevBuf> &b, size_t need) -> int {
        if (!b || b->cap <= 4 * need + (64u << 20)) return MC3D_OK;
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

// The output stages of a solid with host-evaluated atoms (a call of one):
// the mesh size is read first (one more synchronisation), then emit, then
// the bisection in rounds with the host answering its atoms between them.
// Leaves the final control block in ctx->ctl_host.
int run_output_host(mc3d_ctx *ctx, const Call &c, const Pipe &p, Bufs &b, const Job &j, Upload &U, MeshBuffers &mb,
                    double &ms_cb) {
    Ctl *hc = (Ctl *)ctx->ctl_host.p;
    HIP_TRY(hipMemcpyAsync(hc, U.ctl, U.back_bytes, hipMemcpyDeviceToHost, ctx->stream));
    HIP_TRY(hipStreamSynchronize(ctx->stream));
    if (hc->scan_error) {
        set_error("internal error: classify gave up waiting for the blocks before it");
        return MC3D_E_INTERNAL;
    }
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
    b.vkeys = ctx->vkeys.as<uint64_t>();
    HIP_TRY(hipEventRecord(ctx->ev[EV_EMIT0], ctx->stream));
    HIP_TRY(launch_emit(p, b, mb.tris->as<uint32_t>(), vcap, tcap, ctx->stream));
    HIP_TRY(hipEventRecord(ctx->ev[EV_EMIT1], ctx->stream));
    double *verts = mb.verts->as<double>(), *state = ctx->hstate.as<double>(), *probe = ctx->hprobe.as<double>();
    double *inter = c.interior ? mb.interior->as<double>() : nullptr;
    for (int round = 0; round <= j.iters; round++) {
        if (round > 0 && (rc = host_round(ctx, j, U, 0, probe, nv, ms_cb))) return rc;
        HIP_TRY(launch_bisect_host(c.mode, U.hsegs[0].P, U.hsegs[0].g, b.bits, b.vkeys, state, probe, verts, inter,
                                   nv, round > 0, ctx->stream));
    }
    if (U.conj) HIP_TRY(launch_inverse_apply(U.conj, U.conj_nops, verts, U.ctl, nv, ctx->stream));
    HIP_TRY(hipEventRecord(ctx->ev[EV_OUT], ctx->stream));
    HIP_TRY(hipMemcpyAsync(hc, U.ctl, U.back_bytes, hipMemcpyDeviceToHost, ctx->stream));
    HIP_TRY(hipStreamSynchronize(ctx->stream));
    return MC3D_OK;
}

// Pinned host memory for a call's results: vertices, triangles, interior points.
size_t host_result_bytes(uint64_t nv, uint64_t nt, bool interior) {
    return align_up(nv * 24 + nt * 12, 256) + (interior ? nv * 24 : 0) + 16;
}

// Copies of the results into mb.host on the copy stream: the triangles once
// emit is done (while the bisection runs), the vertices and interior points
// once the output stages are.
int enqueue_host_copy(mc3d_ctx *ctx, MeshBuffers &mb, uint64_t nv, uint64_t nt, bool interior) {
    int rc;
    if (!mb.host) mb.host = ctx->hpool->take(host_result_bytes(nv, nt, interior));
    if ((rc = mb.host->ensure(host_result_bytes(nv, nt, interior)))) return rc;
    mb.nv = nv, mb.nt = nt;
    char *h = (char *)mb.host->p;
    HIP_TRY(hipStreamWaitEvent(ctx->copy, ctx->ev[EV_EMIT1], 0));
    if (nt) HIP_TRY(hipMemcpyAsync(h + nv * 24, mb.tris->p, nt * 12, hipMemcpyDeviceToHost, ctx->copy));
    HIP_TRY(hipStreamWaitEvent(ctx->copy, ctx->ev[EV_OUT], 0));
    if (nv) HIP_TRY(hipMemcpyAsync(h, mb.verts->p, nv * 24, hipMemcpyDeviceToHost, ctx->copy));
    if (interior && nv)
        HIP_TRY(hipMemcpyAsync(h + align_up(nv * 24 + nt * 12, 256), mb.interior->p, nv * 24, hipMemcpyDeviceToHost,
                               ctx->copy));
    return MC3D_OK;
}

// Everything a captured chain depends on: the kernels' arguments and launch
// shapes (Pipe, Bufs, outputs and capacities) and the copies' addresses and
// sizes.  Two calls with equal keys enqueue exactly the same work.
struct GraphKey {
    Pipe p;
    Bufs b;
    const void *verts, *tris, *interior, *staging, *blob, *ctl_host;
    const double *conj;
    uint64_t vcap, tcap, total, back;
    uint32_t conj_nops, pad;
};

// The upload, every stage and the control block's copy back, as one HIP
// graph launch: captured once per call shape and replayed by later calls of
// the same shape (hipGraphLaunch replaces a dozen enqueues).  No memset is
// captured: the control block arrives zeroed with the upload.
int graph_launch(mc3d_ctx *ctx, const Pipe &p, const Bufs &b, const Upload &U, double *verts, uint32_t *tris,
                 double *interior, uint64_t vcap, uint64_t tcap) {
    GraphKey k;
    memset((void *)&k, 0, sizeof k);
    k.p = p, k.b = b, k.verts = verts, k.tris = tris, k.interior = interior;
    k.staging = ctx->staging.p, k.blob = ctx->blob.p, k.ctl_host = ctx->ctl_host.p;
    k.conj = U.conj, k.conj_nops = U.conj_nops;
    k.vcap = vcap, k.tcap = tcap, k.total = U.total, k.back = U.back_bytes;
    const uint8_t *kb = (const uint8_t *)&k;
    hipGraphExec_t exec = nullptr;
    for (size_t i = 0; i < ctx->graphs.size(); i++) {
        if (memcmp(ctx->graphs[i].key.data(), kb, sizeof k) == 0) {
            mc3d_ctx::Graph hit = std::move(ctx->graphs[i]);  // now the most recent
            ctx->graphs.erase(ctx->graphs.begin() + i);
            exec = hit.exec;
            ctx->graphs.push_back(std::move(hit));
            break;
        }
    }
    if (!exec) {
        HIP_TRY(hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeRelaxed));
        int rc = enqueue_upload(ctx, U);
        if (!rc) rc = enqueue_front(p, b, ctx->stream, nullptr);
        if (!rc) rc = enqueue_output(p, b, U, verts, tris, interior, vcap, tcap, ctx->stream, nullptr);
        if (!rc && hipMemcpyAsync(ctx->ctl_host.p, U.ctl, U.back_bytes, hipMemcpyDeviceToHost, ctx->stream) != hipSuccess)
            rc = MC3D_E_HIP;
        hipGraph_t graph = nullptr;
        const hipError_t e = hipStreamEndCapture(ctx->stream, &graph);
        if (!rc && e == hipSuccess && graph && hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0) != hipSuccess)
            exec = nullptr;
        if (graph) (void)hipGraphDestroy(graph);
        if (!exec) {
            (void)hipGetLastError();
            ctx->graphs_off = true;  // enqueue stage by stage from now on
            return MC3D_E_UNSUPPORTED;
        }
        mc3d_ctx::Graph g;
        g.key.assign(kb, kb + sizeof k);
        g.exec = exec;
        ctx->graphs.push_back(std::move(g));
        while (ctx->graphs.size() > 4) {
            (void)hipGraphExecDestroy(ctx->graphs.front().exec);
            ctx->graphs.erase(ctx->graphs.begin());
        }
    }
    HIP_TRY(hipEventRecord(ctx->ev[EV_PLAN0], ctx->stream));
    HIP_TRY(hipGraphLaunch(exec, ctx->stream));
    HIP_TRY(hipEventRecord(ctx->ev[EV_OUT], ctx->stream));
    return MC3D_OK;
}

// Blocks at each level of mcBlock.Pieces(64, ...) (mc.go:760-797) from a
// root of l[0] x l[1] x l[2] cubes when every block is accepted: the sizes
// the level lists need.  The halves of a split are floor and ceil of the
// longest axis, so few distinct shapes occur per level.
std::vector<uint64_t> filter_level_sizes(int lx, int ly, int lz) {
    std::map<std::array<int, 3>, uint64_t> cur{{{lx, ly, lz}, 1}};
    std::vector<uint64_t> sizes;
    while (!cur.empty()) {
        uint64_t total = 0;
        std::map<std::array<int, 3>, uint64_t> next;
        for (const auto &e : cur) {
            total += e.second;
            const std::array<int, 3> &d = e.first;
            if ((int64_t)d[0] * d[1] * d[2] / 2 < 64) continue;
            int axis = 0;
            if (d[1] >= d[0] && d[1] >= d[2]) axis = 1;
            else if (d[2] >= d[0] && d[2] >= d[1]) axis = 2;
            std::array<int, 3> a = d, b = d;
            a[axis] = d[axis] / 2, b[axis] = d[axis] - d[axis] / 2;
            next[a] += e.second;
            next[b] += e.second;
        }
        sizes.push_back(total);
        cur.swap(next);
    }
    return sizes;
}

// The block filter of job j on the device, into the cube mask `keep`
// (after the upload: it reads the grid's axes there).  The collider's
// triangles are binned in a grid over the search grid's bounds (cells of
// about F.cell, at most 2^20 of them; triangles outside fall in the edge
// cells), then the recursion runs one launch per level.
int enqueue_block_filter(mc3d_ctx *ctx, const Job &j, const Upload &U, uint64_t *keep) {
    const DeviceRectFilter &F = *j.dfilter;
    hipStream_t st = ctx->stream;
    if (F.ntris >= (1ull << 32) / 9) {
        set_error("block filter: too many collider triangles");
        return MC3D_E_UNSUPPORTED;
    }
    int rc;
    const size_t tbytes = F.ntris * 9 * sizeof(double);
    if ((rc = ctx->rf_tris.ensure(tbytes + 8))) return rc;
    if (F.host_tris) {
        if (tbytes) HIP_TRY(hipMemcpyAsync(ctx->rf_tris.p, F.host_tris, tbytes, hipMemcpyHostToDevice, st));
    } else {
        HIP_TRY(launch_rgrid_gather(F.dev_verts, F.dev_tris, F.ntris, ctx->rf_tris.as<double>(), st));
    }
    RectGridDev R;
    R.tris = ctx->rf_tris.as<double>();
    R.ntris = (uint32_t)F.ntris;
    double ext[3], emax = 0;
    for (int a = 0; a < 3; a++) {
        R.o[a] = j.ax[a].front();
        ext[a] = j.ax[a].back() - j.ax[a].front();
        emax = std::max(emax, ext[a]);
    }
    double cell = F.cell > 0 && std::isfinite(F.cell) ? F.cell : emax / 64;
    cell = std::max(cell, emax / 1e6);
    if (!(cell > 0) || !std::isfinite(cell)) cell = 1;
    uint64_t ncell;
    for (;;) {
        ncell = 1;
        for (int a = 0; a < 3; a++) {
            R.n[a] = (int)std::min(ext[a] / cell + 1, 1e7);
            ncell *= (uint64_t)R.n[a];
        }
        if (ncell <= (1u << 20)) break;
        cell *= 1.25;
    }
    R.inv = 1 / cell;
    if ((rc = ctx->rf_counts.ensure(ncell * 4)) || (rc = ctx->rf_start.ensure(ncell * 4 + 4)) ||
        (rc = ctx->rf_cursor.ensure(ncell * 4)) || (rc = ctx->rf_tiles.ensure(1024 * 4)) ||
        (rc = ctx->rf_total.ensure(4)))
        return rc;
    HIP_TRY(hipMemsetAsync(ctx->rf_counts.p, 0, ncell * 4, st));
    HIP_TRY(launch_rgrid_count(R, ctx->rf_counts.as<uint32_t>(), st));
    HIP_TRY(launch_rgrid_scan(ctx->rf_counts.as<uint32_t>(), (uint32_t)ncell, ctx->rf_tiles.as<uint32_t>(),
                              ctx->rf_start.as<uint32_t>(), ctx->rf_cursor.as<uint32_t>(), st));
    HIP_TRY(hipMemcpyAsync(ctx->rf_total.p, ctx->rf_start.as<uint32_t>() + ncell, 4, hipMemcpyDeviceToHost, st));
    HIP_TRY(hipStreamSynchronize(st));
    const uint32_t total = *ctx->rf_total.as<uint32_t>();
    if ((rc = ctx->rf_cells.ensure((size_t)total * 4 + 4))) return rc;
    R.cell_start = ctx->rf_start.as<uint32_t>();
    R.cell_tris = ctx->rf_cells.as<uint32_t>();
    HIP_TRY(launch_rgrid_fill(R, ctx->rf_cursor.as<uint32_t>(), ctx->rf_cells.as<uint32_t>(), st));
    // the vertex table: cells of half the edge, at most 2^21
    double vcell = cell / 2;
    uint64_t nvcell;
    for (;;) {
        nvcell = 1;
        for (int a = 0; a < 3; a++) {
            R.vn[a] = (int)std::min(ext[a] / vcell + 1, 1e7);
            nvcell *= (uint64_t)R.vn[a];
        }
        if (nvcell <= (1u << 21)) break;
        vcell *= 1.25;
    }
    for (int a = 0; a < 3; a++) R.vo[a] = R.o[a];
    R.vinv = 1 / vcell;
    if ((rc = ctx->rf_vsat.ensure(nvcell * 4))) return rc;
    HIP_TRY(hipMemsetAsync(ctx->rf_vsat.p, 0, nvcell * 4, st));
    HIP_TRY(launch_vgrid(R, ctx->rf_vsat.as<uint32_t>(), st));
    R.vsat = ctx->rf_vsat.as<uint32_t>();

    // the recursion from the root block (mc.go:712-716), level by level
    const FBlock root{{0, 0, 0}, {j.g.nx - 1, j.g.ny - 1, j.g.nz - 1}};
    const std::vector<uint64_t> sizes = filter_level_sizes(root.mx[0], root.mx[1], root.mx[2]);
    const uint64_t cap = *std::max_element(sizes.begin(), sizes.end());
    if (cap >= (1ull << 32)) {
        set_error("block filter: too many blocks");
        return MC3D_E_UNSUPPORTED;
    }
    const uint32_t nlevels = (uint32_t)sizes.size();
    if ((rc = ctx->rf_front[0].ensure(cap * sizeof(FBlock))) || (rc = ctx->rf_front[1].ensure(cap * sizeof(FBlock))) ||
        (rc = ctx->rf_levels.ensure((nlevels + 1) * 4)))
        return rc;
    HIP_TRY(hipMemsetAsync(keep, 0, j.nwords * 8, st));
    uint32_t *counts = ctx->rf_levels.as<uint32_t>();
    FBlock *front[2] = {ctx->rf_front[0].as<FBlock>(), ctx->rf_front[1].as<FBlock>()};
    HIP_TRY(launch_filter_root(root, front[0], counts, nlevels, st));
    const Grid &g = U.hsegs[0].g;
    const BlockFilter BF{R, g.xs, g.ys, g.zs, j.delta * 1e-3, F.extra, keep, g.W, g.ny, U.ctl};
    for (uint32_t L = 0; L < nlevels; L++)
        HIP_TRY(launch_filter_level(BF, front[L & 1], counts + L, front[(L + 1) & 1], counts + L + 1, (uint32_t)cap,
                                    sizes[L], st));
    return MC3D_OK;
}

// Meshes `jobs` (prepared) in one pass of the pipeline; out[i] gets job i's
// mesh.  A job with host-evaluated atoms is only ever run alone.
int run_call(mc3d_ctx *ctx, std::vector<Job> &jobs, const mc3d_options &o, std::vector<mc3d_mesh *> &out,
             std::chrono::steady_clock::time_point t_call, double ms_prepare) {
    const auto t_enqueue = std::chrono::steady_clock::now();
    double ms_enqueue = 0, ms_wait = 0, ms_d2h = 0;
    Call c;
    int rc;
    if ((rc = plan_call(jobs, o, c))) return rc;
    Bufs b;
    uint32_t epoch = 0;
    if ((rc = ctx_buffers(ctx, c, b, epoch))) return rc;
    const bool dfilter = c.nsegs == 1 && jobs[0].dfilter;
    if (c.nsegs == 1 && (!jobs[0].keep.empty() || dfilter)) {
        if ((rc = ctx->keep.ensure(jobs[0].nwords * 8))) return rc;
        if (!dfilter)
            HIP_TRY(hipMemcpyAsync(ctx->keep.p, jobs[0].keep.data(), jobs[0].keep.size() * 8, hipMemcpyHostToDevice,
                                   ctx->stream));
        b.keep = ctx->keep.as<uint64_t>();
    }
    Upload U;
    U.epoch = epoch;
    if ((rc = prepare_upload(ctx, jobs, c.nwords, c.nbricks, b, U))) return rc;
    if ((rc = ctx->ctl_host.ensure(U.back_bytes)) || (rc = ctx->ctl_early.ensure(U.back_bytes))) return rc;

    std::shared_ptr<MeshBuffers> mb = std::make_shared<MeshBuffers>();
    mb->pool = ctx->pool;
    mb->hpool = ctx->hpool;
    mc3d_stats st{};
    Ctl *hc = (Ctl *)ctx->ctl_host.p;
    const SegOut *so = (const SegOut *)((const char *)hc + U.o_sout);
    const bool to_host = !o.device_only;
    // Results left on the device, no host-evaluated atoms, no statistics, no
    // block filter: the whole chain is one graph launch.
    bool graphed = o.device_only && !c.host && !c.stats && !b.keep && !o.no_graph && !ctx->graphs_off;
    if (c.host) {
        if ((rc = enqueue_upload(ctx, U))) return rc;
        if (dfilter && (rc = enqueue_block_filter(ctx, jobs[0], U, ctx->keep.as<uint64_t>()))) return rc;
        if ((rc = host_round(ctx, jobs[0], U, 0, nullptr, (uint64_t)jobs[0].g.nx * jobs[0].g.ny * jobs[0].g.nz,
                             st.ms_host_callbacks)))
            return rc;
        // The kernels take segment 0 from their arguments: the pipe is made
        // only now, with the host answers host_round placed in the segment.
        const Pipe p = make_pipe(c, U);
        if (!p.seg0.P.host_vals) {
            set_error("internal error: host answers missing from the segment");
            return MC3D_E_INTERNAL;
        }
        if ((rc = enqueue_front(p, b, ctx->stream, ctx->ev))) return rc;
        if ((rc = run_output_host(ctx, c, p, b, jobs[0], U, *mb, st.ms_host_callbacks))) return rc;
        if (to_host) {
            const auto t0 = std::chrono::steady_clock::now();
            if ((rc = enqueue_host_copy(ctx, *mb, hc->totals[0], hc->totals[1], c.interior))) return rc;
            HIP_TRY(hipStreamSynchronize(ctx->copy));
            ms_d2h = ms_since(t0);
        }
    } else {
        const Pipe p = make_pipe(c, U);
        uint64_t vcap = ctx->v_hint, tcap = ctx->t_hint;
        if ((rc = take_outputs(ctx, *mb, c, vcap, tcap))) return rc;
        b.vkeys = ctx->vkeys.as<uint64_t>();
        b.vseg = c.nsegs > 1 ? ctx->vseg.as<uint32_t>() : nullptr;
        double *verts = mb->verts->as<double>(), *inter = c.interior ? mb->interior->as<double>() : nullptr;
        uint32_t *tris = mb->tris->as<uint32_t>();
        if (graphed) {
            rc = graph_launch(ctx, p, b, U, verts, tris, inter, vcap, tcap);
            if (rc == MC3D_E_UNSUPPORTED) graphed = false;  // the capture failed: enqueue stage by stage
            else if (rc) return rc;
        }
        // The totals after the scan, copied while the output stages run:
        // when the meshes fit, their copies to the host are enqueued at once.
        Ctl *he = (Ctl *)ctx->ctl_early.p;
        if (!graphed) {
            if ((rc = enqueue_upload(ctx, U))) return rc;
            if (dfilter && (rc = enqueue_block_filter(ctx, jobs[0], U, ctx->keep.as<uint64_t>()))) return rc;
            if ((rc = enqueue_front(p, b, ctx->stream, ctx->ev))) return rc;
            if (to_host) {
                HIP_TRY(hipMemcpyAsync(he, U.ctl, sizeof(Ctl), hipMemcpyDeviceToHost, ctx->stream));
                HIP_TRY(hipEventRecord(ctx->ev[EV_CTL], ctx->stream));
            }
            if ((rc = enqueue_output(p, b, U, verts, tris, inter, vcap, tcap, ctx->stream, ctx->ev))) return rc;
            HIP_TRY(hipMemcpyAsync(hc, U.ctl, U.back_bytes, hipMemcpyDeviceToHost, ctx->stream));
        }
        bool copying = false;
        if (to_host) {
            HIP_TRY(hipEventSynchronize(ctx->ev[EV_CTL]));
            const uint64_t nv = he->totals[0], nt = he->totals[1];
            if (!he->inside && nv <= vcap && nt <= tcap) {
                if ((rc = enqueue_host_copy(ctx, *mb, nv, nt, c.interior))) return rc;
                copying = true;
            }
        }
        ms_enqueue = ms_since(t_enqueue);
        const auto t_wait = std::chrono::steady_clock::now();
        HIP_TRY(hipStreamSynchronize(ctx->stream));
        ms_wait = ms_since(t_wait);
        if (hc->scan_error) {
            if (copying) (void)hipStreamSynchronize(ctx->copy);
            set_error("internal error: classify gave up waiting for the blocks before it");
            return MC3D_E_INTERNAL;
        }
        if (hc->inside) {
            if (copying) (void)hipStreamSynchronize(ctx->copy);
            size_t bad = 0;
            while (bad + 1 < jobs.size() && !so[bad].inside) bad++;
            set_error(c.nsegs == 1 ? std::string("solid is true outside of bounds")
                                   : "solid " + std::to_string(bad) + " of the batch is true outside of bounds");
            return MC3D_E_INSIDE_BOUNDARY;
        }
        const uint64_t nv = hc->totals[0], nt = hc->totals[1];
        if (nv >= (1ull << 32) || nt >= (1ull << 32)) {
            if (copying) (void)hipStreamSynchronize(ctx->copy);
            set_error("mesh too large for 32-bit vertex indices");
            return MC3D_E_UNSUPPORTED;
        }
        if (nv > vcap || nt > tcap) {  // grow and run the output stages again
            ctx->v_hint = std::max(ctx->v_hint, nv + nv / 4 + 1024);
            ctx->t_hint = std::max(ctx->t_hint, nt + nt / 4 + 1024);
            vcap = ctx->v_hint, tcap = ctx->t_hint;
            if ((rc = take_outputs(ctx, *mb, c, vcap, tcap))) return rc;
            b.vkeys = ctx->vkeys.as<uint64_t>();
            b.vseg = c.nsegs > 1 ? ctx->vseg.as<uint32_t>() : nullptr;
            // the bisection's operation counts start again with the second run
            HIP_TRY(hipMemsetAsync(U.ctl->flops + 2, 0, 2 * sizeof(unsigned long long), ctx->stream));
            if ((rc = enqueue_output(p, b, U, mb->verts->as<double>(), mb->tris->as<uint32_t>(),
                                     c.interior ? mb->interior->as<double>() : nullptr, vcap, tcap, ctx->stream,
                                     ctx->ev)))
                return rc;
            HIP_TRY(hipMemcpyAsync(hc, U.ctl, U.back_bytes, hipMemcpyDeviceToHost, ctx->stream));
            if (to_host && (rc = enqueue_host_copy(ctx, *mb, nv, nt, c.interior))) return rc;
            copying = to_host;
            HIP_TRY(hipStreamSynchronize(ctx->stream));
        }
        if (copying) {
            const auto t0 = std::chrono::steady_clock::now();
            HIP_TRY(hipStreamSynchronize(ctx->copy));
            ms_d2h = ms_since(t0);
        }
    }
    if (hc->filter_overflow) {  // the level lists are sized for every block accepted
        set_error("internal error: block filter level list overflow");
        return MC3D_E_INTERNAL;
    }
    const uint64_t nv = hc->totals[0], nt = hc->totals[1];
    if (!to_host && (rc = fit_outputs(ctx, *mb, nv, nt))) return rc;

    st.bricks_active = hc->nact_bricks;
    st.guard_evals = hc->gstats[0];
    st.guard_rechecks = hc->gstats[1];
    st.guard_mismatches = hc->gstats[2];
    st.guard_violations = hc->gstats[3];
    st.corner_flops64 = hc->flops[0];
    st.corner_flops32 = hc->flops[1];
    st.bisect_flops64 = hc->flops[2];
    st.bisect_flops32 = hc->flops[3];
    st.words_listed = hc->totals[2];
    st.brick_code_total = hc->code_total;
    st.brick_code_max = hc->code_max;
    st.bricks_full = hc->full_bricks;
    st.ms_fill = 0;  // the plan writes the constant bricks' words
    if (!graphed) {  // (a graph launch has no events between its stages)
        st.ms_plan = elapsed(ctx->ev[EV_PLAN0], ctx->ev[EV_PLAN1]);
        st.ms_corners = elapsed(ctx->ev[EV_PLAN1], ctx->ev[EV_CORNERS]);
        st.ms_classify = elapsed(ctx->ev[EV_CORNERS], ctx->ev[EV_CLASSIFY]);
        st.ms_scan = elapsed(ctx->ev[EV_CLASSIFY], ctx->ev[EV_SCAN]);
        st.ms_emit = elapsed(ctx->ev[EV_EMIT0], ctx->ev[EV_EMIT1]);
        st.ms_bisect = elapsed(ctx->ev[EV_EMIT1], ctx->ev[EV_OUT]);
    }
    st.ms_device = elapsed(ctx->ev[EV_PLAN0], ctx->ev[EV_OUT]);
    st.ms_d2h = ms_d2h;

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
        m->on_host = to_host;
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
        ms.program_length = j.low->P.code.size();
    }
    if (to_host) {  // the device outputs go back to the pool now
        ctx->pool->give(std::move(mb->verts));
        ctx->pool->give(std::move(mb->tris));
        ctx->pool->give(std::move(mb->interior));
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

// ---------------------------------------------------------------------------
// The allocation-free search: every buffer from `scratch`, outputs in the
// caller's device arrays.
// ---------------------------------------------------------------------------
// Carves the call's buffers out of [base, base + cap): the blob first, then
// the pipeline's arrays, the vertex keys last with whatever room is left
// (up to vcap keys).  Returns the key capacity, or 0 if the rest does not fit.
uint64_t carve_scratch(char *base, size_t cap, const Call &c, size_t blob_bytes, uint64_t vcap, Bufs &b,
                       size_t &need) {
    size_t off = 0;
    auto take = [&](size_t bytes) {
        char *p = base + off;
        off = align_up(off + bytes, 256);
        return p;
    };
    (void)take(blob_bytes);  // the program blob sits at the start
    b.bits = (uint64_t *)take(c.nwords * 8);
    b.counts = (uint32_t *)take(c.nscan * sizeof(uint32_t));
    b.offs = (uint2 *)take(c.nscan * sizeof(uint2));
    b.actw = (uint32_t *)take(c.nwords * sizeof(uint32_t));
    b.tiles = (uint4 *)take(classify_blocks(c) * sizeof(uint4));
    if (c.cull) {
        b.binfo = (uint2 *)take(c.nbricks * sizeof(uint2));
        b.actb = (uint32_t *)take(c.nbricks * sizeof(uint32_t));
        b.bcode = (Insn *)take(c.nbricks * c.cap * sizeof(Insn) + 16);
    }
    need = off + 8 * std::min<uint64_t>(vcap, 1);
    if (off + 256 > cap) return 0;
    b.vkeys = (uint64_t *)take(0);
    const uint64_t room = (cap - off) / 8;
    return std::min<uint64_t>(vcap, room);
}

int search_noalloc(mc3d_ctx *ctx, const mc3d_solid *s, double delta, int iters, const mc3d_options *opts,
                   double *verts, size_t vcap, uint32_t *tris, size_t tcap, size_t *nv_out, size_t *nt_out) {
    const mc3d_options o = default_options(opts);
    HIP_TRY(hipSetDevice(ctx->device));
    std::lock_guard<std::mutex> lk(g_noalloc_mu[ctx->device]);
    std::vector<Job> jobs(1);
    jobs[0].s = s, jobs[0].delta = delta, jobs[0].iters = iters;
    int rc = prepare_job(jobs[0], o);
    if (rc) return rc;
    if (jobs[0].low->host || !jobs[0].low->P.colliders.empty()) {
        set_error("mc3d_marching_cubes_search_noalloc: host-evaluated and collider solids need buffers of their "
                  "own; use mc3d_marching_cubes_search");
        return MC3D_E_UNSUPPORTED;
    }
    Call c;
    if ((rc = plan_call(jobs, o, c))) return rc;
    Upload U;
    U.cols.assign(1, {});
    layout_upload(jobs, c.nwords, c.nbricks, U);
    if (U.back_bytes > ctx->ctl_host.cap) {
        set_error("control block larger than its pinned copy");
        return MC3D_E_INTERNAL;
    }
    Bufs b;
    size_t need = 0;
    char *base = (char *)ctx->scratch;
    const uint64_t kcap = carve_scratch(base, SCRATCH_BYTES - SCRATCH_AGG_BYTES, c, U.total, vcap, b, need);
    if (!kcap && vcap) {
        set_error("mc3d_marching_cubes_search_noalloc: the grid needs " + std::to_string(need) +
                  " bytes of scratch, more than the " + std::to_string(SCRATCH_BYTES) + " the library holds");
        return MC3D_E_UNSUPPORTED;
    }
    b.table = ctx->table.as<uint8_t>();
    b.etable = ctx->etable.as<uint8_t>();
    b.agg = (ScanAgg *)(base + SCRATCH_BYTES - SCRATCH_AGG_BYTES);
    uint64_t gen;
    next_epoch(U.epoch, gen);
    if (g_scratch_agg_gen[ctx->device] != gen) {  // the epochs started over: zero the sums again
        HIP_TRY(launch_zero(b.agg, SCRATCH_AGG_BYTES, ctx->stream));
        g_scratch_agg_gen[ctx->device] = gen;
    }
    // The blob is built in host memory and copied from there (pageable:
    // the copy has taken the bytes when hipMemcpyAsync returns).
    std::vector<char> h(U.total);
    fill_upload(jobs, c.nwords, c.nbricks, b, h.data(), base, U);
    HIP_TRY(hipMemcpyAsync(base, h.data(), U.total, hipMemcpyHostToDevice, ctx->stream));
    const Pipe p = make_pipe(c, U);
    if ((rc = enqueue_front(p, b, ctx->stream, nullptr))) return rc;
    // emit and bisect keep within the capacities (extra vertices and
    // triangles are dropped and flagged), so a mesh too large is an error,
    // never a write past the caller's arrays.
    if ((rc = enqueue_output(p, b, U, verts, tris, nullptr, kcap, tcap, ctx->stream, nullptr))) return rc;
    Ctl *hc = (Ctl *)ctx->ctl_host.p;
    HIP_TRY(hipMemcpyAsync(hc, U.ctl, U.back_bytes, hipMemcpyDeviceToHost, ctx->stream));
    HIP_TRY(hipStreamSynchronize(ctx->stream));
    if (hc->scan_error) {
        set_error("internal error: classify gave up waiting for the blocks before it");
        return MC3D_E_INTERNAL;
    }
    if (hc->inside) {
        set_error("solid is true outside of bounds");
        return MC3D_E_INSIDE_BOUNDARY;
    }
    const uint64_t nv = hc->totals[0], nt = hc->totals[1];
    if (nv_out) *nv_out = (size_t)nv;
    if (nt_out) *nt_out = (size_t)nt;
    if (nv > vcap || nt > tcap) {
        set_error("mc3d_marching_cubes_search_noalloc: the mesh has " + std::to_string(nv) + " vertices and " +
                  std::to_string(nt) + " triangles, more than the output arrays hold");
        return MC3D_E_CAPACITY;
    }
    if (nv > kcap) {
        set_error("mc3d_marching_cubes_search_noalloc: " + std::to_string(nv) +
                  " vertex keys do not fit the scratch left after the grid");
        return MC3D_E_UNSUPPORTED;
    }
    return MC3D_OK;
}

}  // namespace

extern "C" {

int mc3d_ctx_create(int device, mc3d_ctx **out) {
    if (!out) return MC3D_E_INVALID_ARG;
    *out = nullptr;
    return guarded([&]() -> int {
        int n = 0;
        HIP_TRY(hipGetDeviceCount(&n));
        if (device < 0 || device >= n) {
            set_error("mc3d_ctx_create: no such HIP device");
            return MC3D_E_INVALID_ARG;
        }
        HIP_TRY(hipSetDevice(device));
        std::unique_ptr<mc3d_ctx> c(new mc3d_ctx);
        c->device = device;
        int rc = scratch_address(device, &c->scratch);
        if (rc) return rc;
        hipError_t e = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking);
        if (e == hipSuccess) e = hipStreamCreateWithFlags(&c->copy, hipStreamNonBlocking);
        if (e != hipSuccess) {
            set_error(std::string("hipStreamCreate: ") + hipGetErrorString(e));
            mc3d_ctx_destroy(c.release());
            return MC3D_E_HIP;
        }
        for (auto &ev : c->ev) {
            if (hipEventCreate(&ev) != hipSuccess) {
                set_error("hipEventCreate failed");
                mc3d_ctx_destroy(c.release());
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
        if (c->ctl_host.ensure(sizeof(Ctl) + 4096) || c->ctl_early.ensure(sizeof(Ctl) + 4096) ||
            c->table.ensure(256 * 16) || c->etable.ensure(256 * 16) ||
            hipMemcpy(c->table.p, tab, 256 * 16, hipMemcpyHostToDevice) != hipSuccess ||
            hipMemcpy(c->etable.p, etab, 256 * 16, hipMemcpyHostToDevice) != hipSuccess) {
            set_error("mc3d_ctx_create: setup failed");
            mc3d_ctx_destroy(c.release());
            return MC3D_E_HIP;
        }
        *out = c.release();
        return MC3D_OK;
    });
}

void mc3d_ctx_destroy(mc3d_ctx *ctx) {
    if (!ctx) return;
    (void)hipSetDevice(ctx->device);
    if (ctx->stream) (void)hipStreamSynchronize(ctx->stream);
    if (ctx->copy) (void)hipStreamSynchronize(ctx->copy);
    for (auto &g : ctx->graphs) (void)hipGraphExecDestroy(g.exec);
    for (auto &ev : ctx->ev)
        if (ev) (void)hipEventDestroy(ev);
    if (ctx->stream) (void)hipStreamDestroy(ctx->stream);
    if (ctx->copy) (void)hipStreamDestroy(ctx->copy);
    delete ctx;
}

}  // extern "C"

namespace {
// One solid, one call: MarchingCubesSearch, with interior points
// (MarchingCubesInterior) or with an inverse chain for the vertices
// (MarchingCubesConj).
using RectFilter = std::function<bool(const double *, const double *)>;

// The x and y edges of slab plane z among the last call's nv vertices
// (ctx->vkeys): their keys without z and the slab's vertex indices, sorted
// by key; `global` is left for their indices in the stitched mesh.
struct PlaneList {
    std::unique_ptr<DevBuf> keys{new DevBuf}, idx{new DevBuf}, global{new DevBuf};
    uint32_t n = 0;
};

int plane_list(mc3d_ctx *ctx, uint32_t nv, uint32_t z, PlaneList &out) {
    hipStream_t st = ctx->stream;
    int rc;
    size_t scan_bytes = 0, sort_bytes = 0;
    HIP_TRY(grp_scan(nullptr, scan_bytes, nullptr, nullptr, std::max<uint32_t>(nv, 1), st));
    HIP_TRY(grp_sort(nullptr, sort_bytes, nullptr, nullptr, nullptr, nullptr, std::max<uint32_t>(nv, 1), st));
    if ((rc = ctx->s_flags.ensure(nv * 4 + 16)) || (rc = ctx->s_scan.ensure(nv * 4 + 16)) ||
        (rc = ctx->s_keys.ensure(nv * 8 + 16)) || (rc = ctx->s_idx.ensure(nv * 4 + 16)) ||
        (rc = ctx->s_count.ensure(16)) || (rc = ctx->rf_total.ensure(4)) ||
        (rc = ctx->g_tmp.ensure(std::max(scan_bytes, sort_bytes) + 16)))
        return rc;
    HIP_TRY(launch_plane_list(ctx->vkeys.as<uint64_t>(), nv, z, ctx->s_flags.as<uint32_t>(),
                              ctx->s_scan.as<uint32_t>(), ctx->g_tmp.p, ctx->g_tmp.cap, ctx->s_keys.as<uint64_t>(),
                              ctx->s_idx.as<uint32_t>(), ctx->s_count.as<uint32_t>(), st));
    HIP_TRY(hipMemcpyAsync(ctx->rf_total.p, ctx->s_count.p, 4, hipMemcpyDeviceToHost, st));
    HIP_TRY(hipStreamSynchronize(st));
    const uint32_t m = *ctx->rf_total.as<uint32_t>();
    if ((rc = out.keys->ensure((size_t)m * 8 + 16)) || (rc = out.idx->ensure((size_t)m * 4 + 16)) ||
        (rc = out.global->ensure((size_t)m * 4 + 16)))
        return rc;
    if (m) {
        size_t bytes = ctx->g_tmp.cap;
        HIP_TRY(grp_sort(ctx->g_tmp.p, bytes, ctx->s_keys.as<uint64_t>(), out.keys->as<uint64_t>(),
                         ctx->s_idx.as<uint32_t>(), out.idx->as<uint32_t>(), m, st));
    }
    out.n = m;
    return MC3D_OK;
}

// MarchingCubesSearch of one solid in z slabs of `planes` cube layers
// (mc.go:614-645 keeps only a few planes alive; here device memory for the
// grid is bounded by one slab): each slab runs the pipeline on corner planes
// [c0, c1] with its shared planes open (no boundary check there), its mesh
// left on the device; then the slabs' meshes are stitched into one, the
// upper copy of each shared plane's x and y edge vertices dropped for the
// lower one.
int search_slabs(mc3d_ctx *ctx, Job &base, const mc3d_options &o, int planes, mc3d_mesh **out,
                 std::chrono::steady_clock::time_point t_call, double ms_prepare) {
    const std::vector<double> zs = base.ax[2];
    const int ncube = (int)zs.size() - 1;
    mc3d_options so = o;
    so.device_only = 1;
    // each slab's outputs are buffers of their own, so its chain would be
    // captured anew: enqueue the stages instead
    so.no_graph = 1;
    struct Piece {
        std::unique_ptr<mc3d_mesh, void (*)(mc3d_mesh *)> m{nullptr, mc3d_mesh_free};
        PlaneList top, bot;
    };
    std::vector<Piece> pieces;
    pieces.reserve((size_t)(ncube + planes - 1) / planes);
    int rc;
    for (int c0 = 0; c0 < ncube; c0 += planes) {
        const int c1 = std::min(ncube, c0 + planes);
        std::vector<Job> jobs(1);
        Job &j = jobs[0];
        j.s = base.s, j.delta = base.delta, j.iters = base.iters, j.low = base.low;
        j.ax[0] = base.ax[0], j.ax[1] = base.ax[1];
        j.ax[2].assign(zs.begin() + c0, zs.begin() + c1 + 1);
        j.interior = base.interior, j.conj = base.conj;
        j.g = Grid{};
        j.g.zlo_open = c0 > 0, j.g.zhi_open = c1 < ncube;
        if ((rc = layout_job(j, so, true))) return rc;
        std::vector<mc3d_mesh *> ms;
        if ((rc = run_call(ctx, jobs, so, ms, t_call, ms_prepare))) return rc;
        pieces.emplace_back();
        Piece &pc = pieces.back();
        pc.m.reset(ms[0]);
        if (pc.m->nv >= (1ull << 32)) {
            set_error("mesh too large for 32-bit vertex indices");
            return MC3D_E_UNSUPPORTED;
        }
        const uint32_t nv = (uint32_t)pc.m->nv;
        if (c0 > 0 && (rc = plane_list(ctx, nv, 0, pc.bot))) return rc;
        if (c1 < ncube && (rc = plane_list(ctx, nv, (uint32_t)(c1 - c0), pc.top))) return rc;
    }
    // the stitched mesh's sizes and buffers
    uint64_t nv_all = 0, nt_all = 0, nv_max = 0;
    for (size_t k = 0; k < pieces.size(); k++) {
        if (k && pieces[k].bot.n != pieces[k - 1].top.n) {
            set_error("internal error: z slabs disagree on a shared plane's vertices");
            return MC3D_E_INTERNAL;
        }
        nv_all += pieces[k].m->nv - pieces[k].bot.n;
        nt_all += pieces[k].m->nt;
        nv_max = std::max<uint64_t>(nv_max, pieces[k].m->nv);
    }
    if (nv_all >= (1ull << 32) || nt_all >= (1ull << 32)) {
        set_error("mesh too large for 32-bit vertex indices");
        return MC3D_E_UNSUPPORTED;
    }
    const bool interior = base.interior;
    std::shared_ptr<MeshBuffers> mb = std::make_shared<MeshBuffers>();
    mb->pool = ctx->pool;
    mb->hpool = ctx->hpool;
    mb->verts = ctx->pool->take(nv_all * 24 + 16);
    mb->tris = ctx->pool->take(nt_all * 12 + 16);
    if ((rc = mb->verts->ensure(nv_all * 24 + 16)) || (rc = mb->tris->ensure(nt_all * 12 + 16))) return rc;
    if (interior) {
        mb->interior = ctx->pool->take(nv_all * 24 + 16);
        if ((rc = mb->interior->ensure(nv_all * 24 + 16))) return rc;
    }
    size_t scan_bytes = 0;
    hipStream_t st = ctx->stream;
    HIP_TRY(grp_scan(nullptr, scan_bytes, nullptr, nullptr, (uint32_t)std::max<uint64_t>(nv_max, 1), st));
    if ((rc = ctx->s_keep.ensure(nv_max * 4 + 16)) || (rc = ctx->s_rank.ensure(nv_max * 4 + 16)) ||
        (rc = ctx->s_newidx.ensure(nv_max * 4 + 16)) || (rc = ctx->s_bad.ensure(16)) ||
        (rc = ctx->g_tmp.ensure(scan_bytes + 16)) || (rc = ctx->rf_total.ensure(4)))
        return rc;
    HIP_TRY(hipMemsetAsync(ctx->s_bad.p, 0, 4, st));
    uint32_t vbase = 0;
    uint64_t tbase = 0;
    uint32_t *keep = ctx->s_keep.as<uint32_t>(), *rank = ctx->s_rank.as<uint32_t>();
    uint32_t *newidx = ctx->s_newidx.as<uint32_t>();
    for (size_t k = 0; k < pieces.size(); k++) {
        const mc3d_mesh &m = *pieces[k].m;
        const uint32_t nv = (uint32_t)m.nv;
        const PlaneList &bot = pieces[k].bot;
        const double *v = m.buf->verts->as<double>() + 3 * m.voff;
        const double *in = interior ? m.buf->interior->as<double>() + 3 * m.voff : nullptr;
        const uint32_t *t = m.buf->tris->as<uint32_t>() + 3 * m.toff;
        HIP_TRY(launch_stitch_keep(keep, nv, bot.idx->as<uint32_t>(), bot.n, st));
        if (nv) {
            size_t bytes = ctx->g_tmp.cap;
            HIP_TRY(grp_scan(ctx->g_tmp.p, bytes, keep, rank, nv, st));
        }
        HIP_TRY(launch_stitch_verts(v, in, nv, keep, rank, vbase, mb->verts->as<double>(),
                                    interior ? mb->interior->as<double>() : nullptr, newidx, st));
        if (k)
            HIP_TRY(launch_stitch_pairs(pieces[k - 1].top.keys->as<uint64_t>(), pieces[k - 1].top.global->as<uint32_t>(),
                                        bot.keys->as<uint64_t>(), bot.idx->as<uint32_t>(), bot.n, newidx,
                                        ctx->s_bad.as<uint32_t>(), st));
        HIP_TRY(launch_stitch_tris(t, m.nt, newidx, mb->tris->as<uint32_t>() + 3 * tbase, st));
        const PlaneList &top = pieces[k].top;
        HIP_TRY(launch_stitch_globals(top.idx->as<uint32_t>(), top.n, newidx, top.global->as<uint32_t>(), st));
        vbase += nv - bot.n;
        tbase += m.nt;
    }
    HIP_TRY(hipMemcpyAsync(ctx->rf_total.p, ctx->s_bad.p, 4, hipMemcpyDeviceToHost, st));
    HIP_TRY(hipEventRecord(ctx->ev[EV_EMIT1], st));
    HIP_TRY(hipEventRecord(ctx->ev[EV_OUT], st));
    const bool to_host = !o.device_only;
    if (to_host && (rc = enqueue_host_copy(ctx, *mb, nv_all, nt_all, interior))) return rc;
    HIP_TRY(hipStreamSynchronize(st));
    if (*ctx->rf_total.as<uint32_t>()) {
        if (to_host) (void)hipStreamSynchronize(ctx->copy);
        set_error("internal error: z slabs disagree on a shared plane's edges");
        return MC3D_E_INTERNAL;
    }
    double ms_d2h = 0;
    if (to_host) {
        const auto t0 = std::chrono::steady_clock::now();
        HIP_TRY(hipStreamSynchronize(ctx->copy));
        ms_d2h = ms_since(t0);
    }
    // one mesh: the slabs' statistics summed, the whole grid's sizes
    std::unique_ptr<mc3d_mesh> res(new (std::nothrow) mc3d_mesh);
    if (!res) return MC3D_E_OOM;
    res->device = ctx->device;
    res->buf = mb;
    res->has_interior = interior;
    res->on_host = to_host;
    res->nv = nv_all, res->nt = nt_all;
    mc3d_stats &st_all = res->stats;
    st_all = pieces[0].m->stats;
    for (size_t k = 1; k < pieces.size(); k++) {
        const mc3d_stats &a = pieces[k].m->stats;
        st_all.guard_evals += a.guard_evals, st_all.guard_rechecks += a.guard_rechecks;
        st_all.guard_mismatches += a.guard_mismatches, st_all.guard_violations += a.guard_violations;
        st_all.bricks_total += a.bricks_total, st_all.bricks_active += a.bricks_active;
        st_all.corner_flops64 += a.corner_flops64, st_all.corner_flops32 += a.corner_flops32;
        st_all.bisect_flops64 += a.bisect_flops64, st_all.bisect_flops32 += a.bisect_flops32;
        st_all.brick_code_total += a.brick_code_total, st_all.bricks_full += a.bricks_full;
        st_all.brick_code_max = std::max(st_all.brick_code_max, a.brick_code_max);
        st_all.ms_device += a.ms_device, st_all.ms_plan += a.ms_plan, st_all.ms_corners += a.ms_corners;
        st_all.ms_classify += a.ms_classify, st_all.ms_scan += a.ms_scan, st_all.ms_emit += a.ms_emit;
        st_all.ms_bisect += a.ms_bisect, st_all.ms_enqueue += a.ms_enqueue, st_all.ms_wait += a.ms_wait;
        st_all.ms_host_callbacks += a.ms_host_callbacks, st_all.words_listed += a.words_listed;
    }
    st_all.corners = (uint64_t)base.ax[0].size() * base.ax[1].size() * zs.size();
    st_all.grid_nz = zs.size();
    st_all.vertices = nv_all, st_all.triangles = nt_all;
    st_all.bisection_evals = base.iters > 0 ? nv_all * (uint64_t)(base.iters + 1) : 0;
    st_all.ms_d2h = ms_d2h;
    st_all.ms_prepare = ms_prepare;
    st_all.ms_call = ms_since(t_call);
    if (to_host) {  // the device outputs go back to the pool now
        ctx->pool->give(std::move(mb->verts));
        ctx->pool->give(std::move(mb->tris));
        ctx->pool->give(std::move(mb->interior));
    }
    *out = res.release();
    return MC3D_OK;
}

int search_one(mc3d_ctx *ctx, const mc3d_solid *s, double delta, int iters, const mc3d_options *opts,
               mc3d_mesh **out, bool interior, std::vector<double> conj, const RectFilter *filter = nullptr,
               int filter_threads = 1, const DeviceRectFilter *dfilter = nullptr) {
    *out = nullptr;
    const auto t_call = std::chrono::steady_clock::now();
    const mc3d_options o = default_options(opts);
    HIP_TRY(hipSetDevice(ctx->device));
    std::vector<Job> jobs(1);
    jobs[0].s = s, jobs[0].delta = delta, jobs[0].iters = iters;
    jobs[0].interior = interior;
    jobs[0].conj = std::move(conj);
    const bool filtered = filter || dfilter;
    int rc = prepare_job(jobs[0], o, filtered || o.slab_planes < 0);
    if (rc) return rc;
    // z slabs: asked for, or a grid past 2^32 occupancy words
    const Job &j0 = jobs[0];
    int planes = 0;
    if (!filtered && o.slab_planes > 0) planes = o.slab_planes;
    else if (!filtered && o.slab_planes == 0 && j0.nwords >= (1ull << 32))
        planes = (int)std::max<uint64_t>(1, (1ull << 30) / ((uint64_t)j0.g.W * j0.g.ny) - 1);
    if (planes > 0 && planes < j0.g.nz - 1)
        return search_slabs(ctx, jobs[0], o, planes, out, t_call, ms_since(t_call));
    if (j0.nwords >= (1ull << 32)) {
        set_error("grid too large: more than 2^32 occupancy words");
        return MC3D_E_UNSUPPORTED;
    }
    if (filter && (rc = filter_keep_mask(jobs[0].ax, delta, *filter, filter_threads, jobs[0].keep))) return rc;
    jobs[0].dfilter = dfilter;
    const double ms_prepare = ms_since(t_call);
    std::vector<mc3d_mesh *> meshes;
    if ((rc = run_call(ctx, jobs, o, meshes, t_call, ms_prepare))) return rc;
    *out = meshes[0];
    return MC3D_OK;
}
}  // namespace

struct mc3d_rect_collider {
    RectMesh mesh;
    double cell = 0;  // the device filter's grid cell: twice the mean triangle extent
};

extern "C" {

int mc3d_marching_cubes_search_filter(mc3d_ctx *ctx, const mc3d_solid *s, mc3d_rect_filter_fn f, void *user,
                                      double delta, int iters, const mc3d_options *opts, mc3d_mesh **out) {
    if (!ctx || !s || !f || !out) {
        set_error("mc3d_marching_cubes_search_filter: null argument");
        return MC3D_E_INVALID_ARG;
    }
    return guarded([&]() -> int {
        if (f == mc3d_rect_collider_collides) {
            // the library's own rect collider: the recursion runs on the device
            const mc3d_rect_collider *rcol = (const mc3d_rect_collider *)user;
            DeviceRectFilter df;
            df.host_tris = rcol->mesh.tris.data();
            df.ntris = rcol->mesh.tris.size() / 9;
            df.cell = rcol->cell;
            return search_one(ctx, s, delta, iters, opts, out, false, {}, nullptr, 1, &df);
        }
        // a caller's filter is asked on this thread
        const RectFilter filter = [f, user](const double *mn, const double *mx) { return f(user, mn, mx) != 0; };
        return search_one(ctx, s, delta, iters, opts, out, false, {}, &filter, 1);
    });
}

mc3d_rect_collider *mc3d_rect_collider_create(const double *triangles, size_t n) {
    if (n && !triangles) {
        set_error("mc3d_rect_collider_create: null triangles");
        return nullptr;
    }
    mc3d_rect_collider *c = nullptr;
    const int rc = guarded([&]() -> int {
        std::unique_ptr<mc3d_rect_collider> m(new mc3d_rect_collider);
        m->mesh.build(triangles, n);
        double ext = 0;
        size_t finite = 0;
        for (size_t i = 0; i < n; i++) {
            const double *t = triangles + 9 * i;
            double e = 0;
            for (int a = 0; a < 3; a++)
                e = std::max(e, std::max(t[a], std::max(t[3 + a], t[6 + a])) - std::min(t[a], std::min(t[3 + a], t[6 + a])));
            if (std::isfinite(e)) ext += e, finite++;
        }
        m->cell = finite ? 2 * ext / (double)finite : 0;
        c = m.release();
        return MC3D_OK;
    });
    return rc ? nullptr : c;
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
    return guarded([&]() -> int {
        double extra = extra_space + 2 * big_delta * sqrt(3.0);  // mc.go:167
        // the coarse mesh (mc.go:169), left on the device
        mc3d_options co = default_options(opts);
        co.device_only = 1;
        mc3d_mesh *coarse = nullptr;
        int rc = search_one(ctx, s, big_delta, iters, &co, &coarse, false, {});
        if (rc) return rc;
        std::unique_ptr<mc3d_mesh, void (*)(mc3d_mesh *)> keep_coarse(coarse, mc3d_mesh_free);
        // filter = MeshToCollider(coarseMesh).RectCollision(r.Expand(extraSpace))
        // (mc.go:170-174), on the device; the coarse triangles' extents are
        // at most big_delta
        DeviceRectFilter df;
        if ((rc = mc3d_mesh_device_ptrs(coarse, &df.dev_verts, &df.dev_tris))) return rc;
        df.ntris = coarse->nt;
        df.extra = extra;
        df.cell = 2 * big_delta;
        if (df.ntris && (!df.dev_verts || !df.dev_tris)) {
            set_error("internal error: coarse mesh not on the device");
            return MC3D_E_INTERNAL;
        }
        return search_one(ctx, s, small_delta, iters, opts, out, false, {}, nullptr, 1, &df);
    });
}

int mc3d_marching_cubes_search(mc3d_ctx *ctx, const mc3d_solid *s, double delta, int iters,
                               const mc3d_options *opts, mc3d_mesh **out) {
    if (!ctx || !s || !out) {
        set_error("mc3d_marching_cubes_search: null argument");
        return MC3D_E_INVALID_ARG;
    }
    return guarded([&]() { return search_one(ctx, s, delta, iters, opts, out, false, {}); });
}

int mc3d_marching_cubes_search_noalloc(mc3d_ctx *ctx, const mc3d_solid *s, double delta, int iters,
                                       const mc3d_options *opts, double *verts, size_t vcap, uint32_t *tris,
                                       size_t tcap, size_t *n_vertices, size_t *n_triangles) {
    if (!ctx || !s || (vcap && !verts) || (tcap && !tris)) {
        set_error("mc3d_marching_cubes_search_noalloc: null argument");
        return MC3D_E_INVALID_ARG;
    }
    if (opts && opts->count_mismatch) {
        set_error("mc3d_marching_cubes_search_noalloc: statistics runs are not supported");
        return MC3D_E_UNSUPPORTED;
    }
    return guarded([&]() { return search_noalloc(ctx, s, delta, iters, opts, verts, vcap, tris, tcap, n_vertices,
                                                 n_triangles); });
}

int mc3d_marching_cubes_interior(mc3d_ctx *ctx, const mc3d_solid *s, double delta, int iters,
                                 const mc3d_options *opts, mc3d_mesh **out) {
    if (!ctx || !s || !out) {
        set_error("mc3d_marching_cubes_interior: null argument");
        return MC3D_E_INVALID_ARG;
    }
    return guarded([&]() { return search_one(ctx, s, delta, iters, opts, out, true, {}); });
}

int mc3d_marching_cubes_conj(mc3d_ctx *ctx, const mc3d_solid *s, double delta, int iters,
                             const mc3d_xform *xforms, size_t n, const mc3d_options *opts, mc3d_mesh **out) {
    if (!ctx || !s || !out || (n && !xforms)) {
        set_error("mc3d_marching_cubes_conj: null argument");
        return MC3D_E_INVALID_ARG;
    }
    *out = nullptr;
    return guarded([&]() -> int {
        // TransformSolid(JoinedTransform(xforms), s), mc.go:67-68
        mc3d_solid *t = mc3d_transform(const_cast<mc3d_solid *>(s), xforms, n);
        if (!t) return MC3D_E_INVALID_ARG;  // the error is set
        std::unique_ptr<mc3d_solid, void (*)(mc3d_solid *)> hold(t, mc3d_solid_release);
        return search_one(ctx, t, delta, iters, opts, out, false, inverse_records(xforms, n));
    });
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
    std::vector<mc3d_mesh *> got(n, nullptr);
    const int rc = guarded([&]() -> int {
        const auto t_call = std::chrono::steady_clock::now();
        const mc3d_options o = default_options(opts);
        HIP_TRY(hipSetDevice(ctx->device));
        // Bounds, spacers and lowering of every solid, on up to 16 host threads
        // (solid trees are immutable; errors are kept per solid).
        std::vector<Job> all(n);
        std::vector<int> rcs(n, MC3D_OK);
        std::vector<std::string> errs(n);
        HostPool::get().run(n, [&](size_t i) {
            all[i].s = solids[i], all[i].delta = deltas[i], all[i].iters = iters[i];
            rcs[i] = guarded([&]() { return prepare_job(all[i], o); });
            if (rcs[i]) errs[i] = mc3d_last_error();
        });
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
            if (j.low->host) {
                alone.push_back(std::move(j));
                alone_at.push_back(i);
            } else {
                batched.push_back(std::move(j));
                batched_at.push_back(i);
            }
        }
        const double ms_prepare = ms_since(t_call);
        if (!batched.empty()) {
            std::vector<mc3d_mesh *> ms;
            int r = run_call(ctx, batched, o, ms, t_call, ms_prepare);
            if (r) return r;
            for (size_t k = 0; k < ms.size(); k++) got[batched_at[k]] = ms[k];
        }
        for (size_t k = 0; k < alone.size(); k++) {
            std::vector<Job> one;
            one.push_back(std::move(alone[k]));
            std::vector<mc3d_mesh *> ms;
            int r = run_call(ctx, one, o, ms, std::chrono::steady_clock::now(), 0.0);
            if (r) {
                set_error("solid " + std::to_string(alone_at[k]) + " of the batch: " + mc3d_last_error());
                return r;
            }
            got[alone_at[k]] = ms[0];
        }
        return MC3D_OK;
    });
    if (rc) {
        for (mc3d_mesh *m : got) delete m;
        return rc;
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

int mc3d_mesh_host_ptrs(const mc3d_mesh *m, const double **verts, const uint32_t **tris) {
    if (!m || !m->on_host) {
        set_error("mc3d_mesh_host_ptrs: not a mesh fetched to the host (device_only was set)");
        return MC3D_E_INVALID_ARG;
    }
    if (verts) *verts = m->hverts();
    if (tris) *tris = m->htris();
    return MC3D_OK;
}

int mc3d_mesh_copy(mc3d_mesh *m, double *verts, uint32_t *tris) {
    if (!m) return MC3D_E_INVALID_ARG;
    return guarded([&]() -> int {
        if (m->on_host) {
            if (verts && m->nv) parallel_copy(verts, m->hverts(), m->nv * 24);
            if (tris && m->nt) parallel_copy(tris, m->htris(), m->nt * 12);
            return MC3D_OK;
        }
        std::vector<CopySpan> spans;
        if (verts && m->nv) spans.push_back({verts, m->buf->verts->as<double>() + 3 * m->voff, m->nv * 24});
        if (tris && m->nt) spans.push_back({tris, m->buf->tris->as<uint32_t>() + 3 * m->toff, m->nt * 12});
        return copy_to_caller(m->device, spans);
    });
}

size_t mc3d_stl_bytes(size_t n_triangles) { return 84 + 50 * n_triangles; }

}  // extern "C"

namespace {
// The binary STL of nt device triangles into out (host): the body by
// k_encode_stl into device scratch, one copy through pinned memory, the
// 84-byte header written here.
// GroupTriangles of n triangles on this context's device (vertices and
// triangle indices, or 9 doubles per triangle when tris is null): *perm
// receives a device array of n triangle indices in the grouped order
// (valid until the next grouping on this context).
int group_order(mc3d_ctx *ctx, const double *verts, const uint32_t *tris, uint64_t n, const uint32_t **perm) {
    *perm = nullptr;
    if (n >= (1ull << 31)) {
        set_error("GroupTriangles: more than 2^31 triangles");
        return MC3D_E_UNSUPPORTED;
    }
    hipStream_t st = ctx->stream;
    int rc;
    const size_t n4 = n * 4 + 16, n8 = n * 8 + 16;
    if ((rc = ctx->g_box.ensure(n * sizeof(GroupBox) + 16)) || (rc = ctx->g_flag.ensure(n + 16)) ||
        (rc = ctx->g_f.ensure(n4)) || (rc = ctx->g_scan.ensure(n4)) || (rc = ctx->g_axis.ensure(n + 16)))
        return rc;
    for (int a = 0; a < 4; a++)
        if ((rc = ctx->g_keys[a].ensure(n8)) || (rc = ctx->g_perm[a].ensure(n4))) return rc;
    if (n == 0) {
        *perm = ctx->g_perm[0].as<uint32_t>();
        return MC3D_OK;
    }
    const uint32_t n32 = (uint32_t)n;
    // the hipCUB temporary storage for the sorts and the scans
    size_t sort_bytes = 0, scan_bytes = 0;
    HIP_TRY(grp_sort(nullptr, sort_bytes, nullptr, nullptr, nullptr, nullptr, n32, st));
    HIP_TRY(grp_scan(nullptr, scan_bytes, nullptr, nullptr, n32, st));
    if ((rc = ctx->g_tmp.ensure(std::max(sort_bytes, scan_bytes) + 16))) return rc;
    uint64_t *keys[3] = {ctx->g_keys[0].as<uint64_t>(), ctx->g_keys[1].as<uint64_t>(), ctx->g_keys[2].as<uint64_t>()};
    uint32_t *iota = ctx->g_perm[3].as<uint32_t>();
    HIP_TRY(launch_grp_boxes(verts, tris, n32, ctx->g_box.as<GroupBox>(), keys, iota, st));
    uint32_t *P[4] = {ctx->g_perm[0].as<uint32_t>(), ctx->g_perm[1].as<uint32_t>(), ctx->g_perm[2].as<uint32_t>(),
                      iota};
    for (int a = 0; a < 3; a++) {
        size_t bytes = ctx->g_tmp.cap;
        HIP_TRY(grp_sort(ctx->g_tmp.p, bytes, keys[a], ctx->g_keys[3].as<uint64_t>(), iota, P[a], n32, st));
    }
    // the levels, until every node holds at most one triangle
    int levels = 0;
    while ((n32 + (1u << levels) - 1) >> levels > 1) levels++;
    const uint32_t big_nodes = std::max<uint32_t>(1, n32 / 64 + 1);  // nodes of a level of nodes above 64
    if ((rc = ctx->g_slots.ensure((size_t)big_nodes * 36 * 8 + 16))) return rc;
    GroupLevel G;
    G.n = n32;
    G.box = ctx->g_box.as<GroupBox>();
    G.axis = ctx->g_axis.as<uint8_t>();
    G.flag = ctx->g_flag.as<uint8_t>();
    uint32_t *spare = P[3];
    for (int L = 0; L < levels; L++) {
        G.L = L;
        for (int a = 0; a < 3; a++) G.perm[a] = P[a];
        HIP_TRY(launch_grp_axis(G, ctx->g_slots.as<unsigned long long>(), st));
        HIP_TRY(launch_grp_flags(G, st));
        for (int a = 0; a < 3; a++) {
            HIP_TRY(launch_grp_gather_flags(G, a, ctx->g_f.as<uint32_t>(), st));
            size_t bytes = ctx->g_tmp.cap;
            HIP_TRY(grp_scan(ctx->g_tmp.p, bytes, ctx->g_f.as<uint32_t>(), ctx->g_scan.as<uint32_t>(), n32, st));
            HIP_TRY(launch_grp_scatter(G, a, ctx->g_scan.as<uint32_t>(), spare, st));
            std::swap(P[a], spare);
        }
    }
    *perm = P[0];
    return MC3D_OK;
}

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
    HIP_TRY(launch_encode_stl(verts, tris, nt, ctx->stl.as<uint16_t>(), ctx->stream));
    HIP_TRY(hipMemcpyAsync(ctx->outstage.p, ctx->stl.p, body, hipMemcpyDeviceToHost, ctx->stream));
    HIP_TRY(hipStreamSynchronize(ctx->stream));
    parallel_copy(out + 84, ctx->outstage.p, body);
    return MC3D_OK;
}
}  // namespace

extern "C" {

int mc3d_mesh_encode_stl(mc3d_ctx *ctx, const mc3d_mesh *m, uint8_t *out, size_t out_bytes) {
    if (!ctx || !m || !out) {
        set_error("mc3d_mesh_encode_stl: null argument");
        return MC3D_E_INVALID_ARG;
    }
    if (m->device != ctx->device) {
        set_error("mc3d_mesh_encode_stl: the mesh is not on this context's device");
        return MC3D_E_INVALID_ARG;
    }
    return guarded([&]() -> int {
        HIP_TRY(hipSetDevice(ctx->device));
        if (m->on_host)  // the mesh's pinned copy goes up again
            return mc3d_encode_stl(ctx, m->hverts(), m->nv, m->htris(), m->nt, out, out_bytes);
        // triangle indices count from the mesh's first vertex
        return encode_stl_device(ctx, m->buf->verts->as<double>() + 3 * m->voff,
                                 m->buf->tris->as<uint32_t>() + 3 * m->toff, m->nt, out, out_bytes);
    });
}

int mc3d_group_triangles(mc3d_ctx *ctx, const double *verts, size_t nv, const uint32_t *tris, size_t nt,
                         uint32_t *perm) {
    if (!ctx || (nt && !perm) || (nv && !verts) || (nt && !tris)) {
        set_error("mc3d_group_triangles: null argument");
        return MC3D_E_INVALID_ARG;
    }
    for (size_t i = 0; i < 3 * nt; i++)
        if (tris[i] >= nv) {
            set_error("mc3d_group_triangles: triangle " + std::to_string(i / 3) + " refers to vertex " +
                      std::to_string(tris[i]) + " of " + std::to_string(nv));
            return MC3D_E_INVALID_ARG;
        }
    return guarded([&]() -> int {
        HIP_TRY(hipSetDevice(ctx->device));
        int rc;
        if ((rc = ctx->stl_in.ensure(nv * 24 + nt * 12 + 16))) return rc;
        double *dv = ctx->stl_in.as<double>();
        uint32_t *dt = (uint32_t *)(ctx->stl_in.as<char>() + align_up(nv * 24, 16));
        if (nv) HIP_TRY(hipMemcpyAsync(dv, verts, nv * 24, hipMemcpyHostToDevice, ctx->stream));
        if (nt) HIP_TRY(hipMemcpyAsync(dt, tris, nt * 12, hipMemcpyHostToDevice, ctx->stream));
        const uint32_t *dp = nullptr;
        if ((rc = group_order(ctx, dv, dt, nt, &dp))) return rc;
        if (nt) HIP_TRY(hipMemcpyAsync(perm, dp, nt * 4, hipMemcpyDeviceToHost, ctx->stream));
        HIP_TRY(hipStreamSynchronize(ctx->stream));
        return MC3D_OK;
    });
}

int mc3d_mesh_group_triangles(mc3d_ctx *ctx, mc3d_mesh *m) {
    if (!ctx || !m) {
        set_error("mc3d_mesh_group_triangles: null argument");
        return MC3D_E_INVALID_ARG;
    }
    if (m->device != ctx->device) {
        set_error("mc3d_mesh_group_triangles: the mesh is not on this context's device");
        return MC3D_E_INVALID_ARG;
    }
    if (m->has_interior) {
        set_error("mc3d_mesh_group_triangles: a mesh with interior points keeps its order");
        return MC3D_E_INVALID_ARG;
    }
    return guarded([&]() -> int {
        HIP_TRY(hipSetDevice(ctx->device));
        if (m->on_host) {
            // the pinned arrays: grouped on the device, the indices permuted here
            std::vector<uint32_t> perm(m->nt);
            int rc = mc3d_group_triangles(ctx, m->hverts(), m->nv, m->htris(), m->nt, perm.data());
            if (rc) return rc;
            uint32_t *t = (uint32_t *)m->htris();
            std::vector<uint32_t> old(t, t + 3 * m->nt);
            for (size_t k = 0; k < m->nt; k++)
                for (int c = 0; c < 3; c++) t[3 * k + c] = old[3 * (size_t)perm[k] + c];
            return MC3D_OK;
        }
        double *dv = m->buf->verts->as<double>() + 3 * m->voff;
        uint32_t *dt = m->buf->tris->as<uint32_t>() + 3 * m->toff;
        const uint32_t *dp = nullptr;
        int rc;
        if ((rc = group_order(ctx, dv, dt, m->nt, &dp))) return rc;
        if (!m->nt) return MC3D_OK;
        if ((rc = ctx->g_tris.ensure(m->nt * 12 + 16))) return rc;
        HIP_TRY(launch_grp_apply(dt, dp, (uint32_t)m->nt, ctx->g_tris.as<uint32_t>(), ctx->stream));
        HIP_TRY(hipMemcpyAsync(dt, ctx->g_tris.p, m->nt * 12, hipMemcpyDeviceToDevice, ctx->stream));
        HIP_TRY(hipStreamSynchronize(ctx->stream));
        return MC3D_OK;
    });
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
    return guarded([&]() -> int {
        HIP_TRY(hipSetDevice(ctx->device));
        int rc;
        if ((rc = ctx->stl_in.ensure(nv * 24 + nt * 12 + 16))) return rc;
        double *dv = ctx->stl_in.as<double>();
        uint32_t *dt = (uint32_t *)(ctx->stl_in.as<char>() + align_up(nv * 24, 16));
        if (nv) HIP_TRY(hipMemcpyAsync(dv, verts, nv * 24, hipMemcpyHostToDevice, ctx->stream));
        if (nt) HIP_TRY(hipMemcpyAsync(dt, tris, nt * 12, hipMemcpyHostToDevice, ctx->stream));
        return encode_stl_device(ctx, dv, dt, nt, out, out_bytes);
    });
}

int mc3d_mesh_copy_interior(mc3d_mesh *m, double *interior) {
    if (!m || !m->has_interior) {
        set_error("mc3d_mesh_copy_interior: not a mesh of mc3d_marching_cubes_interior");
        return MC3D_E_INVALID_ARG;
    }
    if (!interior || !m->nv) return MC3D_OK;
    return guarded([&]() -> int {
        if (m->on_host) {
            parallel_copy(interior, m->hinterior(), m->nv * 24);
            return MC3D_OK;
        }
        return copy_to_caller(m->device, {{interior, m->buf->interior->as<double>() + 3 * m->voff, m->nv * 24}});
    });
}

int mc3d_mesh_device_ptrs(const mc3d_mesh *m, const double **verts, const uint32_t **tris) {
    if (!m) return MC3D_E_INVALID_ARG;
    const bool have = !m->on_host && m->buf && m->buf->verts && m->buf->tris;
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

int mc3d_diag_stream(mc3d_ctx *ctx, size_t bytes, int32_t write, double *ms) {
    if (!ctx || bytes < 16) {
        set_error("mc3d_diag_stream: null context or fewer than 16 bytes");
        return MC3D_E_INVALID_ARG;
    }
    return guarded([&]() -> int {
        HIP_TRY(hipSetDevice(ctx->device));
        const uint64_t items = bytes / 16;
        int rc;
        // the stream buffer reuses the STL body scratch; the read form's
        // per-block words go to the STL input scratch
        if ((rc = ctx->stl.ensure(items * 16)) || (rc = ctx->stl_in.ensure(8192 * 4))) return rc;
        // one untimed pass writes the buffer (the read pass reads defined data)
        HIP_TRY(launch_stream(true, ctx->stl.p, items, nullptr, ctx->stream));
        HIP_TRY(hipEventRecord(ctx->ev[EV_PLAN0], ctx->stream));
        HIP_TRY(launch_stream(write != 0, ctx->stl.p, items, ctx->stl_in.as<uint32_t>(), ctx->stream));
        HIP_TRY(hipEventRecord(ctx->ev[EV_PLAN1], ctx->stream));
        HIP_TRY(hipStreamSynchronize(ctx->stream));
        if (ms) *ms = elapsed(ctx->ev[EV_PLAN0], ctx->ev[EV_PLAN1]);
        return MC3D_OK;
    });
}

int mc3d_selftest_plan_fold(uint64_t seed, uint32_t programs, uint64_t *checked, uint64_t *mismatches) {
    if (checked) *checked = 0;
    if (mismatches) *mismatches = 0;
    return guarded([&]() -> int {
        uint64_t state = seed, nchecked = 0, nbad = 0;
        auto next = [&]() {  // splitmix64
            uint64_t z = (state += 0x9E3779B97F4A7C15ull);
            z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
            z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
            return z ^ (z >> 31);
        };
        auto uni = [&](double lo, double hi) { return lo + (hi - lo) * (double)(next() >> 11) * 0x1.0p-53; };
        for (uint32_t prog = 0; prog < programs; prog++) {
            // a random expression over up to 24 leaves: spheres and boxes
            // (some in a transformed frame), constants, the five operators
            Program P;
            std::vector<AtomBox> abox;
            const int leaves = 1 + (int)(next() % 24);
            int pending = 0, made = 0;
            while (made < leaves || pending > 1) {
                const bool leaf = made < leaves && (pending < 2 || next() % 3 != 0);
                if (leaf) {
                    made++, pending++;
                    const uint32_t r = (uint32_t)(next() % 16);
                    if (r == 0 || r == 1) {
                        P.code.push_back(Insn{r ? OP_TRUE : OP_FALSE, 0});
                        continue;
                    }
                    const uint32_t frame = next() % 5 == 0 ? 1u : 0u;
                    const double c[3] = {uni(-1, 1), uni(-1, 1), uni(-1, 1)}, e = uni(0.01, 0.6);
                    AtomBox b;
                    const uint32_t at = (uint32_t)P.f64.size();
                    if (next() & 1) {  // sphere: centre, radius, threshold
                        P.f64.insert(P.f64.end(), {c[0], c[1], c[2], e, e * e, 0, 0});
                        P.code.push_back(Insn{OP_ATOM | (A_SPHERE << 8) | (frame << 16), at});
                    } else {  // box
                        P.f64.insert(P.f64.end(), {c[0] - e, c[1] - e, c[2] - e, c[0] + e, c[1] + e, c[2] + e});
                        P.code.push_back(Insn{OP_ATOM | (A_RECT << 8) | (frame << 16), at});
                    }
                    for (int k = 0; k < 3; k++) b.lo[k] = c[k] - e * 1.01, b.hi[k] = c[k] + e * 1.01;
                    P.atom_pc.push_back((uint32_t)P.code.size() - 1);
                    P.atom_box.push_back(b);
                } else if (next() % 6 == 0) {
                    P.code.push_back(Insn{OP_NOT, 0});
                } else {
                    P.code.push_back(Insn{OP_OR + (uint32_t)(next() % 4), 0});
                    pending--;
                }
            }
            std::vector<AtomBox> boxes(P.code.size());
            for (size_t i = 0; i < P.atom_pc.size(); i++) boxes[P.atom_pc[i]] = P.atom_box[i];
            std::vector<PlanNode> nodes;
            plan_nodes(P, boxes, nodes);
            const uint32_t ncode = (uint32_t)P.code.size();
            // groups of bricks near each other, as a wave's tile: a subtree is
            // skipped only when every brick of the group misses its box
            for (int group = 0; group < 4; group++) {
                constexpr int NB = 4;
                double lo[NB][3], hi[NB][3];
                const double w = group < 2 ? uni(0.001, 0.1) : uni(0.1, 1.5);
                const double c[3] = {uni(-1.6, 1.4), uni(-1.6, 1.4), uni(-1.6, 1.4)};
                for (int k = 0; k < NB; k++)
                    for (int a = 0; a < 3; a++) lo[k][a] = c[a] + uni(0, 2 * w), hi[k][a] = lo[k][a] + w;
                auto all_miss = [&](const PlanNode &t) {
                    for (int k = 0; k < NB; k++)
                        if (!node_misses(t, lo[k], hi[k])) return false;
                    return true;
                };
                for (int k = 0; k < NB; k++) {
                    const uint32_t cap = (group + k) % 4 == 0 ? (uint32_t)(next() % 4) : 64u;
                    uint32_t S1[PLAN_STACK], S2[PLAN_STACK];
                    Insn o1[64], o2[64];
                    uint32_t n1 = 0, n2 = 0;
                    bool over1 = false, over2 = false;
                    plan_fold(P.code.data(), boxes.data(), P.f64.data(), ncode, S1, lo[k], hi[k], o1, cap, n1, over1);
                    plan_fold_nodes(nodes.data(), ncode, S2, lo[k], hi[k], o2, cap, n2, over2, all_miss);
                    bool same = S1[0] == S2[0] && n1 == n2 && over1 == over2;
                    for (uint32_t i = 0; same && i < std::min(n1, cap); i++)
                        same = o1[i].x == o2[i].x && o1[i].y == o2[i].y;
                    nchecked++;
                    nbad += same ? 0 : 1;
                }
            }
        }
        if (checked) *checked = nchecked;
        if (mismatches) *mismatches = nbad;
        return MC3D_OK;
    });
}

int mc3d_contains(mc3d_ctx *ctx, const mc3d_solid *s, const double *pts, size_t n, int32_t precision,
                  uint8_t *out) {
    if (!ctx || !s || (n && (!pts || !out))) {
        set_error("mc3d_contains: null argument");
        return MC3D_E_INVALID_ARG;
    }
    if (n == 0) return MC3D_OK;
    return guarded([&]() -> int {
        HIP_TRY(hipSetDevice(ctx->device));
        std::vector<Job> jobs(1);
        Job &j = jobs[0];
        j.s = s;
        int rc = lowered(s, j.low);
        if (rc) return rc;
        j.g = Grid{nullptr, nullptr, nullptr, 1, 1, 1, 1};
        Upload U;
        if ((rc = upload(ctx, jobs, 0, 0, Bufs{}, U))) return rc;
        if ((rc = ctx->pts.ensure(n * 24)) || (rc = ctx->out8.ensure(n))) return rc;
        HIP_TRY(hipMemcpyAsync(ctx->pts.p, pts, n * 24, hipMemcpyHostToDevice, ctx->stream));
        if (!j.low->P.host_atoms.empty()) {
            double ms_cb = 0;
            if ((rc = host_round(ctx, j, U, 0, ctx->pts.as<double>(), n, ms_cb))) return rc;
        }
        HIP_TRY(launch_contains(precision == MC3D_PREC_FP32_GUARD ? 1 : 0, U.hsegs[0].P, ctx->pts.as<double>(),
                                ctx->out8.as<uint8_t>(), n, ctx->stream));
        HIP_TRY(hipMemcpyAsync(out, ctx->out8.p, n, hipMemcpyDeviceToHost, ctx->stream));
        HIP_TRY(hipStreamSynchronize(ctx->stream));
        return MC3D_OK;
    });
}

}  // extern "C"
