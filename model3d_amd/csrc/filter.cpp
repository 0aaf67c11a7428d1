// MarchingCubesFilter's block recursion and MarchingCubesC2F's coarse-mesh
// filter, on the host.  The result of the recursion is a bit per cube (the
// cubes of accepted blocks); the device pipeline then emits triangles only
// for those cubes and vertices only for the edges they use.
//
//   MarchingCubesFilter / SearchFilter / C2F   model3d/mc.go:83-180
//   mcBlock (Bounds, Volume, Pieces, Split)    model3d/mc.go:724-797
//   MeshToCollider(...).RectCollision          model3d/collisions.go:138-142, 169-179, 344-356
//   Triangle.RectCollision / SegmentCollision  model3d/primitives.go:181-249, 440-495
//   Segment.RectCollision, NewSegment          model3d/primitives.go:547-554, 632-638
//   rayCollisionWithBounds                     model3d/bvh.go:322-351
#include <math.h>
#include <string.h>

#include <algorithm>
#include <mutex>
#include <thread>

#include "mc3d_internal.h"
#include "rect_collide.h"

namespace mc3d {

using namespace rc;

// A bounding-volume tree over the coarse mesh's triangles: a rect collides
// with the mesh if it collides with a triangle whose box it overlaps.
void RectMesh::build(const double *tris9, size_t n) { tris.assign(tris9, tris9 + 9 * n); }

void RectMesh::build_tree() const {
    const size_t n = tris.size() / 9;
    order.resize(n);
    for (size_t i = 0; i < n; i++) order[i] = (uint32_t)i;
    nodes.clear();
    if (n) split(0, (uint32_t)n);
}

uint32_t RectMesh::split(uint32_t lo, uint32_t hi) const {
    Node nd;
    for (int a = 0; a < 3; a++) nd.mn[a] = INFINITY, nd.mx[a] = -INFINITY;
    for (uint32_t i = lo; i < hi; i++) {
        const double *t = &tris[9 * (size_t)order[i]];
        for (int k = 0; k < 3; k++)
            for (int a = 0; a < 3; a++) nd.mn[a] = go_min(nd.mn[a], t[3 * k + a]), nd.mx[a] = go_max(nd.mx[a], t[3 * k + a]);
    }
    nd.lo = lo, nd.hi = hi;
    const uint32_t id = (uint32_t)nodes.size();
    nodes.push_back(nd);
    if (hi - lo > 4) {
        int axis = 0;
        for (int a = 1; a < 3; a++)
            if (nd.mx[a] - nd.mn[a] > nd.mx[axis] - nd.mn[axis]) axis = a;
        const uint32_t mid = (lo + hi) / 2;
        std::nth_element(order.begin() + lo, order.begin() + mid, order.begin() + hi, [&](uint32_t p, uint32_t q) {
            const double *a = &tris[9 * (size_t)p], *b = &tris[9 * (size_t)q];
            return a[axis] + a[3 + axis] + a[6 + axis] < b[axis] + b[3 + axis] + b[6 + axis];
        });
        const uint32_t l = split(lo, mid);
        const uint32_t r = split(mid, hi);
        nodes[id].left = l, nodes[id].right = r;
    }
    return id;
}

bool RectMesh::collides(const double *mn, const double *mx) const {
    std::call_once(tree_once, [this] { build_tree(); });
    if (nodes.empty()) return false;
    uint32_t stack[128];
    int sp = 0;
    stack[sp++] = 0;
    while (sp) {
        const Node &nd = nodes[stack[--sp]];
        if (!box_overlap(mn, mx, nd.mn, nd.mx)) continue;
        if (nd.left == NONE) {
            for (uint32_t i = nd.lo; i < nd.hi; i++) {
                if (triangle_hits_rect(&tris[9 * (size_t)order[i]], mn, mx)) return true;
            }
        } else {
            stack[sp++] = nd.left;
            stack[sp++] = nd.right;
        }
    }
    return false;
}

namespace {

struct Block {
    int mn[3], mx[3];  // corner indices; cubes [mn, mx)
    long volume() const { return (long)(mx[0] - mn[0]) * (mx[1] - mn[1]) * (mx[2] - mn[2]); }
    // mcBlock.Split, mc.go:777-797
    void split(Block &b1, Block &b2) const {
        const int lx = mx[0] - mn[0], ly = mx[1] - mn[1], lz = mx[2] - mn[2];
        int axis = 0;
        if (ly >= lx && ly >= lz) axis = 1;
        else if (lz >= lx && lz >= ly) axis = 2;
        b1 = *this, b2 = *this;
        b1.mx[axis] = (mx[axis] + mn[axis]) / 2;
        b2.mn[axis] = b1.mx[axis];
    }
};

struct Recursion {
    const std::vector<double> *ax;
    double eps;
    const std::function<bool(const double *, const double *)> *f;
    // mcBlock.Bounds(epsilon), mc.go:736-740
    bool accept(const Block &b) const {
        const double mn[3] = {ax[0][b.mn[0]] + -eps, ax[1][b.mn[1]] + -eps, ax[2][b.mn[2]] + -eps};
        const double mx[3] = {ax[0][b.mx[0]] + eps, ax[1][b.mx[1]] + eps, ax[2][b.mx[2]] + eps};
        return (*f)(mn, mx);
    }
    // mcBlock.Pieces, mc.go:760-775
    template <class Leaf>
    void pieces(const Block &b, long min_volume, Leaf &&leaf) const {
        if (!accept(b)) return;
        if (b.volume() / 2 < min_volume) {
            leaf(b);
        } else {
            Block b1, b2;
            b.split(b1, b2);
            pieces(b1, min_volume, leaf);
            pieces(b2, min_volume, leaf);
        }
    }
};

}  // namespace

int filter_keep_mask(const std::vector<double> ax[3], double delta,
                     const std::function<bool(const double *, const double *)> &f, int threads,
                     std::vector<uint64_t> &keep) {
    const int nx = (int)ax[0].size(), ny = (int)ax[1].size(), nz = (int)ax[2].size();
    const int W = (nx + 63) / 64;
    keep.assign((size_t)W * ny * nz, 0);
    Recursion R{ax, delta * 1e-3, &f};
    const Block root{{0, 0, 0}, {nx - 1, ny - 1, nz - 1}};
    const long sub_volume = 64;
    const long divide = std::max(root.volume() / 4096, sub_volume);  // mc.go:128
    std::vector<Block> queue;
    R.pieces(root, divide, [&](const Block &b) { queue.push_back(b); });
    // each queued block split further (mc.go:105-107), on `threads` threads;
    // the cubes of every accepted leaf become bits of the occupancy word
    // layout (atomic ORs: leaves of different blocks can share a word)
    auto set_cubes = [&](const Block &b) {
        for (int z = b.mn[2]; z < b.mx[2]; z++)
            for (int y = b.mn[1]; y < b.mx[1]; y++) {
                uint64_t *row = &keep[((size_t)z * ny + y) * W];
                for (int x = b.mn[0]; x < b.mx[0];) {
                    const int w = x >> 6, lo = x & 63, hi = std::min(64, b.mx[0] - (w << 6));
                    const uint64_t m = (hi >= 64 ? ~0ull : ((1ull << hi) - 1)) & (~0ull << lo);
                    __atomic_fetch_or(&row[w], m, __ATOMIC_RELAXED);
                    x = (w << 6) + hi;
                }
            }
    };
    auto work = [&](size_t i0, size_t step) {
        for (size_t i = i0; i < queue.size(); i += step) R.pieces(queue[i], sub_volume, set_cubes);
    };
    const size_t nt = std::max<size_t>(1, std::min<size_t>((size_t)threads, queue.size()));
    if (nt == 1) {
        work(0, 1);
    } else {
        std::vector<std::thread> th;
        for (size_t t = 0; t < nt; t++) th.emplace_back(work, t, nt);
        for (auto &t : th) t.join();
    }
    return MC3D_OK;
}

}  // namespace mc3d
