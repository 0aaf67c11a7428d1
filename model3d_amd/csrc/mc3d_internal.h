// mc3d_internal.h -- shared between the host-side solid lowering (solid.cpp)
// and the device pipeline (mc3d.hip).
#pragma once

#include <stdint.h>

#include <functional>
#include <memory>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/mc3d.h"

namespace mc3d {

// ---------------------------------------------------------------------------
// Device program: a postfix boolean program over "atoms".  An atom is one
// primitive Contains (or a bounds check) evaluated at the query point pushed
// through the atom's transform frame.  Ops are pure, so the evaluation order
// is free; the host orders children Sethi-Ullman style to keep the boolean
// stack shallow (it is a bit stack in one 32-bit register).
// ---------------------------------------------------------------------------
enum Op : uint32_t {
    OP_ATOM = 1,    // push atom(kind, params@y, frame)
    OP_OR = 2,      // a b -> a|b
    OP_AND = 3,     // a b -> a&b
    OP_ANDNOT = 4,  // a b -> a & !b      (SubtractedSolid: positive below negative)
    OP_RANDNOT = 5, // a b -> !a & b      (SubtractedSolid with children swapped)
    OP_NOT = 6,     // a -> !a
    OP_TRUE = 7,
    OP_FALSE = 8,
};

enum AtomKind : uint32_t {
    A_SPHERE = 1,     // cx cy cz r S_r          (S_r: largest s with sqrt(s) <= r)
    A_RECT = 2,       // minx miny minz maxx maxy maxz
    A_CAPSULE = 3,    // s0(3) s1(3) v(3) norm r S_r
    A_CYLINDER = 4,   // p2(3) dir(3) norm r S_r
    A_CONE = 5,       // base(3) dir(3) norm r
    A_CONESLICE = 6,  // p1(3) dir(3) norm r1 r2
    A_TORUS = 7,      // center(3) b1(3) b2(3) outer inner
    A_COLLIDER = 8,   // collider index (as double), min(3) max(3)
    A_HOST = 9,       // host atom index (as double)
};

// Instruction: x = op | kind << 8 | frame << 16; y = param offset (doubles).
struct Insn {
    uint32_t x, y;
};
#if defined(__HIPCC__)
#define MC3D_HD __host__ __device__
#else
#define MC3D_HD
#endif
MC3D_HD static inline uint32_t insn_op(uint32_t x) { return x & 0xffu; }
MC3D_HD static inline uint32_t insn_kind(uint32_t x) { return (x >> 8) & 0xffu; }
MC3D_HD static inline uint32_t insn_frame(uint32_t x) { return x >> 16; }

// Frame: chain of inverse elementary transforms from world space, stored in
// the double param array as records of 10 doubles {kind, v[9]}.
struct Frame {
    uint32_t offset;  // in doubles
    uint32_t nops;
};

// Node i of a program's expression tree (the subtree whose postfix code
// ends at i) as the brick plan walks it, everything a step needs in one
// 128-byte record: the instruction; `link`, which takes a leaf to the
// outermost node whose subtree starts at it and an operator to its child
// whose subtree starts where its own does; `cfalse`, the subtree's value
// when all its atoms are false; the node's box (an atom's own, an
// operator's union of its atoms', empty for constants); and for
// untransformed boxes and spheres the params of the "certainly true" test
// (rect: min, max; sphere: centre, threshold).
struct PlanNode {
    uint32_t x, y;  // the instruction
    uint32_t link, cfalse;
    double lo[3], hi[3];
    double prm[6];
    double pad[2];
};
static_assert(sizeof(PlanNode) == 128, "a plan node is one 128-byte fetch");

// Conservative world-space AABB of where an atom can be true (used only to
// cull atoms that are certainly false in a brick; never changes a result).
struct AtomBox {
    double lo[3], hi[3];
};

// ColliderContains casts every ray along one fixed direction
// (model3d/collisions.go:120-125), so the triangles can be binned once in
// the plane perpendicular to it: a query tests only the triangles whose
// (widened) projected box holds its projection, with the reference's
// float64 Moller-Trumbore leaf test (primitives.go:207-249).
constexpr int RAY_REC = 13;  // t0(3) cross1(3) invDet v1(3) v2(3)
struct RayGrid {
    std::vector<double> rec;            // RAY_REC doubles per kept triangle
    std::vector<uint32_t> cell_start;   // nu * nv + 1
    std::vector<uint32_t> cell_tris;    // record indices
    double u[3], v[3];                  // projection basis, perpendicular to the ray
    double u0, v0, inv_cell;            // grid origin and 1 / cell size
    int nu, nv;
};

// SphereCollision of a mesh collider (JoinedCollider, collisions.go:292-302,
// over Triangle.SphereCollision, primitives.go:251-279) for one radius R:
// the triangles binned in a uniform 3D grid by their boxes grown by R (far
// beyond rounding), so that a query tests only the triangles of its cell.
// Per triangle its vertices and the point-independent values of the ray
// test along its own normal.
constexpr int SPH_REC = 24;  // t0 t1 t2 (9), normal (3), hit (1: 0 never), invDet (1), cross1 (3), v1 (3), v2 (3), pad
struct SphereGrid {
    double R = 0;
    std::vector<double> rec;           // SPH_REC doubles per triangle
    std::vector<uint32_t> cell_start;  // n[0] * n[1] * n[2] + 1
    std::vector<uint32_t> cell_tris;   // triangle indices
    double o[3] = {0, 0, 0}, inv_cell = 1;
    int n[3] = {1, 1, 1};
};

struct ColliderData {
    std::vector<double> tris;   // 9 doubles per triangle
    double mn[3], mx[3];
    uint64_t id = 0;            // unique per collider (device caches)
    mutable std::once_flag grid_once;
    mutable RayGrid grid;
    const RayGrid &ray_grid() const;  // built on first use
    mutable std::mutex sphere_mu;
    mutable std::vector<std::unique_ptr<SphereGrid>> spheres;
    const SphereGrid &sphere_grid(double R) const;  // built on first use per radius
};

// Device view of a RayGrid and (insets, hollow shells) a SphereGrid.
struct DevCollider {
    const double *rec;
    const uint32_t *cell_start;
    const uint32_t *cell_tris;
    double u[3], v[3];
    double u0, v0, inv_cell;
    int nu, nv;
    const double *srec;  // null: no sphere queries
    const uint32_t *scell_start;
    const uint32_t *scell_tris;
    double so[3], sinv;
    int sn[3], pad;
};

struct HostAtom {
    mc3d_contains_fn fn;
    void *user;
};

struct Program {
    std::vector<Insn> code;
    std::vector<double> f64;
    std::vector<float> f32;        // float copies of f64 (guard-band path)
    std::vector<std::pair<uint32_t, int>> f32_dir;  // f32 slots rounded down (-1) / up (+1)
    std::vector<Frame> frames;
    std::vector<uint32_t> atom_pc; // pc of each atom instruction
    std::vector<AtomBox> atom_box; // per atom, world space
    std::vector<std::shared_ptr<const ColliderData>> colliders;
    std::vector<double> collider_R;  // per collider: the radius of its sphere queries (0: none)
    std::vector<HostAtom> host_atoms;
    int max_stack = 0;
};

struct Solid;
using SolidPtr = std::shared_ptr<const Solid>;

// Lowers a solid tree to a device program.  Returns false (with msg) when a
// node cannot be represented.
bool lower_solid(const Solid &s, Program &out, std::string &msg);
// JoinedTransform(xforms).Inverse() as frame records {kind, v[9]} per op.
std::vector<double> inverse_records(const mc3d_xform *xforms, size_t n);

// MarchingCubesFilter's block recursion (mc.go:83-135, 724-797) over the grid
// axes: keep gets one bit per cube of an accepted block, in the occupancy
// word layout (cube (x, y, z) is bit x % 64 of word (x / 64, y, z)).  f is
// asked for blocks' bounds, from `threads` threads below the first level.
int filter_keep_mask(const std::vector<double> ax[3], double delta,
                     const std::function<bool(const double *, const double *)> &f, int threads,
                     std::vector<uint64_t> &keep);

// MeshToCollider(mesh).RectCollision (collisions.go:138-179, 344-356;
// primitives.go:448-495) over a triangle list: true when a triangle whose
// bounding box overlaps the rect collides with it.
struct RectMesh {
    static constexpr uint32_t NONE = 0xffffffffu;
    struct Node {
        double mn[3], mx[3];
        uint32_t lo = 0, hi = 0, left = NONE, right = NONE;
    };
    std::vector<double> tris;  // 9 doubles per triangle
    void build(const double *tris9, size_t n);  // keeps the triangles
    bool collides(const double *mn, const double *mx) const;  // builds the tree on first use

  private:
    mutable std::once_flag tree_once;
    mutable std::vector<uint32_t> order;
    mutable std::vector<Node> nodes;
    void build_tree() const;
    uint32_t split(uint32_t lo, uint32_t hi) const;
};
void solid_bounds(const Solid &s, double mn[3], double mx[3]);
bool bounds_valid(const Solid &s);

// 256-case table (mcLookupTable, model3d/mc.go:434-456), 16 bytes per case:
// count, then up to 15 cube-edge ids.
const uint8_t *case_table();

// Square spacer axis values (newSquareSpacer, model3d/mc.go:594-608).
std::vector<double> spacer_axis(double lo, double hi, double delta);

void set_error(const std::string &msg);

}  // namespace mc3d

namespace mc3d {
// A solid's program as the pipeline takes it: the lowered postfix program,
// the frame of each host-evaluated atom and the brick plan's node records.
struct Lowered {
    Program P;
    std::vector<uint32_t> frames;  // frame of each host-evaluated atom
    std::vector<PlanNode> nodes;   // the brick plan's node records of P (plan_nodes)
    bool host = false;             // some atom is evaluated on the host
};
}  // namespace mc3d

struct mc3d_solid {
    mc3d::SolidPtr ptr;
    int refs;
    // The solid's lowered program, made by its first search and shared by
    // the later ones (a solid does not change once made).
    mutable std::mutex low_mu;
    mutable std::shared_ptr<const mc3d::Lowered> low;
};
