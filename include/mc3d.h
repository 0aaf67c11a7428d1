/*
 * mc3d.h -- C-ABI of the MI355X marching-cubes library (libmc3d.so).
 *
 * Drop-in boundary for model3d's solid-to-mesh path:
 *   func MarchingCubesSearch(s Solid, delta float64, iters int) *Mesh
 *     (model3d/mc.go:45-49)
 * A caller lowers its Solid tree through the constructors below (one per
 * reference Solid type), then calls mc3d_marching_cubes_search, which runs
 * the whole path on the GPU and returns vertices (float64 xyz) and
 * triangles (uint32 vertex-index triples).  Triangle i of the reference mesh
 * is {verts[tris[3i]], verts[tris[3i+1]], verts[tris[3i+2]]}.
 *
 * All entry points return an int status (MC3D_OK == 0) or a handle (NULL on
 * failure); none throws or aborts across the ABI.  mc3d_last_error() gives
 * the message of the calling thread's last failure.  A context is not
 * thread-safe: use one per host thread (one per goroutine in Go).  Solid
 * handles are immutable and may be shared between threads and contexts.
 */
#ifndef MC3D_H
#define MC3D_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes ------------------------------------------------------ */
enum {
    MC3D_OK = 0,
    MC3D_E_INVALID_ARG = 1,
    /* panic("invalid bounds: min=%#v max=%#v"), model3d/bounder.go:48-52 */
    MC3D_E_INVALID_BOUNDS = 2,
    /* panic("solid is true outside of bounds"), model3d/mc.go:687-689 */
    MC3D_E_INSIDE_BOUNDARY = 3,
    /* panic("vertex not on edge"), model3d/mc.go:659 (unreachable: every
     * vertex is produced on its own grid edge) */
    MC3D_E_NOT_ON_EDGE = 4,
    MC3D_E_HIP = 5,
    MC3D_E_OOM = 6,
    MC3D_E_UNSUPPORTED = 7,
    /* a host Contains callback reported failure */
    MC3D_E_CALLBACK = 8,
    /* an exception inside the library (reported, never thrown across the ABI) */
    MC3D_E_INTERNAL = 9,
    /* the caller's output arrays are too small (the counts needed are returned) */
    MC3D_E_CAPACITY = 10,
};

typedef struct mc3d_ctx mc3d_ctx;
typedef struct mc3d_solid mc3d_solid;
typedef struct mc3d_mesh mc3d_mesh;

/* ---- context ----------------------------------------------------------- */
/* Binds a HIP device and a stream; owns grow-only device work buffers. */
int mc3d_ctx_create(int device, mc3d_ctx **out);
void mc3d_ctx_destroy(mc3d_ctx *ctx);

/* ---- solids (model3d's Solid implementations) -------------------------- */
/* Each constructor returns a new handle with reference count 1.  Composite
 * constructors take their own reference on each child, so callers may
 * release children right after.  Vectors are double[3] {x, y, z}. */

/* &Sphere{Center, Radius}, model3d/shapes.go:13-31 */
mc3d_solid *mc3d_sphere(const double center[3], double radius);
/* &Rect{MinVal, MaxVal}, model3d/shapes.go:145-173 */
mc3d_solid *mc3d_rect(const double min[3], const double max[3]);
/* &Capsule{P1, P2, Radius}, model3d/shapes.go:348-368 */
mc3d_solid *mc3d_capsule(const double p1[3], const double p2[3], double radius);
/* &Cylinder{P1, P2, Radius}, model3d/shapes.go:542-597 */
mc3d_solid *mc3d_cylinder(const double p1[3], const double p2[3], double radius);
/* &Cone{Tip, Base, Radius}, model3d/shapes.go:859-898 */
mc3d_solid *mc3d_cone(const double tip[3], const double base[3], double radius);
/* &ConeSlice{P1, P2, R1, R2}, model3d/shapes.go:1039-1084 */
mc3d_solid *mc3d_cone_slice(const double p1[3], const double p2[3], double r1, double r2);
/* &Torus{Center, Axis, OuterRadius, InnerRadius}, model3d/shapes.go:1333-1365 */
mc3d_solid *mc3d_torus(const double center[3], const double axis[3], double outer, double inner);
/* JoinedSolid{...}, model3d/solid.go:70-96 (n >= 1) */
mc3d_solid *mc3d_joined(mc3d_solid *const *children, size_t n);
/* &SubtractedSolid{Positive, Negative}, model3d/solid.go:118-140 */
mc3d_solid *mc3d_subtracted(mc3d_solid *positive, mc3d_solid *negative);
/* IntersectedSolid{...}, model3d/solid.go:142-170 (n >= 1) */
mc3d_solid *mc3d_intersected(mc3d_solid *const *children, size_t n);

/* Elementary transforms of model3d/transform.go.  A list of them is a
 * JoinedTransform (transform.go:191-219), applied left to right. */
enum {
    MC3D_XF_TRANSLATE = 1, /* &Translate{Offset}: v[0..2] */
    MC3D_XF_SCALE = 2,     /* &Scale{Scale}: v[0] */
    MC3D_XF_VECSCALE = 3,  /* &VecScale{Scale}: v[0..2] */
    MC3D_XF_MATRIX = 4,    /* &Matrix3Transform{Matrix} (also Rotation, Mirror): v[0..8] row-major */
};
typedef struct {
    int32_t kind;
    int32_t pad;
    double v[9];
} mc3d_xform;
/* TransformSolid(JoinedTransform(xforms), s), model3d/transform.go:240-246 */
mc3d_solid *mc3d_transform(mc3d_solid *child, const mc3d_xform *xforms, size_t n);
/* ForceSolidBounds(s, min, max), model3d/solid.go:307-309; CacheSolidBounds
 * (solid.go:316-318) is mc3d_force_bounds(s, s.Min(), s.Max()). */
mc3d_solid *mc3d_force_bounds(mc3d_solid *child, const double min[3], const double max[3]);
/* NewColliderSolid(MeshToCollider(mesh)), model3d/solid.go:256-258 and
 * model3d/collisions.go:138-142: ray-parity Contains over the triangles
 * (9 doubles each: three xyz vertices). */
mc3d_solid *mc3d_collider_solid(const double *triangles, size_t n_triangles);
/* NewColliderSolidInset(MeshToCollider(mesh), inset), model3d/solid.go:260-270:
 * Contains is ColliderContains(c, p, inset) (collisions.go:119-134) -- inside
 * by ray parity and at least `inset` from the surface (SphereCollision,
 * collisions.go:292-302, primitives.go:251-279); a negative inset also takes
 * the points within -inset outside. */
mc3d_solid *mc3d_collider_solid_inset(const double *triangles, size_t n_triangles, double inset);
/* NewColliderSolidHollow(MeshToCollider(mesh), r), model3d/solid.go:272-279:
 * Contains is SphereCollision(p, r), the shell within r of the surface. */
mc3d_solid *mc3d_collider_solid_hollow(const double *triangles, size_t n_triangles, double r);

/* A Solid the library cannot lower (a Go closure, a user type): Contains is
 * asked of the host in batches.  fn must write out[i] = Contains(p_i) for
 * the n points pts[3i..3i+2] and return 0, or nonzero to abort the search
 * with MC3D_E_CALLBACK.  It is called only from the thread that called
 * mc3d_marching_cubes_search / mc3d_contains. */
typedef int (*mc3d_contains_fn)(void *user, const double *pts, size_t n, uint8_t *out);
mc3d_solid *mc3d_host_solid(const double min[3], const double max[3], mc3d_contains_fn fn, void *user);

/* A solid does not change once made: its first search lowers it to the
 * device program, which the handle keeps (and shares with later calls, from
 * any thread) until it is released. */
mc3d_solid *mc3d_solid_retain(mc3d_solid *s);
void mc3d_solid_release(mc3d_solid *s);
/* Solid.Min() / Solid.Max() with the reference's float64 arithmetic. */
int mc3d_solid_bounds(const mc3d_solid *s, double min[3], double max[3]);

/* ---- meshing ----------------------------------------------------------- */
enum {
    MC3D_PREC_FP64 = 0,       /* every primitive evaluated in float64 */
    MC3D_PREC_FP32_GUARD = 1, /* float32 with an error band; in-band points re-decided in float64 */
};
typedef struct {
    int32_t precision;   /* MC3D_PREC_* */
    int32_t cull;        /* 1: per-brick interval culling of primitives */
    int32_t device_only; /* 1: leave results on the device; 0: copy them to pinned host memory
                          * (mc3d_mesh_host_ptrs), the triangles while the bisection runs */
    int32_t count_mismatch; /* 1: also count float32 sign mismatches (slower; statistics only) */
    int32_t brick;       /* culling brick: 64 x brick x brick corners (0: default 4) */
    int32_t rows_per_walk; /* grid rows evaluated per walk of a brick program: 1, 2 or 4 (0: default 1) */
    int32_t waves_per_brick; /* wavefronts sharing one brick: 1, 2 or 4 (0: default 4) */
    int32_t no_graph;    /* 1: enqueue the stages one by one (per-stage timings in mc3d_stats); 0: a call
                          * left on the device runs its chain as one HIP graph launch, captured once per
                          * call shape and replayed (only ms_device is timed then) */
    int32_t slab_planes; /* z-slab streaming of one solid's search: the grid runs through the pipeline
                          * this many cube layers at a time (device memory for the grid bounded by one
                          * slab; the slabs' meshes are stitched along the shared planes, so the mesh
                          * is the whole grid's).  0: whole grid unless it passes 2^32 occupancy words
                          * (then slabs of at most 2^30); < 0: never.  Not with a block filter. */
} mc3d_options;

typedef struct {
    uint64_t corners;          /* |Xs| * |Ys| * |Zs| */
    uint64_t vertices;         /* V */
    uint64_t triangles;        /* T */
    uint64_t bisection_evals;  /* V * (iters + 1) when iters > 0, else 0 */
    uint64_t guard_evals;      /* float32 primitive evaluations */
    uint64_t guard_rechecks;   /* of those, re-decided in float64 */
    uint64_t guard_mismatches; /* plain float32 decision != float64 decision (count_mismatch only) */
    uint64_t guard_violations; /* band-sure decision != float64 decision (count_mismatch only; must be 0) */
    uint64_t bricks_total;     /* culling bricks */
    uint64_t bricks_active;    /* bricks whose simplified program is not a constant */
    /* Algorithmic operation counts of the corner and bisection stages, float64
     * and float32 (count_mismatch only): per evaluated primitive a fixed count
     * of add/sub/mul/div/sqrt/compare, plus transforms and midpoints. */
    uint64_t corner_flops64, corner_flops32, bisect_flops64, bisect_flops32;
    uint64_t grid_nx, grid_ny, grid_nz; /* |Xs|, |Ys|, |Zs| */
    uint64_t brick_code_total;  /* instructions over all active bricks' folded programs */
    uint64_t brick_code_max;    /* longest folded brick program */
    uint64_t bricks_full;       /* active bricks whose folded program overflowed (full program used) */
    uint64_t program_length;    /* instructions of the whole lowered program */
    double ms_device;          /* GPU time of the whole pipeline (HIP events) */
    /* Per kernel (HIP events on the context's stream): k_plan (which also
     * writes the occupancy words of constant bricks), ms_fill (kept for the
     * layout; 0), the corner kernel, k_classify, the two scan kernels,
     * k_emit, the bisection (with host-evaluated atoms: its rounds,
     * callbacks included). */
    double ms_plan, ms_fill, ms_corners, ms_classify, ms_scan, ms_emit, ms_bisect;
    double ms_d2h;             /* copy of results to host */
    double ms_host_callbacks;  /* time spent in mc3d_contains_fn */
    double ms_call;            /* host wall time of the whole call */
    double ms_prepare;         /* host: bounds, grid axes and lowering of the call's solids */
    double ms_enqueue;         /* host: upload and kernel launches, up to the wait for the results */
    double ms_wait;            /* host: waiting for the device at the end of the call */
    uint64_t words_listed;     /* occupancy words that emit a vertex or a triangle (the call's) */
} mc3d_stats;

/* MarchingCubesSearch(s, delta, iters), model3d/mc.go:45-49 (iters == 0 is
 * MarchingCubes, mc.go:14-37).  opts may be NULL (FP32 guard + culling). */
int mc3d_marching_cubes_search(mc3d_ctx *ctx, const mc3d_solid *s, double delta, int iters,
                               const mc3d_options *opts, mc3d_mesh **out);
/* The same search with no allocation at all: no counterpart in the
 * reference (MarchingCubesSearch, mc.go:45-49, returns a new *Mesh); it is
 * for callers that cannot allocate device memory (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc).  Every
 * buffer the pipeline needs comes from a fixed 64 MiB device array per GPU,
 * held by the library until the process ends; grids that need more (about
 * 100M corners with culling at the default brick) are MC3D_E_UNSUPPORTED.
 * The caller passes device arrays for the outputs: verts (3 * vcap doubles)
 * and tris (3 * tcap uint32, vertex indices).  *n_vertices / *n_triangles
 * receive the mesh's counts; a mesh larger than the arrays is
 * MC3D_E_CAPACITY (nothing is written beyond them; zero capacities ask for
 * the counts).  The result is bit-identical to mc3d_marching_cubes_search
 * with the same options.  Host-evaluated and collider solids are
 * MC3D_E_UNSUPPORTED (their tables need buffers of their own).  Calls on one
 * GPU share its array: the library runs them one at a time (calls from
 * other threads wait), and each call keeps the array's contents only until
 * the next one. */
int mc3d_marching_cubes_search_noalloc(mc3d_ctx *ctx, const mc3d_solid *s, double delta, int iters,
                                       const mc3d_options *opts, double *verts, size_t vcap, uint32_t *tris,
                                       size_t tcap, size_t *n_vertices, size_t *n_triangles);
/* MarchingCubesInterior(s, delta, iters), model3d/mc.go:55-60: the mesh of
 * MarchingCubesSearch plus, for every vertex, a nearby point known to be
 * inside the solid -- the bisection bracket's inside end (mc.go:252-255; with
 * iters == 0 the inside corner of the vertex's edge).  The reference returns
 * the points as a CoordMap keyed by vertex; here they come aligned with the
 * vertices, from mc3d_mesh_copy_interior. */
int mc3d_marching_cubes_interior(mc3d_ctx *ctx, const mc3d_solid *s, double delta, int iters,
                                 const mc3d_options *opts, mc3d_mesh **out);
/* MarchingCubesConj(s, delta, iters, xforms...), model3d/mc.go:66-71: the
 * search of TransformSolid(JoinedTransform(xforms), s), its vertices then
 * mapped by the joined transform's Inverse() (Mesh.Transform, mesh.go:613-642)
 * on the device. */
int mc3d_marching_cubes_conj(mc3d_ctx *ctx, const mc3d_solid *s, double delta, int iters,
                             const mc3d_xform *xforms, size_t n, const mc3d_options *opts, mc3d_mesh **out);
/* A block filter for MarchingCubesFilter: nonzero when the rect [min, max]
 * may touch the surface (it may be conservative, never wrong the other way). */
typedef int (*mc3d_rect_filter_fn)(void *user, const double min[3], const double max[3]);
/* MeshToCollider(mesh).RectCollision (collisions.go:138-179, 344-356;
 * Triangle.RectCollision, primitives.go:448-495) over n triangles (9 doubles
 * each): true when a triangle whose bounding box overlaps the rect collides
 * with it.  mc3d_rect_collider_collides has the mc3d_rect_filter_fn shape
 * (user = the collider) and may be called from several threads. */
typedef struct mc3d_rect_collider mc3d_rect_collider;
mc3d_rect_collider *mc3d_rect_collider_create(const double *triangles, size_t n);
int mc3d_rect_collider_collides(void *collider, const double min[3], const double max[3]);
void mc3d_rect_collider_free(mc3d_rect_collider *c);
/* MarchingCubesFilter(s, f, delta), model3d/mc.go:73-135: marching cubes
 * only in the blocks f accepts -- the reference's block recursion (mcBlock
 * Pieces/Split, mc.go:724-797), its triangles those of the cubes of accepted
 * blocks.  A caller's f is asked on the calling thread; with
 * f == mc3d_rect_collider_collides the recursion and the collider's test run
 * on the device (one wave per block, the triangles binned in a grid).  Every
 * corner is still evaluated, so "solid is true outside of bounds" covers the
 * whole grid. */
int mc3d_marching_cubes_filter(mc3d_ctx *ctx, const mc3d_solid *s, mc3d_rect_filter_fn f, void *user, double delta,
                               const mc3d_options *opts, mc3d_mesh **out);
/* MarchingCubesSearchFilter(s, f, delta, iters), mc.go:137-143 */
int mc3d_marching_cubes_search_filter(mc3d_ctx *ctx, const mc3d_solid *s, mc3d_rect_filter_fn f, void *user,
                                      double delta, int iters, const mc3d_options *opts, mc3d_mesh **out);
/* MarchingCubesC2F(s, bigDelta, smallDelta, extraSpace, iters), mc.go:145-180:
 * the coarse search, its mesh as the rect collider (MeshToCollider(...)
 * .RectCollision of each block's bounds expanded by extraSpace +
 * 2 sqrt(3) bigDelta) of the block recursion, then the filtered fine search,
 * all on the device.  bigDelta < smallDelta is MC3D_E_INVALID_ARG with the
 * reference's panic text. */
int mc3d_marching_cubes_c2f(mc3d_ctx *ctx, const mc3d_solid *s, double big_delta, double small_delta,
                            double extra_space, int iters, const mc3d_options *opts, mc3d_mesh **out);
/* MarchingCubesSearch(solids[i], deltas[i], iters[i]) for i < n in one pass
 * of the pipeline (one launch per stage for all solids; solids with
 * host-evaluated atoms run alone).  out[i] receives solid i's mesh, an
 * independent mc3d_mesh.  The reference meshes one solid per call
 * (mc.go:45-49); this is that call over a list, for callers with many
 * solids.  On error no mesh is returned and the message names the solid. */
int mc3d_marching_cubes_search_batch(mc3d_ctx *ctx, const mc3d_solid *const *solids, const double *deltas,
                                     const int32_t *iters, size_t n, const mc3d_options *opts, mc3d_mesh **out);
int mc3d_mesh_counts(const mc3d_mesh *m, size_t *n_vertices, size_t *n_triangles);
/* Copies the results to host arrays of 3*V doubles and 3*T uint32 (either
 * may be NULL).  For a mesh left on the device (device_only == 1) the copy
 * runs in 4 MiB chunks through pinned buffers of the library, each chunk's
 * host copy overlapping the next chunks' DMA. */
int mc3d_mesh_copy(mc3d_mesh *m, double *verts, uint32_t *tris);
/* The results of a mesh searched with device_only == 0, in the library's
 * pinned host memory (3*V doubles, 3*T uint32; valid until mc3d_mesh_free):
 * a caller that builds its own mesh objects (the reference's *Mesh of
 * *Triangle, mc.go:20-35) reads them in place, with no further copy.
 * MC3D_E_INVALID_ARG for a mesh left on the device. */
int mc3d_mesh_host_ptrs(const mc3d_mesh *m, const double **verts, const uint32_t **tris);
/* Device pointers of the results (owned by the mesh; valid until mc3d_mesh_free). */
int mc3d_mesh_device_ptrs(const mc3d_mesh *m, const double **verts, const uint32_t **tris);
int mc3d_mesh_stats(const mc3d_mesh *m, mc3d_stats *out);
/* The interior point of every vertex (3 doubles each, in mc3d_mesh_copy's
 * vertex order) of a mesh from mc3d_marching_cubes_interior;
 * MC3D_E_INVALID_ARG for any other mesh. */
int mc3d_mesh_copy_interior(mc3d_mesh *m, double *interior);
void mc3d_mesh_free(mc3d_mesh *m);

/* Binary STL, as EncodeSTL / WriteSTL write it (model3d/export.go:20-66,
 * fileformats/stl.go:17-58): 80 zero bytes, the triangle count (uint32),
 * then per triangle the float32 cast of Triangle.Normal() computed in
 * float64 (primitives.go:27-33), the float32 casts of its three vertices and
 * two zero bytes; little-endian.  Triangles in the given order (the
 * reference writes Mesh.TriangleSlice's map order; for SaveGroupedSTL's,
 * mc3d_mesh_group_triangles / mc3d_group_triangles first).  The encoding
 * runs on the device.  out must hold mc3d_stl_bytes(T) bytes; more than 2^32 - 1
 * triangles is MC3D_E_UNSUPPORTED ("too many triangles for STL format"). */
size_t mc3d_stl_bytes(size_t n_triangles);
/* a mesh of this context's device, from its device buffers */
int mc3d_mesh_encode_stl(mc3d_ctx *ctx, const mc3d_mesh *m, uint8_t *out, size_t out_bytes);
/* host arrays: verts 3 doubles per vertex, tris 3 vertex indices per triangle */
int mc3d_encode_stl(mc3d_ctx *ctx, const double *verts, size_t nv, const uint32_t *tris, size_t nt, uint8_t *out,
                    size_t out_bytes);

/* GroupTriangles(faces) (model3d/bvh.go:118-156, 219-292) on the device:
 * perm[k] = the index of the k-th triangle of the grouped order (sorted by
 * box midpoints per axis, recursively cut at n / 2 along the axis whose two
 * halves' box areas sum least).  The reference's sort.Slice orders equal
 * midpoints arbitrarily; here they keep their index order.  Triangles with
 * NaN coordinates get an unspecified (but complete) order.  Host arrays as
 * for mc3d_encode_stl. */
int mc3d_group_triangles(mc3d_ctx *ctx, const double *verts, size_t nv, const uint32_t *tris, size_t nt,
                         uint32_t *perm);
/* The mesh's triangles reordered in place by GroupTriangles, as
 * SaveGroupedSTL (mesh.go:721-743) writes them: then
 * mc3d_mesh_encode_stl / mc3d_mesh_copy give the grouped order.  Not for a
 * MarchingCubesInterior mesh. */
int mc3d_mesh_group_triangles(mc3d_ctx *ctx, mc3d_mesh *m);

/* Solid.Contains for n host points on the GPU (out[i] = 0/1). */
int mc3d_contains(mc3d_ctx *ctx, const mc3d_solid *s, const double *pts, size_t n, int32_t precision,
                  uint8_t *out);

/* The 256-case marching cubes table the kernels use (mcLookupTable,
 * model3d/mc.go:434-456): out[16*k] = triangle count of case k, then up to
 * 15 cube-edge ids (3 per triangle, CCW from outside).  Edge id e: axis
 * e/4 (0=x, 1=y, 2=z), and the two remaining corner bits in e%4. */
int mc3d_case_table(uint8_t out[256 * 16]);

/* Diagnostics: one pass of a plain streaming kernel over `bytes` of device
 * memory -- k_stream_read reads every byte once, k_stream_write (write != 0)
 * writes every byte once -- after an untimed pass that fills the buffer;
 * *ms gets the timed pass's duration (HIP events).  Calibrates the HBM
 * traffic counters (FETCH_SIZE, WRITE_SIZE) against known bytes and
 * measures the bandwidth a plain stream reaches.  No counterpart in the
 * reference. */
int mc3d_diag_stream(mc3d_ctx *ctx, size_t bytes, int32_t write, double *ms);

/* Diagnostics (host only, no GPU): folds random programs against random
 * brick boxes with the brick plan's two folds -- every instruction in turn,
 * and with subtrees whose atoms all miss the brick skipped -- and counts the
 * cases whose folded code, root or overflow differ (must be 0).  No
 * counterpart in the reference; it checks the library's own culling. */
int mc3d_selftest_plan_fold(uint64_t seed, uint32_t programs, uint64_t *checked, uint64_t *mismatches);

const char *mc3d_last_error(void);
const char *mc3d_version(void);

#ifdef __cplusplus
}
#endif
#endif
