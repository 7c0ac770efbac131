/*
 * fclx.h — C ABI of the MI355X proximity-query engine.
 *
 * This is the boundary the FCL mesh hot path crosses.  It replaces, for
 * BVHModel<OBBRSS>/BVHModel<RSS> geometry:
 *   - fcl::collide  (include/fcl/narrowphase/collision-inl.h:93-149, 178-207)
 *     through the dispatch-table cell collision_matrix[BV_OBBRSS][BV_OBBRSS]
 *     (narrowphase/detail/collision_func_matrix-inl.h:571-626, 851);
 *   - fcl::distance (include/fcl/narrowphase/distance-inl.h:90-193, 222-246)
 *     through distance_matrix[BV_RSS][BV_RSS] and [BV_OBBRSS][BV_OBBRSS]
 *     (narrowphase/detail/distance_func_matrix-inl.h:386-458, 661-663);
 *   - BVHModel<BV>::beginModel/addSubModel/endModel
 *     (include/fcl/geometry/bvh/BVH_model-inl.h:206-253, 382-517), whose
 *     built hierarchy is flattened and kept resident in device memory.
 * Batched entry points run many queries in one launch on a caller stream.
 *
 * Conventions: plain pointers and sizes only.  A pose is 12 doubles: the
 * rotation row-major (9) followed by the translation (3).  Pointers named
 * "device" must point to device memory of the context's GPU; all others are
 * host memory.  Every function returns FCLX_OK (0) or a negative FCLX_ERR_*
 * code; no exception crosses the boundary.  fclx_last_error() gives text.
 * A context may be used from several threads.  Single-query calls
 * (fclx_collide, fclx_distance) share staging buffers and serialise on an
 * internal lock; FAST-engine collide batches with num_max_contacts above 64
 * share the context's work buffers and serialise on a second lock (they
 * return once the batch is done); other batched calls only enqueue work.
 * The status word that fclx_sync reports and clears is per context.
 */
#ifndef FCLX_H
#define FCLX_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define FCLX_OK 0
#define FCLX_ERR_INVALID_ARGUMENT (-1)
#define FCLX_ERR_OUT_OF_MEMORY (-2)
#define FCLX_ERR_CAPACITY (-3)       /* output buffer too small; see fclx_sync */
#define FCLX_ERR_DEVICE (-4)         /* HIP runtime error */
#define FCLX_ERR_UNSUPPORTED (-5)
#define FCLX_ERR_EMPTY_MODEL (-6)    /* cf. BVH_ERR_BUILD_EMPTY_MODEL (BVH_internal.h:58-69) */
#define FCLX_ERR_STACK_OVERFLOW (-7) /* a traversal exceeded its stack bound */

#define FCLX_BV_OBBRSS 0
#define FCLX_BV_RSS 1

/* Traversal engines.  PLAIN: one query per lane, FP64 throughout, the
 * reference's depth-first order (contacts come out in the reference's order
 * for every num_max_contacts).  FAST: wavefront-cooperative traversal with
 * FP32 conservative culling and FP64 leaves; the same contact set as PLAIN
 * when a query's contact count stays below num_max_contacts (records of such
 * a query come out in no particular order).  A query that reaches the cap
 * gets the reference's depth-first prefix, in order: its records are ordered
 * by depth-first path keys and cut, or, when per-query statistics are
 * requested, the query is re-run by PLAIN.  Batches over a context holding
 * a mesh of 2^20 or more triangles, and batches of 2^22 or more queries, run
 * PLAIN for caps above 64 (the FAST engine's task encoding has 21 bits per
 * node index and 22 per query). */
#define FCLX_ENGINE_PLAIN 0
#define FCLX_ENGINE_FAST 1
/* BOOL: only whether each query's pair collides -- counts[q] = 1 or 0, the
 * same as min(collide() count, 1) for any num_max_contacts -- with no
 * records and no statistics (bv_tests / leaf_tests are zeroed); a wavefront
 * keeps 4 or 8 queries in flight, each descending nearly depth first on a
 * stack of its own, and stops a query at its first intersecting triangle
 * pair.  Batches on different streams may run at once (up to eight without
 * waiting for each other's work areas).  Batches
 * beyond 2^22 queries or over meshes of 2^20 or more triangles run the
 * PLAIN kernel with a cap of 1 (same answers). */
#define FCLX_ENGINE_BOOL 2

typedef struct fclx_context fclx_context;

typedef struct {
  int32_t num_bvs;
  int32_t num_tris;
  int32_t num_vertices;
  int32_t depth;          /* levels, root = 1 */
  int32_t bv_type;
  int32_t reserved;
  int64_t device_bytes;   /* HBM held by the mesh */
} fclx_mesh_info;

typedef struct {
  int32_t num_queries;
  const int32_t* query_meshes; /* device: 2 mesh ids per query */
  const double* tf1;           /* device: 12 per query */
  const double* tf2;           /* device: 12 per query */
  int32_t num_max_contacts;    /* CollisionRequest::num_max_contacts (>= 1) */
  int32_t engine;              /* FCLX_ENGINE_* */
  int32_t* counts;             /* device: contacts per query */
  int32_t* pairs;              /* device: 3 int32 per record (query, b1, b2); may be NULL */
  int64_t pair_capacity;       /* records that fit in pairs */
  int64_t* num_pairs;          /* device: 1 int64, records produced (set by the call) */
  int64_t* bv_tests;           /* device, optional: per-query num_bv_tests */
  int64_t* leaf_tests;         /* device, optional: per-query num_leaf_tests */
  /* CollisionRequest::enable_contact (mesh_collision_traversal_node-inl.h:
   * 578-598): each record is one contact (a triangle pair gives 0-2), counts
   * and num_max_contacts count contacts, and the geometry below is written
   * per record in the world frame.  Runs the reference-order kernel whatever
   * the engine.  The three arrays are required when pairs is set. */
  int32_t enable_contact;
  int32_t reserved;
  double* contact_pos;         /* device: 3 per record */
  double* contact_normal;      /* device: 3 per record */
  double* contact_depth;       /* device: 1 per record */
} fclx_collide_batch_args;

typedef struct {
  int32_t num_queries;
  const int32_t* query_meshes; /* device: 2 mesh ids per query */
  const double* tf1;           /* device */
  const double* tf2;           /* device */
  int32_t enable_nearest_points;
  int32_t engine;              /* FCLX_ENGINE_PLAIN: one query per lane, the
                                  reference's visit order, every output and
                                  statistic as the reference's.
                                  FCLX_ENGINE_FAST: a wavefront searches
                                  several queries 64 bounds at a time; the
                                  same min_distance, ids and points as the
                                  reference (its visit order is kept as path
                                  keys for ties); the statistics count this
                                  engine's own tests */
  double* distance;            /* device: min_distance per query */
  double* points;              /* device, optional: 6 per query, world frame */
  int32_t* ids;                /* device, optional: (b1, b2) per query */
  int64_t* bv_tests;           /* device, optional */
  int64_t* leaf_tests;         /* device, optional */
} fclx_distance_batch_args;

/* Primitive shapes (geometry/shape/{box,sphere,capsule,cylinder,convex}.h),
 * in their own local frames as the reference defines them: a box centred
 * at the origin with the given side lengths, capsule and cylinder along z
 * with total length lz (capsule: of its segment), a convex as its vertex
 * list (the GJK support scans the vertices, minkowski_diff-inl.h:182-196,
 * so faces are not needed). */
#define FCLX_SHAPE_BOX 0
#define FCLX_SHAPE_SPHERE 1
#define FCLX_SHAPE_CAPSULE 2
#define FCLX_SHAPE_CYLINDER 3
#define FCLX_SHAPE_CONVEX 4

typedef struct {
  int32_t type;           /* FCLX_SHAPE_* */
  int32_t num_vertices;   /* CONVEX: vertices in the pool */
  double dims[3];         /* BOX: side x, y, z; SPHERE: radius; CAPSULE and
                             CYLINDER: radius, lz */
  int64_t vertex_offset;  /* CONVEX: index of its first vertex in the pool */
} fclx_shape;

typedef struct {
  int32_t num_pairs;
  int32_t num_shapes;
  const fclx_shape* shapes;    /* device: num_shapes records */
  const double* vertices;      /* device: convex vertex pool, 3 doubles each (NULL
                                  when no shape is CONVEX) */
  int64_t num_vertices;        /* vertices in the pool */
  const int32_t* pair_shapes;  /* device: 2 shape ids per pair */
  const double* tf1;           /* device: 12 per pair */
  const double* tf2;           /* device: 12 per pair */
  int32_t num_max_contacts;    /* CollisionRequest::num_max_contacts (>= 1) */
  int32_t enable_contact;      /* CollisionRequest::enable_contact */
  double gjk_tolerance;        /* CollisionRequest::gjk_tolerance (> 0; the
                                  reference's default is 1e-6), used by GJK and
                                  EPA alike (collision-inl.h:166-169) */
  int32_t* counts;             /* device: contacts per pair */
  /* device, required with enable_contact: slot = pair * num_max_contacts + i
   * for i < counts[pair]; world-frame point (3), normal (3, pointing from
   * the first shape to the second) and penetration depth (1) */
  double* contact_pos;
  double* contact_normal;
  double* contact_depth;
  /* device, optional: 2 per pair, the passes of GJK's loop and of EPA's
   * loop (gjk-inl.h:101-196, epa-inl.h:283-321); 0 for closed-form pairs,
   * and EPA's is 0 without enable_contact, where it is not needed */
  int32_t* passes;
} fclx_shape_collide_batch_args;

/* Context on one GPU (device ordinal).  No counterpart in the reference,
 * which has no device state. */
int fclx_init(int32_t device, fclx_context** out);
void fclx_shutdown(fclx_context* ctx);
const char* fclx_last_error(const fclx_context* ctx);

/* Build the BVH on the host exactly as BVHModel<BV>::endModel() does
 * (BVH_model-inl.h:449-517, 832-938; mean split, default fitter) and upload
 * it.  vertices: 3 doubles each; tris: 3 int32 vertex indices each.
 * *mesh_id receives the id that queries use. */
int fclx_mesh_create(fclx_context* ctx, int32_t bv_type, const double* vertices,
                     int32_t num_vertices, const int32_t* tris, int32_t num_tris,
                     int32_t* mesh_id);
int fclx_mesh_destroy(fclx_context* ctx, int32_t mesh_id);
int fclx_mesh_get_info(fclx_context* ctx, int32_t mesh_id, fclx_mesh_info* info);
/* Host copy of the built hierarchy (BVHModel::getBV, BVH_model.h:86-92):
 * per node 31 doubles = OBB axis(9, row-major) To(3) extent(3), RSS axis(9)
 * To(3) l(2) r(1); links = first_child, first_primitive, num_primitives;
 * sizes = BV::size().  Any output pointer may be NULL. */
int fclx_mesh_export_nodes(fclx_context* ctx, int32_t mesh_id, double* bv_data,
                           int32_t* links, double* sizes);

/* BVHModel::beginReplaceModel / replaceSubModel(vertices) /
 * endReplaceModel(refit = true, bottomup) (BVH_model-inl.h:521-620) on the
 * device: the mesh's vertices are replaced by `vertices` (device memory, 3
 * doubles per vertex, the mesh's vertex count) and its hierarchy refit in
 * place, keeping the topology (refitTree, :941-1076).  bottomup = 1 (the
 * reference's default): leaves fitted to their triangles and parents merged
 * from their children, level by level; 0: every node fitted to its
 * triangles as the build does.  Runs on `stream` and returns once done (the
 * mesh's cull bound is updated on the host); queries on the mesh must not
 * run meanwhile.  fclx_mesh_export_nodes then reports the refit BVs.
 * The reference's merges (OBB::operator+, RSS::operator+) need not contain
 * both children, so after a bottom-up refit a BV's RSS distance is not
 * always a lower bound of the triangle pairs below it, and distance()
 * depends on the visit order: while any mesh of the context holds merged
 * BVs, FCLX_ENGINE_FAST distance batches run the reference-order PLAIN
 * kernel (a top-down refit clears this).  Collide batches are unaffected:
 * their results do not depend on the visit order. */
int fclx_mesh_refit(fclx_context* ctx, int32_t mesh_id, const double* vertices,
                    int32_t bottomup, void* stream);

/* Host-only BVH build with no device and no context: the same hierarchy
 * fclx_mesh_create uploads, written to caller arrays sized for 2*num_tris-1
 * nodes in the fclx_mesh_export_nodes layout.  *depth may be NULL. */
int fclx_bvh_build(int32_t bv_type, const double* vertices, int32_t num_vertices,
                   const int32_t* tris, int32_t num_tris, double* bv_data, int32_t* links,
                   double* sizes, int32_t* depth);

/* Batched collide()/distance() on a HIP stream (NULL = default stream).
 * Asynchronous: completion and overflow status come from fclx_sync. */
int fclx_collide_batch(fclx_context* ctx, const fclx_collide_batch_args* args,
                       void* stream);
int fclx_distance_batch(fclx_context* ctx, const fclx_distance_batch_args* args,
                        void* stream);
/* No counterpart in the reference.  For callers that cannot allocate
 * device memory during a call -- code under HIP graph capture, tight
 * serving loops, harnesses that forbid hipMalloc: fclx_collide_batch with
 * FCLX_ENGINE_BOOL (counts[q] = min(collide() count, 1); *num_pairs set to
 * 0; bv_tests / leaf_tests zeroed when given) that allocates nothing.  Its
 * per-query stacks live in a fixed device array of
 * FCLX_NOALLOC_SCRATCH_BYTES that the library holds on every GPU it runs
 * on until the process ends (fclx_init resolves its address); the grid is
 * cut to the wavefronts whose stacks fit, so large batches take longer
 * than with fclx_collide_batch, with the same answers.  Limits: engine
 * FCLX_ENGINE_BOOL, enable_contact 0, fewer than 2^22 queries per call,
 * meshes of fewer than 2^20 triangles and hierarchies within the traversal
 * stack bound; anything else returns FCLX_ERR_UNSUPPORTED (or
 * FCLX_ERR_INVALID_ARGUMENT) without launching.  Every call on one GPU
 * shares that GPU's array: calls that overlap on one GPU (other streams,
 * other contexts, other threads) race and give wrong answers; calls
 * ordered on one stream, or separated by a synchronisation, are safe.
 * The array keeps the last call's contents; each call clears the counters
 * it needs. */
#define FCLX_NOALLOC_SCRATCH_BYTES (64LL << 20)
int fclx_collide_batch_noalloc(fclx_context* ctx, const fclx_collide_batch_args* args,
                               void* stream);
/* Batched fcl::collide(mesh, tf1, shape, tf2) for a BVHModel<OBBRSS> and a
 * primitive (BOX, SPHERE, CAPSULE, CYLINDER, or CONVEX with >= 4 vertices)
 * with the GST_INDEP solver, enable_cost off: per query what
 * BVHShapeCollider<OBBRSS> (orientedBVHShapeCollide,
 * collision_func_matrix-inl.h:375-419, 500-521) adds to an empty result --
 * contacts (b1 = triangle id, b2 = Contact::NONE) in the reference's order,
 * at most num_max_contacts, from meshShapeCollisionOrientedNodeLeafTesting
 * (mesh_shape_collision_traversal_node-inl.h:205-283).  Asynchronous; a
 * query naming a missing mesh or shape, an RSS mesh, a mesh deeper than 4096
 * levels or a convex with fewer than 4 vertices gets count 0 and makes
 * fclx_sync return FCLX_ERR_INVALID_ARGUMENT.  While the context holds a
 * mesh deeper than 64 levels, batches keep their descent stacks in
 * stream-ordered device memory (4 bytes per level per resident lane) instead
 * of LDS.  With enable_contact the call
 * takes stream-ordered device memory for EPA (about 12 KB per resident lane,
 * freed when the batch ends). */
typedef struct {
  int32_t num_queries;
  const int32_t* query_meshes; /* device: mesh id per query */
  const int32_t* query_shapes; /* device: shape id per query (into shapes) */
  int32_t num_shapes;
  const fclx_shape* shapes;    /* device: num_shapes records */
  const double* vertices;      /* device: convex vertex pool (NULL if none) */
  int64_t num_vertices;
  const double* tf1;           /* device: the mesh's pose, 12 per query */
  const double* tf2;           /* device: the shape's pose, 12 per query */
  int32_t num_max_contacts;    /* >= 1 */
  int32_t enable_contact;
  double gjk_tolerance;        /* > 0 (the reference's default is 1e-6) */
  int32_t* counts;             /* device: contacts per query */
  int32_t* primitives;         /* device, optional: slot = query *
                                  num_max_contacts + i: the triangle (b1) */
  /* device, required with enable_contact, per slot: the reference's
   * Contact::pos (3), normal (3) and penetration_depth (1).  As in the
   * reference, a sphere's normal points from the triangle to the sphere and
   * the other shapes' from the shape to the triangle, and depths are <= 0. */
  double* contact_pos;
  double* contact_normal;
  double* contact_depth;
  int64_t* stats;              /* device, optional: 2 per query, num_bv_tests
                                  and num_leaf_tests */
} fclx_mesh_shape_collide_batch_args;
int fclx_mesh_shape_collide_batch(fclx_context* ctx,
                                  const fclx_mesh_shape_collide_batch_args* args, void* stream);

/* Batched fcl::collide between primitive shapes with the GST_INDEP solver:
 * per pair what ShapeShapeCollide (collision_func_matrix-inl.h:81-112) and
 * ShapeCollisionTraversalNode::leafTesting
 * (shape_collision_traversal_node-inl.h:69-134, enable_cost off) give, via
 * GJKSolver_indep::shapeIntersect (gjk_solver_indep-inl.h:110-305).
 * Asynchronous; a pair naming a missing or malformed shape gets count 0 and
 * makes fclx_sync return FCLX_ERR_INVALID_ARGUMENT.  The context's sort and
 * EPA buffers serve one batch at a time: a batch on another stream waits
 * on the previous one. */
int fclx_shape_collide_batch(fclx_context* ctx, const fclx_shape_collide_batch_args* args,
                             void* stream);
/* Breakdown of the last FCLX_ENGINE_FAST collide batch with a cap above the
 * ordered limit (no counterpart in the reference, which only keeps
 * num_bv_tests/num_leaf_tests per traversal, bvh_collision_traversal_node.h:
 * 91-94).  Times come from HIP events on the caller's stream and are filled
 * only while profiling is enabled; the work counts always are. */
typedef struct {
  double setup_ms;
  double expand_ms;          /* all breadth-first levels */
  double leaves_ms;          /* leaf kernel over the breadth-first candidates */
  double dfs_ms;             /* persistent depth-first kernel */
  double fixup_ms;           /* truncation of queries that reached the cap */
  int64_t expand_bv_tests;   /* BV tests done by the breadth-first levels */
  int64_t expand_leaf_tests; /* triangle tests of the leaf kernel */
  int64_t dfs_bv_tests;      /* BV tests done by the depth-first kernel */
  int64_t dfs_leaf_tests;    /* triangle tests done by the depth-first kernel */
  int64_t fp64_fallbacks;    /* BV tests the FP32 cull left to FP64 */
  int64_t front_size;        /* tasks handed to the depth-first kernel */
  int32_t levels;            /* breadth-first levels run */
  int32_t flagged;           /* queries whose contact count reached the cap */
  /* depth-first load balance (profiling only): wavefronts that did work in
   * the first round, mean busy time over the span of every round, and the
   * times at which 50 % / 90 % / all wavefronts of the first round had
   * finished, from its first wavefront's start */
  int32_t dfs_waves;
  int32_t dfs_rounds;        /* depth-first rounds (work counts always) */
  double dfs_wave_busy;
  double dfs_end_p50_ms;
  double dfs_end_p90_ms;
  double dfs_end_max_ms;
  /* last BOOL-engine batch (profiling only): wavefront steps and the lanes
   * they used, per kind (0: FP32 BV tests, 1: FP64 triangle tests, 2: FP64
   * BV tests); queued pairs dropped because their query was already
   * answered; times a wavefront took new queries into free slots */
  int64_t bool_steps[3];
  int64_t bool_lanes[3];
  int64_t bool_skipped;
  int64_t bool_rounds;
  int64_t bool_stack_max; /* deepest query stack, entries */
  double bool_wave_ms; /* mean wavefront lifetime, ms */
  double bool_end_p50_ms; /* wavefront ends after the first start: median, */
  double bool_end_p90_ms; /* 90th percentile */
  double bool_span_ms;    /* and last, ms */
  int64_t dfs_inline_leaf_tests; /* of dfs_leaf_tests: those the depth-first kernel ran
                                  * itself (the rest ran in the deferred leaf kernel) */
  int64_t dfs_wave_steps;        /* BV-test steps of depth-first wavefronts (dfs_bv_tests /
                                  * dfs_wave_steps: lanes a step used, of 64) */
} fclx_engine_report;

int fclx_set_profiling(fclx_context* ctx, int32_t enable);
/* Tunables of the FAST engine's exhaustive path, per context (no counterpart
 * in the reference).  name / accepted range:
 *   "max_levels"        1..64   breadth-first levels before a depth-first round
 *   "front_target"      >= 64   stop expanding once the front holds this many tasks
 *                               (default 1 << 18)
 *   "dfs_blocks_per_cu" 1..32   cap on resident depth-first blocks per CU
 *   "defer_leaf"        >= 64   deferred leaf-pair buffer, entries
 *   "defer_uncertain"   >= 64   deferred uncertain-task buffer, entries
 *   "idle_stop"         0..1    share of idle wavefronts that ends a round (0: never)
 *   "stop_poll"         1..1024 depth-first steps between checks for a round's end
 *   "dfs_waves_per_simd" 4..6   register budget of the depth-first kernel
 *   "dfs_inline_leaves" 0, 1    1: at 4 wavefronts per SIMD the depth-first
 *                               kernel runs the triangle tests of its leaf
 *                               pairs itself and keeps the top of its stack
 *                               in LDS (default 1); 0: leaf pairs deferred
 *                               to a separate kernel, stack in device memory
 *   "dfs_stack_top"     128, 256, 512  stack entries per wavefront kept in
 *                               LDS with inline triangle tests (default 512)
 *   "bool_queries_per_wave" 0, 4, 8, 16, 32  queries a BOOL-engine wavefront
 *                               keeps in flight (default 8, measured best
 *                               under serving load; 0: 4 for batches below
 *                               65536 queries, else 8)
 *   "bool_waves_per_simd" 2..4  register budget of the BOOL-engine kernel
 *                               (32 queries per wavefront: at most 3)
 *   "bool_leaf_batch"   1..64   leaf pairs a BOOL-engine wavefront queues
 *                               before testing them
 *   "bool_width"        1..64   most tasks of one query a BOOL-engine step
 *                               tests (at least 64 / queries per wavefront;
 *                               default 16)
 *   "bool_refill"       1..32   free query slots that make a BOOL-engine
 *                               wavefront take new queries (default 4)
 *   "dist_queries_per_wave" 2, 4, 8  queries a FAST distance wavefront keeps
 *                               in flight (default 4)
 *   "dist_waves_per_simd" 2, 3  register budget of the FAST distance kernel
 *   "meshshape_waves_per_simd" 1..3  register budget of the mesh-vs-primitive
 *                               kernel (default 2)
 *   "meshshape_leaf_share" 0..100  percent of a wavefront's live queries
 *                               waiting at leaves that starts their leaf
 *                               tests (default 0: 50 for booleans, 75
 *                               with contact geometry)
 *   "dist_width"        1..64   most bounds of one query per distance step
 *                               (default 32)
 *   "dist_leaf_batch"   1..64   leaf pairs queued before their distances are
 *                               taken (default 16)
 *   "fast_persistent"   0, 1    1: FAST collide batches with caps above 64 run
 *                               as one persistent launch (the BOOL kernel's
 *                               exhaustive form: per-query slots fed from
 *                               global query counters, FP64 steps inline)
 *                               instead of depth-first rounds; same
 *                               outputs (default 0: rounds, faster here)
 *   "fast_graph"        0, 1    1: that launch sequence is captured once in a
 *                               HIP graph and replayed while its arguments
 *                               stay the same (default 1)
 *   "dist_fp32_cull"    0, 1    1: FAST distance takes certified FP32 lower
 *                               bounds of the RSS distances first (64-B
 *                               float nodes) and FP64 bounds only where both
 *                               children of a node stay; same outputs
 *                               (default 0: FP64 bounds throughout, faster
 *                               on MI355X)
 *   "epa_blocks_per_cu" 0..32   resident blocks per CU of the shape path's EPA
 *                               kernel (0: as many as fit); each lane holds
 *                               a 13 KiB EPA block in device memory
 *   "shape_waves_per_simd" 2..4 register budget of the shape path's GJK and
 *                               EPA kernels (default 2)
 * Results do not depend on them; smaller buffers mean more rounds. */
int fclx_set_engine_param(fclx_context* ctx, const char* name, double value);
/* The current value of a tunable above (for tuners); no reference
 * counterpart. */
int fclx_get_engine_param(fclx_context* ctx, const char* name, double* value);
int fclx_engine_last_report(fclx_context* ctx, fclx_engine_report* out);

/* Wait for the stream, then report (and clear) the device status: FCLX_OK,
 * FCLX_ERR_CAPACITY (pair records beyond pair_capacity were dropped) or
 * FCLX_ERR_STACK_OVERFLOW. */
int fclx_sync(fclx_context* ctx, void* stream);

/* One query with host buffers: the drop-in path behind fcl::collide for two
 * BVHModel<OBBRSS> (orientedMeshCollide, collision_func_matrix-inl.h:571-590).
 * cap = num_max_contacts minus contacts already in the caller's result.
 * *num_contacts receives the number of contacts (<= cap); the first
 * min(*num_contacts, pairs_capacity) (b1, b2) pairs, in the reference's
 * order, go to pairs.  Returns FCLX_ERR_CAPACITY when pairs was too small
 * (call again with room for *num_contacts pairs).
 * stats (optional) = {num_bv_tests, num_leaf_tests}. */
int fclx_collide(fclx_context* ctx, int32_t mesh1, const double tf1[12], int32_t mesh2,
                 const double tf2[12], int32_t cap, int32_t* pairs, int32_t pairs_capacity,
                 int32_t* num_contacts, int64_t* stats);
/* fclx_collide with enable_contact (mesh_collision_traversal_node-inl.h:
 * 578-598): up to cap contacts in the reference's order, each with its
 * (b1, b2) in pairs, world-frame point (3 doubles) in pos, normal (3) in
 * normal and penetration depth in depth; capacity counts contacts. */
int fclx_collide_contact(fclx_context* ctx, int32_t mesh1, const double tf1[12], int32_t mesh2,
                         const double tf2[12], int32_t cap, int32_t* pairs, double* pos,
                         double* normal, double* depth, int32_t capacity,
                         int32_t* num_contacts, int64_t* stats);
/* One primitive pair with host buffers: the drop-in path behind
 * fcl::collide for two shapes (ShapeShapeCollide,
 * collision_func_matrix-inl.h:278-311, with GJKSolver_indep).  verts1 /
 * verts2: a CONVEX shape's vertex list (3 doubles each; NULL otherwise; the
 * records' vertex_offset is not used).  cap = num_max_contacts minus the
 * contacts already in the caller's result.  *num_contacts receives the
 * contacts added (<= cap, at most 4); with enable_contact their world-frame
 * point, normal and depth go to pos (3 each), normal (3) and depth (1),
 * which need room for min(cap, 4) contacts.  stats (optional) = {GJK loop
 * passes, EPA loop passes}. */
int fclx_shape_collide(fclx_context* ctx, const fclx_shape* s1, const double* verts1,
                       const double tf1[12], const fclx_shape* s2, const double* verts2,
                       const double tf2[12], int32_t cap, int32_t enable_contact,
                       double gjk_tolerance, double* pos, double* normal, double* depth,
                       int32_t* num_contacts, int32_t* stats);
/* One query: the drop-in path behind fcl::distance for two BVHModel<RSS> or
 * BVHModel<OBBRSS> (orientedMeshDistance, distance_func_matrix-inl.h:386-404).
 * points (optional) = nearest points on mesh1 and mesh2 in the world frame;
 * ids (optional) = (b1, b2). */
int fclx_distance(fclx_context* ctx, int32_t mesh1, const double tf1[12], int32_t mesh2,
                  const double tf2[12], int32_t enable_nearest_points, double* min_distance,
                  double* points, int32_t* ids, int64_t* stats);

/* ------------------------------------------------------ BVTT front lists
 * collide() with a front list (detail::collide(node, front_list),
 * collision_node-inl.h:71-82; BVHFrontList, geometry/bvh/detail/
 * BVH_front.h:49-72; propagateBVHFrontListCollisionRecurse,
 * traversal_recurse-inl.h:463-521), for temporally coherent queries such as
 * the waypoints of a trajectory.  A front set holds, per sequence, the BVTT
 * nodes where the sequence's last query stopped (its front, up to
 * `capacity` (b1, b2) entries, 16 B of device memory per entry).  A batch
 * runs each sequence's queries in order on one lane: a query whose front is
 * empty traverses from the roots recording the front (no early stop: the
 * reference skips the stop rule whenever a front list is given), one whose
 * front holds entries propagates it.  Per query: counts (contacts, at most
 * num_max_contacts -- a pair reached by opening a front entry is tested
 * twice and may count twice, as in the reference), records (query, b1, b2)
 * in the reference's order per query, and the reference's num_bv_tests /
 * num_leaf_tests.  Only whether the count is nonzero is the same as a fresh
 * collide() (test_fcl_frontlist.cpp checks exactly this).  A front
 * belongs to one mesh pair: a query naming another pair than its
 * sequence's front starts from the roots.  A front that outgrows the
 * capacity is dropped (the next query starts from the roots) and fclx_sync
 * reports FCLX_ERR_CAPACITY.  enable_contact is not supported here. */
typedef struct fclx_front_set fclx_front_set;

int fclx_front_set_create(fclx_context* ctx, int32_t num_sequences, int32_t capacity,
                          fclx_front_set** out);
int fclx_front_set_destroy(fclx_front_set* fronts);
/* Empties every front (asynchronous on the stream). */
int fclx_front_set_clear(fclx_front_set* fronts, void* stream);
/* Host copy of one sequence's front: *n = its size, the first min(n, room)
 * entries (b1, b2) to pairs.  Synchronous. */
int fclx_front_set_read(fclx_front_set* fronts, int32_t sequence, int32_t* pairs, int32_t room,
                        int32_t* n);

typedef struct {
  int32_t num_sequences;        /* <= the set's num_sequences; sequence s uses front s */
  const int32_t* seq_offsets;   /* device: num_sequences + 1; sequence s is queries
                                   [seq_offsets[s], seq_offsets[s+1]) in order */
  const int32_t* query_meshes;  /* device: 2 mesh ids (OBBRSS) per query */
  const double* tf1;            /* device: 12 per query */
  const double* tf2;            /* device: 12 per query */
  int32_t num_max_contacts;     /* >= 1 */
  int32_t* counts;              /* device: per query */
  int32_t* pairs;               /* device: 3 int32 per record (query, b1, b2); may be NULL */
  int64_t pair_capacity;
  int64_t* num_pairs;           /* device: records produced */
  int64_t* bv_tests;            /* device, optional: per query */
  int64_t* leaf_tests;          /* device, optional: per query */
} fclx_collide_front_batch_args;

int fclx_collide_front_batch(fclx_context* ctx, const fclx_collide_front_batch_args* args,
                             fclx_front_set* fronts, void* stream);

/* -------------------------------------------------------------- broadphase
 * The candidate pairs of a broadphase manager (BroadPhaseCollisionManager,
 * broadphase/broadphase_collision_manager.h:77-122, whose collide() hands
 * every pair of objects with overlapping AABBs to its callback): all pairs
 * (i, j), i < j, of objects whose world AABBs overlap (AABB::overlap,
 * closed intervals, math/bv/AABB-inl.h), by sort and sweep on the device.
 * aabbs: device, 6 doubles per object (lower x, y, z, upper x, y, z);
 * groups: device, optional, a group id per object -- only pairs across
 * groups are reported (a query set against a scene, or two managers);
 * pairs: device, 2 int32 per pair, sorted ascending by (i, j);
 * *num_pairs (device) receives the number of overlapping pairs -- when it
 * exceeds pair_capacity (at most 2^31 - 1) only that many are stored (which
 * ones is unspecified) and the call is repeated with more room.
 * Asynchronous on the stream. */
int fclx_broadphase_pairs(fclx_context* ctx, int32_t num_objects, const double* aabbs,
                          const int32_t* groups, int32_t* pairs, int64_t pair_capacity,
                          int64_t* num_pairs, void* stream);

/* ---------------------------------------------------------------- serving
 * A request loop over one resident scene (no counterpart in the reference,
 * whose collide() is one synchronous call): planner requests -- a robot's
 * link meshes and a trajectory of link poses -- are queued as they arrive
 * and run as first-contact collide() queries (CollisionRequest(1, false),
 * the scene as the first object), each link against the scene at each
 * waypoint.  A worker thread forms a batch for a free stream slot from the
 * waiting requests -- the smallest first, the oldest first once one has
 * waited 1.5 ms -- that fit the batch size and the device-memory budget, so
 * requests that arrive while others are on the GPU join the next batch and
 * batches on different streams overlap on the device.  A request's result
 * does not depend on what it is batched with. */
typedef struct fclx_server fclx_server;

typedef struct {
  int32_t scene_mesh;          /* OBBRSS mesh id of the scene */
  int32_t num_streams;         /* batches in flight at once, 1..8 (6 measured best
                                  on the robot-trajectory workload) */
  double scene_pose[12];       /* the scene's pose */
  int64_t hbm_budget;          /* device bytes of requests in flight at once; a
                                  query takes 204 B (0: 2 GiB) */
  int32_t max_batch_queries;   /* queries per batch at most (0: 65536), or a
                                  stream's share of the waiting queries when
                                  that is more; a request larger than this
                                  runs alone */
  int32_t engine;              /* FCLX_ENGINE_BOOL (the natural choice: only
                                  booleans are returned), _FAST or _PLAIN */
} fclx_server_config;

typedef struct {
  int32_t num_links;
  int32_t num_waypoints;
  const int32_t* link_meshes;  /* host: a mesh id (OBBRSS) per link */
  const double* link_poses;    /* host: num_waypoints x num_links poses (12 each),
                                  waypoint-major, world frame; pinned memory
                                  (hipHostMalloc, 16-B aligned) is read by the
                                  device directly, pageable memory is packed
                                  into the loop's pinned staging first */
  uint8_t* result;             /* host: num_waypoints x num_links, written before
                                  the completion is reported: 1 = collides */
  int64_t tag;                 /* returned with the completion */
} fclx_request;
/* The request's arrays must stay valid until its completion is reported. */

typedef struct {
  int64_t tag;
  int32_t status;              /* FCLX_OK or an FCLX_ERR_* code */
  int32_t num_queries;
  double queue_ms;             /* submission to launch */
  double latency_ms;           /* submission to results in host memory */
} fclx_completion;

int fclx_server_create(fclx_context* ctx, const fclx_server_config* cfg, fclx_server** out);
/* Queues a request; returns at once. */
int fclx_server_submit(fclx_server* srv, const fclx_request* req);
/* Up to max completions; waits up to wait_ms (0: no wait) for the first. */
int fclx_server_poll(fclx_server* srv, fclx_completion* out, int32_t max, int32_t wait_ms,
                     int32_t* n);
/* Load replay for measurement: submits requests[i] at arrival_ms[i] after
 * the call (ascending), waits for all of them, and gives each request's
 * latency (submission to results), optionally its time in the queue
 * (submission to launch; queue_ms may be NULL) and the wall time of the
 * whole replay. */
int fclx_server_replay(fclx_server* srv, const fclx_request* requests,
                       const double* arrival_ms, int32_t n, double* latency_ms,
                       double* elapsed_ms, double* queue_ms);
/* Totals since creation: batches launched, queries run. */
int fclx_server_stats(fclx_server* srv, int64_t* batches, int64_t* queries);
/* Finishes every queued request, then stops the worker and frees it. */
int fclx_server_destroy(fclx_server* srv);

#ifdef __cplusplus
}
#endif

#endif
