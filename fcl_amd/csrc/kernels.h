// Launch descriptors shared by the C ABI layer and the kernels.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "fclx.h"
#include "fclx_layout.h"

namespace fclx {

// bits of the context's device status word
enum : int32_t {
  kStatusStackOverflow = 1,
  kStatusPairOverflow = 2,
  kStatusQueueOverflow = 4,
  kStatusBadQuery = 8,  // a query named a missing mesh or an unsupported BV pair
  kStatusFrontOverflow = 16,  // a BVTT front outgrew its room
};

// Largest per-lane stack the traversal kernels are instantiated for; the
// host rejects meshes whose summed depths exceed it.
constexpr int kMaxStack = 256;

struct CollideLaunch {
  const MeshDev* meshes;       // device table indexed by mesh id
  int32_t num_meshes;
  int32_t num_queries;
  const int32_t* query_meshes; // 2 per query
  const double* tf1;           // 12 per query
  const double* tf2;
  int32_t cap;                 // num_max_contacts
  int32_t* counts;             // per query
  int32_t* pairs;              // 3 per record: query, b1, b2
  long long pair_capacity;
  long long* num_pairs;        // running record count (may exceed capacity)
  long long* bv_tests;         // optional, per query
  long long* leaf_tests;       // optional, per query
  int32_t* status;             // device status word (kStatus* bits)
  // enable_contact (plain kernel only): a record per contact, with its
  // world-frame point (3), normal (3) and penetration depth
  int32_t contact;
  double* contact_pos;
  double* contact_normal;
  double* contact_depth;
};

struct DistanceLaunch {
  const MeshDev* meshes;
  int32_t num_meshes;
  int32_t num_queries;
  const int32_t* query_meshes;
  const double* tf1;
  const double* tf2;
  double* distance;            // per query
  double* points;              // optional, 6 per query (world frame)
  int32_t* ids;                // optional, 2 per query
  long long* bv_tests;
  long long* leaf_tests;
  int32_t* status;
};

// Primitive-shape pairs (records in the fclx_shape layout of fclx.h)
struct ShapeLaunch {
  int32_t num_pairs;
  int32_t num_shapes;
  const fclx_shape* shapes;    // device, num_shapes records
  const double* vertices;      // device, convex vertex pool (3 per vertex)
  long long num_vertices;
  const int32_t* pair_shapes;  // 2 shape ids per pair
  const double* tf1;           // 12 per pair
  const double* tf2;
  int32_t cap;                 // num_max_contacts
  int32_t contact;             // enable_contact
  double tol;                  // GJK and EPA tolerance
  int32_t* counts;             // per pair
  double* contact_pos;         // 3 per slot (slot = pair * cap + i), with contact
  double* contact_normal;      // 3 per slot
  double* contact_depth;       // 1 per slot
  int32_t* passes;             // optional, 2 per pair: GJK and EPA loop passes
  int32_t* status;
};

// Mesh (BVHModel<OBBRSS>) against a primitive shape, one query per lane.
struct MeshShapeLaunch {
  const MeshDev* meshes;        // device table indexed by mesh id
  int32_t num_meshes;
  int32_t num_queries;
  const int32_t* query_meshes;  // per query: the mesh
  const int32_t* query_shapes;  // per query: the shape record
  int32_t num_shapes;
  const fclx_shape* shapes;     // device, num_shapes records
  const double* vertices;       // device, convex vertex pool (3 per vertex)
  long long num_vertices;
  const double* tf1;            // 12 per query: the mesh's pose
  const double* tf2;            // 12 per query: the shape's pose
  int32_t cap;                  // num_max_contacts
  int32_t contact;              // enable_contact
  double tol;                   // GJK and EPA tolerance
  int32_t* counts;              // per query
  int32_t* prims;               // optional, cap per query: the contacts' triangle (b1)
  double* contact_pos;          // 3 per slot (slot = query * cap + i), with contact
  double* contact_normal;       // 3 per slot
  double* contact_depth;        // 1 per slot
  long long* stats;             // optional, 2 per query: BV tests, leaf tests
  int32_t* status;
  int32_t leaf_share;           // percent of the live lanes waiting at leaves that starts a
                                // leaf round (1..100; 0: 50, or 100 with contact)
  int32_t max_depth;            // deepest mesh of the context (levels)
};
// waves: register budget, 1..3 waves per SIMD
hipError_t launch_mesh_shape_collide(const MeshShapeLaunch& L, int waves, hipStream_t s);
size_t mesh_shape_epa_lane_bytes();

hipError_t launch_collide_plain(const CollideLaunch& L, int stack_need, hipStream_t s);

// collide() with BVTT front lists: one lane per sequence of queries
struct FrontLaunch {
  const MeshDev* meshes;
  int32_t num_meshes;
  int32_t num_sequences;
  const int32_t* seq_offsets;  // num_sequences + 1: sequence s = queries [off[s], off[s+1])
  const int32_t* query_meshes; // 2 per query
  const double* tf1;
  const double* tf2;
  int32_t cap;                 // num_max_contacts
  int32_t* counts;
  int32_t* pairs;              // 3 per record: query, b1, b2
  long long pair_capacity;
  long long* num_pairs;
  long long* bv_tests;
  long long* leaf_tests;
  int2* fronts;                // front_cap entries per sequence
  int2* second;                // front_cap entries per sequence (scratch)
  int32_t* front_sizes;        // per sequence
  int32_t* front_meshes;       // 2 per sequence: the mesh pair of its front
  int32_t front_cap;
  int32_t* status;
};
hipError_t launch_collide_front(const FrontLaunch& L, int stack_need, hipStream_t s);

// Broadphase pairs of overlapping world AABBs (kernels_broadphase.hip):
// aabbs 6 doubles per object (lower x y z, upper x y z); groups optional
// (pairs only across groups); pairs 2 per pair, ascending (lower, higher);
// *num_pairs = all pairs found (those beyond capacity are not stored).
hipError_t launch_broadphase(int n, const double* aabbs, const int32_t* groups, int32_t* pairs,
                             long long capacity, long long* num_pairs, hipStream_t s);

// Device refit of one mesh (kernels_refit.hip)
struct RefitLaunch {
  const double* new_verts;     // 3 per vertex
  double* verts;               // the mesh's vertex array (replaced)
  double* tri_verts;           // the mesh's per-triangle vertex records (rebuilt)
  const int32_t* tris;
  const int32_t* span;         // 2 per node
  const int32_t* prim_pos;     // per triangle: position in the build order
  int32_t* order;              // scratch, per position: the triangle
  ObbNode64* obb64;
  RssNode64* rss64;
  double* size64;
  ObbNode32* obb32;
  RssNode32* rss32;
  double* cmax;                // device scalar: largest |OBB centre|_1
  double* rcmax;               // device scalar: largest |RSS origin|_1
  int32_t num_verts;
  int32_t num_tris;
  int32_t num_bvs;
  int32_t bv_type;
};
// bottomup: level_nodes holds the node ids level by level (root level 0),
// level_offsets_host[l] .. [l+1] (host array of levels + 1) bounding each.
hipError_t launch_refit(const RefitLaunch& L, bool bottomup, const int32_t* level_nodes,
                        const int32_t* level_offsets_host, int levels, hipStream_t s);
// Sort buffers, the EPA queue and per-lane EPA memory of the shape path; one
// per context, used by one batch at a time (the caller serialises calls;
// batches on different streams are ordered through an event).
struct ShapeWorkspace;
ShapeWorkspace* shape_workspace_create(int device);
void shape_workspace_destroy(ShapeWorkspace* w);
void shape_workspace_set_epa_blocks_per_cu(ShapeWorkspace* w, int blocks);
void shape_workspace_set_waves(ShapeWorkspace* w, int waves);
size_t shape_epa_lane_bytes();
hipError_t shape_collide_run(ShapeWorkspace* w, const ShapeLaunch& L, hipStream_t s);
hipError_t launch_distance_plain(const DistanceLaunch& L, int stack_need, hipStream_t s);
// FAST engine: wavefront-wide search with the reference's visit order kept
// as path keys (engine_dist.hip); BVTTs deeper than 62 levels go to the
// plain kernel
struct DistanceTuning {
  int queries_per_wave = 4;  // 2, 4, 8
  int waves_per_simd = 2;    // 2, 3
  int width = 32;            // most bounds of one query per step
  int leaf_batch = 16;       // leaf pairs queued before their distances are taken
  int fp32_cull = 0;         // 1: FP32 lower bounds first, FP64 where both children stay
};
hipError_t launch_distance_wave(const DistanceLaunch& L, int stack_need, const DistanceTuning& tu,
                                hipStream_t s);

}  // namespace fclx
