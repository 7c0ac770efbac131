// Device-resident layout of an uploaded mesh.  Shared by the host upload code
// and the kernels; plain structs only.
//
// A mesh lives in HBM as structure-of-node-records:
//   obb64[n]  128 B: OBB axis (9 doubles, row-major, column j = box axis j),
//                    centre To (3), half extents (3), first_child link.
//   rss64[n]  128 B: RSS axis (9), origin To (3), rectangle l (2), radius r,
//                    first_child link.
//   size64[n]   8 B: BV::size() of the model's BV type (OBBRSS: squared norm
//                    of the OBB extents; RSS: sqrt(l0^2+l1^2)+2r), computed on
//                    the host so that no device sqrt enters a descent decision.
//   obb32[n]   64 B: the OBB rounded to float, for conservative culling.
//   rss32[n]   64 B: the RSS rounded to float, for conservative distance
//                    lower bounds.
//   verts      24 B each (x,y,z doubles); tris 12 B each (3 x int32).
//   tri_verts 128 B per triangle: its three vertices (9 doubles) followed by
//              padding, so a triangle test reads one line per triangle.
// Node records are 128-B aligned so that one lane's node fetch is one L2
// line; a wavefront of lanes on different nodes then issues 64 whole-line
// requests instead of 16x as many partial ones.
#pragma once
#include <stdint.h>

namespace fclx {

struct alignas(128) ObbNode64 {
  double axis[9];
  double to[3];
  double extent[3];
  int32_t first_child;  // >= 0: left child (right = left+1); < 0: -(tri+1)
  int32_t pad;
};
static_assert(sizeof(ObbNode64) == 128, "ObbNode64 must be one 128-B line");

struct alignas(128) RssNode64 {
  double axis[9];
  double to[3];
  double l[2];
  double r;
  int32_t first_child;
  int32_t pad;
};
static_assert(sizeof(RssNode64) == 128, "RssNode64 must be one 128-B line");

struct alignas(64) ObbNode32 {
  float axis[9];
  float to[3];
  float extent[3];
  int32_t first_child;
};
static_assert(sizeof(ObbNode32) == 64, "ObbNode32 must be 64 B");

struct alignas(64) RssNode32 {
  float axis[9];
  float to[3];
  float l[2];
  float r;
  int32_t first_child;
};
static_assert(sizeof(RssNode32) == 64, "RssNode32 must be 64 B");

// Per-mesh device descriptor (pointers into HBM).  Kernels receive an array
// of these and index it with the query's mesh ids.
constexpr int kTriVertsStride = 16;  // doubles per triangle in tri_verts

struct MeshDev {
  const ObbNode64* obb64;
  const RssNode64* rss64;
  const double* size64;
  const ObbNode32* obb32;
  const double* verts;  // 3 per vertex
  const int32_t* tris;  // 3 per triangle
  int32_t num_bvs;
  int32_t num_tris;
  int32_t num_verts;
  int32_t depth;        // levels from root to deepest leaf (root = 1)
  int32_t bv_type;      // 0 = OBBRSS, 1 = RSS
  int32_t pad;
  double cmax;          // max over nodes of |OBB centre|_1 (bounds the FP32 cull error)
  const int32_t* span;      // 2 per node: first_primitive, num_primitives
  const int32_t* prim_pos;  // per triangle: its position in the build's primitive order
  const double* tri_verts;  // kTriVertsStride per triangle: P1, P2, P3
  const RssNode32* rss32;
  double rcmax;         // max over nodes of |RSS origin|_1 (bounds the FP32 distance error)
};

}  // namespace fclx
