// Per-query set-up shared by the one-query-per-lane kernels and the
// distance engines: poses, the relative transform of a mesh pair, query
// validation and vertex access.
#pragma once

#include "../kernels.h"
#include "dist64.h"

namespace fclx {

__device__ __forceinline__ void load_pose(const double* tf, M3& R, V3& t) {
#pragma unroll
  for (int i = 0; i < 9; ++i) R.m[i] = tf[i];
  t = v3(tf[9], tf[10], tf[11]);
}

// R = R1^T R2, T = R1^T (t2 - t1)   (math/geometry-inl.h:653-682)
__device__ __forceinline__ void relative_pose(const double* tf1, const double* tf2, M3& R,
                                              V3& T) {
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j)
      R.m[3 * i + j] = (tf1[i] * tf2[j] + tf1[3 + i] * tf2[3 + j]) + tf1[6 + i] * tf2[6 + j];
  const V3 d = v3(tf2[9] - tf1[9], tf2[10] - tf1[10], tf2[11] - tf1[11]);
  T = v3(dot(col(tf1, 0), d), dot(col(tf1, 1), d), dot(col(tf1, 2), d));
}

// tf1^-1 * tf2 for isometries: (R1^T R2, R1^T t2 + (-(R1^T t1)))
// (mesh_distance_traversal_node-inl.h:632)
__device__ __forceinline__ void inverse_compose(const double* tf1, const double* tf2, M3& R,
                                                V3& T) {
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j)
      R.m[3 * i + j] = (tf1[i] * tf2[j] + tf1[3 + i] * tf2[3 + j]) + tf1[6 + i] * tf2[6 + j];
  const V3 t1 = v3(tf1[9], tf1[10], tf1[11]), t2 = v3(tf2[9], tf2[10], tf2[11]);
  const V3 a = v3(dot(col(tf1, 0), t2), dot(col(tf1, 1), t2), dot(col(tf1, 2), t2));
  const V3 b = v3(-dot(col(tf1, 0), t1), -dot(col(tf1, 1), t1), -dot(col(tf1, 2), t1));
  T = a + b;
}

// Both ids name live meshes of one BV type (OBBRSS when collision needs the
// OBB half); anything else has no cell in the reference's dispatch tables.
__device__ __forceinline__ bool query_ok(const MeshDev* meshes, int n, int i1, int i2,
                                         bool need_obbrss) {
  if (i1 < 0 || i2 < 0 || i1 >= n || i2 >= n) return false;
  const MeshDev& a = meshes[i1];
  const MeshDev& b = meshes[i2];
  if (a.num_bvs <= 0 || b.num_bvs <= 0 || a.bv_type != b.bv_type) return false;
  return !need_obbrss || a.bv_type == 0;
}

__device__ __forceinline__ V3 vertex(const MeshDev& m, int i) {
  const double* p = m.verts + 3 * static_cast<size_t>(i);
  return v3(p[0], p[1], p[2]);
}

}  // namespace fclx
