// Batched fcl::collide between a BVHModel<OBBRSS> and a primitive shape (Box,
// Sphere, Capsule, Cylinder, Convex) with FCL's GST_INDEP solver, in FP64:
//   orientedBVHShapeCollide (collision_func_matrix-inl.h:375-419, 500-521);
//   the shape's bounding box: computeBV<OBBRSS> by the generic fit over the
//   shape's bound vertices (geometry/shape/utility-inl.h:124-133, OBB fitn
//   bv/utility-inl.h:131-145), of which the BV test uses the OBB;
//   collisionRecurse (traversal_recurse-inl.h:82-129) over the mesh alone --
//   the shape is one leaf -- with BVTesting = OBB overlap of the shape's box
//   and the node's box posed by the mesh's transform
//   (mesh_shape_collision_traversal_node.h, OBBRSS node);
//   leaf test meshShapeCollisionOrientedNodeLeafTesting
//   (mesh_shape_collision_traversal_node-inl.h:205-283): sphereTriangleIntersect
//   (sphere_triangle-inl.h:138-230) for a sphere, GJK + EPA on the shape and
//   the posed TriangleP for the others (gjk_solver_indep-inl.h:448-505).
//
// One query per lane.  A lane fits the shape's box (Jacobi eigenvectors of
// the bound vertices' covariance), descends the mesh depth first with the
// reference's child order and stops at num_max_contacts.  Without
// enable_contact a GJK "inside" is the answer (EPA's status is never Failed
// after it).  With it, GJK still decides the hit and the record's slot
// (count, cap and order are GJK's), and EPA only supplies the slot's
// contact: the lane queues GJK's final simplex with the slot, and a second
// kernel runs EPA over the queue, each lane in a block of device memory of
// its own, taking the next queued pair when its own ends.  (When the queue
// for num_queries x num_max_contacts slots would not fit its budget, or
// for meshes deeper than the LDS stack, the descending lane runs EPA
// itself, in a grid sized to the EPA memory it holds.)
#include <hip/hip_runtime.h>

#include <algorithm>

#include "device/exact64.h"
#include "device/fit64.h"
#include "device/shape64.h"
#include "kernels.h"

namespace fclx {

namespace {

constexpr int kBlk = 256;
constexpr int kMsStack = 64;  // pending right children in LDS: meshes up to this depth
// deeper meshes (up to kMsDeepMax levels) keep the stack in device memory
constexpr int kMsDeepMax = 4096;
constexpr int kMsDeepBlocksPerCu = 2;
constexpr int kEpaBlocksPerCu = 2;
constexpr size_t kEpaQueueBudget = size_t(4) << 30;  // bytes of queued EPA seeds at most

// A hit whose contact EPA supplies: GJK's final simplex, the query, the
// triangle and the record slot the contact goes to.
struct MsEpaSeed {
  shp::D3 w[4], d[4];
  int32_t q, rank, prim, pad;
  long long slot;
};

using shp::D3;
using shp::mk;


__device__ __forceinline__ double pick3(const double* a, int i) {
  return i == 0 ? a[0] : (i == 1 ? a[1] : a[2]);
}

// axisFromEigen (math/geometry-inl.h:561-598): columns = the eigenvectors of
// the largest and middle eigenvalues and their cross product.
__device__ void axis_from_eigen(const double E[9], const double s[3], double axis[9]) {
  int mn, mid, mx;
  if (s[0] > s[1]) {
    mx = 0;
    mn = 1;
  } else {
    mn = 0;
    mx = 1;
  }
  if (s[2] < pick3(s, mn)) {
    mid = mn;
    mn = 2;
  } else if (s[2] > pick3(s, mx)) {
    mid = mx;
    mx = 2;
  } else {
    mid = 2;
  }
  D3 c0, c1;
#pragma unroll
  for (int r = 0; r < 3; ++r) {
    c0.v[r] = mx == 0 ? E[r] : (mx == 1 ? E[3 + r] : E[6 + r]);
    c1.v[r] = mid == 0 ? E[r] : (mid == 1 ? E[3 + r] : E[6 + r]);
  }
  const D3 c2 = shp::cross(c0, c1);
#pragma unroll
  for (int r = 0; r < 3; ++r) {
    axis[3 * r + 0] = c0.v[r];
    axis[3 * r + 1] = c1.v[r];
    axis[3 * r + 2] = c2.v[r];
  }
}

// getBoundVertices in the shape's frame (box-inl.h:122-138,
// sphere-inl.h:96-120, capsule-inl.h:108-162, cylinder-inl.h:102-126); the
// same points in the same order.
__device__ int bound_points(const shp::Shape& s, D3* p) {
  switch (s.type) {
    case shp::kBox: {
      const double a = s.dims[0] / 2, b = s.dims[1] / 2, c = s.dims[2] / 2;
      int k = 0;
#pragma unroll
      for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j)
#pragma unroll
          for (int l = 0; l < 2; ++l) p[k++] = mk(i ? -a : a, j ? -b : b, l ? -c : c);
      return 8;
    }
    case shp::kSphere: {
      const double m = (1 + sqrt(5.0)) / 2.0;
      const double e = s.dims[0] * 6 / (sqrt(27.0) + sqrt(15.0));
      const double a = e, b = m * e;
      p[0] = mk(0, a, b);
      p[1] = mk(0, -a, b);
      p[2] = mk(0, a, -b);
      p[3] = mk(0, -a, -b);
      p[4] = mk(a, b, 0);
      p[5] = mk(-a, b, 0);
      p[6] = mk(a, -b, 0);
      p[7] = mk(-a, -b, 0);
      p[8] = mk(b, 0, a);
      p[9] = mk(b, 0, -a);
      p[10] = mk(-b, 0, a);
      p[11] = mk(-b, 0, -a);
      return 12;
    }
    case shp::kCapsule: {
      const double r = s.dims[0], hl = s.dims[1] * 0.5;
      const double m = (1 + sqrt(5.0)) / 2.0;
      const double e = r * 6 / (sqrt(27.0) + sqrt(15.0));
      const double a = e, b = m * e, r2 = r * 2 / sqrt(3.0);
#pragma unroll
      for (int h = 0; h < 2; ++h) {
        // b - hl written as b + (-hl): the same double
        const double z = h ? -hl : hl;
        D3* q = p + 12 * h;
        q[0] = mk(0, a, b + z);
        q[1] = mk(0, -a, b + z);
        q[2] = mk(0, a, -b + z);
        q[3] = mk(0, -a, -b + z);
        q[4] = mk(a, b, z);
        q[5] = mk(-a, b, z);
        q[6] = mk(a, -b, z);
        q[7] = mk(-a, -b, z);
        q[8] = mk(b, 0, a + z);
        q[9] = mk(b, 0, -a + z);
        q[10] = mk(-b, 0, a + z);
        q[11] = mk(-b, 0, -a + z);
      }
      const double c = 0.5 * r2, d = r;
#pragma unroll
      for (int h = 0; h < 2; ++h) {
        const double z = h ? -hl : hl;
        D3* q = p + 24 + 6 * h;
        q[0] = mk(r2, 0, z);
        q[1] = mk(c, d, z);
        q[2] = mk(-c, d, z);
        q[3] = mk(-r2, 0, z);
        q[4] = mk(-c, -d, z);
        q[5] = mk(c, -d, z);
      }
      return 36;
    }
    case shp::kCylinder: {
      const double r = s.dims[0], hl = s.dims[1] * 0.5;
      const double r2 = r * 2 / sqrt(3.0), a = 0.5 * r2, b = r;
#pragma unroll
      for (int h = 0; h < 2; ++h) {
        const double z = h ? hl : -hl;
        D3* q = p + 6 * h;
        q[0] = mk(r2, 0, z);
        q[1] = mk(a, b, z);
        q[2] = mk(-a, b, z);
        q[3] = mk(-r2, 0, z);
        q[4] = mk(-a, -b, z);
        q[5] = mk(a, -b, z);
      }
      return 12;
    }
  }
  return 0;
}

// The shape's box in the world frame: OBB fitn over the posed bound
// vertices (a convex's own vertices): point covariance
// (geometry-inl.h:1335-1421), eigen_old, axisFromEigen,
// getExtentAndCenter_pointcloud (geometry-inl.h:228-287).
__device__ void shape_obb(const shp::Shape& s, const shp::Xf& T, double axis[9], double to[3],
                          double extent[3]) {
  D3 local[36];
  const bool convex = s.type == shp::kConvex;
  const int n = convex ? s.nv : bound_points(s, local);
  double S1[3] = {0, 0, 0};
  double s00 = 0, s11 = 0, s22 = 0, s01 = 0, s02 = 0, s12 = 0;
  for (int i = 0; i < n; ++i) {
    const D3 q = convex ? mk(s.verts[3 * i], s.verts[3 * i + 1], s.verts[3 * i + 2]) : local[i];
    const D3 p = shp::apply(T, q);
#pragma unroll
    for (int k = 0; k < 3; ++k) S1[k] += p.v[k];
    s00 += p.v[0] * p.v[0];
    s11 += p.v[1] * p.v[1];
    s22 += p.v[2] * p.v[2];
    s01 += p.v[0] * p.v[1];
    s02 += p.v[0] * p.v[2];
    s12 += p.v[1] * p.v[2];
  }
  const double np = static_cast<double>(n);
  double M[9], E[9], ev[3];
  M[0] = s00 - S1[0] * S1[0] / np;
  M[4] = s11 - S1[1] * S1[1] / np;
  M[8] = s22 - S1[2] * S1[2] / np;
  M[1] = s01 - S1[0] * S1[1] / np;
  M[5] = s12 - S1[1] * S1[2] / np;
  M[2] = s02 - S1[0] * S1[2] / np;
  M[3] = M[1];
  M[6] = M[2];
  M[7] = M[5];
  eigen_old(M, ev, E);
  axis_from_eigen(E, ev, axis);
  double mn[3] = {1.7976931348623157e308, 1.7976931348623157e308, 1.7976931348623157e308};
  double mx[3] = {-1.7976931348623157e308, -1.7976931348623157e308, -1.7976931348623157e308};
  for (int i = 0; i < n; ++i) {
    const D3 q = convex ? mk(s.verts[3 * i], s.verts[3 * i + 1], s.verts[3 * i + 2]) : local[i];
    const D3 p = shp::apply(T, q);
#pragma unroll
    for (int k = 0; k < 3; ++k) {
      const double pr = shp::dot(shp::col(axis, k), p);
      if (pr > mx[k]) mx[k] = pr;
      if (pr < mn[k]) mn[k] = pr;
    }
  }
  const D3 o = mk((mx[0] + mn[0]) / 2, (mx[1] + mn[1]) / 2, (mx[2] + mn[2]) / 2);
  const D3 c = shp::rot(axis, o);
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    to[k] = c.v[k];
    extent[k] = (mx[k] - mn[k]) * 0.5;
  }
}

// segmentSqrDistance (sphere_triangle-inl.h:84-108)
__device__ __forceinline__ double segment_sqr_distance(const D3& from, const D3& to, const D3& p,
                                                       D3& nearest) {
  D3 diff = shp::sub(p, from);
  const D3 v = shp::sub(to, from);
  double t = shp::dot(v, diff);
  if (t > 0) {
    const double dotVV = shp::dot(v, v);
    if (t < dotVV) {
      t /= dotVV;
      diff = shp::sub(diff, shp::mul(v, t));
    } else {
      t = 1;
      diff = shp::sub(diff, v);
    }
  } else {
    t = 0;
  }
  nearest = shp::add(from, shp::mul(v, t));
  return shp::dot(diff, diff);
}

// projectInTriangle (sphere_triangle-inl.h:112-135)
__device__ __forceinline__ bool project_in_triangle(const D3& p1, const D3& p2, const D3& p3,
                                                    const D3& n, const D3& p) {
  const double r1 = shp::dot(shp::cross(shp::sub(p2, p1), n), shp::sub(p, p1));
  const double r2 = shp::dot(shp::cross(shp::sub(p3, p2), n), shp::sub(p, p2));
  const double r3 = shp::dot(shp::cross(shp::sub(p1, p3), n), shp::sub(p, p3));
  return (r1 > 0 && r2 > 0 && r3 > 0) || (r1 <= 0 && r2 <= 0 && r3 <= 0);
}

// sphereTriangleIntersect (sphere_triangle-inl.h:138-230): the solver's
// outputs (normal from the centre towards the contact, depth <= 0).
__device__ bool sphere_triangle(double radius, const D3& center, const D3& P1, const D3& P2,
                                const D3& P3, D3& point, double& depth, D3& normal_out) {
  D3 normal = shp::normalized(shp::cross(shp::sub(P2, P1), shp::sub(P3, P1)));
  const double rt = radius + 2.220446049250313e-16;
  double dist = shp::dot(shp::sub(center, P1), normal);
  if (dist < 0) {
    dist *= -1;
    normal = shp::mul(normal, -1.0);
  }
  if (!(dist < rt)) return false;
  bool has = false;
  D3 cp = mk(0, 0, 0);
  if (project_in_triangle(P1, P2, P3, normal, center)) {
    has = true;
    cp = shp::sub(center, shp::mul(normal, dist));
  } else {
    const double r2 = rt * rt;
    D3 ne;
    if (segment_sqr_distance(P1, P2, center, ne) < r2) {
      has = true;
      cp = ne;
    }
    if (segment_sqr_distance(P2, P3, center, ne) < r2) {
      has = true;
      cp = ne;
    }
    if (segment_sqr_distance(P3, P1, center, ne) < r2) {
      has = true;
      cp = ne;
    }
  }
  if (!has) return false;
  const D3 cc = shp::sub(cp, center);
  const double dsq = shp::sqn(cc);
  if (!(dsq < rt * rt)) return false;
  if (dsq > 0) {
    normal_out = shp::normalized(cc);
    depth = -(radius - sqrt(dsq));
  } else {
    normal_out = shp::neg(normal);
    depth = -radius;
  }
  point = cp;
  return true;
}

// A usable query: a mesh of OBBRSS nodes no deeper than the stack (depth_max
// levels), and a shape of a known type (a convex with at least 4 vertices
// inside the pool).
__device__ __forceinline__ bool load_query(const MeshShapeLaunch& L, int q, shp::Shape& s,
                                           int depth_max) {
  const int mi = L.query_meshes[q], si = L.query_shapes[q];
  if (mi < 0 || mi >= L.num_meshes || si < 0 || si >= L.num_shapes) return false;
  const MeshDev& M = L.meshes[mi];
  if (M.num_bvs <= 0 || M.bv_type != 0 || M.depth > depth_max) return false;
  const fclx_shape& r = L.shapes[si];
  s.type = r.type;
  s.nv = r.num_vertices;
  s.dims[0] = r.dims[0];
  s.dims[1] = r.dims[1];
  s.dims[2] = r.dims[2];
  s.verts = nullptr;
  switch (r.type) {
    case shp::kBox:
    case shp::kSphere:
    case shp::kCapsule:
    case shp::kCylinder:
      return true;
    case shp::kConvex:
      if (r.num_vertices < 4 || r.vertex_offset < 0 ||
          r.vertex_offset + r.num_vertices > L.num_vertices || L.vertices == nullptr)
        return false;
      s.verts = L.vertices + 3 * r.vertex_offset;
      return true;
  }
  return false;
}

// Queries grouped by shape kind, so that a wavefront's lanes run the same
// leaf test (the order of queries does not change any query's answer).
// Keys: the five kinds, then 5 for a query whose shape id is out of range
// (answered as a bad query by the main kernel).  ctr[0..5]: counts,
// ctr[8..13]: placement cursors.
constexpr int kKinds = 6;

__device__ __forceinline__ int kind_of(const MeshShapeLaunch& L, int q) {
  const int si = L.query_shapes[q];
  if (si < 0 || si >= L.num_shapes) return kKinds - 1;
  const int t = L.shapes[si].type;
  return t >= 0 && t < kKinds - 1 ? t : kKinds - 1;
}

__global__ void ms_zero(int32_t* ctr) {
  if (threadIdx.x < 16) ctr[threadIdx.x] = 0;
}

__global__ void __launch_bounds__(kBlk) ms_count(MeshShapeLaunch L, int32_t* ctr) {
  __shared__ int cnt[kKinds];
  if (threadIdx.x < kKinds) cnt[threadIdx.x] = 0;
  __syncthreads();
  const int i = blockIdx.x * kBlk + threadIdx.x;
  if (i < L.num_queries) atomicAdd(&cnt[kind_of(L, i)], 1);
  __syncthreads();
  if (threadIdx.x < kKinds && cnt[threadIdx.x]) atomicAdd(&ctr[threadIdx.x], cnt[threadIdx.x]);
}

__global__ void __launch_bounds__(kBlk) ms_place(MeshShapeLaunch L, int32_t* ctr,
                                                 int32_t* order) {
  __shared__ int cnt[kKinds];
  __shared__ int base[kKinds];
  if (threadIdx.x < kKinds) cnt[threadIdx.x] = 0;
  __syncthreads();
  const int i = blockIdx.x * kBlk + threadIdx.x;
  int key = 0, off = 0;
  if (i < L.num_queries) {
    key = kind_of(L, i);
    off = atomicAdd(&cnt[key], 1);
  }
  __syncthreads();
  if (threadIdx.x < kKinds) {
    int start = 0;
    for (int k = 0; k < static_cast<int>(threadIdx.x); ++k) start += ctr[k];
    base[threadIdx.x] =
        start + (cnt[threadIdx.x] ? atomicAdd(&ctr[8 + threadIdx.x], cnt[threadIdx.x]) : 0);
  }
  __syncthreads();
  if (i < L.num_queries) order[base[key] + off] = i;
}

// The leaf test of one (shape, triangle) pair, kept out of line so that the
// descent loop's registers are not sized for GJK and EPA.
template <bool kContact>
__device__ __noinline__ bool leaf_test(const shp::Shape& s, const shp::Xf& T1, const shp::Xf& T2,
                                       const double* tv, double tol, shp::Epa& E,
                                       shp::Contact& c) {
  bool hit;
  if (s.type == shp::kSphere) {
    // gjk_solver_indep-inl.h:527-545: the triangle posed by the mesh's tf
    D3 P[3];
#pragma unroll
    for (int k = 0; k < 3; ++k) P[k] = shp::apply(T1, mk(tv[3 * k], tv[3 * k + 1], tv[3 * k + 2]));
    D3 p, n;
    double d;
    hit = sphere_triangle(s.dims[0], T2.t, P[0], P[1], P[2], p, d, n);
    if (hit) {
      c.pos = p;
      c.normal = shp::neg(n);  // the leaf stores minus the solver's normal
      c.depth = d;
    }
  } else {
    shp::Shape tri;
    tri.type = shp::kTriangle;
    tri.nv = 3;
    tri.dims[0] = tri.dims[1] = tri.dims[2] = 0;
    tri.verts = tv;
    shp::MDiff m;
    shp::make_mdiff(s, T2, tri, T1, m);
    shp::Gjk g;
    hit = shp::gjk_evaluate(g, m, tol) == shp::kGjkInside;
    if constexpr (kContact) {
      if (hit) {
        // the leaf's -normal undoes the solver's: EPA's normal stays
        bool more = shp::epa_start(E, g, m, tol);
        while (more) more = shp::epa_step(E, m, tol);
        c = shp::epa_finish(E, m, T2);
      }
    }
  }
  return hit;
}

// The queued form of leaf_test<true>: a sphere's contact from the closed
// form; for the other shapes GJK alone, its final simplex in g and
// need_epa set on a hit.
__device__ __noinline__ bool leaf_test_gjk(const shp::Shape& s, const shp::Xf& T1,
                                           const shp::Xf& T2, const double* tv, double tol,
                                           shp::Contact& c, shp::Gjk& g, bool& need_epa) {
  need_epa = false;
  if (s.type == shp::kSphere) {
    D3 P[3];
#pragma unroll
    for (int k = 0; k < 3; ++k) P[k] = shp::apply(T1, mk(tv[3 * k], tv[3 * k + 1], tv[3 * k + 2]));
    D3 p, n;
    double d;
    const bool hit = sphere_triangle(s.dims[0], T2.t, P[0], P[1], P[2], p, d, n);
    if (hit) {
      c.pos = p;
      c.normal = shp::neg(n);
      c.depth = d;
    }
    return hit;
  }
  shp::Shape tri;
  tri.type = shp::kTriangle;
  tri.nv = 3;
  tri.dims[0] = tri.dims[1] = tri.dims[2] = 0;
  tri.verts = tv;
  shp::MDiff m;
  shp::make_mdiff(s, T2, tri, T1, m);
  const bool hit = shp::gjk_evaluate(g, m, tol) == shp::kGjkInside;
  need_epa = hit;
  return hit;
}

// Lanes of a wavefront step together: a lane whose descent reached a leaf
// that passed its BV test waits there, the others keep descending, and the
// waiting lanes test their leaves together once they are a set share of the
// live lanes (or nobody can descend).  The costly leaf test (GJK) then
// runs once for many lanes instead of once per descent step.  Each lane's
// own order of tests is the reference's, so counts, triangles and test
// statistics are unchanged.
//
// kDeep: the pending right children live in device memory (deep_stack, one
// column of L.max_depth entries per lane, lane-interleaved) instead of LDS,
// for contexts holding a mesh deeper than kMsStack levels.
//
// kQueue (with kContact): hits that need EPA are queued (seeds, one
// reservation per wavefront on nseed) instead of running EPA here.
template <bool kContact, int kWaves, bool kDeep, bool kQueue = false>
__global__ void __launch_bounds__(kBlk, kWaves) meshshape_collide(MeshShapeLaunch L,
                                                                  const int32_t* order,
                                                                  shp::EpaLane* lanes,
                                                                  shp::EpaWave* waves,
                                                                  int* deep_stack,
                                                                  MsEpaSeed* seeds = nullptr,
                                                                  unsigned long long* nseed =
                                                                      nullptr) {
  const size_t gid = static_cast<size_t>(blockIdx.x) * kBlk + threadIdx.x;
  const size_t stride = static_cast<size_t>(gridDim.x) * kBlk;
  // the horizon walk's stack in LDS, one column per thread
  __shared__ int16_t xs_lds[(kContact && !kQueue) ? shp::kNfEpa : 1][kBlk];
  shp::Epa E;
  if constexpr (kContact && !kQueue) {
    E.M = lanes + gid;
    E.S = waves + gid / 64;
    E.lane = threadIdx.x & 63;
    E.xs = &xs_lds[0][threadIdx.x];
    E.xs_stride = kBlk;
  }
  size_t qq = gid;  // position in the (sorted) order
  size_t q = 0;     // the query at that position
  shp::Shape s;
  shp::Xf T1, T2;
  double sa[9], sto[3], sext[3];
  const ObbNode64* nodes = nullptr;
  const double* tris = nullptr;
  // pending right children in LDS, one column per lane (lane-interleaved:
  // conflict-free whatever depth each lane is at); a deep stack holds at
  // most L.max_depth - 1 entries
  __shared__ int stk[kDeep ? 1 : kMsStack][kBlk];
  const int col = threadIdx.x;
  const int depth_max = kDeep ? L.max_depth : kMsStack;
  int sp = 0, b = 0, count = 0, pend = -1;
  long long nbv = 0, nleaf = 0;
  bool live = false;  // a query in progress
  auto push = [&](int v) {
    if constexpr (kDeep)
      deep_stack[static_cast<size_t>(sp) * stride + gid] = v;
    else
      stk[sp][col] = v;
    ++sp;
  };
  auto pop = [&]() -> int {
    --sp;
    if constexpr (kDeep)
      return deep_stack[static_cast<size_t>(sp) * stride + gid];
    else
      return stk[sp][col];
  };
  // take queries until one is usable (the others are answered at once)
  auto start = [&]() {
    while (qq < static_cast<size_t>(L.num_queries)) {
      q = order ? static_cast<size_t>(order[qq]) : qq;
      if (load_query(L, static_cast<int>(q), s, depth_max)) {
        const MeshDev& M = L.meshes[L.query_meshes[q]];
        nodes = M.obb64;
        tris = M.tri_verts;
        T1 = shp::xf_load(L.tf1 + 12 * q);
        T2 = shp::xf_load(L.tf2 + 12 * q);
        shape_obb(s, T2, sa, sto, sext);
        sp = 0;
        b = 0;
        count = 0;
        pend = -1;
        nbv = 0;
        nleaf = 0;
        live = true;
        return;
      }
      L.counts[q] = 0;
      if (L.stats) {
        L.stats[2 * q] = 0;
        L.stats[2 * q + 1] = 0;
      }
      atomicOr(L.status, kStatusBadQuery);
      qq += stride;
    }
    live = false;
  };
  auto finish = [&]() {
    L.counts[q] = count;
    if (L.stats) {
      L.stats[2 * q] = nbv;
      L.stats[2 * q + 1] = nleaf;
    }
    qq += stride;
    start();
  };
  auto next = [&]() {  // the next node of the descent, or the query's end
    if (sp == 0)
      finish();
    else
      b = pop();
  };
  start();
  while (true) {
    const bool walk = live && pend < 0;
    const unsigned long long mw = __ballot(walk);
    const unsigned long long ml = __ballot(live && pend >= 0);
    if (!mw && !ml) break;
    if (ml && (!mw || 100 * __popcll(ml) >= L.leaf_share * __popcll(mw | ml))) {
      if (pend < 0) continue;
      const int prim = pend;
      pend = -1;
      ++nleaf;
      shp::Contact c;
      bool hit, need_epa = false;
      shp::Gjk g;
      const double* tv = tris + static_cast<size_t>(prim) * kTriVertsStride;
      if constexpr (kQueue)
        hit = leaf_test_gjk(s, T1, T2, tv, L.tol, c, g, need_epa);
      else
        hit = leaf_test<kContact>(s, T1, T2, tv, L.tol, E, c);
      if constexpr (kQueue) {
        // the seeds of this leaf round: one reservation for the wavefront
        const unsigned long long em = __ballot(need_epa);
        if (em != 0) {
          const int leader = __ffsll(static_cast<long long>(em)) - 1;
          unsigned long long b = 0;
          if (static_cast<int>(threadIdx.x & 63) == leader)
            b = atomicAdd(nseed, static_cast<unsigned long long>(__popcll(em)));
          b = __shfl(b, leader);
          if (need_epa) {
            MsEpaSeed& sd =
                seeds[b + __popcll(em & ((1ULL << (threadIdx.x & 63)) - 1ULL))];
#pragma unroll
            for (int j = 0; j < 4; ++j) {
              sd.w[j] = g.w[j];
              sd.d[j] = g.d[j];
            }
            sd.q = static_cast<int32_t>(q);
            sd.rank = g.rank;
            sd.prim = prim;
            sd.slot = static_cast<long long>(q * L.cap + count);
          }
        }
      }
      if (hit) {
        const size_t k = q * L.cap + count;
        if (L.prims) L.prims[k] = prim;
        if constexpr (kContact) {
          if (!need_epa) {
#pragma unroll
            for (int j = 0; j < 3; ++j) {
              L.contact_pos[3 * k + j] = c.pos.v[j];
              L.contact_normal[3 * k + j] = c.normal.v[j];
            }
            L.contact_depth[k] = c.depth;
          }
        }
        if (++count >= L.cap) {
          finish();
          continue;
        }
      }
      next();
    } else if (walk) {
      const ObbNode64& nd = nodes[b];
      ++nbv;
      M3 R0;
#pragma unroll
      for (int i = 0; i < 9; ++i) R0.m[i] = T1.R[i];
      double B[9], Tr[3];
      box_relative(R0, v3(T1.t.v[0], T1.t.v[1], T1.t.v[2]), sa, sto, nd.axis, nd.to, B, Tr);
      if (obb_disjoint64(B, Tr, sext, nd.extent)) {
        next();
      } else if (nd.first_child >= 0) {
        push(nd.first_child + 1);
        b = nd.first_child;
      } else {
        pend = -(nd.first_child + 1);  // tested in the next leaf round
      }
    }
  }
}

// EPA over the queued hits, each lane on its own pair one EPA loop pass per
// trip, taking the wavefront's next entry when its own ends (the entries
// come 64 at a time per wavefront from one counter), so that lanes do not
// wait for the slowest pair of their wavefront.  The contact goes to the
// slot the descent reserved, as leaf_test<true> writes it.
template <int kWaves>
__global__ void __launch_bounds__(kBlk, kWaves) meshshape_epa(MeshShapeLaunch L,
                                                              const MsEpaSeed* seeds,
                                                              const unsigned long long* nseed,
                                                              unsigned long long* grab,
                                                              shp::EpaLane* lanes,
                                                              shp::EpaWave* waves) {
  constexpr int kRefillAt = 16;
  const int lane = threadIdx.x & 63;
  const size_t gid = static_cast<size_t>(blockIdx.x) * kBlk + threadIdx.x;
  const long long total = static_cast<long long>(*nseed);
  long long cursor = 0, stop = 0;  // the wavefront's entries not yet started
  bool exhausted = total == 0;
  __shared__ int16_t xs_lds[shp::kNfEpa][kBlk];  // the horizon walk's stack
  shp::Epa E;
  E.M = lanes + gid;
  E.S = waves + gid / 64;
  E.lane = lane;
  E.xs = &xs_lds[0][threadIdx.x];
  E.xs_stride = kBlk;
  shp::MDiff m;
  shp::Xf T2;
  long long slot = 0;
  bool active = false;
  while (true) {
    bool finish = false, started = false;
    const unsigned long long idle = __ballot(!active);
    const int nidle = __popcll(idle);
    if (cursor >= stop && !exhausted && (nidle == 64 || nidle >= kRefillAt)) {
      unsigned long long base = 0;
      if (lane == 0) base = atomicAdd(grab, 64ULL);
      base = __shfl(base, 0);
      const long long b = static_cast<long long>(base);
      cursor = b < total ? b : total;
      stop = b + 64 < total ? b + 64 : total;
      exhausted = b + 64 >= total;
    }
    if (cursor < stop && (nidle == 64 || nidle >= kRefillAt)) {
      const long long v = cursor + __popcll(idle & ((1ULL << lane) - 1ULL));
      cursor = cursor + nidle < stop ? cursor + nidle : stop;
      if (!active && v < stop) {
        const MsEpaSeed& sd = seeds[v];
        const int q = sd.q;
        shp::Shape s;
        (void)load_query(L, q, s, kMsDeepMax);  // checked by the descent
        const MeshDev& M = L.meshes[L.query_meshes[q]];
        shp::Shape tri;
        tri.type = shp::kTriangle;
        tri.nv = 3;
        tri.dims[0] = tri.dims[1] = tri.dims[2] = 0;
        tri.verts = M.tri_verts + static_cast<size_t>(sd.prim) * kTriVertsStride;
        const shp::Xf T1 = shp::xf_load(L.tf1 + 12 * static_cast<size_t>(q));
        T2 = shp::xf_load(L.tf2 + 12 * static_cast<size_t>(q));
        shp::make_mdiff(s, T2, tri, T1, m);
        shp::Gjk g;
#pragma unroll
        for (int j = 0; j < 4; ++j) {
          g.w[j] = sd.w[j];
          g.d[j] = sd.d[j];
        }
        g.rank = sd.rank;
        slot = sd.slot;
        active = shp::epa_start(E, g, m, L.tol);
        finish = !active;
        started = true;
      }
    }
    if (active && !started) {
      active = shp::epa_step(E, m, L.tol);
      finish = !active;
    }
    if (finish) {
      // the leaf's -normal undoes the solver's: EPA's normal stays
      const shp::Contact c = shp::epa_finish(E, m, T2);
#pragma unroll
      for (int j = 0; j < 3; ++j) {
        L.contact_pos[3 * slot + j] = c.pos.v[j];
        L.contact_normal[3 * slot + j] = c.normal.v[j];
      }
      L.contact_depth[slot] = c.depth;
    }
    if (exhausted && cursor >= stop && !__ballot(active)) break;
  }
}

}  // namespace

size_t mesh_shape_epa_lane_bytes() { return sizeof(shp::EpaLane) + sizeof(shp::EpaWave) / 64; }

// the queued-EPA descent for a register budget of 1, 2 or 3 waves per SIMD
void launch_queued(int waves, unsigned grid, hipStream_t s, const MeshShapeLaunch& L,
                   const int32_t* order, MsEpaSeed* seeds, unsigned long long* nseed) {
  switch (waves) {
    case 1:
      meshshape_collide<true, 1, false, true>
          <<<grid, kBlk, 0, s>>>(L, order, nullptr, nullptr, nullptr, seeds, nseed);
      break;
    case 3:
      meshshape_collide<true, 3, false, true>
          <<<grid, kBlk, 0, s>>>(L, order, nullptr, nullptr, nullptr, seeds, nseed);
      break;
    default:
      meshshape_collide<true, 2, false, true>
          <<<grid, kBlk, 0, s>>>(L, order, nullptr, nullptr, nullptr, seeds, nseed);
  }
}

__global__ void ms_zero_u64(unsigned long long* p) {
  if (threadIdx.x < 2) p[threadIdx.x] = 0;
}

template <bool kContact>
void launch_variant(int waves, unsigned grid, hipStream_t s, const MeshShapeLaunch& L,
                    const int32_t* order, shp::EpaLane* lanes, shp::EpaWave* wv,
                    int* deep_stack) {
  if (deep_stack) {
    meshshape_collide<kContact, 1, true><<<grid, kBlk, 0, s>>>(L, order, lanes, wv, deep_stack);
    return;
  }
  switch (waves) {
    case 1:
      meshshape_collide<kContact, 1, false><<<grid, kBlk, 0, s>>>(L, order, lanes, wv, nullptr);
      break;
    case 3:
      meshshape_collide<kContact, 3, false><<<grid, kBlk, 0, s>>>(L, order, lanes, wv, nullptr);
      break;
    default:
      meshshape_collide<kContact, 2, false><<<grid, kBlk, 0, s>>>(L, order, lanes, wv, nullptr);
  }
}

// Batches at least this large are grouped by shape kind first.
constexpr int kSortMin = 4096;

hipError_t launch_mesh_shape_collide(const MeshShapeLaunch& in, int waves, hipStream_t s) {
  if (in.num_queries <= 0) return hipSuccess;
  MeshShapeLaunch L = in;
  // measured: half the live lanes for booleans, three quarters with contact
  // geometry (whose leaf rounds queue EPA's work for a pass of its own)
  if (L.leaf_share <= 0) L.leaf_share = L.contact ? 75 : 50;
  const long long blocks = (static_cast<long long>(L.num_queries) + kBlk - 1) / kBlk;
  // the order of queries grouped by shape kind (stream-ordered memory)
  int32_t* order = nullptr;
  hipError_t e = hipSuccess;
  if (L.num_queries >= kSortMin) {
    e = hipMallocAsync(reinterpret_cast<void**>(&order),
                       (static_cast<size_t>(L.num_queries) + 16) * sizeof(int32_t), s);
    if (e != hipSuccess) return e;
    int32_t* ctr = order + L.num_queries;
    ms_zero<<<1, 64, 0, s>>>(ctr);
    ms_count<<<static_cast<unsigned>(blocks), kBlk, 0, s>>>(L, ctr);
    ms_place<<<static_cast<unsigned>(blocks), kBlk, 0, s>>>(L, ctr, order);
  }
  int dev = 0, cus = 0;
  e = hipGetDevice(&dev);
  if (e == hipSuccess) e = hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
  if (e != hipSuccess) {
    if (order) (void)hipFreeAsync(order, s);
    return e;
  }
  // a context with a mesh deeper than the LDS stack: a persistent grid whose
  // lanes keep their stacks in device memory (queries on deeper meshes
  // than kMsDeepMax are refused by load_query)
  const bool deep = L.max_depth > kMsStack;
  if (deep) L.max_depth = std::min(L.max_depth, kMsDeepMax);
  long long grid = blocks;
  if (deep) grid = std::min<long long>(grid, static_cast<long long>(cus) * kMsDeepBlocksPerCu);
  if (L.contact) grid = std::min<long long>(grid, static_cast<long long>(cus) * kEpaBlocksPerCu);
  int* deep_stack = nullptr;
  if (deep) {
    e = hipMallocAsync(reinterpret_cast<void**>(&deep_stack),
                       static_cast<size_t>(grid) * kBlk * L.max_depth * sizeof(int), s);
    if (e != hipSuccess) {
      if (order) (void)hipFreeAsync(order, s);
      return e;
    }
  }
  if (!L.contact) {
    launch_variant<false>(waves, static_cast<unsigned>(grid), s, L, order, nullptr, nullptr,
                          deep_stack);
    e = hipGetLastError();
    if (deep_stack) (void)hipFreeAsync(deep_stack, s);
    if (order) (void)hipFreeAsync(order, s);
    return e;
  }
  // contacts with EPA queued: the descent on every query at once, then EPA
  // on a persistent grid sized to its memory
  const size_t nslots = static_cast<size_t>(L.num_queries) * static_cast<size_t>(L.cap);
  if (!deep && nslots * sizeof(MsEpaSeed) <= kEpaQueueBudget) {
    MsEpaSeed* seeds = nullptr;
    unsigned long long* ctr = nullptr;  // [0] seeds queued, [1] seeds handed out
    e = hipMallocAsync(reinterpret_cast<void**>(&seeds), nslots * sizeof(MsEpaSeed), s);
    if (e == hipSuccess) {
      e = hipMallocAsync(reinterpret_cast<void**>(&ctr), 2 * sizeof(unsigned long long), s);
      if (e != hipSuccess) (void)hipFreeAsync(seeds, s);
    }
    if (e != hipSuccess) {
      if (order) (void)hipFreeAsync(order, s);
      return e;
    }
    ms_zero_u64<<<1, 64, 0, s>>>(ctr);
    launch_queued(waves, static_cast<unsigned>(blocks), s, L, order, seeds, ctr);
    e = hipGetLastError();
    const long long egrid = static_cast<long long>(cus) * kEpaBlocksPerCu;
    const size_t nl = static_cast<size_t>(egrid) * kBlk;
    shp::EpaLane* lanes = nullptr;
    shp::EpaWave* wv = nullptr;
    if (e == hipSuccess) e = hipMallocAsync(reinterpret_cast<void**>(&lanes), nl * sizeof(shp::EpaLane), s);
    if (e == hipSuccess) {
      e = hipMallocAsync(reinterpret_cast<void**>(&wv), nl / 64 * sizeof(shp::EpaWave), s);
      if (e != hipSuccess) (void)hipFreeAsync(lanes, s);
    }
    if (e == hipSuccess) {
      meshshape_epa<2><<<static_cast<unsigned>(egrid), kBlk, 0, s>>>(L, seeds, ctr, ctr + 1,
                                                                        lanes, wv);
      e = hipGetLastError();
      (void)hipFreeAsync(lanes, s);
      (void)hipFreeAsync(wv, s);
    }
    (void)hipFreeAsync(seeds, s);
    (void)hipFreeAsync(ctr, s);
    if (order) (void)hipFreeAsync(order, s);
    return e;
  }
  const size_t nl = static_cast<size_t>(grid) * kBlk;
  shp::EpaLane* lanes = nullptr;
  shp::EpaWave* wv = nullptr;
  e = hipMallocAsync(reinterpret_cast<void**>(&lanes), nl * sizeof(shp::EpaLane), s);
  if (e == hipSuccess) {
    e = hipMallocAsync(reinterpret_cast<void**>(&wv), nl / 64 * sizeof(shp::EpaWave), s);
    if (e != hipSuccess) (void)hipFreeAsync(lanes, s);
  }
  if (e != hipSuccess) {
    if (deep_stack) (void)hipFreeAsync(deep_stack, s);
    if (order) (void)hipFreeAsync(order, s);
    return e;
  }
  launch_variant<true>(waves, static_cast<unsigned>(grid), s, L, order, lanes, wv, deep_stack);
  e = hipGetLastError();
  (void)hipFreeAsync(lanes, s);
  (void)hipFreeAsync(wv, s);
  if (deep_stack) (void)hipFreeAsync(deep_stack, s);
  if (order) (void)hipFreeAsync(order, s);
  return e;
}

}  // namespace fclx
