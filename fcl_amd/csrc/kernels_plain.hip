// Plain FP64 BVH-vs-BVH traversal: one query per lane, depth-first with an
// explicit per-lane stack, exactly the reference's visit order.
//
// collide: traversal_recurse-inl.h:82-129 (collisionRecurse).  The recursion
// checks canStop() only between a node's first and second child; with an
// explicit stack the second child is popped right after the first child's
// subtree finishes, so testing the stop rule at every pop is the same rule.
// Leaf pairs run the OBB test before the triangle test, and every test is
// counted as the reference's num_bv_tests / num_leaf_tests.
//
// distance: traversal_recurse-inl.h:257-316 (distanceRecurse).  Both child
// bounds are computed, the nearer is visited first, and the farther is
// re-checked against the current minimum when it is popped.
#include <hip/hip_runtime.h>

#include "device/dist64.h"
#include "device/pose64.h"
#include "kernels.h"

namespace fclx {

template <int STACK>
__global__ void __launch_bounds__(256) collide_plain_kernel(CollideLaunch L) {
  const int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= L.num_queries) return;
  const int i1 = L.query_meshes[2 * q], i2 = L.query_meshes[2 * q + 1];
  if (!query_ok(L.meshes, L.num_meshes, i1, i2, /*need_obbrss=*/true)) {
    L.counts[q] = 0;
    if (L.bv_tests) L.bv_tests[q] = 0;
    if (L.leaf_tests) L.leaf_tests[q] = 0;
    atomicOr(L.status, kStatusBadQuery);
    return;
  }
  const MeshDev m1 = L.meshes[i1];
  const MeshDev m2 = L.meshes[i2];
  M3 R;
  V3 T;
  relative_pose(L.tf1 + 12 * static_cast<size_t>(q), L.tf2 + 12 * static_cast<size_t>(q), R, T);
  const int cap = L.cap;
  int count = 0;
  long long bv_tests = 0, leaf_tests = 0;
  int2 stack[STACK];
  int sp = 0;
  int b1 = 0, b2 = 0;
  bool overflow = false;
  while (true) {
    // visit (b1,b2)
    const ObbNode64& n1 = m1.obb64[b1];
    const ObbNode64& n2 = m2.obb64[b2];
    const int fc1 = n1.first_child, fc2 = n2.first_child;
    ++bv_tests;
    bool descend = false;
    if (obb_overlap64(R, T, n1, n2)) {
      if (fc1 < 0 && fc2 < 0) {
        ++leaf_tests;
        const int id1 = -(fc1 + 1), id2 = -(fc2 + 1);
        const int32_t* t1 = m1.tris + 3 * static_cast<size_t>(id1);
        const int32_t* t2 = m2.tris + 3 * static_cast<size_t>(id2);
        const V3 P1 = vertex(m1, t1[0]), P2 = vertex(m1, t1[1]), P3 = vertex(m1, t1[2]);
        const V3 Q1 = vertex(m2, t2[0]), Q2 = vertex(m2, t2[1]), Q3 = vertex(m2, t2[2]);
        if (L.contact) {
          if (tri_intersect64(P1, P2, P3, Q1, Q2, Q3, R, T)) {
            // mesh_collision_traversal_node-inl.h:578-598: the pair adds as
            // many of its 0-2 contacts as the cap leaves room for
            V3 pts[2], nrm;
            double depth;
            unsigned n = tri_contact64(P1, P2, P3, mul(R, Q1) + T, mul(R, Q2) + T,
                                       mul(R, Q3) + T, pts, nrm, depth);
            if (static_cast<unsigned>(cap) < static_cast<unsigned>(count) + n)
              n = cap > count ? static_cast<unsigned>(cap - count) : 0u;
            const double* w = L.tf1 + 12 * static_cast<size_t>(q);
            for (unsigned i = 0; i < n; ++i) {
              const long long slot = static_cast<long long>(
                  atomicAdd(reinterpret_cast<unsigned long long*>(L.num_pairs), 1ULL));
              if (slot < L.pair_capacity) {
                L.pairs[3 * slot] = q;
                L.pairs[3 * slot + 1] = id1;
                L.pairs[3 * slot + 2] = id2;
                const V3 pw = v3(dot(v3(w[0], w[1], w[2]), pts[i]) + w[9],
                                 dot(v3(w[3], w[4], w[5]), pts[i]) + w[10],
                                 dot(v3(w[6], w[7], w[8]), pts[i]) + w[11]);
                const V3 nw = v3(dot(v3(w[0], w[1], w[2]), nrm), dot(v3(w[3], w[4], w[5]), nrm),
                                 dot(v3(w[6], w[7], w[8]), nrm));
                double* pp = L.contact_pos + 3 * slot;
                double* np = L.contact_normal + 3 * slot;
                pp[0] = pw.x;
                pp[1] = pw.y;
                pp[2] = pw.z;
                np[0] = nw.x;
                np[1] = nw.y;
                np[2] = nw.z;
                L.contact_depth[slot] = depth;
              } else if (L.pairs) {  // no buffer: counts only
                atomicOr(L.status, kStatusPairOverflow);
              }
            }
            count += static_cast<int>(n);
          }
        } else if (tri_intersect64(P1, P2, P3, Q1, Q2, Q3, R, T)) {
          if (count < cap) {
            const long long slot = static_cast<long long>(
                atomicAdd(reinterpret_cast<unsigned long long*>(L.num_pairs), 1ULL));
            if (slot < L.pair_capacity) {
              L.pairs[3 * slot] = q;
              L.pairs[3 * slot + 1] = id1;
              L.pairs[3 * slot + 2] = id2;
            } else if (L.pairs) {  // no buffer: counts only
              atomicOr(L.status, kStatusPairOverflow);
            }
            ++count;
          }
        }
      } else {
        descend = true;
      }
    }
    if (descend) {
      // firstOverSecond (bvh_collision_traversal_node-inl.h:77-89)
      const bool split1 = (fc2 < 0) || (fc1 >= 0 && (m1.size64[b1] > m2.size64[b2]));
      if (sp >= STACK) {
        overflow = true;
        break;
      }
      if (split1) {
        stack[sp++] = make_int2(fc1 + 1, b2);
        b1 = fc1;
      } else {
        stack[sp++] = make_int2(b1, fc2 + 1);
        b2 = fc2;
      }
      continue;
    }
    // pop the next pending second child, unless the request is satisfied
    if (sp == 0 || (count > 0 && count >= cap)) break;
    const int2 e = stack[--sp];
    b1 = e.x;
    b2 = e.y;
  }
  L.counts[q] = count;
  if (L.bv_tests) L.bv_tests[q] = bv_tests;
  if (L.leaf_tests) L.leaf_tests[q] = leaf_tests;
  if (overflow) atomicOr(L.status, kStatusStackOverflow);
}

struct DistEntry {
  int b1, b2;
  double d;
};

template <int STACK>
__global__ void __launch_bounds__(256) distance_plain_kernel(DistanceLaunch L) {
  const int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= L.num_queries) return;
  const int i1 = L.query_meshes[2 * q], i2 = L.query_meshes[2 * q + 1];
  if (!query_ok(L.meshes, L.num_meshes, i1, i2, /*need_obbrss=*/false)) {
    L.distance[q] = 1.7976931348623157e308;
    if (L.ids) {
      L.ids[2 * q] = -1;
      L.ids[2 * q + 1] = -1;
    }
    if (L.bv_tests) L.bv_tests[q] = 0;
    if (L.leaf_tests) L.leaf_tests[q] = 0;
    atomicOr(L.status, kStatusBadQuery);
    return;
  }
  const MeshDev m1 = L.meshes[i1];
  const MeshDev m2 = L.meshes[i2];
  const double* tf1 = L.tf1 + 12 * static_cast<size_t>(q);
  const double* tf2 = L.tf2 + 12 * static_cast<size_t>(q);
  M3 R;
  V3 T;
  inverse_compose(tf1, tf2, R, T);
  long long bv_tests = 0, leaf_tests = 0;

  double best = 1.7976931348623157e308;  // DistanceResult::min_distance starts at max()
  int best1 = -1, best2 = -1;
  V3 bestP = v3(0, 0, 0), bestQ = v3(0, 0, 0);
  // triangle-pair distance with model 2's triangle moved into model 1's frame
  auto leaf = [&](int id1, int id2) {
    const int32_t* t1 = m1.tris + 3 * static_cast<size_t>(id1);
    const int32_t* t2 = m2.tris + 3 * static_cast<size_t>(id2);
    const V3 S[3] = {vertex(m1, t1[0]), vertex(m1, t1[1]), vertex(m1, t1[2])};
    const V3 Tt[3] = {mul(R, vertex(m2, t2[0])) + T, mul(R, vertex(m2, t2[1])) + T,
                      mul(R, vertex(m2, t2[2])) + T};
    V3 P, Q;
    const double d = tri_distance64(S, Tt, P, Q);
    if (best > d) {  // DistanceResult::update keeps the first minimum
      best = d;
      best1 = id1;
      best2 = id2;
      bestP = P;
      bestQ = Q;
    }
  };
  // seed with triangle 0 of each model (distancePreprocessOrientedNode)
  leaf(0, 0);

  DistEntry stack[STACK];
  int sp = 0;
  int b1 = 0, b2 = 0;
  bool overflow = false;
  while (true) {
    const RssNode64& n1 = m1.rss64[b1];
    const RssNode64& n2 = m2.rss64[b2];
    const int fc1 = n1.first_child, fc2 = n2.first_child;
    bool have_next = false;
    if (fc1 < 0 && fc2 < 0) {
      ++leaf_tests;
      leaf(-(fc1 + 1), -(fc2 + 1));
    } else {
      int a1, a2, c1, c2;
      if ((fc2 < 0) || (fc1 >= 0 && (m1.size64[b1] > m2.size64[b2]))) {
        a1 = fc1; a2 = b2; c1 = fc1 + 1; c2 = b2;
      } else {
        a1 = b1; a2 = fc2; c1 = b1; c2 = fc2 + 1;
      }
      const double d1 = rss_distance64(R, T, m1.rss64[a1], m2.rss64[a2]);
      const double d2 = rss_distance64(R, T, m1.rss64[c1], m2.rss64[c2]);
      bv_tests += 2;
      // visit the nearer child now (if it can improve), keep the other
      int n1a, n1b, f1a, f1b;
      double dn, df;
      if (d2 < d1) {
        n1a = c1; n1b = c2; dn = d2; f1a = a1; f1b = a2; df = d1;
      } else {
        n1a = a1; n1b = a2; dn = d1; f1a = c1; f1b = c2; df = d2;
      }
      if (sp >= STACK) {
        overflow = true;
        break;
      }
      stack[sp++] = DistEntry{f1a, f1b, df};
      if (!(dn >= best)) {  // canStop(c) with rel_err = abs_err = 0
        b1 = n1a;
        b2 = n1b;
        have_next = true;
      }
    }
    if (have_next) continue;
    bool found = false;
    while (sp > 0) {
      const DistEntry e = stack[--sp];
      if (!(e.d >= best)) {
        b1 = e.b1;
        b2 = e.b2;
        found = true;
        break;
      }
    }
    if (!found) break;
  }
  L.distance[q] = best;
  if (L.ids) {
    L.ids[2 * q] = best1;
    L.ids[2 * q + 1] = best2;
  }
  if (L.points) {
    // nearest points back to the world frame through tf1
    M3 R1;
    V3 t1;
    load_pose(tf1, R1, t1);
    const V3 wp = mul(R1, bestP) + t1, wq = mul(R1, bestQ) + t1;
    double* o = L.points + 6 * static_cast<size_t>(q);
    o[0] = wp.x; o[1] = wp.y; o[2] = wp.z;
    o[3] = wq.x; o[4] = wq.y; o[5] = wq.z;
  }
  if (L.bv_tests) L.bv_tests[q] = bv_tests;
  if (L.leaf_tests) L.leaf_tests[q] = leaf_tests;
  if (overflow) atomicOr(L.status, kStatusStackOverflow);
}

hipError_t launch_collide_plain(const CollideLaunch& L, int stack_need, hipStream_t s) {
  const int block = 256;
  const int grid = (L.num_queries + block - 1) / block;
  if (grid == 0) return hipSuccess;
  if (stack_need <= 64)
    collide_plain_kernel<64><<<grid, block, 0, s>>>(L);
  else if (stack_need <= 128)
    collide_plain_kernel<128><<<grid, block, 0, s>>>(L);
  else
    collide_plain_kernel<256><<<grid, block, 0, s>>>(L);
  return hipGetLastError();
}

hipError_t launch_distance_plain(const DistanceLaunch& L, int stack_need, hipStream_t s) {
  const int block = 256;
  const int grid = (L.num_queries + block - 1) / block;
  if (grid == 0) return hipSuccess;
  if (stack_need <= 64)
    distance_plain_kernel<64><<<grid, block, 0, s>>>(L);
  else if (stack_need <= 128)
    distance_plain_kernel<128><<<grid, block, 0, s>>>(L);
  else
    distance_plain_kernel<256><<<grid, block, 0, s>>>(L);
  return hipGetLastError();
}

}  // namespace fclx
