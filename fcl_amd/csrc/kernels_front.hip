// Mesh-vs-mesh collide() with BVTT front lists (geometry/bvh/detail/
// BVH_front.h:49-72): one lane per sequence of queries over one mesh pair,
// the queries of a sequence run in order, each starting from the front the
// previous one left.
//
// detail::collide(node, front_list) (collision_node-inl.h:71-82): an empty
// front means a traversal from the roots that records the front --
// collisionRecurse (traversal_recurse-inl.h:82-129) adds every leaf pair
// before its BV test and every internal pair found disjoint, and with a
// front list never stops early.  A front with entries is propagated
// (propagateBVHFrontListCollisionRecurse, traversal_recurse-inl.h:463-521):
// each entry in list order -- a leaf pair is invalidated and re-traversed
// into a second list; an internal pair that now overlaps is invalidated and
// its two children are traversed, appending to the list being walked, so
// the walk reaches those entries too (a leaf pair reached that way is
// tested a second time).  Then the invalid entries are dropped (the rest
// keep their order) and the second list is appended.
//
// Layout per sequence: `front` holds front_cap (b1, b2) entries; the walk
// appends behind the current front in the same array, and the second list
// goes to `second` (front_cap entries).  An entry is invalidated by storing
// b1 as -(b1 + 1).  A sequence whose front outgrows front_cap flags
// kStatusFrontOverflow, and its front is emptied (the next query starts from
// the roots).
#include <hip/hip_runtime.h>

#include "device/exact64.h"
#include "device/pose64.h"
#include "kernels.h"

namespace fclx {

namespace {

struct Lane {
  const MeshDev* m1;
  const MeshDev* m2;
  M3 R;
  V3 T;
  int q;
  int cap;
  int count;
  long long bv, leaf;
  bool overflow;
};

__device__ __forceinline__ bool disjoint(Lane& c, int b1, int b2) {
  ++c.bv;
  return !obb_overlap64(c.R, c.T, c.m1->obb64[b1], c.m2->obb64[b2]);
}

// firstOverSecond (bvh_collision_traversal_node-inl.h:77-89)
__device__ __forceinline__ bool first_over_second(const Lane& c, int b1, int b2) {
  const int fc1 = c.m1->obb64[b1].first_child, fc2 = c.m2->obb64[b2].first_child;
  return fc2 < 0 || (fc1 >= 0 && c.m1->size64[b1] > c.m2->size64[b2]);
}

// Leaf pair that passed its BV test (mesh_collision_traversal_node-inl.h:
// 526-620 without contact geometry): a contact while the cap has room.
__device__ void leaf(Lane& c, const FrontLaunch& L, int b1, int b2) {
  ++c.leaf;
  const int id1 = -(c.m1->obb64[b1].first_child + 1);
  const int id2 = -(c.m2->obb64[b2].first_child + 1);
  const int32_t* t1 = c.m1->tris + 3 * static_cast<size_t>(id1);
  const int32_t* t2 = c.m2->tris + 3 * static_cast<size_t>(id2);
  if (!tri_intersect64(vertex(*c.m1, t1[0]), vertex(*c.m1, t1[1]), vertex(*c.m1, t1[2]),
                       vertex(*c.m2, t2[0]), vertex(*c.m2, t2[1]), vertex(*c.m2, t2[2]), c.R,
                       c.T))
    return;
  if (c.count >= c.cap) return;
  const long long slot = static_cast<long long>(
      atomicAdd(reinterpret_cast<unsigned long long*>(L.num_pairs), 1ULL));
  if (slot < L.pair_capacity) {
    L.pairs[3 * slot] = c.q;
    L.pairs[3 * slot + 1] = id1;
    L.pairs[3 * slot + 2] = id2;
  } else if (L.pairs) {
    atomicOr(L.status, kStatusPairOverflow);
  }
  ++c.count;
}

// collisionRecurse from (b1, b2) with a front list: the front entries go
// to dst[*n] (room `room`).  The explicit stack pops a node's first child's
// subtree before its second child, the recursion's order.
template <int STACK>
__device__ void recurse(Lane& c, const FrontLaunch& L, int b1, int b2, int2* dst, int& n,
                        int room) {
  int2 stack[STACK];
  int sp = 0;
  stack[sp++] = make_int2(b1, b2);
  while (sp > 0) {
    const int2 e = stack[--sp];
    const int fc1 = c.m1->obb64[e.x].first_child, fc2 = c.m2->obb64[e.y].first_child;
    if (fc1 < 0 && fc2 < 0) {
      if (n < room) dst[n] = e; else c.overflow = true;
      ++n;
      if (!disjoint(c, e.x, e.y)) leaf(c, L, e.x, e.y);
      continue;
    }
    if (disjoint(c, e.x, e.y)) {
      if (n < room) dst[n] = e; else c.overflow = true;
      ++n;
      continue;
    }
    if (sp + 2 > STACK) {
      c.overflow = true;
      atomicOr(L.status, kStatusStackOverflow);
      return;
    }
    if (first_over_second(c, e.x, e.y)) {
      stack[sp++] = make_int2(fc1 + 1, e.y);
      stack[sp++] = make_int2(fc1, e.y);
    } else {
      stack[sp++] = make_int2(e.x, fc2 + 1);
      stack[sp++] = make_int2(e.x, fc2);
    }
  }
}

template <int STACK>
__global__ void __launch_bounds__(256) collide_front_kernel(FrontLaunch L) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= L.num_sequences) return;
  const int room = L.front_cap;
  int2* F = L.fronts + static_cast<size_t>(s) * room;
  int2* P = L.second + static_cast<size_t>(s) * room;
  int n = L.front_sizes[s];
  int pm1 = L.front_meshes[2 * s], pm2 = L.front_meshes[2 * s + 1];
  bool lost = false;  // a front outgrew its room: emptied at the end
  for (int q = L.seq_offsets[s]; q < L.seq_offsets[s + 1]; ++q) {
    const int i1 = L.query_meshes[2 * q], i2 = L.query_meshes[2 * q + 1];
    if (!query_ok(L.meshes, L.num_meshes, i1, i2, /*need_obbrss=*/true)) {
      L.counts[q] = 0;
      if (L.bv_tests) L.bv_tests[q] = 0;
      if (L.leaf_tests) L.leaf_tests[q] = 0;
      atomicOr(L.status, kStatusBadQuery);
      continue;
    }
    // a front belongs to one mesh pair: another pair starts from the roots
    if (i1 != pm1 || i2 != pm2) n = 0;
    pm1 = i1;
    pm2 = i2;
    Lane c;
    c.m1 = L.meshes + i1;
    c.m2 = L.meshes + i2;
    relative_pose(L.tf1 + 12 * static_cast<size_t>(q), L.tf2 + 12 * static_cast<size_t>(q), c.R,
                  c.T);
    c.q = q;
    c.cap = L.cap;
    c.count = 0;
    c.bv = c.leaf = 0;
    c.overflow = false;
    if (n == 0) {
      recurse<STACK>(c, L, 0, 0, F, n, room);
    } else {
      int len = n, np = 0;
      for (int i = 0; i < len && i < room; ++i) {
        const int2 e = F[i];
        const int fc1 = c.m1->obb64[e.x].first_child, fc2 = c.m2->obb64[e.y].first_child;
        if (fc1 < 0 && fc2 < 0) {
          F[i].x = -(e.x + 1);
          recurse<STACK>(c, L, e.x, e.y, P, np, room);
        } else if (!disjoint(c, e.x, e.y)) {
          F[i].x = -(e.x + 1);
          if (first_over_second(c, e.x, e.y)) {
            recurse<STACK>(c, L, fc1, e.y, F, len, room);
            recurse<STACK>(c, L, fc1 + 1, e.y, F, len, room);
          } else {
            recurse<STACK>(c, L, e.x, fc2, F, len, room);
            recurse<STACK>(c, L, e.x, fc2 + 1, F, len, room);
          }
        }
      }
      // drop the invalid entries, then append the second list
      int k = 0;
      const int kept = len < room ? len : room;
      for (int i = 0; i < kept; ++i) {
        const int2 e = F[i];
        if (e.x >= 0) F[k++] = e;
      }
      for (int i = 0; i < np && i < room; ++i) {
        if (k < room) F[k] = P[i]; else c.overflow = true;
        ++k;
      }
      n = k;
    }
    if (c.overflow || n > room) {
      lost = true;
      n = 0;
    }
    L.counts[q] = c.count;
    if (L.bv_tests) L.bv_tests[q] = c.bv;
    if (L.leaf_tests) L.leaf_tests[q] = c.leaf;
  }
  if (lost) atomicOr(L.status, kStatusFrontOverflow);
  L.front_sizes[s] = n;
  L.front_meshes[2 * s] = pm1;
  L.front_meshes[2 * s + 1] = pm2;
}

}  // namespace

hipError_t launch_collide_front(const FrontLaunch& L, int stack_need, hipStream_t s) {
  const int block = 256;
  const int grid = (L.num_sequences + block - 1) / block;
  if (grid == 0) return hipSuccess;
  if (stack_need <= 64)
    collide_front_kernel<64><<<grid, block, 0, s>>>(L);
  else if (stack_need <= 128)
    collide_front_kernel<128><<<grid, block, 0, s>>>(L);
  else
    collide_front_kernel<256><<<grid, block, 0, s>>>(L);
  return hipGetLastError();
}

}  // namespace fclx
