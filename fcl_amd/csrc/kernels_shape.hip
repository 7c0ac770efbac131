// Batched primitive-shape collide(): one shape pair per lane, FP64, FCL's
// GST_INDEP solver (closed forms for the five special pairs, GJK + EPA for
// every other pair).  The per-pair bookkeeping follows
// shape_collision_traversal_node-inl.h:69-134: with enable_contact the
// pair's contacts (deepest first when more than num_max_contacts), without
// it one contact and no geometry.
//
// GJK/EPA state (68 support vertices, 128 EPA faces, the horizon stack) is
// per-lane private memory; a wavefront runs one pair per lane, so the lanes
// of a wave take different branches when their pairs differ in type or in
// GJK/EPA iteration count.
#include <hip/hip_runtime.h>

#include "device/shape64.h"
#include "kernels.h"

namespace fclx {

static_assert(sizeof(fclx_shape) == 40, "fclx_shape layout");

namespace {

// A shape record is usable when its type is known, its dimensions are
// finite and positive where the reference's constructors require it, and a
// convex names a vertex range inside the pool.
__device__ __forceinline__ bool load_shape(const ShapeLaunch& L, int id, shp::Shape& s) {
  if (id < 0 || id >= L.num_shapes) return false;
  const fclx_shape& r = L.shapes[id];
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
      if (r.num_vertices < 1 || r.vertex_offset < 0 ||
          r.vertex_offset + r.num_vertices > L.num_vertices || L.vertices == nullptr)
        return false;
      s.verts = L.vertices + 3 * r.vertex_offset;
      return true;
  }
  return false;
}

__global__ void __launch_bounds__(128) shape_collide_kernel(ShapeLaunch L) {
  const int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= L.num_pairs) return;
  shp::Shape s1, s2;
  if (!load_shape(L, L.pair_shapes[2 * q], s1) || !load_shape(L, L.pair_shapes[2 * q + 1], s2)) {
    L.counts[q] = 0;
    atomicOr(L.status, kStatusBadQuery);
    return;
  }
  const shp::Xf a = shp::xf_load(L.tf1 + 12 * static_cast<size_t>(q));
  const shp::Xf b = shp::xf_load(L.tf2 + 12 * static_cast<size_t>(q));
  shp::Contact c[4];
  const int n = shp::shape_collide(s1, a, s2, b, L.cap, L.contact != 0, L.tol, c);
  L.counts[q] = n;
  if (L.contact) {
    const size_t base = static_cast<size_t>(q) * L.cap;
    for (int i = 0; i < n; ++i) {
      const size_t k = base + i;
#pragma unroll
      for (int j = 0; j < 3; ++j) {
        L.contact_pos[3 * k + j] = c[i].pos.v[j];
        L.contact_normal[3 * k + j] = c[i].normal.v[j];
      }
      L.contact_depth[k] = c[i].depth;
    }
  }
}

}  // namespace

hipError_t launch_shape_collide(const ShapeLaunch& L, hipStream_t s) {
  if (L.num_pairs <= 0) return hipSuccess;
  const int block = 128;
  const int grid = (L.num_pairs + block - 1) / block;
  shape_collide_kernel<<<grid, block, 0, s>>>(L);
  return hipGetLastError();
}

}  // namespace fclx
