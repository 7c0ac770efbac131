// fcl::collide / fcl::distance / BVHModel drop-in over the C ABI.
//
// Dispatch mirrors collide<S,Solver> (narrowphase/collision-inl.h:93-149)
// and distance<Solver> (distance-inl.h:90-193): an invalid solver type
// returns (size_t)-1 / -1, num_max_contacts == 0 warns and returns 0, a
// pre-satisfied result short-circuits (collision_func_matrix-inl.h:580,
// distance_func_matrix-inl.h:396), and a geometry pair outside the mesh
// path prints the reference's "not supported" warning with the neutral
// return.  The mesh-mesh cells and the primitive-primitive cells of the
// five supported shapes run on the GPU through one process-wide engine
// context (device 0, or FCLX_DEVICE).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <climits>
#include <cstdio>
#include <cstdlib>
#include <iostream>
#include <mutex>

#include "../../../include/fcl/fcl.h"
#include "../../../include/fclx.h"

namespace fcl {
namespace detail {

namespace {
std::once_flag g_once;
fclx_context* g_ctx = nullptr;
std::atomic<bool> g_warned_cost{false};

fclx_context* engine() {
  std::call_once(g_once, [] {
    const char* d = std::getenv("FCLX_DEVICE");
    const int dev = d ? std::atoi(d) : 0;
    if (fclx_init(dev, &g_ctx) != FCLX_OK) {
      std::cerr << "fcl: GPU engine initialisation failed on device " << dev << "\n";
      g_ctx = nullptr;
    }
  });
  return g_ctx;
}
}  // namespace

struct DeviceMesh {
  int32_t id = -1;
  ~DeviceMesh() {
    if (g_ctx && id >= 0) fclx_mesh_destroy(g_ctx, id);
  }
};

int device_mesh_id(const DeviceMesh* m) { return m ? m->id : -1; }

int upload_mesh(int bv_type, const std::vector<Vector3d>& vertices,
                const std::vector<Triangle>& tris, std::shared_ptr<DeviceMesh>& out) {
  fclx_context* ctx = engine();
  if (!ctx) return BVH_ERR_UNKNOWN;
  std::vector<double> v(3 * vertices.size());
  for (size_t i = 0; i < vertices.size(); ++i)
    for (int k = 0; k < 3; ++k) v[3 * i + k] = vertices[i][k];
  std::vector<int32_t> t(3 * tris.size());
  for (size_t i = 0; i < tris.size(); ++i)
    for (int k = 0; k < 3; ++k) {
      if (tris[i][k] >= vertices.size()) return BVH_ERR_INCORRECT_DATA;
      t[3 * i + k] = static_cast<int32_t>(tris[i][k]);
    }
  auto m = std::make_shared<DeviceMesh>();
  const int rc = fclx_mesh_create(ctx, bv_type, v.data(), static_cast<int32_t>(vertices.size()),
                                  t.data(), static_cast<int32_t>(tris.size()), &m->id);
  switch (rc) {
    case FCLX_OK:
      out = m;
      return BVH_OK;
    case FCLX_ERR_EMPTY_MODEL:
      return BVH_ERR_BUILD_EMPTY_MODEL;
    case FCLX_ERR_OUT_OF_MEMORY:
      return BVH_ERR_MODEL_OUT_OF_MEMORY;
    case FCLX_ERR_INVALID_ARGUMENT:
      return BVH_ERR_INCORRECT_DATA;
    default:
      std::cerr << "fcl: mesh upload failed: " << fclx_last_error(ctx) << "\n";
      return BVH_ERR_UNKNOWN;
  }
}

namespace {

void to_pose(const Transform3d& tf, double out[12]) {
  for (int i = 0; i < 9; ++i) out[i] = tf.R.m[i];
  for (int i = 0; i < 3; ++i) out[9 + i] = tf.t[i];
}

// built BVHModel of the given node type, or nullptr
int mesh_id_of(const CollisionGeometry<double>* g, NODE_TYPE want) {
  if (!g || g->getObjectType() != OT_BVH || g->getNodeType() != want) return -1;
  if (want == BV_OBBRSS) return static_cast<const BVHModel<OBBRSSd>*>(g)->deviceMeshId();
  if (want == BV_RSS) return static_cast<const BVHModel<RSSd>*>(g)->deviceMeshId();
  return -1;
}

// One of the five primitives as an engine shape record (with the vertex
// list of a convex); false for any other geometry.
bool to_shape(const CollisionGeometry<double>* g, fclx_shape& r, std::vector<double>& verts) {
  if (!g || g->getObjectType() != OT_GEOM) return false;
  r = fclx_shape{};
  switch (g->getNodeType()) {
    case GEOM_BOX: {
      const auto* b = static_cast<const Boxd*>(g);
      r.type = FCLX_SHAPE_BOX;
      for (int k = 0; k < 3; ++k) r.dims[k] = b->side[k];
      return true;
    }
    case GEOM_SPHERE:
      r.type = FCLX_SHAPE_SPHERE;
      r.dims[0] = static_cast<const Sphered*>(g)->radius;
      return true;
    case GEOM_CAPSULE: {
      const auto* c = static_cast<const Capsuled*>(g);
      r.type = FCLX_SHAPE_CAPSULE;
      r.dims[0] = c->radius;
      r.dims[1] = c->lz;
      return true;
    }
    case GEOM_CYLINDER: {
      const auto* c = static_cast<const Cylinderd*>(g);
      r.type = FCLX_SHAPE_CYLINDER;
      r.dims[0] = c->radius;
      r.dims[1] = c->lz;
      return true;
    }
    case GEOM_CONVEX: {
      const auto& v = static_cast<const Convexd*>(g)->getVertices();
      r.type = FCLX_SHAPE_CONVEX;
      r.num_vertices = static_cast<int32_t>(v.size());
      verts.resize(3 * v.size());
      for (size_t i = 0; i < v.size(); ++i)
        for (int k = 0; k < 3; ++k) verts[3 * i + k] = v[i][k];
      return true;
    }
    default:
      return false;
  }
}

// the pairs both solvers answer in closed form (gjk_solver_libccd-inl.h:
// 244-252, gjk_solver_indep-inl.h:248-256)
bool closed_form_pair(int t1, int t2) {
  if (t1 == t2) return t1 == FCLX_SHAPE_SPHERE || t1 == FCLX_SHAPE_BOX;
  const int o = t1 == FCLX_SHAPE_SPHERE ? t2 : (t2 == FCLX_SHAPE_SPHERE ? t1 : -1);
  return o == FCLX_SHAPE_BOX || o == FCLX_SHAPE_CAPSULE || o == FCLX_SHAPE_CYLINDER;
}

std::atomic<bool> g_warned_libccd{false};
std::atomic<bool> g_warned_guess{false};

// ShapeShapeCollide (collision_func_matrix-inl.h:278-311) through the GPU
std::size_t shape_collide(const CollisionGeometry<double>* o1, const fclx_shape& r1,
                          const std::vector<double>& v1, const Transform3d& tf1,
                          const CollisionGeometry<double>* o2, const fclx_shape& r2,
                          const std::vector<double>& v2, const Transform3d& tf2,
                          const CollisionRequest<double>& request,
                          CollisionResult<double>& result) {
  if (request.isSatisfied(result)) return result.numContacts();
  if (request.gjk_solver_type == GST_LIBCCD && !closed_form_pair(r1.type, r2.type) &&
      !g_warned_libccd.exchange(true))
    std::cerr << "fcl: the libccd solver is not available; GJK/EPA pairs use GST_INDEP\n";
  if (request.enable_cached_gjk_guess && !g_warned_guess.exchange(true))
    std::cerr << "fcl: cached GJK guesses are not supported; GJK starts from (1, 0, 0)\n";
  fclx_context* ctx = engine();
  if (!ctx) return result.numContacts();
  const std::size_t room = request.num_max_contacts - result.numContacts();
  const int32_t cap = room > 4 ? 4 : static_cast<int32_t>(room);
  double p1[12], p2[12];
  to_pose(tf1, p1);
  to_pose(tf2, p2);
  double pos[12], nrm[12], dep[4];
  int32_t n = 0;
  const int rc = fclx_shape_collide(ctx, &r1, v1.empty() ? nullptr : v1.data(), p1, &r2,
                                    v2.empty() ? nullptr : v2.data(), p2, cap,
                                    request.enable_contact ? 1 : 0, request.gjk_tolerance, pos,
                                    nrm, dep, &n, nullptr);
  if (rc != FCLX_OK) {
    std::cerr << "fcl: shape collide failed: " << fclx_last_error(ctx) << "\n";
    return result.numContacts();
  }
  for (int32_t i = 0; i < n; ++i) {
    if (request.enable_contact)
      result.addContact(Contact<double>(
          o1, o2, Contact<double>::NONE, Contact<double>::NONE,
          Vector3d(pos[3 * i], pos[3 * i + 1], pos[3 * i + 2]),
          Vector3d(nrm[3 * i], nrm[3 * i + 1], nrm[3 * i + 2]), dep[i]));
    else
      result.addContact(Contact<double>(o1, o2, Contact<double>::NONE, Contact<double>::NONE));
  }
  return result.numContacts();
}

// BVHShapeCollider<OBBRSS> (collision_func_matrix-inl.h:500-521,
// orientedBVHShapeCollide 375-419) through the batch kernel with one query:
// contacts (triangle, NONE) in the reference's order, appended to result.
std::size_t mesh_shape_collide(const CollisionGeometry<double>* o1, int mesh,
                               const Transform3d& tf1, const CollisionGeometry<double>* o2,
                               const fclx_shape& r2, const std::vector<double>& v2,
                               const Transform3d& tf2, const CollisionRequest<double>& request,
                               CollisionResult<double>& result) {
  if (request.isSatisfied(result)) return result.numContacts();
  if (request.gjk_solver_type == GST_LIBCCD && !g_warned_libccd.exchange(true))
    std::cerr << "fcl: the libccd solver is not available; GJK/EPA pairs use GST_INDEP\n";
  if (request.enable_cached_gjk_guess && !g_warned_guess.exchange(true))
    std::cerr << "fcl: cached GJK guesses are not supported; GJK starts from (1, 0, 0)\n";
  if (request.enable_cost && !g_warned_cost.exchange(true))
    std::cerr << "fcl: cost sources are not computed on the GPU mesh path\n";
  fclx_context* ctx = engine();
  if (!ctx) return result.numContacts();
  fclx_mesh_info info{};
  if (fclx_mesh_get_info(ctx, mesh, &info) != FCLX_OK) return result.numContacts();
  // a mesh-shape query has at most one contact per triangle
  const std::size_t room = request.num_max_contacts - result.numContacts();
  const int32_t cap = static_cast<int32_t>(
      std::max<std::size_t>(1, std::min<std::size_t>(room, static_cast<std::size_t>(info.num_tris))));
  fclx_shape rec = r2;
  rec.vertex_offset = 0;
  // one device block: ids, record, vertices, poses, counts, triangles, geometry
  const size_t nv = v2.size() / 3;
  const size_t doubles = 24 + 3 * nv + 7 * static_cast<size_t>(cap);
  const size_t ints = 3 + static_cast<size_t>(cap);
  const size_t bytes = 8 * doubles + sizeof(fclx_shape) + 4 * ints + 64;
  char* d = nullptr;
  if (hipMalloc(reinterpret_cast<void**>(&d), bytes) != hipSuccess) {
    std::cerr << "fcl: mesh-shape collide: device allocation failed\n";
    return result.numContacts();
  }
  double* d_tf = reinterpret_cast<double*>(d);
  double* d_verts = d_tf + 24;
  double* d_pos = d_verts + 3 * nv;
  double* d_nrm = d_pos + 3 * static_cast<size_t>(cap);
  double* d_dep = d_nrm + 3 * static_cast<size_t>(cap);
  fclx_shape* d_rec = reinterpret_cast<fclx_shape*>(d_dep + cap);
  int32_t* d_int = reinterpret_cast<int32_t*>(d_rec + 1);
  int32_t* d_prims = d_int + 3;
  int32_t hin[3] = {mesh, 0, -1};
  double tfs[24];
  to_pose(tf1, tfs);
  to_pose(tf2, tfs + 12);
  bool ok = hipMemcpy(d_tf, tfs, sizeof(tfs), hipMemcpyHostToDevice) == hipSuccess &&
            hipMemcpy(d_rec, &rec, sizeof(rec), hipMemcpyHostToDevice) == hipSuccess &&
            hipMemcpy(d_int, hin, sizeof(hin), hipMemcpyHostToDevice) == hipSuccess &&
            (nv == 0 || hipMemcpy(d_verts, v2.data(), 8 * v2.size(), hipMemcpyHostToDevice) ==
                            hipSuccess);
  int32_t n = 0;
  std::vector<int32_t> prims(cap);
  std::vector<double> geo(7 * static_cast<size_t>(cap));
  if (ok) {
    fclx_mesh_shape_collide_batch_args a{};
    a.num_queries = 1;
    a.query_meshes = d_int;
    a.query_shapes = d_int + 1;
    a.num_shapes = 1;
    a.shapes = d_rec;
    a.vertices = nv ? d_verts : nullptr;
    a.num_vertices = static_cast<int64_t>(nv);
    a.tf1 = d_tf;
    a.tf2 = d_tf + 12;
    a.num_max_contacts = cap;
    a.enable_contact = request.enable_contact ? 1 : 0;
    a.gjk_tolerance = request.gjk_tolerance;
    a.counts = d_int + 2;
    a.primitives = d_prims;
    a.contact_pos = d_pos;
    a.contact_normal = d_nrm;
    a.contact_depth = d_dep;
    ok = fclx_mesh_shape_collide_batch(ctx, &a, nullptr) == FCLX_OK &&
         fclx_sync(ctx, nullptr) == FCLX_OK &&
         hipMemcpy(&n, d_int + 2, 4, hipMemcpyDeviceToHost) == hipSuccess && n >= 0 &&
         n <= cap &&
         hipMemcpy(prims.data(), d_prims, 4 * static_cast<size_t>(n), hipMemcpyDeviceToHost) ==
             hipSuccess &&
         (!request.enable_contact ||
          hipMemcpy(geo.data(), d_pos, 8 * 7 * static_cast<size_t>(cap),
                    hipMemcpyDeviceToHost) == hipSuccess);
  }
  (void)hipFree(d);
  if (!ok) {
    std::cerr << "fcl: mesh-shape collide failed: " << fclx_last_error(ctx) << "\n";
    return result.numContacts();
  }
  const double* pos = geo.data();
  const double* nrm = pos + 3 * static_cast<size_t>(cap);
  const double* dep = nrm + 3 * static_cast<size_t>(cap);
  for (int32_t i = 0; i < n; ++i) {
    if (request.enable_contact)
      result.addContact(Contact<double>(
          o1, o2, prims[i], Contact<double>::NONE,
          Vector3d(pos[3 * i], pos[3 * i + 1], pos[3 * i + 2]),
          Vector3d(nrm[3 * i], nrm[3 * i + 1], nrm[3 * i + 2]), dep[i]));
    else
      result.addContact(Contact<double>(o1, o2, prims[i], Contact<double>::NONE));
  }
  return result.numContacts();
}

void unsupported(const char* what, const CollisionGeometry<double>* o1,
                 const CollisionGeometry<double>* o2) {
  std::cerr << "Warning: " << what << " function between node type "
            << (o1 ? o1->getNodeType() : BV_UNKNOWN) << " and node type "
            << (o2 ? o2->getNodeType() : BV_UNKNOWN) << " is not supported\n";
}

}  // namespace
}  // namespace detail

template <>
std::size_t collide<double>(const CollisionGeometry<double>* o1, const Transform3d& tf1,
                            const CollisionGeometry<double>* o2, const Transform3d& tf2,
                            const CollisionRequest<double>& request,
                            CollisionResult<double>& result) {
  if (request.gjk_solver_type != GST_LIBCCD && request.gjk_solver_type != GST_INDEP) {
    std::cerr << "Warning! Invalid GJK solver\n";
    return static_cast<std::size_t>(-1);
  }
  if (request.num_max_contacts == 0) {
    std::cerr << "Warning: should stop early as num_max_contact is " << request.num_max_contacts
              << " !\n";
    return 0;
  }
  {
    fclx_shape r1, r2;
    std::vector<double> v1, v2;
    if (detail::to_shape(o1, r1, v1) && detail::to_shape(o2, r2, v2))
      return detail::shape_collide(o1, r1, v1, tf1, o2, r2, v2, tf2, request, result);
  }
  const int id1 = detail::mesh_id_of(o1, BV_OBBRSS);
  {
    fclx_shape r2;
    std::vector<double> v2;
    if (id1 >= 0 && detail::to_shape(o2, r2, v2))
      return detail::mesh_shape_collide(o1, id1, tf1, o2, r2, v2, tf2, request, result);
  }
  const int id2 = detail::mesh_id_of(o2, BV_OBBRSS);
  {
    // a shape first and a mesh second: the reference calls the mesh-shape
    // cell with the two swapped (collision-inl.h:120-129), so the contacts
    // name the mesh first
    fclx_shape r1;
    std::vector<double> v1;
    if (id2 >= 0 && detail::to_shape(o1, r1, v1))
      return detail::mesh_shape_collide(o2, id2, tf2, o1, r1, v1, tf1, request, result);
  }
  if (id1 < 0 || id2 < 0) {
    detail::unsupported("collision", o1, o2);
    return 0;
  }
  if (request.isSatisfied(result)) return result.numContacts();
  if (request.enable_cost && !detail::g_warned_cost.exchange(true))
    std::cerr << "fcl: cost sources are not computed on the GPU mesh path\n";
  const std::size_t have = result.numContacts();
  const std::size_t room = request.num_max_contacts - have;
  const int32_t cap = room > static_cast<std::size_t>(INT_MAX) ? INT_MAX : static_cast<int32_t>(room);
  double p1[12], p2[12];
  detail::to_pose(tf1, p1);
  detail::to_pose(tf2, p2);
  fclx_context* ctx = detail::engine();
  int32_t room_n = std::min<int32_t>(cap, 4096);
  std::vector<int32_t> pairs(2 * static_cast<size_t>(room_n));
  std::vector<double> geo;  // enable_contact: pos (3), normal (3), depth per contact
  auto run = [&](int32_t capacity, int32_t& n) {
    if (!request.enable_contact)
      return fclx_collide(ctx, id1, p1, id2, p2, cap, pairs.data(), capacity, &n, nullptr);
    geo.resize(7 * static_cast<size_t>(capacity));
    return fclx_collide_contact(ctx, id1, p1, id2, p2, cap, pairs.data(), geo.data(),
                                geo.data() + 3 * static_cast<size_t>(capacity),
                                geo.data() + 6 * static_cast<size_t>(capacity), capacity, &n,
                                nullptr);
  };
  int32_t n = 0;
  int rc = run(room_n, n);
  if (rc == FCLX_ERR_CAPACITY) {
    room_n = n;
    pairs.resize(2 * static_cast<size_t>(n));
    rc = run(room_n, n);
  }
  if (rc != FCLX_OK) {
    std::cerr << "fcl: GPU collide failed: " << fclx_last_error(ctx) << "\n";
    return result.numContacts();
  }
  for (int32_t i = 0; i < n; ++i) {
    if (!request.enable_contact) {
      result.addContact(Contact<double>(o1, o2, pairs[2 * i], pairs[2 * i + 1]));
      continue;
    }
    const double* pos = geo.data() + 3 * static_cast<size_t>(i);
    const double* nrm = geo.data() + 3 * static_cast<size_t>(room_n) + 3 * static_cast<size_t>(i);
    const double depth = geo[6 * static_cast<size_t>(room_n) + i];
    result.addContact(Contact<double>(o1, o2, pairs[2 * i], pairs[2 * i + 1],
                                      Vector3d(pos[0], pos[1], pos[2]),
                                      Vector3d(nrm[0], nrm[1], nrm[2]), depth));
  }
  return result.numContacts();
}

template <>
double distance<double>(const CollisionGeometry<double>* o1, const Transform3d& tf1,
                        const CollisionGeometry<double>* o2, const Transform3d& tf2,
                        const DistanceRequest<double>& request, DistanceResult<double>& result) {
  if (request.gjk_solver_type != GST_LIBCCD && request.gjk_solver_type != GST_INDEP) return -1;
  const NODE_TYPE nt = o1 ? o1->getNodeType() : BV_UNKNOWN;
  const int id1 = detail::mesh_id_of(o1, nt == BV_RSS ? BV_RSS : BV_OBBRSS);
  const int id2 = detail::mesh_id_of(o2, nt == BV_RSS ? BV_RSS : BV_OBBRSS);
  if (id1 < 0 || id2 < 0) {
    detail::unsupported("distance", o1, o2);
    return std::numeric_limits<double>::max();
  }
  if (request.isSatisfied(result)) return result.min_distance;
  double p1[12], p2[12];
  detail::to_pose(tf1, p1);
  detail::to_pose(tf2, p2);
  fclx_context* ctx = detail::engine();
  double d = 0, pts[6];
  int32_t ids[2];
  const int rc = fclx_distance(ctx, id1, p1, id2, p2, request.enable_nearest_points ? 1 : 0, &d,
                               pts, ids, nullptr);
  if (rc != FCLX_OK) {
    std::cerr << "fcl: GPU distance failed: " << fclx_last_error(ctx) << "\n";
    return result.min_distance;
  }
  if (request.enable_nearest_points)
    result.update(d, o1, o2, ids[0], ids[1], Vector3d(pts[0], pts[1], pts[2]),
                  Vector3d(pts[3], pts[4], pts[5]));
  else
    result.update(d, o1, o2, ids[0], ids[1]);
  return result.min_distance;
}

std::size_t collideBatch(const std::vector<CollideQuery>& queries,
                         const CollisionRequest<double>& request,
                         std::vector<CollisionResult<double>>& results) {
  results.assign(queries.size(), CollisionResult<double>());
  if (queries.empty()) return 0;
  if (request.num_max_contacts == 0) return 0;
  fclx_context* ctx = detail::engine();
  if (!ctx) return static_cast<std::size_t>(-1);
  const size_t n = queries.size();
  std::vector<int32_t> qm(2 * n);
  std::vector<double> tf(24 * n);
  for (size_t i = 0; i < n; ++i) {
    qm[2 * i] = detail::mesh_id_of(queries[i].o1, BV_OBBRSS);
    qm[2 * i + 1] = detail::mesh_id_of(queries[i].o2, BV_OBBRSS);
    if (qm[2 * i] < 0 || qm[2 * i + 1] < 0) {
      detail::unsupported("collision", queries[i].o1, queries[i].o2);
      return static_cast<std::size_t>(-1);
    }
    detail::to_pose(queries[i].tf1, &tf[12 * i]);
    detail::to_pose(queries[i].tf2, &tf[12 * (n + i)]);
  }
  const int32_t cap = request.num_max_contacts > static_cast<std::size_t>(INT_MAX)
                          ? INT_MAX
                          : static_cast<int32_t>(request.num_max_contacts);
  int32_t *d_qm = nullptr, *d_counts = nullptr, *d_pairs = nullptr;
  double* d_tf = nullptr;
  int64_t* d_num = nullptr;
  long long capacity = std::max<long long>(1024, 64LL * n);
  std::size_t ret = static_cast<std::size_t>(-1);
  hipStream_t s = nullptr;
  do {
    if (hipMalloc(&d_qm, sizeof(int32_t) * 2 * n) != hipSuccess) break;
    if (hipMalloc(&d_tf, sizeof(double) * 24 * n) != hipSuccess) break;
    if (hipMalloc(&d_counts, sizeof(int32_t) * n) != hipSuccess) break;
    if (hipMalloc(&d_num, sizeof(int64_t)) != hipSuccess) break;
    if (hipMemcpy(d_qm, qm.data(), sizeof(int32_t) * 2 * n, hipMemcpyHostToDevice) != hipSuccess)
      break;
    if (hipMemcpy(d_tf, tf.data(), sizeof(double) * 24 * n, hipMemcpyHostToDevice) != hipSuccess)
      break;
    int rc = FCLX_ERR_CAPACITY;
    int64_t produced = 0;
    for (int attempt = 0; attempt < 3 && rc == FCLX_ERR_CAPACITY; ++attempt) {
      (void)hipFree(d_pairs);
      d_pairs = nullptr;
      if (hipMalloc(&d_pairs, sizeof(int32_t) * 3 * capacity) != hipSuccess) break;
      fclx_collide_batch_args a{};
      a.num_queries = static_cast<int32_t>(n);
      a.query_meshes = d_qm;
      a.tf1 = d_tf;
      a.tf2 = d_tf + 12 * n;
      a.num_max_contacts = cap;
      a.engine = FCLX_ENGINE_FAST;
      a.counts = d_counts;
      a.pairs = d_pairs;
      a.pair_capacity = capacity;
      a.num_pairs = d_num;
      rc = fclx_collide_batch(ctx, &a, s);
      if (rc == FCLX_OK) rc = fclx_sync(ctx, s);
      if (hipMemcpy(&produced, d_num, sizeof(int64_t), hipMemcpyDeviceToHost) != hipSuccess) {
        rc = FCLX_ERR_DEVICE;
        break;
      }
      if (rc == FCLX_ERR_CAPACITY) capacity = produced + 1024;
    }
    if (rc != FCLX_OK) break;
    std::vector<int32_t> rec(3 * static_cast<size_t>(produced));
    if (produced > 0 && hipMemcpy(rec.data(), d_pairs, sizeof(int32_t) * rec.size(),
                                  hipMemcpyDeviceToHost) != hipSuccess)
      break;
    for (int64_t i = 0; i < produced; ++i) {
      const size_t q = static_cast<size_t>(rec[3 * i]);
      results[q].addContact(
          Contact<double>(queries[q].o1, queries[q].o2, rec[3 * i + 1], rec[3 * i + 2]));
    }
    ret = n;
  } while (false);
  (void)hipFree(d_qm);
  (void)hipFree(d_tf);
  (void)hipFree(d_counts);
  (void)hipFree(d_num);
  (void)hipFree(d_pairs);
  return ret;
}

namespace detail {

bool broadphase_pairs(const std::vector<AABB<double>>& boxes, const std::vector<int32_t>* groups,
                      std::vector<std::pair<int32_t, int32_t>>& out) {
  out.clear();
  const size_t n = boxes.size();
  if (n < 2) return true;
  fclx_context* ctx = engine();
  if (!ctx) return false;
  std::vector<double> b(6 * n);
  for (size_t i = 0; i < n; ++i)
    for (int k = 0; k < 3; ++k) {
      b[6 * i + k] = boxes[i].min_[k];
      b[6 * i + 3 + k] = boxes[i].max_[k];
    }
  double* d_b = nullptr;
  int32_t *d_g = nullptr, *d_p = nullptr;
  int64_t* d_n = nullptr;
  int64_t cap = std::max<int64_t>(1024, 8 * static_cast<int64_t>(n)), found = 0;
  bool ok = hipMalloc(&d_b, b.size() * sizeof(double)) == hipSuccess &&
            hipMalloc(&d_n, sizeof(int64_t)) == hipSuccess &&
            hipMemcpy(d_b, b.data(), b.size() * sizeof(double), hipMemcpyHostToDevice) ==
                hipSuccess;
  if (ok && groups)
    ok = hipMalloc(&d_g, n * sizeof(int32_t)) == hipSuccess &&
         hipMemcpy(d_g, groups->data(), n * sizeof(int32_t), hipMemcpyHostToDevice) ==
             hipSuccess;
  for (int attempt = 0; ok && attempt < 2; ++attempt) {
    (void)hipFree(d_p);
    d_p = nullptr;
    ok = hipMalloc(&d_p, 2 * cap * sizeof(int32_t)) == hipSuccess &&
         fclx_broadphase_pairs(ctx, static_cast<int32_t>(n), d_b, d_g, d_p, cap, d_n, nullptr) ==
             FCLX_OK &&
         hipStreamSynchronize(nullptr) == hipSuccess &&
         hipMemcpy(&found, d_n, sizeof(int64_t), hipMemcpyDeviceToHost) == hipSuccess;
    if (!ok || found <= cap) break;
    cap = found;  // room for every pair, then once more
  }
  if (ok && found > cap) ok = false;
  if (ok && found > 0) {
    std::vector<int32_t> p(2 * static_cast<size_t>(found));
    ok = hipMemcpy(p.data(), d_p, p.size() * sizeof(int32_t), hipMemcpyDeviceToHost) ==
         hipSuccess;
    for (int64_t k = 0; ok && k < found; ++k) out.emplace_back(p[2 * k], p[2 * k + 1]);
  }
  (void)hipFree(d_b);
  (void)hipFree(d_g);
  (void)hipFree(d_p);
  (void)hipFree(d_n);
  if (!ok) std::cerr << "fcl: device broadphase failed: " << fclx_last_error(ctx) << "\n";
  return ok;
}

void default_collide_pairs(const std::vector<ObjectPair>& pairs, DefaultCollisionData<double>* data) {
  if (data->done || pairs.empty()) return;
  const CollisionRequest<double>& req = data->request;
  // the mesh pairs' narrow phase in one batch (each pair as if its result
  // were empty, capped at num_max_contacts)
  std::vector<CollideQuery> qs;
  std::vector<int> batched(pairs.size(), -1);
  if (!req.enable_contact && !req.enable_cost && req.num_max_contacts > 0) {
    for (size_t k = 0; k < pairs.size(); ++k) {
      const auto* g1 = pairs[k].first->collisionGeometry().get();
      const auto* g2 = pairs[k].second->collisionGeometry().get();
      if (mesh_id_of(g1, BV_OBBRSS) >= 0 && mesh_id_of(g2, BV_OBBRSS) >= 0) {
        batched[k] = static_cast<int>(qs.size());
        qs.push_back({g1, pairs[k].first->getTransform(), g2, pairs[k].second->getTransform()});
      }
    }
  }
  std::vector<CollisionResult<double>> res;
  if (!qs.empty() && collideBatch(qs, req, res) != qs.size()) {
    res.clear();
    std::fill(batched.begin(), batched.end(), -1);
  }
  // fold in pair order, exactly as DefaultCollisionFunction pair by pair:
  // a pair adds its contacts until the shared result holds num_max_contacts
  for (size_t k = 0; k < pairs.size() && !data->done; ++k) {
    const int q = batched[k];
    const std::size_t have = data->result.numContacts();
    if (q >= 0 && have + res[q].numContacts() < req.num_max_contacts) {
      // the whole list fits: the same contacts as collide() would add
      for (const auto& c : res[q].contacts) data->result.addContact(c);
    } else if (q < 0 || res[q].numContacts() > 0) {
      // a list the cap cuts (or geometry outside the batch): collide()
      // itself, which adds the reference's depth-first prefix
      DefaultCollisionFunction(pairs[k].first, pairs[k].second, data);
      continue;
    }
    if (data->result.isCollision() && data->result.numContacts() >= req.num_max_contacts)
      data->done = true;
  }
}

}  // namespace detail

}  // namespace fcl
