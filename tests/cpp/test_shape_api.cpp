// fcl::collide between primitives through the drop-in C++ API, used the way
// FCL's shape tests use it (test_fcl_geometric_shapes.cpp: Box / Sphere /
// Capsule / Cylinder / Convex objects, CollisionRequest with
// enable_contact and gjk_solver_type, contacts read back from the result),
// checked against the oracle's restatement (fclo_shape_collide).
//
// Bar: contact counts equal; every contact's position, normal and depth
// bit-identical; o1/o2 are the geometries passed in and b1/b2 are NONE.
#include <cmath>
#include <cstdio>
#include <memory>
#include <random>
#include <vector>

#include "fcl/geometry/shape/box.h"
#include "fcl/geometry/shape/capsule.h"
#include "fcl/geometry/shape/convex.h"
#include "fcl/geometry/shape/cylinder.h"
#include "fcl/geometry/shape/sphere.h"
#include "fcl/geometry/bvh/BVH_model.h"
#include "fcl/narrowphase/collision.h"
#include "fclo.h"
#include "fclx_synth.h"

using namespace fcl;

static int failures = 0;
#define CHECK(cond)                                                          \
  do {                                                                       \
    if (!(cond)) {                                                           \
      std::printf("FAIL %s:%d: %s\n", __FILE__, __LINE__, #cond);            \
      ++failures;                                                            \
    }                                                                        \
  } while (0)

struct Shape {
  std::shared_ptr<CollisionGeometry<double>> geom;
  fclo_shape o;
  std::vector<double> verts;  // oracle copy of a convex's vertices
};

static Shape make_shape(int type, std::mt19937_64& rng) {
  std::uniform_real_distribution<double> u(std::log(0.1), 0.0);
  auto d = [&] { return std::exp(u(rng)); };
  Shape s;
  s.o = fclo_shape{};
  s.o.type = type;
  switch (type) {
    case FCLO_SHAPE_BOX: {
      const double x = d(), y = d(), z = d();
      s.geom = std::make_shared<Boxd>(x, y, z);
      s.o.dims[0] = x;
      s.o.dims[1] = y;
      s.o.dims[2] = z;
      break;
    }
    case FCLO_SHAPE_SPHERE: {
      const double r = d();
      s.geom = std::make_shared<Sphered>(r);
      s.o.dims[0] = r;
      break;
    }
    case FCLO_SHAPE_CAPSULE: {
      const double r = d(), lz = d();
      s.geom = std::make_shared<Capsuled>(r, lz);
      s.o.dims[0] = r;
      s.o.dims[1] = lz;
      break;
    }
    case FCLO_SHAPE_CYLINDER: {
      const double r = d(), lz = d();
      s.geom = std::make_shared<Cylinderd>(r, lz);
      s.o.dims[0] = r;
      s.o.dims[1] = lz;
      break;
    }
    default: {
      std::normal_distribution<double> g(0.0, 1.0);
      const double sc = 0.5 * d();
      auto v = std::make_shared<std::vector<Vector3d>>();
      for (int i = 0; i < 32; ++i) {
        const double x = g(rng) * sc, y = g(rng) * sc, z = g(rng) * sc;
        v->push_back(Vector3d(x, y, z));
        s.verts.insert(s.verts.end(), {x, y, z});
      }
      s.geom = std::make_shared<Convexd>(v, 0, std::make_shared<std::vector<int>>());
      s.o.vertices = s.verts.data();
      s.o.num_vertices = 32;
      break;
    }
  }
  return s;
}

static Transform3d random_pose(std::mt19937_64& rng, double h) {
  std::normal_distribution<double> g(0.0, 1.0);
  std::uniform_real_distribution<double> u(-h, h);
  double q[4];
  double n = 0;
  for (double& x : q) {
    x = g(rng);
    n += x * x;
  }
  n = std::sqrt(n);
  const double w = q[0] / n, x = q[1] / n, y = q[2] / n, z = q[3] / n;
  Transform3d tf = Transform3d::Identity();
  const double R[9] = {1 - 2 * (y * y + z * z), 2 * (x * y - w * z), 2 * (x * z + w * y),
                       2 * (x * y + w * z), 1 - 2 * (x * x + z * z), 2 * (y * z - w * x),
                       2 * (x * z - w * y), 2 * (y * z + w * x), 1 - 2 * (x * x + y * y)};
  for (int i = 0; i < 9; ++i) tf.R.m[i] = R[i];
  tf.t = Vector3d(u(rng), u(rng), u(rng));
  return tf;
}

static void to_pose(const Transform3d& tf, double out[12]) {
  for (int i = 0; i < 9; ++i) out[i] = tf.R.m[i];
  for (int i = 0; i < 3; ++i) out[9 + i] = tf.t[i];
}

int main() {
  std::mt19937_64 rng(0x5A9E);
  long hits = 0, contacts = 0, pairs = 0;
  for (int t1 = 0; t1 < 5; ++t1)
    for (int t2 = 0; t2 < 5; ++t2)
      for (int rep = 0; rep < 40; ++rep) {
        Shape a = make_shape(t1, rng), b = make_shape(t2, rng);
        const Transform3d tf1 = random_pose(rng, 0.4), tf2 = random_pose(rng, 0.4);
        double p1[12], p2[12];
        to_pose(tf1, p1);
        to_pose(tf2, p2);
        for (int cap : {1, 4}) {
          CollisionRequest<double> req(cap, true);
          req.gjk_solver_type = GST_INDEP;
          CollisionResult<double> res;
          const std::size_t n = collide(a.geom.get(), tf1, b.geom.get(), tf2, req, res);
          fclo_contact oc[8];
          const int on = fclo_shape_collide(&a.o, p1, &b.o, p2, cap, 1, 1e-6, oc);
          ++pairs;
          CHECK(n == static_cast<std::size_t>(on) && res.numContacts() == n);
          for (std::size_t i = 0; i < n && static_cast<int>(i) < on; ++i) {
            const Contact<double>& c = res.getContact(i);
            CHECK(c.o1 == a.geom.get() && c.o2 == b.geom.get());
            CHECK(c.b1 == Contact<double>::NONE && c.b2 == Contact<double>::NONE);
            for (int k = 0; k < 3; ++k) {
              CHECK(c.pos[k] == oc[i].pos[k]);
              CHECK(c.normal[k] == oc[i].normal[k]);
            }
            CHECK(c.penetration_depth == oc[i].depth);
          }
          hits += n > 0;
          contacts += n;
          // without enable_contact: one contact, no geometry
          CollisionResult<double> bare;
          const std::size_t m =
              collide(a.geom.get(), tf1, b.geom.get(), tf2, CollisionRequest<double>(cap), bare);
          CHECK(m == (on > 0 ? 1u : 0u));
          if (m) CHECK(bare.getContact(0).penetration_depth == 0);
        }
        // CollisionObject overload, and a result that is already satisfied
        CollisionObject<double> oa(a.geom, tf1), ob(b.geom, tf2);
        CollisionRequest<double> req(1, true);
        req.gjk_solver_type = GST_INDEP;
        CollisionResult<double> r1;
        const std::size_t n1 = collide(&oa, &ob, req, r1);
        CollisionResult<double> r2 = r1;
        CHECK(collide(&oa, &ob, req, r2) == n1);
        if (n1) CHECK(r2.numContacts() == n1);
      }
  CHECK(hits > 300);
  std::printf("%ld shape pairs (x2 caps), %ld colliding, %ld contacts\n", pairs, hits, contacts);
  // a closed-form pair with the default request (the libccd solver answers
  // the five special pairs with the same closed forms)
  {
    Sphered s1(0.5), s2(0.5);
    Transform3d tf = Transform3d::Identity();
    tf.t = Vector3d(0.75, 0, 0);
    CollisionRequest<double> req(1, true);
    CollisionResult<double> res;
    CHECK(collide(&s1, Transform3d::Identity(), &s2, tf, req, res) == 1);
    CHECK(res.getContact(0).penetration_depth == 0.25);
    CHECK(res.getContact(0).normal[0] == 1 && res.getContact(0).pos[0] == 0.375);
  }
  // a mesh (BVHModel<OBBRSS>) against each primitive kind: contacts, their
  // triangles (b1) and geometry equal to the oracle's mesh-shape collide
  {
    int32_t nv = 0, nt = 0;
    fclx_synth_icosphere_size(8, &nv, &nt);
    std::vector<double> V(3 * static_cast<size_t>(nv));
    std::vector<int32_t> T(3 * static_cast<size_t>(nt));
    CHECK(fclx_synth_icosphere(8, 0.15, 0xF0, V.data(), T.data()) == 0);
    std::vector<Vector3d> ps;
    std::vector<Triangle> ts;
    for (int32_t i = 0; i < nv; ++i) ps.emplace_back(V[3 * i], V[3 * i + 1], V[3 * i + 2]);
    for (int32_t i = 0; i < nt; ++i) ts.emplace_back(T[3 * i], T[3 * i + 1], T[3 * i + 2]);
    auto mesh = std::make_shared<BVHModel<OBBRSSd>>();
    CHECK(mesh->beginModel() == BVH_OK);
    CHECK(mesh->addSubModel(ps, ts) == BVH_OK);
    CHECK(mesh->endModel() == BVH_OK);
    fclo_model om;
    CHECK(fclo_model_build(&om, FCLO_BV_OBBRSS, V.data(), nv, T.data(), nt) == 0);
    long ms_queries = 0, ms_hits = 0;
    for (int type = 0; type < 5; ++type)
      for (int rep = 0; rep < 30; ++rep) {
        Shape b = make_shape(type, rng);
        const Transform3d tf1 = random_pose(rng, 0.2), tf2 = random_pose(rng, 1.0);
        double p1[12], p2[12];
        to_pose(tf1, p1);
        to_pose(tf2, p2);
        for (int cap : {1, 5}) {
          for (bool geometry : {false, true}) {
            CollisionRequest<double> req(cap, geometry);
            req.gjk_solver_type = GST_INDEP;
            CollisionResult<double> res;
            const std::size_t n = collide(mesh.get(), tf1, b.geom.get(), tf2, req, res);
            int prims[8];
            fclo_contact oc[8];
            fclo_collide_stats st;
            const int on = fclo_mesh_shape_collide(&om, p1, &b.o, p2, cap, geometry ? 1 : 0,
                                                   1e-6, prims, oc, &st);
            ++ms_queries;
            CHECK(n == static_cast<std::size_t>(on) && res.numContacts() == n);
            for (std::size_t i = 0; i < n && static_cast<int>(i) < on; ++i) {
              const Contact<double>& c = res.getContact(i);
              CHECK(c.o1 == mesh.get() && c.o2 == b.geom.get());
              CHECK(c.b1 == prims[i] && c.b2 == Contact<double>::NONE);
              if (!geometry) continue;
              for (int k = 0; k < 3; ++k) {
                CHECK(c.pos[k] == oc[i].pos[k]);
                CHECK(c.normal[k] == oc[i].normal[k]);
              }
              CHECK(c.penetration_depth == oc[i].depth);
            }
            ms_hits += n > 0;
            // shape first: the same contacts, the mesh still named first
            CollisionResult<double> sw;
            CHECK(collide(b.geom.get(), tf2, mesh.get(), tf1, req, sw) == n);
            for (std::size_t i = 0; i < n && i < sw.numContacts(); ++i) {
              const Contact<double>& c = sw.getContact(i);
              CHECK(c.o1 == mesh.get() && c.b1 == prims[i]);
              if (geometry) CHECK(c.penetration_depth == oc[i].depth);
            }
          }
        }
      }
    fclo_model_free(&om);
    CHECK(ms_hits > 40);
    std::printf("%ld mesh-shape queries, %ld colliding\n", ms_queries, ms_hits);
  }
  if (failures == 0) std::printf("PASS\n");
  return failures == 0 ? 0 : 1;
}
