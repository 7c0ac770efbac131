// GPU test of the C++ drop-in API (fcl::collide, fcl::distance,
// fcl::BVHModel, fcl::collideBatch) in the shape of FCL's own mesh tests
// (test/test_fcl_collision.cpp:313-888 mesh_mesh, test/test_fcl_distance.cpp:
// 86-304), checked against the CPU oracle on the same inputs.
// Usage: test_mesh_api <dir with env.obj and rob.obj>
#include <algorithm>
#include <climits>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <random>
#include <sstream>
#include <string>
#include <vector>

#include "fcl/narrowphase/collision.h"
#include "fcl/narrowphase/distance.h"
#include "fclo.h"

using namespace fcl;

static int failures = 0;
#define CHECK(cond)                                                          \
  do {                                                                       \
    if (!(cond)) {                                                           \
      std::printf("FAIL %s:%d: %s\n", __FILE__, __LINE__, #cond);            \
      ++failures;                                                            \
    }                                                                        \
  } while (0)

struct Mesh {
  std::vector<Vector3d> v;
  std::vector<Triangle> t;
};

static Mesh load_obj(const std::string& path) {
  Mesh m;
  std::ifstream in(path);
  std::string line;
  while (std::getline(in, line)) {
    std::istringstream ss(line);
    std::string tok;
    if (!(ss >> tok)) continue;
    if (tok == "v") {
      double x, y, z;
      ss >> x >> y >> z;
      m.v.emplace_back(x, y, z);
    } else if (tok == "f") {
      std::vector<long> idx;
      std::string w;
      while (ss >> w) idx.push_back(std::stol(w.substr(0, w.find('/'))) - 1);
      for (size_t k = 0; k + 2 < idx.size(); ++k) m.t.emplace_back(idx[0], idx[k + 1], idx[k + 2]);
    }
  }
  return m;
}

// poses like test::generateRandomTransforms (test_fcl_utility.h:348-373),
// from an explicit seed
static std::vector<Transform3d> random_transforms(int n, unsigned seed) {
  std::mt19937_64 rng(seed);
  std::uniform_real_distribution<double> U(0, 1);
  std::vector<Transform3d> out(n);
  for (auto& tf : out) {
    const double x = -3000 + 6000 * U(rng), y = -3000 + 6000 * U(rng), z = 3000 * U(rng);
    const double a = 2 * M_PI * U(rng), b = 2 * M_PI * U(rng), c = 2 * M_PI * U(rng);
    const double c1 = std::cos(a), c2 = std::cos(b), c3 = std::cos(c);
    const double s1 = std::sin(a), s2 = std::sin(b), s3 = std::sin(c);
    const double R[9] = {c1 * c2, -c2 * s1, s2,
                         c3 * s1 + c1 * s2 * s3, c1 * c3 - s1 * s2 * s3, -c2 * s3,
                         s1 * s3 - c1 * c3 * s2, c3 * s1 * s2 + c1 * s3, c2 * c3};
    for (int i = 0; i < 9; ++i) tf.R.m[i] = R[i];
    tf.t = Vector3d(x, y, z);
  }
  return out;
}

static void to_pose(const Transform3d& tf, double out[12]) {
  for (int i = 0; i < 9; ++i) out[i] = tf.R.m[i];
  for (int i = 0; i < 3; ++i) out[9 + i] = tf.t[i];
}

struct Oracle {
  fclo_model m;
  Oracle(const Mesh& mesh, int bv) {
    std::vector<double> v(3 * mesh.v.size());
    std::vector<int> t(3 * mesh.t.size());
    for (size_t i = 0; i < mesh.v.size(); ++i)
      for (int k = 0; k < 3; ++k) v[3 * i + k] = mesh.v[i][k];
    for (size_t i = 0; i < mesh.t.size(); ++i)
      for (int k = 0; k < 3; ++k) t[3 * i + k] = static_cast<int>(mesh.t[i][k]);
    fclo_model_build(&m, bv, v.data(), static_cast<int>(mesh.v.size()), t.data(),
                     static_cast<int>(mesh.t.size()));
  }
  ~Oracle() { fclo_model_free(&m); }
};

template <typename BV>
static void build(BVHModel<BV>& m, const Mesh& mesh) {
  CHECK(m.beginModel() == BVH_OK);
  CHECK(m.addSubModel(mesh.v, mesh.t) == BVH_OK);
  CHECK(m.endModel() == BVH_OK);
}

int main(int argc, char** argv) {
  const std::string dir = argc > 1 ? argv[1] : ".";
  const Mesh env = load_obj(dir + "/env.obj"), rob = load_obj(dir + "/rob.obj");
  CHECK(env.t.size() == 2180 && rob.t.size() == 216);
  BVHModel<OBBRSSd> m1, m2;
  build(m1, env);
  build(m2, rob);
  CHECK(m1.getNumBVs() == 2 * 2180 - 1);
  Oracle o1(env, FCLO_BV_OBBRSS), o2(rob, FCLO_BV_OBBRSS);
  const auto tfs = random_transforms(10, 2024);
  const Transform3d I = Transform3d::Identity();

  // mesh_mesh: exhaustive contacts, every pair in the reference's order
  int total = 0;
  for (const auto& tf : tfs) {
    CollisionRequest<double> req(INT_MAX, false);
    CollisionResult<double> res;
    const std::size_t n = collide(&m1, tf, &m2, I, req, res);
    double p1[12], p2[12];
    to_pose(tf, p1);
    to_pose(I, p2);
    std::vector<int> pairs(2 * 100000);
    fclo_collide_stats st;
    fclo_collide(&o1.m, p1, &o2.m, p2, 100000, pairs.data(), &st);
    CHECK(n == static_cast<std::size_t>(st.num_contacts));
    for (std::size_t i = 0; i < n && i < static_cast<std::size_t>(st.num_contacts); ++i) {
      CHECK(res.getContact(i).b1 == pairs[2 * i] && res.getContact(i).b2 == pairs[2 * i + 1]);
      CHECK(res.getContact(i).o1 == &m1 && res.getContact(i).o2 == &m2);
    }
    total += static_cast<int>(n);

    // default request: the first contact in depth-first order
    CollisionResult<double> first;
    collide(&m1, tf, &m2, I, CollisionRequest<double>(), first);
    CHECK(first.numContacts() == (n > 0 ? 1u : 0u));
    if (n > 0) CHECK(first.getContact(0).b1 == pairs[0] && first.getContact(0).b2 == pairs[1]);

    // a result that already holds contacts: appended to, capped in total
    if (n >= 3) {
      CollisionResult<double> pre;
      pre.addContact(Contact<double>(&m1, &m2, -5, -5));
      collide(&m1, tf, &m2, I, CollisionRequest<double>(3), pre);
      CHECK(pre.numContacts() == 3);
      CHECK(pre.getContact(1).b1 == pairs[0] && pre.getContact(2).b1 == pairs[2]);
      // and a satisfied result short-circuits
      CHECK(collide(&m1, tf, &m2, I, CollisionRequest<double>(3), pre) == 3);
    }
    // the same through CollisionObject
    auto g1 = std::make_shared<BVHModel<OBBRSSd>>(m1);
    CollisionObject<double> a(g1, tf), b(std::make_shared<BVHModel<OBBRSSd>>(m2), I);
    CollisionResult<double> r2;
    CHECK(collide(&a, &b, req, r2) == n);
  }
  CHECK(total > 0);
  std::printf("mesh_mesh: %d contacts over %zu poses\n", total, tfs.size());

  // batched entry point: same sets as one-by-one
  {
    std::vector<CollideQuery> qs;
    for (const auto& tf : tfs) qs.push_back({&m1, tf, &m2, I});
    std::vector<CollisionResult<double>> rs;
    CHECK(collideBatch(qs, CollisionRequest<double>(100000), rs) == qs.size());
    for (size_t q = 0; q < qs.size(); ++q) {
      CollisionResult<double> one;
      collide(&m1, tfs[q], &m2, I, CollisionRequest<double>(100000), one);
      auto A = rs[q].contacts, B = one.contacts;
      std::sort(A.begin(), A.end());
      std::sort(B.begin(), B.end());
      CHECK(A.size() == B.size());
      for (size_t i = 0; i < A.size() && i < B.size(); ++i) CHECK(A[i] == B[i]);
    }
  }

  // distance: RSS and OBBRSS hierarchies, bit-identical to the oracle
  {
    BVHModel<RSSd> r1, r2;
    build(r1, env);
    build(r2, rob);
    Oracle q1(env, FCLO_BV_RSS), q2(rob, FCLO_BV_RSS);
    for (const auto& tf : random_transforms(10, 2025)) {
      DistanceRequest<double> req(true);
      DistanceResult<double> dr, doo;
      const double d = distance(&r1, tf, &r2, I, req, dr);
      const double d2 = distance(&m1, tf, &m2, I, req, doo);
      double p1[12], p2[12];
      to_pose(tf, p1);
      to_pose(I, p2);
      fclo_distance_out od;
      fclo_distance(&q1.m, p1, &q2.m, p2, 1, &od);
      CHECK(d == od.min_distance && dr.min_distance == d);
      CHECK(d2 == d);  // OBBRSS traverses differently, same minimum
      CHECK(dr.b1 == od.b1 && dr.b2 == od.b2);
      for (int k = 0; k < 3; ++k)
        CHECK(dr.nearest_points[0][k] == od.p1[k] && dr.nearest_points[1][k] == od.p2[k]);
    }
  }

  // enable_contact: contact points, normals and depths as the oracle computes them
  {
    int checked = 0;
    for (const auto& tf : tfs) {
      for (std::size_t cap : {std::size_t(3), std::size_t(100000)}) {
        CollisionRequest<double> req(cap, true);
        CollisionResult<double> res;
        const std::size_t n = collide(&m1, tf, &m2, I, req, res);
        double p1[12], p2[12];
        to_pose(tf, p1);
        to_pose(I, p2);
        const int store = 20000;
        std::vector<int> pairs(2 * store);
        std::vector<double> pos(3 * store), nrm(3 * store), dep(store);
        fclo_collide_stats st;
        fclo_collide_contact(&o1.m, p1, &o2.m, p2, static_cast<int>(cap), store, pairs.data(),
                             pos.data(), nrm.data(), dep.data(), &st);
        CHECK(n == static_cast<std::size_t>(st.num_contacts));
        for (std::size_t i = 0; i < n && i < static_cast<std::size_t>(store); ++i) {
          const Contact<double>& c = res.getContact(i);
          CHECK(c.b1 == pairs[2 * i] && c.b2 == pairs[2 * i + 1]);
          for (int k = 0; k < 3; ++k)
            CHECK(c.pos[k] == pos[3 * i + k] && c.normal[k] == nrm[3 * i + k]);
          CHECK(c.penetration_depth == dep[i]);
          ++checked;
        }
      }
    }
    CHECK(checked > 0);
    std::printf("enable_contact: %d contacts checked\n", checked);
  }

  // error conventions of collide<S,Solver> / distance<Solver>
  {
    CollisionResult<double> res;
    CHECK(collide(&m1, I, &m2, I, CollisionRequest<double>(0), res) == 0);
    CollisionRequest<double> bad;
    bad.gjk_solver_type = static_cast<GJKSolverType>(7);
    CHECK(collide(&m1, I, &m2, I, bad, res) == static_cast<std::size_t>(-1));
    BVHModel<RSSd> r1;
    build(r1, rob);
    CHECK(collide(&r1, I, &r1, I, CollisionRequest<double>(), res) == 0);  // unsupported cell
    DistanceResult<double> dres;
    CHECK(distance(&m1, I, &r1, I, DistanceRequest<double>(), dres) ==
          std::numeric_limits<double>::max());
    BVHModel<OBBRSSd> empty;
    CHECK(empty.beginModel() == BVH_OK);
    CHECK(empty.endModel() == BVH_ERR_BUILD_EMPTY_MODEL);
    CHECK(empty.addVertex(Vector3d()) == BVH_OK);
    BVHModel<OBBRSSd> seq;
    CHECK(seq.endModel() == BVH_ERR_BUILD_OUT_OF_SEQUENCE);
  }

  std::printf(failures ? "FAILED (%d)\n" : "PASS\n", failures);
  return failures ? 1 : 0;
}
