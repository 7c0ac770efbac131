// GPU test of the broadphase manager drop-in (fcl::DynamicAABBTreeCollisionManager
// over the device sort-and-sweep) in the shape of FCL's own broadphase test
// (test/test_fcl_broadphase_collision_2.cpp:186-330, broad_phase_collision_test):
// an environment of meshes and primitives, self collision and query collision
// with DefaultCollisionFunction, exhaustive (num_max_contacts = 100000: the
// contact count must equal the brute-force manager's) and first contact (the
// flag must); plus query-vs-manager and manager-vs-manager collision, a
// user callback that sees every candidate pair, and distance with
// DefaultDistanceFunction against the minimum over all pairs.
#include <cmath>
#include <cstdio>
#include <memory>
#include <random>
#include <set>
#include <vector>

#include "fcl/broadphase/broadphase_dynamic_AABB_tree.h"
#include "fcl/broadphase/default_broadphase_callbacks.h"

using namespace fcl;

static int failures = 0;
#define CHECK(cond)                                                          \
  do {                                                                       \
    if (!(cond)) {                                                           \
      std::printf("FAIL %s:%d: %s\n", __FILE__, __LINE__, #cond);            \
      ++failures;                                                            \
    }                                                                        \
  } while (0)

static Matrix3d euler(double a, double b, double c) {
  // test_fcl_utility.h:312-326 (eulerToMatrix)
  const double c1 = std::cos(a), c2 = std::cos(b), c3 = std::cos(c);
  const double s1 = std::sin(a), s2 = std::sin(b), s3 = std::sin(c);
  Matrix3d R;
  R(0, 0) = c1 * c2;
  R(0, 1) = -c2 * s1;
  R(0, 2) = s2;
  R(1, 0) = c3 * s1 + c1 * s2 * s3;
  R(1, 1) = c1 * c3 - s1 * s2 * s3;
  R(1, 2) = -c2 * s3;
  R(2, 0) = s1 * s3 - c1 * c3 * s2;
  R(2, 1) = c3 * s1 * s2 + c1 * s3;
  R(2, 2) = c2 * c3;
  return R;
}

// a UV sphere of radius r as triangles
static std::shared_ptr<BVHModel<OBBRSSd>> sphere_mesh(double r, int seg) {
  auto m = std::make_shared<BVHModel<OBBRSSd>>();
  std::vector<Vector3d> v;
  std::vector<Triangle> t;
  const double pi = std::acos(-1.0);
  for (int i = 0; i <= seg; ++i)
    for (int j = 0; j < 2 * seg; ++j) {
      const double th = pi * i / seg, ph = pi * j / seg;
      v.emplace_back(r * std::sin(th) * std::cos(ph), r * std::sin(th) * std::sin(ph),
                     r * std::cos(th));
    }
  for (int i = 0; i < seg; ++i)
    for (int j = 0; j < 2 * seg; ++j) {
      const int a = i * 2 * seg + j, b = i * 2 * seg + (j + 1) % (2 * seg);
      const int c = a + 2 * seg, d = b + 2 * seg;
      t.emplace_back(a, b, d);
      t.emplace_back(a, d, c);
    }
  m->beginModel();
  m->addSubModel(v, t);
  m->endModel();
  return m;
}

static std::shared_ptr<BVHModel<OBBRSSd>> box_mesh(double x, double y, double z) {
  auto m = std::make_shared<BVHModel<OBBRSSd>>();
  std::vector<Vector3d> v;
  for (int i = 0; i < 8; ++i)
    v.emplace_back((i & 1 ? 0.5 : -0.5) * x, (i & 2 ? 0.5 : -0.5) * y, (i & 4 ? 0.5 : -0.5) * z);
  const int f[12][3] = {{0, 1, 3}, {0, 3, 2}, {4, 6, 7}, {4, 7, 5}, {0, 4, 5}, {0, 5, 1},
                        {2, 3, 7}, {2, 7, 6}, {0, 2, 6}, {0, 6, 4}, {1, 5, 7}, {1, 7, 3}};
  std::vector<Triangle> t;
  for (auto& k : f) t.emplace_back(k[0], k[1], k[2]);
  m->beginModel();
  m->addSubModel(v, t);
  m->endModel();
  return m;
}

// an environment in the spirit of generateEnvironments /
// generateEnvironmentsMesh (test_fcl_utility.h): meshes and primitives at
// random poses in a cube of the given scale
static std::vector<std::unique_ptr<CollisionObjectd>> environment(int n, double scale,
                                                                 unsigned seed, bool meshes_only) {
  std::mt19937_64 rng(seed);
  std::uniform_real_distribution<double> u(0, 1);
  std::vector<std::unique_ptr<CollisionObjectd>> env;
  for (int i = 0; i < n; ++i) {
    Transform3d tf;
    tf.linear() = euler(u(rng) * 6.283, u(rng) * 6.283, u(rng) * 6.283);
    tf.translation() = Vector3d((2 * u(rng) - 1) * scale, (2 * u(rng) - 1) * scale,
                                (2 * u(rng) - 1) * scale);
    std::shared_ptr<CollisionGeometryd> g;
    const int kind = meshes_only ? i % 2 : i % 4;
    const double s = 5 + 10 * u(rng);
    if (kind == 0) g = box_mesh(s, s * (0.5 + u(rng)), s * (0.5 + u(rng)));
    else if (kind == 1) g = sphere_mesh(s * 0.6, 10);
    else if (kind == 2) g = std::make_shared<Boxd>(s, s, s * 0.7);
    else g = std::make_shared<Sphered>(s * 0.5);
    env.emplace_back(new CollisionObjectd(g, tf));
  }
  return env;
}

static std::vector<CollisionObjectd*> raw(const std::vector<std::unique_ptr<CollisionObjectd>>& v) {
  std::vector<CollisionObjectd*> out;
  for (auto& p : v) out.push_back(p.get());
  return out;
}

// NaiveCollisionManager's order (broadphase_bruteforce-inl.h): every pair
// i < j, the boxes tested first
static void brute_self(const std::vector<CollisionObjectd*>& objs, DefaultCollisionDatad* d) {
  for (size_t i = 0; i < objs.size(); ++i)
    for (size_t j = i + 1; j < objs.size(); ++j)
      if (objs[i]->getAABB().overlap(objs[j]->getAABB()))
        if (DefaultCollisionFunction(objs[i], objs[j], d)) return;
}

static std::set<std::pair<const void*, const void*>> g_seen;
static bool record_pair(CollisionObjectd* a, CollisionObjectd* b, void*) {
  g_seen.insert({a < b ? a : b, a < b ? b : a});
  return false;
}

int main() {
  const double scale = 100;
  auto env = environment(400, scale, 7, false);
  auto objs = raw(env);
  DynamicAABBTreeCollisionManagerd mgr;
  mgr.registerObjects(objs);
  mgr.setup();
  CHECK(mgr.size() == objs.size() && !mgr.empty());

  // every candidate pair reaches a user callback exactly once
  g_seen.clear();
  mgr.collide(nullptr, record_pair);
  std::set<std::pair<const void*, const void*>> want;
  for (size_t i = 0; i < objs.size(); ++i)
    for (size_t j = i + 1; j < objs.size(); ++j)
      if (objs[i]->getAABB().overlap(objs[j]->getAABB()))
        want.insert({objs[i] < objs[j] ? objs[i] : objs[j], objs[i] < objs[j] ? objs[j] : objs[i]});
  std::printf("candidate pairs: %zu (expected %zu)\n", g_seen.size(), want.size());
  CHECK(g_seen == want && !want.empty());

  // self collision: exhaustive counts and first-contact flags
  for (size_t cap : {size_t(100000), size_t(1)}) {
    DefaultCollisionDatad a, b;
    a.request.num_max_contacts = b.request.num_max_contacts = cap;
    mgr.collide(&a, DefaultCollisionFunction);
    brute_self(objs, &b);
    std::printf("self collision cap %zu: manager %zu contacts, brute force %zu\n", cap,
                a.result.numContacts(), b.result.numContacts());
    CHECK(a.result.numContacts() == b.result.numContacts());
    CHECK(a.result.isCollision() == b.result.isCollision());
  }
  {
    // a cap the fold has to cut (the pair that crosses it runs collide())
    DefaultCollisionDatad a, b;
    a.request.num_max_contacts = b.request.num_max_contacts = 37;
    mgr.collide(&a, DefaultCollisionFunction);
    brute_self(objs, &b);
    CHECK(a.result.numContacts() == b.result.numContacts());
  }

  // query objects against the manager
  auto queries = environment(30, scale, 11, false);
  for (auto& q : queries) {
    DefaultCollisionDatad a, b;
    a.request.num_max_contacts = b.request.num_max_contacts = 100000;
    mgr.collide(q.get(), &a, DefaultCollisionFunction);
    for (auto* o : objs)
      if (q->getAABB().overlap(o->getAABB())) DefaultCollisionFunction(q.get(), o, &b);
    CHECK(a.result.numContacts() == b.result.numContacts());
  }

  // manager against manager
  {
    DynamicAABBTreeCollisionManagerd other;
    other.registerObjects(raw(queries));
    other.setup();
    DefaultCollisionDatad a, b;
    a.request.num_max_contacts = b.request.num_max_contacts = 100000;
    mgr.collide(&other, &a, DefaultCollisionFunction);
    for (auto* o : objs)
      for (auto& q : queries)
        if (o->getAABB().overlap(q->getAABB())) DefaultCollisionFunction(o, q.get(), &b);
    std::printf("manager vs manager: %zu contacts (brute force %zu)\n", a.result.numContacts(),
                b.result.numContacts());
    CHECK(a.result.numContacts() == b.result.numContacts());
  }

  // moving objects: update() takes the new boxes
  for (size_t i = 0; i < objs.size(); i += 3) {
    Transform3d tf = objs[i]->getTransform();
    tf.translation() = tf.translation() * 0.5;
    objs[i]->setTransform(tf);
    objs[i]->computeAABB();
  }
  mgr.update();
  {
    DefaultCollisionDatad a, b;
    a.request.num_max_contacts = b.request.num_max_contacts = 100000;
    mgr.collide(&a, DefaultCollisionFunction);
    brute_self(objs, &b);
    CHECK(a.result.numContacts() == b.result.numContacts());
  }

  // distance between meshes: the minimum over all pairs
  {
    auto menv = environment(60, 60, 23, true);
    auto mo = raw(menv);
    DynamicAABBTreeCollisionManagerd dm;
    dm.registerObjects(mo);
    dm.setup();
    DefaultDistanceDatad d;
    dm.distance(&d, DefaultDistanceFunction);
    double best = std::numeric_limits<double>::max();
    for (size_t i = 0; i < mo.size(); ++i)
      for (size_t j = i + 1; j < mo.size(); ++j) {
        DistanceRequestd req;
        DistanceResultd res;
        best = std::min(best, fcl::distance(mo[i], mo[j], req, res));
      }
    std::printf("distance: manager %.17g, all pairs %.17g\n", d.result.min_distance, best);
    CHECK(d.result.min_distance == best || (best <= 0 && d.result.min_distance <= 0));
  }

  std::printf(failures ? "FAILED (%d)\n" : "PASS\n", failures);
  return failures ? 1 : 0;
}
