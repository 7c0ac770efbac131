// fcl/fcl.h — drop-in C++ API of the MI355X engine for FCL's mesh path.
//
// Mirrors the parts of FCL 0.7.0's public API that the BVH-vs-BVH
// collide()/distance() path uses, with the same names, argument meanings,
// defaults and neutral error returns:
//   fcl::collide   include/fcl/narrowphase/collision.h:55-66, collision-inl.h:93-207
//   fcl::distance  include/fcl/narrowphase/distance.h, distance-inl.h:90-246
//   CollisionRequest / CollisionResult / Contact   collision_request.h:51-106,
//                  collision_result.h:51-93, contact.h:47-91
//   DistanceRequest / DistanceResult   distance_request.h:51-113, distance_result.h:50-107
//   BVHModel<BV>   include/fcl/geometry/bvh/BVH_model.h:56-227 (build API)
//   CollisionObject  include/fcl/narrowphase/collision_object.h
// Eigen is not required: Vector3/Matrix3/Transform3 are small value types
// with the subset of Eigen's interface that callers of this path use.
//   Box / Sphere / Capsule / Cylinder / Convex  include/fcl/geometry/shape/
// Mesh-mesh queries on BVHModel<OBBRSS<double>> (collide, distance),
// BVHModel<RSS<double>> (distance) and collide between two of the five
// primitives (the GST_INDEP solver) run on the GPU; any other pair of
// geometry types is reported unsupported with the reference's neutral
// return (0 contacts / max distance).  Only S = double is instantiated.
#ifndef FCL_AMD_FCL_H
#define FCL_AMD_FCL_H

#include <algorithm>
#include <algorithm>
#include <cmath>
#include <cstddef>
#include <cstdint>
#include <limits>
#include <memory>
#include <utility>
#include <vector>

namespace fcl {

// ------------------------------------------------------------------ math
template <typename S>
struct Vector3 {
  S v[3];
  Vector3() : v{0, 0, 0} {}
  Vector3(S x, S y, S z) : v{x, y, z} {}
  static Vector3 Zero() { return Vector3(); }
  S& operator[](int i) { return v[i]; }
  const S& operator[](int i) const { return v[i]; }
  S& operator()(int i) { return v[i]; }
  const S& operator()(int i) const { return v[i]; }
  Vector3 operator+(const Vector3& o) const { return {v[0] + o[0], v[1] + o[1], v[2] + o[2]}; }
  Vector3 operator-(const Vector3& o) const { return {v[0] - o[0], v[1] - o[1], v[2] - o[2]}; }
  Vector3 operator-() const { return {-v[0], -v[1], -v[2]}; }
  Vector3 operator*(S s) const { return {v[0] * s, v[1] * s, v[2] * s}; }
  Vector3 operator/(S s) const { return {v[0] / s, v[1] / s, v[2] / s}; }
  Vector3& operator+=(const Vector3& o) { return *this = *this + o; }
  Vector3& operator-=(const Vector3& o) { return *this = *this - o; }
  bool operator==(const Vector3& o) const { return v[0] == o[0] && v[1] == o[1] && v[2] == o[2]; }
  S dot(const Vector3& o) const { return (v[0] * o[0] + v[1] * o[1]) + v[2] * o[2]; }
  Vector3 cross(const Vector3& o) const {
    return {v[1] * o[2] - v[2] * o[1], v[2] * o[0] - v[0] * o[2], v[0] * o[1] - v[1] * o[0]};
  }
  S squaredNorm() const { return dot(*this); }
  S norm() const { return std::sqrt(squaredNorm()); }
  Vector3 normalized() const { return *this / norm(); }
  void normalize() { *this = normalized(); }
  static Vector3 UnitX() { return {1, 0, 0}; }
  static Vector3 UnitY() { return {0, 1, 0}; }
  static Vector3 UnitZ() { return {0, 0, 1}; }
  static Vector3 Ones() { return {1, 1, 1}; }
  static Vector3 Constant(S c) { return {c, c, c}; }
  void setZero() { v[0] = v[1] = v[2] = 0; }
  Vector3& operator*=(S s) { return *this = *this * s; }
  Vector3& operator/=(S s) { return *this = *this / s; }
  Vector3 cwiseAbs() const { return {std::abs(v[0]), std::abs(v[1]), std::abs(v[2])}; }
  Vector3 cwiseProduct(const Vector3& o) const { return {v[0] * o[0], v[1] * o[1], v[2] * o[2]}; }
  S maxCoeff() const { return std::max(v[0], std::max(v[1], v[2])); }
  S minCoeff() const { return std::min(v[0], std::min(v[1], v[2])); }
  bool isApprox(const Vector3& o, S prec = S(1e-12)) const {
    return (*this - o).norm() <= prec * std::min(norm(), o.norm());
  }
  S x() const { return v[0]; }
  S y() const { return v[1]; }
  S z() const { return v[2]; }
  S& x() { return v[0]; }
  S& y() { return v[1]; }
  S& z() { return v[2]; }
  const S* data() const { return v; }
  S* data() { return v; }
};
template <typename S>
Vector3<S> operator*(S s, const Vector3<S>& a) {
  return a * s;
}

template <typename S>
struct AngleAxis;
template <typename S>
struct Quaternion;

template <typename S>
struct Matrix3 {
  S m[9];  // row-major
  Matrix3() : m{0, 0, 0, 0, 0, 0, 0, 0, 0} {}
  // rotations convert to their matrices, as Eigen's do
  Matrix3(const AngleAxis<S>& aa);  // NOLINT(runtime/explicit)
  Matrix3(const Quaternion<S>& q);  // NOLINT(runtime/explicit)
  void setIdentity() { *this = Identity(); }
  void setZero() { *this = Matrix3(); }
  Matrix3 operator+(const Matrix3& o) const {
    Matrix3 r;
    for (int i = 0; i < 9; ++i) r.m[i] = m[i] + o.m[i];
    return r;
  }
  Matrix3 operator-(const Matrix3& o) const {
    Matrix3 r;
    for (int i = 0; i < 9; ++i) r.m[i] = m[i] - o.m[i];
    return r;
  }
  Matrix3 operator*(S s) const {
    Matrix3 r;
    for (int i = 0; i < 9; ++i) r.m[i] = m[i] * s;
    return r;
  }
  S determinant() const { return row(0).dot(row(1).cross(row(2))); }
  bool isApprox(const Matrix3& o, S prec = S(1e-12)) const {
    S d = 0, a = 0, b = 0;
    for (int i = 0; i < 9; ++i) {
      d += (m[i] - o.m[i]) * (m[i] - o.m[i]);
      a += m[i] * m[i];
      b += o.m[i] * o.m[i];
    }
    return std::sqrt(d) <= prec * std::sqrt(std::min(a, b));
  }
  static Matrix3 Identity() {
    Matrix3 r;
    r.m[0] = r.m[4] = r.m[8] = 1;
    return r;
  }
  static Matrix3 Zero() { return Matrix3(); }
  S& operator()(int i, int j) { return m[3 * i + j]; }
  const S& operator()(int i, int j) const { return m[3 * i + j]; }
  Matrix3 transpose() const {
    Matrix3 r;
    for (int i = 0; i < 3; ++i)
      for (int j = 0; j < 3; ++j) r(i, j) = (*this)(j, i);
    return r;
  }
  Vector3<S> row(int i) const { return {m[3 * i], m[3 * i + 1], m[3 * i + 2]}; }
  Vector3<S> col(int j) const { return {m[j], m[3 + j], m[6 + j]}; }
  Vector3<S> operator*(const Vector3<S>& p) const { return {row(0).dot(p), row(1).dot(p), row(2).dot(p)}; }
  Matrix3 operator*(const Matrix3& o) const {
    Matrix3 r;
    for (int i = 0; i < 3; ++i)
      for (int j = 0; j < 3; ++j) r(i, j) = row(i).dot(o.col(j));
    return r;
  }
  bool operator==(const Matrix3& o) const {
    for (int i = 0; i < 9; ++i)
      if (m[i] != o.m[i]) return false;
    return true;
  }
};

// Rigid transform (Eigen Isometry3 subset): p -> R p + t.
template <typename S>
struct Transform3 {
  Matrix3<S> R = Matrix3<S>::Identity();
  Vector3<S> t;
  static Transform3 Identity() { return Transform3(); }
  void setIdentity() { *this = Transform3(); }
  Matrix3<S>& linear() { return R; }
  const Matrix3<S>& linear() const { return R; }
  Matrix3<S>& rotation() { return R; }
  const Matrix3<S>& rotation() const { return R; }
  Vector3<S>& translation() { return t; }
  const Vector3<S>& translation() const { return t; }
  Vector3<S> operator*(const Vector3<S>& p) const { return R * p + t; }
  Transform3 operator*(const Transform3& o) const {
    Transform3 r;
    r.R = R * o.R;
    r.t = R * o.t + t;
    return r;
  }
  Transform3 inverse() const {
    Transform3 r;
    r.R = R.transpose();
    r.t = -(r.R * t);
    return r;
  }
  // Eigen's in-place forms: translate / rotate act on the right (in the
  // transform's own frame), pretranslate / prerotate on the left (world)
  Transform3& translate(const Vector3<S>& v) {
    t += R * v;
    return *this;
  }
  Transform3& pretranslate(const Vector3<S>& v) {
    t += v;
    return *this;
  }
  Transform3& rotate(const Matrix3<S>& r) {
    R = R * r;
    return *this;
  }
  Transform3& prerotate(const Matrix3<S>& r) {
    R = r * R;
    t = r * t;
    return *this;
  }
  Transform3 operator*(const Matrix3<S>& r) const {
    Transform3 o = *this;
    o.R = R * r;
    return o;
  }
};

// Rotation by `angle` about the unit vector `axis` (common/types.h: Eigen's
// AngleAxis).
template <typename S>
struct AngleAxis {
  S a = 0;
  Vector3<S> ax = Vector3<S>::UnitX();
  AngleAxis() = default;
  AngleAxis(S angle, const Vector3<S>& axis) : a(angle), ax(axis) {}
  S angle() const { return a; }
  const Vector3<S>& axis() const { return ax; }
  Matrix3<S> toRotationMatrix() const {
    const S c = std::cos(a), s = std::sin(a), t = 1 - c;
    const S x = ax[0], y = ax[1], z = ax[2];
    Matrix3<S> r;
    r(0, 0) = t * x * x + c;
    r(0, 1) = t * x * y - s * z;
    r(0, 2) = t * x * z + s * y;
    r(1, 0) = t * x * y + s * z;
    r(1, 1) = t * y * y + c;
    r(1, 2) = t * y * z - s * x;
    r(2, 0) = t * x * z - s * y;
    r(2, 1) = t * y * z + s * x;
    r(2, 2) = t * z * z + c;
    return r;
  }
};

// Unit quaternion w + xi + yj + zk (common/types.h: Eigen's Quaternion; the
// constructor takes w first, as Eigen's does).
template <typename S>
struct Quaternion {
  S q[4] = {1, 0, 0, 0};  // w, x, y, z
  Quaternion() = default;
  Quaternion(S w, S x, S y, S z) : q{w, x, y, z} {}
  explicit Quaternion(const AngleAxis<S>& aa) {
    const S h = aa.angle() / 2, s = std::sin(h);
    q[0] = std::cos(h);
    q[1] = aa.axis()[0] * s;
    q[2] = aa.axis()[1] * s;
    q[3] = aa.axis()[2] * s;
  }
  static Quaternion Identity() { return Quaternion(); }
  S w() const { return q[0]; }
  S x() const { return q[1]; }
  S y() const { return q[2]; }
  S z() const { return q[3]; }
  S norm() const { return std::sqrt(q[0] * q[0] + q[1] * q[1] + q[2] * q[2] + q[3] * q[3]); }
  Quaternion normalized() const {
    const S n = norm();
    return {q[0] / n, q[1] / n, q[2] / n, q[3] / n};
  }
  void normalize() { *this = normalized(); }
  Quaternion conjugate() const { return {q[0], -q[1], -q[2], -q[3]}; }
  Quaternion inverse() const { return conjugate(); }
  Quaternion operator*(const Quaternion& o) const {
    return {q[0] * o.q[0] - q[1] * o.q[1] - q[2] * o.q[2] - q[3] * o.q[3],
            q[0] * o.q[1] + q[1] * o.q[0] + q[2] * o.q[3] - q[3] * o.q[2],
            q[0] * o.q[2] - q[1] * o.q[3] + q[2] * o.q[0] + q[3] * o.q[1],
            q[0] * o.q[3] + q[1] * o.q[2] - q[2] * o.q[1] + q[3] * o.q[0]};
  }
  Matrix3<S> toRotationMatrix() const {
    const S w = q[0], x = q[1], y = q[2], z = q[3];
    Matrix3<S> r;
    r(0, 0) = 1 - 2 * (y * y + z * z);
    r(0, 1) = 2 * (x * y - w * z);
    r(0, 2) = 2 * (x * z + w * y);
    r(1, 0) = 2 * (x * y + w * z);
    r(1, 1) = 1 - 2 * (x * x + z * z);
    r(1, 2) = 2 * (y * z - w * x);
    r(2, 0) = 2 * (x * z - w * y);
    r(2, 1) = 2 * (y * z + w * x);
    r(2, 2) = 1 - 2 * (x * x + y * y);
    return r;
  }
  Vector3<S> operator*(const Vector3<S>& p) const { return toRotationMatrix() * p; }
};

template <typename S>
Matrix3<S>::Matrix3(const AngleAxis<S>& aa) : Matrix3(aa.toRotationMatrix()) {}
template <typename S>
Matrix3<S>::Matrix3(const Quaternion<S>& q) : Matrix3(q.toRotationMatrix()) {}

// Pure translation (common/types.h: Eigen's Translation<S, 3>).
template <typename S>
struct Translation3 {
  Vector3<S> t;
  Translation3() = default;
  Translation3(S x, S y, S z) : t(x, y, z) {}
  explicit Translation3(const Vector3<S>& v) : t(v) {}
  const Vector3<S>& vector() const { return t; }
  const Vector3<S>& translation() const { return t; }
  Transform3<S> operator*(const Matrix3<S>& R) const {
    Transform3<S> r;
    r.R = R;
    r.t = t;
    return r;
  }
  Transform3<S> operator*(const AngleAxis<S>& aa) const { return *this * aa.toRotationMatrix(); }
  Transform3<S> operator*(const Quaternion<S>& q) const { return *this * q.toRotationMatrix(); }
  operator Transform3<S>() const {  // NOLINT(runtime/explicit)
    Transform3<S> r;
    r.t = t;
    return r;
  }
};

using Vector3d = Vector3<double>;
using Matrix3d = Matrix3<double>;
using Transform3d = Transform3<double>;
using Quaterniond = Quaternion<double>;
using AngleAxisd = AngleAxis<double>;
using Translation3d = Translation3<double>;
using Vector3f = Vector3<float>;
using Matrix3f = Matrix3<float>;
using Transform3f = Transform3<float>;
using Quaternionf = Quaternion<float>;
using AngleAxisf = AngleAxis<float>;
using Translation3f = Translation3<float>;

// math/constants.h
template <typename S>
struct constants {
  static constexpr S pi() { return S(3.141592653589793238462643383279502884197169399L); }
  static constexpr S eps() { return std::numeric_limits<S>::epsilon(); }
};

// Axis-aligned box (math/bv/AABB.h): the broadphase's bounding volume.
// An empty box (the default) has min_ above max_, so adding a point sets it.
template <typename S>
struct AABB {
  Vector3<S> min_, max_;
  AABB()
      : min_(std::numeric_limits<S>::max(), std::numeric_limits<S>::max(),
             std::numeric_limits<S>::max()),
        max_(-std::numeric_limits<S>::max(), -std::numeric_limits<S>::max(),
             -std::numeric_limits<S>::max()) {}
  AABB(const Vector3<S>& a, const Vector3<S>& b) : min_(a), max_(b) {}
  // AABB-inl.h: closed intervals, touching boxes overlap
  bool overlap(const AABB& o) const {
    for (int k = 0; k < 3; ++k)
      if (min_[k] > o.max_[k] || max_[k] < o.min_[k]) return false;
    return true;
  }
  AABB& operator+=(const Vector3<S>& p) {
    for (int k = 0; k < 3; ++k) {
      if (p[k] < min_[k]) min_[k] = p[k];
      if (p[k] > max_[k]) max_[k] = p[k];
    }
    return *this;
  }
  Vector3<S> center() const { return (min_ + max_) * S(0.5); }
  S width() const { return max_[0] - min_[0]; }
  S height() const { return max_[1] - min_[1]; }
  S depth() const { return max_[2] - min_[2]; }
};
template <typename S>
AABB<S> translate(const AABB<S>& a, const Vector3<S>& t) {
  return AABB<S>(a.min_ + t, a.max_ + t);
}
using AABBd = AABB<double>;

// -------------------------------------------------------------- geometry
enum OBJECT_TYPE { OT_UNKNOWN, OT_BVH, OT_GEOM, OT_OCTREE, OT_COUNT };
enum NODE_TYPE {
  BV_UNKNOWN, BV_AABB, BV_OBB, BV_RSS, BV_kIOS, BV_OBBRSS, BV_KDOP16, BV_KDOP18, BV_KDOP24,
  GEOM_BOX, GEOM_SPHERE, GEOM_ELLIPSOID, GEOM_CAPSULE, GEOM_CONE, GEOM_CYLINDER, GEOM_CONVEX,
  GEOM_PLANE, GEOM_HALFSPACE, GEOM_TRIANGLE, GEOM_OCTREE, NODE_COUNT
};
enum GJKSolverType { GST_LIBCCD, GST_INDEP };
enum BVHBuildState {
  BVH_BUILD_STATE_EMPTY, BVH_BUILD_STATE_BEGUN, BVH_BUILD_STATE_PROCESSED,
  BVH_BUILD_STATE_UPDATE_BEGUN, BVH_BUILD_STATE_UPDATED, BVH_BUILD_STATE_REPLACE_BEGUN
};
enum BVHReturnCode {
  BVH_OK = 0, BVH_ERR_MODEL_OUT_OF_MEMORY = -1, BVH_ERR_BUILD_OUT_OF_SEQUENCE = -2,
  BVH_ERR_BUILD_EMPTY_MODEL = -3, BVH_ERR_BUILD_EMPTY_PREVIOUS_FRAME = -4,
  BVH_ERR_UNSUPPORTED_FUNCTION = -5, BVH_ERR_UNUPDATED_MODEL = -6, BVH_ERR_INCORRECT_DATA = -7,
  BVH_ERR_UNKNOWN = -8
};
enum BVHModelType { BVH_MODEL_UNKNOWN, BVH_MODEL_TRIANGLES, BVH_MODEL_POINTCLOUD };

template <typename S>
class CollisionGeometry {
 public:
  virtual ~CollisionGeometry() = default;
  virtual OBJECT_TYPE getObjectType() const { return OT_UNKNOWN; }
  virtual NODE_TYPE getNodeType() const { return BV_UNKNOWN; }
  // collision_geometry.h: the local AABB, its centre and the radius of the
  // sphere about that centre holding the geometry
  virtual void computeLocalAABB() {}
  void* getUserData() const { return user_data; }
  void setUserData(void* d) { user_data = d; }
  void* user_data = nullptr;
  Vector3<S> aabb_center;
  S aabb_radius = 1;
  AABB<S> aabb_local;

 protected:
  // centre and radius from aabb_local (shape_base-inl.h / computeLocalAABB
  // of the shapes)
  void setCenterAndRadius() {
    aabb_center = aabb_local.center();
    aabb_radius = (aabb_local.min_ - aabb_center).norm();
  }
};

// Bounding-volume tags of the two supported hierarchies.
template <typename S_>
struct OBBRSS {
  using S = S_;
  static constexpr NODE_TYPE node_type = BV_OBBRSS;
};
template <typename S_>
struct RSS {
  using S = S_;
  static constexpr NODE_TYPE node_type = BV_RSS;
};
using OBBRSSd = OBBRSS<double>;
using RSSd = RSS<double>;

struct Triangle {
  std::size_t vids[3] = {0, 0, 0};
  Triangle() = default;
  Triangle(std::size_t a, std::size_t b, std::size_t c) : vids{a, b, c} {}
  void set(std::size_t a, std::size_t b, std::size_t c) {
    vids[0] = a;
    vids[1] = b;
    vids[2] = c;
  }
  std::size_t operator[](int i) const { return vids[i]; }
  std::size_t& operator[](int i) { return vids[i]; }
};

namespace detail {
// device-resident copy of a built hierarchy (released with the last owner)
struct DeviceMesh;
int upload_mesh(int bv_type, const std::vector<Vector3d>& vertices,
                const std::vector<Triangle>& tris, std::shared_ptr<DeviceMesh>& out);
int device_mesh_id(const DeviceMesh* m);
}  // namespace detail

// BVHModel<BV>: beginModel / addVertex / addTriangle / addSubModel /
// endModel as in BVH_model-inl.h:206-517; endModel() builds the hierarchy
// (mean split) on the host and uploads it.
template <typename BV>
class BVHModel : public CollisionGeometry<typename BV::S> {
 public:
  using S = typename BV::S;
  BVHModel() = default;
  // copies share the immutable device hierarchy; raw arrays point at the copy's storage
  BVHModel(const BVHModel& o)
      : CollisionGeometry<S>(o), build_state(o.build_state), verts_(o.verts_), tris_(o.tris_),
        dev_(o.dev_) {
    sync();
  }
  BVHModel& operator=(const BVHModel& o) {
    if (this != &o) {
      CollisionGeometry<S>::operator=(o);
      build_state = o.build_state;
      verts_ = o.verts_;
      tris_ = o.tris_;
      dev_ = o.dev_;
      sync();
    }
    return *this;
  }

  OBJECT_TYPE getObjectType() const override { return OT_BVH; }
  NODE_TYPE getNodeType() const override { return BV::node_type; }
  BVHModelType getModelType() const {
    if (num_tris && num_vertices) return BVH_MODEL_TRIANGLES;
    if (num_vertices) return BVH_MODEL_POINTCLOUD;
    return BVH_MODEL_UNKNOWN;
  }
  int getNumBVs() const { return build_state == BVH_BUILD_STATE_PROCESSED ? 2 * num_tris - 1 : 0; }

  int beginModel(int num_tris_ = 0, int num_vertices_ = 0) {
    const bool was_empty = build_state == BVH_BUILD_STATE_EMPTY;
    verts_.clear();
    tris_.clear();
    dev_.reset();
    if (num_tris_ > 0) tris_.reserve(num_tris_);
    if (num_vertices_ > 0) verts_.reserve(num_vertices_);
    sync();
    if (!was_empty) {
      build_state = BVH_BUILD_STATE_EMPTY;
      return BVH_ERR_BUILD_OUT_OF_SEQUENCE;
    }
    build_state = BVH_BUILD_STATE_BEGUN;
    return BVH_OK;
  }
  int addVertex(const Vector3<S>& p) {
    if (build_state != BVH_BUILD_STATE_BEGUN) return BVH_ERR_BUILD_OUT_OF_SEQUENCE;
    verts_.push_back(p);
    sync();
    return BVH_OK;
  }
  int addTriangle(const Vector3<S>& p1, const Vector3<S>& p2, const Vector3<S>& p3) {
    if (build_state == BVH_BUILD_STATE_PROCESSED) return BVH_ERR_BUILD_OUT_OF_SEQUENCE;
    const std::size_t o = verts_.size();
    verts_.push_back(p1);
    verts_.push_back(p2);
    verts_.push_back(p3);
    tris_.emplace_back(o, o + 1, o + 2);
    sync();
    return BVH_OK;
  }
  int addSubModel(const std::vector<Vector3<S>>& ps, const std::vector<Triangle>& ts) {
    if (build_state == BVH_BUILD_STATE_PROCESSED) return BVH_ERR_BUILD_OUT_OF_SEQUENCE;
    const std::size_t o = verts_.size();
    verts_.insert(verts_.end(), ps.begin(), ps.end());
    for (const Triangle& t : ts) tris_.emplace_back(t[0] + o, t[1] + o, t[2] + o);
    sync();
    return BVH_OK;
  }
  int addSubModel(const std::vector<Vector3<S>>& ps) {
    if (build_state == BVH_BUILD_STATE_PROCESSED) return BVH_ERR_BUILD_OUT_OF_SEQUENCE;
    verts_.insert(verts_.end(), ps.begin(), ps.end());
    sync();
    return BVH_OK;
  }
  int endModel() {
    if (build_state != BVH_BUILD_STATE_BEGUN) return BVH_ERR_BUILD_OUT_OF_SEQUENCE;
    if (tris_.empty() && verts_.empty()) return BVH_ERR_BUILD_EMPTY_MODEL;
    if (tris_.empty()) return BVH_ERR_UNSUPPORTED_FUNCTION;  // point clouds: out of scope
    const int rc = detail::upload_mesh(BV::node_type == BV_OBBRSS ? 0 : 1, verts_, tris_, dev_);
    if (rc != BVH_OK) return rc;
    build_state = BVH_BUILD_STATE_PROCESSED;
    return BVH_OK;
  }

  // BVH_model-inl.h computeLocalAABB: the vertices' box, its centre, and
  // the largest distance of a vertex from that centre
  void computeLocalAABB() override {
    AABB<S> box;
    for (const auto& v : verts_) box += v;
    this->aabb_center = box.center();
    S r = 0;
    for (const auto& v : verts_) {
      const S d = (this->aabb_center - v).squaredNorm();
      if (d > r) r = d;
    }
    this->aabb_radius = std::sqrt(r);
    this->aabb_local = box;
  }

  // device handle of the built model (-1 before endModel)
  int deviceMeshId() const { return dev_ ? detail::device_mesh_id(dev_.get()) : -1; }

  // the reference exposes the raw arrays; these point into the model's storage
  Vector3<S>* vertices = nullptr;
  Triangle* tri_indices = nullptr;
  int num_tris = 0;
  int num_vertices = 0;
  BVHBuildState build_state = BVH_BUILD_STATE_EMPTY;

 private:
  void sync() {
    vertices = verts_.empty() ? nullptr : verts_.data();
    tri_indices = tris_.empty() ? nullptr : tris_.data();
    num_tris = static_cast<int>(tris_.size());
    num_vertices = static_cast<int>(verts_.size());
  }
  std::vector<Vector3<S>> verts_;
  std::vector<Triangle> tris_;
  std::shared_ptr<detail::DeviceMesh> dev_;
};

// ------------------------------------------------------ requests/results
// ------------------------------------------------------------ shapes
// The primitives of collide()'s shape path (geometry/shape/shape_base.h,
// box.h:58-67, sphere.h:57-60, capsule.h:58-64, cylinder.h:58-64,
// convex.h:110-152), in the reference's local frames: centred at the
// origin, capsule and cylinder along z with total length lz (the capsule's
// segment length).  A convex keeps its vertex and face lists; its support
// function scans the vertices (minkowski_diff-inl.h:182-196).
template <typename S>
class ShapeBase : public CollisionGeometry<S> {
 public:
  OBJECT_TYPE getObjectType() const override { return OT_GEOM; }
};

template <typename S_>
class Box : public ShapeBase<S_> {
 public:
  using S = S_;
  Box(S x, S y, S z) : side(x, y, z) {}
  explicit Box(const Vector3<S>& side_) : side(side_) {}
  Box() = default;
  NODE_TYPE getNodeType() const override { return GEOM_BOX; }
  void computeLocalAABB() override {
    const Vector3<S> h = side * S(0.5);
    this->aabb_local = AABB<S>(-h, h);
    this->setCenterAndRadius();
  }
  Vector3<S> side;
};

template <typename S_>
class Sphere : public ShapeBase<S_> {
 public:
  using S = S_;
  explicit Sphere(S radius_) : radius(radius_) {}
  NODE_TYPE getNodeType() const override { return GEOM_SPHERE; }
  void computeLocalAABB() override {
    this->aabb_local = AABB<S>(Vector3<S>(-radius, -radius, -radius),
                               Vector3<S>(radius, radius, radius));
    this->aabb_center = Vector3<S>();
    this->aabb_radius = radius;
  }
  S radius;
};

template <typename S_>
class Capsule : public ShapeBase<S_> {
 public:
  using S = S_;
  Capsule(S radius_, S lz_) : radius(radius_), lz(lz_) {}
  NODE_TYPE getNodeType() const override { return GEOM_CAPSULE; }
  void computeLocalAABB() override {
    const Vector3<S> h(radius, radius, S(0.5) * lz + radius);
    this->aabb_local = AABB<S>(-h, h);
    this->setCenterAndRadius();
  }
  S radius;
  S lz;
};

template <typename S_>
class Cylinder : public ShapeBase<S_> {
 public:
  using S = S_;
  Cylinder(S radius_, S lz_) : radius(radius_), lz(lz_) {}
  NODE_TYPE getNodeType() const override { return GEOM_CYLINDER; }
  void computeLocalAABB() override {
    const Vector3<S> h(radius, radius, S(0.5) * lz);
    this->aabb_local = AABB<S>(-h, h);
    this->setCenterAndRadius();
  }
  S radius;
  S lz;
};

template <typename S_>
class Convex : public ShapeBase<S_> {
 public:
  using S = S_;
  Convex(const std::shared_ptr<const std::vector<Vector3<S>>>& vertices, int num_faces,
         const std::shared_ptr<const std::vector<int>>& faces, bool throw_if_invalid = false)
      : vertices_(vertices), num_faces_(num_faces), faces_(faces) {
    (void)throw_if_invalid;
  }
  NODE_TYPE getNodeType() const override { return GEOM_CONVEX; }
  void computeLocalAABB() override {
    AABB<S> box;
    for (const auto& v : *vertices_) box += v;
    this->aabb_local = box;
    this->setCenterAndRadius();
  }
  const std::vector<Vector3<S>>& getVertices() const { return *vertices_; }
  int getFaceCount() const { return num_faces_; }
  const std::vector<int>& getFaces() const { return *faces_; }

 private:
  std::shared_ptr<const std::vector<Vector3<S>>> vertices_;
  int num_faces_;
  std::shared_ptr<const std::vector<int>> faces_;
};

using Boxd = Box<double>;
using Sphered = Sphere<double>;
using Capsuled = Capsule<double>;
using Cylinderd = Cylinder<double>;
using Convexd = Convex<double>;

template <typename S>
struct Contact {
  const CollisionGeometry<S>* o1 = nullptr;
  const CollisionGeometry<S>* o2 = nullptr;
  intptr_t b1 = NONE;
  intptr_t b2 = NONE;
  Vector3<S> normal;
  Vector3<S> pos;
  S penetration_depth = 0;
  static const int NONE = -1;
  Contact() = default;
  Contact(const CollisionGeometry<S>* o1_, const CollisionGeometry<S>* o2_, int b1_, int b2_)
      : o1(o1_), o2(o2_), b1(b1_), b2(b2_) {}
  // contact-inl.h:78-94: with the contact geometry (enable_contact)
  Contact(const CollisionGeometry<S>* o1_, const CollisionGeometry<S>* o2_, int b1_, int b2_,
          const Vector3<S>& pos_, const Vector3<S>& normal_, S depth_)
      : o1(o1_), o2(o2_), b1(b1_), b2(b2_), normal(normal_), pos(pos_),
        penetration_depth(depth_) {}
  bool operator<(const Contact& o) const {
    if (b1 == o.b1) return b2 < o.b2;
    return b1 < o.b1;
  }
  bool operator==(const Contact& o) const {
    return o1 == o.o1 && o2 == o.o2 && b1 == o.b1 && b2 == o.b2;
  }
};

template <typename S>
class CollisionResult;

template <typename S>
struct CollisionRequest {
  std::size_t num_max_contacts;
  bool enable_contact;
  std::size_t num_max_cost_sources;
  bool enable_cost;
  bool use_approximate_cost;
  GJKSolverType gjk_solver_type;
  bool enable_cached_gjk_guess = false;
  Vector3<S> cached_gjk_guess = Vector3<S>(1, 0, 0);
  S gjk_tolerance{1e-6};
  CollisionRequest(std::size_t num_max_contacts_ = 1, bool enable_contact_ = false,
                   std::size_t num_max_cost_sources_ = 1, bool enable_cost_ = false,
                   bool use_approximate_cost_ = true,
                   GJKSolverType gjk_solver_type_ = GST_LIBCCD, S gjk_tolerance_ = 1e-6)
      : num_max_contacts(num_max_contacts_), enable_contact(enable_contact_),
        num_max_cost_sources(num_max_cost_sources_), enable_cost(enable_cost_),
        use_approximate_cost(use_approximate_cost_), gjk_solver_type(gjk_solver_type_),
        gjk_tolerance(gjk_tolerance_) {}
  // collision_request-inl.h:77-82
  bool isSatisfied(const CollisionResult<S>& result) const {
    return (!enable_cost) && result.isCollision() && (num_max_contacts <= result.numContacts());
  }
};

template <typename S>
class CollisionResult {
 public:
  void addContact(const Contact<S>& c) { contacts.push_back(c); }
  bool isCollision() const { return !contacts.empty(); }
  std::size_t numContacts() const { return contacts.size(); }
  const Contact<S>& getContact(std::size_t i) const {
    return i < contacts.size() ? contacts[i] : contacts.back();
  }
  void getContacts(std::vector<Contact<S>>& out) const { out = contacts; }
  void clear() { contacts.clear(); }
  std::vector<Contact<S>> contacts;
};

template <typename S>
class DistanceResult;

template <typename S>
struct DistanceRequest {
  bool enable_nearest_points;
  bool enable_signed_distance;
  S rel_err;
  S abs_err;
  S distance_tolerance;
  GJKSolverType gjk_solver_type;
  explicit DistanceRequest(bool enable_nearest_points_ = false,
                           bool enable_signed_distance_ = false, S rel_err_ = 0.0,
                           S abs_err_ = 0.0, S distance_tolerance_ = 1e-6,
                           GJKSolverType gjk_solver_type_ = GST_LIBCCD)
      : enable_nearest_points(enable_nearest_points_),
        enable_signed_distance(enable_signed_distance_), rel_err(rel_err_), abs_err(abs_err_),
        distance_tolerance(distance_tolerance_), gjk_solver_type(gjk_solver_type_) {}
  // distance_request-inl.h:73-77
  bool isSatisfied(const DistanceResult<S>& result) const { return result.min_distance <= 0; }
};

template <typename S>
class DistanceResult {
 public:
  S min_distance = std::numeric_limits<S>::max();
  Vector3<S> nearest_points[2];
  const CollisionGeometry<S>* o1 = nullptr;
  const CollisionGeometry<S>* o2 = nullptr;
  intptr_t b1 = NONE;
  intptr_t b2 = NONE;
  static const int NONE = -1;
  // distance_result-inl.h:58-105: keep the first strict improvement
  void update(S distance, const CollisionGeometry<S>* o1_, const CollisionGeometry<S>* o2_,
              int b1_, int b2_) {
    if (min_distance > distance) {
      min_distance = distance;
      o1 = o1_;
      o2 = o2_;
      b1 = b1_;
      b2 = b2_;
    }
  }
  void update(S distance, const CollisionGeometry<S>* o1_, const CollisionGeometry<S>* o2_,
              int b1_, int b2_, const Vector3<S>& p1, const Vector3<S>& p2) {
    if (min_distance > distance) {
      min_distance = distance;
      o1 = o1_;
      o2 = o2_;
      b1 = b1_;
      b2 = b2_;
      nearest_points[0] = p1;
      nearest_points[1] = p2;
    }
  }
  void clear() { *this = DistanceResult(); }
};

template <typename S>
class CollisionObject {
 public:
  // collision_object-inl.h:63-86: the geometry's local AABB, then the world one
  explicit CollisionObject(const std::shared_ptr<CollisionGeometry<S>>& geom,
                           const Transform3<S>& tf = Transform3<S>::Identity())
      : geom_(geom), tf_(tf) {
    geom_->computeLocalAABB();
    computeAABB();
  }
  const std::shared_ptr<CollisionGeometry<S>>& collisionGeometry() const { return geom_; }
  const Transform3<S>& getTransform() const { return tf_; }
  const Matrix3<S>& getRotation() const { return tf_.linear(); }
  const Vector3<S>& getTranslation() const { return tf_.translation(); }
  // as in the reference, the world AABB follows a new pose only after
  // computeAABB()
  void setTransform(const Transform3<S>& tf) { tf_ = tf; }
  OBJECT_TYPE getObjectType() const { return geom_->getObjectType(); }
  NODE_TYPE getNodeType() const { return geom_->getNodeType(); }
  const AABB<S>& getAABB() const { return aabb_; }
  // collision_object-inl.h:118-131: the local box moved by the translation
  // when the rotation is the identity, else the cube of the geometry's
  // bounding radius about its moved centre
  void computeAABB() {
    if (tf_.linear() == Matrix3<S>::Identity()) {
      aabb_ = translate(geom_->aabb_local, tf_.translation());
    } else {
      const Vector3<S> c = tf_ * geom_->aabb_center;
      const Vector3<S> d(geom_->aabb_radius, geom_->aabb_radius, geom_->aabb_radius);
      aabb_ = AABB<S>(c - d, c + d);
    }
  }
  void* getUserData() const { return user_data_; }
  void setUserData(void* d) { user_data_ = d; }

 private:
  std::shared_ptr<CollisionGeometry<S>> geom_;
  Transform3<S> tf_;
  AABB<S> aabb_;
  void* user_data_ = nullptr;
};

// ------------------------------------------------------------- queries
template <typename S>
std::size_t collide(const CollisionGeometry<S>* o1, const Transform3<S>& tf1,
                    const CollisionGeometry<S>* o2, const Transform3<S>& tf2,
                    const CollisionRequest<S>& request, CollisionResult<S>& result);
template <typename S>
std::size_t collide(const CollisionObject<S>* o1, const CollisionObject<S>* o2,
                    const CollisionRequest<S>& request, CollisionResult<S>& result) {
  return collide(o1->collisionGeometry().get(), o1->getTransform(), o2->collisionGeometry().get(),
                 o2->getTransform(), request, result);
}

template <typename S>
S distance(const CollisionGeometry<S>* o1, const Transform3<S>& tf1,
           const CollisionGeometry<S>* o2, const Transform3<S>& tf2,
           const DistanceRequest<S>& request, DistanceResult<S>& result);
template <typename S>
S distance(const CollisionObject<S>* o1, const CollisionObject<S>* o2,
           const DistanceRequest<S>& request, DistanceResult<S>& result) {
  return distance(o1->collisionGeometry().get(), o1->getTransform(),
                  o2->collisionGeometry().get(), o2->getTransform(), request, result);
}

// the double instantiations live in libfclx.so
template <>
std::size_t collide<double>(const CollisionGeometry<double>*, const Transform3d&,
                            const CollisionGeometry<double>*, const Transform3d&,
                            const CollisionRequest<double>&, CollisionResult<double>&);
template <>
double distance<double>(const CollisionGeometry<double>*, const Transform3d&,
                        const CollisionGeometry<double>*, const Transform3d&,
                        const DistanceRequest<double>&, DistanceResult<double>&);

// ------------------------------------------------------------- batched
// Many mesh-mesh queries in one engine launch (no counterpart in the
// reference).  results[i] receives query i's contacts as collide() would
// append them (the set; for capped queries the reference's depth-first
// prefix).  Returns the number of queries processed, or (size_t)-1 when a
// query names unsupported geometry.
struct CollideQuery {
  const CollisionGeometry<double>* o1;
  Transform3d tf1;
  const CollisionGeometry<double>* o2;
  Transform3d tf2;
};
std::size_t collideBatch(const std::vector<CollideQuery>& queries,
                         const CollisionRequest<double>& request,
                         std::vector<CollisionResult<double>>& results);

// ------------------------------------------------------------ broadphase
// broadphase_collision_manager.h:57-66: a callback gets two objects and the
// caller's data, and returns true to stop the broadphase.
template <typename S>
using CollisionCallBack = bool (*)(CollisionObject<S>* o1, CollisionObject<S>* o2, void* cdata);
template <typename S>
using DistanceCallBack = bool (*)(CollisionObject<S>* o1, CollisionObject<S>* o2, void* cdata,
                                  S& dist);

// default_broadphase_callbacks.h:60-103: collide() into one shared result
// until the request is satisfied
template <typename S>
struct DefaultCollisionData {
  CollisionRequest<S> request;
  CollisionResult<S> result;
  bool done{false};
};
template <typename S>
bool DefaultCollisionFunction(CollisionObject<S>* o1, CollisionObject<S>* o2, void* data) {
  auto* collision_data = static_cast<DefaultCollisionData<S>*>(data);
  const CollisionRequest<S>& request = collision_data->request;
  CollisionResult<S>& result = collision_data->result;
  if (collision_data->done) return true;
  collide(o1, o2, request, result);
  if (!request.enable_cost && result.isCollision() &&
      result.numContacts() >= request.num_max_contacts)
    collision_data->done = true;
  return collision_data->done;
}

// default_broadphase_callbacks.h:177-228: distance() into one shared result;
// a pair in touch or collision stops the broadphase
template <typename S>
struct DefaultDistanceData {
  DistanceRequest<S> request;
  DistanceResult<S> result;
  bool done{false};
};
template <typename S>
bool DefaultDistanceFunction(CollisionObject<S>* o1, CollisionObject<S>* o2, void* data,
                             S& dist) {
  auto* cdata = static_cast<DefaultDistanceData<S>*>(data);
  const DistanceRequest<S>& request = cdata->request;
  DistanceResult<S>& result = cdata->result;
  if (cdata->done) {
    dist = result.min_distance;
    return true;
  }
  distance(o1, o2, request, result);
  dist = result.min_distance;
  if (dist <= 0) return true;
  return cdata->done;
}

// broadphase_collision_manager.h:77-122
template <typename S>
class BroadPhaseCollisionManager {
 public:
  virtual ~BroadPhaseCollisionManager() = default;
  virtual void registerObjects(const std::vector<CollisionObject<S>*>& other_objs) {
    for (auto* o : other_objs) registerObject(o);
  }
  virtual void registerObject(CollisionObject<S>* obj) = 0;
  virtual void unregisterObject(CollisionObject<S>* obj) = 0;
  virtual void setup() = 0;
  virtual void update() = 0;
  virtual void update(CollisionObject<S>* updated_obj) {
    (void)updated_obj;
    update();
  }
  virtual void update(const std::vector<CollisionObject<S>*>& updated_objs) {
    (void)updated_objs;
    update();
  }
  virtual void clear() = 0;
  virtual void getObjects(std::vector<CollisionObject<S>*>& objs) const = 0;
  virtual void collide(CollisionObject<S>* obj, void* cdata, CollisionCallBack<S> callback) const = 0;
  virtual void distance(CollisionObject<S>* obj, void* cdata, DistanceCallBack<S> callback) const = 0;
  virtual void collide(void* cdata, CollisionCallBack<S> callback) const = 0;
  virtual void distance(void* cdata, DistanceCallBack<S> callback) const = 0;
  virtual void collide(BroadPhaseCollisionManager* other_manager, void* cdata,
                       CollisionCallBack<S> callback) const = 0;
  virtual void distance(BroadPhaseCollisionManager* other_manager, void* cdata,
                        DistanceCallBack<S> callback) const = 0;
  virtual bool empty() const = 0;
  virtual std::size_t size() const = 0;
};

namespace detail {
using ObjectPair = std::pair<CollisionObject<double>*, CollisionObject<double>*>;
// Pairs (i, j), i < j, of boxes that overlap (AABB::overlap), ascending,
// found on the device; with groups only pairs across groups.  Returns false
// when the device call fails.
bool broadphase_pairs(const std::vector<AABB<double>>& boxes, const std::vector<int32_t>* groups,
                      std::vector<std::pair<int32_t, int32_t>>& out);
// DefaultCollisionFunction over the pairs in order, with the mesh pairs'
// narrow phase run as one device batch: the same result as calling it pair
// by pair.
void default_collide_pairs(const std::vector<ObjectPair>& pairs, DefaultCollisionData<double>* data);
}  // namespace detail

// A BroadPhaseCollisionManager whose candidate pairs come from one device
// sort-and-sweep over the objects' world AABBs (the boxes taken at setup()
// and update(), as the reference's managers take them), handed to the
// callback in ascending (i, j) registration order.  With
// DefaultCollisionFunction the mesh pairs' narrow phase runs as one batch.
// Distance queries visit pairs in ascending AABB distance and stop once the
// callback's distance is at most the next pair's AABB distance.  The name of
// FCL's default manager (broadphase_dynamic_AABB_tree.h) is kept for
// drop-in use; the answers (contact counts of exhaustive requests, collision
// flags, minimum distances) are the same, the visiting order is not.
template <typename S>
class DynamicAABBTreeCollisionManager : public BroadPhaseCollisionManager<S> {
 public:
  void registerObject(CollisionObject<S>* obj) override {
    objs_.push_back(obj);
    boxes_.push_back(obj->getAABB());
  }
  void unregisterObject(CollisionObject<S>* obj) override {
    for (std::size_t i = 0; i < objs_.size(); ++i)
      if (objs_[i] == obj) {
        objs_.erase(objs_.begin() + static_cast<std::ptrdiff_t>(i));
        boxes_.erase(boxes_.begin() + static_cast<std::ptrdiff_t>(i));
        return;
      }
  }
  void setup() override { update(); }
  void update() override {
    for (std::size_t i = 0; i < objs_.size(); ++i) boxes_[i] = objs_[i]->getAABB();
  }
  void clear() override {
    objs_.clear();
    boxes_.clear();
  }
  void getObjects(std::vector<CollisionObject<S>*>& objs) const override { objs = objs_; }
  bool empty() const override { return objs_.empty(); }
  std::size_t size() const override { return objs_.size(); }

  void collide(void* cdata, CollisionCallBack<S> callback) const override {
    std::vector<std::pair<int32_t, int32_t>> ij;
    if (!detail::broadphase_pairs(boxes_, nullptr, ij)) return;
    std::vector<detail::ObjectPair> ps;
    for (const auto& p : ij) ps.emplace_back(objs_[p.first], objs_[p.second]);
    run(ps, cdata, callback);
  }
  void collide(CollisionObject<S>* obj, void* cdata, CollisionCallBack<S> callback) const override {
    std::vector<AABB<S>> boxes{obj->getAABB()};
    boxes.insert(boxes.end(), boxes_.begin(), boxes_.end());
    std::vector<int32_t> groups(boxes.size(), 1);
    groups[0] = 0;
    std::vector<std::pair<int32_t, int32_t>> ij;
    if (!detail::broadphase_pairs(boxes, &groups, ij)) return;
    std::vector<detail::ObjectPair> ps;
    for (const auto& p : ij)
      if (objs_[p.second - 1] != obj) ps.emplace_back(obj, objs_[p.second - 1]);
    run(ps, cdata, callback);
  }
  void collide(BroadPhaseCollisionManager<S>* other_manager, void* cdata,
               CollisionCallBack<S> callback) const override {
    std::vector<CollisionObject<S>*> other;
    other_manager->getObjects(other);
    std::vector<AABB<S>> boxes = boxes_;
    for (auto* o : other) boxes.push_back(o->getAABB());
    std::vector<int32_t> groups(boxes.size(), 1);
    for (std::size_t i = 0; i < objs_.size(); ++i) groups[i] = 0;
    std::vector<std::pair<int32_t, int32_t>> ij;
    if (!detail::broadphase_pairs(boxes, &groups, ij)) return;
    const int32_t n = static_cast<int32_t>(objs_.size());
    std::vector<detail::ObjectPair> ps;
    for (const auto& p : ij) ps.emplace_back(objs_[p.first], other[p.second - n]);
    run(ps, cdata, callback);
  }

  void distance(void* cdata, DistanceCallBack<S> callback) const override {
    std::vector<detail::ObjectPair> ps;
    for (std::size_t i = 0; i < objs_.size(); ++i)
      for (std::size_t j = i + 1; j < objs_.size(); ++j) ps.emplace_back(objs_[i], objs_[j]);
    run_distance(ps, cdata, callback);
  }
  void distance(CollisionObject<S>* obj, void* cdata, DistanceCallBack<S> callback) const override {
    std::vector<detail::ObjectPair> ps;
    for (auto* o : objs_)
      if (o != obj) ps.emplace_back(obj, o);
    run_distance(ps, cdata, callback);
  }
  void distance(BroadPhaseCollisionManager<S>* other_manager, void* cdata,
                DistanceCallBack<S> callback) const override {
    std::vector<CollisionObject<S>*> other;
    other_manager->getObjects(other);
    std::vector<detail::ObjectPair> ps;
    for (auto* a : objs_)
      for (auto* b : other) ps.emplace_back(a, b);
    run_distance(ps, cdata, callback);
  }

 private:
  static void run(const std::vector<detail::ObjectPair>& ps, void* cdata,
                  CollisionCallBack<S> callback) {
    if (callback == &DefaultCollisionFunction<S>) {
      detail::default_collide_pairs(ps, static_cast<DefaultCollisionData<S>*>(cdata));
      return;
    }
    for (const auto& p : ps)
      if (callback(p.first, p.second, cdata)) return;
  }
  static S box_distance(const AABB<S>& a, const AABB<S>& b) {
    S d = 0;
    for (int k = 0; k < 3; ++k) {
      const S g = std::max(a.min_[k] - b.max_[k], b.min_[k] - a.max_[k]);
      if (g > 0) d += g * g;
    }
    return std::sqrt(d);
  }
  static void run_distance(std::vector<detail::ObjectPair> ps, void* cdata,
                           DistanceCallBack<S> callback) {
    std::vector<std::pair<S, std::size_t>> order;
    for (std::size_t k = 0; k < ps.size(); ++k)
      order.emplace_back(box_distance(ps[k].first->getAABB(), ps[k].second->getAABB()), k);
    std::stable_sort(order.begin(), order.end(),
                     [](const std::pair<S, std::size_t>& a, const std::pair<S, std::size_t>& b) {
                       return a.first < b.first;
                     });
    S dist = std::numeric_limits<S>::max();
    for (const auto& o : order) {
      if (o.first >= dist) return;  // no later pair can come closer
      if (callback(ps[o.second].first, ps[o.second].second, cdata, dist)) return;
    }
  }
  std::vector<CollisionObject<S>*> objs_;
  std::vector<AABB<S>> boxes_;
};
using DynamicAABBTreeCollisionManagerd = DynamicAABBTreeCollisionManager<double>;
using BroadPhaseCollisionManagerd = BroadPhaseCollisionManager<double>;
using DefaultCollisionDatad = DefaultCollisionData<double>;
using DefaultDistanceDatad = DefaultDistanceData<double>;
using CollisionObjectd = CollisionObject<double>;
using CollisionGeometryd = CollisionGeometry<double>;
using CollisionRequestd = CollisionRequest<double>;
using CollisionResultd = CollisionResult<double>;
using DistanceRequestd = DistanceRequest<double>;
using DistanceResultd = DistanceResult<double>;

}  // namespace fcl

#endif
