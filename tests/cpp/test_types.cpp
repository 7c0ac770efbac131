// The drop-in's math types against the Eigen idioms FCL callers use to
// build poses (common/types.h aliases: AngleAxis, Quaternion, Translation3,
// Transform3 as an isometry), checked against closed forms; host only.
#include <cmath>
#include <cstdio>

#include "fcl/common/types.h"

using namespace fcl;

static int failures = 0;
#define CHECK(cond)                                                \
  do {                                                             \
    if (!(cond)) {                                                 \
      std::printf("FAIL %s:%d: %s\n", __FILE__, __LINE__, #cond);  \
      ++failures;                                                  \
    }                                                              \
  } while (0)

static bool near(double a, double b, double tol = 1e-12) { return std::fabs(a - b) <= tol; }
static bool near(const Vector3d& a, const Vector3d& b, double tol = 1e-12) {
  return (a - b).norm() <= tol;
}

int main() {
  const double pi = constants<double>::pi();
  CHECK(near(pi, std::acos(-1.0), 0));
  // a quarter turn about z takes x to y
  const Matrix3d Rz = AngleAxisd(pi / 2, Vector3d::UnitZ()).toRotationMatrix();
  CHECK(near(Rz * Vector3d::UnitX(), Vector3d::UnitY()));
  CHECK(near(Rz.determinant(), 1.0));
  // the same rotation as a quaternion, and quaternion products compose
  const Quaterniond qz(AngleAxisd(pi / 2, Vector3d::UnitZ()));
  CHECK(qz.toRotationMatrix().isApprox(Rz));
  const Quaterniond qx(AngleAxisd(pi / 3, Vector3d::UnitX()));
  CHECK((qz * qx).toRotationMatrix().isApprox(Rz * AngleAxisd(pi / 3, Vector3d::UnitX()).toRotationMatrix()));
  CHECK(near(Quaterniond(2, 0, 0, 0).normalized().w(), 1.0));
  // implicit conversion of a rotation to its matrix, as Eigen allows
  Matrix3d M = AngleAxisd(0.3, Vector3d(1, 2, 2).normalized());
  CHECK((M * M.transpose()).isApprox(Matrix3d::Identity()));
  // an axis is fixed by its own rotation
  const Vector3d ax = Vector3d(1, 2, 2).normalized();
  CHECK(near(M * ax, ax));
  // poses built the usual ways agree: p -> R p + t
  Transform3d a = Transform3d::Identity();
  a.linear() = AngleAxisd(0.7, Vector3d::UnitY()).toRotationMatrix();
  a.translation() = Vector3d(1, -2, 3);
  const Transform3d b = Translation3d(1, -2, 3) * AngleAxisd(0.7, Vector3d::UnitY());
  const Vector3d p(0.5, 0.25, -1);
  CHECK(near(a * p, b * p));
  Transform3d c = Transform3d::Identity();
  c.pretranslate(Vector3d(1, -2, 3));
  c.rotate(AngleAxisd(0.7, Vector3d::UnitY()));
  CHECK(near(c * p, a * p));
  Transform3d d = Transform3d::Identity();
  d.rotate(AngleAxisd(0.7, Vector3d::UnitY()));
  d.translate(a.linear().transpose() * Vector3d(1, -2, 3));
  CHECK(near(d * p, a * p));
  CHECK(near((a * a.inverse()) * p, p));
  Transform3d e = Translation3d(Vector3d(4, 5, 6));
  CHECK(near(e * Vector3d::Zero(), Vector3d(4, 5, 6)));
  // vector helpers
  CHECK(near(Vector3d(3, 4, 0).normalized().norm(), 1.0));
  CHECK(near(Vector3d(-1, 2, -3).cwiseAbs().maxCoeff(), 3.0));
  CHECK(near(Vector3d::Ones().squaredNorm(), 3.0));
  if (failures == 0) std::printf("types ok\n");
  return failures == 0 ? 0 : 1;
}
