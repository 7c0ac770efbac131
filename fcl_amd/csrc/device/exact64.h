// FP64 device arithmetic of the mesh hot path: OBB separating-axis test,
// triangle-triangle SAT, RSS rectangle distance and triangle distance.
//
// These produce bit-identical results to FCL's double-precision routines as
// long as each small vector expression is evaluated in one fixed order and no
// multiply-add is fused.  The order used throughout (and by the oracle the
// tests compare against) is
//   dot(a,b)   = (a0*b0 + a1*b1) + a2*b2
//   (M v)_i    = dot(row i, v);  (A B)_ij = (A_i0 B_0j + A_i1 B_1j) + A_i2 B_2j
//   cross(a,b) = (a1 b2 - a2 b1, a2 b0 - a0 b2, a0 b1 - a1 b0)
// and the library is compiled with -ffp-contract=off.  Division and sqrt
// are the IEEE correctly rounded operations.
#pragma once
#include <hip/hip_runtime.h>
#include "../fclx_layout.h"

namespace fclx {

struct V3 {
  double x, y, z;
};

__device__ __forceinline__ V3 v3(double x, double y, double z) { return V3{x, y, z}; }
__device__ __forceinline__ V3 operator-(const V3& a, const V3& b) {
  return V3{a.x - b.x, a.y - b.y, a.z - b.z};
}
__device__ __forceinline__ V3 operator+(const V3& a, const V3& b) {
  return V3{a.x + b.x, a.y + b.y, a.z + b.z};
}
__device__ __forceinline__ double dot(const V3& a, const V3& b) {
  return (a.x * b.x + a.y * b.y) + a.z * b.z;
}
__device__ __forceinline__ V3 cross(const V3& a, const V3& b) {
  return V3{a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
// p + a*s, per component
__device__ __forceinline__ V3 madd(const V3& p, const V3& a, double s) {
  return V3{p.x + a.x * s, p.y + a.y * s, p.z + a.z * s};
}
__device__ __forceinline__ double sq(const V3& a) { return dot(a, a); }

// 3x3 matrix, row-major
struct M3 {
  double m[9];
};
__device__ __forceinline__ V3 row(const M3& A, int i) {
  return V3{A.m[3 * i], A.m[3 * i + 1], A.m[3 * i + 2]};
}
__device__ __forceinline__ V3 col(const double* A, int j) {
  return V3{A[j], A[3 + j], A[6 + j]};
}
// A*v
__device__ __forceinline__ V3 mul(const M3& A, const V3& v) {
  return V3{dot(row(A, 0), v), dot(row(A, 1), v), dot(row(A, 2), v)};
}

// std::min / std::max semantics ((b<a)?b:a and (a<b)?b:a)
__device__ __forceinline__ double smin(double a, double b) { return (b < a) ? b : a; }
__device__ __forceinline__ double smax(double a, double b) { return (a < b) ? b : a; }

// Relative pose of box 2 in box 1: R = A1^T (R0 A2), T = (R0 c2 + T0 - c1)^T A1.
// (math/bv/OBB-inl.h:383-395; math/bv/RSS-inl.h:1956-1974)
__device__ __forceinline__ void box_relative(const M3& R0, const V3& T0, const double* A1,
                                             const double* c1, const double* A2,
                                             const double* c2, double R[9], double T[3]) {
  double R0A2[9];
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j)
      R0A2[3 * i + j] =
          (R0.m[3 * i] * A2[j] + R0.m[3 * i + 1] * A2[3 + j]) + R0.m[3 * i + 2] * A2[6 + j];
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j)
      R[3 * i + j] = (A1[i] * R0A2[j] + A1[3 + i] * R0A2[3 + j]) + A1[6 + i] * R0A2[6 + j];
  const V3 C2 = v3(c2[0], c2[1], c2[2]);
  double Tt[3];
  Tt[0] = (dot(row(R0, 0), C2) + T0.x) - c1[0];
  Tt[1] = (dot(row(R0, 1), C2) + T0.y) - c1[1];
  Tt[2] = (dot(row(R0, 2), C2) + T0.z) - c1[2];
#pragma unroll
  for (int j = 0; j < 3; ++j) T[j] = (Tt[0] * A1[j] + Tt[1] * A1[3 + j]) + Tt[2] * A1[6 + j];
}

// 15-axis OBB SAT, math/bv/OBB-inl.h:398-523.  true = separated.
__device__ __forceinline__ bool obb_disjoint64(const double B[9], const double T[3],
                                               const double a[3], const double b[3]) {
  const double reps = 1e-6;
  double F[9];
#pragma unroll
  for (int i = 0; i < 9; ++i) F[i] = fabs(B[i]) + reps;
  // face axes of box 1 and box 2, in the reference's order A0,B0,A1,A2,B1,B2
  if (fabs(T[0]) > a[0] + ((F[0] * b[0] + F[1] * b[1]) + F[2] * b[2])) return true;
  {
    double s = (B[0] * T[0] + B[3] * T[1]) + B[6] * T[2];
    if (fabs(s) > b[0] + ((F[0] * a[0] + F[3] * a[1]) + F[6] * a[2])) return true;
  }
  if (fabs(T[1]) > a[1] + ((F[3] * b[0] + F[4] * b[1]) + F[5] * b[2])) return true;
  if (fabs(T[2]) > a[2] + ((F[6] * b[0] + F[7] * b[1]) + F[8] * b[2])) return true;
  {
    double s = (B[1] * T[0] + B[4] * T[1]) + B[7] * T[2];
    if (fabs(s) > b[1] + ((F[1] * a[0] + F[4] * a[1]) + F[7] * a[2])) return true;
  }
  {
    double s = (B[2] * T[0] + B[5] * T[1]) + B[8] * T[2];
    if (fabs(s) > b[2] + ((F[2] * a[0] + F[5] * a[1]) + F[8] * a[2])) return true;
  }
  // edge-edge axes Ai x Bj: with (i1,i2) the remaining box-1 indices in
  // cyclic order and j1<j2 the remaining box-2 indices,
  //   |T[i2] B(i1,j) - T[i1] B(i2,j)| >
  //       a[i1] F(i2,j) + a[i2] F(i1,j) + b[j1] F(i,j2) + b[j2] F(i,j1)
#pragma unroll
  for (int i = 0; i < 3; ++i) {
    const int i1 = (i + 1) % 3, i2 = (i + 2) % 3;
#pragma unroll
    for (int j = 0; j < 3; ++j) {
      const int j1 = (j == 0) ? 1 : 0, j2 = (j == 2) ? 1 : 2;
      double s = T[i2] * B[3 * i1 + j] - T[i1] * B[3 * i2 + j];
      double r = a[i1] * F[3 * i2 + j] + a[i2] * F[3 * i1 + j] + b[j1] * F[3 * i + j2] +
                 b[j2] * F[3 * i + j1];
      if (fabs(s) > r) return true;
    }
  }
  return false;
}

__device__ __forceinline__ bool obb_overlap64(const M3& R0, const V3& T0, const ObbNode64& n1,
                                              const ObbNode64& n2) {
  double B[9], T[3];
  box_relative(R0, T0, n1.axis, n1.to, n2.axis, n2.to, B, T);
  return !obb_disjoint64(B, T, n1.extent, n2.extent);
}

// projections of both triangles on ax overlap (touching counts),
// traversal/collision/intersect-inl.h:1085-1106
__device__ __forceinline__ bool project6(const V3& ax, const V3& p1, const V3& p2, const V3& p3,
                                         const V3& q1, const V3& q2, const V3& q3) {
  const double P1 = dot(ax, p1), P2 = dot(ax, p2), P3 = dot(ax, p3);
  const double Q1 = dot(ax, q1), Q2 = dot(ax, q2), Q3 = dot(ax, q3);
  const double mn1 = smin(P1, smin(P2, P3));
  const double mx2 = smax(Q1, smax(Q2, Q3));
  if (mn1 > mx2) return false;
  const double mx1 = smax(P1, smax(P2, P3));
  const double mn2 = smin(Q1, smin(Q2, Q3));
  if (mn2 > mx1) return false;
  return true;
}

// Triangle-triangle SAT with Q given in its own frame and (R,T) moving it
// into P's frame.  traversal/collision/intersect-inl.h:597-617, 727-845.
__device__ __forceinline__ bool tri_intersect64(const V3& P1, const V3& P2, const V3& P3,
                                                const V3& Q1, const V3& Q2, const V3& Q3,
                                                const M3& R, const V3& T) {
  const V3 Q1w = mul(R, Q1) + T, Q2w = mul(R, Q2) + T, Q3w = mul(R, Q3) + T;
  const V3 p1 = P1 - P1, p2 = P2 - P1, p3 = P3 - P1;
  const V3 q1 = Q1w - P1, q2 = Q2w - P1, q3 = Q3w - P1;
  const V3 e1 = p2 - p1, e2 = p3 - p2, e3 = p1 - p3;
  const V3 f1 = q2 - q1, f2 = q3 - q2, f3 = q1 - q3;
  const V3 n1 = cross(e1, e2);
  if (!project6(n1, p1, p2, p3, q1, q2, q3)) return false;
  const V3 m1 = cross(f1, f2);
  if (!project6(m1, p1, p2, p3, q1, q2, q3)) return false;
  // nine edge-edge axes e_i x f_j, row by row
  if (!project6(cross(e1, f1), p1, p2, p3, q1, q2, q3)) return false;
  if (!project6(cross(e1, f2), p1, p2, p3, q1, q2, q3)) return false;
  if (!project6(cross(e1, f3), p1, p2, p3, q1, q2, q3)) return false;
  if (!project6(cross(e2, f1), p1, p2, p3, q1, q2, q3)) return false;
  if (!project6(cross(e2, f2), p1, p2, p3, q1, q2, q3)) return false;
  if (!project6(cross(e2, f3), p1, p2, p3, q1, q2, q3)) return false;
  if (!project6(cross(e3, f1), p1, p2, p3, q1, q2, q3)) return false;
  if (!project6(cross(e3, f2), p1, p2, p3, q1, q2, q3)) return false;
  if (!project6(cross(e3, f3), p1, p2, p3, q1, q2, q3)) return false;
  // in-plane edge normals of each triangle
  if (!project6(cross(e1, n1), p1, p2, p3, q1, q2, q3)) return false;
  if (!project6(cross(e2, n1), p1, p2, p3, q1, q2, q3)) return false;
  if (!project6(cross(e3, n1), p1, p2, p3, q1, q2, q3)) return false;
  if (!project6(cross(f1, m1), p1, p2, p3, q1, q2, q3)) return false;
  if (!project6(cross(f2, m1), p1, p2, p3, q1, q2, q3)) return false;
  if (!project6(cross(f3, m1), p1, p2, p3, q1, q2, q3)) return false;
  return true;
}

// Plane of a triangle, n = normalised (v2-v1) x (v3-v1), t = n.v1
// (intersect-inl.h buildTrianglePlane with geometry-inl.h normalize).  A
// degenerate triangle has no plane; it gets n = 0, t = 0.
__device__ __forceinline__ void triangle_plane64(const V3& a, const V3& b, const V3& c, V3& n,
                                                 double& t) {
  n = cross(b - a, c - a);
  const double s = sq(n);
  if (s > 0) {
    const double l = sqrt(s);
    n = v3(n.x / l, n.y / l, n.z / l);
    t = dot(n, a);
  } else {
    n = v3(0, 0, 0);
    t = 0;
  }
}

// computeDeepestPoints (intersect-inl.h:848-885, epsilon 1e-5): the
// vertices of a triangle deepest below the plane (n,t), within 1e-6 of the
// deepest, none if all lie strictly on one side or all above -epsilon.
__device__ __forceinline__ void deepest_points64(const V3 (&p)[3], const V3& n, double t,
                                                 double& depth, V3 (&out)[3], unsigned& num) {
  const double eps = 1e-5;
  double max_depth = -1.7976931348623157e308;
  unsigned k = 0, neg = 0, pos = 0, zero = 0;
#pragma unroll
  for (int i = 0; i < 3; ++i) {
    const double dist = -(dot(n, p[i]) - t);
    if (dist > eps) ++pos;
    else if (dist < -eps) ++neg;
    else ++zero;
    if (dist > max_depth) {
      max_depth = dist;
      k = 1;
      out[0] = p[i];
    } else if (dist + 1e-6 >= max_depth) {
      ++k;
      out[k - 1] = p[i];
    }
  }
  if (max_depth < -eps) k = 0;
  if (zero == 0 && (neg == 0 || pos == 0)) k = 0;
  depth = max_depth;
  num = k;
}

// Contact output of intersect_Triangle (intersect-inl.h:802-842) for a pair
// that passed the SAT: P in its own frame, Q already moved into it (R Q + T
// as tri_intersect64 forms it).  Up to 2 points (P's frame), normal, depth.
__device__ __forceinline__ unsigned tri_contact64(const V3& P1, const V3& P2, const V3& P3,
                                                  const V3& Q1, const V3& Q2, const V3& Q3,
                                                  V3 (&pts)[2], V3& normal, double& depth) {
  V3 n1, n2;
  double t1, t2;
  triangle_plane64(P1, P2, P3, n1, t1);
  triangle_plane64(Q1, Q2, Q3, n2, t2);
  const V3 P[3] = {P1, P2, P3};
  const V3 Q[3] = {Q1, Q2, Q3};
  V3 dp1[3], dp2[3];
  double pd1, pd2;
  unsigned nd1, nd2;
  deepest_points64(Q, n1, t1, pd2, dp2, nd2);
  deepest_points64(P, n2, t2, pd1, dp1, nd1);
  unsigned n;
  if (pd1 > pd2) {
    n = nd2 < 2 ? nd2 : 2;
    pts[0] = dp2[0];
    pts[1] = dp2[1];
    normal = n1;
    depth = pd2;
  } else {
    n = nd1 < 2 ? nd1 : 2;
    pts[0] = dp1[0];
    pts[1] = dp1[1];
    normal = v3(-n2.x, -n2.y, -n2.z);
    depth = pd1;
  }
  return n;
}

}  // namespace fclx
