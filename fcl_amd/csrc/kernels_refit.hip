// BVH refit on the device: BVHModel::beginReplaceModel / replaceSubModel /
// endReplaceModel(refit = true, bottomup) (BVH_model-inl.h:521-620) for a
// resident mesh, keeping its topology (refitTree, :941-1076).
//
// Top-down (refitTree_topdown, :1049-1061): every node is fitted as the
// build fits it -- covariance of its triangles' vertices, Jacobi
// eigenvectors, extents, RSS rectangle and radius -- over its primitive
// range in the build's final primitive order.  One lane per node; each
// node's sums run in the oracle's sequential order, so the bits equal the
// oracle's.
//
// Bottom-up (refitTree_bottomup, :957-1017): a leaf's BV is fitted to its
// triangle (fit3: the normal and the longest edge as axes,
// math/bv/utility-inl.h:92-118, 208-232); an internal node's BV is
// left + right (OBB::operator+, OBB-inl.h:161-174, 254-371; RSS::operator+,
// RSS-inl.h:313-377).  One launch per tree level, deepest first, one lane per
// node of the level; a launch boundary orders each parent after its
// children.
//
// Afterwards BV::size(), the FP32 copies used for culling and the mesh's
// largest |OBB centre|_1 (the FP32 cull margin's scale) are recomputed.
#include <hip/hip_runtime.h>

#include "device/fit64.h"
#include "kernels.h"

namespace fclx {

namespace {

__global__ void refit_vertices(RefitLaunch L) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < L.num_verts) {
    for (int k = 0; k < 3; ++k) L.verts[3 * i + k] = L.new_verts[3 * i + k];
  }
  if (i < L.num_tris) {
    double* tv = L.tri_verts + static_cast<size_t>(kTriVertsStride) * i;
    for (int j = 0; j < 3; ++j) {
      const int v = L.tris[3 * i + j];
      for (int k = 0; k < 3; ++k) tv[3 * j + k] = L.new_verts[3 * static_cast<size_t>(v) + k];
    }
  }
}

__global__ void refit_zero(RefitLaunch L) {
  *L.cmax = 0.0;
  *L.rcmax = 0.0;
}

__global__ void refit_order(RefitLaunch L) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t < L.num_tris) L.order[L.prim_pos[t]] = t;
}

struct TriPoints {
  const double* v;
  const int32_t* tris;
  const int32_t* order;
  int first;
  __device__ void operator()(int i, double p[3]) const {
    const int t = order[first + i / 3];
    const double* q = v + 3 * static_cast<size_t>(tris[3 * t + i % 3]);
    p[0] = q[0];
    p[1] = q[1];
    p[2] = q[2];
  }
};

struct ArrayPoints {
  const double* p;
  __device__ void operator()(int i, double o[3]) const {
    o[0] = p[3 * i];
    o[1] = p[3 * i + 1];
    o[2] = p[3 * i + 2];
  }
};

// Covariance of the vertices of triangles order[first .. first + n)
// (getCovariance, triangle branch, math/geometry-inl.h:1333-1421).
__device__ void tri_covariance(const RefitLaunch& L, int first, int n, double M[9]) {
  double S1[3] = {0, 0, 0};
  double s00 = 0, s11 = 0, s22 = 0, s01 = 0, s02 = 0, s12 = 0;
  for (int i = 0; i < n; ++i) {
    const int t = L.order[first + i];
    const double* p1 = L.verts + 3 * static_cast<size_t>(L.tris[3 * t]);
    const double* p2 = L.verts + 3 * static_cast<size_t>(L.tris[3 * t + 1]);
    const double* p3 = L.verts + 3 * static_cast<size_t>(L.tris[3 * t + 2]);
    for (int k = 0; k < 3; ++k) S1[k] += (p1[k] + p2[k]) + p3[k];
    s00 += (p1[0] * p1[0] + p2[0] * p2[0] + p3[0] * p3[0]);
    s11 += (p1[1] * p1[1] + p2[1] * p2[1] + p3[1] * p3[1]);
    s22 += (p1[2] * p1[2] + p2[2] * p2[2] + p3[2] * p3[2]);
    s01 += (p1[0] * p1[1] + p2[0] * p2[1] + p3[0] * p3[1]);
    s02 += (p1[0] * p1[2] + p2[0] * p2[2] + p3[0] * p3[2]);
    s12 += (p1[1] * p1[2] + p2[1] * p2[2] + p3[1] * p3[2]);
  }
  const double np = static_cast<double>(3 * n);
  M[0] = s00 - S1[0] * S1[0] / np;
  M[4] = s11 - S1[1] * S1[1] / np;
  M[8] = s22 - S1[2] * S1[2] / np;
  M[1] = s01 - S1[0] * S1[1] / np;
  M[5] = s12 - S1[1] * S1[2] / np;
  M[2] = s02 - S1[0] * S1[2] / np;
  M[3] = M[1];
  M[6] = M[2];
  M[7] = M[5];
}

// getCovariance over a point array (point branch).
__device__ void pts_covariance(const double* ps, int n, double M[9]) {
  double S1[3] = {0, 0, 0};
  double s00 = 0, s11 = 0, s22 = 0, s01 = 0, s02 = 0, s12 = 0;
  for (int i = 0; i < n; ++i) {
    const double* p = ps + 3 * i;
    for (int k = 0; k < 3; ++k) S1[k] += p[k];
    s00 += p[0] * p[0];
    s11 += p[1] * p[1];
    s22 += p[2] * p[2];
    s01 += p[0] * p[1];
    s02 += p[0] * p[2];
    s12 += p[1] * p[2];
  }
  const double np = static_cast<double>(n);
  M[0] = s00 - S1[0] * S1[0] / np;
  M[4] = s11 - S1[1] * S1[1] / np;
  M[8] = s22 - S1[2] * S1[2] / np;
  M[1] = s01 - S1[0] * S1[1] / np;
  M[5] = s12 - S1[1] * S1[2] / np;
  M[2] = s02 - S1[0] * S1[2] / np;
  M[3] = M[1];
  M[6] = M[2];
  M[7] = M[5];
}

// BV fields of a node (links are left as they are)
struct Bv {
  double oa[9], oto[3], oe[3];   // OBB
  double ra[9], rto[3], rl[2], rr;  // RSS
};

__device__ void load_bv(const RefitLaunch& L, int b, Bv& o) {
  const ObbNode64& n = L.obb64[b];
  const RssNode64& r = L.rss64[b];
  for (int k = 0; k < 9; ++k) {
    o.oa[k] = n.axis[k];
    o.ra[k] = r.axis[k];
  }
  for (int k = 0; k < 3; ++k) {
    o.oto[k] = n.to[k];
    o.oe[k] = n.extent[k];
    o.rto[k] = r.to[k];
  }
  o.rl[0] = r.l[0];
  o.rl[1] = r.l[1];
  o.rr = r.r;
}

__device__ void store_bv(const RefitLaunch& L, int b, const Bv& o) {
  ObbNode64& n = L.obb64[b];
  RssNode64& r = L.rss64[b];
  for (int k = 0; k < 9; ++k) {
    n.axis[k] = o.oa[k];
    r.axis[k] = o.ra[k];
  }
  for (int k = 0; k < 3; ++k) {
    n.to[k] = o.oto[k];
    n.extent[k] = o.oe[k];
    r.to[k] = o.rto[k];
  }
  r.l[0] = o.rl[0];
  r.l[1] = o.rl[1];
  r.r = o.rr;
}

__global__ void __launch_bounds__(256) refit_topdown(RefitLaunch L) {
  const int b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= L.num_bvs) return;
  const int first = L.span[2 * b], n = L.span[2 * b + 1];
  double M[9], E[9], s[3];
  tri_covariance(L, first, n, M);
  eigen_old(M, s, E);
  Bv o;
  axis_from_eigen(E, s, o.oa);
  for (int k = 0; k < 9; ++k) o.ra[k] = o.oa[k];
  const TriPoints pts{L.verts, L.tris, L.order, first};
  extent_and_center(pts, 3 * n, o.oa, o.oto, o.oe);
  rss_fit(pts, 3 * n, o.ra, o.rto, o.rl, &o.rr);
  store_bv(L, b, o);
}

__device__ void normalize3(double v[3]) {
  const double nrm = sqrt((v[0] * v[0] + v[1] * v[1]) + v[2] * v[2]);
  v[0] /= nrm;
  v[1] /= nrm;
  v[2] /= nrm;
}

// fit3 axes (math/bv/utility-inl.h:92-118): the normal as axis 2, the
// longest edge as axis 0 (the first on ties), axis 1 = axis 2 x axis 0
__device__ void fit3_axes(const double p[9], double axis[9]) {
  double e[3][3], len[3];
  for (int k = 0; k < 3; ++k) {
    e[0][k] = p[k] - p[3 + k];
    e[1][k] = p[3 + k] - p[6 + k];
    e[2][k] = p[6 + k] - p[k];
  }
  for (int i = 0; i < 3; ++i) len[i] = (e[i][0] * e[i][0] + e[i][1] * e[i][1]) + e[i][2] * e[i][2];
  int imax = 0;
  if (len[1] > len[0]) imax = 1;
  if (len[2] > len[imax]) imax = 2;
  double c2[3], c0[3], c1[3];
  fit_cross(e[0], e[1], c2);
  normalize3(c2);
  for (int k = 0; k < 3; ++k) c0[k] = e[imax][k];
  normalize3(c0);
  fit_cross(c2, c0, c1);
  for (int r = 0; r < 3; ++r) {
    axis[3 * r + 0] = c0[r];
    axis[3 * r + 1] = c1[r];
    axis[3 * r + 2] = c2[r];
  }
}

// computeVertices (math/bv/OBB-inl.h:233-250)
__device__ void obb_corners(const Bv& b, double v[24]) {
  double e0[3], e1[3], e2[3];
  for (int r = 0; r < 3; ++r) {
    e0[r] = b.oa[3 * r + 0] * b.oe[0];
    e1[r] = b.oa[3 * r + 1] * b.oe[1];
    e2[r] = b.oa[3 * r + 2] * b.oe[2];
  }
  for (int i = 0; i < 8; ++i) {
    // signs of (e0, e1, e2) in the reference's vertex order
    const bool s0 = (i == 1 || i == 2 || i == 5 || i == 6);
    const bool s1 = (i == 2 || i == 3 || i == 6 || i == 7);
    const bool s2 = i >= 4;
    for (int r = 0; r < 3; ++r) {
      double x = b.oto[r];
      x = s0 ? x + e0[r] : x - e0[r];
      x = s1 ? x + e1[r] : x - e1[r];
      x = s2 ? x + e2[r] : x - e2[r];
      v[3 * i + r] = x;
    }
  }
}

// The quaternion (x, y, z, w) of a rotation matrix: the trace form when the
// trace is positive, else pivoting on the largest diagonal element
__device__ void quat_from_matrix(const double m[9], double q[4]) {
  const double tr = (m[0] + m[4]) + m[8];
  if (tr > 0) {
    double t = sqrt(tr + 1.0);
    q[3] = 0.5 * t;
    t = 0.5 / t;
    q[0] = (m[7] - m[5]) * t;
    q[1] = (m[2] - m[6]) * t;
    q[2] = (m[3] - m[1]) * t;
  } else {
    int i = 0;
    if (m[4] > m[0]) i = 1;
    if (m[8] > m[3 * i + i]) i = 2;
    const int j = (i + 1) % 3, k = (j + 1) % 3;
    double t = sqrt(((m[3 * i + i] - m[3 * j + j]) - m[3 * k + k]) + 1.0);
    q[i] = 0.5 * t;
    t = 0.5 / t;
    q[3] = (m[3 * k + j] - m[3 * j + k]) * t;
    q[j] = (m[3 * j + i] + m[3 * i + j]) * t;
    q[k] = (m[3 * k + i] + m[3 * i + k]) * t;
  }
}

__device__ void quat_to_matrix(const double q[4], double m[9]) {
  const double tx = 2 * q[0], ty = 2 * q[1], tz = 2 * q[2];
  const double twx = tx * q[3], twy = ty * q[3], twz = tz * q[3];
  const double txx = tx * q[0], txy = ty * q[0], txz = tz * q[0];
  const double tyy = ty * q[1], tyz = tz * q[1], tzz = tz * q[2];
  m[0] = 1 - (tyy + tzz);
  m[1] = txy - twz;
  m[2] = txz + twy;
  m[3] = txy + twz;
  m[4] = 1 - (txx + tzz);
  m[5] = tyz - twx;
  m[6] = txz - twy;
  m[7] = tyz + twx;
  m[8] = 1 - (txx + tyy);
}

__device__ double max3(const double e[3]) {
  const double a = (e[0] < e[1]) ? e[1] : e[0];
  return (a < e[2]) ? e[2] : a;
}

// OBB::operator+: boxes far apart (centre distance above twice the summed
// largest extents) take the centre line as axis 0 and the other two from
// the corners projected off it; otherwise the frames are averaged as
// quaternions.  Extents cover the 16 corners.
__device__ void obb_merge(const Bv& b1, const Bv& b2, Bv& out) {
  double d[3];
  for (int k = 0; k < 3; ++k) d[k] = b1.oto[k] - b2.oto[k];
  const double dn = sqrt((d[0] * d[0] + d[1] * d[1]) + d[2] * d[2]);
  double v[48];
  obb_corners(b1, v);
  obb_corners(b2, v + 24);
  if (dn > 2 * (max3(b1.oe) + max3(b2.oe))) {
    double a0[3] = {d[0], d[1], d[2]};
    normalize3(a0);
    double vp[48];
    for (int i = 0; i < 16; ++i) {
      const double dd = fit_dot(v + 3 * i, a0);
      for (int k = 0; k < 3; ++k) vp[3 * i + k] = v[3 * i + k] - a0[k] * dd;
    }
    double M[9], E[9], s[3];
    pts_covariance(vp, 16, M);
    eigen_old(M, s, E);
    int mx, mid;
    eigen_order(s, mx, mid);
    for (int r = 0; r < 3; ++r) {
      out.oa[3 * r + 0] = a0[r];
      out.oa[3 * r + 1] = E[3 * mx + r];
      out.oa[3 * r + 2] = E[3 * mid + r];
    }
    extent_and_center(ArrayPoints{v}, 16, out.oa, out.oto, out.oe);
    return;
  }
  double to[3];
  for (int k = 0; k < 3; ++k) to[k] = (b1.oto[k] + b2.oto[k]) * 0.5;
  double q0[4], q1[4], q[4];
  quat_from_matrix(b1.oa, q0);
  quat_from_matrix(b2.oa, q1);
  if (((q0[0] * q1[0] + q0[1] * q1[1]) + q0[2] * q1[2]) + q0[3] * q1[3] < 0)
    for (int k = 0; k < 4; ++k) q1[k] = -q1[k];
  for (int k = 0; k < 4; ++k) q[k] = q0[k] + q1[k];
  const double qn = sqrt(((q[0] * q[0] + q[1] * q[1]) + q[2] * q[2]) + q[3] * q[3]);
  for (int k = 0; k < 4; ++k) q[k] /= qn;
  double ax[9];
  quat_to_matrix(q, ax);
  double pmin[3] = {1.7976931348623157e308, 1.7976931348623157e308, 1.7976931348623157e308};
  double pmax[3] = {-1.7976931348623157e308, -1.7976931348623157e308, -1.7976931348623157e308};
  for (int i = 0; i < 16; ++i) {
    double diff[3];
    for (int k = 0; k < 3; ++k) diff[k] = v[3 * i + k] - to[k];
    for (int j = 0; j < 3; ++j) {
      const double dd = (diff[0] * ax[j] + diff[1] * ax[3 + j]) + diff[2] * ax[6 + j];
      if (dd > pmax[j]) pmax[j] = dd;
      else if (dd < pmin[j]) pmin[j] = dd;
    }
  }
  for (int j = 0; j < 3; ++j) {
    const double h = 0.5 * (pmax[j] + pmin[j]);
    for (int r = 0; r < 3; ++r) to[r] += ax[3 * r + j] * h;
    out.oe[j] = 0.5 * (pmax[j] - pmin[j]);
  }
  for (int k = 0; k < 9; ++k) out.oa[k] = ax[k];
  for (int k = 0; k < 3; ++k) out.oto[k] = to[k];
}

// the 8 corners of an RSS's box (rectangle grown by r), RSS-inl.h:319-351
__device__ void rss_corners(const Bv& b, double v[24]) {
  double d0p[3], d1p[3], d0n[3], d1n[3], d2p[3], d2n[3];
  for (int r = 0; r < 3; ++r) {
    d0p[r] = b.ra[3 * r + 0] * (b.rl[0] + b.rr);
    d1p[r] = b.ra[3 * r + 1] * (b.rl[1] + b.rr);
    d0n[r] = b.ra[3 * r + 0] * (-b.rr);
    d1n[r] = b.ra[3 * r + 1] * (-b.rr);
    d2p[r] = b.ra[3 * r + 2] * b.rr;
    d2n[r] = b.ra[3 * r + 2] * (-b.rr);
  }
  for (int i = 0; i < 8; ++i)
    for (int r = 0; r < 3; ++r) {
      const double a = (i >> 2) ? d0n[r] : d0p[r];
      const double bb = ((i >> 1) & 1) ? d1n[r] : d1p[r];
      const double c = (i & 1) ? d2n[r] : d2p[r];
      v[3 * i + r] = ((b.rto[r] + a) + bb) + c;
    }
}

// RSS::operator+: the corners of both (right operand first), axes 0 and 1
// the columns of the eigenvector matrix for the largest and middle
// eigenvalue, axis 2 the cross product of the left operand's axes 0 and 1
// (all as the reference computes them), then the rectangle and radius of
// the 16 points
__device__ void rss_merge(const Bv& left, const Bv& right, Bv& out) {
  double v[48];
  rss_corners(right, v);
  rss_corners(left, v + 24);
  double M[9], E[9], s[3];
  pts_covariance(v, 16, M);
  eigen_old(M, s, E);
  int mx, mid;
  eigen_order(s, mx, mid);
  double a0[3], a1[3], c[3];
  for (int r = 0; r < 3; ++r) {
    a0[r] = left.ra[3 * r + 0];
    a1[r] = left.ra[3 * r + 1];
  }
  fit_cross(a0, a1, c);
  for (int r = 0; r < 3; ++r) {
    out.ra[3 * r + 0] = E[3 * r + mx];
    out.ra[3 * r + 1] = E[3 * r + mid];
    out.ra[3 * r + 2] = c[r];
  }
  rss_fit(ArrayPoints{v}, 16, out.ra, out.rto, out.rl, &out.rr);
}

__global__ void __launch_bounds__(256) refit_bottomup_level(RefitLaunch L, const int32_t* nodes,
                                                            int count) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= count) return;
  const int b = nodes[k];
  const int fc = L.obb64[b].first_child;
  Bv o;
  if (fc < 0) {
    const int t = -(fc + 1);
    double p[9];
    for (int j = 0; j < 3; ++j)
      for (int c = 0; c < 3; ++c) p[3 * j + c] = L.verts[3 * static_cast<size_t>(L.tris[3 * t + j]) + c];
    fit3_axes(p, o.oa);
    extent_and_center(ArrayPoints{p}, 3, o.oa, o.oto, o.oe);
    for (int i = 0; i < 9; ++i) o.ra[i] = o.oa[i];
    rss_fit(ArrayPoints{p}, 3, o.ra, o.rto, o.rl, &o.rr);
  } else {
    Bv a, c;
    load_bv(L, fc, a);
    load_bv(L, fc + 1, c);
    obb_merge(a, c, o);
    rss_merge(a, c, o);
  }
  store_bv(L, b, o);
}

// BV::size() (OBBRSS-inl.h:133-136 / OBB-inl.h:206-209; RSS-inl.h:409-412),
// FP32 copies of the OBBs and RSSs, the largest |OBB centre|_1 and the
// largest |RSS origin|_1
__global__ void refit_finish(RefitLaunch L) {
  const int b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= L.num_bvs) return;
  const ObbNode64& n = L.obb64[b];
  const RssNode64& r = L.rss64[b];
  L.size64[b] = L.bv_type == 0
                    ? (n.extent[0] * n.extent[0] + n.extent[1] * n.extent[1]) +
                          n.extent[2] * n.extent[2]
                    : sqrt(r.l[0] * r.l[0] + r.l[1] * r.l[1]) + 2 * r.r;
  ObbNode32& f = L.obb32[b];
  for (int k = 0; k < 9; ++k) f.axis[k] = static_cast<float>(n.axis[k]);
  for (int k = 0; k < 3; ++k) {
    f.to[k] = static_cast<float>(n.to[k]);
    f.extent[k] = static_cast<float>(n.extent[k]);
  }
  RssNode32& g = L.rss32[b];
  for (int k = 0; k < 9; ++k) g.axis[k] = static_cast<float>(r.axis[k]);
  for (int k = 0; k < 3; ++k) g.to[k] = static_cast<float>(r.to[k]);
  g.l[0] = static_cast<float>(r.l[0]);
  g.l[1] = static_cast<float>(r.l[1]);
  g.r = static_cast<float>(r.r);
  const double c = (fabs(n.to[0]) + fabs(n.to[1])) + fabs(n.to[2]);
  const double rc = (fabs(r.to[0]) + fabs(r.to[1])) + fabs(r.to[2]);
  // non-negative doubles order as their bit patterns
  atomicMax(reinterpret_cast<unsigned long long*>(L.cmax),
            static_cast<unsigned long long>(__double_as_longlong(c)));
  atomicMax(reinterpret_cast<unsigned long long*>(L.rcmax),
            static_cast<unsigned long long>(__double_as_longlong(rc)));
}

}  // namespace

hipError_t launch_refit(const RefitLaunch& L, bool bottomup, const int32_t* level_nodes,
                        const int32_t* level_offsets_host, int levels, hipStream_t s) {
  const int block = 256;
  const int nv = L.num_verts > L.num_tris ? L.num_verts : L.num_tris;
  refit_zero<<<1, 1, 0, s>>>(L);
  refit_vertices<<<(nv + block - 1) / block, block, 0, s>>>(L);
  if (bottomup) {
    // deepest level first; level_offsets_host[l] .. [l+1] are level l's nodes
    for (int l = levels - 1; l >= 0; --l) {
      const int a = level_offsets_host[l], c = level_offsets_host[l + 1] - a;
      if (c > 0)
        refit_bottomup_level<<<(c + block - 1) / block, block, 0, s>>>(L, level_nodes + a, c);
    }
  } else {
    refit_order<<<(L.num_tris + block - 1) / block, block, 0, s>>>(L);
    refit_topdown<<<(L.num_bvs + block - 1) / block, block, 0, s>>>(L);
  }
  refit_finish<<<(L.num_bvs + block - 1) / block, block, 0, s>>>(L);
  return hipGetLastError();
}

}  // namespace fclx
