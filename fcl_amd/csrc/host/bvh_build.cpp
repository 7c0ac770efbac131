// Host-side top-down BVH build: covariance -> Jacobi eigenvectors -> OBB
// extents + RSS rectangle/radius, split at the mean centroid projection on
// the principal axis.  Follows geometry/bvh/BVH_model-inl.h:832-938,
// geometry/bvh/detail/BV_fitter-inl.h:330-356,447-477,
// geometry/bvh/detail/BV_splitter-inl.h:539-599 and the fitting routines of
// math/geometry-inl.h (292-362, 475-598, 707-987, 1333-1421).
//
// Every expression uses one fixed evaluation order (dot = (x0y0+x1y1)+x2y2,
// no fused multiply-add: built with -ffp-contract=off), so the tree and
// every node coordinate are reproducible bit for bit.
#include "bvh_build.h"

#include <cfloat>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <utility>

namespace fclx {
namespace {

struct Vec {
  double x, y, z;
  double operator[](int i) const { return i == 0 ? x : (i == 1 ? y : z); }
};
inline double dot3(const double* a, const double* b) {
  return (a[0] * b[0] + a[1] * b[1]) + a[2] * b[2];
}
inline void cross3(const double* a, const double* b, double* o) {
  const double x = a[1] * b[2] - a[2] * b[1];
  const double y = a[2] * b[0] - a[0] * b[2];
  const double z = a[0] * b[1] - a[1] * b[0];
  o[0] = x;
  o[1] = y;
  o[2] = z;
}
// (x < 0) ? 0 : x, i.e. std::max<double>(x, 0)
inline double nonneg(double x) { return (x < 0) ? 0.0 : x; }

// Cyclic Jacobi eigen-decomposition of a symmetric 3x3 matrix (at most 50
// sweeps).  vecs receives eigenvector k as row k (the transpose of the
// accumulated rotation), vals the eigenvalues.
void jacobi3(const double m[9], double vals[3], double vecs[9]) {
  double a[3][3];
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) a[i][j] = m[3 * i + j];
  double v[3][3] = {{1, 0, 0}, {0, 1, 0}, {0, 0, 1}};
  double d[3], b[3], z[3];
  for (int p = 0; p < 3; ++p) {
    b[p] = d[p] = a[p][p];
    z[p] = 0;
  }
  auto emit = [&]() {
    for (int k = 0; k < 3; ++k)
      for (int r = 0; r < 3; ++r) vecs[3 * k + r] = v[r][k];
    for (int k = 0; k < 3; ++k) vals[k] = d[k];
  };
  const int n = 3;
  for (int sweep = 0; sweep < 50; ++sweep) {
    double off = 0;
    for (int p = 0; p < n; ++p)
      for (int q = p + 1; q < n; ++q) off += std::fabs(a[p][q]);
    if (off == 0.0) {
      emit();
      return;
    }
    const double tresh = (sweep < 3) ? 0.2 * off / (n * n) : 0.0;
    for (int p = 0; p < n; ++p) {
      for (int q = p + 1; q < n; ++q) {
        const double g = 100.0 * std::fabs(a[p][q]);
        if (sweep > 3 && std::fabs(d[p]) + g == std::fabs(d[p]) &&
            std::fabs(d[q]) + g == std::fabs(d[q])) {
          a[p][q] = 0.0;
          continue;
        }
        if (!(std::fabs(a[p][q]) > tresh)) continue;
        double h = d[q] - d[p];
        double t;
        if (std::fabs(h) + g == std::fabs(h)) {
          t = (a[p][q]) / h;
        } else {
          const double theta = 0.5 * h / (a[p][q]);
          t = 1.0 / (std::fabs(theta) + std::sqrt(1.0 + theta * theta));
          if (theta < 0.0) t = -t;
        }
        const double c = 1.0 / std::sqrt(1 + t * t);
        const double s = t * c;
        const double tau = s / (1.0 + c);
        h = t * a[p][q];
        z[p] -= h;
        z[q] += h;
        d[p] -= h;
        d[q] += h;
        a[p][q] = 0.0;
        auto rot = [&](double& x, double& y) {
          const double gx = x, hy = y;
          x = gx - s * (hy + gx * tau);
          y = hy + s * (gx - hy * tau);
        };
        for (int j = 0; j < p; ++j) rot(a[j][p], a[j][q]);
        for (int j = p + 1; j < q; ++j) rot(a[p][j], a[j][q]);
        for (int j = q + 1; j < n; ++j) rot(a[p][j], a[q][j]);
        for (int j = 0; j < n; ++j) rot(v[j][p], v[j][q]);
      }
    }
    for (int p = 0; p < n; ++p) {
      b[p] += z[p];
      d[p] = b[p];
      z[p] = 0.0;
    }
  }
  std::fprintf(stderr, "eigen: too many iterations in Jacobi transform.\n");
  emit();
}

class Builder {
 public:
  Builder(HostBVH& out) : out_(out), v_(out.vertices.data()), t_(out.tris.data()) {}

  void run() {
    const int n = static_cast<int>(out_.tris.size() / 3);
    out_.nodes.assign(2 * n - 1, HostNode{});
    out_.primitive_indices.resize(2 * n - 1);
    for (int i = 0; i < n; ++i) out_.primitive_indices[i] = static_cast<uint32_t>(i);
    proj_.resize(9 * static_cast<size_t>(n));
    next_ = 1;
    build(0, 0, n);
  }

 private:
  const double* vert(int i) const { return v_ + 3 * static_cast<size_t>(i); }
  const int32_t* tri(uint32_t i) const { return t_ + 3 * static_cast<size_t>(i); }

  // first/second moments of the node's vertices -> covariance (unnormalised)
  void covariance(const uint32_t* idx, int n, double M[9]) const {
    double s1[3] = {0, 0, 0};
    double sxx = 0, syy = 0, szz = 0, sxy = 0, sxz = 0, syz = 0;
    for (int i = 0; i < n; ++i) {
      const int32_t* t = tri(idx[i]);
      const double *a = vert(t[0]), *b = vert(t[1]), *c = vert(t[2]);
      for (int k = 0; k < 3; ++k) s1[k] += (a[k] + b[k]) + c[k];
      sxx += (a[0] * a[0] + b[0] * b[0] + c[0] * c[0]);
      syy += (a[1] * a[1] + b[1] * b[1] + c[1] * c[1]);
      szz += (a[2] * a[2] + b[2] * b[2] + c[2] * c[2]);
      sxy += (a[0] * a[1] + b[0] * b[1] + c[0] * c[1]);
      sxz += (a[0] * a[2] + b[0] * b[2] + c[0] * c[2]);
      syz += (a[1] * a[2] + b[1] * b[2] + c[1] * c[2]);
    }
    const double cnt = static_cast<double>(3 * n);
    M[0] = sxx - s1[0] * s1[0] / cnt;
    M[4] = syy - s1[1] * s1[1] / cnt;
    M[8] = szz - s1[2] * s1[2] / cnt;
    M[1] = M[3] = sxy - s1[0] * s1[1] / cnt;
    M[5] = M[7] = syz - s1[1] * s1[2] / cnt;
    M[2] = M[6] = sxz - s1[0] * s1[2] / cnt;
  }

  // principal axes: columns 0,1 = eigenvectors of the largest and middle
  // eigenvalue, column 2 = their cross product
  static void principal_axes(const double vals[3], const double vecs[9], double axis[9]) {
    int lo, mid, hi;
    if (vals[0] > vals[1]) {
      hi = 0;
      lo = 1;
    } else {
      lo = 0;
      hi = 1;
    }
    if (vals[2] < vals[lo]) {
      mid = lo;
      lo = 2;
    } else if (vals[2] > vals[hi]) {
      mid = hi;
      hi = 2;
    } else {
      mid = 2;
    }
    double c2[3];
    cross3(vecs + 3 * hi, vecs + 3 * mid, c2);
    for (int r = 0; r < 3; ++r) {
      axis[3 * r + 0] = vecs[3 * hi + r];
      axis[3 * r + 1] = vecs[3 * mid + r];
      axis[3 * r + 2] = c2[r];
    }
  }

  // coordinates of the node's 3n vertices in the axis frame -> proj_
  void project(const uint32_t* idx, int n, const double axis[9]) {
    double c0[3] = {axis[0], axis[3], axis[6]};
    double c1[3] = {axis[1], axis[4], axis[7]};
    double c2[3] = {axis[2], axis[5], axis[8]};
    double* P = proj_.data();
    for (int i = 0; i < n; ++i) {
      const int32_t* t = tri(idx[i]);
      for (int j = 0; j < 3; ++j) {
        const double* p = vert(t[j]);
        *P++ = dot3(c0, p);
        *P++ = dot3(c1, p);
        *P++ = dot3(c2, p);
      }
    }
  }

  static void to_world(const double axis[9], const double o[3], double w[3]) {
    for (int i = 0; i < 3; ++i)
      w[i] = (axis[3 * i + 0] * o[0] + axis[3 * i + 1] * o[1]) + axis[3 * i + 2] * o[2];
  }

  void fit_obb(int np, const double axis[9], HostNode& nd) const {
    double lo[3] = {DBL_MAX, DBL_MAX, DBL_MAX};
    double hi[3] = {-DBL_MAX, -DBL_MAX, -DBL_MAX};
    const double* P = proj_.data();
    for (int i = 0; i < np; ++i, P += 3) {
      for (int k = 0; k < 3; ++k) {
        if (P[k] > hi[k]) hi[k] = P[k];
        if (P[k] < lo[k]) lo[k] = P[k];
      }
    }
    double o[3];
    for (int k = 0; k < 3; ++k) o[k] = (hi[k] + lo[k]) / 2;
    to_world(axis, o, nd.obb_to);
    for (int k = 0; k < 3; ++k) nd.obb_extent[k] = (hi[k] - lo[k]) / 2;
  }

  // swept-sphere rectangle: radius from the z spread, then the rectangle's
  // x and y ranges shrunk by the sphere caps and grown until every point is
  // covered, then the four corners grown along the diagonals
  void fit_rss(int np, const double axis[9], HostNode& nd) const {
    const double* P = proj_.data();
    auto X = [&](int i) { return P[3 * i + 0]; };
    auto Y = [&](int i) { return P[3 * i + 1]; };
    auto Z = [&](int i) { return P[3 * i + 2]; };
    double zmin = Z(0), zmax = Z(0);
    for (int i = 1; i < np; ++i) {
      const double z = Z(i);
      if (z < zmin) zmin = z;
      else if (z > zmax) zmax = z;
    }
    const double r = 0.5 * (zmax - zmin);
    const double r2 = r * r;
    const double cz = 0.5 * (zmax + zmin);
    double range[2][2];
    for (int ax = 0; ax < 2; ++ax) {
      int imin = 0, imax = 0;
      double vmin = P[ax], vmax = P[ax];
      for (int i = 1; i < np; ++i) {
        const double val = P[3 * i + ax];
        if (val < vmin) {
          imin = i;
          vmin = val;
        } else if (val > vmax) {
          imax = i;
          vmax = val;
        }
      }
      double dz = Z(imin) - cz;
      double lo = P[3 * imin + ax] + std::sqrt(nonneg(r2 - dz * dz));
      dz = Z(imax) - cz;
      double hi = P[3 * imax + ax] - std::sqrt(nonneg(r2 - dz * dz));
      for (int i = 0; i < np; ++i) {
        if (P[3 * i + ax] < lo) {
          dz = Z(i) - cz;
          const double c = P[3 * i + ax] + std::sqrt(nonneg(r2 - dz * dz));
          if (c < lo) lo = c;
        }
      }
      for (int i = 0; i < np; ++i) {
        if (P[3 * i + ax] > hi) {
          dz = Z(i) - cz;
          const double c = P[3 * i + ax] - std::sqrt(nonneg(r2 - dz * dz));
          if (c > hi) hi = c;
        }
      }
      range[ax][0] = lo;
      range[ax][1] = hi;
    }
    double minx = range[0][0], maxx = range[0][1];
    double miny = range[1][0], maxy = range[1][1];
    const double k = std::sqrt(0.5);
    for (int i = 0; i < np; ++i) {
      const double dzc = cz - Z(i);
      if (X(i) > maxx) {
        if (Y(i) > maxy) {
          const double dx = X(i) - maxx, dy = Y(i) - maxy;
          double u = dx * k + dy * k;
          const double t = (k * u - dx) * (k * u - dx) + (k * u - dy) * (k * u - dy) + dzc * dzc;
          u = u - std::sqrt(nonneg(r2 - t));
          if (u > 0) {
            maxx += u * k;
            maxy += u * k;
          }
        } else if (Y(i) < miny) {
          const double dx = X(i) - maxx, dy = Y(i) - miny;
          double u = dx * k - dy * k;
          const double t =
              (k * u - dx) * (k * u - dx) + (-k * u - dy) * (-k * u - dy) + dzc * dzc;
          u = u - std::sqrt(nonneg(r2 - t));
          if (u > 0) {
            maxx += u * k;
            miny -= u * k;
          }
        }
      } else if (X(i) < minx) {
        if (Y(i) > maxy) {
          const double dx = X(i) - minx, dy = Y(i) - maxy;
          double u = dy * k - dx * k;
          const double t =
              (-k * u - dx) * (-k * u - dx) + (k * u - dy) * (k * u - dy) + dzc * dzc;
          u = u - std::sqrt(nonneg(r2 - t));
          if (u > 0) {
            minx -= u * k;
            maxy += u * k;
          }
        } else if (Y(i) < miny) {
          const double dx = X(i) - minx, dy = Y(i) - miny;
          double u = -dx * k - dy * k;
          const double t =
              (-k * u - dx) * (-k * u - dx) + (-k * u - dy) * (-k * u - dy) + dzc * dzc;
          u = u - std::sqrt(nonneg(r2 - t));
          if (u > 0) {
            minx -= u * k;
            miny -= u * k;
          }
        }
      }
    }
    const double o[3] = {minx, miny, cz};
    to_world(axis, o, nd.rss_to);
    nd.rss_l[0] = maxx - minx;
    if (nd.rss_l[0] < 0) nd.rss_l[0] = 0;
    nd.rss_l[1] = maxy - miny;
    if (nd.rss_l[1] < 0) nd.rss_l[1] = 0;
    nd.rss_r = r;
  }

  void build(int id, int first, int n) {
    HostNode& nd = out_.nodes[id];
    uint32_t* idx = out_.primitive_indices.data() + first;
    double M[9], vals[3], vecs[9];
    covariance(idx, n, M);
    jacobi3(M, vals, vecs);
    principal_axes(vals, vecs, nd.obb_axis);
    std::memcpy(nd.rss_axis, nd.obb_axis, sizeof(nd.obb_axis));
    project(idx, n, nd.obb_axis);
    fit_obb(3 * n, nd.obb_axis, nd);
    fit_rss(3 * n, nd.rss_axis, nd);
    nd.first_primitive = first;
    nd.num_primitives = n;
    if (n == 1) {
      nd.first_child = -(static_cast<int32_t>(idx[0]) + 1);
      return;
    }
    // mean split on the principal axis
    const double sv[3] = {nd.obb_axis[0], nd.obb_axis[3], nd.obb_axis[6]};
    double csum[3] = {0, 0, 0};
    for (int i = 0; i < n; ++i) {
      const int32_t* t = tri(idx[i]);
      const double *a = vert(t[0]), *b = vert(t[1]), *c = vert(t[2]);
      for (int k = 0; k < 3; ++k) csum[k] += (a[k] + b[k] + c[k]);
    }
    const double split =
        (csum[0] * sv[0] + csum[1] * sv[1] + csum[2] * sv[2]) / static_cast<double>(3 * n);
    const int left = next_;
    nd.first_child = left;
    next_ += 2;
    int nleft = 0;
    for (int i = 0; i < n; ++i) {
      const int32_t* t = tri(idx[i]);
      const double *a = vert(t[0]), *b = vert(t[1]), *c = vert(t[2]);
      double centroid[3];
      for (int k = 0; k < 3; ++k) centroid[k] = ((a[k] + b[k]) + c[k]) / 3.0;
      if (!(dot3(sv, centroid) > split)) std::swap(idx[i], idx[nleft++]);
    }
    if (nleft == 0 || nleft == n) nleft = n / 2;
    build(left, first, nleft);
    build(left + 1, first + nleft, n - nleft);
  }

  HostBVH& out_;
  const double* v_;
  const int32_t* t_;
  std::vector<double> proj_;
  int next_ = 1;
};

int tree_depth(const std::vector<HostNode>& nodes) {
  // iterative: (node, level)
  std::vector<std::pair<int, int>> stack{{0, 1}};
  int best = 0;
  while (!stack.empty()) {
    auto [b, lvl] = stack.back();
    stack.pop_back();
    if (lvl > best) best = lvl;
    if (nodes[b].first_child >= 0) {
      stack.push_back({nodes[b].first_child, lvl + 1});
      stack.push_back({nodes[b].first_child + 1, lvl + 1});
    }
  }
  return best;
}

}  // namespace

double HostBVH::node_size(int b) const {
  const HostNode& n = nodes[b];
  if (bv_type == kOBBRSS)
    return (n.obb_extent[0] * n.obb_extent[0] + n.obb_extent[1] * n.obb_extent[1]) +
           n.obb_extent[2] * n.obb_extent[2];
  return std::sqrt(n.rss_l[0] * n.rss_l[0] + n.rss_l[1] * n.rss_l[1]) + 2 * n.rss_r;
}

int build_bvh(HostBVH& out, int32_t bv_type, const double* vertices, int num_vertices,
              const int32_t* tris, int num_tris) {
  if (num_vertices <= 0 || num_tris <= 0) return -1;
  out.bv_type = bv_type;
  out.vertices.assign(vertices, vertices + 3 * static_cast<size_t>(num_vertices));
  out.tris.assign(tris, tris + 3 * static_cast<size_t>(num_tris));
  Builder b(out);
  b.run();
  out.depth = tree_depth(out.nodes);
  return 0;
}

}  // namespace fclx
