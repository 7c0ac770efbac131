// Device bounding-volume fitting in FP64, with the oracle's evaluation
// order (no fused multiply-add; the files that include this are built with
// -ffp-contract=off): the Jacobi eigen decomposition of a 3x3 covariance,
// the axis choice from it, and the OBB / RSS fits and merges of
// BVHModel::refitTree (BVH_model-inl.h:941-1076).
#pragma once
#include <hip/hip_runtime.h>

namespace fclx {

// eigen_old (math/geometry-inl.h:475-558): cyclic Jacobi, at most 50 sweeps;
// vout(r, k) = v[k][r] as the reference lays it out.  All index loops are
// unrolled so the matrices stay in registers.
__device__ inline void eigen_old(const double m[9], double dout[3], double vout[9]) {
  double R[3][3];
  double v[3][3] = {{1, 0, 0}, {0, 1, 0}, {0, 0, 1}};
  double b[3], z[3], d[3];
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j) R[i][j] = m[3 * i + j];
#pragma unroll
  for (int ip = 0; ip < 3; ++ip) {
    b[ip] = d[ip] = R[ip][ip];
    z[ip] = 0;
  }
  for (int it = 0; it < 50; ++it) {
    double sm = 0;
#pragma unroll
    for (int ip = 0; ip < 3; ++ip)
#pragma unroll
      for (int iq = ip + 1; iq < 3; ++iq) sm += fabs(R[ip][iq]);
    if (sm == 0.0) break;
    const double tresh = it < 3 ? 0.2 * sm / (3 * 3) : 0.0;
#pragma unroll
    for (int ip = 0; ip < 3; ++ip) {
#pragma unroll
      for (int iq = ip + 1; iq < 3; ++iq) {
        double g = 100.0 * fabs(R[ip][iq]);
        if (it > 3 && fabs(d[ip]) + g == fabs(d[ip]) && fabs(d[iq]) + g == fabs(d[iq])) {
          R[ip][iq] = 0.0;
        } else if (fabs(R[ip][iq]) > tresh) {
          double h = d[iq] - d[ip];
          double t;
          if (fabs(h) + g == fabs(h)) {
            t = (R[ip][iq]) / h;
          } else {
            const double theta = 0.5 * h / (R[ip][iq]);
            t = 1.0 / (fabs(theta) + sqrt(1.0 + theta * theta));
            if (theta < 0.0) t = -t;
          }
          const double c = 1.0 / sqrt(1 + t * t);
          const double s = t * c;
          const double tau = s / (1.0 + c);
          h = t * R[ip][iq];
          z[ip] -= h;
          z[iq] += h;
          d[ip] -= h;
          d[iq] += h;
          R[ip][iq] = 0.0;
#pragma unroll
          for (int j = 0; j < ip; ++j) {
            g = R[j][ip];
            h = R[j][iq];
            R[j][ip] = g - s * (h + g * tau);
            R[j][iq] = h + s * (g - h * tau);
          }
#pragma unroll
          for (int j = ip + 1; j < iq; ++j) {
            g = R[ip][j];
            h = R[j][iq];
            R[ip][j] = g - s * (h + g * tau);
            R[j][iq] = h + s * (g - h * tau);
          }
#pragma unroll
          for (int j = iq + 1; j < 3; ++j) {
            g = R[ip][j];
            h = R[iq][j];
            R[ip][j] = g - s * (h + g * tau);
            R[iq][j] = h + s * (g - h * tau);
          }
#pragma unroll
          for (int j = 0; j < 3; ++j) {
            g = v[j][ip];
            h = v[j][iq];
            v[j][ip] = g - s * (h + g * tau);
            v[j][iq] = h + s * (g - h * tau);
          }
        }
      }
    }
#pragma unroll
    for (int ip = 0; ip < 3; ++ip) {
      b[ip] += z[ip];
      d[ip] = b[ip];
      z[ip] = 0.0;
    }
  }
  // converged, or 50 sweeps done (the accumulator as it stands)
#pragma unroll
  for (int k = 0; k < 3; ++k)
#pragma unroll
    for (int r = 0; r < 3; ++r) vout[3 * r + k] = v[k][r];
#pragma unroll
  for (int k = 0; k < 3; ++k) dout[k] = d[k];
}

// axisFromEigen (math/geometry-inl.h:561-598): columns 0 and 1 the
// eigenvectors (rows of E) of the largest and middle eigenvalue, column 2
// their cross product.  axis row-major.
__device__ inline void eigen_order(const double s[3], int& mx, int& mid) {
  int mn;
  if (s[0] > s[1]) { mx = 0; mn = 1; } else { mn = 0; mx = 1; }
  if (s[2] < s[mn]) { mid = mn; mn = 2; }
  else if (s[2] > s[mx]) { mid = mx; mx = 2; }
  else { mid = 2; }
}

__device__ inline void fit_cross(const double a[3], const double b[3], double o[3]) {
  const double x = a[1] * b[2] - a[2] * b[1];
  const double y = a[2] * b[0] - a[0] * b[2];
  const double z = a[0] * b[1] - a[1] * b[0];
  o[0] = x;
  o[1] = y;
  o[2] = z;
}

__device__ inline double fit_dot(const double a[3], const double b[3]) {
  return (a[0] * b[0] + a[1] * b[1]) + a[2] * b[2];
}

__device__ inline void axis_from_eigen(const double E[9], const double s[3], double axis[9]) {
  int mx, mid;
  eigen_order(s, mx, mid);
  double c0[3], c1[3], c2[3];
  for (int r = 0; r < 3; ++r) {
    c0[r] = E[3 * mx + r];
    c1[r] = E[3 * mid + r];
  }
  fit_cross(c0, c1, c2);
  for (int r = 0; r < 3; ++r) {
    axis[3 * r + 0] = c0[r];
    axis[3 * r + 1] = c1[r];
    axis[3 * r + 2] = c2[r];
  }
}

// the coordinates of p along the axis columns
__device__ inline void proj_axes(const double axis[9], const double p[3], double out[3]) {
#pragma unroll
  for (int k = 0; k < 3; ++k)
    out[k] = (axis[0 + k] * p[0] + axis[3 + k] * p[1]) + axis[6 + k] * p[2];
}

// getExtentAndCenter (math/geometry-inl.h:229-362) over np points given by
// pt(i, p): extents and centre along the axes.
template <typename Pts>
__device__ void extent_and_center(const Pts& pt, int np, const double axis[9], double center[3],
                                  double extent[3]) {
  double mn[3] = {1.7976931348623157e308, 1.7976931348623157e308, 1.7976931348623157e308};
  double mx[3] = {-1.7976931348623157e308, -1.7976931348623157e308, -1.7976931348623157e308};
  for (int i = 0; i < np; ++i) {
    double p[3], pr[3];
    pt(i, p);
    proj_axes(axis, p, pr);
    for (int k = 0; k < 3; ++k) {
      if (pr[k] > mx[k]) mx[k] = pr[k];
      if (pr[k] < mn[k]) mn[k] = pr[k];
    }
  }
  double o[3];
  for (int k = 0; k < 3; ++k) o[k] = (mx[k] + mn[k]) / 2;
  for (int i = 0; i < 3; ++i)
    center[i] = (axis[3 * i + 0] * o[0] + axis[3 * i + 1] * o[1]) + axis[3 * i + 2] * o[2];
  for (int k = 0; k < 3; ++k) extent[k] = (mx[k] - mn[k]) / 2;
}

__device__ inline double clamp0(double x) { return (x < 0) ? 0.0 : x; }

// getRadiusAndOriginAndRectangleSize (math/geometry-inl.h:707-987) over np
// points given by pt(i, p).  The points' projections are taken again in
// every pass (the same arithmetic, so the same values) instead of being
// stored.
template <typename Pts>
__device__ void rss_fit(const Pts& pt, int np, const double axis[9], double origin[3],
                        double l[2], double* r) {
  auto P = [&](int i, double out[3]) {
    double p[3];
    pt(i, p);
    proj_axes(axis, p, out);
  };
  double q[3];
  P(0, q);
  double minz = q[2], maxz = q[2];
  for (int i = 1; i < np; ++i) {
    P(i, q);
    if (q[2] < minz) minz = q[2];
    else if (q[2] > maxz) maxz = q[2];
  }
  const double rr = 0.5 * (maxz - minz);
  const double radsqr = rr * rr;
  const double cz = 0.5 * (maxz + minz);
  *r = rr;
  double lo[2], hi[2];
  for (int ax = 0; ax < 2; ++ax) {
    int minindex = 0, maxindex = 0;
    P(0, q);
    double mintmp = q[ax], maxtmp = q[ax];
    for (int i = 1; i < np; ++i) {
      P(i, q);
      if (q[ax] < mintmp) { minindex = i; mintmp = q[ax]; }
      else if (q[ax] > maxtmp) { maxindex = i; maxtmp = q[ax]; }
    }
    P(minindex, q);
    double dz = q[2] - cz;
    double vmin = q[ax] + sqrt(clamp0(radsqr - dz * dz));
    P(maxindex, q);
    dz = q[2] - cz;
    double vmax = q[ax] - sqrt(clamp0(radsqr - dz * dz));
    for (int i = 0; i < np; ++i) {
      P(i, q);
      if (q[ax] < vmin) {
        dz = q[2] - cz;
        const double x = q[ax] + sqrt(clamp0(radsqr - dz * dz));
        if (x < vmin) vmin = x;
      }
    }
    for (int i = 0; i < np; ++i) {
      P(i, q);
      if (q[ax] > vmax) {
        dz = q[2] - cz;
        const double x = q[ax] - sqrt(clamp0(radsqr - dz * dz));
        if (x > vmax) vmax = x;
      }
    }
    lo[ax] = vmin;
    hi[ax] = vmax;
  }
  double minx = lo[0], maxx = hi[0], miny = lo[1], maxy = hi[1];
  const double a = sqrt(0.5);
  for (int i = 0; i < np; ++i) {
    P(i, q);
    const double px = q[0], py = q[1], pz = q[2];
    double dx, dy, u, t;
    if (px > maxx) {
      if (py > maxy) {
        dx = px - maxx;
        dy = py - maxy;
        u = dx * a + dy * a;
        t = (a * u - dx) * (a * u - dx) + (a * u - dy) * (a * u - dy) + (cz - pz) * (cz - pz);
        u = u - sqrt(clamp0(radsqr - t));
        if (u > 0) { maxx += u * a; maxy += u * a; }
      } else if (py < miny) {
        dx = px - maxx;
        dy = py - miny;
        u = dx * a - dy * a;
        t = (a * u - dx) * (a * u - dx) + (-a * u - dy) * (-a * u - dy) + (cz - pz) * (cz - pz);
        u = u - sqrt(clamp0(radsqr - t));
        if (u > 0) { maxx += u * a; miny -= u * a; }
      }
    } else if (px < minx) {
      if (py > maxy) {
        dx = px - minx;
        dy = py - maxy;
        u = dy * a - dx * a;
        t = (-a * u - dx) * (-a * u - dx) + (a * u - dy) * (a * u - dy) + (cz - pz) * (cz - pz);
        u = u - sqrt(clamp0(radsqr - t));
        if (u > 0) { minx -= u * a; maxy += u * a; }
      } else if (py < miny) {
        dx = px - minx;
        dy = py - miny;
        u = -dx * a - dy * a;
        t = (-a * u - dx) * (-a * u - dx) + (-a * u - dy) * (-a * u - dy) + (cz - pz) * (cz - pz);
        u = u - sqrt(clamp0(radsqr - t));
        if (u > 0) { minx -= u * a; miny -= u * a; }
      }
    }
  }
  for (int i = 0; i < 3; ++i)
    origin[i] = (axis[3 * i + 0] * minx + axis[3 * i + 1] * miny) + axis[3 * i + 2] * cz;
  l[0] = maxx - minx;
  if (l[0] < 0) l[0] = 0;
  l[1] = maxy - miny;
  if (l[1] < 0) l[1] = 0;
}

}  // namespace fclx
