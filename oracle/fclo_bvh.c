/* fclo_bvh.c — top-down BVH build of the oracle (test infrastructure only).
 *
 * Restates BVHModel<BV>::endModel/buildTree/recursiveBuildTree
 * (geometry/bvh/BVH_model-inl.h:449-517, 832-938) with the OBBRSS/RSS fitter
 * (geometry/bvh/detail/BV_fitter-inl.h:330-356, 447-477) and the mean
 * splitter (geometry/bvh/detail/BV_splitter-inl.h:421-435, 539-599,
 * 491-500).  OBB and RSS of a node share the axes from the covariance's
 * eigenvectors, so one build serves both BVHModel<OBBRSS> and BVHModel<RSS>;
 * the two differ only in BV::size(). */
#include <float.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>
#include "fclo.h"

typedef struct {
  const double *v;
  const int *t;
} mesh_ref;

static const double *vtx(const mesh_ref *m, int i) { return m->v + 3 * i; }

/* math/geometry-inl.h:1333-1421 (getCovariance, triangle branch). */
static void covariance(const mesh_ref *m, const unsigned *idx, int n,
                       double M[9]) {
  double S1[3] = {0, 0, 0};
  double s00 = 0, s11 = 0, s22 = 0, s01 = 0, s02 = 0, s12 = 0;
  for (int i = 0; i < n; ++i) {
    const int *t = m->t + 3 * idx[i];
    const double *p1 = vtx(m, t[0]), *p2 = vtx(m, t[1]), *p3 = vtx(m, t[2]);
    for (int k = 0; k < 3; ++k) S1[k] += (p1[k] + p2[k]) + p3[k];
    s00 += (p1[0] * p1[0] + p2[0] * p2[0] + p3[0] * p3[0]);
    s11 += (p1[1] * p1[1] + p2[1] * p2[1] + p3[1] * p3[1]);
    s22 += (p1[2] * p1[2] + p2[2] * p2[2] + p3[2] * p3[2]);
    s01 += (p1[0] * p1[1] + p2[0] * p2[1] + p3[0] * p3[1]);
    s02 += (p1[0] * p1[2] + p2[0] * p2[2] + p3[0] * p3[2]);
    s12 += (p1[1] * p1[2] + p2[1] * p2[2] + p3[1] * p3[2]);
  }
  double np = (double)(3 * n);
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

/* math/geometry-inl.h:561-598 (axisFromEigen): columns 0,1 are the
 * eigenvectors of the largest and middle eigenvalue, column 2 their cross. */
void fclo_axis_from_eigen(const double E[9], const double s[3], double axis[9]) {
  int mn, mid, mx;
  if (s[0] > s[1]) { mx = 0; mn = 1; } else { mn = 0; mx = 1; }
  if (s[2] < s[mn]) { mid = mn; mn = 2; }
  else if (s[2] > s[mx]) { mid = mx; mx = 2; }
  else { mid = 2; }
  double c0[3], c1[3], c2[3];
  for (int r = 0; r < 3; ++r) {
    c0[r] = E[3 * mx + r];  /* eigenV.row(max) */
    c1[r] = E[3 * mid + r];
  }
  fclo_cross(c0, c1, c2);
  for (int r = 0; r < 3; ++r) {
    axis[3 * r + 0] = c0[r];
    axis[3 * r + 1] = c1[r];
    axis[3 * r + 2] = c2[r];
  }
}

static void proj_axes(const double axis[9], const double p[3], double out[3]) {
  for (int k = 0; k < 3; ++k)
    out[k] = (axis[0 + k] * p[0] + axis[3 + k] * p[1]) + axis[6 + k] * p[2];
}

/* math/geometry-inl.h:292-362 (getExtentAndCenter_mesh). */
static void extent_and_center(const mesh_ref *m, const unsigned *idx, int n,
                              const double axis[9], double center[3],
                              double extent[3]) {
  double mn[3] = {DBL_MAX, DBL_MAX, DBL_MAX};
  double mx[3] = {-DBL_MAX, -DBL_MAX, -DBL_MAX};
  for (int i = 0; i < n; ++i) {
    const int *t = m->t + 3 * idx[i];
    for (int j = 0; j < 3; ++j) {
      double pr[3];
      proj_axes(axis, vtx(m, t[j]), pr);
      for (int k = 0; k < 3; ++k) {
        if (pr[k] > mx[k]) mx[k] = pr[k];
        if (pr[k] < mn[k]) mn[k] = pr[k];
      }
    }
  }
  double o[3];
  for (int k = 0; k < 3; ++k) o[k] = (mx[k] + mn[k]) / 2;
  for (int i = 0; i < 3; ++i)
    center[i] = (axis[3 * i + 0] * o[0] + axis[3 * i + 1] * o[1]) +
                axis[3 * i + 2] * o[2];
  for (int k = 0; k < 3; ++k) extent[k] = (mx[k] - mn[k]) / 2;
}

/* std::max<S>(x, 0) as the reference calls it: (x < 0) ? 0 : x. */
static double clamp0(double x) { return (x < 0) ? 0.0 : x; }

/* math/geometry-inl.h:707-987 (getRadiusAndOriginAndRectangleSize). */
static void rss_fit_core(double *P, int np, const double axis[9], double origin[3],
                         double l[2], double *r);

static void rss_fit(const mesh_ref *m, const unsigned *idx, int n,
                    const double axis[9], double origin[3], double l[2],
                    double *r) {
  int np = 3 * n;
  double *P = (double *)malloc(sizeof(double) * 3 * (size_t)np);
  int pid = 0;
  for (int i = 0; i < n; ++i) {
    const int *t = m->t + 3 * idx[i];
    for (int j = 0; j < 3; ++j) proj_axes(axis, vtx(m, t[j]), P + 3 * pid++);
  }
  rss_fit_core(P, np, axis, origin, l, r);
}

/* The rectangle and radius of np points already projected on the axes
 * (P: 3 per point); frees P. */
static void rss_fit_core(double *P, int np, const double axis[9], double origin[3],
                         double l[2], double *r) {
#define PX(i) P[3 * (i) + 0]
#define PY(i) P[3 * (i) + 1]
#define PZ(i) P[3 * (i) + 2]
  if (np <= 0) {
    free(P);
    return;
  }
  double minz = PZ(0), maxz = PZ(0);
  for (int i = 1; i < np; ++i) {
    double zv = PZ(i);
    if (zv < minz) minz = zv;
    else if (zv > maxz) maxz = zv;
  }
  double rr = 0.5 * (maxz - minz);
  double radsqr = rr * rr;
  double cz = 0.5 * (maxz + minz);
  *r = rr;

  /* rectangle extent along x, then along y: start from the extreme points,
   * shrink by the sphere cap, then grow over every point. */
  double lo[2], hi[2];
  for (int ax = 0; ax < 2; ++ax) {
    int minindex = 0, maxindex = 0;
    double mintmp = P[ax], maxtmp = P[ax];
    for (int i = 1; i < np; ++i) {
      double val = P[3 * i + ax];
      if (val < mintmp) { minindex = i; mintmp = val; }
      else if (val > maxtmp) { maxindex = i; maxtmp = val; }
    }
    double dz = PZ(minindex) - cz;
    double vmin = P[3 * minindex + ax] + sqrt(clamp0(radsqr - dz * dz));
    dz = PZ(maxindex) - cz;
    double vmax = P[3 * maxindex + ax] - sqrt(clamp0(radsqr - dz * dz));
    for (int i = 0; i < np; ++i) {
      if (P[3 * i + ax] < vmin) {
        dz = PZ(i) - cz;
        double x = P[3 * i + ax] + sqrt(clamp0(radsqr - dz * dz));
        if (x < vmin) vmin = x;
      }
    }
    for (int i = 0; i < np; ++i) {
      if (P[3 * i + ax] > vmax) {
        dz = PZ(i) - cz;
        double x = P[3 * i + ax] - sqrt(clamp0(radsqr - dz * dz));
        if (x > vmax) vmax = x;
      }
    }
    lo[ax] = vmin;
    hi[ax] = vmax;
  }
  double minx = lo[0], maxx = hi[0], miny = lo[1], maxy = hi[1];

  /* corner growth (geometry-inl.h:902-976) */
  double a = sqrt(0.5);
  for (int i = 0; i < np; ++i) {
    double dx, dy, u, t;
    if (PX(i) > maxx) {
      if (PY(i) > maxy) {
        dx = PX(i) - maxx;
        dy = PY(i) - maxy;
        u = dx * a + dy * a;
        t = (a * u - dx) * (a * u - dx) + (a * u - dy) * (a * u - dy) +
            (cz - PZ(i)) * (cz - PZ(i));
        u = u - sqrt(clamp0(radsqr - t));
        if (u > 0) { maxx += u * a; maxy += u * a; }
      } else if (PY(i) < miny) {
        dx = PX(i) - maxx;
        dy = PY(i) - miny;
        u = dx * a - dy * a;
        t = (a * u - dx) * (a * u - dx) + (-a * u - dy) * (-a * u - dy) +
            (cz - PZ(i)) * (cz - PZ(i));
        u = u - sqrt(clamp0(radsqr - t));
        if (u > 0) { maxx += u * a; miny -= u * a; }
      }
    } else if (PX(i) < minx) {
      if (PY(i) > maxy) {
        dx = PX(i) - minx;
        dy = PY(i) - maxy;
        u = dy * a - dx * a;
        t = (-a * u - dx) * (-a * u - dx) + (a * u - dy) * (a * u - dy) +
            (cz - PZ(i)) * (cz - PZ(i));
        u = u - sqrt(clamp0(radsqr - t));
        if (u > 0) { minx -= u * a; maxy += u * a; }
      } else if (PY(i) < miny) {
        dx = PX(i) - minx;
        dy = PY(i) - miny;
        u = -dx * a - dy * a;
        t = (-a * u - dx) * (-a * u - dx) + (-a * u - dy) * (-a * u - dy) +
            (cz - PZ(i)) * (cz - PZ(i));
        u = u - sqrt(clamp0(radsqr - t));
        if (u > 0) { minx -= u * a; miny -= u * a; }
      }
    }
  }
#undef PX
#undef PY
#undef PZ
  for (int i = 0; i < 3; ++i)
    origin[i] = (axis[3 * i + 0] * minx + axis[3 * i + 1] * miny) +
                axis[3 * i + 2] * cz;
  l[0] = maxx - minx;
  if (l[0] < 0) l[0] = 0;
  l[1] = maxy - miny;
  if (l[1] < 0) l[1] = 0;
  free(P);
}

static void fit_node(const mesh_ref *m, const unsigned *idx, int n,
                     fclo_node *node) {
  double M[9], E[9], s[3];
  covariance(m, idx, n, M);
  fclo_eigen_old(M, s, E);
  fclo_axis_from_eigen(E, s, node->obb_axis);
  memcpy(node->rss_axis, node->obb_axis, sizeof(node->obb_axis));
  extent_and_center(m, idx, n, node->obb_axis, node->obb_to, node->obb_extent);
  rss_fit(m, idx, n, node->rss_axis, node->rss_to, node->rss_l, &node->rss_r);
}

typedef struct {
  mesh_ref mesh;
  fclo_node *bvs;
  unsigned *prim;
  int num_bvs;
} builder;

/* BVH_model-inl.h:867-938 with BV_splitter-inl.h:539-599 (mean split). */
static void build_rec(builder *b, int bv_id, int first, int num) {
  fclo_node *node = b->bvs + bv_id;
  unsigned *cur = b->prim + first;
  fit_node(&b->mesh, cur, num, node);
  /* split rule: projection of the summed vertex coordinates on axis col 0,
   * divided by 3*num (computeSplitValue_mean) */
  double sv[3] = {node->obb_axis[0], node->obb_axis[3], node->obb_axis[6]};
  double c[3] = {0, 0, 0};
  for (int i = 0; i < num; ++i) {
    const int *t = b->mesh.t + 3 * cur[i];
    const double *p1 = vtx(&b->mesh, t[0]), *p2 = vtx(&b->mesh, t[1]),
                 *p3 = vtx(&b->mesh, t[2]);
    for (int k = 0; k < 3; ++k) c[k] += (p1[k] + p2[k] + p3[k]);
  }
  double split_value = (c[0] * sv[0] + c[1] * sv[1] + c[2] * sv[2]) /
                       (double)(3 * num);
  node->first_primitive = first;
  node->num_primitives = num;
  if (num == 1) {
    node->first_child = -((int)(*cur) + 1);
    return;
  }
  node->first_child = b->num_bvs;
  b->num_bvs += 2;
  int c1 = 0;
  for (int i = 0; i < num; ++i) {
    const int *t = b->mesh.t + 3 * cur[i];
    const double *p1 = vtx(&b->mesh, t[0]), *p2 = vtx(&b->mesh, t[1]),
                 *p3 = vtx(&b->mesh, t[2]);
    double p[3];
    for (int k = 0; k < 3; ++k) p[k] = ((p1[k] + p2[k]) + p3[k]) / 3.0;
    if (!(fclo_dot(sv, p) > split_value)) {
      unsigned tmp = cur[i];
      cur[i] = cur[c1];
      cur[c1] = tmp;
      c1++;
    }
  }
  if (c1 == 0 || c1 == num) c1 = num / 2;
  int left = node->first_child;
  build_rec(b, left, first, c1);
  build_rec(b, left + 1, first + c1, num - c1);
}

int fclo_model_build(fclo_model *m, int bv_type, const double *vertices,
                     int num_vertices, const int *tris, int num_tris) {
  memset(m, 0, sizeof(*m));
  if (num_tris <= 0 || num_vertices <= 0) return -1;
  m->bv_type = bv_type;
  m->num_vertices = num_vertices;
  m->num_tris = num_tris;
  m->vertices = (double *)malloc(sizeof(double) * 3 * (size_t)num_vertices);
  m->tris = (int *)malloc(sizeof(int) * 3 * (size_t)num_tris);
  memcpy(m->vertices, vertices, sizeof(double) * 3 * (size_t)num_vertices);
  memcpy(m->tris, tris, sizeof(int) * 3 * (size_t)num_tris);
  int nb = 2 * num_tris - 1;
  m->bvs = (fclo_node *)calloc((size_t)nb, sizeof(fclo_node));
  m->primitive_indices = (unsigned *)malloc(sizeof(unsigned) * (size_t)nb);
  for (int i = 0; i < num_tris; ++i) m->primitive_indices[i] = (unsigned)i;
  builder b;
  b.mesh.v = m->vertices;
  b.mesh.t = m->tris;
  b.bvs = m->bvs;
  b.prim = m->primitive_indices;
  b.num_bvs = 1;
  build_rec(&b, 0, 0, num_tris);
  m->num_bvs = b.num_bvs;
  return 0;
}

void fclo_model_free(fclo_model *m) {
  free(m->vertices);
  free(m->tris);
  free(m->bvs);
  free(m->primitive_indices);
  memset(m, 0, sizeof(*m));
}

/* ---- refit (BVHModel::endReplaceModel(true, bottomup), BVH_model-inl.h:
 * 593-620, refitTree :941-1076) ------------------------------------------- */

/* getExtentAndCenter_pointcloud (math/geometry-inl.h:229-290). */
static void extent_and_center_pts(const double *ps, int n, const double axis[9],
                                  double center[3], double extent[3]) {
  double mn[3] = {DBL_MAX, DBL_MAX, DBL_MAX};
  double mx[3] = {-DBL_MAX, -DBL_MAX, -DBL_MAX};
  for (int i = 0; i < n; ++i) {
    double pr[3];
    proj_axes(axis, ps + 3 * i, pr);
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

/* getCovariance over a point array (math/geometry-inl.h:1333-1421, point
 * branch). */
static void covariance_pts(const double *ps, int n, double M[9]) {
  double S1[3] = {0, 0, 0};
  double s00 = 0, s11 = 0, s22 = 0, s01 = 0, s02 = 0, s12 = 0;
  for (int i = 0; i < n; ++i) {
    const double *p = ps + 3 * i;
    for (int k = 0; k < 3; ++k) S1[k] += p[k];
    s00 += p[0] * p[0];
    s11 += p[1] * p[1];
    s22 += p[2] * p[2];
    s01 += p[0] * p[1];
    s02 += p[0] * p[2];
    s12 += p[1] * p[2];
  }
  const double np = (double)n;
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

/* getRadiusAndOriginAndRectangleSize over a point array (point-cloud
 * branch: the same algorithm on the points' projections). */
static void rss_fit_pts(const double *ps, int n, const double axis[9], double origin[3],
                        double l[2], double *r) {
  double *P = (double *)malloc(sizeof(double) * 3 * (size_t)n);
  for (int i = 0; i < n; ++i) proj_axes(axis, ps + 3 * i, P + 3 * i);
  rss_fit_core(P, n, axis, origin, l, r);
}

static void normalize3(double v[3]) {
  const double nrm = sqrt((v[0] * v[0] + v[1] * v[1]) + v[2] * v[2]);
  v[0] /= nrm;
  v[1] /= nrm;
  v[2] /= nrm;
}

/* fit3 of OBB and RSS (math/bv/utility-inl.h:92-118, 208-232): axis 2 the
 * triangle's normal, axis 0 its longest edge (the first on ties), axis 1
 * their cross product; then the point-set extents (OBB) or rectangle and
 * radius (RSS) of the three points. */
static void fit3_axes(const double *p, double axis[9]) {
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
  fclo_cross(e[0], e[1], c2);
  normalize3(c2);
  memcpy(c0, e[imax], sizeof(c0));
  normalize3(c0);
  fclo_cross(c2, c0, c1);
  for (int r = 0; r < 3; ++r) {
    axis[3 * r + 0] = c0[r];
    axis[3 * r + 1] = c1[r];
    axis[3 * r + 2] = c2[r];
  }
}

static void obb_corners(const fclo_node *b, double v[24]) {
  double e0[3], e1[3], e2[3];
  for (int r = 0; r < 3; ++r) {
    e0[r] = b->obb_axis[3 * r + 0] * b->obb_extent[0];
    e1[r] = b->obb_axis[3 * r + 1] * b->obb_extent[1];
    e2[r] = b->obb_axis[3 * r + 2] * b->obb_extent[2];
  }
  /* computeVertices (math/bv/OBB-inl.h:233-250): signs of (e0, e1, e2) */
  static const int sg[8][3] = {{-1, -1, -1}, {1, -1, -1}, {1, 1, -1}, {-1, 1, -1},
                               {-1, -1, 1},  {1, -1, 1},  {1, 1, 1},  {-1, 1, 1}};
  for (int i = 0; i < 8; ++i)
    for (int r = 0; r < 3; ++r) {
      double x = b->obb_to[r];
      x = sg[i][0] > 0 ? x + e0[r] : x - e0[r];
      x = sg[i][1] > 0 ? x + e1[r] : x - e1[r];
      x = sg[i][2] > 0 ? x + e2[r] : x - e2[r];
      v[3 * i + r] = x;
    }
}

/* Quaternion of a rotation matrix (the trace-based method with the largest
 * diagonal element as pivot when the trace is not positive), normalised sum,
 * and back to a matrix -- merge_smalldist's averaging of two frames. */
static void quat_from_matrix(const double m[9], double q[4] /* x y z w */) {
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

static void quat_to_matrix(const double q[4], double m[9]) {
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

/* OBB::operator+ (math/bv/OBB-inl.h:161-174, merge_largedist :254-313,
 * merge_smalldist :316-371). */
static void obb_merge(const fclo_node *b1, const fclo_node *b2, fclo_node *out) {
  double d[3];
  for (int k = 0; k < 3; ++k) d[k] = b1->obb_to[k] - b2->obb_to[k];
  const double dn = sqrt((d[0] * d[0] + d[1] * d[1]) + d[2] * d[2]);
  const double e1 = fmax(fmax(b1->obb_extent[0], b1->obb_extent[1]), b1->obb_extent[2]);
  const double e2 = fmax(fmax(b2->obb_extent[0], b2->obb_extent[1]), b2->obb_extent[2]);
  double v[48];
  obb_corners(b1, v);
  obb_corners(b2, v + 24);
  if (dn > 2 * (e1 + e2)) {
    double a0[3] = {d[0], d[1], d[2]};
    normalize3(a0);
    double vp[48];
    for (int i = 0; i < 16; ++i) {
      const double dd = fclo_dot(v + 3 * i, a0);
      for (int k = 0; k < 3; ++k) vp[3 * i + k] = v[3 * i + k] - a0[k] * dd;
    }
    double M[9], E[9], s[3];
    covariance_pts(vp, 16, M);
    fclo_eigen_old(M, s, E);
    int mn, mid, mx;
    if (s[0] > s[1]) { mx = 0; mn = 1; } else { mn = 0; mx = 1; }
    if (s[2] < s[mn]) { mid = mn; mn = 2; }
    else if (s[2] > s[mx]) { mid = mx; mx = 2; }
    else { mid = 2; }
    (void)mn;
    for (int r = 0; r < 3; ++r) {
      out->obb_axis[3 * r + 0] = a0[r];
      out->obb_axis[3 * r + 1] = E[3 * mx + r];
      out->obb_axis[3 * r + 2] = E[3 * mid + r];
    }
    extent_and_center_pts(v, 16, out->obb_axis, out->obb_to, out->obb_extent);
    return;
  }
  double to[3];
  for (int k = 0; k < 3; ++k) to[k] = (b1->obb_to[k] + b2->obb_to[k]) * 0.5;
  double q0[4], q1[4], q[4];
  quat_from_matrix(b1->obb_axis, q0);
  quat_from_matrix(b2->obb_axis, q1);
  if (((q0[0] * q1[0] + q0[1] * q1[1]) + q0[2] * q1[2]) + q0[3] * q1[3] < 0)
    for (int k = 0; k < 4; ++k) q1[k] = -q1[k];
  for (int k = 0; k < 4; ++k) q[k] = q0[k] + q1[k];
  const double qn = sqrt(((q[0] * q[0] + q[1] * q[1]) + q[2] * q[2]) + q[3] * q[3]);
  for (int k = 0; k < 4; ++k) q[k] /= qn;
  double ax[9];
  quat_to_matrix(q, ax);
  double pmin[3] = {DBL_MAX, DBL_MAX, DBL_MAX}, pmax[3] = {-DBL_MAX, -DBL_MAX, -DBL_MAX};
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
    out->obb_extent[j] = 0.5 * (pmax[j] - pmin[j]);
  }
  memcpy(out->obb_axis, ax, sizeof(ax));
  memcpy(out->obb_to, to, sizeof(to));
}

/* RSS::operator+ (math/bv/RSS-inl.h:313-377): the 8 corners of each
 * operand's box (the rectangle grown by r), their covariance's eigenvectors
 * for axes 0 and 1, axis 2 as the cross product of the first operand's own
 * axes 0 and 1, then the rectangle and radius of the 16 points. */
static void rss_corners(const fclo_node *b, double v[24]) {
  double d0p[3], d1p[3], d0n[3], d1n[3], d2p[3], d2n[3];
  for (int r = 0; r < 3; ++r) {
    d0p[r] = b->rss_axis[3 * r + 0] * (b->rss_l[0] + b->rss_r);
    d1p[r] = b->rss_axis[3 * r + 1] * (b->rss_l[1] + b->rss_r);
    d0n[r] = b->rss_axis[3 * r + 0] * (-b->rss_r);
    d1n[r] = b->rss_axis[3 * r + 1] * (-b->rss_r);
    d2p[r] = b->rss_axis[3 * r + 2] * b->rss_r;
    d2n[r] = b->rss_axis[3 * r + 2] * (-b->rss_r);
  }
  const double *A[2] = {d0p, d0n}, *B[2] = {d1p, d1n}, *C[2] = {d2p, d2n};
  for (int i = 0; i < 8; ++i)
    for (int r = 0; r < 3; ++r)
      v[3 * i + r] = ((b->rss_to[r] + A[i >> 2][r]) + B[(i >> 1) & 1][r]) + C[i & 1][r];
}

static void rss_merge(const fclo_node *left, const fclo_node *right, fclo_node *out) {
  double v[48];
  rss_corners(right, v);
  rss_corners(left, v + 24);
  double M[9], E[9], s[3];
  covariance_pts(v, 16, M);
  fclo_eigen_old(M, s, E);
  int mn, mid, mx;
  if (s[0] > s[1]) { mx = 0; mn = 1; } else { mn = 0; mx = 1; }
  if (s[2] < s[mn]) { mid = mn; mn = 2; }
  else if (s[2] > s[mx]) { mid = mx; mx = 2; }
  else { mid = 2; }
  (void)mn;
  double a0[3], a1[3], c[3];
  for (int r = 0; r < 3; ++r) {
    a0[r] = left->rss_axis[3 * r + 0];
    a1[r] = left->rss_axis[3 * r + 1];
  }
  fclo_cross(a0, a1, c);
  double ax[9];
  for (int r = 0; r < 3; ++r) {
    ax[3 * r + 0] = E[3 * r + mx];
    ax[3 * r + 1] = E[3 * r + mid];
    ax[3 * r + 2] = c[r];
  }
  memcpy(out->rss_axis, ax, sizeof(ax));
  rss_fit_pts(v, 16, ax, out->rss_to, out->rss_l, &out->rss_r);
}

static void refit_bottomup_rec(fclo_model *m, int b) {
  fclo_node *n = m->bvs + b;
  if (n->first_child < 0) {
    const int *t = m->tris + 3 * (-(n->first_child + 1));
    double p[9];
    for (int j = 0; j < 3; ++j) memcpy(p + 3 * j, m->vertices + 3 * t[j], 3 * sizeof(double));
    fit3_axes(p, n->obb_axis);
    extent_and_center_pts(p, 3, n->obb_axis, n->obb_to, n->obb_extent);
    memcpy(n->rss_axis, n->obb_axis, sizeof(n->obb_axis));
    rss_fit_pts(p, 3, n->rss_axis, n->rss_to, n->rss_l, &n->rss_r);
    return;
  }
  refit_bottomup_rec(m, n->first_child);
  refit_bottomup_rec(m, n->first_child + 1);
  const fclo_node *a = m->bvs + n->first_child, *c = a + 1;
  obb_merge(a, c, n);
  rss_merge(a, c, n);
}

int fclo_model_refit(fclo_model *m, const double *vertices, int bottomup) {
  memcpy(m->vertices, vertices, sizeof(double) * 3 * (size_t)m->num_vertices);
  if (bottomup) {
    refit_bottomup_rec(m, 0);
    return 0;
  }
  mesh_ref mr = {m->vertices, m->tris};
  for (int i = 0; i < m->num_bvs; ++i) {
    fclo_node *n = m->bvs + i;
    fit_node(&mr, m->primitive_indices + n->first_primitive, n->num_primitives, n);
  }
  return 0;
}

static int depth_rec(const fclo_model *m, int b) {
  if (m->bvs[b].first_child < 0) return 1;
  int l = depth_rec(m, m->bvs[b].first_child);
  int r = depth_rec(m, m->bvs[b].first_child + 1);
  return 1 + (l > r ? l : r);
}

int fclo_model_depth(const fclo_model *m) { return depth_rec(m, 0); }

/* OBBRSS::size() = obb.extent.squaredNorm() (OBBRSS-inl.h:133-136,
 * OBB-inl.h:206-209); RSS::size() = sqrt(l0^2 + l1^2) + 2r (RSS-inl.h:409-412). */
double fclo_node_size(const fclo_model *m, int b) {
  const fclo_node *n = m->bvs + b;
  if (m->bv_type == FCLO_BV_OBBRSS)
    return (n->obb_extent[0] * n->obb_extent[0] +
            n->obb_extent[1] * n->obb_extent[1]) +
           n->obb_extent[2] * n->obb_extent[2];
  return sqrt(n->rss_l[0] * n->rss_l[0] + n->rss_l[1] * n->rss_l[1]) +
         2 * n->rss_r;
}
