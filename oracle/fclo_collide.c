/* fclo_collide.c — mesh-mesh collision of the oracle (test infrastructure
 * only): the 15-axis OBB separating-axis test, the 17-axis triangle SAT and
 * the recursive BVTT depth-first traversal with the reference's descent,
 * early-stop and contact-cap rules. */
#include <float.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>
#include "fclo.h"

/* math/bv/OBB-inl.h:398-523 (obbDisjoint).  B is the box-2-in-box-1 rotation,
 * T the box-2 centre in box-1 coordinates, a and b the half extents.  Returns
 * 1 when one of the 15 axes separates the boxes.  Each |B| entry is inflated
 * by reps = 1e-6 and every comparison is strict, so touching boxes overlap. */
int fclo_obb_disjoint(const double B[9], const double T[3], const double a[3],
                      const double b[3]) {
  const double reps = 1e-6;
  double Bf[9];
  for (int i = 0; i < 9; ++i) Bf[i] = fabs(B[i]) + reps;
#define b_(i, j) B[3 * (i) + (j)]
#define f_(i, j) Bf[3 * (i) + (j)]
  double t, s;
  /* box-1 face axes A0, then box-2 face axis B0, then A1, A2, B1, B2 —
   * the reference's test order */
  t = fabs(T[0]);
  if (t > (a[0] + ((f_(0, 0) * b[0] + f_(0, 1) * b[1]) + f_(0, 2) * b[2])))
    return 1;
  s = (b_(0, 0) * T[0] + b_(1, 0) * T[1]) + b_(2, 0) * T[2];
  t = fabs(s);
  if (t > (b[0] + ((f_(0, 0) * a[0] + f_(1, 0) * a[1]) + f_(2, 0) * a[2])))
    return 1;
  t = fabs(T[1]);
  if (t > (a[1] + ((f_(1, 0) * b[0] + f_(1, 1) * b[1]) + f_(1, 2) * b[2])))
    return 1;
  t = fabs(T[2]);
  if (t > (a[2] + ((f_(2, 0) * b[0] + f_(2, 1) * b[1]) + f_(2, 2) * b[2])))
    return 1;
  s = (b_(0, 1) * T[0] + b_(1, 1) * T[1]) + b_(2, 1) * T[2];
  t = fabs(s);
  if (t > (b[1] + ((f_(0, 1) * a[0] + f_(1, 1) * a[1]) + f_(2, 1) * a[2])))
    return 1;
  s = (b_(0, 2) * T[0] + b_(1, 2) * T[1]) + b_(2, 2) * T[2];
  t = fabs(s);
  if (t > (b[2] + ((f_(0, 2) * a[0] + f_(1, 2) * a[1]) + f_(2, 2) * a[2])))
    return 1;

  /* the nine edge-edge axes Ai x Bj.  For axis Ai x Bj, with (i1,i2) the
   * other two box-1 axes in cyclic order and (j1,j2) likewise for box 2:
   *   s = T[i2]*B(i1,j) - T[i1]*B(i2,j)
   *   r = a[i1]*Bf(i2,j) + a[i2]*Bf(i1,j) + b[j1]*Bf(i,j2) + b[j2]*Bf(i,j1)
   * evaluated left to right, tested in the order A0xB0, A0xB1, ..., A2xB2. */
  for (int i = 0; i < 3; ++i) {
    int i1 = (i + 1) % 3, i2 = (i + 2) % 3;
    for (int j = 0; j < 3; ++j) {
      /* the reference lists b terms with the smaller box-2 index first */
      int j1 = (j == 0) ? 1 : 0, j2 = (j == 2) ? 1 : 2;
      s = T[i2] * b_(i1, j) - T[i1] * b_(i2, j);
      t = fabs(s);
      double r = a[i1] * f_(i2, j) + a[i2] * f_(i1, j) + b[j1] * f_(i, j2) +
                 b[j2] * f_(i, j1);
      if (t > r) return 1;
    }
  }
#undef b_
#undef f_
  return 0;
}

/* math/bv/OBB-inl.h:383-395 (overlap(R0,T0,b1,b2)) via OBBRSS-inl.h:154-160:
 *   R = b1.axis^T (R0 b2.axis),  T = (R0 b2.To + T0 - b1.To)^T b1.axis. */
int fclo_obb_overlap(const double R0[9], const double T0[3],
                     const fclo_node *n1, const fclo_node *n2) {
  const double *A1 = n1->obb_axis, *A2 = n2->obb_axis;
  double R0b2[9], R[9], Tt[3], T[3];
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j)
      R0b2[3 * i + j] = (R0[3 * i + 0] * A2[0 + j] + R0[3 * i + 1] * A2[3 + j]) +
                        R0[3 * i + 2] * A2[6 + j];
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j)
      R[3 * i + j] = (A1[0 + i] * R0b2[0 + j] + A1[3 + i] * R0b2[3 + j]) +
                     A1[6 + i] * R0b2[6 + j];
  for (int i = 0; i < 3; ++i)
    Tt[i] = (((R0[3 * i + 0] * n2->obb_to[0] + R0[3 * i + 1] * n2->obb_to[1]) +
              R0[3 * i + 2] * n2->obb_to[2]) + T0[i]) - n1->obb_to[i];
  for (int j = 0; j < 3; ++j)
    T[j] = (Tt[0] * A1[0 + j] + Tt[1] * A1[3 + j]) + Tt[2] * A1[6 + j];
  return !fclo_obb_disjoint(R, T, n1->obb_extent, n2->obb_extent);
}

/* std::min / std::max as the reference calls them (NaN-propagation and
 * signed-zero choice included). */
static double smin(double a, double b) { return (b < a) ? b : a; }
static double smax(double a, double b) { return (a < b) ? b : a; }

/* traversal/collision/intersect-inl.h:1085-1106 (project6): 1 when the
 * projections of the two triangles on ax overlap (touching counts). */
static int project6(const double ax[3], const double p1[3], const double p2[3],
                    const double p3[3], const double q1[3], const double q2[3],
                    const double q3[3]) {
  double P1 = fclo_dot(ax, p1), P2 = fclo_dot(ax, p2), P3 = fclo_dot(ax, p3);
  double Q1 = fclo_dot(ax, q1), Q2 = fclo_dot(ax, q2), Q3 = fclo_dot(ax, q3);
  double mn1 = smin(P1, smin(P2, P3));
  double mx2 = smax(Q1, smax(Q2, Q3));
  if (mn1 > mx2) return 0;
  double mx1 = smax(P1, smax(P2, P3));
  double mn2 = smin(Q1, smin(Q2, Q3));
  if (mn2 > mx1) return 0;
  return 1;
}

static void sub3(const double a[3], const double b[3], double o[3]) {
  o[0] = a[0] - b[0];
  o[1] = a[1] - b[1];
  o[2] = a[2] - b[2];
}

/* traversal/collision/intersect-inl.h:597-617 then 727-845
 * (intersect_Triangle without contact output): move Q into P's frame,
 * translate both by -P1, and test the 17 candidate axes in the reference's
 * order: n1, m1, e_i x f_j (i,j = 1..3 with f3 introduced after e1xf2),
 * e_i x n1, f_i x m1. */
int fclo_intersect_triangle(const double P1[3], const double P2[3],
                            const double P3[3], const double Q1[3],
                            const double Q2[3], const double Q3[3],
                            const double R[9], const double T[3]) {
  double Qw[3][3];
  const double *Qs[3] = {Q1, Q2, Q3};
  for (int k = 0; k < 3; ++k)
    for (int i = 0; i < 3; ++i)
      Qw[k][i] = fclo_dot(R + 3 * i, Qs[k]) + T[i];
  double p1[3], p2[3], p3[3], q1[3], q2[3], q3[3];
  sub3(P1, P1, p1);
  sub3(P2, P1, p2);
  sub3(P3, P1, p3);
  sub3(Qw[0], P1, q1);
  sub3(Qw[1], P1, q2);
  sub3(Qw[2], P1, q3);
  double e1[3], e2[3], e3[3], f1[3], f2[3], f3[3], n1[3], m1[3], ax[3];
  sub3(p2, p1, e1);
  sub3(p3, p2, e2);
  fclo_cross(e1, e2, n1);
  if (!project6(n1, p1, p2, p3, q1, q2, q3)) return 0;
  sub3(q2, q1, f1);
  sub3(q3, q2, f2);
  fclo_cross(f1, f2, m1);
  if (!project6(m1, p1, p2, p3, q1, q2, q3)) return 0;
  fclo_cross(e1, f1, ax);
  if (!project6(ax, p1, p2, p3, q1, q2, q3)) return 0;
  fclo_cross(e1, f2, ax);
  if (!project6(ax, p1, p2, p3, q1, q2, q3)) return 0;
  sub3(q1, q3, f3);
  fclo_cross(e1, f3, ax);
  if (!project6(ax, p1, p2, p3, q1, q2, q3)) return 0;
  fclo_cross(e2, f1, ax);
  if (!project6(ax, p1, p2, p3, q1, q2, q3)) return 0;
  fclo_cross(e2, f2, ax);
  if (!project6(ax, p1, p2, p3, q1, q2, q3)) return 0;
  fclo_cross(e2, f3, ax);
  if (!project6(ax, p1, p2, p3, q1, q2, q3)) return 0;
  sub3(p1, p3, e3);
  fclo_cross(e3, f1, ax);
  if (!project6(ax, p1, p2, p3, q1, q2, q3)) return 0;
  fclo_cross(e3, f2, ax);
  if (!project6(ax, p1, p2, p3, q1, q2, q3)) return 0;
  fclo_cross(e3, f3, ax);
  if (!project6(ax, p1, p2, p3, q1, q2, q3)) return 0;
  fclo_cross(e1, n1, ax);
  if (!project6(ax, p1, p2, p3, q1, q2, q3)) return 0;
  fclo_cross(e2, n1, ax);
  if (!project6(ax, p1, p2, p3, q1, q2, q3)) return 0;
  fclo_cross(e3, n1, ax);
  if (!project6(ax, p1, p2, p3, q1, q2, q3)) return 0;
  fclo_cross(f1, m1, ax);
  if (!project6(ax, p1, p2, p3, q1, q2, q3)) return 0;
  fclo_cross(f2, m1, ax);
  if (!project6(ax, p1, p2, p3, q1, q2, q3)) return 0;
  fclo_cross(f3, m1, ax);
  if (!project6(ax, p1, p2, p3, q1, q2, q3)) return 0;
  return 1;
}

/* intersect-inl.h:1054-1068 buildTrianglePlane with normalize
 * (math/geometry-inl.h:413-426).  For a degenerate triangle the reference
 * leaves the plane uninitialised; here it is n = 0, t = 0. */
static void triangle_plane(const double v1[3], const double v2[3], const double v3[3],
                           double n[3], double *t) {
  double a[3], b[3];
  sub3(v2, v1, a);
  sub3(v3, v1, b);
  fclo_cross(a, b, n);
  double sq = fclo_dot(n, n);
  if (sq > 0) {
    double l = sqrt(sq);
    n[0] /= l;
    n[1] /= l;
    n[2] /= l;
    *t = fclo_dot(n, v1);
  } else {
    n[0] = n[1] = n[2] = 0;
    *t = 0;
  }
}

/* intersect-inl.h:848-885 computeDeepestPoints (epsilon 1e-5, :1116-1120). */
static void deepest_points(const double pts[3][3], const double n[3], double t,
                           double *depth, double out[3][3], unsigned *num) {
  const double eps = 1e-5;
  double max_depth = -DBL_MAX;
  unsigned k = 0, neg = 0, pos = 0, zero = 0;
  for (int i = 0; i < 3; ++i) {
    double dist = -(fclo_dot(n, pts[i]) - t);
    if (dist > eps) pos++;
    else if (dist < -eps) neg++;
    else zero++;
    if (dist > max_depth) {
      max_depth = dist;
      k = 1;
      memcpy(out[0], pts[i], sizeof(out[0]));
    } else if (dist + 1e-6 >= max_depth) {
      k++;
      memcpy(out[k - 1], pts[i], sizeof(out[0]));
    }
  }
  if (max_depth < -eps) k = 0;
  if (zero == 0 && (neg == 0 || pos == 0)) k = 0;
  *depth = max_depth;
  *num = k;
}

/* intersect-inl.h:597-617 then 727-845 with contact output: the SAT of
 * fclo_intersect_triangle, then planes of P and of Q' = R Q + T, the deepest
 * points of each triangle below the other's plane, and the shallower
 * penetration's points (at most 2), normal and depth.  Points are in the
 * frame of P.  Returns 1 if the triangles intersect. */
int fclo_intersect_triangle_contact(const double P1[3], const double P2[3],
                                    const double P3[3], const double Q1[3],
                                    const double Q2[3], const double Q3[3],
                                    const double R[9], const double T[3],
                                    double contacts[6], unsigned *num_contacts,
                                    double *depth, double normal[3]) {
  if (!fclo_intersect_triangle(P1, P2, P3, Q1, Q2, Q3, R, T)) return 0;
  double Q[3][3];
  const double *Qs[3] = {Q1, Q2, Q3};
  for (int k = 0; k < 3; ++k)
    for (int i = 0; i < 3; ++i) Q[k][i] = fclo_dot(R + 3 * i, Qs[k]) + T[i];
  double P[3][3];
  memcpy(P[0], P1, sizeof(P[0]));
  memcpy(P[1], P2, sizeof(P[0]));
  memcpy(P[2], P3, sizeof(P[0]));
  double n1[3], n2[3], t1, t2;
  triangle_plane(P[0], P[1], P[2], n1, &t1);
  triangle_plane(Q[0], Q[1], Q[2], n2, &t2);
  double dp1[3][3], dp2[3][3], pd1, pd2;
  unsigned nd1, nd2;
  deepest_points(Q, n1, t1, &pd2, dp2, &nd2);
  deepest_points(P, n2, t2, &pd1, dp1, &nd1);
  if (pd1 > pd2) {
    *num_contacts = nd2 < 2 ? nd2 : 2;
    for (unsigned i = 0; i < *num_contacts; ++i) memcpy(contacts + 3 * i, dp2[i], 3 * sizeof(double));
    memcpy(normal, n1, 3 * sizeof(double));
    *depth = pd2;
  } else {
    *num_contacts = nd1 < 2 ? nd1 : 2;
    for (unsigned i = 0; i < *num_contacts; ++i) memcpy(contacts + 3 * i, dp1[i], 3 * sizeof(double));
    normal[0] = -n2[0];
    normal[1] = -n2[1];
    normal[2] = -n2[2];
    *depth = pd1;
  }
  return 1;
}

typedef struct {
  const fclo_model *m1, *m2;
  double R[9], T[3];
  int cap;      /* contacts that may be added */
  int store;    /* pair slots available in pairs */
  int count;
  int *pairs;
  int64_t bv_tests, leaf_tests;
  /* enable_contact: geometry per stored contact, world frame */
  int contact;
  const double *tf1;
  double *pos, *nrm, *depth;
} collide_ctx;

/* isSatisfied (collision_request-inl.h:77-82) with enable_cost off: at least
 * one contact and the cap reached. */
static int can_stop(const collide_ctx *c) {
  return c->count > 0 && c->count >= c->cap;
}

static int bv_disjoint(collide_ctx *c, int b1, int b2) {
  c->bv_tests++;
  return !fclo_obb_overlap(c->R, c->T, c->m1->bvs + b1, c->m2->bvs + b2);
}

/* mesh_collision_traversal_node-inl.h:526-620 (leaf).  With enable_contact a
 * pair adds its 0-2 contacts, as many as the cap leaves room for, with
 * pos = tf1 * point, normal = tf1.linear() * normal. */
static void leaf_test(collide_ctx *c, int b1, int b2) {
  c->leaf_tests++;
  int id1 = -(c->m1->bvs[b1].first_child + 1);
  int id2 = -(c->m2->bvs[b2].first_child + 1);
  const int *t1 = c->m1->tris + 3 * id1, *t2 = c->m2->tris + 3 * id2;
  const double *v1 = c->m1->vertices, *v2 = c->m2->vertices;
  if (c->contact) {
    double pts[6], nrm[3], depth;
    unsigned n = 0;
    if (!fclo_intersect_triangle_contact(v1 + 3 * t1[0], v1 + 3 * t1[1], v1 + 3 * t1[2],
                                         v2 + 3 * t2[0], v2 + 3 * t2[1], v2 + 3 * t2[2],
                                         c->R, c->T, pts, &n, &depth, nrm))
      return;
    if ((unsigned)c->cap < (unsigned)c->count + n)
      n = c->cap > c->count ? (unsigned)(c->cap - c->count) : 0;
    const double *tf = c->tf1;
    for (unsigned i = 0; i < n; ++i) {
      if (c->count < c->store) {
        int k = c->count;
        c->pairs[2 * k] = id1;
        c->pairs[2 * k + 1] = id2;
        for (int r = 0; r < 3; ++r) {
          c->pos[3 * k + r] = fclo_dot(tf + 3 * r, pts + 3 * i) + tf[9 + r];
          c->nrm[3 * k + r] = fclo_dot(tf + 3 * r, nrm);
        }
        c->depth[k] = depth;
      }
      c->count++;
    }
    return;
  }
  if (fclo_intersect_triangle(v1 + 3 * t1[0], v1 + 3 * t1[1], v1 + 3 * t1[2],
                              v2 + 3 * t2[0], v2 + 3 * t2[1], v2 + 3 * t2[2],
                              c->R, c->T)) {
    if (c->count < c->cap) {
      if (c->count < c->store) {
        c->pairs[2 * c->count] = id1;
        c->pairs[2 * c->count + 1] = id2;
      }
      c->count++;
    }
  }
}

/* narrowphase/detail/traversal/traversal_recurse-inl.h:82-129
 * (collisionRecurse) with firstOverSecond from
 * bvh_collision_traversal_node-inl.h:77-89. */
static void collide_rec(collide_ctx *c, int b1, int b2) {
  const fclo_node *n1 = c->m1->bvs + b1, *n2 = c->m2->bvs + b2;
  int l1 = n1->first_child < 0, l2 = n2->first_child < 0;
  if (l1 && l2) {
    if (bv_disjoint(c, b1, b2)) return;
    leaf_test(c, b1, b2);
    return;
  }
  if (bv_disjoint(c, b1, b2)) return;
  double sz1 = fclo_node_size(c->m1, b1), sz2 = fclo_node_size(c->m2, b2);
  if (l2 || (!l1 && (sz1 > sz2))) {
    collide_rec(c, n1->first_child, b2);
    if (can_stop(c)) return;
    collide_rec(c, n1->first_child + 1, b2);
  } else {
    collide_rec(c, b1, n2->first_child);
    if (can_stop(c)) return;
    collide_rec(c, b1, n2->first_child + 1);
  }
}

/* ---- BVTT front lists (geometry/bvh/detail/BVH_front.h:49-72) ------------
 * A front is the list of (b1, b2) BVTT nodes where the last traversal
 * stopped: leaf pairs (recorded before their BV test) and internal pairs
 * found disjoint.  An entry is invalidated by storing b1 as -(b1 + 1). */
typedef struct {
  int *p;  /* 2 ints per entry */
  int n, cap;
} front_list;

static void front_add(front_list *f, int b1, int b2) {
  if (f->n == f->cap) {
    f->cap = f->cap ? 2 * f->cap : 64;
    f->p = (int *)realloc(f->p, 2 * (size_t)f->cap * sizeof(int));
  }
  f->p[2 * f->n] = b1;
  f->p[2 * f->n + 1] = b2;
  f->n++;
}

static int first_over_second(const collide_ctx *c, int b1, int b2) {
  const int l1 = c->m1->bvs[b1].first_child < 0, l2 = c->m2->bvs[b2].first_child < 0;
  return l2 || (!l1 && (fclo_node_size(c->m1, b1) > fclo_node_size(c->m2, b2)));
}

/* traversal_recurse-inl.h:82-129 (collisionRecurse) with a front list:
 * every leaf pair is added to the front before its BV test, an internal
 * pair when it is found disjoint, and there is no early stop (the stop rule
 * is skipped whenever a front list is given). */
static void collide_rec_front(collide_ctx *c, int b1, int b2, front_list *f) {
  const fclo_node *n1 = c->m1->bvs + b1, *n2 = c->m2->bvs + b2;
  const int l1 = n1->first_child < 0, l2 = n2->first_child < 0;
  if (l1 && l2) {
    front_add(f, b1, b2);
    if (bv_disjoint(c, b1, b2)) return;
    leaf_test(c, b1, b2);
    return;
  }
  if (bv_disjoint(c, b1, b2)) {
    front_add(f, b1, b2);
    return;
  }
  if (first_over_second(c, b1, b2)) {
    collide_rec_front(c, n1->first_child, b2, f);
    collide_rec_front(c, n1->first_child + 1, b2, f);
  } else {
    collide_rec_front(c, b1, n2->first_child, f);
    collide_rec_front(c, b1, n2->first_child + 1, f);
  }
}

/* traversal_recurse-inl.h:463-521 (propagateBVHFrontListCollisionRecurse).
 * The loop also visits the entries that descents append to the list while
 * it runs (std::list iteration reaches them): a leaf pair found that way is
 * tested once more, and its contact counted again.  Afterwards the invalid
 * entries are erased (the rest keep their order) and the re-recorded leaf
 * pairs are appended. */
static void propagate_front(collide_ctx *c, front_list *f) {
  front_list append = {0, 0, 0};
  for (int i = 0; i < f->n; ++i) {
    const int b1 = f->p[2 * i], b2 = f->p[2 * i + 1];
    const int l1 = c->m1->bvs[b1].first_child < 0, l2 = c->m2->bvs[b2].first_child < 0;
    if (l1 && l2) {
      f->p[2 * i] = -(b1 + 1);
      collide_rec_front(c, b1, b2, &append);
    } else if (!bv_disjoint(c, b1, b2)) {
      f->p[2 * i] = -(b1 + 1);
      const fclo_node *n1 = c->m1->bvs + b1, *n2 = c->m2->bvs + b2;
      if (first_over_second(c, b1, b2)) {
        collide_rec_front(c, n1->first_child, b2, f);
        collide_rec_front(c, n1->first_child + 1, b2, f);
      } else {
        collide_rec_front(c, b1, n2->first_child, f);
        collide_rec_front(c, b1, n2->first_child + 1, f);
      }
    }
  }
  int k = 0;
  for (int i = 0; i < f->n; ++i)
    if (f->p[2 * i] >= 0) {
      f->p[2 * k] = f->p[2 * i];
      f->p[2 * k + 1] = f->p[2 * i + 1];
      ++k;
    }
  f->n = k;
  for (int i = 0; i < append.n; ++i) front_add(f, append.p[2 * i], append.p[2 * i + 1]);
  free(append.p);
}

/* detail::collide(node, front_list) (collision_node-inl.h:71-82) with a
 * non-null front list, for a mesh pair: propagate the front when it holds
 * entries, else traverse from the roots recording it.  front/n_front: the
 * front in (2 ints per entry); front_out (room front_room entries) gets the
 * new front and *n_out its full size.  Contacts are capped at cap (the
 * traversal goes on: no early stop), their first min(count, store) pairs go
 * to pairs. */
void fclo_collide_front(const fclo_model *m1, const double tf1[12], const fclo_model *m2,
                        const double tf2[12], int cap, int store, int *pairs, const int *front,
                        int n_front, int *front_out, int front_room, int *n_out,
                        fclo_collide_stats *stats) {
  collide_ctx c;
  memset(&c, 0, sizeof(c));
  c.m1 = m1;
  c.m2 = m2;
  c.cap = cap;
  c.store = store;
  c.pairs = pairs;
  fclo_relative_transform(tf1, tf2, c.R, c.T);
  front_list f = {0, 0, 0};
  for (int i = 0; i < n_front; ++i) front_add(&f, front[2 * i], front[2 * i + 1]);
  if (f.n > 0)
    propagate_front(&c, &f);
  else
    collide_rec_front(&c, 0, 0, &f);
  *n_out = f.n;
  for (int i = 0; i < f.n && i < front_room; ++i) {
    front_out[2 * i] = f.p[2 * i];
    front_out[2 * i + 1] = f.p[2 * i + 1];
  }
  free(f.p);
  stats->num_contacts = c.count;
  stats->num_bv_tests = c.bv_tests;
  stats->num_leaf_tests = c.leaf_tests;
}

static void collide_store(const fclo_model *m1, const double tf1[12],
                          const fclo_model *m2, const double tf2[12], int cap,
                          int store, int *pairs, double *pos, double *nrm,
                          double *depth, fclo_collide_stats *stats) {
  collide_ctx c;
  memset(&c, 0, sizeof(c));
  c.m1 = m1;
  c.m2 = m2;
  c.cap = cap;
  c.store = store;
  c.pairs = pairs;
  c.contact = pos != 0;
  c.tf1 = tf1;
  c.pos = pos;
  c.nrm = nrm;
  c.depth = depth;
  /* mesh_collision_traversal_node-inl.h:742 */
  fclo_relative_transform(tf1, tf2, c.R, c.T);
  if (cap > 0) collide_rec(&c, 0, 0);
  stats->num_contacts = c.count;
  stats->num_bv_tests = c.bv_tests;
  stats->num_leaf_tests = c.leaf_tests;
}

void fclo_collide(const fclo_model *m1, const double tf1[12],
                  const fclo_model *m2, const double tf2[12], int cap,
                  int *pairs, fclo_collide_stats *stats) {
  collide_store(m1, tf1, m2, tf2, cap, cap, pairs, 0, 0, 0, stats);
}

void fclo_collide_contact(const fclo_model *m1, const double tf1[12],
                          const fclo_model *m2, const double tf2[12], int cap,
                          int store, int *pairs, double *pos, double *normal,
                          double *depth, fclo_collide_stats *stats) {
  collide_store(m1, tf1, m2, tf2, cap, store, pairs, pos, normal, depth, stats);
}

void fclo_collide_batch(const fclo_model *const *models, int num_queries,
                        const int *q_m1, const int *q_m2, const double *tf1s,
                        const double *tf2s, int cap, int store, int *counts,
                        int *pairs, int64_t *bv_tests, int64_t *leaf_tests,
                        int threads) {
#pragma omp parallel for schedule(dynamic, 64) num_threads(threads)
  for (int q = 0; q < num_queries; ++q) {
    fclo_collide_stats st;
    collide_store(models[q_m1[q]], tf1s + 12 * (size_t)q, models[q_m2[q]],
                  tf2s + 12 * (size_t)q, cap, store,
                  pairs ? pairs + 2 * (size_t)q * (size_t)store : 0, 0, 0, 0, &st);
    counts[q] = st.num_contacts;
    if (bv_tests) bv_tests[q] = st.num_bv_tests;
    if (leaf_tests) leaf_tests[q] = st.num_leaf_tests;
  }
}

/* fclo_collide_batch with every query's records in one flat array: query
 * q's first min(count, offsets[q+1] - offsets[q]) pairs, in the reference's
 * order, go to pairs + 2 * offsets[q].  Callers size the room per query
 * from a counting pass (store = 0). */
void fclo_collide_batch_flat(const fclo_model *const *models, int num_queries,
                             const int *q_m1, const int *q_m2, const double *tf1s,
                             const double *tf2s, int cap, const int64_t *offsets,
                             int *counts, int *pairs, int threads) {
#pragma omp parallel for schedule(dynamic, 16) num_threads(threads)
  for (int q = 0; q < num_queries; ++q) {
    fclo_collide_stats st;
    const int64_t room = offsets[q + 1] - offsets[q];
    collide_store(models[q_m1[q]], tf1s + 12 * (size_t)q, models[q_m2[q]],
                  tf2s + 12 * (size_t)q, cap, (int)room, pairs + 2 * offsets[q], 0, 0, 0,
                  &st);
    counts[q] = st.num_contacts;
  }
}

/* All triangle pairs with no hierarchy (test of the traversal's culling):
 * writes every intersecting (id1, id2) in row-major order, returns the count
 * (pairs beyond store are counted but not written). */
int64_t fclo_collide_all_pairs(const fclo_model *m1, const double tf1[12],
                               const fclo_model *m2, const double tf2[12],
                               int *pairs, int64_t store) {
  double R[9], T[3];
  fclo_relative_transform(tf1, tf2, R, T);
  int64_t n = 0;
  const double *v1 = m1->vertices, *v2 = m2->vertices;
  for (int i = 0; i < m1->num_tris; ++i) {
    const int *t1 = m1->tris + 3 * i;
    for (int j = 0; j < m2->num_tris; ++j) {
      const int *t2 = m2->tris + 3 * j;
      if (fclo_intersect_triangle(v1 + 3 * t1[0], v1 + 3 * t1[1], v1 + 3 * t1[2],
                                  v2 + 3 * t2[0], v2 + 3 * t2[1], v2 + 3 * t2[2], R,
                                  T)) {
        if (n < store) {
          pairs[2 * n] = i;
          pairs[2 * n + 1] = j;
        }
        n++;
      }
    }
  }
  return n;
}
