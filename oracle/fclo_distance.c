/* fclo_distance.c — mesh-mesh distance of the oracle (test infrastructure
 * only): PQP rectangle-rectangle distance for RSS bounds, PQP
 * triangle-triangle distance for leaves, and the recursive distance
 * traversal with nearer-child-first descent and lower-bound pruning. */
#include <float.h>
#include <math.h>
#include <string.h>
#include "fclo.h"

/* math/bv/RSS-inl.h:449-454 */
static void clip_to_range(double *val, double a, double b) {
  if (*val < a) *val = a;
  else if (*val > b) *val = b;
}

/* math/bv/RSS-inl.h:457-482 (segCoords) */
static void seg_coords(double *t, double *u, double a, double b, double AdB,
                       double AdT, double BdT) {
  double denom = 1 - AdB * AdB;
  if (denom == 0) {
    *t = 0;
  } else {
    *t = (AdT - BdT * AdB) / denom;
    clip_to_range(t, 0.0, a);
  }
  *u = *t * AdB - BdT;
  if (*u < 0) {
    *u = 0;
    *t = AdT;
    clip_to_range(t, 0.0, a);
  } else if (*u > b) {
    *u = b;
    *t = *u * AdB + AdT;
    clip_to_range(t, 0.0, a);
  }
}

/* math/bv/RSS-inl.h:485-509 (inVoronoi) */
static int in_voronoi(double a, double b, double AnB, double AnT, double AdB,
                      double AdT, double BdT) {
  if (fabs(AnB) < 1e-7) return 0;
  double u = -AnT / AnB;
  clip_to_range(&u, 0.0, b);
  double t = u * AdB + AdT;
  clip_to_range(&t, 0.0, a);
  double v = t * AdB - BdT;
  if (AnB > 0) {
    if (v > (u + 1e-7)) return 1;
  } else {
    if (v < (u - 1e-7)) return 1;
  }
  return 0;
}

typedef struct {
  const double *Rab, *Tab, *a, *b;
} rect_pair;

/* Closest points of one A-edge / B-edge pair (the tail shared by the 16
 * edge-pair cases of RSS-inl.h:512-1137).  The A edge runs along A axis ia
 * and sits at A-coordinate a[1-ia] when ua is set (0 otherwise); the B edge
 * runs along B axis ib at B-coordinate b[1-ib] when ub is set.
 *   D_k = ((Tab_k [+ Rab(k,1-ib)*b[1-ib]]) + Rab(k,ib)*u) [- a[1-ia] | - t] */
static double edge_pair(const rect_pair *r, int ia, int ua, int ib, int ub,
                        double AdB, double AdT, double BdT, double P[3],
                        double Q[3]) {
  double t, u, D[3];
  seg_coords(&t, &u, r->a[ia], r->b[ib], AdB, AdT, BdT);
  for (int k = 0; k < 3; ++k) {
    double d = r->Tab[k];
    if (ub) d = d + r->Rab[3 * k + (1 - ib)] * r->b[1 - ib];
    d = d + r->Rab[3 * k + ib] * u;
    if (k == ia) d = d - t;
    else if (k == 1 - ia && ua) d = d - r->a[1 - ia];
    D[k] = d;
  }
  if (P && Q) {
    double pa[3];
    pa[ia] = t;
    pa[1 - ia] = ua ? r->a[1 - ia] : 0;
    pa[2] = 0;
    for (int k = 0; k < 3; ++k) {
      P[k] = pa[k];
      Q[k] = D[k] + pa[k];
    }
  }
  return sqrt((D[0] * D[0] + D[1] * D[1]) + D[2] * D[2]);
}

/* math/bv/RSS-inl.h:512-1224 (rectDistance).  Distance between rectangle A
 * = [0,a0]x[0,a1] and rectangle B placed by (Rab, Tab).  The 16 edge-pair
 * cases are tried in the reference's order with its exact argument
 * expressions, then the face-normal separation fallback. */
double fclo_rect_distance(const double Rab[9], const double Tab[3],
                          const double a[2], const double b[2], double P[3],
                          double Q[3]) {
  const double A0B0 = Rab[0], A0B1 = Rab[1], A1B0 = Rab[3], A1B1 = Rab[4];
  const double aA0B0 = a[0] * A0B0, aA0B1 = a[0] * A0B1;
  const double aA1B0 = a[1] * A1B0, aA1B1 = a[1] * A1B1;
  const double bA0B0 = b[0] * A0B0, bA1B0 = b[0] * A1B0;
  const double bA0B1 = b[1] * A0B1, bA1B1 = b[1] * A1B1;
  double Tba[3];
  for (int i = 0; i < 3; ++i)
    Tba[i] = (Rab[0 + i] * Tab[0] + Rab[3 + i] * Tab[1]) + Rab[6 + i] * Tab[2];
  rect_pair rp = {Rab, Tab, a, b};

  /* x-extents of A's corners projected on B0, and of B's corners on A0 */
  double ALL_x = -Tba[0];
  double ALU_x = ALL_x + aA1B0;
  double AUL_x = ALL_x + aA0B0;
  double AUU_x = ALU_x + aA0B0;
  double LA1_lx, LA1_ux, UA1_lx, UA1_ux;
  if (ALL_x < ALU_x) {
    LA1_lx = ALL_x; LA1_ux = ALU_x; UA1_lx = AUL_x; UA1_ux = AUU_x;
  } else {
    LA1_lx = ALU_x; LA1_ux = ALL_x; UA1_lx = AUU_x; UA1_ux = AUL_x;
  }
  double BLL_x = Tab[0];
  double BLU_x = BLL_x + bA0B1;
  double BUL_x = BLL_x + bA0B0;
  double BUU_x = BLU_x + bA0B0;
  double LB1_lx, LB1_ux, UB1_lx, UB1_ux;
  if (BLL_x < BLU_x) {
    LB1_lx = BLL_x; LB1_ux = BLU_x; UB1_lx = BUL_x; UB1_ux = BUU_x;
  } else {
    LB1_lx = BLU_x; LB1_ux = BLL_x; UB1_lx = BUU_x; UB1_ux = BUL_x;
  }

  /* A1-edges vs B1-edges */
  if ((UA1_ux > b[0]) && (UB1_ux > a[0])) {
    if (((UA1_lx > b[0]) ||
         in_voronoi(b[1], a[1], A1B0, aA0B0 - b[0] - Tba[0], A1B1,
                    aA0B1 - Tba[1], -Tab[1] - bA1B0)) &&
        ((UB1_lx > a[0]) ||
         in_voronoi(a[1], b[1], A0B1, Tab[0] + bA0B0 - a[0], A1B1,
                    Tab[1] + bA1B0, Tba[1] - aA0B1)))
      return edge_pair(&rp, 1, 1, 1, 1, A1B1, Tab[1] + bA1B0, Tba[1] - aA0B1,
                       P, Q);
  }
  if ((UA1_lx < 0) && (LB1_ux > a[0])) {
    if (((UA1_ux < 0) ||
         in_voronoi(b[1], a[1], -A1B0, Tba[0] - aA0B0, A1B1, aA0B1 - Tba[1],
                    -Tab[1])) &&
        ((LB1_lx > a[0]) ||
         in_voronoi(a[1], b[1], A0B1, Tab[0] - a[0], A1B1, Tab[1],
                    Tba[1] - aA0B1)))
      return edge_pair(&rp, 1, 1, 1, 0, A1B1, Tab[1], Tba[1] - aA0B1, P, Q);
  }
  if ((LA1_ux > b[0]) && (UB1_lx < 0)) {
    if (((LA1_lx > b[0]) ||
         in_voronoi(b[1], a[1], A1B0, -Tba[0] - b[0], A1B1, -Tba[1],
                    -Tab[1] - bA1B0)) &&
        ((UB1_ux < 0) ||
         in_voronoi(a[1], b[1], -A0B1, -Tab[0] - bA0B0, A1B1, Tab[1] + bA1B0,
                    Tba[1])))
      return edge_pair(&rp, 1, 0, 1, 1, A1B1, Tab[1] + bA1B0, Tba[1], P, Q);
  }
  if ((LA1_lx < 0) && (LB1_lx < 0)) {
    if (((LA1_ux < 0) ||
         in_voronoi(b[1], a[1], -A1B0, Tba[0], A1B1, -Tba[1], -Tab[1])) &&
        ((LB1_ux < 0) ||
         in_voronoi(a[1], b[1], -A0B1, -Tab[0], A1B1, Tab[1], Tba[1])))
      return edge_pair(&rp, 1, 0, 1, 0, A1B1, Tab[1], Tba[1], P, Q);
  }

  double ALL_y = -Tba[1];
  double ALU_y = ALL_y + aA1B1;
  double AUL_y = ALL_y + aA0B1;
  double AUU_y = ALU_y + aA0B1;
  double LA1_ly, LA1_uy, UA1_ly, UA1_uy, LB0_lx, LB0_ux, UB0_lx, UB0_ux;
  if (ALL_y < ALU_y) {
    LA1_ly = ALL_y; LA1_uy = ALU_y; UA1_ly = AUL_y; UA1_uy = AUU_y;
  } else {
    LA1_ly = ALU_y; LA1_uy = ALL_y; UA1_ly = AUU_y; UA1_uy = AUL_y;
  }
  if (BLL_x < BUL_x) {
    LB0_lx = BLL_x; LB0_ux = BUL_x; UB0_lx = BLU_x; UB0_ux = BUU_x;
  } else {
    LB0_lx = BUL_x; LB0_ux = BLL_x; UB0_lx = BUU_x; UB0_ux = BLU_x;
  }

  /* A1-edges vs B0-edges */
  if ((UA1_uy > b[1]) && (UB0_ux > a[0])) {
    if (((UA1_ly > b[1]) ||
         in_voronoi(b[0], a[1], A1B1, aA0B1 - Tba[1] - b[1], A1B0,
                    aA0B0 - Tba[0], -Tab[1] - bA1B1)) &&
        ((UB0_lx > a[0]) ||
         in_voronoi(a[1], b[0], A0B0, Tab[0] - a[0] + bA0B1, A1B0,
                    Tab[1] + bA1B1, Tba[0] - aA0B0)))
      return edge_pair(&rp, 1, 1, 0, 1, A1B0, Tab[1] + bA1B1, Tba[0] - aA0B0,
                       P, Q);
  }
  if ((UA1_ly < 0) && (LB0_ux > a[0])) {
    if (((UA1_uy < 0) ||
         in_voronoi(b[0], a[1], -A1B1, Tba[1] - aA0B1, A1B0, aA0B0 - Tba[0],
                    -Tab[1])) &&
        ((LB0_lx > a[0]) ||
         in_voronoi(a[1], b[0], A0B0, Tab[0] - a[0], A1B0, Tab[1],
                    Tba[0] - aA0B0)))
      return edge_pair(&rp, 1, 1, 0, 0, A1B0, Tab[1], Tba[0] - aA0B0, P, Q);
  }
  if ((LA1_uy > b[1]) && (UB0_lx < 0)) {
    if (((LA1_ly > b[1]) ||
         in_voronoi(b[0], a[1], A1B1, -Tba[1] - b[1], A1B0, -Tba[0],
                    -Tab[1] - bA1B1)) &&
        ((UB0_ux < 0) ||
         in_voronoi(a[1], b[0], -A0B0, -Tab[0] - bA0B1, A1B0, Tab[1] + bA1B1,
                    Tba[0])))
      return edge_pair(&rp, 1, 0, 0, 1, A1B0, Tab[1] + bA1B1, Tba[0], P, Q);
  }
  if ((LA1_ly < 0) && (LB0_lx < 0)) {
    if (((LA1_uy < 0) ||
         in_voronoi(b[0], a[1], -A1B1, Tba[1], A1B0, -Tba[0], -Tab[1])) &&
        ((LB0_ux < 0) ||
         in_voronoi(a[1], b[0], -A0B0, -Tab[0], A1B0, Tab[1], Tba[0])))
      return edge_pair(&rp, 1, 0, 0, 0, A1B0, Tab[1], Tba[0], P, Q);
  }

  double BLL_y = Tab[1];
  double BLU_y = BLL_y + bA1B1;
  double BUL_y = BLL_y + bA1B0;
  double BUU_y = BLU_y + bA1B0;
  double LA0_lx, LA0_ux, UA0_lx, UA0_ux, LB1_ly, LB1_uy, UB1_ly, UB1_uy;
  if (ALL_x < AUL_x) {
    LA0_lx = ALL_x; LA0_ux = AUL_x; UA0_lx = ALU_x; UA0_ux = AUU_x;
  } else {
    LA0_lx = AUL_x; LA0_ux = ALL_x; UA0_lx = AUU_x; UA0_ux = ALU_x;
  }
  if (BLL_y < BLU_y) {
    LB1_ly = BLL_y; LB1_uy = BLU_y; UB1_ly = BUL_y; UB1_uy = BUU_y;
  } else {
    LB1_ly = BLU_y; LB1_uy = BLL_y; UB1_ly = BUU_y; UB1_uy = BUL_y;
  }

  /* A0-edges vs B1-edges */
  if ((UA0_ux > b[0]) && (UB1_uy > a[1])) {
    if (((UA0_lx > b[0]) ||
         in_voronoi(b[1], a[0], A0B0, aA1B0 - Tba[0] - b[0], A0B1,
                    aA1B1 - Tba[1], -Tab[0] - bA0B0)) &&
        ((UB1_ly > a[1]) ||
         in_voronoi(a[0], b[1], A1B1, Tab[1] - a[1] + bA1B0, A0B1,
                    Tab[0] + bA0B0, Tba[1] - aA1B1)))
      return edge_pair(&rp, 0, 1, 1, 1, A0B1, Tab[0] + bA0B0, Tba[1] - aA1B1,
                       P, Q);
  }
  if ((UA0_lx < 0) && (LB1_uy > a[1])) {
    if (((UA0_ux < 0) ||
         in_voronoi(b[1], a[0], -A0B0, Tba[0] - aA1B0, A0B1, aA1B1 - Tba[1],
                    -Tab[0])) &&
        ((LB1_ly > a[1]) ||
         in_voronoi(a[0], b[1], A1B1, Tab[1] - a[1], A0B1, Tab[0],
                    Tba[1] - aA1B1)))
      return edge_pair(&rp, 0, 1, 1, 0, A0B1, Tab[0], Tba[1] - aA1B1, P, Q);
  }
  if ((LA0_ux > b[0]) && (UB1_ly < 0)) {
    if (((LA0_lx > b[0]) ||
         in_voronoi(b[1], a[0], A0B0, -b[0] - Tba[0], A0B1, -Tba[1],
                    -bA0B0 - Tab[0])) &&
        ((UB1_uy < 0) ||
         in_voronoi(a[0], b[1], -A1B1, -Tab[1] - bA1B0, A0B1, Tab[0] + bA0B0,
                    Tba[1])))
      return edge_pair(&rp, 0, 0, 1, 1, A0B1, Tab[0] + bA0B0, Tba[1], P, Q);
  }
  if ((LA0_lx < 0) && (LB1_ly < 0)) {
    if (((LA0_ux < 0) ||
         in_voronoi(b[1], a[0], -A0B0, Tba[0], A0B1, -Tba[1], -Tab[0])) &&
        ((LB1_uy < 0) ||
         in_voronoi(a[0], b[1], -A1B1, -Tab[1], A0B1, Tab[0], Tba[1])))
      return edge_pair(&rp, 0, 0, 1, 0, A0B1, Tab[0], Tba[1], P, Q);
  }

  double LA0_ly, LA0_uy, UA0_ly, UA0_uy, LB0_ly, LB0_uy, UB0_ly, UB0_uy;
  if (ALL_y < AUL_y) {
    LA0_ly = ALL_y; LA0_uy = AUL_y; UA0_ly = ALU_y; UA0_uy = AUU_y;
  } else {
    LA0_ly = AUL_y; LA0_uy = ALL_y; UA0_ly = AUU_y; UA0_uy = ALU_y;
  }
  if (BLL_y < BUL_y) {
    LB0_ly = BLL_y; LB0_uy = BUL_y; UB0_ly = BLU_y; UB0_uy = BUU_y;
  } else {
    LB0_ly = BUL_y; LB0_uy = BLL_y; UB0_ly = BUU_y; UB0_uy = BLU_y;
  }

  /* A0-edges vs B0-edges */
  if ((UA0_uy > b[1]) && (UB0_uy > a[1])) {
    if (((UA0_ly > b[1]) ||
         in_voronoi(b[0], a[0], A0B1, aA1B1 - Tba[1] - b[1], A0B0,
                    aA1B0 - Tba[0], -Tab[0] - bA0B1)) &&
        ((UB0_ly > a[1]) ||
         in_voronoi(a[0], b[0], A1B0, Tab[1] - a[1] + bA1B1, A0B0,
                    Tab[0] + bA0B1, Tba[0] - aA1B0)))
      return edge_pair(&rp, 0, 1, 0, 1, A0B0, Tab[0] + bA0B1, Tba[0] - aA1B0,
                       P, Q);
  }
  if ((UA0_ly < 0) && (LB0_uy > a[1])) {
    if (((UA0_uy < 0) ||
         in_voronoi(b[0], a[0], -A0B1, Tba[1] - aA1B1, A0B0, aA1B0 - Tba[0],
                    -Tab[0])) &&
        ((LB0_ly > a[1]) ||
         in_voronoi(a[0], b[0], A1B0, Tab[1] - a[1], A0B0, Tab[0],
                    Tba[0] - aA1B0)))
      return edge_pair(&rp, 0, 1, 0, 0, A0B0, Tab[0], Tba[0] - aA1B0, P, Q);
  }
  if ((LA0_uy > b[1]) && (UB0_ly < 0)) {
    if (((LA0_ly > b[1]) ||
         in_voronoi(b[0], a[0], A0B1, -Tba[1] - b[1], A0B0, -Tba[0],
                    -Tab[0] - bA0B1)) &&
        ((UB0_uy < 0) ||
         in_voronoi(a[0], b[0], -A1B0, -Tab[1] - bA1B1, A0B0, Tab[0] + bA0B1,
                    Tba[0])))
      return edge_pair(&rp, 0, 0, 0, 1, A0B0, Tab[0] + bA0B1, Tba[0], P, Q);
  }
  if ((LA0_ly < 0) && (LB0_ly < 0)) {
    if (((LA0_uy < 0) ||
         in_voronoi(b[0], a[0], -A0B1, Tba[1], A0B0, -Tba[0], -Tab[0])) &&
        ((LB0_uy < 0) ||
         in_voronoi(a[0], b[0], -A1B0, -Tab[1], A0B0, Tab[0], Tba[0])))
      return edge_pair(&rp, 0, 0, 0, 0, A0B0, Tab[0], Tba[0], P, Q);
  }

  /* no edge pair holds the closest points: separation along face normals
   * (RSS-inl.h:1139-1224) */
  double sep1, sep2;
  if (Tab[2] > 0.0) {
    sep1 = Tab[2];
    if (Rab[6] < 0.0) sep1 += b[0] * Rab[6];
    if (Rab[7] < 0.0) sep1 += b[1] * Rab[7];
  } else {
    sep1 = -Tab[2];
    if (Rab[6] > 0.0) sep1 -= b[0] * Rab[6];
    if (Rab[7] > 0.0) sep1 -= b[1] * Rab[7];
  }
  if (Tba[2] < 0) {
    sep2 = -Tba[2];
    if (Rab[2] < 0.0) sep2 += a[0] * Rab[2];
    if (Rab[5] < 0.0) sep2 += a[1] * Rab[5];
  } else {
    sep2 = Tba[2];
    if (Rab[2] > 0.0) sep2 -= a[0] * Rab[2];
    if (Rab[5] > 0.0) sep2 -= a[1] * Rab[5];
  }
  if (sep1 >= sep2 && sep1 >= 0) {
    if (P && Q) {
      Q[0] = 0;
      Q[1] = 0;
      Q[2] = (Tab[2] > 0) ? sep1 : -sep1;
      P[0] = P[1] = P[2] = 0;
    }
  }
  if (sep2 >= sep1 && sep2 >= 0) {
    double Pf[3];
    double sg = (Tba[2] < 0) ? 1.0 : -1.0;
    for (int k = 0; k < 3; ++k) Pf[k] = (sg * Rab[3 * k + 2]) * sep2 + Tab[k];
    if (P && Q) {
      for (int k = 0; k < 3; ++k) {
        P[k] = Pf[k];
        Q[k] = Tab[k];
      }
    }
  }
  double sep = (sep1 > sep2 ? sep1 : sep2);
  return (sep > 0 ? sep : 0);
}

/* math/bv/RSS-inl.h:1956-1974 (distance(R0,T0,b1,b2)): relative frame of
 * the two rectangles, rectangle distance minus both radii, clamped at 0. */
double fclo_rss_distance(const double R0[9], const double T0[3],
                         const fclo_node *n1, const fclo_node *n2) {
  const double *A1 = n1->rss_axis, *A2 = n2->rss_axis;
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
    Tt[i] = (((R0[3 * i + 0] * n2->rss_to[0] + R0[3 * i + 1] * n2->rss_to[1]) +
              R0[3 * i + 2] * n2->rss_to[2]) + T0[i]) - n1->rss_to[i];
  for (int j = 0; j < 3; ++j)
    T[j] = (Tt[0] * A1[0 + j] + Tt[1] * A1[3 + j]) + Tt[2] * A1[6 + j];
  double dist = fclo_rect_distance(R, T, n1->rss_l, n2->rss_l, 0, 0);
  dist -= (n1->rss_r + n2->rss_r);
  return (dist < 0.0) ? 0.0 : dist;
}

static void sub3(const double a[3], const double b[3], double o[3]) {
  o[0] = a[0] - b[0];
  o[1] = a[1] - b[1];
  o[2] = a[2] - b[2];
}
static void add3(const double a[3], const double b[3], double o[3]) {
  o[0] = a[0] + b[0];
  o[1] = a[1] + b[1];
  o[2] = a[2] + b[2];
}
static void axpy3(const double p[3], const double a[3], double s, double o[3]) {
  o[0] = p[0] + a[0] * s;
  o[1] = p[1] + a[1] * s;
  o[2] = p[2] + a[2] * s;
}
static void copy3(const double a[3], double o[3]) {
  o[0] = a[0];
  o[1] = a[1];
  o[2] = a[2];
}

/* primitive_shape_algorithm/triangle_distance-inl.h:54-167 (segPoints):
 * closest points X on segment P+sA and Y on Q+tB, and the vector VEC used
 * for the slab test. */
static void seg_points(const double P[3], const double A[3], const double Q[3],
                       const double B[3], double VEC[3], double X[3],
                       double Y[3]) {
  double T[3], TMP[3];
  sub3(Q, P, T);
  double AdA = fclo_dot(A, A), BdB = fclo_dot(B, B), AdB = fclo_dot(A, B);
  double AdT = fclo_dot(A, T), BdT = fclo_dot(B, T);
  double denom = AdA * BdB - AdB * AdB;
  double t = (AdT * BdB - BdT * AdB) / denom;
  if ((t < 0) || isnan(t)) t = 0;
  else if (t > 1) t = 1;
  double u = (t * AdB - BdT) / BdB;
  if ((u <= 0) || isnan(u)) {
    copy3(Q, Y);
    t = AdT / AdA;
    if ((t <= 0) || isnan(t)) {
      copy3(P, X);
      sub3(Q, P, VEC);
    } else if (t >= 1) {
      add3(P, A, X);
      sub3(Q, X, VEC);
    } else {
      axpy3(P, A, t, X);
      fclo_cross(T, A, TMP);
      fclo_cross(A, TMP, VEC);
    }
  } else if (u >= 1) {
    add3(Q, B, Y);
    t = (AdB + AdT) / AdA;
    if ((t <= 0) || isnan(t)) {
      copy3(P, X);
      sub3(Y, P, VEC);
    } else if (t >= 1) {
      add3(P, A, X);
      sub3(Y, X, VEC);
    } else {
      axpy3(P, A, t, X);
      sub3(Y, P, T);
      fclo_cross(T, A, TMP);
      fclo_cross(A, TMP, VEC);
    }
  } else {
    axpy3(Q, B, u, Y);
    if ((t <= 0) || isnan(t)) {
      copy3(P, X);
      fclo_cross(T, B, TMP);
      fclo_cross(B, TMP, VEC);
    } else if (t >= 1) {
      add3(P, A, X);
      sub3(Q, X, T);
      fclo_cross(T, B, TMP);
      fclo_cross(B, TMP, VEC);
    } else {
      axpy3(P, A, t, X);
      fclo_cross(A, B, VEC);
      if (fclo_dot(VEC, T) < 0) {
        VEC[0] = VEC[0] * (-1);
        VEC[1] = VEC[1] * (-1);
        VEC[2] = VEC[2] * (-1);
      }
    }
  }
}

/* Vertex-face check of triDistance (triangle_distance-inl.h:258-379): if the
 * normal n of triangle F separates it from triangle G, the G vertex nearest
 * to F's plane may project inside F.  sgn_first selects the reference's
 * orientation (first block: T1 face / T2 vertex, second block: T2 face / T1
 * vertex).  Returns 1 and sets P,Q,*dist when the vertex projects inside. */
static int vertex_face(const double F[9], const double G[9], const double E[9],
                       int face_is_first, int *shown_disjoint, double P[3],
                       double Q[3], double *dist) {
  double n[3];
  fclo_cross(E + 0, E + 3, n);
  double nl = fclo_dot(n, n);
  if (!(nl > 1e-15)) return 0;
  double pr[3], V[3], Z[3];
  for (int k = 0; k < 3; ++k) {
    if (face_is_first) sub3(F + 0, G + 3 * k, V);  /* T1[0] - T2[k] */
    else sub3(F + 0, G + 3 * k, V);                /* T2[0] - T1[k] */
    pr[k] = fclo_dot(V, n);
  }
  int point = -1;
  if ((pr[0] > 0) && (pr[1] > 0) && (pr[2] > 0)) {
    if (pr[0] < pr[1]) point = 0; else point = 1;
    if (pr[2] < pr[point]) point = 2;
  } else if ((pr[0] < 0) && (pr[1] < 0) && (pr[2] < 0)) {
    if (pr[0] > pr[1]) point = 0; else point = 1;
    if (pr[2] > pr[point]) point = 2;
  }
  if (point < 0) return 0;
  *shown_disjoint = 1;
  const double *g = G + 3 * point;
  for (int e = 0; e < 3; ++e) {
    sub3(g, F + 3 * e, V);
    fclo_cross(n, E + 3 * e, Z);
    if (!(fclo_dot(V, Z) > 0)) return 0;
  }
  double onf[3];
  axpy3(g, n, pr[point] / nl, onf);
  if (face_is_first) {
    copy3(onf, P);
    copy3(g, Q);
  } else {
    copy3(g, P);
    copy3(onf, Q);
  }
  double d[3];
  sub3(P, Q, d);
  *dist = sqrt(fclo_dot(d, d));
  return 1;
}

/* primitive_shape_algorithm/triangle_distance-inl.h:170-394 (triDistance):
 * S and T are 3 vertices each (9 doubles).  Returns the distance and the
 * closest points P (on S) and Q (on T); 0 with P,Q unspecified when the
 * triangles overlap. */
double fclo_tri_distance(const double S[9], const double T[9], double P[3],
                         double Q[3]) {
  double Sv[9], Tv[9], VEC[3], V[3], Z[3];
  double minP[3] = {0, 0, 0}, minQ[3] = {0, 0, 0};
  sub3(S + 3, S + 0, Sv + 0);
  sub3(S + 6, S + 3, Sv + 3);
  sub3(S + 0, S + 6, Sv + 6);
  sub3(T + 3, T + 0, Tv + 0);
  sub3(T + 6, T + 3, Tv + 3);
  sub3(T + 0, T + 6, Tv + 6);
  int shown_disjoint = 0;
  sub3(S, T, V);
  double mindd = fclo_dot(V, V) + 1;
  for (int i = 0; i < 3; ++i) {
    for (int j = 0; j < 3; ++j) {
      seg_points(S + 3 * i, Sv + 3 * i, T + 3 * j, Tv + 3 * j, VEC, P, Q);
      sub3(Q, P, V);
      double dd = fclo_dot(V, V);
      if (dd <= mindd) {
        copy3(P, minP);
        copy3(Q, minQ);
        mindd = dd;
        sub3(S + 3 * ((i + 2) % 3), P, Z);
        double a = fclo_dot(Z, VEC);
        sub3(T + 3 * ((j + 2) % 3), Q, Z);
        double b = fclo_dot(Z, VEC);
        if ((a <= 0) && (b >= 0)) return sqrt(dd);
        double p = fclo_dot(V, VEC);
        if (a < 0) a = 0;
        if (b > 0) b = 0;
        if ((p - a + b) > 0) shown_disjoint = 1;
      }
    }
  }
  double dist;
  if (vertex_face(S, T, Sv, 1, &shown_disjoint, P, Q, &dist)) return dist;
  if (vertex_face(T, S, Tv, 0, &shown_disjoint, P, Q, &dist)) return dist;
  if (shown_disjoint) {
    copy3(minP, P);
    copy3(minQ, Q);
    return sqrt(mindd);
  }
  return 0;
}

typedef struct {
  const fclo_model *m1, *m2;
  double tf[12];    /* model-2 frame expressed in model-1 frame */
  int nearest;
  double min_distance;
  double p1[3], p2[3];
  int b1, b2;
  int64_t bv_tests, leaf_tests;
} dist_ctx;

/* DistanceResult::update (distance_result-inl.h:84-105): strict improvement. */
static void update(dist_ctx *c, double d, int id1, int id2, const double P[3],
                   const double Q[3]) {
  if (c->min_distance > d) {
    c->min_distance = d;
    c->b1 = id1;
    c->b2 = id2;
    if (c->nearest) {
      copy3(P, c->p1);
      copy3(Q, c->p2);
    }
  }
}

static void tri_pts(const fclo_model *m, int id, double out[9]) {
  const int *t = m->tris + 3 * id;
  for (int k = 0; k < 3; ++k) copy3(m->vertices + 3 * t[k], out + 3 * k);
}

/* tf * p for an isometry: R p + t */
static void xform(const double tf[12], const double p[3], double o[3]) {
  double r[3];
  for (int i = 0; i < 3; ++i) r[i] = fclo_dot(tf + 3 * i, p) + tf[9 + i];
  copy3(r, o);
}

/* mesh_distance_traversal_node-inl.h:452-499: triangle pair distance with
 * model-2's triangle moved into model-1's frame. */
static void dist_leaf(dist_ctx *c, int id1, int id2) {
  double S[9], T[9], Tw[9], P[3], Q[3];
  tri_pts(c->m1, id1, S);
  tri_pts(c->m2, id2, T);
  for (int k = 0; k < 3; ++k) xform(c->tf, T + 3 * k, Tw + 3 * k);
  double d = fclo_tri_distance(S, Tw, P, Q);
  update(c, d, id1, id2, P, Q);
}

static double bv_dist(dist_ctx *c, int b1, int b2) {
  c->bv_tests++;
  return fclo_rss_distance(c->tf, c->tf + 9, c->m1->bvs + b1, c->m2->bvs + b2);
}

/* canStop with rel_err = abs_err = 0 (the reference ignores the request's
 * values: mesh_distance_traversal_node-inl.h:103-104, 147-153). */
static int dist_can_stop(const dist_ctx *c, double d) {
  return (d >= c->min_distance - 0.0) && (d * (1 + 0.0) >= c->min_distance);
}

/* traversal_recurse-inl.h:257-316 (distanceRecurse) */
static void distance_rec(dist_ctx *c, int b1, int b2) {
  const fclo_node *n1 = c->m1->bvs + b1, *n2 = c->m2->bvs + b2;
  int l1 = n1->first_child < 0, l2 = n2->first_child < 0;
  if (l1 && l2) {
    c->leaf_tests++;
    dist_leaf(c, -(n1->first_child + 1), -(n2->first_child + 1));
    return;
  }
  int a1, a2, c1, c2;
  double sz1 = fclo_node_size(c->m1, b1), sz2 = fclo_node_size(c->m2, b2);
  if (l2 || (!l1 && (sz1 > sz2))) {
    a1 = n1->first_child; a2 = b2;
    c1 = n1->first_child + 1; c2 = b2;
  } else {
    a1 = b1; a2 = n2->first_child;
    c1 = b1; c2 = n2->first_child + 1;
  }
  double d1 = bv_dist(c, a1, a2);
  double d2 = bv_dist(c, c1, c2);
  if (d2 < d1) {
    if (!dist_can_stop(c, d2)) distance_rec(c, c1, c2);
    if (!dist_can_stop(c, d1)) distance_rec(c, a1, a2);
  } else {
    if (!dist_can_stop(c, d1)) distance_rec(c, a1, a2);
    if (!dist_can_stop(c, d2)) distance_rec(c, c1, c2);
  }
}

/* distance-inl.h:90-193 -> orientedMeshDistance (distance_func_matrix-inl.h:
 * 386-404) -> detail::distance (collision_node-inl.h:136-147) with the RSS /
 * OBBRSS oriented node: seed with triangle 0 vs triangle 0, traverse, then
 * map the nearest points to the world frame through tf1. */
void fclo_distance(const fclo_model *m1, const double tf1[12],
                   const fclo_model *m2, const double tf2[12],
                   int enable_nearest_points, fclo_distance_out *out) {
  dist_ctx c;
  memset(&c, 0, sizeof(c));
  c.m1 = m1;
  c.m2 = m2;
  c.nearest = enable_nearest_points;
  c.min_distance = DBL_MAX;
  c.b1 = c.b2 = -1;
  fclo_inverse_compose(tf1, tf2, c.tf);
  dist_leaf(&c, 0, 0);
  distance_rec(&c, 0, 0);
  out->min_distance = c.min_distance;
  out->b1 = c.b1;
  out->b2 = c.b2;
  if (enable_nearest_points) {
    xform(tf1, c.p1, out->p1);
    xform(tf1, c.p2, out->p2);
  } else {
    memset(out->p1, 0, sizeof(out->p1));
    memset(out->p2, 0, sizeof(out->p2));
  }
  out->num_bv_tests = c.bv_tests;
  out->num_leaf_tests = c.leaf_tests;
}

void fclo_distance_batch(const fclo_model *const *models, int num_queries,
                         const int *q_m1, const int *q_m2, const double *tf1s,
                         const double *tf2s, int enable_nearest_points,
                         double *dist, double *points, int *ids,
                         int64_t *bv_tests, int64_t *leaf_tests, int threads) {
#pragma omp parallel for schedule(dynamic, 64) num_threads(threads)
  for (int q = 0; q < num_queries; ++q) {
    fclo_distance_out o;
    fclo_distance(models[q_m1[q]], tf1s + 12 * (size_t)q, models[q_m2[q]],
                  tf2s + 12 * (size_t)q, enable_nearest_points, &o);
    dist[q] = o.min_distance;
    if (points) {
      copy3(o.p1, points + 6 * (size_t)q);
      copy3(o.p2, points + 6 * (size_t)q + 3);
    }
    if (ids) {
      ids[2 * q] = o.b1;
      ids[2 * q + 1] = o.b2;
    }
    if (bv_tests) bv_tests[q] = o.num_bv_tests;
    if (leaf_tests) leaf_tests[q] = o.num_leaf_tests;
  }
}

/* Minimum triangle-pair distance over all pairs with no hierarchy, with the
 * pair moved into model 1's frame as the traversal does.  Ties keep the
 * first pair in row-major order. */
double fclo_distance_all_pairs(const fclo_model *m1, const double tf1[12],
                               const fclo_model *m2, const double tf2[12],
                               int ids[2]) {
  dist_ctx c;
  memset(&c, 0, sizeof(c));
  c.m1 = m1;
  c.m2 = m2;
  c.min_distance = DBL_MAX;
  c.b1 = c.b2 = -1;
  fclo_inverse_compose(tf1, tf2, c.tf);
  for (int i = 0; i < m1->num_tris; ++i)
    for (int j = 0; j < m2->num_tris; ++j) dist_leaf(&c, i, j);
  ids[0] = c.b1;
  ids[1] = c.b2;
  return c.min_distance;
}
