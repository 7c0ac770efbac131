// FP64 device distance arithmetic: PQP rectangle-rectangle distance (RSS
// lower bounds) and PQP triangle-triangle distance (leaves), with the same
// fixed evaluation order as exact64.h.
#pragma once
#include "exact64.h"

namespace fclx {

template <typename S>
__device__ __forceinline__ S clip(S v, S a, S b) {
  return (v < a) ? a : ((v > b) ? b : v);
}

// math/bv/RSS-inl.h:457-482
template <typename S>
__device__ __forceinline__ void seg_coords(S& t, S& u, S a, S b, S AdB, S AdT, S BdT) {
  const S denom = S(1) - AdB * AdB;
  if (denom == 0) {
    t = 0;
  } else {
    t = clip((AdT - BdT * AdB) / denom, S(0), a);
  }
  u = t * AdB - BdT;
  if (u < 0) {
    u = 0;
    t = clip(AdT, S(0), a);
  } else if (u > b) {
    u = b;
    t = clip(u * AdB + AdT, S(0), a);
  }
}

// math/bv/RSS-inl.h:485-509
template <typename S>
__device__ __forceinline__ bool in_voronoi(S a, S b, S AnB, S AnT, S AdB, S AdT, S BdT) {
  if (fabs(AnB) < S(1e-7)) return false;
  const S u = clip(-AnT / AnB, S(0), b);
  const S t = clip(u * AdB + AdT, S(0), a);
  const S v = t * AdB - BdT;
  if (AnB > 0) return v > (u + S(1e-7));
  return v < (u - S(1e-7));
}

// Closest points of one A-edge / B-edge pair: the A edge runs along A axis
// IA at A-coordinate a[1-IA] when UA (0 otherwise); the B edge along B axis
// IB at B-coordinate b[1-IB] when UB.  dir: B's point minus A's point.
template <int IA, int UA, int IB, int UB, typename S>
__device__ __forceinline__ S edge_pair(const S* Rab, const S* Tab, const S* a, const S* b, S AdB,
                                       S AdT, S BdT, S* dir) {
  S t, u;
  seg_coords(t, u, a[IA], b[IB], AdB, AdT, BdT);
  S* D = dir;
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    S d = Tab[k];
    if (UB) d = d + Rab[3 * k + (1 - IB)] * b[1 - IB];
    d = d + Rab[3 * k + IB] * u;
    if (k == IA) d = d - t;
    else if (k == 1 - IA && UA) d = d - a[1 - IA];
    D[k] = d;
  }
  return sqrt((D[0] * D[0] + D[1] * D[1]) + D[2] * D[2]);
}

// math/bv/RSS-inl.h:512-1224 (rectDistance), the distance and the
// direction it was measured along (dir: from A towards B; the difference of
// the closest points for an edge pair, a face normal for the fallback).
// The 16 edge-pair cases run in the reference's order with its argument
// expressions; then the face-normal fallback.  S = double is the FP64 bound;
// S = float is the FP32 estimate that only proposes a direction.
template <typename S>
__device__ inline S rect_distance_t(const S* Rab, const S* Tab, const S* a, const S* b, S* dir) {
  const S A0B0 = Rab[0], A0B1 = Rab[1], A1B0 = Rab[3], A1B1 = Rab[4];
  const S aA0B0 = a[0] * A0B0, aA0B1 = a[0] * A0B1;
  const S aA1B0 = a[1] * A1B0, aA1B1 = a[1] * A1B1;
  const S bA0B0 = b[0] * A0B0, bA1B0 = b[0] * A1B0;
  const S bA0B1 = b[1] * A0B1, bA1B1 = b[1] * A1B1;
  S Tba[3];
#pragma unroll
  for (int i = 0; i < 3; ++i) Tba[i] = (Rab[i] * Tab[0] + Rab[3 + i] * Tab[1]) + Rab[6 + i] * Tab[2];

  const S ALL_x = -Tba[0];
  const S ALU_x = ALL_x + aA1B0;
  const S AUL_x = ALL_x + aA0B0;
  const S AUU_x = ALU_x + aA0B0;
  const bool ax1 = ALL_x < ALU_x;
  const S LA1_lx = ax1 ? ALL_x : ALU_x, LA1_ux = ax1 ? ALU_x : ALL_x;
  const S UA1_lx = ax1 ? AUL_x : AUU_x, UA1_ux = ax1 ? AUU_x : AUL_x;
  const S BLL_x = Tab[0];
  const S BLU_x = BLL_x + bA0B1;
  const S BUL_x = BLL_x + bA0B0;
  const S BUU_x = BLU_x + bA0B0;
  const bool bx1 = BLL_x < BLU_x;
  const S LB1_lx = bx1 ? BLL_x : BLU_x, LB1_ux = bx1 ? BLU_x : BLL_x;
  const S UB1_lx = bx1 ? BUL_x : BUU_x, UB1_ux = bx1 ? BUU_x : BUL_x;

  // A1 edges vs B1 edges
  if ((UA1_ux > b[0]) && (UB1_ux > a[0])) {
    if (((UA1_lx > b[0]) || in_voronoi(b[1], a[1], A1B0, aA0B0 - b[0] - Tba[0], A1B1,
                                       aA0B1 - Tba[1], -Tab[1] - bA1B0)) &&
        ((UB1_lx > a[0]) || in_voronoi(a[1], b[1], A0B1, Tab[0] + bA0B0 - a[0], A1B1,
                                       Tab[1] + bA1B0, Tba[1] - aA0B1)))
      return edge_pair<1, 1, 1, 1>(Rab, Tab, a, b, A1B1, Tab[1] + bA1B0, Tba[1] - aA0B1, dir);
  }
  if ((UA1_lx < 0) && (LB1_ux > a[0])) {
    if (((UA1_ux < 0) ||
         in_voronoi(b[1], a[1], -A1B0, Tba[0] - aA0B0, A1B1, aA0B1 - Tba[1], -Tab[1])) &&
        ((LB1_lx > a[0]) ||
         in_voronoi(a[1], b[1], A0B1, Tab[0] - a[0], A1B1, Tab[1], Tba[1] - aA0B1)))
      return edge_pair<1, 1, 1, 0>(Rab, Tab, a, b, A1B1, Tab[1], Tba[1] - aA0B1, dir);
  }
  if ((LA1_ux > b[0]) && (UB1_lx < 0)) {
    if (((LA1_lx > b[0]) ||
         in_voronoi(b[1], a[1], A1B0, -Tba[0] - b[0], A1B1, -Tba[1], -Tab[1] - bA1B0)) &&
        ((UB1_ux < 0) ||
         in_voronoi(a[1], b[1], -A0B1, -Tab[0] - bA0B0, A1B1, Tab[1] + bA1B0, Tba[1])))
      return edge_pair<1, 0, 1, 1>(Rab, Tab, a, b, A1B1, Tab[1] + bA1B0, Tba[1], dir);
  }
  if ((LA1_lx < 0) && (LB1_lx < 0)) {
    if (((LA1_ux < 0) || in_voronoi(b[1], a[1], -A1B0, Tba[0], A1B1, -Tba[1], -Tab[1])) &&
        ((LB1_ux < 0) || in_voronoi(a[1], b[1], -A0B1, -Tab[0], A1B1, Tab[1], Tba[1])))
      return edge_pair<1, 0, 1, 0>(Rab, Tab, a, b, A1B1, Tab[1], Tba[1], dir);
  }

  const S ALL_y = -Tba[1];
  const S ALU_y = ALL_y + aA1B1;
  const S AUL_y = ALL_y + aA0B1;
  const S AUU_y = ALU_y + aA0B1;
  const bool ay1 = ALL_y < ALU_y;
  const S LA1_ly = ay1 ? ALL_y : ALU_y, LA1_uy = ay1 ? ALU_y : ALL_y;
  const S UA1_ly = ay1 ? AUL_y : AUU_y, UA1_uy = ay1 ? AUU_y : AUL_y;
  const bool bx0 = BLL_x < BUL_x;
  const S LB0_lx = bx0 ? BLL_x : BUL_x, LB0_ux = bx0 ? BUL_x : BLL_x;
  const S UB0_lx = bx0 ? BLU_x : BUU_x, UB0_ux = bx0 ? BUU_x : BLU_x;

  // A1 edges vs B0 edges
  if ((UA1_uy > b[1]) && (UB0_ux > a[0])) {
    if (((UA1_ly > b[1]) || in_voronoi(b[0], a[1], A1B1, aA0B1 - Tba[1] - b[1], A1B0,
                                       aA0B0 - Tba[0], -Tab[1] - bA1B1)) &&
        ((UB0_lx > a[0]) || in_voronoi(a[1], b[0], A0B0, Tab[0] - a[0] + bA0B1, A1B0,
                                       Tab[1] + bA1B1, Tba[0] - aA0B0)))
      return edge_pair<1, 1, 0, 1>(Rab, Tab, a, b, A1B0, Tab[1] + bA1B1, Tba[0] - aA0B0, dir);
  }
  if ((UA1_ly < 0) && (LB0_ux > a[0])) {
    if (((UA1_uy < 0) ||
         in_voronoi(b[0], a[1], -A1B1, Tba[1] - aA0B1, A1B0, aA0B0 - Tba[0], -Tab[1])) &&
        ((LB0_lx > a[0]) ||
         in_voronoi(a[1], b[0], A0B0, Tab[0] - a[0], A1B0, Tab[1], Tba[0] - aA0B0)))
      return edge_pair<1, 1, 0, 0>(Rab, Tab, a, b, A1B0, Tab[1], Tba[0] - aA0B0, dir);
  }
  if ((LA1_uy > b[1]) && (UB0_lx < 0)) {
    if (((LA1_ly > b[1]) ||
         in_voronoi(b[0], a[1], A1B1, -Tba[1] - b[1], A1B0, -Tba[0], -Tab[1] - bA1B1)) &&
        ((UB0_ux < 0) ||
         in_voronoi(a[1], b[0], -A0B0, -Tab[0] - bA0B1, A1B0, Tab[1] + bA1B1, Tba[0])))
      return edge_pair<1, 0, 0, 1>(Rab, Tab, a, b, A1B0, Tab[1] + bA1B1, Tba[0], dir);
  }
  if ((LA1_ly < 0) && (LB0_lx < 0)) {
    if (((LA1_uy < 0) || in_voronoi(b[0], a[1], -A1B1, Tba[1], A1B0, -Tba[0], -Tab[1])) &&
        ((LB0_ux < 0) || in_voronoi(a[1], b[0], -A0B0, -Tab[0], A1B0, Tab[1], Tba[0])))
      return edge_pair<1, 0, 0, 0>(Rab, Tab, a, b, A1B0, Tab[1], Tba[0], dir);
  }

  const S BLL_y = Tab[1];
  const S BLU_y = BLL_y + bA1B1;
  const S BUL_y = BLL_y + bA1B0;
  const S BUU_y = BLU_y + bA1B0;
  const bool ax0 = ALL_x < AUL_x;
  const S LA0_lx = ax0 ? ALL_x : AUL_x, LA0_ux = ax0 ? AUL_x : ALL_x;
  const S UA0_lx = ax0 ? ALU_x : AUU_x, UA0_ux = ax0 ? AUU_x : ALU_x;
  const bool by1 = BLL_y < BLU_y;
  const S LB1_ly = by1 ? BLL_y : BLU_y, LB1_uy = by1 ? BLU_y : BLL_y;
  const S UB1_ly = by1 ? BUL_y : BUU_y, UB1_uy = by1 ? BUU_y : BUL_y;

  // A0 edges vs B1 edges
  if ((UA0_ux > b[0]) && (UB1_uy > a[1])) {
    if (((UA0_lx > b[0]) || in_voronoi(b[1], a[0], A0B0, aA1B0 - Tba[0] - b[0], A0B1,
                                       aA1B1 - Tba[1], -Tab[0] - bA0B0)) &&
        ((UB1_ly > a[1]) || in_voronoi(a[0], b[1], A1B1, Tab[1] - a[1] + bA1B0, A0B1,
                                       Tab[0] + bA0B0, Tba[1] - aA1B1)))
      return edge_pair<0, 1, 1, 1>(Rab, Tab, a, b, A0B1, Tab[0] + bA0B0, Tba[1] - aA1B1, dir);
  }
  if ((UA0_lx < 0) && (LB1_uy > a[1])) {
    if (((UA0_ux < 0) ||
         in_voronoi(b[1], a[0], -A0B0, Tba[0] - aA1B0, A0B1, aA1B1 - Tba[1], -Tab[0])) &&
        ((LB1_ly > a[1]) ||
         in_voronoi(a[0], b[1], A1B1, Tab[1] - a[1], A0B1, Tab[0], Tba[1] - aA1B1)))
      return edge_pair<0, 1, 1, 0>(Rab, Tab, a, b, A0B1, Tab[0], Tba[1] - aA1B1, dir);
  }
  if ((LA0_ux > b[0]) && (UB1_ly < 0)) {
    if (((LA0_lx > b[0]) ||
         in_voronoi(b[1], a[0], A0B0, -b[0] - Tba[0], A0B1, -Tba[1], -bA0B0 - Tab[0])) &&
        ((UB1_uy < 0) ||
         in_voronoi(a[0], b[1], -A1B1, -Tab[1] - bA1B0, A0B1, Tab[0] + bA0B0, Tba[1])))
      return edge_pair<0, 0, 1, 1>(Rab, Tab, a, b, A0B1, Tab[0] + bA0B0, Tba[1], dir);
  }
  if ((LA0_lx < 0) && (LB1_ly < 0)) {
    if (((LA0_ux < 0) || in_voronoi(b[1], a[0], -A0B0, Tba[0], A0B1, -Tba[1], -Tab[0])) &&
        ((LB1_uy < 0) || in_voronoi(a[0], b[1], -A1B1, -Tab[1], A0B1, Tab[0], Tba[1])))
      return edge_pair<0, 0, 1, 0>(Rab, Tab, a, b, A0B1, Tab[0], Tba[1], dir);
  }

  const bool ay0 = ALL_y < AUL_y;
  const S LA0_ly = ay0 ? ALL_y : AUL_y, LA0_uy = ay0 ? AUL_y : ALL_y;
  const S UA0_ly = ay0 ? ALU_y : AUU_y, UA0_uy = ay0 ? AUU_y : ALU_y;
  const bool by0 = BLL_y < BUL_y;
  const S LB0_ly = by0 ? BLL_y : BUL_y, LB0_uy = by0 ? BUL_y : BLL_y;
  const S UB0_ly = by0 ? BLU_y : BUU_y, UB0_uy = by0 ? BUU_y : BLU_y;

  // A0 edges vs B0 edges
  if ((UA0_uy > b[1]) && (UB0_uy > a[1])) {
    if (((UA0_ly > b[1]) || in_voronoi(b[0], a[0], A0B1, aA1B1 - Tba[1] - b[1], A0B0,
                                       aA1B0 - Tba[0], -Tab[0] - bA0B1)) &&
        ((UB0_ly > a[1]) || in_voronoi(a[0], b[0], A1B0, Tab[1] - a[1] + bA1B1, A0B0,
                                       Tab[0] + bA0B1, Tba[0] - aA1B0)))
      return edge_pair<0, 1, 0, 1>(Rab, Tab, a, b, A0B0, Tab[0] + bA0B1, Tba[0] - aA1B0, dir);
  }
  if ((UA0_ly < 0) && (LB0_uy > a[1])) {
    if (((UA0_uy < 0) ||
         in_voronoi(b[0], a[0], -A0B1, Tba[1] - aA1B1, A0B0, aA1B0 - Tba[0], -Tab[0])) &&
        ((LB0_ly > a[1]) ||
         in_voronoi(a[0], b[0], A1B0, Tab[1] - a[1], A0B0, Tab[0], Tba[0] - aA1B0)))
      return edge_pair<0, 1, 0, 0>(Rab, Tab, a, b, A0B0, Tab[0], Tba[0] - aA1B0, dir);
  }
  if ((LA0_uy > b[1]) && (UB0_ly < 0)) {
    if (((LA0_ly > b[1]) ||
         in_voronoi(b[0], a[0], A0B1, -Tba[1] - b[1], A0B0, -Tba[0], -Tab[0] - bA0B1)) &&
        ((UB0_uy < 0) ||
         in_voronoi(a[0], b[0], -A1B0, -Tab[1] - bA1B1, A0B0, Tab[0] + bA0B1, Tba[0])))
      return edge_pair<0, 0, 0, 1>(Rab, Tab, a, b, A0B0, Tab[0] + bA0B1, Tba[0], dir);
  }
  if ((LA0_ly < 0) && (LB0_ly < 0)) {
    if (((LA0_uy < 0) || in_voronoi(b[0], a[0], -A0B1, Tba[1], A0B0, -Tba[0], -Tab[0])) &&
        ((LB0_uy < 0) || in_voronoi(a[0], b[0], -A1B0, -Tab[1], A0B0, Tab[0], Tba[0])))
      return edge_pair<0, 0, 0, 0>(Rab, Tab, a, b, A0B0, Tab[0], Tba[0], dir);
  }

  // face-normal separation: along A's normal (sep1) or B's (sep2)
  S sep1, sep2;
  if (Tab[2] > S(0)) {
    sep1 = Tab[2];
    if (Rab[6] < S(0)) sep1 += b[0] * Rab[6];
    if (Rab[7] < S(0)) sep1 += b[1] * Rab[7];
  } else {
    sep1 = -Tab[2];
    if (Rab[6] > S(0)) sep1 -= b[0] * Rab[6];
    if (Rab[7] > S(0)) sep1 -= b[1] * Rab[7];
  }
  if (Tba[2] < 0) {
    sep2 = -Tba[2];
    if (Rab[2] < S(0)) sep2 += a[0] * Rab[2];
    if (Rab[5] < S(0)) sep2 += a[1] * Rab[5];
  } else {
    sep2 = Tba[2];
    if (Rab[2] > S(0)) sep2 -= a[0] * Rab[2];
    if (Rab[5] > S(0)) sep2 -= a[1] * Rab[5];
  }
  if (sep1 > sep2) {
    // B lies above A's plane when Tab[2] > 0
    const S sg = Tab[2] > S(0) ? S(1) : S(-1);
    dir[0] = S(0);
    dir[1] = S(0);
    dir[2] = sg;
  } else {
    // A lies on the positive side of B's plane when Tba[2] < 0
    const S sg = Tba[2] < 0 ? S(-1) : S(1);
    dir[0] = sg * Rab[2];
    dir[1] = sg * Rab[5];
    dir[2] = sg * Rab[8];
  }
  const S sep = (sep1 > sep2 ? sep1 : sep2);
  return (sep > 0 ? sep : 0);
}

__device__ inline double rect_distance64(const double* Rab, const double* Tab, const double* a,
                                         const double* b) {
  double dir[3];
  return rect_distance_t<double>(Rab, Tab, a, b, dir);
}

// RSS lower bound of the BVTT node (b1,b2): rectangle distance minus both
// radii, clamped at 0 (math/bv/RSS-inl.h:1956-1974).
__device__ __forceinline__ double rss_distance64(const M3& R0, const V3& T0, const RssNode64& n1,
                                                 const RssNode64& n2) {
  double R[9], T[3];
  box_relative(R0, T0, n1.axis, n1.to, n2.axis, n2.to, R, T);
  double d = rect_distance64(R, T, n1.l, n2.l);
  d -= (n1.r + n2.r);
  return (d < 0.0) ? 0.0 : d;
}

// triangle_distance-inl.h:54-167 (segPoints)
__device__ inline void seg_points(const V3& P, const V3& A, const V3& Q, const V3& B, V3& VEC,
                                  V3& X, V3& Y) {
  V3 T = Q - P;
  const double AdA = dot(A, A), BdB = dot(B, B), AdB = dot(A, B);
  const double AdT = dot(A, T), BdT = dot(B, T);
  const double denom = AdA * BdB - AdB * AdB;
  double t = (AdT * BdB - BdT * AdB) / denom;
  if ((t < 0) || isnan(t)) t = 0;
  else if (t > 1) t = 1;
  const double u = (t * AdB - BdT) / BdB;
  if ((u <= 0) || isnan(u)) {
    Y = Q;
    t = AdT / AdA;
    if ((t <= 0) || isnan(t)) {
      X = P;
      VEC = Q - P;
    } else if (t >= 1) {
      X = P + A;
      VEC = Q - X;
    } else {
      X = madd(P, A, t);
      VEC = cross(A, cross(T, A));
    }
  } else if (u >= 1) {
    Y = Q + B;
    t = (AdB + AdT) / AdA;
    if ((t <= 0) || isnan(t)) {
      X = P;
      VEC = Y - P;
    } else if (t >= 1) {
      X = P + A;
      VEC = Y - X;
    } else {
      X = madd(P, A, t);
      T = Y - P;
      VEC = cross(A, cross(T, A));
    }
  } else {
    Y = madd(Q, B, u);
    if ((t <= 0) || isnan(t)) {
      X = P;
      VEC = cross(B, cross(T, B));
    } else if (t >= 1) {
      X = P + A;
      T = Q - X;
      VEC = cross(B, cross(T, B));
    } else {
      X = madd(P, A, t);
      VEC = cross(A, B);
      if (dot(VEC, T) < 0) VEC = V3{VEC.x * (-1), VEC.y * (-1), VEC.z * (-1)};
    }
  }
}

// Vertex-to-face stage of triDistance (triangle_distance-inl.h:258-379) for
// face triangle F (edges E) against the vertices of G.  FACE_FIRST selects
// which output receives the face point.
template <bool FACE_FIRST>
__device__ __forceinline__ bool vertex_face(const V3* F, const V3* G, const V3* E,
                                            bool& shown_disjoint, V3& P, V3& Q,
                                            double& dist) {
  const V3 n = cross(E[0], E[1]);
  const double nl = dot(n, n);
  if (!(nl > 1e-15)) return false;
  double pr[3];
#pragma unroll
  for (int k = 0; k < 3; ++k) pr[k] = dot(F[0] - G[k], n);
  int point = -1;
  if ((pr[0] > 0) && (pr[1] > 0) && (pr[2] > 0)) {
    point = (pr[0] < pr[1]) ? 0 : 1;
    if (pr[2] < pr[point]) point = 2;
  } else if ((pr[0] < 0) && (pr[1] < 0) && (pr[2] < 0)) {
    point = (pr[0] > pr[1]) ? 0 : 1;
    if (pr[2] > pr[point]) point = 2;
  }
  if (point < 0) return false;
  shown_disjoint = true;
  const V3 g = G[point];
#pragma unroll
  for (int e = 0; e < 3; ++e)
    if (!(dot(g - F[e], cross(n, E[e])) > 0)) return false;
  const V3 onf = madd(g, n, pr[point] / nl);
  if (FACE_FIRST) {
    P = onf;
    Q = g;
  } else {
    P = g;
    Q = onf;
  }
  dist = sqrt(sq(P - Q));
  return true;
}

// triangle_distance-inl.h:170-394 (triDistance): distance between triangles
// S and T (already in one frame), closest points P on S and Q on T.
__device__ inline double tri_distance64(const V3* S, const V3* T, V3& P, V3& Q) {
  const V3 Sv[3] = {S[1] - S[0], S[2] - S[1], S[0] - S[2]};
  const V3 Tv[3] = {T[1] - T[0], T[2] - T[1], T[0] - T[2]};
  V3 VEC, minP = v3(0, 0, 0), minQ = v3(0, 0, 0);
  bool shown_disjoint = false;
  double mindd = sq(S[0] - T[0]) + 1;
  for (int i = 0; i < 3; ++i) {
    for (int j = 0; j < 3; ++j) {
      seg_points(S[i], Sv[i], T[j], Tv[j], VEC, P, Q);
      const V3 V = Q - P;
      const double dd = dot(V, V);
      if (dd <= mindd) {
        minP = P;
        minQ = Q;
        mindd = dd;
        double a = dot(S[(i + 2) % 3] - P, VEC);
        double b = dot(T[(j + 2) % 3] - Q, VEC);
        if ((a <= 0) && (b >= 0)) return sqrt(dd);
        const double p = dot(V, VEC);
        if (a < 0) a = 0;
        if (b > 0) b = 0;
        if ((p - a + b) > 0) shown_disjoint = true;
      }
    }
  }
  double dist;
  if (vertex_face<true>(S, T, Sv, shown_disjoint, P, Q, dist)) return dist;
  if (vertex_face<false>(T, S, Tv, shown_disjoint, P, Q, dist)) return dist;
  if (shown_disjoint) {
    P = minP;
    Q = minQ;
    return sqrt(mindd);
  }
  return 0;
}

}  // namespace fclx
