// FP32 conservative lower bound of the RSS distance of a BVTT node
// (distance(R0, T0, RSS, RSS), math/bv/RSS-inl.h:1956-1974: rectangle
// distance minus both radii, clamped at 0).
//
// The bound is certified by weak duality, not by trusting the FP32 case
// analysis.  For any direction u with |u| <= 1, the rectangles
//   A = {s e0 + t e1 : 0 <= s <= a0, 0 <= t <= a1}          (A's frame)
//   B = {Tab + Rab (s e0 + t e1) : 0 <= s <= b0, 0 <= t <= b1}
// are at least
//   gap(u) = u.Tab + min(0, b0 u.Rab_c0) + min(0, b1 u.Rab_c1)
//            - max(0, a0 u0) - max(0, a1 u1)
// apart (min over B of u.x minus max over A of u.x), with equality for the
// direction of the closest points.  The FP32 restatement of rectDistance
// only proposes u (its closest-point difference, or the face normal it
// separated along); gap(u) is then evaluated in FP32 and lowered by a
// rounding margin
//   M = KS*u32*S_T + KR*u32*(a0+a1+b0+b1) + 4*u32*(r1+r2),  u32 = 2^-24,
// with S_T = |c1|_1 + |T0|_1 + |c2|_1 (c: the RSS origins).  Derivation
// (first order, fmaf chains with error <= gamma_3 of their terms): Rab and
// Tab are formed as in the OBB cull (cull32.h), each Rab entry within 13u of
// the exact product of the FP64 inputs and each Tab entry within 16.5u S_T;
// so u.Tab errs by <= sqrt(3)*16.5u S_T + 3u*sqrt(3) S_T (its own rounding),
// each b u.Rab_c term by <= (sqrt(3)*13 + 3*sqrt(3) + 2)u b, each a u term
// by <= 2u a, the four additions by <= 4u (S_T + a0+a1+b0+b1), and the
// radius subtraction by <= 2u (|gap| + r1 + r2 + M).  That is < 42u S_T +
// 36u (b0+b1) + 8u (a0+a1) + 3u (r1+r2); KS = 72 and KR = 60 also cover the
// second-order terms and the FP64 rounding of the reference's own bound and
// of the RSS fit (~1e-15 relative).  The result never exceeds the exact
// distance between the stored FP64 volumes, hence never exceeds the
// distance of any triangle pair they contain: a subtree whose bound is
// above the best distance found holds no pair that could be reported.
#pragma once
#include <hip/hip_runtime.h>

#include "cull32.h"
#include "dist64.h"

namespace fclx {

constexpr float kRssKS = 72.0f;
constexpr float kRssKR = 60.0f;

// KS*u32*(rcmax1 + |T0|_1 + rcmax2) for a query, rounded up to float.
__device__ __forceinline__ float rss_margin32(double rcmax1, double rcmax2, const V3& T0) {
  const double st = (rcmax1 + rcmax2) + ((fabs(T0.x) + fabs(T0.y)) + fabs(T0.z));
  return static_cast<float>(st * (static_cast<double>(kRssKS) * 5.9604644775390625e-08) *
                            (1.0 + 1e-6));
}

// Lower bound (>= 0) of the RSS distance of (n1, n2) under the query's
// relative pose R0 (9, row-major), T0 (3) in float, ms = rss_margin32(...).
__device__ __forceinline__ float rss_lb32(const float* R0, const float* T0, float ms,
                                          const RssNode32& n1, const RssNode32& n2) {
  const float* A1 = n1.axis;
  const float* A2 = n2.axis;
  float P[9];
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j)
      P[3 * i + j] = fdot3(R0[3 * i], R0[3 * i + 1], R0[3 * i + 2], A2[j], A2[3 + j], A2[6 + j]);
  float B[9];
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j)
      B[3 * i + j] = fdot3(A1[i], A1[3 + i], A1[6 + i], P[j], P[3 + j], P[6 + j]);
  float Tt[3];
#pragma unroll
  for (int i = 0; i < 3; ++i)
    Tt[i] = (fdot3(R0[3 * i], R0[3 * i + 1], R0[3 * i + 2], n2.to[0], n2.to[1], n2.to[2]) +
             T0[i]) -
            n1.to[i];
  float T[3];
#pragma unroll
  for (int j = 0; j < 3; ++j) T[j] = fdot3(Tt[0], Tt[1], Tt[2], A1[j], A1[3 + j], A1[6 + j]);

  float dir[3];
  (void)rect_distance_t<float>(B, T, n1.l, n2.l, dir);
  const float nn = fdot3(dir[0], dir[1], dir[2], dir[0], dir[1], dir[2]);
  // a proposal with no direction (touching, or NaN) proves nothing
  if (!(nn > 1e-30f) || !(nn < 1e30f)) return 0.0f;
  // |u| <= 1 with room for the rounding of sqrt, the division and the products
  const float sc = 0.99998f / sqrtf(nn);
  const float u0 = dir[0] * sc, u1 = dir[1] * sc, u2 = dir[2] * sc;
  const float g = fdot3(u0, u1, u2, T[0], T[1], T[2]);
  const float p0 = fdot3(u0, u1, u2, B[0], B[3], B[6]);
  const float p1 = fdot3(u0, u1, u2, B[1], B[4], B[7]);
  const float gap = (((g + fminf(0.0f, n2.l[0] * p0)) + fminf(0.0f, n2.l[1] * p1)) -
                     fmaxf(0.0f, n1.l[0] * u0)) -
                    fmaxf(0.0f, n1.l[1] * u1);
  const float M =
      ms + kCullU * (kRssKR * ((n1.l[0] + n1.l[1]) + (n2.l[0] + n2.l[1])) + 4.0f * (n1.r + n2.r));
  const float lb = (gap - (n1.r + n2.r)) - M;
  return lb > 0.0f ? lb : 0.0f;  // NaN gives 0
}

}  // namespace fclx
