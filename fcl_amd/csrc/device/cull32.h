// FP32 conservative OBB separating-axis test with a three-way outcome.
//
// The FP64 test (exact64.h, OBB-inl.h:383-523) reports "disjoint" when some
// axis k has |s_k| > r_k.  Here the same 15 quantities are formed in FP32
// from float copies of the node data and of the query transform.  For every
// axis the FP32 difference d'_k = |s'_k| - r'_k differs from the FP64 one by
// at most
//   M = KS*u*S_T + KR*u*(a0+a1+a2+b0+b1+b2),   u = 2^-24,
// where S_T = |c1|_1 + |T0|_1 + |c2|_1 bounds every centre/translation term
// and a, b are the half extents.  Derivation (first order, every FP32 dot
// product a chain of fmaf with error <= gamma_3): the float conversions
// perturb each input by <= u relative; R0*A2 then carries <= 5.01u per entry
// (entries of |R0||A2| <= 1 for rotations), B = A1^T(R0 A2) <= 13u, the
// centre offset R0 c2 + T0 - c1 <= 7.02u S_T, its projection T <= 16.5u S_T;
// the box-2 face projections then err by <= 54.1u S_T and the edge-axis
// terms T[i2]B(i1,j) - T[i1]B(i2,j) by <= 63u S_T; |B|+1e-6 carries <= 14u,
// so every radius sum errs by <= 20.1u (sum of extents).  The FP64 values
// themselves are within ~1e-14 relative of exact.  KS = 68 and KR = 22 add
// 5-8% over 63 and 20.1, which also covers the rounding of |s'|-r' and of M.
//   d'_k >  M  for some k  =>  the FP64 test finds that axis separating;
//   d'_k <= -M for every k =>  the FP64 test finds no separating axis;
//   otherwise the caller runs the FP64 test.
// Either sure outcome is the FP64 outcome, so culling visits exactly the
// reference's set of BVTT nodes.
#pragma once
#include <hip/hip_runtime.h>

#include "../fclx_layout.h"

namespace fclx {

constexpr float kCullU = 5.9604644775390625e-08f;  // 2^-24
constexpr float kCullKS = 68.0f;
constexpr float kCullKR = 22.0f;

enum CullResult : int { kCullDisjoint = 0, kCullOverlap = 1, kCullUnsure = 2 };

// Query transform in float plus its translation-margin term KS*u*S_T
// (S_T includes the largest |c|_1 of either mesh's nodes).
struct alignas(64) QueryXf32 {
  float R[9];
  float T[3];
  float ms;
  int32_t m1, m2;
  int32_t valid;
};

__device__ __forceinline__ float fdot3(float a0, float a1, float a2, float b0, float b1,
                                       float b2) {
  return fmaf(a2, b2, fmaf(a1, b1, a0 * b0));
}

__device__ __forceinline__ int obb_cull32(const QueryXf32& X, const ObbNode32& n1,
                                          const ObbNode32& n2) {
  const float* R0 = X.R;
  const float* A1 = n1.axis;
  const float* A2 = n2.axis;
  // P = R0 A2
  float P[9];
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j)
      P[3 * i + j] = fdot3(R0[3 * i], R0[3 * i + 1], R0[3 * i + 2], A2[j], A2[3 + j], A2[6 + j]);
  // B = A1^T P
  float B[9];
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j)
      B[3 * i + j] = fdot3(A1[i], A1[3 + i], A1[6 + i], P[j], P[3 + j], P[6 + j]);
  // T = (R0 c2 + T0 - c1)^T A1
  float Tt[3];
#pragma unroll
  for (int i = 0; i < 3; ++i)
    Tt[i] = (fdot3(R0[3 * i], R0[3 * i + 1], R0[3 * i + 2], n2.to[0], n2.to[1], n2.to[2]) +
             X.T[i]) -
            n1.to[i];
  float T[3];
#pragma unroll
  for (int j = 0; j < 3; ++j) T[j] = fdot3(Tt[0], Tt[1], Tt[2], A1[j], A1[3 + j], A1[6 + j]);

  const float* a = n1.extent;
  const float* b = n2.extent;
  const float M =
      X.ms + (kCullKR * kCullU) * (((a[0] + a[1]) + a[2]) + ((b[0] + b[1]) + b[2]));
  const float reps = 1e-6f;
  float F[9];
#pragma unroll
  for (int i = 0; i < 9; ++i) F[i] = fabsf(B[i]) + reps;

  // d_k = |s_k| - r_k for the 15 axes, reduced to their maximum; a NaN
  // d_k (NaN input) is recorded separately and forces the FP64 test
  float dmax = -INFINITY;
  bool nan = false;
  auto acc = [&](float d) {
    nan |= (d != d);
    dmax = fmaxf(dmax, d);
  };
#pragma unroll
  for (int i = 0; i < 3; ++i)
    acc(fabsf(T[i]) - fmaf(F[3 * i + 2], b[2], fmaf(F[3 * i + 1], b[1], fmaf(F[3 * i], b[0], a[i]))));
#pragma unroll
  for (int j = 0; j < 3; ++j) {
    const float s = fdot3(B[j], B[3 + j], B[6 + j], T[0], T[1], T[2]);
    acc(fabsf(s) - fmaf(F[6 + j], a[2], fmaf(F[3 + j], a[1], fmaf(F[j], a[0], b[j]))));
  }
#pragma unroll
  for (int i = 0; i < 3; ++i) {
    const int i1 = (i + 1) % 3, i2 = (i + 2) % 3;
#pragma unroll
    for (int j = 0; j < 3; ++j) {
      const int j1 = (j == 0) ? 1 : 0, j2 = (j == 2) ? 1 : 2;
      const float s = fmaf(T[i2], B[3 * i1 + j], -(T[i1] * B[3 * i2 + j]));
      const float r = fmaf(b[j2], F[3 * i + j1],
                           fmaf(b[j1], F[3 * i + j2],
                                fmaf(a[i2], F[3 * i1 + j], a[i1] * F[3 * i2 + j])));
      acc(fabsf(s) - r);
    }
  }
  if (nan || !(M == M)) return kCullUnsure;
  if (dmax > M) return kCullDisjoint;
  if (dmax <= -M) return kCullOverlap;
  return kCullUnsure;
}

}  // namespace fclx
