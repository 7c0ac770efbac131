// FP64 primitive-shape collision with FCL's GST_INDEP solver, one shape pair
// per lane: supports of box / sphere / capsule / cylinder / convex, GJK
// with the reference's simplex bookkeeping, EPA with its face lists kept in
// the reference's order, and the five closed forms the solver dispatches
// to.  Evaluation order as in exact64.h (no fused multiply-add).
//
//   supports          convexity_based_algorithm/minkowski_diff-inl.h:64-273
//   GJK               convexity_based_algorithm/gjk-inl.h:82-315
//   EPA               convexity_based_algorithm/epa-inl.h:54-396
//   projections       math/detail/project-inl.h:123-316
//   closed forms      primitive_shape_algorithm/{sphere_sphere,sphere_box,
//                     sphere_capsule,sphere_cylinder,box_box}-inl.h
//   dispatch          gjk_solver_indep-inl.h:110-305
//
// EPA's recursive horizon walk (epa-inl.h:347-396) runs on an explicit
// stack, and the recursive simplex completion (gjk-inl.h:230-290) as one
// function per simplex rank; both visit in the reference's order.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace fclx {
namespace shp {

struct D3 {
  double v[3];
};

__device__ __forceinline__ D3 mk(double a, double b, double c) { return D3{{a, b, c}}; }
__device__ __forceinline__ D3 add(const D3& a, const D3& b) {
  return mk(a.v[0] + b.v[0], a.v[1] + b.v[1], a.v[2] + b.v[2]);
}
__device__ __forceinline__ D3 sub(const D3& a, const D3& b) {
  return mk(a.v[0] - b.v[0], a.v[1] - b.v[1], a.v[2] - b.v[2]);
}
__device__ __forceinline__ D3 neg(const D3& a) { return mk(-a.v[0], -a.v[1], -a.v[2]); }
__device__ __forceinline__ D3 mul(const D3& a, double s) {
  return mk(a.v[0] * s, a.v[1] * s, a.v[2] * s);
}
__device__ __forceinline__ D3 dvd(const D3& a, double s) {
  return mk(a.v[0] / s, a.v[1] / s, a.v[2] / s);
}
__device__ __forceinline__ double dot(const D3& a, const D3& b) {
  return (a.v[0] * b.v[0] + a.v[1] * b.v[1]) + a.v[2] * b.v[2];
}
__device__ __forceinline__ D3 cross(const D3& a, const D3& b) {
  return mk(a.v[1] * b.v[2] - a.v[2] * b.v[1], a.v[2] * b.v[0] - a.v[0] * b.v[2],
            a.v[0] * b.v[1] - a.v[1] * b.v[0]);
}
__device__ __forceinline__ double sqn(const D3& a) { return dot(a, a); }
__device__ __forceinline__ double nrm(const D3& a) { return sqrt(sqn(a)); }
__device__ __forceinline__ D3 normalized(const D3& a) {  // Eigen: unchanged at zero
  const double z = sqn(a);
  return z > 0 ? dvd(a, sqrt(z)) : a;
}
__device__ __forceinline__ double triple(const D3& a, const D3& b, const D3& c) {
  return dot(a, cross(b, c));
}

struct Xf {
  double R[9];  // row-major
  D3 t;
};

__device__ __forceinline__ Xf xf_load(const double* tf) {
  Xf r;
#pragma unroll
  for (int i = 0; i < 9; ++i) r.R[i] = tf[i];
  r.t = mk(tf[9], tf[10], tf[11]);
  return r;
}
__device__ __forceinline__ D3 rot(const double* R, const D3& p) {
  return mk(dot(mk(R[0], R[1], R[2]), p), dot(mk(R[3], R[4], R[5]), p),
            dot(mk(R[6], R[7], R[8]), p));
}
__device__ __forceinline__ D3 rotT(const double* R, const D3& p) {
  return mk(dot(mk(R[0], R[3], R[6]), p), dot(mk(R[1], R[4], R[7]), p),
            dot(mk(R[2], R[5], R[8]), p));
}
__device__ __forceinline__ D3 apply(const Xf& T, const D3& p) { return add(rot(T.R, p), T.t); }
__device__ __forceinline__ D3 col(const double* R, int j) { return mk(R[j], R[3 + j], R[6 + j]); }
__device__ __forceinline__ void mulTN(const double* A, const double* B, double* out) {
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j)
      out[3 * i + j] = (A[i] * B[j] + A[3 + i] * B[3 + j]) + A[6 + i] * B[6 + j];
}
// a^-1 b for isometries: (Ra^T Rb, Ra^T tb + (-(Ra^T ta)))
__device__ __forceinline__ Xf inv_compose(const Xf& a, const Xf& b) {
  Xf r;
  mulTN(a.R, b.R, r.R);
  r.t = add(rotT(a.R, b.t), neg(rotT(a.R, a.t)));
  return r;
}

enum : int { kBox = 0, kSphere = 1, kCapsule = 2, kCylinder = 3, kConvex = 4, kTriangle = 5 };

// One primitive: type, dims (box sides / radius / radius+lz) and, for a
// convex, its vertices (local frame).
struct Shape {
  int type;
  int nv;
  double dims[3];
  const double* verts;
};

struct Contact {
  D3 normal, pos;
  double depth;
};

// ------------------------------------------------------------ supports
__device__ inline D3 support(const Shape& s, const D3& d) {
  switch (s.type) {
    case kBox:
      return mk(d.v[0] > 0 ? s.dims[0] / 2 : -s.dims[0] / 2,
                d.v[1] > 0 ? s.dims[1] / 2 : -s.dims[1] / 2,
                d.v[2] > 0 ? s.dims[2] / 2 : -s.dims[2] / 2);
    case kSphere:
      return mul(d, s.dims[0]);
    case kCapsule: {
      const double half_h = s.dims[1] * 0.5;
      const D3 v = mul(d, s.dims[0]);
      const D3 p1 = add(mk(0, 0, half_h), v), p2 = add(mk(0, 0, -half_h), v);
      return dot(d, p1) > dot(d, p2) ? p1 : p2;
    }
    case kCylinder: {
      const double zdist = sqrt(d.v[0] * d.v[0] + d.v[1] * d.v[1]);
      const double half_h = s.dims[1] * 0.5;
      if (zdist == 0.0) return mk(0, 0, d.v[2] > 0 ? half_h : -half_h);
      const double k = s.dims[0] / zdist;
      return mk(k * d.v[0], k * d.v[1], d.v[2] > 0 ? half_h : -half_h);
    }
    case kTriangle: {
      // TriangleP (minkowski_diff-inl.h:79-100): verts = a, b, c; the
      // first of the larger dot products, c only when strictly above
      const D3 a = mk(s.verts[0], s.verts[1], s.verts[2]);
      const D3 b = mk(s.verts[3], s.verts[4], s.verts[5]);
      const D3 c = mk(s.verts[6], s.verts[7], s.verts[8]);
      const double da = dot(d, a), db = dot(d, b), dc = dot(d, c);
      if (da > db) return dc > da ? c : a;
      return dc > db ? c : b;
    }
    default: {  // convex: the first vertex of largest d . v
      double maxdot = -1.7976931348623157e308;
      D3 best = mk(0, 0, 0);
      // four vertices' loads at a time (indices clamped to the last vertex,
      // so every load stays in the convex's range), compared in order
      for (int i0 = 0; i0 < s.nv; i0 += 4) {
        D3 p[4];
#pragma unroll
        for (int k = 0; k < 4; ++k) {
          const int i = i0 + k < s.nv ? i0 + k : s.nv - 1;
          p[k] = mk(s.verts[3 * i], s.verts[3 * i + 1], s.verts[3 * i + 2]);
        }
#pragma unroll
        for (int k = 0; k < 4; ++k) {
          if (i0 + k >= s.nv) break;
          const double dt = dot(d, p[k]);
          if (dt > maxdot) {
            best = p[k];
            maxdot = dt;
          }
        }
      }
      return best;
    }
  }
}

// The Minkowski difference of the pair in shape 0's frame
// (minkowski_diff-inl.h:209-273): toshape1 = R2^T R1 turns a direction into
// shape 1's frame, toshape0 = tf1^-1 tf2 maps shape 1's points back.
// toshape0's rotation R1^T R2 is toshape1 transposed entry for entry (the
// same products summed in the same order), so only toshape1 is kept.
struct MDiff {
  Shape s0, s1;
  double r[9];  // R2^T R1
  D3 t;         // translation of tf1^-1 tf2
};

__device__ __forceinline__ D3 support0(const MDiff& m, const D3& d) { return support(m.s0, d); }
__device__ __forceinline__ D3 support1(const MDiff& m, const D3& d) {
  return add(rotT(m.r, support(m.s1, rot(m.r, d))), m.t);
}
__device__ __forceinline__ D3 msupport(const MDiff& m, const D3& d) {
  return sub(support0(m, d), support1(m, neg(d)));
}

// -------------------------------------------------- simplex projection
struct Proj {
  double p[4];
  double sqr;
  int encode;
};

__device__ __forceinline__ Proj proj_init() {
  Proj r;
  r.p[0] = r.p[1] = r.p[2] = r.p[3] = 0;
  r.sqr = -1;
  r.encode = 0;
  return r;
}

__device__ inline Proj project_line_origin(const D3& a, const D3& b) {
  Proj r = proj_init();
  const D3 d = sub(b, a);
  const double l = sqn(d);
  if (l > 0) {
    const double t = -dot(a, d);
    r.p[1] = (t >= l) ? 1 : ((t <= 0) ? 0 : (t / l));
    r.p[0] = 1 - r.p[1];
    if (t >= l) {
      r.sqr = sqn(b);
      r.encode = 2;
    } else if (t <= 0) {
      r.sqr = sqn(a);
      r.encode = 1;
    } else {
      r.sqr = sqn(add(a, mul(d, r.p[1])));
      r.encode = 3;
    }
  }
  return r;
}

__device__ inline Proj project_triangle_origin(const D3& a, const D3& b, const D3& c) {
  const int nexti[3] = {1, 2, 0};
  Proj r = proj_init();
  const D3 vt[3] = {a, b, c};
  const D3 dl[3] = {sub(a, b), sub(b, c), sub(c, a)};
  const D3 n = cross(dl[0], dl[1]);
  const double l = sqn(n);
  if (l > 0) {
    double mindist = -1;
#pragma unroll
    for (int i = 0; i < 3; ++i) {
      if (dot(vt[i], cross(dl[i], n)) > 0) {
        const int j = nexti[i];
        const Proj rl = project_line_origin(vt[i], vt[j]);
        if (mindist < 0 || rl.sqr < mindist) {
          mindist = rl.sqr;
          r.encode = ((rl.encode & 1) ? 1 << i : 0) + ((rl.encode & 2) ? 1 << j : 0);
          r.p[i] = rl.p[0];
          r.p[j] = rl.p[1];
          r.p[nexti[j]] = 0;
        }
      }
    }
    if (mindist < 0) {
      const double d = dot(a, n);
      const double s = sqrt(l);
      const D3 o = mul(n, d / l);
      mindist = sqn(o);
      r.encode = 7;
      r.p[0] = nrm(cross(dl[1], sub(b, o))) / s;
      r.p[1] = nrm(cross(dl[2], sub(c, o))) / s;
      r.p[2] = 1 - r.p[0] - r.p[1];
    }
    r.sqr = mindist;
  }
  return r;
}

__device__ inline Proj project_tetra_origin(const D3& a, const D3& b, const D3& c, const D3& d) {
  const int nexti[3] = {1, 2, 0};
  Proj r = proj_init();
  const D3 vt[4] = {a, b, c, d};
  const D3 dl[3] = {sub(a, d), sub(b, d), sub(c, d)};
  const double vl = triple(dl[0], dl[1], dl[2]);
  const bool ng = (vl * dot(a, cross(sub(b, c), sub(a, b)))) <= 0;
  if (ng && fabs(vl) > 0) {
    double mindist = -1;
#pragma unroll
    for (int i = 0; i < 3; ++i) {
      const int j = nexti[i];
      const double s = vl * dot(d, cross(dl[i], dl[j]));
      if (s > 0) {
        const Proj rt = project_triangle_origin(vt[i], vt[j], d);
        if (mindist < 0 || rt.sqr < mindist) {
          mindist = rt.sqr;
          r.encode = ((rt.encode & 1) ? 1 << i : 0) + ((rt.encode & 2) ? 1 << j : 0) +
                     ((rt.encode & 4) ? 8 : 0);
          r.p[i] = rt.p[0];
          r.p[j] = rt.p[1];
          r.p[nexti[j]] = 0;
          r.p[3] = rt.p[2];
        }
      }
    }
    if (mindist < 0) {
      mindist = 0;
      r.encode = 15;
      r.p[0] = triple(c, b, d) / vl;
      r.p[1] = triple(a, c, d) / vl;
      r.p[2] = triple(b, a, d) / vl;
      r.p[3] = 1 - (r.p[0] + r.p[1] + r.p[2]);
    }
    r.sqr = mindist;
  } else if (!ng) {
    r = project_triangle_origin(a, b, c);
    r.p[3] = 0;
  }
  return r;
}

// ---------------------------------------------------------------- GJK
enum : int { kGjkValid = 0, kGjkInside = 1, kGjkFailed = 2 };
constexpr int kNvGjk = 4;
constexpr int kNvEpa = 64;
constexpr int kNfEpa = 128;

// GJK's state (gjk.h:60-120) by value: the current simplex's vertices in
// order, with their search directions, and the ray.  The reference's
// vertex store, free list and second simplex only decide which storage
// slot holds a vertex, which changes no result; the barycentric weights it
// stores are never read back.
struct Gjk {
  D3 w[4], d[4];
  int rank;
  D3 ray;
};

// a[k] = v for a position k known only at run time, as a select per
// component so that the array stays in registers
__device__ __forceinline__ void put(D3 (&a)[4], int k, const D3& v) {
#pragma unroll
  for (int j = 0; j < 4; ++j) {
    const bool hit = j == k;
#pragma unroll
    for (int c = 0; c < 3; ++c) a[j].v[c] = hit ? v.v[c] : a[j].v[c];
  }
}

// gjk-inl.h:323-335 appendVertex: direction and support point at the end
__device__ __forceinline__ D3 gjk_append(Gjk& g, const MDiff& m, const D3& v) {
  const D3 d = normalized(v);
  const D3 w = msupport(m, d);
  put(g.d, g.rank, d);
  put(g.w, g.rank, w);
  ++g.rank;
  return w;
}

// gjk-inl.h:82-200 evaluate from the guess (1, 0, 0) (a fresh solver per
// collide() call, collision-inl.h:164-170, so no cached guess carries over;
// evaluate is called with -guess, gjk_solver_indep-inl.h:133)
__device__ inline int gjk_evaluate(Gjk& g, const MDiff& m, double tol, int* passes = nullptr) {
  int iterations = 0;
  double alpha = 0;
  D3 lastw[4];
  int clastw = 0;
  int status = kGjkValid;
  g.rank = 0;
  g.ray = mk(-1, 0, 0);
  gjk_append(g, m, sqn(g.ray) > 0 ? neg(g.ray) : mk(1, 0, 0));
  g.ray = g.w[0];
  lastw[0] = lastw[1] = lastw[2] = lastw[3] = g.ray;
  do {
    const double rl = nrm(g.ray);
    if (rl < tol) {
      status = kGjkInside;
      break;
    }
    const D3 w = gjk_append(g, m, neg(g.ray));
    bool found = false;
#pragma unroll
    for (int i = 0; i < 4; ++i) found |= sqn(sub(w, lastw[i])) < tol;
    if (found) {
      --g.rank;
      break;
    }
    clastw = (clastw + 1) & 3;
    put(lastw, clastw, w);
    const double omega = dot(g.ray, w) / rl;
    alpha = (alpha < omega) ? omega : alpha;  // std::max
    if ((rl - alpha) - tol * rl <= 0) {
      --g.rank;
      break;
    }
    Proj pr = proj_init();
    if (g.rank == 2)
      pr = project_line_origin(g.w[0], g.w[1]);
    else if (g.rank == 3)
      pr = project_triangle_origin(g.w[0], g.w[1], g.w[2]);
    else if (g.rank == 4)
      pr = project_tetra_origin(g.w[0], g.w[1], g.w[2], g.w[3]);
    if (pr.sqr >= 0) {
      // the next simplex: the vertices the projection kept, in order
      D3 nw[4] = {g.w[0], g.w[1], g.w[2], g.w[3]};
      D3 nd[4] = {g.d[0], g.d[1], g.d[2], g.d[3]};
      int nr = 0;
      D3 ray = mk(0, 0, 0);
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        if (i < g.rank && (pr.encode & (1 << i))) {
          put(nw, nr, g.w[i]);
          put(nd, nr, g.d[i]);
          ++nr;
          ray = add(ray, mul(g.w[i], pr.p[i]));
        }
      }
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        g.w[i] = nw[i];
        g.d[i] = nd[i];
      }
      g.rank = nr;
      g.ray = ray;
      if (pr.encode == 15) status = kGjkInside;
    } else {
      --g.rank;
      break;
    }
    status = (++iterations < 128) ? status : kGjkFailed;
  } while (status == kGjkValid);
  if (passes) *passes = iterations;
  return status;
}

// gjk_solver_indep-inl.h:127-130: the pair's Minkowski difference
__device__ inline void make_mdiff(const Shape& s1, const Xf& tf1, const Shape& s2, const Xf& tf2,
                                  MDiff& m) {
  m.s0 = s1;
  m.s1 = s2;
  mulTN(tf2.R, tf1.R, m.r);
  m.t = add(rotT(tf1.R, tf2.t), neg(rotT(tf1.R, tf1.t)));
}

// encloseOrigin (gjk-inl.h:230-290) as a loop: a depth-first search over
// the vertex appended at positions rank .. 3, trying the reference's
// candidate directions in its order (rank 1: +-x, +-y, +-z; rank 2:
// +-(w1 - w0) x axis for the axes where that is non-zero; rank 3: +-n of
// the triangle if non-zero) and accepting the first tetrahedron of non-zero
// volume.  On failure the simplex is back at its starting rank.
__device__ __forceinline__ void set_vertex(Gjk& g, const MDiff& m, int k, const D3& v) {
  const D3 d = normalized(v);
  put(g.d, k, d);
  put(g.w, k, msupport(m, d));
}
__device__ __forceinline__ bool enclose4(const Gjk& g) {
  const D3 w3 = g.w[3];
  return fabs(triple(sub(g.w[0], w3), sub(g.w[1], w3), sub(g.w[2], w3))) > 0;
}
__device__ __forceinline__ D3 unit_axis(int i) {
  return mk(i == 0 ? 1.0 : 0.0, i == 1 ? 1.0 : 0.0, i == 2 ? 1.0 : 0.0);
}
__device__ inline bool enclose_origin(Gjk& g, const MDiff& m) {
  const int r0 = g.rank;
  if (r0 == 4) return enclose4(g);
  if (r0 < 1) return false;
  int t1 = 0, t2 = 0, t3 = 0;  // candidates tried at positions 1, 2, 3
  int k = r0;                  // position being filled
  while (true) {
    bool have = false;
    D3 dir = mk(0, 0, 0);
    if (k == 1) {
      if (t1 < 6) {
        const D3 ax = unit_axis(t1 >> 1);
        dir = (t1 & 1) ? neg(ax) : ax;
        have = true;
      }
    } else if (k == 2) {
      while (t2 < 6) {
        const D3 p = cross(sub(g.w[1], g.w[0]), unit_axis(t2 >> 1));
        if (sqn(p) > 0) {
          dir = (t2 & 1) ? neg(p) : p;
          have = true;
          break;
        }
        t2 += 2;  // neither sign of a zero candidate is tried
      }
    } else if (t3 < 2) {
      const D3 n = cross(sub(g.w[1], g.w[0]), sub(g.w[2], g.w[0]));
      if (sqn(n) > 0) {
        dir = t3 ? neg(n) : n;
        have = true;
      } else {
        t3 = 2;
      }
    }
    if (!have) {  // position k exhausted: the vertex at k - 1 goes
      if (k == r0) {
        g.rank = r0;
        return false;
      }
      --k;
      if (k == 1)
        ++t1;
      else
        ++t2;
      continue;
    }
    set_vertex(g, m, k, dir);
    if (k == 3) {
      if (enclose4(g)) {
        g.rank = 4;
        return true;
      }
      ++t3;
      continue;
    }
    ++k;
    if (k == 2)
      t2 = 0;
    else
      t3 = 0;
  }
}

// ---------------------------------------------------------------- EPA
enum : int {
  kEpaValid,
  kEpaTouching,
  kEpaDegenerated,
  kEpaNonConvex,
  kEpaInvalidHull,
  kEpaOutOfFaces,
  kEpaOutOfVertices,
  kEpaAccuracyReached,
  kEpaFallBack,
  kEpaFailed
};

// One EPA face (epa.h SimplexF): normal, distance, vertices, neighbour
// faces and the neighbour's edge index, and the pass stamp of the horizon
// walk.  One cache line.
struct alignas(64) EpaFace {
  D3 n;
  double d;
  int32_t pass;
  int8_t c[3];
  int8_t f[3];
  int8_t e[3];
};

// The reference keeps EPA's faces in two linked lists (hull and stock,
// epa.h SimplexList); here:
//  - the hull list is only ever prepended to (epa-inl.h:71-81), so its
//    order from the root is the order of decreasing append time; a stamp
//    per face records that time (-1: not on the hull), and the root is the
//    hull face with the largest stamp;
//  - the stock is a last-in-first-out store of free slots: the freed slots
//    on a stack, then the never-used slots in index order, which is the
//    order the reference's stock list hands them out;
//  - d * d and the stamp of every face sit in arrays of their own, scanned
//    by every lane of a wavefront in step.
// Per-lane memory: one contiguous block (a face or vertex is one line).
struct EpaLane {
  EpaFace fc[kNfEpa];
  D3 w[kNvGjk + kNvEpa];    // support points (GJK's simplex first)
  D3 dir[kNvGjk + kNvEpa];  // their search directions
  int8_t freed[kNfEpa];     // stack of freed face slots
  int16_t xs[kNfEpa];       // horizon walk stack: face | edge << 7 | stage << 9
};
// Per-wavefront memory, lane-interleaved so that the scan's loads coalesce.
struct EpaWave {
  double dsq[kNfEpa][64];
  int32_t stamp[kNfEpa][64];
};

struct Epa {
  EpaLane* M;
  EpaWave* S;
  int lane;
  // the horizon walk's stack: frame i at xs[i * xs_stride] (the lane's own
  // M->xs with stride 1, or a lane's column of an LDS array with stride 64)
  int16_t* xs = nullptr;
  int xs_stride = 1;
  int nfreed;      // entries on the freed stack
  int hi;          // slots handed out so far: [0, hi)
  int hull_count;
  int next_stamp;
  int nextsv;
  int status;
  int best;        // current best face (the reference's outer is its copy)
  int pass;
  int it;          // loop passes done
  bool fallback;   // no tetrahedron: the FallBack result
  int fb_vertex;   // the simplex's first vertex then (1 once the winding swap moved it)
};

__device__ __forceinline__ void hull_add(Epa& E, int f) {
  E.S->stamp[f][E.lane] = E.next_stamp++;
  ++E.hull_count;
}
__device__ __forceinline__ void hull_drop(Epa& E, int f) {
  E.S->stamp[f][E.lane] = -1;
  --E.hull_count;
}
__device__ __forceinline__ void stock_put(Epa& E, int f) { E.M->freed[E.nfreed++] = (int8_t)f; }
__device__ __forceinline__ void bind(EpaLane& M, int fa, int ea, int fb, int eb) {
  M.fc[fa].e[ea] = static_cast<int8_t>(eb);
  M.fc[fa].f[ea] = static_cast<int8_t>(fb);
  M.fc[fb].e[eb] = static_cast<int8_t>(ea);
  M.fc[fb].f[eb] = static_cast<int8_t>(fa);
}

// epa-inl.h:109-141 getEdgeDist (with the reference's distance formula),
// on the edge's two support points
__device__ inline bool edge_dist(const D3& wa, const D3& wb, const D3& n, double& dist) {
  const D3 ba = sub(wb, wa);
  const D3 n_ab = cross(ba, n);
  if (dot(wa, n_ab) < 0) {
    const double a_dot_ba = dot(wa, ba), b_dot_ba = dot(wb, ba);
    if (a_dot_ba > 0)
      dist = nrm(wa);
    else if (b_dot_ba < 0)
      dist = nrm(wb);
    else {
      const double a_dot_b = dot(wa, wb);
      const double v = sqn(wa) * sqn(wb) - a_dot_b * a_dot_b;
      dist = sqrt(v > 0 ? v : 0.0);
    }
    return true;
  }
  return false;
}

// epa-inl.h:145-189 newFace
__device__ inline int new_face(Epa& E, int a, int b, int c, bool forced, double tol) {
  EpaLane& M = *E.M;
  int f;
  if (E.nfreed > 0)
    f = M.freed[--E.nfreed];
  else if (E.hi < kNfEpa)
    f = E.hi++;
  else {
    E.status = kEpaOutOfFaces;  // the reference's test of an empty stock
    return -1;
  }
  hull_add(E, f);
  EpaFace& F = M.fc[f];
  F.pass = 0;
  F.c[0] = static_cast<int8_t>(a);
  F.c[1] = static_cast<int8_t>(b);
  F.c[2] = static_cast<int8_t>(c);
  // the three support points once (the face writes above could alias them
  // for the compiler, which would read them again for every use)
  const D3 wa = M.w[a], wb = M.w[b], wc = M.w[c];
  const D3 n = cross(sub(wb, wa), sub(wc, wa));
  const double l = nrm(n);
  if (l > tol) {
    double d;
    if (!(edge_dist(wa, wb, n, d) || edge_dist(wb, wc, n, d) || edge_dist(wc, wa, n, d)))
      d = dot(wa, n) / l;
    F.d = d;
    E.S->dsq[f][E.lane] = d * d;
    F.n = dvd(n, l);
    if (forced || d >= -tol) return f;
    E.status = kEpaNonConvex;
  } else {
    E.status = kEpaDegenerated;
  }
  hull_drop(E, f);
  stock_put(E, f);
  return -1;
}

// epa-inl.h:193-209 findBest: the first face, in list order from the root,
// whose d * d is below all before it -- the smallest d * d, ties to the
// latest append; a NaN root is never displaced.  The scan reads eight
// faces' stamps and distances at a time (all in bounds: kNfEpa is a
// multiple of 8) before it compares them in slot order, so that their
// loads are in flight together instead of one after another.
__device__ inline int find_best(const Epa& E) {
  static_assert(kNfEpa % 8 == 0, "the scan reads whole groups of eight faces");
  const EpaWave& S = *E.S;
  int root = -1, root_stamp = -1;
  double root_dsq = 0;
  int best = -1, best_stamp = -1;
  double best_dsq = 0;
  for (int f0 = 0; f0 < E.hi; f0 += 8) {
    int st[8];
    double dv[8];
#pragma unroll
    for (int k = 0; k < 8; ++k) {
      st[k] = S.stamp[f0 + k][E.lane];
      dv[k] = S.dsq[f0 + k][E.lane];
    }
#pragma unroll
    for (int k = 0; k < 8; ++k) {
      const int f = f0 + k;
      const int s = f < E.hi ? st[k] : -1;
      if (s < 0) continue;
      const double v = dv[k];
      if (s > root_stamp) {
        root = f;
        root_stamp = s;
        root_dsq = v;
      }
      if (v == v && (best < 0 || v < best_dsq || (v == best_dsq && s > best_stamp))) {
        best = f;
        best_dsq = v;
        best_stamp = s;
      }
    }
  }
  return (root_dsq != root_dsq || best < 0) ? root : best;
}

struct Horizon {
  int cf, ff, nf;
};

// epa-inl.h:347-396 expand(pass, w, f, e, horizon), recursion unrolled onto
// a stack of frames; a frame in stage 1 waits for its first child, stage 2
// for its second.  Same visit order and results as the recursive form.
__device__ __forceinline__ int next3(int e) { return e == 2 ? 0 : e + 1; }
__device__ __forceinline__ int prev3(int e) { return e == 0 ? 2 : e - 1; }

__device__ inline bool expand(Epa& E, int pass, int w, int f0, int e0, Horizon& h,
                              double tol) {
  EpaLane& M = *E.M;
  const D3 ww = M.w[w];
  int sp = 0;
  int f = f0, e = e0;
  bool result = false;
  bool call = true;  // evaluate (f, e) as a fresh call
  while (true) {
    if (call) {
      // the whole face record at once (one round trip, not one per field)
      const EpaFace F = M.fc[f];
      result = false;
      if (F.pass != pass) {
        const int e1 = next3(e);
        if (dot(F.n, ww) - F.d < -tol) {
          const int nf = new_face(E, F.c[e1], F.c[e], w, false, tol);
          if (nf >= 0) {
            bind(M, nf, 0, f, e);
            if (h.cf >= 0)
              bind(M, nf, 2, h.cf, 1);
            else
              h.ff = nf;
            h.cf = nf;
            ++h.nf;
            result = true;
          }
        } else {
          M.fc[f].pass = pass;
          // every frame marks a distinct face with this pass: at most kNfEpa
          if (sp >= kNfEpa) return false;
          E.xs[sp++ * E.xs_stride] = static_cast<int16_t>(f | (e << 7) | (1 << 9));
          const int fe1 = F.f[e1], ee1 = F.e[e1];
          f = fe1;
          e = ee1;
          continue;  // call the first child
        }
      }
      call = false;
    }
    // return `result` to the parent frame
    if (sp == 0) return result;
    const int top = E.xs[(sp - 1) * E.xs_stride];
    const int pf = top & 127, pe = (top >> 7) & 3;
    if (top >> 9 == 1) {
      if (!result) {
        --sp;  // parent returns false
        continue;
      }
      E.xs[(sp - 1) * E.xs_stride] = static_cast<int16_t>(pf | (pe << 7) | (2 << 9));
      const int e2 = prev3(pe);
      f = M.fc[pf].f[e2];
      e = M.fc[pf].e[e2];
      call = true;
      continue;  // call the second child
    }
    // stage 2: both children done
    --sp;
    if (result) {
      hull_drop(E, pf);
      stock_put(E, pf);
    }
  }
}

// epa-inl.h:225-343 evaluate, in three parts so that a lane can run one
// loop pass at a time: epa_start (enclosing step and the tetrahedron; false
// when the FallBack result applies), epa_step (one pass of the loop at
// epa-inl.h:283-321; false once the loop has ended) and epa_finish (the
// result of epa-inl.h:323-343 and its contact, gjk_solver_indep-inl.h:
// 141-157).  The GJK simplex's vertices take slots 0-3 after the enclosing
// step; c names them in the reference's order.
__device__ inline bool epa_start(Epa& E, Gjk& g, const MDiff& m, double tol) {
  EpaLane& M = *E.M;
  E.nfreed = 0;
  E.hi = 0;
  E.hull_count = 0;
  E.next_stamp = 0;
  E.status = kEpaFailed;
  E.nextsv = 0;
  E.fallback = true;
  E.fb_vertex = 0;
  const bool enclosed = g.rank > 1 && enclose_origin(g, m);
#pragma unroll
  for (int i = 0; i < 4; ++i)
    if (i < g.rank) {
      M.w[i] = g.w[i];
      M.dir[i] = g.d[i];
    }
  if (enclosed) {
    E.status = kEpaValid;
    int c[4] = {0, 1, 2, 3};
    if (dot(sub(g.w[0], g.w[3]), cross(sub(g.w[1], g.w[3]), sub(g.w[2], g.w[3]))) < 0) {
      c[0] = 1;
      c[1] = 0;
      // the reference swaps the simplex's first two vertices in place, so a
      // FallBack after a failed tetrahedron reports the swapped first one
      E.fb_vertex = 1;
    }
    int tet[4];
    tet[0] = new_face(E, c[0], c[1], c[2], true, tol);
    tet[1] = new_face(E, c[1], c[0], c[3], true, tol);
    tet[2] = new_face(E, c[2], c[1], c[3], true, tol);
    tet[3] = new_face(E, c[0], c[2], c[3], true, tol);
    if (E.hull_count == 4) {
      E.best = find_best(E);
      E.pass = 0;
      bind(M, tet[0], 0, tet[1], 0);
      bind(M, tet[0], 1, tet[2], 0);
      bind(M, tet[0], 2, tet[3], 0);
      bind(M, tet[1], 1, tet[3], 2);
      bind(M, tet[1], 2, tet[2], 1);
      bind(M, tet[2], 2, tet[3], 1);
      E.status = kEpaValid;
      E.it = 0;
      E.fallback = false;
      return true;
    }
  }
  E.status = kEpaFallBack;
  return false;
}

__device__ inline bool epa_step(Epa& E, const MDiff& m, double tol) {
  if (E.it >= 255) return false;
  ++E.it;
  EpaLane& M = *E.M;
  if (E.nextsv >= kNvEpa) {
    E.status = kEpaOutOfVertices;
    return false;
  }
  Horizon h{-1, -1, 0};
  const int w = kNvGjk + E.nextsv++;
  bool valid = true;
  const int best = E.best;
  const EpaFace B = M.fc[best];  // the whole record at once
  M.fc[best].pass = ++E.pass;
  const D3 bn = B.n;
  const D3 wd = normalized(bn);
  const D3 ww = msupport(m, wd);
  M.dir[w] = wd;
  M.w[w] = ww;
  const double wdist = dot(bn, ww) - B.d;
  if (!(wdist > tol)) {
    E.status = kEpaAccuracyReached;
    return false;
  }
  // (expand changes other faces' links, never the best face's own)
  for (int j = 0; j < 3 && valid; ++j) valid &= expand(E, E.pass, w, B.f[j], B.e[j], h, tol);
  if (!(valid && h.nf >= 3)) {
    E.status = kEpaInvalidHull;
    return false;
  }
  bind(M, h.ff, 2, h.cf, 1);
  hull_drop(E, best);
  stock_put(E, best);
  E.best = find_best(E);
  return true;
}

// The contact of a pair GJK found inside.  EPA's status is never Failed
// once GJK reports Inside (the enclosing path ends in one of its loop
// states, every other path in FallBack), so the pair always collides and
// EPA only supplies the contact.
__device__ inline Contact epa_finish(const Epa& E, const MDiff& m, const Xf& tf1) {
  const EpaLane& M = *E.M;
  D3 normal;
  double depth;
  int rank;
  int rc[3] = {0, 0, 0};
  double rp[3] = {1, 0, 0};
  if (E.fallback) {
    // -guess with guess = -(1, 0, 0) (evaluate is called with -guess),
    // normalised
    normal = mk(1, 0, 0);
    depth = 0;
    rank = 1;
    rc[0] = E.fb_vertex;
  } else {
    const EpaFace& outer = M.fc[E.best];
    const D3 projection = mul(outer.n, outer.d);
    normal = outer.n;
    depth = outer.d;
    rank = 3;
    for (int k = 0; k < 3; ++k) rc[k] = outer.c[k];
    const D3 a = M.w[rc[0]], b = M.w[rc[1]], c2 = M.w[rc[2]];
    rp[0] = nrm(cross(sub(b, projection), sub(c2, projection)));
    rp[1] = nrm(cross(sub(c2, projection), sub(a, projection)));
    rp[2] = nrm(cross(sub(a, projection), sub(b, projection)));
    const double sum = rp[0] + rp[1] + rp[2];
    rp[0] /= sum;
    rp[1] /= sum;
    rp[2] /= sum;
  }
  D3 w0 = mk(0, 0, 0);
#pragma unroll
  for (int i = 0; i < 3; ++i)
    if (i < rank) w0 = add(w0, mul(support0(m, M.dir[rc[i]]), rp[i]));
  Contact c;
  c.normal = normal;
  c.pos = apply(tf1, sub(w0, mul(normal, depth * 0.5)));
  c.depth = -depth;
  return c;
}

// GJK's final simplex of a pair queued for EPA
struct EpaSeed {
  D3 w[4], d[4];
  int32_t q;
  int32_t rank;
  int32_t gjk_passes;
};

// -------------------------------------------------------- closed forms
__device__ inline bool sphere_sphere(const Shape& s1, const Xf& tf1, const Shape& s2,
                                     const Xf& tf2, bool want, Contact* out, int& n) {
  const D3 diff = sub(tf2.t, tf1.t);
  const double len = nrm(diff);
  const double r1 = s1.dims[0], r2 = s2.dims[0];
  if (len > r1 + r2) return false;
  if (want) {
    out[0].normal = len > 0 ? dvd(diff, len) : diff;
    out[0].pos = add(tf1.t, dvd(mul(diff, r1), r1 + r2));
    out[0].depth = r1 + r2 - len;
    n = 1;
  }
  return true;
}

__device__ inline bool sphere_box(const Shape& sp, const Xf& X_FS, const Shape& bx,
                                  const Xf& X_FB, bool want, Contact* out, int& n) {
  const double r = sp.dims[0];
  const Xf X_BS = inv_compose(X_FB, X_FS);
  const D3 p_BC = X_BS.t;
  D3 p_BN;
  bool clamped = false;
  for (int i = 0; i < 3; ++i) {
    const double half = bx.dims[i] / 2;
    p_BN.v[i] = p_BC.v[i];
    if (p_BC.v[i] < -half) {
      clamped = true;
      p_BN.v[i] = -half;
    }
    if (p_BC.v[i] > half) {
      clamped = true;
      p_BN.v[i] = half;
    }
  }
  const D3 p_CN = sub(p_BN, p_BC);
  const double sq = sqn(p_CN);
  if (sq > r * r) return false;
  if (want) {
    const double eps = 16 * 2.220446049250313e-16;
    double depth;
    D3 nv, p_BP;
    if (clamped && sq > eps * eps) {
      const double dist = sqrt(sq);
      nv = dvd(p_CN, dist);
      depth = r - dist;
      p_BP = add(p_BN, mul(nv, depth * 0.5));
    } else {
      double min_distance = __builtin_inf();
      int min_axis = -1;
      for (int i = 0; i < 3; ++i) {
        const double half = bx.dims[i] / 2;
        const double dist = p_BC.v[i] >= 0 ? half - p_BC.v[i] : p_BC.v[i] + half;
        if (dist + eps < min_distance) {
          min_distance = dist;
          min_axis = i;
        }
      }
      nv = mk(0, 0, 0);
      nv.v[min_axis] = p_BC.v[min_axis] >= 0 ? -1 : 1;
      depth = min_distance + r;
      p_BP = add(p_BC, mul(nv, (r - min_distance) / 2));
    }
    out[0].normal = rot(X_FB.R, nv);
    out[0].pos = apply(X_FB, p_BP);
    out[0].depth = depth;
    n = 1;
  }
  return true;
}

__device__ inline bool sphere_capsule(const Shape& sp, const Xf& tf1, const Shape& cap,
                                      const Xf& tf2, bool want, Contact* out, int& n) {
  const D3 pos1 = mk(0., 0., 0.5 * cap.dims[1]);
  const D3 pos2 = mk(0., 0., -0.5 * cap.dims[1]);
  const D3 s_c = add(rotT(tf2.R, tf1.t), neg(rotT(tf2.R, tf2.t)));
  const D3 v = sub(pos2, pos1), w = sub(s_c, pos1);
  const double c1 = dot(w, v), c2 = dot(v, v);
  D3 seg;
  if (c1 <= 0)
    seg = pos1;
  else if (c2 <= c1)
    seg = pos2;
  else
    seg = add(pos1, mul(v, c1 / c2));
  const D3 diff = sub(s_c, seg);
  const double distance = nrm(diff) - sp.dims[0] - cap.dims[0];
  if (distance > 0) return false;
  const D3 ln = neg(normalized(diff));
  if (want) {
    out[0].normal = rot(tf2.R, ln);
    out[0].pos = apply(tf2, add(seg, mul(ln, distance)));
    out[0].depth = -distance;
    n = 1;
  }
  return true;
}

__device__ inline bool sphere_cylinder(const Shape& sp, const Xf& X_FS, const Shape& cy,
                                       const Xf& X_FC, bool want, Contact* out, int& n) {
  const double r_s = sp.dims[0], radius = cy.dims[0], height = cy.dims[1];
  const Xf X_CS = inv_compose(X_FC, X_FS);
  const D3 p_CS = X_CS.t;
  D3 p_CN = p_CS;
  bool outside = false;
  const double half_h = height / 2;
  if (p_CS.v[2] > half_h) {
    outside = true;
    p_CN.v[2] = half_h;
  } else if (p_CS.v[2] < -half_h) {
    outside = true;
    p_CN.v[2] = -half_h;
  }
  double rx = p_CS.v[0], ry = p_CS.v[1];
  const double sqd = rx * rx + ry * ry;
  if (sqd > radius * radius) {
    outside = true;
    const double k = radius / sqrt(sqd);
    rx *= k;
    ry *= k;
    p_CN.v[0] = rx;
    p_CN.v[1] = ry;
  }
  const D3 p_SN = sub(p_CN, p_CS);
  const double sq = sqn(p_SN);
  if (sq > r_s * r_s) return false;
  if (want) {
    const double eps = 16 * 2.220446049250313e-16;
    double depth;
    D3 nv, p_CP;
    if (outside && sq > eps * eps) {
      const double d = sqrt(sq);
      nv = dvd(p_SN, d);
      depth = r_s - d;
      p_CP = add(p_CN, mul(nv, depth * 0.5));
    } else {
      const double face = p_CS.v[2] >= 0 ? height / 2 - p_CS.v[2] : p_CS.v[2] + height / 2;
      const double nx = p_CS.v[0], ny = p_CS.v[1];
      const double d_xy = sqrt(nx * nx + ny * ny);
      const double barrel = radius - d_xy;
      if (barrel < face - eps) {
        if (d_xy > eps) {
          nv = mk(-nx / d_xy, -ny / d_xy, 0);
          depth = r_s + barrel;
          p_CP = add(p_CS, mul(nv, (r_s - barrel) / 2));
        } else {
          nv = mk(-1, 0, 0);
          depth = r_s + radius;
          p_CP = add(p_CS, mul(nv, (r_s - barrel) / 2));
        }
      } else {
        nv = mk(0, 0, p_CS.v[2] >= 0 ? -1 : 1);
        depth = face + r_s;
        p_CP = add(p_CS, mul(nv, (r_s - face) / 2));
      }
    }
    out[0].normal = rot(X_FC.R, nv);
    out[0].pos = apply(X_FC, p_CP);
    out[0].depth = depth;
    n = 1;
  }
  return true;
}

// ------------------------------------------------------------- box-box
// The box-box closed form FCL carries from ODE (box_box-inl.h:110-862,
// boxBox2): a 15-axis separation search, then either one point between the
// two closest edges or the face contact -- the incident face clipped to
// the reference face's rectangle, points kept while behind the reference
// face, at most four of them spread around their centroid.  Organised here
// as small steps; each step's arithmetic and constants are the reference's,
// so the contacts are bit-identical to the oracle's.

// Parameters (ta, tb) of the closest points pa + ta ua and pb + tb ub of two
// lines with unit directions; 0, 0 when they are nearly parallel (1 - c^2
// at most 1e-4 in single precision).
__device__ inline void closest_line_params(const D3& pa, const D3& ua, const D3& pb, const D3& ub,
                                           double& ta, double& tb) {
  const D3 gap = sub(pb, pa);
  const double c = dot(ua, ub);
  const double ga = dot(ua, gap);
  const double gb = -dot(ub, gap);
  const double det = 1 - c * c;
  if (det <= static_cast<double>(0.0001f)) {
    ta = 0;
    tb = 0;
    return;
  }
  const double inv = 1 / det;
  ta = (ga + c * gb) * inv;
  tb = (c * ga + gb) * inv;
}

// One Sutherland-Hodgman pass over a polygon of n (x, y) points: the part
// with sign * v[axis] < lim, each crossing edge contributing its crossing
// point.  The output stops at 8 points (full is then set).
__device__ inline int clip_pass(const double* in, int n, int axis, double sign, double lim,
                                double* out, bool& full) {
  const int other = 1 - axis;
  int m = 0;
  for (int i = 0; i < n; ++i) {
    const double* cur = in + 2 * i;
    const double* nxt = in + 2 * (i + 1 < n ? i + 1 : 0);
    const bool cur_in = sign * cur[axis] < lim;
    if (cur_in) {
      out[2 * m] = cur[0];
      out[2 * m + 1] = cur[1];
      if (++m == 8) {
        full = true;
        return m;
      }
    }
    if (cur_in != (sign * nxt[axis] < lim)) {
      out[2 * m + other] =
          cur[other] + (nxt[other] - cur[other]) / (nxt[axis] - cur[axis]) * (sign * lim - cur[axis]);
      out[2 * m + axis] = sign * lim;
      if (++m == 8) {
        full = true;
        return m;
      }
    }
  }
  return m;
}

// The quad's part inside the rectangle [-h0, h0] x [-h1, h1]: passes
// against -x, +x, -y, +y in turn (at most 8 points).
__device__ inline int clip_quad_to_rect(const double h[2], const double quad[8], double out[16]) {
  double a[16], b[16];
  for (int k = 0; k < 8; ++k) a[k] = quad[k];
  double* src = a;
  double* dst = b;
  int n = 4;
  for (int pass = 0; pass < 4; ++pass) {
    const int axis = pass >> 1;
    bool full = false;
    n = clip_pass(src, n, axis, (pass & 1) ? 1.0 : -1.0, h[axis], dst, full);
    double* t = src;
    src = dst;
    dst = t;
    if (full) break;
  }
  for (int k = 0; k < 2 * n; ++k) out[k] = src[k];
  return n;
}

// Centroid of a contact polygon of n (x, y) points: the point itself, the
// midpoint of two, or the area-weighted centroid of the closed polygon (a
// weight of 1e18 in single precision when its area vanishes).
__device__ inline void polygon_centroid(int n, const double* p, double& cx, double& cy) {
  if (n == 1) {
    cx = p[0];
    cy = p[1];
    return;
  }
  if (n == 2) {
    cx = 0.5 * (p[0] + p[2]);
    cy = 0.5 * (p[1] + p[3]);
    return;
  }
  double area = 0, sx = 0, sy = 0, w;
  for (int i = 0; i + 1 < n; ++i) {
    const double* u = p + 2 * i;
    w = u[0] * u[3] - u[2] * u[1];
    area += w;
    sx += w * (u[0] + u[2]);
    sy += w * (u[1] + u[3]);
  }
  const double* last = p + 2 * (n - 1);
  w = last[0] * p[1] - p[0] * last[1];
  const double scale = fabs(area + w) > 2.220446049250313e-16 ? 1 / (3 * (area + w))
                                                               : static_cast<double>(1e18f);
  cx = scale * (sx + w * (last[0] + p[0]));
  cy = scale * (sy + w * (last[1] + p[1]));
}

// Of n contact points keep m: `first` (the deepest), then for each step of
// 2 pi / m around the centroid the unused point whose angle is nearest.
__device__ inline void spread_contacts(int n, const double* p, int m, int first, int* keep) {
  double cx, cy;
  polygon_centroid(n, p, cx, cy);
  double ang[8];
  bool used[8];
  for (int i = 0; i < n; ++i) {
    ang[i] = atan2(p[2 * i + 1] - cy, p[2 * i] - cx);
    used[i] = false;
  }
  const double pi = 3.14159265358979323846;
  used[first] = true;
  keep[0] = first;
  for (int j = 1; j < m; ++j) {
    double target = j * (2 * pi / m) + ang[first];
    if (target > pi) target -= 2 * pi;
    double best = HUGE_VAL;  // every angular gap (at most pi) is below it
    int pick = first;
    for (int i = 0; i < n; ++i) {
      if (used[i]) continue;
      double gap = fabs(ang[i] - target);
      if (gap > pi) gap = 2 * pi - gap;
      if (gap < best) {
        best = gap;
        pick = i;
      }
    }
    used[pick] = true;
    keep[j] = pick;
  }
}

// The two boxes in box 1's frame: rel = R1^T R2, its absolute values, and
// box 2's centre.
struct BoxFrames {
  const double* R1;
  const double* R2;
  D3 t1, t2, d;   // centres, d = t2 - t1 (world)
  D3 d1;          // d in box 1's frame
  D3 h1, h2;      // half extents
  double rel[9], arel[9];
};

// The separating-axis search: false when some axis separates the boxes.
// Otherwise the axis of least penetration (edge axes favoured by 1.05):
// `feature` 1-3 box 1's faces, 4-6 box 2's, 7-15 edge pairs; `depth` its
// penetration; `normal` in the world frame, from box 1 towards box 2.
__device__ inline bool least_penetration(BoxFrames& f, int& feature, double& depth, D3& normal) {
  const double* R = f.rel;
  double* Q = f.arel;
  const D3& d1 = f.d1;
  double best = -1.7976931348623157e308;
  int face = -1;
  bool face_of_1 = true, flip = false;
  feature = 0;
  for (int k = 0; k < 3; ++k) {  // box 1's face normals
    const double along = d1.v[k];
    const double sep = fabs(along) - (dot(mk(Q[3 * k], Q[3 * k + 1], Q[3 * k + 2]), f.h2) + f.h1.v[k]);
    if (sep > 0) return false;
    if (sep > best) {
      best = sep;
      face = k;
      face_of_1 = true;
      flip = along < 0;
      feature = 1 + k;
    }
  }
  for (int k = 0; k < 3; ++k) {  // box 2's face normals
    const double along = dot(col(f.R2, k), f.d);
    const double sep = fabs(along) - (dot(mk(Q[k], Q[3 + k], Q[6 + k]), f.h1) + f.h2.v[k]);
    if (sep > 0) return false;
    if (sep > best) {
      best = sep;
      face = k;
      face_of_1 = false;
      flip = along < 0;
      feature = 4 + k;
    }
  }
  // edge axes: |rel| grown by 10 eps times the largest half extent (at least 1)
  const double eps = 2.220446049250313e-16;
  {
    const double* a = f.h1.v;
    const double* b = f.h2.v;
    const double ma = a[0] > a[1] ? (a[0] > a[2] ? a[0] : a[2]) : (a[1] > a[2] ? a[1] : a[2]);
    const double mb = b[0] > b[1] ? (b[0] > b[2] ? b[0] : b[2]) : (b[1] > b[2] ? b[1] : b[2]);
    double mm = ma > mb ? ma : mb;
    if (mm < 1.0) mm = 1.0;
    const double grow = mm * 10.0 * eps;
    for (int i = 0; i < 9; ++i) Q[i] += grow;
  }
  const double* a = f.h1.v;
  const double* b = f.h2.v;
  D3 edge_axis = mk(0, 0, 0);
  for (int i = 0; i < 3; ++i) {    // box 1's edge i
    const int i1 = (i + 1) % 3, i2 = (i + 2) % 3;
    for (int j = 0; j < 3; ++j) {  // box 2's edge j
      const int j1 = j == 0 ? 1 : 0, j2 = j == 2 ? 1 : 2;
      const double along = d1.v[i2] * R[3 * i1 + j] - d1.v[i1] * R[3 * i2 + j];
      const double reach = ((a[i1] * Q[3 * i2 + j] + a[i2] * Q[3 * i1 + j]) + b[j1] * Q[3 * i + j2]) +
                           b[j2] * Q[3 * i + j1];
      double sep = fabs(along) - reach;
      if (sep > 0) return false;
      D3 ax;
      ax.v[i] = 0;
      ax.v[i1] = -R[3 * i2 + j];
      ax.v[i2] = R[3 * i1 + j];
      const double len = nrm(ax);
      if (len > eps) {
        sep /= len;
        if (sep * 1.05 > best) {
          best = sep;
          face = -1;
          edge_axis = dvd(ax, len);
          flip = along < 0;
          feature = 7 + 3 * i + j;
        }
      }
    }
  }
  if (!feature) return false;
  normal = face != -1 ? col(face_of_1 ? f.R1 : f.R2, face) : rot(f.R1, edge_axis);
  if (flip) normal = neg(normal);
  depth = -best;
  return true;
}

// An edge-edge feature: one contact half way between the closest points of
// the two edge lines (each edge through the box corner furthest along the
// normal's side).
__device__ inline Contact edge_edge_contact(const BoxFrames& f, int feature, const D3& normal,
                                            double depth) {
  D3 pa = f.t1;
  for (int j = 0; j < 3; ++j) {
    const double sign = dot(col(f.R1, j), normal) > 0 ? 1 : -1;
    pa = add(pa, mul(col(f.R1, j), f.h1.v[j] * sign));
  }
  D3 pb = f.t2;
  for (int j = 0; j < 3; ++j) {
    const double sign = dot(col(f.R2, j), normal) > 0 ? -1 : 1;
    pb = add(pb, mul(col(f.R2, j), f.h2.v[j] * sign));
  }
  const D3 ua = col(f.R1, (feature - 7) / 3), ub = col(f.R2, (feature - 7) % 3);
  double ta, tb;
  closest_line_params(pa, ua, pb, ub, ta, tb);
  pa = add(pa, mul(ua, ta));
  pb = add(pb, mul(ub, tb));
  Contact c;
  c.normal = normal;
  c.pos = mul(add(pa, pb), 0.5);
  c.depth = depth;
  return c;
}

// A face feature: the incident box's face most anti-parallel to the normal,
// projected onto the reference face's plane, clipped to its rectangle; the
// points behind the reference face are contacts (their depth below it), at
// most 4 kept, the deepest first.
__device__ inline int face_contacts(const BoxFrames& f, int feature, const D3& normal, bool want,
                                    Contact* out) {
  const bool ref1 = feature <= 3;  // reference face on box 1
  const double* Ra = ref1 ? f.R1 : f.R2;
  const double* Rb = ref1 ? f.R2 : f.R1;
  const D3 Pa = ref1 ? f.t1 : f.t2, Pb = ref1 ? f.t2 : f.t1;
  const D3 Sa = ref1 ? f.h1 : f.h2, Sb = ref1 ? f.h2 : f.h1;
  const D3 nref = ref1 ? normal : neg(normal);
  const D3 nb = rotT(Rb, nref);
  const D3 anb = mk(fabs(nb.v[0]), fabs(nb.v[1]), fabs(nb.v[2]));
  // the incident box's axis closest to the normal (lanr) and the other two
  int lanr, e1, e2;
  if (anb.v[1] > anb.v[0]) {
    if (anb.v[1] > anb.v[2]) {
      e1 = 0;
      lanr = 1;
      e2 = 2;
    } else {
      e1 = 0;
      e2 = 1;
      lanr = 2;
    }
  } else {
    if (anb.v[0] > anb.v[2]) {
      lanr = 0;
      e1 = 1;
      e2 = 2;
    } else {
      e1 = 0;
      e2 = 1;
      lanr = 2;
    }
  }
  const D3 face_c = nb.v[lanr] < 0 ? add(sub(Pb, Pa), mul(col(Rb, lanr), Sb.v[lanr]))
                                   : sub(sub(Pb, Pa), mul(col(Rb, lanr), Sb.v[lanr]));
  const int ref_axis = ref1 ? feature - 1 : feature - 4;
  const int u = ref_axis == 0 ? 1 : 0, v = ref_axis == 2 ? 1 : 2;  // the reference face's axes
  const double c1 = dot(col(Ra, u), face_c), c2 = dot(col(Ra, v), face_c);
  double m11 = dot(col(Rb, e1), col(Ra, u)), m12 = dot(col(Rb, e2), col(Ra, u));
  double m21 = dot(col(Rb, e1), col(Ra, v)), m22 = dot(col(Rb, e2), col(Ra, v));
  const double k1 = m11 * Sb.v[e1], k2 = m21 * Sb.v[e1], k3 = m12 * Sb.v[e2], k4 = m22 * Sb.v[e2];
  const double quad[8] = {c1 - k1 - k3, c2 - k2 - k4, c1 - k1 + k3, c2 - k2 + k4,
                          c1 + k1 + k3, c2 + k2 + k4, c1 + k1 - k3, c2 + k2 - k4};
  const double rect[2] = {Sa.v[u], Sa.v[v]};
  double clipped[16];
  const int nclip = clip_quad_to_rect(rect, quad, clipped);
  if (nclip < 1) return 0;
  // back from the face's (u, v) coordinates to points on the incident face
  const double inv_det = 1.f / (m11 * m22 - m12 * m21);
  m11 *= inv_det;
  m12 *= inv_det;
  m21 *= inv_det;
  m22 *= inv_det;
  D3 pts[8];
  double dep[8];
  int cnt = 0;
  for (int j = 0; j < nclip; ++j) {
    const double du = clipped[2 * j] - c1, dv = clipped[2 * j + 1] - c2;
    const double s1 = m22 * du - m12 * dv;
    const double s2 = -m21 * du + m11 * dv;
    pts[cnt] = add(add(face_c, mul(col(Rb, e1), s1)), mul(col(Rb, e2), s2));
    dep[cnt] = Sa.v[ref_axis] - dot(nref, pts[cnt]);
    if (dep[cnt] >= 0) {
      clipped[2 * cnt] = clipped[2 * j];
      clipped[2 * cnt + 1] = clipped[2 * j + 1];
      ++cnt;
    }
  }
  if (cnt < 1) return 0;
  int order[8] = {0, 1, 2, 3, 4, 5, 6, 7};
  if (cnt > 4) {
    int deepest = 0;
    for (int i = 1; i < cnt; ++i)
      if (dep[i] > dep[deepest]) deepest = i;
    spread_contacts(cnt, clipped, 4, deepest, order);
    cnt = 4;
  }
  if (!want) return 0;
  for (int j = 0; j < cnt; ++j) {
    const int i = order[j];
    out[j].normal = normal;
    out[j].pos = feature < 4 ? add(add(pts[i], Pa), mul(normal, dep[i] / 2))
                             : sub(add(pts[i], Pa), mul(normal, dep[i] / 2));
    out[j].depth = dep[i];
  }
  return cnt;
}

// box_box-inl.h:253-862 (boxBox2, world frame, up to 4 contacts).
__device__ inline bool box_box(const Shape& b1, const Xf& tf1, const Shape& b2, const Xf& tf2,
                               bool want, Contact* out, int& n) {
  BoxFrames f;
  f.R1 = tf1.R;
  f.R2 = tf2.R;
  f.t1 = tf1.t;
  f.t2 = tf2.t;
  f.d = sub(f.t2, f.t1);
  f.d1 = rotT(f.R1, f.d);
  f.h1 = mul(mk(b1.dims[0], b1.dims[1], b1.dims[2]), 0.5);
  f.h2 = mul(mk(b2.dims[0], b2.dims[1], b2.dims[2]), 0.5);
  mulTN(f.R1, f.R2, f.rel);
  for (int i = 0; i < 9; ++i) f.arel[i] = fabs(f.rel[i]);
  int feature;
  double depth;
  D3 normal;
  n = 0;
  if (!least_penetration(f, feature, depth, normal)) return false;
  if (feature > 6) {
    if (want) {
      out[0] = edge_edge_contact(f, feature, normal, depth);
      n = 1;
    }
    return true;
  }
  n = face_contacts(f, feature, normal, want, out);
  return true;
}

// ------------------------------------------------------------ dispatch
__device__ __forceinline__ void flip(Contact* c, int n) {
  for (int i = 0; i < n; ++i) c[i].normal = neg(c[i].normal);
}

// The pairs GJKSolver_indep::shapeIntersect answers in closed form
// (gjk_solver_indep-inl.h:250-305): two spheres, two boxes, and a sphere
// with a box, capsule or cylinder in either order.
__device__ __forceinline__ bool is_closed_form(int t1, int t2) {
  if (t1 == t2) return t1 == kSphere || t1 == kBox;
  const int o = t1 == kSphere ? t2 : (t2 == kSphere ? t1 : -1);
  return o == kBox || o == kCapsule || o == kCylinder;
}

// Closed-form shapeIntersect: up to 4 contacts (two boxes), else 1.
__device__ inline bool closed_form(const Shape& s1, const Xf& a, const Shape& s2, const Xf& b,
                                   bool want, Contact* out, int& n) {
  const int t1 = s1.type, t2 = s2.type;
  n = 0;
  bool r;
  if (t1 == kSphere && t2 == kSphere)
    r = sphere_sphere(s1, a, s2, b, want, out, n);
  else if (t1 == kBox && t2 == kBox)
    r = box_box(s1, a, s2, b, want, out, n);
  else if (t1 == kSphere && t2 == kCapsule)
    r = sphere_capsule(s1, a, s2, b, want, out, n);
  else if (t1 == kCapsule && t2 == kSphere) {
    r = sphere_capsule(s2, b, s1, a, want, out, n);
    if (want) flip(out, n);
  } else if (t1 == kSphere && t2 == kBox)
    r = sphere_box(s1, a, s2, b, want, out, n);
  else if (t1 == kBox && t2 == kSphere) {
    r = sphere_box(s2, b, s1, a, want, out, n);
    if (want) flip(out, n);
  } else if (t1 == kSphere && t2 == kCylinder)
    r = sphere_cylinder(s1, a, s2, b, want, out, n);
  else {
    r = sphere_cylinder(s2, b, s1, a, want, out, n);
    if (want) flip(out, n);
  }
  if (!want) n = 0;
  return r;
}

// shape_collision_traversal_node-inl.h:69-134 with enable_cost off, given
// the pair's intersection result: the contacts added for cap free slots
// (deepest first when they do not all fit; one contact without geometry
// without enable_contact).
__device__ inline int keep_contacts(bool hit, Contact* c, int n, int cap, bool enable_contact) {
  if (!hit || cap <= 0) return 0;
  if (!enable_contact) {
    c[0].normal = mk(0, 0, 0);
    c[0].pos = mk(0, 0, 0);
    c[0].depth = 0;
    return 1;
  }
  if (cap >= n) return n;
  for (int i = 0; i < cap; ++i) {
    int best = i;
    for (int j = i + 1; j < n; ++j)
      if (c[j].depth > c[best].depth) best = j;
    const Contact t = c[i];
    c[i] = c[best];
    c[best] = t;
  }
  return cap;
}

}  // namespace shp
}  // namespace fclx
