/*
 * fclo — CPU oracle for the FCL mesh collide()/distance() hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This is a plain-C restatement of FCL 0.7.0's
 * double-precision BVH build, BVH-vs-BVH collision traversal and BVH-vs-BVH
 * distance traversal.  It exists to check the HIP engine in fcl_amd/; only
 * the test suite, the smoke check and the CPU-baseline leg of the benchmark
 * may link or call it.  The product never does.
 *
 * Every function names the reference file:line it restates (paths relative to
 * the FCL source tree).  The reference evaluates small vector expressions
 * through Eigen, whose association order is not fixed by FCL; this
 * restatement fixes one order and uses it everywhere:
 *   dot(a,b)       = (a0*b0 + a1*b1) + a2*b2
 *   (M*v)_i        = dot(row i of M, v)
 *   (A*B)_ij       = (A_i0*B_0j + A_i1*B_1j) + A_i2*B_2j
 *   (A^T*B)_ij     = (A_0i*B_0j + A_1i*B_1j) + A_2i*B_2j
 *   cross(a,b)     = (a1*b2 - a2*b1, a2*b0 - a0*b2, a0*b1 - a1*b0)
 *   squaredNorm(v) = (v0*v0 + v1*v1) + v2*v2
 * with no fused multiply-add (built with -ffp-contract=off).
 *
 * Matrices are 9 doubles, row-major: M(i,j) = m[3*i+j].  A pose is 12
 * doubles: rotation (row-major) followed by translation.
 */
#ifndef FCLO_H
#define FCLO_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum { FCLO_BV_OBBRSS = 0, FCLO_BV_RSS = 1 };

/* One BVH node: an OBB and an RSS sharing the same fitted axes, plus the
 * reference's child/primitive link (BV_node_base.h:49-76). */
typedef struct {
  double obb_axis[9];
  double obb_to[3];
  double obb_extent[3];
  double rss_axis[9];
  double rss_to[3];
  double rss_l[2];
  double rss_r;
  int first_child;      /* >=0: left child index; <0: -(primitive+1) */
  int first_primitive;
  int num_primitives;
} fclo_node;

typedef struct {
  int bv_type;          /* FCLO_BV_OBBRSS or FCLO_BV_RSS */
  int num_vertices;
  int num_tris;
  int num_bvs;
  double *vertices;     /* 3*num_vertices */
  int *tris;            /* 3*num_tris */
  fclo_node *bvs;       /* 2*num_tris-1 */
  unsigned *primitive_indices;
} fclo_model;

/* ---- math ---------------------------------------------------------------- */
double fclo_dot(const double a[3], const double b[3]);
void fclo_cross(const double a[3], const double b[3], double out[3]);
void fclo_eigen_old(const double m[9], double dout[3], double vout[9]);
void fclo_relative_transform(const double tf1[12], const double tf2[12],
                             double R[9], double T[3]);
void fclo_inverse_compose(const double tf1[12], const double tf2[12],
                          double out[12]);

/* ---- BVH build ----------------------------------------------------------- */
/* Returns 0 on success; negative for an empty model. */
int fclo_model_build(fclo_model *m, int bv_type, const double *vertices,
                     int num_vertices, const int *tris, int num_tris);
void fclo_model_free(fclo_model *m);
/* BVHModel::beginReplaceModel / replaceSubModel(vertices) /
 * endReplaceModel(refit = true, bottomup) (BVH_model-inl.h:521-620): the
 * vertices replaced (same count), the hierarchy refit keeping its topology
 * -- top-down refits every node as the build fits it, bottom-up fits each
 * leaf's triangle (fit3) and merges children (operator+) up to the root
 * (refitTree, :941-1076). */
int fclo_model_refit(fclo_model *m, const double *vertices, int bottomup);
int fclo_model_depth(const fclo_model *m);
double fclo_node_size(const fclo_model *m, int b);

/* ---- primitive tests ----------------------------------------------------- */
int fclo_obb_disjoint(const double B[9], const double T[3], const double a[3],
                      const double b[3]);
int fclo_obb_overlap(const double R0[9], const double T0[3],
                     const fclo_node *b1, const fclo_node *b2);
int fclo_intersect_triangle(const double P1[3], const double P2[3],
                            const double P3[3], const double Q1[3],
                            const double Q2[3], const double Q3[3],
                            const double R[9], const double T[3]);
/* The same test with the reference's contact output (enable_contact):
 * up to 2 contact points (frame of P), the normal and the depth. */
int fclo_intersect_triangle_contact(const double P1[3], const double P2[3],
                                    const double P3[3], const double Q1[3],
                                    const double Q2[3], const double Q3[3],
                                    const double R[9], const double T[3],
                                    double contacts[6], unsigned *num_contacts,
                                    double *depth, double normal[3]);
double fclo_rect_distance(const double Rab[9], const double Tab[3],
                          const double a[2], const double b[2], double P[3],
                          double Q[3]);
double fclo_rss_distance(const double R0[9], const double T0[3],
                         const fclo_node *b1, const fclo_node *b2);
double fclo_tri_distance(const double S[9], const double T[9], double P[3],
                         double Q[3]);

/* ---- collision ----------------------------------------------------------- */
typedef struct {
  int num_contacts;
  int64_t num_bv_tests;
  int64_t num_leaf_tests;
} fclo_collide_stats;

/* One mesh-mesh collide() query (collision-inl.h:93-149 -> orientedMeshCollide).
 * cap: number of contacts that may still be added (num_max_contacts minus the
 * contacts already in the result).  Pairs (b1,b2) are written in traversal
 * order into pairs[2*k]; at most cap pairs are written. */
void fclo_collide(const fclo_model *m1, const double tf1[12],
                  const fclo_model *m2, const double tf2[12], int cap,
                  int *pairs, fclo_collide_stats *stats);

/* collide() with a BVTT front list (detail::collide(node, front_list),
 * collision_node-inl.h:71-82; traversal_recurse-inl.h:82-129, 463-521): a
 * front of n_front entries in (2 ints each; 0 entries: traverse from the
 * roots), the new front out (front_room entries, *n_out its size). */
void fclo_collide_front(const fclo_model *m1, const double tf1[12], const fclo_model *m2,
                        const double tf2[12], int cap, int store, int *pairs, const int *front,
                        int n_front, int *front_out, int front_room, int *n_out,
                        fclo_collide_stats *stats);

/* collide() with enable_contact: cap counts contacts (a triangle pair adds
 * 0-2); the first min(count, store) contacts are written as (b1,b2) pairs
 * plus world-frame pos[3], normal[3] and depth. */
void fclo_collide_contact(const fclo_model *m1, const double tf1[12],
                          const fclo_model *m2, const double tf2[12], int cap,
                          int store, int *pairs, double *pos, double *normal,
                          double *depth, fclo_collide_stats *stats);

/* ---- distance ------------------------------------------------------------ */
typedef struct {
  double min_distance;
  double p1[3], p2[3];   /* nearest points, world frame */
  int b1, b2;
  int64_t num_bv_tests;
  int64_t num_leaf_tests;
} fclo_distance_out;

void fclo_distance(const fclo_model *m1, const double tf1[12],
                   const fclo_model *m2, const double tf2[12],
                   int enable_nearest_points, fclo_distance_out *out);

/* ---- brute force over all triangle pairs (no hierarchy) ------------------ */
int64_t fclo_collide_all_pairs(const fclo_model *m1, const double tf1[12],
                               const fclo_model *m2, const double tf2[12],
                               int *pairs, int64_t store);
double fclo_distance_all_pairs(const fclo_model *m1, const double tf1[12],
                               const fclo_model *m2, const double tf2[12],
                               int ids[2]);

/* ---- batch drivers (OpenMP over queries; the CPU baseline) ---------------- */
/* Query q collides models[q_m1[q]] vs models[q_m2[q]].  counts[q] receives
 * the number of contacts (<= cap); the first min(counts[q], store) pairs are
 * written to pairs[2*q*store ...].  store may be 0. */
void fclo_collide_batch(const fclo_model *const *models, int num_queries,
                        const int *q_m1, const int *q_m2, const double *tf1s,
                        const double *tf2s, int cap, int store, int *counts,
                        int *pairs, int64_t *bv_tests, int64_t *leaf_tests,
                        int threads);
void fclo_collide_batch_flat(const fclo_model *const *models, int num_queries,
                             const int *q_m1, const int *q_m2, const double *tf1s,
                             const double *tf2s, int cap, const int64_t *offsets,
                             int *counts, int *pairs, int threads);
void fclo_distance_batch(const fclo_model *const *models, int num_queries,
                         const int *q_m1, const int *q_m2, const double *tf1s,
                         const double *tf2s, int enable_nearest_points,
                         double *dist, double *points, int *ids,
                         int64_t *bv_tests, int64_t *leaf_tests, int threads);

/* ---- primitive shapes, GST_INDEP solver (fclo_shape.c) -------------------
 * Pinned by the reference's shapeIntersectionGJK_* known answers (booleans,
 * contact counts, normals at the reference's tolerances, closed-form points
 * and depths), an LP intersection test on polytope pairs, and EPA's bound
 * 0 <= depth <= true penetration.  Parity unpinned: EPA's depth and contact
 * point beyond that bound (the reference's tests never check them for this
 * solver, and its EPA can stop early on near-coplanar supports), the order
 * of equally deep contacts when a cap cuts them (std::partial_sort), and
 * box-box contact culling's atan2 (libm here). */
enum {
  FCLO_SHAPE_BOX = 0,      /* dims = side lengths x, y, z */
  FCLO_SHAPE_SPHERE = 1,   /* dims[0] = radius */
  FCLO_SHAPE_CAPSULE = 2,  /* dims[0] = radius, dims[1] = lz */
  FCLO_SHAPE_CYLINDER = 3, /* dims[0] = radius, dims[1] = lz */
  FCLO_SHAPE_CONVEX = 4,   /* vertices (local frame), num_vertices */
  FCLO_SHAPE_TRIANGLE = 5  /* TriangleP: vertices = a, b, c (only as the mesh
                              side of the mesh-shape leaf test) */
};

typedef struct {
  int type;
  int num_vertices;
  double dims[3];
  const double *vertices;
} fclo_shape;

typedef struct {
  double normal[3];
  double pos[3];
  double depth;
} fclo_contact;

/* GJKSolver_indep::shapeIntersect with the contact vector (up to 4 for two
 * boxes, else 1); out may be NULL for the boolean. */
int fclo_shape_intersect(const fclo_shape *s1, const double tf1[12], const fclo_shape *s2,
                         const double tf2[12], double tol, fclo_contact *out, int *n);
/* collide() of two shapes: contacts added (<= cap) into out (room for 8). */
int fclo_shape_collide(const fclo_shape *s1, const double tf1[12], const fclo_shape *s2,
                       const double tf2[12], int cap, int enable_contact, double tol,
                       fclo_contact *out);
/* the same per query (OpenMP): counts[q]; out gets cap slots per query;
 * passes (optional) gets the passes of GJK's loop and of EPA's loop per
 * query (0 where a closed form answers or EPA falls back) */
void fclo_shape_collide_batch(const fclo_shape *shapes, int num_queries, const int *qs,
                              const double *tf1s, const double *tf2s, int cap,
                              int enable_contact, double tol, int *counts, fclo_contact *out,
                              int threads, int *passes);
/* a shape as a 40-byte record (the engine's shape-table layout) */
typedef struct {
  int type;
  int num_vertices;
  double dims[3];
  long long vertex_offset; /* CONVEX: first vertex in the pool */
} fclo_shape_rec;
/* fclo_shape_collide_batch over such records and their vertex pool */
void fclo_shape_collide_records(const fclo_shape_rec *recs, const double *pool,
                                int num_queries, const int *qs, const double *tf1s,
                                const double *tf2s, int cap, int enable_contact, double tol,
                                int *counts, fclo_contact *out, int threads, int *passes);

/* ---- mesh vs primitive shape, BVHModel<OBBRSS> (fclo_shape.c) ------------
 * fcl::collide(mesh, tf1, shape, tf2) for an OBBRSS mesh and a Box, Sphere,
 * Capsule, Cylinder or Convex (>= 4 vertices) with the GST_INDEP solver:
 * orientedBVHShapeCollide (collision_func_matrix-inl.h:375-419, 500-521),
 * the shape's OBBRSS from its bound vertices (utility-inl.h:124-133 with
 * BV_fitter fitn), descent of the mesh only (traversal_recurse-inl.h:82-129,
 * BVTesting = OBB overlap of the shape's box with the node's box posed by
 * tf1), leaf test meshShapeCollisionOrientedNodeLeafTesting
 * (mesh_shape_collision_traversal_node-inl.h:205-283): sphereTriangleIntersect
 * for a sphere, GJK + EPA on (shape, TriangleP) for the rest.
 * Contacts are (b1 = triangle id, b2 = NONE) in the reference's order, at
 * most cap; with enable_contact each has the reference's point, normal and
 * penetration depth.  Returns the count; stats optional. */
void fclo_shape_obb(const fclo_shape *s, const double tf[12], double axis[9], double to[3],
                    double extent[3]);
int fclo_mesh_shape_collide(const fclo_model *m, const double tf1[12], const fclo_shape *s,
                            const double tf2[12], int cap, int enable_contact, double tol,
                            int *prims, fclo_contact *out, fclo_collide_stats *st);
/* per query q: mesh models[mesh_ids[q]], shape recs[shape_ids[q]]; prims and
 * out get cap slots per query (either may be NULL), stats 2 per query */
void fclo_mesh_shape_collide_batch(const fclo_model *const *models, const int *mesh_ids,
                                   const fclo_shape_rec *recs, const double *pool,
                                   int num_queries, const int *shape_ids, const double *tf1s,
                                   const double *tf2s, int cap, int enable_contact, double tol,
                                   int *counts, int *prims, fclo_contact *out, int64_t *stats,
                                   int threads);
/* sphereTriangleIntersect (sphere_triangle-inl.h:138-230) on its own */
int fclo_sphere_triangle(double radius, const double center[3], const double P1[3],
                         const double P2[3], const double P3[3], double point[3],
                         double *depth, double normal[3]);
void fclo_axis_from_eigen(const double E[9], const double s[3], double axis[9]);

#ifdef __cplusplus
}
#endif

#endif
