/*
 * fclo_shape.c — CPU oracle for FCL's primitive-shape collide() with the
 * GST_INDEP solver.  TEST INFRASTRUCTURE ONLY (see fclo.h).
 *
 * Restates, in the evaluation order fixed in fclo.h:
 *   getSupport          narrowphase/detail/convexity_based_algorithm/minkowski_diff-inl.h:64-205
 *   MinkowskiDiff       minkowski_diff-inl.h:209-273
 *   GJK                 convexity_based_algorithm/gjk-inl.h:82-315
 *   EPA                 convexity_based_algorithm/epa-inl.h:54-396
 *   simplex projection  math/detail/project-inl.h:123-316 (the *Origin forms)
 *   closed forms        primitive_shape_algorithm/sphere_sphere-inl.h:66-90,
 *                       sphere_box-inl.h:69-160, sphere_capsule-inl.h:59-118,
 *                       sphere_cylinder-inl.h:63-200, box_box-inl.h:82-862
 *   dispatch            narrowphase/detail/gjk_solver_indep-inl.h:110-305
 *   leaf bookkeeping    traversal/collision/shape_collision_traversal_node-inl.h:69-134
 */
#include <float.h>
#include <math.h>
#include <string.h>

#include "fclo.h"

typedef struct {
  double x[3];
} v3;

static v3 V(double a, double b, double c) {
  v3 r = {{a, b, c}};
  return r;
}
static v3 vadd(v3 a, v3 b) { return V(a.x[0] + b.x[0], a.x[1] + b.x[1], a.x[2] + b.x[2]); }
static v3 vsub(v3 a, v3 b) { return V(a.x[0] - b.x[0], a.x[1] - b.x[1], a.x[2] - b.x[2]); }
static v3 vneg(v3 a) { return V(-a.x[0], -a.x[1], -a.x[2]); }
static v3 vmul(v3 a, double s) { return V(a.x[0] * s, a.x[1] * s, a.x[2] * s); }
static v3 vdiv(v3 a, double s) { return V(a.x[0] / s, a.x[1] / s, a.x[2] / s); }
static double vdot(v3 a, v3 b) { return (a.x[0] * b.x[0] + a.x[1] * b.x[1]) + a.x[2] * b.x[2]; }
static v3 vcross(v3 a, v3 b) {
  return V(a.x[1] * b.x[2] - a.x[2] * b.x[1], a.x[2] * b.x[0] - a.x[0] * b.x[2],
           a.x[0] * b.x[1] - a.x[1] * b.x[0]);
}
static double vsq(v3 a) { return vdot(a, a); }
static double vnorm(v3 a) { return sqrt(vsq(a)); }
/* Eigen normalized(): unchanged when the norm is zero */
static v3 vnormalized(v3 a) {
  double z = vsq(a);
  return z > 0 ? vdiv(a, sqrt(z)) : a;
}
static double triple(v3 a, v3 b, v3 c) { return vdot(a, vcross(b, c)); }

/* rigid transform: R (row-major) and t */
typedef struct {
  double R[9];
  v3 t;
} xf;

static xf xf_from(const double tf[12]) {
  xf r;
  memcpy(r.R, tf, 9 * sizeof(double));
  r.t = V(tf[9], tf[10], tf[11]);
  return r;
}
static v3 rot(const double R[9], v3 p) {
  return V(vdot(V(R[0], R[1], R[2]), p), vdot(V(R[3], R[4], R[5]), p),
           vdot(V(R[6], R[7], R[8]), p));
}
static v3 rotT(const double R[9], v3 p) {
  return V(vdot(V(R[0], R[3], R[6]), p), vdot(V(R[1], R[4], R[7]), p),
           vdot(V(R[2], R[5], R[8]), p));
}
static v3 xf_apply(const xf *T, v3 p) { return vadd(rot(T->R, p), T->t); }
static v3 col(const double R[9], int j) { return V(R[j], R[3 + j], R[6 + j]); }
/* A^T B */
static void mulTN(const double A[9], const double B[9], double out[9]) {
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j)
      out[3 * i + j] = (A[i] * B[j] + A[3 + i] * B[3 + j]) + A[6 + i] * B[6 + j];
}
/* a^-1 * b for isometries: (Ra^T Rb, Ra^T tb + (-(Ra^T ta))) */
static xf inv_compose(const xf *a, const xf *b) {
  xf r;
  mulTN(a->R, b->R, r.R);
  r.t = vadd(rotT(a->R, b->t), vneg(rotT(a->R, a->t)));
  return r;
}

/* ------------------------------------------------------------- supports */
static v3 get_support(const fclo_shape *s, v3 d) {
  switch (s->type) {
    case FCLO_SHAPE_BOX:
      return V(d.x[0] > 0 ? s->dims[0] / 2 : -s->dims[0] / 2,
               d.x[1] > 0 ? s->dims[1] / 2 : -s->dims[1] / 2,
               d.x[2] > 0 ? s->dims[2] / 2 : -s->dims[2] / 2);
    case FCLO_SHAPE_SPHERE:
      return vmul(d, s->dims[0]);
    case FCLO_SHAPE_CAPSULE: {
      double half_h = s->dims[1] * 0.5;
      v3 v = vmul(d, s->dims[0]);
      v3 p1 = vadd(V(0, 0, half_h), v), p2 = vadd(V(0, 0, -half_h), v);
      return vdot(d, p1) > vdot(d, p2) ? p1 : p2;
    }
    case FCLO_SHAPE_CYLINDER: {
      double zdist = sqrt(d.x[0] * d.x[0] + d.x[1] * d.x[1]);
      double half_h = s->dims[1] * 0.5;
      if (zdist == 0.0) return V(0, 0, d.x[2] > 0 ? half_h : -half_h);
      double k = s->dims[0] / zdist;
      return V(k * d.x[0], k * d.x[1], d.x[2] > 0 ? half_h : -half_h);
    }
    case FCLO_SHAPE_TRIANGLE: {
      /* minkowski_diff-inl.h:79-100 (TriangleP): the first of the larger
       * dot products, c only when strictly above */
      v3 a = V(s->vertices[0], s->vertices[1], s->vertices[2]);
      v3 b = V(s->vertices[3], s->vertices[4], s->vertices[5]);
      v3 c = V(s->vertices[6], s->vertices[7], s->vertices[8]);
      double da = vdot(d, a), db = vdot(d, b), dc = vdot(d, c);
      if (da > db) return dc > da ? c : a;
      return dc > db ? c : b;
    }
    case FCLO_SHAPE_CONVEX: {
      double maxdot = -DBL_MAX;
      v3 best = V(0, 0, 0);
      for (int i = 0; i < s->num_vertices; ++i) {
        v3 p = V(s->vertices[3 * i], s->vertices[3 * i + 1], s->vertices[3 * i + 2]);
        double dt = vdot(d, p);
        if (dt > maxdot) {
          best = p;
          maxdot = dt;
        }
      }
      return best;
    }
  }
  return V(0, 0, 0);
}

typedef struct {
  const fclo_shape *s0, *s1;
  double toshape1[9]; /* R2^T R1 */
  xf toshape0;        /* tf1^-1 tf2 */
} mdiff;

static v3 support0(const mdiff *m, v3 d) { return get_support(m->s0, d); }
static v3 support1(const mdiff *m, v3 d) {
  return xf_apply(&m->toshape0, get_support(m->s1, rot(m->toshape1, d)));
}
static v3 msupport(const mdiff *m, v3 d) { return vsub(support0(m, d), support1(m, vneg(d))); }

/* ---------------------------------------------------- simplex projection */
typedef struct {
  double p[4];
  double sqr;
  int encode;
} proj;

static proj proj_init(void) {
  proj r;
  r.p[0] = r.p[1] = r.p[2] = r.p[3] = 0;
  r.sqr = -1;
  r.encode = 0;
  return r;
}

static proj project_line_origin(v3 a, v3 b) {
  proj r = proj_init();
  v3 d = vsub(b, a);
  double l = vsq(d);
  if (l > 0) {
    double t = -vdot(a, d);
    r.p[1] = (t >= l) ? 1 : ((t <= 0) ? 0 : (t / l));
    r.p[0] = 1 - r.p[1];
    if (t >= l) {
      r.sqr = vsq(b);
      r.encode = 2;
    } else if (t <= 0) {
      r.sqr = vsq(a);
      r.encode = 1;
    } else {
      r.sqr = vsq(vadd(a, vmul(d, r.p[1])));
      r.encode = 3;
    }
  }
  return r;
}

static proj project_triangle_origin(v3 a, v3 b, v3 c) {
  static const int nexti[3] = {1, 2, 0};
  proj r = proj_init();
  v3 vt[3] = {a, b, c};
  v3 dl[3] = {vsub(a, b), vsub(b, c), vsub(c, a)};
  v3 n = vcross(dl[0], dl[1]);
  double l = vsq(n);
  if (l > 0) {
    double mindist = -1;
    for (int i = 0; i < 3; ++i) {
      if (vdot(vt[i], vcross(dl[i], n)) > 0) {
        int j = nexti[i];
        proj rl = project_line_origin(vt[i], vt[j]);
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
      double d = vdot(a, n);
      double s = sqrt(l);
      v3 o = vmul(n, d / l);
      mindist = vsq(o);
      r.encode = 7;
      r.p[0] = vnorm(vcross(dl[1], vsub(b, o))) / s;
      r.p[1] = vnorm(vcross(dl[2], vsub(c, o))) / s;
      r.p[2] = 1 - r.p[0] - r.p[1];
    }
    r.sqr = mindist;
  }
  return r;
}

static proj project_tetra_origin(v3 a, v3 b, v3 c, v3 d) {
  static const int nexti[3] = {1, 2, 0};
  proj r = proj_init();
  v3 vt[4] = {a, b, c, d};
  v3 dl[3] = {vsub(a, d), vsub(b, d), vsub(c, d)};
  double vl = triple(dl[0], dl[1], dl[2]);
  int ng = (vl * vdot(a, vcross(vsub(b, c), vsub(a, b)))) <= 0;
  if (ng && fabs(vl) > 0) {
    double mindist = -1;
    for (int i = 0; i < 3; ++i) {
      int j = nexti[i];
      double s = vl * vdot(d, vcross(dl[i], dl[j]));
      if (s > 0) {
        proj rt = project_triangle_origin(vt[i], vt[j], d);
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

/* ------------------------------------------------------------------- GJK */
enum { GJK_VALID = 0, GJK_INSIDE = 1, GJK_FAILED = 2 };
#define NV_GJK 4
#define NV_EPA 64
#define NF_EPA 128

typedef struct {
  v3 d, w;
} svert;

typedef struct {
  int c[4]; /* vertex indices */
  double p[4];
  int rank;
} simplex;

/* where gjk_epa_intersect reports the passes of GJK's and EPA's loops for
 * the batch (per thread; NULL: not wanted) */
static __thread int *t_passes;

typedef struct {
  const mdiff *shape;
  svert vstore[NV_GJK + NV_EPA]; /* 0..3: GJK's store, 4..: EPA's */
  int free_v[4];
  int nfree;
  simplex simplices[2];
  int current;
  v3 ray;
  double tol;
  int status;
  int max_iter;
} gjk_state;

static void gjk_get_support(gjk_state *g, v3 d, svert *sv) {
  sv->d = vnormalized(d);
  sv->w = msupport(g->shape, sv->d);
}
static void gjk_remove(gjk_state *g, simplex *s) { g->free_v[g->nfree++] = s->c[--s->rank]; }
static void gjk_append(gjk_state *g, simplex *s, v3 v) {
  s->p[s->rank] = 0;
  s->c[s->rank] = g->free_v[--g->nfree];
  gjk_get_support(g, v, &g->vstore[s->c[s->rank++]]);
}

static int gjk_evaluate(gjk_state *g, v3 guess) {
  int iterations = 0;
  double alpha = 0;
  v3 lastw[4];
  int clastw = 0;
  for (int i = 0; i < 4; ++i) g->free_v[i] = i;
  g->nfree = 4;
  g->current = 0;
  g->status = GJK_VALID;
  g->simplices[0].rank = 0;
  g->ray = guess;
  gjk_append(g, &g->simplices[0], vsq(g->ray) > 0 ? vneg(g->ray) : V(1, 0, 0));
  g->simplices[0].p[0] = 1;
  g->ray = g->vstore[g->simplices[0].c[0]].w;
  lastw[0] = lastw[1] = lastw[2] = lastw[3] = g->ray;
  do {
    int next = 1 - g->current;
    simplex *cs = &g->simplices[g->current];
    simplex *ns = &g->simplices[next];
    double rl = vnorm(g->ray);
    if (rl < g->tol) {
      g->status = GJK_INSIDE;
      break;
    }
    gjk_append(g, cs, vneg(g->ray));
    v3 w = g->vstore[cs->c[cs->rank - 1]].w;
    int found = 0;
    for (int i = 0; i < 4; ++i)
      if (vsq(vsub(w, lastw[i])) < g->tol) {
        found = 1;
        break;
      }
    if (found) {
      gjk_remove(g, cs);
      break;
    }
    lastw[clastw = (clastw + 1) & 3] = w;
    double omega = vdot(g->ray, w) / rl;
    alpha = (alpha < omega) ? omega : alpha; /* std::max */
    if ((rl - alpha) - g->tol * rl <= 0) {
      gjk_remove(g, cs);
      break;
    }
    proj pr = proj_init();
    switch (cs->rank) {
      case 2:
        pr = project_line_origin(g->vstore[cs->c[0]].w, g->vstore[cs->c[1]].w);
        break;
      case 3:
        pr = project_triangle_origin(g->vstore[cs->c[0]].w, g->vstore[cs->c[1]].w,
                                     g->vstore[cs->c[2]].w);
        break;
      case 4:
        pr = project_tetra_origin(g->vstore[cs->c[0]].w, g->vstore[cs->c[1]].w,
                                  g->vstore[cs->c[2]].w, g->vstore[cs->c[3]].w);
        break;
    }
    if (pr.sqr >= 0) {
      ns->rank = 0;
      g->ray = V(0, 0, 0);
      g->current = next;
      for (int i = 0; i < cs->rank; ++i) {
        if (pr.encode & (1 << i)) {
          ns->c[ns->rank] = cs->c[i];
          ns->p[ns->rank++] = pr.p[i];
          g->ray = vadd(g->ray, vmul(g->vstore[cs->c[i]].w, pr.p[i]));
        } else {
          g->free_v[g->nfree++] = cs->c[i];
        }
      }
      if (pr.encode == 15) g->status = GJK_INSIDE;
    } else {
      gjk_remove(g, cs);
      break;
    }
    g->status = (++iterations < g->max_iter) ? g->status : GJK_FAILED;
  } while (g->status == GJK_VALID);
  if (t_passes) t_passes[0] = iterations;
  return g->status;
}

/* gjk-inl.h:230-290 encloseOrigin */
static int enclose_origin(gjk_state *g, simplex *s) {
  switch (s->rank) {
    case 1:
      for (int i = 0; i < 3; ++i) {
        v3 axis = V(0, 0, 0);
        axis.x[i] = 1;
        gjk_append(g, s, axis);
        if (enclose_origin(g, s)) return 1;
        gjk_remove(g, s);
        gjk_append(g, s, vneg(axis));
        if (enclose_origin(g, s)) return 1;
        gjk_remove(g, s);
      }
      break;
    case 2: {
      v3 d = vsub(g->vstore[s->c[1]].w, g->vstore[s->c[0]].w);
      for (int i = 0; i < 3; ++i) {
        v3 axis = V(0, 0, 0);
        axis.x[i] = 1;
        v3 p = vcross(d, axis);
        if (vsq(p) > 0) {
          gjk_append(g, s, p);
          if (enclose_origin(g, s)) return 1;
          gjk_remove(g, s);
          gjk_append(g, s, vneg(p));
          if (enclose_origin(g, s)) return 1;
          gjk_remove(g, s);
        }
      }
      break;
    }
    case 3: {
      v3 n = vcross(vsub(g->vstore[s->c[1]].w, g->vstore[s->c[0]].w),
                    vsub(g->vstore[s->c[2]].w, g->vstore[s->c[0]].w));
      if (vsq(n) > 0) {
        gjk_append(g, s, n);
        if (enclose_origin(g, s)) return 1;
        gjk_remove(g, s);
        gjk_append(g, s, vneg(n));
        if (enclose_origin(g, s)) return 1;
        gjk_remove(g, s);
      }
      break;
    }
    case 4: {
      v3 w3 = g->vstore[s->c[3]].w;
      if (fabs(triple(vsub(g->vstore[s->c[0]].w, w3), vsub(g->vstore[s->c[1]].w, w3),
                      vsub(g->vstore[s->c[2]].w, w3))) > 0)
        return 1;
      break;
    }
  }
  return 0;
}

/* ------------------------------------------------------------------- EPA */
enum {
  EPA_VALID,
  EPA_TOUCHING,
  EPA_DEGENERATED,
  EPA_NONCONVEX,
  EPA_INVALID_HULL,
  EPA_OUT_OF_FACES,
  EPA_OUT_OF_VERTICES,
  EPA_ACCURACY_REACHED,
  EPA_FALLBACK,
  EPA_FAILED
};

typedef struct {
  v3 n;
  double d;
  int c[3]; /* vertex indices */
  int f[3]; /* adjacent faces */
  int e[3];
  int l[2]; /* list links, -1 = none */
  int pass;
} sface;

typedef struct {
  int root, count;
} slist;

typedef struct {
  gjk_state *g;
  sface fc[NF_EPA];
  slist hull, stock;
  int nextsv;
  int status;
  v3 normal;
  double depth;
  int rank;
  int rc[3];
  double rp[3];
  double tol;
  int max_iter;
} epa_state;

static void list_append(epa_state *E, slist *L, int f) {
  E->fc[f].l[0] = -1;
  E->fc[f].l[1] = L->root;
  if (L->root >= 0) E->fc[L->root].l[0] = f;
  L->root = f;
  ++L->count;
}
static void list_remove(epa_state *E, slist *L, int f) {
  if (E->fc[f].l[1] >= 0) E->fc[E->fc[f].l[1]].l[0] = E->fc[f].l[0];
  if (E->fc[f].l[0] >= 0) E->fc[E->fc[f].l[0]].l[1] = E->fc[f].l[1];
  if (f == L->root) L->root = E->fc[f].l[1];
  --L->count;
}
static void bind(epa_state *E, int fa, int ea, int fb, int eb) {
  E->fc[fa].e[ea] = eb;
  E->fc[fa].f[ea] = fb;
  E->fc[fb].e[eb] = ea;
  E->fc[fb].f[eb] = fa;
}

static int edge_dist(epa_state *E, int f, int a, int b, double *dist) {
  v3 wa = E->g->vstore[a].w, wb = E->g->vstore[b].w;
  v3 ba = vsub(wb, wa);
  v3 n_ab = vcross(ba, E->fc[f].n);
  double a_dot_nab = vdot(wa, n_ab);
  if (a_dot_nab < 0) {
    double a_dot_ba = vdot(wa, ba);
    double b_dot_ba = vdot(wb, ba);
    if (a_dot_ba > 0)
      *dist = vnorm(wa);
    else if (b_dot_ba < 0)
      *dist = vnorm(wb);
    else {
      double a_dot_b = vdot(wa, wb);
      double v = vsq(wa) * vsq(wb) - a_dot_b * a_dot_b;
      *dist = sqrt(v > 0 ? v : 0);
    }
    return 1;
  }
  return 0;
}

static int new_face(epa_state *E, int a, int b, int c, int forced) {
  if (E->stock.root >= 0) {
    int f = E->stock.root;
    list_remove(E, &E->stock, f);
    list_append(E, &E->hull, f);
    sface *F = &E->fc[f];
    F->pass = 0;
    F->c[0] = a;
    F->c[1] = b;
    F->c[2] = c;
    v3 wa = E->g->vstore[a].w, wb = E->g->vstore[b].w, wc = E->g->vstore[c].w;
    F->n = vcross(vsub(wb, wa), vsub(wc, wa));
    double l = vnorm(F->n);
    if (l > E->tol) {
      if (!(edge_dist(E, f, a, b, &F->d) || edge_dist(E, f, b, c, &F->d) ||
            edge_dist(E, f, c, a, &F->d)))
        F->d = vdot(wa, F->n) / l;
      F->n = vdiv(F->n, l);
      if (forced || F->d >= -E->tol) return f;
      E->status = EPA_NONCONVEX;
    } else {
      E->status = EPA_DEGENERATED;
    }
    list_remove(E, &E->hull, f);
    list_append(E, &E->stock, f);
    return -1;
  }
  E->status = E->stock.root >= 0 ? EPA_OUT_OF_VERTICES : EPA_OUT_OF_FACES;
  return -1;
}

static int find_best(epa_state *E) {
  int minf = E->hull.root;
  double mind = E->fc[minf].d * E->fc[minf].d;
  for (int f = E->fc[minf].l[1]; f >= 0; f = E->fc[f].l[1]) {
    double sqd = E->fc[f].d * E->fc[f].d;
    if (sqd < mind) {
      minf = f;
      mind = sqd;
    }
  }
  return minf;
}

typedef struct {
  int cf, ff, nf;
} horizon;

static int epa_expand(epa_state *E, int pass, int w, int f, int e, horizon *h) {
  static const int nexti[] = {1, 2, 0};
  static const int previ[] = {2, 0, 1};
  sface *F = &E->fc[f];
  if (F->pass != pass) {
    const int e1 = nexti[e];
    if (vdot(F->n, E->g->vstore[w].w) - F->d < -E->tol) {
      int nf = new_face(E, F->c[e1], F->c[e], w, 0);
      if (nf >= 0) {
        bind(E, nf, 0, f, e);
        if (h->cf >= 0)
          bind(E, nf, 2, h->cf, 1);
        else
          h->ff = nf;
        h->cf = nf;
        ++h->nf;
        return 1;
      }
    } else {
      const int e2 = previ[e];
      F->pass = pass;
      if (epa_expand(E, pass, w, F->f[e1], F->e[e1], h) &&
          epa_expand(E, pass, w, F->f[e2], F->e[e2], h)) {
        list_remove(E, &E->hull, f);
        list_append(E, &E->stock, f);
        return 1;
      }
    }
  }
  return 0;
}

static int epa_evaluate(epa_state *E, gjk_state *g, v3 guess) {
  E->g = g;
  E->hull.root = -1;
  E->hull.count = 0;
  E->stock.root = -1;
  E->stock.count = 0;
  for (int i = 0; i < NF_EPA; ++i) list_append(E, &E->stock, NF_EPA - i - 1);
  E->status = EPA_FAILED;
  E->normal = V(0, 0, 0);
  E->depth = 0;
  E->nextsv = 0;
  simplex *s = &g->simplices[g->current];
  if (s->rank > 1 && enclose_origin(g, s)) {
    while (E->hull.root >= 0) {
      int f = E->hull.root;
      list_remove(E, &E->hull, f);
      list_append(E, &E->stock, f);
    }
    E->status = EPA_VALID;
    E->nextsv = 0;
    v3 w0 = g->vstore[s->c[0]].w, w1 = g->vstore[s->c[1]].w, w2 = g->vstore[s->c[2]].w,
       w3 = g->vstore[s->c[3]].w;
    if (vdot(vsub(w0, w3), vcross(vsub(w1, w3), vsub(w2, w3))) < 0) {
      int t = s->c[0];
      s->c[0] = s->c[1];
      s->c[1] = t;
      double tp = s->p[0];
      s->p[0] = s->p[1];
      s->p[1] = tp;
    }
    int tet[4];
    tet[0] = new_face(E, s->c[0], s->c[1], s->c[2], 1);
    tet[1] = new_face(E, s->c[1], s->c[0], s->c[3], 1);
    tet[2] = new_face(E, s->c[2], s->c[1], s->c[3], 1);
    tet[3] = new_face(E, s->c[0], s->c[2], s->c[3], 1);
    if (E->hull.count == 4) {
      int best = find_best(E);
      sface outer = E->fc[best];
      int pass = 0;
      bind(E, tet[0], 0, tet[1], 0);
      bind(E, tet[0], 1, tet[2], 0);
      bind(E, tet[0], 2, tet[3], 0);
      bind(E, tet[1], 1, tet[3], 2);
      bind(E, tet[1], 2, tet[2], 1);
      bind(E, tet[2], 2, tet[3], 1);
      E->status = EPA_VALID;
      for (int it = 0; it < E->max_iter; ++it) {
        if (t_passes) t_passes[1] = it + 1;
        if (E->nextsv < NV_EPA) {
          horizon h = {-1, -1, 0};
          int w = NV_GJK + E->nextsv++;
          int valid = 1;
          E->fc[best].pass = ++pass;
          gjk_get_support(g, E->fc[best].n, &g->vstore[w]);
          double wdist = vdot(E->fc[best].n, g->vstore[w].w) - E->fc[best].d;
          if (wdist > E->tol) {
            for (int j = 0; j < 3 && valid; ++j)
              valid &= epa_expand(E, pass, w, E->fc[best].f[j], E->fc[best].e[j], &h);
            if (valid && h.nf >= 3) {
              bind(E, h.ff, 2, h.cf, 1);
              list_remove(E, &E->hull, best);
              list_append(E, &E->stock, best);
              best = find_best(E);
              outer = E->fc[best];
            } else {
              E->status = EPA_INVALID_HULL;
              break;
            }
          } else {
            E->status = EPA_ACCURACY_REACHED;
            break;
          }
        } else {
          E->status = EPA_OUT_OF_VERTICES;
          break;
        }
      }
      v3 projection = vmul(outer.n, outer.d);
      E->normal = outer.n;
      E->depth = outer.d;
      E->rank = 3;
      for (int k = 0; k < 3; ++k) E->rc[k] = outer.c[k];
      v3 a = g->vstore[outer.c[0]].w, b = g->vstore[outer.c[1]].w, c = g->vstore[outer.c[2]].w;
      E->rp[0] = vnorm(vcross(vsub(b, projection), vsub(c, projection)));
      E->rp[1] = vnorm(vcross(vsub(c, projection), vsub(a, projection)));
      E->rp[2] = vnorm(vcross(vsub(a, projection), vsub(b, projection)));
      double sum = E->rp[0] + E->rp[1] + E->rp[2];
      E->rp[0] /= sum;
      E->rp[1] /= sum;
      E->rp[2] /= sum;
      return E->status;
    }
  }
  E->status = EPA_FALLBACK;
  E->normal = vneg(guess);
  double nl = vnorm(E->normal);
  if (nl > 0)
    E->normal = vdiv(E->normal, nl);
  else
    E->normal = V(1, 0, 0);
  E->depth = 0;
  E->rank = 1;
  E->rc[0] = s->c[0];
  E->rp[0] = 1;
  return E->status;
}

/* gjk_solver_indep-inl.h:110-165: GJK, then EPA when the origin is inside */
static int gjk_epa_intersect(const fclo_shape *s1, const xf *tf1, const fclo_shape *s2,
                             const xf *tf2, double tol, fclo_contact *out, int *n) {
  v3 guess = V(1, 0, 0);
  mdiff m;
  m.s0 = s1;
  m.s1 = s2;
  mulTN(tf2->R, tf1->R, m.toshape1);
  m.toshape0 = inv_compose(tf1, tf2);
  gjk_state g;
  g.shape = &m;
  g.tol = tol;
  g.max_iter = 128;
  if (gjk_evaluate(&g, vneg(guess)) != GJK_INSIDE) return 0;
  epa_state E;
  E.tol = tol;
  E.max_iter = 255;
  int st = epa_evaluate(&E, &g, vneg(guess));
  if (st == EPA_FAILED) return 0;
  v3 w0 = V(0, 0, 0);
  for (int i = 0; i < E.rank; ++i)
    w0 = vadd(w0, vmul(support0(&m, g.vstore[E.rc[i]].d), E.rp[i]));
  if (out) {
    v3 p = xf_apply(tf1, vsub(w0, vmul(E.normal, E.depth * 0.5)));
    for (int k = 0; k < 3; ++k) {
      out[0].normal[k] = E.normal.x[k];
      out[0].pos[k] = p.x[k];
    }
    out[0].depth = -E.depth;
    *n = 1;
  }
  return 1;
}

/* ----------------------------------------------------------- closed forms */
static void put(fclo_contact *c, v3 normal, v3 pos, double depth) {
  for (int k = 0; k < 3; ++k) {
    c->normal[k] = normal.x[k];
    c->pos[k] = pos.x[k];
  }
  c->depth = depth;
}

static int sphere_sphere(const fclo_shape *s1, const xf *tf1, const fclo_shape *s2,
                         const xf *tf2, fclo_contact *out, int *n) {
  v3 diff = vsub(tf2->t, tf1->t);
  double len = vnorm(diff);
  double r1 = s1->dims[0], r2 = s2->dims[0];
  if (len > r1 + r2) return 0;
  if (out) {
    v3 normal = len > 0 ? vdiv(diff, len) : diff;
    v3 point = vadd(tf1->t, vdiv(vmul(diff, r1), r1 + r2));
    put(out, normal, point, r1 + r2 - len);
    *n = 1;
  }
  return 1;
}

static int sphere_box(const fclo_shape *sp, const xf *X_FS, const fclo_shape *bx,
                      const xf *X_FB, fclo_contact *out, int *n) {
  const double r = sp->dims[0];
  xf X_BS = inv_compose(X_FB, X_FS);
  v3 p_BC = X_BS.t, p_BN;
  int clamped = 0;
  for (int i = 0; i < 3; ++i) {
    double half = bx->dims[i] / 2;
    p_BN.x[i] = p_BC.x[i];
    if (p_BC.x[i] < -half) {
      clamped = 1;
      p_BN.x[i] = -half;
    }
    if (p_BC.x[i] > half) {
      clamped = 1;
      p_BN.x[i] = half;
    }
  }
  v3 p_CN = vsub(p_BN, p_BC);
  double sq = vsq(p_CN);
  if (sq > r * r) return 0;
  if (out) {
    const double eps = 16 * DBL_EPSILON;
    double depth;
    v3 nrm, p_BP;
    if (clamped && sq > eps * eps) {
      double dist = sqrt(sq);
      nrm = vdiv(p_CN, dist);
      depth = r - dist;
      p_BP = vadd(p_BN, vmul(nrm, depth * 0.5));
    } else {
      double min_distance = INFINITY;
      int min_axis = -1;
      for (int i = 0; i < 3; ++i) {
        double half = bx->dims[i] / 2;
        double dist = p_BC.x[i] >= 0 ? half - p_BC.x[i] : p_BC.x[i] + half;
        if (dist + eps < min_distance) {
          min_distance = dist;
          min_axis = i;
        }
      }
      nrm = V(0, 0, 0);
      nrm.x[min_axis] = p_BC.x[min_axis] >= 0 ? -1 : 1;
      depth = min_distance + r;
      p_BP = vadd(p_BC, vmul(nrm, (r - min_distance) / 2));
    }
    put(out, rot(X_FB->R, nrm), xf_apply(X_FB, p_BP), depth);
    *n = 1;
  }
  return 1;
}

static int sphere_capsule(const fclo_shape *sp, const xf *tf1, const fclo_shape *cap,
                          const xf *tf2, fclo_contact *out, int *n) {
  v3 pos1 = V(0., 0., 0.5 * cap->dims[1]);
  v3 pos2 = V(0., 0., -0.5 * cap->dims[1]);
  /* tf2.inverse() * tf1.translation() */
  v3 s_c = vadd(rotT(tf2->R, tf1->t), vneg(rotT(tf2->R, tf2->t)));
  v3 v = vsub(pos2, pos1), w = vsub(s_c, pos1), seg;
  double c1 = vdot(w, v), c2 = vdot(v, v);
  if (c1 <= 0)
    seg = pos1;
  else if (c2 <= c1)
    seg = pos2;
  else
    seg = vadd(pos1, vmul(v, c1 / c2));
  v3 diff = vsub(s_c, seg);
  double distance = vnorm(diff) - sp->dims[0] - cap->dims[0];
  if (distance > 0) return 0;
  v3 ln = vneg(vnormalized(diff));
  if (out) {
    put(out, rot(tf2->R, ln), xf_apply(tf2, vadd(seg, vmul(ln, distance))), -distance);
    *n = 1;
  }
  return 1;
}

static int sphere_cylinder(const fclo_shape *sp, const xf *X_FS, const fclo_shape *cy,
                           const xf *X_FC, fclo_contact *out, int *n) {
  const double r_s = sp->dims[0], radius = cy->dims[0], height = cy->dims[1];
  xf X_CS = inv_compose(X_FC, X_FS);
  v3 p_CS = X_CS.t, p_CN = p_CS;
  int outside = 0;
  const double half_h = height / 2;
  if (p_CS.x[2] > half_h) {
    outside = 1;
    p_CN.x[2] = half_h;
  } else if (p_CS.x[2] < -half_h) {
    outside = 1;
    p_CN.x[2] = -half_h;
  }
  double rx = p_CS.x[0], ry = p_CS.x[1];
  double sqd = rx * rx + ry * ry;
  if (sqd > radius * radius) {
    outside = 1;
    double k = radius / sqrt(sqd);
    rx *= k;
    ry *= k;
    p_CN.x[0] = rx;
    p_CN.x[1] = ry;
  }
  v3 p_SN = vsub(p_CN, p_CS);
  double sq = vsq(p_SN);
  if (sq > r_s * r_s) return 0;
  if (out) {
    const double eps = 16 * DBL_EPSILON;
    double depth;
    v3 nrm, p_CP;
    if (outside && sq > eps * eps) {
      double d = sqrt(sq);
      nrm = vdiv(p_SN, d);
      depth = r_s - d;
      p_CP = vadd(p_CN, vmul(nrm, depth * 0.5));
    } else {
      double face = p_CS.x[2] >= 0 ? height / 2 - p_CS.x[2] : p_CS.x[2] + height / 2;
      double nx = p_CS.x[0], ny = p_CS.x[1];
      double d_xy = sqrt(nx * nx + ny * ny);
      double barrel = radius - d_xy;
      if (barrel < face - eps) {
        if (d_xy > eps) {
          nrm = V(-nx / d_xy, -ny / d_xy, 0);
          depth = r_s + barrel;
          p_CP = vadd(p_CS, vmul(nrm, (r_s - barrel) / 2));
        } else {
          nrm = V(-1, 0, 0);
          depth = r_s + radius;
          p_CP = vadd(p_CS, vmul(nrm, (r_s - barrel) / 2));
        }
      } else {
        nrm = V(0, 0, p_CS.x[2] >= 0 ? -1 : 1);
        depth = face + r_s;
        p_CP = vadd(p_CS, vmul(nrm, (r_s - face) / 2));
      }
    }
    put(out, rot(X_FC->R, nrm), xf_apply(X_FC, p_CP), depth);
    *n = 1;
  }
  return 1;
}

/* box_box-inl.h:82-107 */
static void line_closest(v3 pa, v3 ua, v3 pb, v3 ub, double *alpha, double *beta) {
  v3 p = vsub(pb, pa);
  double uaub = vdot(ua, ub);
  double q1 = vdot(ua, p);
  double q2 = -vdot(ub, p);
  double d = 1 - uaub * uaub;
  if (d <= (double)(0.0001f)) {
    *alpha = 0;
    *beta = 0;
  } else {
    d = 1 / d;
    *alpha = (q1 + uaub * q2) * d;
    *beta = (uaub * q1 + q2) * d;
  }
}

/* box_box-inl.h:110-170 */
static int intersect_rect_quad(double h[2], double p[8], double ret[16]) {
  int nq = 4, nr = 0;
  double buffer[16];
  double *q = p;
  double *r = ret;
  for (int dir = 0; dir <= 1; ++dir) {
    for (int sign = -1; sign <= 1; sign += 2) {
      double *pq = q;
      double *pr = r;
      nr = 0;
      for (int i = nq; i > 0; --i) {
        if (sign * pq[dir] < h[dir]) {
          pr[0] = pq[0];
          pr[1] = pq[1];
          pr += 2;
          nr++;
          if (nr & 8) {
            q = r;
            goto done;
          }
        }
        double *nextq = (i > 1) ? pq + 2 : q;
        if ((sign * pq[dir] < h[dir]) ^ (sign * nextq[dir] < h[dir])) {
          pr[1 - dir] = pq[1 - dir] +
                        (nextq[1 - dir] - pq[1 - dir]) / (nextq[dir] - pq[dir]) *
                            (sign * h[dir] - pq[dir]);
          pr[dir] = sign * h[dir];
          pr += 2;
          nr++;
          if (nr & 8) {
            q = r;
            goto done;
          }
        }
        pq += 2;
      }
      q = r;
      r = (q == ret) ? buffer : ret;
      nq = nr;
    }
  }
done:
  if (q != ret) memcpy(ret, q, nr * 2 * sizeof(double));
  return nr;
}

/* box_box-inl.h:173-250 */
static void cull_points(int n, double p[], int m, int i0, int iret[]) {
  double a, cx, cy, q;
  switch (n) {
    case 1:
      cx = p[0];
      cy = p[1];
      break;
    case 2:
      cx = 0.5 * (p[0] + p[2]);
      cy = 0.5 * (p[1] + p[3]);
      break;
    default:
      a = 0;
      cx = 0;
      cy = 0;
      for (int i = 0; i < n - 1; ++i) {
        q = p[i * 2] * p[i * 2 + 3] - p[i * 2 + 2] * p[i * 2 + 1];
        a += q;
        cx += q * (p[i * 2] + p[i * 2 + 2]);
        cy += q * (p[i * 2 + 1] + p[i * 2 + 3]);
      }
      q = p[n * 2 - 2] * p[1] - p[0] * p[n * 2 - 1];
      if (fabs(a + q) > DBL_EPSILON)
        a = 1 / (3 * (a + q));
      else
        a = 1e18f;
      cx = a * (cx + q * (p[n * 2 - 2] + p[0]));
      cy = a * (cy + q * (p[n * 2 - 1] + p[1]));
  }
  double A[8];
  for (int i = 0; i < n; ++i) A[i] = atan2(p[i * 2 + 1] - cy, p[i * 2] - cx);
  int avail[8];
  for (int i = 0; i < n; ++i) avail[i] = 1;
  avail[i0] = 0;
  iret[0] = i0;
  iret++;
  const double pi = M_PI;
  for (int j = 1; j < m; ++j) {
    a = j * (2 * pi / m) + A[i0];
    if (a > pi) a -= 2 * pi;
    double maxdiff = 1e9, diff;
    *iret = i0;
    for (int i = 0; i < n; ++i) {
      if (avail[i]) {
        diff = fabs(A[i] - a);
        if (diff > pi) diff = 2 * pi - diff;
        if (diff < maxdiff) {
          maxdiff = diff;
          *iret = i;
        }
      }
    }
    avail[*iret] = 0;
    iret++;
  }
}

static double R_(const double M[9], int i, int j) { return M[3 * i + j]; }

/* box_box-inl.h:253-848 (boxBox2, world-frame) and 851-862 */
static int box_box(const fclo_shape *b1, const xf *tf1, const fclo_shape *b2, const xf *tf2,
                   fclo_contact *out, int *n) {
  const double fudge = 1.05;
  const double *R1 = tf1->R, *R2 = tf2->R;
  v3 T1 = tf1->t, T2 = tf2->t;
  v3 normalC = V(0, 0, 0);
  double s, s2, l;
  int invert_normal, code;
  v3 p = vsub(T2, T1);
  v3 pp = rotT(R1, p);
  v3 A = vmul(V(b1->dims[0], b1->dims[1], b1->dims[2]), 0.5);
  v3 B = vmul(V(b2->dims[0], b2->dims[1], b2->dims[2]), 0.5);
  double R[9], Q[9];
  mulTN(R1, R2, R);
  for (int i = 0; i < 9; ++i) Q[i] = fabs(R[i]);
  int best_col = -1;
  const double *normalR = 0;
  double tmp;
  s = -DBL_MAX;
  invert_normal = 0;
  code = 0;
  /* face axes of box 1 */
  for (int k = 0; k < 3; ++k) {
    tmp = pp.x[k];
    s2 = fabs(tmp) - (vdot(V(Q[3 * k], Q[3 * k + 1], Q[3 * k + 2]), B) + A.x[k]);
    if (s2 > 0) return 0;
    if (s2 > s) {
      s = s2;
      best_col = k;
      normalR = R1;
      invert_normal = tmp < 0;
      code = 1 + k;
    }
  }
  /* face axes of box 2 */
  for (int k = 0; k < 3; ++k) {
    tmp = vdot(col(R2, k), p);
    s2 = fabs(tmp) - (vdot(V(Q[k], Q[3 + k], Q[6 + k]), A) + B.x[k]);
    if (s2 > 0) return 0;
    if (s2 > s) {
      s = s2;
      best_col = k;
      normalR = R2;
      invert_normal = tmp < 0;
      code = 4 + k;
    }
  }
  const double eps = DBL_EPSILON;
  {
    double ma = A.x[0] > A.x[1] ? (A.x[0] > A.x[2] ? A.x[0] : A.x[2]) : (A.x[1] > A.x[2] ? A.x[1] : A.x[2]);
    double mb = B.x[0] > B.x[1] ? (B.x[0] > B.x[2] ? B.x[0] : B.x[2]) : (B.x[1] > B.x[2] ? B.x[1] : B.x[2]);
    double m = ma > mb ? ma : mb;
    if (m < 1.0) m = 1.0;
    double sf = m * 10.0 * eps;
    for (int i = 0; i < 9; ++i) Q[i] += sf;
  }
  v3 nv;
  /* edge-edge axes, in the reference's order */
#define EDGE_AXIS(TMP, RAD, NX, NY, NZ, CODE)                                   \
  tmp = (TMP);                                                                  \
  s2 = fabs(tmp) - (RAD);                                                       \
  if (s2 > 0) return 0;                                                         \
  nv = V(NX, NY, NZ);                                                           \
  l = vnorm(nv);                                                                \
  if (l > eps) {                                                                \
    s2 /= l;                                                                    \
    if (s2 * fudge > s) {                                                       \
      s = s2;                                                                   \
      best_col = -1;                                                            \
      normalC = vdiv(nv, l);                                                    \
      invert_normal = tmp < 0;                                                  \
      code = CODE;                                                              \
    }                                                                           \
  }
  const double *a = A.x, *b = B.x;
  EDGE_AXIS(pp.x[2] * R_(R, 1, 0) - pp.x[1] * R_(R, 2, 0),
            ((a[1] * R_(Q, 2, 0) + a[2] * R_(Q, 1, 0)) + b[1] * R_(Q, 0, 2)) + b[2] * R_(Q, 0, 1),
            0, -R_(R, 2, 0), R_(R, 1, 0), 7)
  EDGE_AXIS(pp.x[2] * R_(R, 1, 1) - pp.x[1] * R_(R, 2, 1),
            ((a[1] * R_(Q, 2, 1) + a[2] * R_(Q, 1, 1)) + b[0] * R_(Q, 0, 2)) + b[2] * R_(Q, 0, 0),
            0, -R_(R, 2, 1), R_(R, 1, 1), 8)
  EDGE_AXIS(pp.x[2] * R_(R, 1, 2) - pp.x[1] * R_(R, 2, 2),
            ((a[1] * R_(Q, 2, 2) + a[2] * R_(Q, 1, 2)) + b[0] * R_(Q, 0, 1)) + b[1] * R_(Q, 0, 0),
            0, -R_(R, 2, 2), R_(R, 1, 2), 9)
  EDGE_AXIS(pp.x[0] * R_(R, 2, 0) - pp.x[2] * R_(R, 0, 0),
            ((a[0] * R_(Q, 2, 0) + a[2] * R_(Q, 0, 0)) + b[1] * R_(Q, 1, 2)) + b[2] * R_(Q, 1, 1),
            R_(R, 2, 0), 0, -R_(R, 0, 0), 10)
  EDGE_AXIS(pp.x[0] * R_(R, 2, 1) - pp.x[2] * R_(R, 0, 1),
            ((a[0] * R_(Q, 2, 1) + a[2] * R_(Q, 0, 1)) + b[0] * R_(Q, 1, 2)) + b[2] * R_(Q, 1, 0),
            R_(R, 2, 1), 0, -R_(R, 0, 1), 11)
  EDGE_AXIS(pp.x[0] * R_(R, 2, 2) - pp.x[2] * R_(R, 0, 2),
            ((a[0] * R_(Q, 2, 2) + a[2] * R_(Q, 0, 2)) + b[0] * R_(Q, 1, 1)) + b[1] * R_(Q, 1, 0),
            R_(R, 2, 2), 0, -R_(R, 0, 2), 12)
  EDGE_AXIS(pp.x[1] * R_(R, 0, 0) - pp.x[0] * R_(R, 1, 0),
            ((a[0] * R_(Q, 1, 0) + a[1] * R_(Q, 0, 0)) + b[1] * R_(Q, 2, 2)) + b[2] * R_(Q, 2, 1),
            -R_(R, 1, 0), R_(R, 0, 0), 0, 13)
  EDGE_AXIS(pp.x[1] * R_(R, 0, 1) - pp.x[0] * R_(R, 1, 1),
            ((a[0] * R_(Q, 1, 1) + a[1] * R_(Q, 0, 1)) + b[0] * R_(Q, 2, 2)) + b[2] * R_(Q, 2, 0),
            -R_(R, 1, 1), R_(R, 0, 1), 0, 14)
  EDGE_AXIS(pp.x[1] * R_(R, 0, 2) - pp.x[0] * R_(R, 1, 2),
            ((a[0] * R_(Q, 1, 2) + a[1] * R_(Q, 0, 2)) + b[0] * R_(Q, 2, 1)) + b[1] * R_(Q, 2, 0),
            -R_(R, 1, 2), R_(R, 0, 2), 0, 15)
#undef EDGE_AXIS
  if (!code) return 0;
  v3 normal = best_col != -1 ? col(normalR, best_col) : rot(R1, normalC);
  if (invert_normal) normal = vneg(normal);
  const double depth = -s;
  *n = 0;
  if (code > 6) {
    v3 pa = T1;
    for (int j = 0; j < 3; ++j) {
      double sign = vdot(col(R1, j), normal) > 0 ? 1 : -1;
      pa = vadd(pa, vmul(col(R1, j), A.x[j] * sign));
    }
    v3 pb = T2;
    for (int j = 0; j < 3; ++j) {
      double sign = vdot(col(R2, j), normal) > 0 ? -1 : 1;
      pb = vadd(pb, vmul(col(R2, j), B.x[j] * sign));
    }
    double alpha, beta;
    v3 ua = col(R1, (code - 7) / 3), ub = col(R2, (code - 7) % 3);
    line_closest(pa, ua, pb, ub, &alpha, &beta);
    pa = vadd(pa, vmul(ua, alpha));
    pb = vadd(pb, vmul(ub, beta));
    if (out) {
      put(&out[0], normal, vmul(vadd(pa, pb), 0.5), depth);
      *n = 1;
    }
    return 1;
  }
  const double *Ra, *Rb;
  v3 Pa, Pb, Sa, Sb;
  if (code <= 3) {
    Ra = R1;
    Rb = R2;
    Pa = T1;
    Pb = T2;
    Sa = A;
    Sb = B;
  } else {
    Ra = R2;
    Rb = R1;
    Pa = T2;
    Pb = T1;
    Sa = B;
    Sb = A;
  }
  v3 normal2 = code <= 3 ? normal : vneg(normal);
  v3 nr = rotT(Rb, normal2);
  v3 anr = V(fabs(nr.x[0]), fabs(nr.x[1]), fabs(nr.x[2]));
  int lanr, a1, a2;
  if (anr.x[1] > anr.x[0]) {
    if (anr.x[1] > anr.x[2]) {
      a1 = 0;
      lanr = 1;
      a2 = 2;
    } else {
      a1 = 0;
      a2 = 1;
      lanr = 2;
    }
  } else {
    if (anr.x[0] > anr.x[2]) {
      lanr = 0;
      a1 = 1;
      a2 = 2;
    } else {
      a1 = 0;
      a2 = 1;
      lanr = 2;
    }
  }
  v3 center;
  if (nr.x[lanr] < 0)
    center = vadd(vsub(Pb, Pa), vmul(col(Rb, lanr), Sb.x[lanr]));
  else
    center = vsub(vsub(Pb, Pa), vmul(col(Rb, lanr), Sb.x[lanr]));
  int codeN, code1, code2;
  codeN = code <= 3 ? code - 1 : code - 4;
  if (codeN == 0) {
    code1 = 1;
    code2 = 2;
  } else if (codeN == 1) {
    code1 = 0;
    code2 = 2;
  } else {
    code1 = 0;
    code2 = 1;
  }
  double quad[8];
  double c1 = vdot(col(Ra, code1), center), c2 = vdot(col(Ra, code2), center);
  double m11 = vdot(col(Rb, a1), col(Ra, code1)), m12 = vdot(col(Rb, a2), col(Ra, code1));
  double m21 = vdot(col(Rb, a1), col(Ra, code2)), m22 = vdot(col(Rb, a2), col(Ra, code2));
  double k1 = m11 * Sb.x[a1], k2 = m21 * Sb.x[a1], k3 = m12 * Sb.x[a2], k4 = m22 * Sb.x[a2];
  quad[0] = c1 - k1 - k3;
  quad[1] = c2 - k2 - k4;
  quad[2] = c1 - k1 + k3;
  quad[3] = c2 - k2 + k4;
  quad[4] = c1 + k1 + k3;
  quad[5] = c2 + k2 + k4;
  quad[6] = c1 + k1 - k3;
  quad[7] = c2 + k2 - k4;
  double rect[2] = {Sa.x[code1], Sa.x[code2]};
  double ret[16];
  int n_intersect = intersect_rect_quad(rect, quad, ret);
  if (n_intersect < 1) return 1; /* return_code = code: intersecting, no contact */
  v3 points[8];
  double dep[8];
  double det1 = 1.f / (m11 * m22 - m12 * m21);
  m11 *= det1;
  m12 *= det1;
  m21 *= det1;
  m22 *= det1;
  int cnum = 0;
  for (int j = 0; j < n_intersect; ++j) {
    double kk1 = m22 * (ret[j * 2] - c1) - m12 * (ret[j * 2 + 1] - c2);
    double kk2 = -m21 * (ret[j * 2] - c1) + m11 * (ret[j * 2 + 1] - c2);
    points[cnum] = vadd(vadd(center, vmul(col(Rb, a1), kk1)), vmul(col(Rb, a2), kk2));
    dep[cnum] = Sa.x[codeN] - vdot(normal2, points[cnum]);
    if (dep[cnum] >= 0) {
      ret[cnum * 2] = ret[j * 2];
      ret[cnum * 2 + 1] = ret[j * 2 + 1];
      cnum++;
    }
  }
  if (cnum < 1) return 1;
  int maxc = 4;
  if (maxc > cnum) maxc = cnum;
  if (maxc < 1) maxc = 1;
  int iret[8] = {0, 1, 2, 3, 4, 5, 6, 7};
  if (cnum > maxc) {
    int i1 = 0;
    double maxdepth = dep[0];
    for (int i = 1; i < cnum; ++i)
      if (dep[i] > maxdepth) {
        maxdepth = dep[i];
        i1 = i;
      }
    cull_points(cnum, ret, maxc, i1, iret);
    cnum = maxc;
  }
  if (out) {
    for (int j = 0; j < cnum; ++j) {
      int i = iret[j];
      v3 pw = code < 4 ? vadd(vadd(points[i], Pa), vmul(normal, dep[i] / 2))
                       : vsub(vadd(points[i], Pa), vmul(normal, dep[i] / 2));
      put(&out[j], normal, pw, dep[i]);
    }
    *n = cnum;
  }
  return 1;
}

/* ---------------------------------------------------------------- dispatch */
static void flip_normals(fclo_contact *c, int n) {
  for (int i = 0; i < n; ++i)
    for (int k = 0; k < 3; ++k) c[i].normal[k] = -c[i].normal[k];
}

int fclo_shape_intersect(const fclo_shape *s1, const double tf1[12], const fclo_shape *s2,
                         const double tf2[12], double tol, fclo_contact *out, int *n) {
  xf a = xf_from(tf1), b = xf_from(tf2);
  int t1 = s1->type, t2 = s2->type, r = 0, k = 0;
  *n = 0;
  if (t1 == FCLO_SHAPE_SPHERE && t2 == FCLO_SHAPE_SPHERE)
    r = sphere_sphere(s1, &a, s2, &b, out, &k);
  else if (t1 == FCLO_SHAPE_BOX && t2 == FCLO_SHAPE_BOX)
    r = box_box(s1, &a, s2, &b, out, &k);
  else if (t1 == FCLO_SHAPE_SPHERE && t2 == FCLO_SHAPE_CAPSULE)
    r = sphere_capsule(s1, &a, s2, &b, out, &k);
  else if (t1 == FCLO_SHAPE_CAPSULE && t2 == FCLO_SHAPE_SPHERE) {
    r = sphere_capsule(s2, &b, s1, &a, out, &k);
    if (out) flip_normals(out, k);
  } else if (t1 == FCLO_SHAPE_SPHERE && t2 == FCLO_SHAPE_BOX)
    r = sphere_box(s1, &a, s2, &b, out, &k);
  else if (t1 == FCLO_SHAPE_BOX && t2 == FCLO_SHAPE_SPHERE) {
    r = sphere_box(s2, &b, s1, &a, out, &k);
    if (out) flip_normals(out, k);
  } else if (t1 == FCLO_SHAPE_SPHERE && t2 == FCLO_SHAPE_CYLINDER)
    r = sphere_cylinder(s1, &a, s2, &b, out, &k);
  else if (t1 == FCLO_SHAPE_CYLINDER && t2 == FCLO_SHAPE_SPHERE) {
    r = sphere_cylinder(s2, &b, s1, &a, out, &k);
    if (out) flip_normals(out, k);
  } else
    r = gjk_epa_intersect(s1, &a, s2, &b, tol, out, &k);
  *n = out ? k : 0;
  return r;
}

/* shape_collision_traversal_node-inl.h:69-134 with enable_cost off: cap =
 * num_max_contacts - contacts already in the result.  With enable_contact
 * every contact of the pair is a candidate, deepest first when they do not
 * all fit; otherwise one contact without geometry.  Returns the number of
 * contacts added; out receives them. */
int fclo_shape_collide(const fclo_shape *s1, const double tf1[12], const fclo_shape *s2,
                       const double tf2[12], int cap, int enable_contact, double tol,
                       fclo_contact *out) {
  if (cap <= 0) return 0;
  if (enable_contact) {
    fclo_contact c[8];
    int n = 0;
    if (!fclo_shape_intersect(s1, tf1, s2, tf2, tol, c, &n)) return 0;
    int take = n;
    if (cap < n) {
      /* partial_sort by depth, deepest first (ties: the order chosen here is
       * the first-seen deepest; the reference's partial_sort leaves ties in
       * an unspecified order) */
      for (int i = 0; i < cap; ++i) {
        int best = i;
        for (int j = i + 1; j < n; ++j)
          if (c[j].depth > c[best].depth) best = j;
        fclo_contact t = c[i];
        c[i] = c[best];
        c[best] = t;
      }
      take = cap;
    }
    for (int i = 0; i < take; ++i) out[i] = c[i];
    return take;
  }
  int n = 0;
  if (!fclo_shape_intersect(s1, tf1, s2, tf2, tol, 0, &n)) return 0;
  memset(&out[0], 0, sizeof(out[0]));
  return 1;
}

void fclo_shape_collide_batch(const fclo_shape *shapes, int num_queries, const int *qs,
                              const double *tf1s, const double *tf2s, int cap,
                              int enable_contact, double tol, int *counts, fclo_contact *out,
                              int threads, int *passes) {
#pragma omp parallel for schedule(dynamic, 256) num_threads(threads)
  for (int q = 0; q < num_queries; ++q) {
    fclo_contact c[8];
    int pp[2] = {0, 0};
    t_passes = passes ? pp : 0;
    int n = fclo_shape_collide(shapes + qs[2 * q], tf1s + 12 * (size_t)q, shapes + qs[2 * q + 1],
                               tf2s + 12 * (size_t)q, cap, enable_contact, tol, c);
    t_passes = 0;
    if (passes) {
      passes[2 * q] = pp[0];
      passes[2 * q + 1] = pp[1];
    }
    counts[q] = n;
    if (out)
      for (int i = 0; i < n && i < cap; ++i) out[(size_t)q * cap + i] = c[i];
  }
}

/* The batch over a table of 40-byte shape records (type, num_vertices,
 * dims[3], vertex_offset -- the layout of the engine's shape records) and
 * a vertex pool, so that large batches need no per-shape conversion. */
void fclo_shape_collide_records(const fclo_shape_rec *recs, const double *pool,
                                int num_queries, const int *qs, const double *tf1s,
                                const double *tf2s, int cap, int enable_contact, double tol,
                                int *counts, fclo_contact *out, int threads, int *passes) {
#pragma omp parallel for schedule(dynamic, 256) num_threads(threads)
  for (int q = 0; q < num_queries; ++q) {
    fclo_shape s[2];
    for (int k = 0; k < 2; ++k) {
      const fclo_shape_rec *r = recs + qs[2 * q + k];
      s[k].type = r->type;
      s[k].num_vertices = r->num_vertices;
      for (int i = 0; i < 3; ++i) s[k].dims[i] = r->dims[i];
      s[k].vertices = r->type == FCLO_SHAPE_CONVEX ? pool + 3 * r->vertex_offset : 0;
    }
    fclo_contact c[8];
    int pp[2] = {0, 0};
    t_passes = passes ? pp : 0;
    int n = fclo_shape_collide(&s[0], tf1s + 12 * (size_t)q, &s[1], tf2s + 12 * (size_t)q, cap,
                               enable_contact, tol, c);
    t_passes = 0;
    if (passes) {
      passes[2 * q] = pp[0];
      passes[2 * q + 1] = pp[1];
    }
    counts[q] = n;
    if (out)
      for (int i = 0; i < n && i < cap; ++i) out[(size_t)q * cap + i] = c[i];
  }
}

/* ============================================ mesh vs shape (BVHModel<OBBRSS>) */

/* getBoundVertices in the shape's frame (box-inl.h:122-138,
 * sphere-inl.h:96-120, capsule-inl.h:108-162, cylinder-inl.h:102-126,
 * convex-inl.h:252-262): the points whose fit bounds the shape.  Returns
 * their count (<= 36 for the closed shapes). */
static int bound_points(const fclo_shape *s, v3 *p) {
  switch (s->type) {
    case FCLO_SHAPE_BOX: {
      double a = s->dims[0] / 2, b = s->dims[1] / 2, c = s->dims[2] / 2;
      int k = 0;
      for (int i = 0; i < 2; ++i)
        for (int j = 0; j < 2; ++j)
          for (int l = 0; l < 2; ++l)
            p[k++] = V(i ? -a : a, j ? -b : b, l ? -c : c);
      return 8;
    }
    case FCLO_SHAPE_SPHERE: {
      double m = (1 + sqrt(5.0)) / 2.0;
      double e = s->dims[0] * 6 / (sqrt(27.0) + sqrt(15.0));
      double a = e, b = m * e;
      p[0] = V(0, a, b);   p[1] = V(0, -a, b);  p[2] = V(0, a, -b);  p[3] = V(0, -a, -b);
      p[4] = V(a, b, 0);   p[5] = V(-a, b, 0);  p[6] = V(a, -b, 0);  p[7] = V(-a, -b, 0);
      p[8] = V(b, 0, a);   p[9] = V(b, 0, -a);  p[10] = V(-b, 0, a); p[11] = V(-b, 0, -a);
      return 12;
    }
    case FCLO_SHAPE_CAPSULE: {
      double r = s->dims[0], hl = s->dims[1] * 0.5;
      double m = (1 + sqrt(5.0)) / 2.0;
      double e = r * 6 / (sqrt(27.0) + sqrt(15.0));
      double a = e, b = m * e, r2 = r * 2 / sqrt(3.0);
      for (int h = 0; h < 2; ++h) {
        double z = h ? -hl : hl;
        v3 *q = p + 12 * h;
        q[0] = V(0, a, b + z);   q[1] = V(0, -a, b + z);
        q[2] = V(0, a, -b + z);  q[3] = V(0, -a, -b + z);
        q[4] = V(a, b, z);       q[5] = V(-a, b, z);
        q[6] = V(a, -b, z);      q[7] = V(-a, -b, z);
        q[8] = V(b, 0, a + z);   q[9] = V(b, 0, -a + z);
        q[10] = V(-b, 0, a + z); q[11] = V(-b, 0, -a + z);
      }
      /* the capsule writes b - hl and -b - hl: x + (-hl) == x - hl exactly */
      double c = 0.5 * r2, d = r;
      for (int h = 0; h < 2; ++h) {
        double z = h ? -hl : hl;
        v3 *q = p + 24 + 6 * h;
        q[0] = V(r2, 0, z);  q[1] = V(c, d, z);   q[2] = V(-c, d, z);
        q[3] = V(-r2, 0, z); q[4] = V(-c, -d, z); q[5] = V(c, -d, z);
      }
      return 36;
    }
    case FCLO_SHAPE_CYLINDER: {
      double r = s->dims[0], hl = s->dims[1] * 0.5;
      double r2 = r * 2 / sqrt(3.0), a = 0.5 * r2, b = r;
      for (int h = 0; h < 2; ++h) {
        double z = h ? hl : -hl;
        v3 *q = p + 6 * h;
        q[0] = V(r2, 0, z);  q[1] = V(a, b, z);   q[2] = V(-a, b, z);
        q[3] = V(-r2, 0, z); q[4] = V(-a, -b, z); q[5] = V(a, -b, z);
      }
      return 12;
    }
  }
  return 0;
}

/* The OBB part of computeBV<OBBRSS>(shape, tf) (utility-inl.h:124-133: the
 * generic path, there being no OBBRSS specialisation for these shapes):
 * OBB fitn (BV_fitter / bv/utility-inl.h:131-145) over the posed bound
 * vertices -- point covariance (geometry-inl.h:1335-1421, point branch),
 * eigen_old, axisFromEigen, getExtentAndCenter_pointcloud
 * (geometry-inl.h:228-287).  Only the OBB enters the BV test. */
void fclo_shape_obb(const fclo_shape *s, const double tf[12], double axis[9], double to[3],
                    double extent[3]) {
  v3 local[36];
  xf T = xf_from(tf);
  int n = s->type == FCLO_SHAPE_CONVEX ? s->num_vertices : bound_points(s, local);
  double S1[3] = {0, 0, 0};
  double s00 = 0, s11 = 0, s22 = 0, s01 = 0, s02 = 0, s12 = 0;
  for (int i = 0; i < n; ++i) {
    v3 q = s->type == FCLO_SHAPE_CONVEX
               ? V(s->vertices[3 * i], s->vertices[3 * i + 1], s->vertices[3 * i + 2])
               : local[i];
    v3 p = xf_apply(&T, q);
    for (int k = 0; k < 3; ++k) S1[k] += p.x[k];
    s00 += p.x[0] * p.x[0];
    s11 += p.x[1] * p.x[1];
    s22 += p.x[2] * p.x[2];
    s01 += p.x[0] * p.x[1];
    s02 += p.x[0] * p.x[2];
    s12 += p.x[1] * p.x[2];
  }
  double M[9], E[9], ev[3], np = (double)n;
  M[0] = s00 - S1[0] * S1[0] / np;
  M[4] = s11 - S1[1] * S1[1] / np;
  M[8] = s22 - S1[2] * S1[2] / np;
  M[1] = s01 - S1[0] * S1[1] / np;
  M[5] = s12 - S1[1] * S1[2] / np;
  M[2] = s02 - S1[0] * S1[2] / np;
  M[3] = M[1];
  M[6] = M[2];
  M[7] = M[5];
  fclo_eigen_old(M, ev, E);
  fclo_axis_from_eigen(E, ev, axis);
  double mn[3] = {DBL_MAX, DBL_MAX, DBL_MAX}, mx[3] = {-DBL_MAX, -DBL_MAX, -DBL_MAX};
  for (int i = 0; i < n; ++i) {
    v3 q = s->type == FCLO_SHAPE_CONVEX
               ? V(s->vertices[3 * i], s->vertices[3 * i + 1], s->vertices[3 * i + 2])
               : local[i];
    v3 p = xf_apply(&T, q);
    for (int k = 0; k < 3; ++k) {
      double pr = vdot(col(axis, k), p);
      if (pr > mx[k]) mx[k] = pr;
      if (pr < mn[k]) mn[k] = pr;
    }
  }
  v3 o = V((mx[0] + mn[0]) / 2, (mx[1] + mn[1]) / 2, (mx[2] + mn[2]) / 2);
  v3 c = rot(axis, o);
  for (int k = 0; k < 3; ++k) {
    to[k] = c.x[k];
    extent[k] = (mx[k] - mn[k]) * 0.5;
  }
}

/* segmentSqrDistance (sphere_triangle-inl.h:84-108) */
static double segment_sqr_distance(v3 from, v3 to, v3 p, v3 *nearest) {
  v3 diff = vsub(p, from);
  v3 v = vsub(to, from);
  double t = vdot(v, diff);
  if (t > 0) {
    double dotVV = vdot(v, v);
    if (t < dotVV) {
      t /= dotVV;
      diff = vsub(diff, vmul(v, t));
    } else {
      t = 1;
      diff = vsub(diff, v);
    }
  } else {
    t = 0;
  }
  *nearest = vadd(from, vmul(v, t));
  return vdot(diff, diff);
}

/* projectInTriangle (sphere_triangle-inl.h:112-135) */
static int project_in_triangle(v3 p1, v3 p2, v3 p3, v3 normal, v3 p) {
  double r1 = vdot(vcross(vsub(p2, p1), normal), vsub(p, p1));
  double r2 = vdot(vcross(vsub(p3, p2), normal), vsub(p, p2));
  double r3 = vdot(vcross(vsub(p1, p3), normal), vsub(p, p3));
  return (r1 > 0 && r2 > 0 && r3 > 0) || (r1 <= 0 && r2 <= 0 && r3 <= 0);
}

/* sphereTriangleIntersect (sphere_triangle-inl.h:138-230); the outputs are
 * the solver's (normal from the sphere's centre towards the contact). */
static int sphere_triangle(double radius, v3 center, v3 P1, v3 P2, v3 P3, v3 *point,
                           double *depth, v3 *normal_out) {
  v3 normal = vnormalized(vcross(vsub(P2, P1), vsub(P3, P1)));
  double rt = radius + DBL_EPSILON;
  double dist = vdot(vsub(center, P1), normal);
  if (dist < 0) {
    dist *= -1;
    normal = vmul(normal, -1);
  }
  if (!(dist < rt)) return 0;
  int has = 0;
  v3 cp = V(NAN, NAN, NAN);
  if (project_in_triangle(P1, P2, P3, normal, center)) {
    has = 1;
    cp = vsub(center, vmul(normal, dist));
  } else {
    double r2 = rt * rt;
    v3 ne;
    if (segment_sqr_distance(P1, P2, center, &ne) < r2) { has = 1; cp = ne; }
    if (segment_sqr_distance(P2, P3, center, &ne) < r2) { has = 1; cp = ne; }
    if (segment_sqr_distance(P3, P1, center, &ne) < r2) { has = 1; cp = ne; }
  }
  if (!has) return 0;
  v3 cc = vsub(cp, center);
  double dsq = vsq(cc);
  if (!(dsq < rt * rt)) return 0;
  if (dsq > 0) {
    *normal_out = vnormalized(cc);
    *depth = -(radius - sqrt(dsq));
  } else {
    *normal_out = vneg(normal);
    *depth = -radius;
  }
  *point = cp;
  return 1;
}

int fclo_sphere_triangle(double radius, const double center[3], const double P1[3],
                         const double P2[3], const double P3[3], double point[3],
                         double *depth, double normal[3]) {
  v3 p, n;
  double d = 0;
  int hit = sphere_triangle(radius, V(center[0], center[1], center[2]), V(P1[0], P1[1], P1[2]),
                            V(P2[0], P2[1], P2[2]), V(P3[0], P3[1], P3[2]), &p, &d, &n);
  if (hit) {
    for (int k = 0; k < 3; ++k) {
      point[k] = p.x[k];
      normal[k] = n.x[k];
    }
    *depth = d;
  }
  return hit;
}

typedef struct {
  const fclo_model *m;
  const fclo_shape *s;
  const double *tf1, *tf2;
  xf T1, T2;
  fclo_node sbv;  /* the shape's OBB (world frame) */
  int cap, enable_contact, count;
  double tol;
  int *prims;
  fclo_contact *out;
  fclo_collide_stats *st;
} ms_state;

/* meshShapeCollisionOrientedNodeLeafTesting
 * (mesh_shape_collision_traversal_node-inl.h:205-283, enable_cost off) */
static void ms_leaf(ms_state *S, int prim) {
  S->st->num_leaf_tests++;
  const int *t = S->m->tris + 3 * prim;
  double tv[9];
  for (int k = 0; k < 3; ++k)
    for (int j = 0; j < 3; ++j) tv[3 * k + j] = S->m->vertices[3 * t[k] + j];
  fclo_contact c;
  int hit;
  if (S->s->type == FCLO_SHAPE_SPHERE) {
    /* gjk_solver_indep-inl.h:527-545: the triangle posed by tf1 */
    v3 P[3];
    for (int k = 0; k < 3; ++k) P[k] = xf_apply(&S->T1, V(tv[3 * k], tv[3 * k + 1], tv[3 * k + 2]));
    v3 p, n;
    double d;
    hit = sphere_triangle(S->s->dims[0], S->T2.t, P[0], P[1], P[2], &p, &d, &n);
    if (hit) {
      /* the leaf stores -normal and the solver's depth */
      for (int k = 0; k < 3; ++k) {
        c.pos[k] = p.x[k];
        c.normal[k] = -n.x[k];
      }
      c.depth = d;
    }
  } else {
    /* gjk_solver_indep-inl.h:448-505: GJK/EPA on (shape at tf2, TriangleP at
     * tf1); the leaf's -normal undoes the solver's, leaving EPA's normal */
    fclo_shape tri;
    tri.type = FCLO_SHAPE_TRIANGLE;
    tri.num_vertices = 3;
    tri.dims[0] = tri.dims[1] = tri.dims[2] = 0;
    tri.vertices = tv;
    int n = 0;
    hit = gjk_epa_intersect(S->s, &S->T2, &tri, &S->T1, S->tol, &c, &n);
  }
  if (!hit || S->count >= S->cap) return;
  int i = S->count++;
  if (S->prims) S->prims[i] = prim;
  if (S->out) {
    if (S->enable_contact)
      S->out[i] = c;
    else
      memset(&S->out[i], 0, sizeof(S->out[i]));
  }
}

/* collisionRecurse (traversal_recurse-inl.h:82-129) with the second node
 * always a leaf and firstOverSecond always true: BV test, then the leaf
 * test or both children, stopping once isSatisfied (count == cap). */
static void ms_recurse(ms_state *S, int b) {
  const fclo_node *nd = &S->m->bvs[b];
  S->st->num_bv_tests++;
  if (!fclo_obb_overlap(S->T1.R, S->T1.t.x, &S->sbv, nd)) return;
  if (nd->first_child < 0) {
    ms_leaf(S, -(nd->first_child + 1));
    return;
  }
  ms_recurse(S, nd->first_child);
  if (S->count >= S->cap) return;
  ms_recurse(S, nd->first_child + 1);
}

int fclo_mesh_shape_collide(const fclo_model *m, const double tf1[12], const fclo_shape *s,
                            const double tf2[12], int cap, int enable_contact, double tol,
                            int *prims, fclo_contact *out, fclo_collide_stats *st) {
  fclo_collide_stats dummy;
  ms_state S;
  memset(&S, 0, sizeof(S));
  S.m = m;
  S.s = s;
  S.T1 = xf_from(tf1);
  S.T2 = xf_from(tf2);
  S.cap = cap;
  S.enable_contact = enable_contact;
  S.tol = tol;
  S.prims = prims;
  S.out = out;
  S.st = st ? st : &dummy;
  S.st->num_contacts = 0;
  S.st->num_bv_tests = 0;
  S.st->num_leaf_tests = 0;
  if (cap <= 0 || m->num_bvs <= 0) return 0;
  fclo_shape_obb(s, tf2, S.sbv.obb_axis, S.sbv.obb_to, S.sbv.obb_extent);
  ms_recurse(&S, 0);
  S.st->num_contacts = S.count;
  return S.count;
}

void fclo_mesh_shape_collide_batch(const fclo_model *const *models, const int *mesh_ids,
                                   const fclo_shape_rec *recs, const double *pool,
                                   int num_queries, const int *shape_ids, const double *tf1s,
                                   const double *tf2s, int cap, int enable_contact, double tol,
                                   int *counts, int *prims, fclo_contact *out, int64_t *stats,
                                   int threads) {
#pragma omp parallel for schedule(dynamic, 16) num_threads(threads)
  for (int q = 0; q < num_queries; ++q) {
    const fclo_shape_rec *r = recs + shape_ids[q];
    fclo_shape s;
    s.type = r->type;
    s.num_vertices = r->num_vertices;
    for (int i = 0; i < 3; ++i) s.dims[i] = r->dims[i];
    s.vertices = r->type == FCLO_SHAPE_CONVEX ? pool + 3 * r->vertex_offset : 0;
    fclo_collide_stats st;
    counts[q] = fclo_mesh_shape_collide(
        models[mesh_ids[q]], tf1s + 12 * (size_t)q, &s, tf2s + 12 * (size_t)q, cap,
        enable_contact, tol, prims ? prims + (size_t)q * cap : 0,
        out ? out + (size_t)q * cap : 0, &st);
    if (stats) {
      stats[2 * q] = st.num_bv_tests;
      stats[2 * q + 1] = st.num_leaf_tests;
    }
  }
}
