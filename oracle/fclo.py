"""ctypes binding of the CPU oracle (oracle/libfclo.so).

TEST INFRASTRUCTURE ONLY: the checker for the HIP engine.  Only tests/,
__graft_entry__.smoke() and bench.py's cpu_baseline leg use it.
"""
import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libfclo.so")

c_i32 = ctypes.c_int32
c_i64 = ctypes.c_int64
c_dbl = ctypes.c_double
P_i32 = ctypes.POINTER(c_i32)
P_i64 = ctypes.POINTER(c_i64)
P_dbl = ctypes.POINTER(c_dbl)

OBBRSS = 0
RSS = 1


class Node(ctypes.Structure):
    _fields_ = [("obb_axis", c_dbl * 9), ("obb_to", c_dbl * 3), ("obb_extent", c_dbl * 3),
                ("rss_axis", c_dbl * 9), ("rss_to", c_dbl * 3), ("rss_l", c_dbl * 2),
                ("rss_r", c_dbl), ("first_child", c_i32), ("first_primitive", c_i32),
                ("num_primitives", c_i32)]


class Model(ctypes.Structure):
    _fields_ = [("bv_type", c_i32), ("num_vertices", c_i32), ("num_tris", c_i32),
                ("num_bvs", c_i32), ("vertices", P_dbl), ("tris", P_i32),
                ("bvs", ctypes.POINTER(Node)), ("primitive_indices", ctypes.POINTER(ctypes.c_uint))]


class CollideStats(ctypes.Structure):
    _fields_ = [("num_contacts", c_i32), ("num_bv_tests", c_i64), ("num_leaf_tests", c_i64)]


class DistanceOut(ctypes.Structure):
    _fields_ = [("min_distance", c_dbl), ("p1", c_dbl * 3), ("p2", c_dbl * 3), ("b1", c_i32),
                ("b2", c_i32), ("num_bv_tests", c_i64), ("num_leaf_tests", c_i64)]


_lib = None


def build():
    subprocess.run(["make", "-s", "-C", HERE, "-B"], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = ctypes.CDLL(LIB_PATH)
        PM = ctypes.POINTER(Model)
        sig = {
            "fclo_dot": (c_dbl, [P_dbl, P_dbl]),
            "fclo_eigen_old": (None, [P_dbl, P_dbl, P_dbl]),
            "fclo_relative_transform": (None, [P_dbl, P_dbl, P_dbl, P_dbl]),
            "fclo_inverse_compose": (None, [P_dbl, P_dbl, P_dbl]),
            "fclo_model_build": (c_i32, [PM, c_i32, P_dbl, c_i32, P_i32, c_i32]),
            "fclo_model_free": (None, [PM]),
            "fclo_model_refit": (c_i32, [PM, P_dbl, c_i32]),
            "fclo_model_depth": (c_i32, [PM]),
            "fclo_node_size": (c_dbl, [PM, c_i32]),
            "fclo_obb_disjoint": (c_i32, [P_dbl, P_dbl, P_dbl, P_dbl]),
            "fclo_intersect_triangle": (c_i32, [P_dbl] * 8),
            "fclo_intersect_triangle_contact": (c_i32, [P_dbl] * 9 + [ctypes.POINTER(ctypes.c_uint),
                                                                      P_dbl, P_dbl]),
            "fclo_collide_contact": (None, [PM, P_dbl, PM, P_dbl, c_i32, c_i32, P_i32, P_dbl,
                                            P_dbl, P_dbl, ctypes.POINTER(CollideStats)]),
            "fclo_rect_distance": (c_dbl, [P_dbl, P_dbl, P_dbl, P_dbl, P_dbl, P_dbl]),
            "fclo_rss_distance": (c_dbl, [P_dbl, P_dbl, ctypes.POINTER(Node),
                                          ctypes.POINTER(Node)]),
            "fclo_tri_distance": (c_dbl, [P_dbl, P_dbl, P_dbl, P_dbl]),
            "fclo_collide": (None, [PM, P_dbl, PM, P_dbl, c_i32, P_i32,
                                    ctypes.POINTER(CollideStats)]),
            "fclo_distance": (None, [PM, P_dbl, PM, P_dbl, c_i32, ctypes.POINTER(DistanceOut)]),
            "fclo_collide_all_pairs": (c_i64, [PM, P_dbl, PM, P_dbl, P_i32, c_i64]),
            "fclo_distance_all_pairs": (c_dbl, [PM, P_dbl, PM, P_dbl, P_i32]),
            "fclo_collide_batch": (None, [ctypes.POINTER(PM), c_i32, P_i32, P_i32, P_dbl, P_dbl,
                                          c_i32, c_i32, P_i32, P_i32, P_i64, P_i64, c_i32]),
            "fclo_collide_front": (None, [PM, P_dbl, PM, P_dbl, c_i32, c_i32, P_i32, P_i32, c_i32,
                                          P_i32, c_i32, P_i32, ctypes.POINTER(CollideStats)]),
            "fclo_collide_batch_flat": (None, [ctypes.POINTER(PM), c_i32, P_i32, P_i32, P_dbl,
                                               P_dbl, c_i32, P_i64, P_i32, P_i32, c_i32]),
            "fclo_distance_batch": (None, [ctypes.POINTER(PM), c_i32, P_i32, P_i32, P_dbl,
                                           P_dbl, c_i32, P_dbl, P_dbl, P_i32, P_i64, P_i64,
                                           c_i32]),
            "fclo_shape_intersect": (c_i32, [ctypes.POINTER(Shape), P_dbl, ctypes.POINTER(Shape),
                                             P_dbl, c_dbl, ctypes.POINTER(Contact),
                                             ctypes.POINTER(c_i32)]),
            "fclo_shape_collide": (c_i32, [ctypes.POINTER(Shape), P_dbl, ctypes.POINTER(Shape),
                                           P_dbl, c_i32, c_i32, c_dbl, ctypes.POINTER(Contact)]),
            "fclo_shape_collide_batch": (None, [ctypes.POINTER(Shape), c_i32, P_i32, P_dbl, P_dbl,
                                                c_i32, c_i32, c_dbl, P_i32,
                                                ctypes.POINTER(Contact), c_i32, P_i32]),
            "fclo_shape_collide_records": (None, [ctypes.c_void_p, P_dbl, c_i32, P_i32, P_dbl,
                                                  P_dbl, c_i32, c_i32, c_dbl, P_i32,
                                                  ctypes.c_void_p, c_i32, P_i32]),
            "fclo_shape_obb": (None, [ctypes.POINTER(Shape), P_dbl, P_dbl, P_dbl, P_dbl]),
            "fclo_sphere_triangle": (c_i32, [c_dbl, P_dbl, P_dbl, P_dbl, P_dbl, P_dbl, P_dbl,
                                             P_dbl]),
            "fclo_mesh_shape_collide_batch": (None, [ctypes.POINTER(PM), P_i32, ctypes.c_void_p,
                                                     P_dbl, c_i32, P_i32, P_dbl, P_dbl, c_i32,
                                                     c_i32, c_dbl, P_i32, P_i32,
                                                     ctypes.c_void_p, P_i64, c_i32]),
        }
        for name, (res, args) in sig.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


BOX, SPHERE, CAPSULE, CYLINDER, CONVEX, TRIANGLE = 0, 1, 2, 3, 4, 5


class Shape(ctypes.Structure):
    _fields_ = [("type", c_i32), ("num_vertices", c_i32), ("dims", c_dbl * 3),
                ("vertices", P_dbl)]


class Contact(ctypes.Structure):
    _fields_ = [("normal", c_dbl * 3), ("pos", c_dbl * 3), ("depth", c_dbl)]


def make_shape(kind, dims=(0, 0, 0), vertices=None):
    """A primitive for the shape functions; vertices (n,3) for CONVEX (kept
    alive on the returned object)."""
    s = Shape()
    s.type = kind
    d = list(dims) + [0] * (3 - len(dims))
    for i in range(3):
        s.dims[i] = float(d[i])
    if vertices is not None:
        v = np.ascontiguousarray(vertices, np.float64)
        s._v = v
        s.vertices = v.ctypes.data_as(ctypes.POINTER(c_dbl))
        s.num_vertices = v.shape[0]
    return s


def shapes_from_table(rec, pool):
    """Oracle shapes for fclx_shape records (fields type, num_vertices, dims,
    vertex_offset) and their convex vertex pool (V,3)."""
    pool = _d(pool).reshape(-1, 3)
    out = []
    for r in rec:
        t = int(r["type"])
        if t == CONVEX:
            o = int(r["vertex_offset"])
            out.append(make_shape(t, (0, 0, 0), pool[o:o + int(r["num_vertices"])]))
        else:
            out.append(make_shape(t, tuple(float(x) for x in r["dims"])))
    return out


def shape_intersect(s1, tf1, s2, tf2, tol=1e-6, contacts=True):
    """GJKSolver_indep::shapeIntersect: (hit, [(normal, pos, depth), ...])."""
    tf1 = _d(tf1).reshape(12)
    tf2 = _d(tf2).reshape(12)
    out = (Contact * 8)()
    n = c_i32(0)
    hit = lib().fclo_shape_intersect(ctypes.byref(s1), _p(tf1, c_dbl), ctypes.byref(s2),
                                     _p(tf2, c_dbl), tol, out if contacts else None,
                                     ctypes.byref(n))
    cs = [(np.array(out[i].normal[:]), np.array(out[i].pos[:]), out[i].depth)
          for i in range(n.value)]
    return bool(hit), cs


def shape_collide(s1, tf1, s2, tf2, cap=1, enable_contact=True, tol=1e-6):
    """collide() of two primitives: the contacts added (normal, pos, depth)."""
    tf1 = _d(tf1).reshape(12)
    tf2 = _d(tf2).reshape(12)
    out = (Contact * 8)()
    n = lib().fclo_shape_collide(ctypes.byref(s1), _p(tf1, c_dbl), ctypes.byref(s2),
                                 _p(tf2, c_dbl), cap, 1 if enable_contact else 0, tol, out)
    return [(np.array(out[i].normal[:]), np.array(out[i].pos[:]), out[i].depth)
            for i in range(n)]


def shape_collide_batch(shapes, qs, tf1, tf2, cap=1, enable_contact=True, tol=1e-6,
                        threads=None, passes=False):
    """Per query q: shapes[qs[q,0]] vs shapes[qs[q,1]].  Returns counts (n,)
    and contacts (n, cap, 7) = normal, pos, depth; with passes=True also the
    (n, 2) passes of GJK's and EPA's loops."""
    n = qs.shape[0]
    arr = (Shape * len(shapes))(*shapes)
    qs = np.ascontiguousarray(qs, np.int32)
    tf1 = _d(tf1)
    tf2 = _d(tf2)
    counts = np.zeros(n, np.int32)
    out = (Contact * max(1, n * cap))()
    pp = np.zeros((n, 2), np.int32) if passes else None
    lib().fclo_shape_collide_batch(arr, n, _p(qs, c_i32), _p(tf1, c_dbl), _p(tf2, c_dbl),
                                   cap, 1 if enable_contact else 0, tol, _p(counts, c_i32),
                                   out, threads or (os.cpu_count() or 1),
                                   _p(pp, c_i32) if passes else None)
    c = np.frombuffer(out, dtype=np.float64).reshape(-1, 7)[: n * cap].reshape(n, cap, 7).copy()
    return (counts, c, pp) if passes else (counts, c)


def shape_collide_records(rec, pool, qs, tf1, tf2, cap=1, enable_contact=True, tol=1e-6,
                          threads=None, passes=False, contacts=True):
    """shape_collide_batch over 40-byte shape records (fields type,
    num_vertices, dims, vertex_offset) and their (V,3) vertex pool."""
    rec = np.ascontiguousarray(rec)
    assert rec.dtype.itemsize == 40
    pool = _d(pool).reshape(-1, 3) if len(pool) else np.zeros((1, 3))
    n = qs.shape[0]
    qs = np.ascontiguousarray(qs, np.int32)
    tf1 = _d(tf1)
    tf2 = _d(tf2)
    counts = np.zeros(n, np.int32)
    out = np.zeros((max(1, n * cap), 7)) if contacts else None
    pp = np.zeros((n, 2), np.int32) if passes else None
    lib().fclo_shape_collide_records(rec.ctypes.data, _p(pool, c_dbl), n, _p(qs, c_i32),
                                     _p(tf1, c_dbl), _p(tf2, c_dbl), cap,
                                     1 if enable_contact else 0, tol, _p(counts, c_i32),
                                     out.ctypes.data if contacts else None,
                                     threads or (os.cpu_count() or 1),
                                     _p(pp, c_i32) if passes else None)
    c = out[: n * cap].reshape(n, cap, 7) if contacts else None
    return (counts, c, pp) if passes else (counts, c)


def _p(a, t):
    return a.ctypes.data_as(ctypes.POINTER(t))


def _d(a):
    return np.ascontiguousarray(a, np.float64)


class OracleMesh:
    """A BVHModel built by the oracle's restatement of the reference builder."""

    def __init__(self, V, T, bv_type=OBBRSS):
        self.V = _d(V)
        self.T = np.ascontiguousarray(T, np.int32)
        self.m = Model()
        rc = lib().fclo_model_build(ctypes.byref(self.m), bv_type, _p(self.V, c_dbl),
                                    self.V.shape[0], _p(self.T, c_i32), self.T.shape[0])
        if rc != 0:
            raise ValueError("empty model")

    def __del__(self):
        try:
            lib().fclo_model_free(ctypes.byref(self.m))
        except Exception:
            pass

    @property
    def num_bvs(self):
        return self.m.num_bvs

    def depth(self):
        return lib().fclo_model_depth(ctypes.byref(self.m))

    def refit(self, V, bottomup=True):
        """endReplaceModel(refit = true, bottomup) after replacing the
        vertices by V (same count)."""
        V = _d(V)
        if V.shape != self.V.shape:
            raise ValueError("a replaced model keeps its vertex count")
        self.V = V
        lib().fclo_model_refit(ctypes.byref(self.m), _p(V, c_dbl), 1 if bottomup else 0)

    def nodes(self):
        """(bv[nodes,31], links[nodes,3], sizes[nodes]) in the layout of
        fclx_mesh_export_nodes."""
        n = self.m.num_bvs
        raw = np.ctypeslib.as_array(ctypes.cast(self.m.bvs, ctypes.POINTER(c_dbl)),
                                    shape=(n * ctypes.sizeof(Node) // 8,))
        rec = raw.reshape(n, ctypes.sizeof(Node) // 8)
        bv = np.zeros((n, 31), np.float64)
        bv[:, :30] = rec[:, :30]
        ints = np.ctypeslib.as_array(ctypes.cast(self.m.bvs, P_i32),
                                     shape=(n * ctypes.sizeof(Node) // 4,))
        ints = ints.reshape(n, ctypes.sizeof(Node) // 4)
        links = ints[:, 60:63].copy()
        sizes = np.array([lib().fclo_node_size(ctypes.byref(self.m), i) for i in range(n)])
        return bv.copy(), links, sizes


def collide(m1, tf1, m2, tf2, cap=1):
    tf1 = _d(tf1).reshape(12)
    tf2 = _d(tf2).reshape(12)
    pairs = np.zeros((max(cap, 1), 2), np.int32)
    st = CollideStats()
    lib().fclo_collide(ctypes.byref(m1.m), _p(tf1, c_dbl), ctypes.byref(m2.m), _p(tf2, c_dbl),
                       cap, _p(pairs, c_i32), ctypes.byref(st))
    return pairs[: st.num_contacts].copy(), (st.num_bv_tests, st.num_leaf_tests)


def collide_contact(m1, tf1, m2, tf2, cap=1, store=None):
    """collide() with enable_contact: (pairs (n,2), pos (n,3), normal (n,3),
    depth (n,), (num_bv_tests, num_leaf_tests)); n = contacts."""
    tf1 = _d(tf1).reshape(12)
    tf2 = _d(tf2).reshape(12)
    store = cap if store is None else store
    k = max(store, 1)
    pairs = np.zeros((k, 2), np.int32)
    pos = np.zeros((k, 3))
    nrm = np.zeros((k, 3))
    dep = np.zeros(k)
    st = CollideStats()
    lib().fclo_collide_contact(ctypes.byref(m1.m), _p(tf1, c_dbl), ctypes.byref(m2.m),
                               _p(tf2, c_dbl), cap, store, _p(pairs, c_i32), _p(pos, c_dbl),
                               _p(nrm, c_dbl), _p(dep, c_dbl), ctypes.byref(st))
    n = min(st.num_contacts, store)
    return (pairs[:n].copy(), pos[:n].copy(), nrm[:n].copy(), dep[:n].copy(),
            (st.num_bv_tests, st.num_leaf_tests), st.num_contacts)


def distance(m1, tf1, m2, tf2, nearest=True):
    tf1 = _d(tf1).reshape(12)
    tf2 = _d(tf2).reshape(12)
    o = DistanceOut()
    lib().fclo_distance(ctypes.byref(m1.m), _p(tf1, c_dbl), ctypes.byref(m2.m),
                        _p(tf2, c_dbl), 1 if nearest else 0, ctypes.byref(o))
    pts = np.array([list(o.p1), list(o.p2)])
    return o.min_distance, pts, (o.b1, o.b2), (o.num_bv_tests, o.num_leaf_tests)


def collide_all_pairs(m1, tf1, m2, tf2, store=1 << 20):
    tf1 = _d(tf1).reshape(12)
    tf2 = _d(tf2).reshape(12)
    pairs = np.zeros((store, 2), np.int32)
    n = lib().fclo_collide_all_pairs(ctypes.byref(m1.m), _p(tf1, c_dbl), ctypes.byref(m2.m),
                                     _p(tf2, c_dbl), _p(pairs, c_i32), store)
    if n > store:
        raise ValueError("more intersecting pairs than store")
    return pairs[:n].copy()


def distance_all_pairs(m1, tf1, m2, tf2):
    tf1 = _d(tf1).reshape(12)
    tf2 = _d(tf2).reshape(12)
    ids = np.zeros(2, np.int32)
    d = lib().fclo_distance_all_pairs(ctypes.byref(m1.m), _p(tf1, c_dbl), ctypes.byref(m2.m),
                                      _p(tf2, c_dbl), _p(ids, c_i32))
    return d, (int(ids[0]), int(ids[1]))


def _model_array(meshes):
    PM = ctypes.POINTER(Model)
    arr = (PM * len(meshes))(*[ctypes.pointer(m.m) for m in meshes])
    return arr


def collide_batch(meshes, qm, tf1, tf2, cap, store=0, threads=None):
    """Batch collide over queries; qm (n,2) indexes meshes.  Returns counts,
    pairs (n, store, 2), bv_tests, leaf_tests."""
    qm = np.ascontiguousarray(qm, np.int32)
    n = qm.shape[0]
    q1 = np.ascontiguousarray(qm[:, 0])
    q2 = np.ascontiguousarray(qm[:, 1])
    tf1 = _d(tf1)
    tf2 = _d(tf2)
    counts = np.zeros(n, np.int32)
    pairs = np.zeros((n, max(store, 1), 2), np.int32)
    bvt = np.zeros(n, np.int64)
    lft = np.zeros(n, np.int64)
    arr = _model_array(meshes)
    lib().fclo_collide_batch(arr, n, _p(q1, c_i32), _p(q2, c_i32), _p(tf1, c_dbl),
                             _p(tf2, c_dbl), cap, store, _p(counts, c_i32), _p(pairs, c_i32),
                             _p(bvt, c_i64), _p(lft, c_i64), threads or os.cpu_count())
    return counts, pairs[:, :store], bvt, lft


def collide_front(m1, tf1, m2, tf2, front, cap=1 << 30):
    """One collide() with a BVTT front list: front (k,2) int32 from the
    previous call (k = 0: traverse from the roots).  Returns (pairs (n,2) in
    order, new front (k',2), (num_contacts, num_bv_tests, num_leaf_tests))."""
    tf1 = _d(tf1).reshape(12)
    tf2 = _d(tf2).reshape(12)
    fin = np.ascontiguousarray(front, np.int32).reshape(-1, 2)
    store = 4096
    room = max(2 * len(fin), 1024)
    while True:
        pairs = np.zeros((min(cap, store), 2), np.int32)
        fout = np.zeros((room, 2), np.int32)
        n = c_i32(0)
        st = CollideStats()
        lib().fclo_collide_front(ctypes.byref(m1.m), _p(tf1, c_dbl), ctypes.byref(m2.m),
                                 _p(tf2, c_dbl), cap, len(pairs), _p(pairs, c_i32), _p(fin, c_i32),
                                 len(fin), _p(fout, c_i32), room, ctypes.byref(n),
                                 ctypes.byref(st))
        if n.value <= room and min(st.num_contacts, cap) <= len(pairs):
            return (pairs[:min(st.num_contacts, cap)].copy(), fout[:n.value].copy(),
                    (st.num_contacts, st.num_bv_tests, st.num_leaf_tests))
        room = max(room, n.value)
        store = max(store, min(st.num_contacts, cap))


def collide_batch_flat(meshes, qm, tf1, tf2, cap, threads=None):
    """Batch collide with every contact kept: a counting pass, then a pass
    that writes each query's pairs (reference order) at its offset.
    Returns counts (n,), offsets (n+1,) int64 and records (total, 2)."""
    qm = np.ascontiguousarray(qm, np.int32)
    n = qm.shape[0]
    counts = collide_batch(meshes, qm, tf1, tf2, cap, store=0, threads=threads)[0]
    offsets = np.zeros(n + 1, np.int64)
    np.cumsum(counts, out=offsets[1:])
    recs = np.zeros((max(int(offsets[-1]), 1), 2), np.int32)
    again = np.zeros(n, np.int32)
    q1 = np.ascontiguousarray(qm[:, 0])
    q2 = np.ascontiguousarray(qm[:, 1])
    lib().fclo_collide_batch_flat(_model_array(meshes), n, _p(q1, c_i32), _p(q2, c_i32),
                                  _p(_d(tf1), c_dbl), _p(_d(tf2), c_dbl), cap,
                                  _p(offsets, c_i64), _p(again, c_i32), _p(recs, c_i32),
                                  threads or os.cpu_count())
    assert np.array_equal(again, counts)
    return counts, offsets, recs[: offsets[-1]]


def distance_batch(meshes, qm, tf1, tf2, nearest=True, threads=None):
    qm = np.ascontiguousarray(qm, np.int32)
    n = qm.shape[0]
    q1 = np.ascontiguousarray(qm[:, 0])
    q2 = np.ascontiguousarray(qm[:, 1])
    tf1 = _d(tf1)
    tf2 = _d(tf2)
    dist = np.zeros(n)
    pts = np.zeros((n, 6))
    ids = np.zeros((n, 2), np.int32)
    bvt = np.zeros(n, np.int64)
    lft = np.zeros(n, np.int64)
    arr = _model_array(meshes)
    lib().fclo_distance_batch(arr, n, _p(q1, c_i32), _p(q2, c_i32), _p(tf1, c_dbl),
                              _p(tf2, c_dbl), 1 if nearest else 0, _p(dist, c_dbl),
                              _p(pts, c_dbl), _p(ids, c_i32), _p(bvt, c_i64), _p(lft, c_i64),
                              threads or os.cpu_count())
    return dist, pts, ids, bvt, lft


# primitive tests
def obb_disjoint(B, T, a, b):
    return lib().fclo_obb_disjoint(_p(_d(B), c_dbl), _p(_d(T), c_dbl), _p(_d(a), c_dbl),
                                   _p(_d(b), c_dbl))


def intersect_triangle(P, Q, R=None, T=None):
    """P, Q: (3,3) vertex arrays; (R,T) moves Q into P's frame (identity by default)."""
    P = _d(P)
    Q = _d(Q)
    R = _d(np.eye(3) if R is None else R)
    T = _d(np.zeros(3) if T is None else T)
    args = [_p(np.ascontiguousarray(P[i]), c_dbl) for i in range(3)]
    qs = [np.ascontiguousarray(Q[i]) for i in range(3)]
    args += [_p(q, c_dbl) for q in qs] + [_p(R, c_dbl), _p(T, c_dbl)]
    return lib().fclo_intersect_triangle(*args)


def intersect_triangle_contact(P, Q, R=None, T=None):
    """Contact form: (intersects, points (n,3) in P's frame, normal, depth)."""
    P = _d(P)
    Q = _d(Q)
    R = _d(np.eye(3) if R is None else R)
    T = _d(np.zeros(3) if T is None else T)
    args = [_p(np.ascontiguousarray(P[i]), c_dbl) for i in range(3)]
    qs = [np.ascontiguousarray(Q[i]) for i in range(3)]
    args += [_p(q, c_dbl) for q in qs] + [_p(R, c_dbl), _p(T, c_dbl)]
    pts = np.zeros(6)
    n = ctypes.c_uint(0)
    depth = np.zeros(1)
    normal = np.zeros(3)
    hit = lib().fclo_intersect_triangle_contact(*args, _p(pts, c_dbl), ctypes.byref(n),
                                                _p(depth, c_dbl), _p(normal, c_dbl))
    return bool(hit), pts[: 3 * n.value].reshape(-1, 3), normal, float(depth[0])


def rect_distance(Rab, Tab, a, b):
    P = np.zeros(3)
    Q = np.zeros(3)
    d = lib().fclo_rect_distance(_p(_d(Rab), c_dbl), _p(_d(Tab), c_dbl), _p(_d(a), c_dbl),
                                 _p(_d(b), c_dbl), _p(P, c_dbl), _p(Q, c_dbl))
    return d, P, Q


def rss_distance_nodes(axis1, to1, l1, r1, axis2, to2, l2, r2, R0=None, T0=None):
    """RSS-RSS distance (RSS-inl.h:1956-1974) for hand-built rectangles."""
    n1, n2 = Node(), Node()
    for n, ax, to, l, r in ((n1, axis1, to1, l1, r1), (n2, axis2, to2, l2, r2)):
        n.rss_axis[:] = list(_d(ax).reshape(9))
        n.rss_to[:] = list(_d(to))
        n.rss_l[:] = list(_d(l))
        n.rss_r = float(r)
    R0 = _d(np.eye(3) if R0 is None else R0)
    T0 = _d(np.zeros(3) if T0 is None else T0)
    return lib().fclo_rss_distance(_p(R0, c_dbl), _p(T0, c_dbl), ctypes.byref(n1),
                                   ctypes.byref(n2))


def tri_distance(S, T):
    S = _d(S).reshape(9)
    T = _d(T).reshape(9)
    P = np.zeros(3)
    Q = np.zeros(3)
    d = lib().fclo_tri_distance(_p(S, c_dbl), _p(T, c_dbl), _p(P, c_dbl), _p(Q, c_dbl))
    return d, P, Q


def eigen_old(M):
    M = _d(M).reshape(9)
    d = np.zeros(3)
    v = np.zeros(9)
    lib().fclo_eigen_old(_p(M, c_dbl), _p(d, c_dbl), _p(v, c_dbl))
    return d, v.reshape(3, 3)


def shape_obb(shape, tf):
    """(axis (3,3), centre, extent) of the OBB computeBV<OBBRSS> gives a posed
    shape (the generic fit over its bound vertices)."""
    axis = np.zeros(9)
    to = np.zeros(3)
    ext = np.zeros(3)
    lib().fclo_shape_obb(ctypes.byref(shape), _p(_d(tf), c_dbl), _p(axis, c_dbl),
                         _p(to, c_dbl), _p(ext, c_dbl))
    return axis.reshape(3, 3), to, ext


def sphere_triangle(radius, center, P1, P2, P3):
    """sphereTriangleIntersect: (hit, point, depth, normal) as the solver
    returns them."""
    point = np.zeros(3)
    normal = np.zeros(3)
    depth = np.zeros(1)
    hit = lib().fclo_sphere_triangle(float(radius), _p(_d(center), c_dbl), _p(_d(P1), c_dbl),
                                     _p(_d(P2), c_dbl), _p(_d(P3), c_dbl), _p(point, c_dbl),
                                     _p(depth, c_dbl), _p(normal, c_dbl))
    return bool(hit), point, float(depth[0]), normal


def mesh_shape_collide_batch(meshes, mesh_ids, rec, pool, shape_ids, tf1, tf2, cap=1,
                             enable_contact=False, tol=1e-6, threads=None):
    """collide(mesh, tf1, shape, tf2) per query (BVHModel<OBBRSS> vs a
    primitive, GST_INDEP).  Returns counts (n,), prims (n, cap) -- the
    contacts' b1 --, contacts (n, cap, 7) = normal, pos, depth, and stats
    (n, 2) = num_bv_tests, num_leaf_tests."""
    rec = np.ascontiguousarray(rec)
    assert rec.dtype.itemsize == 40
    pool = _d(pool).reshape(-1, 3) if len(pool) else np.zeros((1, 3))
    mesh_ids = np.ascontiguousarray(mesh_ids, np.int32)
    shape_ids = np.ascontiguousarray(shape_ids, np.int32)
    n = shape_ids.shape[0]
    tf1 = _d(tf1)
    tf2 = _d(tf2)
    counts = np.zeros(n, np.int32)
    prims = np.full((max(1, n * cap),), -1, np.int32)
    out = np.zeros((max(1, n * cap), 7))
    stats = np.zeros((max(1, n), 2), np.int64)
    arr = _model_array(meshes)
    lib().fclo_mesh_shape_collide_batch(arr, _p(mesh_ids, c_i32), rec.ctypes.data,
                                        _p(pool, c_dbl), n, _p(shape_ids, c_i32),
                                        _p(tf1, c_dbl), _p(tf2, c_dbl), cap,
                                        1 if enable_contact else 0, tol, _p(counts, c_i32),
                                        _p(prims, c_i32), out.ctypes.data, _p(stats, c_i64),
                                        threads or (os.cpu_count() or 1))
    return (counts, prims[: n * cap].reshape(n, cap), out[: n * cap].reshape(n, cap, 7),
            stats[:n])
