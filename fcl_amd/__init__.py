"""fcl_amd — MI355X engine for FCL's BVH-vs-BVH collide()/distance() path.

Python plumbing over the C ABI in include/fclx.h: contexts, mesh upload,
single-query and batched calls, and the seeded workload generators.  The
compute runs in libfclx.so's HIP kernels; nothing here computes a result.
Device buffers for batched calls are torch tensors (device memory and streams
only).
"""
import ctypes

import numpy as np

from . import _native as N

OBBRSS = N.FCLX_BV_OBBRSS
RSS = N.FCLX_BV_RSS
ENGINE_PLAIN = N.FCLX_ENGINE_PLAIN
ENGINE_FAST = N.FCLX_ENGINE_FAST
ENGINE_BOOL = N.FCLX_ENGINE_BOOL
BOX = N.FCLX_SHAPE_BOX
SPHERE = N.FCLX_SHAPE_SPHERE
CAPSULE = N.FCLX_SHAPE_CAPSULE
CYLINDER = N.FCLX_SHAPE_CYLINDER
CONVEX = N.FCLX_SHAPE_CONVEX

# one fclx_shape record (include/fclx.h), 40 bytes
SHAPE_DTYPE = np.dtype([("type", "<i4"), ("num_vertices", "<i4"), ("dims", "<f8", (3,)),
                        ("vertex_offset", "<i8")])
assert SHAPE_DTYPE.itemsize == ctypes.sizeof(N.ShapeRec) and all(
    SHAPE_DTYPE.fields[f][1] == getattr(N.ShapeRec, f).offset for f, _ in N.ShapeRec._fields_)


def shape_table(types, dims, convex_vertices=()):
    """Records for fclx_shape_collide_batch.  types (n,), dims (n,3); the k-th
    CONVEX entry takes convex_vertices[k] ((m,3) array).  Returns (records,
    vertex pool (V,3) float64)."""
    types = np.asarray(types, np.int32)
    rec = np.zeros(len(types), SHAPE_DTYPE)
    rec["type"] = types
    rec["dims"] = np.asarray(dims, np.float64).reshape(len(types), 3)
    conv = np.flatnonzero(types == CONVEX)
    if len(conv) != len(convex_vertices):
        raise ValueError("one vertex array per CONVEX shape is needed")
    counts = np.array([len(v) for v in convex_vertices], np.int64)
    offs = np.concatenate([[0], np.cumsum(counts)[:-1]]) if len(counts) else counts
    rec["num_vertices"][conv] = counts
    rec["vertex_offset"][conv] = offs
    pool = (np.concatenate([np.asarray(v, np.float64).reshape(-1, 3) for v in convex_vertices])
            if len(convex_vertices) else np.zeros((0, 3)))
    return rec, np.ascontiguousarray(pool)


class FclxError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__("fclx error %d: %s" % (code, msg))
        self.code = code


# ---------------------------------------------------------------- generators
def icosphere(k, noise=0.15, seed=0):
    """Frequency-k geodesic icosphere, radii perturbed by U(-noise, noise)."""
    L = N.lib()
    nv, nt = N.c_i32(), N.c_i32()
    L.fclx_synth_icosphere_size(k, ctypes.byref(nv), ctypes.byref(nt))
    V = np.empty((nv.value, 3), np.float64)
    T = np.empty((nt.value, 3), np.int32)
    rc = L.fclx_synth_icosphere(k, noise, seed, N.ptr(V, N.c_dbl), N.ptr(T, N.c_i32))
    if rc != 0:
        raise ValueError("bad icosphere parameters")
    return V, T


def heightfield(g, height_frac=0.05, seed=0):
    """g x g heightfield over [-1,1]^2, heights U(-h,h), h = 2*height_frac."""
    L = N.lib()
    nv, nt = N.c_i32(), N.c_i32()
    L.fclx_synth_heightfield_size(g, ctypes.byref(nv), ctypes.byref(nt))
    V = np.empty((nv.value, 3), np.float64)
    T = np.empty((nt.value, 3), np.int32)
    rc = L.fclx_synth_heightfield(g, height_frac, seed, N.ptr(V, N.c_dbl), N.ptr(T, N.c_i32))
    if rc != 0:
        raise ValueError("bad heightfield parameters")
    return V, T


def scale_to_radius(V, r=1.0):
    V = np.ascontiguousarray(V, np.float64)
    N.lib().fclx_synth_scale_to_radius(N.ptr(V, N.c_dbl), V.shape[0], r)
    return V


def poses(n, kind=0, p0=2.2, p1=0.0, seed=0):
    """(tf1, tf2) arrays of shape (n, 12); tf1 = identity.  kind 0: translation
    uniform in [-p0, p0]^3; kind 1: |t| ~ U[p0, p1] along a uniform direction."""
    tf1 = np.empty((n, 12), np.float64)
    tf2 = np.empty((n, 12), np.float64)
    rc = N.lib().fclx_synth_poses(n, kind, p0, p1, seed, N.ptr(tf1, N.c_dbl),
                                  N.ptr(tf2, N.c_dbl))
    if rc != 0:
        raise ValueError("bad pose parameters")
    return tf1, tf2


def uniform(n, seed):
    out = np.empty(n, np.float64)
    N.lib().fclx_synth_uniform(n, seed, N.ptr(out, N.c_dbl))
    return out


# ------------------------------------------------------------- host builder
def bvh_build(V, T, bv_type=OBBRSS):
    """Host-only hierarchy build (no device): returns dict of numpy arrays
    bv (nodes x 31), links (nodes x 3), sizes (nodes), depth."""
    V = np.ascontiguousarray(V, np.float64)
    T = np.ascontiguousarray(T, np.int32)
    nb = max(2 * T.shape[0] - 1, 1)
    bv = np.zeros((nb, 31), np.float64)
    links = np.zeros((nb, 3), np.int32)
    sizes = np.zeros(nb, np.float64)
    depth = N.c_i32()
    rc = N.lib().fclx_bvh_build(bv_type, N.ptr(V, N.c_dbl), V.shape[0], N.ptr(T, N.c_i32),
                                T.shape[0], N.ptr(bv, N.c_dbl), N.ptr(links, N.c_i32),
                                N.ptr(sizes, N.c_dbl), ctypes.byref(depth))
    if rc != 0:
        raise FclxError(rc, "bvh build failed")
    return {"bv": bv, "links": links, "sizes": sizes, "depth": depth.value}


# ------------------------------------------------------------------- context
_DTYPES = {"i32": "int32", "i64": "int64", "f64": "float64", "u8": "uint8"}


def _req(t, name, dtype, min_numel, optional=False):
    """Checks a tensor handed to a batched entry point: contiguous, of the
    given dtype, holding at least min_numel elements, so that the device code
    never reads or writes outside it.  Returns its address (None for an
    absent optional tensor).  _on_device checks the device afterwards."""
    if t is None:
        if optional:
            return None
        raise ValueError("%s is required" % name)
    want = _DTYPES[dtype]
    if str(t.dtype) != "torch." + want:
        raise ValueError("%s must be %s, got %s" % (name, want, t.dtype))
    if not t.is_contiguous():
        raise ValueError("%s must be contiguous" % name)
    if t.numel() < min_numel:
        raise ValueError("%s holds %d elements, at least %d are needed"
                         % (name, t.numel(), min_numel))
    return t.data_ptr()


def _on_device(*tensors):
    """Every tensor given (None skipped) must live in device memory."""
    for t in tensors:
        if t is not None and not t.is_cuda:
            raise ValueError("batched calls take device tensors, got one on %s" % t.device)


class Context:
    """One GPU.  Meshes are uploaded once and named by integer ids."""

    def __init__(self, device=0):
        self._L = N.lib()
        h = ctypes.c_void_p()
        rc = self._L.fclx_init(device, ctypes.byref(h))
        if rc != 0:
            raise FclxError(rc, "fclx_init failed (is a GPU visible?)")
        self._h = h
        self.device = device

    def close(self):
        if self._h:
            self._L.fclx_shutdown(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _check(self, rc):
        if rc != 0:
            raise FclxError(rc, (self._L.fclx_last_error(self._h) or b"").decode())
        return rc

    # meshes
    def mesh(self, V, T, bv_type=OBBRSS):
        V = np.ascontiguousarray(V, np.float64)
        T = np.ascontiguousarray(T, np.int32)
        mid = N.c_i32()
        self._check(self._L.fclx_mesh_create(self._h, bv_type, N.ptr(V, N.c_dbl), V.shape[0],
                                             N.ptr(T, N.c_i32), T.shape[0],
                                             ctypes.byref(mid)))
        return mid.value

    def mesh_refit(self, mid, vertices, bottomup=True, stream=None):
        """endReplaceModel(refit = true, bottomup) on the device with the
        mesh's vertices replaced by `vertices` ((n,3) float64 device tensor,
        the mesh's vertex count); returns once the refit is done."""
        info = self.mesh_info(mid)
        ptr = _req(vertices, "vertices", "f64", 3 * info["num_vertices"])
        _on_device(vertices)
        self._check(self._L.fclx_mesh_refit(self._h, int(mid), ctypes.c_void_p(ptr),
                                            1 if bottomup else 0, stream))

    def mesh_destroy(self, mid):
        self._check(self._L.fclx_mesh_destroy(self._h, mid))

    def mesh_info(self, mid):
        info = N.MeshInfo()
        self._check(self._L.fclx_mesh_get_info(self._h, mid, ctypes.byref(info)))
        return {f: getattr(info, f) for f, _ in N.MeshInfo._fields_}

    def export_nodes(self, mid):
        info = self.mesh_info(mid)
        nb = info["num_bvs"]
        bv = np.zeros((nb, 31), np.float64)
        links = np.zeros((nb, 3), np.int32)
        sizes = np.zeros(nb, np.float64)
        self._check(self._L.fclx_mesh_export_nodes(self._h, mid, N.ptr(bv, N.c_dbl),
                                                   N.ptr(links, N.c_i32),
                                                   N.ptr(sizes, N.c_dbl)))
        return {"bv": bv, "links": links, "sizes": sizes, "depth": info["depth"]}

    # single queries (host buffers)
    def collide(self, m1, tf1, m2, tf2, cap=1):
        tf1 = np.ascontiguousarray(tf1, np.float64).reshape(12)
        tf2 = np.ascontiguousarray(tf2, np.float64).reshape(12)
        room = min(cap, 4096)
        while True:
            pairs = np.zeros((room, 2), np.int32)
            n = N.c_i32()
            stats = np.zeros(2, np.int64)
            rc = self._L.fclx_collide(self._h, m1, N.ptr(tf1, N.c_dbl), m2, N.ptr(tf2, N.c_dbl),
                                      cap, N.ptr(pairs, N.c_i32), room, ctypes.byref(n),
                                      N.ptr(stats, N.c_i64))
            if rc == N.FCLX_ERR_CAPACITY:
                room = n.value
                continue
            self._check(rc)
            return pairs[: n.value].copy(), stats

    def collide_contact(self, m1, tf1, m2, tf2, cap=1):
        """enable_contact form: (pairs (n,2), pos (n,3), normal (n,3), depth (n,), stats)."""
        tf1 = np.ascontiguousarray(tf1, np.float64).reshape(12)
        tf2 = np.ascontiguousarray(tf2, np.float64).reshape(12)
        room = min(cap, 4096)
        while True:
            pairs = np.zeros((room, 2), np.int32)
            pos = np.zeros((room, 3))
            nrm = np.zeros((room, 3))
            dep = np.zeros(room)
            n = N.c_i32()
            stats = np.zeros(2, np.int64)
            rc = self._L.fclx_collide_contact(
                self._h, m1, N.ptr(tf1, N.c_dbl), m2, N.ptr(tf2, N.c_dbl), cap,
                N.ptr(pairs, N.c_i32), N.ptr(pos, N.c_dbl), N.ptr(nrm, N.c_dbl),
                N.ptr(dep, N.c_dbl), room, ctypes.byref(n), N.ptr(stats, N.c_i64))
            if rc == N.FCLX_ERR_CAPACITY:
                room = n.value
                continue
            self._check(rc)
            k = n.value
            return pairs[:k].copy(), pos[:k].copy(), nrm[:k].copy(), dep[:k].copy(), stats

    def distance(self, m1, tf1, m2, tf2, nearest=True):
        tf1 = np.ascontiguousarray(tf1, np.float64).reshape(12)
        tf2 = np.ascontiguousarray(tf2, np.float64).reshape(12)
        d = N.c_dbl()
        pts = np.zeros(6, np.float64)
        ids = np.zeros(2, np.int32)
        stats = np.zeros(2, np.int64)
        self._check(self._L.fclx_distance(self._h, m1, N.ptr(tf1, N.c_dbl), m2,
                                          N.ptr(tf2, N.c_dbl), 1 if nearest else 0,
                                          ctypes.byref(d), N.ptr(pts, N.c_dbl),
                                          N.ptr(ids, N.c_i32), N.ptr(stats, N.c_i64)))
        return d.value, pts.reshape(2, 3), ids, stats

    # batched calls on device tensors
    def collide_batch_async(self, query_meshes, tf1, tf2, num_max_contacts, counts, pairs,
                            num_pairs, engine=ENGINE_PLAIN, bv_tests=None, leaf_tests=None,
                            stream=None, contact=None, noalloc=False):
        """contact: None, or (pos (K,3), normal (K,3), depth (K,)) float64 device
        tensors for enable_contact (a record per contact).  noalloc: call
        fclx_collide_batch_noalloc (ENGINE_BOOL only; no device allocation,
        its work area is shared by every call on the GPU)."""
        a = N.CollideBatchArgs()
        n = int(query_meshes.shape[0])
        a.num_queries = n
        a.query_meshes = _req(query_meshes, "query_meshes", "i32", 2 * n)
        a.tf1 = _req(tf1, "tf1", "f64", 12 * n)
        a.tf2 = _req(tf2, "tf2", "f64", 12 * n)
        a.num_max_contacts = int(num_max_contacts)
        a.engine = int(engine)
        a.counts = _req(counts, "counts", "i32", n)
        a.pairs = _req(pairs, "pairs", "i32", 0, optional=True)
        a.pair_capacity = int(pairs.numel() // 3) if pairs is not None else 0
        a.num_pairs = _req(num_pairs, "num_pairs", "i64", 1)
        a.bv_tests = _req(bv_tests, "bv_tests", "i64", n, optional=True)
        a.leaf_tests = _req(leaf_tests, "leaf_tests", "i64", n, optional=True)
        if contact is not None:
            a.enable_contact = 1
            cap = a.pair_capacity
            a.contact_pos = _req(contact[0], "contact pos", "f64", 3 * cap)
            a.contact_normal = _req(contact[1], "contact normal", "f64", 3 * cap)
            a.contact_depth = _req(contact[2], "contact depth", "f64", cap)
        _on_device(query_meshes, tf1, tf2, counts, pairs, num_pairs, bv_tests, leaf_tests,
                   *(contact or ()))
        fn = self._L.fclx_collide_batch_noalloc if noalloc else self._L.fclx_collide_batch
        self._check(fn(self._h, ctypes.byref(a), stream))

    def distance_batch_async(self, query_meshes, tf1, tf2, distance, points=None, ids=None,
                             engine=ENGINE_PLAIN, bv_tests=None, leaf_tests=None,
                             stream=None, nearest=True):
        a = N.DistanceBatchArgs()
        n = int(query_meshes.shape[0])
        a.num_queries = n
        a.query_meshes = _req(query_meshes, "query_meshes", "i32", 2 * n)
        a.tf1 = _req(tf1, "tf1", "f64", 12 * n)
        a.tf2 = _req(tf2, "tf2", "f64", 12 * n)
        a.enable_nearest_points = 1 if nearest else 0
        a.engine = int(engine)
        a.distance = _req(distance, "distance", "f64", n)
        a.points = _req(points, "points", "f64", 6 * n, optional=True)
        a.ids = _req(ids, "ids", "i32", 2 * n, optional=True)
        a.bv_tests = _req(bv_tests, "bv_tests", "i64", n, optional=True)
        a.leaf_tests = _req(leaf_tests, "leaf_tests", "i64", n, optional=True)
        _on_device(query_meshes, tf1, tf2, distance, points, ids, bv_tests, leaf_tests)
        self._check(self._L.fclx_distance_batch(self._h, ctypes.byref(a), stream))

    def shape_collide_batch_async(self, shapes, vertices, pair_shapes, tf1, tf2,
                                  num_max_contacts, counts, contact=None, tol=1e-6,
                                  stream=None, passes=None):
        """fclx_shape_collide_batch.  shapes: device uint8 tensor of fclx_shape
        records (shape_table's, viewed as bytes); vertices: device float64
        (V,3) or None; contact: None, or (pos, normal, depth) device tensors of
        (n*cap,3), (n*cap,3), (n*cap,) for enable_contact; passes: None or an
        (n,2) int32 device tensor for the GJK / EPA loop passes per pair."""
        a = N.ShapeCollideBatchArgs()
        n = int(pair_shapes.shape[0])
        cap = int(num_max_contacts)
        a.num_pairs = n
        a.num_shapes = int(shapes.numel() // SHAPE_DTYPE.itemsize)
        a.shapes = _req(shapes, "shapes", "u8", 0)
        has_v = vertices is not None and vertices.numel() > 0
        a.vertices = _req(vertices, "vertices", "f64", 0) if has_v else None
        a.num_vertices = int(vertices.numel() // 3) if has_v else 0
        a.pair_shapes = _req(pair_shapes, "pair_shapes", "i32", 2 * n)
        a.tf1 = _req(tf1, "tf1", "f64", 12 * n)
        a.tf2 = _req(tf2, "tf2", "f64", 12 * n)
        a.num_max_contacts = cap
        a.gjk_tolerance = float(tol)
        a.counts = _req(counts, "counts", "i32", n)
        if contact is not None:
            a.enable_contact = 1
            a.contact_pos = _req(contact[0], "contact pos", "f64", 3 * n * cap)
            a.contact_normal = _req(contact[1], "contact normal", "f64", 3 * n * cap)
            a.contact_depth = _req(contact[2], "contact depth", "f64", n * cap)
        a.passes = _req(passes, "passes", "i32", 2 * n, optional=True)
        _on_device(shapes, vertices, pair_shapes, tf1, tf2, counts, passes, *(contact or ()))
        self._check(self._L.fclx_shape_collide_batch(self._h, ctypes.byref(a), stream))

    def mesh_shape_collide_batch_async(self, query_meshes, shapes, vertices, query_shapes, tf1,
                                       tf2, num_max_contacts, counts, primitives=None,
                                       contact=None, tol=1e-6, stats=None, stream=None):
        """fclx_mesh_shape_collide_batch: collide(mesh, tf1, shape, tf2) per
        query for BVHModel<OBBRSS> meshes and primitive shapes.
        query_meshes, query_shapes: (n,) int32 device tensors; shapes: device
        uint8 tensor of fclx_shape records; vertices: device float64 (V,3) or
        None; primitives: None or an (n*cap,) int32 device tensor for the
        contacts' triangles; contact: None, or (pos, normal, depth) device
        tensors for enable_contact; stats: None or (n,2) int64 device tensor
        (BV tests, leaf tests)."""
        a = N.MeshShapeCollideBatchArgs()
        n = int(query_meshes.shape[0])
        cap = int(num_max_contacts)
        a.num_queries = n
        a.query_meshes = _req(query_meshes, "query_meshes", "i32", n)
        a.query_shapes = _req(query_shapes, "query_shapes", "i32", n)
        a.num_shapes = int(shapes.numel() // SHAPE_DTYPE.itemsize)
        a.shapes = _req(shapes, "shapes", "u8", 0)
        has_v = vertices is not None and vertices.numel() > 0
        a.vertices = _req(vertices, "vertices", "f64", 0) if has_v else None
        a.num_vertices = int(vertices.numel() // 3) if has_v else 0
        a.tf1 = _req(tf1, "tf1", "f64", 12 * n)
        a.tf2 = _req(tf2, "tf2", "f64", 12 * n)
        a.num_max_contacts = cap
        a.gjk_tolerance = float(tol)
        a.counts = _req(counts, "counts", "i32", n)
        a.primitives = _req(primitives, "primitives", "i32", n * cap, optional=True)
        if contact is not None:
            a.enable_contact = 1
            a.contact_pos = _req(contact[0], "contact pos", "f64", 3 * n * cap)
            a.contact_normal = _req(contact[1], "contact normal", "f64", 3 * n * cap)
            a.contact_depth = _req(contact[2], "contact depth", "f64", n * cap)
        a.stats = _req(stats, "stats", "i64", 2 * n, optional=True)
        _on_device(query_meshes, shapes, vertices, query_shapes, tf1, tf2, counts, primitives,
                   stats, *(contact or ()))
        self._check(self._L.fclx_mesh_shape_collide_batch(self._h, ctypes.byref(a), stream))

    def shape_collide(self, s1, v1, tf1, s2, v2, tf2, cap=1, contact=True, tol=1e-6):
        """fclx_shape_collide: one primitive pair with host arrays.  s1/s2 are
        shape_table records (one row each), v1/v2 a convex's (m,3) vertices
        or None.  Returns ([(normal, pos, depth), ...], (gjk_passes, epa_passes))."""
        recs = []
        for rec in (s1, s2):
            r = N.ShapeRec()
            r.type = int(rec["type"])
            r.num_vertices = int(rec["num_vertices"])
            for k in range(3):
                r.dims[k] = float(rec["dims"][k])
            recs.append(r)
        vs = [None if v is None else np.ascontiguousarray(v, np.float64).reshape(-1, 3)
              for v in (v1, v2)]
        t1 = np.ascontiguousarray(tf1, np.float64).reshape(12)
        t2 = np.ascontiguousarray(tf2, np.float64).reshape(12)
        pos = np.zeros((4, 3))
        nrm = np.zeros((4, 3))
        dep = np.zeros(4)
        n = ctypes.c_int32(0)
        stats = np.zeros(2, np.int32)
        self._check(self._L.fclx_shape_collide(
            self._h, ctypes.byref(recs[0]), None if vs[0] is None else N.ptr(vs[0], N.c_dbl),
            N.ptr(t1, N.c_dbl), ctypes.byref(recs[1]),
            None if vs[1] is None else N.ptr(vs[1], N.c_dbl), N.ptr(t2, N.c_dbl), int(cap),
            1 if contact else 0, float(tol), N.ptr(pos, N.c_dbl), N.ptr(nrm, N.c_dbl),
            N.ptr(dep, N.c_dbl), ctypes.byref(n), N.ptr(stats, N.c_i32)))
        return [(nrm[i].copy(), pos[i].copy(), float(dep[i])) for i in range(n.value)], tuple(stats)

    def broadphase_pairs_async(self, aabbs, pairs, num_pairs, groups=None, stream=None):
        """fclx_broadphase_pairs: aabbs (n,6) float64 device tensor (lower
        xyz, upper xyz); pairs (cap,2) int32; num_pairs (1,) int64."""
        n = int(aabbs.shape[0])
        a = _req(aabbs, "aabbs", "f64", 6 * n)
        g = _req(groups, "groups", "i32", n, optional=True)
        p = _req(pairs, "pairs", "i32", 0)
        cnt = _req(num_pairs, "num_pairs", "i64", 1)
        _on_device(aabbs, groups, pairs, num_pairs)
        self._check(self._L.fclx_broadphase_pairs(self._h, n, ctypes.c_void_p(a),
                                                  ctypes.c_void_p(g) if g else None,
                                                  ctypes.c_void_p(p), int(pairs.numel() // 2),
                                                  ctypes.c_void_p(cnt), stream))

    def set_profiling(self, enable=True):
        self._check(self._L.fclx_set_profiling(self._h, 1 if enable else 0))

    def set_engine_param(self, name, value):
        """Engine tunables (fclx_set_engine_param); results do not depend on them."""
        self._check(self._L.fclx_set_engine_param(self._h, name.encode(), float(value)))

    def get_engine_param(self, name):
        v = N.c_dbl(0)
        self._check(self._L.fclx_get_engine_param(self._h, name.encode(), ctypes.byref(v)))
        return v.value

    def engine_report(self):
        r = N.EngineReport()
        self._check(self._L.fclx_engine_last_report(self._h, ctypes.byref(r)))
        out = {}
        for f, _ in N.EngineReport._fields_:
            v = getattr(r, f)
            out[f] = list(v) if isinstance(v, ctypes.Array) else v
        return out

    def sync(self, stream=None):
        return self._check(self._L.fclx_sync(self._h, stream))

    def sync_status(self, stream=None):
        """Like sync() but returns the status code instead of raising."""
        return self._L.fclx_sync(self._h, stream)


class FrontSet:
    """fclx_front_set_*: device-resident BVTT front lists, one per sequence
    of temporally coherent queries (e.g. a link along a trajectory)."""

    def __init__(self, ctx, num_sequences, capacity=4096):
        self._ctx = ctx
        self._L = ctx._L
        h = ctypes.c_void_p()
        ctx._check(self._L.fclx_front_set_create(ctx._h, int(num_sequences), int(capacity),
                                                 ctypes.byref(h)))
        self._h = h
        self.num_sequences = int(num_sequences)
        self.capacity = int(capacity)

    def clear(self, stream=None):
        self._ctx._check(self._L.fclx_front_set_clear(self._h, stream))

    def read(self, sequence):
        """The sequence's front as an (n, 2) int32 array (b1, b2)."""
        n = ctypes.c_int32(0)
        buf = np.zeros((self.capacity, 2), np.int32)
        self._ctx._check(self._L.fclx_front_set_read(self._h, int(sequence), N.ptr(buf, N.c_i32),
                                                     self.capacity, ctypes.byref(n)))
        return buf[:min(n.value, self.capacity)].copy()

    def collide_async(self, seq_offsets, query_meshes, tf1, tf2, num_max_contacts, counts,
                      pairs, num_pairs, bv_tests=None, leaf_tests=None, stream=None):
        """fclx_collide_front_batch: sequence s is queries
        [seq_offsets[s], seq_offsets[s+1]), run in order from front s."""
        a = N.FrontBatchArgs()
        ns = int(seq_offsets.shape[0]) - 1
        n = int(query_meshes.shape[0])
        a.num_sequences = ns
        a.seq_offsets = _req(seq_offsets, "seq_offsets", "i32", ns + 1)
        a.query_meshes = _req(query_meshes, "query_meshes", "i32", 2 * n)
        a.tf1 = _req(tf1, "tf1", "f64", 12 * n)
        a.tf2 = _req(tf2, "tf2", "f64", 12 * n)
        a.num_max_contacts = int(num_max_contacts)
        a.counts = _req(counts, "counts", "i32", n)
        a.pairs = _req(pairs, "pairs", "i32", 0, optional=True)
        a.pair_capacity = int(pairs.numel() // 3) if pairs is not None else 0
        a.num_pairs = _req(num_pairs, "num_pairs", "i64", 1)
        a.bv_tests = _req(bv_tests, "bv_tests", "i64", n, optional=True)
        a.leaf_tests = _req(leaf_tests, "leaf_tests", "i64", n, optional=True)
        _on_device(seq_offsets, query_meshes, tf1, tf2, counts, pairs, num_pairs, bv_tests,
                   leaf_tests)
        self._ctx._check(self._L.fclx_collide_front_batch(self._ctx._h, ctypes.byref(a), self._h,
                                                          stream))

    def close(self):
        if self._h:
            self._L.fclx_front_set_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Server:
    """fclx_server_*: the request loop over one resident scene mesh.  A
    request is a robot's link meshes and its link poses along a trajectory
    ((waypoints, links, 12) float64); its result is a (waypoints, links)
    uint8 array, 1 where the link collides with the scene."""

    def __init__(self, ctx, scene_mesh, scene_pose=None, num_streams=6, hbm_budget=0,
                 max_batch_queries=0, engine=ENGINE_BOOL):
        self._ctx = ctx
        self._L = ctx._L
        cfg = N.ServerConfig()
        cfg.scene_mesh = int(scene_mesh)
        cfg.num_streams = int(num_streams)
        pose = (np.array([1, 0, 0, 0, 1, 0, 0, 0, 1, 0, 0, 0], np.float64) if scene_pose is None
                else np.ascontiguousarray(scene_pose, np.float64).reshape(12))
        for i in range(12):
            cfg.scene_pose[i] = float(pose[i])
        cfg.hbm_budget = int(hbm_budget)
        cfg.max_batch_queries = int(max_batch_queries)
        cfg.engine = int(engine)
        h = ctypes.c_void_p()
        ctx._check(self._L.fclx_server_create(ctx._h, ctypes.byref(cfg), ctypes.byref(h)))
        self._h = h
        self._live = {}

    def _request(self, link_meshes, poses, tag):
        lm = np.ascontiguousarray(link_meshes, np.int32)
        P = np.ascontiguousarray(poses, np.float64)
        assert P.ndim == 3 and P.shape[1] == len(lm) and P.shape[2] == 12
        res = np.zeros(P.shape[:2], np.uint8)
        r = N.Request()
        r.num_links = len(lm)
        r.num_waypoints = P.shape[0]
        r.link_meshes = lm.ctypes.data
        r.link_poses = P.ctypes.data
        r.result = res.ctypes.data
        r.tag = int(tag)
        return r, (lm, P, res)

    def submit(self, link_meshes, poses, tag):
        """Queues a request; its result array is returned by poll()."""
        r, keep = self._request(link_meshes, poses, tag)
        self._live[int(tag)] = keep
        self._ctx._check(self._L.fclx_server_submit(self._h, ctypes.byref(r)))

    def poll(self, max_n=256, wait_ms=0):
        """Completions: a list of (tag, status, num_queries, queue_ms,
        latency_ms, result array)."""
        buf = (N.Completion * max_n)()
        n = ctypes.c_int32(0)
        self._ctx._check(self._L.fclx_server_poll(self._h, buf, max_n, int(wait_ms),
                                                  ctypes.byref(n)))
        out = []
        for c in buf[:n.value]:
            keep = self._live.pop(int(c.tag), None)
            out.append((c.tag, c.status, c.num_queries, c.queue_ms, c.latency_ms,
                        keep[2] if keep else None))
        return out

    def replay(self, requests, arrival_ms):
        """requests: list of (link_meshes, poses); submits request i at
        arrival_ms[i] (ms after the call) and waits for all.  Returns (results,
        latencies_ms, elapsed_ms)."""
        n = len(requests)
        arr = (N.Request * max(1, n))()
        keep = []
        for i, (lm, P) in enumerate(requests):
            r, k = self._request(lm, P, i)
            arr[i] = r
            keep.append(k)
        t = np.ascontiguousarray(arrival_ms, np.float64)
        lat = np.zeros(n)
        que = np.zeros(n)
        el = ctypes.c_double(0)
        self._ctx._check(self._L.fclx_server_replay(self._h, arr, N.ptr(t, N.c_dbl), n,
                                                    N.ptr(lat, N.c_dbl), ctypes.byref(el),
                                                    N.ptr(que, N.c_dbl)))
        self.last_queue_ms = que  # each request's submission-to-launch time
        return [k[2] for k in keep], lat, el.value

    def stats(self):
        b = ctypes.c_int64(0)
        q = ctypes.c_int64(0)
        self._ctx._check(self._L.fclx_server_stats(self._h, ctypes.byref(b), ctypes.byref(q)))
        return b.value, q.value

    def close(self):
        if self._h:
            self._L.fclx_server_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
