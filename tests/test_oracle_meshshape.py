"""Pins the oracle's mesh-vs-primitive collide (BVHModel<OBBRSS> against a
Box, Sphere, Capsule, Cylinder or Convex with the GST_INDEP solver) with
checks that do not retype its formulas:

* the shape's fitted box contains the shape (support points in many
  directions) and, for a box with distinct sides, is the box itself;
* sphereTriangleIntersect against point-triangle distance by brute force and
  against closed-form answers;
* the sphere's closed form against GJK/EPA on the same (sphere, triangle);
* the traversal against a loop over every triangle with the same leaf test
  (the BV test only prunes), the cap as a prefix of the full list, and the
  reference's statistics bounds.

The reference's own tests for this path (test_fcl_shape_mesh_consistency.cpp)
compare mesh-shape results with shape-shape ones on tessellated shapes;
the brute-force loop here checks the same property at the level where the
two can differ, the pruning.
"""
import numpy as np
import pytest

import fclo
import util
from fcl_amd import BOX, CAPSULE, CONVEX, CYLINDER, SPHERE, icosphere, scale_to_radius, \
    shape_table

TOL = 1e-6


def _shape(kind, dims, verts=None):
    return fclo.make_shape(kind, dims, verts)


def _support(kind, dims, verts, d):
    d = d / np.linalg.norm(d)
    if kind == BOX:
        return np.sign(d) * np.asarray(dims) / 2
    if kind == SPHERE:
        return d * dims[0]
    if kind == CAPSULE:
        return d * dims[0] + np.array([0, 0, np.sign(d[2]) * dims[1] / 2])
    if kind == CYLINDER:
        r = np.hypot(d[0], d[1])
        xy = d[:2] / r * dims[0] if r > 0 else np.zeros(2)
        return np.array([xy[0], xy[1], np.sign(d[2]) * dims[1] / 2])
    return verts[np.argmax(verts @ d)]


HULL = np.array([[1, 0, 0], [-1, 0, 0], [0, 1.5, 0], [0, -1, 0], [0, 0, 0.7], [0.2, 0.3, -1.1],
                 [0.5, 0.5, 0.5]], np.float64) * 0.4

SHAPES = [(BOX, (0.6, 0.9, 0.4), None), (SPHERE, (0.45, 0, 0), None),
          (CAPSULE, (0.25, 0.8, 0), None), (CYLINDER, (0.3, 0.7, 0), None),
          (CONVEX, (0, 0, 0), HULL)]


@pytest.mark.parametrize("kind,dims,verts", SHAPES)
def test_shape_obb_contains_shape(kind, dims, verts):
    rng = np.random.default_rng(kind)
    tfs = util.random_transforms(20, [-3, -3, -3, 3, 3, 3], seed=kind + 10)
    s = _shape(kind, dims, verts)
    dirs = rng.normal(size=(400, 3))
    for tf in tfs:
        axis, to, ext = fclo.shape_obb(s, tf)
        assert np.allclose(axis.T @ axis, np.eye(3), atol=1e-9)
        R = tf[:9].reshape(3, 3)
        for d in dirs:
            p = R @ _support(kind, dims, verts, d) + tf[9:]
            local = axis.T @ (p - to)
            assert np.all(np.abs(local) <= ext + 1e-9), (kind, local, ext)


def test_box_obb_is_the_box():
    dims = (0.6, 1.4, 0.3)
    s = _shape(BOX, dims)
    for tf in util.random_transforms(10, [-2, -2, -2, 2, 2, 2], seed=3):
        axis, to, ext = fclo.shape_obb(s, tf)
        R = tf[:9].reshape(3, 3)
        # axes sorted by extent: largest side first
        order = np.argsort(dims)[::-1]
        for k in range(3):
            assert abs(abs(axis[:, k] @ R[:, order[k]]) - 1) < 1e-9
        assert np.allclose(ext, np.sort(dims)[::-1] / 2, atol=1e-12)
        assert np.allclose(to, tf[9:], atol=1e-12)


def test_sphere_triangle_closed_forms():
    A, B, C = np.array([-1.0, -1, 0]), np.array([1.0, -1, 0]), np.array([0.0, 1, 0])
    # centre above the interior: contact at the foot, depth -(r - h)
    hit, p, d, n = fclo.sphere_triangle(0.5, [0.1, 0, 0.3], A, B, C)
    assert hit and np.allclose(p, [0.1, 0, 0]) and abs(d + 0.2) < 1e-15
    assert np.allclose(n, [0, 0, -1])
    # below the plane: the normal flips
    hit, p, d, n = fclo.sphere_triangle(0.5, [0.1, 0, -0.3], A, B, C)
    assert hit and np.allclose(n, [0, 0, 1]) and abs(d + 0.2) < 1e-15
    # beside edge AB: nearest point on the edge
    hit, p, d, n = fclo.sphere_triangle(0.5, [0.2, -1.3, 0.1], A, B, C)
    dist = np.hypot(0.3, 0.1)
    assert hit and np.allclose(p, [0.2, -1, 0]) and abs(d + (0.5 - dist)) < 1e-12
    # centre in the plane, inside: depth -radius, normal -plane normal
    hit, p, d, n = fclo.sphere_triangle(0.5, [0.0, 0, 0], A, B, C)
    assert hit and d == -0.5 and np.allclose(p, [0, 0, 0])
    # too far
    assert not fclo.sphere_triangle(0.5, [0.1, 0, 0.6], A, B, C)[0]
    assert not fclo.sphere_triangle(0.5, [3.0, 0, 0.0], A, B, C)[0]


def test_sphere_triangle_matches_distance():
    rng = np.random.default_rng(5)
    n = 0
    for _ in range(3000):
        T = rng.uniform(-1, 1, (3, 3))
        c = rng.uniform(-1.5, 1.5, 3)
        r = rng.uniform(0.1, 1.0)
        dist = util.point_tri_dist(c, *T)
        if abs(dist - r) < 1e-9:
            continue
        hit, p, d, nrm = fclo.sphere_triangle(r, c, *T)
        assert hit == (dist < r), (dist, r)
        if hit:
            n += 1
            # the point is the foot on the face, else on the last edge within
            # reach (edges AB, BC, CA in turn), not always the nearest point
            pc = np.linalg.norm(p - c)
            assert dist - 1e-9 <= pc < r
            assert min(util.point_tri_dist(p, *T), 1.0) < 1e-9
            assert abs(d + (r - pc)) < 1e-12
    assert n > 300


def _tri_shape(T):
    return fclo.make_shape(fclo.TRIANGLE, (0, 0, 0), T)


def test_sphere_closed_form_agrees_with_gjk_epa():
    """The sphere's closed form and GJK/EPA on (sphere, TriangleP) agree on
    the boolean away from touching; for face contacts where EPA's depth is
    the closed form's, their normals point opposite ways (the closed form's leaf stores minus the solver's normal,
    GJK/EPA's leaf keeps EPA's), which the mesh path reproduces."""
    rng = np.random.default_rng(7)
    I = util.identity_pose()
    agree = faces = same = 0
    for _ in range(1500):
        T = rng.uniform(-1, 1, (3, 3))
        r = rng.uniform(0.2, 0.8)
        c = rng.uniform(-1, 1, 3)
        dist = util.point_tri_dist(c, *T)
        if abs(dist - r) < 1e-4:
            continue
        hit, p, d, n = fclo.sphere_triangle(r, c, *T)
        tf = I.copy()
        tf[9:] = c
        ok, cs = fclo.shape_intersect(_shape(SPHERE, (r, 0, 0)), tf, _tri_shape(T), I)
        assert ok == hit
        agree += 1
        if hit:
            nrm = np.cross(T[1] - T[0], T[2] - T[0])
            nrm /= np.linalg.norm(nrm)
            foot = c - (c - T[0]) @ nrm * nrm
            inside = util.point_tri_dist(foot, *T) < 1e-12 and dist > 0.05
            if inside and r - dist > 0.05:
                faces += 1
                normal, _, depth = cs[0]
                # EPA can stop on a face that is not the nearest (flat
                # triangle, see the shape oracle's unpinned depth); where it
                # finds the plane's push, the two answers coincide
                if abs(depth - d) < 1e-7:
                    same += 1
                    assert np.dot(normal, -n) < -1 + 1e-6
    assert agree > 1000 and faces > 20 and same > 5


def _mesh(k=6, r=1.0, seed=0):
    V, T = icosphere(k, 0.15, seed)
    return scale_to_radius(V, r), T


def _brute(model, V, T, kind, dims, verts, tf1, tf2):
    """Triangles whose leaf test hits, in index order."""
    hits = []
    s = _shape(kind, dims, verts)
    for t in range(T.shape[0]):
        P = V[T[t]]
        if kind == SPHERE:
            Pw = util.pose_apply(tf1, P)
            if fclo.sphere_triangle(dims[0], tf2[9:], *Pw)[0]:
                hits.append(t)
        elif fclo.shape_intersect(s, tf2, _tri_shape(np.ascontiguousarray(P)), tf1,
                                  contacts=False)[0]:
            hits.append(t)
    return hits


@pytest.mark.parametrize("kind,dims,verts", SHAPES)
def test_traversal_equals_all_triangles(kind, dims, verts):
    V, T = _mesh(6)
    m = fclo.OracleMesh(V, T)
    rec, pool = shape_table([kind], [dims], [verts] if kind == CONVEX else [])
    n = 12
    tf1 = util.random_transforms(n, [-0.3, -0.3, -0.3, 0.3, 0.3, 0.3], seed=20 + kind)
    tf2 = util.random_transforms(n, [-1.2, -1.2, -1.2, 1.2, 1.2, 1.2], seed=40 + kind)
    cap = T.shape[0]
    counts, prims, _, stats = fclo.mesh_shape_collide_batch(
        [m], np.zeros(n), rec, pool, np.zeros(n), tf1, tf2, cap=cap)
    some = 0
    for q in range(n):
        want = _brute(m, V, T, kind, dims, verts, tf1[q], tf2[q])
        got = prims[q, :counts[q]]
        assert sorted(got.tolist()) == want
        some += bool(want)
        assert stats[q, 1] <= T.shape[0] and stats[q, 0] <= m.num_bvs
        # leaves are tested only after their own BV test passed
        assert stats[q, 1] <= stats[q, 0]
    assert some >= 3
    # cap = k keeps the first k of the full list, and the count stops there
    c3, p3, _, s3 = fclo.mesh_shape_collide_batch(
        [m], np.zeros(n), rec, pool, np.zeros(n), tf1, tf2, cap=3)
    for q in range(n):
        k = min(3, counts[q])
        assert c3[q] == k and (p3[q, :k] == prims[q, :k]).all()
        assert s3[q, 0] <= stats[q, 0]


def test_contact_mode_same_hits_and_geometry():
    V, T = _mesh(5)
    m = fclo.OracleMesh(V, T)
    kinds = [BOX, SPHERE, CAPSULE, CYLINDER, CONVEX]
    rec, pool = shape_table(kinds, [d for _, d, _ in SHAPES], [HULL])
    n = 40
    rng = np.random.default_rng(3)
    sid = rng.integers(0, 5, n).astype(np.int32)
    tf1 = util.random_transforms(n, [-0.2, -0.2, -0.2, 0.2, 0.2, 0.2], seed=1)
    tf2 = util.random_transforms(n, [-1.1, -1.1, -1.1, 1.1, 1.1, 1.1], seed=2)
    cap = 64
    cb, pb, _, _ = fclo.mesh_shape_collide_batch([m], np.zeros(n), rec, pool, sid, tf1, tf2,
                                                 cap=cap)
    cc, pc, geo, _ = fclo.mesh_shape_collide_batch([m], np.zeros(n), rec, pool, sid, tf1, tf2,
                                                   cap=cap, enable_contact=True)
    assert (cb == cc).all() and (pb == pc).all() and cc.sum() > 20
    for q in range(n):
        for i in range(cc[q]):
            nrm, pos, depth = geo[q, i, :3], geo[q, i, 3:6], geo[q, i, 6]
            # the reference stores minus EPA's / the closed form's depth
            assert depth <= 1e-12
            assert abs(np.linalg.norm(nrm) - 1) < 1e-9
            # the contact point lies within the mesh's triangle's reach
            P = util.pose_apply(tf1[q], V[T[pc[q, i]]])
            assert util.point_tri_dist(pos, *P) < 1.0
