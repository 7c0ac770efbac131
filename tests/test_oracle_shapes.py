"""Pins the oracle's primitive-shape path (GST_INDEP: GJK/EPA and the
closed forms) against FCL's own known answers in
test/test_fcl_geometric_shapes.cpp (the shapeIntersectionGJK_* cases, and the
closed-form sphere/box cases the two solvers share), against an independent
linear-programming intersection test for polytope pairs, and against a
closed-form penetration the EPA path must reproduce."""
import numpy as np
import pytest
from scipy.optimize import linprog

import fclo
import util

I = util.identity_pose()


def tr(x=0.0, y=0.0, z=0.0, R=None):
    t = I.copy()
    if R is not None:
        t[:9] = np.asarray(R).reshape(9)
    t[9:] = [x, y, z]
    return t


def compose(a, b):
    Ra, Rb = a[:9].reshape(3, 3), b[:9].reshape(3, 3)
    out = np.zeros(12)
    out[:9] = (Ra @ Rb).reshape(9)
    out[9:] = Ra @ b[9:] + a[9:]
    return out


def approx(a, b, tol):
    """Eigen isApprox: |a - b| <= tol * min(|a|, |b|)."""
    a, b = np.asarray(a, float), np.asarray(b, float)
    return np.linalg.norm(a - b) <= tol * min(np.linalg.norm(a), np.linalg.norm(b))


def check(s1, tf1, s2, tf2, expect, contacts=(), pos=True, depth=True, normal=True,
          tol=1e-9):
    """testShapeIntersection (test_fcl_geometric_shapes.cpp:427-516): the
    boolean from shapeIntersect and from collide(), and with contacts the
    same number of contacts, each expected one matched by an actual one."""
    hit, _ = fclo.shape_intersect(s1, tf1, s2, tf2, contacts=False)
    assert hit == expect
    hit, got = fclo.shape_intersect(s1, tf1, s2, tf2)
    assert hit == expect
    assert (len(fclo.shape_collide(s1, tf1, s2, tf2, cap=1, enable_contact=False)) > 0) == expect
    if not expect:
        return
    assert len(got) == len(contacts)
    used = set()
    for en, ep, ed in contacts:
        ok = False
        for j, (n, p, d) in enumerate(got):
            if j in used:
                continue
            if pos and not approx(ep, p, tol):
                continue
            if depth and not abs(ed - d) <= tol:
                continue
            if normal and not approx(en, n, tol):
                continue
            used.add(j)
            ok = True
            break
        assert ok, (en, ep, ed, got)
    # collide() with an unlimited cap reports the same contacts
    all_c = fclo.shape_collide(s1, tf1, s2, tf2, cap=1 << 20)
    assert len(all_c) == len(got)


@pytest.fixture(scope="module")
def transform():
    rng = np.random.default_rng(5)
    return tr(*rng.uniform(-1000, 1000, 3), R=util.random_rotation(rng))


def test_sphere_sphere_kats(transform):
    """test_fcl_geometric_shapes.cpp:4073-4163 (shapeIntersectionGJK_spheresphere)."""
    s1, s2 = fclo.make_shape(fclo.SPHERE, (20,)), fclo.make_shape(fclo.SPHERE, (10,))
    T = transform
    R = T[:9].reshape(3, 3)
    check(s1, I, s2, tr(40), False)
    check(s1, T, s2, compose(T, tr(40)), False)
    check(s1, I, s2, tr(30), True, [((1, 0, 0), (20, 0, 0), 0.0)])
    check(s1, I, s2, tr(30.01), False)
    check(s1, T, s2, compose(T, tr(30.01)), False)
    x = 20.0 - 0.1 * 20.0 / (20.0 + 10.0)
    check(s1, I, s2, tr(29.9), True, [((1, 0, 0), (x, 0, 0), 0.1)])
    check(s1, T, s2, compose(T, tr(29.9)), True,
          [(R @ [1, 0, 0], R @ [x, 0, 0] + T[9:], 0.1)])
    check(s1, I, s2, I, True, [((0, 0, 0), (0, 0, 0), 30.0)], normal=False, pos=False)
    check(s1, I, s2, tr(-29.9), True, [((-1, 0, 0), (-x, 0, 0), 0.1)])
    check(s1, I, s2, tr(-30.0), True, [((-1, 0, 0), (-20, 0, 0), 0.0)])
    check(s1, I, s2, tr(-30.01), False)
    # centres coincide: the normal is the zero vector (sphere_sphere-inl.h:79)
    hit, got = fclo.shape_intersect(s1, I, s2, I)
    assert hit and np.array_equal(got[0][0], [0, 0, 0]) and got[0][2] == 30.0


def test_box_box_kats(transform):
    """test_fcl_geometric_shapes.cpp:4168-4244 (shapeIntersectionGJK_boxbox):
    four contacts, normal +x, in each configuration."""
    s1, s2 = fclo.make_shape(fclo.BOX, (20, 40, 50)), fclo.make_shape(fclo.BOX, (10, 10, 10))
    T = transform
    R = T[:9].reshape(3, 3)
    a = 3.140 / 6
    q = np.array([[np.cos(a), -np.sin(a), 0], [np.sin(a), np.cos(a), 0], [0, 0, 1]])
    four = lambda n: [(n, (0, 0, 0), 0)] * 4
    check(s1, I, s2, I, True, four((1, 0, 0)), pos=False, depth=False)
    check(s1, T, s2, T, True, four(R @ [1, 0, 0]), pos=False, depth=False)
    check(s1, I, s2, tr(15), True, four((1, 0, 0)), pos=False, depth=False)
    check(s1, T, s2, compose(T, tr(15.01)), False)
    check(s1, I, s2, tr(R=q), True, four((1, 0, 0)), pos=False, depth=False)
    check(s1, T, s2, compose(T, tr(R=q)), True, four(R @ [1, 0, 0]), pos=False, depth=False)


def test_sphere_box_kats(transform):
    """test_fcl_geometric_shapes.cpp:4249-4313 (shapeIntersectionGJK_spherebox)
    and the shared closed-form cases at :834-897."""
    s1, s2 = fclo.make_shape(fclo.SPHERE, (20,)), fclo.make_shape(fclo.BOX, (5, 5, 5))
    T = transform
    check(s1, I, s2, I, True, [((-1, 0, 0), (0, 0, 0), 0)], pos=False, depth=False)
    check(s1, T, s2, T, True, [((0, 0, 0), (0, 0, 0), 0)], pos=False, depth=False, normal=False)
    check(s1, I, s2, tr(22.5), True, [((1, 0, 0), (0, 0, 0), 0)], pos=False, depth=False,
          tol=1e-7)
    check(s1, T, s2, compose(T, tr(22.51)), False)
    check(s1, I, s2, tr(22.4), True, [((1, 0, 0), (0, 0, 0), 0)], pos=False, depth=False,
          tol=1e-2)
    check(s1, T, s2, compose(T, tr(22.501)), False)


def test_sphere_capsule_kats(transform):
    """test_fcl_geometric_shapes.cpp:4318-4388 (shapeIntersectionGJK_spherecapsule)."""
    s1, s2 = fclo.make_shape(fclo.SPHERE, (20,)), fclo.make_shape(fclo.CAPSULE, (5, 10))
    T = transform
    R = T[:9].reshape(3, 3)
    check(s1, I, s2, I, True, [((0, 0, 0), (0, 0, 0), 0)], pos=False, depth=False, normal=False)
    check(s1, T, s2, T, True, [((0, 0, 0), (0, 0, 0), 0)], pos=False, depth=False, normal=False)
    check(s1, I, s2, tr(24.9), True, [((1, 0, 0), (0, 0, 0), 0)], pos=False, depth=False)
    check(s1, T, s2, compose(T, tr(24.9)), True, [(R @ [1, 0, 0], (0, 0, 0), 0)], pos=False,
          depth=False)
    check(s1, I, s2, tr(25), True, [((1, 0, 0), (0, 0, 0), 0)], pos=False, depth=False)
    check(s1, T, s2, compose(T, tr(25.1)), False)


def test_cylinder_cylinder_kats(transform):
    """test_fcl_geometric_shapes.cpp:4393-4439 (shapeIntersectionGJK_
    cylindercylinder): the GJK/EPA path, with the reference's tolerances
    (0.3 on the normal at 9.9: the built-in solver's normal is approximate)."""
    s = fclo.make_shape(fclo.CYLINDER, (5, 10))
    T = transform
    one = lambda n: [(n, (0, 0, 0), 0)]
    check(s, I, s, I, True, one((0, 0, 0)), pos=False, depth=False, normal=False)
    check(s, T, s, T, True, one((0, 0, 0)), pos=False, depth=False, normal=False)
    check(s, I, s, tr(9.9), True, one((1, 0, 0)), pos=False, depth=False, tol=3e-1)
    check(s, T, s, compose(T, tr(9.9)), True, one((0, 0, 0)), pos=False, depth=False,
          normal=False)
    check(s, I, s, tr(10), True, one((1, 0, 0)), pos=False, depth=False)
    check(s, T, s, compose(T, tr(10.01)), False)


def _box_vertices(a, b, c):
    return np.array([[x, y, z] for x in (-a / 2, a / 2) for y in (-b / 2, b / 2)
                     for z in (-c / 2, c / 2)])


def test_epa_result_bounds_closed_form_penetration():
    """A box against the convex hull of an equal box (so the pair takes the
    GJK/EPA path) overlapping by d along x (true penetration depth d), the
    second turned by 1e-7 rad: EPA's answer is a face of a polytope inside
    the Minkowski difference, so its distance lies in [0, d] and its normal
    is a unit vector; the reference reports depth = -epa.depth
    (gjk_solver_indep-inl.h:145-149).  When the expansion converges the
    face is the true one; the reference's EPA may also stop early
    (InvalidHull on near-coplanar supports), which its own tests allow for
    by never checking INDEP depths, so only the bound is asserted here, and
    equality is required for most of the cases."""
    rng = np.random.default_rng(11)
    exact = 0
    for _ in range(60):
        sides = rng.uniform(0.5, 2.0, 3)
        box = fclo.make_shape(fclo.BOX, sides)
        conv = fclo.make_shape(fclo.CONVEX, vertices=_box_vertices(*sides))
        d = rng.uniform(0.01, 0.2) * sides[0]
        R = util.random_rotation(rng)
        t = rng.uniform(-5, 5, 3)
        tf1 = tr(*t, R=R)
        ax = rng.normal(size=3)
        ax /= np.linalg.norm(ax)
        K = np.array([[0, -ax[2], ax[1]], [ax[2], 0, -ax[0]], [-ax[1], ax[0], 0]])
        turn = np.eye(3) + 1e-7 * K
        tf2 = compose(tf1, tr(sides[0] - d, 0.1 * rng.uniform(-1, 1),
                              0.1 * rng.uniform(-1, 1), R=turn))
        hit, got = fclo.shape_intersect(box, tf1, conv, tf2)
        assert hit and len(got) == 1
        n, p, depth = got[0]
        assert abs(np.linalg.norm(n) - 1) < 1e-12
        assert -1e-6 <= -depth <= d + 1e-6
        if abs(-depth - d) < 1e-6:
            exact += 1
            np.testing.assert_allclose(np.abs(n), [1, 0, 0], atol=1e-5)
    assert exact > 30


def _hull_intersect(A, B):
    """Do conv(A) and conv(B) intersect?  LP feasibility of
    sum l_i a_i = sum m_j b_j, l, m >= 0, sum l = sum m = 1."""
    na, nb = len(A), len(B)
    Aeq = np.zeros((5, na + nb))
    Aeq[:3, :na] = A.T
    Aeq[:3, na:] = -B.T
    Aeq[3, :na] = 1
    Aeq[4, na:] = 1
    beq = np.array([0, 0, 0, 1, 1.0])
    r = linprog(np.zeros(na + nb), A_eq=Aeq, b_eq=beq, bounds=(0, None), method="highs")
    return r.status == 0


def _separation(A, B):
    """Largest margin of a separating plane, negative if they overlap
    (support-function search over a dense direction set)."""
    rng = np.random.default_rng(0)
    D = rng.normal(size=(4000, 3))
    D /= np.linalg.norm(D, axis=1)[:, None]
    return np.max((A @ D.T).min(0) - (B @ D.T).max(0))


def test_polytope_booleans_match_lp():
    """Box/convex pairs (GJK path, and the box-box closed form): the
    boolean equals the LP intersection test whenever the pair is not within
    1e-3 of touching."""
    rng = np.random.default_rng(13)
    checked = 0
    for _ in range(300):
        k1, k2 = rng.integers(0, 2, 2)
        shapes, verts = [], []
        for k in (k1, k2):
            if k == 0:
                sides = rng.uniform(0.2, 1.0, 3)
                shapes.append(fclo.make_shape(fclo.BOX, sides))
                verts.append(_box_vertices(*sides))
            else:
                P = rng.normal(size=(12, 3)) * 0.4
                shapes.append(fclo.make_shape(fclo.CONVEX, vertices=P))
                verts.append(P)
        tf1 = tr(*rng.uniform(-0.6, 0.6, 3), R=util.random_rotation(rng))
        tf2 = tr(*rng.uniform(-0.6, 0.6, 3), R=util.random_rotation(rng))
        A = util.pose_apply(tf1, verts[0])
        B = util.pose_apply(tf2, verts[1])
        if abs(_separation(A, B)) < 1e-3 or (-1e-3 < _separation(B, A) < 1e-3):
            continue
        checked += 1
        hit, _ = fclo.shape_intersect(shapes[0], tf1, shapes[1], tf2, contacts=False)
        assert hit == _hull_intersect(A, B)
    assert checked > 200


def test_contact_cap_takes_deepest():
    """shape_collision_traversal_node-inl.h:80-96: with room for fewer
    contacts than the pair has, the deepest come first; with room for all,
    they come in the solver's order."""
    s1, s2 = fclo.make_shape(fclo.BOX, (20, 40, 50)), fclo.make_shape(fclo.BOX, (10, 10, 10))
    a = 0.3
    q = np.array([[np.cos(a), 0, np.sin(a)], [0, 1, 0], [-np.sin(a), 0, np.cos(a)]])
    tf2 = tr(13.0, R=q)
    _, allc = fclo.shape_intersect(s1, I, s2, tf2)
    assert len(allc) >= 2
    depths = sorted((c[2] for c in allc), reverse=True)
    for cap in range(1, len(allc)):
        got = fclo.shape_collide(s1, I, s2, tf2, cap=cap)
        assert [g[2] for g in got] == depths[:cap]
    got = fclo.shape_collide(s1, I, s2, tf2, cap=len(allc))
    assert [g[2] for g in got] == [c[2] for c in allc]
    assert fclo.shape_collide(s1, I, s2, tf2, cap=0) == []
