"""Pins the oracle's primitive tests (OBB SAT, triangle SAT, RSS rectangle
distance, triangle distance, Jacobi eigen-solver) against FCL's own known
answers and against independent brute force / LP / eigh computations."""
import numpy as np
import pytest
from scipy.optimize import linprog

import fclo
import util

EPS78 = np.finfo(np.float64).eps ** (7.0 / 8.0)  # constants<double>::eps_78()


# ------------------------------------------------------------ RSS distance
def test_rss_distance_kats():
    """test/test_fcl_math.cpp:263-290 (test_rss_distance)."""
    I = np.eye(3)
    # two rectangles in the planes y = 1 and y = -1 spanning x in [0,1],
    # z in [-1,1]: the hand fit of the reference's first case (distance 2)
    ax = np.array([[1, 0, 0], [0, 0, -1], [0, 1, 0]], float)  # cols: x, z, -y... right-handed
    ax[:, 2] = np.cross(ax[:, 0], ax[:, 1])
    d = fclo.rss_distance_nodes(ax, [0, 1, -1], [1, 2], 0, ax, [0, -1, -1], [1, 2], 0)
    assert abs(d - 2.0) <= EPS78
    # rss.To=(0,0,0.5) axis=I l=(1,1) r=0.5 ; rss2.To=(-1,-1,2.5) -> 1.0
    d = fclo.rss_distance_nodes(I, [0, 0, 0.5], [1, 1], 0.5, I, [-1, -1, 2.5], [1, 1], 0.5)
    assert abs(d - 1.0) <= EPS78
    # rss2.axis = -I -> sqrt(6) - 1
    d = fclo.rss_distance_nodes(I, [0, 0, 0.5], [1, 1], 0.5, -I, [-1, -1, 2.5], [1, 1], 0.5)
    assert abs(d - (np.sqrt(6.0) - 1.0)) <= EPS78
    # rss2.To = (0,0,2.5) -> 1.0
    d = fclo.rss_distance_nodes(I, [0, 0, 0.5], [1, 1], 0.5, -I, [0, 0, 2.5], [1, 1], 0.5)
    assert abs(d - 1.0) <= EPS78


def test_rect_distance_matches_brute_force():
    """rectDistance == min distance between the two rectangles (as unions of
    triangles, by brute force), for random poses and sizes."""
    rng = np.random.default_rng(11)
    n_sep = n_zero = 0
    for _ in range(250):
        A1 = util.random_rotation(rng)
        A2 = util.random_rotation(rng)
        l1 = rng.uniform(0.05, 2.0, 2)
        l2 = rng.uniform(0.05, 2.0, 2)
        to1 = rng.uniform(-1, 1, 3)
        to2 = rng.uniform(-1, 1, 3) * rng.choice([0.3, 1.0, 3.0])
        d = fclo.rss_distance_nodes(A1, to1, l1, 0, A2, to2, l2, 0)
        bf = util.rect_rect_dist(util.rect_corners(A1, to1, l1), util.rect_corners(A2, to2, l2))
        assert abs(d - bf) <= 1e-9 * max(1.0, bf), (d, bf)
        n_sep += bf > 1e-6
        n_zero += bf <= 1e-12
    assert n_sep > 50 and n_zero > 10


def test_rss_radius_and_clamp():
    """distance = max(0, rect distance - r1 - r2) (RSS-inl.h:1956-1974)."""
    I = np.eye(3)
    d = fclo.rss_distance_nodes(I, [0, 0, 0], [1, 1], 0.25, I, [0, 0, 2], [1, 1], 0.5)
    assert d == pytest.approx(2.0 - 0.75, abs=1e-15)
    d = fclo.rss_distance_nodes(I, [0, 0, 0], [1, 1], 1.5, I, [0, 0, 2], [1, 1], 0.5)
    assert d == 0.0


# --------------------------------------------------------- triangle distance
def test_tri_distance_matches_brute_force():
    rng = np.random.default_rng(12)
    n_sep = n_ov = 0
    for i in range(400):
        S = rng.normal(size=(3, 3))
        T = rng.normal(size=(3, 3)) + rng.normal(size=3) * (0.3 if i % 2 else 2.0)
        d, P, Q = fclo.tri_distance(S, T)
        bf = util.tri_tri_dist(S, T)
        assert abs(d - bf) <= 1e-12 * max(1.0, bf), (d, bf)
        if bf > 1e-9:
            n_sep += 1
            # closest points lie on their triangles and realise the distance
            assert np.linalg.norm(P - Q) == pytest.approx(d, rel=1e-12, abs=1e-14)
            assert util.point_tri_dist(P, *S) <= 1e-12
            assert util.point_tri_dist(Q, *T) <= 1e-12
        else:
            n_ov += 1
    assert n_sep > 100 and n_ov > 20


def test_tri_distance_kats():
    S = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0]], float)
    # parallel triangle 2 above: distance 2
    d, P, Q = fclo.tri_distance(S, S + [0, 0, 2])
    assert d == 2.0
    # vertex of T above the interior of S
    T = np.array([[0.2, 0.2, 0.5], [0.2, 0.2, 3], [1, 1, 3]], float)
    d, P, Q = fclo.tri_distance(S, T)
    assert d == pytest.approx(0.5, abs=1e-15)
    assert np.allclose(P, [0.2, 0.2, 0]) and np.allclose(Q, [0.2, 0.2, 0.5])
    # crossing triangles: 0
    T = np.array([[0.2, 0.2, -1], [0.2, 0.2, 1], [0.3, -1, 0]], float)
    assert fclo.tri_distance(S, T)[0] == 0.0


# ------------------------------------------------------------ triangle SAT
def _tri(*rows):
    return np.array(rows, float)


@pytest.mark.parametrize("P,Q,expect", [
    # shared vertex
    (_tri([0, 0, 0], [1, 0, 0], [0, 1, 0]), _tri([0, 0, 0], [-1, 0, 1], [0, -1, 1]), 1),
    # shared edge, different planes
    (_tri([0, 0, 0], [1, 0, 0], [0, 1, 0]), _tri([0, 0, 0], [1, 0, 0], [0, 0, 1]), 1),
    # vertex of Q touching the interior of P
    (_tri([0, 0, 0], [2, 0, 0], [0, 2, 0]), _tri([0.5, 0.5, 0], [0.5, 0.5, 1], [1, 1, 1]), 1),
    # coplanar overlapping
    (_tri([0, 0, 0], [2, 0, 0], [0, 2, 0]), _tri([0.5, 0.5, 0], [3, 0.5, 0], [0.5, 3, 0]), 1),
    # coplanar, separated by a gap along an in-plane edge normal
    (_tri([0, 0, 0], [1, 0, 0], [0, 1, 0]), _tri([1, 1, 0], [2, 1, 0], [1, 2, 0]), 0),
    # edge-edge crossing at one point (non-coplanar)
    (_tri([-1, 0, 0], [1, 0, 0], [0, 0, -1]), _tri([0, -1, 0], [0, 1, 0], [0, 0, 1]), 1),
    # parallel planes
    (_tri([0, 0, 0], [1, 0, 0], [0, 1, 0]), _tri([0, 0, 1e-9], [1, 0, 1e-9], [0, 1, 1e-9]), 0),
    # piercing
    (_tri([0, 0, 0], [4, 0, 0], [0, 4, 0]), _tri([1, 1, -1], [1, 1, 1], [2, 5, 0]), 1),
])
def test_triangle_sat_kats(P, Q, expect):
    assert fclo.intersect_triangle(P, Q) == expect
    assert fclo.intersect_triangle(Q, P) == expect
    # integer translations keep every value exact: the answer is invariant
    sh = np.array([3.0, -5.0, 7.0])
    assert fclo.intersect_triangle(P + sh, Q + sh) == expect


def test_triangle_sat_against_independent_predicates():
    """Disjoint by a clear distance -> 0; an edge crossing the other
    triangle's interior unambiguously -> 1."""
    rng = np.random.default_rng(13)
    n0 = n1 = 0
    for i in range(1500):
        P = rng.normal(size=(3, 3))
        Q = rng.normal(size=(3, 3)) + rng.normal(size=3) * 0.8
        R = util.random_rotation(rng) if i % 3 == 0 else np.eye(3)
        T = rng.normal(size=3) * 0.2 if i % 3 == 0 else np.zeros(3)
        got = fclo.intersect_triangle(P, Q, R, T)
        Qw = Q @ R.T + T
        if util.tri_tri_dist(P, Qw) > 1e-9:
            assert got == 0
            n0 += 1
        elif any(util.seg_crosses_tri(Qw[k], Qw[(k + 1) % 3], *P, 1e-9) for k in range(3)) or \
                any(util.seg_crosses_tri(P[k], P[(k + 1) % 3], *Qw, 1e-9) for k in range(3)):
            assert got == 1
            n1 += 1
    assert n0 > 250 and n1 > 150


# ------------------------------------------------------------------ OBB SAT
def _boxes_overlap_lp(c1, A1, e1, c2, A2, e2):
    """Feasibility of c1 + A1 u = c2 + A2 v with |u| <= e1, |v| <= e2."""
    Aeq = np.hstack([A1, -A2])
    beq = c2 - c1
    bounds = [(-e, e) for e in e1] + [(-e, e) for e in e2]
    r = linprog(np.zeros(6), A_eq=Aeq, b_eq=beq, bounds=bounds, method="highs")
    return r.status == 0


def test_obb_disjoint_matches_lp():
    """OBB overlap agrees with exact box-box overlap (the property
    test/test_fcl_collision.cpp:151-191 checks against box-box collide):
    boxes that still overlap when shrunk by 1e-4 are reported overlapping,
    boxes still apart when grown by 1e-4 are reported disjoint."""
    rng = np.random.default_rng(14)
    n_ov = n_dis = 0
    for _ in range(600):
        A1 = util.random_rotation(rng)
        A2 = util.random_rotation(rng)
        e1 = rng.uniform(0.1, 1.0, 3)
        e2 = rng.uniform(0.1, 1.0, 3)
        c1 = np.zeros(3)
        c2 = rng.uniform(-1.6, 1.6, 3)
        B = A1.T @ A2
        T = A1.T @ (c2 - c1)
        got_disjoint = fclo.obb_disjoint(B, T, e1, e2)
        if _boxes_overlap_lp(c1, A1, e1 * (1 - 1e-4), c2, A2, e2 * (1 - 1e-4)):
            assert got_disjoint == 0
            n_ov += 1
        elif not _boxes_overlap_lp(c1, A1, e1 * (1 + 1e-4), c2, A2, e2 * (1 + 1e-4)):
            assert got_disjoint == 1
            n_dis += 1
    assert n_ov > 100 and n_dis > 100


def test_obb_touching_counts_as_overlap():
    I = np.eye(3)
    e = np.array([1.0, 1.0, 1.0])
    # face-to-face contact: centres 2 apart along x
    assert fclo.obb_disjoint(I, np.array([2.0, 0, 0]), e, e) == 0
    # the 1e-6 inflation of |R| keeps a gap below ~1e-6 overlapping
    assert fclo.obb_disjoint(I, np.array([2.0 + 1e-7, 0, 0]), e, e) == 0
    assert fclo.obb_disjoint(I, np.array([2.0 + 1e-3, 0, 0]), e, e) == 1


# ------------------------------------------------------------ eigen solver
def test_eigen_old_matches_eigh():
    rng = np.random.default_rng(15)
    for _ in range(200):
        A = rng.normal(size=(3, 3))
        M = A @ A.T
        d, V = fclo.eigen_old(M)
        w, U = np.linalg.eigh(M)
        assert np.allclose(np.sort(d), w, rtol=1e-12, atol=1e-12 * np.abs(w).max())
        # row k of the output is the eigenvector of d[k] (axisFromEigen reads
        # eigenV.row(max), math/geometry-inl.h:594-596)
        for k in range(3):
            assert np.allclose(M @ V[k], d[k] * V[k], atol=1e-10 * np.abs(w).max())
        assert np.allclose(V.T @ V, np.eye(3), atol=1e-12)


def test_eigen_old_diagonal_is_exact():
    d, V = fclo.eigen_old(np.diag([3.0, 1.0, 2.0]))
    assert list(d) == [3.0, 1.0, 2.0]
    assert np.array_equal(V, np.eye(3))


# ------------------------------------------------ triangle contact output
def _plane(T):
    n = np.cross(T[1] - T[0], T[2] - T[0])
    n = n / np.linalg.norm(n)
    return n, n @ T[0]


def test_triangle_contact_piercing_vertex_closed_form():
    """A triangle Q dipping one vertex a depth d below triangle P's plane
    (its other vertices far above, so P barely penetrates Q's plane): one
    contact at that vertex, depth d, normal = P's unit normal.  The same
    holds in any rigid frame (intersect-inl.h:802-842)."""
    rng = np.random.default_rng(41)
    P0 = np.array([[0, 0, 0], [2, 0, 0], [0, 2, 0.0]])
    for _ in range(200):
        d = rng.uniform(0.01, 0.4)
        x, y = rng.uniform(0.2, 0.8, 2)
        Q0 = np.array([[x, y, -d], [x + 0.1, y, 1.5], [x, y + 0.1, 1.5]])
        Rm = util.random_rotation(rng)
        t = rng.uniform(-3, 3, 3)
        P, Q = P0 @ Rm.T + t, Q0 @ Rm.T + t
        hit, pts, normal, depth = fclo.intersect_triangle_contact(P, Q)
        assert hit and len(pts) == 1
        assert abs(depth - d) < 1e-12
        np.testing.assert_allclose(pts[0], Q[0], atol=1e-12)
        np.testing.assert_allclose(normal, Rm @ np.array([0, 0, 1.0]), atol=1e-12)


def test_triangle_contact_edge_gives_two_points():
    """An edge of Q parallel to P's plane at depth d: two contacts (both
    edge vertices) at that depth."""
    P = np.array([[0, 0, 0], [3, 0, 0], [0, 3, 0.0]])
    Q = np.array([[0.5, 0.5, -0.2], [1.0, 0.6, -0.2], [0.7, 0.5, 2.0]])
    hit, pts, normal, depth = fclo.intersect_triangle_contact(P, Q)
    assert hit and len(pts) == 2 and abs(depth - 0.2) < 1e-12
    np.testing.assert_allclose(np.sort(pts[:, 0]), [0.5, 1.0])
    np.testing.assert_allclose(normal, [0, 0, 1.0])


def _intersecting_pairs(rng, n):
    out = []
    while len(out) < n:
        P = rng.uniform(-1, 1, (3, 3))
        Q = rng.uniform(-1, 1, (3, 3))
        if fclo.intersect_triangle(P, Q):
            out.append((P, Q))
    return out


def test_triangle_contact_points_are_deepest_vertices():
    """On random intersecting pairs: every contact point is a vertex of one
    triangle lying at the reported depth below the other's plane, no vertex
    of that triangle lies deeper, and the normal is the other's unit normal
    (sign: Q's points use +n_P, P's points use -n_Q).  Planes and depths are
    recomputed here with numpy."""
    rng = np.random.default_rng(43)
    seen = 0
    for P, Q in _intersecting_pairs(rng, 400):
        hit, pts, normal, depth = fclo.intersect_triangle_contact(P, Q)
        assert hit
        if len(pts) == 0:
            continue
        seen += 1
        nP, tP = _plane(P)
        nQ, tQ = _plane(Q)
        if np.allclose(normal, nP, atol=1e-9):
            tri, n, t = Q, nP, tP
        else:
            np.testing.assert_allclose(normal, -nQ, atol=1e-9)
            tri, n, t = P, nQ, tQ
        d_all = -(tri @ n - t)
        assert abs(d_all.max() - depth) < 1e-9
        for p in pts:
            assert min(np.abs(tri - p).sum(1)) < 1e-12
            assert abs(-(p @ n - t) - depth) < 1e-6 + 1e-9
        assert len(pts) <= 2
    assert seen > 300


def test_triangle_contact_swap_symmetry():
    """Swapping the triangles keeps the points and depth and flips the
    normal, whenever the two penetrations differ."""
    rng = np.random.default_rng(47)
    checked = 0
    for P, Q in _intersecting_pairs(rng, 300):
        _, pa, na, da = fclo.intersect_triangle_contact(P, Q)
        _, pb, nb, db = fclo.intersect_triangle_contact(Q, P)
        nP, tP = _plane(P)
        nQ, tQ = _plane(Q)
        pen_p = (-(P @ nQ - tQ)).max()
        pen_q = (-(Q @ nP - tP)).max()
        if abs(pen_p - pen_q) < 1e-9:
            continue
        checked += 1
        assert da == db
        np.testing.assert_array_equal(pa, pb)
        np.testing.assert_allclose(na, -nb, atol=1e-15)
    assert checked > 250
