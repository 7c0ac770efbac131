"""Pins the oracle's traversals: the collision traversal must find exactly
the intersecting triangle pairs a brute-force all-pairs scan finds (the
invariant test/test_fcl_collision.cpp:313-888 asserts across BV types), in
an order where every capped result is a prefix of the exhaustive one; the
distance traversal must reach the brute-force minimum and agree across BV
types (test/test_fcl_distance.cpp:86-304)."""
import numpy as np
import pytest

import fcl_amd as F
import fclo
import util


@pytest.fixture(scope="module")
def spheres():
    V1, T1 = F.icosphere(5, 0.15, 1)
    V2, T2 = F.icosphere(4, 0.15, 2)
    return (V1, T1), (V2, T2)


@pytest.fixture(scope="module")
def env_rob():
    return util.load_obj("env.obj"), util.load_obj("rob.obj")


def _sorted(p):
    p = np.asarray(p).reshape(-1, 2)
    return p[np.lexsort((p[:, 1], p[:, 0]))]


def test_collide_exhaustive_equals_brute_force_spheres(spheres):
    (V1, T1), (V2, T2) = spheres
    m1, m2 = fclo.OracleMesh(V1, T1), fclo.OracleMesh(V2, T2)
    tf1, tf2 = F.poses(60, 0, 1.6, 0, 77)
    hit = 0
    for q in range(60):
        got, _ = fclo.collide(m1, tf1[q], m2, tf2[q], cap=100000)
        bf = fclo.collide_all_pairs(m1, tf1[q], m2, tf2[q])
        assert np.array_equal(_sorted(got), _sorted(bf))
        hit += len(bf) > 0
    assert hit > 10


def test_collide_env_rob_matches_brute_force(env_rob):
    """env.obj vs rob.obj with poses drawn like the reference's mesh_mesh test
    (extents [-3000,3000]^2 x [0,3000], Euler angles in [0,2pi))."""
    (V1, T1), (V2, T2) = env_rob
    m1, m2 = fclo.OracleMesh(V1, T1), fclo.OracleMesh(V2, T2)
    tfs = util.random_transforms(10, [-3000, -3000, 0, 3000, 3000, 3000], 2024)
    I = util.identity_pose()
    total = 0
    for q in range(10):
        got, _ = fclo.collide(m1, tfs[q], m2, I, cap=1 << 30)
        bf = fclo.collide_all_pairs(m1, tfs[q], m2, I)
        assert np.array_equal(_sorted(got), _sorted(bf))
        total += len(bf)
    assert total > 0


@pytest.mark.parametrize("cap", [1, 2, 5, 17])
def test_capped_result_is_prefix_of_exhaustive(spheres, cap):
    """With num_max_contacts = cap the reference stops after the cap-th
    contact in depth-first order: the capped list is the exhaustive list's
    prefix, and it does no more BV tests than the exhaustive run."""
    (V1, T1), (V2, T2) = spheres
    m1, m2 = fclo.OracleMesh(V1, T1), fclo.OracleMesh(V2, T2)
    tf1, tf2 = F.poses(40, 0, 1.5, 0, 78)
    for q in range(40):
        full, (bv_full, _) = fclo.collide(m1, tf1[q], m2, tf2[q], cap=100000)
        part, (bv_part, _) = fclo.collide(m1, tf1[q], m2, tf2[q], cap=cap)
        assert np.array_equal(part, full[: min(cap, len(full))])
        assert bv_part <= bv_full
        if len(full) <= cap:
            assert bv_part == bv_full


def test_collide_statistics_shape(spheres):
    """Every leaf test follows a passing BV test; disjoint poses stop at the
    root test."""
    (V1, T1), (V2, T2) = spheres
    m1, m2 = fclo.OracleMesh(V1, T1), fclo.OracleMesh(V2, T2)
    far = util.identity_pose()
    far[9:] = [10, 0, 0]
    pairs, (bv, leaf) = fclo.collide(m1, util.identity_pose(), m2, far, cap=1)
    assert len(pairs) == 0 and bv == 1 and leaf == 0
    tf1, tf2 = F.poses(20, 0, 1.0, 0, 79)
    for q in range(20):
        _, (bv, leaf) = fclo.collide(m1, tf1[q], m2, tf2[q], cap=100000)
        assert leaf <= bv


def test_distance_equals_brute_force(spheres):
    (V1, T1), (V2, T2) = spheres
    for bv_type in (fclo.RSS, fclo.OBBRSS):
        m1, m2 = fclo.OracleMesh(V1, T1, bv_type), fclo.OracleMesh(V2, T2, bv_type)
        tf1, tf2 = F.poses(25, 1, 2.1, 4.0, 80)
        for q in range(25):
            d, pts, ids, (bv, leaf) = fclo.distance(m1, tf1[q], m2, tf2[q])
            bf, bf_ids = fclo.distance_all_pairs(m1, tf1[q], m2, tf2[q])
            assert d == bf
            # the reported pair realises the minimum and the points are that far apart
            assert np.linalg.norm(pts[0] - pts[1]) == pytest.approx(d, rel=1e-12)
            assert leaf < m1.m.num_tris * m2.m.num_tris  # pruning happened


def test_distance_consistent_across_bv_types(env_rob):
    """RSS and OBBRSS hierarchies give the same distance (they traverse with
    different size rules, so this also checks that pruning never loses the
    minimum), like test_fcl_distance.cpp's cross-BV comparison."""
    (V1, T1), (V2, T2) = env_rob
    r1, r2 = fclo.OracleMesh(V1, T1, fclo.RSS), fclo.OracleMesh(V2, T2, fclo.RSS)
    o1, o2 = fclo.OracleMesh(V1, T1, fclo.OBBRSS), fclo.OracleMesh(V2, T2, fclo.OBBRSS)
    tfs = util.random_transforms(10, [-3000, -3000, 0, 3000, 3000, 3000], 2025)
    I = util.identity_pose()
    for q in range(10):
        dr = fclo.distance(r1, tfs[q], r2, I)[0]
        do = fclo.distance(o1, tfs[q], o2, I)[0]
        assert dr == do
        bf, _ = fclo.distance_all_pairs(r1, tfs[q], r2, I)
        assert dr == bf


def test_distance_seed_and_overlap(spheres):
    """Overlapping meshes give 0; the traversal is seeded with triangle 0 vs
    triangle 0 so a pose where that pair is the unique minimum reports it."""
    (V1, T1), (V2, T2) = spheres
    m1, m2 = fclo.OracleMesh(V1, T1), fclo.OracleMesh(V2, T2)
    d, _, _, _ = fclo.distance(m1, util.identity_pose(), m2, util.identity_pose())
    assert d == 0.0


@pytest.mark.parametrize("cap", [1, 3, 8, 100000])
def test_contact_mode_matches_pairwise_recomposition(spheres, cap):
    """collide() with enable_contact (mesh_collision_traversal_node-inl.h:
    578-598): rebuilt here from the exhaustive pair list in depth-first order
    and the per-pair contact function, mapped to the world frame with numpy;
    the traversal's output is the first `cap` contacts of that list (the cap
    counts contacts, and a pair crossing it adds only what fits)."""
    (V1, T1), (V2, T2) = spheres
    m1, m2 = fclo.OracleMesh(V1, T1), fclo.OracleMesh(V2, T2)
    tf1, tf2 = F.poses(30, 0, 1.5, 0, 79)
    R1s = tf1[:, :9].reshape(-1, 3, 3)
    total = 0
    for q in range(30):
        pairs, _ = fclo.collide(m1, tf1[q], m2, tf2[q], cap=100000)
        R1 = R1s[q]
        R2 = tf2[q][:9].reshape(3, 3)
        R = R1.T @ R2
        T = R1.T @ (tf2[q][9:] - tf1[q][9:])
        want = []
        for a, b in pairs:
            P = V1[T1[a]]
            Q = V2[T2[b]]
            hit, pts, nrm, dep = fclo.intersect_triangle_contact(P, Q, R, T)
            assert hit
            for p in pts:
                want.append((a, b, R1 @ p + tf1[q][9:], R1 @ nrm, dep))
        gp, gpos, gnrm, gdep, _, n = fclo.collide_contact(m1, tf1[q], m2, tf2[q], cap=cap,
                                                          store=min(cap, 4096))
        want = want[:cap]
        assert n == len(want)
        total += n
        for i, (a, b, p, nv, d) in enumerate(want):
            assert (gp[i] == (a, b)).all()
            np.testing.assert_allclose(gpos[i], p, atol=1e-12)
            np.testing.assert_allclose(gnrm[i], nv, atol=1e-12)
            assert gdep[i] == d
    assert total > 0
