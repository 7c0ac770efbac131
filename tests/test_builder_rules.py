"""Pins the builder's rules (the oracle's restatement of BVHModel::endModel
and the product builder, which equals it bit for bit) with checks that a
plausible mistake fails:

- Each node's OBB axis 0 is the top eigenvector of the scatter matrix of
  its triangles' vertices (getCovariance + eigen_old + axisFromEigen,
  geometry-inl.h:475-598, 1333-1421), compared with numpy.linalg.eigh.
- Each internal node's children split its triangles at the mean centroid
  projection on that axis (computeSplitValue_mean, BV_splitter-inl.h:
  539-599; partition in recursiveBuildTree, BVH_model-inl.h:867-938): the
  left child holds exactly the centroids at or below the mean, the right
  child those above; a one-sided split falls back to halves.
- OBB extents are tight: on each axis some vertex touches each face
  (getExtentAndCenter_mesh, geometry-inl.h:292-362).
A small numpy builder that follows these rules passes the same checks, and
its variants with a median split or the second axis fail them."""
import numpy as np
import pytest

import fcl_amd as F
import fclo
import util


def _order(links):
    """Triangles in build order, from the pre-order leaves."""
    order, stack = [], [0]
    while stack:
        i = stack.pop()
        fc = links[i, 0]
        if fc < 0:
            order.append(-(fc + 1))
        else:
            stack += [fc + 1, fc]
    return np.array(order)


def violations(V, T, bv, links, gap=1e-6, tie=1e-9, any_sign=False):
    """Rule violations of a built hierarchy (bv: nodes x >= 15, OBB axis
    row-major, centre, extent; links: first_child, first_primitive,
    num_primitives).  Returns a list of strings (empty when every rule
    holds).  any_sign: the left child may hold the centroids above the mean
    instead (a refit's eigenvector can point the other way than the
    build's, which fixed the order of the children)."""
    out = []
    order = _order(links)
    scale = max(np.abs(V).max(), 1.0)
    cent = V[T].sum(axis=1)  # p1 + p2 + p3 per triangle
    for i in range(links.shape[0]):
        fc, fp, npr = links[i]
        prims = order[fp:fp + npr]
        pts = V[T[prims]].reshape(-1, 3)
        ax = bv[i, 0:9].reshape(3, 3)
        c, e = bv[i, 9:12], bv[i, 12:15]
        loc = (pts - c) @ ax
        hi, lo = loc.max(axis=0), loc.min(axis=0)
        if np.any(np.abs(hi - e) > 1e-9 * scale) or np.any(np.abs(lo + e) > 1e-9 * scale):
            out.append("node %d: extents not tight" % i)
        if fc < 0:
            continue
        S1 = pts.sum(axis=0)
        M = pts.T @ pts - np.outer(S1, S1) / len(pts)
        w, U = np.linalg.eigh(M)
        if w[2] - w[1] > gap * max(abs(w[2]), 1e-300):
            if abs(abs(U[:, 2] @ ax[:, 0]) - 1) > 1e-7:
                out.append("node %d: axis 0 is not the top eigenvector" % i)
        proj = cent[prims] @ ax[:, 0] / 3.0
        split = (cent[prims].sum(axis=0) @ ax[:, 0]) / (3 * npr)
        nleft = links[fc, 2]
        left = set(order[fp:fp + nleft].tolist())
        below = proj <= split
        if np.any(np.abs(proj - split) <= tie * scale):
            continue  # a centroid on the split: rounding decides its side
        if below.all() or not below.any():
            if nleft != npr // 2:
                out.append("node %d: one-sided split not halved" % i)
            continue
        want = set(prims[below].tolist())
        if left != want and not (any_sign and left == set(prims[~below].tolist())):
            out.append("node %d: children do not split at the mean" % i)
    return out


def py_build(V, T, split="mean", axis=0):
    """A plain numpy builder with the rules above (eigh for the axes), for
    checking the checker; split "median" or axis 1 plant a mistake."""
    n = len(T)
    bv = np.zeros((2 * n - 1, 15))
    links = np.zeros((2 * n - 1, 3), np.int64)
    prims = np.arange(n)
    cent = V[T].sum(axis=1)
    count = [1]

    def rec(node, first, num):
        idx = prims[first:first + num]
        pts = V[T[idx]].reshape(-1, 3)
        S1 = pts.sum(axis=0)
        w, U = np.linalg.eigh(pts.T @ pts - np.outer(S1, S1) / len(pts))
        ax = np.stack([U[:, 2], U[:, 1], np.cross(U[:, 2], U[:, 1])], 1)
        loc = pts @ ax
        lo, hi = loc.min(axis=0), loc.max(axis=0)
        bv[node, 0:9] = ax.reshape(9)
        bv[node, 9:12] = ax @ ((lo + hi) / 2)
        bv[node, 12:15] = (hi - lo) / 2
        links[node, 1:] = first, num
        if num == 1:
            links[node, 0] = -(idx[0] + 1)
            return
        links[node, 0] = count[0]
        child = count[0]
        count[0] += 2
        p = cent[idx] @ ax[:, axis] / 3.0
        s = np.median(p) if split == "median" else (cent[idx].sum(axis=0) @ ax[:, axis]) / (3 * num)
        left = idx[p <= s]
        right = idx[p > s]
        if len(left) in (0, num):
            left, right = idx[: num // 2], idx[num // 2:]
        prims[first:first + num] = np.concatenate([left, right])
        rec(child, first, len(left))
        rec(child + 1, first + len(left), num - len(left))

    rec(0, 0, n)
    return bv, links


MESHES = {"ico5": F.icosphere(5, 0.15, 3), "hf12": F.heightfield(12, 0.05, 4),
          "rob": util.load_obj("rob.obj")}


@pytest.mark.parametrize("name", sorted(MESHES))
def test_oracle_builder_follows_the_rules(name):
    V, T = MESHES[name]
    bv, links, _ = fclo.OracleMesh(V, T).nodes()
    assert violations(V, T, bv, links) == []


@pytest.mark.parametrize("name", sorted(MESHES))
def test_product_builder_follows_the_rules(name):
    V, T = MESHES[name]
    b = F.bvh_build(V, T)
    assert violations(V, T, b["bv"], b["links"]) == []


@pytest.mark.parametrize("name", sorted(MESHES))
def test_the_checks_accept_a_correct_build_and_reject_planted_mistakes(name):
    V, T = MESHES[name]
    bv, links = py_build(V, T)
    assert violations(V, T, bv, links) == []
    bad = violations(V, T, *py_build(V, T, split="median"))
    assert any("mean" in v for v in bad), "a median split must be caught"
    bad = violations(V, T, *py_build(V, T, axis=1))
    assert any("mean" in v for v in bad), "a split on the second axis must be caught"
    # an OBB that is not tight, or axes given in another order
    loose = bv.copy()
    loose[0, 12] *= 1.01
    assert any("tight" in v for v in violations(V, T, loose, links))
    swapped = bv.copy()
    ax = swapped[:, 0:9].reshape(-1, 3, 3)
    ax[:, :, [0, 1]] = ax[:, :, [1, 0]]
    swapped[:, 12:15] = swapped[:, [13, 12, 14]]
    assert any("eigenvector" in v for v in violations(V, T, swapped, links))
