"""The product's host BVH builder (libfclx.so, no device needed) against the
oracle's restatement of BVHModel::endModel, plus invariants of the built
hierarchy (test/test_fcl_bvh_models.cpp checks build states and counts)."""
import numpy as np
import pytest

import fcl_amd as F
import fclo
import util


def _meshes():
    out = {"ico11": F.icosphere(11, 0.15, 0xA0), "ico3": F.icosphere(3, 0.15, 5),
           "hf17": F.heightfield(17, 0.05, 9)}
    out["env"] = util.load_obj("env.obj")
    out["rob"] = util.load_obj("rob.obj")
    return out


MESHES = _meshes()


@pytest.mark.parametrize("name", sorted(MESHES))
@pytest.mark.parametrize("bv_type", [F.OBBRSS, F.RSS])
def test_builder_bit_exact_vs_oracle(name, bv_type):
    V, T = MESHES[name]
    prod = F.bvh_build(V, T, bv_type)
    bv, links, sizes = fclo.OracleMesh(V, T, bv_type).nodes()
    assert np.array_equal(prod["links"], links)
    assert np.array_equal(prod["bv"].view(np.uint64), bv.view(np.uint64))
    assert np.array_equal(prod["sizes"].view(np.uint64), sizes.view(np.uint64))


@pytest.mark.parametrize("name", sorted(MESHES))
def test_hierarchy_invariants(name):
    V, T = MESHES[name]
    b = F.bvh_build(V, T, F.OBBRSS)
    n = T.shape[0]
    links = b["links"]
    assert links.shape[0] == 2 * n - 1
    leaves = links[links[:, 0] < 0, 0]
    assert sorted((-(leaves + 1)).tolist()) == list(range(n))  # each triangle once
    # children of node i are first_child and first_child+1, each with the
    # parent's primitive range split in two
    for i in range(links.shape[0]):
        fc, fp, npr = links[i]
        if fc >= 0:
            assert links[fc, 1] == fp and links[fc + 1, 1] == fp + links[fc, 2]
            assert links[fc, 2] + links[fc + 1, 2] == npr
        else:
            assert npr == 1
    # containment: every vertex of every triangle under a node lies in its
    # OBB and within r of its RSS rectangle (tolerance: rounding only)
    scale = np.abs(V).max()
    tol = 1e-12 * max(scale, 1.0)
    rng = np.random.default_rng(3)
    order = []  # primitive order of the build, reconstructed from leaves
    stack = [0]
    while stack:
        i = stack.pop()
        fc = links[i, 0]
        if fc < 0:
            order.append(-(fc + 1))
        else:
            stack += [fc + 1, fc]
    order = np.array(order)
    for i in rng.choice(links.shape[0], size=min(300, links.shape[0]), replace=False):
        fp, npr = links[i, 1], links[i, 2]
        pts = V[T[order[fp:fp + npr]].reshape(-1)]
        ax = b["bv"][i, 0:9].reshape(3, 3)
        c = b["bv"][i, 9:12]
        e = b["bv"][i, 12:15]
        loc = (pts - c) @ ax
        assert np.all(np.abs(loc) <= e + tol)
        rax = b["bv"][i, 15:24].reshape(3, 3)
        to = b["bv"][i, 24:27]
        l = b["bv"][i, 27:29]
        r = b["bv"][i, 29]
        rl = (pts - to) @ rax
        dx = np.maximum(0, np.maximum(-rl[:, 0], rl[:, 0] - l[0]))
        dy = np.maximum(0, np.maximum(-rl[:, 1], rl[:, 1] - l[1]))
        dist = np.sqrt(dx * dx + dy * dy + rl[:, 2] ** 2)
        assert np.all(dist <= r + 1e-9 * max(scale, 1.0))


def test_build_errors():
    from fcl_amd import _native as N
    with pytest.raises(F.FclxError) as e:
        F.bvh_build(np.zeros((0, 3)), np.zeros((0, 3), np.int32))
    assert e.value.code == N.FCLX_ERR_EMPTY_MODEL
    with pytest.raises(F.FclxError) as e:
        F.bvh_build(np.zeros((3, 3)), np.array([[0, 1, 5]], np.int32))
    assert e.value.code == N.FCLX_ERR_INVALID_ARGUMENT


def test_single_triangle_model():
    V = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0]], float)
    b = F.bvh_build(V, np.array([[0, 1, 2]], np.int32))
    assert b["links"].tolist() == [[-1, 0, 1]] and b["depth"] == 1
