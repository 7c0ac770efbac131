"""GPU parity of fclx_mesh_shape_collide_batch (BVHModel<OBBRSS> against a
primitive shape) with the oracle: counts, the contacts' triangles, the
reference's BV / leaf test counts and, with enable_contact, every contact
point, normal and depth bit for bit."""
import numpy as np
import pytest
import torch

import fcl_amd as F
import fclo
import util
from fcl_amd import BOX, CAPSULE, CONVEX, CYLINDER, SPHERE, icosphere, scale_to_radius, \
    shape_table

pytestmark = pytest.mark.gpu

DEV = torch.device("cuda:0")

HULL = np.array([[1, 0, 0], [-1, 0, 0], [0, 1.5, 0], [0, -1, 0], [0, 0, 0.7], [0.2, 0.3, -1.1],
                 [0.5, 0.5, 0.5]], np.float64) * 0.4
KINDS = [BOX, SPHERE, CAPSULE, CYLINDER, CONVEX]
DIMS = [(0.6, 0.9, 0.4), (0.45, 0, 0), (0.25, 0.8, 0), (0.3, 0.7, 0), (0, 0, 0)]


def _t(a, dtype):
    return torch.from_numpy(np.ascontiguousarray(a)).to(DEV, dtype)


def ms_gpu(ctx, qmesh, rec, pool, qshape, tf1, tf2, cap, contact):
    n = len(qshape)
    d_rec = torch.from_numpy(np.ascontiguousarray(rec).view(np.uint8)).to(DEV)
    d_pool = _t(np.asarray(pool, np.float64).reshape(-1, 3), torch.float64)
    counts = torch.full((n,), -1, dtype=torch.int32, device=DEV)
    prims = torch.full((max(1, n * cap),), -7, dtype=torch.int32, device=DEV)
    stats = torch.full((max(1, n), 2), -1, dtype=torch.int64, device=DEV)
    geo = None
    if contact:
        geo = (torch.full((max(1, n * cap), 3), float("nan"), dtype=torch.float64, device=DEV),
               torch.full((max(1, n * cap), 3), float("nan"), dtype=torch.float64, device=DEV),
               torch.full((max(1, n * cap),), float("nan"), dtype=torch.float64, device=DEV))
    s = torch.cuda.current_stream().cuda_stream
    ctx.mesh_shape_collide_batch_async(_t(qmesh, torch.int32), d_rec, d_pool,
                                       _t(qshape, torch.int32), _t(tf1, torch.float64),
                                       _t(tf2, torch.float64), cap, counts, primitives=prims,
                                       contact=geo, stats=stats, stream=s)
    rc = ctx.sync_status(s)
    out = np.zeros((n, cap, 7))
    if contact:
        pos, nrm, dep = (t.cpu().numpy()[: n * cap] for t in geo)
        out[:, :, 0:3] = nrm.reshape(n, cap, 3)
        out[:, :, 3:6] = pos.reshape(n, cap, 3)
        out[:, :, 6] = dep.reshape(n, cap)
    return (rc, counts.cpu().numpy(), prims.cpu().numpy()[: n * cap].reshape(n, cap), out,
            stats.cpu().numpy()[:n])


def _mesh(k, seed=0):
    V, T = icosphere(k, 0.15, seed)
    return scale_to_radius(V, 1.0), T


def _case(n, seed, near=1.2):
    tf1 = util.random_transforms(n, [-0.3, -0.3, -0.3, 0.3, 0.3, 0.3], seed=seed)
    tf2 = util.random_transforms(n, [-near, -near, -near, near, near, near], seed=seed + 1)
    return tf1, tf2


def check(ctx, meshes, mids, qm, rec, pool, qs, tf1, tf2, cap, contact, label):
    print("case: %s n=%d cap=%d contact=%d" % (label, len(qs), cap, contact), flush=True)
    rc, gc, gp, gg, gs = ms_gpu(ctx, np.asarray(mids)[qm], rec, pool, qs, tf1, tf2, cap, contact)
    assert rc == 0
    oc, op, og, os_ = fclo.mesh_shape_collide_batch(meshes, qm, rec, pool, qs, tf1, tf2, cap=cap,
                                                    enable_contact=contact)
    assert np.array_equal(gc, oc), label
    assert np.array_equal(gs, os_), label
    for q in range(len(qs)):
        k = oc[q]
        assert np.array_equal(gp[q, :k], op[q, :k]), (label, q)
        if contact:
            assert np.array_equal(gg[q, :k], og[q, :k]), (label, q, gg[q, :k], og[q, :k])
    return oc


@pytest.mark.parametrize("kind", KINDS)
@pytest.mark.parametrize("cap", [1, 4096])
def test_mesh_shape_parity(ctx, kind, cap):
    V, T = _mesh(12)
    mid = ctx.mesh(V, T)
    m = fclo.OracleMesh(V, T)
    k = KINDS.index(kind)
    rec, pool = shape_table([kind], [DIMS[k]], [HULL] if kind == CONVEX else [])
    n = 4500  # above the size from which batches are grouped by shape kind
    tf1, tf2 = _case(n, 100 + k)
    counts = check(ctx, [m], [mid], np.zeros(n, np.int32), rec, pool, np.zeros(n, np.int32),
                   tf1, tf2, cap, False, "kind %d" % kind)
    assert (counts > 0).sum() > n // 5
    ctx.mesh_destroy(mid)


def test_mesh_shape_contact_parity(ctx):
    meshes = [_mesh(8, 1), _mesh(14, 2)]
    mids = [ctx.mesh(V, T) for V, T in meshes]
    ms = [fclo.OracleMesh(V, T) for V, T in meshes]
    rec, pool = shape_table(KINDS, DIMS, [HULL])
    n = 4200  # mixed kinds, grouped by kind on the device
    rng = np.random.default_rng(9)
    qm = rng.integers(0, 2, n).astype(np.int32)
    qs = rng.integers(0, 5, n).astype(np.int32)
    tf1, tf2 = _case(n, 300)
    for cap in (1, 6, 64):
        counts = check(ctx, ms, mids, qm, rec, pool, qs, tf1, tf2, cap, True, "contact")
        assert counts.sum() > n // 4
    # small batches are not grouped: the same answers
    check(ctx, ms, mids, qm[:900], rec, pool, qs[:900], tf1[:900], tf2[:900], 6, True,
          "contact, ungrouped")
    for mid in mids:
        ctx.mesh_destroy(mid)


def test_mesh_shape_edge_cases(ctx):
    V, T = _mesh(6)
    mid = ctx.mesh(V, T)
    rss = ctx.mesh(V, T, F.RSS)
    m = fclo.OracleMesh(V, T)
    thin = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0]], np.float64)
    rec, pool = shape_table([SPHERE, CONVEX], [(0.5, 0, 0), (0, 0, 0)], [thin])
    tf1, tf2 = _case(6, 7, near=0.3)
    tf2[:, 9:] = [1.0, 0.0, 0.0]  # on the mesh's surface
    # empty batch
    rc, gc, _, _, _ = ms_gpu(ctx, np.zeros(0), rec, pool, np.zeros(0), tf1[:0], tf2[:0], 1, False)
    assert rc == 0 and len(gc) == 0
    # a missing mesh, an RSS mesh, a missing shape and a 3-vertex convex get
    # count 0 and the batch reports invalid arguments; valid queries answer
    qm = np.array([mid, 999, rss, mid, mid, mid], np.int32)
    qs = np.array([0, 0, 0, 5, 1, 0], np.int32)
    print("case: edge queries", flush=True)
    rc, gc, gp, _, gs = ms_gpu(ctx, qm, rec, pool, qs, tf1, tf2, 8, False)
    assert rc == F.N.FCLX_ERR_INVALID_ARGUMENT
    assert list(gc[1:5]) == [0, 0, 0, 0]
    oc, op, _, os_ = fclo.mesh_shape_collide_batch([m], np.zeros(2), rec, pool,
                                                   np.zeros(2, np.int32), tf1[[0, 5]],
                                                   tf2[[0, 5]], cap=8)
    assert gc[0] == oc[0] and gc[5] == oc[1] and oc[0] > 0
    assert np.array_equal(gp[0, :oc[0]], op[0, :oc[0]])
    assert np.array_equal(gs[[0, 5]], os_)
    # the status is cleared by the sync: a valid batch afterwards succeeds
    rc, gc, _, _, _ = ms_gpu(ctx, qm[:1], rec, pool, qs[:1], tf1[:1], tf2[:1], 8, False)
    assert rc == 0 and gc[0] == oc[0]
    ctx.mesh_destroy(mid)
    ctx.mesh_destroy(rss)


def chain_mesh(n, s=0.05):
    """Small triangles along the x axis whose positions grow so fast that
    every mean split peels off the farthest one: a hierarchy n levels deep
    (x_k = 2 ((k + 1) x_{k-1} - sum of the earlier x))."""
    xs = [0.0, 1.0]
    total = 1.0
    while len(xs) < n:
        x = 2.0 * ((len(xs) + 1) * xs[-1] - total)
        xs.append(x)
        total += x
    V = np.array([[x + dx, dy, dz] for x in xs for dx, dy, dz in
                  ((0, 0, 0), (0, s, 0), (0, 0, s))], np.float64)
    T = np.arange(3 * n, dtype=np.int32).reshape(n, 3)
    return V, T, xs


def test_mesh_shape_deep_hierarchy():
    """A mesh 72 levels deep (a chain of triangles that the mean split
    peels off one at a time): the descent stack leaves LDS for device
    memory and every count, triangle, statistic and contact equals the
    oracle's.  Own context: a deep mesh switches every mesh-shape batch of
    its context to the device-memory stack."""
    V, T, xs = chain_mesh(72)
    depth = F.bvh_build(V, T)["depth"]
    assert depth == 72, depth
    c2 = F.Context(0)
    try:
        mid = c2.mesh(V, T)
        assert c2.mesh_info(mid)["depth"] == depth
        m = fclo.OracleMesh(V, T)
        rec, pool = shape_table(KINDS, DIMS, [HULL])
        n = 600
        rng = np.random.default_rng(21)
        qs = rng.integers(0, 5, n).astype(np.int32)
        tf1 = np.tile(util.identity_pose(), (n, 1))
        tf2 = util.random_transforms(n, [-0.3, -0.3, -0.3, 0.3, 0.3, 0.3], seed=22)
        # each shape next to one of the chain's ten nearest triangles, the
        # deepest leaves of the hierarchy
        tf2[:, 9] += np.array(xs)[rng.integers(0, 10, n)]
        # the FP64 box tests of the upper levels (extents near 1e125) let few
        # queries through to the leaves, but those descend past 64 levels
        _, _, _, os_ = fclo.mesh_shape_collide_batch([m], np.zeros(n, np.int32), rec, pool, qs,
                                                     tf1, tf2, cap=1)
        assert (os_[:, 0] >= 66).sum() >= 5  # 65 or more right children pending
        for cap, contact in ((1, False), (8, False), (4, True)):
            counts = check(c2, [m], [mid], np.zeros(n, np.int32), rec, pool, qs, tf1, tf2, cap,
                           contact, "deep mesh")
            assert (counts > 0).sum() >= 5
        c2.mesh_destroy(mid)
    finally:
        c2.close()
