"""Device refit (fclx_mesh_refit: endReplaceModel with refit, top-down and
bottom-up) against the oracle's refit: every node's OBB and RSS and
BV::size() bit for bit, then collide() and distance() on the refit meshes
equal to the oracle's (counts, pairs, statistics, distances), and the
reference's front-list-with-refit flow (test_fcl_frontlist.cpp
collide_front_list_Test)."""
import numpy as np
import pytest
import torch

import fcl_amd as F
import fclo
import util

pytestmark = pytest.mark.gpu

DEV = torch.device("cuda:0")


def _t(a, dtype):
    return torch.from_numpy(np.ascontiguousarray(a)).to(DEV, dtype)


def _deform(V, seed, amp=0.03):
    rng = np.random.default_rng(seed)
    return V + rng.normal(0, amp * np.abs(V).max(), V.shape)


@pytest.mark.parametrize("bottomup", [True, False])
@pytest.mark.parametrize("bv_type", [F.OBBRSS, F.RSS])
def test_refit_nodes_bit_exact(ctx, bottomup, bv_type):
    V, T = F.icosphere(9, 0.15, 41)
    mid = ctx.mesh(V, T, bv_type)
    om = fclo.OracleMesh(V, T, bv_type)
    for round_, seed in enumerate((1, 2)):  # two refits in a row
        V2 = _deform(V, seed)
        print("case: refit bottomup=%d bv=%d round %d" % (bottomup, bv_type, round_), flush=True)
        ctx.mesh_refit(mid, _t(V2, torch.float64), bottomup=bottomup)
        om.refit(V2, bottomup=bottomup)
        got = ctx.export_nodes(mid)
        bv, links, sizes = om.nodes()
        assert np.array_equal(got["links"], links)
        assert np.array_equal(got["bv"][:, :30].view(np.uint64), bv[:, :30].view(np.uint64))
        assert np.array_equal(got["sizes"].view(np.uint64), sizes.view(np.uint64))
    ctx.mesh_destroy(mid)


@pytest.mark.parametrize("bottomup", [True, False])
def test_collide_after_refit(ctx, bottomup):
    """FAST (FP32 cull against the refit's FP32 nodes and cull bound),
    BOOL and PLAIN engines on refit meshes: counts, pair sets and
    statistics equal to the oracle's on its own refit meshes."""
    from test_gpu_parity import _set, collide_gpu
    (V1, T1), (V2, T2) = F.icosphere(8, 0.15, 51), F.heightfield(20, 0.05, 52)
    V2 = F.scale_to_radius(V2, 1.0)
    g = [ctx.mesh(V1, T1), ctx.mesh(V2, T2)]
    o = [fclo.OracleMesh(V1, T1), fclo.OracleMesh(V2, T2)]
    W1, W2 = _deform(V1, 3) * 1.3, _deform(V2, 4)
    ctx.mesh_refit(g[0], _t(W1, torch.float64), bottomup=bottomup)
    ctx.mesh_refit(g[1], _t(W2, torch.float64), bottomup=bottomup)
    o[0].refit(W1, bottomup=bottomup)
    o[1].refit(W2, bottomup=bottomup)
    n = 600
    tf1, tf2 = F.poses(n, 0, 1.8, 0, 0xC9)
    qm = np.tile(g, (n, 1))
    oc, op, obv, olf = fclo.collide_batch(o, np.tile([0, 1], (n, 1)), tf1, tf2, 100000,
                                          store=4000)
    for engine in (F.ENGINE_FAST, F.ENGINE_PLAIN):
        print("case: collide after refit, bottomup=%d engine=%d" % (bottomup, engine), flush=True)
        rc, gc, gp, gbv, glf, nrec = collide_gpu(ctx, qm, tf1, tf2, 100000, engine=engine,
                                                 capacity=1 << 22)
        assert rc == 0 and np.array_equal(gc, oc) and nrec == oc.sum()
        assert np.array_equal(gbv, obv) and np.array_equal(glf, olf)
        for q in range(n):
            got = gp.get(q, np.zeros((0, 2), np.int32))
            assert _set(got) == _set(op[q, : oc[q]])
    counts = torch.zeros(n, dtype=torch.int32, device=DEV)
    num = torch.zeros(1, dtype=torch.int64, device=DEV)
    s = torch.cuda.current_stream().cuda_stream
    ctx.collide_batch_async(_t(qm, torch.int32), _t(tf1, torch.float64), _t(tf2, torch.float64),
                            1, counts, None, num, engine=F.ENGINE_BOOL, stream=s)
    assert ctx.sync_status(s) == 0
    assert np.array_equal(counts.cpu().numpy(), (oc > 0).astype(np.int32))
    assert (oc > 0).mean() > 0.2
    for m in g:
        ctx.mesh_destroy(m)


def test_distance_after_refit(ctx):
    (V1, T1), (V2, T2) = F.icosphere(10, 0.15, 61), F.icosphere(10, 0.15, 62)
    g = [ctx.mesh(V1, T1, F.RSS), ctx.mesh(V2, T2, F.RSS)]
    o = [fclo.OracleMesh(V1, T1, fclo.RSS), fclo.OracleMesh(V2, T2, fclo.RSS)]
    W1 = _deform(V1, 7)
    ctx.mesh_refit(g[0], _t(W1, torch.float64), bottomup=True)
    o[0].refit(W1, bottomup=True)
    n = 300
    tf1, tf2 = F.poses(n, 1, 2.1, 4.0, 0xCA)
    od, opts, oids, _, _ = fclo.distance_batch(o, np.tile([0, 1], (n, 1)), tf1, tf2)
    for engine in (F.ENGINE_PLAIN, F.ENGINE_FAST):
        print("case: distance after refit, engine=%d" % engine, flush=True)
        dist = torch.zeros(n, dtype=torch.float64, device=DEV)
        pts = torch.zeros((n, 6), dtype=torch.float64, device=DEV)
        ids = torch.zeros((n, 2), dtype=torch.int32, device=DEV)
        s = torch.cuda.current_stream().cuda_stream
        ctx.distance_batch_async(_t(np.tile(g, (n, 1)), torch.int32), _t(tf1, torch.float64),
                                 _t(tf2, torch.float64), dist, pts, ids, engine=engine, stream=s)
        assert ctx.sync_status(s) == 0
        assert np.array_equal(dist.cpu().numpy().view(np.uint64), od.view(np.uint64))
        assert np.array_equal(ids.cpu().numpy(), oids)
    for m in g:
        ctx.mesh_destroy(m)


def test_reference_frontlist_refit_flow(ctx):
    """collide_front_list_Test (test_fcl_frontlist.cpp:150-200): env.obj
    moved by pose A and built, the front recorded (identity poses), the
    vertices replaced by pose B's and refit top-down on the device, the
    front propagated: flags equal a fresh collide() at B, and every count,
    record and statistic equals the oracle's same flow."""
    from test_gpu_frontlist import run_gpu
    (V1, T1), (V2, T2) = util.load_obj("env.obj"), util.load_obj("rob.obj")
    A, B = util.random_transform_pairs(6, [-3000, -3000, 0, 3000, 3000, 3000], [1, 1, 1],
                                       0.005 * 2 * np.pi, 31)
    I = np.tile(util.identity_pose(), (1, 1))
    plain = [fclo.OracleMesh(V1, T1), fclo.OracleMesh(V2, T2)]
    for k, (a, b) in enumerate(zip(A, B)):
        Va, Vb = util.pose_apply(a, V1), util.pose_apply(b, V1)
        g = [ctx.mesh(Va, T1), ctx.mesh(V2, T2)]
        o = [fclo.OracleMesh(Va, T1), fclo.OracleMesh(V2, T2)]
        fs = F.FrontSet(ctx, 1, 1 << 16)
        try:
            print("case: front + refit flow, pose %d" % k, flush=True)
            qm = np.array([g], np.int32)
            off = np.array([0, 1], np.int32)
            rc, c0, _, _, _ = run_gpu(ctx, fs, off, qm, I, I, 1 << 30, 1 << 16)
            assert rc == 0
            _, ofront, ost = fclo.collide_front(o[0], I[0], o[1], I[0], [], cap=1 << 30)
            assert c0[0] == ost[0] and np.array_equal(fs.read(0), ofront)
            ctx.mesh_refit(g[0], _t(Vb, torch.float64), bottomup=False)
            ctx.mesh_refit(g[1], _t(V2, torch.float64), bottomup=False)
            o[0].refit(Vb, bottomup=False)
            o[1].refit(V2, bottomup=False)
            rc, c1, rec, bvt, lft = run_gpu(ctx, fs, off, qm, I, I, 1 << 30, 1 << 16)
            assert rc == 0
            op, ofront, ost = fclo.collide_front(o[0], I[0], o[1], I[0], ofront, cap=1 << 30)
            assert (c1[0], bvt[0], lft[0]) == ost
            assert np.array_equal(rec[:, 1:], op) and np.array_equal(fs.read(0), ofront)
            fresh, _ = fclo.collide(plain[0], b, plain[1], I[0], cap=1)
            assert (c1[0] > 0) == (len(fresh) > 0)
        finally:
            fs.close()
            for m in g:
                ctx.mesh_destroy(m)
