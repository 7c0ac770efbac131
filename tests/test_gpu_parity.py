"""HIP engine vs the CPU oracle on the same seeded inputs, through the C ABI.

Bar: collision counts, contact pairs (in the reference's order) and the
traversal statistics num_bv_tests / num_leaf_tests are bit-identical;
distances, nearest points and triangle ids are bit-identical too (the FP64
device path uses the oracle's evaluation order with no fused multiply-add).
"""
import numpy as np
import pytest
import torch

import fcl_amd as F
import fclo
import util
from fcl_amd import _native as N

pytestmark = pytest.mark.gpu

DEV = torch.device("cuda:0")


def _t(a, dtype):
    return torch.from_numpy(np.ascontiguousarray(a)).to(DEV, dtype)


def collide_gpu(ctx, qm, tf1, tf2, cap, engine=F.ENGINE_PLAIN, capacity=None, stats=True):
    n = qm.shape[0]
    capacity = capacity if capacity is not None else max(1, n * min(cap, 64))
    d_qm, d_tf1, d_tf2 = _t(qm, torch.int32), _t(tf1, torch.float64), _t(tf2, torch.float64)
    counts = torch.zeros(n, dtype=torch.int32, device=DEV)
    pairs = torch.zeros((capacity, 3), dtype=torch.int32, device=DEV)
    num = torch.zeros(1, dtype=torch.int64, device=DEV)
    bvt = torch.zeros(n, dtype=torch.int64, device=DEV)
    lft = torch.zeros(n, dtype=torch.int64, device=DEV)
    s = torch.cuda.current_stream().cuda_stream
    ctx.collide_batch_async(d_qm, d_tf1, d_tf2, cap, counts, pairs, num, engine=engine,
                            bv_tests=bvt if stats else None, leaf_tests=lft if stats else None,
                            stream=s)
    rc = ctx.sync_status(s)
    nrec = int(num.item())
    rec = pairs[: min(nrec, capacity)].cpu().numpy()
    # per-query lists in production order (records of one query keep order)
    order = np.argsort(rec[:, 0], kind="stable")
    rec = rec[order]
    per_q = {}
    if len(rec):
        qs, starts = np.unique(rec[:, 0], return_index=True)
        ends = list(starts[1:]) + [len(rec)]
        for q, a, b in zip(qs, starts, ends):
            per_q[int(q)] = rec[a:b, 1:]
    return rc, counts.cpu().numpy(), per_q, bvt.cpu().numpy(), lft.cpu().numpy(), nrec


@pytest.fixture(scope="module")
def c0_meshes(ctx):
    """The plain mesh-vs-mesh collide workload at reduced query count:
    two perturbed icospheres, 2,420 triangles each."""
    (V1, T1), (V2, T2) = F.icosphere(11, 0.15, 0xA0), F.icosphere(11, 0.15, 0xB0)
    o = [fclo.OracleMesh(V1, T1), fclo.OracleMesh(V2, T2)]
    g = [ctx.mesh(V1, T1), ctx.mesh(V2, T2)]
    return o, g


@pytest.mark.parametrize("engine", [F.ENGINE_PLAIN, F.ENGINE_FAST])
def test_first_contact_parity(ctx, c0_meshes, engine):
    o, g = c0_meshes
    n = 3000
    tf1, tf2 = F.poses(n, 0, 2.2, 0, 0xC0)
    oc, op, obv, olf = fclo.collide_batch(o, np.tile([0, 1], (n, 1)), tf1, tf2, 1, store=1)
    print("case: first contact, 2x2420 tris, n=%d engine=%d" % (n, engine), flush=True)
    rc, gc, gp, gbv, glf, _ = collide_gpu(ctx, np.tile(g, (n, 1)), tf1, tf2, 1, engine=engine)
    assert rc == 0
    assert np.array_equal(gc, oc)
    assert np.array_equal(gbv, obv) and np.array_equal(glf, olf)
    for q in range(n):
        if oc[q]:
            assert np.array_equal(gp[q], op[q, :1])
    assert 0.3 < (oc > 0).mean() < 0.7


@pytest.mark.parametrize("cap", [3, 100000])
def test_exhaustive_and_capped_parity(ctx, cap):
    (V1, T1), (V2, T2) = F.icosphere(7, 0.15, 11), F.heightfield(30, 0.05, 12)
    V2 = F.scale_to_radius(V2, 1.0)
    o = [fclo.OracleMesh(V1, T1), fclo.OracleMesh(V2, T2)]
    g = [ctx.mesh(V1, T1), ctx.mesh(V2, T2)]
    n = 600
    tf1, tf2 = F.poses(n, 0, 1.8, 0, 0xC1)
    store = min(cap, 4000)
    oc, op, obv, olf = fclo.collide_batch(o, np.tile([0, 1], (n, 1)), tf1, tf2, cap, store=store)
    print("case: exhaustive cap=%d n=%d" % (cap, n), flush=True)
    rc, gc, gp, gbv, glf, nrec = collide_gpu(ctx, np.tile(g, (n, 1)), tf1, tf2, cap,
                                             capacity=int(oc.sum()) + 16)
    assert rc == 0 and nrec == oc.sum()
    assert np.array_equal(gc, oc)
    assert np.array_equal(gbv, obv) and np.array_equal(glf, olf)
    for q in range(n):
        if oc[q]:
            assert np.array_equal(gp[q][:store], op[q, : min(oc[q], store)])
    assert oc.max() == cap if cap < 10 else oc.max() > 10


@pytest.mark.parametrize("engine", [F.ENGINE_PLAIN, F.ENGINE_FAST])
@pytest.mark.parametrize("cap", [1, 5, 100000])
def test_contact_mode_parity(ctx, cap, engine):
    """enable_contact: per query the same contacts as the oracle, in order,
    each with a bit-identical world-frame point, normal and depth; the cap
    counts contacts (a triangle pair gives 0-2)."""
    (V1, T1), (V2, T2) = F.icosphere(7, 0.15, 51), F.heightfield(30, 0.05, 52)
    V2 = F.scale_to_radius(V2, 1.0)
    o = [fclo.OracleMesh(V1, T1), fclo.OracleMesh(V2, T2)]
    g = [ctx.mesh(V1, T1), ctx.mesh(V2, T2)]
    n = 300
    tf1, tf2 = F.poses(n, 0, 1.8, 0, 0xC7)
    store = min(cap, 20000)
    want = [fclo.collide_contact(o[0], tf1[q], o[1], tf2[q], cap=cap, store=store)
            for q in range(n)]
    total = sum(w[5] for w in want)
    cap_rec = max(1, total + 16)
    print("case: contact mode cap=%d engine=%d n=%d" % (cap, engine, n), flush=True)
    d_qm = _t(np.tile(g, (n, 1)), torch.int32)
    counts = torch.zeros(n, dtype=torch.int32, device=DEV)
    pairs = torch.zeros((cap_rec, 3), dtype=torch.int32, device=DEV)
    pos = torch.zeros((cap_rec, 3), dtype=torch.float64, device=DEV)
    nrm = torch.zeros((cap_rec, 3), dtype=torch.float64, device=DEV)
    dep = torch.zeros(cap_rec, dtype=torch.float64, device=DEV)
    num = torch.zeros(1, dtype=torch.int64, device=DEV)
    s = torch.cuda.current_stream().cuda_stream
    ctx.collide_batch_async(d_qm, _t(tf1, torch.float64), _t(tf2, torch.float64), cap, counts,
                            pairs, num, engine=engine, stream=s, contact=(pos, nrm, dep))
    assert ctx.sync_status(s) == 0
    nrec = int(num.item())
    assert nrec == total
    gc = counts.cpu().numpy()
    rec, gpos, gnrm, gdep = (pairs[:nrec].cpu().numpy(), pos[:nrec].cpu().numpy(),
                             nrm[:nrec].cpu().numpy(), dep[:nrec].cpu().numpy())
    order = np.argsort(rec[:, 0], kind="stable")  # one lane per query: in order
    rec, gpos, gnrm, gdep = rec[order], gpos[order], gnrm[order], gdep[order]
    start = 0
    for q in range(n):
        wp, wpos, wn, wd, _, wcount = want[q]
        assert gc[q] == wcount
        k = wcount
        assert (rec[start:start + k, 0] == q).all()
        assert np.array_equal(rec[start:start + k, 1:], wp[:k])
        assert np.array_equal(gpos[start:start + k], wpos[:k])
        assert np.array_equal(gnrm[start:start + k], wn[:k])
        assert np.array_equal(gdep[start:start + k], wd[:k])
        start += k
    assert total > 0 and (cap > 1 or (gc <= 1).all())


def _set(p):
    return sorted(map(tuple, np.asarray(p).reshape(-1, 2).tolist()))


@pytest.mark.parametrize("persistent", [0, 1], ids=["rounds", "persistent"])
@pytest.mark.parametrize("cap,stats", [(100000, True), (70, True), (70, False)])
def test_fast_engine_exhaustive(ctx, cap, stats, persistent):
    """FAST engine, large caps: per query the same contact set and the same
    num_bv_tests / num_leaf_tests as the oracle; queries that reach the cap
    (cap=70 here) come back as the reference's depth-first prefix, by the
    exact re-run when statistics are requested and by depth-first keys
    otherwise.  Both forms: depth-first rounds, and one persistent launch
    (fast_persistent)."""
    ctx.set_engine_param("fast_persistent", persistent)
    try:
        _fast_engine_case(ctx, cap, stats)
    finally:
        ctx.set_engine_param("fast_persistent", 0)


def test_fast_persistent_graph_replays(ctx):
    """The persistent exhaustive form replayed from its captured HIP graph:
    three replays of one batch, with other eager work (PLAIN and BOOL
    batches, synchronisations) between them, then a different batch (a new
    capture) and the first again, each equal to the same form run without
    the graph (which the parity tests above compare with the oracle)."""
    (V1, T1), (V2, T2) = F.icosphere(9, 0.15, 21), F.heightfield(24, 0.05, 22)
    V2 = F.scale_to_radius(V2, 1.0)
    g = [ctx.mesh(V1, T1), ctx.mesh(V2, T2)]
    n = 700
    qm = np.tile(np.array(g, np.int32), (n, 1))
    A = F.poses(n, 0, 1.6, 0, 0xA7)
    B = F.poses(n, 0, 1.6, 0, 0xA8)
    d_qm = _t(qm, torch.int32)
    d = {k: (_t(p[0], torch.float64), _t(p[1], torch.float64)) for k, p in (("A", A), ("B", B))}
    counts = torch.zeros(n, dtype=torch.int32, device=DEV)
    pairs = torch.zeros((1 << 20, 3), dtype=torch.int32, device=DEV)
    num = torch.zeros(1, dtype=torch.int64, device=DEV)
    s = torch.cuda.current_stream().cuda_stream

    def run(key, graph):
        ctx.set_engine_param("fast_graph", graph)
        counts.fill_(-1)
        ctx.collide_batch_async(d_qm, d[key][0], d[key][1], 100000, counts, pairs, num,
                                engine=F.ENGINE_FAST, stream=s)
        assert ctx.sync_status(s) == 0
        k = int(num.item())
        rec = pairs[:k].cpu().numpy()
        return counts.cpu().numpy(), rec[np.lexsort((rec[:, 2], rec[:, 1], rec[:, 0]))]

    def other_work():
        c = torch.zeros(n, dtype=torch.int32, device=DEV)
        p = torch.zeros((n * 64, 3), dtype=torch.int32, device=DEV)
        m = torch.zeros(1, dtype=torch.int64, device=DEV)
        for e in (F.ENGINE_PLAIN, F.ENGINE_BOOL):
            ctx.collide_batch_async(d_qm, d["B"][0], d["B"][1], 64, c, p, m, engine=e, stream=s)
            ctx.sync(s)

    ctx.set_engine_param("fast_persistent", 1)
    try:
        print("case: persistent form, eager reference runs", flush=True)
        want = {k: run(k, 0) for k in ("A", "B")}
        assert want["A"][0].sum() > 0
        for i, key in enumerate(("A", "A", "A", "B", "A")):
            print("case: persistent form, graph replay %d of batch %s" % (i, key), flush=True)
            got = run(key, 1)
            assert np.array_equal(got[0], want[key][0])
            assert np.array_equal(got[1], want[key][1])
            other_work()
    finally:
        ctx.set_engine_param("fast_persistent", 0)
        ctx.set_engine_param("fast_graph", 1)


# "full": a tiny breadth-first front and deferred buffers of a few hundred
# entries, so depth-first rounds stop constantly because a buffer is full;
# "idle": a wide front and one resident block per CU, and a round ends as
# soon as one wavefront idles, so the other stacks are redistributed.
_FULL = {"defer_leaf": 256, "defer_uncertain": 64, "front_target": 64, "max_levels": 2,
         "dfs_blocks_per_cu": 1, "stop_poll": 1}
STRESS = {
    # leaf pairs all deferred, so the small leaf buffer fills too
    "full": dict(_FULL, dfs_inline_leaves=0),
    # leaf pairs tested inside the depth-first kernel: only the uncertain
    # buffer ends rounds, and inline hits and deferred ones share the output
    # (and a small LDS stack top, so stacks spill to device memory and are
    # read back, and stack hand-overs take both parts)
    "full_inline": dict(_FULL, dfs_inline_leaves=1, dfs_stack_top=128),
    "idle": {"front_target": 100000, "idle_stop": 0.0005, "dfs_blocks_per_cu": 1,
             "stop_poll": 1},
}


@pytest.mark.parametrize("mode", ["full", "full_inline", "idle"])
@pytest.mark.parametrize("cap,stats", [(100000, True), (70, False)])
def test_fast_engine_round_boundaries(mode, cap, stats):
    """The FAST-engine checks with tunables that end depth-first rounds early:
    stack hand-over, unconsumed input, uncertain-task children and queue
    dumps cross round boundaries, and results stay exact."""
    c = F.Context(0)
    try:
        for k, v in STRESS[mode].items():
            c.set_engine_param(k, v)
        print("case: round boundaries, mode=%s" % mode, flush=True)
        rep = _fast_engine_case(c, cap, stats)
    finally:
        c.close()
    if mode == "full":
        assert rep["dfs_rounds"] >= 5  # (how many depends on timing)
    elif mode == "full_inline":
        assert rep["dfs_rounds"] >= 2
    else:
        # rebalancing is armed only when the front gives every wavefront a
        # full step (one block of 4 wavefronts per CU here)
        armed = rep["front_size"] >= 64 * 4 * torch.cuda.get_device_properties(0).multi_processor_count
        assert rep["dfs_rounds"] >= 2 or not armed


def _fast_engine_case(ctx, cap, stats):
    (V1, T1), (V2, T2) = F.icosphere(9, 0.15, 21), F.heightfield(24, 0.05, 22)
    V2 = F.scale_to_radius(V2, 1.0)
    (V3, T3) = F.icosphere(6, 0.15, 23)
    meshes = [(V1, T1), (V2, T2), (V3, T3)]
    o = [fclo.OracleMesh(V, T) for V, T in meshes]
    g = [ctx.mesh(V, T) for V, T in meshes]
    n = 900
    rng = np.random.default_rng(31)
    pairs = np.array([(0, 1), (1, 2), (0, 2), (2, 2)], np.int32)[rng.integers(0, 4, n)]
    tf1, tf2 = F.poses(n, 0, 1.6, 0, 0xC3)
    store = min(cap, 3000)
    oc, op, obv, olf = fclo.collide_batch(o, pairs, tf1, tf2, cap, store=store)
    gq = np.array(g, np.int32)[pairs]
    print("case: fast exhaustive cap=%d n=%d stats=%s" % (cap, n, stats), flush=True)
    # before the fix-up a capped query has all its pairs recorded: room for all
    rc, gc, gp, gbv, glf, nrec = collide_gpu(ctx, gq, tf1, tf2, cap, engine=F.ENGINE_FAST,
                                             capacity=1 << 22, stats=stats)
    assert rc == 0
    assert np.array_equal(gc, oc)
    assert nrec == oc.sum()
    if stats:
        assert np.array_equal(gbv, obv) and np.array_equal(glf, olf)
    flagged = 0
    for q in range(n):
        got = gp.get(q, np.zeros((0, 2), np.int32))
        want = op[q, : min(oc[q], store)]
        if oc[q] >= cap:
            flagged += 1
            assert np.array_equal(got, want)  # depth-first prefix, in order
        else:
            assert _set(got) == _set(want)
    if cap < 1000:
        assert flagged > 0
    assert oc.max() > 50
    return ctx.engine_report()


def test_env_rob_exhaustive(ctx):
    """The reference's mesh_mesh fixture (env.obj vs rob.obj) with its pose
    distribution: GPU contact sets equal the oracle's and brute force."""
    (V1, T1), (V2, T2) = util.load_obj("env.obj"), util.load_obj("rob.obj")
    o = [fclo.OracleMesh(V1, T1), fclo.OracleMesh(V2, T2)]
    g = [ctx.mesh(V1, T1), ctx.mesh(V2, T2)]
    n = 10
    tf1 = util.random_transforms(n, [-3000, -3000, 0, 3000, 3000, 3000], 2024)
    tf2 = np.tile(util.identity_pose(), (n, 1))
    print("case: env/rob exhaustive", flush=True)
    rc, gc, gp, _, _, _ = collide_gpu(ctx, np.tile(g, (n, 1)), tf1, tf2, 1 << 30,
                                      capacity=1 << 20)
    assert rc == 0
    for q in range(n):
        bf = fclo.collide_all_pairs(o[0], tf1[q], o[1], tf2[q])
        got = gp.get(q, np.zeros((0, 2), np.int32))
        key = lambda p: sorted(map(tuple, p))
        assert key(got) == key(bf)


def test_fast_engine_fresh_context():
    """FAST engine in a new context, whose work buffers start empty and grow
    while the breadth-first phase accumulates leaf candidates over several
    levels: the env/rob contact sets equal the oracle's (regression for
    candidates lost when that buffer was reallocated)."""
    (V1, T1), (V2, T2) = util.load_obj("env.obj"), util.load_obj("rob.obj")
    o = [fclo.OracleMesh(V1, T1), fclo.OracleMesh(V2, T2)]
    ctx = F.Context(0)
    try:
        g = [ctx.mesh(V1, T1), ctx.mesh(V2, T2)]
        n = 10
        tf1 = util.random_transforms(n, [-3000, -3000, 0, 3000, 3000, 3000], 2024)
        tf2 = np.tile(util.identity_pose(), (n, 1))
        print("case: fast engine, fresh context, env/rob", flush=True)
        rc, gc, gp, gbv, glf, nrec = collide_gpu(ctx, np.tile(g, (n, 1)), tf1, tf2, 100000,
                                                 engine=F.ENGINE_FAST, capacity=1 << 16)
    finally:
        ctx.close()
    assert rc == 0
    oc, op, obv, olf = fclo.collide_batch(o, np.tile([0, 1], (n, 1)), tf1, tf2, 100000,
                                          store=4096)
    assert np.array_equal(gc, oc) and nrec == oc.sum() and oc.sum() > 0
    assert np.array_equal(gbv, obv) and np.array_equal(glf, olf)
    for q in range(n):
        got = gp.get(q, np.zeros((0, 2), np.int32))
        assert _set(got) == _set(op[q, : oc[q]])


def test_single_query_entry_point(ctx, c0_meshes):
    o, g = c0_meshes
    tf1, tf2 = F.poses(40, 0, 1.5, 0, 0xC2)
    for q in range(40):
        for cap in (1, 4):
            gp, gst = ctx.collide(g[0], tf1[q], g[1], tf2[q], cap=cap)
            op, ost = fclo.collide(o[0], tf1[q], o[1], tf2[q], cap=cap)
            assert np.array_equal(gp, op) and tuple(gst) == tuple(ost)


@pytest.mark.parametrize("bv_type", [F.RSS, F.OBBRSS])
def test_distance_parity(ctx, bv_type):
    """Mesh distance at reduced size (the distance workload uses 20,480-tri
    spheres; here 3,920): every output bit-identical to the oracle."""
    (V1, T1), (V2, T2) = F.icosphere(14, 0.15, 0xD0), F.icosphere(14, 0.15, 0xD1)
    o = [fclo.OracleMesh(V1, T1, bv_type), fclo.OracleMesh(V2, T2, bv_type)]
    g = [ctx.mesh(V1, T1, bv_type), ctx.mesh(V2, T2, bv_type)]
    n = 400
    tf1, tf2 = F.poses(n, 1, 2.1, 4.0, 0xD2)
    od, opts, oids, obv, olf = fclo.distance_batch(o, np.tile([0, 1], (n, 1)), tf1, tf2)
    print("case: distance bv=%d n=%d" % (bv_type, n), flush=True)
    qm = np.tile(g, (n, 1))
    dist = torch.zeros(n, dtype=torch.float64, device=DEV)
    pts = torch.zeros((n, 6), dtype=torch.float64, device=DEV)
    ids = torch.zeros((n, 2), dtype=torch.int32, device=DEV)
    bvt = torch.zeros(n, dtype=torch.int64, device=DEV)
    lft = torch.zeros(n, dtype=torch.int64, device=DEV)
    s = torch.cuda.current_stream().cuda_stream
    ctx.distance_batch_async(_t(qm, torch.int32), _t(tf1, torch.float64), _t(tf2, torch.float64),
                             dist, pts, ids, bv_tests=bvt, leaf_tests=lft, stream=s)
    assert ctx.sync_status(s) == 0
    gd = dist.cpu().numpy()
    assert np.array_equal(gd.view(np.uint64), od.view(np.uint64))
    assert np.array_equal(ids.cpu().numpy(), oids)
    assert np.array_equal(pts.cpu().numpy().view(np.uint64), opts.view(np.uint64))
    assert np.array_equal(bvt.cpu().numpy(), obv) and np.array_equal(lft.cpu().numpy(), olf)
    # single-query entry point
    for q in range(0, n, 50):
        d, p, i, st = ctx.distance(g[0], tf1[q], g[1], tf2[q])
        assert d == od[q] and tuple(i) == tuple(oids[q]) and tuple(st) == (obv[q], olf[q])


DIST_DEFAULTS = {"dist_queries_per_wave": 4, "dist_waves_per_simd": 2, "dist_width": 32,
                 "dist_leaf_batch": 16, "dist_fp32_cull": 0}


@pytest.mark.parametrize("params", [
    {},
    {"dist_queries_per_wave": 2, "dist_waves_per_simd": 3, "dist_width": 64},
    {"dist_queries_per_wave": 8, "dist_width": 8, "dist_leaf_batch": 1},
    {"dist_fp32_cull": 1},
    {"dist_fp32_cull": 1, "dist_queries_per_wave": 2, "dist_waves_per_simd": 3, "dist_width": 64},
    {"dist_fp32_cull": 1, "dist_queries_per_wave": 8, "dist_width": 8, "dist_leaf_batch": 1},
], ids=["default", "q2w3", "q8narrow", "fp32", "fp32_q2w3", "fp32_q8narrow"])
def test_distance_fast_engine_tunables(ctx, params):
    """The FAST distance engine under other slot counts, widths and leaf
    batches, with FP64 bounds and with FP32 culling: the same outputs as the
    oracle's (OBBRSS meshes, 300 poses, a third of them touching)."""
    (V1, T1), (V2, T2) = F.icosphere(10, 0.15, 0xD6), F.icosphere(8, 0.15, 0xD7)
    o = [fclo.OracleMesh(V1, T1), fclo.OracleMesh(V2, T2)]
    g = [ctx.mesh(V1, T1), ctx.mesh(V2, T2)]
    n = 300
    tf1, tf2 = F.poses(n, 1, 1.2, 3.5, 0xD8)
    od, opts, oids, _, _ = fclo.distance_batch(o, np.tile([0, 1], (n, 1)), tf1, tf2)
    print("case: fast distance tunables %s" % params, flush=True)
    try:
        for k, v in params.items():
            ctx.set_engine_param(k, v)
        dist = torch.zeros(n, dtype=torch.float64, device=DEV)
        pts = torch.zeros((n, 6), dtype=torch.float64, device=DEV)
        ids = torch.zeros((n, 2), dtype=torch.int32, device=DEV)
        s = torch.cuda.current_stream().cuda_stream
        ctx.distance_batch_async(_t(np.tile(g, (n, 1)), torch.int32), _t(tf1, torch.float64),
                                 _t(tf2, torch.float64), dist, pts, ids, engine=F.ENGINE_FAST,
                                 stream=s)
        assert ctx.sync_status(s) == 0
        assert np.array_equal(dist.cpu().numpy().view(np.uint64), od.view(np.uint64))
        assert np.array_equal(ids.cpu().numpy(), oids)
        assert np.array_equal(pts.cpu().numpy().view(np.uint64), opts.view(np.uint64))
    finally:
        for k, v in DIST_DEFAULTS.items():
            ctx.set_engine_param(k, v)


@pytest.mark.parametrize("bv_type", [F.RSS, F.OBBRSS])
def test_distance_fast_engine(ctx, bv_type):
    """FCLX_ENGINE_FAST distance (wavefront-wide search of several queries
    at once): min_distance, the reported pair and the nearest points
    bit-identical to the oracle's, on separated and on overlapping poses
    (ties between neighbouring triangles are common; the engine reports the
    pair the reference finds first)."""
    (V1, T1), (V2, T2) = F.icosphere(14, 0.15, 0xD0), F.icosphere(9, 0.15, 0xD3)
    o = [fclo.OracleMesh(V1, T1, bv_type), fclo.OracleMesh(V2, T2, bv_type)]
    g = [ctx.mesh(V1, T1, bv_type), ctx.mesh(V2, T2, bv_type)]
    n = 600
    a1, a2 = F.poses(n // 2, 1, 2.1, 4.0, 0xD4)
    b1, b2 = F.poses(n - n // 2, 1, 0.3, 2.2, 0xD5)
    tf1, tf2 = np.concatenate([a1, b1]), np.concatenate([a2, b2])
    od, opts, oids, _, _ = fclo.distance_batch(o, np.tile([0, 1], (n, 1)), tf1, tf2)
    print("case: fast distance bv=%d n=%d" % (bv_type, n), flush=True)
    qm = np.tile(g, (n, 1))
    dist = torch.zeros(n, dtype=torch.float64, device=DEV)
    pts = torch.zeros((n, 6), dtype=torch.float64, device=DEV)
    ids = torch.zeros((n, 2), dtype=torch.int32, device=DEV)
    bvt = torch.zeros(n, dtype=torch.int64, device=DEV)
    lft = torch.zeros(n, dtype=torch.int64, device=DEV)
    s = torch.cuda.current_stream().cuda_stream
    ctx.distance_batch_async(_t(qm, torch.int32), _t(tf1, torch.float64), _t(tf2, torch.float64),
                             dist, pts, ids, engine=F.ENGINE_FAST, bv_tests=bvt, leaf_tests=lft,
                             stream=s)
    assert ctx.sync_status(s) == 0
    gd = dist.cpu().numpy()
    gi = ids.cpu().numpy()
    gp = pts.cpu().numpy()
    assert np.array_equal(gd.view(np.uint64), od.view(np.uint64))
    pos = od > 0
    assert 0.2 < pos.mean() < 0.9, "both separated and touching poses: %g" % pos.mean()
    bad = np.flatnonzero(np.any(gi != oids, axis=1))
    assert len(bad) == 0, "queries %s report another pair than the oracle" % bad[:10]
    assert np.array_equal(gp.view(np.uint64), opts.view(np.uint64))
    assert np.all(bvt.cpu().numpy() > 0) and np.all(lft.cpu().numpy() >= 0)


def test_bad_queries_are_reported(ctx, c0_meshes):
    o, g = c0_meshes
    tf1, tf2 = F.poses(4, 0, 2.2, 0, 1)
    qm = np.array([[g[0], g[1]], [g[0], 999], [-1, g[1]], [g[0], g[1]]], np.int32)
    rc, gc, _, _, _, _ = collide_gpu(ctx, qm, tf1, tf2, 1)
    assert rc == N.FCLX_ERR_INVALID_ARGUMENT
    assert gc[1] == 0 and gc[2] == 0
    # the status was cleared: a good batch is clean again
    rc, _, _, _, _, _ = collide_gpu(ctx, np.tile(g, (4, 1)), tf1, tf2, 1)
    assert rc == 0


def test_pair_capacity_overflow_reported(ctx, c0_meshes):
    o, g = c0_meshes
    n = 200
    tf1, tf2 = F.poses(n, 0, 0.5, 0, 2)
    oc, _, _, _ = fclo.collide_batch(o, np.tile([0, 1], (n, 1)), tf1, tf2, 100000)
    rc, gc, _, _, _, nrec = collide_gpu(ctx, np.tile(g, (n, 1)), tf1, tf2, 100000, capacity=10)
    assert rc == N.FCLX_ERR_CAPACITY
    assert np.array_equal(gc, oc) and nrec == oc.sum()


def test_empty_model_rejected(ctx):
    with pytest.raises(F.FclxError) as e:
        ctx.mesh(np.zeros((0, 3)), np.zeros((0, 3), np.int32))
    assert e.value.code == N.FCLX_ERR_EMPTY_MODEL


def _bool_gpu(ctx, qm, tf1, tf2, stream=None):
    n = qm.shape[0]
    counts = torch.full((n,), -1, dtype=torch.int32, device=DEV)
    num = torch.zeros(1, dtype=torch.int64, device=DEV)
    d = (_t(qm, torch.int32), _t(tf1, torch.float64), _t(tf2, torch.float64))
    s = stream if stream is not None else torch.cuda.current_stream()
    with torch.cuda.stream(s):
        ctx.collide_batch_async(d[0], d[1], d[2], 1, counts, None, num, engine=F.ENGINE_BOOL,
                                stream=s.cuda_stream)
    return counts, num, d


@pytest.mark.parametrize("case", ["c0", "uneven"])
def test_bool_engine_matches_oracle(ctx, c0_meshes, case):
    """FCLX_ENGINE_BOOL: counts are exactly (oracle count > 0), on the plain
    mesh-vs-mesh workload and on uneven pairs (sphere, terrain, small
    sphere; self pairs included)."""
    if case == "c0":
        o, g = c0_meshes
        n = 3000
        pairs = np.tile(np.array([0, 1], np.int32), (n, 1))
        tf1, tf2 = F.poses(n, 0, 2.2, 0, 0xC0)
    else:
        (V1, T1), (V2, T2) = F.icosphere(9, 0.15, 21), F.heightfield(24, 0.05, 22)
        V2 = F.scale_to_radius(V2, 1.0)
        (V3, T3) = F.icosphere(6, 0.15, 23)
        meshes = [(V1, T1), (V2, T2), (V3, T3)]
        o = [fclo.OracleMesh(V, T) for V, T in meshes]
        g = [ctx.mesh(V, T) for V, T in meshes]
        n = 4000
        rng = np.random.default_rng(41)
        pairs = np.array([(0, 1), (1, 2), (0, 2), (2, 2), (1, 1)], np.int32)[rng.integers(0, 5, n)]
        tf1, tf2 = F.poses(n, 0, 1.6, 0, 0xC4)
    oc = fclo.collide_batch(o, pairs, tf1, tf2, 1, store=0)[0]
    print("case: bool engine %s n=%d" % (case, n), flush=True)
    counts, num, _ = _bool_gpu(ctx, np.array(g, np.int32)[pairs], tf1, tf2)
    assert ctx.sync_status(torch.cuda.current_stream().cuda_stream) == 0
    assert np.array_equal(counts.cpu().numpy(), (oc > 0).astype(np.int32))
    assert int(num.item()) == 0
    assert 0.1 < (oc > 0).mean() < 0.9


BOOL_DEFAULTS = {"bool_queries_per_wave": 8, "bool_waves_per_simd": 4, "bool_leaf_batch": 64,
                 "bool_width": 16, "bool_refill": 4}


@pytest.mark.parametrize("params", [
    {"bool_queries_per_wave": 8, "bool_waves_per_simd": 3},
    {"bool_queries_per_wave": 16, "bool_width": 8},
    {"bool_queries_per_wave": 4, "bool_width": 16, "bool_refill": 2},
    {"bool_queries_per_wave": 32, "bool_width": 2, "bool_waves_per_simd": 3, "bool_refill": 1},
    {"bool_queries_per_wave": 32, "bool_width": 64, "bool_waves_per_simd": 2, "bool_leaf_batch": 7},
    {"bool_width": 64, "bool_leaf_batch": 1, "bool_refill": 16},
    {"bool_width": 1, "bool_waves_per_simd": 2, "bool_leaf_batch": 17, "bool_refill": 32},
    {"bool_queries_per_wave": 0, "bool_width": 32},
], ids=["q8w3", "q16w8", "q4", "q32narrow", "q32wide", "wide_lb1", "narrow_lb17", "auto"])
def test_bool_engine_tunables(ctx, c0_meshes, params):
    """Every slot count, width, leaf batch and refill threshold gives the
    oracle's booleans, on 6000 and 1000 queries of the plain pair and on a
    batch smaller than one wavefront's slots."""
    o, g = c0_meshes
    print("case: bool engine tunables %s" % params, flush=True)
    try:
        for k, v in params.items():
            ctx.set_engine_param(k, v)
        for n, seed in ((6000, 0xE0), (1000, 0xE1), (5, 0xE2)):
            tf1, tf2 = F.poses(n, 0, 2.2, 0, seed)
            oc = fclo.collide_batch(o, np.tile([0, 1], (n, 1)), tf1, tf2, 1, store=0)[0]
            counts, _, _ = _bool_gpu(ctx, np.tile(np.array(g, np.int32), (n, 1)), tf1, tf2)
            assert ctx.sync_status(torch.cuda.current_stream().cuda_stream) == 0
            assert np.array_equal(counts.cpu().numpy(), (oc > 0).astype(np.int32))
    finally:
        for k, v in BOOL_DEFAULTS.items():
            ctx.set_engine_param(k, v)


def test_bool_engine_bad_queries(ctx, c0_meshes):
    """Bad mesh ids in a BOOL batch: the batch reports an invalid argument,
    those queries count 0 and the others are still answered."""
    o, g = c0_meshes
    n = 300
    tf1, tf2 = F.poses(n, 0, 2.2, 0, 0xE3)
    qm = np.tile(np.array(g, np.int32), (n, 1))
    bad = np.arange(3, n, 37)
    qm[bad, 1] = 999
    print("case: bool engine, %d bad queries of %d" % (len(bad), n), flush=True)
    counts, _, _ = _bool_gpu(ctx, qm, tf1, tf2)
    assert ctx.sync_status(torch.cuda.current_stream().cuda_stream) == N.FCLX_ERR_INVALID_ARGUMENT
    oc = fclo.collide_batch(o, np.tile([0, 1], (n, 1)), tf1, tf2, 1, store=0)[0]
    want = (oc > 0).astype(np.int32)
    want[bad] = 0
    assert np.array_equal(counts.cpu().numpy(), want)
    counts, _, _ = _bool_gpu(ctx, np.tile(np.array(g, np.int32), (n, 1)), tf1, tf2)
    assert ctx.sync_status(torch.cuda.current_stream().cuda_stream) == 0


def test_bool_engine_concurrent_streams(ctx, c0_meshes):
    """Boolean batches on several streams at once give what they give one
    after the other."""
    o, g = c0_meshes
    n = 2000
    qm = np.tile(np.array(g, np.int32), (n, 1))
    streams = [torch.cuda.Stream() for _ in range(4)]
    poses = [F.poses(n, 0, 2.2, 0, 0xD0 + k) for k in range(4)]
    print("case: bool engine, 4 streams x %d queries" % n, flush=True)
    alone = []
    for k in range(4):
        c, _, _ = _bool_gpu(ctx, qm, poses[k][0], poses[k][1], streams[k])
        streams[k].synchronize()
        alone.append(c.cpu().numpy())
    outs = [_bool_gpu(ctx, qm, poses[k][0], poses[k][1], streams[k]) for k in range(4)]
    for s in streams:
        s.synchronize()
    assert ctx.sync_status(None) == 0
    for k in range(4):
        assert np.array_equal(outs[k][0].cpu().numpy(), alone[k])
        oc = fclo.collide_batch(o, np.tile([0, 1], (n, 1)), poses[k][0], poses[k][1], 1)[0]
        assert np.array_equal(alone[k], (oc > 0).astype(np.int32))
