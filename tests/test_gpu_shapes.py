"""Primitive-shape collide (GJK/EPA and the closed forms) on the GPU vs the
CPU oracle, through fclx_shape_collide_batch.

Bar: per-pair contact counts identical; every contact's normal, point and
depth bit-identical (the device runs the oracle's FP64 operations in the
same order with no fused multiply-add; division and square root are
correctly rounded on both).
"""
import numpy as np
import pytest
import torch

import fcl_amd as F
import fclo
from fcl_amd import workloads as W

pytestmark = pytest.mark.gpu

DEV = torch.device("cuda:0")


def _t(a, dtype):
    return torch.from_numpy(np.ascontiguousarray(a)).to(DEV, dtype)


def shapes_gpu(ctx, rec, pool, ps, tf1, tf2, cap, contact, tol=1e-6, passes=None):
    """Returns (status, counts (n,), contacts (n, cap, 7) = normal, pos, depth);
    passes: None, or a list that receives the (n, 2) GJK / EPA loop passes."""
    n = len(ps)
    d_rec = torch.from_numpy(np.ascontiguousarray(rec).view(np.uint8)).to(DEV)
    d_pool = _t(np.asarray(pool, np.float64).reshape(-1, 3), torch.float64)
    counts = torch.full((n,), -1, dtype=torch.int32, device=DEV)
    geo = None
    if contact:
        geo = (torch.full((n * cap, 3), float("nan"), dtype=torch.float64, device=DEV),
               torch.full((n * cap, 3), float("nan"), dtype=torch.float64, device=DEV),
               torch.full((n * cap,), float("nan"), dtype=torch.float64, device=DEV))
    s = torch.cuda.current_stream().cuda_stream
    d_passes = torch.full((n, 2), -1, dtype=torch.int32, device=DEV) if passes is not None else None
    ctx.shape_collide_batch_async(d_rec, d_pool, _t(ps, torch.int32), _t(tf1, torch.float64),
                                  _t(tf2, torch.float64), cap, counts, contact=geo, tol=tol,
                                  stream=s, passes=d_passes)
    rc = ctx.sync_status(s)
    if passes is not None:
        passes.append(d_passes.cpu().numpy())
    out = np.zeros((n, cap, 7))
    if contact:
        pos, nrm, dep = (t.cpu().numpy() for t in geo)
        out[:, :, 0:3] = nrm.reshape(n, cap, 3)
        out[:, :, 3:6] = pos.reshape(n, cap, 3)
        out[:, :, 6] = dep.reshape(n, cap)
    return rc, counts.cpu().numpy(), out


def check_parity(ctx, rec, pool, ps, tf1, tf2, cap, contact, label):
    print("case: %s n=%d cap=%d contact=%d" % (label, len(ps), cap, contact), flush=True)
    shapes = fclo.shapes_from_table(rec, pool)
    oc, ocon, opass = fclo.shape_collide_batch(shapes, ps, tf1, tf2, cap, contact, passes=True)
    gp = []
    rc, gc, gcon = shapes_gpu(ctx, rec, pool, ps, tf1, tf2, cap, contact, passes=gp)
    assert rc == 0
    # loop passes: GJK's always; EPA's where it ran for the contact
    assert np.array_equal(gp[0][:, 0], opass[:, 0])
    if contact:
        assert np.array_equal(gp[0][:, 1], opass[:, 1])
    bad = np.flatnonzero(gc != oc)
    assert len(bad) == 0, "counts differ at %s (gpu %s, oracle %s)" % (
        bad[:8], gc[bad[:8]], oc[bad[:8]])
    if contact:
        mask = np.arange(cap)[None, :] < oc[:, None]
        g, o = gcon[mask], ocon[mask]
        diff = np.flatnonzero(np.any(g != o, axis=1))
        assert len(diff) == 0, "contacts differ in %d of %d slots; first gpu %s oracle %s" % (
            len(diff), len(g), g[diff[0]], o[diff[0]])
    return oc


@pytest.mark.parametrize("cap,contact", [(1, True), (4, True), (1, False)])
def test_primitive_batch_parity(ctx, cap, contact):
    """The 4,000,000-pair primitive workload's generator at 20,000 pairs."""
    b = W.primitive_batch(20000, seed=0x5A9E)
    oc = check_parity(ctx, b["shapes"], b["vertices"], b["pair_shapes"], b["tf1"], b["tf2"],
                      cap, contact, "primitive batch")
    assert 0.1 < (oc > 0).mean() < 0.9


def test_every_type_pair(ctx):
    """All 25 ordered type pairs (closed forms both ways round, GJK/EPA for
    the rest), close poses so most pairs penetrate, cap 4 so box-box
    returns its whole contact manifold."""
    rng = np.random.default_rng(0x7E57)
    per = 300
    t1, t2 = np.meshgrid(np.arange(5), np.arange(5), indexing="ij")
    types = np.stack([np.repeat(t1.ravel(), per), np.repeat(t2.ravel(), per)], 1).ravel()
    n = len(types) // 2
    dims = np.exp(rng.uniform(np.log(0.1), np.log(1.0), (len(types), 3)))
    conv = [rng.standard_normal((32, 3)) * rng.uniform(0.1, 0.6) for _ in
            range(int((types == F.CONVEX).sum()))]
    rec, pool = F.shape_table(types, dims, conv)
    tf1 = W.random_poses(n, rng, 0.4)
    tf2 = W.random_poses(n, rng, 0.4)
    ps = np.arange(2 * n, dtype=np.int32).reshape(n, 2)
    oc = check_parity(ctx, rec, pool, ps, tf1, tf2, 4, True, "type grid")
    bb = (types[0::2] == F.BOX) & (types[1::2] == F.BOX)
    assert (oc[bb] > 1).any(), "no multi-contact box-box case was exercised"
    assert (oc > 0).mean() > 0.5


def _pose(R=None, t=(0, 0, 0)):
    tf = np.zeros(12)
    tf[:9] = (np.eye(3) if R is None else np.asarray(R)).ravel()
    tf[9:] = t
    return tf


def test_degenerate_configurations(ctx):
    """Coincident centres, exact touching, centres on an axis, a one-point
    and a flat convex, identical boxes at one pose."""
    rz = np.array([[0.0, -1.0, 0.0], [1.0, 0.0, 0.0], [0.0, 0.0, 1.0]])
    c45 = np.sqrt(0.5)
    r45 = np.array([[c45, -c45, 0.0], [c45, c45, 0.0], [0.0, 0.0, 1.0]])
    flat = np.array([[0, 0, 0], [0.5, 0, 0], [0, 0.5, 0], [0.5, 0.5, 0]], float)
    shapes = [
        (F.SPHERE, (0.5, 0, 0)), (F.SPHERE, (0.5, 0, 0)),        # 0 1
        (F.BOX, (1, 1, 1)), (F.BOX, (1, 1, 1)),                  # 2 3
        (F.SPHERE, (0.25, 0, 0)), (F.CYLINDER, (0.5, 1.0, 0)),   # 4 5
        (F.CAPSULE, (0.3, 1.0, 0)), (F.CONVEX, None),            # 6 7 (one point)
        (F.CONVEX, None), (F.BOX, (2, 0.5, 0.25)),               # 8 (flat) 9
    ]
    types = [t for t, _ in shapes]
    dims = [d if d is not None else (0, 0, 0) for _, d in shapes]
    rec, pool = F.shape_table(types, dims, [np.zeros((1, 3)), flat])
    I = _pose()
    cases = [
        (0, 1, I, I),                                  # coincident spheres
        (0, 1, I, _pose(t=(1.0, 0, 0))),               # exactly touching
        (2, 3, I, I),                                  # identical boxes
        (2, 3, I, _pose(r45, (0.3, 0.2, 0.9))),        # face-face, rotated: 8-point clip
        (2, 3, I, _pose(rz, (0.99, 0, 0))),            # shallow face contact
        (4, 2, I, I),                                  # sphere centre at box centre
        (2, 4, I, _pose(t=(0.5, 0.5, 0.5))),           # sphere centre on a box corner
        (4, 5, I, I),                                  # sphere on the cylinder axis
        (5, 4, _pose(t=(0, 0, 0.2)), I),
        (4, 6, _pose(t=(0, 0, 0.5)), I),               # sphere centre on a capsule end
        (7, 2, I, I),                                  # a point inside a box (GJK/EPA)
        (7, 5, _pose(t=(0, 0, 2)), I),                 # a point outside a cylinder
        (8, 2, _pose(t=(0, 0, 0.5)), I),               # flat convex on a box face
        (8, 8, I, _pose(t=(0.1, 0.1, 0))),             # two coplanar flat convexes
        (9, 2, I, _pose(r45, (0.2, 0.1, 0.3))),        # box-box edge/face mix
        (6, 6, I, _pose(rz, (0.1, 0, 0))),             # crossing capsules (GJK/EPA)
        (5, 5, I, I),                                  # coincident cylinders
    ]
    ps = np.array([[a, b] for a, b, _, _ in cases], np.int32)
    tf1 = np.stack([c[2] for c in cases])
    tf2 = np.stack([c[3] for c in cases])
    for cap in (1, 4):
        check_parity(ctx, rec, pool, ps, tf1, tf2, cap, True, "degenerate cap %d" % cap)


def test_malformed_shapes_are_rejected(ctx):
    """Pairs naming a missing shape, an unknown type or a convex whose
    vertices leave the pool get count 0 and fclx_sync reports
    FCLX_ERR_INVALID_ARGUMENT; the valid pairs of the batch are unaffected
    and the next batch is clean."""
    rec, pool = F.shape_table([F.SPHERE, F.SPHERE, F.CONVEX, F.BOX],
                              [(0.5, 0, 0), (0.5, 0, 0), (0, 0, 0), (1, 1, 1)],
                              [np.zeros((4, 3))])
    bad = rec.copy()
    bad[3]["type"] = 9
    bad[2]["num_vertices"] = 5  # offset 0 + 5 > 4 vertices in the pool
    ps = np.array([[0, 1], [0, 7], [-1, 0], [0, 3], [2, 0], [1, 0]], np.int32)
    I = np.tile(_pose(), (len(ps), 1))
    rc, gc, _ = shapes_gpu(ctx, bad, pool, ps, I, I, 1, True)
    assert rc == F.N.FCLX_ERR_INVALID_ARGUMENT
    assert list(gc) == [1, 0, 0, 0, 0, 1]
    rc, gc, _ = shapes_gpu(ctx, rec, pool, ps[[0, 3, 4, 5]], I[:4], I[:4], 1, True)
    assert rc == 0 and list(gc) == [1, 1, 1, 1]


def test_empty_batch(ctx):
    rec, pool = F.shape_table([F.SPHERE], [(1, 0, 0)])
    rc, gc, _ = shapes_gpu(ctx, rec, pool, np.zeros((0, 2), np.int32), np.zeros((0, 12)),
                           np.zeros((0, 12)), 1, True)
    assert rc == 0 and len(gc) == 0


def test_single_pair_entry(ctx):
    """fclx_shape_collide (host buffers, one pair: the path behind
    fcl::collide for two primitives) gives the batch's contacts and passes."""
    b = W.primitive_batch(400, seed=0x51)
    rec, pool, ps = b["shapes"], b["vertices"], b["pair_shapes"]
    print("case: single-pair entry, 400 pairs", flush=True)
    oc, ocon, opass = fclo.shape_collide_records(rec, pool, ps, b["tf1"], b["tf2"], 4, True,
                                                 passes=True)
    for q in range(len(ps)):
        sv = []
        for k in range(2):
            r = rec[ps[q, k]]
            o = int(r["vertex_offset"])
            sv.append(pool[o:o + int(r["num_vertices"])] if r["type"] == F.CONVEX else None)
        cs, passes = ctx.shape_collide(rec[ps[q, 0]], sv[0], b["tf1"][q], rec[ps[q, 1]], sv[1],
                                       b["tf2"][q], cap=4)
        assert len(cs) == oc[q]
        assert passes == tuple(opass[q])
        for i, (nrm, pos, dep) in enumerate(cs):
            assert np.array_equal(nrm, ocon[q, i, 0:3]) and np.array_equal(pos, ocon[q, i, 3:6])
            assert dep == ocon[q, i, 6]
