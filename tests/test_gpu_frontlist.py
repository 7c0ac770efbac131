"""BVTT front lists on the GPU (fclx_collide_front_batch) against the
oracle's restatement of detail::collide(node, front_list): per query the
count, the records (b1, b2) in the reference's order, num_bv_tests /
num_leaf_tests, and the front each sequence is left with, entry for entry.
Booleans also equal a fresh collide() per query (the property the
reference's test_fcl_frontlist.cpp checks)."""
import numpy as np
import pytest
import torch

import fcl_amd as F
import fclo
import util
from fcl_amd import workloads as W

pytestmark = pytest.mark.gpu

DEV = torch.device("cuda:0")


def _t(a, dtype):
    return torch.from_numpy(np.ascontiguousarray(a)).to(DEV, dtype)


def run_gpu(ctx, fs, offsets, qm, tf1, tf2, cap, capacity):
    n = len(qm)
    counts = torch.zeros(n, dtype=torch.int32, device=DEV)
    pairs = torch.zeros((capacity, 3), dtype=torch.int32, device=DEV)
    num = torch.zeros(1, dtype=torch.int64, device=DEV)
    bvt = torch.zeros(n, dtype=torch.int64, device=DEV)
    lft = torch.zeros(n, dtype=torch.int64, device=DEV)
    s = torch.cuda.current_stream().cuda_stream
    fs.collide_async(_t(offsets, torch.int32), _t(qm, torch.int32), _t(tf1, torch.float64),
                     _t(tf2, torch.float64), cap, counts, pairs, num, bv_tests=bvt,
                     leaf_tests=lft, stream=s)
    rc = ctx.sync_status(s)
    nrec = int(num.item())
    rec = pairs[:min(nrec, capacity)].cpu().numpy()
    rec = rec[np.argsort(rec[:, 0], kind="stable")]  # one lane per sequence: in order
    return rc, counts.cpu().numpy(), rec, bvt.cpu().numpy(), lft.cpu().numpy()


def oracle_sequences(om, offsets, qm, tf1, tf2, cap, fronts=None):
    n = len(qm)
    counts = np.zeros(n, np.int32)
    bvt = np.zeros(n, np.int64)
    lft = np.zeros(n, np.int64)
    recs = []
    out_fronts = []
    for s in range(len(offsets) - 1):
        front = np.zeros((0, 2), np.int32) if fronts is None else fronts[s]
        for q in range(offsets[s], offsets[s + 1]):
            p, front, st = fclo.collide_front(om[qm[q, 0]], tf1[q], om[qm[q, 1]], tf2[q], front,
                                              cap=cap)
            counts[q], bvt[q], lft[q] = st
            recs.append(np.concatenate([np.full((len(p), 1), q, np.int32), p], 1))
        out_fronts.append(front)
    return counts, np.concatenate(recs) if recs else np.zeros((0, 3), np.int32), bvt, lft, \
        out_fronts


def check(ctx, om, ids, offsets, qm, tf1, tf2, cap, label, capacity=1 << 16, calls=1):
    """Runs the sequences in `calls` consecutive batches (the fronts carry
    over between calls) and compares everything with the oracle; returns
    the oracle's counts and num_bv_tests of every query (in qm order)."""
    fs = F.FrontSet(ctx, len(offsets) - 1, capacity)
    all_c = np.zeros(len(qm), np.int32)
    all_bv = np.zeros(len(qm), np.int64)
    try:
        fronts = None
        cuts = np.linspace(0, max(np.diff(offsets)), calls + 1).astype(int)
        for c in range(calls):
            # the c-th slice of every sequence
            sub_q = []
            sub_off = [0]
            for s in range(len(offsets) - 1):
                a, b = offsets[s] + cuts[c], min(offsets[s] + cuts[c + 1], offsets[s + 1])
                sub_q += list(range(a, max(a, b)))
                sub_off.append(len(sub_q))
            sub_q = np.array(sub_q, np.int64)
            sub_off = np.array(sub_off, np.int32)
            print("case: %s, call %d, %d sequences, %d queries" % (label, c, len(sub_off) - 1,
                                                                  len(sub_q)), flush=True)
            rc, gc, grec, gbv, glf = run_gpu(ctx, fs, sub_off, ids[qm[sub_q]], tf1[sub_q],
                                            tf2[sub_q], cap, 1 << 20)
            assert rc == 0
            oc, orec, obv, olf, fronts = oracle_sequences(om, sub_off, qm[sub_q], tf1[sub_q],
                                                          tf2[sub_q], cap, fronts)
            assert np.array_equal(gc, oc), label
            assert np.array_equal(gbv, obv) and np.array_equal(glf, olf), label
            assert np.array_equal(grec, orec), label
            for s in range(len(sub_off) - 1):
                assert np.array_equal(fs.read(s), fronts[s]), (label, s)
            all_c[sub_q] = oc
            all_bv[sub_q] = obv
        return all_c, all_bv
    finally:
        fs.close()


def test_env_rob_reference_pairs(ctx):
    """The reference's fixture and pose pairs (test_fcl_frontlist.cpp):
    sequences of two queries, the front of pose A reused at pose B."""
    (V1, T1), (V2, T2) = util.load_obj("env.obj"), util.load_obj("rob.obj")
    om = [fclo.OracleMesh(V1, T1), fclo.OracleMesh(V2, T2)]
    ids = np.array([ctx.mesh(V1, T1), ctx.mesh(V2, T2)], np.int32)
    A, B = util.random_transform_pairs(40, [-3000, -3000, 0, 3000, 3000, 3000], [1, 1, 1],
                                       0.005 * 2 * np.pi, 2718)
    n = 40
    tf1 = np.empty((2 * n, 12))
    tf1[0::2], tf1[1::2] = A, B
    tf2 = np.tile(util.identity_pose(), (2 * n, 1))
    qm = np.tile([0, 1], (2 * n, 1))
    offsets = np.arange(0, 2 * n + 1, 2, dtype=np.int32)
    oc, _ = check(ctx, om, ids, offsets, qm, tf1, tf2, 1 << 30, "env/rob pairs")
    fresh = fclo.collide_batch(om, qm[1::2], B, tf2[1::2], 1)[0]
    assert np.array_equal(oc[1::2] > 0, fresh > 0)
    for m in ids:
        ctx.mesh_destroy(int(m))


def test_trajectory_sequences(ctx):
    """Link meshes along random-walk trajectories against a heightfield and
    boxes scene, first contact (num_max_contacts = 1, the serving request):
    24 sequences of 40 waypoints in two consecutive batches."""
    sV, sT = W.serving_scene(grid=48, n_boxes=80, extent=1.0, seed=0x71)
    rng = np.random.default_rng(0x72)
    links = [W.capsule_mesh(rng.uniform(0.03, 0.06), rng.uniform(0.1, 0.2),
                            int(rng.integers(200, 800)), 0.1, int(rng.integers(1, 2**62)))
             for _ in range(4)]
    om = [fclo.OracleMesh(sV, sT)] + [fclo.OracleMesh(V, T) for V, T in links]
    ids = np.array([ctx.mesh(sV, sT)] + [ctx.mesh(V, T) for V, T in links], np.int32)
    L = 40
    reqs = [W.serving_request(4, L, rng, extent=1.0) for _ in range(6)]
    qm, tf2, off = [], [], [0]
    for P in reqs:
        for link in range(4):  # one sequence per link and request
            qm += [[0, 1 + link]] * L
            tf2.append(P[:, link])
            off.append(off[-1] + L)
    qm = np.array(qm, np.int32)
    tf2 = np.concatenate(tf2)
    tf1 = np.tile(util.identity_pose(), (len(qm), 1))
    offsets = np.array(off, np.int32)
    oc, obv = check(ctx, om, ids, offsets, qm, tf1, tf2, 1, "trajectories", calls=2)
    fresh = fclo.collide_batch(om, qm, tf1, tf2, 1)[0]
    assert np.array_equal(oc > 0, fresh > 0)
    assert 0.05 < (fresh > 0).mean() < 0.95
    for m in ids:
        ctx.mesh_destroy(int(m))


def test_front_overflow_is_reported(ctx):
    """A front larger than the set's capacity: the batch reports
    FCLX_ERR_CAPACITY and the front is dropped, so the next batch starts
    that sequence from the roots (and equals the oracle's fresh start)."""
    (V1, T1), (V2, T2) = F.icosphere(8, 0.15, 3), F.icosphere(8, 0.15, 4)
    om = [fclo.OracleMesh(V1, T1), fclo.OracleMesh(V2, T2)]
    ids = np.array([ctx.mesh(V1, T1), ctx.mesh(V2, T2)], np.int32)
    tf1, tf2 = F.poses(4, 0, 0.3, 0, 17)
    qm = np.tile([0, 1], (4, 1))
    fs = F.FrontSet(ctx, 4, 8)
    try:
        print("case: front overflow", flush=True)
        rc, gc, _, _, _ = run_gpu(ctx, fs, np.arange(5, dtype=np.int32), ids[qm], tf1, tf2, 1,
                                  1024)
        assert rc == F.N.FCLX_ERR_CAPACITY
        assert all(len(fs.read(s)) == 0 for s in range(4))
        oc = fclo.collide_batch(om, qm, tf1, tf2, 1)[0]
        assert np.array_equal(gc > 0, oc > 0)
    finally:
        fs.close()
    for m in ids:
        ctx.mesh_destroy(int(m))
