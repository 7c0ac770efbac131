"""The request loop (fclx_server_*) on the GPU.

Every request's result -- one boolean per link and waypoint, collide() with
CollisionRequest(1, false) and the scene as the first object -- must be the
same alone and in a mix of concurrent requests, whatever the batching
(streams, device-memory budget, batch size, engine), and equal to the
oracle's booleans.  Booleans are compared exactly.
"""
import numpy as np
import pytest

import fcl_amd as F
import fclo
from fcl_amd import workloads as W

pytestmark = pytest.mark.gpu

EXTENT = 1.0


@pytest.fixture(scope="module")
def setup(ctx):
    """A small scene (64x64 heightfield over [-1,1]^2 and 150 boxes), a
    4-link robot of capsule meshes and 24 trajectories of 8-96 waypoints;
    the oracle's booleans per request."""
    sV, sT = W.serving_scene(grid=64, n_boxes=150, extent=EXTENT, seed=0x51)
    rng = np.random.default_rng(0x52)
    links = [W.capsule_mesh(rng.uniform(0.02, 0.05), rng.uniform(0.08, 0.2),
                            int(rng.integers(200, 900)), 0.1, int(rng.integers(1, 2**62)))
             for _ in range(4)]
    scene = ctx.mesh(sV, sT)
    link_ids = np.array([ctx.mesh(V, T) for V, T in links], np.int32)
    om = [fclo.OracleMesh(sV, sT)] + [fclo.OracleMesh(V, T) for V, T in links]
    lens = np.exp(rng.uniform(np.log(8), np.log(96), 24)).astype(int)
    reqs = [np.ascontiguousarray(W.serving_request(4, int(L), rng, extent=EXTENT)) for L in lens]
    expect = []
    for P in reqs:
        nq = P.shape[0] * 4
        qm = np.stack([np.zeros(nq, np.int32), 1 + np.tile(np.arange(4, dtype=np.int32),
                                                         P.shape[0])], 1)
        tf1 = np.tile(np.array([1, 0, 0, 0, 1, 0, 0, 0, 1, 0, 0, 0], float), (nq, 1))
        c = fclo.collide_batch(om, qm, tf1, P.reshape(nq, 12), 1, store=0)[0]
        expect.append((c > 0).astype(np.uint8).reshape(P.shape[0], 4))
    frac = np.mean(np.concatenate([e.ravel() for e in expect]))
    assert 0.02 < frac < 0.9, "the trajectories should touch the scene sometimes: %g" % frac
    return scene, link_ids, reqs, expect


def _check(results, expect):
    for i, (r, e) in enumerate(zip(results, expect)):
        assert np.array_equal(r, e), "request %d differs from the oracle" % i


@pytest.mark.parametrize("streams", [1, 3])
def test_same_alone_and_in_the_mix(ctx, setup, streams):
    scene, link_ids, reqs, expect = setup
    print("case: serving mix, %d streams, %d requests" % (streams, len(reqs)), flush=True)
    srv = F.Server(ctx, scene, num_streams=streams)
    try:
        # a burst: all requests arrive at once, so batches mix them
        arrivals = np.zeros(len(reqs))
        mix, lat, _ = srv.replay([(link_ids, P) for P in reqs], arrivals)
        assert np.all(lat > 0)
        _check(mix, expect)
        batches, queries = srv.stats()
        assert queries == sum(P.shape[0] * 4 for P in reqs)
        assert batches < len(reqs), "a burst should be batched together"
        for i, P in enumerate(reqs[:8]):
            alone, _, _ = srv.replay([(link_ids, P)], [0.0])
            assert np.array_equal(alone[0], mix[i])
    finally:
        srv.close()


def test_pinned_and_pageable_requests(ctx, setup):
    """Requests whose poses sit in pinned host memory are copied to the
    device straight from it, the others through the loop's staging; a burst
    mixing both kinds gives every request the oracle's booleans."""
    import torch

    scene, link_ids, reqs, expect = setup
    pinned = []
    for i, P in enumerate(reqs):
        if i % 2 == 0:
            t = torch.empty(P.shape, dtype=torch.float64, pin_memory=True)
            t.numpy()[...] = P
            pinned.append(t.numpy())
        else:
            pinned.append(P)
    print("case: serving mix of pinned and pageable requests", flush=True)
    srv = F.Server(ctx, scene, num_streams=3)
    try:
        mix, _, _ = srv.replay([(link_ids, P) for P in pinned], np.zeros(len(reqs)))
        _check(mix, expect)
    finally:
        srv.close()


def test_memory_budget_and_batch_size(ctx, setup):
    """A budget of about two average requests and batches of at most 100
    queries: requests wait for memory, large ones run alone, results stay
    exact."""
    scene, link_ids, reqs, expect = setup
    avg = np.mean([P.shape[0] * 4 for P in reqs])
    print("case: serving budget %d B, batch 100" % int(2 * avg * 204), flush=True)
    srv = F.Server(ctx, scene, num_streams=2, hbm_budget=int(2 * avg * 204),
                   max_batch_queries=100)
    try:
        res, _, _ = srv.replay([(link_ids, P) for P in reqs], np.zeros(len(reqs)))
        _check(res, expect)
        batches, _ = srv.stats()
        assert batches >= len(reqs) // 2
    finally:
        srv.close()


def test_plain_engine_and_poll(ctx, setup):
    """The PLAIN engine behind the loop, requests through submit / poll."""
    scene, link_ids, reqs, expect = setup
    print("case: serving submit/poll, PLAIN engine", flush=True)
    srv = F.Server(ctx, scene, num_streams=2, engine=F.ENGINE_PLAIN)
    try:
        for i, P in enumerate(reqs):
            srv.submit(link_ids, P, tag=100 + i)
        got = {}
        for _ in range(200):
            for tag, status, nq, qms, lms, res in srv.poll(64, wait_ms=50):
                assert status == 0 and lms >= qms >= 0
                got[tag] = res
            if len(got) == len(reqs):
                break
        assert len(got) == len(reqs)
        _check([got[100 + i] for i in range(len(reqs))], expect)
    finally:
        srv.close()


def test_rejects_bad_requests(ctx, setup):
    scene, link_ids, reqs, _ = setup
    srv = F.Server(ctx, scene)
    try:
        with pytest.raises(F.FclxError):
            srv.submit(np.array([link_ids[0], 9999], np.int32), reqs[0][:, :2], tag=1)
        with pytest.raises(F.FclxError):
            srv.submit(link_ids, np.zeros((0, 4, 12)), tag=2)
    finally:
        srv.close()
    with pytest.raises(F.FclxError):
        F.Server(ctx, 9999)


def test_failed_launch_is_reported_and_destroy_returns(ctx, setup):
    """A batch whose launch fails (forced from the host through the
    "serve_fail_launches" parameter, before any device work) completes with
    an error status instead of holding its slot; the loop goes on serving
    the next requests exactly and destroy() returns."""
    scene, link_ids, reqs, expect = setup
    print("case: serving with one injected launch failure", flush=True)
    srv = F.Server(ctx, scene, num_streams=2)
    try:
        ctx.set_engine_param("serve_fail_launches", 1)
        srv.submit(link_ids, reqs[0], tag=1)
        got = []
        for _ in range(100):
            got += srv.poll(8, wait_ms=50)
            if got:
                break
        assert len(got) == 1 and got[0][0] == 1
        assert got[0][1] == F.N.FCLX_ERR_DEVICE, "the failed batch must carry its error"
        assert ctx.get_engine_param("serve_fail_launches") == 0
        res, _, _ = srv.replay([(link_ids, P) for P in reqs[:4]], np.zeros(4))
        _check(res, expect[:4])
    finally:
        srv.close()
