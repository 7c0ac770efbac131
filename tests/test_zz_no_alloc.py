"""fclx_collide_batch_noalloc (the boolean collide that allocates nothing,
its work area a fixed device array per GPU) against fclx_collide_batch with
FCLX_ENGINE_BOOL, through the C ABI: the same counts on two calls in a row,
the second smaller, and on a batch large enough that the grid is cut to the
stacks that fit the array."""
import ctypes

import numpy as np
import pytest
import torch

import fcl_amd as F
import fclo
from fcl_amd import _native as N

pytestmark = pytest.mark.gpu

DEV = torch.device("cuda:0")


def _t(a, dtype):
    return torch.from_numpy(np.ascontiguousarray(a)).to(DEV, dtype)


@pytest.fixture(scope="module")
def scene():
    """The plain mesh-vs-mesh pair (two perturbed icospheres, 2,420 triangles
    each) and an uneven set: sphere, terrain and a small sphere."""
    c = F.Context(0)
    (V1, T1), (V2, T2) = F.icosphere(11, 0.15, 0xA0), F.icosphere(11, 0.15, 0xB0)
    (V3, T3), (V4, T4) = F.icosphere(9, 0.15, 21), F.heightfield(24, 0.05, 22)
    V4 = F.scale_to_radius(V4, 1.0)
    meshes = [(V1, T1), (V2, T2), (V3, T3), (V4, T4)]
    g = np.array([c.mesh(V, T) for V, T in meshes], np.int32)
    yield c, g, meshes
    c.close()


def _queries(g, n, seed):
    rng = np.random.default_rng(seed)
    pairs = np.array([(0, 1), (2, 3), (0, 3), (1, 2), (3, 3)], np.int32)[rng.integers(0, 5, n)]
    tf1, tf2 = F.poses(n, 0, 1.8, 0, seed)
    return pairs, g[pairs], tf1, tf2


def _run(c, qm, tf1, tf2, noalloc):
    n = qm.shape[0]
    counts = torch.full((n,), -1, dtype=torch.int32, device=DEV)
    num = torch.full((1,), 7, dtype=torch.int64, device=DEV)
    s = torch.cuda.current_stream().cuda_stream
    c.collide_batch_async(_t(qm, torch.int32), _t(tf1, torch.float64), _t(tf2, torch.float64), 1,
                          counts, None, num, engine=F.ENGINE_BOOL, stream=s, noalloc=noalloc)
    assert c.sync_status(s) == 0
    assert int(num.item()) == 0
    return counts.cpu().numpy()


def test_noalloc_matches_bool_batch(scene):
    c, g, meshes = scene
    for n, seed in ((70000, 0xF1), (700, 0xF2)):
        print("case: no-allocation collide, %d queries, seed %#x" % (n, seed), flush=True)
        _, qm, tf1, tf2 = _queries(g, n, seed)
        got = _run(c, qm, tf1, tf2, noalloc=True)
        want = _run(c, qm, tf1, tf2, noalloc=False)
        assert np.array_equal(got, want)
        assert set(np.unique(got)) <= {0, 1} and 0.05 < got.mean() < 0.95


def test_noalloc_matches_oracle(scene):
    c, g, meshes = scene
    o = [fclo.OracleMesh(V, T) for V, T in meshes]
    for n, seed in ((3000, 0xF3), (257, 0xF4)):
        print("case: no-allocation collide vs oracle, %d queries" % n, flush=True)
        pairs, qm, tf1, tf2 = _queries(g, n, seed)
        oc = fclo.collide_batch(o, pairs, tf1, tf2, 1, store=0)[0]
        assert np.array_equal(_run(c, qm, tf1, tf2, noalloc=True), (oc > 0).astype(np.int32))


def test_noalloc_rejects_what_it_does_not_run(scene):
    """Other engines and contact geometry are refused before any launch."""
    c, g, _ = scene
    n = 8
    _, qm, tf1, tf2 = _queries(g, n, 0xF5)
    counts = torch.zeros(n, dtype=torch.int32, device=DEV)
    num = torch.zeros(1, dtype=torch.int64, device=DEV)
    args = (_t(qm, torch.int32), _t(tf1, torch.float64), _t(tf2, torch.float64), 1, counts)
    with pytest.raises(F.FclxError) as e:
        c.collide_batch_async(*args, None, num, engine=F.ENGINE_PLAIN, noalloc=True)
    assert e.value.code == N.FCLX_ERR_INVALID_ARGUMENT
    pairs = torch.zeros((16, 3), dtype=torch.int32, device=DEV)
    geo = (torch.zeros((16, 3), dtype=torch.float64, device=DEV),
           torch.zeros((16, 3), dtype=torch.float64, device=DEV),
           torch.zeros(16, dtype=torch.float64, device=DEV))
    with pytest.raises(F.FclxError) as e:
        c.collide_batch_async(*args, pairs, num, engine=F.ENGINE_BOOL, contact=geo, noalloc=True)
    assert e.value.code == N.FCLX_ERR_UNSUPPORTED
    a = N.CollideBatchArgs()
    a.num_queries = 1 << 22
    a.query_meshes = a.tf1 = a.tf2 = ctypes.c_void_p(1)
    a.num_max_contacts = 1
    a.engine = N.FCLX_ENGINE_BOOL
    a.counts = a.num_pairs = ctypes.c_void_p(1)
    assert N.lib().fclx_collide_batch_noalloc(c._h, ctypes.byref(a), None) == \
        N.FCLX_ERR_UNSUPPORTED
