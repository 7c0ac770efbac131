"""fcl_amd.autotune on the GPU: a few candidate values of FAST- and
BOOL-engine tunables on a small batch.  Every candidate must give the same
results (the tuner checks the batch's return value), and the best setting
is left applied to the context (the test restores the defaults after)."""
import numpy as np
import pytest
import torch

import fcl_amd as F
from fcl_amd import autotune

pytestmark = pytest.mark.gpu
DEV = "cuda:0"


def test_tune_fast_and_bool(ctx):
    (V1, T1), (V2, T2) = F.icosphere(11, 0.15, 0xA0), F.icosphere(11, 0.15, 0xB0)
    g = np.array([ctx.mesh(V1, T1), ctx.mesh(V2, T2)], np.int32)
    n = 4000
    tf1, tf2 = F.poses(n, 0, 1.5, 0, 0xF1)
    print("case: autotune, %d queries" % n, flush=True)
    qm = torch.from_numpy(np.tile(g, (n, 1))).to(DEV)
    d1, d2 = torch.from_numpy(tf1).to(DEV), torch.from_numpy(tf2).to(DEV)
    counts = torch.zeros(n, dtype=torch.int32, device=DEV)
    num = torch.zeros(1, dtype=torch.int64, device=DEV)
    pairs = torch.zeros((1 << 22, 3), dtype=torch.int32, device=DEV)
    s = torch.cuda.current_stream().cuda_stream

    def exhaustive():
        ctx.collide_batch_async(qm, d1, d2, 100000, counts, pairs, num, engine=F.ENGINE_FAST,
                                stream=s)
        assert ctx.sync_status(s) == 0
        return int(num.item()), int(counts.sum().item())

    def boolean():
        ctx.collide_batch_async(qm, d1, d2, 1, counts, None, num, engine=F.ENGINE_BOOL, stream=s)
        assert ctx.sync_status(s) == 0
        return counts.cpu().numpy().tobytes()

    saved = {name: ctx.get_engine_param(name)
             for name, _ in autotune.FAST_SPACE + autotune.BOOL_SPACE}
    try:
        fs = (("front_target", (1 << 12, 1 << 16)), ("max_levels", (6, 24)))
        best, ms, trials = autotune.tune(ctx, exhaustive, fs, repeats=1)
        assert ms > 0 and 1 <= len(trials) <= 5
        for name, values in fs:
            assert best[name] in [float(v) for v in values] + [saved[name]]
            assert ctx.get_engine_param(name) == best[name]
        bs = (("bool_queries_per_wave", (4, 16)), ("bool_width", (8, 64)))
        best, ms, trials = autotune.tune(ctx, boolean, bs, repeats=1)
        for name, _ in bs:
            assert ctx.get_engine_param(name) == best[name]
    finally:
        for k, v in saved.items():
            ctx.set_engine_param(k, v)
    assert ctx.get_engine_param("bool_width") == saved["bool_width"]
