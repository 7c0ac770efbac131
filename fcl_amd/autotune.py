"""Autotuner for the engine's tunables (fclx_set_engine_param).

The engines' results do not depend on their tunables (front size, levels
of breadth-first expansion, grid size, round-stop share, query slots per
wavefront, per-query width, ...), only their speed does, and the best
setting depends on the workload: mesh sizes, how deep the poses
interpenetrate, batch size.  `tune` runs coordinate descent over a small
grid per parameter: it times a caller-supplied batch (the workload itself,
or a sample of it) under each candidate value of one parameter while the
others stay at their best so far, keeps a value only when it is faster by
a margin, and leaves the best setting applied to the context.  If the batch
function returns a value (a checksum of the results), every candidate must
return the same one.
"""
import time

# Candidate values per parameter, in the order the descent visits them.
FAST_SPACE = (
    ("front_target", (1 << 16, 1 << 18, 1 << 20, 1 << 22)),
    ("max_levels", (16, 20, 24, 28)),
    ("dfs_blocks_per_cu", (4, 6, 8)),
    ("idle_stop", (0.25, 0.5, 0.75)),
    ("stop_poll", (8, 16, 32)),
    ("dfs_inline_leaves", (0, 1)),
    ("dfs_stack_top", (256, 512)),
    ("fast_persistent", (0, 1)),
)

DIST_SPACE = (
    ("dist_queries_per_wave", (2, 4, 8)),
    ("dist_waves_per_simd", (2, 3)),
    ("dist_width", (16, 32, 64)),
    ("dist_leaf_batch", (8, 16, 32, 64)),
    ("dist_fp32_cull", (0, 1)),
)

MESHSHAPE_SPACE = (
    ("meshshape_waves_per_simd", (1, 2, 3)),
    ("meshshape_leaf_share", (0, 25, 50, 75, 100)),
)

BOOL_SPACE = (
    ("bool_queries_per_wave", (0, 4, 8, 16)),
    ("bool_width", (16, 32, 64)),
    ("bool_leaf_batch", (32, 64)),
    ("bool_refill", (2, 4, 8)),
)


class ResultChanged(AssertionError):
    """A tunable changed what the batch returned: never expected."""


def tune(ctx, run, space, repeats=3, rounds=1, margin=0.02, log=None):
    """Coordinate descent over `space` ((name, values) pairs).

    run: executes one batch on the context and waits for it; may return a
    value that every setting must reproduce.  Each candidate is run once to
    warm up, then `repeats` times; its time is the minimum.  A candidate
    replaces the current value only if it is faster by `margin` (relative).

    Returns (best: {name: value}, best_ms, trials: [({name: value}, ms)]);
    the best setting is left applied.
    """
    best = {name: ctx.get_engine_param(name) for name, _ in space}
    reference = []

    def measure():
        out = run()
        if out is not None:
            if not reference:
                reference.append(out)
            elif out != reference[0]:
                raise ResultChanged("results differ under %r" % (current(),))
        ts = []
        for _ in range(repeats):
            t0 = time.perf_counter()
            run()
            ts.append(time.perf_counter() - t0)
        return 1e3 * min(ts)

    def current():
        return {name: ctx.get_engine_param(name) for name, _ in space}

    best_ms = measure()
    trials = [(dict(best), best_ms)]
    if log:
        log("tune: start %s -> %.2f ms" % (best, best_ms))
    for _ in range(rounds):
        for name, values in space:
            keep = best[name]
            for v in values:
                if float(v) == float(keep):
                    continue
                ctx.set_engine_param(name, v)
                ms = measure()
                trials.append(({**best, name: float(v)}, ms))
                if log:
                    log("tune: %s=%g -> %.2f ms" % (name, v, ms))
                if ms < best_ms * (1.0 - margin):
                    best_ms, keep = ms, float(v)
            ctx.set_engine_param(name, keep)
            best[name] = keep
    if log:
        log("tune: best %s -> %.2f ms" % (best, best_ms))
    return best, best_ms, trials
