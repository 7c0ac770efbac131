#!/usr/bin/env python3
"""Benchmark: batched uneven mesh collision (exhaustive contacts) on one GPU.

(--workload primitives instead measures batched primitive-shape collide:
4,000,000 pairs of box / sphere / capsule / cylinder / convex with
enable_contact, GJK/EPA and the closed forms in FP64; see bench_primitives.)

Workload: 64 mesh pairs (perturbed icospheres / heightfield terrains,
500-200,000 triangles log-uniform), 1,000,000 pose queries split by a
Dirichlet(1) draw, num_max_contacts = 100,000 (the full overlapping
triangle-pair set per query).  A step is one FAST-engine collide batch over
all queries with inputs resident in HBM.  Reported beside it: the plain
one-query-per-lane FP64 GPU traversal on the same batch, with exact
agreement of per-query contact counts and pair-set checksums, and the CPU
oracle (the C restatement of FCL's path, OpenMP over host cores) on a
bounded sample of the same queries.

Prints one JSON line.
"""
import argparse
import concurrent.futures as cf
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

PEAK_FP32 = 157.3e12   # MI355X vector FP32, FLOP/s
PEAK_FP64 = 78.6e12    # MI355X vector FP64, FLOP/s
PEAK_HBM = 8.0e12      # B/s
OPS_BV = 315           # FP32 ops of one OBB test (126 transform + 189 SAT)
OPS_LEAF = 923         # FP64 ops of one triangle SAT
BYTES_BV = 128         # two 64-B FP32 node records
BYTES_LEAF = 144       # six FP64 vertices
BYTES_BV64 = 128       # one FP64 OBB node record (mesh-vs-primitive descent)
BYTES_LEAF_TRI = 128   # one triangle's three FP64 vertices (128-B record)
OPS_RSS = 380          # ops of one RSS lower bound (126 transform + ~250 rectDistance)
OPS_TRI_DIST = 900     # FP64 ops of one triangle distance

# HBM bytes the depth-first kernel moves per step (summed over its launches,
# one per round) come from rocprofv3 --pmc FETCH_SIZE and WRITE_SIZE passes
# over this workload, each scaled by the reading of the same counter for a
# device copy of known size (python bench.py --copy-calibration).  The
# summary with the commands is profiles/engine_dfs_traffic.md; its
# machine-readable line, in profiles/engine_dfs_traffic.txt, is parsed by
# measured_traffic().
TRAFFIC_FILE = os.path.join(ROOT, "profiles", "engine_dfs_traffic.txt")


def measured_traffic(queries, pairs):
    """{"bytes", "kernel_ms", "source"} of the profiled engine_dfs traffic for
    this workload size, or None: from the line
    "traffic: queries=Q pairs=P bytes_per_step=B kernel_ms_per_step=T" of the
    profile summary (T: the kernel's summed time per step in the profiled
    run)."""
    import re

    try:
        text = open(TRAFFIC_FILE).read()
    except OSError:
        return None
    for m in re.finditer(r"^traffic: queries=(\d+) pairs=(\d+) bytes_per_step=(\S+) "
                         r"kernel_ms_per_step=(\S+)\s*$", text, re.M):
        if int(m.group(1)) == queries and int(m.group(2)) == pairs:
            return {"bytes": float(m.group(3)), "kernel_ms": float(m.group(4)),
                    "source": "profiles/engine_dfs_traffic.txt: rocprofv3 --pmc FETCH_SIZE and "
                              "WRITE_SIZE, scaled by a copy of known size"}
    return None


def measured_valu(queries, pairs):
    """{"insts_per_step", "busy"} of engine_dfs for this workload size, or
    None: from the line "valu: queries=Q pairs=P insts_per_step=I
    busy_profiled=B" of the profile summary (rocprofv3 --pmc SQ_INSTS_VALU,
    SQ_ACTIVE_INST_VALU and GRBM_GUI_ACTIVE; B is AMD's VALUBusy in that
    run, as a fraction)."""
    import re

    try:
        text = open(TRAFFIC_FILE).read()
    except OSError:
        return None
    for m in re.finditer(r"^valu: queries=(\d+) pairs=(\d+) insts_per_step=(\S+) "
                         r"busy_profiled=(\S+)\s*$", text, re.M):
        if int(m.group(1)) == queries and int(m.group(2)) == pairs:
            return {"insts_per_step": float(m.group(3)), "busy": float(m.group(4))}
    return None


def cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return None


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def host_threads():
    """Host threads for the CPU baseline: the CPUs this process may run on,
    capped by OMP_NUM_THREADS when it is set."""
    n = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else (os.cpu_count() or 1)
    omp = os.environ.get("OMP_NUM_THREADS", "")
    if omp.isdigit() and int(omp) > 0:
        n = min(n, int(omp))
    return max(1, n)


def pair_checksum(torch, recs, nq):
    """Per-query sum of a hash of (b1, b2): equal sums with equal counts
    identify equal pair sets for the parity check between engines."""
    q = recs[:, 0].long()
    b1 = recs[:, 1].long()
    b2 = recs[:, 2].long()
    h = b1 * 1000003 + b2 * 998244353 + (b1 * b2) * 7 + 1
    out = torch.zeros(nq, dtype=torch.int64, device=recs.device)
    out.scatter_add_(0, q, h)
    return out


# FP64 operations of one loop pass, counted along the restatement's common
# path (adds, multiplies, divisions and square roots one each; compares not
# counted).  A support point is 39 + the two shapes' supports, a shape's
# support on this workload's type mix 38 on average (box 3, sphere 3,
# capsule 16, cylinder 9, convex of 32 points 160).  GJK pass: ray norm 6,
# normalisation 9, support point 115, repeat test 36, progress test 11,
# projection about 130, new ray 18.  EPA pass: normalisation 9, support
# point 115, distance test 6, horizon walk 4.6 faces x 7, 3.5 new faces x 77
# (edge distances, normal, offset).  Closed-form pair: about 120 on average.
OPS_GJK_PASS = 330
OPS_EPA_PASS = 460
OPS_CLOSED_PAIR = 120
# Algorithmic bytes of one pair: its two shape ids (8), two 40-B shape
# records, two 96-B poses, the vertices of its convex shapes (32 x 24 B each;
# 0.4 convexes per pair on this workload), and the results: count (4) and one
# 56-B contact slot.
BYTES_PRIM_PAIR = 8 + 80 + 192 + 0.4 * 768 + 4 + 56


def bench_primitives(args):
    """Batched primitive-shape collide: 4,000,000 pairs, enable_contact,
    num_max_contacts 1; a step is one fclx_shape_collide_batch over all
    pairs with inputs resident in HBM."""
    import numpy as np
    import torch

    import fcl_amd as F
    from fcl_amd import workloads

    dev = torch.device("cuda:0")
    n = args.prim_pairs
    t0 = time.time()
    b = workloads.primitive_batch(n)
    log("workload: %d primitive pairs generated in %.1fs" % (n, time.time() - t0))
    ctx = F.Context(0)
    for kv in args.engine_param:
        k, v = kv.split("=", 1)
        ctx.set_engine_param(k, float(v))
    rec = torch.from_numpy(b["shapes"].view(np.uint8)).to(dev)
    pool = torch.from_numpy(b["vertices"]).to(dev)
    ps = torch.from_numpy(b["pair_shapes"]).to(dev)
    tf1 = torch.from_numpy(b["tf1"]).to(dev)
    tf2 = torch.from_numpy(b["tf2"]).to(dev)
    cap = b["cap"]
    counts = torch.zeros(n, dtype=torch.int32, device=dev)
    geo = (torch.empty((n * cap, 3), dtype=torch.float64, device=dev),
           torch.empty((n * cap, 3), dtype=torch.float64, device=dev),
           torch.empty(n * cap, dtype=torch.float64, device=dev))
    passes = torch.zeros((n, 2), dtype=torch.int32, device=dev)
    stream = torch.cuda.current_stream()
    s = stream.cuda_stream

    def run(with_passes=False):
        ctx.shape_collide_batch_async(rec, pool, ps, tf1, tf2, cap, counts, contact=geo,
                                      stream=s, passes=passes if with_passes else None)

    for _ in range(max(args.warmup, 1)):
        run()
    ctx.sync(s)
    run(with_passes=True)
    ctx.sync(s)
    gjk_passes = int(passes[:, 0].sum().item())
    epa_passes = int(passes[:, 1].sum().item())
    types = b["shapes"]["type"]
    t1 = types[b["pair_shapes"][:, 0]]
    t2 = types[b["pair_shapes"][:, 1]]
    closed = int((((t1 == 1) & (t2 == 1)) | ((t1 == 0) & (t2 == 0)) |
                  ((t1 == 1) & np.isin(t2, [0, 2, 3])) |
                  ((t2 == 1) & np.isin(t1, [0, 2, 3]))).sum())
    hits = int((counts > 0).sum().item())
    # timed steps, HIP events on the stream the batch runs on
    ev = [torch.cuda.Event(enable_timing=True) for _ in range(2)]
    torch.cuda.synchronize()
    t_start = time.perf_counter()
    ev[0].record(stream)
    for _ in range(args.steps):
        run()
    ev[1].record(stream)
    ctx.sync(s)
    elapsed = time.perf_counter() - t_start
    gpu_ms = ev[0].elapsed_time(ev[1]) / args.steps
    ms_per_step = elapsed * 1e3 / args.steps
    value = n / (elapsed / args.steps)
    ops = gjk_passes * OPS_GJK_PASS + epa_passes * OPS_EPA_PASS + closed * OPS_CLOSED_PAIR
    alg_bytes = n * BYTES_PRIM_PAIR
    t_k = gpu_ms * 1e-3
    roofline = {
        "kernel": "shape pipeline (shape_gjk + shape_epa dominate)",
        "bound": "valu",
        "achieved": ops / t_k / 1e12,
        "peak": PEAK_FP64 / 1e12,
        "unit": "TFLOP/s",
        "frac": ops / t_k / PEAK_FP64,
        "traffic": None,
        "algorithmic_ops": ops,
        "ops_per_pass": {"gjk": OPS_GJK_PASS, "epa": OPS_EPA_PASS,
                         "closed_form_pair": OPS_CLOSED_PAIR},
        "gjk_passes": gjk_passes, "epa_passes": epa_passes, "closed_form_pairs": closed,
        "hbm_frac": alg_bytes / t_k / PEAK_HBM,
        "algorithmic_bytes": alg_bytes,
        "batch_ms_events": gpu_ms,
    }
    log("primitives: %.2f ms/step (events %.2f), %.3g pairs/s, %d colliding" %
        (ms_per_step, gpu_ms, value, hits))
    cpu = None
    if not args.no_cpu:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import fclo

        threads = host_threads()
        rng = np.random.default_rng(7)
        probe = np.sort(rng.choice(n, size=min(n, 20000), replace=False))
        qs = b["pair_shapes"]
        a = time.perf_counter()
        fclo.shape_collide_records(b["shapes"], b["vertices"], qs[probe], b["tf1"][probe],
                                   b["tf2"][probe], cap, True, threads=threads, contacts=False)
        probe_s = time.perf_counter() - a
        n_s = int(min(n, max(len(probe), len(probe) * args.cpu_seconds / max(probe_s, 1e-3))))
        sample = np.sort(rng.choice(n, size=n_s, replace=False))
        a = time.perf_counter()
        sc, _ = fclo.shape_collide_records(b["shapes"], b["vertices"], qs[sample],
                                           b["tf1"][sample], b["tf2"][sample], cap, True,
                                           threads=threads, contacts=False)
        cpu_s = time.perf_counter() - a
        match = bool(np.array_equal(sc, counts.cpu().numpy()[sample]))
        cpu = {"value": n_s / cpu_s, "unit": "pairs/s", "cores": threads, "kind": "port",
               "sample": "%d of the %d pairs, drawn uniformly (seed 7), OpenMP dynamic "
                         "schedule; contact counts equal the GPU's: %s" % (n_s, n, match)}
        log("cpu oracle: %d pairs in %.1fs on %d threads" % (n_s, cpu_s, threads))
    out = {
        "metric": "primitive collide pairs/sec (GJK/EPA + closed forms, enable_contact)",
        "value": value,
        "unit": "pairs/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded primitive shapes and poses)",
        "config": {"workload": "batched primitive GJK/EPA: 4M pairs of box/sphere/capsule/"
                               "cylinder/convex(32 pts), dims log-uniform [0.05,1], t in "
                               "[-1,1]^3, enable_contact, GST_INDEP",
                   "pairs": n, "num_max_contacts": cap, "colliding": hits},
        "roofline": roofline,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)
    ctx.close()


def bench_distance(args):
    """Mesh distance: two perturbed icospheres of 20,480 triangles in
    BVHModel<RSS>, 10,000 separated poses (|t| uniform in [2.1, 4] along a
    uniform direction), DistanceRequest with nearest points; one step = one
    distance batch over all poses, inputs resident."""
    import numpy as np
    import torch

    import fcl_amd as F

    n = args.dist_queries
    t0 = time.time()
    (V1, T1), (V2, T2) = F.icosphere(32, 0.15, 0xD0), F.icosphere(32, 0.15, 0xD1)
    tf1, tf2 = F.poses(n, 1, 2.1, 4.0, 0xD2)
    ctx = F.Context(0)
    for kv in args.engine_param:
        k, v = kv.split("=", 1)
        ctx.set_engine_param(k, float(v))
    g = np.array([ctx.mesh(V1, T1, F.RSS), ctx.mesh(V2, T2, F.RSS)], np.int32)
    log("distance workload: 2 x %d triangles (RSS), %d poses, built in %.1fs"
        % (len(T1), n, time.time() - t0))
    dev = "cuda:0"
    qm = torch.from_numpy(np.tile(g, (n, 1))).to(dev)
    d1, d2 = torch.from_numpy(tf1).to(dev), torch.from_numpy(tf2).to(dev)
    dist = torch.zeros(n, dtype=torch.float64, device=dev)
    pts = torch.zeros((n, 6), dtype=torch.float64, device=dev)
    ids = torch.zeros((n, 2), dtype=torch.int32, device=dev)
    bvt = torch.zeros(n, dtype=torch.int64, device=dev)
    lft = torch.zeros(n, dtype=torch.int64, device=dev)
    stream = torch.cuda.current_stream()
    s = stream.cuda_stream

    def run(stats=False, engine=F.ENGINE_FAST):
        ctx.distance_batch_async(qm, d1, d2, dist, pts, ids, engine=engine,
                                 bv_tests=bvt if stats else None,
                                 leaf_tests=lft if stats else None, stream=s)

    # the plain kernel: the reference's visit order; its test counts are the
    # reference's (the roofline's work), its distances the comparison
    a = time.perf_counter()
    run(True, F.ENGINE_PLAIN)
    ctx.sync(s)
    plain_s = time.perf_counter() - a
    plain_d = dist.clone()
    w_bv, w_leaf = int(bvt.sum().item()), int(lft.sum().item())
    run(True)
    ctx.sync(s)
    fast_bv, fast_leaf = int(bvt.sum().item()), int(lft.sum().item())
    exact = bool(torch.equal(dist, plain_d))
    tuned = None
    if args.autotune:
        from fcl_amd import autotune

        def checked():
            run()
            ctx.sync(s)
            return dist.cpu().numpy().tobytes()

        tuned, _, _ = autotune.tune(ctx, checked, autotune.DIST_SPACE, log=log)
    for _ in range(args.warmup):
        run()
    ctx.sync(s)
    ev = [torch.cuda.Event(enable_timing=True) for _ in range(2)]
    t_start = time.perf_counter()
    ev[0].record(stream)
    for _ in range(args.steps):
        run()
    ev[1].record(stream)
    ctx.sync(s)
    elapsed = time.perf_counter() - t_start
    gpu_ms = ev[0].elapsed_time(ev[1]) / args.steps
    ms_per_step = elapsed * 1e3 / args.steps
    value = n / (elapsed / args.steps)
    t_k = gpu_ms * 1e-3
    ops = w_bv * OPS_RSS + w_leaf * OPS_TRI_DIST
    alg_bytes = w_bv * BYTES_BV + w_leaf * BYTES_LEAF
    roofline = {
        "kernel": "engine_distance_wave (wavefront-wide, FP64)",
        "bound": "valu",
        "achieved": ops / t_k / 1e12,
        "peak": PEAK_FP64 / 1e12,
        "unit": "TFLOP/s",
        "frac": ops / t_k / PEAK_FP64,
        "traffic": None,
        "algorithmic_ops": ops,
        "ops_per_test": {"rss_bound": OPS_RSS, "triangle_distance": OPS_TRI_DIST},
        "bv_tests": w_bv, "leaf_tests": w_leaf,
        "executed_bv_tests": fast_bv, "executed_leaf_tests": fast_leaf,
        "hbm_frac": alg_bytes / t_k / PEAK_HBM,
        "algorithmic_bytes": alg_bytes,
        "batch_ms_events": gpu_ms,
    }
    # the same batch with FP32 culling (certified FP32 lower bounds first,
    # FP64 where both children of a node stay): same outputs, timed beside
    fp32 = None
    if tuned is None:
        d_ref = dist.clone()
        ctx.set_engine_param("dist_fp32_cull", 1)
        run(True)
        ctx.sync(s)
        f_bv, f_leaf = int(bvt.sum().item()), int(lft.sum().item())
        same = bool(torch.equal(dist, d_ref))
        e2 = [torch.cuda.Event(enable_timing=True) for _ in range(2)]
        e2[0].record(stream)
        for _ in range(args.steps):
            run()
        e2[1].record(stream)
        ctx.sync(s)
        ctx.set_engine_param("dist_fp32_cull", 0)
        fp32 = {"batch_ms_events": e2[0].elapsed_time(e2[1]) / args.steps,
                "bounds_per_query": f_bv / n, "triangle_tests_per_query": f_leaf / n,
                "distances_equal": same}
        log("fp32 culling: %.2f ms per batch, %.0f bounds (FP32 and FP64) and %.0f triangle "
            "tests per query, distances equal %s" % (fp32["batch_ms_events"], f_bv / n,
                                                     f_leaf / n, same))
        roofline["fp32_cull"] = fp32
    d = dist.cpu().numpy()
    log("distance: %.2f ms/step (events %.2f), %.3g queries/s; %.0f RSS and %.0f triangle "
        "tests per query (the reference's: %.0f and %.0f); %d of %d touching" %
        (ms_per_step, gpu_ms, value, fast_bv / n, fast_leaf / n, w_bv / n, w_leaf / n,
         int((d == 0).sum()), n))
    log("plain: %.1f ms, %.3g queries/s, distances equal %s" % (plain_s * 1e3, n / plain_s, exact))
    cpu = None
    if not args.no_cpu:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import fclo

        threads = host_threads()
        om = [fclo.OracleMesh(V1, T1, fclo.RSS), fclo.OracleMesh(V2, T2, fclo.RSS)]
        rng = np.random.default_rng(7)
        qpairs = np.tile(np.array([0, 1], np.int32), (n, 1))
        probe = np.sort(rng.choice(n, size=min(n, 200), replace=False))
        a = time.perf_counter()
        fclo.distance_batch(om, qpairs[probe], tf1[probe], tf2[probe], threads=threads)
        probe_s = time.perf_counter() - a
        n_s = int(min(n, max(len(probe), len(probe) * args.cpu_seconds / max(probe_s, 1e-3))))
        sample = np.sort(rng.choice(n, size=n_s, replace=False))
        a = time.perf_counter()
        od = fclo.distance_batch(om, qpairs[sample], tf1[sample], tf2[sample], threads=threads)[0]
        cpu_s = time.perf_counter() - a
        match = bool(np.array_equal(od, d[sample]))
        cpu = {"value": n_s / cpu_s, "unit": "queries/s", "cores": threads, "kind": "port",
               "sample": "%d of the %d poses, drawn uniformly (seed 7), OpenMP dynamic "
                         "schedule; distances bit-equal to the GPU's: %s" % (n_s, n, match)}
        log("cpu oracle: %d queries in %.1fs on %d threads, equal %s" % (n_s, cpu_s, threads,
                                                                         match))
    out = {
        "metric": "mesh distance queries/sec (DistanceRequest with nearest points)",
        "value": value,
        "unit": "queries/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded perturbed icospheres, separated poses)",
        "config": {"workload": "mesh distance: 2 icospheres of 20,480 tris, BVHModel<RSS>, "
                               "|t| in [2.1, 4], enable_nearest_points",
                   "queries": n, "engine_params": args.engine_param, "autotuned": tuned},
        "plain_gpu": {"value": n / plain_s, "ms": plain_s * 1e3},
        "speedup_vs_plain": value / (n / plain_s),
        "exact_match_vs_plain": exact,
        "roofline": roofline,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)
    ctx.close()


def bench_meshshape(args):
    """Mesh against primitives: a perturbed icosphere of 20,480 triangles in
    BVHModel<OBBRSS> (an environment) against robot links modelled as boxes,
    spheres, capsules, cylinders and a 7-vertex convex, posed with |t|
    uniform in [0.8, 1.2] along a uniform direction so that most of them
    straddle the surface; CollisionRequest defaults (num_max_contacts 1, no
    contact geometry, GST_INDEP).  One step = one batch over all queries,
    inputs resident."""
    import numpy as np
    import torch

    import fcl_amd as F

    n = args.ms_queries
    t0 = time.time()
    V, T = F.icosphere(32, 0.15, 0xE0)
    hull = np.array([[1, 0, 0], [-1, 0, 0], [0, 1.5, 0], [0, -1, 0], [0, 0, 0.7],
                     [0.2, 0.3, -1.1], [0.5, 0.5, 0.5]], np.float64) * 0.12
    kinds = [F.BOX, F.SPHERE, F.CAPSULE, F.CYLINDER, F.CONVEX]
    dims = [(0.3, 0.2, 0.4), (0.15, 0, 0), (0.08, 0.3, 0), (0.1, 0.3, 0), (0, 0, 0)]
    rec, pool = F.shape_table(kinds, dims, [hull])
    tf1, tf2 = F.poses(n, 1, 0.8, 1.2, 0xE1)
    names = ["box", "sphere", "capsule", "cylinder", "convex"]
    use = [names.index(k) for k in args.ms_kinds.split(",")]
    qs = np.array(use, np.int32)[np.arange(n) % len(use)]
    ctx = F.Context(0)
    for kv in args.engine_param:
        k, v = kv.split("=", 1)
        ctx.set_engine_param(k, float(v))
    mid = ctx.mesh(V, T)
    log("mesh-shape workload: %d triangles (OBBRSS) vs %d primitives, built in %.1fs"
        % (len(T), n, time.time() - t0))
    dev = "cuda:0"
    d_qm = torch.full((n,), mid, dtype=torch.int32, device=dev)
    d_qs = torch.from_numpy(qs).to(dev)
    d_rec = torch.from_numpy(np.ascontiguousarray(rec).view(np.uint8)).to(dev)
    d_pool = torch.from_numpy(pool).to(dev)
    d1, d2 = torch.from_numpy(tf1).to(dev), torch.from_numpy(tf2).to(dev)
    counts = torch.zeros(n, dtype=torch.int32, device=dev)
    stats = torch.zeros((n, 2), dtype=torch.int64, device=dev)
    stream = torch.cuda.current_stream()
    s = stream.cuda_stream

    def run(with_stats=False):
        ctx.mesh_shape_collide_batch_async(d_qm, d_rec, d_pool, d_qs, d1, d2, 1, counts,
                                           stats=stats if with_stats else None, stream=s)

    run(True)
    ctx.sync(s)
    w_bv, w_leaf = (int(x) for x in stats.sum(0).tolist())
    tuned = None
    if args.autotune:
        from fcl_amd import autotune

        def checked():
            run()
            ctx.sync(s)
            return counts.cpu().numpy().tobytes()

        tuned, _, _ = autotune.tune(ctx, checked, autotune.MESHSHAPE_SPACE, log=log)
    for _ in range(args.warmup):
        run()
    ctx.sync(s)
    ev = [torch.cuda.Event(enable_timing=True) for _ in range(2)]
    t_start = time.perf_counter()
    ev[0].record(stream)
    for _ in range(args.steps):
        run()
    ev[1].record(stream)
    ctx.sync(s)
    elapsed = time.perf_counter() - t_start
    gpu_ms = ev[0].elapsed_time(ev[1]) / args.steps
    ms_per_step = elapsed * 1e3 / args.steps
    value = n / (elapsed / args.steps)
    t_k = gpu_ms * 1e-3
    c = counts.cpu().numpy()
    # the same batch with contact geometry (enable_contact, up to 4 contacts
    # per query): EPA after GJK at every hit, timed once after a warm-up
    cap_c = 4
    geo = (torch.empty((n * cap_c, 3), dtype=torch.float64, device=dev),
           torch.empty((n * cap_c, 3), dtype=torch.float64, device=dev),
           torch.empty(n * cap_c, dtype=torch.float64, device=dev))
    counts_c = torch.zeros(n, dtype=torch.int32, device=dev)

    def run_contact():
        ctx.mesh_shape_collide_batch_async(d_qm, d_rec, d_pool, d_qs, d1, d2, cap_c, counts_c,
                                           contact=geo, stream=s)

    run_contact()
    ctx.sync(s)
    ev[0].record(stream)
    run_contact()
    ev[1].record(stream)
    ctx.sync(s)
    contact_ms = ev[0].elapsed_time(ev[1])
    log("with contact geometry (cap %d): %.2f ms, %d contacts" %
        (cap_c, contact_ms, int(counts_c.sum().item())))
    # FP64 work: an OBB test per BV test, and per leaf test one GJK loop
    # pass as a floor (a leaf runs one to a few tens of passes)
    ops = w_bv * OPS_BV + w_leaf * OPS_GJK_PASS
    alg_bytes = w_bv * BYTES_BV64 + w_leaf * BYTES_LEAF_TRI
    roofline = {
        "kernel": "meshshape_collide<false> (one query per lane, FP64)",
        "bound": "valu",
        "achieved": ops / t_k / 1e12,
        "peak": PEAK_FP64 / 1e12,
        "unit": "TFLOP/s",
        "frac": ops / t_k / PEAK_FP64,
        "traffic": None,
        "algorithmic_ops": ops,
        "ops_per_test": {"obb": OPS_BV, "leaf_gjk_pass_floor": OPS_GJK_PASS},
        "bv_tests": w_bv, "leaf_tests": w_leaf,
        "hbm_frac": alg_bytes / t_k / PEAK_HBM,
        "algorithmic_bytes": alg_bytes,
        "batch_ms_events": gpu_ms,
    }
    log("mesh-shape: %.2f ms/step (events %.2f), %.3g queries/s; %.1f BV and %.1f leaf tests "
        "per query; %d of %d collide" % (ms_per_step, gpu_ms, value, w_bv / n, w_leaf / n,
                                         int((c > 0).sum()), n))
    cpu = None
    if not args.no_cpu:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import fclo

        threads = host_threads()
        om = fclo.OracleMesh(V, T)
        rng = np.random.default_rng(7)
        zeros = np.zeros(n, np.int32)
        probe = np.sort(rng.choice(n, size=min(n, 2000), replace=False))
        a = time.perf_counter()
        fclo.mesh_shape_collide_batch([om], zeros[probe], rec, pool, qs[probe], tf1[probe],
                                      tf2[probe], threads=threads)
        probe_s = time.perf_counter() - a
        n_s = int(min(n, max(len(probe), len(probe) * args.cpu_seconds / max(probe_s, 1e-3))))
        sample = np.sort(rng.choice(n, size=n_s, replace=False))
        a = time.perf_counter()
        oc = fclo.mesh_shape_collide_batch([om], zeros[sample], rec, pool, qs[sample],
                                           tf1[sample], tf2[sample], threads=threads)[0]
        cpu_s = time.perf_counter() - a
        match = bool(np.array_equal(oc, c[sample]))
        cpu = {"value": n_s / cpu_s, "unit": "queries/s", "cores": threads, "kind": "port",
               "sample": "%d of the %d queries, drawn uniformly (seed 7), OpenMP dynamic "
                         "schedule; counts equal to the GPU's: %s" % (n_s, n, match)}
        log("cpu oracle: %d queries in %.1fs on %d threads, equal %s" % (n_s, cpu_s, threads,
                                                                         match))
    out = {
        "metric": "mesh-vs-primitive collide queries/sec (CollisionRequest defaults)",
        "value": value,
        "unit": "queries/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded perturbed icosphere, primitives at |t| in [0.8, 1.2])",
        "config": {"workload": "mesh-shape collide: icosphere of 20,480 tris (OBBRSS) vs "
                               "box/sphere/capsule/cylinder/convex links, num_max_contacts 1",
                   "queries": n, "kinds": args.ms_kinds, "engine_params": args.engine_param,
                   "autotuned": tuned},
        "colliding": int((c > 0).sum()),
        "contact_mode": {"num_max_contacts": cap_c, "ms": contact_ms,
                         "queries_per_s": n / (contact_ms * 1e-3)},
        "roofline": roofline,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


def first_contact_measure(ctx, steps, no_cpu, log_prefix="first contact"):
    """Plain mesh-vs-mesh collide, CollisionRequest(1, false): two perturbed
    icospheres of 2,420 triangles, 10,000 poses.  The FAST engine runs this
    cap through its ordered kernel (one query per lane, FP32 cull, FP64
    fallback and leaves, the reference's order); beside it the plain FP64
    kernel and the CPU oracle.  Returns a dict for the JSON line."""
    import numpy as np
    import torch

    import fcl_amd as F
    from fcl_amd import workloads

    W = workloads.mesh_collide_pair(10000)
    g = [ctx.mesh(V, T) for V, T in W["meshes"]]
    n = len(W["tf1"])
    dev = torch.device("cuda:0")
    d_qm = torch.from_numpy(np.tile(np.array(g, np.int32), (n, 1))).to(dev)
    d1 = torch.from_numpy(W["tf1"]).to(dev)
    d2 = torch.from_numpy(W["tf2"]).to(dev)
    counts = torch.zeros(n, dtype=torch.int32, device=dev)
    pairs = torch.zeros((n, 3), dtype=torch.int32, device=dev)
    num = torch.zeros(1, dtype=torch.int64, device=dev)
    stream = torch.cuda.current_stream()

    def timed(engine, reps):
        for _ in range(2):
            ctx.collide_batch_async(d_qm, d1, d2, 1, counts, pairs, num, engine=engine,
                                    stream=stream.cuda_stream)
        ctx.sync(stream.cuda_stream)
        a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        a.record(stream)
        for _ in range(reps):
            ctx.collide_batch_async(d_qm, d1, d2, 1, counts, pairs, num, engine=engine,
                                    stream=stream.cuda_stream)
        b.record(stream)
        ctx.sync(stream.cuda_stream)
        ms = a.elapsed_time(b) / reps
        rec = pairs[:int(num.item())].cpu().numpy()
        return ms, counts.cpu().numpy(), rec[np.argsort(rec[:, 0])]

    fast_ms, fc, fp = timed(F.ENGINE_FAST, max(steps, 5) * 4)
    plain_ms, pc, pp = timed(F.ENGINE_PLAIN, max(steps, 5))
    out = {"workload": "2 x 2420-tri icospheres, 10000 poses, num_max_contacts 1",
           "queries": n, "colliding": int((fc > 0).sum()),
           "ordered_ms": fast_ms, "ordered_queries_per_s": n / (fast_ms * 1e-3),
           "plain_ms": plain_ms, "plain_queries_per_s": n / (plain_ms * 1e-3),
           "exact_match_vs_plain": bool(np.array_equal(fc, pc) and np.array_equal(fp, pp))}
    log("%s: ordered %.3f ms, plain %.3f ms for %d queries" % (log_prefix, fast_ms, plain_ms, n))
    if not no_cpu:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import fclo

        om = [fclo.OracleMesh(V, T) for V, T in W["meshes"]]
        threads = host_threads()
        a = time.perf_counter()
        oc, op, obv, olf = fclo.collide_batch(om, W["query_meshes"], W["tf1"], W["tf2"], 1,
                                              store=1, threads=threads)
        cpu_s = time.perf_counter() - a
        a = time.perf_counter()
        fclo.collide_batch(om, W["query_meshes"][:2000], W["tf1"][:2000], W["tf2"][:2000], 1,
                           threads=1)
        one_s = time.perf_counter() - a
        hit = oc > 0
        out["exact_match_vs_oracle"] = bool(
            np.array_equal(fc, oc) and np.array_equal(fp[:, 1:], op[hit, 0]))
        out["cpu_baseline"] = {"value": n / cpu_s, "unit": "queries/s", "cores": threads,
                               "kind": "port", "sample": "all 10000 poses",
                               "one_thread": {"value": 2000 / one_s, "queries": 2000}}
        # roofline of the ordered kernel on the oracle's work counts: FP32
        # OBB tests (315 ops, 128 B) and FP64 triangle tests (923 ops, 144 B)
        t = fast_ms * 1e-3
        bv, lf = float(obv.sum()), float(olf.sum())
        valu = (bv * OPS_BV / PEAK_FP32 + lf * OPS_LEAF / PEAK_FP64) / t
        hbm = (bv * BYTES_BV + lf * BYTES_LEAF) / t / PEAK_HBM
        out["roofline"] = {"kernel": "ordered_collide", "bound": "valu", "valu_frac": valu,
                           "hbm_frac": hbm, "bv_tests_per_query": bv / n,
                           "leaf_tests_per_query": lf / n,
                           "note": "2 x 4839 nodes stay in cache: bound by VALU and latency"}
    for m in g:
        ctx.mesh_destroy(m)
    return out


def bench_first_contact(args):
    import fcl_amd as F

    ctx = F.Context(0)
    fc = first_contact_measure(ctx, args.steps, args.no_cpu)
    out = {
        "metric": "mesh collide queries/sec, first contact (CollisionRequest(1, false))",
        "value": fc["ordered_queries_per_s"], "unit": "queries/s", "n_gpus": 1,
        "steps": max(args.steps, 5) * 4, "warmup": 2, "ms_per_step": fc["ordered_ms"],
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32 cull + f64 leaves", "data": "synthetic (seeded perturbed icospheres)",
        "config": {"workload": fc["workload"]},
        "plain_gpu": {"value": fc["plain_queries_per_s"], "ms": fc["plain_ms"]},
        "speedup_vs_plain": fc["ordered_queries_per_s"] / fc["plain_queries_per_s"],
        "exact_match_vs_plain": fc["exact_match_vs_plain"],
        "exact_match_vs_oracle": fc.get("exact_match_vs_oracle"),
        "roofline": fc.get("roofline"), "cpu_baseline": fc.get("cpu_baseline"),
    }
    print(json.dumps(out), flush=True)
    ctx.close()


def serving_measure(ctx, scene, link_ids, pool, rates, seconds, streams, engine, seed=11,
                    max_batch=0):
    """Replays Poisson request streams (one per rate, `seconds` long) from
    the pool through a fresh request loop; per rate the p50 / p99 request
    latency (arrival to results in host memory) and the queries/s."""
    import numpy as np

    import fcl_amd as F

    nlink = len(link_ids)
    srv = F.Server(ctx, scene, num_streams=streams, engine=engine, max_batch_queries=max_batch)
    # warm-up: every pool request once, spread over 100 ms, so that first-use
    # costs (pool buffers, launch shapes, clocks) fall before the timed loads
    srv.replay([(link_ids, P) for P in pool], np.linspace(0, 100, len(pool)))
    rng = np.random.default_rng(seed)
    loads = []
    for rate in rates:
        gaps = rng.exponential(1000.0 / rate, int(rate * seconds * 1.5) + 16)
        arrivals = np.cumsum(gaps)
        arrivals = arrivals[arrivals < seconds * 1000.0]
        n = len(arrivals)
        idx = rng.integers(0, len(pool), n)
        reqs = [(link_ids, pool[i]) for i in idx]
        nq = np.array([pool[i].shape[0] * nlink for i in idx])
        b0, q0 = srv.stats()
        _, lat, elapsed = srv.replay(reqs, arrivals)
        que = srv.last_queue_ms
        b1, q1 = srv.stats()
        p50, p99 = np.percentile(lat, [50, 99])
        svc = lat - que
        worst = np.argsort(lat)[-4:][::-1]
        loads.append({"rate_per_s": rate, "requests": n, "queries": int(nq.sum()),
                      "p50_ms": float(p50), "p99_ms": float(p99), "max_ms": float(lat.max()),
                      "first_ms": [float(x) for x in lat[:4]],
                      "queue_p50_ms": float(np.percentile(que, 50)),
                      "queue_p99_ms": float(np.percentile(que, 99)),
                      "service_p99_ms": float(np.percentile(svc, 99)),
                      "worst": [{"index": int(i), "queries": int(nq[i]), "latency_ms": float(lat[i]),
                                 "queue_ms": float(que[i])} for i in worst],
                      "queries_per_s": float(nq.sum() / (elapsed / 1e3)),
                      "offered_queries_per_s": float(rate * nq.mean()),
                      "batches": b1 - b0, "elapsed_ms": elapsed})
        log("rate %5d/s: %5d requests, p50 %.2f ms, p99 %.2f ms, %.3g queries/s, %d batches"
            % (rate, n, p50, p99, nq.sum() / (elapsed / 1e3), b1 - b0))
    return srv, loads


def pinned_pool(pool):
    """The request pool's pose arrays in pinned host memory (as a serving
    front end keeps its request buffers): the loop copies them to the device
    straight from there."""
    import torch

    out = []
    for P in pool:
        t = torch.empty(P.shape, dtype=torch.float64, pin_memory=True)
        t.numpy()[...] = P
        out.append(t.numpy())
    return out


def serving_summary(rates, seconds, streams, no_cpu):
    """The serving workload inside the default benchmark line: the 500k-tri
    scene and 20-link robot, p50 / p99 at the given loads, and the booleans
    of four pool requests against the oracle's."""
    import numpy as np

    import fcl_amd as F
    from fcl_amd import workloads

    sV, sT = workloads.serving_scene()
    links = workloads.serving_robot()
    pool = pinned_pool(workloads.serving_requests(64))
    ctx = F.Context(0)
    scene = ctx.mesh(sV, sT)
    link_ids = np.array([ctx.mesh(V, T) for V, T in links], np.int32)
    srv, loads = serving_measure(ctx, scene, link_ids, pool, rates, seconds, streams,
                                 F.ENGINE_BOOL)
    out = {"workload": "500k-tri scene, 20 links of 1k-50k tris, pool of 64 trajectories of "
                       "64-4096 waypoints, Poisson arrivals, %d streams, BOOL engine" % streams,
           "seconds_per_rate": seconds, "loads": loads}
    if not no_cpu:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import fclo

        with cf.ThreadPoolExecutor(max_workers=min(16, os.cpu_count() or 1)) as ex:
            om = list(ex.map(lambda m: fclo.OracleMesh(m[0], m[1]), [(sV, sT)] + links))
        reqs = pool[:4]
        got, _, _ = srv.replay([(link_ids, P) for P in reqs], np.zeros(len(reqs)))
        ok = True
        for P, r in zip(reqs, got):
            nq = P.shape[0] * len(links)
            qm = np.stack([np.zeros(nq, np.int32),
                           1 + np.tile(np.arange(len(links), dtype=np.int32), P.shape[0])], 1)
            tf1 = np.tile(np.array([1, 0, 0, 0, 1, 0, 0, 0, 1, 0, 0, 0], float), (nq, 1))
            c = fclo.collide_batch(om, qm, tf1, P.reshape(nq, 12), 1,
                                   threads=host_threads())[0]
            ok &= bool(np.array_equal((c > 0).astype(np.uint8).reshape(r.shape), r))
        out["booleans_equal_oracle"] = ok
        out["checked_queries"] = int(sum(P.shape[0] for P in reqs) * len(links))
    srv.close()
    ctx.close()
    return out


def bench_serving(args):
    """Robot-trajectory serving: one resident ~500k-triangle scene, a
    20-link robot, requests of 64-4096 waypoints arriving as Poisson
    streams; per load the p50 / p99 request latency (arrival to results in
    host memory) and the queries/s the loop sustained."""
    import numpy as np

    import fcl_amd as F
    from fcl_amd import workloads

    t0 = time.time()
    sV, sT = workloads.serving_scene()
    links = workloads.serving_robot()
    pool = workloads.serving_requests(args.serve_pool)
    if not args.serve_pageable:
        pool = pinned_pool(pool)
    log("serving workload: scene %d tris, %d links (%d-%d tris), pool of %d requests "
        "(mean %.0f waypoints), generated in %.1fs"
        % (len(sT), len(links), min(len(l[1]) for l in links), max(len(l[1]) for l in links),
           len(pool), np.mean([P.shape[0] for P in pool]), time.time() - t0))
    ctx = F.Context(0)
    for kv in args.engine_param:
        k, v = kv.split("=", 1)
        ctx.set_engine_param(k, float(v))
    t0 = time.time()
    scene = ctx.mesh(sV, sT)
    link_ids = np.array([ctx.mesh(V, T) for V, T in links], np.int32)
    log("uploaded BVHs in %.1fs" % (time.time() - t0))
    nlink = len(links)
    # the BOOL engine on one batch of 16 pool requests, inputs resident: time
    # and, in a profiled run, its work (fclx_engine_last_report)
    engine_batch = None
    tuned = None
    if args.serve_engine == "bool":
        import torch

        P = np.ascontiguousarray(np.concatenate([p.reshape(-1, 12) for p in pool[:16]]))
        nq = len(P)
        dev = "cuda:0"
        dqm = torch.from_numpy(np.stack([np.full(nq, scene, np.int32),
                                         np.tile(link_ids, nq // nlink)], 1)).to(dev)
        dtf1 = torch.from_numpy(np.tile(np.array([1, 0, 0, 0, 1, 0, 0, 0, 1, 0, 0, 0], float),
                                        (nq, 1))).to(dev)
        dtf2 = torch.from_numpy(P).to(dev)
        cnt = torch.zeros(nq, dtype=torch.int32, device=dev)
        num = torch.zeros(1, dtype=torch.int64, device=dev)

        def one():
            ctx.collide_batch_async(dqm, dtf1, dtf2, 1, cnt, None, num, engine=F.ENGINE_BOOL)
            ctx.sync()

        for _ in range(3):
            one()
        if args.autotune:
            from fcl_amd import autotune

            def checked():
                one()
                return int(cnt.sum().item())

            tuned, _, _ = autotune.tune(ctx, checked, autotune.BOOL_SPACE, log=log)
        ts = []
        for _ in range(5):
            a = time.perf_counter()
            one()
            ts.append(time.perf_counter() - a)
        ctx.set_profiling(True)
        one()
        r = ctx.engine_report()
        ctx.set_profiling(False)
        engine_batch = {"queries": nq, "ms": 1e3 * min(ts), "queries_per_s": nq / min(ts),
                  "hit_fraction": float((cnt > 0).float().mean().item()),
                  "fp32_bv_tests_per_query": r["bool_lanes"][0] / nq,
                  "fp64_triangle_tests_per_query": r["bool_lanes"][1] / nq,
                  "fp64_bv_tests_per_query": r["bool_lanes"][2] / nq,
                  "lanes_per_bv_step": r["bool_lanes"][0] / max(1, r["bool_steps"][0]),
                  "wavefront_ms": r["bool_wave_ms"], "wavefront_end_p50_ms": r["bool_end_p50_ms"],
                  "wavefront_end_p90_ms": r["bool_end_p90_ms"], "last_end_ms": r["bool_span_ms"],
                  "deepest_stack": r["bool_stack_max"]}
        e = engine_batch
        log("bool engine, %d queries resident: %.2f ms (%.3g queries/s), %.0f FP32 BV tests and "
            "%.0f triangle tests per query" % (nq, e["ms"], e["queries_per_s"],
                                              e["fp32_bv_tests_per_query"],
                                              e["fp64_triangle_tests_per_query"]))
    engine = {"bool": F.ENGINE_BOOL, "fast": F.ENGINE_FAST, "plain": F.ENGINE_PLAIN}[
        args.serve_engine]
    srv, loads = serving_measure(ctx, scene, link_ids, pool, args.serve_rates,
                                 args.serve_seconds, args.serve_streams, engine,
                                 max_batch=args.serve_batch)
    # one request alone, the largest of the pool
    big = max(pool, key=lambda P: P.shape[0])
    _, lat1, _ = srv.replay([(link_ids, big)], [0.0])
    srv.close()
    cpu = None
    if not args.no_cpu:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import fclo

        threads = host_threads()
        om = [fclo.OracleMesh(sV, sT)] + [fclo.OracleMesh(V, T) for V, T in links]
        P = np.concatenate([p.reshape(-1, 12) for p in pool[:4]])
        nq = len(P)
        qm = np.stack([np.zeros(nq, np.int32),
                       1 + np.tile(np.arange(nlink, dtype=np.int32), nq // nlink)], 1)
        tf1 = np.tile(np.array([1, 0, 0, 0, 1, 0, 0, 0, 1, 0, 0, 0], float), (nq, 1))
        a = time.perf_counter()
        fclo.collide_batch(om, qm, tf1, P, 1, store=0, threads=threads)
        cpu_s = time.perf_counter() - a
        cpu = {"value": nq / cpu_s, "unit": "queries/s", "cores": threads, "kind": "port",
               "sample": "the %d link-waypoint queries of 4 pool requests (collide, cap 1), "
                         "OpenMP dynamic schedule" % nq}
        log("cpu oracle: %d queries in %.1fs on %d threads" % (nq, cpu_s, threads))
    top = loads[-1]
    out = {
        "metric": "robot-trajectory serving: p99 request latency at Poisson load",
        "value": top["p99_ms"],
        "unit": "ms",
        "n_gpus": 1,
        "higher_is_better": False,
        "vs_baseline": None,
        "dtype": "f32 cull + f64 leaves",
        "data": "synthetic (seeded heightfield + boxes scene, capsule-like links, random-walk "
                "trajectories)",
        "config": {"workload": "serving: 500k-tri scene, 20 links of 1k-50k tris, 64-4096 "
                               "waypoints log-uniform, CollisionRequest(1, false) per link and "
                               "waypoint", "streams": args.serve_streams,
                   "engine": args.serve_engine, "engine_params": args.engine_param,
                   "rate_per_s": top["rate_per_s"], "pool": len(pool)},
        "loads": loads,
        "largest_request_alone_ms": float(lat1[0]),
        "largest_request_queries": int(big.shape[0] * nlink),
        "bool_engine_batch": engine_batch,
        "autotuned": tuned,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)
    ctx.close()


def main():
    ap = argparse.ArgumentParser(description=__doc__)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--workload",
                    choices=["mesh", "primitives", "serving", "distance", "meshshape",
                             "first-contact"],
                    default="mesh",
                    help="mesh: the batched uneven mesh collision (default); primitives: "
                         "4M primitive-shape pairs; serving: robot-trajectory requests at "
                         "Poisson loads; distance: 10k mesh distance queries (20k-tri RSS); "
                         "meshshape: 200k primitives against a 20k-tri mesh; first-contact: "
                         "10k first-contact queries between 2420-tri icospheres")
    ap.add_argument("--skip-extras", action="store_true",
                    help="mesh workload: leave out the first-contact and serving sub-objects")
    ap.add_argument("--serve-rates", type=float, nargs="+",
                    default=[100, 500, 1000, 2000, 5000])
    ap.add_argument("--serve-seconds", type=float, default=1.0,
                    help="arrival window per load")
    ap.add_argument("--serve-streams", type=int, default=6)
    ap.add_argument("--serve-batch", type=int, default=0,
                    help="queries per serving batch at most (0: the loop's default, 65536)")
    ap.add_argument("--serve-pool", type=int, default=64)
    ap.add_argument("--serve-pageable", action="store_true",
                    help="keep the request poses in pageable memory (the loop then packs "
                         "them through its pinned staging) instead of pinned memory")
    ap.add_argument("--serve-engine", choices=["bool", "fast", "plain"], default="bool",
                    help="engine behind the request loop (bool: booleans only)")
    ap.add_argument("--prim-pairs", type=int, default=4_000_000)
    ap.add_argument("--dist-queries", type=int, default=10_000)
    ap.add_argument("--ms-queries", type=int, default=200_000)
    ap.add_argument("--ms-kinds", default="box,sphere,capsule,cylinder,convex",
                    help="primitive kinds of the meshshape workload, used in turn")
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--queries", type=int, default=1_000_000)
    ap.add_argument("--pairs", type=int, default=64)
    ap.add_argument("--plain", choices=["full", "skip"], default="full",
                    help="time the plain FP64 kernel on the same batch")
    ap.add_argument("--cpu-seconds", type=float, default=15.0,
                    help="target CPU time of the oracle sample")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--engine-param", action="append", default=[], metavar="NAME=VALUE",
                    help="FAST-engine tunable (fclx_set_engine_param); repeatable")
    ap.add_argument("--autotune", action="store_true",
                    help="tune the engine's tunables on the workload before timing "
                         "(fcl_amd.autotune; the setting found goes into the JSON config)")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="only run three device copies of MIB MiB (a known-size kernel "
                         "for scaling rocprofv3 FETCH_SIZE / WRITE_SIZE) and exit")
    args = ap.parse_args()

    world = int(os.environ.get("WORLD_SIZE", "1"))
    if args.gpus != 1 or world != 1:
        print("bench.py measures one GPU only: this build has no multi-GPU path "
              "(queries would shard by index across GPUs).", file=sys.stderr)
        sys.exit(2)

    if args.workload == "primitives":
        bench_primitives(args)
        return
    if args.workload == "serving":
        bench_serving(args)
        return
    if args.workload == "distance":
        bench_distance(args)
        return
    if args.workload == "meshshape":
        bench_meshshape(args)
        return
    if args.workload == "first-contact":
        bench_first_contact(args)
        return

    import numpy as np
    import torch

    if args.copy_calibration > 0:
        n = args.copy_calibration << 20
        src = torch.ones(n, dtype=torch.uint8, device="cuda:0")
        dst = torch.empty_like(src)
        for _ in range(3):
            dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"copy_bytes": n, "launches": 3}), flush=True)
        return

    import fcl_amd as F
    from fcl_amd import workloads

    dev = torch.device("cuda:0")
    t0 = time.time()
    W = workloads.uneven_batch(args.queries, args.pairs)
    log("workload: %d meshes, %d queries, generated in %.1fs"
        % (len(W["meshes"]), args.queries, time.time() - t0))
    ctx = F.Context(0)
    for kv in args.engine_param:
        k, v = kv.split("=", 1)
        ctx.set_engine_param(k, float(v))
    t0 = time.time()
    with cf.ThreadPoolExecutor(max_workers=min(16, os.cpu_count() or 1)) as ex:
        ids = list(ex.map(lambda m: ctx.mesh(m[0], m[1]), W["meshes"]))
    log("uploaded %d BVHs in %.1fs" % (len(ids), time.time() - t0))
    ids = np.array(ids, np.int32)
    qm = ids[W["query_meshes"]]
    nq = qm.shape[0]
    d_qm = torch.from_numpy(qm).to(dev)
    d_tf1 = torch.from_numpy(W["tf1"]).to(dev)
    d_tf2 = torch.from_numpy(W["tf2"]).to(dev)
    counts = torch.zeros(nq, dtype=torch.int32, device=dev)
    num = torch.zeros(1, dtype=torch.int64, device=dev)
    cap = W["cap"]
    stream = torch.cuda.current_stream()
    s = stream.cuda_stream

    capacity = [max(1 << 20, 200 * nq)]
    pairs = [torch.empty((capacity[0], 3), dtype=torch.int32, device=dev)]

    def run(engine):
        while True:
            ctx.collide_batch_async(d_qm, d_tf1, d_tf2, cap, counts, pairs[0], num,
                                    engine=engine, stream=s)
            rc = ctx.sync_status(s)
            if rc == F._native.FCLX_ERR_CAPACITY:
                need = int(num.item())
                capacity[0] = int(need * 1.1) + 1024
                log("growing pair buffer to %d records" % capacity[0])
                pairs[0] = None
                torch.cuda.empty_cache()
                pairs[0] = torch.empty((capacity[0], 3), dtype=torch.int32, device=dev)
                continue
            if rc != 0:
                raise RuntimeError("collide batch failed: %d" % rc)
            return

    # ---------------------------------------------------------------- FAST
    for _ in range(max(args.warmup, 1)):
        run(F.ENGINE_FAST)
    tuned = None
    if args.autotune:
        from fcl_amd import autotune

        def one():
            run(F.ENGINE_FAST)
            return int(num.item()), int(counts.sum().item())

        tuned, tuned_ms, trials = autotune.tune(ctx, one, autotune.FAST_SPACE, repeats=1, log=log)
        log("autotune: %d trials, best %s at %.1f ms" % (len(trials), tuned, tuned_ms))
    ctx.set_profiling(True)
    run(F.ENGINE_FAST)
    rep = ctx.engine_report()
    ctx.set_profiling(False)
    total_pairs = int(num.item())
    fast_counts = counts.clone()
    fast_sum = pair_checksum(torch, pairs[0][:total_pairs], nq)
    step_ms = []
    torch.cuda.synchronize()
    t_start = time.perf_counter()
    for _ in range(args.steps):
        a = time.perf_counter()
        run(F.ENGINE_FAST)
        step_ms.append((time.perf_counter() - a) * 1e3)
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t_start
    ms_per_step = elapsed * 1e3 / args.steps
    value = nq / (elapsed / args.steps)
    log("FAST: %.1f ms/step, %.3g queries/s, %d pairs, report %s" % (ms_per_step, value,
                                                                    total_pairs, rep))

    # --------------------------------------------------------------- plain
    plain = None
    exact = None
    if args.plain == "full":
        torch.cuda.synchronize()
        a = time.perf_counter()
        run(F.ENGINE_PLAIN)
        plain_s = time.perf_counter() - a
        plain_pairs = int(num.item())
        exact = bool(torch.equal(counts, fast_counts)) and plain_pairs == total_pairs and \
            bool(torch.equal(pair_checksum(torch, pairs[0][:plain_pairs], nq), fast_sum))
        plain = {"value": nq / plain_s, "ms": plain_s * 1e3}
        log("plain: %.1f ms, %.3g queries/s, exact match %s" % (plain_s * 1e3, nq / plain_s, exact))

    # ------------------------------------------------------------- roofline
    # The depth-first kernel does the FP32 BV tests of the exact traversal
    # that the breadth-first phase left, and the FP64 triangle tests of the
    # leaf pairs it queues, 64 at a time (the remainder of each round's
    # queues goes to engine_leaves and is not charged here).  Its bound is
    # HBM: the achieved rate is the bytes it moves (rocprofv3 counters of the
    # same workload, profiles/engine_dfs_traffic.md) over its time in this
    # run (HIP events, summed over the step's launches).  The algorithmic
    # bytes (128 B per BV test, 144 B per triangle test) are given beside it:
    # node records are re-used from L2, so their rate is not bounded by the
    # HBM peak.  The VALU fraction prices the BV tests at 315 FP32 operations
    # and the triangle tests at 923 FP64 operations, each against its peak.
    t_dfs = rep["dfs_ms"] * 1e-3
    inline_leaf = rep.get("dfs_inline_leaf_tests", 0)
    alg_bytes = rep["dfs_bv_tests"] * BYTES_BV + inline_leaf * BYTES_LEAF
    valu_frac = None
    if t_dfs > 0:
        valu_frac = (rep["dfs_bv_tests"] * OPS_BV / PEAK_FP32 +
                     inline_leaf * OPS_LEAF / PEAK_FP64) / t_dfs
    traffic = measured_traffic(nq, args.pairs)
    valu = measured_valu(nq, args.pairs)
    achieved = traffic["bytes"] / t_dfs if traffic and t_dfs > 0 else None
    frac = achieved / PEAK_HBM if achieved else None
    if frac is not None and frac > 1:
        log("measured traffic exceeds the HBM peak: profile out of date, roofline left null")
        achieved = frac = None
    roofline = {
        "kernel": "engine_dfs",
        "bound": "hbm",
        "achieved": achieved / 1e9 if achieved else None,
        "peak": PEAK_HBM / 1e9,
        "unit": "GB/s",
        "frac": frac,
        "traffic": traffic["bytes"] if traffic else None,
        "traffic_source": traffic["source"] if traffic else None,
        # the same bytes over the kernel's time inside the profiled run
        "frac_in_profiled_run": (traffic["bytes"] / (traffic["kernel_ms"] * 1e-3) / PEAK_HBM
                                 if traffic else None),
        "algorithmic_bytes": alg_bytes,
        "algorithmic_rate_gbs": alg_bytes / t_dfs / 1e9 if t_dfs > 0 else None,
        "valu_frac": valu_frac,
        "valu_peak_tflops": PEAK_FP32 / 1e12,
        # vector issue, the limit the kernel is closest to (counters of the
        # same workload): wave instructions per step, and AMD's VALUBusy in
        # the counter run (a kernel of independent FP32 FMAs reads 0.87)
        "valu_wave_insts_per_step": valu["insts_per_step"] if valu else None,
        "valu_busy_in_profiled_run": valu["busy"] if valu else None,
        "kernel_ms": rep["dfs_ms"],
        "bv_tests": rep["dfs_bv_tests"],
        "lanes_per_step": (rep["dfs_bv_tests"] / rep["dfs_wave_steps"]
                           if rep.get("dfs_wave_steps") else None),
        "leaf_tests_in_kernel": inline_leaf,
        "leaf_tests_in_engine_leaves": rep["dfs_leaf_tests"] - inline_leaf,
        "fp64_fallback_rate": rep["fp64_fallbacks"] / max(1, rep["dfs_bv_tests"] +
                                                           rep["expand_bv_tests"]),
        "phase_ms": {k: rep[k] for k in ("setup_ms", "expand_ms", "leaves_ms", "dfs_ms",
                                         "fixup_ms")},
        "dfs_balance": {k: rep[k] for k in ("dfs_waves", "dfs_wave_busy", "dfs_end_p50_ms",
                                            "dfs_end_p90_ms", "dfs_end_max_ms")},
    }

    # ------------------------------------------------------------ CPU oracle
    cpu = None
    if not args.no_cpu:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import fclo

        threads = host_threads()
        rng = np.random.default_rng(7)
        # size the sample from a probe so the run takes about cpu_seconds
        probe = np.sort(rng.choice(nq, size=min(nq, 2000), replace=False))
        t0 = time.time()
        with cf.ThreadPoolExecutor(max_workers=min(16, threads)) as ex:
            omeshes = list(ex.map(lambda m: fclo.OracleMesh(m[0], m[1]), W["meshes"]))
        log("oracle BVHs built in %.1fs" % (time.time() - t0))
        Q = W["query_meshes"]
        a = time.perf_counter()
        pc, _, _, _ = fclo.collide_batch(omeshes, Q[probe], W["tf1"][probe], W["tf2"][probe], cap,
                                         threads=threads)
        probe_s = time.perf_counter() - a
        n_s = int(min(nq, max(len(probe), len(probe) * args.cpu_seconds / max(probe_s, 1e-3))))
        sample = np.sort(rng.choice(nq, size=n_s, replace=False))
        a = time.perf_counter()
        sc, _, _, _ = fclo.collide_batch(omeshes, Q[sample], W["tf1"][sample], W["tf2"][sample],
                                         cap, threads=threads)
        cpu_s = time.perf_counter() - a
        # one thread, on the first queries of the same sample (about a third
        # of the multi-threaded run's CPU time)
        n1 = int(min(n_s, max(200, n_s * args.cpu_seconds / 3 / max(cpu_s * threads, 1e-3))))
        s1 = sample[:n1]
        a = time.perf_counter()
        c1, _, _, _ = fclo.collide_batch(omeshes, Q[s1], W["tf1"][s1], W["tf2"][s1], cap,
                                         threads=1)
        one_s = time.perf_counter() - a
        gpu_c = fast_counts.cpu().numpy()
        cpu_match = bool(np.array_equal(sc, gpu_c[sample])) and \
            bool(np.array_equal(pc, gpu_c[probe])) and bool(np.array_equal(c1, gpu_c[s1]))
        cpu = {"value": n_s / cpu_s, "unit": "queries/s", "cores": threads, "kind": "port",
               "cpu_model": cpu_model(),
               "sample": "%d of the %d queries, drawn uniformly (seed 7), OpenMP dynamic "
                         "schedule; contact counts equal the GPU's: %s" % (n_s, nq, cpu_match),
               "one_thread": {"value": n1 / one_s, "unit": "queries/s", "queries": n1}}
        log("cpu oracle: %d queries in %.1fs on %d threads; %d in %.1fs on one"
            % (n_s, cpu_s, threads, n1, one_s))

    # ------------------------------------------- first contact and serving
    first_contact = serving = None
    if not args.skip_extras:
        first_contact = first_contact_measure(ctx, args.steps, args.no_cpu)
        serving = serving_summary([1000, 2000], 1.0, args.serve_streams, args.no_cpu)

    p = np.percentile(step_ms, [50, 99]) if step_ms else [None, None]
    out = {
        "metric": "mesh collide queries/sec at exact-match vs plain GPU; p50/p99 latency",
        "value": value,
        "unit": "queries/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 cull + f64 leaves",
        "data": "synthetic (seeded perturbed icospheres and heightfields)",
        "config": {"workload": "batched uneven mesh collision: 64 pairs, 500-200k tris "
                               "log-uniform, Dirichlet(1) query split, exhaustive contacts",
                   "queries": nq, "pairs": args.pairs, "num_max_contacts": cap,
                   "contact_pairs": total_pairs, "engine_params": args.engine_param,
                   "autotuned": tuned},
        "p50_ms": p[0],
        "p99_ms": p[1],
        "plain_gpu": plain,
        "speedup_vs_plain": (value / plain["value"]) if plain else None,
        "exact_match_vs_plain": exact,
        "roofline": roofline,
        "cpu_baseline": cpu,
        # request latency of the serving workload (arrival to booleans in
        # host memory) at 1000 and 2000 requests/s
        "request_latency_ms": ({str(int(l["rate_per_s"])): {"p50": l["p50_ms"], "p99": l["p99_ms"]}
                                for l in serving["loads"]} if serving else None),
        "first_contact": first_contact,
        "serving": serving,
    }
    print(json.dumps(out), flush=True)
    ctx.close()


if __name__ == "__main__":
    main()
