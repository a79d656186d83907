"""Benchmark: exact embeddings/s and p50 time per query on one MI355X.

Workload: the mid-size power-law Chung-Lu data graph (n=36,692,
m=183,831, exponent 2.1, degree cap 1,400, seed 0) resident in HBM, and its
two 6-vertex queries (a 4-cycle with two pendants; K4 minus an edge with a
2-edge tail).  One step = matching both queries exactly (uint64 counts) with
the heuristic score order.  value = total embeddings / total device time of
the timed steps.

The dominant kernel's roofline comes from HIP events recorded around every
launch on the library's own stream during the timed steps
(CUTS_FLAG_PROFILE).  Its algorithmic bytes are the loads its algorithm
issues, counted on the device per work item (list words, signatures,
membership probes, table entries, two-hop lookups).  The merge model
(8 + 4*len per distinct backward list of every prefix, 16 per trie node;
the sum over all kernels equals the CPU trie walk's figure) is reported for
the whole path beside it.  The CPU baseline times the OpenMP port of the
same trie algorithm (oracle/, test infrastructure) on a bounded sample of
level 0.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu]

Other workloads (one JSON line each; the default above is the headline):
    --workload chunglu4k   triangle + 2-edge tail on Chung-Lu n=4,096 (full count)
    --workload rmat20      5-cycle + chord + 2 tails on R-MAT scale 20: the full
                           count exceeds 2^64, so the query is matched exactly
                           on a nested root sample of position 0 sized to 2^36
                           candidates
    --workload batch256    256 random-walk queries, batched level queue in one hipGraph
                           launch, learned orders, autotuned launch shape
    --workload learned     train the table-size network on R-MAT scale 18,
                           evaluate learned vs score order on R-MAT scale 20
"""
import argparse
import json
import os
import statistics
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
BATCH_ARENA = 48 << 30   # two halves of 24 GiB: the largest trie level of the whole batch
BATCH_QUEUE = 1 << 26    # work items per level


def cpu_baseline(g, shapes, budget_s=20.0):
    """OpenMP trie baseline (same algorithm, same order) on a level-0 sample:
    every 64th level-0 vertex, stopped between row groups after
    budget_s/len(shapes) seconds per query; the rate counts only embeddings
    of sub-tries that were finished."""
    import oracle
    import cuts_amd
    from cuts_amd import workloads as W
    og = oracle.OGraph.from_view(g.view())
    # the GPU box grants 16 host cores per GPU whatever os.cpu_count() says
    threads = min(16, os.cpu_count() or 1)
    total_emb, total_s, samples = 0, 0.0, []
    for shape in shapes:
        qa = oracle.qmatrix(shape.k, W.shape_arcs(shape))
        order = cuts_amd.Plan(W.query_graph(shape)).order
        stride = 64
        r = oracle.trie_cpu(og, qa, order, sample_stride=stride, max_rows=1 << 22, nthreads=threads,
                            max_seconds=budget_s / len(shapes))
        total_emb += r.count
        total_s += r.seconds
        samples.append(f"{shape.name}: every {stride}th level-0 vertex, {'all' if r.complete else 'time-capped'}; "
                       f"{r.count} embeddings in {r.seconds:.2f}s")
    return dict(value=total_emb / total_s if total_s else 0.0, unit="embeddings/s", cores=threads, kind="port",
                sample="; ".join(samples))


def traffic_for(kernel):
    """HBM bytes per dispatch from the committed rocprofv3 --pmc summary."""
    path = os.path.join(ROOT, "profiles", "traffic.json")
    if not os.path.exists(path):
        return None
    with open(path) as f:
        t = json.load(f)
    e = t.get(kernel)
    return None if e is None else e.get("bytes_per_dispatch")


def emit(line):
    print(json.dumps(line), flush=True)


def log(msg):
    print(f"bench: {msg}", file=sys.stderr, flush=True)


def base_line(metric, value, unit, steps, warmup, ms_per_step, workload, extra):
    d = {"metric": metric, "value": value, "unit": unit, "n_gpus": 1, "steps": steps, "warmup": warmup,
         "ms_per_step": ms_per_step, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
         "dtype": "u32", "data": "synthetic (seeded generators; queries located by seeded random walk)",
         "config": {"workload": workload}}
    d["config"].update(extra)
    return d


def run_chunglu4k(args):
    import cuts_amd
    from cuts_amd import workloads as W
    ctx = cuts_amd.Context(0)
    dg = ctx.upload(W.config1_graph())
    p = cuts_amd.Plan(W.query_graph(W.TRI_TAIL2))
    for _ in range(args.warmup):
        ctx.match(dg, p)
    ms, counts = [], []
    for _ in range(args.steps):
        r = ctx.match(dg, p)
        ms.append(r.device_ms)
        counts.append(r.count)
    emit(base_line("exact embeddings/sec + p50 time/query", sum(counts) / (sum(ms) / 1e3), "embeddings/s",
                   args.steps, args.warmup, sum(ms) / args.steps,
                   "chung-lu n=4096 m=32768 gamma=2.3 seed=0; triangle+2-edge-tail; heuristic order",
                   {"count": counts[0], "p50_ms_per_query": statistics.median(ms)}))


def run_rmat20(args, max_work=1 << 36):
    """Config 3: the full count exceeds 2^64, so the query is matched exactly
    on the densest nested root sample of position 0 (v = s-1 mod s, s a
    power of two) whose run stays within max_work candidates -- found by a
    binary search of capped runs, then timed."""
    import cuts_amd
    from cuts_amd import workloads as W
    g = W.config3_graph()
    ctx = cuts_amd.Context(0)
    dg = ctx.upload(g)
    p = cuts_amd.Plan(W.query_graph(W.C5CHORD_2TAIL))
    log(f"R-MAT 20 uploaded ({g.arcs} arcs)")
    lo, hi = 0, 20
    while lo < hi:  # smallest stride exponent whose sample fits (nested samples: monotone)
        mid = (lo + hi) // 2
        r = ctx.match(dg, p, roots=W.root_sample(g.V, 1 << mid), max_work=max_work)
        log(f"stride 2^{mid}: {'over the cap' if r.work_capped else f'{r.work} candidates'}")
        if r.work_capped:
            lo = mid + 1
        else:
            hi = mid
    roots = W.root_sample(g.V, 1 << hi)
    for _ in range(args.warmup):
        ctx.match(dg, p, roots=roots)
    runs = [ctx.match(dg, p, roots=roots, flags=cuts_amd.FLAG_PROFILE) for _ in range(max(1, args.steps))]
    r = runs[-1]
    ms = statistics.median(x.device_ms for x in runs)
    dname, (dms, dn, dbytes) = max(r.kernels.items(), key=lambda kv: kv[1][0])
    ach = dbytes / dn / (dms / dn / 1e3) / 1e9 if dn else 0.0  # demand model, as for the headline
    line = base_line("exact embeddings/sec + p50 time/query", r.count / (ms / 1e3), "embeddings/s",
                     len(runs), args.warmup, ms,
                     "rmat scale=20 ef=16 a,b,c=.57,.19,.19 seed=0; 5-cycle+chord+2-tails; score order; exact "
                     "count on the densest nested root sample of position 0 within 2^36 candidates",
                     {"roots": list(roots), "count": r.count, "level_sizes": r.level_sizes, "work": r.work,
                      "p50_ms_per_query": ms})
    line["roofline"] = {"bound": "hbm", "kernel": dname, "achieved": ach, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                        "frac": ach / HBM_PEAK_GBS, "traffic": None, "avg_launch_ms": dms / max(dn, 1),
                        "launches": dn, "bytes_model": "demand model (loads the kernel's algorithm issues)",
                        "merge_model_GBs": r.merge_bytes[dname] / dn / (dms / dn / 1e3) / 1e9 if dn else 0.0}
    emit(line)


def load_batch_fixture():
    path = os.path.join(ROOT, "tests", "golden", "batch_queries.json")
    with open(path) as f:
        return json.load(f)


def run_batch256(args):
    import numpy as np
    import cuts_amd
    from cuts_amd import autotune, learned as LN, workloads as W
    g = W.config2_graph()
    ctx = cuts_amd.Context(0)
    dg = ctx.upload(g)
    fx = load_batch_fixture()
    qs = [(q["k"], [tuple(e) for e in q["edges"]]) for q in fx["queries"]]
    roots = [tuple(q["roots"]) for q in fx["queries"]]
    heur = [cuts_amd.Plan(cuts_amd.Graph.from_edges(k, e)) for k, e in qs]
    # heuristic order, one query at a time (each on its root sample)
    for p, rt in zip(heur[:8], roots):
        ctx.match(dg, p, roots=rt)
    single = [ctx.match(dg, p, roots=rt) for p, rt in zip(heur, roots)]
    s_ms = [r.device_ms for r in single]
    counts = [r.count for r in single]
    assert counts == [q["count"] for q in fx["queries"]], "single-query counts differ from the fixture"
    log(f"single heuristic: p50 {statistics.median(s_ms):.3f} ms, total {sum(s_ms):.1f} ms")
    if args.score_only:
        # score orders only, default launch shape (for profiling the level kernel)
        bh = cuts_amd.Batch(ctx, dg, heur, arena_bytes=BATCH_ARENA, queue_cap=BATCH_QUEUE, roots=roots,
                            use_graph=not args.plain_launches)
        for _ in range(args.warmup):
            bh.run()
        runs = [bh.run() for _ in range(args.steps)]
        assert all(r.counts == counts for r in runs), "batched counts differ from single-query counts"
        ms = statistics.median(r.device_ms for r in runs)
        emit(base_line("exact embeddings/sec + p50 time/query", sum(counts) / (ms / 1e3), "embeddings/s",
                       args.steps, args.warmup, ms,
                       "256 random-walk queries (k 4-12 in equal quotas) on chung-lu n=36692, each exact on its "
                       "admitted root sample of position 0; score orders; level-synchronous batch queue",
                       {"queries": len(qs), "sum_count": sum(counts), "batch_ms": ms,
                        "amortised_ms_per_query": ms / len(qs), "p50_ms_single_heuristic": statistics.median(s_ms),
                        "plain_launches": bool(args.plain_launches)}))
        return
    # learned orders from a network trained on tables logged on other queries
    # of the same graph (same admission rule, another seed)
    train, tried = W.admit(ctx, dg, g, min(args.train_queries, 1024), seed=1)
    log(f"admitted {len(train)} training queries (tries per size {tried})")
    net, info = train_on_queries(ctx, dg, g, [(q["k"], q["edges"], q["work"], tuple(q["roots"])) for q in train],
                                 seed=1, log=log)
    gfeat = LN.graph_features(g)
    choices = []
    for i in range(0, len(qs), 32):
        choices += LN.choose_orders(net, qs[i:i + 32], gfeat, [p.order for p in heur[i:i + 32]], keep_first=True)
    # the score order judges every proposal (one run capped at its work)
    judged = LN.judge(ctx, dg, qs, heur, single, choices, roots=roots)
    plans = [j.plan for j in judged]
    proposed = sum(1 for c in choices if c.learned)
    kept = sum(1 for j in judged if j.learned)
    work_l = sum(j.work for j in judged)
    work_h = sum(r.work for r in single)
    fallback = 1.0 - kept / len(qs)
    log(f"learned proposals {proposed}, kept by the judge {kept}; candidates {work_l} vs score order {work_h}")
    key = f"batch256:{W.csr_fingerprint(g)}:learned"
    tb = autotune.tune_batch(ctx, dg, plans, key, use_cache=False, log=log, arena_bytes=BATCH_ARENA,
                             queue_cap=BATCH_QUEUE, roots=roots)
    best = tb["best"]
    b = cuts_amd.Batch(ctx, dg, plans, blocks_per_cu=best["blocks_per_cu"], item_work=best["item_work"],
                       arena_bytes=BATCH_ARENA, queue_cap=BATCH_QUEUE, roots=roots)
    runs = []
    for _ in range(args.warmup):
        b.run()
    for _ in range(args.steps):
        r = b.run()
        assert r.counts == counts, "batched learned-order counts differ from single-query score-order counts"
        runs.append(r)
    bh = cuts_amd.Batch(ctx, dg, heur, blocks_per_cu=best["blocks_per_cu"], item_work=best["item_work"],
                        arena_bytes=BATCH_ARENA, queue_cap=BATCH_QUEUE, roots=roots)
    rh = [bh.run() for _ in range(2)][-1]
    assert rh.counts == counts
    ms = statistics.median([r.device_ms for r in runs])
    lat = runs[-1].latency_ms
    cum = list(np.cumsum(s_ms))
    amort = ms / len(qs)
    emit(base_line("exact embeddings/sec + p50 time/query; learned vs heuristic order, same GPU",
                   sum(counts) / (ms / 1e3), "embeddings/s", args.steps, args.warmup, ms,
                   "256 random-walk queries (k 4-12 in equal quotas, edge/vertex 1.0-2.5) on chung-lu n=36692, "
                   "each exact on its admitted root sample of order position 0; level-synchronous batch queue "
                   "replayed as one hipGraph; learned orders (same first vertex); autotuned launch shape",
                   {"queries": len(qs), "sum_count": sum(counts),
                    "batch_learned_ms": ms, "batch_heuristic_ms": rh.device_ms,
                    "amortised_ms_per_query_batch_learned": amort,
                    "p50_ms_single_heuristic": statistics.median(s_ms),
                    "p99_ms_single_heuristic": sorted(s_ms)[int(0.99 * len(s_ms))],
                    "speedup_p50_amortised_batch_learned_vs_single_heuristic": statistics.median(s_ms) / amort,
                    "p50_latency_ms_batch_learned_all_submitted_at_t0": statistics.median(lat),
                    "p99_latency_ms_batch_learned_all_submitted_at_t0": sorted(lat)[int(0.99 * len(lat))],
                    "p50_latency_ms_single_heuristic_sequential_all_submitted_at_t0": statistics.median(cum),
                    "speedup_p50_latency": statistics.median(cum) / statistics.median(lat),
                    "fallback_rate": fallback, "learned_proposed": proposed, "learned_kept": kept,
                    "candidates_learned": work_l, "candidates_heuristic": work_h,
                    "tuned": best, "training": info}))


def train_on_queries(ctx, dg, g, queries, seed, log):
    """Tables logged by the solver itself: for each (k, edges, work, roots)
    the score order's run on its root sample plus one random connected
    order's run with the same first vertex (capped at 4x the score order's
    work).  Sampled table sizes are scaled by the root stride, so targets
    estimate full tables."""
    import numpy as np
    import cuts_amd
    from cuts_amd import learned as LN
    rng = np.random.default_rng(seed)
    samples, targets = [], []
    t0 = time.perf_counter()
    for k, edges, work, rt in queries:
        q = cuts_amd.Graph.from_edges(k, edges)
        adj = LN.adjacency(k, edges)
        h = cuts_amd.Plan(q)
        extra = [LN.random_connected_order(k, adj, rng, first=h.order[0]) for _ in range(3)]
        for order, cap in [(None, 0)] + [(o, 4 * work + 4096) for o in extra]:
            p = cuts_amd.Plan(q, order)
            r = ctx.match(dg, p, max_work=cap, roots=rt)
            if r.work_capped:
                continue
            for kk, a, sm, size in LN.table_samples(k, edges, p.order, r.level_sizes):
                samples.append((kk, a, sm))
                targets.append(LN.target(size * rt[2]))
    t1 = time.perf_counter()
    x, A, meta = LN.encode(samples, LN.graph_features(g))
    net = LN.Mpnn(ctx, seed=seed)
    hist = net.fit(x, A, meta, np.asarray(targets, np.float32), epochs=60, batch=2048, lr=2e-3)
    t2 = time.perf_counter()
    log(f"trained on {len(samples)} samples of {len(queries)} queries: mse {hist[0]:.4f} -> {hist[-1]:.4f}")
    return net, {"queries": len(queries), "samples": len(samples), "mse_first": hist[0], "mse_last": hist[-1],
                 "log_s": t1 - t0, "train_s": t2 - t1}


def run_learned(args):
    """Train on R-MAT scale-18 graphs, evaluate on the R-MAT scale-20 graph.
    Queries of both sets are admitted on nested root samples of position 0
    (cuts_amd.workloads.admit); learned orders keep the score order's first
    vertex and are judged by it."""
    import numpy as np
    import cuts_amd
    from cuts_amd import learned as LN, workloads as W
    ctx = cuts_amd.Context(0)
    samples, targets = [], []
    rng = np.random.default_rng(0)
    per_graph = max(1, args.train_queries // 4)
    for gs in (1, 2, 3, 4):
        g18 = W.rmat18_graph(gs)
        dg18 = ctx.upload(g18)
        gf = LN.graph_features(g18)
        qs, tried = W.admit(ctx, dg18, g18, per_graph, seed=100 + gs, kmin=5, kmax=10)
        for q in qs:
            k, edges, rt = q["k"], q["edges"], tuple(q["roots"])
            adj = LN.adjacency(k, edges)
            qg = cuts_amd.Graph.from_edges(k, edges)
            for order, cap in ((None, 0), (LN.random_connected_order(k, adj, rng, first=q["order"][0]),
                                           4 * q["work"] + 4096)):
                p = cuts_amd.Plan(qg, order)
                r = ctx.match(dg18, p, roots=rt, max_work=cap)
                if r.work_capped:
                    continue
                for kk, a, sm, size in LN.table_samples(k, edges, p.order, r.level_sizes):
                    x1, A1, m1 = LN.encode([(kk, a, sm)], gf)
                    samples.append((x1[0], A1[0], m1[0]))
                    targets.append(LN.target(size * rt[2]))
        log(f"R-MAT 18 seed {gs}: {len(qs)} queries admitted (tries {tried}), {len(samples)} samples so far")
        del dg18
    x = np.stack([s_[0] for s_ in samples])
    A = np.stack([s_[1] for s_ in samples])
    meta = np.asarray([s_[2] for s_ in samples], np.uint32)
    net = LN.Mpnn(ctx, seed=0)
    hist = net.fit(x, A, meta, np.asarray(targets, np.float32), epochs=60, batch=2048, lr=2e-3)
    log(f"trained: mse {hist[0]:.4f} -> {hist[-1]:.4f} on {len(meta)} samples")
    if args.save_weights:
        np.save(args.save_weights, net.params())
    g = W.config3_graph()
    dg = ctx.upload(g)
    gf = LN.graph_features(g)
    evq, tried = W.admit(ctx, dg, g, args.eval_queries, seed=7, kmin=6, kmax=10)
    log(f"admitted {len(evq)} evaluation queries (tries {tried})")
    qs = [(q["k"], [tuple(e) for e in q["edges"]]) for q in evq]
    roots = [tuple(q["roots"]) for q in evq]
    heur = [cuts_amd.Plan(cuts_amd.Graph.from_edges(k, e)) for k, e in qs]
    single = [ctx.match(dg, p, roots=rt) for p, rt in zip(heur, roots)]
    choices = []
    for j in range(0, len(qs), 32):
        choices += LN.choose_orders(net, qs[j:j + 32], gf, [p.order for p in heur[j:j + 32]], keep_first=True)
    judged = LN.judge(ctx, dg, qs, heur, single, choices, roots=roots)
    h_ms, l_ms = [], []
    for p, jd, rt in zip(heur, judged, roots):
        rh = ctx.match(dg, p, roots=rt)  # timed again, warm
        h_ms.append(rh.device_ms)
        l_ms.append(ctx.match(dg, jd.plan, roots=rt).device_ms if jd.learned else rh.device_ms)
    ratio = [a / b for a, b in zip(h_ms, l_ms) if b > 0]
    geo = float(np.exp(np.mean(np.log(ratio)))) if ratio else 0.0
    counts = [r.count for r in single]
    proposed = sum(1 for c in choices if c.learned)
    kept = sum(1 for jd in judged if jd.learned)
    line = base_line("exact embeddings/sec + p50 time/query; learned vs heuristic order, same GPU",
                     sum(counts) / (sum(l_ms) / 1e3), "embeddings/s", 1, 0, sum(l_ms),
                     "train: 4 R-MAT scale-18 graphs (seeds 1-4), 5-10 vertex walk queries; evaluate: 6-10 vertex "
                     "walk queries on R-MAT scale 20 (seed 0); every query exact on its admitted root sample of "
                     "position 0; learned orders keep the first vertex and are judged by the score order",
                     {"train_samples": int(len(meta)), "mse_first": hist[0], "mse_last": hist[-1],
                      "eval_queries": len(evq), "tried": tried, "sum_count": sum(counts),
                      "p50_ms_heuristic": statistics.median(h_ms), "p50_ms_learned": statistics.median(l_ms),
                      "speedup_p50": statistics.median(h_ms) / statistics.median(l_ms),
                      "speedup_total": sum(h_ms) / sum(l_ms), "speedup_geomean": geo,
                      "learned_proposed": proposed, "learned_kept": kept,
                      "fallback_rate": 1.0 - kept / max(1, len(evq)),
                      "candidates_learned": sum(jd.work for jd in judged),
                      "candidates_heuristic": sum(r.work for r in single)})
    emit(line)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--copy-probe", type=int, default=0,
                    help="also run N dispatches of a 1 GiB plain copy (byte-counter yardstick)")
    ap.add_argument("--workload", default="chunglu36k",
                    choices=["chunglu36k", "chunglu4k", "rmat20", "batch256", "learned"])
    ap.add_argument("--train-queries", type=int, default=2000)
    ap.add_argument("--eval-queries", type=int, default=500)
    ap.add_argument("--save-weights", default="")
    ap.add_argument("--score-only", action="store_true", help="batch256: score orders only, no training")
    ap.add_argument("--plain-launches", action="store_true", help="batch256 --score-only: no hipGraph")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: only one GPU is used by this build (asked for --gpus {args.gpus}); "
              "the matcher shards by level-0 chunks, see DESIGN.md", file=sys.stderr)
        sys.exit(2)
    if args.workload != "chunglu36k":
        {"chunglu4k": run_chunglu4k, "rmat20": run_rmat20, "batch256": run_batch256,
         "learned": run_learned}[args.workload](args)
        return

    import cuts_amd
    from cuts_amd import workloads as W

    g = W.config2_graph()
    shapes = [W.C4_2PEND, W.K4E_TAIL2]
    ctx = cuts_amd.Context(0)
    t_up = time.perf_counter()
    dg = ctx.upload(g)
    upload_s = time.perf_counter() - t_up
    plans = [cuts_amd.Plan(W.query_graph(s)) for s in shapes]

    for i in range(args.warmup):
        for p in plans:
            ctx.match(dg, p)
        print(f"bench: warmup step {i + 1}/{args.warmup} done", file=sys.stderr, flush=True)

    per_query_ms, counts, work = [], [], []
    kern = {}
    t0 = time.perf_counter()
    for _ in range(args.steps):
        for p in plans:
            r = ctx.match(dg, p, flags=cuts_amd.FLAG_PROFILE)
            per_query_ms.append(r.device_ms)
            counts.append(r.count)
            work.append(r.work)
            for name, (ms, n, b) in r.kernels.items():
                k = kern.setdefault(name, [0.0, 0, 0.0, 0.0])
                k[0] += ms
                k[1] += n
                k[2] += b
                k[3] += r.merge_bytes[name]
        print(f"bench: timed step done, last query {per_query_ms[-1]:.2f} ms", file=sys.stderr, flush=True)
    wall = time.perf_counter() - t0
    total_ms = sum(per_query_ms)
    total_emb = sum(counts)
    dom = max(kern.items(), key=lambda kv: kv[1][0])
    dname, (dms, dn, dbytes, dmerge) = dom
    avg_launch_s = dms / dn / 1e3
    # algorithmic bytes of the dominant kernel: the loads its algorithm
    # issues, counted on the device per work item (demand model)
    bytes_per_launch = dbytes / dn
    model = ("demand model, counted on the device: per candidate its list word and signature (12 B); per "
             "membership test a hub's id and bitmap word (12 B) or the offsets and words of a lower-bound search; "
             "per prefix its row, selector and scan entry; per closed form its table entries; per two-hop lookup "
             "the offsets and the search's words")
    note = ("the graph (1.6 MB CSR, signatures, hub bitmaps) and the two-hop index stay largely cache-resident, so "
            "the kernel is bound by on-chip latency rather than HBM bandwidth (traffic: HBM bytes from rocprofv3); "
            "merge_model_* restate the whole path in the merge model, whose closed-form levels are never read")
    achieved = bytes_per_launch / avg_launch_s / 1e9
    line = {
        "metric": "exact embeddings/sec + p50 time/query; learned vs heuristic order, same GPU",
        "value": total_emb / (total_ms / 1e3),
        "unit": "embeddings/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": total_ms / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u32",
        "data": "synthetic (seeded Chung-Lu generator; queries located by seeded random walk)",
        "config": {
            "workload": "chung-lu n=36692 m=183831 gamma=2.1 maxdeg=1400 seed=0; "
                        "queries 4-cycle+2-pendants and K4-e+2-tail; heuristic order; one step = both queries",
            "counts": counts[: len(plans)],
            "p50_ms_per_query": statistics.median(per_query_ms),
            "p99_ms_per_query": sorted(per_query_ms)[min(len(per_query_ms) - 1, int(0.99 * len(per_query_ms)))],
            "candidates_per_step": sum(work) / args.steps,
            "candidates_per_s": sum(work) / (total_ms / 1e3),
            "wall_s": wall,
            "upload_s": upload_s,
        },
        "roofline": {
            "bound": "hbm",
            "kernel": dname,
            "achieved": achieved,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS,
            "traffic": traffic_for(dname),
            "algorithmic_bytes_per_launch": bytes_per_launch,
            "bytes_model": model,
            "avg_launch_ms": avg_launch_s * 1e3,
            "launches": dn,
            "merge_model_bytes_per_launch": dmerge / dn,
            "note": note,
        },
        "merge_bytes_total": sum(v[3] for v in kern.values()),
        "merge_model_GBs": sum(v[3] for v in kern.values()) / (total_ms / 1e3) / 1e9,
        "kernels_ms": {k: v[0] for k, v in kern.items() if v[1]},
    }
    if args.copy_probe:
        line["copy_probe_GBs"] = ctx.copy_probe(1 << 30, args.copy_probe)
    if not args.no_cpu:
        print("bench: CPU baseline (about 20 s)", file=sys.stderr, flush=True)
        line["cpu_baseline"] = cpu_baseline(g, shapes)
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
