"""Batched matcher: per-query counts and level sizes equal the single-query
matcher and the oracle; hipGraph replays agree with each other and with
plain launches; capacity errors are reported cleanly; the config-4 workload
matches its fixture.

The level-synchronous queue (the default) is what runs.  The tests at the
end of the file exercise the asynchronous ticket-queue kernel, which stalled
on MI355X (no progress, no fault) in its first test and again in the one
re-test after a fix.  Its block loop had the shape that made the first
version of the level-synchronous kernel stall too: a `threadIdx.x == 0`
region on each side of the loop's back edge, which the compiler merged into
a lane-0 latch of its own, so the other lanes passed the next barrier
without lane 0's new item.  It stays disabled (creation fails) and its tests
stay skipped, not deleted."""
import json
import os
import random

import numpy as np
import pytest

import cuts_amd
import oracle
from cuts_amd import workloads as W

pytestmark = pytest.mark.gpu
GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
TICKET_SKIP = pytest.mark.skip(reason="asynchronous ticket-queue batch kernel stalled on MI355X in its first run and "
                                      "in the one re-test after a fix; disabled until the cause is shown")


def small_queries(g, n, seed, kmin=2, kmax=7):
    rng = random.Random(seed)
    out = []
    for i in range(n):
        k = rng.randint(kmin, kmax)
        e, _ = g.walk_query(k, rng.randint(k - 1, 2 * k), seed * 1000 + i)
        out.append((k, e.tolist()))
    return out


@pytest.mark.parametrize("seed", [0, 1, 2])
def test_level_batch_equals_single_and_oracle(ctx, seed):
    g = cuts_amd.Graph.chung_lu(150, 500, 2.6, 0, 100 + seed)
    dg = ctx.upload(g)
    og = oracle.OGraph.from_view(g.view())
    qs = small_queries(g, 24, seed)
    plans = [cuts_amd.Plan(cuts_amd.Graph.from_edges(k, e)) for k, e in qs]
    for i, (k, e) in enumerate(qs[:8]):
        print(f"oracle seed={seed} query={i} k={k}", flush=True)
        want = oracle.count(og, oracle.qmatrix(k, oracle.undirected(e)))
        assert ctx.match(dg, plans[i]).count == want, i
    singles = [ctx.match(dg, p) for p in plans]
    for flags, kw in ((0, {}), (cuts_amd.FLAG_NO_FUSE, {}), (cuts_amd.FLAG_NO_SHORTCUT, {}),
                      (cuts_amd.FLAG_NO_BITMAP, {}), (cuts_amd.FLAG_NO_PAIR, {}),
                      (0, dict(item_work=64, blocks_per_cu=1)),
                      (0, dict(vchunk=7, item_work=1))):
        print(f"batch seed={seed} flags={flags} {kw}", flush=True)
        b = cuts_amd.Batch(ctx, dg, plans, flags=flags, vchunk=kw.pop("vchunk", 64), **kw)
        r = b.run()
        for i in range(len(qs)):
            assert r.counts[i] == singles[i].count, (flags, i)
            assert r.level_sizes[i] == singles[i].level_sizes, (flags, i)


@pytest.mark.parametrize("seed", [0, 1])
def test_level_batch_pair_level_shapes(ctx, seed):
    """Shapes whose last two positions are counted together as a pair or a
    chain (and some whose are not) in one batch: per-query counts and level sizes equal the
    single-query matcher's with and without the pair level, and the
    oracle's."""
    from test_pair import SHAPES
    g = cuts_amd.Graph.chung_lu(120, 420, 2.4, 0, 300 + seed)
    dg = ctx.upload(g)
    og = oracle.OGraph.from_view(g.view())
    qs = [(1 + max(max(e) for e in edges), edges) for edges in SHAPES.values()]
    plans = [cuts_amd.Plan(cuts_amd.Graph.from_edges(k, e)) for k, e in qs]
    singles = [ctx.match(dg, p) for p in plans]
    for (k, e), x in zip(qs, singles):
        assert x.count == oracle.count(og, oracle.qmatrix(k, oracle.undirected(e))), e
        assert ctx.match(dg, cuts_amd.Plan(cuts_amd.Graph.from_edges(k, e)), flags=cuts_amd.FLAG_NO_PAIR).level_sizes \
            == x.level_sizes, e
    for flags, kw in ((0, {}), (cuts_amd.FLAG_NO_PAIR, {}), (cuts_amd.FLAG_NO_CHAIN, {}), (cuts_amd.FLAG_NO_BITMAP, {}),
                      (cuts_amd.FLAG_NO_TWO_HOP, {}), (0, dict(item_work=32, vchunk=5))):
        print(f"batch seed={seed} flags={flags} {kw}", flush=True)
        r = cuts_amd.Batch(ctx, dg, plans, flags=flags, **kw).run()
        assert r.counts == [x.count for x in singles], flags
        assert r.level_sizes == [x.level_sizes for x in singles], flags


def test_level_batch_wave_sums_beyond_32_bits(ctx):
    """Per-wave partial counts far above 2^31 (a 4-vertex star plus tail
    on a 4,096-vertex power-law graph): a 64-bit wave broadcast that
    sign-extended its low word once lost multiples of 2^32 here."""
    g = cuts_amd.Graph.chung_lu(4096, 32768, 2.3, 0, 5)
    dg = ctx.upload(g)
    edges = [(0, 1), (1, 2), (2, 3), (2, 4)]
    plans = [cuts_amd.Plan(cuts_amd.Graph.from_edges(5, edges), o) for o in ([2, 1, 0, 3, 4], [2, 3, 4, 1, 0])]
    plans.append(cuts_amd.Plan(cuts_amd.Graph.from_edges(4, [(0, 1), (0, 2), (0, 3)])))
    singles = [ctx.match(dg, p) for p in plans]
    assert singles[0].count == singles[1].count > 1 << 36
    for flags in (0, cuts_amd.FLAG_NO_TWO_HOP, cuts_amd.FLAG_NO_PAIR,
                  cuts_amd.FLAG_NO_PAIR | cuts_amd.FLAG_NO_PREFIX_BASE):
        print(f"batch flags={flags}", flush=True)
        r = cuts_amd.Batch(ctx, dg, plans, flags=flags, arena_bytes=4 << 30).run()
        assert r.counts == [x.count for x in singles], flags
        assert r.level_sizes == [x.level_sizes for x in singles], flags


def test_level_batch_mixed_sizes_and_one_vertex_queries(ctx):
    g = cuts_amd.Graph.chung_lu(400, 2000, 2.4, 0, 21)
    dg = ctx.upload(g)
    plans = [cuts_amd.Plan(cuts_amd.Graph.from_edges(1, [])), cuts_amd.Plan(cuts_amd.Graph.from_edges(2, [(0, 1)]))]
    plans += [cuts_amd.Plan(cuts_amd.Graph.from_edges(k, e)) for k, e in small_queries(g, 30, 5, 3, 9)]
    single = [ctx.match(dg, p, max_work=1 << 26) for p in plans]
    keep = [p for p, r in zip(plans, single) if not r.work_capped and W.materialised_rows(r.level_sizes) < 1 << 20]
    want = [r for r, p in zip(single, plans) if p in keep]
    assert len(keep) >= 10
    print(f"batch of {len(keep)} queries", flush=True)
    r = cuts_amd.Batch(ctx, dg, keep, arena_bytes=1 << 30).run()
    assert r.counts == [x.count for x in want]
    assert r.level_sizes == [x.level_sizes for x in want]
    assert r.counts[0] == g.V  # a single vertex matches every data vertex
    assert r.counts[1] == int(g.view()["out_off"][-1])  # an edge matches every arc


def test_level_batch_root_samples_equal_single_and_oracle(ctx):
    g = cuts_amd.Graph.chung_lu(400, 2000, 2.4, 0, 31)
    dg = ctx.upload(g)
    og = oracle.OGraph.from_view(g.view())
    qs = small_queries(g, 20, 9, 3, 7)
    plans = [cuts_amd.Plan(cuts_amd.Graph.from_edges(k, e)) for k, e in qs]
    roots = [W.root_sample(g.V, 1 << (i % 5)) for i in range(len(qs))]
    singles = [ctx.match(dg, p, roots=rt) for p, rt in zip(plans, roots)]
    r = cuts_amd.Batch(ctx, dg, plans, roots=roots, vchunk=16).run()
    assert r.counts == [x.count for x in singles]
    assert r.level_sizes == [x.level_sizes for x in singles]
    for (k, e), p, rt, x in list(zip(qs, plans, roots, singles))[:8]:
        print(f"oracle k={k} roots={rt}", flush=True)
        assert x.count == oracle.count_ex(og, oracle.qmatrix(k, oracle.undirected(e)), root_q=p.order[0], roots=rt)


def test_level_batch_graph_replays_agree(ctx):
    g = cuts_amd.Graph.chung_lu(300, 1200, 2.4, 0, 9)
    dg = ctx.upload(g)
    qs = small_queries(g, 40, 7)
    plans = [cuts_amd.Plan(cuts_amd.Graph.from_edges(k, e)) for k, e in qs]
    want = [ctx.match(dg, p).count for p in plans]
    b = cuts_amd.Batch(ctx, dg, plans, use_graph=True)
    plain = cuts_amd.Batch(ctx, dg, plans, use_graph=False)
    for rep in range(4):
        print(f"replay {rep}", flush=True)
        r = b.run()
        assert r.counts == want, rep
        ctx.match(dg, plans[rep])  # eager work between replays
        assert plain.run().counts == want
        assert all(x > 0 for x in r.latency_ms)


def test_level_batch_directed(ctx):
    rng = random.Random(4)
    n = 80
    arcs = sorted({(rng.randrange(n), rng.randrange(n)) for _ in range(500)})
    a = np.asarray(arcs, dtype=np.uint32)
    g = cuts_amd.Graph.from_arcs(n, a[:, 0], a[:, 1])
    dg = ctx.upload(g)
    og = oracle.OGraph.from_view(g.view())
    plans, qm = [], []
    for i in range(12):
        k = rng.randint(2, 5)
        q = set()
        for v in range(1, k):
            u = rng.randrange(v)
            q.add((u, v) if rng.random() < 0.5 else (v, u))
        qa = np.asarray(sorted(q), dtype=np.uint32)
        plans.append(cuts_amd.Plan(cuts_amd.Graph.from_arcs(k, qa[:, 0], qa[:, 1])))
        qm.append(oracle.qmatrix(k, sorted(q)))
    r = cuts_amd.Batch(ctx, dg, plans, item_work=32).run()
    assert r.counts == [oracle.count(og, m) for m in qm]
    assert r.counts == [ctx.match(dg, p).count for p in plans]


def test_level_batch_capacity_errors_are_clean(ctx):
    g = cuts_amd.Graph.chung_lu(300, 1500, 2.4, 0, 3)
    dg = ctx.upload(g)
    plans = [cuts_amd.Plan(cuts_amd.Graph.from_edges(k, e)) for k, e in small_queries(g, 8, 3, 4, 6)]
    # 40 level-0 items fit in 48 slots; the level-1 items of 64-work chunks do not
    with pytest.raises(cuts_amd.CutsError):
        cuts_amd.Batch(ctx, dg, plans, queue_cap=48, vchunk=64, item_work=64).run()
    with pytest.raises(cuts_amd.CutsError):
        cuts_amd.Batch(ctx, dg, plans, arena_bytes=4096).run()
    with pytest.raises(cuts_amd.CutsError):
        cuts_amd.Batch(ctx, dg, plans, queue_cap=8, vchunk=64)  # level-0 items alone do not fit
    # the context stays usable and a batch with room still agrees
    want = [ctx.match(dg, p).count for p in plans]
    assert cuts_amd.Batch(ctx, dg, plans).run().counts == want


def test_ticket_queue_is_disabled(ctx):
    g = cuts_amd.Graph.chung_lu(100, 300, 2.6, 0, 1)
    dg = ctx.upload(g)
    p = cuts_amd.Plan(cuts_amd.Graph.from_edges(3, [(0, 1), (1, 2)]))
    with pytest.raises(cuts_amd.CutsError, match="disabled"):
        cuts_amd.Batch(ctx, dg, [p], queue="ticket")


@pytest.mark.skipif(not os.path.exists(os.path.join(GOLD, "batch_queries.json")), reason="fixture not generated")
def test_config4_batch_matches_fixture(ctx):
    fx = json.load(open(os.path.join(GOLD, "batch_queries.json")))
    g = W.config2_graph()
    dg = ctx.upload(g)
    plans = [cuts_amd.Plan(cuts_amd.Graph.from_edges(q["k"], q["edges"])) for q in fx["queries"]]
    assert [p.order for p in plans] == [q["order"] for q in fx["queries"]]
    r = cuts_amd.Batch(ctx, dg, plans, arena_bytes=48 << 30, queue_cap=1 << 26,
                       roots=[tuple(q["roots"]) for q in fx["queries"]]).run()
    assert r.counts == [q["count"] for q in fx["queries"]]
    assert r.level_sizes == [q["level_sizes"] for q in fx["queries"]]


# ---- asynchronous ticket-queue kernel (disabled) -------------------------

@TICKET_SKIP
@pytest.mark.parametrize("seed", [0, 1, 2])
def test_batch_equals_single_and_oracle(ctx, seed):
    g = cuts_amd.Graph.chung_lu(150, 500, 2.6, 0, 100 + seed)
    dg = ctx.upload(g)
    og = oracle.OGraph.from_view(g.view())
    qs = small_queries(g, 24, seed)
    plans = [cuts_amd.Plan(cuts_amd.Graph.from_edges(k, e)) for k, e in qs]
    singles = [ctx.match(dg, p) for p in plans]
    for flags in (0, cuts_amd.FLAG_NO_FUSE, cuts_amd.FLAG_NO_SHORTCUT):
        b = cuts_amd.Batch(ctx, dg, plans, flags=flags, vchunk=64, queue="ticket")
        r = b.run()
        for i, (k, e) in enumerate(qs):
            assert r.counts[i] == singles[i].count, (flags, i)
            assert r.level_sizes[i] == singles[i].level_sizes, (flags, i)
    for i, (k, e) in enumerate(qs[:8]):
        assert singles[i].count == oracle.count(og, oracle.qmatrix(k, oracle.undirected(e)))


@TICKET_SKIP
def test_batch_graph_replays_agree(ctx):
    g = cuts_amd.Graph.chung_lu(300, 1200, 2.4, 0, 9)
    dg = ctx.upload(g)
    qs = small_queries(g, 40, 7)
    plans = [cuts_amd.Plan(cuts_amd.Graph.from_edges(k, e)) for k, e in qs]
    want = [ctx.match(dg, p).count for p in plans]
    b = cuts_amd.Batch(ctx, dg, plans, use_graph=True, queue="ticket")
    plain = cuts_amd.Batch(ctx, dg, plans, use_graph=False, queue="ticket")
    for rep in range(4):
        r = b.run()
        assert r.counts == want, rep
        ctx.match(dg, plans[rep])  # eager work between replays
        assert plain.run().counts == want
        assert all(x > 0 for x in r.latency_ms)


@TICKET_SKIP
def test_batch_directed(ctx):
    rng = random.Random(4)
    n = 80
    arcs = sorted({(rng.randrange(n), rng.randrange(n)) for _ in range(500)})
    a = np.asarray(arcs, dtype=np.uint32)
    g = cuts_amd.Graph.from_arcs(n, a[:, 0], a[:, 1])
    dg = ctx.upload(g)
    plans = []
    for i in range(12):
        k = rng.randint(2, 5)
        q = set()
        for v in range(1, k):
            u = rng.randrange(v)
            q.add((u, v) if rng.random() < 0.5 else (v, u))
        qa = np.asarray(sorted(q), dtype=np.uint32)
        plans.append(cuts_amd.Plan(cuts_amd.Graph.from_arcs(k, qa[:, 0], qa[:, 1])))
    r = cuts_amd.Batch(ctx, dg, plans, queue="ticket").run()
    assert r.counts == [ctx.match(dg, p).count for p in plans]


@TICKET_SKIP
def test_batch_tiny_queue_fails_cleanly(ctx):
    g = cuts_amd.Graph.chung_lu(300, 1500, 2.4, 0, 3)
    dg = ctx.upload(g)
    plans = [cuts_amd.Plan(cuts_amd.Graph.from_edges(k, e)) for k, e in small_queries(g, 8, 3)]
    b = cuts_amd.Batch(ctx, dg, plans, queue_cap=48, vchunk=64, queue="ticket")  # 40 level-0 items fit, children do not
    with pytest.raises(cuts_amd.CutsError):
        b.run()
    # the context stays usable
    assert ctx.match(dg, plans[0]).count >= 0
