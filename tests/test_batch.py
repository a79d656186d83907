"""Batched persistent matcher: per-query counts and level sizes equal the
single-query matcher and the oracle; hipGraph replays agree with each other
and with plain launches; the config-4 workload matches its fixture."""
import json
import os
import random

import numpy as np
import pytest

import cuts_amd
import oracle
from cuts_amd import workloads as W

pytestmark = pytest.mark.gpu
# any stall inside the persistent kernel becomes an error within this bound
WAIT = 60.0
GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def small_queries(g, n, seed):
    rng = random.Random(seed)
    out = []
    for i in range(n):
        k = rng.randint(2, 7)
        e, _ = g.walk_query(k, rng.randint(k - 1, 2 * k), seed * 1000 + i)
        out.append((k, e.tolist()))
    return out


@pytest.mark.parametrize("seed", [0, 1, 2])
def test_batch_equals_single_and_oracle(ctx, seed):
    g = cuts_amd.Graph.chung_lu(150, 500, 2.6, 0, 100 + seed)
    dg = ctx.upload(g)
    og = oracle.OGraph.from_view(g.view())
    qs = small_queries(g, 24, seed)
    plans = [cuts_amd.Plan(cuts_amd.Graph.from_edges(k, e)) for k, e in qs]
    singles = [ctx.match(dg, p) for p in plans]
    for flags in (0, cuts_amd.FLAG_NO_FUSE, cuts_amd.FLAG_NO_SHORTCUT):
        b = cuts_amd.Batch(ctx, dg, plans, flags=flags, vchunk=64, wait_limit_s=WAIT)
        r = b.run()
        for i, (k, e) in enumerate(qs):
            assert r.counts[i] == singles[i].count, (flags, i)
            assert r.level_sizes[i] == singles[i].level_sizes, (flags, i)
    for i, (k, e) in enumerate(qs[:8]):
        assert singles[i].count == oracle.count(og, oracle.qmatrix(k, oracle.undirected(e)))


def test_batch_graph_replays_agree(ctx):
    g = cuts_amd.Graph.chung_lu(300, 1200, 2.4, 0, 9)
    dg = ctx.upload(g)
    qs = small_queries(g, 40, 7)
    plans = [cuts_amd.Plan(cuts_amd.Graph.from_edges(k, e)) for k, e in qs]
    want = [ctx.match(dg, p).count for p in plans]
    b = cuts_amd.Batch(ctx, dg, plans, use_graph=True, wait_limit_s=WAIT)
    plain = cuts_amd.Batch(ctx, dg, plans, use_graph=False, wait_limit_s=WAIT)
    for rep in range(4):
        r = b.run()
        assert r.counts == want, rep
        ctx.match(dg, plans[rep])  # eager work between replays
        assert plain.run().counts == want
        assert all(x > 0 for x in r.latency_ms)


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
    r = cuts_amd.Batch(ctx, dg, plans, wait_limit_s=WAIT).run()
    assert r.counts == [ctx.match(dg, p).count for p in plans]


def test_batch_tiny_queue_fails_cleanly(ctx):
    g = cuts_amd.Graph.chung_lu(300, 1500, 2.4, 0, 3)
    dg = ctx.upload(g)
    plans = [cuts_amd.Plan(cuts_amd.Graph.from_edges(k, e)) for k, e in small_queries(g, 8, 3)]
    b = cuts_amd.Batch(ctx, dg, plans, queue_cap=48, vchunk=64, wait_limit_s=WAIT)  # 40 level-0 items fit, children do not
    with pytest.raises(cuts_amd.CutsError):
        b.run()
    # the context stays usable
    assert ctx.match(dg, plans[0]).count >= 0


@pytest.mark.skipif(not os.path.exists(os.path.join(GOLD, "batch_queries.json")), reason="fixture not generated")
def test_config4_batch_matches_fixture(ctx):
    fx = json.load(open(os.path.join(GOLD, "batch_queries.json")))
    g = W.config2_graph()
    dg = ctx.upload(g)
    plans = [cuts_amd.Plan(cuts_amd.Graph.from_edges(q["k"], q["edges"])) for q in fx["queries"]]
    assert [p.order for p in plans] == [q["order"] for q in fx["queries"]]
    r = cuts_amd.Batch(ctx, dg, plans, arena_bytes=32 << 30, wait_limit_s=WAIT).run()
    assert r.counts == [q["count"] for q in fx["queries"]]
    assert r.level_sizes == [q["level_sizes"] for q in fx["queries"]]
