"""A query matched again on the same device graph, plan, roots and flags
takes the level sizes of its last complete run as a hint: a writing level
that had at most one planning block's rows hands its size to the next level
on the device instead of through the host.  Every repeat must give the
counts, level sizes and work of the run that reads every level back
(CUTS_FLAG_SYNC_LEVELS), and so must interleaved queries whose hints belong
to other root samples."""
import random

import pytest

import cuts_amd
from cuts_amd import workloads as W

pytestmark = pytest.mark.gpu

SYNC = cuts_amd.FLAG_SYNC_LEVELS


def same(a, b):
    return (a.count, a.level_sizes, a.work) == (b.count, b.level_sizes, b.work)


def test_repeated_queries_agree_with_synchronous_levels(ctx):
    g = cuts_amd.Graph.chung_lu(4096, 32768, 2.3, 0, 5)
    dg = ctx.upload(g)
    rng = random.Random(77)
    qs = [(s.k, list(s.edges)) for s in (W.TRI_TAIL2, W.C4_2PEND, W.K4E_TAIL2, W.C5CHORD_2TAIL)]
    for i in range(12):
        k = rng.randint(5, 7)
        e, _ = g.walk_query(k, rng.randint(k - 1, k + 2), 300 + i)
        qs.append((k, e.tolist()))
    plans = [cuts_amd.Plan(cuts_amd.Graph.from_edges(k, e)) for k, e in qs]
    # root samples of 1/32 .. 1/128 of the vertices (all levels but the
    # closed-form ones stay small; every level expanded would be 10^10
    # candidates and more on this graph)
    roots = [W.root_sample(g.V, 1 << rng.randint(5, 7)) for _ in plans]
    for flags in (0, cuts_amd.FLAG_NO_TWO_HOP, cuts_amd.FLAG_NO_PAIR | cuts_amd.FLAG_NO_CHAIN):
        want = []
        for i, (p, rt) in enumerate(zip(plans, roots)):
            want.append(ctx.match(dg, p, roots=rt, flags=flags | SYNC))
            print(f"flags={flags} query {i} k={p.k} roots={rt}: {want[-1].count} in {want[-1].device_ms:.1f} ms",
                  flush=True)
        for rep in range(3):
            print(f"flags={flags} repeat {rep}", flush=True)
            for i in rng.sample(range(len(plans)), len(plans)):
                got = ctx.match(dg, plans[i], roots=roots[i], flags=flags)
                assert same(got, want[i]), (flags, rep, qs[i], got, want[i])
        # the same plans on other root samples (other hints), then back
        other = [W.root_sample(g.V, 1 << rng.randint(4, 6)) for _ in plans]
        for p, rt in zip(plans, other):
            assert same(ctx.match(dg, p, roots=rt, flags=flags), ctx.match(dg, p, roots=rt, flags=flags | SYNC))
        for i in range(len(plans)):
            assert same(ctx.match(dg, plans[i], roots=roots[i], flags=flags), want[i])


def test_repeats_on_the_headline_graph(ctx):
    """The first 60 held-out queries and the 5-cycle query three times over:
    the fixture's counts every time, and the first run's level sizes and work
    on every repeat."""
    import json
    import os
    fx = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "rmat20_queries.json")))
    g = W.config3_graph()
    dg = ctx.upload(g)
    items = [fx["config3"]] + fx["queries"][:60]
    plans = [cuts_amd.Plan(cuts_amd.Graph.from_edges(q["k"], [tuple(e) for e in q["edges"]])) for q in items]
    first = None
    for rep in range(3):
        print(f"R-MAT 20 repeat {rep}", flush=True)
        got = [ctx.match(dg, p, roots=tuple(q["roots"])) for p, q in zip(plans, items)]
        assert [r.count for r in got] == [q["count"] for q in items], rep
        if first is None:
            first = got
        assert all(same(a, b) for a, b in zip(got, first)), rep
