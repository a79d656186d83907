"""The persistent batch kernel (queue="persistent"): one cooperative launch
runs every trie level of every query, pulling (query, level, chunk) items
from the device queue, with a device-wide barrier between levels in place of
the level-synchronous queue's launch boundaries.  Per-query counts and level
sizes must equal the single-query matcher's (and so the oracle's, which the
single-query suite checks) in every mode; hipGraph replays must agree with
plain launches; capacity errors and the barrier's own bounded wait must fail
the run cleanly and leave the context usable."""
import json
import os

import pytest

import cuts_amd
from cuts_amd import workloads as W
from test_batch import GOLD, small_queries

pytestmark = pytest.mark.gpu

NO_WAIT = 0xFFFFFFFF  # barrier_timeout_ms: the first barrier aborts the run
WAIT_MS = 10000       # these batches take milliseconds: a longer wait at a barrier fails the run


def persistent(ctx, dg, plans, **kw):
    kw.setdefault("barrier_timeout_ms", WAIT_MS)
    return cuts_amd.Batch(ctx, dg, plans, queue="persistent", **kw)


@pytest.mark.parametrize("seed", [0, 1, 2])
def test_persistent_batch_equals_single(ctx, seed):
    g = cuts_amd.Graph.chung_lu(150, 500, 2.6, 0, 100 + seed)
    dg = ctx.upload(g)
    qs = small_queries(g, 24, seed)
    plans = [cuts_amd.Plan(cuts_amd.Graph.from_edges(k, e)) for k, e in qs]
    singles = [ctx.match(dg, p) for p in plans]
    for flags, kw in ((0, {}), (cuts_amd.FLAG_NO_FUSE, {}), (cuts_amd.FLAG_NO_SHORTCUT, {}),
                      (cuts_amd.FLAG_NO_BITMAP, {}), (cuts_amd.FLAG_NO_PAIR, {}),
                      (0, dict(item_work=64, blocks_per_cu=1)),
                      (0, dict(vchunk=7, item_work=1))):
        print(f"persistent batch seed={seed} flags={flags} {kw}", flush=True)
        r = persistent(ctx, dg, plans, flags=flags, vchunk=kw.pop("vchunk", 64), **kw).run()
        assert r.counts == [x.count for x in singles], flags
        assert r.level_sizes == [x.level_sizes for x in singles], flags


@pytest.mark.parametrize("seed", [0, 1])
def test_persistent_batch_pair_and_chain_shapes(ctx, seed):
    from test_pair import SHAPES
    g = cuts_amd.Graph.chung_lu(120, 420, 2.4, 0, 300 + seed)
    dg = ctx.upload(g)
    qs = [(1 + max(max(e) for e in edges), edges) for edges in SHAPES.values()]
    plans = [cuts_amd.Plan(cuts_amd.Graph.from_edges(k, e)) for k, e in qs]
    singles = [ctx.match(dg, p) for p in plans]
    for flags in (0, cuts_amd.FLAG_NO_PAIR, cuts_amd.FLAG_NO_CHAIN, cuts_amd.FLAG_NO_TWO_HOP):
        print(f"persistent batch seed={seed} flags={flags}", flush=True)
        r = persistent(ctx, dg, plans, flags=flags).run()
        assert r.counts == [x.count for x in singles], flags
        assert r.level_sizes == [x.level_sizes for x in singles], flags


def test_persistent_batch_mixed_sizes_roots_and_wide_rows(ctx):
    """Queries of 1..12 vertices (row widths of all three kernel instances
    across the test), on root samples, against the level-synchronous queue."""
    g = cuts_amd.Graph.chung_lu(400, 2000, 2.4, 0, 31)
    dg = ctx.upload(g)
    qs = small_queries(g, 30, 9, 3, 12)
    plans = [cuts_amd.Plan(cuts_amd.Graph.from_edges(1, []))]
    plans += [cuts_amd.Plan(cuts_amd.Graph.from_edges(k, e)) for k, e in qs]
    roots = [(0, g.V, 1)] + [W.root_sample(g.V, 1 << (3 + i % 4)) for i in range(len(qs))]
    single = [ctx.match(dg, p, roots=rt, max_work=1 << 24) for p, rt in zip(plans, roots)]
    keep = [i for i, r in enumerate(single) if not r.work_capped and W.materialised_rows(r.level_sizes) < 1 << 20]
    assert len(keep) >= 12
    for sub in (keep, [i for i in keep if plans[i].k <= 8]):
        ps, rts = [plans[i] for i in sub], [roots[i] for i in sub]
        print(f"persistent batch of {len(ps)} queries, k <= {max(p.k for p in ps)}", flush=True)
        r = persistent(ctx, dg, ps, roots=rts, vchunk=16, arena_bytes=1 << 30).run()
        lv = cuts_amd.Batch(ctx, dg, ps, roots=rts, vchunk=16, arena_bytes=1 << 30).run()
        assert r.counts == [single[i].count for i in sub] == lv.counts
        assert r.level_sizes == [single[i].level_sizes for i in sub] == lv.level_sizes


def test_persistent_batch_graph_replays_agree(ctx):
    g = cuts_amd.Graph.chung_lu(300, 1200, 2.4, 0, 9)
    dg = ctx.upload(g)
    plans = [cuts_amd.Plan(cuts_amd.Graph.from_edges(k, e)) for k, e in small_queries(g, 40, 7)]
    want = [ctx.match(dg, p).count for p in plans]
    b = persistent(ctx, dg, plans, use_graph=True)
    plain = persistent(ctx, dg, plans, use_graph=False)
    for rep in range(4):
        print(f"persistent replay {rep}", flush=True)
        r = b.run()
        assert r.counts == want, rep
        ctx.match(dg, plans[rep])  # eager work between replays
        assert plain.run().counts == want
        assert all(x > 0 for x in r.latency_ms)


def test_persistent_batch_errors_are_clean(ctx):
    g = cuts_amd.Graph.chung_lu(300, 1500, 2.4, 0, 3)
    dg = ctx.upload(g)
    plans = [cuts_amd.Plan(cuts_amd.Graph.from_edges(k, e)) for k, e in small_queries(g, 8, 3, 4, 6)]
    want = [ctx.match(dg, p).count for p in plans]
    print("queue and arena too small", flush=True)
    with pytest.raises(cuts_amd.CutsError, match="queue full"):
        persistent(ctx, dg, plans, queue_cap=48, vchunk=64, item_work=64).run()
    with pytest.raises(cuts_amd.CutsError, match="arena full"):
        persistent(ctx, dg, plans, arena_bytes=4096).run()
    # the barrier's bounded wait, forced from the host: a zero-length wait
    # aborts at the first barrier, and every block leaves
    print("zero-length barrier wait", flush=True)
    for use_graph in (False, True):
        with pytest.raises(cuts_amd.CutsError, match="barrier"):
            persistent(ctx, dg, plans, barrier_timeout_ms=NO_WAIT, use_graph=use_graph).run()
    # the context stays usable and a batch with room still agrees
    assert persistent(ctx, dg, plans).run().counts == want
    assert cuts_amd.Batch(ctx, dg, plans).run().counts == want


@pytest.mark.skipif(not os.path.exists(os.path.join(GOLD, "batch_queries.json")), reason="fixture not generated")
def test_persistent_config4_batch_matches_fixture(ctx):
    fx = json.load(open(os.path.join(GOLD, "batch_queries.json")))
    g = W.config2_graph()
    dg = ctx.upload(g)
    plans = [cuts_amd.Plan(cuts_amd.Graph.from_edges(q["k"], q["edges"])) for q in fx["queries"]]
    print(f"persistent batch of the {len(plans)} config-4 queries", flush=True)
    r = persistent(ctx, dg, plans, arena_bytes=48 << 30, queue_cap=1 << 26,
                   roots=[tuple(q["roots"]) for q in fx["queries"]]).run()
    assert r.counts == [q["count"] for q in fx["queries"]]
    assert r.level_sizes == [q["level_sizes"] for q in fx["queries"]]
