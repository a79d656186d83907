"""Hub rows: on symmetric graphs too large for the two-hop index, the upload
gives the highest-degree vertices a dense row of |N(u) & N(v)| for every v
(k_hub_rows), and the pair and chain levels take an intersection from a row
whenever u or w has one, walking lists only for the others.

The rows must equal the rows of the sparse product A*A (diagonal removed)
for the vertices the selection rule names; counts with rows for none, some
and every vertex must equal the oracle's (the index is switched off with
CUTS_H2_MAX_WEDGES=0 so that small graphs take this path); on the R-MAT
scale-20 graph, where the rows are built by default, the fixture's counts
must hold with and without them."""
import json
import os

import numpy as np
import pytest
import scipy.sparse as sp

import cuts_amd
import oracle
from cuts_amd import workloads as W
from test_pair import SHAPES, graph_for

pytestmark = pytest.mark.gpu


def force(monkeypatch, rows, min_deg):
    monkeypatch.setenv("CUTS_H2_MAX_WEDGES", "0")
    monkeypatch.setenv("CUTS_H2_HUB_ROWS", str(rows))
    monkeypatch.setenv("CUTS_H2_HUB_MIN_DEG", str(min_deg))


def expected_hubs(g, rows, min_deg):
    d = np.diff(g.view()["out_off"].astype(np.int64))
    cand = sorted((u for u in range(g.V) if d[u] >= min_deg), key=lambda u: (-d[u], u))
    return cand[:rows]


@pytest.mark.parametrize("seed", [0, 1])
def test_hub_rows_equal_sparse_square(ctx, monkeypatch, seed):
    force(monkeypatch, 60, 16)
    g = cuts_amd.Graph.chung_lu(3000 + seed, 15000, 2.2, 0, 77 + seed)
    dg = ctx.upload(g)
    info = dg.info()
    hubs = expected_hubs(g, 60, 16)
    assert info["two_hop_entries"] == 0 and info["hub_rows"] == len(hubs) > 20
    rid, rows = dg.hub_rows()
    assert [int(u) for u in np.argsort(np.where(rid >= 0, rid, 1 << 30))[:len(hubs)]] == hubs
    assert (rid >= 0).sum() == len(hubs)
    vw = g.view()
    o, adj = vw["out_off"].astype(np.int64), vw["out_adj"].astype(np.int64)
    A = sp.csr_matrix((np.ones(len(adj), dtype=np.int64), (np.repeat(np.arange(g.V), np.diff(o)), adj)),
                      shape=(g.V, g.V))
    want = (A[hubs] @ A).toarray()
    want[np.arange(len(hubs)), hubs] = 0
    assert np.array_equal(rows.astype(np.int64), want)
    print(f"seed {seed}: {len(hubs)} hub rows built in {info['hub_rows_build_s'] * 1e3:.1f} ms", flush=True)


@pytest.mark.parametrize("rows", [0, 12, 100000])
def test_pair_and_chain_counts_with_hub_rows(ctx, monkeypatch, rows):
    """No rows (list walks only), rows for the dozen longest lists (a mix of
    lookups and walks in one pass), and rows for every vertex."""
    force(monkeypatch, rows, 1)
    names = ["star4", "triangle+2pend-two", "4cycle+2pend", "k4+2pend", "5cycle+3pend", "triangle+tail2",
             "k4e+tail2", "path5", "4cycle+tail2", "star3+tail2", "k4+tail2", "5cycle+chord+2tails"]
    for seed in (9401, 9402):
        g = graph_for(seed)
        og = oracle.OGraph.from_view(g.view())
        dg = ctx.upload(g)
        assert dg.info()["hub_rows"] == min(rows, g.V - int((np.diff(g.view()["out_off"]) == 0).sum()))
        for name in names:
            edges = SHAPES[name]
            k = 1 + max(max(e) for e in edges)
            want = oracle.count(og, oracle.qmatrix(k, oracle.undirected(edges)))
            p = cuts_amd.Plan(cuts_amd.Graph.from_edges(k, edges))
            for flags in (0, cuts_amd.FLAG_NO_SHARED, cuts_amd.FLAG_NO_BITMAP, cuts_amd.FLAG_NO_ARC_COMMON,
                          cuts_amd.FLAG_SYNC_LEVELS, cuts_amd.FLAG_NO_PAIR | cuts_amd.FLAG_NO_CHAIN):
                assert ctx.match(dg, p, flags=flags).count == want, (rows, seed, name, flags)
            assert ctx.match(dg, p, budget=2048).count == want, (rows, seed, name)


def test_rmat20_fixture_counts_with_and_without_hub_rows(ctx):
    fx = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "rmat20_queries.json")))
    g = W.config3_graph()
    dg = ctx.upload(g)
    info = dg.info()
    assert info["two_hop_entries"] == 0 and info["hub_rows"] > 0
    print(f"R-MAT 20: {info['hub_rows']} hub rows built in {info['hub_rows_build_s']:.2f} s", flush=True)
    for q in [fx["config3"]] + fx["queries"][:40]:
        p = cuts_amd.Plan(cuts_amd.Graph.from_edges(q["k"], [tuple(e) for e in q["edges"]]))
        rt = tuple(q["roots"])
        assert ctx.match(dg, p, roots=rt).count == q["count"], q["edges"]
        assert ctx.match(dg, p, roots=rt, flags=cuts_amd.FLAG_NO_TWO_HOP).count == q["count"], q["edges"]
