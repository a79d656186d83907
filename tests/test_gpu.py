"""GPU parity: the HIP matcher (through the C ABI) against the oracle.

Exact, bit-for-bit equality of uint64 counts and of embedding sets:
networkx-pinned cases, random graphs vs the backtracking oracle under every
execution mode (closed-form last level on/off, fused last two levels on/off,
last two positions counted together on/off, XCD tile mapping on/off, manual
trunks, tiny HBM budgets that force many chunks), per-level table sizes vs the CPU trie baseline, the full config-1
count, root-sampled config-2/3 counts, and the reference-compatible CLI.
"""
import json
import os
import random
import re
import subprocess

import numpy as np
import pytest

import cuts_amd
import oracle
from cuts_amd import workloads as W

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLD = os.path.join(ROOT, "tests", "golden")


def load(name):
    with open(os.path.join(GOLD, name)) as f:
        return json.load(f)


MODES = [
    dict(),
    dict(flags=cuts_amd.FLAG_NO_FUSE),
    dict(flags=cuts_amd.FLAG_NO_SHORTCUT),
    dict(flags=cuts_amd.FLAG_NO_XCD),
    dict(flags=cuts_amd.FLAG_NO_BITMAP),
    dict(flags=cuts_amd.FLAG_NO_PREFIX_BASE | cuts_amd.FLAG_NO_BITMAP),
    dict(trunks=3),
    dict(flags=cuts_amd.FLAG_NO_PAIR),
    dict(flags=cuts_amd.FLAG_NO_TWO_HOP),
    dict(budget=4096),  # a few hundred rows per level: many chunks
    dict(expand_threads=64),  # counting expansions with 64-thread blocks, 16 items per lane
    dict(expand_threads=128, flags=cuts_amd.FLAG_NO_PAIR),
    dict(flags=cuts_amd.FLAG_SYNC_LEVELS),  # every level's work read back before it is expanded
]


def arcs_graph(V, arcs):
    a = np.asarray(arcs, dtype=np.uint32).reshape(-1, 2)
    return cuts_amd.Graph.from_arcs(V, a[:, 0], a[:, 1])


NX = load("nx_counts.json")["cases"]


@pytest.mark.parametrize("case", NX, ids=[c["name"] for c in NX])
def test_networkx_cases(ctx, case):
    g = arcs_graph(case["V"], case["arcs"])
    dg = ctx.upload(g)
    p = cuts_amd.Plan(arcs_graph(case["k"], case["qarcs"]))
    for m in MODES:
        assert ctx.match(dg, p, **m).count == case["count"], m


def random_case(seed):
    """Small power-law graph and a random-walk query; sized so the oracle
    finishes in well under a second."""
    rng = random.Random(seed)
    n = rng.randint(40, 160)
    g = cuts_amd.Graph.chung_lu(n, rng.randint(2 * n, 4 * n), rng.uniform(2.3, 2.9), 0, seed)
    k = rng.randint(2, 7)
    e, _ = g.walk_query(k, rng.randint(k - 1, 2 * k), seed + 1)
    return g, k, e.tolist()


def modes_for(count):
    """Every execution mode; the tiny-budget mode (1024-row chunks, one
    host round trip each) only where the tables stay small."""
    return [m for m in MODES if "budget" not in m or count <= 20_000_000]


@pytest.mark.parametrize("seed", range(24))
def test_random_undirected_vs_oracle(ctx, seed):
    g, k, edges = random_case(seed)
    og = oracle.OGraph.from_view(g.view())
    qa = oracle.qmatrix(k, oracle.undirected(edges))
    want = oracle.count(og, qa)
    dg = ctx.upload(g)
    p = cuts_amd.Plan(cuts_amd.Graph.from_edges(k, edges))
    for m in modes_for(want):
        r = ctx.match(dg, p, **m)
        assert r.count == want, (m, r)


@pytest.mark.parametrize("seed", range(12))
def test_random_directed_vs_oracle(ctx, seed):
    rng = random.Random(1000 + seed)
    n = rng.randint(30, 120)
    arcs = sorted({(rng.randrange(n), rng.randrange(n)) for _ in range(rng.randint(3 * n, 8 * n))})
    k = rng.randint(2, 6)
    q = set()
    for v in range(1, k):
        u = rng.randrange(v)
        q.add((u, v) if rng.random() < 0.5 else (v, u))
        if rng.random() < 0.3:
            q.add((v, u))
    q = sorted(q)
    og = oracle.OGraph.from_arcs(n, arcs)
    want = oracle.count(og, oracle.qmatrix(k, q))
    dg = ctx.upload(arcs_graph(n, arcs))
    p = cuts_amd.Plan(arcs_graph(k, q))
    for m in modes_for(want):
        assert ctx.match(dg, p, **m).count == want, m


@pytest.mark.parametrize("seed", range(8))
def test_embedding_sets_vs_oracle(ctx, seed):
    g, k, edges = random_case(200 + seed)
    og = oracle.OGraph.from_view(g.view())
    qa = oracle.qmatrix(k, oracle.undirected(edges))
    if oracle.count(og, qa) > 200000:
        pytest.skip("too many embeddings for a set comparison")
    want = oracle.enumerate_embeddings(og, qa, cap=1 << 21)
    p = cuts_amd.Plan(cuts_amd.Graph.from_edges(k, edges))
    dg = ctx.upload(g)
    for budget in (0, 2048):
        r = ctx.match(dg, p, write=True, emb_cap=len(want) + 16, budget=budget)
        assert r.count == len(want)
        got = {tuple(row) for row in r.rows.tolist()}
        # oracle rows are indexed by query vertex; GPU rows by order position
        exp = {tuple(row[u] for u in p.order) for row in want.tolist()}
        assert got == exp and len(got) == len(want)


@pytest.mark.parametrize("seed", range(10))
def test_level_sizes_vs_cpu_trie(ctx, seed):
    g, k, edges = random_case(400 + seed)
    og = oracle.OGraph.from_view(g.view())
    qa = oracle.qmatrix(k, oracle.undirected(edges))
    p = cuts_amd.Plan(cuts_amd.Graph.from_edges(k, edges))
    cpu = oracle.trie_cpu(og, qa, p.order, nthreads=4)
    dg = ctx.upload(g)
    for m in modes_for(cpu.count)[:3] + [dict(flags=cuts_amd.FLAG_NO_FUSE | cuts_amd.FLAG_NO_SHORTCUT)]:
        r = ctx.match(dg, p, **m)
        assert r.count == cpu.count
        assert r.level_sizes == [cpu.level[i] for i in range(k)], m


def _profiled(m):
    mm = dict(m)
    mm["flags"] = mm.get("flags", 0) | cuts_amd.FLAG_PROFILE
    return mm


@pytest.mark.parametrize("seed", range(10))
def test_merge_bytes_equal_cpu_trie_undirected(ctx, seed):
    """The profiled merge-model bytes depend on (graph, query, order) only:
    every execution mode reports exactly the CPU trie walk's figure."""
    g, k, edges = random_case(500 + seed)
    og = oracle.OGraph.from_view(g.view())
    p = cuts_amd.Plan(cuts_amd.Graph.from_edges(k, edges))
    cpu = oracle.trie_cpu(og, oracle.qmatrix(k, oracle.undirected(edges)), p.order, nthreads=4)
    dg = ctx.upload(g)
    for m in modes_for(cpu.count):
        r = ctx.match(dg, p, **_profiled(m))
        assert r.count == cpu.count, m
        assert sum(r.merge_bytes.values()) == cpu.bytes, (m, r.merge_bytes)


@pytest.mark.parametrize("seed", range(6))
def test_merge_bytes_equal_cpu_trie_directed(ctx, seed):
    rng = random.Random(2000 + seed)
    n = rng.randint(30, 120)
    arcs = sorted({(rng.randrange(n), rng.randrange(n)) for _ in range(rng.randint(3 * n, 8 * n))})
    k = rng.randint(2, 6)
    q = set()
    for v in range(1, k):
        u = rng.randrange(v)
        q.add((u, v) if rng.random() < 0.5 else (v, u))
    q = sorted(q)
    g = arcs_graph(n, arcs)
    og = oracle.OGraph.from_view(g.view())
    p = cuts_amd.Plan(arcs_graph(k, q))
    cpu = oracle.trie_cpu(og, oracle.qmatrix(k, q), p.order, nthreads=4)
    dg = ctx.upload(g)
    for m in (dict(), dict(trunks=3), dict(flags=cuts_amd.FLAG_NO_XCD)):
        r = ctx.match(dg, p, **_profiled(m))
        assert r.count == cpu.count, m
        assert sum(r.merge_bytes.values()) == cpu.bytes, (m, r.merge_bytes)


def test_root_restriction_partitions(ctx):
    g, k, edges = random_case(77)
    dg = ctx.upload(g)
    p = cuts_amd.Plan(cuts_amd.Graph.from_edges(k, edges))
    total = ctx.match(dg, p).count
    V = g.V
    assert sum(ctx.match(dg, p, roots=(r, V, 4)).count for r in range(4)) == total


@pytest.mark.parametrize("seed", [3, 9, 17])
def test_time_limit_counts_exact_root_prefix(ctx, seed):
    g, k, edges = random_case(seed)
    dg = ctx.upload(g)
    p = cuts_amd.Plan(cuts_amd.Graph.from_edges(k, edges))
    full = ctx.match(dg, p)
    n0 = full.level_sizes[0]
    part = ctx.match(dg, p, trunks=n0, time_limit_ms=0.05)
    assert part.roots_done <= n0
    if part.roots_done == 0:
        assert part.count == 0 and not part.complete
        return
    again = ctx.match(dg, p, roots=(0, part.last_root + 1, 1))
    assert again.level_sizes[0] == part.roots_done
    assert part.count == again.count
    assert part.complete == (part.roots_done == n0)


CFG = load("config_counts.json")
# every closed form above the last level off: pair, chain, the two-hop index,
# shared chain candidates and the per-arc table
INDEPENDENT = (cuts_amd.FLAG_NO_PAIR | cuts_amd.FLAG_NO_CHAIN | cuts_amd.FLAG_NO_TWO_HOP | cuts_amd.FLAG_NO_SHARED |
               cuts_amd.FLAG_NO_ARC_COMMON)


def test_config1_full_count(ctx):
    c = CFG["config1"]
    g = W.config1_graph()
    assert W.csr_fingerprint(g) == c["fingerprint"]
    dg = ctx.upload(g)
    p = cuts_amd.Plan(W.query_graph(W.TRI_TAIL2))
    for m in (dict(), dict(flags=cuts_amd.FLAG_NO_FUSE), dict(trunks=4), dict(budget=1 << 24)):
        assert ctx.match(dg, p, **m).count == c["count"], m


@pytest.mark.parametrize("key", ["config2a", "config2b", "config2a_hubs", "config2b_hubs", "config3"])
def test_config_sampled_counts(ctx, key):
    """Oracle counts on root samples of the benchmark graphs (low-degree
    roots, and hub roots of config 2), in the default mode and in the mode
    without pair, chain and index levels."""
    c = CFG[key]
    cfg = key.split("_")[0]
    g = {"config2a": W.config2_graph, "config2b": W.config2_graph, "config3": W.config3_graph}[cfg]()
    assert W.csr_fingerprint(g) == c["fingerprint"]
    shape = {"config2a": W.C4_2PEND, "config2b": W.K4E_TAIL2, "config3": W.C5CHORD_2TAIL}[cfg]
    p = cuts_amd.Plan(W.query_graph(shape))
    assert p.order[0] == c["root_q"]
    dg = ctx.upload(g)
    r = ctx.match(dg, p, roots=tuple(c["roots"]))
    assert r.count == c["count"]
    assert ctx.match(dg, p, roots=tuple(c["roots"]), flags=INDEPENDENT).count == c["count"]


@pytest.mark.parametrize("key", ["config2a", "config2b"])
def test_config2_full_counts_in_two_independent_modes(ctx, key):
    """The mid-size headline counts at full size (1.4e14 and 1.2e13
    embeddings, far beyond the oracle): the default mode (pair or chain
    level, two-hop index, shared candidates, dense two-hop rows at the hubs)
    and a mode that expands every level but the last (counted by the plain
    last-level closed form, itself oracle-pinned on every random case) must
    both give the committed count."""
    c = CFG[key + "_full"]
    g = W.config2_graph()
    assert W.csr_fingerprint(g) == c["fingerprint"]
    shape = {"config2a": W.C4_2PEND, "config2b": W.K4E_TAIL2}[key]
    p = cuts_amd.Plan(W.query_graph(shape))
    dg = ctx.upload(g)
    fast = ctx.match(dg, p)
    assert fast.count == c["count"]
    slow = ctx.match(dg, p, flags=INDEPENDENT)
    assert slow.count == c["count"]
    assert slow.level_sizes == fast.level_sizes


PLAIN = INDEPENDENT | cuts_amd.FLAG_NO_SHORTCUT | cuts_amd.FLAG_NO_FUSE


@pytest.mark.parametrize("key", ["config2a", "config2b"])
def test_config2_top_hub_roots_default_vs_plain_expansion(ctx, key):
    """The mid-size graph's two heaviest vertices (degrees 1,198 and 1,184)
    as the only roots: 3.5e12 (4-cycle) and 1.9e11 (K4-e) embeddings each,
    beyond the oracle.  The default mode (collapsed pair level or shared
    chain candidates, dense two-hop rows of the hubs) must equal the plain
    trie walk that expands and enumerates every level with no closed form at
    all -- the reference's own algorithm -- and the mode in between."""
    g = W.config2_graph()
    deg = np.diff(g.view()["out_off"].astype(np.int64))
    top = [int(v) for v in np.argsort(-deg, kind="stable")[:2]]
    assert deg[top].tolist() == [1198, 1184]
    shape = {"config2a": W.C4_2PEND, "config2b": W.K4E_TAIL2}[key]
    p = cuts_amd.Plan(W.query_graph(shape))
    dg = ctx.upload(g)
    v = top[0] if key == "config2b" else top[1]
    rt = (v, v + 1, 1)
    fast = ctx.match(dg, p, roots=rt)
    assert fast.count > 10 ** 11
    assert ctx.match(dg, p, roots=rt, flags=INDEPENDENT).count == fast.count
    plain = ctx.match(dg, p, roots=rt, flags=PLAIN)
    assert plain.complete and plain.count == fast.count


def test_cli_report_line_and_trunks(tmp_path):
    g = W.config1_graph()
    data = tmp_path / "cl4k.g"
    query = tmp_path / "tri2.g"
    g.save(str(data))
    W.query_graph(W.TRI_TAIL2).save(str(query))
    want = CFG["config1"]["count"]
    for extra in ([], ["4"], ["0"]):
        r = subprocess.run([cuts_amd.CLI_PATH, str(data), str(query)] + extra, capture_output=True, text=True,
                           timeout=120)
        assert r.returncode == 0, r.stdout + r.stderr
        lines = r.stdout.strip().splitlines()
        assert lines[0] == "start loading graph file from disk to memory..."
        assert lines[-2] == ("end copying graph to gpu..." if not extra else "finish copying graph to gpu...")
        m = re.fullmatch(r"cl4k\.g,tri2\.g,([0-9.e+-]+)ms,(\d+)", lines[-1])
        assert m, lines[-1]
        assert int(m.group(2)) == want


def test_cli_write_ans_txt(tmp_path):
    g, k, edges = random_case(31)
    og = oracle.OGraph.from_view(g.view())
    qa = oracle.qmatrix(k, oracle.undirected(edges))
    if oracle.count(og, qa) > 300000:
        pytest.skip("too many embeddings")
    want = oracle.enumerate_embeddings(og, qa, cap=1 << 21)
    g.save(str(tmp_path / "d.g"))
    q = cuts_amd.Graph.from_edges(k, edges)
    q.save(str(tmp_path / "q.g"))
    p = cuts_amd.Plan(q)
    r = subprocess.run([cuts_amd.CLI_PATH, "d.g", "q.g", "--write"], capture_output=True, text=True,
                       cwd=str(tmp_path), timeout=120)
    assert r.returncode == 0, r.stdout
    # basename of a path without '/' is "" (src/common.cpp:19-26)
    assert re.fullmatch(r",,[0-9.e+-]+ms,%d" % len(want), r.stdout.strip().splitlines()[-3])
    lines = (tmp_path / "ans.txt").read_text().splitlines()
    exp = sorted("".join(f"({u},{row[u]})" for u in p.order) for row in want.tolist())
    assert sorted(lines) == exp


def test_search_entry_points(tmp_path):
    g, k, edges = random_case(5)
    og = oracle.OGraph.from_view(g.view())
    want = oracle.count(og, oracle.qmatrix(k, oracle.undirected(edges)))
    g.save(str(tmp_path / "d.g"))
    cuts_amd.Graph.from_edges(k, edges).save(str(tmp_path / "q.g"))
    assert cuts_amd.search(str(tmp_path / "q.g"), str(tmp_path / "d.g")) == want
    assert cuts_amd.search_dfs_bfs_strategy(str(tmp_path / "q.g"), str(tmp_path / "d.g"), False, 3) == want


@pytest.mark.parametrize("seed", range(6))
def test_sparse_graphs_vs_oracle(ctx, seed):
    """E/V <= 1: the reference switches to its 8-lane virtual-warp kernel
    (src/host_funcs.cu:74,145; src/device_funcs.cu:618-733) for these; the
    flattened kernels need no separate path."""
    rng = random.Random(500 + seed)
    n = rng.randint(50, 400)
    edges = set()
    for v in range(1, n):
        if rng.random() < 0.9:
            edges.add((rng.randrange(v), v))
    g = cuts_amd.Graph.from_edges(n, sorted(edges))
    assert g.arcs <= 2 * n  # E/V <= 1 in undirected edges
    dg = ctx.upload(g)
    og = oracle.OGraph.from_view(g.view())
    for q in ([(0, 1), (1, 2)], [(0, 1), (1, 2), (2, 3)], [(0, 1), (0, 2), (0, 3)], [(0, 1), (1, 2), (1, 3), (3, 4)]):
        k = 1 + max(max(e) for e in q)
        want = oracle.count(og, oracle.qmatrix(k, oracle.undirected(q)))
        p = cuts_amd.Plan(cuts_amd.Graph.from_edges(k, q))
        for m in MODES:
            assert ctx.match(dg, p, **m).count == want, (q, m)


def test_table_log_from_cli_and_library(tmp_path):
    g, k, edges = random_case(12)
    g.save(str(tmp_path / "d.g"))
    q = cuts_amd.Graph.from_edges(k, edges)
    q.save(str(tmp_path / "q.g"))
    log = tmp_path / "tables.jsonl"
    env = dict(os.environ, CUTS_TABLE_LOG=str(log))
    r = subprocess.run([cuts_amd.CLI_PATH, "d.g", "q.g"], cwd=str(tmp_path), env=env, capture_output=True,
                       text=True, timeout=120)
    assert r.returncode == 0, r.stdout
    rec = json.loads(log.read_text().splitlines()[-1])
    ctx = cuts_amd.Context(0)
    p = cuts_amd.Plan(q)
    m = ctx.match(ctx.upload(g), p)
    assert rec["order"] == p.order and rec["level_sizes"] == m.level_sizes and rec["count"] == m.count
    cuts_amd.log_tables(str(log), "library", p, m)
    rec2 = json.loads(log.read_text().splitlines()[-1])
    assert rec2["label"] == "library" and rec2["level_sizes"] == rec["level_sizes"]
