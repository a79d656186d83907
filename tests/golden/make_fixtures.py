"""Writes the golden fixtures under tests/golden/.  Run in the build
container (it needs /root/reference and oracle/_ref built by
``make -C oracle ref``); the committed JSON files are what the tests read.

* nx_counts.json   -- monomorphism counts from networkx 3.4.2
  (DiGraphMatcher.subgraph_monomorphisms_iter) on seeded random graphs, plus
  the reference's 33 query graphs matched into small random digraphs.
* ref_plans.json   -- the reference's 33 query files (data) and the order /
  P / C arrays that the reference's own Graph::sort_search_order computes
  for them and for seeded random queries (oracle/_ref).
* ref_csr.json     -- CSR + signatures the reference's Graph ctor builds for
  hand-written edge-list files exercising the reader's quirks.
* ref_ans.json     -- ans.txt text produced by the reference's writer.
"""
import json
import os
import random
import sys
import tempfile

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

import networkx as nx  # noqa: E402
import numpy as np  # noqa: E402

import oracle  # noqa: E402

REF = "/root/reference"
QDIR = os.path.join(REF, "data_set/query/ours_format")


def read_edge_list(path):
    with open(path) as f:
        V = int(f.readline())
        arcs = []
        for line in f:
            if "\t" in line:
                s, t = line.split("\t")[:2]
                arcs.append((int(s), int(t)))
    return V, sorted(set(arcs))


def nx_count(V, arcs, k, qarcs):
    G = nx.DiGraph()
    G.add_nodes_from(range(V))
    G.add_edges_from(arcs)
    Q = nx.DiGraph()
    Q.add_nodes_from(range(k))
    Q.add_edges_from(qarcs)
    gm = nx.algorithms.isomorphism.DiGraphMatcher(G, Q)
    return sum(1 for _ in gm.subgraph_monomorphisms_iter())


def rand_graph(rng, n, p, directed):
    arcs = set()
    for u in range(n):
        for v in range(n):
            if u == v:
                continue
            if not directed and v < u:
                continue
            if rng.random() < p:
                arcs.add((u, v))
                if not directed:
                    arcs.add((v, u))
    return sorted(arcs)


def rand_query(rng, k, extra, directed):
    """Connected random query: a random spanning tree plus extra edges."""
    edges = set()
    for v in range(1, k):
        u = rng.randrange(v)
        edges.add((u, v))
    pairs = [(u, v) for u in range(k) for v in range(u + 1, k) if (u, v) not in edges]
    rng.shuffle(pairs)
    edges.update(pairs[:extra])
    # random relabelling so that vertex 0 is not always the tree root
    perm = list(range(k))
    rng.shuffle(perm)
    arcs = set()
    for u, v in edges:
        a, b = perm[u], perm[v]
        if directed:
            r = rng.random()
            if r < 0.4:
                arcs.add((a, b))
            elif r < 0.8:
                arcs.add((b, a))
            else:
                arcs.add((a, b))
                arcs.add((b, a))
        else:
            arcs.add((a, b))
            arcs.add((b, a))
    return sorted(arcs)


def write_graph(path, V, arcs):
    with open(path, "w") as f:
        f.write(f"{V}\n")
        for s, t in arcs:
            f.write(f"{s}\t{t}\n")


def make_nx():
    rng = random.Random(20211114)
    cases = []
    for i in range(60):
        directed = i % 3 == 0
        n = rng.randint(6, 22)
        p = rng.uniform(0.15, 0.55)
        arcs = rand_graph(rng, n, p, directed)
        k = rng.randint(2, 5)
        extra = rng.randint(0, k)
        q = rand_query(rng, k, extra, directed)
        c = nx_count(n, arcs, k, q)
        cases.append(dict(name=f"random{i}", V=n, arcs=arcs, k=k, qarcs=q, count=c))
    # the reference's directed query graphs into small random digraphs
    for qi in range(1, 34):
        k, q = read_edge_list(os.path.join(QDIR, f"{qi}.g"))
        n = rng.randint(9, 13)
        arcs = rand_graph(rng, n, rng.uniform(0.55, 0.8), True)
        c = nx_count(n, arcs, k, q)
        cases.append(dict(name=f"refquery{qi}", V=n, arcs=arcs, k=k, qarcs=q, count=c))
    with open(os.path.join(HERE, "nx_counts.json"), "w") as f:
        json.dump(dict(source="networkx " + nx.__version__ + " DiGraphMatcher.subgraph_monomorphisms_iter",
                       cases=cases), f, separators=(",", ":"))
    print("nx_counts.json:", len(cases), "cases")


def make_plans():
    rng = random.Random(7)
    out = []
    tmp = tempfile.mkdtemp()
    for qi in range(1, 34):
        path = os.path.join(QDIR, f"{qi}.g")
        k, q = read_edge_list(path)
        plan = oracle.ref_query_plan(path)
        out.append(dict(name=f"ref{qi}.g", k=k, qarcs=q, **{x: plan[x] for x in
                                                           ("order", "parents_offset", "parents",
                                                            "children_offset", "children")}))
    for i in range(150):
        directed = i % 2 == 0
        k = rng.randint(2, 12)
        q = rand_query(rng, k, rng.randint(0, 2 * k), directed)
        path = os.path.join(tmp, f"q{i}.g")
        write_graph(path, k, q)
        plan = oracle.ref_query_plan(path)
        out.append(dict(name=f"random{i}", k=k, qarcs=q, **{x: plan[x] for x in
                                                           ("order", "parents_offset", "parents",
                                                            "children_offset", "children")}))
    with open(os.path.join(HERE, "ref_plans.json"), "w") as f:
        json.dump(dict(source="reference Graph(0, path) compiled by oracle/Makefile", plans=out), f,
                  separators=(",", ":"))
    print("ref_plans.json:", len(out), "plans")


CSR_FILES = {
    "tabs_and_noise": "6\n0\t1\n0\t1\n1\t2\n# comment\n2 3\n3\t0\n5\t4\n4\t5\n\n1\t1\n",
    "crlf_and_spaces": "5\r\n0\t1\r\n 1\t 2\r\n2\t3 extra\r\n3\t4\t9\r\n4\t0\r\n",
    "isolated": "8\n7\t0\n0\t7\n3\t3\n",
    "header_only": "4\n",
}


def make_csr():
    tmp = tempfile.mkdtemp()
    out = []
    for name, text in CSR_FILES.items():
        path = os.path.join(tmp, name + ".g")
        with open(path, "w", newline="") as f:
            f.write(text)
        csr = oracle.ref_graph_csr(path)
        out.append(dict(name=name, text=text, V=csr["V"], arcs=csr["arcs"],
                        **{x: csr[x].tolist() for x in ("out_off", "out_adj", "in_off", "in_adj", "sig")}))
    with open(os.path.join(HERE, "ref_csr.json"), "w") as f:
        json.dump(dict(source="reference Graph(1, path) compiled by oracle/Makefile", files=out), f,
                  separators=(",", ":"))
    print("ref_csr.json:", len(out), "files")


def make_ans():
    tmp = tempfile.mkdtemp()
    rng = np.random.default_rng(3)
    cases = []
    for k in (2, 5, 7):
        rows = rng.integers(0, 10**6, size=(6, k), dtype=np.uint32)
        order = rng.permutation(k).astype(np.uint32)
        path = oracle.ref_write_matches(tmp, rows, order)
        with open(path) as f:
            text = f.read()
        cases.append(dict(k=k, order=order.tolist(), rows=rows.tolist(), text=text))
    with open(os.path.join(HERE, "ref_ans.json"), "w") as f:
        json.dump(dict(source="reference write_match_to_disk compiled by oracle/Makefile", cases=cases), f,
                  separators=(",", ":"))
    print("ref_ans.json:", len(cases), "cases")


def make_gsi():
    """The reference's own query files in both formats (data): ours_format
    {1..33}.g and the matching gsi_format/{5,6,7}nodes/{0..10}.g."""
    cases = []
    for q in range(1, 34):
        d = "5nodes" if q <= 11 else ("6nodes" if q <= 22 else "7nodes")
        i = (q - 1) % 11
        with open(os.path.join(QDIR, f"{q}.g")) as f:
            ours = f.read()
        with open(os.path.join(REF, "data_set/query/gsi_format", d, f"{i}.g")) as f:
            gsi = f.read()
        cases.append(dict(ours=ours, gsi=gsi, name=f"{q}.g"))
    with open(os.path.join(HERE, "ref_gsi.json"), "w") as f:
        json.dump(dict(source="reference data_set/query ours_format and gsi_format files", cases=cases), f,
                  separators=(",", ":"))
    print("ref_gsi.json:", len(cases), "pairs")


if __name__ == "__main__":
    if not oracle.ref_available():
        sys.exit("build oracle/_ref first: make -C oracle ref")
    make_nx()
    make_plans()
    make_csr()
    make_ans()
    make_gsi()
