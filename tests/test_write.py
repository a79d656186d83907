"""Enumeration at scale: 4.8e8 embeddings of a 3-edge path written to pinned
host memory in chunks whose copies overlap the next chunk's expansion.
Pinned by a closed form for the count (ordered injective 3-paths:
sum over arcs (b,c) of (d_b - 1)(d_c - 1) minus 6 x triangles), by the
oracle's enumeration of a root sample, and by the synchronous-copy mode."""
import numpy as np
import pytest
import scipy.sparse as sp

import cuts_amd
import oracle
from cuts_amd import workloads as W

pytestmark = pytest.mark.gpu


def three_paths(g):
    v = g.view()
    V = v["V"]
    d = np.diff(v["out_off"].astype(np.int64))
    src = np.repeat(np.arange(V), d)
    dst = v["out_adj"].astype(np.int64)
    A = sp.csr_matrix((np.ones(len(src), np.int64), (src, dst)), shape=(V, V))
    tri6 = int(A.multiply(A @ A).sum())
    return int(((d[src] - 1) * (d[dst] - 1)).sum()) - tri6


def test_write_half_a_billion_embeddings_streamed(ctx):
    g = W.config1_graph()
    V = g.V
    edges = [(0, 1), (1, 2), (2, 3)]
    want = three_paths(g)
    assert want > 10 ** 8
    dg = ctx.upload(g)
    p = cuts_amd.Plan(cuts_amd.Graph.from_edges(4, edges))
    r = ctx.match(dg, p, write=True, emb_cap=want)
    assert r.count == want and r.rows.shape == (want, 4)
    assert r.chunks > 10  # the last level went out in many overlapped chunks
    rows = r.rows
    # every row injective; a root sample equals the oracle's enumeration
    assert not np.any(rows[:, 0] == rows[:, 2]) and not np.any(rows[:, 1] == rows[:, 3])
    # (position 0's images v = 400 mod 401: 251,785 rows; the class of vertex
    # 0, the heaviest hub, alone holds 5.7e7)
    s = 401
    sample = rows[rows[:, 0] % s == s - 1]
    og = oracle.OGraph.from_view(g.view())
    qa = oracle.qmatrix(4, oracle.undirected(edges))
    emb = oracle.enumerate_ex(og, qa, root_q=p.order[0], roots=(s - 1, V, s), cap=4 << 20)
    assert len(emb) > 10 ** 5
    mine = sample[np.lexsort(sample.T[::-1])]
    ref = emb[:, p.order]
    ref = ref[np.lexsort(ref.T[::-1])]
    assert mine.shape == ref.shape and np.array_equal(mine, ref)
    # the synchronous-copy mode writes the same number of rows, and a capped
    # buffer receives exactly its capacity
    r2 = ctx.match(dg, p, write=True, emb_cap=want, flags=cuts_amd.FLAG_SYNC_WRITE)
    assert r2.count == want and len(r2.rows) == want
    del r2
    r3 = ctx.match(dg, p, write=True, emb_cap=1000)
    assert r3.count == want and len(r3.rows) == 1000
