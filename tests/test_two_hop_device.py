"""The two-hop index an upload builds on the device (k_two_hop, graphs of up
to 2^17 vertices) against the host build of Graph.two_hop() -- which the
CPU suite pins to the sparse product A*A with its diagonal removed -- entry
for entry, and its dense uint16 rows against rows rebuilt from the host
lists with the same selection rule (every vertex with at least V/64
entries, longest first).  Larger graphs keep the host build."""
import numpy as np
import pytest
import scipy.sparse as sp

import cuts_amd
from cuts_amd import workloads as W

pytestmark = pytest.mark.gpu


def host_dense(off, v, c, V):
    n = np.diff(off.astype(np.int64))
    cand = [u for u in range(V) if n[u] * 64 >= V]
    cand.sort(key=lambda u: (-n[u], u))
    did = np.full(V, -1, dtype=np.int32)
    rows = np.zeros((len(cand), V), dtype=np.uint16)
    for r, u in enumerate(cand):
        did[u] = r
        rows[r, v[off[u]:off[u + 1]]] = c[off[u]:off[u + 1]]
    return did, rows


def check(ctx, g, on_device=True):
    dg = ctx.upload(g)
    info = dg.info()
    assert info["two_hop_on_device"] == int(on_device), info
    off, v, c = g.two_hop()
    doff, dv, dc, did, drows = dg.two_hop(dense=True)
    assert info["two_hop_entries"] == len(v)
    assert np.array_equal(doff, off)
    assert np.array_equal(dv, v)
    assert np.array_equal(dc, c)
    hid, hrows = host_dense(off, v, c, g.V)
    assert info["two_hop_dense_rows"] == len(hrows)
    assert np.array_equal(did, hid)
    assert np.array_equal(drows, hrows)
    print(f"V={g.V} entries={len(v)} dense rows={len(hrows)} built in {info['two_hop_build_s'] * 1e3:.1f} ms "
          f"({'device' if on_device else 'host'})", flush=True)
    return dg


@pytest.mark.parametrize("seed", [0, 1, 2])
def test_device_two_hop_equals_host_chung_lu(ctx, seed):
    check(ctx, cuts_amd.Graph.chung_lu(2000 + 37 * seed, 9000, 2.2, 0, 50 + seed))


def test_device_two_hop_equals_sparse_square(ctx):
    """Directly against A*A (diagonal removed) on an R-MAT graph."""
    g = cuts_amd.Graph.rmat(11, 8, seed=4)
    dg = check(ctx, g)
    vw = g.view()
    o, adj = vw["out_off"].astype(np.int64), vw["out_adj"].astype(np.int64)
    A = sp.csr_matrix((np.ones(len(adj), dtype=np.int64), (np.repeat(np.arange(g.V), np.diff(o)), adj)),
                      shape=(g.V, g.V))
    A2 = (A @ A).tolil()
    A2.setdiag(0)
    A2 = A2.tocsr()
    A2.eliminate_zeros()
    A2.sort_indices()
    off, v, c = dg.two_hop()
    assert np.array_equal(off.astype(np.int64), A2.indptr.astype(np.int64))
    assert np.array_equal(v.astype(np.int64), A2.indices.astype(np.int64))
    assert np.array_equal(c.astype(np.int64), A2.data.astype(np.int64))


def test_device_two_hop_small_cases(ctx):
    """A star (the centre has no two-hop vertex, every leaf has all other
    leaves: dense rows for every leaf), K5, and a graph with isolated
    vertices."""
    check(ctx, cuts_amd.Graph.from_edges(70, [(0, i) for i in range(1, 70)]))
    check(ctx, cuts_amd.Graph.from_edges(5, [(a, b) for a in range(5) for b in range(a + 1, 5)]))
    check(ctx, cuts_amd.Graph.from_edges(300, [(3 * i, 3 * i + 3) for i in range(60)] + [(1, 250), (250, 299)]))


def test_single_edge_has_an_empty_index(ctx):
    dg = ctx.upload(cuts_amd.Graph.from_edges(2, [(0, 1)]))
    assert dg.info()["two_hop_entries"] == 0
    with pytest.raises(cuts_amd.CutsError, match="two-hop"):
        dg.two_hop()


def test_device_two_hop_config_graphs(ctx):
    """The 4,096- and 36,692-vertex workload graphs: built on the device."""
    check(ctx, W.config1_graph())
    check(ctx, W.config2_graph())


def test_large_graph_keeps_the_host_build(ctx):
    check(ctx, cuts_amd.Graph.chung_lu((1 << 17) + 5000, 400000, 2.5, 0, 8), on_device=False)
