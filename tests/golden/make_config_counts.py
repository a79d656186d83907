"""Writes tests/golden/config_counts.json with the oracle's counts for the
benchmark workloads: the full config-1 count, and for the larger configs the
count restricted to a sample of position-0 images (root_begin + i*stride),
small enough for the CPU oracle.  Each entry records the data-graph
fingerprint so a generator change cannot pass silently.

Run from the repository root after building cuts_amd and the oracle.
"""
import json
import os
import sys
import time

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

import cuts_amd  # noqa: E402
import oracle  # noqa: E402
from cuts_amd import workloads as W  # noqa: E402

SAMPLES = {
    "config2a": (W.config2_graph, W.C4_2PEND, (2000, 36692, 211)),
    "config2b": (W.config2_graph, W.K4E_TAIL2, (2000, 36692, 211)),
    # hub roots: vertices 500, 1000, 1500, 2000 -- degrees 71, 42, 28 and 31
    # against a mean of 10, among the top 6% by degree (Chung-Lu puts the
    # heaviest weights first); the heaviest vertices (degree ~1,150) have
    # 1e12+ embeddings each, beyond the oracle, and are pinned on the GPU by
    # two independent execution modes instead
    "config2a_hubs": (W.config2_graph, W.C4_2PEND, (500, 2001, 500)),
    "config2b_hubs": (W.config2_graph, W.K4E_TAIL2, (500, 2001, 500)),
    # v = 511 (mod 512): 2,048 roots of mostly low degree; the full count
    # exceeds 2^64, this sample has 3.3e8 embeddings (the oracle needs ~20 s)
    "config3": (W.config3_graph, W.C5CHORD_2TAIL, (511, 1 << 20, 512)),
}


def main(keys):
    path = os.path.join(HERE, "config_counts.json")
    out = json.load(open(path)) if os.path.exists(path) else {}
    for key in keys:
        t0 = time.time()
        if key == "config1":
            g = W.config1_graph()
            og = oracle.OGraph.from_view(g.view())
            qa = oracle.qmatrix(5, W.shape_arcs(W.TRI_TAIL2))
            c = oracle.count(og, qa)
            out[key] = dict(fingerprint=W.csr_fingerprint(g), shape=W.TRI_TAIL2.name, count=c)
        else:
            gf, shape, roots = SAMPLES[key]
            g = gf()
            og = oracle.OGraph.from_view(g.view())
            qa = oracle.qmatrix(shape.k, W.shape_arcs(shape))
            root_q = cuts_amd.Plan(W.query_graph(shape)).order[0]
            c = oracle.count_ex(og, qa, root_q=root_q, roots=roots)
            out[key] = dict(fingerprint=W.csr_fingerprint(g), shape=shape.name, root_q=root_q, roots=list(roots),
                            count=c)
        print(key, out[key], f"{time.time() - t0:.1f}s", flush=True)
        with open(path, "w") as f:
            json.dump(out, f, indent=1)


if __name__ == "__main__":
    main(sys.argv[1:] or ["config1", "config2a", "config2b", "config2a_hubs", "config2b_hubs", "config3"])
