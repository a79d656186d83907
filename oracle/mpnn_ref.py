"""TEST INFRASTRUCTURE ONLY -- FP64 restatement of the table-size network.

The same computation as the HIP kernels in cuts_amd/csrc/mpnn.hip, written
with torch tensors in float64 on the CPU so that torch's autograd gives an
independent gradient for the hand-written backward kernel.  Weights are
rounded to BF16 first (the HIP forward reads BF16 weights), so the only
difference left is FP32 vs FP64 accumulation.  Pinned by central finite
differences in tests/test_learned.py.
"""
from __future__ import annotations

import numpy as np

H, F, N = 64, 24, 16


def bf16_round(p: np.ndarray) -> np.ndarray:
    """Round-to-nearest-even to BF16, returned as float64."""
    u = np.ascontiguousarray(p, dtype=np.float32).view(np.uint32).astype(np.uint64)
    r = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    return r.astype(np.uint32).view(np.float32).astype(np.float64)


def split(p):
    """Views of the flat parameter vector (input-major matrices)."""
    import torch
    o = 0
    out = {}

    def take(name, *shape):
        nonlocal o
        n = int(np.prod(shape))
        out[name] = p[o:o + n].reshape(*shape)
        o += n

    take("ws0", F, H), take("wn0", F, H), take("b0", H)
    take("ws1", H, H), take("wn1", H, H), take("b1", H)
    take("ws2", H, H), take("wn2", H, H), take("b2", H)
    take("wr", 2 * H, H), take("br", H), take("wo", H), take("bo", 1)
    assert o == p.shape[0]
    return out


def forward(p, x, adj, meta):
    """p: torch float64 [P]; x: [n,16,24]; adj: [n,16] masks; meta: [n].
    Returns y [n] (torch)."""
    import torch
    W = split(p)
    ys = []
    for s in range(len(meta)):
        k = int(meta[s]) & 0xFF
        sm = int(meta[s]) >> 16
        A = torch.zeros((k, k), dtype=torch.float64)
        for v in range(k):
            for u in range(k):
                if (int(adj[s][v]) >> u) & 1:
                    A[v, u] = 1.0
        h = torch.as_tensor(x[s, :k, :], dtype=torch.float64)
        for l in range(3):
            agg = A @ h
            h = torch.relu(W[f"b{l}"] + h @ W[f"ws{l}"] + agg @ W[f"wn{l}"])
        mask = torch.tensor([(sm >> v) & 1 for v in range(k)], dtype=torch.float64)
        c = torch.cat([(h * mask[:, None]).sum(0), h.mean(0)])
        r = torch.relu(W["br"] + c @ W["wr"])
        ys.append(W["bo"][0] + r @ W["wo"])
    return torch.stack(ys)


def loss_and_grad(params: np.ndarray, x, adj, meta, t):
    """Mean squared error and its gradient w.r.t. the BF16-rounded weights."""
    import torch
    p = torch.tensor(bf16_round(params), dtype=torch.float64, requires_grad=True)
    y = forward(p, x, adj, meta)
    tt = torch.as_tensor(np.asarray(t, dtype=np.float64))
    loss = ((y - tt) ** 2).mean()
    loss.backward()
    return float(loss.detach()), p.grad.detach().numpy().copy(), y.detach().numpy()
