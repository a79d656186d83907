"""TEST INFRASTRUCTURE ONLY -- ctypes binding of the CPU oracle.

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg import
this package.  The product (cuts_amd) never does.

``liboracle.so`` (oracle.c) holds the independent backtracking counter, the
OpenMP CPU trie baseline and a second restatement of the score order.
``_ref/libcutsref.so`` (when built from the reference tree by
``make -C oracle ref``) exposes the reference's own host C++: its matching
order, CSR construction and ans.txt writer.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess
from typing import Optional, Sequence

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(_HERE, "liboracle.so")
REF_LIB = os.path.join(_HERE, "_ref", "libcutsref.so")
_u32p = C.POINTER(C.c_uint32)


class _OGraph(C.Structure):
    _fields_ = [("V", C.c_uint32), ("off_out", _u32p), ("adj_out", _u32p), ("off_in", _u32p),
                ("adj_in", _u32p), ("sig", _u32p)]


class TrieResult(C.Structure):
    _fields_ = [("count", C.c_uint64), ("level", C.c_uint64 * 32), ("bytes", C.c_double),
                ("seconds", C.c_double), ("threads", C.c_int), ("complete", C.c_int)]


_lib = None
_ref = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            subprocess.run(["make", "-C", _HERE, "liboracle.so"], check=True, capture_output=True)
        L = C.CDLL(LIB)
        L.oracle_count.restype = C.c_uint64
        L.oracle_count.argtypes = [C.POINTER(_OGraph), C.c_uint32, C.POINTER(C.c_uint8), C.c_int]
        L.oracle_enumerate.restype = C.c_uint64
        L.oracle_enumerate.argtypes = [C.POINTER(_OGraph), C.c_uint32, C.POINTER(C.c_uint8), _u32p, C.c_uint64]
        L.oracle_heuristic_order.restype = C.c_int
        L.oracle_heuristic_order.argtypes = [C.c_uint32, C.POINTER(C.c_uint8), _u32p]
        L.trie_cpu_run.restype = C.c_int
        L.trie_cpu_run.argtypes = [C.POINTER(_OGraph), C.c_uint32, C.POINTER(C.c_uint8), _u32p, C.c_uint32,
                                   C.c_uint64, C.c_int, C.POINTER(TrieResult)]
        L.trie_cpu_run_timed.restype = C.c_int
        L.trie_cpu_run_timed.argtypes = [C.POINTER(_OGraph), C.c_uint32, C.POINTER(C.c_uint8), _u32p, C.c_uint32,
                                         C.c_uint64, C.c_int, C.c_double, C.POINTER(TrieResult)]
        L.trie_cpu_run_roots.restype = C.c_int
        L.trie_cpu_run_roots.argtypes = [C.POINTER(_OGraph), C.c_uint32, C.POINTER(C.c_uint8), _u32p, C.c_uint32,
                                         C.c_uint32, C.c_uint32, C.c_uint32, C.c_uint64, C.c_int, C.c_double,
                                         C.POINTER(TrieResult)]
        L.oracle_max_threads.restype = C.c_int
        L.oracle_max_threads.argtypes = []
        _lib = L
    return _lib


class OGraph:
    """Keeps the numpy CSR arrays alive next to the C view."""

    def __init__(self, V, out_off, out_adj, in_off, in_adj, sig):
        self.arrs = [np.ascontiguousarray(a, dtype=np.uint32) for a in (out_off, out_adj, in_off, in_adj, sig)]
        p = [a.ctypes.data_as(_u32p) for a in self.arrs]
        self.c = _OGraph(V, *p)
        self.V = V

    @classmethod
    def from_view(cls, v: dict) -> "OGraph":
        return cls(v["V"], v["out_off"], v["out_adj"], v["in_off"], v["in_adj"], v["sig"])

    @classmethod
    def from_arcs(cls, V: int, arcs: Sequence[Sequence[int]]) -> "OGraph":
        """CSR built with numpy (sorted, duplicate-free), independent of cuts_amd."""
        a = np.unique(np.asarray(arcs, dtype=np.int64).reshape(-1, 2), axis=0) if len(arcs) else \
            np.zeros((0, 2), np.int64)
        out_off = np.zeros(V + 1, np.int64)
        np.add.at(out_off, a[:, 0] + 1, 1)
        out_off = np.cumsum(out_off)
        ra = a[np.lexsort((a[:, 0], a[:, 1]))]
        in_off = np.zeros(V + 1, np.int64)
        np.add.at(in_off, ra[:, 1] + 1, 1)
        in_off = np.cumsum(in_off)
        sig = np.zeros(2 * V, np.int64)
        sig[0::2] = np.diff(in_off)
        sig[1::2] = np.diff(out_off)
        return cls(V, out_off, a[:, 1], in_off, ra[:, 0], sig)


def qmatrix(k: int, arcs: Sequence[Sequence[int]]) -> np.ndarray:
    m = np.zeros((k, k), dtype=np.uint8)
    for s, t in arcs:
        m[s, t] = 1
    return m


def undirected(edges):
    out = []
    for s, t in edges:
        out.append((s, t))
        out.append((t, s))
    return out


def _qptr(qa: np.ndarray):
    qa = np.ascontiguousarray(qa, dtype=np.uint8)
    return qa, qa.ctypes.data_as(C.POINTER(C.c_uint8))


def count(g: OGraph, qa: np.ndarray, nthreads: int = 0) -> int:
    qa, p = _qptr(qa)
    return int(lib().oracle_count(C.byref(g.c), qa.shape[0], p, nthreads))


def enumerate_embeddings(g: OGraph, qa: np.ndarray, cap: int = 1 << 20) -> np.ndarray:
    """Rows indexed by query vertex: row[u] = f(u)."""
    qa, p = _qptr(qa)
    k = qa.shape[0]
    rows = np.zeros((cap, k), dtype=np.uint32)
    n = int(lib().oracle_enumerate(C.byref(g.c), k, p, rows.ctypes.data_as(_u32p), cap))
    if n > cap:
        raise ValueError(f"{n} embeddings exceed cap {cap}")
    return rows[:n]


def max_threads() -> int:
    """omp_get_max_threads() inside the oracle library."""
    return int(lib().oracle_max_threads())


def heuristic_order(qa: np.ndarray) -> Optional[list]:
    qa, p = _qptr(qa)
    k = qa.shape[0]
    o = np.zeros(k, np.uint32)
    rc = lib().oracle_heuristic_order(k, p, o.ctypes.data_as(_u32p))
    return None if rc else o.tolist()


def trie_cpu(g: OGraph, qa: np.ndarray, order: Sequence[int], sample_stride: int = 1, max_rows: int = 0,
             nthreads: int = 0, max_seconds: float = 0.0, roots=None) -> TrieResult:
    """roots: (first, end, stride) restricts level 0 as the matcher's root
    samples do (before sample_stride applies)."""
    qa, p = _qptr(qa)
    o = np.ascontiguousarray(order, dtype=np.uint32)
    r = TrieResult()
    rb, re, rs = roots if roots is not None else (0, g.V, 1)
    rc = lib().trie_cpu_run_roots(C.byref(g.c), qa.shape[0], p, o.ctypes.data_as(_u32p), rb, re, rs,
                                  sample_stride, max_rows, nthreads, max_seconds, C.byref(r))
    if rc:
        raise ValueError("trie_cpu_run failed")
    return r


# ---- the reference's own host C++ (oracle/_ref) -------------------------

def ref_available() -> bool:
    return os.path.exists(REF_LIB)


def ref():
    global _ref
    if _ref is None:
        L = C.CDLL(REF_LIB)
        L.ref_query_plan.argtypes = [C.c_char_p] + [_u32p] * 6
        L.ref_graph_size.argtypes = [C.c_char_p, _u32p, _u32p]
        L.ref_graph_csr.argtypes = [C.c_char_p] + [_u32p] * 5
        L.ref_write_matches.argtypes = [C.c_uint64, C.POINTER(C.c_uint64), C.c_uint32, _u32p, _u32p]
        L.ref_write_matches.restype = None
        _ref = L
    return _ref


def ref_query_plan(path: str) -> dict:
    """Order and P/C CSR that the reference computes for a query file."""
    with open(path) as f:
        V = int(f.readline())
    a = [np.zeros(n, np.uint32) for n in (1, V, V + 1, V * V + 1, V + 1, V * V + 1)]
    ref().ref_query_plan(path.encode(), *[x.ctypes.data_as(_u32p) for x in a])
    Vv, order, po, par, co, chi = a
    return dict(V=int(Vv[0]), order=order.tolist(), parents_offset=po.tolist(), parents=par[: po[-1]].tolist(),
                children_offset=co.tolist(), children=chi[: co[-1]].tolist())


def ref_graph_csr(path: str) -> dict:
    V, E = np.zeros(1, np.uint32), np.zeros(1, np.uint32)
    ref().ref_graph_size(path.encode(), V.ctypes.data_as(_u32p), E.ctypes.data_as(_u32p))
    n, e = int(V[0]), int(E[0])
    a = [np.zeros(max(x, 1), np.uint32) for x in (n + 1, e, n + 1, e, 2 * n)]
    ref().ref_graph_csr(path.encode(), *[x.ctypes.data_as(_u32p) for x in a])
    return dict(V=n, arcs=e, out_off=a[0], out_adj=a[1][:e], in_off=a[2], in_adj=a[3][:e], sig=a[4][: 2 * n])


def ref_write_matches(workdir: str, rows: np.ndarray, order: Sequence[int]) -> str:
    """Runs the reference writer on rows (one embedding per row, matching
    order positions), each row encoded as its own prefix+candidate record."""
    rows = np.ascontiguousarray(rows, dtype=np.uint32)
    n, k = rows.shape
    ptrs = np.zeros(2 * n, np.uint64)
    ptrs[0::2] = np.arange(n, dtype=np.uint64) * k
    ptrs[1::2] = ptrs[0::2] + k
    o = np.ascontiguousarray(order, dtype=np.uint32)
    cwd = os.getcwd()
    try:
        os.chdir(workdir)
        ref().ref_write_matches(n, ptrs.ctypes.data_as(C.POINTER(C.c_uint64)), k, o.ctypes.data_as(_u32p),
                                rows.reshape(-1).ctypes.data_as(_u32p))
    finally:
        os.chdir(cwd)
    return os.path.join(workdir, "ans.txt")


class _OOpts(C.Structure):
    _fields_ = [("thr_in", _u32p), ("thr_out", _u32p), ("root_q", C.c_int32), ("root_begin", C.c_uint32),
                ("root_end", C.c_uint32), ("root_stride", C.c_uint32)]


def enumerate_ex(g: OGraph, qa: np.ndarray, root_q: int, roots, cap: int = 1 << 20) -> np.ndarray:
    """enumerate_embeddings restricted to query vertex root_q's images in
    range(*roots).  Rows indexed by query vertex."""
    L = lib()
    L.oracle_enumerate_ex.restype = C.c_uint64
    L.oracle_enumerate_ex.argtypes = [C.POINTER(_OGraph), C.c_uint32, C.POINTER(C.c_uint8), C.POINTER(_OOpts),
                                      _u32p, C.c_uint64]
    qa, p = _qptr(qa)
    k = qa.shape[0]
    o = _OOpts()
    o.root_q = root_q
    o.root_begin, o.root_end, o.root_stride = roots
    rows = np.zeros((cap, k), dtype=np.uint32)
    n = int(L.oracle_enumerate_ex(C.byref(g.c), k, p, C.byref(o), rows.ctypes.data_as(_u32p), cap))
    if n > cap:
        raise ValueError(f"{n} embeddings exceed cap {cap}")
    return rows[:n]


def count_ex(g: OGraph, qa: np.ndarray, thr_in=None, thr_out=None, root_q: int = -1, roots=(0, 0, 1),
             nthreads: int = 0) -> int:
    """oracle_count with explicit per-vertex signature thresholds and/or a
    restriction of query vertex root_q's image to range(*roots)."""
    L = lib()
    L.oracle_count_ex.restype = C.c_uint64
    L.oracle_count_ex.argtypes = [C.POINTER(_OGraph), C.c_uint32, C.POINTER(C.c_uint8), C.POINTER(_OOpts),
                                  C.c_int]
    qa, p = _qptr(qa)
    keep = []
    o = _OOpts()
    if thr_in is not None:
        a = np.ascontiguousarray(thr_in, dtype=np.uint32)
        keep.append(a)
        o.thr_in = a.ctypes.data_as(_u32p)
    if thr_out is not None:
        a = np.ascontiguousarray(thr_out, dtype=np.uint32)
        keep.append(a)
        o.thr_out = a.ctypes.data_as(_u32p)
    o.root_q = root_q
    o.root_begin, o.root_end, o.root_stride = roots
    return int(L.oracle_count_ex(C.byref(g.c), qa.shape[0], p, C.byref(o), nthreads))
