"""cuts_amd -- MI355X trie-based subgraph matcher, Python binding.

A thin ctypes layer over ``libcuts_amd.so`` (C ABI in ``include/cuts.h``).
The matcher itself runs in hand-written HIP kernels; this module only moves
arguments across the boundary.  It raises ``RuntimeError`` when the native
library has not been built -- there is no Python or CPU fallback.

Mirrors of the reference's host functions (``inc/host_funcs.h:21-25``):
``search(query_file, data_file, write_to_disk)`` and
``search_dfs_bfs_strategy(query_file, data_file, write_to_disk, trunks)``.
"""
from __future__ import annotations

import ctypes as C
import os
from dataclasses import dataclass, field
from typing import List, Optional, Sequence, Tuple

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libcuts_amd.so")
CLI_PATH = os.path.join(_HERE, "cuts")

FLAG_NO_SHORTCUT = 1
FLAG_NO_FUSE = 2
FLAG_WRITE = 4
FLAG_NO_XCD = 8
FLAG_PROFILE = 16
FLAG_NO_BITMAP = 32
FLAG_NO_PREFIX_BASE = 64
FLAG_NO_PAIR = 128
FLAG_NO_CHAIN = 256
FLAG_NO_TWO_HOP = 512
FLAG_NO_SHARED = 1024
FLAG_NO_ARC_COMMON = 2048
FLAG_SYNC_WRITE = 4096
FLAG_SYNC_LEVELS = 8192   # read every level's work back before expanding it
MAX_QUERY = 16
KERNEL_CLASSES = ["level0", "plan", "scan", "tile_index", "expand_write", "expand_count", "expand_fused",
                  "closed_form", "pair", "common", "chain"]
_NCLASS = len(KERNEL_CLASSES)

_u32p = C.POINTER(C.c_uint32)


class _GraphView(C.Structure):
    _fields_ = [("V", C.c_uint32), ("arcs", C.c_uint64), ("symmetric", C.c_int),
                ("out_off", _u32p), ("out_adj", _u32p), ("in_off", _u32p), ("in_adj", _u32p),
                ("sig", _u32p)]


class _Opts(C.Structure):
    _fields_ = [("trunks", C.c_uint32), ("flags", C.c_uint32), ("budget_bytes", C.c_uint64),
                ("root_begin", C.c_uint32), ("root_end", C.c_uint32), ("root_stride", C.c_uint32),
                ("blocks_per_cu", C.c_uint32), ("emb_rows", _u32p), ("emb_cap", C.c_uint64),
                ("time_limit_ms", C.c_double), ("max_work", C.c_uint64), ("expand_threads", C.c_uint32),
                ("class_blocks_per_cu", C.c_uint8 * 16)]


class _Result(C.Structure):
    _fields_ = [("count", C.c_uint64), ("level_sizes", C.c_uint64 * 16), ("device_ms", C.c_double),
                ("emb_written", C.c_uint64), ("chunks", C.c_uint32), ("complete", C.c_uint32),
                ("kernel_ms", C.c_double * _NCLASS), ("kernel_launches", C.c_uint64 * _NCLASS),
                ("kernel_bytes", C.c_double * _NCLASS),
                ("roots_done", C.c_uint64), ("last_root", C.c_uint32), ("work_capped", C.c_uint32),
                ("work", C.c_uint64), ("merge_bytes", C.c_double * _NCLASS)]


class _DGInfo(C.Structure):
    _fields_ = [("V", C.c_uint32), ("arcs", C.c_uint64), ("symmetric", C.c_int), ("self_loops", C.c_uint64),
                ("hubs", C.c_uint32), ("hub_min_deg", C.c_uint32), ("two_hop_entries", C.c_uint64),
                ("two_hop_dense_rows", C.c_uint32), ("arc_common", C.c_int), ("two_hop_build_s", C.c_double),
                ("arc_common_build_s", C.c_double), ("two_hop_on_device", C.c_int), ("hub_rows", C.c_uint32),
                ("hub_rows_build_s", C.c_double)]


class _BatchOpts(C.Structure):
    _fields_ = [("flags", C.c_uint32), ("blocks_per_cu", C.c_uint32), ("vchunk", C.c_uint32),
                ("no_graph", C.c_uint32), ("queue_cap", C.c_uint64), ("arena_bytes", C.c_uint64),
                ("ticket_queue", C.c_uint32), ("item_work", C.c_uint32), ("roots", C.POINTER(C.c_uint32)),
                ("persistent", C.c_uint32), ("barrier_timeout_ms", C.c_uint32)]


class _BatchResult(C.Structure):
    _fields_ = [("count", C.c_uint64), ("level_sizes", C.c_uint64 * 16), ("latency_ms", C.c_double)]


_lib = None


def lib():
    """Loads libcuts_amd.so (once).  Fails loudly if it is missing."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is not built; run __graft_entry__.build() or make -C cuts_amd")
        L = C.CDLL(LIB_PATH)
        vp = C.c_void_p
        pp = C.POINTER(C.c_void_p)
        sig = {
            "cuts_last_error": (C.c_char_p, []),
            "cuts_version": (C.c_char_p, []),
            "cuts_graph_load": (C.c_int, [C.c_char_p, pp]),
            "cuts_graph_from_arcs": (C.c_int, [C.c_uint32, _u32p, _u32p, C.c_uint64, C.c_int, pp]),
            "cuts_graph_save": (C.c_int, [vp, C.c_char_p]),
            "cuts_graph_view_of": (C.c_int, [vp, C.POINTER(_GraphView)]),
            "cuts_graph_two_hop": (C.c_int, [vp, C.POINTER(C.c_uint64), _u32p, _u32p, C.c_uint64,
                                             C.POINTER(C.c_uint64)]),
            "cuts_graph_free": (None, [vp]),
            "cuts_gen_chung_lu": (C.c_int, [C.c_uint32, C.c_uint64, C.c_double, C.c_uint32, C.c_uint64, pp]),
            "cuts_gen_rmat": (C.c_int, [C.c_uint32, C.c_uint32, C.c_double, C.c_double, C.c_double,
                                        C.c_uint64, pp]),
            "cuts_find_pattern": (C.c_int, [vp, C.c_uint32, _u32p, C.c_uint32, C.c_uint64, _u32p]),
            "cuts_walk_query": (C.c_int, [vp, C.c_uint32, C.c_uint32, C.c_uint64, _u32p, _u32p, _u32p]),
            "cuts_plan_create": (C.c_int, [vp, _u32p, pp]),
            "cuts_plan_info": (C.c_int, [vp, _u32p, _u32p, _u32p, _u32p, _u32p, _u32p]),
            "cuts_plan_free": (None, [vp]),
            "cuts_context_create": (C.c_int, [C.c_int, pp]),
            "cuts_context_destroy": (None, [vp]),
            "cuts_device_graph_upload": (C.c_int, [vp, vp, pp]),
            "cuts_device_graph_free": (None, [vp]),
            "cuts_device_graph_query": (C.c_int, [vp, C.POINTER(_DGInfo)]),
            "cuts_device_graph_arc_common": (C.c_int, [vp, _u32p, C.c_uint64]),
            "cuts_device_graph_two_hop": (C.c_int, [vp, C.POINTER(C.c_uint64), _u32p, _u32p, C.c_uint64,
                                                    C.POINTER(C.c_int32), C.POINTER(C.c_uint16), C.c_uint64]),
            "cuts_device_graph_hub_rows": (C.c_int, [vp, C.POINTER(C.c_int32), C.POINTER(C.c_uint16), C.c_uint64]),
            "cuts_match_options_default": (None, [C.POINTER(_Opts)]),
            "cuts_match": (C.c_int, [vp, vp, vp, C.POINTER(_Opts), C.POINTER(_Result)]),
            "cuts_match_noalloc": (C.c_int, [vp, vp, vp, C.POINTER(_Opts), C.POINTER(_Result)]),
            "cuts_match_many": (C.c_int, [vp, vp, C.POINTER(C.c_void_p), C.c_uint32, C.POINTER(_Opts), C.c_uint32,
                                          C.c_uint32, C.POINTER(_Result), C.POINTER(C.c_double)]),
            "cuts_noalloc_scratch_bytes": (C.c_uint64, []),
            "cuts_host_alloc": (C.c_int, [C.c_uint64, C.POINTER(C.c_void_p)]),
            "cuts_host_free": (None, [C.c_void_p]),
            "cuts_search": (C.c_int, [C.c_char_p, C.c_char_p, C.c_int, C.POINTER(C.c_uint64)]),
            "cuts_search_dfs_bfs_strategy": (C.c_int, [C.c_char_p, C.c_char_p, C.c_int, C.c_uint32,
                                                       C.POINTER(C.c_uint64)]),
            "cuts_write_matches": (C.c_int, [C.c_char_p, vp, _u32p, C.c_uint64]),
            "cuts_copy_probe": (C.c_int, [vp, C.c_uint64, C.c_uint32, C.POINTER(C.c_double)]),
            "cuts_log_tables": (C.c_int, [C.c_char_p, C.c_char_p, vp, C.POINTER(_Result)]),
            "cuts_batch_options_default": (None, [C.POINTER(_BatchOpts)]),
            "cuts_batch_create": (C.c_int, [vp, vp, C.POINTER(C.c_void_p), C.c_uint32, C.POINTER(_BatchOpts), pp]),
            "cuts_batch_run": (C.c_int, [vp, C.POINTER(_BatchResult), C.POINTER(C.c_double),
                                         C.POINTER(C.c_uint64)]),
            "cuts_batch_free": (None, [vp]),
        }
        for name, (res, args) in sig.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


class CutsError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"cuts error {code}: {msg}")
        self.code = code


def _check(rc: int):
    if rc != 0:
        raise CutsError(rc, lib().cuts_last_error().decode())


class _PinnedBlock:
    """One cuts_host_alloc allocation, freed with its last reference."""

    def __init__(self, nbytes: int):
        p = C.c_void_p()
        _check(lib().cuts_host_alloc(nbytes, C.byref(p)))
        self.ptr = p

    def __del__(self):
        if getattr(self, "ptr", None) is not None and self.ptr.value and _lib is not None:
            _lib.cuts_host_free(self.ptr)
            self.ptr = None


def pinned_array(shape, dtype=np.uint32) -> np.ndarray:
    """A numpy array in pinned host memory (cuts_host_alloc); the memory is
    released when the array and its views are gone."""
    n = int(np.prod(shape)) * np.dtype(dtype).itemsize
    block = _PinnedBlock(max(n, 1))
    buf = (C.c_char * max(n, 1)).from_address(block.ptr.value)
    buf._block = block
    return np.frombuffer(buf, dtype=dtype, count=int(np.prod(shape))).reshape(shape)


def _u32(a) -> np.ndarray:
    return np.ascontiguousarray(np.asarray(a, dtype=np.uint32))


def _ptr(a: np.ndarray):
    return a.ctypes.data_as(_u32p)


class Graph:
    """Host CSR graph (``Graph`` of src/graph.cpp:18-61)."""

    def __init__(self, handle):
        self._h = C.c_void_p(handle)

    def __del__(self):
        if getattr(self, "_h", None) and self._h.value and _lib is not None:
            _lib.cuts_graph_free(self._h)
            self._h = None

    @classmethod
    def load(cls, path: str) -> "Graph":
        h = C.c_void_p()
        _check(lib().cuts_graph_load(path.encode(), C.byref(h)))
        return cls(h.value)

    @classmethod
    def from_arcs(cls, V: int, src, dst, symmetrize: bool = False) -> "Graph":
        s, d = _u32(src), _u32(dst)
        h = C.c_void_p()
        _check(lib().cuts_graph_from_arcs(V, _ptr(s), _ptr(d), len(s), int(symmetrize), C.byref(h)))
        return cls(h.value)

    @classmethod
    def from_edges(cls, V: int, edges: Sequence[Sequence[int]]) -> "Graph":
        """Undirected graph: every edge becomes both arcs."""
        e = np.asarray(edges, dtype=np.uint32).reshape(-1, 2)
        return cls.from_arcs(V, e[:, 0], e[:, 1], symmetrize=True)

    @classmethod
    def chung_lu(cls, n, m, gamma, max_deg=0, seed=0) -> "Graph":
        h = C.c_void_p()
        _check(lib().cuts_gen_chung_lu(n, m, gamma, max_deg, seed, C.byref(h)))
        return cls(h.value)

    @classmethod
    def rmat(cls, scale, edge_factor=16, a=0.57, b=0.19, c=0.19, seed=0) -> "Graph":
        h = C.c_void_p()
        _check(lib().cuts_gen_rmat(scale, edge_factor, a, b, c, seed, C.byref(h)))
        return cls(h.value)

    def save(self, path: str):
        _check(lib().cuts_graph_save(self._h, path.encode()))

    def view(self):
        v = _GraphView()
        _check(lib().cuts_graph_view_of(self._h, C.byref(v)))
        n, a = v.V, v.arcs

        def arr(p, cnt):
            if cnt == 0:
                return np.zeros(0, dtype=np.uint32)
            return np.ctypeslib.as_array(p, shape=(cnt,)).copy()
        return dict(V=n, arcs=a, symmetric=bool(v.symmetric), out_off=arr(v.out_off, n + 1),
                    out_adj=arr(v.out_adj, a), in_off=arr(v.in_off, n + 1), in_adj=arr(v.in_adj, a),
                    sig=arr(v.sig, 2 * n))

    def two_hop(self):
        """The two-hop index a device upload builds (symmetric graphs):
        (off[V+1], v, c) -- for every vertex u, the vertices v != u sharing a
        neighbour with u, ascending, and the number of shared neighbours."""
        n = C.c_uint64()
        _check(lib().cuts_graph_two_hop(self._h, None, None, None, 0, C.byref(n)))
        V = self.V
        off = np.zeros(V + 1, dtype=np.uint64)
        v = np.zeros(max(n.value, 1), dtype=np.uint32)
        c = np.zeros(max(n.value, 1), dtype=np.uint32)
        _check(lib().cuts_graph_two_hop(self._h, off.ctypes.data_as(C.POINTER(C.c_uint64)), _ptr(v), _ptr(c),
                                        n.value, C.byref(n)))
        return off, v[: n.value], c[: n.value]

    @property
    def V(self) -> int:
        v = _GraphView()
        _check(lib().cuts_graph_view_of(self._h, C.byref(v)))
        return v.V

    @property
    def arcs(self) -> int:
        v = _GraphView()
        _check(lib().cuts_graph_view_of(self._h, C.byref(v)))
        return v.arcs

    def find_pattern(self, k: int, edges, seed: int = 0) -> np.ndarray:
        e = _u32(edges).reshape(-1)
        w = np.zeros(k, dtype=np.uint32)
        _check(lib().cuts_find_pattern(self._h, k, _ptr(e), len(e) // 2, seed, _ptr(w)))
        return w

    def walk_query(self, k: int, n_edges: int, seed: int):
        cap = k * (k - 1) // 2
        e = np.zeros(2 * cap, dtype=np.uint32)
        n = C.c_uint32()
        w = np.zeros(k, dtype=np.uint32)
        _check(lib().cuts_walk_query(self._h, k, n_edges, seed, _ptr(e), C.byref(n), _ptr(w)))
        return e[: 2 * n.value].reshape(-1, 2).copy(), w


class Plan:
    """Query + matching order (``Graph::sort_search_order``, src/graph.cpp:62-134)."""

    def __init__(self, query: Graph, order: Optional[Sequence[int]] = None):
        h = C.c_void_p()
        o = None if order is None else _u32(order)
        _check(lib().cuts_plan_create(query._h, None if o is None else _ptr(o), C.byref(h)))
        self._h = h
        self.query = query
        k = C.c_uint32()
        _check(lib().cuts_plan_info(self._h, C.byref(k), None, None, None, None, None))
        self.k = k.value
        order_a = np.zeros(self.k, np.uint32)
        po = np.zeros(self.k + 1, np.uint32)
        co = np.zeros(self.k + 1, np.uint32)
        _check(lib().cuts_plan_info(self._h, None, _ptr(order_a), _ptr(po), None, _ptr(co), None))
        par = np.zeros(max(int(po[-1]), 1), np.uint32)
        chi = np.zeros(max(int(co[-1]), 1), np.uint32)
        _check(lib().cuts_plan_info(self._h, None, None, None, _ptr(par), None, _ptr(chi)))
        self.order = order_a.tolist()
        self.parents_offset = po.tolist()
        self.children_offset = co.tolist()
        self.parents = par[: po[-1]].tolist()
        self.children = chi[: co[-1]].tolist()

    def __del__(self):
        if getattr(self, "_h", None) and self._h.value and _lib is not None:
            _lib.cuts_plan_free(self._h)
            self._h = None


@dataclass
class MatchResult:
    count: int
    level_sizes: List[int]
    device_ms: float
    chunks: int
    rows: Optional[np.ndarray] = None  # (n, k) data vertices per matching-order position
    kernels: Optional[dict] = None     # FLAG_PROFILE: class -> (ms, launches, bytes)
    complete: bool = True              # False if time_limit_ms stopped it between trunks
    roots_done: int = 0                # level-0 vertices whose subtrees were counted
    last_root: int = 0                 # the last of them (ascending ids)
    work_capped: bool = False          # max_work stopped the run; count invalid
    work: int = 0                      # candidates examined (sum of scan totals)
    merge_bytes: Optional[dict] = None # FLAG_PROFILE: class -> merge-model algorithmic bytes


class DeviceGraph:
    def __init__(self, ctx: "Context", handle):
        self.ctx = ctx
        self._h = C.c_void_p(handle)

    def __del__(self):
        if getattr(self, "_h", None) and self._h.value and _lib is not None:
            _lib.cuts_device_graph_free(self._h)
            self._h = None

    def info(self) -> dict:
        """What the upload built: hubs, two-hop index, per-arc table, build times."""
        i = _DGInfo()
        _check(lib().cuts_device_graph_query(self._h, C.byref(i)))
        return {name: getattr(i, name) for name, _ in _DGInfo._fields_}

    def arc_common(self) -> np.ndarray:
        """Per-arc |N(u) & N(w)|, aligned with the graph's out-lists."""
        n = self.info()["arcs"]
        out = np.zeros(max(n, 1), dtype=np.uint32)
        _check(lib().cuts_device_graph_arc_common(self._h, _ptr(out), n))
        return out[:n]

    def two_hop(self, dense: bool = False):
        """The two-hop index the upload built, read back: (off, v, c) as
        Graph.two_hop() returns them, plus (dense_id, dense rows as a
        rows x V uint16 array) when dense is true."""
        i = self.info()
        n, V, rows = i["two_hop_entries"], i["V"], i["two_hop_dense_rows"]
        if n == 0:
            raise CutsError(-1, "the device graph has no two-hop index")
        off = np.zeros(V + 1, dtype=np.uint64)
        v = np.zeros(max(n, 1), dtype=np.uint32)
        c = np.zeros(max(n, 1), dtype=np.uint32)
        did = np.zeros(max(V, 1), dtype=np.int32) if dense else None
        rowbuf = np.zeros(max(rows * V, 1), dtype=np.uint16) if dense else None
        _check(lib().cuts_device_graph_two_hop(
            self._h, off.ctypes.data_as(C.POINTER(C.c_uint64)), _ptr(v), _ptr(c), n,
            did.ctypes.data_as(C.POINTER(C.c_int32)) if dense else None,
            rowbuf.ctypes.data_as(C.POINTER(C.c_uint16)) if dense else None, rows * V if dense else 0))
        if not dense:
            return off, v[:n], c[:n]
        return off, v[:n], c[:n], did[:V], rowbuf[:rows * V].reshape(rows, V)

    def hub_rows(self):
        """The hub rows the upload built on a graph too large for the
        two-hop index: (row_id, rows x V uint16 array of |N(hub) & N(v)|)."""
        i = self.info()
        H, V = i["hub_rows"], i["V"]
        if H == 0:
            raise CutsError(-1, "the device graph has no hub rows")
        rid = np.zeros(V, dtype=np.int32)
        rows = np.zeros(H * V, dtype=np.uint16)
        _check(lib().cuts_device_graph_hub_rows(self._h, rid.ctypes.data_as(C.POINTER(C.c_int32)),
                                                rows.ctypes.data_as(C.POINTER(C.c_uint16)), H * V))
        return rid, rows.reshape(H, V)


class Context:
    """One GPU: stream, events and the trie-table arena."""

    def __init__(self, device: int = 0):
        h = C.c_void_p()
        _check(lib().cuts_context_create(device, C.byref(h)))
        self._h = h

    def __del__(self):
        if getattr(self, "_h", None) and self._h.value and _lib is not None:
            _lib.cuts_context_destroy(self._h)
            self._h = None

    def match_many(self, dg: "DeviceGraph", plans, roots=None, flags: int = 0, workers: int = 4, max_work: int = 0,
                   blocks_per_cu: int = 0, expand_threads: int = 0, class_blocks_per_cu=None):
        return match_many(self, dg, plans, roots=roots, flags=flags, workers=workers, max_work=max_work,
                          blocks_per_cu=blocks_per_cu, expand_threads=expand_threads,
                          class_blocks_per_cu=class_blocks_per_cu)

    def copy_probe(self, nbytes: int, iters: int = 10) -> float:
        """Read+write GB/s of a plain copy of nbytes (HBM yardstick)."""
        gbps = C.c_double()
        _check(lib().cuts_copy_probe(self._h, nbytes, iters, C.byref(gbps)))
        return gbps.value

    def upload(self, g: Graph) -> DeviceGraph:
        h = C.c_void_p()
        _check(lib().cuts_device_graph_upload(self._h, g._h, C.byref(h)))
        return DeviceGraph(self, h.value)

    def match(self, dg: DeviceGraph, plan: Plan, trunks: int = 0, flags: int = 0, budget: int = 0,
              roots: Optional[tuple] = None, write: bool = False, emb_cap: int = 0,
              blocks_per_cu: int = 0, time_limit_ms: float = 0.0, max_work: int = 0,
              noalloc: bool = False, rows_out: Optional[np.ndarray] = None,
              expand_threads: int = 0, class_blocks_per_cu: Optional[Sequence[int]] = None) -> MatchResult:
        """One query (cuts_match).  noalloc=True calls cuts_match_noalloc: the
        run's buffers come from the device's static scratch array.  With
        write=True the rows go to rows_out (C-contiguous uint32, (cap, k);
        pinned_array for overlapped copies) or to a new pinned array of
        emb_cap rows."""
        o = _Opts()
        lib().cuts_match_options_default(C.byref(o))
        o.trunks, o.flags, o.budget_bytes, o.blocks_per_cu = trunks, flags, budget, blocks_per_cu
        o.time_limit_ms = time_limit_ms
        o.max_work = max_work
        o.expand_threads = expand_threads
        _set_class_bpc(o, class_blocks_per_cu)
        if roots is not None:
            o.root_begin, o.root_end, o.root_stride = roots
        rows = None
        if write:
            o.flags |= FLAG_WRITE
            if rows_out is not None:
                if rows_out.dtype != np.uint32 or not rows_out.flags.c_contiguous or rows_out.ndim != 2 or \
                        rows_out.shape[1] != plan.k:
                    raise ValueError("rows_out must be a C-contiguous uint32 array of shape (cap, k)")
                rows = rows_out
                emb_cap = rows_out.shape[0]
            else:
                # pinned, so that the copies of one chunk overlap the next chunk
                rows = pinned_array((max(emb_cap, 1), plan.k))
            o.emb_rows = _ptr(rows)
            o.emb_cap = emb_cap
        r = _Result()
        entry = lib().cuts_match_noalloc if noalloc else lib().cuts_match
        _check(entry(self._h, dg._h, plan._h, C.byref(o), C.byref(r)))
        res = MatchResult(int(r.count), [int(x) for x in r.level_sizes[: plan.k]], float(r.device_ms),
                          int(r.chunks), complete=bool(r.complete), roots_done=int(r.roots_done),
                          last_root=int(r.last_root), work_capped=bool(r.work_capped), work=int(r.work))
        res._raw = r
        if write:
            # a view of the pinned buffer when it is mostly used, else a copy
            res.rows = rows[: r.emb_written] if (rows_out is not None or 2 * r.emb_written >= len(rows)) \
                else rows[: r.emb_written].copy()
        if o.flags & FLAG_PROFILE:
            res.kernels = {KERNEL_CLASSES[i]: (float(r.kernel_ms[i]), int(r.kernel_launches[i]),
                                               float(r.kernel_bytes[i])) for i in range(_NCLASS)}
            res.merge_bytes = {KERNEL_CLASSES[i]: float(r.merge_bytes[i]) for i in range(_NCLASS)}
        return res


def _result_of(r, k: int) -> "MatchResult":
    res = MatchResult(int(r.count), [int(x) for x in r.level_sizes[:k]], float(r.device_ms), int(r.chunks),
                      complete=bool(r.complete), roots_done=int(r.roots_done), last_root=int(r.last_root),
                      work_capped=bool(r.work_capped), work=int(r.work))
    res._raw = r
    return res


def _set_class_bpc(o, table):
    """class_blocks_per_cu: None, a sequence indexed by kernel class, or a
    {class name: blocks} dict (names as in MatchResult.kernels)."""
    if not table:
        return
    if isinstance(table, dict):
        for name, b in table.items():
            o.class_blocks_per_cu[KERNEL_CLASSES.index(name)] = int(b)
    else:
        for i, b in enumerate(table):
            o.class_blocks_per_cu[i] = int(b)


def match_many(ctx: "Context", dg: "DeviceGraph", plans: Sequence["Plan"], roots=None, flags: int = 0,
               workers: int = 4, max_work: int = 0, blocks_per_cu: int = 0, expand_threads: int = 0,
               class_blocks_per_cu=None):
    """Many queries against one resident graph on `workers` concurrent streams
    (cuts_match_many).  roots: None or one (first, end, stride) per query.
    Returns (results in query order, host wall ms of the whole set)."""
    n = len(plans)
    opts = (_Opts * max(n, 1))()
    for i in range(n):
        lib().cuts_match_options_default(C.byref(opts[i]))
        opts[i].flags = flags
        opts[i].max_work = max_work
        opts[i].blocks_per_cu = blocks_per_cu
        opts[i].expand_threads = expand_threads
        _set_class_bpc(opts[i], class_blocks_per_cu)
        if roots is not None and roots[i] is not None:
            opts[i].root_begin, opts[i].root_end, opts[i].root_stride = roots[i]
    res = (_Result * max(n, 1))()
    arr = (C.c_void_p * max(n, 1))(*[p._h.value for p in plans])
    wall = C.c_double()
    _check(lib().cuts_match_many(ctx._h, dg._h, arr, n, opts, n, workers, res, C.byref(wall)))
    return [_result_of(res[i], plans[i].k) for i in range(n)], float(wall.value)


@dataclass
class BatchResult:
    counts: List[int]
    level_sizes: List[List[int]]
    latency_ms: List[float]
    device_ms: float
    arena_bytes: int


class Batch:
    """Many queries against one resident data graph: a level-synchronous
    device work queue of (query, level, chunk) items, the launch sequence
    captured once in a hipGraph.  ``queue="level"``: one persistent-grid
    launch per trie level.  ``queue="persistent"``: one cooperative launch
    runs every level, its blocks meeting at a device-wide barrier between
    levels (a wait longer than ``barrier_timeout_ms`` aborts the run).
    ``queue="ticket"`` selects the asynchronous ticket-queue kernel, which is
    disabled (creation fails)."""

    def __init__(self, ctx: Context, dg: DeviceGraph, plans: Sequence[Plan], flags: int = 0,
                 blocks_per_cu: int = 0, vchunk: int = 0, use_graph: bool = True, queue_cap: int = 0,
                 arena_bytes: int = 0, item_work: int = 0, queue: str = "level",
                 roots: Optional[Sequence[Tuple[int, int, int]]] = None, barrier_timeout_ms: int = 0):
        """roots: per query (first, end, stride) of the data vertices that
        order position 0 may take (as Context.match's roots); None = all."""
        if queue not in ("level", "persistent", "ticket"):
            raise ValueError("queue must be 'level', 'persistent' or 'ticket'")
        self.ctx, self.dg, self.plans = ctx, dg, list(plans)
        o = _BatchOpts()
        lib().cuts_batch_options_default(C.byref(o))
        o.flags, o.blocks_per_cu, o.vchunk, o.no_graph = flags, blocks_per_cu, vchunk, 0 if use_graph else 1
        o.queue_cap, o.arena_bytes, o.item_work = queue_cap, arena_bytes, item_work
        o.ticket_queue = 1 if queue == "ticket" else 0
        o.persistent = 1 if queue == "persistent" else 0
        o.barrier_timeout_ms = barrier_timeout_ms
        if roots is not None:
            if len(roots) != len(self.plans):
                raise ValueError("roots: one (first, end, stride) per query")
            self._roots = np.ascontiguousarray(np.asarray(roots, dtype=np.uint32).reshape(-1, 3))
            o.roots = self._roots.ctypes.data_as(C.POINTER(C.c_uint32))
        arr = (C.c_void_p * len(self.plans))(*[p._h.value for p in self.plans])
        h = C.c_void_p()
        _check(lib().cuts_batch_create(ctx._h, dg._h, arr, len(self.plans), C.byref(o), C.byref(h)))
        self._h = h

    def __del__(self):
        if getattr(self, "_h", None) and self._h.value and _lib is not None:
            _lib.cuts_batch_free(self._h)
            self._h = None

    def run(self) -> BatchResult:
        n = len(self.plans)
        res = (_BatchResult * n)()
        ms = C.c_double()
        ab = C.c_uint64()
        _check(lib().cuts_batch_run(self._h, res, C.byref(ms), C.byref(ab)))
        return BatchResult([int(r.count) for r in res],
                           [[int(x) for x in r.level_sizes[: p.k]] for r, p in zip(res, self.plans)],
                           [float(r.latency_ms) for r in res], float(ms.value), int(ab.value))


def noalloc_scratch_bytes() -> int:
    """Size of the static scratch array behind Context.match(noalloc=True)."""
    return int(lib().cuts_noalloc_scratch_bytes())


def search(query_file: str, data_file: str, write_to_disk: bool = False) -> int:
    c = C.c_uint64()
    _check(lib().cuts_search(query_file.encode(), data_file.encode(), int(write_to_disk), C.byref(c)))
    return c.value


def search_dfs_bfs_strategy(query_file: str, data_file: str, write_to_disk: bool, trunks: int) -> int:
    c = C.c_uint64()
    _check(lib().cuts_search_dfs_bfs_strategy(query_file.encode(), data_file.encode(), int(write_to_disk),
                                              trunks, C.byref(c)))
    return c.value


def log_tables(path: str, label: str, plan: Plan, result: MatchResult):
    """Appends the run's per-level table sizes as one JSON line (the solver's
    own table log; see cuts_log_tables in include/cuts.h)."""
    _check(lib().cuts_log_tables(path.encode(), label.encode(), plan._h, C.byref(result._raw)))


def write_matches(path: str, plan: Plan, rows: np.ndarray):
    r = _u32(rows)
    _check(lib().cuts_write_matches(path.encode(), plan._h, _ptr(r), r.shape[0] if r.ndim == 2 else 0))
