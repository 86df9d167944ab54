"""Python binding of the C ABI in include/pac.h (ctypes over libpac.so).

Mirrors the reference's C++ interface of Byy-0024/PartitionAndConquer for the
intersection path: ``Graph`` (src/Graph.hpp:38-103) with
``solve_triangle_counting_{merge,galloping}``, ``solve_lcc_{merge,galloping}``,
``solve_jsim_{merge,galloping}``, and ``PartitionedGraph`` (Partitioned_Graph,
src/Graph.hpp:105-155).  The GPU-native calls (``tc``, ``lcc``, ``jaccard_edges``,
``tc_batch``) expose the dispatch and launch options.

The library is required: importing works without it, but every call raises
``PacError`` when libpac.so is missing -- there is no CPU fallback.
"""
from __future__ import annotations

import ctypes as C
import os
import weakref
from dataclasses import dataclass

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libpac.so")

PAC_OK = 0
ERRORS = {-1: "ARG", -2: "HIP", -3: "OOM", -4: "FORMAT", -5: "IO", -6: "STATE",
          -7: "UNSUPPORTED", -8: "TIMEOUT"}

DISPATCH_FIXED, DISPATCH_LEARNED, FORCE_MERGE, FORCE_GALLOP, FORCE_BITMAP = range(5)
LAUNCH_PLAIN, LAUNCH_PERSISTENT, LAUNCH_GRAPH = range(3)
MODE_TC, MODE_LCC, MODE_JACCARD = range(3)
PACKED_OFF, PACKED_AUTO, PACKED_ALWAYS = range(3)
STRATEGIES = ("merge", "gallop", "bitmap")
NUM_TILES = 4
MLP_IN, MLP_HIDDEN, MLP_OUT = 16, 64, 3
MLP_NPARAMS = MLP_HIDDEN * MLP_IN + MLP_HIDDEN + MLP_HIDDEN * MLP_HIDDEN + MLP_HIDDEN + \
    MLP_OUT * MLP_HIDDEN + MLP_OUT


class PacError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"PAC_ERR_{ERRORS.get(code, code)} ({code}): {msg}")
        self.code = code


class GraphInfo(C.Structure):
    _fields_ = [("num_vertices", C.c_uint64), ("num_entries", C.c_uint64),
                ("num_edges", C.c_uint64), ("max_degree", C.c_uint64),
                ("max_out_degree", C.c_uint64), ("work_W", C.c_uint64),
                ("num_partitions", C.c_uint64), ("num_items", C.c_uint64),
                ("prepare_ms", C.c_double), ("num_pairs", C.c_uint64),
                ("num_rows", C.c_uint64), ("streamed_elems", C.c_uint64),
                ("staged_elems", C.c_uint64), ("staged_item_elems", C.c_uint64),
                ("stream_elem_bytes", C.c_uint64)]

    def algorithmic_bytes(self, mode: int = 1) -> int:
        """Bytes the hash path's intersection launch must move in `mode` (0 TC, 1 LCC,
        2 Jaccard): every pair's stream, the suffix of its partner's out-list ranked above
        its owner (stream_elem_bytes per element in TC and LCC:
        3 with packed lists, else 4; 4 in Jaccard) with the list's bounds stored
        beside the pair (8 B), every row's staged list once per work item holding it, the
        row's list and pair bounds and its row-list entry (20 B), the items (32 B).  LCC adds
        each pair's partner ID (4 B) and the per-vertex counts read and written once (16 B
        per vertex); Jaccard adds the partner ID, edge index and the pair's first hit-word
        index (16 B per pair), one hit bit per streamed element written, and the per-edge
        supports read and written once (8 B per edge)."""
        es = 4 if mode == 2 else self.stream_elem_bytes  # per-edge supports: 4-byte lists
        b = (es * self.streamed_elems + 4 * self.staged_item_elems
             + 8 * self.num_pairs
             + 20 * self.num_rows + 32 * self.num_items)
        if mode == 1:
            b += 4 * self.num_pairs + 16 * self.num_vertices
        elif mode == 2:
            b += 16 * self.num_pairs + self.streamed_elems // 8 + 8 * self.num_edges
        return int(b)


class Timing(C.Structure):
    _fields_ = [("total_ms", C.c_double), ("isect_ms", C.c_double),
                ("epilogue_ms", C.c_double), ("kernel_ms", C.c_double),
                ("launches", C.c_uint64), ("strategy_partitions", C.c_uint64 * 3)]


_lib = None
vp = C.c_void_p


def lib():
    """Load libpac.so (raises PacError if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise PacError(-7, f"{LIB_PATH} not built: run python -m partitionandconquer_amd.build")
    L = C.CDLL(LIB_PATH)
    sig = {
        "pac_last_error": ([], C.c_char_p),
        "pac_ctx_create": ([C.c_int, C.POINTER(vp)], C.c_int),
        "pac_ctx_destroy": ([vp], None),
        "pac_ctx_last_timing": ([vp, C.POINTER(Timing)], C.c_int),
        "pac_ctx_num_cus": ([vp], C.c_int),
        "pac_stream_copy": ([vp, C.c_uint64, C.c_int, C.POINTER(C.c_double)], C.c_int),
        "pac_graph_upload": ([vp, vp, C.c_uint64, vp, C.c_uint64, C.POINTER(vp)], C.c_int),
        "pac_graph_load_csr_dir": ([vp, C.c_char_p, C.POINTER(vp)], C.c_int),
        "pac_graph_load_text": ([vp, C.c_char_p, C.POINTER(vp)], C.c_int),
        "pac_graph_save_csr_dir": ([vp, C.c_char_p], C.c_int),
        "pac_graph_from_edges": ([vp, vp, vp, C.c_uint64, C.c_uint64, C.POINTER(vp)], C.c_int),
        "pac_graph_generate_rmat": ([vp, C.c_uint32, C.c_uint32, C.c_double, C.c_double,
                                     C.c_double, C.c_uint64, C.POINTER(vp)], C.c_int),
        "pac_graph_generate_er": ([vp, C.c_uint64, C.c_double, C.c_uint64, C.POINTER(vp)], C.c_int),
        "pac_graph_download_csr": ([vp, vp, vp], C.c_int),
        "pac_graph_get_info": ([vp, C.POINTER(GraphInfo)], C.c_int),
        "pac_graph_validate": ([vp, C.c_int], C.c_int),
        "pac_graph_free": ([vp], None),
        "pac_graph_prepare": ([vp, C.c_uint32, vp], C.c_int),
        "pac_graph_partition_sizes": ([vp, vp], C.c_int),
        "pac_random_partition": ([C.c_uint64, C.c_uint32, C.c_uint64, vp], C.c_int),
        "pac_greedy_partition": ([vp, C.c_uint32, C.c_uint64, vp], C.c_int),
        "pac_greedy_partition_2": ([vp, C.c_uint32, C.c_uint64, vp], C.c_int),
        "pac_evaluate_partition": ([vp, vp, C.c_uint32, C.POINTER(C.c_double)], C.c_int),
        "pac_tc": ([vp, C.c_int, C.c_int, vp, vp], C.c_int),
        "pac_tc_noalloc": ([vp, C.c_int, vp, vp], C.c_int),
        "pac_ctx_set_timeout": ([vp, C.c_double], C.c_int),
        "pac_ctx_set_packed_lists": ([vp, C.c_int], C.c_int),
        "pac_lcc": ([vp, C.c_int, C.c_int, vp, vp, vp], C.c_int),
        "pac_jaccard_edges": ([vp, C.c_int, C.c_int, vp, vp, vp, vp, vp], C.c_int),
        "pac_jaccard_queries": ([vp, vp, C.c_uint32, vp], C.c_int),
        "pac_cnt_common_pairs": ([vp, vp, vp, C.c_uint64, vp], C.c_int),
        "pac_cnt_common_kway": ([vp, vp, vp, C.c_uint64, vp], C.c_int),
        "pac_cnt_common_pairs_filtered": ([vp, vp, vp, vp, C.c_uint64, C.c_int32, vp], C.c_int),
        "pac_cnt_common_kway_filtered": ([vp, vp, vp, C.c_uint64, C.c_int32, vp], C.c_int),
        "pac_common_neis": ([vp, C.c_int32, C.c_int32, vp, C.c_uint64, C.c_int32, vp,
                             C.c_uint64, C.POINTER(C.c_uint64)], C.c_int),
        "pac_friend_triangle": ([vp, C.c_int32, vp, C.c_uint64, vp], C.c_int),
        "pac_square_count": ([vp, vp], C.c_int),
        "pac_tc_batch": ([vp, vp, C.c_uint32, C.c_int, C.c_int, vp], C.c_int),
        "pac_autotune": ([vp, vp, C.c_uint32, C.c_int, C.c_int, C.c_char_p, C.POINTER(C.c_int),
                          vp], C.c_int),
        "pac_set_tile": ([vp, C.c_int], C.c_int),
        "pac_get_tile": ([vp, C.POINTER(C.c_int)], C.c_int),
        "pac_partition_features": ([vp, vp], C.c_int),
        "pac_profile_partitions": ([vp, C.c_int, vp], C.c_int),
        "pac_fixed_choice": ([vp, vp], C.c_int),
        "pac_mlp_choice": ([vp, vp], C.c_int),
        "pac_mlp_set_params": ([vp, vp], C.c_int),
        "pac_mlp_get_params": ([vp, vp], C.c_int),
        "pac_mlp_init": ([vp, C.c_uint64], C.c_int),
        "pac_mlp_forward": ([vp, vp, C.c_uint64, vp], C.c_int),
        "pac_mlp_loss_grad": ([vp, vp, vp, vp, C.c_uint64, C.POINTER(C.c_float), vp], C.c_int),
        "pac_mlp_train": ([vp, vp, vp, vp, C.c_uint64, C.c_uint32, C.c_float,
                           C.POINTER(C.c_float)], C.c_int),
        "pac_mlp_save": ([vp, C.c_char_p], C.c_int),
        "pac_mlp_load": ([vp, C.c_char_p], C.c_int),
    }
    for name, (args, res) in sig.items():
        f = getattr(L, name)
        f.argtypes = args
        f.restype = res
    _lib = L
    return L


def _check(rc: int):
    if rc != PAC_OK:
        raise PacError(rc, lib().pac_last_error().decode(errors="replace"))


def _ptr(a):
    return None if a is None else a.ctypes.data_as(vp)


@dataclass
class LastTiming:
    total_ms: float
    isect_ms: float
    epilogue_ms: float
    kernel_ms: float
    launches: int
    strategy_partitions: tuple


class Context:
    """One HIP device + stream (pac_ctx)."""

    def __init__(self, device: int = 0):
        h = vp()
        _check(lib().pac_ctx_create(device, C.byref(h)))
        self._h = h
        self.device = device
        self._graphs = weakref.WeakSet()

    def close(self):
        """Destroy the context; graphs still alive on it are freed with it."""
        if getattr(self, "_h", None):
            for g in list(self._graphs):
                g._h = None
            lib().pac_ctx_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    @property
    def num_cus(self) -> int:
        return lib().pac_ctx_num_cus(self._h)

    def last_timing(self) -> LastTiming:
        t = Timing()
        _check(lib().pac_ctx_last_timing(self._h, C.byref(t)))
        return LastTiming(t.total_ms, t.isect_ms, t.epilogue_ms, t.kernel_ms, t.launches,
                          tuple(t.strategy_partitions))

    def set_timeout(self, ms: float):
        """Bound (ms, 0 = none) on waiting for a solve; PacError -8 when it passes."""
        _check(lib().pac_ctx_set_timeout(self._h, ms))

    def set_packed_lists(self, mode: int = 1):
        """When TC and LCC launches stream the 3-byte lists of graphs with every ID below
        2^24: PACKED_AUTO (the default; the launch's largest graph has at least 2^24 edges),
        PACKED_ALWAYS or PACKED_OFF.  Applies from the next launch."""
        _check(lib().pac_ctx_set_packed_lists(self._h, int(mode)))

    def stream_copy(self, nbytes: int, reps: int = 5) -> float:
        """Median time (ms) of a plain device-to-device streaming copy of nbytes."""
        ms = C.c_double()
        _check(lib().pac_stream_copy(self._h, nbytes, reps, C.byref(ms)))
        return ms.value

    # ---- tiles / autotuning ----
    def set_tile(self, tile: int):
        _check(lib().pac_set_tile(self._h, tile))

    def get_tile(self) -> int:
        t = C.c_int()
        _check(lib().pac_get_tile(self._h, C.byref(t)))
        return t.value

    def autotune(self, graphs, mode=MODE_TC, dispatch=DISPATCH_FIXED, cache_path=None):
        """Time every tile on the graphs' persistent launch in `mode` with `dispatch`, keep
        the fastest for this context; returns (tile, intersection ms per tile).  A cache
        file line for the same shapes, mode and dispatch is used without timing."""
        arr = (vp * len(graphs))(*[g._h for g in graphs])
        t = C.c_int()
        ms = np.zeros(NUM_TILES, np.float64)
        _check(lib().pac_autotune(self._h, arr, len(graphs), mode, dispatch,
                                  cache_path.encode() if cache_path else None, C.byref(t),
                                  _ptr(ms)))
        return t.value, ms

    def tc_batch(self, graphs, dispatch=DISPATCH_FIXED, launch=LAUNCH_PERSISTENT) -> np.ndarray:
        arr = (vp * len(graphs))(*[g._h for g in graphs])
        out = np.zeros(len(graphs), np.uint64)
        _check(lib().pac_tc_batch(self._h, arr, len(graphs), dispatch, launch, _ptr(out)))
        return out

    # ---- MLP ----
    def mlp_init(self, seed=1):
        _check(lib().pac_mlp_init(self._h, seed))

    def mlp_set_params(self, params):
        p = np.ascontiguousarray(params, np.float32)
        assert p.size == MLP_NPARAMS
        _check(lib().pac_mlp_set_params(self._h, _ptr(p)))

    def mlp_get_params(self) -> np.ndarray:
        p = np.zeros(MLP_NPARAMS, np.float32)
        _check(lib().pac_mlp_get_params(self._h, _ptr(p)))
        return p

    def mlp_forward(self, x) -> np.ndarray:
        x = np.ascontiguousarray(x, np.float32)
        out = np.zeros((x.shape[0], MLP_OUT), np.float32)
        _check(lib().pac_mlp_forward(self._h, _ptr(x), x.shape[0], _ptr(out)))
        return out

    def mlp_loss_grad(self, x, labels, weights=None):
        x = np.ascontiguousarray(x, np.float32)
        y = np.ascontiguousarray(labels, np.int32)
        w = None if weights is None else np.ascontiguousarray(weights, np.float32)
        g = np.zeros(MLP_NPARAMS, np.float32)
        loss = C.c_float()
        _check(lib().pac_mlp_loss_grad(self._h, _ptr(x), _ptr(y), _ptr(w), x.shape[0],
                                       C.byref(loss), _ptr(g)))
        return loss.value, g

    def mlp_train(self, x, labels, weights=None, epochs=500, lr=3e-3) -> float:
        x = np.ascontiguousarray(x, np.float32)
        y = np.ascontiguousarray(labels, np.int32)
        w = None if weights is None else np.ascontiguousarray(weights, np.float32)
        loss = C.c_float()
        _check(lib().pac_mlp_train(self._h, _ptr(x), _ptr(y), _ptr(w), x.shape[0], epochs, lr,
                                   C.byref(loss)))
        return loss.value

    def mlp_save(self, path):
        _check(lib().pac_mlp_save(self._h, path.encode()))

    def mlp_load(self, path):
        _check(lib().pac_mlp_load(self._h, path.encode()))


class Graph:
    """A symmetric graph on the device (pac_graph); reference Graph, Graph.hpp:38-103."""

    def __init__(self, ctx: Context, handle):
        self.ctx = ctx
        self._h = handle
        self._prepared_P = 0
        ctx._graphs.add(self)

    # ---- constructors ----
    @classmethod
    def from_csr(cls, ctx: Context, vlist, elist) -> "Graph":
        vl = np.ascontiguousarray(vlist, np.uint64)
        el = np.ascontiguousarray(elist, np.int32)
        h = vp()
        _check(lib().pac_graph_upload(ctx._h, _ptr(vl), len(vl) - 1, _ptr(el), len(el), C.byref(h)))
        return cls(ctx, h)

    @classmethod
    def from_edges(cls, ctx: Context, src, dst, num_vertices: int = 0) -> "Graph":
        """Undirected edge list -> symmetric CSR, built on the device."""
        s = np.ascontiguousarray(src, np.int32)
        d = np.ascontiguousarray(dst, np.int32)
        h = vp()
        _check(lib().pac_graph_from_edges(ctx._h, _ptr(s), _ptr(d), len(s), num_vertices,
                                          C.byref(h)))
        return cls(ctx, h)

    @classmethod
    def load_csr_dir(cls, ctx: Context, path: str) -> "Graph":
        h = vp()
        _check(lib().pac_graph_load_csr_dir(ctx._h, path.encode(), C.byref(h)))
        return cls(ctx, h)

    @classmethod
    def load_text(cls, ctx: Context, path: str) -> "Graph":
        h = vp()
        _check(lib().pac_graph_load_text(ctx._h, path.encode(), C.byref(h)))
        return cls(ctx, h)

    @classmethod
    def rmat(cls, ctx: Context, scale: int, edge_factor=16, a=0.57, b=0.19, c=0.19, seed=1):
        h = vp()
        _check(lib().pac_graph_generate_rmat(ctx._h, scale, edge_factor, a, b, c, seed,
                                             C.byref(h)))
        return cls(ctx, h)

    @classmethod
    def er(cls, ctx: Context, n: int, avg_degree: float, seed=1):
        h = vp()
        _check(lib().pac_graph_generate_er(ctx._h, n, avg_degree, seed, C.byref(h)))
        return cls(ctx, h)

    def free(self):
        if getattr(self, "_h", None) and getattr(self.ctx, "_h", None):
            lib().pac_graph_free(self._h)
        self._h = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass

    # ---- reference accessors (Graph.cpp:151-157, 288-298) ----
    def info(self) -> GraphInfo:
        i = GraphInfo()
        _check(lib().pac_graph_get_info(self._h, C.byref(i)))
        return i

    def get_vertex_num(self) -> int:
        return self.info().num_vertices

    def get_edge_num(self) -> int:
        """Directed entries 2m, as Graph::get_edge_num (Graph.cpp:288-290)."""
        return self.info().num_entries

    def csr(self):
        i = self.info()
        vl = np.zeros(i.num_vertices + 1, np.uint64)
        el = np.zeros(i.num_entries, np.int32)
        _check(lib().pac_graph_download_csr(self._h, _ptr(vl), _ptr(el)))
        return vl, el

    def save_csr_dir(self, path: str):
        _check(lib().pac_graph_save_csr_dir(self._h, path.encode()))

    def validate(self, level=1):
        _check(lib().pac_graph_validate(self._h, level))

    # ---- partitioning ----
    def prepare(self, num_partitions=1, v2p=None):
        a = None if v2p is None else np.ascontiguousarray(v2p, np.int32)
        _check(lib().pac_graph_prepare(self._h, num_partitions, _ptr(a)))
        self._prepared_P = self.info().num_partitions
        return self

    def _ensure_prepared(self):
        if not self._prepared_P:
            self.prepare(1)

    def partition_sizes(self) -> np.ndarray:
        out = np.zeros(self.info().num_partitions, np.uint64)
        _check(lib().pac_graph_partition_sizes(self._h, _ptr(out)))
        return out

    def greedy_partition(self, P: int, seed=1) -> np.ndarray:
        out = np.zeros(self.get_vertex_num(), np.int32)
        _check(lib().pac_greedy_partition(self._h, P, seed, _ptr(out)))
        return out

    def greedy_partition_2(self, P: int, seed=1) -> np.ndarray:
        """Graph::greedy_partition_2 (Graph.cpp:969-1015), seeded."""
        out = np.zeros(self.get_vertex_num(), np.int32)
        _check(lib().pac_greedy_partition_2(self._h, P, seed, _ptr(out)))
        return out

    def evaluate_partition(self, v2p, P: int) -> float:
        a = np.ascontiguousarray(v2p, np.int32)
        r = C.c_double()
        _check(lib().pac_evaluate_partition(self._h, _ptr(a), P, C.byref(r)))
        return r.value

    # ---- GPU-native solvers ----
    def tc(self, dispatch=DISPATCH_FIXED, launch=LAUNCH_PERSISTENT, per_vertex=False):
        self._ensure_prepared()
        tot = np.zeros(1, np.uint64)
        pv = np.zeros(self.get_vertex_num(), np.uint64) if per_vertex else None
        _check(lib().pac_tc(self._h, dispatch, launch, _ptr(tot), _ptr(pv)))
        return (int(tot[0]), pv) if per_vertex else int(tot[0])

    def lcc(self, dispatch=DISPATCH_FIXED, launch=LAUNCH_PERSISTENT, want64=True):
        self._ensure_prepared()
        V = self.get_vertex_num()
        l32 = np.zeros(V, np.float32)
        l64 = np.zeros(V, np.float64) if want64 else None
        tot = np.zeros(1, np.uint64)
        _check(lib().pac_lcc(self._h, dispatch, launch, _ptr(l32), _ptr(l64), _ptr(tot)))
        return l32, l64, int(tot[0])

    def lcc_device(self, dispatch=DISPATCH_FIXED, launch=LAUNCH_PERSISTENT) -> int:
        """TC + LCC with the per-vertex results left in device memory; returns T."""
        self._ensure_prepared()
        tot = np.zeros(1, np.uint64)
        _check(lib().pac_lcc(self._h, dispatch, launch, None, None, _ptr(tot)))
        return int(tot[0])

    def jaccard_edges(self, dispatch=DISPATCH_FIXED, launch=LAUNCH_PERSISTENT):
        self._ensure_prepared()
        m = self.info().num_edges
        src = np.zeros(m, np.int32)
        dst = np.zeros(m, np.int32)
        inter = np.zeros(m, np.uint32)
        uni = np.zeros(m, np.uint32)
        ratio = np.zeros(m, np.float32)
        _check(lib().pac_jaccard_edges(self._h, dispatch, launch, _ptr(src), _ptr(dst),
                                       _ptr(inter), _ptr(uni), _ptr(ratio)))
        return src, dst, inter, uni, ratio

    def jaccard_edges_device(self, dispatch=DISPATCH_FIXED, launch=LAUNCH_PERSISTENT):
        """Per-edge Jaccard with every result left in device memory."""
        self._ensure_prepared()
        _check(lib().pac_jaccard_edges(self._h, dispatch, launch, None, None, None, None, None))

    def cnt_common_pairs(self, u, v, thr=None, partition=-1) -> np.ndarray:
        """|N(u[i]) & N(v[i])|, optionally only common neighbours w < thr[i] and/or in
        `partition` of the map of the last prepare (-1: all)."""
        u = np.ascontiguousarray(u, np.int32)
        v = np.ascontiguousarray(v, np.int32)
        if len(v) != len(u):
            raise ValueError("u and v differ in length")
        out = np.zeros(len(u), np.uint64)
        if thr is None and partition == -1:
            _check(lib().pac_cnt_common_pairs(self._h, _ptr(u), _ptr(v), len(u), _ptr(out)))
            return out
        t = None
        if thr is not None:
            t = np.ascontiguousarray(np.broadcast_to(np.asarray(thr, np.int32), u.shape))
        _check(lib().pac_cnt_common_pairs_filtered(self._h, _ptr(u), _ptr(v), _ptr(t), len(u),
                                                   int(partition), _ptr(out)))
        return out

    def cnt_common_kway(self, queries, partition=-1) -> np.ndarray:
        """|N(q_0) & ... & N(q_k-1)| for each query (a sequence of vertex IDs), optionally
        only the common neighbours in `partition`."""
        qoff = np.zeros(len(queries) + 1, np.uint64)
        qoff[1:] = np.cumsum([len(q) for q in queries])
        qv = np.ascontiguousarray(np.concatenate([np.asarray(q, np.int32) for q in queries])
                                  if queries else np.zeros(0, np.int32), np.int32)
        out = np.zeros(len(queries), np.uint64)
        _check(lib().pac_cnt_common_kway_filtered(self._h, _ptr(qoff), _ptr(qv), len(queries),
                                                  int(partition), _ptr(out)))
        return out

    def common_neis(self, u, v=None, nodes=None, partition=-1) -> np.ndarray:
        """The common neighbours of u and v (or of u and the sorted list `nodes`), ascending,
        optionally restricted to `partition`."""
        if (v is None) == (nodes is None):
            raise ValueError("give exactly one of v and nodes")
        nd = None if nodes is None else np.ascontiguousarray(nodes, np.int32)
        n = 0 if nd is None else len(nd)
        cnt = C.c_uint64(0)
        _check(lib().pac_common_neis(self._h, int(u), -1 if v is None else int(v), _ptr(nd), n,
                                     int(partition), None, 0, C.byref(cnt)))
        out = np.zeros(max(1, cnt.value), np.int32)
        _check(lib().pac_common_neis(self._h, int(u), -1 if v is None else int(v), _ptr(nd), n,
                                     int(partition), _ptr(out), len(out), C.byref(cnt)))
        return out[:cnt.value].copy()

    # reference method names (Graph.cpp:241-253, 272-285, 413-447, 468-493)
    def cnt_common_neis_merge(self, u, v) -> int:
        return int(self.cnt_common_pairs([u], [v])[0])

    cnt_common_neis_galloping = cnt_common_neis_merge

    def cnt_common_neis_less_than_merge(self, u, v, threshold) -> int:
        return int(self.cnt_common_pairs([u], [v], thr=[threshold])[0])

    def cnt_common_neis_merge_nodes(self, u, nodes) -> int:
        return len(self.common_neis(u, nodes=nodes))

    def get_common_neis(self, u, v=None, nodes=None) -> np.ndarray:
        return self.common_neis(u, v=v, nodes=nodes)

    def batch_cnt_common_neis_merge(self, queries, num_threads=0) -> np.ndarray:
        """Pairs (a sequence of (u, v)) or k-way queries (sequences of any length)."""
        if len(queries) and all(len(q) == 2 for q in queries):
            q = np.asarray(queries, np.int32).reshape(-1, 2)
            return self.cnt_common_pairs(q[:, 0], q[:, 1])
        return self.cnt_common_kway(queries)

    batch_cnt_common_neis_galloping = batch_cnt_common_neis_merge

    def friend_triangle(self, root, nodes) -> int:
        nodes = np.ascontiguousarray(nodes, np.int32)
        out = np.zeros(1, np.uint64)
        _check(lib().pac_friend_triangle(self._h, int(root), _ptr(nodes), len(nodes), _ptr(out)))
        return int(out[0])

    def square_count(self) -> int:
        out = np.zeros(1, np.uint64)
        _check(lib().pac_square_count(self._h, _ptr(out)))
        return int(out[0])

    # reference method names (Graph.cpp:200-225, 255-270, 309-338, 1291-1324)
    def cnt_common_neis_merge_k(self, q) -> int:
        return int(self.cnt_common_kway([q])[0])

    def solve_friend_triangle_merge(self, root, nodes, num_threads=0) -> int:
        return self.friend_triangle(root, nodes)

    solve_friend_triangle_galloping = solve_friend_triangle_merge

    def solve_square_counting(self) -> int:
        return self.square_count()

    def features(self) -> np.ndarray:
        P = self.info().num_partitions
        out = np.zeros((P, MLP_IN), np.float32)
        _check(lib().pac_partition_features(self._h, _ptr(out)))
        return out

    def profile_partitions(self, reps=3) -> np.ndarray:
        P = self.info().num_partitions
        out = np.zeros((P, 3), np.float64)
        _check(lib().pac_profile_partitions(self._h, reps, _ptr(out)))
        return out

    def fixed_choice(self) -> np.ndarray:
        out = np.zeros(self.info().num_partitions, np.uint8)
        _check(lib().pac_fixed_choice(self._h, _ptr(out)))
        return out

    def mlp_choice(self) -> np.ndarray:
        out = np.zeros(self.info().num_partitions, np.uint8)
        _check(lib().pac_mlp_choice(self._h, _ptr(out)))
        return out

    # ---- reference method names (Graph.hpp:66-67, 90-93) ----
    def solve_triangle_counting_merge(self, num_threads: int = 0) -> int:
        """Graph::solve_triangle_counting_merge (Graph.cpp:355-368): T."""
        return self.tc()

    def solve_triangle_counting_galloping(self, num_threads: int = 0) -> int:
        """Graph::solve_triangle_counting_galloping (Graph.cpp:342-353): 6T."""
        return 6 * self.tc(dispatch=FORCE_GALLOP)

    def solve_lcc_merge(self, num_threads: int = 0) -> np.ndarray:
        """Graph::solve_lcc_merge (Graph.cpp:828-840): double per vertex."""
        return self.lcc()[1]

    def solve_lcc_galloping(self, num_threads: int = 0) -> np.ndarray:
        return self.lcc(dispatch=FORCE_GALLOP)[1]

    def solve_jsim_merge(self, queries, num_threads: int = 0) -> np.ndarray:
        """Graph::solve_jsim_merge (Graph.cpp:798-812): n*n doubles."""
        q = np.ascontiguousarray(queries, np.int32)
        out = np.zeros(len(q) * len(q), np.float64)
        _check(lib().pac_jaccard_queries(self._h, _ptr(q), len(q), _ptr(out)))
        return out

    solve_jsim_galloping = solve_jsim_merge


class PartitionedGraph:
    """Partitioned_Graph(g, vertex2partition, thread_num), Graph.cpp:1017-1069: records
    the partition map and prepares the device work queue for it."""

    def __init__(self, g: Graph, vertex2partition=None, thread_num: int = 0, num_partitions=0):
        self.g = g
        v2p = None if vertex2partition is None else np.ascontiguousarray(vertex2partition, np.int32)
        g.prepare(num_partitions, v2p)
        self.v2p = v2p

    def get_partition_num(self) -> int:
        return self.g.info().num_partitions

    def get_vertex_num(self) -> int:
        return self.g.get_vertex_num()

    def get_edge_num(self) -> int:
        return self.g.get_edge_num()

    def solve_triangle_counting_merge(self) -> int:
        return self.g.tc()

    def solve_triangle_counting_galloping(self) -> int:
        return 6 * self.g.tc(dispatch=FORCE_GALLOP)

    def solve_lcc_merge(self) -> np.ndarray:
        return self.g.lcc()[1]

    def solve_lcc_galloping(self) -> np.ndarray:
        return self.g.lcc(dispatch=FORCE_GALLOP)[1]

    # per-partition queries (Graph.cpp:1071-1195, 1475-1498): partition_id restricts the
    # common neighbours to that partition's vertices
    def cnt_common_neis_merge(self, u, v, partition_id) -> int:
        return int(self.g.cnt_common_pairs([u], [v], partition=partition_id)[0])

    cnt_common_neis_galloping = cnt_common_neis_merge

    def cnt_common_neis_less_than_merge(self, u, v, threshold, partition_id) -> int:
        return int(self.g.cnt_common_pairs([u], [v], thr=[threshold], partition=partition_id)[0])

    def cnt_common_neis_merge_k(self, q, partition_id) -> int:
        return int(self.g.cnt_common_kway([q], partition=partition_id)[0])

    cnt_common_neis_galloping_k = cnt_common_neis_merge_k

    def cnt_common_neis_merge_nodes(self, u, nodes, partition_id) -> int:
        return len(self.g.common_neis(u, nodes=nodes, partition=partition_id))

    def get_common_neis(self, u, nodes, partition_id) -> np.ndarray:
        return self.g.common_neis(u, nodes=nodes, partition=partition_id)

    def get_neis(self, u, partition_id) -> np.ndarray:
        return self.g.common_neis(u, v=u, partition=partition_id)

    def batch_cnt_common_neis_merge(self, queries) -> np.ndarray:
        """Sum over partitions of the per-partition counts = the whole-graph counts."""
        return self.g.batch_cnt_common_neis_merge(queries)

    batch_cnt_common_neis_galloping = batch_cnt_common_neis_merge
    batch_cnt_common_neis_merge_fine_grained = batch_cnt_common_neis_merge
    batch_cnt_common_neis_galloping_fine_grained = batch_cnt_common_neis_merge

    def solve_friend_triangle_merge(self, root, nodes) -> int:
        return self.g.friend_triangle(root, nodes)

    solve_friend_triangle_galloping = solve_friend_triangle_merge

    def solve_square_counting(self) -> int:
        return self.g.square_count()

    def solve_jsim_merge(self, queries) -> np.ndarray:
        """The Graph:: formula (Graph.cpp:798-812), not the partitioned variant's
        mis-indexed union (Graph.cpp:2192)."""
        return self.g.solve_jsim_merge(queries)

    solve_jsim_galloping = solve_jsim_merge

    def evaluate_partition(self) -> float:
        """Balance ratio P * max partition degree / max degree (Graph.cpp:1460-1473)."""
        P = self.get_partition_num()
        v2p = self.v2p if self.v2p is not None else contiguous_partition(self.get_vertex_num(), P)
        return self.g.evaluate_partition(v2p, P)


def random_partition(V: int, P: int, seed=1) -> np.ndarray:
    out = np.zeros(V, np.int32)
    _check(lib().pac_random_partition(V, P, seed, _ptr(out)))
    return out


def contiguous_partition(V: int, P: int) -> np.ndarray:
    w = max(1, -(-V // P))
    return (np.arange(V, dtype=np.int64) // w).astype(np.int32)
