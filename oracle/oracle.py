"""ctypes wrapper of the CPU oracle (oracle/liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, by __graft_entry__.smoke() (as the
checker) and by bench.py's cpu_baseline leg.  The product package never imports
this module.  Every function restates a reference function of
Byy-0024/PartitionAndConquer (src/Graph.cpp); the C source names the line ranges.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle.so")
_lib = None

u64p = np.ctypeslib.ndpointer(np.uint64, flags="C_CONTIGUOUS")
i32p = np.ctypeslib.ndpointer(np.int32, flags="C_CONTIGUOUS")
u32p = np.ctypeslib.ndpointer(np.uint32, flags="C_CONTIGUOUS")
f32p = np.ctypeslib.ndpointer(np.float32, flags="C_CONTIGUOUS")
f64p = np.ctypeslib.ndpointer(np.float64, flags="C_CONTIGUOUS")


class Pairs(C.Structure):
    _fields_ = [(n, C.c_uint64) for n in ("pairs", "rows", "streamed", "staged")]


class Stats(C.Structure):
    _fields_ = [(n, C.c_uint64) for n in
                ("V", "nnz", "m", "max_deg", "max_outdeg", "W", "sum_outdeg_sq", "sum_deg_sq")]


def build() -> str:
    subprocess.run(["make", "-s", "-C", _HERE], check=True)
    return _LIB_PATH


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = C.CDLL(_LIB_PATH)
        L.orc_text_to_csr.argtypes = [C.c_char_p, C.POINTER(C.c_uint64), C.POINTER(C.c_void_p),
                                      C.POINTER(C.c_void_p), C.POINTER(C.c_uint64)]
        L.orc_edges_to_csr.argtypes = [C.c_uint64, u32p, u32p, C.c_uint64, C.POINTER(C.c_void_p),
                                       C.POINTER(C.c_void_p), C.POINTER(C.c_uint64)]
        L.orc_free.argtypes = [C.c_void_p]
        for f in ("orc_cnt_merge", "orc_cnt_galloping"):
            getattr(L, f).argtypes = [u64p, i32p, C.c_int32, C.c_int32]
            getattr(L, f).restype = C.c_uint64
        L.orc_cnt_less_than_merge.argtypes = [u64p, i32p, C.c_int32, C.c_int32, C.c_int32]
        L.orc_cnt_less_than_merge.restype = C.c_uint64
        for f in ("orc_tc_merge", "orc_tc_galloping"):
            getattr(L, f).argtypes = [u64p, i32p, C.c_uint64, C.c_int]
            getattr(L, f).restype = C.c_uint64
        L.orc_tc_merge_rows.argtypes = [u64p, i32p, i32p, C.c_uint64, C.c_int]
        L.orc_tc_merge_rows.restype = C.c_uint64
        for f in ("orc_lcc_merge", "orc_lcc_galloping"):
            getattr(L, f).argtypes = [u64p, i32p, C.c_uint64, f64p, C.c_int]
        L.orc_lcc_numer.argtypes = [u64p, i32p, C.c_uint64, u64p, C.c_int]
        L.orc_lcc_merge_rows.argtypes = [u64p, i32p, i32p, C.c_uint64, f64p, C.c_int]
        L.orc_lcc_merge_rows.restype = C.c_uint64
        for f in ("orc_jsim_merge", "orc_jsim_galloping"):
            getattr(L, f).argtypes = [u64p, i32p, i32p, C.c_uint64, f64p, C.c_int]
        L.orc_jaccard_edges.argtypes = [u64p, i32p, i32p, i32p, C.c_uint64, u32p, u32p, f32p, C.c_int]
        L.orc_pgraph_build.argtypes = [u64p, i32p, C.c_uint64, i32p, C.c_int]
        L.orc_pgraph_build.restype = C.c_void_p
        for f in ("orc_pgraph_num_partitions", "orc_pgraph_tc_merge", "orc_pgraph_tc_galloping"):
            getattr(L, f).argtypes = [C.c_void_p]
            getattr(L, f).restype = C.c_uint64
        L.orc_pgraph_tc_merge_partial.argtypes = [C.c_void_p, C.c_uint64]
        L.orc_pgraph_tc_merge_partial.restype = C.c_uint64
        L.orc_pgraph_lcc_merge.argtypes = [C.c_void_p, f64p]
        L.orc_pgraph_free.argtypes = [C.c_void_p]
        L.orc_pgraph_evaluate_partition.argtypes = [C.c_void_p, C.POINTER(C.c_uint64),
                                                    C.POINTER(C.c_uint64)]
        L.orc_pgraph_evaluate_partition.restype = C.c_double
        L.orc_random_partition.argtypes = [C.c_uint64, C.c_uint64, C.c_uint64, i32p]
        L.orc_contiguous_partition.argtypes = [C.c_uint64, C.c_uint64, i32p]
        L.orc_greedy_partition_2.argtypes = [u64p, i32p, C.c_uint64, C.c_uint64, C.c_uint64, i32p]
        L.orc_gen_rmat.argtypes = [C.c_uint32, C.c_uint32, C.c_double, C.c_double, C.c_double,
                                   C.c_uint64, C.POINTER(C.c_uint64), C.POINTER(C.c_void_p),
                                   C.POINTER(C.c_void_p), C.POINTER(C.c_uint64)]
        L.orc_gen_er.argtypes = [C.c_uint64, C.c_double, C.c_uint64, C.POINTER(C.c_uint64),
                                 C.POINTER(C.c_void_p), C.POINTER(C.c_void_p), C.POINTER(C.c_uint64)]
        L.orc_oriented_stats.argtypes = [u64p, i32p, C.c_uint64, C.POINTER(Stats)]
        L.orc_pair_stats.argtypes = [u64p, i32p, C.c_uint64, C.POINTER(Pairs)]
        L.orc_tc_oriented.argtypes = [u64p, i32p, C.c_uint64, C.c_int]
        L.orc_tc_oriented.restype = C.c_uint64
        L.orc_mlp_forward.argtypes = [f32p, f32p, C.c_uint64, f32p]
        L.orc_mlp_backward.argtypes = [f32p, f32p, i32p, C.c_void_p, C.c_uint64, f32p]
        L.orc_mlp_backward.restype = C.c_float
        for f in ("orc_kway_merge", "orc_kway_galloping"):
            getattr(L, f).argtypes = [u64p, i32p, i32p, C.c_uint64]
            getattr(L, f).restype = C.c_uint64
        L.orc_friend_triangle_merge.argtypes = [u64p, i32p, C.c_int32, i32p, C.c_uint64, C.c_int]
        L.orc_friend_triangle_merge.restype = C.c_uint64
        L.orc_pgraph_square_counting.argtypes = [C.c_void_p]
        L.orc_pgraph_square_counting.restype = C.c_uint64
        L.orc_pgraph_cnt.argtypes = [C.c_void_p, C.c_int32, C.c_int32, C.c_uint64, C.c_int]
        L.orc_pgraph_cnt.restype = C.c_uint64
        L.orc_pgraph_cnt_less_than.argtypes = [C.c_void_p, C.c_int32, C.c_int32, C.c_int32,
                                               C.c_uint64]
        L.orc_pgraph_cnt_less_than.restype = C.c_uint64
        L.orc_pgraph_cnt_nodes.argtypes = [C.c_void_p, C.c_int32, i32p, C.c_uint64, C.c_uint64]
        L.orc_pgraph_cnt_nodes.restype = C.c_uint64
        L.orc_pgraph_common_neis.argtypes = [C.c_void_p, C.c_int32, i32p, C.c_uint64, C.c_uint64,
                                             i32p]
        L.orc_pgraph_common_neis.restype = C.c_uint64
        L.orc_pgraph_get_neis.argtypes = [C.c_void_p, C.c_int32, C.c_uint64, i32p]
        L.orc_pgraph_get_neis.restype = C.c_uint64
        L.orc_pgraph_kway.argtypes = [C.c_void_p, i32p, C.c_uint64, C.c_uint64, C.c_int]
        L.orc_pgraph_kway.restype = C.c_uint64
        L.orc_cnt_merge_nodes.argtypes = [u64p, i32p, C.c_int32, i32p, C.c_uint64]
        L.orc_cnt_merge_nodes.restype = C.c_uint64
        L.orc_common_neis_nodes.argtypes = [u64p, i32p, C.c_int32, i32p, C.c_uint64, i32p]
        L.orc_common_neis_nodes.restype = C.c_uint64
        _lib = L
    return _lib


def _nt(nt):
    return nt if nt else (os.cpu_count() or 1)


class CSR:
    """Symmetric CSR in the reference's in-memory types: vlist uint64[V+1],
    elist int32[2m] (Graph.hpp:99-100)."""

    def __init__(self, vlist: np.ndarray, elist: np.ndarray):
        self.vlist = np.ascontiguousarray(vlist, dtype=np.uint64)
        self.elist = np.ascontiguousarray(elist, dtype=np.int32)
        self.V = len(self.vlist) - 1

    @property
    def nnz(self) -> int:
        return int(self.vlist[-1])

    def deg(self) -> np.ndarray:
        return np.diff(self.vlist).astype(np.int64)

    def edges(self):
        """Undirected edges (u<v) in CSR order."""
        src = np.repeat(np.arange(self.V, dtype=np.int32), np.diff(self.vlist).astype(np.int64))
        keep = src < self.elist
        return src[keep], self.elist[keep]

    def dense(self) -> np.ndarray:
        A = np.zeros((self.V, self.V), dtype=np.int64)
        src = np.repeat(np.arange(self.V), np.diff(self.vlist).astype(np.int64))
        A[src, self.elist] = 1
        return A


def _take_csr(V, vp, ep, nnz) -> CSR:
    L = lib()
    vl = np.ctypeslib.as_array(C.cast(vp, C.POINTER(C.c_uint64)), shape=(V + 1,)).copy()
    el = (np.ctypeslib.as_array(C.cast(ep, C.POINTER(C.c_int32)), shape=(nnz,)).copy()
          if nnz else np.zeros(0, np.int32))
    L.orc_free(vp)
    L.orc_free(ep)
    return CSR(vl, el)


def text_to_csr(path: str) -> CSR:
    """Graph::Graph(string), Graph.cpp:27-90."""
    V, nnz, vp, ep = C.c_uint64(), C.c_uint64(), C.c_void_p(), C.c_void_p()
    rc = lib().orc_text_to_csr(path.encode(), C.byref(V), C.byref(vp), C.byref(ep), C.byref(nnz))
    if rc != 0:
        raise OSError(f"cannot open {path}")
    return _take_csr(V.value, vp, ep, nnz.value)


def edges_to_csr(V: int, src, dst) -> CSR:
    src = np.ascontiguousarray(src, dtype=np.uint32)
    dst = np.ascontiguousarray(dst, dtype=np.uint32)
    nnz, vp, ep = C.c_uint64(), C.c_void_p(), C.c_void_p()
    rc = lib().orc_edges_to_csr(V, src, dst, len(src), C.byref(vp), C.byref(ep), C.byref(nnz))
    if rc != 0:
        raise ValueError("edge endpoint out of range")
    return _take_csr(V, vp, ep, nnz.value)


def gen_rmat(scale, edge_factor=16, a=0.57, b=0.19, c=0.19, seed=1) -> CSR:
    V, nnz, vp, ep = C.c_uint64(), C.c_uint64(), C.c_void_p(), C.c_void_p()
    rc = lib().orc_gen_rmat(scale, edge_factor, a, b, c, seed, C.byref(V), C.byref(vp),
                            C.byref(ep), C.byref(nnz))
    if rc != 0:
        raise ValueError("rmat generation failed")
    return _take_csr(V.value, vp, ep, nnz.value)


def gen_er(n, avg_deg, seed=1) -> CSR:
    V, nnz, vp, ep = C.c_uint64(), C.c_uint64(), C.c_void_p(), C.c_void_p()
    rc = lib().orc_gen_er(n, avg_deg, seed, C.byref(V), C.byref(vp), C.byref(ep), C.byref(nnz))
    if rc != 0:
        raise ValueError("er generation failed")
    return _take_csr(V.value, vp, ep, nnz.value)


def cnt_merge(g: CSR, u, v) -> int:
    return lib().orc_cnt_merge(g.vlist, g.elist, int(u), int(v))


def cnt_less_than_merge(g: CSR, u, v, thr) -> int:
    return lib().orc_cnt_less_than_merge(g.vlist, g.elist, int(u), int(v), int(thr))


def cnt_galloping(g: CSR, u, v) -> int:
    return lib().orc_cnt_galloping(g.vlist, g.elist, u, v)


def tc_merge(g: CSR, nt=0) -> int:
    return lib().orc_tc_merge(g.vlist, g.elist, g.V, _nt(nt))


def tc_galloping(g: CSR, nt=0) -> int:
    return lib().orc_tc_galloping(g.vlist, g.elist, g.V, _nt(nt))


def tc_merge_rows(g: CSR, rows, nt=0) -> int:
    rows = np.ascontiguousarray(rows, dtype=np.int32)
    return lib().orc_tc_merge_rows(g.vlist, g.elist, rows, len(rows), _nt(nt))


def lcc_merge(g: CSR, nt=0) -> np.ndarray:
    out = np.zeros(g.V, np.float64)
    lib().orc_lcc_merge(g.vlist, g.elist, g.V, out, _nt(nt))
    return out


def lcc_galloping(g: CSR, nt=0) -> np.ndarray:
    out = np.zeros(g.V, np.float64)
    lib().orc_lcc_galloping(g.vlist, g.elist, g.V, out, _nt(nt))
    return out


def lcc_numer(g: CSR, nt=0) -> np.ndarray:
    out = np.zeros(g.V, np.uint64)
    lib().orc_lcc_numer(g.vlist, g.elist, g.V, out, _nt(nt))
    return out


def lcc_merge_rows(g: CSR, rows, nt=0):
    rows = np.ascontiguousarray(rows, dtype=np.int32)
    out = np.zeros(len(rows), np.float64)
    tot = lib().orc_lcc_merge_rows(g.vlist, g.elist, rows, len(rows), out, _nt(nt))
    return tot, out


def per_vertex_triangles(g: CSR, nt=0) -> np.ndarray:
    """t(u) = numerator/2 of the reference LCC (Graph.cpp:834-837)."""
    n = lcc_numer(g, nt)
    assert np.all(n % 2 == 0)
    return n // 2


def jsim_merge(g: CSR, q, nt=0) -> np.ndarray:
    q = np.ascontiguousarray(q, dtype=np.int32)
    out = np.zeros(len(q) * len(q), np.float64)
    lib().orc_jsim_merge(g.vlist, g.elist, q, len(q), out, _nt(nt))
    return out.reshape(len(q), len(q))


def jsim_galloping(g: CSR, q, nt=0) -> np.ndarray:
    q = np.ascontiguousarray(q, dtype=np.int32)
    out = np.zeros(len(q) * len(q), np.float64)
    lib().orc_jsim_galloping(g.vlist, g.elist, q, len(q), out, _nt(nt))
    return out.reshape(len(q), len(q))


def jaccard_edges(g: CSR, src, dst, nt=0):
    src = np.ascontiguousarray(src, dtype=np.int32)
    dst = np.ascontiguousarray(dst, dtype=np.int32)
    n = len(src)
    inter = np.zeros(n, np.uint32)
    uni = np.zeros(n, np.uint32)
    ratio = np.zeros(n, np.float32)
    lib().orc_jaccard_edges(g.vlist, g.elist, src, dst, n, inter, uni, ratio, _nt(nt))
    return inter, uni, ratio


class PGraph:
    """Partitioned_Graph (Graph.cpp:1017-1069) and its TC/LCC solvers."""

    def __init__(self, g: CSR, v2p, nt=0):
        self.v2p = np.ascontiguousarray(v2p, dtype=np.int32)
        self._h = lib().orc_pgraph_build(g.vlist, g.elist, g.V, self.v2p, _nt(nt))
        self.V = g.V
        self.max_deg = int(g.deg().max()) if g.V else 0

    def __del__(self):
        if getattr(self, "_h", None):
            lib().orc_pgraph_free(self._h)
            self._h = None

    @property
    def P(self) -> int:
        return lib().orc_pgraph_num_partitions(self._h)

    def tc_merge(self) -> int:
        return lib().orc_pgraph_tc_merge(self._h)

    def tc_galloping(self) -> int:
        return lib().orc_pgraph_tc_galloping(self._h)

    def tc_merge_partial(self, i) -> int:
        return lib().orc_pgraph_tc_merge_partial(self._h, i)

    def lcc_merge(self) -> np.ndarray:
        out = np.zeros(self.V, np.float64)
        lib().orc_pgraph_lcc_merge(self._h, out)
        return out

    def square_counting(self) -> int:
        return lib().orc_pgraph_square_counting(self._h)

    def evaluate_partition(self):
        """Partitioned_Graph::evaluate_partition, Graph.cpp:1460-1473: (balance ratio,
        max_part_deg, max_deg)."""
        mp, md = C.c_uint64(), C.c_uint64()
        r = lib().orc_pgraph_evaluate_partition(self._h, C.byref(mp), C.byref(md))
        return r, mp.value, md.value

    # per-partition queries, Graph.cpp:1071-1195, 1475-1498
    def cnt_merge(self, u, v, k) -> int:
        return lib().orc_pgraph_cnt(self._h, int(u), int(v), int(k), 0)

    def cnt_galloping(self, u, v, k) -> int:
        return lib().orc_pgraph_cnt(self._h, int(u), int(v), int(k), 1)

    def cnt_less_than_merge(self, u, v, thr, k) -> int:
        return lib().orc_pgraph_cnt_less_than(self._h, int(u), int(v), int(thr), int(k))

    def cnt_merge_nodes(self, u, nodes, k) -> int:
        nodes = np.ascontiguousarray(nodes, dtype=np.int32)
        return lib().orc_pgraph_cnt_nodes(self._h, int(u), nodes, len(nodes), int(k))

    def common_neis(self, u, nodes, k) -> np.ndarray:
        nodes = np.ascontiguousarray(nodes, dtype=np.int32)
        out = np.zeros(max(1, len(nodes)), np.int32)
        c = lib().orc_pgraph_common_neis(self._h, int(u), nodes, len(nodes), int(k), out)
        return out[:c].copy()

    def get_neis(self, u, k) -> np.ndarray:
        out = np.zeros(max(1, self.max_deg), np.int32)
        c = lib().orc_pgraph_get_neis(self._h, int(u), int(k), out)
        return out[:c].copy()

    def kway_merge(self, q, k) -> int:
        q = np.ascontiguousarray(q, dtype=np.int32)
        return lib().orc_pgraph_kway(self._h, q, len(q), int(k), 0)

    def kway_galloping(self, q, k) -> int:
        q = np.ascontiguousarray(q, dtype=np.int32)
        return lib().orc_pgraph_kway(self._h, q, len(q), int(k), 1)


def kway_merge(g: CSR, q) -> int:
    q = np.ascontiguousarray(q, dtype=np.int32)
    return lib().orc_kway_merge(g.vlist, g.elist, q, len(q))


def kway_galloping(g: CSR, q) -> int:
    q = np.ascontiguousarray(q, dtype=np.int32)
    return lib().orc_kway_galloping(g.vlist, g.elist, q, len(q))


def cnt_merge_nodes(g: CSR, u, nodes) -> int:
    nodes = np.ascontiguousarray(nodes, dtype=np.int32)
    return lib().orc_cnt_merge_nodes(g.vlist, g.elist, int(u), nodes, len(nodes))


def common_neis(g: CSR, u, nodes) -> np.ndarray:
    """Graph::get_common_neis(u, nodes) (Graph.cpp:481-493); with nodes = N(v) it is
    Graph::get_common_neis(u, v) (Graph.cpp:468-479)."""
    nodes = np.ascontiguousarray(nodes, dtype=np.int32)
    out = np.zeros(max(1, len(nodes)), np.int32)
    c = lib().orc_common_neis_nodes(g.vlist, g.elist, int(u), nodes, len(nodes), out)
    return out[:c].copy()


def friend_triangle_merge(g: CSR, root, nodes, nt=0) -> int:
    nodes = np.ascontiguousarray(nodes, dtype=np.int32)
    return lib().orc_friend_triangle_merge(g.vlist, g.elist, int(root), nodes, len(nodes), _nt(nt))


def random_partition(V, P, seed=1) -> np.ndarray:
    out = np.zeros(V, np.int32)
    lib().orc_random_partition(V, P, seed, out)
    return out


def greedy_partition_2(g: CSR, P, seed=1) -> np.ndarray:
    """Graph::greedy_partition_2 (Graph.cpp:969-1015), seeded."""
    out = np.zeros(g.V, np.int32)
    lib().orc_greedy_partition_2(g.vlist, g.elist, g.V, P, seed, out)
    return out


def contiguous_partition(V, P) -> np.ndarray:
    out = np.zeros(V, np.int32)
    lib().orc_contiguous_partition(V, P, out)
    return out


def oriented_stats(g: CSR) -> dict:
    s = Stats()
    lib().orc_oriented_stats(g.vlist, g.elist, g.V, C.byref(s))
    return {n: getattr(s, n) for n, _ in Stats._fields_}


def pair_stats(g: CSR) -> dict:
    """Intersection pairs of the degree orientation (the hash path's work model):
    pairs, owning rows, streamed elements (sum of min(d+(t), d+(h)) over oriented edges)
    and staged elements (sum of d+ over the owning rows)."""
    s = Pairs()
    lib().orc_pair_stats(g.vlist, g.elist, g.V, C.byref(s))
    return {n: getattr(s, n) for n, _ in Pairs._fields_}


def tc_oriented(g: CSR, nt=0) -> int:
    """T by the degree-ordered forward algorithm: a fast exact checker, pinned against
    tc_merge (the restatement of Graph.cpp:355-368) by the CPU tests."""
    return int(lib().orc_tc_oriented(g.vlist, g.elist, g.V, _nt(nt)))


N_PARAMS = 64 * 16 + 64 + 64 * 64 + 64 + 3 * 64 + 3


def mlp_forward(params, x) -> np.ndarray:
    params = np.ascontiguousarray(params, np.float32)
    x = np.ascontiguousarray(x, np.float32)
    n = x.shape[0]
    out = np.zeros((n, 3), np.float32)
    lib().orc_mlp_forward(params, x, n, out)
    return out


def mlp_backward(params, x, labels, weights=None):
    params = np.ascontiguousarray(params, np.float32)
    x = np.ascontiguousarray(x, np.float32)
    labels = np.ascontiguousarray(labels, np.int32)
    grads = np.zeros(N_PARAMS, np.float32)
    wp = None
    if weights is not None:
        weights = np.ascontiguousarray(weights, np.float32)
        wp = weights.ctypes.data_as(C.c_void_p)
    loss = lib().orc_mlp_backward(params, x, labels, wp, x.shape[0], grads)
    return float(loss), grads
