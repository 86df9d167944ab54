"""The benchmark configurations at their real sizes, on the GPU through the C ABI.

The small configuration (R-MAT scale 16, 64 partitions) is checked against the full oracle
in test_gpu_parity.test_config1_scale16_64_partitions.  Here:
  * R-MAT scale 22, ef 16, seed 22, 1024 partitions;
  * 256 seeded graphs in one launch after autotuning the tile;
  * R-MAT scale 24, ef 16, seed 24, 4096 partitions, learned dispatch replayed from a
    hipGraph (the headline workload);
  * per-edge Jaccard at R-MAT scale 25, ef 16, seed 25.
At these sizes the oracle's restatement of the reference (ID-ordered merges; LCC costs
2 sum d(u)^2 element steps) runs for minutes to hours, so each result is checked by:
the oracle on a random sample of rows or edges (bit-exact), the oracle's fast
degree-ordered count for whole-graph totals where it finishes in seconds, device paths
that share no lookup code with the hash table (merge path, binary search), and
size-independent identities (sum of t(u) = 3T, sum of edge supports = 3T)."""
import os

import numpy as np
import pytest

import partitionandconquer_amd as pac

pytestmark = pytest.mark.gpu

NT = min(16, os.cpu_count() or 1)
MODEL = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))),
                     "partitionandconquer_amd", "dispatch_mlp.bin")
LAUNCHES = [pac.LAUNCH_PLAIN, pac.LAUNCH_PERSISTENT, pac.LAUNCH_GRAPH]
# T of R-MAT scale 24, ef 16, (0.57, 0.19, 0.19, 0.05), seed 24, counted by the oracle's
# tc_merge on the CSR the device generator produced (the generator equals the oracle's
# gen_rmat bit for bit at the scales the parity tests compare)
RMAT24_TRIANGLES = 10283762775


def log(*a):
    print(*a, flush=True)


def bits_equal(a, b):
    a, b = np.asarray(a), np.asarray(b)
    return a.dtype == b.dtype and np.array_equal(a.view(np.uint8), b.view(np.uint8))


def check_stats(g, ref, orc):
    """The GPU's orientation and pair counts (the roofline's work model) against the
    oracle on the same CSR."""
    info = g.info()
    st = orc.oriented_stats(ref)
    assert (info.num_edges, info.work_W, info.max_out_degree, info.max_degree) == \
        (st["m"], st["W"], st["max_outdeg"], st["max_deg"])
    ps = orc.pair_stats(ref)
    assert (info.num_pairs, info.num_rows, info.streamed_elems, info.staged_elems) == \
        (ps["pairs"], ps["rows"], ps["streamed"], ps["staged"])
    assert info.staged_item_elems >= info.staged_elems
    return info


def sample_rows(ref, n, rng, max_deg=4096, n_heavy=16):
    """n random non-isolated rows of degree <= max_deg plus up to n_heavy rows of degree in
    (max_deg, 8 max_deg] (hub-adjacent rows, whose oracle merges cost d(u)^2)."""
    d = ref.deg()
    light = np.nonzero((d >= 1) & (d <= max_deg))[0]
    heavy = np.nonzero((d > max_deg) & (d <= 8 * max_deg))[0]
    rows = rng.choice(light, size=min(n, len(light)), replace=False)
    if len(heavy):
        rows = np.concatenate([rows, rng.choice(heavy, size=min(n_heavy, len(heavy)),
                                                replace=False)])
    return np.sort(rows).astype(np.int32)


def check_lcc_rows(ref, orc, l32, l64, rows):
    want = orc.lcc_merge_rows(ref, rows, NT)[1]
    assert bits_equal(l64[rows], want)
    assert bits_equal(l32[rows], want.astype(np.float32))


def test_rmat22_1024_partitions(ctx, orc):
    """R-MAT scale 22, 1024 contiguous partitions: T equals the oracle's count; every
    strategy on every launch path gives it; t(u) sums to 3T and agrees between the merge
    path and the hash table; LCC is bit-exact on 2000+ sampled rows."""
    log("case: R-MAT scale 22, ef 16, seed 22, 1024 partitions")
    g = pac.Graph.rmat(ctx, 22, 16, 0.57, 0.19, 0.19, seed=22)
    g.prepare(1024)
    vl, el = g.csr()
    ref = orc.CSR(vl, el)
    check_stats(g, ref, orc)
    T = orc.tc_oriented(ref, NT)
    for d in (pac.FORCE_MERGE, pac.FORCE_GALLOP, pac.FORCE_BITMAP, pac.DISPATCH_FIXED):
        for l in LAUNCHES:
            assert g.lcc_device(d, l) == T, (d, l)
    t1, pv1 = g.tc(pac.FORCE_MERGE, pac.LAUNCH_PERSISTENT, per_vertex=True)
    t2, pv2 = g.tc(pac.FORCE_BITMAP, pac.LAUNCH_GRAPH, per_vertex=True)
    assert t1 == T and t2 == T and int(pv1.sum()) == 3 * T
    assert np.array_equal(pv1, pv2)
    l32, l64, t3 = g.lcc(pac.DISPATCH_FIXED, pac.LAUNCH_GRAPH)
    assert t3 == T
    check_lcc_rows(ref, orc, l32, l64, sample_rows(ref, 2000, np.random.default_rng(22)))


def test_batch_256_graphs_autotuned(ctx, orc, tmp_path):
    """256 seeded graphs of uneven size (half Erdos-Renyi, half R-MAT scale 10-18) in one
    launch: the autotuner's tile is the argmin of its times and survives a cache round
    trip; every graph's count equals the oracle's, on the persistent and hipGraph paths and
    on every tile."""
    from experiments import batch_specs
    log("case: 256-graph batch (batch_specs(256), seeds 3000+g)")
    gs, refs = [], []
    for sp in batch_specs(256):
        if sp["kind"] == "er":
            g = pac.Graph.er(ctx, sp["n"], sp["k"], seed=sp["seed"])
        else:
            g = pac.Graph.rmat(ctx, sp["scale"], 16, 0.57, 0.19, 0.19, seed=sp["seed"])
        g.prepare(max(1, -(-g.get_vertex_num() // 4096)))
        gs.append(g)
    want = np.array([orc.tc_oriented(orc.CSR(*g.csr()), NT) for g in gs], np.uint64)
    cache = str(tmp_path / "tiles.txt")
    tile, ms = ctx.autotune(gs, pac.MODE_TC, pac.DISPATCH_FIXED, cache)
    assert tile == int(np.argmin(ms)) and ctx.get_tile() == tile and (ms > 0).all()
    ctx.set_tile((tile + 1) % pac.NUM_TILES)
    tile2, ms2 = ctx.autotune(gs, pac.MODE_TC, pac.DISPATCH_FIXED, cache)   # from the cache
    assert tile2 == tile and ctx.get_tile() == tile and np.array_equal(ms, ms2)
    for l in (pac.LAUNCH_PERSISTENT, pac.LAUNCH_GRAPH):
        assert np.array_equal(ctx.tc_batch(gs, pac.DISPATCH_FIXED, l), want), l
    for t in range(pac.NUM_TILES):
        ctx.set_tile(t)
        assert np.array_equal(ctx.tc_batch(gs, pac.FORCE_BITMAP, pac.LAUNCH_PERSISTENT), want), t
    ctx.set_tile(tile)
    # a tuning error leaves the context's tile alone
    with pytest.raises(pac.PacError):
        ctx.autotune(gs + [gs[0]], pac.MODE_TC, pac.DISPATCH_FIXED)
    assert ctx.get_tile() == tile
    ctx.set_tile(1)
    for g in gs:
        g.free()


def test_rmat24_learned_dispatch_headline(ctx, orc):
    """The headline workload (R-MAT scale 24, 4096 partitions, learned dispatch, hipGraph
    replays): T from the replays equals T from the merge-path and binary-search kernels
    (no hash table involved); t(u) sums to 3T; LCC is bit-exact on 2000+ sampled rows; the
    orientation and pair counts equal the oracle's."""
    log("case: R-MAT scale 24, ef 16, seed 24, 4096 partitions, learned dispatch")
    g = pac.Graph.rmat(ctx, 24, 16, 0.57, 0.19, 0.19, seed=24)
    g.prepare(4096)
    assert os.path.exists(MODEL)
    ctx.mlp_load(MODEL)
    try:
        T = g.lcc_device(pac.DISPATCH_LEARNED, pac.LAUNCH_GRAPH)
        # the oracle's restated Graph::solve_triangle_counting_merge (Graph.cpp:355-368) over
        # the whole graph, run on the GPU box's host in 376 s (experiments.py cpu --scale 24)
        assert T == RMAT24_TRIANGLES
        assert g.lcc_device(pac.DISPATCH_LEARNED, pac.LAUNCH_GRAPH) == T   # a replay
        assert g.tc(pac.FORCE_MERGE, pac.LAUNCH_PERSISTENT) == T
        assert g.tc(pac.FORCE_GALLOP, pac.LAUNCH_PERSISTENT) == T
        t, pv = g.tc(pac.DISPATCH_LEARNED, pac.LAUNCH_GRAPH, per_vertex=True)
        assert t == T and int(pv.sum()) == 3 * T
        l32, l64, t2 = g.lcc(pac.DISPATCH_LEARNED, pac.LAUNCH_GRAPH)
        assert t2 == T
        vl, el = g.csr()
        ref = orc.CSR(vl, el)
        check_stats(g, ref, orc)
        rows = sample_rows(ref, 2000, np.random.default_rng(24))
        check_lcc_rows(ref, orc, l32, l64, rows)
        # t(u) itself on the sampled rows: 2 t(u) is the reference's LCC numerator
        numer = orc.lcc_merge_rows(ref, rows, NT)[1] * (ref.deg()[rows] * (ref.deg()[rows] - 1))
        ok = ref.deg()[rows] >= 2
        assert np.array_equal(2 * pv[rows][ok], np.rint(numer[ok]).astype(np.uint64))
    finally:
        g.free()


def test_rmat25_jaccard_edges(ctx, orc):
    """Per-edge Jaccard at R-MAT scale 25 (537M edges): the hash path's supports (hit
    words summed after the launch) equal the binary-search path's (one atomic per
    triangle) on every edge; they sum to 3T; 50 000 sampled edges are exact in
    intersection, union and ratio bits against the oracle."""
    log("case: R-MAT scale 25, ef 16, seed 25, 4096 partitions, per-edge Jaccard")
    g = pac.Graph.rmat(ctx, 25, 16, 0.57, 0.19, 0.19, seed=25)
    g.prepare(4096)
    try:
        T = g.tc(pac.FORCE_BITMAP, pac.LAUNCH_PERSISTENT)
        assert g.tc(pac.FORCE_MERGE, pac.LAUNCH_PERSISTENT) == T
        _, _, inter_g, _, _ = g.jaccard_edges(pac.FORCE_GALLOP, pac.LAUNCH_PERSISTENT)
        src, dst, inter, uni, ratio = g.jaccard_edges(pac.DISPATCH_FIXED, pac.LAUNCH_GRAPH)
        assert np.array_equal(inter, inter_g)
        del inter_g
        assert int(inter.sum(dtype=np.uint64)) == 3 * T
        vl, el = g.csr()
        ref = orc.CSR(vl, el)
        d = ref.deg()
        assert np.array_equal(uni.astype(np.int64), d[src] + d[dst] - inter)
        rng = np.random.default_rng(25)
        idx = rng.choice(len(src), size=50000, replace=False)
        ri, ru, rr = orc.jaccard_edges(ref, src[idx], dst[idx], NT)
        assert np.array_equal(ri, inter[idx]) and np.array_equal(ru, uni[idx])
        assert bits_equal(rr, ratio[idx])
    finally:
        g.free()
