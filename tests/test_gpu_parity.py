"""GPU parity: the HIP path through the C ABI against the CPU oracle on the same
inputs.  Counts (T, t(u), intersection and union sizes) must be bit-exact; the float
LCC/Jaccard ratios are computed from those integers and must be bit-identical to
(float)(double ratio) of the reference formula (a tolerance of 0 ulp, stricter than
the 1-ulp budget)."""
import itertools
import json
import math
import os

import numpy as np
import pytest

import partitionandconquer_amd as pac

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")
NAMES = sorted(json.load(open(os.path.join(GOLDEN, "golden_index.json"))))
STRATS = [pac.FORCE_MERGE, pac.FORCE_GALLOP, pac.FORCE_BITMAP, pac.DISPATCH_FIXED]
LAUNCHES = [pac.LAUNCH_PLAIN, pac.LAUNCH_PERSISTENT, pac.LAUNCH_GRAPH]
NT = min(16, os.cpu_count() or 1)


def bits_equal(a, b):
    a, b = np.asarray(a), np.asarray(b)
    return a.dtype == b.dtype and np.array_equal(a.view(np.uint8), b.view(np.uint8))


def check_all(ctx, g, ref, orc, P=16, v2p=None, strategies=STRATS, launches=LAUNCHES):
    T = orc.tc_merge(ref, NT)
    tv = orc.per_vertex_triangles(ref, NT)
    g.prepare(P, v2p)
    for d in strategies:
        for l in launches:
            t, pv = g.tc(d, l, per_vertex=True)
            assert t == T, (d, l)
            assert np.array_equal(pv, tv), (d, l)
            assert g.tc(d, l) == T, (d, l)
    return T


@pytest.mark.parametrize("name", NAMES)
def test_golden_fixtures_on_gpu(ctx, orc, name):
    z = np.load(os.path.join(GOLDEN, f"{name}.npz"))
    g = pac.Graph.from_csr(ctx, z["vlist"], z["elist"])
    g.prepare(3)
    for d in STRATS:
        for l in LAUNCHES:
            t, pv = g.tc(d, l, per_vertex=True)
            assert t == int(z["T"]) and np.array_equal(pv, z["t"])
            l32, l64, t2 = g.lcc(d, l)
            assert t2 == int(z["T"])
            assert bits_equal(l64, z["lcc"])
            assert bits_equal(l32, z["lcc"].astype(np.float32))
    src, dst, inter, uni, ratio = g.jaccard_edges()
    lo, hi = np.minimum(src, dst), np.maximum(src, dst)
    order = np.lexsort((hi, lo))
    assert np.array_equal(lo[order], z["eu"]) and np.array_equal(hi[order], z["ev"])
    assert np.array_equal(inter[order], z["inter"]) and np.array_equal(uni[order], z["uni"])
    assert bits_equal(ratio[order], z["ratio"])
    assert bits_equal(g.solve_jsim_merge(z["queries"]), z["jsim"].ravel())


@pytest.mark.parametrize("scale,seed", [(6, 1), (9, 2), (12, 3), (14, 4)])
def test_rmat_all_strategies_and_launches(ctx, orc, scale, seed):
    g = pac.Graph.rmat(ctx, scale, 16, 0.57, 0.19, 0.19, seed=seed)
    ref = orc.gen_rmat(scale, 16, 0.57, 0.19, 0.19, seed=seed)
    vl, el = g.csr()
    assert np.array_equal(vl, ref.vlist) and np.array_equal(el, ref.elist)
    check_all(ctx, g, ref, orc, P=64)


def test_config1_scale16_64_partitions(ctx, orc):
    """Plain method, small: R-MAT scale 16, ef 16, (0.57,0.19,0.19,0.05), 64 contiguous
    partitions: T and per-vertex LCC bit-exact against the oracle, and the partitioned
    oracle (Partitioned_Graph, Graph.cpp:1346-1362) agrees."""
    g = pac.Graph.rmat(ctx, 16, 16, 0.57, 0.19, 0.19, seed=16)
    ref = orc.gen_rmat(16, 16, 0.57, 0.19, 0.19, seed=16)
    vl, el = g.csr()
    assert np.array_equal(vl, ref.vlist) and np.array_equal(el, ref.elist)
    v2p = orc.contiguous_partition(ref.V, 64)
    pg = orc.PGraph(ref, v2p, NT)
    T = pg.tc_merge()
    assert T == orc.tc_merge(ref, NT)
    g.prepare(64, v2p)
    lref = orc.lcc_merge(ref, NT)
    for d in STRATS:
        for l in (pac.LAUNCH_PLAIN, pac.LAUNCH_PERSISTENT):
            l32, l64, t = g.lcc(d, l)
            assert t == T
            assert bits_equal(l64, lref) and bits_equal(l32, lref.astype(np.float32))
    assert pac.PartitionedGraph(g, v2p).solve_triangle_counting_merge() == T
    assert pac.PartitionedGraph(g, v2p).solve_triangle_counting_galloping() == 6 * T


@pytest.mark.parametrize("n,k,seed", [(500, 8.0, 1), (3000, 40.0, 2), (20000, 12.0, 3)])
def test_erdos_renyi(ctx, orc, n, k, seed):
    g = pac.Graph.er(ctx, n, k, seed=seed)
    ref = orc.gen_er(n, k, seed=seed)
    vl, el = g.csr()
    assert np.array_equal(vl, ref.vlist) and np.array_equal(el, ref.elist)
    check_all(ctx, g, ref, orc, P=7)


@pytest.mark.parametrize("P", [1, 2, 5, 64, 1000])
def test_partition_count_and_map_invariance(ctx, orc, P):
    ref = orc.gen_rmat(11, 16, 0.6, 0.15, 0.15, seed=P)
    g = pac.Graph.from_csr(ctx, ref.vlist, ref.elist)
    T = orc.tc_merge(ref, NT)
    maps = [None, orc.random_partition(ref.V, P, seed=1)]
    if P <= 64:
        maps.append(g.greedy_partition(P, seed=2))
    for v2p in maps:
        g.prepare(P, v2p)
        for d in STRATS:
            assert g.tc(d, pac.LAUNCH_PERSISTENT) == T
            assert g.tc(d, pac.LAUNCH_PLAIN) == T
        sizes = g.partition_sizes()
        assert int(sizes.sum()) == g.info().num_edges


@pytest.mark.parametrize("tile", range(pac.NUM_TILES))
def test_rows_above_lds_capacity(ctx, orc, tile):
    """K_1100 has out-degrees up to 1099, above the 1024-entry LDS list of tile 0: those
    rows take the HBM lookup path."""
    n = 1100
    ref = orc.edges_to_csr(n, *zip(*itertools.combinations(range(n), 2)))
    g = pac.Graph.from_csr(ctx, ref.vlist, ref.elist)
    ctx.set_tile(tile)
    try:
        g.prepare(4)
        for d in STRATS:
            for l in LAUNCHES:
                t, pv = g.tc(d, l, per_vertex=True)
                assert t == math.comb(n, 3)
                assert np.all(pv == math.comb(n - 1, 2))
        l32, l64, _ = g.lcc()
        assert np.all(l64 == 1.0)
        src, dst, inter, uni, ratio = g.jaccard_edges(pac.FORCE_BITMAP)
        assert np.all(inter == n - 2) and np.all(uni == n)
    finally:
        ctx.set_tile(1)


@pytest.mark.parametrize("launch", [pac.LAUNCH_PERSISTENT, pac.LAUNCH_GRAPH])
def test_long_rows_deferred_to_wide_pass(ctx, orc, launch):
    """With the 1024-entry tile, the out-lists of a 1100-clique inside a sparse random
    graph exceed the LDS table: the persistent kernel queues those rows (whole or split
    by pair range) for the second pass with the 4096-entry tile.  TC, per-vertex counts,
    LCC and per-edge supports equal the oracle's."""
    rng = np.random.default_rng(5)
    n, V = 1100, 4000
    cs, cd = zip(*itertools.combinations(range(n), 2))
    rs = rng.integers(0, V, 60000)
    rd = rng.integers(0, V, 60000)
    ref = orc.edges_to_csr(V, np.concatenate([cs, rs]).astype(np.uint32),
                           np.concatenate([cd, rd]).astype(np.uint32))
    T = orc.tc_merge(ref, NT)
    lref = orc.lcc_merge(ref, NT)
    g = pac.Graph.from_csr(ctx, ref.vlist, ref.elist)
    ctx.set_tile(0)
    try:
        g.prepare(8)
        for d in (pac.FORCE_BITMAP, pac.FORCE_GALLOP, pac.DISPATCH_FIXED):
            t, pv = g.tc(d, launch, per_vertex=True)
            assert t == T and int(pv.sum()) == 3 * T
            l32, l64, t2 = g.lcc(d, launch)
            assert t2 == T and bits_equal(l64, lref)
            src, dst, inter, uni, ratio = g.jaccard_edges(d, launch)
            ri, ru, rr = orc.jaccard_edges(ref, src, dst, NT)
            assert np.array_equal(inter, ri) and bits_equal(ratio, rr)
        # batched with graphs whose rows all fit: the queue holds the long rows of one graph
        smalls = [orc.gen_rmat(10, 8, 0.57, 0.19, 0.19, seed=s) for s in (3, 4)]
        g2, g3 = (pac.Graph.from_csr(ctx, x.vlist, x.elist) for x in smalls)
        g2.prepare(4)
        g3.prepare(2)
        got = ctx.tc_batch([g2, g, g3], pac.DISPATCH_FIXED, launch)
        assert got.tolist() == [orc.tc_merge(smalls[0], NT), T, orc.tc_merge(smalls[1], NT)]
        with pytest.raises(pac.PacError):
            ctx.tc_batch([g2, g, g2], pac.DISPATCH_FIXED, launch)  # a graph listed twice
    finally:
        ctx.set_tile(1)


def test_degenerate_graphs(ctx, orc):
    empty = pac.Graph.from_csr(ctx, np.zeros(6, np.uint64), np.zeros(0, np.int32))
    empty.prepare(2)
    for l in LAUNCHES:
        assert empty.tc(pac.DISPATCH_FIXED, l) == 0
    l32, l64, t = empty.lcc()
    assert t == 0 and np.all(l64 == 0.0)
    one = orc.edges_to_csr(4, [0], [1])
    g = pac.Graph.from_csr(ctx, one.vlist, one.elist)
    g.prepare(3)
    l32, l64, t = g.lcc()
    assert t == 0 and np.isnan(l64[0]) and np.isnan(l64[1]) and l64[2] == 0.0
    assert bits_equal(l64, orc.lcc_merge(one, 1))
    src, dst, inter, uni, ratio = g.jaccard_edges()
    assert inter.tolist() == [0] and uni.tolist() == [2] and ratio.tolist() == [0.0]


# The hash multipliers of the LDS tables, in the order a row's builds try them, and the
# two buckets of key x among 2^lg under one of them: a replica of hash_mul() and buckets()
# in pac_isect.hip (h = (x mod 2^24) * M + x; b1 = the top lg bits of h, b2 = b1 xor (bits
# [3, 3 + lg) of h, made odd)), used only to build inputs whose tables cannot be filled.
HASH_MULS = (0x9E3778, 0xC2B2AE, 0x85EBCA, 0x27D4EA)


def _cuckoo_buckets(x, lg, mul):
    h = ((x & 0xFFFFFF) * mul + x) & 0xFFFFFFFF
    b1 = h >> (32 - lg)
    return b1, b1 ^ (((h >> 3) & ((1 << lg) - 1)) | 1)


def _colliding(lg, mul, count, lo, taken):
    """count IDs from lo up, none in taken, sharing both buckets under mul: more than four
    such keys cannot all be placed (the two buckets hold four), whatever the order."""
    by = {}
    x = lo
    while True:
        if x not in taken:
            b = _cuckoo_buckets(x, lg, mul)
            by.setdefault(b, []).append(x)
            if len(by[b]) == count:
                return by[b]
        x += 1


def _owner_row_graph(rows, V):
    """Per (h, K, q): vertex h stages N+(h) = K (a clique) and owns q pairs, one per
    in-neighbour t (adjacent to h and to all of K, d(t) = |K| + 1 < d(h) = |K| + q), each
    streaming all of K (the suffix of N+(t) after h), so row h's table must hold K."""
    src, dst = [], []
    t_next = V - 1
    for h, K, q in rows:
        for i, a in enumerate(K):
            src.append(h)
            dst.append(a)
            for b in K[i + 1:]:
                src.append(a)
                dst.append(b)
        for _ in range(q):
            t = t_next
            t_next -= 1
            src.append(t)
            dst.append(h)
            for a in K:
                src.append(t)
                dst.append(a)
    return src, dst


def _check_hash_rows(ctx, orc, src, dst, V):
    ref = orc.edges_to_csr(V, src, dst)
    g = pac.Graph.from_csr(ctx, ref.vlist, ref.elist)
    T = check_all(ctx, g, ref, orc, P=4, strategies=[pac.FORCE_BITMAP, pac.DISPATCH_FIXED])
    for l in LAUNCHES:
        l32, l64, t = g.lcc(pac.FORCE_BITMAP, l)
        assert t == T and bits_equal(l64, orc.lcc_merge(ref, NT))
        src2, dst2, inter, uni, ratio = g.jaccard_edges(pac.FORCE_BITMAP, l)
        ri, ru, rr = orc.jaccard_edges(ref, src2, dst2, NT)
        assert np.array_equal(inter, ri) and np.array_equal(uni, ru) and bits_equal(ratio, rr)
    return T


def _rows_expected_T(rows):
    # per row: the clique, h with it, and each in-neighbour with h and the clique
    return sum(math.comb(len(K), 3) + (q + 1) * math.comb(len(K), 2) + q * len(K)
               for _, K, q in rows)


def test_hash_table_rebuilt_with_the_next_multiplier(ctx, orc):
    """Rows whose staged lists hold six keys sharing both buckets under the first hash
    multiplier: the first build always leaves keys homeless, and the row is built again
    with the second multiplier, under which (a sequential replay says) the list fits.  Rows
    of 3 pairs go one per wave, rows of 80 and 100 pairs to the whole workgroup."""
    V = 3 << 19
    rng = np.random.default_rng(7)
    rows, taken = [], set()
    for h, n, q, lo in ((5, 16, 3, 1000), (7, 40, 80, 1 << 19), (9, 80, 100, 1 << 20)):
        lg = max(4, (n - 1).bit_length())
        bad = _colliding(lg, HASH_MULS[0], 6, lo, taken)
        pool = np.setdiff1d(np.arange(lo, lo + (1 << 18)), np.array(sorted(taken | set(bad))))
        K = sorted(bad + rng.choice(pool, n - 6, replace=False).tolist())
        assert _replay_build(K, lg, HASH_MULS[0]) is False
        assert _replay_build(K, lg, HASH_MULS[1]) is True
        taken |= set(K)
        rows.append((h, K, q))
    src, dst = _owner_row_graph(rows, V)
    assert _check_hash_rows(ctx, orc, src, dst, V) == _rows_expected_T(rows)


def test_hash_table_build_failure_falls_back(ctx, orc):
    """Rows whose staged lists hold, for each of the four hash multipliers, five keys sharing
    both buckets under it: every build leaves a key homeless and the row is counted by a
    fallback (HBM lookups by the wave for a row of 3 pairs, the sorted-list path for rows of
    80 and 100 pairs).  Counts, LCC and Jaccard must still equal the oracle's."""
    V = 3 << 19
    rng = np.random.default_rng(11)
    rows, taken = [], set()
    for h, n, q, lo in ((5, 32, 3, 1000), (7, 40, 80, 1 << 19), (9, 80, 100, 1 << 20)):
        lg = max(4, (n - 1).bit_length())
        bad = []
        for mul in HASH_MULS:
            grp = _colliding(lg, mul, 5, lo, taken | set(bad))
            bad += grp
        pool = np.setdiff1d(np.arange(lo, lo + (1 << 18)), np.array(sorted(taken | set(bad))))
        K = sorted(bad + rng.choice(pool, n - len(bad), replace=False).tolist())
        assert len(set(K)) == n
        for mul in HASH_MULS:
            assert _replay_build(K, lg, mul) is False
        taken |= set(K)
        rows.append((h, K, q))
    src, dst = _owner_row_graph(rows, V)
    assert _check_hash_rows(ctx, orc, src, dst, V) == _rows_expected_T(rows)


def _replay_build(K, lg, mul, max_kicks=48):
    """A sequential replay of row_hash's table build: True when every key finds a slot."""
    nb = 1 << lg
    tab = [None] * (2 * nb)
    for x in K:
        frm = None
        for kick in range(max_kicks + 1):
            b1, b2 = _cuckoo_buckets(x, lg, mul)
            placed = False
            for s_ in (2 * b1, 2 * b1 + 1, 2 * b2, 2 * b2 + 1):
                if tab[s_] is None:
                    tab[s_] = x
                    placed = True
                    break
            if placed:
                break
            if kick == max_kicks:
                return False
            vb = b2 if frm == b1 else b1
            slot = 2 * vb + ((kick >> 1) & 1)
            tab[slot], x = x, tab[slot]
            frm = vb
    return True


def test_jaccard_edges_scale13(ctx, orc):
    ref = orc.gen_rmat(13, 16, 0.57, 0.19, 0.19, seed=13)
    g = pac.Graph.from_csr(ctx, ref.vlist, ref.elist)
    g.prepare(32)
    T = orc.tc_merge(ref, NT)
    for d in STRATS:
        for l in (pac.LAUNCH_PLAIN, pac.LAUNCH_PERSISTENT, pac.LAUNCH_GRAPH):
            src, dst, inter, uni, ratio = g.jaccard_edges(d, l)
            ri, ru, rr = orc.jaccard_edges(ref, src, dst, NT)
            assert np.array_equal(inter, ri) and np.array_equal(uni, ru)
            assert bits_equal(ratio, rr)
            assert int(inter.sum(dtype=np.uint64)) == 3 * T
    # every undirected edge appears once
    key = np.minimum(src, dst).astype(np.int64) * ref.V + np.maximum(src, dst)
    assert len(np.unique(key)) == ref.nnz // 2


def test_jaccard_edges_without_tails(ctx, orc):
    """A solve asked for no tails (src) takes the elementwise epilogue over the per-edge
    degree sums (as a solve that keeps its results on the device does): its heads,
    intersections, unions and ratios equal those of the solve with tails, and the oracle's."""
    ref = orc.gen_rmat(13, 16, 0.57, 0.19, 0.19, seed=131)
    g = pac.Graph.from_csr(ctx, ref.vlist, ref.elist)
    g.prepare(32)
    src, dst, inter, uni, ratio = g.jaccard_edges(pac.DISPATCH_FIXED, pac.LAUNCH_GRAPH)
    ri, ru, rr = orc.jaccard_edges(ref, src, dst, NT)
    m = len(src)
    dst2 = np.zeros(m, np.int32)
    inter2 = np.zeros(m, np.uint32)
    uni2 = np.zeros(m, np.uint32)
    ratio2 = np.zeros(m, np.float32)
    for l in LAUNCHES:
        rc = pac.lib().pac_jaccard_edges(g._h, pac.DISPATCH_FIXED, l, None, dst2.ctypes.data,
                                         inter2.ctypes.data, uni2.ctypes.data,
                                         ratio2.ctypes.data)
        assert rc == 0, l
        assert np.array_equal(dst2, dst) and np.array_equal(inter2, ri)
        assert np.array_equal(uni2, ru) and bits_equal(ratio2, rr)
    g.jaccard_edges_device(pac.DISPATCH_FIXED, pac.LAUNCH_GRAPH)  # device-only: no copies


def test_jaccard_strategy_switches(ctx, orc):
    """Per-edge supports stay exact when consecutive solves of one graph switch between the
    hash path (hit words of the streamed lists, summed after the launch) and the sorted
    paths (direct credits, hit words cleared), including vertices streamed by more than
    512 pairs (the sliced column sums)."""
    ref = orc.gen_rmat(14, 16, 0.57, 0.19, 0.19, seed=141)
    g = pac.Graph.from_csr(ctx, ref.vlist, ref.elist)
    g.prepare(16)
    want = None
    for d in (pac.FORCE_BITMAP, pac.FORCE_MERGE, pac.FORCE_BITMAP, pac.FORCE_GALLOP,
              pac.DISPATCH_FIXED, pac.FORCE_BITMAP):
        src, dst, inter, uni, ratio = g.jaccard_edges(d, pac.LAUNCH_PERSISTENT)
        if want is None:
            want = orc.jaccard_edges(ref, src, dst, NT)
        assert np.array_equal(inter, want[0]) and np.array_equal(uni, want[1])
        assert bits_equal(ratio, want[2])


def test_query_matrix_and_pairs(ctx, orc):
    ref = orc.gen_rmat(12, 16, 0.57, 0.19, 0.19, seed=5)
    g = pac.Graph.from_csr(ctx, ref.vlist, ref.elist)
    rng = np.random.default_rng(1)
    q = np.concatenate([[0, 1, 2], rng.integers(0, ref.V, 60)]).astype(np.int32)
    assert bits_equal(g.solve_jsim_merge(q), orc.jsim_merge(ref, q, NT).ravel())
    u = rng.integers(0, ref.V, 5000).astype(np.int32)
    v = rng.integers(0, ref.V, 5000).astype(np.int32)
    got = g.cnt_common_pairs(u, v)
    want = [orc.cnt_merge(ref, int(a), int(b)) for a, b in zip(u, v)]
    assert got.tolist() == want


def test_batch_and_graph_replays(ctx, orc):
    refs, gs = [], []
    for i in range(12):
        if i % 2:
            r = orc.gen_rmat(8 + i % 5, 16, 0.57, 0.19, 0.19, seed=100 + i)
        else:
            r = orc.gen_er(500 + 300 * i, 8 + 4 * i, seed=100 + i)
        refs.append(r)
        g = pac.Graph.from_csr(ctx, r.vlist, r.elist)
        g.prepare(1 + i * 3)
        gs.append(g)
    want = np.array([orc.tc_merge(r, NT) for r in refs], np.uint64)
    for d in STRATS:
        assert np.array_equal(ctx.tc_batch(gs, d, pac.LAUNCH_PERSISTENT), want)
    assert np.array_equal(ctx.tc_batch(gs, pac.DISPATCH_FIXED, pac.LAUNCH_PLAIN), want)
    # captured once, replayed several times with other work in between
    for rep in range(4):
        got = ctx.tc_batch(gs, pac.DISPATCH_FIXED, pac.LAUNCH_GRAPH)
        assert np.array_equal(got, want), rep
        assert gs[rep].tc(pac.FORCE_MERGE, pac.LAUNCH_PERSISTENT) == want[rep]


def test_csr_files_round_trip(ctx, orc, tmp_path):
    ref = orc.gen_rmat(10, 8, 0.57, 0.19, 0.19, seed=8)
    g = pac.Graph.from_csr(ctx, ref.vlist, ref.elist)
    g.save_csr_dir(str(tmp_path))
    assert np.array_equal(np.fromfile(tmp_path / "csr_vlist.bin", np.uint64), ref.vlist)
    assert np.array_equal(np.fromfile(tmp_path / "csr_elist.bin", np.int32), ref.elist)
    g2 = pac.Graph.load_csr_dir(ctx, str(tmp_path))
    vl, el = g2.csr()
    assert np.array_equal(vl, ref.vlist) and np.array_equal(el, ref.elist)


def test_text_loader_matches_oracle(ctx, orc):
    exp = json.load(open(os.path.join(GOLDEN, "text_expected.json")))
    for f in exp:
        if f.startswith("_"):
            continue
        g = pac.Graph.load_text(ctx, os.path.join(GOLDEN, f))
        vl, el = g.csr()
        assert vl.tolist() == exp[f]["vlist"] and el.tolist() == exp[f]["elist"]


@pytest.mark.parametrize("one_based", [False, True])
def test_text_loader_device_parse(ctx, orc, tmp_path, one_based):
    """The device parser (one thread per line) on a 60 000-line file with the reference's
    quirks -- comments, tab / double-space / CRLF separators, trailing text, self-loops,
    a last line without newline -- and the host fallback for a line over 511 characters and
    for a file ending in a bare number; every CSR equal to the oracle's restatement of
    Graph.cpp:27-90."""
    rng = np.random.default_rng(11 + one_based)
    lo = 1 if one_based else 0
    lines = []
    for _ in range(60000):
        u, v = (int(x) for x in rng.integers(lo, 5000, 2))
        r = rng.random()
        if r < 0.03:
            lines.append("# comment %d %d\n" % (u, v))
        elif r < 0.05:
            lines.append("%% meta %d\n" % u)
        elif r < 0.08:
            lines.append("%d\t%d\r\n" % (u, v))
        elif r < 0.09 and not one_based:
            lines.append("%d  %d\n" % (u, v))       # second number read as 0
        elif r < 0.11:
            lines.append("%d %d extra 77\n" % (u, v))
        elif r < 0.13:
            lines.append("%d %d\n" % (u, u))        # self-loop
        else:
            lines.append("%d %d\n" % (u, v))
    body = "".join(lines)
    cases = {"plain.txt": body + "%d %d" % (lo + 3, lo + 4),          # no final newline
             "longline.txt": body + "1 2 " + "x" * 600 + "\n",       # host fallback
             "bare.txt": body + "%d" % (lo + 9)}                      # host fallback
    for name, text in cases.items():
        path = tmp_path / name
        path.write_text(text)
        ref = orc.text_to_csr(str(path))
        g = pac.Graph.load_text(ctx, str(path))
        vl, el = g.csr()
        assert np.array_equal(vl, ref.vlist) and np.array_equal(el, ref.elist), name


def test_error_paths(ctx, orc, tmp_path):
    ref = orc.gen_rmat(6, 4, 0.57, 0.19, 0.19, seed=1)
    g = pac.Graph.from_csr(ctx, ref.vlist, ref.elist)
    with pytest.raises(pac.PacError) as e:
        pac.pac._check(pac.lib().pac_tc(g._h, 0, 0, None, None))
    assert e.value.code == -6                                 # not prepared
    g.prepare(2)
    with pytest.raises(pac.PacError) as e:
        g.tc(dispatch=9)
    assert e.value.code == -1
    bad = ref.elist.copy()
    bad[0] = ref.V + 5                                        # ID out of range
    with pytest.raises(pac.PacError) as e:
        pac.Graph.from_csr(ctx, ref.vlist, bad)
    assert e.value.code == -4
    unsorted = ref.elist.copy()
    a, b = int(ref.vlist[0]), int(ref.vlist[1])
    if b - a >= 2:
        unsorted[a], unsorted[a + 1] = unsorted[a + 1], unsorted[a]
        with pytest.raises(pac.PacError) as e:
            pac.Graph.from_csr(ctx, ref.vlist, unsorted)
        assert e.value.code == -4
    vl = ref.vlist.copy()
    vl[3] = vl[4] + 1                                         # non-monotone offsets
    with pytest.raises(pac.PacError) as e:
        pac.Graph.from_csr(ctx, vl, ref.elist)
    assert e.value.code == -4
    with pytest.raises(pac.PacError) as e:
        pac.Graph.load_csr_dir(ctx, str(tmp_path / "missing"))
    assert e.value.code == -5
    (tmp_path / "csr_vlist.bin").write_bytes(b"\x00" * 12)    # not a multiple of 8
    (tmp_path / "csr_elist.bin").write_bytes(b"")
    with pytest.raises(pac.PacError) as e:
        pac.Graph.load_csr_dir(ctx, str(tmp_path))
    assert e.value.code == -5
    # asymmetric input is caught by the orientation count
    asym_vl = np.array([0, 1, 1, 1], np.uint64)
    with pytest.raises(pac.PacError) as e:
        pac.Graph.from_csr(ctx, asym_vl, np.array([2], np.int32))
    assert e.value.code == -4
    with pytest.raises(pac.PacError):
        g.prepare(4, np.full(ref.V, -1, np.int32))


def test_generator_properties_large(ctx, orc):
    """Scale 20: size-independent properties (structure, sums) instead of the oracle."""
    g = pac.Graph.rmat(ctx, 20, 16, 0.57, 0.19, 0.19, seed=20)
    g.validate(1)
    g.prepare(1024)
    T = g.tc(pac.FORCE_BITMAP)
    for d in (pac.FORCE_MERGE, pac.FORCE_GALLOP, pac.DISPATCH_FIXED):
        assert g.tc(d) == T
    t2, pv = g.tc(pac.DISPATCH_FIXED, pac.LAUNCH_PLAIN, per_vertex=True)
    assert t2 == T and int(pv.sum()) == 3 * T
    src, dst, inter, uni, ratio = g.jaccard_edges()
    assert int(inter.sum(dtype=np.uint64)) == 3 * T
    # spot-check a sample of rows against the oracle on the downloaded CSR
    vl, el = g.csr()
    ref = orc.CSR(vl, el)
    rows = np.random.default_rng(0).choice(ref.V, 2000, replace=False).astype(np.int32)
    tv = orc.lcc_merge_rows(ref, rows, NT)[1]
    l32, l64, _ = g.lcc()
    assert bits_equal(l64[rows], tv)


@pytest.mark.parametrize("launch", [pac.LAUNCH_GRAPH, pac.LAUNCH_PERSISTENT, pac.LAUNCH_PLAIN])
def test_first_solve_of_each_mode_on_a_fresh_graph(ctx, orc, launch):
    """Regression: the first LCC / SUP solve of a graph must allocate its per-vertex and
    per-edge outputs before the kernel descriptors that point at them are built (the
    hipGraph path used to capture a descriptor with an unallocated output)."""
    ref = orc.gen_rmat(12, 16, 0.57, 0.19, 0.19, seed=121)
    T = orc.tc_merge(ref, NT)
    lref = orc.lcc_merge(ref, NT)
    g = pac.Graph.from_csr(ctx, ref.vlist, ref.elist)
    g.prepare(8)
    assert g.lcc_device(pac.DISPATCH_FIXED, launch) == T
    l32, l64, t = g.lcc(pac.DISPATCH_FIXED, launch)
    assert t == T and bits_equal(l64, lref)
    g2 = pac.Graph.from_csr(ctx, ref.vlist, ref.elist)
    g2.prepare(8)
    src, dst, inter, uni, ratio = g2.jaccard_edges(pac.DISPATCH_FIXED, launch)
    ri, ru, rr = orc.jaccard_edges(ref, src, dst, NT)
    assert np.array_equal(inter, ri) and bits_equal(ratio, rr)
    g3 = pac.Graph.from_csr(ctx, ref.vlist, ref.elist)
    g3.prepare(8)
    g3.jaccard_edges_device(pac.FORCE_BITMAP, launch)
    t3, pv = g3.tc(pac.FORCE_MERGE, launch, per_vertex=True)
    assert t3 == T and int(pv.sum()) == 3 * T


def test_edge_list_ingestion_on_device(ctx, orc):
    """Device symmetrize / sort / dedup of a raw edge list with duplicates, reversed
    duplicates and self-loops equals the host construction (Graph.cpp:70-85)."""
    rng = np.random.default_rng(9)
    V = 5000
    src = rng.integers(0, V, 60000).astype(np.int32)
    dst = (src + rng.integers(0, 40, 60000)) % V          # includes self-loops (offset 0)
    src = np.concatenate([src, dst[:5000], src[:3000]]).astype(np.int32)
    dst = np.concatenate([dst, src[:5000], dst[:3000]]).astype(np.int32)
    ref = orc.edges_to_csr(V, src.astype(np.uint32), dst.astype(np.uint32))
    g = pac.Graph.from_edges(ctx, src, dst, V)
    vl, el = g.csr()
    assert np.array_equal(vl, ref.vlist) and np.array_equal(el, ref.elist)
    g.prepare(4)
    assert g.tc() == orc.tc_merge(ref, NT)
    g2 = pac.Graph.from_edges(ctx, src, dst)                  # V inferred = max ID + 1
    assert g2.get_vertex_num() == int(max(src.max(), dst.max())) + 1
    with pytest.raises(pac.PacError) as e:
        pac.Graph.from_edges(ctx, src, dst, 10)
    assert e.value.code == -4
    empty = pac.Graph.from_edges(ctx, np.zeros(0, np.int32), np.zeros(0, np.int32), 7)
    assert empty.get_vertex_num() == 7 and empty.get_edge_num() == 0


def test_evaluate_partition_matches_oracle(ctx, orc):
    """pac_evaluate_partition against the oracle's Partitioned_Graph::evaluate_partition
    (Graph.cpp:1460-1473) for contiguous, random and greedy maps, and NaN without edges."""
    ref = orc.gen_rmat(12, 16, 0.57, 0.19, 0.19, seed=77)
    g = pac.Graph.from_csr(ctx, ref.vlist, ref.elist)
    for P in (1, 7, 64):
        for v2p in (orc.contiguous_partition(ref.V, P), orc.random_partition(ref.V, P, seed=P),
                    g.greedy_partition(P, seed=P)):
            want = orc.PGraph(ref, v2p).evaluate_partition()[0]
            assert g.evaluate_partition(v2p, int(v2p.max()) + 1) == want, P
    empty = pac.Graph.from_csr(ctx, np.zeros(5, np.uint64), np.zeros(0, np.int32))
    assert math.isnan(empty.evaluate_partition(np.zeros(4, np.int32), 1))


def test_solve_timeout_forced_from_the_host(ctx, orc):
    """PAC_ERR_TIMEOUT: with a deadline far shorter than the solve, the host stops waiting
    while the kernels run to their normal end; the context stays usable (the next solve
    waits behind them) and a generous deadline changes nothing."""
    ref = orc.gen_rmat(16, 16, 0.57, 0.19, 0.19, seed=61)
    T = orc.tc_oriented(ref, NT)
    g = pac.Graph.from_csr(ctx, ref.vlist, ref.elist)
    g.prepare(64)
    try:
        ctx.set_timeout(1e-6)
        with pytest.raises(pac.PacError) as e:
            g.lcc(pac.FORCE_MERGE, pac.LAUNCH_PLAIN)
        assert e.value.code == -8
        ctx.set_timeout(60000.0)
        assert g.tc(pac.FORCE_BITMAP, pac.LAUNCH_PERSISTENT) == T
        with pytest.raises(pac.PacError) as e:
            ctx.set_timeout(-1.0)
        assert e.value.code == -1
    finally:
        ctx.set_timeout(0.0)
    assert g.lcc_device(pac.DISPATCH_FIXED, pac.LAUNCH_GRAPH) == T


def test_greedy_partition_2_matches_oracle(ctx, orc):
    """pac_greedy_partition_2 against the oracle's restatement of Graph.cpp:969-1015 (same
    tie order and seeded draws), with isolated vertices present, and the partitioned count
    under its map."""
    ref = orc.gen_rmat(10, 8, 0.57, 0.19, 0.19, seed=31)
    g = pac.Graph.from_csr(ctx, ref.vlist, ref.elist)
    assert (ref.deg() == 0).any()
    T = orc.tc_merge(ref, NT)
    for P in (1, 2, 7, 32):
        got = g.greedy_partition_2(P, seed=P)
        assert np.array_equal(got, orc.greedy_partition_2(ref, P, seed=P)), P
        g.prepare(P, got)
        assert g.tc(pac.DISPATCH_FIXED) == T


# 3-byte IDs with every byte pattern that decoding could get wrong (top byte 0xFF, zero
# middle bytes, bytes that straddle a 4-byte word in the packed copy)
WIDE_IDS = [0, 1, 0xFF, 0x100, 0xFFFF, 0x10000, 0x00FFFF, 0xFF00FF, 0x7FFFFF, 0x800000,
            0xFFFF00, 0xFFFFFE, 0xFFFFFF]


def _wide_id_edges(V, seed):
    """Edges among 900 IDs spread over [0, V), the largest three and WIDE_IDS among them:
    G(900, 0.05), a 600-vertex clique (out-lists of up to 599 elements, so streamed lists
    span several 256-element blocks) and three hubs adjacent to all."""
    rng = np.random.default_rng(seed)
    ids = np.unique(np.concatenate([rng.choice(V - 3, 900 - len(WIDE_IDS) - 3, replace=False),
                                    [x for x in WIDE_IDS if x < V], [V - 3, V - 2, V - 1]]))
    n = len(ids)
    a = rng.random((n, n)) < 0.05
    cl = rng.choice(n, 600, replace=False)
    a[np.ix_(cl, cl)] = True
    a[rng.choice(n, 3, replace=False), :] = True
    iu = np.triu_indices(n, 1)
    sel = a[iu] | a.T[iu]
    return ids[iu[0][sel]].astype(np.int32), ids[iu[1][sel]].astype(np.int32)


@pytest.mark.parametrize("V", [1 << 24, (1 << 24) + 1])
def test_packed_lists_wide_ids(ctx, orc, V):
    """TC and LCC stream 3-byte lists when every ID is below 2^24 (V = 2^24: IDs up to
    0xFFFFFF) and packing is forced on, and 4-byte lists otherwise (V = 2^24 + 1, packing
    off, or, in the default mode, a graph of fewer than 2^24 edges): per-vertex counts on
    every launch path equal the oracle's in every mode; per-edge Jaccard (always 4-byte
    lists) equals the oracle's on every edge either way."""
    src, dst = _wide_id_edges(V, seed=V & 0xFF)
    g = pac.Graph.from_edges(ctx, src, dst, V)
    ref = orc.CSR(*g.csr())
    T = orc.tc_merge(ref, NT)
    tv = orc.per_vertex_triangles(ref, NT)
    g.prepare(8)
    res = {}
    try:
        for mode in (pac.PACKED_ALWAYS, pac.PACKED_OFF, pac.PACKED_AUTO):
            ctx.set_packed_lists(mode)
            want = 3 if mode == pac.PACKED_ALWAYS and V <= 1 << 24 else 4
            assert g.info().stream_elem_bytes == want, mode
            for d in (pac.FORCE_BITMAP, pac.DISPATCH_FIXED):
                for l in LAUNCHES:
                    t, pv = g.tc(d, l, per_vertex=True)
                    assert t == T and np.array_equal(pv, tv), (mode, d, l)
            res[mode] = g.jaccard_edges(pac.FORCE_BITMAP, pac.LAUNCH_GRAPH)
    finally:
        ctx.set_packed_lists(pac.PACKED_AUTO)
    for a, b in zip(res[pac.PACKED_ALWAYS], res[pac.PACKED_OFF]):
        assert bits_equal(a, b)
    es, ed, inter, uni, ratio = res[pac.PACKED_ALWAYS]
    ri, ru, rr = orc.jaccard_edges(ref, es, ed, NT)
    assert np.array_equal(ri, inter) and np.array_equal(ru, uni) and bits_equal(rr, ratio)
    assert int(inter.sum(dtype=np.uint64)) == 3 * T
    with pytest.raises(pac.PacError):
        ctx.set_packed_lists(3)
    g.free()


def test_packed_lists_rmat_all_modes(ctx, orc):
    """R-MAT scale 15: TC and LCC from the 3-byte and the 4-byte streamed lists are
    identical, under fixed and forced-hash dispatch, on the persistent (one kernel per list
    width) and plain (both widths in one kernel) launch paths, and per-edge supports do not
    change with the setting."""
    g = pac.Graph.rmat(ctx, 15, 16, 0.57, 0.19, 0.19, seed=15)
    g.prepare(32)
    out = {}
    try:
        for mode in (pac.PACKED_ALWAYS, pac.PACKED_OFF):
            ctx.set_packed_lists(mode)
            r = []
            for d in (pac.FORCE_BITMAP, pac.DISPATCH_FIXED):
                r.append(g.tc(d, pac.LAUNCH_PERSISTENT, per_vertex=True)[1])
                r.append(g.tc(d, pac.LAUNCH_PLAIN, per_vertex=True)[1])
                r.extend(g.lcc(d, pac.LAUNCH_GRAPH)[:2])
                r.extend(g.jaccard_edges(d, pac.LAUNCH_PERSISTENT)[2:])
            out[mode] = r
    finally:
        ctx.set_packed_lists(pac.PACKED_AUTO)
    for a, b in zip(out[pac.PACKED_ALWAYS], out[pac.PACKED_OFF]):
        assert bits_equal(a, b)
    T = orc.tc_merge(orc.CSR(*g.csr()), NT)
    assert int(out[pac.PACKED_ALWAYS][0].sum()) == 3 * T


def test_packed_lists_batch_of_mixed_graphs(ctx, orc):
    """A batch whose graphs all have 3-byte lists streams them (forced on); one graph with
    IDs beyond 2^24 makes the whole launch stream 4-byte lists: both give every graph's
    oracle count."""
    gs = [pac.Graph.rmat(ctx, s, 16, 0.57, 0.19, 0.19, seed=100 + s) for s in (10, 12, 14)]
    src, dst = _wide_id_edges((1 << 24) + 1, seed=7)
    gs.append(pac.Graph.from_edges(ctx, src, dst, (1 << 24) + 1))
    for g in gs:
        g.prepare(4)
    want = np.array([orc.tc_oriented(orc.CSR(*g.csr()), NT) for g in gs], np.uint64)
    try:
        ctx.set_packed_lists(pac.PACKED_ALWAYS)
        for l in (pac.LAUNCH_PERSISTENT, pac.LAUNCH_GRAPH):
            assert np.array_equal(ctx.tc_batch(gs[:3], pac.FORCE_BITMAP, l), want[:3])
            assert np.array_equal(ctx.tc_batch(gs, pac.FORCE_BITMAP, l), want)
    finally:
        ctx.set_packed_lists(pac.PACKED_AUTO)
    for g in gs:
        g.free()
