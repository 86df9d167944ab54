"""GPU parity of the multi-way intersections (k-way common neighbours, friend triangles,
square counting) through the C ABI against the CPU oracle on the same inputs: every
count bit-exact."""
import math
import os

import numpy as np
import pytest

import partitionandconquer_amd as pac

pytestmark = pytest.mark.gpu
NT = min(16, os.cpu_count() or 1)


def gpu_graph(ctx, g):
    return pac.Graph.from_csr(ctx, g.vlist, g.elist)


def complete(n, orc):
    return orc.edges_to_csr(n, *zip(*[(i, j) for i in range(n) for j in range(i + 1, n)]))


def test_pair_counts_chunk_boundaries(ctx, orc):
    """Pair counts where both lists sit on either side of the 64-element chunks the
    wave merges (and of the merge/search choice): hubs 0..11 with neighbour sets of
    the listed sizes drawn from one range, so every pair overlaps; counts are the set
    intersections, also under thresholds, and equal the oracle's merge."""
    rng = np.random.default_rng(7)
    sizes = [0, 1, 63, 64, 65, 127, 128, 129, 191, 300, 640, 2000]
    base, span = 100, 2500
    sets = [np.sort(rng.choice(np.arange(base, base + span), n, replace=False)) for n in sizes]
    src = np.concatenate([np.full(len(S), i) for i, S in enumerate(sets)]).astype(np.int32)
    dst = np.concatenate(sets).astype(np.int32)
    V = base + span
    g = pac.Graph.from_edges(ctx, src, dst, V)
    ref = orc.edges_to_csr(V, src, dst)
    H = len(sizes)
    u = np.repeat(np.arange(H), H).astype(np.int32)
    v = np.tile(np.arange(H), H).astype(np.int32)
    want = [len(np.intersect1d(sets[a], sets[b])) for a, b in zip(u, v)]
    assert g.cnt_common_pairs(u, v).tolist() == want
    assert want == [orc.cnt_merge(ref, int(a), int(b)) for a, b in zip(u, v)]
    thr = rng.integers(base - 1, base + span + 1, len(u)).astype(np.int32)
    want_t = [int(np.sum(np.intersect1d(sets[a], sets[b]) < t)) for a, b, t in zip(u, v, thr)]
    assert g.cnt_common_pairs(u, v, thr=thr).tolist() == want_t


@pytest.mark.parametrize("scale,seed", [(8, 1), (12, 2), (14, 3)])
def test_kway_queries(ctx, orc, scale, seed):
    ref = orc.gen_rmat(scale, 16, seed=seed)
    g = gpu_graph(ctx, ref)
    rng = np.random.default_rng(seed)
    deg = np.diff(ref.vlist.astype(np.int64))
    hubs = np.argsort(-deg)[:64]
    queries = []
    for k in (1, 2, 3, 4, 5, 8, 17):
        for _ in range(60):
            pool = hubs if rng.random() < 0.6 else np.arange(ref.V)
            queries.append(rng.choice(pool, size=k).astype(np.int32))
    queries.append(np.array([hubs[0]] * 3, np.int32))          # a vertex repeated
    got = g.cnt_common_kway(queries)
    want = [orc.kway_merge(ref, q) for q in queries]
    assert got.tolist() == want
    assert [orc.kway_galloping(ref, q) for q in queries[:80]] == want[:80]
    pairs = [q for q in queries if len(q) == 2]                  # 2-way agrees with the pair API
    u = np.array([q[0] for q in pairs], np.int32)
    v = np.array([q[1] for q in pairs], np.int32)
    assert g.cnt_common_pairs(u, v).tolist() == g.cnt_common_kway(pairs).tolist()


def test_friend_triangles(ctx, orc):
    n = 12
    ref = complete(n, orc)
    g = gpu_graph(ctx, ref)
    assert g.friend_triangle(0, list(range(n))) == math.comb(n - 1, 2) * (n - 3)
    ref = orc.gen_rmat(13, 16, seed=7)
    g = gpu_graph(ctx, ref)
    rng = np.random.default_rng(7)
    deg = np.diff(ref.vlist.astype(np.int64))
    for root in (int(np.argmax(deg)), 5, int(np.argsort(deg)[len(deg) // 2])):
        nodes = np.concatenate([rng.choice(ref.V, size=150), np.argsort(-deg)[:40],
                                [root, root]]).astype(np.int32)
        assert g.friend_triangle(root, nodes) == orc.friend_triangle_merge(ref, root, nodes, NT)
    assert g.friend_triangle(3, []) == 0


def test_square_counting(ctx, orc):
    for n in (4, 6, 9):
        g = gpu_graph(ctx, complete(n, orc))
        assert g.square_count() == math.comb(n, 4)
    bip = orc.edges_to_csr(9, *zip(*[(i, j) for i in range(4) for j in range(4, 9)]))
    assert gpu_graph(ctx, bip).square_count() == 0
    for scale, seed in ((8, 1), (11, 2)):
        ref = orc.gen_rmat(scale, 8, seed=seed)
        want = orc.PGraph(ref, orc.contiguous_partition(ref.V, 4)).square_counting()
        assert gpu_graph(ctx, ref).square_count() == want
    er = orc.gen_er(3000, 20.0, seed=4)
    want = orc.PGraph(er, orc.random_partition(er.V, 7, 2)).square_counting()
    assert gpu_graph(ctx, er).square_count() == want


@pytest.mark.parametrize("P,seed", [(1, 1), (5, 2), (16, 3)])
def test_partition_filtered_queries(ctx, orc, P, seed):
    """Per-partition pair / k-way counts, the threshold variant and the common-neighbour
    lists (Graph.cpp:241-285, 468-493, 1071-1195, 1475-1498) against the oracle's
    partition-major restatement."""
    ref = orc.gen_rmat(11, 16, seed=seed)
    g = gpu_graph(ctx, ref)
    v2p = orc.random_partition(ref.V, P, seed) if P > 1 else np.zeros(ref.V, np.int32)
    pg = pac.PartitionedGraph(g, v2p)
    opg = orc.PGraph(ref, v2p)
    rng = np.random.default_rng(seed)
    deg = ref.deg()
    hubs = np.argsort(-deg)[:40]
    n = 300
    u = rng.choice(hubs, n).astype(np.int32)
    v = np.where(rng.random(n) < 0.5, rng.choice(hubs, n), rng.integers(0, ref.V, n)).astype(np.int32)
    # pairs of two vertices with at most 16 neighbours each, through a common neighbour where
    # one exists: the kernel merges such a pair in one lane, filters included
    low = np.nonzero((deg >= 1) & (deg <= 16))[0]
    lu, lv = [], []
    for a in rng.choice(low, 600).tolist():
        w = int(ref.elist[int(ref.vlist[a]) + int(rng.integers(0, deg[a]))])
        nb = ref.elist[int(ref.vlist[w]):int(ref.vlist[w + 1])]
        nb = nb[(deg[nb] <= 16) & (nb != a)]
        if len(nb):
            lu.append(a)
            lv.append(rng.choice(nb))
    assert len(lu) > 100
    u = np.concatenate([u, np.array(lu, np.int32)])
    v = np.concatenate([v, np.array(lv, np.int32)])
    assert np.any(g.cnt_common_pairs(u[n:], v[n:]) > 0)
    n = len(u)
    thr = rng.integers(-1, ref.V + 2, n).astype(np.int32)
    for k in sorted({0, opg.P // 2, opg.P - 1}):
        got = g.cnt_common_pairs(u, v, partition=k)
        assert got.tolist() == [opg.cnt_merge(a, b, k) for a, b in zip(u, v)]
        got = g.cnt_common_pairs(u, v, thr=thr, partition=k)
        assert got.tolist() == [opg.cnt_less_than_merge(a, b, t, k) for a, b, t in zip(u, v, thr)]
        qs = [rng.choice(hubs, int(rng.integers(1, 5))).astype(np.int32) for _ in range(100)]
        assert g.cnt_common_kway(qs, partition=k).tolist() == [opg.kway_merge(q, k) for q in qs]
        for i in range(12):
            nodes = np.sort(rng.choice(ref.V, int(rng.integers(0, 400)))).astype(np.int32)
            assert g.common_neis(u[i], nodes=nodes, partition=k).tolist() == \
                opg.common_neis(u[i], nodes, k).tolist()
            assert pg.get_neis(u[i], k).tolist() == opg.get_neis(u[i], k).tolist()
    got = g.cnt_common_pairs(u, v, thr=thr)                      # Graph::less_than
    assert got.tolist() == [orc.cnt_less_than_merge(ref, a, b, t) for a, b, t in zip(u, v, thr)]
    for i in range(20):                                          # Graph::get_common_neis
        vl = ref.elist[ref.vlist[v[i]]:ref.vlist[v[i] + 1]]
        assert g.common_neis(u[i], v=v[i]).tolist() == orc.common_neis(ref, u[i], vl).tolist()
    # the partitions split every count (the batch forms sum them)
    tot = sum(g.cnt_common_pairs(u, v, partition=k).astype(np.int64) for k in range(opg.P))
    assert np.array_equal(tot, g.cnt_common_pairs(u, v).astype(np.int64))


def test_common_list_edges(ctx, orc):
    """Long lists (several 256-element chunks), duplicates and out-of-range values in the
    node list, and empty rows."""
    n = 1100
    ref = complete(n, orc)
    g = gpu_graph(ctx, ref)
    nodes = np.sort(np.concatenate([np.arange(-5, n + 5), np.arange(0, n, 3)])).astype(np.int32)
    want = [x for x in range(n) if x != 7]
    assert g.common_neis(7, nodes=nodes).tolist() == want
    assert g.common_neis(7, v=8).tolist() == [x for x in range(n) if x not in (7, 8)]
    g.prepare(4)
    width = -(-n // 4)
    assert g.common_neis(7, v=7, partition=3).tolist() == list(range(3 * width, n))
    iso = orc.edges_to_csr(6, [0], [1])                          # vertices 2..5 isolated
    gi = gpu_graph(ctx, iso)
    assert gi.common_neis(3, v=0).tolist() == []
    assert gi.common_neis(0, nodes=np.array([1, 1, 2], np.int32)).tolist() == [1]
    with pytest.raises(pac.PacError):
        gi.common_neis(0, nodes=np.array([2, 1], np.int32))     # unsorted
    with pytest.raises(pac.PacError):
        gi.cnt_common_pairs([0], [1], partition=0)               # not prepared
    gi.prepare(2)
    with pytest.raises(pac.PacError):
        gi.cnt_common_pairs([0], [1], partition=2)               # P = 2


def test_multiway_argument_errors(ctx, orc):
    g = gpu_graph(ctx, complete(5, orc))
    with pytest.raises(pac.PacError):
        g.cnt_common_kway([np.array([0, 9], np.int32)])     # vertex out of range
    with pytest.raises(pac.PacError):
        g.cnt_common_kway([np.zeros(0, np.int32)])          # empty query
    with pytest.raises(pac.PacError):
        g.friend_triangle(7, [0, 1])                         # root out of range
