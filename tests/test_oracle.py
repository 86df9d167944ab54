"""CPU tests that pin the oracle (oracle/pac_oracle.cpp) before anything is checked
against it: golden fixtures computed by dense linear algebra, closed forms,
networkx, brute force on small inputs, and invariants."""
import itertools
import json
import math
import os

import networkx as nx
import numpy as np
import pytest

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")
NAMES = sorted(k for k in json.load(open(os.path.join(GOLDEN, "golden_index.json"))))


def load(name, orc):
    z = np.load(os.path.join(GOLDEN, f"{name}.npz"))
    return orc.CSR(z["vlist"], z["elist"]), z


def same_with_nan(a, b):
    a, b = np.asarray(a), np.asarray(b)
    return a.shape == b.shape and np.array_equal(np.isnan(a), np.isnan(b)) and \
        np.array_equal(a[~np.isnan(a)], b[~np.isnan(b)])


def nx_graph(g):
    G = nx.Graph()
    G.add_nodes_from(range(g.V))
    u, v = g.edges()
    G.add_edges_from(zip(u.tolist(), v.tolist()))
    return G


# ---- golden fixtures -----------------------------------------------------------------

@pytest.mark.parametrize("name", NAMES)
def test_golden_tc(name, orc):
    g, z = load(name, orc)
    assert orc.tc_merge(g, 2) == int(z["T"])
    assert orc.tc_galloping(g, 2) == 6 * int(z["T"])      # Graph.cpp:342-353 returns 6T


@pytest.mark.parametrize("name", NAMES)
def test_golden_per_vertex_and_lcc(name, orc):
    g, z = load(name, orc)
    assert np.array_equal(orc.per_vertex_triangles(g, 2), z["t"])
    assert same_with_nan(orc.lcc_merge(g, 2), z["lcc"])
    assert same_with_nan(orc.lcc_galloping(g, 2), z["lcc"])


@pytest.mark.parametrize("name", NAMES)
def test_golden_jaccard(name, orc):
    g, z = load(name, orc)
    inter, uni, ratio = orc.jaccard_edges(g, z["eu"], z["ev"], 2)
    assert np.array_equal(inter, z["inter"])
    assert np.array_equal(uni, z["uni"])
    assert np.array_equal(ratio.view(np.uint32), z["ratio"].view(np.uint32))
    assert same_with_nan(orc.jsim_merge(g, z["queries"], 2), z["jsim"])
    assert same_with_nan(orc.jsim_galloping(g, z["queries"], 2), z["jsim"])


def test_text_loader_fixtures(orc):
    exp = json.load(open(os.path.join(GOLDEN, "text_expected.json")))
    for f, e in exp.items():
        if f.startswith("_"):
            continue
        g = orc.text_to_csr(os.path.join(GOLDEN, f))
        assert g.V == e["V"]
        assert g.vlist.tolist() == e["vlist"]
        assert g.elist.tolist() == e["elist"]


def test_text_loader_missing_file(orc, tmp_path):
    with pytest.raises(OSError):
        orc.text_to_csr(str(tmp_path / "nope.txt"))


# ---- closed forms ---------------------------------------------------------------------

def complete(n, orc):
    return orc.edges_to_csr(n, *zip(*itertools.combinations(range(n), 2)))


@pytest.mark.parametrize("n", [3, 4, 5, 9, 30])
def test_complete_graph(n, orc):
    g = complete(n, orc)
    assert orc.tc_merge(g, 2) == math.comb(n, 3)
    assert np.all(orc.lcc_merge(g, 2) == 1.0)
    u, v = g.edges()
    inter, uni, ratio = orc.jaccard_edges(g, u, v, 2)
    assert np.all(inter == n - 2) and np.all(uni == n)
    assert np.all(ratio == np.float32((n - 2) / n))


@pytest.mark.parametrize("n", [4, 5, 8, 17])
def test_wheel(n, orc):
    edges = [(0, i) for i in range(1, n + 1)] + [(i, i % n + 1) for i in range(1, n + 1)]
    g = orc.edges_to_csr(n + 1, *zip(*edges))
    T = n if n > 3 else 4
    assert orc.tc_merge(g, 2) == T
    lcc = orc.lcc_merge(g, 2)
    assert lcc[0] == 2.0 / (n - 1)
    assert np.all(lcc[1:] == 2.0 / 3.0)


@pytest.mark.parametrize("n", [3, 4, 10])
def test_cycle(n, orc):
    g = orc.edges_to_csr(n, list(range(n)), [(i + 1) % n for i in range(n)])
    assert orc.tc_merge(g, 2) == (1 if n == 3 else 0)


def test_bipartite_and_friendship(orc):
    g = orc.edges_to_csr(9, *zip(*[(i, j) for i in range(4) for j in range(4, 9)]))
    assert orc.tc_merge(g, 2) == 0
    k = 7
    e = [(0, 2 * i + 1) for i in range(k)] + [(0, 2 * i + 2) for i in range(k)] + \
        [(2 * i + 1, 2 * i + 2) for i in range(k)]
    g = orc.edges_to_csr(2 * k + 1, *zip(*e))
    assert orc.tc_merge(g, 2) == k


def test_degenerate(orc):
    empty = orc.CSR(np.zeros(5, np.uint64), np.zeros(0, np.int32))
    assert orc.tc_merge(empty, 1) == 0
    lcc = orc.lcc_merge(empty, 1)
    assert np.all(lcc == 0.0)                                  # d = 0 keeps 0.0 (Graph.cpp:833)
    one = orc.edges_to_csr(3, [0], [1])
    assert orc.tc_merge(one, 1) == 0
    lcc = orc.lcc_merge(one, 1)
    assert np.isnan(lcc[0]) and np.isnan(lcc[1]) and lcc[2] == 0.0   # d = 1 -> 0/0


# ---- brute force and networkx ---------------------------------------------------------

@pytest.mark.parametrize("seed", [1, 2, 3])
def test_dense_brute_force_rmat(seed, orc):
    g = orc.gen_rmat(8, 8, 0.57, 0.19, 0.19, seed=seed)
    A = g.dense()
    A3 = A @ A @ A
    assert orc.tc_merge(g, 2) == np.trace(A3) // 6
    assert np.array_equal(orc.per_vertex_triangles(g, 2), np.diag(A3) // 2)


@pytest.mark.parametrize("seed", [5, 6])
def test_networkx(seed, orc):
    g = orc.gen_er(300, 12.0, seed=seed)
    G = nx_graph(g)
    tri = nx.triangles(G)
    assert orc.tc_merge(g, 2) == sum(tri.values()) // 3
    assert np.array_equal(orc.per_vertex_triangles(g, 2), [tri[i] for i in range(g.V)])
    cl = nx.clustering(G)
    lcc = orc.lcc_merge(g, 2)
    deg = g.deg()
    for u in range(g.V):
        if deg[u] >= 2:
            assert lcc[u] == pytest.approx(cl[u], rel=1e-12)
    u, v = g.edges()
    _, _, ratio = orc.jaccard_edges(g, u, v, 2)
    ref = {(a, b): r for a, b, r in nx.jaccard_coefficient(G, list(zip(u.tolist(), v.tolist())))}
    for i in range(len(u)):
        assert ratio[i] == np.float32(ref[(int(u[i]), int(v[i]))])


def test_intersections_vs_sets(orc):
    rng = np.random.default_rng(7)
    for _ in range(20):
        g = orc.gen_er(120, rng.uniform(2, 30), seed=int(rng.integers(1 << 30)))
        for _ in range(30):
            a, b = rng.integers(0, g.V, 2)
            na = set(g.elist[g.vlist[a]:g.vlist[a + 1]].tolist())
            nb = set(g.elist[g.vlist[b]:g.vlist[b + 1]].tolist())
            assert orc.cnt_merge(g, a, b) == len(na & nb)
            assert orc.cnt_galloping(g, a, b) == len(na & nb)
            thr = int(rng.integers(0, g.V))
            assert orc.cnt_less_than_merge(g, a, b, thr) == len({x for x in na & nb if x < thr})


# ---- partitioned graph -----------------------------------------------------------------

@pytest.mark.parametrize("P", [1, 4, 64])
def test_partitioned_sum_invariance(P, orc):
    g = orc.gen_rmat(10, 16, 0.57, 0.19, 0.19, seed=10)
    T = orc.tc_merge(g, 4)
    for v2p in (orc.contiguous_partition(g.V, P), orc.random_partition(g.V, P, seed=P)):
        pg = orc.PGraph(g, v2p, 4)
        assert pg.tc_merge() == T
        assert pg.tc_galloping() == 6 * T
        assert sum(pg.tc_merge_partial(i) for i in range(pg.P)) == T
        assert same_with_nan(pg.lcc_merge(), orc.lcc_merge(g, 4))


def test_partitioned_partial_is_smallest_vertex_partition(orc):
    """Partial i counts the triangles whose smallest-ID vertex lies in partition i."""
    g = orc.gen_rmat(8, 16, 0.57, 0.19, 0.19, seed=3)
    P = 5
    v2p = orc.random_partition(g.V, P, seed=11)
    pg = orc.PGraph(g, v2p, 2)
    A = g.dense()
    want = np.zeros(P, np.int64)
    for w in range(g.V):
        nbr = np.nonzero(A[w])[0]
        hi = nbr[nbr > w]
        sub = A[np.ix_(hi, hi)]
        want[v2p[w]] += int(sub.sum()) // 2
    assert [pg.tc_merge_partial(i) for i in range(P)] == want.tolist()


# ---- generators -----------------------------------------------------------------------

def check_structure(g):
    vl, el = g.vlist.astype(np.int64), g.elist.astype(np.int64)
    assert vl[0] == 0 and np.all(np.diff(vl) >= 0) and vl[-1] == len(el)
    src = np.repeat(np.arange(g.V), np.diff(vl))
    assert np.all(el >= 0) and np.all(el < g.V)
    assert not np.any(src == el)
    same_row = src[1:] == src[:-1]
    assert np.all(el[1:][same_row] > el[:-1][same_row])           # sorted, unique
    fwd = set(zip(src.tolist(), el.tolist()))
    assert all((b, a) in fwd for a, b in fwd)                      # symmetric


def test_rmat_generator(orc):
    g = orc.gen_rmat(10, 16, 0.57, 0.19, 0.19, seed=42)
    check_structure(g)
    g2 = orc.gen_rmat(10, 16, 0.57, 0.19, 0.19, seed=42)
    assert np.array_equal(g.elist, g2.elist)
    g3 = orc.gen_rmat(10, 16, 0.57, 0.19, 0.19, seed=43)
    assert not np.array_equal(g.vlist, g3.vlist)
    # power law: the low-ID corner is dense
    assert g.deg()[0] > 20 * np.median(g.deg()[g.deg() > 0])


def test_er_generator(orc):
    n, k = 4000, 10.0
    g = orc.gen_er(n, k, seed=9)
    check_structure(g)
    m = g.nnz // 2
    mean = k * n / 2
    sd = math.sqrt(mean)
    assert abs(m - mean) < 5 * sd


def test_oriented_stats_brute(orc):
    g = orc.gen_rmat(7, 8, 0.57, 0.19, 0.19, seed=4)
    d = g.deg()
    s, t = g.edges()
    dplus = np.zeros(g.V, np.int64)
    ori = []
    for a, b in zip(s.tolist(), t.tolist()):
        lo, hi = (a, b) if (d[a], a) < (d[b], b) else (b, a)
        dplus[lo] += 1
        ori.append((lo, hi))
    W = sum(dplus[a] + dplus[b] for a, b in ori)
    st = orc.oriented_stats(g)
    assert st["m"] == len(ori) and st["W"] == W and st["max_outdeg"] == dplus.max()


def test_evaluate_partition(orc):
    """K_8 in 2 contiguous halves: vertex 0 sees 3 neighbours in its own half and 4 in the
    other, so max_part_deg = 4, max_deg = 7, ratio 8/7; in general the per-partition
    neighbour histogram from numpy."""
    k8 = orc.gen_er(8, 7, seed=1)
    r, mp, md = orc.PGraph(k8, orc.contiguous_partition(8, 2)).evaluate_partition()
    assert (mp, md) == (4, 7) and r == 8 / 7
    r, mp, md = orc.PGraph(k8, np.zeros(8, np.int32)).evaluate_partition()
    assert (mp, md, r) == (7, 7, 1.0)
    for P in (3, 16):
        g = orc.gen_rmat(10, 8, 0.57, 0.19, 0.19, seed=P)
        v2p = orc.random_partition(g.V, P, seed=P)
        src = np.repeat(np.arange(g.V), g.deg())
        hist = np.zeros((g.V, v2p.max() + 1), np.int64)
        np.add.at(hist, (src, v2p[g.elist]), 1)
        r, mp, md = orc.PGraph(g, v2p).evaluate_partition()
        assert mp == hist.max() and md == g.deg().max()
        assert r == (v2p.max() + 1) * hist.max() / g.deg().max()
    empty = orc.edges_to_csr(4, np.zeros(0, np.uint32), np.zeros(0, np.uint32))
    assert math.isnan(orc.PGraph(empty, np.zeros(4, np.int32)).evaluate_partition()[0])


def test_greedy_partition_2_known_answers(orc):
    """Placements derived by hand from Graph.cpp:969-1015.  Path 0-1-2, P=2: vertex 1
    (degree 2) takes partition 0; vertex 0 sees count 0 at vertex 1 in both partitions and
    takes the first, 0; vertex 2 sees count 1 in partition 0 and 0 in partition 1: 1.  Star
    with centre 0 and leaves 1-4, P=2: the centre takes 0; the leaves see the centre's
    counts (0,0) -> 0, then (1,0) -> 1, (1,1) -> 0, (2,1) -> 1."""
    path = orc.edges_to_csr(3, np.array([0, 1], np.uint32), np.array([1, 2], np.uint32))
    assert orc.greedy_partition_2(path, 2).tolist() == [0, 0, 1]
    star = orc.edges_to_csr(5, np.zeros(4, np.uint32), np.arange(1, 5, dtype=np.uint32))
    assert orc.greedy_partition_2(star, 2).tolist() == [0, 0, 1, 0, 1]
    # isolated vertices take the random_partition stream
    iso = orc.edges_to_csr(6, np.array([0], np.uint32), np.array([1], np.uint32))
    got = orc.greedy_partition_2(iso, 4, seed=9)
    assert got[2:].tolist() == orc.random_partition(6, 4, seed=9)[2:].tolist()


@pytest.mark.parametrize("P", [2, 5])
def test_greedy_partition_2_minimax(P, orc):
    """Replaying the placements: every vertex with neighbours went to the first partition
    whose largest neighbour count (at the moment it was placed) is smallest."""
    g = orc.gen_rmat(8, 8, 0.57, 0.19, 0.19, seed=P)
    v2p = orc.greedy_partition_2(g, P)
    d = g.deg()
    order = sorted(range(g.V), key=lambda u: (-d[u], u))
    cnt = np.zeros((P, g.V), np.int64)
    for u in order:
        nb = g.elist[g.vlist[u]:g.vlist[u + 1]]
        if len(nb):
            mx = cnt[:, nb].max(axis=1)
            assert v2p[u] == int(np.argmin(mx)), u
        cnt[v2p[u], nb] += 1


def brute_pairs(g):
    """Pair ownership walked per owner row over its symmetric row, as the device builds it
    (x owns {x, u} when u->x, d+(x) > 0 and some out-neighbour of u ranks above x), the
    suffix counted with Python sets -- a different formulation from the oracle's sorted
    positions per tail."""
    d = g.deg()
    before = lambda a, b: (d[a], a) < (d[b], b)
    nb = [set(g.elist[g.vlist[x]:g.vlist[x + 1]].tolist()) for x in range(g.V)]
    outs = [{w for w in nb[x] if before(x, w)} for x in range(g.V)]
    pairs = rows = streamed = staged = 0
    for x in range(g.V):
        own = 0
        for u in nb[x]:
            if not before(u, x) or not outs[x]:
                continue
            suffix = sum(1 for w in outs[u] if before(x, w))
            if suffix:
                own += 1
                streamed += suffix
        pairs += own
        if own:
            rows += 1
            staged += len(outs[x])
    return {"pairs": pairs, "rows": rows, "streamed": streamed, "staged": staged}


@pytest.mark.parametrize("n", [3, 4, 7, 12])
def test_pair_stats_complete_graph(n, orc):
    """K_n orients by ID (all degrees equal), d+(i) = n-1-i: the pair t->h streams the
    n-1-h IDs above h, every one a hit, so the streamed total is C(n,3); pairs are the
    edges t<h<=n-2 (C(n-1,2)), owned by the heads 1..n-2 staging sum of (n-1-h)."""
    z = np.load(os.path.join(GOLDEN, "K12.npz")) if n == 12 else None
    g = orc.CSR(z["vlist"], z["elist"]) if z is not None else orc.gen_er(n, n - 1, seed=1)
    assert g.nnz == n * (n - 1)
    s = orc.pair_stats(g)
    assert s["streamed"] == math.comb(n, 3)
    assert s["pairs"] == math.comb(n - 1, 2)
    assert s["rows"] == max(0, n - 2)
    assert s["staged"] == sum(n - 1 - h for h in range(1, n - 1))


@pytest.mark.parametrize("seed", [1, 2, 3])
def test_pair_stats_brute(seed, orc):
    for g in (orc.gen_rmat(8, 8, 0.57, 0.19, 0.19, seed=seed), orc.gen_er(300, 9.0, seed=seed)):
        assert orc.pair_stats(g) == brute_pairs(g)
    star = orc.edges_to_csr(9, np.zeros(8, np.uint32), np.arange(1, 9, dtype=np.uint32))
    assert orc.pair_stats(star) == {"pairs": 0, "rows": 0, "streamed": 0, "staged": 0}


@pytest.mark.parametrize("name", NAMES)
def test_tc_oriented_golden(name, orc):
    g, z = load(name, orc)
    assert orc.tc_oriented(g, 2) == int(z["T"])


def test_tc_oriented_vs_restatement(orc):
    """The fast checker against the reference's ID-ordered merge count on R-MAT (hubs at
    low IDs), skewed and flat Erdos-Renyi graphs."""
    for s in (6, 9, 12, 14):
        g = orc.gen_rmat(s, 16, 0.57, 0.19, 0.19, seed=s)
        assert orc.tc_oriented(g, 4) == orc.tc_merge(g, 4)
    for n, k in ((500, 8.0), (3000, 40.0)):
        g = orc.gen_er(n, k, seed=n)
        assert orc.tc_oriented(g, 4) == orc.tc_merge(g, 4)


# ---- MLP oracle vs finite differences of an independent float64 forward ---------------

def np_forward(P, x):
    P = P.astype(np.float64)
    W1 = P[:1024].reshape(64, 16); b1 = P[1024:1088]
    W2 = P[1088:5184].reshape(64, 64); b2 = P[5184:5248]
    W3 = P[5248:5440].reshape(3, 64); b3 = P[5440:5443]
    h1 = np.maximum(x @ W1.T + b1, 0)
    h2 = np.maximum(h1 @ W2.T + b2, 0)
    return h2 @ W3.T + b3


def np_loss(P, x, y, w):
    z = np_forward(P, x.astype(np.float64))
    z = z - z.max(1, keepdims=True)
    lse = np.log(np.exp(z).sum(1))
    return float(np.mean(w * (lse - z[np.arange(len(y)), y])))


def test_mlp_oracle_forward_and_gradient(orc):
    rng = np.random.default_rng(0)
    P = (rng.standard_normal(orc.N_PARAMS) * 0.3).astype(np.float32)
    x = rng.random((16, 16)).astype(np.float32)
    y = rng.integers(0, 3, 16).astype(np.int32)
    w = rng.uniform(0.5, 2.0, 16).astype(np.float32)
    z = orc.mlp_forward(P, x)
    assert np.allclose(z, np_forward(P, x.astype(np.float64)), rtol=1e-5, atol=1e-5)
    loss, g = orc.mlp_backward(P, x, y, w)
    assert loss == pytest.approx(np_loss(P, x, y, w), rel=1e-5)
    idx = rng.choice(orc.N_PARAMS, 60, replace=False)
    for k in idx:
        e = np.zeros(orc.N_PARAMS, np.float64); e[k] = 1e-4
        fd = (np_loss(P + e, x, y, w) - np_loss(P - e, x, y, w)) / 2e-4
        assert g[k] == pytest.approx(fd, rel=2e-3, abs=2e-6)


# ---- k-way intersections, friend triangles, squares (independent definitions) --------

def neighbour_sets(g):
    return [set(g.elist[g.vlist[u]:g.vlist[u + 1]].tolist()) for u in range(g.V)]


def brute_squares(g):
    """4-tuples u > v > w > x with u~v, u~w, x~v, x~w, enumerated directly."""
    N = neighbour_sets(g)
    res = 0
    for u in range(g.V):
        for v in N[u]:
            if v >= u:
                continue
            for w in N[u]:
                if w >= v:
                    continue
                res += sum(1 for x in N[v] & N[w] if x < w)
    return res


@pytest.mark.parametrize("seed", [1, 2, 3])
def test_kway_vs_sets(seed, orc):
    g = orc.gen_rmat(9, 8, seed=seed)
    N = neighbour_sets(g)
    rng = np.random.default_rng(seed)
    hubs = np.argsort(-np.diff(g.vlist.astype(np.int64)))[:40]
    for k in (1, 2, 3, 4, 6):
        for _ in range(40):
            q = rng.choice(hubs if rng.random() < 0.7 else np.arange(g.V), size=k)
            want = len(set.intersection(*[N[int(v)] for v in q]))
            assert orc.kway_merge(g, q) == want
            assert orc.kway_galloping(g, q) == want


@pytest.mark.parametrize("P,seed", [(1, 1), (4, 2), (7, 3)])
def test_partition_queries_vs_sets(P, seed, orc):
    """Per-partition intersections (Graph.cpp:1071-1195, 1475-1498) against Python sets
    restricted to the partition's columns; Graph-level node-list queries (241-253, 468-493)
    against plain sets."""
    g = orc.gen_rmat(9, 8, seed=seed)
    N = neighbour_sets(g)
    v2p = orc.random_partition(g.V, P, seed) if P > 1 else np.zeros(g.V, np.int32)
    pg = orc.PGraph(g, v2p)
    Pn = pg.P
    part = lambda k: {w for w in range(g.V) if v2p[w] == k}
    cols = [part(k) for k in range(Pn)]
    rng = np.random.default_rng(seed)
    hubs = np.argsort(-g.deg())[:30]
    for _ in range(60):
        u, v = (int(x) for x in rng.choice(hubs if rng.random() < 0.6 else np.arange(g.V), 2))
        k = int(rng.integers(Pn))
        common = N[u] & N[v] & cols[k]
        assert pg.cnt_merge(u, v, k) == len(common)
        assert pg.cnt_galloping(u, v, k) == len(common)
        thr = int(rng.integers(0, g.V + 1))
        assert pg.cnt_less_than_merge(u, v, thr, k) == len({w for w in common if w < thr})
        nodes = np.sort(rng.choice(g.V, size=int(rng.integers(0, 60))))   # duplicates kept
        want = sorted({int(x) for x in nodes} & N[u] & cols[k])
        assert pg.cnt_merge_nodes(u, nodes, k) == len(want)
        assert pg.common_neis(u, nodes, k).tolist() == want
        assert pg.get_neis(u, k).tolist() == sorted(N[u] & cols[k])
        assert orc.cnt_merge_nodes(g, u, nodes) == len({int(x) for x in nodes} & N[u])
        assert orc.common_neis(g, u, nodes).tolist() == sorted({int(x) for x in nodes} & N[u])
        vl = g.elist[g.vlist[v]:g.vlist[v + 1]]
        assert orc.common_neis(g, u, vl).tolist() == sorted(N[u] & N[v])
        q = rng.choice(hubs, size=int(rng.integers(1, 5)))
        want = len(set.intersection(*[N[int(x)] for x in q]) & cols[k])
        assert pg.kway_merge(q, k) == want
        assert pg.kway_galloping(q, k) == want
    # the partitions split every count
    u, v = int(hubs[0]), int(hubs[1])
    assert sum(pg.cnt_merge(u, v, k) for k in range(Pn)) == len(N[u] & N[v])


def test_friend_triangle(orc):
    n = 9                                    # K_n: every pair sees n - 3 common neighbours
    g = complete(n, orc)
    assert orc.friend_triangle_merge(g, 0, list(range(n)), 2) == math.comb(n - 1, 2) * (n - 3)
    g = orc.gen_rmat(9, 8, seed=5)
    N = neighbour_sets(g)
    rng = np.random.default_rng(5)
    for root in (0, 3, 17):
        nodes = rng.choice(g.V, size=30).tolist() + [root, root]
        want = sum(len(N[root] & N[a] & N[b])
                   for i, a in enumerate(nodes) for b in nodes[i + 1:]
                   if a != root and b != root)
        assert orc.friend_triangle_merge(g, root, nodes, 2) == want


@pytest.mark.parametrize("P", [1, 3, 8])
def test_square_counting(P, orc):
    for n in (4, 5, 7):                      # K_n: one arrangement in three per 4-set
        g = complete(n, orc)
        assert orc.PGraph(g, orc.random_partition(n, P, 3)).square_counting() == math.comb(n, 4)
    g = orc.edges_to_csr(9, *zip(*[(i, j) for i in range(4) for j in range(4, 9)]))
    assert orc.PGraph(g, orc.random_partition(9, P, 4)).square_counting() == 0   # K_{4,5}
    for seed in (1, 2):
        g = orc.gen_rmat(7, 6, seed=seed)
        assert orc.PGraph(g, orc.random_partition(g.V, P, seed)).square_counting() == brute_squares(g)
