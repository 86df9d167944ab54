"""Measurements of the non-headline configurations (one JSON line each).

    python experiments.py scale22   # plain vs batched, every strategy
    python experiments.py batch     # 256 uneven graphs, autotuned tile
    python experiments.py jaccard   # per-edge Jaccard at R-MAT scale 25
    python experiments.py next      # ingestion, queries, k-way, squares
    python experiments.py cpu       # the oracle's whole solves timed on the host

scale22: R-MAT scale 22, ef 16, (0.57,0.19,0.19,0.05), seed 22, 1024 contiguous
  partitions; TC + LCC device time of the plain launch (one launch per partition
  group) and of the persistent/hipGraph launch for each strategy and dispatch.
batch: 256 graphs, graph g seeded 3000+g; even g: Erdos-Renyi G(n,p) with n = 2^U(10,18)
  and average degree U(8,64); odd g: R-MAT scale U{10..18}, ef 16; the draws come from
  a SplitMix64 stream seeded 3.  Each graph gets ceil(V/4096) contiguous partitions.
  The tile (waves per row, LDS sizes) is autotuned on the batch; reported: one batched
  persistent launch (and its hipGraph replay) vs one plain solve per graph, and every
  graph's count against the oracle.
jaccard: R-MAT scale 25, ef 16, seed 25, 4096 partitions; per-edge |N(u)&N(v)|,
  union and FP32 ratio for every edge with all results left in HBM; sampled edges are
  checked against the oracle; the oracle's per-edge merge is timed on a sample of edges.
cpu: the oracle's restatement of the reference's timed solves, run whole on the host
  (OMP_NUM_THREADS threads, each leg in a child process under a time limit, on the CSR the
  GPU generated): R-MAT scale 16 TC-merge + LCC-merge (Graph.cpp:355-368, 828-840) and the
  64-partition Partitioned_Graph TC (Graph.cpp:1346-1362), median of 3 after a warm-up;
  R-MAT scale 22 TC-merge, median of 3 after a warm-up; R-MAT scale 24 TC-merge, one run,
  its rows in 64 consecutive ranges each run as an OpenMP loop (a line printed per range;
  this spreads the low-ID hub rows over the threads better than the reference's one static
  loop, so it errs on the CPU's side).  --scale S runs only the legs at scale S.
next: the components on either side of the path, on R-MAT scale 22 (seed 22) unless noted:
  edge-list ingestion (the graph's undirected edge list, shuffled, to the symmetric CSR on
  the device; an R-MAT scale 18 edge list as text through the device parser), pairwise
  common-neighbour counts (4M uniform random pairs, as test.cpp's generate_queries draws
  them, and 4M random adjacent pairs), 3-way counts (1M queries u, v, w with v, w drawn
  from N(u)), one friend-triangle query (a root of degree about 1000 with its first 1024
  neighbours) and square counting (R-MAT scale 12 against the oracle, scale 16 timed).
  Device times are HIP events around the device work with the inputs resident; each
  is checked exactly against the oracle, which is timed on the same (sampled) inputs.
"""
from __future__ import annotations

import argparse
import json
import math
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

import partitionandconquer_amd as pac  # noqa: E402

GAMMA = 0x9E3779B97F4A7C15
MASK = (1 << 64) - 1


def mix64(z):
    z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & MASK
    z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & MASK
    return z ^ (z >> 31)


class Stream:
    def __init__(self, seed):
        self.st = mix64((seed + GAMMA) & MASK)

    def uniform(self):
        self.st = (self.st + GAMMA) & MASK
        return (mix64(self.st) >> 11) / float(1 << 53)


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def median_ms(ctx, fn, reps=5, warm=1):
    for _ in range(warm):
        fn()
    v = []
    for _ in range(reps):
        fn()
        v.append(ctx.last_timing().total_ms)
    return float(np.median(v))


def exp_scale22(ctx, scale=22, partitions=1024):
    g = pac.Graph.rmat(ctx, scale, 16, 0.57, 0.19, 0.19, seed=scale)
    g.prepare(partitions)
    info = g.info()
    T = g.tc(pac.FORCE_BITMAP)
    res = {"config": f"rmat-s{scale}-ef16-P{partitions}", "m": info.num_edges, "W": info.work_W,
           "triangles": T, "device_ms": {}}
    for dname, d in (("fixed", pac.DISPATCH_FIXED), ("merge", pac.FORCE_MERGE),
                     ("gallop", pac.FORCE_GALLOP), ("bitmap", pac.FORCE_BITMAP)):
        for lname, l in (("plain", pac.LAUNCH_PLAIN), ("persistent", pac.LAUNCH_PERSISTENT),
                         ("graph", pac.LAUNCH_GRAPH)):
            assert g.lcc_device(d, l) == T
            ms = median_ms(ctx, lambda: g.lcc_device(d, l))
            res["device_ms"][f"{dname}/{lname}"] = round(ms, 3)
    best = min(res["device_ms"]["fixed/graph"], res["device_ms"]["bitmap/graph"])
    res["gteps_batched"] = round(info.num_edges / (best * 1e-3) / 1e9, 4)
    # the intersection kernel alone (HIP events around its launch), against the bytes the
    # hash path must move in LCC mode
    ker = []
    for _ in range(5):
        g.lcc_device(pac.DISPATCH_FIXED, pac.LAUNCH_PERSISTENT)
        ker.append(ctx.last_timing().kernel_ms)
    kms = float(np.median(ker))
    alg = info.algorithmic_bytes(pac.MODE_LCC)
    res["kernel_ms"] = round(kms, 3)
    res["algorithmic_bytes"] = alg
    res["roofline_frac"] = round(alg / (kms * 1e-3) / 8e12, 4)
    res["speedup_batched_vs_plain_fixed"] = round(res["device_ms"]["fixed/plain"] / best, 3)
    return res


def batch_specs(G=256, master=3):
    s = Stream(master)
    out = []
    for g in range(G):
        if g % 2 == 0:
            n = int(round(2 ** (10 + 8 * s.uniform())))
            k = 8 + 56 * s.uniform()
            out.append({"kind": "er", "n": n, "k": k, "seed": 3000 + g})
        else:
            scale = 10 + min(8, int(9 * s.uniform()))
            out.append({"kind": "rmat", "scale": scale, "seed": 3000 + g})
    return out


def exp_batch(ctx, G=256, check=True):
    from oracle import oracle as O
    specs = batch_specs(G)
    gs = []
    t0 = time.time()
    for sp in specs:
        if sp["kind"] == "er":
            g = pac.Graph.er(ctx, sp["n"], sp["k"], seed=sp["seed"])
        else:
            g = pac.Graph.rmat(ctx, sp["scale"], 16, 0.57, 0.19, 0.19, seed=sp["seed"])
        g.prepare(max(1, -(-g.get_vertex_num() // 4096)))
        gs.append(g)
    gen_s = time.time() - t0
    m_total = sum(g.info().num_edges for g in gs)
    tile, tile_ms = ctx.autotune(gs, pac.MODE_TC, pac.DISPATCH_FIXED)
    tot = ctx.tc_batch(gs, pac.DISPATCH_FIXED, pac.LAUNCH_PERSISTENT)
    res = {"config": f"batch-{G}", "graphs": G, "m_total": m_total, "gen_s": round(gen_s, 1),
           "tile": tile, "tile_ms": [round(x, 3) for x in tile_ms]}
    for name, l in (("persistent", pac.LAUNCH_PERSISTENT), ("graph", pac.LAUNCH_GRAPH),
                    ("plain_per_graph", pac.LAUNCH_PLAIN)):
        v = []
        for _ in range(5):
            out = ctx.tc_batch(gs, pac.DISPATCH_FIXED, l)
            assert np.array_equal(out, tot)
            v.append(ctx.last_timing().total_ms)
        res[f"{name}_ms"] = round(float(np.median(v)), 3)
    res["speedup_batched_vs_per_graph"] = round(res["plain_per_graph_ms"] / res["graph_ms"], 3)
    res["gteps_batched"] = round(m_total / (res["graph_ms"] * 1e-3) / 1e9, 4)
    if check:
        bad = 0
        for g, t in zip(gs, tot):
            vl, el = g.csr()
            bad += int(O.tc_merge(O.CSR(vl, el), int(os.environ.get("OMP_NUM_THREADS", 8))) != int(t))
        res["oracle_mismatches"] = bad
    res["triangles_total"] = int(tot.sum())
    return res


def exp_jaccard(ctx, scale=25, partitions=4096, sample=200000, cpu_budget=15.0,
                profile_only=False):
    g = pac.Graph.rmat(ctx, scale, 16, 0.57, 0.19, 0.19, seed=scale)
    g.prepare(partitions)
    info = g.info()
    if profile_only:  # three batched solves, for a kernel trace
        ms = []
        for _ in range(3):
            g.jaccard_edges_device(pac.DISPATCH_FIXED, pac.LAUNCH_PERSISTENT)
            ms.append(round(ctx.last_timing().total_ms, 3))
        return {"config": f"rmat-s{scale}-ef16-P{partitions}-jaccard", "device_ms": ms}
    from oracle import oracle as O
    res = {"config": f"rmat-s{scale}-ef16-P{partitions}-jaccard", "m": info.num_edges,
           "W": info.work_W, "device_ms": {}}
    for name, d, l in (("plain_fixed", pac.DISPATCH_FIXED, pac.LAUNCH_PLAIN),
                       ("batched_fixed", pac.DISPATCH_FIXED, pac.LAUNCH_GRAPH),
                       ("batched_bitmap", pac.FORCE_BITMAP, pac.LAUNCH_GRAPH)):
        res["device_ms"][name] = round(median_ms(ctx, lambda: g.jaccard_edges_device(d, l), 3), 3)
    best = min(res["device_ms"]["batched_fixed"], res["device_ms"]["batched_bitmap"])
    res["gteps"] = round(info.num_edges / (best * 1e-3) / 1e9, 4)
    ker = []
    for _ in range(3):
        g.jaccard_edges_device(pac.DISPATCH_FIXED, pac.LAUNCH_PERSISTENT)
        ker.append(ctx.last_timing().kernel_ms)
    kms = float(np.median(ker))
    alg = info.algorithmic_bytes(pac.MODE_JACCARD)
    res["kernel_ms"] = round(kms, 3)
    res["algorithmic_bytes"] = alg
    res["roofline_frac"] = round(alg / (kms * 1e-3) / 8e12, 4)
    res["speedup_batched_vs_plain_fixed"] = round(res["device_ms"]["plain_fixed"] / best, 3)
    # parity on a sample of edges, and the CPU baseline on the same sample
    src, dst, inter, uni, ratio = g.jaccard_edges(pac.DISPATCH_FIXED, pac.LAUNCH_PERSISTENT)
    T = g.tc(pac.FORCE_BITMAP)
    res["sum_inter_equals_3T"] = int(inter.sum(dtype=np.uint64)) == 3 * T
    vl, el = g.csr()
    ref = O.CSR(vl, el)
    rng = np.random.default_rng(0)
    idx = rng.choice(len(src), size=min(sample, len(src)), replace=False)
    nt = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    t0 = time.perf_counter()
    ri, ru, rr = O.jaccard_edges(ref, src[idx], dst[idx], nt)
    dt = time.perf_counter() - t0
    res["sample_edges"] = int(len(idx))
    res["sample_exact"] = bool(np.array_equal(ri, inter[idx]) and np.array_equal(ru, uni[idx])
                               and np.array_equal(rr.view(np.uint32), ratio[idx].view(np.uint32)))
    res["cpu_baseline"] = {"edges_per_s": round(len(idx) / dt, 1), "threads": nt,
                           "est_full_s": round(dt * info.num_edges / len(idx), 1)}
    return res


CPU_WORKER = r"""
import json, sys, time
import numpy as np
sys.path.insert(0, sys.argv[1])
from oracle import oracle as O
vl, el = np.load(sys.argv[2]), np.load(sys.argv[3])
what, warm, reps, nt = sys.argv[4], int(sys.argv[5]), int(sys.argv[6]), int(sys.argv[7])
g = O.CSR(vl, el)
def run():
    if what == "tc":
        return O.tc_merge(g, nt)
    if what == "tc+lcc":
        return (O.tc_merge(g, nt), float(np.nansum(O.lcc_merge(g, nt))))
    if what == "ptc":
        return O.PGraph(g, O.contiguous_partition(g.V, 64), nt).tc_merge()
    if what == "tc-ranges":
        # the rows in 64 consecutive ranges, each an OpenMP loop over its rows, one line
        # printed per range
        tot = 0
        for k, rows in enumerate(np.array_split(np.arange(g.V, dtype=np.int32), 64)):
            tot += O.tc_merge_rows(g, rows, nt)
            print(json.dumps({"range": k}), flush=True)
        return tot
for _ in range(warm):
    run()
t = []
for _ in range(reps):
    t0 = time.perf_counter()
    out = run()
    t.append(time.perf_counter() - t0)
    print(json.dumps({"s": t[-1]}), flush=True)
print(json.dumps({"times_s": t, "median_s": sorted(t)[len(t) // 2], "result": str(out)}), flush=True)
"""


def exp_cpu(ctx, limit_s=900.0, only=None):
    import subprocess
    import tempfile
    nt = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    res = {"threads": nt, "legs": []}
    tmp = tempfile.mkdtemp(prefix="pac_cpu_")
    legs = ((16, "tc+lcc", 1, 3), (16, "ptc", 1, 3), (22, "tc", 1, 3), (24, "tc-ranges", 0, 1))
    for scale, what, warm, reps in legs:
        if only is not None and scale != only:
            continue
        g = pac.Graph.rmat(ctx, scale, 16, 0.57, 0.19, 0.19, seed=scale)
        vl, el = g.csr()
        m = g.info().num_edges
        g.free()
        fv, fe = os.path.join(tmp, "vl.npy"), os.path.join(tmp, "el.npy")
        np.save(fv, vl)
        np.save(fe, el)
        del vl, el
        log(f"cpu leg: scale {scale} {what}, warm-up {warm}, runs {reps}, limit {limit_s:.0f} s")
        leg = {"scale": scale, "solve": what, "m": m, "warmup": warm, "runs": reps}
        try:
            # the child's lines are passed through as they come (a leg can run for minutes)
            p = subprocess.Popen([sys.executable, "-c", CPU_WORKER, ROOT, fv, fe, what, str(warm),
                                  str(reps), str(nt)], stdout=subprocess.PIPE, text=True)
            t_end = time.time() + limit_s
            lines = []
            for line in p.stdout:
                lines.append(line)
                log(line.rstrip())
                if time.time() > t_end:
                    p.kill()
                    raise subprocess.TimeoutExpired(what, limit_s)
            p.wait()
            last = json.loads(lines[-1])
            leg.update(last)
            leg["gteps"] = round(m / last["median_s"] / 1e9, 6)
        except subprocess.TimeoutExpired:
            leg["timed_out_after_s"] = limit_s
        log(json.dumps(leg))
        res["legs"].append(leg)
    return res


def exp_next(ctx, scale=22, seed=22, npairs=4_000_000, nkway=1_000_000, sample=200_000,
             profile_only=False):
    rng = np.random.default_rng(seed)
    g = pac.Graph.rmat(ctx, scale, 16, 0.57, 0.19, 0.19, seed=seed)
    vl, el = g.csr()
    if profile_only:  # three timed counts of 4M random adjacent pairs, for a kernel trace
        e = rng.integers(0, len(el), npairs)
        u = (np.searchsorted(vl.astype(np.int64), e, side="right") - 1).astype(np.int32)
        v = el[e].astype(np.int32)
        ms = []
        for _ in range(3):
            g.cnt_common_pairs(u, v)
            ms.append(round(ctx.last_timing().total_ms, 3))
        return {"config": f"rmat-s{scale}-ef16-adjacent-pairs", "pairs": npairs, "device_ms": ms}
    from oracle import oracle as O
    nt = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    res = {"config": f"rmat-s{scale}-ef16-next", "cpu_threads": nt}
    ref = O.CSR(vl, el)
    V = len(vl) - 1
    vli = vl.astype(np.int64)
    deg = np.diff(vli)

    # (1) ingestion: undirected edge list (u < v, shuffled) -> symmetric CSR
    src = np.repeat(np.arange(V, dtype=np.int32), deg)
    keep = src < el
    perm = rng.permutation(int(keep.sum()))
    su, sv = src[keep][perm], el[keep][perm]
    del src, keep, perm
    wall = []
    for i in range(3):
        t0 = time.perf_counter()
        g1 = pac.Graph.from_edges(ctx, su, sv, V)
        wall.append((time.perf_counter() - t0) * 1e3)
        if i == 0:
            vl1, el1 = g1.csr()
            same = bool(np.array_equal(vl1, vl) and np.array_equal(el1, el))
            del vl1, el1
        g1.free()
    t0 = time.perf_counter()
    rc = O.edges_to_csr(V, su, sv)
    cpu_s = time.perf_counter() - t0
    same = same and bool(np.array_equal(rc.vlist, vl) and np.array_equal(rc.elist, el))
    del rc
    res["ingest_edges"] = {"edges": int(len(su)), "wall_ms": round(float(np.median(wall)), 1),
                           "edges_per_s": round(len(su) / (np.median(wall) * 1e-3), 1),
                           "includes": "host checks + host-to-device copy + device sort/dedup/CSR",
                           "csr_identical": same, "cpu_s": round(cpu_s, 2), "cpu_threads": 1}
    del su, sv
    gs = O.gen_rmat(18, 16, 0.57, 0.19, 0.19, 18)
    d18 = np.diff(gs.vlist.astype(np.int64))
    s18 = np.repeat(np.arange(gs.V, dtype=np.int32), d18)
    k18 = s18 < gs.elist
    path = os.path.join(os.environ.get("TMPDIR", "/tmp"), f"pac_rmat18_{os.getpid()}.txt")
    np.savetxt(path, np.stack([s18[k18], gs.elist[k18]], 1), fmt="%d %d")
    try:
        wall = []
        for i in range(3):
            t0 = time.perf_counter()
            g1 = pac.Graph.load_text(ctx, path)
            wall.append((time.perf_counter() - t0) * 1e3)
            if i == 0:
                vl1, el1 = g1.csr()
            g1.free()
        t0 = time.perf_counter()
        rt = O.text_to_csr(path)  # the reference's text loader, restated
        cpu_s = time.perf_counter() - t0
        same = bool(np.array_equal(rt.vlist, vl1) and np.array_equal(rt.elist, el1))
        res["ingest_text"] = {"bytes": os.path.getsize(path), "edges": int(k18.sum()),
                              "wall_ms": round(float(np.median(wall)), 1),
                              "includes": "file read + device parse + device sort/dedup/CSR",
                              "csr_identical": same, "cpu_s": round(cpu_s, 2), "cpu_threads": 1}
    finally:
        os.remove(path)

    # (2) pairwise counts; algorithmic bytes 4 (d(u) + d(v)) per pair
    def pairs(name, u, v):
        g.cnt_common_pairs(u[:1000], v[:1000])
        ms = []
        for _ in range(3):
            out = g.cnt_common_pairs(u, v)
            ms.append(ctx.last_timing().total_ms)
        t = float(np.median(ms))
        nbytes = 4.0 * float((deg[u] + deg[v]).sum())
        idx = rng.choice(len(u), size=min(sample, len(u)), replace=False)
        t0 = time.perf_counter()
        ri, _, _ = O.jaccard_edges(ref, u[idx], v[idx], nt)
        dt = time.perf_counter() - t0
        res[name] = {"pairs": int(len(u)), "device_ms": round(t, 3),
                     "pairs_per_s": round(len(u) / (t * 1e-3), 1),
                     "alg_GBps": round(nbytes / (t * 1e-3) / 1e9, 1),
                     "roofline_frac": round(nbytes / (t * 1e-3) / 8e12, 4),
                     "sample_exact": bool(np.array_equal(ri.astype(np.uint64), out[idx])),
                     "cpu_pairs_per_s": round(len(idx) / dt, 1)}
    u = rng.integers(0, V, npairs).astype(np.int32)
    v = rng.integers(0, V, npairs).astype(np.int32)
    pairs("pairs_uniform", u, v)
    e = rng.integers(0, len(el), npairs)
    u = (np.searchsorted(vli, e, side="right") - 1).astype(np.int32)
    v = el[e].astype(np.int32)
    pairs("pairs_adjacent", u, v)

    # (3) 3-way counts: u, and v, w drawn from N(u)
    nz = np.flatnonzero(deg >= 2)
    qu = rng.choice(nz, nkway).astype(np.int64)
    qv = el[vli[qu] + (rng.random(nkway) * deg[qu]).astype(np.int64)]
    qw = el[vli[qu] + (rng.random(nkway) * deg[qu]).astype(np.int64)]
    queries = np.stack([qu.astype(np.int32), qv, qw], 1)
    ql = [q for q in queries]
    g.cnt_common_kway(ql[:1000])
    ms = []
    for _ in range(3):
        out = g.cnt_common_kway(ql)
        ms.append(ctx.last_timing().total_ms)
    t = float(np.median(ms))
    ns = min(20000, nkway)
    t0 = time.perf_counter()
    rk = np.array([O.kway_merge(ref, q) for q in queries[:ns]], np.uint64)
    dt = time.perf_counter() - t0
    res["kway3"] = {"queries": nkway, "device_ms": round(t, 3),
                    "queries_per_s": round(nkway / (t * 1e-3), 1),
                    "sample_exact": bool(np.array_equal(rk, out[:ns])),
                    "cpu_queries_per_s": round(ns / dt, 1), "cpu_threads": 1}

    # (4) friend triangle: a root of degree about 1000, its first 1024 neighbours
    root = int(np.argmin(np.abs(deg - 1000)))
    nodes = el[vli[root]:vli[root] + min(1024, deg[root])].astype(np.int32)
    g.friend_triangle(root, nodes)
    ms = []
    for _ in range(3):
        ft = g.friend_triangle(root, nodes)
        ms.append(ctx.last_timing().total_ms)
    t0 = time.perf_counter()
    rf = O.friend_triangle_merge(ref, root, nodes, nt)
    dt = time.perf_counter() - t0
    res["friend_triangle"] = {"root_degree": int(deg[root]), "nodes": int(len(nodes)),
                              "device_ms": round(float(np.median(ms)), 3), "value": ft,
                              "exact": ft == rf, "cpu_s": round(dt, 3)}

    # (5) squares
    sq = {}
    for sc in (12, 16):
        gq = O.gen_rmat(sc, 16, 0.57, 0.19, 0.19, sc)
        gg = pac.Graph.from_csr(ctx, gq.vlist, gq.elist)
        gg.square_count()
        ms = []
        for _ in range(3):
            c = gg.square_count()
            ms.append(ctx.last_timing().total_ms)
        ent = {"squares": c, "device_ms": round(float(np.median(ms)), 3)}
        if sc == 12:
            t0 = time.perf_counter()
            want = O.PGraph(gq, O.contiguous_partition(gq.V, 4)).square_counting()
            ent["cpu_s"] = round(time.perf_counter() - t0, 3)
            ent["exact"] = c == want
        sq[f"rmat{sc}"] = ent
        gg.free()
    res["squares"] = sq
    return res


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("which", choices=["scale22", "batch", "jaccard", "next", "cpu"])
    ap.add_argument("--graphs", type=int, default=256)
    ap.add_argument("--scale", type=int, default=None)
    ap.add_argument("--no-check", action="store_true")
    ap.add_argument("--profile-only", action="store_true",
                    help="jaccard / next: only three timed solves / pair counts (for rocprofv3)")
    a = ap.parse_args(argv)
    ctx = pac.Context(0)
    if a.which == "scale22":
        r = exp_scale22(ctx, a.scale or 22)
    elif a.which == "batch":
        r = exp_batch(ctx, a.graphs, not a.no_check)
    elif a.which == "next":
        r = exp_next(ctx, a.scale or 22, profile_only=a.profile_only)
    elif a.which == "cpu":
        r = exp_cpu(ctx, only=a.scale)
    else:
        r = exp_jaccard(ctx, a.scale or 25, profile_only=a.profile_only)
    print(json.dumps(r))


if __name__ == "__main__":
    main()
