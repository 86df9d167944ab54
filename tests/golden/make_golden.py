"""Writes the golden fixtures of tests/golden/ (run: python tests/golden/make_golden.py).

The reference ships no tests, fixtures or golden outputs, and it cannot be built in
this environment (src/Graph.hpp:5 includes Boost, which is absent).  The fixtures are
therefore computed here by dense linear algebra, independently of the oracle's merge
loops:

  T      = trace(A^3) / 6
  t(u)   = diag(A^3)_u / 2                          (triangles through u)
  LCC(u) = diag(A^3)_u / (d(d-1)) in double, 0.0 for d=0, NaN for d=1
           (the conventions of Graph::solve_lcc_merge, Graph.cpp:828-840)
  inter(u,v) = (A^2)_uv, union = d_u + d_v - inter    (Graph.cpp:808-809)
  jsim(q_i,q_j) = inter / ((double)(d_i + d_j) - inter)

Graphs: closed-form families plus seeded random graphs built with numpy only.
"""
import json
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))


def csr_from_edges(n, edges):
    e = np.asarray(edges, dtype=np.int64).reshape(-1, 2)
    e = e[e[:, 0] != e[:, 1]]
    both = np.concatenate([e, e[:, ::-1]])
    keys = np.unique(both[:, 0] * n + both[:, 1])
    src, dst = keys // n, keys % n
    vl = np.zeros(n + 1, np.uint64)
    np.add.at(vl, src + 1, 1)
    return np.cumsum(vl).astype(np.uint64), dst.astype(np.int32)


def dense(vl, el):
    n = len(vl) - 1
    A = np.zeros((n, n), np.int64)
    src = np.repeat(np.arange(n), np.diff(vl).astype(np.int64))
    A[src, el] = 1
    return A


def expected(vl, el, queries):
    A = dense(vl, el)
    A2 = A @ A
    A3 = A2 @ A
    d = A.sum(1)
    diag = np.diag(A3)
    T = int(np.trace(A3) // 6)
    lcc = np.zeros(len(d), np.float64)
    with np.errstate(invalid="ignore", divide="ignore"):
        nz = d >= 1
        lcc[nz] = diag[nz].astype(np.float64) / (d[nz] * (d[nz] - 1)).astype(np.float64)
    src = np.repeat(np.arange(len(d)), np.diff(vl).astype(np.int64))
    keep = src < el
    eu, ev = src[keep].astype(np.int32), el[keep]
    inter = A2[eu, ev].astype(np.uint32)
    uni = (d[eu] + d[ev] - inter).astype(np.uint32)
    ratio = (inter.astype(np.float64) / uni.astype(np.float64)).astype(np.float32)
    q = np.asarray(queries, np.int64)
    c = A2[np.ix_(q, q)].astype(np.float64)
    with np.errstate(invalid="ignore", divide="ignore"):
        jsim = c / ((d[q][:, None] + d[q][None, :]).astype(np.float64) - c)
    return dict(T=np.uint64(T), t=(diag // 2).astype(np.uint64), lcc=lcc, eu=eu, ev=ev,
                inter=inter, uni=uni, ratio=ratio, queries=q.astype(np.int32), jsim=jsim)


def families():
    out = {}
    n = 12
    out["K12"] = (n, [(i, j) for i in range(n) for j in range(i + 1, n)])
    w = 9  # wheel: hub 0 + 9-cycle
    out["wheel9"] = (w + 1, [(0, i) for i in range(1, w + 1)] +
                     [(i, i % w + 1) for i in range(1, w + 1)])
    out["cycle7"] = (7, [(i, (i + 1) % 7) for i in range(7)])
    out["K3_5"] = (8, [(i, j) for i in range(3) for j in range(3, 8)])
    k = 6  # friendship graph F_6: hub 0 with 6 triangles
    out["friend6"] = (2 * k + 1, [(0, 2 * i + 1) for i in range(k)] + [(0, 2 * i + 2) for i in range(k)] +
                      [(2 * i + 1, 2 * i + 2) for i in range(k)])
    out["star_path"] = (8, [(0, 1), (0, 2), (0, 3), (3, 4), (4, 5), (5, 3)])  # d=1 leaves, isolated 6,7
    rng = np.random.default_rng(20240601)
    for name, nn, p in (("er60", 60, 0.2), ("er200", 200, 0.08)):
        iu = np.triu_indices(nn, 1)
        mask = rng.random(len(iu[0])) < p
        out[name] = (nn, list(zip(iu[0][mask], iu[1][mask])))
    # power-law-ish: preferential pairs drawn from a Zipf-weighted distribution
    nn = 256
    wts = 1.0 / np.arange(1, nn + 1) ** 0.9
    wts /= wts.sum()
    a = rng.choice(nn, size=3000, p=wts)
    b = rng.choice(nn, size=3000, p=wts)
    out["zipf256"] = (nn, list(zip(a, b)))
    return out


def main():
    meta = {}
    for name, (n, edges) in families().items():
        vl, el = csr_from_edges(n, edges)
        deg = np.diff(vl).astype(np.int64)
        queries = np.unique(np.concatenate([np.argsort(-deg)[:4], np.arange(min(n, 5))]))
        exp = expected(vl, el, queries)
        np.savez_compressed(os.path.join(HERE, f"{name}.npz"), vlist=vl, elist=el, **exp)
        meta[name] = {"V": int(n), "nnz": int(vl[-1]), "T": int(exp["T"])}
    with open(os.path.join(HERE, "golden_index.json"), "w") as f:
        json.dump(meta, f, indent=1, sort_keys=True)
    print(json.dumps(meta, indent=1, sort_keys=True))


if __name__ == "__main__":
    main()
