"""Train the learned strategy selector and evaluate it against the fixed heuristic.

    python -m partitionandconquer_amd.train_dispatch [--graphs 64] [--out dispatch_mlp.bin]
    python -m partitionandconquer_amd.train_dispatch --eval-only --eval-scale 24

Training data: 64 seeded R-MAT graphs, graph i has seed 4000+i, scale 18 + (i mod 5),
a drawn uniformly from [0.45, 0.65] (SplitMix64 stream seeded 4), d = 0.05 and
b = c = (1-a-d)/2; edge factor 16; contiguous partitions of 4096 vertex IDs (the
partition width of the scale-24 / 4096-partition evaluation).  Every strategy is timed
on every partition inside the persistent launch the solves use (pac_profile_partitions:
per-item clock stamps from the kernel, strategies rotated over the partitions so every
launch mixes them, median of 3 rounds); the label is the fastest strategy and the sample
weight is the time spread (max - min) normalised to mean 1, so partitions where the
choice matters weigh more.  The MLP is trained on the GPU by the library's own backward
and Adam kernels (pac_mlp_train).

Evaluation: on held-out R-MAT graphs (scale 20, a in {0.45, 0.5, 0.6, 0.65}, seeds
9000+, not in the training set) and on the headline graph (default R-MAT scale 24, 4096
partitions): misprediction rate (choice != argmin of the measured times) and regret
sum(t_chosen - t_best)/sum(t_best) for the MLP and for the fixed heuristic, plus
end-to-end device times of the batched persistent kernel (hipGraph) under both
dispatches and of the plain per-partition launches.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))

import partitionandconquer_amd as pac  # noqa: E402

GAMMA = 0x9E3779B97F4A7C15
MASK = (1 << 64) - 1


def mix64(z):
    z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & MASK
    z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & MASK
    return z ^ (z >> 31)


def training_specs(n=64, master_seed=4):
    st = mix64((master_seed + GAMMA) & MASK)
    out = []
    for i in range(n):
        st = (st + GAMMA) & MASK
        u = (mix64(st) >> 11) / float(1 << 53)
        a = 0.45 + 0.20 * u
        d = 0.05
        b = c = (1.0 - a - d) / 2.0
        out.append(dict(seed=4000 + i, scale=18 + i % 5, a=a, b=b, c=c))
    return out


def log(*a):
    print(*a, flush=True)


def profile_graph(ctx, spec, width=4096, reps=3):
    g = pac.Graph.rmat(ctx, spec["scale"], 16, spec["a"], spec["b"], spec["c"], seed=spec["seed"])
    V = g.get_vertex_num()
    P = max(16, V // width)
    g.prepare(P)
    feats = g.features()
    ms = g.profile_partitions(reps)
    fixed = g.fixed_choice()
    keep = ms.max(1) > 0
    return g, feats[keep], ms[keep], fixed[keep]


def labels_weights(ms):
    y = ms.argmin(1).astype(np.int32)
    spread = ms.max(1) - ms.min(1)
    w = spread / max(spread.mean(), 1e-12)
    return y, w.astype(np.float32)


def score(choice, ms):
    best = ms.min(1)
    chosen = ms[np.arange(len(ms)), choice]
    return dict(mispredict=float(np.mean(choice != ms.argmin(1))),
                regret=float((chosen - best).sum() / max(best.sum(), 1e-12)),
                chosen_ms=float(chosen.sum()), best_ms=float(best.sum()))


HELD_OUT = [dict(seed=9000 + i, scale=20, a=a, b=(0.95 - a) / 2, c=(0.95 - a) / 2)
            for i, a in enumerate((0.45, 0.5, 0.6, 0.65))]


def evaluate(ctx, scale, partitions, steps=5, abc=(0.57, 0.19, 0.19), seed=None):
    g = pac.Graph.rmat(ctx, scale, 16, *abc, seed=scale if seed is None else seed)
    g.prepare(partitions)
    ms = g.profile_partitions(3)
    keep = ms.max(1) > 0
    fixed = g.fixed_choice()
    learned = g.mlp_choice()
    res = {"scale": scale, "partitions": partitions,
           "fixed": score(fixed[keep], ms[keep]), "learned": score(learned[keep], ms[keep]),
           "learned_strategy_counts": np.bincount(learned[keep], minlength=3).tolist(),
           "fixed_strategy_counts": np.bincount(fixed[keep], minlength=3).tolist(),
           "best_strategy_counts": np.bincount(ms[keep].argmin(1), minlength=3).tolist()}
    T = g.tc(pac.FORCE_BITMAP)
    times = {}
    for name, d, l in (("plain_fixed", pac.DISPATCH_FIXED, pac.LAUNCH_PLAIN),
                       ("batched_fixed", pac.DISPATCH_FIXED, pac.LAUNCH_GRAPH),
                       ("batched_learned", pac.DISPATCH_LEARNED, pac.LAUNCH_GRAPH),
                       ("batched_bitmap", pac.FORCE_BITMAP, pac.LAUNCH_GRAPH)):
        assert g.lcc_device(d, l) == T
        v = []
        for _ in range(steps):
            g.lcc_device(d, l)
            v.append(ctx.last_timing().total_ms)
        times[name] = float(np.median(v))
    res["device_ms"] = times
    res["speedup_learned_batched_vs_plain_fixed"] = times["plain_fixed"] / times["batched_learned"]
    res["speedup_learned_vs_fixed_batched"] = times["batched_fixed"] / times["batched_learned"]
    res["triangles"] = T
    g.free()
    return res


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--graphs", type=int, default=64)
    ap.add_argument("--epochs", type=int, default=3000)
    ap.add_argument("--lr", type=float, default=3e-3)
    ap.add_argument("--out", default=os.path.join(HERE, "dispatch_mlp.bin"))
    ap.add_argument("--report", default=None)
    ap.add_argument("--eval-scale", type=int, default=24)
    ap.add_argument("--eval-partitions", type=int, default=4096)
    ap.add_argument("--eval-only", action="store_true")
    ap.add_argument("--no-eval", action="store_true")
    a = ap.parse_args(argv)
    ctx = pac.Context(0)
    report = {}
    if not a.eval_only:
        X, Y, Wt, MS, FX, gid = [], [], [], [], [], []
        t0 = time.time()
        for i, spec in enumerate(training_specs(a.graphs)):
            g, f, ms, fx = profile_graph(ctx, spec)
            y, w = labels_weights(ms)
            X.append(f); Y.append(y); Wt.append(w); MS.append(ms); FX.append(fx)
            gid.append(np.full(len(y), i))
            log(f"graph {i:2d} scale {spec['scale']} a={spec['a']:.3f}: {len(y)} partitions, "
                f"best counts {np.bincount(y, minlength=3).tolist()}, {time.time() - t0:.0f}s")
            g.free()
        X = np.concatenate(X); Y = np.concatenate(Y); Wt = np.concatenate(Wt)
        MS = np.concatenate(MS); FX = np.concatenate(FX); gid = np.concatenate(gid)
        val = gid % 8 == 7  # every 8th graph held out
        ctx.mlp_init(1)
        loss = ctx.mlp_train(X[~val], Y[~val], Wt[~val], epochs=a.epochs, lr=a.lr)
        z_tr = ctx.mlp_forward(X[~val]).argmax(1)
        z_va = ctx.mlp_forward(X[val]).argmax(1)
        report["train"] = {"samples": int((~val).sum()), "loss": loss,
                           "learned": score(z_tr, MS[~val]), "fixed": score(FX[~val], MS[~val])}
        report["validation"] = {"samples": int(val.sum()),
                                "learned": score(z_va, MS[val]), "fixed": score(FX[val], MS[val])}
        log(json.dumps(report, indent=1))
        # final model on all graphs
        ctx.mlp_init(1)
        report["final_loss"] = ctx.mlp_train(X, Y, Wt, epochs=a.epochs, lr=a.lr)
        ctx.mlp_save(a.out)
        log(f"saved {a.out}")
    else:
        ctx.mlp_load(a.out)
    if not a.no_eval:
        report["held_out"] = []
        for sp in HELD_OUT:
            r = evaluate(ctx, sp["scale"], max(16, (1 << sp["scale"]) // 4096),
                         abc=(sp["a"], sp["b"], sp["c"]), seed=sp["seed"])
            r["a"] = sp["a"]
            report["held_out"].append(r)
            log(json.dumps({k: r[k] for k in ("a", "fixed", "learned",
                                               "speedup_learned_vs_fixed_batched")}))
        report["evaluation"] = evaluate(ctx, a.eval_scale, a.eval_partitions)
        log(json.dumps(report["evaluation"], indent=1))
    if a.report:
        with open(a.report, "w") as f:
            json.dump(report, f, indent=1)
    return report


if __name__ == "__main__":
    main()
