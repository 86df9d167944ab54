"""CPU tests of host-side logic that needs no GPU: the seeded workload specifications of
the batch and learned-dispatch experiments, and the bench's one-GPU contract."""
import os
import subprocess
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_batch_specs_follow_the_config():
    from experiments import batch_specs
    s = batch_specs(256)
    assert s == batch_specs(256)                       # deterministic
    er = [x for x in s if x["kind"] == "er"]
    rm = [x for x in s if x["kind"] == "rmat"]
    assert len(er) == 128 and len(rm) == 128
    assert all(2 ** 10 - 1 <= x["n"] <= 2 ** 18 + 1 for x in er)
    assert all(8 <= x["k"] <= 64 for x in er)
    assert all(10 <= x["scale"] <= 18 for x in rm)
    assert [x["seed"] for x in s] == list(range(3000, 3256))
    # log-uniform sizes spread over the whole range
    logs = np.log2([x["n"] for x in er])
    assert logs.min() < 11 and logs.max() > 17
    assert len({x["scale"] for x in rm}) == 9


def test_training_specs_follow_the_config():
    from partitionandconquer_amd.train_dispatch import training_specs
    s = training_specs(64)
    assert s == training_specs(64)
    assert [x["seed"] for x in s] == list(range(4000, 4064))
    assert [x["scale"] for x in s] == [18 + i % 5 for i in range(64)]
    for x in s:
        assert 0.45 <= x["a"] <= 0.65
        assert abs(x["a"] + x["b"] + x["c"] + 0.05 - 1.0) < 1e-12 and x["b"] == x["c"]


def test_labels_and_scores():
    from partitionandconquer_amd.train_dispatch import labels_weights, score
    ms = np.array([[3.0, 2.0, 1.0], [1.0, 5.0, 2.0], [2.0, 2.0, 2.0]])
    y, w = labels_weights(ms)
    assert y.tolist() == [2, 0, 0]
    assert np.isclose(w.mean(), 1.0) and w[2] == 0.0
    sc = score(np.array([2, 1, 0]), ms)
    assert sc["mispredict"] == 1 / 3
    assert np.isclose(sc["regret"], (0 + 4 + 0) / (1 + 1 + 2))


def test_bench_refuses_more_than_one_gpu():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "8"],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 2 and "one GPU" in r.stderr


def test_bench_traffic_comes_from_the_committed_profile():
    sys.path.insert(0, ROOT)
    import bench
    t = bench.measured_traffic(24, 4096)
    assert t is not None and t > 1e11        # hundreds of GB per launch at scale 24
    assert bench.measured_traffic(22, 1024) is None


def test_cpu_baseline_sampling():
    """The CPU baseline's stratified sample: groups of equal estimated work, every group
    priced, hub-free graphs timed in full when the budget allows; empty graphs cost 0."""
    sys.path.insert(0, ROOT)
    import bench
    from oracle import oracle as O
    g = O.gen_rmat(12, 16, 0.57, 0.19, 0.19, seed=12)
    est, rows, nt, spent = bench.cpu_baseline(g.vlist, g.elist, budget_s=0.5, seed=1)
    assert est > 0 and 0 < rows <= g.V and nt >= 1 and spent > 0
    # a budget far above the graph's cost times every non-isolated row once; the estimate
    # (the work spread evenly over the threads) is then at most the time spent
    nonempty = int((np.diff(g.vlist) > 0).sum())
    est2, rows2, _, spent2 = bench.cpu_baseline(g.vlist, g.elist, budget_s=1e3, seed=2)
    assert rows2 == nonempty and 0 < est2 <= spent2 * (1 + 1e-9)
    empty = O.edges_to_csr(50, [], [])
    assert bench.cpu_baseline(empty.vlist, empty.elist, budget_s=0.1)[0] == 0.0


def test_bounded_wait_on_the_cpu(tmp_path):
    """The solve's bounded wait (pac_wait.hpp) with stand-in completion queries: completion,
    PAC_ERR_TIMEOUT after the deadline, failures passed through, no bound at 0."""
    import subprocess
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    exe = str(tmp_path / "wait_check")
    subprocess.run(["g++", "-std=c++17", "-O1", "-I", os.path.join(root, "partitionandconquer_amd", "csrc"),
                    os.path.join(root, "tests", "wait_check.cpp"), "-o", exe], check=True)
    r = subprocess.run([exe], capture_output=True, text=True)
    assert r.returncode == 0, r.stdout + r.stderr
    assert r.stdout.count("ok ") == 6


def test_oracle_under_address_sanitizer():
    """`make -C oracle asan`: the oracle built with ASan + UBSan, its functions run on small
    R-MAT graphs (including the partitioned threshold merge, where the reference reads past a
    sub-list), and the formulations that must agree do."""
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    r = subprocess.run(["make", "-s", "-B", "-C", os.path.join(root, "oracle"), "asan"],
                       capture_output=True, text=True)
    assert r.returncode == 0, r.stdout[-3000:] + r.stderr[-3000:]
    assert "asan check ok" in r.stdout
