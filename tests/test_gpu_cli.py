"""The pac_main CLI with the reference's argv protocol (src/test.cpp:301-389): same
files, same output lines, counts equal to the oracle's."""
import os
import re
import subprocess

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

CLI = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))),
                   "partitionandconquer_amd", "pac_main")


def run(*args, env=None):
    r = subprocess.run([CLI, *map(str, args)], capture_output=True, text=True, timeout=120,
                       env=env)
    assert r.returncode == 0, r.stderr
    return r.stdout


def test_tc_argv_protocol(orc, tmp_path):
    g = orc.gen_rmat(11, 16, 0.57, 0.19, 0.19, seed=77)
    d = tmp_path / "g"
    (d / "origin").mkdir(parents=True)
    g.vlist.tofile(d / "origin" / "csr_vlist.bin")
    g.elist.tofile(d / "origin" / "csr_elist.bin")
    T = orc.tc_merge(g, 4)
    out = run("tc", d, "merge", 0, "Greedy", 32, 32)
    assert re.search(r"Time used for paralle merge: [0-9.]+ (ns|us|ms|s)\.", out)
    assert f"#Triangles: {T}." in out
    out = run("tc", d, "galloping", 0, "Greedy", 32, 32)
    assert f"#Triangles: {6 * T}." in out                 # Graph.cpp:342-353 counts 6T
    # partition verb writes P<n>.bin; the partitioned tc reads it
    run("partition", d, 8, d / "partition" / "Greedy")
    v2p = np.fromfile(d / "partition" / "Greedy" / "P8.bin", np.int32)
    assert len(v2p) == g.V and v2p.min() >= 0 and v2p.max() < 8
    out = run("tc", d, "merge", 1, "Greedy", 8, 32)
    assert re.search(r"Time used for partitioned merge: ", out)
    assert f"#Triangles: {T}." in out
    # lcc verb: doubles bit-identical to Graph::solve_lcc_merge
    run("lcc", d, "merge", 1, "Greedy", 8, 32, tmp_path / "lcc.bin")
    got = np.fromfile(tmp_path / "lcc.bin", np.float64)
    assert np.array_equal(got.view(np.uint64), orc.lcc_merge(g, 4).view(np.uint64))
    # jaccard verb
    run("jaccard", d, 0, "Greedy", 8, 32, tmp_path / "jac")
    src = np.fromfile(tmp_path / "jac" / "src.bin", np.int32)
    dst = np.fromfile(tmp_path / "jac" / "dst.bin", np.int32)
    inter = np.fromfile(tmp_path / "jac" / "inter.bin", np.uint32)
    ri, _, _ = orc.jaccard_edges(g, src, dst, 4)
    assert np.array_equal(inter, ri)


def test_preprocess_and_gen(orc, tmp_path):
    golden = os.path.join(os.path.dirname(__file__), "golden")
    d = str(tmp_path) + "/"
    (tmp_path / "e.txt").write_text(open(os.path.join(golden, "text_1based.txt")).read())
    run("preprocess", d, "e.txt")                        # graphdir + graphfile (preprocess.cpp:43)
    ref = orc.text_to_csr(os.path.join(golden, "text_1based.txt"))
    assert np.array_equal(np.fromfile(tmp_path / "origin" / "csr_vlist.bin", np.uint64), ref.vlist)
    assert np.array_equal(np.fromfile(tmp_path / "origin" / "csr_elist.bin", np.int32), ref.elist)
    run("gen-rmat", 9, 16, 0.57, 0.19, 0.19, 5, tmp_path / "r")
    want = orc.gen_rmat(9, 16, 0.57, 0.19, 0.19, seed=5)
    assert np.array_equal(np.fromfile(tmp_path / "r" / "origin" / "csr_elist.bin", np.int32), want.elist)


def test_missing_partition_file_fails_cleanly(tmp_path, orc):
    g = orc.gen_rmat(6, 4, 0.57, 0.19, 0.19, seed=1)
    d = tmp_path / "g"
    (d / "origin").mkdir(parents=True)
    g.vlist.tofile(d / "origin" / "csr_vlist.bin")
    g.elist.tofile(d / "origin" / "csr_elist.bin")
    r = subprocess.run([CLI, "tc", str(d), "merge", "1", "Greedy", "4", "1"], capture_output=True,
                       text=True, timeout=60)
    assert r.returncode != 0 and "cannot open" in r.stderr
    r = subprocess.run([CLI, "tc", str(tmp_path / "nope"), "merge", "0", "x", "1", "1"],
                       capture_output=True, text=True, timeout=60)
    assert r.returncode != 0 and "error" in r.stderr
