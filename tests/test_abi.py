"""CPU checks of the C-ABI boundary: the library loads, and every entry point that
include/pac.h declares is exported and bound by the Python binding.  No compute
call is made here (there is no GPU in this environment)."""
import ctypes
import os
import re
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "pac.h")


def declared():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(pac_[a-z0-9_]+)\s*\(", text)))


def test_header_declares_entry_points():
    names = declared()
    for must in ("pac_ctx_create", "pac_graph_upload", "pac_graph_prepare", "pac_tc", "pac_lcc",
                 "pac_jaccard_edges", "pac_jaccard_queries", "pac_tc_batch", "pac_mlp_train",
                 "pac_autotune", "pac_profile_partitions"):
        assert must in names


def test_library_exports_every_declared_symbol():
    import partitionandconquer_amd as pac
    L = pac.lib()
    for name in declared():
        assert hasattr(L, name), name
    out = subprocess.run(["nm", "-D", "--defined-only", pac.pac.LIB_PATH], capture_output=True,
                         text=True, check=True).stdout
    exported = set(re.findall(r" T (pac_[a-z0-9_]+)", out))
    assert set(declared()) <= exported


def test_binding_covers_header():
    import partitionandconquer_amd.pac as P
    src = open(P.__file__).read()
    for name in declared():
        assert f'"{name}"' in src, f"{name} has no ctypes signature"


def test_constants_match_header():
    import partitionandconquer_amd.pac as P
    text = open(HEADER).read()
    assert int(re.search(r"#define PAC_MLP_IN (\d+)", text).group(1)) == P.MLP_IN
    assert int(re.search(r"#define PAC_NUM_TILES (\d+)", text).group(1)) == P.NUM_TILES
    assert P.MLP_NPARAMS == 5443
    for k, v in {"PAC_ERR_ARG": -1, "PAC_ERR_STATE": -6, "PAC_ERR_FORMAT": -4}.items():
        assert int(re.search(rf"#define {k} \((-?\d+)\)", text).group(1)) == v


def test_last_error_callable_without_gpu():
    import partitionandconquer_amd as pac
    assert isinstance(pac.lib().pac_last_error(), bytes)


def test_library_does_not_link_the_oracle():
    import partitionandconquer_amd as pac
    out = subprocess.run(["ldd", pac.pac.LIB_PATH], capture_output=True, text=True).stdout
    assert "oracle" not in out
    syms = subprocess.run(["nm", "-D", pac.pac.LIB_PATH], capture_output=True, text=True).stdout
    assert "orc_" not in syms


def test_random_partition_host_function():
    """pac_random_partition is host-only: seeded, in range, deterministic."""
    import numpy as np
    import partitionandconquer_amd as pac
    a = pac.random_partition(1000, 7, seed=3)
    b = pac.random_partition(1000, 7, seed=3)
    assert np.array_equal(a, b) and a.min() >= 0 and a.max() < 7
    assert len(np.unique(a)) == 7
