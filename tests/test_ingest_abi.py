"""Host-side tests (no GPU): Matrix Market ingest of both the library and the
oracle against hand-derived results, the C ABI's exported symbols, and the
`tribit` command's argument handling (tribit.cu:271-315)."""
import os
import subprocess

import numpy as np
import pytest

import tribit_amd
from tests import graphs
from tests.conftest import GOLDEN

WIKI = os.path.join(GOLDEN, "wiki-Vote.mtx")


@pytest.fixture(scope="module")
def lib_built():
    if not os.path.exists(tribit_amd.LIB_PATH):
        pytest.skip("libtribit_hip.so not built")
    return tribit_amd.load_library()


def write(tmp_path, name, text):
    p = tmp_path / name
    p.write_bytes(text.encode())
    return str(p)


GENERAL = """%%MatrixMarket matrix coordinate real general
% comment line
%
4 4 6
1 2 1.5
2 1 2.0
3 3 1.0
2 3 -1
4 1 0.5e3
1 3 7
"""
# unique undirected edges {0,1},{0,2},{0,3},{1,2}; self-loop (3,3) dropped; 1 triangle
GENERAL_KEYS = [graphs.pack(0, 1), graphs.pack(0, 1), graphs.pack(1, 2), graphs.pack(0, 3), graphs.pack(0, 2)]

SYMMETRIC = """%%MatrixMarket matrix coordinate pattern symmetric
4 4 4
2 1
3 1
3 2
4 1
"""

CASES = {
    "general": (GENERAL, 0, 4, 6, GENERAL_KEYS),
    "symmetric": (SYMMETRIC, 0, 4, 4, [graphs.pack(0, 1), graphs.pack(0, 2), graphs.pack(1, 2), graphs.pack(0, 3)]),
    "crlf_blank": ("%%MatrixMarket matrix coordinate pattern general\r\n%c\r\n\r\n3 3 2\r\n1 2\r\n3 2\r\n", 0, 3, 2,
                   [graphs.pack(0, 1), graphs.pack(1, 2)]),
    "stop_at_percent": ("%%MM\n3 3 3\n1 2\n% trailing comment ends the read\n2 3\n", 0, 3, 1, [graphs.pack(0, 1)]),
    "blank_between_entries": ("%%MM\n3 3 3\n1 2\n\n  \t\r\n2 3\nx 1\n3 1\n", 0, 3, 2,
                              [graphs.pack(0, 1), graphs.pack(1, 2)]),
    "values_multi": ("%%MM complex\n3 3 2\n1 2 1.0 2.0\n3 1 -4 5\n", 0, 3, 2, [graphs.pack(0, 1), graphs.pack(0, 2)]),
    "nonsquare": ("%%MM\n3 4 1\n1 2\n", 3, None, None, None),
    "id_zero": ("%%MM\n3 3 1\n0 2\n", 4, None, None, None),
    "id_too_big": ("%%MM\n3 3 2\n1 2\n1 4\n", 4, None, None, None),
    "no_size": ("%%MM\n% only comments\n", 2, None, None, None),
    "bad_size": ("%%MM\nabc\n", 2, None, None, None),
    "empty": ("", 2, None, None, None),
}


@pytest.mark.parametrize("case", sorted(CASES))
def test_mtx_oracle(oracle_mod, tmp_path, case):
    text, st, n, ne, keys = CASES[case]
    r = oracle_mod.read_mtx(write(tmp_path, case + ".mtx", text))
    assert r[0] == st
    if st == 0:
        assert (r[2], r[3]) == (n, ne)
        assert np.array_equal(r[1], np.array([int(k) for k in keys], np.uint64))


@pytest.mark.parametrize("case", sorted(CASES))
def test_mtx_library(lib_built, tmp_path, case):
    text, st, n, ne, keys = CASES[case]
    path = write(tmp_path, case + ".mtx", text)
    if st != 0:
        with pytest.raises(tribit_amd.TribitError) as ei:
            tribit_amd.read_mtx(path)
        assert ei.value.status == st
        return
    k, nn, e = tribit_amd.read_mtx(path)
    assert (nn, e) == (n, ne)
    assert np.array_equal(k, np.array([int(x) for x in keys], np.uint64))


def test_mtx_missing_file(lib_built, oracle_mod):
    assert oracle_mod.read_mtx("/nonexistent/x.mtx")[0] == 1
    with pytest.raises(tribit_amd.TribitError) as ei:
        tribit_amd.read_mtx("/nonexistent/x.mtx")
    assert ei.value.status == 1


def test_mtx_wiki_vote_library_equals_oracle(lib_built, oracle_mod):
    k, n, ne = tribit_amd.read_mtx(WIKI)
    st, k2, n2, ne2 = oracle_mod.read_mtx(WIKI)
    assert st == 0 and (n, ne) == (n2, ne2) == (8297, 103689)
    assert np.array_equal(k, k2)


def test_mtx_parallel_slices_equal_serial(lib_built, oracle_mod, tmp_path):
    """A file large enough for the multi-threaded parse (> 1 MiB) reads identically."""
    keys = oracle_mod.gen_rmat(14, 16, seed=4)
    keys = keys[keys != np.uint64(0xFFFFFFFFFFFFFFFF)]
    a = (keys >> np.uint64(32)).astype(np.int64) + 1
    b = (keys & np.uint64(0xFFFFFFFF)).astype(np.int64) + 1
    swap = np.arange(a.size) % 3 == 0
    r, c = np.where(swap, b, a), np.where(swap, a, b)
    body = "\n".join(f"{x} {y} 1.0" for x, y in zip(r.tolist(), c.tolist()))
    path = write(tmp_path, "big.mtx", f"%%MatrixMarket matrix coordinate real general\n{1 << 14} {1 << 14} {a.size}\n{body}\n")
    assert os.path.getsize(path) > (1 << 20)
    k, n, ne = tribit_amd.read_mtx(path)
    st, k2, n2, ne2 = oracle_mod.read_mtx(path)
    assert st == 0 and ne == ne2 == a.size and n == n2
    assert np.array_equal(k, k2)
    assert np.array_equal(k, keys)


def test_abi_exports_every_declared_symbol(lib_built):
    syms = tribit_amd.header_symbols()
    assert len(syms) >= 15
    missing = [s for s in syms if not hasattr(lib_built, s)]
    assert not missing, missing


def test_abi_status_strings(lib_built):
    for st in range(10):
        assert lib_built.tribit_status_str(st)
    assert b"gfx950" in lib_built.tribit_build_info()


def run_cli(*args):
    return subprocess.run([tribit_amd.CLI_PATH, *args], capture_output=True, text=True, timeout=60)


@pytest.fixture(scope="module")
def cli_built(lib_built):
    if not os.path.exists(tribit_amd.CLI_PATH):
        pytest.skip("tribit command not built")


def test_cli_no_args_prints_usage(cli_built):
    r = run_cli()
    assert r.returncode == 1 and "Usage:" in r.stderr and "-i <input_file>" in r.stderr


def test_cli_unknown_option(cli_built):
    r = run_cli("-x")
    assert r.returncode == 1 and "unknown option" in r.stderr and "Usage:" in r.stderr


def test_cli_missing_input(cli_built):
    r = run_cli("--sched", "plain")
    assert r.returncode == 1 and "input file is required" in r.stderr


def test_cli_missing_argument(cli_built):
    r = run_cli("-i")
    assert r.returncode == 1 and "Usage:" in r.stderr


def test_cli_bad_file_fails_before_device(cli_built, tmp_path):
    r = run_cli("-i", "/nonexistent/graph.mtx")
    assert r.returncode == 2 and "Reading data from /nonexistent/graph.mtx" in r.stdout
    p = write(tmp_path, "ns.mtx", "%%MM\n3 4 1\n1 2\n")
    r = run_cli("-i", p)
    assert r.returncode == 2 and "not square" in r.stderr


def test_chung_lu_batch_params_match_oracle(lib_built, oracle_mod):
    for i in range(32):
        assert tribit_amd.chung_lu_batch_params(3, i) == oracle_mod.chung_lu_batch_params(3, i)


def test_class_features_host(lib_built):
    """Cost-model features of (outdeg class, slot) from bucket statistics (host code, no device): the
    class's size, outdeg min / median / max by outdeg bin, work share and the one-hot class, kind and
    heuristic-slot markers; a slot the class does not fit is refused."""
    p = tribit_amd.Profile()
    # class 0: 10 lists of outdeg <= 1 (bin 0) and 30 of outdeg 25..32 (bin 9); class 2: 5 lists in bin 20
    for b, lists, w, k in ((0 * 32 + 3, 10, 100, 10), (9 * 32 + 5, 30, 900, 300), (20 * 32 + 7, 5, 5000, 500)):
        p.lists[b], p.sum_w[b], p.sum_k[b] = lists, w, k
        p.sum_d[b] = lists * tribit_amd.bucket_max_outdeg(b)
    f = tribit_amd.class_features(p, 0, 1)
    assert f.shape == (tribit_amd.FEATURES,) and np.all(np.isfinite(f))
    assert f[0] == pytest.approx(np.log2(41) / 8)
    assert f[7] == pytest.approx(np.log2(2) / 4)  # smallest outdeg bin: top 1
    assert f[8] == pytest.approx(np.log2(33) / 4)  # median list in the bin of top 32
    assert f[9] == pytest.approx(np.log2(33) / 4)
    assert f[11] == pytest.approx((1000 + 310) / (1000 + 310 + 5500))
    assert list(f[12:15]) == [1, 0, 0] and list(f[15:20]) == [0, 1, 0, 0, 0] and f[21] == 0
    assert tribit_amd.class_features(p, 0, 0)[21] == 1  # the heuristic's slot of the class
    f2 = tribit_amd.class_features(p, 2, 5)
    assert list(f2[12:15]) == [0, 0, 1] and f2[18] == 1 and f2[21] == 1
    assert f2[7] == f2[8] == f2[9] == pytest.approx(np.log2(1537) / 4)
    empty = tribit_amd.class_features(p, 1, 3)
    assert empty[0] == 0 and empty[11] == 0
    assert sorted(tribit_amd.class_slots(1)) == [2, 3, 4, 5, 6, 7] and tribit_amd.class_slots(2) == [4, 5, 6, 7]
    for cls, slot in ((1, 0), (2, 3), (3, 4), (0, 8)):
        with pytest.raises(tribit_amd.TribitError):
            tribit_amd.class_features(p, cls, slot)
