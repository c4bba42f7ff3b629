"""One rank of the reference's multi-GPU decomposition (main.rs:97-220,
main5_multi.cu:589, main6_multi.cu:595): diagonal and off-diagonal partial
counts through the status-returning C ABI and the reference-named symbols.
Pinned by the decomposition's invariant (every triangle in exactly one task,
so the partial counts sum to the oracle's total) and, on a small graph, by a
brute-force listing of the triangles each task owns."""
import ctypes
from itertools import combinations

import numpy as np
import pytest

import tribit_amd
from tests import graphs

pytestmark = pytest.mark.gpu


def note(msg):
    print(f"[gpu-case] {msg}", flush=True)


def partition(keys, n, P):
    """Reference partitions (main.rs:97-99): block size 32 * ceil((n / P) / 32); by the larger id."""
    bs = 32 * -(-(-(-n // P)) // 32)
    u = np.unique(keys[keys != np.uint64(0xFFFFFFFFFFFFFFFF)])
    hi = (u & np.uint64(0xFFFFFFFF)).astype(np.int64)
    return bs, [u[(hi // bs) == p] for p in range(-(-n // bs))]


def split(a, k):
    return [a[i::k] for i in range(k)] if a.size else [a]


def tasks(ctx, keys, n, P, ref_names):
    bs, parts = partition(keys, n, P)
    lib = ctx.lib
    total, diag, off = 0, [], []
    for p, e in enumerate(parts):
        chunks = split(e, 3)
        if ref_names:
            c = tribit_amd.Chunks(chunks)
            t = lib.count_partial_triangles_diagonal(ctypes.addressof(c.ptrs), c.cum.ctypes.data, c.num, n)
            assert lib.tribit_partial_last_status() == 0
        else:
            t = tribit_amd.count_partial_diagonal(ctx, chunks, n)
        diag.append(t)
        later = [x for x in parts[p + 1:] if x.size]
        if later:
            if ref_names:
                c, x = tribit_amd.Chunks(split(e, 2)), tribit_amd.Chunks(later)
                o = lib.count_partial_triangles_off_diagonal(ctypes.addressof(c.ptrs), c.cum.ctypes.data, c.num,
                                                             ctypes.addressof(x.ptrs), x.cum.ctypes.data, x.num, n)
                assert lib.tribit_partial_last_status() == 0
            else:
                o = tribit_amd.count_partial_off_diagonal(ctx, split(e, 2), later, n)
            off.append(o)
        else:
            off.append(0)
    return bs, diag, off


@pytest.mark.parametrize("ref_names", [False, True])
def test_partial_counts_sum_to_total(gpu_ctx, oracle_mod, ref_names):
    note(f"partial counts rmat s13, P=4, reference names={ref_names}")
    if ref_names:
        gpu_ctx.lib.init_gpu_kernel5(0)
    keys = oracle_mod.gen_rmat(13, 12, seed=1)
    n = 1 << 13
    _, diag, off = tasks(gpu_ctx, keys, n, 4, ref_names)
    assert sum(diag) + sum(off) == oracle_mod.count_keys(keys, n)
    assert min(diag) >= 0 and sum(off) > 0


def test_partial_counts_brute_force(gpu_ctx, oracle_mod):
    note("partial counts vs brute-force ownership, n=300, P=3")
    n = 300
    keys = graphs.random_graph(n, 0.12, 21)
    bs, diag, off = tasks(gpu_ctx, keys, n, 3, False)
    u = np.unique(keys)
    A = np.zeros((n, n), bool)
    a = (u >> np.uint64(32)).astype(np.int64)
    b = (u & np.uint64(0xFFFFFFFF)).astype(np.int64)
    A[a, b] = A[b, a] = True
    want_d = [0] * len(diag)
    want_o = [0] * len(off)
    for x, y, z in combinations(range(n), 3):  # x < y < z
        if A[x, y] and A[y, z] and A[x, z]:
            if y // bs == z // bs:
                want_d[z // bs] += 1
            else:
                want_o[y // bs] += 1
    assert diag == want_d and off == want_o


def test_partial_bad_arguments(gpu_ctx):
    note("partial count argument errors")
    with pytest.raises(tribit_amd.TribitError):
        tribit_amd.count_partial_diagonal(gpu_ctx, [np.zeros(0, np.uint64)], 0)
    bad = [np.array([int(graphs.pack(0, 9))], np.uint64)]
    with pytest.raises(tribit_amd.TribitError) as ei:
        tribit_amd.count_partial_diagonal(gpu_ctx, bad, 5)
    assert ei.value.status == 4


def test_decomposition_rmat20_config(gpu_ctx):
    """The whole decomposition at a configuration's size: R-MAT s20 ef16 seed 7 (the mixed family's
    R-MAT member) in P = 16 partitions; every diagonal task and every off-diagonal group (nnz threshold
    2M keys, so columns split into several groups as main.rs:186-220 splits at 1e9) through the
    reference-named symbols. The partial counts sum to the oracle's whole-graph count committed in
    tests/golden/configs.json."""
    import json
    import os

    from tests.conftest import GOLDEN

    with open(os.path.join(GOLDEN, "configs.json")) as f:
        want = json.load(f)["c4_rmat20"]["triangles"]
    note("decomposition R-MAT s20 ef16 seed 7, P=16")
    lib = gpu_ctx.lib
    lib.init_gpu_kernel5(0)
    n = 1 << 20
    keys = gpu_ctx.gen_rmat_keys(20, 16, seed=7)
    keys = keys[keys != np.uint64(0xFFFFFFFFFFFFFFFF)]
    P, threshold = 16, 2_000_000
    bs = -(-(n // P) // 32) * 32
    pmax = ((keys & np.uint64(0xFFFFFFFF)) // np.uint64(bs)).astype(np.int64)
    pmin = ((keys >> np.uint64(32)) // np.uint64(bs)).astype(np.int64)
    nparts = -(-n // bs)
    total, groups = 0, 0
    for pc in range(nparts):
        cols = [keys[(pmax == pc) & (pmin == j)] for j in range(pc + 1)]
        c = tribit_amd.Chunks(cols)
        t = lib.count_partial_triangles_diagonal(ctypes.addressof(c.ptrs), c.cum.ctypes.data, c.num, n)
        assert lib.tribit_partial_last_status() == 0
        total += t
        below = pmin <= pc
        nnz = np.bincount(pmax[below], minlength=nparts)
        group, acc = [], 0
        for s in list(range(pc + 1, nparts)) + [None]:
            if s is not None:
                acc += int(nnz[s])
            if s is None or (acc >= threshold and group):
                if group:
                    idxs = [keys[(pmax == q) & below & (pmin == j)] for q in group for j in range(pc + 1)]
                    x = tribit_amd.Chunks(idxs)
                    o = lib.count_partial_triangles_off_diagonal(ctypes.addressof(c.ptrs), c.cum.ctypes.data, c.num,
                                                                 ctypes.addressof(x.ptrs), x.cum.ctypes.data, x.num,
                                                                 n)
                    assert lib.tribit_partial_last_status() == 0
                    info = tribit_amd.partial_last_info(gpu_ctx)
                    assert info["triangles"] == o and info["lists"] > 0
                    total += o
                    groups += 1
                if s is None:
                    break
                group, acc = [s], int(nnz[s])
            else:
                group.append(s)
    assert groups > nparts  # some columns split into several groups
    assert total == want, (total, want)
