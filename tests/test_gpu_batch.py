"""Batches (tribit_amd/csrc/batch.hip): per-graph exact counts against the
oracle, several parts, every schedule, the captured-graph replays and the
autotuner (which must keep every count)."""
import numpy as np
import pytest

import tribit_amd
from tests import graphs

pytestmark = pytest.mark.gpu

SCHEDS = ("plain", "directed", "wedge", "learned")


def note(msg):
    print(f"[gpu-case] {msg}", flush=True)


def test_batch_from_host_matches_oracle(gpu_ctx, oracle_mod):
    note("batch from host: 40 mixed graphs, 5+ parts")
    rng = np.random.default_rng(1)
    keys, ns, want = [], [], []
    for j in range(40):
        kind = j % 4
        if kind == 0:
            k, t = graphs.complete(int(rng.integers(3, 60)))
            n = int((k & np.uint64(0xFFFFFFFF)).max()) + 1
        elif kind == 1:
            n = int(rng.integers(2, 300))
            k = graphs.random_graph(n, float(rng.uniform(0.02, 0.6)), 100 + j)
        elif kind == 2:
            s = int(rng.integers(8, 12))
            n = 1 << s
            k = oracle_mod.gen_rmat(s, 8, seed=200 + j)
        else:
            n = int(rng.integers(1, 5))
            k = np.zeros(0, np.uint64)  # empty and edge-free graphs
        keys.append(k)
        ns.append(n)
        want.append(oracle_mod.count_keys(k, n) if k.size else 0)
    gpu_ctx.set_batch_part_keys(3000)  # force several parts
    try:
        with gpu_ctx.batch_from_keys(keys, ns) as b:
            inf = b.info()
            assert inf["graphs"] == 40 and inf["parts"] >= 5
            n, m = b.sizes()
            assert list(n) == ns
            assert [int(x) for x in m] == [oracle_mod.dedup(k).size for k in keys]
            for sched in SCHEDS:
                for fl in (0, tribit_amd.FORCE_MULTI, tribit_amd.FORCE_MULTI | tribit_amd.FORCE_LINEAR_PROBE):
                    c, ci = b.count(sched=sched, repeats=2, flags=fl)
                    assert [int(x) for x in c] == want, (sched, fl)
                    assert ci["repeats_agree"] and ci["triangles"] == sum(want)
    finally:
        gpu_ctx.set_batch_part_keys(1 << 30)


def test_batch_chung_lu_and_autotune(gpu_ctx, oracle_mod):
    note("Chung-Lu batch, 12 graphs, seed 3, autotune")
    nb = 12
    with gpu_ctx.batch_gen_chung_lu(3, 0, nb) as b:
        want = []
        for i in range(nb):
            gam, n, avg = oracle_mod.chung_lu_batch_params(3, i)
            k = oracle_mod.gen_chung_lu(n, gam, avg, (3 << 32) + i)
            want.append(oracle_mod.count_keys(k, n))
        c, ci = b.count(repeats=3)
        assert [int(x) for x in c] == want
        assert ci["launches"] >= 3 and ci["repeats_agree"]
        c, _ = b.count(sched="wedge", repeats=2, flags=tribit_amd.FORCE_MULTI)  # multi-task waves, per-graph counts
        assert [int(x) for x in c] == want
        best = b.autotune(repeats=2)
        assert best["triangles"] == sum(want)
        c2, ci2 = b.count(repeats=2)
        assert [int(x) for x in c2] == want
        assert ci2["group_width"] == best["group_width"] and ci2["waves"] == best["waves"]


def test_batch_rejects_bad_keys(gpu_ctx):
    note("batch key validation")
    bad = [np.array([int(graphs.pack(0, 5))], np.uint64)]  # 5 >= n = 3
    with pytest.raises(tribit_amd.TribitError) as ei:
        gpu_ctx.batch_from_keys(bad, [3])
    assert ei.value.status == 4
