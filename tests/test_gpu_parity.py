"""GPU parity: the HIP path (through the C ABI) against the pinned CPU oracle.

Integer work, so every check is bit-exact: triangle counts, and also the
count-ready structures themselves (degrees, oriented CSR) and W+.
Cases: the reference-held wiki-Vote fixture, closed forms, brute-forced tiny
graphs, the networkx golden set (GPU generators must reproduce the oracle's
key sets), every scheduler class (including rows split into id-range chunks
and rows split by wedge count), error paths, and -- at full sizes -- the
size-independent properties (repeat determinism, relabelling invariance,
W+ and m against the oracle).
"""
import json
import os
import re
import subprocess
import sys
from math import comb

import numpy as np
import pytest

import tribit_amd
from tests import gens, graphs
from tests.conftest import GOLDEN

pytestmark = pytest.mark.gpu
WIKI = os.path.join(GOLDEN, "wiki-Vote.mtx")


def note(msg):
    print(f"[gpu-case] {msg}", flush=True)
    sys.stdout.flush()


SCHEDS = ("plain", "directed", "learned")


def scheds_of(g):
    """Every schedule the graph supports (the wedge schedule needs relabelled ids)."""
    return SCHEDS + (("wedge",) if g.labels()[1] else ())


def count_all(g):
    """Counts with every scheduler; they must agree (each triangle is found exactly once)."""
    got = {s: g.count(sched=s).triangles for s in scheds_of(g)}
    assert len(set(got.values())) == 1, got
    return got["plain"]


def gpu_count(ctx, keys, n):
    with ctx.graph_from_keys(keys, n) as g:
        return count_all(g)


def assert_same_structure(g, keys, n, oracle_mod):
    """Oriented CSR, degrees and labels equal the oracle's (relabelled like the graph)."""
    labels, relabeled = g.labels()
    og = oracle_mod.Graph.from_keys(keys, n, relabel=relabeled)
    if relabeled:
        assert np.array_equal(labels, og.labels)
    else:
        assert np.array_equal(labels, np.arange(n, dtype=np.uint32))
    rp, col, deg = g.download()
    orp, ocol, odeg = og.arrays()
    assert np.array_equal(rp.astype(np.uint64), orp)
    assert np.array_equal(col, ocol)
    assert np.array_equal(deg, odeg)
    inf = g.info()
    assert inf["m"] == og.m and inf["wplus"] == og.wplus()


def test_wiki_vote(gpu_ctx, oracle_mod):
    note("wiki-Vote")
    keys, n, _ = tribit_amd.read_mtx(WIKI)
    with gpu_ctx.graph_from_keys(keys, n) as g:
        r = g.count(repeats=3)
        assert r.triangles == 608389 and r.repeats_agree
        assert_same_structure(g, keys, n, oracle_mod)


def test_relabel_option(gpu_ctx, oracle_mod):
    """Input ids kept (relabelling off), degree-rank ids and octave-class ids give the same counts; every
    CSR equals the oracle's under the same ids."""
    note("relabel off / degree rank / octave classes, rmat s14 ef8 seed 3")
    keys = oracle_mod.gen_rmat(14, 8, seed=3)
    n = 1 << 14
    want = oracle_mod.count_keys(keys, n)
    for on in (False, True, 2):
        gpu_ctx.set_relabel(on)
        try:
            with gpu_ctx.graph_from_keys(keys, n) as g:
                assert g.labels()[1] == on
                assert_same_structure(g, keys, n, oracle_mod)
                assert count_all(g) == want
                if not on:
                    with pytest.raises(tribit_amd.TribitError):
                        g.count(sched="wedge")
        finally:
            gpu_ctx.set_relabel(True)


def test_closed_forms(gpu_ctx):
    note("closed forms")
    for n in (3, 4, 5, 31, 64, 65, 100, 257):
        keys, t = graphs.complete(n)
        assert gpu_count(gpu_ctx, keys, n) == t == comb(n, 3), n
    for k in (4, 5, 100, 1000):
        keys, t = graphs.wheel(k)
        assert gpu_count(gpu_ctx, keys, k + 1) == t
    for k in (1, 7, 500):
        keys, t = graphs.windmill(k)
        assert gpu_count(gpu_ctx, keys, 2 * k + 1) == t
    for n in (3, 4, 1000):
        keys, t = graphs.cycle(n)
        assert gpu_count(gpu_ctx, keys, n) == t
    for parts in ((2, 3, 4), (30, 40, 50), (3, 3, 3, 3)):
        keys, t, n = graphs.complete_multipartite(parts)
        assert gpu_count(gpu_ctx, keys, n) == t
    keys, _, n = graphs.random_bipartite(300, 200, 0.2, 1)
    assert gpu_count(gpu_ctx, keys, n) == 0


def test_bruteforce_tiny(gpu_ctx, oracle_mod):
    note("brute force tiny")
    for seed in range(30):
        rng = np.random.default_rng(100 + seed)
        n = int(rng.integers(2, 80))
        keys = graphs.random_graph(n, float(rng.uniform(0.05, 0.95)), seed)
        assert gpu_count(gpu_ctx, keys, n) == oracle_mod.count_bruteforce(n, keys), seed


def test_golden_set(gpu_ctx, oracle_mod):
    with open(os.path.join(GOLDEN, "golden.json")) as f:
        gold = json.load(f)
    for r in gold:
        note(r["name"])
        g = gens.gpu_graph(gpu_ctx, r)
        okeys = gens.oracle_keys(oracle_mod, r)
        with g:
            assert count_all(g) == r["triangles"], r["name"]
            assert_same_structure(g, okeys, r["n"], oracle_mod)


def test_block_classes_and_chunks(gpu_ctx, oracle_mod):
    """Rows with outdeg > 256 (block tasks) and > 2048 (id-range chunks)."""
    note("K_2100 + K_600 (block tasks, 2 chunks)")
    k1, t1 = graphs.complete(2100)
    k2, t2 = graphs.complete(600, offset=2100)
    keys = np.concatenate([k1, k2])
    with gpu_ctx.graph_from_keys(keys, 2700) as g:
        for sched in scheds_of(g):
            r = g.count(sched=sched)
            assert r.triangles == t1 + t2
            if sched == "learned":
                # the two cliques are dense diagonal blocks: the judge may count them on the matrix
                # cores (candidate 4); otherwise menu slots 5..7 (the block kernels) run
                assert (r.candidate == 4 and r.dense_blocks == 2) or sum(r.tasks_by_variant[5:8]) > 0, r
            else:
                assert r.tasks_by_variant[3] > 0 and r.tasks_by_variant[2] > 0
    note("K_4500 (3 chunks)")
    keys, t = graphs.complete(4500)
    assert gpu_count(gpu_ctx, keys, 4500) == t
    note("dense random n=5000 p=0.5")
    keys = graphs.random_graph(5000, 0.5, 3)
    with gpu_ctx.graph_from_keys(keys, 5000) as g:
        assert count_all(g) == oracle_mod.count_keys(keys, 5000)


def test_cursor_tasks(gpu_ctx, oracle_mod):
    """Owners of more than 2048 ids outside the wedge layout count each 2048-id chunk of their row in
    turn, every streamed list keeping a cursor between chunks: K_2600 (two chunks per owner) with input
    ids spread so that the chunks are LDS bitmaps (one 2^18-id window) or hashed sets (wider), under
    the plain and directed schedules, every set kind and several lane-group widths; a 2,400-vertex
    block of density 0.95 on random ids among sparse noise against the oracle."""
    note("cursor tasks: K_2600 with spread ids, hubs")
    k = 2600
    a, b = np.triu_indices(k, 1)
    want = comb(k, 3)
    gpu_ctx.set_relabel(0)
    try:
        for spread in (1, 97, 1000):  # bitmap chunks; bitmap and hashed; hashed
            ids = np.arange(k, dtype=np.uint64) * np.uint64(spread) + np.uint64(1)
            keys = graphs.pack(ids[a], ids[b])
            n = int(ids[-1]) + 3
            with gpu_ctx.graph_from_keys(keys, n) as g:
                for sched in ("plain", "directed"):
                    for fl in (0, tribit_amd.NO_BITMAP, tribit_amd.FORCE_LINEAR_PROBE):
                        for gw in (None, (1, 2, 4, 4), (2, 8, 64, 32)):
                            r = g.count(sched=sched, flags=fl, group_width=gw)
                            assert r.triangles == want, (spread, sched, fl, gw)
        rng = np.random.default_rng(21)
        n = 60000
        blk = np.sort(rng.choice(n, 2400, replace=False)).astype(np.uint64)  # outdeg up to ~2,280
        i, j = np.triu_indices(blk.size, 1)
        sel = rng.random(i.size) < 0.95
        u = rng.integers(0, n, 200000, dtype=np.uint64)
        v = rng.integers(0, n, 200000, dtype=np.uint64)
        keys = np.concatenate([graphs.pack(blk[i[sel]], blk[j[sel]]),
                               graphs.pack(np.minimum(u, v)[u != v], np.maximum(u, v)[u != v])])
        want = oracle_mod.count_keys(keys, n)
        with gpu_ctx.graph_from_keys(keys, n) as g:
            for sched in ("plain", "directed"):
                for fl in (0, tribit_amd.NO_BITMAP):
                    assert g.count(sched=sched, flags=fl).triangles == want, (sched, fl)
    finally:
        gpu_ctx.set_relabel(1)


def test_bitmap_and_hashed_sets(gpu_ctx, oracle_mod):
    """Block tasks whose owner ids fit a 2^18-id window use LDS bitmaps, wider ones cuckoo sets:
    both, and the hashed sets forced everywhere, give the same counts."""
    note("bitmap vs hashed owner sets")
    wide = graphs.complete(300)[0]  # ids 0, 2000, ... (window 598,000 ids)
    wa = (wide >> np.uint64(32)) * np.uint64(2000)
    wb = (wide & np.uint64(0xFFFFFFFF)) * np.uint64(2000)
    narrow_a = (wide >> np.uint64(32)) * np.uint64(500) + np.uint64(1)  # ids 1, 501, ... (149,501 ids)
    narrow_b = (wide & np.uint64(0xFFFFFFFF)) * np.uint64(500) + np.uint64(1)
    keys = np.concatenate([graphs.pack(wa, wb), graphs.pack(narrow_a, narrow_b)])
    n = 600000
    want = 2 * comb(300, 3)
    for on in (False, True):
        gpu_ctx.set_relabel(on)
        try:
            with gpu_ctx.graph_from_keys(keys, n) as g:
                for sched in scheds_of(g):
                    for fl in (0, tribit_amd.NO_BITMAP, tribit_amd.FORCE_LINEAR_PROBE, tribit_amd.FORCE_TILES):
                        r = g.count(sched=sched, flags=fl)
                        assert r.triangles == want, (on, sched, fl)
                        if sched == "learned":  # the two cliques may go to the matrix cores as dense blocks
                            assert (r.candidate == 4 and r.dense_blocks > 0) or sum(r.tasks_by_variant[5:8]) > 0
                        else:
                            assert r.tasks_by_variant[2] > 0
        finally:
            gpu_ctx.set_relabel(True)
    with gpu_ctx.gen_rmat(20, 16, seed=6) as g:
        a = g.count(sched="plain").triangles  # the plain layout has no narrow ranges
        for sched in scheds_of(g):
            assert g.count(sched=sched, flags=tribit_amd.NO_BITMAP).triangles == a
            # narrow wedge ranges as full ids instead of 16-bit halves
            assert g.count(sched=sched, flags=tribit_amd.WIDE_IDS).triangles == a
            assert g.count(sched=sched, flags=tribit_amd.WIDE_IDS | tribit_amd.NO_BITMAP).triangles == a


def test_narrow_ranges(gpu_ctx):
    """Wedge ranges streamed as 16-bit halves (two ids per load): ranges inside one 64K-id window, ranges
    crossing windows, odd starts and ends, owners cut into 2048-id chunks, hashed and bitmap sets."""
    note("narrow wedge ranges")
    k = 2500  # owners of outdeg > 2048: chunked sets, cut ranges
    a, b = np.triu_indices(k, 1)
    for spread, shift in ((3, 1), (97, 5), (26, 60000)):  # one window; crossing windows; straddling 2^16
        ids = np.arange(k, dtype=np.uint64) * np.uint64(spread) + np.uint64(shift)
        keys = graphs.pack(ids[a], ids[b])
        n = int(ids[-1]) + 7
        want = comb(k, 3)
        with gpu_ctx.graph_from_keys(keys, n) as g:
            for sched in ("wedge", "learned"):
                for fl in (0, tribit_amd.NO_BITMAP, tribit_amd.FORCE_LINEAR_PROBE, tribit_amd.WIDE_IDS):
                    for gw in (None, (1, 4, 4, 4), (2, 8, 64, 32)):
                        r = g.count(sched=sched, flags=fl, group_width=gw if sched == "wedge" else None)
                        assert r.triangles == want, (spread, sched, fl, gw)
    rng = np.random.default_rng(11)
    for n, m in ((70000, 900000), (300000, 2000000)):
        u = rng.integers(0, n, m, dtype=np.uint64)
        v = (u + rng.integers(1, 3000, m, dtype=np.uint64)) % np.uint64(n)
        keys = graphs.pack(np.minimum(u, v), np.maximum(u, v))
        keys = keys[(keys >> np.uint64(32)) != (keys & np.uint64(0xFFFFFFFF))]
        with gpu_ctx.graph_from_keys(keys, n) as g:
            want = g.count(sched="plain").triangles
            for fl in (0, tribit_amd.NO_BITMAP, tribit_amd.WIDE_IDS):
                for gw in (None, (1, 2, 4, 4), (1, 4, 32, 64)):
                    assert g.count(sched="wedge", flags=fl, group_width=gw).triangles == want, (n, fl, gw)


def test_forced_multi_task_waves(gpu_ctx, oracle_mod):
    """Every launch of the 64-slot wave kernel as multi-task waves (the kernel large launches switch to),
    under every schedule, with cuckoo and with fallback sets."""
    note("forced multi-task waves")
    for spec in [(13, 12, 1), (16, 16, 4)]:
        keys = oracle_mod.gen_rmat(spec[0], spec[1], seed=spec[2])
        want = oracle_mod.count_keys(keys, 1 << spec[0])
        with gpu_ctx.gen_rmat(spec[0], spec[1], seed=spec[2]) as g:
            for sched in scheds_of(g):
                for fl in (tribit_amd.FORCE_MULTI, tribit_amd.FORCE_MULTI | tribit_amd.FORCE_LINEAR_PROBE):
                    assert g.count(sched=sched, flags=fl).triangles == want, (spec, sched, fl)


def test_fallback_linear_probe(gpu_ctx, oracle_mod):
    """The fallback sets (used when every cuckoo build of a task fails) count the same."""
    note("forced linear-probe fallback")
    for spec in [(13, 12, 1), (16, 16, 4)]:
        with gpu_ctx.gen_rmat(spec[0], spec[1], seed=spec[2]) as g:
            a = g.count().triangles
            for sched in scheds_of(g):
                assert g.count(sched=sched, flags=tribit_amd.FORCE_LINEAR_PROBE).triangles == a
    k1, t1 = graphs.complete(2100)
    with gpu_ctx.graph_from_keys(k1, 2100) as g:
        for sched in scheds_of(g):
            assert g.count(sched=sched, flags=tribit_amd.FORCE_LINEAR_PROBE).triangles == t1
    keys = graphs.random_graph(3000, 0.3, 5)
    with gpu_ctx.graph_from_keys(keys, 3000) as g:
        want = oracle_mod.count_keys(keys, 3000)
        for sched in scheds_of(g):
            assert g.count(sched=sched, flags=tribit_amd.FORCE_LINEAR_PROBE).triangles == want


def test_global_load_path(gpu_ctx, oracle_mod):
    """The streaming path for graphs whose col array exceeds a 32-bit buffer range counts the same."""
    note("forced global-load streaming")
    with gpu_ctx.gen_rmat(16, 16, seed=4) as g:
        a = g.count().triangles
        for sched in scheds_of(g):
            assert g.count(sched=sched, flags=tribit_amd.FORCE_GLOBAL_LOADS).triangles == a
            assert g.count(sched=sched, flags=tribit_amd.FORCE_GLOBAL_LOADS | tribit_amd.FORCE_LINEAR_PROBE
                           ).triangles == a
    keys, t = graphs.complete(2300)
    with gpu_ctx.graph_from_keys(keys, 2300) as g:
        for sched in scheds_of(g):
            assert g.count(sched=sched, flags=tribit_amd.FORCE_GLOBAL_LOADS).triangles == t


def test_rmat16_repeats(gpu_ctx, oracle_mod):
    note("rmat s16 ef16 seed 4")
    with gpu_ctx.gen_rmat(16, 16, seed=4) as g:
        want = oracle_mod.Graph.from_keys(oracle_mod.gen_rmat(16, 16, seed=4), 1 << 16).count()
        for sched in scheds_of(g):
            r = g.count(sched=sched, repeats=2)
            assert r.triangles == want and r.repeats_agree


def test_edge_cases(gpu_ctx):
    note("edge cases")
    assert gpu_count(gpu_ctx, np.zeros(0, np.uint64), 10) == 0
    assert gpu_count(gpu_ctx, np.array([int(graphs.pack(0, 1))], np.uint64), 2) == 0
    assert gpu_count(gpu_ctx, np.full(7, 0xFFFFFFFFFFFFFFFF, np.uint64), 5) == 0
    tri = np.array([int(graphs.pack(0, 1)), int(graphs.pack(1, 2)), int(graphs.pack(0, 2))], np.uint64)
    assert gpu_count(gpu_ctx, np.concatenate([tri, tri, np.full(3, 0xFFFFFFFFFFFFFFFF, np.uint64)]), 3) == 1
    with pytest.raises(tribit_amd.TribitError) as ei:
        gpu_ctx.graph_from_keys(np.array([(2 << 32) | 1], np.uint64), 3)  # min > max
    assert ei.value.status == 4
    with pytest.raises(tribit_amd.TribitError) as ei:
        gpu_ctx.graph_from_keys(np.array([(0 << 32) | 3], np.uint64), 3)  # max >= n
    assert ei.value.status == 4


def test_alloc_cap(gpu_ctx):
    note("allocation cap")
    keys, _ = graphs.complete(200)
    gpu_ctx.set_alloc_cap(4096)
    try:
        with pytest.raises(tribit_amd.TribitError) as ei:
            gpu_ctx.graph_from_keys(keys, 200)
        assert ei.value.status == 5
    finally:
        gpu_ctx.set_alloc_cap(0)
    assert gpu_count(gpu_ctx, keys, 200) == comb(200, 3)


def test_rmat20_exact(gpu_ctx, oracle_mod):
    note("rmat s20 ef16 seed 5 vs oracle")
    with gpu_ctx.gen_rmat(20, 16, seed=5) as g:
        r = g.count()
        tri = r.triangles
        assert r.tasks > g.info()["rows_active"]  # rows with large W+ were split into several tasks
        assert count_all(g) == tri
        og = oracle_mod.Graph.from_keys(oracle_mod.gen_rmat(20, 16, seed=5), 1 << 20)
        inf = g.info()
        assert (inf["m"], inf["wplus"]) == (og.m, og.wplus())
        assert tri == og.count()


@pytest.mark.parametrize("spec", [
    {"gen": "chung_lu", "n": 262144, "gamma": 2.1, "avg_deg": 64.0, "seed": 13},
    {"gen": "rgg", "n": 200000, "avg_deg": 24.0, "seed": 21},
    {"gen": "sbm", "blocks": 100, "size": 512, "p_in": 0.3, "p_out": 1e-4, "seed": 31},
    {"gen": "er", "n": 1 << 17, "samples": 1 << 21, "seed": 41},
], ids=lambda s: s["gen"])
def test_mixed_family_medium(gpu_ctx, oracle_mod, spec):
    """Scaled-down members of the mixed family (configuration 4) against the oracle: CSR and count."""
    n = spec["blocks"] * spec["size"] if spec["gen"] == "sbm" else spec["n"]
    r = {"spec": spec, "n": n}
    note(f"mixed family {spec}")
    with gens.gpu_graph(gpu_ctx, r) as g:
        okeys = gens.oracle_keys(oracle_mod, r)
        assert_same_structure(g, okeys, n, oracle_mod)
        assert count_all(g) == oracle_mod.count_keys(okeys, n)


def test_rmat22_properties(gpu_ctx, oracle_mod):
    """Configuration 2 (R-MAT s22 ef16 seed 2): determinism, relabelling invariance, m and W+ vs oracle."""
    note("rmat s22 ef16 seed 2 properties")
    with gpu_ctx.gen_rmat(22, 16, seed=2) as g:
        r = g.count(repeats=3)
        assert r.repeats_agree
        assert g.count(sched="directed", repeats=2).triangles == r.triangles
        inf = g.info()
    okeys = oracle_mod.gen_rmat(22, 16, seed=2)
    og = oracle_mod.Graph.from_keys(okeys, 1 << 22)
    assert (inf["m"], inf["wplus"]) == (og.m, og.wplus())
    del og
    perm = graphs.permute(okeys, 1 << 22, 7)
    del okeys
    with gpu_ctx.graph_from_keys(perm, 1 << 22) as g2:
        assert g2.count(sched="directed").triangles == r.triangles


def test_cli_wiki_vote_stdout_contract(tmp_path):
    """`tribit -i wiki-Vote.mtx` output parsed by the reference harness's regexes (run_single_gpu.py:78-93).
    The learned schedule judges its candidates in the warm-up pass only: the measured pass runs the
    judged choice once, so its Kernel (ms) is the whole of its counting (scheduler + every launch)."""
    note("cli wiki-Vote")
    rec = tmp_path / "rec.json"
    out = subprocess.run([tribit_amd.CLI_PATH, "-i", WIKI, "--json", str(rec)], capture_output=True, text=True,
                         timeout=120)
    assert out.returncode == 0, out.stderr
    fields = {
        "n_blocks": r"Blocks of threads\s*:\s*(\d+)",
        "preprocessing_time": r"Preprocessing \(ms\)\s*:\s*([\d.]+)",
        "kernel_time": r"Kernel \(ms\)\s*:\s*([\d.]+)",
        "triangles": r"Triangles\s*:\s*(\d+)",
    }
    vals = {k: re.search(p, out.stdout).group(1) for k, p in fields.items()}
    assert vals["triangles"] == "608389"
    assert out.stdout.count("Triangle Counting Summary") == 1
    assert "Reading data from " + WIKI in out.stdout
    r = json.loads(rec.read_text())
    assert r["sched"] == "learned" and r["triangles"] == 608389
    assert r["warmup_ms_judge"] > 0            # the warm-up pass judged the candidates
    assert r["ms_judge"] == 0                  # the measured pass did not
    assert r["candidate"] == r["warmup_candidate"]
    assert abs(float(vals["kernel_time"]) - r["ms_count"]) < 1e-3


def test_harness_csv_wiki_vote(tmp_path):
    """The CSV harness over a folder holding wiki-Vote.mtx, with the real binary: the reference
    harness's columns, the exact count and the pool high-water mark in bytes."""
    from tribit_amd import harness
    note("harness csv wiki-Vote")
    data = tmp_path / "graphs"
    data.mkdir()
    os.symlink(WIKI, data / "wiki-Vote.mtx")
    out = tmp_path / "results.csv"
    assert harness.main(["--exe_path", tribit_amd.CLI_PATH, "--data_path", str(data), "--out_path", str(out),
                         "--get_memory_consumption"]) == 0
    lines = open(out).read().splitlines()
    assert lines[0] == harness.HEADER and len(lines) == 2
    row = lines[1].split(",")
    assert row[:5] == ["tribit", "wiki-Vote.mtx", "8297", "8297", "103689"]
    assert row[8] == "608389" and int(row[5]) > 0 and float(row[7]) > 0
    assert int(row[9]) >= 8 * 103689  # the pool held at least the edge keys


def test_webgraph_input(tmp_path, gpu_ctx, oracle_mod):
    """A BVGraph file (written by the test encoder, both directions of every edge, as the
    reference's symmetric WebGraph inputs) read by tribit_read_webgraph and by `tribit --webgraph`:
    the count equals the oracle's on the same edges."""
    from tests import webgraph_enc
    note("webgraph input rmat s12")
    keys = oracle_mod.gen_rmat(12, 8, seed=9)
    n = 1 << 12
    g0 = oracle_mod.Graph.from_keys(keys, n)
    want = g0.count()
    u = (keys >> np.uint64(32)).astype(np.int64)
    v = (keys & np.uint64(0xFFFFFFFF)).astype(np.int64)
    succ = [set() for _ in range(n)]
    for a, b in zip(u.tolist(), v.tolist()):
        if a != b:
            succ[a].add(b)
            succ[b].add(a)
    base = str(tmp_path / "rmat12")
    webgraph_enc.write(base, [sorted(s) for s in succ])
    wkeys, wn = tribit_amd.read_webgraph(base)
    assert wn == n
    with gpu_ctx.graph_from_keys(wkeys, wn) as g:
        assert g.count(sched="learned").triangles == want
    out = subprocess.run([tribit_amd.CLI_PATH, "--webgraph", base], capture_output=True, text=True, timeout=120)
    assert out.returncode == 0, out.stderr
    assert re.search(r"Triangles\s*:\s*(\d+)", out.stdout).group(1) == str(want)


def test_cli_generated_config1():
    note("cli rmat s13 ef12 seed 1")
    out = subprocess.run([tribit_amd.CLI_PATH, "--rmat", "13,12,0.57,0.19,0.19,1"], capture_output=True, text=True,
                         timeout=120)
    assert out.returncode == 0, out.stderr
    assert re.search(r"Triangles\s*:\s*(\d+)", out.stdout).group(1) == "689356"


def feasible_slots(bucket):
    top = tribit_amd.bucket_max_outdeg(bucket)
    return [s for s, (_, _, dmax) in enumerate(tribit_amd.menu()) if top <= dmax]


def test_plans_profiles_and_learned_tables(gpu_ctx, oracle_mod):
    """Every feasible menu plan (round-robin over the slots), the heuristic, an
    untrained cost model's plan and the timing-log run give the exact count."""
    note("menu plans / profiles / learned tables, rmat s16 and sbm")
    cases = [(gpu_ctx.gen_rmat(16, 16, seed=4), oracle_mod.count_keys(oracle_mod.gen_rmat(16, 16, seed=4), 1 << 16)),
             (None, None)]
    sbm = {"spec": {"gen": "sbm", "blocks": 20, "size": 300, "p_in": 0.4, "p_out": 1e-3, "seed": 8}, "n": 6000}
    cases[1] = (gens.gpu_graph(gpu_ctx, sbm), oracle_mod.count_keys(gens.oracle_keys(oracle_mod, sbm), 6000))
    for g, want in cases:
        with g:
            assert g.count(sched="wedge").triangles == want
            p = g.profile("wedge")
            assert p.triangles == want and p.tick_ns > 0
            used = [b for b in range(tribit_amd.BUCKETS) if p.lists[b]]
            assert used and sum(p.ticks[p.slot[b]][b] for b in used) > 0
            for r in range(3):
                table = np.array([feasible_slots(b)[(b + r) % len(feasible_slots(b))]
                                  for b in range(tribit_amd.BUCKETS)], np.uint8)
                for layout in ("wedge", "directed", "plain"):
                    assert g.profile(layout, table).triangles == want, (r, layout)
            heur = g.learned_plan()
            assert g.count(sched="learned").triangles == want
            ms, tri = g.count_plan("wedge", None, 2)
            assert tri == want and ms > 0
            for c in range(tribit_amd.CLASSES):
                for s in tribit_amd.class_slots(c):
                    a = tribit_amd.heuristic_slots("wedge")
                    a[c] = s
                    assert g.count_plan("wedge", tribit_amd.class_table(a), 1)[1] == want, (c, s)
                    assert np.all(np.isfinite(tribit_amd.class_features(p, c, s)))
            with gpu_ctx.mlp(seed=3) as m:
                gpu_ctx.set_cost_model(m)
                try:
                    t = g.learned_plan()
                    assert all(t[b] in feasible_slots(b) for b in range(tribit_amd.BUCKETS))
                    assert g.count(sched="learned").triangles == want
                finally:
                    gpu_ctx.set_cost_model(None)
            assert np.array_equal(g.learned_plan(), heur)


@pytest.mark.parametrize("spec", [
    {"gen": "sbm", "blocks": 40, "size": 512, "p_in": 0.5, "p_out": 1e-3, "seed": 5},
    {"gen": "rmat", "scale": 18, "ef": 16, "a": 0.57, "b": 0.19, "c": 0.19, "seed": 3},
    {"gen": "chung_lu", "n": 100000, "gamma": 2.1, "avg_deg": 48.0, "seed": 6},
], ids=lambda s: s["gen"])
def test_bit_tiles(gpu_ctx, oracle_mod, spec):
    """64-column bit tiles (AND + popcount against the LDS bitmap) count the same under every schedule,
    relabelled or not, and through the menu's tile slot."""
    n = spec["blocks"] * spec["size"] if spec["gen"] == "sbm" else (1 << spec["scale"] if spec["gen"] == "rmat"
                                                                    else spec["n"])
    r = {"spec": spec, "n": n}
    note(f"bit tiles {spec}")
    want = oracle_mod.count_keys(gens.oracle_keys(oracle_mod, r), n)
    for on in (True, False):
        gpu_ctx.set_relabel(on)
        try:
            with gens.gpu_graph(gpu_ctx, r) as g:
                for sched in scheds_of(g):
                    assert g.count(sched=sched, flags=tribit_amd.FORCE_TILES).triangles == want, (on, sched)
                if on:
                    tile_table = np.array([7 if tribit_amd.bucket_max_outdeg(b) > 256 else 0 if
                                           tribit_amd.bucket_max_outdeg(b) <= 32 else 3
                                           for b in range(tribit_amd.BUCKETS)], np.uint8)
                    assert tribit_amd.menu()[7][0] == 3
                    assert g.profile("wedge", tile_table).triangles == want
                    assert g.profile("directed", tile_table).triangles == want
        finally:
            gpu_ctx.set_relabel(True)
    keys, t = graphs.complete(2300)
    with gpu_ctx.graph_from_keys(keys, 2300) as g:
        for sched in scheds_of(g):
            assert g.count(sched=sched, flags=tribit_amd.FORCE_TILES).triangles == t


def assert_tiles_decode(g, keys, n, oracle_mod):
    """The GPU's 32- and 64-column tiles equal the oracle's (or_tiles) on the same oriented CSR."""
    labels, relabeled = g.labels()
    og = oracle_mod.Graph.from_keys(keys, n, relabel=relabeled)
    for width in (32, 64):
        tp, blk, mask = g.tiles(width)
        otp, oblk, omask = og.tiles(width)
        assert np.array_equal(tp.astype(np.uint64), otp), width
        assert np.array_equal(blk, oblk), width
        assert np.array_equal(mask.astype(np.uint64), omask), width
    inf = g.info()
    assert (inf["tiles32"], inf["tiles64"]) == (og.tiles(32)[1].size, og.tiles(64)[1].size)


def test_bit_tiles_32(gpu_ctx, oracle_mod):
    """32-column uint32 bit tiles: decode equal to the oracle's, and exact counts when every block
    launch streams them (configuration 1's plain density-threshold path: R-MAT s13 ef12 seed 1), on
    every schedule, relabelled or not, and on the golden set."""
    note("32-column tiles: rmat s13 ef12 seed 1")
    keys = oracle_mod.gen_rmat(13, 12, seed=1)
    want = oracle_mod.count_keys(keys, 1 << 13)
    for on in (True, False):
        gpu_ctx.set_relabel(on)
        try:
            with gpu_ctx.graph_from_keys(keys, 1 << 13) as g:
                assert_tiles_decode(g, keys, 1 << 13, oracle_mod)
                for sched in scheds_of(g):
                    for fl in (tribit_amd.FORCE_TILES | tribit_amd.TILES32, tribit_amd.FORCE_TILES):
                        assert g.count(sched=sched, flags=fl).triangles == want, (on, sched, fl)
        finally:
            gpu_ctx.set_relabel(True)
    with open(os.path.join(GOLDEN, "golden.json")) as f:
        gold = json.load(f)
    for r in gold:
        note(f"32-column tiles: {r['name']}")
        okeys = gens.oracle_keys(oracle_mod, r)
        with gens.gpu_graph(gpu_ctx, r) as g:
            assert_tiles_decode(g, okeys, r["n"], oracle_mod)
            for sched in scheds_of(g):
                assert g.count(sched=sched, flags=tribit_amd.FORCE_TILES | tribit_amd.TILES32).triangles == \
                    r["triangles"], (r["name"], sched)
    keys, t = graphs.complete(2300)
    with gpu_ctx.graph_from_keys(keys, 2300) as g:
        for sched in scheds_of(g):
            assert g.count(sched=sched, flags=tribit_amd.FORCE_TILES | tribit_amd.TILES32).triangles == t


def test_dense_core_mfma(gpu_ctx, oracle_mod):
    """Triangles of the top-id core counted by the int8 MFMA kernel plus the rest streamed: exact for
    every core size (partial tiles, the whole graph) and every schedule."""
    note("dense core: K_300, rmat s16, sbm, rmat s20 auto")
    keys, t = graphs.complete(300)
    with gpu_ctx.graph_from_keys(keys, 300) as g:
        for core in (1, 64, 128, 200, 256, 300):
            for sched in scheds_of(g):
                r = g.count(sched=sched, dense_core=core)
                assert r.triangles == t, (core, sched)
                assert r.dense_core == core
    keys = oracle_mod.gen_rmat(16, 16, seed=4)
    want = oracle_mod.count_keys(keys, 1 << 16)
    with gpu_ctx.graph_from_keys(keys, 1 << 16) as g:
        for core in (128, 1000, 4096, 1 << 16):
            for sched in scheds_of(g):
                assert g.count(sched=sched, dense_core=core).triangles == want, (core, sched)
        assert g.count(sched="wedge", dense_core=4096, flags=tribit_amd.FORCE_TILES).triangles == want
    sbm = {"spec": {"gen": "sbm", "blocks": 16, "size": 256, "p_in": 0.5, "p_out": 2e-3, "seed": 9}, "n": 4096}
    want = oracle_mod.count_keys(gens.oracle_keys(oracle_mod, sbm), 4096)
    with gens.gpu_graph(gpu_ctx, sbm) as g:
        for core in (512, 2048, 4096):
            assert g.count(sched="learned", dense_core=core).triangles == want
    with gpu_ctx.gen_rmat(20, 16, seed=6) as g:
        a = g.count(sched="wedge", dense_core=-1).triangles
        r = g.count(sched="learned")
        assert r.triangles == a
        assert g.count(sched="learned", dense_core=8192).triangles == a


def test_dense_diagonal_blocks(gpu_ctx, oracle_mod):
    """Dense diagonal blocks on the matrix cores (octave-class relabelling keeps each block of a block
    model a run of ids): exact against the oracle on unions of cliques with sparse noise, on a stochastic
    block model at configuration 4's size (count from tests/golden/configs.json), and on graphs with no
    dense block (the flag then changes nothing); every schedule and the learned judge agree."""
    with open(os.path.join(GOLDEN, "configs.json")) as f:
        sbm = json.load(f)["c4_sbm"]
    rng = np.random.default_rng(5)
    parts, t, off = [], 0, 0
    for size in (70, 130, 256, 300, 90):  # cliques of various sizes, consecutive ids
        k, tk = graphs.complete(size, offset=off)
        parts.append(k)
        t += tk
        off += size
    n = off
    noise = graphs.pack(rng.integers(0, n, 400), rng.integers(0, n, 400))
    keys = np.concatenate(parts + [noise[(noise >> np.uint64(32)) != (noise & np.uint64(0xFFFFFFFF))]])
    want = oracle_mod.count_keys(keys, n)
    gpu_ctx.set_relabel(2)
    try:
        note("dense blocks: cliques + noise")
        with gpu_ctx.graph_from_keys(keys, n) as g:
            r = g.count(sched="wedge", flags=tribit_amd.DENSE_BLOCKS)
            assert r.triangles == want and r.dense_blocks >= 3, (r.triangles, want, r.dense_blocks)
            for sched in scheds_of(g):
                assert g.count(sched=sched).triangles == want
            assert g.count(sched="learned").triangles == want
        note("dense blocks: SBM 1000 x 512 (configuration 4)")
        spec = sbm["spec"]
        with gpu_ctx.gen_sbm(spec["blocks"], spec["size"], spec["p_in"], spec["p_out"], spec["seed"]) as g:
            r = g.count(sched="wedge", flags=tribit_amd.DENSE_BLOCKS, repeats=3)
            assert r.triangles == sbm["triangles"] and r.repeats_agree
            assert r.dense_blocks >= 900, r.dense_blocks
            base = g.count(sched="wedge", repeats=3)
            assert base.triangles == sbm["triangles"]
            learned = g.count(sched="learned")
            assert learned.triangles == sbm["triangles"] and learned.ms_candidates[4] > 0
            note(f"sbm: wedge {base.ms_count:.3f} ms, with {r.dense_blocks} dense blocks {r.ms_count:.3f} ms "
                 f"(matrix cores {r.ms_dense:.3f} ms); learned candidate {learned.candidate}")
        note("dense blocks: R-MAT s16 (none expected)")
        rk = oracle_mod.gen_rmat(16, 16, seed=4)
        want_r = oracle_mod.count_keys(rk, 1 << 16)
        with gpu_ctx.graph_from_keys(rk, 1 << 16) as g:
            assert g.count(sched="wedge", flags=tribit_amd.DENSE_BLOCKS).triangles == want_r
            assert g.count(sched="learned").triangles == want_r
    finally:
        gpu_ctx.set_relabel(1)
    for seed in range(6):  # brute force on small graphs with planted blocks
        m = np.random.default_rng(40 + seed)
        nn = int(m.integers(150, 400))
        ks = [graphs.random_graph(nn, 0.02, seed)]
        at = 0
        while at < nn - 64:
            L = int(m.integers(64, 140))
            L = min(L, nn - at)
            k = graphs.random_graph(L, float(m.uniform(0.3, 0.9)), 100 + seed * 10 + at)
            ks.append(k + np.uint64(at) * np.uint64((1 << 32) + 1))
            at += L
        kk = np.concatenate(ks)
        gpu_ctx.set_relabel(2)
        try:
            with gpu_ctx.graph_from_keys(kk, nn) as g:
                r = g.count(sched="wedge", flags=tribit_amd.DENSE_BLOCKS)
                assert r.triangles == oracle_mod.count_bruteforce(nn, kk), seed
        finally:
            gpu_ctx.set_relabel(1)


def auto_mode(oracle_mod, keys, n):
    """The automatic id order's rule, restated on the host: octave classes (2) when the edges joining
    ids less than 512 apart fill at least 1/16 of such id pairs and n >= 4096, else degree rank (1)."""
    if n < 4096:
        return 1
    u = oracle_mod.dedup(keys)
    gap = (u & np.uint64(0xFFFFFFFF)).astype(np.int64) - (u >> np.uint64(32)).astype(np.int64)
    w = min(512, n)
    pairs = (w - 1) * n - w * (w - 1) / 2
    return 2 if np.count_nonzero(gap < 512) * 16 >= pairs else 1


def test_auto_ids(gpu_ctx, oracle_mod):
    """The default id order (TRIBIT_OPT_RELABEL 3) picks octave classes for a block model with dense
    512-id blocks and degree rank for sparse blocks, R-MAT, a cell-ordered geometric graph and graphs
    under 4,096 ids; CSR, degrees, labels and W+ equal the oracle's under the chosen ids, counts are
    exact, and the learned judge times the dense-block candidate on the block model."""
    cliques = np.concatenate([graphs.complete(k, offset=o)[0]
                              for k, o in ((70, 0), (130, 70), (256, 200), (300, 456), (90, 756))])
    cases = [
        ("sbm 40x512 p_in 0.3", oracle_mod.gen_sbm(40, 512, 0.3, 1e-4, 7), 40 * 512, 2),
        ("sbm 40x512 p_in 0.05", oracle_mod.gen_sbm(40, 512, 0.05, 1e-4, 7), 40 * 512, 1),
        ("rmat s16 ef16", oracle_mod.gen_rmat(16, 16, seed=4), 1 << 16, 1),
        ("rgg 2^16 avg 24", oracle_mod.gen_rgg(1 << 16, 24.0, 3), 1 << 16, 1),
        ("cliques on 846 ids", cliques, 846, 1),
    ]
    gpu_ctx.set_relabel("auto")
    try:
        for name, keys, n, want_mode in cases:
            note(f"auto ids: {name}")
            assert auto_mode(oracle_mod, keys, n) == want_mode, name
            want = oracle_mod.count_keys(keys, n)
            with gpu_ctx.graph_from_keys(keys, n) as g:
                assert g.id_order() == want_mode, name
                assert_same_structure(g, keys, n, oracle_mod)
                assert g.count(sched="wedge").triangles == want
                r = g.count(sched="learned")
                assert r.triangles == want
                if want_mode == 2:  # the dense-block candidate is judged
                    assert r.ms_candidates[4] > 0, r.ms_candidates
    finally:
        gpu_ctx.set_relabel(1)
