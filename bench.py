"""Benchmark: exact triangle counting GTEPS on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--scale 25] [--ef 16] [--seed 5]
                    [--sched learned|wedge|directed|plain] [--no-cpu-baseline] [--cpu-full]

Workload (default): R-MAT scale 25 (33,554,432 vertices), edge factor 16,
(a,b,c,d) = (0.57,0.19,0.19,0.05), seed 5 -- generated in HBM by the library,
symmetrised, deduplicated, self-loops removed, vertices renumbered by degree
rank, degree-oriented int32 CSR. A step is one scheduled exact count of the
resident graph (scheduler, including the cost model's choice, + counting
kernels); the count of every step must agree. The selected schedule (default
"learned": wedge ranges, per-bucket kernel variant picked by the trained cost
model in tribit_amd/cost_model.f32) is timed, then the "plain" schedule
(textbook edge iterator with the static outdeg thresholds) on the same graph.

Prints ONE JSON line. value = m / (wall-clock seconds per step) / 1e9 (GTEPS),
with m the undirected simple edges. roofline: the dominant counting kernel's
algorithmic bytes over its own rows, divided by its HIP-event device time,
against 8 TB/s: 4 bytes per owner vertex and per streamed range (edge), and per
wedge check 4 bytes (a 32-bit id) or 2 bytes (an id of a narrow wedge range,
read as its 16-bit low half; each such range also reads one byte, its high
half relative to its owner's, and a range across a 64K-id boundary two more,
the position of the crossing). This launch-level byte model is what the wedge
layout streams; the textbook figure B_alg = 4(n+1) + 4m + 4W+ over the whole
count is reported beside it (textbook_b_alg_gbs, roofline.textbook_frac).
cpu_baseline: the OpenMP CPU oracle on the host cores, its CSR build and its
count (edge iterator, merge intersection over the same oriented CSR) timed
apart, each 1 warm-up + the median of 5 runs, on stated samples extrapolated to
the whole graph (key count for the build, merge work for the count). With
--cpu-full the oracle counts every row once and its count is compared with the
GPU's (cpu_baseline.count_match). learned.ms_judge: device time of the first
count's candidate trials (the learned schedule judges a graph once).
"""
from __future__ import annotations

import argparse
import ctypes
import json
import os
import statistics
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "exact-TC GTEPS on 1 MI355X: learned-sched vs plain; % HBM roofline"
HBM_PEAK_GBS = 8000.0
def variant_name(kind: int, width: int, narrow: bool = False, tile_width: int = 64) -> str:
    """Kernel instantiation of a launch of `kind` at lane-group width `width` (buffer-load path; block
    template <log2 set slots, lanes per list, loads per lane, buffer loads, tile width (0: none), 16-bit
    narrow ranges>)."""
    n16 = "true" if narrow else "false"
    return [f"k_count_wave<6,64,{width},8,true>", f"k_count_wave<9,256,{width},8,true>",
            f"k_count_block<12,{width},8,true,0,{n16}>", f"k_count_block<12,{width},8,true,{tile_width},{n16}>",
            f"k_count_wave_multi<6,64,8,{width},8,true>"][kind]


def log(msg: str) -> None:
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def host_cpus() -> dict:
    """Physical cores of the CPUs this process may run on (distinct (package, core) pairs in its
    affinity set), the OpenMP thread count the environment asks for, and the lscpu model name."""
    cpus = sorted(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else list(range(os.cpu_count() or 1))
    phys = set()
    for c in cpus:
        base = f"/sys/devices/system/cpu/cpu{c}/topology/"
        try:
            with open(base + "physical_package_id") as f:
                pkg = f.read().strip()
            with open(base + "core_id") as f:
                core = f.read().strip()
            phys.add((pkg, core))
        except OSError:
            phys.add(("?", str(c)))
    model = "unknown"
    try:
        import subprocess

        out = subprocess.run(["lscpu"], capture_output=True, text=True, timeout=20).stdout
        for line in out.splitlines():
            if line.strip().startswith("Model name:"):
                model = line.split(":", 1)[1].strip()
                break
    except Exception:
        model = cpu_model()
    # a virtual machine may report only the brand: the family / model / stepping numbers identify the part
    ident = {}
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                k, _, v = line.partition(":")
                k = k.strip()
                if k in ("cpu family", "model", "stepping", "cpu MHz", "cache size") and k not in ident:
                    ident[k] = v.strip()
                if not line.strip():
                    break
    except OSError:
        pass
    omp = int(os.environ.get("OMP_NUM_THREADS", "0") or 0)
    return {"logical": len(cpus), "physical": len(phys), "omp_num_threads": omp or None, "model": model,
            "cpuinfo": ident}


def _median_runs(fn, runs: int = 5, warm: bool = True):
    """One untimed warm-up call (unless warm=False), then the median wall time of `runs` calls (and
    the last result)."""
    if warm:
        fn()
    times = []
    res = None
    for _ in range(runs):
        t0 = time.perf_counter()
        res = fn()
        times.append(time.perf_counter() - t0)
    return statistics.median(times), res


def cpu_baseline(g, n: int, m: int, target_s: float, triangles: int, full: bool = False, keys_fn=None,
                 nkeys: int = 0) -> dict:
    """The CPU oracle on the GPU host, in two phases timed apart, each one warm-up + the median of 5:
    CSR build (sort + unique of the input keys, degree orientation; `keys_fn` regenerates the same input
    keys on the CPU, untimed) and counting (edge iterator, merge intersection over the oriented CSR that
    the GPU built). Either phase runs on a stated sample when the whole graph would take too long per
    run and is extrapolated: the CSR build by key count (a prefix of the keys), the count by merge work
    (a strided row sample). Threads: OMP_NUM_THREADS when set (the host's CPU share), else every
    physical core. When the count sample is every row (small graphs, or full=True), the oracle's count
    is compared with the GPU's."""
    import oracle

    if not os.path.exists(oracle.LIB_PATH):
        oracle.build()
    hc = host_cpus()
    threads = hc["omp_num_threads"] or hc["physical"]
    per_run = max(target_s / 12.0, 0.2)  # 2 phases x (1 warm-up + 5 runs)
    rp32, col, _ = g.download()
    rp = rp32.astype(np.uint64)
    d = np.diff(rp).astype(np.int64)
    # merge work of the edge iterator per row: sum over v in N+(u) of (d(u) + d(v))
    dv = d[col.astype(np.int64)] if col.size else np.zeros(0, np.int64)
    nz = d > 0
    wrow = np.zeros(n, np.int64)
    if col.size:
        wrow[nz] = np.add.reduceat(dv, rp[:-1][nz].astype(np.int64))
    del dv
    work = d * d + wrow
    work[d < 2] = 0
    total_work = float(work.sum())
    og = oracle.Graph.from_csr(rp, col)
    # counting: grow the row sample (smaller stride) until one run takes about per_run seconds
    stride = 1 if full else 1 << 16
    sample_work = total_work
    while not full:
        sample_work = float(work[::stride].sum())
        t0 = time.perf_counter()
        og.count_strided(0, stride, threads)
        elapsed = time.perf_counter() - t0
        if elapsed >= per_run / 2 or stride == 1:
            break
        if elapsed < 0.02 or sample_work <= 0:
            stride = max(1, stride // 16)
            continue
        want = total_work / (sample_work / elapsed * per_run)
        new = 1
        while new * 2 <= want:
            new *= 2
        stride = max(1, min(new, stride // 2))
    runs = 1 if full else 5  # full: one whole-graph run, no warm-up (a check, not a timing protocol)
    t_count_s, sample_count = _median_runs(lambda: og.count_strided(0, stride, threads), runs, not full)
    t_count = t_count_s * total_work / max(sample_work, 1.0)
    del og, work, wrow, d, rp, col, rp32
    # CSR build: sort + unique + orientation of a prefix of the input keys sized to about per_run seconds
    csr = None
    if keys_fn is not None and nkeys > 0:
        keys = keys_fn()
        take = keys.size
        t0 = time.perf_counter()
        oracle.Graph.from_keys(keys[: min(take, 1 << 22)].copy(), n)
        probe = time.perf_counter() - t0
        probe_keys = min(take, 1 << 22)
        if probe > 0 and not full:
            take = int(min(keys.size, max(probe_keys, probe_keys * per_run / probe)))

        def build():
            return oracle.Graph.from_keys(keys[:take], n).m  # dedup works on its own copy

        t_csr_s, _ = _median_runs(build, runs, not full)
        t_csr = t_csr_s * keys.size / take
        csr = {"ms": round(t_csr * 1e3, 1), "keys_sampled": int(take), "keys": int(keys.size),
               "ms_sample_median": round(t_csr_s * 1e3, 1)}
        del keys
    out = {
        "value": round(m / t_count / 1e9, 6),
        "unit": "GTEPS",
        "cores": threads,
        "kind": "port",
        "ms_count": round(t_count * 1e3, 1),
        "ms_csr": csr["ms"] if csr else None,
        "gteps_e2e": round(m / (t_count + csr["ms"] * 1e-3) / 1e9, 6) if csr else None,
        "protocol": "each phase: 1 warm-up + median of 5 runs" if not full else "whole graph, one run per phase",
        "sample": f"count: rows u = 0 mod {stride} ({sample_work:.3g} of {total_work:.3g} merge steps, median "
                  f"{t_count_s:.2f} s per run), extrapolated by merge work; "
                  + (f"CSR build: first {csr['keys_sampled']} of {csr['keys']} input keys (median "
                     f"{csr['ms_sample_median'] / 1e3:.2f} s per run), extrapolated by key count" if csr else
                     "CSR build: not timed"),
        "cpu_model": hc["model"],
        "host_cpus": hc,
    }
    if stride == 1:
        out["oracle_triangles"] = int(sample_count)
        out["count_match"] = int(sample_count) == int(triangles)
    return out


def cpu_model() -> str:
    """The host CPU's model name (first processor in /proc/cpuinfo) and its logical CPUs."""
    name = "unknown"
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    name = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return f"{name} ({os.cpu_count()} logical CPUs visible)"


def measured_traffic(args, variant: int):
    """HBM bytes per launch of `variant` from the committed rocprofv3 counter summary of this
    workload and schedule (FETCH_SIZE x2.0 + WRITE_SIZE x1.0, scaled by a plain device copy),
    or None when no summary matches."""
    path = os.path.join(ROOT, "profiles", f"rmat{args.scale}_{args.sched}_counters.json")
    if args.ef != 16 or args.seed != (5 if args.scale == 25 else -1) or not os.path.exists(path):
        return None
    with open(path) as f:
        summ = json.load(f)
    for launch in summ.get("launches", []):
        if launch.get("slot") == variant:
            return int(launch["traffic_bytes"])
    return None


def measure(g, sched: str, steps: int, warmup: int):
    """W untimed counts, then K timed counts (wall clock around all K). The first count of the graph
    (the first warm-up) carries the learned schedule's judging cost, reported as ms_judge."""
    first = None
    for i in range(warmup):
        r = g.count(sched=sched)
        first = first or r
        log(f"{sched} warmup {i}: {r.triangles} triangles, {r.ms_count:.2f} ms (variants {r.ms_by_variant})"
            + (f"; judged candidates {[round(t, 2) for t in r.ms_candidates]} -> {r.candidate}"
               f" (judging {r.ms_judge:.1f} ms)" if sched == "learned" else ""))
    results = []
    t0 = time.perf_counter()
    for _ in range(steps):
        results.append(g.count(sched=sched))
    wall = time.perf_counter() - t0
    tri = {r.triangles for r in results}
    if len(tri) != 1:
        log(f"{sched}: counts disagree across steps: {tri}")
        return None
    nv = len(results[0].ms_by_variant)
    out = {
        "triangles": tri.pop(),
        "ms_per_step": wall / max(steps, 1) * 1e3,
        "ms_dev": statistics.median(r.ms_count for r in results),
        "ms_sched": statistics.median(r.ms_sched for r in results),
        "ms_var": [statistics.median(r.ms_by_variant[v] for r in results) for v in range(nv)],
        "tasks_by_variant": results[0].tasks_by_variant,
        "group_width": list(results[0].group_width),
        "slot_kind": list(results[0].slot_kind),
        "candidate": results[0].candidate,
        "layout": results[0].layout,
        "dense_core": results[0].dense_core,
        "ms_dense": statistics.median(r.ms_dense for r in results),
        "dense_blocks": results[0].dense_blocks,
        "ms_judge": (first or results[0]).ms_judge,
        "ms_first_count": (first or results[0]).ms_count,
        "ms_candidates": list((first or results[0]).ms_candidates),
    }
    log(f"{sched}: {steps} steps, {out['ms_per_step']:.2f} ms/step wall, {out['ms_dev']:.2f} ms device; "
        f"variants {[round(x, 3) for x in out['ms_var']]}")
    return out


def copy_calibration() -> int:
    """A plain device copy of a known size: FETCH_SIZE/WRITE_SIZE of its kernel scale the counters."""
    import torch

    n = 1 << 28
    src = torch.ones(n, dtype=torch.float32, device="cuda")
    dst = torch.empty_like(src)
    torch.cuda.synchronize()
    dst.copy_(src)
    torch.cuda.synchronize()
    print(json.dumps({"copy_bytes_read": n * 4, "copy_bytes_written": n * 4}), flush=True)
    return 0


MIXED = [
    ("rmat", "R-MAT scale 20, edge factor 16, seed 7", lambda c: c.gen_rmat(20, 16, 0.57, 0.19, 0.19, 7)),
    ("er", "Erdos-Renyi n=1,048,576, avg degree 32 (G(n,M), M=16,777,216), seed 7",
     lambda c: c.gen_er(1 << 20, 1 << 24, 7)),
    ("rgg", "2-D random geometric n=2,097,152, avg degree 24, seed 7", lambda c: c.gen_rgg(1 << 21, 24.0, 7)),
    ("sbm", "SBM 1000 blocks x 512, p_in 0.3, p_out 1e-4, seed 7", lambda c: c.gen_sbm(1000, 512, 0.3, 1e-4, 7)),
]


def warm_up(ctx) -> None:
    """Untimed: the process's first copies and launches (one-time runtime set-up) on a small graph, so
    that a reported preprocessing time is the graph's own whatever ran before it in the process."""
    with ctx.gen_rmat(12, 8, 0.57, 0.19, 0.19, 1) as w:
        w.count(sched="plain")


ID_ORDER = {0: "input ids", 1: "degree-rank ids", 2: "octave-class ids"}


def mixed_family(args) -> int:
    """One JSON line per member of the mixed family (plain vs selected schedule)."""
    import tribit_amd

    ctx = tribit_amd.Context(0)
    if args.sched == "learned":
        log(f"cost model loaded: {ctx.load_cost_model()}")
    warm_up(ctx)
    ctx.set_relabel(args.relabel)
    for name, desc, make in MIXED:
        g = make(ctx)
        inf = g.info()
        log(f"{name}: n={inf['n']} m={inf['m']} W+={inf['wplus']} prep={inf['ms_preprocess']:.1f} ms")
        sel = measure(g, args.sched, args.steps, args.warmup)
        plain = measure(g, "plain", args.steps, args.warmup)
        if sel is None or plain is None or sel["triangles"] != plain["triangles"]:
            return 1
        m = inf["m"]
        after = g.info()
        check = cpu_baseline(g, inf["n"], m, args.cpu_seconds, sel["triangles"], True) if args.cpu_full else None
        print(json.dumps({
            "metric": METRIC, "workload": desc, "n": inf["n"], "m": m, "wplus": inf["wplus"],
            "triangles": sel["triangles"], "sched": args.sched,
            "gteps": round(m / (sel["ms_per_step"] * 1e-3) / 1e9, 4), "ms_count_device": round(sel["ms_dev"], 4),
            "ms_by_variant": [round(x, 4) for x in sel["ms_var"]],
            "plain_gteps": round(m / (plain["ms_per_step"] * 1e-3) / 1e9, 4),
            "plain_ms_count_device": round(plain["ms_dev"], 4),
            "ms_preprocess_device": round(inf["ms_preprocess"], 3),
            # preprocessing of the path that ran: the graph's build, the range layout its plan streams when
            # that one is built on first use, and the dense blocks' structures when it counts them
            "ms_preprocess_selected": round(inf["ms_preprocess"] + after["ms_layout"][sel["layout"]] +
                                            (after["ms_blocks"] if sel.get("dense_blocks", 0) else 0.0), 3),
            "ms_dense_blocks_prepare": round(after["ms_blocks"], 3),
            "relabel": args.relabel, "ids": ID_ORDER[g.id_order()], "candidate": sel["candidate"],
            "ms_judge": round(sel["ms_judge"], 3),
            "ms_candidates": [round(x, 3) for x in sel["ms_candidates"]],
            "dense_blocks": sel.get("dense_blocks", 0), "ms_dense": round(sel["ms_dense"], 4),
            "cpu_check": check,
        }), flush=True)
        g.free()
        if check is not None and not check.get("count_match"):
            log(f"{name}: GPU count differs from the CPU oracle's")
            return 1
    ctx.close()
    return 0


def batch_workload(args) -> int:
    """Configuration 3: 4,096 Chung-Lu graphs (seed 3) in one batch: per-graph exact counts from the
    counting launches of every part captured in one hipGraph, before and after the autotuner."""
    import tribit_amd

    ctx = tribit_amd.Context(0)
    warm_up(ctx)
    t0 = time.perf_counter()
    b = ctx.batch_gen_chung_lu(3, 0, args.batch_graphs)
    inf = b.info()
    log(f"batch of {inf['graphs']} Chung-Lu graphs in {inf['parts']} parts: n={inf['n']} m={inf['m']}, "
        f"built in {time.perf_counter() - t0:.1f} s (device preprocessing {inf['ms_preprocess']:.1f} ms)")
    sched = args.batch_sched
    counts, first = b.count(sched=sched, repeats=args.steps)
    log(f"count: {first['ms_count']:.2f} ms per replay ({first['launches']} launches, capture "
        f"{first['ms_capture']:.1f} ms)")
    tuned = b.autotune(sched=sched, repeats=3)
    counts2, after = b.count(sched=sched, repeats=args.steps)
    if not (np.array_equal(counts, counts2) and first["repeats_agree"] and after["repeats_agree"]):
        log("batch counts disagree")
        return 1
    m = inf["m"]
    # roofline of the whole replay: the wedge byte model of every part's lists over the replay time
    work = b.variant_stats(sched)
    nar = sum(work["narrow"])
    b_all = (4.0 * sum(work["rows"]) + 4.0 * sum(work["edges"]) + 4.0 * (sum(work["wplus"]) - nar) + 2.0 * nar +
             1.0 * sum(work["narrow_ranges"]) + 2.0 * sum(work["narrow2_ranges"]))
    achieved = b_all / (after["ms_count"] * 1e-3) / 1e9
    oracle_check = None
    if args.cpu_full:
        oracle_check = batch_oracle_check(counts, inf["graphs"])
    print(json.dumps({
        "metric": METRIC, "workload": f"{inf['graphs']} Chung-Lu graphs, batch seed 3 (gamma ~ U[2.1,2.8], n "
                                      f"log-uniform in [1024, 262144], avg degree ~ U[8,64])",
        "graphs": inf["graphs"], "parts": inf["parts"], "n": inf["n"], "m": m, "triangles": int(counts.sum()),
        "sched": sched, "launches_per_replay": after["launches"],
        "gteps": round(m / (after["ms_count"] * 1e-3) / 1e9, 4), "ms_per_replay": round(after["ms_count"], 4),
        "gteps_before_autotune": round(m / (first["ms_count"] * 1e-3) / 1e9, 4),
        "ms_per_replay_before_autotune": round(first["ms_count"], 4),
        "autotuned": {"group_width": after["group_width"], "waves_per_cu": after["waves"],
                      "tile_width": after.get("tile_width", 0)},
        "roofline": {"bound": "hbm", "kernel": f"whole replay ({after['launches']} launches)",
                     "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": None, "algorithmic_bytes": b_all,
                     "byte_model": "every part's wedge lists: 4 B per owner list and per streamed range, per "
                                   "streamed element 4 B (32-bit id) or 2 B (16-bit half of a narrow range), 1 B "
                                   "per narrow range, 2 B more per range across a 64K-id boundary; block launches "
                                   "that stream bit tiles (tile_width > 0) read 12 B per tile instead"},
        "ms_preprocess_device": round(inf["ms_preprocess"], 2),
        "cpu_check": oracle_check,
    }), flush=True)
    b.free()
    ctx.close()
    if oracle_check is not None and not oracle_check["count_match"]:
        log(f"per-graph counts differ from the CPU oracle's on {oracle_check['mismatches']} graphs")
        return 1
    return 0


def batch_oracle_check(counts, ngraphs: int) -> dict:
    """Every graph of the batch regenerated and counted by the CPU oracle (same generator streams)."""
    import oracle

    if not os.path.exists(oracle.LIB_PATH):
        oracle.build()
    threads = max(1, min(int(os.environ.get("OMP_NUM_THREADS", "0") or 0) or min(16, os.cpu_count() or 1), 16))
    t0 = time.perf_counter()
    bad = 0
    total = 0
    for i in range(ngraphs):
        gamma, n, avg = oracle.chung_lu_batch_params(3, i)
        keys = oracle.gen_chung_lu(n, gamma, avg, (3 << 32) + i)
        c = oracle.count_keys(keys, n, threads)
        total += c
        bad += int(c != int(counts[i]))
        if (i + 1) % 256 == 0:
            log(f"oracle check: {i + 1}/{ngraphs} graphs, {bad} mismatches, {time.perf_counter() - t0:.0f} s")
    return {"graphs": ngraphs, "mismatches": bad, "count_match": bad == 0, "oracle_triangles": total,
            "seconds": round(time.perf_counter() - t0, 1), "cores": threads}


def rank_tasks(keys: np.ndarray, n: int, P: int, pc: int, threshold: int = 1_000_000_000):
    """The reference's multi-GPU layout (main.rs:97-220): block size 32 * ceil((n / P) / 32), edge
    (min, max) in block (max // bs, min // bs); the diagonal task of partition pc (blocks (pc, 0..pc)) and
    the first off-diagonal group of column partition pc (partitions pc+1.. while their nnz sum stays
    under `threshold`; idxs = blocks (s, 0..pc) of each). Returns (bs, cols chunks, idxs chunks, S)."""
    bs = -(-(n // P) // 32) * 32
    keys = keys[keys != np.uint64(0xFFFFFFFFFFFFFFFF)]
    pmax = ((keys & np.uint64(0xFFFFFFFF)) // np.uint64(bs)).astype(np.int64)
    pmin = ((keys >> np.uint64(32)) // np.uint64(bs)).astype(np.int64)
    nparts = -(-n // bs)
    sel = pmax == pc
    cols = [keys[sel & (pmin == j)] for j in range(pc + 1)]
    group, acc = [], 0
    below = pmin <= pc
    nnz = np.bincount(pmax[below], minlength=nparts)  # edges of blocks (s, 0..pc) per later partition s
    for s in range(pc + 1, nparts):
        acc += int(nnz[s])
        if acc >= threshold and group:
            break
        group.append(s)
    idxs = []
    for s in group:
        m = (pmax == s) & below
        idxs += [keys[m & (pmin == j)] for j in range(pc + 1)]
    return bs, cols, idxs, group


def rank_workload(args) -> int:
    """One rank's share of the reference's multi-GPU decomposition at its real shape: the R-MAT graph of
    --scale/--ef/--seed (default configuration 5) in P = --partitions partitions (the reference's default
    64, run_multi_gpu.py:127-129); the diagonal task of partition --partition and its first off-diagonal
    group (main.rs:186-220), each through the reference-named C ABI (count_partial_triangles_diagonal /
    _off_diagonal, main5_multi.cu:589, main6_multi.cu:595) --steps times after --warmup. One JSON line
    per task: GTEPS over the task's distinct edges and device time, and the textbook roofline of its
    count (4 B per vertex slot, per list and per streamed element)."""
    import tribit_amd

    ctx = tribit_amd.Context(0)
    warm_up(ctx)
    n = 1 << args.scale
    t0 = time.perf_counter()
    keys = ctx.gen_rmat_keys(args.scale, args.ef, 0.57, 0.19, 0.19, args.seed)
    bs, cols, idxs, group = rank_tasks(keys, n, args.partitions, args.partition)
    del keys
    log(f"R-MAT s{args.scale} ef{args.ef} seed {args.seed}: {args.partitions} partitions of {bs} ids; partition "
        f"{args.partition}: {sum(c.size for c in cols)} column keys, off-diagonal group {group[0] if group else '-'}"
        f"..{group[-1] if group else '-'} with {sum(x.size for x in idxs)} index keys "
        f"({time.perf_counter() - t0:.1f} s to generate and partition)")
    lib = ctx.lib
    lib.init_gpu_kernel5(0)
    c = tribit_amd.Chunks(cols)
    x = tribit_amd.Chunks(idxs)
    out_rc = 0
    for kind in ("diagonal", "off_diagonal"):
        if kind == "off_diagonal" and not group:
            continue
        rows = []
        for i in range(args.warmup + args.steps):
            w0 = time.perf_counter()
            if kind == "diagonal":
                t = lib.count_partial_triangles_diagonal(ctypes.addressof(c.ptrs), c.cum.ctypes.data, c.num, n)
            else:
                t = lib.count_partial_triangles_off_diagonal(ctypes.addressof(c.ptrs), c.cum.ctypes.data, c.num,
                                                             ctypes.addressof(x.ptrs), x.cum.ctypes.data, x.num, n)
            wall = (time.perf_counter() - w0) * 1e3
            st = lib.tribit_partial_last_status()
            if st != 0:
                log(f"{kind}: status {st}")
                return 1
            info = tribit_amd.partial_last_info(ctx)
            info["ms_wall"] = wall
            if i >= args.warmup:
                rows.append(info)
        tri = {r["triangles"] for r in rows}
        if len(tri) != 1:
            log(f"{kind}: counts disagree across steps {tri}")
            return 1
        med = {k: statistics.median(r[k] for r in rows) for k in ("ms_h2d", "ms_build", "ms_count", "ms_wall")}
        r0 = rows[0]
        b_alg = 4.0 * (n + 1) + 4.0 * r0["lists"] + 4.0 * r0["streamed"]
        ach = b_alg / (med["ms_count"] * 1e-3) / 1e9
        print(json.dumps({
            "metric": METRIC, "workload": f"one rank's {kind} task: R-MAT scale {args.scale}, edge factor {args.ef}, "
                                          f"seed {args.seed}; P={args.partitions} partitions of {bs} ids; "
                                          f"column partition {args.partition}"
                                          + (f", index partitions {group[0]}..{group[-1]}" if kind != "diagonal"
                                             else ""),
            "task": kind, "keys": r0["keys"], "edges": r0["edges"], "lists": r0["lists"],
            "streamed": r0["streamed"], "triangles": r0["triangles"],
            "gteps": round(r0["edges"] / (med["ms_count"] * 1e-3) / 1e9, 4),
            "gteps_e2e_device": round(r0["edges"] / ((med["ms_build"] + med["ms_count"]) * 1e-3) / 1e9, 4),
            "ms_count_device": round(med["ms_count"], 3), "ms_build_device": round(med["ms_build"], 3),
            "ms_h2d": round(med["ms_h2d"], 2), "ms_call_wall": round(med["ms_wall"], 2),
            "roofline": {"bound": "hbm", "achieved": round(ach, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                         "frac": round(ach / HBM_PEAK_GBS, 4),
                         "byte_model": "textbook: 4 B per vertex slot, per streamed list and per streamed element "
                                       "over the task's count time"},
            "steps": args.steps, "warmup": args.warmup,
        }), flush=True)
    ctx.close()
    return out_rc


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--scale", type=int, default=25)
    ap.add_argument("--ef", type=int, default=16)
    ap.add_argument("--seed", type=int, default=5)
    ap.add_argument("--sched", choices=["plain", "directed", "wedge", "learned"], default="learned")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-plain", action="store_true", help="skip the plain-schedule comparison run")
    ap.add_argument("--cpu-seconds", type=float, default=20.0,
                    help="CPU baseline budget: about this many seconds over both phases' timed runs")
    ap.add_argument("--cpu-full", action="store_true",
                    help="run the CPU oracle over every row (mixed: every graph whole; batch: every graph) and "
                         "compare its counts with the GPU's")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only run a plain 1 GiB device copy (reference kernel for HBM byte counters)")
    ap.add_argument("--batch-graphs", type=int, default=4096, help="graphs of the batch workload")
    ap.add_argument("--batch-sched", choices=["plain", "directed", "wedge", "learned"], default="wedge",
                    help="batch workload: the schedule of every part")
    ap.add_argument("--relabel", type=int, default=3, choices=[0, 1, 2, 3],
                    help="mixed workload: vertex ids (1 degree rank, 2 octave classes of the degree, 0 input ids, "
                         "3 the library's automatic choice between 1 and 2)")
    ap.add_argument("--partitions", type=int, default=64, help="rank workload: partitions P")
    ap.add_argument("--partition", type=int, default=0, help="rank workload: the column partition")
    ap.add_argument("--workload", choices=["rmat", "mixed", "batch", "rank"], default="rmat",
                    help="rmat: one R-MAT graph (--scale/--ef/--seed); mixed: the four-family set "
                         "(R-MAT s20, Erdos-Renyi 2^20 avg 32, 2-D RGG 2^21 avg 24, SBM 1000x512), one line each; "
                         "batch: 4,096 Chung-Lu graphs counted as one batch; rank: one rank's diagonal and "
                         "off-diagonal task of the multi-GPU decomposition (--partitions, --partition)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for {args.gpus}); the multi-GPU path is a design, not a build.",
              file=sys.stderr)
        return 2
    if args.copy_calibration:
        return copy_calibration()
    if args.workload == "mixed":
        return mixed_family(args)
    if args.workload == "batch":
        return batch_workload(args)
    if args.workload == "rank":
        return rank_workload(args)

    import tribit_amd

    tribit_amd.load_library()
    ctx = tribit_amd.Context(0)
    model = ctx.load_cost_model() if args.sched == "learned" else False
    if args.sched == "learned":
        log(f"cost model loaded: {model}")
    dev = ctx.device_info()
    warm_up(ctx)
    log(f"device {dev}; generating R-MAT s{args.scale} ef{args.ef} seed {args.seed}")
    t0 = time.perf_counter()
    g = ctx.gen_rmat(args.scale, args.ef, 0.57, 0.19, 0.19, args.seed)
    inf = g.info()
    log(f"graph ready in {time.perf_counter() - t0:.2f} s: n={inf['n']} m={inf['m']} W+={inf['wplus']} "
        f"prep={inf['ms_preprocess']:.1f} ms")
    n, m = inf["n"], inf["m"]

    sel = measure(g, args.sched, args.steps, args.warmup)
    if sel is None:
        return 1
    # preprocessing of the path that ran: the graph's build, plus the range layout the selected count
    # streams when that one is built on first use (the default wedge layout is part of the build)
    ms_layout = g.info()["ms_layout"]
    ms_prep = inf["ms_preprocess"] + ms_layout[sel["layout"]]
    plain = None
    if args.sched != "plain" and not args.no_plain:
        plain = measure(g, "plain", args.steps, args.warmup)
        if plain is None or plain["triangles"] != sel["triangles"]:
            log("plain and selected schedules disagree")
            return 1
    tri = sel["triangles"]
    ms_per_step = sel["ms_per_step"]
    ms_dev = sel["ms_dev"]
    ms_var = sel["ms_var"]
    work = g.variant_stats(args.sched)
    dom = int(np.argmax(ms_var))

    def launch_bytes(i):
        nar = work["narrow"][i]
        return (4.0 * work["rows"][i] + 4.0 * work["edges"][i] + 4.0 * (work["wplus"][i] - nar) + 2.0 * nar +
                1.0 * work["narrow_ranges"][i] + 2.0 * work["narrow2_ranges"][i])

    nar = work["narrow"][dom]
    b_dom = launch_bytes(dom)
    achieved = b_dom / (ms_var[dom] * 1e-3) / 1e9 if ms_var[dom] > 0 else 0.0
    # every launch of the step with work, by the same byte model
    by_launch = []
    for i, ms in enumerate(ms_var):
        if i < len(sel["tasks_by_variant"]) and sel["tasks_by_variant"][i] and ms > 0:
            b = launch_bytes(i)
            by_launch.append({"kernel": variant_name(sel["slot_kind"][i], sel["group_width"][i], work["narrow"][i] > 0),
                              "ms": round(ms, 4), "algorithmic_bytes": b,
                              "achieved": round(b / (ms * 1e-3) / 1e9, 1),
                              "frac": round(b / (ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4)})
    b_alg = 4.0 * (n + 1) + 4.0 * m + 4.0 * inf["wplus"]

    out = {
        "metric": METRIC,
        "value": round(m / (ms_per_step * 1e-3) / 1e9, 4),
        "unit": "GTEPS",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "int32",
        "data": "synthetic (R-MAT generated in HBM by the library's counter-based generator)",
        "config": {
            "workload": f"R-MAT scale {args.scale}, edge factor {args.ef}, (a,b,c,d)=(0.57,0.19,0.19,0.05), "
                        f"seed {args.seed}; symmetrised, deduplicated, self-loops removed; {ID_ORDER[g.id_order()]}; "
                        f"degree-oriented int32 CSR",
            "n": n, "m": m, "wplus": inf["wplus"], "sched": args.sched, "parallelism": "single GPU",
            "cost_model": "tribit_amd/cost_model.f32" if model else None,
        },
        "triangles": tri,
        "ms_count_device": round(ms_dev, 4),
        "ms_sched_device": round(sel["ms_sched"], 4),
        "ms_preprocess_device": round(ms_prep, 3),
        "ms_layouts_on_first_use": {"plain": round(ms_layout[0], 3), "directed": round(ms_layout[2], 3)},
        "gteps_device": round(m / (ms_dev * 1e-3) / 1e9, 4),
        "gteps_e2e": round(m / ((ms_dev + ms_prep) * 1e-3) / 1e9, 4),
        "textbook_b_alg_gbs": round(b_alg / (ms_dev * 1e-3) / 1e9, 1),
        "ms_by_variant": [round(x, 4) for x in ms_var],
        "tasks_by_variant": sel["tasks_by_variant"],
        "learned": None if args.sched != "learned" else {
            "candidate": ["cost-model plan", "heuristic plan", "heuristic plan, directed layout",
                          "textbook plain plan", "heuristic plan with dense diagonal blocks"][sel["candidate"]],
            "layout": {0: "plain", 2: "directed", 3: "wedge"}[sel["layout"]],
            "dense_core_vertices": sel["dense_core"], "ms_dense_core": round(sel["ms_dense"], 4),
            "ms_judge": round(sel["ms_judge"], 3),
            "ms_candidates": [round(x, 3) for x in sel["ms_candidates"]],
        },
        "wedge_checks": int(sum(work["wplus"])),
        "plain": None if plain is None else {
            "gteps": round(m / (plain["ms_per_step"] * 1e-3) / 1e9, 4),
            "ms_per_step": round(plain["ms_per_step"], 4),
            "ms_count_device": round(plain["ms_dev"], 4),
            "ms_by_variant": [round(x, 4) for x in plain["ms_var"]],
        },
        "speedup_vs_plain": None if plain is None else round(plain["ms_per_step"] / ms_per_step, 3),
        "roofline": {
            "bound": "hbm",
            "kernel": variant_name(sel["slot_kind"][dom], sel["group_width"][dom], nar > 0),
            "achieved": round(achieved, 1),
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4),
            "traffic": measured_traffic(args, dom),
            "algorithmic_bytes": b_dom,
            "byte_model": "this launch's own work: 4 B per owner list and per streamed range, per streamed "
                          "element 4 B (32-bit id) or 2 B (16-bit half of a narrow wedge range), 1 B per "
                          "narrow range, 2 B more per range across a 64K-id boundary (not the textbook B_alg)",
            "ms_per_launch": round(ms_var[dom], 4),
            "textbook_b_alg": b_alg,
            "textbook_frac": round(b_alg / (ms_dev * 1e-3) / 1e9 / HBM_PEAK_GBS, 3),
            "by_launch": by_launch,
        },
        "cpu_baseline": None,
        "device": dev["name"],
    }
    if not args.no_cpu_baseline:
        log("cpu baseline (oracle: CSR build and count, each 1 warm-up + median of 5, on stated samples)")

        def regen_keys():
            import oracle

            return oracle.gen_rmat(args.scale, args.ef, 0.57, 0.19, 0.19, args.seed)

        out["cpu_baseline"] = cpu_baseline(g, n, m, args.cpu_seconds, tri, args.cpu_full, keys_fn=regen_keys,
                                           nkeys=inf["entries"])
    g.free()
    ctx.close()
    print(json.dumps(out), flush=True)
    if out["cpu_baseline"] and out["cpu_baseline"].get("count_match") is False:
        log(f"GPU count {tri} differs from the CPU oracle's {out['cpu_baseline']['oracle_triangles']}")
        return 1
    return 0


if __name__ == "__main__":
    sys.exit(main())
