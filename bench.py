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
the position of the crossing). cpu_baseline: the OpenMP CPU oracle
(edge iterator, merge intersection) timed on a strided row sample of the same
CSR, extrapolated to the whole graph by merge work. When that sample is every
row (small graphs, or --cpu-full), the oracle's count is compared with the
GPU's (cpu_baseline.count_match).
"""
from __future__ import annotations

import argparse
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
def variant_name(kind: int, width: int, narrow: bool = False) -> str:
    """Kernel instantiation of a launch of `kind` at lane-group width `width` (buffer-load path)."""
    n16 = "true" if narrow else "false"
    return [f"k_count_wave<6,64,{width},8,true>", f"k_count_wave<9,256,{width},8,true>",
            f"k_count_block<12,{width},8,true,false,{n16}>", f"k_count_block<12,{width},8,true,true,{n16}>",
            f"k_count_wave_multi<6,64,8,{width},8,true>"][kind]


def log(msg: str) -> None:
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def cpu_baseline(g, n: int, m: int, target_s: float, triangles: int, full: bool = False) -> dict:
    """Times the CPU oracle on a strided sample of rows of the same oriented CSR. When the sample is
    every row (small graphs, or full=True), the oracle's count is compared with the GPU's."""
    import oracle

    if not os.path.exists(oracle.LIB_PATH):
        oracle.build()
    threads = int(os.environ.get("OMP_NUM_THREADS", "0") or 0) or min(16, os.cpu_count() or 1)
    threads = max(1, min(threads, 16))
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
    # grow the sample (smaller stride) until it takes about target_s
    stride = 1 if full else 1 << 16
    while True:
        sample_work = float(work[::stride].sum())
        t0 = time.perf_counter()
        sample_count = og.count_strided(0, stride, threads)
        elapsed = time.perf_counter() - t0
        if elapsed >= target_s / 3 or stride == 1:
            break
        if elapsed < 0.05 or sample_work <= 0:
            stride = max(1, stride // 16)
            continue
        want = total_work / (sample_work / elapsed * target_s)
        new = 1
        while new * 2 <= want:
            new *= 2
        stride = max(1, min(new, stride // 2))
    t_full = elapsed * total_work / max(sample_work, 1.0)
    out = {
        "value": round(m / t_full / 1e9, 6),
        "unit": "GTEPS",
        "cores": threads,
        "kind": "port",
        "sample": f"rows u = 0 mod {stride} ({int(np.count_nonzero(d[::stride] >= 2))} active rows, "
                  f"{sample_work:.3g} of {total_work:.3g} merge steps) counted in {elapsed:.2f} s, "
                  f"extrapolated by merge work to {t_full:.1f} s for the whole graph",
        "cpu_model": cpu_model(),
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
    """W untimed counts, then K timed counts (wall clock around all K)."""
    for i in range(warmup):
        r = g.count(sched=sched)
        log(f"{sched} warmup {i}: {r.triangles} triangles, {r.ms_count:.2f} ms (variants {r.ms_by_variant})"
            + (f"; judged candidates {[round(t, 2) for t in r.ms_candidates]} -> {r.candidate}"
               if sched == "learned" else ""))
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


def mixed_family(args) -> int:
    """One JSON line per member of the mixed family (plain vs selected schedule)."""
    import tribit_amd

    ctx = tribit_amd.Context(0)
    if args.sched == "learned":
        log(f"cost model loaded: {ctx.load_cost_model()}")
    warm_up(ctx)
    for name, desc, make in MIXED:
        g = make(ctx)
        inf = g.info()
        log(f"{name}: n={inf['n']} m={inf['m']} W+={inf['wplus']} prep={inf['ms_preprocess']:.1f} ms")
        sel = measure(g, args.sched, args.steps, args.warmup)
        plain = measure(g, "plain", args.steps, args.warmup)
        if sel is None or plain is None or sel["triangles"] != plain["triangles"]:
            return 1
        m = inf["m"]
        check = cpu_baseline(g, inf["n"], m, args.cpu_seconds, sel["triangles"], True) if args.cpu_full else None
        print(json.dumps({
            "metric": METRIC, "workload": desc, "n": inf["n"], "m": m, "wplus": inf["wplus"],
            "triangles": sel["triangles"], "sched": args.sched,
            "gteps": round(m / (sel["ms_per_step"] * 1e-3) / 1e9, 4), "ms_count_device": round(sel["ms_dev"], 4),
            "ms_by_variant": [round(x, 4) for x in sel["ms_var"]],
            "plain_gteps": round(m / (plain["ms_per_step"] * 1e-3) / 1e9, 4),
            "plain_ms_count_device": round(plain["ms_dev"], 4),
            "ms_preprocess_device": round(inf["ms_preprocess"], 3),
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
    sched = "wedge"
    counts, first = b.count(sched=sched, repeats=args.steps)
    log(f"count: {first['ms_count']:.2f} ms per replay ({first['launches']} launches, capture "
        f"{first['ms_capture']:.1f} ms)")
    tuned = b.autotune(sched=sched, repeats=3)
    counts2, after = b.count(sched=sched, repeats=args.steps)
    if not (np.array_equal(counts, counts2) and first["repeats_agree"] and after["repeats_agree"]):
        log("batch counts disagree")
        return 1
    m = inf["m"]
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
        "autotuned": {"group_width": after["group_width"], "waves_per_cu": after["waves"]},
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
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--cpu-full", action="store_true",
                    help="run the CPU oracle over every row (mixed: every graph whole; batch: every graph) and "
                         "compare its counts with the GPU's")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only run a plain 1 GiB device copy (reference kernel for HBM byte counters)")
    ap.add_argument("--batch-graphs", type=int, default=4096, help="graphs of the batch workload")
    ap.add_argument("--workload", choices=["rmat", "mixed", "batch"], default="rmat",
                    help="rmat: one R-MAT graph (--scale/--ef/--seed); mixed: the four-family set "
                         "(R-MAT s20, Erdos-Renyi 2^20 avg 32, 2-D RGG 2^21 avg 24, SBM 1000x512), one line each; "
                         "batch: 4,096 Chung-Lu graphs counted as one batch")
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
    nar = work["narrow"][dom]
    b_dom = (4.0 * work["rows"][dom] + 4.0 * work["edges"][dom] + 4.0 * (work["wplus"][dom] - nar) + 2.0 * nar +
             1.0 * work["narrow_ranges"][dom] + 2.0 * work["narrow2_ranges"][dom])
    achieved = b_dom / (ms_var[dom] * 1e-3) / 1e9 if ms_var[dom] > 0 else 0.0
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
                        f"seed {args.seed}; symmetrised, deduplicated, self-loops removed; degree-rank ids; "
                        f"degree-oriented int32 CSR",
            "n": n, "m": m, "wplus": inf["wplus"], "sched": args.sched, "parallelism": "single GPU",
            "cost_model": "tribit_amd/cost_model.f32" if model else None,
        },
        "triangles": tri,
        "ms_count_device": round(ms_dev, 4),
        "ms_sched_device": round(sel["ms_sched"], 4),
        "ms_preprocess_device": round(inf["ms_preprocess"], 3),
        "gteps_device": round(m / (ms_dev * 1e-3) / 1e9, 4),
        "gteps_e2e": round(m / ((ms_dev + inf["ms_preprocess"]) * 1e-3) / 1e9, 4),
        "textbook_b_alg_gbs": round(b_alg / (ms_dev * 1e-3) / 1e9, 1),
        "ms_by_variant": [round(x, 4) for x in ms_var],
        "tasks_by_variant": sel["tasks_by_variant"],
        "learned": None if args.sched != "learned" else {
            "candidate": ["cost-model plan", "heuristic plan", "heuristic plan, directed layout",
                          "textbook plain plan"][sel["candidate"]],
            "layout": {0: "plain", 2: "directed", 3: "wedge"}[sel["layout"]],
            "dense_core_vertices": sel["dense_core"], "ms_dense_core": round(sel["ms_dense"], 4),
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
            "ms_per_launch": round(ms_var[dom], 4),
        },
        "cpu_baseline": None,
        "device": dev["name"],
    }
    if not args.no_cpu_baseline:
        log("cpu baseline (oracle on a row sample)")
        out["cpu_baseline"] = cpu_baseline(g, n, m, args.cpu_seconds, tri, args.cpu_full)
    g.free()
    ctx.close()
    print(json.dumps(out), flush=True)
    if out["cpu_baseline"] and out["cpu_baseline"].get("count_match") is False:
        log(f"GPU count {tri} differs from the CPU oracle's {out['cpu_baseline']['oracle_triangles']}")
        return 1
    return 0


if __name__ == "__main__":
    sys.exit(main())
