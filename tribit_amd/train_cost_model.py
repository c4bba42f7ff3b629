"""Trains the learned scheduler's cost model on the GPU from timing logs.

    python -m tribit_amd.train_cost_model [--out tribit_amd/cost_model.f32] [--quick]

1. Builds a set of training graphs on the device (R-MAT at several scales and
   seeds, Chung-Lu, 2-D random geometric, stochastic block model, Erdos-Renyi;
   the R-MAT scale-25 member uses seed 1, not the benchmark's seed 5).
2. Counts each graph under exploration plans: in round r every owner-outdeg
   class (<= 32, <= 256, larger) runs on one of its feasible menu slots, the
   rounds cycling through them, so each (bucket, slot) pair is timed in plans
   of the shape the scheduler chooses from; the kernels log every task's
   wall-clock duration, summed per (slot, bucket).
3. One sample per timed (bucket, slot): the library's feature vector and the
   target log2(nanoseconds of the slot's launch per unit of work), where the
   bucket's share of its launch's device time is the share of the launch's
   summed task durations its tasks took, and work = streamed elements +
   streamed ranges. Launch shares carry what per-task durations alone miss:
   the kernel's occupancy and the contention of the tasks running together.
4. Trains the 3-layer perceptron with the library's HIP kernels (FP32, Adam)
   on all graphs but one held out, reports the held-out fit, then retrains on
   all graphs and writes the parameters (float32, little endian) and a JSON
   summary next to them.
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

import tribit_amd as tb  # noqa: E402

DEFAULT_OUT = os.path.join(HERE, "cost_model.f32")

TRAIN_GRAPHS = [
    ("rmat", dict(scale=18, ef=16, seed=11)),
    ("rmat", dict(scale=20, ef=16, seed=12)),
    ("rmat", dict(scale=22, ef=16, seed=13)),
    ("rmat", dict(scale=23, ef=16, seed=14)),
    ("rmat", dict(scale=24, ef=16, seed=15)),
    ("rmat", dict(scale=25, ef=16, seed=1)),
    ("rmat", dict(scale=21, ef=8, seed=16)),
    ("chung_lu", dict(n=1 << 20, gamma=2.1, avg_deg=48.0, seed=21)),
    ("chung_lu", dict(n=1 << 19, gamma=2.5, avg_deg=16.0, seed=22)),
    ("rgg", dict(n=1 << 20, avg_deg=24.0, seed=23)),
    ("sbm", dict(blocks=400, size=512, p_in=0.3, p_out=1e-4, seed=24)),
    ("er", dict(n=1 << 20, samples=1 << 24, seed=25)),
]
QUICK_GRAPHS = [TRAIN_GRAPHS[0], TRAIN_GRAPHS[1], TRAIN_GRAPHS[7], TRAIN_GRAPHS[10]]


def make_graph(ctx, kind, p):
    if kind == "rmat":
        return ctx.gen_rmat(p["scale"], p["ef"], seed=p["seed"])
    if kind == "er":
        return ctx.gen_er(p["n"], p["samples"], p["seed"])
    if kind == "chung_lu":
        return ctx.gen_chung_lu(p["n"], p["gamma"], p["avg_deg"], p["seed"])
    if kind == "rgg":
        return ctx.gen_rgg(p["n"], p["avg_deg"], p["seed"])
    if kind == "sbm":
        return ctx.gen_sbm(p["blocks"], p["size"], p["p_in"], p["p_out"], p["seed"])
    raise ValueError(kind)


def feasible(bucket: int, menu) -> list:
    top = tb.bucket_max_outdeg(bucket)
    return [s for s, (_, _, dmax) in enumerate(menu) if top <= dmax]


def outdeg_class(bucket: int) -> int:
    """The heuristic's class of a bucket: owner outdeg <= 32, <= 256, larger."""
    top = tb.bucket_max_outdeg(bucket)
    return 0 if top <= 32 else 1 if top <= 256 else 2


def exploration_tables(menu, rounds: int):
    """Round r runs every outdeg class on one slot, its r-th feasible one (the
    classes' slot lists rotated against each other), the shape of the plans
    the scheduler picks from: one launch per class, classes sharing a launch
    when they pick the same slot."""
    cls = [outdeg_class(b) for b in range(tb.BUCKETS)]
    feas = []
    for c in range(3):
        bs = [b for b in range(tb.BUCKETS) if cls[b] == c]
        feas.append([s for s in range(len(menu)) if all(tb.bucket_max_outdeg(b) <= menu[s][2] for b in bs)])
    tables = []
    for r in range(rounds):
        pick = [feas[c][(r + c) % len(feas[c])] for c in range(3)]
        tables.append(np.array([pick[cls[b]] for b in range(tb.BUCKETS)], np.uint8))
    tables.append(np.array([heuristic_slot(b) for b in range(tb.BUCKETS)], np.uint8))  # and the heuristic plan
    return tables


def samples_of_profile(p: tb.Profile, menu, min_lists: int = 1):
    """(X, y, work, bucket, slot) for every (bucket, slot) timed in profile p."""
    X, y, w, bk, sl = [], [], [], [], []
    tsum = [sum(float(p.ticks[s][b]) for b in range(tb.BUCKETS)) for s in range(tb.MENU_SLOTS)]
    for b in range(tb.BUCKETS):
        if p.lists[b] < min_lists:
            continue
        s = p.slot[b]
        ticks = p.ticks[s][b]
        if ticks == 0 or tsum[s] <= 0 or p.ms_slot[s] <= 0:
            continue
        work = float(p.sum_w[b]) + float(p.sum_k[b]) + 1.0
        ns = p.ms_slot[s] * 1e6 * ticks / tsum[s]  # the bucket's share of its launch
        X.append(tb.profile_features(p, b, s))
        y.append(np.log2(ns / work))
        w.append(work)
        bk.append(b)
        sl.append(s)
    return X, y, w, bk, sl


def collect(ctx, graphs, rounds: int, repeats: int, log):
    menu = tb.menu()
    tables = exploration_tables(menu, rounds)
    data = []
    for kind, p in graphs:
        t0 = time.time()
        with make_graph(ctx, kind, p) as g:
            g.count(sched="wedge")  # warm-up
            rows = {"X": [], "y": [], "w": [], "b": [], "s": []}
            for _ in range(repeats):
                for t in tables:
                    X, y, w, bk, sl = samples_of_profile(g.profile("wedge", t), menu)
                    rows["X"] += X
                    rows["y"] += y
                    rows["w"] += w
                    rows["b"] += bk
                    rows["s"] += sl
            info = g.info()
        name = f"{kind}:" + ",".join(f"{k}={v}" for k, v in p.items())
        log(f"{name}: m={info['m']} samples={len(rows['y'])} ({time.time() - t0:.1f}s)")
        data.append((name, {k: np.asarray(v) for k, v in rows.items()}))
    return data


def merge_repeats(d):
    """Mean target per (bucket, slot) over repeats (one sample per pair)."""
    key = d["b"].astype(np.int64) * 16 + d["s"].astype(np.int64)
    uk, inv = np.unique(key, return_inverse=True)
    y = np.bincount(inv, weights=d["y"]) / np.bincount(inv)
    first = np.zeros(uk.size, np.int64)
    first[inv[::-1]] = np.arange(inv.size)[::-1]
    return {"X": d["X"][first], "y": y, "w": d["w"][first], "b": d["b"][first], "s": d["s"][first]}


def train(ctx, X, y, epochs, seed):
    m = ctx.mlp(tb.FEATURES, seed=seed)
    loss = m.train(X.astype(np.float32), y.astype(np.float32), epochs=epochs, batch=128, lr=2e-3, seed=seed)
    return m, loss


def heuristic_slot(bucket: int) -> int:
    """The density-threshold heuristic's menu slot on the wedge layout (outdeg <= 32 / <= 256 / larger)."""
    top = tb.bucket_max_outdeg(bucket)
    return 0 if top <= 32 else 3 if top <= 256 else 5


def choice_quality(pred, d):
    """Measured cost (summed over buckets) of the model's pick and of the
    heuristic's slot, each relative to the best measured slot."""
    cost_pick = cost_best = cost_heur = 0.0
    for b in np.unique(d["b"]):
        sel = d["b"] == b
        meas = 2.0 ** d["y"][sel] * d["w"][sel]
        cost_pick += meas[np.argmin(pred[sel])]
        cost_best += meas.min()
        hs = np.flatnonzero(d["s"][sel] == heuristic_slot(int(b)))
        cost_heur += meas[hs[0]] if hs.size else meas.min()
    return cost_pick / cost_best, cost_heur / cost_best


def main() -> None:
    ap = argparse.ArgumentParser()
    ap.add_argument("--out", default=DEFAULT_OUT)
    ap.add_argument("--quick", action="store_true", help="four small graphs (a smoke run of the pipeline)")
    ap.add_argument("--rounds", type=int, default=8)
    ap.add_argument("--repeats", type=int, default=2)
    ap.add_argument("--epochs", type=int, default=300)
    ap.add_argument("--seed", type=int, default=7)
    a = ap.parse_args()

    def log(msg):
        print(f"[train] {msg}", flush=True)

    graphs = QUICK_GRAPHS if a.quick else TRAIN_GRAPHS
    with tb.Context() as ctx:
        data = [(n, merge_repeats(d)) for n, d in collect(ctx, graphs, a.rounds, a.repeats, log)]
        # held-out check: train without the last R-MAT graph, evaluate on it
        hold = max(i for i, (n, _) in enumerate(data) if n.startswith("rmat"))
        tr = [d for i, (_, d) in enumerate(data) if i != hold]
        Xtr = np.concatenate([d["X"] for d in tr])
        ytr = np.concatenate([d["y"] for d in tr])
        m, loss = train(ctx, Xtr, ytr, a.epochs, a.seed)
        dh = data[hold][1]
        ph = m.infer(dh["X"].astype(np.float32))
        r2 = 1.0 - np.sum((ph - dh["y"]) ** 2) / max(np.sum((dh["y"] - dh["y"].mean()) ** 2), 1e-12)
        ratio, hratio = choice_quality(ph, dh)
        log(f"held out {data[hold][0]}: samples={dh['y'].size} R2={r2:.3f} "
            f"cost/best measured: model {ratio:.4f}, heuristic {hratio:.4f}; train loss={loss:.4f}")
        m.free()
        # final model on every graph
        X = np.concatenate([d["X"] for _, d in data])
        y = np.concatenate([d["y"] for _, d in data])
        m, loss = train(ctx, X, y, a.epochs, a.seed)
        params = m.params()
        m.free()
    params.astype("<f4").tofile(a.out)
    meta = {"features": tb.FEATURES, "params": int(params.size), "samples": int(y.size), "graphs": [n for n, _ in data],
            "held_out": data[hold][0], "held_out_r2": float(r2), "held_out_choice_ratio": float(ratio),
            "held_out_heuristic_ratio": float(hratio),
            "final_loss": float(loss), "epochs": a.epochs, "seed": a.seed, "rounds": a.rounds,
            "repeats": a.repeats, "target": "log2(ns of launch time per (streamed element + range)), the bucket's share of its launch"}
    with open(os.path.splitext(a.out)[0] + ".json", "w") as f:
        json.dump(meta, f, indent=1)
    log(f"wrote {a.out} ({params.size} floats), final loss {loss:.4f}")


if __name__ == "__main__":
    main()
