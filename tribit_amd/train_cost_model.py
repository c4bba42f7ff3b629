"""Trains the learned scheduler's cost model on the GPU from measured plans.

    python -m tribit_amd.train_cost_model [--out tribit_amd/cost_model.f32] [--quick]

The learned schedule runs one menu slot per owner-outdeg class (<= 32, <= 256,
larger), the density-threshold heuristic's unless the model predicts a faster
one. The model is trained on exactly that decision:

1. Builds a set of training graphs on the device (R-MAT at several scales,
   edge factors and seeds, Chung-Lu, 2-D random geometric, stochastic block
   model, Erdos-Renyi; some under both vertex relabellings; the R-MAT
   scale-25 member uses seed 1, not the benchmark's seed 5).
2. Per graph, times the heuristic plan H and, for every class c with lists and
   every menu slot s the whole class fits, the plan H with c moved to s: the
   device time of the counting launches, the median of `repeats` counts, the
   plans interleaved over `rounds` rounds and the median round kept.
3. One sample per (graph, class, slot): the library's class feature vector
   (tribit_class_features) and the target log2(T(H with c -> s) / T(H)),
   clipped at `--clip` (a slot several times slower is as useless as one far
   slower).
4. Trains the 3-layer perceptron with the library's HIP kernels (FP32, Adam):
   first without the held-out graphs' families' last members, reporting the
   fit, then on everything. Then checks the final model on held-out graphs
   (never trained on): the heuristic plan against the model's plan, both timed
   interleaved, and writes the parameters (float32, little endian) and a JSON
   summary with those gains next to them.
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

# (family, parameters, relabel mode)
TRAIN_GRAPHS = [
    ("rmat", dict(scale=18, ef=16, seed=11), 1),
    ("rmat", dict(scale=20, ef=16, seed=12), 1),
    ("rmat", dict(scale=22, ef=16, seed=13), 1),
    ("rmat", dict(scale=23, ef=16, seed=14), 1),
    ("rmat", dict(scale=24, ef=16, seed=15), 1),
    ("rmat", dict(scale=25, ef=16, seed=1), 1),
    ("rmat", dict(scale=21, ef=8, seed=16), 1),
    ("rmat", dict(scale=20, ef=32, seed=17), 1),
    ("rmat", dict(scale=19, ef=4, seed=18), 1),
    ("rmat", dict(scale=22, ef=16, seed=13), 2),
    ("chung_lu", dict(n=1 << 20, gamma=2.1, avg_deg=48.0, seed=21), 1),
    ("chung_lu", dict(n=1 << 19, gamma=2.5, avg_deg=16.0, seed=22), 1),
    ("chung_lu", dict(n=1 << 21, gamma=2.8, avg_deg=8.0, seed=26), 1),
    ("chung_lu", dict(n=1 << 18, gamma=2.0, avg_deg=64.0, seed=27), 1),
    ("chung_lu", dict(n=1 << 20, gamma=3.0, avg_deg=24.0, seed=28), 1),
    ("chung_lu", dict(n=1 << 19, gamma=2.3, avg_deg=32.0, seed=29), 2),
    ("rgg", dict(n=1 << 20, avg_deg=24.0, seed=23), 1),
    ("rgg", dict(n=1 << 19, avg_deg=48.0, seed=30), 1),
    ("rgg", dict(n=1 << 21, avg_deg=12.0, seed=31), 1),
    ("rgg", dict(n=1 << 20, avg_deg=24.0, seed=23), 2),
    ("sbm", dict(blocks=400, size=512, p_in=0.3, p_out=1e-4, seed=24), 1),
    ("sbm", dict(blocks=2000, size=256, p_in=0.2, p_out=5e-5, seed=32), 1),
    ("sbm", dict(blocks=200, size=1024, p_in=0.1, p_out=1e-4, seed=33), 1),
    ("sbm", dict(blocks=400, size=512, p_in=0.3, p_out=1e-4, seed=24), 2),
    ("er", dict(n=1 << 20, samples=1 << 24, seed=25), 1),
    ("er", dict(n=1 << 19, samples=1 << 21, seed=34), 1),
    ("er", dict(n=1 << 21, samples=1 << 25, seed=35), 1),
    ("er", dict(n=1 << 18, samples=1 << 23, seed=36), 1),
    ("er", dict(n=1 << 20, samples=1 << 23, seed=37), 2),
]
QUICK_GRAPHS = [TRAIN_GRAPHS[0], TRAIN_GRAPHS[11], TRAIN_GRAPHS[17], TRAIN_GRAPHS[25]]

# Graphs no training run sees (other seeds and sizes): the gain check.
HELD_OUT = [
    ("rmat", dict(scale=21, ef=16, seed=78), 1),
    ("chung_lu", dict(n=300000, gamma=2.3, avg_deg=32.0, seed=78), 1),
    ("er", dict(n=1 << 18, samples=1 << 22, seed=79), 1),
    ("er", dict(n=1 << 20, samples=1 << 24, seed=7), 1),
    ("rgg", dict(n=1 << 21, avg_deg=24.0, seed=7), 1),
    ("sbm", dict(blocks=1000, size=512, p_in=0.3, p_out=1e-4, seed=7), 1),
]


def make_graph(ctx, kind, p, relabel=1):
    ctx.set_relabel(relabel)
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


def graph_name(kind, p, relabel) -> str:
    return f"{kind}:" + ",".join(f"{k}={v}" for k, v in p.items()) + f",relabel={relabel}"


def time_tables(g, tables, rounds: int, repeats: int):
    """Median (over rounds) of each table's median-of-`repeats` device ms; tables run interleaved."""
    ms = [[] for _ in tables]
    ref = None
    for _ in range(rounds):
        for i, t in enumerate(tables):
            m, tri = g.count_plan("wedge", t, repeats)
            if ref is not None and tri != ref:
                raise RuntimeError("plans disagree on the triangle count")
            ref = tri
            ms[i].append(m)
    return [float(np.median(v)) for v in ms]


def collect_graph(g, rounds: int, repeats: int, clip: float):
    """Samples (X, y, cls, slot) of one graph: every class with lists, every slot the class fits."""
    prof = g.profile("wedge")
    lists = np.zeros(tb.CLASSES)
    for b in range(tb.BUCKETS):
        lists[tb.outdeg_class(b)] += prof.lists[b]
    H = tb.heuristic_slots("wedge")
    plans = [(None, None)]
    for c in range(tb.CLASSES):
        if lists[c] == 0:
            continue
        for s in tb.class_slots(c):
            if s != H[c]:
                plans.append((c, s))
    tables = [tb.class_table(H)]
    for c, s in plans[1:]:
        a = list(H)
        a[c] = s
        tables.append(tb.class_table(a))
    ms = time_tables(g, tables, rounds, repeats)
    X, y, cl, sl = [], [], [], []
    for c in range(tb.CLASSES):
        if lists[c] == 0:
            continue
        X.append(tb.class_features(prof, c, H[c]))
        y.append(0.0)
        cl.append(c)
        sl.append(H[c])
    for (c, s), m in zip(plans[1:], ms[1:]):
        X.append(tb.class_features(prof, c, s))
        y.append(min(float(np.log2(m / ms[0])), clip))
        cl.append(c)
        sl.append(s)
    return np.array(X, np.float32), np.array(y, np.float32), np.array(cl), np.array(sl), ms[0]


def collect(ctx, graphs, rounds, repeats, clip, log):
    data = []
    for kind, p, rl in graphs:
        t0 = time.time()
        with make_graph(ctx, kind, p, rl) as g:
            g.count(sched="wedge")  # warm-up: layouts and caches
            X, y, cl, sl, ms_h = collect_graph(g, rounds, repeats, clip)
        name = graph_name(kind, p, rl)
        best = {int(c): (int(sl[cl == c][np.argmin(y[cl == c])]), round(float(2.0 ** -y[cl == c].min()), 3))
                for c in np.unique(cl)}
        log(f"{name}: heuristic {ms_h:.3f} ms, best slot and gain per class {best} ({time.time() - t0:.1f}s)")
        data.append((name, dict(X=X, y=y, c=cl, s=sl)))
    return data


def train(ctx, X, y, epochs, seed):
    m = ctx.mlp(tb.FEATURES, seed=seed)
    loss = m.train(X.astype(np.float32), y.astype(np.float32), epochs=epochs, batch=64, lr=2e-3, seed=seed)
    return m, loss


def pick_quality(pred, d, switch):
    """Measured gain over the heuristic of the slot the model picks per class (the decision rule of the
    library: switch only when the predicted gain beats `switch` log2 units), and the best measured gain."""
    out = []
    for c in np.unique(d["c"]):
        sel = np.flatnonzero(d["c"] == c)
        ys, ps = d["y"][sel], pred[sel]
        h = sel[np.flatnonzero(d["y"][sel] == 0.0)[0]] if np.any(ys == 0.0) else None
        ph = pred[h] if h is not None else 0.0
        i = int(np.argmin(ps))
        chosen = ys[i] if ps[i] < ph - switch else 0.0
        out.append((float(2.0 ** -chosen), float(2.0 ** -ys.min())))
    return out


def held_out_gains(ctx, rounds, repeats, log):
    rows = []
    for kind, p, rl in HELD_OUT:
        with make_graph(ctx, kind, p, rl) as g:
            g.count(sched="wedge")
            table = g.learned_plan()
            ms_h, ms_m = time_tables(g, [None, table], rounds, repeats)
        slots = sorted({(tb.outdeg_class(b), int(table[b])) for b in range(tb.BUCKETS)})
        rows.append({"graph": graph_name(kind, p, rl), "ms_heuristic": ms_h, "ms_model": ms_m,
                     "gain": ms_h / ms_m, "class_slots": slots})
        log(f"held out {rows[-1]['graph']}: heuristic {ms_h:.3f} ms, model {ms_m:.3f} ms, "
            f"gain {ms_h / ms_m:.3f} (class, slot) {slots}")
    return rows


def main() -> None:
    ap = argparse.ArgumentParser()
    ap.add_argument("--out", default=DEFAULT_OUT)
    ap.add_argument("--quick", action="store_true", help="four small graphs (a smoke run of the pipeline)")
    ap.add_argument("--rounds", type=int, default=3)
    ap.add_argument("--repeats", type=int, default=3)
    ap.add_argument("--epochs", type=int, default=400)
    ap.add_argument("--seed", type=int, default=7)
    ap.add_argument("--clip", type=float, default=1.5, help="largest target (log2 slowdown)")
    ap.add_argument("--save-data", default=None, help="write the collected samples here (.npz)")
    ap.add_argument("--data", default=None, help="train from samples saved by --save-data (no collection)")
    a = ap.parse_args()

    def log(msg):
        print(f"[train] {msg}", flush=True)

    graphs = QUICK_GRAPHS if a.quick else TRAIN_GRAPHS
    switch = 0.04  # the library's switching margin (log2 units)
    with tb.Context() as ctx:
        if a.data:
            z = np.load(a.data)
            names = [str(n) for n in z["names"]]
            data = [(n, dict(X=z["X"][z["g"] == i], y=z["y"][z["g"] == i], c=z["c"][z["g"] == i],
                             s=z["s"][z["g"] == i])) for i, n in enumerate(names)]
        else:
            data = collect(ctx, graphs, a.rounds, a.repeats, a.clip, log)
        if a.save_data:
            np.savez(a.save_data, names=np.array([n for n, _ in data]),
                     X=np.concatenate([d["X"] for _, d in data]), y=np.concatenate([d["y"] for _, d in data]),
                     c=np.concatenate([d["c"] for _, d in data]), s=np.concatenate([d["s"] for _, d in data]),
                     g=np.concatenate([np.full(d["y"].size, i) for i, (_, d) in enumerate(data)]))
        # fit check: train without every fifth graph, score the decisions on those
        hold = set(range(2, len(data), 5))
        Xtr = np.concatenate([d["X"] for i, (_, d) in enumerate(data) if i not in hold])
        ytr = np.concatenate([d["y"] for i, (_, d) in enumerate(data) if i not in hold])
        m, loss = train(ctx, Xtr, ytr, a.epochs, a.seed)
        fit = []
        for i in sorted(hold):
            name, d = data[i]
            pr = m.infer(d["X"])
            q = pick_quality(pr, d, switch)
            fit.append({"graph": name, "picked_gain": [round(x, 4) for x, _ in q],
                        "best_gain": [round(x, 4) for _, x in q]})
            log(f"fit check {name}: per class gain picked {fit[-1]['picked_gain']}, best {fit[-1]['best_gain']}")
        m.free()
        X = np.concatenate([d["X"] for _, d in data])
        y = np.concatenate([d["y"] for _, d in data])
        m, loss = train(ctx, X, y, a.epochs, a.seed)
        params = m.params()
        ctx.set_cost_model(m)
        gains = held_out_gains(ctx, a.rounds, 5, log)
        ctx.set_cost_model(None)
        m.free()
    params.astype("<f4").tofile(a.out)
    geo = float(np.exp(np.mean(np.log([r["gain"] for r in gains]))))
    meta = {"features": tb.FEATURES, "params": int(params.size), "samples": int(y.size),
            "graphs": [n for n, _ in data], "final_loss": float(loss), "epochs": a.epochs, "seed": a.seed,
            "rounds": a.rounds, "repeats": a.repeats, "clip": a.clip, "switch_log2": switch,
            "target": "log2(device ms of the plan with one outdeg class moved to the slot / ms of the heuristic plan)",
            "fit_check": fit, "held_out": gains, "held_out_geomean_gain": geo,
            "held_out_min_gain": float(min(r["gain"] for r in gains))}
    with open(os.path.splitext(a.out)[0] + ".json", "w") as f:
        json.dump(meta, f, indent=1)
    log(f"wrote {a.out} ({params.size} floats), final loss {loss:.4f}, held-out geomean gain {geo:.4f}")


if __name__ == "__main__":
    main()
