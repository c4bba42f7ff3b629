"""Writes tests/golden/configs.json: the CPU oracle's exact results for every
benchmark configuration at its full size, so the GPU parity tests can compare
the HIP path with them without running the oracle on the GPU host.

    python tests/golden/make_configs.py [--only NAME ...]

Per single graph: n, input keys, m (undirected simple edges), W+ (sum over
degree-oriented arcs u->v of outdeg(v), ties by lower id), triangles and the
sha256 of the deduplicated key set (the GPU generator must produce the same
set); the edges joining ids less than 512 apart, the id order the library's
default (TRIBIT_OPT_RELABEL 3) gives the graph by its rule (octave classes, 2,
when those edges fill at least 1/16 of such id pairs and n >= 4096, else
degree rank, 1) and W+ under those ids (octave-class ids orient {a < b} as
a -> b). `--ids-only` adds the last three to existing records without counting. For the batch of 4,096 Chung-Lu graphs (seed 3): every graph's count,
vertices and edges, the sha256 of the uint64 count array and the total.

Generators, dedup, orientation and counting are the oracle's (oracle/oracle.c);
its count is pinned by closed forms, brute force and networkx counts
(tests/test_oracle.py, tests/golden/golden.json).
"""
from __future__ import annotations

import argparse
import hashlib
import json
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

import oracle  # noqa: E402

OUT = os.path.join(HERE, "configs.json")

# name -> (description, generator spec); the specs are those of bench.py's workloads
SINGLE = {
    "c1_rmat13": ("R-MAT scale 13, edge factor 12, (.57,.19,.19,.05), seed 1",
                  {"gen": "rmat", "scale": 13, "ef": 12, "seed": 1}),
    "c2_rmat22": ("R-MAT scale 22, edge factor 16, (.57,.19,.19,.05), seed 2",
                  {"gen": "rmat", "scale": 22, "ef": 16, "seed": 2}),
    "c4_rmat20": ("R-MAT scale 20, edge factor 16, (.57,.19,.19,.05), seed 7",
                  {"gen": "rmat", "scale": 20, "ef": 16, "seed": 7}),
    "c4_er": ("Erdos-Renyi n=2^20, 2^24 samples (average degree 32), seed 7",
              {"gen": "er", "n": 1 << 20, "samples": 1 << 24, "seed": 7}),
    "c4_rgg": ("2-D random geometric n=2^21, average degree 24, seed 7",
               {"gen": "rgg", "n": 1 << 21, "avg_deg": 24.0, "seed": 7}),
    "c4_sbm": ("SBM 1000 blocks x 512, p_in 0.3, p_out 1e-4, seed 7",
               {"gen": "sbm", "blocks": 1000, "size": 512, "p_in": 0.3, "p_out": 1e-4, "seed": 7}),
    "c5_rmat25": ("R-MAT scale 25, edge factor 16, (.57,.19,.19,.05), seed 5",
                  {"gen": "rmat", "scale": 25, "ef": 16, "seed": 5}),
}
BATCH = {"c3_batch": ("4,096 Chung-Lu graphs, batch seed 3", {"seed": 3, "graphs": 4096})}


def keys_of(spec: dict):
    g = spec["gen"]
    if g == "rmat":
        return oracle.gen_rmat(spec["scale"], spec["ef"], 0.57, 0.19, 0.19, spec["seed"]), 1 << spec["scale"]
    if g == "er":
        return oracle.gen_er(spec["n"], spec["samples"], spec["seed"]), spec["n"]
    if g == "rgg":
        return oracle.gen_rgg(spec["n"], spec["avg_deg"], spec["seed"]), spec["n"]
    if g == "sbm":
        return (oracle.gen_sbm(spec["blocks"], spec["size"], spec["p_in"], spec["p_out"], spec["seed"]),
                spec["blocks"] * spec["size"])
    raise ValueError(g)


def id_fields(u: np.ndarray, n: int, wplus: int) -> dict:
    """Band edges, the default id order's choice and W+ under it."""
    gap = (u & np.uint64(0xFFFFFFFF)) - (u >> np.uint64(32))
    band = int(np.count_nonzero(gap < np.uint64(512)))
    w = min(512, n)
    ids = 2 if n >= 4096 and band * 16 >= (w - 1) * n - w * (w - 1) / 2 else 1
    if ids == 2:
        r, _ = oracle.relabel(u, n, 2)
        wplus = int(oracle.Graph(r, n, by_id=True).wplus())
    return {"band_edges": band, "ids": ids, "wplus_ids": wplus}


def single(name: str, desc: str, spec: dict) -> dict:
    t0 = time.perf_counter()
    keys, n = keys_of(spec)
    entries = int(keys.size)
    u = oracle.dedup(keys)
    del keys
    sha = hashlib.sha256(u.tobytes()).hexdigest()
    g = oracle.Graph(u, n)
    m = int(g.m)
    wplus = int(g.wplus())
    ids = id_fields(u, n, wplus)
    del u
    tri = int(g.count())
    rec = {"name": name, "workload": desc, "spec": spec, "n": n, "entries": entries, "m": m, "wplus": wplus,
           "triangles": tri, "keys_sha256": sha, **ids}
    print(f"{name}: n={n} m={m} W+={wplus} triangles={tri} ({time.perf_counter() - t0:.1f} s)", flush=True)
    return rec


def batch(name: str, desc: str, spec: dict) -> dict:
    t0 = time.perf_counter()
    counts = np.zeros(spec["graphs"], np.uint64)
    ns = np.zeros(spec["graphs"], np.uint64)
    ms = np.zeros(spec["graphs"], np.uint64)
    for i in range(spec["graphs"]):
        gamma, n, avg = oracle.chung_lu_batch_params(spec["seed"], i)
        keys = oracle.gen_chung_lu(n, gamma, avg, (spec["seed"] << 32) + i)
        g = oracle.Graph.from_keys(keys, n)
        counts[i] = g.count()
        ns[i] = n
        ms[i] = g.m
        if (i + 1) % 512 == 0:
            print(f"{name}: {i + 1}/{spec['graphs']} graphs ({time.perf_counter() - t0:.0f} s)", flush=True)
    rec = {"name": name, "workload": desc, "spec": spec, "graphs": spec["graphs"],
           "total_n": int(ns.sum()), "total_m": int(ms.sum()), "triangles": int(counts.sum()),
           "counts_sha256": hashlib.sha256(counts.tobytes()).hexdigest(),
           "counts": [int(c) for c in counts], "n": [int(x) for x in ns], "m": [int(x) for x in ms]}
    print(f"{name}: {spec['graphs']} graphs, m={rec['total_m']} triangles={rec['triangles']} "
          f"({time.perf_counter() - t0:.1f} s)", flush=True)
    return rec


def main() -> None:
    ap = argparse.ArgumentParser()
    ap.add_argument("--only", nargs="*", help="configuration names to (re)compute; others are kept")
    ap.add_argument("--ids-only", action="store_true", help="only add the id-order fields to single graphs")
    args = ap.parse_args()
    if not os.path.exists(oracle.LIB_PATH):
        oracle.build()
    out = {}
    if os.path.exists(OUT):
        with open(OUT) as f:
            out = json.load(f)
    todo = set(args.only) if args.only else set(SINGLE) | set(BATCH)
    for name, (desc, spec) in SINGLE.items():
        if name in todo and args.ids_only:
            keys, n = keys_of(spec)
            u = oracle.dedup(keys)
            del keys
            assert hashlib.sha256(u.tobytes()).hexdigest() == out[name]["keys_sha256"], name
            out[name].update(id_fields(u, n, out[name]["wplus"]))
            print(f"{name}: {out[name]['band_edges']} band edges, ids {out[name]['ids']}, "
                  f"W+ {out[name]['wplus_ids']}", flush=True)
        elif name in todo:
            out[name] = single(name, desc, spec)
    for name, (desc, spec) in BATCH.items():
        if name in todo and not args.ids_only:
            out[name] = batch(name, desc, spec)
    with open(OUT, "w") as f:
        json.dump(out, f, indent=1)
        f.write("\n")


if __name__ == "__main__":
    main()
