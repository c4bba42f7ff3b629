"""Writes tests/golden/golden.json: triangle counts from networkx, an
implementation independent of the oracle, for seeded graphs of the oracle's
generators and for the reference-held fixture wiki-Vote.mtx.

    python tests/golden/make_golden.py

The counts pin the oracle's count; the key hashes pin the generators (the GPU
generators must reproduce the same deduplicated key sets).
"""
from __future__ import annotations

import hashlib
import json
import os
import sys

import networkx as nx
import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

import oracle  # noqa: E402


def nx_triangles(ukeys: np.ndarray, n: int) -> int:
    g = nx.Graph()
    g.add_nodes_from(range(n))
    a = (ukeys >> np.uint64(32)).astype(np.int64)
    b = (ukeys & np.uint64(0xFFFFFFFF)).astype(np.int64)
    g.add_edges_from(zip(a.tolist(), b.tolist()))
    return sum(nx.triangles(g).values()) // 3


def record(name: str, spec: dict, keys: np.ndarray, n: int) -> dict:
    u = oracle.dedup(keys)
    g = oracle.Graph(u, n)
    return {
        "name": name,
        "spec": spec,
        "n": n,
        "entries": int(keys.size),
        "m": int(u.size),
        "wplus": int(g.wplus()),
        "triangles": nx_triangles(u, n),
        "keys_sha256": hashlib.sha256(u.tobytes()).hexdigest(),
    }


def main() -> None:
    if not os.path.exists(oracle.LIB_PATH):
        oracle.build()
    out = []
    st, keys, n, ne = oracle.read_mtx(os.path.join(HERE, "wiki-Vote.mtx"))
    assert st == 0
    r = record("wiki-Vote", {"file": "wiki-Vote.mtx"}, keys, n)
    r["entries_file"] = ne
    out.append(r)
    # R-MAT scale 13, edge factor 12, (.57,.19,.19,.05), seed 1
    for (s, ef, seed) in [(13, 12, 1), (10, 8, 7), (12, 16, 3), (14, 4, 11)]:
        spec = {"gen": "rmat", "scale": s, "ef": ef, "a": 0.57, "b": 0.19, "c": 0.19, "seed": seed}
        out.append(record(f"rmat-s{s}-ef{ef}-seed{seed}", spec, oracle.gen_rmat(s, ef, seed=seed), 1 << s))
    for (nn, mm, seed) in [(5000, 40000, 2), (20000, 200000, 9), (300, 20000, 4)]:
        spec = {"gen": "er", "n": nn, "samples": mm, "seed": seed}
        out.append(record(f"er-n{nn}-m{mm}-seed{seed}", spec, oracle.gen_er(nn, mm, seed), nn))
    for (nn, gamma, avg, seed) in [(3000, 2.5, 12.0, 7), (20000, 2.1, 30.0, 11), (1024, 2.8, 8.0, 3)]:
        spec = {"gen": "chung_lu", "n": nn, "gamma": gamma, "avg_deg": avg, "seed": seed}
        out.append(record(f"chunglu-n{nn}-g{gamma}-d{avg}-seed{seed}", spec,
                          oracle.gen_chung_lu(nn, gamma, avg, seed), nn))
    for (nn, avg, seed) in [(20000, 12.0, 3), (5000, 40.0, 8)]:
        spec = {"gen": "rgg", "n": nn, "avg_deg": avg, "seed": seed}
        out.append(record(f"rgg-n{nn}-d{avg}-seed{seed}", spec, oracle.gen_rgg(nn, avg, seed), nn))
    for (nb, bs, pin, pout, seed) in [(20, 64, 0.3, 1e-3, 4), (8, 200, 0.5, 2e-3, 6)]:
        spec = {"gen": "sbm", "blocks": nb, "size": bs, "p_in": pin, "p_out": pout, "seed": seed}
        out.append(record(f"sbm-{nb}x{bs}-seed{seed}", spec, oracle.gen_sbm(nb, bs, pin, pout, seed), nb * bs))
    with open(os.path.join(HERE, "golden.json"), "w") as f:
        json.dump(out, f, indent=1)
    for r in out:
        print(r["name"], r["n"], r["m"], r["wplus"], r["triangles"])


if __name__ == "__main__":
    main()
