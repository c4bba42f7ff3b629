"""Graph families with closed-form triangle counts, as packed edge keys.

Keys use the reference's packing (min << 32) | max (tribit.cu:258-267).
"""
from __future__ import annotations

from math import comb

import numpy as np


def pack(u, v) -> np.ndarray:
    u = np.asarray(u, dtype=np.uint64)
    v = np.asarray(v, dtype=np.uint64)
    lo = np.minimum(u, v)
    hi = np.maximum(u, v)
    return (lo << np.uint64(32)) | hi


def complete(n: int, offset: int = 0):
    i, j = np.triu_indices(n, 1)
    return pack(i + offset, j + offset), comb(n, 3)


def cycle(n: int):
    i = np.arange(n)
    return pack(i, (i + 1) % n), (1 if n == 3 else 0)


def wheel(k: int):
    """Hub 0 joined to a k-cycle on 1..k: k triangles for k >= 4."""
    rim = np.arange(1, k + 1)
    keys = np.concatenate([pack(np.zeros(k, np.int64), rim), pack(rim, (rim % k) + 1)])
    return keys, (k if k >= 4 else comb(k + 1, 3))


def windmill(k: int):
    """Friendship graph F_k: k triangles sharing vertex 0."""
    a = 1 + 2 * np.arange(k)
    b = a + 1
    keys = np.concatenate([pack(np.zeros(k, np.int64), a), pack(np.zeros(k, np.int64), b), pack(a, b)])
    return keys, k


def complete_multipartite(parts):
    """K_{a,b,c,...}: triangles = sum over part triples of products."""
    off = np.cumsum([0] + list(parts))
    us, vs = [], []
    for i in range(len(parts)):
        for j in range(i + 1, len(parts)):
            a = np.arange(off[i], off[i + 1])
            b = np.arange(off[j], off[j + 1])
            aa, bb = np.meshgrid(a, b, indexing="ij")
            us.append(aa.ravel())
            vs.append(bb.ravel())
    keys = pack(np.concatenate(us), np.concatenate(vs))
    t = 0
    p = list(parts)
    for i in range(len(p)):
        for j in range(i + 1, len(p)):
            for k in range(j + 1, len(p)):
                t += p[i] * p[j] * p[k]
    return keys, t, int(off[-1])


def random_bipartite(a: int, b: int, p: float, seed: int):
    rng = np.random.default_rng(seed)
    mask = rng.random((a, b)) < p
    i, j = np.nonzero(mask)
    return pack(i, j + a), 0, a + b


def random_graph(n: int, p: float, seed: int):
    rng = np.random.default_rng(seed)
    i, j = np.triu_indices(n, 1)
    keep = rng.random(i.size) < p
    return pack(i[keep], j[keep])


def permute(keys: np.ndarray, n: int, seed: int) -> np.ndarray:
    """Relabels vertices by a random permutation (count is invariant)."""
    perm = np.random.default_rng(seed).permutation(n).astype(np.uint64)
    k = keys[keys != np.uint64(0xFFFFFFFFFFFFFFFF)]
    a = perm[(k >> np.uint64(32)).astype(np.int64)]
    b = perm[(k & np.uint64(0xFFFFFFFF)).astype(np.int64)]
    return pack(a, b)
