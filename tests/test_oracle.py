"""Pins the CPU oracle (oracle/) before it is trusted as the checker.

Checks, each chosen so that a plausible slip fails one of them:
  - the reference-held fixture wiki-Vote.mtx: SNAP's 608,389 triangles, its
    shape, and the reference pipeline's own metadata (tribit.cu:656-786)
    derived independently when the reference was surveyed;
  - closed forms (K_n, wheels, windmills, cycles, multipartite, bipartite);
  - O(n^3) brute force on tiny random graphs;
  - networkx golden counts (tests/golden/golden.json);
  - invariances (relabelling, duplicates, self-loops, reversed direction);
  - structural invariants of the orientation and of the bit tiles.
"""
import json
import os
from math import comb

import numpy as np
import pytest

from tests import gens, graphs
from tests.conftest import GOLDEN

WIKI = os.path.join(GOLDEN, "wiki-Vote.mtx")


def golden():
    with open(os.path.join(GOLDEN, "golden.json")) as f:
        return json.load(f)


def test_wiki_vote_known_answer(oracle_mod):
    st, keys, n, ne = oracle_mod.read_mtx(WIKI)
    assert st == 0
    assert (n, ne) == (8297, 103689)
    g = oracle_mod.Graph.from_keys(keys, n)
    assert g.m == 100762
    assert g.count() == 608389


def test_wiki_vote_reference_metadata(oracle_mod):
    """Restatement of tribit.cu:656-786 reproduces the surveyed wiki-Vote figures."""
    st, keys, n, _ = oracle_mod.read_mtx(WIKI)
    r = oracle_mod.ref_pipeline(oracle_mod.dedup(keys), n)
    assert r == {"m": 100762, "w_l": 70956, "m_purged": 99318, "w_u": 54880, "r": 2964, "t": 113795,
                 "blocks": 28449, "triangles": 608389}


@pytest.mark.parametrize("n", [3, 4, 5, 10, 31, 64, 100])
def test_complete_graph(oracle_mod, n):
    keys, t = graphs.complete(n)
    assert oracle_mod.count_keys(keys, n) == t == comb(n, 3)
    assert oracle_mod.ref_pipeline(oracle_mod.dedup(keys), n)["triangles"] == t


@pytest.mark.parametrize("n", [3, 4, 5, 10, 64, 300])
def test_wplus_complete_graph(oracle_mod, n):
    """W+ closed form: every vertex of K_n has degree n-1, so ties go to the lower id, vertex j has
    out-degree n-1-j and j in-arcs, and W+ = sum_j j (n-1-j) = C(n,3)."""
    keys, _ = graphs.complete(n)
    assert oracle_mod.Graph.from_keys(keys, n).wplus() == comb(n, 3)


@pytest.mark.parametrize("k", [1, 5, 40])
def test_wplus_stars_and_windmills(oracle_mod, k):
    """Closed forms: a star K_{1,k} orients every leaf to the hub, whose out-degree is 0: W+ = 0. In the
    windmill F_k each blade pair (a<b, degree 2) has a->b, a->hub, b->hub and the hub (degree 2k) has
    out-degree 0 for k>1: W+ = sum over arcs a->b of outdeg(b) = k * 1."""
    star = graphs.pack(np.zeros(k, np.int64), np.arange(1, k + 1))
    assert oracle_mod.Graph.from_keys(star, k + 1).wplus() == 0
    if k > 1:
        keys, _ = graphs.windmill(k)
        assert oracle_mod.Graph.from_keys(keys, 2 * k + 1).wplus() == k


def numpy_wplus(ukeys: np.ndarray, n: int) -> int:
    """W+ computed independently of the oracle: degrees by bincount, each edge oriented from the lower
    (degree, id) rank to the higher, out-degrees by bincount over arc tails, summed over arc heads."""
    a = (ukeys >> np.uint64(32)).astype(np.int64)
    b = (ukeys & np.uint64(0xFFFFFFFF)).astype(np.int64)
    deg = np.bincount(np.concatenate([a, b]), minlength=n)
    a_first = (deg[a] < deg[b]) | ((deg[a] == deg[b]) & (a < b))
    tail = np.where(a_first, a, b)
    head = np.where(a_first, b, a)
    outdeg = np.bincount(tail, minlength=n)
    return int(outdeg[head].sum())


@pytest.mark.parametrize("seed", range(6))
def test_wplus_numpy_recompute(oracle_mod, seed):
    """or_wplus against the numpy recomputation on R-MAT, Chung-Lu and Erdos-Renyi graphs."""
    if seed % 3 == 0:
        keys, n = oracle_mod.gen_rmat(12, 8, seed=seed + 1), 1 << 12
    elif seed % 3 == 1:
        keys, n = oracle_mod.gen_chung_lu(5000, 2.3, 12.0, seed), 5000
    else:
        keys, n = oracle_mod.gen_er(3000, 20000, seed), 3000
    u = oracle_mod.dedup(keys)
    assert oracle_mod.Graph(u, n).wplus() == numpy_wplus(u, n)


def test_wplus_numpy_wiki_vote(oracle_mod):
    st, keys, n, _ = oracle_mod.read_mtx(WIKI)
    u = oracle_mod.dedup(keys)
    assert oracle_mod.Graph(u, n).wplus() == numpy_wplus(u, n)


@pytest.mark.parametrize("k", [3, 4, 5, 17, 100])
def test_wheel(oracle_mod, k):
    keys, t = graphs.wheel(k)
    assert oracle_mod.count_keys(keys, k + 1) == t


@pytest.mark.parametrize("k", [1, 2, 7, 50])
def test_windmill(oracle_mod, k):
    keys, t = graphs.windmill(k)
    assert oracle_mod.count_keys(keys, 2 * k + 1) == t


@pytest.mark.parametrize("n", [3, 4, 5, 1000])
def test_cycle(oracle_mod, n):
    keys, t = graphs.cycle(n)
    assert oracle_mod.count_keys(keys, n) == t


@pytest.mark.parametrize("parts", [(1, 1, 1), (2, 3, 4), (5, 6, 7), (3, 3, 3, 3), (10, 1, 20)])
def test_complete_multipartite(oracle_mod, parts):
    keys, t, n = graphs.complete_multipartite(parts)
    assert oracle_mod.count_keys(keys, n) == t


@pytest.mark.parametrize("seed", range(4))
def test_bipartite_has_no_triangles(oracle_mod, seed):
    keys, t, n = graphs.random_bipartite(40, 60, 0.3, seed)
    assert oracle_mod.count_keys(keys, n) == 0


@pytest.mark.parametrize("seed", range(20))
def test_bruteforce_tiny(oracle_mod, seed):
    rng = np.random.default_rng(seed)
    n = int(rng.integers(2, 60))
    keys = graphs.random_graph(n, float(rng.uniform(0.05, 0.9)), seed)
    bf = oracle_mod.count_bruteforce(n, keys)
    assert oracle_mod.count_keys(keys, n) == bf
    assert oracle_mod.ref_pipeline(oracle_mod.dedup(keys), n)["triangles"] == bf


def test_golden_networkx(oracle_mod):
    for r in golden():
        g = oracle_mod.Graph.from_keys(gens.oracle_keys(oracle_mod, r), r["n"])
        assert (g.m, g.wplus(), g.count()) == (r["m"], r["wplus"], r["triangles"]), r["name"]


def test_golden_generator_hashes(oracle_mod):
    import hashlib

    for r in golden():
        if "gen" not in r["spec"]:
            continue
        keys = gens.oracle_keys(oracle_mod, r)
        assert hashlib.sha256(oracle_mod.dedup(keys).tobytes()).hexdigest() == r["keys_sha256"], r["name"]


def test_chung_lu_generator(oracle_mod):
    """CDF is monotone, ends at 2^62, and hub 0 is sampled most often; batch parameters in range."""
    cdf = oracle_mod.chung_lu_cdf(4096, 2.3)
    assert np.all(np.diff(cdf.astype(np.float64)) >= 0) and int(cdf[-1]) == 1 << 62
    keys = oracle_mod.gen_chung_lu(4096, 2.3, 20.0, 1)
    ok = keys[keys != np.uint64(0xFFFFFFFFFFFFFFFF)]
    ends = np.concatenate([(ok >> np.uint64(32)), ok & np.uint64(0xFFFFFFFF)]).astype(np.int64)
    deg = np.bincount(ends, minlength=4096)
    assert deg[0] == deg.max() and deg[:10].mean() > 5 * deg.mean()
    for i in range(64):
        g, n, d = oracle_mod.chung_lu_batch_params(3, i)
        assert 2.1 <= g < 2.8 and 1024 <= n <= 262144 and 8.0 <= d < 64.0


def test_rgg_generator(oracle_mod):
    """Average degree near the target; ids are spatially ordered (most edges join close ids)."""
    n, avg = 20000, 12.0
    keys = oracle_mod.gen_rgg(n, avg, 3)
    m = oracle_mod.dedup(keys).size
    assert m == keys.size  # every pair emitted once
    assert abs(2 * m / n - avg) / avg < 0.1  # boundary effects lower the degree slightly
    lo = (keys >> np.uint64(32)).astype(np.int64)
    hi = (keys & np.uint64(0xFFFFFFFF)).astype(np.int64)
    assert np.median(hi - lo) < n / 50


def test_sbm_generator(oracle_mod):
    nb, bs = 20, 64
    keys = oracle_mod.dedup(oracle_mod.gen_sbm(nb, bs, 0.3, 1e-3, 4))
    lo = (keys >> np.uint64(32)).astype(np.int64)
    hi = (keys & np.uint64(0xFFFFFFFF)).astype(np.int64)
    inner = (lo // bs) == (hi // bs)
    pairs_in = nb * bs * (bs - 1) // 2
    assert abs(inner.sum() / pairs_in - 0.3) < 0.02
    assert 0 < (~inner).sum() <= round(1e-3 * ((nb * bs) * (nb * bs - 1) / 2 - pairs_in)) + 1


def test_invariances(oracle_mod):
    keys = oracle_mod.gen_rmat(11, 8, seed=5)
    n = 1 << 11
    base = oracle_mod.count_keys(keys, n)
    assert base > 0
    # relabelling
    assert oracle_mod.count_keys(graphs.permute(keys, n, 1), n) == base
    # duplicated entries and shuffled order
    rng = np.random.default_rng(0)
    dup = np.concatenate([keys, keys[rng.permutation(keys.size)[: keys.size // 2]]])
    assert oracle_mod.count_keys(dup[rng.permutation(dup.size)], n) == base
    # sentinels (dropped self-loops) anywhere
    sent = np.concatenate([keys, np.full(100, 0xFFFFFFFFFFFFFFFF, np.uint64)])
    assert oracle_mod.count_keys(sent[rng.permutation(sent.size)], n) == base


def test_orientation_invariants(oracle_mod):
    keys = oracle_mod.gen_rmat(12, 8, seed=2)
    n = 1 << 12
    u = oracle_mod.dedup(keys)
    g = oracle_mod.Graph(u, n)
    rp, col, deg = g.arrays()
    assert rp[0] == 0 and rp[-1] == u.size
    src = np.repeat(np.arange(n, dtype=np.uint64), np.diff(rp).astype(np.int64))
    dst = col.astype(np.uint64)
    # rank order (deg, id) increases along every arc
    assert np.all((deg[src] < deg[dst]) | ((deg[src] == deg[dst]) & (src < dst)))
    # every undirected edge exactly once
    back = graphs.pack(src, dst)
    assert np.array_equal(np.sort(back), u)
    # rows sorted ascending
    for r in range(0, n, 97):
        row = col[rp[r]:rp[r + 1]]
        assert np.all(np.diff(row.astype(np.int64)) > 0)
    # degrees are the undirected degrees
    a = (u >> np.uint64(32)).astype(np.int64)
    b = (u & np.uint64(0xFFFFFFFF)).astype(np.int64)
    assert np.array_equal(deg, np.bincount(np.concatenate([a, b]), minlength=n).astype(np.uint32))


@pytest.mark.parametrize("seed", [1, 6])
def test_degree_labels(oracle_mod, seed):
    """Counting-sort labels equal a numpy lexsort of (degree, id); relabelled
    orientation is the id order and is isomorphic to the degree orientation."""
    keys = oracle_mod.gen_rmat(12, 8, seed=seed)
    n = 1 << 12
    u = oracle_mod.dedup(keys)
    a = (u >> np.uint64(32)).astype(np.int64)
    b = (u & np.uint64(0xFFFFFFFF)).astype(np.int64)
    deg = np.bincount(np.concatenate([a, b]), minlength=n)
    old_of_new, new_of_old = oracle_mod.degree_labels(u, n)
    assert np.array_equal(old_of_new, np.lexsort((np.arange(n), deg)).astype(np.uint32))
    assert np.array_equal(new_of_old[old_of_new], np.arange(n, dtype=np.uint32))
    g0 = oracle_mod.Graph(u, n)
    g1 = oracle_mod.Graph.from_keys(keys, n, relabel=True)
    assert np.array_equal(g1.labels, old_of_new)
    rp, col, d1 = g1.arrays()
    src = np.repeat(np.arange(n, dtype=np.int64), np.diff(rp).astype(np.int64))
    assert np.all(src < col.astype(np.int64))
    assert np.all(np.diff(d1.astype(np.int64)) >= 0)
    assert np.array_equal(d1, deg[old_of_new].astype(np.uint32))
    assert (g1.m, g1.wplus(), g1.count()) == (g0.m, g0.wplus(), g0.count())
    # mapping the relabelled arcs back gives the degree orientation's arcs
    rp0, col0, _ = g0.arrays()
    src0 = np.repeat(np.arange(n, dtype=np.uint64), np.diff(rp0).astype(np.int64))
    arcs = (old_of_new[src].astype(np.uint64) << np.uint64(32)) | old_of_new[col].astype(np.uint64)
    arcs0 = (src0 << np.uint64(32)) | col0.astype(np.uint64)
    assert np.array_equal(np.sort(arcs), np.sort(arcs0))


@pytest.mark.parametrize("seed", [2, 7])
def test_octave_labels(oracle_mod, seed):
    """Octave-class labels equal a Python sort of (int.bit_length(degree), id); the id-oriented CSR
    under them is a -> b iff a < b, with the same m and count, and its arcs map back to the input
    edges."""
    keys = np.concatenate([oracle_mod.gen_sbm(6, 200, 0.3, 0.01, seed), oracle_mod.gen_rmat(10, 4, seed=seed)])
    n = 1200
    keys = keys[(keys >> np.uint64(32)) < np.uint64(n)]
    keys = keys[(keys & np.uint64(0xFFFFFFFF)) < np.uint64(n)]
    u = oracle_mod.dedup(keys)
    a = (u >> np.uint64(32)).astype(np.int64)
    b = (u & np.uint64(0xFFFFFFFF)).astype(np.int64)
    deg = np.bincount(np.concatenate([a, b]), minlength=n)
    want = sorted(range(n), key=lambda v: (int(deg[v]).bit_length(), v))
    old_of_new, new_of_old = oracle_mod.octave_labels(u, n)
    assert old_of_new.tolist() == want
    assert np.array_equal(new_of_old[old_of_new], np.arange(n, dtype=np.uint32))
    g0 = oracle_mod.Graph(u, n)
    g2 = oracle_mod.Graph.from_keys(keys, n, relabel=2)
    assert np.array_equal(g2.labels, old_of_new)
    rp, col, d2 = g2.arrays()
    src = np.repeat(np.arange(n, dtype=np.int64), np.diff(rp).astype(np.int64))
    assert np.all(src < col.astype(np.int64))
    assert np.array_equal(d2, deg[old_of_new].astype(np.uint32))
    assert (g2.m, g2.count()) == (g0.m, g0.count()) == (g0.m, oracle_mod.count_bruteforce(n, u))
    lo = np.minimum(old_of_new[src], old_of_new[col]).astype(np.uint64)
    hi = np.maximum(old_of_new[src], old_of_new[col]).astype(np.uint64)
    assert np.array_equal(np.sort((lo << np.uint64(32)) | hi), u)


@pytest.mark.parametrize("width", [32, 64])
def test_tiles_decode(oracle_mod, width):
    keys = oracle_mod.gen_rmat(11, 12, seed=3)
    n = 1 << 11
    g = oracle_mod.Graph.from_keys(keys, n)
    rp, col, _ = g.arrays()
    tp, blk, mask = g.tiles(width)
    pc = np.array([bin(int(x)).count("1") for x in mask])
    assert pc.sum() == g.m
    for u in range(0, n, 13):
        dec = []
        for k in range(int(tp[u]), int(tp[u + 1])):
            for bit in range(width):
                if (int(mask[k]) >> bit) & 1:
                    dec.append(int(blk[k]) * width + bit)
        assert dec == col[rp[u]:rp[u + 1]].tolist()


def test_rmat_generator_distribution(oracle_mod):
    """Top-level quadrant frequencies follow (a, b, c, d) within 4 sigma."""
    s, ef = 14, 8
    keys = oracle_mod.gen_rmat(s, ef, seed=9)
    assert keys.size == ef << s
    ok = keys != np.uint64(0xFFFFFFFFFFFFFFFF)
    lo = (keys[ok] >> np.uint64(32)).astype(np.int64)
    hi = (keys[ok] & np.uint64(0xFFFFFFFF)).astype(np.int64)
    half = 1 << (s - 1)
    # after min/max packing the quadrants (0,1) and (1,0) merge; (0,0): both < half
    p00 = np.mean((lo < half) & (hi < half))
    p11 = np.mean((lo >= half) & (hi >= half))
    N = ok.sum()
    # self-loops (dropped) only remove same-quadrant pairs; compare loosely
    assert abs(p00 - 0.57 / (1 - 0.0)) < 0.02 + 4 * np.sqrt(0.25 / N)
    assert abs(p11 - 0.05) < 0.01 + 4 * np.sqrt(0.25 / N)
    # determinism
    assert np.array_equal(keys, oracle_mod.gen_rmat(s, ef, seed=9))
    assert not np.array_equal(keys, oracle_mod.gen_rmat(s, ef, seed=10))


def test_mix64_known_values(oracle_mod):
    """SplitMix64 reference outputs (seed 0 stream: first outputs of splitmix64 with state 0)."""
    # splitmix64 with state s returns mix64(s) where mix64 adds the golden gamma first.
    assert oracle_mod.mix64(0) == 0xE220A8397B1DCDAF
    assert oracle_mod.mix64(0x9E3779B97F4A7C15) == 0x6E789E6AA1B965F4
