"""BVGraph (WebGraph) reader, tribit_read_webgraph (tribit_amd/csrc/webgraph.cpp): the reference's
multi-GPU ingest (src/multi_gpu/src/main.rs:88, :150-161) turns every arc (u, v), u != v, of a
BVGraph into the key (min << 32) | max. No BVGraph file ships with the reference, so the reader
is pinned by (1) bit strings assembled by hand from the published code definitions, (2) files
written by the test encoder (tests/webgraph_enc.py, itself pinned to the code definitions) over
random graphs with every coding option, and (3) error cases."""
import numpy as np
import pytest

import tribit_amd
from tests import webgraph_enc as enc


def bits(s: str) -> str:
    return s.replace(" ", "")


@pytest.mark.parametrize("n,want", [(0, "1"), (1, "010"), (2, "011"), (3, "00100"), (6, "00111"), (7, "0001000")])
def test_gamma_definition(n, want):
    w = enc.BitWriter()
    w.gamma(n)
    assert "".join(map(str, w.bits)) == want


@pytest.mark.parametrize("n,want", [(0, "1"), (1, "0100"), (2, "0101"), (3, "01100"), (7, "00100000")])
def test_delta_definition(n, want):
    w = enc.BitWriter()
    w.delta(n)
    assert "".join(map(str, w.bits)) == want


@pytest.mark.parametrize("n,k,want", [(0, 3, "100"), (1, 3, "1010"), (2, 3, "1011"), (6, 3, "1111"),
                                      (7, 3, "0100000"), (0, 1, "1"), (1, 1, "010"), (3, 2, "01000")])
def test_zeta_definition(n, k, want):
    # zeta_k(n): unary(h), n + 1 in [2^(hk), 2^((h+1)k)); minimal binary of n + 1 - 2^(hk) over
    # 2^(hk)(2^k - 1) values (hk + k - 1 bits below 2^(hk), else n + 1 itself in hk + k bits)
    w = enc.BitWriter()
    w.zeta(n, k)
    assert "".join(map(str, w.bits)) == want


def write_raw(base, bitstring, props):
    s = bits(bitstring)
    s += "0" * (-len(s) % 8)
    with open(base + ".graph", "wb") as f:
        f.write(bytes(int(s[i:i + 8], 2) for i in range(0, len(s), 8)))
    with open(base + ".properties", "w") as f:
        f.write("".join(f"{k}={v}\n" for k, v in props.items()))


def keyset(pairs):
    return sorted({(min(a, b) << 32) | max(a, b) for a, b in pairs if a != b})


def test_hand_assembled_residuals(tmp_path):
    # 0 -> {1, 2}, 1 -> {2}, 2 -> {} ; no references (windowsize 0), no intervals; gamma outdegrees,
    # zeta_3 residuals: node 0: gamma(2) zeta3(int2nat(1 - 0) = 2) zeta3(2 - 1 - 1 = 0);
    # node 1: gamma(1) zeta3(int2nat(2 - 1) = 2); node 2: gamma(0)
    base = str(tmp_path / "g")
    write_raw(base, "011 1011 100  010 1011  1", {"nodes": 3, "arcs": 3, "windowsize": 0,
                                                 "minintervallength": 0, "zetak": 3})
    keys, n = tribit_amd.read_webgraph(base)
    assert n == 3
    assert sorted(keys.tolist()) == keyset([(0, 1), (0, 2), (1, 2)])


def test_hand_assembled_reference_and_negative_residual(tmp_path):
    # windowsize 1: 0 -> {1, 2}: gamma(2), unary(0) [no reference], residuals zeta3(2) zeta3(0);
    # 1 -> {0, 2} copying from node 0 (r = 1, unary "01"): blocks over [1, 2] = copy 0, skip 1,
    # (implicit copy): count gamma(2), gamma(0), gamma(1 - 1); residual 0 = x - 1: zeta3(int2nat(-1) = 1);
    # 2 -> {}: gamma(0)
    base = str(tmp_path / "g")
    write_raw(base, "011 1 1011 100   011 01 011 1 1 1010   1",
              {"nodes": 3, "arcs": 4, "windowsize": 1, "minintervallength": 0, "zetak": 3})
    keys, n = tribit_amd.read_webgraph(base)
    assert sorted(keys.tolist()) == sorted(keyset([(0, 1), (0, 2)]) + keyset([(1, 0), (1, 2)]))


def test_hand_assembled_interval(tmp_path):
    # minintervallength 2, windowsize 0: 0 -> {3, 4, 5, 9}: gamma(4); intervals: gamma(1),
    # left 3 = 0 + nat2int(gamma(6)), length gamma(3 - 2); residual 9: zeta3(int2nat(9) = 18)
    # zeta3(18): n + 1 = 19 in [8, 64): unary(1) then 19 - 8 = 11 < 8? no -> 19 in 6 bits
    base = str(tmp_path / "g")
    rest = "1" * 9  # nodes 1..9: gamma(0)
    write_raw(base, "00101 010 00111 010 01 010011" + rest,
              {"nodes": 10, "arcs": 4, "windowsize": 0, "minintervallength": 2, "zetak": 3})
    keys, _ = tribit_amd.read_webgraph(base)
    assert sorted(keys.tolist()) == keyset([(0, 3), (0, 4), (0, 5), (0, 9)])


def random_graph(rng, n, deg, local):
    succ = []
    for x in range(n):
        d = int(rng.integers(0, deg + 1))
        if local and x > 2 and rng.random() < 0.5:  # similar to a neighbour: references pay off
            base = succ[x - int(rng.integers(1, 3))]
            s = set(y for y in base if rng.random() < 0.8)
        else:
            s = set()
        s |= set(int(v) for v in rng.integers(0, n, d))
        if local and rng.random() < 0.3:  # a run of consecutive ids: intervals pay off
            a = int(rng.integers(0, max(1, n - 12)))
            s |= set(range(a, min(n, a + int(rng.integers(2, 12)))))
        succ.append(sorted(s))
    return succ


CASES = [
    dict(window=7, min_interval=4, zetak=3, codes=None),
    dict(window=0, min_interval=0, zetak=3, codes=None),
    dict(window=1, min_interval=2, zetak=1, codes=None),
    dict(window=3, min_interval=3, zetak=5, codes={"OUTDEGREES": "DELTA", "RESIDUALS": "GAMMA"}),
    dict(window=2, min_interval=4, zetak=2, codes={"REFERENCES": "GAMMA", "BLOCKS": "DELTA", "BLOCK_COUNT": "UNARY",
                                                   "INTERVALS": "DELTA", "RESIDUALS": "DELTA"}),
    dict(window=5, min_interval=1, zetak=4, codes={"OUTDEGREES": "UNARY", "REFERENCES": "DELTA",
                                                   "RESIDUALS": "ZETA", "BLOCK_COUNT": "DELTA"}),
]


@pytest.mark.parametrize("case", range(len(CASES)))
def test_encoded_round_trip(tmp_path, case):
    rng = np.random.default_rng(100 + case)
    p = CASES[case]
    for n, deg, local in ((1, 0, False), (40, 6, True), (600, 12, True), (3000, 3, False)):
        succ = random_graph(rng, n, deg, local)
        base = str(tmp_path / f"g{n}")
        enc.write(base, succ, window=p["window"], min_interval=p["min_interval"], zetak=p["zetak"], codes=p["codes"])
        keys, nn = tribit_amd.read_webgraph(base)
        assert nn == n
        want = sorted(((min(x, y) << 32) | max(x, y)) for x, s in enumerate(succ) for y in s if y != x)
        assert sorted(keys.tolist()) == want, (case, n)


def test_errors(tmp_path):
    base = str(tmp_path / "g")
    succ = random_graph(np.random.default_rng(5), 200, 8, True)
    enc.write(base, succ)
    full = open(base + ".graph", "rb").read()
    with open(base + ".graph", "wb") as f:  # truncated bit stream
        f.write(full[: len(full) // 2])
    with pytest.raises(tribit_amd.TribitError) as e:
        tribit_amd.read_webgraph(base)
    assert e.value.status == 2  # TRIBIT_ERR_FORMAT
    with pytest.raises(tribit_amd.TribitError) as e:
        tribit_amd.read_webgraph(str(tmp_path / "missing"))
    assert e.value.status == 1  # TRIBIT_ERR_IO
    enc.write(base, succ, flags_text="RESIDUALS_NIBBLE")  # a code this reader does not decode
    with pytest.raises(tribit_amd.TribitError) as e:
        tribit_amd.read_webgraph(base)
    assert e.value.status == 2


def test_malformed_values(tmp_path):
    # values that no well-formed file holds are format errors, not out-of-range ids or overflows
    base = str(tmp_path / "g")
    props = {"nodes": 3, "arcs": 2, "windowsize": 0, "minintervallength": 0, "zetak": 3,
             "compressionflags": "RESIDUALS_GAMMA"}
    w = enc.BitWriter()
    w.gamma(2)              # node 0: outdegree 2
    w.gamma(2)              # first residual: int2nat(1) -> id 1
    w.gamma((1 << 61) - 1)  # a gap far beyond the node count
    w.gamma(0)
    w.gamma(0)
    with open(base + ".graph", "wb") as f:
        f.write(w.to_bytes())
    with open(base + ".properties", "w") as f:
        f.write("".join(f"{k}={v}\n" for k, v in props.items()))
    with pytest.raises(tribit_amd.TribitError) as e:
        tribit_amd.read_webgraph(base)
    assert e.value.status == 2
    # a reference to a node before the first one (node 0, r = 1)
    props.update({"windowsize": 1, "compressionflags": ""})
    w = enc.BitWriter()
    w.gamma(1)
    w.unary(1)
    w.gamma(0)
    w.gamma(0)
    with open(base + ".graph", "wb") as f:
        f.write(w.to_bytes())
    with open(base + ".properties", "w") as f:
        f.write("".join(f"{k}={v}\n" for k, v in props.items()))
    with pytest.raises(tribit_amd.TribitError) as e:
        tribit_amd.read_webgraph(base)
    assert e.value.status == 2
    # an outdegree above the node count
    props.update({"windowsize": 0})
    w = enc.BitWriter()
    w.gamma(4)
    with open(base + ".graph", "wb") as f:
        f.write(w.to_bytes())
    with open(base + ".properties", "w") as f:
        f.write("".join(f"{k}={v}\n" for k, v in props.items()))
    with pytest.raises(tribit_amd.TribitError) as e:
        tribit_amd.read_webgraph(base)
    assert e.value.status == 2
