"""BVGraph (WebGraph) encoder for tests: writes <base>.properties and <base>.graph
for given successor lists, from the encoding side of the published format
(Boldi & Vigna, "The WebGraph Framework I: Compression Techniques", WWW 2004;
"Codes for the World-Wide Web", Internet Math. 2005). Test infrastructure
only: the library's reader (tribit_amd/csrc/webgraph.cpp) is checked against
files this writes and against hand-assembled bit strings.
"""
from __future__ import annotations


class BitWriter:
    """Most significant bit first."""

    def __init__(self):
        self.bits = []

    def write(self, value: int, n: int) -> None:
        for i in range(n - 1, -1, -1):
            self.bits.append((value >> i) & 1)

    def unary(self, n: int) -> None:
        self.bits.extend([0] * n)
        self.bits.append(1)

    def gamma(self, n: int) -> None:
        x = n + 1
        b = x.bit_length() - 1
        self.unary(b)
        self.write(x, b)  # the b low bits of n + 1

    def delta(self, n: int) -> None:
        x = n + 1
        b = x.bit_length() - 1
        self.gamma(b)
        self.write(x, b)

    def zeta(self, n: int, k: int) -> None:
        x = n + 1
        h = (x.bit_length() - 1) // k  # x in [2^(hk), 2^((h+1)k))
        left = 1 << (h * k)
        s = h * k + k  # minimal binary over 2^(hk) (2^k - 1) values: s - 1 or s bits
        self.unary(h)
        if x - left < left:
            self.write(x - left, s - 1)
        else:
            self.write(x, s)

    def code(self, kind: str, n: int, k: int) -> None:
        {"GAMMA": self.gamma, "DELTA": self.delta, "UNARY": self.unary}.get(kind, lambda v: self.zeta(v, k))(n)

    def to_bytes(self) -> bytes:
        bits = self.bits + [0] * (-len(self.bits) % 8)
        out = bytearray()
        for i in range(0, len(bits), 8):
            v = 0
            for b in bits[i:i + 8]:
                v = (v << 1) | b
            out.append(v)
        return bytes(out)


def int2nat(v: int) -> int:
    return 2 * v if v >= 0 else -2 * v - 1


def encode(succ, window=7, min_interval=4, zetak=3, codes=None, choose_ref=None):
    """Bit stream of the successor lists succ[x] (sorted, distinct ids < len(succ)).
    codes: field -> "GAMMA" | "DELTA" | "UNARY" | "ZETA" for OUTDEGREES, REFERENCES,
    BLOCK_COUNT, BLOCKS, INTERVALS, RESIDUALS. choose_ref(x, candidates) picks the
    reference (default: the candidate with the largest overlap, 0 = none)."""
    c = {"OUTDEGREES": "GAMMA", "REFERENCES": "UNARY", "BLOCK_COUNT": "GAMMA", "BLOCKS": "GAMMA",
         "INTERVALS": "GAMMA", "RESIDUALS": "ZETA"}
    c.update(codes or {})
    w = BitWriter()
    for x, s in enumerate(succ):
        s = sorted(s)
        w.code(c["OUTDEGREES"], len(s), zetak)
        if not s:
            continue
        copied = set()
        if window > 0:
            cands = list(range(0, min(window, x) + 1))
            if choose_ref is None:
                r = max(cands, key=lambda r: (len(set(succ[x - r]) & set(s)) if r else 0, -r))
            else:
                r = choose_ref(x, cands)
            w.code(c["REFERENCES"], r, zetak)
            if r > 0:
                ref = sorted(succ[x - r])
                sset = set(s)
                runs, cur, copy = [], 0, True  # alternating runs, the first one copies
                for y in ref:
                    if (y in sset) == copy:
                        cur += 1
                    else:
                        runs.append(cur)
                        cur, copy = 1, not copy
                runs.append(cur)
                blocks = runs[:-1]  # the last run is implicit
                w.code(c["BLOCK_COUNT"], len(blocks), zetak)
                for i, b in enumerate(blocks):
                    w.code(c["BLOCKS"], b if i == 0 else b - 1, zetak)
                at, copy = 0, True
                for b in runs:
                    if copy:
                        copied.update(ref[at:at + b])
                    at += b
                    copy = not copy
        extra = [y for y in s if y not in copied]
        intervals, residuals = [], []
        if min_interval > 0:
            i = 0
            while i < len(extra):
                j = i
                while j + 1 < len(extra) and extra[j + 1] == extra[j] + 1:
                    j += 1
                if j - i + 1 >= min_interval:
                    intervals.append((extra[i], j - i + 1))
                else:
                    residuals.extend(extra[i:j + 1])
                i = j + 1
        else:
            residuals = extra
        if extra and min_interval > 0:
            w.code(c["INTERVALS"], len(intervals), zetak)
            prev = None  # one past the previous interval's end
            for left, ln in intervals:
                w.code(c["INTERVALS"], int2nat(left - x) if prev is None else left - prev - 1, zetak)
                w.code(c["INTERVALS"], ln - min_interval, zetak)
                prev = left + ln
        prev = None
        for y in residuals:
            w.code(c["RESIDUALS"], int2nat(y - x) if prev is None else y - prev - 1, zetak)
            prev = y
    return w


def write(base: str, succ, window=7, min_interval=4, zetak=3, codes=None, choose_ref=None, flags_text=None):
    w = encode(succ, window, min_interval, zetak, codes, choose_ref)
    with open(base + ".graph", "wb") as f:
        f.write(w.to_bytes())
    flags = flags_text if flags_text is not None else "|".join(f"{k}_{v}" for k, v in (codes or {}).items())
    with open(base + ".properties", "w") as f:
        f.write("#BVGraph properties\n")
        f.write("graphclass=it.unimi.dsi.webgraph.BVGraph\nversion=0\n")
        f.write(f"nodes={len(succ)}\narcs={sum(len(s) for s in succ)}\n")
        f.write(f"windowsize={window}\nmaxrefcount=3\nminintervallength={min_interval}\nzetak={zetak}\n")
        f.write(f"compressionflags={flags}\n")
