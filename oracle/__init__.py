"""CPU oracle for exact triangle counting -- TEST INFRASTRUCTURE ONLY.

ctypes binding of oracle/liboracle.so (plain C + OpenMP, built by
``make -C oracle``). Only tests/, ``__graft_entry__.smoke()`` and the CPU
baseline leg of bench.py use it, always as the checker; the HIP library never
calls it. See oracle/oracle.h for what pins each function.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")
EXE_PATH = os.path.join(HERE, "tc_oracle")
SENTINEL = np.uint64(0xFFFFFFFFFFFFFFFF)

_lib = None
_U64 = ctypes.c_uint64
_U64P = ctypes.POINTER(ctypes.c_uint64)
_P = ctypes.c_void_p


class _Graph(ctypes.Structure):
    _fields_ = [("n", _U64), ("m", _U64), ("row_ptr", _U64P), ("col", ctypes.POINTER(ctypes.c_uint32)),
                ("deg", ctypes.POINTER(ctypes.c_uint32))]


class _RefMeta(ctypes.Structure):
    _fields_ = [(f, _U64) for f in ("m", "w_l", "m_purged", "w_u", "r", "t", "blocks", "triangles")]


def build() -> None:
    subprocess.run(["make", "-s", "-C", HERE], check=True)


def lib() -> ctypes.CDLL:
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} missing: run `make -C oracle`")
        L = ctypes.CDLL(LIB_PATH)
        L.or_read_mtx.argtypes = [ctypes.c_char_p, ctypes.POINTER(_U64P), _U64P, _U64P, _U64P]
        L.or_gen_rmat.argtypes = [ctypes.c_uint32, ctypes.c_uint32, ctypes.c_double, ctypes.c_double,
                                  ctypes.c_double, _U64, _P]
        L.or_gen_er.argtypes = [_U64, _U64, _U64, _P]
        L.or_chung_lu_samples.argtypes = [_U64, ctypes.c_double]
        L.or_chung_lu_samples.restype = _U64
        L.or_chung_lu_cdf.argtypes = [_U64, ctypes.c_double, _P]
        L.or_gen_chung_lu.argtypes = [_U64, ctypes.c_double, ctypes.c_double, _U64, _P]
        L.or_chung_lu_batch_params.argtypes = [_U64, _U64, ctypes.POINTER(ctypes.c_double), _U64P,
                                               ctypes.POINTER(ctypes.c_double)]
        L.or_gen_rgg.argtypes = [_U64, ctypes.c_double, _U64, _P, _U64]
        L.or_gen_rgg.restype = _U64
        L.or_sbm_slots.argtypes = [_U64, _U64, ctypes.c_double]
        L.or_sbm_slots.restype = _U64
        L.or_gen_sbm.argtypes = [_U64, _U64, ctypes.c_double, ctypes.c_double, _U64, _P]
        L.or_dedup.argtypes = [_P, _U64]
        L.or_dedup.restype = _U64
        L.or_orient_csr.argtypes = [_P, _U64, _U64, ctypes.POINTER(_Graph)]
        L.or_id_csr.argtypes = [_P, _U64, _U64, ctypes.POINTER(_Graph)]
        L.or_sort_u64.argtypes = [_P, _U64]
        L.or_sort_u64.restype = None
        L.or_degree_labels.argtypes = [_P, _U64, _U64, _P, _P]
        L.or_octave_labels.argtypes = [_P, _U64, _U64, _P, _P]
        L.or_relabel_keys.argtypes = [_P, _U64, _P, _P]
        L.or_relabel_keys.restype = None
        L.or_graph_free.argtypes = [ctypes.POINTER(_Graph)]
        L.or_wplus.argtypes = [ctypes.POINTER(_Graph)]
        L.or_wplus.restype = _U64
        L.or_tiles.argtypes = [ctypes.POINTER(_Graph), ctypes.c_int, _P, _P, _P]
        L.or_tiles.restype = _U64
        L.or_count.argtypes = [ctypes.POINTER(_Graph), ctypes.c_int]
        L.or_count.restype = _U64
        L.or_count_rows.argtypes = [ctypes.POINTER(_Graph), _U64, _U64, ctypes.c_int]
        L.or_count_rows.restype = _U64
        L.or_count_strided.argtypes = [ctypes.POINTER(_Graph), _U64, _U64, ctypes.c_int]
        L.or_count_strided.restype = _U64
        L.or_count_bruteforce.argtypes = [_U64, _P, _U64]
        L.or_count_bruteforce.restype = _U64
        L.or_ref_pipeline.argtypes = [_P, _U64, _U64, ctypes.POINTER(_RefMeta)]
        L.or_mix64.argtypes = [_U64]
        L.or_mix64.restype = _U64
        L.or_free.argtypes = [_P]
        _lib = L
    return _lib


def _ptr(a: np.ndarray):
    return a.ctypes.data if a.size else None


def read_mtx(path: str):
    L = lib()
    kp = _U64P()
    nk, n, ne = _U64(), _U64(), _U64()
    st = L.or_read_mtx(path.encode(), ctypes.byref(kp), ctypes.byref(nk), ctypes.byref(n), ctypes.byref(ne))
    if st != 0:
        return st, None, 0, 0
    keys = np.ctypeslib.as_array(kp, shape=(nk.value,)).copy() if nk.value else np.zeros(0, np.uint64)
    L.or_free(ctypes.cast(kp, _P))
    return 0, keys, n.value, ne.value


def gen_rmat(scale: int, ef: int, a=0.57, b=0.19, c=0.19, seed: int = 1) -> np.ndarray:
    keys = np.empty(ef << scale, np.uint64)
    assert lib().or_gen_rmat(scale, ef, a, b, c, seed, _ptr(keys)) == 0
    return keys


def gen_er(n: int, nsamples: int, seed: int) -> np.ndarray:
    keys = np.empty(nsamples, np.uint64)
    assert lib().or_gen_er(n, nsamples, seed, _ptr(keys)) == 0
    return keys


def chung_lu_cdf(n: int, gamma: float) -> np.ndarray:
    cdf = np.empty(n, np.uint64)
    assert lib().or_chung_lu_cdf(n, gamma, _ptr(cdf)) == 0
    return cdf


def gen_chung_lu(n: int, gamma: float, avg_deg: float, seed: int) -> np.ndarray:
    keys = np.empty(lib().or_chung_lu_samples(n, avg_deg), np.uint64)
    assert lib().or_gen_chung_lu(n, gamma, avg_deg, seed, _ptr(keys)) == 0
    return keys


def chung_lu_batch_params(seed: int, index: int):
    g, d, n = ctypes.c_double(), ctypes.c_double(), _U64()
    lib().or_chung_lu_batch_params(seed, index, ctypes.byref(g), ctypes.byref(n), ctypes.byref(d))
    return g.value, n.value, d.value


def gen_rgg(n: int, avg_deg: float, seed: int) -> np.ndarray:
    cap = int(n * avg_deg) + 1024
    keys = np.empty(cap, np.uint64)
    got = lib().or_gen_rgg(n, avg_deg, seed, _ptr(keys), cap)
    assert got != 0xFFFFFFFFFFFFFFFF
    return keys[:got].copy()


def gen_sbm(blocks: int, size: int, p_in: float, p_out: float, seed: int) -> np.ndarray:
    keys = np.empty(lib().or_sbm_slots(blocks, size, p_out), np.uint64)
    assert lib().or_gen_sbm(blocks, size, p_in, p_out, seed, _ptr(keys)) == 0
    return keys


def dedup(keys: np.ndarray) -> np.ndarray:
    k = np.array(keys, dtype=np.uint64, copy=True)
    m = lib().or_dedup(_ptr(k), k.size)
    return k[:m].copy()


def degree_labels(ukeys: np.ndarray, n: int):
    """(old_of_new, new_of_old): new id = rank of (degree, id)."""
    k = np.ascontiguousarray(ukeys, dtype=np.uint64)
    o2 = np.zeros(max(n, 1), np.uint32)
    n2 = np.zeros(max(n, 1), np.uint32)
    st = lib().or_degree_labels(_ptr(k), k.size, n, _ptr(o2), _ptr(n2))
    if st != 0:
        raise ValueError(f"or_degree_labels status {st}")
    return o2[:n], n2[:n]


def octave_labels(ukeys: np.ndarray, n: int):
    """(old_of_new, new_of_old): new id = rank of (bit length of the degree, id), the id order of
    TRIBIT_OPT_RELABEL 2 (no counterpart in the reference)."""
    k = np.ascontiguousarray(ukeys, dtype=np.uint64)
    o2 = np.zeros(max(n, 1), np.uint32)
    n2 = np.zeros(max(n, 1), np.uint32)
    st = lib().or_octave_labels(_ptr(k), k.size, n, _ptr(o2), _ptr(n2))
    if st != 0:
        raise ValueError(f"or_octave_labels status {st}")
    return o2[:n], n2[:n]


def relabel(ukeys: np.ndarray, n: int, mode: int = 1):
    """Deduplicated keys renumbered by degree rank (mode 1) or octave class (mode 2), sorted, and
    old_of_new."""
    k = np.ascontiguousarray(ukeys, dtype=np.uint64)
    o2, n2 = octave_labels(k, n) if int(mode) == 2 else degree_labels(k, n)
    out = np.zeros(max(k.size, 1), np.uint64)
    lib().or_relabel_keys(_ptr(k), k.size, _ptr(n2), _ptr(out))
    out = out[:k.size].copy()
    lib().or_sort_u64(_ptr(out), out.size)
    return out, o2


class Graph:
    """Degree-oriented CSR built by the oracle from deduplicated keys (by_id: {a < b} oriented a -> b)."""

    def __init__(self, ukeys: np.ndarray, n: int, by_id: bool = False):
        self._k = np.ascontiguousarray(ukeys, dtype=np.uint64)
        self.g = _Graph()
        fn = lib().or_id_csr if by_id else lib().or_orient_csr
        st = fn(_ptr(self._k), self._k.size, n, ctypes.byref(self.g))
        if st != 0:
            raise ValueError(f"or_orient_csr status {st}")
        self.n, self.m = self.g.n, self.g.m

    @classmethod
    def from_keys(cls, keys: np.ndarray, n: int, relabel=False) -> "Graph":
        """relabel: False / 0 input ids, True / 1 degree rank, 2 octave classes."""
        u = dedup(keys)
        if relabel:
            mode = 2 if relabel == 2 else 1
            u, old_of_new = globals()["relabel"](u, n, mode)
            self = cls(u, n, by_id=mode == 2)
            self.labels = old_of_new
            return self
        return cls(u, n)

    @classmethod
    def from_csr(cls, row_ptr: np.ndarray, col: np.ndarray) -> "Graph":
        """Wraps an existing oriented CSR (arrays are referenced, not copied)."""
        self = cls.__new__(cls)
        self._rp = np.ascontiguousarray(row_ptr, dtype=np.uint64)
        self._col = np.ascontiguousarray(col, dtype=np.uint32)
        self._borrowed = True
        self.g = _Graph()
        self.g.n = self._rp.size - 1
        self.g.m = self._col.size
        self.g.row_ptr = self._rp.ctypes.data_as(_U64P)
        self.g.col = self._col.ctypes.data_as(ctypes.POINTER(ctypes.c_uint32))
        self.n, self.m = self.g.n, self.g.m
        return self

    def count_strided(self, start: int, stride: int, threads: int = 0) -> int:
        return lib().or_count_strided(ctypes.byref(self.g), start, stride, threads)

    def arrays(self):
        rp = np.ctypeslib.as_array(self.g.row_ptr, shape=(self.n + 1,)).copy()
        col = np.ctypeslib.as_array(self.g.col, shape=(self.m,)).copy() if self.m else np.zeros(0, np.uint32)
        deg = np.ctypeslib.as_array(self.g.deg, shape=(self.n,)).copy() if self.n else np.zeros(0, np.uint32)
        return rp, col, deg

    def wplus(self) -> int:
        return lib().or_wplus(ctypes.byref(self.g))

    def count(self, threads: int = 0) -> int:
        return lib().or_count(ctypes.byref(self.g), threads)

    def count_rows(self, r0: int, r1: int, threads: int = 0) -> int:
        return lib().or_count_rows(ctypes.byref(self.g), r0, r1, threads)

    def tiles(self, width: int):
        L = lib()
        w = L.or_tiles(ctypes.byref(self.g), width, None, None, None)
        tp = np.zeros(self.n + 1, np.uint64)
        blk = np.zeros(max(w, 1), np.uint32)
        mask = np.zeros(max(w, 1), np.uint64)
        L.or_tiles(ctypes.byref(self.g), width, _ptr(tp), _ptr(blk), _ptr(mask))
        return tp, blk[:w], mask[:w]

    def __del__(self):
        if getattr(self, "_borrowed", False):
            return
        try:
            lib().or_graph_free(ctypes.byref(self.g))
        except Exception:
            pass


def count_keys(keys: np.ndarray, n: int, threads: int = 0) -> int:
    return Graph.from_keys(keys, n).count(threads)


def count_bruteforce(n: int, keys: np.ndarray) -> int:
    k = np.ascontiguousarray(keys, dtype=np.uint64)
    return lib().or_count_bruteforce(n, _ptr(k), k.size)


def ref_pipeline(ukeys: np.ndarray, n: int) -> dict:
    k = np.ascontiguousarray(ukeys, dtype=np.uint64)
    r = _RefMeta()
    assert lib().or_ref_pipeline(_ptr(k), k.size, n, ctypes.byref(r)) == 0
    return {f: getattr(r, f) for f, _ in _RefMeta._fields_}


def mix64(z: int) -> int:
    return lib().or_mix64(z)
