"""tribit_amd -- ctypes binding of libtribit_hip.so, exact triangle counting on MI355X.

The library (tribit_amd/libtribit_hip.so, built by ``python -m tribit_amd.build``)
holds every GPU kernel; this module only marshals arguments across the C ABI
declared in include/tribit.h. There is no CPU fallback: if the library is
missing, ``load_library`` raises.

Reference surface mirrored (sixt99/TriBIT): ``tribit -i <graph.mtx>``
(src/single_gpu/tribit.cu:840-866) is the ``tribit`` executable next to this
file; ``read_mtx`` (tribit.cu:228-269) is :func:`read_mtx`; ``count_triangles``
(tribit.cu:656-823) is :meth:`Context.graph_from_keys` + :meth:`Graph.count`.
"""
from __future__ import annotations

import ctypes
import os
import re
from dataclasses import dataclass

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libtribit_hip.so")
CLI_PATH = os.path.join(_HERE, "tribit")
HEADER_PATH = os.path.join(os.path.dirname(_HERE), "include", "tribit.h")

STATUS = {
    0: "OK", 1: "ERR_IO", 2: "ERR_FORMAT", 3: "ERR_SHAPE", 4: "ERR_RANGE", 5: "ERR_NOMEM",
    6: "ERR_ARG", 7: "ERR_HIP", 8: "ERR_NODEVICE", 9: "ERR_TOO_LARGE",
}
SCHED = {"plain": 0, "learned": 1, "directed": 2, "wedge": 3}


class TribitError(RuntimeError):
    def __init__(self, status: int, what: str):
        self.status = status
        self.name = STATUS.get(status, f"status {status}")
        super().__init__(f"{what}: {self.name}")


class GraphInfo(ctypes.Structure):
    _fields_ = [
        ("n", ctypes.c_uint64), ("entries", ctypes.c_uint64), ("m", ctypes.c_uint64),
        ("wplus", ctypes.c_uint64), ("rows_active", ctypes.c_uint64), ("max_outdeg", ctypes.c_uint32),
        ("reserved0", ctypes.c_uint32), ("tiles32", ctypes.c_uint64), ("tiles64", ctypes.c_uint64),
        ("ms_preprocess", ctypes.c_double), ("device_bytes", ctypes.c_uint64),
        ("ms_layout", ctypes.c_double * 4), ("ms_blocks", ctypes.c_double),
    ]


class CountOpts(ctypes.Structure):
    _fields_ = [("sched", ctypes.c_int32), ("repeats", ctypes.c_int32), ("flags", ctypes.c_int32),
                ("group_width", ctypes.c_int32 * 4), ("dense_core", ctypes.c_int32)]


FORCE_LINEAR_PROBE = 1
FORCE_GLOBAL_LOADS = 2
NO_TILES = 4
FORCE_TILES = 8
TILES32 = 256
NO_BITMAP = 16
WIDE_IDS = 32
NO_JUDGE = 64
FORCE_MULTI = 128
DENSE_BLOCKS = 512
OPT_RELABEL = 1
OPT_BATCH_PART_KEYS = 2


class CountInfo(ctypes.Structure):
    _fields_ = [
        ("triangles", ctypes.c_uint64), ("ms_count", ctypes.c_double), ("ms_sched", ctypes.c_double),
        ("ms_kernel", ctypes.c_double), ("launch_blocks", ctypes.c_uint64), ("tasks", ctypes.c_uint64),
        ("tasks_by_variant", ctypes.c_uint64 * 8), ("repeats_agree", ctypes.c_int32), ("nvariants", ctypes.c_int32),
        ("ms_by_variant", ctypes.c_double * 8), ("fallback_tasks", ctypes.c_uint64),
        ("group_width", ctypes.c_int32 * 8), ("dense_core", ctypes.c_uint64), ("ms_dense", ctypes.c_double),
        ("candidate", ctypes.c_int32), ("layout", ctypes.c_int32), ("ms_judge", ctypes.c_double),
        ("ms_candidates", ctypes.c_double * 5), ("slot_kind", ctypes.c_int32 * 8),
        ("dense_blocks", ctypes.c_uint32), ("reserved1", ctypes.c_uint32),
    ]


COST_MODEL_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "cost_model.f32")


class BatchCountInfo(ctypes.Structure):
    _fields_ = [
        ("ms_count", ctypes.c_double), ("ms_sched", ctypes.c_double), ("ms_capture", ctypes.c_double),
        ("triangles", ctypes.c_uint64), ("tasks", ctypes.c_uint64), ("parts", ctypes.c_uint32),
        ("launches", ctypes.c_uint32), ("repeats_agree", ctypes.c_int32), ("nslots", ctypes.c_int32),
        ("group_width", ctypes.c_int32 * 8), ("waves", ctypes.c_int32 * 8), ("tile_width", ctypes.c_int32),
    ]


BUCKETS = 800
MENU_SLOTS = 8
FEATURES = 24
CLASSES = 3


class Profile(ctypes.Structure):
    _fields_ = [
        ("lists", ctypes.c_uint64 * BUCKETS), ("sum_d", ctypes.c_uint64 * BUCKETS),
        ("sum_k", ctypes.c_uint64 * BUCKETS), ("sum_w", ctypes.c_uint64 * BUCKETS),
        ("eligible", ctypes.c_uint64 * BUCKETS), ("sum_t", ctypes.c_uint64 * BUCKETS),
        ("ticks", (ctypes.c_uint64 * BUCKETS) * MENU_SLOTS),
        ("slot", ctypes.c_uint8 * BUCKETS), ("tick_ns", ctypes.c_double), ("ms_kernel", ctypes.c_double),
        ("triangles", ctypes.c_uint64), ("ms_slot", ctypes.c_double * MENU_SLOTS),
    ]


class VariantStats(ctypes.Structure):
    _fields_ = [("rows", ctypes.c_uint64 * 8), ("edges", ctypes.c_uint64 * 8), ("wplus", ctypes.c_uint64 * 8),
                ("narrow", ctypes.c_uint64 * 8), ("narrow_ranges", ctypes.c_uint64 * 8),
                ("narrow2_ranges", ctypes.c_uint64 * 8)]


class PartialInfo(ctypes.Structure):
    _fields_ = [("keys", ctypes.c_uint64), ("edges", ctypes.c_uint64), ("lists", ctypes.c_uint64),
                ("streamed", ctypes.c_uint64), ("triangles", ctypes.c_uint64), ("ms_h2d", ctypes.c_double),
                ("ms_build", ctypes.c_double), ("ms_count", ctypes.c_double)]


_lib = None
_P = ctypes.c_void_p
_U64 = ctypes.c_uint64
_U64P = ctypes.POINTER(ctypes.c_uint64)


def header_symbols() -> list[str]:
    """Function names declared in include/tribit.h."""
    src = open(HEADER_PATH).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(tribit_[a-z0-9_]+|init_gpu_kernel[56]|count_partial_triangles_[a-z_]+)\s*\(",
                                 src)))


def load_library(path: str | None = None) -> ctypes.CDLL:
    """Loads libtribit_hip.so (raises if it was not built)."""
    global _lib
    if _lib is not None and path is None:
        return _lib
    p = path or LIB_PATH
    if not os.path.exists(p):
        raise RuntimeError(f"{p} is missing: build it with `python -m tribit_amd.build` (no CPU fallback exists)")
    lib = ctypes.CDLL(p)
    sig = {
        "tribit_init": (ctypes.c_int, [ctypes.c_int, ctypes.POINTER(_P)]),
        "tribit_shutdown": (ctypes.c_int, [_P]),
        "tribit_set_alloc_cap": (ctypes.c_int, [_P, _U64]),
        "tribit_pool_peak": (ctypes.c_int, [_P, _U64P]),
        "tribit_status_str": (ctypes.c_char_p, [ctypes.c_int]),
        "tribit_build_info": (ctypes.c_char_p, []),
        "tribit_device_info": (ctypes.c_int, [_P, ctypes.c_char_p, ctypes.c_size_t, ctypes.POINTER(ctypes.c_int),
                                              ctypes.POINTER(ctypes.c_double)]),
        "tribit_read_mtx": (ctypes.c_int, [ctypes.c_char_p, ctypes.POINTER(_U64P), _U64P, _U64P, _U64P]),
        "tribit_read_webgraph": (ctypes.c_int, [ctypes.c_char_p, ctypes.POINTER(_U64P), _U64P, _U64P]),
        "tribit_free_host": (None, [_P]),
        "tribit_gen_rmat": (ctypes.c_int, [_P, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_double, ctypes.c_double,
                                           ctypes.c_double, _U64, ctypes.POINTER(_P), _U64P]),
        "tribit_gen_er": (ctypes.c_int, [_P, _U64, _U64, _U64, ctypes.POINTER(_P), _U64P]),
        "tribit_device_keys_to_host": (ctypes.c_int, [_P, _P, _U64, _P]),
        "tribit_free_device_keys": (ctypes.c_int, [_P, _P]),
        "tribit_gen_chung_lu": (ctypes.c_int, [_P, _U64, ctypes.c_double, ctypes.c_double, _U64, ctypes.POINTER(_P),
                                               _U64P]),
        "tribit_chung_lu_batch_params": (None, [_U64, _U64, ctypes.POINTER(ctypes.c_double), _U64P,
                                                ctypes.POINTER(ctypes.c_double)]),
        "tribit_gen_rgg": (ctypes.c_int, [_P, _U64, ctypes.c_double, _U64, ctypes.POINTER(_P), _U64P]),
        "tribit_gen_sbm": (ctypes.c_int, [_P, _U64, _U64, ctypes.c_double, ctypes.c_double, _U64, ctypes.POINTER(_P),
                                          _U64P]),
        "tribit_graph_from_host_keys": (ctypes.c_int, [_P, _P, _U64, _U64, ctypes.POINTER(_P)]),
        "tribit_graph_from_device_keys": (ctypes.c_int, [_P, _P, _U64, _U64, ctypes.POINTER(_P)]),
        "tribit_graph_get_info": (ctypes.c_int, [_P, ctypes.POINTER(GraphInfo)]),
        "tribit_graph_download": (ctypes.c_int, [_P, _P, _P, _P]),
        "tribit_graph_labels": (ctypes.c_int, [_P, _P, ctypes.POINTER(ctypes.c_int)]),
        "tribit_graph_tiles": (ctypes.c_int, [_P, _P, ctypes.c_int, _U64P, _P, _P, _P]),
        "tribit_set_option": (ctypes.c_int, [_P, ctypes.c_int, ctypes.c_int64]),
        "tribit_graph_free": (ctypes.c_int, [_P]),
        "tribit_count": (ctypes.c_int, [_P, _P, ctypes.POINTER(CountOpts), ctypes.POINTER(CountInfo)]),
        "tribit_count_noalloc": (ctypes.c_int, [_P, _P, ctypes.POINTER(CountOpts), ctypes.POINTER(CountInfo)]),
        "tribit_noalloc_scratch_bytes": (_U64, []),
        "tribit_count_profile": (ctypes.c_int, [_P, _P, ctypes.c_int, _P, ctypes.POINTER(Profile)]),
        "tribit_class_features": (ctypes.c_int, [ctypes.POINTER(Profile), ctypes.c_int, ctypes.c_int, _P]),
        "tribit_count_plan": (ctypes.c_int, [_P, _P, ctypes.c_int, _P, ctypes.c_int, ctypes.POINTER(ctypes.c_double),
                                             ctypes.POINTER(_U64)]),
        "tribit_menu_slot": (ctypes.c_int, [ctypes.c_int, ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_int),
                                            ctypes.POINTER(ctypes.c_uint32)]),
        "tribit_bucket_max_outdeg": (ctypes.c_uint32, [ctypes.c_int]),
        "tribit_set_cost_model": (ctypes.c_int, [_P, _P]),
        "tribit_learned_plan": (ctypes.c_int, [_P, _P, _P]),
        "tribit_batch_from_host": (ctypes.c_int, [_P, _P, _P, _P, ctypes.c_uint32, ctypes.POINTER(_P)]),
        "tribit_batch_gen_chung_lu": (ctypes.c_int, [_P, _U64, _U64, ctypes.c_uint32, ctypes.POINTER(_P)]),
        "tribit_batch_info": (ctypes.c_int, [_P, ctypes.POINTER(ctypes.c_uint32), ctypes.POINTER(ctypes.c_uint32),
                                             _U64P, _U64P, ctypes.POINTER(ctypes.c_double)]),
        "tribit_batch_graph_sizes": (ctypes.c_int, [_P, _P, _P]),
        "tribit_batch_count": (ctypes.c_int, [_P, _P, ctypes.POINTER(CountOpts), _P,
                                              ctypes.POINTER(BatchCountInfo)]),
        "tribit_batch_autotune": (ctypes.c_int, [_P, _P, ctypes.POINTER(CountOpts),
                                                 ctypes.POINTER(BatchCountInfo)]),
        "tribit_batch_free": (ctypes.c_int, [_P]),
        "tribit_count_partial_diagonal": (ctypes.c_int, [_P, _P, _P, ctypes.c_uint32, _U64, _U64P]),
        "tribit_count_partial_off_diagonal": (ctypes.c_int, [_P, _P, _P, ctypes.c_uint32, _P, _P, ctypes.c_uint32,
                                                             _U64, _U64P]),
        "init_gpu_kernel5": (None, [ctypes.c_int]),
        "init_gpu_kernel6": (None, [ctypes.c_int]),
        "count_partial_triangles_diagonal": (_U64, [_P, _P, ctypes.c_uint32, _U64]),
        "count_partial_triangles_off_diagonal": (_U64, [_P, _P, ctypes.c_uint32, _P, _P, ctypes.c_uint32, _U64]),
        "tribit_partial_last_status": (ctypes.c_int, []),
        "tribit_partial_last_info": (ctypes.c_int, [ctypes.POINTER(PartialInfo)]),
        "tribit_mlp_create": (ctypes.c_int, [_P, ctypes.c_int, _U64, ctypes.POINTER(_P)]),
        "tribit_mlp_free": (ctypes.c_int, [_P]),
        "tribit_mlp_num_params": (_U64, [_P]),
        "tribit_mlp_get_params": (ctypes.c_int, [_P, _P]),
        "tribit_mlp_set_params": (ctypes.c_int, [_P, _P]),
        "tribit_mlp_grad": (ctypes.c_int, [_P, _P, _P, ctypes.c_uint32, _P, ctypes.POINTER(ctypes.c_double)]),
        "tribit_mlp_train": (ctypes.c_int, [_P, _P, _P, _U64, ctypes.c_int, ctypes.c_int, ctypes.c_float, _U64,
                                            ctypes.POINTER(ctypes.c_double)]),
        "tribit_mlp_infer": (ctypes.c_int, [_P, _P, _U64, _P]),
        "tribit_variant_stats": (ctypes.c_int, [_P, _P, ctypes.c_int, ctypes.POINTER(VariantStats)]),
        "tribit_batch_variant_stats": (ctypes.c_int, [_P, _P, ctypes.c_int, ctypes.POINTER(VariantStats)]),
        "tribit_count_keys": (ctypes.c_int, [_P, _P, _U64, _U64, ctypes.POINTER(CountOpts), _U64P,
                                             ctypes.POINTER(ctypes.c_double), ctypes.POINTER(CountInfo)]),
    }
    for name, (res, args) in sig.items():
        f = getattr(lib, name)
        f.restype = res
        f.argtypes = args
    if path is None:
        _lib = lib
    return lib


def _check(st: int, what: str) -> None:
    if st != 0:
        raise TribitError(st, what)


def chung_lu_batch_params(seed: int, index: int):
    """(gamma, n, avg_deg) of graph `index` of a Chung-Lu batch drawn from `seed`."""
    g, d, n = ctypes.c_double(), ctypes.c_double(), _U64()
    load_library().tribit_chung_lu_batch_params(seed, index, ctypes.byref(g), ctypes.byref(n), ctypes.byref(d))
    return g.value, n.value, d.value


class Chunks:
    """Host key chunks in the reference's layout: uint64_t** plus a cumulative uint64_t* (num + 1)."""

    def __init__(self, arrays):
        self.arrays = [np.ascontiguousarray(a, np.uint64) for a in arrays]
        self.ptrs = (ctypes.c_void_p * max(len(self.arrays), 1))(*[a.ctypes.data for a in self.arrays])
        self.cum = np.zeros(len(self.arrays) + 1, np.uint64)
        self.cum[1:] = np.cumsum([a.size for a in self.arrays])
        self.num = len(self.arrays)


def count_partial_diagonal(ctx: "Context", chunks, num_nodes: int) -> int:
    c = Chunks(chunks)
    out = _U64()
    _check(ctx.lib.tribit_count_partial_diagonal(ctx.h, ctypes.addressof(c.ptrs), c.cum.ctypes.data, c.num,
                                                 num_nodes, ctypes.byref(out)), "count_partial_diagonal")
    return out.value


def count_partial_off_diagonal(ctx: "Context", cols, idxs, num_nodes: int) -> int:
    c, x = Chunks(cols), Chunks(idxs)
    out = _U64()
    _check(ctx.lib.tribit_count_partial_off_diagonal(ctx.h, ctypes.addressof(c.ptrs), c.cum.ctypes.data, c.num,
                                                     ctypes.addressof(x.ptrs), x.cum.ctypes.data, x.num, num_nodes,
                                                     ctypes.byref(out)), "count_partial_off_diagonal")
    return out.value


def partial_last_info(ctx: "Context") -> dict:
    """Sizes and device times of this thread's last partial count (tribit_partial_last_info)."""
    pi = PartialInfo()
    _check(ctx.lib.tribit_partial_last_info(ctypes.byref(pi)), "partial_last_info")
    return {f: getattr(pi, f) for f, _ in PartialInfo._fields_}


def class_features(p: Profile, cls: int, slot: int) -> np.ndarray:
    """Cost-model features of (owner-outdeg class, menu slot) from a profile's bucket statistics."""
    f = np.zeros(FEATURES, np.float32)
    _check(load_library().tribit_class_features(ctypes.byref(p), cls, slot, f.ctypes.data), "class_features")
    return f


def outdeg_class(bucket: int) -> int:
    """The heuristic's class of a bucket: owner outdeg <= 32 (0), <= 256 (1), larger (2)."""
    top = bucket_max_outdeg(bucket)
    return 0 if top <= 32 else 1 if top <= 256 else 2


def class_slots(cls: int) -> list:
    """Menu slots that every bucket of the class fits."""
    top = (32, 256, 0xFFFFFFFF)[cls]
    return [s for s, (_, _, dmax) in enumerate(menu()) if top <= dmax]


def heuristic_slots(sched: str = "wedge") -> list:
    """The density-threshold heuristic's menu slot per outdeg class."""
    return [0, 3, 5 if sched == "wedge" else 6]


def class_table(assign) -> np.ndarray:
    """Menu plan (slot per bucket) that runs class c on assign[c]."""
    return np.array([assign[outdeg_class(b)] for b in range(BUCKETS)], np.uint8)


def menu():
    """[(kind, group_width, max_outdeg)] of the learned scheduler's menu slots."""
    out = []
    for s in range(MENU_SLOTS):
        k, g, d = ctypes.c_int(), ctypes.c_int(), ctypes.c_uint32()
        _check(load_library().tribit_menu_slot(s, ctypes.byref(k), ctypes.byref(g), ctypes.byref(d)), "menu_slot")
        out.append((k.value, g.value, d.value))
    return out


def bucket_max_outdeg(bucket: int) -> int:
    return int(load_library().tribit_bucket_max_outdeg(bucket))


def build_info() -> str:
    return load_library().tribit_build_info().decode()


def read_mtx(path: str):
    """Matrix Market -> (keys uint64[(min<<32)|max], n, entries), semantics of tribit.cu:228-269."""
    lib = load_library()
    kp = _U64P()
    nk, n, ne = _U64(), _U64(), _U64()
    _check(lib.tribit_read_mtx(path.encode(), ctypes.byref(kp), ctypes.byref(nk), ctypes.byref(n),
                               ctypes.byref(ne)), f"read_mtx({path})")
    try:
        keys = np.ctypeslib.as_array(kp, shape=(nk.value,)).copy() if nk.value else np.zeros(0, np.uint64)
    finally:
        lib.tribit_free_host(ctypes.cast(kp, _P))
    return keys, n.value, ne.value


def read_webgraph(basename: str):
    """BVGraph <basename>.properties/.graph -> (keys uint64[(min<<32)|max] per arc, n), the reference's
    multi-GPU ingest (src/multi_gpu/src/main.rs:88, :150-161)."""
    lib = load_library()
    kp = _U64P()
    nk, n = _U64(), _U64()
    _check(lib.tribit_read_webgraph(basename.encode(), ctypes.byref(kp), ctypes.byref(nk), ctypes.byref(n)),
           f"read_webgraph({basename})")
    try:
        keys = np.ctypeslib.as_array(kp, shape=(nk.value,)).copy() if nk.value else np.zeros(0, np.uint64)
    finally:
        lib.tribit_free_host(ctypes.cast(kp, _P))
    return keys, n.value


@dataclass
class CountResult:
    triangles: int
    ms_count: float
    ms_sched: float
    ms_kernel: float
    launch_blocks: int
    tasks: int
    tasks_by_variant: list
    repeats_agree: bool
    ms_by_variant: list
    fallback_tasks: int = 0
    group_width: tuple = ()
    dense_core: int = 0
    ms_dense: float = 0.0
    candidate: int = -1
    layout: int = 0
    ms_judge: float = 0.0
    ms_candidates: tuple = ()
    slot_kind: tuple = ()
    dense_blocks: int = 0


class Mlp:
    """The learned scheduler's cost model (F -> 64 -> 64 -> 1), trained by HIP kernels."""

    def __init__(self, ctx: "Context", nfeat: int, seed: int = 1):
        self.ctx = ctx
        self.nfeat = nfeat
        h = _P()
        _check(ctx.lib.tribit_mlp_create(ctx.h, nfeat, seed, ctypes.byref(h)), "mlp_create")
        self.h = h

    def num_params(self) -> int:
        return int(self.ctx.lib.tribit_mlp_num_params(self.h))

    def params(self) -> np.ndarray:
        p = np.zeros(self.num_params(), np.float32)
        _check(self.ctx.lib.tribit_mlp_get_params(self.h, p.ctypes.data), "mlp_get_params")
        return p

    def set_params(self, p: np.ndarray) -> None:
        p = np.ascontiguousarray(p, np.float32)
        assert p.size == self.num_params()
        _check(self.ctx.lib.tribit_mlp_set_params(self.h, p.ctypes.data), "mlp_set_params")

    def grad(self, X: np.ndarray, t: np.ndarray):
        X = np.ascontiguousarray(X, np.float32)
        t = np.ascontiguousarray(t, np.float32)
        g = np.zeros(self.num_params(), np.float32)
        loss = ctypes.c_double()
        _check(self.ctx.lib.tribit_mlp_grad(self.h, X.ctypes.data, t.ctypes.data, t.size, g.ctypes.data,
                                            ctypes.byref(loss)), "mlp_grad")
        return loss.value, g

    def train(self, X: np.ndarray, t: np.ndarray, epochs: int = 10, batch: int = 256, lr: float = 1e-3,
              seed: int = 1) -> float:
        X = np.ascontiguousarray(X, np.float32)
        t = np.ascontiguousarray(t, np.float32)
        loss = ctypes.c_double()
        _check(self.ctx.lib.tribit_mlp_train(self.h, X.ctypes.data, t.ctypes.data, t.size, epochs, batch, lr, seed,
                                             ctypes.byref(loss)), "mlp_train")
        return loss.value

    def infer(self, X: np.ndarray) -> np.ndarray:
        X = np.ascontiguousarray(X, np.float32)
        out = np.zeros(max(X.shape[0], 1), np.float32)
        _check(self.ctx.lib.tribit_mlp_infer(self.h, X.ctypes.data, X.shape[0], out.ctypes.data), "mlp_infer")
        return out[:X.shape[0]]

    def free(self) -> None:
        if self.h:
            self.ctx.lib.tribit_mlp_free(self.h)
            self.h = _P()

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.free()


def _info_dict(ci) -> dict:
    d = {f: getattr(ci, f) for f, _ in ci._fields_}
    n = ci.nslots
    d["group_width"] = list(ci.group_width)[:n]
    d["waves"] = list(ci.waves)[:n]
    d["repeats_agree"] = bool(ci.repeats_agree)
    return d


class Batch:
    """Many graphs resident in HBM, counted by one captured hipGraph launch."""

    def __init__(self, ctx: "Context", handle):
        self.ctx = ctx
        self.h = handle

    def info(self) -> dict:
        ng, npart = ctypes.c_uint32(), ctypes.c_uint32()
        tn, tm, ms = _U64(), _U64(), ctypes.c_double()
        _check(self.ctx.lib.tribit_batch_info(self.h, ctypes.byref(ng), ctypes.byref(npart), ctypes.byref(tn),
                                              ctypes.byref(tm), ctypes.byref(ms)), "batch_info")
        return {"graphs": ng.value, "parts": npart.value, "n": tn.value, "m": tm.value, "ms_preprocess": ms.value}

    def sizes(self):
        ng = self.info()["graphs"]
        n = np.zeros(ng, np.uint64)
        m = np.zeros(ng, np.uint64)
        _check(self.ctx.lib.tribit_batch_graph_sizes(self.h, n.ctypes.data, m.ctypes.data), "batch_graph_sizes")
        return n, m

    def _opts(self, sched, repeats, flags):
        o = CountOpts()
        o.sched = SCHED[sched]
        o.repeats = repeats
        o.flags = flags
        return o

    def count(self, sched: str = "learned", repeats: int = 1, flags: int = 0):
        """(per-graph counts uint64[graphs], info dict)."""
        ng = self.info()["graphs"]
        out = np.zeros(ng, np.uint64)
        o = self._opts(sched, repeats, flags)
        ci = BatchCountInfo()
        _check(self.ctx.lib.tribit_batch_count(self.ctx.h, self.h, ctypes.byref(o), out.ctypes.data,
                                               ctypes.byref(ci)), "batch_count")
        return out, _info_dict(ci)

    def variant_stats(self, sched: str = "wedge") -> dict:
        """Work of every part's schedule summed over the batch (units of Graph.variant_stats)."""
        vs = VariantStats()
        _check(self.ctx.lib.tribit_batch_variant_stats(self.ctx.h, self.h, SCHED[sched], ctypes.byref(vs)),
               "batch_variant_stats")
        return {"rows": list(vs.rows), "edges": list(vs.edges), "wplus": list(vs.wplus),
                "narrow": list(vs.narrow), "narrow_ranges": list(vs.narrow_ranges),
                "narrow2_ranges": list(vs.narrow2_ranges)}

    def autotune(self, sched: str = "learned", repeats: int = 3, flags: int = 0) -> dict:
        o = self._opts(sched, repeats, flags)
        ci = BatchCountInfo()
        _check(self.ctx.lib.tribit_batch_autotune(self.ctx.h, self.h, ctypes.byref(o), ctypes.byref(ci)),
               "batch_autotune")
        return _info_dict(ci)

    def free(self) -> None:
        if self.h:
            self.ctx.lib.tribit_batch_free(self.h)
            self.h = _P()

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.free()


class Graph:
    """Count-ready graph resident in HBM (oriented CSR + per-row wedge counts)."""

    def __init__(self, ctx: "Context", handle: int):
        self.ctx = ctx
        self.h = _P(handle)

    def info(self) -> dict:
        gi = GraphInfo()
        _check(self.ctx.lib.tribit_graph_get_info(self.h, ctypes.byref(gi)), "graph_get_info")
        d = {f: getattr(gi, f) for f, _ in GraphInfo._fields_ if not f.startswith("reserved")}
        d["ms_layout"] = list(gi.ms_layout)
        d["ms_blocks"] = gi.ms_blocks
        return d

    def download(self):
        """(row_ptr uint32[n+1], col uint32[m], deg uint32[n]) of the oriented CSR."""
        inf = self.info()
        n, m = inf["n"], inf["m"]
        rp = np.zeros(n + 1, np.uint32)
        col = np.zeros(max(m, 1), np.uint32)
        deg = np.zeros(max(n, 1), np.uint32)
        _check(self.ctx.lib.tribit_graph_download(self.h, rp.ctypes.data, col.ctypes.data, deg.ctypes.data),
               "graph_download")
        return rp, col[:m], deg[:n]

    def profile(self, sched: str = "wedge", table=None) -> Profile:
        """One count under a menu plan (slot per bucket; None = heuristic) with per-task timing logs."""
        p = Profile()
        tp = None
        if table is not None:
            t = np.ascontiguousarray(table, np.uint8)
            assert t.size == BUCKETS
            tp = t.ctypes.data
        _check(self.ctx.lib.tribit_count_profile(self.ctx.h, self.h, SCHED[sched], tp, ctypes.byref(p)),
               "count_profile")
        return p

    def count_plan(self, sched: str = "wedge", table=None, repeats: int = 3):
        """(median device ms of the counting launches, triangles) under a menu plan (None = heuristic)."""
        tp = None
        if table is not None:
            t = np.ascontiguousarray(table, np.uint8)
            assert t.size == BUCKETS
            tp = t.ctypes.data
        ms, tri = ctypes.c_double(), _U64()
        _check(self.ctx.lib.tribit_count_plan(self.ctx.h, self.h, SCHED[sched], tp, repeats, ctypes.byref(ms),
                                              ctypes.byref(tri)), "count_plan")
        return ms.value, int(tri.value)

    def learned_plan(self) -> np.ndarray:
        t = np.zeros(BUCKETS, np.uint8)
        _check(self.ctx.lib.tribit_learned_plan(self.ctx.h, self.h, t.ctypes.data), "learned_plan")
        return t

    def tiles(self, width: int = 64):
        """(tile_ptr uint32[n+1], blk uint32[T], mask uint32|uint64[T]) of the `width`-column bit tiles."""
        nt = _U64()
        lib = self.ctx.lib
        _check(lib.tribit_graph_tiles(self.ctx.h, self.h, width, ctypes.byref(nt), None, None, None), "graph_tiles")
        n = self.info()["n"]
        tp = np.zeros(n + 1, np.uint32)
        blk = np.zeros(max(nt.value, 1), np.uint32)
        mask = np.zeros(max(nt.value, 1), np.uint32 if width == 32 else np.uint64)
        _check(lib.tribit_graph_tiles(self.ctx.h, self.h, width, ctypes.byref(nt), tp.ctypes.data, blk.ctypes.data,
                                      mask.ctypes.data), "graph_tiles")
        return tp, blk[:nt.value], mask[:nt.value]

    def id_order(self) -> int:
        """The ids the graph got: 0 input ids, 1 degree rank, 2 octave classes."""
        rl = ctypes.c_int()
        _check(self.ctx.lib.tribit_graph_labels(self.h, None, ctypes.byref(rl)), "graph_labels")
        return int(rl.value)

    def labels(self):
        """(input id of every internal id uint32[n], id order: 0 input ids, 1 degree rank, 2 octave classes)."""
        n = self.info()["n"]
        ids = np.zeros(max(n, 1), np.uint32)
        rl = ctypes.c_int()
        _check(self.ctx.lib.tribit_graph_labels(self.h, ids.ctypes.data, ctypes.byref(rl)), "graph_labels")
        return ids[:n], int(rl.value)

    def count(self, sched: str = "plain", repeats: int = 1, flags: int = 0, group_width=None,
              dense_core: int = 0) -> CountResult:
        o = CountOpts()
        o.sched = SCHED[sched]
        o.repeats = repeats
        o.flags = flags
        o.dense_core = dense_core
        for i, w in enumerate(group_width or ()):
            o.group_width[i] = int(w)
        ci = CountInfo()
        _check(self.ctx.lib.tribit_count(self.ctx.h, self.h, ctypes.byref(o), ctypes.byref(ci)), "count")
        nv = ci.nvariants
        return CountResult(ci.triangles, ci.ms_count, ci.ms_sched, ci.ms_kernel, ci.launch_blocks, ci.tasks,
                           list(ci.tasks_by_variant)[:max(nv, 4)], bool(ci.repeats_agree),
                           list(ci.ms_by_variant)[:nv], ci.fallback_tasks, tuple(ci.group_width)[:nv],
                           ci.dense_core, ci.ms_dense, ci.candidate, ci.layout, ci.ms_judge,
                           tuple(ci.ms_candidates), tuple(ci.slot_kind)[:nv], ci.dense_blocks)

    def variant_stats(self, sched: str = "plain") -> dict:
        vs = VariantStats()
        _check(self.ctx.lib.tribit_variant_stats(self.ctx.h, self.h, SCHED[sched], ctypes.byref(vs)), "variant_stats")
        return {"rows": list(vs.rows), "edges": list(vs.edges), "wplus": list(vs.wplus),
                "narrow": list(vs.narrow), "narrow_ranges": list(vs.narrow_ranges),
                "narrow2_ranges": list(vs.narrow2_ranges)}

    def free(self) -> None:
        if self.h:
            self.ctx.lib.tribit_graph_free(self.h)
            self.h = _P()

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.free()


class Context:
    """One device context (stream + pool), as init_gpu_kernel5 (main5_multi.cu:578-587)."""

    def __init__(self, device: int = 0):
        self.lib = load_library()
        h = _P()
        _check(self.lib.tribit_init(device, ctypes.byref(h)), f"init(device={device})")
        self.h = h

    def close(self) -> None:
        if self.h:
            own = getattr(self, "_owned_model", None)
            if own is not None:
                self.set_cost_model(None)
                own.free()
                self._owned_model = None
            self.lib.tribit_shutdown(self.h)
            self.h = _P()

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def device_info(self) -> dict:
        buf = ctypes.create_string_buffer(256)
        cus = ctypes.c_int()
        peak = ctypes.c_double()
        _check(self.lib.tribit_device_info(self.h, buf, 256, ctypes.byref(cus), ctypes.byref(peak)), "device_info")
        return {"name": buf.value.decode(), "cus": cus.value, "hbm_peak_gbs": peak.value}

    def mlp(self, nfeat: int = FEATURES, seed: int = 1) -> Mlp:
        return Mlp(self, nfeat, seed)

    def set_cost_model(self, m) -> None:
        """Model used by sched="learned" (None: the density-threshold heuristic)."""
        _check(self.lib.tribit_set_cost_model(self.h, m.h if m is not None else None), "set_cost_model")
        self._model = m

    def load_cost_model(self, path: str | None = None) -> bool:
        """Loads trained parameters (default: the package's cost_model.f32) and sets the model.
        Returns False when the file does not exist (sched="learned" then keeps the heuristic)."""
        path = path or COST_MODEL_PATH
        if not os.path.exists(path):
            return False
        p = np.fromfile(path, dtype="<f4")
        m = self.mlp(FEATURES, seed=1)
        if p.size != m.num_params():
            m.free()
            raise ValueError(f"{path}: {p.size} parameters, model has {m.num_params()}")
        m.set_params(p)
        self.set_cost_model(m)
        old = getattr(self, "_owned_model", None)
        if old is not None:
            old.free()
        self._owned_model = m
        return True

    def batch_from_keys(self, keys_list, n_list) -> Batch:
        """Batch of graphs given as host key arrays ((min<<32)|max, local ids) and vertex counts."""
        sizes = np.array([np.asarray(k).size for k in keys_list], np.uint64)
        off = np.zeros(len(keys_list) + 1, np.uint64)
        off[1:] = np.cumsum(sizes)
        keys = np.ascontiguousarray(np.concatenate([np.asarray(k, np.uint64) for k in keys_list])
                                    if len(keys_list) else np.zeros(0, np.uint64))
        nv = np.ascontiguousarray(n_list, np.uint64)
        h = _P()
        _check(self.lib.tribit_batch_from_host(self.h, keys.ctypes.data if keys.size else None, off.ctypes.data,
                                               nv.ctypes.data, len(keys_list), ctypes.byref(h)), "batch_from_host")
        return Batch(self, h)

    def batch_gen_chung_lu(self, seed: int, first: int, count: int) -> Batch:
        h = _P()
        _check(self.lib.tribit_batch_gen_chung_lu(self.h, seed, first, count, ctypes.byref(h)), "batch_gen_chung_lu")
        return Batch(self, h)

    def set_batch_part_keys(self, nkeys: int) -> None:
        _check(self.lib.tribit_set_option(self.h, OPT_BATCH_PART_KEYS, nkeys), "set_option")

    def set_relabel(self, on) -> None:
        """Relabelling of graphs built from now on: True / 1 degree rank, 2 octave classes of the degree
        (keeps runs of consecutive ids of similar degree together), 3 / "auto" (default) 2 when the edges
        joining ids less than 512 apart fill at least 1/16 of such id pairs, else 1; False / 0 input ids."""
        if on == "auto":
            on = 3
        mode = 1 if on is True else 0 if on is False else int(on)
        _check(self.lib.tribit_set_option(self.h, OPT_RELABEL, mode), "set_option")

    def set_alloc_cap(self, nbytes: int) -> None:
        _check(self.lib.tribit_set_alloc_cap(self.h, nbytes), "set_alloc_cap")

    def pool_peak(self) -> int:
        v = _U64()
        _check(self.lib.tribit_pool_peak(self.h, ctypes.byref(v)), "pool_peak")
        return v.value

    def graph_from_keys(self, keys: np.ndarray, n: int) -> Graph:
        k = np.ascontiguousarray(keys, dtype=np.uint64)
        g = _P()
        _check(self.lib.tribit_graph_from_host_keys(self.h, k.ctypes.data if k.size else None, k.size, n,
                                                    ctypes.byref(g)), "graph_from_host_keys")
        return Graph(self, g.value)

    def gen_rmat(self, scale: int, ef: int, a=0.57, b=0.19, c=0.19, seed: int = 1) -> Graph:
        d = _P()
        nk = _U64()
        _check(self.lib.tribit_gen_rmat(self.h, scale, ef, a, b, c, seed, ctypes.byref(d), ctypes.byref(nk)),
               "gen_rmat")
        g = _P()
        _check(self.lib.tribit_graph_from_device_keys(self.h, d, nk.value, 1 << scale, ctypes.byref(g)),
               "graph_from_device_keys")
        return Graph(self, g.value)

    def gen_rmat_keys(self, scale: int, ef: int, a=0.57, b=0.19, c=0.19, seed: int = 1) -> np.ndarray:
        """The R-MAT generator's input keys ((min << 32) | max, self-loops as the all-ones sentinel),
        generated on the GPU and copied to a host array."""
        d = _P()
        nk = _U64()
        _check(self.lib.tribit_gen_rmat(self.h, scale, ef, a, b, c, seed, ctypes.byref(d), ctypes.byref(nk)),
               "gen_rmat")
        out = np.empty(nk.value, np.uint64)
        try:
            _check(self.lib.tribit_device_keys_to_host(self.h, d, nk.value, out.ctypes.data), "device_keys_to_host")
        finally:
            self.lib.tribit_free_device_keys(self.h, d)
        return out

    def gen_er(self, n: int, nsamples: int, seed: int) -> Graph:
        d = _P()
        nk = _U64()
        _check(self.lib.tribit_gen_er(self.h, n, nsamples, seed, ctypes.byref(d), ctypes.byref(nk)), "gen_er")
        g = _P()
        _check(self.lib.tribit_graph_from_device_keys(self.h, d, nk.value, n, ctypes.byref(g)),
               "graph_from_device_keys")
        return Graph(self, g.value)

    def _from_generated(self, st: int, d, nk, n: int, what: str) -> Graph:
        _check(st, what)
        g = _P()
        _check(self.lib.tribit_graph_from_device_keys(self.h, d, nk.value, n, ctypes.byref(g)),
               "graph_from_device_keys")
        return Graph(self, g.value)

    def gen_chung_lu(self, n: int, gamma: float, avg_deg: float, seed: int) -> Graph:
        d, nk = _P(), _U64()
        st = self.lib.tribit_gen_chung_lu(self.h, n, gamma, avg_deg, seed, ctypes.byref(d), ctypes.byref(nk))
        return self._from_generated(st, d, nk, n, "gen_chung_lu")

    def gen_rgg(self, n: int, avg_deg: float, seed: int) -> Graph:
        d, nk = _P(), _U64()
        st = self.lib.tribit_gen_rgg(self.h, n, avg_deg, seed, ctypes.byref(d), ctypes.byref(nk))
        return self._from_generated(st, d, nk, n, "gen_rgg")

    def gen_sbm(self, blocks: int, size: int, p_in: float, p_out: float, seed: int) -> Graph:
        d, nk = _P(), _U64()
        st = self.lib.tribit_gen_sbm(self.h, blocks, size, p_in, p_out, seed, ctypes.byref(d), ctypes.byref(nk))
        return self._from_generated(st, d, nk, blocks * size, "gen_sbm")

    def count_keys(self, keys: np.ndarray, n: int, sched: str = "plain") -> tuple[int, float]:
        k = np.ascontiguousarray(keys, dtype=np.uint64)
        o = CountOpts()
        o.sched = SCHED[sched]
        o.repeats = 1
        tri = _U64()
        msp = ctypes.c_double()
        ci = CountInfo()
        _check(self.lib.tribit_count_keys(self.h, k.ctypes.data if k.size else None, k.size, n, ctypes.byref(o),
                                          ctypes.byref(tri), ctypes.byref(msp), ctypes.byref(ci)), "count_keys")
        return tri.value, msp.value
