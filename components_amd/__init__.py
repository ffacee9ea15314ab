"""Python view of libkacomp's C ABI (include/kacomp.h), used by the tests and
bench.py.  The compute path is the HIP library; this module only moves arrays
across the boundary with ctypes.  Loading fails loudly when the library has
not been built: there is no fallback.

The class ``ExponentialContraction`` mirrors the reference's component API
(lib/components/exponential_contraction.h:42-60, 325-335):
``ExponentialContraction(conf, rank, size).FindComponents(graph, labels)``.
"""
from __future__ import annotations

import ctypes
import os
from dataclasses import dataclass, field

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "lib", "libkacomp.so")
BIN_DIR = os.path.join(_HERE, "bin")

KC_OK = 0
KC_ERR_INVALID = -1
KC_ERR_HIP = -2
KC_ERR_OOM = -3
KC_ERR_IO = -4
KC_ERR_UNSUPPORTED = -5

VARIANT_EXP, VARIANT_LOCAL, VARIANT_SHORTCUT, VARIANT_LABELS = 0, 1, 2, 3
VARIANTS = {"exp": 0, "local": 1, "shortcut": 2, "labels": 3}

KERNEL_CLASSES = ["deviates", "sweep", "bin", "count", "scan", "scatter", "dedup", "csr", "undo", "canon",
                  "jump", "other", "sort", "relabel", "round", "learned"]
_NK = len(KERNEL_CLASSES)


class CsrHost(ctypes.Structure):
    _fields_ = [("n", ctypes.c_uint64), ("offsets", ctypes.POINTER(ctypes.c_uint64)),
                ("targets", ctypes.POINTER(ctypes.c_uint32))]


class Csr(ctypes.Structure):
    _fields_ = [("n", ctypes.c_uint64), ("m", ctypes.c_uint64), ("offsets", ctypes.POINTER(ctypes.c_uint64)),
                ("targets", ctypes.POINTER(ctypes.c_uint32))]


class Options(ctypes.Structure):
    _fields_ = [("variant", ctypes.c_int32), ("seed", ctypes.c_int32), ("beta", ctypes.c_double),
                ("seq_limit", ctypes.c_uint64), ("use_bfs", ctypes.c_int32), ("deviate_fma", ctypes.c_int32),
                ("learned", ctypes.c_int32), ("record_trace", ctypes.c_int32), ("time_kernels", ctypes.c_int32),
                ("max_levels", ctypes.c_int32), ("sweep_group_small", ctypes.c_int32),
                ("sweep_group_mid", ctypes.c_int32), ("sweep_group_large", ctypes.c_int32),
                ("deviate_stream", ctypes.c_int32), ("learned_alpha", ctypes.c_float),
                ("input_order", ctypes.c_int32), ("tail_sweeps", ctypes.c_int32),
                ("pull_rounds", ctypes.c_int32),
                ("dedup_huge", ctypes.c_int32), ("time_class", ctypes.c_int32),
                ("pre_contract", ctypes.c_int32), ("pre_blocks", ctypes.c_uint32),
                ("replicate", ctypes.c_int32), ("hd_threshold", ctypes.c_double),
                ("relabel_split", ctypes.c_int32)]


class Stats(ctypes.Structure):
    _fields_ = [("time_ms", ctypes.c_double), ("total_ms", ctypes.c_double), ("comm_time_s", ctypes.c_double),
                ("send_volume", ctypes.c_uint64), ("recv_volume", ctypes.c_uint64), ("n", ctypes.c_uint64),
                ("m", ctypes.c_uint64), ("levels", ctypes.c_uint32), ("components", ctypes.c_uint32),
                ("sweeps", ctypes.c_uint64), ("radius_sum", ctypes.c_uint64), ("alg_bytes", ctypes.c_uint64),
                ("kernel_ms", ctypes.c_double * _NK), ("kernel_launches", ctypes.c_uint64 * _NK),
                ("kernel_bytes", ctypes.c_uint64 * _NK), ("replicas", ctypes.c_uint64),
                ("bin_small", ctypes.c_uint32), ("bin_mid", ctypes.c_uint32), ("bin_large", ctypes.c_uint32),
                ("hubs", ctypes.c_uint32), ("kernel_first_ms", ctypes.c_double * _NK),
                ("kernel_first_bytes", ctypes.c_uint64 * _NK)]


class BinTuning(ctypes.Structure):
    _fields_ = [("small", ctypes.c_int32), ("mid", ctypes.c_int32), ("large", ctypes.c_int32), ("ms", ctypes.c_double),
                ("default_ms", ctypes.c_double), ("candidates", ctypes.c_int32)]


class GenParams(ctypes.Structure):
    _fields_ = [("n", ctypes.c_uint64), ("m", ctypes.c_uint64), ("r", ctypes.c_double), ("p", ctypes.c_double),
                ("gamma", ctypes.c_double), ("d", ctypes.c_double), ("periodic", ctypes.c_int32),
                ("seed", ctypes.c_int32), ("rmat_a", ctypes.c_double), ("rmat_b", ctypes.c_double),
                ("rmat_c", ctypes.c_double), ("nx", ctypes.c_uint64), ("ny", ctypes.c_uint64),
                ("nz", ctypes.c_uint64)]


class BatchHost(ctypes.Structure):
    _fields_ = [("graphs", ctypes.c_uint32), ("vtx_off", ctypes.POINTER(ctypes.c_uint64)),
                ("edge_off", ctypes.POINTER(ctypes.c_uint64)), ("offsets", ctypes.POINTER(ctypes.c_uint64)),
                ("targets", ctypes.POINTER(ctypes.c_uint32))]


class BatchCsr(ctypes.Structure):
    _fields_ = [("graphs", ctypes.c_uint32), ("vtx_off", ctypes.POINTER(ctypes.c_uint64)),
                ("edge_off", ctypes.POINTER(ctypes.c_uint64)), ("offsets", ctypes.POINTER(ctypes.c_uint64)),
                ("targets", ctypes.POINTER(ctypes.c_uint32))]


class BatchGenParams(ctypes.Structure):
    _fields_ = [("graphs", ctypes.c_uint32), ("log_n_min", ctypes.c_double), ("log_n_max", ctypes.c_double),
                ("deg_min", ctypes.c_double), ("deg_max", ctypes.c_double), ("types", ctypes.c_int32),
                ("seed", ctypes.c_int32)]


class BatchTuning(ctypes.Structure):
    _fields_ = [("block", ctypes.c_int32), ("groups_per_cu", ctypes.c_int32), ("use_graph", ctypes.c_int32),
                ("split_vertices", ctypes.c_uint32)]


class BatchStats(ctypes.Structure):
    _fields_ = [("time_ms", ctypes.c_double), ("total_ms", ctypes.c_double), ("n", ctypes.c_uint64),
                ("m", ctypes.c_uint64), ("graphs", ctypes.c_uint32), ("max_levels", ctypes.c_uint32),
                ("levels_sum", ctypes.c_uint64), ("sweeps_sum", ctypes.c_uint64), ("radius_sum", ctypes.c_uint64),
                ("components", ctypes.c_uint64), ("tuning", BatchTuning)]


class TrainParams(ctypes.Structure):
    _fields_ = [("steps", ctypes.c_int32), ("batch", ctypes.c_int32), ("lr", ctypes.c_float),
                ("beta1", ctypes.c_float), ("beta2", ctypes.c_float), ("eps", ctypes.c_float),
                ("seed", ctypes.c_int32)]


class TrainStats(ctypes.Structure):
    _fields_ = [("first_loss", ctypes.c_double), ("last_loss", ctypes.c_double), ("time_ms", ctypes.c_double),
                ("samples", ctypes.c_uint64), ("steps", ctypes.c_int32)]


MLP_PARAMS = 4673
MLP_FEATURES = 6


# Every symbol include/kacomp.h declares, with (restype, argtypes).
_P = ctypes.POINTER
_SIGNATURES = {
    "kc_last_error": (ctypes.c_char_p, []),
    "kc_options_init": (None, [_P(Options)]),
    "kc_graph_create": (ctypes.c_int, [_P(CsrHost), _P(ctypes.c_void_p)]),
    "kc_graph_destroy": (ctypes.c_int, [ctypes.c_void_p]),
    "kc_graph_info": (ctypes.c_int, [ctypes.c_void_p, _P(ctypes.c_uint64), _P(ctypes.c_uint64), _P(ctypes.c_uint64)]),
    "kc_find_components": (ctypes.c_int, [ctypes.c_void_p, _P(Options), _P(ctypes.c_uint32), _P(Stats)]),
    "kc_find_components_noalloc": (ctypes.c_int, [ctypes.c_uint64, ctypes.c_void_p, ctypes.c_void_p, _P(Options),
                                                  ctypes.c_void_p, ctypes.c_void_p, _P(Stats)]),
    "kc_copy_labels": (ctypes.c_int, [ctypes.c_void_p, _P(ctypes.c_uint32)]),
    "kc_graph_autotune": (ctypes.c_int, [ctypes.c_void_p, _P(Options), _P(BinTuning)]),
    "kc_graph_set_bins": (ctypes.c_int, [ctypes.c_void_p, _P(BinTuning)]),
    "kc_component_histogram": (ctypes.c_int, [ctypes.c_void_p, _P(ctypes.c_uint64), ctypes.c_uint64,
                                              _P(ctypes.c_uint64)]),
    "kc_trace_levels": (ctypes.c_int, [ctypes.c_void_p, _P(ctypes.c_uint32)]),
    "kc_trace_level_info": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_uint32, _P(ctypes.c_uint64)]),
    "kc_trace_level_keys": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_uint32, _P(ctypes.c_uint64), _P(ctypes.c_uint32)]),
    "kc_deviates": (ctypes.c_int, [ctypes.c_uint32, ctypes.c_uint32, ctypes.c_int32, ctypes.c_uint32, ctypes.c_uint64,
                                   ctypes.c_double, ctypes.c_int32, ctypes.c_int32, _P(ctypes.c_uint32)]),
    "kc_read_metis": (ctypes.c_int, [ctypes.c_char_p, _P(Csr)]),
    "kc_write_metis": (ctypes.c_int, [ctypes.c_char_p, _P(Csr)]),
    "kc_read_edgefile": (ctypes.c_int, [ctypes.c_char_p, _P(Csr)]),
    "kc_read_partitioned_edgefile": (ctypes.c_int, [ctypes.c_char_p, ctypes.c_int, _P(Csr)]),
    "kc_read_binary": (ctypes.c_int, [ctypes.c_char_p, _P(Csr)]),
    "kc_write_binary": (ctypes.c_int, [ctypes.c_char_p, _P(Csr)]),
    "kc_gen_params_init": (None, [_P(GenParams)]),
    "kc_generate": (ctypes.c_int, [ctypes.c_char_p, _P(GenParams), _P(Csr)]),
    "kc_csr_free": (None, [_P(Csr)]),
    "kc_graph_generate": (ctypes.c_int, [ctypes.c_char_p, _P(GenParams), _P(ctypes.c_void_p)]),
    "kc_graph_download": (ctypes.c_int, [ctypes.c_void_p, _P(Csr)]),
    "kc_graph_read_binary": (ctypes.c_int, [ctypes.c_char_p, _P(ctypes.c_void_p)]),
    "kc_graph_read_metis": (ctypes.c_int, [ctypes.c_char_p, _P(ctypes.c_void_p)]),
    "kc_graph_read_edgefile": (ctypes.c_int, [ctypes.c_char_p, _P(ctypes.c_void_p)]),
    "kc_graph_read_partitioned_edgefile": (ctypes.c_int, [ctypes.c_char_p, ctypes.c_int, _P(ctypes.c_void_p)]),
    "kc_batch_gen_params_init": (None, [_P(BatchGenParams)]),
    "kc_batch_generate": (ctypes.c_int, [_P(BatchGenParams), _P(BatchCsr)]),
    "kc_batch_free": (None, [_P(BatchCsr)]),
    "kc_batch_create": (ctypes.c_int, [_P(BatchHost), _P(ctypes.c_void_p)]),
    "kc_batch_destroy": (ctypes.c_int, [ctypes.c_void_p]),
    "kc_batch_autotune": (ctypes.c_int, [ctypes.c_void_p, _P(Options), _P(BatchTuning)]),
    "kc_batch_set_tuning": (ctypes.c_int, [ctypes.c_void_p, _P(BatchTuning)]),
    "kc_batch_run": (ctypes.c_int, [ctypes.c_void_p, _P(Options), _P(ctypes.c_uint32), _P(BatchStats)]),
    "kc_batch_graph_stats": (ctypes.c_int, [ctypes.c_void_p, _P(ctypes.c_uint32), _P(ctypes.c_uint32),
                                            _P(ctypes.c_uint32)]),
    "kc_batch_find_components": (ctypes.c_int, [_P(BatchHost), _P(Options), _P(ctypes.c_uint32), _P(BatchStats)]),
    "kc_mlp_create": (ctypes.c_int, [ctypes.c_int32, _P(ctypes.c_void_p)]),
    "kc_mlp_destroy": (ctypes.c_int, [ctypes.c_void_p]),
    "kc_mlp_get_weights": (ctypes.c_int, [ctypes.c_void_p, _P(ctypes.c_float)]),
    "kc_mlp_set_weights": (ctypes.c_int, [ctypes.c_void_p, _P(ctypes.c_float)]),
    "kc_mlp_forward": (ctypes.c_int, [ctypes.c_void_p, _P(ctypes.c_float), ctypes.c_uint64, _P(ctypes.c_float)]),
    "kc_mlp_gradients": (ctypes.c_int, [ctypes.c_void_p, _P(ctypes.c_float), _P(ctypes.c_float), ctypes.c_uint64,
                                        _P(ctypes.c_float), _P(ctypes.c_float)]),
    "kc_mlp_collect": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, _P(Options), ctypes.c_uint64]),
    "kc_mlp_dataset_size": (ctypes.c_int, [ctypes.c_void_p, _P(ctypes.c_uint64)]),
    "kc_mlp_clear_dataset": (ctypes.c_int, [ctypes.c_void_p]),
    "kc_mlp_add_dataset": (ctypes.c_int, [ctypes.c_void_p, _P(ctypes.c_float), _P(ctypes.c_float), ctypes.c_uint64]),
    "kc_mlp_get_dataset": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_uint64, ctypes.c_uint64, _P(ctypes.c_float),
                                          _P(ctypes.c_float)]),
    "kc_train_params_init": (None, [_P(TrainParams)]),
    "kc_mlp_train": (ctypes.c_int, [ctypes.c_void_p, _P(TrainParams), _P(TrainStats)]),
    "kc_graph_set_mlp": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p]),
    "kc_graph_features": (ctypes.c_int, [ctypes.c_void_p, _P(ctypes.c_float)]),
}

_lib = None


class KcError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"libkacomp error {code}: {msg}")
        self.code = code


def lib():
    """Loads libkacomp.so (raises when it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is missing: run __graft_entry__.build() (make -C components_amd)")
        L = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in _SIGNATURES.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def _check(rc):
    if rc != KC_OK:
        raise KcError(rc, lib().kc_last_error().decode())


def _u64p(a):
    return a.ctypes.data_as(_P(ctypes.c_uint64))


def _u32p(a):
    return a.ctypes.data_as(_P(ctypes.c_uint32))


@dataclass
class Graph:
    """Host CSR: offsets u64[n+1], targets u32[2m]."""
    offsets: np.ndarray
    targets: np.ndarray

    @property
    def n(self):
        return len(self.offsets) - 1

    @property
    def m(self):
        return int(self.offsets[-1]) // 2

    @staticmethod
    def from_edges(n, edges):
        """Symmetric CSR (no self-loops, no duplicates) from an (k,2) edge array."""
        e = np.asarray(edges, dtype=np.int64).reshape(-1, 2)
        e = e[e[:, 0] != e[:, 1]]
        both = np.concatenate([e, e[:, ::-1]]) if len(e) else np.zeros((0, 2), np.int64)
        key = np.unique(both[:, 0] * (n + 1) + both[:, 1]) if len(both) else np.zeros(0, np.int64)
        src = key // (n + 1)
        dst = key % (n + 1)
        off = np.zeros(n + 1, np.uint64)
        np.add.at(off, src + 1, 1)
        off = np.cumsum(off, dtype=np.uint64)
        return Graph(off, dst.astype(np.uint32))


def _take_csr(c):
    n = int(c.n)
    off = np.ctypeslib.as_array(c.offsets, shape=(n + 1,)).copy()
    m2 = int(off[-1])
    tgt = np.ctypeslib.as_array(c.targets, shape=(max(m2, 1),))[:m2].copy() if m2 else np.zeros(0, np.uint32)
    lib().kc_csr_free(ctypes.byref(c))
    return Graph(off, tgt)


def gen_params(**kw):
    p = GenParams()
    lib().kc_gen_params_init(ctypes.byref(p))
    for k, v in kw.items():
        setattr(p, k, v)
    return p


def generate(name, **kw):
    """kc_generate with KaGen-style parameters (n, m, r, p, gamma, d, periodic, seed, rmat_*, nx, ny, nz)."""
    c = Csr()
    _check(lib().kc_generate(name.encode(), ctypes.byref(gen_params(**kw)), ctypes.byref(c)))
    return _take_csr(c)


def read_metis(path):
    c = Csr()
    _check(lib().kc_read_metis(str(path).encode(), ctypes.byref(c)))
    return _take_csr(c)


def write_metis(path, g):
    off = np.ascontiguousarray(g.offsets, np.uint64)
    tgt = np.ascontiguousarray(g.targets, np.uint32)
    c = Csr(g.n, g.m, _u64p(off), _u32p(tgt) if len(tgt) else _P(ctypes.c_uint32)())
    _check(lib().kc_write_metis(str(path).encode(), ctypes.byref(c)))


def read_edgefile(path):
    """Sorted edge file ("<type> n m" header, "<type> u v" lines, 1-based)."""
    c = Csr()
    _check(lib().kc_read_edgefile(str(path).encode(), ctypes.byref(c)))
    return _take_csr(c)


def read_partitioned_edgefile(path, parts=1):
    """The edge files "<path>_0" .. "<path>_<parts-1>" as one graph."""
    c = Csr()
    _check(lib().kc_read_partitioned_edgefile(str(path).encode(), int(parts), ctypes.byref(c)))
    return _take_csr(c)


def read_binary(path):
    """Sorted binary file (n, m as u64, then m 1-based (u, v) u64 pairs)."""
    c = Csr()
    _check(lib().kc_read_binary(str(path).encode(), ctypes.byref(c)))
    return _take_csr(c)


def write_binary(path, g):
    off = np.ascontiguousarray(g.offsets, np.uint64)
    tgt = np.ascontiguousarray(g.targets, np.uint32)
    c = Csr(g.n, g.m, _u64p(off), _u32p(tgt) if len(tgt) else _P(ctypes.c_uint32)())
    _check(lib().kc_write_binary(str(path).encode(), ctypes.byref(c)))


STREAM_REFERENCE, STREAM_HASHED = 0, 1
# dedup of the huge cluster-edge buckets (kc_options.dedup_huge)
DEDUP_MODES = {"auto": 0, "tables": 1, "bitmaps": 2, "bytemaps": 3, "maps_tables": 4}
# kc_options.pre_contract: none, the default branch's local contraction, -single, -contraction
PRE_MODES = {"none": 0, "local": 1, "single": 2, "double": 3}


def options(variant="exp", seed=1, beta=0.1, seq_limit=4096, fma=True, record_trace=False, time_kernels=False,
            tmax=0, smax=0, wmax=0, use_bfs=False, max_levels=4096, stream=STREAM_REFERENCE, learned=False,
            alpha=1.0, input_order=False, tail_sweeps=0, pull_rounds=0, dedup_huge="auto", time_class=None,
            pre_contract="none", pre_blocks=1, replicate=False, hd_threshold=16.0, relabel_split=0):
    o = Options()
    lib().kc_options_init(ctypes.byref(o))
    o.variant = VARIANTS[variant] if isinstance(variant, str) else int(variant)
    o.seed = seed
    o.beta = beta
    o.seq_limit = seq_limit
    o.deviate_fma = 1 if fma else 0
    o.record_trace = 1 if record_trace else 0
    o.time_kernels = 1 if time_kernels else 0
    o.sweep_group_small, o.sweep_group_mid, o.sweep_group_large = tmax, smax, wmax
    o.use_bfs = 1 if use_bfs else 0
    o.max_levels = max_levels
    o.deviate_stream = stream
    o.learned = 1 if learned else 0
    o.learned_alpha = alpha
    o.input_order = 1 if input_order else 0
    o.tail_sweeps = int(tail_sweeps)
    o.pull_rounds = int(pull_rounds)
    o.dedup_huge = DEDUP_MODES[dedup_huge] if isinstance(dedup_huge, str) else int(dedup_huge)
    # with time_kernels: events on the launches of one class only (a name of KERNEL_CLASSES)
    o.time_class = 0 if time_class is None else KERNEL_CLASSES.index(time_class) + 1
    # local components over pre_blocks blocks of consecutive ids contracted first (a name of PRE_MODES)
    o.pre_contract = PRE_MODES[pre_contract] if isinstance(pre_contract, str) else int(pre_contract)
    o.pre_blocks = int(pre_blocks)
    o.replicate = 1 if replicate else 0
    o.hd_threshold = float(hd_threshold)
    # relabel of the degree-ordered level 0: 1 owner-split, -1 one pass, 0 default
    o.relabel_split = int(relabel_split)
    return o


def stats_dict(st):
    d = {k: getattr(st, k) for k, _ in Stats._fields_ if not k.startswith("kernel_")}
    d["kernel_ms"] = dict(zip(KERNEL_CLASSES, list(st.kernel_ms)))
    d["kernel_launches"] = dict(zip(KERNEL_CLASSES, list(st.kernel_launches)))
    d["kernel_bytes"] = dict(zip(KERNEL_CLASSES, list(st.kernel_bytes)))
    d["kernel_first_ms"] = dict(zip(KERNEL_CLASSES, list(st.kernel_first_ms)))
    d["kernel_first_bytes"] = dict(zip(KERNEL_CLASSES, list(st.kernel_first_bytes)))
    return d


class DeviceGraph:
    """kc_graph handle: the CSR copied once to HBM."""

    def __init__(self, g: Graph = None, _handle=None):
        self.h = ctypes.c_void_p()
        if _handle is not None:
            self.h = _handle
        else:
            off = np.ascontiguousarray(g.offsets, np.uint64)
            tgt = np.ascontiguousarray(g.targets, np.uint32)
            if len(tgt) == 0:
                tgt = np.zeros(1, np.uint32)
            h = CsrHost(g.n, _u64p(off), _u32p(tgt))
            _check(lib().kc_graph_create(ctypes.byref(h), ctypes.byref(self.h)))
        n, m, ni = ctypes.c_uint64(), ctypes.c_uint64(), ctypes.c_uint64()
        _check(lib().kc_graph_info(self.h, ctypes.byref(n), ctypes.byref(m), ctypes.byref(ni)))
        self.n, self.m, self.non_isolated = n.value, m.value, ni.value

    @staticmethod
    def generate(name, **kw):
        """kc_graph_generate: the graph is generated and built in HBM."""
        h = ctypes.c_void_p()
        _check(lib().kc_graph_generate(name.encode(), ctypes.byref(gen_params(**kw)), ctypes.byref(h)))
        return DeviceGraph(_handle=h)

    @staticmethod
    def read_binary(path):
        """kc_graph_read_binary: a sorted binary graph file parsed and built in HBM."""
        h = ctypes.c_void_p()
        _check(lib().kc_graph_read_binary(str(path).encode(), ctypes.byref(h)))
        return DeviceGraph(_handle=h)

    @staticmethod
    def read_metis(path):
        """kc_graph_read_metis: a METIS file parsed and built in HBM."""
        h = ctypes.c_void_p()
        _check(lib().kc_graph_read_metis(str(path).encode(), ctypes.byref(h)))
        return DeviceGraph(_handle=h)

    @staticmethod
    def read_edgefile(path, parts=None):
        """kc_graph_read_edgefile (parts None) or kc_graph_read_partitioned_edgefile: edge-list text
        parsed and built in HBM."""
        h = ctypes.c_void_p()
        if parts is None:
            _check(lib().kc_graph_read_edgefile(str(path).encode(), ctypes.byref(h)))
        else:
            _check(lib().kc_graph_read_partitioned_edgefile(str(path).encode(), int(parts), ctypes.byref(h)))
        return DeviceGraph(_handle=h)

    def download(self):
        """Level-0 CSR copied back to the host (rows unsorted for device-built graphs)."""
        c = Csr()
        _check(lib().kc_graph_download(self.h, ctypes.byref(c)))
        return _take_csr(c)

    def close(self):
        if self.h:
            lib().kc_graph_destroy(self.h)
            self.h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def find_components(self, copy_labels=True, **kw):
        o = options(**kw)
        st = Stats()
        labels = np.empty(max(self.n, 1), np.uint32)
        _check(lib().kc_find_components(self.h, ctypes.byref(o), _u32p(labels) if copy_labels else None,
                                        ctypes.byref(st)))
        return (labels[: self.n] if copy_labels else None), stats_dict(st)

    def autotune(self, **kw):
        """kc_graph_autotune: times the degree-bin threshold candidates with these options and keeps
        the fastest in the handle; returns the choice."""
        o = options(**kw)
        t = BinTuning()
        _check(lib().kc_graph_autotune(self.h, ctypes.byref(o), ctypes.byref(t)))
        return {"small": t.small, "mid": t.mid, "large": t.large, "ms": t.ms, "default_ms": t.default_ms,
                "candidates": t.candidates}

    def set_bins(self, small=None, mid=None, large=None):
        """Thresholds kept in the handle (all None: cleared)."""
        if small is None:
            _check(lib().kc_graph_set_bins(self.h, None))
        else:
            t = BinTuning(small, mid, large, 0.0, 0.0, 0)
            _check(lib().kc_graph_set_bins(self.h, ctypes.byref(t)))

    def labels(self):
        labels = np.empty(max(self.n, 1), np.uint32)
        _check(lib().kc_copy_labels(self.h, _u32p(labels)))
        return labels[: self.n]

    def component_histogram(self):
        """(size, number of components of that size) pairs of the last run, increasing size."""
        np_ = ctypes.c_uint64()
        _check(lib().kc_component_histogram(self.h, None, 0, ctypes.byref(np_)))
        out = np.empty(2 * max(np_.value, 1), np.uint64)
        _check(lib().kc_component_histogram(self.h, _u64p(out), np_.value, ctypes.byref(np_)))
        return out[: 2 * np_.value].reshape(-1, 2)

    def set_mlp(self, mlp):
        """Model used when find_components(learned=True)."""
        _check(lib().kc_graph_set_mlp(self.h, mlp.h if mlp is not None else None))
        self._mlp = mlp

    def features(self):
        """Level-0 features (vertices of degree > 0 in id order), non_isolated x 6."""
        X = np.empty((max(self.non_isolated, 1), MLP_FEATURES), np.float32)
        _check(lib().kc_graph_features(self.h, X.ctypes.data_as(_P(ctypes.c_float))))
        return X[: self.non_isolated]

    def trace(self):
        """Per-level (info, keys, gid) of the last run with record_trace=True."""
        nl = ctypes.c_uint32()
        _check(lib().kc_trace_levels(self.h, ctypes.byref(nl)))
        out = []
        for lvl in range(nl.value):
            info = (ctypes.c_uint64 * 4)()
            _check(lib().kc_trace_level_info(self.h, lvl, info))
            n = int(info[0])
            keys = np.empty(max(n, 1), np.uint64)
            gid = np.empty(max(n, 1), np.uint32)
            _check(lib().kc_trace_level_keys(self.h, lvl, _u64p(keys), _u32p(gid)))
            out.append((tuple(int(x) for x in info), keys[:n], gid[:n]))
        return out


def deviates(gid0, count, seed, iteration, rng_offset, beta=0.1, fma=True, stream=STREAM_REFERENCE):
    out = np.empty(max(count, 1), np.uint32)
    _check(lib().kc_deviates(gid0, count, seed, iteration, rng_offset, beta, 1 if fma else 0, stream, _u32p(out)))
    return out[:count]


NOALLOC_MAX_VERTICES = 1 << 20
NOALLOC_MAX_EDGE_SLOTS = 1 << 24


def find_components_noalloc(n, offsets_ptr, targets_ptr, labels_ptr, stream=None, **kw):
    """kc_find_components_noalloc: the labels or shortcut variant on a CSR already in device
    memory (offsets, targets and labels are device addresses the caller owns; stream a
    hipStream_t handle or None), with no allocation by the library.  Returns the stats dict."""
    o = options(**kw)
    st = Stats()
    _check(lib().kc_find_components_noalloc(n, offsets_ptr, targets_ptr, ctypes.byref(o), labels_ptr, stream,
                                            ctypes.byref(st)))
    return stats_dict(st)


def find_components(g: Graph, **kw):
    dg = DeviceGraph(g)
    try:
        return dg.find_components(**kw)
    finally:
        dg.close()


class Config:
    """Subset of the reference's Config (lib/config.h:29-87) the components read."""

    def __init__(self, seed=1, beta=0.1, sequential_limit=1, fixed_limit=False, use_bfs=False):
        self.seed = seed
        self.beta = beta
        self.sequential_limit = sequential_limit
        self.fixed_limit = fixed_limit
        self.use_bfs = use_bfs


class _Component:
    variant = "exp"

    def __init__(self, conf: Config, rank: int = 0, size: int = 1):
        if size != 1:
            raise ValueError("one GPU holds the whole graph: size must be 1")
        self.conf = conf
        self.stats = None

    def FindComponents(self, g: Graph, labels: np.ndarray):
        c = self.conf
        seq = c.sequential_limit if c.fixed_limit else max((g.m // 256) * c.sequential_limit, 4096)
        out, self.stats = find_components(g, variant=self.variant, seed=c.seed, beta=c.beta, seq_limit=seq,
                                          use_bfs=c.use_bfs)
        labels[: g.n] = out

    def GetCommTime(self):
        return 0.0

    def GetSendVolume(self):
        return 0

    def GetReceiveVolume(self):
        return 0


class ExponentialContraction(_Component):
    variant = "exp"


class LocalContraction(_Component):
    variant = "local"


class ShortcutPropagation(_Component):
    variant = "shortcut"


class Propagation(_Component):
    variant = "labels"


# ---------------------------------------------------------------------------
# Batched form.
@dataclass
class BatchGraphs:
    """Packed CSR segments: graph g owns vertices vtx_off[g]:vtx_off[g+1] and
    edge slots edge_off[g]:edge_off[g+1]; offsets are global, targets local."""
    vtx_off: np.ndarray
    edge_off: np.ndarray
    offsets: np.ndarray
    targets: np.ndarray

    @property
    def graphs(self):
        return len(self.vtx_off) - 1

    def graph(self, g):
        """Graph g as a standalone host CSR with local ids."""
        a, z = int(self.vtx_off[g]), int(self.vtx_off[g + 1])
        e0, e1 = int(self.edge_off[g]), int(self.edge_off[g + 1])
        return Graph((self.offsets[a:z + 1] - np.uint64(e0)).astype(np.uint64), self.targets[e0:e1].copy())

    @staticmethod
    def pack(graphs):
        vtx = np.zeros(len(graphs) + 1, np.uint64)
        eo = np.zeros(len(graphs) + 1, np.uint64)
        for i, g in enumerate(graphs):
            vtx[i + 1] = vtx[i] + g.n
            eo[i + 1] = eo[i] + int(g.offsets[-1])
        offs = [np.asarray(g.offsets[:-1], np.uint64) + eo[i] for i, g in enumerate(graphs)]
        offsets = np.concatenate(offs + [np.array([eo[-1]], np.uint64)])
        targets = np.concatenate([np.asarray(g.targets, np.uint32) for g in graphs] + [np.zeros(0, np.uint32)])
        return BatchGraphs(vtx, eo, offsets, targets)


def batch_generate(graphs=4096, log_n_min=8.0, log_n_max=18.0, deg_min=2.0, deg_max=16.0, types=7, seed=4):
    """kc_batch_generate: a mixed batch of G(n,m), 2D geometric and percolated grids."""
    p = BatchGenParams()
    lib().kc_batch_gen_params_init(ctypes.byref(p))
    p.graphs, p.log_n_min, p.log_n_max = graphs, log_n_min, log_n_max
    p.deg_min, p.deg_max, p.types, p.seed = deg_min, deg_max, types, seed
    c = BatchCsr()
    _check(lib().kc_batch_generate(ctypes.byref(p), ctypes.byref(c)))
    G = int(c.graphs)
    vtx = np.ctypeslib.as_array(c.vtx_off, shape=(G + 1,)).copy()
    eo = np.ctypeslib.as_array(c.edge_off, shape=(G + 1,)).copy()
    n, m2 = int(vtx[-1]), int(eo[-1])
    off = np.ctypeslib.as_array(c.offsets, shape=(n + 1,)).copy()
    tgt = np.ctypeslib.as_array(c.targets, shape=(max(m2, 1),))[:m2].copy()
    lib().kc_batch_free(ctypes.byref(c))
    return BatchGraphs(vtx, eo, off, tgt)


def batch_stats_dict(st):
    d = {k: getattr(st, k) for k, _ in BatchStats._fields_ if k != "tuning"}
    t = st.tuning
    d["tuning"] = {"block": t.block, "groups_per_cu": t.groups_per_cu, "use_graph": t.use_graph,
                   "split_vertices": t.split_vertices}
    return d


class DeviceBatch:
    """kc_batch handle: the packed batch copied once to HBM."""

    def __init__(self, b: BatchGraphs):
        self.h = ctypes.c_void_p()
        self.graphs = b.graphs
        self.n = int(b.vtx_off[-1])
        self._keep = [np.ascontiguousarray(b.vtx_off, np.uint64), np.ascontiguousarray(b.edge_off, np.uint64),
                      np.ascontiguousarray(b.offsets, np.uint64), np.ascontiguousarray(b.targets, np.uint32)]
        if len(self._keep[3]) == 0:
            self._keep[3] = np.zeros(1, np.uint32)
        bh = BatchHost(self.graphs, *(_u64p(a) for a in self._keep[:3]), _u32p(self._keep[3]))
        _check(lib().kc_batch_create(ctypes.byref(bh), ctypes.byref(self.h)))
        self._keep = None

    def close(self):
        if self.h:
            lib().kc_batch_destroy(self.h)
            self.h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_tuning(self, block=256, groups_per_cu=4, use_graph=True, split_vertices=0):
        """split_vertices: graphs with more vertices go through the device-wide path as one union."""
        t = BatchTuning(block, groups_per_cu, 1 if use_graph else 0, split_vertices)
        _check(lib().kc_batch_set_tuning(self.h, ctypes.byref(t)))

    def autotune(self, **kw):
        t = BatchTuning()
        _check(lib().kc_batch_autotune(self.h, ctypes.byref(options(**kw)), ctypes.byref(t)))
        return {"block": t.block, "groups_per_cu": t.groups_per_cu, "use_graph": t.use_graph,
                "split_vertices": t.split_vertices}

    def run(self, copy_labels=True, **kw):
        st = BatchStats()
        labels = np.empty(max(self.n, 1), np.uint32)
        _check(lib().kc_batch_run(self.h, ctypes.byref(options(**kw)), _u32p(labels) if copy_labels else None,
                                  ctypes.byref(st)))
        return (labels[: self.n] if copy_labels else None), batch_stats_dict(st)

    def graph_stats(self):
        lv = np.empty(max(self.graphs, 1), np.uint32)
        sw = np.empty_like(lv)
        rd = np.empty_like(lv)
        _check(lib().kc_batch_graph_stats(self.h, _u32p(lv), _u32p(sw), _u32p(rd)))
        return lv[: self.graphs], sw[: self.graphs], rd[: self.graphs]


# ---------------------------------------------------------------------------
# Learned centre proposals.
def _f32p(a):
    return a.ctypes.data_as(_P(ctypes.c_float))


class Mlp:
    """kc_mlp handle: MLP 6 -> 64 -> 64 -> 1 (bf16 weights, f32 accumulation) in HBM."""

    def __init__(self, seed=1):
        self.h = ctypes.c_void_p()
        _check(lib().kc_mlp_create(seed, ctypes.byref(self.h)))

    def close(self):
        if self.h:
            lib().kc_mlp_destroy(self.h)
            self.h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def get_weights(self):
        w = np.empty(MLP_PARAMS, np.float32)
        _check(lib().kc_mlp_get_weights(self.h, _f32p(w)))
        return w

    def set_weights(self, w):
        w = np.ascontiguousarray(w, np.float32)
        assert len(w) == MLP_PARAMS
        _check(lib().kc_mlp_set_weights(self.h, _f32p(w)))

    def forward(self, X):
        X = np.ascontiguousarray(X, np.float32).reshape(-1, MLP_FEATURES)
        out = np.empty(max(len(X), 1), np.float32)
        _check(lib().kc_mlp_forward(self.h, _f32p(X), len(X), _f32p(out)))
        return out[: len(X)]

    def gradients(self, X, y):
        X = np.ascontiguousarray(X, np.float32).reshape(-1, MLP_FEATURES)
        y = np.ascontiguousarray(y, np.float32)
        g = np.empty(MLP_PARAMS, np.float32)
        loss = ctypes.c_float()
        _check(lib().kc_mlp_gradients(self.h, _f32p(X), _f32p(y), len(X), _f32p(g), ctypes.byref(loss)))
        return g, loss.value

    def collect(self, dg, stride=1, **kw):
        _check(lib().kc_mlp_collect(self.h, dg.h, ctypes.byref(options(**kw)), stride))

    def dataset_size(self):
        n = ctypes.c_uint64()
        _check(lib().kc_mlp_dataset_size(self.h, ctypes.byref(n)))
        return n.value

    def dataset(self, first=0, count=None):
        n = self.dataset_size()
        count = n - first if count is None else count
        X = np.empty((max(count, 1), MLP_FEATURES), np.float32)
        y = np.empty(max(count, 1), np.float32)
        _check(lib().kc_mlp_get_dataset(self.h, first, count, _f32p(X), _f32p(y)))
        return X[:count], y[:count]

    def add_dataset(self, X, y):
        X = np.ascontiguousarray(X, np.float32).reshape(-1, MLP_FEATURES)
        y = np.ascontiguousarray(y, np.float32)
        _check(lib().kc_mlp_add_dataset(self.h, _f32p(X), _f32p(y), len(X)))

    def clear_dataset(self):
        _check(lib().kc_mlp_clear_dataset(self.h))

    def train(self, steps=400, batch=4096, lr=3e-3, seed=1):
        p = TrainParams()
        lib().kc_train_params_init(ctypes.byref(p))
        p.steps, p.batch, p.lr, p.seed = steps, batch, lr, seed
        st = TrainStats()
        _check(lib().kc_mlp_train(self.h, ctypes.byref(p), ctypes.byref(st)))
        return {k: getattr(st, k) for k, _ in TrainStats._fields_}
