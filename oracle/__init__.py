"""ctypes view of the CPU oracle (oracle/build/liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg, to check results and as the CPU baseline.  The product
path never imports it.
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "build", "liboracle.so")
_P = ctypes.POINTER
_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is missing: run make -C oracle")
        L = ctypes.CDLL(LIB_PATH)
        u64, u32, i32, dbl = ctypes.c_uint64, ctypes.c_uint32, ctypes.c_int32, ctypes.c_double
        sig = {
            "or_canonical_labels": (None, [u64, _P(u64), _P(u32), _P(u32)]),
            "or_exp_deviate": (u32, [u32, i32, u32, u64, dbl, i32]),
            "or_std_deviate_nofma": (u64, [u32, i32, u32, u64, dbl]),
            "or_std_deviate_fma": (u64, [u32, i32, u32, u64, dbl]),
            "or_std_fma_compiled": (i32, []),
            "or_lp_fixpoint": (u64, [u64, _P(u64), _P(u32), _P(u64), u64, _P(u64)]),
            "or_lp_closed_form": (None, [u64, _P(u64), _P(u32), _P(u64), u64, _P(u64)]),
            "or_run": (ctypes.c_void_p, [u64, _P(u64), _P(u32), i32, i32, dbl, i32, i32, u64, i32, i32, u64, i32,
                                         dbl]),
            "or_replication": (None, [ctypes.c_void_p, _P(u64)]),
            "or_replicate": (u64, [u64, _P(u64), _P(u32), i32, dbl, _P(u64), _P(u32), _P(u32)]),
            "or_hashed_deviate": (u32, [u32, i32, u32, dbl]),
            "or_free": (None, [ctypes.c_void_p]),
            "or_labels": (None, [ctypes.c_void_p, _P(u32)]),
            "or_levels": (u32, [ctypes.c_void_p]),
            "or_level_info": (None, [ctypes.c_void_p, u32, _P(u64)]),
            "or_level_keys": (None, [ctypes.c_void_p, u32, _P(u64), _P(u32)]),
            "or_total_sweeps": (u64, [ctypes.c_void_p]),
            "or_seconds": (dbl, [ctypes.c_void_p]),
            "or_uniform_deviate_100": (u32, [u32, i32, u32]),
            "or_splitmix64_mix": (u64, [u64]),
            "or_hashed_uniform": (dbl, [u32, i32, u32]),
            "or_mlp_forward": (None, [_P(dbl), _P(ctypes.c_float), u64, i32, _P(dbl)]),
            "or_mlp_grad": (None, [_P(dbl), _P(ctypes.c_float), _P(ctypes.c_float), u64, i32, _P(dbl), _P(dbl)]),
            "or_level_features": (None, [u64, _P(u64), _P(u32), u32, _P(ctypes.c_float)]),
        }
        for k, (r, a) in sig.items():
            f = getattr(L, k)
            f.restype = r
            f.argtypes = a
        _lib = L
    return _lib


def _arrs(offsets, targets):
    off = np.ascontiguousarray(offsets, np.uint64)
    tgt = np.ascontiguousarray(targets, np.uint32)
    if len(tgt) == 0:
        tgt = np.zeros(1, np.uint32)
    return off, tgt


def canonical_labels(offsets, targets):
    off, tgt = _arrs(offsets, targets)
    n = len(off) - 1
    out = np.empty(max(n, 1), np.uint32)
    lib().or_canonical_labels(n, off.ctypes.data_as(_P(ctypes.c_uint64)), tgt.ctypes.data_as(_P(ctypes.c_uint32)),
                              out.ctypes.data_as(_P(ctypes.c_uint32)))
    return out[:n]


def lp_fixpoint(offsets, targets, key0, hop=1 << 32):
    off, tgt = _arrs(offsets, targets)
    k0 = np.ascontiguousarray(key0, np.uint64)
    out = np.empty(max(len(k0), 1), np.uint64)
    r = lib().or_lp_fixpoint(len(off) - 1, off.ctypes.data_as(_P(ctypes.c_uint64)),
                             tgt.ctypes.data_as(_P(ctypes.c_uint32)), k0.ctypes.data_as(_P(ctypes.c_uint64)), hop,
                             out.ctypes.data_as(_P(ctypes.c_uint64)))
    return out[: len(k0)], int(r)


def lp_closed_form(offsets, targets, key0, hop=1 << 32):
    off, tgt = _arrs(offsets, targets)
    k0 = np.ascontiguousarray(key0, np.uint64)
    out = np.empty(max(len(k0), 1), np.uint64)
    lib().or_lp_closed_form(len(off) - 1, off.ctypes.data_as(_P(ctypes.c_uint64)),
                            tgt.ctypes.data_as(_P(ctypes.c_uint32)), k0.ctypes.data_as(_P(ctypes.c_uint64)), hop,
                            out.ctypes.data_as(_P(ctypes.c_uint64)))
    return out[: len(k0)]


def exp_deviate(gid, seed, iteration, rng_offset, beta=0.1, fma=True):
    return int(lib().or_exp_deviate(gid, seed, iteration, rng_offset, beta, 1 if fma else 0))


def hashed_deviate(gid, seed, iteration, beta=0.1):
    return int(lib().or_hashed_deviate(gid, seed, iteration, beta))


def hashed_uniform(gid, seed, iteration):
    return float(lib().or_hashed_uniform(gid, seed, iteration))


def uniform_deviate_100(gid, seed, iteration):
    return int(lib().or_uniform_deviate_100(gid, seed, iteration))


def splitmix64_mix(z):
    return int(lib().or_splitmix64_mix(z))


def replicate(offsets, targets, seed, hd_threshold=16.0):
    """Level 0 of the graph with its hubs split into proxies: (offsets, targets, ids), a proxy's id
    being its hub's."""
    off, tgt = _arrs(offsets, targets)
    n = len(off) - 1
    nr = int(lib().or_replicate(n, off.ctypes.data_as(_P(ctypes.c_uint64)), tgt.ctypes.data_as(_P(ctypes.c_uint32)),
                                seed, float(hd_threshold), None, None, None))
    n0 = int(np.count_nonzero(np.diff(off.astype(np.int64))))
    m2 = int(off[-1]) + 2 * (nr - n0)
    o = np.empty(nr + 1, np.uint64)
    t = np.empty(max(m2, 1), np.uint32)
    gid = np.empty(max(nr, 1), np.uint32)
    lib().or_replicate(n, off.ctypes.data_as(_P(ctypes.c_uint64)), tgt.ctypes.data_as(_P(ctypes.c_uint32)), seed,
                       float(hd_threshold), o.ctypes.data_as(_P(ctypes.c_uint64)), t.ctypes.data_as(_P(ctypes.c_uint32)),
                       gid.ctypes.data_as(_P(ctypes.c_uint32)))
    return o, t[:m2], gid[:nr]


class Run:
    """Result of or_run: labels and per-level trace."""

    VARIANTS = {"exp": 0, "local": 1, "shortcut": 2, "labels": 3}
    PRE_MODES = {"none": 0, "local": 1, "single": 2, "double": 3}

    def __init__(self, offsets, targets, variant="exp", seed=1, beta=0.1, fma=True, seq_limit=4096, keep_keys=False,
                 stream=0, pre_contract="none", pre_blocks=1, replicate=False, hd_threshold=16.0):
        off, tgt = _arrs(offsets, targets)
        self.n = len(off) - 1
        v = self.VARIANTS[variant] if isinstance(variant, str) else int(variant)
        self.h = lib().or_run(self.n, off.ctypes.data_as(_P(ctypes.c_uint64)), tgt.ctypes.data_as(_P(ctypes.c_uint32)),
                              v, seed, beta, 1 if fma else 0, stream, seq_limit, 1 if keep_keys else 0,
                              self.PRE_MODES[pre_contract] if isinstance(pre_contract, str) else int(pre_contract),
                              pre_blocks, 1 if replicate else 0, float(hd_threshold))

    def labels(self):
        out = np.empty(max(self.n, 1), np.uint32)
        lib().or_labels(self.h, out.ctypes.data_as(_P(ctypes.c_uint32)))
        return out[: self.n]

    def levels(self):
        out = []
        for lvl in range(lib().or_levels(self.h)):
            info = (ctypes.c_uint64 * 4)()
            lib().or_level_info(self.h, lvl, info)
            out.append(tuple(int(x) for x in info))
        return out

    def level_keys(self, lvl):
        n = self.levels()[lvl][0]
        keys = np.empty(max(n, 1), np.uint64)
        gid = np.empty(max(n, 1), np.uint32)
        lib().or_level_keys(self.h, lvl, keys.ctypes.data_as(_P(ctypes.c_uint64)),
                            gid.ctypes.data_as(_P(ctypes.c_uint32)))
        return keys[:n], gid[:n]

    def sweeps(self):
        return int(lib().or_total_sweeps(self.h))

    def replication(self):
        """(hubs split, proxies added) by replicate."""
        out = (ctypes.c_uint64 * 2)()
        lib().or_replication(self.h, out)
        return int(out[0]), int(out[1])

    def seconds(self):
        return float(lib().or_seconds(self.h))

    def __del__(self):
        try:
            lib().or_free(self.h)
        except Exception:
            pass


MLP_PARAMS = 4673


def mlp_forward(w, X, bf16=True):
    """Scores of the rows of X (n x 6) for parameters w (public order), FP64."""
    w = np.ascontiguousarray(w, np.float64)
    X = np.ascontiguousarray(X, np.float32)
    out = np.empty(max(len(X), 1), np.float64)
    lib().or_mlp_forward(w.ctypes.data_as(_P(ctypes.c_double)), X.ctypes.data_as(_P(ctypes.c_float)), len(X),
                         1 if bf16 else 0, out.ctypes.data_as(_P(ctypes.c_double)))
    return out[: len(X)]


def mlp_grad(w, X, y, bf16=True):
    """(gradient in public order, loss) of 0.5 * mean((score - y)^2), FP64."""
    w = np.ascontiguousarray(w, np.float64)
    X = np.ascontiguousarray(X, np.float32)
    y = np.ascontiguousarray(y, np.float32)
    g = np.empty(MLP_PARAMS, np.float64)
    loss = ctypes.c_double()
    lib().or_mlp_grad(w.ctypes.data_as(_P(ctypes.c_double)), X.ctypes.data_as(_P(ctypes.c_float)),
                      y.ctypes.data_as(_P(ctypes.c_float)), len(X), 1 if bf16 else 0,
                      g.ctypes.data_as(_P(ctypes.c_double)), ctypes.byref(loss))
    return g, loss.value


def level_features(offsets, targets, round_index=0):
    off, tgt = _arrs(offsets, targets)
    n = len(off) - 1
    X = np.empty((max(n, 1), 6), np.float32)
    lib().or_level_features(n, off.ctypes.data_as(_P(ctypes.c_uint64)), tgt.ctypes.data_as(_P(ctypes.c_uint32)),
                            round_index, X.ctypes.data_as(_P(ctypes.c_float)))
    return X[:n]


MPI_EXP = os.path.join(_HERE, "build", "mpi_exp")
MPIRUN = "/opt/conda/bin/mpirun"


def write_csr_file(path, offsets, targets):
    """CSR file read by mpi_exp: n and 2m as u64, the offsets (u64[n+1]), the targets (u32[2m])."""
    off = np.ascontiguousarray(offsets, np.uint64)
    tgt = np.ascontiguousarray(targets, np.uint32)
    with open(path, "wb") as f:
        np.array([len(off) - 1, int(off[-1])], np.uint64).tofile(f)
        off.tofile(f)
        tgt[: int(off[-1])].tofile(f)


def mpi_available():
    return os.path.exists(MPI_EXP) and os.path.exists(MPIRUN)


def mpi_exp(csr_path, ranks, seed=1, beta=0.1, iterations=1, seq=1, fixed_seq=False, local_phase=True,
            labels_path=None, timeout=600):
    """Runs the MPI form of exponential contraction on `ranks` processes (mpirun -np ranks) over a
    CSR file; returns the GLOBAL RESULT fields (time is the mean over iterations of the maximum over
    the ranks, in seconds) and, with labels_path, the canonical labels."""
    import subprocess
    cmd = [MPIRUN, "-np", str(ranks), MPI_EXP, csr_path, "-seed", str(seed), "-beta", repr(beta), "-i", str(iterations),
           "-seq", str(seq)]
    if fixed_seq:
        cmd.append("-fixed-seq")
    if not local_phase:
        cmd.append("-no-local")
    if labels_path:
        cmd += ["-labels", labels_path]
    env = dict(os.environ)
    env.setdefault("OMP_NUM_THREADS", "1")
    r = subprocess.run(cmd, capture_output=True, text=True, timeout=timeout, env=env)
    if r.returncode != 0:
        raise RuntimeError(f"mpi_exp failed ({r.returncode}): {r.stderr[-2000:]}")
    line = [x for x in r.stdout.splitlines() if x.startswith("GLOBAL RESULT")][-1]
    res = {}
    for tok in line.split()[2:]:
        k, _, v = tok.partition("=")
        res.setdefault(k, v)
    out = {"time": float(res["time"]), "stddev": float(res["stddev"]), "ranks": int(res["ranks"]),
           "n": int(res["n"]), "m": int(res["m"]), "seq_limit": int(res["seq_limit"]), "levels": int(res["levels"]),
           "sweeps": int(res["sweeps"]), "iterations": int(res["iterations"])}
    if labels_path:
        out["labels"] = np.fromfile(labels_path, dtype=np.uint32)
    return out
