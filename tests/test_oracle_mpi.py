"""The MPI form of exponential contraction (the host-CPU baseline) against union-find and against
the single-process oracle's levels (CPU only; skipped where MPICH is absent)."""
import os

import numpy as np
import pytest

import components_amd as ca
import oracle

pytestmark = pytest.mark.skipif(not oracle.mpi_available(), reason="MPICH (mpirun, mpi.h) not available")


def graphs():
    rng = np.random.default_rng(4)
    return [
        ("simple", ca.read_metis(os.path.join(os.path.dirname(__file__), "golden", "simple.graph"))),
        ("gnm", ca.generate("gnm_undirected", n=1 << 11, m=1 << 11, seed=7)),
        ("rmat", ca.generate("rmat", n=1 << 12, m=1 << 15, seed=3)),
        ("grid", ca.generate("grid_2d", nx=40, ny=30, p=0.55, seed=2)),
        ("random", ca.Graph.from_edges(500, rng.integers(0, 500, size=(400, 2)))),
        ("empty", ca.Graph(np.zeros(9, np.uint64), np.zeros(0, np.uint32))),
    ]


@pytest.mark.parametrize("ranks", [1, 2, 3, 4])
def test_mpi_labels_are_canonical(tmp_path, ranks):
    for name, g in graphs():
        path = str(tmp_path / f"{name}.csr")
        oracle.write_csr_file(path, g.offsets, g.targets)
        for seq, fixed in ((1, False), (0, True), (64, True)):
            res = oracle.mpi_exp(path, ranks, seed=3, seq=seq, fixed_seq=fixed, labels_path=str(tmp_path / "lab"),
                                 timeout=120)
            assert np.array_equal(res["labels"], oracle.canonical_labels(g.offsets, g.targets)), (name, ranks, seq)
            assert res["n"] == g.n and res["m"] == g.m


@pytest.mark.parametrize("ranks", [1, 3])
def test_mpi_levels_equal_single_process_oracle(tmp_path, ranks):
    """Without isolated vertices the rank blocks over input ids are the oracle's pre-contraction
    blocks over level indices, so the decomposition levels are the same (the fixed point of every
    level does not depend on the sweep order): the MPI form counts the levels after its local
    phase, the oracle lists the local phase as a level too."""
    for name, g in graphs()[1:4]:
        deg = np.diff(g.offsets.astype(np.int64))
        if (deg == 0).any():  # keep the vertices with an edge
            keep = np.flatnonzero(deg > 0)
            newid = np.full(g.n, -1, np.int64)
            newid[keep] = np.arange(len(keep))
            src = np.repeat(np.arange(g.n), deg)
            g = ca.Graph.from_edges(len(keep), np.stack([newid[src], newid[g.targets.astype(np.int64)]], 1))
        path = str(tmp_path / f"{name}.csr")
        oracle.write_csr_file(path, g.offsets, g.targets)
        res = oracle.mpi_exp(path, ranks, seed=5, seq=0, fixed_seq=True, timeout=120)
        run = oracle.Run(g.offsets, g.targets, variant="exp", seed=5 + 1000, seq_limit=0, pre_contract="local",
                         pre_blocks=ranks)
        lv = run.levels()
        assert lv[0][2] == 0  # the local phase
        assert res["levels"] == len(lv) - 1, name
        # one rank sweeps in place in the oracle's order; across ranks the ghost keys lag a sweep
        # behind, so the same fixed point can take more sweeps
        if ranks == 1:
            assert res["sweeps"] == run.sweeps(), name
        else:
            assert res["sweeps"] >= run.sweeps(), name
