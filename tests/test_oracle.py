"""The oracle, pinned against the reference fixture, brute force, the standard
library and closed forms (CPU only)."""
import itertools
import json
import os
import random
from collections import deque

import numpy as np
import pytest

import components_amd as ca
import oracle

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def bfs_labels(n, adj):
    """Brute force: BFS from every unvisited vertex in id order; the start is the minimum."""
    lab = [-1] * n
    for s in range(n):
        if lab[s] >= 0:
            continue
        lab[s] = s
        q = deque([s])
        while q:
            v = q.popleft()
            for w in adj[v]:
                if lab[w] < 0:
                    lab[w] = s
                    q.append(w)
    return np.array(lab, dtype=np.uint32)


def graph_from_edges(n, edges):
    return ca.Graph.from_edges(n, np.array(edges, dtype=np.int64).reshape(-1, 2))


def adj_of(g):
    return [list(g.targets[g.offsets[v]:g.offsets[v + 1]]) for v in range(g.n)]


def random_graph(rng, n, m):
    e = rng.integers(0, max(n, 1), size=(m, 2)) if n else np.zeros((0, 2), np.int64)
    return graph_from_edges(n, e)


def test_simple_graph_fixture():
    # example/simple.graph of the reference: edges {1-3, 2-3, 4-5, 4-6} (1-based)
    g = ca.read_metis(os.path.join(GOLDEN, "simple.graph"))
    assert g.n == 6 and g.m == 4
    assert list(oracle.canonical_labels(g.offsets, g.targets)) == [0, 0, 0, 3, 3, 3]
    for variant in ("exp", "local", "shortcut", "labels"):
        assert list(oracle.Run(g.offsets, g.targets, variant=variant, seed=1001).labels()) == [0, 0, 0, 3, 3, 3]


@pytest.mark.parametrize("n", [1, 2, 3, 4, 5])
def test_canonical_labels_brute_force_all_graphs(n):
    pairs = list(itertools.combinations(range(n), 2))
    for mask in range(1 << len(pairs)):
        edges = [p for i, p in enumerate(pairs) if mask >> i & 1]
        g = graph_from_edges(n, edges)
        want = bfs_labels(n, adj_of(g))
        assert np.array_equal(oracle.canonical_labels(g.offsets, g.targets), want)


def test_canonical_labels_brute_force_sampled_6_7():
    rng = random.Random(3)
    for n in (6, 7):
        pairs = list(itertools.combinations(range(n), 2))
        for _ in range(1500):
            edges = [p for p in pairs if rng.random() < 0.25]
            g = graph_from_edges(n, edges)
            assert np.array_equal(oracle.canonical_labels(g.offsets, g.targets), bfs_labels(n, adj_of(g)))


def test_canonical_invariants_random():
    rng = np.random.default_rng(11)
    for _ in range(20):
        n = int(rng.integers(1, 3000))
        g = random_graph(rng, n, int(rng.integers(0, 2 * n)))
        lab = oracle.canonical_labels(g.offsets, g.targets)
        v = np.arange(n)
        assert np.all(lab <= v)                       # canon[v] <= v
        assert np.array_equal(lab[lab], lab)          # idempotent
        src = np.repeat(v, np.diff(g.offsets).astype(np.int64))
        assert np.array_equal(lab[src], lab[g.targets])  # constant along every edge
        assert len(np.unique(lab)) == int(np.sum(lab == v))


def test_planted_mistake_is_caught():
    # an oracle that drops the last edge of every row must disagree with BFS somewhere
    rng = np.random.default_rng(5)
    caught = False
    for _ in range(20):
        g = random_graph(rng, 40, 30)
        deg = np.diff(g.offsets).astype(np.int64)
        keep = np.ones(len(g.targets), bool)
        ends = g.offsets[1:].astype(np.int64) - 1
        keep[ends[deg > 0]] = False
        src = np.repeat(np.arange(g.n), deg)[keep]
        broken = graph_from_edges(g.n, np.stack([src, g.targets[keep]], 1))
        caught |= not np.array_equal(oracle.canonical_labels(broken.offsets, broken.targets),
                                     bfs_labels(g.n, adj_of(g)))
    assert caught


def test_deviate_restatement_matches_standard_library():
    L = oracle.lib()
    rng = random.Random(9)
    for _ in range(20000):
        gid = rng.randrange(1 << 32)
        seed = rng.randrange(-3, 1 << 31)
        it = rng.randrange(1, 64)
        off = rng.randrange(1 << 34)
        beta = rng.choice([0.1, 0.02, 0.3, 1.0])
        assert oracle.exp_deviate(gid, seed, it, off, beta, True) == L.or_std_deviate_fma(gid, seed, it, off, beta)
        assert oracle.exp_deviate(gid, seed, it, off, beta, False) == L.or_std_deviate_nofma(gid, seed, it, off, beta)


def test_deviate_golden_vectors():
    with open(os.path.join(GOLDEN, "deviates.json")) as f:
        cases = json.load(f)["cases"]
    for c in cases:
        for i in range(c["count"]):
            args = (c["gid0"] + i, c["seed"], c["iteration"], c["rng_offset"], c["beta"])
            assert oracle.exp_deviate(*args, True) == c["fma"][i]
            assert oracle.exp_deviate(*args, False) == c["nofma"][i]


def test_deviate_distribution():
    # P(dev >= k) = exp(-beta k) for the exponential law (checked loosely)
    devs = np.array([oracle.exp_deviate(g, 1001, 1, 1 << 20, 0.1) for g in range(0, 200000, 7)])
    for k in (1, 5, 10, 20):
        p = np.mean(devs >= k)
        assert abs(p - np.exp(-0.1 * k)) < 0.02


def test_lp_fixpoint_equals_closed_form():
    rng = np.random.default_rng(21)
    for trial in range(30):
        n = int(rng.integers(1, 400))
        g = random_graph(rng, n, int(rng.integers(0, 3 * n)))
        dev = rng.integers(0, 25, size=n).astype(np.uint64)
        key0 = (dev << np.uint64(32)) | np.arange(n, dtype=np.uint64)
        got, rounds = oracle.lp_fixpoint(g.offsets, g.targets, key0)
        want = oracle.lp_closed_form(g.offsets, g.targets, key0)
        assert np.array_equal(got, want)
        assert rounds >= 1


def test_lp_fixpoint_path_closed_form():
    # on a path with deviates 0 at vertex 0 only, key(v) = (v hops, label 0) up to where others win
    n = 50
    g = graph_from_edges(n, [(i, i + 1) for i in range(n - 1)])
    dev = np.full(n, 1000, np.uint64)
    dev[0] = 0
    key0 = (dev << np.uint64(32)) | np.arange(n, dtype=np.uint64)
    got, _ = oracle.lp_fixpoint(g.offsets, g.targets, key0)
    assert np.array_equal(got >> np.uint64(32), np.arange(n, dtype=np.uint64))
    assert np.all((got & np.uint64(0xFFFFFFFF)) == 0)


@pytest.mark.parametrize("variant", ["exp", "local", "shortcut", "labels"])
def test_oracle_runs_are_canonical(variant):
    rng = np.random.default_rng(31)
    for trial in range(12):
        n = int(rng.integers(1, 2500))
        g = random_graph(rng, n, int(rng.integers(0, 2 * n)))
        run = oracle.Run(g.offsets, g.targets, variant=variant, seed=1000 + trial)
        assert np.array_equal(run.labels(), oracle.canonical_labels(g.offsets, g.targets))


def test_oracle_exp_levels_shrink_and_radius():
    g = ca.generate("gnm_undirected", n=1 << 12, m=1 << 13, seed=2)
    run = oracle.Run(g.offsets, g.targets, variant="exp", seed=1001, keep_keys=True)
    lv = run.levels()
    assert len(lv) >= 1
    # level 0 holds the vertices of degree > 0, in id order
    deg = np.diff(g.offsets)
    assert (deg == 0).any()
    assert lv[0][0] == int((deg > 0).sum()) and lv[0][1] == g.m
    for a, b in zip(lv, lv[1:]):
        assert b[0] < a[0]
    # radius = max hops from the centre, recomputed from the keys
    keys, gid = run.level_keys(0)
    assert np.array_equal(gid, np.flatnonzero(deg > 0))
    lab = (keys & np.uint64(0xFFFFFFFF)).astype(np.int64)
    hops = (keys >> np.uint64(32)).astype(np.int64) - (keys[lab] >> np.uint64(32)).astype(np.int64)
    assert hops.max() == lv[0][3]
    assert np.all(lab[lab] == lab)  # centres are their own centre


def _level0(g):
    """Level 0 of the contraction: the vertices of degree > 0, renumbered in id order."""
    off = np.asarray(g.offsets, np.int64)
    deg = np.diff(off)
    keep = np.flatnonzero(deg > 0)
    newid = np.full(len(deg), -1, np.int64)
    newid[keep] = np.arange(len(keep))
    tgt = newid[np.asarray(g.targets, np.int64)]
    return np.concatenate([[0], np.cumsum(deg[keep])]).astype(np.uint64), tgt.astype(np.uint32)


def _block_reference(off, tgt, blocks):
    """Per vertex: smallest member of its component over the edges inside its block
    (scipy's connected_components on those edges), and whether it has an edge leaving its block."""
    import scipy.sparse as sp
    from scipy.sparse.csgraph import connected_components
    n = len(off) - 1
    bs = -(-n // blocks)
    src = np.repeat(np.arange(n), np.diff(off.astype(np.int64)))
    dst = tgt.astype(np.int64)
    inside = src // bs == dst // bs
    a = sp.coo_matrix((np.ones(int(inside.sum())), (src[inside], dst[inside])), shape=(n, n))
    _, comp = connected_components(a, directed=False)
    smallest = np.full(comp.max() + 1, n, np.int64)
    np.minimum.at(smallest, comp, np.arange(n))
    leaves = np.zeros(n, bool)
    leaves[src[~inside]] = True
    return smallest[comp], leaves, src, dst


@pytest.mark.parametrize("mode", ["local", "single", "double"])
def test_oracle_pre_contraction_pinned(mode):
    """The pre-contraction levels of the oracle against closed forms: labels equal union-find
    for every block count; one block leaves no edge (the contracted graph is empty); a block per
    vertex contracts nothing; the level-0 clusters equal scipy's components inside the blocks
    (interface vertices on their own in local mode), and the contracted graph's sizes equal the
    distinct cluster pairs."""
    graphs = [ca.generate("gnm_undirected", n=1 << 11, m=1 << 12, seed=7),
              ca.generate("grid_2d", nx=40, ny=30, p=0.6, seed=2),
              ca.generate("rmat", n=1 << 11, m=1 << 14, seed=3)]
    for g in graphs:
        ref = oracle.canonical_labels(g.offsets, g.targets)
        off, tgt = _level0(g)
        n0, m0 = len(off) - 1, int(off[-1]) // 2
        for blocks in (1, 3, 16, n0):
            run = oracle.Run(g.offsets, g.targets, seed=3, stream=1, seq_limit=0, keep_keys=True,
                             pre_contract=mode, pre_blocks=blocks)
            assert np.array_equal(run.labels(), ref), (mode, blocks)
            lv = run.levels()
            assert lv[0] == (n0, m0, 0, 0)
            if blocks == 1 and mode != "local":
                assert len(lv) == 1
            if blocks == 1 and mode == "local":
                assert len(lv) == 1  # no vertex has a neighbour in another block
            if blocks == n0:
                assert lv[1][:2] == (n0, m0)
            keys, _ = run.level_keys(0)
            assert np.all(keys >> np.uint64(32) == 0)
            smallest, leaves, src, dst = _block_reference(off, tgt, blocks)
            want = np.where(leaves, np.arange(n0), smallest) if mode == "local" else smallest
            assert np.array_equal(keys.astype(np.int64), want), (mode, blocks)
            cl = want
            pairs = np.unique(np.stack([cl[src], cl[dst]], 1)[cl[src] != cl[dst]], axis=0)
            if len(pairs):
                assert lv[1][0] == len(np.unique(pairs[:, 0])) and lv[1][1] == len(pairs) // 2
