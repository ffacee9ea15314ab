"""Pins of the oracle's random streams and of the local variant's two passes (CPU only).

The reference draws exponential deviates with std::exponential_distribution<double>(beta),
truncated (lib/components/exponential_contraction.h:467-471), and the local variant draws
uniform_int_distribution<unsigned>(0, 99) (lib/components/local_contraction.h:173-176), then
takes two synchronous min passes over N(v) (:195-274).  The hashed exponential stream and the
uniform 0..99 stream are this build's own draws of those laws, so they are pinned by what the
mathematics fixes: published SplitMix64 output sequences for the mixer, goodness-of-fit tests
for the laws, independence across ids and iterations, and, for the two passes, the closed form
min over the closed 2-hop neighbourhood.  Every check is also run against a planted mistake
that it must reject.
"""
import math

import numpy as np
import pytest
import scipy.sparse as sp
from scipy import stats

import components_amd as ca
import oracle

M64 = (1 << 64) - 1
GOLDEN_GAMMA = 0x9E3779B97F4A7C15


def splitmix64_outputs(mix, seed, count):
    """The SplitMix64 generator: state += golden gamma, output = mix(state)."""
    return [mix((seed + i * GOLDEN_GAMMA) & M64) for i in range(1, count + 1)]


def mix_shifted_constant(z):
    """Planted mistake: the second multiplier off by one in its last bit."""
    z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & M64
    z = ((z ^ (z >> 27)) * 0x94D049BB133111EA) & M64
    return z ^ (z >> 31)


def mix_without_first_xorshift(z):
    """Planted mistake: the first xor-shift dropped."""
    z = (z * 0xBF58476D1CE4E5B9) & M64
    z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & M64
    return z ^ (z >> 31)


# Published output sequences of SplitMix64 (Vigna's splitmix64.c; java.util.SplittableRandom
# uses the same mixer): seed 0 and seed 1234567.
KAT_SEED0 = [0xE220A8397B1DCDAF, 0x6E789E6AA1B965F4, 0x06C45D188009454F, 0xF88BB8A8724C81EC]
KAT_SEED1234567 = [6457827717110365317, 3203168211198807973, 9817491932198370423, 4593380528125082431,
                   16408922859458223821]


def kat_holds(mix):
    return (splitmix64_outputs(mix, 0, 4) == KAT_SEED0 and
            splitmix64_outputs(mix, 1234567, 5) == KAT_SEED1234567)


def test_splitmix64_known_answers():
    assert kat_holds(oracle.splitmix64_mix)
    assert not kat_holds(mix_shifted_constant)
    assert not kat_holds(mix_without_first_xorshift)


def avalanche_ok(mix, samples=2048, seed=5):
    """Flipping any one input bit flips every output bit with probability close to 1/2."""
    rng = np.random.default_rng(seed)
    xs = [int(x) for x in rng.integers(0, M64, size=samples, dtype=np.uint64, endpoint=True)]
    ys = np.array([mix(x) for x in xs], np.uint64)
    shifts = np.arange(64, dtype=np.uint64)
    p = np.empty((64, 64))
    for i in range(64):
        d = ys ^ np.array([mix(x ^ (1 << i)) for x in xs], np.uint64)
        p[i] = ((d[:, None] >> shifts) & np.uint64(1)).mean(0)
    # binomial sd at p = 1/2 is about 0.011 for 2048 samples: 0.44..0.56 is > 5 sd
    return bool(np.all(np.abs(p - 0.5) < 0.06))


def test_splitmix64_bijective_on_sample_and_avalanche():
    xs = list(range(0, 1 << 16)) + [(i * GOLDEN_GAMMA) & M64 for i in range(1 << 16)]
    ys = {oracle.splitmix64_mix(x) for x in xs}
    assert len(ys) == len(set(xs))
    assert avalanche_ok(oracle.splitmix64_mix)
    assert not avalanche_ok(mix_without_first_xorshift)


def uniform_sample(draw, seed, iteration, count, stride=1):
    return np.array([draw(g, seed, iteration) for g in range(0, count * stride, stride)])


def uniform_law_ok(u):
    return stats.kstest(u, "uniform").pvalue > 1e-4 and u.min() >= 0.0 and u.max() < 1.0


@pytest.mark.parametrize("seed,iteration", [(1001, 1), (3, 2), (-7, 9), (1 << 30, 1)])
def test_hashed_uniform_is_uniform(seed, iteration):
    u = uniform_sample(oracle.hashed_uniform, seed, iteration, 60000)
    assert uniform_law_ok(u)
    # planted mistake: one bit too many shifted out (u in [0, 1/2))
    assert not uniform_law_ok(u / 2)


def independence_ok(a, b):
    """Pearson correlation of two paired samples within 5 standard errors of 0."""
    r = np.corrcoef(a, b)[0, 1]
    return abs(r) < 5 / math.sqrt(len(a))


def test_hashed_uniform_independent_across_ids_and_iterations():
    n = 40000
    u1 = uniform_sample(oracle.hashed_uniform, 1001, 1, n)
    u2 = uniform_sample(oracle.hashed_uniform, 1001, 2, n)
    assert independence_ok(u1[:-1], u1[1:])   # neighbouring ids
    assert independence_ok(u1, u2)            # the same id at the next iteration
    s2 = uniform_sample(oracle.hashed_uniform, 1002, 1, n)
    assert independence_ok(u1, s2)            # the next seed
    # planted mistakes: the iteration left out of the key; the id halved
    same_it = uniform_sample(lambda g, s, i: oracle.hashed_uniform(g, s, 1), 1001, 2, n)
    assert not independence_ok(u1, same_it)
    halved = np.array([oracle.hashed_uniform(g // 2, 1001, 1) for g in range(n)])
    assert not independence_ok(halved[:-1], halved[1:])


def tail_law_ok(devs, beta, ks=(1, 2, 5, 10, 20, 40)):
    """P(dev >= k) = exp(-beta k) within 5 binomial standard errors, for every k with p > 1e-3."""
    n = len(devs)
    for k in ks:
        p = math.exp(-beta * k)
        if p < 1e-3:
            continue
        sd = math.sqrt(p * (1 - p) / n)
        if abs(np.mean(devs >= k) - p) > 5 * sd + 1e-12:
            return False
    return True


@pytest.mark.parametrize("beta", [0.02, 0.1, 0.3, 1.0])
def test_hashed_deviate_tail_law(beta):
    n = 50000
    devs = np.array([oracle.hashed_deviate(g, 1001, 1, beta) for g in range(n)])
    assert tail_law_ok(devs, beta)
    # planted mistake: the rate off by a factor 2
    wrong = np.array([oracle.hashed_deviate(g, 1001, 1, 2 * beta) for g in range(n)])
    assert not tail_law_ok(wrong, beta)


@pytest.mark.parametrize("beta", [0.1, 0.5])
def test_hashed_deviate_is_truncated_continuous_exponential(beta):
    n = 40000
    u = uniform_sample(oracle.hashed_uniform, 77, 3, n)
    x = np.array([-math.log(1.0 - v) / beta for v in u])
    assert stats.kstest(x, "expon", args=(0, 1 / beta)).pvalue > 1e-4
    assert stats.kstest(x / 2, "expon", args=(0, 1 / beta)).pvalue < 1e-6  # planted: rate doubled
    devs = np.array([oracle.hashed_deviate(g, 77, 3, beta) for g in range(n)])
    assert np.array_equal(devs, np.floor(x).astype(np.int64))


def uniform100_law_ok(draws):
    counts = np.bincount(draws, minlength=101)
    if counts[100:].sum() or len(counts) > 101:
        return False
    return stats.chisquare(counts[:100]).pvalue > 1e-4


@pytest.mark.parametrize("seed,iteration", [(1001, 1), (5, 4), (-2, 17)])
def test_uniform_deviate_100_law(seed, iteration):
    n = 100000
    d = uniform_sample(oracle.uniform_deviate_100, seed, iteration, n)
    assert d.min() == 0 and d.max() == 99
    assert uniform100_law_ok(d)
    # planted mistake: the range squeezed to 0..98
    assert not uniform100_law_ok(d * 99 // 100)


def test_uniform_deviate_100_independent():
    n = 40000
    a = uniform_sample(oracle.uniform_deviate_100, 1001, 1, n)
    b = uniform_sample(oracle.uniform_deviate_100, 1001, 2, n)
    assert independence_ok(a[:-1], a[1:]) and independence_ok(a, b)
    # the pair (neighbouring ids) is uniform over the 100 x 100 cells
    assert stats.chisquare(np.bincount(a[:-1] * 100 + a[1:], minlength=10000)).pvalue > 1e-4
    same_it = uniform_sample(lambda g, s, i: oracle.uniform_deviate_100(g, s, 1), 1001, 2, n)
    assert not independence_ok(a, same_it)


# --- the local variant's two synchronous passes --------------------------------------------

def level0(g):
    off = np.asarray(g.offsets, np.int64)
    deg = np.diff(off)
    keep = np.flatnonzero(deg > 0)
    newid = np.full(len(deg), -1, np.int64)
    newid[keep] = np.arange(len(keep))
    tgt = newid[np.asarray(g.targets, np.int64)]
    return np.concatenate([[0], np.cumsum(deg[keep])]).astype(np.int64), tgt, keep


def two_hop_min(off, tgt, key0):
    """Closed form of two synchronous min passes: min of key0 over the closed 2-hop
    neighbourhood, read from the sparsity pattern of (A + I)^2."""
    n = len(off) - 1
    a = sp.csr_matrix((np.ones(len(tgt)), tgt, off), shape=(n, n)) + sp.identity(n, format="csr")
    b = (a @ a).tocsr()
    b.sort_indices()
    return np.minimum.reduceat(key0[b.indices], b.indptr[:-1])


def passes_in_place(off, tgt, key0, passes=2):
    """Planted mistake: each pass updates the keys in place (Gauss-Seidel) instead of from the
    previous pass's keys."""
    k = key0.copy()
    for _ in range(passes):
        for v in range(len(off) - 1):
            for e in range(off[v], off[v + 1]):
                k[v] = min(k[v], k[tgt[e]])
    return k


def test_local_keys_equal_two_hop_closed_form():
    graphs = [ca.generate("gnm_undirected", n=1 << 12, m=1 << 13, seed=7),
              ca.generate("grid_2d", nx=50, ny=40, p=0.7, seed=2),
              ca.generate("rmat", n=1 << 12, m=1 << 15, seed=3),
              ca.generate("rgg_2d", n=1 << 12, r=0.03, seed=4)]
    planted_caught = False
    for gi, g in enumerate(graphs):
        seed = 1001 + gi
        run = oracle.Run(g.offsets, g.targets, variant="local", seed=seed, seq_limit=0, keep_keys=True)
        keys, gid = run.level_keys(0)
        off, tgt, keep = level0(g)
        assert np.array_equal(gid, keep)
        dev = np.array([oracle.uniform_deviate_100(int(x), seed, 1) for x in keep], np.uint64)
        key0 = (dev << np.uint64(32)) | np.arange(len(keep), dtype=np.uint64)
        assert np.array_equal(keys, two_hop_min(off, tgt, key0))
        if gi < 2:  # the pure-Python planted pass runs on the two smaller graphs
            planted_caught |= not np.array_equal(passes_in_place(off, tgt, key0), two_hop_min(off, tgt, key0))
            assert not np.array_equal(passes_in_place(off, tgt, key0, passes=1), two_hop_min(off, tgt, key0))
    assert planted_caught


# --- hub replication (-replicate) ----------------------------------------------------------

def split_ok(g, off_r, tgt_r, gid_r, threshold):
    """What a split of the hubs into proxies must satisfy (DistributeHighDegreeVertices,
    exponential_contraction.h:935-1253, every vertex its own PE): the graph stays symmetric and
    simple; proxies come after the n0 vertices, each joined to its hub, which keeps only its
    proxies; a hub of degree d has at most floor(sqrt(d)) of them; and merging every proxy into
    its hub gives back level 0's edges exactly, each once per direction."""
    off0, tgt0, keep = level0(g)
    n0, nr = len(off0) - 1, len(off_r) - 1
    R = nr - n0
    off_r = off_r.astype(np.int64)
    if R < 0 or off_r[-1] != off0[-1] + 2 * R:
        return False
    src = np.repeat(np.arange(nr), np.diff(off_r))
    dst = tgt_r.astype(np.int64)
    if (src == dst).any() or len(np.unique(src * nr + dst)) != len(src):  # no loops, no duplicates
        return False
    if not np.array_equal(np.sort(src * nr + dst), np.sort(dst * nr + src)):  # symmetric
        return False
    # a proxy's hub: the level-0 vertex with its id
    pos = np.full(int(g.n), -1, np.int64)
    pos[keep] = np.arange(n0)
    home = np.concatenate([np.arange(n0), pos[gid_r[n0:].astype(np.int64)]])
    if not np.array_equal(gid_r[:n0], keep) or (home < 0).any():
        return False
    deg0 = np.diff(off0)
    limit = max(1, int(threshold * math.sqrt(n0)))
    hubs = np.flatnonzero(deg0 >= limit)
    is_hub = np.zeros(n0, bool)
    is_hub[hubs] = True
    proxies_of = np.bincount(home[n0:], minlength=n0)
    if (proxies_of[~is_hub] != 0).any() or (proxies_of[hubs] > np.floor(np.sqrt(deg0[hubs]))).any():
        return False
    if (is_hub & (proxies_of == 0) & (deg0 > 0)).any():
        return False
    for p in range(n0, nr):  # every proxy's row starts at its hub and holds it once
        row = dst[off_r[p]:off_r[p + 1]]
        if row[0] != home[p] or (row == home[p]).sum() != 1 or len(row) < 2:
            return False
    for h in hubs:  # a hub's row is its proxies
        row = dst[off_r[h]:off_r[h + 1]]
        if not (np.all(row >= n0) and np.all(home[row] == h)):
            return False
    merged = np.stack([home[src], home[dst]], 1)
    merged = merged[merged[:, 0] != merged[:, 1]]  # the hub-proxy joins vanish
    orig = np.stack([np.repeat(np.arange(n0), deg0), tgt0], 1)
    a = np.sort(merged[:, 0] * n0 + merged[:, 1])
    b = np.sort(orig[:, 0] * n0 + orig[:, 1])
    return np.array_equal(a, b)


def test_replicated_level0_pinned():
    star = ca.Graph.from_edges(3001, [(0, i) for i in range(1, 3001)])
    hubs = ca.Graph.from_edges(2500, [(0, i) for i in range(1, 1200)] + [(1300, i) for i in range(1301, 2500)]
                               + [(0, 1300), (5, 1305), (7, 8)])
    cases = [(star, 16.0), (hubs, 4.0), (ca.generate("rmat", n=1 << 12, m=1 << 16, seed=3), 2.0),
             (ca.generate("rhg", n=1 << 12, gamma=2.5, d=12.0, seed=5), 1.0),
             (ca.generate("gnm_undirected", n=1 << 10, m=1 << 12, seed=2), 16.0)]  # no hub at all
    for g, thr in cases:
        for seed in (1001, 7):
            o, t, gid = oracle.replicate(g.offsets, g.targets, seed, thr)
            assert split_ok(g, o, t, gid, thr), (g.n, thr, seed)
    # the hub pair (0, 1300) joins two proxies; the split moved something
    o, t, gid = oracle.replicate(hubs.offsets, hubs.targets, 3, 4.0)
    assert len(o) - 1 > len(level0(hubs)[0]) - 1
    # planted mistakes: a hub that keeps its own edges besides its proxies; one proxy edge
    # redirected to the hub itself (the two ends disagree about the proxy)
    off0, tgt0, _ = level0(hubs)
    src = np.repeat(np.arange(len(o) - 1), np.diff(o.astype(np.int64)))
    keep_hub = np.concatenate([np.stack([src, t.astype(np.int64)], 1),
                               np.stack([np.zeros(int(off0[1]), np.int64), tgt0[:int(off0[1])]], 1),
                               np.stack([tgt0[:int(off0[1])], np.zeros(int(off0[1]), np.int64)], 1)])
    bad = ca.Graph.from_edges(len(o) - 1, keep_hub)
    assert not split_ok(hubs, bad.offsets, bad.targets, gid, 4.0)
    t2 = t.copy()
    e = int(np.flatnonzero(t2 >= len(off0) - 1)[-1])  # some edge into a proxy
    t2[e] = 0
    assert not split_ok(hubs, o, t2, gid, 4.0)


def test_replicated_runs_are_canonical():
    for g, thr in [(ca.generate("rmat", n=1 << 12, m=1 << 16, seed=3), 2.0),
                   (ca.generate("rhg", n=1 << 12, gamma=2.5, d=12.0, seed=5), 1.0)]:
        ref = oracle.canonical_labels(g.offsets, g.targets)
        for stream in (0, 1):
            run = oracle.Run(g.offsets, g.targets, seed=1001, stream=stream, seq_limit=0, keep_keys=True,
                             replicate=True, hd_threshold=thr)
            assert np.array_equal(run.labels(), ref)
            hubs, reps = run.replication()
            assert hubs > 0 and reps > 0
            keys, gid = run.level_keys(0)
            n0 = int(np.count_nonzero(np.diff(g.offsets.astype(np.int64))))
            assert run.levels()[0][0] == n0 + reps
            # a proxy draws its hub's deviate: its key's starting distance is the hub's
            dev = np.array([oracle.hashed_deviate(int(x), 1001, 1) if stream else
                            oracle.exp_deviate(int(x), 1001, 1, n0, 0.1, True) for x in gid], np.uint64)
            assert np.all((keys >> np.uint64(32)) <= dev)
