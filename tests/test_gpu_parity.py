"""GPU path through the C ABI against the oracle (bit-exact for all integer work)."""
import itertools
import json
import os
import re
import subprocess

import numpy as np
import pytest

import components_amd as ca
import oracle

pytestmark = pytest.mark.gpu
GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
VARIANTS = ["exp", "local", "shortcut", "labels"]


def union(graphs):
    """Disjoint union: canonical labels of the union are the per-graph labels plus offsets."""
    offs, tgts, base_v, base_e = [np.zeros(1, np.uint64)], [], 0, 0
    for g in graphs:
        offs.append(g.offsets[1:].astype(np.uint64) + np.uint64(base_e))
        tgts.append(g.targets.astype(np.uint32) + np.uint32(base_v))
        base_v += g.n
        base_e += int(g.offsets[-1])
    return ca.Graph(np.concatenate(offs), np.concatenate(tgts) if tgts else np.zeros(0, np.uint32))


def rand_graph(rng, n, m):
    return ca.Graph.from_edges(n, rng.integers(0, n, size=(m, 2)))


def check_labels(g, variants=VARIANTS, **kw):
    ref = oracle.canonical_labels(g.offsets, g.targets)
    dg = ca.DeviceGraph(g)
    try:
        for v in variants:
            lab, st = dg.find_components(variant=v, **kw)
            assert np.array_equal(lab, ref), f"variant {v}"
            assert st["components"] == len(np.unique(ref))
    finally:
        dg.close()


def test_deviates_golden_vectors():
    with open(os.path.join(GOLDEN, "deviates.json")) as f:
        cases = json.load(f)["cases"]
    for c in cases:
        for fma, key in ((True, "fma"), (False, "nofma")):
            got = ca.deviates(c["gid0"], c["count"], c["seed"], c["iteration"], c["rng_offset"], c["beta"], fma)
            assert list(got) == c[key]


def test_deviates_bulk_against_oracle():
    rng = np.random.default_rng(1)
    for trial in range(4):
        gid0 = int(rng.integers(0, 1 << 31))
        seed, it, off = int(rng.integers(0, 1 << 20)), int(rng.integers(1, 50)), int(rng.integers(0, 1 << 32))
        beta = [0.1, 0.05, 0.7, 0.013][trial]
        got = ca.deviates(gid0, 50000, seed, it, off, beta, True)
        want = [oracle.exp_deviate(gid0 + i, seed, it, off, beta, True) for i in range(50000)]
        assert np.array_equal(got, np.array(want, np.uint32))
        got = ca.deviates(gid0, 20000, seed, it, off, beta, True, ca.STREAM_HASHED)
        want = [oracle.hashed_deviate(gid0 + i, seed, it, beta) for i in range(20000)]
        assert np.array_equal(got, np.array(want, np.uint32))


def test_simple_graph_fixture():
    g = ca.read_metis(os.path.join(GOLDEN, "simple.graph"))
    for v in VARIANTS:
        lab, _ = ca.find_components(g, variant=v, seed=1001)
        assert list(lab) == [0, 0, 0, 3, 3, 3]


@pytest.mark.parametrize("n", [1, 2, 3, 4, 5])
def test_all_graphs_up_to_5_vertices(n):
    pairs = list(itertools.combinations(range(n), 2))
    graphs = []
    for mask in range(1 << len(pairs)):
        graphs.append(ca.Graph.from_edges(n, [p for i, p in enumerate(pairs) if mask >> i & 1]))
    g = union(graphs)
    for seq in (0, 4096):
        check_labels(g, seed=1001, seq_limit=seq, stream=ca.STREAM_HASHED)
    check_labels(g, variants=["exp"], seed=1001, seq_limit=4096)


def test_sampled_graphs_6_7_vertices():
    rng = np.random.default_rng(7)
    graphs = []
    for n in (6, 7):
        pairs = list(itertools.combinations(range(n), 2))
        for _ in range(3000):
            graphs.append(ca.Graph.from_edges(n, [p for p in pairs if rng.random() < 0.3]))
    check_labels(union(graphs), seed=1002, seq_limit=0, stream=ca.STREAM_HASHED)


# schedule: pull sweeps; time-stepped rounds (push or pull chosen per early round on large dense
# levels, pull sweeps at the tail); rounds with the direction chosen on every level
@pytest.mark.parametrize("bfs", [False, True, "pull1"])
@pytest.mark.parametrize("stream", [ca.STREAM_REFERENCE, ca.STREAM_HASHED])
@pytest.mark.parametrize("case", ["gnm", "grid2", "grid3", "rmat", "rgg", "path", "star"])
def test_level_trace_matches_oracle(case, stream, bfs):
    rng = np.random.default_rng(abs(hash(case)) % 1000)
    if case == "gnm":
        g = ca.generate("gnm_undirected", n=1 << 14, m=1 << 15, seed=3)
    elif case == "grid2":
        g = ca.generate("grid_2d", nx=128, ny=100, p=0.55, seed=2)
    elif case == "grid3":
        g = ca.generate("grid_3d", nx=24, ny=24, nz=24, p=0.3, seed=2)
    elif case == "rmat":
        g = ca.generate("rmat", n=1 << 13, m=1 << 16, seed=3)
    elif case == "rgg":
        g = ca.generate("rgg_2d", n=1 << 13, r=0.02, seed=2)
    elif case == "path":
        g = ca.Graph.from_edges(5000, [(i, i + 1) for i in range(4999)])
    else:
        g = ca.Graph.from_edges(20001, [(0, i) for i in range(1, 20001)] + [(i, i + 1) for i in range(1, 1000)])
    seq = 64
    run = oracle.Run(g.offsets, g.targets, variant="exp", seed=1005, seq_limit=seq, keep_keys=True, stream=stream)
    dg = ca.DeviceGraph(g)
    lab, st = dg.find_components(variant="exp", seed=1005, seq_limit=seq, record_trace=True, stream=stream,
                                 use_bfs=bool(bfs), pull_rounds=1 if bfs == "pull1" else 0)
    tr = dg.trace()
    dg.close()
    want_levels = run.levels()
    assert len(tr) == len(want_levels)
    for lvl, ((info, keys, gid), winfo) in enumerate(zip(tr, want_levels)):
        assert info[0] == winfo[0] and info[1] == winfo[1], f"level {lvl} sizes"
        assert info[3] == winfo[3], f"level {lvl} max radius"
        wkeys, wgid = run.level_keys(lvl)
        assert np.array_equal(gid, wgid), f"level {lvl} gid"
        assert np.array_equal(keys, wkeys), f"level {lvl} keys"
    assert np.array_equal(lab, run.labels())
    assert np.array_equal(lab, oracle.canonical_labels(g.offsets, g.targets))
    assert st["radius_sum"] == sum(w[3] for w in want_levels)


def replicate_graphs():
    star = ca.Graph.from_edges(3001, [(0, i) for i in range(1, 3001)])
    hubs = ca.Graph.from_edges(2500, [(0, i) for i in range(1, 1200)] + [(1300, i) for i in range(1301, 2500)]
                               + [(0, 1300), (5, 1305), (7, 8)])
    return [("star", star, 16.0), ("hubs", hubs, 4.0), ("rmat", ca.generate("rmat", n=1 << 13, m=1 << 16, seed=3), 2.0),
            ("rhg", ca.generate("rhg", n=1 << 13, gamma=2.5, d=12.0, seed=5), 1.0),
            ("gnm", ca.generate("gnm_undirected", n=1 << 12, m=1 << 13, seed=2), 16.0)]


@pytest.mark.parametrize("bfs", [False, True])
@pytest.mark.parametrize("stream", [ca.STREAM_REFERENCE, ca.STREAM_HASHED])
def test_replicate_trace_matches_oracle(stream, bfs):
    """-replicate: level 0 with its hubs split into proxies that draw their hub's deviate; every
    level's sizes, radius, ids and keys equal the oracle's, and so do the hub and proxy counts."""
    for name, g, thr in replicate_graphs():
        print(f"replicate {name}: n={g.n} m={g.m} threshold={thr}", flush=True)
        for seq in (0, 64):
            run = oracle.Run(g.offsets, g.targets, variant="exp", seed=1007, seq_limit=seq, keep_keys=True,
                             stream=stream, replicate=True, hd_threshold=thr)
            dg = ca.DeviceGraph(g)
            try:
                lab, st = dg.find_components(variant="exp", seed=1007, seq_limit=seq, record_trace=True, stream=stream,
                                             use_bfs=bfs, replicate=True, hd_threshold=thr)
                tr = dg.trace()
            finally:
                dg.close()
            assert (st["hubs"], st["replicas"]) == run.replication(), name
            want = run.levels()
            assert len(tr) == len(want), name
            for lvl, ((info, keys, gid), winfo) in enumerate(zip(tr, want)):
                assert info[0] == winfo[0] and info[1] == winfo[1], (name, lvl)
                assert info[3] == winfo[3], (name, lvl)
                wkeys, wgid = run.level_keys(lvl)
                assert np.array_equal(gid, wgid), (name, lvl)
                assert np.array_equal(keys, wkeys), (name, lvl)
            assert np.array_equal(lab, oracle.canonical_labels(g.offsets, g.targets)), name
            assert st["components"] == len(np.unique(lab))


def test_replicate_option_checks():
    g = ca.generate("rmat", n=1 << 10, m=1 << 13, seed=3)
    dg = ca.DeviceGraph(g)
    try:
        for bad in (dict(variant="local", replicate=True), dict(replicate=True, pre_contract="single", pre_blocks=4),
                    dict(replicate=True, hd_threshold=-1.0)):
            with pytest.raises(ca.KcError):
                dg.find_components(**bad)
        # a threshold no vertex reaches: the plain path, nothing split
        lab, st = dg.find_components(replicate=True, hd_threshold=1e6, seed=3)
        assert st["hubs"] == 0 and st["replicas"] == 0
        assert np.array_equal(lab, oracle.canonical_labels(g.offsets, g.targets))
        # the same handle after a replicated call (its work buffers may have grown)
        lab2, st2 = dg.find_components(replicate=True, hd_threshold=0.5, seed=4)
        assert st2["replicas"] > 0 and np.array_equal(lab2, lab)
        lab3, _ = dg.find_components(seed=5)
        assert np.array_equal(lab3, lab)
    finally:
        dg.close()


def test_autotuned_bins_keep_results():
    """The autotuner only picks how rows are spread over lanes: with its choice kept in the handle
    every level's sizes, radius, ids and keys, and the labels, equal the defaults' (and the
    oracle's labels); the stats report the thresholds used."""
    for g in (ca.generate("rmat", n=1 << 14, m=1 << 18, seed=3), ca.generate("rgg_2d", n=1 << 14, r=0.012, seed=2)):
        dg = ca.DeviceGraph(g)
        try:
            kw = dict(seed=1009, seq_limit=64, stream=ca.STREAM_HASHED, use_bfs=True)
            lab0, st0 = dg.find_components(record_trace=True, **kw)
            tr0 = dg.trace()
            assert (st0["bin_small"], st0["bin_mid"], st0["bin_large"]) == (16, 256, 4096)
            t = dg.autotune(**kw)
            print(f"autotune n={g.n}: {t}", flush=True)
            assert t["candidates"] == 27 and t["ms"] <= t["default_ms"]
            lab, st = dg.find_components(record_trace=True, **kw)
            assert (st["bin_small"], st["bin_mid"], st["bin_large"]) == (t["small"], t["mid"], t["large"])
            tr = dg.trace()
            assert np.array_equal(lab, lab0) and len(tr) == len(tr0)
            for (info, keys, gid), (info0, keys0, gid0) in zip(tr, tr0):
                assert info[0] == info0[0] and info[1] == info0[1] and info[3] == info0[3]
                assert np.array_equal(keys, keys0) and np.array_equal(gid, gid0)
            # explicit thresholds still win over the kept ones; clearing restores the defaults
            _, st = dg.find_components(tmax=4, smax=64, wmax=256, **kw)
            assert (st["bin_small"], st["bin_mid"], st["bin_large"]) == (4, 64, 256)
            dg.set_bins()
            _, st = dg.find_components(**kw)
            assert (st["bin_small"], st["bin_mid"], st["bin_large"]) == (16, 256, 4096)
            assert np.array_equal(lab, oracle.canonical_labels(g.offsets, g.targets))
        finally:
            dg.close()


def test_local_contraction_trace_matches_oracle():
    g = ca.generate("gnm_undirected", n=1 << 13, m=1 << 14, seed=9)
    run = oracle.Run(g.offsets, g.targets, variant="local", seed=1001, seq_limit=64, keep_keys=True)
    dg = ca.DeviceGraph(g)
    lab, _ = dg.find_components(variant="local", seed=1001, seq_limit=64, record_trace=True)
    tr = dg.trace()
    dg.close()
    assert len(tr) == len(run.levels())
    for lvl, (info, keys, gid) in enumerate(tr):
        wkeys, wgid = run.level_keys(lvl)
        assert np.array_equal(keys, wkeys) and np.array_equal(gid, wgid)
    assert np.array_equal(lab, run.labels())


@pytest.mark.parametrize("case", ["empty", "one", "isolated", "pairs", "clique", "star", "path", "twohubs"])
def test_edge_cases(case):
    if case == "empty":
        g = ca.Graph(np.zeros(1, np.uint64), np.zeros(0, np.uint32))
    elif case == "one":
        g = ca.Graph.from_edges(1, [])
    elif case == "isolated":
        g = ca.Graph.from_edges(1000, [])
    elif case == "pairs":
        g = ca.Graph.from_edges(2000, [(2 * i, 2 * i + 1) for i in range(1000)])
    elif case == "clique":
        g = ca.Graph.from_edges(300, list(itertools.combinations(range(300), 2)))
    elif case == "star":
        g = ca.Graph.from_edges(200001, [(200000, i) for i in range(200000)])
    elif case == "path":
        g = ca.Graph.from_edges(100000, [(i, i + 1) for i in range(99999)])
    else:
        e = [(0, i) for i in range(2, 60000)] + [(1, i) for i in range(30000, 90000)]
        g = ca.Graph.from_edges(90000, e)
    for stream in (ca.STREAM_REFERENCE, ca.STREAM_HASHED):
        check_labels(g, seed=1001, stream=stream)
    check_labels(g, variants=["exp"], seed=1001, seq_limit=0, stream=ca.STREAM_HASHED)
    check_labels(g, variants=["exp"], seed=1001, seq_limit=0, stream=ca.STREAM_HASHED, use_bfs=True)


@pytest.mark.parametrize("wmax,smax,tmax", [(256, 64, 4), (4096, 256, 16), (1 << 16, 1024, 32)])
def test_bin_thresholds_do_not_change_results(wmax, smax, tmax):
    g = ca.generate("rmat", n=1 << 15, m=1 << 19, seed=4)
    ref = oracle.canonical_labels(g.offsets, g.targets)
    lab, _ = ca.find_components(g, variant="exp", seed=1001, tmax=tmax, smax=smax, wmax=wmax)
    assert np.array_equal(lab, ref)


def test_random_graphs_many_seeds():
    rng = np.random.default_rng(17)
    for trial in range(8):
        n = int(rng.integers(10, 50000))
        g = rand_graph(rng, n, int(rng.integers(0, 3 * n)))
        check_labels(g, seed=1000 + trial, stream=trial % 2)


def test_config0_full_size():
    g = ca.generate("gnm_undirected", n=1 << 20, m=1 << 22, seed=1)
    ref = oracle.canonical_labels(g.offsets, g.targets)
    dg = ca.DeviceGraph(g)
    for stream in (ca.STREAM_REFERENCE, ca.STREAM_HASHED):
        for bfs in (False, True):
            lab, st = dg.find_components(variant="exp", seed=1001, seq_limit=16384, stream=stream, use_bfs=bfs)
            assert np.array_equal(lab, ref)
    dg.close()


def test_rmat_scale22_schedules_agree():
    dg = ca.DeviceGraph.generate("rmat", n=1 << 22, m=1 << 26, seed=3)
    g = dg.download()
    ref = oracle.canonical_labels(g.offsets, g.targets)
    traces = []
    # sweeps; rounds with directions chosen and pull sweeps at the tail (default); push rounds only;
    # early switch to sweeps; directions chosen on every level; huge buckets in hash tables, bitmaps,
    # byte maps
    for bfs, tail, r1, tab in ((False, 0, 0, "auto"), (True, 0, 0, "auto"), (True, -1, -1, "auto"),
                               (True, 1, 0, "auto"), (True, 0, 1, "auto"), (True, 0, 0, "tables"),
                               (True, 0, 0, "bitmaps"), (True, 0, 0, "bytemaps"), (True, 0, 0, "maps_tables")):
        lab, st = dg.find_components(variant="exp", seed=1001, seq_limit=dg.m // 256, stream=ca.STREAM_HASHED,
                                     use_bfs=bfs, tail_sweeps=tail, pull_rounds=r1, dedup_huge=tab,
                                     record_trace=True)
        assert np.array_equal(lab, ref), (bfs, tail, r1, tab)
        traces.append(dg.trace())
    for other in traces[1:]:
        assert len(traces[0]) == len(other)
        for (i0, k0, g0), (i1, k1, g1) in zip(traces[0], other):
            assert i0[0] == i1[0] and i0[1] == i1[1] and i0[3] == i1[3]
            assert np.array_equal(k0, k1) and np.array_equal(g0, g1)
    dg.close()


def test_rmat_scale20_invariants():
    g = ca.generate("rmat", n=1 << 20, m=1 << 24, seed=3)
    ref = oracle.canonical_labels(g.offsets, g.targets)
    lab, st = ca.find_components(g, variant="exp", seed=1001, seq_limit=(g.m // 256), stream=ca.STREAM_HASHED)
    assert np.array_equal(lab, ref)


def test_app_check_flag():
    exe = os.path.join(ca.BIN_DIR, "exp_contraction")
    r = subprocess.run([exe, "-gen", "gnm_undirected", "-n", "14", "-m", "16", "-i", "2", "-check"],
                       capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "CHECK labels_bit_exact=yes" in r.stdout
    assert "GLOBAL RESULT runner=exp" in r.stdout
    for app in ("local_contraction", "shortcuts", "labels", "all_reduce"):
        r = subprocess.run([os.path.join(ca.BIN_DIR, app), "-input", "file", "-in",
                            os.path.join(GOLDEN, "simple.graph"), "-i", "1", "-check"],
                           capture_output=True, text=True, timeout=300)
        assert r.returncode == 0 and "CHECK labels_bit_exact=yes" in r.stdout, app
    r = subprocess.run([os.path.join(ca.BIN_DIR, "all_reduce"), "-gen", "rmat", "-n", "15", "-m", "19", "-i", "2",
                        "-check"], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0 and "CHECK labels_bit_exact=yes" in r.stdout, r.stdout + r.stderr
    assert "GLOBAL RESULT runner=allreduce" in r.stdout and "n=32768" in r.stdout


def test_app_reads_edge_and_binary_inputs(tmp_path):
    g = ca.generate("gnm_undirected", n=3000, m=2500, seed=9)
    src = np.repeat(np.arange(g.n), np.diff(g.offsets.astype(np.int64)))
    keep = src < g.targets
    u, v = src[keep] + 1, g.targets[keep].astype(np.int64) + 1
    lines = [f"e {a} {b}" for a, b in zip(u.tolist(), v.tolist())]
    (tmp_path / "g.edges").write_text(f"p {g.n} {len(u)}\n" + "\n".join(lines) + "\n")
    half = len(lines) // 2
    (tmp_path / "g.part_0").write_text(f"p {g.n} {len(u)}\n" + "\n".join(lines[:half]) + "\n")
    (tmp_path / "g.part_1").write_text(f"p {g.n} {len(u)}\n" + "\n".join(lines[half:]) + "\n")
    ca.write_binary(tmp_path / "g.bin", g)
    exe = os.path.join(ca.BIN_DIR, "exp_contraction")
    for args in (["-input", "edgefile", "-in", str(tmp_path / "g.edges")],
                 ["-input", "partition", "-in", str(tmp_path / "g.part"), "-parts", "2"],
                 ["-input", "binary", "-in", str(tmp_path / "g.bin")]):
        r = subprocess.run([exe] + args + ["-i", "2", "-check"], capture_output=True, text=True, timeout=300)
        assert r.returncode == 0, r.stdout + r.stderr
        assert "CHECK labels_bit_exact=yes" in r.stdout, args
        assert f"n={g.n} m={g.m}" in r.stdout, args


@pytest.mark.parametrize("name,kw", [
    ("gnm_undirected", dict(n=50000, m=200000, seed=1)),
    ("rmat", dict(n=1 << 15, m=1 << 19, seed=3)),
    ("grid_2d", dict(nx=300, ny=200, p=0.5, seed=2)),
    ("grid_3d", dict(nx=20, ny=20, nz=20, p=0.3, periodic=1, seed=2)),
    ("ba", dict(n=20000, d=4.0, seed=6)),
])
def test_device_generator_matches_host(name, kw):
    host = ca.generate(name, **kw)
    dg = ca.DeviceGraph.generate(name, **kw)
    dev = dg.download()
    assert np.array_equal(dev.offsets, host.offsets)
    for v in range(dev.n):
        a, b = int(dev.offsets[v]), int(dev.offsets[v + 1])
        if b - a > 1:
            dev.targets[a:b] = np.sort(dev.targets[a:b])
    assert np.array_equal(dev.targets, host.targets)
    lab, _ = dg.find_components(seed=1001)
    assert np.array_equal(lab, oracle.canonical_labels(host.offsets, host.targets))
    dg.close()


def histogram_of(labels):
    _, counts = np.unique(labels, return_counts=True)
    sizes, num = np.unique(counts, return_counts=True)
    return np.column_stack([sizes, num]).astype(np.uint64)


@pytest.mark.parametrize("variant", VARIANTS)
def test_component_histogram(variant):
    rng = np.random.default_rng(17)
    g = union([rand_graph(rng, 3000, 2400), ca.generate("rmat", n=1 << 12, m=1 << 14, seed=3),
               ca.Graph(np.zeros(6, np.uint64), np.zeros(0, np.uint32))])
    ref = oracle.canonical_labels(g.offsets, g.targets)
    dg = ca.DeviceGraph(g)
    try:
        dg.find_components(variant=variant, seed=4)
        h = dg.component_histogram()
    finally:
        dg.close()
    want = histogram_of(ref)
    assert np.array_equal(h, want)
    assert int((h[:, 0] * h[:, 1]).sum()) == g.n


def test_app_prints_component_summary(tmp_path):
    """labels and local_contraction print OutputComponents' lines after each
    run: sizes other than 1 (degree-0 vertices are not listed) and their count."""
    rng = np.random.default_rng(3)
    g = union([rand_graph(rng, 400, 300), ca.Graph.from_edges(5, [(0, 1), (2, 3)])])
    p = tmp_path / "g.graph"
    ca.write_metis(p, g)
    ref = oracle.canonical_labels(g.offsets, g.targets)
    want = histogram_of(ref)
    want = want[want[:, 0] != 1]
    for prog in ("labels", "local_contraction"):
        r = subprocess.run([os.path.join(ca.BIN_DIR, prog), "-input", "file", "-in", str(p), "-i", "2", "-seed", "3"],
                           capture_output=True, text=True, timeout=120)
        assert r.returncode == 0, r.stderr
        lines = [l for l in r.stdout.splitlines() if l.startswith("COMPONENTS [")]
        assert len(lines) == 2, r.stdout
        pairs = re.findall(r"size=(\d+) \(num=(\d+)\)", lines[0])
        assert [(int(a), int(b)) for a, b in pairs] == [(int(a), int(b)) for a, b in want]
        num = [l for l in r.stdout.splitlines() if l.startswith("NUM COMPONENTS")]
        assert num and int(num[0].split()[-1]) == int(want[:, 1].sum())


@pytest.mark.parametrize("stream", [ca.STREAM_HASHED, ca.STREAM_REFERENCE])
@pytest.mark.parametrize("bfs", [False, True])
def test_degree_order_matches_input_order(stream, bfs):
    """Level 0 stored in degree order (large graphs) gives the same levels, keys
    and labels as the input order, and both equal the oracle's trace."""
    g = ca.generate("rmat", n=1 << 17, m=1 << 20, seed=5)
    assert int((np.diff(g.offsets.astype(np.int64)) > 0).sum()) >= 1 << 16
    dg = ca.DeviceGraph(g)
    try:
        runs = {}
        for inorder in (False, True):
            lab, st = dg.find_components(seed=9, stream=stream, use_bfs=bfs, seq_limit=4096, record_trace=True,
                                         input_order=inorder)
            runs[inorder] = (lab, dg.trace())
    finally:
        dg.close()
    ref = oracle.canonical_labels(g.offsets, g.targets)
    run = oracle.Run(g.offsets, g.targets, variant="exp", seed=9, seq_limit=4096, stream=stream, keep_keys=True)
    want = run.levels()
    for inorder, (lab, tr) in runs.items():
        assert np.array_equal(lab, ref), inorder
        assert len(tr) == len(want), inorder
        for lvl, ((info, keys, gid), winfo) in enumerate(zip(tr, want)):
            assert info[0] == winfo[0] and info[1] == winfo[1] and info[3] == winfo[3], (inorder, lvl)
            wkeys, wgid = run.level_keys(lvl)
            assert np.array_equal(keys, wkeys), (inorder, lvl)
            assert np.array_equal(gid, wgid), (inorder, lvl)


def test_owner_split_relabel_matches_one_pass():
    """The owner-split relabel of the degree-ordered level 0 (its long rows cut into pieces by the
    owner part of their targets, each part looked up by its own blocks; the shorter rows above the
    16-lane bin kept whole as shared pieces) gives the same levels, keys, ids and labels as the
    one-pass relabel over whole rows, for split bounds of 64 and 1024 slots and the default.  The
    bin thresholds change between calls on one handle, so the pieces are cut again (short pieces
    with smax 16, rows cut every 256 slots with wmax 256), and every run equals the oracle's
    labels."""
    g = ca.generate("rmat", n=1 << 17, m=1 << 21, seed=7)
    ref = oracle.canonical_labels(g.offsets, g.targets)
    dg = ca.DeviceGraph(g)
    try:
        for bins in ((0, 0, 0), (4, 16, 256), (16, 128, 1024), (0, 0, 0)):
            tmax, smax, wmax = bins
            runs = {}
            for split in (1, 64, 1024, -1):
                for bfs in (True, False):
                    lab, st = dg.find_components(seed=11, stream=ca.STREAM_HASHED, use_bfs=bfs, seq_limit=1024,
                                                 record_trace=True, tmax=tmax, smax=smax, wmax=wmax,
                                                 relabel_split=split)
                    runs[split, bfs] = (lab, dg.trace())
            print(f"bins {bins}: levels {len(runs[1, True][1])}", flush=True)
            for split, bfs in itertools.product((1, 64, 1024), (True, False)):
                (lab, tr), (lab1, tr1) = runs[split, bfs], runs[-1, bfs]
                assert np.array_equal(lab, ref) and np.array_equal(lab1, ref), (bins, split, bfs)
                assert len(tr) == len(tr1) >= 2, (bins, split, bfs)
                for lvl, ((info, keys, gid), (info1, keys1, gid1)) in enumerate(zip(tr, tr1)):
                    # (sizes and radius: the count of in-place pull sweeps, the tail sweeps of the
                    # rounds included, depends on the order in which the device runs the rows)
                    assert (info[0], info[1], info[3]) == (info1[0], info1[1], info1[3]), (bins, split, bfs, lvl)
                    assert np.array_equal(keys, keys1) and np.array_equal(gid, gid1), (bins, split, bfs, lvl)
    finally:
        dg.close()


@pytest.mark.parametrize("beta", [0.1, 0.02])
@pytest.mark.parametrize("case", ["rmat", "grid2"])
def test_packed_sweeps_trace_matches_oracle(case, beta):
    """Pull sweeps (the default schedule) on levels of at least 2^16 vertices run over packed 32-bit
    keys (distance << 24 | label) when every distance stays below 255; at beta 0.02 the deviates
    exceed that and the sweeps keep the 64-bit keys.  Both give the oracle's keys, ids, sizes and
    radii at every level, and its labels."""
    if case == "rmat":
        g = ca.generate("rmat", n=1 << 17, m=1 << 20, seed=5)
    else:
        g = ca.generate("grid_2d", nx=400, ny=400, p=0.6, seed=3)
    assert int((np.diff(g.offsets.astype(np.int64)) > 0).sum()) >= 1 << 16
    seq = 1024
    dg = ca.DeviceGraph(g)
    try:
        lab, st = dg.find_components(seed=21, beta=beta, stream=ca.STREAM_HASHED, use_bfs=False, seq_limit=seq,
                                     record_trace=True)
        tr = dg.trace()
    finally:
        dg.close()
    run = oracle.Run(g.offsets, g.targets, variant="exp", seed=21, beta=beta, seq_limit=seq, stream=ca.STREAM_HASHED,
                     keep_keys=True)
    want = run.levels()
    print(f"{case} beta {beta}: levels {len(want)}, sweeps {st['sweeps']}", flush=True)
    assert len(tr) == len(want) >= 1
    for lvl, ((info, keys, gid), winfo) in enumerate(zip(tr, want)):
        assert info[0] == winfo[0] and info[1] == winfo[1] and info[3] == winfo[3], (case, beta, lvl)
        wkeys, wgid = run.level_keys(lvl)
        assert np.array_equal(gid, wgid), (case, beta, lvl)
        assert np.array_equal(keys, wkeys), (case, beta, lvl)
    assert np.array_equal(lab, oracle.canonical_labels(g.offsets, g.targets))


def test_option_checks_and_timed_class():
    g = ca.generate("rmat", n=1 << 16, m=1 << 20, seed=4)
    ref = oracle.canonical_labels(g.offsets, g.targets)
    dg = ca.DeviceGraph(g)
    try:
        for bad in (dict(dedup_huge=5), dict(dedup_huge=-1)):
            with pytest.raises(ca.KcError):
                dg.find_components(variant="exp", seed=1001, **bad)
        # every launch timed, then only the relabel launches: same launches, time on one class only
        # (after a first call, which also builds the level-0 bins that later calls reuse)
        dg.find_components(variant="exp", seed=1001, use_bfs=True)
        lab_all, st_all = dg.find_components(variant="exp", seed=1001, use_bfs=True, time_kernels=True)
        lab_one, st_one = dg.find_components(variant="exp", seed=1001, use_bfs=True, time_kernels=True,
                                             time_class="relabel")
        assert np.array_equal(lab_all, ref) and np.array_equal(lab_one, ref)
        assert st_all["kernel_launches"] == st_one["kernel_launches"]
        assert st_one["kernel_ms"]["relabel"] > 0
        assert all(v == 0 for k, v in st_one["kernel_ms"].items() if k != "relabel")
        assert sum(v > 0 for v in st_all["kernel_ms"].values()) > 1
    finally:
        dg.close()


@pytest.mark.parametrize("mode", ["bitmaps", "bytemaps", "tables", "maps_tables"])
def test_huge_bucket_modes_on_a_hub_graph(mode):
    # two hubs sharing most of their neighbours: their clusters' buckets are huge and full of repeats
    rng = np.random.default_rng(11)
    n = 1 << 17
    src = np.concatenate([np.zeros(60000, np.int64), np.ones(60000, np.int64), rng.integers(2, n, 200000)])
    dst = np.concatenate([rng.integers(2, n, 60000), rng.integers(2, n, 60000), rng.integers(2, n, 200000)])
    g = ca.Graph.from_edges(n, np.stack([src, dst], axis=1))
    ref = oracle.canonical_labels(g.offsets, g.targets)
    for bfs in (False, True):
        lab, _ = ca.find_components(g, variant="exp", seed=1001, seq_limit=0, use_bfs=bfs,
                                    stream=ca.STREAM_HASHED, dedup_huge=mode)
        assert np.array_equal(lab, ref), (mode, bfs)


def test_device_binary_reader_matches_host(tmp_path):
    def sorted_rows(g):
        off = g.offsets.astype(np.int64)
        return [sorted(g.targets[off[v]:off[v + 1]].tolist()) for v in range(g.n)]

    for name, kw in (("rmat", dict(n=1 << 14, m=1 << 18, seed=5)), ("gnm_undirected", dict(n=5000, m=3000, seed=2))):
        g = ca.generate(name, **kw)
        p = tmp_path / f"{name}.bin"
        ca.write_binary(p, g)
        host = ca.read_binary(p)
        dg = ca.DeviceGraph.read_binary(p)
        try:
            dev = dg.download()
            assert dev.n == host.n and np.array_equal(dev.offsets, host.offsets)
            assert sorted_rows(dev) == sorted_rows(host)
            lab, _ = dg.find_components(variant="exp", seed=1001)
            assert np.array_equal(lab, oracle.canonical_labels(host.offsets, host.targets))
        finally:
            dg.close()
    bad = tmp_path / "bad.bin"
    np.array([3, 2, 1, 2, 2, 4], dtype=np.uint64).tofile(bad)  # id 4 outside [1, 3]
    with pytest.raises(ca.KcError) as e:
        ca.DeviceGraph.read_binary(bad)
    assert e.value.code == ca.KC_ERR_INVALID


def _sorted_rows(g):
    off = g.offsets.astype(np.int64)
    return [sorted(g.targets[off[v]:off[v + 1]].tolist()) for v in range(g.n)]


def test_device_metis_reader_matches_host(tmp_path):
    """kc_graph_read_metis against kc_read_metis on written graphs and on hand-made files with
    comments, CR line ends, blank runs, a missing last newline, fewer vertex lines than n,
    lines past the n-th (ignored, whatever they hold) and an empty file body."""
    files = []
    for name, kw in (("rmat", dict(n=1 << 14, m=1 << 18, seed=5)), ("gnm_undirected", dict(n=5000, m=3000, seed=2)),
                     ("grid_2d", dict(nx=60, ny=50, p=0.5, seed=1))):
        p = tmp_path / f"{name}.graph"
        ca.write_metis(p, ca.generate(name, **kw))
        files.append(p)
    hand = {
        "comments.graph": "% a comment\n%another 1 2 x\n6 7\n2 3\n% inside\n1 3\n1 2\n5  6\n4\t6\n4 5",
        "crlf.graph": "4 3\r\n2\r\n1 3 \r\n2 4\r\n3\r\n",
        "short.graph": "10 2\n2\n1\n\n5\n4\n",
        "extra.graph": "3 1\n2\n1\n\nthis line is past n: 7 x 0\n9 9 9\n",
        "nobody.graph": "% only a header\n5 0",
        "big_ids.graph": "3 2\n0000003\n\n1\n",
    }
    for fname, text in hand.items():
        p = tmp_path / fname
        p.write_bytes(text.encode())
        files.append(p)
    files.append(os.path.join(GOLDEN, "simple.graph"))
    for p in files:
        print(f"metis file {os.path.basename(str(p))}", flush=True)
        host = ca.read_metis(p)
        dg = ca.DeviceGraph.read_metis(p)
        try:
            dev = dg.download()
            assert dev.n == host.n and np.array_equal(dev.offsets, host.offsets), p
            assert _sorted_rows(dev) == _sorted_rows(host), p
            lab, _ = dg.find_components(variant="exp", seed=1001)
            assert np.array_equal(lab, oracle.canonical_labels(host.offsets, host.targets)), p
        finally:
            dg.close()
    for text, code in (("3 1\n2\n1 x\n\n", ca.KC_ERR_IO), ("3 1\n2\n1 -3\n\n", ca.KC_ERR_IO),
                       ("3 1\n2\n1 0\n\n", ca.KC_ERR_INVALID), ("3 1\n4\n\n\n", ca.KC_ERR_INVALID),
                       ("3 1\n99999999999\n\n\n", ca.KC_ERR_INVALID), ("% nothing\n", ca.KC_ERR_IO)):
        p = tmp_path / "bad.graph"
        p.write_bytes(text.encode())
        with pytest.raises(ca.KcError) as e:
            ca.DeviceGraph.read_metis(p)
        assert e.value.code == code, text


def test_device_edgefile_readers_match_host(tmp_path):
    """kc_graph_read_edgefile and kc_graph_read_partitioned_edgefile against the host readers:
    generated graphs, comments, CRLF, blank lines, lines past the header's m (ignored by the
    plain reader, read by the partitioned one), a missing last newline, and the errors."""
    def edges_once(g):
        src = np.repeat(np.arange(g.n, dtype=np.int64), np.diff(g.offsets.astype(np.int64)))
        keep = src < g.targets
        return src[keep], g.targets.astype(np.int64)[keep]

    def write(path, n, u, v, header_m=None):
        lines = [f"p {n} {len(u) if header_m is None else header_m}"]
        lines += [f"e {a + 1} {b + 1}" for a, b in zip(u.tolist(), v.tolist())]
        path.write_text("\n".join(lines) + "\n")

    def same(dg, host):
        try:
            dev = dg.download()
            assert dev.n == host.n and np.array_equal(dev.offsets, host.offsets)
            assert _sorted_rows(dev) == _sorted_rows(host)
            lab, _ = dg.find_components(variant="exp", seed=1001)
            assert np.array_equal(lab, oracle.canonical_labels(host.offsets, host.targets))
        finally:
            dg.close()

    files = []
    for name, kw in (("rmat", dict(n=1 << 13, m=1 << 17, seed=5)), ("rgg_2d", dict(n=3000, r=0.03, seed=2))):
        g = ca.generate(name, **kw)
        u, v = edges_once(g)
        p = tmp_path / f"{name}.edges"
        write(p, g.n, u, v)
        files.append(p)
    hand = {"sem.edges": b"% c\r\np 6 4\r\ne 1 3\r\n% inner\r\ne 3 1\r\ne 2 2\r\n\r\ne 4 5\r\ne 5 6\r\nx y z\r\n",
            "tabs.edges": b"%x\np\t5 3\n\t e   1\t2  \n\ne 2 3 trailing words\ne 4 5",
            "empty.edges": b"p 7 0\n"}
    for fname, raw in hand.items():
        (tmp_path / fname).write_bytes(raw)
        files.append(tmp_path / fname)
    for p in files:
        print(f"edge file {os.path.basename(str(p))}", flush=True)
        same(ca.DeviceGraph.read_edgefile(p), ca.read_edgefile(p))
    g = ca.generate("rgg_2d", n=2000, r=0.03, seed=2)
    u, v = edges_once(g)
    cuts = [0, len(u) // 3, 2 * len(u) // 3, len(u)]
    for r in range(3):
        write(tmp_path / f"part_{r}", g.n, u[cuts[r]:cuts[r + 1]], v[cuts[r]:cuts[r + 1]], header_m=len(u))
    for parts in (1, 3):
        same(ca.DeviceGraph.read_edgefile(tmp_path / "part", parts=parts),
             ca.read_partitioned_edgefile(tmp_path / "part", parts=parts))
    (tmp_path / "bad_0").write_text(f"p {g.n} {len(u)}\n")
    (tmp_path / "bad_1").write_text(f"p {g.n + 1} {len(u)}\n")
    for path, parts, code in ((tmp_path / "bad", 2, ca.KC_ERR_INVALID),):
        with pytest.raises(ca.KcError) as e:
            ca.DeviceGraph.read_edgefile(path, parts=parts)
        assert e.value.code == code
    for text, code in (("p 3 1\ne 1 4\n", ca.KC_ERR_INVALID), ("p 3 1\ne 1\n", ca.KC_ERR_IO),
                       ("p 3 1\ne 1 x\n", ca.KC_ERR_IO), ("p 3\ne 1 2\n", ca.KC_ERR_IO),
                       ("% only\n", ca.KC_ERR_IO), ("p 3 1\ne 0 2\n", ca.KC_ERR_INVALID)):
        p = tmp_path / "bad.edges"
        p.write_text(text)
        with pytest.raises(ca.KcError) as e1:
            ca.read_edgefile(p)
        with pytest.raises(ca.KcError) as e2:
            ca.DeviceGraph.read_edgefile(p)
        assert e1.value.code == code and e2.value.code == code, text


# -single / -contraction / the default branch's local contraction, over a vertex-block distribution
@pytest.mark.parametrize("bfs", [False, True])
@pytest.mark.parametrize("blocks", [1, 5, 64, 1 << 20])
@pytest.mark.parametrize("mode", ["local", "single", "double"])
def test_pre_contraction_trace_matches_oracle(mode, blocks, bfs):
    graphs = [ca.generate("gnm_undirected", n=1 << 14, m=1 << 15, seed=3),
              ca.generate("rmat", n=1 << 13, m=1 << 16, seed=3)]
    for g in graphs:
        print(f"pre_contract={mode} blocks={blocks} bfs={bfs} n={g.n}", flush=True)
        run = oracle.Run(g.offsets, g.targets, variant="exp", seed=1005, seq_limit=64, keep_keys=True,
                         stream=ca.STREAM_HASHED, pre_contract=mode, pre_blocks=blocks)
        dg = ca.DeviceGraph(g)
        lab, st = dg.find_components(variant="exp", seed=1005, seq_limit=64, record_trace=True,
                                     stream=ca.STREAM_HASHED, use_bfs=bfs, pre_contract=mode, pre_blocks=blocks)
        tr = dg.trace()
        dg.close()
        want_levels = run.levels()
        assert len(tr) == len(want_levels)
        n_pre = {"local": 1, "single": 1, "double": 2}[mode]
        for lvl, ((info, keys, gid), winfo) in enumerate(zip(tr, want_levels)):
            # sizes and radius are order-independent; sweep counts are not (in-place sweeps)
            assert (info[0], info[1], info[3]) == (winfo[0], winfo[1], winfo[3]), f"level {lvl} info"
            if lvl < n_pre:
                assert info[2] == 0 and info[3] == 0, f"pre-contraction level {lvl}"
            wkeys, wgid = run.level_keys(lvl)
            assert np.array_equal(gid, wgid), f"level {lvl} gid"
            assert np.array_equal(keys, wkeys), f"level {lvl} keys"
        assert np.array_equal(lab, run.labels())
        assert np.array_equal(lab, oracle.canonical_labels(g.offsets, g.targets))
        assert st["radius_sum"] == sum(w[3] for w in want_levels)


def test_pre_contraction_option_checks():
    g = ca.generate("gnm_undirected", n=1 << 10, m=1 << 11, seed=1)
    dg = ca.DeviceGraph(g)
    try:
        with pytest.raises(ca.KcError):
            dg.find_components(pre_contract=7, pre_blocks=4)
        with pytest.raises(ca.KcError):
            dg.find_components(pre_contract="single", pre_blocks=0)
        with pytest.raises(ca.KcError):
            dg.find_components(variant="labels", pre_contract="single", pre_blocks=4)
        lab, _ = dg.find_components(variant="local", pre_contract="double", pre_blocks=4)
        assert np.array_equal(lab, oracle.canonical_labels(g.offsets, g.targets))
    finally:
        dg.close()


def test_app_single_and_contraction_flags():
    exe = os.path.join(ca.BIN_DIR, "exp_contraction")
    for flag in ("-single", "-contraction"):
        r = subprocess.run([exe, "-gen", "rmat", "-n", "15", "-m", "19", "-i", "2", "-pes", "8", flag, "-check"],
                           capture_output=True, text=True, timeout=300)
        assert r.returncode == 0, r.stdout + r.stderr
        assert "CHECK labels_bit_exact=yes" in r.stdout, flag
