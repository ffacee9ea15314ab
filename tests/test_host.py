"""libkacomp host side without a GPU: exported symbols, METIS I/O, generators,
error paths."""
import ctypes
import os
import re
import subprocess

import numpy as np
import pytest

import components_amd as ca
import oracle

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLDEN = os.path.join(ROOT, "tests", "golden")


def declared_symbols():
    text = open(os.path.join(ROOT, "include", "kacomp.h")).read()
    return sorted(set(re.findall(r"\b(kc_[a-z_0-9]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    L = ca.lib()
    syms = declared_symbols()
    assert len(syms) >= 15
    for s in syms:
        assert hasattr(L, s), s
    nm = subprocess.run(["nm", "-D", "--defined-only", ca.LIB_PATH], capture_output=True, text=True, check=True).stdout
    for s in syms:
        assert re.search(rf"\bT {s}$", nm, re.M), s
    assert set(syms) == set(ca._SIGNATURES), "ctypes signatures and header differ"


def test_kernel_classes_match_header():
    """The ctypes stats layout and class names follow the header's class table."""
    text = open(os.path.join(ROOT, "include", "kacomp.h")).read()
    count = int(re.search(r"#define KC_KERNEL_CLASSES (\d+)", text).group(1))
    assert len(ca.KERNEL_CLASSES) == count
    idx = {int(v): k.lower() for k, v in re.findall(r"#define KC_K_([A-Z]+) (\d+)", text)}
    assert sorted(idx) == list(range(count))
    assert [idx[i] for i in range(count)] == ca.KERNEL_CLASSES
    for f in ("kernel_ms", "kernel_launches", "kernel_bytes"):
        assert getattr(ca.Stats, f).size == count * 8


# ctypes mirrors of the header's structs (the Python side of the C ABI)
STRUCTS = {"kc_csr_host": "CsrHost", "kc_csr": "Csr", "kc_gen_params": "GenParams", "kc_options": "Options",
           "kc_stats": "Stats", "kc_bin_tuning": "BinTuning", "kc_batch_host": "BatchHost",
           "kc_batch_csr": "BatchCsr", "kc_batch_gen_params": "BatchGenParams", "kc_batch_tuning": "BatchTuning",
           "kc_batch_stats": "BatchStats", "kc_train_params": "TrainParams", "kc_train_stats": "TrainStats"}


def test_ctypes_structs_match_header(tmp_path):
    """Every ctypes struct has the size of its C struct and every field its C offset, as a C program
    compiled against include/kacomp.h prints them (a field added on one side only fails here)."""
    lines = ["#include <stddef.h>", "#include <stdio.h>", '#include "kacomp.h"', "int main(void) {"]
    for cname, pname in STRUCTS.items():
        lines.append(f'  printf("{cname} size %zu\\n", sizeof({cname}));')
        for field, _ in getattr(ca, pname)._fields_:
            lines.append(f'  printf("{cname} {field} %zu\\n", offsetof({cname}, {field}));')
    lines += ["  return 0;", "}"]
    src = tmp_path / "layout.c"
    src.write_text("\n".join(lines) + "\n")
    exe = tmp_path / "layout"
    subprocess.run(["gcc", "-std=c11", "-I", os.path.join(ROOT, "include"), str(src), "-o", str(exe)], check=True)
    out = subprocess.run([str(exe)], capture_output=True, text=True, check=True).stdout.split("\n")
    got = {}
    for line in out:
        if line:
            cname, what, val = line.split()
            got[cname, what] = int(val)
    for cname, pname in STRUCTS.items():
        cls = getattr(ca, pname)
        assert got[cname, "size"] == ctypes.sizeof(cls), cname
        for field, _ in cls._fields_:
            assert got[cname, field] == getattr(cls, field).offset, (cname, field)


def test_read_metis_simple():
    g = ca.read_metis(os.path.join(GOLDEN, "simple.graph"))
    assert g.n == 6
    assert list(g.offsets) == [0, 1, 2, 4, 6, 7, 8]
    assert list(g.targets) == [2, 2, 0, 1, 4, 5, 3, 3]


def test_metis_roundtrip(tmp_path):
    g = ca.generate("gnm_undirected", n=500, m=900, seed=4)
    p = tmp_path / "g.graph"
    ca.write_metis(p, g)
    h = ca.read_metis(p)
    assert np.array_equal(g.offsets, h.offsets) and np.array_equal(g.targets, h.targets)


def test_metis_semantics(tmp_path):
    p = tmp_path / "c.graph"
    # comment lines, a trailing 0 stops a line, an empty line is an isolated vertex,
    # missing lines are isolated, a one-sided listing is symmetrised
    p.write_text("% comment\n5 3 0\n2 0 4\n% inner comment\n1\n\n")
    g = ca.read_metis(p)
    assert g.n == 5
    assert list(g.offsets) == [0, 1, 2, 2, 2, 2]
    assert list(g.targets) == [1, 0]


def test_metis_errors(tmp_path):
    with pytest.raises(ca.KcError) as e:
        ca.read_metis(tmp_path / "missing.graph")
    assert e.value.code == ca.KC_ERR_IO
    p = tmp_path / "bad.graph"
    p.write_text("3 1\n4\n\n\n")
    with pytest.raises(ca.KcError) as e:
        ca.read_metis(p)
    assert e.value.code == ca.KC_ERR_INVALID


def _edges_once(g):
    """Every undirected edge of a CSR once, as (u, v) with u < v, by u."""
    src = np.repeat(np.arange(g.n, dtype=np.int64), np.diff(g.offsets.astype(np.int64)))
    tgt = g.targets.astype(np.int64)
    keep = src < tgt
    return src[keep], tgt[keep]


def _write_edgefile(path, n, u, v, header_m=None, extra=""):
    lines = [f"p {n} {len(u) if header_m is None else header_m}"]
    lines += [f"e {a + 1} {b + 1}" for a, b in zip(u.tolist(), v.tolist())]
    path.write_text("\n".join(lines) + "\n" + extra)


def test_edgefile_matches_generated_graph(tmp_path):
    g = ca.generate("gnm_undirected", n=700, m=1500, seed=6)
    u, v = _edges_once(g)
    p = tmp_path / "g.edges"
    _write_edgefile(p, g.n, u, v)
    h = ca.read_edgefile(p)
    assert h.n == g.n and np.array_equal(g.offsets, h.offsets) and np.array_equal(g.targets, h.targets)


def test_edgefile_semantics(tmp_path):
    p = tmp_path / "s.edges"
    # comments before the header and between edges, CRLF, one direction per
    # edge, a duplicate and a self-loop; the header's m = 4 bounds the lines read
    p.write_bytes(b"% c\r\np 6 4\r\ne 1 3\r\n% inner\r\ne 3 1\r\ne 2 2\r\ne 4 5\r\ne 5 6\r\n")
    g = ca.read_edgefile(p)
    assert g.n == 6
    assert list(g.offsets) == [0, 1, 1, 2, 3, 4, 4]
    assert list(g.targets) == [2, 0, 4, 3]


def test_edgefile_simple_fixture_labels(tmp_path):
    # the reference's example graph, as an edge file: same CSR, same components
    m = ca.read_metis(os.path.join(GOLDEN, "simple.graph"))
    u, v = _edges_once(m)
    p = tmp_path / "simple.edges"
    _write_edgefile(p, m.n, u, v)
    g = ca.read_edgefile(p)
    assert np.array_equal(g.offsets, m.offsets) and np.array_equal(g.targets, m.targets)
    assert list(oracle.canonical_labels(g.offsets, g.targets)) == [0, 0, 0, 3, 3, 3]


def test_partitioned_edgefile(tmp_path):
    g = ca.generate("rgg_2d", n=2000, r=0.03, seed=2)
    u, v = _edges_once(g)
    cuts = [0, len(u) // 3, 2 * len(u) // 3, len(u)]
    for r in range(3):
        _write_edgefile(tmp_path / f"part_{r}", g.n, u[cuts[r]:cuts[r + 1]], v[cuts[r]:cuts[r + 1]],
                        header_m=len(u))
    h = ca.read_partitioned_edgefile(tmp_path / "part", parts=3)
    assert np.array_equal(g.offsets, h.offsets) and np.array_equal(g.targets, h.targets)
    one = ca.read_partitioned_edgefile(tmp_path / "part", parts=1)  # one rank's share alone
    assert one.m == cuts[1]
    (tmp_path / "bad_0").write_text(f"p {g.n} {len(u)}\n")
    (tmp_path / "bad_1").write_text(f"p {g.n + 1} {len(u)}\n")
    with pytest.raises(ca.KcError) as e:
        ca.read_partitioned_edgefile(tmp_path / "bad", parts=2)
    assert e.value.code == ca.KC_ERR_INVALID


def test_binary_roundtrip_and_layout(tmp_path):
    g = ca.generate("gnm_undirected", n=900, m=2500, seed=8)
    p = tmp_path / "g.bin"
    ca.write_binary(p, g)
    raw = np.fromfile(p, dtype=np.uint64)
    u, v = _edges_once(g)
    assert raw[0] == g.n and raw[1] == len(u)
    assert np.array_equal(raw[2::2], u + 1) and np.array_equal(raw[3::2], v + 1)
    h = ca.read_binary(p)
    assert np.array_equal(g.offsets, h.offsets) and np.array_equal(g.targets, h.targets)


def test_binary_and_edgefile_errors(tmp_path):
    p = tmp_path / "t.bin"
    np.array([4, 3, 1, 2, 2, 3], dtype=np.uint64).tofile(p)  # 3 edges announced, 2 present
    with pytest.raises(ca.KcError) as e:
        ca.read_binary(p)
    assert e.value.code == ca.KC_ERR_IO
    np.array([4, 1, 1, 5], dtype=np.uint64).tofile(p)  # id 5 outside [1, 4]
    with pytest.raises(ca.KcError) as e:
        ca.read_binary(p)
    assert e.value.code == ca.KC_ERR_INVALID
    with pytest.raises(ca.KcError) as e:
        ca.read_binary(tmp_path / "missing.bin")
    assert e.value.code == ca.KC_ERR_IO
    q = tmp_path / "t.edges"
    q.write_text("p 3 1\ne 1 0\n")
    with pytest.raises(ca.KcError) as e:
        ca.read_edgefile(q)
    assert e.value.code == ca.KC_ERR_INVALID
    q.write_text("p 3\n")
    with pytest.raises(ca.KcError) as e:
        ca.read_edgefile(q)
    assert e.value.code == ca.KC_ERR_IO


def check_csr(g):
    n = g.n
    off = g.offsets.astype(np.int64)
    assert off[0] == 0 and np.all(np.diff(off) >= 0)
    src = np.repeat(np.arange(n), np.diff(off))
    tgt = g.targets.astype(np.int64)
    assert np.all(tgt < n)
    assert not np.any(src == tgt)
    fwd = np.sort(src * n + tgt)
    rev = np.sort(tgt * n + src)
    assert np.array_equal(fwd, rev)            # symmetric
    assert len(np.unique(fwd)) == len(fwd)     # no duplicates
    for v in range(0, n, max(1, n // 50)):
        row = g.targets[off[v]:off[v + 1]]
        assert np.all(np.diff(row.astype(np.int64)) > 0)


@pytest.mark.parametrize("name,kw", [
    ("gnm_undirected", dict(n=4096, m=16384, seed=1)),
    ("rmat", dict(n=1 << 12, m=1 << 16, seed=3)),
    ("grid_2d", dict(nx=64, ny=50, p=0.6, seed=2)),
    ("grid_3d", dict(nx=16, ny=16, nz=16, p=0.3, seed=2)),
    ("rgg_2d", dict(n=4096, r=0.03, seed=2)),
    ("rgg_3d", dict(n=4096, r=0.1, seed=2)),
    ("rhg", dict(n=4096, gamma=2.8, d=16.0, seed=5)),
    ("ba", dict(n=3000, d=4.0, seed=6)),
])
def test_generators_valid_and_deterministic(name, kw):
    g = ca.generate(name, **kw)
    check_csr(g)
    h = ca.generate(name, **kw)
    assert np.array_equal(g.offsets, h.offsets) and np.array_equal(g.targets, h.targets)
    kw2 = dict(kw, seed=kw["seed"] + 1)
    k = ca.generate(name, **kw2)
    assert not (np.array_equal(g.offsets, k.offsets) and np.array_equal(g.targets, k.targets))


def test_generator_sizes():
    g = ca.generate("grid_2d", nx=30, ny=20, p=1.0, seed=1)
    assert g.n == 600 and g.m == 29 * 20 + 30 * 19
    g = ca.generate("grid_3d", nx=8, ny=8, nz=8, p=1.0, seed=1)
    assert g.n == 512 and g.m == 3 * 8 * 8 * 7
    g = ca.generate("grid_2d", nx=10, ny=10, p=1.0, periodic=1, seed=1)
    assert g.m == 200
    g = ca.generate("gnm_undirected", n=1 << 16, m=1 << 18, seed=1)
    assert (1 << 18) - 200 <= g.m <= (1 << 18)
    g = ca.generate("rgg_2d", n=1 << 14, r=np.sqrt(8 / (np.pi * (1 << 14))), seed=2)
    assert 6.5 < 2 * g.m / g.n < 8.5  # average degree about 8 (boundary loss)
    g = ca.generate("grid_3d", nx=32, ny=32, nz=32, p=0.3, seed=2)
    assert abs(g.m - 0.3 * 3 * 32 * 32 * 31) < 600


def test_rmat_skew():
    g = ca.generate("rmat", n=1 << 14, m=1 << 18, seed=3)
    deg = np.diff(g.offsets.astype(np.int64))
    assert deg.max() > 50 * max(1.0, deg.mean())
    assert deg[0] == deg.max() or deg[:64].max() == deg.max()  # hubs at low ids (quadrant a)


def test_generator_errors():
    with pytest.raises(ca.KcError) as e:
        ca.generate("rdg_2d", n=100)
    assert e.value.code == ca.KC_ERR_UNSUPPORTED
    with pytest.raises(ca.KcError) as e:
        ca.generate("nonsense", n=100)
    assert e.value.code == ca.KC_ERR_UNSUPPORTED
    with pytest.raises(ca.KcError):
        ca.generate("rmat", n=1000, m=100)


def test_options_defaults_match_reference():
    o = ca.options()
    assert o.seed == 1 and abs(o.beta - 0.1) < 1e-15 and o.variant == 0 and o.deviate_fma == 1


def test_no_device_is_a_clean_error():
    # without a GPU the device entry points fail with KC_ERR_HIP rather than crashing
    if os.environ.get("HIP_VISIBLE_DEVICES") == "" or not os.path.exists("/dev/kfd"):
        g = ca.Graph.from_edges(3, [(0, 1)])
        with pytest.raises(ca.KcError) as e:
            ca.DeviceGraph(g)
        assert e.value.code == ca.KC_ERR_HIP
    else:
        pytest.skip("a GPU is present")


def test_cli_flag_parsing_without_device():
    # the program prints the reference's CONFIG block before touching the device
    exe = os.path.join(ca.BIN_DIR, "exp_contraction")
    r = subprocess.run([exe, "-gen", "gnm_undirected", "-n", "8", "-m", "9", "-i", "1", "-beta", "0.2", "-seed", "5"],
                       capture_output=True, text=True, timeout=120)
    assert "CONFIG" in r.stdout
    assert re.search(r"beta\s+0.2", r.stdout) and re.search(r"seed\s+5", r.stdout)
    assert re.search(r"\bn\s+256\b", r.stdout) and re.search(r"\bm\s+512\b", r.stdout)
    if not os.path.exists("/dev/kfd"):
        assert r.returncode != 0
    # -n and -m are log2 sizes: a shift past 62 is refused before anything runs
    r = subprocess.run([os.path.join(ca.BIN_DIR, "all_reduce"), "-gen", "rmat", "-n", "1048576", "-i", "1"],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 1 and "log2 sizes" in r.stderr


def test_batch_generator_valid_and_deterministic():
    b = ca.batch_generate(graphs=64, log_n_min=2, log_n_max=10, deg_min=1, deg_max=12, seed=3)
    b2 = ca.batch_generate(graphs=64, log_n_min=2, log_n_max=10, deg_min=1, deg_max=12, seed=3)
    assert np.array_equal(b.offsets, b2.offsets) and np.array_equal(b.targets, b2.targets)
    assert b.graphs == 64 and b.vtx_off[0] == 0 and b.edge_off[0] == 0
    assert np.all(np.diff(b.vtx_off.astype(np.int64)) >= 1)
    for g in range(b.graphs):
        gr = b.graph(g)
        assert gr.n >= 2 and gr.n <= 2 ** 10 + 64
        assert int(b.offsets[int(b.vtx_off[g])]) == int(b.edge_off[g])
        src = np.repeat(np.arange(gr.n), np.diff(gr.offsets).astype(np.int64))
        assert np.all(gr.targets < gr.n) and np.all(src != gr.targets)
        a = set(zip(src.tolist(), gr.targets.tolist()))
        assert a == {(v, u) for u, v in a}, f"graph {g} not symmetric"
    b3 = ca.batch_generate(graphs=64, log_n_min=2, log_n_max=10, deg_min=1, deg_max=12, seed=4)
    assert not np.array_equal(b.vtx_off, b3.vtx_off)


def test_batch_generator_types():
    # one type at a time: G(n,m) keeps n exactly; grids have side^dims vertices
    b = ca.batch_generate(graphs=16, log_n_min=6, log_n_max=6, deg_min=4, deg_max=4, types=1, seed=1)
    assert np.all(np.diff(b.vtx_off.astype(np.int64)) == 64)
    g = ca.batch_generate(graphs=16, log_n_min=6, log_n_max=6, deg_min=4, deg_max=4, types=4, seed=1)
    assert set(np.diff(g.vtx_off.astype(np.int64)).tolist()) <= {64, 64}  # 8x8 and 4x4x4
    with pytest.raises(ca.KcError):
        ca.batch_generate(graphs=4, types=0)
    with pytest.raises(ca.KcError):
        ca.batch_generate(graphs=4, log_n_min=5, log_n_max=3)


def test_batch_create_rejects_bad_input():
    good = ca.BatchGraphs.pack([ca.Graph.from_edges(3, [(0, 1)]), ca.Graph.from_edges(2, [(0, 1)])])
    bad_target = ca.BatchGraphs(good.vtx_off, good.edge_off, good.offsets, good.targets.copy())
    bad_target.targets[-1] = 2  # graph 1 has two vertices
    with pytest.raises(ca.KcError) as e:
        ca.DeviceBatch(bad_target)
    assert e.value.code == ca.KC_ERR_INVALID and "out of range" in str(e.value)
    bad_off = ca.BatchGraphs(good.vtx_off, good.edge_off + np.uint64(1), good.offsets, good.targets)
    with pytest.raises(ca.KcError) as e:
        ca.DeviceBatch(bad_off)
    assert e.value.code == ca.KC_ERR_INVALID


def test_create_rejects_one_way_edges():
    # Both contraction paths read every cluster edge from both of its ends, so raw CSR handed over
    # by a caller must hold each edge in both rows; a one-way edge is refused before any device work.
    one_way = ca.Graph(np.array([0, 1, 1, 2], np.uint64), np.array([1, 1], np.uint32))  # 0->1, 2->1
    with pytest.raises(ca.KcError) as e:
        ca.DeviceGraph(one_way)
    assert e.value.code == ca.KC_ERR_INVALID and "not symmetric" in str(e.value) and "0 -> 1" in str(e.value)
    # a larger graph with a single reverse slot redirected
    g = ca.generate("gnm_undirected", n=1 << 12, m=1 << 14, seed=5)
    tgt = g.targets.copy()
    row = int(np.flatnonzero(np.diff(g.offsets.astype(np.int64)) >= 2)[7])
    a = int(g.offsets[row])
    tgt[a] = tgt[a + 1]  # the row now names one neighbour twice and loses the other
    with pytest.raises(ca.KcError) as e:
        ca.DeviceGraph(ca.Graph(g.offsets, tgt))
    assert e.value.code == ca.KC_ERR_INVALID and "not symmetric" in str(e.value)
    # packed batches are checked per graph, with local ids
    good = ca.BatchGraphs.pack([ca.Graph.from_edges(3, [(0, 1), (1, 2)]), ca.Graph.from_edges(3, [(0, 2)])])
    bad = ca.BatchGraphs(good.vtx_off, good.edge_off, good.offsets, good.targets.copy())
    e0 = int(good.edge_off[1])
    bad.targets[e0 + 1] = 1  # graph 1: 2 -> 1 instead of 2 -> 0
    with pytest.raises(ca.KcError) as e:
        ca.DeviceBatch(bad)
    assert e.value.code == ca.KC_ERR_INVALID and "graph 1 is not symmetric" in str(e.value)
    # symmetric input passes the check and fails only for want of a device here
    if not os.path.exists("/dev/kfd"):
        for build in (lambda: ca.DeviceGraph(g), lambda: ca.DeviceBatch(good)):
            with pytest.raises(ca.KcError) as e:
                build()
            assert e.value.code == ca.KC_ERR_HIP
