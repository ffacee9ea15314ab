"""Benchmark: exponential contraction connected components on one MI355X.

Workload: R-MAT scale 24, edge factor 16, (a,b,c,d) = (0.57,0.19,0.19,0.05),
generator seed 3, symmetrised, self-loops and duplicates removed, u32 ids,
u64 offsets; generated and built in HBM before the timed region.  A step is
one kc_find_components call (exponential contraction down to the base case,
undo, canonical min-id labels) on the resident graph; step i draws its
deviates with seed 1 + i + 1000 as the reference harness does
(app/exp_contraction.cpp:75-77).  value = GTEPS = m / step time.

Before the timed steps, --breakdown-steps untimed steps bracket every launch
with HIP events (the per-class breakdown); the timed steps bracket only the
launches of the class that took the most time there, whose average launch
duration gives the roofline.

Prints one JSON line (rank 0, N = 1).
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "GTEPS + contraction rounds, learned/batched vs plain, 1 MI355X, labels bit-exact"
MI355X_HBM_GBS = 8000.0
# kernel timer class -> the kernel a trace shows for it
# kernel classes made of one kernel: the roofline is reported for the one
# among them that takes the most time in a step
CLASS_KERNEL = {"round": "k_round", "relabel": "k_relabel_hot", "scatter": "k_scatter", "csr": "k_gather_uniques"}


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--scale", type=int, default=24)
    ap.add_argument("--edge-factor", type=int, default=16)
    ap.add_argument("--stream", choices=["hashed", "reference"], default="hashed")
    ap.add_argument("--schedule", choices=["rounds", "sweeps"], default="rounds",
                    help="rounds: time-stepped push (each edge pushed once); sweeps: in-place pull sweeps")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline and the label check")
    ap.add_argument("--cpu-scale", type=int, default=24,
                    help="R-MAT scale of the CPU baseline's graph (the bench workload's scale by default)")
    ap.add_argument("--cpu-ranks", type=int, default=0, help="MPI ranks of the CPU baseline (0: min(16, nproc))")
    ap.add_argument("--cpu-iterations", type=int, default=1, help="runs of the CPU baseline (seeds 1001, 1002, ...)")
    ap.add_argument("--tmax", type=int, default=0)
    ap.add_argument("--smax", type=int, default=0)
    ap.add_argument("--wmax", type=int, default=0)
    ap.add_argument("--no-autotune", action="store_true", help="keep the built-in degree-bin thresholds")
    ap.add_argument("--relabel-split", default="default",
                    help="owner-split relabel of the degree-ordered level 0: on, off (one pass over whole rows), "
                         "default, or the row length above which rows are split")
    ap.add_argument("--no-batched", action="store_true", help="skip the batched-form leg")
    ap.add_argument("--no-learned", action="store_true", help="skip the learned-form leg")
    ap.add_argument("--no-configs", action="store_true", help="skip the other plain workloads")
    ap.add_argument("--pull-rounds", type=int, default=0,
                    help="rounds schedule: push or pull chosen per early round (0 default, 1 every level, -1 never)")
    ap.add_argument("--tail-sweeps", type=int, default=0,
                    help="rounds schedule: pull sweeps once <= n>>k rows are pending (0 default k, -1 never)")
    ap.add_argument("--dedup-huge", default="auto", choices=["auto", "tables", "bitmaps", "bytemaps", "maps_tables"],
                    help="dedup of the huge cluster-edge buckets (auto: byte maps, bitmaps, tables by size)")
    ap.add_argument("--untimed-kernels", action="store_true",
                    help="time the steps without HIP events on the dominant kernel (no roofline)")
    ap.add_argument("--breakdown-steps", type=int, default=2,
                    help="untimed steps with every launch bracketed by HIP events (per-class breakdown)")
    ap.add_argument("--learned-rhg-log-n", type=int, default=24)
    ap.add_argument("--learned-rmat-scale", type=int, default=25)
    ap.add_argument("--batch-graphs", type=int, default=4096)
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only copy 1 GiB device-to-device three times (HBM counter calibration) and exit")
    return ap.parse_args()


def copy_calibration():
    """A plain device copy of known size: 2^30 bytes read and 2^30 written per
    launch, for scaling FETCH_SIZE / WRITE_SIZE counter readings."""
    import torch

    src = torch.ones(1 << 28, dtype=torch.int32, device="cuda")
    dst = torch.empty_like(src)
    for _ in range(3):
        dst.copy_(src)
    torch.cuda.synchronize()
    print(json.dumps({"copy_bytes_per_launch": 1 << 30, "launches": 3}), flush=True)


SPLIT_MODES = {"default": 0, "on": 1, "off": -1}
ORDER_MIN_VERTICES = 1 << 16  # level 0 of at least this many non-isolated vertices is stored in degree order


def read_traffic(kernel, workload, key="hbm_bytes_per_launch"):
    """HBM bytes per launch of `kernel` from the committed counter summary
    (profiles/traffic_<workload>.txt, lines `key: value`), or None: `key` is
    hbm_bytes_per_launch (the kernel's level-0 launch) or
    class_hbm_bytes_per_launch (the average over its class's launches in a call)."""
    path = os.path.join(ROOT, "profiles", f"traffic_{workload}.txt")
    if not os.path.exists(path):
        return None
    kv = {}
    with open(path) as f:
        for line in f:
            if ":" in line and not line.startswith("#"):
                k, v = line.split(":", 1)
                kv[k.strip()] = v.strip()
    if kv.get("kernel") != kernel:
        return None
    try:
        return int(float(kv[key]))
    except (KeyError, ValueError):
        return None


def main():
    a = parse()
    if a.copy_calibration:
        copy_calibration()
        return
    if a.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for --gpus {a.gpus}); multi-GPU sharding is a design "
              "only, not a measured path", file=sys.stderr)
        sys.exit(2)
    import numpy as np

    import __graft_entry__ as ge

    if not ge.is_built():
        ge.build()
    import components_amd as ca

    stream = ca.STREAM_HASHED if a.stream == "hashed" else ca.STREAM_REFERENCE
    gp = dict(n=1 << a.scale, m=a.edge_factor << a.scale, seed=3, rmat_a=0.57, rmat_b=0.19, rmat_c=0.19)
    t0 = time.time()
    dg = ca.DeviceGraph.generate("rmat", **gp)
    gen_s = time.time() - t0
    n, m = dg.n, dg.m
    seq_limit = max(m // 256, 4096)  # ComputeSequentialLimit with one PE per compute unit
    run_kw = dict(variant="exp", beta=0.1, seq_limit=seq_limit, stream=stream, tmax=a.tmax, smax=a.smax, wmax=a.wmax,
                  use_bfs=a.schedule == "rounds", tail_sweeps=a.tail_sweeps,
                  pull_rounds=a.pull_rounds, dedup_huge=a.dedup_huge,
                  relabel_split=SPLIT_MODES[a.relabel_split] if a.relabel_split in SPLIT_MODES else int(a.relabel_split))
    print(f"[bench] rmat scale {a.scale}: n={n} m={m} non_isolated={dg.non_isolated} built in {gen_s:.1f}s",
          file=sys.stderr, flush=True)

    # degree-bin thresholds tuned on this graph (kept in the handle for every call below that leaves
    # them at 0); results do not depend on them
    bins = None
    if not (a.tmax or a.smax or a.wmax) and not a.no_autotune:
        bins = dg.autotune(seed=1001, **run_kw)
        print(f"[bench] bins tuned: {bins}", file=sys.stderr, flush=True)
    for i in range(a.warmup):
        dg.find_components(copy_labels=False, seed=1 + i + 1000, **run_kw)
    # per-kernel-class breakdown, outside the timed region: every launch
    # bracketed by HIP events on the library's stream (their recording slows
    # the host's launches, so the timed steps below carry fewer events)
    brk = [dg.find_components(copy_labels=False, seed=1 + i + 1000, time_kernels=True, **run_kw)[1]
           for i in range(a.breakdown_steps)]
    nb = max(len(brk), 1)
    cls_ms = {c: sum(s["kernel_ms"][c] for s in brk) for c in ca.KERNEL_CLASSES}
    cls_bytes = {c: sum(s["kernel_bytes"][c] for s in brk) for c in ca.KERNEL_CLASSES}
    dom = max(CLASS_KERNEL, key=lambda c: cls_ms[c]) if brk else "relabel"
    kernel = CLASS_KERNEL.get(dom, dom)
    # timed region: HIP events around the dominant kernel's launches only
    stats = []
    t0 = time.perf_counter()
    for i in range(a.steps):
        _, st = dg.find_components(copy_labels=False, seed=1 + i + 1000, time_kernels=not a.untimed_kernels,
                                   time_class=dom, **run_kw)
        stats.append(st)
    wall = time.perf_counter() - t0
    ms_per_step = wall * 1e3 / a.steps
    dev_ms = float(np.mean([s["time_ms"] for s in stats]))
    value = m / (ms_per_step / 1e3) / 1e9

    # the dominant kernel's launches over the timed steps
    launches = max(sum(s["kernel_launches"][dom] for s in stats), 1)
    bytes_per_launch = sum(s["kernel_bytes"][dom] for s in stats) / launches
    avg_launch_ms = sum(s["kernel_ms"][dom] for s in stats) / launches
    achieved = bytes_per_launch / (avg_launch_ms / 1e3) / 1e9 if avg_launch_ms > 0 else 0.0
    # the first level's launch alone (the levels' launches differ in shape: level 0 is the large one)
    first_ms = float(np.mean([s["kernel_first_ms"][dom] for s in stats]))
    first_bytes = float(np.mean([s["kernel_first_bytes"][dom] for s in stats]))
    first_gbs = first_bytes / (first_ms / 1e3) / 1e9 if first_ms > 0 else 0.0
    total_kernel_ms = sum(cls_ms.values())
    all_bytes = sum(cls_bytes.values())
    workload = f"rmat{a.scale}_{a.schedule}"
    # The roofline is that of the class's level-0 launch, which holds nearly all of its time: on a
    # degree-ordered level 0 the relabel's is the owner-split kernel, launched once per step (the later
    # levels' relabels are k_relabel_hot); the class average over every launch is reported beside it.
    ordered = dg.non_isolated >= ORDER_MIN_VERTICES
    kernel0 = "k_relabel_split" if dom == "relabel" and ordered and a.relabel_split != "off" else kernel
    traffic0 = read_traffic(kernel0, workload)
    traffic_all = read_traffic(kernel0, workload, "class_hbm_bytes_per_launch")

    out = {
        "metric": METRIC,
        "value": round(value, 4),
        "unit": "GTEPS",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": round(ms_per_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u32/u64 (integer)",
        "data": "synthetic (device-generated R-MAT, counter-based SplitMix64 streams)",
        "config": {"workload": f"rmat-scale{a.scale}-ef{a.edge_factor}-abcd(0.57,0.19,0.19,0.05)-seed3",
                   "n": n, "m": m, "non_isolated": dg.non_isolated, "form": "plain",
                   "deviate_stream": a.stream, "schedule": a.schedule, "beta": 0.1, "seq_limit": seq_limit,
                   "relabel": "one-pass" if a.relabel_split == "off" else "owner-split",
                   "bins": {"small": stats[-1]["bin_small"], "mid": stats[-1]["bin_mid"],
                            "large": stats[-1]["bin_large"],
                            "tuned": None if bins is None else {k: (round(v, 4) if isinstance(v, float) else v)
                                                                for k, v in bins.items()}},
                   "parallelism": "1 GPU"},
        "levels": stats[-1]["levels"],
        "sweeps": stats[-1]["sweeps"],
        "radius_sum": stats[-1]["radius_sum"],
        "device_ms_per_step": round(dev_ms, 4),
        "roofline": {"bound": "hbm", "kernel": kernel0, "achieved": round(first_gbs, 2), "peak": MI355X_HBM_GBS,
                     "unit": "GB/s", "frac": round(first_gbs / MI355X_HBM_GBS, 4), "traffic": traffic0,
                     "alg_bytes_per_launch": int(first_bytes), "avg_launch_ms": round(first_ms, 5),
                     "launches_per_step": 1.0,
                     "class_all_launches": {"class": dom, "kernels": kernel0 if kernel0 == kernel else
                                            f"{kernel0} (level 0) + {kernel} (later levels)",
                                            "achieved": round(achieved, 2),
                                            "frac": round(achieved / MI355X_HBM_GBS, 4), "traffic": traffic_all,
                                            "alg_bytes_per_launch": int(bytes_per_launch),
                                            "avg_launch_ms": round(avg_launch_ms, 5),
                                            "launches_per_step": round(launches / a.steps, 2)}},
        "whole_step": {"alg_bytes": int(all_bytes / nb), "kernel_ms": round(total_kernel_ms / nb, 4),
                       "alg_gbs": round(all_bytes / (total_kernel_ms / 1e3) / 1e9, 2) if total_kernel_ms else 0.0,
                       "breakdown_steps": len(brk)},
        "kernel_ms_per_step": {c: round(cls_ms[c] / nb, 4) for c in ca.KERNEL_CLASSES if cls_ms[c] > 0},
        "cpu_baseline": None,
    }

    out["forms"] = forms_leg(a, dg, run_kw, m)
    if not a.no_cpu:
        out["cpu_baseline"], out["labels_bit_exact"] = cpu_leg(a, dg, run_kw, stream)
    dg.close()
    if not a.no_configs:
        out["plain_configs"] = configs_leg(a, stream, run_kw)
    if not a.no_batched:
        out["batched"] = batched_leg(a, stream)
    if not a.no_learned:
        out["learned"] = learned_leg(a, stream, run_kw)
    print(json.dumps(out), flush=True)


def forms_leg(a, dg, run_kw, m):
    """The headline workload under every (deviate stream, schedule) pair: the
    reference's own draw (exponential_contraction.h:467-471) and the hashed
    one, the in-place pull sweeps of RunContractionLP's default loop (:518-586)
    and the time-stepped rounds.  Both schedules reach the same fixed point, so
    levels and labels depend on the stream only."""
    import numpy as np

    out = {}
    for stream_name, st_id in (("reference", 0), ("hashed", 1)):
        for sched in ("rounds", "sweeps"):
            kw = dict(run_kw, stream=st_id, use_bfs=sched == "rounds")
            for i in range(a.warmup):
                dg.find_components(copy_labels=False, seed=1 + i + 1000, **kw)
            sts = []
            t0 = time.perf_counter()
            for i in range(a.steps):
                sts.append(dg.find_components(copy_labels=False, seed=1 + i + 1000, **kw)[1])
            ms = (time.perf_counter() - t0) * 1e3 / a.steps
            out[f"{stream_name}/{sched}"] = {
                "value": round(m / (ms / 1e3) / 1e9, 4), "unit": "GTEPS", "ms_per_step": round(ms, 4),
                "levels": float(np.mean([s["levels"] for s in sts])),
                "sweeps": float(np.mean([s["sweeps"] for s in sts])),
                "radius_sum": float(np.mean([s["radius_sum"] for s in sts]))}
            print(f"[bench] forms: {stream_name}/{sched}: {ms:.2f} ms, {out[f'{stream_name}/{sched}']['levels']} "
                  f"levels", file=sys.stderr, flush=True)
    return out


def configs_leg(a, stream, run_kw):
    """The plain form on the other workloads: G(n,m) n = 2^20, m = 2^22, seed 1;
    3D grid 256^3 with every bond kept with probability 0.3, seed 2; 2D random
    geometric graph n = 2^24 at average degree 8, seed 2.  Labels are checked
    against the oracle's union-find."""
    import numpy as np

    import components_amd as ca
    import oracle

    out = {}
    cases = [("gnm-n2^20-m2^22-seed1", "device", "gnm_undirected", dict(n=1 << 20, m=1 << 22, seed=1)),
             ("grid3d-256^3-p0.3-seed2", "device", "grid_3d", dict(nx=256, ny=256, nz=256, p=0.3, seed=2)),
             ("rgg2d-n2^24-deg8-seed2", "host", "rgg_2d",
              dict(n=1 << 24, r=float(np.sqrt(8 / (np.pi * (1 << 24)))), seed=2))]
    for label, where, name, kw in cases:
        t0 = time.time()
        if where == "device":
            dg = ca.DeviceGraph.generate(name, **kw)
            host = None
        else:
            host = ca.generate(name, **kw)
            dg = ca.DeviceGraph(host)
        m = dg.m
        seq = max(m // 256, 4096)
        res = {"n": dg.n, "m": m}
        for sched, bfs in (("rounds", True), ("sweeps", False)):
            kw2 = dict(run_kw, seq_limit=seq, use_bfs=bfs)
            for i in range(a.warmup):
                dg.find_components(copy_labels=False, seed=1 + i + 1000, **kw2)
            sts = []
            t1 = time.perf_counter()
            for i in range(a.steps):
                sts.append(dg.find_components(copy_labels=False, seed=1 + i + 1000, **kw2)[1])
            ms = (time.perf_counter() - t1) * 1e3 / a.steps
            res[sched] = {"value": round(m / (ms / 1e3) / 1e9, 4), "unit": "GTEPS", "ms_per_step": round(ms, 4),
                          "levels": float(np.mean([s["levels"] for s in sts])),
                          "sweeps": float(np.mean([s["sweeps"] for s in sts])),
                          "radius_sum": float(np.mean([s["radius_sum"] for s in sts]))}
        lab = dg.find_components(seed=1001, **dict(run_kw, seq_limit=seq))[0]
        if host is None:
            host = dg.download()
        res["labels_bit_exact"] = bool(np.array_equal(lab, oracle.canonical_labels(host.offsets, host.targets)))
        dg.close()
        del host
        out[label] = res
        print(f"[bench] {label}: rounds {res['rounds']['ms_per_step']:.2f} ms, sweeps {res['sweeps']['ms_per_step']:.2f} ms,"
              f" exact {res['labels_bit_exact']} ({time.time() - t0:.1f}s)", file=sys.stderr, flush=True)
    return out


def batched_leg(a, stream):
    """The batched form on the mixed batch (4096 graphs: n log-uniform in
    [2^8, 2^18], average degree uniform in [2, 16], G(n,m) / 2D geometric /
    percolated grid, seed 4), beside the plain path on the same graphs taken as
    one disjoint union.  Labels of every graph are checked against the oracle's
    union-find labels of the union."""
    import numpy as np

    import components_amd as ca
    import oracle

    t0 = time.time()
    b = ca.batch_generate(graphs=a.batch_graphs, seed=4)
    gen_s = time.time() - t0
    n, m = int(b.vtx_off[-1]), int(b.edge_off[-1]) // 2
    print(f"[bench] batch: {b.graphs} graphs, n={n} m={m} generated in {gen_s:.1f}s", file=sys.stderr, flush=True)
    db = ca.DeviceBatch(b)
    tune = db.autotune(seed=1001, stream=stream)
    db.set_tuning(block=tune["block"], groups_per_cu=tune["groups_per_cu"], use_graph=True,
                  split_vertices=tune["split_vertices"])
    print(f"[bench] batch: tuned {tune}", file=sys.stderr, flush=True)
    for i in range(a.warmup):
        db.run(copy_labels=False, seed=1 + i + 1000, stream=stream)
    stats = []
    t0 = time.perf_counter()
    for i in range(a.steps):
        _, st = db.run(copy_labels=False, seed=1 + i + 1000, stream=stream)
        stats.append(st)
    ms = (time.perf_counter() - t0) * 1e3 / a.steps
    dev_ms = float(np.mean([s["time_ms"] for s in stats]))
    lab, st = db.run(seed=1001, stream=stream)
    db.close()
    # the disjoint union with global ids: oracle labels and the plain path
    per_vertex_base = np.repeat(b.vtx_off[:-1].astype(np.uint32), np.diff(b.vtx_off).astype(np.int64))
    tg = b.targets + np.repeat(b.vtx_off[:-1].astype(np.uint32), np.diff(b.edge_off).astype(np.int64))
    ref = oracle.canonical_labels(b.offsets, tg) - per_vertex_base
    exact = bool(np.array_equal(lab, ref))
    del ref
    print(f"[bench] batch: {ms:.2f} ms/step, labels bit-exact {exact}", file=sys.stderr, flush=True)
    dg = ca.DeviceGraph(ca.Graph(b.offsets, tg))
    del tg
    seq = max(m // 256, 4096)
    plain = []
    for i in range(a.warmup):
        dg.find_components(copy_labels=False, seed=1 + i + 1000, seq_limit=seq, stream=stream, use_bfs=True)
    t0 = time.perf_counter()
    for i in range(a.steps):
        plain.append(dg.find_components(copy_labels=False, seed=1 + i + 1000, seq_limit=seq, stream=stream,
                                        use_bfs=True)[1])
    pms = (time.perf_counter() - t0) * 1e3 / a.steps
    dg.close()
    return {"workload": f"batch-{b.graphs}-graphs-n2^8..2^18-deg2..16-gnm/rgg2d/grid-seed4", "graphs": b.graphs,
            "n": n, "m": m, "value": round(m / (ms / 1e3) / 1e9, 4), "unit": "GTEPS", "ms_per_step": round(ms, 4),
            "device_ms_per_step": round(dev_ms, 4), "levels_sum": stats[-1]["levels_sum"],
            "max_levels": stats[-1]["max_levels"], "sweeps_sum": stats[-1]["sweeps_sum"],
            "radius_sum": stats[-1]["radius_sum"], "tuning": stats[-1]["tuning"], "labels_bit_exact": exact,
            "plain_union": {"value": round(m / (pms / 1e3) / 1e9, 4), "unit": "GTEPS", "ms_per_step": round(pms, 4),
                            "levels": plain[-1]["levels"], "sweeps": plain[-1]["sweeps"],
                            "schedule": "rounds", "seq_limit": seq},
            "generation_s": round(gen_s, 1)}


def learned_leg(a, stream, run_kw):
    """Learned centre proposals: the MLP is trained on traces of the plain
    method on the plain workloads (G(n,m) 2^20/2^22, percolated 3D grid 128^3
    p = 0.3, 2D geometric 2^20 at average degree 8, R-MAT scale 22), then the
    plain and learned forms run at the same seeds on the hyperbolic graph
    n = 2^24, gamma 2.8, average degree 16, and on R-MAT scale 25."""
    import numpy as np

    import components_amd as ca
    import oracle

    t0 = time.time()
    mlp = ca.Mlp(seed=1)
    train = [("gnm_undirected", dict(n=1 << 20, m=1 << 22, seed=1), 16),
             ("grid_3d", dict(nx=128, ny=128, nz=128, p=0.3, seed=2), 16),
             ("rmat", dict(n=1 << 22, m=16 << 22, seed=3, rmat_a=0.57, rmat_b=0.19, rmat_c=0.19), 16)]
    for name, kw, stride in train:
        g = ca.DeviceGraph.generate(name, **kw)
        mlp.collect(g, stride=stride, seed=11, seq_limit=0, stream=stream)
        g.close()
    rgg = ca.generate("rgg_2d", n=1 << 20, r=float(np.sqrt(8 / (np.pi * (1 << 20)))), seed=2)
    g = ca.DeviceGraph(rgg)
    mlp.collect(g, stride=16, seed=11, seq_limit=0, stream=stream)
    g.close()
    tr = mlp.train(steps=600, batch=8192, lr=3e-3)
    X, y = mlp.dataset()
    const = 0.5 * float(np.var(y.astype(np.float64)))
    _, full_loss = mlp.gradients(X, y)
    del X, y
    train_s = time.time() - t0
    print(f"[bench] learned: {tr['samples']} trace rows, loss {full_loss:.3f} (constant {const:.3f}), "
          f"{train_s:.1f}s", file=sys.stderr, flush=True)
    evals = {}
    graphs = [("rhg-n2^%d-gamma2.8-deg16-seed5" % a.learned_rhg_log_n, "host", "rhg",
               dict(n=1 << a.learned_rhg_log_n, gamma=2.8, d=16.0, seed=5)),
              ("rmat-scale%d-ef16-seed3" % a.learned_rmat_scale, "device", "rmat",
               dict(n=1 << a.learned_rmat_scale, m=16 << a.learned_rmat_scale, seed=3, rmat_a=0.57, rmat_b=0.19,
                    rmat_c=0.19))]
    for label, where, name, kw in graphs:
        t1 = time.time()
        if where == "host":
            host = ca.generate(name, **kw)
            dg = ca.DeviceGraph(host)
        else:
            host = None
            dg = ca.DeviceGraph.generate(name, **kw)
        dg.set_mlp(mlp)
        m = dg.m
        seq = max(m // 256, 4096)
        res = {"n": dg.n, "m": m}
        labels = {}
        # learned: level 0's scores (a function of the graph and the model) computed by the first call
        # and reused; learned_cold: computed again in every step (the model re-attached before it)
        for form, learned in (("plain", False), ("learned", True), ("learned_cold", True)):
            kw2 = dict(run_kw, seq_limit=seq, learned=learned)
            for i in range(a.warmup):
                dg.find_components(copy_labels=False, seed=1 + i + 1000, **kw2)
            sts = []
            t2 = time.perf_counter()
            for i in range(a.steps):
                if form == "learned_cold":
                    dg.set_mlp(mlp)
                sts.append(dg.find_components(copy_labels=False, seed=1 + i + 1000, **kw2)[1])
            ms = (time.perf_counter() - t2) * 1e3 / a.steps
            labels[form] = dg.find_components(seed=1001, **kw2)[0]
            res[form] = {"value": round(m / (ms / 1e3) / 1e9, 4), "unit": "GTEPS", "ms_per_step": round(ms, 4),
                         "levels": float(np.mean([s["levels"] for s in sts])),
                         "sweeps": float(np.mean([s["sweeps"] for s in sts])),
                         "radius_sum": float(np.mean([s["radius_sum"] for s in sts]))}
        if host is None:
            host = dg.download()
        ref = oracle.canonical_labels(host.offsets, host.targets)
        res["labels_bit_exact"] = bool(all(np.array_equal(x, ref) for x in labels.values()))
        del ref, host
        dg.close()
        evals[label] = res
        print(f"[bench] learned: {label}: plain {res['plain']['ms_per_step']:.2f} ms {res['plain']['levels']} levels, "
              f"learned {res['learned']['ms_per_step']:.2f} ms {res['learned']['levels']} levels, "
              f"cold {res['learned_cold']['ms_per_step']:.2f} ms ({time.time() - t1:.1f}s)", file=sys.stderr, flush=True)
    mlp.close()
    return {"model": "MLP 6-64-64-1, bf16 weights, f32 accumulation (MFMA 16x16x16 bf16)",
            "training": {"trace_rows": tr["samples"], "steps": tr["steps"], "batch": 8192,
                         "full_data_loss": round(full_loss, 4), "constant_predictor_loss": round(const, 4),
                         "seconds": round(train_s, 1)},
            "alpha": 1.0, "eval": evals}


def host_info():
    """CPU model (lscpu) and the number of online cores (nproc) of this host."""
    import subprocess

    model = "unknown"
    try:
        for line in subprocess.run(["lscpu"], capture_output=True, text=True, timeout=30).stdout.splitlines():
            if line.startswith("Model name:"):
                model = line.split(":", 1)[1].strip()
                break
    except (OSError, subprocess.SubprocessError):
        pass
    try:
        nproc = int(subprocess.run(["nproc"], capture_output=True, text=True, timeout=30).stdout.strip())
    except (OSError, ValueError, subprocess.SubprocessError):
        nproc = os.cpu_count() or 1
    return model, nproc


def cpu_leg(a, dg, run_kw, stream):
    """The GPU labels of the full workload against the oracle's union-find, and
    the host-CPU baseline: the MPI form of exponential contraction
    (oracle/mpi_exp.cpp, the reference's default FindComponents path over
    `mpirun -np P` ranks, P = --cpu-ranks) on the same R-MAT graph, timed as the
    reference times it (max over ranks), its labels checked too.  The
    union-find's single-thread time is the floor beside it."""
    import tempfile

    import numpy as np

    import components_amd as ca
    import oracle

    labels, _ = dg.find_components(seed=1001, **run_kw)
    g = dg.download()
    t0 = time.perf_counter()
    ref = oracle.canonical_labels(g.offsets, g.targets)
    uf_s = time.perf_counter() - t0
    exact = bool(np.array_equal(labels, ref))
    model, nproc = host_info()
    ranks = a.cpu_ranks if a.cpu_ranks > 0 else max(1, min(16, nproc))
    print(f"[bench] cpu: {model}, nproc {nproc}; union-find {uf_s:.2f} s", file=sys.stderr, flush=True)
    if a.cpu_scale != a.scale:  # a bounded sample: the same generator at a smaller scale
        del g
        g = ca.generate("rmat", n=1 << a.cpu_scale, m=a.edge_factor << a.cpu_scale, seed=3, rmat_a=0.57,
                        rmat_b=0.19, rmat_c=0.19)
        ref = oracle.canonical_labels(g.offsets, g.targets)
    if not oracle.mpi_available():
        raise RuntimeError("the MPI baseline (oracle/build/mpi_exp, /opt/conda/bin/mpirun) is missing")
    tmp = "/dev/shm" if os.path.isdir("/dev/shm") else tempfile.gettempdir()
    csr = os.path.join(tmp, f"kc_bench_{os.getpid()}.csr")
    lab = os.path.join(tmp, f"kc_bench_{os.getpid()}.labels")
    try:
        oracle.write_csr_file(csr, g.offsets, g.targets)
        t0 = time.perf_counter()
        res = oracle.mpi_exp(csr, ranks, seed=1, iterations=a.cpu_iterations, labels_path=lab, timeout=900)
        wall = time.perf_counter() - t0
    finally:
        for f in (csr, lab):
            if os.path.exists(f):
                os.remove(f)
    ok = bool(np.array_equal(res["labels"], ref))
    del g, ref
    print(f"[bench] cpu: mpirun -np {ranks}: {res['time']:.3f} s per run, {res['levels']} levels, "
          f"{res['sweeps']} sweeps, labels {ok} ({wall:.1f} s)", file=sys.stderr, flush=True)
    cpu = {"value": round(res["m"] / res["time"] / 1e9, 5), "unit": "GTEPS", "cores": ranks, "kind": "port",
           "form": f"mpi: mpirun -np {ranks} (MPICH 3.3.2), one rank per core",
           "sample": f"R-MAT scale {a.cpu_scale}, edge factor {a.edge_factor}, (0.57,0.19,0.19,0.05), seed 3 "
                     f"(n={res['n']}, m={res['m']}), the {'same graph as value' if a.cpu_scale == a.scale else 'a smaller graph'}; "
                     f"oracle/mpi_exp.cpp: block partition, local phase, ghost-key exchange per sweep, reference "
                     f"deviate stream, seq limit m/P = {res['seq_limit']}; {res['iterations']} run(s) with seed 1001, "
                     f"max over ranks",
           "seconds": round(res["time"], 4), "levels": res["levels"], "sweeps": res["sweeps"],
           "labels_bit_exact": ok, "host_cpu": model, "nproc": nproc,
           "union_find_1thread_s": round(uf_s, 3)}
    return cpu, exact


if __name__ == "__main__":
    main()
