"""A sequential replica of the batched kernel's per-graph control flow (k_batch's contract_graph,
components_amd/csrc/batch.hip), run on the CPU over the edge-case batch.

It follows the kernel's level loop step by step -- keys from the deviates of the level's ids, sweeps to
the fixed point, the stall rule, labels and radius, the cluster-edge hash set sized as the kernel sizes
it, the next level numbered in label order from the per-centre counts, the record stack laid out as
the kernel lays it out (level size, previous record, ids, labels, next ids), the undo from the top
record down, canonical labels -- and asserts every index the kernel would use against the buffer the
kernel gives it: the record stack of 10 words per vertex, the per-graph hash set (load at most 1/2,
so its bounded probing always finds a slot), the level CSR buffers.  The results are compared with
the oracle run on each graph alone, as the GPU test of this batch compares the kernel's.
"""
import numpy as np
import pytest

import oracle
from batch_cases import special_graphs

STACK_PER_VERTEX = 10  # kStackPerVertex
MAX_LEVELS = 4096


def replica(g, seed, stream, beta=0.1):
    n0 = g.n
    off0 = np.asarray(g.offsets, np.int64)
    col0 = np.asarray(g.targets, np.int64)
    if n0 == 0:
        return np.zeros(0, np.uint32), 0, 0
    deg0 = np.diff(off0)
    rng_offset = int(np.count_nonzero(deg0))
    if rng_offset == 0:  # no edge: no level runs
        return np.arange(n0, dtype=np.uint32), 0, 0
    slots = int(off0[-1])
    cap = STACK_PER_VERTEX * n0
    tcap = 64
    while tcap < 2 * slots:
        tcap <<= 1
    stack = np.full(cap, -1, np.int64)
    written = np.zeros(cap, bool)

    def put(i, values):
        values = np.atleast_1d(np.asarray(values, np.int64))
        assert i >= 0 and i + len(values) <= cap, f"record stack write [{i}, {i + len(values)}) beyond {cap}"
        stack[i:i + len(values)] = values
        written[i:i + len(values)] = True

    def get(i, count=1):
        assert i >= 0 and i + count <= cap and written[i:i + count].all(), "record stack read of an unwritten word"
        return stack[i:i + count]

    off, col = off0, col0
    nl, rec, last = n0, 0, None
    level0, iteration, levels, radius_sum, status = True, 1, 0, 0, 0
    while True:
        if levels >= MAX_LEVELS:
            status = 2
            break
        hdr = 2 if level0 else 2 + nl
        if rec + hdr + 2 * nl > cap:
            status = 1
            break
        ids = np.arange(nl) if level0 else get(rec + 2, nl)
        dev = [oracle.exp_deviate(int(i), seed, iteration, rng_offset, beta, True) if stream == 0 else
               oracle.hashed_deviate(int(i), seed, iteration, beta) for i in ids]
        key = (np.array(dev, np.uint64) << np.uint64(32)) | np.arange(nl, dtype=np.uint64)
        sweeps, first_change = 0, False
        while True:  # in-place sweeps, one row after another (the same fixed point as the kernel's)
            changed = False
            for v in range(nl):
                for e in range(off[v], off[v + 1]):
                    c = key[col[e]] + np.uint64(1 << 32)
                    if c < key[v]:
                        key[v] = c
                        changed = True
            sweeps += 1
            if sweeps == 1:
                first_change = changed
            if not changed:
                break
        levels += 1
        if not first_change and (not level0 or rng_offset == n0):  # stall: the identity, draw again
            iteration += 1
            continue
        lab = (key & np.uint64(0xFFFFFFFF)).astype(np.int64)
        dist = (key >> np.uint64(32)).astype(np.int64)
        radius_sum += int((dist - dist[lab]).max())
        put(rec + hdr, lab)
        m2l = int(off[nl])
        tsz = 64
        while tsz < 2 * m2l:
            tsz <<= 1
        tsz = min(tsz, tcap)
        src = np.repeat(np.arange(nl), np.diff(off[:nl + 1]))
        cross = lab[src] != lab[col[:m2l]]
        edges = np.unique(lab[src][cross] * nl + lab[col[:m2l]][cross])
        if 2 * len(edges) > tsz:  # the kernel's probing would report a full table
            status = 3
            break
        cnt = np.bincount(edges // nl, minlength=nl)
        centres = np.flatnonzero(cnt)
        newidx = np.full(nl, -1, np.int64)
        newidx[centres] = np.arange(len(centres))
        put(rec + hdr + nl, np.where(newidx < 0, 0xFFFFFFFF, newidx))
        rec_next = rec + hdr + 2 * nl
        room = rec_next + 2 + 3 * nl <= cap
        n_next = len(centres)
        if room and n_next:
            put(rec_next + 2, ids[centres])
        put(rec, nl)
        put(rec + 1, 0xFFFFFFFF if last is None else last)
        last = rec
        if n_next == 0:
            break
        if not room:
            status = 1
            break
        # the next level's CSR fits the level buffers (n_next + 1 offsets, at most the input's slots)
        a, b = edges // nl, edges % nl
        assert len(edges) <= slots and n_next <= n0
        off = np.concatenate([[0], np.cumsum(np.bincount(newidx[a], minlength=n_next))]).astype(np.int64)
        col = newidx[b]
        rec, nl, level0 = rec_next, n_next, False
        iteration += 1
    assert status == 0, f"status {status}"
    # undo, top record down: rep_l[v] = rep_{l+1}[newidx[lab[v]]], or the id of lab[v]
    up, r = None, last
    while r is not None:
        n = int(get(r)[0])
        prev = int(get(r + 1)[0])
        lvl0 = r == 0
        lab = get(r + (2 if lvl0 else 2 + n), n)
        nidx = get(r + (2 if lvl0 else 2 + n) + n, n)
        ids = np.arange(n) if lvl0 else get(r + 2, n)
        out = np.where(nidx[lab] != 0xFFFFFFFF, up[np.minimum(nidx[lab], len(up) - 1)] if up is not None else 0,
                       ids[lab])
        assert up is not None or (nidx == 0xFFFFFFFF).all()
        up = out
        r = None if prev == 0xFFFFFFFF else prev
    minid = np.full(n0, n0, np.int64)
    np.minimum.at(minid, up, np.arange(n0))
    return minid[up].astype(np.uint32), levels, radius_sum


@pytest.mark.parametrize("stream", [0, 1])
def test_batch_replica_on_edge_cases(stream):
    for gi, g in enumerate(special_graphs()):
        labels, levels, radius = replica(g, 7, stream)
        assert np.array_equal(labels, oracle.canonical_labels(g.offsets, g.targets)), gi
        if g.n and g.m:
            run = oracle.Run(g.offsets, g.targets, variant="exp", seed=7, seq_limit=0, stream=stream)
            lv = run.levels()
            assert levels == len(lv), gi
            assert radius == sum(x[3] for x in lv), gi
