"""Exact search + locate on the whole CPU-runnable workload: a 1 Mbp i.i.d.
reference (seed 1, 41% GC) and 100K reads of 100 bp, 90% copied from random
positions and strands without errors and 10% random.  Per read, the
backward search of the + strand and the forward search on backward_rev for
the - strand (bwt.c:1233-1237), then get_SA of the backward index over the
first 64 rows of each interval (bwt.c:1271-1277).

The oracle runs the whole batch here: every read's row count and position
sum per strand equals the oracle's, on the compact index and on the
expanded one, and every copied read is located at the position it was
copied from."""
import os

import numpy as np
import pytest

import fmoracle as fo
from hpgfm import synth

pytestmark = pytest.mark.gpu

CAP = 64


def exact_locate(hp, ix, batch):
    n = batch.nreads
    rows = np.zeros((n, 2), np.uint32)
    sums = np.zeros((n, 2), np.uint64)
    where = []
    for s, view in ((0, hp.VIEW_BACKWARD), (1, hp.VIEW_BACKWARD_REV)):
        rec, _ = hp.exact_batch(ix, batch, view, s)
        d = hp.DeviceBuffer.from_array(rec)
        try:
            p, nr = hp.locate_hits(ix, d, n, CAP, forward=0)
        finally:
            d.free()
        rows[:, s] = nr
        m = np.arange(CAP)[None, :] < np.minimum(nr, CAP)[:, None]
        sums[:, s] = np.where(m, p, 0).sum(axis=1, dtype=np.uint64)
        where.append((p, m))
    return rows, sums, where


def test_exact_locate_whole_batch(gpu_engine):
    hp = gpu_engine
    n = 100_000
    g, _, _ = synth.genome(1_000_000, seed=1, gc=0.41)
    reads, offs, truth = synth.reads(g, n, length=100, seed=2, sub_rate=0.0, random_frac=0.1)
    ix = hp.Index.build_gpu(g, 16)
    batch = hp.ReadBatch(reads, offs)
    o = fo.Oracle(g, ratio=16)
    threads = max(1, min(16, len(os.sched_getaffinity(0))))
    _, orows, osums = o.exact_locate_batch(reads, offs, cap=CAP, threads=threads)
    try:
        rows, sums, where = exact_locate(hp, ix, batch)
        assert np.array_equal(rows, orows) and np.array_equal(sums, osums)
        copied = np.nonzero(~truth["is_random"])[0]
        s = truth["strand"][copied]
        assert np.all(rows[copied, s] >= 1)
        for strand in (0, 1):
            i = copied[s == strand]
            p, m = where[strand]
            hit = m[i] & (p[i] == truth["pos"][i][:, None].astype(np.uint64))
            assert np.all(hit.any(axis=1)), strand
        ix.expand()
        rows2, sums2, _ = exact_locate(hp, ix, batch)
        assert np.array_equal(rows2, orows) and np.array_equal(sums2, osums)
    finally:
        batch.free()
        ix.close()
