"""fmi_search1_batch_noalloc (the allocation-free variant of
fmi_search1_batch) through the C ABI: two calls in a row, the second on a
smaller batch, each equal to fmi_search1_batch on the same reads -- counts
and every stored record, bit for bit -- on the compact index and in text
mode with a k-mer table."""
import ctypes as C

import numpy as np
import pytest

import fmoracle as fo
from hpgfm import synth

pytestmark = pytest.mark.gpu


def _run(hp, fn, ix, batch, cap, mode):
    out = hp.Search1Output(batch.nreads, cap)
    out.d_counts.zero()
    out.d_hits.zero()
    hp.check(fn(ix.h, mode, 3, batch.d_reads.ptr, batch.d_offsets.ptr, batch.nreads, cap,
                out.d_hits.ptr, out.d_counts.ptr, None))
    hp.sync()
    c, h = out.counts().reshape(-1, 2), out.hits()
    out.free()
    return c, h


def _same(a, b, cap):
    ca, ha = a
    cb, hb = b
    if not np.array_equal(ca, cb):
        return False
    n = np.minimum(ca, cap)
    for i in range(ca.shape[0]):
        for s in (0, 1):
            if not fo.hits_equal(ha[i, s, :n[i, s]], hb[i, s, :n[i, s]]):
                return False
    return True


def test_noalloc_equals_search1(gpu_engine):
    hp = gpu_engine
    L = hp.lib()
    g, _, _ = synth.genome(300_000, seed=131, repeat_frac=0.3, fam_len=(300, 2000))
    ix = hp.Index.build(g, 16).to_device()
    big = synth.reads(g, 6000, length=120, seed=132, sub_rate=0.02, n_rate=0.002)[:2]
    small = synth.reads(g, 700, length=80, seed=133, sub_rate=0.03)[:2]
    cap = 12
    try:
        for stage in ("compact", "text mode, 9-mer table"):
            if stage != "compact":
                ix.expand()
                ix.kmer_table(9)
            for mode in (hp.MODE_MISMATCH, hp.MODE_EDIT):
                for reads, offs in (big, small):       # the second call: a smaller batch
                    print(f"[no-alloc] {stage}, mode {mode}, {len(offs) - 1} reads", flush=True)
                    batch = hp.ReadBatch(reads, offs)
                    try:
                        want = _run(hp, L.fmi_search1_batch, ix, batch, cap, mode)
                        got = _run(hp, L.fmi_search1_batch_noalloc, ix, batch, cap, mode)
                    finally:
                        batch.free()
                    assert int(want[0].sum()) > 0
                    assert _same(got, want, cap), (stage, mode, len(offs) - 1)
    finally:
        ix.close()


def test_noalloc_rejects_bad_calls(gpu_engine):
    hp = gpu_engine
    L = hp.lib()
    g, _, _ = synth.genome(50_000, seed=134)
    ix = hp.Index.build(g, 16)
    p = C.c_void_p(1)
    # not on the device yet, bad mode / strand mask, NULL buffers
    assert L.fmi_search1_batch_noalloc(ix.h, hp.MODE_EDIT, 3, p, p, 1, 1, p, p, None) == -7
    ix.to_device()
    assert L.fmi_search1_batch_noalloc(ix.h, 9, 3, p, p, 1, 1, p, p, None) == -1
    assert L.fmi_search1_batch_noalloc(ix.h, hp.MODE_EDIT, 0, p, p, 1, 1, p, p, None) == -1
    assert L.fmi_search1_batch_noalloc(ix.h, hp.MODE_EDIT, 3, None, p, 1, 1, p, p, None) == -1
    assert L.fmi_search1_batch_noalloc(ix.h, hp.MODE_EDIT, 3, p, p, 1 << 31, 1, p, p, None) == -6
    ix.close()
