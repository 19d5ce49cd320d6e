"""Launch chains captured as HIP graphs (fmi_graph_*): the 1-error search in
text mode with a k-mer table (the search and the rows of its text-mode
records) followed by the k-error search, captured once on a stream and
replayed three times, with the outputs cleared and other searches run
eagerly in between, gives the eager records on every replay."""
import numpy as np
import pytest

import fmoracle as fo
from hpgfm import synth

pytestmark = pytest.mark.gpu


def test_graph_replays_equal_eager(gpu_engine):
    hp = gpu_engine
    g, _, _ = synth.genome(300_000, seed=101, repeat_frac=0.3, fam_len=(300, 2000))
    ix = hp.Index.build(g, 16).to_device().expand()
    ix.kmer_table(8)
    lens = np.random.default_rng(102).integers(30, 200, 2000)
    reads, offs, _ = synth.reads(g, 2000, seed=103, sub_rate=0.015, n_rate=0.002, lengths=lens)
    batch = hp.ReadBatch(reads, offs)
    oreads, ooffs, _ = synth.reads(g, 300, length=100, seed=104, sub_rate=0.02)
    other = hp.ReadBatch(oreads, ooffs)
    st = hp.Stream()
    cap, list_cap, anchor_cap = 16, 64, 8
    n = batch.nreads
    out = hp.Search1Output(n, cap)
    fs = hp.DeviceBuffer.from_array(np.maximum(lens // 3, 10).astype(np.int16))
    k_out = [hp.DeviceBuffer(2 * n * list_cap * 64), hp.DeviceBuffer(2 * n * 4),
             hp.DeviceBuffer(2 * n * anchor_cap * 64), hp.DeviceBuffer(2 * n * 4),
             hp.DeviceBuffer(2 * n)]

    def chain():
        hp.search1_launch(ix, batch, out, mode=hp.MODE_EDIT, stream=st.handle)
        hp.search_k_launch(ix, batch, fs, list_cap, anchor_cap, *k_out, stream=st.handle)

    def results():
        st.sync()
        kr = hp.SearchKResult(k_out[1].download(np.uint32, 2 * n, st.handle),
                              k_out[0].download(hp.HITK_DTYPE, 2 * n * list_cap, st.handle),
                              k_out[3].download(np.uint32, 2 * n, st.handle),
                              k_out[2].download(hp.HITK_DTYPE, 2 * n * anchor_cap, st.handle),
                              k_out[4].download(np.uint8, 2 * n, st.handle).reshape(n, 2),
                              list_cap, anchor_cap)
        return out.counts(st.handle), out.hits(st.handle), kr

    for b in [out.d_counts, out.d_hits] + k_out:  # slots past a count are never written
        b.zero(st.handle)
    chain()                              # eager: the reference, and sizes the scratch
    want = results()
    graph = hp.Graph(st)
    with graph:
        chain()
    try:
        for rep in range(3):
            for b in [out.d_counts, out.d_hits] + k_out:
                b.zero(st.handle)
            hp.search1_batch(ix, other, cap=8)          # eager work between replays
            hp.search_k_batch(ix, other, np.full(other.nreads, 30, np.int16), list_cap=16)
            graph.launch()
            got = results()
            assert np.array_equal(got[0], want[0]), rep
            n_stored = np.minimum(want[0].reshape(n, 2), cap)
            for i in range(n):
                for s in (0, 1):
                    assert fo.hits_equal(got[1][i, s, :n_stored[i, s]],
                                         want[1][i, s, :n_stored[i, s]]), (rep, i, s)
            gk, wk = got[2], want[2]
            assert np.array_equal(gk.full, wk.full), rep
            for i in range(n):
                for s in (0, 1):
                    # records compared field by field up to their error count
                    assert fo.hitsk_equal(gk.final(i, s), wk.final(i, s)), (rep, i, s)
                    assert fo.hitsk_equal(gk.anchors(i, s), wk.anchors(i, s)), (rep, i, s)
    finally:
        graph.close()
        for b in [fs] + k_out:
            b.free()
        out.free()
        batch.free()
        other.free()
        st.close()
        ix.close()
