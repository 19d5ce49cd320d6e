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


def _k_buffers(hp, n, list_cap, anchor_cap):
    return [hp.DeviceBuffer(2 * n * list_cap * 64), hp.DeviceBuffer(2 * n * 4),
            hp.DeviceBuffer(2 * n * anchor_cap * 64), hp.DeviceBuffer(2 * n * 4),
            hp.DeviceBuffer(2 * n)]


def test_capture_pins_stream_scratch(gpu_engine):
    """Once a graph is captured on a stream, a call that would need more
    scratch on that stream is refused (FMI_ESTATE) instead of moving the
    scratch the graph writes; the replay still equals the eager records, and
    destroying the graph lifts the pin.  A stream with a live graph cannot be
    destroyed."""
    hp = gpu_engine
    g, _, _ = synth.genome(200_000, seed=111, repeat_frac=0.2, fam_len=(300, 2000))
    ix = hp.Index.build(g, 16).to_device()
    reads, offs, _ = synth.reads(g, 400, length=90, seed=112, sub_rate=0.02)
    batch = hp.ReadBatch(reads, offs)
    n = batch.nreads
    st = hp.Stream()
    list_cap, anchor_cap = 16, 4
    fs = hp.DeviceBuffer.from_array(np.full(n, 30, np.int16))
    k_out = _k_buffers(hp, n, list_cap, anchor_cap)
    out = hp.Search1Output(n, 8)
    big = _k_buffers(hp, n, 1024, anchor_cap)

    def chain():
        hp.search_k_launch(ix, batch, fs, list_cap, anchor_cap, *k_out, stream=st.handle)
        hp.lib().fmi_search1_vector_batch(ix.h, hp.MODE_EDIT, 3, batch.d_reads.ptr,
                                          batch.d_offsets.ptr, None, n, 8, out.d_hits.ptr,
                                          out.d_counts.ptr, st.handle)

    def results():
        st.sync()
        return (k_out[1].download(np.uint32, 2 * n, st.handle),
                k_out[0].download(hp.HITK_DTYPE, 2 * n * list_cap, st.handle),
                out.counts(st.handle), out.hits(st.handle))

    for b in k_out + [out.d_counts, out.d_hits]:
        b.zero(st.handle)
    chain()
    want = results()
    graph = hp.Graph(st)
    with graph:
        chain()
    try:
        with pytest.raises(hp.FmiError) as e:       # 64x the list capacity: more scratch
            hp.search_k_launch(ix, batch, fs, 1024, anchor_cap, *big, stream=st.handle)
        assert e.value.code == -7
        with pytest.raises(hp.FmiError) as e:
            st.close()
        assert e.value.code == -7 and st.handle is not None
        for b in k_out + [out.d_counts, out.d_hits]:
            b.zero(st.handle)
        graph.launch()
        got = results()
        assert np.array_equal(got[0], want[0]) and np.array_equal(got[2], want[2])
        for t in range(2 * n):
            c = min(int(want[0][t]), list_cap)
            assert fo.hitsk_equal(got[1][t * list_cap:t * list_cap + c],
                                  want[1][t * list_cap:t * list_cap + c]), t
        cs = np.minimum(want[2], 8)
        for t in range(2 * n):
            assert fo.hits_equal(got[3].reshape(-1, 8)[t, :cs[t]], want[3].reshape(-1, 8)[t, :cs[t]]), t
    finally:
        graph.close()
    # the pin is gone with the graph
    hp.search_k_launch(ix, batch, fs, 1024, anchor_cap, *big, stream=st.handle)
    st.sync()
    st.close()
    for b in k_out + big + [fs]:
        b.free()
    out.free()
    batch.free()
    ix.close()


def test_memset_in_capture_is_a_kernel(gpu_engine):
    """DeviceBuffer.zero (fmi_memset_d) inside a capture records a fill
    kernel: every replay clears the buffer again after eager writes."""
    hp = gpu_engine
    st = hp.Stream()
    pat = np.arange(1 << 16, dtype=np.uint32) | 1
    buf = hp.DeviceBuffer.from_array(pat, st.handle)
    graph = hp.Graph(st)
    with graph:
        hp.check(hp.lib().fmi_memset_d(buf.ptr, 0x5A, buf.nbytes - 12, st.handle))
    try:
        for rep in range(3):
            buf.upload(pat, st.handle)
            graph.launch()
            got = buf.download(np.uint8, buf.nbytes, st.handle)
            assert np.all(got[:-12] == 0x5A), rep
            assert np.array_equal(got[-12:], pat.view(np.uint8)[-12:]), rep
    finally:
        graph.close()
    buf.free()
    st.close()


def test_stream_destroy_releases_its_memory(gpu_engine):
    """50 cycles of create a stream, run searches that allocate its counters
    and scratch (the k-error search's lists, the vector helper's slices),
    destroy it: the device's free memory stays flat."""
    hp = gpu_engine
    g, _, _ = synth.genome(100_000, seed=121)
    ix = hp.Index.build(g, 16).to_device()
    reads, offs, _ = synth.reads(g, 3000, length=120, seed=122, sub_rate=0.02)
    batch = hp.ReadBatch(reads, offs)
    n = batch.nreads
    fs = hp.DeviceBuffer.from_array(np.full(n, 40, np.int16))
    k_out = _k_buffers(hp, n, 256, 4)
    out = hp.Search1Output(n, 8)

    def cycle():
        st = hp.Stream()
        hp.search_k_launch(ix, batch, fs, 256, 4, *k_out, stream=st.handle)
        hp.check(hp.lib().fmi_search1_vector_batch(ix.h, hp.MODE_EDIT, 3, batch.d_reads.ptr,
                                                   batch.d_offsets.ptr, None, n, 8, out.d_hits.ptr,
                                                   out.d_counts.ptr, st.handle))
        hp.search1_launch(ix, batch, out, stream=st.handle)
        st.sync()
        st.close()

    cycle()
    free0, _ = hp.mem_info()
    for _ in range(50):
        cycle()
    free1, _ = hp.mem_info()
    # one cycle's scratch is hundreds of MB; 50 leaked cycles would be tens of
    # GB.  The runtime's own pools move the free figure by a few hundred MB
    # between calls, so the bound is 1 GiB.
    assert free0 - free1 < (1 << 30), (free0, free1)
    for b in k_out + [fs]:
        b.free()
    out.free()
    batch.free()
    ix.close()


def test_forked_graph_runs_branches_and_pins_their_scratch(gpu_engine):
    """Three independent calls on three streams captured as one graph: the
    exact search on the capturing stream, the 1-error and the k-error search
    on two streams forked from it with fmi_stream_wait and joined back.
    Every replay equals the eager records; while the graph lives a forked
    stream's scratch is pinned as the capturing stream's is (a call needing
    more is refused, the stream cannot be destroyed), and destroying the
    graph lifts it."""
    hp = gpu_engine
    g, _, _ = synth.genome(250_000, seed=121, repeat_frac=0.3, fam_len=(300, 2000))
    ix = hp.Index.build(g, 16).to_device().expand()
    ix.kmer_table(8)
    lens = np.random.default_rng(122).integers(40, 220, 1500)
    reads, offs, _ = synth.reads(g, 1500, seed=123, sub_rate=0.015, n_rate=0.002, lengths=lens)
    batch = hp.ReadBatch(reads, offs)
    n = batch.nreads
    st = [hp.Stream() for _ in range(3)]
    ex = hp.DeviceBuffer(n * 32)
    out = hp.Search1Output(n, 8)
    list_cap, anchor_cap = 32, 8
    fs = hp.DeviceBuffer.from_array(np.maximum(lens // 3, 10).astype(np.int16))
    k_out = _k_buffers(hp, n, list_cap, anchor_cap)
    L = hp.lib()

    def chain():
        hp.check(L.fmi_exact_batch(ix.h, hp.VIEW_BACKWARD, 0, batch.d_reads.ptr,
                                   batch.d_offsets.ptr, n, ex.ptr, None, st[0].handle))
        hp.search1_launch(ix, batch, out, mode=hp.MODE_EDIT, stream=st[1].handle)
        hp.search_k_launch(ix, batch, fs, list_cap, anchor_cap, *k_out, stream=st[2].handle)

    def results():
        for x in st:
            x.sync()
        kr = hp.SearchKResult(k_out[1].download(np.uint32, 2 * n),
                              k_out[0].download(hp.HITK_DTYPE, 2 * n * list_cap),
                              k_out[3].download(np.uint32, 2 * n),
                              k_out[2].download(hp.HITK_DTYPE, 2 * n * anchor_cap),
                              k_out[4].download(np.uint8, 2 * n).reshape(n, 2),
                              list_cap, anchor_cap)
        return ex.download(hp.HIT_DTYPE, n), out.counts(), out.hits(), kr

    def clear():
        for b in [ex, out.d_counts, out.d_hits] + k_out:
            b.zero()
        hp.sync()

    clear()
    chain()                              # eager: the reference, and sizes each stream's scratch
    want = results()
    graph = hp.Graph(st[0])
    with graph:
        st[1].wait(st[0])
        st[2].wait(st[0])
        chain()
        st[0].wait(st[1])
        st[0].wait(st[2])
    big = _k_buffers(hp, n, 1024, anchor_cap)
    try:
        for rep in range(3):
            clear()
            hp.search_k_batch(ix, batch, np.full(n, 40, np.int16), list_cap=8)   # eager work
            graph.launch()
            got = results()
            assert np.array_equal(got[0].view(np.uint8), want[0].view(np.uint8)), rep
            assert np.array_equal(got[1], want[1]), rep
            ns = np.minimum(want[1].reshape(n, 2), 8)
            for i in range(n):
                for s in (0, 1):
                    assert fo.hits_equal(got[2][i, s, :ns[i, s]], want[2][i, s, :ns[i, s]]), (rep, i, s)
            assert np.array_equal(got[3].full, want[3].full), rep
            for i in range(n):
                for s in (0, 1):
                    assert fo.hitsk_equal(got[3].final(i, s), want[3].final(i, s)), (rep, i, s)
                    assert fo.hitsk_equal(got[3].anchors(i, s), want[3].anchors(i, s)), (rep, i, s)
        # the forked stream's scratch is pinned by the graph
        with pytest.raises(hp.FmiError):
            hp.search_k_launch(ix, batch, fs, 1024, anchor_cap, *big, stream=st[2].handle)
        with pytest.raises(hp.FmiError):
            st[2].close()
    finally:
        graph.close()
    # destroying the graph lifts the pin
    hp.search_k_launch(ix, batch, fs, 1024, anchor_cap, *big, stream=st[2].handle)
    st[2].sync()
    for b in [ex, fs] + k_out + big:
        b.free()
    out.free()
    batch.free()
    for x in st:
        x.close()
    ix.close()


def test_exact_locate_chain_as_graph(gpu_engine):
    """The exact search of both strands, each followed by locate (SA of the
    first rows of each interval), captured on one stream and replayed three
    times with cleared outputs and other searches run eagerly in between:
    every replay gives the eager records, rows and positions."""
    hp = gpu_engine
    g, _, _ = synth.genome(200_000, seed=131)
    reads, offs, _ = synth.reads(g, 3000, length=100, seed=132, sub_rate=0.0, random_frac=0.1)
    ix = hp.Index.build(g, 16).to_device().expand()
    ix.kmer_table(8)
    batch = hp.ReadBatch(reads, offs)
    n, cap = batch.nreads, 16
    st = hp.Stream()
    L = hp.lib()
    outs = [hp.DeviceBuffer(n * 32) for _ in range(2)]
    pos = [hp.DeviceBuffer(n * cap * 8) for _ in range(2)]
    rows = [hp.DeviceBuffer(n * 4) for _ in range(2)]

    def chain():
        for s, view in ((0, hp.VIEW_BACKWARD), (1, hp.VIEW_BACKWARD_REV)):
            hp.check(L.fmi_exact_batch(ix.h, view, s, batch.d_reads.ptr, batch.d_offsets.ptr, n,
                                       outs[s].ptr, None, st.handle))
            hp.check(L.fmi_locate_hits(ix.h, 0, outs[s].ptr, n, cap, pos[s].ptr, rows[s].ptr,
                                       st.handle))

    def results():
        st.sync()
        return [(outs[s].download(hp.HIT_DTYPE, n).view(np.uint8), rows[s].download(np.uint32, n),
                 pos[s].download(np.uint64, n * cap).reshape(n, cap)) for s in (0, 1)]

    def clear():
        for b in outs + pos + rows:
            b.zero()
        hp.sync()

    try:
        clear()
        chain()
        want = results()
        assert sum(int(w[1].sum()) for w in want) > 0
        graph = hp.Graph(st)
        with graph:
            chain()
        try:
            for rep in range(3):
                clear()
                hp.exact_batch(ix, batch, hp.VIEW_FORWARD, 1)           # eager work
                graph.launch()
                got = results()
                for s in (0, 1):
                    assert np.array_equal(got[s][0], want[s][0]), (rep, s)
                    assert np.array_equal(got[s][1], want[s][1]), (rep, s)
                    m = np.arange(cap)[None, :] < want[s][1][:, None]
                    assert np.array_equal(np.where(m, got[s][2], 0), np.where(m, want[s][2], 0)), \
                        (rep, s)
        finally:
            graph.close()
    finally:
        st.close()
        for b in outs + pos + rows:
            b.free()
        batch.free()
        ix.close()
