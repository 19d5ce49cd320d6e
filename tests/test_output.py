"""The reference's text result lines (write_results, search/results.c:3-43;
manage_single_result, search/results.h:125-187) from fmi_format_results.

The golden text was written by the reference's own write_results (compiled
from its sources) for BWSearch1CPU's lists of the golden reads, with a
karyotype of named contigs.  CPU: the formatter over the golden records,
their rows resolved with the oracle's get_SA and binsearch, prints exactly
that text.  GPU: the device pipeline -- fmi_search1_batch (edit mode),
fmi_resolve_hits_text, then the formatter -- prints it too."""
import numpy as np
import pytest

import fmoracle as fo
import hpgfm
from conftest import split_golden_hits


def golden_text(g, i, s):
    off = g["wr_off"]
    return g["wr_blob"][int(off[2 * i + s]):int(off[2 * i + s + 1])].tobytes().decode()


def mapping_of(W):
    return "".join("ACGT"[int(x)] for x in W).encode()


def koff_of(g):
    lens = g["wr_kend"] - g["wr_kstart"] + 1
    return np.concatenate([[0], np.cumsum(lens)]).astype(np.uint64)


def named_index(g):
    ix = hpgfm.Index.build(g["genome"], int(g["ratio"]))
    ix.set_karyotype(g["wr_kstart"], g["wr_kend"])
    ix.set_contig_names([f"chr{c + 1}" for c in range(len(g["wr_kstart"]))])
    return ix


def resolve_on_oracle(o, g, hits, strand, row_cap):
    """Rows of each record as fmi_resolve_hits_text lays them out, from the
    oracle's get_SA and binsearch, with the record's own window length."""
    koff, kst = koff_of(g), g["wr_kstart"]
    n1 = o.size_sa
    chrom = np.full((len(hits), row_cap), -1, np.int32)
    coord = np.zeros((len(hits), row_cap), np.uint64)
    for h, r in enumerate(hits):
        lc = int(r["end"]) - int(r["start"]) + 1
        if r["num_mismatches"]:
            lc += {fo.DELETION: -1, fo.INSERTION: 1}.get(int(r["err_kind"]), 0)
        direction = (not r["dir"]) if strand else bool(r["dir"])
        for j in range(min(int(r["l"] - r["k"] + 1), row_cap)):
            row = int(r["k"]) + j
            key = ((n1 - o.get_sa(1, row) - lc - 1) % (1 << 64)) if direction else o.get_sa(0, row)
            idx = fo.binsearch(koff[:-1], key)
            if idx >= 1 and (key + lc) % (1 << 64) <= int(koff[idx]):
                chrom[h, j] = idx - 1
                coord[h, j] = int(kst[idx - 1]) + key - int(koff[idx - 1])
    return chrom, coord


@pytest.mark.parametrize("name", ["golden_small", "golden_repeats"])
def test_format_matches_reference_text(name, request):
    g = request.getfixturevalue(name)
    ix = named_index(g)
    o = fo.Oracle(g["genome"], ratio=int(g["ratio"]))
    hits = split_golden_hits(g)
    reads, offs = g["reads"], g["offsets"]
    lines = 0
    for i in np.nonzero(g["wr_done"])[0]:
        W = reads[offs[i]:offs[i + 1]]
        for s in (0, 1):
            h = hits[(int(i), s)]
            chrom, coord = resolve_on_oracle(o, g, h, s, 64)
            text, miss = hpgfm.format_results(ix, int(i), s, mapping_of(W), h, chrom, coord)
            assert miss == 0
            assert text == golden_text(g, int(i), s), (i, s)
            lines += text.count("\n")
    assert lines > 300


def test_format_arguments_and_truncation(golden_small):
    g = golden_small
    ix = named_index(g)
    hits = split_golden_hits(g)
    i = next(int(i) for i in np.nonzero(g["wr_done"])[0] if len(hits[(int(i), 0)]))
    h = hits[(i, 0)]
    chrom = np.zeros((len(h), 1), np.int32)
    coord = np.arange(len(h), dtype=np.uint64).reshape(-1, 1)
    text, miss = hpgfm.format_results(ix, 7, 0, b"ACGT", h, chrom, coord)
    assert text.startswith("read_7 + chr1 0 ")
    assert miss == int(sum(max(int(r["l"] - r["k"] + 1) - 1, 0) for r in h))
    import ctypes as C
    n, m = C.c_size_t(), C.c_uint64()
    with pytest.raises(hpgfm.FmiError):      # records announced, none given
        hpgfm.check(hpgfm.lib().fmi_format_results(ix.h, 0, 0, b"A", 1, None, 1, None, None, 1,
                                                   None, 0, C.byref(n), C.byref(m)))
    small = C.create_string_buffer(8)        # truncated copy, full length reported
    hpgfm.check(hpgfm.lib().fmi_format_results(ix.h, 7, 0, b"ACGT", 4,
                                               h.ctypes.data_as(C.c_void_p), len(h),
                                               chrom.ctypes.data_as(C.c_void_p),
                                               coord.ctypes.data_as(C.c_void_p), 1, small, 8,
                                               C.byref(n), C.byref(m)))
    assert n.value == len(text) and small.value == text[:7].encode()


@pytest.mark.gpu
@pytest.mark.parametrize("name", ["golden_small", "golden_repeats"])
def test_gpu_pipeline_text(gpu_engine, name, request):
    hp = gpu_engine
    g = request.getfixturevalue(name)
    ix = named_index(g).to_device()
    reads, offs = g["reads"], g["offsets"]
    batch = hp.ReadBatch(reads, offs)
    cap, row_cap = 32, 64
    out = hp.Search1Output(batch.nreads, cap)
    hp.search1_launch(ix, batch, out, mode=hp.MODE_EDIT)
    hp.sync()
    counts = out.counts().reshape(-1, 2)
    hits = out.hits()
    chrom, coord = hp.resolve_hits(ix, batch, out, row_cap, text_rule=True)
    for i in np.nonzero(g["wr_done"])[0]:
        i = int(i)
        W = reads[offs[i]:offs[i + 1]]
        for s in (0, 1):
            c = int(counts[i, s])
            assert c <= cap
            text, miss = hp.format_results(ix, i, s, mapping_of(W), hits[i, s, :c],
                                           chrom[i, s, :c], coord[i, s, :c])
            assert miss == 0
            assert text == golden_text(g, i, s), (i, s)
    out.free()
    batch.free()
    ix.close()
