"""Text mode (fmi_index_expand): the full SA / ISA of both directions and the
2-bit text, derived on the GPU from the samples, and the 1-error searches
that continue single-row intervals on the text.  The arrays must be the
oracle's SA / ISA at every row and the genome itself; every search must give
the golden lists and exactly the records of the unexpanded engine."""
import numpy as np
import pytest

import fmoracle as fo
from conftest import split_golden_hits
from hpgfm import synth

pytestmark = pytest.mark.gpu


def unpack_text(words, n):
    sh = (np.arange(16, dtype=np.uint32) * 2)[None, :]
    return ((words[:, None] >> sh) & 3).astype(np.uint8).reshape(-1)[:n]


@pytest.fixture(scope="module", params=["golden_small", "golden_repeats"])
def tgold(request, gpu_engine):
    g = request.getfixturevalue(request.param)
    ix = gpu_engine.Index.build(g["genome"], int(g["ratio"])).to_device()
    base = ix.device_bytes
    ix.expand()
    assert ix.device_bytes > base
    batch = gpu_engine.ReadBatch(g["reads"], g["offsets"])
    yield gpu_engine, g, ix, batch
    batch.free()
    ix.close()


def test_expanded_arrays(tgold):
    hp, g, ix, batch = tgold
    o = fo.Oracle(g["genome"], ratio=int(g["ratio"]))
    n = len(g["genome"])
    for fwd in (0, 1):
        sa, isa, text = ix.full_arrays(fwd)
        assert [int(x) for x in sa] == [o.get_sa(fwd, m) for m in range(n + 1)]
        assert np.array_equal(isa[sa], np.arange(n + 1, dtype=np.uint32))
        assert np.array_equal(unpack_text(text, n), g["genome"])


@pytest.mark.parametrize("plain", [False, True])
@pytest.mark.parametrize("mode", ["edit", "mismatch"])
def test_textmode_search1_golden(tgold, plain, mode):
    hp, g, ix, batch = tgold
    cap = int(g["cap"])
    m = hp.MODE_EDIT if mode == "edit" else hp.MODE_MISMATCH
    counts, hits = hp.search1_batch(ix, batch, cap=cap, mode=m, plain=plain)
    for (i, s), w in split_golden_hits(g).items():
        if mode == "mismatch":
            w = w[(w["num_mismatches"] == 0) | (w["err_kind"] == fo.MISMATCH)]
        assert int(counts[i, s]) == len(w), (i, s)
        assert fo.hits_equal(hits[i, s, :min(len(w), cap)], w[:cap]), (i, s)


def test_textmode_equals_block_mode(gpu_engine):
    """Reads with N bases, ragged lengths, dense repeats: the expanded and
    the plain index give identical counts and records, at a capacity that
    truncates, on both strands and each strand alone, with either way of
    issuing an iteration's loads (fmi_set_load_mode)."""
    hp = gpu_engine
    g, _, _ = synth.genome(600_000, seed=91, repeat_frac=0.4, fam_len=(300, 3000),
                           divergence=(0.0, 0.05), copies_per_family=12)
    a = hp.Index.build(g, 16).to_device()
    b = hp.Index.build(g, 16).to_device().expand()
    lens = np.random.default_rng(92).integers(0, 251, 4000)
    reads, offs, _ = synth.reads(g, 4000, seed=93, sub_rate=0.015, n_rate=0.002, lengths=lens)
    batch = hp.ReadBatch(reads, offs)
    try:
        for lm in (hp.LOADS_ALL, hp.LOADS_OWN):
            hp.set_load_mode(lm)
            for m in (hp.MODE_EDIT, hp.MODE_MISMATCH):
                for mask in (3, 1, 2):
                    c1, h1 = hp.search1_batch(ix=a, batch=batch, cap=6, mode=m, strand_mask=mask)
                    c2, h2 = hp.search1_batch(ix=b, batch=batch, cap=6, mode=m, strand_mask=mask)
                    assert np.array_equal(c1, c2), (lm, m, mask)
                    n = np.minimum(c1, 6)
                    for i in range(len(lens)):
                        for s in (0, 1):
                            assert fo.hits_equal(h1[i, s, :n[i, s]], h2[i, s, :n[i, s]]), \
                                (lm, m, mask, i, s)
    finally:
        hp.set_load_mode(hp.LOADS_AUTO)
    wa = hp.search1_count_work(a, batch, cap=6)
    wb = hp.search1_count_work(b, batch, cap=6)
    assert wa[2] == 0 and wa[3] == 0
    assert wb[0] < wa[0] and wb[2] > 0 and wb[3] > 0
    batch.free()
    a.close()
    b.close()


def test_textmode_exact_golden(tgold):
    """BWExactSearchBackward/Forward on the four views with anchors, the
    expanded index: the final interval (empty ones included), the position
    where the search stopped and the anchor equal the golden vectors."""
    hp, g, ix, batch = tgold
    for view in range(4):
        for d in (0, 1):
            rec, anc = hp.exact_batch(ix, batch, view, d, anchors=True)
            want = g[f"ex_v{view}_d{d}"]
            wa = g[f"ex_anchor_v{view}_d{d}"]
            for key in ("k", "l", "pos"):
                assert np.array_equal(rec[key], want[key]), (view, d, key)
            for key in ("k", "l", "start", "pos", "end", "dir"):
                assert np.array_equal(anc[key], wa[key]), (view, d, key)


def test_textmode_exact_with_n(gpu_engine):
    """Reads with N bases and ragged lengths: the expanded index gives the
    same exact records and anchors as the compact one on every view."""
    hp = gpu_engine
    g, _, _ = synth.genome(300_000, seed=95, repeat_frac=0.3, fam_len=(300, 3000))
    a = hp.Index.build(g, 16).to_device()
    b = hp.Index.build(g, 16).to_device().expand()
    lens = np.random.default_rng(96).integers(0, 251, 3000)
    reads, offs, _ = synth.reads(g, 3000, seed=97, sub_rate=0.01, n_rate=0.003, lengths=lens)
    batch = hp.ReadBatch(reads, offs)
    for view in range(4):
        for d in (0, 1):
            r1, a1 = hp.exact_batch(a, batch, view, d, anchors=True)
            r2, a2 = hp.exact_batch(b, batch, view, d, anchors=True)
            for key in ("k", "l", "pos"):
                assert np.array_equal(r1[key], r2[key]), (view, d, key)
            for key in ("k", "l", "start", "pos", "end"):
                assert np.array_equal(a1[key], a2[key]), (view, d, key)
    batch.free()
    a.close()
    b.close()


@pytest.mark.parametrize("form", ["compact", "expanded", "kmer8"])
def test_exact_staged_and_long_reads_vs_oracle(gpu_engine, form):
    """Reads of 1-400 bases in one batch, so a wavefront holds reads staged in
    LDS (at most 256 bases) beside reads searched from their bytes, with N
    bases, every length at a staging word or slot boundary, and copies of
    the genome's ends: on every view and direction the record and the anchor
    equal the oracle's (BWExactSearchBackward / Forward,
    search/search.h:9-65)."""
    hp = gpu_engine
    g, _, _ = synth.genome(200_000, seed=141, repeat_frac=0.3, fam_len=(300, 3000))
    # random lengths, then every word and slot boundary of the staging (16
    # bases per code word, 32 per flag word, 256 per slot) four times
    edges = np.array([1, 2, 3, 15, 16, 17, 31, 32, 33, 63, 64, 65, 127, 128, 129,
                      255, 256, 257], np.int64)
    lens = np.concatenate([np.random.default_rng(142).integers(1, 401, 1200), np.tile(edges, 4)])
    reads, offs, _ = synth.reads(g, len(lens), seed=143, sub_rate=0.003, n_rate=0.002,
                                 lengths=lens)
    assert np.any(reads > 3)
    # copies of the genome's first and last bases, both orientations: text
    # runs that reach the text's start (suffix 0) or its last word
    ends = []
    for L in (20, 33, 40, 100, 200, 256, 300):
        for w in (g[:L], g[-L:]):
            ends += [w, synth.revcomp(w)]
    reads = np.concatenate([reads] + ends).astype(np.uint8)
    offs = np.concatenate([offs, offs[-1] + np.cumsum([len(w) for w in ends])]).astype(np.uint64)
    o = fo.Oracle(g, ratio=16)
    lsa = o.size_sa - 1
    ix = hp.Index.build(g, 16).to_device()
    if form != "compact":
        ix.expand()
    if form == "kmer8":
        ix.kmer_table(8)
    batch = hp.ReadBatch(reads, offs)
    try:
        for view in range(4):
            for d in (0, 1):
                rec, anc = hp.exact_batch(ix, batch, view, d, anchors=True)
                for i in range(batch.nreads):
                    W = reads[offs[i]:offs[i + 1]]
                    L = len(W)
                    r, a = o.exact(view, d, W, 0, lsa, 0, 0 if d else L - 1, L - 1, anchor=True)
                    for key in ("k", "l", "pos"):
                        assert rec[i][key] == r[key], (form, view, d, i, L, key)
                    for key in ("k", "l", "start", "pos", "end", "dir"):
                        assert anc[i][key] == a[key], (form, view, d, i, L, "anchor", key)
    finally:
        batch.free()
        ix.close()


def test_textmode_locate_and_inverse(tgold):
    """get_SA / get_ISA read the full arrays on an expanded index: the golden
    vectors, and locate over whole intervals."""
    hp, g, ix, batch = tgold
    rows = g["sa_rows"].astype(np.int64)
    for fwd, ks, ki in ((0, "sa_bwd", "isa_bwd"), (1, "sa_fwd", "isa_fwd")):
        pos, nrows = hp.locate_batch(ix, rows, rows, cap=1, forward=fwd)
        assert np.all(nrows == 1) and np.array_equal(pos[:, 0], g[ks])
        assert np.array_equal(hp.inverse_locate_batch(ix, rows.astype(np.uint64), forward=fwd), g[ki])
    o = fo.Oracle(g["genome"], ratio=int(g["ratio"]))
    k = np.array([0, 5, 100, 7, 3], np.int64)
    l = np.array([40, 4, 103, 7, 2000], np.int64)
    pos, nrows = hp.locate_batch(ix, k, l, cap=64, forward=0)
    for i in range(len(k)):
        want = [o.get_sa(0, m) for m in range(k[i], min(l[i], k[i] + 63) + 1)] if k[i] <= l[i] else []
        assert int(nrows[i]) == len(want) and list(pos[i, :len(want)]) == want


def test_textmode_seed_map(gpu_engine):
    """Seed mapping on an expanded index equals the compact index's, region
    for region (the located positions come from the full suffix array)."""
    hp = gpu_engine
    g, _, _ = synth.genome(400_000, seed=98, repeat_frac=0.3, fam_len=(300, 2000),
                           divergence=(0.0, 0.05))
    cuts = np.array([0, 120_000, 260_000], np.uint64)
    lens = np.diff(np.concatenate([cuts, [len(g)]])).astype(np.uint64)
    starts = np.full(3, 500, np.uint64)
    out = []
    for expand in (False, True):
        ix = hp.Index.build(g, 16)
        ix.set_karyotype(starts, starts + lens - 1)
        ix.to_device()
        if expand:
            ix.expand()
        reads, offs, _ = synth.reads(g, 1200, length=120, seed=99, sub_rate=0.01)
        batch = hp.ReadBatch(reads, offs)
        out.append(hp.seed_map_batch(ix, batch, 18, 5, 1000))
        batch.free()
        ix.close()
    (n1, r1, g1, i1), (n2, r2, g2, i2) = out
    assert n1 == n2 and r1 == r2 and r1 > 0
    for key in g1.dtype.names:
        assert np.array_equal(g1[key], g2[key]), key
