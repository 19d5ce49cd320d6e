"""The counting launches of the exact and k-error searches
(fmi_exact_count_work, fmi_search_k_count_work): their results equal the
plain launches' record for record (field by field: the error entries past
num_mismatches and the padding of a k-error record are not written), and
their counters agree with what the text fixes.  On the compact index without a k-mer table the exact search of
the + strand (the backward search of bwt.c:1233-1237) makes one LF query per
base until the interval empties: min(len, m + 1) queries for a read whose
longest suffix occurring in the text has m bases (a base N ends the search
with the query that meets it), each reading one occurrence block, or two
when k and l fall in different blocks; and no SA / ISA, text or table reads."""
import numpy as np
import pytest

import fmoracle as fo
from hpgfm import synth

pytestmark = pytest.mark.gpu


def longest_occurring_suffix(text, w):
    t = text.tobytes()
    m = 0
    for j in range(1, len(w) + 1):
        s = w[len(w) - j:]
        if s[0] > 3 or t.find(s.tobytes()) < 0:
            break
        m = j
    return m


@pytest.fixture(scope="module")
def wsetup(gpu_engine):
    hp = gpu_engine
    g, _, _ = synth.genome(60_000, seed=91, repeat_frac=0.3, fam_len=(200, 1500),
                           divergence=(0.0, 0.05))
    rng = np.random.default_rng(92)
    lens = rng.integers(0, 140, 600)
    reads, offs, _ = synth.reads(g, 600, seed=93, sub_rate=0.02, n_rate=0.004, lengths=lens,
                                 random_frac=0.1)
    batch = hp.ReadBatch(reads, offs)
    yield hp, g, reads, offs, lens, batch
    batch.free()


def test_exact_work_closed_form(wsetup):
    hp, g, reads, offs, lens, batch = wsetup
    ix = hp.Index.build(g, 16).to_device()
    try:
        w, rec = hp.exact_count_work(ix, batch, hp.VIEW_BACKWARD, 0)
        plain, _ = hp.exact_batch(ix, batch, hp.VIEW_BACKWARD, 0)
        assert np.array_equal(rec.view(np.uint8), plain.view(np.uint8))
        q = sum(min(len(W), longest_occurring_suffix(g, W) + 1)
                for W in (reads[offs[i]:offs[i + 1]] for i in range(batch.nreads)))
        blocks, lf, sa_isa, text, table = w
        assert lf == q
        assert lf <= blocks <= 2 * lf
        assert (sa_isa, text, table) == (0, 0, 0)
        w1, rec1 = hp.exact_count_work(ix, batch, hp.VIEW_BACKWARD_REV, 1)
        plain1, _ = hp.exact_batch(ix, batch, hp.VIEW_BACKWARD_REV, 1)
        assert np.array_equal(rec1.view(np.uint8), plain1.view(np.uint8))
        assert w1[1] <= w1[0] <= 2 * w1[1]
    finally:
        ix.close()


@pytest.mark.parametrize("form", ["expanded", "kmer8"])
def test_exact_work_index_forms(wsetup, form):
    """On the expanded index the searches that narrow to one row go on in
    the text (one SA read, text compares, one ISA read); with a k-mer table
    every read of at least k bases free of N in its first k reads one table
    entry, and one more, of the (k - 4)-table, when that k-mer is absent
    from the text or holds an N and its first k - 4 bases do not."""
    hp, g, reads, offs, lens, batch = wsetup
    ix = hp.Index.build(g, 16).to_device().expand()
    if form == "kmer8":
        ix.kmer_table(8)
    try:
        for view, d in ((hp.VIEW_BACKWARD, 0), (hp.VIEW_BACKWARD_REV, 1)):
            w, rec = hp.exact_count_work(ix, batch, view, d)
            plain, _ = hp.exact_batch(ix, batch, view, d)
            assert np.array_equal(rec.view(np.uint8), plain.view(np.uint8))
            blocks, lf, sa_isa, text, table = w
            assert text > 0 and sa_isa > 0 and sa_isa % 2 == 0
            assert lf <= blocks <= 2 * lf
            if form == "kmer8":
                ws = [reads[offs[i]:offs[i + 1]] for i in range(batch.nreads)]
                clean8 = sum(1 for W in ws if len(W) >= 8 and np.all((W[-8:] if d == 0 else W[:8]) <= 3))
                if d == 0:
                    # the + strand's backward search takes W[-1], W[-2], ...: its
                    # k-mer's interval is empty exactly when W[-8:] is absent
                    t = g.tobytes()
                    want = 0
                    for W in ws:
                        hit8 = False
                        if len(W) >= 8:
                            k8 = W[-8:]
                            ok8 = bool(np.all(k8 <= 3))
                            want += ok8
                            hit8 = ok8 and t.find(k8.tobytes()) >= 0
                        if not hit8 and len(W) >= 4 and np.all(W[-4:] <= 3):
                            want += 1
                    assert table == want
                else:
                    assert table >= clean8
            else:
                assert table == 0
    finally:
        ix.close()


@pytest.mark.parametrize("form", ["compact", "kmer8"])
def test_search_k_work(wsetup, form):
    hp, g, reads, offs, lens, batch = wsetup
    ix = hp.Index.build(g, 16).to_device()
    if form == "kmer8":
        ix.expand()
        ix.kmer_table(8)
    fs = np.maximum(lens // 3, 4).astype(np.int16)
    try:
        for lc, ac in ((128, 16), (5, 2)):
            work = np.zeros(6, np.uint64)
            a = hp.search_k_batch(ix, batch, fs, list_cap=lc, anchor_cap=ac, work=work)
            b = hp.search_k_batch(ix, batch, fs, list_cap=lc, anchor_cap=ac)
            assert np.array_equal(a.full, b.full)
            for i in range(batch.nreads):
                for s in (0, 1):
                    assert fo.hitsk_equal(a.final(i, s), b.final(i, s)), (lc, i, s)
                    assert fo.hitsk_equal(a.anchors(i, s), b.anchors(i, s)), (lc, i, s)
            blocks, lf, sa_isa, text, table, recs = (int(x) for x in work)
            assert lf > 0 and lf <= blocks <= 2 * lf
            stored = sum(len(a.final(i, s)) for i in range(batch.nreads) for s in (0, 1))
            assert recs >= stored
            if form == "compact":
                assert (text, table) == (0, 0)
            else:
                assert text > 0 and table > 0
    finally:
        ix.close()
