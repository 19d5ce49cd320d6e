"""The k-mer interval table (fmi_index_kmer_table): every entry equals the
oracle's exact search of that k-mer from the whole index; the 1-error search
that takes the first k steps of each exact phase from it keeps the golden
records of the compiled reference core and the oracle's (N bases, reads too
short for a lookup, both modes, compact and expanded index); the exact
search keeps the golden intervals, stop positions and anchors (dead ones
included) on all four views; the seed mapping gives the same regions and
per-seed intervals with and without the table."""
import numpy as np
import pytest

import fmoracle as fo
from conftest import split_golden_hits
from hpgfm import synth

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module", params=["golden_small", "golden_repeats"])
def kgold(request, gpu_engine):
    g = request.getfixturevalue(request.param)
    ix = gpu_engine.Index.build(g["genome"], int(g["ratio"])).to_device()
    batch = gpu_engine.ReadBatch(g["reads"], g["offsets"])
    yield gpu_engine, g, ix, batch
    batch.free()
    ix.close()


@pytest.mark.parametrize("expanded", [False, True])
@pytest.mark.parametrize("k", [4, 7, 15])
def test_kmer_golden_search1(kgold, k, expanded):
    hp, g, ix, batch = kgold
    if expanded:
        ix.expand()
    ix.kmer_table(k)
    want = split_golden_hits(g)
    cap = int(g["cap"])
    counts, hits = hp.search1_batch(ix, batch, cap=cap, mode=hp.MODE_EDIT)
    assert np.array_equal(counts, g["s1_counts"])
    for i in range(batch.nreads):
        for s in (0, 1):
            c = int(counts[i, s])
            assert fo.hits_equal(hits[i, s, :c], want[(i, s)]), (k, i, s)
    w = hp.search1_count_work(ix, batch, cap=cap)
    assert w[4] > 0
    for bad in (3, 16):                  # k is 0 or 4..15
        with pytest.raises(hp.FmiError):
            ix.kmer_table(bad)
    ix.kmer_table(0)
    w0 = hp.search1_count_work(ix, batch, cap=cap)
    assert w0[4] == 0 and w0[1] > w[1]


@pytest.mark.parametrize("name", ["golden_small", "golden_repeats"])
def test_kmer_table_before_expand(gpu_engine, name, request):
    """A table built on the compact index, then the index expanded: the
    golden 1-error records, and the same again after the table is rebuilt
    on the expanded index."""
    hp = gpu_engine
    g = request.getfixturevalue(name)
    ix = hp.Index.build(g["genome"], int(g["ratio"])).to_device()
    batch = hp.ReadBatch(g["reads"], g["offsets"])
    want = split_golden_hits(g)
    cap = int(g["cap"])
    ix.kmer_table(6)
    ix.expand()
    for _ in range(2):
        counts, hits = hp.search1_batch(ix, batch, cap=cap, mode=hp.MODE_EDIT)
        assert np.array_equal(counts, g["s1_counts"])
        for i in range(batch.nreads):
            for s in (0, 1):
                assert fo.hits_equal(hits[i, s, :int(counts[i, s])], want[(i, s)]), (i, s)
        ix.kmer_table(7)
    batch.free()
    ix.close()


def test_kmer_entries_and_oracle_with_n(gpu_engine):
    hp = gpu_engine
    g, _, _ = synth.genome(150_000, seed=61, repeat_frac=0.3, fam_len=(200, 1500),
                           divergence=(0.0, 0.04))
    ix = hp.Index.build(g, 16).to_device()
    o = fo.Oracle(g, ratio=16)
    k = 6
    ix.kmer_table(k)
    tables = [ix.kmer_entries(0), ix.kmer_entries(1)]
    assert len(tables[0]) == 4 ** k
    # entries: the exact search of the k-mer in feed order, per direction
    rng = np.random.default_rng(62)
    lsa = o.size_sa - 1
    for code in rng.integers(0, 4 ** k, 200):
        syms = np.array([(int(code) >> (2 * j)) & 3 for j in range(k)], np.uint8)
        # direction 0 (BWT of X) takes W[k-1], W[k-2], ... : W = reversed feed
        r0 = o.exact(fo.VIEW_BACKWARD, 0, syms[::-1].copy(), 0, lsa, 0, k - 1, k - 1)
        # direction 1 (BWT of reverse X) takes W[0], W[1], ...
        r1 = o.exact(fo.VIEW_FORWARD, 1, syms, 0, lsa, 0, 0, k - 1)
        for fwd, r in ((0, r0), (1, r1)):
            e = int(tables[fwd][int(code)])
            want = (int(r["k"]), int(r["l"])) if r["k"] <= r["l"] else (1, 0)
            assert (e & 0xFFFFFFFF, e >> 32) == want, (fwd, int(code))
    lens = rng.integers(1, 40, 500)
    lens[:20] = rng.integers(1, 2 * k, 20)        # too short for a lookup
    reads, offs, _ = synth.reads(g, 500, seed=63, sub_rate=0.02, n_rate=0.01, lengths=lens)
    batch = hp.ReadBatch(reads, offs)
    for mode, mm in ((hp.MODE_EDIT, False), (hp.MODE_MISMATCH, True)):
        counts, hits = hp.search1_batch(ix, batch, cap=64, mode=mode)
        for i in range(batch.nreads):
            W = reads[offs[i]:offs[i + 1]]
            for s in (0, 1):
                n, want = o.search1(s, W, cap=64, mismatch_only=mm)
                assert int(counts[i, s]) == n, (mode, i, s)
                assert fo.hits_equal(hits[i, s, :min(n, 64)], want), (mode, i, s)
    batch.free()
    ix.close()


@pytest.mark.parametrize("expanded", [False, True])
def test_kmer_exact_golden(kgold, expanded):
    hp, g, ix, batch = kgold
    if expanded:
        ix.expand()
    for k in (4, 7):
        ix.kmer_table(k)
        for view in range(4):
            for d in (0, 1):
                rec, anc = hp.exact_batch(ix, batch, view, d, anchors=True)
                want = g[f"ex_v{view}_d{d}"]
                wa = g[f"ex_anchor_v{view}_d{d}"]
                for key in ("k", "l", "pos"):
                    assert np.array_equal(rec[key], want[key]), (k, view, d, key)
                for key in ("k", "l", "start", "pos", "end", "dir"):
                    assert np.array_equal(anc[key], wa[key]), (k, view, d, key)
    ix.kmer_table(0)


def test_kmer_seed_map_unchanged(gpu_engine):
    hp = gpu_engine
    g, _, _ = synth.genome(300_000, seed=66, repeat_frac=0.4, fam_len=(200, 2000),
                           divergence=(0.0, 0.03))
    ix = hp.Index.build(g, 16)
    ix.set_karyotype([0, 0], [149_999, 149_999])
    ix.to_device()
    reads, offs, _ = synth.reads(g, 600, length=100, seed=67, sub_rate=0.02, n_rate=0.005)
    batch = hp.ReadBatch(reads, offs)
    base = hp.seed_map_batch(ix, batch, 18, 5, 1000)
    for k in (6, 12):
        ix.kmer_table(k)
        got = hp.seed_map_batch(ix, batch, 18, 5, 1000)
        assert got[0] == base[0] and got[1] == base[1]
        assert got[2].tobytes() == base[2].tobytes(), k
        assert got[3].tobytes() == base[3].tobytes(), k
    batch.free()
    ix.close()
