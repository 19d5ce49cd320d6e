"""Exact seed mapping, bwt_map_exact_seed (bwt.c:1200-1294).

CPU: the oracle's restatement against closed forms -- a seed cut from a
known position of a known contig yields a region at that contig's 1-based
coordinate on the right strand, with the one-short length of bwt.c:1214;
occurrences crossing a contig end yield none; the cumulative filter skips
exactly the strands the reference's running `int` count exceeds.
GPU: fmi_seed_map_batch equals the oracle region for region, in order."""
import numpy as np
import pytest

import fmoracle as fo
from hpgfm import synth


def karyotype(n, cuts):
    starts = np.concatenate([[0], cuts]).astype(np.uint64)
    lens = np.diff(np.concatenate([starts, [n]])).astype(np.uint64)
    koff = np.concatenate([[0], np.cumsum(lens)]).astype(np.uint64)
    kstart = np.full(len(lens), 1000, np.uint64)          # contigs start at coordinate 1000
    return koff, kstart


@pytest.fixture(scope="module")
def seeded():
    g, _, _ = synth.genome(30_000, seed=17)
    o = fo.Oracle(g, ratio=8)
    koff, kstart = karyotype(len(g), [9000, 20000])
    return g, o, koff, kstart


def test_seed_regions_closed_form(seeded):
    g, o, koff, kstart = seeded
    rng = np.random.default_rng(3)
    L, sl, st = 60, 18, 5
    checked = 0
    for _ in range(200):
        p = int(rng.integers(0, len(g) - L))
        for strand in (0, 1):
            W = g[p:p + L].copy() if strand == 0 else synth.revcomp(g[p:p + L])
            regs, info = o.seed_map(W, sl, st, 1000, koff, kstart)
            for q in range(len(info)):
                s0 = q * st
                tpos = p + s0 if strand == 0 else p + (L - s0 - sl)   # text position of the seed
                c = int(np.searchsorted(koff, tpos, side="right")) - 1
                fits = tpos + (sl - 1) <= int(koff[c + 1])
                mine = regs[(regs["seed_id"] == q) & (regs["strand"] == strand)]
                want_start = int(kstart[c]) + tpos - int(koff[c]) + 1
                hit = mine[(mine["valid"] == 1) & (mine["start"] == want_start)]
                if fits:
                    assert len(hit) == 1, (p, strand, q)
                    assert int(hit[0]["end"]) == want_start + sl - 1
                    assert int(hit[0]["chrom_idx"]) == c + 1
                    exp_ss = s0 if strand == 0 else L - (s0 + sl - 1) - 1
                    assert int(hit[0]["seq_start"]) == exp_ss
                    checked += 1
                else:
                    assert len(hit) == 0
    assert checked > 1000


def test_seed_filter_is_cumulative(seeded):
    g, o, koff, kstart = seeded
    # a seed made of one repeated base occurs many times on one strand
    W = np.zeros(18, np.uint8)
    regs, info = o.seed_map(W, 6, 1, 1 << 30, koff, kstart)
    kp, lp, km, lm = (int(x) for x in info[0, :4])
    cp, cm = lp - kp + 1, lm - km + 1
    assert cp > 2 and cm >= 1
    for filt in (cp - 1, cp, cp + cm - 1, cp + cm):
        _, inf = o.seed_map(W, 6, 1, filt, koff, kstart)
        rows_p, rows_m = int(inf[0, 4]), int(inf[0, 5])
        assert rows_p == (cp if cp <= filt else 0)
        assert rows_m == (cm if cp + cm <= filt else 0)


@pytest.mark.gpu
def test_seed_map_gpu_vs_oracle(gpu_engine):
    hp = gpu_engine
    g, _, _ = synth.genome(500_000, seed=23, repeat_frac=0.3, fam_len=(300, 2000),
                           divergence=(0.0, 0.05), n_frac=0.001, n_run=(5, 50))
    gi = synth.replace_n(g, seed=1)
    ix = hp.Index.build(gi, 8)
    cuts = [100_000, 250_000, 400_000]
    koff, kstart = karyotype(len(gi), cuts)
    ix.set_karyotype(kstart, kstart + np.diff(koff) - 1)
    ix.to_device()
    o = fo.Oracle(gi, ratio=8)
    lengths = np.random.default_rng(24).integers(0, 251, 1500)
    reads, offs, _ = synth.reads(g, 1500, seed=25, sub_rate=0.01, n_rate=0.002, lengths=lengths)
    batch = hp.ReadBatch(reads, offs)
    # seeds of up to 32 bases are searched from registers, longer ones from
    # the read's bytes; then the same on the expanded index with an 8-mer
    # table (the first 8 steps of each orientation from one table read)
    cases = ((18, 5, 1000), (12, 3, 7), (25, 25, 1), (32, 7, 1000), (33, 11, 1000), (40, 40, 50))
    for form in ("compact", "kmer8"):
        if form == "kmer8":
            ix.expand()
            ix.kmer_table(8)
        for sl, st, filt in cases:
            ns, nr, regs, info = hp.seed_map_batch(ix, batch, sl, st, filt)
            want, wi = [], []
            for i in range(batch.nreads):
                W = reads[offs[i]:offs[i + 1]]
                r, inf = o.seed_map(W, sl, st, filt, koff, kstart, read=i)
                want.append(r)
                wi.append(inf)
            want = np.concatenate(want)
            wi = np.concatenate([x for x in wi if len(x)]) if any(len(x) for x in wi) else np.zeros((0, 6))
            assert ns == len(wi) and nr == len(want), (form, sl)
            for key in ("k_plus", "l_plus", "k_minus", "l_minus"):
                assert np.array_equal(info[key], wi[:, ("k_plus", "l_plus", "k_minus", "l_minus").index(key)]), \
                    (form, sl, key)
            assert np.array_equal(info["rows_plus"], wi[:, 4]) and np.array_equal(info["rows_minus"], wi[:, 5])
            for key in want.dtype.names:
                assert np.array_equal(regs[key], want[key]), (form, sl, st, filt, key)
    # a capacity below the row count stores the first rows and reports all
    ns, nr, regs, _ = hp.seed_map_batch(ix, batch, 18, 5, 1000, region_cap=10, with_info=False)
    assert len(regs) == 10 and nr > 10
    batch.free()
