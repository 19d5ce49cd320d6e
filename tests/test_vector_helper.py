"""The mapping layer's 1-error search, BWSearch1VectorHelper
(search/search.c:1491-1709) over the interval vectors of
BWExactSearchVectorBackward/Forward (bwt.c:1940-1950, 1994-2003; seed windows
bwt.c:1447-1479).

CPU: the oracle against the golden vectors of the compiled reference (whole
reads and windows), against the compiled reference live when it is built,
and its mismatch records against a brute-force scan of the text (every
single substitution that occurs, once, with its occurrence count).
GPU: fmi_search1_vector_batch through the C ABI against the golden vectors
and the oracle (reads with N, windows, the mismatch projection, truncation,
strand masks)."""
import numpy as np
import pytest

import fmoracle as fo
from conftest import split_golden_hits
from hpgfm import synth


@pytest.fixture(scope="module", params=["golden_small", "golden_repeats"])
def golden(request):
    g = request.getfixturevalue(request.param)
    return g, fo.Oracle(g["genome"], ratio=int(g["ratio"]))


def test_golden_vector_helper(golden):
    g, o = golden
    reads, offs = g["reads"], g["offsets"]
    cap = int(g["cap"])
    whole = split_golden_hits(g, "s1v")
    win = split_golden_hits(g, "s1w")
    for i in range(len(offs) - 1):
        W = reads[offs[i]:offs[i + 1]]
        a, b = (int(x) for x in g["s1w_bounds"][i])
        for s in (0, 1):
            n, h = o.search1_vector(s, W, 0, len(W) - 1, cap=cap)
            assert n == int(g["s1v_counts"][i, s]) and fo.hits_equal(h, whole[(i, s)]), (i, s)
            n, h = o.search1_vector(s, W, a, b, cap=cap)
            assert n == int(g["s1w_counts"][i, s]) and fo.hits_equal(h, win[(i, s)]), (i, s, a, b)


def count_occ(text, pat):
    t = text.tobytes()
    p = pat.tobytes()
    c, i = 0, t.find(p)
    while i >= 0:
        c += 1
        i = t.find(p, i + 1)
    return c


def substitution_positions(hits, start):
    """Position of the substituted base of every MISMATCH record.  A mismatch
    at W[start] that follows an insertion record in the first block carries
    that record's err_pos, the base b in 0..3 (search.c:1556 writes b into
    it); with start >= 4 no other record of the backward half has err_pos
    below start, so such a record is at W[start]."""
    assert start >= 4
    out = []
    for h in hits:
        if h["num_mismatches"] == 0 or int(h["err_kind"]) != fo.MISMATCH:
            continue
        p = int(h["err_pos"])
        if h["dir"] == 0 and p < start:
            p = start
        out.append((p, int(h["err_base"]), int(h["l"] - h["k"] + 1)))
    return out


@pytest.mark.parametrize("seed", [4, 5])
def test_mismatch_records_are_hamming1(seed):
    """Every single substitution of W[start..end] (of its reverse complement
    on strand 1) that occurs in X appears once, with l - k + 1 = its count;
    the exact record counts W itself."""
    g, _, _ = synth.genome(3000, seed=seed, repeat_frac=0.3, fam_len=(100, 400),
                           divergence=(0.0, 0.03), copies_per_family=6)
    o = fo.Oracle(g, ratio=8)
    reads, offs, _ = synth.reads(g, 100, length=30, seed=seed + 10, sub_rate=0.04)
    rng = np.random.default_rng(seed)
    for i in range(len(offs) - 1):
        R = reads[offs[i]:offs[i + 1]]
        a = int(rng.integers(4, 8))
        b = int(rng.integers(len(R) - 6, len(R)))
        for s in (0, 1):
            n, hits = o.search1_vector(s, R, a, b, cap=4096)
            got = {}
            for (p, base, c) in substitution_positions(hits, a):
                assert (p, base) not in got, (i, s, p, base)
                got[(p, base)] = c
            ex = [int(h["l"] - h["k"] + 1) for h in hits if h["num_mismatches"] == 0]
            W = R[a:b + 1]
            c0 = count_occ(g, W if s == 0 else synth.revcomp(W))
            assert ex == ([c0] if c0 else [])
            want = {}
            for p in range(len(W)):
                for base in range(4):
                    if base == W[p]:
                        continue
                    V = W.copy()
                    V[p] = base
                    c = count_occ(g, V if s == 0 else synth.revcomp(V))
                    if c:
                        want[(a + p, base)] = c
            assert got == want, (i, s)
            # the mismatch projection keeps the same records in order
            n_mm, mm = o.search1_vector(s, R, a, b, cap=4096, mismatch_only=True)
            keep = (hits["num_mismatches"] == 0) | (hits["err_kind"] == fo.MISMATCH)
            assert fo.hits_equal(mm, hits[keep])


ref_only = pytest.mark.skipif(not fo.ref_available(), reason="oracle/_ref not built")


@ref_only
def test_oracle_matches_reference_vector_helper():
    g, _, _ = synth.genome(30000, seed=61, repeat_frac=0.3, fam_len=(200, 1500),
                           divergence=(0.0, 0.05))
    o = fo.Oracle(g, ratio=8)
    r = fo.RefCore(oracle=o)
    lens = np.random.default_rng(62).integers(3, 110, 250)
    reads, offs, _ = synth.reads(g, 250, seed=63, sub_rate=0.02, lengths=lens)
    rng = np.random.default_rng(64)
    for i in range(len(offs) - 1):
        W = reads[offs[i]:offs[i + 1]]
        L = len(W)
        a = int(rng.integers(0, L - 2))
        b = int(rng.integers(a + 2, L))
        for (st, en) in ((0, L - 1), (a, b)):
            for s in (0, 1):
                n1, h1 = o.search1_vector(s, W, st, en)
                n2, h2, full = r.search1_vector(s, W, st, en)
                assert n1 == n2 and not full and fo.hits_equal(h1, h2), (i, st, en, s)
                n3, h3, full3 = r.search1_vector(s, W, st, en, cap=2)
                assert fo.hits_equal(h3, h1[:2]) and full3 == (n1 > 2)
    r.close()


# ------------------------------------------------------------------ GPU

@pytest.fixture(scope="module", params=["golden_small", "golden_repeats"])
def vgold(request, gpu_engine):
    g = request.getfixturevalue(request.param)
    ix = gpu_engine.Index.build(g["genome"], int(g["ratio"])).to_device()
    batch = gpu_engine.ReadBatch(g["reads"], g["offsets"])
    yield gpu_engine, g, ix, batch
    batch.free()
    ix.close()


def check(counts, hits, want, cap, project=False):
    for (i, s), w in want.items():
        if project:
            w = w[(w["num_mismatches"] == 0) | (w["err_kind"] == fo.MISMATCH)]
        assert int(counts[i, s]) == len(w), (i, s)
        assert fo.hits_equal(hits[i, s, :min(len(w), cap)], w[:cap]), (i, s)


@pytest.mark.gpu
@pytest.mark.parametrize("mode", ["edit", "mismatch"])
def test_gpu_vector_helper_golden(vgold, mode):
    hp, g, ix, batch = vgold
    cap = int(g["cap"])
    m = hp.MODE_EDIT if mode == "edit" else hp.MODE_MISMATCH
    counts, hits = hp.search1_vector_batch(ix, batch, cap=cap, mode=m)
    check(counts, hits, split_golden_hits(g, "s1v"), cap, project=mode == "mismatch")
    counts, hits = hp.search1_vector_batch(ix, batch, cap=cap, mode=m, bounds=g["s1w_bounds"])
    check(counts, hits, split_golden_hits(g, "s1w"), cap, project=mode == "mismatch")


@pytest.mark.gpu
def test_gpu_vector_helper_truncation_and_strands(vgold):
    hp, g, ix, batch = vgold
    want = split_golden_hits(g, "s1v")
    counts, hits = hp.search1_vector_batch(ix, batch, cap=3)
    check(counts, hits, want, 3)
    for mask, s in ((hp.STRAND_PLUS, 0), (hp.STRAND_MINUS, 1)):
        counts, hits = hp.search1_vector_batch(ix, batch, cap=64, strand_mask=mask)
        assert np.all(counts[:, 1 - s] == 0)
        for i in range(counts.shape[0]):
            w = want[(i, s)]
            assert int(counts[i, s]) == len(w)
            assert fo.hits_equal(hits[i, s, :min(len(w), 64)], w[:64])


@pytest.mark.gpu
def test_gpu_vector_helper_vs_oracle_with_n(gpu_engine):
    """Reads with N bases (code 4, a symbol that never occurs), ragged
    lengths 0..250 and random windows, against the oracle."""
    hp = gpu_engine
    g, _, _ = synth.genome(300_000, seed=71, repeat_frac=0.3, fam_len=(300, 2000),
                           divergence=(0.0, 0.05))
    ix = hp.Index.build(g, 16).to_device()
    o = fo.Oracle(g, ratio=16)
    lens = np.random.default_rng(72).integers(0, 251, 1500)
    reads, offs, _ = synth.reads(g, 1500, seed=73, sub_rate=0.01, n_rate=0.003, lengths=lens)
    rng = np.random.default_rng(74)
    bounds = np.zeros((1500, 2), np.int16)
    for i in range(1500):
        L = int(lens[i])
        if L >= 3:
            a = int(rng.integers(0, L - 2))
            bounds[i] = (a, int(rng.integers(a + 2, L)))
        else:
            bounds[i] = (0, L - 1)
    batch = hp.ReadBatch(reads, offs)
    cap = 48
    for bnd in (None, bounds):
        counts, hits = hp.search1_vector_batch(ix, batch, cap=cap, bounds=bnd)
        for i in range(1500):
            W = reads[offs[i]:offs[i + 1]]
            a, b = (0, len(W) - 1) if bnd is None else (int(bnd[i, 0]), int(bnd[i, 1]))
            for s in (0, 1):
                n, h = o.search1_vector(s, W, a, b, cap=cap)
                assert int(counts[i, s]) == n, (i, s, a, b)
                assert fo.hits_equal(hits[i, s, :min(n, cap)], h), (i, s)
    batch.free()
    ix.close()
