"""The k-error search BWSearchCPU (search/search.c:1171-1405) and its helpers
(search.c:3-1169, search.h:66-198): W split into len // fragsize fragments,
up to fragments - 1 errors, breadth-first over bounded result lists.

CPU: the oracle against the golden vectors of the compiled reference (final
list, anchors, return value), against the compiled reference live at list
capacities that truncate, and against a brute-force scan of the text: every
substitution-only record is an occurrence of its variant with l - k + 1 =
its count, and with two fragments every one-substitution occurrence is
found.
GPU: fmi_search_k_batch through the C ABI against the golden vectors and the
oracle."""
import itertools

import numpy as np
import pytest

import fmoracle as fo
from hpgfm import synth


def split(g, key, counts_key):
    counts = g[counts_key]
    flat = g[key]
    out, o = {}, 0
    for i in range(counts.shape[0]):
        for s in (0, 1):
            c = int(counts[i, s])
            out[(i, s)] = flat[o:o + c]
            o += c
    return out


@pytest.fixture(scope="module", params=["golden_small", "golden_repeats"])
def kgold(request):
    g = request.getfixturevalue(request.param)
    return g, fo.Oracle(g["genome"], ratio=int(g["ratio"]))


def test_golden_search_k(kgold):
    g, o = kgold
    fin = split(g, "sk_hits", "sk_counts")
    anc = split(g, "sk_anchors", "sk_anchor_counts")
    cap = int(g["sk_cap"])
    reads, offs = g["reads"], g["offsets"]
    for i in range(len(offs) - 1):
        W = reads[offs[i]:offs[i + 1]]
        for s in (0, 1):
            full, f, a = o.search_k(s, W, int(g["sk_fragsize"][i]), cap, 16)
            assert full == bool(g["sk_full"][i, s]), (i, s)
            assert fo.hitsk_equal(f, fin[(i, s)]), (i, s)
            assert fo.hitsk_equal(a, anc[(i, s)]), (i, s)


ref_only = pytest.mark.skipif(not fo.ref_available(), reason="oracle/_ref not built")


@ref_only
def test_oracle_matches_reference_search_k():
    g, _, _ = synth.genome(20000, seed=71, repeat_frac=0.3, fam_len=(200, 1500),
                           divergence=(0.0, 0.05))
    o = fo.Oracle(g, ratio=8)
    r = fo.RefCore(oracle=o)
    lens = np.random.default_rng(72).integers(10, 130, 120)
    reads, offs, _ = synth.reads(g, 120, seed=73, sub_rate=0.03, lengths=lens)
    for i in range(len(offs) - 1):
        W = reads[offs[i]:offs[i + 1]]
        for fs in (max(len(W) // 2, 5), max(len(W) // 4, 5)):
            for lc, ac in ((1500, 16), (5, 2)):
                for s in (0, 1):
                    a = o.search_k(s, W, fs, lc, ac)
                    b = r.search_k(s, W, fs, lc, ac)
                    assert a[0] == b[0] and fo.hitsk_equal(a[1], b[1]) and fo.hitsk_equal(a[2], b[2])
    r.close()


def count_occ(text, pat):
    t, p = text.tobytes(), pat.tobytes()
    c, i = 0, t.find(p)
    while i >= 0:
        c += 1
        i = t.find(p, i + 1)
    return c


def interval(o, view, d, V):
    lsa = o.size_sa - 1
    L = len(V)
    r = o.exact(view, d, V, 0, lsa, 0, L - 1 if d == 0 else 0, L - 1)
    return int(r["k"]), int(r["l"])


@pytest.mark.parametrize("k", [1, 2])
def test_substitutions_found_by_brute_force(k):
    """Every substitution-only record is an occurrence of its variant, with
    l - k + 1 = its count.  With two fragments (k = 1) every variant of W
    with one substitution that occurs in X (in its reverse complement for
    strand 1) is reported: one fragment is exact, and some final record
    holds the variant's interval on the record's direction (as a
    substitution or as an insertion + deletion pair over the same text).
    With three fragments the reference's pruning can miss a two-substitution
    occurrence (the oracle does the same, pinned against the compiled
    reference above), so k = 2 checks soundness only."""
    g, _, _ = synth.genome(4000, seed=80 + k, repeat_frac=0.3, fam_len=(100, 400),
                           divergence=(0.0, 0.03), copies_per_family=6)
    o = fo.Oracle(g, ratio=8)
    L = 18 if k == 2 else 24
    reads, offs, _ = synth.reads(g, 40, length=L, seed=90 + k, sub_rate=0.05)
    checked = 0
    for i in range(len(offs) - 1):
        W = reads[offs[i]:offs[i + 1]]
        for s in (0, 1):
            vb = fo.VIEW_FORWARD_REV if s else fo.VIEW_BACKWARD
            vf = fo.VIEW_BACKWARD_REV if s else fo.VIEW_FORWARD
            _, fin, _ = o.search_k(s, W, L // (k + 1), 5000, 16)
            got = {(int(h["dir"]), int(h["k"]), int(h["l"])) for h in fin}
            for h in fin:
                m = int(h["num_mismatches"])
                if any(int(h["err_kind"][j]) != fo.MISMATCH for j in range(m)):
                    continue
                V = W.copy()
                for j in range(m):
                    V[int(h["err_pos"][j])] = int(h["err_base"][j])
                c = count_occ(g, V if s == 0 else synth.revcomp(V))
                assert c == int(h["l"] - h["k"] + 1), (i, s)
            for e in range(k + 1):
                for ps in itertools.combinations(range(L), e):
                    for bs in itertools.product(range(3), repeat=e):
                        V = W.copy()
                        for p, b in zip(ps, bs):
                            V[p] = (W[p] + 1 + b) % 4
                        if not count_occ(g, V if s == 0 else synth.revcomp(V)):
                            continue
                        if k == 1:
                            assert ((0, *interval(o, vb, 0, V)) in got or
                                    (1, *interval(o, vf, 1, V)) in got), (i, s, ps, bs)
                        checked += 1
    assert checked >= 20


# ------------------------------------------------------------------ GPU

@pytest.fixture(scope="module", params=["golden_small", "golden_repeats"])
def kgpu(request, gpu_engine):
    g = request.getfixturevalue(request.param)
    ix = gpu_engine.Index.build(g["genome"], int(g["ratio"])).to_device()
    batch = gpu_engine.ReadBatch(g["reads"], g["offsets"])
    yield gpu_engine, g, ix, batch
    batch.free()
    ix.close()


@pytest.mark.gpu
@pytest.mark.parametrize("kmer", [0, 5])
@pytest.mark.parametrize("expanded", [False, True])
def test_gpu_search_k_golden(kgpu, expanded, kmer):
    """The golden records and anchors, with and without a k-mer table (the
    exact searches and D walks from the whole index start with its entry)."""
    hp, g, ix, batch = kgpu
    if expanded:
        ix.expand()
    ix.kmer_table(kmer)
    try:
        cap = int(g["sk_cap"])
        res = hp.search_k_batch(ix, batch, g["sk_fragsize"], list_cap=cap, anchor_cap=16)
        fin = split(g, "sk_hits", "sk_counts")
        anc = split(g, "sk_anchors", "sk_anchor_counts")
        for i in range(batch.nreads):
            for s in (0, 1):
                assert bool(res.full[i, s]) == bool(g["sk_full"][i, s]), (i, s)
                assert fo.hitsk_equal(res.final(i, s), fin[(i, s)]), (i, s)
                assert fo.hitsk_equal(res.anchors(i, s), anc[(i, s)]), (i, s)
    finally:
        ix.kmer_table(0)


@pytest.mark.gpu
@pytest.mark.parametrize("kmer", [0, 8])
@pytest.mark.parametrize("expanded", [False, True])
def test_gpu_search_k_vs_oracle(gpu_engine, expanded, kmer):
    """Repeats, N bases, fragments of 2-5, list capacities that truncate;
    on the expanded index one-row stretches run on the text; with a k-mer
    table, searches from the whole index start with its entry."""
    hp = gpu_engine
    g, _, _ = synth.genome(200_000, seed=75, repeat_frac=0.4, fam_len=(300, 2000),
                           divergence=(0.0, 0.05))
    ix = hp.Index.build(g, 16).to_device()
    if expanded:
        ix.expand()
    ix.kmer_table(kmer)
    o = fo.Oracle(g, ratio=16)
    lens = np.random.default_rng(76).integers(8, 160, 800)
    reads, offs, _ = synth.reads(g, 800, seed=77, sub_rate=0.03, n_rate=0.003, lengths=lens)
    fs = np.maximum(lens // np.random.default_rng(78).integers(2, 6, 800), 4).astype(np.int16)
    batch = hp.ReadBatch(reads, offs)
    for lc, ac in ((200, 16), (6, 3)):
        res = hp.search_k_batch(ix, batch, fs, list_cap=lc, anchor_cap=ac)
        for i in range(batch.nreads):
            W = reads[offs[i]:offs[i + 1]]
            for s in (0, 1):
                full, f, a = o.search_k(s, W, int(fs[i]), lc, ac)
                assert bool(res.full[i, s]) == full, (lc, i, s)
                assert fo.hitsk_equal(res.final(i, s), f), (lc, i, s)
                assert fo.hitsk_equal(res.anchors(i, s), a), (lc, i, s)
    batch.free()
    ix.close()


@pytest.mark.gpu
def test_gpu_search_k_front_and_triage_paths(gpu_engine):
    """The paths around the front kernel and the triage: reads of at most
    and of more than 256 bases in one batch (the front covers the first),
    two and three fragments (a task whose every exact fragment search dies is
    finished by the triage), anchor lists of one and two records (the
    triage's add_result on a full list ends the search with full = 1) and
    each strand alone, against the oracle task by task."""
    hp = gpu_engine
    g, _, _ = synth.genome(200_000, seed=81, repeat_frac=0.3, fam_len=(300, 2000),
                           divergence=(0.0, 0.05))
    ix = hp.Index.build(g, 16).to_device().expand()
    ix.kmer_table(9)
    o = fo.Oracle(g, ratio=16)
    rng = np.random.default_rng(82)
    lens = np.concatenate([rng.integers(30, 256, 500), rng.integers(257, 400, 60)])
    rng.shuffle(lens)
    reads, offs, _ = synth.reads(g, len(lens), seed=83, sub_rate=0.02, n_rate=0.003,
                                 lengths=lens, random_frac=0.2)
    fs = np.where(rng.random(len(lens)) < 0.5, lens // 3, lens // 2).astype(np.int16)
    batch = hp.ReadBatch(reads, offs)
    try:
        for lc, ac, mask in ((128, 1, 3), (4, 2, 3), (64, 16, 1), (64, 16, 2)):
            res = hp.search_k_batch(ix, batch, fs, list_cap=lc, anchor_cap=ac, strand_mask=mask)
            for i in range(batch.nreads):
                W = reads[offs[i]:offs[i + 1]]
                for s in (0, 1):
                    if not mask & (1 << s):
                        assert res.full[i, s] == 0 and len(res.final(i, s)) == 0, (lc, ac, i, s)
                        assert len(res.anchors(i, s)) == 0, (lc, ac, i, s)
                        continue
                    full, f, a = o.search_k(s, W, int(fs[i]), lc, ac)
                    assert bool(res.full[i, s]) == full, (lc, ac, mask, i, s)
                    assert fo.hitsk_equal(res.final(i, s), f), (lc, ac, mask, i, s)
                    assert fo.hitsk_equal(res.anchors(i, s), a), (lc, ac, mask, i, s)
    finally:
        batch.free()
        ix.close()
