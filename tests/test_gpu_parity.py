"""Parity of the HIP engine (libhpgfm.so through its C ABI) with the
reference's search path: the golden vectors of the compiled reference core,
the oracle on seeded inputs, and size-independent properties.  Bit-exact:
every record, in the reference's order."""
import numpy as np
import pytest

import fmoracle as fo
from conftest import split_golden_hits
from hpgfm import synth

pytestmark = pytest.mark.gpu


def on_device(hp, genome, ratio):
    return hp.Index.build(genome, ratio).to_device()


@pytest.fixture(scope="module", params=["golden_small", "golden_repeats"])
def gold(request, gpu_engine):
    g = request.getfixturevalue(request.param)
    ix = on_device(gpu_engine, g["genome"], int(g["ratio"]))
    batch = gpu_engine.ReadBatch(g["reads"], g["offsets"])
    yield gpu_engine, g, ix, batch
    batch.free()
    ix.close()


def test_exact_golden(gold):
    hp, g, ix, batch = gold
    for view in range(4):
        for d in (0, 1):
            rec, anc = hp.exact_batch(ix, batch, view, d, anchors=True)
            want = g[f"ex_v{view}_d{d}"]
            wa = g[f"ex_anchor_v{view}_d{d}"]
            for key in ("k", "l", "pos"):
                assert np.array_equal(rec[key], want[key]), (view, d, key)
            for key in ("k", "l", "start", "pos", "end", "dir"):
                assert np.array_equal(anc[key], wa[key]), (view, d, key)


def check_lists(counts, hits, want, cap, project=False):
    n = counts.shape[0]
    for i in range(n):
        for s in (0, 1):
            w = want[(i, s)]
            if project:
                w = w[(w["num_mismatches"] == 0) | (w["err_kind"] == fo.MISMATCH)]
            assert int(counts[i, s]) == len(w), (i, s)
            got = hits[i, s, :min(len(w), cap)]
            assert fo.hits_equal(got, w[:cap]), (i, s)


@pytest.mark.parametrize("plain", [False, True])
def test_search1_edit_golden(gold, plain):
    hp, g, ix, batch = gold
    cap = int(g["cap"])
    counts, hits = hp.search1_batch(ix, batch, cap=cap, mode=hp.MODE_EDIT, plain=plain)
    check_lists(counts, hits, split_golden_hits(g), cap)


@pytest.mark.parametrize("plain", [False, True])
def test_search1_mismatch_golden(gold, plain):
    hp, g, ix, batch = gold
    cap = int(g["cap"])
    counts, hits = hp.search1_batch(ix, batch, cap=cap, mode=hp.MODE_MISMATCH, plain=plain)
    check_lists(counts, hits, split_golden_hits(g), cap, project=True)


def test_search1_truncation(gold):
    """A list of capacity cap keeps the reference's first cap records and
    reports the full count."""
    hp, g, ix, batch = gold
    counts, hits = hp.search1_batch(ix, batch, cap=3, mode=hp.MODE_EDIT)
    check_lists(counts, hits, split_golden_hits(g), 3)


def test_search1_strand_masks(gold):
    hp, g, ix, batch = gold
    want = split_golden_hits(g)
    for mask, s in ((hp.STRAND_PLUS, 0), (hp.STRAND_MINUS, 1)):
        counts, hits = hp.search1_batch(ix, batch, cap=64, mode=hp.MODE_EDIT, strand_mask=mask)
        assert np.all(counts[:, 1 - s] == 0)
        for i in range(counts.shape[0]):
            w = want[(i, s)]
            assert int(counts[i, s]) == len(w)
            assert fo.hits_equal(hits[i, s, :min(len(w), 64)], w[:64])


def test_exact_vector_vs_oracle(gold):
    hp, g, ix, batch = gold
    o = fo.Oracle(g["genome"], ratio=int(g["ratio"]))
    reads, offs = g["reads"], g["offsets"]
    lsa = o.size_sa - 1
    for view in range(4):
        for d in (0, 1):
            vk, vl, lg = hp.exact_vector_batch(ix, batch, view, d)
            for i in range(0, len(offs) - 1, 7):
                W = reads[offs[i]:offs[i + 1]]
                L = len(W)
                good, ok, ol = o.exact_vector(view, d, W, 0, L - 1, 0, lsa)
                assert int(lg[i]) == good, (view, d, i)
                a, b = int(offs[i]), int(offs[i + 1])
                assert np.array_equal(vk[a:b], ok[:L]) and np.array_equal(vl[a:b], ol[:L])


def test_locate_golden(gold):
    hp, g, ix, batch = gold
    rows = g["sa_rows"].astype(np.int64)
    for fwd, key in ((0, "sa_bwd"), (1, "sa_fwd")):
        pos, nrows = hp.locate_batch(ix, rows, rows, cap=1, forward=fwd)
        assert np.all(nrows == 1)
        assert np.array_equal(pos[:, 0], g[key])
    # whole intervals, capped
    k = np.array([0, 5, 100, 7, 3], np.int64)
    l = np.array([40, 4, 103, 7, 2000], np.int64)
    o = fo.Oracle(g["genome"], ratio=int(g["ratio"]))
    pos, nrows = hp.locate_batch(ix, k, l, cap=64, forward=0)
    for i in range(len(k)):
        want = [o.get_sa(0, m) for m in range(k[i], min(l[i], k[i] + 63) + 1)] if k[i] <= l[i] else []
        assert int(nrows[i]) == len(want)
        assert list(pos[i, :len(want)]) == want


def test_locate_hits_after_exact(gold):
    """fmi_locate_hits reads k, l straight from exact-search records: the
    rows of each read's interval equal the oracle's SA over [k, l] (capped),
    and a dead interval (k > l, or the (n + 1, n) of an N read) has none."""
    hp, g, ix, batch = gold
    n = batch.nreads
    o = fo.Oracle(g["genome"], ratio=int(g["ratio"]))
    for view, fwd in ((hp.VIEW_BACKWARD, 0), (hp.VIEW_FORWARD, 1)):
        out = hp.DeviceBuffer(max(n, 1) * 32)
        hp.check(hp.lib().fmi_exact_batch(ix.h, view, 0 if fwd == 0 else 1, batch.d_reads.ptr,
                                          batch.d_offsets.ptr, n, out.ptr, None, None))
        hp.sync()
        rec = out.download(hp.HIT_DTYPE, n)
        cap = 16
        pos, nrows = hp.locate_hits(ix, out, n, cap, forward=fwd)
        out.free()
        for i in range(n):
            k, l = int(rec["k"][i]), int(rec["l"][i])
            want = [o.get_sa(fwd, m) for m in range(k, min(l, k + cap - 1) + 1)] if k <= l else []
            assert int(nrows[i]) == len(want), (view, i)
            assert list(pos[i, :len(want)]) == want, (view, i)


def test_inverse_locate_golden(gold):
    """get_ISA (runtime.h:88) against the golden vectors of the compiled
    reference core, every text position against the oracle, ISA(SA(i)) = i
    through both kernels, and positions beyond n."""
    hp, g, ix, batch = gold
    pos = g["sa_rows"].astype(np.uint64)
    for fwd, key in ((0, "isa_bwd"), (1, "isa_fwd")):
        assert np.array_equal(hp.inverse_locate_batch(ix, pos, forward=fwd), g[key])
    n = ix.size_sa - 1
    o = fo.Oracle(g["genome"], ratio=int(g["ratio"]))
    allpos = np.arange(n + 1, dtype=np.uint64)
    for fwd in (0, 1):
        rows = hp.inverse_locate_batch(ix, allpos, forward=fwd)
        assert [int(x) for x in rows] == [o.get_isa(fwd, m) for m in range(n + 1)]
        sa, nrows = hp.locate_batch(ix, rows.astype(np.int64), rows.astype(np.int64), cap=1,
                                    forward=fwd)
        assert np.array_equal(sa[:, 0], allpos)
    out = hp.inverse_locate_batch(ix, np.array([n + 1, 1 << 40, 0], np.uint64))
    assert list(out[:2]) == [np.uint64(2**64 - 1)] * 2 and int(out[2]) == o.get_isa(0, 0)


def test_resolve_hits(gpu_engine):
    """Position resolution of bwt.c:2158-2166 against the same formula on the
    oracle's get_SA and binsearch."""
    hp = gpu_engine
    g, _, _ = synth.genome(200_000, seed=41, repeat_frac=0.2, fam_len=(300, 2000))
    ix = hp.Index.build(g, 16)
    lens = np.array([50_000, 70_000, 80_000], np.uint64)
    starts = np.array([0, 100, 0], np.uint64)
    ix.set_karyotype(starts, starts + lens - 1)
    ix.to_device()
    reads, offs, _ = synth.reads(g, 500, length=60, seed=42, sub_rate=0.01)
    batch = hp.ReadBatch(reads, offs)
    cap, row_cap = 8, 4
    out = hp.Search1Output(batch.nreads, cap)
    hp.search1_launch(ix, batch, out, mode=hp.MODE_EDIT)
    hp.sync()
    chrom, coord = hp.resolve_hits(ix, batch, out, row_cap)
    counts = out.counts().reshape(-1, 2)
    hits = out.hits()
    o = fo.Oracle(g, ratio=16)
    koff = np.concatenate([[0], np.cumsum(lens)]).astype(np.uint64)
    n1 = o.size_sa
    checked = 0
    for i in range(batch.nreads):
        L = int(offs[i + 1] - offs[i])
        for s in (0, 1):
            for h in range(min(int(counts[i, s]), cap)):
                r = hits[i, s, h]
                lc = L
                if r["num_mismatches"]:
                    lc += {fo.DELETION: -1, fo.INSERTION: 1}.get(int(r["err_kind"]), 0)
                direction = (not r["dir"]) if s else r["dir"]
                for j in range(row_cap):
                    if r["k"] + j > r["l"]:
                        assert chrom[i, s, h, j] == -1
                        continue
                    row = int(r["k"] + j)
                    if direction:
                        key = (n1 - o.get_sa(1, row) - lc - 1) % (1 << 64)
                    else:
                        key = o.get_sa(0, row)
                    # size_t arithmetic of bwt.c:2159-2166 (wraps modulo 2^64)
                    idx = fo.binsearch(koff[:-1], key)
                    if idx >= 1 and (key + lc) % (1 << 64) <= int(koff[idx]):
                        assert chrom[i, s, h, j] == idx - 1
                        want = (int(starts[idx - 1]) + key - int(koff[idx - 1])) % (1 << 64)
                        assert int(coord[i, s, h, j]) == want
                    else:
                        assert chrom[i, s, h, j] == -1
                    checked += 1
    assert checked > 300
    out.free()
    batch.free()


@pytest.mark.parametrize("mode", [0, 1])
def test_search1_vs_oracle_with_n(gpu_engine, mode):
    """1 Mbp with repeats; reads of 40..160 bp with substitutions, random
    reads and N bases; bit-exact against the oracle, both modes."""
    hp = gpu_engine
    g, _, _ = synth.genome(1_000_000, seed=51, repeat_frac=0.2, fam_len=(300, 6000),
                           divergence=(0.02, 0.15))
    rng = np.random.default_rng(52)
    lengths = rng.integers(40, 161, 6000)
    reads, offs, _ = synth.reads(g, 6000, seed=53, sub_rate=0.01, random_frac=0.05,
                                 n_rate=0.003, lengths=lengths)
    ix = on_device(hp, g, 16)
    batch = hp.ReadBatch(reads, offs)
    cap = 32
    counts, hits = hp.search1_batch(ix, batch, cap=cap, mode=mode)
    o = fo.Oracle(bwts=(*ix.bwt(0), *ix.bwt(1)), ratio=16)
    _, ocounts, ohits = o.search1_batch(reads, offs, cap=cap, mismatch_only=bool(mode))
    ocounts = ocounts.reshape(-1, 2)
    ohits = ohits.reshape(-1, 2, cap)
    assert np.array_equal(counts, ocounts)
    for i in range(batch.nreads):
        for s in (0, 1):
            c = min(int(counts[i, s]), cap)
            assert fo.hits_equal(hits[i, s, :c], ohits[i, s, :c]), (i, s)
    batch.free()


def test_results_independent_of_batch_mix(gpu_engine):
    """A read's records are the same alone, in a small batch and inside a
    large mixed batch (serving determinism)."""
    hp = gpu_engine
    g, _, _ = synth.genome(500_000, seed=61, repeat_frac=0.2)
    ix = on_device(hp, g, 16)
    reads, offs, _ = synth.reads(g, 3000, seed=62, sub_rate=0.01,
                                 lengths=np.random.default_rng(63).integers(50, 251, 3000))
    big = hp.ReadBatch(reads, offs)
    cb, hb = hp.search1_batch(ix, big, cap=16, mode=hp.MODE_MISMATCH)
    for lo, hi in ((0, 1), (17, 40), (2500, 3000)):
        sub = hp.ReadBatch(reads[offs[lo]:offs[hi]], offs[lo:hi + 1] - offs[lo])
        cs, hs = hp.search1_batch(ix, sub, cap=16, mode=hp.MODE_MISMATCH)
        assert np.array_equal(cs, cb[lo:hi])
        for i in range(hi - lo):
            for s in (0, 1):
                c = min(int(cs[i, s]), 16)
                assert fo.hits_equal(hs[i, s, :c], hb[lo + i, s, :c])
        sub.free()
    big.free()


def test_results_independent_of_tuning(gpu_engine):
    """Workgroups per CU, queue chunk and refill threshold change the
    schedule, never the records."""
    hp = gpu_engine
    g, _, _ = synth.genome(400_000, seed=81, repeat_frac=0.3)
    ix = on_device(hp, g, 16)
    reads, offs, _ = synth.reads(g, 4000, seed=82, sub_rate=0.01,
                                 lengths=np.random.default_rng(83).integers(30, 200, 4000))
    batch = hp.ReadBatch(reads, offs)
    base_c, base_h = hp.search1_batch(ix, batch, cap=8, mode=hp.MODE_EDIT, plain=True)
    try:
        for wg, chunk, refill in ((1, 1, 1), (8, 64, 16), (2, 1024, 64), (16, 7, 33)):
            hp.set_tuning(wg, chunk, refill)
            c, h = hp.search1_batch(ix, batch, cap=8, mode=hp.MODE_EDIT)
            assert np.array_equal(c, base_c)
            assert np.array_equal(h, base_h)
    finally:
        hp.set_tuning(0, 0, 0)
    batch.free()


def test_count_blocks_bounds(gpu_engine):
    """Counted blocks lie between one and two per LF query."""
    hp = gpu_engine
    g, _, _ = synth.genome(300_000, seed=71)
    ix = on_device(hp, g, 16)
    reads, offs, _ = synth.reads(g, 2000, length=150, seed=72, sub_rate=0.01)
    batch = hp.ReadBatch(reads, offs)
    blocks, steps = hp.search1_count_blocks(ix, batch, mode=hp.MODE_MISMATCH)
    assert steps > 2000 * 150
    assert steps <= blocks <= 2 * steps
    batch.free()
