"""The 1-mismatch search on a 3.1 Gbp synthetic reference (41% GC, 45% repeat
families at 1-20% divergence, 0.1% N runs) with 2M reads of 150 bp from both
strands at 1% substitutions: the index size where SA values, rows and text
positions pass 2^31 and the index no longer fits the caches.

The whole batch is held to the size-independent properties of the 64 Mbp
test: the compact index, the expanded index and the expanded index with the
k = 15 table, under both load modes, give the same counts and records; every
read drawn without a substitution or an N has an exact record on its strand
whose interval holds the row of the drawn position, the row here found by
the engine's own get_ISA (fmi_inverse_locate_batch, runtime.h:88) on the
compact index, which walks LF from the sampled ISA.

Independently of the engine: 2000 reads give the oracle's BWSearch1CPU
records in both modes (the oracle built from this index's BWTs and samples,
on the reference's own occurrence layout), rows past 2^31 among them; and
error-free reads resolved to coordinates (fmi_resolve_hits, bwt.c:2158-2166)
land on their drawn position, every resolved coordinate holding the read in
the indexed text, coordinates past 2^31 among them.  On the same index: the
exact seed regions of the whole 2M x 250 bp seed workload against the
positions its seeds were cut from, and the serving loop's requests under
Poisson load against the batch call alone."""
import os

import numpy as np
import pytest

import fmoracle as fo
from hpgfm import synth
from test_gpu_fullsize import same_records

pytestmark = pytest.mark.gpu

NBP = 3_100_000_000
NREADS = 2_000_000
LEN = 150
CAP = 16


def say(msg):
    print(f"[fullsize 3.1 Gbp] {msg}", flush=True)


@pytest.fixture(scope="module")
def huge(gpu_engine):
    hp = gpu_engine
    say(f"genome {NBP} bp, {NREADS} x {LEN} bp reads")
    g, _, _ = synth.genome(NBP, seed=1, gc=0.41, repeat_frac=0.45, fam_len=(300, 6000),
                           divergence=(0.01, 0.20), n_frac=0.001)
    reads, offs, truth = synth.reads(g, NREADS, length=LEN, seed=2, sub_rate=0.01)
    say("index")
    text = synth.replace_n(g, seed=1)
    ix = hp.Index.build_gpu(text, 16)
    batch = hp.ReadBatch(reads, offs)
    out = hp.Search1Output(NREADS, CAP)
    hp.search1_launch(ix, batch, out, mode=hp.MODE_MISMATCH)
    hp.sync()
    compact = (out.counts().reshape(-1, 2), out.hits())
    state = dict(hp=hp, g=g, text=text, reads=reads, truth=truth, ix=ix, batch=batch, out=out,
                 compact=compact)
    yield state
    hp.set_load_mode(hp.LOADS_AUTO)
    out.free()
    batch.free()
    ix.close()


def test_3g_planted_positions(huge):
    hp, g, ix = huge["hp"], huge["g"], huge["ix"]
    say("planted positions")
    pos = huge["truth"]["pos"].astype(np.int64)
    strand = huge["truth"]["strand"].astype(np.int64)
    mat = huge["reads"].reshape(NREADS, LEN)
    clean = np.zeros(NREADS, bool)
    for lo in range(0, NREADS, 100_000):
        hi = min(lo + 100_000, NREADS)
        src = g[pos[lo:hi, None] + np.arange(LEN)[None, :]]
        rc = strand[lo:hi] == 1
        src[rc] = (3 - src[rc])[:, ::-1]
        m = mat[lo:hi]
        clean[lo:hi] = np.all(src == m, axis=1) & np.all(m < 4, axis=1)
    counts, hits = huge["compact"]
    idx = np.nonzero(clean)[0]
    c = counts[idx, strand[idx]]
    assert len(idx) > 300_000 and np.all(c >= 1)
    idx = idx[c <= CAP]
    assert len(idx) > 0.7 * clean.sum()
    rec = hits[idx, strand[idx], :]
    valid = np.arange(CAP)[None, :] < counts[idx, strand[idx]][:, None]
    exact = valid & (rec["num_mismatches"] == 0)
    assert np.all(exact.any(axis=1))
    assert int(pos[idx].max()) >= 1 << 31
    n1 = NBP + 1
    lc = rec["end"].astype(np.int64) - rec["start"].astype(np.int64) + 1
    p = np.broadcast_to(pos[idx][:, None], rec.shape)
    # BWT of the text: SA[row] = pos; of the reversed text: SA[row] = n1 - lc - 1 - pos
    direction = np.where(strand[idx][:, None] == 1, rec["dir"] == 0, rec["dir"] == 1)
    row = np.full(rec.shape, -1, np.int64)
    for fwd in (0, 1):
        sel = exact & (direction == bool(fwd))
        q = (n1 - lc - 1 - p)[sel] if fwd else p[sel]
        row[sel] = hp.inverse_locate_batch(ix, q.astype(np.uint64), forward=fwd).astype(np.int64)
    found = exact & (rec["k"] <= row) & (row <= rec["l"])
    assert np.all(found.any(axis=1))


def test_3g_seed_regions_planted(huge):
    """Exact seeds (bwt_map_exact_seed, bwt.c:1200-1294) of the whole seed
    workload, 2M reads of 250 bp, 18 bases at stride 5 (94M seeds, one call),
    on the compact index: every seed without a substitution or an N yields,
    on its read's strand, the region of the position it was cut from
    (1-based, one contig from coordinate 0), unless that strand passed the
    1000-row filter, in which case the strand has no region for the seed at
    all; every such region spans the seed's 18 bases minus one (the
    reference's one-short length)."""
    hp, g, ix = huge["hp"], huge["g"], huge["ix"]
    n, L, sl, st = 2_000_000, 250, 18, 5
    say(f"exact seeds of {n} x {L} bp")
    reads, offs, truth = synth.reads(g, n, length=L, seed=4, sub_rate=0.01)
    batch = hp.ReadBatch(reads, offs)
    try:
        nseeds, nrows, regs, _ = hp.seed_map_batch(ix, batch, sl, st, 1000, with_info=False)
    finally:
        batch.free()
    s0 = np.arange(0, L - sl + 1, st)
    assert nseeds == n * len(s0) and len(regs) == nrows
    say(f"{nseeds} seeds, {nrows} regions; checking")
    pos = truth["pos"].astype(np.int64)
    strand = truth["strand"].astype(np.int64)
    mat = reads.reshape(n, L)

    def key(read, seed, s, start):
        return ((read.astype(np.uint64) << np.uint64(39)) | (seed.astype(np.uint64) << np.uint64(33))
                | (s.astype(np.uint64) << np.uint64(32)) | start.astype(np.uint64))

    bounds = np.searchsorted(regs["read"], np.arange(0, n + 1, 250_000))
    hits = misses = 0
    top = 0
    for c, lo in enumerate(range(0, n, 250_000)):
        hi = min(n, lo + 250_000)
        src = g[pos[lo:hi, None] + np.arange(L)[None, :]]
        rc = strand[lo:hi] == 1
        src[rc] = (3 - src[rc])[:, ::-1]
        m = mat[lo:hi]
        bad = np.concatenate([np.zeros((hi - lo, 1), np.int32),
                              np.cumsum((src != m) | (m > 3), axis=1, dtype=np.int32)], axis=1)
        clean = (bad[:, s0 + sl] - bad[:, s0]) == 0
        ri, q = np.nonzero(clean)
        ri = ri + lo
        off = s0[q]
        tpos = np.where(strand[ri] == 0, pos[ri] + off, pos[ri] + L - off - sl)
        top = max(top, int(tpos.max()))
        v = regs[bounds[c]:bounds[c + 1]]
        v = v[v["valid"] == 1]
        have = key(v["read"], v["seed_id"], v["strand"], v["start"])
        want = key(ri, q, strand[ri], tpos + 1)
        hit = np.isin(want, have)
        groups = np.unique(have >> np.uint64(32))
        # a missing planted region only where the seed's strand has no region at all
        assert not np.any(np.isin(want[~hit] >> np.uint64(32), groups)), c
        sel = np.isin(have, want)
        assert np.all(v["end"][sel] - v["start"][sel] == sl - 1), c
        hits += int(hit.sum())
        misses += int((~hit).sum())
    assert top >= 1 << 31
    assert hits / (hits + misses) > 0.9, (hits, misses)


def test_3g_sample_vs_oracle(huge):
    """2000 reads against the oracle's BWSearch1CPU (search/search.c:1876-1933)
    in mismatch mode (the compact index's records of the whole batch) and in
    edit mode; the oracle is built from this index's BWTs and S / R samples
    on the reference's own per-symbol occurrence layout."""
    hp, ix = huge["hp"], huge["ix"]
    say("oracle from the index's BWTs and samples")
    o = fo.Oracle(bwts=(*ix.bwt(0), *ix.bwt(1)), ratio=0)
    o.set_sampled(ix.sa_ratio, *ix.samples(0), *ix.samples(1))
    threads = max(1, min(16, len(os.sched_getaffinity(0))))
    pick = np.sort(np.random.default_rng(33).choice(NREADS, 2000, replace=False))
    rs = huge["reads"].reshape(NREADS, LEN)[pick].reshape(-1)
    os_ = (np.arange(len(pick) + 1) * LEN).astype(np.uint64)
    counts, hits = huge["compact"]
    _, oc, oh = o.search1_batch(rs, os_, cap=CAP, mismatch_only=True, threads=threads)
    assert same_records((counts[pick], hits[pick]), (oc.reshape(-1, 2), oh.reshape(-1, 2, CAP)))
    stored = np.arange(CAP)[None, None, :] < np.minimum(oc.reshape(-1, 2), CAP)[:, :, None]
    assert int(oh.reshape(-1, 2, CAP)[stored]["l"].max()) >= 1 << 31
    say("mismatch mode: ok")
    sub = hp.ReadBatch(rs, os_)
    try:
        ge = hp.search1_batch(ix, sub, cap=CAP, mode=hp.MODE_EDIT)
    finally:
        sub.free()
    _, oc, oh = o.search1_batch(rs, os_, cap=CAP, mismatch_only=False, threads=threads)
    assert same_records(ge, (oc.reshape(-1, 2), oh.reshape(-1, 2, CAP)))
    say("edit mode: ok")
    o.close()


def test_3g_planted_coordinates(huge):
    """Error-free reads resolved to genome coordinates (fmi_resolve_hits,
    bwt.c:2158-2166; one contig from coordinate 0): the exact record on the
    read's strand resolves to the drawn position when its interval fits the
    row capacity, and every coordinate it resolves to holds the read (its
    reverse complement on the - strand) in the indexed text."""
    hp, ix, text = huge["hp"], huge["ix"], huge["text"]
    say("planted coordinates")
    pos = huge["truth"]["pos"].astype(np.int64)
    strand = huge["truth"]["strand"].astype(np.int64)
    mat = huge["reads"].reshape(NREADS, LEN)
    cand = np.random.default_rng(34).choice(NREADS, 150_000, replace=False)
    src = text[pos[cand, None] + np.arange(LEN)[None, :]]
    rc = strand[cand] == 1
    src[rc] = (3 - src[rc])[:, ::-1]
    m = mat[cand]
    pick = np.sort(cand[np.all(src == m, axis=1) & np.all(m < 4, axis=1)][:20_000])
    assert len(pick) == 20_000 and int(pos[pick].max()) >= 1 << 31
    rs = mat[pick].reshape(-1)
    os_ = (np.arange(len(pick) + 1) * LEN).astype(np.uint64)
    row_cap = 16
    sub = hp.ReadBatch(rs, os_)
    out = hp.Search1Output(len(pick), CAP)
    try:
        hp.search1_launch(ix, sub, out, mode=hp.MODE_MISMATCH)
        hp.sync()
        chrom, coord = hp.resolve_hits(ix, sub, out, row_cap)
        counts = out.counts().reshape(-1, 2)
        recs = out.hits()
    finally:
        out.free()
        sub.free()
    found = 0
    for j, i in enumerate(pick):
        s = int(strand[i])
        want = mat[i] if s == 0 else synth.revcomp(mat[i])
        on_target = False
        for h in range(min(int(counts[j, s]), CAP)):
            r = recs[j, s, h]
            if r["num_mismatches"]:
                continue
            nrow = int(r["l"] - r["k"] + 1)
            cs = coord[j, s, h, :min(nrow, row_cap)].astype(np.int64)
            assert np.all(chrom[j, s, h, :min(nrow, row_cap)] == 0), i
            for c in cs:
                assert np.array_equal(text[c:c + LEN], want), (i, c)
            if nrow <= row_cap:
                on_target |= bool(np.any(cs == pos[i]))
            else:
                on_target = True
        assert on_target, i
        found += 1
    assert found == len(pick)


def test_3g_modes_agree(huge):
    hp, ix = huge["hp"], huge["ix"]
    say("expanding")
    ix.expand()
    for k in (0, 15):
        ix.kmer_table(k)
        for loads in (hp.LOADS_ALL, hp.LOADS_OWN):
            say(f"expanded index, k-mer table k={k}, load mode {loads}")
            hp.set_load_mode(loads)
            hp.search1_launch(ix, huge["batch"], huge["out"], mode=hp.MODE_MISMATCH)
            hp.sync()
            got = (huge["out"].counts().reshape(-1, 2), huge["out"].hits())
            assert same_records(huge["compact"], got), (k, loads)
    hp.set_load_mode(hp.LOADS_AUTO)
    ix.kmer_table(0)


def test_3g_served_equal_alone(huge):
    """The serving configuration: the 3.1 Gbp index resident (expanded, k = 15
    table), Poisson arrivals at 200 requests/s of 1-50K reads of 50-250 bp
    with 1% substitutions, at most 4 launches in flight under a 2 GiB
    budget; every request's counts and records equal those of
    fmi_search1_batch on its reads alone."""
    hp, ix = huge["hp"], huge["ix"]
    from hpgfm import serve
    say("serving")
    ix.expand()
    ix.kmer_table(15)
    pool = serve.read_pool(huge["g"], 1_000_000, seed=5)
    srv = hp.Server(ix, budget_bytes=2 << 30, max_inflight=4, max_batch_reads=1 << 20, cap=CAP,
                    mode=hp.MODE_MISMATCH)
    try:
        res = serve.run(srv, pool, rate=200, duration_s=1.0, seed=200, keep=True)
        st = srv.stats()
    finally:
        srv.close()
    assert res["requests"] > 100 and st["requests_done"] == res["requests"]
    reads, offs = pool
    say(f"{res['requests']} requests served; each against the batch call alone")
    for start, size, c, recs in res["kept"]:
        o = offs[start:start + size + 1]
        batch = hp.ReadBatch(reads[int(o[0]):int(o[-1])], o - o[0])
        try:
            c2, h2 = hp.search1_batch(ix, batch, cap=CAP, mode=hp.MODE_MISMATCH)
        finally:
            batch.free()
        assert np.array_equal(c, c2)
        alone = h2[np.arange(CAP)[None, None, :] < np.minimum(c2, CAP)[:, :, None]]
        assert fo.hits_equal(recs, alone)
    ix.kmer_table(0)
