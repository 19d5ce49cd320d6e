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
compact index, which walks LF from the sampled ISA.  On the same index: the
exact seed regions of planted seeds, and the serving loop's requests under
Poisson load against the batch call alone."""
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
    ix = hp.Index.build_gpu(synth.replace_n(g, seed=1), 16)
    batch = hp.ReadBatch(reads, offs)
    out = hp.Search1Output(NREADS, CAP)
    hp.search1_launch(ix, batch, out, mode=hp.MODE_MISMATCH)
    hp.sync()
    compact = (out.counts().reshape(-1, 2), out.hits())
    state = dict(hp=hp, g=g, reads=reads, truth=truth, ix=ix, batch=batch, out=out,
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
    """Exact seeds (bwt_map_exact_seed, bwt.c:1200-1294) of 100K reads of 250
    bp, 18 bases at stride 5, on the compact index: every seed without a
    substitution or an N yields, on its read's strand, the region of the
    position it was cut from (1-based, one contig from coordinate 0), unless
    that strand passed the 1000-row filter, in which case the strand has no
    region for the seed at all."""
    hp, g, ix = huge["hp"], huge["g"], huge["ix"]
    say("exact seeds")
    n, L, sl, st = 100_000, 250, 18, 5
    reads, offs, truth = synth.reads(g, n, length=L, seed=4, sub_rate=0.01)
    batch = hp.ReadBatch(reads, offs)
    try:
        nseeds, nrows, regs, _ = hp.seed_map_batch(ix, batch, sl, st, 1000)
    finally:
        batch.free()
    assert len(regs) == nrows
    pos = truth["pos"].astype(np.int64)
    strand = truth["strand"].astype(np.int64)
    src = g[pos[:, None] + np.arange(L)[None, :]]
    rc = strand == 1
    src[rc] = (3 - src[rc])[:, ::-1]
    mat = reads.reshape(n, L)
    bad = np.concatenate([np.zeros((n, 1), np.int32),
                          np.cumsum((src != mat) | (mat > 3), axis=1, dtype=np.int32)], axis=1)
    s0 = np.arange(0, L - sl + 1, st)
    clean = (bad[:, s0 + sl] - bad[:, s0]) == 0                     # [n, seeds]
    ri, q = np.nonzero(clean)
    off = s0[q]
    tpos = np.where(strand[ri] == 0, pos[ri] + off, pos[ri] + L - off - sl)
    assert int(tpos.max()) >= 1 << 31

    def key(read, seed, s, start):
        return ((read.astype(np.uint64) << np.uint64(39)) | (seed.astype(np.uint64) << np.uint64(33))
                | (s.astype(np.uint64) << np.uint64(32)) | start.astype(np.uint64))

    v = regs[regs["valid"] == 1]
    have = key(v["read"], v["seed_id"], v["strand"], v["start"])
    want = key(ri, q, strand[ri], tpos + 1)
    hit = np.isin(want, have)
    groups = np.unique(have >> np.uint64(32))
    # a missing planted region only where the seed's strand has no region at all
    assert not np.any(np.isin(want[~hit] >> np.uint64(32), groups))
    assert hit.mean() > 0.9, hit.mean()
    ends = v["end"][np.isin(have, want)] - v["start"][np.isin(have, want)]
    assert np.all(ends == sl - 1)


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
