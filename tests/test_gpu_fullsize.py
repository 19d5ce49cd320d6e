"""The 1-mismatch search at the size its throughput is quoted on: a 64 Mbp
synthetic reference (41% GC, 20% mutated copies of 300 bp - 6 kbp repeat
families at 2-15% divergence) and 1M reads of 150 bp from both strands with
1% substitutions.

The oracle cannot search the whole batch in a test's time, so the whole
batch is held to properties that do not depend on its size, and a sample of
it to the oracle:

- the compact index, the expanded index (single-row intervals continued on
  the text) and the expanded index with the k = 15 interval table, under both
  load modes, give the same counts and the same records for every read and
  strand;
- every read drawn without a substitution has an exact record on the strand
  it was drawn from, and that record's interval holds the row of the
  position it was drawn at (ISA of the position, or of its mirror in the
  reversed text for a forward-direction record);
- 2000 of the reads give the oracle's BWSearch1CPU records, mismatch-only and
  edit mode, bit for bit (search/search.c:1876-2138);
- the inexact seed mapping of 200K of the reads finds every 18-base window
  with at most one substitution within 10 bp of where it was cut."""
import os

import numpy as np
import pytest

import fmoracle as fo
from hpgfm import synth

pytestmark = pytest.mark.gpu

NBP = 64_000_000
NREADS = 1_000_000
LEN = 150
CAP = 16


def say(msg):
    print(f"[fullsize] {msg}", flush=True)


@pytest.fixture(scope="module")
def big(gpu_engine):
    hp = gpu_engine
    say(f"genome {NBP} bp, {NREADS} x {LEN} bp reads")
    g, _, _ = synth.genome(NBP, seed=1, gc=0.41, repeat_frac=0.2, fam_len=(300, 6000),
                           divergence=(0.02, 0.15))
    reads, offs, truth = synth.reads(g, NREADS, length=LEN, seed=2, sub_rate=0.01)
    ix = hp.Index.build_gpu(g, 16)
    batch = hp.ReadBatch(reads, offs)
    out = hp.Search1Output(NREADS, CAP)
    state = dict(hp=hp, g=g, reads=reads, offs=offs, truth=truth, ix=ix, batch=batch, out=out)
    yield state
    hp.set_load_mode(hp.LOADS_AUTO)
    out.free()
    batch.free()
    ix.close()


def run(st, mode=None):
    hp = st["hp"]
    hp.search1_launch(st["ix"], st["batch"], st["out"],
                      mode=hp.MODE_MISMATCH if mode is None else mode)
    hp.sync()
    return st["out"].counts().reshape(-1, 2), st["out"].hits()


def same_records(a, b):
    """Counts equal and, where both hold records, hits_equal's fields equal."""
    ca, ha = a
    cb, hb = b
    if not np.array_equal(ca, cb):
        return False
    cap = ha.shape[2]
    m = np.arange(cap)[None, None, :] < np.minimum(ca, cap)[:, :, None]
    x, y = ha[m], hb[m]
    for key in ("k", "l", "start", "pos", "end", "dir", "num_mismatches", "strand"):
        if not np.array_equal(x[key], y[key]):
            return False
    e = x["num_mismatches"] > 0
    return all(np.array_equal(x[key][e], y[key][e]) for key in ("err_kind", "err_base", "err_pos"))


@pytest.fixture(scope="module")
def compact(big):
    say("compact index, mismatch mode")
    return run(big)


def test_fullsize_modes_agree(big, compact):
    hp = big["hp"]
    big["ix"].expand()
    for k in (0, 15):
        big["ix"].kmer_table(k)
        for loads in (hp.LOADS_ALL, hp.LOADS_OWN):
            say(f"expanded index, k-mer table k={k}, load mode {loads}")
            hp.set_load_mode(loads)
            assert same_records(compact, run(big)), (k, loads)
    hp.set_load_mode(hp.LOADS_AUTO)


def test_fullsize_planted_positions(big, compact):
    """Error-free reads are found where they were drawn."""
    say("planted positions")
    g, reads, truth = big["g"], big["reads"], big["truth"]
    pos = truth["pos"].astype(np.int64)
    strand = truth["strand"].astype(np.int64)
    mat = reads.reshape(NREADS, LEN)
    clean = np.zeros(NREADS, bool)
    for lo in range(0, NREADS, 100_000):
        hi = min(lo + 100_000, NREADS)
        src = g[pos[lo:hi, None] + np.arange(LEN)[None, :]]
        rc = strand[lo:hi] == 1
        src[rc] = (3 - src[rc])[:, ::-1]
        clean[lo:hi] = np.all(src == mat[lo:hi], axis=1)
    counts, hits = compact
    idx = np.nonzero(clean)[0]
    c = counts[idx, strand[idx]]
    assert len(idx) > 150_000 and np.all(c >= 1)
    idx = idx[c <= CAP]                     # every record of the strand is held
    assert len(idx) > 0.9 * clean.sum()
    rec = hits[idx, strand[idx], :]                          # [m, CAP]
    valid = np.arange(CAP)[None, :] < counts[idx, strand[idx]][:, None]
    exact = valid & (rec["num_mismatches"] == 0)
    assert np.all(exact.any(axis=1))
    big["ix"].expand()
    _, isa0, _ = big["ix"].full_arrays(0)
    _, isa1, _ = big["ix"].full_arrays(1)
    n1 = NBP + 1
    lc = rec["end"].astype(np.int64) - rec["start"].astype(np.int64) + 1
    p = pos[idx][:, None]
    # BWT of the text: SA[row] = pos; of the reversed text: SA[row] = n1 - lc - 1 - pos
    direction = np.where(strand[idx][:, None] == 1, rec["dir"] == 0, rec["dir"] == 1) & exact
    mirror = np.clip(n1 - lc - 1 - p, 0, NBP)
    row = np.where(direction, isa1[np.where(direction, mirror, 0)],
                   isa0[np.where(direction, 0, p)]).astype(np.int64)
    found = exact & (rec["k"] <= row) & (row <= rec["l"])
    assert np.all(found.any(axis=1))


def test_fullsize_sample_vs_oracle(big, compact):
    hp, ix = big["hp"], big["ix"]
    say("oracle sample")
    ix.expand()
    o = fo.Oracle(bwts=(*ix.bwt(0), *ix.bwt(1)), ratio=0)
    sa0, isa0, _ = ix.full_arrays(0)
    sa1, isa1, _ = ix.full_arrays(1)
    o.set_samples(ix.sa_ratio, sa0, isa0, sa1, isa1)
    del sa0, isa0, sa1, isa1
    rng = np.random.default_rng(3)
    pick = np.sort(rng.choice(NREADS, 2000, replace=False))
    rs = big["reads"].reshape(NREADS, LEN)[pick].reshape(-1)
    os_ = (np.arange(len(pick) + 1) * LEN).astype(np.uint64)
    threads = max(1, min(16, len(os.sched_getaffinity(0))))
    counts, hits = compact
    _, oc, oh = o.search1_batch(rs, os_, cap=CAP, mismatch_only=True, threads=threads)
    assert same_records((counts[pick], hits[pick]),
                        (oc.reshape(-1, 2), oh.reshape(-1, 2, CAP)))
    sub = hp.ReadBatch(rs, os_)
    try:
        ge = hp.search1_batch(ix, sub, cap=CAP, mode=hp.MODE_EDIT)
    finally:
        sub.free()
    _, oc, oh = o.search1_batch(rs, os_, cap=CAP, mismatch_only=False, threads=threads)
    assert same_records(ge, (oc.reshape(-1, 2), oh.reshape(-1, 2, CAP)))


def test_fullsize_inexact_seeds_planted(big):
    """Inexact seeds (bwt_map_inexact_seeds_seq, bwt.c:3262-3302) of the first
    200K reads, 18-base windows: every window with at most one substitution
    has, on its read's strand, a region within 10 bp of the 1-based start it
    was cut from (the one contig starts at coordinate 0), and an error-free
    window exactly one, unless its list overflowed (status set)."""
    hp, g, ix = big["hp"], big["g"], big["ix"]
    say("inexact seeds")
    n, ss_, ms, cap = 200_000, 18, 16, 16
    reads, offs = big["reads"][:n * LEN], big["offs"][:n + 1]
    pos = big["truth"]["pos"][:n].astype(np.int64)
    strand = big["truth"]["strand"][:n].astype(np.int64)
    batch = hp.ReadBatch(reads, offs)
    try:
        res = hp.map_inexact_seeds_batch(ix, batch, ss_, ms, region_cap=cap)
    finally:
        batch.free()
    nw = LEN // ss_                                       # 8 windows; the 6-base rest < 16
    assert np.array_equal(np.diff(res.window_base), np.full(n, nw))
    src = g[pos[:, None] + np.arange(LEN)[None, :]]
    rc = strand == 1
    src[rc] = (3 - src[rc])[:, ::-1]
    diff = np.concatenate([np.zeros((n, 1), np.int32),
                           np.cumsum(src != reads.reshape(n, LEN), axis=1, dtype=np.int32)], axis=1)
    st = np.arange(nw) * ss_
    errs = (diff[:, st + ss_] - diff[:, st]).reshape(-1)            # [n * nw]
    t0 = np.where(strand[:, None] == 0, pos[:, None] + st[None, :],
                  pos[:, None] + (LEN - st[None, :] - ss_)).reshape(-1)
    s = np.repeat(strand, nw)
    w = np.arange(n * nw)
    regs = res.regs.reshape(n * nw, 2, cap)[w, s]
    valid = np.arange(cap)[None, :] < np.minimum(res.nreg.reshape(n * nw, 2)[w, s], cap)[:, None]
    near = valid & (regs["chrom_idx"] == 1) & \
        (np.abs(regs["start"].astype(np.int64) - (t0 + 1)[:, None]) < 10)
    ok = res.status[w, s] == 0
    assert ok.mean() > 0.8
    one = ok & (errs <= 1)
    assert one.sum() > 0.6 * n * nw
    assert np.all(near[one].any(axis=1))
    zero = ok & (errs == 0)
    assert np.all(near[zero].sum(axis=1) == 1)
