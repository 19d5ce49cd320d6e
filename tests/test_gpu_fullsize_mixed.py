"""Mixed error budgets at full size: a 256 Mbp synthetic reference (20% repeat
families at 2-15% divergence) and 4M reads of log-normal length (median 120,
clipped to 50-250 bp) with 1% substitutions and 0.5% N bases, split 50/35/15%
into budget 0 (exact search of both strands, bwt.c:1233-1237), budget 1
(BWSearch1CPU, edit mode) and budget 2 (BWSearchCPU with fragsize len / 3).

Every group's whole output is size-independent under the index's form: the
compact index, the expanded index and the expanded index with the k = 15
interval table (the 1-error search under both load modes) give the same
exact intervals, the same 1-error counts and records, and the same k-error
final lists, anchor lists, counts and overflow flags for every read.

Against the oracle (the C restatement of the reference's search on the
reference's own occurrence layout, built from this index's BWTs and samples):
2000 reads of every group, N reads among them, record for record -- the
exact searches of both strands with their anchors (search/search.h:9-65,
bwt.c:1233-1237), BWSearch1CPU in edit mode (search/search.c:1876-1933) and
BWSearchCPU with fragsize len / 3 at list capacities 32 and 128
(search/search.c:1171-1405).  And against the genome itself: every located
row of an error-free budget-0 read's exact interval holds the read there (or
its reverse complement), and the position the read was drawn from is among
them."""
import os

import numpy as np
import pytest

import fmoracle as fo
from hpgfm import synth
from test_gpu_fullsize import same_records

pytestmark = pytest.mark.gpu

NBP = 256_000_000
NREADS = 4_000_000
CAP = 16
LIST_CAP = 32
ANCHOR_CAP = 16


def say(msg):
    print(f"[fullsize mixed] {msg}", flush=True)


def subset(reads, offs, idx):
    o64 = offs.astype(np.int64)
    ln = o64[1:][idx] - o64[:-1][idx]
    so = np.concatenate([[0], np.cumsum(ln)]).astype(np.uint64)
    gather = np.repeat(o64[:-1][idx] - so[:-1].astype(np.int64), ln) + np.arange(int(so[-1]))
    return reads[gather], so


@pytest.fixture(scope="module")
def mixed(gpu_engine):
    hp = gpu_engine
    say(f"genome {NBP} bp, {NREADS} reads of 50-250 bp")
    g, _, _ = synth.genome(NBP, seed=1, gc=0.41, repeat_frac=0.2, fam_len=(300, 6000),
                           divergence=(0.02, 0.15))
    rng = np.random.default_rng(3)
    lens = synth.lognormal_lengths(rng, NREADS, median=120, lo=50, hi=250)
    reads, offs, truth = synth.reads(g, NREADS, seed=4, sub_rate=0.01, n_rate=0.005, lengths=lens)
    budget = rng.choice(3, size=NREADS, p=[0.5, 0.35, 0.15])
    ix = hp.Index.build_gpu(g, 16)
    batches, host, members = [], [], []
    for b in range(3):
        idx = np.nonzero(budget == b)[0]
        rs, os_ = subset(reads, offs, idx)
        batches.append((hp.ReadBatch(rs, os_), os_))
        host.append((rs, os_))
        members.append(idx)
    fs2 = np.maximum(np.diff(batches[2][1]).astype(np.int64) // 3, 1).astype(np.int16)
    yield dict(hp=hp, ix=ix, batches=[x[0] for x in batches], fs2=fs2, g=g, host=host,
               members=members, truth=truth)
    hp.set_load_mode(hp.LOADS_AUTO)
    for b, _ in batches:
        b.free()
    ix.close()


def run_groups(st):
    hp, ix = st["hp"], st["ix"]
    b0, b1, b2 = st["batches"]
    ex = [hp.exact_batch(ix, b0, hp.VIEW_BACKWARD, 0)[0],
          hp.exact_batch(ix, b0, hp.VIEW_BACKWARD_REV, 1)[0]]
    c1, h1 = hp.search1_batch(ix, b1, cap=CAP, mode=hp.MODE_EDIT)
    rk = hp.search_k_batch(ix, b2, st["fs2"], LIST_CAP, ANCHOR_CAP)
    return ex, (c1, h1), rk


def same_k_lists(na, a, nb, b, cap):
    if not np.array_equal(na, nb):
        return False
    a = a.reshape(-1, cap)
    b = b.reshape(-1, cap)
    m = np.arange(cap)[None, :] < np.minimum(na, cap)[:, None]
    x, y = a[m], b[m]
    for key in ("k", "l", "start", "pos", "end", "dir", "num_mismatches"):
        if not np.array_equal(x[key], y[key]):
            return False
    e = np.arange(8)[None, :] < np.minimum(x["num_mismatches"], 8)[:, None]
    return all(np.array_equal(x[key][e], y[key][e]) for key in ("err_kind", "err_base", "err_pos"))


def same_groups(p, q):
    (ea, sa, ka), (eb, sb, kb) = p, q
    for d in (0, 1):
        for key in ("k", "l", "pos"):
            if not np.array_equal(ea[d][key], eb[d][key]):
                return "exact"
    if not same_records(sa, sb):
        return "search1"
    if not np.array_equal(ka.full, kb.full):
        return "search_k overflow flags"
    if not same_k_lists(ka.nfin, ka.fin, kb.nfin, kb.fin, LIST_CAP):
        return "search_k final lists"
    if not same_k_lists(ka.nanc, ka.anc, kb.nanc, kb.anc, ANCHOR_CAP):
        return "search_k anchors"
    return None


@pytest.fixture(scope="module")
def base(mixed):
    say("compact index")
    return run_groups(mixed)


def test_mixed_budgets_index_forms_agree(mixed, base):
    hp, ix = mixed["hp"], mixed["ix"]
    assert int(base[2].nfin.sum()) > 0 and int(base[1][0].sum()) > 0
    say("expanded index")
    ix.expand()
    assert same_groups(base, run_groups(mixed)) is None
    ix.kmer_table(15)
    for loads in (hp.LOADS_ALL, hp.LOADS_OWN):
        say(f"expanded index, k-mer table k=15, load mode {loads}")
        hp.set_load_mode(loads)
        assert same_groups(base, run_groups(mixed)) is None, loads
    hp.set_load_mode(hp.LOADS_AUTO)
    ix.kmer_table(0)


def pick_sample(rng, reads, offs, n):
    """n reads of a group: three quarters at random, a quarter among the
    reads that hold an N."""
    nr = len(offs) - 1
    o64 = offs.astype(np.int64)
    has_n = np.add.reduceat(reads > 3, o64[:-1]) > 0 if nr else np.zeros(0, bool)
    with_n = np.nonzero(has_n)[0]
    a = rng.choice(nr, n - n // 4, replace=False)
    b = rng.choice(with_n, min(n // 4, len(with_n)), replace=False)
    return np.unique(np.concatenate([a, b]))


def test_mixed_sample_vs_oracle(mixed, base):
    hp, ix = mixed["hp"], mixed["ix"]
    say("oracle from the index's BWTs and samples")
    o = fo.Oracle(bwts=(*ix.bwt(0), *ix.bwt(1)), ratio=0)
    o.set_sampled(ix.sa_ratio, *ix.samples(0), *ix.samples(1))
    threads = max(1, min(16, len(os.sched_getaffinity(0))))
    rng = np.random.default_rng(31)
    lsa = o.size_sa - 1
    (ex_plus, ex_minus), (c1, h1), rk = base
    # budget 0: the exact searches of both strands, with anchors
    rs, os_ = mixed["host"][0]
    pick = pick_sample(rng, rs, os_, 2000)
    sr, so = subset(rs, os_, pick)
    assert np.any(sr > 3)
    sub = hp.ReadBatch(sr, so)
    try:
        got = [hp.exact_batch(ix, sub, hp.VIEW_BACKWARD, 0, anchors=True),
               hp.exact_batch(ix, sub, hp.VIEW_BACKWARD_REV, 1, anchors=True)]
    finally:
        sub.free()
    for j, i in enumerate(pick):
        W = sr[int(so[j]):int(so[j + 1])]
        L = len(W)
        for s, (view, d, full) in enumerate(((fo.VIEW_BACKWARD, 0, ex_plus),
                                             (fo.VIEW_BACKWARD_REV, 1, ex_minus))):
            r, a = o.exact(view, d, W, 0, lsa, 0, 0 if d else L - 1, L - 1, anchor=True)
            for key in ("k", "l", "pos"):
                assert full[i][key] == r[key], (i, s, key)
                assert got[s][0][j][key] == r[key], (i, s, key)
            for key in ("k", "l", "start", "pos", "end", "dir"):
                assert got[s][1][j][key] == a[key], (i, s, "anchor", key)
    say("budget 0 against the oracle: ok")
    # budget 1: BWSearch1CPU, edit mode
    rs, os_ = mixed["host"][1]
    pick = pick_sample(rng, rs, os_, 2000)
    sr, so = subset(rs, os_, pick)
    _, oc, oh = o.search1_batch(sr, so, cap=CAP, mismatch_only=False, threads=threads)
    assert same_records((c1[pick], h1[pick]), (oc.reshape(-1, 2), oh.reshape(-1, 2, CAP)))
    say("budget 1 against the oracle: ok")
    # budget 2: BWSearchCPU with fragsize len / 3, list capacities 32 and 128
    rs, os_ = mixed["host"][2]
    pick = pick_sample(rng, rs, os_, 2000)
    sr, so = subset(rs, os_, pick)
    fs = mixed["fs2"][pick]
    _, nf, fin, na, anc, full = o.search_k_batch(sr, so, fs, LIST_CAP, ANCHOR_CAP, threads=threads)
    t = (2 * pick[:, None] + np.arange(2)[None, :]).reshape(-1)
    assert np.array_equal(rk.full[pick], full)
    assert same_k_lists(rk.nfin[t], rk.fin.reshape(-1, LIST_CAP)[t].reshape(-1), nf, fin, LIST_CAP)
    assert same_k_lists(rk.nanc[t], rk.anc.reshape(-1, ANCHOR_CAP)[t].reshape(-1), na, anc,
                        ANCHOR_CAP)
    sub = hp.ReadBatch(sr, so)
    try:
        g128 = hp.search_k_batch(ix, sub, fs, 128, ANCHOR_CAP)
    finally:
        sub.free()
    _, nf, fin, na, anc, full = o.search_k_batch(sr, so, fs, 128, ANCHOR_CAP, threads=threads)
    assert np.array_equal(g128.full, full)
    assert same_k_lists(g128.nfin, g128.fin, nf, fin, 128)
    assert same_k_lists(g128.nanc, g128.anc, na, anc, ANCHOR_CAP)
    assert int(nf.sum()) > 0
    say("budget 2 against the oracle (list capacities 32 and 128): ok")
    o.close()


def test_mixed_budget0_planted(mixed, base):
    """Error-free budget-0 reads: every located row of the exact interval on
    the read's strand holds the read in the genome (its reverse complement
    on the - strand: the forward search on backward_rev locates through
    backward, bwt.c:1233-1250), and the drawn position is one of them."""
    hp, ix, g = mixed["hp"], mixed["ix"], mixed["g"]
    say("budget 0 planted positions")
    rs, os_ = mixed["host"][0]
    idx = mixed["members"][0]
    pos = mixed["truth"]["pos"][idx].astype(np.int64)
    strand = mixed["truth"]["strand"][idx].astype(np.int64)
    o64 = os_.astype(np.int64)
    lens = np.diff(o64)
    sel = np.random.default_rng(32).choice(len(idx), 200_000, replace=False)
    sel.sort()
    clean = np.zeros(len(sel), bool)
    for j, i in enumerate(sel):
        W = rs[o64[i]:o64[i + 1]]
        src = g[pos[i]:pos[i] + lens[i]]
        if strand[i]:
            src = synth.revcomp(src)
        clean[j] = np.array_equal(src, W)
    sel = sel[clean]
    assert len(sel) > 20_000
    ex = base[0]
    ks = np.where(strand[sel] == 0, ex[0]["k"][sel], ex[1]["k"][sel]).astype(np.int64)
    ls = np.where(strand[sel] == 0, ex[0]["l"][sel], ex[1]["l"][sel]).astype(np.int64)
    assert np.all(ks <= ls)
    cap = 16
    rows, nrows = hp.locate_batch(ix, ks, ls, cap, forward=0)
    rows = rows.reshape(len(sel), cap).astype(np.int64)
    for j, i in enumerate(sel):
        L = int(lens[i])
        W = rs[o64[i]:o64[i + 1]]
        want = W if strand[i] == 0 else synth.revcomp(W)
        q = rows[j, :nrows[j]]
        assert nrows[j] == min(ls[j] - ks[j] + 1, cap)
        for p in q:
            assert np.array_equal(g[p:p + L], want), (i, p)
        if ls[j] - ks[j] + 1 <= cap:
            assert pos[i] in q, i
