"""Mixed error budgets at full size: a 256 Mbp synthetic reference (20% repeat
families at 2-15% divergence) and 4M reads of log-normal length (median 120,
clipped to 50-250 bp) with 1% substitutions and 0.5% N bases, split 50/35/15%
into budget 0 (exact search of both strands, bwt.c:1233-1237), budget 1
(BWSearch1CPU, edit mode) and budget 2 (BWSearchCPU with fragsize len / 3).

Every group's whole output is size-independent under the index's form: the
compact index, the expanded index and the expanded index with the k = 15
interval table (the 1-error search under both load modes) give the same
exact intervals, the same 1-error counts and records, and the same k-error
final lists, anchor lists, counts and overflow flags for every read."""
import numpy as np
import pytest

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
    reads, offs, _ = synth.reads(g, NREADS, seed=4, sub_rate=0.01, n_rate=0.005, lengths=lens)
    budget = rng.choice(3, size=NREADS, p=[0.5, 0.35, 0.15])
    ix = hp.Index.build_gpu(g, 16)
    batches = []
    for b in range(3):
        rs, os_ = subset(reads, offs, np.nonzero(budget == b)[0])
        batches.append((hp.ReadBatch(rs, os_), os_))
    fs2 = np.maximum(np.diff(batches[2][1]).astype(np.int64) // 3, 1).astype(np.int16)
    yield dict(hp=hp, ix=ix, batches=[x[0] for x in batches], fs2=fs2)
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


def test_mixed_budgets_index_forms_agree(mixed):
    hp, ix = mixed["hp"], mixed["ix"]
    say("compact index")
    base = run_groups(mixed)
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
