"""Pins the oracle (oracle/fmoracle.c) before it is trusted as the checker.

- against the golden vectors written by the reference's own compiled search
  core (tests/golden/make_golden.py);
- against the reference core itself where oracle/_ref is built;
- against brute force: suffix ranges by sorting the suffixes in Python, and
  the Hamming-1 projection of BWSearch1CPU by scanning the text.
"""
import hashlib

import numpy as np
import pytest

import fmoracle as fo
from conftest import split_golden_hits
from hpgfm import synth


def arrays_sha(o):
    h = hashlib.sha256()
    for fwd in (0, 1):
        a = o.arrays(fwd)
        for key in ("C", "C1", "desp", "count", "S", "R"):
            h.update(np.ascontiguousarray(a[key]).tobytes())
    return h.hexdigest()


@pytest.fixture(scope="module", params=["golden_small", "golden_repeats"])
def golden(request):
    g = request.getfixturevalue(request.param)
    o = fo.Oracle(g["genome"], ratio=int(g["ratio"]))
    return g, o


def test_golden_index_arrays(golden):
    g, o = golden
    assert arrays_sha(o) == str(g["arrays_sha"])
    b = o.arrays(0)
    sa = fo.sa_naive(g["genome"])
    bwt, dollar = fo.bwt_from_sa(g["genome"], sa)
    assert hashlib.sha256(bwt.tobytes()).hexdigest() == str(g["bwt_sha"])
    assert dollar == int(g["bwt_dollar"]) == b["dollar"]


def test_golden_search1(golden):
    g, o = golden
    want = split_golden_hits(g)
    reads, offs = g["reads"], g["offsets"]
    cap = int(g["cap"])
    for i in range(len(offs) - 1):
        W = reads[offs[i]:offs[i + 1]]
        for s in (0, 1):
            n, h, an = o.search1(s, W, cap=cap, anchors=True)
            assert n == int(g["s1_counts"][i, s]), (i, s)
            assert fo.hits_equal(h, want[(i, s)]), (i, s)
            assert fo.hits_equal(an, g["s1_anchor"][i, s]), (i, s)


def test_golden_exact(golden):
    g, o = golden
    reads, offs = g["reads"], g["offsets"]
    lsa = o.size_sa - 1
    for view in range(4):
        for d in (0, 1):
            rec = g[f"ex_v{view}_d{d}"]
            anc = g[f"ex_anchor_v{view}_d{d}"]
            for i in range(len(offs) - 1):
                W = reads[offs[i]:offs[i + 1]]
                L = len(W)
                r, a = o.exact(view, d, W, 0, lsa, 0, L - 1 if d == 0 else 0, L - 1, anchor=True)
                assert (r["k"], r["l"], r["pos"]) == (rec[i]["k"], rec[i]["l"], rec[i]["pos"])
                for key in ("k", "l", "start", "pos", "end", "dir"):
                    assert a[key] == anc[i][key], (view, d, i, key)


def test_golden_sa_isa(golden):
    g, o = golden
    for j, m in enumerate(g["sa_rows"]):
        m = int(m)
        assert o.get_sa(0, m) == int(g["sa_bwd"][j])
        assert o.get_sa(1, m) == int(g["sa_fwd"][j])
        assert o.get_isa(0, m) == int(g["isa_bwd"][j])
        assert o.get_isa(1, m) == int(g["isa_fwd"][j])


def test_set_samples_equals_lf_walk():
    """S/R sampled from full SA / ISA arrays (the brute-force suffix sort
    here) equal the ones the build-time LF walk writes."""
    g, _, _ = synth.genome(3000, seed=31, repeat_frac=0.3, fam_len=(100, 400))
    walk = fo.Oracle(g, ratio=8)
    o = fo.Oracle(g, ratio=0)
    full = []
    for t in (g, g[::-1]):
        sa = np.array(py_suffix_array(t), np.uint32)
        isa = np.empty_like(sa)
        isa[sa] = np.arange(len(sa), dtype=np.uint32)
        full += [sa, isa]
    o.set_samples(8, *full)
    for fwd in (0, 1):
        a, b = o.arrays(fwd), walk.arrays(fwd)
        assert np.array_equal(a["S"], b["S"]) and np.array_equal(a["R"], b["R"])
    for m in range(0, o.size_sa, 97):
        assert o.get_sa(0, m) == walk.get_sa(0, m) and o.get_isa(1, m) == walk.get_isa(1, m)
    with pytest.raises(ValueError):
        o.set_samples(8, full[0][:-8], *full[1:])


def test_exact_locate_batch_closed_form():
    """Exact search of both strands + get_SA: a read copied from strand s of
    a unique text at position p has one row on strand s, at p, and none on
    the other; random reads have none; the compiled reference agrees."""
    g, _, _ = synth.genome(200_000, seed=1)
    o = fo.Oracle(g, ratio=16)
    reads, offs, truth = synth.reads(g, 2000, length=100, seed=2, random_frac=0.1)
    _, rows, sums = o.exact_locate_batch(reads, offs, cap=8)
    for i in range(len(offs) - 1):
        if truth["is_random"][i]:
            assert rows[i].sum() == 0, i
            continue
        s = int(truth["strand"][i])
        assert rows[i, s] == 1 and int(sums[i, s]) == int(truth["pos"][i]), i
        assert rows[i, 1 - s] == 0, i
    if fo.ref_available():
        r = fo.RefCore(oracle=o)
        _, rr, rs = r.exact_locate_batch(reads, offs, cap=8)
        assert np.array_equal(rr, rows) and np.array_equal(rs, sums)
        r.close()


# ---------------------------------------------------------------- brute force

def py_suffix_array(text):
    s = bytes(int(x) + 1 for x in text) + b"\x00"   # '$' smallest
    return sorted(range(len(s)), key=lambda i: s[i:])


def sa_range(sa, text, pat):
    """Rows whose suffix starts with pat (the exact SA interval)."""
    t = list(int(x) for x in text)
    rows = [r for r, p in enumerate(sa) if t[p:p + len(pat)] == list(int(x) for x in pat)]
    return (rows[0], rows[-1]) if rows else None


def count_occ(text, pat):
    t = bytes(text.tolist())
    p = bytes(np.asarray(pat, dtype=np.uint8).tolist())
    if len(p) == 0:
        return len(t) + 1
    c, i = 0, t.find(p)
    while i >= 0:
        c += 1
        i = t.find(p, i + 1)
    return c


@pytest.mark.parametrize("seed", [1, 2, 3])
def test_exact_is_suffix_range(seed):
    rng = np.random.default_rng(seed)
    g = rng.integers(0, 4, size=600).astype(np.uint8)
    g[100:160] = g[300:360]                     # a repeat
    o = fo.Oracle(g, ratio=8)
    sa = py_suffix_array(g)
    sa_rev = py_suffix_array(g[::-1])
    lsa = o.size_sa - 1
    for _ in range(150):
        L = int(rng.integers(1, 12))
        p = int(rng.integers(0, len(g) - L))
        W = g[p:p + L].copy() if rng.random() < 0.7 else rng.integers(0, 4, L).astype(np.uint8)
        r = o.exact(fo.VIEW_BACKWARD, 0, W, 0, lsa, 0, L - 1, L - 1)
        want = sa_range(sa, g, W)
        if want is None:
            assert r["k"] > r["l"]
        else:
            assert (r["k"], r["l"], r["pos"]) == (want[0], want[1], -1)
        # forward search on the forward index = suffix range of reverse(W) in reverse(X)
        r = o.exact(fo.VIEW_FORWARD, 1, W, 0, lsa, 0, 0, L - 1)
        want = sa_range(sa_rev, g[::-1], W[::-1])
        if want is None:
            assert r["k"] > r["l"]
        else:
            assert (r["k"], r["l"], r["pos"]) == (want[0], want[1], L)
        # reverse complement: forward search on backward_rev (bwt.c:1233-1237)
        r = o.exact(fo.VIEW_BACKWARD_REV, 1, W, 0, lsa, 0, 0, L - 1)
        want = sa_range(sa, g, synth.revcomp(W))
        if want is None:
            assert r["k"] > r["l"]
        else:
            assert (r["k"], r["l"]) == want
    for m in range(len(g) + 1):
        assert o.get_sa(0, m) == sa[m]
        assert o.get_sa(1, m) == sa_rev[m]
        assert o.get_isa(0, sa[m]) == m


@pytest.mark.parametrize("seed", [4, 5])
def test_search1_mismatch_projection_is_hamming1(seed):
    """The {exact, MISMATCH} records of BWSearch1CPU are exactly the single
    substitutions of W (or of its reverse complement for strand 1) that
    occur in X, each once, with l - k + 1 = number of occurrences."""
    rng = np.random.default_rng(seed)
    g, _, _ = synth.genome(3000, seed=seed, repeat_frac=0.3, fam_len=(100, 400),
                           divergence=(0.0, 0.03), copies_per_family=6)
    o = fo.Oracle(g, ratio=8)
    reads, offs, _ = synth.reads(g, 120, length=24, seed=seed + 10, sub_rate=0.04)
    for i in range(len(offs) - 1):
        W = reads[offs[i]:offs[i + 1]]
        for s in (0, 1):
            n, hits = o.search1(s, W, cap=4096, mismatch_only=True)
            n_edit, hits_edit = o.search1(s, W, cap=4096, mismatch_only=False)
            keep = (hits_edit["num_mismatches"] == 0) | (hits_edit["err_kind"] == fo.MISMATCH)
            assert fo.hits_equal(hits, hits_edit[keep])      # projection keeps order
            got = {}
            for h in hits:
                key = (-1, -1) if h["num_mismatches"] == 0 else (int(h["err_pos"]), int(h["err_base"]))
                assert key not in got
                got[key] = int(h["l"] - h["k"] + 1)
            want = {}
            target = W if s == 0 else None
            c = count_occ(g, W if s == 0 else synth.revcomp(W))
            if c:
                want[(-1, -1)] = c
            for p in range(len(W)):
                for b in range(4):
                    if b == W[p]:
                        continue
                    V = W.copy()
                    V[p] = b
                    c = count_occ(g, V if s == 0 else synth.revcomp(V))
                    if c:
                        want[(p, b)] = c
            del target
            assert got == want, (i, s)


def test_n_policy():
    """An N (code 4) is a symbol after T that never occurs: the step gives
    (n + 1, n), so the exact search dies there and every base is a
    mismatch candidate at that position."""
    g, _, _ = synth.genome(2000, seed=9)
    o = fo.Oracle(g, ratio=8)
    k, l = o.lf(fo.VIEW_BACKWARD, 0, o.n, 4)
    assert (k, l) == (o.n + 1, o.n)
    W = g[500:530].copy()
    W[20] = 4
    n, hits = o.search1(0, W, cap=64, mismatch_only=True)
    assert n == 1 and hits[0]["err_kind"] == fo.MISMATCH and hits[0]["err_pos"] == 20
    assert hits[0]["err_base"] == g[520]


def test_comparator_catches_planted_mistakes(golden_repeats):
    """The record comparison used by every parity test fails on a dropped
    record, a swapped order, a wrong error position and a wrong bound."""
    want = split_golden_hits(golden_repeats)
    lst = next(v for v in want.values() if len(v) >= 3)
    assert fo.hits_equal(lst, lst.copy())
    assert not fo.hits_equal(lst[:-1], lst)
    sw = lst.copy()
    sw[[0, 1]] = sw[[1, 0]]
    assert not fo.hits_equal(sw, lst) or lst[0] == lst[1]
    bad = lst.copy()
    m = np.nonzero(bad["num_mismatches"] > 0)[0][0]
    bad["err_pos"][m] += 1
    assert not fo.hits_equal(bad, lst)
    bad = lst.copy()
    bad["l"][0] += 1
    assert not fo.hits_equal(bad, lst)


# ------------------------------------------------------ against the reference

ref_only = pytest.mark.skipif(not fo.ref_available(), reason="oracle/_ref not built")


@ref_only
@pytest.mark.parametrize("seed", [31, 32])
def test_oracle_matches_reference_core(seed):
    g, _, _ = synth.genome(30000, seed=seed, repeat_frac=0.3, fam_len=(300, 1500),
                           divergence=(0.0, 0.05))
    o = fo.Oracle(g, ratio=16)
    r = fo.RefCore(g, ratio=16)
    for fwd in (0, 1):
        a, b = o.arrays(fwd), r.arrays(fwd)
        for key in ("C", "C1", "desp", "count", "S", "R"):
            assert np.array_equal(a[key], b[key]), (fwd, key)
    reads, offs, _ = synth.reads(g, 300, seed=seed + 1, sub_rate=0.02, random_frac=0.1,
                                 lengths=np.random.default_rng(seed).integers(0, 90, 300))
    for i in range(len(offs) - 1):
        W = reads[offs[i]:offs[i + 1]]
        for s in (0, 1):
            n1, h1, a1 = o.search1(s, W, cap=2048, anchors=True)
            n2, h2, full, a2 = r.search1(s, W, cap=2048, anchors=True)
            assert n1 == n2 and not full
            assert fo.hits_equal(h1, h2)
            assert fo.hits_equal(a1, a2)
            # truncation: a results_list of capacity 2 keeps the first two.
            # (BWSearch1CPU drops its helpers' "full" return, search.c:1905
            # and :1930, so the flag only implies truncation.)
            n3, h3, full3 = r.search1(s, W, cap=2)
            assert fo.hits_equal(h3, h1[:2])
            assert (not full3) or n1 > 2
    r.close()
