"""Read-level 1-error mapping of __bwt_map_inexact_read (bwt.c:1890-2236):
only the error-free records of the vector-driven 1-error search become
alignments (bwt.c:2038), the running count of l - k over them is held to
filter_read_mappings (bwt.c:2145-2156), each SA row inside a contig gives an
alignment at its 0-based coordinate (bwt.c:2158-2166), and each strand's
alignments are deduplicated within 10 bp and handed over from last to first
with primary / secondary flags (bwt.c:2190-2233).

CPU: the oracle's restatement against a brute-force computation from the
text alone -- occurrences of the read and of its reverse complement found
by scanning, ordered as the suffix array orders them (suffixes of the text,
and of the reversed text for the - strand), the contig rule, the dedup rule
and the flags -- on a genome with tandem repeats (occurrences 1-7 bp apart),
far-apart copies, copies across contig ends, short reads, reads with N and
filters that are exceeded on either strand.
GPU: fmi_map_reads_batch against the oracle, record for record."""
import numpy as np
import pytest

import fmoracle as fo
from hpgfm import synth

COMP = np.array([3, 2, 1, 0, 4], np.uint8)


def planted_genome(seed=7):
    rng = np.random.default_rng(seed)
    g = rng.integers(0, 4, 24_000).astype(np.uint8)
    unit7 = rng.integers(0, 4, 7).astype(np.uint8)
    g[3000:3210] = np.tile(unit7, 30)                    # period 7: occurrences 7 bp apart
    g[6000:6060] = 0                                     # A-run: occurrences 1 bp apart
    seg = rng.integers(0, 4, 60).astype(np.uint8)
    for p in (9000, 9500, 15000):                        # far-apart copies
        g[p:p + 60] = seg
    g[12000:12060] = COMP[seg][::-1]                     # a reverse-complement copy
    g[7995:8055] = seg                                   # spans the contig end at 8000
    near = rng.integers(0, 4, 40).astype(np.uint8)
    g[18000:18040] = near
    g[18006:18046] = near                                # overlapping copies 6 bp apart
    return g


def brute_force(text, W, filt, koff, kstart):
    """(status, [(position, chromosome, strand, secondary)]) from the text."""
    n = len(text)
    L = len(W)
    if L < 5:
        return 2, []
    tb = text.tobytes()
    rb = text[::-1].tobytes()
    out, status, tot = [], 0, 0
    if np.any(W > 3):
        return 0, []                      # an N never occurs in the index
    for strand in (0, 1):
        pat = W if strand == 0 else COMP[W][::-1]
        pb = pat.tobytes()
        occ = [p for p in range(n - L + 1) if tb[p:p + L] == pb]
        if not occ:
            continue
        tot += len(occ) - 1                   # l - k of the strand's exact record
        if tot > filt:
            status |= 1
            break
        if strand == 0:
            occ.sort(key=lambda p: tb[p:])                    # rows of the backward index
        else:
            occ.sort(key=lambda p: rb[n - p - L:])            # rows of the forward index
        tmp = []
        for p in occ:
            idx = int(np.searchsorted(koff, p, side="right"))   # upper bound
            if p + L <= koff[idx]:
                tmp.append((int(kstart[idx - 1] + p - koff[idx - 1]), idx - 1))
        keep = []
        for i, (pos, c) in enumerate(tmp):
            if not any(c == c2 and abs(pos - p2) < 10 for p2, c2 in tmp[:i]):
                keep.append((pos, c))
        for pos, c in reversed(keep):
            out.append((pos, c, strand, int(len(out) > 0)))
    return status, out


@pytest.fixture(scope="module")
def planted():
    g = planted_genome()
    lens = np.array([8000, 9000, 7000], np.uint64)
    kstart = np.array([0, 100, 0], np.uint64)
    koff = np.concatenate([[0], np.cumsum(lens)]).astype(np.uint64)
    o = fo.Oracle(g, ratio=8)
    rng = np.random.default_rng(8)
    reads = [g[3000:3030], g[3010:3035], g[6000:6020], g[6005:6045], g[9000:9060],
             g[9010:9050], g[12000:12060], COMP[g[9000:9060]][::-1], g[7990:8030],
             g[18000:18040], g[18003:18043], g[100:104], g[200:203], g[300:305]]
    w = g[5000:5050].copy()
    w[20] = 4
    reads.append(w)
    for _ in range(60):
        p = int(rng.integers(0, len(g) - 80))
        L = int(rng.integers(5, 80))
        r = g[p:p + L].copy()
        if rng.random() < 0.5:
            r = COMP[r][::-1]
        reads.append(r)
    return g, o, koff, kstart, reads


@pytest.mark.parametrize("filt", [1, 3, 10, 200])
def test_oracle_map_read_vs_brute_force(planted, filt):
    g, o, koff, kstart, reads = planted
    seen_exceeded = seen_dedup = 0
    for i, W in enumerate(reads):
        st, recs, total = o.map_read(W, filt, koff, kstart, read=i)
        want_st, want = brute_force(g, W, filt, koff, kstart)
        assert st == want_st, (i, filt)
        assert total == len(want) == len(recs), (i, filt)
        got = [(int(r["position"]), int(r["chromosome"]), int(r["strand"]), int(r["secondary"]))
               for r in recs]
        assert got == want, (i, filt)
        assert np.all(recs["cigar_len"] == len(W)) and np.all(recs["read"] == i)
        seen_exceeded += st & 1
        seen_dedup += int(len(W) >= 5 and len(want) < sum(
            (g[p:p + len(W)].tobytes() == W.tobytes()) for p in range(len(g) - len(W) + 1)))
    if filt == 3:
        assert seen_exceeded > 0
    if filt == 200:
        assert seen_dedup > 0


def test_oracle_map_read_strand_order_and_flags(planted):
    """A read with copies on both strands: + strand alignments first, in the
    reverse of the row order; only the very first is primary."""
    g, o, koff, kstart, _ = planted
    W = g[9000:9060]
    st, recs, _ = o.map_read(W, 100, koff, kstart)
    assert st == 0
    assert list(recs["strand"]) == sorted(recs["strand"])
    assert recs["secondary"][0] == 0 and np.all(recs["secondary"][1:] == 1)
    assert {int(p) for p in recs["position"][recs["strand"] == 1]} == {12000 - 8000 + 100}


@pytest.mark.gpu
@pytest.mark.parametrize("expanded", [False, True])
def test_gpu_map_reads_vs_oracle(gpu_engine, planted, expanded):
    hp = gpu_engine
    g, o, koff, kstart, reads = planted
    lens = np.array([8000, 9000, 7000], np.uint64)
    ix = hp.Index.build(g, 8)
    ix.set_karyotype(kstart, kstart + lens - 1)
    ix.to_device()
    if expanded:
        ix.expand()
        ix.kmer_table(6)
    g2, _, _ = synth.genome(24_000, seed=9, repeat_frac=0.3, fam_len=(200, 1500))
    more, moffs, _ = synth.reads(g, 2000, seed=10, sub_rate=0.01, n_rate=0.002,
                                 lengths=np.random.default_rng(11).integers(3, 150, 2000))
    all_reads = reads + [more[int(moffs[i]):int(moffs[i + 1])] for i in range(2000)]
    offs = np.concatenate([[0], np.cumsum([len(r) for r in all_reads])]).astype(np.uint64)
    flat = np.concatenate(all_reads).astype(np.uint8)
    batch = hp.ReadBatch(flat, offs)
    try:
        for filt, cap in ((3, 4), (50, 16), (500, 64)):
            st, cnt, recs = hp.map_reads_batch(ix, batch, filt, cap)
            for i, W in enumerate(all_reads):
                ost, orecs, ototal = o.map_read(W, filt, koff, kstart, read=i, cap=cap)
                assert int(st[i]) == ost, (filt, i)
                assert int(cnt[i]) == ototal, (filt, i)
                m = min(ototal, cap)
                got = recs[i, :m]
                for key in ("position", "chromosome", "read", "strand", "secondary", "cigar_len"):
                    assert np.array_equal(got[key], orecs[key]), (filt, i, key)
    finally:
        batch.free()
        ix.close()
