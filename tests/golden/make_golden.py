"""Writes the golden fixtures of the FM-index search path.

Runs the reference's own search core, compiled from its sources into
oracle/_ref/libfmref_ref.so (`make -C oracle ref`, needs the reference
checkout), on a small seeded synthetic genome and read set, and stores
inputs and outputs as data:

  golden_small.npz
    genome            uint8 codes (64 kbp, 41% GC, 20% repeat families)
    reads, offsets    2000 reads of 20..120 bp, both strands, 2% substitutions,
                      10% random, no N (the reference leaves N undefined)
    s1_counts[i, s]   number of records BWSearch1CPU keeps (cap 512)
    s1_hits           those records, flattened in (read, strand) order
    s1_anchor[i, s]   the two anchors of each call
    ex_*              BWExactSearchBackward/Forward on the four views
    sa_rows, sa_bwd, sa_fwd, isa_bwd, isa_fwd   get_SA / get_ISA samples
    arrays_sha        SHA-256 of the reference index arrays (C, C1, desp,
                      count, S, R for both directions)
    s1v_counts/hits   BWSearch1VectorHelper over the interval vectors of each
                      whole read (bwt.c:1940-1950, 1994-2003), per strand
    s1w_bounds        one window [start, end] per read (a seed, bwt.c:1447)
    s1w_counts/hits   the same search over that window
    sk_*              BWSearchCPU (search.c:1171) per read and strand with
                      fragsize max(len // 3, 8), lists of sk_cap records and
                      anchor lists of 16: the final list, the anchors, the
                      return value (1: the anchor list filled up)
    wr_blob, wr_off   write_results' text lines (search/results.c:3-43) for
                      BWSearch1CPU's list (cap 512) with the contigs
                      wr_kstart / wr_kend named chr1.., the read's ACGT text:
                      the text of (read i, strand s) is
                      wr_blob[wr_off[2i+s]:wr_off[2i+s+1]]; only for reads
                      of 12..99 bp (wr_done): the reference's 6-byte mask
                      buffer (results.h:134) overflows for error positions
                      of 3 digits, and shorter reads print thousands of rows

  golden_repeats.npz
    the same fields for 400 reads of 0..60 bp on 8 kbp made of 40 mutated
    copies of one 200 bp unit (many records per read, all error kinds).

Usage: python tests/golden/make_golden.py   (from the repository root)
"""
import hashlib
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, os.path.join(ROOT, "oracle"))
sys.path.insert(0, os.path.join(ROOT, "hpg-libs_amd"))

import fmoracle as fo  # noqa: E402
from hpgfm import synth  # noqa: E402

OUT = os.path.join(ROOT, "tests", "golden", "golden_small.npz")
RATIO = 16
CAP = 512
KCAP = 300


def arrays_sha(arrs):
    h = hashlib.sha256()
    for fwd in (0, 1):
        a = arrs[fwd]
        for key in ("C", "C1", "desp", "count", "S", "R"):
            h.update(np.ascontiguousarray(a[key]).tobytes())
    return h.hexdigest()


def repeat_genome(seed):
    """8 kbp made of mutated copies of one 200 bp unit (dense repeats)."""
    rng = np.random.default_rng(seed)
    unit = synth.iid_bases(rng, 200)
    parts = []
    for _ in range(40):
        c = unit.copy()
        m = rng.random(200) < 0.01
        c[m] = (c[m] + rng.integers(1, 4, size=int(m.sum()))) % 4
        parts.append(c)
    return np.concatenate(parts).astype(np.uint8)


def main():
    g, _, _ = synth.genome(64_000, seed=11, repeat_frac=0.2, fam_len=(300, 2000), method="v1",
                           divergence=(0.0, 0.08))
    rng = np.random.default_rng(12)
    lengths = rng.integers(20, 121, size=2000)
    reads, offsets, _ = synth.reads(g, 2000, seed=13, sub_rate=0.02, random_frac=0.1,
                                    lengths=lengths)
    write(OUT, g, reads, offsets, rng)
    g2 = repeat_genome(21)
    rng2 = np.random.default_rng(22)
    lengths2 = rng2.integers(1, 61, size=400)
    lengths2[:3] = 0                      # empty reads are an edge the reference defines
    reads2, offsets2, _ = synth.reads(g2, 400, seed=23, sub_rate=0.02, random_frac=0.05,
                                      lengths=lengths2)
    write(OUT.replace("golden_small", "golden_repeats"), g2, reads2, offsets2, rng2)


def write(path, g, reads, offsets, rng):
    ref = fo.RefCore(g, ratio=RATIO)
    n = len(offsets) - 1
    counts = np.zeros((n, 2), np.uint32)
    anchors = np.zeros((n, 2, 2), fo.HIT_DTYPE)
    hits = []
    for i in range(n):
        W = reads[offsets[i]:offsets[i + 1]]
        for s in (0, 1):
            cnt, h, full, an = ref.search1(s, W, cap=CAP, anchors=True)
            assert not full, "raise CAP"
            counts[i, s] = cnt
            anchors[i, s] = an
            hits.append(h)
    s1_hits = np.concatenate(hits) if hits else np.zeros(0, fo.HIT_DTYPE)

    # BWSearch1VectorHelper: whole reads, then one window per read (own rng,
    # so the fields above do not move).  Windows under 3 bases are not run:
    # the reference reads outside its vectors there; they hold no records.
    wrng = np.random.default_rng(int(len(g)) + 7)
    bounds = np.zeros((n, 2), np.int16)
    v_counts = np.zeros((n, 2), np.uint32)
    w_counts = np.zeros((n, 2), np.uint32)
    v_hits, w_hits = [], []
    for i in range(n):
        W = reads[offsets[i]:offsets[i + 1]]
        L = len(W)
        if L >= 3:
            a = int(wrng.integers(0, L - 2))
            b = int(wrng.integers(a + 2, L))
        else:
            a, b = 0, L - 1
        bounds[i] = (a, b)
        for s in (0, 1):
            for (st, en, cnts, lst) in ((0, L - 1, v_counts, v_hits), (a, b, w_counts, w_hits)):
                if en - st < 2:
                    continue
                cnt, h, full = ref.search1_vector(s, W, st, en, cap=CAP)
                assert not full, "raise CAP"
                cnts[i, s] = cnt
                lst.append(h)
    s1v_hits = np.concatenate(v_hits) if v_hits else np.zeros(0, fo.HIT_DTYPE)
    s1w_hits = np.concatenate(w_hits) if w_hits else np.zeros(0, fo.HIT_DTYPE)

    # BWSearchCPU with fragsize max(len // 3, 8) (up to 2 errors on most
    # reads), lists of KCAP records and anchor lists of 16
    k_fs = np.zeros(n, np.int16)
    k_counts = np.zeros((n, 2), np.uint32)
    k_acounts = np.zeros((n, 2), np.uint32)
    k_full = np.zeros((n, 2), np.uint8)
    k_hits, k_anch = [], []
    for i in range(n):
        W = reads[offsets[i]:offsets[i + 1]]
        fs = max(len(W) // 3, 8)
        k_fs[i] = fs
        for s in (0, 1):
            full, fin, anc = ref.search_k(s, W, fs, KCAP, 16)
            k_counts[i, s] = len(fin)
            k_acounts[i, s] = len(anc)
            k_full[i, s] = full
            k_hits.append(fin)
            k_anch.append(anc)
    sk_hits = np.concatenate(k_hits) if k_hits else np.zeros(0, fo.HITK_DTYPE)
    sk_anch = np.concatenate(k_anch) if k_anch else np.zeros(0, fo.HITK_DTYPE)

    # write_results text with a karyotype of 2-3 contigs
    G = len(g)
    if G >= 30000:
        klen = [G // 3, G // 3 + 4000, G - 2 * (G // 3) - 4000]
        kst = [0, 100, 0]
    else:
        klen = [G // 2, G - G // 2]
        kst = [0, 0]
    kstart = np.array(kst, np.uint64)
    kend = kstart + np.array(klen, np.uint64) - 1
    wr_parts = []
    wr_done = np.zeros(n, np.uint8)
    for i in range(n):
        W = reads[offsets[i]:offsets[i + 1]]
        ok = 12 <= len(W) < 100
        wr_done[i] = ok
        mapping = "".join("ACGT"[int(x)] for x in W).encode()
        for s in (0, 1):
            wr_parts.append(ref.write_results(s, W, mapping, i, CAP, kstart, kend).encode()
                            if ok else b"")
    wr_off = np.concatenate([[0], np.cumsum([len(x) for x in wr_parts])]).astype(np.uint64)
    wr_blob = np.frombuffer(b"".join(wr_parts), np.uint8)

    ex = {}
    lsa = ref.size_sa - 1
    for view in range(4):
        for d in (0, 1):
            rec = np.zeros(n, fo.HIT_DTYPE)
            anc = np.zeros(n, fo.HIT_DTYPE)
            for i in range(n):
                W = reads[offsets[i]:offsets[i + 1]]
                L = len(W)
                r, a = ref.exact(view, d, W, 0, lsa, 0, L - 1 if d == 0 else 0, L - 1, anchor=True)
                rec[i] = r
                anc[i] = a
            ex[f"ex_v{view}_d{d}"] = rec
            ex[f"ex_anchor_v{view}_d{d}"] = anc

    rows = np.unique(np.concatenate([np.arange(0, 64), rng.integers(0, len(g) + 1, 2000)]))
    sa_bwd = np.array([ref.get_sa(0, int(m)) for m in rows], np.uint64)
    sa_fwd = np.array([ref.get_sa(1, int(m)) for m in rows], np.uint64)
    isa_bwd = np.array([ref.get_isa(0, int(m)) for m in rows], np.uint64)
    isa_fwd = np.array([ref.get_isa(1, int(m)) for m in rows], np.uint64)
    sha = arrays_sha([ref.arrays(0), ref.arrays(1)])
    bwt, dollar = ref.bwt(0)
    bwti, dollari = ref.bwt(1)
    ref.close()

    np.savez_compressed(
        path, genome=g, reads=reads, offsets=offsets, ratio=np.uint32(RATIO), cap=np.uint32(CAP),
        s1_counts=counts, s1_hits=s1_hits, s1_anchor=anchors, sa_rows=rows.astype(np.uint64),
        sa_bwd=sa_bwd, sa_fwd=sa_fwd, isa_bwd=isa_bwd, isa_fwd=isa_fwd,
        arrays_sha=np.array(sha), bwt_sha=np.array(hashlib.sha256(bwt.tobytes()).hexdigest()),
        bwt_dollar=np.uint64(dollar),
        bwti_sha=np.array(hashlib.sha256(bwti.tobytes()).hexdigest()),
        bwti_dollar=np.uint64(dollari), s1v_counts=v_counts, s1v_hits=s1v_hits,
        s1w_bounds=bounds, s1w_counts=w_counts, s1w_hits=s1w_hits, sk_fragsize=k_fs,
        sk_counts=k_counts, sk_hits=sk_hits, sk_anchor_counts=k_acounts, sk_anchors=sk_anch,
        sk_full=k_full, sk_cap=np.uint32(KCAP), wr_kstart=kstart, wr_kend=kend, wr_blob=wr_blob,
        wr_off=wr_off, wr_done=wr_done, **ex)
    print("wrote", path, os.path.getsize(path), "bytes;", int(counts.sum()), "records")


if __name__ == "__main__":
    main()
