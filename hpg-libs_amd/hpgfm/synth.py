"""Deterministic synthetic genomes and reads for the FM-index search engine.

Bases are the codes 0..3 of the alphabet "ACGT" (the code of a base is its
position in the index's alphabet string, bwt_commons.c:60-91 of the
reference); 4 marks an N.

genome():  i.i.d. bases at a given GC fraction, optionally overlaid with
           repeat families (mutated copies of random consensus sequences)
           and runs of N, split into named contigs.
reads():   fixed-length or variable-length reads sampled from random
           positions and strands, with uniform substitutions, a fraction of
           i.i.d. random reads, and N bases.
"""
import numpy as np

A, Cb, G, T, N = 0, 1, 2, 3, 4


def iid_bases(rng, n, gc=0.41, chunk=1 << 26):
    """n i.i.d. bases with P(C) = P(G) = gc/2, generated in chunks so that
    multi-gigabase genomes need no float temporaries of their own size."""
    at = (1.0 - gc) / 2.0
    cut = np.array([at, 0.5, 0.5 + gc / 2.0], dtype=np.float64)   # A | C | G | T
    out = np.empty(n, dtype=np.uint8)
    for a in range(0, n, chunk):
        b = min(n, a + chunk)
        out[a:b] = np.searchsorted(cut, rng.random(b - a), side="right").astype(np.uint8)
    return out


def replace_n(codes, seed=0):
    """Index policy for N (code 4) in a reference: each N becomes a
    pseudo-random base drawn from a generator seeded with `seed`, so the
    indexed text is over ACGT only and the choice is reproducible."""
    codes = np.asarray(codes, dtype=np.uint8)
    m = codes > 3
    if not m.any():
        return codes
    out = codes.copy()
    out[m] = np.random.default_rng(seed).integers(0, 4, int(m.sum()), dtype=np.uint8)
    return out


def _genome_v1(n, seed, gc, repeat_frac, fam_len, divergence, n_frac, n_run, contigs,
               copies_per_family):
    """The first generator (float draws, one copy at a time): kept so that
    tests/golden/make_golden.py reproduces the committed fixtures."""
    rng = np.random.default_rng(seed)
    g = iid_bases(rng, n, gc)
    if repeat_frac > 0 and n > fam_len[1] * 4:
        target = int(repeat_frac * n)
        covered = 0
        while covered < target:
            L = int(rng.integers(fam_len[0], fam_len[1] + 1))
            cons = iid_bases(rng, L, gc)
            for _ in range(copies_per_family):
                if covered >= target:
                    break
                p = int(rng.integers(0, n - L))
                d = float(rng.uniform(divergence[0], divergence[1]))
                cp = cons.copy()
                m = rng.random(L) < d
                cp[m] = (cp[m] + rng.integers(1, 4, size=int(m.sum()))) % 4
                if rng.random() < 0.5:
                    cp = (3 - cp[::-1]).astype(np.uint8)
                g[p:p + L] = cp
                covered += L
    if n_frac > 0:
        target = int(n_frac * n)
        covered = 0
        while covered < target:
            L = int(rng.integers(n_run[0], n_run[1] + 1))
            p = int(rng.integers(0, max(1, n - L)))
            g[p:p + L] = N
            covered += L
    if contigs > 1:
        cuts = set()
        while len(cuts) < contigs - 1:
            cuts.add(int(rng.integers(1000, n - 1000)))
        starts = np.concatenate([[0], np.sort(np.array(sorted(cuts), dtype=np.int64))]).astype(np.uint64)
    else:
        starts = np.zeros(1, np.uint64)
    names = [f"chr{i + 1}" for i in range(len(starts))]
    return g, starts, names


def _base_lut(gc):
    """Base of every 16-bit draw: P(A) = P(T) = (1 - gc) / 2, P(C) = P(G) = gc / 2
    to within 2^-16."""
    at = (1.0 - gc) / 2.0
    cuts = np.array([at, 0.5, 0.5 + gc / 2.0]) * 65536.0
    return np.searchsorted(cuts, np.arange(65536) + 0.5, side="right").astype(np.uint8)


def iid_bases_par(seed, n, gc=0.41, chunk=1 << 26, threads=None):
    """n i.i.d. bases at GC fraction gc from 16-bit draws; chunk i has its own
    generator spawned from (seed, i), so the output does not depend on how
    many threads fill it."""
    import os
    from concurrent.futures import ThreadPoolExecutor
    lut = _base_lut(gc)
    out = np.empty(n, dtype=np.uint8)
    nch = (n + chunk - 1) // chunk
    kids = np.random.SeedSequence([int(seed), 0x5EED]).spawn(max(nch, 1))

    def fill(i):
        a = i * chunk
        b = min(n, a + chunk)
        r = np.random.Generator(np.random.PCG64(kids[i]))
        np.take(lut, r.integers(0, 65536, size=b - a, dtype=np.uint16), out=out[a:b])

    if threads is None:
        threads = max(1, min(16, len(os.sched_getaffinity(0))))
    if nch <= 1 or threads == 1:
        for i in range(nch):
            fill(i)
    else:
        with ThreadPoolExecutor(threads) as ex:
            list(ex.map(fill, range(nch)))
    return out


def genome(n, seed=1, gc=0.41, repeat_frac=0.0, fam_len=(300, 6000),
           divergence=(0.02, 0.15), n_frac=0.0, n_run=(50, 5000), contigs=1,
           copies_per_family=20, method="v2"):
    """Returns (codes uint8[n], contig_starts uint64[c], contig_names list).

    Repeat families: a random consensus of fam_len bases, copies_per_family
    mutated copies (substitutions at a per-copy divergence drawn from
    `divergence`, half of them reverse-complemented) pasted at random
    positions until repeat_frac of the genome is covered.  method "v1" is
    the first, slower generator (the golden fixtures' genomes)."""
    if method == "v1":
        return _genome_v1(n, seed, gc, repeat_frac, fam_len, divergence, n_frac, n_run, contigs,
                          copies_per_family)
    g = iid_bases_par(seed, n, gc)
    rng = np.random.default_rng([int(seed), 1])
    lut = _base_lut(gc)
    if repeat_frac > 0 and n > fam_len[1] * 4:
        target = int(repeat_frac * n)
        covered = 0
        while covered < target:
            L = int(rng.integers(fam_len[0], fam_len[1] + 1))
            cons = lut[rng.integers(0, 65536, size=L, dtype=np.uint16)]
            k = int(min(copies_per_family, -(-(target - covered) // L)))
            ps = rng.integers(0, n - L, size=k)
            thr = (rng.uniform(divergence[0], divergence[1], size=k) * 65536.0).astype(np.uint32)
            cp = np.broadcast_to(cons, (k, L)).copy()
            m = rng.integers(0, 65536, size=(k, L), dtype=np.uint16) < thr[:, None]
            cp[m] = (cp[m] + rng.integers(1, 4, size=int(m.sum()), dtype=np.uint8)) % 4
            flip = rng.random(k) < 0.5
            cp[flip] = (3 - cp[flip])[:, ::-1]
            for j in range(k):
                g[ps[j]:ps[j] + L] = cp[j]
            covered += k * L
    if n_frac > 0:
        target = int(n_frac * n)
        covered = 0
        while covered < target:
            L = int(rng.integers(n_run[0], n_run[1] + 1))
            p = int(rng.integers(0, max(1, n - L)))
            g[p:p + L] = N
            covered += L
    if contigs > 1:
        cuts = set()
        while len(cuts) < contigs - 1:
            cuts.add(int(rng.integers(1000, n - 1000)))
        starts = np.concatenate([[0], np.sort(np.array(sorted(cuts), dtype=np.int64))]).astype(np.uint64)
    else:
        starts = np.zeros(1, np.uint64)
    names = [f"chr{i + 1}" for i in range(len(starts))]
    return g, starts, names


def revcomp(codes):
    """Reverse complement of codes; N (4) stays N."""
    c = np.asarray(codes, dtype=np.uint8)[::-1]
    out = np.where(c < 4, 3 - c, c).astype(np.uint8)
    return out


def reads(g, nreads, length=100, seed=2, sub_rate=0.0, random_frac=0.0,
          both_strands=True, n_rate=0.0, lengths=None):
    """Samples reads from genome g.

    length: fixed read length, or ignored when `lengths` (an int array) is
    given.  Returns (concatenated codes uint8, offsets uint64[nreads+1],
    truth dict with pos, strand, is_random)."""
    rng = np.random.default_rng(seed)
    n = len(g)
    if lengths is None:
        lengths = np.full(nreads, length, dtype=np.int64)
    lengths = np.asarray(lengths, dtype=np.int64)
    offsets = np.zeros(nreads + 1, dtype=np.uint64)
    offsets[1:] = np.cumsum(lengths)
    total = int(offsets[-1])
    out = np.empty(total, dtype=np.uint8)
    maxlen = int(lengths.max()) if nreads else 0
    pos = rng.integers(0, n - maxlen + 1, size=nreads)
    strand = rng.integers(0, 2, size=nreads) if both_strands else np.zeros(nreads, np.int64)
    is_rand = rng.random(nreads) < random_frac
    if nreads and np.all(lengths == lengths[0]):
        L = int(lengths[0])
        idx = pos[:, None] + np.arange(L)[None, :]
        mat = g[idx]
        rc = strand == 1
        mat[rc] = np.where(mat[rc] < 4, 3 - mat[rc], mat[rc])[:, ::-1]
        nr = int(is_rand.sum())
        if nr:
            mat[is_rand] = rng.integers(0, 4, size=(nr, L), dtype=np.uint8)
        if sub_rate > 0:
            m = rng.random(mat.shape) < sub_rate
            mat[m] = (mat[m] % 4 + rng.integers(1, 4, size=int(m.sum()), dtype=np.uint8)) % 4
        if n_rate > 0:
            m = rng.random(mat.shape) < n_rate
            mat[m] = N
        out[:] = mat.reshape(-1)
    elif nreads:
        # variable lengths, vectorised over all bases: base j of read i comes
        # from pos[i] + j (+ strand) or pos[i] + L_i - 1 - j (- strand)
        starts = offsets[:-1].astype(np.int64)
        local = np.arange(total, dtype=np.int64) - np.repeat(starts, lengths)
        rev = np.repeat(strand == 1, lengths)
        lrep = np.repeat(lengths, lengths)
        src = np.repeat(pos.astype(np.int64), lengths) + np.where(rev, lrep - 1 - local, local)
        s = g[src]
        s = np.where(rev & (s < 4), 3 - s, s).astype(np.uint8)
        rnd = np.repeat(is_rand, lengths)
        nr = int(rnd.sum())
        if nr:
            s[rnd] = rng.integers(0, 4, size=nr, dtype=np.uint8)
        if sub_rate > 0:
            m = rng.random(total) < sub_rate
            s[m] = (s[m] % 4 + rng.integers(1, 4, size=int(m.sum()), dtype=np.uint8)) % 4
        if n_rate > 0:
            s[rng.random(total) < n_rate] = N
        out[:] = s
    return out, offsets, dict(pos=pos, strand=strand, is_random=is_rand)


def lognormal_lengths(rng, nreads, median=120, sigma=0.35, lo=50, hi=250):
    x = rng.lognormal(mean=np.log(median), sigma=sigma, size=nreads)
    return np.clip(np.rint(x), lo, hi).astype(np.int64)
