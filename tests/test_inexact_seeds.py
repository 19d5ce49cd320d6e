"""Inexact seed mapping: bwt_map_inexact_seeds_seq (bwt.c:3262-3302) and
bwt_map_inexact_seed (bwt.c:1401-1644) -- windows of seed_size bases, the
1-error vector helper on both strands, one region per kept SA row, the
within-10-bp dedup (bwt.c:1593-1615) and the reversed hand-over
(bwt.c:1616-1626).

The mapping layer (bwt.c) does not compile here (it includes the
un-vendored samtools submodule), so the glue is pinned by what fixes it:
CPU: windows copied from a known place of a unique text give a region at
that place (1-based, contig-relative, on the reverse complement's
coordinates for the - strand, none across a contig end), every mismatch
region differs from the text in exactly one base, and the kept regions of a
strand are at least 10 apart on a contig (the dedup's fixed point).
GPU: fmi_map_inexact_seeds_batch through the C ABI against the oracle,
record for record, status bits included."""
import numpy as np
import pytest

import fmoracle as fo
from hpgfm import synth


def karyo(n, parts):
    lens = [n // parts] * (parts - 1) + [n - (n // parts) * (parts - 1)]
    koff = np.concatenate([[0], np.cumsum(lens)]).astype(np.uint64)
    kstart = np.zeros(parts, np.uint64)
    kstart[1] = 100            # a contig that starts at coordinate 100 of its chromosome
    return koff, kstart, lens


def contig_of(koff, p):
    return int(np.searchsorted(koff, p, side="right"))      # 1-based idx


def test_planted_windows_closed_form():
    n = 120_000
    g, _, _ = synth.genome(n, seed=21)
    o = fo.Oracle(g, ratio=8)
    koff, kstart, _ = karyo(n, 3)
    L, ss_, ms = 100, 30, 15
    reads, offs, tr = synth.reads(g, 120, length=L, seed=22)
    # reads whose windows straddle the contig ends, on both strands
    extra = []
    for q in (int(koff[1]) - 45, int(koff[2]) - 80, int(koff[1]) - 5):
        for s in (0, 1):
            W = g[q:q + L].copy()
            extra.append((W if s == 0 else synth.revcomp(W), q, s))
    reads = np.concatenate([reads] + [e[0] for e in extra])
    offs = np.concatenate([offs, offs[-1] + L * np.arange(1, len(extra) + 1)]).astype(np.uint64)
    tr = {"pos": np.concatenate([tr["pos"], [e[1] for e in extra]]),
          "strand": np.concatenate([tr["strand"], [e[2] for e in extra]])}
    checked = crossed = 0
    for i in range(len(offs) - 1):
        W = reads[offs[i]:offs[i + 1]]
        res = o.map_inexact_seeds(W, ss_, ms, koff, kstart, read=i)
        assert len(res) == 2 * 3                    # windows 0-29, 30-59, 60-89; 90-99 < 15
        p, s = int(tr["pos"][i]), int(tr["strand"][i])
        for w in range(3):
            st, se = w * ss_, w * ss_ + ss_ - 1
            # text span of the window on its own strand
            t0 = p + st if s == 0 else p + (L - 1 - se)
            c = contig_of(koff, t0)
            inside = t0 + ss_ <= int(koff[c])
            regs, status = res[2 * w + s]
            assert status == 0
            near = [r for r in regs if int(r["chrom_idx"]) == c and
                    abs(int(r["start"]) - (int(kstart[c - 1]) + t0 - int(koff[c - 1]) + 1)) < 10]
            if not inside:
                crossed += 1
                assert not near
                continue
            assert len(near) == 1, (i, w)
            r = near[0]
            assert int(r["end"]) - int(r["start"]) == ss_ - 1
            assert int(r["seed_id"]) == w and int(r["read"]) == i and int(r["strand"]) == s
            assert int(r["seq_len"]) == L
            assert (int(r["seq_start"]), int(r["seq_end"])) == \
                ((st, se) if s == 0 else (L - se - 1, L - st - 1))
            checked += 1
    assert checked > 300 and crossed == 5


def test_mismatch_regions_and_dedup_fixed_point():
    n = 200_000
    g, _, _ = synth.genome(n, seed=31, repeat_frac=0.4, fam_len=(200, 1500),
                           divergence=(0.0, 0.03), copies_per_family=8)
    o = fo.Oracle(g, ratio=16)
    koff, kstart, _ = karyo(n, 2)
    lens = np.random.default_rng(32).integers(20, 160, 150)
    reads, offs, _ = synth.reads(g, 150, seed=33, sub_rate=0.02, lengths=lens)
    mism = 0
    for i in range(len(offs) - 1):
        W = reads[offs[i]:offs[i + 1]]
        res = o.map_inexact_seeds(W, 24, 10, koff, kstart, read=i, region_cap=4096)
        for t, (regs, status) in enumerate(res):
            assert status & 2 == 0
            starts = regs["start"].astype(np.int64)
            for a in range(len(regs)):
                for b in range(a + 1, len(regs)):
                    if regs["chrom_idx"][a] == regs["chrom_idx"][b]:
                        assert abs(starts[a] - starts[b]) >= 10
            for r in regs:
                if int(r["type_err"]) - (int(r["seed_id"]) * 24) != -1:
                    continue                         # mismatch records only
                wlen = int(r["end"]) - int(r["start"]) + 1
                c = int(r["chrom_idx"])
                t0 = int(koff[c - 1]) + int(r["start"]) - 1 - int(kstart[c - 1])
                a, b = int(r["seq_start"]), int(r["seq_end"])
                V = W[a:b + 1] if int(r["strand"]) == 0 else synth.revcomp(W)[a:b + 1]
                assert b - a + 1 == wlen
                assert int(np.count_nonzero(g[t0:t0 + wlen] != V)) == 1, (i, t)
                mism += 1
    assert mism > 50


# ------------------------------------------------------------------ GPU

@pytest.mark.gpu
@pytest.mark.parametrize("expanded", [False, True])
def test_gpu_inexact_seeds_vs_oracle(gpu_engine, expanded):
    hp = gpu_engine
    n = 300_000
    g, _, _ = synth.genome(n, seed=41, repeat_frac=0.4, fam_len=(200, 2000),
                           divergence=(0.0, 0.04), copies_per_family=6)
    koff, kstart, lens = karyo(n, 3)
    ix = hp.Index.build(g, 16)
    ix.set_karyotype(kstart, kstart + np.array(lens, np.uint64) - 1)
    ix.to_device()
    if expanded:
        ix.expand()
    o = fo.Oracle(g, ratio=16)
    rl = np.random.default_rng(42).integers(1, 220, 400)
    reads, offs, _ = synth.reads(g, 400, seed=43, sub_rate=0.015, n_rate=0.002, lengths=rl)
    batch = hp.ReadBatch(reads, offs)
    # windows of up to 32 bases are searched from registers, longer ones
    # from the read's bytes
    for seed_size, min_seed, rec_cap, region_cap in ((30, 15, 256, 64), (16, 8, 6, 3),
                                                     (32, 16, 256, 64), (33, 16, 256, 64),
                                                     (40, 20, 128, 32)):
        res = hp.map_inexact_seeds_batch(ix, batch, seed_size, min_seed, rec_cap, region_cap)
        for i in range(batch.nreads):
            W = reads[offs[i]:offs[i + 1]]
            want = o.map_inexact_seeds(W, seed_size, min_seed, koff, kstart, read=i,
                                       rec_cap=rec_cap, region_cap=region_cap)
            w0 = int(res.window_base[i])
            assert int(res.window_base[i + 1]) - w0 == len(want) // 2
            for t, (regs, st) in enumerate(want):
                w, s = w0 + t // 2, t % 2
                assert int(res.status[w, s]) == st, (seed_size, i, t)
                got = res.regions(w, s)
                assert len(got) == len(regs), (seed_size, i, t)
                assert got.tobytes() == regs.tobytes(), (seed_size, i, t)
    batch.free()
    ix.close()
