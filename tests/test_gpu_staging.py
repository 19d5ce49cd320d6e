"""The staged 1-error search (fmi_set_staging(1), the default): every lane
holds its read in LDS, packed once into a 128-byte record, and reads longer
than 256 bases go through the unstaged kernel from a list afterwards.  Its
counts and records must equal the unstaged search's and the oracle's, bit
for bit, at every read length around the slot's word boundaries (16-base
code words, 32-base flag words, the 256-base slot), with N bases, empty
reads, truncating capacities, every strand mask and both load modes, on the
compact index, the expanded index and the expanded index with a k-mer
table."""
import numpy as np
import pytest

import fmoracle as fo
from hpgfm import synth

pytestmark = pytest.mark.gpu

EDGE_LENGTHS = [0, 1, 2, 5, 15, 16, 17, 31, 32, 33, 47, 48, 63, 64, 65, 95, 96, 127, 128,
                129, 150, 191, 192, 200, 223, 224, 225, 240, 255, 256, 257, 258, 300, 511,
                512, 999, 1000]


@pytest.fixture(scope="module")
def staged_setup(gpu_engine):
    hp = gpu_engine
    g, _, _ = synth.genome(400_000, seed=301, repeat_frac=0.35, fam_len=(300, 3000),
                           divergence=(0.0, 0.05), copies_per_family=10)
    rng = np.random.default_rng(302)
    lens = np.concatenate([np.repeat(EDGE_LENGTHS, 12), rng.integers(0, 400, 1500)])
    rng.shuffle(lens)
    reads, offs, _ = synth.reads(g, len(lens), seed=303, sub_rate=0.015, n_rate=0.004,
                                 lengths=lens)
    compact = hp.Index.build(g, 16).to_device()
    expanded = hp.Index.build(g, 16).to_device().expand()
    tabled = hp.Index.build(g, 16).to_device().expand()
    tabled.kmer_table(11)
    batch = hp.ReadBatch(reads, offs)
    yield hp, g, reads, offs, lens, batch, {"compact": compact, "expanded": expanded,
                                            "kmer11": tabled}
    batch.free()
    for ix in (compact, expanded, tabled):
        ix.close()


def _same(a, b, cap):
    (ca, ha), (cb, hb) = a, b
    if not np.array_equal(ca, cb):
        bad = np.argwhere(ca != cb)[:5]
        return f"counts differ at {bad.tolist()}"
    n = np.minimum(ca, cap)
    for i in range(ca.shape[0]):
        for s in (0, 1):
            if not fo.hits_equal(ha[i, s, :n[i, s]], hb[i, s, :n[i, s]]):
                return f"records differ at read {i} strand {s}"
    return None


@pytest.mark.parametrize("which", ["compact", "expanded", "kmer11"])
def test_staged_equals_unstaged(staged_setup, which):
    hp, g, reads, offs, lens, batch, ixs = staged_setup
    ix = ixs[which]
    assert hp.get_staging()
    try:
        for lm in (hp.LOADS_ALL, hp.LOADS_OWN):
            hp.set_load_mode(lm)
            for m in (hp.MODE_EDIT, hp.MODE_MISMATCH):
                for mask in (3, 1, 2):
                    hp.set_staging(True)
                    staged = hp.search1_batch(ix, batch, cap=6, mode=m, strand_mask=mask)
                    hp.set_staging(False)
                    plain = hp.search1_batch(ix, batch, cap=6, mode=m, strand_mask=mask)
                    err = _same(staged, plain, 6)
                    assert err is None, (which, lm, m, mask, err)
    finally:
        hp.set_staging(True)
        hp.set_load_mode(hp.LOADS_AUTO)


def test_staged_vs_oracle(staged_setup):
    """The staged search on the expanded index with the k-mer table against
    the oracle on every read of up to 300 bases and the long ones too."""
    hp, g, reads, offs, lens, batch, ixs = staged_setup
    ix = ixs["kmer11"]
    cap = 24
    o = fo.Oracle(bwts=(*ix.bwt(0), *ix.bwt(1)), ratio=16)
    for mode in (hp.MODE_EDIT, hp.MODE_MISMATCH):
        counts, hits = hp.search1_batch(ix, batch, cap=cap, mode=mode)
        _, oc, oh = o.search1_batch(reads, offs, cap=cap, mismatch_only=bool(mode))
        oc = oc.reshape(-1, 2)
        oh = oh.reshape(-1, 2, cap)
        assert np.array_equal(counts, oc), mode
        for i in range(batch.nreads):
            for s in (0, 1):
                c = min(int(counts[i, s]), cap)
                assert fo.hits_equal(hits[i, s, :c], oh[i, s, :c]), (mode, i, s, int(lens[i]))


def test_staged_all_long_and_all_short(staged_setup):
    """A batch of only reads longer than a slot (every task goes through the
    list) and a batch of only reads of at most one code word."""
    hp, g, *_ = staged_setup
    ix = staged_setup[6]["kmer11"]
    for lens in (np.full(700, 600), np.random.default_rng(304).integers(0, 17, 3000)):
        reads, offs, _ = synth.reads(g, len(lens), seed=305, sub_rate=0.01, lengths=lens)
        b = hp.ReadBatch(reads, offs)
        try:
            staged = hp.search1_batch(ix, b, cap=8, mode=hp.MODE_EDIT)
            hp.set_staging(False)
            plain = hp.search1_batch(ix, b, cap=8, mode=hp.MODE_EDIT)
        finally:
            hp.set_staging(True)
            b.free()
        assert _same(staged, plain, 8) is None
