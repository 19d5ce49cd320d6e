"""The plain comparison point (fmi_plain_*): the reference's GPU design
re-created on this device.  Its occurrence layout must be the reference's
calculate_O output (search/preprocess.c:74-130) byte for byte, checked
against the oracle's arrays, which the golden SHA-256 of the compiled
reference pins; its 1-error results must equal the golden lists and the
engine's, record for record."""
import numpy as np
import pytest

import fmoracle as fo
from conftest import split_golden_hits
from hpgfm import synth

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module", params=["golden_small", "golden_repeats"])
def pgold(request, gpu_engine):
    g = request.getfixturevalue(request.param)
    ix = gpu_engine.Index.build(g["genome"], int(g["ratio"])).to_device()
    p = gpu_engine.PlainSearch(ix)
    yield gpu_engine, g, ix, p
    p.close()
    ix.close()


def test_plain_layout_is_the_reference_layout(pgold):
    hp, g, ix, p = pgold
    o = fo.Oracle(g["genome"], ratio=int(g["ratio"]))
    for fwd in (0, 1):
        a = o.arrays(fwd)
        d, c = p.layout(fwd)
        assert d.shape == a["desp"].shape
        assert np.array_equal(d, a["desp"]), fwd
        assert np.array_equal(c, a["count"]), fwd


@pytest.mark.parametrize("mode", ["edit", "mismatch"])
def test_plain_search1_golden(pgold, mode):
    hp, g, ix, p = pgold
    cap = int(g["cap"])
    m = hp.MODE_EDIT if mode == "edit" else hp.MODE_MISMATCH
    counts, hits, ms_k, ms_t = p.search1(g["reads"], g["offsets"], cap=cap, mode=m)
    assert ms_t >= ms_k >= 0
    for (i, s), w in split_golden_hits(g).items():
        if mode == "mismatch":
            w = w[(w["num_mismatches"] == 0) | (w["err_kind"] == fo.MISMATCH)]
        assert int(counts[i, s]) == len(w), (i, s)
        assert fo.hits_equal(hits[i, s, :min(len(w), cap)], w[:cap]), (i, s)


def test_plain_equals_engine_with_n(gpu_engine):
    hp = gpu_engine
    g, _, _ = synth.genome(400_000, seed=81, repeat_frac=0.3, fam_len=(300, 3000),
                           divergence=(0.0, 0.1))
    ix = hp.Index.build(g, 16).to_device()
    p = hp.PlainSearch(ix)
    lens = np.random.default_rng(82).integers(0, 251, 3000)
    reads, offs, _ = synth.reads(g, 3000, seed=83, sub_rate=0.01, n_rate=0.002, lengths=lens)
    batch = hp.ReadBatch(reads, offs)
    for m in (hp.MODE_MISMATCH, hp.MODE_EDIT):
        c1, h1 = hp.search1_batch(ix, batch, cap=24, mode=m)
        c2, h2, _, _ = p.search1(reads, offs, cap=24, mode=m)
        assert np.array_equal(c1, c2)
        for i in range(len(lens)):
            for s in (0, 1):
                n = min(int(c1[i, s]), 24)
                assert fo.hits_equal(h1[i, s, :n], h2[i, s, :n]), (i, s)
    batch.free()
    p.close()
    ix.close()
