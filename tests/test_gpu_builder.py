"""The GPU index builder (fmi_index_build_gpu) writes the same index as the
host builder: same BWT and '$' row in both directions, same samples (checked
through locate), same search results; and the golden reference BWT."""
import hashlib

import numpy as np
import pytest

from hpgfm import synth

pytestmark = pytest.mark.gpu

TEXTS = {
    "one": np.array([2], np.uint8),
    "two": np.array([3, 0], np.uint8),
    "allA_97": np.zeros(97, np.uint8),
    "allA_5000": np.zeros(5000, np.uint8),
    "period3": np.tile(np.array([0, 1, 2], np.uint8), 50),
    "desc": np.array([3] * 10 + [2] * 10 + [1] * 10 + [0] * 10, np.uint8),
    "tandem": np.tile(np.random.default_rng(1).integers(0, 4, 37).astype(np.uint8), 300),
}


def same_index(hp, a, b):
    for f in (0, 1):
        ba, da = a.bwt(f)
        bb, db = b.bwt(f)
        assert da == db
        assert np.array_equal(ba, bb)


@pytest.mark.parametrize("name", sorted(TEXTS))
def test_gpu_builder_edge_texts(gpu_engine, name):
    hp = gpu_engine
    t = TEXTS[name]
    same_index(hp, hp.Index.build_gpu(t, 4), hp.Index.build(t, 4))


@pytest.mark.parametrize("n,rep", [(1000, 0.0), (200_000, 0.3), (3_000_000, 0.45)])
def test_gpu_builder_genomes(gpu_engine, n, rep):
    hp = gpu_engine
    g, _, _ = synth.genome(n, seed=n % 97, repeat_frac=rep, fam_len=(300, 3000),
                           divergence=(0.0, 0.05))
    a = hp.Index.build_gpu(g, 16)
    b = hp.Index.build(g, 16).to_device()
    same_index(hp, a, b)
    rows = np.random.default_rng(3).integers(0, n + 1, 4000).astype(np.int64)
    for fwd in (0, 1):
        pa, _ = hp.locate_batch(a, rows, rows, cap=1, forward=fwd)
        pb, _ = hp.locate_batch(b, rows, rows, cap=1, forward=fwd)
        assert np.array_equal(pa, pb)
    reads, offs, _ = synth.reads(g, 3000, length=100, seed=4, sub_rate=0.01)
    batch = hp.ReadBatch(reads, offs)
    ca, ha = hp.search1_batch(a, batch, cap=8, mode=hp.MODE_EDIT)
    cb, hb = hp.search1_batch(b, batch, cap=8, mode=hp.MODE_EDIT)
    assert np.array_equal(ca, cb) and np.array_equal(ha, hb)
    batch.free()


@pytest.mark.parametrize("name", ["golden_small", "golden_repeats"])
def test_gpu_builder_matches_reference_bwt(gpu_engine, name, request):
    g = request.getfixturevalue(name)
    ix = gpu_engine.Index.build_gpu(g["genome"], int(g["ratio"]))
    b, d = ix.bwt(0)
    bi, di = ix.bwt(1)
    assert hashlib.sha256(b.tobytes()).hexdigest() == str(g["bwt_sha"]) and d == int(g["bwt_dollar"])
    assert hashlib.sha256(bi.tobytes()).hexdigest() == str(g["bwti_sha"]) and di == int(g["bwti_dollar"])
