"""Read ingestion from the reference's fastq_batch_t layout
(bioformats/fastq/fastq_batch.h:24-34: sequences with a NUL after each,
int data_indices): the device encoding equals encode_bases
(bwt_commons.c:99-102, fmi_encode_bases on the host) byte for byte with the
offsets the NULs imply; the 2-bit packing holds every code and N flag; and
requests submitted to the serving loop in that layout get the records of
fmi_search1_batch on their encoded reads."""
import numpy as np
import pytest

import fmoracle as fo
from hpgfm import synth

pytestmark = pytest.mark.gpu

LETTERS = np.frombuffer(b"ACGTacgtNnRYKM.-", np.uint8)


def random_texts(rng, n, lo=0, hi=160, odd=0.02):
    out = []
    for _ in range(n):
        L = int(rng.integers(lo, hi))
        t = LETTERS[rng.integers(0, 4, L)].copy()
        low = rng.random(L) < 0.1
        t[low] = LETTERS[rng.integers(4, 8, int(low.sum()))]
        weird = rng.random(L) < odd
        t[weird] = LETTERS[rng.integers(8, 16, int(weird.sum()))]
        out.append(t.tobytes())
    return out


def test_fastq_encode_equals_host_encoding(gpu_engine):
    hp = gpu_engine
    texts = random_texts(np.random.default_rng(141), 3000)
    texts[:3] = [b"", b"A", b""]                     # empty reads keep their NUL
    seq, di = hp.fastq_batch(texts)
    codes, offs = hp.fastq_encode_batch(seq, di)
    want = hp.encode_bases(b"".join(texts))
    assert np.array_equal(codes, want)
    lens = np.array([len(t) for t in texts])
    assert np.array_equal(offs, np.concatenate([[0], np.cumsum(lens)]).astype(np.uint64))
    # data_indices that do not start at 0 (a window of a larger batch)
    codes2, offs2 = hp.fastq_encode_batch(seq, di[1000:2001])
    assert np.array_equal(codes2, want[int(offs[1000]):int(offs[2000])])
    assert np.array_equal(offs2, offs[1000:2001] - offs[1000])


def test_pack_codes_roundtrip(gpu_engine):
    hp = gpu_engine
    rng = np.random.default_rng(142)
    codes = rng.integers(0, 4, 100_003).astype(np.uint8)
    codes[rng.random(len(codes)) < 0.01] = 4
    packed, nmask = hp.pack_codes(codes)
    i = np.arange(len(codes))
    got = (packed[i // 16] >> (2 * (i % 16)).astype(np.uint32)) & 3
    isn = (nmask[i // 16] >> (i % 16).astype(np.uint16)) & 1
    assert np.array_equal(isn.astype(bool), codes > 3)
    assert np.array_equal(got[codes < 4], codes[codes < 4])
    assert np.all(got[codes > 3] == 0)


def test_server_fastq_requests(gpu_engine):
    hp = gpu_engine
    g, _, _ = synth.genome(1_000_000, seed=143, repeat_frac=0.2)
    ix = hp.Index.build(g, 16).to_device()
    rng = np.random.default_rng(144)
    reads, offs, _ = synth.reads(g, 20_000, seed=145, sub_rate=0.01, n_rate=0.002,
                                 lengths=rng.integers(20, 200, 20_000))
    alpha = np.frombuffer(b"ACGTN", np.uint8)
    cap = 6
    srv = hp.Server(ix, budget_bytes=32 << 20, max_inflight=3, max_batch_reads=8000, cap=cap)
    try:
        jobs = []
        a = 0
        while a < 20_000:
            b = min(20_000, a + int(rng.integers(1, 3000)))
            texts = [alpha[reads[int(offs[i]):int(offs[i + 1])]].tobytes() for i in range(a, b)]
            if len(jobs) % 2:
                texts = [t.lower() for t in texts]
            seq, di = hp.fastq_batch(texts)
            if len(jobs) % 3 == 0:
                rid = srv.submit_fastq(seq, di)
            else:
                o = offs[a:b + 1]
                rid = srv.submit(reads[int(o[0]):int(o[-1])], o - o[0])
            jobs.append((a, b, rid))
            a = b
        for a, b, rid in jobs:
            c, h, _ = srv.wait(rid, timeout_us=60_000_000)
            o = offs[a:b + 1]
            batch = hp.ReadBatch(reads[int(o[0]):int(o[-1])], o - o[0])
            c2, h2 = hp.search1_batch(ix, batch, cap=cap, mode=hp.MODE_MISMATCH)
            batch.free()
            assert np.array_equal(c, c2), (a, b)
            alone = h2[np.arange(cap)[None, None, :] < np.minimum(c2, cap)[:, :, None]]
            assert fo.hits_equal(h, alone), (a, b)
    finally:
        srv.close()
        ix.close()
