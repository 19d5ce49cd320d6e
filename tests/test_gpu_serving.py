"""The serving loop (fmi_server_*): every request's records equal the records
fmi_search1_batch gives for its reads alone, under concurrent load with
several launches in flight; admission rejects requests that can never fit."""
import numpy as np
import pytest

import fmoracle as fo
from hpgfm import synth

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def served(gpu_engine):
    hp = gpu_engine
    g, _, _ = synth.genome(2_000_000, seed=91, repeat_frac=0.2)
    ix = hp.Index.build(g, 16).to_device()
    from hpgfm import serve
    pool = serve.read_pool(g, 60_000, seed=92)
    return hp, ix, pool


def test_served_results_equal_alone(served):
    hp, ix, pool = served
    from hpgfm import serve
    cap = 8
    srv = hp.Server(ix, budget_bytes=64 << 20, max_inflight=3, max_batch_reads=20_000, cap=cap,
                    mode=hp.MODE_MISMATCH)
    res = serve.run(srv, pool, rate=400, duration_s=0.5, req_reads=(1, 6000), keep=True)
    st = srv.stats()
    srv.close()
    assert res["requests"] > 50
    assert st["requests_done"] == res["requests"]
    assert st["launches"] >= 2
    reads, offs = pool
    for start, size, c, recs in res["kept"]:
        o = offs[start:start + size + 1]
        batch = hp.ReadBatch(reads[int(o[0]):int(o[-1])], o - o[0])
        c2, h2 = hp.search1_batch(ix, batch, cap=cap, mode=hp.MODE_MISMATCH)
        batch.free()
        assert np.array_equal(c, c2)
        alone = [h2[i, s, :min(int(c2[i, s]), cap)] for i in range(size) for s in (0, 1)]
        alone = np.concatenate(alone) if alone else np.zeros(0, recs.dtype)
        assert fo.hits_equal(recs, alone)


def test_admission_rejects_oversize(served):
    hp, ix, pool = served
    srv = hp.Server(ix, budget_bytes=1 << 20, max_inflight=1, max_batch_reads=100_000, cap=16)
    reads, offs = pool
    with pytest.raises(hp.FmiError) as e:
        srv.submit(reads[:int(offs[5000])], offs[:5001])      # ~ 5000*2*16*32 B > 1 MiB
    assert e.value.code == -6
    rid = srv.submit(reads[:int(offs[10])], offs[:11])         # small ones still served
    c, h, ms = srv.wait(rid)
    assert c.shape == (10, 2) and ms >= 0
    srv.close()


def test_more_reads_than_budget_wait_for_memory(served):
    """Requests that together exceed the arena are admitted as earlier
    launches free their memory (no deadlock, every request completes)."""
    hp, ix, pool = served
    reads, offs = pool
    srv = hp.Server(ix, budget_bytes=3 << 20, max_inflight=2, max_batch_reads=100_000, cap=4)
    ids = []
    for i in range(40):
        a, b = i * 1000, i * 1000 + 1000
        o = offs[a:b + 1]
        ids.append(srv.submit(reads[int(o[0]):int(o[-1])], o - o[0]))
    for rid in ids:
        c, h, ms = srv.wait(rid, timeout_us=60_000_000)
        assert c.shape == (1000, 2)
    assert srv.stats()["requests_done"] == 40
    srv.close()


def test_backpressure_refuses_then_serves(served):
    """A full queue (max_pending) or a full input staging refuses a submit
    with FMI_EAGAIN and queues nothing; every accepted request is served with
    the records of its reads alone, and a refused one goes through once the
    queue drains."""
    hp, ix, pool = served
    reads, offs = pool
    cap = 4
    srv = hp.Server(ix, budget_bytes=8 << 20, max_inflight=1, max_batch_reads=2000, cap=cap,
                    max_pending=3, staging_bytes=2 << 20)
    accepted, refused = [], 0
    for i in range(30):
        a, b = i * 2000, i * 2000 + 2000
        o = offs[a:b + 1]
        rid = srv.submit(reads[int(o[0]):int(o[-1])], o - o[0], block=False)
        if rid is None:
            refused += 1
        else:
            accepted.append((a, rid))
    st = srv.stats()
    assert refused > 0 and st["refused"] == refused and st["pending"] <= 3
    # the same request again, blocking: it waits for room and is served
    o = offs[0:2001]
    late = srv.submit(reads[int(o[0]):int(o[-1])], o - o[0])
    accepted.append((0, late))
    for a, rid in accepted:
        c, h, ms = srv.wait(rid, timeout_us=60_000_000)
        o = offs[a:a + 2001]
        batch = hp.ReadBatch(reads[int(o[0]):int(o[-1])], o - o[0])
        c2, h2 = hp.search1_batch(ix, batch, cap=cap, mode=hp.MODE_MISMATCH)
        batch.free()
        assert np.array_equal(c, c2)
        alone = h2[np.arange(cap)[None, None, :] < np.minimum(c2, cap)[:, :, None]]
        assert fo.hits_equal(h, alone)
    assert srv.stats()["requests_done"] == len(accepted)
    srv.close()


def test_cancel(served):
    """Cancelled requests that were still queued are never served (their ids
    are released); the others are served with the records of their reads
    alone; done + cancelled = submitted."""
    hp, ix, pool = served
    reads, offs = pool
    cap = 4
    srv = hp.Server(ix, budget_bytes=8 << 20, max_inflight=1, max_batch_reads=2000, cap=cap)
    ids = []
    for i in range(24):
        a = i * 2000
        o = offs[a:a + 2001]
        ids.append((a, srv.submit(reads[int(o[0]):int(o[-1])], o - o[0])))
    cancelled = set()
    for j in range(23, 5, -3):          # from the back of the queue: still pending
        srv.cancel(ids[j][1])
        cancelled.add(ids[j][1])
    for a, rid in ids:
        if rid in cancelled:
            with pytest.raises(hp.FmiError) as e:
                srv.wait(rid, timeout_us=1000)
            assert e.value.code in (-1, -9)
            continue
        c, h, ms = srv.wait(rid, timeout_us=60_000_000)
        o = offs[a:a + 2001]
        batch = hp.ReadBatch(reads[int(o[0]):int(o[-1])], o - o[0])
        c2, h2 = hp.search1_batch(ix, batch, cap=cap, mode=hp.MODE_MISMATCH)
        batch.free()
        assert np.array_equal(c, c2)
        alone = h2[np.arange(cap)[None, None, :] < np.minimum(c2, cap)[:, :, None]]
        assert fo.hits_equal(h, alone)
    st = srv.stats()
    assert st["cancelled"] == len(cancelled)
    assert st["requests_done"] + st["cancelled"] == len(ids)
    srv.close()
