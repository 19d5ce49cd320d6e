"""Open-loop load driver for the serving loop (fmi_server_*).

Requests arrive as a Poisson process at `rate` requests/s; each is a random
window of a pre-generated read pool, of U[min_reads, max_reads] reads.  One
thread submits at the scheduled arrival times (never waiting for results), a
second thread collects completions.  Latency is the server's own measure:
from submit to the moment the request's launch was seen complete.
"""
import threading
import time

import numpy as np

import hpgfm
from hpgfm import synth


def read_pool(genome, nreads, seed=5, sub_rate=0.01, len_range=(50, 250)):
    rng = np.random.default_rng(seed)
    lengths = rng.integers(len_range[0], len_range[1] + 1, nreads)
    reads, offs, _ = synth.reads(genome, nreads, seed=seed + 1, sub_rate=sub_rate, lengths=lengths)
    return reads, offs


def run(server, pool, rate, duration_s, req_reads=(1, 50_000), seed=9, keep=False):
    """Returns a dict with p50/p99/mean latency (ms), served reads/s, request
    count and, when keep=True, the list of (offset window, counts, hits)."""
    reads, offs = pool
    npool = len(offs) - 1
    rng = np.random.default_rng(seed)
    t_arr, t = [], 0.0
    while True:
        t += rng.exponential(1.0 / rate)
        if t > duration_s:
            break
        t_arr.append(t)
    sizes = rng.integers(req_reads[0], req_reads[1] + 1, len(t_arr))
    sizes = np.minimum(sizes, npool)
    starts = rng.integers(0, npool - sizes + 1)
    ids = [None] * len(t_arr)
    lat = [None] * len(t_arr)
    kept = [None] * len(t_arr) if keep else None
    ready = threading.Semaphore(0)
    err = []

    def collector():
        for i in range(len(t_arr)):
            ready.acquire()
            try:
                c, h, ms = server.wait(ids[i])
            except Exception as e:   # noqa: BLE001 - surfaced to the caller below
                err.append(e)
                return
            lat[i] = ms
            if keep:
                kept[i] = (int(starts[i]), int(sizes[i]), c, h)

    th = threading.Thread(target=collector)
    th.start()
    t0 = time.perf_counter()
    for i, ta in enumerate(t_arr):
        now = time.perf_counter() - t0
        if ta > now:
            time.sleep(ta - now)
        a, b = int(starts[i]), int(starts[i] + sizes[i])
        o = offs[a:b + 1]
        ids[i] = server.submit(reads[int(o[0]):int(o[-1])], o - o[0])
        ready.release()
    th.join()
    wall = time.perf_counter() - t0
    if err:
        raise err[0]
    lat_a = np.array(lat, dtype=np.float64)
    total = int(sizes.sum())
    out = {
        "rate_req_s": rate,
        "requests": len(t_arr),
        "reads": total,
        "wall_s": round(wall, 3),
        "reads_per_s": round(total / wall) if wall > 0 else 0,
        "p50_ms": round(float(np.percentile(lat_a, 50)), 3) if len(lat_a) else None,
        "p99_ms": round(float(np.percentile(lat_a, 99)), 3) if len(lat_a) else None,
        "mean_ms": round(float(lat_a.mean()), 3) if len(lat_a) else None,
        "server": server.stats(),
    }
    if keep:
        out["kept"] = kept
    return out


def run_closed(server, pool, nrequests, window=16, req_reads=(1, 50_000), seed=11, clients=4):
    """Closed loop: `clients` threads each keep window / clients requests in
    flight until `nrequests` have been served in all, so the server runs at
    its capacity without an unbounded backlog (each client's submits copy
    their reads into the server's staging in parallel with the others').
    Returns reads/s and the latency distribution."""
    reads, offs = pool
    npool = len(offs) - 1
    rng = np.random.default_rng(seed)
    sizes = np.minimum(rng.integers(req_reads[0], req_reads[1] + 1, nrequests), npool)
    starts = rng.integers(0, npool - sizes + 1)
    clients = max(1, min(clients, window))
    lat = [[] for _ in range(clients)]
    err = []

    def client(c):
        mine = list(range(c, nrequests, clients))
        per = max(1, window // clients)
        inflight = []
        nxt = 0
        try:
            while nxt < len(mine) or inflight:
                while nxt < len(mine) and len(inflight) < per:
                    q = mine[nxt]
                    a, b = int(starts[q]), int(starts[q] + sizes[q])
                    o = offs[a:b + 1]
                    inflight.append(server.submit(reads[int(o[0]):int(o[-1])], o - o[0]))
                    nxt += 1
                _, _, ms = server.wait(inflight.pop(0))
                lat[c].append(ms)
        except Exception as e:   # noqa: BLE001 - surfaced to the caller below
            err.append(e)

    th = [threading.Thread(target=client, args=(c,)) for c in range(clients)]
    t0 = time.perf_counter()
    for t in th:
        t.start()
    for t in th:
        t.join()
    wall = time.perf_counter() - t0
    if err:
        raise err[0]
    lat_a = np.array([x for v in lat for x in v], dtype=np.float64)
    total = int(sizes.sum())
    return {"window": window, "clients": clients, "requests": nrequests, "reads": total,
            "wall_s": round(wall, 3), "reads_per_s": round(total / wall) if wall > 0 else 0,
            "p50_ms": round(float(np.percentile(lat_a, 50)), 3),
            "p99_ms": round(float(np.percentile(lat_a, 99)), 3),
            "server": server.stats()}
