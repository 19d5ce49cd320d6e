"""Autotuner of the persistent 1-error search launch.

Sweeps the launch shape of fmi_search1_batch -- workgroups (of 256 lanes)
per CU, tasks a wavefront takes from the device queue per atomic, the number
of idle lanes at which a wavefront refills them, and how an iteration issues
its loads (fmi_set_load_mode) -- on a tuning batch of the workload's shape
(its own seed, not the batch that is then timed), times each point with
per-launch HIP events and installs the fastest with fmi_set_tuning and
fmi_set_load_mode.  Records never depend on the tuning, only the schedule
does.

Results persist in tuning.json beside this file, one entry per device and
workload key, so a later run (and a profiler run of the same command) uses
the same launch shape without re-tuning.
"""
import itertools
import json
import os

import hpgfm

DEFAULT_GRID = {
    "wg_per_cu": (2, 4, 8, 16),
    "chunk": (16, 64, 256),
    "refill_min": (1, 4, 8, 16, 32, 48, 64),
    "loads": (hpgfm.LOADS_ALL, hpgfm.LOADS_OWN),
}

TUNING_FILE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "tuning.json")


def time_point(ix, batch, out, mode, reps=2, plain=False):
    hpgfm.search1_launch(ix, batch, out, mode=mode, plain=plain)
    hpgfm.sync()
    with hpgfm.TimingRegion() as tr:
        for _ in range(reps):
            hpgfm.search1_launch(ix, batch, out, mode=mode, plain=plain)
        hpgfm.sync()
    return tr.min_ms


def autotune(ix, batch, out, mode=hpgfm.MODE_MISMATCH, grid=None, reps=2, log=None):
    """Returns (best (wg_per_cu, chunk, refill_min, loads), best_ms, table)
    and leaves the best tuning installed.  `table` lists every point, plus
    the plain one-thread-per-read launch under key "plain"."""
    grid = grid or DEFAULT_GRID
    table = {}
    best, best_ms = None, float("inf")
    for loads, wg, chunk, refill in itertools.product(grid["loads"], grid["wg_per_cu"],
                                                      grid["chunk"], grid["refill_min"]):
        hpgfm.set_load_mode(loads)
        hpgfm.set_tuning(wg, chunk, refill)
        ms = time_point(ix, batch, out, mode, reps)
        table[(wg, chunk, refill, loads)] = ms
        if ms < best_ms:
            best, best_ms = (wg, chunk, refill, loads), ms
    table["plain"] = time_point(ix, batch, out, mode, reps, plain=True)
    hpgfm.set_tuning(*best[:3])
    hpgfm.set_load_mode(best[3])
    if log:
        log(f"autotune: best wg/cu={best[0]} chunk={best[1]} refill={best[2]} loads={best[3]} "
            f"{best_ms:.3f} ms (plain {table['plain']:.3f} ms, {len(table) - 1} points)")
    return best, best_ms, table


def device_key():
    """The GPU the tuning was measured on (name and CU count)."""
    return f"{hpgfm.device_name()}/{hpgfm.cu_count()}cu"


def load(key, path=TUNING_FILE):
    """The persisted entry for (this device, key), or None."""
    try:
        with open(path) as f:
            db = json.load(f)
    except (OSError, ValueError):
        return None
    return db.get(device_key(), {}).get(key)


def save(key, entry, path=TUNING_FILE):
    try:
        with open(path) as f:
            db = json.load(f)
    except (OSError, ValueError):
        db = {}
    db.setdefault(device_key(), {})[key] = entry
    tmp = path + ".tmp"
    with open(tmp, "w") as f:
        json.dump(db, f, indent=1, sort_keys=True)
    os.replace(tmp, path)


def install(entry):
    hpgfm.set_tuning(entry["wg_per_cu"], entry["chunk"], entry["refill_min"])
    hpgfm.set_load_mode(entry["loads"])
