"""Autotuner of the persistent 1-error search launch.

Sweeps the launch shape of fmi_search1_batch on a sample of the caller's
own batch -- workgroups (of 256 lanes) per CU, tasks a wavefront takes from
the device queue per atomic, the number of idle lanes at which a wavefront
refills them, and how an iteration issues its loads (fmi_set_load_mode) --
times each point with per-launch HIP events and installs the fastest with
fmi_set_tuning and fmi_set_load_mode.  Records never depend on the tuning,
only the schedule does.
"""
import itertools

import hpgfm

DEFAULT_GRID = {
    "wg_per_cu": (2, 4, 8, 16),
    "chunk": (16, 64, 256),
    "refill_min": (8, 16, 32, 48, 64),
    "loads": (hpgfm.LOADS_ALL, hpgfm.LOADS_OWN),
}


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
