"""Random-access and streaming roofline probes of the GPU the index lives on.

    python hpg-libs_amd/hpgfm/probe.py [--sizes-gib 0.0625,0.25,1,4,8,32] [--lines 16,32,64,128,256]

For each buffer size and line size: random aligned line gathers
(fmi_gather_bench, >= 2^27 loads, each lane keeping several lines in flight),
reported in GB/s and in lines/s; then one streaming copy (fmi_copy_bench).
The FM-index search is a chain of dependent random 64-byte block loads, so the
gather rate at the index's size bounds it; the line-size sweep shows whether
a cost is per line or per byte in each cache regime.
"""
import argparse
import json
import os
import sys

sys.path.insert(0, os.path.dirname(os.path.dirname(os.path.abspath(__file__))))

import hpgfm  # noqa: E402


def run(sizes_gib, line_sizes, nloads=1 << 27):
    rows = []
    for gib in sizes_gib:
        nbytes = int(gib * (1 << 30))
        for lb in line_sizes:
            ms, done = hpgfm.gather_bench(nbytes, lb, nloads)
            rows.append({"buffer_gib": gib, "line_bytes": lb, "ms": round(ms, 4),
                         "GBps": round(done * lb / (ms / 1e3) / 1e9, 1),
                         "Glines_per_s": round(done / (ms / 1e3) / 1e9, 2)})
            print(f"gather  buf {gib:>8} GiB  line {lb:>4} B  {rows[-1]['GBps']:8.1f} GB/s  "
                  f"{rows[-1]['Glines_per_s']:6.2f} G lines/s  ({ms:.2f} ms)", flush=True)
    cms = hpgfm.copy_bench(2 << 30, reps=3)
    copy = round(2 * (2 << 30) / (cms / 1e3) / 1e9, 1)
    print(f"copy    2 GiB -> 2 GiB  {copy:8.1f} GB/s (read + write)  ({cms:.3f} ms)", flush=True)
    return {"gather": rows, "copy_GBps": copy}


def main():
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--sizes-gib", default="0.0625,0.25,1,4,8,32")
    ap.add_argument("--lines", default="16,32,64,128,256")
    ap.add_argument("--json", default=None, help="also write the table to this file")
    a = ap.parse_args()
    if hpgfm.device_count() < 1:
        print("probe: no GPU visible", file=sys.stderr)
        sys.exit(1)
    res = run([float(x) for x in a.sizes_gib.split(",")], [int(x) for x in a.lines.split(",")])
    if a.json:
        with open(a.json, "w") as f:
            json.dump(res, f, indent=1)


if __name__ == "__main__":
    main()
