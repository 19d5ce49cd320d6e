"""The reference-named C surface (include/hpg_search.h): a C program
compiled with cc against the header -- not ctypes -- runs the reference's
own call patterns for every read of the golden sets (BWSearch1CPU with its
anchor list on both strand pairings, BWExactSearch* on the four views,
get_SA / get_ISA, BWExactSearchVector* driving BWSearch1VectorHelper on whole
reads and windows, BWSearchCPU with its anchors, binsearch) and its
results_list / result structures equal the lists the compiled reference
wrote into the fixtures (search/search.c, search/search.h,
search/results.h:13-37).

CPU: the program compiles and links with -std=c99 -Wall -Werror, and the
`result` it sees has the reference's field offsets.
GPU: the program's outputs against golden_small and golden_repeats."""
import os
import subprocess

import numpy as np
import pytest

import fmoracle as fo
from conftest import ROOT, split_golden_hits

INCLUDE = os.path.join(ROOT, "include")
LIBDIR = os.path.join(ROOT, "hpg-libs_amd", "hpgfm")
SRC = os.path.join(ROOT, "tests", "c_abi", "golden_surface.c")

# result (search/results.h:13-21) as a C compiler lays it out on x86-64
RESULT_DTYPE = np.dtype({
    "names": ["dir", "k", "l", "start", "pos", "end", "num_mismatches", "err_kind", "err_base",
              "err_pos"],
    "formats": ["u1", "<i8", "<i8", "<i2", "<i2", "<i2", "u1", ("u1", 20), ("u1", 20),
                ("<i2", 20)],
    "offsets": [0, 8, 16, 24, 26, 28, 30, 31, 51, 72],
    "itemsize": 112,
})


def cc(src, out, extra=()):
    cmd = ["gcc", "-std=c99", "-O2", "-Wall", "-Wextra", "-Werror", f"-I{INCLUDE}", src,
           f"-L{LIBDIR}", "-lhpgfm", f"-Wl,-rpath,{LIBDIR}", "-o", out, *extra]
    r = subprocess.run(cmd, capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    return out


def test_c_caller_compiles_against_header(tmp_path):
    cc(SRC, str(tmp_path / "golden_surface"))


def test_result_layout_in_c(tmp_path):
    src = tmp_path / "layout.c"
    src.write_text(
        '#include <stddef.h>\n#include <stdio.h>\n#include "hpg_search.h"\n'
        'int main(void) {\n'
        '  printf("%zu %zu %zu %zu %zu %zu %zu %zu %zu %zu\\n", sizeof(result),\n'
        '         offsetof(result, k), offsetof(result, l), offsetof(result, start),\n'
        '         offsetof(result, pos), offsetof(result, end), offsetof(result, num_mismatches),\n'
        '         offsetof(result, err_kind), offsetof(result, err_base), offsetof(result, err_pos));\n'
        '  printf("%zu %zu\\n", sizeof(results_list), offsetof(results_list, read_index));\n'
        '  return 0;\n}\n')
    exe = cc(str(src), str(tmp_path / "layout"))
    out = subprocess.run([exe], capture_output=True, text=True, check=True).stdout.split()
    got = [int(x) for x in out]
    assert got[:10] == [112, 8, 16, 24, 26, 28, 30, 31, 51, 72]
    assert got[10:] == [32, 24]
    assert [RESULT_DTYPE.fields[f][1] for f in RESULT_DTYPE.names] == [0] + got[1:10]


class Reader:
    def __init__(self, data):
        self.b = data
        self.o = 0

    def u32(self):
        v = int(np.frombuffer(self.b, np.uint32, 1, self.o)[0])
        self.o += 4
        return v

    def arr(self, dtype, count):
        a = np.frombuffer(self.b, dtype, count, self.o).copy()
        self.o += np.dtype(dtype).itemsize * count
        return a

    def results(self):
        return self.arr(RESULT_DTYPE, self.u32())


def as_hits(res, strand):
    h = np.zeros(len(res), fo.HIT_DTYPE)
    for key in ("k", "l", "start", "pos", "end", "dir", "num_mismatches"):
        h[key] = res[key]
    for key in ("err_kind", "err_base", "err_pos"):
        h[key] = res[key][:, 0] if len(res) else 0
    h["strand"] = strand
    return h


def as_hitsk(res):
    h = np.zeros(len(res), fo.HITK_DTYPE)
    for key in ("k", "l", "start", "pos", "end", "dir", "num_mismatches"):
        h[key] = res[key]
    for key in ("err_kind", "err_base", "err_pos"):
        h[key] = res[key][:, :8] if len(res) else 0
    return h


def split_k(g, key, counts_key):
    counts = g[counts_key]
    flat = g[key]
    out, o = {}, 0
    for i in range(counts.shape[0]):
        for s in (0, 1):
            c = int(counts[i, s])
            out[(i, s)] = flat[o:o + c]
            o += c
    return out


@pytest.mark.gpu
@pytest.mark.parametrize("name", ["golden_small", "golden_repeats"])
def test_c_caller_reproduces_golden(name, request, tmp_path, gpu_engine):
    g = request.getfixturevalue(name)
    exe = cc(SRC, str(tmp_path / "golden_surface"))
    reads, offs = g["reads"], g["offsets"].astype(np.uint64)
    n, nreads = len(g["genome"]), len(offs) - 1
    cap, kcap = int(g["cap"]), int(g["sk_cap"])
    rows = g["sa_rows"].astype(np.uint64)
    hdr = np.array([n, int(g["ratio"]) | (cap << 32), nreads, kcap], np.uint64)
    inp = tmp_path / "in.bin"
    with open(inp, "wb") as f:
        for a in (hdr, g["genome"].astype(np.uint8), offs, reads.astype(np.uint8),
                  g["s1w_bounds"].astype(np.int16), g["sk_fragsize"].astype(np.int16),
                  np.array([len(rows)], np.uint64), rows):
            f.write(np.ascontiguousarray(a).tobytes())
    outp = tmp_path / "out.bin"
    print(f"[hpg_search C caller] {name}: {nreads} reads", flush=True)
    r = subprocess.run([exe, str(inp), str(outp)], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stdout + r.stderr
    rd = Reader(open(outp, "rb").read())
    # BWSearch1CPU and its anchors
    want = split_golden_hits(g)
    for i in range(nreads):
        for s in (0, 1):
            full = rd.u32()
            lst = rd.results()
            anc = rd.results()
            assert full == 0, (i, s)
            assert len(lst) == int(g["s1_counts"][i, s]), (i, s)
            assert fo.hits_equal(as_hits(lst, s), want[(i, s)]), (i, s)
            wa = g["s1_anchor"][i, s]
            assert len(anc) == 2
            for key in ("k", "l", "start", "pos", "end", "dir"):
                assert np.array_equal(anc[key], wa[key]), (i, s, key)
    # exact searches on the four views
    for v in range(4):
        for d in (0, 1):
            both = rd.arr(RESULT_DTYPE, 2 * nreads).reshape(nreads, 2)
            wr, wa = g[f"ex_v{v}_d{d}"], g[f"ex_anchor_v{v}_d{d}"]
            for key in ("k", "l", "pos"):
                assert np.array_equal(both[:, 0][key], wr[key]), (v, d, key)
            for key in ("k", "l", "start", "pos", "end", "dir"):
                assert np.array_equal(both[:, 1][key], wa[key]), (v, d, "anchor", key)
    # get_SA / get_ISA
    sa = rd.arr(np.uint64, 4 * len(rows)).reshape(-1, 4)
    for j, key in enumerate(("sa_bwd", "sa_fwd", "isa_bwd", "isa_fwd")):
        assert np.array_equal(sa[:, j], g[key].astype(np.uint64)), key
    # BWSearch1VectorHelper on whole reads, then windows
    for prefix in ("s1v", "s1w"):
        wv = split_golden_hits(g, prefix)
        for i in range(nreads):
            for s in (0, 1):
                full = rd.u32()
                lst = rd.results()
                assert full == 0, (prefix, i, s)
                assert len(lst) == int(g[f"{prefix}_counts"][i, s]), (prefix, i, s)
                assert fo.hits_equal(as_hits(lst, s), wv[(i, s)]), (prefix, i, s)
    # BWSearchCPU with anchors
    fin = split_k(g, "sk_hits", "sk_counts")
    anc = split_k(g, "sk_anchors", "sk_anchor_counts")
    for i in range(nreads):
        for s in (0, 1):
            full = rd.u32()
            lst = rd.results()
            a = rd.results()
            assert full == int(g["sk_full"][i, s]), (i, s)
            assert fo.hitsk_equal(as_hitsk(lst), fin[(i, s)]), (i, s)
            assert fo.hitsk_equal(as_hitsk(a), anc[(i, s)]), (i, s)
    # binsearch: entries <= key of [0, 10, 10, 20, 35]; NULL gives 0
    assert list(rd.arr(np.uint64, 6)) == [1, 1, 3, 4, 5, 0]
    assert rd.o == len(rd.b)
