"""The C-ABI library loads and exports every function include/fmi.h
declares, and rejects bad calls with status codes before touching a GPU."""
import ctypes as C
import os
import re

import numpy as np
import pytest

import hpgfm

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    src = open(os.path.join(ROOT, "include", "fmi.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(fmi_[a-z0-9_]+)\s*\(", src)))


def declared_surface():
    """Functions include/hpg_search.h declares (the reference-named surface)."""
    src = open(os.path.join(ROOT, "include", "hpg_search.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    lines, in_macro = [], False
    for line in src.splitlines():
        if line.lstrip().startswith("#") or in_macro:
            in_macro = line.rstrip().endswith("\\")
            continue
        lines.append(line)
    names = re.findall(r"\b([A-Za-z_][A-Za-z0-9_]*)\s*\(", "\n".join(lines))
    return sorted(set(names))


def test_header_and_binding_agree():
    assert declared_functions() == sorted(hpgfm.EXPORTS)


def test_library_exports_the_reference_surface():
    lib = C.CDLL(hpgfm.LIB_PATH)
    names = declared_surface()
    for want in ("BWExactSearchBackward", "BWExactSearchForward", "BWExactSearchVectorBackward",
                 "BWExactSearchVectorForward", "BWSearch1CPU", "BWSearch1VectorHelper",
                 "BWSearchCPU", "binsearch", "load_bwt_index", "free_bwt_index"):
        assert want in names
    missing = [f for f in names if not hasattr(lib, f)]
    assert not missing, missing


def test_library_exports_every_declared_symbol():
    lib = C.CDLL(hpgfm.LIB_PATH)
    missing = [f for f in declared_functions() if not hasattr(lib, f)]
    assert not missing


def test_status_strings():
    L = hpgfm.lib()
    for code in range(0, -8, -1):
        assert L.fmi_status_str(code) not in (None, b"unknown status")
    assert L.fmi_status_str(-99) == b"unknown status"


def test_calls_before_to_device_are_rejected():
    g = np.random.default_rng(0).integers(0, 4, 1000).astype(np.uint8)
    ix = hpgfm.Index.build(g, 8)
    L = hpgfm.lib()
    rc = L.fmi_exact_batch(ix.h, 0, 0, None, None, 1, None, None, None)
    assert rc == -7                                        # FMI_ESTATE
    rc = L.fmi_search1_batch(ix.h, 1, 3, None, None, 1, 4, None, None, None)
    assert rc == -7
    rc = L.fmi_locate_batch(ix.h, 0, None, None, 1, 1, None, None, None)
    assert rc == -7
    assert b"fmi_index_to_device" in L.fmi_last_error()
    rc = L.fmi_exact_batch(None, 0, 0, None, None, 1, None, None, None)
    assert rc == -1


def test_tuning_range():
    with pytest.raises(hpgfm.FmiError):
        hpgfm.set_tuning(65, 32, 16)
    with pytest.raises(hpgfm.FmiError):
        hpgfm.set_tuning(8, 5000, 16)
    with pytest.raises(hpgfm.FmiError):
        hpgfm.set_tuning(8, 32, 65)
    hpgfm.set_tuning(4, 128, 8)
    assert hpgfm.get_tuning() == (4, 128, 8)
    hpgfm.set_tuning(0, 0, 0)
    wg, chunk, refill = hpgfm.get_tuning()
    assert wg > 0 and chunk > 0 and 1 <= refill <= 64
    with pytest.raises(hpgfm.FmiError):
        hpgfm.set_load_mode(3)
    for m in (hpgfm.LOADS_OWN, hpgfm.LOADS_ALL, hpgfm.LOADS_AUTO):
        hpgfm.set_load_mode(m)
        assert hpgfm.get_load_mode() == m


def test_graph_arguments():
    """The graph entry points reject a missing stream or graph before any
    HIP call (no GPU needed); destroying no graph is a no-op."""
    import ctypes as C
    L = hpgfm.lib()
    assert L.fmi_graph_begin(None) == -1
    g = C.c_void_p()
    assert L.fmi_graph_end(None, C.byref(g)) == -1
    assert L.fmi_graph_launch(None, None) == -1
    assert L.fmi_graph_destroy(None) == 0
    # fork / join: two different created streams, checked before any HIP call
    assert L.fmi_stream_wait(None, None) == -1
    h = C.c_void_p(0x1000)
    assert L.fmi_stream_wait(h, h) == -1


def test_work_count_arguments():
    """The counting calls reject a missing work array, and an index that is
    not on the device, before any HIP call."""
    g = np.random.default_rng(1).integers(0, 4, 1000).astype(np.uint8)
    ix = hpgfm.Index.build(g, 8)
    L = hpgfm.lib()
    assert L.fmi_search_k_count_work(ix.h, 3, None, None, None, 1, 8, 4, None, None, None, None,
                                     None, None, None) == -1
    w = np.zeros(6, np.uint64)
    rc = L.fmi_search_k_count_work(ix.h, 3, None, None, None, 1, 8, 4, None, None, None, None,
                                   None, None, w.ctypes.data)
    assert rc == -7                                        # FMI_ESTATE: not on the device
    rc = L.fmi_exact_count_work(ix.h, 0, 0, None, None, 1, None, None, None, w.ctypes.data)
    assert rc == -7


def test_read_validation():
    reads = np.zeros(10, np.uint8)
    with pytest.raises(ValueError):
        hpgfm.validate_reads(reads, np.array([0, 5, 3], np.uint64))
    with pytest.raises(ValueError):
        hpgfm.validate_reads(reads, np.array([0, 5, 20], np.uint64))
    with pytest.raises(ValueError):
        hpgfm.validate_reads(np.zeros(2000, np.uint8), np.array([0, 1500], np.uint64))
    hpgfm.validate_reads(reads, np.array([0, 5, 10], np.uint64))


def test_encode_bases():
    """encode_bases (bwt_commons.c:99-102) over the table that
    bwt_init_replace_table (bwt_commons.c:58-89) builds for "ACGT": both cases
    of each base map to its index; the bytes the reference leaves unset map
    to 4, the N code of every search entry point."""
    allb = np.arange(256, dtype=np.uint8)
    got = hpgfm.encode_bases(allb)
    want = np.full(256, 4, np.uint8)
    for i, c in enumerate("ACGT"):
        want[ord(c)] = want[ord(c.lower())] = i
    assert np.array_equal(got, want)
    assert list(hpgfm.encode_bases("AcGtNx")) == [0, 1, 2, 3, 4, 4]
    assert len(hpgfm.encode_bases(b"")) == 0
