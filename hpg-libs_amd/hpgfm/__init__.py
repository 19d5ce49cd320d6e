"""hpgfm -- Python binding of libhpgfm.so, the MI355X FM-index search engine.

The engine itself is the C ABI in include/fmi.h (HIP kernels for gfx950);
this module wraps it with ctypes and numpy so that tests, the benchmark and
the serving loop can drive it.  There is no CPU fallback: every search runs
on the GPU, and loading fails loudly when the library is not built.
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libhpgfm.so")

# fmi_hit (include/fmi.h), 32 bytes
HIT_DTYPE = np.dtype([
    ("k", "<i8"), ("l", "<i8"),
    ("start", "<i2"), ("pos", "<i2"), ("end", "<i2"), ("err_pos", "<i2"),
    ("dir", "u1"), ("num_mismatches", "u1"), ("err_kind", "u1"), ("err_base", "u1"),
    ("strand", "u1"), ("pad", "u1", (3,)),
])
assert HIT_DTYPE.itemsize == 32

# fmi_hitk (include/fmi.h), 64 bytes
HITK_DTYPE = np.dtype([
    ("k", "<i8"), ("l", "<i8"), ("start", "<i2"), ("pos", "<i2"), ("end", "<i2"),
    ("dir", "u1"), ("num_mismatches", "u1"), ("err_kind", "u1", (8,)), ("err_base", "u1", (8,)),
    ("err_pos", "<i2", (8,)), ("pad", "u1", (8,)),
])
assert HITK_DTYPE.itemsize == 64

OK = 0
MATCH, DELETION, MISMATCH, INSERTION = 0, 1, 2, 3
VIEW_BACKWARD, VIEW_FORWARD, VIEW_BACKWARD_REV, VIEW_FORWARD_REV = 0, 1, 2, 3
MODE_EDIT, MODE_MISMATCH = 0, 1
STRAND_PLUS, STRAND_MINUS = 1, 2
MAX_READ_LEN = 1000

# Every symbol include/fmi.h declares; tests check the library exports them.
EXPORTS = [
    "fmi_status_str", "fmi_last_error", "fmi_index_build", "fmi_index_build_gpu", "fmi_index_save",
    "fmi_index_load", "fmi_index_import_reference", "fmi_index_free", "fmi_index_size_sa",
    "fmi_index_sa_ratio", "fmi_index_set_contig_names", "fmi_index_contig_count",
    "fmi_index_contig_name",
    "fmi_index_device_bytes", "fmi_index_to_device", "fmi_index_get_bwt", "fmi_index_device_layout",
    "fmi_index_set_karyotype", "fmi_device_count", "fmi_device_alloc", "fmi_device_free",
    "fmi_memcpy_h2d", "fmi_memcpy_d2h", "fmi_memset_d", "fmi_device_mem_info", "fmi_stream_create",
    "fmi_device_name", "fmi_device_cu_count",
    "fmi_stream_destroy", "fmi_stream_sync", "fmi_stream_wait", "fmi_graph_begin", "fmi_graph_end",
    "fmi_graph_launch", "fmi_graph_destroy", "fmi_last_kernel_ms", "fmi_exact_batch",
    "fmi_exact_vector_batch", "fmi_search1_batch", "fmi_search1_batch_plain",
    "fmi_search1_anchors_batch", "fmi_search1_batch_noalloc",
    "fmi_set_tuning", "fmi_get_tuning", "fmi_set_load_mode", "fmi_get_load_mode", "fmi_set_staging", "fmi_get_staging", "fmi_locate_batch", "fmi_locate_hits", "fmi_resolve_hits", "fmi_resolve_hits_text",
    "fmi_gather_bench", "fmi_gather_probe",
    "fmi_search1_count_blocks", "fmi_timing_begin", "fmi_timing_end",
    "fmi_seed_map_batch", "fmi_map_inexact_seeds_batch", "fmi_map_reads_batch", "fmi_format_results", "fmi_inverse_locate_batch", "fmi_encode_bases", "fmi_copy_bench",
    "fmi_search1_vector_batch", "fmi_index_expand", "fmi_index_kmer_table", "fmi_index_get_kmer_table",
    "fmi_index_get_full", "fmi_index_sample_count", "fmi_index_get_samples",
    "fmi_search1_count_work", "fmi_search_k_count_work", "fmi_exact_count_work", "fmi_search_k_batch", "fmi_plain_create", "fmi_plain_search1", "fmi_plain_words",
    "fmi_plain_get_layout", "fmi_plain_device_bytes", "fmi_plain_free",
    "fmi_server_create", "fmi_server_submit", "fmi_server_wait", "fmi_server_fetch",
    "fmi_server_cancel", "fmi_server_submit_fastq", "fmi_fastq_encode_batch", "fmi_pack_codes",
    "fmi_server_get_stats", "fmi_server_destroy",
]


class ServerConfig(C.Structure):
    _fields_ = [("device_budget_bytes", C.c_uint64), ("max_inflight", C.c_uint32),
                ("max_batch_reads", C.c_uint32), ("cap", C.c_uint32), ("mode", C.c_int),
                ("strand_mask", C.c_int), ("max_pending", C.c_uint32),
                ("staging_bytes", C.c_uint64)]


class ServerStats(C.Structure):
    _fields_ = [("requests_done", C.c_uint64), ("reads_done", C.c_uint64),
                ("launches", C.c_uint64), ("max_batch_reads_seen", C.c_uint64),
                ("max_inflight_seen", C.c_uint32), ("pending", C.c_uint32),
                ("refused", C.c_uint64), ("cancelled", C.c_uint64)]


EAGAIN, ECANCELED = -8, -9


class FmiError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"fmi error {code}: {msg}")
        self.code = code


_p = C.c_void_p
_u32 = C.c_uint32
_u64 = C.c_uint64
_i = C.c_int
_LIB = None


def _bind(lib):
    sig = {
        "fmi_status_str": ([_i], C.c_char_p),
        "fmi_last_error": ([], C.c_char_p),
        "fmi_index_build": ([_p, _u64, _u32, C.POINTER(_p)], _i),
        "fmi_index_build_gpu": ([_p, _u64, _u32, C.POINTER(_p)], _i),
        "fmi_index_save": ([_p, C.c_char_p], _i),
        "fmi_index_load": ([C.c_char_p, C.POINTER(_p)], _i),
        "fmi_index_import_reference": ([C.c_char_p, C.POINTER(_p)], _i),
        "fmi_index_set_contig_names": ([_p, _u64, _p], _i),
        "fmi_index_contig_count": ([_p], _u64),
        "fmi_index_contig_name": ([_p, _u64], C.c_char_p),
        "fmi_index_free": ([_p], None),
        "fmi_index_size_sa": ([_p], _u64),
        "fmi_index_sa_ratio": ([_p], _u32),
        "fmi_index_device_bytes": ([_p], _u64),
        "fmi_index_device_layout": ([_p, _p], _i),
        "fmi_index_to_device": ([_p], _i),
        "fmi_index_get_bwt": ([_p, _i, _p, C.POINTER(_u64)], _i),
        "fmi_index_set_karyotype": ([_p, _u64, _p, _p], _i),
        "fmi_device_count": ([C.POINTER(_i)], _i),
        "fmi_device_alloc": ([C.POINTER(_p), C.c_size_t], _i),
        "fmi_device_free": ([_p], _i),
        "fmi_memcpy_h2d": ([_p, _p, C.c_size_t, _p], _i),
        "fmi_memcpy_d2h": ([_p, _p, C.c_size_t, _p], _i),
        "fmi_memset_d": ([_p, _i, C.c_size_t, _p], _i),
        "fmi_device_mem_info": ([C.POINTER(_u64), C.POINTER(_u64)], _i),
        "fmi_device_name": ([C.c_char_p, C.c_size_t], _i),
        "fmi_device_cu_count": ([C.POINTER(_i)], _i),
        "fmi_stream_create": ([C.POINTER(_p)], _i),
        "fmi_stream_destroy": ([_p], _i),
        "fmi_stream_sync": ([_p], _i),
        "fmi_graph_begin": ([C.c_void_p], _i),
        "fmi_stream_wait": ([C.c_void_p, C.c_void_p], _i),
        "fmi_graph_end": ([C.c_void_p, C.POINTER(C.c_void_p)], _i),
        "fmi_graph_launch": ([C.c_void_p, C.c_void_p], _i),
        "fmi_graph_destroy": ([C.c_void_p], _i),
        "fmi_last_kernel_ms": ([C.POINTER(C.c_float)], _i),
        "fmi_timing_begin": ([], _i),
        "fmi_timing_end": ([C.POINTER(_u32), C.POINTER(C.c_double), C.POINTER(C.c_float),
                            C.POINTER(C.c_float)], _i),
        "fmi_exact_batch": ([_p, _i, _i, _p, _p, _u32, _p, _p, _p], _i),
        "fmi_exact_vector_batch": ([_p, _i, _i, _p, _p, _u32, _p, _p, _p, _p], _i),
        "fmi_search1_batch": ([_p, _i, _i, _p, _p, _u32, _u32, _p, _p, _p], _i),
        "fmi_search1_batch_plain": ([_p, _i, _i, _p, _p, _u32, _u32, _p, _p, _p], _i),
        "fmi_search1_anchors_batch": ([_p, _i, _p, _p, _u32, _p, _p], _i),
        "fmi_search1_batch_noalloc": ([_p, _i, _i, _p, _p, _u32, _u32, _p, _p, _p], _i),
        "fmi_set_tuning": ([_u32, _u32, _u32], _i),
        "fmi_get_tuning": ([C.POINTER(_u32), C.POINTER(_u32), C.POINTER(_u32)], _i),
        "fmi_set_load_mode": ([_u32], _i),
        "fmi_get_load_mode": ([C.POINTER(_u32)], _i),
        "fmi_set_staging": ([_i], _i),
        "fmi_get_staging": ([C.POINTER(_i)], _i),
        "fmi_search_k_count_work": ([_p, _i, _p, _p, _p, _u32, _u32, _u32, _p, _p, _p, _p, _p, _p,
                                     _p], _i),
        "fmi_exact_count_work": ([_p, _i, _i, _p, _p, _u32, _p, _p, _p, _p], _i),
        "fmi_locate_batch": ([_p, _i, _p, _p, _u32, _u32, _p, _p, _p], _i),
        "fmi_locate_hits": ([_p, _i, _p, _u32, _u32, _p, _p, _p], _i),
        "fmi_resolve_hits": ([_p, _p, _p, _p, _u32, _u32, _u32, _p, _p, _p], _i),
        "fmi_resolve_hits_text": ([_p, _p, _p, _p, _u32, _u32, _u32, _p, _p, _p], _i),
        "fmi_inverse_locate_batch": ([_p, _i, _p, _u64, _p, _p], _i),
        "fmi_encode_bases": ([_p, _u64, _p], _i),
        "fmi_copy_bench": ([_p, _p, _u64, C.POINTER(C.c_float), _p], _i),
        "fmi_search1_vector_batch": ([_p, _i, _i, _p, _p, _p, _u32, _u32, _p, _p, _p], _i),
        "fmi_index_expand": ([_p], _i),
        "fmi_index_kmer_table": ([_p, _u32], _i),
        "fmi_index_get_kmer_table": ([_p, _i, _p], _i),
        "fmi_search1_count_work": ([_p, _i, _i, _p, _p, _u32, _u32, _p], _i),
        "fmi_search_k_batch": ([_p, _i, _p, _p, _p, _u32, _u32, _u32, _p, _p, _p, _p, _p, _p],
                               _i),
        "fmi_index_get_full": ([_p, _i, _p, _p, _p], _i),
        "fmi_index_sample_count": ([_p], _u64),
        "fmi_index_get_samples": ([_p, _i, _p, _p], _i),
        "fmi_plain_create": ([_p, C.POINTER(_p)], _i),
        "fmi_plain_search1": ([_p, _i, _p, _p, _u32, _u32, _p, _p, C.POINTER(C.c_double),
                               C.POINTER(C.c_double)], _i),
        "fmi_plain_words": ([_p], _u64),
        "fmi_plain_get_layout": ([_p, _i, _p, _p], _i),
        "fmi_plain_device_bytes": ([_p], _u64),
        "fmi_plain_free": ([_p], _i),
        "fmi_gather_bench": ([_p, _u64, _u32, _u64, _p, C.POINTER(C.c_float), C.POINTER(_u64),
                              _p], _i),
        "fmi_gather_probe": ([_p, _u64, _u32, _u32, _i, _u64, _p, C.POINTER(C.c_float),
                              C.POINTER(_u64), _p], _i),
        "fmi_search1_count_blocks": ([_p, _i, _i, _p, _p, _u32, C.POINTER(_u64),
                                      C.POINTER(_u64)], _i),
        "fmi_format_results": ([_p, _u64, _i, C.c_char_p, _u32, _p, _u32, _p, _p, _u32, _p,
                                C.c_size_t, C.POINTER(C.c_size_t), C.POINTER(_u64)], _i),
        "fmi_map_inexact_seeds_batch": ([_p, _p, _p, _u32, _u32, _u32, _u32, _u32, _u64, _p,
                                         _p, _p, C.POINTER(_u64), _p], _i),
        "fmi_map_reads_batch": ([_p, _p, _p, _u32, _u32, _u32, _p, _p, _p, _p], _i),
        "fmi_seed_map_batch": ([_p, _p, _p, _u32, _u32, _u32, _u32, _p, _u64, _p, _u64,
                                C.POINTER(_u64), C.POINTER(_u64), _p], _i),
        "fmi_server_create": ([_p, C.POINTER(ServerConfig), C.POINTER(_p)], _i),
        "fmi_server_submit": ([_p, _p, _p, _u32, C.POINTER(_u64)], _i),
        "fmi_server_wait": ([_p, _u64, C.c_int64, C.POINTER(C.c_double), C.POINTER(_u64)], _i),
        "fmi_server_fetch": ([_p, _u64, _p, _p], _i),
        "fmi_server_cancel": ([_p, _u64], _i),
        "fmi_server_submit_fastq": ([_p, _p, _p, _u32, C.POINTER(_u64)], _i),
        "fmi_fastq_encode_batch": ([_p, _p, _u32, _u64, _p, _p, _p], _i),
        "fmi_pack_codes": ([_p, _u64, _p, _p, _p], _i),
        "fmi_server_get_stats": ([_p, C.POINTER(ServerStats)], _i),
        "fmi_server_destroy": ([_p], _i),
    }
    for name, (args, res) in sig.items():
        f = getattr(lib, name)
        f.argtypes = args
        f.restype = res
    return lib


def lib():
    """The loaded engine library; raises when it is not built."""
    global _LIB
    if _LIB is None:
        if not os.path.exists(LIB_PATH):
            raise FmiError(-5, f"{LIB_PATH} is not built (python -c 'import __graft_entry__ as g; g.build()')")
        _LIB = _bind(C.CDLL(LIB_PATH))
    return _LIB


def check(rc):
    if rc != OK:
        L = lib()
        raise FmiError(rc, f"{L.fmi_status_str(rc).decode()}: {L.fmi_last_error().decode()}")
    return rc


def _ptr(a):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


def mem_info():
    """(free, total) bytes of the current device."""
    f, t = C.c_uint64(), C.c_uint64()
    check(lib().fmi_device_mem_info(C.byref(f), C.byref(t)))
    return int(f.value), int(t.value)


def device_name():
    """Marketing name of the current device as HIP reports it."""
    buf = C.create_string_buffer(256)
    check(lib().fmi_device_name(buf, 256))
    return buf.value.decode(errors="replace")


def cu_count():
    n = C.c_int(0)
    check(lib().fmi_device_cu_count(C.byref(n)))
    return int(n.value)


def device_count():
    n = C.c_int(0)
    rc = lib().fmi_device_count(C.byref(n))
    return n.value if rc == OK else 0


class DeviceBuffer:
    """A device allocation owned by Python."""

    def __init__(self, nbytes):
        self.nbytes = int(nbytes)
        p = C.c_void_p()
        check(lib().fmi_device_alloc(C.byref(p), max(self.nbytes, 1)))
        self.ptr = p

    @classmethod
    def from_array(cls, a, stream=None):
        a = np.ascontiguousarray(a)
        b = cls(a.nbytes)
        b.upload(a, stream)
        return b

    def upload(self, a, stream=None):
        a = np.ascontiguousarray(a)
        assert a.nbytes <= self.nbytes
        check(lib().fmi_memcpy_h2d(self.ptr, _ptr(a), a.nbytes, stream))
        check(lib().fmi_stream_sync(stream))

    def download(self, dtype, count, stream=None):
        out = np.empty(count, dtype=dtype)
        assert out.nbytes <= self.nbytes
        check(lib().fmi_memcpy_d2h(_ptr(out), self.ptr, out.nbytes, stream))
        check(lib().fmi_stream_sync(stream))
        return out

    def zero(self, stream=None):
        check(lib().fmi_memset_d(self.ptr, 0, self.nbytes, stream))

    def free(self):
        if self.ptr is not None and self.ptr.value:
            lib().fmi_device_free(self.ptr)
        self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class Stream:
    def __init__(self):
        s = C.c_void_p()
        check(lib().fmi_stream_create(C.byref(s)))
        self.handle = s

    def sync(self):
        check(lib().fmi_stream_sync(self.handle))

    def wait(self, other):
        """fmi_stream_wait: this stream waits for the work issued so far on
        other; during other's capture this stream joins it (a fork), and
        other.wait(this) joins it back."""
        check(lib().fmi_stream_wait(self.handle, other.handle))

    def close(self):
        """Drains the stream and frees it with the library's counters and
        scratch for it; raises (FMI_ESTATE) while a graph captured on it
        lives."""
        if self.handle is not None:
            check(lib().fmi_stream_destroy(self.handle))
            self.handle = None


def sync(stream=None):
    check(lib().fmi_stream_sync(stream))


class Graph:
    """A chain of launches captured once on a Stream and replayed with one
    call (fmi_graph_*):

        with hpgfm.Graph(stream) as g:
            hpgfm.search1_launch(ix, batch, out, stream=stream.handle)
        g.launch()            # the whole chain again, on the same buffers
    """

    def __init__(self, stream):
        self.stream = stream
        self.handle = None

    def __enter__(self):
        check(lib().fmi_graph_begin(self.stream.handle))
        return self

    def __exit__(self, *exc):
        h = C.c_void_p()
        rc = lib().fmi_graph_end(self.stream.handle, C.byref(h))
        if exc[0] is None:
            check(rc)
            self.handle = h
        return False

    def launch(self, stream=None):
        check(lib().fmi_graph_launch(self.handle, (stream or self.stream).handle))

    def close(self):
        if self.handle is not None:
            lib().fmi_graph_destroy(self.handle)
            self.handle = None


def last_kernel_ms():
    ms = C.c_float(0)
    check(lib().fmi_last_kernel_ms(C.byref(ms)))
    return float(ms.value)


class TimingRegion:
    """Per-launch HIP-event timing of every kernel launched inside the
    `with` block; after it: .launches, .sum_ms, .min_ms, .max_ms."""

    def __enter__(self):
        check(lib().fmi_timing_begin())
        return self

    def __exit__(self, *exc):
        n = C.c_uint32()
        s = C.c_double()
        mn = C.c_float()
        mx = C.c_float()
        check(lib().fmi_timing_end(C.byref(n), C.byref(s), C.byref(mn), C.byref(mx)))
        self.launches, self.sum_ms = int(n.value), float(s.value)
        self.min_ms, self.max_ms = float(mn.value), float(mx.value)
        return False

    @property
    def mean_ms(self):
        return self.sum_ms / max(self.launches, 1)


def set_tuning(wg_per_cu=0, chunk=0, refill_min=0):
    check(lib().fmi_set_tuning(wg_per_cu, chunk, refill_min))


def get_tuning():
    a, b, c = C.c_uint32(), C.c_uint32(), C.c_uint32()
    check(lib().fmi_get_tuning(C.byref(a), C.byref(b), C.byref(c)))
    return int(a.value), int(b.value), int(c.value)


LOADS_AUTO, LOADS_ALL, LOADS_OWN = 0, 1, 2


def set_load_mode(mode):
    """fmi_set_load_mode: LOADS_AUTO, LOADS_ALL or LOADS_OWN."""
    check(lib().fmi_set_load_mode(mode))


def get_load_mode():
    m = C.c_uint32()
    check(lib().fmi_get_load_mode(C.byref(m)))
    return int(m.value)


def set_staging(enable):
    """fmi_set_staging: 1 (the default) holds each lane's read in LDS during
    the persistent 1-error search, 0 reads it from memory at every step."""
    check(lib().fmi_set_staging(1 if enable else 0))


def get_staging():
    v = C.c_int()
    check(lib().fmi_get_staging(C.byref(v)))
    return bool(v.value)


class Index:
    """A two-direction FM index (host arrays, plus an HBM copy once on device)."""

    def __init__(self, handle):
        self.h = handle

    @classmethod
    def build(cls, text, sa_ratio=16):
        text = np.ascontiguousarray(text, dtype=np.uint8)
        h = C.c_void_p()
        check(lib().fmi_index_build(_ptr(text), len(text), sa_ratio, C.byref(h)))
        return cls(h)

    @classmethod
    def build_gpu(cls, text, sa_ratio=16):
        """Builds on the current GPU; the index is left on the device."""
        text = np.ascontiguousarray(text, dtype=np.uint8)
        h = C.c_void_p()
        check(lib().fmi_index_build_gpu(_ptr(text), len(text), sa_ratio, C.byref(h)))
        return cls(h)

    @classmethod
    def load(cls, path):
        h = C.c_void_p()
        check(lib().fmi_index_load(os.fsencode(path), C.byref(h)))
        return cls(h)

    @classmethod
    def import_reference(cls, directory):
        """fmi_index_import_reference: the index directory the reference's
        builder writes (C, C1, O/Oi, S/R, Si/Ri, config.txt, index)."""
        h = C.c_void_p()
        check(lib().fmi_index_import_reference(os.fsencode(directory), C.byref(h)))
        return cls(h)

    def save(self, path):
        check(lib().fmi_index_save(self.h, os.fsencode(path)))

    def contig_names(self):
        L = lib()
        return [L.fmi_index_contig_name(self.h, c).decode()
                for c in range(int(L.fmi_index_contig_count(self.h)))]

    def set_contig_names(self, names):
        arr = (C.c_char_p * max(len(names), 1))(*[x.encode() for x in names])
        check(lib().fmi_index_set_contig_names(self.h, len(names), arr))

    def close(self):
        if self.h is not None:
            lib().fmi_index_free(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def size_sa(self):
        return int(lib().fmi_index_size_sa(self.h))

    @property
    def n(self):
        return self.size_sa - 1

    @property
    def sa_ratio(self):
        return int(lib().fmi_index_sa_ratio(self.h))

    @property
    def device_bytes(self):
        return int(lib().fmi_index_device_bytes(self.h))

    def to_device(self):
        check(lib().fmi_index_to_device(self.h))
        return self

    def expand(self):
        """fmi_index_expand: full SA / ISA of both directions and the 2-bit
        text on the device, for the text-comparison continuation."""
        check(lib().fmi_index_expand(self.h))
        return self

    def kmer_table(self, k):
        """fmi_index_kmer_table: the interval of every k-mer of both
        directions on the device (k = 0 frees it)."""
        check(lib().fmi_index_kmer_table(self.h, k))
        self._kmer_k = k
        return self

    def kmer_entries(self, forward=0):
        """One direction's k-mer table on the host (uint64 k | l << 32)."""
        k = self._kmer_k
        out = np.empty(4 ** k, np.uint64)
        check(lib().fmi_index_get_kmer_table(self.h, forward, _ptr(out)))
        return out

    def full_arrays(self, forward=0):
        """(sa, isa, text words) of one direction after expand()."""
        n1 = self.n + 1
        sa = np.empty(n1, np.uint32)
        isa = np.empty(n1, np.uint32)
        text = np.empty(max((self.n + 15) // 16, 1), np.uint32)
        check(lib().fmi_index_get_full(self.h, forward, _ptr(sa), _ptr(isa), _ptr(text)))
        return sa, isa, text[:(self.n + 15) // 16]

    def layout_bytes(self):
        """Device bytes by part (fmi_index_device_layout)."""
        a = np.zeros(3, np.uint64)
        check(lib().fmi_index_device_layout(self.h, _ptr(a)))
        return {"blocks_and_samples": int(a[0]), "sa_isa_text": int(a[1]),
                "kmer_tables": int(a[2])}

    def samples(self, forward=0):
        """(S, R): the sampled SA and ISA of one direction (S.vec / R.vec)."""
        m = int(lib().fmi_index_sample_count(self.h))
        S = np.empty(m, np.uint32)
        R = np.empty(m, np.uint32)
        check(lib().fmi_index_get_samples(self.h, forward, _ptr(S), _ptr(R)))
        return S, R

    def bwt(self, forward=0):
        out = np.empty(max(self.n, 1), np.uint8)
        d = C.c_uint64()
        check(lib().fmi_index_get_bwt(self.h, forward, _ptr(out), C.byref(d)))
        return out[:self.n], int(d.value)

    def set_karyotype(self, start, end):
        start = np.ascontiguousarray(start, dtype=np.uint64)
        end = np.ascontiguousarray(end, dtype=np.uint64)
        check(lib().fmi_index_set_karyotype(self.h, len(start), _ptr(start), _ptr(end)))


def validate_reads(reads, offsets):
    offsets = np.ascontiguousarray(offsets, dtype=np.uint64)
    if len(offsets) < 1 or offsets[0] != 0:
        raise ValueError("offsets must start at 0")
    lens = np.diff(offsets.astype(np.int64))
    if np.any(lens < 0) or np.any(lens > MAX_READ_LEN):
        raise ValueError(f"read lengths must be in 0..{MAX_READ_LEN}")
    if int(offsets[-1]) > len(reads):
        raise ValueError("offsets run past the reads buffer")
    return offsets


class ReadBatch:
    """Reads resident in HBM (codes as uint8, concatenated, with offsets)."""

    def __init__(self, reads, offsets, stream=None):
        reads = np.ascontiguousarray(reads, dtype=np.uint8)
        offsets = validate_reads(reads, offsets)
        self.nreads = len(offsets) - 1
        self.total = int(offsets[-1])
        self.d_reads = DeviceBuffer.from_array(reads if len(reads) else np.zeros(1, np.uint8), stream)
        self.d_offsets = DeviceBuffer.from_array(offsets, stream)
        self.offsets = offsets
        self.lengths = np.diff(offsets.astype(np.int64))

    def free(self):
        self.d_reads.free()
        self.d_offsets.free()


def exact_batch(ix, batch, view, direction, anchors=False, stream=None):
    """fmi_exact_batch; returns (records, anchors or None) as host arrays."""
    n = batch.nreads
    out = DeviceBuffer(max(n, 1) * 32)
    anc = DeviceBuffer(max(n, 1) * 32) if anchors else None
    check(lib().fmi_exact_batch(ix.h, view, direction, batch.d_reads.ptr, batch.d_offsets.ptr, n,
                                out.ptr, anc.ptr if anc else None, stream))
    sync(stream)
    r = out.download(HIT_DTYPE, n, stream)
    a = anc.download(HIT_DTYPE, n, stream) if anc else None
    return r, a


def exact_count_work(ix, batch, view, direction):
    """fmi_exact_count_work: the exact search of fmi_exact_batch with its
    work counted; returns ((blocks, LF queries, SA / ISA reads, text
    compares, k-mer table lookups), records)."""
    n = batch.nreads
    out = DeviceBuffer(max(n, 1) * 32)
    w = np.zeros(5, np.uint64)
    check(lib().fmi_exact_count_work(ix.h, view, direction, batch.d_reads.ptr, batch.d_offsets.ptr,
                                     n, out.ptr, None, None, _ptr(w)))
    r = out.download(HIT_DTYPE, n)
    out.free()
    return tuple(int(x) for x in w), r


def exact_vector_batch(ix, batch, view, direction, stream=None):
    n = batch.nreads
    tot = max(batch.total, 1)
    vk = DeviceBuffer(tot * 8)
    vl = DeviceBuffer(tot * 8)
    lg = DeviceBuffer(max(n, 1) * 2)
    check(lib().fmi_exact_vector_batch(ix.h, view, direction, batch.d_reads.ptr,
                                       batch.d_offsets.ptr, n, vk.ptr, vl.ptr, lg.ptr, stream))
    sync(stream)
    return (vk.download(np.int64, batch.total, stream), vl.download(np.int64, batch.total, stream),
            lg.download(np.int16, n, stream))


class Search1Output:
    """Device buffers for fmi_search1_batch results."""

    def __init__(self, nreads, cap):
        self.nreads, self.cap = nreads, cap
        self.d_hits = DeviceBuffer(max(nreads * 2 * cap, 1) * 32)
        self.d_counts = DeviceBuffer(max(nreads * 2, 1) * 4)

    def counts(self, stream=None):
        return self.d_counts.download(np.uint32, self.nreads * 2, stream)

    def hits(self, stream=None):
        return self.d_hits.download(HIT_DTYPE, self.nreads * 2 * self.cap, stream).reshape(
            self.nreads, 2, self.cap)

    def free(self):
        self.d_hits.free()
        self.d_counts.free()


def search1_launch(ix, batch, out, mode=MODE_MISMATCH, strand_mask=3, plain=False, stream=None):
    f = lib().fmi_search1_batch_plain if plain else lib().fmi_search1_batch
    check(f(ix.h, mode, strand_mask, batch.d_reads.ptr, batch.d_offsets.ptr, batch.nreads, out.cap,
            out.d_hits.ptr, out.d_counts.ptr, stream))


def search1_batch(ix, batch, cap=64, mode=MODE_MISMATCH, strand_mask=3, plain=False, stream=None):
    """Returns (counts[n, 2], hits[n, 2, cap]) on the host."""
    out = Search1Output(batch.nreads, cap)
    search1_launch(ix, batch, out, mode, strand_mask, plain, stream)
    sync(stream)
    c = out.counts(stream).reshape(batch.nreads, 2)
    h = out.hits(stream)
    out.free()
    return c, h


def search1_anchors_batch(ix, batch, strand_mask=3, stream=None):
    """fmi_search1_anchors_batch: BWSearch1CPU's two anchors per read and
    strand, as a host array [nreads, 2 strands, 2 passes] of HIT_DTYPE."""
    n = batch.nreads
    d = DeviceBuffer(max(4 * n, 1) * 32)
    d.zero(stream)
    check(lib().fmi_search1_anchors_batch(ix.h, strand_mask, batch.d_reads.ptr,
                                          batch.d_offsets.ptr, n, d.ptr, stream))
    sync(stream)
    out = d.download(HIT_DTYPE, 4 * n, stream).reshape(n, 2, 2)
    d.free()
    return out


def search1_vector_batch(ix, batch, cap=64, mode=MODE_EDIT, strand_mask=3, bounds=None,
                         stream=None):
    """fmi_search1_vector_batch (BWSearch1VectorHelper per read and strand);
    bounds: None or int16 [nreads, 2] windows relative to each read.
    Returns (counts[n, 2], hits[n, 2, cap]) on the host."""
    out = Search1Output(batch.nreads, cap)
    db = None
    if bounds is not None:
        b = np.ascontiguousarray(bounds, dtype=np.int16).reshape(batch.nreads, 2)
        db = DeviceBuffer.from_array(b, stream)
    check(lib().fmi_search1_vector_batch(ix.h, mode, strand_mask, batch.d_reads.ptr,
                                         batch.d_offsets.ptr, db.ptr if db else None,
                                         batch.nreads, cap, out.d_hits.ptr, out.d_counts.ptr,
                                         stream))
    sync(stream)
    c = out.counts(stream).reshape(batch.nreads, 2)
    h = out.hits(stream)
    out.free()
    if db:
        db.free()
    return c, h


class PlainSearch:
    """The reference's GPU design re-created (fmi_plain_*): the comparison
    point of the engine, not a product path."""

    def __init__(self, ix):
        h = C.c_void_p()
        check(lib().fmi_plain_create(ix.h, C.byref(h)))
        self.h = h
        self.words = int(lib().fmi_plain_words(h))
        self.device_bytes = int(lib().fmi_plain_device_bytes(h))

    def search1(self, reads, offsets, cap=32, mode=MODE_MISMATCH):
        """Host in, host out: (counts[n, 2], hits[n, 2, cap], ms_kernel, ms_total)."""
        reads = np.ascontiguousarray(reads, dtype=np.uint8)
        offsets = np.ascontiguousarray(offsets, dtype=np.uint64)
        n = len(offsets) - 1
        counts = np.zeros(max(2 * n, 1), np.uint32)
        hits = np.zeros(max(2 * n * cap, 1), HIT_DTYPE)
        mk, mt = C.c_double(0), C.c_double(0)
        check(lib().fmi_plain_search1(self.h, mode, reads.ctypes.data if len(reads) else None,
                                      offsets.ctypes.data, n, cap, hits.ctypes.data,
                                      counts.ctypes.data, C.byref(mk), C.byref(mt)))
        return (counts[:2 * n].reshape(n, 2), hits[:2 * n * cap].reshape(n, 2, cap),
                float(mk.value), float(mt.value))

    def layout(self, forward):
        """(desp[4, words] uint32, count[4, words] uint64) of one direction."""
        d = np.zeros((4, self.words), np.uint32)
        c = np.zeros((4, self.words), np.uint64)
        check(lib().fmi_plain_get_layout(self.h, forward, d.ctypes.data, c.ctypes.data))
        return d, c

    def close(self):
        if self.h:
            lib().fmi_plain_free(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def search1_count_work(ix, batch, cap=32, mode=MODE_MISMATCH, strand_mask=3):
    """(blocks, LF queries, SA/ISA reads, text steps, k-mer table lookups) of
    one fmi_search1_batch."""
    w = np.zeros(5, np.uint64)
    check(lib().fmi_search1_count_work(ix.h, mode, strand_mask, batch.d_reads.ptr,
                                       batch.d_offsets.ptr, batch.nreads, cap, _ptr(w)))
    return tuple(int(x) for x in w)


class SearchKResult:
    """Host copies of one fmi_search_k_batch: full[n, 2] (1: the anchor list
    filled up, 2: fragment size out of range), final(i, s), anchors(i, s)."""

    def __init__(self, nfin, fin, nanc, anc, full, list_cap, anchor_cap):
        self.nfin, self.fin, self.nanc, self.anc = nfin, fin, nanc, anc
        self.full, self.list_cap, self.anchor_cap = full, list_cap, anchor_cap

    def final(self, i, s):
        t = 2 * i + s
        return self.fin[t * self.list_cap:t * self.list_cap + int(self.nfin[t])]

    def anchors(self, i, s):
        t = 2 * i + s
        return self.anc[t * self.anchor_cap:t * self.anchor_cap + int(self.nanc[t])]


def search_k_launch(ix, batch, d_fragsize, list_cap, anchor_cap, d_fin, d_nfin, d_anc, d_nanc,
                    d_full, strand_mask=3, stream=None):
    check(lib().fmi_search_k_batch(ix.h, strand_mask, batch.d_reads.ptr, batch.d_offsets.ptr,
                                   d_fragsize.ptr, batch.nreads, list_cap, anchor_cap, d_fin.ptr,
                                   d_nfin.ptr, d_anc.ptr if d_anc else None,
                                   d_nanc.ptr if d_nanc else None, d_full.ptr, stream))


def search_k_batch(ix, batch, fragsize, list_cap=300, anchor_cap=16, strand_mask=3, stream=None,
                   work=None):
    """fmi_search_k_batch (BWSearchCPU per read and strand); fragsize: int16
    per read.  Returns a SearchKResult on the host.  With work (a uint64
    array of 6) the search runs as fmi_search_k_count_work and adds its
    (blocks, LF queries, SA / ISA reads, text compares, k-mer table lookups,
    list records written or read) to work."""
    n = batch.nreads
    fs = DeviceBuffer.from_array(np.ascontiguousarray(fragsize, dtype=np.int16), stream)
    d_fin = DeviceBuffer(max(2 * n * list_cap, 1) * 64)
    d_nfin = DeviceBuffer(max(2 * n, 1) * 4)
    d_anc = DeviceBuffer(max(2 * n * anchor_cap, 1) * 64)
    d_nanc = DeviceBuffer(max(2 * n, 1) * 4)
    d_full = DeviceBuffer(max(2 * n, 1))
    if work is None:
        search_k_launch(ix, batch, fs, list_cap, anchor_cap, d_fin, d_nfin, d_anc, d_nanc, d_full,
                        strand_mask, stream)
    else:
        w = np.zeros(6, np.uint64)
        check(lib().fmi_search_k_count_work(ix.h, strand_mask, batch.d_reads.ptr,
                                            batch.d_offsets.ptr, fs.ptr, n, list_cap, anchor_cap,
                                            d_fin.ptr, d_nfin.ptr, d_anc.ptr, d_nanc.ptr,
                                            d_full.ptr, stream, _ptr(w)))
        work += w
    sync(stream)
    res = SearchKResult(d_nfin.download(np.uint32, 2 * n, stream),
                        d_fin.download(HITK_DTYPE, 2 * n * list_cap, stream),
                        d_nanc.download(np.uint32, 2 * n, stream),
                        d_anc.download(HITK_DTYPE, 2 * n * anchor_cap, stream),
                        d_full.download(np.uint8, 2 * n, stream).reshape(n, 2),
                        list_cap, anchor_cap)
    for b in (fs, d_fin, d_nfin, d_anc, d_nanc, d_full):
        b.free()
    return res


def search1_count_blocks(ix, batch, mode=MODE_MISMATCH, strand_mask=3):
    b = C.c_uint64()
    s = C.c_uint64()
    check(lib().fmi_search1_count_blocks(ix.h, mode, strand_mask, batch.d_reads.ptr,
                                         batch.d_offsets.ptr, batch.nreads, C.byref(b), C.byref(s)))
    return int(b.value), int(s.value)


def locate_batch(ix, k, l, cap, forward=0, stream=None):
    k = np.ascontiguousarray(k, dtype=np.int64)
    l = np.ascontiguousarray(l, dtype=np.int64)
    n = len(k)
    dk = DeviceBuffer.from_array(k if n else np.zeros(1, np.int64), stream)
    dl = DeviceBuffer.from_array(l if n else np.zeros(1, np.int64), stream)
    dp = DeviceBuffer(max(n * cap, 1) * 8)
    dn = DeviceBuffer(max(n, 1) * 4)
    check(lib().fmi_locate_batch(ix.h, forward, dk.ptr, dl.ptr, n, cap, dp.ptr, dn.ptr, stream))
    sync(stream)
    return dp.download(np.uint64, n * cap, stream).reshape(n, cap), dn.download(np.uint32, n, stream)


def locate_hits(ix, d_hits, count, cap, forward=0, stream=None):
    """fmi_locate_hits over `count` records already on the device (a
    DeviceBuffer of fmi_hit); returns (positions[count, cap], rows[count])."""
    dp = DeviceBuffer(max(count * cap, 1) * 8)
    dn = DeviceBuffer(max(count, 1) * 4)
    check(lib().fmi_locate_hits(ix.h, forward, d_hits.ptr, count, cap, dp.ptr, dn.ptr, stream))
    sync(stream)
    out = dp.download(np.uint64, count * cap, stream).reshape(count, cap), \
        dn.download(np.uint32, count, stream)
    dp.free()
    dn.free()
    return out


def inverse_locate_batch(ix, text_pos, forward=0, stream=None):
    """fmi_inverse_locate_batch: the BWT row of each text position (get_ISA)."""
    t = np.ascontiguousarray(text_pos, dtype=np.uint64)
    n = len(t)
    if n == 0:
        return np.zeros(0, np.uint64)
    dt = DeviceBuffer.from_array(t, stream)
    dr = DeviceBuffer(n * 8)
    check(lib().fmi_inverse_locate_batch(ix.h, forward, dt.ptr, n, dr.ptr, stream))
    sync(stream)
    return dr.download(np.uint64, n, stream)


def format_results(ix, read_index, strand, mapping, hits, chrom, coord):
    """fmi_format_results for one (read, strand) list: hits (HIT_DTYPE),
    chrom / coord [len(hits), row_cap] as resolve_hits gives them for the
    list.  Returns (text, rows beyond row_cap)."""
    hits = np.ascontiguousarray(hits, dtype=HIT_DTYPE)
    chrom = np.ascontiguousarray(chrom, dtype=np.int32)
    coord = np.ascontiguousarray(coord, dtype=np.uint64)
    row_cap = chrom.shape[1] if chrom.ndim == 2 and chrom.shape[1] else 1
    mapping = bytes(mapping)
    n = C.c_size_t()
    miss = C.c_uint64()
    L = lib()
    check(L.fmi_format_results(ix.h, read_index, strand, mapping, len(mapping), _ptr(hits),
                               len(hits), _ptr(chrom), _ptr(coord), row_cap, None, 0,
                               C.byref(n), C.byref(miss)))
    buf = C.create_string_buffer(n.value + 1)
    check(L.fmi_format_results(ix.h, read_index, strand, mapping, len(mapping), _ptr(hits),
                               len(hits), _ptr(chrom), _ptr(coord), row_cap, buf, n.value + 1,
                               C.byref(n), C.byref(miss)))
    return buf.raw[:n.value].decode(), int(miss.value)


def encode_bases(text):
    """fmi_encode_bases: ASCII bases (str, bytes or uint8 array) to codes
    0..3 for A/C/G/T in either case and 4 for every other byte."""
    if isinstance(text, str):
        text = text.encode("latin-1")
    src = np.frombuffer(bytes(text), np.uint8) if not isinstance(text, np.ndarray) else \
        np.ascontiguousarray(text, dtype=np.uint8)
    dst = np.empty(len(src), np.uint8)
    if len(src):
        check(lib().fmi_encode_bases(src.ctypes.data, len(src), dst.ctypes.data))
    return dst


def resolve_hits(ix, batch, out, row_cap, text_rule=False, stream=None):
    """fmi_resolve_hits (or fmi_resolve_hits_text with text_rule): chrom and
    coord [nreads, 2, cap, row_cap] on the host."""
    n, cap = batch.nreads, out.cap
    tot = max(2 * n * cap * row_cap, 1)
    dc = DeviceBuffer(tot * 4)
    dd = DeviceBuffer(tot * 8)
    fn = lib().fmi_resolve_hits_text if text_rule else lib().fmi_resolve_hits
    check(fn(ix.h, out.d_hits.ptr, out.d_counts.ptr, batch.d_offsets.ptr, n, cap, row_cap, dc.ptr,
             dd.ptr, stream))
    sync(stream)
    shape = (n, 2, cap, row_cap)
    res = (dc.download(np.int32, 2 * n * cap * row_cap, stream).reshape(shape),
           dd.download(np.uint64, 2 * n * cap * row_cap, stream).reshape(shape))
    dc.free()
    dd.free()
    return res


SEED_REGION_DTYPE = np.dtype([
    ("start", "<u8"), ("end", "<u8"), ("chrom_idx", "<u4"), ("read", "<u4"),
    ("seq_start", "<u2"), ("seq_end", "<u2"), ("seed_id", "<u2"), ("strand", "u1"),
    ("valid", "u1"),
])
SEED_INFO_DTYPE = np.dtype([
    ("k_plus", "<i8"), ("l_plus", "<i8"), ("k_minus", "<i8"), ("l_minus", "<i8"),
    ("rows_plus", "<u4"), ("rows_minus", "<u4"),
])
assert SEED_REGION_DTYPE.itemsize == 32 and SEED_INFO_DTYPE.itemsize == 40


INEXACT_REGION_DTYPE = np.dtype([
    ("start", "<u8"), ("end", "<u8"), ("chrom_idx", "<u4"), ("read", "<u4"),
    ("seq_start", "<u2"), ("seq_end", "<u2"), ("seq_len", "<u2"), ("seed_id", "<u2"),
    ("pos_err", "<i4"), ("type_err", "<i4"), ("strand", "u1"), ("pad", "u1", (7,)),
])
assert INEXACT_REGION_DTYPE.itemsize == 48


READ_MAPPING_DTYPE = np.dtype([
    ("position", "<u8"), ("chromosome", "<u4"), ("read", "<u4"), ("strand", "u1"),
    ("secondary", "u1"), ("cigar_len", "<u2"), ("pad", "<u4"),
])
assert READ_MAPPING_DTYPE.itemsize == 24


def fastq_batch(texts):
    """The reference's fastq_batch_t sequence layout for a list of read
    strings (bytes): (seq with a NUL after each read, int32 data_indices)."""
    seq = b"".join(t + b"\0" for t in texts)
    di = np.concatenate([[0], np.cumsum([len(t) + 1 for t in texts])]).astype(np.int32)
    return np.frombuffer(seq, np.uint8).copy(), di


def fastq_encode_batch(seq, data_indices, stream=None):
    """fmi_fastq_encode_batch on the device: (codes, offsets) on the host."""
    seq = np.ascontiguousarray(seq, dtype=np.uint8)
    di = np.ascontiguousarray(data_indices, dtype=np.int32)
    n = len(di) - 1
    nbytes = int(di[-1] - di[0])
    ds = DeviceBuffer.from_array(seq if len(seq) else np.zeros(1, np.uint8), stream)
    dd = DeviceBuffer.from_array(di, stream)
    dc = DeviceBuffer(max(nbytes - n, 1))
    do = DeviceBuffer((n + 1) * 8)
    check(lib().fmi_fastq_encode_batch(ds.ptr, dd.ptr, n, nbytes, dc.ptr, do.ptr, stream))
    sync(stream)
    res = dc.download(np.uint8, nbytes - n, stream), do.download(np.uint64, n + 1, stream)
    for b in (ds, dd, dc, do):
        b.free()
    return res


def pack_codes(codes, stream=None):
    """fmi_pack_codes: (packed uint32 words, N masks) on the host."""
    codes = np.ascontiguousarray(codes, dtype=np.uint8)
    nw = (len(codes) + 15) // 16
    dc = DeviceBuffer.from_array(codes if len(codes) else np.zeros(1, np.uint8), stream)
    dp = DeviceBuffer(max(nw, 1) * 4)
    dm = DeviceBuffer(max(nw, 1) * 2)
    check(lib().fmi_pack_codes(dc.ptr, len(codes), dp.ptr, dm.ptr, stream))
    sync(stream)
    res = dp.download(np.uint32, nw, stream), dm.download(np.uint16, nw, stream)
    for b in (dc, dp, dm):
        b.free()
    return res


def map_reads_batch(ix, batch, filter_read_mappings, cap, stream=None):
    """fmi_map_reads_batch (__bwt_map_inexact_read's alignments): (status[n],
    counts[n], alignments[n, cap]) on the host."""
    n = batch.nreads
    d_out = DeviceBuffer(max(n * cap, 1) * 24)
    d_cnt = DeviceBuffer(max(n, 1) * 4)
    d_st = DeviceBuffer(max(n, 1))
    check(lib().fmi_map_reads_batch(ix.h, batch.d_reads.ptr, batch.d_offsets.ptr, n,
                                    filter_read_mappings, cap, d_out.ptr, d_cnt.ptr, d_st.ptr,
                                    stream))
    sync(stream)
    res = (d_st.download(np.uint8, n, stream), d_cnt.download(np.uint32, n, stream),
           d_out.download(READ_MAPPING_DTYPE, n * cap, stream).reshape(n, cap))
    for b in (d_out, d_cnt, d_st):
        b.free()
    return res


def windows_of(offsets, seed_size, min_seed_size):
    """Windows per read of bwt_map_inexact_seeds_seq's tiling (bwt.c:3270-3296)."""
    lens = np.diff(np.asarray(offsets, dtype=np.int64))
    return lens // seed_size + (lens % seed_size >= min_seed_size)


class InexactSeeds:
    """Host copies of one fmi_map_inexact_seeds_batch: regions(w, s) of
    window w (batch numbering) and strand s, status[w, s], window_base[r]
    (first window of read r)."""

    def __init__(self, regs, nreg, status, region_cap, window_base):
        self.regs, self.nreg, self.status = regs, nreg, status
        self.region_cap, self.window_base = region_cap, window_base

    def regions(self, w, s):
        t = 2 * w + s
        return self.regs[t * self.region_cap:t * self.region_cap + int(self.nreg[t])]


def map_inexact_seeds_batch(ix, batch, seed_size, min_seed_size, rec_cap=256, region_cap=64,
                            stream=None):
    """fmi_map_inexact_seeds_batch over a ReadBatch; returns InexactSeeds."""
    nw = windows_of(batch.offsets, seed_size, min_seed_size)
    total = int(nw.sum())
    regs = DeviceBuffer(max(2 * total * region_cap, 1) * 48)
    nreg = DeviceBuffer(max(2 * total, 1) * 4)
    st = DeviceBuffer(max(2 * total, 1))
    got = C.c_uint64()
    check(lib().fmi_map_inexact_seeds_batch(ix.h, batch.d_reads.ptr, batch.d_offsets.ptr,
                                            batch.nreads, seed_size, min_seed_size, rec_cap,
                                            region_cap, total, regs.ptr, nreg.ptr, st.ptr,
                                            C.byref(got), stream))
    assert got.value == total
    sync(stream)
    out = InexactSeeds(regs.download(INEXACT_REGION_DTYPE, 2 * total * region_cap, stream),
                       nreg.download(np.uint32, 2 * total, stream),
                       st.download(np.uint8, 2 * total, stream).reshape(-1, 2), region_cap,
                       np.concatenate([[0], np.cumsum(nw)]).astype(np.int64))
    for b in (regs, nreg, st):
        b.free()
    return out


def seed_map_batch(ix, batch, seed_len, stride, filter_seed_mappings=1000, region_cap=None,
                   with_info=True, stream=None):
    """fmi_seed_map_batch; returns (nseeds, nrows, regions, infos) on the host.
    region_cap defaults to enough for every resolved row (a first call
    counts them)."""
    L = lib()
    ns, nr = C.c_uint64(), C.c_uint64()
    if region_cap is None:
        check(L.fmi_seed_map_batch(ix.h, batch.d_reads.ptr, batch.d_offsets.ptr, batch.nreads,
                                   seed_len, stride, filter_seed_mappings, None, 0, None, 0,
                                   C.byref(ns), C.byref(nr), stream))
        region_cap = int(nr.value)
        info_cap = int(ns.value)
    else:
        info_cap = 0
        if with_info:
            check(L.fmi_seed_map_batch(ix.h, batch.d_reads.ptr, batch.d_offsets.ptr, batch.nreads,
                                       seed_len, stride, filter_seed_mappings, None, 0, None, 0,
                                       C.byref(ns), C.byref(nr), stream))
            info_cap = int(ns.value)
    info = DeviceBuffer(max(info_cap, 1) * 40) if with_info else None
    regs = DeviceBuffer(max(region_cap, 1) * 32)
    check(L.fmi_seed_map_batch(ix.h, batch.d_reads.ptr, batch.d_offsets.ptr, batch.nreads,
                               seed_len, stride, filter_seed_mappings,
                               info.ptr if info else None, info_cap if info else 0, regs.ptr,
                               region_cap, C.byref(ns), C.byref(nr), stream))
    nrows = int(nr.value)
    out = regs.download(SEED_REGION_DTYPE, min(nrows, region_cap), stream)
    inf = info.download(SEED_INFO_DTYPE, int(ns.value), stream) if info else None
    regs.free()
    if info:
        info.free()
    return int(ns.value), nrows, out, inf


class Server:
    """Continuous-batching request loop over a device-resident index
    (fmi_server_*).  submit() copies the reads into the server's staging and
    returns a request id (retrying while the server answers FMI_EAGAIN, or
    returning None with block=False); wait(id) returns (counts[n, 2],
    records, latency_ms) where records holds the stored records of every
    (read, strand) slot, slot after slot."""

    def __init__(self, ix, budget_bytes=1 << 30, max_inflight=4, max_batch_reads=1 << 20,
                 cap=16, mode=MODE_MISMATCH, strand_mask=3, max_pending=0, staging_bytes=0):
        self.cfg = ServerConfig(budget_bytes, max_inflight, max_batch_reads, cap, mode,
                                strand_mask, max_pending, staging_bytes)
        h = C.c_void_p()
        check(lib().fmi_server_create(ix.h, C.byref(self.cfg), C.byref(h)))
        self.h = h
        self.ix = ix
        self.cap = cap
        self._n = {}

    def submit(self, reads, offsets, block=True, max_wait_s=60.0):
        import time as _t
        reads = np.ascontiguousarray(reads, dtype=np.uint8)
        offsets = np.ascontiguousarray(offsets, dtype=np.uint64)
        n = len(offsets) - 1
        rid = C.c_uint64()
        t0 = None
        while True:
            rc = lib().fmi_server_submit(self.h, _ptr(reads) if len(reads) else None,
                                         _ptr(offsets), n, C.byref(rid))
            if rc != EAGAIN:
                break
            if not block:
                return None
            t0 = t0 or _t.perf_counter()
            if _t.perf_counter() - t0 > max_wait_s:
                raise FmiError(rc, "server queue stayed full")
            _t.sleep(5e-5)
        check(rc)
        self._n[rid.value] = n
        return rid.value

    def submit_fastq(self, seq, data_indices, block=True, max_wait_s=60.0):
        """A request in the reference's fastq_batch_t layout: seq bytes with a
        NUL after each read, int32 data_indices[n + 1]."""
        import time as _t
        seq = np.ascontiguousarray(np.frombuffer(seq, np.uint8) if isinstance(seq, bytes) else seq,
                                   dtype=np.uint8)
        di = np.ascontiguousarray(data_indices, dtype=np.int32)
        n = len(di) - 1
        rid = C.c_uint64()
        t0 = None
        while True:
            rc = lib().fmi_server_submit_fastq(self.h, _ptr(seq) if len(seq) else None, _ptr(di),
                                               n, C.byref(rid))
            if rc != EAGAIN:
                break
            if not block:
                return None
            t0 = t0 or _t.perf_counter()
            if _t.perf_counter() - t0 > max_wait_s:
                raise FmiError(rc, "server queue stayed full")
            _t.sleep(5e-5)
        check(rc)
        self._n[rid.value] = n
        return rid.value

    def cancel(self, rid):
        check(lib().fmi_server_cancel(self.h, rid))
        self._n.pop(rid, None)

    def wait(self, rid, timeout_us=-1):
        lat = C.c_double()
        nrec = C.c_uint64()
        rc = lib().fmi_server_wait(self.h, rid, timeout_us, C.byref(lat), C.byref(nrec))
        if rc == EAGAIN:
            return None
        check(rc)
        n = self._n.pop(rid)
        counts = np.empty(max(2 * n, 1), np.uint32)
        recs = np.empty(max(int(nrec.value), 1), HIT_DTYPE)
        check(lib().fmi_server_fetch(self.h, rid, _ptr(counts), _ptr(recs)))
        return counts[:2 * n].reshape(n, 2), recs[:int(nrec.value)], lat.value

    def stats(self):
        s = ServerStats()
        check(lib().fmi_server_get_stats(self.h, C.byref(s)))
        return {f: getattr(s, f) for f, _ in ServerStats._fields_}

    def close(self):
        if self.h is not None:
            lib().fmi_server_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def copy_bench(nbytes=2 << 30, reps=3, stream=None):
    """Streaming copy of nbytes between two HBM buffers (fmi_copy_bench);
    returns the mean kernel ms over reps launches after one warm-up."""
    src = DeviceBuffer(nbytes)
    dst = DeviceBuffer(nbytes)
    src.zero(stream)
    ms = C.c_float(0)
    total = 0.0
    for r in range(reps + 1):
        check(lib().fmi_copy_bench(src.ptr, dst.ptr, nbytes, C.byref(ms), stream))
        if r:
            total += float(ms.value)
    src.free()
    dst.free()
    return total / reps


def gather_bench(nbytes=8 << 30, line_bytes=64, nlines=1 << 28, stream=None):
    """Random line gather over an nbytes HBM buffer; returns (ms, lines loaded)."""
    buf = DeviceBuffer(nbytes)
    buf.zero(stream)
    sink = DeviceBuffer(8)
    sink.zero(stream)
    ms = C.c_float(0)
    done = C.c_uint64(0)
    check(lib().fmi_gather_bench(buf.ptr, nbytes, line_bytes, nlines, sink.ptr, C.byref(ms),
                                 C.byref(done), stream))
    buf.free()
    sink.free()
    return float(ms.value), int(done.value)


class GatherProbe:
    """The random 64-byte line gather (fmi_gather_probe) over one buffer of
    nbytes, kept allocated across sweep points."""

    def __init__(self, nbytes):
        self.nbytes = int(nbytes)
        self.buf = DeviceBuffer(self.nbytes)
        self.buf.zero()
        self.sink = DeviceBuffer(8)
        self.sink.zero()

    def run(self, lines_in_flight, wg_per_cu, cooperative, nlines=1 << 26, stream=None):
        """(GB/s, G lines/s, ms) of one launch."""
        ms = C.c_float(0)
        done = C.c_uint64(0)
        check(lib().fmi_gather_probe(self.buf.ptr, self.nbytes, lines_in_flight, wg_per_cu,
                                     int(bool(cooperative)), nlines, self.sink.ptr, C.byref(ms),
                                     C.byref(done), stream))
        sec = float(ms.value) / 1e3
        return done.value * 64 / sec / 1e9, done.value / sec / 1e9, float(ms.value)

    def close(self):
        self.buf.free()
        self.sink.free()
