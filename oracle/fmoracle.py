"""ctypes loader for the checker libraries (test infrastructure only).

`Oracle` wraps oracle/libfmoracle.so, the C restatement of the reference
search path.  `RefCore` wraps oracle/_ref/libfmref_ref.so, the reference's
own search core compiled from its sources (present only where it was built).
Only tests/, the smoke check and the CPU-baseline leg of bench.py import this
module; the HIP engine never does.
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ORACLE_SO = os.path.join(HERE, "libfmoracle.so")
REF_SO = os.path.join(HERE, "_ref", "libfmref_ref.so")

# Same layout as fmo_hit / fmi_hit (32 bytes).
HIT_DTYPE = np.dtype([
    ("k", "<i8"), ("l", "<i8"),
    ("start", "<i2"), ("pos", "<i2"), ("end", "<i2"), ("err_pos", "<i2"),
    ("dir", "u1"), ("num_mismatches", "u1"), ("err_kind", "u1"), ("err_base", "u1"),
    ("strand", "u1"), ("pad", "u1", (3,)),
])
assert HIT_DTYPE.itemsize == 32

SEED_REGION_DTYPE = np.dtype([
    ("start", "<u8"), ("end", "<u8"), ("chrom_idx", "<u4"), ("read", "<u4"),
    ("seq_start", "<u2"), ("seq_end", "<u2"), ("seed_id", "<u2"), ("strand", "u1"),
    ("valid", "u1"),
])

INEXACT_REGION_DTYPE = np.dtype([
    ("start", "<u8"), ("end", "<u8"), ("chrom_idx", "<u4"), ("read", "<u4"),
    ("seq_start", "<u2"), ("seq_end", "<u2"), ("seq_len", "<u2"), ("seed_id", "<u2"),
    ("pos_err", "<i4"), ("type_err", "<i4"), ("strand", "u1"), ("pad", "u1", (7,)),
])

MATCH, DELETION, MISMATCH, INSERTION = 0, 1, 2, 3
VIEW_BACKWARD, VIEW_FORWARD, VIEW_BACKWARD_REV, VIEW_FORWARD_REV = 0, 1, 2, 3

_p = C.c_void_p
_u8p = C.POINTER(C.c_uint8)


HITK_DTYPE = np.dtype([
    ("k", "<i8"), ("l", "<i8"), ("start", "<i2"), ("pos", "<i2"), ("end", "<i2"),
    ("dir", "u1"), ("num_mismatches", "u1"), ("err_kind", "u1", (8,)), ("err_base", "u1", (8,)),
    ("err_pos", "<i2", (8,)), ("pad", "u1", (8,)),
])
assert HITK_DTYPE.itemsize == 64


# fmo_read_mapping (fmoracle.h) / fmi_read_mapping (include/fmi.h), 24 bytes
READ_MAPPING_DTYPE = np.dtype([
    ("position", "<u8"), ("chromosome", "<u4"), ("read", "<u4"), ("strand", "u1"),
    ("secondary", "u1"), ("cigar_len", "<u2"), ("pad", "<u4"),
])
assert READ_MAPPING_DTYPE.itemsize == 24


def hitsk_equal(a, b):
    """k-error records equal field by field, error entries up to num_mismatches."""
    if len(a) != len(b):
        return False
    for x, y in zip(a, b):
        for key in ("k", "l", "start", "pos", "end", "dir", "num_mismatches"):
            if x[key] != y[key]:
                return False
        m = min(int(x["num_mismatches"]), 8)
        for key in ("err_kind", "err_base", "err_pos"):
            if not np.array_equal(x[key][:m], y[key][:m]):
                return False
    return True


def _ptr(a):
    return a.ctypes.data_as(C.c_void_p) if a is not None else None


class _FmoDir(C.Structure):
    _fields_ = [("n", C.c_uint64), ("dollar", C.c_uint64),
                ("C", C.c_uint32 * 4), ("C1", C.c_uint32 * 4),
                ("words", C.c_uint64),
                ("desp", C.c_void_p * 4), ("count", C.c_void_p * 4),
                ("ratio", C.c_uint32), ("nsamp", C.c_uint64),
                ("S", C.c_void_p), ("R", C.c_void_p)]


class _FmoIndex(C.Structure):
    _fields_ = [("bwd", _FmoDir), ("fwd", _FmoDir)]


def _load_oracle():
    if not os.path.exists(ORACLE_SO):
        raise RuntimeError(f"{ORACLE_SO} is not built (make -C oracle)")
    lib = C.CDLL(ORACLE_SO)
    lib.fmo_sa_naive.argtypes = [_p, C.c_uint64, _p]
    lib.fmo_bwt_from_sa.argtypes = [_p, C.c_uint64, _p, _p, C.POINTER(C.c_uint64)]
    lib.fmo_index_build_naive.argtypes = [_p, _p, C.c_uint64, C.c_uint32]
    lib.fmo_index_from_bwts.argtypes = [_p, _p, C.c_uint64, _p, C.c_uint64, C.c_uint64, C.c_uint32]
    lib.fmo_dir_set_samples.argtypes = [_p, C.c_int, C.c_uint32, _p, _p, C.c_uint64]
    lib.fmo_index_free.argtypes = [_p]
    lib.fmo_dir_words.argtypes = [_p, C.c_int]
    lib.fmo_dir_words.restype = C.c_uint64
    lib.fmo_dir_export.argtypes = [_p, C.c_int, _p, _p, _p, _p, _p, _p,
                                   C.POINTER(C.c_uint64), C.POINTER(C.c_uint64)]
    lib.fmo_size_sa.argtypes = [_p]
    lib.fmo_size_sa.restype = C.c_uint64
    lib.fmo_occ.argtypes = [_p, C.c_int, C.c_uint, C.c_uint64]
    lib.fmo_occ.restype = C.c_uint64
    lib.fmo_lf.argtypes = [_p, C.c_int, C.c_int64, C.c_int64, C.c_uint,
                           C.POINTER(C.c_int64), C.POINTER(C.c_int64)]
    lib.fmo_get_sa.argtypes = [_p, C.c_int, C.c_uint64]
    lib.fmo_get_sa.restype = C.c_uint64
    lib.fmo_get_isa.argtypes = [_p, C.c_int, C.c_uint64]
    lib.fmo_get_isa.restype = C.c_uint64
    lib.fmo_exact.argtypes = [_p, C.c_int, C.c_int, _p, _p, _p]
    lib.fmo_exact_vector.argtypes = [_p, C.c_int, C.c_int, _p, C.c_int16, C.c_int16,
                                     C.c_int64, C.c_int64, _p, _p]
    lib.fmo_exact_vector.restype = C.c_int16
    lib.fmo_search1.argtypes = [_p, C.c_int, C.c_int, _p, C.c_int16, _p, C.c_uint32, _p]
    lib.fmo_search1.restype = C.c_uint32
    lib.fmo_search1_vector.argtypes = [_p, C.c_int, C.c_int, _p, C.c_int16, C.c_int16, _p,
                                       C.c_uint32]
    lib.fmo_search1_vector.restype = C.c_uint32
    lib.fmo_search_k.argtypes = [_p, C.c_int, _p, C.c_int16, C.c_int16, C.c_uint32, C.c_uint32,
                                 _p, C.POINTER(C.c_uint32), _p, C.POINTER(C.c_uint32)]
    lib.fmo_search_k.restype = C.c_int
    lib.fmo_seed_map.argtypes = [_p, _p, C.c_uint32, C.c_uint32, C.c_uint32, C.c_int32, _p, _p,
                                 C.c_uint64, C.c_uint32, _p, C.c_uint64, _p]
    lib.fmo_seed_map.restype = C.c_uint64
    lib.fmo_map_inexact_seeds.argtypes = [_p, _p, C.c_uint32, C.c_uint32, C.c_uint32, _p, _p,
                                          C.c_uint64, C.c_uint32, C.c_uint32, C.c_uint32, _p, _p,
                                          _p, C.c_uint32]
    lib.fmo_map_inexact_seeds.restype = C.c_uint32
    lib.fmo_map_inexact_seeds_batch.argtypes = [_p, _p, _p, C.c_uint64, C.c_uint32, C.c_uint32,
                                                _p, _p, C.c_uint64, C.c_uint32, C.c_uint32,
                                                C.POINTER(C.c_uint64), C.c_int]
    lib.fmo_map_inexact_seeds_batch.restype = C.c_double
    lib.fmo_binsearch.argtypes = [_p, C.c_uint64, C.c_uint64]
    lib.fmo_binsearch.restype = C.c_uint64
    lib.fmo_search1_batch.argtypes = [_p, _p, _p, C.c_uint64, C.c_int, C.c_uint32, _p, _p, C.c_int]
    lib.fmo_search1_batch.restype = C.c_double
    lib.fmo_search_k_batch.argtypes = [_p, _p, _p, C.c_uint64, _p, C.c_uint32, C.c_uint32, _p,
                                       _p, _p, _p, _p, C.c_int]
    lib.fmo_search_k_batch.restype = C.c_double
    lib.fmo_map_read.argtypes = [_p, _p, C.c_uint32, C.c_uint32, _p, _p, C.c_uint64, C.c_uint32,
                                 _p, C.c_uint32, _p]
    lib.fmo_map_read.restype = C.c_uint32
    lib.fmo_exact_batch.argtypes = [_p, _p, _p, C.c_uint64, _p, C.c_int]
    lib.fmo_exact_batch.restype = C.c_double
    lib.fmo_exact_locate_batch.argtypes = [_p, _p, _p, C.c_uint64, C.c_uint32, _p, _p, C.c_int]
    lib.fmo_exact_locate_batch.restype = C.c_double
    return lib


_ORACLE_LIB = None


def oracle_lib():
    global _ORACLE_LIB
    if _ORACLE_LIB is None:
        _ORACLE_LIB = _load_oracle()
    return _ORACLE_LIB


def sa_naive(text):
    text = np.ascontiguousarray(text, dtype=np.uint8)
    sa = np.empty(len(text) + 1, dtype=np.uint64)
    if oracle_lib().fmo_sa_naive(_ptr(text), len(text), _ptr(sa)) != 0:
        raise MemoryError("fmo_sa_naive")
    return sa


def bwt_from_sa(text, sa):
    text = np.ascontiguousarray(text, dtype=np.uint8)
    out = np.empty(max(len(text), 1), dtype=np.uint8)
    d = C.c_uint64(0)
    oracle_lib().fmo_bwt_from_sa(_ptr(text), len(text), _ptr(sa), _ptr(out), C.byref(d))
    return out[:len(text)], int(d.value)


class Oracle:
    """The C restatement of the reference index + search (per-symbol layout)."""

    def __init__(self, text=None, ratio=16, bwts=None):
        self.lib = oracle_lib()
        self._ix = _FmoIndex()
        if bwts is not None:
            b, d, bi, di = bwts
            b = np.ascontiguousarray(b, dtype=np.uint8)
            bi = np.ascontiguousarray(bi, dtype=np.uint8)
            n = len(b)
            rc = self.lib.fmo_index_from_bwts(C.byref(self._ix), _ptr(b), d, _ptr(bi), di, n, ratio)
        else:
            text = np.ascontiguousarray(text, dtype=np.uint8)
            n = len(text)
            rc = self.lib.fmo_index_build_naive(C.byref(self._ix), _ptr(text), n, ratio)
        if rc != 0:
            raise MemoryError("oracle index build failed")
        self.n = n
        self.ratio = ratio
        self.ref = C.byref(self._ix)

    def close(self):
        if self._ix is not None:
            self.lib.fmo_index_free(self.ref)
            self._ix = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_samples(self, ratio, sa_bwd, isa_bwd, sa_fwd, isa_fwd):
        """Samples every ratio-th entry of full SA / ISA arrays (uint32, n + 1,
        one pair per direction) into S and R, instead of the LF walk of
        calculate_S_and_R (preprocess.c:7-42) that ratio > 0 runs at build."""
        for fwd, sa, isa in ((0, sa_bwd, isa_bwd), (1, sa_fwd, isa_fwd)):
            S = np.ascontiguousarray(sa[::ratio], dtype=np.uint32)
            R = np.ascontiguousarray(isa[::ratio], dtype=np.uint32)
            if self.lib.fmo_dir_set_samples(self.ref, fwd, ratio, _ptr(S), _ptr(R), len(S)):
                raise ValueError("sample arrays do not fit the index")
        self.ratio = ratio

    def set_sampled(self, ratio, S_bwd, R_bwd, S_fwd, R_fwd):
        """Installs already-sampled S / R arrays (S.vec, R.vec, Si.vec, Ri.vec:
        every ratio-th entry of SA and ISA of each direction)."""
        for fwd, S, R in ((0, S_bwd, R_bwd), (1, S_fwd, R_fwd)):
            S = np.ascontiguousarray(S, dtype=np.uint32)
            R = np.ascontiguousarray(R, dtype=np.uint32)
            if self.lib.fmo_dir_set_samples(self.ref, fwd, ratio, _ptr(S), _ptr(R), len(S)):
                raise ValueError("sample arrays do not fit the index")
        self.ratio = ratio

    @property
    def size_sa(self):
        return int(self.lib.fmo_size_sa(self.ref))

    def arrays(self, fwd):
        w = int(self.lib.fmo_dir_words(self.ref, fwd))
        d = self._ix.fwd if fwd else self._ix.bwd
        ns = int(d.nsamp)
        Cv = np.empty(4, np.uint32); C1 = np.empty(4, np.uint32)
        desp = np.empty(4 * w, np.uint32); cnt = np.empty(4 * w, np.uint64)
        S = np.empty(ns, np.uint32); R = np.empty(ns, np.uint32)
        dol = C.c_uint64(); nsamp = C.c_uint64()
        self.lib.fmo_dir_export(self.ref, fwd, _ptr(Cv), _ptr(C1), _ptr(desp), _ptr(cnt),
                                _ptr(S), _ptr(R), C.byref(dol), C.byref(nsamp))
        return dict(C=Cv, C1=C1, desp=desp.reshape(4, w), count=cnt.reshape(4, w), S=S, R=R,
                    dollar=int(dol.value))

    def occ(self, view, b, m):
        return int(self.lib.fmo_occ(self.ref, view, b, m))

    def lf(self, view, k, l, b):
        ko = C.c_int64(); lo = C.c_int64()
        self.lib.fmo_lf(self.ref, view, k, l, b, C.byref(ko), C.byref(lo))
        return ko.value, lo.value

    def get_sa(self, fwd, m):
        return int(self.lib.fmo_get_sa(self.ref, fwd, m))

    def get_isa(self, fwd, m):
        return int(self.lib.fmo_get_isa(self.ref, fwd, m))

    def exact(self, view, direction, W, k, l, start, pos, end, anchor=False):
        W = np.ascontiguousarray(W, dtype=np.uint8)
        r = np.zeros(1, HIT_DTYPE)
        r["k"], r["l"], r["start"], r["pos"], r["end"] = k, l, start, pos, end
        a = np.zeros(1, HIT_DTYPE) if anchor else None
        self.lib.fmo_exact(self.ref, view, direction, _ptr(W), _ptr(r), _ptr(a))
        return (r[0], a[0]) if anchor else r[0]

    def exact_vector(self, view, direction, W, start, end, k, l):
        W = np.ascontiguousarray(W, dtype=np.uint8)
        m = max(end - start + 1, 1)
        vk = np.zeros(m, np.int64); vl = np.zeros(m, np.int64)
        g = self.lib.fmo_exact_vector(self.ref, view, direction, _ptr(W), start, end, k, l,
                                      _ptr(vk), _ptr(vl))
        return int(g), vk, vl

    def search1(self, strand, W, cap=1 << 14, mismatch_only=False, anchors=False):
        W = np.ascontiguousarray(W, dtype=np.uint8)
        out = np.zeros(max(cap, 1), HIT_DTYPE)
        an = np.zeros(2, HIT_DTYPE) if anchors else None
        n = self.lib.fmo_search1(self.ref, strand, int(mismatch_only), _ptr(W), len(W),
                                 _ptr(out), cap, _ptr(an))
        res = out[:min(n, cap)].copy()
        return (int(n), res, an) if anchors else (int(n), res)

    def search1_vector(self, strand, W, start=0, end=None, cap=1 << 14, mismatch_only=False):
        """BWSearch1VectorHelper over W[start..end] on one strand: (count, records)."""
        W = np.ascontiguousarray(W, dtype=np.uint8)
        end = len(W) - 1 if end is None else end
        out = np.zeros(max(cap, 1), HIT_DTYPE)
        n = self.lib.fmo_search1_vector(self.ref, strand, int(mismatch_only), _ptr(W), start, end,
                                        _ptr(out), cap)
        return int(n), out[:min(n, cap)].copy()

    def search_k(self, strand, W, fragsize, list_cap=2000, anchor_cap=16):
        """BWSearchCPU on one strand: (full, final records, anchors)."""
        W = np.ascontiguousarray(W, dtype=np.uint8)
        fin = np.zeros(max(list_cap, 1), HITK_DTYPE)
        anc = np.zeros(max(anchor_cap, 1), HITK_DTYPE)
        nf, na = C.c_uint32(0), C.c_uint32(0)
        full = self.lib.fmo_search_k(self.ref, strand, _ptr(W), len(W), fragsize, list_cap,
                                     anchor_cap, _ptr(fin), C.byref(nf), _ptr(anc), C.byref(na))
        return bool(full), fin[:nf.value].copy(), anc[:na.value].copy()

    def seed_map(self, W, seed_len, stride, filter_seed_mappings, koff, kstart, read=0):
        """bwt_map_exact_seed over the tiling of one read: (regions, info[nseeds, 6])."""
        W = np.ascontiguousarray(W, dtype=np.uint8)
        koff = np.ascontiguousarray(koff, dtype=np.uint64)
        kstart = np.ascontiguousarray(kstart, dtype=np.uint64)
        L = len(W)
        ns = (L - seed_len) // stride + 1 if L >= seed_len else 0
        info = np.zeros((max(ns, 1), 6), np.int64)
        n = self.lib.fmo_seed_map(self.ref, _ptr(W), L, seed_len, stride, filter_seed_mappings,
                                  _ptr(koff), _ptr(kstart), len(kstart), read, None, 0, _ptr(info))
        out = np.zeros(max(int(n), 1), SEED_REGION_DTYPE)
        self.lib.fmo_seed_map(self.ref, _ptr(W), L, seed_len, stride, filter_seed_mappings,
                              _ptr(koff), _ptr(kstart), len(kstart), read, _ptr(out), n, None)
        return out[:int(n)], info[:ns]

    def map_inexact_seeds(self, W, seed_size, min_seed_size, koff, kstart, read=0, rec_cap=256,
                          region_cap=64):
        """bwt_map_inexact_seeds_seq on one read: a list of (regions, status)
        per window and strand (+ then -), in the reference's list order."""
        W = np.ascontiguousarray(W, dtype=np.uint8)
        koff = np.ascontiguousarray(koff, dtype=np.uint64)
        kstart = np.ascontiguousarray(kstart, dtype=np.uint64)
        L = len(W)
        nw = L // seed_size + (1 if L % seed_size >= min_seed_size else 0)
        out = np.zeros(max(2 * nw * region_cap, 1), INEXACT_REGION_DTYPE)
        cnt = np.zeros(max(2 * nw, 1), np.uint32)
        st = np.zeros(max(2 * nw, 1), np.uint8)
        got = self.lib.fmo_map_inexact_seeds(self.ref, _ptr(W), L, seed_size, min_seed_size,
                                             _ptr(koff), _ptr(kstart), len(kstart), read, rec_cap,
                                             region_cap, _ptr(out), _ptr(cnt), _ptr(st), nw)
        assert got == nw
        return [(out[t * region_cap:t * region_cap + int(cnt[t])], int(st[t]))
                for t in range(2 * nw)]

    def map_inexact_seeds_batch(self, reads, offsets, seed_size, min_seed_size, koff, kstart,
                                rec_cap=256, region_cap=64, threads=0):
        """(seconds, regions kept) of map_inexact_seeds over a batch."""
        reads = np.ascontiguousarray(reads, dtype=np.uint8)
        offsets = np.ascontiguousarray(offsets, dtype=np.uint64)
        koff = np.ascontiguousarray(koff, dtype=np.uint64)
        kstart = np.ascontiguousarray(kstart, dtype=np.uint64)
        tot = C.c_uint64()
        t = self.lib.fmo_map_inexact_seeds_batch(self.ref, _ptr(reads), _ptr(offsets),
                                                 len(offsets) - 1, seed_size, min_seed_size,
                                                 _ptr(koff), _ptr(kstart), len(kstart), rec_cap,
                                                 region_cap, C.byref(tot), threads)
        return float(t), int(tot.value)

    def map_read(self, W, filter_read_mappings, koff, kstart, read=0, cap=None):
        """__bwt_map_inexact_read's alignments of one read: (status, records
        [READ_MAPPING_DTYPE], total); status bit 0 filter exceeded, bit 1 a
        read under 5 bases."""
        W = np.ascontiguousarray(W, dtype=np.uint8)
        koff = np.ascontiguousarray(koff, dtype=np.uint64)
        kstart = np.ascontiguousarray(kstart, dtype=np.uint64)
        cap = 2 * (filter_read_mappings + 1) if cap is None else cap
        out = np.zeros(max(cap, 1), READ_MAPPING_DTYPE)
        st = np.zeros(1, np.uint8)
        n = self.lib.fmo_map_read(self.ref, _ptr(W), len(W), filter_read_mappings, _ptr(koff),
                                  _ptr(kstart), len(kstart), read, _ptr(out), cap, _ptr(st))
        return int(st[0]), out[:min(n, cap)].copy(), int(n)

    def search1_batch(self, reads, offsets, cap=64, mismatch_only=False, threads=0,
                      keep_hits=True):
        reads = np.ascontiguousarray(reads, dtype=np.uint8)
        offsets = np.ascontiguousarray(offsets, dtype=np.uint64)
        nr = len(offsets) - 1
        hits = np.zeros(nr * 2 * cap, HIT_DTYPE) if keep_hits else None
        counts = np.zeros(nr * 2, np.uint32)
        t = self.lib.fmo_search1_batch(self.ref, _ptr(reads), _ptr(offsets), nr,
                                       int(mismatch_only), cap, _ptr(hits), _ptr(counts), threads)
        return t, counts, hits

    def search_k_batch(self, reads, offsets, fragsize, list_cap=128, anchor_cap=16, threads=0):
        """BWSearchCPU on both strands of every read: (seconds, nfinal[2n],
        final[2n * list_cap], nanchors[2n], anchors[2n * anchor_cap],
        full[n, 2]) in the layout fmi_search_k_batch writes."""
        reads = np.ascontiguousarray(reads, dtype=np.uint8)
        offsets = np.ascontiguousarray(offsets, dtype=np.uint64)
        fragsize = np.ascontiguousarray(fragsize, dtype=np.int16)
        nr = len(offsets) - 1
        fin = np.zeros(max(2 * nr * list_cap, 1), HITK_DTYPE)
        anc = np.zeros(max(2 * nr * anchor_cap, 1), HITK_DTYPE)
        nf = np.zeros(max(2 * nr, 1), np.uint32)
        na = np.zeros(max(2 * nr, 1), np.uint32)
        full = np.zeros(max(2 * nr, 1), np.uint8)
        t = self.lib.fmo_search_k_batch(self.ref, _ptr(reads), _ptr(offsets), nr, _ptr(fragsize),
                                        list_cap, anchor_cap, _ptr(fin), _ptr(nf), _ptr(anc),
                                        _ptr(na), _ptr(full), threads)
        return float(t), nf[:2 * nr], fin, na[:2 * nr], anc, full[:2 * nr].reshape(nr, 2)

    def exact_batch(self, reads, offsets, threads=0):
        reads = np.ascontiguousarray(reads, dtype=np.uint8)
        offsets = np.ascontiguousarray(offsets, dtype=np.uint64)
        nr = len(offsets) - 1
        out = np.zeros(nr * 4, np.int64)
        t = self.lib.fmo_exact_batch(self.ref, _ptr(reads), _ptr(offsets), nr, _ptr(out), threads)
        return t, out.reshape(nr, 4)

    def exact_locate_batch(self, reads, offsets, cap=64, threads=0):
        """(seconds, rows[n, 2], position sums[n, 2]): exact search of both
        strands, then get_SA over the first cap rows of each interval."""
        reads = np.ascontiguousarray(reads, dtype=np.uint8)
        offsets = np.ascontiguousarray(offsets, dtype=np.uint64)
        n = len(offsets) - 1
        rows = np.zeros(max(2 * n, 1), np.uint32)
        sums = np.zeros(max(2 * n, 1), np.uint64)
        t = self.lib.fmo_exact_locate_batch(self.ref, _ptr(reads), _ptr(offsets), n, cap,
                                            _ptr(rows), _ptr(sums), threads)
        return float(t), rows[:2 * n].reshape(n, 2), sums[:2 * n].reshape(n, 2)


def binsearch(array, key):
    array = np.ascontiguousarray(array, dtype=np.uint64)
    return int(oracle_lib().fmo_binsearch(_ptr(array), len(array), key))


def ref_available():
    return os.path.exists(REF_SO)


class _RefDirArrays(C.Structure):
    _fields_ = [("C", C.c_void_p), ("C1", C.c_void_p), ("desp", C.c_void_p * 4),
                ("count", C.c_void_p * 4), ("words", C.c_uint64), ("S", C.c_void_p),
                ("R", C.c_void_p), ("nsamp", C.c_uint64)]


class RefCore:
    """The reference's own search core (compiled from its sources).

    RefCore(text) builds the index with the reference builder (dbwt and
    calculate_C/O/S_and_R).  RefCore(oracle=o) runs the reference search
    code over the oracle's arrays, which equal the reference builder's
    (pinned by the golden SHA of C/C1/desp/count/S/R); the oracle must stay
    alive while this object is used."""

    def __init__(self, text=None, ratio=16, workdir=None, oracle=None):
        import tempfile
        if not ref_available():
            raise RuntimeError(f"{REF_SO} is not built (make -C oracle ref)")
        lib = C.CDLL(REF_SO)
        lib.ref_build.argtypes = [_p, C.c_uint64, C.c_uint32, C.c_char_p]
        lib.ref_build.restype = C.c_void_p
        lib.ref_wrap.argtypes = [_p, _p, C.c_uint64, C.c_uint32]
        lib.ref_wrap.restype = C.c_void_p
        lib.ref_free.argtypes = [_p]
        lib.ref_size_sa.argtypes = [_p]
        lib.ref_size_sa.restype = C.c_uint64
        lib.ref_bwt.argtypes = [_p, C.c_int, _p, C.POINTER(C.c_uint64)]
        lib.ref_words.argtypes = [_p]
        lib.ref_words.restype = C.c_uint64
        lib.ref_nsamp.argtypes = [_p]
        lib.ref_nsamp.restype = C.c_uint64
        lib.ref_arrays.argtypes = [_p, C.c_int, _p, _p, _p, _p, _p, _p]
        lib.ref_exact.argtypes = [_p, C.c_int, C.c_int, _p, _p, _p]
        lib.ref_exact_vector.argtypes = [_p, C.c_int, C.c_int, _p, C.c_int16, C.c_int16,
                                         C.c_int64, C.c_int64, _p, _p]
        lib.ref_exact_vector.restype = C.c_int16
        lib.ref_search1.argtypes = [_p, C.c_int, _p, C.c_int16, C.c_uint32, _p, _p,
                                    C.POINTER(C.c_int)]
        lib.ref_search1.restype = C.c_uint32
        lib.ref_search1_vector.argtypes = [_p, C.c_int, _p, C.c_int16, C.c_int16, C.c_uint32, _p,
                                           C.POINTER(C.c_int)]
        lib.ref_search1_vector.restype = C.c_uint32
        lib.ref_search_k.argtypes = [_p, C.c_int, _p, C.c_int16, C.c_int16, C.c_uint32,
                                     C.c_uint32, _p, C.POINTER(C.c_uint32), _p,
                                     C.POINTER(C.c_uint32)]
        lib.ref_search_k.restype = C.c_int
        lib.ref_search_k_batch.argtypes = [_p, _p, _p, _p, C.c_uint64, C.c_uint32, C.c_uint32,
                                           _p, C.c_int]
        lib.ref_search_k_batch.restype = C.c_double
        lib.ref_get_sa.argtypes = [_p, C.c_int, C.c_uint64]
        lib.ref_get_sa.restype = C.c_uint64
        lib.ref_get_isa.argtypes = [_p, C.c_int, C.c_uint64]
        lib.ref_get_isa.restype = C.c_uint64
        lib.ref_binsearch.argtypes = [_p, C.c_uint64, C.c_uint64]
        lib.ref_binsearch.restype = C.c_uint64
        lib.ref_search1_batch.argtypes = [_p, _p, _p, C.c_uint64, C.c_uint32, _p, C.c_int]
        lib.ref_search1_batch.restype = C.c_double
        lib.ref_exact_batch.argtypes = [_p, _p, _p, C.c_uint64, _p, C.c_int]
        lib.ref_exact_batch.restype = C.c_double
        lib.ref_save_dir.argtypes = [_p, C.c_char_p, C.c_uint64, _p, _p]
        lib.ref_write_results.argtypes = [_p, C.c_int, _p, C.c_char_p, C.c_int16, C.c_uint64,
                                          C.c_uint32, C.c_uint64, _p, _p, C.c_char_p, C.c_size_t]
        lib.ref_write_results.restype = C.c_size_t
        lib.ref_exact_locate_batch.argtypes = [_p, _p, _p, C.c_uint64, C.c_uint32, _p, _p, C.c_int]
        lib.ref_exact_locate_batch.restype = C.c_double
        self.lib = lib
        self._tmp = None
        if oracle is not None:
            self._oracle = oracle
            dirs = []
            for d in (oracle._ix.bwd, oracle._ix.fwd):
                a = _RefDirArrays()
                a.C = C.cast(C.addressof(d.C), C.c_void_p)
                a.C1 = C.cast(C.addressof(d.C1), C.c_void_p)
                for b in range(4):
                    a.desp[b] = d.desp[b]
                    a.count[b] = d.count[b]
                a.words, a.S, a.R, a.nsamp = d.words, d.S, d.R, d.nsamp
                dirs.append(a)
            self._dirs = dirs
            self.n = oracle.n
            self.h = lib.ref_wrap(C.byref(dirs[0]), C.byref(dirs[1]), self.n, oracle.ratio)
            return
        text = np.ascontiguousarray(text, dtype=np.uint8)
        self.n = len(text)
        self._tmp = tempfile.TemporaryDirectory() if workdir is None else None
        wd = workdir or self._tmp.name
        self.h = lib.ref_build(_ptr(text), self.n, ratio, wd.encode())

    def close(self):
        if self.h:
            self.lib.ref_free(self.h)
            self.h = None
        if self._tmp is not None:
            self._tmp.cleanup()
            self._tmp = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def size_sa(self):
        return int(self.lib.ref_size_sa(self.h))

    def bwt(self, fwd):
        out = np.empty(max(self.n, 1), np.uint8)
        d = C.c_uint64()
        self.lib.ref_bwt(self.h, fwd, _ptr(out), C.byref(d))
        return out[:self.n], int(d.value)

    def arrays(self, fwd):
        w = int(self.lib.ref_words(self.h)); ns = int(self.lib.ref_nsamp(self.h))
        Cv = np.empty(4, np.uint32); C1 = np.empty(4, np.uint32)
        desp = np.empty(4 * w, np.uint32); cnt = np.empty(4 * w, np.uint64)
        S = np.empty(ns, np.uint32); R = np.empty(ns, np.uint32)
        self.lib.ref_arrays(self.h, fwd, _ptr(Cv), _ptr(C1), _ptr(desp), _ptr(cnt), _ptr(S), _ptr(R))
        return dict(C=Cv, C1=C1, desp=desp.reshape(4, w), count=cnt.reshape(4, w), S=S, R=R)

    def exact(self, view, direction, W, k, l, start, pos, end, anchor=False):
        W = np.ascontiguousarray(W, dtype=np.uint8)
        r = np.zeros(1, HIT_DTYPE)
        r["k"], r["l"], r["start"], r["pos"], r["end"] = k, l, start, pos, end
        a = np.zeros(1, HIT_DTYPE) if anchor else None
        self.lib.ref_exact(self.h, view, direction, _ptr(W), _ptr(r), _ptr(a))
        return (r[0], a[0]) if anchor else r[0]

    def exact_vector(self, view, direction, W, start, end, k, l):
        W = np.ascontiguousarray(W, dtype=np.uint8)
        m = max(end - start + 1, 1)
        vk = np.zeros(m, np.int64); vl = np.zeros(m, np.int64)
        g = self.lib.ref_exact_vector(self.h, view, direction, _ptr(W), start, end, k, l,
                                      _ptr(vk), _ptr(vl))
        return int(g), vk, vl

    def search1(self, strand, W, cap=1 << 14, anchors=False):
        W = np.ascontiguousarray(W, dtype=np.uint8)
        out = np.zeros(max(cap, 1), HIT_DTYPE)
        an = np.zeros(2, HIT_DTYPE) if anchors else None
        full = C.c_int(0)
        n = self.lib.ref_search1(self.h, strand, _ptr(W), len(W), cap, _ptr(out), _ptr(an),
                                 C.byref(full))
        res = out[:n].copy()
        return (int(n), res, bool(full.value), an) if anchors else (int(n), res, bool(full.value))

    def search1_vector(self, strand, W, start=0, end=None, cap=1 << 14):
        W = np.ascontiguousarray(W, dtype=np.uint8)
        end = len(W) - 1 if end is None else end
        out = np.zeros(max(cap, 1), HIT_DTYPE)
        full = C.c_int(0)
        n = self.lib.ref_search1_vector(self.h, strand, _ptr(W), start, end, cap, _ptr(out),
                                        C.byref(full))
        return int(n), out[:n].copy(), bool(full.value)

    def search_k(self, strand, W, fragsize, list_cap=2000, anchor_cap=16):
        W = np.ascontiguousarray(W, dtype=np.uint8)
        fin = np.zeros(max(list_cap, 1), HITK_DTYPE)
        anc = np.zeros(max(anchor_cap, 1), HITK_DTYPE)
        nf, na = C.c_uint32(0), C.c_uint32(0)
        full = self.lib.ref_search_k(self.h, strand, _ptr(W), len(W), fragsize, list_cap,
                                     anchor_cap, _ptr(fin), C.byref(nf), _ptr(anc), C.byref(na))
        return bool(full), fin[:nf.value].copy(), anc[:na.value].copy()

    def write_results(self, strand, W, mapping, read_index, cap, starts, ends):
        """The reference's write_results text for BWSearch1CPU on one strand
        of W (codes) with the read text `mapping` (bytes)."""
        W = np.ascontiguousarray(W, dtype=np.uint8)
        st = np.ascontiguousarray(starts, dtype=np.uint64)
        en = np.ascontiguousarray(ends, dtype=np.uint64)
        size = 1 << 16
        while True:
            buf = C.create_string_buffer(size)
            n = self.lib.ref_write_results(self.h, strand, _ptr(W), mapping, len(W), read_index,
                                           cap, len(st), _ptr(st), _ptr(en), buf, size)
            if n < size:
                return buf.raw[:n].decode()
            size = n + 1

    def save_dir(self, path, starts=None, ends=None):
        """Writes the reference's index directory (its own writers) to path;
        starts / ends: the karyotype (default: one contig over the text)."""
        if starts is None:
            starts, ends = [0], [self.n - 1]
        st = np.ascontiguousarray(starts, dtype=np.uint64)
        en = np.ascontiguousarray(ends, dtype=np.uint64)
        os.makedirs(path, exist_ok=True)
        if self.lib.ref_save_dir(self.h, path.encode(), len(st), _ptr(st), _ptr(en)) != 0:
            raise ValueError("too many contigs")

    def exact_locate_batch(self, reads, offsets, cap=64, threads=0):
        """(seconds, rows[n, 2], position sums[n, 2]) of the exact search of
        both strands followed by get_SA over the first cap rows."""
        reads = np.ascontiguousarray(reads, dtype=np.uint8)
        offsets = np.ascontiguousarray(offsets, dtype=np.uint64)
        n = len(offsets) - 1
        rows = np.zeros(max(2 * n, 1), np.uint32)
        sums = np.zeros(max(2 * n, 1), np.uint64)
        t = self.lib.ref_exact_locate_batch(self.h, _ptr(reads), _ptr(offsets), n, cap, _ptr(rows),
                                            _ptr(sums), threads)
        return float(t), rows[:2 * n].reshape(n, 2), sums[:2 * n].reshape(n, 2)

    def search_k_batch(self, reads, offsets, fragsize, list_cap=128, anchor_cap=16, threads=0):
        """(seconds, final-list sizes[2n]) of BWSearchCPU over a batch."""
        reads = np.ascontiguousarray(reads, dtype=np.uint8)
        offsets = np.ascontiguousarray(offsets, dtype=np.uint64)
        fragsize = np.ascontiguousarray(fragsize, dtype=np.int16)
        n = len(offsets) - 1
        out = np.zeros(max(2 * n, 1), np.uint32)
        t = self.lib.ref_search_k_batch(self.h, _ptr(reads), _ptr(offsets), _ptr(fragsize), n,
                                        list_cap, anchor_cap, _ptr(out), threads)
        return float(t), out[:2 * n]

    def get_sa(self, fwd, m):
        return int(self.lib.ref_get_sa(self.h, fwd, m))

    def get_isa(self, fwd, m):
        return int(self.lib.ref_get_isa(self.h, fwd, m))

    def search1_batch(self, reads, offsets, cap=64, threads=0):
        reads = np.ascontiguousarray(reads, dtype=np.uint8)
        offsets = np.ascontiguousarray(offsets, dtype=np.uint64)
        nr = len(offsets) - 1
        counts = np.zeros(nr * 2, np.uint32)
        t = self.lib.ref_search1_batch(self.h, _ptr(reads), _ptr(offsets), nr, cap,
                                       _ptr(counts), threads)
        return t, counts

    def exact_batch(self, reads, offsets, threads=0):
        reads = np.ascontiguousarray(reads, dtype=np.uint8)
        offsets = np.ascontiguousarray(offsets, dtype=np.uint64)
        nr = len(offsets) - 1
        out = np.zeros(nr * 4, np.int64)
        t = self.lib.ref_exact_batch(self.h, _ptr(reads), _ptr(offsets), nr, _ptr(out), threads)
        return t, out.reshape(nr, 4)


def hits_equal(a, b):
    """Record-wise equality; error fields compare only where num_mismatches > 0
    (the reference leaves them unset on exact records)."""
    if len(a) != len(b):
        return False
    keys = ["k", "l", "start", "pos", "end", "dir", "num_mismatches", "strand"]
    for key in keys:
        if not np.array_equal(a[key], b[key]):
            return False
    m = a["num_mismatches"] > 0
    for key in ["err_kind", "err_base", "err_pos"]:
        if not np.array_equal(a[key][m], b[key][m]):
            return False
    return True
