"""ctypes binding of libfastbit_hip.so (include/fastbit_hip.h).

The product path: every call goes to the HIP library; if the library is not
built or no GPU is present, calls raise.  There is no CPU fallback.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "lib", "libfastbit_hip.so")

AND, OR, XOR, MINUS = 0, 1, 2, 3
PATHS = {1: "c0", 2: "d1", 3: "d1swap", 4: "c1", 5: "d2", 6: "c2", 7: "shortcut"}

_u32p = C.POINTER(C.c_uint32)
_lib = None


class FastbitError(RuntimeError):
    pass


def lib():
    """Load the HIP library (raises if it was not built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise FastbitError(f"{LIB_PATH} is not built; run __graft_entry__.build()")
        L = C.CDLL(LIB_PATH)
        vp = C.c_void_p
        L.fbh_last_error.restype = C.c_char_p
        L.fbh_bv_free.restype = None
        L.fbh_bv_free.argtypes = [vp]
        for name, args in {
            "fbh_init": [C.c_int],
            "fbh_sync": [],
            "fbh_bv_new": [C.POINTER(vp)],
            "fbh_bv_from_serial": [_u32p, C.c_uint32, C.POINTER(vp)],
            "fbh_bv_set": [vp, C.c_int, C.c_uint32],
            "fbh_bv_copy": [vp, C.POINTER(vp)],
            "fbh_bv_write": [vp, _u32p, C.c_uint32, _u32p],
            "fbh_bv_op": [vp, vp, C.c_int, C.POINTER(C.c_int)],
            "fbh_bv_op_new": [vp, vp, C.c_int, C.POINTER(vp), C.POINTER(C.c_int)],
            "fbh_bv_count": [vp, vp, _u32p],
            "fbh_bv_erase": [vp, C.c_uint32, C.c_uint32],
            "fbh_bv_compress": [vp],
            "fbh_bv_decompress": [vp],
            "fbh_bv_flip": [vp],
            "fbh_bv_cnt": [vp, _u32p],
            "fbh_bv_size": [vp, _u32p],
            "fbh_bv_adjust_size": [vp, C.c_uint32, C.c_uint32],
            "fbh_bv_rids": [vp, _u32p, C.c_uint32, _u32p],
            "fbh_bv_nwords": [vp, _u32p],
            "fbh_bench_op": [vp, vp, C.c_int, C.c_int, C.POINTER(C.c_float), _u32p],
            "fbh_bench_cnt_rot": [C.POINTER(vp), C.c_int, C.c_int, C.POINTER(C.c_float),
                                  C.POINTER(C.c_uint32)],
            "fbh_bench_op_rot": [C.POINTER(vp), C.POINTER(vp), C.c_int, C.c_int, C.c_int,
                                 C.POINTER(C.c_float), _u32p],
        }.items():
            f = getattr(L, name)
            f.argtypes = args
            f.restype = C.c_int
        _lib = L
    return _lib


def check(rc: int, what: str = ""):
    if rc != 0:
        msg = lib().fbh_last_error()
        raise FastbitError(f"{what} failed (rc={rc}): {msg.decode() if msg else ''}")


def _u32(a):
    a = np.ascontiguousarray(a, dtype=np.uint32)
    return a, a.ctypes.data_as(_u32p)


class Bitvector:
    """An ibis::bitvector held in HBM (src/bitvector.h:62-318).

    The in-place operators are the reference's ``a op= b`` followed by
    ``compress()``; ``last_path`` records the routine the dispatch chose.
    """

    __slots__ = ("_h", "last_path")

    def __init__(self, handle=None):
        if handle is None:
            h = C.c_void_p()
            check(lib().fbh_bv_new(C.byref(h)), "fbh_bv_new")
            handle = h
        self._h = handle
        self.last_path = None

    @classmethod
    def from_serial(cls, words) -> "Bitvector":
        a, p = _u32(words)
        h = C.c_void_p()
        check(lib().fbh_bv_from_serial(p, a.size, C.byref(h)), "fbh_bv_from_serial")
        return cls(h)

    def __del__(self):
        try:
            if self._h is not None and _lib is not None:
                _lib.fbh_bv_free(self._h)
        except Exception:
            pass
        self._h = None

    def copy(self) -> "Bitvector":
        h = C.c_void_p()
        check(lib().fbh_bv_copy(self._h, C.byref(h)), "fbh_bv_copy")
        return Bitvector(h)

    def set(self, val: int, n: int) -> None:
        check(lib().fbh_bv_set(self._h, int(val), int(n)), "fbh_bv_set")

    def write(self) -> np.ndarray:
        """Serialized words (write(array_t&), src/bitvector.cpp:2283-2313)."""
        n = C.c_uint32()
        check(lib().fbh_bv_write(self._h, None, 0, C.byref(n)), "fbh_bv_write")
        out = np.zeros(n.value, dtype=np.uint32)
        check(lib().fbh_bv_write(self._h, out.ctypes.data_as(_u32p), n.value, C.byref(n)),
              "fbh_bv_write")
        return out

    serial = write

    def op(self, other: "Bitvector", op: int) -> str:
        path = C.c_int(0)
        check(lib().fbh_bv_op(self._h, other._h, int(op), C.byref(path)), "fbh_bv_op")
        self.last_path = PATHS.get(path.value, "?")
        return self.last_path

    def op_new(self, other: "Bitvector", op: int):
        """The out-of-place operator (a new, uncompressed Bitvector, as
        ibis::bitvector::operator& returns it) and the routine's name."""
        h = C.c_void_p()
        path = C.c_int(0)
        check(lib().fbh_bv_op_new(self._h, other._h, int(op), C.byref(h), C.byref(path)), "fbh_bv_op_new")
        return Bitvector(h), PATHS.get(path.value, "?")

    def __and__(self, o):
        return self.op_new(o, AND)[0]

    def __or__(self, o):
        return self.op_new(o, OR)[0]

    def __xor__(self, o):
        return self.op_new(o, XOR)[0]

    def __sub__(self, o):
        return self.op_new(o, MINUS)[0]

    def count(self, mask: "Bitvector") -> int:
        """count(mask): the ones of self AND mask."""
        c = C.c_uint32()
        check(lib().fbh_bv_count(self._h, mask._h, C.byref(c)), "fbh_bv_count")
        return c.value

    def erase(self, i: int, j: int) -> None:
        check(lib().fbh_bv_erase(self._h, int(i), int(j)), "fbh_bv_erase")

    def __iand__(self, o):
        self.op(o, AND)
        return self

    def __ior__(self, o):
        self.op(o, OR)
        return self

    def __ixor__(self, o):
        self.op(o, XOR)
        return self

    def __isub__(self, o):
        self.op(o, MINUS)
        return self

    def compress(self):
        check(lib().fbh_bv_compress(self._h), "fbh_bv_compress")

    def decompress(self):
        check(lib().fbh_bv_decompress(self._h), "fbh_bv_decompress")

    def flip(self):
        check(lib().fbh_bv_flip(self._h), "fbh_bv_flip")

    def cnt(self) -> int:
        v = C.c_uint32()
        check(lib().fbh_bv_cnt(self._h, C.byref(v)), "fbh_bv_cnt")
        return v.value

    def size(self) -> int:
        v = C.c_uint32()
        check(lib().fbh_bv_size(self._h, C.byref(v)), "fbh_bv_size")
        return v.value

    def adjustSize(self, nv: int, nt: int):
        check(lib().fbh_bv_adjust_size(self._h, int(nv), int(nt)), "fbh_bv_adjust_size")

    @property
    def nwords(self) -> int:
        v = C.c_uint32()
        check(lib().fbh_bv_nwords(self._h, C.byref(v)), "fbh_bv_nwords")
        return v.value

    def rids(self) -> np.ndarray:
        n = C.c_uint32()
        check(lib().fbh_bv_rids(self._h, None, 0, C.byref(n)), "fbh_bv_rids")
        out = np.zeros(max(n.value, 1), dtype=np.uint32)
        check(lib().fbh_bv_rids(self._h, out.ctypes.data_as(_u32p), n.value, C.byref(n)),
              "fbh_bv_rids")
        return out[: n.value]


def bench_op(a: Bitvector, b: Bitvector, op: int, reps: int):
    """(ms per op, result words) for (copy of a) op= b + compress."""
    ms = C.c_float()
    nw = C.c_uint32()
    check(lib().fbh_bench_op(a._h, b._h, int(op), int(reps), C.byref(ms), C.byref(nw)),
          "fbh_bench_op")
    return ms.value, nw.value


def bench_op_rot(a, b, op: int, reps: int):
    """As bench_op, cycling over the operand pairs zip(a, b) so that the
    working set streams from HBM instead of the last-level cache."""
    n = len(a)
    if n == 0 or len(b) != n:
        raise ValueError("bench_op_rot needs equal, non-empty operand lists")
    pa = (C.c_void_p * n)(*[x._h for x in a])
    pb = (C.c_void_p * n)(*[x._h for x in b])
    ms = C.c_float()
    nw = C.c_uint32()
    check(lib().fbh_bench_op_rot(pa, pb, n, int(op), int(reps), C.byref(ms), C.byref(nw)),
          "fbh_bench_op_rot")
    return ms.value, nw.value


def bench_cnt_rot(a, reps: int):
    """cnt() on the device, cycling over the bitvectors in a: (ms per
    count, ones of a[0])."""
    n = len(a)
    if n == 0:
        raise ValueError("bench_cnt_rot needs a non-empty list")
    pa = (C.c_void_p * n)(*[x._h for x in a])
    ms = C.c_float()
    ones = C.c_uint32()
    check(lib().fbh_bench_cnt_rot(pa, n, int(reps), C.byref(ms), C.byref(ones)), "fbh_bench_cnt_rot")
    return ms.value, ones.value


# ---------------------------------------------------------------------------
# Indexes
RELIC, BIN = 1, 2
COL_INT32, COL_UINT32, COL_UINT16, COL_FLOAT, COL_DOUBLE = 0, 1, 2, 3, 4
OP_NONE, OP_LT, OP_LE = 0, 1, 2
PLAN_AUTO, PLAN_LITERAL, PLAN_COMPRESSED, PLAN_SCAN = 0, 1, 2, 3
# two-level encodings and what an evaluation on one did (fastbit_hip.h)
ENC_FUZZ, ENC_ZONA = 1, 2
TL_EMPTY, TL_FINE, TL_ALL, TL_EQUAL, TL_SAME_BIN = 0, 1, 6, 7, 8
PLAN_NAMES = {0: "auto", 1: "literal", 2: "compressed", 3: "scan"}
COST_IAPI, COST_PARTITION = 0, 1
_COLT = {np.dtype(np.int32): COL_INT32, np.dtype(np.uint32): COL_UINT32,
         np.dtype(np.uint16): COL_UINT16, np.dtype(np.float32): COL_FLOAT,
         np.dtype(np.float64): COL_DOUBLE}
_dp = C.POINTER(C.c_double)
_i64p = C.POINTER(C.c_int64)
_u64p = C.POINTER(C.c_uint64)
_ip = C.POINTER(C.c_int)
_fp = C.POINTER(C.c_float)
NFEATURES, NPLANS, NPARAM = 12, 3, 5187


def _index_sigs(L):
    vp = C.c_void_p
    sig = {
        "fbh_index_build_relic": [vp, C.c_int, C.c_uint32, C.POINTER(vp)],
        "fbh_index_build_relic_card": [_u32p, C.c_uint32, C.c_uint32, C.POINTER(vp)],
        "fbh_index_build_bin": [vp, C.c_int, C.c_uint32, _dp, C.c_uint32, C.POINTER(vp)],
        "fbh_index_info": [vp, _u32p, _u32p, _u64p],
        "fbh_index_deconstruct": [vp, C.POINTER(_dp), _u64p, C.POINTER(_i64p), _u64p,
                                  C.POINTER(_u32p), _u64p],
        "fbh_index_resolve_range": [vp, C.c_double, C.c_int, C.c_double, C.c_int, _u32p, _u32p,
                                    _u32p, _u32p],
        "fbh_index_estimate_cost": [vp, C.c_double, C.c_int, C.c_double, C.c_int, _dp, _dp,
                                    C.POINTER(C.c_int)],
        "fbh_index_evaluate": [vp, C.c_double, C.c_int, C.c_double, C.c_int, C.c_int, _u64p,
                               C.POINTER(vp), C.POINTER(C.c_int)],
        "fbh_index_evaluate_rids": [vp, C.c_double, C.c_int, C.c_double, C.c_int, C.c_int, _u32p,
                                    C.c_uint64, _u64p],
        "fbh_index_evaluate_batch": [vp, C.c_uint32, _dp, _ip, _dp, _ip, _ip, C.c_int, _u64p, _ip,
                                     C.POINTER(C.c_float)],
        "fbh_index_query_bytes": [vp, C.c_double, C.c_int, C.c_double, C.c_int, C.c_int, _dp],
        "fbh_index_estimate": [vp, C.c_double, C.c_int, C.c_double, C.c_int, C.POINTER(vp), C.POINTER(vp)],
        "fbh_index_evaluate_noalloc": [vp, C.c_double, C.c_int, C.c_double, C.c_int, C.c_int, vp,
                                       C.c_uint64, _u64p, C.POINTER(C.c_int)],
        "fbh_index_column_bytes": [vp, _u64p],
        "fbh_index_set_cost_model": [vp, C.c_int],
        "fbh_index_build_bin_nbins": [vp, C.c_int, C.c_uint32, C.c_uint32, C.POINTER(vp)],
        "fbh_index_bounds": [vp, _dp, C.c_uint32, _u32p],
        "fbh_index_attach": [vp, C.c_int, C.c_uint32, _dp, C.c_uint64, _i64p, C.c_uint64, _u32p,
                             C.c_uint64, C.POINTER(vp)],
        "fbh_index_read_file": [C.c_char_p, vp, C.c_int, C.c_uint32, C.POINTER(vp)],
        "fbh_index_build_fuzz": [vp, C.c_int, C.c_uint32, C.c_uint32, C.POINTER(vp)],
        "fbh_index_build_zona": [vp, C.c_int, C.c_uint32, C.POINTER(vp)],
        "fbh_index_coarse": [vp, _ip, _u32p, _u32p, C.c_uint32, _i64p, C.c_uint32, _u32p],
        "fbh_index_coarse_bitmap": [vp, C.c_uint32, C.POINTER(vp)],
        "fbh_index_two_level_option": [vp, C.c_double, C.c_int, C.c_double, C.c_int, _ip, _u64p],
        "fbh_index_sum_bins": [vp, C.c_uint32, C.c_uint32, _u64p],
        "fbh_index_plan_features": [vp, C.c_uint32, _dp, _ip, _dp, _ip, _fp],
        "fbh_index_time_plans": [vp, C.c_uint32, _dp, _ip, _dp, _ip, C.c_float, _fp, _u64p],
        "fbh_index_choose_plans": [vp, vp, C.c_uint32, _dp, _ip, _dp, _ip, _ip, _u32p],
        "fbh_planner_new": [C.c_uint64, C.POINTER(vp)],
        "fbh_planner_set_params": [vp, _fp, C.c_uint32],
        "fbh_planner_get_params": [vp, _fp, C.c_uint32],
        "fbh_planner_set_norm": [vp, _fp, _fp],
        "fbh_planner_get_norm": [vp, _fp, _fp],
        "fbh_planner_forward": [vp, _fp, C.c_uint32, _fp],
        "fbh_planner_grad": [vp, _fp, _fp, _fp, C.c_uint32, _fp, _fp],
        "fbh_planner_train": [vp, _fp, _fp, _fp, C.c_uint32, C.c_uint32, C.c_uint32, C.c_float,
                              C.c_uint64, _fp],
        "fbh_planner_step": [vp, _fp, _fp, _fp, C.c_uint32, C.c_float],
        "fbh_planner_choose": [vp, _fp, C.c_uint32, _ip, _ip, _u32p],
    }
    for name, args in sig.items():
        f = getattr(L, name)
        f.argtypes = args
        f.restype = C.c_int
    L.fbh_index_free.argtypes = [vp]
    L.fbh_index_free.restype = None
    L.fbh_free.argtypes = [vp]
    L.fbh_free.restype = None
    L.fbh_planner_free.argtypes = [vp]
    L.fbh_planner_free.restype = None


class Index:
    """ibis::relic / ibis::bin with its raw column, in HBM."""

    def __init__(self, handle):
        self._h = handle

    @staticmethod
    def _lib():
        L = lib()
        if not getattr(L, "_index_sigs", False):
            _index_sigs(L)
            L._index_sigs = True
        return L

    @classmethod
    def relic(cls, col):
        L = cls._lib()
        a = np.ascontiguousarray(col)
        h = C.c_void_p()
        check(L.fbh_index_build_relic(a.ctypes.data_as(C.c_void_p), _COLT[a.dtype], a.size,
                                      C.byref(h)), "fbh_index_build_relic")
        return cls(h)

    @classmethod
    def relic_card(cls, ind, card):
        L = cls._lib()
        a = np.ascontiguousarray(ind, np.uint32)
        h = C.c_void_p()
        check(L.fbh_index_build_relic_card(a.ctypes.data_as(_u32p), a.size, int(card), C.byref(h)),
              "fbh_index_build_relic_card")
        return cls(h)

    @classmethod
    def bin(cls, col, bounds):
        L = cls._lib()
        a = np.ascontiguousarray(col)
        b = np.ascontiguousarray(bounds, np.float64)
        h = C.c_void_p()
        check(L.fbh_index_build_bin(a.ctypes.data_as(C.c_void_p), _COLT[a.dtype], a.size,
                                    b.ctypes.data_as(_dp), b.size, C.byref(h)), "fbh_index_build_bin")
        return cls(h)

    @classmethod
    def bin_equal_weight(cls, col, nbins):
        """bin with nbins equal-weight bins, boundaries chosen as the
        reference's "<binning nbins=N equal-weight/>" does."""
        L = cls._lib()
        a = np.ascontiguousarray(col)
        h = C.c_void_p()
        check(L.fbh_index_build_bin_nbins(a.ctypes.data_as(C.c_void_p), _COLT[a.dtype], a.size, int(nbins),
                                          C.byref(h)), "fbh_index_build_bin_nbins")
        return cls(h)

    @classmethod
    def attach(cls, col, keys, offsets, bms):
        """fastbit_iapi_attach_full_index: an index given as its keys, word
        offsets and serialized bitmaps, over the column it indexes."""
        L = cls._lib()
        a = np.ascontiguousarray(col)
        k = np.ascontiguousarray(keys, np.float64)
        o = np.ascontiguousarray(offsets, np.int64)
        b = np.ascontiguousarray(bms, np.uint32)
        h = C.c_void_p()
        check(L.fbh_index_attach(a.ctypes.data_as(C.c_void_p), _COLT[a.dtype], a.size, k.ctypes.data_as(_dp),
                                 k.size, o.ctypes.data_as(_i64p), o.size, b.ctypes.data_as(_u32p), b.size,
                                 C.byref(h)), "fbh_index_attach")
        return cls(h)

    @classmethod
    def fuzz(cls, col, ncoarse=0):
        """ibis::fuzz over a column: the relic and an interval-encoded coarse
        level (ncoarse: the index spec's "ncoarse=", 0 for the default)."""
        L = cls._lib()
        a = np.ascontiguousarray(col)
        h = C.c_void_p()
        check(L.fbh_index_build_fuzz(a.ctypes.data_as(C.c_void_p), _COLT[a.dtype], a.size, int(ncoarse),
                                     C.byref(h)), "fbh_index_build_fuzz")
        return cls(h)

    @classmethod
    def zona(cls, col):
        """ibis::zona over a column: the relic and 11 equality-encoded coarse
        bins."""
        L = cls._lib()
        a = np.ascontiguousarray(col)
        h = C.c_void_p()
        check(L.fbh_index_build_zona(a.ctypes.data_as(C.c_void_p), _COLT[a.dtype], a.size, C.byref(h)),
              "fbh_index_build_zona")
        return cls(h)

    def coarse(self):
        """(encoding, cbounds, coarse byte offsets, number of coarse bitmaps)."""
        L = self._lib()
        enc, nc, nb = C.c_int(), C.c_uint32(), C.c_uint32()
        check(L.fbh_index_coarse(self._h, C.byref(enc), C.byref(nc), None, 0, None, 0, C.byref(nb)),
              "fbh_index_coarse")
        cb = np.zeros(nc.value + 1, np.uint32)
        co = np.zeros(nb.value + 1, np.int64)
        check(L.fbh_index_coarse(self._h, None, None, cb.ctypes.data_as(_u32p), cb.size if nc.value else 0,
                                 co.ctypes.data_as(_i64p), co.size if nb.value else 0, None), "fbh_index_coarse")
        return enc.value, (cb if nc.value else cb[:0]), (co if nb.value else co[:0]), nb.value

    def coarse_bitmap(self, i):
        h = C.c_void_p()
        check(self._lib().fbh_index_coarse_bitmap(self._h, int(i), C.byref(h)), "fbh_index_coarse_bitmap")
        return Bitvector(h)

    def two_level_option(self, lo, lop, hi, rop):
        """(option, mask of the coarse bitmaps read) of a fuzz / zona range."""
        opt, ld = C.c_int(), C.c_uint64()
        check(self._lib().fbh_index_two_level_option(self._h, lo, lop, hi, rop, C.byref(opt), C.byref(ld)),
              "fbh_index_two_level_option")
        return opt.value, ld.value

    @classmethod
    def read_file(cls, path, col):
        """relic::read / bin::read of an index file, over its column."""
        L = cls._lib()
        a = np.ascontiguousarray(col)
        h = C.c_void_p()
        check(L.fbh_index_read_file(os.fsencode(path), a.ctypes.data_as(C.c_void_p), _COLT[a.dtype], a.size,
                                    C.byref(h)), "fbh_index_read_file")
        return cls(h)

    def sum_bins(self, ib, ie):
        """fastbit_iapi_get_number_of_hits: ones of sumBins(ib, ie)."""
        n = C.c_uint64()
        check(self._lib().fbh_index_sum_bins(self._h, int(ib), int(ie), C.byref(n)), "fbh_index_sum_bins")
        return n.value

    def bounds(self):
        L = self._lib()
        n = C.c_uint32()
        check(L.fbh_index_bounds(self._h, None, 0, C.byref(n)), "fbh_index_bounds")
        out = np.zeros(max(n.value, 1))
        check(L.fbh_index_bounds(self._h, out.ctypes.data_as(_dp), n.value, C.byref(n)), "fbh_index_bounds")
        return out[: n.value]

    def __del__(self):
        try:
            if self._h is not None and _lib is not None:
                _lib.fbh_index_free(self._h)
        except Exception:
            pass
        self._h = None

    def info(self):
        nobs, nrows, nw = C.c_uint32(), C.c_uint32(), C.c_uint64()
        check(self._lib().fbh_index_info(self._h, C.byref(nobs), C.byref(nrows), C.byref(nw)),
              "fbh_index_info")
        return nobs.value, nrows.value, nw.value

    def deconstruct(self):
        L = self._lib()
        k, o, b = _dp(), _i64p(), _u32p()
        nk, no, nb = C.c_uint64(), C.c_uint64(), C.c_uint64()
        check(L.fbh_index_deconstruct(self._h, C.byref(k), C.byref(nk), C.byref(o), C.byref(no),
                                      C.byref(b), C.byref(nb)), "fbh_index_deconstruct")
        out = (np.ctypeslib.as_array(k, (nk.value,)).copy() if nk.value else np.zeros(0),
               np.ctypeslib.as_array(o, (no.value,)).copy(),
               np.ctypeslib.as_array(b, (nb.value,)).copy() if nb.value else np.zeros(0, np.uint32))
        for p in (k, o, b):
            L.fbh_free(p)
        return out

    def resolve_range(self, lo, lop, hi, rop):
        r = [C.c_uint32() for _ in range(4)]
        check(self._lib().fbh_index_resolve_range(self._h, lo, lop, hi, rop, *[C.byref(x) for x in r]),
              "fbh_index_resolve_range")
        return tuple(x.value for x in r)

    def estimate_cost(self, lo, lop, hi, rop):
        ic, sc, plan = C.c_double(), C.c_double(), C.c_int()
        check(self._lib().fbh_index_estimate_cost(self._h, lo, lop, hi, rop, C.byref(ic), C.byref(sc),
                                                  C.byref(plan)), "fbh_index_estimate_cost")
        return ic.value, sc.value, plan.value

    def evaluate(self, lo, lop, hi, rop, plan=PLAN_AUTO, want_bitvector=False):
        """(hits, plan used, result Bitvector or None)."""
        hits = C.c_uint64()
        used = C.c_int()
        bv = C.c_void_p()
        check(self._lib().fbh_index_evaluate(self._h, lo, lop, hi, rop, plan, C.byref(hits),
                                             C.byref(bv) if want_bitvector else None, C.byref(used)),
              "fbh_index_evaluate")
        return hits.value, used.value, (Bitvector(bv) if want_bitvector else None)

    def rids(self, lo, lop, hi, rop, plan=PLAN_LITERAL):
        L = self._lib()
        n = C.c_uint64()
        check(L.fbh_index_evaluate_rids(self._h, lo, lop, hi, rop, plan, None, 0, C.byref(n)),
              "fbh_index_evaluate_rids")
        out = np.zeros(max(n.value, 1), np.uint32)
        check(L.fbh_index_evaluate_rids(self._h, lo, lop, hi, rop, plan, out.ctypes.data_as(_u32p),
                                        n.value, C.byref(n)), "fbh_index_evaluate_rids")
        return out[: n.value]

    def estimate(self, lo, lop, hi, rop):
        """index::estimate(expr, lower, upper): (lower, upper) Bitvectors;
        an upper of size 0 stands for lower."""
        lw, up = C.c_void_p(), C.c_void_p()
        check(self._lib().fbh_index_estimate(self._h, lo, lop, hi, rop, C.byref(lw), C.byref(up)),
              "fbh_index_estimate")
        return Bitvector(lw), Bitvector(up)

    def evaluate_noalloc(self, lo, lop, hi, rop, plan, dev_out, out_words):
        """fbh_index_evaluate_noalloc: the literal result into the device
        buffer at address dev_out (out_words words); (hits, plan used)."""
        hits = C.c_uint64()
        used = C.c_int()
        check(self._lib().fbh_index_evaluate_noalloc(self._h, lo, lop, hi, rop, int(plan), C.c_void_p(dev_out),
                                                     int(out_words), C.byref(hits), C.byref(used)),
              "fbh_index_evaluate_noalloc")
        return hits.value, used.value

    def evaluate_batch(self, lo, lop, hi, rop, plan=PLAN_AUTO, plans=None):
        """Evaluate many ranges back to back on the device.

        lo/hi are arrays, lop/rop arrays or scalars; `plans` (optional) picks
        a plan per query.  Returns (hits uint64[nq], plans used int32[nq],
        device milliseconds from the first to the last query)."""
        lo = np.ascontiguousarray(lo, np.float64)
        nq = lo.size
        hi = np.ascontiguousarray(np.broadcast_to(hi, (nq,)), np.float64)
        lop = np.ascontiguousarray(np.broadcast_to(lop, (nq,)), np.int32)
        rop = np.ascontiguousarray(np.broadcast_to(rop, (nq,)), np.int32)
        pp = None if plans is None else np.ascontiguousarray(np.broadcast_to(plans, (nq,)), np.int32)
        hits = np.zeros(nq, np.uint64)
        used = np.zeros(nq, np.int32)
        ms = C.c_float()
        check(self._lib().fbh_index_evaluate_batch(
            self._h, nq, lo.ctypes.data_as(_dp), lop.ctypes.data_as(_ip), hi.ctypes.data_as(_dp),
            rop.ctypes.data_as(_ip), None if pp is None else pp.ctypes.data_as(_ip), int(plan),
            hits.ctypes.data_as(_u64p), used.ctypes.data_as(_ip), C.byref(ms)),
            "fbh_index_evaluate_batch")
        return hits, used, ms.value

    def query_bytes(self, lo, lop, hi, rop, plan):
        """Algorithmic bytes `plan` moves for the range."""
        v = C.c_double()
        check(self._lib().fbh_index_query_bytes(self._h, lo, lop, hi, rop, int(plan), C.byref(v)),
              "fbh_index_query_bytes")
        return v.value

    def set_cost_model(self, model):
        """COST_IAPI (in-memory arrays, the default) or COST_PARTITION."""
        check(self._lib().fbh_index_set_cost_model(self._h, int(model)), "fbh_index_set_cost_model")

    def column_bytes(self):
        v = C.c_uint64()
        check(self._lib().fbh_index_column_bytes(self._h, C.byref(v)), "fbh_index_column_bytes")
        return v.value

    @staticmethod
    def _queries(lo, lop, hi, rop):
        lo = np.ascontiguousarray(lo, np.float64)
        nq = lo.size
        hi = np.ascontiguousarray(np.broadcast_to(hi, (nq,)), np.float64)
        lop = np.ascontiguousarray(np.broadcast_to(lop, (nq,)), np.int32)
        rop = np.ascontiguousarray(np.broadcast_to(rop, (nq,)), np.int32)
        return nq, lo, lop, hi, rop

    def plan_features(self, lo, lop, hi, rop):
        """float32[nq, NFEATURES] plan-choice features of the ranges."""
        nq, lo, lop, hi, rop = self._queries(lo, lop, hi, rop)
        f = np.zeros((nq, NFEATURES), np.float32)
        check(self._lib().fbh_index_plan_features(
            self._h, nq, lo.ctypes.data_as(_dp), lop.ctypes.data_as(_ip), hi.ctypes.data_as(_dp),
            rop.ctypes.data_as(_ip), f.ctypes.data_as(_fp)), "fbh_index_plan_features")
        return f

    def time_plans(self, lo, lop, hi, rop, cap=2.0):
        """(ms float32[nq, 3], hits uint64[nq, 3]) per plan (literal,
        compressed, scan); a censored compressed time is negative."""
        nq, lo, lop, hi, rop = self._queries(lo, lop, hi, rop)
        ms = np.zeros((nq, NPLANS), np.float32)
        hits = np.zeros((nq, NPLANS), np.uint64)
        check(self._lib().fbh_index_time_plans(
            self._h, nq, lo.ctypes.data_as(_dp), lop.ctypes.data_as(_ip), hi.ctypes.data_as(_dp),
            rop.ctypes.data_as(_ip), float(cap), ms.ctypes.data_as(_fp), hits.ctypes.data_as(_u64p)),
            "fbh_index_time_plans")
        return ms, hits

    def choose_plans(self, planner, lo, lop, hi, rop):
        """(plans int32[nq], number of heuristic fallbacks) from a Planner."""
        nq, lo, lop, hi, rop = self._queries(lo, lop, hi, rop)
        plans = np.zeros(nq, np.int32)
        nf = C.c_uint32()
        check(self._lib().fbh_index_choose_plans(
            self._h, planner._h, nq, lo.ctypes.data_as(_dp), lop.ctypes.data_as(_ip),
            hi.ctypes.data_as(_dp), rop.ctypes.data_as(_ip), plans.ctypes.data_as(_ip), C.byref(nf)),
            "fbh_index_choose_plans")
        return plans, nf.value


# ---------------------------------------------------------------------------
# Conjunctions
class _Term(C.Structure):
    _fields_ = [("ix", C.c_void_p), ("lo", C.c_double), ("lop", C.c_int), ("hi", C.c_double),
                ("rop", C.c_int)]


def _conj_sigs(L):
    vp = C.c_void_p
    for name, args in {
        "fbh_conj_prepare": [C.POINTER(_Term), _u32p, C.c_uint32, C.c_int, C.POINTER(vp)],
        "fbh_conj_run": [vp, _u64p, C.POINTER(C.c_float)],
        "fbh_conj_tune": [vp, C.POINTER(C.c_int), _u32p, _u32p, C.POINTER(C.c_float)],
        "fbh_conj_plans": [vp, C.POINTER(C.c_int)],
        "fbh_conj_rids": [vp, _u64p, _u32p, C.c_uint64],
        "fbh_conj_result": [vp, C.c_uint32, C.POINTER(vp)],
        "fbh_conj_evaluate": [C.POINTER(_Term), C.c_uint32, C.c_int, _u64p],
        "fbh_expr_prepare": [C.POINTER(_Term), _u32p, C.POINTER(C.c_int32), _u32p, C.c_uint32, C.c_int,
                             C.POINTER(vp)],
        "fbh_index_batch_prepare": [vp, C.c_uint32, _dp, _ip, _dp, _ip, _ip, C.c_int, C.POINTER(vp)],
    }.items():
        f = getattr(L, name)
        f.argtypes = args
        f.restype = C.c_int
    L.fbh_conj_free.argtypes = [vp]
    L.fbh_conj_free.restype = None


class Conjunctions:
    """A prepared batch of conjunctive queries (fbh_conj_prepare): queries
    is a list of lists of (Index, lo, lop, hi, rop) terms."""

    def __init__(self, queries, plan=PLAN_AUTO):
        L = Index._lib()
        if not getattr(L, "_conj_sigs", False):
            _conj_sigs(L)
            L._conj_sigs = True
        self._L = L
        self._keep = []
        terms, qoff = [], [0]
        for q in queries:
            for (ix, lo, lop, hi, rop) in q:
                terms.append(_Term(ix._h, float(lo), int(lop), float(hi), int(rop)))
                self._keep.append(ix)
            qoff.append(len(terms))
        self.nq = len(queries)
        self.nterms = len(terms)
        arr = (_Term * len(terms))(*terms)
        off = np.ascontiguousarray(qoff, np.uint32)
        h = C.c_void_p()
        check(L.fbh_conj_prepare(arr, off.ctypes.data_as(_u32p), self.nq, int(plan), C.byref(h)),
              "fbh_conj_prepare")
        self._h = h

    def run(self):
        """(hits uint64[nq], device milliseconds of the replay)."""
        hits = np.zeros(self.nq, np.uint64)
        ms = C.c_float()
        check(self._L.fbh_conj_run(self._h, hits.ctypes.data_as(_u64p), C.byref(ms)), "fbh_conj_run")
        return hits, ms.value

    def tune(self):
        nt, blocks, tpi, ms = C.c_int(), C.c_uint32(), C.c_uint32(), C.c_float()
        check(self._L.fbh_conj_tune(self._h, C.byref(nt), C.byref(blocks), C.byref(tpi), C.byref(ms)),
              "fbh_conj_tune")
        return dict(threads=nt.value, blocks=blocks.value, tiles_per_item=tpi.value, ms=ms.value)

    def plans(self):
        out = np.zeros(self.nterms, np.int32)
        check(self._L.fbh_conj_plans(self._h, out.ctypes.data_as(C.POINTER(C.c_int))), "fbh_conj_plans")
        return out

    def rids(self):
        """(offsets uint64[nq+1], rids uint32[offsets[-1]]): the row ids of
        every query, query q's at rids[offsets[q]:offsets[q+1]]."""
        off = np.zeros(self.nq + 1, np.uint64)
        check(self._L.fbh_conj_rids(self._h, off.ctypes.data_as(_u64p), None, 0), "fbh_conj_rids")
        out = np.zeros(max(int(off[-1]), 1), np.uint32)
        check(self._L.fbh_conj_rids(self._h, off.ctypes.data_as(_u64p), out.ctypes.data_as(_u32p),
                                    int(off[-1])), "fbh_conj_rids")
        return off, out[:int(off[-1])]

    def result(self, q):
        """The hit vector of query q as a Bitvector (canonical WAH)."""
        h = C.c_void_p()
        check(self._L.fbh_conj_result(self._h, int(q), C.byref(h)), "fbh_conj_result")
        return Bitvector(h)

    def __del__(self):
        try:
            if self._h is not None:
                self._L.fbh_conj_free(self._h)
        except Exception:
            pass
        self._h = None


# ---------------------------------------------------------------------------
# Learned plan choice
def _f32(a, shape):
    return np.ascontiguousarray(np.asarray(a, np.float32).reshape(shape))



EXPR_AND, EXPR_OR, EXPR_XOR, EXPR_NAND, EXPR_NOR, EXPR_NOT, EXPR_MINUS = -1, -2, -3, -4, -5, -6, -7


class RangeBatch(Conjunctions):
    """A prepared batch of single range queries on one index
    (fbh_index_batch_prepare): one persistent, graph-captured launch;
    run / tune / plans as for Conjunctions."""

    def __init__(self, ix, lo, lop, hi, rop, plan=PLAN_AUTO, plans=None):
        L = Index._lib()
        if not getattr(L, "_conj_sigs", False):
            _conj_sigs(L)
            L._conj_sigs = True
        self._L = L
        self._keep = [ix]
        nq, lo, lop, hi, rop = Index._queries(lo, lop, hi, rop)
        pp = None if plans is None else np.ascontiguousarray(np.broadcast_to(plans, (nq,)), np.int32)
        self._arrays = (lo, lop, hi, rop, pp)
        self.nq = nq
        self.nterms = nq
        h = C.c_void_p()
        self._h = None
        check(L.fbh_index_batch_prepare(ix._h, nq, lo.ctypes.data_as(_dp), lop.ctypes.data_as(_ip),
                                        hi.ctypes.data_as(_dp), rop.ctypes.data_as(_ip),
                                        None if pp is None else pp.ctypes.data_as(_ip), int(plan), C.byref(h)),
              "fbh_index_batch_prepare")
        self._h = h


class Expressions(Conjunctions):
    """A prepared batch of boolean expressions over range terms
    (fbh_expr_prepare).  queries is a list of (terms, program): terms a list
    of (Index, lo, lop, hi, rop), program a postfix list of term positions
    (>= 0) and EXPR_* operators."""

    def __init__(self, queries, plan=PLAN_AUTO):
        L = Index._lib()
        if not getattr(L, "_conj_sigs", False):
            _conj_sigs(L)
            L._conj_sigs = True
        self._L = L
        self._keep = []
        terms, toff, prog, poff = [], [0], [], [0]
        for qterms, qprog in queries:
            for (ix, lo, lop, hi, rop) in qterms:
                terms.append(_Term(ix._h, float(lo), int(lop), float(hi), int(rop)))
                self._keep.append(ix)
            toff.append(len(terms))
            prog.extend(int(x) for x in qprog)
            poff.append(len(prog))
        self.nq = len(queries)
        self.nterms = len(terms)
        arr = (_Term * max(len(terms), 1))(*terms)
        to = np.ascontiguousarray(toff, np.uint32)
        pr = np.ascontiguousarray(prog if prog else [0], np.int32)
        po = np.ascontiguousarray(poff, np.uint32)
        h = C.c_void_p()
        self._h = None
        check(L.fbh_expr_prepare(arr, to.ctypes.data_as(_u32p), pr.ctypes.data_as(C.POINTER(C.c_int32)),
                                 po.ctypes.data_as(_u32p), self.nq, int(plan), C.byref(h)), "fbh_expr_prepare")
        self._h = h

class Planner:
    """The FP32 MLP cost model of the plan choice (fbh_planner_*)."""

    def __init__(self, seed=0):
        self._L = Index._lib()
        h = C.c_void_p()
        check(self._L.fbh_planner_new(int(seed), C.byref(h)), "fbh_planner_new")
        self._h = h

    def __del__(self):
        try:
            if self._h is not None:
                self._L.fbh_planner_free(self._h)
        except Exception:
            pass
        self._h = None

    @property
    def params(self):
        out = np.zeros(NPARAM, np.float32)
        check(self._L.fbh_planner_get_params(self._h, out.ctypes.data_as(_fp), NPARAM),
              "fbh_planner_get_params")
        return out

    @params.setter
    def params(self, v):
        a = _f32(v, (-1,))
        check(self._L.fbh_planner_set_params(self._h, a.ctypes.data_as(_fp), a.size),
              "fbh_planner_set_params")

    def set_norm(self, mu, sd):
        mu, sd = _f32(mu, (NFEATURES,)), _f32(sd, (NFEATURES,))
        check(self._L.fbh_planner_set_norm(self._h, mu.ctypes.data_as(_fp), sd.ctypes.data_as(_fp)),
              "fbh_planner_set_norm")

    def norm(self):
        mu, sd = np.zeros(NFEATURES, np.float32), np.zeros(NFEATURES, np.float32)
        check(self._L.fbh_planner_get_norm(self._h, mu.ctypes.data_as(_fp), sd.ctypes.data_as(_fp)),
              "fbh_planner_get_norm")
        return mu, sd

    def forward(self, x):
        x = _f32(x, (-1, NFEATURES))
        out = np.zeros((x.shape[0], NPLANS), np.float32)
        check(self._L.fbh_planner_forward(self._h, x.ctypes.data_as(_fp), x.shape[0],
                                          out.ctypes.data_as(_fp)), "fbh_planner_forward")
        return out

    def grad(self, x, y, mask):
        x = _f32(x, (-1, NFEATURES))
        n = x.shape[0]
        y, m = _f32(y, (n, NPLANS)), _f32(mask, (n, NPLANS))
        g = np.zeros(NPARAM, np.float32)
        loss = C.c_float()
        check(self._L.fbh_planner_grad(self._h, x.ctypes.data_as(_fp), y.ctypes.data_as(_fp),
                                       m.ctypes.data_as(_fp), n, C.byref(loss), g.ctypes.data_as(_fp)),
              "fbh_planner_grad")
        return loss.value, g

    def step(self, x, y, mask, lr):
        x = _f32(x, (-1, NFEATURES))
        n = x.shape[0]
        y, m = _f32(y, (n, NPLANS)), _f32(mask, (n, NPLANS))
        check(self._L.fbh_planner_step(self._h, x.ctypes.data_as(_fp), y.ctypes.data_as(_fp),
                                       m.ctypes.data_as(_fp), n, float(lr)), "fbh_planner_step")

    def train(self, x, y, mask, epochs=40, batch=256, lr=1e-3, seed=0):
        """Trains on the device; returns float32[epochs] mean losses."""
        x = _f32(x, (-1, NFEATURES))
        n = x.shape[0]
        y, m = _f32(y, (n, NPLANS)), _f32(mask, (n, NPLANS))
        hist = np.zeros(epochs, np.float32)
        check(self._L.fbh_planner_train(self._h, x.ctypes.data_as(_fp), y.ctypes.data_as(_fp),
                                        m.ctypes.data_as(_fp), n, int(epochs), int(batch), float(lr),
                                        int(seed), hist.ctypes.data_as(_fp)), "fbh_planner_train")
        return hist

    def choose(self, x, fallback):
        """(plans int32[n], fallbacks) for feature rows x."""
        x = _f32(x, (-1, NFEATURES))
        n = x.shape[0]
        fb = np.ascontiguousarray(np.broadcast_to(fallback, (n,)), np.int32)
        plans = np.zeros(n, np.int32)
        nf = C.c_uint32()
        check(self._L.fbh_planner_choose(self._h, x.ctypes.data_as(_fp), n, fb.ctypes.data_as(_ip),
                                         plans.ctypes.data_as(_ip), C.byref(nf)), "fbh_planner_choose")
        return plans, nf.value


def plan_labels(ms):
    """Training targets of timed queries: y = log2(microseconds) per plan,
    mask 1 for a measured time, -1 for a censored (lower-bound) one."""
    ms = np.asarray(ms, np.float64)
    us = np.maximum(np.abs(ms) * 1000.0, 1e-3)
    y = np.log2(us).astype(np.float32)
    mask = np.where(ms < 0, -1.0, 1.0).astype(np.float32)
    return y, mask
