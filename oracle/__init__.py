"""CPU oracle for the FastBit query hot path.

TEST INFRASTRUCTURE ONLY.  Only ``tests/``, ``__graft_entry__.smoke()`` and the
``cpu_baseline`` leg of ``bench.py`` may import this package; the product
library (``libfastlib_amd``) never calls it.

Two checkers live here:

* ``liboracle.so`` -- a C++ restatement of the reference algorithms
  (``wah.cpp`` restates ``src/bitvector.{h,cpp}``; ``index.cpp`` restates the
  relic/bin build, locate, sumBins and edge-check paths).  Each function cites
  the reference file:line it follows.
* ``_ref/libfbref.so`` -- the reference library itself, compiled from its
  sources by ``oracle/Makefile`` (target ``ref``) where the source tree is
  present.  It pins the restatement; it is optional at run time.

Bitvectors cross the ctypes boundary in the reference's serialized form
(``src/bitvector.cpp:2283-2313``): the words, then ``active.val`` when
``active.nbits > 0``, then ``active.nbits``.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
ORACLE_SO = os.path.join(_HERE, "liboracle.so")
REF_SO = os.path.join(_HERE, "_ref", "libfbref.so")

OP_AND, OP_OR, OP_XOR, OP_MINUS = 0, 1, 2, 3
PATH_NAMES = {0: "none", 1: "c0", 2: "d1", 3: "d1swap", 4: "c1", 5: "d2",
              6: "c2", 7: "shortcut"}

_u32p = C.POINTER(C.c_uint32)


def _arr(a):
    a = np.ascontiguousarray(a, dtype=np.uint32)
    return a, a.ctypes.data_as(_u32p), C.c_size_t(a.size)


class _Lib:
    prefix = ""

    def __init__(self, path):
        self.path = path
        self.lib = C.CDLL(path)
        self._free = getattr(self.lib, self.prefix + "free")
        self._free.argtypes = [C.c_void_p]

    def _take(self, out, n):
        res = np.ctypeslib.as_array(out, shape=(n.value,)).copy() if n.value else \
            np.zeros(0, dtype=np.uint32)
        self._free(out)
        return res.astype(np.uint32)

    def _call_out(self, name, *args):
        fn = getattr(self.lib, self.prefix + name)
        out = _u32p()
        n = C.c_size_t()
        rc = fn(*args, C.byref(out), C.byref(n))
        if rc != 0:
            raise RuntimeError(f"{self.prefix}{name} failed rc={rc}")
        return self._take(out, n)

    def compress(self, a):
        a, p, n = _arr(a)
        return self._call_out("wah_compress", p, n)

    def decompress(self, a):
        a, p, n = _arr(a)
        return self._call_out("wah_decompress", p, n)

    def flip(self, a):
        a, p, n = _arr(a)
        return self._call_out("wah_flip", p, n)

    def rids(self, a):
        a, p, n = _arr(a)
        return self._call_out("wah_rids", p, n)

    def from_bits(self, bits):
        b = np.ascontiguousarray(bits, dtype=np.uint8)
        return self._call_out("wah_from_bits", b.ctypes.data_as(C.POINTER(C.c_uint8)),
                              C.c_size_t(b.size))

    def count(self, a, mask):
        """bitvector::count(mask): the ones of a AND mask."""
        a, pa, na = _arr(a)
        b, pb, nb = _arr(mask)
        c = C.c_uint32()
        rc = getattr(self.lib, self.prefix + "wah_count")(pa, na, pb, nb, C.byref(c))
        if rc != 0:
            raise RuntimeError(f"{self.prefix}wah_count failed rc={rc}")
        return c.value

    def erase(self, a, i, j):
        """bitvector::erase(i, j): the words after removing bits [i, j)."""
        a, p, n = _arr(a)
        return self._call_out("wah_erase", p, n, C.c_uint32(i), C.c_uint32(j))

    def cnt(self, a):
        """Return (cnt, size, words after the call)."""
        a, p, n = _arr(a)
        c = C.c_uint32()
        s = C.c_uint32()
        fn = getattr(self.lib, self.prefix + "wah_cnt")
        out = _u32p()
        nn = C.c_size_t()
        rc = fn(p, n, C.byref(c), C.byref(s), C.byref(out), C.byref(nn))
        if rc != 0:
            raise RuntimeError("wah_cnt failed")
        return c.value, s.value, self._take(out, nn)


class Oracle(_Lib):
    """The C++ restatement (always available once ``make -C oracle`` ran)."""
    prefix = "orc_"

    def op(self, a, b, op, compress=True):
        """a OP= b (then compress); returns (serialized words, path name)."""
        a, pa, na = _arr(a)
        b, pb, nb = _arr(b)
        path = C.c_int(-1)
        fn = self.lib.orc_wah_op
        out = _u32p()
        n = C.c_size_t()
        rc = fn(pa, na, pb, nb, C.c_int(op), C.c_int(int(compress)), C.byref(out),
                C.byref(n), C.byref(path))
        if rc != 0:
            raise RuntimeError(f"orc_wah_op failed rc={rc}")
        return self._take(out, n), PATH_NAMES.get(path.value, "?")

    def op_new(self, a, b, op):
        """The out-of-place operator (a new bitvector, not compressed);
        returns (serialized words, path name)."""
        a, pa, na = _arr(a)
        b, pb, nb = _arr(b)
        path = C.c_int(-1)
        out = _u32p()
        n = C.c_size_t()
        rc = self.lib.orc_wah_op_new(pa, na, pb, nb, C.c_int(op), C.byref(out), C.byref(n), C.byref(path))
        if rc != 0:
            raise RuntimeError(f"orc_wah_op_new failed rc={rc}")
        return self._take(out, n), PATH_NAMES.get(path.value, "?")

    def from_runs(self, first_bit, lens):
        lens = np.ascontiguousarray(lens, dtype=np.uint32)
        return self._call_out("wah_from_runs", C.c_int(int(first_bit)),
                              lens.ctypes.data_as(_u32p), C.c_size_t(lens.size))

    def literal(self, a):
        a, p, n = _arr(a)
        return self._call_out("wah_literal", p, n)


class Reference(_Lib):
    """The reference FastBit library compiled from source (optional)."""
    prefix = "ref_"

    def op(self, a, b, op, compress=True):
        a, pa, na = _arr(a)
        b, pb, nb = _arr(b)
        return self._call_out("wah_op", pa, na, pb, nb, C.c_int(op), C.c_int(int(compress)))

    def op_new(self, a, b, op):
        a, pa, na = _arr(a)
        b, pb, nb = _arr(b)
        return self._call_out("wah_op_new", pa, na, pb, nb, C.c_int(op))

    def time_op(self, a, b, op, reps=3):
        """(mean ms of a OP= b; a.compress() on copies of a, result words)."""
        a, pa, na = _arr(a)
        b, pb, nb = _arr(b)
        ms = C.c_double()
        nw = C.c_uint32()
        rc = self.lib.ref_wah_time_op(pa, na, pb, nb, C.c_int(op), C.c_int(reps), C.byref(ms), C.byref(nw))
        if rc != 0:
            raise RuntimeError(f"ref_wah_time_op failed rc={rc}")
        return ms.value, nw.value


_oracle = None
_ref = None


def oracle() -> Oracle:
    global _oracle
    if _oracle is None:
        if not os.path.exists(ORACLE_SO):
            raise RuntimeError("oracle/liboracle.so is not built (run make -C oracle)")
        _oracle = Oracle(ORACLE_SO)
    return _oracle


def reference():
    """The compiled reference, or None where it was not built."""
    global _ref
    if _ref is None and os.path.exists(REF_SO):
        _ref = Reference(REF_SO)
    return _ref


# ---------------------------------------------------------------------------
# relic / bin oracle (index.cpp) and the reference's index API.
COL_I32, COL_U32, COL_U16, COL_F32, COL_F64 = 0, 1, 2, 3, 4
_COL_NP = {COL_I32: np.int32, COL_U32: np.uint32, COL_U16: np.uint16, COL_F32: np.float32,
           COL_F64: np.float64}
_dp = C.POINTER(C.c_double)
_i64p = C.POINTER(C.c_int64)


def col_type_of(arr):
    a = np.asarray(arr)
    for k, t in _COL_NP.items():
        if a.dtype == t:
            return k
    raise TypeError(a.dtype)


# Cost models of the plan choice: arrays registered through the in-memory
# API (src/bord.cpp:5825-6058) and partitions with index files
# (src/column.cpp:6029-6207).
COST_IAPI, COST_PARTITION = 0, 1


class OracleIndex:
    """ibis::relic / ibis::bin restated on the CPU (oracle/index.cpp)."""

    def __init__(self, handle, col):
        self._h = handle
        self.col = np.ascontiguousarray(col)
        self.type = col_type_of(self.col)

    @classmethod
    def relic_card(cls, ind, card):
        L = oracle().lib
        L.orc_relic_card.restype = C.c_void_p
        a = np.ascontiguousarray(ind, np.uint32)
        return cls(L.orc_relic_card(a.ctypes.data_as(_u32p), C.c_uint32(a.size), C.c_uint32(card)), a)

    @classmethod
    def relic_values(cls, col):
        L = oracle().lib
        L.orc_relic_values.restype = C.c_void_p
        v = np.ascontiguousarray(col, np.float64)
        return cls(L.orc_relic_values(v.ctypes.data_as(_dp), C.c_uint32(v.size)), col)

    @classmethod
    def bin_bounds(cls, col, bounds):
        L = oracle().lib
        L.orc_bin_bounds.restype = C.c_void_p
        v = np.ascontiguousarray(col, np.float64)
        b = np.ascontiguousarray(bounds, np.float64)
        return cls(L.orc_bin_bounds(v.ctypes.data_as(_dp), C.c_uint32(v.size), b.ctypes.data_as(_dp),
                                    C.c_uint32(b.size)), col)

    def as_file(self):
        """Marks the index as one read from a file: its bitmaps are measured
        by their serialized sizes (the file's positions)."""
        oracle().lib.orc_index_set_offsets_built(C.c_void_p(self._h), C.c_int(1))
        return self

    def __del__(self):
        try:
            oracle().lib.orc_index_free(C.c_void_p(self._h))
        except Exception:
            pass

    def write(self):
        """(keys, offsets, bitmaps) as relic::write / bin::write give them."""
        L = oracle().lib
        k, o, b = _dp(), _i64p(), _u32p()
        nk, no, nb = C.c_size_t(), C.c_size_t(), C.c_size_t()
        L.orc_index_write(C.c_void_p(self._h), C.byref(k), C.byref(nk), C.byref(o), C.byref(no),
                          C.byref(b), C.byref(nb))
        out = (np.ctypeslib.as_array(k, (nk.value,)).copy() if nk.value else np.zeros(0),
               np.ctypeslib.as_array(o, (no.value,)).copy() if no.value else np.zeros(0, np.int64),
               np.ctypeslib.as_array(b, (nb.value,)).copy() if nb.value else np.zeros(0, np.uint32))
        for p in (k, o, b):
            L.orc_free(p)
        return out

    def bounds(self):
        L = oracle().lib
        p, n = _dp(), C.c_size_t()
        L.orc_index_bounds(C.c_void_p(self._h), C.byref(p), C.byref(n))
        out = np.ctypeslib.as_array(p, (n.value,)).copy()
        L.orc_free(p)
        return out

    def evaluate(self, lo, lop, hi, rop):
        """(hit count, ascending rids)."""
        L = oracle().lib
        cnt = C.c_uint64()
        r, nr = _u32p(), C.c_size_t()
        rc = L.orc_index_eval(C.c_void_p(self._h), self.col.ctypes.data_as(C.c_void_p), self.type,
                              C.c_double(lo), C.c_int(lop), C.c_double(hi), C.c_int(rop),
                              C.byref(cnt), C.byref(r), C.byref(nr))
        if rc != 0:
            raise RuntimeError("orc_index_eval failed")
        rids = np.ctypeslib.as_array(r, (nr.value,)).copy() if nr.value else np.zeros(0, np.uint32)
        L.orc_free(r)
        return int(cnt.value), rids

    def estimate(self, lo, lop, hi, rop):
        """index::estimate(expr, lower, upper): the serialized words of lower
        and upper (a cleared upper is [0])."""
        L = oracle().lib
        lw, uw = _u32p(), _u32p()
        nl, nu = C.c_size_t(), C.c_size_t()
        rc = L.orc_index_estimate(C.c_void_p(self._h), self.col.ctypes.data_as(C.c_void_p), self.type,
                                  C.c_double(lo), C.c_int(lop), C.c_double(hi), C.c_int(rop),
                                  C.byref(lw), C.byref(nl), C.byref(uw), C.byref(nu))
        if rc != 0:
            raise RuntimeError("orc_index_estimate failed")
        o = oracle()
        return o._take(lw, nl), o._take(uw, nu)

    def evaluate_plan(self, lo, lop, hi, rop, plan=0, model=COST_IAPI):
        """(hit count, ascending rids, plan used) under `plan`: 0 the
        reference's own choice under the cost `model`, 1 the index, 3 the
        raw scan."""
        L = oracle().lib
        cnt = C.c_uint64()
        used = C.c_int()
        r, nr = _u32p(), C.c_size_t()
        rc = L.orc_index_eval_plan(C.c_void_p(self._h), self.col.ctypes.data_as(C.c_void_p), self.type,
                                   C.c_double(lo), C.c_int(lop), C.c_double(hi), C.c_int(rop),
                                   C.c_int(plan), C.c_int(model), C.byref(cnt), C.byref(r),
                                   C.byref(nr), C.byref(used))
        if rc != 0:
            raise RuntimeError("orc_index_eval_plan failed")
        rids = np.ctypeslib.as_array(r, (nr.value,)).copy() if nr.value else np.zeros(0, np.uint32)
        L.orc_free(r)
        return int(cnt.value), rids, int(used.value)

    def cost(self, lo, lop, hi, rop, model=COST_IAPI):
        """(index cost, scan cost, plan, empty) of the reference's choice."""
        L = oracle().lib
        ic, sc = C.c_double(), C.c_double()
        plan, empty = C.c_int(), C.c_int()
        L.orc_index_cost(C.c_void_p(self._h), self.col.ctypes.data_as(C.c_void_p), self.type,
                         C.c_double(lo), C.c_int(lop), C.c_double(hi), C.c_int(rop), C.c_int(model),
                         C.byref(ic), C.byref(sc), C.byref(plan), C.byref(empty))
        return ic.value, sc.value, plan.value, bool(empty.value)


# FastBitDataType codes of the reference's iapi (src/iapi.h)
_FB_TYPE = {COL_I32: 5, COL_U32: 6, COL_U16: 4, COL_F32: 9, COL_F64: 10}


def ref_relic_card(ind, card):
    r = reference()
    L = r.lib
    a = np.ascontiguousarray(ind, np.uint32)
    k, o, b = _dp(), _i64p(), _u32p()
    nk, no, nb = C.c_size_t(), C.c_size_t(), C.c_size_t()
    rc = L.ref_relic_card(a.ctypes.data_as(_u32p), C.c_size_t(a.size), C.c_uint32(card), C.byref(k),
                          C.byref(nk), C.byref(o), C.byref(no), C.byref(b), C.byref(nb))
    if rc != 0:
        raise RuntimeError("ref_relic_card failed")
    out = (np.ctypeslib.as_array(k, (nk.value,)).copy(), np.ctypeslib.as_array(o, (no.value,)).copy(),
           np.ctypeslib.as_array(b, (nb.value,)).copy())
    for p in (k, o, b):
        L.ref_free(p)
    return out


# The reference's array registry (src/iapi.cpp:1644-1665) identifies arrays
# by name and by buffer address and never forgets them: registered buffers
# are kept alive for the life of the process so that no later array reuses
# an address the registry still maps to an older name.
_REGISTERED = []


class RefIapi:
    """One array registered in the reference's in-memory API."""

    def __init__(self, name, col, spec):
        self.name = name.encode()
        self.col = np.ascontiguousarray(col)
        _REGISTERED.append(self.col)
        L = reference().lib
        L.ref_iapi_query.restype = C.c_int64
        k, o, b = _dp(), _i64p(), _u32p()
        nk, no, nb = C.c_uint64(), C.c_uint64(), C.c_uint64()
        rc = L.ref_iapi_index(self.name, _FB_TYPE[col_type_of(self.col)],
                              self.col.ctypes.data_as(C.c_void_p), C.c_uint64(self.col.size),
                              spec.encode(), C.byref(k), C.byref(nk), C.byref(o), C.byref(no),
                              C.byref(b), C.byref(nb))
        if rc < 0:
            raise RuntimeError(f"ref_iapi_index failed rc={rc}")
        self.keys = np.ctypeslib.as_array(k, (nk.value,)).copy()
        self.offsets = np.ctypeslib.as_array(o, (no.value,)).copy()
        self.bitmaps = np.ctypeslib.as_array(b, (nb.value,)).copy()
        for p in (k, o, b):
            L.ref_free(p)

    def estimate(self, lo, lop, hi, rop):
        return ref_iapi_estimate(self.name, lo, lop, hi, rop)

    def query_range(self, lo, lop, hi, rop, want_rids=True):
        """lo OP1 x OP2 hi as one range term (as a where clause gives it)."""
        L = reference().lib
        L.ref_iapi_query_range.restype = C.c_int64
        cap = self.col.size if want_rids else 0
        buf = np.zeros(max(cap, 1), np.uint64)
        n = L.ref_iapi_query_range(self.name, C.c_double(lo), C.c_int(lop), C.c_double(hi),
                                   C.c_int(rop),
                                   buf.ctypes.data_as(C.POINTER(C.c_uint64)) if want_rids else None,
                                   C.c_uint64(cap))
        if n < 0:
            raise RuntimeError(f"ref_iapi_query_range failed {n}")
        return int(n), (buf[:n].astype(np.uint32) if want_rids else None)

    def query(self, lo, lop, hi, rop, want_rids=True):
        """lo OP1 x AND x OP2 hi: two one-sided selections combined with
        fastbit_selection_combine (each term picks its own plan)."""
        L = reference().lib
        cap = self.col.size if want_rids else 0
        buf = np.zeros(max(cap, 1), np.uint64)
        n = L.ref_iapi_query(self.name, C.c_double(lo), C.c_int(lop), C.c_double(hi), C.c_int(rop),
                             buf.ctypes.data_as(C.POINTER(C.c_uint64)) if want_rids else None,
                             C.c_uint64(cap))
        if n < 0:
            raise RuntimeError(f"ref_iapi_query failed {n}")
        return int(n), (buf[:n].astype(np.uint32) if want_rids else None)


def ref_iapi_estimate(name, lo, lop, hi, rop):
    """The reference's index::estimate(expr, lower, upper) on the index of a
    registered array: the serialized words of lower and upper."""
    R = reference()
    L = R.lib
    lw, uw = _u32p(), _u32p()
    nl, nu = C.c_size_t(), C.c_size_t()
    nm = name if isinstance(name, bytes) else name.encode()
    rc = L.ref_iapi_estimate(nm, C.c_double(lo), C.c_int(lop), C.c_double(hi), C.c_int(rop),
                             C.byref(lw), C.byref(nl), C.byref(uw), C.byref(nu))
    if rc != 0:
        raise RuntimeError(f"ref_iapi_estimate failed rc={rc}")
    return R._take(lw, nl), R._take(uw, nu)


def ref_bin_bounds(name):
    """The boundaries of the bin index the reference built for the
    registered array `name` (bin::binBoundaries)."""
    L = reference().lib
    p, n = _dp(), C.c_uint64()
    rc = L.ref_iapi_bin_bounds(name.encode() if isinstance(name, str) else name, C.byref(p), C.byref(n))
    if rc != 0:
        raise RuntimeError(f"ref_iapi_bin_bounds failed rc={rc}")
    out = np.ctypeslib.as_array(p, (n.value,)).copy() if n.value else np.zeros(0)
    L.ref_free(p)
    return out


def bounds_from_keys(keys):
    """bin bounds as bin(..., keys, offsets) rebuilds them
    (src/ibin.cpp:358-372): compactValue(maxval[j], minval[j+1]), DBL_MAX."""
    nb = len(keys) // 2
    mn, mx = keys[:nb], keys[nb:]
    out = [compact_value(mx[j], mn[j + 1]) for j in range(nb - 1)]
    out.append(np.finfo(np.float64).max)
    return np.array(out, np.float64)


def compact_value(left, right, start=0.0):
    """util::compactValue (src/util.cpp:1227-1312)."""
    import math
    FLT_EPS = 1.1920928955078125e-07
    if left == right:
        return left
    if not (left != right):
        return 0.0
    if left > right:
        left, right = right, left
    if left < 0 and right >= 0:
        return 0.0
    if left < 1.0 and right >= 1.0:
        return 1.0
    if left < -1.0 and right >= -1.0:
        return -1.0
    if left == 0.0:
        diff = math.floor(math.log10(right))
        sep = 10.0 ** diff
        if sep > right:
            sep *= 0.5 if (-3.0 <= diff < 3.0) else 0.1
    elif right < 0.0 and right * 10 > left:
        diff = math.ceil(math.log10(-right))
        sep = -(10.0 ** diff)
        if sep > right:
            sep = sep + sep if (-3.0 <= diff <= 3.0) else sep * 10
    elif left > 0.0 and right > 10 * left:
        diff = math.ceil(math.log10(left))
        sep = 10.0 ** diff
        if sep <= left:
            sep = sep + sep if (sep >= -3.0 and diff <= 3.0) else sep * 10.0
    else:
        diff = 10.0 ** math.ceil(FLT_EPS + math.log10(right - left))
        if not (start > left and start <= right):
            start = 0.5 * (right + left)
        sep = math.floor(0.5 + start / diff) * diff
        for div in (2, 5, 2, 2):
            if sep > left and sep <= right:
                break
            diff /= div
            sep = math.floor(0.5 + start / diff) * diff
    if not (sep > left and sep <= right):
        sep = right
    return sep


class RefAttached:
    """A column registered in the reference's in-memory API with an index
    attached from keys/offsets/bitmaps (no rebuild)."""

    def __init__(self, name, col, keys, offsets, bitmaps):
        self.name = name.encode()
        self.col = np.ascontiguousarray(col)
        self.keys = np.ascontiguousarray(keys, np.float64)
        self.offsets = np.ascontiguousarray(offsets, np.int64)
        self.bitmaps = np.ascontiguousarray(bitmaps, np.uint32)
        _REGISTERED.append((self.col, self.keys, self.offsets, self.bitmaps))
        L = reference().lib
        L.ref_iapi_query.restype = C.c_int64
        rc = L.ref_iapi_attach(self.name, _FB_TYPE[col_type_of(self.col)],
                               self.col.ctypes.data_as(C.c_void_p), C.c_uint64(self.col.size),
                               self.keys.ctypes.data_as(_dp), C.c_uint64(self.keys.size),
                               self.offsets.ctypes.data_as(_i64p), C.c_uint64(self.offsets.size),
                               self.bitmaps.ctypes.data_as(_u32p), C.c_uint64(self.bitmaps.size))
        if rc < 0:
            raise RuntimeError(f"ref_iapi_attach failed rc={rc}")

    def estimate(self, lo, lop, hi, rop):
        return ref_iapi_estimate(self.name, lo, lop, hi, rop)

    def count_range(self, lo, lop, hi, rop):
        """Hits of lo OP1 x OP2 hi as one range term (one qContinuousRange
        through fastbit_selection_evaluate)."""
        L = reference().lib
        L.ref_iapi_query_range.restype = C.c_int64
        n = L.ref_iapi_query_range(self.name, C.c_double(lo), C.c_int(lop), C.c_double(hi), C.c_int(rop), None,
                                   C.c_uint64(0))
        if n < 0:
            raise RuntimeError(f"ref_iapi_query_range failed {n}")
        return int(n)

    def count(self, lo, lop, hi, rop):
        n = reference().lib.ref_iapi_query(self.name, C.c_double(lo), C.c_int(lop), C.c_double(hi),
                                           C.c_int(rop), None, C.c_uint64(0))
        if n < 0:
            raise RuntimeError(f"ref_iapi_query failed {n}")
        return int(n)


# ---------------------------------------------------------------------------
# Index files written and read by the reference (oracle/refdrv.cpp).
def ref_write_index_file(name, path):
    """relic::write / bin::write of the index the reference built for the
    registered array `name` (see RefIapi)."""
    rc = reference().lib.ref_iapi_write_index_file(name.encode(), os.fsencode(path))
    if rc < 0:
        raise RuntimeError(f"ref_iapi_write_index_file failed rc={rc}")


def ref_relic_card_file(ind, card, path):
    """relic(0, card, ind) written to an index file."""
    a = np.ascontiguousarray(ind, np.uint32)
    rc = reference().lib.ref_relic_card_file(a.ctypes.data_as(_u32p), C.c_size_t(a.size), C.c_uint32(card),
                                             os.fsencode(path))
    if rc < 0:
        raise RuntimeError(f"ref_relic_card_file failed rc={rc}")


def ref_file_triple(path):
    """(kind, keys, offsets, bitmaps) of an index file as the reference reads
    it; kind is "relic" or "bin"."""
    L = reference().lib
    kind = C.c_int()
    k, o, b = _dp(), _i64p(), _u32p()
    nk, no, nb = C.c_uint64(), C.c_uint64(), C.c_uint64()
    rc = L.ref_file_triple(os.fsencode(path), C.byref(kind), C.byref(k), C.byref(nk), C.byref(o), C.byref(no),
                           C.byref(b), C.byref(nb))
    if rc < 0:
        raise RuntimeError(f"ref_file_triple failed rc={rc}")
    out = ("relic" if kind.value == 1 else "bin",
           np.ctypeslib.as_array(k, (nk.value,)).copy() if nk.value else np.zeros(0),
           np.ctypeslib.as_array(o, (no.value,)).copy(),
           np.ctypeslib.as_array(b, (nb.value,)).copy() if nb.value else np.zeros(0, np.uint32))
    for p in (k, o, b):
        L.ref_free(p)
    return out


def ref_file_locate(path, lo, lop, hi, rop):
    """(cand0, hit0, hit1, cand1) of lo OP1 x OP2 hi on an index file."""
    out = np.zeros(4, np.uint32)
    rc = reference().lib.ref_file_locate(os.fsencode(path), C.c_double(lo), C.c_int(lop), C.c_double(hi),
                                         C.c_int(rop), out.ctypes.data_as(_u32p))
    if rc < 0:
        raise RuntimeError(f"ref_file_locate failed rc={rc}")
    return tuple(int(x) for x in out)


def ref_file_sum_bins(path, ib, ie):
    """Ones of index::sumBins(ib, ie) on an index file."""
    L = reference().lib
    L.ref_file_sum_bins.restype = C.c_int64
    n = L.ref_file_sum_bins(os.fsencode(path), C.c_uint32(ib), C.c_uint32(ie))
    if n < 0:
        raise RuntimeError(f"ref_file_sum_bins failed {n}")
    return int(n)


def ref_iapi_expr(terms, prog):
    """Hit count of a postfix boolean program over range terms on arrays
    registered with RefIapi: terms is a list of (name, lo, lop, hi, rop),
    prog items >= 0 push a term, -1..-7 are AND OR XOR NAND NOR NOT MINUS;
    built with fastbit_selection_combine, counted by
    fastbit_selection_evaluate."""
    L = reference().lib
    L.ref_iapi_expr.restype = C.c_int64
    n = len(terms)
    names = (C.c_char_p * n)(*[t[0].encode() for t in terms])
    lo = np.array([t[1] for t in terms], np.float64)
    lop = np.array([t[2] for t in terms], np.int32)
    hi = np.array([t[3] for t in terms], np.float64)
    rop = np.array([t[4] for t in terms], np.int32)
    pr = np.ascontiguousarray(prog, np.int32)
    cnt = L.ref_iapi_expr(names, lo.ctypes.data_as(_dp), lop.ctypes.data_as(C.POINTER(C.c_int)),
                          hi.ctypes.data_as(_dp), rop.ctypes.data_as(C.POINTER(C.c_int)),
                          pr.ctypes.data_as(C.POINTER(C.c_int32)), C.c_uint32(pr.size))
    if cnt < 0:
        raise RuntimeError(f"ref_iapi_expr failed {cnt}")
    return int(cnt)


def expr_rows(masks, prog):
    """The rows a postfix program selects, from the boolean row masks of its
    terms (NOT complements over every row)."""
    st = []
    for op in prog:
        if op >= 0:
            st.append(masks[op])
        elif op == -6:
            st.append(~st.pop())
        else:
            b, a = st.pop(), st.pop()
            st.append({-1: lambda: a & b, -2: lambda: a | b, -3: lambda: a ^ b, -4: lambda: ~(a & b),
                       -5: lambda: ~(a | b), -7: lambda: a & ~b}[op]())
    assert len(st) == 1
    return st[0]


# ---------------------------------------------------------------------------
# Two-level encodings over a relic (ibis::fuzz, ibis::zona;
# src/irelic.h:540-712) restated in oracle/index.cpp.
ENC_FUZZ, ENC_ZONA = 1, 2
IDX_FUZZ, IDX_ZONA = 24, 25          # ibis::index::INDEX_TYPE (src/index.h:139-143)
TL_EMPTY, TL_FINE, TL_ALL, TL_EQUAL, TL_SAME_BIN = 0, 1, 6, 7, 8


def two_level_file_parts(path_or_bytes):
    """The parts of a fuzz / zona index file (src/ixfuzz.cpp:46-143,
    src/ixzona.cpp:40-100): (encoding, fine positions, nc, cbounds, coarse
    positions, coarse words, word offsets of the coarse bitmaps)."""
    import struct
    b = path_or_bytes if isinstance(path_or_bytes, (bytes, bytearray)) else open(path_or_bytes, "rb").read()
    if b[5] not in (IDX_FUZZ, IDX_ZONA):
        raise ValueError("not a fuzz / zona index file")
    w = b[6]
    it = np.int32 if w == 4 else np.int64
    nobs = struct.unpack_from("<I", b, 12)[0]
    fine = np.frombuffer(b, it, nobs + 1, 24 + 8 * nobs).astype(np.int64)
    p = int(fine[-1])
    nc = struct.unpack_from("<I", b, p)[0]
    cb = np.frombuffer(b, np.uint32, nc + 1, p + 4).copy()
    ncb = nc - (nc + 1) // 2 + 1 if b[5] == IDX_FUZZ else nc
    co = np.frombuffer(b, it, ncb + 1, p + 4 + 4 * (nc + 1)).astype(np.int64)
    words = np.frombuffer(b[int(co[0]):int(co[-1])], np.uint32).copy()
    enc = ENC_FUZZ if b[5] == IDX_FUZZ else ENC_ZONA
    return enc, fine, nc, cb, co, words, (co - co[0]) // 4


def two_level_widen(src, dst):
    """The same fuzz / zona index file with 8-byte bitmap positions: the relic
    part as relic::write64 lays it out (src/irelic.cpp:456-642), then nc,
    cbounds, 8-byte coarse positions and the coarse bitmaps
    (src/ixfuzz.cpp:1272-1331)."""
    import struct
    b = open(src, "rb").read()
    assert b[6] == 4
    nobs = struct.unpack_from("<I", b, 12)[0]
    obeg = 24 + 8 * nobs
    pos = np.frombuffer(b, np.int32, nobs + 1, obeg).astype(np.int64)
    p = int(pos[-1])
    nc = struct.unpack_from("<I", b, p)[0]
    ncb = nc - (nc + 1) // 2 + 1 if b[5] == IDX_FUZZ else nc
    cpos_at = p + 4 + 4 * (nc + 1)
    co = np.frombuffer(b, np.int32, ncb + 1, cpos_at).astype(np.int64)
    s1 = 4 * (nobs + 1)                 # growth of the fine positions
    s2 = s1 + 4 * (ncb + 1)             # growth up to the coarse bitmaps
    head = bytearray(b[:obeg])
    head[6] = 8
    out = (bytes(head) + (pos + s1).astype("<i8").tobytes() + b[obeg + s1:p]
           + b[p:cpos_at] + (co + s2).astype("<i8").tobytes() + b[int(co[0]):])
    with open(dst, "wb") as f:
        f.write(out)


class OracleTwoLevel:
    """The coarse level of ibis::fuzz / ibis::zona over an OracleIndex relic."""

    def __init__(self, fine, handle):
        self.fine = fine
        self._h = handle
        if not handle:
            raise RuntimeError("two-level oracle failed")

    @classmethod
    def build(cls, fine, enc, ncoarse=0):
        """coarsen() on the index built in memory."""
        L = oracle().lib
        L.orc_two_level_build.restype = C.c_void_p
        return cls(fine, L.orc_two_level_build(C.c_void_p(fine._h), C.c_int(enc), C.c_uint32(ncoarse)))

    @classmethod
    def from_file(cls, fine, path):
        """The coarse level as the index file at path holds it; `fine` is the
        relic over the same column with file offsets (offsets_built)."""
        L = oracle().lib
        L.orc_two_level_file.restype = C.c_void_p
        enc, fp, nc, cb, co, words, woff = two_level_file_parts(path)
        fp = np.ascontiguousarray(fp, np.int64)
        cb = np.ascontiguousarray(cb, np.uint32)
        co = np.ascontiguousarray(co, np.int64)
        words = np.ascontiguousarray(words, np.uint32)
        woff = np.ascontiguousarray(woff, np.int64)
        h = L.orc_two_level_file(C.c_void_p(fine._h), C.c_int(enc), fp.ctypes.data_as(_i64p),
                                 C.c_uint32(fp.size - 1), cb.ctypes.data_as(_u32p), C.c_uint32(nc),
                                 co.ctypes.data_as(_i64p), C.c_uint32(co.size - 1),
                                 words.ctypes.data_as(_u32p), woff.ctypes.data_as(_i64p))
        return cls(fine, h)

    def __del__(self):
        try:
            oracle().lib.orc_two_level_free(C.c_void_p(self._h))
        except Exception:
            pass

    def info(self):
        """(cbounds, coarse offsets, fine offsets) in bytes."""
        L = oracle().lib
        a, b, c = _u32p(), _i64p(), _i64p()
        na, nb, nc = C.c_size_t(), C.c_size_t(), C.c_size_t()
        L.orc_two_level_info(C.c_void_p(self._h), C.byref(a), C.byref(na), C.byref(b), C.byref(nb), C.byref(c),
                             C.byref(nc))
        out = (np.ctypeslib.as_array(a, (na.value,)).copy() if na.value else np.zeros(0, np.uint32),
               np.ctypeslib.as_array(b, (nb.value,)).copy() if nb.value else np.zeros(0, np.int64),
               np.ctypeslib.as_array(c, (nc.value,)).copy() if nc.value else np.zeros(0, np.int64))
        for q in (a, b, c):
            L.orc_free(q)
        return out

    def coarse_bitmap(self, i):
        """Serialized words of coarse bitmap i."""
        L = oracle().lib
        p, n = _u32p(), C.c_size_t()
        if L.orc_two_level_cbit(C.c_void_p(self._h), C.c_uint32(i), C.byref(p), C.byref(n)) != 0:
            raise IndexError(i)
        out = np.ctypeslib.as_array(p, (n.value,)).copy()
        L.orc_free(p)
        return out

    def cost(self, lo, lop, hi, rop):
        L = oracle().lib
        L.orc_two_level_cost.restype = C.c_double
        return L.orc_two_level_cost(C.c_void_p(self.fine._h), C.c_void_p(self._h), self.fine.type, C.c_double(lo),
                                    C.c_int(lop), C.c_double(hi), C.c_int(rop))

    def evaluate(self, lo, lop, hi, rop):
        """(hit count, ascending rids, option, coarse bitmaps read as a mask)."""
        L = oracle().lib
        cnt, opt, ld = C.c_uint64(), C.c_int(), C.c_uint64()
        r, nr = _u32p(), C.c_size_t()
        rc = L.orc_two_level_eval(C.c_void_p(self.fine._h), C.c_void_p(self._h), C.c_double(lo), C.c_int(lop),
                                  C.c_double(hi), C.c_int(rop), C.byref(cnt), C.byref(r), C.byref(nr),
                                  C.byref(opt), C.byref(ld))
        if rc != 0:
            raise RuntimeError("orc_two_level_eval failed")
        rids = np.ctypeslib.as_array(r, (nr.value,)).copy() if nr.value else np.zeros(0, np.uint32)
        L.orc_free(r)
        return int(cnt.value), rids, int(opt.value), int(ld.value)
