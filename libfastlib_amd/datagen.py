"""Deterministic synthetic workloads (ctypes binding of lib/libfbgen.so).

The generators use their own seeded xoshiro256** stream and IEEE basic
arithmetic only, so the same seed gives the same data on every host.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
GEN_PATH = os.path.join(HERE, "lib", "libfbgen.so")
_g = None
_u32p = C.POINTER(C.c_uint32)


def _lib():
    global _g
    if _g is None:
        if not os.path.exists(GEN_PATH):
            raise RuntimeError(f"{GEN_PATH} is not built")
        g = C.CDLL(GEN_PATH)
        g.fbg_free.argtypes = [C.c_void_p]
        g.fbg_markov_wah.argtypes = [C.c_uint64, C.c_uint32, C.c_double, C.c_double,
                                     C.POINTER(_u32p), C.POINTER(C.c_size_t)]
        g.fbg_uniform_i32.argtypes = [C.c_uint64, C.c_uint32, C.c_int32, C.c_int32, C.c_void_p]
        g.fbg_normal_f32.argtypes = [C.c_uint64, C.c_uint32, C.c_void_p]
        g.fbg_zipf_i32.argtypes = [C.c_uint64, C.c_uint32, C.c_double, C.c_uint32, C.c_void_p]
        g.fbg_float_mix.argtypes = [C.c_uint64, C.c_uint32, C.c_uint32, C.c_double, C.c_double,
                                    C.c_void_p, C.c_void_p]
        g.fbg_sorted_runs_i32.argtypes = [C.c_uint64, C.c_uint32, C.c_double, C.c_void_p]
        g.fbg_uniform_f64.argtypes = [C.c_uint64, C.c_uint32, C.c_void_p]
        g.fbg_uniform_incl_i64.argtypes = [C.c_uint64, C.c_uint32, C.c_int64, C.c_int64,
                                           C.c_void_p]
        _g = g
    return _g


def markov_wah(seed: int, nbits: int, density: float, mean_run: float) -> np.ndarray:
    """Canonical serialized WAH of a two-state Markov bit string."""
    out = _u32p()
    n = C.c_size_t()
    rc = _lib().fbg_markov_wah(seed, nbits, density, mean_run, C.byref(out), C.byref(n))
    if rc != 0:
        raise RuntimeError("fbg_markov_wah failed")
    arr = np.ctypeslib.as_array(out, shape=(n.value,)).copy()
    _lib().fbg_free(out)
    return arr.astype(np.uint32)


def uniform_i32(seed, n, lo, hi):
    out = np.empty(n, np.int32)
    _lib().fbg_uniform_i32(seed, n, lo, hi, out.ctypes.data)
    return out


def normal_f32(seed, n):
    out = np.empty(n, np.float32)
    _lib().fbg_normal_f32(seed, n, out.ctypes.data)
    return out


def zipf_i32(seed, n, s, K):
    out = np.empty(n, np.int32)
    _lib().fbg_zipf_i32(seed, n, s, K, out.ctypes.data)
    return out


def float_mix(seed, n, pool=100000, wnorm=0.7, s=1.2):
    out = np.empty(n, np.float32)
    _lib().fbg_float_mix(seed, n, pool, wnorm, s, out.ctypes.data, None)
    return out


def sorted_runs_i32(seed, n, p=1.0 / 512):
    out = np.empty(n, np.int32)
    _lib().fbg_sorted_runs_i32(seed, n, p, out.ctypes.data)
    return out


def uniform_f64(seed, n):
    out = np.empty(n, np.float64)
    _lib().fbg_uniform_f64(seed, n, out.ctypes.data)
    return out


def uniform_incl_i64(seed, n, lo, hi):
    out = np.empty(n, np.int64)
    _lib().fbg_uniform_incl_i64(seed, n, lo, hi, out.ctypes.data)
    return out


def table4(seed, n):
    """The 4-column table of the conjunctive workload: uint16 uniform in
    [0, 65536), int32 Zipf(1.1) over 10^6 values, float32 N(0,1), int32
    sorted runs of geometric(1/512) length."""
    return [uniform_i32(seed, n, 0, 65536).astype(np.uint16),
            zipf_i32(seed + 1, n, 1.1, 1 << 20),
            normal_f32(seed + 2, n),
            sorted_runs_i32(seed + 3, n, 1.0 / 512)]


def conj_queries(seed, cols, nq, min_terms=2, max_terms=4, smin=1e-6, smax=0.3):
    """nq conjunctions of 2..4 range terms lo <= x < hi on distinct columns,
    each term's selectivity drawn log-uniform in [smin, smax]; endpoints are
    the column's empirical quantiles.  Returns [[(col_index, lo, hi), ...]]."""
    srt = [np.sort(c) for c in cols]
    n = len(cols[0])
    u = uniform_f64(seed, nq * 16)
    k = 0
    out = []
    for _ in range(nq):
        nt = min_terms + int(u[k] * (max_terms - min_terms + 1)); k += 1
        nt = min(nt, max_terms)
        order = np.argsort(u[k:k + len(cols)]); k += len(cols)
        q = []
        for ci in order[:nt]:
            sel = np.exp(np.log(smin) + u[k] * (np.log(smax) - np.log(smin))); k += 1
            start = u[k] * (1.0 - sel); k += 1
            a = srt[ci][min(int(start * n), n - 1)]
            b = srt[ci][min(int((start + sel) * n), n - 1)]
            if b <= a:
                b = a + (1 if srt[ci].dtype.kind in "iu" else np.spacing(np.float32(abs(a) + 1)))
            q.append((int(ci), float(a), float(b)))
        out.append(q)
    return out


def quantile_ranges(srt, seed, k):
    """k ranges lo <= x < hi whose endpoints are empirical quantiles of a
    column at uniform draws: srt is the sorted column; lo and hi are the
    values at ranks floor(u * n) of 2k uniform draws, ordered per pair."""
    n = len(srt)
    u = uniform_f64(seed, 2 * k)
    a = srt[np.minimum((u[:k] * n).astype(np.int64), n - 1)].astype(np.float64)
    b = srt[np.minimum((u[k:] * n).astype(np.int64), n - 1)].astype(np.float64)
    return np.minimum(a, b), np.maximum(a, b)


# The binned float workload: a float32 column of 2^27 rows from a pool of
# 10^5 distinct values (70% N(0,1) draws, 30% Zipf(1.2) support), binned
# with 10,000 equal-weight bins; 1024 quantile ranges for the range-query
# batch, and 50,000 training / 8192 test ranges for the learned plan choice.
BIN_F32_SEED = 3
BIN_F32_QUERY_SEED = 1003
BIN_F32_TRAIN_SEED = 5005
BIN_F32_TEST_SEED = 5006


def bin_f32_column(log2n=27):
    return float_mix(BIN_F32_SEED, 1 << log2n, pool=100000, wnorm=0.7, s=1.2)


# The conjunctive workload: table4(CONJ_TABLE_SEED, 2^26) and
# conj_queries(CONJ_QUERY_SEED, ..., 4096).
CONJ_TABLE_SEED = 41
CONJ_QUERY_SEED = 42
