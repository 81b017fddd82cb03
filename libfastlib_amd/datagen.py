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
