"""Equal-weight bin boundaries (bin::scanAndPartition + index::divideCounts
+ util::compactValue) restated in the library's host code, checked against
the boundaries the reference library computes for the same column.  CPU
only: fbh_bin_equal_weight touches no device."""
import ctypes as C
import itertools
import os

import numpy as np
import pytest

import oracle as O

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
_names = itertools.count()


@pytest.fixture(scope="module")
def lib():
    path = os.path.join(ROOT, "libfastlib_amd", "lib", "libfastbit_hip.so")
    if not os.path.exists(path):
        pytest.skip("libfastbit_hip.so not built")
    L = C.CDLL(path)
    dp = C.POINTER(C.c_double)
    up = C.POINTER(C.c_uint32)
    L.fbh_bin_equal_weight.argtypes = [dp, up, C.c_uint32, C.c_uint32, C.c_int, C.c_double, dp,
                                       C.c_uint32, up]
    return L


def ours(lib, col, nbins):
    v, c = np.unique(col.astype(np.float64), return_counts=True)
    v = np.ascontiguousarray(v)
    c = np.ascontiguousarray(c, np.uint32)
    n = C.c_uint32()
    is_float = 1 if col.dtype.kind == "f" else 0
    out = np.zeros(len(v) * 2 + 8)
    rc = lib.fbh_bin_equal_weight(v.ctypes.data_as(C.POINTER(C.c_double)),
                                  c.ctypes.data_as(C.POINTER(C.c_uint32)), len(v), nbins, is_float,
                                  -np.finfo(np.float64).max, out.ctypes.data_as(C.POINTER(C.c_double)),
                                  out.size, C.byref(n))
    assert rc == 0
    return out[: n.value]


COLS = {
    "normal": lambda rng, n: rng.normal(size=n).astype(np.float32),
    "normal_round": lambda rng, n: (np.round(rng.normal(size=n) * 100) / 100).astype(np.float32),
    "heavy_float": lambda rng, n: np.where(rng.random(n) < 0.3, np.float32(0.25),
                                           rng.normal(size=n)).astype(np.float32),
    "zipf_int": lambda rng, n: np.minimum(rng.zipf(1.3, size=n), 100000).astype(np.int32),
    "uniform_int": lambda rng, n: rng.integers(-5000, 5000, size=n).astype(np.int32),
    "few_int": lambda rng, n: rng.integers(0, 40, size=n).astype(np.int32),
    "u16": lambda rng, n: rng.integers(0, 65536, size=n).astype(np.uint16),
    "sorted_runs": lambda rng, n: np.repeat(np.arange(n // 37 + 1), 37)[:n].astype(np.int32),
}


@pytest.mark.parametrize("kind", sorted(COLS))
@pytest.mark.parametrize("n,nbins", [(5000, 10), (31 * 997 + 3, 100), (200003, 1000)])
def test_equal_weight_bounds_match_reference(ref, lib, kind, n, nbins):
    rng = np.random.default_rng(n + nbins + len(kind))
    col = COLS[kind](rng, n)
    name = f"bb{next(_names)}"
    O.RefIapi(name, col, f"<binning nbins={nbins} equal-weight/>")
    want = O.ref_bin_bounds(name)
    got = ours(lib, col, nbins)
    assert len(got) == len(want), (len(got), len(want))
    assert np.array_equal(got, want)


@pytest.mark.slow
def test_config3_bounds_match_reference(lib):
    """The 2^27-row float column, nbins=10000: the reference's boundaries
    (tests/golden/config3_2p27_bounds.npz) from the same histogram."""
    from libfastlib_amd import datagen as G
    path = os.path.join(ROOT, "tests", "golden", "config3_2p27_bounds.npz")
    z = np.load(path)
    if "ref_bounds" not in z:
        pytest.skip("fixture without the reference's boundaries")
    col = G.float_mix(3, 1 << 27, pool=100000, wnorm=0.7, s=1.2)
    got = ours(lib, col, 10000)
    assert np.array_equal(got, z["ref_bounds"])
