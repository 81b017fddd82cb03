"""Two-level indexes over a relic (ibis::fuzz, ibis::zona; src/irelic.h:
540-712): the oracle's restatement (oracle/index.cpp) against index files
the reference wrote (tests/golden/idx/{fuzz,zona}_*, made by
tests/golden/make_two_level.py) and against the column itself.

What pins what:
  * cbounds          the file's coarse bins (coarsen's byte split and, for
                     fuzz, its default count or "ncoarse=8")
  * coarse bitmaps   their rows, from the column: fuzz bitmap i holds the
                     coarse bins [i, i + ceil(nc/2)), zona bitmap i bin i;
                     their words, after canonical compression, the file's
  * costs            estimateCost of the reference on the index it built in
                     memory and on the file it wrote (96 ranges each)
  * options          the coarse bitmaps evaluate reads: the reference index
                     read from the file loads each bitmap on first use, so
                     the mask names the option it took (1..5, same bin)
  * answers          counts and row ids against the column
"""
from __future__ import annotations

import os

import numpy as np
import pytest

import oracle

IDX = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "idx")
NAMES = ["fuzz_i32", "fuzz_i32_nc8", "zona_i32", "zona_f32"]


def fixture(name):
    with np.load(os.path.join(IDX, name + ".npz")) as z:
        return {k: z[k] for k in z.files}


def enc_of(name):
    return oracle.ENC_FUZZ if name.startswith("fuzz") else oracle.ENC_ZONA


def spec_nc(name):
    return 8 if name.endswith("nc8") else 0


def brute(col, lo, lop, hi, rop):
    x = col.astype(np.float64)
    m = np.ones(col.size, bool)
    if lop == 1:
        m &= lo < x
    elif lop == 2:
        m &= lo <= x
    if rop == 1:
        m &= x < hi
    elif rop == 2:
        m &= x <= hi
    return m


def terms(f):
    return [(float(f["lo"][i]), int(f["lop"][i]), float(f["hi"][i]), int(f["rop"][i])) for i in range(len(f["lo"]))]


def canonical(words):
    o = oracle.oracle()
    return o.compress(o.decompress(words))


@pytest.fixture(scope="module", params=NAMES)
def built(request):
    name = request.param
    f = fixture(name)
    fine = oracle.OracleIndex.relic_values(f["col"])
    return name, f, fine, oracle.OracleTwoLevel.build(fine, enc_of(name), spec_nc(name))


@pytest.fixture(scope="module", params=NAMES)
def from_file(request):
    name = request.param
    f = fixture(name)
    fine = oracle.OracleIndex.relic_values(f["col"]).as_file()
    return name, f, fine, oracle.OracleTwoLevel.from_file(fine, os.path.join(IDX, name + ".idx"))


def test_file_parts_roundtrip_8byte(tmp_path):
    """The 8-byte layout holds the same parts (positions shifted)."""
    for name in NAMES:
        src = os.path.join(IDX, name + ".idx")
        dst = str(tmp_path / (name + "_w8.idx"))
        oracle.two_level_widen(src, dst)
        a = oracle.two_level_file_parts(src)
        b = oracle.two_level_file_parts(dst)
        assert a[0] == b[0] and a[2] == b[2]
        assert np.array_equal(a[3], b[3]) and np.array_equal(a[5], b[5]) and np.array_equal(a[6], b[6])
        assert np.array_equal(np.diff(a[1]), np.diff(b[1])) and np.array_equal(np.diff(a[4]), np.diff(b[4]))


def test_coarse_bins_match_file(built):
    name, f, fine, tl = built
    cb, coff, off = tl.info()
    assert np.array_equal(cb, f["cbounds"])
    assert cb.size - 1 == int(f["nc"])
    # coffset64 of the built index sums bitvector::bytes() (4 per word + 40);
    # the file's positions step by the serialized size (4 per word + 8)
    assert np.array_equal(np.diff(coff), np.diff(f["coffsets"]) + 32)


def test_coarse_bitmaps_rows_and_words(built):
    name, f, fine, tl = built
    col = f["col"]
    cb = f["cbounds"]
    vals = np.unique(col)                      # relic: fine bitmap k holds value vals[k]
    pos = np.searchsorted(vals, col)
    nc = cb.size - 1
    nc2 = (nc + 1) // 2
    nb = nc - nc2 + 1 if name.startswith("fuzz") else nc
    woff = f["cbits_off"]
    o = oracle.oracle()
    for i in range(nb):
        a, b = (cb[i], cb[i + nc2]) if name.startswith("fuzz") else (cb[i], cb[i + 1])
        want = ((pos >= a) & (pos < b)).astype(np.uint8)
        mine = tl.coarse_bitmap(i)
        assert np.array_equal(o.rids(mine), np.nonzero(want)[0].astype(np.uint32)), (name, i)
        ref = f["cbits"][woff[i]:woff[i + 1]]
        assert np.array_equal(canonical(ref), canonical(mine)), (name, i)


def test_costs_match_reference_built_in_memory(built):
    name, f, fine, tl = built
    got = np.array([tl.cost(*t) for t in terms(f)])
    assert np.array_equal(got, f["cost_mem"])


def test_costs_options_and_counts_match_reference_file(from_file):
    name, f, fine, tl = from_file
    opts = set()
    for i, t in enumerate(terms(f)):
        assert tl.cost(*t) == f["cost_file"][i], (name, i, t)
        cnt, rids, opt, loads = tl.evaluate(*t)
        assert cnt == f["counts"][i], (name, i, t)
        assert loads == int(f["loaded"][i]), (name, i, t, opt, bin(loads), bin(int(f["loaded"][i])))
        opts.add(opt)
    # the fixtures reach the fine level and every coarse option
    assert {1, 2, 3, 4, 5} <= opts, opts


def test_answers_are_the_columns(built):
    name, f, fine, tl = built
    col = f["col"]
    for i, t in enumerate(terms(f)):
        cnt, rids, opt, loads = tl.evaluate(*t)
        m = brute(col, *t)
        assert cnt == int(m.sum()) == f["counts"][i]
        if i % 8 == 0:
            assert np.array_equal(rids, np.nonzero(m)[0].astype(np.uint32))


def test_no_coarse_level_below_32_values():
    col = (np.arange(5000) % 31).astype(np.int32)
    fine = oracle.OracleIndex.relic_values(col)
    for enc in (oracle.ENC_FUZZ, oracle.ENC_ZONA):
        tl = oracle.OracleTwoLevel.build(fine, enc)
        cb, coff, off = tl.info()
        assert cb.size == 0 and coff.size == 0
        cnt, _, opt, loads = tl.evaluate(3.0, 2, 20.0, 1)
        assert cnt == int(((col >= 3) & (col < 20)).sum()) and loads == 0 and opt == oracle.TL_FINE


def test_planted_mistake_in_coarse_bounds_is_caught():
    """A coarse level shifted by one fine bitmap answers some ranges wrong:
    the answer checks above would see it."""
    f = fixture("zona_i32")
    col = f["col"]
    path = os.path.join(IDX, "zona_i32.idx")
    enc, fp, nc, cb, co, words, woff = oracle.two_level_file_parts(path)
    bad = cb.copy()
    bad[3] += 1
    import ctypes as C
    L = oracle.oracle().lib
    L.orc_two_level_file.restype = C.c_void_p
    fine = oracle.OracleIndex.relic_values(col).as_file()
    args = [np.ascontiguousarray(x) for x in (fp.astype(np.int64), bad.astype(np.uint32), co.astype(np.int64),
                                               words.astype(np.uint32), woff.astype(np.int64))]
    h = L.orc_two_level_file(C.c_void_p(fine._h), C.c_int(enc), args[0].ctypes.data_as(C.POINTER(C.c_int64)),
                             C.c_uint32(fp.size - 1), args[1].ctypes.data_as(C.POINTER(C.c_uint32)), C.c_uint32(nc),
                             args[2].ctypes.data_as(C.POINTER(C.c_int64)), C.c_uint32(co.size - 1),
                             args[3].ctypes.data_as(C.POINTER(C.c_uint32)), args[4].ctypes.data_as(C.POINTER(C.c_int64)))
    tl = oracle.OracleTwoLevel(fine, h)
    wrong = sum(tl.evaluate(*t)[0] != f["counts"][i] for i, t in enumerate(terms(f)))
    assert wrong > 0
