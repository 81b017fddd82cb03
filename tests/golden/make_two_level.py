"""Writes the two-level index fixtures (ibis::fuzz, ibis::zona) with the
reference FastBit library.

Each fixture is an index file the reference wrote (fuzz::write /
zona::write after building the index for a registered array) next to an
.npz holding:

  col                the column
  lo, lop, hi, rop   96 range terms
  counts             the reference's hit counts (evaluate on the index it
                     built in memory; the file read back gives the same)
  cost_mem           estimateCost on the index built in memory (before the
                     write, which replaces its bitmap offsets by file
                     positions)
  cost_file          estimateCost on the index read from the file
  loaded             bit i set when evaluate on the index read from the
                     file loaded coarse bitmap i (the file's bitmaps load on
                     first use, so this names the option evaluate took)
  nc, cbounds        the coarse bins as the file holds them
  coffsets           coarse bitmap positions (bytes)
  cbits, cbits_off   the coarse bitmaps' serialized words (cbits_off in
                     words from the first coarse bitmap)

  fuzz_i32        fuzz over 20,000 int32 values in [0, 500): a sorted half,
                  a uniform half; the default number of coarse bins
  fuzz_i32_nc8    the same column with "ncoarse=8"
  zona_i32        zona over the same column (11 coarse bins)
  zona_f32        zona over 20,000 float32 values, normal(0, 40) rounded to
                  halves

The 8-byte-position form of each file is written here too and read back by
the reference to the same counts, costs and loads; the tests make it from
the 4-byte file with the same rewrite (oracle.two_level_widen).

    python tests/golden/make_two_level.py
"""
from __future__ import annotations

import ctypes as C
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

import oracle  # noqa: E402

OUT = os.path.join(HERE, "idx")
FUZZ, ZONA = oracle.IDX_FUZZ, oracle.IDX_ZONA


def queries(rng, vmin, vmax, n=96):
    span = vmax - vmin
    lo = rng.uniform(vmin - 0.05 * span, vmax, n)
    width = span * rng.choice([0.01, 0.05, 0.2, 0.5, 0.9], n) * rng.uniform(0.5, 1.5, n)
    hi = lo + width
    lop = rng.integers(0, 3, n)
    rop = rng.integers(0, 3, n)
    rop[(lop == 0) & (rop == 0)] = 1
    k = rng.random(n) < 0.5
    lo[k] = np.round(lo[k])
    hi[k] = np.round(hi[k])
    return lo, lop.astype(np.int32), hi, rop.astype(np.int32)


def evaluate(L, name, path, q):
    cost, ld = C.c_double(), C.c_uint64()
    cnt, costs, loads = [], [], []
    for a, b, c, d in zip(*q):
        n = L.ref_two_level_eval(name, os.fsencode(path) if path else None, C.c_double(a), C.c_int(int(b)),
                                 C.c_double(c), C.c_int(int(d)), C.byref(cost), C.byref(ld))
        assert n >= 0, n
        cnt.append(n)
        costs.append(cost.value)
        loads.append(ld.value)
    return np.array(cnt, np.int64), np.array(costs), np.array(loads, np.uint64)


def make(L, rng, fname, col, spec):
    R = oracle.RefIapi("tl_" + fname, col, spec)
    q = queries(rng, float(col.min()), float(col.max()))
    counts, cost_mem, _ = evaluate(L, R.name, None, q)
    path = os.path.join(OUT, fname + ".idx")
    rc = L.ref_iapi_write_coarse_file(R.name, os.fsencode(path))
    assert rc == 0, rc
    b = open(path, "rb").read()
    assert b[5] in (FUZZ, ZONA), b[:8]
    fc, cost_file, loaded = evaluate(L, R.name, path, q)
    assert np.array_equal(fc, counts)
    w8 = os.path.join("/tmp", fname + "_w8.idx")
    oracle.two_level_widen(path, w8)
    c8, k8, l8 = evaluate(L, R.name, w8, q)
    assert np.array_equal(c8, counts) and np.array_equal(k8, cost_file) and np.array_equal(l8, loaded)
    os.remove(w8)
    # every count is the column's own
    lo, lop, hi, rop = q
    for i in range(len(lo)):
        m = np.ones(col.size, bool)
        if lop[i] == 1: m &= lo[i] < col
        if lop[i] == 2: m &= lo[i] <= col
        if rop[i] == 1: m &= col < hi[i]
        if rop[i] == 2: m &= col <= hi[i]
        assert int(m.sum()) == counts[i], (i, int(m.sum()), counts[i])
    _, _, nc, cb, co, words, woff = oracle.two_level_file_parts(b)
    np.savez_compressed(os.path.join(OUT, fname + ".npz"), col=col, lo=lo, lop=lop, hi=hi, rop=rop,
                        counts=counts, cost_mem=cost_mem, cost_file=cost_file, loaded=loaded,
                        nc=np.array(nc), cbounds=cb, coffsets=co, cbits=words, cbits_off=woff)
    opts = len(set(int(x) for x in loaded))
    print(fname, "type", b[5], "nc", nc, "file", len(b), "distinct coarse loads", opts)


def main():
    os.makedirs(OUT, exist_ok=True)
    L = oracle.reference().lib
    L.ref_two_level_eval.restype = C.c_int64
    rng = np.random.default_rng(20261204)
    n = 20000
    col = np.concatenate([np.sort(rng.integers(0, 500, n // 2)), rng.integers(0, 500, n - n // 2)]).astype(np.int32)
    make(L, rng, "fuzz_i32", col.copy(), "<binning none/><encoding interval-equality/>")
    make(L, rng, "fuzz_i32_nc8", col.copy(), "<binning none/><encoding interval-equality ncoarse=8/>")
    make(L, rng, "zona_i32", col.copy(), "<binning none/><encoding equality-equality/>")
    colf = (np.round(rng.standard_normal(n) * 80.0) / 2.0).astype(np.float32)
    make(L, rng, "zona_f32", colf, "<binning none/><encoding equality-equality/>")


if __name__ == "__main__":
    main()
