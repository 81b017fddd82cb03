"""Writes index-file fixtures with the reference FastBit library.

Each fixture is an index file the reference wrote (relic::write /
bin::write) next to an .npz holding what the reference reads back from it:

  relic_i32.idx   relic over 9300 int32 values (9300 = 300 * 31 rows, no
                  active word): the first half in sorted runs over [0, 300),
                  the second half uniform in [0, 300) -- fills and literals
  bin_f32.idx     bin with "<binning nbins=16 equal-weight/>" over 5000
                  float32 values, normal(0, 1)
  relic_card.idx  relic(0, 32, ind) over 4001 values, geometric(0.1) - 1
                  capped at 39 (bitmaps of many sizes):
                  rows with a value >= 32 are in no bitmap; its terms add
                  the ranges where sumBins' choice between the bitmaps of
                  the range and the complement of the others depends on
                  measuring bitmaps by their serialized sizes
  *_w8.idx        the same files with 8-byte bitmap positions, the layout
                  the reference writes for indexes beyond 2 GiB (rewritten
                  here; the reference reads them back to the same triple)

The .npz of a fixture holds the column, the reference's in-memory form of
the file (keys, offsets in words, serialized bitmaps), 48 range terms with
relic::locate / bin::locate on the file's index (cand0, hit0, hit1, cand1)
and the reference's hit counts, and index::sumBins counts for pairs (ib, ie).

    python tests/golden/make_idx_files.py
"""
from __future__ import annotations

import os
import struct
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

import oracle  # noqa: E402

OUT = os.path.join(HERE, "idx")


def widen(src, dst):
    """Rewrite an index file with 4-byte bitmap positions into the 8-byte
    layout: relic keeps its values first, bin its positions first."""
    b = open(src, "rb").read()
    assert b[6] == 4
    kind = b[5]
    nobs = struct.unpack_from("<I", b, 12)[0]
    if kind == 7:      # relic: header, nrows nobs card, vals at 24, positions
        obeg = 24 + 8 * nobs
        pos = np.frombuffer(b, np.int32, nobs + 1, obeg).astype(np.int64)
        shift = 4 * (nobs + 1)
        head = bytearray(b[:obeg])
        head[6] = 8
        out = bytes(head) + (pos + shift).astype("<i8").tobytes() + b[obeg + 4 * (nobs + 1):]
    else:              # bin: header, nrows nobs, positions, pad, 3 key arrays
        oend4 = 16 + 4 * (nobs + 1)
        kbeg4 = 8 * ((7 + oend4) // 8)
        oend8 = 16 + 8 * (nobs + 1)
        kbeg8 = 8 * ((7 + oend8) // 8)
        pos = np.frombuffer(b, np.int32, nobs + 1, 16).astype(np.int64)
        head = bytearray(b[:16])
        head[6] = 8
        out = (bytes(head) + (pos + kbeg8 - kbeg4).astype("<i8").tobytes() + b"\0" * (kbeg8 - oend8)
               + b[kbeg4:])
    with open(dst, "wb") as f:
        f.write(out)


def queries(rng, vmin, vmax, n=48):
    span = vmax - vmin
    lo = rng.uniform(vmin - 0.1 * span, vmax + 0.1 * span, n)
    hi = lo + rng.uniform(-0.1 * span, 0.6 * span, n)
    lop = rng.integers(0, 3, n)
    rop = rng.integers(0, 3, n)
    both = (lop == 0) & (rop == 0)
    rop[both] = 2
    if np.issubdtype(type(vmin), np.integer) or float(vmin).is_integer():
        # half the bounds on integers, so that < and <= differ
        k = rng.random(n) < 0.5
        lo[k] = np.round(lo[k])
        hi[k] = np.round(hi[k])
    return lo, lop.astype(np.int32), hi, rop.astype(np.int32)


def sum_pairs(rng, nobs):
    p = [(0, 0), (0, nobs), (0, 1), (nobs - 1, nobs), (3, 7), (2, nobs - 2), (nobs // 2, nobs + 5),
         (5, 3), (1, nobs), (0, nobs - 1)]
    for _ in range(14):
        a, b = sorted(rng.integers(0, nobs + 1, 2))
        p.append((int(a), int(b)))
    return np.array(p, np.uint32)


def record(name, col, path, lo, lop, hi, rop, counts, rng):
    kind, keys, offs, bms = oracle.ref_file_triple(path)
    nobs = offs.size - 1
    loc = np.array([oracle.ref_file_locate(path, float(a), int(b), float(c), int(d))
                    for a, b, c, d in zip(lo, lop, hi, rop)], np.uint32)
    pairs = sum_pairs(rng, nobs)
    sb = np.array([oracle.ref_file_sum_bins(path, int(a), int(b)) for a, b in pairs], np.int64)
    w8 = path[:-4] + "_w8.idx"
    widen(path, w8)
    k8, keys8, offs8, bms8 = oracle.ref_file_triple(w8)
    assert k8 == kind and np.array_equal(keys8, keys) and np.array_equal(offs8, offs) \
        and np.array_equal(bms8, bms)
    np.savez_compressed(os.path.join(OUT, name + ".npz"), col=col, kind=np.array(kind), keys=keys,
                        offsets=offs, bitmaps=bms, lo=lo, lop=lop, hi=hi, rop=rop, locate=loc,
                        counts=np.array(counts, np.int64), pairs=pairs, sum_bins=sb)
    print(name, kind, nobs, bms.size, int(np.sum(counts)))


def main():
    os.makedirs(OUT, exist_ok=True)
    rng = np.random.default_rng(20261203)

    # relic over int32 (fastbit_iapi_build_index, equality encoding)
    n = 9300
    col = np.concatenate([np.arange(n // 2) * 300 // (n // 2),
                          rng.integers(0, 300, n - n // 2)]).astype(np.int32)
    R = oracle.RefIapi("idx_relic", col, "<binning none/><encoding equality/>")
    path = os.path.join(OUT, "relic_i32.idx")
    oracle.ref_write_index_file("idx_relic", path)
    q = queries(rng, 0, 299)
    counts = []
    for a, b, c, d in zip(*q):
        h = oracle.ref_file_locate(path, float(a), int(b), float(c), int(d))
        cnt = oracle.ref_file_sum_bins(path, h[1], h[2])
        assert cnt == R.query_range(float(a), int(b), float(c), int(d), want_rids=False)[0]
        counts.append(cnt)
    record("relic_i32", col, path, *q, counts, rng)

    # bin over float32, equal-weight
    col = rng.standard_normal(5000).astype(np.float32)
    R = oracle.RefIapi("idx_bin", col, "<binning nbins=16 equal-weight/>")
    path = os.path.join(OUT, "bin_f32.idx")
    oracle.ref_write_index_file("idx_bin", path)
    q = queries(rng, -3.0, 3.0)
    counts = [R.query_range(float(a), int(b), float(c), int(d), want_rids=False)[0] for a, b, c, d in zip(*q)]
    record("bin_f32", col, path, *q, counts, rng)

    # relic(0, card, ind) with rows in no bitmap
    ind = np.minimum(rng.geometric(0.1, 4001) - 1, 39).astype(np.uint32)
    path = os.path.join(OUT, "relic_card.idx")
    oracle.ref_relic_card_file(ind, 32, path)
    q = queries(rng, 0, 31)
    # ranges [ib, ie) where sumBins' straight/complement choice depends on
    # the measure (serialized sizes, as the file's offsets give, against
    # bitvector::bytes()): with rows in no bitmap the two choices differ
    _, _, offs, _ = oracle.ref_file_triple(path)
    ser = 4.0 * np.diff(offs)
    byt = np.where(np.diff(offs) > 0, 4.0 * (np.diff(offs) - 2) + 40.0, 0.0)
    flips = [(a, e) for a in range(32) for e in range(a + 1, 33)
             if (ser[a:e].sum() <= ser.sum() // 2) != (byt[a:e].sum() <= byt.sum() // 2)][:8]
    assert flips, "no range where the measure decides"
    q = (np.concatenate([q[0], [float(a) for a, _ in flips]]),
         np.concatenate([q[1], [2] * len(flips)]).astype(np.int32),
         np.concatenate([q[2], [float(e) for _, e in flips]]),
         np.concatenate([q[3], [1] * len(flips)]).astype(np.int32))
    counts = []
    for a, b, c, d in zip(*q):
        h = oracle.ref_file_locate(path, float(a), int(b), float(c), int(d))
        counts.append(oracle.ref_file_sum_bins(path, h[1], h[2]))
    record("relic_card", ind, path, *q, counts, rng)


if __name__ == "__main__":
    main()
