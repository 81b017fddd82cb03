"""Index files written by the reference (tests/golden/idx, made by
tests/golden/make_idx_files.py): the file layout of relic::write and
bin::write (src/irelic.cpp:354-642, src/ibin.cpp:5530-5760), checked
against the reference's own reading of the same files, and the oracle's
sumBins against the reference's counts on them."""
from __future__ import annotations

import os
import struct

import numpy as np
import pytest

import oracle

IDX = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "idx")
NAMES = ["relic_i32", "bin_f32", "relic_card"]


def fixture(name):
    with np.load(os.path.join(IDX, name + ".npz")) as z:
        return {k: z[k] for k in z.files}


def parse(path):
    """The in-memory form (keys, word offsets, words) of an index file, read
    by the layout alone: relic = header, nrows nobs card, values at byte 24,
    nobs+1 positions; bin = header, nrows nobs, positions, then at the next
    8-byte boundary bounds, maxval, minval (keys = minval ++ maxval)."""
    b = open(path, "rb").read()
    assert b[:5] == b"#IBIS" and b[7] == 0
    kind, w = b[5], b[6]
    nrows, nobs = struct.unpack_from("<II", b, 8)
    pt = np.int32 if w == 4 else np.int64
    if kind == 7:
        assert struct.unpack_from("<I", b, 16)[0] == nobs
        keys = np.frombuffer(b, np.float64, nobs, 24)
        pos = np.frombuffer(b, pt, nobs + 1, 24 + 8 * nobs).astype(np.int64)
        bounds = None
    else:
        assert kind == 0
        pos = np.frombuffer(b, pt, nobs + 1, 16).astype(np.int64)
        kb = 8 * ((7 + 16 + w * (nobs + 1)) // 8)
        bounds = np.frombuffer(b, np.float64, nobs, kb)
        mx = np.frombuffer(b, np.float64, nobs, kb + 8 * nobs)
        mn = np.frombuffer(b, np.float64, nobs, kb + 16 * nobs)
        keys = np.concatenate([mn, mx])
    words = np.frombuffer(b, np.uint32, (pos[-1] - pos[0]) // 4, pos[0])
    return ("relic" if kind == 7 else "bin"), nrows, keys, (pos - pos[0]) // 4, words, bounds


@pytest.mark.parametrize("w8", [False, True])
@pytest.mark.parametrize("name", NAMES)
def test_layout_matches_reference_reading(name, w8):
    f = fixture(name)
    kind, nrows, keys, offs, words, _ = parse(os.path.join(IDX, name + ("_w8" if w8 else "") + ".idx"))
    assert kind == str(f["kind"])
    assert nrows == f["col"].size
    assert np.array_equal(keys, f["keys"])
    assert np.array_equal(offs, f["offsets"])
    assert np.array_equal(words, f["bitmaps"])


@pytest.mark.parametrize("name", NAMES)
def test_bitmaps_cover_every_row(name):
    """Every non-null bitmap is code words ++ [active value] ++ active bit
    count and covers the column's rows; relic bitmaps are disjoint and their
    ones are the rows of each value."""
    f = fixture(name)
    n = f["col"].size
    g, anb = divmod(n, 31)
    offs, bms = f["offsets"], f["bitmaps"]
    total = 0
    for b in range(offs.size - 1):
        s, e = int(offs[b]), int(offs[b + 1])
        if s == e:
            continue
        assert bms[e - 1] == anb
        code = bms[s:e - (2 if anb else 1)]
        fill = (code >> 31) == 1
        groups = int(np.sum(np.where(fill, code & 0x3FFFFFFF, 1)))
        assert groups == g
        ones = int(np.sum(np.where(fill, ((code >> 30) & 1) * 31 * (code & 0x3FFFFFFF),
                                   [bin(int(x)).count("1") for x in code])))
        if anb:
            ones += bin(int(bms[e - 2])).count("1")
        if str(f["kind"]) == "relic":
            v = f["keys"][b]
            assert ones == int(np.sum(f["col"] == v))
        total += ones
    if name == "relic_card":
        assert total == int(np.sum(f["col"] < 32))
    else:
        assert total == n


def test_relic_card_counts_oracle(orc):
    """relic(0, card, ind) leaves rows with ind >= card in no bitmap, so
    sumBins' choice between the bitmaps of the range and the complement of
    the others shows in the count.  The reference chose by the serialized
    sizes of the file's offsets; the oracle's relic_card must agree."""
    f = fixture("relic_card")
    ox = oracle.OracleIndex.relic_card(f["col"], 32)
    for lo, lop, hi, rop, c in zip(f["lo"], f["lop"], f["hi"], f["rop"], f["counts"]):
        assert ox.evaluate(float(lo), int(lop), float(hi), int(rop))[0] == c


def test_fixture_consistency_with_reference(ref):
    """The committed files still read back, through the reference, to the
    committed triples, locate results and sumBins counts."""
    for name in NAMES:
        f = fixture(name)
        path = os.path.join(IDX, name + ".idx")
        kind, keys, offs, bms = oracle.ref_file_triple(path)
        assert kind == str(f["kind"])
        assert np.array_equal(keys, f["keys"]) and np.array_equal(offs, f["offsets"])
        assert np.array_equal(bms, f["bitmaps"])
        for i in range(0, f["lo"].size, 7):
            assert oracle.ref_file_locate(path, float(f["lo"][i]), int(f["lop"][i]), float(f["hi"][i]),
                                          int(f["rop"][i])) == tuple(int(x) for x in f["locate"][i])
        for (a, b), c in zip(f["pairs"], f["sum_bins"]):
            assert oracle.ref_file_sum_bins(path, int(a), int(b)) == c
