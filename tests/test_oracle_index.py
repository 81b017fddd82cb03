"""Pins the oracle's relic / bin restatement against the reference FastBit
library built from source (oracle/_ref): index words (keys, offsets,
serialized bitmaps) and range-query answers on the same columns.  Runs on
the CPU; skipped where the reference library is not built."""
import itertools

import numpy as np
import pytest

import oracle as O

_names = itertools.count()


def _name(tag):
    return f"t{next(_names)}_{tag}"


def _same(oi, keys, offs, bms):
    k, o, b = oi.write()
    assert np.array_equal(k, keys)
    assert np.array_equal(o, offs)
    assert np.array_equal(b, bms)


def _ranges(rng, vmin, vmax, n, frac=True):
    out = []
    for _ in range(n):
        a, b = sorted(rng.uniform(vmin - 2, vmax + 2, size=2))
        if not frac:
            a, b = float(np.floor(a)), float(np.ceil(b))
        lop = int(rng.integers(0, 3))
        rop = int(rng.integers(0, 3))
        if lop == 0 and rop == 0:
            rop = 2
        out.append((float(a), lop, float(b), rop))
    return out


@pytest.mark.parametrize("n,card", [(1, 5), (30, 7), (31, 7), (62, 3), (1000, 50), (31 * 700 + 11, 300),
                                    (200003, 40)])
def test_relic_card_words(ref, n, card):
    rng = np.random.default_rng(n + card)
    ind = rng.integers(0, card + 3, size=n).astype(np.uint32)   # some rows >= card: NULL
    k, o, b = O.ref_relic_card(ind, card)
    _same(O.OracleIndex.relic_card(ind, card), k, o, b)


REL_COLS = {
    "int32": lambda rng, n: rng.integers(-150, 150, size=n).astype(np.int32),
    "uint16": lambda rng, n: rng.integers(100, 400, size=n).astype(np.uint16),
    "float32": lambda rng, n: np.round(rng.normal(size=n) * 8).astype(np.float32) / 4,
    # one value holds most rows: its bitmap is stored decompressed
    "heavy": lambda rng, n: np.where(rng.random(n) < 0.8, 7, rng.integers(-40, 40, size=n)).astype(np.int32),
}


@pytest.mark.parametrize("kind", sorted(REL_COLS))
@pytest.mark.parametrize("n", [1000, 31 * 2000 + 13, 200003])
def test_relic_values_words_and_queries(ref, kind, n):
    rng = np.random.default_rng(n)
    col = REL_COLS[kind](rng, n)
    R = O.RefIapi(_name("relic"), col, "<binning none/><encoding equality/>")
    oi = O.OracleIndex.relic_values(col)
    _same(oi, R.keys, R.offsets, R.bitmaps)
    for q in _ranges(rng, float(col.min()), float(col.max()), 12, frac=kind != "int32"):
        c, r, _ = oi.evaluate_plan(*q)
        rc, rr = R.query_range(*q)
        assert c == rc, q
        assert np.array_equal(r, rr), q


BIN_COLS = {
    "normal": lambda rng, n: rng.normal(size=n).astype(np.float32),
    "mix": lambda rng, n: np.where(rng.random(n) < 0.3, np.float32(0.5),
                                   rng.normal(size=n)).astype(np.float32),
    "double": lambda rng, n: rng.uniform(-1e3, 1e3, size=n),
}


@pytest.mark.parametrize("kind", sorted(BIN_COLS))
@pytest.mark.parametrize("n,nbins", [(1000, 17), (31 * 3000 + 5, 100), (300001, 1000)])
def test_bin_words_and_queries(ref, kind, n, nbins):
    rng = np.random.default_rng(n + nbins)
    col = BIN_COLS[kind](rng, n)
    R = O.RefIapi(_name("bin"), col, f"<binning nbins={nbins} equal-weight/>")
    bounds = O.bounds_from_keys(R.keys)
    oi = O.OracleIndex.bin_bounds(col, bounds)
    _same(oi, R.keys, R.offsets, R.bitmaps)
    for q in _ranges(rng, float(col.min()), float(col.max()), 12):
        c, r, _ = oi.evaluate_plan(*q)
        rc, rr = R.query_range(*q)
        assert c == rc, q
        assert np.array_equal(r, rr), q


def test_bin_int_queries(ref):
    """Integer column under a bin index: integral bounds agree with the
    reference's answers (fractional bounds are covered against the
    restated edge rule in the GPU tests)."""
    rng = np.random.default_rng(5)
    n = 31 * 900 + 7
    col = rng.integers(-500, 500, size=n).astype(np.int32)
    R = O.RefIapi(_name("binint"), col, "<binning nbins=40 equal-weight/>")
    oi = O.OracleIndex.bin_bounds(col, O.bounds_from_keys(R.keys))
    _same(oi, R.keys, R.offsets, R.bitmaps)
    for q in _ranges(rng, -500, 500, 20, frac=False):
        c, r, _ = oi.evaluate_plan(*q)
        rc, rr = R.query_range(*q)
        assert c == rc, q
        assert np.array_equal(r, rr), q


def _fractional(rng, vmin, vmax, n):
    """Ranges with fractional bounds, where part::doScan's conversion of
    the bounds to an integer type changes the answer."""
    out = []
    for _ in range(n):
        a, b = sorted(rng.uniform(vmin - 3, vmax + 3, size=2))
        lop = int(rng.integers(0, 3))
        rop = int(rng.integers(0, 3))
        if lop == 0 and rop == 0:
            lop = 2
        out.append((float(np.round(a) - 0.25 * rng.integers(1, 4)), lop,
                    float(np.round(b) + 0.25 * rng.integers(0, 4)), rop))
    return out


INT_COLS = {
    "uniform": lambda rng, n: rng.integers(-60, 60, size=n).astype(np.int32),
    "heavy": lambda rng, n: np.where(rng.random(n) < 0.8, 7, rng.integers(-40, 40, size=n)).astype(np.int32),
    "wide": lambda rng, n: rng.integers(-5000, 5000, size=n).astype(np.int32),
    "u16": lambda rng, n: rng.integers(100, 700, size=n).astype(np.uint16),
}


def _conjunction(oi, lo, lop, hi, rop):
    """The two one-sided terms of lo OP1 x AND x OP2 hi, each under its own
    plan choice, intersected (countQuery::doEvaluate, LOGICAL_AND,
    src/countQuery.cpp:953-963)."""
    r = None
    if lop:
        r = oi.evaluate_plan(lo, lop, 0.0, 0)[1]
    if rop:
        r2 = oi.evaluate_plan(0.0, 0, hi, rop)[1]
        r = r2 if r is None else np.intersect1d(r, r2)
    return r


@pytest.mark.parametrize("kind", sorted(INT_COLS))
@pytest.mark.parametrize("n", [1000, 31 * 1500 + 7, 100003])
def test_relic_plan_choice(ref, kind, n):
    """On integer columns with fractional bounds the reference's answer
    follows whichever of index and scan its cost rule picked for the range
    term; the oracle's restated rule must pick the same."""
    rng = np.random.default_rng(n + len(kind))
    col = INT_COLS[kind](rng, n)
    R = O.RefIapi(_name("rplan"), col, "<binning none/><encoding equality/>")
    oi = O.OracleIndex.relic_values(col)
    seen = set()
    for q in _fractional(rng, float(col.min()), float(col.max()), 40):
        c, r, used = oi.evaluate_plan(*q)
        rc, rr = R.query_range(*q)
        assert c == rc, (q, used, oi.cost(*q))
        assert np.array_equal(r, rr), q
        ci = oi.evaluate_plan(*q, plan=1)[0]
        cs = oi.evaluate_plan(*q, plan=3)[0]
        if ci != cs:
            seen.add(used)
    if n > 1000 and kind != "u16":   # truncation of positive bounds is exact
        assert seen, "no query told the two plans apart"


@pytest.mark.parametrize("nbins", [2, 5, 40])
def test_bin_plan_choice(ref, nbins):
    rng = np.random.default_rng(nbins)
    n = 31 * 1200 + 9
    col = rng.integers(-300, 300, size=n).astype(np.int32)
    R = O.RefIapi(_name("bplan"), col, f"<binning nbins={nbins} equal-weight/>")
    oi = O.OracleIndex.bin_bounds(col, O.bounds_from_keys(R.keys))
    _same(oi, R.keys, R.offsets, R.bitmaps)
    for q in _fractional(rng, -300, 300, 40):
        c, r, used = oi.evaluate_plan(*q)
        rc, rr = R.query_range(*q)
        assert c == rc, (q, used, oi.cost(*q))
        assert np.array_equal(r, rr), q


@pytest.mark.parametrize("kind", ["float32", "int32"])
def test_conjunction_of_one_sided_terms(ref, kind):
    """lo OP1 x AND x OP2 hi built with fastbit_selection_combine equals the
    intersection of the two one-sided terms (float bounds on a float column
    and integral bounds on an integer column, where index and scan agree)."""
    rng = np.random.default_rng(17)
    n = 31 * 2500 + 3
    if kind == "float32":
        col = np.round(rng.normal(size=n) * 16).astype(np.float32) / 8
        R = O.RefIapi(_name("conjf"), col, "<binning nbins=64 equal-weight/>")
        oi = O.OracleIndex.bin_bounds(col, O.bounds_from_keys(R.keys))
        qs = _ranges(rng, float(col.min()), float(col.max()), 30)
    else:
        col = rng.integers(-200, 200, size=n).astype(np.int32)
        R = O.RefIapi(_name("conji"), col, "<binning none/><encoding equality/>")
        oi = O.OracleIndex.relic_values(col)
        qs = _ranges(rng, -200, 200, 30, frac=False)
    for q in qs:
        want = _conjunction(oi, *q)
        rc, rr = R.query(*q)
        assert np.array_equal(want, rr), q
        assert np.array_equal(oi.evaluate_plan(*q)[1], rr), q


def test_config1_relic_fixture():
    """The oracle reproduces the reference's relic index and answers on the
    2^20-row int32 workload (tests/golden/config1_relic_2p20.json)."""
    import hashlib
    import json
    import os
    import sys
    here = os.path.join(os.path.dirname(__file__), "golden")
    sys.path.insert(0, here)
    from make_config1 import workload
    meta = json.load(open(os.path.join(here, "config1_relic_2p20.json")))
    col, lo, hi = workload()
    oi = O.OracleIndex.relic_values(col)
    k, o, b = oi.write()
    assert hashlib.sha256(k.tobytes()).hexdigest() == meta["keys_sha256"]
    assert hashlib.sha256(o.tobytes()).hexdigest() == meta["offsets_sha256"]
    assert hashlib.sha256(b.tobytes()).hexdigest() == meta["bitmaps_sha256"]
    for i in range(0, 256, 8):
        c, r, _ = oi.evaluate_plan(float(lo[i]), 2, float(hi[i]), 1)
        assert c == meta["counts"][i]
        assert hashlib.sha256(r.astype(np.uint32).tobytes()).hexdigest() == meta["rids_sha256"][i]


def _kat():
    import json
    import os
    d = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "kat_query_count.json")))
    cols = {}
    for t, cs in d["tables"].items():
        for c, v in cs.items():
            cols[(t, c)] = np.array(v["values"], np.float32 if v["type"] == "float32" else np.int32)
    return cols, d["queries"]


def test_query_count_kat():
    """The reference's tests/scripts/query-count.list entries that are
    conjunctions of range terms: the oracle's relic indexes give the listed
    counts (each term under the reference's own plan choice)."""
    cols, queries = _kat()
    idx = {k: O.OracleIndex.relic_values(v) for k, v in cols.items()}
    for q in queries:
        r = None
        for (c, lo, lop, hi, rop) in q["terms"]:
            rr = idx[(q["table"], c)].evaluate_plan(lo, lop, hi, rop)[1]
            r = rr if r is None else np.intersect1d(r, rr)
        assert len(r) == q["count"], q["text"]


def test_estimate_matches_reference(orc, ref):
    """index::estimate(expr, lower, upper): the oracle's lower and upper rows
    against the reference's bin::estimate on the same bins (float and int32
    columns, equal-weight bins, every branch: no candidate, the cost
    give-up, hits with edge bins, edge bins only).  The reference index is
    first warmed by one estimate, so that sumBins has filled offset64 (the
    state the restatement describes)."""
    import oracle as O
    rng = np.random.default_rng(3)

    def bits(words, n):
        b = np.zeros(n, np.uint8)
        b[orc.rids(words)] = 1
        return b
    kinds = set()
    for t in range(24):
        n = int(rng.integers(1000, 40000))
        col = (rng.normal(size=n) if t % 2 else rng.integers(0, 500, size=n)).astype(np.float32 if t % 2 else np.int32)
        R = O.RefIapi(f"est{t}", col, "<binning nbins=%d equal-weight/>" % int(rng.integers(5, 200)))
        m = float(np.median(col))
        R.estimate(m, 2, m + 1e-3, 1)
        oi = O.OracleIndex.bin_bounds(col, O.bounds_from_keys(R.keys))
        for q in range(8):
            a, b = sorted(rng.choice(col, 2).astype(np.float64))
            if t % 3 == 0:
                a -= 0.25
                b += 0.5
            if q == 7:
                b = a                      # an empty range
            lop, rop = int(rng.integers(1, 3)), int(rng.integers(1, 3))
            rl, ru = R.estimate(a, lop, b, rop)
            ol, ou = oi.estimate(a, lop, b, rop)
            assert np.array_equal(bits(rl, n), bits(ol, n)), (t, q)
            assert (len(ru) == 1) == (len(ou) == 1), (t, q)
            if len(ru) > 1:
                assert np.array_equal(bits(ru, n), bits(ou, n)), (t, q)
                kinds.add("all" if orc.cnt(ru)[0] == n else "upper")
            else:
                kinds.add("cleared")
    assert kinds == {"all", "upper", "cleared"}
