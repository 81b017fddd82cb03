"""relic / bin build and range evaluation on the MI355X against the oracle
(words of the index, hit counts, row ids) for every plan."""
import numpy as np
import pytest

import oracle as O

pytestmark = pytest.mark.gpu
PLANS = (1, 2, 3)   # literal, compressed, scan


@pytest.fixture(scope="module")
def L():
    import libfastlib_amd as L
    L.lib()
    return L


def _same_words(gi, oi):
    gk, go, gb = gi.deconstruct()
    ok, oo, ob = oi.write()
    assert np.array_equal(gk, ok)
    assert np.array_equal(go, oo)
    assert np.array_equal(gb, ob)


def _queries(rng, lo_max, n, ints=True):
    out = []
    for _ in range(n):
        if ints:
            lo = int(rng.integers(-2, lo_max + 2))
            hi = lo + int(rng.integers(1, max(2, lo_max // 3)))
        else:
            lo, hi = sorted(rng.normal(size=2) * 1.5)
        lop = int(rng.integers(0, 3))
        rop = int(rng.integers(0, 3))
        if lop == 0 and rop == 0:
            lop = 2
        out.append((float(lo), lop, float(hi), rop))
    return out


@pytest.mark.parametrize("n", [1, 30, 31, 62, 1000, 31 * 300 + 7, 200003])
def test_relic_card(L, n):
    rng = np.random.default_rng(n)
    card = 50
    ind = rng.integers(0, card, size=n).astype(np.uint32)
    gi = L.Index.relic_card(ind, card)
    oi = O.OracleIndex.relic_card(ind, card)
    _same_words(gi, oi)
    for (lo, lop, hi, rop) in _queries(rng, card, 25):
        c, r = oi.evaluate(lo, lop, hi, rop)
        for plan in PLANS:
            h, used, _ = gi.evaluate(lo, lop, hi, rop, plan)
            assert h == c, (lo, lop, hi, rop, plan)
        assert np.array_equal(gi.rids(lo, lop, hi, rop, 1), r)
        assert np.array_equal(gi.rids(lo, lop, hi, rop, 3), r)


@pytest.mark.parametrize("dtype", [np.int32, np.uint16, np.float32])
def test_relic_values(L, dtype):
    rng = np.random.default_rng(7)
    n = 31 * 2000 + 13
    col = rng.integers(0, 300, size=n).astype(dtype)
    gi = L.Index.relic(col)
    oi = O.OracleIndex.relic_values(col)
    _same_words(gi, oi)
    for (lo, lop, hi, rop) in _queries(rng, 300, 20):
        c, r = oi.evaluate(lo, lop, hi, rop)
        for plan in PLANS:
            h, _, bv = gi.evaluate(lo, lop, hi, rop, plan, want_bitvector=True)
            assert h == c
            assert np.array_equal(bv.rids(), r)


@pytest.mark.parametrize("n", [1000, 31 * 1000 + 5, 300001])
def test_bin_float(L, n):
    rng = np.random.default_rng(n)
    col = rng.normal(size=n).astype(np.float32)
    # bounds: the reference's equal-weight boundaries when the reference is
    # here, else quantile boundaries (the evaluation is exact either way)
    qs = np.quantile(col.astype(np.float64), np.linspace(0, 1, 60)[1:-1])
    bounds = np.concatenate([np.unique(qs), [np.finfo(np.float64).max]])
    gi = L.Index.bin(col, bounds)
    oi = O.OracleIndex.bin_bounds(col, bounds)
    _same_words(gi, oi)
    for (lo, lop, hi, rop) in _queries(rng, 0, 30, ints=False):
        c, r = oi.evaluate(lo, lop, hi, rop)
        for plan in PLANS:
            h, _, _ = gi.evaluate(lo, lop, hi, rop, plan)
            assert h == c, (lo, lop, hi, rop, plan)
        assert np.array_equal(gi.rids(lo, lop, hi, rop, 1), r)


def test_bin_int_edges(L):
    """Integer column under a bin index: the edge check uses part::doScan's
    cast-to-T bound rules (fractional and negative bounds)."""
    rng = np.random.default_rng(3)
    n = 31 * 800 + 3
    col = rng.integers(-500, 500, size=n).astype(np.int32)
    bounds = np.array([-400.5, -250.0, -100.25, 0.0, 77.0, 150.5, 333.0, np.finfo(np.float64).max])
    gi = L.Index.bin(col, bounds)
    oi = O.OracleIndex.bin_bounds(col, bounds)
    _same_words(gi, oi)
    for lo, hi in ((-300.5, 100.5), (-3.5, 3.5), (-250.0, -100.25), (10.2, 10.8), (-1e12, 1e12)):
        for lop in (1, 2):
            for rop in (1, 2):
                c, r = oi.evaluate(lo, lop, hi, rop)
                for plan in PLANS:
                    assert gi.evaluate(lo, lop, hi, rop, plan)[0] == c
                assert np.array_equal(gi.rids(lo, lop, hi, rop, 1), r)
