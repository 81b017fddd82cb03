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


def test_relic_float_signed_zeros(L):
    """-0.0 and +0.0 are one value of index::mapValues' std::map, keyed by
    the zero that comes first (both orders)."""
    rng = np.random.default_rng(11)
    for first in (-0.0, 0.0):
        col = (np.round(rng.standard_normal(31 * 500 + 7) * 4.0) / 2.0).astype(np.float32)
        col[col == 0] = np.float32(0.0)
        z = np.nonzero(col == 0)[0]
        col[z[::2]] = np.float32(-0.0)
        col[z[0]] = np.float32(first)
        gi = L.Index.relic(col)
        oi = O.OracleIndex.relic_values(col)
        _same_words(gi, oi)
        keys = gi.deconstruct()[0]
        k0 = keys[keys == 0]
        assert k0.size == 1 and np.signbit(k0[0]) == np.signbit(first)
        c, r = oi.evaluate(-0.5, 2, 0.0, 2)
        assert gi.evaluate(-0.5, 2, 0.0, 2, 1)[0] == c == int(((col >= -0.5) & (col <= 0)).sum())


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
    cast-to-T bound rules (fractional and negative bounds); the scan plan
    is part::doScan over every row; AUTO answers as the reference does."""
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
                for plan in (1, 2):
                    assert gi.evaluate(lo, lop, hi, rop, plan)[0] == c
                assert np.array_equal(gi.rids(lo, lop, hi, rop, 1), r)
                cs, rs, _ = oi.evaluate_plan(lo, lop, hi, rop, plan=3)
                assert gi.evaluate(lo, lop, hi, rop, 3)[0] == cs
                assert np.array_equal(gi.rids(lo, lop, hi, rop, 3), rs)
                ca, _, used = oi.evaluate_plan(lo, lop, hi, rop)
                h, gused, _ = gi.evaluate(lo, lop, hi, rop, 0)
                assert h == ca
                assert (gused == 3) == (used == 3)


@pytest.mark.parametrize("model", [0, 1])
@pytest.mark.parametrize("kind", ["heavy", "wide", "relic_card"])
def test_auto_plan_matches_reference_rule(L, kind, model):
    """AUTO picks index or scan by the reference's cost rule (both cost
    models); on integer columns with fractional bounds the two disagree, so
    the counts show the choice."""
    rng = np.random.default_rng(len(kind) + 10 * model)
    n = 31 * 3000 + 17
    if kind == "heavy":
        col = np.where(rng.random(n) < 0.8, 7, rng.integers(-40, 40, size=n)).astype(np.int32)
        gi, oi = L.Index.relic(col), O.OracleIndex.relic_values(col)
    elif kind == "wide":
        col = rng.integers(-3000, 3000, size=n).astype(np.int32)
        gi, oi = L.Index.relic(col), O.OracleIndex.relic_values(col)
    else:
        col = rng.integers(0, 70, size=n).astype(np.uint32)
        gi, oi = L.Index.relic_card(col, 60), O.OracleIndex.relic_card(col, 60)
    gi.set_cost_model(model)
    lo_v, hi_v = float(col.min()), float(col.max())
    for _ in range(40):
        a, b = sorted(rng.uniform(lo_v - 3, hi_v + 3, size=2))
        q = (float(np.round(a) - 0.25), int(rng.integers(0, 3)), float(np.round(b) + 0.5),
             int(rng.integers(1, 3)))
        c, r, used = oi.evaluate_plan(*q, model=model)
        ic, sc, _, _ = oi.cost(*q, model=model)
        gic, gsc, gplan = gi.estimate_cost(*q)
        assert (gic, gsc) == (ic, sc), q
        h, gused, bv = gi.evaluate(*q, 0, want_bitvector=True)
        assert h == c, (q, used, gused)
        assert (gused == 3) == (used == 3), q
        assert np.array_equal(bv.rids(), r)


def test_batch_matches_single(L):
    """evaluate_batch gives the counts of one-at-a-time evaluation, per plan
    and with the heuristic's per-query choice."""
    rng = np.random.default_rng(11)
    n = 31 * 5000 + 9
    col = rng.normal(size=n).astype(np.float32)
    qs = np.quantile(col.astype(np.float64), np.linspace(0, 1, 200)[1:-1])
    gi = L.Index.bin(col, np.concatenate([np.unique(qs), [np.finfo(np.float64).max]]))
    oi = O.OracleIndex.bin_bounds(col, np.concatenate([np.unique(qs), [np.finfo(np.float64).max]]))
    qq = _queries(rng, 0, 64, ints=False)
    lo = np.array([q[0] for q in qq]); lop = np.array([q[1] for q in qq])
    hi = np.array([q[2] for q in qq]); rop = np.array([q[3] for q in qq])
    want = np.array([oi.evaluate(*q)[0] for q in qq], np.uint64)
    for plan in (0, 1, 2, 3):
        hits, used, ms = gi.evaluate_batch(lo, lop, hi, rop, plan)
        assert np.array_equal(hits, want), plan
        assert ms >= 0
        if plan:
            assert (used == plan).all()
        else:
            for q, u in zip(qq, used):
                assert gi.estimate_cost(*q)[2] == u
    mixed = rng.integers(1, 4, size=len(qq))
    hits, used, _ = gi.evaluate_batch(lo, lop, hi, rop, plans=mixed)
    assert np.array_equal(hits, want) and np.array_equal(used, mixed)


@pytest.mark.parametrize("kind", ["normal", "heavy", "zipf_int"])
def test_bin_equal_weight(L, kind):
    """Equal-weight bins built on the device (histogram on the GPU,
    divideCounts on the host) give the oracle's words for those bounds."""
    rng = np.random.default_rng(len(kind))
    n = 31 * 6000 + 11
    if kind == "normal":
        col = rng.normal(size=n).astype(np.float32)
    elif kind == "heavy":
        col = np.where(rng.random(n) < 0.3, np.float32(0.25), rng.normal(size=n)).astype(np.float32)
    else:
        col = np.minimum(rng.zipf(1.3, size=n), 100000).astype(np.int32)
    gi = L.Index.bin_equal_weight(col, 300)
    b = gi.bounds()
    assert b[-1] == np.finfo(np.float64).max and np.all(np.diff(b) > 0)
    oi = O.OracleIndex.bin_bounds(col, b)
    _same_words(gi, oi)
    for (lo, lop, hi, rop) in _queries(rng, 0, 10, ints=False):
        assert gi.evaluate(lo, lop, hi, rop, 1)[0] == oi.evaluate(lo, lop, hi, rop)[0]


@pytest.mark.slow
def test_bin_2p27_equal_weight_build(L):
    """The 2^27-row float column binned with nbins=10000 on the device:
    boundaries and words equal the reference's (tests/golden digests)."""
    import hashlib
    import json
    import os
    from libfastlib_amd import datagen as G
    here = os.path.join(os.path.dirname(__file__), "golden")
    meta = json.load(open(os.path.join(here, "config3_2p27.json")))
    ref_bounds = np.load(os.path.join(here, "config3_2p27_bounds.npz"))["ref_bounds"]
    col = G.float_mix(meta["seed"], meta["rows"], pool=100000, wnorm=0.7, s=1.2)
    gi = L.Index.bin_equal_weight(col, 10000)
    k, o, b = gi.deconstruct()
    assert hashlib.sha256(k.tobytes()).hexdigest() == meta["keys_sha256"]
    assert hashlib.sha256(o.tobytes()).hexdigest() == meta["offsets_sha256"]
    assert hashlib.sha256(b.tobytes()).hexdigest() == meta["bitmaps_sha256"]
    # the index drops empty inner bins, so compare the bounds it kept
    kept = gi.bounds()
    assert np.all(np.isin(kept, ref_bounds))


@pytest.mark.slow
def test_bin_2p27_reference_words(L):
    """The 2^27-row float index (10,000 equal-weight bins) built on the GPU is
    word-identical to the reference's (digests in tests/golden)."""
    import hashlib
    import json
    import os
    from libfastlib_amd import datagen as G
    here = os.path.join(os.path.dirname(__file__), "golden")
    meta = json.load(open(os.path.join(here, "config3_2p27.json")))
    bounds = np.load(os.path.join(here, "config3_2p27_bounds.npz"))["bounds"]
    col = G.float_mix(meta["seed"], meta["rows"], pool=100000, wnorm=0.7, s=1.2)
    gi = L.Index.bin(col, bounds)
    nobs, nrows, nw = gi.info()
    assert (nobs, nrows) == (meta["nobs"], meta["rows"])
    k, o, b = gi.deconstruct()
    assert hashlib.sha256(k.tobytes()).hexdigest() == meta["keys_sha256"]
    assert hashlib.sha256(o.tobytes()).hexdigest() == meta["offsets_sha256"]
    assert len(b) == meta["nwords"]
    assert hashlib.sha256(b.tobytes()).hexdigest() == meta["bitmaps_sha256"]


def test_config1_relic_2p20(L):
    """2^20 int32 values uniform in [0,1024), relic index, 256 range terms:
    index words, hit counts and row ids equal the reference's
    (tests/golden/config1_relic_2p20.json) under every plan."""
    import hashlib
    import json
    import os
    import sys
    here = os.path.join(os.path.dirname(__file__), "golden")
    sys.path.insert(0, here)
    from make_config1 import workload
    meta = json.load(open(os.path.join(here, "config1_relic_2p20.json")))
    col, lo, hi = workload()
    gi = L.Index.relic(col)
    k, o, b = gi.deconstruct()
    assert hashlib.sha256(k.tobytes()).hexdigest() == meta["keys_sha256"]
    assert hashlib.sha256(o.tobytes()).hexdigest() == meta["offsets_sha256"]
    assert hashlib.sha256(b.tobytes()).hexdigest() == meta["bitmaps_sha256"]
    want = np.array(meta["counts"], np.uint64)
    for plan in (0, 1, 2, 3):
        hits, _, _ = gi.evaluate_batch(lo, 2, hi, 1, plan)
        assert np.array_equal(hits, want), plan
    for i in range(0, 256, 4):
        r = gi.rids(float(lo[i]), 2, float(hi[i]), 1, 1)
        assert hashlib.sha256(r.astype(np.uint32).tobytes()).hexdigest() == meta["rids_sha256"][i]


def _table_indexes(L, cols):
    return [L.Index.relic(cols[0]), L.Index.bin_equal_weight(cols[1], 500),
            L.Index.bin_equal_weight(cols[2], 500), L.Index.relic(cols[3])]


def _oracle_indexes(gis, cols):
    out = []
    for gi, c in zip(gis, cols):
        k, o, b = gi.deconstruct()
        if len(k) == len(o) - 1:
            out.append(O.OracleIndex.relic_values(c))
        else:
            out.append(O.OracleIndex.bin_bounds(c, gi.bounds()))
    return out


@pytest.mark.parametrize("plan", [0, 1, 3])
def test_conjunctions(L, plan):
    """AND of 2-4 range terms over four columns (relic and bin indexes):
    counts equal the intersection of the oracle's per-term answers under
    the plan each term ran."""
    from libfastlib_amd import datagen as G
    n = 31 * 3000 + 17
    cols = G.table4(21, n)
    gis = _table_indexes(L, cols)
    ois = _oracle_indexes(gis, cols)
    qs = G.conj_queries(22, cols, 48)
    batch = L.Conjunctions([[(gis[ci], lo, 2, hi, 1) for (ci, lo, hi) in q] for q in qs], plan)
    hits, _ = batch.run()
    plans = batch.plans()
    k = 0
    for qi, q in enumerate(qs):
        r = None
        for (ci, lo, hi) in q:
            p = int(plans[k]); k += 1
            assert p in (1, 3)
            rr = ois[ci].evaluate_plan(lo, 2, hi, 1, plan=p)[1]
            r = rr if r is None else np.intersect1d(r, rr)
        assert hits[qi] == len(r), (qi, q)
    cfg = batch.tune()
    assert cfg["threads"] in (128, 256, 512)
    hits2, _ = batch.run()
    assert np.array_equal(hits, hits2)


@pytest.mark.parametrize("plan", [0, 1, 3])
def test_conjunction_rids_and_hit_vectors(L, orc, plan):
    """Row ids of every query of a conjunction batch (one more persistent
    launch after the counts) equal the intersection of the oracle's
    per-term rows; the hit vector of a query is the canonical WAH of those
    rows.  Four tiles of rows, with empty and non-empty tiles per query."""
    from libfastlib_amd import datagen as G
    n = 31 * 2048 * 3 + 31 * 700 + 13
    cols = G.table4(31, n)
    gis = _table_indexes(L, cols)
    ois = _oracle_indexes(gis, cols)
    qs = G.conj_queries(32, cols, 40, smax=0.6)
    batch = L.Conjunctions([[(gis[ci], lo, 2, hi, 1) for (ci, lo, hi) in q] for q in qs], plan)
    hits, _ = batch.run()
    off, rids = batch.rids()
    assert np.array_equal(np.diff(off), hits)
    plans = batch.plans()
    k = 0
    for qi, q in enumerate(qs):
        r = None
        for (ci, lo, hi) in q:
            p = int(plans[k]); k += 1
            rr = ois[ci].evaluate_plan(lo, 2, hi, 1, plan=p)[1]
            r = rr if r is None else np.intersect1d(r, rr)
        got = rids[off[qi]:off[qi + 1]]
        assert np.array_equal(got, r.astype(np.uint32)), (qi, len(got), len(r))
        if qi % 5 == 0:
            bits = np.zeros(n, np.uint8)
            bits[r] = 1
            assert np.array_equal(batch.result(qi).write(), orc.from_bits(bits)), qi


def _term_rows(col, lo, lop, hi, rop):
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


@pytest.mark.parametrize("plan", [0, 1, 3])
def test_conjunction_one_term_and_empty_tiles(L, orc, plan):
    """Queries of one term (counted in the term's own pass), queries with
    no hit (every tile empty after some term), and tiles emptied only by
    the last term, over a clustered column (narrow ranges touch few of the
    five tiles): counts, row ids and hit vectors against the column, before
    and after tuning, and the same terms as AND programs."""
    n = 31 * 2048 * 4 + 31 * 100 + 7
    rng = np.random.default_rng(41)
    a = np.sort(rng.integers(0, 1000, n)).astype(np.int32)
    b = rng.integers(0, 500, n).astype(np.int32)
    ia, ib = L.Index.relic(a), L.Index.relic(b)
    cols = {id(ia): a, id(ib): b}
    qs = [
        [(ia, 10.0, 2, 20.0, 1)],
        [(ib, 0.0, 2, 250.0, 1)],
        [(ia, 100.0, 2, 110.0, 1), (ia, 500.0, 2, 510.0, 1)],
        [(ib, 0.0, 2, 500.0, 1), (ia, 300.0, 2, 330.0, 1)],
        [(ia, 2000.0, 2, 3000.0, 1)],
        [(ib, 10.0, 1, 20.0, 2), (ib, 15.0, 2, 400.0, 1), (ia, 0.0, 2, 999.0, 2)],
        [(ia, 0.0, 2, 999.0, 2), (ib, 7.0, 2, 7.0, 2)],
    ]
    want = []
    for q in qs:
        m = np.ones(n, bool)
        for (ix, lo, lop, hi, rop) in q:
            m &= _term_rows(cols[id(ix)], lo, lop, hi, rop)
        want.append(np.nonzero(m)[0].astype(np.uint32))
    assert want[2].size == 0 and want[4].size == 0 and want[3].size > 0
    print("conjunction shapes", plan, flush=True)
    batch = L.Conjunctions(qs, plan)
    hits, _ = batch.run()
    assert [int(h) for h in hits] == [w.size for w in want]
    off, rids = batch.rids()
    for qi, w in enumerate(want):
        assert np.array_equal(rids[off[qi]:off[qi + 1]], w), qi
        bits = np.zeros(n, np.uint8)
        bits[w] = 1
        assert np.array_equal(batch.result(qi).write(), orc.from_bits(bits)), qi
    batch.tune()
    assert np.array_equal(batch.run()[0], hits)
    progs = [[0] if len(q) == 1 else [0] + sum(([k, -1] for k in range(1, len(q))), []) for q in qs]
    E = L.Expressions([(q, p) for q, p in zip(qs, progs)], plan)
    assert np.array_equal(E.run()[0], hits)
    eoff, erids = E.rids()
    assert np.array_equal(eoff, off) and np.array_equal(erids, rids)


@pytest.mark.parametrize("n", [31 * 4096, 31 * 4096 + 5, 31 * 2048 * 3 + 31 * 9 + 29, 1000])
def test_compressed_plan_words(L, orc, n):
    """The compressed plan (OR merged per tile, canonical WAH written with
    a look-back between tiles) gives the canonical words of the rows the
    range selects: narrow ranges (0-fills across tiles), wide ones (1-fills
    across tiles), complemented OR sets, no hit and every hit; tiles that
    end on a group boundary and a last tile holding only the active bits."""
    rng = np.random.default_rng(n)
    col = rng.normal(size=n).astype(np.float32)
    ints = np.sort(rng.integers(0, 300, n)).astype(np.int32)      # long runs per bitmap
    ixs = [(L.Index.bin_equal_weight(col, 120), col), (L.Index.relic(ints), ints)]
    for ix, c in ixs:
        vals = np.unique(c.astype(np.float64))
        cases = [(vals[0] - 1, vals[-1] + 1), (vals[-1] + 1, vals[-1] + 2), (vals[1], vals[2]),
                 (vals[len(vals) // 3], vals[-2]), (vals[2], vals[len(vals) // 2])]
        for _ in range(8):
            a, b = sorted(rng.choice(vals, 2, replace=False))
            cases.append((float(a), float(b) + 0.5))
        los = np.array([x for x, _ in cases])
        his = np.array([y for _, y in cases])
        hb, used, _ = ix.evaluate_batch(los, 2, his, 1, plan=2)
        assert np.all(used == 2)
        for (lo, hi), h in zip(cases, hb):
            bits = ((c.astype(np.float64) >= lo) & (c.astype(np.float64) < hi)).astype(np.uint8)
            want = orc.from_bits(bits)
            hc, u, bv = ix.evaluate(float(lo), 2, float(hi), 1, plan=2, want_bitvector=True)
            assert u == 2 and hc == int(bits.sum()) == int(h), (n, lo, hi)
            assert np.array_equal(bv.write(), want), (n, lo, hi)
            assert bv.cnt() == hc
            _, _, lv = ix.evaluate(float(lo), 2, float(hi), 1, plan=1, want_bitvector=True)
            assert np.array_equal(lv.write(), want)
            x = bv.copy()
            x.op(lv, 2)          # XOR of equal vectors: words from the compressed plan's chunk index
            assert x.cnt() == 0


def test_range_batch_rids(L):
    """Row ids of a range batch (the persistent range kernel) under the
    literal, mixed and scan-only plans equal the single-query row ids."""
    rng = np.random.default_rng(8)
    n = 31 * 2048 * 2 + 31 * 5 + 7
    col = rng.normal(size=n).astype(np.float32)
    ix = L.Index.bin_equal_weight(col, 300)
    lo = rng.normal(size=24)
    hi = lo + np.abs(rng.normal(size=24))
    mixed = np.where(np.arange(24) % 3 == 0, 3, 1).astype(np.int32)
    for kw in (dict(plan=1), dict(plans=mixed), dict(plan=3)):
        rb = L.RangeBatch(ix, lo, 2, hi, 1, **kw)
        off, rids = rb.rids()
        for i in range(24):
            want = np.nonzero((col >= lo[i]) & (col < hi[i]))[0].astype(np.uint32)
            assert np.array_equal(rids[off[i]:off[i + 1]], want), (kw, i)
            assert np.array_equal(ix.rids(float(lo[i]), 2, float(hi[i]), 1, 1), want)


def test_query_count_kat(L):
    """The reference's tests/scripts/query-count.list conjunctions of range
    terms (tests/golden/kat_query_count.json) through the batched
    conjunction engine, relic and bin indexes."""
    import json
    import os
    d = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "kat_query_count.json")))
    for kind in ("relic", "bin"):
        idx = {}
        for t, cs in d["tables"].items():
            for c, v in cs.items():
                a = np.array(v["values"], np.float32 if v["type"] == "float32" else np.int32)
                idx[(t, c)] = L.Index.relic(a) if kind == "relic" else L.Index.bin_equal_weight(a, 7)
        for table in d["tables"]:      # a batch covers the rows of one table
            sel = [q for q in d["queries"] if q["table"] == table]
            qs = [[(idx[(table, c)], lo, lop, hi, rop) for (c, lo, lop, hi, rop) in q["terms"]] for q in sel]
            hits, _ = L.Conjunctions(qs).run()
            assert [int(h) for h in hits] == [q["count"] for q in sel], (kind, table)


@pytest.mark.parametrize("seed", [0, 1])
def test_heavy_bitmaps_block_encoder(L, seed):
    """Bitmaps of more than 2^15 rows go through the block encoder: long
    runs of full groups, runs crossing its chunks, active-word rows, and
    partial groups between runs; words identical to the oracle's."""
    rng = np.random.default_rng(seed)
    n = 31 * 9000 + 17
    col = rng.integers(0, 6, size=n).astype(np.int32)
    col[: n * 2 // 5] = 1                         # a long run of full groups
    col[150000:152500] = 2                        # a shorter run, not group aligned
    col[200000:260000:2] = 3                      # dense but never full
    col[-40:] = 4                                 # the active word and the last groups
    gi = L.Index.relic(col)
    oi = O.OracleIndex.relic_values(col)
    _same_words(gi, oi)
    fcol = col.astype(np.float32) + 0.25
    bounds = np.array([0.5, 1.5, 2.5, 3.5, 4.5, np.finfo(np.float64).max])
    gb = L.Index.bin(fcol, bounds)
    ob = O.OracleIndex.bin_bounds(fcol, bounds)
    _same_words(gb, ob)
    for (lo, lop, hi, rop) in [(0.0, 2, 2.0, 1), (1.0, 2, 5.0, 2), (3.2, 1, 4.6, 1)]:
        c, r = ob.evaluate(lo, lop, hi, rop)
        h, _, _ = gb.evaluate(lo, lop, hi, rop, 1)
        assert h == c


def test_tiled_copy_beyond_2p32_work_items(L):
    """65,536 bitmaps over 2^26 rows: the tiled copy has 69M segments, more
    than 2^32 work-items at one wave per segment; every tile must be filled
    (the literal plan agrees with a direct count in every tile)."""
    n = 1 << 26
    col = np.random.default_rng(26).integers(0, 65536, size=n).astype(np.uint16)
    ix = L.Index.relic(col)
    cnt = np.bincount(col, minlength=65536)
    cum = np.concatenate([[0], np.cumsum(cnt)])
    lo = np.array([0, 11143, 33960, 51824, 100, 65000], np.float64)
    hi = np.array([7, 11165, 53459, 51825, 60000, 65536], np.float64)
    want = cum[hi.astype(np.int64)] - cum[lo.astype(np.int64)]
    for plan in (1, 3):
        h, _, _ = ix.evaluate_batch(lo, 2, hi, 1, plan=plan)
        np.testing.assert_array_equal(h, want.astype(np.uint64))
    r = ix.rids(11143.0, 2, 11165.0, 1, 1)
    np.testing.assert_array_equal(r, np.nonzero((col >= 11143) & (col < 11165))[0])


def test_estimate_matches_oracle(L):
    """fbh_index_estimate against the oracle's bin::estimate / relic::estimate
    (pinned against the reference): the rows of lower and upper, and a
    cleared upper where the reference clears it."""
    rng = np.random.default_rng(61)

    def rows(words):
        return O.oracle().rids(words)
    seen = set()
    for t in range(12):
        n = int(rng.integers(2000, 60000))
        col = rng.normal(size=n).astype(np.float32) if t % 2 else rng.integers(0, 300, size=n).astype(np.int32)
        if t % 4 == 3:
            gi, oi = L.Index.relic(col), O.OracleIndex.relic_values(col)
        else:
            gi = L.Index.bin_equal_weight(col, int(rng.integers(5, 150)))
            oi = O.OracleIndex.bin_bounds(col, gi.bounds())
        for q in range(8):
            a, b = sorted(rng.choice(col, 2).astype(np.float64))
            if q % 3 == 0:
                a -= 0.25
                b += 0.5
            lop, rop = int(rng.integers(1, 3)), int(rng.integers(1, 3))
            glo, gup = gi.estimate(a, lop, b, rop)
            olo, oup = oi.estimate(a, lop, b, rop)
            assert np.array_equal(glo.rids(), rows(olo)), (t, q, "lower")
            assert glo.size() == n
            if len(oup) == 1:
                assert gup.size() == 0, (t, q)
                seen.add("cleared")
            else:
                assert gup.size() == n and np.array_equal(gup.rids(), rows(oup)), (t, q, "upper")
                seen.add("all" if len(rows(oup)) == n else "upper")
    assert seen == {"cleared", "all", "upper"}


def test_range_batch_matches_evaluate_batch(L):
    """fbh_index_batch_prepare (one persistent graph-captured launch) gives
    the counts of the per-query path on relic and bin indexes of every
    column type, under the heuristic, literal and scan plans, per-query
    plan mixes (the kernel with scan staging), after tuning, and on
    repeated replays."""
    rng = np.random.default_rng(71)
    cols = [rng.integers(-50, 400, size=31 * 3000 + 17).astype(np.int32),
            rng.integers(0, 600, size=31 * 2500 + 3).astype(np.uint32),
            rng.integers(0, 65535, size=31 * 4000 + 29).astype(np.uint16),
            rng.normal(size=31 * 5000 + 11).astype(np.float32),
            (rng.normal(size=31 * 2200 + 5) * 10).astype(np.float64)]
    for ci, col in enumerate(cols):
        if ci < 3:
            ix = L.Index.relic(col)
        elif ci == 3:
            ix = L.Index.bin_equal_weight(col, 200)
        else:                                   # double: explicit bounds
            qb = np.unique(np.quantile(col, np.linspace(0, 1, 150)[1:-1]))
            ix = L.Index.bin(col, np.concatenate([qb, [np.finfo(np.float64).max]]))
        srt = np.sort(col.astype(np.float64))
        a = srt[rng.integers(0, len(col), 96)]
        b = srt[rng.integers(0, len(col), 96)]
        lo, hi = np.minimum(a, b), np.maximum(a, b) + (0.5 if ci < 3 else 0.0)
        want, _, _ = ix.evaluate_batch(lo, 2, hi, 1, plan=1)
        for plan in (0, 1, 3):
            rb = L.RangeBatch(ix, lo, 2, hi, 1, plan=plan)
            h, _ = rb.run()
            assert np.array_equal(h, want), (ci, plan)
        mix = rng.integers(0, 4, size=len(lo)).astype(np.int32)
        rb = L.RangeBatch(ix, lo, 2, hi, 1, plans=mix)
        assert np.array_equal(rb.run()[0], want), (ci, "mix")
        cfg = rb.tune()
        for _ in range(3):
            assert np.array_equal(rb.run()[0], want), (ci, "tuned", cfg)
