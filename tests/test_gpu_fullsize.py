"""The full-size query workloads on the MI355X against committed answers
(tests/golden/make_fullsize.py).

Binned float workload: 2^27 float32 rows, 10,000 equal-weight bins built on
the device (the index is word-identical to the reference's, see
test_gpu_index.py), 1024 quantile ranges lo <= x < hi.  At 2^27 rows the
index has 2114 tiles of 2048 groups, so the literal plan runs the
192-thread form of the literal kernel that the benchmark times.  The 8192
test ranges of the learned plan choice run under the heuristic, each fixed
plan and the learned choice.

Conjunctive workload: 4 columns of 2^26 rows (uint16 relic, three
equal-weight bin indexes), 4096 conjunctions of 2-4 range terms, under the
heuristic, literal and scan plans, after the autotuner, and as AND programs
of the expression kernel.
"""
import hashlib
import os

import numpy as np
import pytest

pytestmark = [pytest.mark.gpu, pytest.mark.slow]
GOLDEN = os.path.join(os.path.dirname(__file__), "golden")


def _say(*a):
    print(*a, flush=True)


@pytest.fixture(scope="module")
def L():
    import libfastlib_amd as L
    L.lib()
    return L


@pytest.fixture(scope="module")
def binf32(L):
    from libfastlib_amd import datagen as G
    g = np.load(os.path.join(GOLDEN, "bin_f32_2p27_queries.npz"))
    col = G.bin_f32_column(27)
    srt = np.sort(col)
    lo, hi = G.quantile_ranges(srt, G.BIN_F32_QUERY_SEED, 1024)
    assert np.array_equal(lo, g["lo"]) and np.array_equal(hi, g["hi"]), "query generator changed"
    del srt
    ix = L.Index.bin_equal_weight(col, 10000)
    nobs, nrows, _ = ix.info()
    assert nrows == 1 << 27
    _say("bin_f32 index:", nobs, "bitmaps")
    yield col, ix, g
    del ix


def test_bin_f32_batch_counts(binf32):
    """All 1024 ranges under the heuristic, literal and scan plans give the
    reference's hit counts."""
    col, ix, g = binf32
    want = g["counts"].astype(np.uint64)
    for plan in (0, 1, 3):
        _say("bin_f32 batch, plan", plan)
        hits, used, _ = ix.evaluate_batch(g["lo"], 2, g["hi"], 1, plan=plan)
        bad = np.nonzero(hits != want)[0]
        assert bad.size == 0, (plan, bad[:8], hits[bad[:8]], want[bad[:8]])
        if plan:
            assert np.all(used == plan)


def test_bin_f32_range_batch(L, binf32):
    """The 1024 ranges as one persistent graph-captured launch
    (fbh_index_batch_prepare) under the heuristic and scan plans, untuned
    and tuned: the reference's counts."""
    col, ix, g = binf32
    want = g["counts"].astype(np.uint64)
    for plan in (0, 3):
        rb = L.RangeBatch(ix, g["lo"], 2, g["hi"], 1, plan=plan)
        h, _ = rb.run()
        assert np.array_equal(h, want), plan
        if plan == 0:
            rb.tune()
            assert np.array_equal(rb.run()[0], want)


def test_bin_f32_rids(binf32):
    """Row ids of every 16th range equal the reference's (SHA-256), through
    the literal plan; the scan plan on a quarter of them."""
    col, ix, g = binf32
    for k, (i, d) in enumerate(zip(g["rid_index"], g["rid_sha256"])):
        plans = (1, 3) if k % 4 == 0 else (1,)
        for plan in plans:
            r = ix.rids(float(g["lo"][i]), 2, float(g["hi"][i]), 1, plan=plan)
            assert len(r) == int(g["counts"][i]), (i, plan)
            assert hashlib.sha256(r.astype(np.uint32).tobytes()).hexdigest() == str(d), (i, plan)


def test_bin_f32_plans_same_words(binf32):
    """The result bitvector of a range is the same words under the literal,
    compressed and scan plans (canonical after compress()), on the
    narrowest ranges, the widest, and ranges whose bitmaps sumBins
    complements."""
    col, ix, g = binf32
    lo, hi, cnt = g["lo"], g["hi"], g["counts"]
    order = np.argsort(cnt)
    pick = list(order[:4]) + list(order[-2:]) + list(order[len(order) // 2: len(order) // 2 + 2])
    for i in pick:
        words = []
        for plan in (1, 2, 3):
            h, used, bv = ix.evaluate(float(lo[i]), 2, float(hi[i]), 1, plan=plan, want_bitvector=True)
            assert h == int(cnt[i]) and used == plan, (i, plan, h)
            words.append(bv.write())
        assert np.array_equal(words[0], words[1]), (i, "compressed")
        assert np.array_equal(words[0], words[2]), (i, "scan")


def test_bin_f32_learned_plans(L, binf32):
    """8192 test ranges: the heuristic, literal, scan and learned plans give
    the same counts as a scan of the column (256 of them checked against
    the reference when the fixture was written).  The planner is trained
    on the device from 2048 training ranges the engine times itself."""
    from libfastlib_amd import datagen as G
    col, ix, g = binf32
    tlo, thi, want = g["test_lo"], g["test_hi"], g["test_counts"].astype(np.uint64)
    for plan in (0, 1, 3):
        _say("bin_f32 test ranges, plan", plan)
        hits, _, _ = ix.evaluate_batch(tlo, 2, thi, 1, plan=plan)
        assert np.array_equal(hits, want), plan
    srt = np.sort(col)
    trlo, trhi = G.quantile_ranges(srt, G.BIN_F32_TRAIN_SEED, 2048)
    del srt
    ms, th = ix.time_plans(trlo, 2, trhi, 1, cap=2.0)
    assert np.array_equal(th[:, 0], th[:, 2])
    done = ms[:, 1] > 0
    assert np.array_equal(th[done, 1], th[done, 0])
    P = L.native.Planner(seed=11)
    y, m = L.native.plan_labels(ms)
    P.train(ix.plan_features(trlo, 2, trhi, 1), y, m, epochs=20, batch=256, lr=2e-3, seed=3)
    learned, _ = ix.choose_plans(P, tlo, 2, thi, 1)
    assert set(np.unique(learned)) <= {1, 2, 3}
    hits, used, _ = ix.evaluate_batch(tlo, 2, thi, 1, plans=learned)
    assert np.array_equal(used, learned)
    assert np.array_equal(hits, want)


@pytest.fixture(scope="module")
def table4(L):
    from libfastlib_amd import datagen as G
    g = np.load(os.path.join(GOLDEN, "conj_4x2p26_counts.npz"))
    cols = G.table4(G.CONJ_TABLE_SEED, 1 << 26)
    qs = G.conj_queries(G.CONJ_QUERY_SEED, cols, 4096)
    flat = np.array([(ci, a, b) for q in qs for (ci, a, b) in q], np.float64)
    assert np.array_equal(flat, g["terms"]), "conjunction generator changed"
    _say("conj: building 4 indexes")
    ixs = [L.Index.relic(cols[0]), L.Index.bin_equal_weight(cols[1], 10000),
           L.Index.bin_equal_weight(cols[2], 10000), L.Index.bin_equal_weight(cols[3], 10000)]
    yield ixs, qs, g
    del ixs


def test_conj_fullsize_counts(L, table4):
    """4096 conjunctions over 4 x 2^26 rows: counts under the heuristic,
    literal and scan plans and after autotuning equal the committed
    answers (48 of them checked against the reference when written)."""
    ixs, qs, g = table4
    want = g["counts"].astype(np.uint64)
    terms = [[(ixs[ci], a, 2, b, 1) for (ci, a, b) in q] for q in qs]
    for plan in (0, 1, 3):
        _say("conj batch, plan", plan)
        batch = L.Conjunctions(terms, plan=plan)
        hits, _ = batch.run()
        bad = np.nonzero(hits != want)[0]
        assert bad.size == 0, (plan, bad[:8])
        if plan == 0:
            cfg = batch.tune()
            hits, _ = batch.run()
            assert np.array_equal(hits, want), ("tuned", cfg)
        del batch


def test_conj_fullsize_and_programs(L, table4):
    """The same conjunctions as AND programs through the expression kernel
    (short-circuit on empty tiles, cheapest term first)."""
    ixs, qs, g = table4
    want = g["counts"].astype(np.uint64)
    progs = [([(ixs[ci], a, 2, b, 1) for (ci, a, b) in q], [0] + [x for k in range(1, len(q)) for x in (k, -1)])
             for q in qs]
    batch = L.Expressions(progs)
    hits, _ = batch.run()
    assert np.array_equal(hits, want)


def test_conj_fullsize_rids(L, table4):
    """Row ids of all 4096 conjunctions from one more persistent launch:
    per-query lengths equal the committed counts, ids ascend within each
    query, and every 256th query's ids equal a scan of the columns."""
    from libfastlib_amd import datagen as G
    ixs, qs, g = table4
    want = g["counts"].astype(np.uint64)
    batch = L.Conjunctions([[(ixs[ci], a, 2, b, 1) for (ci, a, b) in q] for q in qs])
    _say("conj rids")
    off, rids = batch.rids()
    assert np.array_equal(np.diff(off), want)
    d = np.diff(rids.astype(np.int64))
    starts = off[1:-1][off[1:-1] < len(rids)].astype(np.int64)
    ok = d > 0
    ok[starts[starts > 0] - 1] = True          # a new query may restart lower
    assert ok.all()
    cols = G.table4(G.CONJ_TABLE_SEED, 1 << 26)
    for i in range(0, len(qs), 256):
        m = None
        for ci, a, b in qs[i]:
            c = cols[ci]
            v = c.astype(np.int64) if c.dtype.kind in "iu" else c.astype(np.float64)
            t = (v >= (int(a) if c.dtype.kind in "iu" else a)) & (v < (int(b) if c.dtype.kind in "iu" else b))
            m = t if m is None else (m & t)
        assert np.array_equal(rids[off[i]:off[i + 1]], np.nonzero(m)[0].astype(np.uint32)), i
