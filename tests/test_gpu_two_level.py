"""Two-level indexes on the GPU (fbh_index_build_fuzz / _zona and
fbh_index_read_file on fuzz / zona files; src/irelic.h:540-712): coarse
bins and bitmaps, the option and the coarse bitmaps each range reads, the
cost, and the answers (counts, row ids, hit vectors, batches) against the
reference's files and answers (tests/golden/idx/{fuzz,zona}_*) and against
the oracle on larger columns of several tiles."""
from __future__ import annotations

import os

import numpy as np
import pytest

import oracle

pytestmark = pytest.mark.gpu
IDX = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "idx")
NAMES = ["fuzz_i32", "fuzz_i32_nc8", "zona_i32", "zona_f32"]


@pytest.fixture(scope="module")
def L():
    import libfastlib_amd as L
    L.lib()
    return L


def fixture(name):
    with np.load(os.path.join(IDX, name + ".npz")) as z:
        return {k: z[k] for k in z.files}


def terms(f):
    return [(float(f["lo"][i]), int(f["lop"][i]), float(f["hi"][i]), int(f["rop"][i])) for i in range(len(f["lo"]))]


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


def canonical(words):
    o = oracle.oracle()
    return o.compress(o.decompress(words))


def build(L, name, col):
    return L.Index.fuzz(col, 8 if name.endswith("nc8") else 0) if name.startswith("fuzz") else L.Index.zona(col)


@pytest.mark.parametrize("name", NAMES)
def test_built_index_matches_reference(L, name):
    f = fixture(name)
    col = f["col"]
    print("two-level build", name, col.size, flush=True)
    ix = build(L, name, col)
    enc, cb, coff, nb = ix.coarse()
    assert enc == (L.ENC_FUZZ if name.startswith("fuzz") else L.ENC_ZONA)
    assert np.array_equal(cb, f["cbounds"])
    assert np.array_equal(np.diff(coff), np.diff(f["coffsets"]) + 32)
    woff = f["cbits_off"]
    for i in range(nb):
        got = ix.coarse_bitmap(i).write()
        assert np.array_equal(got, canonical(f["cbits"][woff[i]:woff[i + 1]])), (name, i)
    fine = oracle.OracleIndex.relic_values(col)
    tl = oracle.OracleTwoLevel.build(fine, oracle.ENC_FUZZ if name.startswith("fuzz") else oracle.ENC_ZONA,
                                     8 if name.endswith("nc8") else 0)
    for i, t in enumerate(terms(f)):
        ic, sc, plan = ix.estimate_cost(*t)
        assert ic == f["cost_mem"][i], (name, i, t)
        opt, loads = ix.two_level_option(*t)
        _, _, oopt, oloads = tl.evaluate(*t)
        assert (opt, loads) == (oopt, oloads), (name, i, t)
        hits, used, _ = ix.evaluate(*t, plan=L.PLAN_LITERAL)
        assert hits == f["counts"][i], (name, i, t, opt)


@pytest.mark.parametrize("w8", [False, True])
@pytest.mark.parametrize("name", NAMES)
def test_read_file_matches_reference(L, name, w8, tmp_path):
    f = fixture(name)
    col = f["col"]
    path = os.path.join(IDX, name + ".idx")
    if w8:
        oracle.two_level_widen(path, str(tmp_path / "w8.idx"))
        path = str(tmp_path / "w8.idx")
    print("two-level file", name, w8, flush=True)
    ix = L.Index.read_file(path, col)
    enc, cb, coff, nb = ix.coarse()
    _, _, nc, fcb, fco, words, woff = oracle.two_level_file_parts(path)
    assert np.array_equal(cb, fcb) and np.array_equal(coff, fco)
    for i in range(nb):
        assert np.array_equal(ix.coarse_bitmap(i).write(), words[woff[i]:woff[i + 1]]), (name, i)
    for i, t in enumerate(terms(f)):
        ic, _, _ = ix.estimate_cost(*t)
        assert ic == f["cost_file"][i], (name, i, t)
        opt, loads = ix.two_level_option(*t)
        assert loads == int(f["loaded"][i]), (name, i, t, opt)
        hits, _, _ = ix.evaluate(*t, plan=L.PLAN_LITERAL)
        assert hits == f["counts"][i], (name, i, t, opt)


@pytest.mark.parametrize("name", ["fuzz_i32", "zona_f32"])
def test_rids_hit_vectors_and_plans(L, name):
    f = fixture(name)
    col = f["col"]
    ix = build(L, name, col)
    o = oracle.oracle()
    for i, t in enumerate(terms(f)):
        m = brute(col, *t)
        want = np.nonzero(m)[0].astype(np.uint32)
        if i % 3 == 0:
            assert np.array_equal(ix.rids(*t, plan=L.PLAN_LITERAL), want), (name, i, t)
        if i % 5 == 0:
            hits, _, bv = ix.evaluate(*t, plan=L.PLAN_LITERAL, want_bitvector=True)
            assert hits == want.size
            assert np.array_equal(o.rids(bv.write()), want), (name, i, t)
        for plan in (L.PLAN_AUTO, L.PLAN_SCAN, L.PLAN_COMPRESSED):
            hits, _, _ = ix.evaluate(*t, plan=plan)
            assert hits == want.size, (name, i, t, plan)


@pytest.mark.parametrize("name", ["fuzz_i32", "zona_i32"])
def test_batches(L, name):
    f = fixture(name)
    col = f["col"]
    ix = build(L, name, col)
    T = terms(f)
    lo = np.array([t[0] for t in T]); lop = np.array([t[1] for t in T], np.int32)
    hi = np.array([t[2] for t in T]); rop = np.array([t[3] for t in T], np.int32)
    hits, used, _ = ix.evaluate_batch(lo, lop, hi, rop, plan=L.PLAN_AUTO)
    assert np.array_equal(hits, f["counts"])
    assert used.size == len(T)
    b = L.RangeBatch(ix, lo, lop, hi, rop, plan=L.PLAN_LITERAL)
    assert np.array_equal(b.run()[0], f["counts"])
    b.tune()
    assert np.array_equal(b.run()[0], f["counts"])
    offs, rids = b.rids()
    for i in range(0, len(T), 7):
        want = np.nonzero(brute(col, *T[i]))[0].astype(np.uint32)
        assert np.array_equal(rids[offs[i]:offs[i + 1]], want), (name, i)


@pytest.mark.parametrize("enc", ["fuzz", "zona"])
def test_several_tiles_against_oracle(L, enc):
    """300,001 rows (5 tiles of 2048 groups) over 3,000 values, clustered
    and uniform halves, 64 ranges: options, coarse reads, counts and row ids
    as the oracle's."""
    rng = np.random.default_rng(77)
    n = 300001
    col = np.concatenate([np.sort(rng.integers(0, 3000, n // 2)), rng.integers(0, 3000, n - n // 2)]).astype(np.int32)
    print("two-level tiles", enc, n, flush=True)
    ix = L.Index.fuzz(col) if enc == "fuzz" else L.Index.zona(col)
    fine = oracle.OracleIndex.relic_values(col)
    tl = oracle.OracleTwoLevel.build(fine, oracle.ENC_FUZZ if enc == "fuzz" else oracle.ENC_ZONA)
    ocb, ocoff, _ = tl.info()
    _, cb, coff, nb = ix.coarse()
    assert np.array_equal(cb, ocb) and np.array_equal(coff, ocoff)
    opts = set()
    for q in range(64):
        lo = float(rng.integers(-50, 3000))
        hi = lo + float(rng.choice([3, 40, 400, 1500, 2900]))
        t = (lo, int(rng.integers(1, 3)), hi, int(rng.integers(1, 3)))
        cnt, rids, oopt, oloads = tl.evaluate(*t)
        opt, loads = ix.two_level_option(*t)
        assert (opt, loads) == (oopt, oloads), (q, t)
        opts.add(opt)
        got = ix.rids(*t, plan=L.PLAN_LITERAL)
        assert got.size == cnt and np.array_equal(got, rids), (q, t, opt)
    assert len(opts & {2, 3, 4, 5}) >= 2, opts
