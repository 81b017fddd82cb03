"""Indexes that arrive built: fbh_index_read_file on index files the
reference wrote (tests/golden/idx), fbh_index_attach on the deconstructed
form (fastbit_iapi_attach_full_index), and fbh_index_sum_bins
(fastbit_iapi_get_number_of_hits) -- words, locate, hit counts and row ids
against the reference's own reading of the same files."""
import os
import shutil

import numpy as np
import pytest

import oracle as O

pytestmark = pytest.mark.gpu
IDX = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "idx")
NAMES = ["relic_i32", "bin_f32", "relic_card"]


@pytest.fixture(scope="module")
def L():
    import libfastlib_amd as L
    L.lib()
    return L


def fixture(name):
    with np.load(os.path.join(IDX, name + ".npz")) as z:
        return {k: z[k] for k in z.files}


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
    return np.flatnonzero(m).astype(np.uint32)


def plans_for(name):
    # relic_card leaves rows with a value >= card in no bitmap: a scan of the
    # raw values would select them, the index (the reference) does not
    return (1, 2) if name == "relic_card" else (0, 1, 2, 3)


def check_index(ix, f, name, locate=True):
    k, o, b = ix.deconstruct()
    assert np.array_equal(k, f["keys"])
    assert np.array_equal(o, f["offsets"])
    assert np.array_equal(b, f["bitmaps"])
    nobs, nrows, nw = ix.info()
    assert (nobs, nrows, nw) == (f["offsets"].size - 1, f["col"].size, f["bitmaps"].size)
    for i, (lo, lop, hi, rop) in enumerate(zip(f["lo"], f["lop"], f["hi"], f["rop"])):
        args = (float(lo), int(lop), float(hi), int(rop))
        if locate:
            assert ix.resolve_range(*args) == tuple(int(x) for x in f["locate"][i]), (name, args)
        for plan in plans_for(name):
            assert ix.evaluate(*args, plan)[0] == f["counts"][i], (name, args, plan)
        r = ix.rids(*args, 1)
        assert r.size == f["counts"][i]
        if name != "relic_card":
            assert np.array_equal(r, brute(f["col"], *args)), (name, args)
    for (a, e), c in zip(f["pairs"], f["sum_bins"]):
        assert ix.sum_bins(int(a), int(e)) == c, (name, a, e)


@pytest.mark.parametrize("w8", [False, True])
@pytest.mark.parametrize("name", NAMES)
def test_read_reference_file(L, name, w8):
    f = fixture(name)
    print(f"read_file {name} w8={w8}", flush=True)
    ix = L.Index.read_file(os.path.join(IDX, name + ("_w8" if w8 else "") + ".idx"), f["col"])
    check_index(ix, f, name)


@pytest.mark.parametrize("name", NAMES)
def test_attach_reference_triple(L, name):
    f = fixture(name)
    print(f"attach {name}", flush=True)
    ix = L.Index.attach(f["col"], f["keys"], f["offsets"], f["bitmaps"])
    # an attached bin rebuilds its boundaries from the minima and maxima, so
    # its locate differs from the file's; counts and rows do not
    check_index(ix, f, name, locate=str(f["kind"]) == "relic")
    if str(f["kind"]) == "bin":
        assert np.array_equal(ix.bounds(), O.bounds_from_keys(f["keys"]))


def _roundtrip(gi, gj, queries):
    a, b = gi.deconstruct(), gj.deconstruct()
    for x, y in zip(a, b):
        assert np.array_equal(x, y)
    for q in queries:
        for plan in (0, 1, 2, 3):
            assert gi.evaluate(*q, plan)[0] == gj.evaluate(*q, plan)[0], (q, plan)


def test_attach_roundtrip_relic_2p20(L):
    """relic over 2^20 int32 in [0, 1024): built, deconstructed, attached."""
    from libfastlib_amd import datagen as G
    col = G.uniform_i32(1, 1 << 20, 0, 1024)
    print("roundtrip relic 2^20", flush=True)
    gi = L.Index.relic(col)
    gj = L.Index.attach(col, *gi.deconstruct())
    rng = np.random.default_rng(5)
    qs = [(float(lo), 2, float(lo + w), 1) for lo, w in zip(rng.integers(-5, 1024, 24), rng.integers(1, 400, 24))]
    _roundtrip(gi, gj, qs)
    for lo, lop, hi, rop in qs[:6]:
        assert gj.evaluate(lo, lop, hi, rop, 1)[0] == brute(col, lo, lop, hi, rop).size


def test_attach_roundtrip_bin_2p22(L):
    """bin with 1000 equal-weight bins over 2^22 float32 normal(0, 1)."""
    from libfastlib_amd import datagen as G
    col = G.normal_f32(7, 1 << 22)
    print("roundtrip bin 2^22", flush=True)
    gi = L.Index.bin_equal_weight(col, 1000)
    gj = L.Index.attach(col, *gi.deconstruct())
    a, b = gi.deconstruct(), gj.deconstruct()
    for x, y in zip(a, b):
        assert np.array_equal(x, y)
    rng = np.random.default_rng(6)
    for _ in range(16):
        lo, hi = sorted(rng.normal(size=2) * 1.5)
        q = (float(lo), 2, float(hi), 1)
        want = brute(col, *q).size
        for plan in (0, 1, 2, 3):
            assert gj.evaluate(*q, plan)[0] == want, (q, plan)


def test_read_reference_file_2p20(L, tmp_path):
    """Index files of 2^20 rows written by the reference at run time."""
    if O.reference() is None:
        pytest.skip("reference library (oracle/_ref) not built here")
    from libfastlib_amd import datagen as G
    col = G.uniform_i32(11, 1 << 20, -500, 500)
    R = O.RefIapi("att_relic_2p20", col, "<binning none/><encoding equality/>")
    path = str(tmp_path / "relic.idx")
    O.ref_write_index_file("att_relic_2p20", path)
    print("read_file relic 2^20", flush=True)
    ix = L.Index.read_file(path, col)
    k, o, b = ix.deconstruct()
    assert np.array_equal(k, R.keys) and np.array_equal(o, R.offsets) and np.array_equal(b, R.bitmaps)
    rng = np.random.default_rng(12)
    for lo, w in zip(rng.integers(-520, 520, 16), rng.integers(1, 300, 16)):
        q = (float(lo), 2, float(lo + w), 1)
        c = R.query_range(*q, want_rids=False)[0]
        assert ix.evaluate(*q, 1)[0] == c
        assert ix.evaluate(*q, 2)[0] == c
        h = O.ref_file_locate(path, *q)
        assert ix.resolve_range(*q) == h
    fcol = G.normal_f32(13, 1 << 20)
    R = O.RefIapi("att_bin_2p20", fcol, "<binning nbins=1000 equal-weight/>")
    path = str(tmp_path / "bin.idx")
    O.ref_write_index_file("att_bin_2p20", path)
    print("read_file bin 2^20", flush=True)
    ix = L.Index.read_file(path, fcol)
    for _ in range(16):
        lo, hi = sorted(rng.normal(size=2) * 1.5)
        q = (float(lo), 2, float(hi), 1)
        assert ix.resolve_range(*q) == O.ref_file_locate(path, *q)
        assert ix.evaluate(*q, 1)[0] == R.query_range(*q, want_rids=False)[0]


def _expect_error(L, fn):
    with pytest.raises(L.FastbitError):
        fn()


def test_read_file_rejects_damaged(L, tmp_path):
    """Files that do not describe the column are errors found on the host,
    before any kernel reads the words."""
    f = fixture("relic_i32")
    src = os.path.join(IDX, "relic_i32.idx")
    raw = open(src, "rb").read()
    col = f["col"]

    def write(name, data):
        p = str(tmp_path / name)
        with open(p, "wb") as fh:
            fh.write(data)
        return p

    print("damaged files", flush=True)
    _expect_error(L, lambda: L.Index.read_file(str(tmp_path / "missing.idx"), col))
    _expect_error(L, lambda: L.Index.read_file(write("magic.idx", b"#IBIX" + raw[5:]), col))
    _expect_error(L, lambda: L.Index.read_file(write("short.idx", raw[: len(raw) // 2]), col))
    _expect_error(L, lambda: L.Index.read_file(write("type.idx", raw[:5] + bytes([1]) + raw[6:]), col))
    _expect_error(L, lambda: L.Index.read_file(src, col[:-1]))
    # one fill word of the first bitmap one group longer
    bms = f["bitmaps"].copy()
    fills = np.flatnonzero((bms[: f["offsets"][1]] >> 31) == 1)
    assert fills.size
    bms[fills[0]] += 1
    _expect_error(L, lambda: L.Index.attach(col, f["keys"], f["offsets"], bms))
    pos0 = int(np.frombuffer(raw, np.int32, 1, 24 + 8 * 300)[0])
    bad = bytearray(raw)
    bad[pos0 + 4 * int(fills[0]): pos0 + 4 * int(fills[0]) + 4] = np.uint32(bms[fills[0]]).tobytes()
    _expect_error(L, lambda: L.Index.read_file(write("fill.idx", bytes(bad)), col))
    # offsets beyond the words, descending offsets, mismatching key count
    o = f["offsets"].copy()
    o[-1] += 1
    _expect_error(L, lambda: L.Index.attach(col, f["keys"], o, f["bitmaps"]))
    o = f["offsets"].copy()
    o[1], o[2] = o[2], o[1]
    _expect_error(L, lambda: L.Index.attach(col, f["keys"], o, f["bitmaps"]))
    _expect_error(L, lambda: L.Index.attach(col, f["keys"][:-1], f["offsets"], f["bitmaps"]))
    # the damaged copies leave the good file readable
    shutil.copy(src, str(tmp_path / "good.idx"))
    assert L.Index.read_file(str(tmp_path / "good.idx"), col).info()[0] == 300
