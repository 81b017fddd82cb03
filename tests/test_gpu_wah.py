"""WAH bitvector operators on the MI355X, bit-exact against the oracle and
the reference's golden words."""
import hashlib
import json
import os

import numpy as np
import pytest

from helpers import markov_bits, random_pair, split_serial

pytestmark = pytest.mark.gpu
GOLDEN = os.path.join(os.path.dirname(__file__), "golden")


@pytest.fixture(scope="module")
def L():
    import libfastlib_amd as L
    L.lib()
    return L


def gpu_op(L, a, b, op):
    x = L.Bitvector.from_serial(a)
    y = L.Bitvector.from_serial(b)
    path = x.op(y, op)
    return x.write(), path


def test_random_ops_match_oracle(L, orc):
    rng = np.random.default_rng(31)
    seen = set()
    for t in range(120):
        a, b, n = random_pair(rng, orc)
        if t % 4 == 0:
            a = orc.op(a, orc.from_bits(markov_bits(n, 0.1, 7, rng)), 1, compress=False)[0]
        for op in range(4):
            want, wpath = orc.op(a, b, op)
            got, gpath = gpu_op(L, a, b, op)
            assert gpath == wpath, (t, op)
            assert np.array_equal(got, want), (t, op, gpath, len(got), len(want))
            seen.add(gpath)
    assert {"c2", "d2", "c0"} <= seen


def test_literal_operand_paths(L, orc):
    rng = np.random.default_rng(32)
    for _ in range(30):
        n = int(rng.integers(62, 5000))
        a = orc.decompress(orc.from_bits(markov_bits(n, 0.5, 1.5, rng)))
        b = orc.from_bits(markov_bits(n, 0.01, 3, rng))
        for x, y in ((a, b), (b, a), (a, a)):
            for op in range(4):
                want, wpath = orc.op(x, y, op)
                got, gpath = gpu_op(L, x, y, op)
                assert gpath == wpath
                assert np.array_equal(got, want), (op, gpath)


def test_golden_ops(L):
    g = np.load(os.path.join(GOLDEN, "wah_ops.npz"))
    for i in range(int(g["ncase"])):
        for op in range(4):
            got, _ = gpu_op(L, g[f"a{i}"], g[f"b{i}"], op)
            assert np.array_equal(got, g[f"r{i}_{op}"]), (i, op)


def test_edge_cases(L, orc):
    cases = []
    for n in (1, 30, 31, 32, 61, 62, 63, 93, 124):
        for bits in (np.zeros(n, np.uint8), np.ones(n, np.uint8),
                     (np.arange(n) % 2).astype(np.uint8), (np.arange(n) % 37 == 0).astype(np.uint8)):
            cases.append(orc.from_bits(bits))
    for i, a in enumerate(cases):
        for b in cases[i - (i % 4): i - (i % 4) + 4]:
            for op in range(4):
                want, wp = orc.op(a, b, op)
                got, gp = gpu_op(L, a, b, op)
                assert gp == wp and np.array_equal(got, want), (i, op, wp)


def test_unequal_sizes_extend(L, orc):
    rng = np.random.default_rng(33)
    for _ in range(10):
        a = orc.from_bits(markov_bits(int(rng.integers(40, 900)), 0.3, 5, rng))
        b = orc.from_bits(markov_bits(int(rng.integers(900, 2000)), 0.3, 5, rng))
        for x, y in ((a, b), (b, a)):
            for op in range(4):
                want, _ = orc.op(x, y, op)
                got, _ = gpu_op(L, x, y, op)
                assert np.array_equal(got, want)


def test_unary_ops(L, orc):
    rng = np.random.default_rng(34)
    for t in range(40):
        a, b, n = random_pair(rng, orc)
        u = orc.op(a, b, int(rng.integers(0, 4)), compress=False)[0]   # non-canonical words
        x = L.Bitvector.from_serial(u)
        c, s, after = orc.cnt(u)
        assert x.size() == s and x.cnt() == c
        assert np.array_equal(x.write(), after)
        x = L.Bitvector.from_serial(u)
        x.compress()
        assert np.array_equal(x.write(), orc.compress(u))
        x = L.Bitvector.from_serial(u)
        x.decompress()
        assert np.array_equal(x.write(), orc.decompress(u))
        x = L.Bitvector.from_serial(u)
        x.size()          # the oracle's flip entry counts first (do_cnt may collapse)
        x.flip()
        assert np.array_equal(x.write(), orc.flip(u))
        x = L.Bitvector.from_serial(u)
        assert np.array_equal(x.rids(), orc.rids(u))


def test_set_and_adjust(L, orc):
    for val in (0, 1):
        for n in (0, 5, 31, 62, 100, 1000):
            x = L.Bitvector()
            x.set(val, n)
            bits = np.full(n, val, np.uint8)
            assert np.array_equal(x.write(), orc.from_bits(bits))
            x.adjustSize(n + 40, n + 100)
            bits2 = np.concatenate([bits, np.ones(40, np.uint8), np.zeros(60, np.uint8)])
            assert np.array_equal(x.write(), orc.from_bits(bits2))


def _digest_rows(maxbits):
    with open(os.path.join(GOLDEN, "wah_config2_digests.json")) as f:
        return [r for r in json.load(f)["rows"] if r["nbits"] <= maxbits]


def _check_rows(L, rows):
    from libfastlib_amd import datagen as G
    cache = {}
    for r in rows:
        key = (r["seed_a"], r["seed_b"])
        if key not in cache:
            cache.clear()
            a = G.markov_wah(r["seed_a"], r["nbits"], r["density"], r["mean_run"])
            b = G.markov_wah(r["seed_b"], r["nbits"], r["density"], r["mean_run"])
            assert [len(a), len(b)] == r["in_words"]
            cache[key] = (L.Bitvector.from_serial(a), L.Bitvector.from_serial(b))
        x, y = cache[key]
        z = x.copy()
        path = z.op(y, r["op"])
        out = z.write()
        assert path == r["path"], r
        assert len(out) == r["words"], r
        assert hashlib.sha256(out.tobytes()).hexdigest() == r["sha256"], r
        assert z.cnt() == r["cnt"], r


def test_config2_digests_2p24(L):
    _check_rows(L, _digest_rows(1 << 24))


def test_config2_digests_2p28(L):
    """All nine Markov cells at 2^28 bits, four operators each: identical
    words to the reference after compress()."""
    _check_rows(L, [r for r in _digest_rows(1 << 28) if r["nbits"] == (1 << 28)])


def test_canonical_result_spanning_tiles(L, orc):
    """A canonical (c0) result that is one run across many kernel tiles:
    middle tiles own no output word (regression: their store count was
    computed as 0 - 1)."""
    n = 31 * 40000 + 7
    ones = np.ones(n, np.uint8)
    a = orc.decompress(orc.from_bits(ones))                 # literal-only all ones
    alt = orc.decompress(orc.from_bits((np.arange(n) % 2).astype(np.uint8)))
    for x, y in ((a, a), (a, alt), (alt, alt)):
        for op in range(4):
            want, wpath = orc.op(x, y, op)
            got, gpath = gpu_op(L, x, y, op)
            assert gpath == wpath == "c0"
            assert np.array_equal(got, want), op


def test_c0_runs_across_wave_spans(L, orc):
    """Literal operands of ~2.1M groups whose op results hold fill runs of up
    to 3000 groups: runs cross the rounds and the per-wave spans the literal
    kernels split the groups into; group counts around multiples of 4 and
    64 exercise partial chunks and rounds."""
    rng = np.random.default_rng(41)
    for groups in ((1 << 21) + 1, (1 << 21) + 63, 1 << 21):
        lens = rng.integers(1, 3000, size=groups // 500 + 8)
        cls = np.repeat(rng.integers(0, 3, size=lens.size), lens)[:groups]
        bits = np.where(np.repeat(cls, 31) == 2, rng.integers(0, 2, size=groups * 31),
                        np.repeat(cls, 31)).astype(np.uint8)
        a = orc.decompress(orc.from_bits(bits))
        b = orc.decompress(orc.from_bits(np.where(rng.random(bits.size) < 0.002, 0, 1).astype(np.uint8)))
        for op in range(4):
            want, wpath = orc.op(a, b, op)
            got, gpath = gpu_op(L, a, b, op)
            assert gpath == wpath == "c0", (groups, op)
            assert np.array_equal(got, want), (groups, op, len(got), len(want))


def test_c0_repeated_ops_tag_cycle(L, orc):
    """300 literal ops in a row: the one-launch literal kernel tags its block
    totals per launch (1..255) and clears the table when the tags wrap, so
    every call sees only its own totals."""
    rng = np.random.default_rng(43)
    n = 31 * 50000 + 11
    a = orc.decompress(orc.from_bits(markov_bits(n, 0.5, 1.5, rng)))
    b = orc.decompress(orc.from_bits(markov_bits(n, 0.3, 40, rng)))
    want = {op: orc.op(a, b, op)[0] for op in range(4)}
    x = L.Bitvector.from_serial(a)
    y = L.Bitvector.from_serial(b)
    for i in range(300):
        z = x.copy()
        assert z.op(y, i % 4) == "c0"
        if i % 37 == 0 or i >= 250:
            assert np.array_equal(z.write(), want[i % 4]), i


def test_c0_beyond_one_batch(L, orc):
    """11M groups: more than the one-launch kernel holds in registers, so
    the two-launch literal kernels (count, then write) take it."""
    rng = np.random.default_rng(44)
    groups = 11_000_000
    wa = rng.integers(0, 1 << 31, size=groups, dtype=np.uint32) & np.uint32(0x7FFFFFFF)
    wb = rng.integers(0, 1 << 31, size=groups, dtype=np.uint32) & np.uint32(0x7FFFFFFF)
    wa[1000:90000] = 0                       # runs for the fill words
    wb[5_000_000:5_400_000] = 0x7FFFFFFF
    wa[5_200_000:5_300_000] = 0x7FFFFFFF
    a = np.concatenate([wa, np.array([0], np.uint32)])        # words, then 0 active bits
    b = np.concatenate([wb, np.array([0], np.uint32)])
    for op in (0, 1, 3):
        want, wpath = orc.op(a, b, op)
        got, gpath = gpu_op(L, a, b, op)
        assert gpath == wpath == "c0"
        assert np.array_equal(got, want), op


def test_d2_beyond_one_batch(L, orc):
    """Compressed operands of 11.3M groups (more rounds than the one-launch
    kernel holds per resident wave): the tile pass takes them; its words
    equal the oracle's for every op."""
    from libfastlib_amd import datagen as G
    nbits = 31 * 11_300_000
    a = G.markov_wah(7, nbits, 0.5, 64)
    b = G.markov_wah(8, nbits, 0.5, 64)
    for op in range(4):
        want, wpath = orc.op(a, b, op)
        got, gpath = gpu_op(L, a, b, op)
        assert gpath == wpath == "d2", (op, gpath, wpath)
        assert np.array_equal(got, want), op


def test_d2_mixed_kernel_shapes(L, orc):
    """d1 / d2 ops at group counts around round (256), chunk (32 words) and
    segment boundaries, operands from runs of every kind."""
    from libfastlib_amd import datagen as G
    for k, groups in enumerate((255, 256, 257, 4096 + 3, 256 * 64 * 4 + 1, 700_001)):
        nbits = 31 * groups + (k % 3) * 7
        a = G.markov_wah(100 + k, nbits, 0.5, 48)
        b = G.markov_wah(200 + k, nbits, 0.4, 20)
        lit = orc.decompress(b)
        for x, y in ((a, b), (a, lit), (lit, a)):
            for op in range(4):
                want, wpath = orc.op(x, y, op)
                got, gpath = gpu_op(L, x, y, op)
                assert gpath == wpath, (groups, op)
                assert np.array_equal(got, want), (groups, op, gpath)


def test_ops_from_several_host_threads(L, orc):
    """Four host threads run operators on handles of their own at the same
    time (ctypes releases the interpreter lock during each call); the
    library's lock keeps their use of the shared stream and scratch apart,
    so every result matches the oracle."""
    import threading
    rng = np.random.default_rng(77)
    work = []
    for t in range(4):
        cases = []
        for _ in range(12):
            a, b, n = random_pair(rng, orc, 2000, 40000)
            op = int(rng.integers(0, 4))
            cases.append((a, b, op, orc.op(a, b, op)[0]))
        work.append(cases)
    bad = []

    def run(cases, tag):
        try:
            for rep in range(3):
                for i, (a, b, op, want) in enumerate(cases):
                    got, _ = gpu_op(L, a, b, op)
                    if not np.array_equal(got, want):
                        bad.append((tag, rep, i, op))
        except Exception as e:          # surfaced below, in the test's thread
            bad.append((tag, repr(e)))
    ths = [threading.Thread(target=run, args=(c, k)) for k, c in enumerate(work)]
    for th in ths:
        th.start()
    for th in ths:
        th.join()
    assert not bad, bad[:5]


def _forms(orc, a, b):
    """a canonical, decompressed, and as uncompressed c2 outputs (count-1
    fills, raw-pushed words)."""
    return [a, orc.decompress(a), orc.op(a, b, 0, compress=False)[0], orc.op(a, b, 1, compress=False)[0]]


def test_out_of_place_operators_match_oracle(L, orc):
    """a & b, a | b, a ^ b, a - b through fbh_bv_op_new: the words of the
    reference's out-of-place dispatch (c0 / c1 / d2 with the strict size
    test / c2 loop words / shortcut copies, nothing compressed; unequal
    sizes through the extended copy and the in-place operator), word for
    word against the oracle, which is pinned against the reference."""
    rng = np.random.default_rng(41)
    paths = set()
    for t in range(150):
        a, b, n = random_pair(rng, orc, 1, 6000)
        if t % 5 == 0 and n > 40:
            b = orc.from_bits(markov_bits(n + int(rng.integers(-40, 40)), 0.2, 5, rng))
        x = _forms(orc, a, b)[int(rng.integers(0, 4))]
        y = _forms(orc, b, a)[int(rng.integers(0, 4))]
        X, Y = L.Bitvector.from_serial(x), L.Bitvector.from_serial(y)
        for op in range(4):
            want, wpath = orc.op_new(x, y, op)
            r, gpath = X.op_new(Y, op)
            assert gpath == wpath, (t, op, gpath, wpath)
            assert np.array_equal(r.write(), want), (t, op, gpath)
            paths.add(gpath)
        # the operands are not modified
        assert np.array_equal(X.write(), x) or orc.cnt(x)[2].tolist() == X.write().tolist()
    assert {"c0", "c1", "c2", "d2", "shortcut"} <= paths


def test_out_of_place_size_boundary(L, orc):
    """Pairs with (|a|w + |b|w) * 31 == nbits exactly: the out-of-place
    operator takes the c2 loop where the in-place one takes d2."""
    rng = np.random.default_rng(42)
    hit = 0
    for t in range(4000):
        g = int(rng.integers(4, 60))
        n = 31 * g
        a = orc.from_bits(markov_bits(n, 0.3, 40, rng))
        b = orc.from_bits(markov_bits(n, 0.3, 40, rng))
        na, nb = len(a) - 1, len(b) - 1          # no active word: n is whole groups
        if (na + nb) != g:
            continue
        hit += 1
        X, Y = L.Bitvector.from_serial(a), L.Bitvector.from_serial(b)
        for op in range(4):
            want, wpath = orc.op_new(a, b, op)
            r, gpath = X.op_new(Y, op)
            assert gpath == wpath and np.array_equal(r.write(), want), (t, op, gpath)
            _, ipath = orc.op(a, b, op)
            assert ipath in ("d2", "shortcut"), ipath
        if hit >= 40:
            break
    assert hit >= 10


def test_count_mask_and_erase_match_oracle(L, orc):
    """count(mask) against the oracle (pinned against the reference and the
    popcount of the AND), every operand form; erase(i, j) word for word,
    cuts inside literals, inside fills and in the active word."""
    rng = np.random.default_rng(43)
    for t in range(150):
        a, b, n = random_pair(rng, orc, 1, 8000)
        x = _forms(orc, a, b)[t % 4]
        y = _forms(orc, b, a)[(t // 4) % 4]
        X, Y = L.Bitvector.from_serial(x), L.Bitvector.from_serial(y)
        assert X.count(Y) == orc.count(x, y), t
        sz = orc.cnt(x)[1]
        for _ in range(3):
            i = int(rng.integers(0, sz))
            j = int(rng.integers(i, sz + 40)) if t % 3 else int(rng.integers(i, min(sz, i + 64) + 1))
            Z = L.Bitvector.from_serial(x)
            Z.erase(i, j)
            assert np.array_equal(Z.write(), orc.erase(x, i, j)), (t, i, j, sz)
