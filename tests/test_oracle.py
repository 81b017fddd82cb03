"""The CPU oracle pinned against the compiled reference, the committed
golden fixtures and closed forms (CPU only)."""
import json
import os

import numpy as np
import pytest

from helpers import markov_bits, random_pair, split_serial, groups_of

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")


def _wah_of_bits_closed_form(bits):
    """Canonical WAH by the closed form: 31-bit groups, runs of >= 2 equal
    all-0/all-1 groups become fills (MSB-first bit order)."""
    n = len(bits)
    g = n // 31
    words = []
    lit = []
    for k in range(g):
        v = 0
        for b in bits[31 * k:31 * k + 31]:
            v = (v << 1) | int(b)
        lit.append(v)
    i = 0
    while i < g:
        w = lit[i]
        if w in (0, 0x7FFFFFFF):
            j = i
            while j < g and lit[j] == w:
                j += 1
            words.append(((0xC0000000 if w else 0x80000000) | (j - i)) if j - i >= 2 else w)
            i = j
        else:
            words.append(w)
            i += 1
    anb = n - 31 * g
    if anb:
        v = 0
        for b in bits[31 * g:]:
            v = (v << 1) | int(b)
        words.append(v)
    words.append(anb)
    return np.array(words, np.uint32)


def test_from_bits_closed_form(orc):
    rng = np.random.default_rng(11)
    for n in (0, 1, 30, 31, 32, 61, 62, 63, 500, 2000):
        for d, L in ((0.5, 1), (0.01, 1), (0.9, 300), (0.3, 40)):
            bits = markov_bits(n, d, L, rng) if n else np.zeros(0, np.uint8)
            assert np.array_equal(orc.from_bits(bits), _wah_of_bits_closed_form(bits))


def test_ops_bits_brute_force(orc):
    """Every op's result bits equal the op on the decoded bits."""
    rng = np.random.default_rng(12)
    fns = [np.bitwise_and, np.bitwise_or, np.bitwise_xor, lambda x, y: x & (1 - y)]
    for _ in range(60):
        n = int(rng.integers(1, 2500))
        ba = markov_bits(n, 0.2, 20, rng)
        bb = markov_bits(n, 0.6, 5, rng)
        a, b = orc.from_bits(ba), orc.from_bits(bb)
        for op in range(4):
            r, _ = orc.op(a, b, op)
            bits = np.zeros(n, np.uint8)
            bits[orc.rids(r)] = 1
            assert np.array_equal(bits, fns[op](ba, bb).astype(np.uint8))


def test_cnt_closed_form_and_invariants(orc):
    rng = np.random.default_rng(13)
    for _ in range(40):
        a, b, n = random_pair(rng, orc)
        ca, sa, _ = orc.cnt(a)
        cb, _, _ = orc.cnt(b)
        assert sa == n
        words, aval, anb = split_serial(a)
        w = words.astype(np.uint64)
        closed = int(sum(bin(int(x)).count("1") for x in w if x <= 0x7FFFFFFF)) + \
            int(sum(31 * (int(x) & 0x3FFFFFFF) for x in w if x >= 0xC0000000)) + bin(aval).count("1")
        assert ca == closed
        cor = orc.cnt(orc.op(a, b, 1)[0])[0]
        cand = orc.cnt(orc.op(a, b, 0)[0])[0]
        assert cor + cand == ca + cb
        # A | B == ~(~A & ~B) at the bit level
        na, nb = orc.flip(a), orc.flip(b)
        x = orc.flip(orc.op(na, nb, 0)[0])
        assert np.array_equal(orc.rids(x), orc.rids(orc.op(a, b, 1)[0]))


def test_do_cnt_collapses_all_zero(orc):
    # two literal zero words + active: do_cnt turns m_vec into one 0-fill
    c, s, after = orc.cnt(np.array([0, 0, 5, 3], np.uint32))
    assert (c, s) == (2, 65)
    assert list(after) == [0x80000002, 5, 3]


def test_oracle_matches_reference_random(orc, ref):
    rng = np.random.default_rng(14)
    paths = {}
    for _ in range(150):
        a, b, n = random_pair(rng, orc)
        if rng.random() < 0.3:
            # non-canonical operand: an uncompressed c2 result
            a = orc.op(a, orc.from_bits(markov_bits(n, 0.1, 7, rng)), 1, compress=False)[0]
        for op in range(4):
            want = ref.op(a, b, op)
            got, path = orc.op(a, b, op)
            paths[path] = paths.get(path, 0) + 1
            assert np.array_equal(want, got), (op, path, n)
    assert {"c2", "d2", "c0", "shortcut"} <= set(paths)


def test_oracle_matches_reference_literal_paths(orc, ref):
    """d1 / c1 / d1swap: one literal-only operand."""
    rng = np.random.default_rng(15)
    for _ in range(60):
        n = int(rng.integers(62, 4000))
        a = orc.from_bits(markov_bits(n, 0.5, 1.5, rng))     # mostly literal
        a = orc.decompress(a)                                 # literal-only
        b = orc.from_bits(markov_bits(n, 0.01, 3, rng))       # compressed
        for x, y in ((a, b), (b, a)):
            for op in range(4):
                assert np.array_equal(ref.op(x, y, op), orc.op(x, y, op)[0])


def test_oracle_unary_vs_reference(orc, ref):
    rng = np.random.default_rng(16)
    for _ in range(40):
        a, b, n = random_pair(rng, orc)
        u = orc.op(a, b, int(rng.integers(0, 4)), compress=False)[0]
        assert np.array_equal(ref.compress(u), orc.compress(u))
        assert np.array_equal(ref.decompress(u), orc.decompress(u))
        assert np.array_equal(ref.flip(u), orc.flip(u))
        assert np.array_equal(ref.rids(u), orc.rids(u))
        assert ref.cnt(u)[:2] == orc.cnt(u)[:2]
        assert np.array_equal(ref.cnt(u)[2], orc.cnt(u)[2])


@pytest.mark.skipif(not os.path.exists(os.path.join(GOLDEN, "wah_ops.npz")),
                    reason="golden fixture not generated")
def test_oracle_matches_golden_ops(orc):
    g = np.load(os.path.join(GOLDEN, "wah_ops.npz"))
    ncase = int(g["ncase"])
    for i in range(ncase):
        a, b = g[f"a{i}"], g[f"b{i}"]
        for op in range(4):
            assert np.array_equal(orc.op(a, b, op)[0], g[f"r{i}_{op}"]), (i, op)


@pytest.mark.skipif(not os.path.exists(os.path.join(GOLDEN, "wah_config2_digests.json")),
                    reason="golden digests not generated")
def test_oracle_matches_config2_digests_small(orc):
    """The 2^20-bit rows of the committed digests (fast enough for CPU)."""
    import hashlib
    from libfastlib_amd import datagen as G
    with open(os.path.join(GOLDEN, "wah_config2_digests.json")) as f:
        rows = json.load(f)["rows"]
    checked = 0
    for r in rows:
        if r["nbits"] > (1 << 20):
            continue
        a = G.markov_wah(r["seed_a"], r["nbits"], r["density"], r["mean_run"])
        b = G.markov_wah(r["seed_b"], r["nbits"], r["density"], r["mean_run"])
        out, path = orc.op(a, b, r["op"])
        assert hashlib.sha256(out.tobytes()).hexdigest() == r["sha256"], r
        assert path == r["path"]
        checked += 1
    assert checked > 0


def _bits_of(orc, words, n):
    bits = np.zeros(n, np.uint8)
    bits[orc.rids(words)] = 1
    return bits


def _operand_forms(orc, a, b):
    """a in its canonical, decompressed and uncompressed-c2-output forms."""
    return [a, orc.decompress(a), orc.op(a, b, 0, compress=False)[0], orc.op(a, b, 1, compress=False)[0]]


def test_op_new_bits_and_count_brute_force(orc):
    """Out-of-place results decode to the op's bits (sizes equal or not: the
    shorter operand is extended with zeros); count(mask) is the popcount of
    the AND; erase removes exactly the bits [i, j)."""
    rng = np.random.default_rng(21)
    fns = [np.bitwise_and, np.bitwise_or, np.bitwise_xor, lambda x, y: x & (1 - y)]
    for t in range(60):
        n = int(rng.integers(1, 2500))
        m = n if t % 3 else max(1, n + int(rng.integers(-70, 70)))
        ba, bb = markov_bits(n, 0.2, 20, rng), markov_bits(m, 0.6, 5, rng)
        a, b = orc.from_bits(ba), orc.from_bits(bb)
        # forms 2 and 3 are a AND c / a OR c, uncompressed: their own bits
        x = _operand_forms(orc, a, orc.from_bits(markov_bits(n, 0.3, 3, rng)))[t % 4]
        ba = _bits_of(orc, x, n)
        k = max(n, m)
        ea, eb = np.zeros(k, np.uint8), np.zeros(k, np.uint8)
        ea[:n], eb[:m] = ba, bb
        for op in range(4):
            r, _ = orc.op_new(x, b, op)
            assert np.array_equal(_bits_of(orc, r, k), fns[op](ea, eb).astype(np.uint8)), (t, op)
        if n == m:
            assert orc.count(x, b) == int(np.sum(ba & bb))
        i = int(rng.integers(0, n))
        j = int(rng.integers(i, n + 40))
        e = orc.erase(x, i, j)
        want = np.concatenate([ba[:i], ba[j:]])
        assert orc.cnt(e)[1] == len(want) and np.array_equal(_bits_of(orc, e, len(want)), want), (t, i, j)


def test_op_new_count_erase_match_reference(orc, ref):
    """The out-of-place operators (every routine, including the c1 loops and
    unequal sizes), count(mask) and erase, word for word against the
    reference on canonical, decompressed and uncompressed operands."""
    rng = np.random.default_rng(22)
    paths = set()
    for t in range(300):
        a, b, n = random_pair(rng, orc, 1, 3000)
        if t % 5 == 0 and n > 40:
            b = orc.from_bits(markov_bits(n + int(rng.integers(-40, 40)), 0.2, 5, rng))
        va, vb = _operand_forms(orc, a, b), _operand_forms(orc, b, a)
        x, y = va[int(rng.integers(0, 4))], vb[int(rng.integers(0, 4))]
        for op in range(4):
            w, p = orc.op_new(x, y, op)
            paths.add(p)
            assert np.array_equal(w, ref.op_new(x, y, op)), (t, op, p)
        if orc.cnt(x)[1] == orc.cnt(y)[1]:
            assert orc.count(x, y) == ref.count(x, y), t
        sz = orc.cnt(x)[1]
        if sz > 1:
            i = int(rng.integers(0, sz))
            j = int(rng.integers(i, sz + 40))
            assert np.array_equal(orc.erase(x, i, j), ref.erase(x, i, j)), (t, i, j)
    assert {"c0", "c1", "c2", "d2", "shortcut"} <= paths
