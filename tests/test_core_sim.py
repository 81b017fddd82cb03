"""The partitioned step algorithm that the gfx950 op kernel runs, exercised
on the CPU (libfbsim.so compiles the same host/device core): arbitrary cut
points, resynchronisation at common boundaries, walk, run-summary stitching,
word write.  Checked against the oracle's *_c2 routines and against the
canonical compression of the literal op."""
import ctypes as C
import os

import numpy as np
import pytest

import oracle as O
from helpers import markov_bits, random_pair, split_serial, groups_of

SIM = os.path.join(os.path.dirname(os.path.dirname(__file__)), "libfastlib_amd", "lib", "libfbsim.so")
u32p = C.POINTER(C.c_uint32)


@pytest.fixture(scope="module")
def sim():
    if not os.path.exists(SIM):
        pytest.skip("libfbsim.so not built")
    L = C.CDLL(SIM)
    L.fbsim_op.restype = C.c_long
    return L


def run_sim(sim, a, b, op, canon, splits):
    wa, av, an = split_serial(a)
    wb, bv, bn = split_serial(b)
    g = groups_of(wa)
    cap = len(wa) + len(wb) + 8
    out = np.zeros(cap, np.uint32)
    hd = np.zeros(cap, np.uint32)
    sp = np.asarray(splits, np.uint32)
    wa = np.ascontiguousarray(wa)
    wb = np.ascontiguousarray(wb)
    n = sim.fbsim_op(wa.ctypes.data_as(u32p), len(wa), wb.ctypes.data_as(u32p), len(wb), g, op,
                     canon, sp.ctypes.data_as(u32p), len(sp), out.ctypes.data_as(u32p),
                     hd.ctypes.data_as(u32p), cap)
    assert n >= 0, n
    act = [av & bv, av | bv, av ^ bv, av & ~bv & 0xFFFFFFFF][op]
    res = list(out[:n]) + ([act] if an else []) + [an]
    return np.array(res, np.uint32), hd[:n]


def c2_oracle(orc, a, b, op):
    pa, na = O._arr(a)[1:]
    pb, nb = O._arr(b)[1:]
    return orc._call_out("wah_c2", pa, na, pb, nb, C.c_int(op))


def c2_raw_oracle(orc, a, b, op):
    pa, na = O._arr(a)[1:]
    pb, nb = O._arr(b)[1:]
    return orc._call_out("wah_c2_raw", pa, na, pb, nb, C.c_int(op))


def test_sim_c2_matches_oracle(sim, orc):
    rng = np.random.default_rng(21)
    for t in range(120):
        a, b, n = random_pair(rng, orc)
        if t % 3 == 0:
            a = orc.op(a, orc.from_bits(markov_bits(n, 0.1, 7, rng)), 1, compress=False)[0]
            a = orc.compress(a) if t % 2 else a
        g = groups_of(split_serial(a)[0])
        for op in range(4):
            splits = rng.integers(0, g + 1, size=int(rng.integers(0, 25)))
            got, heads = run_sim(sim, a, b, op, 0, splits)
            assert np.array_equal(got, c2_oracle(orc, a, b, op)), (t, op)
            # heads are the start groups of the result words
            words = split_serial(got)[0]
            lens = np.where(words > 0x7FFFFFFF, words & 0x3FFFFFFF, 1)
            assert np.array_equal(heads, np.concatenate([[0], np.cumsum(lens)[:-1]]).astype(np.uint32))


def test_sim_canonical_matches_compressed_literal_op(sim, orc):
    rng = np.random.default_rng(22)
    for t in range(80):
        a, b, n = random_pair(rng, orc)
        g = groups_of(split_serial(a)[0])
        for op in range(4):
            lit = orc.op(orc.decompress(a), orc.decompress(b), op, compress=False)[0]
            want = orc.compress(lit)
            splits = rng.integers(0, g + 1, size=int(rng.integers(0, 25)))
            got, _ = run_sim(sim, a, b, op, 1, splits)
            assert np.array_equal(got, want), (t, op)


def test_sim_every_cut_point(sim, orc):
    """Cut at every single group of a small pair."""
    rng = np.random.default_rng(23)
    for _ in range(6):
        a = orc.from_bits(markov_bits(620, 0.3, 40, rng))
        b = orc.from_bits(markov_bits(620, 0.05, 3, rng))
        g = groups_of(split_serial(a)[0])
        for op in range(4):
            got, _ = run_sim(sim, a, b, op, 0, np.arange(g + 1))
            assert np.array_equal(got, c2_oracle(orc, a, b, op))


def test_sim_every_cut_noncanonical_chains(sim, orc):
    """Operands that are results of earlier *_c2 ops (count-1 fills, raw
    pushed words) and sparse operands with long fills: every cut point, all
    four ops, against the oracle's sequential routine."""
    rng = np.random.default_rng(24)
    for t in range(20):
        n = int(rng.integers(300, 30000))
        parts = [orc.from_bits(markov_bits(n, float(rng.choice([0.002, 0.02, 0.2])),
                                           float(rng.choice([1, 3, 40])), rng)) for _ in range(4)]
        a = parts[0]
        for q in parts[1:3]:
            a = c2_oracle(orc, a, q, int(rng.integers(0, 4)))
        b = parts[3] if t % 2 else c2_oracle(orc, parts[3], parts[1], 1)
        g = groups_of(split_serial(a)[0])
        for op in range(4):
            want = c2_oracle(orc, a, b, op)
            got, _ = run_sim(sim, a, b, op, 0, np.arange(g + 1))
            assert np.array_equal(got, want), (t, op)
            got, _ = run_sim(sim, b, a, op, 0, rng.integers(0, g + 1, size=40))
            assert np.array_equal(got, c2_oracle(orc, b, a, op)), (t, op, "swapped")


def test_sim_raw_c2_words(sim, orc):
    """The raw mode (MODE_RAW = 2): the *_c2 loop's own words, not
    compressed (the out-of-place operators return them), from walks that
    start only where both operands start a code word; every cut point on
    small pairs, random cuts on larger and non-canonical ones."""
    rng = np.random.default_rng(25)
    for t in range(60):
        n = int(rng.integers(100, 20000))
        a = orc.from_bits(markov_bits(n, float(rng.choice([0.002, 0.02, 0.2, 0.6])),
                                      float(rng.choice([1, 3, 40, 300])), rng))
        b = orc.from_bits(markov_bits(n, float(rng.choice([0.002, 0.02, 0.2, 0.6])),
                                      float(rng.choice([1, 3, 40, 300])), rng))
        if t % 3 == 0:
            a = c2_raw_oracle(orc, a, b, int(rng.integers(0, 4)))
        g = groups_of(split_serial(a)[0])
        cuts = np.arange(g + 1) if g < 700 else rng.integers(0, g + 1, size=200)
        for op in range(4):
            want = c2_raw_oracle(orc, a, b, op)
            got, heads = run_sim(sim, a, b, op, 2, cuts)
            assert np.array_equal(got, want), (t, op)
            words = split_serial(got)[0]
            lens = np.where(words > 0x7FFFFFFF, words & 0x3FFFFFFF, 1)
            assert np.array_equal(heads, np.concatenate([[0], np.cumsum(lens)[:-1]]).astype(np.uint32))


def test_sim_c2_canonical_with_representations(sim, orc):
    """The *_c2 result from the group domain (canonical words of the op's
    values, a single all-0 / all-1 group written as a count-1 fill where
    c2_rep_at walks to a fill token from the last synchronisation point)
    equals the oracle's *_c2 words, on operands with count-1 fills and
    literal 0 / all-ones words inside runs of fills."""
    sim.fbsim_c2_canon_rep.restype = C.c_long
    rng = np.random.default_rng(23)
    for t in range(160):
        a, b, n = random_pair(rng, orc)
        if t % 3 == 0:
            a = orc.op(a, orc.from_bits(markov_bits(n, 0.1, 7, rng)), 1, compress=False)[0]
            a = orc.compress(a) if t % 2 else a
        if t % 4 == 1:
            b = orc.op(b, orc.from_bits(markov_bits(n, 0.3, 3, rng)), 0, compress=False)[0]
        wa, av, an = split_serial(a)
        wb, bv, bn = split_serial(b)
        g = groups_of(wa)
        wa = np.ascontiguousarray(wa)
        wb = np.ascontiguousarray(wb)
        for op in range(4):
            cap = g + 8
            out = np.zeros(cap, np.uint32)
            k = sim.fbsim_c2_canon_rep(wa.ctypes.data_as(u32p), len(wa), wb.ctypes.data_as(u32p), len(wb), g,
                                       op, out.ctypes.data_as(u32p), cap)
            assert k >= 0, k
            act = [av & bv, av | bv, av ^ bv, av & ~bv & 0xFFFFFFFF][op]
            got = np.array(list(out[:k]) + ([act] if an else []) + [an], np.uint32)
            assert np.array_equal(got, c2_oracle(orc, a, b, op)), (t, op)


@pytest.mark.parametrize("d,run", [(0.01, 64), (0.01, 1), (0.1, 8), (0.5, 64)])
def test_sim_c2_representations_markov(sim, orc, d, run):
    """The same on 2^20-bit Markov operands, where the *_c2 loops emit
    count-1 fills (checked to occur in the expected words)."""
    sim.fbsim_c2_canon_rep.restype = C.c_long
    rng = np.random.default_rng(int(d * 1000) + run)
    n = 1 << 20
    a = orc.from_bits(markov_bits(n, d, run, rng))
    b = orc.from_bits(markov_bits(n, d, run, rng))
    wa, av, an = split_serial(a)
    wb, bv, bn = split_serial(b)
    g = groups_of(wa)
    wa = np.ascontiguousarray(wa)
    wb = np.ascontiguousarray(wb)
    ones1 = 0
    for op in range(4):
        want = c2_oracle(orc, a, b, op)
        ww = split_serial(want)[0]
        ones1 += int(np.sum((ww == 0x80000001) | (ww == 0xC0000001)))
        out = np.zeros(g + 8, np.uint32)
        k = sim.fbsim_c2_canon_rep(wa.ctypes.data_as(u32p), len(wa), wb.ctypes.data_as(u32p), len(wb), g, op,
                                   out.ctypes.data_as(u32p), g + 8)
        act = [av & bv, av | bv, av ^ bv, av & ~bv & 0xFFFFFFFF][op]
        got = np.array(list(out[:k]) + ([act] if an else []) + [an], np.uint32)
        assert np.array_equal(got, want), op
    if (d, run) == (0.01, 64):
        assert ones1 > 0
