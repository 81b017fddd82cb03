"""Boolean expressions over range terms on the MI355X (fbh_expr_prepare):
counts against the rows the same postfix program selects from brute-force
term masks, and against the reference's fastbit_selection_combine +
fastbit_selection_evaluate where that library is present."""
import numpy as np
import pytest

import oracle as O
from expr_cases import columns, rand_prog, rand_term, term_mask

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def L():
    import libfastlib_amd as L
    L.lib()
    return L


def _indexes(L, cols):
    return {"ef": L.Index.bin_equal_weight(cols["ef"], 64), "ei": L.Index.relic(cols["ei"]),
            "eu": L.Index.relic(cols["eu"])}


def _cases(rng, cols, nq):
    names = list(cols)
    out = []
    for _ in range(nq):
        nt = int(rng.integers(1, 6))
        terms = []
        for _ in range(nt):
            nm = names[int(rng.integers(len(names)))]
            terms.append(rand_term(rng, nm, cols[nm]))
        out.append((terms, rand_prog(rng, nt)))
    return out


@pytest.mark.parametrize("n", [40001, 31 * 2048 * 3, (1 << 20) + 17])
def test_expressions_brute_force(L, n):
    cols = columns(n, 3)
    ixs = _indexes(L, cols)
    rng = np.random.default_rng(n)
    cases = _cases(rng, cols, 40)
    want = [int(O.expr_rows([term_mask(cols[t[0]], *t[1:]) for t in terms], prog).sum())
            for terms, prog in cases]
    for plan in (0, 1, 3):
        print(f"expressions n={n} plan={plan}", flush=True)
        E = L.Expressions([([(ixs[t[0]],) + t[1:] for t in terms], prog) for terms, prog in cases], plan)
        hits, _ = E.run()
        assert [int(h) for h in hits] == want, plan
        hits2, _ = E.run()
        assert np.array_equal(hits, hits2)


def test_expressions_match_reference(L):
    if O.reference() is None:
        pytest.skip("reference library (oracle/_ref) not built here")
    n = 40001
    cols = columns(n, 3)
    O.RefIapi("ef", cols["ef"], "<binning nbins=64 equal-weight/>")
    O.RefIapi("ei", cols["ei"], "<binning none/><encoding equality/>")
    O.RefIapi("eu", cols["eu"], "<binning none/><encoding equality/>")
    ixs = _indexes(L, cols)
    rng = np.random.default_rng(99)
    cases = _cases(rng, cols, 48)
    print("expressions vs reference", flush=True)
    E = L.Expressions([([(ixs[t[0]],) + t[1:] for t in terms], prog) for terms, prog in cases])
    hits, _ = E.run()
    for (terms, prog), h in zip(cases, hits):
        assert int(h) == O.ref_iapi_expr(terms, prog), (terms, prog)


def test_expressions_tune_and_and_only(L):
    """An AND-only program gives the conjunction batch's counts; the tuned
    configuration gives the same counts."""
    n = 1 << 20
    cols = columns(n, 5)
    ixs = _indexes(L, cols)
    rng = np.random.default_rng(8)
    qs = []
    for _ in range(32):
        terms = [rand_term(rng, nm, cols[nm]) for nm in ("ef", "ei", "eu")]
        qs.append(terms)
    print("expressions tune", flush=True)
    C = L.Conjunctions([[(ixs[t[0]],) + t[1:] for t in terms] for terms in qs])
    E = L.Expressions([([(ixs[t[0]],) + t[1:] for t in terms], [0, 1, -1, 2, -1]) for terms in qs])
    hc, _ = C.run()
    he, _ = E.run()
    assert np.array_equal(hc, he)
    E.tune()
    he2, _ = E.run()
    assert np.array_equal(he, he2)


def test_expression_program_errors(L):
    cols = columns(4001, 1)
    ix = L.Index.relic(cols["ei"])
    t = [(ix, -10.0, 2, 10.0, 1), (ix, 0.0, 2, 50.0, 1)]
    for prog in ([0, 1], [-1], [0, -6, -1], [0, 2, -1], [0, 1, -9], [0, 0, 0, 0, 0, -1, -1, -1, -1]):
        with pytest.raises(L.FastbitError):
            L.Expressions([(t, prog)])


def test_expression_short_circuit_and_rids(L):
    """Programs whose left operand of AND / MINUS is empty on some tiles
    only (its term is a range of a sorted column, so its rows are one
    contiguous run): counts and row ids against brute force, as a batch
    that needs a stack of two results and as one that needs four."""
    n = 31 * 2048 * 4 + 99
    rng = np.random.default_rng(17)
    cols = {"ef": np.sort(rng.standard_normal(n)).astype(np.float32),
            "ei": rng.integers(-300, 300, n).astype(np.int32),
            "eu": rng.integers(0, 2000, n).astype(np.uint16)}
    ixs = _indexes(L, cols)
    progs2 = ([0, 1, -1], [0, 1, -7], [0, 1, -1, -6])
    progs4 = progs2 + ([0, 1, 2, -1, -1],)
    cases = []
    for lo in (-2.5, -0.3, 0.9):
        terms = [("ef", lo, 2, lo + 0.4, 1), ("ei", -100.0, 2, 150.0, 1), ("eu", 10.0, 1, 1500.0, 2)]
        cases.append(terms)
    for progs in (progs2, progs4):
        qs = [(terms, p) for terms in cases for p in progs]
        want = [O.expr_rows([term_mask(cols[t[0]], *t[1:]) for t in terms], p) for terms, p in qs]
        E = L.Expressions([([(ixs[t[0]],) + t[1:] for t in terms], p) for terms, p in qs])
        hits, _ = E.run()
        assert [int(h) for h in hits] == [int(w.sum()) for w in want], len(progs)
        off, rids = E.rids()
        for i, w in enumerate(want):
            assert np.array_equal(rids[off[i]:off[i + 1]], np.nonzero(w)[0].astype(np.uint32)), (len(progs), i)
        bits = want[1].astype(np.uint8)
        assert int(E.result(1).cnt()) == int(bits.sum())
