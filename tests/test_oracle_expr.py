"""Boolean expressions over range terms in the reference
(fastbit_selection_combine + fastbit_selection_evaluate, i.e. countQuery
over the LOGICAL_* tree, src/countQuery.cpp:944-1003) against the rows
that the same postfix program selects from brute-force term masks."""
import numpy as np

import oracle
from expr_cases import columns, rand_prog, rand_term, term_mask


def test_reference_expressions_match_brute_force(ref):
    n = 40001
    cols = columns(n, 3)
    oracle.RefIapi("ef", cols["ef"], "<binning nbins=64 equal-weight/>")
    oracle.RefIapi("ei", cols["ei"], "<binning none/><encoding equality/>")
    oracle.RefIapi("eu", cols["eu"], "<binning none/><encoding equality/>")
    rng = np.random.default_rng(4)
    names = list(cols)
    for case in range(60):
        nt = int(rng.integers(1, 6))
        terms = []
        for _ in range(nt):
            nm = names[int(rng.integers(len(names)))]
            terms.append(rand_term(rng, nm, cols[nm]))
        prog = rand_prog(rng, nt)
        want = int(oracle.expr_rows([term_mask(cols[t[0]], *t[1:]) for t in terms], prog).sum())
        assert oracle.ref_iapi_expr(terms, prog) == want, (case, terms, prog)
