"""Random boolean programs over range terms, shared by the CPU and GPU
expression tests."""
import numpy as np

OPS = (-1, -2, -3, -4, -5, -7)      # AND OR XOR NAND NOR MINUS; -6 is NOT


def rand_prog(rng, nterms, depth=4):
    """A postfix program that pushes terms 0..nterms-1 in order, combines
    them with random operators and NOTs, and never holds more than depth
    results."""
    prog, sp, used = [], 0, 0
    while used < nterms or sp > 1:
        choices = []
        if used < nterms and sp < depth:
            choices.append(0)
        if sp >= 2:
            choices.append(1)
        if sp >= 1 and (not prog or prog[-1] != -6):
            choices.append(2)
        c = choices[int(rng.integers(len(choices)))]
        if c == 0:
            prog.append(used)
            used += 1
            sp += 1
        elif c == 1:
            prog.append(int(OPS[int(rng.integers(len(OPS)))]))
            sp -= 1
        else:
            prog.append(-6)
    return prog


def columns(n, seed):
    rng = np.random.default_rng(seed)
    return {
        "ef": rng.standard_normal(n).astype(np.float32),
        "ei": rng.integers(-300, 300, n).astype(np.int32),
        "eu": rng.integers(0, 2000, n).astype(np.uint16),
    }


def rand_term(rng, name, col):
    """lo OP1 col OP2 hi with integral bounds on integer columns."""
    if col.dtype == np.float32:
        lo, hi = sorted(rng.normal(size=2) * 1.2)
    else:
        a, b = int(col.min()), int(col.max())
        lo = int(rng.integers(a - 5, b + 5))
        hi = lo + int(rng.integers(0, (b - a) // 2 + 1))
    lop = int(rng.integers(0, 3))
    rop = int(rng.integers(0, 3))
    if lop == 0 and rop == 0:
        rop = 2
    return (name, float(lo), lop, float(hi), rop)


def term_mask(col, lo, lop, hi, rop):
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
