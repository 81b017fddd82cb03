"""Shared test helpers: small seeded bit strings and WAH utilities."""
from __future__ import annotations

import numpy as np

HEADER0 = 0x80000000
ALLONES = 0x7FFFFFFF


def markov_bits(n, d, L, rng):
    """Two-state Markov bit string (numpy, for small sizes)."""
    p10 = 1.0 / L
    p01 = min(1.0, d * p10 / (1.0 - d))
    out = np.empty(n, np.uint8)
    s = rng.random() < d
    r = rng.random(n)
    for i in range(n):
        out[i] = s
        s = (r[i] >= p10) if s else (r[i] < p01)
    return out


def split_serial(s):
    s = np.asarray(s, np.uint32)
    anb = int(s[-1])
    if anb:
        return s[:-2], int(s[-2]), anb
    return s[:-1], 0, 0


def groups_of(words):
    w = np.asarray(words, np.uint64)
    fill = w > ALLONES
    return int(np.where(fill, w & 0x3FFFFFFF, 1).sum())


CASES = [(d, L) for d in (1e-3, 0.02, 0.3, 0.5, 0.9, 0.99) for L in (1, 3, 40, 200, 2000)]


def random_pair(rng, orc, nmin=32, nmax=6000):
    n = int(rng.integers(nmin, nmax))
    d, L = CASES[int(rng.integers(0, len(CASES)))]
    d = min(d, 0.999)
    a = orc.from_bits(markov_bits(n, d, L, rng))
    b = orc.from_bits(markov_bits(n, d, L, rng))
    return a, b, n
