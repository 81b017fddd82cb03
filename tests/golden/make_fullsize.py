"""Writes the answers of the full-size query workloads.

Binned float workload (2^27 float32 rows, datagen.bin_f32_column):
  - the 1024 quantile ranges lo <= x < hi of the range-query batch
    (datagen.quantile_ranges, seed BIN_F32_QUERY_SEED): the reference
    FastBit library's hit counts for every one of them and the SHA-256 of
    its row ids for every 16th, each range evaluated as one
    qContinuousRange through fastbit_selection_evaluate over the index the
    reference built with "<binning nbins=10000 equal-weight/>"; every count
    and digest is also checked equal to a brute-force scan of the column;
  - the 8192 test ranges of the learned plan choice (seed
    BIN_F32_TEST_SEED): brute-force counts, of which 256 are checked
    against the reference.
Conjunctive workload (4 columns of 2^26 rows, datagen.table4 and
datagen.conj_queries): brute-force counts of the 4096 conjunctions, of
which 48 are checked against the reference (fastbit_selection_combine of
the terms, counted by fastbit_selection_evaluate).

The reference runs in worker processes (it evaluates one query at a time
on one thread); each worker builds its own copy of the index.

    python tests/golden/make_fullsize.py [bin|conj|all]
"""
from __future__ import annotations

import hashlib
import multiprocessing as mp
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

import oracle  # noqa: E402
from libfastlib_amd import datagen as G  # noqa: E402

SPEC = "<binning nbins=10000 equal-weight/>"
RID_EVERY = 16
WORKERS = 4
_COL = None


def rid_digest(rids):
    return hashlib.sha256(np.ascontiguousarray(rids, np.uint32).tobytes()).hexdigest()


def _bin_worker(args):
    lo, hi, want_rids = args
    R = oracle.RefIapi("c3w", _COL, SPEC)
    out = []
    for a, b, w in zip(lo, hi, want_rids):
        n, rids = R.query_range(float(a), 2, float(b), 1, want_rids=bool(w))
        out.append((n, rid_digest(rids) if w else ""))
    return out


def bin_workload():
    global _COL
    t = time.time()
    col = G.bin_f32_column(27)
    srt = np.sort(col)
    lo, hi = G.quantile_ranges(srt, G.BIN_F32_QUERY_SEED, 1024)
    tlo, thi = G.quantile_ranges(srt, G.BIN_F32_TEST_SEED, 8192)
    print("column and queries", round(time.time() - t, 1), flush=True)
    brute = (np.searchsorted(srt, hi, "left") - np.searchsorted(srt, lo, "left")).astype(np.int64)
    tbrute = (np.searchsorted(srt, thi, "left") - np.searchsorted(srt, tlo, "left")).astype(np.int64)
    del srt
    want = (np.arange(1024) % RID_EVERY) == 0
    brute_dig = {int(i): rid_digest(np.nonzero((col >= lo[i]) & (col < hi[i]))[0]) for i in np.nonzero(want)[0]}
    # reference: the 1024 batch ranges, then 256 of the test ranges
    sample = np.arange(0, 8192, 32)
    jobs_lo = np.concatenate([lo, tlo[sample]])
    jobs_hi = np.concatenate([hi, thi[sample]])
    jobs_w = np.concatenate([want, np.zeros(len(sample), bool)])
    parts = np.array_split(np.arange(len(jobs_lo)), WORKERS)
    _COL = col
    t = time.time()
    with mp.get_context("fork").Pool(WORKERS) as pool:
        res = pool.map(_bin_worker, [(jobs_lo[p], jobs_hi[p], jobs_w[p]) for p in parts])
    print("reference queries", round(time.time() - t, 1), flush=True)
    flat = [x for r in res for x in r]
    ref_counts = np.array([c for c, _ in flat[:1024]], np.int64)
    ref_test = np.array([c for c, _ in flat[1024:]], np.int64)
    if not np.array_equal(ref_counts, brute):
        bad = np.nonzero(ref_counts != brute)[0]
        raise SystemExit(f"reference counts differ from the scan at {bad[:10]}")
    if not np.array_equal(ref_test, tbrute[sample]):
        raise SystemExit("reference counts of the test ranges differ from the scan")
    digs = []
    for i in np.nonzero(want)[0]:
        d = flat[int(i)][1]
        if d != brute_dig[int(i)]:
            raise SystemExit(f"reference row ids of range {i} differ from the scan's")
        digs.append(d)
    np.savez_compressed(os.path.join(HERE, "bin_f32_2p27_queries.npz"),
                        lo=lo, hi=hi, counts=ref_counts, rid_index=np.nonzero(want)[0].astype(np.int32),
                        rid_sha256=np.array(digs), test_lo=tlo, test_hi=thi, test_counts=tbrute,
                        test_ref_index=sample.astype(np.int32))
    print("bin workload written: hits", int(ref_counts.sum()), flush=True)


_TABLE = None


def _conj_count(q):
    m = None
    for ci, a, b in q:
        c = _TABLE[ci]
        if c.dtype.kind in "iu":
            v = c.astype(np.int64)
            t = (v >= int(a)) & (v < int(b))
        else:
            v = c.astype(np.float64)
            t = (v >= a) & (v < b)
        m = t if m is None else (m & t)
    return int(np.count_nonzero(m))


def conj_workload():
    global _TABLE
    t = time.time()
    n = 1 << 26
    cols = G.table4(G.CONJ_TABLE_SEED, n)
    qs = G.conj_queries(G.CONJ_QUERY_SEED, cols, 4096)
    for c, q in ((c, q) for q in qs for c in q):
        ci, a, b = c
        if cols[ci].dtype.kind in "iu" and (a != int(a) or b != int(b)):
            raise SystemExit("a fractional bound on an integer column")
    print("table and queries", round(time.time() - t, 1), flush=True)
    _TABLE = cols
    t = time.time()
    with mp.get_context("fork").Pool(8) as pool:
        counts = np.array(pool.map(_conj_count, qs, chunksize=16), np.int64)
    print("brute-force conjunctions", round(time.time() - t, 1), flush=True)
    # the reference on a sample: the four columns registered and indexed
    # (the bounds are integral on the integer columns, so every index and
    # plan selects the same rows), the terms ANDed by selection trees
    names = [f"t4c{i}" for i in range(4)]
    refs = [oracle.RefIapi(nm, c, SPEC) for nm, c in zip(names, cols)]
    sample = np.arange(0, 4096, 4096 // 48)
    for i in sample:
        q = qs[i]
        terms = [(names[ci], a, 2, b, 1) for ci, a, b in q]
        prog = [0] + [x for k in range(1, len(q)) for x in (k, -1)]
        r = oracle.ref_iapi_expr(terms, prog)
        if r != counts[i]:
            raise SystemExit(f"reference count of conjunction {i} is {r}, the scan's {counts[i]}")
    del refs
    nterm = np.array([len(q) for q in qs], np.int32)
    flat = np.array([(ci, a, b) for q in qs for (ci, a, b) in q], np.float64)
    np.savez_compressed(os.path.join(HERE, "conj_4x2p26_counts.npz"), counts=counts, nterms=nterm,
                        terms=flat, ref_index=sample.astype(np.int32))
    print("conjunctions written: hits", int(counts.sum()), flush=True)


if __name__ == "__main__":
    what = sys.argv[1] if len(sys.argv) > 1 else "all"
    if oracle.reference() is None:
        raise SystemExit("the reference library (oracle/_ref) is not built")
    if what in ("bin", "all"):
        bin_workload()
    if what in ("conj", "all"):
        conj_workload()
