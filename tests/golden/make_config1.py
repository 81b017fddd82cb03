"""Writes the relic fixture of the small, CPU-runnable workload from the
reference FastBit library.

Column: 2^20 int32 values uniform in [0, 1024) (libfastlib_amd.datagen
.uniform_i32, seed 1).  The reference builds an equality-encoded relic index
(fastbit_iapi_build_index with "<binning none/><encoding equality/>"), then
answers 256 range terms lo <= x < hi with lo uniform in [0, 1023] and the
width uniform in [1, 256] (datagen.uniform_incl_i64, seeds 2 and 3).  The
fixture keeps the SHA-256 of the index triple, the 256 hit counts and the
SHA-256 of every query's row ids (uint32, ascending).

    python tests/golden/make_config1.py
"""
from __future__ import annotations

import hashlib
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

import oracle  # noqa: E402
from libfastlib_amd import datagen as G  # noqa: E402


def workload():
    col = G.uniform_i32(1, 1 << 20, 0, 1024)
    lo = G.uniform_incl_i64(2, 256, 0, 1023)
    w = G.uniform_incl_i64(3, 256, 1, 256)
    return col, lo.astype(np.float64), (lo + w).astype(np.float64)


def main():
    col, lo, hi = workload()
    R = oracle.RefIapi("c1", col, "<binning none/><encoding equality/>")
    counts, digests = [], []
    for a, b in zip(lo, hi):
        c, r = R.query_range(float(a), 2, float(b), 1)
        counts.append(c)
        digests.append(hashlib.sha256(r.astype(np.uint32).tobytes()).hexdigest())
    meta = dict(rows=int(col.size), nobs=int(len(R.keys)), nwords=int(len(R.bitmaps)),
                keys_sha256=hashlib.sha256(R.keys.tobytes()).hexdigest(),
                offsets_sha256=hashlib.sha256(R.offsets.tobytes()).hexdigest(),
                bitmaps_sha256=hashlib.sha256(R.bitmaps.tobytes()).hexdigest(),
                counts=counts, rids_sha256=digests)
    with open(os.path.join(HERE, "config1_relic_2p20.json"), "w") as f:
        json.dump(meta, f, indent=0)
    print(meta["nobs"], meta["nwords"], sum(counts))


if __name__ == "__main__":
    main()
