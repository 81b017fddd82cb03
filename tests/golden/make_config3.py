"""Writes the binned-float-index fixture from the reference FastBit library.

Column: 2^27 float32 values from a pool of 10^5 distinct values (70% N(0,1)
draws, 30% Zipf(1.2) support), generated by libfastlib_amd.datagen.float_mix
with seed 3.  The reference builds its index with the spec
``<binning nbins=10000 equal-weight/>`` through fastbit_iapi_build_index;
the fixture keeps the bin boundaries the reference computed
(bin::binBoundaries), the boundaries as the reference rebuilds them from the
deconstructed keys (src/ibin.cpp:358-372), the bin count, and the SHA-256 of
the serialized bitmaps.

    python tests/golden/make_config3.py [log2_rows]
"""
from __future__ import annotations

import hashlib
import json
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

import oracle  # noqa: E402
from libfastlib_amd import datagen as G  # noqa: E402

SEED = 3


def main(log2n=27):
    n = 1 << log2n
    t = time.time()
    col = G.float_mix(SEED, n, pool=100000, wnorm=0.7, s=1.2)
    print("generated", round(time.time() - t, 1), flush=True)
    t = time.time()
    R = oracle.RefIapi("c3", col, "<binning nbins=10000 equal-weight/>")
    print("reference build", round(time.time() - t, 1), flush=True)
    bounds = oracle.bounds_from_keys(R.keys)
    ref_bounds = oracle.ref_bin_bounds("c3")
    tag = f"config3_2p{log2n}"
    np.savez_compressed(os.path.join(HERE, tag + "_bounds.npz"), bounds=bounds, ref_bounds=ref_bounds)
    meta = dict(rows=n, seed=SEED, spec="<binning nbins=10000 equal-weight/>",
                nobs=int(len(R.keys) // 2), nwords=int(len(R.bitmaps)),
                keys_sha256=hashlib.sha256(R.keys.tobytes()).hexdigest(),
                offsets_sha256=hashlib.sha256(R.offsets.tobytes()).hexdigest(),
                bitmaps_sha256=hashlib.sha256(R.bitmaps.tobytes()).hexdigest(),
                reference_build_s=round(time.time() - t, 1))
    with open(os.path.join(HERE, tag + ".json"), "w") as f:
        json.dump(meta, f, indent=1)
    print(meta)


if __name__ == "__main__":
    main(int(sys.argv[1]) if len(sys.argv) > 1 else 27)
