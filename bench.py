#!/usr/bin/env python3
"""Benchmark of the FastBit query hot path on one MI355X.

Workload (the headline metric, range queries/s at 2^27 rows): a float32
column of 2^27 rows drawn from a pool of 10^5 distinct values (70% N(0,1)
draws, 30% Zipf(1.2) support; seed 3), an ibis::bin index with 10,000
equal-weight bins built on the GPU (the boundaries FastBit's
"<binning nbins=10000 equal-weight/>" chooses), and 1024 range queries
lo <= x < hi whose endpoints are the column's empirical quantiles at uniform
draws.  One step evaluates the 1024 queries (hit counts) with the
plan the ibis heuristic picks per query; inputs are resident in HBM.

The same JSON line reports
  - per-plan throughput (literal / scan) and the algorithmic HBM bytes they
    move, the roofline of the dominant kernel;
  - WAH AND / OR at 2^28 bits on the bandwidth-bound Markov cells
    (density 0.5 / mean run 1 and 64, density 0.01 / run 1), in GB/s of
    algorithmic bytes 4(|A|w + |B|w + |C|w), operands rotated over several
    copies so that they stream from HBM;
  - the CPU baseline: the reference FastBit library (oracle/_ref) answering
    a sample of the same queries on the host with the same index attached.

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W]
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def config3(log2n, nq, extra=()):
    """The column, its nq quantile queries (lo, hi), and one more (lo, hi)
    query set per (seed, count) in `extra`."""
    from libfastlib_amd import datagen as G
    n = 1 << log2n
    t = time.time()
    col = G.bin_f32_column(log2n)
    log(f"column: {n} rows in {time.time() - t:.1f}s")
    t = time.time()
    srt = np.sort(col)
    lo, hi = G.quantile_ranges(srt, G.BIN_F32_QUERY_SEED, nq)
    more = [G.quantile_ranges(srt, s, k) for s, k in extra]
    log(f"queries: {nq} + {sum(k for _, k in extra)} in {time.time() - t:.1f}s")
    del srt
    return col, lo, hi, more


def measured_traffic(kernel, summary="traffic_bin2p27_1024q.txt"):
    """HBM bytes per launch of `kernel` on this workload, from the committed
    rocprofv3 FETCH_SIZE / WRITE_SIZE summary (counter-scaled), or None."""
    path = os.path.join(ROOT, "profiles", summary)
    if not os.path.exists(path):
        return None
    for line in open(path):
        f = line.split()
        if f and f[0] == kernel:
            kv = dict(x.split("=") for x in f[1:])
            return float(kv["fetch_bytes"]) + float(kv["write_bytes"])
    return None


def query_bytes(ix, lo, hi, plan):
    return np.array([ix.query_bytes(float(lo[i]), 2, float(hi[i]), 1, plan) for i in range(len(lo))])


WAH_CELLS = [(d, run) for d in (1e-4, 1e-2, 0.5) for run in (1, 64, 4096)]
OP_NAMES = ("AND", "OR", "XOR", "MINUS")


def wah_cells(L, reps=40, copies=4):
    """The nine Markov cells (density d, mean 1-run length L) of two 2^28-bit
    vectors: AND / OR / XOR / MINUS in place plus compress(), and cnt(), each
    timed with inputs resident over `copies` rotating copies (working set
    beyond the last-level cache).  Algorithmic bytes: 4 x (words of A + of B
    + of the result) per op, 4 x words of A per cnt."""
    from libfastlib_amd import datagen as G
    out = []
    for d, run in WAH_CELLS:
        wa = G.markov_wah(100, 1 << 28, d, run)
        wb = G.markov_wah(101, 1 << 28, d, run)
        A = [L.Bitvector.from_serial(wa) for _ in range(copies)]
        B = [L.Bitvector.from_serial(wb) for _ in range(copies)]
        cell = f"d={d:g},L={run}"
        for op, name in enumerate(OP_NAMES):
            z = A[0].copy()
            path = z.op(B[0], op)
            outw = z.nwords + 1 + (1 if z.write()[-1] else 0)
            ms, _ = L.bench_op_rot(A, B, op, reps)
            byts = 4.0 * (len(wa) + len(wb) + outw)
            gbs = byts / (ms * 1e-3) / 1e9 if ms > 0 else None
            out.append(dict(cell=cell, op=name, path=path, us=round(ms * 1e3, 2),
                            in_words=[int(len(wa)), int(len(wb))], out_words=int(outw),
                            GBs=None if gbs is None else round(gbs, 1),
                            hbm_frac=None if gbs is None else round(gbs / HBM_PEAK_GBS, 3)))
            log("wah", out[-1])
        ms, ones = L.bench_cnt_rot(A, reps)
        gbs = 4.0 * len(wa) / (ms * 1e-3) / 1e9 if ms > 0 else None
        out.append(dict(cell=cell, op="cnt", path="cnt", us=round(ms * 1e3, 2), in_words=[int(len(wa))],
                        ones=int(ones), GBs=None if gbs is None else round(gbs, 1),
                        hbm_frac=None if gbs is None else round(gbs / HBM_PEAK_GBS, 3)))
        log("wah", out[-1])
    return out


def relic_ranges(seed, k, card=65536):
    """k ranges lo <= x < hi over integer values [0, card): lo uniform,
    width log-uniform in [1, card]."""
    r = np.random.default_rng(seed)
    lo = r.integers(0, card, size=k)
    w = np.floor(2.0 ** r.uniform(0, np.log2(card), size=k)).astype(np.int64)
    hi = np.minimum(lo + np.maximum(w, 1), card)
    return lo.astype(np.float64), hi.astype(np.float64)


def conj_section(L, log2n, nq, steps, learned=None, cpu=None):
    """4 columns of 2^log2n rows (uint16 uniform -> relic; int32 Zipf(1.1),
    float32 N(0,1), int32 sorted runs -> 10000 equal-weight bins), nq
    queries ANDing 2-4 range terms of selectivity log-uniform in
    [1e-6, 0.3]; one prepared, autotuned, graph-captured batch.  With
    learned = (train, test, epochs), the learned plan choice also runs on
    the uint16 relic column (65,536 bitmaps; uniform lo, log-uniform
    widths), where the ibis rule prices wide ranges as scans."""
    from libfastlib_amd import datagen as G
    n = 1 << log2n
    t = time.time()
    cols = G.table4(G.CONJ_TABLE_SEED, n)
    qs = G.conj_queries(G.CONJ_QUERY_SEED, cols, nq)
    log(f"conj table: {time.time() - t:.1f}s")
    t = time.time()
    ixs = [L.Index.relic(cols[0]), L.Index.bin_equal_weight(cols[1], 10000),
           L.Index.bin_equal_weight(cols[2], 10000), L.Index.bin_equal_weight(cols[3], 10000)]
    L.lib().fbh_sync()
    build_s = time.time() - t
    batch = L.Conjunctions([[(ixs[ci], lo, 2, hi, 1) for (ci, lo, hi) in q] for q in qs])
    cfg = batch.tune()
    batch.run()
    ms = []
    for _ in range(steps):
        hits, m = batch.run()
        ms.append(m)
    plans = batch.plans()
    # bytes the terms' plans read with no early exit (bitmap words of the OR
    # sets and edge bins, raw values of edge rows or of the scanned column):
    # the kernel stops a tile at its first empty term, so it reads less
    res_bytes = 4.0 * (n // 31 + 1)
    flat = [(ci, lo, hi) for q in qs for (ci, lo, hi) in q]
    nominal = sum(ixs[ci].query_bytes(float(lo), 2, float(hi), 1, int(p)) - res_bytes
                  for (ci, lo, hi), p in zip(flat, plans))
    # row ids of every query: a count replay, one more launch writing the
    # ids, the copy to host memory
    t = time.perf_counter()
    roff, rids = batch.rids()
    rid_s = time.perf_counter() - t
    if not np.array_equal(np.diff(roff), hits):
        raise RuntimeError("row-id offsets differ from the counts")
    # the same batch with every term scanned gives the same counts
    ref, _ = L.Conjunctions([[(ixs[ci], lo, 2, hi, 1) for (ci, lo, hi) in q] for q in qs], plan=3).run()
    if not np.array_equal(ref, hits):
        raise RuntimeError("conjunction counts differ from the scan plan's")
    out = dict(rows=n, queries=nq, terms=int(batch.nterms), index_build_s=round(build_s, 2),
               qps=round(nq / (np.median(ms) * 1e-3), 1), ms_per_batch=round(float(np.median(ms)), 3),
               tuned=cfg, term_plans={"literal": int(np.sum(plans == 1)), "scan": int(np.sum(plans == 3))},
               hits_total=int(hits.sum()),
               nominal_term_GB=round(nominal / 1e9, 3),
               nominal_GBs=round(nominal / (np.median(ms) * 1e-3) / 1e9, 1),
               traffic_per_launch=measured_traffic("k_conj", "traffic_conj_4x2p26.txt"),
               rid_output=dict(rids=int(len(rids)), seconds=round(rid_s, 3),
                               rids_per_s=round(len(rids) / rid_s, 0),
                               note="fbh_conj_rids: count replay, row-id launch, ids copied to host "
                                    "memory (PCIe-inclusive wall time)"))
    tr = out["traffic_per_launch"]
    if tr:
        out["hbm_GBs_measured"] = round(tr / (np.median(ms) * 1e-3) / 1e9, 1)
        out["hbm_frac_measured"] = round(out["hbm_GBs_measured"] / HBM_PEAK_GBS, 4)
    log("conj", out)
    if cpu:
        out["cpu_baseline"] = cpu_baseline_conj(ixs, cols, qs, hits, *cpu)
        log("conj cpu", out["cpu_baseline"])
    # the same terms combined by boolean programs (fastbit_selection_combine
    # trees): the first two terms ORed, a third one ANDed NOT, a fourth ANDed
    def program(k):
        p = [0] if k == 1 else [0, 1, -2]
        if k >= 3:
            p += [2, -7]
        if k >= 4:
            p += [3, -1]
        return p
    eq = [([(ixs[ci], lo, 2, hi, 1) for (ci, lo, hi) in q], program(len(q))) for q in qs]
    ebatch = L.Expressions(eq)
    ecfg = ebatch.tune()
    ebatch.run()
    ems = []
    for _ in range(steps):
        ehits, m = ebatch.run()
        ems.append(m)
    eref, _ = L.Expressions(eq, plan=3).run()
    if not np.array_equal(eref, ehits):
        raise RuntimeError("expression counts differ from the scan plan's")
    out["expressions"] = dict(program="(t0 OR t1) AND NOT t2 AND t3", qps=round(nq / (np.median(ems) * 1e-3), 1),
                              ms_per_batch=round(float(np.median(ems)), 3), tuned=ecfg,
                              hits_total=int(ehits.sum()))
    log("expressions", out["expressions"])
    # the conjunctions themselves as programs (t0 AND t1 AND ...): the
    # expression kernel's short-circuit against the conjunction kernel
    aq = [([(ixs[ci], lo, 2, hi, 1) for (ci, lo, hi) in q], [0] + [x for k in range(1, len(q)) for x in (k, -1)])
          for q in qs]
    abatch = L.Expressions(aq)
    abatch.tune()
    ams = []
    for _ in range(steps):
        ahits, m = abatch.run()
        ams.append(m)
    if not np.array_equal(ahits, hits):
        raise RuntimeError("AND programs differ from the conjunction batch")
    out["expressions_and"] = dict(program="t0 AND t1 AND ...", qps=round(nq / (np.median(ams) * 1e-3), 1),
                                  ms_per_batch=round(float(np.median(ams)), 3))
    log("expressions_and", out["expressions_and"])
    rel = None
    if learned:
        ntr, nte, epochs = learned
        rel = learned_section(L, ixs[0], relic_ranges(7001, ntr), relic_ranges(7002, nte), epochs)
        rel["workload"] = (f"relic over the uint16 column of 2^{log2n} rows (65,536 bitmaps), ranges "
                           f"lo <= x < hi with lo uniform and width log-uniform in [1, 65536]")
    del ixs
    return out, rel


def two_level_section(L, log2n, nq, steps):
    """The same ranges on a relic and on its two-level forms: an int32
    column of 2^log2n rows over 10,000 values (uniform), nq ranges
    lo <= x < hi with lo uniform and width log-uniform in [1, 10000].  The
    relic batch runs the range kernel (each query ORs its bitmap range or
    the complement); the fuzz and zona batches run each query's option
    program (one or two coarse bitmaps and the fine bitmaps at the edges)
    in the expression kernel.  All three are autotuned graph replays with
    the same counts, checked against the scan plan."""
    card = 10000
    rng = np.random.default_rng(9101)
    col = rng.integers(0, card, size=1 << log2n).astype(np.int32)
    lo, hi = relic_ranges(9102, nq, card)
    res_bytes = 4.0 * ((1 << log2n) // 31 + 1)
    out = dict(rows=1 << log2n, values=card, queries=nq)
    ref = None
    for name, make in (("relic", lambda: L.Index.relic(col)), ("fuzz", lambda: L.Index.fuzz(col)),
                       ("zona", lambda: L.Index.zona(col))):
        t = time.time()
        ix = make()
        L.lib().fbh_sync()
        build_s = time.time() - t
        b = L.RangeBatch(ix, lo, 2, hi, 1, plan=1)
        cfg = b.tune()
        b.run()
        ms = []
        for _ in range(steps):
            h, m = b.run()
            ms.append(m)
        if ref is None:
            ref, _ = L.RangeBatch(ix, lo, 2, hi, 1, plan=3).run()
        if not np.array_equal(h, ref):
            raise RuntimeError(f"{name} counts differ from the scan plan's")
        qb = sum(ix.query_bytes(float(a), 2, float(c), 1, 1) for a, c in zip(lo, hi)) - res_bytes * nq
        gm = float(np.median(ms))
        r = dict(qps=round(nq / (gm * 1e-3), 1), ms_per_batch=round(gm, 3), build_s=round(build_s, 2),
                 algo_GB=round(qb / 1e9, 3), GBs=round(qb / (gm * 1e-3) / 1e9, 1), tuned=cfg)
        if name != "relic":
            enc, cb, coff, nb = ix.coarse()
            opts = [ix.two_level_option(float(a), 2, float(c), 1)[0] for a, c in zip(lo, hi)]
            r.update(coarse_bins=int(cb.size - 1), coarse_bitmaps=int(nb),
                     options={str(k): int(v) for k, v in zip(*np.unique(opts, return_counts=True))})
        out[name] = r
        log("two-level", name, r)
        del b, ix
    out["hits_total"] = int(ref.sum())
    return out


def learned_section(L, ix, train, test, epochs):
    """The learned plan choice on the same index: every query of a training
    set and of a test set timed under the three plans by the engine, the
    MLP trained on the device, then regret (label times) and end-to-end
    batch throughput of the learned plans against the ibis heuristic, each
    fixed plan and the per-query best."""
    N = L.native
    (lo_tr, hi_tr), (lo_te, hi_te) = train, test
    t = time.time()
    ms_tr, _ = ix.time_plans(lo_tr, 2, hi_tr, 1, cap=2.0)
    label_s = time.time() - t
    log(f"learned: timed {len(lo_tr)} training queries x 3 plans in {label_s:.1f}s")
    x_tr = ix.plan_features(lo_tr, 2, hi_tr, 1)
    y, m = N.plan_labels(ms_tr)
    P = N.Planner(seed=2024)
    t = time.time()
    hist = P.train(x_tr, y, m, epochs=epochs, batch=256, lr=2e-3, seed=7)
    train_s = time.time() - t
    log(f"learned: trained {epochs} epochs in {train_s:.2f}s, loss {hist[0]:.3f} -> {hist[-1]:.3f}")

    ms, hits = ix.time_plans(lo_te, 2, hi_te, 1, cap=2.0)
    nq = len(lo_te)
    done = ms[:, 1] > 0
    identical = bool(np.array_equal(hits[:, 0], hits[:, 2]) and np.array_equal(hits[done, 1], hits[done, 0]))
    if not identical:
        raise RuntimeError("plans disagree on the test queries' counts")
    t = time.perf_counter()
    learned, nfall = ix.choose_plans(P, lo_te, 2, hi_te, 1)
    choose_ms = (time.perf_counter() - t) * 1e3
    heur = np.array([ix.estimate_cost(float(a), 2, float(b), 1)[2] for a, b in zip(lo_te, hi_te)], np.int32)
    T = np.abs(ms.astype(np.float64))          # censored: the lower bound
    best = np.argmin(T, axis=1).astype(np.int32) + 1
    rows = np.arange(nq)

    def label_total(plans):
        return float(T[rows, plans - 1].sum())
    tb = label_total(best)
    lab = {"learned": label_total(learned), "heuristic": label_total(heur), "best": tb,
           "literal": label_total(np.full(nq, 1)), "compressed_lower_bound": label_total(np.full(nq, 2)),
           "scan": label_total(np.full(nq, 3))}
    censored_learned = int(np.sum(ms[rows, learned - 1] < 0))

    # end-to-end batches (warm second run); the learned wall includes the
    # feature extraction and inference
    def batch(plans=None, plan=0):
        ix.evaluate_batch(lo_te, 2, hi_te, 1, plan=plan, plans=plans)
        t0 = time.perf_counter()
        h, _, _ = ix.evaluate_batch(lo_te, 2, hi_te, 1, plan=plan, plans=plans)
        return h, time.perf_counter() - t0
    h_ref, _ = batch(plan=1)
    wall = {}
    for name, kw in (("literal", dict(plan=1)), ("scan", dict(plan=3)), ("heuristic", dict(plan=0)),
                     ("best", dict(plans=best)), ("learned", dict(plans=learned))):
        h, w = batch(**kw)
        if not np.array_equal(h, h_ref):
            raise RuntimeError(f"{name} plans give other counts")
        wall[name] = w
    wall["learned"] += choose_ms * 1e-3
    qps = {k: round(nq / v, 1) for k, v in wall.items()}
    fixed_best = max(qps["literal"], qps["scan"])
    out = dict(
        train_queries=len(lo_tr), test_queries=nq, label_timing_s=round(label_s, 1),
        epochs=epochs, train_s=round(train_s, 2), loss_first=round(float(hist[0]), 4),
        loss_last=round(float(hist[-1]), 4), choose_ms=round(choose_ms, 2), fallbacks=int(nfall),
        plans_learned={N.PLAN_NAMES[k]: int(np.sum(learned == k)) for k in (1, 2, 3)},
        plans_heuristic={N.PLAN_NAMES[k]: int(np.sum(heur == k)) for k in (1, 2, 3)},
        plans_best={N.PLAN_NAMES[k]: int(np.sum(best == k)) for k in (1, 2, 3)},
        censored_compressed=int(np.sum(~done)), learned_picks_censored=censored_learned,
        label_ms=({k: round(v, 2) for k, v in lab.items()}),
        regret_learned=round(lab["learned"] / tb - 1, 4),
        regret_heuristic=round(lab["heuristic"] / tb - 1, 4),
        batch_qps=qps,
        learned_vs_heuristic=round(qps["learned"] / qps["heuristic"], 3),
        learned_vs_best_fixed=round(qps["learned"] / fixed_best, 3),
        learned_vs_scan_only=round(qps["learned"] / qps["scan"], 3),
        counts_identical_all_plans=identical)
    log("learned", out)
    return out


def host_cpu():
    """(CPU model, host threads the benchmark may use)."""
    model = "unknown"
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    n = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else (os.cpu_count() or 1)
    return model, n


_RW = None


def _ref_worker_init(d, kind):
    """A host worker of the CPU baseline: the reference library with the
    GPU-built index(es) attached from the arrays in directory d."""
    global _RW
    import oracle
    if kind == "range":
        a = {k: np.load(os.path.join(d, k + ".npy")) for k in ("col", "keys", "offs", "bms")}
        _RW = oracle.RefAttached("cpu_range", a["col"], a["keys"], a["offs"], a["bms"])
        _RW.count_range(float(a["col"][0]), 2, float(a["col"][0]) + 1.0, 1)     # warm the caches
    else:
        _RW = []
        for i in range(4):
            a = {k: np.load(os.path.join(d, f"{k}{i}.npy")) for k in ("col", "keys", "offs", "bms")}
            _RW.append(oracle.RefAttached(f"cpu_t{i}", a["col"], a["keys"], a["offs"], a["bms"]))


def _ref_worker_ranges(q):
    lo, hi = q
    t = time.perf_counter()
    c = [_RW.count_range(float(a), 2, float(b), 1) for a, b in zip(lo, hi)]
    return c, time.perf_counter() - t


def _ref_worker_conj(qs):
    import oracle
    t = time.perf_counter()
    c = []
    for q in qs:
        terms = [(f"cpu_t{ci}", a, 2, b, 1) for (ci, a, b) in q]
        c.append(oracle.ref_iapi_expr(terms, [0] + [x for k in range(1, len(q)) for x in (k, -1)]))
    return c, time.perf_counter() - t


def _spawn_pool(arrays, kind, workers):
    """A pool of spawned host processes, each attaching the arrays (written
    to a temporary directory) to the reference library."""
    import multiprocessing as mp
    import tempfile
    d = tempfile.mkdtemp(prefix="fbh_cpu_")
    for k, v in arrays.items():
        np.save(os.path.join(d, k + ".npy"), v)
    pool = mp.get_context("spawn").Pool(workers, initializer=_ref_worker_init, initargs=(d, kind))
    return pool, d


def cpu_baseline(ix, col, lo, hi, hits_gpu, sample, workers):
    """The reference library answering `sample` of the same ranges, each as
    one two-sided range (a qContinuousRange through
    fastbit_selection_evaluate) with the GPU-built index attached, the
    queries spread over host worker processes (one thread each)."""
    import oracle
    if oracle.reference() is None:
        return None
    import shutil
    keys, offs, bms = ix.deconstruct()
    pool, d = _spawn_pool(dict(col=col, keys=keys, offs=offs, bms=bms), "range", workers)
    try:
        parts = np.array_split(np.arange(sample), workers)
        t = time.perf_counter()
        res = pool.map(_ref_worker_ranges, [(lo[p], hi[p]) for p in parts])
        dt = time.perf_counter() - t
    finally:
        pool.close()
        pool.join()
        shutil.rmtree(d, ignore_errors=True)
    counts = [c for r in res for c in r[0]]
    if not np.array_equal(np.array(counts, np.uint64), hits_gpu[:sample]):
        raise RuntimeError("reference counts differ from the GPU's")
    model, _ = host_cpu()
    busy = sum(r[1] for r in res)
    return dict(value=round(sample / dt, 3), unit="queries/s", cores=workers, kind="reference",
                cpu=model, per_thread_qps=round(sample / busy, 4),
                sample=f"{sample} of the 1024 ranges lo <= x < hi, each one qContinuousRange through "
                       f"fastbit_selection_evaluate on the same 2^27-row column with the GPU-built index "
                       f"attached (fastbit_iapi_attach_full_index), spread over {workers} worker processes; "
                       f"counts checked equal to the GPU's; {dt:.1f}s wall")


def cpu_baseline_wah(cells):
    """The reference's in-place operator plus compress() (one host thread) on
    the same 2^28-bit Markov operands as the GPU's bandwidth cells."""
    import oracle
    R = oracle.reference()
    if R is None:
        return None
    from libfastlib_amd import datagen as G
    model, _ = host_cpu()
    out = []
    for d, run in cells:
        wa = G.markov_wah(100, 1 << 28, d, run)
        wb = G.markov_wah(101, 1 << 28, d, run)
        for op, name in ((0, "AND"), (1, "OR")):
            ms, nw = R.time_op(wa, wb, op, reps=2)
            byts = 4.0 * (len(wa) + len(wb) + nw)
            out.append(dict(cell=f"d={d:g},L={run}", op=name, ms=round(ms, 3), GBs=round(byts / (ms * 1e-3) / 1e9, 2)))
            log("cpu wah", out[-1])
    return dict(unit="ms per op (in place + compress)", cores=1, kind="reference", cpu=model, cells=out)


def cpu_baseline_conj(ixs, cols, qs, hits_gpu, sample, workers):
    """The reference answering `sample` of the conjunctions (terms ANDed by
    fastbit_selection_combine, counted by fastbit_selection_evaluate) with
    the four GPU-built indexes attached, over host worker processes."""
    import oracle
    if oracle.reference() is None:
        return None
    import shutil
    arrays = {}
    for i, (ix, c) in enumerate(zip(ixs, cols)):
        k, o, b = ix.deconstruct()
        arrays.update({f"col{i}": c, f"keys{i}": k, f"offs{i}": o, f"bms{i}": b})
    pool, d = _spawn_pool(arrays, "conj", workers)
    try:
        pick = np.arange(0, len(qs), max(1, len(qs) // sample))[:sample]
        parts = np.array_split(pick, workers)
        t = time.perf_counter()
        res = pool.map(_ref_worker_conj, [[qs[i] for i in p] for p in parts])
        dt = time.perf_counter() - t
    finally:
        pool.close()
        pool.join()
        shutil.rmtree(d, ignore_errors=True)
    counts = np.array([c for r in res for c in r[0]], np.uint64)
    if not np.array_equal(counts, hits_gpu[pick]):
        raise RuntimeError("reference conjunction counts differ from the GPU's")
    model, _ = host_cpu()
    return dict(value=round(len(pick) / dt, 3), unit="queries/s", cores=workers, kind="reference", cpu=model,
                sample=f"{len(pick)} of the 4096 conjunctions, the four GPU-built indexes attached, spread over "
                       f"{workers} worker processes; counts checked equal to the GPU's; {dt:.1f}s wall")


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--log2-rows", type=int, default=27)
    ap.add_argument("--queries", type=int, default=1024)
    ap.add_argument("--cpu-sample", type=int, default=64)
    ap.add_argument("--cpu-workers", type=int, default=0, help="host processes of the CPU baseline "
                    "(default: the host threads available, at most 16)")
    ap.add_argument("--no-wah", action="store_true")
    ap.add_argument("--wah-only", type=int, default=0, metavar="REPS",
                    help="time only the WAH AND/OR cells (REPS ops each) and print them as JSON")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-conj", action="store_true")
    ap.add_argument("--conj-only", action="store_true",
                    help="run only the conjunction / expression batches (no CPU baseline) and print them as JSON")
    ap.add_argument("--no-learned", action="store_true")
    ap.add_argument("--no-two-level", action="store_true")
    ap.add_argument("--two-level-only", action="store_true",
                    help="only the relic / fuzz / zona comparison at 2^26 rows")
    ap.add_argument("--train-queries", type=int, default=50000)
    ap.add_argument("--test-queries", type=int, default=8192)
    ap.add_argument("--epochs", type=int, default=40)
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this benchmark runs on one GPU; sharding over several GPUs "
                                   "is designed but not built", "n_gpus": args.gpus}))
        sys.exit(2)

    import libfastlib_amd as L
    from libfastlib_amd import datagen as G
    L.lib()
    if args.wah_only:
        print(json.dumps({"wah_ops_2p28": wah_cells(L, reps=args.wah_only)}))
        return
    if args.two_level_only:
        print(json.dumps({"two_level_2p26": two_level_section(L, 26, 1024, args.steps)}))
        return
    if args.conj_only:
        conj, _ = conj_section(L, 26, 4096, args.steps)
        print(json.dumps({"conjunctions_4x2p26": conj}))
        return
    extra = () if args.no_learned else ((G.BIN_F32_TRAIN_SEED, args.train_queries),
                                        (G.BIN_F32_TEST_SEED, args.test_queries))
    col, lo, hi, more = config3(args.log2_rows, args.queries, extra=extra)
    t = time.time()
    ix = L.Index.bin_equal_weight(col, 10000)     # <binning nbins=10000 equal-weight/>
    L.lib().fbh_sync()
    build_s = time.time() - t
    nobs, nrows, nwords = ix.info()
    log(f"index: {nobs} bins, {nwords} words, GPU build {build_s:.2f}s")

    # the headline: the ranges as one persistent, graph-captured launch
    # (fbh_index_batch_prepare) under the ibis heuristic's plans, autotuned
    # once before the timed steps; a step is one replay (counts to host)
    rb = L.RangeBatch(ix, lo, 2, hi, 1, plan=0)
    tuned = rb.tune()
    log(f"range batch tuned: {tuned}")
    for _ in range(args.warmup):
        rb.run()
    walls, gms = [], []
    hits = None
    for _ in range(args.steps):
        t = time.perf_counter()
        hits, gm = rb.run()
        walls.append(time.perf_counter() - t)
        gms.append(gm)
    wall = float(np.sum(walls))
    qps = args.queries * args.steps / wall
    used = rb.plans()
    plan_mix = {L.native.PLAN_NAMES[k]: int(np.sum(used == k)) for k in (1, 2, 3)}
    log(f"heuristic plans, one launch: {qps:.1f} q/s, mix {plan_mix}")
    # the per-query path (a launch per query over three streams) gives the
    # same counts
    h_pq, _, gm_pq = ix.evaluate_batch(lo, 2, hi, 1, 0)
    if not np.array_equal(h_pq, hits):
        raise RuntimeError("the per-query path's counts differ from the batch's")

    # algorithmic bytes of a counting query: the bitmap words its plan reads
    # (OR set and edge bitmaps) and the raw values it reads (edge rows, or
    # the whole column for the scan); the batch writes no literal result
    groups = nrows // 31
    res_bytes = 4.0 * (groups + 1)
    bl = query_bytes(ix, lo, hi, 1) - res_bytes
    bs = query_bytes(ix, lo, hi, 3) - res_bytes
    per_plan = {}
    for plan, qb in ((1, bl), (3, bs)):
        pb = L.RangeBatch(ix, lo, 2, hi, 1, plan=plan)
        pb.tune()
        h, _ = pb.run()
        if not np.array_equal(h, hits):
            raise RuntimeError(f"plan {plan} counts differ from the heuristic plan's")
        gm = float(np.median([pb.run()[1] for _ in range(3)]))
        per_plan[L.native.PLAN_NAMES[plan]] = dict(
            qps=round(len(lo) / (gm * 1e-3), 1), gpu_ms=round(gm, 3),
            algo_GB=round(qb.sum() / 1e9, 3), GBs=round(qb.sum() / (gm * 1e-3) / 1e9, 1))
        log("plan", L.native.PLAN_NAMES[plan], per_plan[L.native.PLAN_NAMES[plan]])
        del pb
    per_query_launches = dict(qps=round(len(lo) / (gm_pq * 1e-3), 1), gpu_ms=round(gm_pq, 3),
                              note="fbh_index_evaluate_batch: one launch per query rotating over three "
                                   "streams, literal results written")
    # roofline of the kernel of the timed step: k_range, one launch per
    # step, its bytes those of each query's chosen plan, its time the HIP
    # events around the replay on the library stream (the replay is k_range
    # and the few-microsecond k_sum_parts)
    qb_auto = np.where(used == 3, bs, bl)
    gm_med = float(np.median(gms))
    achieved = round(qb_auto.sum() / (gm_med * 1e-3) / 1e9, 1)
    roofline = dict(bound="hbm", kernel="k_range", achieved=achieved, peak=HBM_PEAK_GBS, unit="GB/s",
                    frac=round(achieved / HBM_PEAK_GBS, 4), traffic=measured_traffic("k_range"),
                    avg_launch_us=round(gm_med * 1e3, 1), algo_bytes_per_launch=round(float(qb_auto.sum()), 0))

    # row ids (countQuery::getHitRows): every 64th range through the
    # literal plan, ids copied to host memory (PCIe included)
    rid_q = np.arange(0, len(lo), 64)
    ix.rids(float(lo[0]), 2, float(hi[0]), 1)
    t = time.perf_counter()
    nr = sum(len(ix.rids(float(lo[i]), 2, float(hi[i]), 1)) for i in rid_q)
    dt = time.perf_counter() - t
    rid_output = dict(queries=int(len(rid_q)), rids=int(nr), rids_per_s=round(nr / dt, 0),
                      GBs=round(4.0 * nr / dt / 1e9, 2),
                      note="fbh_index_evaluate_rids: literal plan, row-id kernel, ids copied to host "
                           "memory (PCIe-inclusive wall time)")
    log("rids", rid_output)

    _, ncpu = host_cpu()
    workers = max(1, min(args.cpu_workers or ncpu, 16))
    learned = None if args.no_learned else learned_section(L, ix, more[0], more[1], args.epochs)
    wah = None if args.no_wah else wah_cells(L)
    conj, relic_learned = (None, None) if args.no_conj else conj_section(
        L, 26, 4096, 5, None if args.no_learned else (args.train_queries, args.test_queries, args.epochs),
        cpu=None if args.no_cpu else (32, workers))
    two_level = None if args.no_two_level else two_level_section(L, 26, 1024, 5)
    cpu = None if args.no_cpu else cpu_baseline(ix, col, lo, hi, hits, args.cpu_sample, workers)
    cpu_wah = None if (args.no_cpu or args.no_wah) else cpu_baseline_wah(
        [(0.5, 1), (0.5, 64), (0.01, 1)])

    out = {
        "metric": "range queries/s at 2^27 rows; WAH op GB/s (% HBM peak); learned vs fixed plan",
        "value": round(qps, 2),
        "unit": "queries/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(1e3 * wall / args.steps, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u32",
        "data": "synthetic (deterministic generators, seeds in bench.py)",
        "config": {"workload": f"binned float index: 2^{args.log2_rows} float32 rows, 10000 equal-weight "
                               f"bins ({nobs} bitmaps, {nwords} words), {args.queries} quantile range "
                               f"queries lo <= x < hi, plan chosen per query by the ibis heuristic",
                   "rows": int(nrows), "bins": int(nobs), "queries": args.queries,
                   "gpu_index_build_s": round(build_s, 2)},
        "plan_mix": plan_mix,
        "tuned": tuned,
        "per_plan": per_plan,
        "per_query_launches": per_query_launches,
        "rid_output": rid_output,
        "best_plan_bytes_model_GB": round(float(np.minimum(bl, bs).sum()) / 1e9, 3),
        "roofline": roofline,
        "learned_plan": learned,
        "learned_plan_relic_2p26": relic_learned,
        "wah_ops_2p28": wah,
        "conjunctions_4x2p26": conj,
        "two_level_2p26": two_level,
        "cpu_baseline": cpu,
        "cpu_baseline_wah": cpu_wah,
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
