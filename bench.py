#!/usr/bin/env python3
"""Benchmark: the fused bitmap-index query path on one MI355X.

Workload (a "step"): the 12 star-schema queries of config 2 over a 120M-row
fact table, one fused launch per query (bitmap program -> compaction ->
group-by into int64), inputs and bitmaps resident in HBM.  Prints ONE JSON
line: fact rows/s over the 12 queries, the roofline of the fused kernel
(algorithmic bytes / measured launch time vs 8 TB/s), and the CPU oracle
timed on the host cores on a bounded sample.  Progress goes to stderr.

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W]
"""
import argparse
import ctypes as C
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = ("fact rows/sec (G rows/s) + % HBM roofline; batched vs plain; "
          "p99 query latency")
HBM_PEAK_GBS = 8000.0


def log(*a):
    print("[bench]", *a, file=sys.stderr, flush=True)


PROFILE = os.path.join(ROOT, "profiles", "fused_query_config2.md")


def profiled_traffic():
    """Calibrated HBM bytes (FETCH_SIZE x2 + WRITE_SIZE x1, rocprofv3 --pmc)
    per fused launch of this workload, averaged over the 12 queries, as
    committed with the profile; None if the profile is absent."""
    try:
        with open(PROFILE) as f:
            for line in f:
                if line.startswith("traffic_bytes_per_launch:"):
                    return int(float(line.split(":")[1]))
    except OSError:
        pass
    return None


def copy_calibration(nbytes, reps):
    """Plain device copy of nbytes, reps times (for byte-counter scaling)."""
    from merle_amd.host import Device, Timer, check
    dev = Device(0)
    a, b = dev.alloc(nbytes), dev.alloc(nbytes)
    check(dev.lib.mbx_fill_u32(dev.ctx, a.ptr, 7, nbytes // 4))
    t = Timer(dev)
    t.start()
    for _ in range(reps):
        check(dev.lib.mbx_copy_kernel(dev.ctx, b.ptr, a.ptr, nbytes))
    ms = t.stop()
    print(json.dumps({"copy_bytes": nbytes, "reps": reps,
                      "GB_per_s": round(2 * nbytes * reps / ms / 1e6, 1)}))
    dev.close()
    return 0


def _cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return None


def cpu_baseline(dev, sch, specs, target_s=5.0, reps=3):
    """The oracle (raw-column scan, OpenMP) on the host cores, on the first
    R fact rows, R sized from a short probe to ~target_s of CPU work per
    pass; the median of `reps` passes on all threads, then one pass on one
    thread over a sample sized the same way."""
    from merle_amd import spec as S
    from oracle import cpuref
    threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    n = sch.nrows
    host = S.Schema(n)
    fact = [k for k in sch.cols if k.startswith("lo_")]
    for k, c in sch.cols.items():
        if k not in fact:
            host.add(k, c.host if c.host is not None else c.dev.host())
    probe = min(n, 4_000_000)
    for k in fact:
        host.add(k, sch[k].dev.host(probe))
    host.nrows = probe

    def rows_for(nthreads):
        t0 = time.perf_counter()
        for s in specs:
            cpuref.query(s, host, nthreads=nthreads, row0=0, row1=min(probe, 1_000_000))
        per_row = (time.perf_counter() - t0) / (min(probe, 1_000_000) * len(specs))
        r = int(min(n, max(64, target_s / (per_row * len(specs)))))
        return r - r % 64

    def run(nthreads, r):
        res = []
        t0 = time.perf_counter()
        for s in specs:
            out, _ = cpuref.query(s, host, nthreads=nthreads, row0=0, row1=r)
            res.append(out)
        return time.perf_counter() - t0, res

    R = max(probe, rows_for(threads))
    if R > probe:
        for k in fact:
            host.cols[k].host = sch[k].dev.host(R)
            host.cols[k].n = R
    host.nrows = R
    passes = [run(threads, R) for _ in range(reps)]
    dt = float(np.median([p[0] for p in passes]))
    results = passes[0][1]
    R1 = min(R, rows_for(1))
    dt1, _ = run(1, R1)
    rate = R * len(specs) / dt / 1e9
    return dict(value=round(rate, 4), unit="G rows/s", cores=threads,
                kind="port",
                sample=f"the 12 config-2 queries over the first {R:,} of "
                       f"{n:,} fact rows (oracle raw-column scan, "
                       f"{threads} OpenMP threads, median of {reps} passes "
                       f"of {dt:.1f} s)",
                one_core={"value": round(R1 * len(specs) / dt1 / 1e9, 4),
                          "unit": "G rows/s",
                          "sample": f"the same queries over the first "
                                    f"{R1:,} rows, one thread, {dt1:.1f} s"},
                cpu_model=_cpu_model(), nproc=os.cpu_count()), R, results


def run_batched(dev, reps=5, train_programs=1024):
    """Config 3: 256 queries over the config-1 tables.  Plain = one fused
    launch per query, back to back on one stream (the reference's
    one-query-per-launch driver); batched = one persistent launch of all 256,
    replayed from a hipGraph, launch shape from the autotuner.  Every query's
    int64 output must be identical."""
    from merle_amd import abi, configs
    from merle_amd.host import Timer, check
    n = configs.C1_ROWS
    sch = configs.config1_device(dev, n)
    bins, slabs = configs.build_bins(dev, sch, {c: configs.C1_BOUNDS for c in configs.C1_COLS})
    specs = configs.config3_queries(256)
    qs = configs.query_array(sch, specs, bins, configs.C1_BOUNDS)
    offs = np.cumsum([0] + [q.nr_grp for q in specs])
    plain_out = dev.alloc(8 * int(offs[-1]), np.int64, int(offs[-1]))
    batch_out = dev.alloc(8 * int(offs[-1]), np.int64, int(offs[-1]))
    for i in range(len(specs)):
        qs[i].d_out = batch_out.ptr + 8 * int(offs[i])
    t = Timer(dev)

    pq = configs.heuristic_queries(dev, qs)

    def plain(arr):
        check(dev.lib.mbx_fill_u32(dev.ctx, plain_out.ptr, 0, 2 * int(offs[-1])))
        for i in range(len(specs)):
            check(dev.lib.mbx_prog_grpby(dev.ctx, C.byref(arr[i].prg),
                                         C.byref(arr[i].op), arr[i].nr_grp,
                                         plain_out.ptr + 8 * int(offs[i]),
                                         abi.PLAN_FUSED | abi.ACCUMULATE |
                                         arr[i].flags))
    tph = []
    plain(pq)
    for _ in range(reps):
        t.start()
        plain(pq)
        tph.append(t.stop())
    planned_same = plain_out.host()
    plain(qs)
    tp = []
    for _ in range(reps):
        t.start()
        plain(qs)
        tp.append(t.stop())
    written = qs
    qs = pq          # the batch runs the planned queries
    tune, tms = abi.Tune(), C.c_float()
    check(dev.lib.mbx_autotune(dev.ctx, qs, len(specs), C.byref(tune), C.byref(tms)))

    def batched(arr):
        b = C.c_void_p()
        check(dev.lib.mbx_batch_create(dev.ctx, arr, len(specs), C.byref(tune),
                                       C.byref(b)))
        check(dev.lib.mbx_batch_run(dev.ctx, b))
        ts = []
        for _ in range(reps):
            t.start()
            check(dev.lib.mbx_batch_run(dev.ctx, b))
            ts.append(t.stop())
        check(dev.lib.mbx_batch_destroy(dev.ctx, b))
        return float(np.median(ts)), batch_out.host()

    mb, got = batched(qs)
    ref = plain_out.host()
    same = bool(np.array_equal(ref, got) and np.array_equal(ref, planned_same))

    # the same batch under learned plans: a cost model trained on runtimes
    # the engine measures for other random queries over the same tables
    # picks each query's column order among its fused candidates (the
    # batched launch is fused-only); the evaluation form stays the batch's
    # own, eager, so conjunctions run in its conjunction kernel
    from merle_amd import learned
    t0 = time.time()
    train = configs.config3_queries(train_programs, seed=configs.SEED + 33)
    X, Y = learned.collect(dev, sch, bins, 0, specs=train, bounds=configs.C1_BOUNDS)
    model = learned.CostModel(dev)
    model.train(X, Y)
    lq = (abi.Query * len(specs))()
    C.memmove(lq, written, C.sizeof(lq))
    for i, s in enumerate(specs):
        plan = model.choose_fused(learned.Prepared(dev, sch, bins, s, configs.C1_BOUNDS))
        applied = abi.Vprg()
        check(dev.lib.mbx_plan_apply(C.byref(lq[i].prg), C.byref(plan), C.byref(applied)))
        lq[i].prg = applied
        lq[i].flags = abi.PLAN_EAGER
    model.close()
    learn_s = time.time() - t0
    ml, got = batched(lq)
    same = same and bool(np.array_equal(ref, got))
    nbytes, nsel = C.c_uint64(), C.c_uint64()
    check(dev.lib.mbx_alg_bytes_batch(dev.ctx, qs, len(specs), C.byref(nbytes),
                                      C.byref(nsel)))
    # the same bytes counted per query (a line two queries read counts twice):
    # what the launch reads, most of it served from L2 / Infinity Cache
    per_query = 0
    for i in range(len(specs)):
        qb, qk = C.c_uint64(), C.c_uint64()
        check(dev.lib.mbx_alg_bytes(dev.ctx, C.byref(qs[i].prg), C.byref(qs[i].op),
                                    qs[i].nr_grp, C.byref(qb), C.byref(qk)))
        per_query += qb.value
    t.close()
    mp, mh = float(np.median(tp)), float(np.median(tph))
    for a in [plain_out, batch_out] + slabs + [c.dev for c in sch.cols.values()]:
        a.free()
    rows = n * len(specs)
    return {
        "workload": f"{len(specs)} queries (AND/OR/ANDNOT, selectivity 1e-6..1e-1, "
                    f"1..10,000 groups) over {n:,} rows x 3 Zipf columns",
        "plain_ms": round(mp, 3), "plain_heuristic_plan_ms": round(mh, 3),
        "batched_ms": round(mb, 3),
        "plain_G_rows_per_s": round(rows / mp / 1e6, 2),
        "batched_G_rows_per_s": round(rows / mb / 1e6, 2),
        "batched_vs_plain": round(mp / mb, 3),
        "batched_learned_ms": round(ml, 3),
        "batched_learned_vs_plain": round(mp / ml, 3),
        "learned_training": f"{len(train)} other random queries over the same "
                            f"tables, {X.shape[0]} timed plans, {learn_s:.1f} s",
        "alg_GB_per_s_batched": round(nbytes.value / mb / 1e6, 1),
        "roofline_frac_batched": round(nbytes.value / mb / 1e6 / HBM_PEAK_GBS, 4),
        "per_query_alg_GB_per_s_batched": round(per_query / mb / 1e6, 1),
        "tune": {"tile_words": tune.tile_words, "wg_per_cu": tune.wg_per_cu,
                 "lds_groups": tune.lds_groups},
        "identical_to_plain": same,
        "note": "plain = as-written programs, one fused launch each; batched = "
                "the same queries with the static heuristic plan in one "
                "persistent hipGraph launch; batched_learned = the same launch "
                "with the learned chooser's column orders (eager form)",
    }


def learned_queries(dev, sch, bins, bounds, qs, specs, train_specs):
    """A copy of a query array with each request's plan picked by a cost
    model trained on the engine-timed candidates of other requests over the
    same tables (`train_specs`).  -> (queries, training note, plan kinds)."""
    from merle_amd import abi, learned
    from merle_amd.host import check
    t0 = time.time()
    X, Y = learned.collect(dev, sch, bins, 0, specs=train_specs, bounds=bounds)
    model = learned.CostModel(dev)
    model.train(X, Y)
    out = (abi.Query * len(qs))()
    C.memmove(out, qs, C.sizeof(out))
    kinds = {"fused": 0, "eager": 0, "scan": 0, "materialise": 0}
    for i, s in enumerate(specs):
        p = learned.Prepared(dev, sch, bins, s, bounds)
        plan = model.choose(p)
        applied = abi.Vprg()
        check(dev.lib.mbx_plan_apply(C.byref(p.prg), C.byref(plan), C.byref(applied)))
        out[i].prg = applied
        out[i].flags = plan.flags & (abi.PLAN_MASK | abi.PLAN_EAGER)
        k = plan.flags & abi.PLAN_MASK
        kinds["scan" if k == abi.PLAN_SCAN else "materialise"
              if k == abi.PLAN_MATERIALISE else "eager"
              if plan.flags & abi.PLAN_EAGER else "fused"] += 1
    model.close()
    note = (f"{len(train_specs)} other requests over the same tables, "
            f"{X.shape[0]} timed plans, {time.time() - t0:.1f} s")
    return out, note, kinds


def run_serving_alone(dev, nreq=1000, reps=3):
    """The config-5 requests one by one, each alone on the device under its
    static heuristic plan (best of `reps` launches between an event pair):
    the service times the serving replays queue behind, and the slowest
    requests -- those whose edge bin covers a Zipf head -- with their
    algorithmic bytes."""
    from merle_amd import abi, configs
    from merle_amd.host import Timer, check
    sch = configs.config5_device(dev)
    bins, slabs = configs.build_bins(dev, sch, configs.C5_BOUNDS)
    specs = configs.config5_queries(nreq)
    hq = configs.heuristic_queries(
        dev, configs.query_array(sch, specs, bins, configs.C5_BOUNDS))
    out = dev.alloc(8 * 10_000, np.int64, 10_000)
    t = Timer(dev)
    ms = np.zeros(nreq)
    for i in range(nreq):
        q = hq[i]

        def run():
            check(dev.lib.mbx_prog_grpby(dev.ctx, C.byref(q.prg), C.byref(q.op),
                                         q.nr_grp, out.ptr,
                                         abi.PLAN_FUSED | q.flags))
        run()
        best = np.inf
        for _ in range(reps):
            t.start()
            run()
            best = min(best, t.stop())
        ms[i] = best
    slow = []
    for i in np.argsort(-ms)[:5]:
        q = hq[int(i)]
        b, k = C.c_uint64(), C.c_uint64()
        check(dev.lib.mbx_alg_bytes(dev.ctx, C.byref(q.prg), C.byref(q.op),
                                    q.nr_grp, C.byref(b), C.byref(k)))
        slow.append({"request": int(i), "ms": round(float(ms[i]), 3),
                     "alg_GB": round(b.value / 1e9, 3), "selected": k.value,
                     "alg_TB_per_s": round(b.value / ms[i] / 1e9, 2)})
    t.close()
    out.free()
    for a in slabs + [c.dev for c in sch.cols.values()]:
        a.free()
    return {"workload": f"{nreq} config-5 requests over {sch.nrows:,} rows, each "
                        "alone under its heuristic plan",
            "p50_ms": round(float(np.percentile(ms, 50)), 3),
            "p99_ms": round(float(np.percentile(ms, 99)), 3),
            "max_ms": round(float(ms.max()), 3),
            "mean_ms": round(float(ms.mean()), 4), "slowest": slow}


def run_serving(dev, nreq=1000, rates=(2000.0, 20000.0), budget=256 << 20,
                tight=256 << 10, train_requests=600):
    """Config 5: 600M rows x 10 Zipf columns + 640 bucket bitmaps resident;
    Poisson arrivals replayed against the server at a moderate and at a
    saturating rate, once with continuous batching (4 streams; up to 4 small
    requests per dispatch, large and scan requests alone as launches of
    their own) and once one request at a time, under the plans a cost model
    trained on other requests picks and under the static heuristic plans;
    then once more at the
    saturating rate with a budget of the server's fixed memory plus `tight`
    bytes, so admission binds.  Every request's result must be identical in
    every replay."""
    from merle_amd import abi, configs
    from merle_amd.host import check
    n = configs.C5_ROWS
    t0 = time.time()
    sch = configs.config5_device(dev, n)
    bins, slabs = configs.build_bins(dev, sch, configs.C5_BOUNDS)
    log(f"config-5 tables and {sum(len(v) for v in bins.values())} bitmaps "
        f"ready in {time.time() - t0:.1f} s")
    specs = configs.config5_queries(nreq)
    written = configs.query_array(sch, specs, bins, configs.C5_BOUNDS)
    hq = configs.heuristic_queries(dev, written)
    lq, lnote, lkinds = learned_queries(
        dev, sch, bins, configs.C5_BOUNDS, written, specs,
        configs.config5_training_queries(train_requests))
    log(f"serving plans learned: {lkinds}")
    tot = sum(q.nr_grp for q in specs)
    # algorithmic bytes from a sample of the requests (the accounting kernel
    # is slow), scaled to all of them
    ns = min(nreq, 64)
    nbytes, nsel = C.c_uint64(), C.c_uint64()
    check(dev.lib.mbx_alg_bytes_batch(dev.ctx, hq, ns, C.byref(nbytes),
                                      C.byref(nsel)))
    est = nbytes.value * nreq / ns

    def replay(qs, arrivals, cfg):
        res = np.zeros(tot, np.int64)
        lat = np.zeros(nreq)
        st = abi.ServeStats()
        check(dev.lib.mbx_serve_replay(
            dev.ctx, qs, nreq, arrivals.ctypes.data_as(C.POINTER(C.c_double)),
            C.byref(cfg), res.ctypes.data,
            lat.ctypes.data_as(C.POINTER(C.c_double)), C.byref(st)))
        gbs = est / st.wall_s / 1e9
        return res, {
            "completed": st.completed, "rejected": st.rejected,
            "wall_s": round(st.wall_s, 4),
            "G_rows_per_s": round(n * st.completed / st.wall_s / 1e9, 1),
            "p50_ms": round(st.p50_s * 1e3, 3),
            "p99_ms": round(st.p99_s * 1e3, 3),
            "mean_ms": round(st.mean_s * 1e3, 3),
            "mean_batch": round(st.mean_batch, 2),
            "admission_waits": st.admission_waits,
            "peak_mem_MB": round(st.peak_mem / 1e6, 3),
            "fixed_mem_MB": round(st.fixed_mem / 1e6, 3),
            "alg_GB_per_s": round(gbs, 1),
            "roofline_frac": round(gbs / HBM_PEAK_GBS, 4),
        }

    batched_cfg = abi.ServeCfg(4, 4, budget, 0, 0, abi.Tune(8192, 4, 4096, 0))
    plain_cfg = abi.ServeCfg(1, 1, budget, 1, 0, abi.Tune())
    out, same, ref = {}, True, None
    for rate in rates:
        rng = np.random.default_rng(5)
        arrivals = np.cumsum(rng.exponential(1.0 / rate, nreq))
        arrivals -= arrivals[0]
        block = {}
        for plans, qs in (("learned", lq), ("heuristic", hq)):
            sub = {}
            for mode, cfg in (("batched", batched_cfg),
                              ("one_at_a_time", plain_cfg)):
                res, sub[mode] = replay(qs, arrivals, cfg)
                ref = res if ref is None else ref
                same &= bool(np.array_equal(res, ref))
            sub["batched_vs_one_at_a_time"] = round(
                sub["one_at_a_time"]["wall_s"] / sub["batched"]["wall_s"], 3)
            sub["p99_batched_vs_one_at_a_time"] = round(
                sub["batched"]["p99_ms"] / sub["one_at_a_time"]["p99_ms"], 3)
            block[f"{plans}_plans"] = sub
        out[f"rate_{int(rate)}_per_s"] = block
    # admission binding: the fixed memory plus a few outputs
    fixed = C.c_uint64()
    check(dev.lib.mbx_serve_fixed_mem(C.byref(batched_cfg), C.byref(fixed)))
    tight_cfg = abi.ServeCfg(4, 4, fixed.value + tight, 0, 0,
                             abi.Tune(8192, 4, 4096, 0))
    rng = np.random.default_rng(5)
    arrivals = np.cumsum(rng.exponential(1.0 / max(rates), nreq))
    arrivals -= arrivals[0]
    res, tb = replay(lq, arrivals, tight_cfg)
    same &= bool(np.array_equal(res, ref))
    tb["budget_MB"] = round((fixed.value + tight) / 1e6, 3)
    out[f"rate_{int(max(rates))}_per_s_tight_budget_learned_plans"] = tb
    for a in slabs + [c.dev for c in sch.cols.values()]:
        a.free()
    return {
        "workload": f"{n:,} rows x 10 Zipf(1.2) columns (1,000..100,000 values), "
                    f"64 bucket bitmaps per column resident; {nreq} requests per "
                    f"replay, Poisson arrivals, memory budget {budget >> 20} MiB "
                    f"(tight: fixed + {tight >> 10} KiB)",
        "learned_plans": {"training": lnote, "kinds": lkinds},
        **out,
        "identical_results": same,
        "alg_bytes_note": f"estimated from the first {ns} requests",
    }


def run_wah(dev, sch, specs, qs, fused_us, reps=20):
    """The WAH baseline (merle/wahGpu*.cu, ssb/wah-demo.cu) on the config-2
    queries: each predicate's WAH bitmap (groups of the predicate through
    its foreign key, compressed, scanned; untimed), then the index-only step
    the reference's report times -- the predicates ANDed left to right (the
    operand with more words as the AND's down side), the result scanned and
    decompressed -- per query (median of `reps`).  Every scan writes a skip
    table beside it, which the AND and the decompression read.  A
    negated clause (d_year != 6) is its complement range.  The decompressed
    result is checked against the groups of the fused program's bitmap.
    The roofline is the decompression kernel's (the largest writer), its
    launches back to back under one event pair."""
    from merle_amd import abi, configs, spec as S
    from merle_amd.host import Timer, check
    n = sch.nrows
    ng = int(dev.lib.mbx_wah_groups_of(n))
    groups = dev.alloc(4 * ng, np.uint32, ng)
    res = [dev.alloc(4 * ng, np.uint32, ng) for _ in range(2)]
    rex = dev.alloc(4 * (ng + 1), np.uint32, ng + 1)
    nsk = int(dev.lib.mbx_wah_skip_len(ng))
    rsk = dev.alloc(4 * nsk, np.uint32, nsk)
    tmp = dev.alloc(4 * ng, np.uint32, ng)
    nw = ((n + 63) // 64) * 2
    poss = dev.alloc(4 * nw, np.uint32, nw)

    def wah_of(src, lo, hi):
        ref = S.colref(sch, src)
        check(dev.lib.mbx_wah_select(dev.ctx, C.byref(ref), n, lo, hi, groups.ptr))
        k = C.c_size_t()
        check(dev.lib.mbx_wah_compress(dev.ctx, groups.ptr, ng, tmp.ptr, C.byref(k)))
        w = dev.alloc(4 * max(1, k.value), np.uint32, k.value)
        ex = dev.alloc(4 * (k.value + 1), np.uint32, k.value + 1)
        sk = dev.alloc(4 * nsk, np.uint32, nsk)
        if k.value:
            check(dev.lib.mbx_memcpy_d2d(dev.ctx, w.ptr, tmp.ptr, 4 * k.value))
        check(dev.lib.mbx_wah_scan_skip(dev.ctx, w.ptr, k.value, ex.ptr, sk.ptr,
                                        nsk))
        lf = C.c_uint64()
        check(dev.lib.mbx_wah_longest_1fill(dev.ctx, w.ptr, k.value, C.byref(lf)))
        return w, ex, k.value, lf.value, sk

    def wah_arg(words, excl, k, skip):
        r = abi.Wah()
        r.words, r.excl, r.skip, r.n, r.nskip = words, excl, skip, k, nsk
        return r

    def ranges(spec):
        neg, loaded = set(), {}
        for op, d, c in spec.instrs:
            if op == abi.ORM:
                loaded[d] = c
            elif op == abi.NOT:
                neg.add(loaded[d])
        out = []
        for i, p in enumerate(spec.preds):
            lo, hi = p.lo, p.hi
            if i in neg:
                card = configs.C2_ATTRS[p.src.col][2]
                assert lo == 0 or hi == card, "complement is not one range"
                lo, hi = (hi, card) if lo == 0 else (0, lo)
            out.append((p.src, lo, hi))
        return out

    per_q, sizes, identical, steps = {}, {}, True, []
    t = Timer(dev)
    for s, (sp, prg, op, out) in zip(specs, qs):
        ws = [wah_of(src, lo, hi) for src, lo, hi in ranges(s)]
        sizes[s.name] = [w[2] for w in ws]
        # each AND's result keeps its down operand's words: down is the
        # operand with more words (its literals then find the up words
        # covering them staged in LDS), unless it holds a long 1-fill
        order = [(0, None)]
        acc1, acc_n = ws[0][3], ws[0][2]
        for i in range(1, len(ws)):
            cand = [x for x in ((ws[i][2], True, ws[i][3]), (acc_n, False, acc1))
                    if x[2] <= 1]
            assert cand, "no operand can be the AND's down side"
            down = max(cand)[1]
            order.append((i, down))
            acc1, acc_n = 1, (ws[i][2] if down else acc_n)

        # the step's arguments, built once: an AND's result keeps its down
        # operand's words, so their count scan and skip table serve it
        ands = []
        cur = wah_arg(ws[0][0].ptr, ws[0][1].ptr, ws[0][2], ws[0][4].ptr)
        b = 0
        for i, down in order[1:]:
            w, ex, k, _, sk = ws[i]
            mine = wah_arg(w.ptr, ex.ptr, k, sk.ptr)
            up, dn = (cur, mine) if down else (mine, cur)
            ands.append((up, dn, res[b].ptr))
            cur = wah_arg(res[b].ptr, dn.excl, dn.n, dn.skip)
            b ^= 1
        final = wah_arg(cur.words, rex.ptr, cur.n, rsk.ptr)

        def step():
            for up, dn, o in ands:
                check(dev.lib.mbx_wah_and_no1_skip(dev.ctx, C.byref(up),
                                                   C.byref(dn), o))
            check(dev.lib.mbx_wah_scan_skip(dev.ctx, final.words, final.n,
                                            rex.ptr, rsk.ptr, nsk))
            check(dev.lib.mbx_wah_decompress_skip(dev.ctx, C.byref(final),
                                                  groups.ptr, ng))
            return final.words, final.n

        cur, cur_n = step()
        dev.sync()
        got = groups.host()
        check(dev.lib.mbx_prog_eval(dev.ctx, C.byref(prg), poss.ptr, None))
        check(dev.lib.mbx_wah_groups(dev.ctx, poss.ptr, n, groups.ptr))
        identical &= bool(np.array_equal(got, groups.host()))
        times = []
        for _ in range(reps):
            t.start()
            step()
            times.append(t.stop() * 1e3)
        per_q[s.name] = round(float(np.median(times)), 1)
        # keep the result for the isolated decompression roofline
        keep = dev.alloc(4 * max(1, cur_n), np.uint32, cur_n)
        if cur_n:
            check(dev.lib.mbx_memcpy_d2d(dev.ctx, keep.ptr, cur, 4 * cur_n))
        kex = dev.alloc(4 * (cur_n + 1), np.uint32, cur_n + 1)
        ksk = dev.alloc(4 * nsk, np.uint32, nsk)
        check(dev.lib.mbx_wah_scan_skip(dev.ctx, keep.ptr, cur_n, kex.ptr,
                                        ksk.ptr, nsk))
        steps.append((keep, kex, cur_n, ksk,
                      wah_arg(keep.ptr, kex.ptr, cur_n, ksk.ptr)))
        for w, ex, _, _, sk in ws:
            w.free()
            ex.free()
            sk.free()
    # decompression in isolation: every query's result back to back
    dev.sync()
    t.start()
    for _ in range(reps):
        for _, _, _, _, arg in steps:
            check(dev.lib.mbx_wah_decompress_skip(dev.ctx, C.byref(arg),
                                                  groups.ptr, ng))
    ms = t.stop()
    t.close()
    launch_us = ms * 1e3 / (reps * len(steps))
    alg = sum(4 * (k + k + 1 + ng + nsk) for _, _, k, _, _ in steps) / len(steps)
    for keep, kex, _, ksk, _ in steps:
        ksk.free()
        keep.free()
        kex.free()
    for a in [groups, tmp, rex, rsk, poss] + res:
        a.free()
    wah_sum = sum(per_q.values())
    return {
        "workload": f"the 12 config-2 queries over {n:,} rows: each "
                    "predicate's WAH bitmap (31-row groups, compressed); timed: "
                    "the index-only step (ANDs, count scan, decompression)",
        "index_only_us": per_q,
        "index_only_sum_us": round(wah_sum, 1),
        "fused_query_sum_us": round(sum(fused_us), 1),
        "fused_vs_wah_index_only": round(wah_sum / sum(fused_us), 2),
        "wah_words": sizes,
        "plain_words_per_bitmap": ng,
        "identical_to_fused_bitmaps": identical,
        "roofline_decompress": {
            "bound": "hbm", "achieved": round(alg / launch_us / 1e3, 1),
            "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(alg / launch_us / 1e3 / HBM_PEAK_GBS, 4),
            "alg_bytes_per_launch": int(alg), "launch_us": round(launch_us, 2),
            "note": "4 B per WAH word, count-scan and skip-table entry read, "
                    "4 B per group written; launches back to back under one "
                    "event pair"},
    }


def run_learned(dev, nprog=20000, neval=200):
    """Config 4: runtimes of random programs' candidate plans (evaluation
    orders x fuse / eager / materialise) measured by the engine, the
    16-64-1 fp32 MLP trained on them with the library's kernels, then
    held-out programs run with the as-written plan, the static heuristic and
    the learned choice (identical sums required)."""
    from merle_amd import learned
    t0 = time.time()
    sch = learned.config4_device(dev)
    cols = [f"c{c}" for c in range(learned.C4_COLS)]
    bins = learned.build_bins(dev, sch, cols)
    t1 = time.time()
    X, Y = learned.collect(dev, sch, bins, nprog, log=log)
    t2 = time.time()
    model = learned.CostModel(dev)
    loss = model.train(X, Y)
    t3 = time.time()
    ev = learned.evaluate(dev, sch, bins, model, neval)
    model.close()
    for c in sch.cols.values():
        c.dev.free()
    return {
        "workload": f"{learned.C4_ROWS:,} rows x 8 Zipf columns (s 0.5/1.0/1.5), "
                    f"{nprog} training programs, {neval} held-out programs",
        "samples": int(X.shape[0]), "setup_s": round(t1 - t0, 2),
        "collect_s": round(t2 - t1, 2), "train_s": round(t3 - t2, 2),
        "train_mse_log_ms": round(loss, 4), **ev,
    }


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=100)
    ap.add_argument("--streams", type=int, default=4,
                    help="concurrent streams the 12 queries are issued on")
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--rows", type=int, default=None,
                    help="fact rows (default: config 2's 120M)")
    ap.add_argument("--no-extras", action="store_true",
                    help="skip the batched / learned / serving sections")
    ap.add_argument("--serve-requests", type=int, default=1000)
    ap.add_argument("--serve-rates", default="2000,20000",
                    help="Poisson arrival rates of the serving replays (req/s)")
    ap.add_argument("--learned-programs", type=int, default=20000,
                    help="training programs of the plan chooser")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--copy-calibration", type=int, default=0,
                    help="run only a plain device copy of this many bytes")
    ap.add_argument("--batched-only", action="store_true",
                    help="run only the config-3 batched section (profiling)")
    ap.add_argument("--wah-only", action="store_true",
                    help="config 2 and the WAH section alone (profiling)")
    ap.add_argument("--serving-alone", action="store_true",
                    help="only the config-5 requests timed one by one (profiling)")
    args = ap.parse_args()
    if args.copy_calibration:
        return copy_calibration(args.copy_calibration, max(1, args.steps))
    if args.batched_only or args.serving_alone:
        from merle_amd.host import Device
        dev = Device(0)
        if args.batched_only:
            print(json.dumps({"batched": run_batched(dev)}), flush=True)
        else:
            print(json.dumps({"serving_alone": run_serving_alone(dev)}), flush=True)
        dev.close()
        return 0
    if args.gpus != 1:
        print(f"bench.py measures one GPU only (asked for {args.gpus}); the "
              "path shards by row range with one int64 group reduction at the "
              "end, which is not run here", file=sys.stderr)
        return 2

    from merle_amd import abi, configs, spec as S
    from merle_amd.host import Device, Timer, check

    dev = Device(0)
    n = args.rows or configs.C2_ROWS
    log(f"generating config-2 tables: {n:,} fact rows")
    sch = configs.config2_device(dev, n)
    specs = configs.config2_queries()
    log("building bitmap index (one bitmap per attribute value)")
    t = Timer(dev)
    idx, slabs, cms, cbytes = configs.config2_index(dev, sch, timer=t)
    t.close()
    nbm = sum(len(v) for v in idx.values())
    log(f"index: {nbm} bitmaps in {cms:.1f} ms ({cbytes / cms / 1e6:.0f} GB/s)")

    # outputs of the 12 queries in one buffer, zeroed by one fill per step
    offs = np.cumsum([0] + [q.nr_grp for q in specs])
    outbuf = dev.alloc(8 * int(offs[-1]), np.int64, int(offs[-1]))
    # static plan of each query (outside the timed region): column densities
    # -> sparsest column first, eager loads unless the early-out pays
    qs, plans = [], []
    for i, (s, (planned, op, flags)) in enumerate(
            zip(specs, configs.config2_plans(dev, sch, idx, specs))):
        qs.append((s, planned, op, outbuf.ptr + 8 * int(offs[i])))
        plans.append(flags)

    # algorithmic bytes of each query (device accounting kernel)
    balg, nsel = [], []
    for s, prg, op, out in qs:
        b, k = C.c_uint64(), C.c_uint64()
        check(dev.lib.mbx_alg_bytes(dev.ctx, C.byref(prg), C.byref(op),
                                    s.nr_grp, C.byref(b), C.byref(k)))
        balg.append(b.value)
        nsel.append(k.value)
    log("selectivity:", ", ".join(f"{s.name}={k / n:.1e}" for s, k in zip(specs, nsel)))

    # The step: the 12 queries, one fused launch each, issued round robin on
    # `streams` concurrent streams of the context (queries are independent;
    # each zeroes its own output on its stream before its launch), then the
    # context's stream joins the others.
    nstreams = max(1, args.streams)
    own = C.c_void_p(dev.lib.mbx_stream(dev.ctx))
    extra = [C.c_void_p() for _ in range(nstreams - 1)]
    for x in extra:
        check(dev.lib.mbx_stream_create(C.byref(x)))
    streams = [own] + extra
    joinv = (C.c_void_p * max(1, len(extra)))(*[x.value for x in extra])

    def step(ns):
        for i, (s, prg, op, out) in enumerate(qs):
            check(dev.lib.mbx_set_stream(dev.ctx, streams[i % ns]))
            check(dev.lib.mbx_prog_grpby(dev.ctx, C.byref(prg), C.byref(op),
                                         s.nr_grp, out, plans[i]))
        check(dev.lib.mbx_set_stream(dev.ctx, None))
        if ns > 1:
            check(dev.lib.mbx_stream_join(dev.ctx, joinv, ns - 1))

    def isolated(i):
        """Query i alone on the context's stream, adding into its output."""
        s, prg, op, out = qs[i]
        check(dev.lib.mbx_prog_grpby(dev.ctx, C.byref(prg), C.byref(op),
                                     s.nr_grp, out, plans[i] | abi.ACCUMULATE))

    def step_isolated(timers):
        """One stream; one fill zeroes all outputs, then each launch adds
        into its output between its own event pair."""
        check(dev.lib.mbx_fill_u32(dev.ctx, outbuf.ptr, 0, 2 * int(offs[-1])))
        for i in range(len(qs)):
            timers[i].start()
            isolated(i)
            timers[i].mark_stop()

    def sequence_isolated(reps):
        """The 12 launches back to back on one stream (each waits for the
        previous one), `reps` times inside one event pair: ms per sequence."""
        check(dev.lib.mbx_fill_u32(dev.ctx, outbuf.ptr, 0, 2 * int(offs[-1])))
        t = Timer(dev)
        dev.sync()
        t.start()
        for _ in range(reps):
            for i in range(len(qs)):
                isolated(i)
        ms = t.stop()
        t.close()
        return ms / reps

    def timed_steps(ns, k):
        whole = Timer(dev)
        dev.sync()
        whole.start()
        for _ in range(k):
            step(ns)
        ms = whole.stop()
        whole.close()
        return ms / k

    for _ in range(args.warmup):
        step(nstreams)
    dev.sync()
    log(f"warmup done; timing {args.steps} steps on {nstreams} streams")
    ms_per_step = timed_steps(nstreams, args.steps)
    # the same step with every launch on one stream, for comparison
    for _ in range(args.warmup):
        step(1)
    ms_one_stream = timed_steps(1, args.steps)
    # The fused kernel in isolation: its launches back to back on one stream
    # under one event pair, so the mean launch duration holds the kernels and
    # the gaps between them but no event records.  Per query, event pairs
    # around each launch (each record adds a few microseconds).
    ms_seq = sequence_isolated(max(3, args.steps))
    timers = [Timer(dev) for _ in qs]
    per_q = [[] for _ in qs]
    for _ in range(max(3, args.steps)):
        step_isolated(timers)
        for i, t in enumerate(timers):
            per_q[i].append(t.elapsed())
    for t in timers:
        t.close()
    dev.sync()
    for x in extra:
        check(dev.lib.mbx_stream_destroy(x))
    qms = [float(np.median(v)) for v in per_q]
    rows_per_s = n * len(qs) / (ms_per_step / 1e3)
    achieved = sum(balg) / (ms_seq / 1e3) / 1e9
    step_gbs = sum(balg) / (ms_per_step / 1e3) / 1e9
    log(f"{ms_per_step:.3f} ms/step on {nstreams} streams ({ms_one_stream:.3f} on "
        f"one), {rows_per_s / 1e9:.1f} G rows/s; fused kernel {achieved:.0f} GB/s "
        f"algorithmic in isolation, {step_gbs:.0f} GB/s over the step")
    for (s, *_), b, q in zip(qs, balg, qms):
        log(f"  {s.name}: {q * 1e3:.1f} us, {b / 1e6:.1f} MB alg, "
            f"{b / q / 1e6:.0f} GB/s")

    result = {
        "metric": METRIC,
        "value": round(rows_per_s / 1e9, 3),
        "unit": "G rows/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "int64",
        "data": "synthetic (seeded star schema, uniform keys)",
        "config": {
            "workload": f"star schema, {n:,} fact rows, 12 queries of 2-4 "
                        "predicates, one fused launch per query, the "
                        f"launches issued round robin on {nstreams} "
                        "concurrent streams",
            "fact_rows": n, "queries": len(qs),
            "bitmaps": nbm, "streams": nstreams,
            "plan": "fused, static heuristic (sparsest first; eager loads "
                    f"for {sum(1 for f in plans if f & abi.PLAN_EAGER)} of 12)",
        },
        "roofline": {
            "bound": "hbm", "achieved": round(achieved, 1),
            "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4),
            "traffic": profiled_traffic(),
            "alg_bytes_per_launch": int(sum(balg) / len(balg)),
            "launch_us": round(ms_seq * 1e3 / len(qs), 2),
            "note": "fused kernel in isolation: the 12 launches back to back "
                    "on one stream under one event pair; launch_us = that "
                    "time / 12 (kernels plus the gaps between them)",
        },
        "roofline_step": {
            "achieved": round(step_gbs, 1), "peak": HBM_PEAK_GBS,
            "unit": "GB/s", "frac": round(step_gbs / HBM_PEAK_GBS, 4),
            "note": "algorithmic bytes of the 12 launches / ms_per_step "
                    "(launches overlapping on the streams)",
        },
        "one_stream": {"ms_per_step": round(ms_one_stream, 4),
                       "value": round(n * len(qs) / ms_one_stream / 1e6, 3)},
        "index_build": {"ms": round(cms, 2),
                        "GB_per_s": round(cbytes / cms / 1e6, 1)},
        "per_query_us": {s.name: round(q * 1e3, 2) for (s, *_), q in zip(qs, qms)},
        "per_query_note": "event pair around each isolated launch (the "
                          "records add a few microseconds to each)",
    }

    if not args.no_cpu and not args.wah_only:
        log("CPU baseline (oracle) on a bounded sample")
        cb, R, cres = cpu_baseline(dev, sch, specs)
        result["cpu_baseline"] = cb
        # parity on the sample: the same queries, under the plans the timed
        # step runs, restricted to R rows
        ok = True
        for (s, prg, op, out), want, o, fl in zip(qs, cres, offs, plans):
            prg.factsz = R
            check(dev.lib.mbx_prog_grpby(dev.ctx, C.byref(prg), C.byref(op),
                                         s.nr_grp, out, fl & ~abi.ACCUMULATE))
            dev.sync()
            got = outbuf.host()[int(o):int(o) + s.nr_grp]
            ok &= bool(np.array_equal(got, want))
            prg.factsz = n
        result["parity_on_cpu_sample"] = ok
        log(f"cpu {cb['value']} G rows/s on {cb['cores']} threads; parity {ok}")
    if not args.no_extras or args.wah_only:
        log("WAH baseline on the config-2 queries")
        result["wah"] = run_wah(dev, sch, specs, qs, [q * 1e3 for q in qms])
        log(f"  {result['wah']}")
    if not args.no_extras and not args.wah_only:
        outbuf.free()
        for a in slabs + [c.dev for c in sch.cols.values()]:
            a.free()
        log("config 3: batched vs plain")
        result["batched"] = run_batched(dev)
        log(f"  {result['batched']}")
        log("config 4: learned plan chooser")
        result["learned"] = run_learned(dev, args.learned_programs)
        log(f"  {result['learned']}")
        log("config 5: serving")
        result["serving"] = run_serving(
            dev, args.serve_requests,
            tuple(float(r) for r in args.serve_rates.split(",")))
        log(f"  {result['serving']}")
    print(json.dumps(result), flush=True)
    dev.close()
    return 0


if __name__ == "__main__":
    sys.exit(main())
