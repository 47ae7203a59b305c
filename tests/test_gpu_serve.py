"""The server: memory-based admission, continuous batching over several
streams; every result the same alone and in the mix; the live
submit / poll / wait / cancel API; ordering after the caller's own work on
the context's stream."""
import ctypes as C

import numpy as np
import pytest

from merle_amd import abi, configs, learned, spec as S
from merle_amd.host import check
from oracle import cpuref

import gpuq

pytestmark = pytest.mark.gpu

N = 1_500_001


@pytest.fixture(scope="module")
def tables(dev):
    host = configs.config1_host(N, seed=777)
    sch = S.Schema(N)
    for k, v in host.items():
        sch.add(k, v, dev.upload(v))
    b = list(configs.C1_BOUNDS)
    bins = {c: gpuq.create_bin(dev, sch, S.Src(c), b[:-1], b[1:])
            for c in configs.C1_COLS}
    yield sch, bins
    for c in configs.C1_COLS:
        check(dev.lib.mbx_free(dev.ctx, bins[c][0]))
    for c in sch.cols.values():
        c.dev.free()


def _queries(sch, bins, specs):
    qs = (abi.Query * len(specs))()
    for i, s in enumerate(specs):
        qs[i].prg = S.to_vprg(configs.bin_cols(s, bins), N, s.instrs)
        qs[i].op = S.to_rowop(s, sch)
        qs[i].nr_grp = s.nr_grp
    return qs


def _alone(dev, qs, specs):
    res = []
    for i, s in enumerate(specs):
        out = dev.alloc(8 * s.nr_grp, np.int64, s.nr_grp)
        check(dev.lib.mbx_prog_grpby(dev.ctx, C.byref(qs[i].prg), C.byref(qs[i].op),
                                     s.nr_grp, out.ptr, abi.PLAN_FUSED))
        dev.sync()
        res.append(out.host())
        out.free()
    return res


def _replay(dev, qs, arrivals, cfg, specs):
    tot = sum(s.nr_grp for s in specs)
    res = np.zeros(tot, np.int64)
    lat = np.zeros(len(specs))
    st = abi.ServeStats()
    check(dev.lib.mbx_serve_replay(dev.ctx, qs, len(specs),
                                   arrivals.ctypes.data_as(C.POINTER(C.c_double)),
                                   C.byref(cfg), res.ctypes.data,
                                   lat.ctypes.data_as(C.POINTER(C.c_double)),
                                   C.byref(st)))
    offs = np.cumsum([0] + [s.nr_grp for s in specs])
    return [res[offs[i]:offs[i + 1]] for i in range(len(specs))], lat, st


@pytest.mark.parametrize("mode", ["batched", "fused", "plain"])
def test_same_alone_and_in_the_mix(dev, tables, mode):
    sch, bins = tables
    specs = configs.config3_queries(80, seed=31)
    qs = _queries(sch, bins, specs)
    alone = _alone(dev, qs, specs)
    for i in range(0, 80, 10):
        np.testing.assert_array_equal(alone[i], cpuref.query(specs[i], sch)[0])
    rng = np.random.default_rng(5)
    arrivals = np.cumsum(rng.exponential(1 / 200000.0, len(specs)))  # 200k req/s
    # "fused": every request large enough for its own fused launch
    cfg = abi.ServeCfg(4, 16, 64 << 20, 1 if mode == "plain" else 0,
                       1 if mode == "fused" else 0, abi.Tune())
    mixed, lat, st = _replay(dev, qs, arrivals, cfg, specs)
    assert st.completed == len(specs) and st.rejected == 0
    assert np.all(lat > 0) and st.p99_s >= st.p50_s > 0
    if mode == "batched":
        assert st.mean_batch > 1.0          # requests actually shared launches
    else:
        # plain, and large requests (fused launches of their own) go alone
        assert st.mean_batch == 1.0
    for i in range(len(specs)):
        np.testing.assert_array_equal(mixed[i], alone[i], err_msg=specs[i].name)


def test_admission_by_memory(dev, tables):
    """A tight budget makes requests wait for memory (outputs are freed as
    batches finish); a request larger than the whole budget is refused.  The
    context's stream is kept busy first, so no request can finish (and free
    its output) before all of them have arrived: the budget binds for sure."""
    sch, bins = tables
    specs = configs.config3_queries(40, seed=32)
    huge = specs[3]
    specs[3] = S.QuerySpec(huge.preds, huge.instrs, huge.keys, 40_000,
                           huge.measure, huge.ma, huge.mb, "huge")
    qs = _queries(sch, bins, specs)
    alone = _alone(dev, qs, specs)
    arrivals = np.zeros(len(specs))           # everything arrives at once
    # the server's fixed memory plus 64 KiB: below the 320,000 B of the
    # 40,000-group request and below the 86,528 B the other 39 outputs need
    # together, above any one of them
    cfg = abi.ServeCfg(2, 4, 0, 0, 0, abi.Tune())
    fixed = C.c_uint64()
    check(dev.lib.mbx_serve_fixed_mem(C.byref(cfg), C.byref(fixed)))
    budget = fixed.value + 64 * 1024
    cfg.mem_budget = budget
    buf = dev.alloc(1 << 30)
    _busy_stream(dev, buf, 64)
    mixed, lat, st = _replay(dev, qs, arrivals, cfg, specs)
    buf.free()
    assert lat[3] < 0 and st.rejected == 1
    assert st.completed == len(specs) - 1
    assert 0 < st.admission_waits < len(specs)   # requests, not passes
    assert st.fixed_mem == fixed.value
    assert fixed.value < st.peak_mem <= budget
    for i in range(len(specs)):
        if i != 3:
            np.testing.assert_array_equal(mixed[i], alone[i])


def test_budget_below_fixed_memory_refused(dev):
    cfg = abi.ServeCfg(2, 4, 4096, 0, 0, abi.Tune())
    srv = C.c_void_p()
    assert dev.lib.mbx_serve_open(dev.ctx, C.byref(cfg), C.byref(srv)) == \
        abi.ERR_ADMISSION


def _busy_stream(dev, buf, reps):
    """Work on the context's stream that keeps it busy for a few ms."""
    for _ in range(reps):
        check(dev.lib.mbx_fill_u32(dev.ctx, buf.ptr, 7, buf.nbytes // 4))


def test_submit_poll_wait_cancel(dev, tables):
    """The live API.  The context's stream is kept busy, so the first request
    (whose output fills the budget) cannot finish before the others are
    submitted: the second is then deterministically queued, the third is
    cancelled while queued, and both others end with the results they get
    alone.  The requests also wait for the caller's earlier work."""
    sch, bins = tables
    specs = configs.config3_queries(3, seed=41)
    qs = _queries(sch, bins, specs)
    alone = _alone(dev, qs, specs)
    cfg = abi.ServeCfg(2, 4, 0, 0, 0, abi.Tune())
    fixed = C.c_uint64()
    check(dev.lib.mbx_serve_fixed_mem(C.byref(cfg), C.byref(fixed)))
    need = [(8 * s.nr_grp + 255) // 256 * 256 for s in specs]
    cfg.mem_budget = fixed.value + max(need[:2])   # one request at a time
    srv = C.c_void_p()
    check(dev.lib.mbx_serve_open(dev.ctx, C.byref(cfg), C.byref(srv)))
    buf = dev.alloc(1 << 30)
    try:
        _busy_stream(dev, buf, 64)
        outs = [np.full(s.nr_grp, -7, np.int64) for s in specs]
        tk = [C.c_uint64() for _ in specs]
        for i in range(3):
            check(dev.lib.mbx_serve_submit(srv, C.byref(qs[i]), outs[i].ctypes.data,
                                           C.byref(tk[i])))
        state, lat = C.c_int(), C.c_double()
        check(dev.lib.mbx_serve_poll(srv, tk[1].value, C.byref(state), C.byref(lat)))
        assert state.value == abi.REQ_QUEUED and lat.value < 0
        check(dev.lib.mbx_serve_cancel(srv, tk[2].value))
        check(dev.lib.mbx_serve_poll(srv, tk[2].value, C.byref(state), None))
        assert state.value == abi.REQ_CANCELLED
        for i in (0, 1):
            check(dev.lib.mbx_serve_wait(srv, tk[i].value, -1.0, C.byref(state),
                                         C.byref(lat)))
            assert state.value == abi.REQ_DONE and lat.value > 0
            np.testing.assert_array_equal(outs[i], alone[i])
        # a finished request cannot be cancelled
        assert dev.lib.mbx_serve_cancel(srv, tk[0].value) == abi.ERR_INVALID
        assert np.all(outs[2] == -7)             # never written
    finally:
        st = abi.ServeStats()
        check(dev.lib.mbx_serve_close(srv, C.byref(st)))
        buf.free()
    assert st.completed == 2 and st.cancelled == 1 and st.admission_waits >= 1


def test_serve_after_unsynced_bitmap_creation(dev):
    """Bitmaps built on the context's stream without a sync, then served at
    once: every request must see the finished bitmaps (oracle)."""
    n = 2_000_003
    host = configs.config1_host(n, seed=4242)
    sch = S.Schema(n)
    for k, v in host.items():
        sch.add(k, v, dev.upload(v))
    dev.sync()
    b = list(configs.C1_BOUNDS)
    ptrs = {}
    for c in configs.C1_COLS:      # consecutive bins, no FK: no wait inside
        ref = S.colref(sch, S.Src(c))
        outs = (C.c_void_p * 40)()
        check(dev.lib.mbx_create_bin(dev.ctx, C.byref(ref), n,
                                     (C.c_uint64 * 40)(*b[:-1]),
                                     (C.c_uint64 * 40)(*b[1:]), 40, outs, None))
        ptrs[c] = [outs[j] for j in range(40)]
    specs = configs.config3_queries(24, seed=43)
    qs = (abi.Query * len(specs))()
    for i, s in enumerate(specs):
        qs[i].prg = S.to_vprg(configs.bin_cols(s, ptrs), n, s.instrs)
        qs[i].op = S.to_rowop(s, sch)
        qs[i].nr_grp = s.nr_grp
    cfg = abi.ServeCfg(2, 4, 0, 0, 0, abi.Tune())
    mixed, lat, st = _replay(dev, qs, np.zeros(len(specs)), cfg, specs)
    assert st.completed == len(specs)
    for i, s in enumerate(specs):
        np.testing.assert_array_equal(mixed[i], cpuref.query(s, sch)[0],
                                      err_msg=s.name)
    for c in configs.C1_COLS:
        check(dev.lib.mbx_free(dev.ctx, ptrs[c][0]))
    for c in sch.cols.values():
        c.dev.free()


def test_two_streams_one_context(dev, tables):
    """Queries of one context on two streams (mbx_set_stream), never
    synchronised in between, run concurrently; each stream's launches use
    their own flush replicas, so every sum equals the oracle's."""
    sch, bins = tables
    hip = C.CDLL("libamdhip64.so.7")
    streams = [C.c_void_p(), C.c_void_p()]
    for s in streams:
        assert hip.hipStreamCreateWithFlags(C.byref(s), 1) == 0   # non-blocking
    specs = configs.config3_queries(40, seed=44)
    qs = _queries(sch, bins, specs)
    outs = [dev.alloc(8 * s.nr_grp, np.int64, s.nr_grp) for s in specs]
    try:
        for rep in range(3):
            for i, s in enumerate(specs):
                check(dev.lib.mbx_set_stream(dev.ctx, streams[i % 2]))
                check(dev.lib.mbx_prog_grpby(dev.ctx, C.byref(qs[i].prg),
                                             C.byref(qs[i].op), s.nr_grp,
                                             outs[i].ptr,
                                             abi.PLAN_FUSED | abi.PLAN_EAGER))
            for st_ in streams:
                check(dev.lib.mbx_set_stream(dev.ctx, st_))
                dev.sync()
            for i, s in enumerate(specs):
                np.testing.assert_array_equal(outs[i].host(),
                                              cpuref.query(s, sch)[0],
                                              err_msg=f"{s.name} rep {rep}")
    finally:
        check(dev.lib.mbx_set_stream(dev.ctx, None))
        for o in outs:
            o.free()
        for s in streams:
            hip.hipStreamDestroy(s)


def test_config5_full(dev):
    """Config 5 at full size: 600M rows x 10 Zipf(1.2) columns and their 640
    bucket bitmaps resident; 64 requests from the config-3 program space
    under the static heuristic plans, Poisson arrivals at 2,000 req/s, two
    slots of up to four: every result in the mix equals the request run
    alone, and eight of them equal the oracle over all 600M rows."""
    n = configs.C5_ROWS
    print(f"config 5: {n:,} rows", flush=True)
    sch = configs.config5_device(dev)
    bins, slabs = configs.build_bins(dev, sch, configs.C5_BOUNDS)
    specs = configs.config5_queries(64)
    qs = configs.heuristic_queries(
        dev, configs.query_array(sch, specs, bins, configs.C5_BOUNDS))
    try:
        alone = []
        for i, s in enumerate(specs):
            out = dev.alloc(8 * s.nr_grp, np.int64, s.nr_grp)
            check(dev.lib.mbx_prog_grpby(dev.ctx, C.byref(qs[i].prg),
                                         C.byref(qs[i].op), s.nr_grp, out.ptr,
                                         abi.PLAN_FUSED | qs[i].flags))
            dev.sync()
            alone.append(out.host())
            out.free()
        rng = np.random.default_rng(5)
        arrivals = np.cumsum(rng.exponential(1 / 2000.0, len(specs)))
        cfg = abi.ServeCfg(2, 4, 0, 0, 0, abi.Tune(8192, 4, 4096, 0))
        mixed, lat, st = _replay(dev, qs, arrivals, cfg, specs)
        assert st.completed == len(specs)
        for i, s in enumerate(specs):
            np.testing.assert_array_equal(mixed[i], alone[i], err_msg=s.name)
        check_idx = list(range(0, len(specs), 8))
        names = set().union(*(configs.spec_columns(specs[i]) for i in check_idx))
        host = configs.host_subset(sch, names)
        for i in check_idx:
            print(f"  oracle {specs[i].name}", flush=True)
            np.testing.assert_array_equal(alone[i],
                                          cpuref.query(specs[i], host)[0],
                                          err_msg=specs[i].name)
    finally:
        for a in slabs + [c.dev for c in sch.cols.values()]:
            a.free()
