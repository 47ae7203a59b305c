"""Batched persistent launch (hipGraph) vs the plain one-query-per-launch
path, on config-3-style queries over config-1-shaped tables."""
import ctypes as C

import numpy as np
import pytest

from merle_amd import abi, configs, spec as S
from merle_amd.host import check
from oracle import cpuref

import gpuq

pytestmark = pytest.mark.gpu

N = 2_000_003   # smaller than config 1, not a multiple of 64


@pytest.fixture(scope="module")
def tables(dev):
    host = configs.config1_host(N, seed=4242)
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


def make_queries(dev, sch, bins, specs):
    qs = (abi.Query * len(specs))()
    outs = []
    for i, s in enumerate(specs):
        qs[i].prg = S.to_vprg(configs.bin_cols(s, bins), N, s.instrs)
        qs[i].op = S.to_rowop(s, sch)
        qs[i].nr_grp = s.nr_grp
        out = dev.alloc(8 * s.nr_grp, np.int64, s.nr_grp)
        qs[i].d_out = out.ptr
        outs.append(out)
    return qs, outs


def plain_results(dev, qs, specs):
    res = []
    for i, s in enumerate(specs):
        out = dev.alloc(8 * s.nr_grp, np.int64, s.nr_grp)
        check(dev.lib.mbx_prog_grpby(dev.ctx, C.byref(qs[i].prg), C.byref(qs[i].op),
                                     s.nr_grp, out.ptr, abi.PLAN_FUSED))
        dev.sync()
        res.append(out.host())
        out.free()
    return res


@pytest.mark.parametrize("eager", [False, True])
def test_batch_equals_plain_and_oracle(dev, tables, eager):
    """Early-out queries all run in the general kernel; with eager loads the
    conjunctions with up to 3,000 groups run in the conjunction kernel and
    the rest (larger LDS tables, HBM tables) in the general one, in the same
    graph."""
    sch, bins = tables
    specs = configs.config3_queries(64, seed=12)
    qs, outs = make_queries(dev, sch, bins, specs)
    plain = plain_results(dev, qs, specs)
    if eager:
        for i in range(len(specs)):
            qs[i].flags = abi.PLAN_EAGER
        assert any(s.nr_grp > 4096 for s in specs)
        assert any(3000 < s.nr_grp <= 4096 for s in specs)
    for i in range(0, len(specs), 8):           # oracle on a subset
        want, _ = cpuref.query(specs[i], sch)
        np.testing.assert_array_equal(plain[i], want, err_msg=specs[i].name)
    b = C.c_void_p()
    check(dev.lib.mbx_batch_create(dev.ctx, qs, len(specs), None, C.byref(b)))
    # graph replays with eager work in between; every replay == plain
    for rep in range(3):
        check(dev.lib.mbx_batch_run(dev.ctx, b))
        dev.sync()
        for i, o in enumerate(outs):
            np.testing.assert_array_equal(o.host(), plain[i],
                                          err_msg=f"replay {rep} {specs[i].name}")
        gpuq.run_query(dev, sch, specs[0], configs.bin_cols(specs[0], bins))
    check(dev.lib.mbx_batch_run_eager(dev.ctx, b))
    dev.sync()
    for i, o in enumerate(outs):
        np.testing.assert_array_equal(o.host(), plain[i])
    check(dev.lib.mbx_batch_destroy(dev.ctx, b))
    for o in outs:
        o.free()


@pytest.mark.parametrize("tune", [(512, 2, 1024), (8192, 8, 4096), (1024, 1, 0)])
def test_batch_launch_shapes_agree(dev, tables, tune):
    """Results do not depend on tile size, grid or LDS group capacity."""
    sch, bins = tables
    specs = configs.config3_queries(24, seed=12)
    qs, outs = make_queries(dev, sch, bins, specs)
    plain = plain_results(dev, qs, specs)
    for i in range(0, len(specs), 2):       # eager loads for every other query
        qs[i].flags = abi.PLAN_EAGER
    t = abi.Tune(*tune)
    b = C.c_void_p()
    check(dev.lib.mbx_batch_create(dev.ctx, qs, len(specs), C.byref(t), C.byref(b)))
    check(dev.lib.mbx_batch_run(dev.ctx, b))
    dev.sync()
    for i, o in enumerate(outs):
        np.testing.assert_array_equal(o.host(), plain[i])
    check(dev.lib.mbx_batch_destroy(dev.ctx, b))
    for o in outs:
        o.free()


def test_config3_full(dev):
    """Config 3 at full size: 256 queries over the 12M-row config-1 tables in
    one persistent graph launch under the static heuristic plans (as the
    benchmark times it): batch == plain == oracle for all 256 queries."""
    n = configs.C1_ROWS
    host = configs.config1_host(n)
    sch = S.Schema(n)
    for k, v in host.items():
        sch.add(k, v, dev.upload(v))
    b = list(configs.C1_BOUNDS)
    bins = {c: gpuq.create_bin(dev, sch, S.Src(c), b[:-1], b[1:])
            for c in configs.C1_COLS}
    specs = configs.config3_queries(256)
    qs = (abi.Query * len(specs))()
    outs = []
    for i, s in enumerate(specs):
        prg = S.to_vprg(configs.bin_cols(s, bins), n, s.instrs)
        dens = (C.c_double * 4)()
        check(dev.lib.mbx_col_density(dev.ctx, C.byref(prg), dens))
        plan = abi.Plan()
        check(dev.lib.mbx_plan_heuristic(C.byref(prg), dens, C.byref(plan)))
        check(dev.lib.mbx_plan_apply(C.byref(prg), C.byref(plan), C.byref(qs[i].prg)))
        qs[i].op = S.to_rowop(s, sch)
        qs[i].nr_grp = s.nr_grp
        qs[i].flags = plan.flags & abi.PLAN_EAGER
        outs.append(dev.alloc(8 * s.nr_grp, np.int64, s.nr_grp))
        qs[i].d_out = outs[-1].ptr
    plain = []
    for i, s in enumerate(specs):
        out = dev.alloc(8 * s.nr_grp, np.int64, s.nr_grp)
        check(dev.lib.mbx_prog_grpby(dev.ctx, C.byref(qs[i].prg), C.byref(qs[i].op),
                                     s.nr_grp, out.ptr,
                                     abi.PLAN_FUSED | qs[i].flags))
        dev.sync()
        plain.append(out.host())
        out.free()
    tune, ms = abi.Tune(), C.c_float()
    check(dev.lib.mbx_autotune(dev.ctx, qs, len(specs), C.byref(tune), C.byref(ms)))
    bt = C.c_void_p()
    check(dev.lib.mbx_batch_create(dev.ctx, qs, len(specs), C.byref(tune),
                                   C.byref(bt)))
    for rep in range(2):
        check(dev.lib.mbx_batch_run(dev.ctx, bt))
        dev.sync()
        for i, s in enumerate(specs):
            np.testing.assert_array_equal(outs[i].host(), plain[i],
                                          err_msg=f"replay {rep} {s.name}")
    check(dev.lib.mbx_batch_destroy(dev.ctx, bt))
    for i, s in enumerate(specs):
        np.testing.assert_array_equal(plain[i], cpuref.query(s, sch)[0],
                                      err_msg=s.name)
    for o in outs:
        o.free()
    for c in configs.C1_COLS:
        check(dev.lib.mbx_free(dev.ctx, bins[c][0]))
    for c in sch.cols.values():
        c.dev.free()


def _time_batch(dev, qs, n, tune, reps=7):
    from merle_amd.host import Timer
    b = C.c_void_p()
    check(dev.lib.mbx_batch_create(dev.ctx, qs, n, C.byref(tune), C.byref(b)))
    t = Timer(dev)
    check(dev.lib.mbx_batch_run(dev.ctx, b))
    ts = []
    for _ in range(reps):
        t.start()
        check(dev.lib.mbx_batch_run(dev.ctx, b))
        ts.append(t.stop())
    t.close()
    check(dev.lib.mbx_batch_destroy(dev.ctx, b))
    return float(np.median(ts))


def test_autotune(dev, tables):
    """The autotuner returns the fastest of the shapes it timed, and that
    shape, timed again as a graph launch, is no slower than the default
    shape or the slowest candidate (within timing noise)."""
    sch, bins = tables
    specs = configs.config3_queries(64, seed=13)
    qs, outs = make_queries(dev, sch, bins, specs)
    for i in range(len(specs)):
        qs[i].flags = abi.PLAN_EAGER
    best, ms = abi.Tune(), C.c_float()
    cands, cms, cnt = (abi.Tune * 64)(), (C.c_float * 64)(), C.c_int()
    check(dev.lib.mbx_autotune_report(dev.ctx, qs, len(specs), C.byref(best),
                                      C.byref(ms), cands, cms, 64, C.byref(cnt)))
    assert cnt.value == 40
    times = [cms[i] for i in range(cnt.value)]
    k = int(np.argmin(times))
    assert (best.tile_words, best.wg_per_cu, best.lds_groups) == \
        (cands[k].tile_words, cands[k].wg_per_cu, cands[k].lds_groups)
    assert ms.value == min(times) > 0
    assert best.tile_words % 512 == 0 and 1 <= best.wg_per_cu <= 16
    slow = cands[int(np.argmax(times))]
    t_best = _time_batch(dev, qs, len(specs), best)
    t_def = _time_batch(dev, qs, len(specs), abi.Tune(2048, 4, 4096, 0))
    t_slow = _time_batch(dev, qs, len(specs), slow)
    print(f"autotune: best {t_best:.3f} ms, default {t_def:.3f}, slowest "
          f"candidate {t_slow:.3f}")
    assert t_best <= 1.1 * t_def and t_best <= 1.1 * t_slow
    for o in outs:
        o.free()


def test_bad_tune_rejected(dev, tables):
    sch, bins = tables
    specs = configs.config3_queries(2, seed=14)
    qs, outs = make_queries(dev, sch, bins, specs)
    b = C.c_void_p()
    t = abi.Tune(100, 4, 4096)
    assert dev.lib.mbx_batch_create(dev.ctx, qs, 2, C.byref(t), C.byref(b)) == abi.ERR_INVALID
    for o in outs:
        o.free()
