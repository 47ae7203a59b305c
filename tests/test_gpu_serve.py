"""The serving loop: memory-based admission, continuous batching over several
streams; every result the same alone and in the mix."""
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
    if mode != "plain":
        assert st.mean_batch > 1.0          # requests actually shared launches
    for i in range(len(specs)):
        np.testing.assert_array_equal(mixed[i], alone[i], err_msg=specs[i].name)


def test_admission_by_memory(dev, tables):
    """A tight budget makes requests wait for memory (outputs are freed as
    batches finish); a request larger than the whole budget is refused."""
    sch, bins = tables
    specs = configs.config3_queries(40, seed=32)
    huge = specs[3]
    specs[3] = S.QuerySpec(huge.preds, huge.instrs, huge.keys, 40_000,
                           huge.measure, huge.ma, huge.mb, "huge")
    qs = _queries(sch, bins, specs)
    alone = _alone(dev, qs, specs)
    arrivals = np.zeros(len(specs))           # everything arrives at once
    # 64 KiB: below the 320,000 B of the 40,000-group request and below the
    # 86,528 B the other 39 outputs need together, above any one of them
    budget = 64 * 1024
    cfg = abi.ServeCfg(2, 4, budget, 0, 0, abi.Tune())
    mixed, lat, st = _replay(dev, qs, arrivals, cfg, specs)
    assert lat[3] < 0 and st.rejected == 1
    assert st.completed == len(specs) - 1
    assert st.admission_waits > 0
    assert st.peak_mem <= budget
    for i in range(len(specs)):
        if i != 3:
            np.testing.assert_array_equal(mixed[i], alone[i])
