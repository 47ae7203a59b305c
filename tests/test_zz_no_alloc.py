"""mbx_prog_grpby_noalloc: the entry point for callers that cannot allocate.

Through the C ABI, against mbx_prog_grpby on the same queries: every plan
(fused early-out and eager, materialise, scan), group tables in LDS and in
HBM, on two calls in a row where the second has a smaller input (fewer rows,
other data, fewer groups) -- the scratch keeps the first call's contents, so
whatever must start at zero has to be cleared or left clear.  The pool's
use, live count and capacity do not change across the calls (nothing is
allocated), and a materialise call above the scratch's row limit is refused
before any work.
"""
import ctypes as C

import numpy as np
import pytest

from merle_amd import abi, configs, spec as S
from merle_amd.host import check

import gpuq

pytestmark = pytest.mark.gpu

PLANS = [abi.PLAN_FUSED, abi.PLAN_FUSED | abi.PLAN_EAGER,
         abi.PLAN_MATERIALISE, abi.PLAN_SCAN]


def _tables(dev, n, seed):
    host = configs.config1_host(n, seed=seed)
    sch = S.Schema(n)
    for k, v in host.items():
        sch.add(k, v, dev.upload(v))
    b = list(configs.C1_BOUNDS)
    bins = {c: gpuq.create_bin(dev, sch, S.Src(c), b[:-1], b[1:])
            for c in configs.C1_COLS}
    return sch, bins


def _free(dev, sch, bins):
    for c in configs.C1_COLS:
        check(dev.lib.mbx_free(dev.ctx, bins[c][0]))
    for c in sch.cols.values():
        c.dev.free()


def _pick(specs, big):
    """A query with an LDS group table and, if big, one beyond it."""
    small = next(s for s in reversed(specs) if 1 < s.nr_grp <= 4096)
    if not big:
        return [small]
    return [small, next(s for s in reversed(specs) if s.nr_grp > 4096)]


def test_noalloc_matches_prog_grpby(dev):
    first = _tables(dev, 3_000_017, seed=71)
    second = _tables(dev, 1_000_003, seed=72)
    specs = configs.config3_queries(96, seed=73)
    try:
        for flags in PLANS:
            for (sch, bins), big in ((first, True), (second, False)):
                for s in _pick(specs, big):
                    prg = S.to_vprg(configs.bin_cols(s, bins), sch.nrows, s.instrs)
                    op = S.to_rowop(s, sch)
                    want = dev.alloc(8 * s.nr_grp, np.int64, s.nr_grp)
                    got = dev.alloc(8 * s.nr_grp, np.int64, s.nr_grp)
                    check(dev.lib.mbx_prog_grpby(dev.ctx, C.byref(prg), C.byref(op),
                                                 s.nr_grp, want.ptr, flags))
                    dev.sync()
                    # stale contents in the output must not leak into it
                    check(dev.lib.mbx_fill_u32(dev.ctx, got.ptr, 0xdeadbeef,
                                               2 * s.nr_grp))
                    dev.sync()
                    before = dev.pool_stats()
                    check(dev.lib.mbx_prog_grpby_noalloc(
                        dev.ctx, C.byref(prg), C.byref(op), s.nr_grp, got.ptr,
                        flags))
                    dev.sync()
                    assert dev.pool_stats() == before
                    np.testing.assert_array_equal(
                        got.host(), want.host(),
                        err_msg=f"{s.name} rows {sch.nrows} flags {flags}")
                    assert want.host().any()
                    want.free()
                    got.free()
    finally:
        _free(dev, *first)
        _free(dev, *second)


def test_noalloc_materialise_row_limit(dev):
    """Above the row limit the materialise plan is refused on the host; the
    fused plan of the same query is not limited by the scratch."""
    n = 16_000_000
    sch, bins = _tables(dev, n, seed=74)
    try:
        s = configs.config1_spec()
        prg = S.to_vprg(configs.bin_cols(s, bins), n, s.instrs)
        op = S.to_rowop(s, sch)
        out = dev.alloc(8 * s.nr_grp, np.int64, s.nr_grp)
        assert dev.lib.mbx_prog_grpby_noalloc(
            dev.ctx, C.byref(prg), C.byref(op), s.nr_grp, out.ptr,
            abi.PLAN_MATERIALISE) == abi.ERR_CAPACITY
        check(dev.lib.mbx_prog_grpby_noalloc(dev.ctx, C.byref(prg), C.byref(op),
                                             s.nr_grp, out.ptr, abi.PLAN_FUSED))
        dev.sync()
        want = dev.alloc(8 * s.nr_grp, np.int64, s.nr_grp)
        check(dev.lib.mbx_prog_grpby(dev.ctx, C.byref(prg), C.byref(op),
                                     s.nr_grp, want.ptr, abi.PLAN_FUSED))
        dev.sync()
        np.testing.assert_array_equal(out.host(), want.host())
        out.free()
        want.free()
    finally:
        _free(dev, sch, bins)
