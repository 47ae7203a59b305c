"""Config 2 at full size: the 120M-row star schema, its 1,597 per-value
bitmaps projected through the foreign keys, and the 12 queries, against the
oracle over all rows.

The queries run under the plans the benchmark times (the static heuristic:
sparsest column first, eager loads for all 12), and two of them -- Q2.3
(sparsest, LDS group table) and Q4.2 (437,500 groups in HBM) -- under every
other plan.  Bitmaps of four attributes are compared byte for byte with the
oracle's create_bin.
"""
import ctypes as C

import numpy as np
import pytest

from merle_amd import abi, configs, spec as S
from merle_amd.host import check, stride_words
from oracle import cpuref

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def star(dev):
    print(f"config 2: {configs.C2_ROWS:,} rows", flush=True)
    sch = configs.config2_device(dev)
    idx, slabs, _, _ = configs.config2_index(dev, sch)
    host = configs.host_schema(sch)
    specs = configs.config2_queries()
    want = {s.name: cpuref.query(s, host)[0] for s in specs}
    yield sch, idx, host, specs, want
    for a in slabs + [c.dev for c in sch.cols.values()]:
        a.free()


def _run(dev, prg, op, nr_grp, flags):
    out = dev.alloc(8 * nr_grp, np.int64, nr_grp)
    check(dev.lib.mbx_prog_grpby(dev.ctx, C.byref(prg), C.byref(op), nr_grp,
                                 out.ptr, flags))
    dev.sync()
    got = out.host()
    out.free()
    return got


def test_config2_timed_plans(dev, star):
    sch, idx, host, specs, want = star
    plans = configs.config2_plans(dev, sch, idx, specs)
    assert all(fl & abi.PLAN_EAGER for _, _, fl in plans)
    for s, (prg, op, fl) in zip(specs, plans):
        print(f"  {s.name} flags {fl}", flush=True)
        np.testing.assert_array_equal(_run(dev, prg, op, s.nr_grp, fl),
                                      want[s.name], err_msg=s.name)
        assert want[s.name].any(), s.name


@pytest.mark.parametrize("name", ["Q2.3", "Q4.2"])
@pytest.mark.parametrize("flags", [abi.PLAN_FUSED, abi.PLAN_MATERIALISE,
                                   abi.PLAN_MATERIALISE | abi.PLAN_EAGER,
                                   abi.PLAN_SCAN, abi.PLAN_DIRECT])
def test_config2_every_plan(dev, star, name, flags):
    sch, idx, host, specs, want = star
    s = next(q for q in specs if q.name == name)
    prg = S.to_vprg(configs.config2_cols(s, idx), sch.nrows, s.instrs)
    op = S.to_rowop(s, sch)
    print(f"  {name} flags {flags}", flush=True)
    np.testing.assert_array_equal(_run(dev, prg, op, s.nr_grp, flags),
                                  want[name], err_msg=f"{name} flags {flags}")


@pytest.mark.parametrize("attr", ["d_year", "c_region", "s_nation", "p_mfgr"])
def test_config2_bitmaps_byte_identical(dev, star, attr):
    sch, idx, host, specs, want = star
    dim, fk, card = configs.C2_ATTRS[attr]
    n = sch.nrows
    ref = cpuref.create_bin(host[attr].host, list(range(card)),
                            list(range(1, card + 1)), fk=host[fk].host,
                            dimsz=host[attr].host.size)
    sw = stride_words(n)
    got = np.empty(sw, np.uint32)
    for v in range(card):
        check(dev.lib.mbx_memcpy_d2h(dev.ctx, got.ctypes.data, idx[attr][v],
                                     sw * 4))
        np.testing.assert_array_equal(got, ref[v], err_msg=f"{attr}={v}")
    # the bins partition the rows: popcounts add up to N
    assert int(np.bitwise_count(ref).sum()) == n
