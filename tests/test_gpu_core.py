"""GPU parity of the core path through the C ABI: bitmap creation, the fused
program + group-by, the materialise and scan plans, against the oracle."""
import ctypes as C

import numpy as np
import pytest

from merle_amd import abi, configs, spec as S
from merle_amd.host import check, stride_words
from oracle import cpuref

import datagen
import gpuq

pytestmark = pytest.mark.gpu

SIZES = [1, 31, 32, 33, 63, 64, 65, 1000, 1023, 1024, 1025, 4099, 65537,
         1 << 20 | 17]


@pytest.mark.parametrize("n", SIZES)
@pytest.mark.parametrize("dt", [np.uint8, np.uint16, np.uint32, np.uint64])
def test_bins_all_modes(dev, n, dt):
    """Byte-identical bitmaps vs ref_create_bin for every bin mode: width-1,
    equal-width, sorted (uneven) and arbitrary (overlapping) bins, each with
    more than 8 bins (lines stored through the LDS transpose) and with up to
    8 (one lane per line)."""
    rng = np.random.default_rng(n + np.dtype(dt).itemsize)
    v = rng.integers(0, 300, n).astype(dt)
    d = dev.upload(v)
    ref = abi.ColRef(d.ptr, None, 0, v.dtype.itemsize)
    uneven = np.unique(rng.integers(0, 300, 70))
    sets = [(list(range(10, 110)), list(range(11, 111))),          # unit, 100 bins
            (list(range(0, 300, 30)), list(range(30, 330, 30))),   # equal
            (list(uneven[:-1]), list(uneven[1:])),                 # sorted
            (list(rng.integers(0, 150, 7)), list(rng.integers(100, 300, 7))),
            (list(range(40, 45)), list(range(41, 46))),            # unit, 5 bins
            (list(rng.integers(0, 150, 70)),                       # arbitrary, 70
             list(rng.integers(100, 300, 70)))]
    for mins, maxs in sets:
        nb = len(mins)
        outs = (C.c_void_p * nb)()
        check(dev.lib.mbx_create_bin(dev.ctx, C.byref(ref), n,
                                     (C.c_uint64 * nb)(*map(int, mins)),
                                     (C.c_uint64 * nb)(*map(int, maxs)), nb,
                                     outs, None))
        got = gpuq.bitmaps_host(dev, list(outs), n)
        want = cpuref.create_bin(v, mins, maxs)
        np.testing.assert_array_equal(got, want)
        check(dev.lib.mbx_free(dev.ctx, outs[0]))
    d.free()


@pytest.mark.parametrize("n", [1, 77, 4099, 300001])
def test_bins_fk_projection(dev, n):
    """FK projection (create_bin with fk, bmpidx_create.cu:25-34) incl.
    fk >= dimsz -> INT64_MAX, and create_join / create_select."""
    rng = np.random.default_rng(n)
    dim = rng.integers(0, 50, 1000).astype(np.uint16)
    fk = rng.integers(0, 1100, n).astype(np.uint32)
    dd, df = dev.upload(dim), dev.upload(fk)
    ref = abi.ColRef(dd.ptr, df.ptr, 1000, 2)
    mins = list(range(50)) + [1 << 62]
    maxs = list(range(1, 51)) + [(1 << 64) - 1]
    nb = len(mins)
    outs = (C.c_void_p * nb)()
    check(dev.lib.mbx_create_bin(dev.ctx, C.byref(ref), n,
                                 (C.c_uint64 * nb)(*mins),
                                 (C.c_uint64 * nb)(*maxs), nb, outs, None))
    got = gpuq.bitmaps_host(dev, list(outs), n)
    np.testing.assert_array_equal(got, cpuref.create_bin(dim, mins, maxs, fk, 1000))
    out = C.c_void_p()
    check(dev.lib.mbx_create_join(dev.ctx, df.ptr, dd.ptr, 2, n, 1000, 7, 31,
                                  C.byref(out)))
    got = gpuq.bitmaps_host(dev, [out.value], n)[0]
    np.testing.assert_array_equal(got, cpuref.create_join(dim, 7, 31, fk, 1000))
    check(dev.lib.mbx_create_select(dev.ctx, df.ptr, 4, n, 100, 900, C.byref(out)))
    got = gpuq.bitmaps_host(dev, [out.value], n)[0]
    np.testing.assert_array_equal(got, cpuref.create_join(fk, 100, 900))


@pytest.mark.parametrize("n", [1, 4099, 300001])
@pytest.mark.parametrize("dt", [np.uint16, np.uint32, np.uint64])
@pytest.mark.parametrize("dimsz", [3000, 40000, 70000])
def test_bins_fk_radix_table(dev, n, dt, dimsz):
    """FK projection on the radix path, through the dimension's bin-index
    table (uint8 entries below 255 bins, uint16 from 255), read from LDS
    (tables up to 64 KB: 3,000 rows, and 40,000 rows of uint8) or from L2:
    width-1, equal-width and sorted bins, 5 to 400 bins, keys past the
    dimension (INT64_MAX: caught by a last bin reaching 2^64-1, else by
    none)."""
    rng = np.random.default_rng(n * 7 + np.dtype(dt).itemsize + dimsz)
    dim = rng.integers(0, 2000, dimsz).astype(dt)
    fk = rng.integers(0, dimsz + dimsz // 10, n).astype(np.uint32)
    dd, df = dev.upload(dim), dev.upload(fk)
    ref = abi.ColRef(dd.ptr, df.ptr, dimsz, np.dtype(dt).itemsize)
    top = (1 << 64) - 1
    uneven = [int(x) for x in np.unique(rng.integers(0, 2000, 500))]
    sets = [(list(range(100, 105)), list(range(101, 106))),        # unit, 5
            (list(range(0, 254)), list(range(1, 255))),           # unit, 254
            (list(range(0, 255)), list(range(1, 256))),           # unit, 255
            (list(range(0, 300)), list(range(1, 301))),           # unit, 300
            (list(range(0, 2000, 50)), list(range(50, 2050, 50))),  # equal, 40
            (uneven[:30], uneven[1:30] + [top]),                  # sorted, to 2^64-1
            (uneven[:400], uneven[1:401])]                        # sorted, 400
    for mins, maxs in sets:
        nb = len(mins)
        outs = (C.c_void_p * nb)()
        check(dev.lib.mbx_create_bin(dev.ctx, C.byref(ref), n,
                                     (C.c_uint64 * nb)(*mins),
                                     (C.c_uint64 * nb)(*maxs), nb, outs, None))
        got = gpuq.bitmaps_host(dev, list(outs), n)
        np.testing.assert_array_equal(
            got, cpuref.create_bin(dim, mins, maxs, fk, dimsz), err_msg=f"{nb} bins")
        check(dev.lib.mbx_free(dev.ctx, outs[0]))
    dd.free()
    df.free()


def test_bins_fk_table_pool_full(dev):
    """A context whose fixed pool has no room for the dimension's bin table
    (2.5 MiB pool, 2.1 MiB of it context scratch; 1,000,000-row dimension)
    reads the dimension values directly and gives the same bitmaps."""
    rng = np.random.default_rng(17)
    n, dimsz = 300001, 1_000_000
    dim = rng.integers(0, 64, dimsz).astype(np.uint32)
    fk = rng.integers(0, dimsz + 5000, n).astype(np.uint32)
    dd, df = dev.upload(dim), dev.upload(fk)
    ctx = C.c_void_p()
    check(dev.lib.mbx_init(0, 5 << 19, C.byref(ctx)))
    try:
        for nb in (5, 64):
            mins, maxs = list(range(nb)), list(range(1, nb + 1))
            sw = stride_words(n)
            slab = dev.alloc(sw * 4 * nb)
            ref = abi.ColRef(dd.ptr, df.ptr, dimsz, 4)
            check(dev.lib.mbx_create_bin_into(ctx, C.byref(ref), n,
                                              (C.c_uint64 * nb)(*mins),
                                              (C.c_uint64 * nb)(*maxs), nb,
                                              slab.ptr, sw))
            check(dev.lib.mbx_sync(ctx))
            got = gpuq.bitmaps_host(dev, [slab.ptr + j * sw * 4 for j in range(nb)], n)
            np.testing.assert_array_equal(
                got, cpuref.create_bin(dim, mins, maxs, fk, dimsz), err_msg=f"{nb} bins")
            slab.free()
    finally:
        dev.lib.mbx_destroy(ctx)
    dd.free()
    df.free()


@pytest.mark.parametrize("n", [70001, 1_000_003])
def test_dense_candidate_ranges(dev, n):
    """Ranges where most rows are candidates take the word-by-word row op
    instead of the row list: two-valued (75% x 80% of rows) and three-valued
    (an edge bin over the head of the Zipf column a32 leaves most rows
    uncertain; AND NOT of a misaligned range), plain and through a foreign
    key, with LDS and HBM group tables, every plan == oracle."""
    rng = np.random.default_rng(n)
    schema = gpuq.upload_schema(dev, datagen.small_schema(n, 77))
    a16 = gpuq.exact_bitmap(dev, schema, S.Pred(S.Src("a16"), 0, 3000))
    b32 = gpuq.exact_bitmap(dev, schema, S.Pred(S.Src("b32"), 0, 40))
    b_a32 = list(range(0, 1001, 250))
    a32 = gpuq.create_bin(dev, schema, S.Src("a32"), b_a32[:-1], b_a32[1:])
    b_a8 = list(range(0, 257, 8))
    a8 = gpuq.create_bin(dev, schema, S.Src("a8"), b_a8[:-1], b_a8[1:])
    b_d32 = list(range(0, 61, 7))
    d32 = gpuq.create_bin(dev, schema, S.Src("d32", "fk"), b_d32[:-1], b_d32[1:])

    def binned(ptrs, bounds, lo, hi):
        mid, l, r = S.from_bin(bounds, lo, hi)
        return S.BitmapCol([ptrs[i] for i in mid], None if l is None else ptrs[l],
                           None if r is None else ptrs[r])

    O, A, M, N_ = abi.ORM, abi.ANDM, abi.AND, abi.NOT
    cases = [
        ([S.Pred(S.Src("a16"), 0, 3000), S.Pred(S.Src("b32"), 0, 40)],
         [S.BitmapCol([a16]), S.BitmapCol([b32])], [(O, 0, 0), (A, 0, 1)]),
        ([S.Pred(S.Src("a32"), 1, 500), S.Pred(S.Src("a8"), 10, 20)],
         [binned(a32, b_a32, 1, 500), binned(a8, b_a8, 10, 20)],
         [(O, 0, 0), (O, 1, 1), (N_, 1, 1), (M, 0, 1)]),
        ([S.Pred(S.Src("a32"), 3, 700), S.Pred(S.Src("d32", "fk"), 2, 30)],
         [binned(a32, b_a32, 3, 700), binned(d32, b_d32, 2, 30)],
         [(O, 0, 0), (A, 0, 1)]),
    ]
    for preds, cols, ins in cases:
        for extra in (0, 5000):
            keys, total = datagen.random_keys(rng)
            m, ma, mb = datagen.random_measure(rng)
            spec = S.QuerySpec(preds, ins, keys, total + extra, m, ma, mb)
            want, nsel = cpuref.query(spec, schema)
            assert nsel > n // 8, spec   # candidates dense enough to skip the row list
            for plan in (abi.PLAN_FUSED, abi.PLAN_FUSED | abi.PLAN_EAGER,
                         abi.PLAN_MATERIALISE, abi.PLAN_SCAN):
                got = gpuq.run_query(dev, schema, spec, cols, plan)
                np.testing.assert_array_equal(got, want, err_msg=f"plan {plan} {spec}")
            # the accounting kernel walks the same ranges: its count of
            # selected rows is the oracle's
            prg = S.to_vprg(cols, schema.nrows, spec.instrs)
            op = S.to_rowop(spec, schema)
            nb, ns = C.c_uint64(), C.c_uint64()
            check(dev.lib.mbx_alg_bytes(dev.ctx, C.byref(prg), C.byref(op), spec.nr_grp,
                                        C.byref(nb), C.byref(ns)))
            assert ns.value == nsel, spec
    for ptr in (a16, b32, a32[0], a8[0], d32[0]):
        check(dev.lib.mbx_free(dev.ctx, ptr))
    for c in schema.cols.values():
        c.dev.free()


@pytest.mark.parametrize("seed", range(16))
def test_random_queries_all_plans(dev, seed):
    """Random programs (AND/OR/NOT/ANDM/ORM, exact and edge bins), group
    counts from 1 to >4096 (LDS table and HBM atomics): fused (early-out,
    eager -- the conjunction form for two- and three-valued conjunctions --
    and the no-compaction and registers-in-memory ablations) == materialise
    == scan == oracle, bit-exact int64; the first stage with its registers
    in device memory writes the same bitmaps as the product one."""
    rng = np.random.default_rng(1000 + seed)
    n = int(rng.choice([1, 63, 64, 65, 5000, 70001, 300007]))
    schema = gpuq.upload_schema(dev, datagen.small_schema(n, seed))
    nw = ((n + 63) // 64) * 2
    poss = [dev.alloc(8 * nw, np.uint32, 2 * nw) for _ in range(4)]
    for _ in range(6):
        spec = datagen.random_spec(rng, schema, extra_groups=int(rng.integers(0, 3)))
        want, _ = cpuref.query(spec, schema)
        cols, allocs = gpuq.plan_cols(dev, schema, spec, rng)
        for plan in (abi.PLAN_FUSED, abi.PLAN_FUSED | abi.PLAN_EAGER,
                     abi.PLAN_FUSED | abi.PLAN_DIRECT,
                     abi.PLAN_FUSED | abi.PLAN_BASEFUSE,
                     abi.PLAN_MATERIALISE, abi.PLAN_SCAN):
            got = gpuq.run_query(dev, schema, spec, cols, plan)
            np.testing.assert_array_equal(got, want, err_msg=f"plan {plan} {spec}")
        prg = S.to_vprg(cols, n, spec.instrs)
        for i, fl in ((0, 0), (2, abi.PLAN_BASEFUSE)):
            check(dev.lib.mbx_prog_eval_ex(dev.ctx, C.byref(prg), poss[i].ptr,
                                           poss[i + 1].ptr, fl))
        dev.sync()
        np.testing.assert_array_equal(poss[2].host()[:nw], poss[0].host()[:nw])
        if any(c.leftmost or c.rightmost for c in cols):
            np.testing.assert_array_equal(poss[3].host()[:nw], poss[1].host()[:nw])
        gpuq.free_all(dev, allocs)
    for p in poss:
        p.free()
    for c in schema.cols.values():
        c.dev.free()


def test_dense_and_empty(dev):
    """All rows selected (many list rounds per tile) and none selected."""
    n = 1 << 21
    rng = np.random.default_rng(7)
    schema = gpuq.upload_schema(dev, datagen.small_schema(n, 7))
    for lo, hi in ((0, 256), (300, 400)):
        spec = S.QuerySpec([S.Pred(S.Src("a8"), lo, hi)], [(abi.ORM, 0, 0)],
                           [S.Key(S.Src("b32"))], 50, abi.MEAS_COL, S.Src("m32"))
        want, _ = cpuref.query(spec, schema)
        cols, allocs = gpuq.plan_cols(dev, schema, spec, rng, binned_frac=0)
        for plan in (abi.PLAN_FUSED, abi.PLAN_FUSED | abi.PLAN_EAGER,
                     abi.PLAN_FUSED | abi.PLAN_DIRECT,
                     abi.PLAN_MATERIALISE, abi.PLAN_SCAN):
            np.testing.assert_array_equal(gpuq.run_query(dev, schema, spec, cols, plan), want)
        gpuq.free_all(dev, allocs)


def test_errors(dev):
    """Misaligned bitmaps and bad programs fail with a status, not a fault."""
    n = 1000
    v = dev.upload(np.arange(n, dtype=np.uint32))
    prg = abi.Vprg()
    prg.factsz = n
    prg.cols[0].middle[0] = v.ptr + 4
    prg.instrs[0].opcode = abi.ORM
    op = abi.RowOp()
    out = dev.alloc(8)
    assert dev.lib.mbx_prog_grpby(dev.ctx, C.byref(prg), C.byref(op), 1,
                                  out.ptr, 0) == abi.ERR_ALIGN
    prg.cols[0].middle[0] = v.ptr
    prg.instrs[0].dstreg = 5
    assert dev.lib.mbx_prog_grpby(dev.ctx, C.byref(prg), C.byref(op), 1,
                                  out.ptr, 0) == abi.ERR_INVALID
    prg.instrs[0].dstreg = 0
    prg.factsz = 1 << 31
    assert dev.lib.mbx_prog_grpby(dev.ctx, C.byref(prg), C.byref(op), 1,
                                  out.ptr, 0) == abi.ERR_CAPACITY
    # the dedicated plan: a shape without a kernel (one clause of one
    # middle bin) is refused before any launch, as is its use with any plan
    # other than the fused one
    prg.factsz = n
    assert dev.lib.mbx_prog_grpby(dev.ctx, C.byref(prg), C.byref(op), 1,
                                  out.ptr, abi.PLAN_DEDICATED) == abi.ERR_UNSUPPORTED
    assert dev.lib.mbx_prog_grpby(
        dev.ctx, C.byref(prg), C.byref(op), 1, out.ptr,
        abi.PLAN_MATERIALISE | abi.PLAN_DEDICATED) == abi.ERR_INVALID
    v.free()
    out.free()


@pytest.fixture(scope="module")
def config1(dev):
    n = configs.C1_ROWS
    sch = configs.config1_device(dev, n)
    host = configs.config1_host(n)
    for k, a in host.items():
        sch[k].host = a
    yield sch
    for c in sch.cols.values():
        c.dev.free()


def test_config1_generator_matches_host(dev, config1):
    for k in ("p0", "p1", "p2", "m"):
        np.testing.assert_array_equal(config1[k].dev.host(), config1[k].host)


def test_config1_full(dev, config1):
    """Config 1 at full size: 40-bin bitmaps byte-identical to the oracle,
    the fused query (aligned and misaligned) bit-exact vs the raw-column
    oracle, invariants sum(groups) and count."""
    b = list(configs.C1_BOUNDS)
    bins = {}
    for c in configs.C1_COLS:
        bins[c] = gpuq.create_bin(dev, config1, S.Src(c), b[:-1], b[1:])
    got = gpuq.bitmaps_host(dev, bins["p1"], config1.nrows)
    np.testing.assert_array_equal(got, cpuref.create_bin(config1["p1"].host, b[:-1], b[1:]))
    for mis in (False, True):
        spec = configs.config1_spec(mis)
        want, nsel = cpuref.query(spec, config1)
        cols = []
        for p in spec.preds:
            mid, l, r = S.from_bin(b, p.lo, p.hi)
            ptrs = bins[p.src.col]
            cols.append(S.BitmapCol([ptrs[i] for i in mid],
                                    None if l is None else ptrs[l],
                                    None if r is None else ptrs[r]))
        for plan in (abi.PLAN_FUSED, abi.PLAN_FUSED | abi.PLAN_EAGER,
                     abi.PLAN_FUSED | abi.PLAN_DIRECT,
                     abi.PLAN_MATERIALISE, abi.PLAN_SCAN):
            got = gpuq.run_query(dev, config1, spec, cols, plan)
            np.testing.assert_array_equal(got, want, err_msg=f"{spec.name} plan {plan}")
        # COUNT version == number of selected rows
        cnt = S.QuerySpec(spec.preds, spec.instrs, [], 1, abi.MEAS_COUNT)
        assert gpuq.run_query(dev, config1, cnt, cols)[0] == nsel
        assert 0.005 < nsel / config1.nrows < 0.02
    for c in configs.C1_COLS:
        check(dev.lib.mbx_free(dev.ctx, bins[c][0]))


def test_multipass_launch_shapes(dev):
    """70M rows, past one resident pass: the eager conjunction form on 4
    words per lane over the resident grid (group table in LDS) and on 2
    words in even passes (table in HBM), two- and three-valued, next to the
    early-out form; every result equals the oracle."""
    n = 70_000_000
    sch = configs.config1_device(dev, n, seed=5)
    for c in sch.cols.values():
        c.host = c.dev.host()
    b = list(configs.C1_BOUNDS)
    bins = {c: gpuq.create_bin(dev, sch, S.Src(c), b[:-1], b[1:])
            for c in configs.C1_COLS}
    conj_shapes = {((abi.ORM, 0), (abi.ANDM, 0)),
                   ((abi.ORM, 0), (abi.ORM, 1), (abi.NOT, 1), (abi.AND, 0))}
    picked = {}
    for q in configs.config3_queries(160, seed=31):
        shape = tuple((op, d) for op, d, _ in q.instrs[:4])
        if not any(shape[:len(s)] == tuple(s) for s in conj_shapes):
            continue
        cols = configs.bin_cols(q, bins)
        chk = any(c.leftmost is not None or c.rightmost is not None for c in cols)
        kind = (chk, q.nr_grp > 4096)
        if len(picked.setdefault(kind, [])) < 2:
            picked[kind].append((q, cols))
    assert {(False, False), (True, False)} <= set(picked)
    # dense candidates at this size (the word-by-word row op with the words
    # kept in registers: three-valued at 4 words per lane, two-valued at 6):
    # an edge bin over the Zipf head of p0 (half the rows uncertain), and a
    # program that is not a conjunction, (p0 & p1) | p2, two-valued
    O, A, R = abi.ORM, abi.ANDM, abi.OR
    dense = [S.QuerySpec([S.Pred(S.Src("p0"), 3, 100), S.Pred(S.Src("p1"), 0, 100)],
                         [(O, 0, 0), (A, 0, 1)],
                         [S.Key(S.Src("p2"), abi.KEY_DIV, 100, 0)], 10,
                         abi.MEAS_COUNT, None, None, "dense-chk"),
             S.QuerySpec([S.Pred(S.Src("p0"), 0, 25), S.Pred(S.Src("p1"), 0, 25),
                          S.Pred(S.Src("p2"), 0, 100)],
                         [(O, 0, 0), (A, 0, 1), (O, 1, 2), (R, 0, 1)],
                         [S.Key(S.Src("p1"), abi.KEY_DIV, 250, 0)], 4,
                         abi.MEAS_COUNT, None, None, "dense-or")]
    picked["dense"] = [(q, configs.bin_cols(q, bins)) for q in dense]
    for kind, qs in sorted(picked.items(), key=str):
        for q, cols in qs:
            want, nsel = cpuref.query(q, sch)
            if kind == "dense":
                assert nsel > n // 5, q.name
            for plan in (abi.PLAN_FUSED | abi.PLAN_EAGER, abi.PLAN_FUSED):
                got = gpuq.run_query(dev, sch, q, cols, plan)
                np.testing.assert_array_equal(got, want,
                                              err_msg=f"{kind} {q.name} plan {plan}")
    for c in configs.C1_COLS:
        check(dev.lib.mbx_free(dev.ctx, bins[c][0]))
    for c in sch.cols.values():
        c.dev.free()
        c.host = None


def test_accumulate_and_group_count(dev):
    """MBX_ACCUMULATE adds into the output instead of zeroing it (a second
    run doubles every sum), for the LDS table and the HBM atomics; a query
    with no groups is refused with a status."""
    n = 200_003
    rng = np.random.default_rng(5)
    schema = gpuq.upload_schema(dev, datagen.small_schema(n, 5))
    for extra in (0, 5000):   # LDS table, then HBM atomics
        spec = datagen.random_spec(rng, schema, extra_groups=extra)
        want, _ = cpuref.query(spec, schema)
        cols, allocs = gpuq.plan_cols(dev, schema, spec, rng)
        prg = S.to_vprg(cols, n, spec.instrs)
        op = S.to_rowop(spec, schema)
        out = dev.alloc(8 * spec.nr_grp, np.int64, spec.nr_grp)
        for plan in (abi.PLAN_FUSED, abi.PLAN_FUSED | abi.PLAN_EAGER,
                     abi.PLAN_MATERIALISE, abi.PLAN_SCAN):
            check(dev.lib.mbx_prog_grpby(dev.ctx, C.byref(prg), C.byref(op),
                                         spec.nr_grp, out.ptr, plan))
            check(dev.lib.mbx_prog_grpby(dev.ctx, C.byref(prg), C.byref(op),
                                         spec.nr_grp, out.ptr,
                                         plan | abi.ACCUMULATE))
            dev.sync()
            np.testing.assert_array_equal(out.host(), 2 * want,
                                          err_msg=f"plan {plan} groups {spec.nr_grp}")
        assert dev.lib.mbx_prog_grpby(dev.ctx, C.byref(prg), C.byref(op), 0,
                                      out.ptr, abi.PLAN_FUSED) == abi.ERR_INVALID
        out.free()
        gpuq.free_all(dev, allocs)
    for c in schema.cols.values():
        c.dev.free()


@pytest.mark.parametrize("n,dimsz", [(1_000_003, 5_000), (333, 77), (2_097_152, 300_000)])
def test_create_bin_multi_equals_single(dev, n, dimsz):
    """Several attributes through one foreign-key column in one pass
    (mbx_create_bin_multi) write the same bytes as create_bin_into per
    attribute (whose bytes the oracle pins): unit, equal-width and sorted
    bins, a 300-bin set (16-bit index), 1,000 bins beside 5 and 25 (a
    few-bin set through ballots beside several LDS chunks), keys past the
    dimension; and the
    cases the packed path does not take (bins that are not consecutive,
    indexes wider than 32 bits together) fall back to one pass per set."""
    rng = np.random.default_rng(n)
    fk = rng.integers(0, dimsz + dimsz // 10 + 1, n).astype(np.uint32)
    a8 = rng.integers(0, 250, dimsz).astype(np.uint8)
    a16 = rng.integers(0, 300, dimsz).astype(np.uint16)
    a32 = rng.integers(0, 1 << 20, dimsz).astype(np.uint32)
    d_fk, d8, d16, d32 = (dev.upload(x) for x in (fk, a8, a16, a32))
    sw = (((n + 31) // 32 + 63) // 64) * 64
    sorted_b = sorted(set(int(x) for x in rng.integers(0, 1 << 20, 40))) + [1 << 20]
    cases = {
        "packed": [(d8, 1, list(range(250)), list(range(1, 251))),
                   (d16, 2, list(range(300)), list(range(1, 301))),
                   (d32, 4, sorted_b[:-1], sorted_b[1:])],
        "equal": [(d8, 1, list(range(0, 250, 10)), list(range(10, 260, 10))),
                  (d32, 4, list(range(0, 1 << 20, 1 << 16)),
                   list(range(1 << 16, (1 << 20) + 1, 1 << 16)))],
        "generic": [(d8, 1, [0, 5, 100], [3, 50, 250]),
                    (d16, 2, list(range(300)), list(range(1, 301)))],
        "mixed": [(d8, 1, [0, 50, 100, 150, 200], [50, 100, 150, 200, 250]),
                  (d16, 2, list(range(1000)), list(range(1, 1001))),
                  (d8, 1, list(range(0, 250, 10)), list(range(10, 260, 10)))],
        "wide": [(d16, 2, list(range(300)), list(range(1, 301))),
                 (d16, 2, list(range(0, 300, 1)), list(range(1, 301))),
                 (d8, 1, list(range(250)), list(range(1, 251)))],
    }
    try:
        for name, sets in cases.items():
            outs = [dev.alloc(sw * 4 * len(lo)) for _, _, lo, _ in sets]
            arr = (abi.BinSet * len(sets))()
            keep = []
            for i, ((col, eb, lo, hi), o) in enumerate(zip(sets, outs)):
                L, H = (C.c_uint64 * len(lo))(*lo), (C.c_uint64 * len(hi))(*hi)
                keep += [L, H]
                arr[i].data, arr[i].elem_bytes, arr[i].nbins = col.ptr, eb, len(lo)
                arr[i].mins = C.cast(L, C.POINTER(C.c_uint64))
                arr[i].maxs = C.cast(H, C.POINTER(C.c_uint64))
                arr[i].slab = o.ptr
            check(dev.lib.mbx_create_bin_multi(dev.ctx, d_fk.ptr, dimsz, n, arr,
                                               len(sets), sw))
            for (col, eb, lo, hi), o in zip(sets, outs):
                ref = abi.ColRef(col.ptr, d_fk.ptr, dimsz, eb)
                single = dev.alloc(sw * 4 * len(lo))
                check(dev.lib.mbx_create_bin_into(
                    dev.ctx, C.byref(ref), n, (C.c_uint64 * len(lo))(*lo),
                    (C.c_uint64 * len(hi))(*hi), len(lo), single.ptr, sw))
                dev.sync()
                np.testing.assert_array_equal(o.host(dtype=np.uint32),
                                              single.host(dtype=np.uint32),
                                              err_msg=f"{name} n={n}")
                single.free()
            for o in outs:
                o.free()
    finally:
        for x in (d_fk, d8, d16, d32):
            x.free()
