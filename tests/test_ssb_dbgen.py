"""The SSB flights on data written by the reference's own generator.

ssb-dbgen-lite.c (the reference's deterministic SSB generator) is compiled
from its source into oracle/_ref/ by `make -C oracle ref` (git-ignored; the
tests skip where it is absent).  Its binary column files (4-byte header =
element size in bytes, ssb-dbgen-lite.c:192-200) are read through the
library's loader and the 13 flights (ssb/cpuref.c:244-302) are checked:

- CPU: the oracle's straight loops (s1ref..s4ref restated) == the
  descriptor form the GPU path uses, on the generated scale-1 data;
- GPU: the reference's sparse and dense bin sets (ssb/bmpidx-create.cu:8-30)
  are built with create_bin (FK projection for customer city and part
  mfgr); each flight runs its Sparse, Medium and Dense programs of
  ssb/fixedlayout.cu (col::from_bin / from_bin2, edge bins -> candidate
  checks) as fused, eager, materialise and scan plans and through its
  dedicated kernel; every result equals the oracle's.
"""
import ctypes as C
import os
import subprocess

import numpy as np
import pytest

from merle_amd import abi, load, spec as S, ssb as ssbq
from merle_amd.host import check

from oracle import cpuref

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DBGEN = os.path.join(ROOT, "oracle", "_ref", "dbgen-lite")
FACT = ("loCustKey", "loPartKey", "loOrderDate", "loExtendedPrice",
        "loRevenue", "loSupplyCost", "loQuantity", "loDiscount", "loSuppCity")
DIMS = ("custCity", "partMfgr")


def _dbgen_dir(tmp_path_factory):
    if not os.path.exists(DBGEN) and os.path.isdir("/root/reference"):
        subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "oracle"), "ref"],
                       check=False, capture_output=True)
    if not os.path.exists(DBGEN):
        pytest.skip("oracle/_ref/dbgen-lite not built (needs the reference source)")
    d = tmp_path_factory.mktemp("ssb1")
    subprocess.run([DBGEN, "-s", "1", "-r", "7", "-b", "-d", str(d)], check=True,
                   capture_output=True)
    return str(d)


def _read(path):
    lib = load()
    p, nb, h = C.c_void_p(), C.c_size_t(), C.c_uint32()
    check(lib.mbx_colfile_read(path.encode(), C.byref(p), C.byref(nb), C.byref(h)))
    dt = {1: np.uint8, 2: np.uint16, 4: np.uint32, 8: np.uint64}[h.value]
    a = np.frombuffer(C.string_at(p.value, nb.value), dtype=dt).copy()
    lib.mbx_colfile_free(p)
    return a


@pytest.fixture(scope="module")
def ssb_host(tmp_path_factory):
    d = _dbgen_dir(tmp_path_factory)
    cols = {}
    for name in FACT + DIMS:
        a = _read(os.path.join(d, name))
        raw = np.fromfile(os.path.join(d, name), dtype=np.uint8)
        assert int.from_bytes(raw[:4].tobytes(), "little") == a.itemsize
        assert raw.size == 4 + a.nbytes
        cols[name] = a
    n = cols["loCustKey"].size
    assert n == 6_000_000
    s = S.Schema(n)
    for k, v in cols.items():
        s.add(k, v)
    return s


def _ssb_struct(s):
    r = cpuref.RefSSB()
    for name, _ in cpuref.RefSSB._fields_:
        setattr(r, name, s[name].host.ctypes.data)
    return r


def _oracle(s, q):
    """The flight as a straight loop (s1ref..s4ref restated)."""
    L = cpuref.lib()
    st = _ssb_struct(s)
    out = np.zeros(q.ngrp, dtype=np.int64)
    if q.kind == 1:
        out[0] = L.ref_s1(C.byref(st), s.nrows, *q.args)
    elif q.kind == 2:
        L.ref_s2(C.byref(st), s.nrows, *q.args, out.ctypes.data)
    elif q.kind == 3:
        L.ref_s3(C.byref(st), s.nrows, *q.args, out.ctypes.data)
    else:
        L.ref_s4(C.byref(st), s.nrows, *q.args, out.ctypes.data)
    return out


def test_dbgen_flights_loops_equal_descriptors(ssb_host):
    """On the reference generator's data: straight loops == descriptor form."""
    for q in ssbq.FLIGHTS:
        want = _oracle(ssb_host, q)
        got, _ = cpuref.query(ssbq.flight_spec(q), ssb_host)
        np.testing.assert_array_equal(got, want, err_msg=q.name)
        if q.kind == 1:
            assert want[0] > 0, q.name   # the flight selects rows at all


def test_dbgen_fixed_programs_equal_loops(ssb_host):
    """Each fixed-layout program, run by the oracle over the exact predicates
    of its columns, selects what the straight loop selects -- including
    flight 3's date < d as NOT(date in [d, 1999)) (fixedlayout.cu:93-97)."""
    for q in ssbq.FLIGHTS:
        want = _oracle(ssb_host, q)
        for fp in ssbq.fixed_plans(q):
            got, _ = cpuref.query(ssbq.fixed_spec(q, fp), ssb_host)
            np.testing.assert_array_equal(got, want, err_msg=f"{q.name} {fp.label}")


def test_fixed_plans_cover_every_range(lib):
    """The Sparse / Medium / Dense programs of ssb/fixedlayout.cu cover every
    predicate with the prebuilt bins alone: no column comes out empty, the
    middle bins lie inside the range and middle plus edges cover it."""
    for f in ssbq.FLIGHTS:
        plans = ssbq.fixed_plans(f)
        assert [p.label for p in plans] == ["Sparse", "Medium", "Dense"]
        for p in plans:
            for col, lo, hi, lvl in p.cols:
                r = ssbq.fixed_bins(col, lo, hi, lvl)
                assert r is not None, (f.name, p.label, col)
                b = ssbq.BIN_SETS[col][r[0]]
                mid, left, right = r[1]
                for i in mid:
                    assert lo <= b[i] and b[i + 1] <= hi
                covered = sorted(list(mid) + [j for j in (left, right) if j is not None])
                assert b[covered[0]] <= lo and b[covered[-1] + 1] >= hi
                assert all(b[j + 1] == b[k] for j, k in zip(covered, covered[1:]))


def test_fixed_programs_have_dedicated_kernels(lib):
    """Every SSB fixed-layout program, as written, is a conjunction of
    single-column clauses whose shape has a dedicated kernel
    (mbx_prog_shape, host only); the shape code spells out the program's
    bins: SSB11 Sparse = date (1 middle bin) AND discount (left and right
    edge bins) AND quantity (2 middle bins and a right edge bin)."""
    fake = {}

    def ptr(key):
        return fake.setdefault(key, 0x100000 * (len(fake) + 1))

    for f in ssbq.FLIGHTS:
        for fp in ssbq.fixed_plans(f):
            spec = ssbq.fixed_spec(f, fp)
            cols = []
            for cn, lo, hi, lvl in fp.cols:
                d, (mid, l, r) = ssbq.fixed_bins(cn, lo, hi, lvl)
                cols.append(S.BitmapCol([ptr((cn, d, j)) for j in mid],
                                        None if l is None else ptr((cn, d, l)),
                                        None if r is None else ptr((cn, d, r))))
            prg = S.to_vprg(cols, 6_000_000, spec.instrs)
            code, built = C.c_uint32(), C.c_int()
            check(lib.mbx_prog_shape(C.byref(prg), C.byref(code), C.byref(built)))
            assert code.value and built.value == 1, (f.name, fp.label, hex(code.value))
            # one clause per program column, in instruction order, the
            # negated one (flight 3's date < d) last
            clauses = [(code.value >> (6 * k)) & 63 for k in range(4)]
            assert sum(1 for c in clauses if c) == len(fp.cols)
            negated = any(op == abi.NOT for op, _, _ in spec.instrs)
            assert [bool(c & 32) for c in clauses if c][-1] == negated
            assert sum(bool(c & 32) for c in clauses) == int(negated)
            if f.name == "SSB11" and fp.label == "Sparse":
                assert clauses[:3] == [1, 8 | 16, 2 | 16]


@pytest.mark.gpu
def test_dbgen_flights_on_gpu(dev, ssb_host):
    """Every flight under its three fixed-layout programs and four plans."""
    s = ssb_host
    for c in s.cols.values():
        c.dev = dev.upload(c.host)
    srcs = {"loOrderDate": S.Src("loOrderDate"), "loDiscount": S.Src("loDiscount"),
            "loQuantity": S.Src("loQuantity"), "loSuppCity": S.Src("loSuppCity"),
            "custCity": S.Src("custCity", "loCustKey"),
            "partMfgr": S.Src("partMfgr", "loPartKey")}
    bins, allocs = {}, []
    for col, sets in ssbq.BIN_SETS.items():
        for d, b in enumerate(sets):
            ref = S.colref(s, srcs[col])
            nb = len(b) - 1
            outs = (C.c_void_p * nb)()
            single = C.c_size_t()
            check(dev.lib.mbx_create_bin(dev.ctx, C.byref(ref), s.nrows,
                                         (C.c_uint64 * nb)(*b[:-1]),
                                         (C.c_uint64 * nb)(*b[1:]), nb, outs,
                                         C.byref(single)))
            bins[col, d] = [outs[j] for j in range(nb)]
            allocs.append(outs[0])
    out = dev.alloc(8 * 1024, np.int64, 1024)
    nplans = 0
    for q in ssbq.FLIGHTS:
        want = _oracle(s, q)
        for fp in ssbq.fixed_plans(q):
            spec = ssbq.fixed_spec(q, fp)
            op = S.to_rowop(spec, s)
            cols = []
            for col, lo, hi, lvl in fp.cols:
                d, (mid, l, r) = ssbq.fixed_bins(col, lo, hi, lvl)
                ptrs = bins[col, d]
                cols.append(S.BitmapCol([ptrs[i] for i in mid],
                                        None if l is None else ptrs[l],
                                        None if r is None else ptrs[r]))
            prg = S.to_vprg(cols, s.nrows, spec.instrs)
            for plan in (abi.PLAN_FUSED, abi.PLAN_FUSED | abi.PLAN_EAGER,
                         abi.PLAN_MATERIALISE, abi.PLAN_SCAN,
                         abi.PLAN_FUSED | abi.PLAN_DEDICATED):
                check(dev.lib.mbx_prog_grpby(dev.ctx, C.byref(prg), C.byref(op),
                                             spec.nr_grp, out.ptr, plan))
                dev.sync()
                np.testing.assert_array_equal(out.host(spec.nr_grp), want,
                                              err_msg=f"{q.name} {fp.label} plan {plan}")
                nplans += 1
    assert nplans == 3 * 5 * len(ssbq.FLIGHTS)
    out.free()
    for p in allocs:
        check(dev.lib.mbx_free(dev.ctx, p))
    for c in s.cols.values():
        c.dev.free()
        c.dev = None


@pytest.mark.gpu
def test_demo_driver_report(ssb_host, tmp_path_factory):
    """The C++ driver (merle_amd/lib/mbx_demo, the reference's ssb_demoall
    on the mybmpidx mirror) runs every flight's Sparse / Medium / Dense
    programs and its Perfect / AllAlig / MisAlig controlled layouts
    (recipes built per query), method 1 (no compaction) against method 2 and
    the synthetic suite (perfect / many-OR / candidate-check recipes on
    seeded Zipf columns), checks each plan against the scan plan or the other
    method (exit status 0) and prints the five-block report.  The sums it
    writes for every SSB row and plan (the dedicated kernels' included)
    equal the oracle's straight loops (s1ref..s4ref restated) on the same
    data."""
    demo = os.path.join(ROOT, "merle_amd", "lib", "mbx_demo")
    if not os.path.exists(demo):
        pytest.skip("mbx_demo not built")
    d = _dbgen_dir(tmp_path_factory)
    sums = str(tmp_path_factory.mktemp("demo") / "sums.tsv")
    r = subprocess.run([demo, d, "3", "1000000", sums], capture_output=True,
                       text=True, timeout=300)
    assert r.returncode == 0, r.stderr
    blocks = r.stdout.strip("\n").split("\n\n")
    assert len(blocks) == 5
    # block 0: each flight's WAH index-only time (its predicates' WAH
    # bitmaps ANDed and the result decompressed, checked by the driver
    # against the Perfect program's bitmap); RTScan has no counterpart: nan
    wah = blocks[0].splitlines()
    assert wah[0].split("\t") == ["Case", "WAH", "RTScan"]
    assert [row.split("\t")[0] for row in wah[1:]] == [q.name for q in ssbq.FLIGHTS]
    assert all(float(row.split("\t")[1]) > 0 and row.split("\t")[2] == "nan"
               for row in wah[1:])
    # block 1: every bin set's creation time and size, for the library's
    # bitmaps and for their WAH compression
    idx = blocks[1].splitlines()
    assert idx[0].split("\t") == ["Column", "My(ms)", "WAH(ms)", "My(MB)", "WAH(MB)"]
    assert len(idx) == 1 + 2 * 7
    for row in idx[1:]:
        f = row.split("\t")
        assert len(f) == 5 and all(float(x) > 0 for x in f[1:]), row
        # compressed bitmaps are never larger than the plain ones here
        assert float(f[4]) <= float(f[3]) * 1.05, row
    # every SSB row's sums (scan, fused, materialise) == the straight loops
    want = {q.name: _oracle(ssb_host, q) for q in ssbq.FLIGHTS}
    seen = set()
    with open(sums) as fh:
        for line in fh:
            name, bt, plan, *vals = line.rstrip("\n").split("\t")
            np.testing.assert_array_equal(np.array(vals, np.int64), want[name],
                                          err_msg=f"{name} {bt} {plan}")
            seen.add((name, bt, plan))
    bts = ("Sparse", "Medium", "Dense", "Perfect", "AllAlig", "MisAlig")
    # the fixed layouts also run their dedicated kernels (every planned
    # shape of these programs on this data is built)
    assert seen == {(q.name, bt, p) for q in ssbq.FLIGHTS for bt in bts
                    for p in ("join", "ours", "stg")} | \
        {(q.name, bt, "ded") for q in ssbq.FLIGHTS for bt in bts[:3]}
    # method 1 / method 2 (synth/cpuref.c:194-205): 7 selectivities
    meth = blocks[3].splitlines()
    assert meth[0].split("\t") == ["Seletiv", "Method1", "Method2"]
    assert len(meth) == 1 + 7
    for row in meth[1:]:
        f = row.split("\t")
        assert len(f) == 3 and float(f[1]) > 0 and float(f[2]) > 0
    # the synthetic block (synth/cpuref.c:207-256): 10 cases x 10 selectivities
    syn = blocks[4].splitlines()
    assert syn[0].split("\t")[:4] == ["Skew", "BitW", "Seletiv", "Join"]
    assert len(syn) == 1 + 100
    for row in syn[1:]:
        f = row.split("\t")
        assert len(f) == 17 and f[1] in ("16", "32")
        assert all(float(x) > 0 for x in f[3:16])
    ssb_rows = blocks[2].splitlines()
    assert ssb_rows[0].split("\t")[:4] == ["Case", "BinType", "Join", "Ours"]
    nf = len(ssbq.FLIGHTS)
    assert len(ssb_rows) == 1 + 6 * nf
    for i, row in enumerate(ssb_rows[1:]):
        f = row.split("\t")
        fixed = i < 3 * nf
        assert f[0] == ssbq.FLIGHTS[(i if fixed else i - 3 * nf) // 3].name
        assert f[1] == (("Sparse", "Medium", "Dense") if fixed else
                        ("Perfect", "AllAlig", "MisAlig"))[i % 3]
        assert len(f) == 10 and float(f[3]) > 0
        assert all(float(x) > 0 for x in f[4:9])   # ablations and stages
        # dedicated kernels on the fixed layouts, none on the controlled
        # ones (as the reference's report)
        assert (float(f[9]) > 0) if fixed else f[9] == "0"
        # DOWORK times the scan on the fixed layouts' Sparse row only
        # (ssb/lambdas.cuh:157); the controlled rows print 0 there
        assert (float(f[2]) > 0) if fixed and i % 3 == 0 else f[2] == "0"
