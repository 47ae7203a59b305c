"""Pins for the CPU oracle (oracle/cpuref.c) -- CPU only.

The reference ships no golden vectors or fixtures for this path (its only
checks are run-time GPU-vs-CPU comparisons, ssb/cpuref.c:331-344,
synth/cpuref.c:236-241), so every oracle function is pinned here by closed
forms, invariants and brute force written independently in numpy (npref.py),
plus the fixtures in tests/golden/.
"""
import datetime
import os

import numpy as np
import pytest

from merle_amd import abi
from merle_amd.spec import Key, Pred, QuerySpec, Schema, Src
from oracle import cpuref

import datagen
import npref

TAIL_SIZES = [1, 2, 31, 32, 33, 63, 64, 65, 127, 128, 129, 300, 1000, 4099]


def test_oracle_builds():
    assert cpuref.lib() is not None


@pytest.mark.parametrize("n", TAIL_SIZES)
@pytest.mark.parametrize("dt", [np.uint8, np.uint16, np.uint32, np.uint64])
def test_create_bin_closed_form(n, dt):
    """bit b of word k == [lo <= v(32k+b) < hi], tail words zero
    (bmpidx_create.cu:14-56, primitive.cuh:102-118)."""
    rng = np.random.default_rng(n * 7 + np.dtype(dt).itemsize)
    hi_v = min(np.iinfo(dt).max, 5000)
    v = rng.integers(0, hi_v, n).astype(dt)
    # consecutive bins and an arbitrary overlapping set
    bounds = np.unique(rng.integers(0, hi_v, 9))
    sets = [(bounds[:-1], bounds[1:]),
            (rng.integers(0, hi_v // 2, 5), rng.integers(hi_v // 4, hi_v, 5))]
    for mins, maxs in sets:
        out = cpuref.create_bin(v, list(map(int, mins)), list(map(int, maxs)))
        stride = out.shape[1]
        for j in range(len(mins)):
            want = npref.pack((v >= mins[j]) & (v < maxs[j]), stride)
            np.testing.assert_array_equal(out[j], want)


@pytest.mark.parametrize("n", [1, 33, 1000, 4099])
def test_create_bin_fk_int64max(n):
    """fk >= dimsz reads as INT64_MAX in create_bin (bmpidx_create.cu:26-27):
    a bin reaching past 2^63-1 catches those rows, others never do."""
    rng = np.random.default_rng(n)
    dim = rng.integers(0, 100, 37).astype(np.uint16)
    fk = rng.integers(0, 45, n).astype(np.uint32)
    mins = [0, 50, 1 << 62]
    maxs = [50, 100, (1 << 64) - 1]
    out = cpuref.create_bin(dim, mins, maxs, fk=fk, dimsz=37)
    eff = np.where(fk < 37, dim[np.minimum(fk, 36)].astype(np.uint64),
                   np.uint64(npref.INT64_MAX))
    for j in range(3):
        want = npref.pack((eff >= np.uint64(mins[j])) & (eff < np.uint64(maxs[j])),
                          out.shape[1])
        np.testing.assert_array_equal(out[j], want)
    assert npref.unpack(out[2], n).sum() == (fk >= 37).sum()


@pytest.mark.parametrize("n", [7, 64, 1001])
def test_create_join_closed_form(n):
    """create_select/create_join (bmpidx_create.cu:58-136): key >= dimsz -> 0."""
    rng = np.random.default_rng(n + 1)
    dim = rng.integers(0, 200, 50).astype(np.uint8)
    fk = rng.integers(0, 60, n).astype(np.uint32)
    out = cpuref.create_join(dim, 20, 120, fk=fk, dimsz=50)
    ok = fk < 50
    eff = np.where(ok, dim[np.minimum(fk, 49)], 0)
    want = npref.pack(ok & (eff >= 20) & (eff < 120), out.size)
    np.testing.assert_array_equal(out, want)
    col = rng.integers(0, 1 << 30, n).astype(np.uint32)
    out = cpuref.create_join(col, 1 << 20, 1 << 29)
    np.testing.assert_array_equal(
        out, npref.pack((col >= (1 << 20)) & (col < (1 << 29)), out.size))


def test_bin_invariants():
    """popcount(bin j) == histogram count; consecutive bins are disjoint and
    their union covers every row."""
    rng = np.random.default_rng(5)
    n = 12345
    v = np.minimum(rng.zipf(1.2, n) - 1, 999).astype(np.uint32)
    bounds = list(range(0, 1001, 25))
    out = cpuref.create_bin(v, bounds[:-1], bounds[1:])
    hist, _ = np.histogram(v, bins=bounds)
    pop = [int(npref.unpack(out[j], n).sum()) for j in range(40)]
    assert pop == list(hist)
    acc = np.zeros(out.shape[1], dtype=np.uint32)
    for j in range(40):
        assert not np.any(acc & out[j])
        acc |= out[j]
    assert npref.unpack(acc, n).all()


def _exact_cols(spec, schema):
    """One exact ("perfect") bitmap per predicate, built by numpy."""
    masks = npref.pred_masks(spec, schema)
    return [([npref.pack(m, ((schema.nrows + 31) // 32 + 63) // 64 * 64)], None, None)
            for m in masks], masks


@pytest.mark.parametrize("seed", range(12))
def test_vprg_nochk_equals_predicates(seed):
    """Word-wise program evaluation (primitive.cuh:201-252) == the program as
    a boolean expression over exact predicate masks."""
    rng = np.random.default_rng(100 + seed)
    n = int(rng.integers(1, 3000))
    schema = datagen.small_schema(n, seed)
    for _ in range(8):
        spec = datagen.random_spec(rng, schema)
        cols, masks = _exact_cols(spec, schema)
        p = cpuref.vprg_of(cols, n, spec.instrs)
        poss, unc = cpuref.vprg_eval(p, chk=0)
        want = npref.run_program(spec.instrs, masks, n)
        np.testing.assert_array_equal(npref.unpack(poss, n), want)
        assert not np.any(unc)
        # tail bits cleared even after NOT
        if n % 32:
            assert (int(poss[-1]) >> (n % 32)) == 0


def _edge_cols(rng, spec, schema):
    """Bins with a candidate check: middle bins hold some certain rows, edge
    bins hold the rest of the predicate's rows plus false positives."""
    n = schema.nrows
    stride = ((n + 31) // 32 + 63) // 64 * 64
    masks = npref.pred_masks(spec, schema)
    cols = []
    for m in masks:
        certain = m & (rng.random(n) < 0.6)
        extra = (~m) & (rng.random(n) < 0.2)
        edge = (m & ~certain) | extra
        mid = [npref.pack(certain, stride)] if rng.random() < 0.8 else []
        if not mid:
            edge = m | extra
        left = npref.pack(edge, stride)
        cols.append((mid, left, None))
    return cols, masks


@pytest.mark.parametrize("seed", range(12))
def test_vprg_chk_invariants(seed):
    """Three-valued evaluation (primitive.cuh:255-332, NOT corrected):
    truth <= possible, and possible-but-not-uncertain rows are true."""
    rng = np.random.default_rng(200 + seed)
    n = int(rng.integers(1, 3000))
    schema = datagen.small_schema(n, seed)
    for _ in range(8):
        spec = datagen.random_spec(rng, schema)
        cols, masks = _edge_cols(rng, spec, schema)
        p = cpuref.vprg_of(cols, n, spec.instrs)
        poss, unc = cpuref.vprg_eval(p, chk=1)
        P, U = npref.unpack(poss, n), npref.unpack(unc, n)
        T = npref.run_program(spec.instrs, masks, n)
        assert not np.any(T & ~P), "a true row is not possible"
        assert not np.any(P & ~U & ~T), "a certain row is false"
        assert not np.any(U & ~P)


@pytest.mark.parametrize("seed", range(10))
def test_query_equals_numpy_groupby(seed):
    """ref_query (raw columns) == an independent numpy keyed sum."""
    rng = np.random.default_rng(300 + seed)
    n = int(rng.integers(1, 5000))
    schema = datagen.small_schema(n, seed)
    for _ in range(6):
        spec = datagen.random_spec(rng, schema)
        got, sel = cpuref.query(spec, schema)
        want, wsel = npref.group_sum(spec, schema)
        np.testing.assert_array_equal(got, want)
        assert sel == wsel


@pytest.mark.parametrize("seed", range(10))
def test_vprg_grpby_equals_query(seed):
    """Bitmap-driven group-by (primitive.cuh:340-445) with uncertain-row
    re-checks == the raw-column answer, for exact and candidate-check bins."""
    rng = np.random.default_rng(400 + seed)
    n = int(rng.integers(1, 5000))
    schema = datagen.small_schema(n, seed)
    for _ in range(6):
        spec = datagen.random_spec(rng, schema)
        want, wsel = cpuref.query(spec, schema)
        for cols in (_exact_cols(spec, schema)[0], _edge_cols(rng, spec, schema)[0]):
            p = cpuref.vprg_of(cols, n, spec.instrs)
            got, sel = cpuref.vprg_grpby(p, spec, schema)
            np.testing.assert_array_equal(got, want)
            assert sel == wsel


def test_ssb_date_to_year_closed_form():
    """ssb/ssbdemo.h:10-39: day offsets from 1992-01-01 (1992, 1996 leap)."""
    d0 = datetime.date(1992, 1, 1)
    for day in range(0, 2600, 3):
        y = (d0 + datetime.timedelta(days=day)).year
        assert cpuref.lib().ref_ssb_date_to_year(day) == min(y, 1998) - 1992
    assert (datetime.date(1997, 12, 1) - d0).days == 2161
    assert (datetime.date(1994, 2, 11) - d0).days == 772


def _ssb_schema(n, seed):
    rng = np.random.default_rng(seed)
    s = Schema(n)
    ncust, npart = 300, 500
    s.add("loCustKey", rng.integers(0, ncust, n).astype(np.uint32))
    s.add("loPartKey", rng.integers(0, npart, n).astype(np.uint32))
    s.add("loOrderDate", rng.integers(0, 2557, n).astype(np.uint16))
    s.add("loExtendedPrice", rng.integers(0, 55450, n).astype(np.uint32))
    disc = rng.integers(0, 11, n).astype(np.uint8)
    s.add("loDiscount", disc)
    s.add("loRevenue", (s["loExtendedPrice"].host * (100 - disc.astype(np.uint32)) // 100).astype(np.uint32))
    s.add("loSupplyCost", rng.integers(0, 131072, n).astype(np.uint32))
    s.add("loQuantity", rng.integers(1, 51, n).astype(np.uint8))
    s.add("loSuppCity", rng.integers(0, 250, n).astype(np.uint8))
    s.add("custCity", rng.integers(0, 250, ncust).astype(np.uint8))
    s.add("partMfgr", rng.integers(0, 1000, npart).astype(np.uint16))
    return s


def _ssb_struct(s):
    r = cpuref.RefSSB()
    for name, _ in cpuref.RefSSB._fields_:
        setattr(r, name, s[name].host.ctypes.data)
    return r


YEAR_BOUNDS = (366, 731, 1096, 1461, 1827, 2192)


def test_ssb_flights_vs_descriptors():
    """The SSB flights restated as straight loops (ssb/cpuref.c:141-242) ==
    the same queries written as row-op descriptors (the encoding the SSB
    driver uses)."""
    import ctypes as C
    from merle_amd import ssb as ssbq
    n = 20000
    s = _ssb_schema(n, 11)
    L = cpuref.lib()
    st = _ssb_struct(s)
    for q in ssbq.FLIGHTS:
        spec = ssbq.flight_spec(q)
        want, _ = cpuref.query(spec, s)
        out = np.zeros(spec.nr_grp, dtype=np.int64)
        args = q.args
        if q.kind == 1:
            out[0] = L.ref_s1(C.byref(st), n, *args)
        elif q.kind == 2:
            L.ref_s2(C.byref(st), n, *args, out.ctypes.data)
        elif q.kind == 3:
            L.ref_s3(C.byref(st), n, *args, out.ctypes.data)
        else:
            L.ref_s4(C.byref(st), n, *args, out.ctypes.data)
        np.testing.assert_array_equal(out, want, err_msg=q.name)


def test_synth_ref_vs_descriptor():
    """synth_ref (synth/cpuref.c:125-169) == its descriptor form:
    (a1 in R OR a2 in R) AND dim1[fk] in R, COUNT by dim2[fk]."""
    import ctypes as C
    rng = np.random.default_rng(3)
    n, dimsz = 50000, 500
    a1 = np.minimum(rng.zipf(1.1, n) - 1, 65535).astype(np.uint16)
    a2 = np.minimum(rng.zipf(1.1, n) - 1, 65535).astype(np.uint16)
    fk = rng.integers(0, dimsz, n).astype(np.uint32)
    d1 = np.arange(dimsz, dtype=np.uint32)
    d2 = rng.integers(0, 256, dimsz).astype(np.uint8)
    want = np.zeros(256, dtype=np.int64)
    cpuref.lib().ref_synth(a1.ctypes.data, a2.ctypes.data, 16, fk.ctypes.data,
                           d1.ctypes.data, d2.ctypes.data, n, 3, 63, 5, 65,
                           20, 400, want.ctypes.data)
    s = Schema(n)
    for k, v in dict(a1=a1, a2=a2, fk=fk, d1=d1, d2=d2).items():
        s.add(k, v)
    spec = QuerySpec([Pred(Src("a1"), 3, 63), Pred(Src("a2"), 5, 65),
                      Pred(Src("d1", "fk"), 20, 400)],
                     [(abi.ORM, 0, 0), (abi.ORM, 0, 1), (abi.ANDM, 0, 2)],
                     [Key(Src("d2", "fk"))], 256, abi.MEAS_COUNT)
    got, _ = cpuref.query(spec, s)
    np.testing.assert_array_equal(got, want)
    assert want.sum() > 0


def test_mlp_grad_finite_difference():
    """Backward of the cost-model MLP vs central differences of its forward;
    forward vs a numpy matmul formulation."""
    rng = np.random.default_rng(9)
    n, din, dh, dout = 17, 16, 64, 3
    x = rng.normal(size=(n, din)).astype(np.float32)
    t = rng.normal(size=(n, dout)).astype(np.float32)
    W1 = (rng.normal(size=(dh, din)) * 0.3).astype(np.float32)
    b1 = (rng.normal(size=dh) * 0.1).astype(np.float32)
    W2 = (rng.normal(size=(dout, dh)) * 0.3).astype(np.float32)
    b2 = (rng.normal(size=dout) * 0.1).astype(np.float32)
    loss, g, y = cpuref.mlp_loss_grad(x, t, W1, b1, W2, b2)
    h = np.maximum(x.astype(np.float64) @ W1.T.astype(np.float64) + b1, 0)
    yy = h @ W2.T.astype(np.float64) + b2
    np.testing.assert_allclose(y, yy, rtol=1e-5, atol=1e-5)
    assert abs(loss - np.mean((yy - t) ** 2)) < 1e-9 * max(1, loss)
    # analytic backward in numpy
    pre = x.astype(np.float64) @ W1.T.astype(np.float64) + b1
    dy = 2 * (yy - t) / (n * dout)
    dpre = (dy @ W2.astype(np.float64)) * (pre > 0)
    for got, want in zip(g, (dpre.T @ x, dpre.sum(0), dy.T @ h, dy.sum(0))):
        np.testing.assert_allclose(got, want, rtol=1e-4, atol=1e-7)
    # central differences, away from ReLU kinks (a perturbation of eps can
    # move a pre-activation across 0 only if it is within eps*|x|_1 of it)
    params = [W1, b1, W2, b2]
    eps = 1e-3
    margin = eps * (np.abs(x).sum(1).max() + 1) * 4
    safe_h = np.all(np.abs(pre) > margin, axis=0)
    for pi, P in enumerate(params):
        flat = P.reshape(-1)
        cand = np.arange(flat.size)
        if pi in (0, 1):
            cand = cand[safe_h[cand // (din if pi == 0 else 1)]]
        for idx in rng.choice(cand, size=min(12, cand.size), replace=False):
            old = flat[idx]
            flat[idx] = old + eps
            lp, _, _ = cpuref.mlp_loss_grad(x, t, *params)
            flat[idx] = old - eps
            lm, _, _ = cpuref.mlp_loss_grad(x, t, *params)
            flat[idx] = old
            fd = (lp - lm) / (2 * eps)
            assert abs(fd - g[pi].reshape(-1)[idx]) < 2e-3 * max(1.0, abs(fd)), (pi, idx)


GOLDEN = os.path.join(os.path.dirname(__file__), "golden")


def test_golden_fixtures():
    """Committed fixtures (tests/golden/make_golden.py): small seeded inputs
    with expected bitmaps and group sums."""
    path = os.path.join(GOLDEN, "config1_small.npz")
    z = np.load(path)
    v = z["p0"]
    bounds = list(z["bounds"])
    out = cpuref.create_bin(v, bounds[:-1], bounds[1:])
    np.testing.assert_array_equal(out, z["bins_p0"])
    from merle_amd import configs
    spec = configs.config1_spec()
    s = Schema(v.size)
    for k in ("p0", "p1", "p2", "m"):
        s.add(k, z[k])
    got, _ = cpuref.query(spec, s)
    np.testing.assert_array_equal(got, z["sums"])


# ------------------------------------------------------------------ WAH
# The WAH baseline (merle/wahGpuCmprs.cu, merle/wahGpuDirect.cu).  Pinned by
# the group layout's closed form, hand-worked known answers, round trips,
# the verbatim AND, and -- where the reference's CPU codec
# (merle/host/wahCpu.c, wahConv.c) is built into oracle/_ref -- by that
# codec's own outputs.
WAHREF = os.path.join(os.path.dirname(cpuref.__file__), "_ref", "libwahref.so")


def _wahref():
    import ctypes as C
    if not os.path.exists(WAHREF):
        pytest.skip("oracle/_ref/libwahref.so not built (needs the reference tree)")
    L = C.CDLL(WAHREF)
    vp, sz = C.c_void_p, C.c_size_t
    L.wahCmprsCPU.argtypes = [vp, sz, vp]
    L.wahCmprsCPU.restype = sz
    L.wahAndCPU.argtypes = [vp] * 6
    L.wahAndCPU.restype = vp
    L.wahLongest1Fill.argtypes = [vp, sz]
    L.wahLongest1Fill.restype = sz
    return L


def _runs(rng, ng, p_lit=0.3, max_run=40, p_one=0.3):
    """Groups made of zero runs, one runs and random literals."""
    out = []
    while len(out) < ng:
        r = rng.random()
        if r < p_lit:
            out.append(int(rng.integers(1, 0x7fffffff)))
        elif r < p_lit + (1 - p_lit) * p_one:
            out += [0x7fffffff] * int(rng.integers(1, max_run))
        else:
            out += [0] * int(rng.integers(1, max_run))
    return np.array(out[:ng], np.uint32)


@pytest.mark.parametrize("n", [1, 30, 31, 32, 61, 62, 63, 64, 65, 1000, 12345])
def test_wah_groups_closed_form(n):
    """Row 31g+j is bit 30-j of group g (dbjoinFlatWah's ballot,
    wahGpuCmprs.cu:256-300), ceil(n/31) groups, bit 31 and rows >= n zero;
    the groups of a bitmap equal those of the predicate."""
    rng = np.random.default_rng(n)
    v = rng.integers(0, 100, n).astype(np.uint32)
    lo, hi = 20, 70
    g = cpuref.wah_select(v, lo, hi)
    assert len(g) == (n + 30) // 31
    want = np.zeros(len(g), np.uint32)
    for row in range(n):
        if lo <= v[row] < hi:
            want[row // 31] |= np.uint32(1 << (30 - row % 31))
    np.testing.assert_array_equal(g, want)
    assert not np.any(g >> 31)
    bm = cpuref.create_join(v, lo, hi)
    np.testing.assert_array_equal(cpuref.wah_groups(bm, n), want)
    # through a foreign key, keys past the dimension select nothing
    dim = rng.integers(0, 100, 50).astype(np.uint32)
    fk = rng.integers(0, 60, n).astype(np.uint32)
    gj = cpuref.wah_select(dim, lo, hi, fk=fk, dimsz=50)
    wj = np.zeros(len(g), np.uint32)
    for row in range(n):
        if fk[row] < 50 and lo <= dim[fk[row]] < hi:
            wj[row // 31] |= np.uint32(1 << (30 - row % 31))
    np.testing.assert_array_equal(gj, wj)


def test_wah_compress_known_answers():
    """wahCompress (wahGpuCmprs.cu:215-245) worked by hand: a run of zero
    or one groups is one fill carrying its length, a single one a fill of
    one, literals stay, a trailing zero run is dropped, the last group stays
    when it is not zero."""
    L, O = 0x7fffffff, 0x80000000
    cases = [
        ([0, 0, 0, 5, L, L, 0, 7, 0, 0], [O | 3, 5, 0xc0000002, O | 1, 7]),
        ([5], [5]), ([0], []), ([L], [0xc0000001]), ([0, 0, L], [O | 2, 0xc0000001]),
        ([L, L, L, 0], [0xc0000003]), ([9, 9, 9], [9, 9, 9]),
        ([0, 3, 0, 3, 0], [O | 1, 3, O | 1, 3]), ([], []),
    ]
    for groups, want in cases:
        got = cpuref.wah_compress(np.array(groups, np.uint32))
        assert list(got) == want, (groups, list(map(hex, got)))
        # scan: groups before each word, then all groups
        s = cpuref.wah_scan(got)
        cnt = [w & 0x3fffffff if w >> 31 else 1 for w in want]
        assert list(s) == list(np.concatenate([[0], np.cumsum(cnt)]).astype(int))


def test_wah_compress_round_trip_and_reference_codec():
    """decompress(compress(x)) == x (zero groups past the last word); the
    compression equals the reference's CPU codec (wahCmprsCPU,
    merle/host/wahConv.c:36-57) up to the trailing zero fill that the GPU
    compression drops."""
    rng = np.random.default_rng(5)
    ref = None
    try:
        ref = _wahref()
    except pytest.skip.Exception:
        pass
    for trial in range(40):
        ng = int(rng.integers(1, 3000))
        x = _runs(rng, ng, p_lit=float(rng.choice([0.02, 0.3, 0.9])))
        w = cpuref.wah_compress(x)
        np.testing.assert_array_equal(cpuref.wah_decompress(w, ng), x)
        s = cpuref.wah_scan(w)
        nz = np.nonzero(x)[0]
        assert s[-1] == (nz[-1] + 1 if len(nz) else 0)
        if ref is not None:
            buf = np.full(ng + 2, 0xffffffff, np.uint32)
            ref.wahCmprsCPU(x.ctypes.data, ng, buf.ctypes.data)
            k = int(np.argmax(buf == 0xffffffff))
            r = buf[:k]
            if len(r) and r[-1] >> 30 == 2:
                r = r[:-1]
            np.testing.assert_array_equal(w, r)


def test_wah_and_no1_against_verbatim_and():
    """wahAndNo1 (wahGpuDirect.cu:8-34): down's words kept, each literal
    ANDed with the up word covering its group; decompressed it is the AND
    of the decompressed inputs.  A down literal inside an up zero fill that
    starts before it reads zero (a search that took the next up word would
    AND it with a literal).  The reference's CPU AND (wahAndCPU,
    merle/host/wahCpu.c:23-78, fully merged output) decompresses to the
    same groups."""
    O = 0x80000000
    up = np.array([O | 5, 0x7ffffffe], np.uint32)
    down = np.array([O | 2, 0x12345, 0x33, 0x44, 0x55], np.uint32)
    got = cpuref.wah_and_no1(up, down)
    assert list(got) == [O | 2, O | 1, O | 1, O | 1, 0x55 & 0x7ffffffe]
    rng = np.random.default_rng(9)
    ref = None
    try:
        ref = _wahref()
    except pytest.skip.Exception:
        pass
    for trial in range(40):
        ng = int(rng.integers(1, 2000))
        a = _runs(rng, ng, p_lit=float(rng.choice([0.05, 0.5])))
        b = _runs(rng, ng, p_lit=float(rng.choice([0.05, 0.5])), p_one=0.0)
        b[b == 0x7fffffff] = 0x7ffffffe   # down: no fill of ones
        wa, wb = cpuref.wah_compress(a), cpuref.wah_compress(b)
        assert cpuref.wah_longest_1fill(wb) <= 1
        r = cpuref.wah_and_no1(wa, wb)
        assert len(r) == len(wb)
        np.testing.assert_array_equal(cpuref.wah_decompress(r, ng), a & b)
        if ref is not None:
            pa = np.concatenate([wa, [0, 0]]).astype(np.uint32)
            pb = np.concatenate([wb, [0, 0]]).astype(np.uint32)
            out = np.zeros(len(wa) + len(wb) + 4, np.uint32)
            if len(wa) == 0 or len(wb) == 0:
                continue
            end = ref.wahAndCPU(pa.ctypes.data, pa.ctypes.data + 4 * len(wa),
                                pb.ctypes.data, pb.ctypes.data + 4 * len(wb),
                                out.ctypes.data, out.ctypes.data + 4 * len(out))
            k = (end - out.ctypes.data) // 4
            np.testing.assert_array_equal(cpuref.wah_decompress(out[:k], ng), a & b)


def test_wah_longest_1fill_reference_codec():
    """wahLongest1Fill (merle/host/wahConv.c:6-20): consecutive one fills
    add up, a 0x7fffffff literal counts as a fill of one."""
    w = np.array([0xc0000003, 0xc0000002, 5, 0x7fffffff, 0xc0000004, 0x80000009,
                  0xc0000001], np.uint32)
    assert cpuref.wah_longest_1fill(w) == 5
    ref = _wahref()
    rng = np.random.default_rng(3)
    for _ in range(30):
        x = _runs(rng, int(rng.integers(1, 2000)))
        w = cpuref.wah_compress(x)
        assert cpuref.wah_longest_1fill(w) == ref.wahLongest1Fill(w.ctypes.data, len(w))
