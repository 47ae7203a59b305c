"""WAH compressed-bitmap baseline on the GPU (merle/wahGpuCmprs.cu,
merle/wahGpuDirect.cu) through the C ABI, bit-exact against the oracle
(oracle/cpuref.c, whose WAH functions are pinned in test_oracle.py)."""
import ctypes as C

import numpy as np
import pytest

from merle_amd import abi
from merle_amd.host import check
from oracle import cpuref

pytestmark = pytest.mark.gpu

O, ONES = 0x80000000, 0x7fffffff


def _runs(rng, ng, p_lit=0.3, max_run=40, p_one=0.3):
    out = []
    while len(out) < ng:
        r = rng.random()
        if r < p_lit:
            out.append(int(rng.integers(1, ONES)))
        elif r < p_lit + (1 - p_lit) * p_one:
            out += [ONES] * int(rng.integers(1, max_run))
        else:
            out += [0] * int(rng.integers(1, max_run))
    return np.array(out[:ng], np.uint32)


def _colref(arr, fk=None, dimsz=0):
    r = abi.ColRef()
    r.data = arr.ptr
    r.elem_bytes = arr.dtype.itemsize
    r.fk = fk.ptr if fk is not None else None
    r.dimsz = dimsz
    return r


def _compress(dev, groups_dev, ng):
    out = dev.alloc(4 * max(ng, 1), np.uint32, max(ng, 1))
    k = C.c_size_t()
    check(dev.lib.mbx_wah_compress(dev.ctx, groups_dev.ptr, ng, out.ptr,
                                   C.byref(k)))
    return out, k.value


def _scan(dev, wah_dev, n):
    ex = dev.alloc(4 * (n + 1), np.uint32, n + 1)
    check(dev.lib.mbx_wah_scan(dev.ctx, wah_dev.ptr if n else None, n, ex.ptr))
    return ex


@pytest.mark.parametrize("n", [1, 30, 31, 32, 63, 64, 1983, 1984, 1985, 70001,
                               300007])
@pytest.mark.parametrize("dt", [np.uint8, np.uint16, np.uint32, np.uint64])
def test_wah_select_and_groups(dev, n, dt):
    """Groups of a predicate (direct and through foreign keys, keys past
    the dimension included) and of the library's own bitmaps == oracle."""
    rng = np.random.default_rng(n + np.dtype(dt).itemsize)
    v = rng.integers(0, 200, n).astype(dt)
    lo, hi = 30, 90
    ng = int(dev.lib.mbx_wah_groups_of(n))
    assert ng == (n + 30) // 31
    dv = dev.upload(v)
    g = dev.alloc(4 * ng, np.uint32, ng)
    check(dev.lib.mbx_wah_select(dev.ctx, C.byref(_colref(dv)), n, lo, hi, g.ptr))
    want = cpuref.wah_select(v, lo, hi)
    np.testing.assert_array_equal(g.host(), want)
    # through a foreign key
    dim = rng.integers(0, 200, 777).astype(dt)
    fk = rng.integers(0, 800, n).astype(np.uint32)
    ddim, dfk = dev.upload(dim), dev.upload(fk)
    check(dev.lib.mbx_wah_select(dev.ctx, C.byref(_colref(ddim, dfk, 777)), n,
                                 lo, hi, g.ptr))
    np.testing.assert_array_equal(g.host(), cpuref.wah_select(dim, lo, hi, fk=fk,
                                                              dimsz=777))
    # from a bitmap of the library
    if dt == np.uint32:
        bm = C.c_void_p()
        check(dev.lib.mbx_create_select(dev.ctx, dv.ptr, 4, n, lo, hi,
                                        C.byref(bm)))
        g2 = dev.alloc(4 * ng, np.uint32, ng)
        check(dev.lib.mbx_wah_groups(dev.ctx, bm.value, n, g2.ptr))
        np.testing.assert_array_equal(g2.host(), want)
        check(dev.lib.mbx_free(dev.ctx, bm.value))
        g2.free()
    for a in (dv, g, ddim, dfk):
        a.free()


@pytest.mark.parametrize("seed", range(8))
def test_wah_compress_scan_decompress(dev, seed):
    """compress (run structure, trailing zero run dropped), the count scan
    and decompression == oracle; decompression to more groups pads zeros,
    to fewer truncates."""
    rng = np.random.default_rng(seed)
    for ng in (1, 2, 17, 4095, 4096, 4097, 100_003, int(rng.integers(5e5, 2e6))):
        x = _runs(rng, ng, p_lit=float(rng.choice([0.0, 0.02, 0.3, 0.9, 1.0])),
                  p_one=float(rng.choice([0.0, 0.3])))
        if seed == 0 and ng == 17:
            x[:] = 0
        dx = dev.upload(x)
        out, k = _compress(dev, dx, ng)
        want = cpuref.wah_compress(x)
        assert k == len(want), (ng, k, len(want))
        got = out.host(k)
        np.testing.assert_array_equal(got, want)
        ex = _scan(dev, out, k)
        np.testing.assert_array_equal(ex.host(), cpuref.wah_scan(want))
        for cap in (ng, ng + 100, max(1, ng // 2)):
            dec = dev.alloc(4 * cap, np.uint32, cap)
            check(dev.lib.mbx_wah_decompress(dev.ctx, out.ptr, ex.ptr, k, dec.ptr,
                                             cap))
            np.testing.assert_array_equal(dec.host(),
                                          cpuref.wah_decompress(want, cap))
            dec.free()
        for a in (dx, out, ex):
            a.free()


@pytest.mark.parametrize("n", [2048 * 4096 - 1, 2048 * 4096, 2048 * 4096 + 1,
                               9_000_001])
def test_wah_scan_both_forms(dev, n):
    """The count scan == oracle on both sides of the tile count up to which
    each tile sums its predecessors itself (2,048 tiles of 4,096 words);
    the words mix literals with fills of up to 300 groups."""
    rng = np.random.default_rng(n)
    w = rng.integers(1, ONES, n, dtype=np.uint32)
    fill = rng.random(n) < 0.1
    w[fill] = (O | (rng.integers(0, 2, n, dtype=np.uint32) << 30)
               | rng.integers(1, 300, n, dtype=np.uint32))[fill]
    dw = dev.upload(w)
    ex = _scan(dev, dw, n)
    np.testing.assert_array_equal(ex.host(), cpuref.wah_scan(w))
    dw.free()
    ex.free()


@pytest.mark.parametrize("seed", range(6))
def test_wah_and_no1(dev, seed):
    """up AND down in down's word structure == oracle (which decompresses
    to the verbatim AND), including a down literal inside an up 0-fill that
    starts before it and down words past the end of up."""
    rng = np.random.default_rng(100 + seed)
    cases = [(np.array([O | 5, 0x7ffffffe], np.uint32),
              np.array([O | 2, 0x12345, 0x33, 0x44, 0x55, 0x66, 0x77], np.uint32))]
    for _ in range(5):
        ng = int(rng.choice([50, 5000, 300_000]))
        a = _runs(rng, ng, p_lit=float(rng.choice([0.05, 0.5])))
        b = _runs(rng, int(ng * rng.choice([0.5, 1.0, 1.5])),
                  p_lit=float(rng.choice([0.05, 0.5])), p_one=0.0)
        cases.append((cpuref.wah_compress(a), cpuref.wah_compress(b)))
    for up, down in cases:
        if len(down) == 0 or len(up) == 0:
            continue
        du, dd = dev.upload(up), dev.upload(down)
        su, sd = _scan(dev, du, len(up)), _scan(dev, dd, len(down))
        out = dev.alloc(4 * len(down), np.uint32, len(down))
        check(dev.lib.mbx_wah_and_no1(dev.ctx, du.ptr, su.ptr, len(up), dd.ptr,
                                      sd.ptr, len(down), out.ptr))
        np.testing.assert_array_equal(out.host(), cpuref.wah_and_no1(up, down))
        for x in (du, dd, su, sd, out):
            x.free()


def _scan_skip(dev, wah_dev, n, nskip):
    ex = dev.alloc(4 * (n + 1), np.uint32, n + 1)
    sk = dev.alloc(4 * max(nskip, 1), np.uint32, max(nskip, 1))
    check(dev.lib.mbx_wah_scan_skip(dev.ctx, wah_dev.ptr if n else None, n,
                                    ex.ptr, sk.ptr, nskip))
    return ex, sk


def _wah_ref(w_dev, ex, sk, n, nskip):
    r = abi.Wah()
    r.words = w_dev.ptr if n else None
    r.excl = ex.ptr
    r.skip = sk.ptr if (sk is not None and nskip) else None
    r.n = n
    r.nskip = nskip if sk is not None else 0
    return r


def _with_empty_words(rng, w, k):
    """w with k words that count no group (0-fills of 0) at random places,
    some of them in one run longer than a skip window's 1,024 words."""
    at = np.sort(rng.integers(0, len(w) + 1, k))
    out = np.insert(w, at, np.uint32(O))
    i = int(rng.integers(0, len(out) + 1))
    return np.insert(out, i, np.full(1500, O, np.uint32))


def _skip_expected(excl, n, nskip):
    """Entry i: the last word starting at or before group 1,024 i (bit 31:
    it starts there), for i below nskip and the bitmap's groups / 1,024."""
    total = int(excl[n])
    m = min(nskip, -(-total // 1024))
    g = 1024 * np.arange(m, dtype=np.int64)
    u = np.searchsorted(excl[:n].astype(np.int64), g, side="right") - 1
    return (u.astype(np.uint32)
            | np.where(excl[u] == g, np.uint32(O), np.uint32(0))), m


def _skip_cases(rng):
    cases = [np.array([O | 1024, 0x1234, O | 3000, 0x55], np.uint32),
             np.array([0x77] * 1024 + [O | 0x40000000 | 2048, 0x9], np.uint32)]
    for ng in (1, 1023, 1024, 1025, 70_001, 2_000_003):
        x = _runs(rng, ng, p_lit=float(rng.choice([0.02, 0.3, 0.9])),
                  max_run=int(rng.choice([40, 5000])))
        cases.append(cpuref.wah_compress(x))
    return cases


@pytest.mark.parametrize("seed", range(3))
def test_wah_scan_skip(dev, seed):
    """The skip table: entry i = the last word starting at or before group
    1,024 i, bit 31 when it starts there (searchsorted on the oracle's
    scan), full, short and empty tables, words counting no group among the
    others; the scan beside it == oracle."""
    rng = np.random.default_rng(700 + seed)
    for w in _skip_cases(rng):
        if seed == 2:
            w = _with_empty_words(rng, w, 50)
        n = len(w)
        dw = dev.upload(w) if n else None
        want_ex = cpuref.wah_scan(w)
        full = -(-int(want_ex[n]) // 1024)
        assert dev.lib.mbx_wah_skip_len(int(want_ex[n])) == full
        for nskip in (full, full // 2, full + 3, 0):
            ex, sk = _scan_skip(dev, dw, n, nskip)
            np.testing.assert_array_equal(ex.host(), want_ex)
            want, m = _skip_expected(want_ex, n, nskip)
            np.testing.assert_array_equal(sk.host()[:m], want)
            ex.free()
            sk.free()
        if dw is not None:
            dw.free()


@pytest.mark.parametrize("seed", range(4))
def test_wah_decompress_skip(dev, seed):
    """Decompression through the skip table == oracle with full, short and
    no tables, words counting no group among the others (seeds 2, 3)."""
    rng = np.random.default_rng(800 + seed)
    for w in _skip_cases(rng):
        if len(w) == 0:
            continue
        if seed >= 2:
            w = _with_empty_words(rng, w, 50)
        n = len(w)
        dw = dev.upload(w)
        total = int(cpuref.wah_scan(w)[n])
        full = -(-total // 1024)
        for nskip in (full, full // 3, 0):
            ex, sk = _scan_skip(dev, dw, n, nskip)
            r = _wah_ref(dw, ex, sk if nskip else None, n, nskip)
            for cap in (total, total + 2000, max(1, total // 2)):
                dec = dev.alloc(4 * cap, np.uint32, cap)
                check(dev.lib.mbx_wah_decompress_skip(dev.ctx, C.byref(r),
                                                      dec.ptr, cap))
                np.testing.assert_array_equal(dec.host(),
                                              cpuref.wah_decompress(w, cap))
                dec.free()
            ex.free()
            sk.free()
        dw.free()


@pytest.mark.parametrize("seed", range(4))
def test_wah_and_no1_skip(dev, seed):
    """The AND by skip windows == oracle: both tables, a short table on
    either side, no up table, no down table (the word-mapped kernel), down
    1-fills of one group, down words past the end of up and words counting
    no group on either side (seeds 2, 3: runs of 1,500 of them push a
    window past its one-staging bounds)."""
    rng = np.random.default_rng(900 + seed)
    cases = [(np.array([O | 5, 0x7ffffffe], np.uint32),
              np.array([O | 2, 0x12345, O | 0x40000000 | 1, 0x44, O | 3000, 0x66],
                       np.uint32))]
    for _ in range(4):
        ng = int(rng.choice([50, 5000, 300_000, 2_000_003]))
        a = _runs(rng, ng, p_lit=float(rng.choice([0.05, 0.5, 0.95])),
                  max_run=int(rng.choice([40, 5000])))
        b = _runs(rng, int(ng * rng.choice([0.5, 1.0, 1.5])),
                  p_lit=float(rng.choice([0.05, 0.5, 0.95])), p_one=0.0,
                  max_run=int(rng.choice([40, 5000])))
        cases.append((cpuref.wah_compress(a), cpuref.wah_compress(b)))
    for up, down in cases:
        if len(down) == 0 or len(up) == 0:
            continue
        if seed >= 2:
            up = _with_empty_words(rng, up, 30)
            down = _with_empty_words(rng, down, 30)
        want = cpuref.wah_and_no1(up, down)
        du, dd = dev.upload(up), dev.upload(down)
        tu = -(-int(cpuref.wah_scan(up)[-1]) // 1024)
        td = -(-int(cpuref.wah_scan(down)[-1]) // 1024)
        for nu, nd in ((tu, td), (tu // 2, td), (tu, td // 2), (0, td), (tu, 0)):
            exu, sku = _scan_skip(dev, du, len(up), nu)
            exd, skd = _scan_skip(dev, dd, len(down), nd)
            ru = _wah_ref(du, exu, sku if nu else None, len(up), nu)
            rd = _wah_ref(dd, exd, skd if nd else None, len(down), nd)
            out = dev.alloc(4 * len(down), np.uint32, len(down))
            check(dev.lib.mbx_wah_and_no1_skip(dev.ctx, C.byref(ru), C.byref(rd),
                                               out.ptr))
            np.testing.assert_array_equal(out.host(), want, err_msg=str((nu, nd)))
            for x in (exu, sku, exd, skd, out):
                x.free()
        du.free()
        dd.free()


def test_wah_longest_1fill(dev):
    """Consecutive one fills add up; a 0x7fffffff literal counts as one."""
    rng = np.random.default_rng(7)
    arrays = [np.array([0xc0000003, 0xc0000002, 5, ONES, 0xc0000004, O | 9,
                        0xc0000001], np.uint32)]
    for _ in range(6):
        w = rng.choice(np.array([0xc0000001, 0xc0000007, ONES, O | 3, 12345],
                                np.uint32), int(rng.integers(1, 100_000)))
        arrays.append(w.astype(np.uint32))
    for w in arrays:
        dw = dev.upload(w)
        got = C.c_uint64()
        check(dev.lib.mbx_wah_longest_1fill(dev.ctx, dw.ptr, len(w),
                                            C.byref(got)))
        assert got.value == cpuref.wah_longest_1fill(w)
        dw.free()


def test_wah_full_size_chain(dev):
    """120M rows (3,870,968 groups): two predicates' groups, compressed,
    ANDed, scanned and decompressed == the oracle's groups of the AND."""
    n = 120_000_000
    rng = np.random.default_rng(11)
    a = rng.integers(0, 1000, n, dtype=np.uint32)
    b = rng.integers(0, 1000, n, dtype=np.uint32)
    ng = (n + 30) // 31
    da, db = dev.upload(a), dev.upload(b)
    ga, gb = (dev.alloc(4 * ng, np.uint32, ng) for _ in range(2))
    check(dev.lib.mbx_wah_select(dev.ctx, C.byref(_colref(da)), n, 0, 10, ga.ptr))
    check(dev.lib.mbx_wah_select(dev.ctx, C.byref(_colref(db)), n, 100, 400, gb.ptr))
    wa, ka = _compress(dev, ga, ng)
    wb, kb = _compress(dev, gb, ng)
    want_a = cpuref.wah_compress(cpuref.wah_select(a, 0, 10))
    want_b = cpuref.wah_compress(cpuref.wah_select(b, 100, 400))
    np.testing.assert_array_equal(wa.host(ka), want_a)
    np.testing.assert_array_equal(wb.host(kb), want_b)
    lf = C.c_uint64()
    check(dev.lib.mbx_wah_longest_1fill(dev.ctx, wa.ptr, ka, C.byref(lf)))
    assert lf.value <= 1
    sa, sb = _scan(dev, wa, ka), _scan(dev, wb, kb)
    r = dev.alloc(4 * ka, np.uint32, ka)
    check(dev.lib.mbx_wah_and_no1(dev.ctx, wb.ptr, sb.ptr, kb, wa.ptr, sa.ptr, ka,
                                  r.ptr))
    sr = _scan(dev, r, ka)
    dec = dev.alloc(4 * ng, np.uint32, ng)
    check(dev.lib.mbx_wah_decompress(dev.ctx, r.ptr, sr.ptr, ka, dec.ptr, ng))
    want = cpuref.wah_select(a, 0, 10) & cpuref.wah_select(b, 100, 400)
    np.testing.assert_array_equal(dec.host(), want)
    # the same chain through skip tables: the same words and groups
    nsk = int(dev.lib.mbx_wah_skip_len(ng))
    ea, ka_ = _scan_skip(dev, wa, ka, nsk)
    eb, kb_ = _scan_skip(dev, wb, kb, nsk)
    np.testing.assert_array_equal(ea.host(), sa.host())
    r2 = dev.alloc(4 * ka, np.uint32, ka)
    check(dev.lib.mbx_wah_and_no1_skip(dev.ctx, C.byref(_wah_ref(wb, eb, kb_, kb, nsk)),
                                       C.byref(_wah_ref(wa, ea, ka_, ka, nsk)), r2.ptr))
    np.testing.assert_array_equal(r2.host(), r.host())
    er, kr = _scan_skip(dev, r2, ka, nsk)
    check(dev.lib.mbx_fill_u32(dev.ctx, dec.ptr, 0xdeadbeef, ng))
    check(dev.lib.mbx_wah_decompress_skip(dev.ctx, C.byref(_wah_ref(r2, er, kr, ka, nsk)),
                                          dec.ptr, ng))
    np.testing.assert_array_equal(dec.host(), want)
    for x in (da, db, ga, gb, wa, wb, sa, sb, r, sr, dec, ea, ka_, eb, kb_, r2, er, kr):
        x.free()


def test_wah_errors(dev):
    """Misaligned arrays and sizes past what a fill can count are refused
    before any launch."""
    g = dev.alloc(4096, np.uint32, 1024)
    k = C.c_size_t()
    assert dev.lib.mbx_wah_compress(dev.ctx, g.ptr + 2, 10, g.ptr, C.byref(k)) \
        == abi.ERR_ALIGN
    assert dev.lib.mbx_wah_compress(dev.ctx, g.ptr, 1 << 30, g.ptr, C.byref(k)) \
        == abi.ERR_CAPACITY
    assert dev.lib.mbx_wah_scan(dev.ctx, g.ptr, 1 << 30, g.ptr) == abi.ERR_CAPACITY
    assert dev.lib.mbx_wah_and_no1(dev.ctx, None, g.ptr, 5, g.ptr, g.ptr, 5,
                                   g.ptr) == abi.ERR_INVALID
    # the skip-table entry points: a table of entries without a pointer,
    # misaligned tables, sizes past a fill's count, null arguments
    assert dev.lib.mbx_wah_scan_skip(dev.ctx, g.ptr, 10, g.ptr, None, 4) \
        == abi.ERR_INVALID
    assert dev.lib.mbx_wah_scan_skip(dev.ctx, g.ptr, 10, g.ptr, g.ptr + 2, 4) \
        == abi.ERR_ALIGN
    assert dev.lib.mbx_wah_scan_skip(dev.ctx, g.ptr, 10, g.ptr, g.ptr,
                                     1 << 30) == abi.ERR_CAPACITY
    ok = abi.Wah()
    ok.words, ok.excl, ok.skip, ok.n, ok.nskip = g.ptr, g.ptr, g.ptr, 5, 1
    bad = abi.Wah()
    bad.words, bad.excl, bad.skip, bad.n, bad.nskip = g.ptr, g.ptr, None, 5, 1
    mis = abi.Wah()
    mis.words, mis.excl, mis.skip, mis.n, mis.nskip = g.ptr, g.ptr + 2, g.ptr, 5, 1
    for up, dn in ((bad, ok), (ok, bad)):
        assert dev.lib.mbx_wah_and_no1_skip(dev.ctx, C.byref(up), C.byref(dn),
                                            g.ptr) == abi.ERR_INVALID
    assert dev.lib.mbx_wah_and_no1_skip(dev.ctx, C.byref(ok), C.byref(mis),
                                        g.ptr) == abi.ERR_ALIGN
    assert dev.lib.mbx_wah_and_no1_skip(dev.ctx, C.byref(ok), None, g.ptr) \
        == abi.ERR_INVALID
    assert dev.lib.mbx_wah_decompress_skip(dev.ctx, C.byref(bad), g.ptr, 10) \
        == abi.ERR_INVALID
    assert dev.lib.mbx_wah_decompress_skip(dev.ctx, C.byref(ok), None, 10) \
        == abi.ERR_INVALID
    assert dev.lib.mbx_wah_decompress_skip(dev.ctx, C.byref(mis), g.ptr, 10) \
        == abi.ERR_ALIGN
    g.free()


def test_wah_empty_operands(dev):
    """An empty up operand (an all-zero bitmap compresses to no words)
    turns every down literal into a 0-fill of one group; an empty down
    operand writes nothing; an empty WAH decompresses to zeros."""
    down = np.array([O | 3, 0x1234, ONES, 0xc0000001, 0x55], np.uint32)
    dd, sd = dev.upload(down), None
    sd = _scan(dev, dd, len(down))
    up = dev.alloc(64, np.uint32, 16)
    su = _scan(dev, up, 0)
    out = dev.alloc(4 * len(down), np.uint32, len(down))
    check(dev.lib.mbx_wah_and_no1(dev.ctx, up.ptr, su.ptr, 0, dd.ptr, sd.ptr,
                                  len(down), out.ptr))
    assert list(out.host()) == [O | 3, O | 1, O | 1, O | 1, O | 1]
    assert list(out.host()) == list(cpuref.wah_and_no1(np.zeros(0, np.uint32), down))
    check(dev.lib.mbx_wah_and_no1(dev.ctx, dd.ptr, sd.ptr, len(down), up.ptr,
                                  su.ptr, 0, out.ptr))
    g = dev.alloc(4 * 100, np.uint32, 100)
    check(dev.lib.mbx_fill_u32(dev.ctx, g.ptr, 7, 100))
    check(dev.lib.mbx_wah_decompress(dev.ctx, up.ptr, su.ptr, 0, g.ptr, 100))
    assert not g.host().any()
    for x in (dd, sd, up, su, out, g):
        x.free()
