"""CPU tests of the library's host side: exports, the pool allocator, plan
helpers (col::from_bin, recipe), generators, column files, error paths."""
import ctypes as C
import os
import re

import numpy as np
import pytest

import merle_amd
from merle_amd import abi, spec as S

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    names = set()
    for fn in os.listdir(os.path.join(ROOT, "include")):
        txt = open(os.path.join(ROOT, "include", fn)).read()
        txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
        txt = re.sub(r"//[^\n]*", "", txt)
        names |= set(re.findall(r"\b(mbx_[a-z0-9_]+)\s*\(", txt))
    return names


def test_every_declared_symbol_is_exported(lib):
    names = declared_symbols()
    assert len(names) > 30
    missing = [n for n in sorted(names) if not hasattr(lib, n)]
    assert not missing, missing


def test_abi_layout(lib):
    abi.check_layout(lib)


def test_no_device_is_an_error_not_a_crash(lib):
    n = C.c_int(-1)
    assert lib.mbx_device_count(C.byref(n)) == abi.OK
    if n.value == 0:
        ctx = C.c_void_p()
        assert lib.mbx_init(0, 0, C.byref(ctx)) == abi.ERR_NODEVICE
        assert b"no HIP device" in lib.mbx_last_error()


def _pool(lib, cap):
    p = C.c_void_p()
    assert lib.mbx_hostpool_create(cap, C.byref(p)) == abi.OK
    return p


def _stats(lib, p):
    v = [C.c_size_t() for _ in range(5)]
    lib.mbx_hostpool_stats(p, *[C.byref(x) for x in v])
    return [x.value for x in v]


def test_pool_alloc_free_coalesce(lib):
    """256-B alignment, exhaustion is a status, frees coalesce back to one
    block (dumb_pool.cuh:39-95 semantics, minus its 48-slot limit)."""
    p = _pool(lib, 1 << 20)
    ptrs = []
    for i in range(100):          # more than the reference's 48 live slots
        q = C.c_void_p()
        assert lib.mbx_hostpool_alloc(p, 1000 + i, C.byref(q)) == abi.OK
        assert q.value % 256 == 0
        ptrs.append(q.value)
    used, cap, peak, nlive, largest = _stats(lib, p)
    want_used = sum((1000 + i + 255) // 256 * 256 for i in range(100))
    assert nlive == 100 and used == want_used and cap == 1 << 20
    q = C.c_void_p()
    assert lib.mbx_hostpool_alloc(p, 1 << 20, C.byref(q)) == abi.ERR_OOM
    rng = np.random.default_rng(0)
    for i in rng.permutation(100):
        assert lib.mbx_hostpool_free(p, ptrs[i]) == abi.OK
    used, cap, peak, nlive, largest = _stats(lib, p)
    assert used == 0 and nlive == 0 and largest == cap and peak == want_used
    assert lib.mbx_hostpool_free(p, 12345) == abi.ERR_INVALID
    assert lib.mbx_hostpool_alloc(p, 1 << 20, C.byref(q)) == abi.OK
    lib.mbx_hostpool_destroy(p)


def test_pool_random_workload(lib):
    """Random alloc/free never overlaps live blocks."""
    p = _pool(lib, 1 << 22)
    rng = np.random.default_rng(1)
    live = {}
    for _ in range(3000):
        if live and rng.random() < 0.45:
            k = list(live)[rng.integers(len(live))]
            assert lib.mbx_hostpool_free(p, k) == abi.OK
            del live[k]
        else:
            sz = int(rng.integers(1, 50000))
            q = C.c_void_p()
            st = lib.mbx_hostpool_alloc(p, sz, C.byref(q))
            if st == abi.ERR_OOM:
                continue
            assert st == abi.OK
            live[q.value] = sz
        iv = sorted((a, a + s) for a, s in live.items())
        assert all(b[0] >= a[1] for a, b in zip(iv, iv[1:]))
    lib.mbx_hostpool_destroy(p)


def _covers(bounds, idxs):
    return set().union(*[set(range(bounds[i], bounds[i + 1])) for i in idxs]) if idxs else set()


def test_from_bin_properties(lib):
    """col::from_bin (bmpidx_create.cu:138-163), brute force over small bin
    sets: middle bins lie inside [min,max), middle+edges cover it, edges
    straddle its ends, at most 4 middle bins (else a capacity error)."""
    rng = np.random.default_rng(2)
    for _ in range(300):
        nb = int(rng.integers(1, 9))
        bounds = sorted(set(rng.integers(0, 40, nb + 1).tolist()))
        if len(bounds) < 2:
            continue
        for lo in range(-1, 42, 3):
            for hi in range(lo, 43, 4):
                r = S.from_bin(bounds, max(lo, 0), max(hi, 0))
                lo_, hi_ = max(lo, 0), max(hi, 0)
                want_range = set(range(lo_, hi_))
                if r is None:
                    ok_cover = bounds[0] <= lo_ and bounds[-1] >= hi_ and lo_ < hi_
                    if ok_cover:    # only a >4-middle overflow may fail
                        inside = [i for i in range(len(bounds) - 1)
                                  if bounds[i] >= lo_ and bounds[i + 1] <= hi_]
                        assert len(inside) > 3
                    continue
                mid, l, rr = r
                assert len(mid) <= 4
                assert _covers(bounds, mid) <= want_range
                edges = [x for x in (l, rr) if x is not None]
                assert want_range <= _covers(bounds, mid + edges)
                for e in edges:
                    s = _covers(bounds, [e])
                    assert s & want_range and s - want_range


def test_recipe_plans(lib):
    """recipe::perfect/many_or/candchk bin plans (bmpidx_create.cu:228-316):
    perfect and many_or bins partition [min,max) exactly; candchk's bins
    cover it and its edge bins stick out by step/2 (or 1)."""
    lo_a, hi_a = (C.c_uint32 * 6)(), (C.c_uint32 * 6)()
    slot, n = (C.c_int * 6)(), C.c_int()
    for mn in range(0, 30):
        for mx in range(mn + 1, mn + 40):
            for kind in (0, 1, 2):
                for col in (0, 1, 2):
                    assert lib.mbx_recipe_plan(kind, col, mn, mx, lo_a, hi_a,
                                               slot, C.byref(n)) == abi.OK
                    bins = [(lo_a[i], hi_a[i], slot[i]) for i in range(n.value)]
                    cover = set()
                    for a, b, _ in bins:
                        cover |= set(range(a, b))
                    rng_ = set(range(mn, mx))
                    if kind < 2 or col >= 2:
                        assert cover == rng_
                        assert sum(b - a for a, b, _ in bins) == mx - mn
                        assert all(s < 4 for _, _, s in bins)
                    else:
                        # one edge bin; the union is an interval that contains
                        # [min,max) and sticks out on one side by <= step/2 (or 1)
                        step = (mx - mn) // 3
                        assert rng_ <= cover
                        assert cover == set(range(min(cover), max(cover) + 1))
                        out = cover - rng_
                        lim = max(step // 2, 1)
                        if col == 0:
                            assert all(mn - lim <= v < mn for v in out) or \
                                all(v == mx for v in out)
                        else:
                            assert all(mx <= v < mx + lim for v in out) or \
                                all(v == mn - 1 for v in out)
                        edge = [b for b in bins if b[2] >= 4]
                        assert len(edge) == 1


# Known answers of the recipe bin plans, worked out by hand from
# bmpidx_create.cu:228-316 (step = (max - min) / 3 in integer arithmetic):
# many_or splits [min, max) into [min, min+s), [min+s, min+2s), [min+2s, max)
# (the remainder goes to the last bin); candchk widens column 0's first bin
# to the left by s/2 (down to 0: max(min, s/2) - s/2) as its leftmost edge
# bin, and column 1's last bin to the right by s/2 as its rightmost edge bin,
# whose middle bins then come first; a step of 0 gives one bin, widened by 1
# for candchk.  Slots: 0..3 middle, 4 leftmost, 5 rightmost.
RECIPE_CASES = [
    # kind, column, min, max, [(lo, hi, slot)]
    (0, 0, 3, 9, [(3, 9, 0)]),
    (0, 1, 0, 1, [(0, 1, 0)]),
    (1, 0, 10, 40, [(10, 20, 0), (20, 30, 1), (30, 40, 2)]),
    (1, 3, 10, 41, [(10, 20, 0), (20, 30, 1), (30, 41, 2)]),
    (1, 0, 10, 42, [(10, 20, 0), (20, 30, 1), (30, 42, 2)]),
    (1, 2, 5, 7, [(5, 7, 0)]),
    (1, 0, 0, 21, [(0, 7, 0), (7, 14, 1), (14, 21, 2)]),
    (2, 0, 10, 40, [(5, 20, 4), (20, 30, 0), (30, 40, 1)]),
    (2, 0, 10, 41, [(5, 20, 4), (20, 30, 0), (30, 41, 1)]),
    (2, 0, 2, 32, [(0, 12, 4), (12, 22, 0), (22, 32, 1)]),
    (2, 0, 0, 21, [(0, 7, 4), (7, 14, 0), (14, 21, 1)]),
    (2, 0, 100, 121, [(97, 107, 4), (107, 114, 0), (114, 121, 1)]),
    (2, 0, 4, 6, [(4, 7, 4)]),
    (2, 1, 10, 40, [(30, 45, 5), (10, 20, 0), (20, 30, 1)]),
    (2, 1, 10, 41, [(30, 46, 5), (10, 20, 0), (20, 30, 1)]),
    (2, 1, 0, 21, [(14, 24, 5), (0, 7, 0), (7, 14, 1)]),
    (2, 1, 4, 6, [(3, 6, 5)]),
    (2, 1, 0, 2, [(0, 2, 5)]),    # deviation: min - 1 clamped at 0
    (2, 2, 10, 40, [(10, 20, 0), (20, 30, 1), (30, 40, 2)]),
    (2, 3, 5, 7, [(5, 7, 0)]),
]


@pytest.mark.parametrize("kind,col,mn,mx,want", RECIPE_CASES)
def test_recipe_known_answers(lib, kind, col, mn, mx, want):
    lo_a, hi_a = (C.c_uint32 * 6)(), (C.c_uint32 * 6)()
    slot, n = (C.c_int * 6)(), C.c_int()
    assert lib.mbx_recipe_plan(kind, col, mn, mx, lo_a, hi_a, slot,
                               C.byref(n)) == abi.OK
    got = [(lo_a[i], hi_a[i], slot[i]) for i in range(n.value)]
    assert got == want


def test_generators(lib):
    """Deterministic, in range; Zipf(1.0) frequencies follow (v+1)^-1/H by a
    chi-square test; uniform is flat."""
    n = 400000
    a = np.empty(n, np.uint32)
    b = np.empty(n, np.uint32)
    assert lib.mbx_gen_zipf_u32_host(a.ctypes.data, n, 99, 1000, 1.0) == 0
    assert lib.mbx_gen_zipf_u32_host(b.ctypes.data, n, 99, 1000, 1.0) == 0
    np.testing.assert_array_equal(a, b)
    assert a.max() < 1000
    H = np.sum(1.0 / np.arange(1, 1001))
    assert abs(H - 7.485471) < 1e-6
    p = (1.0 / np.arange(1, 1001)) / H
    obs = np.bincount(a, minlength=1000)[:100]
    exp = p[:100] * n
    chi2 = np.sum((obs - exp) ** 2 / exp)
    assert chi2 < 160, chi2          # 99 dof; p ~ 1e-4 at 150
    assert lib.mbx_gen_uniform_u32_host(a.ctypes.data, n, 5, 10, 20) == 0
    assert a.min() == 10 and a.max() == 19
    cnt = np.bincount(a - 10)
    assert np.all(np.abs(cnt - n / 10) < 5 * np.sqrt(n / 10))
    # splitmix64 output on known inputs (public reference algorithm)
    def mix(z):
        M = (1 << 64) - 1
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9 & M
        z = (z ^ (z >> 27)) * 0x94D049BB133111EB & M
        return z ^ (z >> 31)
    for seed, i in ((0, 0), (1, 5), (123456789, 1 << 40)):
        want = mix((seed + (i + 1) * 0x9E3779B97F4A7C15) & ((1 << 64) - 1))
        assert lib.mbx_rand_at(seed, i) == want


def test_colfile_roundtrip(lib, tmp_path):
    """4-byte LE header (bytes for SSB, ssb-dbgen-lite.c:192-201; bits for
    the synthetic files, onegen_zipf.py:73-84), N from the file size."""
    for dt in (np.uint8, np.uint16, np.uint32, np.uint64):
        for kind, hdr in ((0, np.dtype(dt).itemsize), (1, 8 * np.dtype(dt).itemsize)):
            v = np.arange(1000, dtype=dt)
            path = str(tmp_path / f"c{kind}").encode()
            assert lib.mbx_colfile_write(path, v.ctypes.data, v.size, v.itemsize, kind) == 0
            raw = open(path, "rb").read()
            assert int.from_bytes(raw[:4], "little") == hdr
            assert raw[4:] == v.tobytes()
            p, nb, h = C.c_void_p(), C.c_size_t(), C.c_uint32()
            assert lib.mbx_colfile_read(path, C.byref(p), C.byref(nb), C.byref(h)) == 0
            assert nb.value == v.nbytes and h.value == hdr
            got = np.ctypeslib.as_array((C.c_uint8 * nb.value).from_address(p.value))
            assert got.tobytes() == v.tobytes()
            lib.mbx_colfile_free(p)
    assert lib.mbx_colfile_read(b"/nonexistent/x", C.byref(p), C.byref(nb), None) == abi.ERR_INVALID


def test_library_missing_fails_loudly(monkeypatch, tmp_path):
    import importlib
    monkeypatch.setattr(merle_amd.abi, "LIB_PATH", str(tmp_path / "nope.so"))
    monkeypatch.setattr(merle_amd, "_lib", None)
    with pytest.raises(merle_amd.LibraryMissing):
        merle_amd.load()


def _prg(ncols, instrs):
    """A vprg whose columns 0..ncols-1 each have one (fake, never read)
    middle bin, with the given instructions."""
    p = abi.Vprg()
    for c in range(ncols):
        p.cols[c].middle[0] = 0x1000 * (c + 1)
    p.factsz = 1 << 20
    for i, (op, d, s1) in enumerate(instrs):
        p.instrs[i].opcode, p.instrs[i].dstreg, p.instrs[i].src1 = op, d, s1
    return p


def test_prog_shape_codes(lib):
    """mbx_prog_shape: clause k of a conjunction of single-column clauses at
    bits 6k..6k+5 as middle bins | left << 3 | right << 4 | negated << 5,
    in the order the program reads them; 0 for other programs."""
    def prg_of(cols, ins):
        p = abi.Vprg()
        p.factsz = 1000
        nxt = [0x1000]

        def ptr():
            nxt[0] += 0x1000
            return nxt[0]
        for c, (nm, l, r) in enumerate(cols):
            for k in range(nm):
                p.cols[c].middle[k] = ptr()
            if l:
                p.cols[c].leftmost = ptr()
            if r:
                p.cols[c].rightmost = ptr()
        for i, (o, d, s1) in enumerate(ins):
            p.instrs[i].opcode, p.instrs[i].dstreg, p.instrs[i].src1 = o, d, s1
        return p

    def shape(p):
        code, built = C.c_uint32(), C.c_int()
        assert lib.mbx_prog_shape(C.byref(p), C.byref(code), C.byref(built)) == abi.OK
        return code.value, built.value

    ORM, ANDM = abi.ORM, abi.ANDM
    # a AND b AND c, clause order = instruction order
    p = prg_of([(1, 0, 0), (0, 1, 1), (2, 0, 1)],
               [(ORM, 0, 0), (ANDM, 0, 1), (ANDM, 0, 2)])
    assert shape(p) == (1 | (24 << 6) | (18 << 12), 1)   # SSB11 Sparse's shape
    p = prg_of([(1, 0, 0), (0, 1, 1), (2, 0, 1)],
               [(ORM, 0, 2), (ANDM, 0, 0), (ANDM, 0, 1)])
    assert shape(p)[0] == 18 | (1 << 6) | (24 << 12)
    # a AND b AND NOT c: the negated clause carries bit 5
    p = prg_of([(1, 0, 0), (4, 0, 0), (2, 0, 0)],
               [(ORM, 0, 0), (ANDM, 0, 1), (ORM, 1, 2), (abi.NOT, 1, 1),
                (abi.AND, 0, 1)])
    assert shape(p)[0] == 1 | (4 << 6) | ((2 | 32) << 12)
    # four clauses of four middle bins: a conjunction, no kernel built
    p = prg_of([(4, 0, 0)] * 4,
               [(ORM, 0, 0), (ANDM, 0, 1), (ANDM, 0, 2), (ANDM, 0, 3)])
    assert shape(p) == (4 | (4 << 6) | (4 << 12) | (4 << 18), 0)
    # (a OR b) AND c: a clause of two columns -- no shape code
    p = prg_of([(1, 0, 0), (1, 0, 0), (1, 0, 0)],
               [(ORM, 0, 0), (ORM, 0, 1), (ANDM, 0, 2)])
    assert shape(p) == (0, 0)
    # not a conjunction (a OR (b AND c)) -- no shape code
    p = prg_of([(1, 0, 0), (1, 0, 0), (1, 0, 0)],
               [(ORM, 1, 1), (ANDM, 1, 2), (ORM, 0, 0), (abi.OR, 0, 1)])
    assert shape(p) == (0, 0)


def test_plan_features_and_candidates(lib):
    """The cost-model features price a program by what it selects: the
    program run over the column densities as independent probabilities
    (AND: product, OR: a + b - ab, NOT: 1 - a), and the re-check share as
    edge[c] x selectivity / dens[c]; plan-kind flags are one-hot.  The
    candidates of a 3-column conjunction are the early-out plan for each of
    the 6 ordered pairs of leading columns (the rest sparsest first), then
    eager, materialise and scan in sparsest-first order."""
    def feats(prg, dens, edge, plan):
        f = (C.c_float * abi.NFEAT)()
        assert lib.mbx_plan_features_ex(C.byref(prg), (C.c_double * 4)(*dens),
                                        (C.c_double * 4)(*edge), 10,
                                        C.byref(plan), f) == abi.OK
        return np.array(f[:])

    lg = lambda x: np.log10(x) / 9 + 1
    conj = _prg(3, [(abi.ORM, 0, 0), (abi.ANDM, 0, 1), (abi.ANDM, 0, 2)])
    dens, edge = [0.5, 0.2, 0.4, 0], [0.1, 0, 0, 0]
    plan = abi.Plan()
    assert lib.mbx_plan_heuristic(C.byref(conj), (C.c_double * 4)(*dens),
                                  C.byref(plan)) == abi.OK
    assert list(plan.order[:plan.norder]) == [1, 2, 0]
    f = feats(conj, dens, edge, plan)
    np.testing.assert_allclose(f[:4], [0.2, 0.4, 0.5, 0], rtol=1e-6)
    np.testing.assert_allclose(f[4:8], [0.2, 0.08, 0.04, 0], rtol=1e-6)
    np.testing.assert_allclose(f[12], lg(0.04), rtol=1e-5)
    np.testing.assert_allclose(f[13], lg(0.1 * 0.04 / 0.5), rtol=1e-5)
    # ANDNOT: c0 AND NOT c1 with an empty c1 selects all of c0
    andnot = _prg(2, [(abi.ORM, 0, 0), (abi.ORM, 1, 1), (abi.NOT, 1, 1),
                      (abi.AND, 0, 1)])
    d2 = [0.9, 0.0, 0, 0]
    desc = abi.Plan()
    assert lib.mbx_plan_describe(C.byref(andnot), C.byref(desc)) == abi.OK
    f = feats(andnot, d2, [0.8, 0, 0, 0], desc)
    np.testing.assert_allclose(f[12], lg(0.9), rtol=1e-5)
    np.testing.assert_allclose(f[13], lg(0.8), rtol=1e-5)
    # OR: a + b - ab
    orp = _prg(2, [(abi.ORM, 0, 0), (abi.ORM, 0, 1)])
    f = feats(orp, [0.3, 0.5, 0, 0], [0, 0, 0, 0], desc)
    np.testing.assert_allclose(f[12], lg(0.3 + 0.5 - 0.15), rtol=1e-5)
    # candidates and the one-hot plan kinds
    cand, cnt = (abi.Plan * 32)(), C.c_int()
    assert lib.mbx_plan_candidates(C.byref(conj), (C.c_double * 4)(*dens), cand,
                                   32, C.byref(cnt)) == abi.OK
    got = [(list(c.order[:c.norder]), c.flags) for c in cand[:cnt.value]]
    assert got == [([1, 2, 0], 0), ([1, 0, 2], 0), ([2, 1, 0], 0),
                   ([2, 0, 1], 0), ([0, 1, 2], 0), ([0, 2, 1], 0),
                   ([1, 2, 0], abi.PLAN_EAGER), ([1, 2, 0], abi.PLAN_MATERIALISE),
                   ([1, 2, 0], abi.PLAN_SCAN)]
    kinds = [feats(conj, dens, edge, c)[[11, 14, 15]] for c in cand[6:9]]
    np.testing.assert_array_equal(kinds, [[0, 1, 0], [1, 0, 0], [0, 0, 1]])
    assert lib.mbx_plan_candidates(C.byref(conj), (C.c_double * 4)(*dens), cand,
                                   8, C.byref(cnt)) == abi.ERR_CAPACITY
