"""The benchmark workloads as concrete, seeded synthetic inputs.

config 1  plain method: 12M-row fact table (scale 2 x 6M rows, the
          batchgen.sh:13 convention), columns p0..p2 ~ finite Zipf(s=1.0) on
          {0..999}, measure m ~ uniform [1, 100000]; 40 equal-width bins per
          predicate column; one conjunctive 3-predicate program, SUM(m) into
          int64 grouped by p2 (25 groups).
config 2  star schema: 120M-row fact table, int32 foreign keys into date
          (2,556), customer (600K), supplier (40K) and part (1.4M)
          dimensions; one bitmap per attribute value projected onto the fact
          rows; 12 SSB-flight-shaped queries (conjunctions of 2-4
          predicates, selectivity 1e-2..1e-5, up to 7x250x250 groups).
config 3  the config-1 tables, 256 random AND / OR / ANDNOT programs with
          log-uniform selectivity in [1e-6, 1e-1] and 1..10,000 groups.
config 5  serving: 600M-row fact table of 10 int32 columns z0..z9, finite
          Zipf(s=1.2) over 1,000..100,000 values (log-spaced), 64
          equal-width bucket bitmaps per column; requests drawn from the
          config-3 program space over these columns.
"""
import ctypes as C

import numpy as np

from . import abi, load
from .host import check
from .spec import Key, Pred, QuerySpec, Schema, Src

SEED = 0x6d65726c65  # "merle"


def col_seed(base, idx):
    return load().mbx_rand_at(base, 1000 + idx)


# ---------------------------------------------------------------- config 1
C1_ROWS = 12_000_000
C1_DOMAIN = 1000
C1_BOUNDS = tuple(range(0, 1001, 25))        # 40 bins of width 25
C1_COLS = ("p0", "p1", "p2")


def config1_spec(misaligned=False):
    """ORM p0 in [25,75) ; ANDM p1 in [25,75) ; ANDM p2 in [0,25);
    SUM(m) GROUP BY p2.  The misaligned variant asks p0 in [30,80), whose
    edge bins need a candidate check."""
    lo0, hi0 = (30, 80) if misaligned else (25, 75)
    preds = [Pred(Src("p0"), lo0, hi0), Pred(Src("p1"), 25, 75),
             Pred(Src("p2"), 0, 25)]
    ins = [(abi.ORM, 0, 0), (abi.ANDM, 0, 1), (abi.ANDM, 0, 2)]
    return QuerySpec(preds, ins, [Key(Src("p2"))], 25, abi.MEAS_COL, Src("m"),
                     None, "config1" + ("-misaligned" if misaligned else ""))


def config1_host(n=C1_ROWS, seed=SEED, s=1.0):
    """Host copy of the config-1 columns (the same values the device
    generator writes)."""
    lib = load()
    out = {}
    for i, name in enumerate(C1_COLS):
        a = np.empty(n, np.uint32)
        check(lib.mbx_gen_zipf_u32_host(a.ctypes.data, n, col_seed(seed, i),
                                        C1_DOMAIN, s))
        out[name] = a
    m = np.empty(n, np.uint32)
    check(lib.mbx_gen_uniform_u32_host(m.ctypes.data, n, col_seed(seed, 9), 1,
                                       100001))
    out["m"] = m
    return out


def config1_device(dev, n=C1_ROWS, seed=SEED, s=1.0):
    """Device columns of config 1 -> Schema (host copies left empty)."""
    lib = dev.lib
    sch = Schema(n)
    for i, name in enumerate(C1_COLS):
        a = dev.alloc(4 * n, np.uint32, n)
        check(lib.mbx_gen_zipf_u32(dev.ctx, a.ptr, n, col_seed(seed, i),
                                   C1_DOMAIN, s))
        sch.add(name, None, a)
    m = dev.alloc(4 * n, np.uint32, n)
    check(lib.mbx_gen_uniform_u32(dev.ctx, m.ptr, n, col_seed(seed, 9), 1,
                                  100001))
    sch.add("m", None, m)
    dev.sync()
    return sch


# ---------------------------------------------------------------- config 2
C2_ROWS = 120_000_000
C2_DIMS = {"date": 2556, "cust": 600_000, "supp": 40_000, "part": 1_400_000}
# dimension attribute -> (dimension, fact fk column, cardinality)
C2_ATTRS = {
    "d_year": ("date", "lo_date", 7),
    "c_region": ("cust", "lo_cust", 5), "c_nation": ("cust", "lo_cust", 25),
    "c_city": ("cust", "lo_cust", 250),
    "s_region": ("supp", "lo_supp", 5), "s_nation": ("supp", "lo_supp", 25),
    "s_city": ("supp", "lo_supp", 250),
    "p_mfgr": ("part", "lo_part", 5), "p_category": ("part", "lo_part", 25),
    "p_brand": ("part", "lo_part", 1000),
}
YEAR_DAY0 = (0, 366, 731, 1096, 1461, 1827, 2192)   # 1992..1998 (ssbdemo.h)


def _src(attr):
    return Src(attr, C2_ATTRS[attr][1])


def _eq(attr, v):
    return Pred(_src(attr), v, v + 1)


def _key(attr, base, stride, div=1):
    kind = abi.KEY_DIV if div > 1 else abi.KEY_IDENT
    return Key(_src(attr), kind, div, base, stride)


REV, COST = Src("lo_revenue"), Src("lo_supplycost")
O, A, M, N_, R = abi.ORM, abi.ANDM, abi.AND, abi.NOT, abi.OR


def config2_queries():
    """12 queries modeled on the SSB flights (ssb/cpuref.c:252-298).  Each
    predicate i is program column i; every bin is exact (one per value)."""
    Q = []
    # flight 1: no group-by, SUM(revenue)
    Q.append(QuerySpec([_eq("d_year", 1), _eq("c_nation", 7)],
                       [(O, 0, 0), (A, 0, 1)], [], 1, abi.MEAS_COL, REV,
                       None, "Q1.1"))
    Q.append(QuerySpec([_eq("d_year", 2), _eq("c_nation", 7), _eq("s_region", 2)],
                       [(O, 0, 0), (A, 0, 1), (A, 0, 2)], [], 1, abi.MEAS_COL,
                       REV, None, "Q1.2"))
    # flight 2: part x supplier, GROUP BY year, brand
    Q.append(QuerySpec([_eq("p_category", 12), _eq("s_region", 1)],
                       [(O, 0, 0), (A, 0, 1)],
                       [_key("d_year", 0, 40), _key("p_brand", 480, 1)], 280,
                       abi.MEAS_COL, REV, None, "Q2.1"))
    Q.append(QuerySpec([Pred(_src("p_brand"), 521, 525), _eq("s_region", 2)],
                       [(O, 0, 0), (A, 0, 1)],
                       [_key("d_year", 0, 4), _key("p_brand", 521, 1)], 28,
                       abi.MEAS_COL, REV, None, "Q2.2"))
    Q.append(QuerySpec([_eq("p_brand", 777), _eq("s_nation", 3)],
                       [(O, 0, 0), (A, 0, 1)], [_key("d_year", 0, 1)], 7,
                       abi.MEAS_COL, REV, None, "Q2.3"))
    # flight 3: customer x supplier x date; year != 1998 as NOT + AND
    # (the reference's program for Q3.x, ssb/fixedlayout.cu:93-97)
    Q.append(QuerySpec([_eq("c_region", 2), _eq("s_nation", 11), _eq("d_year", 6)],
                       [(O, 0, 0), (A, 0, 1), (O, 1, 2), (N_, 1, 1), (M, 0, 1)],
                       [_key("c_nation", 10, 70), _key("s_city", 110, 7),
                        _key("d_year", 0, 1)], 350, abi.MEAS_COL, REV, None,
                       "Q3.1"))
    Q.append(QuerySpec([_eq("c_nation", 14), _eq("s_nation", 14), _eq("d_year", 6)],
                       [(O, 0, 0), (A, 0, 1), (O, 1, 2), (N_, 1, 1), (M, 0, 1)],
                       [_key("c_city", 140, 70), _key("s_city", 140, 7),
                        _key("d_year", 0, 1)], 700, abi.MEAS_COL, REV, None,
                       "Q3.2"))
    Q.append(QuerySpec([Pred(_src("c_city"), 141, 143),
                        Pred(_src("s_city"), 141, 143)],
                       [(O, 0, 0), (A, 0, 1)],
                       [_key("c_city", 141, 14), _key("s_city", 141, 7),
                        _key("d_year", 0, 1)], 28, abi.MEAS_COL, REV, None,
                       "Q3.3"))
    Q.append(QuerySpec([Pred(_src("c_city"), 141, 143),
                        Pred(_src("s_city"), 141, 143), _eq("d_year", 5)],
                       [(O, 0, 0), (A, 0, 1), (A, 0, 2)],
                       [_key("c_city", 141, 2), _key("s_city", 141, 1)], 4,
                       abi.MEAS_COL, REV, None, "Q3.4"))
    # flight 4: SUM(revenue - supplycost)
    Q.append(QuerySpec([_eq("c_region", 1), _eq("s_nation", 6),
                        Pred(_src("p_mfgr"), 0, 2)],
                       [(O, 0, 0), (A, 0, 1), (A, 0, 2)],
                       [_key("d_year", 0, 5), _key("c_nation", 5, 1)], 35,
                       abi.MEAS_SUB, REV, COST, "Q4.1"))
    Q.append(QuerySpec([_eq("c_region", 1), _eq("s_region", 1),
                        Pred(_src("p_mfgr"), 0, 2), Pred(_src("d_year"), 5, 7)],
                       [(O, 0, 0), (A, 0, 1), (A, 0, 2), (A, 0, 3)],
                       [_key("d_year", 0, 62500), _key("c_city", 0, 250),
                        _key("s_city", 0, 1)], 437_500, abi.MEAS_SUB, REV, COST,
                       "Q4.2"))
    Q.append(QuerySpec([Pred(_src("d_year"), 5, 7), _eq("c_region", 1),
                        _eq("s_nation", 6), _eq("p_category", 3)],
                       [(O, 0, 0), (A, 0, 1), (A, 0, 2), (A, 0, 3)],
                       [_key("d_year", 5, 400), _key("s_city", 60, 40),
                        _key("p_brand", 120, 1)], 800, abi.MEAS_SUB, REV, COST,
                       "Q4.3"))
    return Q


def config2_device(dev, n=C2_ROWS, seed=SEED):
    """Device tables of config 2 -> Schema.  Values are uniform; nested
    attributes follow ssb-dbgen-lite.c:24-27 (city -> nation = city/10 ->
    region = nation/5) and brand -> category = brand/40 -> mfgr = brand/200;
    revenue and supply cost follow the ranges of ssb-dbgen-lite.c:132-137."""
    lib = dev.lib
    sch = Schema(n)

    def gen(name, count, lo, hi, idx):
        a = dev.alloc(4 * count, np.uint32, count)
        check(lib.mbx_gen_uniform_u32(dev.ctx, a.ptr, count, col_seed(seed, idx),
                                      lo, hi))
        sch.add(name, None, a)
        return a

    gen("lo_date", n, 0, C2_DIMS["date"], 20)
    gen("lo_cust", n, 0, C2_DIMS["cust"], 21)
    gen("lo_supp", n, 0, C2_DIMS["supp"], 22)
    gen("lo_part", n, 0, C2_DIMS["part"], 23)
    gen("lo_revenue", n, 0, 55450, 24)
    gen("lo_supplycost", n, 0, 131072, 25)
    # date dimension: year of each day offset
    days = np.arange(C2_DIMS["date"], dtype=np.int64)
    year = (np.searchsorted(np.array(YEAR_DAY0[1:]), days, side="right")).astype(np.uint32)
    sch.add("d_year", year, dev.upload(year))
    for pre, dim, idx in (("c", "cust", 26), ("s", "supp", 27)):
        city = gen(f"{pre}_city", C2_DIMS[dim], 0, 250, idx)
        for name, div in ((f"{pre}_nation", 10), (f"{pre}_region", 50)):
            a = dev.alloc(4 * C2_DIMS[dim], np.uint32, C2_DIMS[dim])
            check(lib.mbx_gen_map_u32(dev.ctx, a.ptr, city.ptr, C2_DIMS[dim], div, 0))
            sch.add(name, None, a)
    brand = gen("p_brand", C2_DIMS["part"], 0, 1000, 28)
    for name, div in (("p_category", 40), ("p_mfgr", 200)):
        a = dev.alloc(4 * C2_DIMS["part"], np.uint32, C2_DIMS["part"])
        check(lib.mbx_gen_map_u32(dev.ctx, a.ptr, brand.ptr, C2_DIMS["part"], div, 0))
        sch.add(name, None, a)
    dev.sync()
    # Dimension attributes are stored in the narrowest unsigned type that
    # holds their values, as the reference generator stores its own
    # (uint8 cities, uint16 part codes, ssb-dbgen-lite.c:16,42-43): the
    # 1,000 brands in 2 bytes, every other attribute in 1.  The 1.4M-row
    # part table then fits an XCD's 4 MB L2 (2.8 MB instead of 5.6).
    for name in C2_ATTRS:
        c = sch[name]
        h = c.host if c.host is not None else c.dev.host()
        hn = h.astype(np.uint8 if int(h.max()) < 256 else np.uint16)
        if c.dev is not None:
            c.dev.free()
        c.dev = dev.upload(hn)
        c.host = hn if c.host is not None else None
    return sch


def config2_index(dev, sch, attrs=None, timer=None):
    """One bitmap per attribute value, projected through the fact FK (bins
    [v, v+1)).  The attributes of one dimension are built in one pass over
    its foreign keys (mbx_create_bin_multi: c_region, c_nation and c_city
    from one read of lo_cust); a dimension with one attribute takes
    create_bin_into.  Slabs are allocated first, so a timer brackets only
    the creation (the reference times cudaMalloc too,
    ssb/bmpidx-create.cu:64-68).
    -> ({attr: [device ptr per value]}, [slabs], ms, algorithmic bytes:
    each key column once, plus every bitmap)."""
    from .host import stride_words
    from .spec import colref
    attrs = list(attrs or C2_ATTRS)
    idx, slabs, ms, nbytes = {}, [], 0.0, 0
    sw = stride_words(sch.nrows)
    by_fk = {}
    for a in attrs:
        by_fk.setdefault(C2_ATTRS[a][1], []).append(a)
    for fk, group in by_fk.items():
        keep = []
        for a in group:
            card = C2_ATTRS[a][2]
            slab = dev.alloc(sw * 4 * card)
            slabs.append(slab)
            idx[a] = [slab.ptr + j * sw * 4 for j in range(card)]
            lo = (C.c_uint64 * card)(*range(card))
            hi = (C.c_uint64 * card)(*range(1, card + 1))
            keep.append((a, card, slab, lo, hi))
        if timer:
            dev.sync()
            timer.start()
        if len(group) == 1:
            a, card, slab, lo, hi = keep[0]
            ref = colref(sch, Src(a, fk))
            check(dev.lib.mbx_create_bin_into(dev.ctx, C.byref(ref), sch.nrows,
                                              lo, hi, card, slab.ptr, sw))
        else:
            sets = (abi.BinSet * len(group))()
            for i, (a, card, slab, lo, hi) in enumerate(keep):
                col = sch[a]
                sets[i].data = col.dev.ptr
                sets[i].elem_bytes = col.elem_bytes
                sets[i].nbins = card
                sets[i].mins = C.cast(lo, C.POINTER(C.c_uint64))
                sets[i].maxs = C.cast(hi, C.POINTER(C.c_uint64))
                sets[i].slab = slab.ptr
            check(dev.lib.mbx_create_bin_multi(dev.ctx, sch[fk].dev.ptr,
                                               sch[group[0]].n, sch.nrows, sets,
                                               len(group), sw))
        if timer:
            ms += timer.stop()
        nbytes += sch.nrows * 4 + sum(card for _, card, *_ in keep) * sch.nrows // 8
    dev.sync()
    return idx, slabs, ms, nbytes


def config2_cols(spec, idx):
    """BitmapCols of a config-2 query: every predicate covers whole values,
    so each column is the OR of the exact per-value bitmaps."""
    from .spec import BitmapCol
    cols = []
    for p in spec.preds:
        ptrs = idx[p.src.col]
        assert 0 <= p.lo < p.hi <= len(ptrs) and p.hi - p.lo <= 4
        cols.append(BitmapCol([ptrs[v] for v in range(p.lo, p.hi)]))
    return cols


def config2_plans(dev, sch, idx, specs):
    """The static heuristic plan of each query (column densities on the
    device -> sparsest column first, eager loads unless the ANDM early-out
    pays).  -> [(planned vprg, row op, plan flags)]."""
    from .spec import to_rowop, to_vprg
    out = []
    for s in specs:
        prg = to_vprg(config2_cols(s, idx), sch.nrows, s.instrs)
        dens = (C.c_double * 4)()
        check(dev.lib.mbx_col_density(dev.ctx, C.byref(prg), dens))
        plan = abi.Plan()
        check(dev.lib.mbx_plan_heuristic(C.byref(prg), dens, C.byref(plan)))
        planned = abi.Vprg()
        check(dev.lib.mbx_plan_apply(C.byref(prg), C.byref(plan), C.byref(planned)))
        out.append((planned, to_rowop(s, sch), int(plan.flags)))
    return out


def host_schema(sch, rows=None):
    """A host copy of a device schema (for the oracle): every column, the
    fact columns (those with sch.nrows entries) cut to `rows`."""
    h = Schema(sch.nrows if rows is None else rows)
    for k, c in sch.cols.items():
        if c.host is not None:
            h.add(k, c.host if c.n != sch.nrows or rows is None else c.host[:rows])
        elif c.n == sch.nrows and rows is not None:
            h.add(k, c.dev.host(rows))
        else:
            h.add(k, c.dev.host())
    return h


# ---------------------------------------------------------------- config 3
def zipf_pmf(n=C1_DOMAIN, s=1.0):
    w = 1.0 / np.arange(1, n + 1, dtype=np.float64) ** s
    return w / w.sum()


def _range_density(pmf, lo, hi):
    return float(pmf[lo:hi].sum())


def config3_queries(nq=256, seed=SEED + 3, allow_edges=True):
    """nq random queries over the config-1 tables: AND, OR and ANDNOT
    programs over 2-4 range predicates on p0..p2 (bin-aligned, or with
    partly covered edge bins that need a candidate check), 1..10,000 groups
    from bucketed keys, SUM(m) or COUNT.  Queries are drawn so that their
    estimated selectivity (independent columns, exact Zipf pmf) is spread
    log-uniformly over [1e-6, 1e-1]."""
    rng = np.random.default_rng(seed)
    pmf = zipf_pmf()
    bounds = C1_BOUNDS
    nb = len(bounds) - 1
    key_choices = [
        (lambda c: Key(Src(c), abi.KEY_DIV, 25, 0), 40),
        (lambda c: Key(Src(c), abi.KEY_DIV, 100, 0), 10),
        (lambda c: Key(Src(c), abi.KEY_BUCKET, 1, 0, 1, (5, 25, 100, 300)), 5),
        (lambda c: Key(Src("m"), abi.KEY_DIV, 10000, 0), 11),
    ]
    pool = []
    while len(pool) < nq * 40:
        npred = int(rng.integers(2, 5))
        preds, dens = [], []
        for _ in range(npred):
            col = C1_COLS[rng.integers(3)]
            j = int(rng.integers(nb))
            w = int(rng.integers(1, 5))
            lo, hi = bounds[j], bounds[min(nb, j + w)]
            if allow_edges and rng.random() < 0.3:   # misaligned range
                lo = min(hi - 1, lo + int(rng.integers(1, 20)))
            preds.append(Pred(Src(col), lo, hi))
            dens.append(_range_density(pmf, lo, hi))
        shape = rng.integers(3)
        if shape == 0:        # AND chain
            ins = [(abi.ORM, 0, 0)] + [(abi.ANDM, 0, i) for i in range(1, npred)]
            sel = float(np.prod(dens))
        elif shape == 1:      # (c0 OR c1) AND rest
            ins = [(abi.ORM, 0, 0), (abi.ORM, 0, 1)] + \
                  [(abi.ANDM, 0, i) for i in range(2, npred)]
            sel = (dens[0] + dens[1] - dens[0] * dens[1]) * float(np.prod(dens[2:]))
        else:                 # c0 AND NOT c1 AND rest
            ins = [(abi.ORM, 0, 0), (abi.ORM, 1, 1), (abi.NOT, 1, 1),
                   (abi.AND, 0, 1)] + [(abi.ANDM, 0, i) for i in range(2, npred)]
            sel = dens[0] * (1 - dens[1]) * float(np.prod(dens[2:]))
        if not (1e-6 <= sel <= 1e-1):
            continue
        keys, total, stride = [], 1, 1
        for _ in range(int(rng.integers(0, 4))):
            mk, card = key_choices[rng.integers(len(key_choices))]
            if total * card > 10_000:
                break
            k = mk(C1_COLS[rng.integers(3)])
            keys.append(Key(k.src, k.kind, k.div, k.base, stride, k.bounds))
            stride *= card
            total *= card
        meas = abi.MEAS_COL if rng.random() < 0.8 else abi.MEAS_COUNT
        pool.append((np.log10(sel), QuerySpec(
            preds, ins, keys, total, meas, Src("m") if meas else None, None,
            f"b{len(pool)}")))
    # stratify: equal counts per log-selectivity band
    pool.sort(key=lambda t: t[0])
    edges = np.linspace(-6, -1, nq + 1)
    out, i = [], 0
    for b in range(nq):
        while i < len(pool) and pool[i][0] < edges[b]:
            i += 1
        if i < len(pool):
            out.append(pool[i][1])
            i += 1
    while len(out) < nq:
        out.append(pool[int(rng.integers(len(pool)))][1])
    return out


def bin_cols(spec, bin_ptrs, bounds=C1_BOUNDS):
    """BitmapCols of a spec over per-column bin slabs {col: [ptrs]} through
    col::from_bin (edges when a range is misaligned).  bounds: one bound
    list for every column, or {col: bounds}."""
    from .spec import BitmapCol, from_bin
    cols = []
    for p in spec.preds:
        b = bounds[p.src.col] if isinstance(bounds, dict) else bounds
        r = from_bin(list(b), p.lo, p.hi)
        if r is None:
            raise ValueError(f"range [{p.lo},{p.hi}) not coverable by bins")
        mid, l, rr = r
        ptrs = bin_ptrs[p.src.col]
        cols.append(BitmapCol([ptrs[i] for i in mid],
                              None if l is None else ptrs[l],
                              None if rr is None else ptrs[rr]))
    return cols


# ---------------------------------------------------------------- config 5
C5_ROWS = 600_000_000
C5_COLS = tuple(f"z{c}" for c in range(10))
C5_SKEW = 1.2
C5_DOMAINS = tuple(int(round(1000 * 100 ** (c / 9))) for c in range(10))
C5_NBINS = 64


def c5_bounds(c):
    """64 equal-width buckets covering [0, domain) of column c."""
    w = -(-C5_DOMAINS[c] // C5_NBINS)
    return tuple(range(0, w * C5_NBINS + 1, w))


C5_BOUNDS = {name: c5_bounds(c) for c, name in enumerate(C5_COLS)}


def config5_device(dev, n=C5_ROWS, seed=SEED + 5):
    """The 10 columns, generated on the device (no host copy)."""
    sch = Schema(n)
    for c, name in enumerate(C5_COLS):
        a = dev.alloc(4 * n, np.uint32, n)
        check(dev.lib.mbx_gen_zipf_u32(dev.ctx, a.ptr, n, col_seed(seed, 50 + c),
                                       C5_DOMAINS[c], C5_SKEW))
        sch.add(name, None, a)
    dev.sync()
    return sch


def config5_queries(nq, seed=SEED + 55):
    """nq requests from the config-3 program space over z0..z9: 2-4 range
    predicates on distinct columns (1-4 buckets, 30% misaligned), AND / OR /
    ANDNOT, estimated selectivity log-uniform over [1e-6, 1e-1], 1..10,000
    groups from bucketed keys, SUM(z0) or COUNT."""
    rng = np.random.default_rng(seed)
    pmfs = [zipf_pmf(d, C5_SKEW) for d in C5_DOMAINS]
    pool = []
    while len(pool) < nq * 20:
        npred = int(rng.integers(2, 5))
        cols = rng.choice(len(C5_COLS), size=npred, replace=False)
        preds, dens = [], []
        for c in cols:
            b = C5_BOUNDS[C5_COLS[c]]
            nb = len(b) - 1
            j = int(rng.integers(nb))
            w = int(rng.integers(1, 5))
            lo, hi = b[j], b[min(nb, j + w)]
            if rng.random() < 0.3 and hi - lo > 1:
                lo = min(hi - 1, lo + int(rng.integers(1, b[1])))
            preds.append(Pred(Src(C5_COLS[c]), lo, hi))
            dens.append(_range_density(pmfs[c], lo, min(hi, C5_DOMAINS[c])))
        shape = rng.integers(3)
        if shape == 0:
            ins = [(abi.ORM, 0, 0)] + [(abi.ANDM, 0, i) for i in range(1, npred)]
            sel = float(np.prod(dens))
        elif shape == 1:
            ins = [(abi.ORM, 0, 0), (abi.ORM, 0, 1)] + \
                  [(abi.ANDM, 0, i) for i in range(2, npred)]
            sel = (dens[0] + dens[1] - dens[0] * dens[1]) * float(np.prod(dens[2:]))
        else:
            ins = [(abi.ORM, 0, 0), (abi.ORM, 1, 1), (abi.NOT, 1, 1),
                   (abi.AND, 0, 1)] + [(abi.ANDM, 0, i) for i in range(2, npred)]
            sel = dens[0] * (1 - dens[1]) * float(np.prod(dens[2:]))
        if not (1e-6 <= sel <= 1e-1):
            continue
        keys, total = [], 1
        for _ in range(int(rng.integers(0, 4))):
            c = int(rng.integers(len(C5_COLS)))
            wdt = C5_BOUNDS[C5_COLS[c]][1]
            div = wdt * int(rng.choice([1, 4, 16]))
            card = -(-C5_DOMAINS[c] // div)
            if total * card > 10_000:
                break
            keys.append(Key(Src(C5_COLS[c]), abi.KEY_DIV, div, 0, total))
            total *= card
        meas = abi.MEAS_COL if rng.random() < 0.8 else abi.MEAS_COUNT
        pool.append((np.log10(sel), QuerySpec(
            preds, ins, keys, total, meas, Src("z0") if meas else None, None,
            f"r{len(pool)}")))
    pool.sort(key=lambda t: t[0])
    edges = np.linspace(-6, -1, nq + 1)
    out, i = [], 0
    for b in range(nq):
        while i < len(pool) and pool[i][0] < edges[b]:
            i += 1
        if i < len(pool):
            out.append(pool[i][1])
            i += 1
    while len(out) < nq:
        out.append(pool[int(rng.integers(len(pool)))][1])
    order = rng.permutation(len(out))       # arrival order: mixed selectivities
    return [out[k] for k in order]


def config5_training_queries(nq, seed=SEED + 56, heavy_share=0.2):
    """Requests for training a plan chooser over the config-5 tables: the
    config-5 program space, and a share of heavy ones -- a range starting
    inside the first bucket of a column (its edge bin holds the Zipf head,
    so most rows need a re-check) AND 1-3 wide ranges -- which the static
    heuristic runs fused although the scan is cheaper."""
    rng = np.random.default_rng(seed + 1)
    nheavy = int(round(nq * heavy_share))
    out = config5_queries(nq - nheavy, seed)
    for h in range(nheavy):
        npred = int(rng.integers(2, 5))
        cols = rng.choice(len(C5_COLS), size=npred, replace=False)
        preds = []
        for i, c in enumerate(cols):
            b = C5_BOUNDS[C5_COLS[c]]
            if i == 0:
                preds.append(Pred(Src(C5_COLS[c]), int(rng.integers(1, b[1])),
                                  b[int(rng.integers(1, 4))]))
            else:
                preds.append(Pred(Src(C5_COLS[c]), 0, b[int(rng.integers(2, 5))]))
        ins = [(abi.ORM, 0, 0)] + [(abi.ANDM, 0, i) for i in range(1, npred)]
        c = int(rng.integers(len(C5_COLS)))
        div = C5_BOUNDS[C5_COLS[c]][1] * 16
        card = -(-C5_DOMAINS[c] // div)
        out.append(QuerySpec(preds, ins, [Key(Src(C5_COLS[c]), abi.KEY_DIV, div, 0, 1)],
                             card, abi.MEAS_COL, Src("z0"), None, f"h{h}"))
    order = rng.permutation(len(out))
    return [out[k] for k in order]


# ------------------------------------------------------------ shared helpers
def build_bins(dev, sch, bounds_of):
    """One create_bin_into slab per column: {col: [bin ptrs]}, [slabs]."""
    from .host import stride_words
    sw = stride_words(sch.nrows)
    out, slabs = {}, []
    for col, b in bounds_of.items():
        nb = len(b) - 1
        slab = dev.alloc(sw * 4 * nb)
        ref = abi.ColRef(sch[col].dev.ptr, None, 0, 4)
        check(dev.lib.mbx_create_bin_into(dev.ctx, C.byref(ref), sch.nrows,
                                          (C.c_uint64 * nb)(*b[:-1]),
                                          (C.c_uint64 * nb)(*b[1:]), nb,
                                          slab.ptr, sw))
        out[col] = [slab.ptr + j * sw * 4 for j in range(nb)]
        slabs.append(slab)
    dev.sync()
    return out, slabs


def query_array(sch, specs, bins, bounds):
    """mbx_query array of specs over bin slabs (programs as written)."""
    from .spec import to_rowop, to_vprg
    qs = (abi.Query * len(specs))()
    for i, s in enumerate(specs):
        qs[i].prg = to_vprg(bin_cols(s, bins, bounds), sch.nrows, s.instrs)
        qs[i].op = to_rowop(s, sch)
        qs[i].nr_grp = s.nr_grp
    return qs


def heuristic_queries(dev, qs):
    """A copy of a query array with the static heuristic plan applied to
    every query (sparsest column first; eager loads unless the early-out
    pays), as the batched and serving paths run them."""
    out = (abi.Query * len(qs))()
    C.memmove(out, qs, C.sizeof(out))
    for i in range(len(qs)):
        dens = (C.c_double * 4)()
        check(dev.lib.mbx_col_density(dev.ctx, C.byref(qs[i].prg), dens))
        plan = abi.Plan()
        check(dev.lib.mbx_plan_heuristic(C.byref(qs[i].prg), dens, C.byref(plan)))
        check(dev.lib.mbx_plan_apply(C.byref(qs[i].prg), C.byref(plan),
                                     C.byref(out[i].prg)))
        out[i].flags = plan.flags & abi.PLAN_EAGER
    return out


def spec_columns(spec):
    """Names of every column a query reads (predicates, keys, measures,
    their foreign keys)."""
    names = set()
    for src in [p.src for p in spec.preds] + [k.src for k in spec.keys] + \
            [spec.ma, spec.mb]:
        if src is not None:
            names.add(src.col)
            if src.fk:
                names.add(src.fk)
    return names


def host_subset(sch, names):
    """Host copies of the named columns of a device schema."""
    h = Schema(sch.nrows)
    for k in names:
        c = sch[k]
        h.add(k, c.host if c.host is not None else c.dev.host())
    return h
