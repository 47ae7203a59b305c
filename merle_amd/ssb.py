"""The SSB flights of the reference as row-op descriptors.

Arguments are those of ssb_cpujoin (ssb/cpuref.c:244-302); group formulas
follow s1ref..s4ref (ssb/cpuref.c:141-242) and the device ops s1op..s4op
(ssb/lambdas.cuh:19-152), including their quirks (Q4's key uses the supplier
city range as the stride of the city term).  Bin tables are those of
ssb/bmpidx-create.cu:8-30.
"""
from dataclasses import dataclass
from typing import Tuple

from . import abi
from .spec import Key, Pred, QuerySpec, Src

# ssb/ssbdemo.h:10-26 -- day offsets from 1992-01-01
D920101, D930101, D940101, D940201, D940204, D940211 = 0, 366, 731, 762, 765, 772
D940301, D950101, D960101, D970101, D971201 = 790, 1096, 1461, 1827, 2161
D980101, D990101 = 2192, 2557
# ssbDateToYear (ssb/ssbdemo.h:29-39) as a bucketization: year = #{b <= date}
YEAR_BOUNDS = (D930101, D940101, D950101, D960101, D970101, D980101)


def date_to_year(d):
    return sum(1 for b in YEAR_BOUNDS if b <= d)


# ssb/bmpidx-create.cu:8-30
DATE_SP = (D920101, D930101, D940101, D950101, D960101, D970101, D980101, D990101)
DATE_DE = (D940101, D940201, D940301, D971201, D980101)
DISCNT_SP = (0, 2, 5, 8, 11)
DISCNT_DE = tuple(range(12))
QTY_SP = (1, 11, 21, 31, 41, 51)
QTY_DE = tuple(range(1, 52, 5))
MFGR_SP = (0, 200, 400, 600, 800, 1000)
MFGR_DE = (40, 80, 120, 160, 240, 280)
CITY_SP = (0, 50, 100, 150, 200, 250)
CITY_DE = (50, 60, 190, 200)


@dataclass
class Flight:
    name: str
    kind: int                 # 1..4 -> s1ref..s4ref
    args: Tuple[int, ...]
    ngrp: int                 # NGRP_S* (ssb/ssbdemo.h:79-87)


FLIGHTS = [
    Flight("SSB11", 1, (D930101, D940101, 1, 4, 0, 25), 1),
    Flight("SSB12", 1, (D940101, D940201, 4, 7, 26, 36), 1),
    Flight("SSB13", 1, (D940204, D940211, 5, 8, 26, 36), 1),
    Flight("SSB21", 2, (40, 80, 150, 200), 280),
    Flight("SSB22", 2, (260, 268, 200, 250), 56),
    Flight("SSB23", 2, (260, 261, 50, 100), 7),
    Flight("SSB31", 3, (200, 250, 200, 250, D920101, D980101), 150),
    Flight("SSB32", 3, (190, 200, 190, 200, D920101, D980101), 600),
    Flight("SSB33", 3, (51, 55, 51, 55, D920101, D980101), 96),
    Flight("SSB34", 3, (51, 55, 51, 55, D971201, D980101), 16),
    Flight("SSB41", 4, (150, 200, 150, 200, 0, 400, D920101, D990101), 350),
    Flight("SSB42", 4, (150, 200, 150, 200, 0, 400, D970101, D990101), 100),
    Flight("SSB43", 4, (150, 200, 190, 200, 120, 160, D970101, D990101), 800),
]

CONJ = [(abi.ORM, 0, 0), (abi.ANDM, 0, 1), (abi.ANDM, 0, 2), (abi.ANDM, 0, 3)]


def _scaled(lo, hi, span, div):
    """(transform kind, div, base, width) of a scaled key term."""
    if hi - lo >= span:
        return abi.KEY_DIV, div, lo // div, hi // div - lo // div
    return abi.KEY_IDENT, 1, lo, hi - lo


def flight_spec(f: Flight) -> QuerySpec:
    a = f.args
    date, disc, qty = Src("loOrderDate"), Src("loDiscount"), Src("loQuantity")
    scity, ccity = Src("loSuppCity"), Src("custCity", "loCustKey")
    mfgr = Src("partMfgr", "loPartKey")
    if f.kind == 1:   # s1ref: SUM(extprice * discount)
        preds = [Pred(date, a[0], a[1]), Pred(disc, a[2], a[3]), Pred(qty, a[4], a[5])]
        return QuerySpec(preds, CONJ[:3], [], 1, abi.MEAS_MUL,
                         Src("loExtendedPrice"), disc, f.name)
    if f.kind == 2:   # s2ref: (mfgr - min) + year * (max - min)
        pmin, pmax, smin, smax = a
        preds = [Pred(mfgr, pmin, pmax), Pred(scity, smin, smax)]
        keys = [Key(mfgr, abi.KEY_IDENT, 1, pmin, 1),
                Key(date, abi.KEY_BUCKET, 1, 0, pmax - pmin, YEAR_BOUNDS)]
        return QuerySpec(preds, CONJ[:2], keys, f.ngrp, abi.MEAS_COL,
                         Src("loRevenue"), None, f.name)
    if f.kind == 3:   # s3ref
        cmin, cmax, smin, smax, dmin, dmax = a
        if cmax - cmin >= 50:   # both cities scale by 10 when the c-range does
            ck, cd, cb, ca = abi.KEY_DIV, 10, cmin // 10, cmax // 10 - cmin // 10
            sk, sd, sb, sbw = abi.KEY_DIV, 10, smin // 10, smax // 10 - smin // 10
        else:
            ck, cd, cb, ca = abi.KEY_IDENT, 1, cmin, cmax - cmin
            sk, sd, sb, sbw = abi.KEY_IDENT, 1, smin, smax - smin
        preds = [Pred(ccity, cmin, cmax), Pred(scity, smin, smax), Pred(date, dmin, dmax)]
        keys = [Key(date, abi.KEY_BUCKET, 1, date_to_year(dmin), ca * sbw, YEAR_BOUNDS),
                Key(ccity, ck, cd, cb, ca),
                Key(scity, sk, sd, sb, 1)]
        return QuerySpec(preds, CONJ[:3], keys, f.ngrp, abi.MEAS_COL,
                         Src("loRevenue"), None, f.name)
    # s4ref: SUM(revenue - supplycost)
    cmin, cmax, smin, smax, pmin, pmax, dmin, dmax = a
    sk, sd, sb, sa = _scaled(smin, smax, 50, 10)
    pk, pd, pb, pw = _scaled(pmin, pmax, 200, 40)
    preds = [Pred(ccity, cmin, cmax), Pred(scity, smin, smax),
             Pred(mfgr, pmin, pmax), Pred(date, dmin, dmax)]
    keys = [Key(date, abi.KEY_BUCKET, 1, date_to_year(dmin), sa * pw, YEAR_BOUNDS),
            Key(scity, sk, sd, sb, sa),
            Key(mfgr, pk, pd, pb, 1)]
    return QuerySpec(preds, CONJ[:4], keys, f.ngrp, abi.MEAS_SUB,
                     Src("loRevenue"), Src("loSupplyCost"), f.name)


# ---------------------------------------------------------------------------
# The fixed-layout plans of ssb/fixedlayout.cu (s1fix..s4fix): per flight,
# three programs over the prebuilt bin sets.  SPARSE(c) = col::from_bin on the
# sparse bins; DENSE(c) = col::from_bin2 with the sparse bins as the coarse
# level and the dense bins as the fine one (fixedlayout.cu:6-12).

SPARSE, DENSE = 0, 1
BIN_SETS = {
    "loOrderDate": (DATE_SP, DATE_DE),
    "loDiscount": (DISCNT_SP, DISCNT_DE),
    "loQuantity": (QTY_SP, QTY_DE),
    "loSuppCity": (CITY_SP, CITY_DE),
    "custCity": (CITY_SP, CITY_DE),       # cCitySpBin = sCitySpBin (:4)
    "partMfgr": (MFGR_SP, MFGR_DE),
}
# fixedlayout.cu:147 passes quantity [1, 25) for SSB11 where ssb_cpujoin
# passes [0, 25) (cpuref.c:252); both select the same rows since quantity is
# at least 1 (ssb-dbgen-lite.c:139).
_FIXED_ARGS = {"SSB11": (D930101, D940101, 1, 4, 1, 25)}


@dataclass
class FixedPlan:
    label: str                               # Sparse / Medium / Dense
    cols: Tuple[Tuple[str, int, int, int], ...]   # (source, lo, hi, SPARSE|DENSE)
    instrs: Tuple[Tuple[int, int, int], ...]


def fixed_plans(f: Flight):
    """The Sparse, Medium and Dense programs of s1fix..s4fix
    (ssb/fixedlayout.cu:35-140).  The row op stays the flight's own
    (flight_spec): for flight 3 the op keeps the original date range while
    the program loads the complement (:93-104)."""
    a = _FIXED_ARGS.get(f.name, f.args)
    LOAD, ANDM = abi.ORM, abi.ANDM
    if f.kind == 1:
        cols = [("loOrderDate", a[0], a[1]), ("loDiscount", a[2], a[3]),
                ("loQuantity", a[4], a[5])]
        ins = [(LOAD, 0, 0), (ANDM, 0, 1), (ANDM, 0, 2)]
        medium, dense = {1, 2}, {0, 1, 2}
    elif f.kind == 2:
        cols = [("partMfgr", a[0], a[1]), ("loSuppCity", a[2], a[3])]
        ins = [(LOAD, 0, 0), (ANDM, 0, 1)]
        medium, dense = {0}, {0, 1}
    elif f.kind == 3:
        dmin, dmax = a[4], a[5]
        ins = [(LOAD, 0, 0), (ANDM, 0, 1)]
        if dmin == D920101:      # date < dmax as NOT(date in [dmax, 1999))
            dmin, dmax = dmax, D990101
            ins += [(LOAD, 1, 2), (abi.NOT, 1, 1), (abi.AND, 0, 1)]
        else:
            ins += [(ANDM, 0, 2)]
        cols = [("custCity", a[0], a[1]), ("loSuppCity", a[2], a[3]),
                ("loOrderDate", dmin, dmax)]
        medium, dense = {0}, {0, 1, 2}
    else:
        cols = [("custCity", a[0], a[1]), ("loSuppCity", a[2], a[3]),
                ("partMfgr", a[4], a[5])]
        ins = [(LOAD, 0, 0), (ANDM, 0, 1), (ANDM, 0, 2)]
        if a[7] < D990101 or a[6] > D920101:   # a date column only if bounded
            cols.append(("loOrderDate", a[6], a[7]))
            ins.append((ANDM, 0, 3))
        medium, dense = {0, 2}, {0, 1, 2}
    out = []
    for label, lvl in (("Sparse", set()), ("Medium", medium), ("Dense", dense)):
        out.append(FixedPlan(label, tuple((c, lo, hi, DENSE if i in lvl else SPARSE)
                                          for i, (c, lo, hi) in enumerate(cols)),
                             tuple(ins)))
    return out


def fixed_bins(col: str, lo: int, hi: int, level: int):
    """-> (set index 0 sparse / 1 dense, (middle, left, right)) of SPARSE(c)
    or DENSE(c), or None when the range is not coverable."""
    from .spec import from_bin, from_bin2
    sp, de = BIN_SETS[col]
    if level == SPARSE:
        r = from_bin(list(sp), lo, hi)
        return None if r is None else (0, r)
    return from_bin2(list(sp), list(de), lo, hi)


_SRC = {"loOrderDate": Src("loOrderDate"), "loDiscount": Src("loDiscount"),
        "loQuantity": Src("loQuantity"), "loSuppCity": Src("loSuppCity"),
        "custCity": Src("custCity", "loCustKey"),
        "partMfgr": Src("partMfgr", "loPartKey")}


def fixed_spec(f: Flight, plan: FixedPlan) -> QuerySpec:
    """The flight under one fixed-layout program: the program's instructions
    and, as the row op's predicates, the exact predicate behind each program
    column (an uncertain row is re-checked by running the program over them,
    so flight 3's complemented date column carries its complemented range)."""
    import dataclasses
    spec = flight_spec(f)
    return dataclasses.replace(
        spec, preds=[Pred(_SRC[c], lo, hi) for c, lo, hi, _ in plan.cols],
        instrs=list(plan.instrs))
