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
