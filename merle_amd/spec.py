"""Logical query descriptions and their translation to the C ABI structs.

A query is a bitmap program over up to four predicate columns (vprg,
primitive.cuh:164-198) plus the row operation that the reference writes as a
device functor (ssb/lambdas.cuh:19-152): exact predicates for re-checks, a
mixed-radix group key and a measure.  QuerySpec holds that description with
column names; ``to_rowop``/``to_vprg`` bind it to device pointers.
"""
from dataclasses import dataclass, field
from typing import Dict, List, Optional, Sequence, Tuple

import numpy as np

from . import abi


@dataclass
class Column:
    host: Optional[np.ndarray]          # host copy (may be None on device-only data)
    dev: object = None                   # DevArray or None
    n: int = 0

    @property
    def elem_bytes(self):
        if self.host is not None:
            return self.host.dtype.itemsize
        return self.dev.dtype.itemsize


class Schema:
    """Named columns: fact columns and dimension attributes alike."""

    def __init__(self, nrows):
        self.nrows = nrows
        self.cols: Dict[str, Column] = {}

    def add(self, name, host=None, dev=None):
        n = host.size if host is not None else dev.n
        self.cols[name] = Column(host, dev, n)
        return self.cols[name]

    def __getitem__(self, k):
        return self.cols[k]


@dataclass
class Src:
    col: str                      # fact column, or dimension attribute if fk
    fk: Optional[str] = None      # fact foreign-key column


@dataclass
class Pred:
    src: Src
    lo: int
    hi: int


@dataclass
class Key:
    src: Src
    kind: int = abi.KEY_IDENT
    div: int = 1
    base: int = 0
    stride: int = 1
    bounds: Tuple[int, ...] = ()


@dataclass
class QuerySpec:
    preds: List[Pred]                       # predicate i <-> program column i
    instrs: List[Tuple[int, int, int]]      # (opcode, dstreg, src1)
    keys: List[Key]
    nr_grp: int
    measure: int = abi.MEAS_COUNT
    ma: Optional[Src] = None
    mb: Optional[Src] = None
    name: str = ""


def _dev_ptr(col):
    if col.dev is None:
        raise ValueError("column has no device copy")
    return col.dev.ptr


def colref(schema: Schema, src: Src) -> abi.ColRef:
    c = schema[src.col]
    r = abi.ColRef()
    r.data = _dev_ptr(c)
    r.elem_bytes = c.elem_bytes
    if src.fk:
        r.fk = _dev_ptr(schema[src.fk])
        r.dimsz = c.n
    else:
        r.fk = None
        r.dimsz = 0
    return r


def to_rowop(spec: QuerySpec, schema: Schema) -> abi.RowOp:
    op = abi.RowOp()
    op.npreds = len(spec.preds)
    for i, p in enumerate(spec.preds):
        op.preds[i].src = colref(schema, p.src)
        op.preds[i].lo = p.lo
        op.preds[i].hi = p.hi
    op.nkeys = len(spec.keys)
    for i, k in enumerate(spec.keys):
        kp = op.keys[i]
        kp.src = colref(schema, k.src)
        kp.kind = k.kind
        kp.div = k.div
        kp.base = k.base
        kp.stride = k.stride
        kp.nbounds = len(k.bounds)
        for j, b in enumerate(k.bounds):
            kp.bounds[j] = b
    op.measure = spec.measure
    if spec.ma is not None:
        op.ma = colref(schema, spec.ma)
    if spec.mb is not None:
        op.mb = colref(schema, spec.mb)
    return op


@dataclass
class BitmapCol:
    """Bitmap handles of one program column (mbx_col): device pointers."""
    middle: List[int] = field(default_factory=list)
    leftmost: Optional[int] = None
    rightmost: Optional[int] = None


def to_vprg(cols: Sequence[BitmapCol], factsz: int,
            instrs: Sequence[Tuple[int, int, int]]) -> abi.Vprg:
    p = abi.Vprg()
    for i, c in enumerate(cols):
        for k, m in enumerate(c.middle):
            p.cols[i].middle[k] = m
        p.cols[i].leftmost = c.leftmost
        p.cols[i].rightmost = c.rightmost
    p.factsz = factsz
    for i, ins in enumerate(instrs):
        p.instrs[i].opcode, p.instrs[i].dstreg, p.instrs[i].src1 = ins[:3]
    return p


def from_bin(bounds: Sequence[int], lo: int, hi: int):
    """col::from_bin over bin indices (host-only C helper).  Returns
    (middle idx list, leftmost idx|None, rightmost idx|None) or None."""
    import ctypes as C
    from . import load
    lib = load()
    n = len(bounds) - 1
    ids = (C.c_uint64 * n)(*range(n))
    b = (C.c_uint64 * (n + 1))(*bounds)
    slots = (C.c_uint64 * 6)()
    st = lib.mbx_col_from_bin(n, ids, b, lo, hi, slots)
    if st != abi.OK:
        return None
    none = 2 ** 64 - 1
    mid = [s for s in list(slots)[:4] if s != none]
    left = None if slots[4] == none else slots[4]
    right = None if slots[5] == none else slots[5]
    if not mid and left is None and right is None:
        return None
    return mid, left, right


def from_bin2(coarse: Sequence[int], fine: Sequence[int], lo: int, hi: int):
    """col::from_bin2 (bmpidx_create.cu:165-174): the coarse bins unless they
    need a candidate check, then the fine ones.  -> (0 coarse | 1 fine,
    (middle, left, right)) or None."""
    r = from_bin(coarse, lo, hi)
    if r is not None and r[1] is None and r[2] is None:
        return 0, r
    if r is None:
        return None
    f = from_bin(fine, lo, hi)
    return None if f is None else (1, f)
