"""Seeded test data: small schemas, random programs and row operations."""
import numpy as np

from merle_amd import abi
from merle_amd.spec import Key, Pred, QuerySpec, Schema, Src

# (column name, dtype, value range) of the small synthetic fact table
FACT_COLS = [("a8", np.uint8, 256), ("a16", np.uint16, 4000),
             ("a32", np.uint32, 1000), ("a64", np.uint64, 1 << 40),
             ("b32", np.uint32, 50)]


def small_schema(n, seed, dim_n=97, fk_overflow=3):
    """Fact table of n rows + one dimension of dim_n rows.  fk values reach
    dim_n + fk_overflow - 1, so some rows point past the dimension."""
    rng = np.random.default_rng(seed)
    s = Schema(n)
    for name, dt, hi in FACT_COLS:
        if name == "a32":   # skewed, so bins have uneven densities
            v = np.minimum(rng.zipf(1.3, n) - 1, hi - 1)
        else:
            v = rng.integers(0, hi, n, dtype=np.uint64)
        s.add(name, np.ascontiguousarray(v.astype(dt)))
    s.add("fk", rng.integers(0, dim_n + fk_overflow, n).astype(np.uint32))
    s.add("fk_ok", rng.integers(0, dim_n, n).astype(np.uint32))
    s.add("d32", rng.integers(0, 60, dim_n).astype(np.uint32))
    s.add("d16", rng.integers(0, 9, dim_n).astype(np.uint16))
    s.add("m32", rng.integers(1, 100001, n).astype(np.uint32))
    s.add("c32", rng.integers(0, 131072, n).astype(np.uint32))
    return s


SOURCES = [("a8", None, 256), ("a16", None, 4000), ("a32", None, 1000),
           ("b32", None, 50), ("d32", "fk", 60), ("d16", "fk_ok", 9)]


def random_pred(rng, schema):
    col, fk, hi = SOURCES[rng.integers(len(SOURCES))]
    lo = int(rng.integers(0, hi))
    width = int(rng.integers(1, max(2, hi // 2)))
    return Pred(Src(col, fk), lo, min(hi + 1, lo + width))


def random_program(rng, npreds, allow_not=True):
    """Random instruction list over registers r0..r2 and columns
    0..npreds-1, ending with the result in r0 (<= 6 instructions)."""
    ops = [abi.ORM, abi.ANDM, abi.AND, abi.OR]
    if allow_not:
        ops.append(abi.NOT)
    n = int(rng.integers(1, 7))
    ins = [(abi.ORM, 0, int(rng.integers(npreds)))]
    for _ in range(n - 1):
        op = ops[rng.integers(len(ops))]
        d = int(rng.integers(3))
        s = int(rng.integers(npreds)) if op in (abi.ANDM, abi.ORM) else int(rng.integers(3))
        ins.append((op, d, s))
    return ins[:6]


def random_keys(rng):
    keys, stride, total = [], 1, 1
    choices = [
        (Key(Src("b32"), abi.KEY_IDENT, 1, 0), 50),
        (Key(Src("a8"), abi.KEY_DIV, 16, 0), 16),
        (Key(Src("d16", "fk_ok"), abi.KEY_IDENT, 1, 0), 9),
        (Key(Src("a32"), abi.KEY_BUCKET, 1, 0, 1, (10, 100, 300)), 4),
        (Key(Src("d32", "fk"), abi.KEY_DIV, 10, 0), 6),
    ]
    for _ in range(int(rng.integers(0, 4))):
        k, card = choices[rng.integers(len(choices))]
        k = Key(k.src, k.kind, k.div, k.base, stride, k.bounds)
        keys.append(k)
        stride *= card
        total *= card
    return keys, total


def random_measure(rng):
    m = int(rng.integers(4))
    if m == abi.MEAS_COUNT:
        return m, None, None
    if m == abi.MEAS_COL:
        return m, Src("m32"), None
    if m == abi.MEAS_MUL:
        return m, Src("m32"), Src("a8")
    return m, Src("m32"), Src("c32")


def random_spec(rng, schema, npreds=None, allow_not=True, extra_groups=0):
    npreds = npreds or int(rng.integers(1, 5))
    preds = [random_pred(rng, schema) for _ in range(npreds)]
    ins = random_program(rng, npreds, allow_not)
    keys, total = random_keys(rng)
    m, ma, mb = random_measure(rng)
    return QuerySpec(preds, ins, keys, total + extra_groups, m, ma, mb)
