"""Independent numpy formulations used to pin the C oracle.

These restate the *definitions* (a bit is set iff lo <= v < hi; a program is
a boolean expression over predicate masks; a group-by is a keyed sum), not
the oracle's word-at-a-time algorithms, so a mistake in either side shows up
as a mismatch.
"""
import numpy as np

from merle_amd import abi

INT64_MAX = (1 << 63) - 1


def pack(bits, stride_words=None):
    """bool[n] -> uint32 words, bit b of word k = row 32k+b (LSB first)."""
    bits = np.asarray(bits, dtype=bool)
    n = bits.size
    W = (n + 31) // 32
    padded = np.zeros(W * 32, dtype=bool)
    padded[:n] = bits
    words = np.packbits(padded, bitorder="little").view("<u4").astype(np.uint32)
    if stride_words is not None:
        out = np.zeros(stride_words, dtype=np.uint32)
        out[:W] = words
        return out
    return words


def unpack(words, n):
    b = np.unpackbits(np.ascontiguousarray(words, dtype="<u4").view(np.uint8),
                      bitorder="little")
    return b[:n].astype(bool)


def values(schema, src):
    """(value uint64[n], valid bool[n]) of a spec Src on the fact rows."""
    data = schema[src.col].host
    if src.fk is None:
        return data.astype(np.uint64), np.ones(data.size, dtype=bool)
    fk = schema[src.fk].host.astype(np.int64)
    ok = fk < data.size
    v = np.zeros(fk.size, dtype=np.uint64)
    v[ok] = data[fk[ok]].astype(np.uint64)
    return v, ok


def pred_masks(spec, schema):
    out = []
    for p in spec.preds:
        v, ok = values(schema, p.src)
        out.append(ok & (v >= np.uint64(p.lo)) & (v < np.uint64(p.hi)))
    return out


def run_program(instrs, masks, n):
    r = [np.zeros(n, dtype=bool) for _ in range(3)]
    for op, d, s in instrs:
        if op == abi.AND:
            r[d] = r[d] & r[s]
        elif op == abi.OR:
            r[d] = r[d] | r[s]
        elif op == abi.NOT:
            r[d] = ~r[s]
        elif op == abi.ANDM:
            r[d] = r[d] & masks[s]
        elif op == abi.ORM:
            r[d] = r[d] | masks[s]
    return r[0]


def group_sum(spec, schema, selected=None):
    n = schema.nrows
    if selected is None:
        selected = run_program(spec.instrs, pred_masks(spec, schema), n)
    key = np.zeros(n, dtype=np.int64)
    ok = selected.copy()
    for k in spec.keys:
        v, vok = values(schema, k.src)
        ok &= vok
        if k.kind == abi.KEY_DIV:
            f = v // np.uint64(k.div)
        elif k.kind == abi.KEY_BUCKET:
            f = np.zeros(n, dtype=np.uint64)
            for b in k.bounds:
                f += (v >= np.uint64(b)).astype(np.uint64)
        else:
            f = v
        key += np.int64(k.stride) * (f.astype(np.int64) - np.int64(k.base))
    if spec.measure == abi.MEAS_COUNT:
        val = np.ones(n, dtype=np.int64)
    else:
        a, aok = values(schema, spec.ma)
        ok &= aok
        if spec.measure == abi.MEAS_COL:
            val = a.view(np.int64)
        else:
            b, bok = values(schema, spec.mb)
            ok &= bok
            if spec.measure == abi.MEAS_MUL:
                val = (a * b).view(np.int64)
            else:
                val = a.view(np.int64) - b.view(np.int64)
    ok &= (key >= 0) & (key < spec.nr_grp)
    out = np.zeros(spec.nr_grp, dtype=np.int64)
    np.add.at(out, key[ok], val[ok])
    return out, int(ok.sum())
