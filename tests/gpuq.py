"""Helpers for GPU tests: device copies of schemas and bitmap plans."""
import ctypes as C

import numpy as np

from merle_amd import abi, spec as S
from merle_amd.host import check, stride_words


def upload_schema(dev, schema):
    for name, col in schema.cols.items():
        col.dev = dev.upload(col.host)
    return schema


def create_bin(dev, schema, src, bounds_lo, bounds_hi):
    """mbx_create_bin over a spec Src -> (list of bitmap pointers, slab)."""
    ref = S.colref(schema, src)
    n = schema.nrows
    nb = len(bounds_lo)
    outs = (C.c_void_p * nb)()
    single = C.c_size_t()
    check(dev.lib.mbx_create_bin(dev.ctx, C.byref(ref), n,
                                 (C.c_uint64 * nb)(*map(int, bounds_lo)),
                                 (C.c_uint64 * nb)(*map(int, bounds_hi)), nb,
                                 outs, C.byref(single)))
    assert single.value == (n + 31) // 32 * 4
    return [outs[j] for j in range(nb)]


def bitmaps_host(dev, ptrs, n):
    sw = stride_words(n)
    out = np.zeros((len(ptrs), sw), np.uint32)
    for j, p in enumerate(ptrs):
        check(dev.lib.mbx_memcpy_d2h(dev.ctx, out[j].ctypes.data, p, sw * 4))
    return out


def exact_bitmap(dev, schema, pred):
    """create_join over the predicate's exact range."""
    c = schema[pred.src.col]
    out = C.c_void_p()
    if pred.src.fk:
        check(dev.lib.mbx_create_join(dev.ctx, schema[pred.src.fk].dev.ptr,
                                      c.dev.ptr, c.elem_bytes, schema.nrows,
                                      c.n, pred.lo, pred.hi, C.byref(out)))
    else:
        check(dev.lib.mbx_create_select(dev.ctx, c.dev.ptr, c.elem_bytes,
                                        schema.nrows, pred.lo, pred.hi,
                                        C.byref(out)))
    return out.value


def plan_cols(dev, schema, spec, rng, binned_frac=0.6, nbins=12):
    """One BitmapCol per predicate: exact bitmaps, or equal-width bins mapped
    through col::from_bin (edge bins -> candidate check)."""
    cols, allocs = [], []
    for p in spec.preds:
        col = None
        if rng.random() < binned_frac:
            hi_v = int(max(p.hi, 2)) * 2
            width = max(1, hi_v // nbins)
            bounds = list(range(0, width * (nbins + 1), width))
            ptrs = create_bin(dev, schema, p.src, bounds[:-1], bounds[1:])
            allocs.append(ptrs[0])
            fb = S.from_bin(bounds, p.lo, p.hi)
            if fb is not None:
                mid, l, r = fb
                col = S.BitmapCol([ptrs[i] for i in mid],
                                  None if l is None else ptrs[l],
                                  None if r is None else ptrs[r])
        if col is None:
            b = exact_bitmap(dev, schema, p)
            allocs.append(b)
            col = S.BitmapCol([b])
        cols.append(col)
    return cols, allocs


def run_query(dev, schema, spec, cols, plan=abi.PLAN_FUSED):
    prg = S.to_vprg(cols, schema.nrows, spec.instrs)
    op = S.to_rowop(spec, schema)
    out = dev.alloc(8 * spec.nr_grp, np.int64, spec.nr_grp)
    check(dev.lib.mbx_prog_grpby(dev.ctx, C.byref(prg), C.byref(op),
                                 spec.nr_grp, out.ptr, plan))
    dev.sync()
    res = out.host()
    out.free()
    return res


def free_all(dev, ptrs):
    for p in ptrs:
        check(dev.lib.mbx_free(dev.ctx, p))
