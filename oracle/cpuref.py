"""ctypes binding of the CPU oracle (oracle/libcpuref.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg, as the checker.  Nothing in merle_amd/ imports
it.  Inputs are host numpy arrays; the query description is the same
merle_amd.spec.QuerySpec the product consumes.
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libcpuref.so")

END, AND, OR, NOT, ANDM, ORM = range(6)


class RefSrc(C.Structure):
    _fields_ = [("data", C.c_void_p), ("fk", C.c_void_p),
                ("dimsz", C.c_uint32), ("elem_bytes", C.c_uint32)]


class RefCol(C.Structure):
    _fields_ = [("middle", C.c_void_p * 4), ("leftmost", C.c_void_p),
                ("rightmost", C.c_void_p)]


class RefVprg(C.Structure):
    _fields_ = [("cols", RefCol * 4), ("factsz", C.c_uint32),
                ("instrs", (C.c_uint8 * 4) * 6)]


class RefPred(C.Structure):
    _fields_ = [("src", RefSrc), ("lo", C.c_uint64), ("hi", C.c_uint64)]


class RefKey(C.Structure):
    _fields_ = [("src", RefSrc), ("kind", C.c_uint32), ("div", C.c_uint32),
                ("base", C.c_int64), ("stride", C.c_int64),
                ("nbounds", C.c_uint32), ("bounds", C.c_uint64 * 8)]


class RefRowop(C.Structure):
    _fields_ = [("preds", RefPred * 4), ("npreds", C.c_uint32),
                ("keys", RefKey * 3), ("nkeys", C.c_uint32),
                ("measure", C.c_uint32), ("ma", RefSrc), ("mb", RefSrc)]


class RefSSB(C.Structure):
    _fields_ = [(n, C.c_void_p) for n in (
        "loCustKey", "loPartKey", "loOrderDate", "loExtendedPrice",
        "loRevenue", "loSupplyCost", "loQuantity", "loDiscount",
        "loSuppCity", "custCity", "partMfgr")]


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} missing: run `make -C oracle`")
        L = C.CDLL(LIB_PATH)
        vp, u32, u64, sz, i32 = C.c_void_p, C.c_uint32, C.c_uint64, C.c_size_t, C.c_int
        P = C.POINTER
        L.ref_create_bin.argtypes = [P(RefSrc), u32, P(u64), P(u64), sz, vp, sz]
        L.ref_create_bin.restype = None
        L.ref_create_join.argtypes = [P(RefSrc), u32, u32, u32, vp]
        L.ref_create_join.restype = None
        L.ref_vprg_is_nochk.argtypes = [P(RefVprg)]
        L.ref_vprg_eval.argtypes = [P(RefVprg), i32, vp, vp]
        L.ref_vprg_eval.restype = u64
        L.ref_query.argtypes = [P(RefRowop), vp, u32, u32, vp, i32]
        L.ref_query.restype = u64
        L.ref_query_range.argtypes = [P(RefRowop), vp, u32, u32, u32, vp, i32]
        L.ref_query_range.restype = u64
        L.ref_vprg_grpby.argtypes = [P(RefVprg), P(RefRowop), u32, vp]
        L.ref_vprg_grpby.restype = u64
        L.ref_ssb_date_to_year.argtypes = [C.c_uint16]
        L.ref_ssb_date_to_year.restype = u32
        L.ref_s1.argtypes = [P(RefSSB), sz] + [C.c_uint16] * 2 + [C.c_uint8] * 4
        L.ref_s1.restype = C.c_int64
        L.ref_s2.argtypes = [P(RefSSB), sz, u32, u32, C.c_uint8, C.c_uint8, vp]
        L.ref_s2.restype = None
        L.ref_s3.argtypes = [P(RefSSB), sz] + [C.c_uint8] * 4 + [C.c_uint16] * 2 + [vp]
        L.ref_s3.restype = None
        L.ref_s4.argtypes = [P(RefSSB), sz] + [C.c_uint8] * 4 + [C.c_uint16] * 4 + [vp]
        L.ref_s4.restype = None
        L.ref_synth.argtypes = [vp, vp, u32, vp, vp, vp, sz] + [u64] * 6 + [vp]
        L.ref_synth.restype = None
        L.ref_mlp_loss_grad.argtypes = [vp, vp, i32, i32, i32, i32] + [vp] * 9
        L.ref_mlp_loss_grad.restype = C.c_double
        L.ref_wah_select.argtypes = [P(RefSrc), u32, u64, u64, vp]
        L.ref_wah_select.restype = None
        L.ref_wah_groups.argtypes = [vp, u32, vp]
        L.ref_wah_groups.restype = None
        L.ref_wah_compress.argtypes = [vp, sz, vp]
        L.ref_wah_compress.restype = sz
        L.ref_wah_scan.argtypes = [vp, sz, vp]
        L.ref_wah_scan.restype = None
        L.ref_wah_and_no1.argtypes = [vp, vp, sz, vp, vp, sz, vp]
        L.ref_wah_and_no1.restype = None
        L.ref_wah_decompress.argtypes = [vp, vp, sz, vp, sz]
        L.ref_wah_decompress.restype = None
        L.ref_wah_longest_1fill.argtypes = [vp, sz]
        L.ref_wah_longest_1fill.restype = u64
        _lib = L
    return _lib


def _ptr(a):
    return None if a is None else a.ctypes.data


def src_of(schema, s):
    """RefSrc of a merle_amd.spec.Src over the schema's host arrays."""
    c = schema[s.col]
    r = RefSrc()
    r.data = _ptr(c.host)
    r.elem_bytes = c.host.dtype.itemsize
    if s.fk:
        r.fk = _ptr(schema[s.fk].host)
        r.dimsz = c.host.size
    return r


def raw_src(data, fk=None, dimsz=0):
    r = RefSrc()
    r.data = _ptr(data)
    r.elem_bytes = data.dtype.itemsize
    r.fk = _ptr(fk)
    r.dimsz = dimsz
    return r


def rowop_of(spec, schema):
    op = RefRowop()
    op.npreds = len(spec.preds)
    for i, p in enumerate(spec.preds):
        op.preds[i].src = src_of(schema, p.src)
        op.preds[i].lo, op.preds[i].hi = p.lo, p.hi
    op.nkeys = len(spec.keys)
    for i, k in enumerate(spec.keys):
        kp = op.keys[i]
        kp.src = src_of(schema, k.src)
        kp.kind, kp.div, kp.base, kp.stride = k.kind, k.div, k.base, k.stride
        kp.nbounds = len(k.bounds)
        for j, b in enumerate(k.bounds):
            kp.bounds[j] = b
    op.measure = spec.measure
    if spec.ma is not None:
        op.ma = src_of(schema, spec.ma)
    if spec.mb is not None:
        op.mb = src_of(schema, spec.mb)
    return op


def instrs_of(instrs):
    arr = np.zeros((6, 4), dtype=np.uint8)
    for i, ins in enumerate(instrs):
        arr[i, :3] = ins[:3]
    return arr


def create_bin(col, mins, maxs, fk=None, dimsz=0):
    """-> uint32 array (nb, stride_words) in the reference bitmap layout."""
    n = col.size if fk is None else fk.size
    nb = len(mins)
    W = (n + 31) // 32
    stride = ((W + 63) // 64) * 64
    out = np.zeros((nb, stride), dtype=np.uint32)
    s = raw_src(col, fk, dimsz)
    lib().ref_create_bin(C.byref(s), n, (C.c_uint64 * nb)(*mins),
                         (C.c_uint64 * nb)(*maxs), nb, out.ctypes.data, stride)
    return out


def create_join(col, lo, hi, fk=None, dimsz=0):
    n = col.size if fk is None else fk.size
    W = (n + 31) // 32
    out = np.zeros(((W + 63) // 64) * 64, dtype=np.uint32)
    s = raw_src(col, fk, dimsz)
    lib().ref_create_join(C.byref(s), n, lo, hi, out.ctypes.data)
    return out


def vprg_of(cols, factsz, instrs):
    """cols: list of (middle [arrays], leftmost array|None, rightmost|None)."""
    p = RefVprg()
    keep = []
    for i, (mid, l, r) in enumerate(cols):
        for k, m in enumerate(mid):
            p.cols[i].middle[k] = _ptr(m)
            keep.append(m)
        p.cols[i].leftmost = _ptr(l)
        p.cols[i].rightmost = _ptr(r)
    p.factsz = factsz
    for i, ins in enumerate(instrs):
        for j in range(3):
            p.instrs[i][j] = ins[j]
    return p


def vprg_eval(p, chk=-1):
    W = (p.factsz + 31) // 32
    poss = np.zeros(W + 2, dtype=np.uint32)
    unc = np.zeros(W + 2, dtype=np.uint32)
    lib().ref_vprg_eval(C.byref(p), chk, poss.ctypes.data, unc.ctypes.data)
    return poss[:W], unc[:W]


def query(spec, schema, nthreads=0, row0=None, row1=None):
    out = np.zeros(spec.nr_grp, dtype=np.int64)
    op = rowop_of(spec, schema)
    ins = instrs_of(spec.instrs)
    if row0 is None:
        sel = lib().ref_query(C.byref(op), ins.ctypes.data, schema.nrows,
                              spec.nr_grp, out.ctypes.data, nthreads)
    else:
        sel = lib().ref_query_range(C.byref(op), ins.ctypes.data, row0, row1,
                                    spec.nr_grp, out.ctypes.data, nthreads)
    return out, sel


def vprg_grpby(p, spec, schema):
    out = np.zeros(spec.nr_grp, dtype=np.int64)
    op = rowop_of(spec, schema)
    sel = lib().ref_vprg_grpby(C.byref(p), C.byref(op), spec.nr_grp,
                               out.ctypes.data)
    return out, sel


def mlp_loss_grad(x, t, W1, b1, W2, b2):
    x = np.ascontiguousarray(x, np.float32)
    t = np.ascontiguousarray(t, np.float32)
    n, din = x.shape
    dh, dout = W1.shape[0], W2.shape[0]
    g = [np.zeros_like(a, dtype=np.float32) for a in (W1, b1, W2, b2)]
    y = np.zeros((n, dout), np.float32)
    loss = lib().ref_mlp_loss_grad(
        x.ctypes.data, t.ctypes.data, n, din, dh, dout,
        *[np.ascontiguousarray(a, np.float32).ctypes.data for a in (W1, b1, W2, b2)],
        *[a.ctypes.data for a in g], y.ctypes.data)
    return loss, g, y


# ---------------------------------------------------------------- WAH
def wah_ngroups(n):
    """31-row groups of an n-row bitmap."""
    return (n + 30) // 31


def wah_select(col, lo, hi, fk=None, dimsz=0):
    """Groups (31 rows per uint32, first row at bit 30) of [lo <= v < hi]."""
    n = len(fk) if fk is not None else len(col)
    src = raw_src(col, fk, dimsz)
    out = np.zeros(wah_ngroups(n), np.uint32)
    lib().ref_wah_select(C.byref(src), n, lo, hi, _ptr(out))
    return out


def wah_groups(bitmap, n):
    out = np.zeros(wah_ngroups(n), np.uint32)
    lib().ref_wah_groups(_ptr(np.ascontiguousarray(bitmap, np.uint32)), n,
                         _ptr(out))
    return out


def wah_compress(groups):
    g = np.ascontiguousarray(groups, np.uint32)
    out = np.zeros(max(1, len(g)), np.uint32)
    k = lib().ref_wah_compress(_ptr(g), len(g), _ptr(out))
    return out[:k].copy()


def wah_scan(wah):
    w = np.ascontiguousarray(wah, np.uint32)
    out = np.zeros(len(w) + 1, np.uint32)
    lib().ref_wah_scan(_ptr(w), len(w), _ptr(out))
    return out


def wah_and_no1(up, down):
    up = np.ascontiguousarray(up, np.uint32)
    down = np.ascontiguousarray(down, np.uint32)
    su, sd = wah_scan(up), wah_scan(down)
    out = np.zeros(max(1, len(down)), np.uint32)
    lib().ref_wah_and_no1(_ptr(up), _ptr(su), len(up), _ptr(down), _ptr(sd),
                          len(down), _ptr(out))
    return out[:len(down)].copy()


def wah_decompress(wah, ng):
    w = np.ascontiguousarray(wah, np.uint32)
    s = wah_scan(w)
    out = np.zeros(max(1, ng), np.uint32)
    lib().ref_wah_decompress(_ptr(w), _ptr(s), len(w), _ptr(out), ng)
    return out[:ng].copy()


def wah_longest_1fill(wah):
    w = np.ascontiguousarray(wah, np.uint32)
    return int(lib().ref_wah_longest_1fill(_ptr(w), len(w)))

