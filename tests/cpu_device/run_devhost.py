"""Random programs through the host build of the device query functions
(devhost.cpp) against the oracle.  Run as a child process (the library may
be built with AddressSanitizer, which must be preloaded):

    python run_devhost.py <libdevhost.so> <seeds>
"""
import ctypes as C
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path[:0] = [os.path.dirname(os.path.dirname(HERE)), os.path.dirname(HERE)]

import datagen  # noqa: E402
from merle_amd import abi, spec as S  # noqa: E402
from oracle import cpuref  # noqa: E402

L = C.CDLL(sys.argv[1])
L.dh_alloc.restype = C.c_void_p
L.dh_alloc.argtypes = [C.c_size_t]
for f in (L.dh_scan, L.dh_fused):
    f.argtypes = [C.POINTER(abi.Vprg), C.POINTER(abi.RowOp), C.c_uint32, C.c_void_p]
L.dh_conj_words.restype = C.c_long
L.dh_chk_conj_words.restype = C.c_long
L.dh_key_div.restype = C.c_uint64
L.dh_key_div.argtypes = [C.c_uint32, C.c_uint64]

# The key division by multiplication equals integer division for divisors
# 1..300, powers of two and their neighbours, random 32-bit divisors, and
# dividends at every boundary of 32- and 64-bit ranges plus random ones.
rng0 = np.random.default_rng(5)
divs = list(range(1, 301)) + [1 << k for k in range(32)] + \
    [(1 << k) + 1 for k in range(1, 32)] + [(1 << k) - 1 for k in range(2, 33)] + \
    [int(x) for x in rng0.integers(1, 1 << 32, 200)]
ndiv = 0
for d in divs:
    xs = [0, 1, d - 1, d, d + 1, 2 * d - 1, 2 * d, (1 << 32) - 1, 1 << 32,
          (1 << 63) - 1, 1 << 63, (1 << 64) - 1, (1 << 64) - d] + \
        [int(x) for x in rng0.integers(0, 1 << 63, 8, dtype=np.uint64)] + \
        [int(x) for x in rng0.integers(0, 1 << 32, 8)]
    for x in xs:
        got = L.dh_key_div(d, x)
        if got != x // d:
            print(f"key division wrong: {x} / {d} = {got}, want {x // d}")
            sys.exit(96)
        ndiv += 1
print(f"key division: {ndiv} quotients exact")

# The re-check of a candidate row looks the program up in a truth table and
# does not load the columns the host marks as known.  Over 256 rows (row
# m + 16e: column c's exact predicate is bit c of m; for a column with edge
# bins, bit c of e puts the row in its edge bin) the oracle evaluates each
# random program three-valued (P, U): the truth table must equal the
# program's two-valued result on the 16 predicate patterns, and for every
# row in P the result with the known columns substituted must equal it.
L.dh_prg_masks.argtypes = [C.POINTER(abi.Vprg), C.c_void_p]
dummy = L.dh_alloc(256)
rows = np.arange(256)
mpat, epat = rows % 16, rows // 16


def bits(mask):
    w = np.zeros(8, np.uint32)
    for r in np.nonzero(mask)[0]:
        w[r // 32] |= np.uint32(1 << (r % 32))
    return w


def row_bits(words):
    return np.array([(int(words[r // 32]) >> (r % 32)) & 1 for r in range(256)], bool)


CNF_SHAPES = [[(abi.ORM, 0, 0), (abi.ANDM, 0, 1), (abi.ANDM, 0, 2)],
              [(abi.ORM, 0, 0), (abi.ORM, 0, 1), (abi.ANDM, 0, 2)],
              [(abi.ORM, 0, 0), (abi.ORM, 1, 1), (abi.NOT, 1, 1), (abi.AND, 0, 1)],
              [(abi.ORM, 0, 0), (abi.ORM, 1, 1), (abi.NOT, 1, 1), (abi.AND, 0, 1),
               (abi.ANDM, 0, 2), (abi.ANDM, 0, 3)]]
nprog = nconj = 0
for it in range(600):
    ins = CNF_SHAPES[it % 4] if it % 2 else datagen.random_program(rng0, int(rng0.integers(1, 5)))
    edge = [bool(rng0.random() < 0.5) for _ in range(4)]
    ref_cols = []
    prg = abi.Vprg()
    for c in range(4):
        pred = ((mpat >> c) & 1).astype(bool)
        inedge = ((epat >> c) & 1).astype(bool) & edge[c]
        mid = bits(pred & ~inedge)
        ref_cols.append(([mid], bits(inedge) if edge[c] else None, None))
        prg.cols[c].middle[0] = dummy
        if edge[c]:
            prg.cols[c].leftmost = dummy
    prg.factsz = 256
    for i, (o, d, s_) in enumerate(ins):
        prg.instrs[i] = abi.Instr(o, d, s_, 0)
    out4 = (C.c_uint32 * 4)()
    if L.dh_prg_masks(C.byref(prg), out4) != 0:
        print("dh_prg_masks failed", ins)
        sys.exit(95)
    truth, known, kval, conj = list(out4)
    exact = row_bits(cpuref.vprg_eval(cpuref.vprg_of(
        [([bits((mpat >> c) & 1 == 1)], None, None) for c in range(4)], 256, ins), 0)[0])
    for m in range(16):
        if ((truth >> m) & 1) != exact[m]:
            print(f"truth table wrong: {ins} m={m}")
            sys.exit(94)
    poss, _ = cpuref.vprg_eval(cpuref.vprg_of(ref_cols, 256, ins), 1)
    for r in np.nonzero(row_bits(poss))[0]:
        m = int(mpat[r])
        if ((truth >> m) & 1) != ((truth >> ((m & ~known) | kval)) & 1):
            print(f"unloaded columns change a candidate row: {ins} edges={edge} row={r}")
            sys.exit(93)
    nprog += 1
    nconj += conj
print(f"re-check tables: {nprog} programs ({nconj} conjunctions) agree with the oracle")


class HostArray:
    """A column in a heap block of exactly its size (bounds-checked)."""

    def __init__(self, a):
        a = np.ascontiguousarray(a)
        self.n, self.dtype = a.size, a.dtype
        self.ptr = L.dh_alloc(a.nbytes)
        C.memmove(self.ptr, a.ctypes.data, a.nbytes)


def bitmap(words, n):
    w64 = (n + 63) // 64
    w = np.zeros(w64 * 2, np.uint32)
    w[:min(len(words), w64 * 2)] = words[:w64 * 2]
    return HostArray(w).ptr


for seed in range(int(sys.argv[2])):
    rng = np.random.default_rng(1000 + seed)
    n = int(rng.choice([1, 63, 64, 65, 5000, 70001, 300007]))
    schema = datagen.small_schema(n, seed)
    for c in schema.cols.values():
        c.dev = HostArray(c.host)
    for it in range(6):
        spec = datagen.random_spec(rng, schema, extra_groups=int(rng.integers(0, 3)))
        want, _ = cpuref.query(spec, schema)
        cols = []
        for p in spec.preds:
            col = None
            if rng.random() < 0.6:   # equal-width bins, edges -> candidate check
                hi_v = int(max(p.hi, 2)) * 2
                width = max(1, hi_v // 12)
                bounds = list(range(0, width * 13, width))
                fk = schema[p.src.fk].host if p.src.fk else None
                bins = cpuref.create_bin(schema[p.src.col].host, bounds[:-1], bounds[1:],
                                         fk, schema[p.src.col].n if fk is not None else 0)
                fb = S.from_bin(bounds, p.lo, p.hi)
                if fb is not None:
                    mid, left, right = fb
                    col = S.BitmapCol([bitmap(bins[i], n) for i in mid],
                                      None if left is None else bitmap(bins[left], n),
                                      None if right is None else bitmap(bins[right], n))
            if col is None:
                fk = schema[p.src.fk].host if p.src.fk else None
                ex = cpuref.create_join(schema[p.src.col].host, p.lo, min(p.hi, 2**32 - 1),
                                        fk, schema[p.src.col].n if fk is not None else 0)
                col = S.BitmapCol([bitmap(ex, n)])
            cols.append(col)
        prg = S.to_vprg(cols, n, spec.instrs)
        op = S.to_rowop(spec, schema)
        for name, f in (("fused", L.dh_fused), ("scan", L.dh_scan)):
            out = np.zeros(spec.nr_grp, np.int64)
            st = f(C.byref(prg), C.byref(op), spec.nr_grp, out.ctypes.data)
            if st != 0:
                print("STATUS", st, seed, it, name, spec)
                sys.exit(1)
            if not np.array_equal(out, want):
                print("MISMATCH", seed, it, name, spec)
                sys.exit(1)
    print("seed", seed, "n", n, "ok; conjunction-form words",
          L.dh_conj_words(), "three-valued", L.dh_chk_conj_words(), flush=True)
