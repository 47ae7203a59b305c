"""The learned plan chooser: data, training and evaluation.

Config 4: a scale-20 synthetic Zipf schema (120M rows, 8 int32 columns,
s cycling over 0.5 / 1.0 / 1.5 on values 0..999, 32 equal-width bins per
column), seeded random bitmap programs, every candidate plan of each
(mbx_plan_candidates: early-out orders, eager, materialise, scan) timed on
the GPU, and a 16 -> 64 -> 1 fp32 MLP (the library's HIP kernels) trained on
the log runtimes.  The chooser scores every candidate plan of a query and
runs the cheapest; the static heuristic (sparsest column first, fused) is
the fallback and the yardstick: every plan returns identical sums.  A share
of the programs put a misaligned range over the head of a skewed column, so
that an edge bin leaves most rows to re-check -- the case where the
heuristic's fused plan lost to the scan before the fused kernels re-checked
dense ranges into exact words ahead of the row op; now the two are close.
"""
import ctypes as C
import math

import numpy as np

from . import abi
from .configs import col_seed, SEED
from .host import Timer, check, stride_words
from .spec import BitmapCol, Key, Pred, QuerySpec, Schema, Src, from_bin, \
    to_rowop, to_vprg

C4_ROWS = 120_000_000
C4_COLS = 8
C4_SKEW = (0.5, 1.0, 1.5)
C4_DOMAIN = 1000
C4_BOUNDS = tuple(range(0, 1025, 32))          # 32 bins of width 32


def config4_device(dev, n=C4_ROWS, seed=SEED + 4):
    lib = dev.lib
    sch = Schema(n)
    for c in range(C4_COLS):
        a = dev.alloc(4 * n, np.uint32, n)
        check(lib.mbx_gen_zipf_u32(dev.ctx, a.ptr, n, col_seed(seed, 40 + c),
                                   C4_DOMAIN, C4_SKEW[c % 3]))
        sch.add(f"c{c}", None, a)
    m = dev.alloc(4 * n, np.uint32, n)
    check(lib.mbx_gen_uniform_u32(dev.ctx, m.ptr, n, col_seed(seed, 49), 1, 100001))
    sch.add("m", None, m)
    dev.sync()
    return sch


def build_bins(dev, sch, cols, bounds=C4_BOUNDS):
    """{col: [bitmap ptr per bin]} via one create_bin pass per column."""
    sw = stride_words(sch.nrows)
    nb = len(bounds) - 1
    out = {}
    for c in cols:
        slab = dev.alloc(sw * 4 * nb)
        ref = abi.ColRef(sch[c].dev.ptr, None, 0, 4)
        check(dev.lib.mbx_create_bin_into(dev.ctx, C.byref(ref), sch.nrows,
                                          (C.c_uint64 * nb)(*bounds[:-1]),
                                          (C.c_uint64 * nb)(*bounds[1:]), nb,
                                          slab.ptr, sw))
        out[c] = [slab.ptr + j * sw * 4 for j in range(nb)]
    dev.sync()
    return out


HEAVY_SHARE = 0.15
# a plan replaces the heuristic's only if predicted cheaper by this much
# (log runtime: about 10 %)
MARGIN = 0.1


def random_program(rng, ncols=C4_COLS, bounds=C4_BOUNDS, heavy_share=HEAVY_SHARE):
    """A conjunction of 2-4 range predicates on distinct columns (80%), or
    an OR / ANDNOT shape; 1-4 bins per range, 30% misaligned.  With
    probability heavy_share (named "heavy...") the first predicate is a
    misaligned range starting inside the first bin of a column with s = 1.5
    (its edge bin holds most rows) and the others are wide."""
    heavy = rng.random() < heavy_share
    npred = int(rng.integers(2, 5))
    if heavy:
        skewed = [c for c in range(ncols) if C4_SKEW[c % 3] == 1.5]
        first = int(rng.choice(skewed))
        rest = rng.choice([c for c in range(ncols) if c != first],
                          size=npred - 1, replace=False)
        cols = [first] + [int(c) for c in rest]
    else:
        cols = [int(c) for c in rng.choice(ncols, size=npred, replace=False)]
    nb = len(bounds) - 1
    preds = []
    for i, c in enumerate(cols):
        if heavy and i == 0:
            lo = int(rng.integers(1, bounds[1]))
            hi = bounds[int(rng.integers(1, 4))]
        elif heavy:
            lo, hi = 0, bounds[int(rng.integers(2, 5))]   # 2-4 bins from 0
        else:
            j = int(rng.integers(nb))
            w = int(rng.integers(1, 5))
            lo, hi = bounds[j], bounds[min(nb, j + w)]
            if rng.random() < 0.3:
                lo = min(hi - 1, lo + int(rng.integers(1, 16)))
        preds.append(Pred(Src(f"c{c}"), lo, hi))
    u = rng.random()
    if heavy or u < 0.8:
        ins = [(abi.ORM, 0, 0)] + [(abi.ANDM, 0, i) for i in range(1, npred)]
    elif u < 0.9:
        ins = [(abi.ORM, 0, 0), (abi.ORM, 0, 1)] + \
              [(abi.ANDM, 0, i) for i in range(2, npred)]
    else:
        ins = [(abi.ORM, 0, 0), (abi.ORM, 1, 1), (abi.NOT, 1, 1),
               (abi.AND, 0, 1)] + [(abi.ANDM, 0, i) for i in range(2, npred)]
    keys, total = [], 1
    for _ in range(int(rng.integers(0, 3))):
        c = int(rng.integers(ncols))
        div = int(rng.choice([10, 100]))
        card = C4_DOMAIN // div
        if total * card > 10_000:
            break
        keys.append(Key(Src(f"c{c}"), abi.KEY_DIV, div, 0, total))
        total *= card
    return QuerySpec(preds, ins, keys, total, abi.MEAS_COL, Src("m"),
                     name="heavy" if heavy else "")


def bitmap_cols(spec, bins, bounds=C4_BOUNDS):
    """bounds: one bound list for every column, or {column: bounds}."""
    cols = []
    for p in spec.preds:
        b = bounds[p.src.col] if isinstance(bounds, dict) else bounds
        mid, l, r = from_bin(list(b), p.lo, p.hi)
        b = bins[p.src.col]
        cols.append(BitmapCol([b[i] for i in mid], None if l is None else b[l],
                              None if r is None else b[r]))
    return cols


class Prepared:
    """A program bound to the device: vprg, row op, column statistics,
    candidate plans."""

    def __init__(self, dev, sch, bins, spec, bounds=C4_BOUNDS):
        self.spec = spec
        self.prg = to_vprg(bitmap_cols(spec, bins, bounds), sch.nrows, spec.instrs)
        self.op = to_rowop(spec, sch)
        self.dens = (C.c_double * 4)()
        self.edge = (C.c_double * 4)()
        check(dev.lib.mbx_col_stats(dev.ctx, C.byref(self.prg), self.dens,
                                    self.edge))
        cand = (abi.Plan * 32)()
        cnt = C.c_int()
        check(dev.lib.mbx_plan_candidates(C.byref(self.prg), self.dens, cand, 32,
                                          C.byref(cnt)))
        self.cands = list(cand[:cnt.value])
        self.as_written = abi.Plan()
        check(dev.lib.mbx_plan_describe(C.byref(self.prg), C.byref(self.as_written)))
        self.heuristic = abi.Plan()
        check(dev.lib.mbx_plan_heuristic(C.byref(self.prg), self.dens,
                                         C.byref(self.heuristic)))

    def features(self, dev, plan):
        f = (C.c_float * abi.NFEAT)()
        check(dev.lib.mbx_plan_features_ex(C.byref(self.prg), self.dens, self.edge,
                                           self.spec.nr_grp, C.byref(plan), f))
        return np.frombuffer(f, dtype=np.float32).copy()

    def run(self, dev, plan, out, timer=None):
        if timer:
            timer.start()
        check(dev.lib.mbx_prog_grpby_plan(dev.ctx, C.byref(self.prg),
                                          C.byref(self.op), self.spec.nr_grp,
                                          out.ptr, C.byref(plan)))
        return timer.stop() if timer else None

    def time_all(self, dev, plans, out, timers):
        """Every plan once, back to back, each between its own event pair;
        one wait at the end.  -> [ms]."""
        for plan, t in zip(plans, timers):
            t.start()
            check(dev.lib.mbx_prog_grpby_plan(dev.ctx, C.byref(self.prg),
                                              C.byref(self.op), self.spec.nr_grp,
                                              out.ptr, C.byref(plan)))
            t.mark_stop()
        return [t.elapsed() for t in timers[:len(plans)]]


def collect(dev, sch, bins, nprog, plans_per_prog=None, seed=7, log=None,
            specs=None, bounds=C4_BOUNDS, heavy_share=HEAVY_SHARE):
    """Times the candidate plans of nprog programs (random config-4
    programs, or the given specs over bins of `bounds`) -- every candidate,
    or plans_per_prog random ones: -> (features [n,16] float32,
    log runtime [n] float32)."""
    rng = np.random.default_rng(seed)
    if specs is not None:
        nprog = len(specs)
    out = dev.alloc(8 * 10_000, np.int64, 10_000)
    timers = [Timer(dev) for _ in range(32)]
    X, Y = [], []
    for i in range(nprog):
        spec = specs[i] if specs is not None else \
            random_program(rng, heavy_share=heavy_share)
        p = Prepared(dev, sch, bins, spec, bounds)
        plans = p.cands
        if plans_per_prog is not None and plans_per_prog < len(plans):
            picks = rng.choice(len(plans), size=plans_per_prog, replace=False)
            plans = [plans[int(k)] for k in picks]
        for plan, ms in zip(plans, p.time_all(dev, plans, out, timers)):
            X.append(p.features(dev, plan))
            Y.append(math.log(max(ms, 1e-4)))
        if log and (i + 1) % 2000 == 0:
            log(f"collected {i + 1}/{nprog} programs ({len(Y)} timed plans)")
    for t in timers:
        t.close()
    out.free()
    return np.stack(X).astype(np.float32), np.array(Y, np.float32)[:, None]


class CostModel:
    """16 -> 64 -> 1 fp32 MLP on the device (mbx_mlp_*)."""

    def __init__(self, dev, seed=11):
        self.dev = dev
        m = C.c_void_p()
        check(dev.lib.mbx_mlp_create(dev.ctx, abi.NFEAT, 64, 1, seed, C.byref(m)))
        self.m = m
        self.ymean = 0.0

    def train(self, X, Y, steps=1500, lr=3e-3):
        self.ymean = float(Y.mean())
        dx = self.dev.upload(np.ascontiguousarray(X, np.float32))
        dy = self.dev.upload(np.ascontiguousarray(Y - self.ymean, np.float32))
        loss = C.c_double()
        check(self.dev.lib.mbx_mlp_train(self.dev.ctx, self.m, dx.ptr, dy.ptr,
                                         X.shape[0], steps, C.c_float(lr),
                                         C.byref(loss)))
        dx.free()
        dy.free()
        return loss.value

    def choose(self, p, margin=MARGIN):
        """The model's plan if it predicts it cheaper than the heuristic's
        by more than `margin` (log runtime), else the heuristic's."""
        best = abi.Plan()
        pred = C.c_float()
        check(self.dev.lib.mbx_plan_choose_ex(self.dev.ctx, self.m, C.byref(p.prg),
                                              p.dens, p.edge, p.spec.nr_grp,
                                              C.c_float(margin), C.byref(best),
                                              C.byref(pred)))
        return best

    def choose_fused(self, p):
        """The cheapest fused candidate by the model (the batched launch
        runs fused plans only)."""
        cands = [c for c in p.cands if (c.flags & 3) == abi.PLAN_FUSED]
        if not cands:
            return p.heuristic
        X = np.stack([p.features(self.dev, c) for c in cands]).astype(np.float32)
        dx = self.dev.upload(X)
        dy = self.dev.alloc(4 * len(cands), np.float32, len(cands))
        check(self.dev.lib.mbx_mlp_forward(self.dev.ctx, self.m, dx.ptr, len(cands),
                                           dy.ptr))
        y = dy.host()
        dx.free()
        dy.free()
        return cands[int(np.argmin(y))]

    def close(self):
        self.dev.lib.mbx_mlp_destroy(self.dev.ctx, self.m)


def evaluate(dev, sch, bins, model, nprog, seed=99, reps=3,
             heavy_share=HEAVY_SHARE):
    """Held-out programs: median runtime of the as-written plan, the static
    heuristic and the learned choice; all three must return identical sums.
    The gain is also reported without the heavy programs, and the plans the
    chooser picked are counted by kind."""
    rng = np.random.default_rng(seed)
    out = dev.alloc(8 * 10_000, np.int64, 10_000)
    t = Timer(dev)
    tot = {"all": [0.0, 0.0, 0.0], "heavy": [0.0, 0.0, 0.0],
           "other": [0.0, 0.0, 0.0]}
    kinds = {"fused": 0, "eager": 0, "materialise": 0, "scan": 0}
    same = True
    nheavy = 0
    for _ in range(nprog):
        p = Prepared(dev, sch, bins, random_program(rng, heavy_share=heavy_share))
        chosen = model.choose(p)
        k = chosen.flags & abi.PLAN_MASK
        kinds["scan" if k == abi.PLAN_SCAN else "materialise"
              if k == abi.PLAN_MATERIALISE else "eager"
              if chosen.flags & abi.PLAN_EAGER else "fused"] += 1
        res, times = [], []
        for plan in (p.as_written, p.heuristic, chosen):
            ms = sorted(p.run(dev, plan, out, t) for _ in range(reps))[reps // 2]
            times.append(ms)
            res.append(out.host(p.spec.nr_grp))
        same &= all(np.array_equal(res[0], r) for r in res[1:])
        heavy = p.spec.name == "heavy"
        nheavy += heavy
        for key in ("all", "heavy" if heavy else "other"):
            for j in range(3):
                tot[key][j] += times[j]
    t.close()
    out.free()
    tw, th, tl = tot["all"]
    ow, oh, ol = tot["other"]
    return dict(programs=nprog, heavy_programs=nheavy,
                as_written_ms=round(tw, 3), heuristic_ms=round(th, 3),
                learned_ms=round(tl, 3),
                learned_vs_as_written=round(tw / tl, 4),
                learned_vs_heuristic=round(th / tl, 4),
                without_heavy={"heuristic_ms": round(oh, 3),
                               "learned_ms": round(ol, 3),
                               "learned_vs_heuristic": round(oh / ol, 4) if ol else None},
                learned_plan_kinds=kinds, identical_results=same)
