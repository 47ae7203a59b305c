// mbx_demo -- the SSB driver of the reference's demo API
// (ssb_demoall, main.c:6-12; ssb/fixedlayout.cu, ssb/lambdas.cuh:154-243)
// written against the mybmpidx mirror.
//
//   mbx_demo <ssb_dir> [reps] [synth_rows] [sums_file]
//
// <ssb_dir> holds the binary column files of ssb-dbgen-lite -b (4-byte
// header = element size in bytes, ssb-dbgen-lite.c:192-200).  The driver
// loads the fact and dimension columns it needs, builds the sparse and dense
// bin sets of ssb/bmpidx-create.cu:8-30 with create_bin (customer city and
// part mfgr projected through their foreign keys), then runs the 13 flights
// (the arguments of ssb_cpujoin, ssb/cpuref.c:244-302) under the Sparse,
// Medium and Dense programs of ssb/fixedlayout.cu (col::from_bin and
// col::from_bin2; edge bins -> candidate checks) and then under the
// controlled layouts of ssb/controlled-demo.cu (MOREWORK: recipe::perfect,
// many_or and candchk -- exact, three-way-split and misaligned bins built
// per query with create_join), timed per plan as the mean of `reps` runs.
// It prints the reference's five-block report (drawFigs/drawFigs.py:15-46):
// index creation and the SSB block in full (BinType Sparse / Medium / Dense
// / Perfect / AllAlig / MisAlig), the method and synthetic blocks in full
// when synth_rows > 0 (seeded columns of that many rows, see synth_suite),
// the WAH / RTScan block with each flight's WAH index-only time (its
// predicates' WAH bitmaps ANDed, the result decompressed; RTScan has no
// counterpart and reads nan), and the index block's WAH columns (every
// bin's WAH bitmap built and compressed, and its size).  SSB columns: Join
// = the bitmap-free scan (Sparse rows
// only, as DOWORK), Ours = fused under the library's static plan, PrgStg1 /
// Stg2 / Stg3 = the three materialise stages, Basefus = the base_fuse
// ablation (method 1 with the program's registers in device memory),
// NfuStg1 = the first stage with its registers in device memory, Dedicat
// = the kernel built for the program's shape (MBX_PLAN_DEDICATED; fixed
// layouts only, as in the reference, and 0 where the library holds no
// kernel for the shape).  Every
// plan's int64 sums must equal the scan's, or the driver exits with status
// 1.  With a
// sums_file the driver also writes every SSB row's sums there, one line per
// (flight, bin type, plan): name, bin type, plan (join / ours / stg / ded), then
// the sums, tab-separated.
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "mbx_bmpidx.hpp"

using namespace mybmpidx;

namespace {

struct Col {
  void *d = nullptr;
  size_t n = 0;
  uint32_t eb = 0;
};

bool load(mbx_ctx *ctx, const std::string &dir, const char *name, Col &c) {
  size_t nbytes = 0;
  uint32_t header = 0;
  const std::string path = dir + "/" + name;
  if (mbx_colfile_load(ctx, path.c_str(), &c.d, &nbytes, &header) != MBX_OK) {
    std::fprintf(stderr, "cannot load %s: %s\n", path.c_str(), mbx_last_error());
    return false;
  }
  c.eb = header;
  c.n = nbytes / header;
  return true;
}

// ssb/ssbdemo.h:10-26, day offsets from 1992-01-01
enum : uint64_t {
  D920101 = 0, D930101 = 366, D940101 = 731, D940201 = 762, D940204 = 765,
  D940211 = 772, D940301 = 790, D950101 = 1096, D960101 = 1461,
  D970101 = 1827, D971201 = 2161, D980101 = 2192, D990101 = 2557
};
// ssbDateToYear (ssb/ssbdemo.h:29-39) as bucket bounds: year = #{b <= date}
const uint64_t kYear[] = {D930101, D940101, D950101, D960101, D970101, D980101};

struct Flight {
  const char *name;
  int kind;            // 1..4: s1ref..s4ref
  uint64_t a[8];
  uint32_t ngrp;       // NGRP_S* (ssb/ssbdemo.h:79-87)
};
const Flight kFlights[] = {
    {"SSB11", 1, {D930101, D940101, 1, 4, 0, 25}, 1},
    {"SSB12", 1, {D940101, D940201, 4, 7, 26, 36}, 1},
    {"SSB13", 1, {D940204, D940211, 5, 8, 26, 36}, 1},
    {"SSB21", 2, {40, 80, 150, 200}, 280},
    {"SSB22", 2, {260, 268, 200, 250}, 56},
    {"SSB23", 2, {260, 261, 50, 100}, 7},
    {"SSB31", 3, {200, 250, 200, 250, D920101, D980101}, 150},
    {"SSB32", 3, {190, 200, 190, 200, D920101, D980101}, 600},
    {"SSB33", 3, {51, 55, 51, 55, D920101, D980101}, 96},
    {"SSB34", 3, {51, 55, 51, 55, D971201, D980101}, 16},
    {"SSB41", 4, {150, 200, 150, 200, 0, 400, D920101, D990101}, 350},
    {"SSB42", 4, {150, 200, 150, 200, 0, 400, D970101, D990101}, 100},
    {"SSB43", 4, {150, 200, 190, 200, 120, 160, D970101, D990101}, 800},
};

// predicate sources, in bin-set order
enum Src { DATE, DISC, QTY, SCITY, CCITY, MFGR, NSRC };
const char *kSrcName[NSRC] = {"loOrderDate", "loDiscount", "loQuantity",
                              "loSuppCity",  "custCity",   "partMfgr"};
// ssb/bmpidx-create.cu:8-30: sparse and dense bin bounds per source
const std::vector<size_t> kBounds[2][NSRC] = {
    {{D920101, D930101, D940101, D950101, D960101, D970101, D980101, D990101},
     {0, 2, 5, 8, 11},
     {1, 11, 21, 31, 41, 51},
     {0, 50, 100, 150, 200, 250},
     {0, 50, 100, 150, 200, 250},
     {0, 200, 400, 600, 800, 1000}},
    {{D940101, D940201, D940301, D971201, D980101},
     {0, 1, 2, 3, 4, 5, 6, 7, 8, 9, 10, 11},
     {1, 6, 11, 16, 21, 26, 31, 36, 41, 46, 51},
     {50, 60, 190, 200},
     {50, 60, 190, 200},
     {40, 80, 120, 160, 240, 280}}};

struct Tables {
  Col c[NSRC], fk_cust, fk_part, extprice, revenue, supplycost;
  uint32_t nrows = 0;
  mbx_colref ref(int s) const {
    mbx_colref r{};
    r.data = c[s].d;
    r.elem_bytes = c[s].eb;
    if (s == CCITY) {
      r.fk = static_cast<const uint32_t *>(fk_cust.d);
      r.dimsz = (uint32_t)c[s].n;
    } else if (s == MFGR) {
      r.fk = static_cast<const uint32_t *>(fk_part.d);
      r.dimsz = (uint32_t)c[s].n;
    }
    return r;
  }
  mbx_colref plain(const Col &col) const {
    mbx_colref r{};
    r.data = col.d;
    r.elem_bytes = col.eb;
    return r;
  }
};

struct Pred {
  int src;
  uint64_t lo, hi;
};

void key(mbx_keypart &k, const mbx_colref &src, uint32_t kind, uint32_t div,
         int64_t base, int64_t stride) {
  std::memset(&k, 0, sizeof(k));
  k.src = src;
  k.kind = kind;
  k.div = div;
  k.base = base;
  k.stride = stride;
}
void year_key(mbx_keypart &k, const mbx_colref &date, uint64_t dmin,
              int64_t stride) {
  key(k, date, MBX_KEY_BUCKET, 1, 0, stride);
  int y = 0;
  for (uint64_t b : kYear) y += b <= dmin;
  k.base = y;
  k.nbounds = 6;
  for (int i = 0; i < 6; ++i) k.bounds[i] = kYear[i];
}
// (kind, div, base, width) of a scaled key term (s3ref/s4ref)
void scaled(uint64_t lo, uint64_t hi, uint64_t span, uint32_t div,
            uint32_t &kind, uint32_t &d, int64_t &base, int64_t &width) {
  if (hi - lo >= span) {
    kind = MBX_KEY_DIV;
    d = div;
    base = lo / div;
    width = hi / div - lo / div;
  } else {
    kind = MBX_KEY_IDENT;
    d = 1;
    base = lo;
    width = hi - lo;
  }
}

// The flight as predicates + row op: group formulas of s1ref..s4ref
// (ssb/cpuref.c:141-242) and s1op..s4op (ssb/lambdas.cuh:19-152).
std::vector<Pred> flight_op(const Flight &f, const Tables &t, mbx_rowop &op) {
  std::memset(&op, 0, sizeof(op));
  const uint64_t *a = f.a;
  std::vector<Pred> p;
  if (f.kind == 1) {
    p = {{DATE, a[0], a[1]}, {DISC, a[2], a[3]}, {QTY, a[4], a[5]}};
    op.measure = MBX_MEAS_MUL;
    op.ma = t.plain(t.extprice);
    op.mb = t.ref(DISC);
  } else if (f.kind == 2) {
    p = {{MFGR, a[0], a[1]}, {SCITY, a[2], a[3]}};
    op.nkeys = 2;
    key(op.keys[0], t.ref(MFGR), MBX_KEY_IDENT, 1, (int64_t)a[0], 1);
    year_key(op.keys[1], t.ref(DATE), 0, (int64_t)(a[1] - a[0]));
    op.keys[1].base = 0;
    op.measure = MBX_MEAS_COL;
    op.ma = t.plain(t.revenue);
  } else if (f.kind == 3) {
    p = {{CCITY, a[0], a[1]}, {SCITY, a[2], a[3]}, {DATE, a[4], a[5]}};
    uint32_t ck, cd, sk, sd;
    int64_t cb, ca, sb, sw;
    if (a[1] - a[0] >= 50) {   // both cities scale by 10 when the c-range does
      ck = sk = MBX_KEY_DIV;
      cd = sd = 10;
      cb = a[0] / 10; ca = a[1] / 10 - a[0] / 10;
      sb = a[2] / 10; sw = a[3] / 10 - a[2] / 10;
    } else {
      ck = sk = MBX_KEY_IDENT;
      cd = sd = 1;
      cb = a[0]; ca = a[1] - a[0];
      sb = a[2]; sw = a[3] - a[2];
    }
    op.nkeys = 3;
    year_key(op.keys[0], t.ref(DATE), a[4], ca * sw);
    key(op.keys[1], t.ref(CCITY), ck, cd, cb, ca);
    key(op.keys[2], t.ref(SCITY), sk, sd, sb, 1);
    op.measure = MBX_MEAS_COL;
    op.ma = t.plain(t.revenue);
  } else {
    p = {{CCITY, a[0], a[1]}, {SCITY, a[2], a[3]}, {MFGR, a[4], a[5]},
         {DATE, a[6], a[7]}};
    uint32_t sk, sd, pk, pd;
    int64_t sb, sa, pb, pw;
    scaled(a[2], a[3], 50, 10, sk, sd, sb, sa);
    scaled(a[4], a[5], 200, 40, pk, pd, pb, pw);
    op.nkeys = 3;
    year_key(op.keys[0], t.ref(DATE), a[6], sa * pw);
    key(op.keys[1], t.ref(SCITY), sk, sd, sb, sa);
    key(op.keys[2], t.ref(MFGR), pk, pd, pb, 1);
    op.measure = MBX_MEAS_SUB;
    op.ma = t.plain(t.revenue);
    op.mb = t.plain(t.supplycost);
  }
  op.npreds = (uint32_t)p.size();
  for (size_t i = 0; i < p.size(); ++i) {
    op.preds[i].src = t.ref(p[i].src);
    op.preds[i].lo = p[i].lo;
    op.preds[i].hi = p[i].hi;
  }
  return p;
}

// The programs of s1fix..s4fix (ssb/fixedlayout.cu:35-140): columns,
// instructions, and the columns that DENSE() (col::from_bin2, sparse bins as
// the coarse level) replaces in the Medium and Dense programs.
struct FixedCol {
  int src;
  uint64_t lo, hi;
};
struct Fixed {
  std::vector<FixedCol> cols;
  std::vector<vprg::instr> ins;
  unsigned medium = 0, dense = 0;   // column bit masks
};
vprg::instr ins(vprg::ops o, uint8_t d, uint8_t s) {
  vprg::instr i;
  i.opcode = o;
  i.dstreg = d;
  i.src1 = s;
  return i;
}
Fixed fixed_layout(const Flight &f) {
  Fixed x;
  const uint64_t *a = f.a;
  const vprg::ops LOAD = vprg::ORM, ANDM = vprg::ANDM;
  if (f.kind == 1) {
    // fixedlayout.cu:147 passes quantity [1, 25) for SSB11 where ssb_cpujoin
    // passes [0, 25); quantity is at least 1 (ssb-dbgen-lite.c:139)
    x.cols = {{DATE, a[0], a[1]}, {DISC, a[2], a[3]},
              {QTY, a[4] ? a[4] : 1, a[5]}};
    x.ins = {ins(LOAD, 0, 0), ins(ANDM, 0, 1), ins(ANDM, 0, 2)};
    x.medium = 0b110;
    x.dense = 0b111;
  } else if (f.kind == 2) {
    x.cols = {{MFGR, a[0], a[1]}, {SCITY, a[2], a[3]}};
    x.ins = {ins(LOAD, 0, 0), ins(ANDM, 0, 1)};
    x.medium = 0b01;
    x.dense = 0b11;
  } else if (f.kind == 3) {
    uint64_t dmin = a[4], dmax = a[5];
    x.ins = {ins(LOAD, 0, 0), ins(ANDM, 0, 1)};
    if (dmin == D920101) {   // date < dmax as NOT(date in [dmax, 1999))
      dmin = dmax;
      dmax = D990101;
      x.ins.push_back(ins(LOAD, 1, 2));
      x.ins.push_back(ins(vprg::NOT, 1, 1));
      x.ins.push_back(ins(vprg::AND, 0, 1));
    } else {
      x.ins.push_back(ins(ANDM, 0, 2));
    }
    x.cols = {{CCITY, a[0], a[1]}, {SCITY, a[2], a[3]}, {DATE, dmin, dmax}};
    x.medium = 0b001;
    x.dense = 0b111;
  } else {
    x.cols = {{CCITY, a[0], a[1]}, {SCITY, a[2], a[3]}, {MFGR, a[4], a[5]}};
    x.ins = {ins(LOAD, 0, 0), ins(ANDM, 0, 1), ins(ANDM, 0, 2)};
    if (a[7] < D990101 || a[6] > D920101) {   // a date column only if bounded
      x.cols.push_back({DATE, a[6], a[7]});
      x.ins.push_back(ins(ANDM, 0, 3));
    }
    x.medium = 0b0101;
    x.dense = 0b0111;
  }
  return x;
}


// The library's static plan of a program (column densities -> sparsest
// column first; eager loads unless the early-out pays), made once outside
// the timed loop.  "Ours" runs the fused query under it.
bool static_plan(mbx_ctx *ctx, const mbx_vprg &ap, mbx_plan &plan) {
  double dens[MBX_MAXCOLS];
  if (mbx_col_density(ctx, &ap, dens) != MBX_OK ||
      mbx_plan_heuristic(&ap, dens, &plan) != MBX_OK) {
    std::fprintf(stderr, "plan: %s\n", mbx_last_error());
    return false;
  }
  return true;
}

// ------------------------------------------------ the synthetic suite
// synth_demoall (synth/cpuref.c:184-257).  The query (synth/synthdemo.h:
// 11-16): COUNT(*) GROUP BY dim.attr2 over the rows with (a1 in [lo, hi) OR
// a2 in [lo, hi)) AND dim.attr1[fk] in [lo, hi).  Data per case: a1, a2 =
// Zipf(s) on [0, 65535) stored at the case's bit width, fk = Zipf(s) on
// [0, N/100), dim.attr1[i] = i, dim.attr2 = uniform bytes.  The reference
// reads files of unseeded numpy zipf mod 65535 (synth/onegen_zipf.py:5-10,
// batchgen.sh:12-25); these columns are seeded and generated here.
// [lo, lo + 60) is the first window whose probability under the generator's
// pmf is at most the column selectivity of synth/cpuref.c:191-195.
constexpr uint32_t kSynthVals = 65535, kSynthWin = 60, kSynthGroups = 256;
const double kSynthColSel[11] = {.02, .04, .06, .09, .12, .18,
                                 .25, .39, .51, .67, .82};
const double kSynthFulSel[11] = {1 / 512.0, 1 / 256.0, 1 / 128.0, 1 / 64.0,
                                 1 / 32.0,  1 / 16.0,  1 / 8.0,   1 / 4.0,
                                 .4,        .6,        .8};

struct SynthCase {
  uint32_t n = 0, dimsz = 0, bw = 0;
  double s = 0;
  void *col[3] = {nullptr, nullptr, nullptr};   // a1, a2, fk
  void *dim1 = nullptr, *dim2 = nullptr;
  std::vector<double> pre;                       // prefix sums of the pmf
};

void synth_free(mbx_ctx *ctx, SynthCase &c) {
  for (void *&q : c.col) if (q) mbx_free(ctx, q), q = nullptr;
  if (c.dim1) mbx_free(ctx, c.dim1), c.dim1 = nullptr;
  if (c.dim2) mbx_free(ctx, c.dim2), c.dim2 = nullptr;
}

bool synth_case(mbx_ctx *ctx, uint32_t n, double s, uint32_t bw, int idx,
                SynthCase &c) {
  c.n = n;
  c.dimsz = n / 100;
  c.bw = bw;
  c.s = s;
  std::vector<uint32_t> h(n);
  for (int k = 0; k < 3; ++k) {
    const uint64_t seed = 0x5eed0000u + 16u * (uint64_t)idx + (uint64_t)k;
    mbx_gen_zipf_u32_host(h.data(), n, seed, k < 2 ? kSynthVals : c.dimsz, s);
    const size_t eb = k < 2 ? bw / 8 : 4;
    if (mbx_alloc(ctx, (size_t)n * eb, &c.col[k])) return false;
    if (eb == 2) {
      std::vector<uint16_t> h16(h.begin(), h.end());
      mbx_memcpy_h2d(ctx, c.col[k], h16.data(), (size_t)n * 2);
    } else {
      mbx_memcpy_h2d(ctx, c.col[k], h.data(), (size_t)n * 4);
    }
  }
  std::vector<uint32_t> iota(c.dimsz), u(c.dimsz);
  for (uint32_t i = 0; i < c.dimsz; ++i) iota[i] = i;
  mbx_gen_uniform_u32_host(u.data(), c.dimsz, 13458786, 0, 256);
  std::vector<uint8_t> b(u.begin(), u.end());
  if (mbx_alloc(ctx, (size_t)c.dimsz * 4, &c.dim1) ||
      mbx_alloc(ctx, c.dimsz, &c.dim2))
    return false;
  mbx_memcpy_h2d(ctx, c.dim1, iota.data(), (size_t)c.dimsz * 4);
  mbx_memcpy_h2d(ctx, c.dim2, b.data(), c.dimsz);
  c.pre.assign(kSynthVals + 1, 0.0);
  for (uint32_t v = 0; v < kSynthVals; ++v)
    c.pre[v + 1] = c.pre[v] + std::pow((double)v + 1.0, -s);
  return true;
}

// The query of selectivity index o: its recipe inputs and row op.
void synth_query(const SynthCase &c, int o, recipe &r, mbx_rowop &op) {
  uint32_t lo = 0;
  while (lo + kSynthWin < kSynthVals &&
         (c.pre[lo + kSynthWin] - c.pre[lo]) / c.pre[kSynthVals] > kSynthColSel[o])
    ++lo;
  const uint32_t hi = lo + kSynthWin;
  const uint32_t *fk = static_cast<const uint32_t *>(c.col[2]);
  r = recipe();
  r.factsz = c.n;
  std::memset(&op, 0, sizeof(op));
  op.npreds = 3;
  for (int k = 0; k < 3; ++k) {
    r.nbit[k] = k < 2 ? c.bw : 32;
    r.min[k] = lo;
    r.max[k] = hi;
    r.dimsz[k] = k < 2 ? recipe::nodim : c.dimsz;
    r.fk[k] = k < 2 ? nullptr : fk;
    r.attr[k] = k < 2 ? c.col[k] : c.dim1;
    op.preds[k].src.data = k < 2 ? c.col[k] : c.dim1;
    op.preds[k].src.fk = k < 2 ? nullptr : fk;
    op.preds[k].src.dimsz = k < 2 ? 0 : c.dimsz;
    op.preds[k].src.elem_bytes = k < 2 ? c.bw / 8 : 4;
    op.preds[k].lo = lo;
    op.preds[k].hi = hi;
  }
  op.nkeys = 1;
  key(op.keys[0], mbx_colref{c.dim2, fk, c.dimsz, 1}, MBX_KEY_IDENT, 1, 0, 1);
  op.measure = MBX_MEAS_COUNT;
}

void synth_instrs(vprg::instr (&ins3)[MAXNINSTR]) {
  ins3[0] = ins(vprg::ORM, 0, 0);
  ins3[1] = ins(vprg::ORM, 0, 1);
  ins3[2] = ins(vprg::ANDM, 0, 2);
}

// Block 3: method 1 against method 2 (synth_method, synth/synthdemo.cu:
// 147-204) on the skew-1.04, 32-bit case, the perfect recipe and the seven
// smallest selectivities.  Method 1 is the fused plan without compaction
// (MBX_PLAN_DIRECT), method 2 the fused plan.  Returns mismatches, -1 on an
// error.
template <class Timed>
int synth_methods(mbx_ctx *ctx, uint32_t n, Timed &&timed) {
  SynthCase c;
  if (!synth_case(ctx, n, 1.04, 32, 1, c)) {
    std::fprintf(stderr, "synthetic data: %s\n", mbx_last_error());
    synth_free(ctx, c);
    return -1;
  }
  void *d1 = nullptr, *d2 = nullptr;
  if (mbx_alloc(ctx, 8 * kSynthGroups, &d1) || mbx_alloc(ctx, 8 * kSynthGroups, &d2)) {
    synth_free(ctx, c);
    return -1;
  }
  int bad = 0;
  std::vector<int64_t> h1(kSynthGroups), h2(kSynthGroups);
  vprg::instr ins3[MAXNINSTR];
  synth_instrs(ins3);
  for (int o = 0; o < 7; ++o) {
    recipe r;
    mbx_rowop op;
    synth_query(c, o, r, op);
    vprg p = r.perfect(ins3);
    const mbx_vprg ap = p.to_abi();
    const float m1 = timed([&] {
      return mbx_prog_grpby(ctx, &ap, &op, kSynthGroups, static_cast<int64_t *>(d1),
                            MBX_PLAN_FUSED | MBX_PLAN_DIRECT);
    });
    const float m2 = timed([&] {
      return mbx_prog_grpby(ctx, &ap, &op, kSynthGroups, static_cast<int64_t *>(d2),
                            MBX_PLAN_FUSED);
    });
    mbx_memcpy_d2h(ctx, h1.data(), d1, 8 * kSynthGroups);
    mbx_memcpy_d2h(ctx, h2.data(), d2, 8 * kSynthGroups);
    bad += std::memcmp(h1.data(), h2.data(), 8 * kSynthGroups) != 0;
    p.release();
    std::printf("%.5f\t%.4f\t%.4f\n", kSynthFulSel[o], m1, m2);
    std::fflush(stdout);
  }
  mbx_free(ctx, d1);
  mbx_free(ctx, d2);
  synth_free(ctx, c);
  return bad;
}

// Block 4: one row per case and selectivity as synth_bmp times it
// (synth/synthdemo.cu:79-145): the scan, the perfect, many-OR and
// candidate-check recipes fused, and each recipe's three materialise
// stages.  Returns the number of plans whose sums differ from the scan's,
// or -1 on an error.
template <class Timed>
int synth_suite(mbx_ctx *ctx, uint32_t n, Timed &&timed) {
  static const double kSkew[5] = {1.04, 1.08, 1.12, 1.16, 1.20};
  const size_t nw = mbx_bitmap_stride_words(n);
  void *poss = nullptr, *unc = nullptr, *rows = nullptr, *count = nullptr,
       *out = nullptr, *want = nullptr;
  if (mbx_alloc(ctx, nw * 4, &poss) || mbx_alloc(ctx, nw * 4, &unc) ||
      mbx_alloc(ctx, (size_t)n * 4, &rows) || mbx_alloc(ctx, 256, &count) ||
      mbx_alloc(ctx, 8 * kSynthGroups, &out) ||
      mbx_alloc(ctx, 8 * kSynthGroups, &want)) {
    std::fprintf(stderr, "alloc: %s\n", mbx_last_error());
    return -1;
  }
  int bad = 0;
  std::vector<int64_t> h_want(kSynthGroups), h_got(kSynthGroups);
  vprg::instr ins3[MAXNINSTR];
  synth_instrs(ins3);
  for (int ci = 0; ci < 10; ++ci) {
    SynthCase c;
    if (!synth_case(ctx, n, kSkew[ci / 2], ci % 2 ? 32 : 16, ci, c)) {
      std::fprintf(stderr, "synthetic data: %s\n", mbx_last_error());
      synth_free(ctx, c);
      return -1;
    }
    for (int o = 1; o < 11; ++o) {
      recipe r;
      mbx_rowop op;
      synth_query(c, o, r, op);
      float t[4][4] = {};   // [recipe][fused, stage 1, 2, 3]; scan in t[3][0]
      for (int kind = 0; kind < 3; ++kind) {
        vprg p = kind == 0 ? r.perfect(ins3)
                           : kind == 1 ? r.many_or(ins3) : r.candchk(ins3);
        const mbx_vprg ap = p.to_abi();
        if (kind == 0) {
          t[3][0] = timed([&] {
            return mbx_prog_grpby(ctx, &ap, &op, kSynthGroups,
                                  static_cast<int64_t *>(want), MBX_PLAN_SCAN);
          });
          mbx_memcpy_d2h(ctx, h_want.data(), want, 8 * kSynthGroups);
        }
        mbx_plan plan;
        if (!static_plan(ctx, ap, plan)) return -1;
        t[kind][0] = timed([&] {
          return mbx_prog_grpby_plan(ctx, &ap, &op, kSynthGroups,
                                     static_cast<int64_t *>(out), &plan);
        });
        mbx_memcpy_d2h(ctx, h_got.data(), out, 8 * kSynthGroups);
        bad += std::memcmp(h_want.data(), h_got.data(), 8 * kSynthGroups) != 0;
        uint32_t *pu = p.is_nochk() ? nullptr : static_cast<uint32_t *>(unc);
        t[kind][1] = timed([&] {
          return mbx_prog_eval(ctx, &ap, static_cast<uint32_t *>(poss), pu);
        });
        t[kind][2] = timed([&] {
          return mbx_compact(ctx, static_cast<uint32_t *>(poss), pu, n,
                             static_cast<uint32_t *>(rows),
                             static_cast<uint32_t *>(count));
        });
        t[kind][3] = timed([&] {
          return mbx_rows_grpby(ctx, static_cast<uint32_t *>(rows),
                                static_cast<uint32_t *>(count), &ap, &op,
                                kSynthGroups, static_cast<int64_t *>(out), 0);
        });
        mbx_memcpy_d2h(ctx, h_got.data(), out, 8 * kSynthGroups);
        bad += std::memcmp(h_want.data(), h_got.data(), 8 * kSynthGroups) != 0;
        p.release();
      }
      std::printf("%.2f\t%u\t%.5f\t%.4f\t%.4f\t%.4f\t%.4f", c.s, c.bw,
                  kSynthFulSel[o], t[3][0], t[0][0], t[1][0], t[2][0]);
      for (int kind = 0; kind < 3; ++kind)
        std::printf("\t%.4f\t%.4f\t%.4f", t[kind][1], t[kind][2], t[kind][3]);
      std::printf("\t0.0\n");
      std::fflush(stdout);
    }
    synth_free(ctx, c);
  }
  for (void *q : {poss, unc, rows, count, out, want}) mbx_free(ctx, q);
  return bad;
}

double now_ms() {
  return std::chrono::duration<double, std::milli>(
             std::chrono::steady_clock::now().time_since_epoch())
      .count();
}

// One predicate as a WAH bitmap (device words, their count scan and skip
// table).
struct Wah {
  uint32_t *w = nullptr, *ex = nullptr, *skip = nullptr;
  size_t n = 0, nskip = 0;
  uint64_t longest1 = 0;
  mbx_wah arg() const { return mbx_wah{w, ex, skip, n, nskip}; }
};

// The predicate's groups (mbx_wah_select, the analogue of dbjoinFlatWah)
// compressed into a pool buffer, then scanned; `groups` is ng words of
// scratch.  False on a library error (message printed).
bool wah_of(mbx_ctx *ctx, const mbx_colref &src, uint32_t nrows, uint64_t lo,
            uint64_t hi, uint32_t *groups, Wah &out) {
  const size_t ng = mbx_wah_groups_of(nrows);
  size_t k = 0;
  uint32_t *tmp = nullptr;
  if (mbx_wah_select(ctx, &src, nrows, lo, hi, groups) != MBX_OK ||
      mbx_alloc(ctx, 4 * (ng ? ng : 1), (void **)&tmp) != MBX_OK ||
      mbx_wah_compress(ctx, groups, ng, tmp, &k) != MBX_OK) {
    std::fprintf(stderr, "wah: %s\n", mbx_last_error());
    return false;
  }
  out.n = k;
  out.nskip = mbx_wah_skip_len(ng);
  if (mbx_alloc(ctx, 4 * (k ? k : 1), (void **)&out.w) != MBX_OK ||
      mbx_alloc(ctx, 4 * (k + 1), (void **)&out.ex) != MBX_OK ||
      mbx_alloc(ctx, 4 * (out.nskip ? out.nskip : 1), (void **)&out.skip) !=
          MBX_OK ||
      (k && mbx_memcpy_d2d(ctx, out.w, tmp, 4 * k) != MBX_OK) ||
      mbx_wah_scan_skip(ctx, out.w, k, out.ex, out.skip, out.nskip) != MBX_OK ||
      mbx_wah_longest_1fill(ctx, out.w, k, &out.longest1) != MBX_OK) {
    std::fprintf(stderr, "wah: %s\n", mbx_last_error());
    return false;
  }
  mbx_sync(ctx);
  mbx_free(ctx, tmp);
  return true;
}

void wah_free(mbx_ctx *ctx, Wah &w) {
  if (w.w) mbx_free(ctx, w.w);
  if (w.ex) mbx_free(ctx, w.ex);
  if (w.skip) mbx_free(ctx, w.skip);
  w = Wah();
}

}  // namespace

int main(int argc, char **argv) {
  if (argc < 2) {
    std::fprintf(stderr, "usage: %s <ssb_dir> [reps] [synth_rows]\n", argv[0]);
    return 2;
  }
  const std::string dir = argv[1];
  const int reps = argc > 2 ? std::atoi(argv[2]) : 20;
  const uint32_t synth_rows =
      argc > 3 ? (uint32_t)std::strtoul(argv[3], nullptr, 10) : 0;
  FILE *sums = nullptr;
  if (argc > 4 && !(sums = std::fopen(argv[4], "w"))) {
    std::fprintf(stderr, "cannot write %s\n", argv[4]);
    return 2;
  }
  auto write_sums = [&](const char *name, const char *bt, const char *plan,
                        const int64_t *v, uint32_t g) {
    if (!sums) return;
    std::fprintf(sums, "%s\t%s\t%s", name, bt, plan);
    for (uint32_t i = 0; i < g; ++i) std::fprintf(sums, "\t%lld", (long long)v[i]);
    std::fprintf(sums, "\n");
  };
  mbx_ctx *ctx = context();
  if (!ctx) {
    std::fprintf(stderr, "no device: %s\n", mbx_last_error());
    return 2;
  }
  Tables t;
  bool ok = true;
  for (int s = 0; s < NSRC; ++s) ok = ok && load(ctx, dir, kSrcName[s], t.c[s]);
  ok = ok && load(ctx, dir, "loCustKey", t.fk_cust) &&
       load(ctx, dir, "loPartKey", t.fk_part) &&
       load(ctx, dir, "loExtendedPrice", t.extprice) &&
       load(ctx, dir, "loRevenue", t.revenue) &&
       load(ctx, dir, "loSupplyCost", t.supplycost);
  if (!ok) return 2;
  t.nrows = (uint32_t)t.c[DATE].n;
  mbx_timer *tm = nullptr;
  mbx_timer_create(ctx, &tm);
  auto timed = [&](auto &&fn) {
    float ms = 0, sum = 0;
    for (int r = 0; r < reps; ++r) {
      mbx_timer_start(ctx, tm);
      if (fn() != MBX_OK) {
        std::fprintf(stderr, "launch: %s\n", mbx_last_error());
        std::exit(2);
      }
      mbx_timer_stop(ctx, tm, &ms);
      sum += ms;
    }
    return sum / reps;
  };

  // ------------------------------------------ block 0: the WAH baseline
  // ssb/wah-demo.cu:24-292: each flight's predicates (its exact ranges, as
  // the Perfect layout below) as WAH bitmaps, built and scanned untimed;
  // timed per rep, index only: the predicates ANDed left to right
  // (mbx_wah_and_no1_skip; the down side, which must hold no long 1-fill,
  // is the operand with more words, and a flight where neither operand
  // can be down prints nan), the result scanned and decompressed (every
  // scan with its skip table).
  // RTScan (OptiX ray tracing) has no counterpart here: nan.
  const size_t ngroups = mbx_wah_groups_of(t.nrows);
  const size_t nskip = mbx_wah_skip_len(ngroups);
  uint32_t *wgroups = nullptr, *wres[2] = {nullptr, nullptr}, *wex = nullptr,
           *wsk = nullptr;
  if (mbx_alloc(ctx, 4 * ngroups, (void **)&wgroups) != MBX_OK ||
      mbx_alloc(ctx, 4 * ngroups, (void **)&wres[0]) != MBX_OK ||
      mbx_alloc(ctx, 4 * ngroups, (void **)&wres[1]) != MBX_OK ||
      mbx_alloc(ctx, 4 * (ngroups + 1), (void **)&wex) != MBX_OK ||
      mbx_alloc(ctx, 4 * (nskip ? nskip : 1), (void **)&wsk) != MBX_OK) {
    std::fprintf(stderr, "alloc: %s\n", mbx_last_error());
    return 2;
  }
  std::vector<float> wah_ms;
  std::vector<std::vector<uint32_t>> wah_res;
  for (const Flight &f : kFlights) {
    mbx_rowop op;
    std::vector<Pred> pr = flight_op(f, t, op);
    if (f.kind == 1 && pr[2].lo == 0) pr[2].lo = 1;   // quantity >= 1
    if (f.kind == 4 && f.a[6] <= D920101 && f.a[7] >= D990101) pr.pop_back();
    std::vector<Wah> w(pr.size());
    for (size_t i = 0; i < pr.size(); ++i)
      if (!wah_of(ctx, t.ref(pr[i].src), t.nrows, pr[i].lo, pr[i].hi, wgroups,
                  w[i]))
        return 2;
    // operand order: acc = w[0]; the result of each AND keeps its down
    // operand's words, so down is the operand with more words (its
    // literals then find the up words covering them staged in LDS) unless
    // it holds a long 1-fill
    bool ok = true;
    std::vector<std::pair<int, bool>> steps;   // (operand, operand is down)
    uint64_t acc1 = w[0].longest1;
    size_t acc_n = w[0].n;
    for (size_t i = 1; i < w.size() && ok; ++i) {
      const bool next_ok = w[i].longest1 <= 1, acc_ok = acc1 <= 1;
      if (!next_ok && !acc_ok) {
        ok = false;
        break;
      }
      const bool down = next_ok && (!acc_ok || w[i].n >= acc_n);
      steps.push_back({(int)i, down});
      if (down) acc_n = w[i].n;
      acc1 = 1;   // a result keeps down's words: no long 1-fill
    }
    float ms = NAN;
    if (ok) {
      ms = timed([&] {
        mbx_wah cur = w[0].arg();
        int buf = 0;
        for (const auto &st : steps) {
          const mbx_wah o = w[st.first].arg();
          const mbx_wah &up = st.second ? cur : o;
          const mbx_wah dn = st.second ? o : cur;
          int rc = mbx_wah_and_no1_skip(ctx, &up, &dn, wres[buf]);
          if (rc != MBX_OK) return rc;
          // the result keeps down's word structure: its scan and table
          cur = dn;
          cur.words = wres[buf];
          buf ^= 1;
        }
        int rc = mbx_wah_scan_skip(ctx, cur.words, cur.n, wex, wsk, nskip);
        if (rc != MBX_OK) return rc;
        const mbx_wah res{cur.words, wex, wsk, cur.n, nskip};
        return mbx_wah_decompress_skip(ctx, &res, wgroups, ngroups);
      });
    }
    wah_ms.push_back(ms);
    // the result's groups, checked against the Perfect program's bitmap in
    // block 2
    wah_res.emplace_back(ok ? ngroups : 0);
    if (ok) mbx_memcpy_d2h(ctx, wah_res.back().data(), wgroups, 4 * ngroups);
    for (Wah &x : w) wah_free(ctx, x);
  }
  std::printf("Case\tWAH\tRTScan\n");
  for (size_t i = 0; i < wah_ms.size(); ++i)
    std::printf("%s\t%.4f\tnan\n", kFlights[i].name, wah_ms[i]);

  // ---------------------------------------------- block 1: index creation
  // My: create_bin of the column's bins; WAH: each bin's groups
  // (mbx_wah_select) compressed, one bin at a time, and the compressed size
  // (ssb/bmpidx-create.cu:56-116)
  std::vector<uint *> bins[2][NSRC];
  std::printf("\nColumn\tMy(ms)\tWAH(ms)\tMy(MB)\tWAH(MB)\n");
  const char *kSet[2] = {"Sparse", "Dense"};
  {
    // one untimed create_bin per source first: the first launch of each
    // bitmap-creation kernel also loads its code, which is not creation
    for (int s = 0; s < NSRC; ++s) {
      const std::vector<size_t> &b = kBounds[0][s];
      std::vector<uint64_t> lo(b.begin(), b.end() - 1), hi(b.begin() + 1, b.end());
      std::vector<uint *> tmp(b.size() - 1, nullptr);
      const mbx_colref r = t.ref(s);
      if (!create_bin(r.fk, r.data, 8 * r.elem_bytes, t.nrows, r.dimsz, lo.data(),
                      hi.data(), tmp.size(), tmp.data())) {
        std::fprintf(stderr, "create_bin %s: %s\n", kSrcName[s], mbx_last_error());
        return 2;
      }
      mbx_sync(ctx);
      mbx_free(ctx, tmp[0]);
    }
  }
  for (int d = 0; d < 2; ++d) {
    double tot_ms = 0, tot_mb = 0, tot_wms = 0, tot_wmb = 0;
    for (int s = 0; s < NSRC; ++s) {
      const std::vector<size_t> &b = kBounds[d][s];
      const size_t nb = b.size() - 1;
      std::vector<uint64_t> lo(b.begin(), b.end() - 1), hi(b.begin() + 1, b.end());
      bins[d][s].assign(nb, nullptr);
      const mbx_colref r = t.ref(s);
      mbx_sync(ctx);
      const double t0 = now_ms();   // like the reference, allocation included
      const size_t bytes =
          create_bin(r.fk, r.data, 8 * r.elem_bytes, t.nrows, r.dimsz,
                     lo.data(), hi.data(), nb, bins[d][s].data());
      mbx_sync(ctx);
      const double ms = now_ms() - t0;
      if (!bytes) {
        std::fprintf(stderr, "create_bin %s: %s\n", kSrcName[s], mbx_last_error());
        return 2;
      }
      const double mb = (double)bytes * nb / (1 << 20);
      const double w0 = now_ms();
      size_t wah_words = 0;
      for (size_t j = 0; j < nb; ++j) {
        size_t k = 0;
        if (mbx_wah_select(ctx, &r, t.nrows, lo[j], hi[j], wgroups) != MBX_OK ||
            mbx_wah_compress(ctx, wgroups, ngroups, wres[0], &k) != MBX_OK) {
          std::fprintf(stderr, "wah %s: %s\n", kSrcName[s], mbx_last_error());
          return 2;
        }
        wah_words += k;
      }
      mbx_sync(ctx);
      const double wms = now_ms() - w0, wmb = 4.0 * wah_words / (1 << 20);
      tot_ms += ms;
      tot_mb += mb;
      tot_wms += wms;
      tot_wmb += wmb;
      std::printf("%s_%s\t%.2f\t%.2f\t%.2f\t%.2f\n", kSrcName[s], kSet[d], ms,
                  wms, mb, wmb);
    }
    std::printf("Total_%s\t%.2f\t%.2f\t%.2f\t%.2f\n", kSet[d], tot_ms, tot_wms,
                tot_mb, tot_wmb);
  }

  // ------------------------------------------------------ block 2: SSB
  std::printf("\nCase\tBinType\tJoin\tOurs\tBasefus\tNfuStg1\tPrgStg1\tStg2\tStg3\tDedicat");
  const size_t nw = mbx_bitmap_stride_words(t.nrows);
  uint32_t *poss = nullptr, *unc = nullptr, *rows = nullptr, *count = nullptr;
  int64_t *out = nullptr, *want = nullptr;
  if (mbx_alloc(ctx, nw * 4, (void **)&poss) || mbx_alloc(ctx, nw * 4, (void **)&unc) ||
      mbx_alloc(ctx, (size_t)t.nrows * 4, (void **)&rows) ||
      mbx_alloc(ctx, 256, (void **)&count) || mbx_alloc(ctx, 8 * 1024, (void **)&out) ||
      mbx_alloc(ctx, 8 * 1024, (void **)&want)) {
    std::fprintf(stderr, "alloc: %s\n", mbx_last_error());
    return 2;
  }
  int bad = 0;
  std::vector<int64_t> h_want(1024), h_got(1024);
  // Basefus: the base_fuse ablation (method 1, the program's registers in
  // device memory); NfuStg1: the first materialise stage with its registers
  // in device memory (the reference's _stg1 with idxbuf).  Both under the
  // static plan's order; Basefus's sums must equal the scan's too.
  auto ablations = [&](const mbx_vprg &ap0, const mbx_rowop &op0, uint32_t g,
                       const mbx_plan &plan, float &basefus, float &nfu1) {
    mbx_vprg ap;
    if (mbx_plan_apply(&ap0, &plan, &ap) != MBX_OK) {
      std::fprintf(stderr, "plan: %s\n", mbx_last_error());
      return false;
    }
    basefus = timed([&] {
      return mbx_prog_grpby(ctx, &ap, &op0, g, out,
                            MBX_PLAN_FUSED | MBX_PLAN_BASEFUSE);
    });
    mbx_memcpy_d2h(ctx, h_got.data(), out, 8 * g);
    bad += std::memcmp(h_want.data(), h_got.data(), 8 * g) != 0;
    bool chk = false;
    for (int c = 0; c < MBX_MAXCOLS; ++c)
      chk |= ap.cols[c].leftmost || ap.cols[c].rightmost;
    nfu1 = timed([&] {
      return mbx_prog_eval_ex(ctx, &ap, poss, chk ? unc : nullptr,
                              MBX_PLAN_BASEFUSE);
    });
    return true;
  };
  const char *kPlan[3] = {"Sparse", "Medium", "Dense"};
  for (const Flight &f : kFlights) {
    mbx_rowop op;
    flight_op(f, t, op);
    const Fixed x = fixed_layout(f);
    // the row op re-checks an uncertain row by running the program over the
    // exact predicates of its columns, so they follow the program's columns
    op.npreds = (uint32_t)x.cols.size();
    for (size_t i = 0; i < x.cols.size(); ++i) {
      op.preds[i].src = t.ref(x.cols[i].src);
      op.preds[i].lo = x.cols[i].lo;
      op.preds[i].hi = x.cols[i].hi;
    }
    for (int level = 0; level < 3; ++level) {
      const unsigned dense = level == 0 ? 0u : level == 1 ? x.medium : x.dense;
      vprg p;
      p.factsz = t.nrows;
      for (size_t i = 0; i < x.cols.size(); ++i) {
        const FixedCol &c = x.cols[i];
        const std::vector<size_t> &sb = kBounds[0][c.src], &db = kBounds[1][c.src];
        p.cols[i] = (dense >> i) & 1
                        ? col::from_bin2(sb.size() - 1, bins[0][c.src].data(),
                                         sb.data(), db.size() - 1,
                                         bins[1][c.src].data(), db.data(), c.lo,
                                         c.hi)
                        : col::from_bin(sb.size() - 1, bins[0][c.src].data(),
                                        sb.data(), c.lo, c.hi);
        if (p.cols[i].empty()) {
          std::fprintf(stderr, "%s %s: column %zu not coverable\n", f.name,
                       kPlan[level], i);
          return 2;
        }
      }
      for (size_t i = 0; i < x.ins.size(); ++i) p.instrs[i] = x.ins[i];
      const mbx_vprg ap = p.to_abi();
      const uint32_t g = f.ngrp;
      // the scan is timed once per flight, as DOWORK(..., join = true) does
      // on the Sparse program only (ssb/lambdas.cuh:157,168)
      float join = 0;
      if (level == 0)
        join = timed([&] {
          return mbx_prog_grpby(ctx, &ap, &op, g, want, MBX_PLAN_SCAN);
        });
      mbx_plan plan;
      if (!static_plan(ctx, ap, plan)) return 2;
      const float ours = timed([&] {
        return mbx_prog_grpby_plan(ctx, &ap, &op, g, out, &plan);
      });
      mbx_memcpy_d2h(ctx, h_want.data(), want, 8 * g);
      mbx_memcpy_d2h(ctx, h_got.data(), out, 8 * g);
      bad += std::memcmp(h_want.data(), h_got.data(), 8 * g) != 0;
      write_sums(f.name, kPlan[level], "join", h_want.data(), g);
      write_sums(f.name, kPlan[level], "ours", h_got.data(), g);
      float basefus = 0, nfu1 = 0;
      if (!ablations(ap, op, g, plan, basefus, nfu1)) return 2;
      // Dedicat: the kernel built for the planned program's shape (the
      // reference's hand-written c::s** functions, ssb/fixedlayout.cu:
      // 15-32); 0 where the library holds none, as DEDICATE reads 0 for a
      // missing function
      float ded = 0;
      {
        mbx_vprg dp;
        if (mbx_plan_apply(&ap, &plan, &dp) != MBX_OK) {
          std::fprintf(stderr, "plan: %s\n", mbx_last_error());
          return 2;
        }
        const uint32_t dfl = MBX_PLAN_FUSED | MBX_PLAN_DEDICATED;
        const int st = mbx_prog_grpby(ctx, &dp, &op, g, out, dfl);
        if (st == MBX_OK) {
          ded = timed([&] { return mbx_prog_grpby(ctx, &dp, &op, g, out, dfl); });
          mbx_memcpy_d2h(ctx, h_got.data(), out, 8 * g);
          bad += std::memcmp(h_want.data(), h_got.data(), 8 * g) != 0;
          write_sums(f.name, kPlan[level], "ded", h_got.data(), g);
        } else if (st == MBX_ERR_UNSUPPORTED) {
          std::fprintf(stderr, "%s %s: %s\n", f.name, kPlan[level],
                       mbx_last_error());
        } else {
          std::fprintf(stderr, "dedicated: %s\n", mbx_last_error());
          return 2;
        }
      }
      const float s1 = timed([&] {
        return mbx_prog_eval(ctx, &ap, poss, p.is_nochk() ? nullptr : unc);
      });
      const float s2 = timed([&] {
        return mbx_compact(ctx, poss, p.is_nochk() ? nullptr : unc, t.nrows,
                           rows, count);
      });
      const float s3 = timed([&] {
        return mbx_rows_grpby(ctx, rows, count, &ap, &op, g, out, 0);
      });
      mbx_memcpy_d2h(ctx, h_got.data(), out, 8 * g);
      bad += std::memcmp(h_want.data(), h_got.data(), 8 * g) != 0;
      write_sums(f.name, kPlan[level], "stg", h_got.data(), g);
      if (level == 0)
        std::printf("\n%s\t%s\t%.4f", f.name, kPlan[level], join);
      else
        std::printf("\n%s\t%s\t0", f.name, kPlan[level]);
      std::printf("\t%.4f\t%.4f\t%.4f\t%.4f\t%.4f\t%.4f\t%.4f", ours, basefus,
                  nfu1, s1, s2, s3, ded);
    }
  }
  // the controlled layouts (ssb/controlled-demo.cu:7-129): each flight's
  // predicates get bins made for the query by the recipes -- Perfect (one
  // exact bitmap per column), AllAlig (many_or: three aligned bins) and
  // MisAlig (candchk: edge bins widened by step / 2 on columns 0 and 1) --
  // as one conjunction of all columns; flight 4 drops its date column when
  // the range spans every date (controlled-demo.cu:104-105)
  const char *kRecipe[3] = {"Perfect", "AllAlig", "MisAlig"};
  for (const Flight &f : kFlights) {
    mbx_rowop op;
    std::vector<Pred> pr = flight_op(f, t, op);
    if (f.kind == 1 && pr[2].lo == 0) pr[2].lo = 1;   // quantity >= 1
    if (f.kind == 4 && f.a[6] <= D920101 && f.a[7] >= D990101) pr.pop_back();
    recipe r;
    r.factsz = t.nrows;
    std::vector<vprg::instr> rins;
    for (size_t i = 0; i < pr.size(); ++i) {
      const mbx_colref c = t.ref(pr[i].src);
      r.nbit[i] = 8 * c.elem_bytes;
      r.min[i] = (uint)pr[i].lo;
      r.max[i] = (uint)pr[i].hi;
      r.dimsz[i] = recipe::nodim;   // keys are in range (ssb-dbgen-lite)
      r.fk[i] = c.fk;
      r.attr[i] = c.data;
      rins.push_back(ins(i == 0 ? vprg::ORM : vprg::ANDM, 0, (uint8_t)i));
      op.preds[i].src = c;
      op.preds[i].lo = pr[i].lo;
      op.preds[i].hi = pr[i].hi;
    }
    op.npreds = (uint32_t)pr.size();
    vprg::instr instrs[MAXNINSTR];
    for (size_t i = 0; i < rins.size(); ++i) instrs[i] = rins[i];
    const uint32_t g = f.ngrp;
    for (int kind = 0; kind < 3; ++kind) {
      vprg p = kind == 0 ? r.perfect(instrs)
               : kind == 1 ? r.many_or(instrs) : r.candchk(instrs);
      for (size_t i = 0; i < pr.size(); ++i)
        if (p.cols[i].empty()) {
          std::fprintf(stderr, "%s %s: recipe failed: %s\n", f.name,
                       kRecipe[kind], mbx_last_error());
          return 2;
        }
      const mbx_vprg ap = p.to_abi();
      if (mbx_prog_grpby(ctx, &ap, &op, g, want, MBX_PLAN_SCAN) != MBX_OK) {
        std::fprintf(stderr, "scan: %s\n", mbx_last_error());
        return 2;
      }
      mbx_plan plan;
      if (!static_plan(ctx, ap, plan)) return 2;
      const float ours = timed([&] {
        return mbx_prog_grpby_plan(ctx, &ap, &op, g, out, &plan);
      });
      mbx_memcpy_d2h(ctx, h_want.data(), want, 8 * g);
      mbx_memcpy_d2h(ctx, h_got.data(), out, 8 * g);
      bad += std::memcmp(h_want.data(), h_got.data(), 8 * g) != 0;
      write_sums(f.name, kRecipe[kind], "join", h_want.data(), g);
      write_sums(f.name, kRecipe[kind], "ours", h_got.data(), g);
      float basefus = 0, nfu1 = 0;
      if (!ablations(ap, op, g, plan, basefus, nfu1)) return 2;
      const float s1 = timed([&] {
        return mbx_prog_eval(ctx, &ap, poss, p.is_nochk() ? nullptr : unc);
      });
      if (kind == 0 && !wah_res[&f - kFlights].empty()) {
        // the WAH baseline's result (block 0) == this program's bitmap
        std::vector<uint32_t> gh(ngroups);
        if (mbx_wah_groups(ctx, poss, t.nrows, wgroups) != MBX_OK ||
            mbx_memcpy_d2h(ctx, gh.data(), wgroups, 4 * ngroups) != MBX_OK) {
          std::fprintf(stderr, "wah check: %s\n", mbx_last_error());
          return 2;
        }
        if (gh != wah_res[&f - kFlights]) {
          std::fprintf(stderr, "%s: WAH result differs from the bitmap\n", f.name);
          ++bad;
        }
      }
      const float s2 = timed([&] {
        return mbx_compact(ctx, poss, p.is_nochk() ? nullptr : unc, t.nrows,
                           rows, count);
      });
      const float s3 = timed([&] {
        return mbx_rows_grpby(ctx, rows, count, &ap, &op, g, out, 0);
      });
      mbx_memcpy_d2h(ctx, h_got.data(), out, 8 * g);
      bad += std::memcmp(h_want.data(), h_got.data(), 8 * g) != 0;
      write_sums(f.name, kRecipe[kind], "stg", h_got.data(), g);
      std::printf("\n%s\t%s\t0\t%.4f\t%.4f\t%.4f\t%.4f\t%.4f\t%.4f\t0",
                  f.name, kRecipe[kind], ours, basefus, nfu1, s1, s2, s3);
      mbx_sync(ctx);
      p.release();
    }
  }
  if (sums) std::fclose(sums);
  // blocks 3 (method 1 / method 2) and 4 (the synthetic suite) when
  // synth_rows > 0, else their headers alone
  std::printf("\n\nSeletiv\tMethod1\tMethod2\n");
  std::fflush(stdout);
  if (synth_rows) {
    const int mb = synth_methods(ctx, synth_rows, timed);
    if (mb < 0) return 2;
    bad += mb;
  }
  std::printf("\nSkew\tBitW\tSeletiv\tJoin\tPerfect\tManyOrs\tCandChk\t"
              "PftStg1\tPftStg2\tPftStg3\tOrsStg1\tOrsStg2\tOrsStg3\tChkStg1\t"
              "ChkStg2\tChkStg3\tWAH\n");
  std::fflush(stdout);
  if (synth_rows) {
    const int sb = synth_suite(ctx, synth_rows, timed);
    if (sb < 0) return 2;
    bad += sb;
  }
  for (int d = 0; d < 2; ++d)
    for (int s = 0; s < NSRC; ++s)
      if (!bins[d][s].empty()) free_bitmap(bins[d][s][0]);
  mbx_timer_destroy(tm);
  if (bad) std::fprintf(stderr, "%d plan results differ from the scan\n", bad);
  return bad ? 1 : 0;
}
