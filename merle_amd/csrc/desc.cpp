// desc.cpp -- validation and conversion of the ABI query descriptors
// (mbx_vprg, mbx_rowop) into their kernel-argument forms (DevPrg, DevOp).
// Host-only code.
#include <cstring>
#include <string>

#include "mbx_internal.hpp"

namespace mbx {

int check_src(const mbx_colref &c, const char *what) {
  if (c.data == nullptr)
    return fail(MBX_ERR_INVALID, std::string(what) + ": null column");
  if (c.elem_bytes != 1 && c.elem_bytes != 2 && c.elem_bytes != 4 &&
      c.elem_bytes != 8)
    return fail(MBX_ERR_INVALID,
                std::string(what) + ": elem_bytes must be 1, 2, 4 or 8");
  if ((reinterpret_cast<uintptr_t>(c.data) % c.elem_bytes) != 0 ||
      (reinterpret_cast<uintptr_t>(c.fk) % 4) != 0)
    return fail(MBX_ERR_ALIGN, std::string(what) + ": misaligned column");
  return MBX_OK;
}

static DevSrc to_dev(const mbx_colref &c) {
  DevSrc d;
  d.data = c.data;
  d.fk = c.fk;
  d.dimsz = c.dimsz;
  d.eb = c.elem_bytes;
  return d;
}

static int check_bmp(const uint32_t *p) {
  if (reinterpret_cast<uintptr_t>(p) % 8 != 0)
    return fail(MBX_ERR_ALIGN, "bitmap pointer is not 8-byte aligned");
  return MBX_OK;
}

// vprg validation follows the reference's limits (primitive.cuh:121,164-198):
// registers r0..r2, column operands 0..3, at most 6 instructions.
// Symbolic run of the program.  Each register holds 0 (empty), a
// conjunction of clauses, or "other".  A clause is an OR of plain columns
// (a column mask) or NOT of one column.  ORM into an empty register opens a
// clause and ORM / OR into a register that is one plain clause extends it;
// ANDM and AND add clauses; NOT turns a one-column clause into its negation.
// The program is a CNF of column literals when r0 ends as a non-empty
// conjunction whose clauses use every memory column exactly once (each with
// bins).  Then r0 = AND over clauses of (OR of the clause's bins),
// complemented for negated clauses, word for word, in the two-valued form
// and, with edge bins, in the three-valued one.
namespace {
struct Cnf {
  enum { kZero = 0, kConj = 1, kOther = 2 };
  int kind = kZero;
  uint32_t n = 0;
  uint8_t mask[8] = {};
  bool neg[8] = {};
  bool add(uint8_t m, bool ng) {
    for (uint32_t i = 0; i < n; ++i)
      if (mask[i] == m && neg[i] == ng) return true;   // x & x = x
    if (n == 8) return false;
    mask[n] = m;
    neg[n] = ng;
    ++n;
    return true;
  }
  bool one_plain() const { return kind == kConj && n == 1 && !neg[0]; }
};
}  // namespace

static bool cnf_program(const DevPrg &P, Cnf &r0) {
  Cnf r[3];
  for (uint32_t i = 0; i < P.nins; ++i) {
    const uint32_t op = P.ins[i][0], d = P.ins[i][1], s = P.ins[i][2];
    Cnf &rd = r[d];
    switch (op) {
    case MBX_ORM:
      if (rd.kind == Cnf::kZero) {
        rd = Cnf();
        rd.kind = Cnf::kConj;
        rd.add((uint8_t)(1u << s), false);
      } else if (rd.one_plain()) {
        rd.mask[0] |= (uint8_t)(1u << s);
      } else {
        rd.kind = Cnf::kOther;
      }
      break;
    case MBX_ANDM:   // 0 & x stays 0
      if (rd.kind == Cnf::kConj && !rd.add((uint8_t)(1u << s), false))
        rd.kind = Cnf::kOther;
      break;
    case MBX_OR: {
      const Cnf rs = r[s];
      if (rs.kind == Cnf::kZero) break;
      if (rd.kind == Cnf::kZero) rd = rs;
      else if (rd.one_plain() && rs.one_plain()) rd.mask[0] |= rs.mask[0];
      else rd.kind = Cnf::kOther;
      break;
    }
    case MBX_AND: {
      const Cnf rs = r[s];
      if (rd.kind == Cnf::kZero || rs.kind == Cnf::kZero) {
        rd = Cnf();
      } else if (rd.kind == Cnf::kConj && rs.kind == Cnf::kConj) {
        for (uint32_t k = 0; k < rs.n && rd.kind == Cnf::kConj; ++k)
          if (!rd.add(rs.mask[k], rs.neg[k])) rd.kind = Cnf::kOther;
      } else {
        rd.kind = Cnf::kOther;
      }
      break;
    }
    default: {   // NOT of one literal
      const Cnf rs = r[s];
      if (rs.kind == Cnf::kConj && rs.n == 1 &&
          (rs.mask[0] & (rs.mask[0] - 1)) == 0) {
        rd = rs;
        rd.neg[0] = !rs.neg[0];
      } else {
        rd.kind = Cnf::kOther;
      }
      break;
    }
    }
  }
  r0 = r[0];
  if (r0.kind != Cnf::kConj || r0.n == 0) return false;
  // the flat list holds exactly the columns of memory instructions; each
  // must sit in exactly one clause of r0 (none read only elsewhere)
  uint32_t seen = 0;
  for (uint32_t k = 0; k < r0.n; ++k) {
    if (seen & r0.mask[k]) return false;
    seen |= r0.mask[k];
  }
  return seen == P.memcols;
}

int make_dev_prg(const mbx_vprg *p, DevPrg *out) {
  if (!p || !out) return fail(MBX_ERR_INVALID, "vprg: null");
  if (p->factsz == 0) return fail(MBX_ERR_INVALID, "vprg: factsz == 0");
  if (p->factsz > 0x7fffffffu)
    return fail(MBX_ERR_CAPACITY, "vprg: factsz must be < 2^31");
  std::memset(out, 0, sizeof(*out));
  bool chk = false;
  for (int c = 0; c < MBX_MAXCOLS; ++c) {
    const mbx_col &col = p->cols[c];
    DevCol &d = out->cols[c];
    uint32_t nm = 0;
    while (nm < MBX_MAXBIN_PERCOL && col.middle[nm]) {
      int st = check_bmp(col.middle[nm]);
      if (st) return st;
      d.m[nm] = reinterpret_cast<const uint64_t *>(col.middle[nm]);
      ++nm;
    }
    d.nm = nm;
    if (col.leftmost) {
      int st = check_bmp(col.leftmost);
      if (st) return st;
    }
    if (col.rightmost) {
      int st = check_bmp(col.rightmost);
      if (st) return st;
    }
    d.l = reinterpret_cast<const uint64_t *>(col.leftmost);
    d.r = reinterpret_cast<const uint64_t *>(col.rightmost);
    d.has_edge = (col.leftmost || col.rightmost) ? 1u : 0u;
    chk = chk || d.has_edge;
  }
  uint32_t n = 0;
  for (; n < MBX_MAXNINSTR; ++n) {
    const mbx_instr &in = p->instrs[n];
    if (in.opcode == MBX_END) break;
    if (in.opcode > MBX_ORM)
      return fail(MBX_ERR_INVALID, "vprg: bad opcode " +
                                       std::to_string(in.opcode));
    if (in.dstreg > 2) return fail(MBX_ERR_INVALID, "vprg: dstreg > 2");
    const bool mem = in.opcode == MBX_ANDM || in.opcode == MBX_ORM;
    if (mem ? in.src1 >= MBX_MAXCOLS : in.src1 > 2)
      return fail(MBX_ERR_INVALID, "vprg: src1 out of range");
    out->ins[n][0] = in.opcode;
    out->ins[n][1] = in.dstreg;
    out->ins[n][2] = in.src1;
    out->ins[n][3] = in.src2;
    if (mem) out->memcols |= 1u << in.src1;
  }
  out->nins = n;
  // Truth table of the program over the exact column predicates, with the
  // row re-check's semantics (registers start false; ANDM / ORM read
  // column src1, AND / OR / NOT registers): an uncertain row is then one
  // lookup instead of an interpretation of the instruction list.
  uint32_t truth = 0;
  for (uint32_t m = 0; m < (1u << MBX_MAXCOLS); ++m) {
    bool r[3] = {false, false, false};
    for (uint32_t i = 0; i < n; ++i) {
      const uint8_t op = out->ins[i][0], d = out->ins[i][1], s = out->ins[i][2];
      bool v;
      if (op == MBX_AND) v = r[d] && r[s];
      else if (op == MBX_OR) v = r[d] || r[s];
      else if (op == MBX_NOT) v = !r[s];
      else {
        const bool es = (m >> s) & 1u;
        v = op == MBX_ANDM ? (r[d] && es) : (r[d] || es);
      }
      r[d] = v;
    }
    if (r[0]) truth |= 1u << m;
  }
  out->truth = truth;
  // The flat operand list: column blocks (middle bins, then edge bins),
  // clause by clause when the program is a CNF of column literals, else in
  // column order; flend marks the last entry of each clause (each column
  // when not a CNF).
  Cnf cnf;
  bool conj = cnf_program(*out, cnf);
  for (uint32_t c = 0; c < MBX_MAXCOLS; ++c)   // an empty column: generic path
    if (((out->memcols >> c) & 1u) && out->cols[c].nm == 0 &&
        !out->cols[c].has_edge)
      conj = false;
  auto add_col = [&](uint32_t c) {
    const DevCol &d = out->cols[c];
    out->beg[c] = (uint8_t)out->nfl;
    auto add = [&](const uint64_t *b, bool edge) {
      out->fl[out->nfl] = b;
      out->flcol[out->nfl] = (uint8_t)c;
      out->fledge[out->nfl] = edge ? 1 : 0;
      ++out->nfl;
    };
    for (uint32_t k = 0; k < d.nm; ++k) add(d.m[k], false);
    if (d.l) add(d.l, true);
    if (d.r) add(d.r, true);   // an empty column reads as all-zero
    out->end[c] = (uint8_t)out->nfl;
  };
  for (uint32_t c = 0; c < MBX_MAXCOLS; ++c) out->beg[c] = out->end[c] = 0;
  if (conj) {
    for (uint32_t k = 0; k < cnf.n; ++k) {
      for (uint32_t c = 0; c < MBX_MAXCOLS; ++c)
        if ((cnf.mask[k] >> c) & 1u) add_col(c);
      out->flend[out->nfl - 1] = 1;
      out->flneg[out->nfl - 1] = cnf.neg[k] ? 1 : 0;
    }
  } else {
    for (uint32_t c = 0; c < MBX_MAXCOLS; ++c) {
      if ((out->memcols >> c) & 1u) add_col(c);
      else out->beg[c] = out->end[c] = (uint8_t)out->nfl;
      if (out->end[c] > out->beg[c]) out->flend[out->end[c] - 1] = 1;
    }
  }
  // Columns the re-check of a candidate row need not load: those no memory
  // instruction reads (the program does not depend on them), and in a
  // conjunction the columns of clauses without edge bins, which every
  // candidate row satisfies -- a plain clause is its exact OR, true (its
  // columns are set true), a negated one its column's exact predicate,
  // false.  Each column sits in one clause, so the others are unaffected.
  uint32_t known = ~out->memcols & ((1u << MBX_MAXCOLS) - 1u), known_val = 0;
  if (conj)
    for (uint32_t k = 0; k < cnf.n; ++k) {
      bool edge = false;
      for (uint32_t c = 0; c < MBX_MAXCOLS; ++c)
        if ((cnf.mask[k] >> c) & 1u) edge = edge || out->cols[c].has_edge;
      if (edge) continue;
      known |= cnf.mask[k];
      if (!cnf.neg[k]) known_val |= cnf.mask[k];
    }
  out->known = known;
  out->known_val = known_val;
  out->nrows = p->factsz;
  out->nwords = (uint32_t)ceil_div(p->factsz, 64);
  out->chk = chk ? 1u : 0u;
  out->conj = conj ? 1u : 0u;
  return MBX_OK;
}

// Clause k of a conjunction, in flat-list order, as nm | left << 3 |
// right << 4 | negated << 5 at bits 6k..6k+5 (the dedicated kernels'
// template argument); 0 unless every clause holds exactly one column.
uint32_t conj_shape_code(const DevPrg &P) {
  if (!P.conj || P.nfl == 0) return 0;
  uint32_t code = 0, k = 0, e = 0;
  while (e < P.nfl) {
    if (k == MBX_MAXCOLS) return 0;
    const uint32_t c = P.flcol[e];
    const DevCol &d = P.cols[c];
    const uint32_t n = d.nm + (d.l ? 1u : 0u) + (d.r ? 1u : 0u);
    if (n == 0 || e + n > P.nfl) return 0;
    for (uint32_t j = e; j < e + n; ++j)
      if (P.flcol[j] != c || (P.flend[j] != 0) != (j == e + n - 1)) return 0;
    const uint32_t cl = d.nm | (d.l ? 8u : 0u) | (d.r ? 16u : 0u) |
                        (P.flneg[e + n - 1] ? 32u : 0u);
    code |= cl << (6 * k);
    ++k;
    e += n;
  }
  return code;
}

void key_div_magic(uint32_t d, uint64_t *m, uint8_t *sh1, uint8_t *sh2) {
  uint32_t l = 0;
  while ((1ull << l) < d) ++l;
  const unsigned __int128 num = (unsigned __int128)((1ull << l) - d) << 64;
  *m = (uint64_t)(num / d) + 1;
  *sh1 = (uint8_t)(l < 1 ? l : 1);
  *sh2 = (uint8_t)(l > 1 ? l - 1 : 0);
}

int make_dev_op(const mbx_rowop *op, uint32_t npreds_hint, DevOp *out) {
  if (!op || !out) return fail(MBX_ERR_INVALID, "rowop: null");
  std::memset(out, 0, sizeof(*out));
  (void)npreds_hint;
  if (op->npreds > MBX_MAXCOLS || op->nkeys > MBX_MAXKEYS)
    return fail(MBX_ERR_INVALID, "rowop: too many predicates or keys");
  out->npreds = op->npreds;
  for (uint32_t i = 0; i < op->npreds; ++i) {
    int st = check_src(op->preds[i].src, "rowop predicate");
    if (st) return st;
    out->pred[i] = to_dev(op->preds[i].src);
    out->lo[i] = op->preds[i].lo;
    out->hi[i] = op->preds[i].hi;
  }
  out->nkeys = op->nkeys;
  for (uint32_t i = 0; i < op->nkeys; ++i) {
    const mbx_keypart &k = op->keys[i];
    int st = check_src(k.src, "rowop key");
    if (st) return st;
    if (k.kind > MBX_KEY_BUCKET)
      return fail(MBX_ERR_INVALID, "rowop key: bad kind");
    if (k.kind == MBX_KEY_DIV && k.div == 0)
      return fail(MBX_ERR_INVALID, "rowop key: div == 0");
    if (k.nbounds > MBX_MAXBOUNDS)
      return fail(MBX_ERR_INVALID, "rowop key: too many bounds");
    DevKey &d = out->keys[i];
    d.src = to_dev(k.src);
    d.kind = k.kind;
    d.div = k.div;
    if (k.kind == MBX_KEY_DIV) key_div_magic(k.div, &d.magic, &d.sh1, &d.sh2);
    d.base = k.base;
    d.stride = k.stride;
    d.nb = k.nbounds;
    for (uint32_t j = 0; j < MBX_MAXBOUNDS; ++j) d.bounds[j] = k.bounds[j];
  }
  if (op->measure > MBX_MEAS_SUB)
    return fail(MBX_ERR_INVALID, "rowop: bad measure");
  out->measure = op->measure;
  if (op->measure != MBX_MEAS_COUNT) {
    int st = check_src(op->ma, "rowop measure a");
    if (st) return st;
    out->ma = to_dev(op->ma);
  }
  if (op->measure == MBX_MEAS_MUL || op->measure == MBX_MEAS_SUB) {
    int st = check_src(op->mb, "rowop measure b");
    if (st) return st;
    out->mb = to_dev(op->mb);
  }
  return MBX_OK;
}

}  // namespace mbx
