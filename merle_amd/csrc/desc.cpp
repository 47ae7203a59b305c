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
// Symbolic run of the program: each register holds 0 (empty), a
// conjunction of column literals (bit c: column c, bit 4+c: NOT column c),
// or "other".  True (and flneg set per column) when r0 ends as a non-empty
// conjunction in which no column appears both plain and negated and every
// column has bins -- then r0 = AND of the literals, word for word, in the
// two-valued form and, with edge bins, in the three-valued one.
static bool conj_literals(DevPrg &P) {
  enum { kZero = 0, kConj = 1, kOther = 2 };
  int kind[3] = {kZero, kZero, kZero};
  uint32_t lit[3] = {0, 0, 0};
  for (uint32_t i = 0; i < P.nins; ++i) {
    const uint32_t op = P.ins[i][0], d = P.ins[i][1], s = P.ins[i][2];
    switch (op) {
    case MBX_ORM:
      if (kind[d] == kZero) { kind[d] = kConj; lit[d] = 1u << s; }
      else { kind[d] = kOther; }
      break;
    case MBX_ANDM:
      if (kind[d] == kConj) lit[d] |= 1u << s;
      else if (kind[d] != kZero) kind[d] = kOther;   // 0 & x stays 0
      break;
    case MBX_NOT:
      if (kind[s] == kConj && (lit[s] & (lit[s] - 1)) == 0) {
        kind[d] = kConj;   // NOT of one literal
        lit[d] = (lit[s] & 0xfu) ? lit[s] << 4 : lit[s] >> 4;
      } else {
        kind[d] = kOther;
      }
      break;
    case MBX_AND:
      if (kind[d] == kConj && kind[s] == kConj) lit[d] |= lit[s];
      else if (kind[d] == kZero || kind[s] == kZero) { kind[d] = kZero; lit[d] = 0; }
      else kind[d] = kOther;
      break;
    default:   // OR: a disjunction
      kind[d] = kOther;
      break;
    }
  }
  if (kind[0] != kConj) return false;
  const uint32_t pos = lit[0] & 0xfu, neg = lit[0] >> 4;
  if (pos & neg) return false;
  // the flat list holds exactly the columns of memory instructions; every
  // one of them must be a literal of r0 (none may be read only elsewhere)
  if ((pos | neg) != P.memcols) return false;
  for (uint32_t c = 0; c < MBX_MAXCOLS; ++c) {
    if (!((P.memcols >> c) & 1u)) continue;
    if (P.end[c] <= P.beg[c]) return false;   // empty column: generic path
    P.flneg[P.end[c] - 1] = (neg >> c) & 1u;
  }
  return true;
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
  for (uint32_t c = 0; c < MBX_MAXCOLS; ++c) {
    out->beg[c] = (uint8_t)out->nfl;
    if ((out->memcols >> c) & 1u) {
      const DevCol &d = out->cols[c];
      auto add = [&](const uint64_t *b, bool edge) {
        out->fl[out->nfl] = b;
        out->flcol[out->nfl] = (uint8_t)c;
        out->fledge[out->nfl] = edge ? 1 : 0;
        ++out->nfl;
      };
      for (uint32_t k = 0; k < d.nm; ++k) add(d.m[k], false);
      if (d.l) add(d.l, true);
      if (d.r) add(d.r, true);   // an empty column reads as all-zero
    }
    out->end[c] = (uint8_t)out->nfl;
    if (out->end[c] > out->beg[c]) out->flend[out->end[c] - 1] = 1;
  }
  out->nrows = p->factsz;
  out->nwords = (uint32_t)ceil_div(p->factsz, 64);
  out->chk = chk ? 1u : 0u;
  out->conj = conj_literals(*out) ? 1u : 0u;
  return MBX_OK;
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
