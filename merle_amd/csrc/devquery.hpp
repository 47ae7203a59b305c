// devquery.hpp -- device side of a query: program evaluation over 64-row
// words, the row operation, and the tile pipeline shared by the single-query
// and batched kernels.
#pragma once
#include "devfuncs.hpp"

namespace mbx {

// ------------------------------------------------------------ accounting
// Algorithmic-byte accounting (mbx_alg_bytes): when a kernel is instantiated
// with ACCT = true, every load of the plan marks its 128-byte line in a
// per-array bitset instead of being timed.  The product instantiations use
// ACCT = false and carry no trace of it.
struct AcctEntry {
  uint64_t base, nbytes;
  uint32_t off;       // first bitset word
  uint32_t pad;
};
constexpr int kAcctMax = 64;
struct Acct {
  AcctEntry e[kAcctMax];
  uint32_t n, pad;
  uint32_t *bits;
  unsigned long long *selected;
};

template <bool ACCT>
__device__ __forceinline__ void mark(const Acct *acct, const void *p) {
  if (!ACCT) return;
  const uint64_t a = reinterpret_cast<uint64_t>(p);
  for (uint32_t i = 0; i < acct->n; ++i) {
    const AcctEntry &e = acct->e[i];
    if (a >= e.base && a < e.base + e.nbytes) {
      const uint64_t line = (a - e.base) >> 7;
      atomicOr(acct->bits + e.off + (line >> 5), 1u << (line & 31));
      return;
    }
  }
}

template <bool ACCT>
__device__ __forceinline__ uint64_t elem_at(const void *p, uint32_t eb,
                                            size_t i, const Acct *acct) {
  mark<ACCT>(acct, static_cast<const char *>(p) + i * eb);
  return ld_elem(p, eb, i, false);
}

template <bool ACCT>
__device__ __forceinline__ bool fetch_a(const DevSrc &s, uint32_t row,
                                        uint64_t *v, const Acct *acct) {
  if (s.fk == nullptr) {
    *v = elem_at<ACCT>(s.data, s.eb, row, acct);
    return true;
  }
  mark<ACCT>(acct, s.fk + row);
  const uint32_t k = s.fk[row];
  if (k >= s.dimsz) return false;
  *v = elem_at<ACCT>(s.data, s.eb, k, acct);
  return true;
}

// --------------------------------------------------------- program words
// Virtual registers r0..r2 live in VGPRs; instruction fields are kernel
// arguments, so the selects below are on wave-uniform values (the reference
// keeps the registers in shared memory, primitive.cuh:345-362).
// Branch-free selects keep the three registers in VGPRs; an if-chain here is
// turned into a private array indexed by the register number, i.e. scratch.
__device__ __forceinline__ uint64_t rget(uint64_t a, uint64_t b, uint64_t c,
                                         uint32_t i) {
  uint64_t r = c;
  r = i == 1 ? b : r;
  r = i == 0 ? a : r;
  return r;
}
__device__ __forceinline__ void rset(uint64_t &a, uint64_t &b, uint64_t &c,
                                     uint32_t i, uint64_t v) {
  a = i == 0 ? v : a;
  b = i == 1 ? v : b;
  c = i == 2 ? v : c;
}

__device__ __forceinline__ uint64_t sel4(uint64_t a, uint64_t b, uint64_t c,
                                         uint64_t d, uint32_t i) {
  uint64_t r = d;
  r = i == 2 ? c : r;
  r = i == 1 ? b : r;
  r = i == 0 ? a : r;
  return r;
}

template <bool ACCT>
__device__ __forceinline__ uint64_t ld_bmp(const uint64_t *p, uint32_t w,
                                           const Acct *acct) {
  mark<ACCT>(acct, p + w);
  return ld_stream(p + w);
}

// Loads the operand bitmaps e0 <= e < e1 of the program's flat list at this
// lane's words w0 + k * ws whose flag ld[k] is set, and ORs them into
// c[col][k] (middle bins) / u[col][k] (edge bins) -- c = OR of middle bins,
// u = leftmost | rightmost (primitive.cuh:216-226, 294-300).  kOpGroup
// bitmaps at a time, every load of a group issued before any is used, so a
// group costs one memory round trip.  A load is skipped (uniform branch)
// when no lane of the wave wants its word; otherwise every lane loads, an
// unwanted lane at the address of the first wanted lane (a line the wave
// fetches anyway), and its value is dropped.  Loads behind lane-divergent
// branches would each be waited for inside their branch.
constexpr int kOpGroup = 4;
// Bitmaps per load group of the conjunction form (EAGER 3).
constexpr int kConjGroup = 4;

// A bitmap word: streaming (non-temporal) for plans that read each word once,
// cached where other queries read the same lines soon after (the batched
// launch, whose items of one row range run back to back on one XCD).
template <bool NT>
__device__ __forceinline__ uint64_t ld_word(const uint64_t *p) {
  return NT ? ld_stream(p) : *p;
}

template <bool ACCT, bool CHK, int WPL, int NC, bool NT = true>
__device__ __forceinline__ void load_ops(const DevPrg &P, uint32_t e0,
                                         uint32_t e1, uint32_t w0, uint32_t ws,
                                         const bool (&ld)[WPL],
                                         uint64_t (&c)[NC][WPL],
                                         uint64_t (&u)[NC][WPL],
                                         const Acct *acct) {
  uint32_t wa[WPL];
  bool any[WPL];
#pragma unroll
  for (int k = 0; k < WPL; ++k) {
    const uint32_t w = w0 + ws * k;
    const uint64_t m = __ballot(ld[k]);
    any[k] = m != 0;
    const uint32_t wf = __shfl(w, any[k] ? (int)__builtin_ctzll(m) : 0, 64);
    wa[k] = ld[k] ? w : wf;
  }
  for (uint32_t g = e0; g < e1; g += kOpGroup) {
    uint64_t v[kOpGroup][WPL];
#pragma unroll
    for (int j = 0; j < kOpGroup; ++j) {
#pragma unroll
      for (int k = 0; k < WPL; ++k) v[j][k] = 0;
      if (g + j < e1) {
        const uint64_t *b = P.fl[g + j];
#pragma unroll
        for (int k = 0; k < WPL; ++k)
          if (any[k]) {
            if (ACCT && ld[k]) mark<true>(acct, b + wa[k]);
            v[j][k] = ld_word<NT>(b + wa[k]);
          }
      }
    }
#pragma unroll
    for (int j = 0; j < kOpGroup; ++j) {
      if (g + j >= e1) break;
      const uint32_t col = P.flcol[g + j];
      const bool edge = P.fledge[g + j] != 0;
#pragma unroll
      for (int k = 0; k < WPL; ++k) {
        const uint64_t x = ld[k] ? v[j][k] : 0ull;
#pragma unroll
        for (int cc = 0; cc < NC; ++cc) {
          const bool here = NC == 1 || col == (uint32_t)cc;
          c[cc][k] |= (here && !edge) ? x : 0ull;
          if (CHK) u[cc][k] |= (here && edge) ? x : 0ull;
        }
      }
    }
  }
}

// One instruction on one 64-row word; (cs, us) is the memory operand of
// ANDM / ORM (certain middle bins, uncertain edge bins).  nochk form:
// primitive.cuh:201-252; chk form: primitive.cuh:255-332, with the
// three-valued NOT corrected to P' = ~P_s | U_s, U' = U_s (the reference's
// P' = ~P_s drops uncertain rows).
template <bool CHK>
__device__ __forceinline__ void apply_ins(uint32_t op, uint32_t d, uint32_t s,
                                          uint64_t cs, uint64_t us,
                                          uint64_t &p0, uint64_t &p1,
                                          uint64_t &p2, uint64_t &u0,
                                          uint64_t &u1, uint64_t &u2) {
  const uint64_t pd = rget(p0, p1, p2, d);
  if (op == MBX_AND) {
    const uint64_t np = pd & rget(p0, p1, p2, s);
    rset(p0, p1, p2, d, np);
    if (CHK)
      rset(u0, u1, u2, d, (rget(u0, u1, u2, d) | rget(u0, u1, u2, s)) & np);
  } else if (op == MBX_OR) {
    const uint64_t ps = rget(p0, p1, p2, s);
    if (CHK) {
      const uint64_t ust = rget(u0, u1, u2, s), cst = ps & ~ust;
      rset(u0, u1, u2, d, (rget(u0, u1, u2, d) & ~cst) | (ust & ~pd));
    }
    rset(p0, p1, p2, d, pd | ps);
  } else if (op == MBX_NOT) {
    const uint64_t ps = rget(p0, p1, p2, s);
    if (CHK) {
      const uint64_t ust = rget(u0, u1, u2, s);
      rset(p0, p1, p2, d, ~ps | ust);
      rset(u0, u1, u2, d, ust);
    } else {
      rset(p0, p1, p2, d, ~ps);
    }
  } else if (op == MBX_ANDM) {
    const uint64_t np = pd & (us | cs);
    rset(p0, p1, p2, d, np);
    if (CHK) rset(u0, u1, u2, d, (rget(u0, u1, u2, d) | us) & np);
  } else if (op == MBX_ORM) {
    if (CHK) rset(u0, u1, u2, d, (rget(u0, u1, u2, d) & ~cs) | (us & ~pd));
    rset(p0, p1, p2, d, pd | us | cs);
  }
}

// Where the conjunction form reads the shape of its flat operand list:
// from the program descriptor (RuntimeShape, every product launch) or from
// compile-time constants (the dedicated kernels, whose shape is a template
// argument).
struct RuntimeShape {
  static constexpr bool kNT = true;   // bitmap words: streaming loads
  // words a round of the dense re-check (recheck_words)
  static constexpr uint32_t kRecheck = 4;
  static __device__ __forceinline__ uint32_t nfl(const DevPrg &P) { return P.nfl; }
  static __device__ __forceinline__ bool edge(const DevPrg &P, uint32_t e) {
    return P.fledge[e] != 0;
  }
  static __device__ __forceinline__ bool end(const DevPrg &P, uint32_t e) {
    return P.flend[e] != 0;
  }
  static __device__ __forceinline__ bool neg(const DevPrg &P, uint32_t e) {
    return P.flneg[e] != 0;
  }
};
// The same, bitmap words through cached loads (the batched launch), and a
// word a round of the dense re-check (its conjunction kernel is held to 80
// VGPRs: two words spill).
struct RuntimeShapeCached : RuntimeShape {
  static constexpr bool kNT = false;
  static constexpr uint32_t kRecheck = 1;
};

// EAGER 3: a conjunction of clauses (P.conj).  in[k] / keep[k]: word k is
// inside the range / its surviving row bits.
template <bool CHK, bool ACCT, int WPL, class F>
__device__ __forceinline__ void eval_conj(const DevPrg &P, uint32_t w0,
                                          uint32_t ws, uint32_t wend,
                                          const bool (&in)[WPL],
                                          const uint64_t (&keep)[WPL],
                                          uint64_t (&po)[WPL],
                                          uint64_t (&uo)[WPL],
                                          const Acct *acct) {
  (void)wend;
  // EAGER 3: a conjunction of clauses (P.conj): r0 = AND over clauses of
  // the OR of their bins (a clause is an OR of columns, or NOT of one
  // column; the flat list holds it contiguously, flend on its last entry),
  // complemented where flneg.  The flat list is walked in groups of
  // kConjGroup, every load of a group in flight at once; per word the work
  // is one OR per bitmap and one AND per clause -- no program
  // interpretation.  Three-valued (CHK), with c = OR of a clause's middle
  // bins and u = OR of its edge bins, the instructions reduce to
  //   p = AND(plain clauses: c | u) & AND(negated clauses: ~c)
  //   u = OR(clauses: u & ~c) & p
  // (ORM into an empty register gives p = c | u, u = u; ORM / OR onto a
  // clause give p |= p_s, u = (u | u_s) & ~(c | c_s) for disjoint bins;
  // ANDM and AND give p &= p_s, u = (u | u_s) & p; NOT gives
  // p = ~p_s | u_s = ~c, u = u_s).
  uint32_t wa[WPL];
  bool any[WPL];
  uint64_t acc[WPL], cur[WPL], accu[WPL], curu[WPL];
#pragma unroll
  for (int k = 0; k < WPL; ++k) {
    const uint32_t w = w0 + ws * k;
    const uint64_t m = __ballot(in[k]);
    any[k] = m != 0;
    const uint32_t wf = __shfl(w, any[k] ? (int)__builtin_ctzll(m) : 0, 64);
    wa[k] = in[k] ? w : wf;
    acc[k] = keep[k];
    cur[k] = 0;
    accu[k] = curu[k] = 0;
  }
  for (uint32_t g = 0; g < F::nfl(P); g += kConjGroup) {
    uint64_t v[kConjGroup][WPL];
#pragma unroll
    for (int j = 0; j < kConjGroup; ++j) {
#pragma unroll
      for (int k = 0; k < WPL; ++k) v[j][k] = 0;
      if (g + j < F::nfl(P)) {
        const uint64_t *b = P.fl[g + j];
#pragma unroll
        for (int k = 0; k < WPL; ++k)
          if (any[k]) {
            if (ACCT && in[k]) mark<true>(acct, b + wa[k]);
            v[j][k] = ld_word<F::kNT>(b + wa[k]);
          }
      }
    }
#pragma unroll
    for (int j = 0; j < kConjGroup; ++j) {
      if (g + j >= F::nfl(P)) break;
      if (CHK && F::edge(P, g + j)) {
#pragma unroll
        for (int k = 0; k < WPL; ++k) curu[k] |= v[j][k];
      } else {
#pragma unroll
        for (int k = 0; k < WPL; ++k) cur[k] |= v[j][k];
      }
      if (F::end(P, g + j)) {
        // plain: c | u; negated: ~c = (c | (u & 0)) ^ ~0
        const uint64_t flip = F::neg(P, g + j) ? ~0ull : 0ull;
#pragma unroll
        for (int k = 0; k < WPL; ++k) {
          const uint64_t lit = CHK ? (cur[k] | (curu[k] & ~flip)) : cur[k];
          acc[k] &= lit ^ flip;
          if (CHK) {
            accu[k] |= curu[k] & ~cur[k];
            curu[k] = 0;
          }
          cur[k] = 0;
        }
      }
    }
    if (g + kConjGroup < F::nfl(P)) {
      // Early-out between groups: a word whose running AND is empty in
      // every lane skips the later groups' loads (uniform branch), and a
      // lane whose word is empty loads the first live lane's address, so
      // a 128-byte line with no live word is not fetched.
#pragma unroll
      for (int k = 0; k < WPL; ++k) {
        const bool live = acc[k] != 0;
        const uint64_t m = __ballot(live);
        any[k] = m != 0;
        const uint32_t w = w0 + ws * k;
        const uint32_t wf = __shfl(w, any[k] ? (int)__builtin_ctzll(m) : 0, 64);
        wa[k] = live ? w : wf;
      }
    }
  }
#pragma unroll
  for (int k = 0; k < WPL; ++k) {
    po[k] = acc[k];
    uo[k] = CHK ? (accu[k] & acc[k]) : 0;
  }
}

// Evaluates the program on this lane's words w0 + k * ws (k < WPL, word <
// wend).  EAGER = false keeps the ANDM early-out of primitive.cuh:215/292: a
// word whose destination is already empty loads nothing, so a 128-byte line
// of the operand is fetched only if one of its 16 words is live -- at the
// cost of one memory round trip per memory instruction (the WPL words of an
// instruction share it).  EAGER = true loads the operand of every column a
// memory instruction reads up front (one round trip in all), then runs the
// program word by word, so the register state is not replicated per word.
// Both give the same bits: an empty destination has no uncertain rows (U is
// a subset of P after every instruction), so ANDM on it yields the same
// empty pair whether or not the operand is read.
// Eager evaluation with every operand load of the program in flight at
// once, straight into registers (no per-column accumulators): for programs
// whose flat operand list has at most kDirectOps bitmaps.
constexpr uint32_t kDirectOps = 5;

template <bool CHK, bool ACCT, int EAGER, int WPL, class F = RuntimeShape>
__device__ __forceinline__ void eval_words(const DevPrg &P, uint32_t w0,
                                           uint32_t ws, uint32_t wend,
                                           uint64_t (&po)[WPL],
                                           uint64_t (&uo)[WPL],
                                           const Acct *acct) {
  bool in[WPL];
  uint64_t keep[WPL];
#pragma unroll
  for (int k = 0; k < WPL; ++k) {
    const uint32_t w = w0 + ws * k;
    in[k] = w < wend;
    // rows >= N never survive (the reference has no such guard,
    // primitive.cuh:357-360 only bounds the word index)
    keep[k] = in[k] ? ~0ull : 0ull;
    if (in[k] && w == P.nwords - 1 && (P.nrows & 63u))
      keep[k] = (1ull << (P.nrows & 63u)) - 1ull;
  }
  if (EAGER == 3) {
    eval_conj<CHK, ACCT, WPL, F>(P, w0, ws, wend, in, keep, po, uo, acct);
    return;
  }
  if (EAGER == 4) {
    // EAGER 4, the ablation of where the registers live (abla::base_fuse,
    // primitive_abla.cuh:15-59): the early-out interpreter of EAGER 0, with
    // the virtual registers of each word in device memory (P.greg) -- every
    // instruction reads its registers from there and writes them back.
    uint64_t *gr[WPL];
#pragma unroll
    for (int k = 0; k < WPL; ++k) {
      gr[k] = P.greg + (size_t)(in[k] ? w0 + ws * k : 0) * 6;
      if (in[k])
#pragma unroll
        for (int r = 0; r < 6; ++r) gr[k][r] = 0;
    }
#pragma unroll
    for (uint32_t i = 0; i < MBX_MAXNINSTR; ++i) {
      if (i >= P.nins) break;
      const uint32_t op = P.ins[i][0], d = P.ins[i][1], s = P.ins[i][2];
      uint64_t cs[1][WPL], us[1][WPL];
#pragma unroll
      for (int k = 0; k < WPL; ++k) cs[0][k] = us[0][k] = 0;
      if (op == MBX_ANDM || op == MBX_ORM) {
        bool ld[WPL];
#pragma unroll
        for (int k = 0; k < WPL; ++k)
          ld[k] = in[k] && (op == MBX_ORM || gr[k][d % 3] != 0);
        load_ops<ACCT, CHK, WPL, 1, F::kNT>(P, P.beg[s], P.end[s], w0, ws, ld, cs, us,
                                    acct);
      }
#pragma unroll
      for (int k = 0; k < WPL; ++k) {
        if (!in[k]) continue;
        uint64_t p0 = gr[k][0], p1 = gr[k][1], p2 = gr[k][2];
        uint64_t u0 = gr[k][3], u1 = gr[k][4], u2 = gr[k][5];
        apply_ins<CHK>(op, d, s, cs[0][k], us[0][k], p0, p1, p2, u0, u1, u2);
        gr[k][0] = p0, gr[k][1] = p1, gr[k][2] = p2;
        gr[k][3] = u0, gr[k][4] = u1, gr[k][5] = u2;
      }
    }
#pragma unroll
    for (int k = 0; k < WPL; ++k) {
      po[k] = in[k] ? gr[k][0] & keep[k] : 0;
      uo[k] = (CHK && in[k]) ? gr[k][3] & keep[k] : 0;
    }
    return;
  }
  if (EAGER == 2) {
    // EAGER 2: the flat list holds <= kDirectOps bitmaps; v[e][k] is entry e
    // at word k; a column's operand is the OR of its entries [beg, end)
    uint64_t v[kDirectOps][WPL];
    uint32_t wa[WPL];
    bool any[WPL];
#pragma unroll
    for (int k = 0; k < WPL; ++k) {
      const uint32_t w = w0 + ws * k;
      const uint64_t m = __ballot(in[k]);
      any[k] = m != 0;
      const uint32_t wf = __shfl(w, any[k] ? (int)__builtin_ctzll(m) : 0, 64);
      wa[k] = in[k] ? w : wf;
    }
#pragma unroll
    for (uint32_t e = 0; e < kDirectOps; ++e) {
#pragma unroll
      for (int k = 0; k < WPL; ++k) v[e][k] = 0;
      if (e < P.nfl) {
        const uint64_t *b = P.fl[e];
#pragma unroll
        for (int k = 0; k < WPL; ++k)
          if (any[k]) {
            if (ACCT && in[k]) mark<true>(acct, b + wa[k]);
            v[e][k] = ld_word<F::kNT>(b + wa[k]);
          }
      }
    }
#pragma unroll
    for (int k = 0; k < WPL; ++k) {
      uint64_t p0 = 0, p1 = 0, p2 = 0, u0 = 0, u1 = 0, u2 = 0;
#pragma unroll
      for (uint32_t i = 0; i < MBX_MAXNINSTR; ++i) {
        if (i >= P.nins) break;
        const uint32_t op = P.ins[i][0], d = P.ins[i][1], s = P.ins[i][2];
        uint64_t cs = 0, us = 0;
        if (op == MBX_ANDM || op == MBX_ORM) {
          const uint32_t eb = P.beg[s], ee = P.end[s];
#pragma unroll
          for (uint32_t e = 0; e < kDirectOps; ++e) {
            const bool sel = e >= eb && e < ee;
            const bool edge = P.fledge[e] != 0;
            cs |= (sel && !edge) ? v[e][k] : 0ull;
            if (CHK) us |= (sel && edge) ? v[e][k] : 0ull;
          }
        }
        apply_ins<CHK>(op, d, s, cs, us, p0, p1, p2, u0, u1, u2);
      }
      po[k] = p0 & keep[k];
      uo[k] = CHK ? (u0 & keep[k]) : 0;
    }
    return;
  }
  if (EAGER) {
    uint64_t pc[MBX_MAXCOLS][WPL], pu[MBX_MAXCOLS][WPL];
#pragma unroll
    for (uint32_t c = 0; c < MBX_MAXCOLS; ++c)
#pragma unroll
      for (int k = 0; k < WPL; ++k) pc[c][k] = pu[c][k] = 0;
    load_ops<ACCT, CHK, WPL, MBX_MAXCOLS, F::kNT>(P, 0, P.nfl, w0, ws, in, pc, pu,
                                          acct);
#pragma unroll
    for (int k = 0; k < WPL; ++k) {
      uint64_t p0 = 0, p1 = 0, p2 = 0, u0 = 0, u1 = 0, u2 = 0;
#pragma unroll
      for (uint32_t i = 0; i < MBX_MAXNINSTR; ++i) {
        if (i >= P.nins) break;
        const uint32_t op = P.ins[i][0], d = P.ins[i][1], s = P.ins[i][2];
        const uint64_t cs = sel4(pc[0][k], pc[1][k], pc[2][k], pc[3][k], s);
        const uint64_t us =
            CHK ? sel4(pu[0][k], pu[1][k], pu[2][k], pu[3][k], s) : 0;
        apply_ins<CHK>(op, d, s, cs, us, p0, p1, p2, u0, u1, u2);
      }
      po[k] = p0 & keep[k];
      uo[k] = CHK ? (u0 & keep[k]) : 0;
    }
    return;
  }
  uint64_t p0[WPL], p1[WPL], p2[WPL], u0[WPL], u1[WPL], u2[WPL];
#pragma unroll
  for (int k = 0; k < WPL; ++k) p0[k] = p1[k] = p2[k] = u0[k] = u1[k] = u2[k] = 0;
#pragma unroll
  for (uint32_t i = 0; i < MBX_MAXNINSTR; ++i) {
    if (i >= P.nins) break;
    const uint32_t op = P.ins[i][0], d = P.ins[i][1], s = P.ins[i][2];
    uint64_t cs[1][WPL], us[1][WPL];
#pragma unroll
    for (int k = 0; k < WPL; ++k) cs[0][k] = us[0][k] = 0;
    if (op == MBX_ANDM || op == MBX_ORM) {
      bool ld[WPL];
#pragma unroll
      for (int k = 0; k < WPL; ++k)
        ld[k] = in[k] && (op == MBX_ORM || rget(p0[k], p1[k], p2[k], d) != 0);
      load_ops<ACCT, CHK, WPL, 1, F::kNT>(P, P.beg[s], P.end[s], w0, ws, ld, cs, us,
                                  acct);
    }
#pragma unroll
    for (int k = 0; k < WPL; ++k)
      apply_ins<CHK>(op, d, s, cs[0][k], us[0][k], p0[k], p1[k], p2[k], u0[k],
                     u1[k], u2[k]);
  }
#pragma unroll
  for (int k = 0; k < WPL; ++k) {
    po[k] = p0[k] & keep[k];
    uo[k] = CHK ? (u0[k] & keep[k]) : 0;
  }
}

template <bool CHK, bool ACCT, int EAGER = 0>
__device__ __forceinline__ void eval_word(const DevPrg &P, uint32_t w,
                                          uint64_t &po, uint64_t &uo,
                                          const Acct *acct) {
  uint64_t p[1], u[1];
  eval_words<CHK, ACCT, EAGER, 1>(P, w, 64u, P.nwords, p, u, acct);
  po = p[0];
  uo = u[0];
}

// ----------------------------------------------------------- row operation
// Every value the row op needs is fetched in two rounds: first the foreign
// keys of all its sources, then all values.  Sources never wait on one
// another, so a row costs two memory round trips whatever its number of
// predicates, key parts and measures (the reference's ops fetch one source
// after the other, ssb/lambdas.cuh:104-151).
template <bool ACCT>
__device__ __forceinline__ uint32_t src_index(const DevSrc &s, uint32_t row,
                                              const Acct *acct) {
  if (s.fk == nullptr) return row;
  mark<ACCT>(acct, s.fk + row);
  return s.fk[row];
}

template <bool ACCT>
__device__ __forceinline__ uint64_t src_value(const DevSrc &s, uint32_t idx,
                                              bool &ok, const Acct *acct) {
  ok = s.fk == nullptr || idx < s.dimsz;   // fk >= dimsz: no value
  return ok ? elem_at<ACCT>(s.data, s.eb, idx, acct) : 0;
}

// Re-check of an uncertain row: the program over the exact predicates of
// its columns (the op's chk branch, ssb/lambdas.cuh:27-33), looked up in the
// program's truth table (bit c of the index = column c's predicate).  For
// a candidate row (CAND: a row the bitmaps left in P or U) the columns in
// P.known are not loaded: their value is P.known_val (no bit at or above
// npreds: a program with edge bins reads only columns below npreds, which
// the launch checks).  The scan plan re-checks every row and loads every
// column.
template <bool ACCT, bool CAND = true>
__device__ __forceinline__ bool row_passes(const DevPrg &P, const DevOp &O,
                                           uint32_t row, const Acct *acct) {
  const uint32_t known = CAND ? P.known : 0u;
  uint32_t ix[MBX_MAXCOLS];
#pragma unroll
  for (int c = 0; c < MBX_MAXCOLS; ++c)
    if (c < (int)O.npreds && !((known >> c) & 1u))
      ix[c] = src_index<ACCT>(O.pred[c], row, acct);
  uint32_t m = CAND ? P.known_val : 0u;
#pragma unroll
  for (int c = 0; c < MBX_MAXCOLS; ++c)
    if (c < (int)O.npreds && !((known >> c) & 1u)) {
      bool ok;
      const uint64_t v = src_value<ACCT>(O.pred[c], ix[c], ok, acct);
      m |= (ok && v >= O.lo[c] && v < O.hi[c]) ? 1u << c : 0u;
    }
  return (P.truth >> m) & 1u;
}

// row_passes for E rows at once: every predicate load of the E rows is in
// flight together (a row costs two round trips however many there are).
// live[e] in: the row is to be checked; out: it passes.
template <bool ACCT, bool CAND, int E>
__device__ __forceinline__ void rows_pass(const DevPrg &P, const DevOp &O,
                                          const uint32_t *row, bool *live,
                                          const Acct *acct) {
  const uint32_t known = CAND ? P.known : 0u;
  uint32_t ix[E][MBX_MAXCOLS];
#pragma unroll
  for (int e = 0; e < E; ++e)
#pragma unroll
    for (int c = 0; c < MBX_MAXCOLS; ++c)
      if (live[e] && c < (int)O.npreds && !((known >> c) & 1u))
        ix[e][c] = src_index<ACCT>(O.pred[c], row[e], acct);
#pragma unroll
  for (int e = 0; e < E; ++e) {
    if (!live[e]) continue;
    uint32_t m = CAND ? P.known_val : 0u;
#pragma unroll
    for (int c = 0; c < MBX_MAXCOLS; ++c)
      if (c < (int)O.npreds && !((known >> c) & 1u)) {
        bool ok;
        const uint64_t v = src_value<ACCT>(O.pred[c], ix[e][c], ok, acct);
        m |= (ok && v >= O.lo[c] && v < O.hi[c]) ? 1u << c : 0u;
      }
    live[e] = (P.truth >> m) & 1u;
  }
}

// x / kp.div with key_div_magic's multiplier and shifts: exact for every
// 64-bit x, a few multiplies instead of the 64-bit division sequence.
__device__ __forceinline__ uint64_t key_div(const DevKey &kp, uint64_t x) {
  const uint64_t t = __umul64hi(kp.magic, x);
  return (t + ((x - t) >> kp.sh1)) >> kp.sh2;
}

// Group key and value of one row from its fetched sources v (key parts,
// then measures a and b); ok is false for a dropped row (a source past its
// dimension, or a key outside [0, nr_grp)).
__device__ __forceinline__ void keyval_of(const DevOp &O,
                                          const uint64_t (&v)[MBX_MAXKEYS + 2],
                                          bool vok, uint32_t nr_grp,
                                          uint32_t &key, int64_t &val,
                                          bool &ok) {
  int64_t k = 0;
#pragma unroll
  for (int i = 0; i < MBX_MAXKEYS; ++i)
    if (i < (int)O.nkeys) {
      const DevKey &kp = O.keys[i];
      const uint64_t x = v[i];
      uint64_t f = x;
      if (kp.kind == MBX_KEY_DIV) {
        f = key_div(kp, x);
      } else if (kp.kind == MBX_KEY_BUCKET) {
        f = 0;
        for (uint32_t j = 0; j < kp.nb; ++j) f += kp.bounds[j] <= x;
      }
      k += kp.stride * ((int64_t)f - kp.base);
    }
  const uint64_t a = v[3], b = v[4];
  val = O.measure == MBX_MEAS_MUL   ? (int64_t)(a * b)
        : O.measure == MBX_MEAS_SUB ? (int64_t)a - (int64_t)b
                                    : (int64_t)a;
  key = (uint32_t)k;
  ok = vok && k >= 0 && k < (int64_t)nr_grp;
}

// {value, group} of E selected rows at once (primitive.cuh:10-15); ok[e] is
// false for a dropped row (fk out of range or key outside [0, nr_grp)).
template <bool ACCT, int E>
__device__ __forceinline__ void rows_keyval(const DevOp &O, const uint32_t *row,
                                            const bool *live, uint32_t nr_grp,
                                            uint32_t *key, int64_t *val,
                                            bool *ok, const Acct *acct) {
  constexpr int NS = MBX_MAXKEYS + 2;   // key parts, then measure a, b
  const bool has_a = O.measure != MBX_MEAS_COUNT;
  const bool has_b = O.measure == MBX_MEAS_MUL || O.measure == MBX_MEAS_SUB;
  uint32_t ix[E][NS];
#pragma unroll
  for (int e = 0; e < E; ++e) {
    if (!live[e]) continue;
#pragma unroll
    for (int i = 0; i < MBX_MAXKEYS; ++i)
      if (i < (int)O.nkeys) ix[e][i] = src_index<ACCT>(O.keys[i].src, row[e], acct);
    if (has_a) ix[e][3] = src_index<ACCT>(O.ma, row[e], acct);
    if (has_b) ix[e][4] = src_index<ACCT>(O.mb, row[e], acct);
  }
  uint64_t v[E][NS];
  bool vok[E];
#pragma unroll
  for (int e = 0; e < E; ++e) {
    vok[e] = live[e];
    if (!live[e]) continue;
#pragma unroll
    for (int i = 0; i < MBX_MAXKEYS; ++i)
      if (i < (int)O.nkeys) {
        bool o;
        v[e][i] = src_value<ACCT>(O.keys[i].src, ix[e][i], o, acct);
        vok[e] = vok[e] && o;
      }
    v[e][3] = 1;
    v[e][4] = 0;
    if (has_a) {
      bool o;
      v[e][3] = src_value<ACCT>(O.ma, ix[e][3], o, acct);
      vok[e] = vok[e] && o;
    }
    if (has_b) {
      bool o;
      v[e][4] = src_value<ACCT>(O.mb, ix[e][4], o, acct);
      vok[e] = vok[e] && o;
    }
  }
#pragma unroll
  for (int e = 0; e < E; ++e) {
    int64_t k = 0;
#pragma unroll
    for (int i = 0; i < MBX_MAXKEYS; ++i)
      if (i < (int)O.nkeys) {
        const DevKey &kp = O.keys[i];
        const uint64_t x = v[e][i];
        uint64_t f = x;
        if (kp.kind == MBX_KEY_DIV) {
          f = key_div(kp, x);
        } else if (kp.kind == MBX_KEY_BUCKET) {
          f = 0;
          for (uint32_t j = 0; j < kp.nb; ++j) f += kp.bounds[j] <= x;
        }
        k += kp.stride * ((int64_t)f - kp.base);
      }
    const uint64_t a = v[e][3], b = v[e][4];
    val[e] = O.measure == MBX_MEAS_MUL   ? (int64_t)(a * b)
             : O.measure == MBX_MEAS_SUB ? (int64_t)a - (int64_t)b
                                         : (int64_t)a;
    key[e] = (uint32_t)k;
    ok[e] = vok[e] && k >= 0 && k < (int64_t)nr_grp;
  }
}

template <bool ACCT>
__device__ __forceinline__ bool row_keyval(const DevOp &O, uint32_t row,
                                           uint32_t nr_grp, uint32_t &key,
                                           int64_t &val, const Acct *acct) {
  const bool live = true;
  bool ok;
  rows_keyval<ACCT, 1>(O, &row, &live, nr_grp, &key, &val, &ok, acct);
  return ok;
}

// Aggregation target: the workgroup's LDS table (small group counts) or the
// output in HBM (global int64 atomics).  The LDS table is the dynamic LDS of
// the launch, named here so that its atomics stay in the LDS address space
// (they then never order against the global loads of other rows).
extern __shared__ unsigned long long s_grp_dyn[];

struct GroupSink {
  bool lds;                  // true: LDS table of nr_grp slots
  unsigned long long *out;   // HBM output
};

// The key indexes memory, so it is bounds-checked right here, behind an
// empty asm that hides its provenance from the optimizer: gfx950 code built
// from an earlier shape of row_keyval used a stale register as the key on
// some measure paths (the live key was overwritten by the measure load),
// and a wild key must cost a wrong sum in a test, never a memory fault.
__device__ __forceinline__ void sink_add(const GroupSink &s, uint32_t key,
                                         int64_t val, uint32_t nr_grp) {
  asm volatile("" : "+v"(key));
  if (key >= nr_grp) return;
  if (s.lds) atomicAdd(&s_grp_dyn[key], (unsigned long long)val);
  else atomicAdd(s.out + key, (unsigned long long)val);
}

// ------------------------------------------------------------ table flush
// Device-scope atomics on one address serialise (about 12 ns apiece on
// MI355X), so ~2000 persistent workgroups adding their LDS tables straight
// into the output spend tens of microseconds on each hot group.  Instead
// member m of nmembers adds its non-zero groups into replica m % R of the
// table (R = min(nmembers, kSpreadRep), so about nmembers / R adds per
// address); one counter per replica finds the replica's last member, which
// moves the replica into the output (R adds per output address) and leaves
// replica and counter zero for the next launch.  Callers: every workgroup of
// the grid, with its LDS table complete; nmembers must be the same in all.
//
// Ordering uses returning device-scope atomics only: a returned atomic has
// been performed at the device's coherence point, so a member's adds are
// performed before its counter increment is issued, and the last member's
// exchanges, issued after it saw the final count, read every add.  No
// agent-scope fence is needed (on gfx950 one writes back and invalidates the
// L2, which costs far more than the atomics it would order).
struct Spread {
  unsigned long long *rep;   // kSpreadRep tables of gpad int64, zero on entry
  uint32_t *done;            // kSpreadRep counters, kSpreadDoneStride apart
  uint32_t gpad;             // table stride: nr_grp rounded up to 16
};

// A thread's groups are i = tid + k * blockDim.x; kFlushU of them go out
// together, every atomic issued before any result is used, so a thread's
// adds cost one round trip per kFlushU groups (a result consumed inside the
// lane-divergent `if (v)` would be waited for there, one round trip each).
constexpr uint32_t kFlushU = 4;

__device__ __forceinline__ void flush_groups(const Spread &sp,
                                             unsigned long long *out,
                                             uint32_t nr_grp, uint32_t member,
                                             uint32_t nmembers) {
  __shared__ uint32_t s_last;
  const uint32_t R = nmembers < kSpreadRep ? nmembers : kSpreadRep;
  const uint32_t r = member % R;
  unsigned long long *rep = sp.rep + (size_t)r * sp.gpad;
  const uint32_t nt = blockDim.x;
  __syncthreads();
  unsigned long long seen = 0;
  for (uint32_t i0 = threadIdx.x; i0 < nr_grp; i0 += kFlushU * nt) {
    unsigned long long got[kFlushU];
#pragma unroll
    for (uint32_t u = 0; u < kFlushU; ++u) {
      const uint32_t i = i0 + u * nt;
      const unsigned long long v = i < nr_grp ? s_grp_dyn[i] : 0ull;
      got[u] = 0;
      if (v)
        got[u] = __hip_atomic_fetch_add(rep + i, v, __ATOMIC_RELAXED,
                                        __HIP_MEMORY_SCOPE_AGENT);
    }
#pragma unroll
    for (uint32_t u = 0; u < kFlushU; ++u) seen |= got[u];
  }
  asm volatile("" ::"v"(seen));   // every add has returned
  __syncthreads();
  if (threadIdx.x == 0) {
    const uint32_t members = nmembers / R + (r < nmembers % R ? 1u : 0u);
    const uint32_t before = __hip_atomic_fetch_add(
        sp.done + (size_t)r * kSpreadDoneStride, 1u, __ATOMIC_RELAXED,
        __HIP_MEMORY_SCOPE_AGENT);
    s_last = before == members - 1;
  }
  __syncthreads();
  if (!s_last) return;
  for (uint32_t i0 = threadIdx.x; i0 < nr_grp; i0 += kFlushU * nt) {
    unsigned long long v[kFlushU];
#pragma unroll
    for (uint32_t u = 0; u < kFlushU; ++u) {
      const uint32_t i = i0 + u * nt;
      v[u] = 0;
      if (i < nr_grp)
        v[u] = __hip_atomic_exchange(rep + i, 0ull, __ATOMIC_RELAXED,
                                     __HIP_MEMORY_SCOPE_AGENT);
    }
#pragma unroll
    for (uint32_t u = 0; u < kFlushU; ++u)
      if (v[u]) atomicAdd(out + i0 + u * nt, v[u]);
  }
  if (threadIdx.x == 0)
    (void)__hip_atomic_exchange(sp.done + (size_t)r * kSpreadDoneStride, 0u,
                                __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// One row per lane into the sink, where atomics on one address serialise:
// the rows of the wave's two leading keys are summed across the wave first
// (one atomic per key); the remaining rows add directly.
__device__ __forceinline__ void sink_wave(const GroupSink &sink, uint32_t key,
                                          int64_t val, bool ok,
                                          uint32_t nr_grp) {
  bool pend = ok;
#pragma unroll
  for (int round = 0; round < 2; ++round) {
    const unsigned long long act = __ballot(pend);
    if (!act) break;
    const int leader = (int)__builtin_ctzll(act);
    const uint32_t lk = __shfl(key, leader, 64);
    const bool mine = pend && key == lk;
    const unsigned long long v =
        wave_sum_u64(mine ? (unsigned long long)val : 0ull);
    if (lane_id() == (uint32_t)leader) sink_add(sink, lk, (int64_t)v, nr_grp);
    pend = pend && !mine;
  }
  if (pend) sink_add(sink, key, val, nr_grp);
}

// List entries: row id, MSB = uncertain (primitive.cuh:406-407).  E entries
// per lane share their memory round trips.
template <bool CHK, bool ACCT, int E>
__device__ __forceinline__ void process_entries(const DevPrg &P,
                                                const DevOp &O,
                                                uint32_t nr_grp,
                                                const uint32_t *ent,
                                                const bool *valid,
                                                const GroupSink &sink,
                                                const Acct *acct) {
  uint32_t key[E];
  int64_t val[E];
  bool ok[E];
  uint32_t row[E];
  bool live[E];
#pragma unroll
  for (int e = 0; e < E; ++e) {
    row[e] = ent[e] & 0x7fffffffu;
    live[e] = valid[e];
    if (CHK && live[e] && (ent[e] >> 31))
      live[e] = row_passes<ACCT>(P, O, row[e], acct);
  }
  rows_keyval<ACCT, E>(O, row, live, nr_grp, key, val, ok, acct);
#pragma unroll
  for (int e = 0; e < E; ++e) {
    if (ACCT) {
      if (ok[e]) atomicAdd(acct->selected, 1ull);
      continue;
    }
    if (sink.lds) {
      if (ok[e]) sink_add(sink, key[e], val[e], nr_grp);
      continue;
    }
    // HBM output: atomics on one address serialise
    sink_wave(sink, key[e], val[e], ok[e], nr_grp);
  }
}

template <bool CHK, bool ACCT>
__device__ __forceinline__ void process_entry(const DevPrg &P, const DevOp &O,
                                              uint32_t nr_grp, uint32_t e,
                                              const GroupSink &sink,
                                              const Acct *acct) {
  const bool valid = true;
  process_entries<CHK, ACCT, 1>(P, O, nr_grp, &e, &valid, sink, acct);
}

// ------------------------------------------------------- block primitives
// Exclusive scan of one value per thread over a 256-thread workgroup.
// s_wave holds 4 wave totals.  Returns the exclusive prefix; *total gets the
// workgroup sum.  Contains two barriers (the second keeps s_wave safe for the
// next call).
__device__ __forceinline__ uint32_t block_excl_scan(uint32_t x,
                                                    uint32_t *s_wave,
                                                    uint32_t *total) {
  const uint32_t lane = lane_id(), wv = threadIdx.x >> 6;
  const uint32_t incl = wave_incl_scan(x);
  if (lane == 63) s_wave[wv] = incl;
  __syncthreads();
  uint32_t before = 0, sum = 0;
#pragma unroll
  for (uint32_t j = 0; j < 4; ++j) {
    const uint32_t t = s_wave[j];
    before += j < wv ? t : 0;
    sum += t;
  }
  __syncthreads();
  *total = sum;
  return before + incl - x;
}

// ------------------------------------------------------------- scan words
// The scan plan (grpby, primitive.cuh:16-48) as 64-row words: every row's
// exact predicates, no bitmaps.  For each word of its range the wave loads
// the 64 rows' predicate values lane by lane (one coalesced 256-byte access
// per column) and a ballot of the rows' results is the word; lane j keeps
// word j of each 64-word chunk, the layout eval_words produces, so the
// compaction and row op that follow are the fused plan's.  kScanU words
// have their loads in flight together, and no load sits behind a
// lane-divergent branch (rows past the end read row 0 and are dropped).
constexpr int kScanU = 4;

template <int WPL>
__device__ __forceinline__ void scan_words(const DevPrg &P, const DevOp &O,
                                           uint32_t w0, uint32_t ws,
                                           uint32_t wend, uint64_t (&po)[WPL]) {
  const uint32_t lane = lane_id();
#pragma unroll
  for (int k = 0; k < WPL; ++k) {
    uint64_t mine = 0;
    const uint32_t wk = w0 + ws * k;   // the chunk's first word
    for (uint32_t j0 = 0; j0 < 64u && wk + j0 < wend; j0 += kScanU) {
      uint32_t row[kScanU];
      bool in[kScanU];
      uint32_t ix[kScanU][MBX_MAXCOLS];
#pragma unroll
      for (int u = 0; u < kScanU; ++u) {
        const uint32_t w = wk + j0 + (uint32_t)u;
        row[u] = w * 64u + lane;
        in[u] = w < wend && row[u] < P.nrows;
        const uint32_t r = in[u] ? row[u] : 0u;
#pragma unroll
        for (int c = 0; c < MBX_MAXCOLS; ++c)
          if (c < (int)O.npreds)
            ix[u][c] = O.pred[c].fk ? O.pred[c].fk[r] : r;
      }
#pragma unroll
      for (int u = 0; u < kScanU; ++u) {
        uint32_t m = 0;
#pragma unroll
        for (int c = 0; c < MBX_MAXCOLS; ++c) {
          if (c >= (int)O.npreds) continue;
          const DevSrc &sc = O.pred[c];
          if (sc.fk && sc.dimsz == 0) continue;   // no dimension row at all
          const bool ok = sc.fk == nullptr || ix[u][c] < sc.dimsz;
          const uint64_t v = ld_elem(sc.data, sc.eb, ok ? ix[u][c] : 0u, false);
          m |= (ok && v >= O.lo[c] && v < O.hi[c]) ? 1u << c : 0u;
        }
        const bool pass = in[u] && ((P.truth >> m) & 1u);
        const uint64_t b = __ballot(pass);
        if (lane == j0 + (uint32_t)u) mine = b;
      }
    }
    po[k] = mine;
  }
}

// ------------------------------------------------------------ tile pipeline

// One wave's share: lane l holds the words w0 + l + k * ws (k < WPL) below
// wend -- WPL chunks of 64 consecutive words, ws apart.  Program -> popcount
// -> wave scan -> the wave's own LDS row list (rounds of kWaveList entries)
// -> row op -> sink.
// Waves never wait for each other (no workgroup barrier), so a CU overlaps
// the bitmap loads of some waves with the gathers of others.  The
// reference's method 1 / method 2 switch (primitive.cuh:373-397) becomes
// rounds: a dense range takes more rounds, each using all lanes.
constexpr uint32_t kWaveList = kListCap / 4;
// A wave range with at least this many candidate rows per word on average
// (a quarter of its rows) skips the row list.  The list is filled round by
// round, and in a dense range only the two or three lanes whose rows fall
// in the round write it, one row at a time: for ranges where most rows are
// uncertain (an edge bin over a skewed column's head) that serial fill cost
// more than the row op itself.
constexpr uint32_t kDenseRowsPerWord = 16;

// The three-valued words of a dense range (an edge bin over a skewed
// column's head: most rows uncertain, few passing) made exact: a word with
// uncertain rows is re-checked, lane b taking bit b, so a predicate
// column's loads for the word are one coalesced 256-B access; p[k] becomes
// the word's passing rows and u[k] 0.
// The row op then runs on the passing rows only, through the list path.
// Kept apart from the row op so that the loop holds the predicates'
// descriptors only.
// R words a round: their loads are all in flight before the first compare.
template <int WPL, uint32_t R>
__device__ __forceinline__ void recheck_words(const DevPrg &P, const DevOp &O,
                                              uint32_t w0, uint32_t ws,
                                              uint64_t *p, uint64_t *u) {
  const uint32_t lane = lane_id();
  const uint32_t need = ~P.known & ((1u << O.npreds) - 1u);
#pragma unroll
  for (int k = 0; k < WPL; ++k) {
    const uint32_t wk = w0 + ws * k;   // lane s holds word wk + s
    uint64_t mine = 0;                 // lane s: word wk + s's passing rows
    for (uint32_t s = 0; s < 64u; s += R) {
      uint64_t pw[R], uw[R];
      bool anyu = false;
#pragma unroll
      for (uint32_t e = 0; e < R; ++e) {
        pw[e] = wave_read64(p[k], s + e);
        uw[e] = wave_read64(u[k], s + e);
        anyu = anyu || (pw[e] & uw[e]) != 0;   // wave-uniform
      }
      uint64_t m[R];
#pragma unroll
      for (uint32_t e = 0; e < R; ++e) m[e] = pw[e];
      if (anyu) {
        // every lane reads its rows' predicate columns, so the loads are
        // branch-free (a row past the table reads the last row; it is no
        // candidate)
        uint32_t rc[R];
#pragma unroll
        for (uint32_t e = 0; e < R; ++e)
          rc[e] = min((wk + s + e) * 64u + lane, P.nrows - 1u);
        uint32_t bits[R];
#pragma unroll
        for (uint32_t e = 0; e < R; ++e) bits[e] = P.known_val;
#pragma unroll
        for (int c = 0; c < MBX_MAXCOLS; ++c) {
          if (!((need >> c) & 1u)) continue;
          const DevSrc &sc = O.pred[c];
          uint64_t v[R];
          bool okv[R];
#pragma unroll
          for (uint32_t e = 0; e < R; ++e) {
            const uint32_t ix = sc.fk ? sc.fk[rc[e]] : rc[e];
            okv[e] = !(sc.fk && (sc.dimsz == 0 || ix >= sc.dimsz));
            v[e] = ld_elem(sc.data, sc.eb, okv[e] ? ix : 0u, false);
          }
#pragma unroll
          for (uint32_t e = 0; e < R; ++e)
            bits[e] |= (okv[e] && v[e] >= O.lo[c] && v[e] < O.hi[c]) ? 1u << c : 0u;
        }
#pragma unroll
        for (uint32_t e = 0; e < R; ++e) {
          const bool cand = (pw[e] >> lane) & 1ull;
          const bool unc = (uw[e] >> lane) & 1ull;
          m[e] = __ballot(cand && (!unc || ((P.truth >> bits[e]) & 1u)));
        }
      }
#pragma unroll
      for (uint32_t e = 0; e < R; ++e)
        if (lane == s + e) mine = m[e];
    }
    p[k] = mine;
    u[k] = 0;
  }
}

// SCAN: the words come from the rows' exact predicates (scan_words), not
// from bitmaps.
template <bool CHK, bool ACCT, int WPL, int EAGER, int E = 2, bool DIRECT = false,
          bool SCAN = false, class F = RuntimeShape>
__device__ __forceinline__ void run_wave_range(const DevPrg &P, const DevOp &O,
                                               uint32_t nr_grp, uint32_t w0,
                                               uint32_t ws, uint32_t wend,
                                               const GroupSink &sink,
                                               uint32_t *s_list,
                                               const Acct *acct) {
  const uint32_t lane = lane_id();
  uint64_t p[WPL], u[WPL];
  if constexpr (SCAN) {
    static_assert(!CHK && !ACCT && !DIRECT, "the scan words are exact");
    scan_words<WPL>(P, O, w0, ws, wend, p);
#pragma unroll
    for (int k = 0; k < WPL; ++k) u[k] = 0;
  } else {
    eval_words<CHK, ACCT, EAGER, WPL, F>(P, w0 + lane, ws, wend, p, u, acct);
  }
  if (DIRECT) {
    // The ablation without compaction (abla::method1, primitive_abla.cuh:
    // 85-98): each lane walks the set bits of its own words, re-checks an
    // uncertain row and adds the row straight into the sink; the loop
    // diverges across lanes, so there is no wave aggregation.
#pragma unroll
    for (int k = 0; k < WPL; ++k) {
      uint64_t m = p[k];
      const uint32_t row0 = (w0 + ws * k + lane) * 64u;
      while (m) {
        const uint32_t b = (uint32_t)__builtin_ctzll(m);
        m &= m - 1;
        const uint32_t row = row0 + b;
        if (CHK && ((u[k] >> b) & 1ull) && !row_passes<ACCT>(P, O, row, acct))
          continue;
        uint32_t key;
        int64_t val;
        if (row_keyval<ACCT>(O, row, nr_grp, key, val, acct))
          sink_add(sink, key, val, nr_grp);
      }
    }
    return;
  }
  uint32_t cnt = 0;
#pragma unroll
  for (int k = 0; k < WPL; ++k) cnt += __popcll(p[k]);
  uint32_t incl = wave_incl_scan(cnt);
  uint32_t total = __shfl(incl, 63, 64);
  if constexpr (CHK && !ACCT) {
    if (total >= kDenseRowsPerWord * 64u * WPL) {
      // dense three-valued: exact words first, then the row op on the
      // passing rows (the list path, or the dense path if they are dense)
      recheck_words<WPL, F::kRecheck>(P, O, w0, ws, p, u);
      cnt = 0;
#pragma unroll
      for (int k = 0; k < WPL; ++k) cnt += __popcll(p[k]);
      incl = wave_incl_scan(cnt);
      total = __shfl(incl, 63, 64);
    }
  }
  const uint32_t off = incl - cnt;
  if (total >= kDenseRowsPerWord * 64u * WPL) {
    // Dense: word by word (E at a time), lane b takes bit b, so a column's
    // loads for a word's rows are one coalesced 256-B access.  Where the
    // wave's words fit its list space they wait there (read back as
    // broadcasts), so they hold no registers during the row op.
    constexpr int NV = CHK ? 2 : 1;
    constexpr bool kStage = NV * WPL * 64 * 8 <= (int)(kWaveList * 4);
    uint64_t *sw = reinterpret_cast<uint64_t *>(s_list);
    if constexpr (kStage) {
#pragma unroll
      for (int k = 0; k < WPL; ++k) {
        sw[(NV * k) * 64 + lane] = p[k];
        if (CHK) sw[(NV * k + 1) * 64 + lane] = u[k];
      }
      wave_lds_sync();
    }
    constexpr int ED = E;
#pragma unroll
    for (int k = 0; k < WPL; ++k) {
      const uint32_t wk = w0 + ws * k;   // lane s holds word wk + s
      for (uint32_t s = 0; s < 64u; s += ED) {
        uint32_t ent[ED];
        bool valid[ED];
        bool anyv = false;
#pragma unroll
        for (int e = 0; e < ED; ++e) {
          const uint32_t src = s + (uint32_t)e;
          const uint64_t pw = kStage ? sw[(NV * k) * 64 + src]
                                     : __shfl(p[k], (int)src, 64);
          valid[e] = (pw >> lane) & 1ull;
          ent[e] = (wk + src) * 64u + lane;
          if (CHK) {
            const uint64_t uw = kStage ? sw[(NV * k + 1) * 64 + src]
                                       : __shfl(u[k], (int)src, 64);
            if ((uw >> lane) & 1ull) ent[e] |= 0x80000000u;
          }
          anyv = anyv || valid[e];
        }
        if (__ballot(anyv) == 0) continue;
        process_entries<CHK, ACCT, ED>(P, O, nr_grp, ent, valid, sink, acct);
      }
    }
    if (kStage) wave_lds_sync();   // the list space is rewritten next range
    return;
  }
  for (uint32_t base = 0; base < total; base += kWaveList) {
    const uint32_t end = base + kWaveList;
    if (cnt && off < end && off + cnt > base) {
      uint32_t pos = off;
#pragma unroll
      for (int k = 0; k < WPL; ++k) {
        uint64_t m = p[k];
        const uint32_t row0 = (w0 + ws * k + lane) * 64u;
        while (m && pos < end) {
          const uint32_t b = (uint32_t)__builtin_ctzll(m);
          m &= m - 1;
          if (pos >= base) {
            uint32_t e = row0 + b;
            if (CHK && ((u[k] >> b) & 1ull)) e |= 0x80000000u;
            s_list[pos - base] = e;
          }
          ++pos;
        }
      }
    }
    wave_lds_sync();
    const uint32_t nl = min(kWaveList, total - base);
    // wave-uniform trip count: every lane takes part in the sink's wave
    // reductions, also in a partly filled round
    for (uint32_t j0 = 0; j0 < nl; j0 += 64 * E) {
      const uint32_t j = j0 + lane;
      uint32_t ent[E];
      bool valid[E];
#pragma unroll
      for (int e = 0; e < E; ++e) {
        valid[e] = j + 64u * e < nl;
        ent[e] = valid[e] ? s_list[j + 64u * e] : 0u;
      }
      process_entries<CHK, ACCT, E>(P, O, nr_grp, ent, valid, sink, acct);
    }
    wave_lds_sync();
  }
}

// Wave tile wt of 64 x WPL words (the batched kernel's unit).
template <bool CHK, bool ACCT, int WPL, int EAGER>
__device__ __forceinline__ void run_wave_tile(const DevPrg &P, const DevOp &O,
                                              uint32_t nr_grp, uint32_t wt,
                                              const GroupSink &sink,
                                              uint32_t *s_list,
                                              const Acct *acct) {
  const uint32_t w0 = wt * WPL * 64u;
  const uint32_t wend = min(w0 + WPL * 64u, P.nwords);
  run_wave_range<CHK, ACCT, WPL, EAGER>(P, O, nr_grp, w0, 64u, wend, sink,
                                        s_list, acct);
}

}  // namespace mbx
