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
constexpr int kAcctMax = 48;
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

template <bool ACCT>
__device__ __forceinline__ uint64_t ld_bmp(const uint64_t *p, uint32_t w,
                                           const Acct *acct) {
  mark<ACCT>(acct, p + w);
  return ld_stream(p + w);
}

// Memory operand of one column at word w: c = OR of middle bins,
// u = leftmost | rightmost (primitive.cuh:216-226, 294-300).
template <bool ACCT>
__device__ __forceinline__ uint64_t memop_c(const DevCol &c, uint32_t w,
                                            const Acct *acct) {
  uint64_t m = 0;
  for (uint32_t k = 0; k < c.nm; ++k) m |= ld_bmp<ACCT>(c.m[k], w, acct);
  return m;
}
template <bool ACCT>
__device__ __forceinline__ uint64_t memop_u(const DevCol &c, uint32_t w,
                                            const Acct *acct) {
  uint64_t e = 0;
  if (c.l) e |= ld_bmp<ACCT>(c.l, w, acct);
  if (c.r) e |= ld_bmp<ACCT>(c.r, w, acct);
  return e;
}

// Evaluates the program on 64-row word w.  nochk form: primitive.cuh:201-252;
// chk form: primitive.cuh:255-332, with the three-valued NOT corrected to
// P' = ~P_s | U_s, U' = U_s (the reference's P' = ~P_s drops uncertain rows).
// ANDM skips its loads for lanes whose destination is already empty, the
// early-out of primitive.cuh:215/292: a 128-byte line of the operand is
// fetched only if one of its 16 words is still live.
template <bool CHK, bool ACCT>
__device__ __forceinline__ void eval_word(const DevPrg &P, uint32_t w,
                                          uint64_t &po, uint64_t &uo,
                                          const Acct *acct) {
  uint64_t p0 = 0, p1 = 0, p2 = 0, u0 = 0, u1 = 0, u2 = 0;
#pragma unroll
  for (uint32_t i = 0; i < MBX_MAXNINSTR; ++i) {
    if (i >= P.nins) break;
    const uint32_t op = P.ins[i][0], d = P.ins[i][1], s = P.ins[i][2];
    const uint64_t pd = rget(p0, p1, p2, d);
    if (op == MBX_AND) {
      const uint64_t np = pd & rget(p0, p1, p2, s);
      rset(p0, p1, p2, d, np);
      if (CHK)
        rset(u0, u1, u2, d,
             (rget(u0, u1, u2, d) | rget(u0, u1, u2, s)) & np);
    } else if (op == MBX_OR) {
      const uint64_t ps = rget(p0, p1, p2, s);
      if (CHK) {
        const uint64_t us = rget(u0, u1, u2, s), cs = ps & ~us;
        rset(u0, u1, u2, d, (rget(u0, u1, u2, d) & ~cs) | (us & ~pd));
      }
      rset(p0, p1, p2, d, pd | ps);
    } else if (op == MBX_NOT) {
      const uint64_t ps = rget(p0, p1, p2, s);
      if (CHK) {
        const uint64_t us = rget(u0, u1, u2, s);
        rset(p0, p1, p2, d, ~ps | us);
        rset(u0, u1, u2, d, us);
      } else {
        rset(p0, p1, p2, d, ~ps);
      }
    } else if (op == MBX_ANDM) {
      const DevCol &c = P.cols[s];
      if (pd != 0) {
        if (CHK) {
          const uint64_t us = memop_u<ACCT>(c, w, acct);
          const uint64_t cs = memop_c<ACCT>(c, w, acct);
          const uint64_t np = pd & (us | cs);
          rset(p0, p1, p2, d, np);
          rset(u0, u1, u2, d, (rget(u0, u1, u2, d) | us) & np);
        } else {
          rset(p0, p1, p2, d, pd & memop_c<ACCT>(c, w, acct));
        }
      }
    } else if (op == MBX_ORM) {
      const DevCol &c = P.cols[s];
      const uint64_t cs = memop_c<ACCT>(c, w, acct);
      if (CHK) {
        const uint64_t us = memop_u<ACCT>(c, w, acct);
        rset(u0, u1, u2, d, (rget(u0, u1, u2, d) & ~cs) | (us & ~pd));
        rset(p0, p1, p2, d, pd | us | cs);
      } else {
        rset(p0, p1, p2, d, pd | cs);
      }
    }
  }
  // rows >= N never survive (the reference has no such guard, primitive.cuh
  // :357-360 only bounds the word index)
  if (w == P.nwords - 1 && (P.nrows & 63u)) {
    const uint64_t tail = (1ull << (P.nrows & 63u)) - 1ull;
    p0 &= tail;
    u0 &= tail;
  }
  po = p0;
  uo = CHK ? u0 : 0;
}

// ----------------------------------------------------------- row operation
// Re-check of an uncertain row: the program's instruction list over the
// exact predicates of its columns (the op's chk branch, ssb/lambdas.cuh:27-33).
template <bool ACCT>
__device__ __forceinline__ bool row_passes(const DevPrg &P, const DevOp &O,
                                        uint32_t row, const Acct *acct) {
  bool e0 = false, e1 = false, e2 = false, e3 = false;
  for (uint32_t c = 0; c < O.npreds; ++c) {
    uint64_t v;
    const bool ok = fetch_a<ACCT>(O.pred[c], row, &v, acct) && v >= O.lo[c] &&
                    v < O.hi[c];
    if (c == 0) e0 = ok;
    else if (c == 1) e1 = ok;
    else if (c == 2) e2 = ok;
    else e3 = ok;
  }
  bool r0 = false, r1 = false, r2 = false;
  for (uint32_t i = 0; i < P.nins; ++i) {
    const uint32_t op = P.ins[i][0], d = P.ins[i][1], s = P.ins[i][2];
    const bool rd = d == 0 ? r0 : (d == 1 ? r1 : r2);
    bool v;
    if (op == MBX_AND || op == MBX_OR || op == MBX_NOT) {
      const bool rs = s == 0 ? r0 : (s == 1 ? r1 : r2);
      v = op == MBX_AND ? (rd && rs) : (op == MBX_OR ? (rd || rs) : !rs);
    } else {
      const bool es = s == 0 ? e0 : (s == 1 ? e1 : (s == 2 ? e2 : e3));
      v = op == MBX_ANDM ? (rd && es) : (rd || es);
    }
    if (d == 0) r0 = v;
    else if (d == 1) r1 = v;
    else r2 = v;
  }
  return r0;
}

// {value, group} of a selected row (primitive.cuh:10-15).  false drops it.
template <bool ACCT>
__device__ __forceinline__ bool row_keyval(const DevOp &O, uint32_t row,
                                           uint32_t nr_grp, uint32_t &key,
                                           int64_t &val, const Acct *acct) {
  int64_t k = 0;
  for (uint32_t i = 0; i < O.nkeys; ++i) {
    const DevKey &kp = O.keys[i];
    uint64_t v;
    if (!fetch_a<ACCT>(kp.src, row, &v, acct)) return false;
    uint64_t f = v;
    if (kp.kind == MBX_KEY_DIV) {
      f = v / kp.div;
    } else if (kp.kind == MBX_KEY_BUCKET) {
      f = 0;
      for (uint32_t j = 0; j < kp.nb; ++j) f += kp.bounds[j] <= v;
    }
    k += kp.stride * ((int64_t)f - kp.base);
  }
  if (k < 0 || k >= (int64_t)nr_grp) return false;
  key = (uint32_t)k;
  // measure: 1, a, a*b or a-b; one exit for every kind
  uint64_t a = 1, b = 0;
  bool ok = true;
  if (O.measure != MBX_MEAS_COUNT) ok = fetch_a<ACCT>(O.ma, row, &a, acct);
  if (ok && O.measure >= MBX_MEAS_MUL) ok = fetch_a<ACCT>(O.mb, row, &b, acct);
  val = O.measure == MBX_MEAS_MUL   ? (int64_t)(a * b)
        : O.measure == MBX_MEAS_SUB ? (int64_t)a - (int64_t)b
                                    : (int64_t)a;
  return ok;
}

// Aggregation target: an LDS table (small group counts) or the output in
// HBM (global int64 atomics).
struct GroupSink {
  unsigned long long *lds;   // non-null: LDS table of nr_grp slots
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
  if (s.lds) atomicAdd(s.lds + key, (unsigned long long)val);
  else atomicAdd(s.out + key, (unsigned long long)val);
}

// One list entry: row id, MSB = uncertain (primitive.cuh:406-407).
template <bool CHK, bool ACCT>
__device__ __forceinline__ void process_entry(const DevPrg &P, const DevOp &O,
                                              uint32_t nr_grp, uint32_t e,
                                              const GroupSink &sink,
                                              const Acct *acct) {
  const uint32_t row = e & 0x7fffffffu;
  if (CHK && (e >> 31) && !row_passes<ACCT>(P, O, row, acct)) return;
  uint32_t key;
  int64_t val;
  if (!row_keyval<ACCT>(O, row, nr_grp, key, val, acct)) return;
  if (ACCT) {
    atomicAdd(acct->selected, 1ull);
    return;
  }
  sink_add(sink, key, val, nr_grp);
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

// ------------------------------------------------------------ tile pipeline
// One tile = 256 threads x WPL 64-row words.  Program -> popcount -> block
// scan -> LDS row list (rounds of kListCap entries) -> row op -> sink.
// The reference's method 1/method 2 switch (primitive.cuh:373-397) becomes
// rounds: a dense tile simply takes more rounds, every round using all lanes.
template <bool CHK, bool ACCT, int WPL>
__device__ __forceinline__ void run_tile(const DevPrg &P, const DevOp &O,
                                         uint32_t nr_grp, uint32_t tile,
                                         const GroupSink &sink,
                                         uint32_t *s_list, uint32_t *s_wave,
                                         const Acct *acct) {
  const uint32_t tid = threadIdx.x;
  uint64_t p[WPL], u[WPL];
  uint32_t cnt = 0;
#pragma unroll
  for (int k = 0; k < WPL; ++k) {
    const uint32_t w = (tile * WPL + k) * 256 + tid;
    p[k] = 0;
    u[k] = 0;
    if (w < P.nwords) eval_word<CHK, ACCT>(P, w, p[k], u[k], acct);
    cnt += __popcll(p[k]);
  }
  uint32_t total;
  const uint32_t off = block_excl_scan(cnt, s_wave, &total);
  for (uint32_t base = 0; base < total; base += kListCap) {
    const uint32_t end = base + kListCap;
    if (cnt && off < end && off + cnt > base) {
      uint32_t pos = off;
#pragma unroll
      for (int k = 0; k < WPL; ++k) {
        uint64_t m = p[k];
        const uint32_t row0 = ((tile * WPL + k) * 256 + tid) * 64;
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
    __syncthreads();
    const uint32_t nl = min(kListCap, total - base);
    for (uint32_t j = tid; j < nl; j += 256)
      process_entry<CHK, ACCT>(P, O, nr_grp, s_list[j], sink, acct);
    __syncthreads();
  }
}

}  // namespace mbx
