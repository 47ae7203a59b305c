// The device query functions of merle_amd/csrc/devquery.hpp built for the
// host (one lane, see hip/hip_runtime.h), exported for a CPU check:
//   dh_fused: every word of the program through every evaluation form the
//     kernels use -- early-out (the interpreter of primitive.cuh:200-332,
//     restated), eager with per-column accumulators, straight into
//     registers, and the conjunction form (two- and three-valued) -- which
//     must agree bit for bit (status 97-99 otherwise), then the row op and
//     group-by on the selected rows;
//   dh_scan: the bitmap-free scan (every row, exact predicates);
//   dh_key_div: the key division by an invariant divisor;
//   dh_prg_masks: the re-check's truth table and unloaded columns.
#include "../../merle_amd/csrc/devquery.hpp"
thread_local host_dim3 threadIdx, blockIdx, blockDim, gridDim;
unsigned long long mbx::s_grp_dyn[1 << 16];
using namespace mbx;
static long g_conj = 0, g_chk_conj = 0;
extern "C" {
long dh_conj_words() { return g_conj; }
long dh_chk_conj_words() { return g_chk_conj; }
void *dh_alloc(size_t n) { return malloc(n ? n : 1); }
// The host-built re-check tables of a program: truth table, columns a
// candidate row need not load and their value, conjunction flag.
int dh_prg_masks(const mbx_vprg *prg, uint32_t *out) {
  DevPrg P;
  int st = make_dev_prg(prg, &P);
  if (st) return st;
  out[0] = P.truth;
  out[1] = P.known;
  out[2] = P.known_val;
  out[3] = P.conj;
  return 0;
}
// x / d through the kernels' key division (key_div_magic + key_div)
uint64_t dh_key_div(uint32_t d, uint64_t x) {
  DevKey k{};
  k.div = d;
  key_div_magic(d, &k.magic, &k.sh1, &k.sh2);
  return key_div(k, x);
}
void dh_free(void *p) { free(p); }
int dh_scan(const mbx_vprg *prg, const mbx_rowop *op, uint32_t nr_grp, int64_t *out) {
  DevPrg P;
  DevOp O;
  int st = make_dev_prg(prg, &P);
  if (st) return st;
  st = make_dev_op(op, 0, &O);
  if (st) return st;
  for (uint32_t row = 0; row < P.nrows; ++row) {
    if (!row_passes<false, false>(P, O, row, nullptr)) continue;
    uint32_t key;
    int64_t val;
    if (row_keyval<false>(O, row, nr_grp, key, val, nullptr)) out[key] += val;
  }
  return 0;
}
int dh_fused(const mbx_vprg *prg, const mbx_rowop *op, uint32_t nr_grp, int64_t *out) {
  DevPrg P;
  DevOp O;
  int st = make_dev_prg(prg, &P);
  if (st) return st;
  st = make_dev_op(op, 0, &O);
  if (st) return st;
  GroupSink sink;
  sink.lds = false;
  sink.out = (unsigned long long *)out;
  for (uint32_t w = 0; w < P.nwords; ++w) {
    uint64_t p, u, p2, u2;
    if (P.chk) {
      eval_word<true, false>(P, w, p, u, nullptr);
      eval_word<true, false, 1>(P, w, p2, u2, nullptr);
    } else {
      eval_word<false, false>(P, w, p, u, nullptr);
      eval_word<false, false, 2>(P, w, p2, u2, nullptr);
    }
    if (p != p2 || u != u2) return 99;
    if (P.conj) ++g_conj;
    if (P.conj && P.chk) ++g_chk_conj;
    {
      // four words at stride 64, the last ones past an early end on odd words
      uint64_t a[4], b[4], c[4], d[4], e[4], f[4];
      const uint32_t wend =
          (w & 1) ? (w + 129 < P.nwords ? w + 129 : P.nwords) : P.nwords;
      if (P.chk) {
        eval_words<true, false, 0, 4>(P, w, 64, wend, a, b, nullptr);
        eval_words<true, false, 1, 4>(P, w, 64, wend, c, d, nullptr);
        if (P.conj) eval_words<true, false, 3, 4>(P, w, 64, wend, e, f, nullptr);
      } else {
        eval_words<false, false, 0, 4>(P, w, 64, wend, a, b, nullptr);
        eval_words<false, false, 2, 4>(P, w, 64, wend, c, d, nullptr);
        if (P.conj) eval_words<false, false, 3, 4>(P, w, 64, wend, e, f, nullptr);
      }
      for (int k = 0; k < 4; ++k) {
        uint64_t q = 0, v = 0;
        if (w + 64 * k < wend) {
          if (P.chk) eval_word<true, false>(P, w + 64 * k, q, v, nullptr);
          else eval_word<false, false>(P, w + 64 * k, q, v, nullptr);
        }
        if (a[k] != q || b[k] != v || c[k] != q || d[k] != v) return 98;
        if (P.conj && (e[k] != q || f[k] != v)) return 97;
      }
    }
    while (p) {
      const uint32_t bit = __builtin_ctzll(p);
      p &= p - 1;
      uint32_t e = w * 64 + bit;
      if (P.chk && ((u >> bit) & 1)) e |= 0x80000000u;
      if (P.chk) process_entry<true, false>(P, O, nr_grp, e, sink, nullptr);
      else process_entry<false, false>(P, O, nr_grp, e, sink, nullptr);
    }
  }
  return 0;
}
}
