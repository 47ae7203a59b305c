/*
 * cpuref.c -- CPU oracle for the fused bitmap-index query path.
 *
 * TEST INFRASTRUCTURE ONLY (see cpuref.h): the checker, never the product.
 * Plain C99 + OpenMP over host arrays.  Each function cites the reference
 * file:line it restates.  Parity pins for every function live in
 * tests/test_oracle.py (closed forms, invariants, brute force).
 */
#include "cpuref.h"
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

/* ------------------------------------------------------------------------ */
/* value fetch: bmpidx_create.cu:17-35 (create_bin's fetch lambda)           */
/* ------------------------------------------------------------------------ */
static inline uint64_t load_elem(const void *p, uint32_t eb, size_t i) {
  switch (eb) {
  case 1: return ((const uint8_t *)p)[i];
  case 2: return ((const uint16_t *)p)[i];
  case 4: return ((const uint32_t *)p)[i];
  default: return ((const uint64_t *)p)[i];
  }
}

/* Returns 0 and leaves *v untouched when the foreign key is out of range. */
static inline int fetch(const ref_src *s, size_t row, uint64_t *v) {
  if (s->fk == NULL) {
    *v = load_elem(s->data, s->elem_bytes, row);
    return 1;
  }
  uint32_t k = s->fk[row];
  if (k >= s->dimsz) return 0;
  *v = load_elem(s->data, s->elem_bytes, k);
  return 1;
}

static inline size_t nwords32(uint32_t n) { return ((size_t)n + 31) / 32; }

/* create_bin (bmpidx_create.cu:14-56; kernel primitive.cuh:102-118):
 * value = fetch(row), INT64_MAX when fk >= dimsz (:26-27); bit j set iff
 * mins[j] <= value < maxes[j] (:36-38). */
void ref_create_bin(const ref_src *src, uint32_t n, const uint64_t *mins,
                    const uint64_t *maxs, size_t nb, uint32_t *out,
                    size_t stride_words) {
  const size_t W = nwords32(n);
  for (size_t j = 0; j < nb; ++j)
    memset(out + j * stride_words, 0, stride_words * sizeof(uint32_t));
  #pragma omp parallel for schedule(static)
  for (long w = 0; w < (long)W; ++w) {
    for (uint32_t b = 0; b < 32; ++b) {
      size_t row = (size_t)w * 32 + b;
      if (row >= n) break;
      uint64_t v;
      if (!fetch(src, row, &v)) v = (uint64_t)INT64_MAX;
      for (size_t j = 0; j < nb; ++j)
        if (v >= mins[j] && v < maxs[j])
          out[j * stride_words + w] |= 1u << b;
    }
  }
}

/* create_select / create_join (bmpidx_create.cu:58-136): fk == NULL reads the
 * column directly; otherwise key >= dimsz -> 0 (:99); 32-bit bounds. */
void ref_create_join(const ref_src *src, uint32_t n, uint32_t lo, uint32_t hi,
                     uint32_t *out) {
  const size_t W = nwords32(n);
  #pragma omp parallel for schedule(static)
  for (long w = 0; w < (long)W; ++w) {
    uint32_t word = 0;
    for (uint32_t b = 0; b < 32; ++b) {
      size_t row = (size_t)w * 32 + b;
      if (row >= n) break;
      uint64_t v;
      if (!fetch(src, row, &v)) continue;
      if (v >= lo && v < hi) word |= 1u << b;
    }
    out[w] = word;
  }
}

/* ------------------------------------------------------------------------ */
/* vprg program (primitive.cuh:164-332)                                      */
/* ------------------------------------------------------------------------ */
int ref_vprg_is_nochk(const ref_vprg *p) {
  for (int c = 0; c < 4; ++c)
    if (p->cols[c].leftmost || p->cols[c].rightmost) return 0;
  return 1;
}

/* Memory operand of a column at word w: c = OR of the middle bins up to the
 * first NULL, u = leftmost | rightmost (primitive.cuh:216-226, 294-300).  In
 * nochk mode the reference's `goto` skips the edge bins when a middle slot is
 * NULL; nochk implies there are no edge bins, so both forms agree. */
static inline void memop(const ref_col *c, size_t w, uint32_t *cs,
                         uint32_t *us) {
  uint32_t m = 0, e = 0;
  for (int k = 0; k < 4 && c->middle[k]; ++k) m |= c->middle[k][w];
  if (c->leftmost) e |= c->leftmost[w];
  if (c->rightmost) e |= c->rightmost[w];
  *cs = m;
  *us = e;
}

/* One 32-row word, nochk form (primitive.cuh:201-252). */
static uint32_t eval_nochk(const ref_vprg *p, size_t w) {
  uint32_t r[3] = {0, 0, 0};
  for (int i = 0; i < 6; ++i) {
    const uint8_t *in = p->instrs[i];
    const uint8_t op = in[0], d = in[1], s = in[2];
    if (op == REF_END) break;
    uint32_t cs, us;
    switch (op) {
    case REF_AND: r[d] &= r[s]; break;
    case REF_OR: r[d] |= r[s]; break;
    case REF_NOT: r[d] = ~r[s]; break;
    case REF_ANDM:
      if (r[d] == 0) break; /* early out (:215) */
      memop(&p->cols[s], w, &cs, &us);
      r[d] &= cs;
      break;
    case REF_ORM:
      memop(&p->cols[s], w, &cs, &us);
      r[d] |= cs;
      break;
    default: break;
    }
  }
  return r[0];
}

/* One 32-row word, three-valued form (primitive.cuh:255-332).  P = possible,
 * U = uncertain (certain = P & ~U).  NOT is the corrected three-valued form:
 * P' = ~(P_s & ~U_s) = ~P_s | U_s, U' = U_s; the reference's P' = ~P_s drops
 * uncertain rows (primitive.cuh:283-288), which its demos never exercise. */
static void eval_chk(const ref_vprg *p, size_t w, uint32_t *po, uint32_t *uo) {
  uint32_t P[3] = {0, 0, 0}, U[3] = {0, 0, 0};
  for (int i = 0; i < 6; ++i) {
    const uint8_t *in = p->instrs[i];
    const uint8_t op = in[0], d = in[1], s = in[2];
    if (op == REF_END) break;
    uint32_t cs, us;
    switch (op) {
    case REF_AND: /* :264-269 */
      P[d] &= P[s];
      U[d] = (U[d] | U[s]) & P[d];
      break;
    case REF_OR: { /* :270-282 */
      uint32_t ps = P[s], pp = P[d], usrc = U[s], csrc = ps & ~usrc;
      U[d] = (U[d] & ~csrc) | (usrc & ~pp);
      P[d] = pp | ps;
      break;
    }
    case REF_NOT: { /* :283-288, corrected */
      uint32_t ps = P[s], usrc = U[s];
      P[d] = ~ps | usrc;
      U[d] = usrc;
      break;
    }
    case REF_ANDM: /* :290-309 */
      if (P[d] == 0) break;
      memop(&p->cols[s], w, &cs, &us);
      P[d] &= (us | cs);
      U[d] = (U[d] | us) & P[d];
      break;
    case REF_ORM: /* :311-327 */
      memop(&p->cols[s], w, &cs, &us);
      U[d] = (U[d] & ~cs) | (us & ~P[d]);
      P[d] |= us | cs;
      break;
    default: break;
    }
  }
  *po = P[0];
  *uo = U[0];
}

uint64_t ref_vprg_eval(const ref_vprg *p, int chk, uint32_t *poss,
                       uint32_t *unc) {
  if (chk < 0) chk = !ref_vprg_is_nochk(p);
  const uint32_t n = p->factsz;
  const size_t W = nwords32(n);
  const uint32_t tailmask = (n % 32) ? ((1u << (n % 32)) - 1u) : 0xffffffffu;
  uint64_t total = 0;
  #pragma omp parallel for schedule(static) reduction(+ : total)
  for (long w = 0; w < (long)W; ++w) {
    uint32_t P, U = 0;
    if (chk) eval_chk(p, (size_t)w, &P, &U);
    else P = eval_nochk(p, (size_t)w);
    if ((size_t)w == W - 1) { P &= tailmask; U &= tailmask; }
    poss[w] = P;
    if (unc) unc[w] = U;
    total += (uint64_t)__builtin_popcount(P);
  }
  return total;
}

/* ------------------------------------------------------------------------ */
/* row op (op_t contract primitive.cuh:7-15; ops ssb/lambdas.cuh:19-152)    */
/* ------------------------------------------------------------------------ */
/* Exact predicate program on raw columns: the reference's op re-checks every
 * predicate for an uncertain row (ssb/lambdas.cuh:27-33, synth/synthdemo.cu:
 * 325-362).  Here the check runs the query's own instruction list over the
 * exact predicate booleans, which is the same conjunction/disjunction. */
static int row_passes(const ref_rowop *op, const uint8_t instrs[6][4],
                      size_t row) {
  int e[4] = {0, 0, 0, 0};
  for (uint32_t c = 0; c < op->npreds && c < 4; ++c) {
    uint64_t v;
    e[c] = fetch(&op->preds[c].src, row, &v) && v >= op->preds[c].lo &&
           v < op->preds[c].hi;
  }
  int r[3] = {0, 0, 0};
  for (int i = 0; i < 6; ++i) {
    const uint8_t o = instrs[i][0], d = instrs[i][1], s = instrs[i][2];
    if (o == REF_END) break;
    switch (o) {
    case REF_AND: r[d] = r[d] && r[s]; break;
    case REF_OR: r[d] = r[d] || r[s]; break;
    case REF_NOT: r[d] = !r[s]; break;
    case REF_ANDM: r[d] = r[d] && e[s]; break;
    case REF_ORM: r[d] = r[d] || e[s]; break;
    default: break;
    }
  }
  return r[0];
}

/* Group key and value of a selected row (the `{value, group}` pair of
 * primitive.cuh:10-15).  Returns 0 when the row is dropped: a key fetch with
 * fk >= dimsz, or a key outside [0, nr_grp) (primitive.cuh:385, 433). */
static int row_keyval(const ref_rowop *op, size_t row, uint32_t nr_grp,
                      uint32_t *key, int64_t *val) {
  int64_t k = 0;
  for (uint32_t i = 0; i < op->nkeys && i < 3; ++i) {
    const ref_key *kp = &op->keys[i];
    uint64_t v;
    if (!fetch(&kp->src, row, &v)) return 0;
    uint64_t f = v;
    if (kp->kind == REF_KEY_DIV) f = v / kp->div;
    else if (kp->kind == REF_KEY_BUCKET) {
      f = 0;
      for (uint32_t j = 0; j < kp->nbounds && j < 8; ++j)
        f += kp->bounds[j] <= v;
    }
    k += kp->stride * ((int64_t)f - kp->base);
  }
  if (k < 0 || k >= (int64_t)nr_grp) return 0;
  uint64_t a = 0, b = 0;
  switch (op->measure) {
  case REF_MEAS_COUNT: *val = 1; break;
  case REF_MEAS_COL:
    if (!fetch(&op->ma, row, &a)) return 0;
    *val = (int64_t)a;
    break;
  case REF_MEAS_MUL:
    if (!fetch(&op->ma, row, &a) || !fetch(&op->mb, row, &b)) return 0;
    *val = (int64_t)(a * b);
    break;
  case REF_MEAS_SUB:
    if (!fetch(&op->ma, row, &a) || !fetch(&op->mb, row, &b)) return 0;
    *val = (int64_t)a - (int64_t)b;
    break;
  default: return 0;
  }
  *key = (uint32_t)k;
  return 1;
}

static int set_threads(int nthreads) {
#ifdef _OPENMP
  return nthreads > 0 ? nthreads : omp_get_max_threads();
#else
  (void)nthreads;
  return 1;
#endif
}

uint64_t ref_query_range(const ref_rowop *op, const uint8_t instrs[6][4],
                         uint32_t row0, uint32_t row1, uint32_t nr_grp,
                         int64_t *out, int nthreads) {
  const int nt = set_threads(nthreads);
  memset(out, 0, (size_t)nr_grp * sizeof(int64_t));
  uint64_t total = 0;
  /* per-thread private group arrays, reduced at the end: no atomics */
  int64_t *priv = (int64_t *)calloc((size_t)nt * nr_grp, sizeof(int64_t));
  if (!priv) return UINT64_MAX;
  #pragma omp parallel num_threads(nt) reduction(+ : total)
  {
#ifdef _OPENMP
    int64_t *mine = priv + (size_t)omp_get_thread_num() * nr_grp;
#else
    int64_t *mine = priv;
#endif
    #pragma omp for schedule(static)
    for (long r = (long)row0; r < (long)row1; ++r) {
      if (!row_passes(op, instrs, (size_t)r)) continue;
      uint32_t k;
      int64_t v;
      if (!row_keyval(op, (size_t)r, nr_grp, &k, &v)) continue;
      mine[k] += v;
      ++total;
    }
  }
  for (int t = 0; t < nt; ++t)
    for (uint32_t g = 0; g < nr_grp; ++g) out[g] += priv[(size_t)t * nr_grp + g];
  free(priv);
  return total;
}

uint64_t ref_query(const ref_rowop *op, const uint8_t instrs[6][4], uint32_t n,
                   uint32_t nr_grp, int64_t *out, int nthreads) {
  return ref_query_range(op, instrs, 0, n, nr_grp, out, nthreads);
}

/* vprg_grpby (primitive.cuh:340-445): evaluate the program, then call the op
 * on each possible row with its uncertain flag (:383-384, 431-432); the op
 * checks predicates only for uncertain rows (ssb/lambdas.cuh:27). */
uint64_t ref_vprg_grpby(const ref_vprg *p, const ref_rowop *op,
                        uint32_t nr_grp, int64_t *out) {
  const uint32_t n = p->factsz;
  const size_t W = nwords32(n);
  uint32_t *poss = (uint32_t *)malloc(W * sizeof(uint32_t) + 4);
  uint32_t *unc = (uint32_t *)malloc(W * sizeof(uint32_t) + 4);
  if (!poss || !unc) { free(poss); free(unc); return UINT64_MAX; }
  const int chk = !ref_vprg_is_nochk(p);
  ref_vprg_eval(p, chk, poss, unc);
  const int nt = set_threads(0);
  int64_t *priv = (int64_t *)calloc((size_t)nt * nr_grp, sizeof(int64_t));
  uint64_t total = 0;
  #pragma omp parallel num_threads(nt) reduction(+ : total)
  {
#ifdef _OPENMP
    int64_t *mine = priv + (size_t)omp_get_thread_num() * nr_grp;
#else
    int64_t *mine = priv;
#endif
    #pragma omp for schedule(static)
    for (long w = 0; w < (long)W; ++w) {
      uint32_t m = poss[w];
      while (m) {
        const int b = __builtin_ctz(m);
        m &= m - 1;
        const size_t row = (size_t)w * 32 + (size_t)b;
        if (chk && ((unc[w] >> b) & 1u) && !row_passes(op, p->instrs, row))
          continue;
        uint32_t k;
        int64_t v;
        if (!row_keyval(op, row, nr_grp, &k, &v)) continue;
        mine[k] += v;
        ++total;
      }
    }
  }
  memset(out, 0, (size_t)nr_grp * sizeof(int64_t));
  for (int t = 0; t < nt; ++t)
    for (uint32_t g = 0; g < nr_grp; ++g) out[g] += priv[(size_t)t * nr_grp + g];
  free(priv);
  free(poss);
  free(unc);
  return total;
}

/* ------------------------------------------------------------------------ */
/* SSB flights (ssb/cpuref.c:141-242), int64 sums                            */
/* ------------------------------------------------------------------------ */
/* ssb/ssbdemo.h:10-39: day offsets from 1992-01-01. */
uint32_t ref_ssb_date_to_year(uint16_t date) {
  return date >= 2192 ? 6 : date >= 1827 ? 5 : date >= 1461 ? 4
       : date >= 1096 ? 3 : date >= 731 ? 2 : date >= 366 ? 1 : 0;
}

/* s1ref, ssb/cpuref.c:142-156 */
int64_t ref_s1(const ref_ssb *d, size_t n, uint16_t dateMin, uint16_t dateMax,
               uint8_t discntMin, uint8_t discntMax, uint8_t qtyMin,
               uint8_t qtyMax) {
  int64_t res = 0;
  #pragma omp parallel for reduction(+ : res)
  for (long i = 0; i < (long)n; ++i) {
    if (d->loOrderDate[i] < dateMin || d->loOrderDate[i] >= dateMax) continue;
    if (d->loDiscount[i] < discntMin || d->loDiscount[i] >= discntMax) continue;
    if (d->loQuantity[i] < qtyMin || d->loQuantity[i] >= qtyMax) continue;
    res += (int64_t)d->loExtendedPrice[i] * (int64_t)d->loDiscount[i];
  }
  return res;
}

/* s2ref, ssb/cpuref.c:158-175 (serial: groups are written directly) */
void ref_s2(const ref_ssb *d, size_t n, uint32_t pMfgrMin, uint32_t pMfgrMax,
            uint8_t sCityMin, uint8_t sCityMax, int64_t *out) {
  for (size_t i = 0; i < n; ++i) {
    if (d->loSuppCity[i] < sCityMin || d->loSuppCity[i] >= sCityMax) continue;
    uint16_t mfgr = d->partMfgr[d->loPartKey[i]];
    if (mfgr < pMfgrMin || mfgr >= pMfgrMax) continue;
    uint32_t year = ref_ssb_date_to_year(d->loOrderDate[i]);
    uint32_t grp = (mfgr - pMfgrMin) + year * (pMfgrMax - pMfgrMin);
    out[grp] += (int64_t)d->loRevenue[i];
  }
}

/* s3ref, ssb/cpuref.c:177-206 */
void ref_s3(const ref_ssb *d, size_t n, uint8_t cCityMin, uint8_t cCityMax,
            uint8_t sCityMin, uint8_t sCityMax, uint16_t dateMin,
            uint16_t dateMax, int64_t *out) {
  for (size_t i = 0; i < n; ++i) {
    uint16_t sCity = d->loSuppCity[i];
    if (sCity < sCityMin || sCity >= sCityMax) continue;
    uint16_t date = d->loOrderDate[i];
    if (date < dateMin || date >= dateMax) continue;
    uint16_t cCity = d->custCity[d->loCustKey[i]];
    if (cCity < cCityMin || cCity >= cCityMax) continue;
    uint8_t cMinS = cCityMin, cMaxS = cCityMax, sMinS = sCityMin, sMaxS = sCityMax;
    if (cCityMax - cCityMin >= 50) {
      cCity /= 10; cMinS /= 10; cMaxS /= 10;
      sCity /= 10; sMinS /= 10; sMaxS /= 10;
    }
    const uint32_t year = ref_ssb_date_to_year(date);
    const uint32_t yearMin = ref_ssb_date_to_year(dateMin);
    const size_t a = (size_t)(cMaxS - cMinS), b = (size_t)(sMaxS - sMinS);
    const size_t grp = a * b * (year - yearMin) + a * (size_t)(cCity - cMinS) +
                       (size_t)(sCity - sMinS);
    out[grp] += (int64_t)d->loRevenue[i];
  }
}

/* s4ref, ssb/cpuref.c:208-242; revenue - supplycost is signed here. */
void ref_s4(const ref_ssb *d, size_t n, uint8_t cCityMin, uint8_t cCityMax,
            uint8_t sCityMin, uint8_t sCityMax, uint16_t pMfgrMin,
            uint16_t pMfgrMax, uint16_t dateMin, uint16_t dateMax,
            int64_t *out) {
  for (size_t i = 0; i < n; ++i) {
    uint16_t sCity = d->loSuppCity[i];
    if (sCity < sCityMin || sCity >= sCityMax) continue;
    uint16_t date = d->loOrderDate[i];
    if (date < dateMin || date >= dateMax) continue;
    uint16_t cCity = d->custCity[d->loCustKey[i]];
    if (cCity < cCityMin || cCity >= cCityMax) continue;
    uint16_t pMfgr = d->partMfgr[d->loPartKey[i]];
    if (pMfgr < pMfgrMin || pMfgr >= pMfgrMax) continue;
    uint8_t sMinS = sCityMin, sMaxS = sCityMax;
    if (sCityMax - sCityMin >= 50) { sCity /= 10; sMinS /= 10; sMaxS /= 10; }
    uint16_t pMinS = pMfgrMin, pMaxS = pMfgrMax;
    if (pMfgrMax - pMfgrMin >= 200) { pMfgr /= 40; pMinS /= 40; pMaxS /= 40; }
    const uint32_t year = ref_ssb_date_to_year(date);
    const uint32_t yearMin = ref_ssb_date_to_year(dateMin);
    const size_t a = (size_t)(sMaxS - sMinS), b = (size_t)(pMaxS - pMinS);
    const size_t grp = a * b * (year - yearMin) + a * (size_t)(sCity - sMinS) +
                       (size_t)(pMfgr - pMinS);
    out[grp] += (int64_t)d->loRevenue[i] - (int64_t)d->loSupplyCost[i];
  }
}

/* ------------------------------------------------------------------------ */
/* synth_ref (synth/cpuref.c:125-169): COUNT grouped by dimattr2[fk]         */
/* ------------------------------------------------------------------------ */
void ref_synth(const void *fa1, const void *fa2, uint32_t bitwidth,
               const uint32_t *fk, const uint32_t *dimattr1,
               const uint8_t *dimattr2, size_t n, uint64_t fa1lo,
               uint64_t fa1hi, uint64_t fa2lo, uint64_t fa2hi, uint64_t da1lo,
               uint64_t da1hi, int64_t grpout[256]) {
  memset(grpout, 0, 256 * sizeof(int64_t));
  const uint32_t eb = bitwidth / 8;
  for (size_t i = 0; i < n; ++i) {
    uint64_t a1 = load_elem(fa1, eb, i), a2 = load_elem(fa2, eb, i);
    if (!((a1 >= fa1lo && a1 < fa1hi) || (a2 >= fa2lo && a2 < fa2hi))) continue;
    uint32_t k = fk[i];
    if (dimattr1[k] < da1lo || dimattr1[k] >= da1hi) continue;
    grpout[dimattr2[k]]++;
  }
}

/* ------------------------------------------------------------------------ */
/* cost-model MLP: fp32 parameters, double accumulation                      */
/* ------------------------------------------------------------------------ */
double ref_mlp_loss_grad(const float *x, const float *t, int n, int din,
                         int dh, int dout, const float *W1, const float *b1,
                         const float *W2, const float *b2, float *gW1,
                         float *gb1, float *gW2, float *gb2, float *y) {
  double *pre = (double *)malloc(sizeof(double) * (size_t)n * dh);
  double *dy = (double *)malloc(sizeof(double) * (size_t)n * dout);
  double *gw1 = (double *)calloc((size_t)dh * din, sizeof(double));
  double *gbb1 = (double *)calloc((size_t)dh, sizeof(double));
  double *gw2 = (double *)calloc((size_t)dout * dh, sizeof(double));
  double *gbb2 = (double *)calloc((size_t)dout, sizeof(double));
  double loss = 0.0;
  const double scale = 1.0 / ((double)n * (double)dout);
  for (int i = 0; i < n; ++i) {
    for (int h = 0; h < dh; ++h) {
      double s = b1[h];
      for (int k = 0; k < din; ++k) s += (double)W1[h * din + k] * x[i * din + k];
      pre[(size_t)i * dh + h] = s;
    }
    for (int o = 0; o < dout; ++o) {
      double s = b2[o];
      for (int h = 0; h < dh; ++h) {
        double a = pre[(size_t)i * dh + h];
        s += (double)W2[o * dh + h] * (a > 0 ? a : 0);
      }
      if (y) y[i * dout + o] = (float)s;
      double e = s - t[i * dout + o];
      loss += e * e;
      dy[(size_t)i * dout + o] = 2.0 * e * scale;
    }
  }
  for (int i = 0; i < n; ++i) {
    for (int o = 0; o < dout; ++o) {
      double g = dy[(size_t)i * dout + o];
      gbb2[o] += g;
      for (int h = 0; h < dh; ++h) {
        double a = pre[(size_t)i * dh + h];
        gw2[o * dh + h] += g * (a > 0 ? a : 0);
      }
    }
    for (int h = 0; h < dh; ++h) {
      if (pre[(size_t)i * dh + h] <= 0) continue;
      double g = 0;
      for (int o = 0; o < dout; ++o) g += dy[(size_t)i * dout + o] * W2[o * dh + h];
      gbb1[h] += g;
      for (int k = 0; k < din; ++k) gw1[h * din + k] += g * x[i * din + k];
    }
  }
  if (gW1) for (int i = 0; i < dh * din; ++i) gW1[i] = (float)gw1[i];
  if (gb1) for (int i = 0; i < dh; ++i) gb1[i] = (float)gbb1[i];
  if (gW2) for (int i = 0; i < dout * dh; ++i) gW2[i] = (float)gw2[i];
  if (gb2) for (int i = 0; i < dout; ++i) gb2[i] = (float)gbb2[i];
  free(pre); free(dy); free(gw1); free(gbb1); free(gw2); free(gbb2);
  return loss * scale;
}

/* ------------------------------------------------------------------------ */
/* WAH compressed-bitmap baseline                                           */
/* ------------------------------------------------------------------------ */
#define WAH_FILL 0x80000000u
#define WAH_ONE 0x40000000u
#define WAH_CNT 0x3fffffffu
#define WAH_ROWS 0x7fffffffu

static inline size_t wah_ngroups(uint32_t n) { return ((size_t)n + 30) / 31; }

/* dbjoinFlatWah (wahGpuCmprs.cu:247-320): lane 30-j of a warp tests row
 * base+j, so the ballot puts the group's first row at bit 30. */
void ref_wah_select(const ref_src *src, uint32_t n, uint64_t lo, uint64_t hi,
                    uint32_t *groups) {
  const size_t G = wah_ngroups(n);
  #pragma omp parallel for schedule(static)
  for (long g = 0; g < (long)G; ++g) {
    uint32_t w = 0;
    for (uint32_t j = 0; j < 31; ++j) {
      const size_t row = (size_t)g * 31 + j;
      if (row >= n) break;
      uint64_t v;
      if (fetch(src, row, &v) && v >= lo && v < hi) w |= 1u << (30 - j);
    }
    groups[g] = w;
  }
}

void ref_wah_groups(const uint32_t *bitmap, uint32_t n, uint32_t *groups) {
  const size_t G = wah_ngroups(n);
  #pragma omp parallel for schedule(static)
  for (long g = 0; g < (long)G; ++g) {
    uint32_t w = 0;
    for (uint32_t j = 0; j < 31; ++j) {
      const size_t row = (size_t)g * 31 + j;
      if (row >= n) break;
      if ((bitmap[row >> 5] >> (row & 31)) & 1u) w |= 1u << (30 - j);
    }
    groups[g] = w;
  }
}

/* wahCompress (wahGpuCmprs.cu:215-245): a group is kept when it is a
 * literal, or the last of its fill run; the last group only when it is not
 * zero (so a trailing zero run leaves nothing).  A kept fill carries the
 * groups since the previous kept group. */
size_t ref_wah_compress(const uint32_t *groups, size_t ng, uint32_t *out) {
  size_t k = 0, prev_end = 0;
  for (size_t i = 0; i < ng; ++i) {
    const uint32_t v = groups[i] & WAH_ROWS;
    const int fill = v == 0 || v == WAH_ROWS;
    int keep;
    if (i + 1 == ng) keep = v != 0;
    else keep = !fill || (groups[i + 1] & WAH_ROWS) != v;
    if (!keep) continue;
    if (fill)
      out[k++] = WAH_FILL | (v ? WAH_ONE : 0u) | (uint32_t)(i + 1 - prev_end);
    else
      out[k++] = v;
    prev_end = i + 1;
  }
  return k;
}

static inline uint32_t wah_count(uint32_t w) { return (w & WAH_FILL) ? (w & WAH_CNT) : 1u; }
/* the 31 rows a word stands for in each of its groups */
static inline uint32_t wah_rows(uint32_t w) {
  if (w & WAH_FILL) return (w & WAH_ONE) ? WAH_ROWS : 0u;
  return w & WAH_ROWS;
}

void ref_wah_scan(const uint32_t *wah, size_t n, uint32_t *excl) {
  uint32_t acc = 0;
  for (size_t i = 0; i < n; ++i) {
    excl[i] = acc;
    acc += wah_count(wah[i]);
  }
  excl[n] = acc;
}

/* the word of (wah, excl, n) holding group g; n when g is past the end */
static size_t wah_find(const uint32_t *excl, size_t n, uint32_t g) {
  if (n == 0 || g >= excl[n]) return n;
  size_t lo = 0, hi = n;   /* excl[lo] <= g < excl[hi] */
  while (hi - lo > 1) {
    const size_t mid = lo + (hi - lo) / 2;
    if (excl[mid] <= g) lo = mid;
    else hi = mid;
  }
  return lo;
}

/* wahAndNo1 (wahGpuDirect.cu:8-34) with the up word taken as the one that
 * covers the down word's group.  The reference finds it by a lower-bound
 * search of inclusive scans (its parameters upIncScan / downIncScan), but
 * its SSB driver passes the exclusive scans of wahCntExcScan
 * (ssb/wah-demo.cu:48-61), so there the search names the word after a
 * covering fill that starts before the group. */
void ref_wah_and_no1(const uint32_t *up, const uint32_t *up_excl, size_t nup,
                     const uint32_t *down, const uint32_t *down_excl,
                     size_t ndown, uint32_t *out) {
  #pragma omp parallel for schedule(static)
  for (long d = 0; d < (long)ndown; ++d) {
    uint32_t dw = down[d];
    if (dw == (WAH_FILL | WAH_ONE | 1u)) dw = WAH_ROWS;
    uint32_t o;
    if (dw & WAH_FILL) {
      o = dw;
    } else {
      const size_t u = wah_find(up_excl, nup, down_excl[d]);
      o = dw & (u < nup ? wah_rows(up[u]) : 0u);
      if (o == 0) o = WAH_FILL | 1u;
    }
    out[d] = o;
  }
}

void ref_wah_decompress(const uint32_t *wah, const uint32_t *excl, size_t n,
                        uint32_t *groups, size_t ng) {
  size_t g = 0;
  for (size_t i = 0; i < n && g < ng; ++i) {
    const uint32_t c = wah_count(wah[i]), r = wah_rows(wah[i]);
    for (uint32_t k = 0; k < c && g < ng; ++k) groups[g++] = r;
  }
  (void)excl;
  for (; g < ng; ++g) groups[g] = 0;
}

uint64_t ref_wah_longest_1fill(const uint32_t *wah, size_t n) {
  uint64_t best = 0, cur = 0;
  for (size_t i = 0; i < n; ++i) {
    uint32_t w = wah[i];
    if (w == WAH_ROWS) w = WAH_FILL | WAH_ONE | 1u;
    if ((w >> 30) == 3u) {
      cur += w & WAH_CNT;
      if (cur > best) best = cur;
    } else {
      cur = 0;
    }
  }
  return best;
}

