/*
 * cpuref.h -- CPU oracle for the fused bitmap-index query path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker: only tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg may load it.  The
 * product (merle_amd/, libmbx.so) never links, loads or calls it.
 *
 * It restates, in plain C99 + OpenMP over host arrays, the reference's
 * algorithms for this path:
 *   - bitmap creation            bmpidx_create.cu:14-136, primitive.cuh:88-118
 *   - vprg program evaluation    primitive.cuh:164-332 (nochk and chk)
 *   - bitmap-driven group-by     primitive.cuh:340-445 (row set, op contract)
 *   - raw-column query answers   ssb/cpuref.c:141-242, synth/cpuref.c:125-169
 * with two deliberate fixes recorded beside the code: rows >= N are never
 * selected (the reference leaves tail bits after NOT), and NOT over a
 * three-valued register keeps uncertain rows possible.
 *
 * Layout facts shared with the product: a bitmap is an array of uint32 words,
 * bit b of word k is row 32k+b (LSB first); ranges are half-open [lo, hi);
 * a foreign key >= dimsz selects nothing; sums are int64.
 */
#ifndef MBX_ORACLE_CPUREF_H
#define MBX_ORACLE_CPUREF_H
#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* A value source: fact column (fk == NULL) or dimension attribute via fk. */
typedef struct ref_src {
  const void *data;
  const uint32_t *fk;
  uint32_t dimsz;
  uint32_t elem_bytes; /* 1, 2, 4, 8 */
} ref_src;

/* ---- bitmap creation (bmpidx_create.cu:14-56 create_bin) ---------------- */
/* out holds nb bitmaps, bitmap j at out + j*stride_words; words past
 * ceil(n/32) up to stride_words are written as 0. */
void ref_create_bin(const ref_src *src, uint32_t n, const uint64_t *mins,
                    const uint64_t *maxs, size_t nb, uint32_t *out,
                    size_t stride_words);
/* create_select / create_join (bmpidx_create.cu:58-136): one bitmap. */
void ref_create_join(const ref_src *src, uint32_t n, uint32_t lo, uint32_t hi,
                     uint32_t *out);

/* ---- vprg program (primitive.cuh:164-198) -------------------------------- */
enum { REF_END = 0, REF_AND, REF_OR, REF_NOT, REF_ANDM, REF_ORM };
typedef struct ref_col {
  const uint32_t *middle[4];
  const uint32_t *leftmost, *rightmost;
} ref_col;
typedef struct ref_vprg {
  ref_col cols[4];
  uint32_t factsz;
  uint8_t instrs[6][4]; /* opcode, dstreg, src1, src2 */
} ref_vprg;

/* 1 if no column has edge bins (primitive.cuh:190-195). */
int ref_vprg_is_nochk(const ref_vprg *p);
/* Word-wise program evaluation (primitive.cuh:200-332).  Writes the possible
 * bitmap (and, in chk mode, the uncertain bitmap; unc may be NULL in nochk
 * mode).  chk < 0 picks the mode from ref_vprg_is_nochk. Rows >= factsz are
 * cleared. Returns the number of possible rows. */
uint64_t ref_vprg_eval(const ref_vprg *p, int chk, uint32_t *poss,
                       uint32_t *unc);

/* ---- row operation (the op_t contract, primitive.cuh:7-15) --------------- */
enum { REF_KEY_IDENT = 0, REF_KEY_DIV = 1, REF_KEY_BUCKET = 2 };
enum { REF_MEAS_COUNT = 0, REF_MEAS_COL = 1, REF_MEAS_MUL = 2,
       REF_MEAS_SUB = 3 };
typedef struct ref_pred {
  ref_src src;
  uint64_t lo, hi;
} ref_pred;
typedef struct ref_key {
  ref_src src;
  uint32_t kind;    /* REF_KEY_* */
  uint32_t div;     /* REF_KEY_DIV: v / div */
  int64_t base;     /* subtracted after the transform */
  int64_t stride;   /* multiplier of this key part */
  uint32_t nbounds; /* REF_KEY_BUCKET: f(v) = #{j : bounds[j] <= v} */
  uint64_t bounds[8];
} ref_key;
typedef struct ref_rowop {
  ref_pred preds[4]; /* exact predicate behind vprg column i */
  uint32_t npreds;
  ref_key keys[3];
  uint32_t nkeys;
  uint32_t measure; /* REF_MEAS_* */
  ref_src ma, mb;
} ref_rowop;

/* Raw-column answer: every row is checked by running the instruction list of
 * `instrs` over exact predicate booleans (column i <-> preds[i]); selected rows
 * add their measure to out[key]; keys outside [0, nr_grp) are dropped.  out is
 * zeroed first.  nthreads <= 0 uses the OpenMP default. Returns selected rows. */
uint64_t ref_query(const ref_rowop *op, const uint8_t instrs[6][4], uint32_t n,
                   uint32_t nr_grp, int64_t *out, int nthreads);
/* Same, restricted to rows [row0, row1): used for bounded CPU timing. */
uint64_t ref_query_range(const ref_rowop *op, const uint8_t instrs[6][4],
                         uint32_t row0, uint32_t row1, uint32_t nr_grp,
                         int64_t *out, int nthreads);

/* Bitmap-driven answer (vprg_grpby, primitive.cuh:340-445): program over the
 * bitmaps, then the op on every possible row; uncertain rows are re-checked
 * against all predicates. out is zeroed first. Returns rows aggregated. */
uint64_t ref_vprg_grpby(const ref_vprg *p, const ref_rowop *op,
                        uint32_t nr_grp, int64_t *out);

/* ---- SSB flights, restated as straight loops (ssb/cpuref.c:141-242) ------ */
typedef struct ref_ssb {
  const uint32_t *loCustKey, *loPartKey;
  const uint16_t *loOrderDate;
  const uint32_t *loExtendedPrice, *loRevenue, *loSupplyCost;
  const uint8_t *loQuantity, *loDiscount, *loSuppCity;
  const uint8_t *custCity;
  const uint16_t *partMfgr;
} ref_ssb;
uint32_t ref_ssb_date_to_year(uint16_t date); /* ssb/ssbdemo.h:29-39 */
int64_t ref_s1(const ref_ssb *d, size_t n, uint16_t dateMin, uint16_t dateMax,
               uint8_t discntMin, uint8_t discntMax, uint8_t qtyMin,
               uint8_t qtyMax);
void ref_s2(const ref_ssb *d, size_t n, uint32_t pMfgrMin, uint32_t pMfgrMax,
            uint8_t sCityMin, uint8_t sCityMax, int64_t *out);
void ref_s3(const ref_ssb *d, size_t n, uint8_t cCityMin, uint8_t cCityMax,
            uint8_t sCityMin, uint8_t sCityMax, uint16_t dateMin,
            uint16_t dateMax, int64_t *out);
void ref_s4(const ref_ssb *d, size_t n, uint8_t cCityMin, uint8_t cCityMax,
            uint8_t sCityMin, uint8_t sCityMax, uint16_t pMfgrMin,
            uint16_t pMfgrMax, uint16_t dateMin, uint16_t dateMax,
            int64_t *out);

/* ---- synthetic suite (synth/cpuref.c:125-169) ----------------------------- */
void ref_synth(const void *fa1, const void *fa2, uint32_t bitwidth,
               const uint32_t *fk, const uint32_t *dimattr1,
               const uint8_t *dimattr2, size_t n, uint64_t fa1lo,
               uint64_t fa1hi, uint64_t fa2lo, uint64_t fa2hi, uint64_t da1lo,
               uint64_t da1hi, int64_t grpout[256]);

/* ---- WAH compressed-bitmap baseline (merle/wahGpuCmprs.cu,
 *      merle/wahGpuDirect.cu, merle/host/wahConv.c) ------------------------ */
/* Group layout (dbjoinFlatWah, wahGpuCmprs.cu:247-320): 31 rows per uint32,
 * row 31g+j at bit 30-j of group g, bit 31 zero; ceil(n/31) groups, rows
 * >= n zero (the reference makes floor(n/31) groups and drops the tail). */
void ref_wah_select(const ref_src *src, uint32_t n, uint64_t lo, uint64_t hi,
                    uint32_t *groups);
/* The same groups from a bitmap of this layout family (uint32 words, bit b
 * of word k = row 32k+b). */
void ref_wah_groups(const uint32_t *bitmap, uint32_t n, uint32_t *groups);
/* WAH words: bit 31 set = fill (bit 30 = fill bit, bits 0-29 = groups),
 * else a literal of 31 rows.  wahCompress (wahGpuCmprs.cu:215-245): each
 * maximal run of all-zero or all-one groups -> one fill, every other group a
 * literal, a trailing zero run dropped.  Returns the number of words. */
size_t ref_wah_compress(const uint32_t *groups, size_t ng, uint32_t *out);
/* wahCntExcScan (wahGpuCmprs.cu:6-14): excl[i] = groups before word i,
 * excl[n] = all groups. */
void ref_wah_scan(const uint32_t *wah, size_t n, uint32_t *excl);
/* wahAndNo1 (wahGpuDirect.cu:8-34): the AND in `down`'s word structure: a
 * 0-fill of down stays, a literal (or 1-fill of one group) is ANDed with
 * the up word covering its group (groups past up read 0), a zero result
 * becomes a 0-fill of one group.  down must hold no 1-fill longer than one
 * group.  out has ndown words. */
void ref_wah_and_no1(const uint32_t *up, const uint32_t *up_excl, size_t nup,
                     const uint32_t *down, const uint32_t *down_excl,
                     size_t ndown, uint32_t *out);
/* wahDecomp (wahGpuCmprs.cu:37-51): the groups of the words, then zero
 * groups up to ng. */
void ref_wah_decompress(const uint32_t *wah, const uint32_t *excl, size_t n,
                        uint32_t *groups, size_t ng);
/* wahLongest1Fill (merle/host/wahConv.c:6-20): the longest run of one
 * groups held by consecutive 1-fill words (a 0x7fffffff literal counts as
 * a 1-fill of one group). */
uint64_t ref_wah_longest_1fill(const uint32_t *wah, size_t n);

/* ---- cost-model MLP (fp32 forward/backward, double accumulation) --------- */
/* y = W2 relu(W1 x + b1) + b2 ; x [n,din], W1 [dh,din], W2 [dout,dh].
 * loss = mean over n*dout of (y - t)^2.  Gradients are written if non-NULL.
 * Returns the loss. */
double ref_mlp_loss_grad(const float *x, const float *t, int n, int din,
                         int dh, int dout, const float *W1, const float *b1,
                         const float *W2, const float *b2, float *gW1,
                         float *gb1, float *gW2, float *gb2, float *y);

#ifdef __cplusplus
}
#endif
#endif
