/*
 * mbx.h -- C ABI of the MI355X-native fused bitmap-index query path.
 *
 * Plain pointers and sizes only.  Every call returns an int status (MBX_OK = 0)
 * and never exits the process; mbx_last_error() gives the detail.  Device
 * pointers are gfx950 global memory obtained from mbx_alloc (or any hipMalloc
 * allocation meeting the alignment rules below).
 *
 * Bitmap layout is the reference's (primitive.cuh:94-98): uint32 words, bit b
 * of word k is row 32k+b, tail bits zero.  The kernels read two reference
 * words at a time (one 64-row word per lane), so every bitmap pointer must be
 * 8-byte aligned and readable up to ceil(N/64)*8 bytes; bitmaps made by this
 * library are 256-byte aligned and padded with zero words.
 *
 * Ranges are half-open [lo, hi).  Values are unsigned little-endian of
 * 1, 2, 4 or 8 bytes.  Sums are int64 (the reference sums into uint32 and
 * wraps; the low 32 bits of every sum here equal that wrapped value).
 */
#ifndef MBX_H
#define MBX_H
#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Limits of a program, as primitive.cuh:121 (MAXBIN_PERCOL, MAXCOLS,
 * MAXNINSTR). */
#define MBX_MAXBIN_PERCOL 4
#define MBX_MAXCOLS 4
#define MBX_MAXNINSTR 6
#define MBX_MAXKEYS 3
#define MBX_MAXBOUNDS 8

enum mbx_status {
  MBX_OK = 0,
  MBX_ERR_INVALID = 1,     /* bad argument */
  MBX_ERR_OOM = 2,         /* device pool or host memory exhausted */
  MBX_ERR_HIP = 3,         /* a HIP runtime call failed */
  MBX_ERR_ALIGN = 4,       /* a bitmap pointer is not 8-byte aligned */
  MBX_ERR_CAPACITY = 5,    /* a fixed capacity was exceeded */
  MBX_ERR_NODEVICE = 6,    /* no usable GPU */
  MBX_ERR_ADMISSION = 7,   /* serving: request refused by memory admission */
  MBX_ERR_UNSUPPORTED = 8  /* MBX_PLAN_DEDICATED: no kernel built for the
                              program's shape */
};

typedef struct mbx_ctx mbx_ctx;   /* device, stream, memory pool */

/* ---------------------------------------------------------------- context */
/* Replaces the implicit CUDA context and merle/dumb_pool.cuh:39-95.
 * pool_bytes = 0 lets the pool grow in 256 MiB arenas; otherwise it is one
 * arena of exactly pool_bytes (rounded up to 256 B) and exhaustion is
 * MBX_ERR_OOM. */
int mbx_init(int device, size_t pool_bytes, mbx_ctx **out);
int mbx_destroy(mbx_ctx *ctx);
const char *mbx_strerror(int status);
const char *mbx_last_error(void);     /* per host thread */
int mbx_device_count(int *n);
int mbx_sync(mbx_ctx *ctx);
/* The HIP stream all launches of this context go to (hipStream_t). */
void *mbx_stream(mbx_ctx *ctx);
int mbx_set_stream(mbx_ctx *ctx, void *hip_stream);  /* NULL = own stream */
/* No reference counterpart: streams for running independent queries of one
 * context concurrently (mbx_set_stream before each call).  Non-blocking,
 * default priority.  mbx_stream_join makes the context's current stream
 * wait on the device for the work enqueued so far on the given streams. */
int mbx_stream_create(void **hip_stream);
int mbx_stream_destroy(void *hip_stream);
int mbx_stream_join(mbx_ctx *ctx, void *const *hip_streams, int n);

/* ----------------------------------------------------------------- memory */
/* Pool-backed, 256-byte aligned.  Replaces __cum (bmpidx_create.cu:4-9) and
 * dumb_pool_t::alloc/free (merle/dumb_pool.cuh:57-92). */
int mbx_alloc(mbx_ctx *ctx, size_t bytes, void **dptr);
int mbx_free(mbx_ctx *ctx, void *dptr);
int mbx_pool_stats(mbx_ctx *ctx, size_t *used, size_t *capacity,
                   size_t *peak, size_t *nlive);
int mbx_memcpy_h2d(mbx_ctx *ctx, void *dst, const void *src, size_t bytes);
int mbx_memcpy_d2h(mbx_ctx *ctx, void *dst, const void *src, size_t bytes);
int mbx_memcpy_d2d(mbx_ctx *ctx, void *dst, const void *src, size_t bytes);
/* Fills with a kernel (no hipMemset), so it is safe inside graph capture. */
int mbx_fill_u32(mbx_ctx *ctx, void *dst, uint32_t value, size_t count);

/* A host-backed pool with the same allocator, for tests without a GPU and
 * for admission planning.  Addresses are host memory. */
typedef struct mbx_hostpool mbx_hostpool;
int mbx_hostpool_create(size_t capacity, mbx_hostpool **out);
int mbx_hostpool_alloc(mbx_hostpool *p, size_t bytes, void **ptr);
int mbx_hostpool_free(mbx_hostpool *p, void *ptr);
int mbx_hostpool_stats(mbx_hostpool *p, size_t *used, size_t *capacity,
                       size_t *peak, size_t *nlive, size_t *largest_free);
int mbx_hostpool_destroy(mbx_hostpool *p);

/* ------------------------------------------------------- bitmap creation */
/* A value source: a fact column (fk == NULL) or a dimension attribute read
 * through a foreign-key column of the fact table. */
typedef struct mbx_colref {
  const void *data;       /* device: fact column or dimension attribute */
  const uint32_t *fk;     /* device: foreign keys, or NULL */
  uint32_t dimsz;         /* dimension rows; fk >= dimsz matches nothing */
  uint32_t elem_bytes;    /* 1, 2, 4 or 8 */
} mbx_colref;

/* ceil(N/32): the reference's words per bitmap. */
size_t mbx_bitmap_words(uint32_t nrows);
/* Words between consecutive bitmaps of one slab: ceil(N/32) rounded up to 64
 * words (256 B). */
size_t mbx_bitmap_stride_words(uint32_t nrows);

/* create_bin (bmpidx_create.cu:14-56): one pass over the column writes nbins
 * bitmaps, bit j set iff mins[j] <= value < maxs[j]; a foreign key >= dimsz
 * reads as INT64_MAX (:26-27).  One slab is allocated from the pool;
 * out[j] = out[0] + j*stride; free only out[0].  *single_bytes (optional)
 * gets ceil(N/32)*4 as the reference returns.  Host arrays mins/maxs/out.
 * The call waits for its kernels when it uses temporary pool memory: bounds
 * on the device (bins of unequal widths or not consecutive), or the bin
 * index of every dimension row for a foreign-key column with consecutive
 * bins (dimsz+1 entries of 1 byte, 2 from 255 bins; read directly instead
 * when the pool has no room). */
int mbx_create_bin(mbx_ctx *ctx, const mbx_colref *src, uint32_t nrows,
                   const uint64_t *mins, const uint64_t *maxs, size_t nbins,
                   uint32_t **out, size_t *single_bytes);
/* Same, into caller memory: nbins bitmaps at slab + j*stride_words. */
int mbx_create_bin_into(mbx_ctx *ctx, const mbx_colref *src, uint32_t nrows,
                        const uint64_t *mins, const uint64_t *maxs,
                        size_t nbins, uint32_t *slab, size_t stride_words);
/* Several attributes of one dimension read through one foreign-key column
 * (c_region, c_nation and c_city through lo_custkey): the bitmaps of every
 * set as mbx_create_bin_into(ctx, {set.data, fk, dimsz, set.elem_bytes},
 * nrows, set.mins, set.maxs, set.nbins, set.slab, stride_words) would write
 * them, in one pass over the keys (one key load and one lookup of a table of
 * packed per-row bin indexes per fact row).  Sets of consecutive bins, at
 * most 4 sets whose indexes fit 32 bits together (8 bits below 255 bins,
 * else 16); other cases take one pass per set.  No reference counterpart
 * beyond create_bin itself (bmpidx_create.cu:14-56).  Waits for its
 * kernels (the table is temporary pool memory). */
typedef struct mbx_binset {
  const void *data;          /* device: dimension attribute */
  uint32_t elem_bytes;       /* 1, 2, 4 or 8 */
  uint32_t nbins;
  const uint64_t *mins, *maxs;   /* host, nbins each */
  uint32_t *slab;            /* device: nbins bitmaps, stride_words apart */
} mbx_binset;
int mbx_create_bin_multi(mbx_ctx *ctx, const uint32_t *fk, uint32_t dimsz,
                         uint32_t nrows, const mbx_binset *sets, size_t nsets,
                         size_t stride_words);
/* create_select / create_join (bmpidx_create.cu:58-136): one bitmap;
 * *out = NULL and MBX_ERR_OOM on allocation failure. */
int mbx_create_select(mbx_ctx *ctx, const void *src, uint32_t elem_bytes,
                      uint32_t nrows, uint32_t lo, uint32_t hi,
                      uint32_t **out);
int mbx_create_join(mbx_ctx *ctx, const uint32_t *fk, const void *dimattr,
                    uint32_t elem_bytes, uint32_t nrows, uint32_t dimsz,
                    uint32_t lo, uint32_t hi, uint32_t **out);

/* --------------------------------------------------------------- program */
/* vprg (primitive.cuh:135-198).  Opcodes as vprg::ops. */
enum mbx_opcode { MBX_END = 0, MBX_AND, MBX_OR, MBX_NOT, MBX_ANDM, MBX_ORM };
typedef struct mbx_col {
  const uint32_t *middle[MBX_MAXBIN_PERCOL]; /* bins inside the range */
  const uint32_t *leftmost, *rightmost;      /* partly covered edge bins */
} mbx_col;
typedef struct mbx_instr {
  uint8_t opcode, dstreg, src1, src2;        /* src1 also names a column */
} mbx_instr;
typedef struct mbx_vprg {
  mbx_col cols[MBX_MAXCOLS];
  uint32_t factsz;
  mbx_instr instrs[MBX_MAXNINSTR];
} mbx_vprg;

/* The row operation: a data descriptor standing in for the reference's
 * per-query device functor op_t (primitive.cuh:10-15, ssb/lambdas.cuh:19-152).
 *  - preds[i] is the exact predicate behind vprg column i; an uncertain row
 *    is kept iff the program's instruction list, run over these exact
 *    predicate booleans, leaves r0 true.
 *  - group key = sum_i stride_i * (f_i(v_i) - base_i), f = identity, v/div or
 *    #{bounds <= v}; rows whose key falls outside [0, nr_grp) are dropped.
 *  - measure = 1, a, a*b or a-b, in int64. */
enum mbx_keykind { MBX_KEY_IDENT = 0, MBX_KEY_DIV = 1, MBX_KEY_BUCKET = 2 };
enum mbx_measure { MBX_MEAS_COUNT = 0, MBX_MEAS_COL = 1, MBX_MEAS_MUL = 2,
                   MBX_MEAS_SUB = 3 };
typedef struct mbx_pred {
  mbx_colref src;
  uint64_t lo, hi;
} mbx_pred;
typedef struct mbx_keypart {
  mbx_colref src;
  uint32_t kind;
  uint32_t div;
  int64_t base;
  int64_t stride;
  uint32_t nbounds;
  uint32_t pad_;
  uint64_t bounds[MBX_MAXBOUNDS];
} mbx_keypart;
typedef struct mbx_rowop {
  mbx_pred preds[MBX_MAXCOLS];
  uint32_t npreds;
  uint32_t nkeys;
  mbx_keypart keys[MBX_MAXKEYS];
  uint32_t measure;
  uint32_t pad_;
  mbx_colref ma, mb;
} mbx_rowop;

/* Execution plans of one query. */
enum mbx_flags {
  MBX_PLAN_FUSED = 0,        /* vprg_grpby (primitive.cuh:340-445) */
  MBX_PLAN_MATERIALISE = 1,  /* _stg1/_stg2/_stg3 (primitive_abla.cuh:262-453) */
  MBX_PLAN_SCAN = 2,         /* grpby on every row (primitive.cuh:16-48) */
  MBX_PLAN_MASK = 3,
  MBX_ACCUMULATE = 16,       /* add into d_out instead of zeroing it first */
  MBX_PLAN_EAGER = 32,       /* fused/materialise: load every bitmap operand
                                up front instead of skipping ANDM loads for
                                empty words (same bits, fewer round trips) */
  MBX_PLAN_DIRECT = 64,      /* fused, ablation: no compaction -- each lane
                                walks the set bits of its own words and runs
                                the row op on them (abla::method1,
                                primitive_abla.cuh:62-108) */
  MBX_PLAN_BASEFUSE = 128,   /* fused, ablation: method 1 with the program's
                                virtual registers in device memory
                                (abla::base_fuse, primitive_abla.cuh:15-59);
                                waits for its kernel (pool scratch) */
  MBX_PLAN_DEDICATED = 256   /* fused, ablation: a kernel built for the
                                program's shape, the flat operand list fixed
                                at compile time (abla::hardcoded and the
                                dedicated c::s** kernels,
                                primitive_abla.cuh:188-260,
                                ssb/inst/1c.cu:5-22); MBX_ERR_UNSUPPORTED
                                when the library holds none for the shape
                                (see mbx_prog_shape) */
};

/* ---------------------------------------------- WAH compressed bitmaps */
/* The compressed-bitmap baseline the reference's report sets beside the
 * fused path (merle/wahGpuCmprs.cu, merle/wahGpuDirect.cu,
 * ssb/wah-demo.cu); the WAH format is the reference's, byte for byte.
 * Groups: 31 rows per uint32, row 31g+j at bit 30-j of group g, bit 31
 * zero; mbx_wah_groups_of(N) = ceil(N/31) groups, rows >= N zero (the
 * reference's dbjoinFlatWah makes floor(N/31) and drops the last rows).
 * WAH words: bit 31 set = a fill (bit 30 = the fill bit, bits 0-29 = its
 * groups), else a literal holding one group.  All arrays are device uint32
 * and at most 2^30 - 1 words or groups long. */
size_t mbx_wah_groups_of(uint32_t nrows);
/* dbjoinFlatWah (wahGpuCmprs.cu:247-320): groups of [lo <= v < hi] over a
 * column or a dimension attribute through foreign keys (keys >= dimsz
 * select nothing). */
int mbx_wah_select(mbx_ctx *ctx, const mbx_colref *src, uint32_t nrows,
                   uint64_t lo, uint64_t hi, uint32_t *groups);
/* The groups of a bitmap of this library (bit b of word k = row 32k+b). */
int mbx_wah_groups(mbx_ctx *ctx, const uint32_t *bitmap, uint32_t nrows,
                   uint32_t *groups);
/* wahCompress (wahGpuCmprs.cu:215-245): each maximal run of all-zero or
 * all-one groups becomes one fill, every other group a literal, and a
 * trailing zero run is dropped.  out holds up to ngroups words; *nout gets
 * the count (the call waits for its kernels). */
int mbx_wah_compress(mbx_ctx *ctx, const uint32_t *groups, size_t ngroups,
                     uint32_t *out, size_t *nout);
/* wahCntExcScan (wahGpuCmprs.cu:6-14): excl[i] = groups before word i,
 * excl[n] = all groups (n + 1 entries).  Its tile sums live in pool memory
 * kept per stream of the context (grown on demand, after a wait). */
int mbx_wah_scan(mbx_ctx *ctx, const uint32_t *wah, size_t n, uint32_t *excl);
/* wahAndNo1 (wahGpuDirect.cu:8-34): up AND down in down's word structure
 * (out has ndown words): a 0-fill of down stays, a literal (or a 1-fill of
 * one group) is ANDed with the up word covering its group -- groups past up
 * read zero -- and a zero result becomes a 0-fill of one group.  down must
 * hold no 1-fill longer than one group (mbx_wah_longest_1fill <= 1).  The
 * reference's SSB driver passes exclusive scans where the function expects
 * inclusive ones (ssb/wah-demo.cu:48-61), so its lookup can name the word
 * after a covering fill; here the covering word is always the one used. */
int mbx_wah_and_no1(mbx_ctx *ctx, const uint32_t *up, const uint32_t *up_excl,
                    size_t nup, const uint32_t *down,
                    const uint32_t *down_excl, size_t ndown, uint32_t *out);
/* wahDecomp (wahGpuCmprs.cu:37-51): the first ngroups groups of the words,
 * zero past their last group. */
int mbx_wah_decompress(mbx_ctx *ctx, const uint32_t *wah, const uint32_t *excl,
                       size_t n, uint32_t *groups, size_t ngroups);
/* wahLongest1Fill (merle/host/wahConv.c:6-20): the longest run of one
 * groups held by consecutive 1-fill words (a 0x7fffffff literal counts as
 * one).  Waits for its kernel. */
int mbx_wah_longest_1fill(mbx_ctx *ctx, const uint32_t *wah, size_t n,
                          uint64_t *longest);
/* No reference counterpart: a skip table beside the count scan, so that
 * the AND and the decompression find the word covering a group with one
 * read instead of a search through the scan.  skip[i] = the word covering
 * group 1024*i, with bit 31 set when that word starts at the group, for i
 * below nskip and below ceil(groups / 1024); mbx_wah_skip_len(G) =
 * ceil(G / 1024) entries cover a bitmap of G groups.  mbx_wah_scan_skip is
 * mbx_wah_scan plus the table (mbx_wah_scan = a table of no entries). */
size_t mbx_wah_skip_len(size_t ngroups);
int mbx_wah_scan_skip(mbx_ctx *ctx, const uint32_t *wah, size_t n,
                      uint32_t *excl, uint32_t *skip, size_t nskip);
/* A WAH bitmap on the device: n words, their count scan (n + 1 entries)
 * and their skip table (nskip entries of mbx_wah_scan_skip, or null and
 * 0); a table shorter than the bitmap's groups need is completed by
 * searches. */
typedef struct mbx_wah {
  const uint32_t *words;
  const uint32_t *excl;
  const uint32_t *skip;
  size_t n;
  size_t nskip;
} mbx_wah;
/* mbx_wah_and_no1 and mbx_wah_decompress on bitmaps with skip tables, the
 * same results: the AND takes the down bitmap's groups 1,024 at a time
 * (the down words starting there, the up words covering them) when down
 * has a table, else it is mbx_wah_and_no1. */
int mbx_wah_and_no1_skip(mbx_ctx *ctx, const mbx_wah *up, const mbx_wah *down,
                         uint32_t *out);
int mbx_wah_decompress_skip(mbx_ctx *ctx, const mbx_wah *w, uint32_t *groups,
                            size_t ngroups);

/* Program + group-by in one call.  d_out[nr_grp] int64 on the device.
 * The fused plan never materialises a bitmap. */
int mbx_prog_grpby(mbx_ctx *ctx, const mbx_vprg *prg, const mbx_rowop *op,
                   uint32_t nr_grp, int64_t *d_out, uint32_t flags);

/* No reference counterpart: for callers that cannot allocate device memory
 * (code under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc).  Same arguments, plans and results as mbx_prog_grpby, but it
 * calls no hipMalloc, hipFree or pool allocation: its scratch (flush
 * replicas; the materialise plan's bitmaps and row list) is one 64 MiB
 * device array per GPU, set up by mbx_init.  The materialise plan takes at
 * most 15,294,884 rows (MBX_ERR_CAPACITY above).  Every call on a GPU
 * shares that GPU's array: calls that overlap on one GPU (other streams,
 * other contexts, other threads) race -- serialise them. */
int mbx_prog_grpby_noalloc(mbx_ctx *ctx, const mbx_vprg *prg,
                           const mbx_rowop *op, uint32_t nr_grp, int64_t *d_out,
                           uint32_t flags);

/* Shape of a program for MBX_PLAN_DEDICATED: a conjunction of clauses of
 * one column each (at most 4), clause k in bits 6k..6k+5 as
 * nm | left << 3 | right << 4 | negated << 5 (nm = middle bins, left /
 * right = edge bins present), clauses in the order the fused kernel reads
 * them; *code = 0 for any other program.  *built = 1 when the library holds
 * a dedicated kernel for the code.  Host only.  Replaces the choice of a
 * hand-written c::s** function per query and layout (ssb/fixedlayout.cu:
 * 15-32, where a missing one reads as 0). */
int mbx_prog_shape(const mbx_vprg *prg, uint32_t *code, int *built);

/* The materialise stages, separately callable (primitive_abla.cuh:262-394).
 * mbx_prog_eval writes ceil(N/64)*2 words each of possible/uncertain
 * (d_unc may be NULL for a program without edge bins); it returns nothing
 * else.  mbx_compact appends possible row ids (MSB = uncertain) to d_rows and
 * counts them in *d_count (device uint32, zeroed by the call).
 * mbx_rows_grpby reads the count on the device (no host round trip). */
int mbx_prog_eval(mbx_ctx *ctx, const mbx_vprg *prg, uint32_t *d_poss,
                  uint32_t *d_unc);
/* flags & MBX_PLAN_BASEFUSE: the ablation with the registers in device
 * memory (the reference report's NfuStg1: _stg1 with idxbuf,
 * primitive_abla.cuh:262-288); waits for its kernel. */
int mbx_prog_eval_ex(mbx_ctx *ctx, const mbx_vprg *prg, uint32_t *d_poss,
                     uint32_t *d_unc, uint32_t flags);
int mbx_compact(mbx_ctx *ctx, const uint32_t *d_poss, const uint32_t *d_unc,
                uint32_t nrows, uint32_t *d_rows, uint32_t *d_count);
int mbx_rows_grpby(mbx_ctx *ctx, const uint32_t *d_rows,
                   const uint32_t *d_count, const mbx_vprg *prg,
                   const mbx_rowop *op, uint32_t nr_grp, int64_t *d_out,
                   uint32_t flags);

/* Algorithmic bytes of a fused run (compulsory 128-byte lines of every
 * bitmap, column and dimension the plan must touch, plus 8*nr_grp), counted
 * on the device by a separate accounting kernel. */
int mbx_alg_bytes(mbx_ctx *ctx, const mbx_vprg *prg, const mbx_rowop *op,
                  uint32_t nr_grp, uint64_t *bytes, uint64_t *selected);

/* ------------------------------------------------------------ batching */
/* No reference counterpart: the reference launches one query per kernel
 * (ssb/lambdas.cuh:179-184, synth/synthdemo.cu:122-129).  One query of a
 * batch: */
typedef struct mbx_query {
  mbx_vprg prg;
  mbx_rowop op;
  uint32_t nr_grp;
  uint32_t flags;          /* 0 or MBX_PLAN_EAGER (all operand loads up front);
                              serving also takes MBX_PLAN_SCAN (a scan launch
                              of its own; other plans run fused) */
  int64_t *d_out;          /* device, nr_grp int64 */
} mbx_query;

/* Launch shape of the persistent batched kernel (chosen by mbx_autotune). */
typedef struct mbx_tune {
  uint32_t tile_words;     /* 64-row words per work item (multiple of 256) */
  uint32_t wg_per_cu;      /* resident workgroups per CU for the grid */
  uint32_t lds_groups;     /* group-table capacity in LDS (int64 slots) */
  uint32_t pad_;
} mbx_tune;

/* Algorithmic bytes of a set of queries (compulsory 128-B lines of their
 * bitmaps and gathered columns, a line shared by several queries counted
 * once, + 8 B per output group).  No reference counterpart: it prices a
 * batch for the roofline. */
int mbx_alg_bytes_batch(mbx_ctx *ctx, const mbx_query *qs, uint32_t nq,
                        uint64_t *bytes, uint64_t *selected);

typedef struct mbx_batch mbx_batch;
/* Builds device descriptors and captures the whole batch (output zeroing,
 * work-queue reset, one persistent launch) into a hipGraph. */
int mbx_batch_create(mbx_ctx *ctx, const mbx_query *qs, uint32_t nq,
                     const mbx_tune *tune, mbx_batch **out);
int mbx_batch_run(mbx_ctx *ctx, mbx_batch *b);        /* graph launch */
int mbx_batch_run_eager(mbx_ctx *ctx, mbx_batch *b);  /* same, no graph */
int mbx_batch_destroy(mbx_ctx *ctx, mbx_batch *b);
/* Times candidate launch shapes on the batch and returns the fastest. */
int mbx_autotune(mbx_ctx *ctx, const mbx_query *qs, uint32_t nq,
                 mbx_tune *best, float *best_ms);
/* Same, also reporting every shape timed and its median time (the first
 * `max` of them; *count = how many were timed). */
int mbx_autotune_report(mbx_ctx *ctx, const mbx_query *qs, uint32_t nq,
                        mbx_tune *best, float *best_ms, mbx_tune *cands,
                        float *cand_ms, int max, int *count);
int mbx_default_tune(mbx_tune *t);


/* ------------------------------------------------------------- serving */
/* No reference counterpart (the reference runs one query per launch from a
 * single host thread).  A server keeps the caller's columns and bitmaps
 * where they are and runs submitted queries on nstreams slots (a HIP stream
 * each): memory-based admission, then continuous batching -- whenever a slot
 * is free, up to max_batch admitted requests go out on it as one persistent
 * batched launch, while the other slots are still running; a request of
 * >= fused_rows fact rows, or with the scan plan, goes out alone as a launch
 * of its own.  plain = 1 is the
 * one-at-a-time baseline (one slot, one fused launch per request).
 *
 * Memory: the server's fixed working memory (per slot: descriptors, work
 * counters, flush replicas) is taken from the pool at open and charged
 * against mem_budget first; each request then reserves its int64 output
 * (nr_grp * 8 bytes, rounded to 256) at admission.  Admission is FIFO: a
 * request that can never fit is rejected; one that does not fit now waits
 * until running requests release memory; if the pool itself refuses it
 * while nothing of the server's is running or admitted, it is rejected.
 * Serving runs the fused and batched plans only (no materialise scratch).
 *
 * Ordering: every request is ordered after the work enqueued on the
 * context's stream before its submit call (the slot waits on an event
 * recorded there).  Results are copied to the host by per-slot completion
 * threads, never by the dispatching thread. */
typedef struct mbx_serve_cfg {
  uint32_t nstreams;        /* 1..16 slots */
  uint32_t max_batch;       /* requests per dispatch, >= 1 */
  uint64_t mem_budget;      /* bytes; 0 = unbounded */
  uint32_t plain;           /* 1: one request at a time, one stream */
  uint32_t fused_rows;      /* requests with >= this many fact rows run as
                               their own fused launches; 0 = 32Mi */
  mbx_tune tune;            /* tile_words == 0: default shape */
} mbx_serve_cfg;
typedef struct mbx_serve_stats {
  uint32_t completed, rejected;
  double wall_s, p50_s, p99_s, mean_s, max_s;
  uint32_t batches;         /* dispatches (launch groups) */
  uint32_t admission_waits; /* requests delayed at least once by admission */
  double mean_batch;
  uint64_t peak_mem;        /* bytes held at the peak, fixed part included */
  uint32_t cancelled, failed;
  uint64_t fixed_mem;       /* the server's fixed working memory */
} mbx_serve_stats;

enum mbx_req_state {
  MBX_REQ_QUEUED = 0,       /* waiting for admission */
  MBX_REQ_ADMITTED = 1,     /* output reserved, waiting for a slot */
  MBX_REQ_RUNNING = 2,      /* on a slot */
  MBX_REQ_DONE = 3,         /* result in h_result */
  MBX_REQ_REJECTED = 4,     /* refused by admission */
  MBX_REQ_CANCELLED = 5,
  MBX_REQ_FAILED = 6        /* a HIP error; see mbx_serve_close's status */
};

typedef struct mbx_server mbx_server;
/* The fixed working memory a server of this configuration takes at open
 * (so a caller can size mem_budget). */
int mbx_serve_fixed_mem(const mbx_serve_cfg *cfg, uint64_t *bytes);
int mbx_serve_open(mbx_ctx *ctx, const mbx_serve_cfg *cfg, mbx_server **out);
/* q->d_out is ignored; the result (q->nr_grp int64) lands in h_result, host
 * memory that must stay valid until the request ends.  Returns at once with
 * a ticket. */
int mbx_serve_submit(mbx_server *s, const mbx_query *q, int64_t *h_result,
                     uint64_t *ticket);
/* State of a request and, once it is done, its latency (result on the host
 * minus submit time; -1 otherwise). */
int mbx_serve_poll(mbx_server *s, uint64_t ticket, int *state,
                   double *latency_s);
/* Waits until the request ends or timeout_s passes (< 0: no limit). */
int mbx_serve_wait(mbx_server *s, uint64_t ticket, double timeout_s,
                   int *state, double *latency_s);
/* Cancels a queued or admitted request; MBX_ERR_INVALID once it runs. */
int mbx_serve_cancel(mbx_server *s, uint64_t ticket);
/* Waits for every submitted request to end, then frees the server.  Returns
 * the first HIP error any request met, else MBX_OK. */
int mbx_serve_close(mbx_server *s, mbx_serve_stats *stats);

/* A client of the server: replays a request trace (arrival times in
 * seconds, sorted), submitting each request at its arrival time.
 * reqs[i].d_out is ignored: results land in h_results (requests' outputs
 * concatenated in request order); latency_s[i] = result on the host -
 * arrival, or -1 if refused. */
int mbx_serve_replay(mbx_ctx *ctx, const mbx_query *reqs, uint32_t nq,
                     const double *arrival_s, const mbx_serve_cfg *cfg,
                     int64_t *h_results, double *latency_s,
                     mbx_serve_stats *stats);

/* --------------------------------------------------------------- timing */
/* Event pairs on the context's stream, as DOWORK times its launches with
 * cudaEventRecord / cudaEventElapsedTime (ssb/lambdas.cuh:158-240). */
typedef struct mbx_timer mbx_timer;
int mbx_timer_create(mbx_ctx *ctx, mbx_timer **t);
int mbx_timer_start(mbx_ctx *ctx, mbx_timer *t);
int mbx_timer_stop(mbx_ctx *ctx, mbx_timer *t, float *ms);  /* syncs */
/* Records the end event without waiting; mbx_timer_elapsed waits for it. */
int mbx_timer_record_stop(mbx_ctx *ctx, mbx_timer *t);
int mbx_timer_elapsed(mbx_timer *t, float *ms);
int mbx_timer_destroy(mbx_timer *t);
/* Plain device copy of `bytes` (for byte-counter calibration). */
int mbx_copy_kernel(mbx_ctx *ctx, void *dst, const void *src, size_t bytes);


/* ------------------------------------------------------------ plans */
/* No reference counterpart beyond the static bin choice of col::from_bin2
 * (bmpidx_create.cu:165-174); these pick evaluation order and fused vs
 * materialised execution per query.  A plan of one query: the column order of a conjunctive program
 * (ORM r0 <- c ; ANDM r0 <- c ; ...) and the execution plan (MBX_PLAN_*).
 * Every plan of a query selects the same rows and gives the same sums. */
#define MBX_NFEAT 16
typedef struct mbx_plan {
  uint8_t order[MBX_MAXCOLS];
  uint8_t norder;
  uint8_t reorderable;      /* program is a pure conjunction */
  uint16_t pad_;
  uint32_t flags;           /* MBX_PLAN_FUSED (| MBX_PLAN_EAGER),
                               MBX_PLAN_MATERIALISE or MBX_PLAN_SCAN */
} mbx_plan;
typedef struct mbx_mlp mbx_mlp;

/* Fraction of rows in each column's operand (OR of its bins); columns
 * without bitmaps get 0.  Popcounts run on the device. */
int mbx_col_density(mbx_ctx *ctx, const mbx_vprg *prg, double dens[MBX_MAXCOLS]);
/* The same, plus edge[c]: the fraction of rows only in column c's edge bins
 * (rows a candidate check must re-check); edge may be NULL. */
int mbx_col_stats(mbx_ctx *ctx, const mbx_vprg *prg, double dens[MBX_MAXCOLS],
                  double edge[MBX_MAXCOLS]);
int mbx_plan_describe(const mbx_vprg *prg, mbx_plan *out);
int mbx_plan_apply(const mbx_vprg *in, const mbx_plan *plan, mbx_vprg *out);
int mbx_plan_features(const mbx_vprg *prg, const double *dens, uint32_t nr_grp,
                      const mbx_plan *plan, float *feat /* MBX_NFEAT */);
/* With the edge fractions of mbx_col_stats (NULL = none known). */
int mbx_plan_features_ex(const mbx_vprg *prg, const double *dens,
                         const double *edge, uint32_t nr_grp,
                         const mbx_plan *plan, float *feat /* MBX_NFEAT */);
/* Every plan: all orders x {fused, fused eager, materialise}, and the scan. */
int mbx_plan_enumerate(const mbx_vprg *prg, mbx_plan *out, int max, int *count);
/* The plans the learned chooser scores (at most 15): the early-out plan for
 * every ordered pair of leading columns (the rest sparsest first), and
 * eager, materialise and scan, sparsest column first. */
int mbx_plan_candidates(const mbx_vprg *prg, const double *dens, mbx_plan *out,
                        int max, int *count);
/* Static heuristic: ascending density, fused. */
int mbx_plan_heuristic(const mbx_vprg *prg, const double *dens, mbx_plan *out);
/* Learned chooser (cost-model MLP over every candidate); mlp == NULL falls
 * back to the heuristic. */
int mbx_plan_choose(mbx_ctx *ctx, mbx_mlp *mlp, const mbx_vprg *prg,
                    const double *dens, uint32_t nr_grp, mbx_plan *best,
                    float *predicted);
/* With edge fractions, and the heuristic kept unless the model predicts a
 * plan cheaper by more than `margin` (log runtime; 0 = always the model's). */
int mbx_plan_choose_ex(mbx_ctx *ctx, mbx_mlp *mlp, const mbx_vprg *prg,
                       const double *dens, const double *edge,
                       uint32_t nr_grp, float margin, mbx_plan *best,
                       float *predicted);
int mbx_prog_grpby_plan(mbx_ctx *ctx, const mbx_vprg *prg, const mbx_rowop *op,
                        uint32_t nr_grp, int64_t *d_out, const mbx_plan *plan);

/* ------------------------------------------------------ cost-model MLP */
/* No reference counterpart.  y = W2 relu(W1 x + b1) + b2 in fp32;
 * parameters flat W1|b1|W2|b2.
 * Forward, backward and Adam run as HIP kernels on the device. */
int mbx_mlp_create(mbx_ctx *ctx, int din, int dh, int dout, uint64_t seed,
                   mbx_mlp **out);
int mbx_mlp_destroy(mbx_ctx *ctx, mbx_mlp *m);
int mbx_mlp_nparams(const mbx_mlp *m);
int mbx_mlp_set_params(mbx_ctx *ctx, mbx_mlp *m, const float *w);
int mbx_mlp_get_params(mbx_ctx *ctx, mbx_mlp *m, float *w);
int mbx_mlp_forward(mbx_ctx *ctx, mbx_mlp *m, const float *d_x, int n,
                    float *d_y);
/* loss = mean (y-t)^2; grad (host, nparams) may be NULL */
int mbx_mlp_loss_grad(mbx_ctx *ctx, mbx_mlp *m, const float *d_x,
                      const float *d_t, int n, double *loss, float *grad);
int mbx_mlp_train(mbx_ctx *ctx, mbx_mlp *m, const float *d_x, const float *d_t,
                  int n, int steps, float lr, double *loss);

/* ------------------------------------------------- data: generators, files */
/* Seeded synthetic columns, bit-identical on host and device: value i is a
 * pure function of (seed, i) (splitmix64 mix; uniform by multiply-high;
 * finite Zipf(s) on {0..nvals-1} by an integer inverse-CDF table). */
uint64_t mbx_rand_at(uint64_t seed, uint64_t i);
int mbx_zipf_table(uint32_t nvals, double s, uint64_t *thresholds);
int mbx_gen_uniform_u32_host(uint32_t *out, size_t n, uint64_t seed,
                             uint32_t lo, uint32_t hi);
int mbx_gen_zipf_u32_host(uint32_t *out, size_t n, uint64_t seed,
                          uint32_t nvals, double s);
int mbx_gen_uniform_u32(mbx_ctx *ctx, uint32_t *d_out, size_t n,
                        uint64_t seed, uint32_t lo, uint32_t hi);
int mbx_gen_zipf_u32(mbx_ctx *ctx, uint32_t *d_out, size_t n, uint64_t seed,
                     uint32_t nvals, double s);
/* d_out[i] = d_in[i] / div (% mod when mod != 0): nested attributes. */
int mbx_gen_map_u32(mbx_ctx *ctx, uint32_t *d_out, const uint32_t *d_in,
                    size_t n, uint32_t div, uint32_t mod);

/* Column files: 4-byte LE header (element BYTES when header_kind = 0, as
 * ssb-dbgen-lite.c:192-201 writes; BITS when 1, as synth/onegen_zipf.py:73-84
 * writes) + raw LE values; readers size by the file (ssb/cpuref.c:8-27). */
int mbx_colfile_write(const char *path, const void *data, size_t n,
                      uint32_t elem_bytes, int header_kind);
int mbx_colfile_read(const char *path, void **data, size_t *nbytes,
                     uint32_t *header);
void mbx_colfile_free(void *p);
int mbx_colfile_load(mbx_ctx *ctx, const char *path, void **dptr,
                     size_t *nbytes, uint32_t *header);

/* sizeof of every ABI struct in declaration order (binding checks). */
int mbx_abi_sizes(size_t *out, int n);

#ifdef __cplusplus
}
#endif
#endif
