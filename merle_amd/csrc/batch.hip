// batch.hip -- many queries in one persistent launch, captured in a hipGraph.
//
// The reference runs one query per launch (ssb/lambdas.cuh:179-184).  Here a
// batch of queries becomes a set of (query, row range) work items.  The
// queries of a batch mostly read the same columns and bitmaps, so items are
// ordered row range first: the items of one row range -- that range of every
// query -- run close together in time, on one XCD, and its lines are fetched
// from HBM once and then served by that XCD's L2 to the other queries.  Row
// range r belongs to XCD r mod 8 (workgroup b runs on XCD b mod 8); each XCD
// drains its own list through its own counter.  For each item a workgroup
// zeroes its LDS group table, runs the same wave pipeline as the
// single-query kernel (program -> compaction -> row op) over the range, and
// flushes the table to the query's int64 output.  Queries come in two
// classes, each drained by its own persistent kernel: eager conjunctions
// with a small LDS table (most queries), whose kernel holds only that
// evaluation form and sink (held to 80 VGPRs: 6 waves per SIMD), and the
// rest (every form, LDS or HBM sink: 111 VGPRs, 4 waves).  The waves spend
// most of their time waiting on loads (L2 hits, dependent rounds), so the
// two kernels run side by side on two streams, and the conjunction kernel
// trades words per lane for resident waves.  Output zeroing, counter reset
// and the two launches are captured once into a hipGraph (fill kernels only:
// no memset nodes) and replayed per run.
#include <algorithm>
#include <cstring>
#include <vector>

#include "batch.hpp"
#include "devquery.hpp"
#include "query_launch.hpp"

namespace mbx {

struct BatchArgs {
  const BatchQ *qs;
  uint32_t nq;
  uint32_t max_items;   // row ranges of the longest query
  uint32_t nx;          // XCD lists (8, fewer for a tiny grid)
  uint32_t *counter;    // nx counters, kXcdCounterStride apart
};

// Bitmap words through cached loads, R words a round of the dense re-check
// (the conjunction kernel, held to 80 VGPRs, takes one).
template <uint32_t R>
struct BatchItemShape : RuntimeShapeCached {
  static constexpr uint32_t kRecheck = R;
};

template <bool CHK, int EAGER, uint32_t R = 1>
__device__ __forceinline__ void batch_item(const BatchQ &q, uint32_t w_begin,
                                           uint32_t w_end, const GroupSink &sink,
                                           uint32_t *s_list) {
  // the item's 64-word chunks, interleaved over the workgroup's four waves,
  // kBWPL per wave and pass (a 512-word item in one pass)
  constexpr int W = kBWPL;
  const uint32_t wv = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  for (uint32_t w0 = w_begin + wv * 64; w0 < w_end; w0 += 4 * 64 * W)
    run_wave_range<CHK, false, W, EAGER, 2, false, false, BatchItemShape<R>>(
        q.prg, q.op, q.nr_grp, w0, 4 * 64, w_end, sink, s_list, nullptr);
}

// CONJ: the class-0 kernel (eager conjunctions, LDS table); else every form.
template <bool CONJ>
__global__ __launch_bounds__(256)
__attribute__((amdgpu_waves_per_eu(CONJ ? 6 : 1, 8))) void batch_kernel(BatchArgs a) {
  __shared__ __align__(16) uint32_t s_list[4][kWaveList];
  __shared__ uint32_t s_item;
  // uniform, so the row-list addresses stay scalar (4 VGPRs fewer)
  const uint32_t wv = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const uint32_t x = blockIdx.x % a.nx;
  // this XCD's row ranges x, x + nx, ...: nq items each
  const uint32_t nranges = a.max_items > x ? (a.max_items - x + a.nx - 1) / a.nx : 0;
  const uint32_t nitems = nranges * a.nq;
  for (;;) {
    if (threadIdx.x == 0)
      s_item = atomicAdd(a.counter + x * kXcdCounterStride, 1u);
    __syncthreads();
    const uint32_t j = __builtin_amdgcn_readfirstlane(s_item);
    if (j >= nitems) break;   // every thread sees the same item
    const uint32_t r = x + a.nx * (j / a.nq);
    const BatchQ &q = a.qs[j % a.nq];
    if (r >= q.nitems) {      // a shorter query: no such range
      __syncthreads();        // s_item is rewritten next round
      continue;
    }
    GroupSink sink;
    sink.lds = CONJ || q.lds != 0;
    sink.out = q.out;
    if (CONJ || q.lds)
      for (uint32_t i = threadIdx.x; i < q.nr_grp; i += 256) s_grp_dyn[i] = 0;
    __syncthreads();
    const uint32_t w_begin = r * q.words_per_item;
    const uint32_t w_end = min(w_begin + q.words_per_item, q.prg.nwords);
    // eager evaluation for conjunctions (no program interpretation, two- or
    // three-valued) and for two-valued programs with small operand lists
    // (straight into registers); the rest take the early-out form (same
    // bits, fewer registers for the whole kernel)
    if (CONJ) {
      if (q.chk) batch_item<true, 3>(q, w_begin, w_end, sink, s_list[wv]);
      else batch_item<false, 3>(q, w_begin, w_end, sink, s_list[wv]);
    } else if (q.chk) {
      if (q.eager && q.prg.conj)
        batch_item<true, 3, 4>(q, w_begin, w_end, sink, s_list[wv]);
      else
        batch_item<true, 0, 4>(q, w_begin, w_end, sink, s_list[wv]);
    } else if (q.eager && q.prg.conj) {
      batch_item<false, 3>(q, w_begin, w_end, sink, s_list[wv]);
    } else if (q.eager && q.prg.nfl <= kDirectOps) {
      batch_item<false, 2>(q, w_begin, w_end, sink, s_list[wv]);
    } else {
      batch_item<false, 0>(q, w_begin, w_end, sink, s_list[wv]);
    }
    if (CONJ || q.lds) {
      __syncthreads();
      for (uint32_t i = threadIdx.x; i < q.nr_grp; i += 256) {
        const unsigned long long v = s_grp_dyn[i];
        if (v) atomicAdd(q.out + i, v);
      }
    }
    __syncthreads();   // s_item and the LDS table are reused next round
  }
}

__global__ void zero_outputs_kernel(const BatchQ *qs, uint32_t nq) {
  for (uint32_t q = blockIdx.x; q < nq; q += gridDim.x) {
    unsigned long long *o = qs[q].out;
    const uint32_t g = qs[q].nr_grp;
    for (uint32_t i = threadIdx.x; i < g; i += blockDim.x) o[i] = 0;
  }
}

}  // namespace mbx

using namespace mbx;

struct mbx_batch {
  BatchQ *d_qs = nullptr;     // pool memory: descriptors then the counters
  uint32_t *d_counter = nullptr;
  uint32_t nq = 0;
  mbx::BatchShape shape{};
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  mbx_tune tune{};
  mbx::BatchFork fork{};      // side stream of the general kernel
};

namespace mbx {

static unsigned batch_grid(mbx_ctx *ctx, const mbx_tune &t,
                           uint32_t total_items) {
  uint64_t g = std::min<uint64_t>((uint64_t)ctx->num_cus * t.wg_per_cu, total_items);
  if (g > kXcdLists) g -= g % kXcdLists;   // whole rounds over the XCDs
  return (unsigned)std::max<uint64_t>(1, g);
}

int batch_launch(mbx_ctx *ctx, hipStream_t s, const BatchQ *d_qs,
                 const BatchShape &sh, uint32_t *d_counter, const mbx_tune &t,
                 const BatchFork *fork) {
  const uint32_t nq = sh.nq[0] + sh.nq[1];
  hipLaunchKernelGGL(zero_outputs_kernel, dim3(std::min(nq, 1024u)),
                     dim3(256), 0, s, d_qs, nq);
  MBX_HIP(hipGetLastError());
  int st = launch_fill_u32(s, d_counter, 0, kCounterBytes / 4);
  if (st) return st;
  // both classes: the general kernel on the side stream, after the zeroing
  const bool split = fork && sh.nq[0] && sh.nq[1];
  if (split) {
    MBX_HIP(hipEventRecord(fork->ev_fork, s));
    MBX_HIP(hipStreamWaitEvent(fork->side, fork->ev_fork, 0));
  }
  for (int c = 0; c < 2; ++c) {
    if (sh.nq[c] == 0) continue;
    const hipStream_t ks = (split && c == 1) ? fork->side : s;
    const unsigned grid = batch_grid(ctx, t, sh.items[c]);
    BatchArgs a;
    a.qs = d_qs + (c ? sh.nq[0] : 0);
    a.nq = sh.nq[c];
    a.max_items = sh.max_items[c];
    a.nx = grid < kXcdLists ? grid : kXcdLists;
    a.counter = d_counter + c * kXcdLists * kXcdCounterStride;
    if (c == 0)
      hipLaunchKernelGGL(batch_kernel<true>, dim3(grid), dim3(256),
                         sh.lds_bytes[0], ks, a);
    else
      hipLaunchKernelGGL(batch_kernel<false>, dim3(grid), dim3(256),
                         sh.lds_bytes[1], ks, a);
    MBX_HIP(hipGetLastError());
  }
  if (split) {
    MBX_HIP(hipEventRecord(fork->ev_join, fork->side));
    MBX_HIP(hipStreamWaitEvent(s, fork->ev_join, 0));
  }
  return MBX_OK;
}

static int enqueue_batch(mbx_ctx *ctx, mbx_batch *b) {
  return batch_launch(ctx, ctx->stream, b->d_qs, b->shape, b->d_counter,
                      b->tune, b->fork.side ? &b->fork : nullptr);
}

int batch_prepare(mbx_ctx *ctx, const mbx_query *qs, uint32_t nq,
                  const mbx_tune &t, BatchQ *h, BatchShape *sh) {
  std::memset(sh, 0, sizeof(*sh));
  std::vector<BatchQ> cls[2];
  for (uint32_t i = 0; i < nq; ++i) {
    QueryPlan p;
    int st = prepare(ctx, &qs[i].prg, &qs[i].op, qs[i].nr_grp, qs[i].d_out, &p);
    if (st) {
      set_error("batch query " + std::to_string(i) + ": " + mbx_last_error());
      return st;
    }
    BatchQ b;
    std::memset(&b, 0, sizeof(b));
    b.prg = p.prg;
    b.op = p.op;
    b.out = p.out;
    b.nr_grp = p.nr_grp;
    b.lds = p.nr_grp <= t.lds_groups ? 1u : 0u;
    b.chk = p.prg.chk;
    b.eager = (qs[i].flags & MBX_PLAN_EAGER) ? 1u : 0u;
    b.words_per_item = t.tile_words;
    b.nitems = (uint32_t)ceil_div(p.prg.nwords, b.words_per_item);
    const int c = (b.eager && p.prg.conj && b.lds && p.nr_grp <= kConjLdsGroups) ? 0 : 1;
    sh->items[c] += b.nitems;
    sh->max_items[c] = std::max(sh->max_items[c], b.nitems);
    if (b.lds) sh->lds_bytes[c] = std::max<size_t>(sh->lds_bytes[c], (size_t)p.nr_grp * 8);
    cls[c].push_back(b);
  }
  for (int c = 0; c < 2; ++c) sh->nq[c] = (uint32_t)cls[c].size();
  std::copy(cls[0].begin(), cls[0].end(), h);
  std::copy(cls[1].begin(), cls[1].end(), h + cls[0].size());
  return MBX_OK;
}

static void destroy_batch(mbx_ctx *ctx, mbx_batch *b) {
  if (!b) return;
  if (b->exec) (void)hipGraphExecDestroy(b->exec);
  if (b->graph) (void)hipGraphDestroy(b->graph);
  if (b->fork.side) {
    (void)hipStreamSynchronize(b->fork.side);
    (void)hipStreamDestroy(b->fork.side);
  }
  if (b->fork.ev_fork) (void)hipEventDestroy(b->fork.ev_fork);
  if (b->fork.ev_join) (void)hipEventDestroy(b->fork.ev_join);
  if (b->d_qs && ctx) ctx->pool->free(b->d_qs);
  delete b;
}

static int valid_tune(const mbx_tune *t) {
  if (t->tile_words == 0 || t->tile_words % kItemQuantum)
    return fail(MBX_ERR_INVALID, "tune: tile_words must be a multiple of 512");
  if (t->wg_per_cu == 0 || t->wg_per_cu > 16)
    return fail(MBX_ERR_INVALID, "tune: wg_per_cu must be in 1..16");
  if (t->lds_groups > kLdsGroups)
    return fail(MBX_ERR_INVALID, "tune: lds_groups above 4096");
  return MBX_OK;
}

static int build_batch(mbx_ctx *ctx, const mbx_query *qs, uint32_t nq,
                       const mbx_tune *tune, bool capture, mbx_batch **out) {
  if (!ctx || !qs || !out || nq == 0)
    return fail(MBX_ERR_INVALID, "batch: bad argument");
  *out = nullptr;
  mbx_tune t;
  if (tune) t = *tune;
  else mbx_default_tune(&t);
  int st = valid_tune(&t);
  if (st) return st;
  std::vector<BatchQ> h(nq);
  BatchShape shape;
  st = batch_prepare(ctx, qs, nq, t, h.data(), &shape);
  if (st) return st;
  auto *b = new mbx_batch;
  b->nq = nq;
  b->shape = shape;
  b->tune = t;
  void *mem = nullptr;
  const size_t qbytes = ceil_div(sizeof(BatchQ) * nq, 256) * 256;
  st = ctx->pool->alloc(qbytes + kCounterBytes, &mem);
  if (st) {
    delete b;
    return st;
  }
  b->d_qs = static_cast<BatchQ *>(mem);
  b->d_counter = reinterpret_cast<uint32_t *>(static_cast<char *>(mem) + qbytes);
  if (shape.nq[0] && shape.nq[1]) {
    hipError_t fe = hipStreamCreateWithFlags(&b->fork.side, hipStreamNonBlocking);
    if (fe == hipSuccess)
      fe = hipEventCreateWithFlags(&b->fork.ev_fork, hipEventDisableTiming);
    if (fe == hipSuccess)
      fe = hipEventCreateWithFlags(&b->fork.ev_join, hipEventDisableTiming);
    if (fe != hipSuccess) {
      destroy_batch(ctx, b);
      return hip_fail(fe, "batch: side stream");
    }
  }
  hipError_t e = hipMemcpyAsync(b->d_qs, h.data(), sizeof(BatchQ) * nq,
                                hipMemcpyHostToDevice, ctx->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
  if (e != hipSuccess) {
    destroy_batch(ctx, b);
    return hip_fail(e, "batch: descriptor upload");
  }
  if (capture) {
    e = hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeRelaxed);
    if (e != hipSuccess) {
      destroy_batch(ctx, b);
      return hip_fail(e, "batch: begin capture");
    }
    st = enqueue_batch(ctx, b);
    hipGraph_t g = nullptr;
    e = hipStreamEndCapture(ctx->stream, &g);
    b->graph = g;
    if (st || e != hipSuccess) {
      destroy_batch(ctx, b);
      return st ? st : hip_fail(e, "batch: end capture");
    }
    e = hipGraphInstantiate(&b->exec, g, nullptr, nullptr, 0);
    if (e != hipSuccess) {
      destroy_batch(ctx, b);
      return hip_fail(e, "batch: graph instantiate");
    }
  }
  *out = b;
  return MBX_OK;
}

// Project-written autotuner: times every candidate launch shape on the
// batch itself (eager launches, median of 3 after one warm-up) and keeps the
// fastest.  Results do not depend on the shape: each candidate computes the
// same sums.  cands/cand_ms (optional) receive every shape timed.
static int autotune(mbx_ctx *ctx, const mbx_query *qs, uint32_t nq,
                    mbx_tune *best, float *best_ms, mbx_tune *cands,
                    float *cand_ms, int max, int *count) {
  if (!ctx || !qs || !best) return fail(MBX_ERR_INVALID, "autotune: null");
  static const uint32_t tiles[] = {512, 1024, 2048, 4096, 8192};
  static const uint32_t wgs[] = {2, 4, 5, 8};
  static const uint32_t ldsg[] = {1024, 4096};
  hipEvent_t e0, e1;
  MBX_HIP(hipEventCreate(&e0));
  MBX_HIP(hipEventCreate(&e1));
  float best_t = 1e30f;
  mbx_tune bt;
  mbx_default_tune(&bt);
  int st = MBX_OK, k = 0;
  for (uint32_t tw : tiles)
    for (uint32_t wg : wgs)
      for (uint32_t lg : ldsg) {
        mbx_tune t{tw, wg, lg, 0};
        mbx_batch *b = nullptr;
        st = build_batch(ctx, qs, nq, &t, false, &b);
        if (st) goto done;
        float reps[3];
        for (int r = -1; r < 3 && !st; ++r) {
          (void)hipEventRecord(e0, ctx->stream);
          st = enqueue_batch(ctx, b);
          (void)hipEventRecord(e1, ctx->stream);
          if (hipEventSynchronize(e1) != hipSuccess)
            st = fail(MBX_ERR_HIP, "autotune: sync");
          float ms = 0;
          (void)hipEventElapsedTime(&ms, e0, e1);
          if (r >= 0) reps[r] = ms;
        }
        destroy_batch(ctx, b);
        if (st) goto done;
        std::sort(reps, reps + 3);
        if (cands && cand_ms && k < max) {
          cands[k] = t;
          cand_ms[k] = reps[1];
        }
        ++k;
        if (reps[1] < best_t) {
          best_t = reps[1];
          bt = t;
        }
      }
done:
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  if (st) return st;
  *best = bt;
  if (best_ms) *best_ms = best_t;
  if (count) *count = k;
  return MBX_OK;
}

}  // namespace mbx

extern "C" {

int mbx_default_tune(mbx_tune *t) {
  if (!t) return fail(MBX_ERR_INVALID, "null tune");
  t->tile_words = 2048;
  t->wg_per_cu = 4;
  t->lds_groups = kLdsGroups;
  t->pad_ = 0;
  return MBX_OK;
}

int mbx_batch_create(mbx_ctx *ctx, const mbx_query *qs, uint32_t nq,
                     const mbx_tune *tune, mbx_batch **out) {
  return build_batch(ctx, qs, nq, tune, true, out);
}

int mbx_batch_run(mbx_ctx *ctx, mbx_batch *b) {
  if (!ctx || !b) return fail(MBX_ERR_INVALID, "batch_run: null");
  if (!b->exec) return enqueue_batch(ctx, b);
  MBX_HIP(hipGraphLaunch(b->exec, ctx->stream));
  return MBX_OK;
}

int mbx_batch_run_eager(mbx_ctx *ctx, mbx_batch *b) {
  if (!ctx || !b) return fail(MBX_ERR_INVALID, "batch_run_eager: null");
  return enqueue_batch(ctx, b);
}

int mbx_batch_destroy(mbx_ctx *ctx, mbx_batch *b) {
  if (ctx) (void)hipStreamSynchronize(ctx->stream);
  destroy_batch(ctx, b);
  return MBX_OK;
}

int mbx_autotune(mbx_ctx *ctx, const mbx_query *qs, uint32_t nq,
                 mbx_tune *best, float *best_ms) {
  return mbx::autotune(ctx, qs, nq, best, best_ms, nullptr, nullptr, 0,
                       nullptr);
}

int mbx_autotune_report(mbx_ctx *ctx, const mbx_query *qs, uint32_t nq,
                        mbx_tune *best, float *best_ms, mbx_tune *cands,
                        float *cand_ms, int max, int *count) {
  if (!cands || !cand_ms || !count || max < 0)
    return fail(MBX_ERR_INVALID, "autotune_report: null");
  return mbx::autotune(ctx, qs, nq, best, best_ms, cands, cand_ms, max, count);
}

}  // extern "C"
