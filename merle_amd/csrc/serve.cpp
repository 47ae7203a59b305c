// serve.cpp -- a request loop over resident columns and bitmaps:
// memory-based admission and continuous batching on several HIP streams.
//
// Requests arrive at given times (a trace, e.g. Poisson arrivals).  Each
// arrived request is admitted when its output fits the memory budget (the
// output is then allocated from the pool); admitted requests wait for a free
// slot.  A slot is a HIP stream with a descriptor buffer: whenever one is
// free, up to max_batch admitted requests go out on it -- large ones as
// fused launches of their own, the rest as one persistent batched launch
// (the same kernel as mbx_batch_*) -- while the other slots are still
// running, so work of requests that arrived later overlaps on the device
// with work already in flight.  When a slot's completion event fires,
// its results are copied to the host, their memory is released and the next
// admitted requests take the slot.  Latency = result on the host - arrival.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <deque>
#include <thread>
#include <vector>

#include "batch.hpp"
#include "query_launch.hpp"

using namespace mbx;

namespace {

struct Slot {
  hipStream_t s = nullptr;
  hipEvent_t done = nullptr;
  BatchQ *h = nullptr;       // pinned staging of the descriptors
  BatchQ *d = nullptr;       // device descriptors (+ counter after them)
  uint32_t *counter = nullptr;
  unsigned long long *spread = nullptr;   // flush replicas of fused launches
  uint32_t *spread_done = nullptr;
  std::vector<uint32_t> qs;  // requests in flight on this slot
  bool busy = false;
};

// By default, requests at least this large (fact rows) go out as their own
// fused launches -- a whole-GPU grid each -- rather than as items of a
// batched launch, whose per-item passes suit many small queries.
constexpr uint32_t kServeFusedRows = 32u << 20;

double now_s(std::chrono::steady_clock::time_point t0) {
  return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0)
      .count();
}

size_t round256(size_t b) { return (b + 255) / 256 * 256; }

}  // namespace

extern "C" {

int mbx_serve_replay(mbx_ctx *ctx, const mbx_query *reqs, uint32_t nq,
                     const double *arrival_s, const mbx_serve_cfg *cfg,
                     int64_t *h_results, double *latency_s,
                     mbx_serve_stats *stats) {
  if (!ctx || !reqs || !arrival_s || !cfg || !h_results || !latency_s || !nq)
    return fail(MBX_ERR_INVALID, "serve: null argument");
  if (cfg->nstreams == 0 || cfg->nstreams > 16 || cfg->max_batch == 0)
    return fail(MBX_ERR_INVALID, "serve: need 1..16 streams and max_batch > 0");
  for (uint32_t i = 1; i < nq; ++i)
    if (arrival_s[i] < arrival_s[i - 1])
      return fail(MBX_ERR_INVALID, "serve: arrivals must be sorted");
  mbx_tune tune = cfg->tune;
  if (tune.tile_words == 0) mbx_default_tune(&tune);
  const bool plain = cfg->plain != 0;
  const uint32_t fused_rows = cfg->fused_rows ? cfg->fused_rows : kServeFusedRows;
  const uint32_t nslot = plain ? 1 : cfg->nstreams;
  const uint32_t maxb = plain ? 1 : cfg->max_batch;

  std::vector<size_t> off(nq + 1, 0);
  for (uint32_t i = 0; i < nq; ++i) off[i + 1] = off[i] + reqs[i].nr_grp;

  // slots
  std::vector<Slot> slots(nslot);
  const size_t dbytes = round256(sizeof(BatchQ) * maxb) + kCounterBytes;
  int st = MBX_OK;
  size_t used = 0, peak = 0;
  for (auto &sl : slots) {
    hipError_t e = hipStreamCreateWithFlags(&sl.s, hipStreamNonBlocking);
    if (e == hipSuccess) e = hipEventCreateWithFlags(&sl.done, hipEventDisableTiming);
    if (e == hipSuccess)
      e = hipHostMalloc(reinterpret_cast<void **>(&sl.h), sizeof(BatchQ) * maxb, 0);
    if (e != hipSuccess) {
      st = hip_fail(e, "serve: slot setup");
      break;
    }
    void *p = nullptr;
    st = ctx->pool->alloc(dbytes, &p);
    if (st) break;
    sl.d = static_cast<BatchQ *>(p);
    sl.counter = reinterpret_cast<uint32_t *>(static_cast<char *>(p) +
                                              round256(sizeof(BatchQ) * maxb));
    used += dbytes;
    // fixed working memory of the loop (like the context's own scratch),
    // outside the request budget
    if (!plain) {
      st = ctx->pool->alloc(spread_bytes(), &p);
      if (st) break;
      sl.spread = static_cast<unsigned long long *>(p);
      sl.spread_done = reinterpret_cast<uint32_t *>(
          static_cast<char *>(p) + (size_t)kSpreadRep * kLdsGroups * 8);
      st = launch_fill_u32(sl.s, sl.spread, 0, spread_bytes() / 4);
      if (st) break;
    }
  }
  const size_t budget = cfg->mem_budget ? cfg->mem_budget : (size_t)-1;
  std::vector<void *> outp(nq, nullptr);
  std::vector<size_t> need(nq, 0);
  std::deque<uint32_t> pending, admitted;
  uint32_t next = 0, finished = 0, rejected = 0, batches = 0, waited = 0;
  uint64_t batched_q = 0;
  const auto t0 = std::chrono::steady_clock::now();
  std::vector<mbx_query> tmp(maxb);
  std::vector<double> done_at(nq, -1.0);

  while (!st && finished + rejected < nq) {
    bool progress = false;
    const double t = now_s(t0);
    while (next < nq && arrival_s[next] <= t) {
      pending.push_back(next++);
      progress = true;
    }
    // memory-based admission (FIFO)
    while (!pending.empty()) {
      const uint32_t q = pending.front();
      need[q] = round256((size_t)reqs[q].nr_grp * 8);
      if (need[q] + slots.size() * dbytes > budget) {
        latency_s[q] = -1.0;   // can never fit
        ++rejected;
        pending.pop_front();
        continue;
      }
      if (used + need[q] > budget) {
        ++waited;
        break;
      }
      if (ctx->pool->alloc(need[q], &outp[q]) != MBX_OK) {
        ++waited;
        break;   // pool full right now: retry when memory is released
      }
      used += need[q];
      peak = std::max(peak, used);
      admitted.push_back(q);
      pending.pop_front();
      progress = true;
    }
    // completions
    for (auto &sl : slots) {
      if (!sl.busy) continue;
      hipError_t e = hipEventQuery(sl.done);
      if (e == hipErrorNotReady) continue;
      if (e != hipSuccess) {
        st = hip_fail(e, "serve: batch");
        break;
      }
      for (uint32_t q : sl.qs) {
        e = hipMemcpyAsync(h_results + off[q], outp[q],
                           (size_t)reqs[q].nr_grp * 8, hipMemcpyDeviceToHost,
                           sl.s);
        if (e != hipSuccess) break;
      }
      if (e == hipSuccess) e = hipStreamSynchronize(sl.s);
      if (e != hipSuccess) {
        st = hip_fail(e, "serve: results");
        break;
      }
      const double tc = now_s(t0);
      for (uint32_t q : sl.qs) {
        latency_s[q] = tc - arrival_s[q];
        done_at[q] = tc;
        ctx->pool->free(outp[q]);
        used -= need[q];
        ++finished;
      }
      sl.qs.clear();
      sl.busy = false;
      progress = true;
    }
    if (st) break;
    // dispatch admitted requests onto free slots
    for (auto &sl : slots) {
      if (sl.busy || admitted.empty()) continue;
      const uint32_t k = std::min<uint32_t>(maxb, (uint32_t)admitted.size());
      sl.qs.assign(admitted.begin(), admitted.begin() + k);
      admitted.erase(admitted.begin(), admitted.begin() + k);
      for (uint32_t i = 0; i < k; ++i) {
        tmp[i] = reqs[sl.qs[i]];
        tmp[i].d_out = static_cast<int64_t *>(outp[sl.qs[i]]);
      }
      hipStream_t keep = ctx->stream;
      if (plain) {
        ctx->stream = sl.s;
        st = mbx_prog_grpby(ctx, &tmp[0].prg, &tmp[0].op, tmp[0].nr_grp,
                            tmp[0].d_out,
                            MBX_PLAN_FUSED | (tmp[0].flags & MBX_PLAN_EAGER));
        ctx->stream = keep;
      } else {
        // large requests: one fused launch each (slot-private flush
        // replicas, so slots overlap safely); the rest: one batched launch
        uint32_t nsmall = 0;
        ctx->stream = sl.s;
        for (uint32_t i = 0; i < k && !st; ++i) {
          if (tmp[i].prg.factsz < fused_rows) {
            tmp[nsmall++] = tmp[i];
            continue;
          }
          QueryPlan qp;
          st = prepare(ctx, &tmp[i].prg, &tmp[i].op, tmp[i].nr_grp,
                       tmp[i].d_out, &qp);
          if (st) break;
          qp.eager = (tmp[i].flags & MBX_PLAN_EAGER) != 0;
          qp.spread = sl.spread;
          qp.spread_done = sl.spread_done;
          st = launch_fill_u32(sl.s, tmp[i].d_out, 0, 2ull * tmp[i].nr_grp);
          if (!st) st = launch_fused(ctx, qp, nullptr);
        }
        ctx->stream = keep;
        BatchShape shape;
        const uint32_t kb = nsmall;
        if (!st && kb)
          st = batch_prepare(ctx, tmp.data(), kb, tune, sl.h, &shape);
        if (!st && kb) {
          hipError_t e = hipMemcpyAsync(sl.d, sl.h, sizeof(BatchQ) * kb,
                                        hipMemcpyHostToDevice, sl.s);
          if (e != hipSuccess) st = hip_fail(e, "serve: descriptors");
        }
        if (!st && kb)
          st = batch_launch(ctx, sl.s, sl.d, shape, sl.counter, tune);
      }
      if (!st) {
        hipError_t e = hipEventRecord(sl.done, sl.s);
        if (e != hipSuccess) st = hip_fail(e, "serve: event");
      }
      if (st) break;
      sl.busy = true;
      ++batches;
      batched_q += k;
      progress = true;
    }
    if (!progress) std::this_thread::yield();
  }

  for (auto &sl : slots) {
    if (sl.s) (void)hipStreamSynchronize(sl.s);
    if (sl.d) ctx->pool->free(sl.d);
    if (sl.spread) ctx->pool->free(sl.spread);
    if (sl.h) (void)hipHostFree(sl.h);
    if (sl.done) (void)hipEventDestroy(sl.done);
    if (sl.s) (void)hipStreamDestroy(sl.s);
  }
  for (uint32_t q = 0; q < nq; ++q)
    if (outp[q] && done_at[q] < 0) ctx->pool->free(outp[q]);
  if (st) return st;
  if (stats) {
    std::memset(stats, 0, sizeof(*stats));
    std::vector<double> lat;
    double last = 0;
    for (uint32_t q = 0; q < nq; ++q)
      if (latency_s[q] >= 0) {
        lat.push_back(latency_s[q]);
        last = std::max(last, done_at[q]);
      }
    std::sort(lat.begin(), lat.end());
    auto pct = [&](double p) {
      if (lat.empty()) return 0.0;
      size_t i = (size_t)std::ceil(p * lat.size()) - 1;
      return lat[std::min(i, lat.size() - 1)];
    };
    double sum = 0;
    for (double v : lat) sum += v;
    stats->completed = (uint32_t)lat.size();
    stats->rejected = rejected;
    stats->wall_s = last - (nq ? arrival_s[0] : 0.0);
    stats->p50_s = pct(0.50);
    stats->p99_s = pct(0.99);
    stats->mean_s = lat.empty() ? 0 : sum / lat.size();
    stats->max_s = lat.empty() ? 0 : lat.back();
    stats->batches = batches;
    stats->mean_batch = batches ? (double)batched_q / batches : 0;
    stats->peak_mem = peak;
    stats->admission_waits = waited;
  }
  return MBX_OK;
}

}  // extern "C"
