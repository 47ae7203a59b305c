// serve.cpp -- a request server over resident columns and bitmaps:
// memory-based admission and continuous batching on several HIP streams.
//
// Threads of a server:
//  - callers submit, poll, wait and cancel from any thread;
//  - one dispatcher runs the scheduler (sched.hpp): admission by memory,
//    then, whenever a slot is free, admitted requests go out on it -- a
//    large or scan request alone, as a launch of its own; else up to
//    max_batch small ones as one persistent batched launch (the kernel of
//    mbx_batch_*) -- while the other slots are still running, so work of
//    requests that arrived later overlaps on the device with work already in
//    flight.  Enqueueing is all
//    it does with the GPU: it never waits for a kernel or a copy;
//  - one completion thread per slot waits for the slot's event, copies the
//    results (already moved by the slot's stream into pinned staging) into
//    the callers' buffers, and hands the slot back to the scheduler.
// Latency = result in the caller's buffer - arrival.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <mutex>
#include <thread>
#include <vector>

#include "batch.hpp"
#include "query_launch.hpp"
#include "sched.hpp"

using namespace mbx;

namespace {

// By default, requests at least this large (fact rows) go out as their own
// fused launches -- a whole-GPU grid each -- rather than as items of a
// batched launch, whose per-item passes suit many small queries.
constexpr uint32_t kServeFusedRows = 32u << 20;

size_t round256(size_t b) { return (b + 255) / 256 * 256; }

using Clock = std::chrono::steady_clock;

struct ReqRec {
  mbx_query q;
  int64_t *h_result = nullptr;
  Clock::time_point arrival;
  double latency = -1.0;
  double done_at = -1.0;            // seconds since the server opened
  hipEvent_t ready = nullptr;       // recorded on the context's stream
  void *mem = nullptr;              // pool block of the output
};

struct SlotRec {
  hipStream_t s = nullptr;
  hipEvent_t done = nullptr;
  BatchQ *h = nullptr;              // pinned staging of the descriptors
  BatchQ *d = nullptr;              // device descriptors, counters after them
  uint32_t *counter = nullptr;
  unsigned long long *spread = nullptr;   // flush replicas of fused launches
  uint32_t *spread_done = nullptr;
  int64_t *stage = nullptr;         // pinned staging of the results
  size_t stage_cap = 0;             // int64 entries
  std::vector<uint32_t> reqs;       // requests of the work in flight
  std::vector<size_t> offs;         // their offsets in stage
  bool launched = false;            // work in flight for the completion thread
  std::thread waiter;
  std::condition_variable cv;
};

}  // namespace

struct mbx_server : SchedDevice {
  mbx_ctx *ctx = nullptr;
  mbx_tune tune{};
  bool plain = false;
  uint32_t fused_rows = kServeFusedRows;
  uint32_t maxb = 1;
  size_t fixed = 0;
  Clock::time_point t0;

  std::mutex mu;
  std::condition_variable cv_disp;  // dispatcher: new work or a free slot
  std::condition_variable cv_done;  // waiters: a request ended
  Scheduler *sched = nullptr;
  std::deque<ReqRec> reqs;          // by ticket (stable addresses)
  std::vector<SlotRec> slots;
  std::thread dispatcher;
  uint64_t gen = 0;                 // bumped on every event the dispatcher acts on
  bool stop = false;
  int error = MBX_OK;               // first HIP error
  std::string error_msg;
  std::vector<mbx_query> tmp;

  double since_open(Clock::time_point t) const {
    return std::chrono::duration<double>(t - t0).count();
  }

  // ---- SchedDevice (called by the scheduler, mu held)
  int reserve(uint32_t q, size_t bytes) override {
    void *p = nullptr;
    const int st = ctx->pool->alloc(bytes, &p);
    if (st != MBX_OK) return MBX_ERR_OOM;
    reqs[q].mem = p;
    reqs[q].q.d_out = static_cast<int64_t *>(p);
    return MBX_OK;
  }

  void release(uint32_t q) override {
    if (reqs[q].mem) ctx->pool->free(reqs[q].mem);
    reqs[q].mem = nullptr;
  }

  int launch(uint32_t slot, const uint32_t *ids, uint32_t k) override {
    SlotRec &sl = slots[slot];
    int st = enqueue(sl, ids, k);
    if (st != MBX_OK) {
      // nothing may still run on the slot when its requests are failed
      (void)hipStreamSynchronize(sl.s);
      note_error(st);
      return st;
    }
    sl.reqs.assign(ids, ids + k);
    sl.launched = true;
    sl.cv.notify_one();
    return MBX_OK;
  }

  void note_error(int st) {
    if (error == MBX_OK) {
      error = st;
      error_msg = mbx_last_error();
    }
  }

  // Enqueues the work of requests ids[0..k) on slot sl: the wait for the
  // caller's earlier work, the launches, the result copies and the event.
  int enqueue(SlotRec &sl, const uint32_t *ids, uint32_t k) {
    uint32_t newest = ids[0];
    size_t total = 0;
    sl.offs.assign(k, 0);
    for (uint32_t i = 0; i < k; ++i) {
      newest = std::max(newest, ids[i]);
      sl.offs[i] = total;
      total += reqs[ids[i]].q.nr_grp;
    }
    // requests are ordered after the caller's work on the context's stream
    // (events recorded in submit order, so the newest covers the others)
    MBX_HIP(hipStreamWaitEvent(sl.s, reqs[newest].ready, 0));
    if (total > sl.stage_cap) {
      if (sl.stage) (void)hipHostFree(sl.stage);
      sl.stage = nullptr;
      sl.stage_cap = 0;
      const size_t cap = std::max<size_t>(total, 4096);
      MBX_HIP(hipHostMalloc(reinterpret_cast<void **>(&sl.stage), cap * 8, 0));
      sl.stage_cap = cap;
    }
    tmp.resize(k);
    for (uint32_t i = 0; i < k; ++i) tmp[i] = reqs[ids[i]].q;
    int st = MBX_OK;
    uint32_t nsmall = 0;
    for (uint32_t i = 0; i < k && !st; ++i) {
      // a scan request as a scan launch of its own; plain: every other
      // request as one fused launch; else large ones only
      const bool scan = (tmp[i].flags & MBX_PLAN_MASK) == MBX_PLAN_SCAN;
      if (!scan && !plain && tmp[i].prg.factsz < fused_rows) {
        tmp[nsmall++] = tmp[i];
        continue;
      }
      QueryPlan qp;
      st = prepare(ctx, &tmp[i].prg, &tmp[i].op, tmp[i].nr_grp, tmp[i].d_out,
                   &qp);
      if (st) break;
      qp.eager = (tmp[i].flags & MBX_PLAN_EAGER) != 0;
      qp.spread = sl.spread;
      qp.spread_done = sl.spread_done;
      qp.stream = sl.s;
      st = launch_fill_u64(sl.s, tmp[i].d_out, 0, tmp[i].nr_grp);
      if (!st) st = scan ? launch_scan(ctx, qp) : launch_fused(ctx, qp, nullptr);
    }
    if (!st && nsmall) {
      BatchShape shape;
      st = batch_prepare(ctx, tmp.data(), nsmall, tune, sl.h, &shape);
      if (!st) {
        hipError_t e = hipMemcpyAsync(sl.d, sl.h, sizeof(BatchQ) * nsmall,
                                      hipMemcpyHostToDevice, sl.s);
        if (e != hipSuccess) st = hip_fail(e, "serve: descriptors");
      }
      if (!st) st = batch_launch(ctx, sl.s, sl.d, shape, sl.counter, tune);
    }
    for (uint32_t i = 0; i < k && !st; ++i) {
      const ReqRec &r = reqs[ids[i]];
      hipError_t e = hipMemcpyAsync(sl.stage + sl.offs[i], r.q.d_out,
                                    (size_t)r.q.nr_grp * 8,
                                    hipMemcpyDeviceToHost, sl.s);
      if (e != hipSuccess) st = hip_fail(e, "serve: result copy");
    }
    if (!st) {
      hipError_t e = hipEventRecord(sl.done, sl.s);
      if (e != hipSuccess) st = hip_fail(e, "serve: event");
    }
    return st;
  }

  // ---- threads
  void dispatch_loop() {
    std::unique_lock<std::mutex> lk(mu);
    for (;;) {
      if (stop) break;
      const uint64_t seen = gen;
      (void)sched->step(*this);   // errors fail their requests
      cv_done.notify_all();       // rejections and failures end requests
      if (stop) break;
      // a request waiting for the pool (not the budget) is retried now and
      // then: the caller may free pool memory without telling the server
      const bool waiting = sched->has_queued();
      auto pred = [&] { return stop || gen != seen; };
      if (waiting)
        cv_disp.wait_for(lk, std::chrono::milliseconds(1), pred);
      else
        cv_disp.wait(lk, pred);
    }
  }

  void slot_loop(uint32_t s) {
    SlotRec &sl = slots[s];
    std::unique_lock<std::mutex> lk(mu);
    for (;;) {
      sl.cv.wait(lk, [&] { return sl.launched || stop; });
      if (!sl.launched) break;    // stop, and nothing in flight
      const std::vector<uint32_t> ids = sl.reqs;
      std::vector<std::pair<int64_t *, size_t>> dst;   // caller buffer, bytes
      for (uint32_t q : ids)
        dst.emplace_back(reqs[q].h_result, (size_t)reqs[q].q.nr_grp * 8);
      const std::vector<size_t> offs = sl.offs;
      const int64_t *stage = sl.stage;
      lk.unlock();
      hipError_t e = hipEventSynchronize(sl.done);
      if (e == hipSuccess)
        for (size_t i = 0; i < ids.size(); ++i)
          std::memcpy(dst[i].first, stage + offs[i], dst[i].second);
      const Clock::time_point tc = Clock::now();
      lk.lock();
      sl.launched = false;
      if (e != hipSuccess) {
        note_error(hip_fail(e, "serve: slot work"));
        sched->fail_slot(s, *this);
      } else {
        for (uint32_t q : ids) {
          reqs[q].latency =
              std::chrono::duration<double>(tc - reqs[q].arrival).count();
          reqs[q].done_at = since_open(tc);
        }
        sched->finish(s, *this, nullptr);
      }
      ++gen;
      cv_disp.notify_one();
      cv_done.notify_all();
    }
  }

  int submit_at(const mbx_query *q, int64_t *h_result, Clock::time_point at,
                uint64_t *ticket) {
    if (!q || !h_result || !ticket)
      return fail(MBX_ERR_INVALID, "serve_submit: null argument");
    if (q->nr_grp == 0) return fail(MBX_ERR_INVALID, "serve_submit: nr_grp == 0");
    hipEvent_t ev = nullptr;
    MBX_HIP(hipEventCreateWithFlags(&ev, hipEventDisableTiming));
    std::lock_guard<std::mutex> g(mu);
    hipError_t e = hipEventRecord(ev, ctx->stream);
    if (e != hipSuccess) {
      (void)hipEventDestroy(ev);
      return hip_fail(e, "serve_submit: event");
    }
    ReqRec r;
    r.q = *q;
    r.q.d_out = nullptr;
    r.h_result = h_result;
    r.arrival = at;
    r.ready = ev;
    reqs.push_back(r);
    // large and scan requests run as launches of their own: dispatched
    // alone, they hold up no batch mates on their slot
    const bool solo = plain || (q->flags & MBX_PLAN_MASK) == MBX_PLAN_SCAN ||
                      q->prg.factsz >= fused_rows;
    const uint32_t id = sched->add(round256((size_t)q->nr_grp * 8), solo);
    *ticket = id;
    ++gen;
    cv_disp.notify_one();
    return MBX_OK;
  }
};

namespace {

// Fixed working memory of one slot: descriptors + work counters, then the
// flush replicas of its fused launches.
size_t slot_fixed_bytes(size_t maxb) {
  return round256(sizeof(BatchQ) * maxb) + kCounterBytes +
         round256(spread_bytes());
}

int check_cfg(const mbx_serve_cfg *cfg) {
  if (!cfg) return fail(MBX_ERR_INVALID, "serve: null config");
  if (cfg->nstreams == 0 || cfg->nstreams > 16 || cfg->max_batch == 0)
    return fail(MBX_ERR_INVALID, "serve: need 1..16 streams and max_batch > 0");
  return MBX_OK;
}

void free_slot(mbx_server *s, SlotRec &sl) {
  if (sl.s) (void)hipStreamSynchronize(sl.s);
  if (sl.d) s->ctx->pool->free(sl.d);
  if (sl.spread) s->ctx->pool->free(sl.spread);
  if (sl.h) (void)hipHostFree(sl.h);
  if (sl.stage) (void)hipHostFree(sl.stage);
  if (sl.done) (void)hipEventDestroy(sl.done);
  if (sl.s) (void)hipStreamDestroy(sl.s);
  sl.s = nullptr;
  sl.done = nullptr;
  sl.h = sl.d = nullptr;
  sl.spread = nullptr;
  sl.stage = nullptr;
  sl.stage_cap = 0;
}

void destroy_server(mbx_server *s) {
  {
    std::lock_guard<std::mutex> g(s->mu);
    s->stop = true;
    s->cv_disp.notify_all();
    for (auto &sl : s->slots) sl.cv.notify_all();
  }
  if (s->dispatcher.joinable()) s->dispatcher.join();
  for (auto &sl : s->slots)
    if (sl.waiter.joinable()) sl.waiter.join();
  for (auto &sl : s->slots) free_slot(s, sl);
  for (auto &r : s->reqs) {
    if (r.mem) s->ctx->pool->free(r.mem);
    if (r.ready) (void)hipEventDestroy(r.ready);
  }
  delete s->sched;
  delete s;
}

ReqRec *find(mbx_server *s, uint64_t t) {
  return t < s->reqs.size() ? &s->reqs[t] : nullptr;
}

// Waits until every request has ended, fills stats and the latencies of
// tickets 0..nlat, then frees the server.
int close_server(mbx_server *s, mbx_serve_stats *stats, double *lat_out,
                 size_t nlat) {
  {
    std::unique_lock<std::mutex> lk(s->mu);
    s->cv_done.wait(lk, [&] { return s->sched->idle(); });
  }
  const int st = s->error;
  const std::string msg = s->error_msg;
  for (size_t i = 0; i < nlat && i < s->reqs.size(); ++i)
    lat_out[i] = s->reqs[i].latency;
  if (stats) {
    std::memset(stats, 0, sizeof(*stats));
    std::vector<double> lat;
    double first = 1e300, last = 0;
    for (const auto &r : s->reqs) {
      first = std::min(first, s->since_open(r.arrival));
      if (r.latency >= 0) {
        lat.push_back(r.latency);
        last = std::max(last, r.done_at);
      }
    }
    std::sort(lat.begin(), lat.end());
    auto pct = [&](double p) {
      if (lat.empty()) return 0.0;
      size_t i = (size_t)std::ceil(p * lat.size()) - 1;
      return lat[std::min(i, lat.size() - 1)];
    };
    double sum = 0;
    for (double v : lat) sum += v;
    const Scheduler &sc = *s->sched;
    stats->completed = sc.completed();
    stats->rejected = sc.rejected();
    stats->cancelled = sc.cancelled();
    stats->failed = sc.failed();
    stats->wall_s = lat.empty() ? 0.0 : last - first;
    stats->p50_s = pct(0.50);
    stats->p99_s = pct(0.99);
    stats->mean_s = lat.empty() ? 0 : sum / lat.size();
    stats->max_s = lat.empty() ? 0 : lat.back();
    stats->batches = sc.launches();
    stats->mean_batch =
        sc.launches() ? (double)sc.launched_reqs() / sc.launches() : 0;
    stats->peak_mem = sc.peak();
    stats->admission_waits = sc.delayed();
    stats->fixed_mem = s->fixed;
  }
  destroy_server(s);
  if (st) set_error(msg);
  return st;
}

}  // namespace

extern "C" {

int mbx_serve_fixed_mem(const mbx_serve_cfg *cfg, uint64_t *bytes) {
  int st = check_cfg(cfg);
  if (st) return st;
  if (!bytes) return fail(MBX_ERR_INVALID, "serve_fixed_mem: null");
  const uint64_t nslot = cfg->plain ? 1 : cfg->nstreams;
  const uint64_t maxb = cfg->plain ? 1 : cfg->max_batch;
  *bytes = nslot * slot_fixed_bytes(maxb);
  return MBX_OK;
}

int mbx_serve_open(mbx_ctx *ctx, const mbx_serve_cfg *cfg, mbx_server **out) {
  if (!ctx || !out) return fail(MBX_ERR_INVALID, "serve_open: null argument");
  *out = nullptr;
  int st = check_cfg(cfg);
  if (st) return st;
  auto *s = new mbx_server;
  s->ctx = ctx;
  s->tune = cfg->tune;
  if (s->tune.tile_words == 0) mbx_default_tune(&s->tune);
  s->plain = cfg->plain != 0;
  s->fused_rows = cfg->fused_rows ? cfg->fused_rows : kServeFusedRows;
  const uint32_t nslot = s->plain ? 1 : cfg->nstreams;
  s->maxb = s->plain ? 1 : cfg->max_batch;
  s->slots = std::vector<SlotRec>(nslot);
  const size_t dbytes = round256(sizeof(BatchQ) * s->maxb) + kCounterBytes;
  for (auto &sl : s->slots) {
    hipError_t e = hipStreamCreateWithFlags(&sl.s, hipStreamNonBlocking);
    if (e == hipSuccess)
      e = hipEventCreateWithFlags(&sl.done, hipEventDisableTiming);
    if (e == hipSuccess)
      e = hipHostMalloc(reinterpret_cast<void **>(&sl.h),
                        sizeof(BatchQ) * s->maxb, 0);
    if (e != hipSuccess) {
      st = hip_fail(e, "serve_open: slot setup");
      break;
    }
    void *p = nullptr;
    st = ctx->pool->alloc(dbytes, &p);
    if (st) break;
    sl.d = static_cast<BatchQ *>(p);
    sl.counter = reinterpret_cast<uint32_t *>(static_cast<char *>(p) +
                                              round256(sizeof(BatchQ) * s->maxb));
    st = ctx->pool->alloc(spread_bytes(), &p);
    if (st) break;
    sl.spread = static_cast<unsigned long long *>(p);
    sl.spread_done = reinterpret_cast<uint32_t *>(
        static_cast<char *>(p) + (size_t)kSpreadRep * kLdsGroups * 8);
    st = launch_fill_u32(sl.s, sl.spread, 0, spread_bytes() / 4);
    if (st) break;
    s->fixed += slot_fixed_bytes(s->maxb);
  }
  if (!st && cfg->mem_budget && s->fixed > cfg->mem_budget)
    st = fail(MBX_ERR_ADMISSION,
              "serve_open: the server's fixed memory (" +
                  std::to_string(s->fixed) + " B) exceeds mem_budget");
  if (st) {
    for (auto &sl : s->slots) free_slot(s, sl);
    delete s;
    return st;
  }
  s->sched = new Scheduler(nslot, s->maxb, cfg->mem_budget, s->fixed);
  s->t0 = Clock::now();
  s->dispatcher = std::thread([s] { s->dispatch_loop(); });
  for (uint32_t i = 0; i < nslot; ++i)
    s->slots[i].waiter = std::thread([s, i] { s->slot_loop(i); });
  *out = s;
  return MBX_OK;
}

int mbx_serve_submit(mbx_server *s, const mbx_query *q, int64_t *h_result,
                     uint64_t *ticket) {
  if (!s) return fail(MBX_ERR_INVALID, "serve_submit: null server");
  return s->submit_at(q, h_result, Clock::now(), ticket);
}

int mbx_serve_poll(mbx_server *s, uint64_t ticket, int *state,
                   double *latency_s) {
  if (!s) return fail(MBX_ERR_INVALID, "serve_poll: null server");
  std::lock_guard<std::mutex> g(s->mu);
  ReqRec *r = find(s, ticket);
  if (!r) return fail(MBX_ERR_INVALID, "serve_poll: unknown ticket");
  if (state) *state = s->sched->state((uint32_t)ticket);
  if (latency_s) *latency_s = r->latency;
  return MBX_OK;
}

int mbx_serve_wait(mbx_server *s, uint64_t ticket, double timeout_s,
                   int *state, double *latency_s) {
  if (!s) return fail(MBX_ERR_INVALID, "serve_wait: null server");
  std::unique_lock<std::mutex> lk(s->mu);
  if (!find(s, ticket)) return fail(MBX_ERR_INVALID, "serve_wait: unknown ticket");
  auto ended = [&] {
    return s->sched->state((uint32_t)ticket) >= MBX_REQ_DONE;
  };
  if (timeout_s < 0)
    s->cv_done.wait(lk, ended);
  else
    s->cv_done.wait_for(lk, std::chrono::duration<double>(timeout_s), ended);
  if (state) *state = s->sched->state((uint32_t)ticket);
  if (latency_s) *latency_s = s->reqs[ticket].latency;
  return MBX_OK;
}

int mbx_serve_cancel(mbx_server *s, uint64_t ticket) {
  if (!s) return fail(MBX_ERR_INVALID, "serve_cancel: null server");
  std::lock_guard<std::mutex> g(s->mu);
  if (!find(s, ticket)) return fail(MBX_ERR_INVALID, "serve_cancel: unknown ticket");
  if (s->sched->cancel((uint32_t)ticket, *s) != MBX_OK)
    return fail(MBX_ERR_INVALID, "serve_cancel: the request is running or over");
  ++s->gen;
  s->cv_disp.notify_one();
  s->cv_done.notify_all();
  return MBX_OK;
}

int mbx_serve_close(mbx_server *s, mbx_serve_stats *stats) {
  if (!s) return fail(MBX_ERR_INVALID, "serve_close: null server");
  return close_server(s, stats, nullptr, 0);
}

int mbx_serve_replay(mbx_ctx *ctx, const mbx_query *reqs, uint32_t nq,
                     const double *arrival_s, const mbx_serve_cfg *cfg,
                     int64_t *h_results, double *latency_s,
                     mbx_serve_stats *stats) {
  if (!ctx || !reqs || !arrival_s || !cfg || !h_results || !latency_s || !nq)
    return fail(MBX_ERR_INVALID, "serve: null argument");
  for (uint32_t i = 1; i < nq; ++i)
    if (arrival_s[i] < arrival_s[i - 1])
      return fail(MBX_ERR_INVALID, "serve: arrivals must be sorted");
  for (uint32_t i = 0; i < nq; ++i)
    if (reqs[i].nr_grp == 0) return fail(MBX_ERR_INVALID, "serve: nr_grp == 0");
  mbx_server *s = nullptr;
  int st = mbx_serve_open(ctx, cfg, &s);
  if (st) return st;
  const Clock::time_point t0 = Clock::now();
  size_t off = 0;
  std::vector<uint64_t> tickets(nq);
  for (uint32_t i = 0; i < nq && !st; ++i) {
    const auto at = t0 + std::chrono::duration_cast<Clock::duration>(
                             std::chrono::duration<double>(arrival_s[i]));
    for (;;) {   // submit at the arrival time
      const auto now = Clock::now();
      if (now >= at) break;
      if (at - now > std::chrono::microseconds(200))
        std::this_thread::sleep_for(std::chrono::microseconds(50));
      else
        std::this_thread::yield();
    }
    st = s->submit_at(&reqs[i], h_results + off, at, &tickets[i]);
    off += reqs[i].nr_grp;
  }
  const int cst = close_server(s, stats, latency_s, nq);
  return st ? st : cst;
}

}  // extern "C"
