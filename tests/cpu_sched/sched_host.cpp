// The serving scheduler (merle_amd/csrc/sched.hpp) against a fake device on
// the CPU: reservations come from a fixed mbx_hostpool (the library's own
// allocator), launches only record what they were given, and the test
// decides when each slot finishes.  Each scenario prints one line; the exit
// status is the number of failed checks.
#include <cstdio>
#include <cstdlib>
#include <algorithm>
#include <map>
#include <random>
#include <set>
#include <vector>

#include "../../merle_amd/csrc/sched.hpp"

using mbx::SchedDevice;
using mbx::Scheduler;

static int g_fail = 0;
#define CHECK(c)                                                             \
  do {                                                                       \
    if (!(c)) {                                                              \
      std::printf("  check failed at line %d: %s\n", __LINE__, #c);         \
      ++g_fail;                                                              \
    }                                                                        \
  } while (0)

struct FakeDevice : SchedDevice {
  mbx_hostpool *pool = nullptr;
  std::map<uint32_t, void *> held;          // request -> reservation
  std::vector<std::vector<uint32_t>> launches_on;   // per slot
  std::set<uint32_t> launched_ever;
  int fail_slot = -1;                       // launches on this slot fail
  int twice = 0;                            // a request launched twice

  FakeDevice(size_t cap, uint32_t nslots) : launches_on(nslots) {
    mbx_hostpool_create(cap, &pool);
  }
  ~FakeDevice() override { mbx_hostpool_destroy(pool); }
  int reserve(uint32_t q, size_t bytes) override {
    void *p = nullptr;
    if (mbx_hostpool_alloc(pool, bytes, &p) != MBX_OK) return MBX_ERR_OOM;
    held[q] = p;
    return MBX_OK;
  }
  void release(uint32_t q) override {
    auto it = held.find(q);
    if (it == held.end()) return;
    mbx_hostpool_free(pool, it->second);
    held.erase(it);
  }
  int launch(uint32_t slot, const uint32_t *r, uint32_t n) override {
    if ((int)slot == fail_slot) return MBX_ERR_HIP;
    for (uint32_t i = 0; i < n; ++i) {
      if (!launched_ever.insert(r[i]).second) ++twice;
      launches_on[slot].push_back(r[i]);
    }
    return MBX_OK;
  }
  size_t pool_used() {
    size_t used, cap, peak, nlive, largest;
    mbx_hostpool_stats(pool, &used, &cap, &peak, &nlive, &largest);
    return used;
  }
};

// The pool refuses a request while nothing is running or admitted: the
// request is rejected (it would otherwise wait forever), and the next one,
// which fits, is served.
static void pool_full_nothing_running() {
  FakeDevice d(4096, 2);
  void *hog = nullptr;
  mbx_hostpool_alloc(d.pool, 3584, &hog);   // memory the caller keeps
  Scheduler s(2, 4, 0, 0);
  const uint32_t a = s.add(1024);
  const uint32_t b = s.add(256);
  s.step(d);
  CHECK(s.state(a) == MBX_REQ_REJECTED);
  CHECK(s.state(b) == MBX_REQ_RUNNING);
  CHECK(s.rejected() == 1);
  s.finish(s.slot_of(b), d, nullptr);
  CHECK(s.idle() && d.held.empty());
  mbx_hostpool_free(d.pool, hog);
  std::printf("pool_full_nothing_running: rejected %u, completed %u\n",
              s.rejected(), s.completed());
}

// The pool refuses a request while another runs: it waits (counted once,
// however many passes it waits) and is served when memory is released.
static void pool_full_waits_for_running() {
  FakeDevice d(2048, 1);
  Scheduler s(1, 1, 0, 0);
  const uint32_t a = s.add(1536);
  const uint32_t b = s.add(1024);
  for (int i = 0; i < 5; ++i) s.step(d);
  CHECK(s.state(a) == MBX_REQ_RUNNING);
  CHECK(s.state(b) == MBX_REQ_QUEUED);
  CHECK(s.delayed() == 1);
  s.finish(0, d, nullptr);
  s.step(d);
  CHECK(s.state(b) == MBX_REQ_RUNNING);
  s.finish(0, d, nullptr);
  CHECK(s.idle() && s.completed() == 2 && s.rejected() == 0);
  std::printf("pool_full_waits_for_running: delayed %u\n", s.delayed());
}

// The budget (fixed memory first) binds: FIFO admission, requests that can
// never fit are rejected at once, the others wait for memory.
static void budget_fifo() {
  FakeDevice d(1 << 20, 1);
  const size_t fixed = 1000;
  Scheduler s(1, 4, fixed + 1024, fixed);
  const uint32_t big = s.add(2048);      // never fits
  const uint32_t r1 = s.add(512), r2 = s.add(512), r3 = s.add(512);
  const uint32_t r4 = s.add(256);
  s.step(d);
  s.step(d);
  CHECK(s.state(big) == MBX_REQ_REJECTED);
  CHECK(s.state(r1) == MBX_REQ_RUNNING && s.state(r2) == MBX_REQ_RUNNING);
  CHECK(s.state(r3) == MBX_REQ_QUEUED);
  CHECK(s.state(r4) == MBX_REQ_QUEUED);  // FIFO: no overtaking r3
  CHECK(d.launches_on[0].size() == 2);
  CHECK(s.used() == fixed + 1024 && s.peak() == fixed + 1024);
  CHECK(s.delayed() == 1);
  s.finish(0, d, nullptr);
  s.step(d);
  CHECK(s.state(r3) == MBX_REQ_RUNNING && s.state(r4) == MBX_REQ_RUNNING);
  s.finish(0, d, nullptr);
  CHECK(s.idle() && s.used() == fixed && s.completed() == 4);
  std::printf("budget_fifo: rejected %u, delayed %u, peak %zu\n",
              s.rejected(), s.delayed(), s.peak());
}

// Slot assignment: every free slot takes up to max_batch admitted requests
// in admission order.
static void slots_and_batches() {
  FakeDevice d(1 << 20, 3);
  Scheduler s(3, 2, 0, 0);
  for (int i = 0; i < 7; ++i) s.add(256);
  s.step(d);
  CHECK(d.launches_on[0] == std::vector<uint32_t>({0, 1}));
  CHECK(d.launches_on[1] == std::vector<uint32_t>({2, 3}));
  CHECK(d.launches_on[2] == std::vector<uint32_t>({4, 5}));
  CHECK(s.state(6) == MBX_REQ_ADMITTED);
  std::vector<uint32_t> done;
  s.finish(1, d, &done);
  CHECK(done == std::vector<uint32_t>({2, 3}));
  s.step(d);
  CHECK(s.state(6) == MBX_REQ_RUNNING && s.slot_of(6) == 1);
  s.finish(0, d, nullptr);
  s.finish(1, d, nullptr);
  s.finish(2, d, nullptr);
  CHECK(s.idle() && s.launches() == 4 && s.launched_reqs() == 7);
  std::printf("slots_and_batches: %u launches\n", s.launches());
}

// Solo requests (launches of their own) go out alone, in FIFO order, and
// batches stop in front of them.
static void solo_requests() {
  FakeDevice d(1 << 20, 2);
  Scheduler s(2, 4, 0, 0);
  s.add(256);
  s.add(256);
  s.add(256, true);
  s.add(256);
  s.add(256, true);
  s.add(256, true);
  s.step(d);
  CHECK(d.launches_on[0] == std::vector<uint32_t>({0, 1}));
  CHECK(d.launches_on[1] == std::vector<uint32_t>({2}));
  s.finish(1, d, nullptr);
  s.step(d);
  CHECK(d.launches_on[1] == std::vector<uint32_t>({2, 3}));
  s.finish(0, d, nullptr);
  s.step(d);
  CHECK(d.launches_on[0] == std::vector<uint32_t>({0, 1, 4}));
  s.finish(1, d, nullptr);
  s.step(d);
  CHECK(d.launches_on[1] == std::vector<uint32_t>({2, 3, 5}));
  s.finish(0, d, nullptr);
  s.finish(1, d, nullptr);
  CHECK(s.idle() && s.launches() == 5);
  std::printf("solo_requests: %u launches\n", s.launches());
}

// Cancel: a queued or admitted request ends cancelled (an admitted one
// gives its memory back); a running one is left alone.
static void cancel_states() {
  FakeDevice d(1 << 20, 1);
  Scheduler s(1, 1, 0, 0);
  const uint32_t a = s.add(256), b = s.add(512), c = s.add(768);
  s.step(d);                     // a runs, b and c admitted
  CHECK(s.state(b) == MBX_REQ_ADMITTED);
  CHECK(s.cancel(a, d) == MBX_ERR_INVALID);
  const size_t before = d.pool_used();
  CHECK(s.cancel(b, d) == MBX_OK && s.state(b) == MBX_REQ_CANCELLED);
  CHECK(d.pool_used() < before);
  const uint32_t e = s.add(256);
  CHECK(s.cancel(e, d) == MBX_OK);
  s.finish(0, d, nullptr);
  s.step(d);
  CHECK(s.state(c) == MBX_REQ_RUNNING);
  s.finish(0, d, nullptr);
  CHECK(s.idle() && s.cancelled() == 2 && s.completed() == 2);
  CHECK(d.held.empty());
  std::printf("cancel_states: cancelled %u\n", s.cancelled());
}

// A launch that fails ends its requests as failed and frees the slot.
static void launch_failure() {
  FakeDevice d(1 << 20, 2);
  d.fail_slot = 0;
  Scheduler s(2, 2, 0, 0);
  for (int i = 0; i < 4; ++i) s.add(256);
  const int st = s.step(d);
  CHECK(st == MBX_ERR_HIP);
  CHECK(s.state(0) == MBX_REQ_FAILED && s.state(1) == MBX_REQ_FAILED);
  CHECK(s.state(2) == MBX_REQ_RUNNING && s.state(3) == MBX_REQ_RUNNING);
  CHECK(!s.slot_busy(0) && s.failed() == 2);
  s.finish(1, d, nullptr);
  CHECK(s.idle() && d.held.empty());
  std::printf("launch_failure: failed %u\n", s.failed());
}

// Random workload: arrivals, completions in any order, cancels, a small
// fragmenting pool and a binding budget.  After every pass: memory in use
// never exceeds the budget and matches the reservations; a request never
// runs twice; and whenever requests are still queued, some slot is busy --
// the scheduler never sits idle with work it will never start.  At the end
// every request has ended.
static void random_workload(uint64_t seed) {
  std::mt19937_64 rng(seed);
  const uint32_t nslots = 1 + rng() % 4, maxb = 1 + rng() % 6;
  const size_t fixed = 4096, budget = fixed + 64 * 1024;
  FakeDevice d(48 * 1024, nslots);          // the pool binds before the budget
  Scheduler s(nslots, maxb, budget, fixed);
  std::vector<size_t> need;
  std::vector<void *> hogs;                 // pool memory the caller holds
  uint32_t steps = 0;
  const uint32_t nreq = 3000;
  while (need.size() < nreq || !s.idle()) {
    if (need.size() < nreq && rng() % 3) {
      const size_t n = 256 * (1 + rng() % 96);
      need.push_back(n);
      s.add(n, rng() % 5 == 0);
    }
    if (rng() % 11 == 0 && !need.empty()) {
      const size_t back = std::min<size_t>(need.size(), 20);
      (void)s.cancel((uint32_t)(need.size() - 1 - rng() % back), d);
    }
    if (rng() % 40 == 0) {
      void *p = nullptr;
      if (mbx_hostpool_alloc(d.pool, 256 * (1 + rng() % 64), &p) == MBX_OK)
        hogs.push_back(p);
    }
    if (!hogs.empty() && rng() % 30 == 0) {
      mbx_hostpool_free(d.pool, hogs.back());
      hogs.pop_back();
    }
    s.step(d);
    // invariants
    size_t held = 0;
    for (auto &kv : d.held) held += need[kv.first];
    CHECK(s.used() <= budget);
    CHECK(s.used() == fixed + held);
    CHECK(!s.has_queued() || s.any_busy());
    if (g_fail) return;
    // finish a random busy slot
    std::vector<uint32_t> busy;
    for (uint32_t k = 0; k < nslots; ++k)
      if (s.slot_busy(k)) busy.push_back(k);
    if (!busy.empty() && rng() % 2) s.finish(busy[rng() % busy.size()], d, nullptr);
    if (++steps > 100000) {
      CHECK(!"no progress");
      return;
    }
  }
  for (void *p : hogs) mbx_hostpool_free(d.pool, p);
  uint32_t ended = 0;
  for (uint32_t q = 0; q < nreq; ++q) ended += s.state(q) >= MBX_REQ_DONE;
  CHECK(ended == nreq);
  CHECK(d.twice == 0 && d.held.empty() && s.used() == fixed);
  CHECK(s.completed() + s.rejected() + s.cancelled() == nreq);
  std::printf("random_workload seed %llu: %u slots x %u, completed %u, "
              "rejected %u, cancelled %u, delayed %u, %u passes\n",
              (unsigned long long)seed, nslots, maxb, s.completed(),
              s.rejected(), s.cancelled(), s.delayed(), steps);
}

int main(int argc, char **argv) {
  const int seeds = argc > 1 ? std::atoi(argv[1]) : 8;
  pool_full_nothing_running();
  pool_full_waits_for_running();
  budget_fifo();
  slots_and_batches();
  solo_requests();
  cancel_states();
  launch_failure();
  for (int i = 0; i < seeds && !g_fail; ++i) random_workload(1000 + i);
  std::printf("%s: %d failed checks\n", g_fail ? "FAIL" : "OK", g_fail);
  return g_fail;
}
