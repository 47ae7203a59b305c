// sched.hpp -- the serving loop's scheduling logic, free of HIP.
//
// Requests go through: queued -> admitted (memory reserved) -> running (on a
// slot) -> done, or end as rejected / cancelled / failed.  Admission is FIFO
// and by memory: a request whose reservation can never fit the budget is
// rejected at once; one that does not fit now waits at the head of the queue
// (it is counted once as delayed) until running requests finish and release
// memory.  When the allocator itself refuses a reservation while nothing is
// running or admitted -- so nothing the scheduler owns will ever be released
// -- the request is rejected instead of waiting forever.  A free slot takes
// up to max_batch admitted requests, in admission order; a request marked
// solo (one that runs as a launch of its own: a large or a scan request)
// goes out alone, so the requests behind it do not wait for it on its slot.
//
// The device side is behind SchedDevice, so the same code runs the GPU
// server (serve.cpp) and a fake device in CPU tests.  The class is not
// thread-safe: its owner serialises calls.
#pragma once
#include <cstddef>
#include <cstdint>
#include <deque>
#include <vector>

#include "../../include/mbx.h"

namespace mbx {

struct SchedDevice {
  virtual ~SchedDevice() = default;
  // Reserves `bytes` of device memory for request `req`: MBX_OK, or
  // MBX_ERR_OOM when the allocator cannot serve it right now.
  virtual int reserve(uint32_t req, size_t bytes) = 0;
  virtual void release(uint32_t req) = 0;
  // Starts requests reqs[0..n) on slot `slot` (asynchronously).  An error
  // fails those requests; the slot stays free.
  virtual int launch(uint32_t slot, const uint32_t *reqs, uint32_t n) = 0;
};

class Scheduler {
 public:
  // fixed_bytes: memory the server holds for its whole life (charged against
  // the budget before any request); budget 0 = unbounded.
  Scheduler(uint32_t nslots, uint32_t max_batch, size_t budget,
            size_t fixed_bytes)
      : max_batch_(max_batch ? max_batch : 1),
        budget_(budget ? budget : (size_t)-1),
        fixed_(fixed_bytes),
        used_(fixed_bytes),
        peak_(fixed_bytes),
        slots_(nslots ? nslots : 1) {}

  // A new request needing `need` bytes; returns its id (0, 1, 2, ...).
  uint32_t add(size_t need, bool solo = false) {
    Req r;
    r.need = need;
    r.solo = solo;
    reqs_.push_back(r);
    const uint32_t id = (uint32_t)(reqs_.size() - 1);
    pending_.push_back(id);
    return id;
  }

  // Admission, then dispatch onto free slots.  Returns the first launch
  // error (its requests are failed), else MBX_OK.
  int step(SchedDevice &d) {
    admit(d);
    return dispatch(d);
  }

  // Slot `slot` has finished: its requests are done and their memory is
  // released.  `out` (optional) receives the request ids.
  void finish(uint32_t slot, SchedDevice &d, std::vector<uint32_t> *out) {
    Slot &s = slots_[slot];
    for (uint32_t q : s.reqs) {
      reqs_[q].state = MBX_REQ_DONE;
      d.release(q);
      used_ -= reqs_[q].need;
      ++completed_;
      if (out) out->push_back(q);
    }
    s.reqs.clear();
    s.busy = false;
  }

  // A slot whose work failed: its requests end as failed (memory released).
  void fail_slot(uint32_t slot, SchedDevice &d) {
    Slot &s = slots_[slot];
    for (uint32_t q : s.reqs) {
      reqs_[q].state = MBX_REQ_FAILED;
      d.release(q);
      used_ -= reqs_[q].need;
      ++failed_;
    }
    s.reqs.clear();
    s.busy = false;
  }

  // Cancels a queued or admitted request (MBX_OK); a running or finished one
  // is left alone (MBX_ERR_INVALID).
  int cancel(uint32_t id, SchedDevice &d) {
    if (id >= reqs_.size()) return MBX_ERR_INVALID;
    Req &r = reqs_[id];
    if (r.state == MBX_REQ_QUEUED) {
      erase(pending_, id);
    } else if (r.state == MBX_REQ_ADMITTED) {
      erase(admitted_, id);
      d.release(id);
      used_ -= r.need;
    } else {
      return MBX_ERR_INVALID;
    }
    r.state = MBX_REQ_CANCELLED;
    ++cancelled_;
    return MBX_OK;
  }

  int state(uint32_t id) const { return reqs_[id].state; }
  uint32_t slot_of(uint32_t id) const { return reqs_[id].slot; }
  bool slot_busy(uint32_t s) const { return slots_[s].busy; }
  const std::vector<uint32_t> &slot_reqs(uint32_t s) const {
    return slots_[s].reqs;
  }
  uint32_t nslots() const { return (uint32_t)slots_.size(); }
  size_t nreqs() const { return reqs_.size(); }
  // every request has reached a terminal state
  bool idle() const {
    return pending_.empty() && admitted_.empty() && !any_busy();
  }
  bool any_busy() const {
    for (const Slot &s : slots_)
      if (s.busy) return true;
    return false;
  }
  bool has_queued() const { return !pending_.empty() || !admitted_.empty(); }

  size_t used() const { return used_; }
  size_t peak() const { return peak_; }
  uint32_t delayed() const { return delayed_; }
  uint32_t rejected() const { return rejected_; }
  uint32_t completed() const { return completed_; }
  uint32_t cancelled() const { return cancelled_; }
  uint32_t failed() const { return failed_; }
  uint32_t launches() const { return launches_; }
  uint64_t launched_reqs() const { return launched_reqs_; }

 private:
  struct Req {
    size_t need = 0;
    int state = MBX_REQ_QUEUED;
    uint32_t slot = 0;
    bool delayed = false;
    bool solo = false;
  };
  struct Slot {
    bool busy = false;
    std::vector<uint32_t> reqs;
  };

  static void erase(std::deque<uint32_t> &q, uint32_t id) {
    for (auto it = q.begin(); it != q.end(); ++it)
      if (*it == id) {
        q.erase(it);
        return;
      }
  }

  void admit(SchedDevice &d) {
    while (!pending_.empty()) {
      const uint32_t q = pending_.front();
      Req &r = reqs_[q];
      if (r.need > budget_ || fixed_ > budget_ - r.need) {  // never fits
        r.state = MBX_REQ_REJECTED;
        ++rejected_;
        pending_.pop_front();
        continue;
      }
      if (used_ + r.need > budget_) {       // wait for memory
        mark_delayed(r);
        return;
      }
      if (d.reserve(q, r.need) != MBX_OK) {
        if (!any_busy() && admitted_.empty()) {
          // nothing of ours will be released: waiting could last forever
          r.state = MBX_REQ_REJECTED;
          ++rejected_;
          pending_.pop_front();
          continue;
        }
        mark_delayed(r);
        return;
      }
      used_ += r.need;
      if (used_ > peak_) peak_ = used_;
      r.state = MBX_REQ_ADMITTED;
      admitted_.push_back(q);
      pending_.pop_front();
    }
  }

  int dispatch(SchedDevice &d) {
    int first_err = MBX_OK;
    for (uint32_t s = 0; s < slots_.size() && !admitted_.empty(); ++s) {
      Slot &sl = slots_[s];
      if (sl.busy) continue;
      // a solo head goes alone; otherwise the leading non-solo requests
      uint32_t k = 1;
      if (!reqs_[admitted_.front()].solo)
        while (k < max_batch_ && k < admitted_.size() &&
               !reqs_[admitted_[k]].solo)
          ++k;
      sl.reqs.assign(admitted_.begin(), admitted_.begin() + k);
      admitted_.erase(admitted_.begin(), admitted_.begin() + k);
      for (uint32_t q : sl.reqs) {
        reqs_[q].state = MBX_REQ_RUNNING;
        reqs_[q].slot = s;
      }
      sl.busy = true;
      const int st = d.launch(s, sl.reqs.data(), k);
      if (st != MBX_OK) {
        fail_slot(s, d);
        if (first_err == MBX_OK) first_err = st;
        continue;
      }
      ++launches_;
      launched_reqs_ += k;
    }
    return first_err;
  }

  void mark_delayed(Req &r) {
    if (!r.delayed) {
      r.delayed = true;
      ++delayed_;
    }
  }

  uint32_t max_batch_;
  size_t budget_;
  size_t fixed_, used_, peak_;
  std::vector<Req> reqs_;
  std::vector<Slot> slots_;
  std::deque<uint32_t> pending_, admitted_;
  uint32_t delayed_ = 0, rejected_ = 0, completed_ = 0, cancelled_ = 0,
           failed_ = 0, launches_ = 0;
  uint64_t launched_reqs_ = 0;
};

}  // namespace mbx
