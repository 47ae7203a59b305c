// batch.hpp -- descriptors of the persistent batched launch (host + device).
#pragma once
#include "mbx_internal.hpp"

namespace mbx {

constexpr int kBWPL = 2;          // 64-word chunks per wave per pass
constexpr uint32_t kItemQuantum = 512;   // item sizes are multiples of this
constexpr uint32_t kXcdLists = 8;        // one work list per XCD
constexpr uint32_t kXcdCounterStride = 32;   // uint32: counters 128 B apart
// Counters of the two launches of a batch (one set per query class).
constexpr size_t kCounterBytes = 2 * kXcdLists * kXcdCounterStride * 4;
// Largest LDS group table of the conjunction kernel: with its 8 KB of row
// lists, five 4-wave workgroups fit a CU's 160 KB of LDS at the largest
// table, six at tables up to 2,300 groups, as its 80 VGPRs allow (6 waves
// per SIMD; the general kernel's 111 allow 4).
constexpr uint32_t kConjLdsGroups = 3000;

struct BatchQ {
  DevPrg prg;
  DevOp op;
  unsigned long long *out;
  uint32_t nr_grp, lds, chk, eager;
  uint32_t nitems, words_per_item, pad[2];
};

// The two query classes of a batch, each run by its own persistent kernel:
// class 0 are eager conjunctions with an LDS table of at most
// kConjLdsGroups (the conjunction kernel), class 1 the rest (the general
// kernel).  Descriptors of class 0 come first.
struct BatchShape {
  uint32_t nq[2];
  uint32_t items[2];       // work items (row ranges of every query)
  uint32_t max_items[2];   // row ranges of the class's longest query
  size_t lds_bytes[2];     // dynamic LDS: the class's largest LDS table
};

// Builds the device descriptors of nq queries for launch shape t: fills
// h[0..nq) (class 0 first) and *shape.  Host only, no device calls.
int batch_prepare(mbx_ctx *ctx, const mbx_query *qs, uint32_t nq,
                  const mbx_tune &t, BatchQ *h, BatchShape *shape);
// Side stream and events for running the two classes' kernels side by side.
struct BatchFork {
  hipStream_t side;
  hipEvent_t ev_fork, ev_join;
};

// Enqueues output zeroing, counter reset and the persistent kernels on s
// (with fork and both classes present, the general kernel on fork->side,
// joined back into s).  d_counter: kCounterBytes of device memory.
int batch_launch(mbx_ctx *ctx, hipStream_t s, const BatchQ *d_qs,
                 const BatchShape &shape, uint32_t *d_counter,
                 const mbx_tune &t, const BatchFork *fork = nullptr);

}  // namespace mbx
