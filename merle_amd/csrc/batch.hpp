// batch.hpp -- descriptors of the persistent batched launch (host + device).
#pragma once
#include "mbx_internal.hpp"

namespace mbx {

constexpr int kBWPL = 4;          // 64-word chunks per wave per pass
constexpr uint32_t kItemQuantum = 512;   // item sizes are multiples of this
constexpr uint32_t kXcdLists = 8;        // one work list per XCD
constexpr uint32_t kXcdCounterStride = 32;   // uint32: counters 128 B apart
constexpr size_t kCounterBytes = kXcdLists * kXcdCounterStride * 4;

struct BatchQ {
  DevPrg prg;
  DevOp op;
  unsigned long long *out;
  uint32_t nr_grp, lds, chk, eager;
  uint32_t item0, nitems, words_per_item, pad;
};

// Builds the device descriptors of nq queries for launch shape t: fills
// h[0..nq), *total_items and *lds_bytes.  Host only, no device calls.
int batch_prepare(mbx_ctx *ctx, const mbx_query *qs, uint32_t nq,
                  const mbx_tune &t, BatchQ *h, uint32_t *total_items,
                  size_t *lds_bytes, uint32_t *max_items);
// Enqueues output zeroing, counter reset and the persistent kernel on s.
// d_counter: kCounterBytes of device memory.
int batch_launch(hipStream_t s, const BatchQ *d_qs, uint32_t nq,
                 uint32_t max_items, uint32_t *d_counter, size_t lds_bytes,
                 unsigned grid);
unsigned batch_grid(mbx_ctx *ctx, const mbx_tune &t, uint32_t total_items);

}  // namespace mbx
