// query_launch.hpp -- host-side launch interface of the single-query plans.
#pragma once
#include "mbx_internal.hpp"

namespace mbx {

struct Acct;

struct QueryPlan {
  DevPrg prg;
  DevOp op;
  unsigned long long *out = nullptr;
  uint32_t nr_grp = 0;
  bool lds = false;   // group table fits in LDS
  bool eager = false; // all operand loads up front (no ANDM early-out)
  bool direct = false; // no compaction: lanes walk their own set bits
  // replica tables + counters of the LDS-table flush (zero between
  // launches); null = the context's set for the launch stream.  Launches
  // that may overlap on the device need their own.
  unsigned long long *spread = nullptr;
  uint32_t *spread_done = nullptr;
  // stream of the launch; null = the context's stream
  hipStream_t stream = nullptr;
};

inline hipStream_t stream_of(mbx_ctx *ctx, const QueryPlan &q) {
  return q.stream ? q.stream : ctx->stream;
}

// The context's flush replicas for stream s: one set per stream, so
// launches on different streams of one context never share replicas (a
// set is taken from the pool and zeroed on the stream's first use).
int ctx_spread(mbx_ctx *ctx, hipStream_t s, unsigned long long **rep,
               uint32_t **done);

// Bytes of one set of flush replicas (tables, then counters).
size_t spread_bytes();

// The no-allocation entry point's device scratch: found and cleared once per
// device (mbx_init calls it with the device current).
int noalloc_setup(mbx_ctx *ctx);
// Rows the materialise plan can take in a scratch of that many bytes.
size_t noalloc_max_rows(size_t scratch_bytes);

// Validates the query and builds its device descriptors.
int prepare(mbx_ctx *ctx, const mbx_vprg *prg, const mbx_rowop *op,
            uint32_t nr_grp, int64_t *d_out, QueryPlan *q);
int launch_fused(mbx_ctx *ctx, const QueryPlan &q, const Acct *acct);
// The fused plan through the kernel built for the program's shape
// (MBX_PLAN_DEDICATED); MBX_ERR_UNSUPPORTED when there is none.
int launch_dedicated(mbx_ctx *ctx, const QueryPlan &q);
int launch_scan(mbx_ctx *ctx, const QueryPlan &q);
int launch_stage1(mbx_ctx *ctx, const DevPrg &P, uint64_t *po, uint64_t *uo);
int launch_stage2(mbx_ctx *ctx, const uint64_t *po, const uint64_t *uo,
                  uint32_t nwords, uint32_t *rows, uint32_t *count);
int launch_stage3(mbx_ctx *ctx, const QueryPlan &q, const uint32_t *rows,
                  const uint32_t *count);

}  // namespace mbx
