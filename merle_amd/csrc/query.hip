// query.hip -- single-query plans on gfx950:
//   fused        vprg_grpby            (primitive.cuh:340-445)
//   materialise  _stg1 / _stg2 / _stg3 (primitive_abla.cuh:262-453)
//   scan         grpby                 (primitive.cuh:16-48)
// plus the algorithmic-byte accounting kernel behind mbx_alg_bytes.
#include <algorithm>
#include <cstring>
#include <map>
#include <mutex>
#include <vector>

#include "devquery.hpp"
#include "query_launch.hpp"

namespace mbx {

struct QArgs {
  DevPrg prg;
  DevOp op;
  unsigned long long *out;
  uint32_t nr_grp;
  const Acct *acct;
  Spread sp;
};

// ------------------------------------------------------------ fused plan
// Wave r of the R = 4 * gridDim.x waves takes the 64-word chunks r, r + R,
// r + 2R, ... -- WPL of them per pass, all loaded at once, and normally one
// pass: the launch sizes R so that WPL chunks per wave cover the bitmaps.
// At any moment the waves together read one contiguous window of each
// bitmap (DRAM-page friendly), not R scattered ranges.  The workgroup keeps
// its group table in LDS until the end; the table then goes out through the
// replicated flush (flush_groups).
template <bool CHK, bool LDSG, bool ACCT, int WPL, int EAGER, bool DIRECT = false>
__global__ __launch_bounds__(256) void fused_kernel(QArgs a) {
  __shared__ __align__(16) uint32_t s_list[4][kWaveList];
  GroupSink sink;
  sink.lds = LDSG;
  sink.out = a.out;
  if (LDSG)
    for (uint32_t i = threadIdx.x; i < a.nr_grp; i += 256) s_grp_dyn[i] = 0;
  __syncthreads();
  const uint32_t wv = threadIdx.x >> 6;
  const uint32_t ws = gridDim.x * 4 * 64;   // words between a wave's chunks
  for (uint32_t w0 = (blockIdx.x * 4 + wv) * 64; w0 < a.prg.nwords;
       w0 += ws * WPL)
    run_wave_range<CHK, ACCT, WPL, EAGER, 2, DIRECT>(a.prg, a.op, a.nr_grp, w0,
                                                     ws, a.prg.nwords, sink,
                                                     s_list[wv], a.acct);
  if (LDSG) flush_groups(a.sp, a.out, a.nr_grp, blockIdx.x, gridDim.x);
}

// ------------------------------------------------------------- scan plan
// Every row, all predicates checked (op(row, chk=true)), no bitmaps.  The
// rows are evaluated 64 at a time into words (scan_words: coalesced column
// loads, a ballot per word), then compacted and run through the row op like
// the fused plan's -- wave r takes the 64-word chunks r, r + R, ... as there.
template <bool LDSG, int WPL>
__global__ __launch_bounds__(256) void scan_kernel(QArgs a) {
  __shared__ __align__(16) uint32_t s_list[4][kWaveList];
  GroupSink sink;
  sink.lds = LDSG;
  sink.out = a.out;
  if (LDSG)
    for (uint32_t i = threadIdx.x; i < a.nr_grp; i += 256) s_grp_dyn[i] = 0;
  __syncthreads();
  const uint32_t wv = threadIdx.x >> 6;
  const uint32_t ws = gridDim.x * 4 * 64;   // words between a wave's chunks
  for (uint32_t w0 = (blockIdx.x * 4 + wv) * 64; w0 < a.prg.nwords;
       w0 += ws * WPL)
    run_wave_range<false, false, WPL, 0, 2, false, true>(
        a.prg, a.op, a.nr_grp, w0, ws, a.prg.nwords, sink, s_list[wv], nullptr);
  if (LDSG) flush_groups(a.sp, a.out, a.nr_grp, blockIdx.x, gridDim.x);
}

// ------------------------------------------------------- materialise plan
// Stage 1: program -> possible/uncertain bitmaps in HBM.
// GREG: the ablation with the registers in device memory (NfuStg1 of the
// reference's report: _stg1 with idxbuf, primitive_abla.cuh:262-288).
template <bool CHK, bool GREG = false>
__global__ __launch_bounds__(256) void stage1_kernel(DevPrg P,
                                                     uint64_t *__restrict__ po,
                                                     uint64_t *__restrict__ uo) {
  for (uint32_t w = blockIdx.x * 256 + threadIdx.x; w < P.nwords;
       w += gridDim.x * 256) {
    uint64_t p, u;
    eval_word<CHK, false, GREG ? 4 : 1>(P, w, p, u, nullptr);
    po[w] = p;
    if (CHK) uo[w] = u;
  }
}

// Stage 2: popcount -> list offsets -> row list (primitive_abla.cuh:290-354).
// Workgroup b owns the contiguous chunks [b*cpw, (b+1)*cpw) of 256 x WPL
// words: a first pass counts its rows, one atomicAdd reserves its segment of
// the list (one per workgroup of a grid of ~num_cus workgroups, so few adds
// meet on the counter), and a second pass -- its reads mostly from L2 --
// block-scans each chunk and writes the row ids in row order.
template <bool CHK, int WPL>
__global__ __launch_bounds__(256) void stage2_kernel(
    const uint64_t *__restrict__ po, const uint64_t *__restrict__ uo,
    uint32_t nwords, uint32_t *__restrict__ rows, uint32_t *count,
    uint32_t cpw) {
  __shared__ uint32_t s_wave[4];
  __shared__ uint32_t s_base;
  const uint32_t tid = threadIdx.x;
  const uint32_t nchunks = (nwords + 256 * WPL - 1) / (256 * WPL);
  const uint32_t c0 = blockIdx.x * cpw;
  const uint32_t c1 = c0 + cpw < nchunks ? c0 + cpw : nchunks;
  uint32_t mine = 0;
  for (uint32_t c = c0; c < c1; ++c)
#pragma unroll
    for (int k = 0; k < WPL; ++k) {
      const uint32_t w = (c * WPL + k) * 256 + tid;
      if (w < nwords) mine += __popcll(po[w]);
    }
  uint32_t total;
  (void)block_excl_scan(mine, s_wave, &total);
  if (tid == 0) s_base = total ? atomicAdd(count, total) : 0;
  __syncthreads();
  if (total == 0) return;   // workgroup-uniform
  uint32_t base = s_base;
  for (uint32_t c = c0; c < c1; ++c) {
    uint64_t p[WPL], u[WPL];
    uint32_t cnt = 0;
#pragma unroll
    for (int k = 0; k < WPL; ++k) {
      const uint32_t w = (c * WPL + k) * 256 + tid;
      p[k] = w < nwords ? po[w] : 0;
      u[k] = (CHK && w < nwords) ? uo[w] : 0;
      cnt += __popcll(p[k]);
    }
    uint32_t ctot;
    uint32_t off = base + block_excl_scan(cnt, s_wave, &ctot);
    base += ctot;
#pragma unroll
    for (int k = 0; k < WPL; ++k) {
      uint64_t m = p[k];
      const uint32_t row0 = ((c * WPL + k) * 256 + tid) * 64;
      while (m) {
        const uint32_t b = (uint32_t)__builtin_ctzll(m);
        m &= m - 1;
        uint32_t e = row0 + b;
        if (CHK && ((u[k] >> b) & 1ull)) e |= 0x80000000u;
        rows[off++] = e;
      }
    }
  }
}

// Stage 3: op + group-by over the list.  The list length is read on the
// device, so no host round trip sits between stages 2 and 3 (the reference
// copies it to the host, primitive_abla.cuh:437-444).
template <bool CHK, bool LDSG>
__global__ __launch_bounds__(256) void stage3_kernel(
    QArgs a, const uint32_t *__restrict__ rows, const uint32_t *count) {
  GroupSink sink;
  sink.lds = LDSG;
  sink.out = a.out;
  if (LDSG)
    for (uint32_t i = threadIdx.x; i < a.nr_grp; i += 256) s_grp_dyn[i] = 0;
  __syncthreads();
  const uint32_t n = *count;
  // workgroup-uniform trip count: every lane takes part in the sink's wave
  // reductions
  for (uint32_t j0 = blockIdx.x * 256; j0 < n; j0 += gridDim.x * 256) {
    const uint32_t j = j0 + threadIdx.x;
    const bool valid = j < n;
    const uint32_t e = valid ? rows[j] : 0u;
    process_entries<CHK, false, 1>(a.prg, a.op, a.nr_grp, &e, &valid, sink,
                                   nullptr);
  }
  if (LDSG) flush_groups(a.sp, a.out, a.nr_grp, blockIdx.x, gridDim.x);
}

// Popcount of accounting bitsets: lines[e] = set bits of entry e.
__global__ void acct_count_kernel(const uint32_t *__restrict__ bits,
                                  const uint32_t *__restrict__ offs,
                                  uint32_t nent,
                                  unsigned long long *__restrict__ lines) {
  for (uint32_t e = blockIdx.x; e < nent; e += gridDim.x) {
    unsigned long long c = 0;
    for (uint32_t i = offs[e] + threadIdx.x; i < offs[e + 1]; i += blockDim.x)
      c += __popc(bits[i]);
    atomicAdd(lines + e, c);
  }
}


// Rows in the OR of a column's bins (plan densities).
// out[0] += rows in the column's operand (OR of its bins), out[1] += rows
// only in its edge bins (the rows a candidate check must re-check).
__global__ void density_kernel(DevCol c, uint32_t nwords, uint32_t nrows,
                               unsigned long long *out) {
  __shared__ unsigned long long s[2][256];
  unsigned long long cnt = 0, cnte = 0;
  for (uint32_t w = blockIdx.x * 256 + threadIdx.x; w < nwords;
       w += gridDim.x * 256) {
    uint64_t x = 0, e = 0;
    for (uint32_t k = 0; k < c.nm; ++k) x |= c.m[k][w];
    if (c.l) e |= c.l[w];
    if (c.r) e |= c.r[w];
    uint64_t keep = ~0ull;
    if (w == nwords - 1 && (nrows & 63u)) keep = (1ull << (nrows & 63u)) - 1ull;
    cnt += __popcll((x | e) & keep);
    cnte += __popcll(e & ~x & keep);
  }
  s[0][threadIdx.x] = cnt;
  s[1][threadIdx.x] = cnte;
  __syncthreads();
  for (int k = 128; k > 0; k >>= 1) {
    if ((int)threadIdx.x < k) {
      s[0][threadIdx.x] += s[0][threadIdx.x + k];
      s[1][threadIdx.x] += s[1][threadIdx.x + k];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0 && s[0][0]) atomicAdd(out, s[0][0]);
  if (threadIdx.x == 0 && s[1][0]) atomicAdd(out + 1, s[1][0]);
}

// ---------------------------------------------------------------- launches
static constexpr int kWPL = 2;   // 64-row words per thread per tile

int prepare(mbx_ctx *ctx, const mbx_vprg *prg, const mbx_rowop *op,
            uint32_t nr_grp, int64_t *d_out, QueryPlan *q) {
  if (!ctx) return fail(MBX_ERR_INVALID, "null ctx");
  if (!d_out) return fail(MBX_ERR_INVALID, "null d_out");
  if (nr_grp == 0) return fail(MBX_ERR_INVALID, "nr_grp == 0");
  if (reinterpret_cast<uintptr_t>(d_out) % 8)
    return fail(MBX_ERR_ALIGN, "d_out not 8-byte aligned");
  int st = make_dev_prg(prg, &q->prg);
  if (st) return st;
  st = make_dev_op(op, 0, &q->op);
  if (st) return st;
  if (q->prg.chk && q->op.npreds == 0)
    return fail(MBX_ERR_INVALID,
                "program has edge bins but the row op has no predicates to "
                "re-check uncertain rows");
  for (uint32_t i = 0; i < q->prg.nins; ++i) {
    const uint32_t o = q->prg.ins[i][0], s = q->prg.ins[i][2];
    if (q->prg.chk && (o == MBX_ANDM || o == MBX_ORM) && s >= q->op.npreds)
      return fail(MBX_ERR_INVALID,
                  "instruction reads a column without an exact predicate");
  }
  q->nr_grp = nr_grp;
  q->out = reinterpret_cast<unsigned long long *>(d_out);
  q->lds = nr_grp <= kLdsGroups;
  return MBX_OK;
}

static unsigned persistent_grid(mbx_ctx *ctx, uint32_t work, unsigned per_cu) {
  uint64_t g = (uint64_t)ctx->num_cus * per_cu;
  if (g > work) g = work;
  return g ? (unsigned)g : 1u;
}

size_t spread_bytes() {
  return (size_t)kSpreadRep * kLdsGroups * 8 +
         (size_t)kSpreadRep * kSpreadDoneStride * 4;
}

int ctx_spread(mbx_ctx *ctx, hipStream_t s, unsigned long long **rep,
               uint32_t **done) {
  if (s == ctx->own_stream) {
    *rep = ctx->spread;
    *done = ctx->spread_done;
    return MBX_OK;
  }
  std::lock_guard<std::mutex> g(ctx->spread_mu);
  auto it = ctx->spreads.find(s);
  if (it == ctx->spreads.end()) {
    void *p = nullptr;
    int st = ctx->pool->alloc(spread_bytes(), &p);
    if (st) return st;
    st = launch_fill_u32(s, p, 0, spread_bytes() / 4);
    if (st) {
      ctx->pool->free(p);
      return st;
    }
    auto *r = static_cast<unsigned long long *>(p);
    auto *d = reinterpret_cast<uint32_t *>(static_cast<char *>(p) +
                                           (size_t)kSpreadRep * kLdsGroups * 8);
    it = ctx->spreads.emplace(s, std::make_pair(r, d)).first;
  }
  *rep = it->second.first;
  *done = it->second.second;
  return MBX_OK;
}

static int spread_of(mbx_ctx *ctx, const QueryPlan &q, Spread *sp) {
  sp->gpad = (uint32_t)((q.nr_grp + 15) / 16 * 16);
  if (q.spread) {
    sp->rep = q.spread;
    sp->done = q.spread_done;
    return MBX_OK;
  }
  return ctx_spread(ctx, stream_of(ctx, q), &sp->rep, &sp->done);
}

static size_t grp_lds(const QueryPlan &q) {
  return q.lds ? (size_t)q.nr_grp * 8 : 0;
}

// Resident workgroups per CU for a kernel at a dynamic LDS size (cached).
static unsigned occupancy(const void *fn, size_t lds) {
  static std::mutex mu;
  static std::map<std::pair<const void *, size_t>, unsigned> cache;
  std::lock_guard<std::mutex> g(mu);
  auto key = std::make_pair(fn, lds);
  auto it = cache.find(key);
  if (it != cache.end()) return it->second;
  int nb = 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, fn, 256, lds) !=
          hipSuccess ||
      nb < 1) {
    (void)hipGetLastError();
    nb = 2;
  }
  cache[key] = (unsigned)nb;
  return (unsigned)nb;
}

// Grid: enough waves that WPL chunks each cover the program's words in one
// pass, no more than are resident; small WPL is tried first (fewer
// registers, more resident waves).  Past the resident limit the waves make
// several passes.
template <bool C, bool L, bool A, int E, int W>
static bool fused_try(mbx_ctx *ctx, hipStream_t s, const QArgs &a, size_t lds,
                      bool last) {
  const void *fn = reinterpret_cast<const void *>(&fused_kernel<C, L, A, W, E>);
  const unsigned per_cu = occupancy(fn, lds);
  const uint64_t chunks = ceil_div(a.prg.nwords, 64);
  const uint64_t waves = ceil_div(chunks, W);
  const uint64_t resident = (uint64_t)ctx->num_cus * per_cu * 4;
  if (waves > resident && !last) return false;
  const unsigned grid = (unsigned)ceil_div(waves < resident ? waves : resident, 4);
  hipLaunchKernelGGL((fused_kernel<C, L, A, W, E>), dim3(grid), dim3(256),
                     lds, s, a);
  return true;
}

template <bool C, bool L, bool A, int E>
static void fused_launch(mbx_ctx *ctx, hipStream_t s, const QArgs &a,
                         size_t lds) {
  if (A) {
    (void)fused_try<C, L, A, E, 2>(ctx, s, a, lds, true);
    return;
  }
  if (fused_try<C, L, A, E, 2>(ctx, s, a, lds, false)) return;
  if (C) {   // three-valued state doubles the registers: stop at 4 words
    (void)fused_try<C, L, A, E, 4>(ctx, s, a, lds, true);
    return;
  }
  if (fused_try<C, L, A, E, 4>(ctx, s, a, lds, false)) return;
  (void)fused_try<C, L, A, E, C ? 4 : 6>(ctx, s, a, lds, true);
}

// W words per lane and an even share of passes: the grid is cut so that
// every wave makes the same number of full passes.
template <bool C, bool L, bool A, int E, int W>
static void fused_even(mbx_ctx *ctx, hipStream_t s, const QArgs &a,
                       size_t lds) {
  const void *fn = reinterpret_cast<const void *>(&fused_kernel<C, L, A, W, E>);
  const unsigned per_cu = occupancy(fn, lds);
  const uint64_t chunks = ceil_div(a.prg.nwords, 64);
  const uint64_t waves = ceil_div(chunks, W);
  const uint64_t resident = (uint64_t)ctx->num_cus * per_cu * 4;
  const uint64_t passes = ceil_div(waves, resident);
  const unsigned grid = (unsigned)ceil_div(ceil_div(waves, passes), 4);
  hipLaunchKernelGGL((fused_kernel<C, L, A, W, E>), dim3(grid), dim3(256),
                     lds, s, a);
}

// The conjunction form (EAGER 3).  Its shape comes from a sweep of words per
// lane (1..8) and grids on 120M-row star-schema queries: one pass at 2 words
// when the resident waves cover the words; otherwise, with the group table
// in LDS, 4 words on the full resident grid (6 words hold more words per
// pass but leave a third fewer waves for the gathers: 4-9 % slower); with
// the table in HBM, 2 words in even passes, which spreads the HBM atomics
// over the whole kernel (9 % faster than 4 words on 437,500 groups).
// Three-valued (candidate-check) conjunctions follow the same rule: on 60
// random programs over a 120M-row, 8-column Zipf table, 4 words on the
// resident grid took 2.75 ms, 2 words in even passes 2.87-2.89, 2 words on
// the resident grid 3.04, 1 word 3.51.
template <bool C, bool L>
static void fused_conj(mbx_ctx *ctx, hipStream_t s, const QArgs &a,
                       size_t lds) {
  if (fused_try<C, L, false, 3, 2>(ctx, s, a, lds, false)) return;
  if constexpr (L) (void)fused_try<C, L, false, 3, 4>(ctx, s, a, lds, true);
  else fused_even<C, L, false, 3, 2>(ctx, s, a, lds);
}

// The no-compaction ablation: early-out evaluation, 2 words per lane, the
// resident grid.
// E = 4: the base_fuse ablation (registers in device memory, P.greg).
template <bool C, bool L, int E = 0>
static void fused_direct(mbx_ctx *ctx, hipStream_t s, const QArgs &a,
                         size_t lds) {
  const void *fn =
      reinterpret_cast<const void *>(&fused_kernel<C, L, false, 2, E, true>);
  const unsigned per_cu = occupancy(fn, lds);
  const uint64_t waves = ceil_div(ceil_div(a.prg.nwords, 64), 2);
  const uint64_t resident = (uint64_t)ctx->num_cus * per_cu * 4;
  const unsigned grid = (unsigned)ceil_div(waves < resident ? waves : resident, 4);
  hipLaunchKernelGGL((fused_kernel<C, L, false, 2, E, true>), dim3(grid),
                     dim3(256), lds, s, a);
}

template <bool C, bool L>
static void fused_modes(mbx_ctx *ctx, hipStream_t s, const QArgs &a,
                        size_t lds, int mode) {
  if (mode == 3) fused_conj<C, L>(ctx, s, a, lds);
  else if (mode != 0) fused_launch<C, L, false, 1>(ctx, s, a, lds);
  else fused_launch<C, L, false, 0>(ctx, s, a, lds);
}

int launch_fused(mbx_ctx *ctx, const QueryPlan &q, const Acct *acct) {
  QArgs a;
  std::memset(&a, 0, sizeof(a));
  a.prg = q.prg;
  a.op = q.op;
  a.out = q.out;
  a.nr_grp = q.nr_grp;
  a.acct = acct;
  int st = spread_of(ctx, q, &a.sp);
  if (st) return st;
  const size_t lds = acct ? 0 : grp_lds(q);
  const bool chk = q.prg.chk != 0;
  // evaluation: 0 early-out; eager: 3 for conjunctions of column literals
  // (ANDs of column ORs, no program interpretation), else 1 (per-column
  // accumulators, then the program word by word)
  const int mode = !q.eager ? 0 : (q.prg.conj ? 3 : 1);
  // accounting runs the early-out form: the compulsory-line model
  const hipStream_t s = stream_of(ctx, q);
  if (acct) {
    if (chk) fused_launch<true, false, true, 0>(ctx, s, a, lds);
    else fused_launch<false, false, true, 0>(ctx, s, a, lds);
  } else if (q.direct && q.prg.greg) {
    if (chk && q.lds) fused_direct<true, true, 4>(ctx, s, a, lds);
    else if (chk) fused_direct<true, false, 4>(ctx, s, a, lds);
    else if (q.lds) fused_direct<false, true, 4>(ctx, s, a, lds);
    else fused_direct<false, false, 4>(ctx, s, a, lds);
  } else if (q.direct) {
    if (chk && q.lds) fused_direct<true, true>(ctx, s, a, lds);
    else if (chk) fused_direct<true, false>(ctx, s, a, lds);
    else if (q.lds) fused_direct<false, true>(ctx, s, a, lds);
    else fused_direct<false, false>(ctx, s, a, lds);
  } else if (chk) {
    if (q.lds) fused_modes<true, true>(ctx, s, a, lds, mode);
    else fused_modes<true, false>(ctx, s, a, lds, mode);
  } else {
    if (q.lds) fused_modes<false, true>(ctx, s, a, lds, mode);
    else fused_modes<false, false>(ctx, s, a, lds, mode);
  }
  MBX_HIP(hipGetLastError());
  return MBX_OK;
}

// Scan launch: 2 words per lane on the resident grid (several passes over
// a large table).
int launch_scan(mbx_ctx *ctx, const QueryPlan &q) {
  QArgs a;
  std::memset(&a, 0, sizeof(a));
  a.prg = q.prg;
  a.op = q.op;
  a.out = q.out;
  a.nr_grp = q.nr_grp;
  int st = spread_of(ctx, q, &a.sp);
  if (st) return st;
  const hipStream_t s = stream_of(ctx, q);
  constexpr int W = 2;
  const size_t lds = grp_lds(q);
  const void *fn = q.lds ? reinterpret_cast<const void *>(&scan_kernel<true, W>)
                         : reinterpret_cast<const void *>(&scan_kernel<false, W>);
  const unsigned per_cu = occupancy(fn, lds);
  const uint64_t waves = ceil_div(ceil_div(q.prg.nwords, 64), W);
  const uint64_t resident = (uint64_t)ctx->num_cus * per_cu * 4;
  const unsigned grid =
      (unsigned)std::max<uint64_t>(1, ceil_div(std::min(waves, resident), 4));
  if (q.lds)
    hipLaunchKernelGGL((scan_kernel<true, W>), dim3(grid), dim3(256), lds, s, a);
  else
    hipLaunchKernelGGL((scan_kernel<false, W>), dim3(grid), dim3(256), 0, s, a);
  MBX_HIP(hipGetLastError());
  return MBX_OK;
}

int launch_stage1(mbx_ctx *ctx, const DevPrg &P, uint64_t *po, uint64_t *uo) {
  const unsigned grid =
      persistent_grid(ctx, (uint32_t)ceil_div(P.nwords, 256), 16);
  if (P.greg && P.chk)
    hipLaunchKernelGGL((stage1_kernel<true, true>), dim3(grid), dim3(256), 0,
                       ctx->stream, P, po, uo);
  else if (P.greg)
    hipLaunchKernelGGL((stage1_kernel<false, true>), dim3(grid), dim3(256), 0,
                       ctx->stream, P, po, uo);
  else if (P.chk)
    hipLaunchKernelGGL(stage1_kernel<true>, dim3(grid), dim3(256), 0,
                       ctx->stream, P, po, uo);
  else
    hipLaunchKernelGGL(stage1_kernel<false>, dim3(grid), dim3(256), 0,
                       ctx->stream, P, po, uo);
  MBX_HIP(hipGetLastError());
  return MBX_OK;
}

int launch_stage2(mbx_ctx *ctx, const uint64_t *po, const uint64_t *uo,
                  uint32_t nwords, uint32_t *rows, uint32_t *count) {
  int st = launch_fill_u32(ctx->stream, count, 0, 1);
  if (st) return st;
  const uint32_t nchunks = (uint32_t)ceil_div(nwords, 256 * kWPL);
  const uint32_t g0 = nchunks < (uint32_t)ctx->num_cus ? nchunks : ctx->num_cus;
  const uint32_t cpw = (uint32_t)ceil_div(nchunks, g0 ? g0 : 1);
  const unsigned grid = cpw ? (unsigned)ceil_div(nchunks, cpw) : 1u;
  if (uo)
    hipLaunchKernelGGL((stage2_kernel<true, kWPL>), dim3(grid ? grid : 1),
                       dim3(256), 0, ctx->stream, po, uo, nwords, rows, count,
                       cpw);
  else
    hipLaunchKernelGGL((stage2_kernel<false, kWPL>), dim3(grid ? grid : 1),
                       dim3(256), 0, ctx->stream, po, uo, nwords, rows, count,
                       cpw);
  MBX_HIP(hipGetLastError());
  return MBX_OK;
}

int launch_stage3(mbx_ctx *ctx, const QueryPlan &q, const uint32_t *rows,
                  const uint32_t *count) {
  QArgs a;
  std::memset(&a, 0, sizeof(a));
  a.prg = q.prg;
  a.op = q.op;
  a.out = q.out;
  a.nr_grp = q.nr_grp;
  int st = spread_of(ctx, q, &a.sp);
  if (st) return st;
  const unsigned grid = (unsigned)ctx->num_cus * 4;
  const size_t lds = grp_lds(q);
#define MBX_S3(C, L)                                                           \
  hipLaunchKernelGGL((stage3_kernel<C, L>), dim3(grid), dim3(256), lds,        \
                     ctx->stream, a, rows, count)
  if (q.prg.chk) {
    if (q.lds) MBX_S3(true, true);
    else MBX_S3(true, false);
  } else {
    if (q.lds) MBX_S3(false, true);
    else MBX_S3(false, false);
  }
#undef MBX_S3
  MBX_HIP(hipGetLastError());
  return MBX_OK;
}

// Materialise plan with scratch from the pool (freed after a sync).
static int run_materialise(mbx_ctx *ctx, const QueryPlan &q) {
  const size_t nw = q.prg.nwords;
  void *buf = nullptr;
  const size_t bytes = nw * 8 * 2 + (size_t)q.prg.nrows * 4 + 256;
  int st = ctx->pool->alloc(bytes, &buf);
  if (st) return st;
  uint64_t *po = static_cast<uint64_t *>(buf);
  uint64_t *uo = po + nw;
  uint32_t *count = reinterpret_cast<uint32_t *>(uo + nw);
  uint32_t *rows = count + 64;
  st = launch_stage1(ctx, q.prg, po, q.prg.chk ? uo : nullptr);
  if (!st)
    st = launch_stage2(ctx, po, q.prg.chk ? uo : nullptr, (uint32_t)nw, rows,
                       count);
  if (!st) st = launch_stage3(ctx, q, rows, count);
  hipError_t e = hipStreamSynchronize(ctx->stream);
  ctx->pool->free(buf);
  if (!st && e != hipSuccess) st = hip_fail(e, "materialise plan");
  return st;
}

// ------------------------------------------------- no-allocation variant
// Callers that cannot allocate device memory (code under graph capture,
// tight serving loops, harnesses that forbid it) run mbx_prog_grpby_noalloc:
// its scratch -- the flush replicas of the LDS group tables and the
// materialise plan's possible / uncertain bitmaps, row count and row list --
// lives in one file-scope device array per GPU instead of the pool.
constexpr size_t kNoAllocScratch = size_t(64) << 20;

// Scratch layout: flush replicas (tables, then counters) first, then the
// materialise region.
static size_t noalloc_rep_bytes() { return ceil_div(spread_bytes(), 256) * 256; }

// Bytes of the materialise region for n rows: possible and uncertain
// bitmaps, the row count (256 B), the row list.
static size_t noalloc_mat_bytes(size_t n) {
  return ceil_div(n, 64) * 8 * 2 + 256 + n * 4;
}

size_t noalloc_max_rows(size_t scratch_bytes) {
  const size_t room = scratch_bytes - noalloc_rep_bytes();
  size_t lo = 0, hi = room / 4;   // largest n with noalloc_mat_bytes(n) <= room
  while (lo < hi) {
    const size_t mid = (lo + hi + 1) / 2;
    if (noalloc_mat_bytes(mid) <= room) lo = mid;
    else hi = mid - 1;
  }
  return lo;
}

}  // namespace mbx

static __device__ __align__(256) unsigned char scratch[mbx::kNoAllocScratch];

namespace {
// device address of `scratch` on each GPU (the host-side name is only a
// stand-in), found once per device by noalloc_setup
struct NoAllocDev {
  std::once_flag once;
  unsigned char *ptr = nullptr;
  int status = MBX_ERR_HIP;
};
constexpr int kMaxDevices = 64;
NoAllocDev g_noalloc[kMaxDevices];
}  // namespace

namespace mbx {

// Finds this device's copy of `scratch` and clears its replica tables (they
// must start at zero; kernels leave them zero).  Called by mbx_init with the
// context's device current; once per device.
int noalloc_setup(mbx_ctx *ctx) {
  if (ctx->device < 0 || ctx->device >= kMaxDevices) return MBX_ERR_INVALID;
  NoAllocDev &d = g_noalloc[ctx->device];
  std::call_once(d.once, [&] {
    void *p = nullptr;
    hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p) {
      (void)hipGetLastError();
      d.status = MBX_ERR_HIP;
      return;
    }
    int st = launch_fill_u32(ctx->stream, p, 0, noalloc_rep_bytes() / 4);
    if (st == MBX_OK) {
      e = hipStreamSynchronize(ctx->stream);
      if (e != hipSuccess) st = MBX_ERR_HIP;
    }
    d.status = st;
    if (st == MBX_OK) d.ptr = static_cast<unsigned char *>(p);
  });
  return d.status;
}

}  // namespace mbx

using namespace mbx;

extern "C" {

int mbx_prog_grpby(mbx_ctx *ctx, const mbx_vprg *prg, const mbx_rowop *op,
                   uint32_t nr_grp, int64_t *d_out, uint32_t flags) {
  QueryPlan q;
  int st = prepare(ctx, prg, op, nr_grp, d_out, &q);
  if (st) return st;
  if (!(flags & MBX_ACCUMULATE)) {
    st = launch_fill_u64(ctx->stream, d_out, 0, nr_grp);
    if (st) return st;
  }
  q.eager = (flags & MBX_PLAN_EAGER) != 0;
  q.direct = (flags & MBX_PLAN_DIRECT) != 0;
  if ((flags & MBX_PLAN_BASEFUSE) && (flags & MBX_PLAN_MASK) == MBX_PLAN_FUSED) {
    // the ablation: method 1 with the registers in device memory, which
    // are pool scratch freed after a wait
    void *greg = nullptr;
    st = ctx->pool->alloc((size_t)q.prg.nwords * 48, &greg);
    if (st) return st;
    q.prg.greg = static_cast<uint64_t *>(greg);
    q.direct = true;
    st = launch_fused(ctx, q, nullptr);
    hipError_t e = hipStreamSynchronize(ctx->stream);
    ctx->pool->free(greg);
    if (!st && e != hipSuccess) st = hip_fail(e, "base_fuse ablation");
    return st;
  }
  if (flags & MBX_PLAN_DEDICATED) {
    if ((flags & MBX_PLAN_MASK) != MBX_PLAN_FUSED ||
        (flags & (MBX_PLAN_DIRECT | MBX_PLAN_BASEFUSE)))
      return fail(MBX_ERR_INVALID,
                  "MBX_PLAN_DEDICATED is a form of the fused plan alone");
    return launch_dedicated(ctx, q);
  }
  switch (flags & MBX_PLAN_MASK) {
  case MBX_PLAN_FUSED: return launch_fused(ctx, q, nullptr);
  case MBX_PLAN_MATERIALISE: return run_materialise(ctx, q);
  case MBX_PLAN_SCAN: return launch_scan(ctx, q);
  default: return fail(MBX_ERR_INVALID, "unknown plan");
  }
}

int mbx_prog_grpby_noalloc(mbx_ctx *ctx, const mbx_vprg *prg,
                           const mbx_rowop *op, uint32_t nr_grp, int64_t *d_out,
                           uint32_t flags) {
  QueryPlan q;
  int st = prepare(ctx, prg, op, nr_grp, d_out, &q);
  if (st) return st;
  const NoAllocDev &d = g_noalloc[ctx->device];
  if (!d.ptr)
    return fail(MBX_ERR_HIP, "prog_grpby_noalloc: no scratch on this device");
  unsigned char *base = d.ptr;   // the device address, never `scratch`
  q.spread = reinterpret_cast<unsigned long long *>(base);
  q.spread_done = reinterpret_cast<uint32_t *>(
      base + (size_t)kSpreadRep * kLdsGroups * 8);
  q.eager = (flags & MBX_PLAN_EAGER) != 0;
  q.direct = (flags & MBX_PLAN_DIRECT) != 0;
  const uint32_t plan = flags & MBX_PLAN_MASK;
  if (plan == MBX_PLAN_MATERIALISE && q.prg.nrows > noalloc_max_rows(kNoAllocScratch))
    return fail(MBX_ERR_CAPACITY,
                "prog_grpby_noalloc: the materialise plan takes at most " +
                    std::to_string(noalloc_max_rows(kNoAllocScratch)) +
                    " rows in its " + std::to_string(kNoAllocScratch >> 20) +
                    " MiB scratch");
  if (plan > MBX_PLAN_SCAN) return fail(MBX_ERR_INVALID, "unknown plan");
  if (flags & MBX_PLAN_BASEFUSE)
    return fail(MBX_ERR_INVALID,
                "prog_grpby_noalloc: the base_fuse ablation needs scratch");
  if ((flags & MBX_PLAN_DEDICATED) &&
      (plan != MBX_PLAN_FUSED || (flags & MBX_PLAN_DIRECT)))
    return fail(MBX_ERR_INVALID,
                "MBX_PLAN_DEDICATED is a form of the fused plan alone");
  if (!(flags & MBX_ACCUMULATE)) {
    st = launch_fill_u64(ctx->stream, d_out, 0, nr_grp);
    if (st) return st;
  }
  if (flags & MBX_PLAN_DEDICATED) return launch_dedicated(ctx, q);
  if (plan == MBX_PLAN_FUSED) return launch_fused(ctx, q, nullptr);
  if (plan == MBX_PLAN_SCAN) return launch_scan(ctx, q);
  const size_t nw = q.prg.nwords;
  uint64_t *po = reinterpret_cast<uint64_t *>(base + noalloc_rep_bytes());
  uint64_t *uo = po + nw;
  uint32_t *count = reinterpret_cast<uint32_t *>(uo + nw);
  uint32_t *rows = count + 64;
  st = launch_stage1(ctx, q.prg, po, q.prg.chk ? uo : nullptr);
  if (!st)
    st = launch_stage2(ctx, po, q.prg.chk ? uo : nullptr, (uint32_t)nw, rows,
                       count);
  if (!st) st = launch_stage3(ctx, q, rows, count);
  return st;
}

int mbx_prog_eval_ex(mbx_ctx *ctx, const mbx_vprg *prg, uint32_t *d_poss,
                     uint32_t *d_unc, uint32_t flags) {
  if (!ctx || !d_poss) return fail(MBX_ERR_INVALID, "prog_eval: null");
  DevPrg P;
  int st = make_dev_prg(prg, &P);
  if (st) return st;
  if (reinterpret_cast<uintptr_t>(d_poss) % 8 ||
      reinterpret_cast<uintptr_t>(d_unc) % 8)
    return fail(MBX_ERR_ALIGN, "prog_eval: outputs not 8-byte aligned");
  if (P.chk && !d_unc)
    return fail(MBX_ERR_INVALID, "prog_eval: program needs d_unc");
  if (!(flags & MBX_PLAN_BASEFUSE))
    return launch_stage1(ctx, P, reinterpret_cast<uint64_t *>(d_poss),
                         reinterpret_cast<uint64_t *>(d_unc));
  void *greg = nullptr;   // the ablation's registers, freed after a wait
  st = ctx->pool->alloc((size_t)P.nwords * 48, &greg);
  if (st) return st;
  P.greg = static_cast<uint64_t *>(greg);
  st = launch_stage1(ctx, P, reinterpret_cast<uint64_t *>(d_poss),
                     reinterpret_cast<uint64_t *>(d_unc));
  hipError_t e = hipStreamSynchronize(ctx->stream);
  ctx->pool->free(greg);
  if (!st && e != hipSuccess) st = hip_fail(e, "prog_eval ablation");
  return st;
}

int mbx_prog_eval(mbx_ctx *ctx, const mbx_vprg *prg, uint32_t *d_poss,
                  uint32_t *d_unc) {
  return mbx_prog_eval_ex(ctx, prg, d_poss, d_unc, 0);
}

int mbx_compact(mbx_ctx *ctx, const uint32_t *d_poss, const uint32_t *d_unc,
                uint32_t nrows, uint32_t *d_rows, uint32_t *d_count) {
  if (!ctx || !d_poss || !d_rows || !d_count)
    return fail(MBX_ERR_INVALID, "compact: null");
  if (nrows == 0 || nrows > 0x7fffffffu)
    return fail(MBX_ERR_INVALID, "compact: nrows must be in [1, 2^31)");
  if (reinterpret_cast<uintptr_t>(d_poss) % 8 ||
      reinterpret_cast<uintptr_t>(d_unc) % 8)
    return fail(MBX_ERR_ALIGN, "compact: bitmaps not 8-byte aligned");
  return launch_stage2(ctx, reinterpret_cast<const uint64_t *>(d_poss),
                       reinterpret_cast<const uint64_t *>(d_unc),
                       (uint32_t)ceil_div(nrows, 64), d_rows, d_count);
}

int mbx_rows_grpby(mbx_ctx *ctx, const uint32_t *d_rows,
                   const uint32_t *d_count, const mbx_vprg *prg,
                   const mbx_rowop *op, uint32_t nr_grp, int64_t *d_out,
                   uint32_t flags) {
  if (!d_rows || !d_count) return fail(MBX_ERR_INVALID, "rows_grpby: null");
  QueryPlan q;
  int st = prepare(ctx, prg, op, nr_grp, d_out, &q);
  if (st) return st;
  if (!(flags & MBX_ACCUMULATE)) {
    st = launch_fill_u64(ctx->stream, d_out, 0, nr_grp);
    if (st) return st;
  }
  return launch_stage3(ctx, q, d_rows, d_count);
}

}  // extern "C"

namespace mbx {

// Compulsory-line accounting of a set of queries: the ACCT form of the fused
// kernel marks every 128-byte line its loads touch, in one bitset per
// distinct memory range, shared by all the queries (a line several queries
// read counts once).  Ranges are collected per query, sorted and merged
// (adjacent bins of one slab become one range).
static int alg_bytes_multi(mbx_ctx *ctx, const mbx_query *qs, uint32_t nq,
                           uint64_t *bytes, uint64_t *selected) {
  if (!ctx || !qs || !nq) return fail(MBX_ERR_INVALID, "alg_bytes: null");
  void *dummy_out = nullptr;
  int st = ctx->pool->alloc(8 * 4096, &dummy_out);
  if (st) return st;
  std::vector<QueryPlan> plans(nq);
  std::vector<std::pair<uint64_t, uint64_t>> ranges;   // [base, end)
  auto add = [&](const void *p, uint64_t nbytes) {
    if (!p || !nbytes) return;
    const uint64_t b = reinterpret_cast<uint64_t>(p);
    ranges.emplace_back(b, b + nbytes);
  };
  uint64_t groups = 0;
  for (uint32_t qi = 0; qi < nq && !st; ++qi) {
    QueryPlan &q = plans[qi];
    st = prepare(ctx, &qs[qi].prg, &qs[qi].op, qs[qi].nr_grp,
                 static_cast<int64_t *>(dummy_out), &q);
    if (st) break;
    groups += qs[qi].nr_grp;
    const uint64_t N = q.prg.nrows, bmp_bytes = (uint64_t)q.prg.nwords * 8;
    for (uint32_t e = 0; e < q.prg.nfl; ++e) add(q.prg.fl[e], bmp_bytes);
    auto add_src = [&](const DevSrc &sr) {
      if (!sr.data) return;
      if (sr.fk) {
        add(sr.fk, N * 4);
        add(sr.data, (uint64_t)sr.dimsz * sr.eb);
      } else {
        add(sr.data, N * sr.eb);
      }
    };
    if (q.prg.chk)
      for (uint32_t k = 0; k < q.op.npreds; ++k) add_src(q.op.pred[k]);
    for (uint32_t k = 0; k < q.op.nkeys; ++k) add_src(q.op.keys[k].src);
    if (q.op.measure != MBX_MEAS_COUNT) add_src(q.op.ma);
    if (q.op.measure == MBX_MEAS_MUL || q.op.measure == MBX_MEAS_SUB)
      add_src(q.op.mb);
  }
  if (st) {
    ctx->pool->free(dummy_out);
    return st;
  }
  // Ranges closer than a gap merge (adjacent bins of one slab; bins of one
  // slab read by different queries); a gap costs 1 bit per 128 bytes, so the
  // smallest gap that leaves at most kAcctMax ranges is used.
  std::sort(ranges.begin(), ranges.end());
  std::vector<std::pair<uint64_t, uint64_t>> merged;
  for (uint64_t gap = 1ull << 20;; gap <<= 2) {
    merged.clear();
    for (const auto &r : ranges) {
      if (!merged.empty() && r.first <= merged.back().second + gap)
        merged.back().second = std::max(merged.back().second, r.second);
      else
        merged.push_back(r);
    }
    if (merged.size() <= (size_t)kAcctMax || gap >= (8ull << 30)) break;
  }
  uint64_t span = 0;
  for (const auto &r : merged) span += r.second - r.first;
  if (span / 128 / 8 > (1ull << 30)) {
    ctx->pool->free(dummy_out);
    return fail(MBX_ERR_CAPACITY, "alg_bytes: address span too large");
  }
  if (merged.size() > (size_t)kAcctMax) {
    ctx->pool->free(dummy_out);
    return fail(MBX_ERR_CAPACITY, "alg_bytes: too many distinct arrays");
  }
  Acct h;
  std::memset(&h, 0, sizeof(h));
  std::vector<uint32_t> offs(1, 0);
  for (const auto &r : merged) {
    AcctEntry &e = h.e[h.n++];
    e.base = r.first;
    e.nbytes = r.second - r.first;
    e.off = offs.back();
    offs.push_back(offs.back() + (uint32_t)ceil_div(ceil_div(e.nbytes, 128), 32));
  }
  const size_t nbits_words = offs.back();
  const size_t need = sizeof(Acct) + 256 + nbits_words * 4 + 8 +
                      offs.size() * 4 + h.n * 8 + 256;
  void *buf = nullptr;
  st = ctx->pool->alloc(need, &buf);
  if (st) {
    ctx->pool->free(dummy_out);
    return st;
  }
  char *b = static_cast<char *>(buf);
  Acct *d_acct = reinterpret_cast<Acct *>(b);
  char *cur = b + ((sizeof(Acct) + 255) / 256) * 256;
  uint32_t *d_bits = reinterpret_cast<uint32_t *>(cur);
  cur += nbits_words * 4;
  cur = reinterpret_cast<char *>((reinterpret_cast<uintptr_t>(cur) + 7) & ~7ull);
  unsigned long long *d_sel = reinterpret_cast<unsigned long long *>(cur);
  cur += 8;
  unsigned long long *d_lines = reinterpret_cast<unsigned long long *>(cur);
  cur += h.n * 8;
  uint32_t *d_offs = reinterpret_cast<uint32_t *>(cur);
  h.bits = d_bits;
  h.selected = d_sel;
  if (nbits_words) st = launch_fill_u32(ctx->stream, d_bits, 0, nbits_words);
  if (!st) st = launch_fill_u64(ctx->stream, d_sel, 0, 1 + h.n);
  if (!st) {
    hipError_t e = hipMemcpyAsync(d_acct, &h, sizeof(Acct),
                                  hipMemcpyHostToDevice, ctx->stream);
    if (e == hipSuccess)
      e = hipMemcpyAsync(d_offs, offs.data(), offs.size() * 4,
                         hipMemcpyHostToDevice, ctx->stream);
    if (e != hipSuccess) st = hip_fail(e, "alg_bytes upload");
  }
  for (uint32_t qi = 0; qi < nq && !st; ++qi) st = launch_fused(ctx, plans[qi], d_acct);
  if (!st && h.n) {
    hipLaunchKernelGGL(acct_count_kernel, dim3(h.n), dim3(256), 0, ctx->stream,
                       d_bits, d_offs, h.n, d_lines);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) st = hip_fail(e, "alg_bytes count");
  }
  std::vector<unsigned long long> lines(h.n + 1, 0);
  if (!st) {
    hipError_t e = hipMemcpyAsync(lines.data(), d_sel, (h.n + 1) * 8,
                                  hipMemcpyDeviceToHost, ctx->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
    if (e != hipSuccess) st = hip_fail(e, "alg_bytes readback");
  } else {
    (void)hipStreamSynchronize(ctx->stream);
  }
  ctx->pool->free(buf);
  ctx->pool->free(dummy_out);
  if (st) return st;
  uint64_t tot = 0;
  for (uint32_t k = 0; k < h.n; ++k) tot += lines[1 + k] * 128ull;
  tot += 8ull * groups;
  if (bytes) *bytes = tot;
  if (selected) *selected = lines[0];
  return MBX_OK;
}

}  // namespace mbx

extern "C" {

int mbx_alg_bytes(mbx_ctx *ctx, const mbx_vprg *prg, const mbx_rowop *op,
                  uint32_t nr_grp, uint64_t *bytes, uint64_t *selected) {
  if (!prg || !op) return fail(MBX_ERR_INVALID, "alg_bytes: null");
  mbx_query q;
  std::memset(&q, 0, sizeof(q));
  q.prg = *prg;
  q.op = *op;
  q.nr_grp = nr_grp;
  return alg_bytes_multi(ctx, &q, 1, bytes, selected);
}

int mbx_alg_bytes_batch(mbx_ctx *ctx, const mbx_query *qs, uint32_t nq,
                        uint64_t *bytes, uint64_t *selected) {
  return alg_bytes_multi(ctx, qs, nq, bytes, selected);
}

int mbx_col_stats(mbx_ctx *ctx, const mbx_vprg *prg, double dens[MBX_MAXCOLS],
                  double edge[MBX_MAXCOLS]) {
  if (!ctx || !dens) return fail(MBX_ERR_INVALID, "col_stats: null");
  DevPrg P;
  int st = make_dev_prg(prg, &P);
  if (st) return st;
  void *buf = nullptr;
  st = ctx->pool->alloc(16 * MBX_MAXCOLS, &buf);
  if (st) return st;
  unsigned long long *d = static_cast<unsigned long long *>(buf);
  st = launch_fill_u64(ctx->stream, d, 0, 2 * MBX_MAXCOLS);
  const unsigned grid = persistent_grid(ctx, (uint32_t)ceil_div(P.nwords, 256), 4);
  for (int c = 0; c < MBX_MAXCOLS && !st; ++c) {
    if (!P.cols[c].nm && !P.cols[c].has_edge) continue;
    hipLaunchKernelGGL(density_kernel, dim3(grid), dim3(256), 0, ctx->stream,
                       P.cols[c], P.nwords, P.nrows, d + 2 * c);
  }
  unsigned long long h[2 * MBX_MAXCOLS] = {};
  hipError_t e = hipGetLastError();
  if (!st && e == hipSuccess)
    e = hipMemcpyAsync(h, d, sizeof(h), hipMemcpyDeviceToHost, ctx->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
  ctx->pool->free(buf);
  if (st) return st;
  if (e != hipSuccess) return hip_fail(e, "col_stats");
  for (int c = 0; c < MBX_MAXCOLS; ++c) {
    dens[c] = (double)h[2 * c] / P.nrows;
    if (edge) edge[c] = (double)h[2 * c + 1] / P.nrows;
  }
  return MBX_OK;
}

int mbx_col_density(mbx_ctx *ctx, const mbx_vprg *prg, double dens[MBX_MAXCOLS]) {
  return mbx_col_stats(ctx, prg, dens, nullptr);
}

}  // extern "C"
