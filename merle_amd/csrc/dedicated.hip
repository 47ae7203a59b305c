// dedicated.hip -- the hardcoded-program ablation (abla::hardcoded,
// primitive_abla.cuh:188-260, and the per-query c::s** functions of
// ssb/inst/*c.cu): kernels built for one program shape each, the flat
// operand list of the conjunction form fixed at compile time.
//
// The reference writes the bitmap expression of each SSB case by hand as a
// device lambda (e.g. x = d1; x &= q0 | q1 | q2; x &= di0 | di1,
// ssb/inst/1c.cu:16-19) and keeps the row op of the interpreted path.  Here
// the same split holds: a dedicated kernel runs the fused kernel's pipeline
// -- compaction, row op, sink and flush unchanged -- with the conjunction
// form reading its shape (entries per clause, edge bins, negations) from a
// template argument instead of the program descriptor, so every loop over
// the operand list is unrolled and every flag test folded away.  The bitmap
// pointers stay kernel arguments, and the group table stays in LDS (at most
// 4,096 groups, as in every SSB flight).  A shape is the code mbx_prog_shape
// computes; the library holds kernels for the shapes of the SSB flights'
// fixed-layout programs (ssb/fixedlayout.cu), as written and in the column
// order of the static plan, and MBX_PLAN_DEDICATED on any other program
// returns MBX_ERR_UNSUPPORTED (the reference's DEDICATE reads 0 where no
// dedicated function exists, ssb/fixedlayout.cu:19-22).
#include <algorithm>
#include <cstring>
#include <map>
#include <mutex>

#include "devquery.hpp"
#include "query_launch.hpp"

// Shape codes with a dedicated kernel: clause k at bits 6k..6k+5 =
// nm | left << 3 | right << 4 | negated << 5.  These are the 26 shapes of
// the 39 SSB fixed-layout programs (13 flights x Sparse / Medium / Dense,
// ssb/fixedlayout.cu:35-140, over the bin sets of ssb/bmpidx-create.cu:8-30)
// as written and in the column order the static plan gives them on the
// reference generator's data (ssb-dbgen-lite.c); e.g. 0x12601 is SSB11
// Sparse: date (1 middle bin), discount (left and right edge bins only),
// quantity (2 middle bins and a right edge bin).
#define MBX_DEDICATED_SHAPES(X) \
  X(0x41u) X(0x48u) X(0x201u) X(0x1088u) \
  X(0x1208u) X(0x2041u) X(0x2048u) X(0x20c1u) \
  X(0x20c8u) X(0x3081u) X(0x3088u) X(0x8201u) \
  X(0x8208u) X(0x12601u) X(0x140c1u) X(0x18048u) \
  X(0x18608u) X(0x21041u) X(0x21201u) X(0x21208u) \
  X(0x81041u) X(0x81201u) X(0x82041u) X(0x88041u) \
  X(0x88048u) X(0x88201u)

namespace mbx {

// The shape of a code as compile-time functions of the flat-list entry.
template <uint32_t CODE>
struct StaticShape {
  static constexpr bool kNT = true;   // bitmap words: streaming loads
  static constexpr uint32_t kRecheck = 4;   // words a round of the dense re-check
  static constexpr uint32_t clause(uint32_t k) { return (CODE >> (6 * k)) & 63u; }
  static constexpr uint32_t entries(uint32_t k) {
    return (clause(k) & 7u) + ((clause(k) >> 3) & 1u) + ((clause(k) >> 4) & 1u);
  }
  static constexpr uint32_t count() {
    uint32_t n = 0;
    for (uint32_t k = 0; k < MBX_MAXCOLS; ++k) n += entries(k);
    return n;
  }
  static constexpr bool any_edge() {
    bool e = false;
    for (uint32_t k = 0; k < MBX_MAXCOLS; ++k) e = e || (clause(k) & 24u) != 0;
    return e;
  }
  // entry e: (clause, position in the clause)
  static constexpr uint32_t clause_of(uint32_t e) {
    uint32_t k = 0;
    while (k + 1 < MBX_MAXCOLS && e >= entries(k)) e -= entries(k++);
    return k;
  }
  static constexpr uint32_t pos_of(uint32_t e) {
    uint32_t k = 0;
    while (k + 1 < MBX_MAXCOLS && e >= entries(k)) e -= entries(k++);
    return e;
  }
  static constexpr uint32_t kNfl = count();
  static constexpr bool kChk = any_edge();

  static constexpr bool edge_at(uint32_t e) {
    return pos_of(e) >= (clause(clause_of(e)) & 7u);
  }
  static constexpr bool end_at(uint32_t e) {
    return pos_of(e) + 1 == entries(clause_of(e));
  }
  static constexpr bool neg_at(uint32_t e) {
    return end_at(e) && ((clause(clause_of(e)) >> 5) & 1u) != 0;
  }

  static __device__ __forceinline__ uint32_t nfl(const DevPrg &) { return kNfl; }
  static __device__ __forceinline__ bool edge(const DevPrg &, uint32_t e) {
    return edge_at(e);
  }
  static __device__ __forceinline__ bool end(const DevPrg &, uint32_t e) {
    return end_at(e);
  }
  static __device__ __forceinline__ bool neg(const DevPrg &, uint32_t e) {
    return neg_at(e);
  }
};

struct DedArgs {
  DevPrg prg;
  DevOp op;
  unsigned long long *out;
  uint32_t nr_grp;
  Spread sp;
};

// The fused kernel (query.hip) with the conjunction form's shape fixed and
// the group table in LDS.
template <uint32_t CODE, int W>
__global__ __launch_bounds__(256) void dedicated_kernel(DedArgs a) {
  using S = StaticShape<CODE>;
  __shared__ __align__(16) uint32_t s_list[4][kWaveList];
  GroupSink sink;
  sink.lds = true;
  sink.out = a.out;
  for (uint32_t i = threadIdx.x; i < a.nr_grp; i += 256) s_grp_dyn[i] = 0;
  __syncthreads();
  const uint32_t wv = threadIdx.x >> 6;
  const uint32_t ws = gridDim.x * 4 * 64;
  for (uint32_t w0 = (blockIdx.x * 4 + wv) * 64; w0 < a.prg.nwords;
       w0 += ws * W)
    run_wave_range<S::kChk, false, W, 3, 2, false, false, S>(
        a.prg, a.op, a.nr_grp, w0, ws, a.prg.nwords, sink, s_list[wv], nullptr);
  flush_groups(a.sp, a.out, a.nr_grp, blockIdx.x, gridDim.x);
}

static unsigned ded_occupancy(const void *fn, size_t lds) {
  static std::mutex mu;
  static std::map<std::pair<const void *, size_t>, unsigned> cache;
  std::lock_guard<std::mutex> g(mu);
  const auto key = std::make_pair(fn, lds);
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

// Launch shape of the product's conjunction launches with an LDS group
// table (query.hip, fused_conj): one pass at 2 words per lane when the
// resident waves cover the words, else 4 words on the resident grid.
template <uint32_t CODE>
static int ded_launch(mbx_ctx *ctx, hipStream_t s, const DedArgs &a,
                      size_t lds) {
  const uint64_t chunks = ceil_div(a.prg.nwords, 64);
  const void *f2 = reinterpret_cast<const void *>(&dedicated_kernel<CODE, 2>);
  const uint64_t res2 = (uint64_t)ctx->num_cus * ded_occupancy(f2, lds) * 4;
  const uint64_t w2 = ceil_div(chunks, 2);
  if (w2 <= res2) {
    hipLaunchKernelGGL((dedicated_kernel<CODE, 2>),
                       dim3((unsigned)ceil_div(w2, 4)), dim3(256), lds, s, a);
  } else {
    const void *f4 = reinterpret_cast<const void *>(&dedicated_kernel<CODE, 4>);
    const uint64_t res4 = (uint64_t)ctx->num_cus * ded_occupancy(f4, lds) * 4;
    const uint64_t w4 = ceil_div(chunks, 4);
    hipLaunchKernelGGL((dedicated_kernel<CODE, 4>),
                       dim3((unsigned)ceil_div(std::min(w4, res4), 4)),
                       dim3(256), lds, s, a);
  }
  MBX_HIP(hipGetLastError());
  return MBX_OK;
}

template <uint32_t CODE>
static int ded_run(mbx_ctx *ctx, const QueryPlan &q) {
  DedArgs a;
  std::memset(&a, 0, sizeof(a));
  a.prg = q.prg;
  a.op = q.op;
  a.out = q.out;
  a.nr_grp = q.nr_grp;
  a.sp.gpad = (q.nr_grp + 15) / 16 * 16;
  const hipStream_t s = stream_of(ctx, q);
  if (q.spread) {
    a.sp.rep = q.spread;
    a.sp.done = q.spread_done;
  } else {
    const int st = ctx_spread(ctx, s, &a.sp.rep, &a.sp.done);
    if (st) return st;
  }
  return ded_launch<CODE>(ctx, s, a, (size_t)q.nr_grp * 8);
}

static bool ded_built(uint32_t code) {
  switch (code) {
#define MBX_DED_CASE(c) \
  case c:               \
    return c != 0u;
    MBX_DEDICATED_SHAPES(MBX_DED_CASE)
#undef MBX_DED_CASE
  default: return false;
  }
}

int launch_dedicated(mbx_ctx *ctx, const QueryPlan &q) {
  if (!q.lds)
    return fail(MBX_ERR_UNSUPPORTED,
                "dedicated kernels keep the group table in LDS (at most " +
                    std::to_string(kLdsGroups) + " groups)");
  const uint32_t code = conj_shape_code(q.prg);
  switch (code) {
#define MBX_DED_CASE(c) \
  case c:               \
    if constexpr (c != 0u) return ded_run<c>(ctx, q); \
    break;
    MBX_DEDICATED_SHAPES(MBX_DED_CASE)
#undef MBX_DED_CASE
  default: break;
  }
  return fail(MBX_ERR_UNSUPPORTED,
              code ? "no dedicated kernel for shape code " + std::to_string(code)
                   : std::string("the program is not a conjunction of "
                                 "single-column clauses"));
}

}  // namespace mbx

extern "C" int mbx_prog_shape(const mbx_vprg *prg, uint32_t *code, int *built) {
  if (!prg || !code) return mbx::fail(MBX_ERR_INVALID, "prog_shape: null");
  mbx::DevPrg P;
  const int st = mbx::make_dev_prg(prg, &P);
  if (st) return st;
  *code = mbx::conj_shape_code(P);
  if (built) *built = mbx::ded_built(*code) ? 1 : 0;
  return MBX_OK;
}
