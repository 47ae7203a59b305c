// mbx_internal.hpp -- shared host-side declarations of libmbx (not installed).
#pragma once
#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>
#include <map>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/mbx.h"

namespace mbx {

// ------------------------------------------------------------------ errors
// Per-thread detail message behind mbx_last_error().
void set_error(const std::string &msg);
int fail(int status, const std::string &msg);
int hip_fail(hipError_t e, const char *what);

#define MBX_HIP(call)                                                          \
  do {                                                                         \
    hipError_t e_ = (call);                                                    \
    if (e_ != hipSuccess) return ::mbx::hip_fail(e_, #call);                   \
  } while (0)

// ------------------------------------------------------------------- pool
// First-fit allocator with coalescing over one or more arenas, 256-byte
// granules.  The bookkeeping is host-only; the arena memory is hipMalloc'd
// (device pool) or aligned_alloc'd (host pool used by CPU tests).  Replaces
// merle/dumb_pool.cuh:39-95, which is a bump allocator with 48 live slots.
class Pool {
 public:
  static constexpr size_t kAlign = 256;
  static constexpr size_t kArena = size_t(256) << 20;
  // fixed_capacity == 0: grow in kArena arenas; else one arena of that size.
  Pool(bool device, size_t fixed_capacity);
  ~Pool();
  Pool(const Pool &) = delete;
  Pool &operator=(const Pool &) = delete;

  int alloc(size_t bytes, void **out);
  int free(void *p);
  void stats(size_t *used, size_t *capacity, size_t *peak, size_t *nlive,
             size_t *largest_free) const;
  int init_status() const { return init_status_; }

 private:
  struct Arena {
    char *base;
    size_t size;
  };
  int add_arena(size_t size);
  void *raw_alloc(size_t size);
  void raw_free(void *p);

  bool device_;
  bool fixed_;
  int init_status_ = MBX_OK;
  std::vector<Arena> arenas_;
  std::map<char *, size_t> free_;   // start -> size (coalesced)
  std::map<char *, size_t> live_;   // start -> size
  size_t used_ = 0, peak_ = 0, capacity_ = 0;
  mutable std::mutex mu_;
};

}  // namespace mbx

struct mbx_ctx {
  int device = 0;
  hipStream_t own_stream = nullptr;
  hipStream_t stream = nullptr;
  int num_cus = 256;
  size_t lds_per_cu = 160 * 1024;
  size_t lds_per_block = 64 * 1024;
  mbx::Pool *pool = nullptr;
  // scratch reused by the plans (one query at a time per context)
  uint32_t *d_counter = nullptr;   // 4 uint32 work/counter slots
  // replicated flush targets of the LDS group tables: kSpreadRep tables of
  // kLdsGroups int64 and kSpreadRep counters 128 B apart; all zero between
  // launches (the last workgroup of each replica clears what it used)
  unsigned long long *spread = nullptr;
  uint32_t *spread_done = nullptr;
  // the same for every other stream the context has launched a query on
  // (mbx_set_stream), so queries on different streams never share replicas
  std::mutex spread_mu;
  std::map<hipStream_t, std::pair<unsigned long long *, uint32_t *>> spreads;
  // mbx_stream_join's event per joined stream (guarded by spread_mu)
  std::map<hipStream_t, hipEvent_t> join_events;
  // the WAH count scan's tile sums per stream: (pool memory, bytes), grown
  // on demand (guarded by spread_mu)
  std::map<hipStream_t, std::pair<void *, size_t>> wah_tiles;
};

struct mbx_timer {
  hipEvent_t a = nullptr, b = nullptr;
};

namespace mbx {

// ------------------------------------------------- device-side descriptors
// Kernel-argument forms of mbx_vprg / mbx_rowop: pointers are typed as
// 64-row words, middle-bin counts and instruction counts are precomputed, so
// the kernels branch only on wave-uniform values.
// Operand bitmaps a program can read: 4 columns x (4 middle + 2 edge bins).
constexpr uint32_t kMaxOperands = MBX_MAXCOLS * (MBX_MAXBIN_PERCOL + 2);

struct DevCol {
  const uint64_t *m[MBX_MAXBIN_PERCOL];
  const uint64_t *l, *r;
  uint32_t nm;      // leading non-null middle bins
  uint32_t has_edge;
};
struct DevPrg {
  DevCol cols[MBX_MAXCOLS];
  uint8_t ins[MBX_MAXNINSTR][4];
  uint32_t nins;
  uint32_t nrows;
  uint32_t nwords;  // ceil(nrows / 64)
  uint32_t chk;     // any edge bin -> three-valued evaluation
  uint32_t memcols; // bit c: some ANDM/ORM reads column c
  // Flat list of the operand bitmaps of the columns in memcols, column by
  // column (middle bins, then edge bins); column c owns [beg[c], end[c]).
  const uint64_t *fl[kMaxOperands];
  uint8_t flcol[kMaxOperands];
  uint8_t fledge[kMaxOperands];
  uint8_t beg[MBX_MAXCOLS], end[MBX_MAXCOLS];
  uint8_t flend[kMaxOperands];   // entry is the last of its column
  uint8_t flneg[kMaxOperands];   // (last entries) the column enters negated
  uint32_t nfl;
  // the program computes a conjunction of column literals: r0 = AND over
  // columns c of (OR of c's bins), complemented where flneg
  // (e.g. ORM r0<-a; ANDM r0<-b; ORM r1<-c; NOT r1; AND r0,r1 = a & b & ~c)
  uint32_t conj;
  // the program as a boolean function of its columns' exact predicates:
  // bit m is r0 when column c's predicate is bit c of m (at most 4 columns)
  uint32_t truth;
  // columns whose predicate the re-check of a candidate row need not load
  // (bit c), and the value it takes for them (known_val bit c)
  uint32_t known, known_val;
  // ablation only (abla::base_fuse / _stg1 with idxbuf): the program's
  // registers in device memory, 6 uint64 per 64-row word (P r0..r2, then
  // U r0..r2); null in every product plan
  uint64_t *greg;
};
struct DevSrc {
  const void *data;
  const uint32_t *fk;
  uint32_t dimsz;
  uint32_t eb;
};
struct DevKey {
  DevSrc src;
  uint32_t kind, div;
  int64_t base, stride;
  uint32_t nb;
  uint8_t sh1, sh2, pad[2];   // KEY_DIV: shifts of the division by div
  uint64_t magic;             // KEY_DIV: multiplier of the division by div
  uint64_t bounds[MBX_MAXBOUNDS];
};
struct DevOp {
  DevSrc pred[MBX_MAXCOLS];
  uint64_t lo[MBX_MAXCOLS], hi[MBX_MAXCOLS];
  uint32_t npreds, nkeys, measure, pad;
  DevKey keys[MBX_MAXKEYS];
  DevSrc ma, mb;
};

// Validation + conversion; returns MBX_OK or an error (with message).
int make_dev_prg(const mbx_vprg *p, DevPrg *out);
// Compile-time shape code of a conjunction of single-column clauses (the
// dedicated kernels' template argument; mbx_prog_shape documents the
// encoding), 0 for any other program.
uint32_t conj_shape_code(const DevPrg &P);
// Division of any 64-bit x by an invariant d >= 1 as
//   t = mulhi64(m, x);  q = (t + ((x - t) >> sh1)) >> sh2
// (Granlund and Montgomery, "Division by Invariant Integers using
// Multiplication", PLDI 1994, figure 4.1, N = 64): l = ceil(log2 d),
// m = floor(2^64 (2^l - d) / d) + 1, sh1 = min(l, 1), sh2 = max(l - 1, 0).
void key_div_magic(uint32_t d, uint64_t *m, uint8_t *sh1, uint8_t *sh2);
int make_dev_op(const mbx_rowop *op, uint32_t npreds_hint, DevOp *out);
int check_src(const mbx_colref &c, const char *what);

inline size_t ceil_div(size_t a, size_t b) { return (a + b - 1) / b; }

// ------------------------------------------------------- launch helpers
int launch_fill_u32(hipStream_t s, void *dst, uint32_t v, size_t count);
int launch_fill_u64(hipStream_t s, void *dst, uint64_t v, size_t count);

// Group-table capacity kept in LDS by the query kernels (int64 slots).
constexpr uint32_t kLdsGroups = 4096;
// Row-list capacity of one compaction round (uint32 entries).
constexpr uint32_t kListCap = 2048;
// Replicas of the group table that workgroups flush into (see flush_groups).
constexpr uint32_t kSpreadRep = 64;
constexpr uint32_t kSpreadDoneStride = 32;   // uint32 counters, 128 B apart

}  // namespace mbx
