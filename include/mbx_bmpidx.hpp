// mbx_bmpidx.hpp -- C++ host API mirroring the reference's namespace
// mybmpidx (primitive.cuh:120-474, bmpidx_create.cu:1-317), built on the
// C ABI in mbx.h.  Same names, argument meanings and failure behaviour:
// create_select/create_join return NULL on failure, col::from_bin returns an
// empty col, release() is explicit ("NOT AN AUTO FREE DTOR").
//
// Differences, all deliberate:
//  - memory comes from the context's pool (mbx_alloc) instead of cudaMalloc,
//    and a failed allocation is reported, never exit()ed;
//  - the per-query device functor op_t is replaced by the data descriptor
//    mbx_rowop (see mbx.h), so vprg_grpby is a host call, not a template;
//  - sums are int64.
#ifndef MBX_BMPIDX_HPP
#define MBX_BMPIDX_HPP
#include <cstddef>
#include <cstdint>

#include "mbx.h"

namespace mybmpidx {

using uint = uint32_t;
static constexpr size_t MAXBIN_PERCOL = MBX_MAXBIN_PERCOL,
                        MAXCOLS = MBX_MAXCOLS, MAXNINSTR = MBX_MAXNINSTR;

// The context every call below uses.  The first call creates one on device 0
// (the reference relies on the implicit CUDA context the same way).
mbx_ctx *context();
void use_context(mbx_ctx *ctx);

// bmpidx_create.cu:58-136.  ALL ranges are left inclusive.
uint *create_select(const void *src, size_t nbit, uint factsz, uint min,
                    uint max);
uint *create_join(const uint32_t *fk, const void *dimattr, size_t nbit,
                  uint factsz, uint dimsz, uint min, uint max);
// bmpidx_create.cu:14-56.  ONLY FREE OUT[0]!  Returns the single-bitmap byte
// size, or 0 (out[] untouched) on failure.
size_t create_bin(const uint *fk, const void *attr, size_t nbit, uint factsz,
                  uint dimsz, const uint64_t *min, const uint64_t *max,
                  size_t ncol, uint **out);
void free_bitmap(const uint *p);

// Bitmap index of a column (primitive.cuh:135-162).
struct col {
  uint *middle[MAXBIN_PERCOL] = {};
  uint *leftmost = nullptr, *rightmost = nullptr;

  void release() noexcept;
  bool is_nochk() const noexcept {
    return leftmost == nullptr && rightmost == nullptr;
  }
  bool empty() const noexcept {
    return !middle[0] && !leftmost && !rightmost;
  }
  // Bins [bounds[i], bounds[i+1]) -> the range [min, max).  Empty on failure
  // (including more than MAXBIN_PERCOL middle bins, which the reference
  // silently overruns, bmpidx_create.cu:160-161).
  static col from_bin(size_t n, uint *bins[], const size_t bounds[],
                      size_t min, size_t max);
  // Coarse bins unless they need a candidate check, then fine bins
  // (bmpidx_create.cu:165-173).
  static col from_bin2(size_t nc, uint *coarse_bins[], const size_t cbounds[],
                       size_t nf, uint *fine_bins[], const size_t fbounds[],
                       size_t min, size_t max);
  mbx_col to_abi() const;
};

// Virtual bitmap program (primitive.cuh:164-198).
struct vprg {
  col cols[MAXCOLS];
  uint factsz = 0;
  enum ops : uint8_t { END, AND, OR, NOT, ANDM, ORM };
  struct instr {
    ops opcode = END;
    uint8_t dstreg = 0;
    uint8_t src1 = 0;
    uint8_t src2 = 0;
  };
  instr instrs[MAXNINSTR];

  void release() noexcept;
  bool is_nochk() const noexcept;
  mbx_vprg to_abi() const;
};

// Per-query index synthesis (primitive.cuh:447-474, bmpidx_create.cu:228-316).
struct recipe {
  static constexpr uint nodim = 0xffffffffu;
  uint factsz = 0;
  uint nbit[MAXCOLS] = {}, min[MAXCOLS] = {}, max[MAXCOLS] = {},
       dimsz[MAXCOLS] = {};
  const uint32_t *fk[MAXCOLS] = {};
  const void *attr[MAXCOLS] = {};

  vprg perfect(const vprg::instr instrs[MAXNINSTR]) const;
  vprg many_or(const vprg::instr instrs[MAXNINSTR]) const;
  vprg candchk(const vprg::instr instrs[MAXNINSTR]) const;
};

// The fused query (vprg_grpby, primitive.cuh:340-445) and its alternative
// plans.  out: device int64[nr_grp], zeroed by the call.  Returns an mbx
// status.
int vprg_grpby(const vprg &p, const mbx_rowop &op, uint nr_grp, int64_t *out,
               uint32_t plan = MBX_PLAN_FUSED);

}  // namespace mybmpidx

extern "C" {
// Host-only planning helpers exported for tests and other front ends.
// Slot codes: 0..3 middle[k], 4 leftmost, 5 rightmost.
int mbx_col_from_bin(size_t nbin, const uint64_t *bin_ids,
                     const uint64_t *bounds, uint64_t min, uint64_t max,
                     uint64_t *slots /* 6 ids, UINT64_MAX = none */);
int mbx_col_from_bin2(size_t nc, const uint64_t *cids, const uint64_t *cb,
                      size_t nf, const uint64_t *fids, const uint64_t *fb,
                      uint64_t min, uint64_t max, uint64_t *slots);
// recipe bin plan of one column: kind 0 perfect, 1 many_or, 2 candchk.
// Writes up to 6 (lo, hi, slot) triples; returns their count in *n.
int mbx_recipe_plan(int kind, int colidx, uint32_t min, uint32_t max,
                    uint32_t *lo, uint32_t *hi, int *slot, int *n);
}

#endif
