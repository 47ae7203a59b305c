// devfuncs.hpp -- device helpers shared by the gfx950 kernels.
#pragma once
#include "mbx_internal.hpp"

namespace mbx {

// Streaming (non-temporal) loads for data read once per query -- bitmaps and
// fact columns; ordinary cached loads for dimension tables, which are small
// and re-read through foreign keys (the reference makes the same split,
// bmpidx_create.cu:98-101).
template <typename T>
__device__ __forceinline__ T ld_stream(const T *p) {
  return __builtin_nontemporal_load(p);
}

__device__ __forceinline__ uint64_t ld_elem(const void *p, uint32_t eb,
                                            size_t i, bool stream) {
  switch (eb) {
  case 1: return static_cast<const uint8_t *>(p)[i];
  case 2: return static_cast<const uint16_t *>(p)[i];
  case 4:
    return stream ? ld_stream(static_cast<const uint32_t *>(p) + i)
                  : static_cast<const uint32_t *>(p)[i];
  default:
    return stream ? ld_stream(static_cast<const uint64_t *>(p) + i)
                  : static_cast<const uint64_t *>(p)[i];
  }
}

// Value of a source for fact row `row`.  false when fk >= dimsz.
__device__ __forceinline__ bool fetch_src(const DevSrc &s, size_t row,
                                          uint64_t *v) {
  if (s.fk == nullptr) {
    *v = ld_elem(s.data, s.eb, row, false);
    return true;
  }
  const uint32_t k = s.fk[row];
  if (k >= s.dimsz) return false;
  *v = ld_elem(s.data, s.eb, k, false);
  return true;
}

// wave64 helpers ---------------------------------------------------------
__device__ __forceinline__ uint32_t lane_id() { return threadIdx.x & 63u; }

// Lane src's v for a wave-uniform src: two scalar lane reads, no LDS trip.
__device__ __forceinline__ uint64_t wave_read64(uint64_t v, uint32_t src) {
  const uint32_t lo =
      (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)v, (int)src);
  const uint32_t hi =
      (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)(v >> 32), (int)src);
  return ((uint64_t)hi << 32) | lo;
}

// Sum of x over the 64 lanes of a wave (every lane gets it).
__device__ __forceinline__ unsigned long long wave_sum_u64(unsigned long long x) {
#pragma unroll
  for (int d = 32; d > 0; d >>= 1) x += __shfl_xor(x, d, 64);
  return x;
}

// Inclusive prefix sum across the 64 lanes of a wave.
__device__ __forceinline__ uint32_t wave_incl_scan(uint32_t x) {
  const uint32_t lane = lane_id();
#pragma unroll
  for (int d = 1; d < 64; d <<= 1) {
    uint32_t y = __shfl_up(x, d, 64);
    if (lane >= (uint32_t)d) x += y;
  }
  return x;
}

// Orders this wave's LDS writes before its own later LDS reads: the LDS
// executes a wave's instructions in issue order, so only the compiler must
// be kept from moving accesses across this point.
__device__ __forceinline__ void wave_lds_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

}  // namespace mbx
