// Host stand-in for <hip/hip_runtime.h>: lets the device query functions of
// merle_amd/csrc/devquery.hpp compile as plain C++ for a CPU check of their
// logic.  One lane: __ballot returns the lane's own bit, shuffles return the
// lane's own value, a wave reduction adds nothing from other lanes, and
// atomics are plain read-modify-writes.
#pragma once
#include <algorithm>
#include <cstddef>
#include <cstdint>
#include <cstdlib>
typedef int hipError_t;
typedef void *hipStream_t;
typedef void *hipEvent_t;
#define hipSuccess 0
#define hipErrorOutOfMemory 2
inline const char *hipGetErrorString(hipError_t) { return "host"; }
inline hipError_t hipMalloc(void **p, size_t n) { *p = malloc(n); return *p ? 0 : 2; }
inline hipError_t hipFree(void *p) { free(p); return 0; }
inline hipError_t hipGetLastError() { return 0; }
#define __device__
#define __host__
#define __global__
#define __forceinline__ inline
#define __noinline__
#define __shared__
#define __launch_bounds__(x)
struct host_dim3 { unsigned x, y, z; };
extern thread_local host_dim3 threadIdx, blockIdx, blockDim, gridDim;
template <class T> T host_ntload(const T *p) { return *p; }
#define __builtin_nontemporal_load(p) host_ntload(p)
inline int __popcll(unsigned long long x) { return __builtin_popcountll(x); }
inline unsigned long long __umul64hi(unsigned long long a, unsigned long long b) {
  return (unsigned long long)(((unsigned __int128)a * b) >> 64);
}
inline unsigned __shfl_up(unsigned v, int, int) { return v; }
inline unsigned __shfl(unsigned v, int, int) { return v; }
#define __builtin_amdgcn_readlane(v, l) (v)
inline void __syncthreads() {}
#define __builtin_amdgcn_fence(a, b)
#define __builtin_amdgcn_wave_barrier()
inline unsigned long long atomicAdd(unsigned long long *p, unsigned long long v) {
  auto o = *p; *p += v; return o;
}
inline unsigned atomicOr(unsigned *p, unsigned v) { auto o = *p; *p |= v; return o; }
using std::max;
using std::min;
#define __HIP_MEMORY_SCOPE_AGENT 3
template <typename T, typename U>
inline T __hip_atomic_fetch_add(T *p, U v, int, int) { T o = *p; *p = o + (T)v; return o; }
template <typename T, typename U>
inline T __hip_atomic_exchange(T *p, U v, int, int) { T o = *p; *p = (T)v; return o; }
inline unsigned long long __ballot(bool p) { return p ? 1ull : 0ull; }
inline unsigned long long __shfl_xor(unsigned long long, int, int) { return 0ull; }
