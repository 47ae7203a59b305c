// gen.hip -- seeded synthetic column generators, bit-identical on host and
// device.
//
// Value i of a column is a pure function of (seed, i): splitmix64's output
// mix (Steele, Lea, Flood 2014; Vigna's public reference constants) applied to
// seed + (i+1)*golden, then
//   uniform [lo, hi): lo + floor(u * (hi - lo) / 2^64)  (Lemire's
//                     multiply-high mapping), and
//   finite Zipf(s) on {0..n-1}, P(v) = (v+1)^-s / H_{n,s}: the smallest v with
//                     u < T[v], where T[v] = floor(2^64 * CDF(v)) is computed
//                     once on the host in long double and shared by both
//                     sides, so the comparison is integer-exact.
// The reference's generator (synth/onegen_zipf.py:5-10) is unseeded numpy
// zipf, which needs s > 1; configs here use s in {0.5, 1.0, 1.2, 1.5}.
#include <cmath>
#include <cstring>
#include <vector>

#include "mbx_internal.hpp"

namespace mbx {

__host__ __device__ __forceinline__ uint64_t mix64(uint64_t z) {
  z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
  z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
  return z ^ (z >> 31);
}

__host__ __device__ __forceinline__ uint64_t rand_at(uint64_t seed,
                                                     uint64_t i) {
  return mix64(seed + (i + 1) * 0x9e3779b97f4a7c15ull);
}

__host__ __device__ __forceinline__ uint64_t mulhi64(uint64_t a, uint64_t b) {
#ifdef __HIP_DEVICE_COMPILE__
  return __umul64hi(a, b);
#else
  return (uint64_t)(((unsigned __int128)a * b) >> 64);
#endif
}

__host__ __device__ __forceinline__ uint32_t uniform_at(uint64_t seed,
                                                        uint64_t i,
                                                        uint32_t lo,
                                                        uint32_t hi) {
  return lo + (uint32_t)mulhi64(rand_at(seed, i), (uint64_t)(hi - lo));
}

// T has n entries, T[n-1] = UINT64_MAX (u < T is then always true except
// for u == UINT64_MAX, which also maps to n-1 below).
__host__ __device__ __forceinline__ uint32_t zipf_at(uint64_t seed, uint64_t i,
                                                     const uint64_t *T,
                                                     uint32_t n) {
  const uint64_t u = rand_at(seed, i);
  uint32_t lo = 0, hi = n - 1;   // answer in [lo, hi]
  while (lo < hi) {
    const uint32_t mid = (lo + hi) >> 1;
    if (u < T[mid]) hi = mid;
    else lo = mid + 1;
  }
  return lo;
}

__global__ void gen_uniform_kernel(uint32_t *out, size_t n, uint64_t seed,
                                   uint32_t lo, uint32_t hi) {
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (size_t)gridDim.x * blockDim.x)
    out[i] = uniform_at(seed, i, lo, hi);
}

__global__ void gen_zipf_kernel(uint32_t *out, size_t n, uint64_t seed,
                                const uint64_t *T, uint32_t nv) {
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (size_t)gridDim.x * blockDim.x)
    out[i] = zipf_at(seed, i, T, nv);
}

// out[i] = table[in[i]] (dimension attribute projection, used to derive
// nested attributes such as city -> nation)
__global__ void gen_map_kernel(uint32_t *out, const uint32_t *in, size_t n,
                               uint32_t div, uint32_t mod) {
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (size_t)gridDim.x * blockDim.x) {
    uint32_t v = in[i] / div;
    out[i] = mod ? v % mod : v;
  }
}

static int zipf_table(uint32_t n, double s, std::vector<uint64_t> &T) {
  if (n == 0 || !(s >= 0.0) || s > 20.0)
    return fail(MBX_ERR_INVALID, "zipf: need n > 0 and 0 <= s <= 20");
  T.resize(n);
  long double H = 0;
  for (uint32_t v = 0; v < n; ++v) H += powl((long double)(v + 1), -s);
  long double c = 0;
  const long double two64 = 18446744073709551616.0L;
  for (uint32_t v = 0; v < n; ++v) {
    c += powl((long double)(v + 1), -s);
    long double t = c / H * two64;
    T[v] = (t >= two64) ? UINT64_MAX : (uint64_t)t;
  }
  T[n - 1] = UINT64_MAX;
  return MBX_OK;
}

static unsigned gen_grid(mbx_ctx *ctx, size_t n) {
  size_t g = ceil_div(n, 256);
  const size_t cap = (size_t)ctx->num_cus * 16;
  if (g > cap) g = cap;
  return g ? (unsigned)g : 1u;
}

}  // namespace mbx

using namespace mbx;

extern "C" {

uint64_t mbx_rand_at(uint64_t seed, uint64_t i) { return rand_at(seed, i); }

int mbx_zipf_table(uint32_t nvals, double s, uint64_t *thresholds) {
  std::vector<uint64_t> T;
  int st = zipf_table(nvals, s, T);
  if (st) return st;
  if (thresholds) std::memcpy(thresholds, T.data(), T.size() * 8);
  return MBX_OK;
}

int mbx_gen_uniform_u32_host(uint32_t *out, size_t n, uint64_t seed,
                             uint32_t lo, uint32_t hi) {
  if (!out || hi <= lo) return fail(MBX_ERR_INVALID, "gen_uniform: bad range");
  for (size_t i = 0; i < n; ++i) out[i] = uniform_at(seed, i, lo, hi);
  return MBX_OK;
}

int mbx_gen_zipf_u32_host(uint32_t *out, size_t n, uint64_t seed,
                          uint32_t nvals, double s) {
  if (!out) return fail(MBX_ERR_INVALID, "gen_zipf: null");
  std::vector<uint64_t> T;
  int st = zipf_table(nvals, s, T);
  if (st) return st;
  for (size_t i = 0; i < n; ++i) out[i] = zipf_at(seed, i, T.data(), nvals);
  return MBX_OK;
}

int mbx_gen_uniform_u32(mbx_ctx *ctx, uint32_t *d_out, size_t n,
                        uint64_t seed, uint32_t lo, uint32_t hi) {
  if (!ctx || !d_out || hi <= lo)
    return fail(MBX_ERR_INVALID, "gen_uniform: bad argument");
  if (n == 0) return MBX_OK;
  hipLaunchKernelGGL(gen_uniform_kernel, dim3(gen_grid(ctx, n)), dim3(256), 0,
                     ctx->stream, d_out, n, seed, lo, hi);
  MBX_HIP(hipGetLastError());
  return MBX_OK;
}

int mbx_gen_zipf_u32(mbx_ctx *ctx, uint32_t *d_out, size_t n, uint64_t seed,
                     uint32_t nvals, double s) {
  if (!ctx || !d_out) return fail(MBX_ERR_INVALID, "gen_zipf: null");
  std::vector<uint64_t> T;
  int st = zipf_table(nvals, s, T);
  if (st) return st;
  if (n == 0) return MBX_OK;
  void *dT = nullptr;
  st = ctx->pool->alloc(T.size() * 8, &dT);
  if (st) return st;
  hipError_t e = hipMemcpyAsync(dT, T.data(), T.size() * 8,
                                hipMemcpyHostToDevice, ctx->stream);
  if (e == hipSuccess) {
    hipLaunchKernelGGL(gen_zipf_kernel, dim3(gen_grid(ctx, n)), dim3(256), 0,
                       ctx->stream, d_out, n, seed,
                       static_cast<const uint64_t *>(dT), nvals);
    e = hipGetLastError();
  }
  hipError_t e2 = hipStreamSynchronize(ctx->stream);
  ctx->pool->free(dT);
  if (e == hipSuccess) e = e2;
  if (e != hipSuccess) return hip_fail(e, "gen_zipf");
  return MBX_OK;
}

int mbx_gen_map_u32(mbx_ctx *ctx, uint32_t *d_out, const uint32_t *d_in,
                    size_t n, uint32_t div, uint32_t mod) {
  if (!ctx || !d_out || !d_in || div == 0)
    return fail(MBX_ERR_INVALID, "gen_map: bad argument");
  if (n == 0) return MBX_OK;
  hipLaunchKernelGGL(gen_map_kernel, dim3(gen_grid(ctx, n)), dim3(256), 0,
                     ctx->stream, d_out, d_in, n, div, mod);
  MBX_HIP(hipGetLastError());
  return MBX_OK;
}

}  // extern "C"
