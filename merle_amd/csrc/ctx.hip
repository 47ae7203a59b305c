// ctx.hip -- context, memory, fill/copy kernels, timers, descriptor checks.
#include <cstring>

#include "mbx_internal.hpp"
#include "query_launch.hpp"

namespace mbx {

__global__ void fill_u32_kernel(uint32_t *__restrict__ dst, uint32_t v,
                                size_t n) {
  size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  const size_t step = (size_t)gridDim.x * blockDim.x;
  for (; i < n; i += step) dst[i] = v;
}

__global__ void fill_u64_kernel(uint64_t *__restrict__ dst, uint64_t v,
                                size_t n) {
  size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  const size_t step = (size_t)gridDim.x * blockDim.x;
  for (; i < n; i += step) dst[i] = v;
}

// Plain 16-byte copy, the calibration kernel for HBM byte counters.
__global__ void copy_kernel(uint4 *__restrict__ dst,
                            const uint4 *__restrict__ src, size_t n16) {
  size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  const size_t step = (size_t)gridDim.x * blockDim.x;
  for (; i < n16; i += step) dst[i] = src[i];
}

static unsigned grid_for(size_t n, unsigned per_block) {
  size_t g = ceil_div(n, per_block);
  if (g > 8192) g = 8192;
  return g ? (unsigned)g : 1u;
}

int launch_fill_u32(hipStream_t s, void *dst, uint32_t v, size_t count) {
  if (count == 0) return MBX_OK;
  hipLaunchKernelGGL(fill_u32_kernel, dim3(grid_for(count, 256)), dim3(256), 0,
                     s, static_cast<uint32_t *>(dst), v, count);
  MBX_HIP(hipGetLastError());
  return MBX_OK;
}

int launch_fill_u64(hipStream_t s, void *dst, uint64_t v, size_t count) {
  if (count == 0) return MBX_OK;
  hipLaunchKernelGGL(fill_u64_kernel, dim3(grid_for(count, 256)), dim3(256), 0,
                     s, static_cast<uint64_t *>(dst), v, count);
  MBX_HIP(hipGetLastError());
  return MBX_OK;
}

}  // namespace mbx

using namespace mbx;

extern "C" {

int mbx_device_count(int *n) {
  if (!n) return fail(MBX_ERR_INVALID, "null");
  int c = 0;
  hipError_t e = hipGetDeviceCount(&c);
  if (e != hipSuccess) {
    (void)hipGetLastError();
    *n = 0;
    return MBX_OK;
  }
  *n = c;
  return MBX_OK;
}

int mbx_init(int device, size_t pool_bytes, mbx_ctx **out) {
  if (!out) return fail(MBX_ERR_INVALID, "mbx_init: null out");
  *out = nullptr;
  int n = 0;
  hipError_t e = hipGetDeviceCount(&n);
  if (e != hipSuccess || n == 0) {
    (void)hipGetLastError();
    return fail(MBX_ERR_NODEVICE, "mbx_init: no HIP device visible");
  }
  if (device < 0 || device >= n)
    return fail(MBX_ERR_INVALID, "mbx_init: bad device index");
  MBX_HIP(hipSetDevice(device));
  auto *c = new mbx_ctx;
  c->device = device;
  hipDeviceProp_t prop;
  if (hipGetDeviceProperties(&prop, device) == hipSuccess) {
    c->num_cus = prop.multiProcessorCount > 0 ? prop.multiProcessorCount : 256;
    c->lds_per_cu = prop.maxSharedMemoryPerMultiProcessor
                        ? prop.maxSharedMemoryPerMultiProcessor
                        : 160 * 1024;
    if (prop.sharedMemPerBlock) c->lds_per_block = prop.sharedMemPerBlock;
  }
  e = hipStreamCreateWithFlags(&c->own_stream, hipStreamNonBlocking);
  if (e != hipSuccess) {
    delete c;
    return hip_fail(e, "hipStreamCreate");
  }
  c->stream = c->own_stream;
  c->pool = new Pool(true, pool_bytes);
  if (c->pool->init_status() != MBX_OK) {
    (void)hipStreamDestroy(c->own_stream);
    delete c->pool;
    delete c;
    return MBX_ERR_OOM;
  }
  void *p = nullptr;
  int st = c->pool->alloc(256, &p);
  if (st != MBX_OK) {
    delete c->pool;
    (void)hipStreamDestroy(c->own_stream);
    delete c;
    return st;
  }
  c->d_counter = static_cast<uint32_t *>(p);
  const size_t rep_bytes = (size_t)kSpreadRep * kLdsGroups * 8;
  const size_t done_bytes = (size_t)kSpreadRep * kSpreadDoneStride * 4;
  st = c->pool->alloc(rep_bytes + done_bytes, &p);
  if (st == MBX_OK)
    st = launch_fill_u32(c->stream, p, 0, (rep_bytes + done_bytes) / 4);
  if (st == MBX_OK) {
    hipError_t se = hipStreamSynchronize(c->stream);
    if (se != hipSuccess) st = hip_fail(se, "mbx_init: scratch");
  }
  if (st != MBX_OK) {
    delete c->pool;
    (void)hipStreamDestroy(c->own_stream);
    delete c;
    return st;
  }
  c->spread = static_cast<unsigned long long *>(p);
  c->spread_done = reinterpret_cast<uint32_t *>(static_cast<char *>(p) + rep_bytes);
  // the no-allocation entry point's scratch on this device (a failure here
  // only disables that entry point, which then reports it)
  (void)noalloc_setup(c);
  *out = c;
  return MBX_OK;
}

int mbx_destroy(mbx_ctx *ctx) {
  if (!ctx) return MBX_OK;
  (void)hipSetDevice(ctx->device);
  (void)hipStreamSynchronize(ctx->stream);
  for (auto &kv : ctx->join_events)
    if (kv.second) (void)hipEventDestroy(kv.second);
  delete ctx->pool;
  if (ctx->own_stream) (void)hipStreamDestroy(ctx->own_stream);
  delete ctx;
  return MBX_OK;
}

int mbx_sync(mbx_ctx *ctx) {
  if (!ctx) return fail(MBX_ERR_INVALID, "null ctx");
  MBX_HIP(hipStreamSynchronize(ctx->stream));
  return MBX_OK;
}

void *mbx_stream(mbx_ctx *ctx) { return ctx ? (void *)ctx->stream : nullptr; }

int mbx_set_stream(mbx_ctx *ctx, void *s) {
  if (!ctx) return fail(MBX_ERR_INVALID, "null ctx");
  ctx->stream = s ? static_cast<hipStream_t>(s) : ctx->own_stream;
  return MBX_OK;
}

int mbx_stream_create(void **stream) {
  if (!stream) return fail(MBX_ERR_INVALID, "stream_create: null");
  hipStream_t s = nullptr;
  MBX_HIP(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
  *stream = s;
  return MBX_OK;
}

int mbx_stream_destroy(void *stream) {
  if (!stream) return MBX_OK;
  MBX_HIP(hipStreamDestroy(static_cast<hipStream_t>(stream)));
  return MBX_OK;
}

// The context's current stream waits (on the device) for the work enqueued
// so far on each of the given streams.
int mbx_stream_join(mbx_ctx *ctx, void *const *streams, int n) {
  if (!ctx || (n > 0 && !streams)) return fail(MBX_ERR_INVALID, "stream_join: null");
  std::lock_guard<std::mutex> g(ctx->spread_mu);
  for (int i = 0; i < n; ++i) {
    hipStream_t s = static_cast<hipStream_t>(streams[i]);
    if (!s || s == ctx->stream) continue;
    // one event per stream, re-recorded: a wait takes the record made
    // before it
    hipEvent_t &e = ctx->join_events[s];
    if (!e) MBX_HIP(hipEventCreateWithFlags(&e, hipEventDisableTiming));
    MBX_HIP(hipEventRecord(e, s));
    MBX_HIP(hipStreamWaitEvent(ctx->stream, e, 0));
  }
  return MBX_OK;
}

int mbx_alloc(mbx_ctx *ctx, size_t bytes, void **dptr) {
  if (!ctx) return fail(MBX_ERR_INVALID, "null ctx");
  return ctx->pool->alloc(bytes, dptr);
}

int mbx_free(mbx_ctx *ctx, void *dptr) {
  if (!ctx) return fail(MBX_ERR_INVALID, "null ctx");
  return ctx->pool->free(dptr);
}

int mbx_pool_stats(mbx_ctx *ctx, size_t *used, size_t *capacity, size_t *peak,
                   size_t *nlive) {
  if (!ctx) return fail(MBX_ERR_INVALID, "null ctx");
  ctx->pool->stats(used, capacity, peak, nlive, nullptr);
  return MBX_OK;
}

int mbx_memcpy_h2d(mbx_ctx *ctx, void *dst, const void *src, size_t bytes) {
  if (!ctx) return fail(MBX_ERR_INVALID, "null ctx");
  if (bytes == 0) return MBX_OK;
  MBX_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, ctx->stream));
  MBX_HIP(hipStreamSynchronize(ctx->stream));
  return MBX_OK;
}

int mbx_memcpy_d2h(mbx_ctx *ctx, void *dst, const void *src, size_t bytes) {
  if (!ctx) return fail(MBX_ERR_INVALID, "null ctx");
  if (bytes == 0) return MBX_OK;
  MBX_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, ctx->stream));
  MBX_HIP(hipStreamSynchronize(ctx->stream));
  return MBX_OK;
}

int mbx_memcpy_d2d(mbx_ctx *ctx, void *dst, const void *src, size_t bytes) {
  if (!ctx) return fail(MBX_ERR_INVALID, "null ctx");
  if (bytes == 0) return MBX_OK;
  MBX_HIP(
      hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToDevice, ctx->stream));
  return MBX_OK;
}

int mbx_fill_u32(mbx_ctx *ctx, void *dst, uint32_t value, size_t count) {
  if (!ctx) return fail(MBX_ERR_INVALID, "null ctx");
  return launch_fill_u32(ctx->stream, dst, value, count);
}

int mbx_timer_create(mbx_ctx *ctx, mbx_timer **t) {
  if (!ctx || !t) return fail(MBX_ERR_INVALID, "null");
  auto *x = new mbx_timer;
  if (hipEventCreate(&x->a) != hipSuccess ||
      hipEventCreate(&x->b) != hipSuccess) {
    delete x;
    return fail(MBX_ERR_HIP, "hipEventCreate failed");
  }
  *t = x;
  return MBX_OK;
}

int mbx_timer_start(mbx_ctx *ctx, mbx_timer *t) {
  if (!ctx || !t) return fail(MBX_ERR_INVALID, "null");
  MBX_HIP(hipEventRecord(t->a, ctx->stream));
  return MBX_OK;
}

int mbx_timer_stop(mbx_ctx *ctx, mbx_timer *t, float *ms) {
  if (!ctx || !t) return fail(MBX_ERR_INVALID, "null");
  MBX_HIP(hipEventRecord(t->b, ctx->stream));
  MBX_HIP(hipEventSynchronize(t->b));
  float v = 0;
  MBX_HIP(hipEventElapsedTime(&v, t->a, t->b));
  if (ms) *ms = v;
  return MBX_OK;
}

int mbx_timer_record_stop(mbx_ctx *ctx, mbx_timer *t) {
  if (!ctx || !t) return fail(MBX_ERR_INVALID, "null");
  MBX_HIP(hipEventRecord(t->b, ctx->stream));
  return MBX_OK;
}

int mbx_timer_elapsed(mbx_timer *t, float *ms) {
  if (!t || !ms) return fail(MBX_ERR_INVALID, "null");
  MBX_HIP(hipEventSynchronize(t->b));
  MBX_HIP(hipEventElapsedTime(ms, t->a, t->b));
  return MBX_OK;
}

int mbx_timer_destroy(mbx_timer *t) {
  if (!t) return MBX_OK;
  (void)hipEventDestroy(t->a);
  (void)hipEventDestroy(t->b);
  delete t;
  return MBX_OK;
}

int mbx_copy_kernel(mbx_ctx *ctx, void *dst, const void *src, size_t bytes) {
  if (!ctx || !dst || !src) return fail(MBX_ERR_INVALID, "null");
  if (bytes % 16 || reinterpret_cast<uintptr_t>(dst) % 16 ||
      reinterpret_cast<uintptr_t>(src) % 16)
    return fail(MBX_ERR_ALIGN, "copy kernel needs 16-byte sizes/pointers");
  const size_t n16 = bytes / 16;
  size_t g = ceil_div(n16, 256 * 4);
  if (g > (size_t)ctx->num_cus * 16) g = (size_t)ctx->num_cus * 16;
  if (g == 0) g = 1;
  hipLaunchKernelGGL(copy_kernel, dim3((unsigned)g), dim3(256), 0, ctx->stream,
                     static_cast<uint4 *>(dst),
                     static_cast<const uint4 *>(src), n16);
  MBX_HIP(hipGetLastError());
  return MBX_OK;
}

}  // extern "C"

extern "C" {
// sizeof of every ABI struct, in declaration order, for binding checks.
int mbx_abi_sizes(size_t *out, int n) {
  const size_t s[] = {sizeof(mbx_colref), sizeof(mbx_col),   sizeof(mbx_instr),
                      sizeof(mbx_vprg),   sizeof(mbx_pred),  sizeof(mbx_keypart),
                      sizeof(mbx_rowop),  sizeof(mbx_query), sizeof(mbx_tune),
                      sizeof(mbx_plan), sizeof(mbx_serve_cfg),
                      sizeof(mbx_serve_stats), sizeof(mbx_wah)};
  const int m = (int)(sizeof(s) / sizeof(s[0]));
  for (int i = 0; i < n && i < m; ++i) out[i] = s[i];
  return m;
}
}
