// ctx.hip -- context, memory, fill/copy kernels, timers, descriptor checks.
#include <cstring>

#include "mbx_internal.hpp"

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

int check_src(const mbx_colref &c, const char *what) {
  if (c.data == nullptr)
    return fail(MBX_ERR_INVALID, std::string(what) + ": null column");
  if (c.elem_bytes != 1 && c.elem_bytes != 2 && c.elem_bytes != 4 &&
      c.elem_bytes != 8)
    return fail(MBX_ERR_INVALID,
                std::string(what) + ": elem_bytes must be 1, 2, 4 or 8");
  if ((reinterpret_cast<uintptr_t>(c.data) % c.elem_bytes) != 0 ||
      (reinterpret_cast<uintptr_t>(c.fk) % 4) != 0)
    return fail(MBX_ERR_ALIGN, std::string(what) + ": misaligned column");
  return MBX_OK;
}

static DevSrc to_dev(const mbx_colref &c) {
  DevSrc d;
  d.data = c.data;
  d.fk = c.fk;
  d.dimsz = c.dimsz;
  d.eb = c.elem_bytes;
  return d;
}

static int check_bmp(const uint32_t *p) {
  if (reinterpret_cast<uintptr_t>(p) % 8 != 0)
    return fail(MBX_ERR_ALIGN, "bitmap pointer is not 8-byte aligned");
  return MBX_OK;
}

// vprg validation follows the reference's limits (primitive.cuh:121,164-198):
// registers r0..r2, column operands 0..3, at most 6 instructions.
int make_dev_prg(const mbx_vprg *p, DevPrg *out) {
  if (!p || !out) return fail(MBX_ERR_INVALID, "vprg: null");
  if (p->factsz == 0) return fail(MBX_ERR_INVALID, "vprg: factsz == 0");
  if (p->factsz > 0x7fffffffu)
    return fail(MBX_ERR_CAPACITY, "vprg: factsz must be < 2^31");
  std::memset(out, 0, sizeof(*out));
  bool chk = false;
  for (int c = 0; c < MBX_MAXCOLS; ++c) {
    const mbx_col &col = p->cols[c];
    DevCol &d = out->cols[c];
    uint32_t nm = 0;
    while (nm < MBX_MAXBIN_PERCOL && col.middle[nm]) {
      int st = check_bmp(col.middle[nm]);
      if (st) return st;
      d.m[nm] = reinterpret_cast<const uint64_t *>(col.middle[nm]);
      ++nm;
    }
    d.nm = nm;
    if (col.leftmost) {
      int st = check_bmp(col.leftmost);
      if (st) return st;
    }
    if (col.rightmost) {
      int st = check_bmp(col.rightmost);
      if (st) return st;
    }
    d.l = reinterpret_cast<const uint64_t *>(col.leftmost);
    d.r = reinterpret_cast<const uint64_t *>(col.rightmost);
    d.has_edge = (col.leftmost || col.rightmost) ? 1u : 0u;
    chk = chk || d.has_edge;
  }
  uint32_t n = 0;
  for (; n < MBX_MAXNINSTR; ++n) {
    const mbx_instr &in = p->instrs[n];
    if (in.opcode == MBX_END) break;
    if (in.opcode > MBX_ORM)
      return fail(MBX_ERR_INVALID, "vprg: bad opcode " +
                                       std::to_string(in.opcode));
    if (in.dstreg > 2) return fail(MBX_ERR_INVALID, "vprg: dstreg > 2");
    const bool mem = in.opcode == MBX_ANDM || in.opcode == MBX_ORM;
    if (mem ? in.src1 >= MBX_MAXCOLS : in.src1 > 2)
      return fail(MBX_ERR_INVALID, "vprg: src1 out of range");
    out->ins[n][0] = in.opcode;
    out->ins[n][1] = in.dstreg;
    out->ins[n][2] = in.src1;
    out->ins[n][3] = in.src2;
  }
  out->nins = n;
  out->nrows = p->factsz;
  out->nwords = (uint32_t)ceil_div(p->factsz, 64);
  out->chk = chk ? 1u : 0u;
  return MBX_OK;
}

int make_dev_op(const mbx_rowop *op, uint32_t npreds_hint, DevOp *out) {
  if (!op || !out) return fail(MBX_ERR_INVALID, "rowop: null");
  std::memset(out, 0, sizeof(*out));
  (void)npreds_hint;
  if (op->npreds > MBX_MAXCOLS || op->nkeys > MBX_MAXKEYS)
    return fail(MBX_ERR_INVALID, "rowop: too many predicates or keys");
  out->npreds = op->npreds;
  for (uint32_t i = 0; i < op->npreds; ++i) {
    int st = check_src(op->preds[i].src, "rowop predicate");
    if (st) return st;
    out->pred[i] = to_dev(op->preds[i].src);
    out->lo[i] = op->preds[i].lo;
    out->hi[i] = op->preds[i].hi;
  }
  out->nkeys = op->nkeys;
  for (uint32_t i = 0; i < op->nkeys; ++i) {
    const mbx_keypart &k = op->keys[i];
    int st = check_src(k.src, "rowop key");
    if (st) return st;
    if (k.kind > MBX_KEY_BUCKET)
      return fail(MBX_ERR_INVALID, "rowop key: bad kind");
    if (k.kind == MBX_KEY_DIV && k.div == 0)
      return fail(MBX_ERR_INVALID, "rowop key: div == 0");
    if (k.nbounds > MBX_MAXBOUNDS)
      return fail(MBX_ERR_INVALID, "rowop key: too many bounds");
    DevKey &d = out->keys[i];
    d.src = to_dev(k.src);
    d.kind = k.kind;
    d.div = k.div;
    d.base = k.base;
    d.stride = k.stride;
    d.nb = k.nbounds;
    for (uint32_t j = 0; j < MBX_MAXBOUNDS; ++j) d.bounds[j] = k.bounds[j];
  }
  if (op->measure > MBX_MEAS_SUB)
    return fail(MBX_ERR_INVALID, "rowop: bad measure");
  out->measure = op->measure;
  if (op->measure != MBX_MEAS_COUNT) {
    int st = check_src(op->ma, "rowop measure a");
    if (st) return st;
    out->ma = to_dev(op->ma);
  }
  if (op->measure == MBX_MEAS_MUL || op->measure == MBX_MEAS_SUB) {
    int st = check_src(op->mb, "rowop measure b");
    if (st) return st;
    out->mb = to_dev(op->mb);
  }
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
  *out = c;
  return MBX_OK;
}

int mbx_destroy(mbx_ctx *ctx) {
  if (!ctx) return MBX_OK;
  (void)hipSetDevice(ctx->device);
  (void)hipStreamSynchronize(ctx->stream);
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
                      sizeof(mbx_rowop),  sizeof(mbx_query), sizeof(mbx_tune)};
  const int m = (int)(sizeof(s) / sizeof(s[0]));
  for (int i = 0; i < n && i < m; ++i) out[i] = s[i];
  return m;
}
}
