// mlp.hip -- the plan chooser's cost model: a 2-layer fp32 MLP
// (din -> dh ReLU -> dout) with hand-written forward, backward and Adam
// kernels for gfx950.
//
//   y = W2 relu(W1 x + b1) + b2,  loss = mean over n*dout of (y - t)^2
//
// The model is tiny (16x64 + 64xk weights), so every kernel is a plain
// one-thread-per-output loop; weight gradients are reduced deterministically
// (per-chunk partial sums, then a fixed-order sum over chunks), so training
// is reproducible run to run.
#include <cmath>
#include <cstring>
#include <vector>

#include "mbx_internal.hpp"

namespace mbx {

constexpr int kChunk = 256;   // samples per partial-sum chunk

// pre[i,h] = b1[h] + sum_k W1[h,k] x[i,k]; y[i,o] = b2[o] + sum_h W2[o,h] relu
__global__ void mlp_fwd_kernel(const float *__restrict__ x, int n, int din,
                               int dh, int dout, const float *__restrict__ W1,
                               const float *__restrict__ b1,
                               const float *__restrict__ W2,
                               const float *__restrict__ b2,
                               float *__restrict__ pre, float *__restrict__ y) {
  extern __shared__ float s_w[];   // W1 | b1 | W2 | b2
  float *sW1 = s_w, *sb1 = sW1 + dh * din, *sW2 = sb1 + dh,
        *sb2 = sW2 + dout * dh;
  const int nw = dh * din + dh + dout * dh + dout;
  for (int i = threadIdx.x; i < nw; i += blockDim.x) {
    float v;
    if (i < dh * din) v = W1[i];
    else if (i < dh * din + dh) v = b1[i - dh * din];
    else if (i < dh * din + dh + dout * dh) v = W2[i - dh * din - dh];
    else v = b2[i - dh * din - dh - dout * dh];
    s_w[i] = v;
  }
  __syncthreads();
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  float acc[8];
  for (int o = 0; o < dout; ++o) acc[o] = sb2[o];
  for (int h = 0; h < dh; ++h) {
    float s = sb1[h];
    for (int k = 0; k < din; ++k) s = fmaf(sW1[h * din + k], x[i * din + k], s);
    if (pre) pre[i * dh + h] = s;
    const float a = s > 0.f ? s : 0.f;
    for (int o = 0; o < dout; ++o) acc[o] = fmaf(sW2[o * dh + h], a, acc[o]);
  }
  for (int o = 0; o < dout; ++o) y[i * dout + o] = acc[o];
}

// dy[i,o] = 2 (y - t) / (n*dout); per-sample squared error for the loss
__global__ void mlp_dy_kernel(const float *__restrict__ y,
                              const float *__restrict__ t, int m, float scale,
                              float *__restrict__ dy, float *__restrict__ sq) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= m) return;
  const float e = y[i] - t[i];
  dy[i] = 2.f * e * scale;
  sq[i] = e * e;
}

// dpre[i,h] = (sum_o dy[i,o] W2[o,h]) * [pre > 0]
__global__ void mlp_dpre_kernel(const float *__restrict__ dy,
                                const float *__restrict__ pre,
                                const float *__restrict__ W2, int n, int dh,
                                int dout, float *__restrict__ dpre) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= n * dh) return;
  const int i = j / dh, h = j % dh;
  float g = 0.f;
  if (pre[j] > 0.f)
    for (int o = 0; o < dout; ++o) g = fmaf(dy[i * dout + o], W2[o * dh + h], g);
  dpre[j] = g;
}

// Partial sums of one chunk of samples for every parameter:
//   gW1[h,k] = sum_i dpre[i,h] x[i,k]   gb1[h] = sum_i dpre[i,h]
//   gW2[o,h] = sum_i dy[i,o] relu(pre[i,h])   gb2[o] = sum_i dy[i,o]
// Sums run in double, in sample order.
__global__ void mlp_grad_partial_kernel(
    const float *__restrict__ x, const float *__restrict__ pre,
    const float *__restrict__ dpre, const float *__restrict__ dy, int n,
    int din, int dh, int dout, double *__restrict__ part) {
  const int np = dh * din + dh + dout * dh + dout;
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= np) return;
  const int c = blockIdx.y;
  const int i0 = c * kChunk, i1 = min(n, i0 + kChunk);
  double s = 0.0;
  if (p < dh * din) {
    const int h = p / din, k = p % din;
    for (int i = i0; i < i1; ++i) s += (double)dpre[i * dh + h] * x[i * din + k];
  } else if (p < dh * din + dh) {
    const int h = p - dh * din;
    for (int i = i0; i < i1; ++i) s += dpre[i * dh + h];
  } else if (p < dh * din + dh + dout * dh) {
    const int q = p - dh * din - dh, o = q / dh, h = q % dh;
    for (int i = i0; i < i1; ++i) {
      const float a = pre[i * dh + h];
      s += (double)dy[i * dout + o] * (a > 0.f ? a : 0.f);
    }
  } else {
    const int o = p - dh * din - dh - dout * dh;
    for (int i = i0; i < i1; ++i) s += dy[i * dout + o];
  }
  part[(size_t)c * np + p] = s;
}

__global__ void mlp_grad_reduce_kernel(const double *__restrict__ part,
                                       int nchunk, int np,
                                       float *__restrict__ g) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= np) return;
  double s = 0.0;
  for (int c = 0; c < nchunk; ++c) s += part[(size_t)c * np + p];
  g[p] = (float)s;
}

// Adam (Kingma & Ba 2015), bias-corrected; params and moments are flat.
__global__ void mlp_adam_kernel(float *__restrict__ w,
                                const float *__restrict__ g,
                                float *__restrict__ m, float *__restrict__ v,
                                int np, float lr, float b1, float b2,
                                float eps, float c1, float c2) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= np) return;
  const float gp = g[p];
  const float mp = b1 * m[p] + (1.f - b1) * gp;
  const float vp = b2 * v[p] + (1.f - b2) * gp * gp;
  m[p] = mp;
  v[p] = vp;
  w[p] -= lr * (mp / c1) / (sqrtf(vp / c2) + eps);
}

__global__ void sum_kernel(const float *__restrict__ a, int n,
                           double *__restrict__ out) {
  // one block: fixed-order strided sums then a tree in LDS
  __shared__ double s[256];
  double acc = 0.0;
  for (int i = threadIdx.x; i < n; i += 256) acc += a[i];
  s[threadIdx.x] = acc;
  __syncthreads();
  for (int k = 128; k > 0; k >>= 1) {
    if ((int)threadIdx.x < k) s[threadIdx.x] += s[threadIdx.x + k];
    __syncthreads();
  }
  if (threadIdx.x == 0) *out = s[0];
}

}  // namespace mbx

using namespace mbx;

// Parameters live in one flat device array in the order W1 | b1 | W2 | b2,
// followed by the gradient and the two Adam moments of the same layout.
struct mbx_mlp {
  int din, dh, dout, np;
  float *d_w = nullptr, *d_g, *d_m, *d_v;
  long step = 0;
  // scratch sized for the largest batch seen
  int cap = 0;
  float *d_pre = nullptr, *d_dpre, *d_y, *d_dy, *d_sq;
  double *d_part = nullptr, *d_sum;
};

static int ensure_scratch(mbx_ctx *ctx, mbx_mlp *m, int n) {
  if (n <= m->cap) return MBX_OK;
  if (m->d_pre) ctx->pool->free(m->d_pre);
  if (m->d_part) ctx->pool->free(m->d_part);
  m->d_pre = nullptr;
  m->d_part = nullptr;
  const size_t f = (size_t)n * (2 * m->dh + 3 * m->dout);
  void *p = nullptr;
  int st = ctx->pool->alloc(f * 4 + 1024, &p);
  if (st) return st;
  m->d_pre = static_cast<float *>(p);
  m->d_dpre = m->d_pre + (size_t)n * m->dh;
  m->d_y = m->d_dpre + (size_t)n * m->dh;
  m->d_dy = m->d_y + (size_t)n * m->dout;
  m->d_sq = m->d_dy + (size_t)n * m->dout;
  const int nchunk = (n + kChunk - 1) / kChunk;
  st = ctx->pool->alloc(((size_t)nchunk * m->np + 16) * 8, &p);
  if (st) return st;
  m->d_part = static_cast<double *>(p);
  m->d_sum = m->d_part + (size_t)nchunk * m->np;
  m->cap = n;
  return MBX_OK;
}

static size_t fwd_lds(const mbx_mlp *m) { return (size_t)m->np * 4; }

static int forward(mbx_ctx *ctx, mbx_mlp *m, const float *x, int n, float *pre,
                   float *y) {
  const float *W1 = m->d_w, *b1 = W1 + m->dh * m->din, *W2 = b1 + m->dh,
              *b2 = W2 + m->dout * m->dh;
  hipLaunchKernelGGL(mlp_fwd_kernel, dim3((n + 255) / 256), dim3(256),
                     fwd_lds(m), ctx->stream, x, n, m->din, m->dh, m->dout, W1,
                     b1, W2, b2, pre, y);
  MBX_HIP(hipGetLastError());
  return MBX_OK;
}

// loss and gradient (into m->d_g) of the batch; *loss on the host.
static int loss_grad(mbx_ctx *ctx, mbx_mlp *m, const float *x, const float *t,
                     int n, double *loss) {
  int st = ensure_scratch(ctx, m, n);
  if (st) return st;
  st = forward(ctx, m, x, n, m->d_pre, m->d_y);
  if (st) return st;
  const int ny = n * m->dout;
  const float scale = 1.f / (float)ny;
  hipLaunchKernelGGL(mlp_dy_kernel, dim3((ny + 255) / 256), dim3(256), 0,
                     ctx->stream, m->d_y, t, ny, scale, m->d_dy, m->d_sq);
  const float *W2 = m->d_w + m->dh * m->din + m->dh;
  hipLaunchKernelGGL(mlp_dpre_kernel, dim3((n * m->dh + 255) / 256), dim3(256),
                     0, ctx->stream, m->d_dy, m->d_pre, W2, n, m->dh, m->dout,
                     m->d_dpre);
  const int nchunk = (n + kChunk - 1) / kChunk;
  hipLaunchKernelGGL(mlp_grad_partial_kernel,
                     dim3((m->np + 255) / 256, nchunk), dim3(256), 0,
                     ctx->stream, x, m->d_pre, m->d_dpre, m->d_dy, n, m->din,
                     m->dh, m->dout, m->d_part);
  hipLaunchKernelGGL(mlp_grad_reduce_kernel, dim3((m->np + 255) / 256),
                     dim3(256), 0, ctx->stream, m->d_part, nchunk, m->np,
                     m->d_g);
  MBX_HIP(hipGetLastError());
  if (loss) {
    hipLaunchKernelGGL(sum_kernel, dim3(1), dim3(256), 0, ctx->stream,
                       m->d_sq, ny, m->d_sum);
    double s = 0;
    MBX_HIP(hipMemcpyAsync(&s, m->d_sum, 8, hipMemcpyDeviceToHost,
                           ctx->stream));
    MBX_HIP(hipStreamSynchronize(ctx->stream));
    *loss = s / ny;
  }
  return MBX_OK;
}

extern "C" {

int mbx_mlp_create(mbx_ctx *ctx, int din, int dh, int dout, uint64_t seed,
                   mbx_mlp **out) {
  if (!ctx || !out || din <= 0 || dh <= 0 || dout <= 0 || dout > 8 ||
      din > 64 || dh > 256)
    return fail(MBX_ERR_INVALID, "mlp_create: need din<=64, dh<=256, dout<=8");
  auto *m = new mbx_mlp;
  m->din = din;
  m->dh = dh;
  m->dout = dout;
  m->np = dh * din + dh + dout * dh + dout;
  void *p = nullptr;
  int st = ctx->pool->alloc((size_t)m->np * 4 * 4 + 1024, &p);
  if (st) {
    delete m;
    return st;
  }
  m->d_w = static_cast<float *>(p);
  m->d_g = m->d_w + m->np;
  m->d_m = m->d_g + m->np;
  m->d_v = m->d_m + m->np;
  // He-style init from the library's own generator: uniform in
  // [-sqrt(6/fan_in), sqrt(6/fan_in)], biases 0
  std::vector<float> h((size_t)m->np * 4, 0.f);
  uint64_t s = seed;
  auto uni = [&](float a) {
    s = s * 6364136223846793005ull + 1442695040888963407ull;
    const double u = (double)(s >> 11) * (1.0 / 9007199254740992.0);
    return (float)((2.0 * u - 1.0) * a);
  };
  const float a1 = std::sqrt(6.f / din), a2 = std::sqrt(6.f / dh);
  for (int i = 0; i < dh * din; ++i) h[i] = uni(a1);
  for (int i = 0; i < dout * dh; ++i) h[dh * din + dh + i] = uni(a2);
  hipError_t e = hipMemcpyAsync(m->d_w, h.data(), h.size() * 4,
                                hipMemcpyHostToDevice, ctx->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
  if (e != hipSuccess) {
    ctx->pool->free(p);
    delete m;
    return hip_fail(e, "mlp_create upload");
  }
  *out = m;
  return MBX_OK;
}

int mbx_mlp_destroy(mbx_ctx *ctx, mbx_mlp *m) {
  if (!m) return MBX_OK;
  if (ctx) {
    (void)hipStreamSynchronize(ctx->stream);
    ctx->pool->free(m->d_w);
    if (m->d_pre) ctx->pool->free(m->d_pre);
    if (m->d_part) ctx->pool->free(m->d_part);
  }
  delete m;
  return MBX_OK;
}

// host <-> device parameters, flat W1|b1|W2|b2
int mbx_mlp_set_params(mbx_ctx *ctx, mbx_mlp *m, const float *w) {
  if (!ctx || !m || !w) return fail(MBX_ERR_INVALID, "mlp_set_params: null");
  MBX_HIP(hipMemcpyAsync(m->d_w, w, (size_t)m->np * 4, hipMemcpyHostToDevice,
                         ctx->stream));
  MBX_HIP(hipStreamSynchronize(ctx->stream));
  return launch_fill_u32(ctx->stream, m->d_m, 0, 2 * (size_t)m->np);
}

int mbx_mlp_get_params(mbx_ctx *ctx, mbx_mlp *m, float *w) {
  if (!ctx || !m || !w) return fail(MBX_ERR_INVALID, "mlp_get_params: null");
  MBX_HIP(hipMemcpyAsync(w, m->d_w, (size_t)m->np * 4, hipMemcpyDeviceToHost,
                         ctx->stream));
  MBX_HIP(hipStreamSynchronize(ctx->stream));
  return MBX_OK;
}

int mbx_mlp_nparams(const mbx_mlp *m) { return m ? m->np : 0; }

// y[n, dout] = model(x[n, din]); device pointers.
int mbx_mlp_forward(mbx_ctx *ctx, mbx_mlp *m, const float *d_x, int n,
                    float *d_y) {
  if (!ctx || !m || !d_x || !d_y || n <= 0)
    return fail(MBX_ERR_INVALID, "mlp_forward: bad argument");
  return forward(ctx, m, d_x, n, nullptr, d_y);
}

// loss and flat gradient (host array of nparams floats) of one batch.
int mbx_mlp_loss_grad(mbx_ctx *ctx, mbx_mlp *m, const float *d_x,
                      const float *d_t, int n, double *loss, float *grad) {
  if (!ctx || !m || !d_x || !d_t || n <= 0)
    return fail(MBX_ERR_INVALID, "mlp_loss_grad: bad argument");
  int st = loss_grad(ctx, m, d_x, d_t, n, loss);
  if (st) return st;
  if (grad) {
    MBX_HIP(hipMemcpyAsync(grad, m->d_g, (size_t)m->np * 4,
                           hipMemcpyDeviceToHost, ctx->stream));
    MBX_HIP(hipStreamSynchronize(ctx->stream));
  }
  return MBX_OK;
}

// Full-batch Adam for `steps` steps; *loss gets the final batch loss.
int mbx_mlp_train(mbx_ctx *ctx, mbx_mlp *m, const float *d_x, const float *d_t,
                  int n, int steps, float lr, double *loss) {
  if (!ctx || !m || !d_x || !d_t || n <= 0 || steps < 0)
    return fail(MBX_ERR_INVALID, "mlp_train: bad argument");
  const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
  for (int s = 0; s < steps; ++s) {
    int st = loss_grad(ctx, m, d_x, d_t, n, nullptr);
    if (st) return st;
    ++m->step;
    const float c1 = 1.f - std::pow(b1, (float)m->step);
    const float c2 = 1.f - std::pow(b2, (float)m->step);
    hipLaunchKernelGGL(mlp_adam_kernel, dim3((m->np + 255) / 256), dim3(256),
                       0, ctx->stream, m->d_w, m->d_g, m->d_m, m->d_v, m->np,
                       lr, b1, b2, eps, c1, c2);
    MBX_HIP(hipGetLastError());
  }
  if (loss) return loss_grad(ctx, m, d_x, d_t, n, loss);
  MBX_HIP(hipStreamSynchronize(ctx->stream));
  return MBX_OK;
}

}  // extern "C"
