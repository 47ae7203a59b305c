// plan.cpp -- query plans: evaluation order and fuse-versus-materialise.
//
// A plan rewrites only the instruction list of a conjunctive program
// (ORM r0 <- c_a ; ANDM r0 <- c_b ; ...): AND is commutative, so every order
// selects the same rows and gives the same int64 sums, but the order decides
// how many bitmap lines the ANDM early-out (primitive.cuh:215) skips, and the
// plan picks the fused kernel or the three materialise stages.  The static
// heuristic puts the sparsest column first and always fuses; the learned
// chooser scores every candidate with the cost-model MLP and takes the
// cheapest, falling back to the heuristic when no model is given.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <numeric>
#include <vector>

#include "mbx_internal.hpp"

using namespace mbx;

extern "C" {

// Columns of a pure conjunction, in program order; false for other shapes.
static bool conjunction(const mbx_vprg *p, uint8_t cols[4], int *ncol) {
  int n = 0;
  for (int i = 0; i < MBX_MAXNINSTR; ++i) {
    const mbx_instr &in = p->instrs[i];
    if (in.opcode == MBX_END) break;
    if (in.dstreg != 0 || in.src1 >= MBX_MAXCOLS) return false;
    if (i == 0 ? in.opcode != MBX_ORM : in.opcode != MBX_ANDM) return false;
    for (int k = 0; k < n; ++k)
      if (cols[k] == in.src1) return false;
    cols[n++] = in.src1;
  }
  *ncol = n;
  return n >= 1;
}

int mbx_plan_describe(const mbx_vprg *p, mbx_plan *out) {
  if (!p || !out) return fail(MBX_ERR_INVALID, "plan_describe: null");
  std::memset(out, 0, sizeof(*out));
  uint8_t cols[4];
  int n = 0;
  if (conjunction(p, cols, &n)) {
    out->reorderable = 1;
    out->norder = (uint8_t)n;
    std::memcpy(out->order, cols, 4);
  }
  out->flags = MBX_PLAN_FUSED;
  return MBX_OK;
}

int mbx_plan_apply(const mbx_vprg *in, const mbx_plan *plan, mbx_vprg *out) {
  if (!in || !plan || !out) return fail(MBX_ERR_INVALID, "plan_apply: null");
  *out = *in;
  if (!plan->reorderable) return MBX_OK;
  uint8_t cols[4];
  int n = 0;
  if (!conjunction(in, cols, &n) || n != plan->norder)
    return fail(MBX_ERR_INVALID, "plan_apply: plan does not fit the program");
  bool seen[4] = {false, false, false, false};
  for (int k = 0; k < n; ++k) {
    const uint8_t c = plan->order[k];
    bool member = false;
    for (int j = 0; j < n; ++j) member = member || cols[j] == c;
    if (!member || seen[c])
      return fail(MBX_ERR_INVALID, "plan_apply: order is not a permutation");
    seen[c] = true;
    out->instrs[k].opcode = k == 0 ? MBX_ORM : MBX_ANDM;
    out->instrs[k].dstreg = 0;
    out->instrs[k].src1 = c;
  }
  return MBX_OK;
}

// The 16 cost-model features of (program, densities, plan).  dens[c] is the
// fraction of rows in column c's operand (OR of its bins).
int mbx_plan_features(const mbx_vprg *p, const double *dens, uint32_t nr_grp,
                      const mbx_plan *plan, float *f) {
  if (!p || !dens || !plan || !f) return fail(MBX_ERR_INVALID, "features: null");
  std::memset(f, 0, sizeof(float) * MBX_NFEAT);
  uint8_t order[4];
  int n = 0;
  if (plan->reorderable) {
    n = plan->norder;
    std::memcpy(order, plan->order, 4);
  } else {
    for (int c = 0; c < MBX_MAXCOLS; ++c)
      if (p->cols[c].middle[0] || p->cols[c].leftmost || p->cols[c].rightmost)
        order[n++] = (uint8_t)c;
  }
  double run = 1.0, lines = 0.0;
  int nbins = 0, chk = 0, nins = 0;
  for (int k = 0; k < n && k < 4; ++k) {
    const double d = std::min(1.0, std::max(0.0, dens[order[k]]));
    // share of 1024-row bitmap lines still live when this operand is read
    lines += k == 0 ? 1.0 : 1.0 - std::pow(1.0 - run, 1024.0);
    run *= d;
    f[k] = (float)d;
    f[4 + k] = (float)run;
  }
  for (int c = 0; c < MBX_MAXCOLS; ++c) {
    for (int k = 0; k < MBX_MAXBIN_PERCOL; ++k) nbins += p->cols[c].middle[k] != 0;
    nbins += (p->cols[c].leftmost != 0) + (p->cols[c].rightmost != 0);
    chk |= (p->cols[c].leftmost || p->cols[c].rightmost);
  }
  while (nins < MBX_MAXNINSTR && p->instrs[nins].opcode != MBX_END) ++nins;
  f[8] = nins / 6.f + 0.5f * chk;
  f[9] = (float)(std::log2((double)nr_grp + 1.0) / 20.0);
  f[10] = (float)(std::log2((double)p->factsz + 1.0) / 32.0);
  f[11] = (plan->flags & MBX_PLAN_MASK) == MBX_PLAN_MATERIALISE ? 1.f : 0.f;
  f[12] = (float)(std::log10(std::max(run, 1e-9)) / 9.0 + 1.0);
  f[13] = nbins / 24.f;
  f[14] = (plan->flags & MBX_PLAN_EAGER) ? 1.f : 0.f;
  f[15] = n ? (float)(lines / n) : 0.f;
  return MBX_OK;
}

// Every order of a conjunction x {fused, fused eager, materialise}; other
// programs get their own order only.
int mbx_plan_enumerate(const mbx_vprg *p, mbx_plan *out, int max, int *count) {
  if (!p || !out || !count) return fail(MBX_ERR_INVALID, "enumerate: null");
  mbx_plan base;
  int st = mbx_plan_describe(p, &base);
  if (st) return st;
  std::vector<mbx_plan> all;
  std::vector<uint8_t> perm(base.order, base.order + base.norder);
  std::sort(perm.begin(), perm.end());
  do {
    for (uint32_t fl : {(uint32_t)MBX_PLAN_FUSED,
                        (uint32_t)(MBX_PLAN_FUSED | MBX_PLAN_EAGER),
                        (uint32_t)MBX_PLAN_MATERIALISE}) {
      mbx_plan q = base;
      std::copy(perm.begin(), perm.end(), q.order);
      q.flags = fl;
      all.push_back(q);
    }
  } while (base.reorderable && std::next_permutation(perm.begin(), perm.end()));
  *count = (int)all.size();
  if ((int)all.size() > max) return fail(MBX_ERR_CAPACITY, "enumerate: max too small");
  std::copy(all.begin(), all.end(), out);
  return MBX_OK;
}

int mbx_plan_heuristic(const mbx_vprg *p, const double *dens, mbx_plan *out) {
  if (!p || !dens || !out) return fail(MBX_ERR_INVALID, "heuristic: null");
  int st = mbx_plan_describe(p, out);
  if (st) return st;
  if (out->reorderable)
    std::stable_sort(out->order, out->order + out->norder,
                     [&](uint8_t a, uint8_t b) { return dens[a] < dens[b]; });
  out->flags = MBX_PLAN_FUSED;
  // Load every operand up front unless the ANDM early-out would skip most
  // lines: a 1024-row line of the next operand is skipped only when all its
  // rows are already empty, probability (1-d)^1024 for a first operand of
  // density d; below one half the early-out saves too little to pay for a
  // round trip per instruction (eager conjunctions are also evaluated
  // without program interpretation).
  const double d0 = out->reorderable ? dens[out->order[0]] : 1.0;
  if (std::pow(1.0 - std::min(1.0, std::max(0.0, d0)), 1024.0) < 0.5)
    out->flags |= MBX_PLAN_EAGER;
  return MBX_OK;
}

// Learned chooser: scores all candidates with the MLP (output 0 = predicted
// log runtime) on the device and returns the cheapest.  mlp == NULL falls
// back to the static heuristic.
int mbx_plan_choose(mbx_ctx *ctx, mbx_mlp *mlp, const mbx_vprg *p,
                    const double *dens, uint32_t nr_grp, mbx_plan *best,
                    float *pred) {
  if (!ctx || !p || !dens || !best) return fail(MBX_ERR_INVALID, "choose: null");
  if (!mlp) return mbx_plan_heuristic(p, dens, best);
  mbx_plan cand[80];
  int n = 0;
  int st = mbx_plan_enumerate(p, cand, 80, &n);
  if (st) return st;
  std::vector<float> feat((size_t)n * MBX_NFEAT);
  for (int i = 0; i < n; ++i) {
    st = mbx_plan_features(p, dens, nr_grp, &cand[i], &feat[(size_t)i * MBX_NFEAT]);
    if (st) return st;
  }
  void *buf = nullptr;
  st = ctx->pool->alloc(feat.size() * 4 + (size_t)n * 4 * 8 + 256, &buf);
  if (st) return st;
  float *dx = static_cast<float *>(buf);
  float *dy = dx + ((feat.size() + 63) / 64) * 64;
  std::vector<float> y((size_t)n * 8);
  hipError_t e = hipMemcpyAsync(dx, feat.data(), feat.size() * 4,
                                hipMemcpyHostToDevice, ctx->stream);
  if (e == hipSuccess) {
    st = mbx_mlp_forward(ctx, mlp, dx, n, dy);
    if (!st)
      e = hipMemcpyAsync(y.data(), dy, (size_t)n * 4, hipMemcpyDeviceToHost,
                         ctx->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
  }
  ctx->pool->free(buf);
  if (st) return st;
  if (e != hipSuccess) return hip_fail(e, "choose");
  int bi = 0;
  for (int i = 1; i < n; ++i)
    if (y[i] < y[bi]) bi = i;
  *best = cand[bi];
  if (pred) *pred = y[bi];
  return MBX_OK;
}

// Runs a query under a plan (same results for every plan).
int mbx_prog_grpby_plan(mbx_ctx *ctx, const mbx_vprg *prg, const mbx_rowop *op,
                        uint32_t nr_grp, int64_t *d_out, const mbx_plan *plan) {
  if (!plan) return mbx_prog_grpby(ctx, prg, op, nr_grp, d_out, MBX_PLAN_FUSED);
  mbx_vprg q;
  int st = mbx_plan_apply(prg, plan, &q);
  if (st) return st;
  return mbx_prog_grpby(ctx, &q, op, nr_grp, d_out,
                        plan->flags &
                            (MBX_PLAN_MASK | MBX_PLAN_EAGER | MBX_ACCUMULATE));
}

}  // extern "C"
