// plan.cpp -- query plans: evaluation order and fuse-versus-materialise.
//
// A plan rewrites only the instruction list of a conjunctive program
// (ORM r0 <- c_a ; ANDM r0 <- c_b ; ...): AND is commutative, so every order
// selects the same rows and gives the same int64 sums, but the order decides
// how many bitmap lines the ANDM early-out (primitive.cuh:215) skips, and the
// plan picks the fused kernel (early-out or eager loads), the three
// materialise stages, or the bitmap-free scan (grpby, primitive.cuh:16-48),
// which wins when edge bins leave most rows to re-check.  The static
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

// The 16 cost-model features of (program, column statistics, plan).
// dens[c] is the fraction of rows in column c's operand (OR of its bins),
// edge[c] (NULL = none known) the fraction only in its edge bins.
//   0-3   densities in plan order        4-7  running products (the share
//                                              of rows still live)
//   8     bins read / 24 + 0.5 if the program has edge bins
//   9     log2(groups + 1) / 20          10   log2(rows + 1) / 32
//   11    plan is materialise            12   log10(selectivity) / 9 + 1
//   13    log10(uncertain share) / 9 + 1: rows an edge bin leaves to
//         re-check, sum over columns of edge[c] x selectivity / dens[c]
//   14    plan loads eagerly             15   plan is the scan
// The selectivity is the program run over the densities as independent
// probabilities (AND: product, OR: a + b - ab, NOT: 1 - a), so OR and
// ANDNOT programs are priced by what they select.
int mbx_plan_features_ex(const mbx_vprg *p, const double *dens,
                         const double *edge, uint32_t nr_grp,
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
  auto clamp01 = [](double x) { return std::min(1.0, std::max(0.0, x)); };
  double run = 1.0, unc = 0.0;   // run: product in plan order
  int nbins = 0, chk = 0;
  for (int k = 0; k < n && k < 4; ++k) {
    const double d = clamp01(dens[order[k]]);
    run *= d;
    f[k] = (float)d;
    f[4 + k] = (float)run;
  }
  // the program over the densities: registers hold probabilities
  double r[3] = {0, 0, 0};
  for (int i = 0; i < MBX_MAXNINSTR && p->instrs[i].opcode != MBX_END; ++i) {
    const mbx_instr &in = p->instrs[i];
    const int d = in.dstreg % 3, sr = in.src1 % 3;
    const double m = in.src1 < MBX_MAXCOLS ? clamp01(dens[in.src1]) : 0.0;
    switch (in.opcode) {
    case MBX_AND: r[d] *= r[sr]; break;
    case MBX_OR: r[d] = r[d] + r[sr] - r[d] * r[sr]; break;
    case MBX_NOT: r[d] = 1.0 - r[sr]; break;
    case MBX_ANDM: r[d] *= m; break;
    case MBX_ORM: r[d] = r[d] + m - r[d] * m; break;
    default: break;
    }
  }
  const double sel = clamp01(r[0]);
  if (edge)
    for (int c = 0; c < MBX_MAXCOLS; ++c) {
      const double e = clamp01(edge[c]), d = clamp01(dens[c]);
      if (e > 0 && d > 0) unc += e * std::min(1.0, sel / d);
    }
  for (int c = 0; c < MBX_MAXCOLS; ++c) {
    for (int k = 0; k < MBX_MAXBIN_PERCOL; ++k) nbins += p->cols[c].middle[k] != 0;
    nbins += (p->cols[c].leftmost != 0) + (p->cols[c].rightmost != 0);
    chk |= (p->cols[c].leftmost || p->cols[c].rightmost);
  }
  const uint32_t kind = plan->flags & MBX_PLAN_MASK;
  f[8] = nbins / 24.f + 0.5f * chk;
  f[9] = (float)(std::log2((double)nr_grp + 1.0) / 20.0);
  f[10] = (float)(std::log2((double)p->factsz + 1.0) / 32.0);
  f[11] = kind == MBX_PLAN_MATERIALISE ? 1.f : 0.f;
  f[12] = (float)(std::log10(std::max(sel, 1e-9)) / 9.0 + 1.0);
  f[13] = (float)(std::log10(std::max(std::min(unc, 1.0), 1e-9)) / 9.0 + 1.0);
  f[14] = (plan->flags & MBX_PLAN_EAGER) ? 1.f : 0.f;
  f[15] = kind == MBX_PLAN_SCAN ? 1.f : 0.f;
  return MBX_OK;
}

int mbx_plan_features(const mbx_vprg *p, const double *dens, uint32_t nr_grp,
                      const mbx_plan *plan, float *f) {
  return mbx_plan_features_ex(p, dens, nullptr, nr_grp, plan, f);
}

// Every order of a conjunction x {fused, fused eager, materialise}, plus
// the scan (order-free); other programs get their own order only.
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
  mbx_plan sc = base;
  sc.flags = MBX_PLAN_SCAN;
  all.push_back(sc);
  *count = (int)all.size();
  if ((int)all.size() > max) return fail(MBX_ERR_CAPACITY, "enumerate: max too small");
  std::copy(all.begin(), all.end(), out);
  return MBX_OK;
}

// The candidates the chooser scores (and the engine times to train it).
// Only the ANDM early-out depends on the order, and mostly through its
// first two columns, so: the early-out plan for every ordered pair of
// leading columns (the rest by ascending density), and one plan each --
// sparsest column first -- for eager loads (the conjunction form does not
// depend on the order), materialise (its first stage loads eagerly) and
// the scan.  At most 4 x 3 + 3 = 15 plans.
int mbx_plan_candidates(const mbx_vprg *p, const double *dens, mbx_plan *out,
                        int max, int *count) {
  if (!p || !dens || !out || !count) return fail(MBX_ERR_INVALID, "candidates: null");
  mbx_plan base;
  int st = mbx_plan_heuristic(p, dens, &base);
  if (st) return st;
  base.flags = MBX_PLAN_FUSED;
  std::vector<mbx_plan> all;
  const int n = base.reorderable ? base.norder : 0;
  if (n >= 2) {
    for (int a = 0; a < n; ++a)
      for (int b = 0; b < n; ++b) {
        if (a == b) continue;
        mbx_plan q = base;
        int k = 0;
        q.order[k++] = base.order[a];
        q.order[k++] = base.order[b];
        for (int j = 0; j < n; ++j)
          if (j != a && j != b) q.order[k++] = base.order[j];
        all.push_back(q);
      }
  } else {
    all.push_back(base);
  }
  for (uint32_t fl : {(uint32_t)(MBX_PLAN_FUSED | MBX_PLAN_EAGER),
                      (uint32_t)MBX_PLAN_MATERIALISE, (uint32_t)MBX_PLAN_SCAN}) {
    mbx_plan q = base;
    q.flags = fl;
    all.push_back(q);
  }
  *count = (int)all.size();
  if ((int)all.size() > max) return fail(MBX_ERR_CAPACITY, "candidates: max too small");
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

// Learned chooser: scores the candidates (mbx_plan_candidates) with the MLP
// (output 0 = predicted log runtime) on the device and returns the
// cheapest -- unless the static heuristic's plan is predicted within
// `margin` of it (log units: 0.1 = about 10 %), in which case the heuristic
// plan stays: the heuristic is the fallback and the yardstick, and a
// prediction that barely beats it is not worth a different plan.  mlp ==
// NULL returns the heuristic plan.
int mbx_plan_choose_ex(mbx_ctx *ctx, mbx_mlp *mlp, const mbx_vprg *p,
                       const double *dens, const double *edge, uint32_t nr_grp,
                       float margin, mbx_plan *best, float *pred) {
  if (!ctx || !p || !dens || !best) return fail(MBX_ERR_INVALID, "choose: null");
  mbx_plan heur;
  int st = mbx_plan_heuristic(p, dens, &heur);
  if (st) return st;
  if (!mlp) {
    *best = heur;
    return MBX_OK;
  }
  mbx_plan cand[33];
  int n = 0;
  st = mbx_plan_candidates(p, dens, cand, 32, &n);
  if (st) return st;
  cand[n++] = heur;   // scored too, as the yardstick
  std::vector<float> feat((size_t)n * MBX_NFEAT);
  for (int i = 0; i < n; ++i) {
    st = mbx_plan_features_ex(p, dens, edge, nr_grp, &cand[i],
                              &feat[(size_t)i * MBX_NFEAT]);
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
  if (y[bi] > y[n - 1] - margin) bi = n - 1;   // not clearly better
  *best = cand[bi];
  if (pred) *pred = y[bi];
  return MBX_OK;
}

int mbx_plan_choose(mbx_ctx *ctx, mbx_mlp *mlp, const mbx_vprg *p,
                    const double *dens, uint32_t nr_grp, mbx_plan *best,
                    float *pred) {
  return mbx_plan_choose_ex(ctx, mlp, p, dens, nullptr, nr_grp, 0.f, best,
                            pred);
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
