// bmpidx.cpp -- the C++ host API (namespace mybmpidx) over the C ABI.
#include "../../include/mbx_bmpidx.hpp"

#include <algorithm>
#include <cstring>
#include <mutex>
#include <string>

#include "mbx_internal.hpp"

namespace {

mbx_ctx *g_ctx = nullptr;
std::mutex g_ctx_mu;

// from_bin over any bin handle type (bmpidx_create.cu:138-163).  Returns
// false (col left empty) on an uncovered range or more than 4 middle bins.
template <typename T>
bool from_bin_impl(size_t nbin, const T *bins, const uint64_t *bounds,
                   uint64_t min, uint64_t max, T middle[4], T *left, T *right,
                   T none) {
  for (int k = 0; k < 4; ++k) middle[k] = none;
  *left = *right = none;
  if (nbin == 0 || min >= max || bounds[0] > min || bounds[nbin] < max)
    return false;
  while (bounds[1] <= min) --nbin, ++bins, ++bounds;   // bins left of range
  while (bounds[nbin - 1] >= max) --nbin;              // bins right of range
  if (nbin == 0) return false;
  if (nbin == 1) {
    if (bounds[0] == min && bounds[1] == max) middle[0] = bins[0];
    else *left = bins[0];
    return true;
  }
  if (bounds[0] != min) {
    *left = bins[0];
    ++bounds, ++bins, --nbin;
  }
  if (bounds[nbin] != max) *right = bins[--nbin];
  if (nbin > 4) {
    for (int k = 0; k < 4; ++k) middle[k] = none;
    *left = *right = none;
    mbx::set_error("from_bin: more than 4 middle bins");
    return false;
  }
  for (size_t k = 0; k < nbin; ++k) middle[k] = bins[k];
  return true;
}

// recipe bin plans (bmpidx_create.cu:228-316), uint32 arithmetic as there.
struct BinReq {
  uint32_t lo, hi;
  int slot;  // 0..3 middle, 4 leftmost, 5 rightmost
};

int recipe_plan(int kind, int i, uint32_t mn, uint32_t mx, BinReq *r) {
  int n = 0;
  const uint32_t step = (mx - mn) / 3;
  if (kind == 0) {                      // perfect (:228-239)
    r[n++] = {mn, mx, 0};
    return n;
  }
  if (kind == 2 && i == 0) {            // candchk column 0 (:268-282)
    if (step == 0) {
      r[n++] = {mn, mx + 1, 4};
    } else {
      const uint32_t l = std::max(mn, step / 2) - step / 2;
      r[n++] = {l, mn + step, 4};
      r[n++] = {mn + step, mn + 2 * step, 0};
      r[n++] = {mn + 2 * step, mx, 1};
    }
    return n;
  }
  if (kind == 2 && i == 1) {            // candchk column 1 (:284-298)
    if (step == 0) {
      // the reference computes min - 1 in uint and so builds an empty bin
      // when min == 0 (bmpidx_create.cu:289); clamp at 0 instead
      r[n++] = {mn ? mn - 1 : 0, mx, 5};
    } else {
      r[n++] = {mn + 2 * step, mx + step / 2, 5};
      r[n++] = {mn, mn + step, 0};
      r[n++] = {mn + step, mn + 2 * step, 1};
    }
    return n;
  }
  // many_or, also candchk columns >= 2 (:241-262, :300-314)
  if (step == 0) {
    r[n++] = {mn, mx, 0};
    return n;
  }
  r[n++] = {mn, mn + step, 0};
  r[n++] = {mn + step, mn + 2 * step, 1};
  r[n++] = {mn + 2 * step, mx, 2};
  return n;
}

}  // namespace

namespace mybmpidx {

mbx_ctx *context() {
  std::lock_guard<std::mutex> g(g_ctx_mu);
  if (!g_ctx) {
    if (mbx_init(0, 0, &g_ctx) != MBX_OK) g_ctx = nullptr;
  }
  return g_ctx;
}

void use_context(mbx_ctx *ctx) {
  std::lock_guard<std::mutex> g(g_ctx_mu);
  g_ctx = ctx;
}

static uint32_t elem_bytes_of(size_t nbit) {
  return nbit <= 8 ? 1 : nbit <= 16 ? 2 : nbit <= 32 ? 4 : 8;
}

uint *create_select(const void *src, size_t nbit, uint factsz, uint min,
                    uint max) {
  uint32_t *out = nullptr;
  mbx_ctx *c = context();
  if (!c) return nullptr;
  if (mbx_create_select(c, src, elem_bytes_of(nbit), factsz, min, max, &out))
    return nullptr;
  return out;
}

uint *create_join(const uint32_t *fk, const void *dimattr, size_t nbit,
                  uint factsz, uint dimsz, uint min, uint max) {
  if (fk == nullptr) return create_select(dimattr, nbit, factsz, min, max);
  uint32_t *out = nullptr;
  mbx_ctx *c = context();
  if (!c) return nullptr;
  if (mbx_create_join(c, fk, dimattr, elem_bytes_of(nbit), factsz, dimsz, min,
                      max, &out))
    return nullptr;
  return out;
}

size_t create_bin(const uint *fk, const void *attr, size_t nbit, uint factsz,
                  uint dimsz, const uint64_t *min, const uint64_t *max,
                  size_t ncol, uint **out) {
  mbx_ctx *c = context();
  if (!c) return 0;
  mbx_colref src;
  src.data = attr;
  src.fk = fk;
  src.dimsz = dimsz;
  src.elem_bytes = elem_bytes_of(nbit);
  size_t single = 0;
  if (mbx_create_bin(c, &src, factsz, min, max, ncol, out, &single)) return 0;
  return single;
}

void free_bitmap(const uint *p) {
  if (p && context()) mbx_free(context(), const_cast<uint *>(p));
}

void col::release() noexcept {
  free_bitmap(leftmost);
  free_bitmap(rightmost);
  for (size_t i = 0; i < MAXBIN_PERCOL; ++i) free_bitmap(middle[i]);
  *this = col();
}

col col::from_bin(size_t n, uint *bins[], const size_t bounds[], size_t min,
                  size_t max) {
  col res;
  std::vector<uint64_t> b(bounds, bounds + n + 1);
  if (!from_bin_impl<uint *>(n, bins, b.data(), min, max, res.middle,
                             &res.leftmost, &res.rightmost, nullptr))
    return col();
  return res;
}

col col::from_bin2(size_t nc, uint *coarse_bins[], const size_t cbounds[],
                   size_t nf, uint *fine_bins[], const size_t fbounds[],
                   size_t min, size_t max) {
  col res = from_bin(nc, coarse_bins, cbounds, min, max);
  if (res.leftmost || res.rightmost)
    return from_bin(nf, fine_bins, fbounds, min, max);
  return res;
}

mbx_col col::to_abi() const {
  mbx_col c;
  for (size_t k = 0; k < MAXBIN_PERCOL; ++k) c.middle[k] = middle[k];
  c.leftmost = leftmost;
  c.rightmost = rightmost;
  return c;
}

void vprg::release() noexcept {
  for (size_t i = 0; i < MAXCOLS; ++i) cols[i].release();
}

bool vprg::is_nochk() const noexcept {
  for (size_t i = 0; i < MAXCOLS; ++i)
    if (!cols[i].is_nochk()) return false;
  return true;
}

mbx_vprg vprg::to_abi() const {
  mbx_vprg p;
  std::memset(&p, 0, sizeof(p));
  for (size_t i = 0; i < MAXCOLS; ++i) p.cols[i] = cols[i].to_abi();
  p.factsz = factsz;
  for (size_t i = 0; i < MAXNINSTR; ++i) {
    p.instrs[i].opcode = instrs[i].opcode;
    p.instrs[i].dstreg = instrs[i].dstreg;
    p.instrs[i].src1 = instrs[i].src1;
    p.instrs[i].src2 = instrs[i].src2;
  }
  return p;
}

static vprg build(const recipe &r, const vprg::instr *instrs, int kind) {
  vprg ret;
  ret.factsz = r.factsz;
  std::memcpy(ret.instrs, instrs, sizeof(vprg::instr) * MAXNINSTR);
  for (int i = 0; i < (int)MAXCOLS; ++i) {
    if (r.attr[i] == nullptr) {
      // candchk stops at the first absent column (bmpidx_create.cu:269,285,302)
      if (kind == 2) break;
      continue;
    }
    BinReq req[6];
    const int n = recipe_plan(kind, i, r.min[i], r.max[i], req);
    for (int k = 0; k < n; ++k) {
      uint *b = create_join(r.fk[i], r.attr[i], r.nbit[i], r.factsz,
                            r.dimsz[i], req[k].lo, req[k].hi);
      col &c = ret.cols[i];
      if (req[k].slot < 4) c.middle[req[k].slot] = b;
      else if (req[k].slot == 4) c.leftmost = b;
      else c.rightmost = b;
    }
  }
  return ret;
}

vprg recipe::perfect(const vprg::instr instrs[MAXNINSTR]) const {
  return build(*this, instrs, 0);
}
vprg recipe::many_or(const vprg::instr instrs[MAXNINSTR]) const {
  return build(*this, instrs, 1);
}
vprg recipe::candchk(const vprg::instr instrs[MAXNINSTR]) const {
  return build(*this, instrs, 2);
}

int vprg_grpby(const vprg &p, const mbx_rowop &op, uint nr_grp, int64_t *out,
               uint32_t plan) {
  mbx_ctx *c = context();
  if (!c) return MBX_ERR_NODEVICE;
  const mbx_vprg a = p.to_abi();
  return mbx_prog_grpby(c, &a, &op, nr_grp, out, plan);
}

}  // namespace mybmpidx

extern "C" {

int mbx_col_from_bin(size_t nbin, const uint64_t *ids, const uint64_t *bounds,
                     uint64_t min, uint64_t max, uint64_t *slots) {
  if (!ids || !bounds || !slots)
    return mbx::fail(MBX_ERR_INVALID, "col_from_bin: null");
  const uint64_t none = UINT64_MAX;
  uint64_t mid[4], l, r;
  mbx::set_error("");
  const bool ok =
      from_bin_impl<uint64_t>(nbin, ids, bounds, min, max, mid, &l, &r, none);
  for (int k = 0; k < 4; ++k) slots[k] = mid[k];
  slots[4] = l;
  slots[5] = r;
  if (!ok && std::string(mbx_last_error()).find("middle") != std::string::npos)
    return MBX_ERR_CAPACITY;
  return MBX_OK;
}

int mbx_col_from_bin2(size_t nc, const uint64_t *cids, const uint64_t *cb,
                      size_t nf, const uint64_t *fids, const uint64_t *fb,
                      uint64_t min, uint64_t max, uint64_t *slots) {
  int st = mbx_col_from_bin(nc, cids, cb, min, max, slots);
  if (st) return st;
  if (slots[4] != UINT64_MAX || slots[5] != UINT64_MAX)
    return mbx_col_from_bin(nf, fids, fb, min, max, slots);
  return MBX_OK;
}

int mbx_recipe_plan(int kind, int colidx, uint32_t min, uint32_t max,
                    uint32_t *lo, uint32_t *hi, int *slot, int *n) {
  if (!lo || !hi || !slot || !n || kind < 0 || kind > 2 || colidx < 0 ||
      colidx >= (int)MBX_MAXCOLS)
    return mbx::fail(MBX_ERR_INVALID, "recipe_plan: bad argument");
  BinReq req[6];
  const int m = recipe_plan(kind, colidx, min, max, req);
  for (int k = 0; k < m; ++k) lo[k] = req[k].lo, hi[k] = req[k].hi,
                              slot[k] = req[k].slot;
  *n = m;
  return MBX_OK;
}

}  // extern "C"
