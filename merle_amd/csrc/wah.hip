// wah.hip -- the WAH compressed-bitmap baseline on gfx950: the format and
// operations the reference's report compares the fused path against
// (merle/wahGpuCmprs.cu, merle/wahGpuDirect.cu, ssb/wah-demo.cu).
//
// Group layout: 31 rows per uint32, row 31g+j at bit 30-j of group g, bit 31
// zero (dbjoinFlatWah, wahGpuCmprs.cu:247-320); ceil(N/31) groups, rows >= N
// zero.  WAH words: bit 31 set = a fill (bit 30 = the fill bit, bits 0-29 =
// groups), else a literal holding one group.
//
// Kernels: select (a predicate straight into groups: each wave ballots 31
// row blocks of 64, stages the 31 masks in LDS and lane g cuts group g out
// of them), groups (the same from a 64-row-word bitmap), compress (keep
// flags -> tile sums -> tile offsets -> ordered scatter -> fill lengths),
// count scan (the same three-pass tile scan), AND in the down operand's
// word structure, decompress, and the longest 1-fill.  The AND and the
// decompression stage, per wave, the other operand's words that cover the
// wave's groups in LDS and search them there.
#include <algorithm>
#include <cstring>
#include <mutex>

#include "devfuncs.hpp"
#include "mbx_internal.hpp"

namespace mbx {

constexpr uint32_t kWahFill = 0x80000000u;
constexpr uint32_t kWahOne = 0x40000000u;
constexpr uint32_t kWahCnt = 0x3fffffffu;
constexpr uint32_t kWahRows = 0x7fffffffu;
// a tile of the scans: 256 threads x kWahItems consecutive items
constexpr uint32_t kWahItems = 16;
constexpr uint32_t kWahTile = 256 * kWahItems;
// groups between consecutive entries of a skip table: entry i names the
// word covering group kWahSkip * i, bit 31 set when the word starts there
constexpr uint32_t kWahSkip = 1024;
constexpr uint32_t kWahSkipAt = 0x80000000u;

__device__ __forceinline__ uint32_t wah_count(uint32_t w) {
  return (w & kWahFill) ? (w & kWahCnt) : 1u;
}
__device__ __forceinline__ uint32_t wah_rows(uint32_t w) {
  if (w & kWahFill) return (w & kWahOne) ? kWahRows : 0u;
  return w & kWahRows;
}
// 31 rows with row j at bit j -> row j at bit 30 - j
__device__ __forceinline__ uint32_t wah_order(uint32_t x) {
  return __builtin_bitreverse32(x & kWahRows) >> 1;
}

// Exclusive scan over the 256 threads of a workgroup; *total = the sum.
__device__ __forceinline__ uint32_t wah_block_scan(uint32_t x, uint32_t *s_w,
                                                   uint32_t *total) {
  const uint32_t lane = lane_id(), wv = threadIdx.x >> 6;
  const uint32_t incl = wave_incl_scan(x);
  if (lane == 63) s_w[wv] = incl;
  __syncthreads();
  uint32_t before = 0, sum = 0;
#pragma unroll
  for (uint32_t j = 0; j < 4; ++j) {
    const uint32_t t = s_w[j];
    before += j < wv ? t : 0;
    sum += t;
  }
  __syncthreads();
  *total = sum;
  return before + incl - x;
}

// ------------------------------------------------------------ groups
template <typename T>
__device__ __forceinline__ uint64_t wah_elem(const void *p, uint64_t i) {
  return static_cast<const T *>(p)[i];
}

// Wave w covers rows [1984w, 1984w + 1984): 31 ballots of 64 rows, then
// lane g holds group 64w + g = bits 31g..31g+30 of the 1984 ballot bits.
// T: the element type of the column (a load per row with no branch on the
// width, so the 31 loads of a lane are in flight together).
template <typename T>
__global__ __launch_bounds__(256) void wah_select_kernel(DevSrc src,
                                                         uint32_t nrows,
                                                         uint64_t lo,
                                                         uint64_t hi,
                                                         uint32_t *groups,
                                                         uint32_t ngroups) {
  __shared__ uint64_t s_m[4][32];
  const uint32_t lane = lane_id(), wv = threadIdx.x >> 6;
  const uint64_t wave = (uint64_t)blockIdx.x * 4 + wv;
  const uint64_t r0 = wave * 1984u;
  uint64_t v[31];
  bool ok[31];
  if (src.fk && src.dimsz == 0) {   // no dimension row: nothing selected
#pragma unroll
    for (int b = 0; b < 31; ++b) {
      ok[b] = false;
      v[b] = 0;
    }
  } else if (src.fk) {
    uint32_t k[31];
#pragma unroll
    for (int b = 0; b < 31; ++b) {
      const uint64_t row = r0 + 64u * b + lane;
      k[b] = row < nrows ? src.fk[row] : 0xffffffffu;
    }
#pragma unroll
    for (int b = 0; b < 31; ++b) {
      ok[b] = k[b] < src.dimsz;
      v[b] = wah_elem<T>(src.data, ok[b] ? k[b] : 0u);
    }
  } else {
#pragma unroll
    for (int b = 0; b < 31; ++b) {
      const uint64_t row = r0 + 64u * b + lane;
      ok[b] = row < nrows;
      v[b] = wah_elem<T>(src.data, ok[b] ? row : 0u);
    }
  }
#pragma unroll
  for (int b = 0; b < 31; ++b) {
    const uint64_t m = __ballot(ok[b] && v[b] >= lo && v[b] < hi);
    if (lane == 0) s_m[wv][b] = m;
  }
  if (lane == 0) s_m[wv][31] = 0;
  wave_lds_sync();
  const uint32_t bit = 31u * lane, w = bit >> 6, sh = bit & 63u;
  uint64_t x = s_m[wv][w] >> sh;
  if (sh > 33) x |= s_m[wv][w + 1] << (64 - sh);
  const uint64_t g = wave * 64u + lane;
  if (g < ngroups) groups[g] = wah_order((uint32_t)x);
}

// From a bitmap (uint32 words, bit b of word k = row 32k+b): group g is
// bits 31g..31g+30, within two words.
__global__ __launch_bounds__(256) void wah_groups_kernel(
    const uint32_t *__restrict__ bm, uint32_t nrows, uint32_t *groups,
    uint32_t ngroups) {
  const uint32_t nw = (uint32_t)(((uint64_t)nrows + 31) / 32);
  for (uint64_t g = (uint64_t)blockIdx.x * 256 + threadIdx.x; g < ngroups;
       g += (uint64_t)gridDim.x * 256) {
    const uint64_t bit = 31u * g;
    const uint32_t w = (uint32_t)(bit >> 5), sh = (uint32_t)(bit & 31u);
    uint32_t x = bm[w] >> sh;
    if (sh > 1 && w + 1 < nw) x |= bm[w + 1] << (32 - sh);
    const uint64_t left = (uint64_t)nrows - bit;   // rows of this group < N
    if (left < 31) x &= (1u << left) - 1u;
    groups[g] = wah_order(x);
  }
}

// ---------------------------------------------------------- compress
// A group is kept when it is a literal or the last of its fill run; the
// last group only when it is not zero, so a trailing zero run leaves
// nothing (wahCompress, wahGpuCmprs.cu:215-245).
__device__ __forceinline__ bool wah_keep(const uint32_t *groups, uint32_t g,
                                         uint32_t ng) {
  const uint32_t v = groups[g] & kWahRows;
  if (g + 1 == ng) return v != 0;
  if (v != 0 && v != kWahRows) return true;
  return (groups[g + 1] & kWahRows) != v;
}

__global__ __launch_bounds__(256) void wah_keep_count_kernel(
    const uint32_t *__restrict__ groups, uint32_t ng, uint32_t *tile_sums) {
  __shared__ uint32_t s_w[4];
  const uint32_t base = blockIdx.x * kWahTile;
  uint32_t c = 0;
#pragma unroll
  for (uint32_t k = 0; k < kWahItems; ++k) {
    const uint32_t g = base + k * 256 + threadIdx.x;
    c += (g < ng && wah_keep(groups, g, ng)) ? 1u : 0u;
  }
  uint32_t total;
  (void)wah_block_scan(c, s_w, &total);
  if (threadIdx.x == 0) tile_sums[blockIdx.x] = total;
}

// Exclusive scan of ntiles tile sums in place by one workgroup;
// tile_sums[ntiles] = the total.
__global__ __launch_bounds__(256) void wah_tiles_scan_kernel(
    uint32_t *tile_sums, uint32_t ntiles) {
  __shared__ uint32_t s_w[4];
  uint32_t carry = 0;
  for (uint32_t b = 0; b < ntiles; b += 256) {
    const uint32_t i = b + threadIdx.x;
    const uint32_t x = i < ntiles ? tile_sums[i] : 0u;
    uint32_t total;
    const uint32_t ex = wah_block_scan(x, s_w, &total);
    if (i < ntiles) tile_sums[i] = carry + ex;
    carry += total;
  }
  if (threadIdx.x == 0) tile_sums[ntiles] = carry;
}

// LDS index of item j of a tile that is written coalesced (item k * 256 +
// thread) and read by thread t as its 16 consecutive items: j + j / 16
// keeps both patterns free of bank conflicts (thread t's items start 17t
// apart).
__device__ __forceinline__ uint32_t wah_pad(uint32_t j) { return j + (j >> 4); }

// Kept groups in order: out[pos] = the literal or the fill's kind, and
// ends[pos] = g + 1, the end of the run the word closes.  The tile's
// groups (and the next tile's first) are read coalesced into LDS; thread t
// then takes its 16 consecutive groups there.

__global__ __launch_bounds__(256) void wah_scatter_kernel(
    const uint32_t *__restrict__ groups, uint32_t ng,
    const uint32_t *__restrict__ tile_offs, uint32_t *out, uint32_t *ends) {
  __shared__ uint32_t s_w[4];
  __shared__ uint32_t s_g[kWahTile + kWahTile / 16 + 1];
  const uint32_t base = blockIdx.x * kWahTile, t = threadIdx.x;
#pragma unroll
  for (uint32_t k = 0; k < kWahItems; ++k) {
    const uint32_t j = k * 256 + t, g = base + j;
    s_g[wah_pad(j)] = g < ng ? groups[g] & kWahRows : 0u;
  }
  if (t == 0)
    s_g[wah_pad(kWahTile)] = base + kWahTile < ng ? groups[base + kWahTile] & kWahRows
                                                  : 0u;
  __syncthreads();
  uint32_t keep = 0;
#pragma unroll
  for (uint32_t i = 0; i < kWahItems; ++i) {
    const uint32_t j = t * kWahItems + i, g = base + j;
    if (g >= ng) break;
    const uint32_t v = s_g[wah_pad(j)];
    bool kp;
    if (g + 1 == ng) kp = v != 0;
    else if (v != 0 && v != kWahRows) kp = true;
    else kp = s_g[wah_pad(j + 1)] != v;
    if (kp) keep |= 1u << i;
  }
  uint32_t total;
  uint32_t pos = tile_offs[blockIdx.x] +
                 wah_block_scan((uint32_t)__popc(keep), s_w, &total);
#pragma unroll
  for (uint32_t i = 0; i < kWahItems; ++i) {
    if (!((keep >> i) & 1u)) continue;
    const uint32_t j = t * kWahItems + i;
    const uint32_t v = s_g[wah_pad(j)];
    out[pos] = v == 0 ? kWahFill : v == kWahRows ? (kWahFill | kWahOne) : v;
    ends[pos] = base + j + 1;
    ++pos;
  }
}

// A fill's length: its run is every group since the previous kept word.
__global__ __launch_bounds__(256) void wah_fill_len_kernel(
    uint32_t *out, const uint32_t *__restrict__ ends,
    const uint32_t *__restrict__ count) {
  const uint32_t n = *count;
  for (uint32_t i = blockIdx.x * 256 + threadIdx.x; i < n;
       i += gridDim.x * 256) {
    const uint32_t w = out[i];
    if (w & kWahFill) out[i] = w | (ends[i] - (i ? ends[i - 1] : 0u));
  }
}

// ------------------------------------------------------- count scan
// Tile sums read the words coalesced (word base + k * 256 + thread);
// the scan itself stages its tile in LDS the same way, then thread t scans
// its 16 consecutive words there.
__global__ __launch_bounds__(256) void wah_count_sum_kernel(
    const uint32_t *__restrict__ wah, uint32_t n, uint32_t *tile_sums) {
  __shared__ uint32_t s_w[4];
  const uint32_t base = blockIdx.x * kWahTile;
  uint32_t c = 0;
#pragma unroll
  for (uint32_t k = 0; k < kWahItems; ++k) {
    const uint32_t i = base + k * 256 + threadIdx.x;
    c += i < n ? wah_count(wah[i]) : 0u;
  }
  uint32_t total;
  (void)wah_block_scan(c, s_w, &total);
  if (threadIdx.x == 0) tile_sums[blockIdx.x] = total;
}

// Tile counts up to which the scan's tiles sum their predecessors' sums
// themselves (at most 8 loads a thread) instead of a scan launch between.
constexpr uint32_t kWahSelfPrefixTiles = 2048;

// kSums: tiles holds the tiles' sums and each tile adds up those before it;
// otherwise tiles holds their exclusive scan and the total at [gridDim.x].
template <bool kSums>
__global__ __launch_bounds__(256) void wah_count_scan_kernel(
    const uint32_t *__restrict__ wah, uint32_t n,
    const uint32_t *__restrict__ tiles, uint32_t *excl, uint32_t *skip,
    uint32_t nskip) {
  __shared__ uint32_t s_w[4];
  __shared__ uint32_t s_c[kWahTile + kWahTile / 16];
  const uint32_t base = blockIdx.x * kWahTile, t = threadIdx.x;
#pragma unroll
  for (uint32_t k = 0; k < kWahItems; ++k) {
    const uint32_t j = k * 256 + t, i = base + j;
    s_c[wah_pad(j)] = i < n ? wah_count(wah[i]) : 0u;
  }
  uint32_t off;
  if (kSums) {
    uint32_t p = 0;
    for (uint32_t i = t; i < blockIdx.x; i += 256) p += tiles[i];
    (void)wah_block_scan(p, s_w, &off);   // its barriers order s_c too
  } else {
    off = tiles[blockIdx.x];
    __syncthreads();
  }
  uint32_t sum = 0;
#pragma unroll
  for (uint32_t i = 0; i < kWahItems; ++i) sum += s_c[wah_pad(t * kWahItems + i)];
  uint32_t total;
  uint32_t acc = off + wah_block_scan(sum, s_w, &total);
#pragma unroll
  for (uint32_t i = 0; i < kWahItems; ++i) {
    const uint32_t p = wah_pad(t * kWahItems + i), v = s_c[p];
    s_c[p] = acc;
    acc += v;
  }
  __syncthreads();
#pragma unroll
  for (uint32_t k = 0; k < kWahItems; ++k) {
    const uint32_t j = k * 256 + t, i = base + j;
    if (i < n) excl[i] = s_c[wah_pad(j)];
  }
  if (blockIdx.x == gridDim.x - 1 && t == 255)
    excl[n] = kSums ? off + total : tiles[gridDim.x];
  if (!skip) return;
  // the skip entries whose group lies in the tile's groups [off, off +
  // total): the last of the tile's words starting at or before it
  const uint32_t tn = min(n - base, kWahTile);
  const uint64_t hi = (uint64_t)off + total;
  const uint64_t iend = min<uint64_t>((hi + kWahSkip - 1) / kWahSkip, nskip);
  for (uint64_t i = ((uint64_t)off + kWahSkip - 1) / kWahSkip + t; i < iend;
       i += 256) {
    const uint32_t g = (uint32_t)(i * kWahSkip);
    uint32_t a = 0, b = tn;   // s_c[a] <= g (item 0 starts at off <= g)
    while (b - a > 1) {
      const uint32_t m = (a + b) >> 1;
      if (s_c[wah_pad(m)] <= g) a = m;
      else b = m;
    }
    skip[i] = (base + a) | (s_c[wah_pad(a)] == g ? kWahSkipAt : 0u);
  }
}

// ------------------------------------------------------ AND, decompress
// The word of (excl, n) holding group g: excl[u] <= g < excl[u + 1]; n when
// g is past the last group.
__device__ __forceinline__ uint32_t wah_find(const uint32_t *__restrict__ excl,
                                             uint32_t n, uint32_t g) {
  if (n == 0 || g >= excl[n]) return n;
  uint32_t lo = 0, hi = n;
  while (hi - lo > 1) {
    const uint32_t mid = lo + ((hi - lo) >> 1);
    if (excl[mid] <= g) lo = mid;
    else hi = mid;
  }
  return lo;
}

// A wave's items (down words of the AND, groups of the decompression) are
// kWahSpan consecutive ones, lane l taking items l, l + 64, ... so that
// every load and store is coalesced.  The words of the other operand that
// cover the wave's groups are staged in LDS, kWahSpan words at a time
// (their group offsets and their 31 rows), from the word covering the
// wave's first group on.  When one staging covers every item's group, the
// owner scan (wah_owners) names each group's word at once; otherwise an
// item costs a 10-step binary search in LDS, and as the items' groups
// increase with the item, an item the staged words do not reach is
// reached by a later chunk of words.
constexpr uint32_t kWahSpan = kWahSkip;
constexpr uint32_t kWahPerLane = kWahSpan / 64;

struct WahStage {
  uint32_t ex[4][kWahSpan + 1];
  uint32_t rows[4][kWahSpan];
};

// wah_find by the whole wave (g wave-uniform, below excl[n]): each round
// the 64 lanes test 64 evenly spaced words of the range and the range
// shrinks to the span between the last one at or before g and the next, so
// a few million words take 4 dependent rounds instead of 22.
__device__ __forceinline__ uint32_t wah_wave_find(
    const uint32_t *__restrict__ excl, uint32_t n, uint32_t g) {
  uint32_t lo = 0, hi = n;   // excl[lo] <= g < excl[hi]
  while (hi - lo > 1) {
    const uint32_t stride = (hi - lo + 63) / 64;
    const uint32_t idx = lo + lane_id() * stride;
    const bool ok = idx < hi && excl[idx] <= g;   // a prefix of the lanes
    const uint64_t b = __ballot(ok);
    const uint32_t last = 63u - (uint32_t)__builtin_clzll(b);   // lane 0 is ok
    const uint32_t nlo = lo + last * stride;
    hi = min(hi, nlo + stride);
    lo = nlo;
  }
  return lo;
}

// The word covering group kWahSkip * i (below excl[n]): from the skip table
// when it holds entry i, else by the wave search.  Wave-uniform i.  A table
// entry is kept below n, so a table of another bitmap gives wrong words,
// never reads out of bounds.
__device__ __forceinline__ uint32_t wah_cover_of(
    const uint32_t *__restrict__ excl, const uint32_t *__restrict__ skip,
    uint32_t nskip, uint32_t n, uint32_t i) {
  if (skip && i < nskip) return min(skip[i] & ~kWahSkipAt, n - 1);
  return wah_wave_find(excl, n, i * kWahSkip);
}

// The first word starting at or after group kWahSkip * i (below excl[n]).
__device__ __forceinline__ uint32_t wah_first_of(
    const uint32_t *__restrict__ excl, const uint32_t *__restrict__ skip,
    uint32_t nskip, uint32_t n, uint32_t i) {
  if (skip && i < nskip) {
    const uint32_t e = skip[i], u = min(e & ~kWahSkipAt, n - 1);
    return (e & kWahSkipAt) ? u : u + 1;
  }
  const uint32_t u = wah_wave_find(excl, n, i * kWahSkip);
  return excl[u] == i * kWahSkip ? u : u + 1;
}

// Stages words [w0, w0 + c), c = min(n - w0, kWahSpan), of (wah, excl, n):
// group offsets, rows, and the offset after the last (returned).  Every
// load is in flight before the first store (a lane past c reads word w0
// again, dropped).
__device__ __forceinline__ uint32_t wah_stage(const uint32_t *__restrict__ wah,
                                              const uint32_t *__restrict__ excl,
                                              uint32_t n, uint32_t w0,
                                              uint32_t *s_ex, uint32_t *s_rows,
                                              uint32_t *cnt) {
  const uint32_t c = min(n - w0, kWahSpan);
  *cnt = c;
  uint32_t e[kWahPerLane], r[kWahPerLane];
#pragma unroll
  for (uint32_t k = 0; k < kWahPerLane; ++k) {
    const uint32_t j = lane_id() + 64u * k, at = w0 + (j < c ? j : 0u);
    e[k] = excl[at];
    r[k] = wah[at];
  }
  const uint32_t end = excl[w0 + c];   // w0 + c <= n
  wave_lds_sync();   // the previous chunk's searches are done
#pragma unroll
  for (uint32_t k = 0; k < kWahPerLane; ++k) {
    const uint32_t j = lane_id() + 64u * k;
    if (j < c) {
      s_ex[j] = e[k];
      s_rows[j] = wah_rows(r[k]);
    }
  }
  if (lane_id() == 0) s_ex[c] = end;
  wave_lds_sync();
  return end;
}

// The staged word covering group g (s_ex[0] <= g < s_ex[cnt]).
__device__ __forceinline__ uint32_t wah_staged(const uint32_t *s_ex,
                                               const uint32_t *s_rows,
                                               uint32_t cnt, uint32_t g) {
  uint32_t lo = 0, hi = cnt;
  while (hi - lo > 1) {
    const uint32_t mid = (lo + hi) >> 1;
    if (s_ex[mid] <= g) lo = mid;
    else hi = mid;
  }
  return s_rows[lo];
}

// Inclusive max over the wave's lanes.
__device__ __forceinline__ uint32_t wave_max_scan(uint32_t x) {
  const uint32_t lane = lane_id();
#pragma unroll
  for (uint32_t d = 1; d < 64; d <<= 1) {
    const uint32_t y = __shfl_up(x, d, 64);
    if (lane >= d) x = max(x, y);
  }
  return x;
}

// The staged word covering each of the kWahSpan groups [gb, gb + kWahSpan)
// (the staged words start at or before gb and cover the groups asked for):
// every staged word's index is written at its first group's offset, then a
// running maximum over the offsets, 64 at a time across the lanes, gives
// each offset its word.  Offset k * 64 + lane's word is left in own[k] and
// in s_ex[k * 64 + lane] (s_ex is overwritten).  Independent LDS accesses
// and shuffles instead of a 10-step dependent search per group.
__device__ __forceinline__ void wah_owners(uint32_t *s_ex, uint32_t cnt,
                                           uint32_t gb,
                                           uint32_t (&own)[kWahPerLane]) {
  const uint32_t lane = lane_id();
  // each word's start offset from gb, two 16-bit offsets a register;
  // 0xffff for word 0 (it covers gb and starts at or before it: offset 0
  // keeps 0), words past cnt and starts outside the span
  uint32_t offs[kWahPerLane / 2];
#pragma unroll
  for (uint32_t k = 0; k < kWahPerLane; ++k) {
    const uint32_t j = lane + 64u * k;
    const uint32_t st = j < cnt ? s_ex[j] : 0u;
    const uint32_t o =
        j > 0 && j < cnt && st >= gb && st - gb < kWahSpan ? st - gb : 0xffffu;
    if (k & 1) offs[k >> 1] |= o << 16;
    else offs[k >> 1] = o;
  }
  wave_lds_sync();
#pragma unroll
  for (uint32_t k = 0; k < kWahPerLane; ++k) s_ex[lane + 64u * k] = 0;
  wave_lds_sync();
#pragma unroll
  for (uint32_t k = 0; k < kWahPerLane; ++k) {
    // of words starting at one offset (all but the last count no group)
    // the last is kept
    const uint32_t o = (offs[k >> 1] >> (16u * (k & 1))) & 0xffffu;
    if (o != 0xffffu) atomicMax(&s_ex[o], lane + 64u * k);
  }
  wave_lds_sync();
  uint32_t carry = 0;
#pragma unroll
  for (uint32_t k = 0; k < kWahPerLane; ++k) {
    const uint32_t x = max(wave_max_scan(s_ex[64u * k + lane]), carry);
    carry = __shfl(x, 63, 64);
    own[k] = x;
    s_ex[64u * k + lane] = x;
  }
  wave_lds_sync();
}

// wahAndNo1 (wahGpuDirect.cu:8-34): down's word structure; a 0-fill stays,
// a literal (a 1-fill of one group reads as one) is ANDed with the up word
// covering its group (past up: 0), a zero result becomes a 0-fill of one.
// A word is stored as soon as it is known.
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(4)))
void wah_and_no1_kernel(
    const uint32_t *__restrict__ up, const uint32_t *__restrict__ up_excl,
    uint32_t nup, const uint32_t *__restrict__ down,
    const uint32_t *__restrict__ down_excl, uint32_t ndown, uint32_t *out) {
  __shared__ WahStage st;
  const uint32_t wv = threadIdx.x >> 6, lane = lane_id();
  const uint64_t d0 = ((uint64_t)blockIdx.x * 4 + wv) * kWahSpan;
  if (d0 >= ndown) return;   // the whole wave
  const uint32_t utot = nup ? up_excl[nup] : 0u;
  // the wave's down words and their groups, every load in flight at once
  // (a lane past the end reads word d0 again, dropped)
  uint32_t dwv[kWahPerLane], gv[kWahPerLane];
#pragma unroll
  for (uint32_t k = 0; k < kWahPerLane; ++k) {
    const uint64_t d = d0 + lane + 64u * k;
    const uint64_t at = d < ndown ? d : d0;
    const uint32_t dw = down[at];
    gv[k] = down_excl[at];
    dwv[k] = dw == (kWahFill | kWahOne | 1u) ? kWahRows : dw;
  }
  // the up words from the one covering the wave's first group
  const uint32_t g_first = __shfl(gv[0], 0, 64);   // item 0 of lane 0
  const uint64_t dl = min<uint64_t>(d0 + kWahSpan, ndown) - 1;
  const uint32_t g_last = down_excl[dl];            // the wave's last group
  uint32_t w0 = 0, cnt = 0, end = 0, pend = 0;
  if (g_first < utot) {
    w0 = wah_wave_find(up_excl, nup, g_first);
    end = wah_stage(up, up_excl, nup, w0, st.ex[wv], st.rows[wv], &cnt);
    if (g_last - g_first < kWahSpan && end > min(g_last, utot - 1)) {
      // every literal's group is within kWahSpan of the first and covered:
      // the covering words by offset, then one LDS read per literal
      uint32_t own[kWahPerLane];
      wah_owners(st.ex[wv], cnt, g_first, own);
      for (uint32_t k = 0; k < kWahPerLane; ++k) {
        const uint64_t d = d0 + lane + 64u * k;
        if (d >= ndown) break;
        const uint32_t dw = dwv[k], g = gv[k];
        uint32_t o = dw;
        if (!(dw & kWahFill)) {
          const uint32_t v =
              g < utot ? dw & st.rows[wv][st.ex[wv][g - g_first]] : 0u;
          o = v ? v : (kWahFill | 1u);
        }
        out[d] = o;
      }
      return;
    }
  }
  for (uint32_t k = 0; k < kWahPerLane; ++k) {
    const uint64_t d = d0 + lane + 64u * k;
    if (d >= ndown) break;
    const uint32_t dw = dwv[k], g = gv[k];
    uint32_t o = dw;
    if (!(dw & kWahFill)) {
      if (g < utot && g >= end) {
        pend |= 1u << k;
        continue;
      }
      const uint32_t v =
          g < utot ? dw & wah_staged(st.ex[wv], st.rows[wv], cnt, g) : 0u;
      o = v ? v : (kWahFill | 1u);
    }
    out[d] = o;
  }
  while (__ballot(pend != 0)) {
    w0 += cnt;
    end = wah_stage(up, up_excl, nup, w0, st.ex[wv], st.rows[wv], &cnt);
    for (uint32_t k = 0; k < kWahPerLane; ++k) {
      if (((pend >> k) & 1u) && gv[k] < end) {
        const uint32_t v = dwv[k] & wah_staged(st.ex[wv], st.rows[wv], cnt, gv[k]);
        out[d0 + lane + 64u * k] = v ? v : (kWahFill | 1u);
        pend &= ~(1u << k);
      }
    }
  }
}

// wahAndNo1 over bitmaps with skip tables: a wave takes the down words
// starting in one window of kWahSkip groups (at most kWahSkip words when
// each counts a group) and stages the up words covering the window (at
// most kWahSkip likewise) from the up table's entry, so one staging and
// the owner scan resolve every literal of the window.  A window whose
// words exceed those bounds (words counting no group) takes a binary
// search in device memory per literal.  Persistent: the window count is
// the down bitmap's groups / kWahSkip, known on the device only.
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(4)))
void wah_and_no1_win_kernel(
    const uint32_t *__restrict__ up, const uint32_t *__restrict__ up_excl,
    const uint32_t *__restrict__ up_skip, uint32_t nsu, uint32_t nup,
    const uint32_t *__restrict__ down, const uint32_t *__restrict__ down_excl,
    const uint32_t *__restrict__ down_skip, uint32_t nsd, uint32_t ndown,
    uint32_t *out) {
  __shared__ WahStage st;
  const uint32_t wv = threadIdx.x >> 6, lane = lane_id();
  const uint32_t dtot = down_excl[ndown];   // ndown >= 1
  const uint32_t utot = nup ? up_excl[nup] : 0u;
  const uint32_t nwin = (uint32_t)(((uint64_t)dtot + kWahSkip - 1) / kWahSkip);
  for (uint32_t w = blockIdx.x * 4 + wv; w < nwin; w += gridDim.x * 4) {
    const uint32_t gb = w * kWahSkip;
    // (window 0 also takes the words before the first word covering a
    // group, which count none)
    const uint32_t a = w ? wah_first_of(down_excl, down_skip, nsd, ndown, w) : 0u;
    const uint32_t b = (uint64_t)gb + kWahSkip < dtot
                           ? wah_first_of(down_excl, down_skip, nsd, ndown, w + 1)
                           : ndown;
    if (a >= b) continue;   // no word starts in the window
    if (b - a <= kWahSpan) {
      // the down words, their loads in flight with the up staging
      uint32_t dwv[kWahPerLane], gv[kWahPerLane];
#pragma unroll
      for (uint32_t k = 0; k < kWahPerLane; ++k) {
        const uint32_t d = a + lane + 64u * k, at = d < b ? d : a;
        const uint32_t dw = down[at];
        gv[k] = down_excl[at];
        dwv[k] = dw == (kWahFill | kWahOne | 1u) ? kWahRows : dw;
      }
      bool fast = true;   // no up group in the window: every literal reads 0
      if (gb < utot) {
        uint32_t cnt;
        const uint32_t c = wah_cover_of(up_excl, up_skip, nsu, nup, w);
        const uint32_t end =
            wah_stage(up, up_excl, nup, c, st.ex[wv], st.rows[wv], &cnt);
        fast = st.ex[wv][0] <= gb && end > min(gb + kWahSkip, utot) - 1;
        if (fast) {
          uint32_t own[kWahPerLane];
          wah_owners(st.ex[wv], cnt, gb, own);
        }
      }
      if (fast) {
#pragma unroll
        for (uint32_t k = 0; k < kWahPerLane; ++k) {
          const uint32_t d = a + lane + 64u * k;
          const uint32_t dw = dwv[k], g = gv[k];
          uint32_t o = dw;
          if (!(dw & kWahFill)) {
            const uint32_t v =
                g < utot
                    ? dw & st.rows[wv][st.ex[wv][min(g - gb, kWahSkip - 1)]]
                    : 0u;
            o = v ? v : (kWahFill | 1u);
          }
          if (d < b) out[d] = o;
        }
        continue;
      }
    }
    for (uint32_t d = a + lane; d < b; d += 64) {
      const uint32_t dw0 = down[d], g = down_excl[d];
      const uint32_t dw = dw0 == (kWahFill | kWahOne | 1u) ? kWahRows : dw0;
      uint32_t o = dw;
      if (!(dw & kWahFill)) {
        const uint32_t v =
            g < utot ? dw & wah_rows(up[wah_find(up_excl, nup, g)]) : 0u;
        o = v ? v : (kWahFill | 1u);
      }
      out[d] = o;
    }
  }
}

// wahDecomp (wahGpuCmprs.cu:37-51): groups of the words, zero past them.
// A group is stored as soon as a chunk reaches it (the first chunk reaches
// every group unless some word counts no group).
__global__ __launch_bounds__(256) void wah_decompress_kernel(
    const uint32_t *__restrict__ wah, const uint32_t *__restrict__ excl,
    const uint32_t *__restrict__ skip, uint32_t nskip, uint32_t n,
    uint32_t *groups, uint32_t ng) {
  __shared__ WahStage st;
  const uint32_t wv = threadIdx.x >> 6, lane = lane_id();
  const uint64_t g0 = ((uint64_t)blockIdx.x * 4 + wv) * kWahSpan;
  if (g0 >= ng) return;   // the whole wave
  const uint32_t total = n ? excl[n] : 0u;
  uint32_t w0 = 0, cnt = 0, end = 0, pend = 0;
  if (g0 < total) {
    w0 = wah_cover_of(excl, skip, nskip, n, (uint32_t)(g0 / kWahSkip));
    end = wah_stage(wah, excl, n, w0, st.ex[wv], st.rows[wv], &cnt);
    const uint32_t glast = (uint32_t)min<uint64_t>(g0 + kWahSpan, total) - 1;
    if (end > glast) {
      // the staged words cover every group of the wave below the total:
      // the covering word of each group by its offset
      uint32_t own[kWahPerLane];
      wah_owners(st.ex[wv], cnt, (uint32_t)g0, own);
#pragma unroll
      for (uint32_t k = 0; k < kWahPerLane; ++k) {
        const uint64_t g = g0 + lane + 64u * k;
        if (g < ng) groups[g] = g < total ? st.rows[wv][own[k]] : 0u;
      }
      return;
    }
  }
  for (uint32_t k = 0; k < kWahPerLane; ++k) {
    const uint64_t g = g0 + lane + 64u * k;
    if (g >= ng) break;
    if (g < total && g >= end) {
      pend |= 1u << k;
      continue;
    }
    groups[g] = g < total ? wah_staged(st.ex[wv], st.rows[wv], cnt, (uint32_t)g)
                          : 0u;
  }
  while (__ballot(pend != 0)) {
    w0 += cnt;
    end = wah_stage(wah, excl, n, w0, st.ex[wv], st.rows[wv], &cnt);
    for (uint32_t k = 0; k < kWahPerLane; ++k) {
      const uint32_t g = (uint32_t)(g0 + lane + 64u * k);
      if (((pend >> k) & 1u) && g < end) {
        groups[g] = wah_staged(st.ex[wv], st.rows[wv], cnt, g);
        pend &= ~(1u << k);
      }
    }
  }
}

// wahLongest1Fill (merle/host/wahConv.c:6-20): a thread starts at each word
// that opens a run of consecutive one fills and walks it.
__device__ __forceinline__ bool wah_is_one(uint32_t w) {
  return w == kWahRows || (w >> 30) == 3u;
}
__global__ __launch_bounds__(256) void wah_longest_1fill_kernel(
    const uint32_t *__restrict__ wah, uint32_t n, unsigned long long *best) {
  for (uint32_t i = blockIdx.x * 256 + threadIdx.x; i < n;
       i += gridDim.x * 256) {
    if (!wah_is_one(wah[i]) || (i > 0 && wah_is_one(wah[i - 1]))) continue;
    unsigned long long run = 0;
    for (uint32_t j = i; j < n && wah_is_one(wah[j]); ++j)
      run += wah[j] == kWahRows ? 1u : (wah[j] & kWahCnt);
    atomicMax(best, run);
  }
}

// ------------------------------------------------------------- host
static unsigned grid_of(size_t items, size_t per_block) {
  const size_t g = ceil_div(items, per_block);
  return g ? (unsigned)g : 1u;
}

// Pool scratch freed after the context's stream has drained.
struct Scratch {
  mbx_ctx *ctx;
  void *p = nullptr;
  explicit Scratch(mbx_ctx *c) : ctx(c) {}
  ~Scratch() {
    if (p) {
      (void)hipStreamSynchronize(ctx->stream);
      ctx->pool->free(p);
    }
  }
};

static int wah_check(mbx_ctx *ctx, const void *p, const char *what) {
  if (!ctx) return fail(MBX_ERR_INVALID, std::string(what) + ": null ctx");
  if (reinterpret_cast<uintptr_t>(p) % 4)
    return fail(MBX_ERR_ALIGN, std::string(what) + ": misaligned array");
  return MBX_OK;
}

}  // namespace mbx

using namespace mbx;

extern "C" {

size_t mbx_wah_groups_of(uint32_t nrows) { return ((size_t)nrows + 30) / 31; }

int mbx_wah_select(mbx_ctx *ctx, const mbx_colref *src, uint32_t nrows,
                   uint64_t lo, uint64_t hi, uint32_t *groups) {
  if (!src || !groups) return fail(MBX_ERR_INVALID, "wah_select: null");
  int st = wah_check(ctx, groups, "wah_select");
  if (st) return st;
  st = check_src(*src, "wah_select");
  if (st) return st;
  const uint32_t ng = (uint32_t)mbx_wah_groups_of(nrows);
  if (!ng) return MBX_OK;
  DevSrc d;
  d.data = src->data;
  d.fk = src->fk;
  d.dimsz = src->dimsz;
  d.eb = src->elem_bytes;
  const dim3 grid(grid_of(ng, 256)), block(256);
  switch (d.eb) {
  case 1:
    hipLaunchKernelGGL(wah_select_kernel<uint8_t>, grid, block, 0, ctx->stream,
                       d, nrows, lo, hi, groups, ng);
    break;
  case 2:
    hipLaunchKernelGGL(wah_select_kernel<uint16_t>, grid, block, 0, ctx->stream,
                       d, nrows, lo, hi, groups, ng);
    break;
  case 4:
    hipLaunchKernelGGL(wah_select_kernel<uint32_t>, grid, block, 0, ctx->stream,
                       d, nrows, lo, hi, groups, ng);
    break;
  default:
    hipLaunchKernelGGL(wah_select_kernel<uint64_t>, grid, block, 0, ctx->stream,
                       d, nrows, lo, hi, groups, ng);
    break;
  }
  MBX_HIP(hipGetLastError());
  return MBX_OK;
}

int mbx_wah_groups(mbx_ctx *ctx, const uint32_t *bitmap, uint32_t nrows,
                   uint32_t *groups) {
  if (!bitmap || !groups) return fail(MBX_ERR_INVALID, "wah_groups: null");
  int st = wah_check(ctx, groups, "wah_groups");
  if (!st) st = wah_check(ctx, bitmap, "wah_groups");
  if (st) return st;
  const size_t ng = mbx_wah_groups_of(nrows);
  if (!ng) return MBX_OK;
  const unsigned grid =
      (unsigned)std::min<size_t>(ceil_div(ng, 256), (size_t)ctx->num_cus * 16);
  hipLaunchKernelGGL(wah_groups_kernel, dim3(grid), dim3(256), 0, ctx->stream,
                     bitmap, nrows, groups, (uint32_t)ng);
  MBX_HIP(hipGetLastError());
  return MBX_OK;
}

int mbx_wah_compress(mbx_ctx *ctx, const uint32_t *groups, size_t ngroups,
                     uint32_t *out, size_t *nout) {
  if (!groups || !out || !nout) return fail(MBX_ERR_INVALID, "wah_compress: null");
  int st = wah_check(ctx, groups, "wah_compress");
  if (!st) st = wah_check(ctx, out, "wah_compress");
  if (st) return st;
  *nout = 0;
  if (ngroups == 0) return MBX_OK;
  if (ngroups > kWahCnt)
    return fail(MBX_ERR_CAPACITY, "wah_compress: more groups than a fill can count");
  const uint32_t ng = (uint32_t)ngroups;
  const uint32_t ntiles = (uint32_t)ceil_div(ng, kWahTile);
  Scratch s(ctx);
  st = ctx->pool->alloc(((size_t)ntiles + 1) * 4 + (size_t)ng * 4 + 256, &s.p);
  if (st) return st;
  uint32_t *tiles = static_cast<uint32_t *>(s.p);
  uint32_t *ends = tiles + ((ntiles + 1 + 63) / 64) * 64;
  hipLaunchKernelGGL(wah_keep_count_kernel, dim3(ntiles), dim3(256), 0,
                     ctx->stream, groups, ng, tiles);
  hipLaunchKernelGGL(wah_tiles_scan_kernel, dim3(1), dim3(256), 0, ctx->stream,
                     tiles, ntiles);
  hipLaunchKernelGGL(wah_scatter_kernel, dim3(ntiles), dim3(256), 0,
                     ctx->stream, groups, ng, tiles, out, ends);
  hipLaunchKernelGGL(wah_fill_len_kernel,
                     dim3((unsigned)std::min<size_t>(ceil_div(ng, 256),
                                                     (size_t)ctx->num_cus * 16)),
                     dim3(256), 0, ctx->stream, out, ends, tiles + ntiles);
  MBX_HIP(hipGetLastError());
  uint32_t k = 0;
  MBX_HIP(hipMemcpyAsync(&k, tiles + ntiles, 4, hipMemcpyDeviceToHost,
                         ctx->stream));
  MBX_HIP(hipStreamSynchronize(ctx->stream));
  *nout = k;
  return MBX_OK;
}

// The count scan's tile sums for the context's stream: kept per stream and
// grown on demand (the old buffer is freed once the stream has drained), so
// the scan neither allocates nor waits in steady state.
static int wah_tiles_of(mbx_ctx *ctx, size_t bytes, uint32_t **out) {
  std::lock_guard<std::mutex> g(ctx->spread_mu);
  auto &slot = ctx->wah_tiles[ctx->stream];
  if (slot.second < bytes) {
    if (slot.first) {
      MBX_HIP(hipStreamSynchronize(ctx->stream));
      ctx->pool->free(slot.first);
      slot = {nullptr, 0};
    }
    const size_t want = std::max<size_t>(bytes, 64 << 10);
    void *p = nullptr;
    int st = ctx->pool->alloc(want, &p);
    if (st) return st;
    slot = {p, want};
  }
  *out = static_cast<uint32_t *>(slot.first);
  return MBX_OK;
}

size_t mbx_wah_skip_len(size_t ngroups) { return ceil_div(ngroups, kWahSkip); }

// A WAH bitmap argument: its arrays on the context's device, the sizes
// within the format's limits.
static int wah_ref_check(mbx_ctx *ctx, const mbx_wah &w, const char *what) {
  if (!w.excl || (!w.words && w.n) || (!w.skip && w.nskip))
    return fail(MBX_ERR_INVALID, std::string(what) + ": null");
  int st = wah_check(ctx, w.excl, what);
  if (!st && w.n) st = wah_check(ctx, w.words, what);
  if (!st && w.nskip) st = wah_check(ctx, w.skip, what);
  if (st) return st;
  if (w.n > kWahCnt || w.nskip > kWahCnt)
    return fail(MBX_ERR_CAPACITY, std::string(what) + ": too many words");
  return MBX_OK;
}

int mbx_wah_scan(mbx_ctx *ctx, const uint32_t *wah, size_t n, uint32_t *excl) {
  return mbx_wah_scan_skip(ctx, wah, n, excl, nullptr, 0);
}

int mbx_wah_scan_skip(mbx_ctx *ctx, const uint32_t *wah, size_t n,
                      uint32_t *excl, uint32_t *skip, size_t nskip) {
  if ((!wah && n) || !excl || (!skip && nskip))
    return fail(MBX_ERR_INVALID, "wah_scan: null");
  int st = wah_check(ctx, excl, "wah_scan");
  if (!st && n) st = wah_check(ctx, wah, "wah_scan");
  if (!st && nskip) st = wah_check(ctx, skip, "wah_scan");
  if (st) return st;
  if (n > kWahCnt || nskip > kWahCnt)
    return fail(MBX_ERR_CAPACITY, "wah_scan: too many words");
  if (!nskip) skip = nullptr;
  if (n == 0) return launch_fill_u32(ctx->stream, excl, 0, 1);
  const uint32_t ntiles = (uint32_t)ceil_div(n, kWahTile);
  uint32_t *tiles = nullptr;
  st = wah_tiles_of(ctx, ((size_t)ntiles + 1) * 4, &tiles);
  if (st) return st;
  hipLaunchKernelGGL(wah_count_sum_kernel, dim3(ntiles), dim3(256), 0,
                     ctx->stream, wah, (uint32_t)n, tiles);
  if (ntiles <= kWahSelfPrefixTiles) {
    hipLaunchKernelGGL(wah_count_scan_kernel<true>, dim3(ntiles), dim3(256), 0,
                       ctx->stream, wah, (uint32_t)n, tiles, excl, skip,
                       (uint32_t)nskip);
  } else {
    hipLaunchKernelGGL(wah_tiles_scan_kernel, dim3(1), dim3(256), 0,
                       ctx->stream, tiles, ntiles);
    hipLaunchKernelGGL(wah_count_scan_kernel<false>, dim3(ntiles), dim3(256),
                       0, ctx->stream, wah, (uint32_t)n, tiles, excl, skip,
                       (uint32_t)nskip);
  }
  MBX_HIP(hipGetLastError());
  return MBX_OK;
}

int mbx_wah_and_no1(mbx_ctx *ctx, const uint32_t *up, const uint32_t *up_excl,
                    size_t nup, const uint32_t *down,
                    const uint32_t *down_excl, size_t ndown, uint32_t *out) {
  if (!up_excl || !down_excl || !out || (!up && nup) || (!down && ndown))
    return fail(MBX_ERR_INVALID, "wah_and_no1: null");
  for (const void *p : {(const void *)up, (const void *)up_excl,
                        (const void *)down, (const void *)down_excl,
                        (const void *)out}) {
    int st = wah_check(ctx, p, "wah_and_no1");
    if (st) return st;
  }
  if (nup > kWahCnt || ndown > kWahCnt)
    return fail(MBX_ERR_CAPACITY, "wah_and_no1: too many words");
  if (ndown == 0) return MBX_OK;
  hipLaunchKernelGGL(wah_and_no1_kernel,
                     dim3(grid_of(ceil_div(ndown, kWahSpan), 4)), dim3(256), 0,
                     ctx->stream, up, up_excl, (uint32_t)nup, down, down_excl,
                     (uint32_t)ndown, out);
  MBX_HIP(hipGetLastError());
  return MBX_OK;
}

int mbx_wah_and_no1_skip(mbx_ctx *ctx, const mbx_wah *up, const mbx_wah *down,
                         uint32_t *out) {
  if (!up || !down || !out) return fail(MBX_ERR_INVALID, "wah_and_no1: null");
  int st = wah_ref_check(ctx, *up, "wah_and_no1");
  if (!st) st = wah_ref_check(ctx, *down, "wah_and_no1");
  if (!st) st = wah_check(ctx, out, "wah_and_no1");
  if (st) return st;
  if (down->n == 0) return MBX_OK;
  if (!down->skip || !down->nskip)   // no down windows: the word-mapped kernel
    return mbx_wah_and_no1(ctx, up->words, up->excl, up->n, down->words,
                           down->excl, down->n, out);
  hipLaunchKernelGGL(wah_and_no1_win_kernel, dim3((unsigned)ctx->num_cus * 4),
                     dim3(256), 0, ctx->stream, up->words, up->excl,
                     up->nskip ? up->skip : nullptr, (uint32_t)up->nskip,
                     (uint32_t)up->n, down->words, down->excl, down->skip,
                     (uint32_t)down->nskip, (uint32_t)down->n, out);
  MBX_HIP(hipGetLastError());
  return MBX_OK;
}

int mbx_wah_decompress(mbx_ctx *ctx, const uint32_t *wah, const uint32_t *excl,
                       size_t n, uint32_t *groups, size_t ngroups) {
  const mbx_wah w{wah, excl, nullptr, n, 0};
  return mbx_wah_decompress_skip(ctx, &w, groups, ngroups);
}

int mbx_wah_decompress_skip(mbx_ctx *ctx, const mbx_wah *w, uint32_t *groups,
                            size_t ngroups) {
  if (!w) return fail(MBX_ERR_INVALID, "wah_decompress: null");
  int st = wah_ref_check(ctx, *w, "wah_decompress");
  if (st) return st;
  if (!groups) return fail(MBX_ERR_INVALID, "wah_decompress: null");
  st = wah_check(ctx, groups, "wah_decompress");
  if (st) return st;
  if (ngroups > 0xffffffffull)
    return fail(MBX_ERR_CAPACITY, "wah_decompress: too many words or groups");
  if (ngroups == 0) return MBX_OK;
  hipLaunchKernelGGL(wah_decompress_kernel,
                     dim3(grid_of(ceil_div(ngroups, kWahSpan), 4)), dim3(256), 0,
                     ctx->stream, w->words, w->excl,
                     w->nskip ? w->skip : nullptr, (uint32_t)w->nskip,
                     (uint32_t)w->n, groups, (uint32_t)ngroups);
  MBX_HIP(hipGetLastError());
  return MBX_OK;
}

int mbx_wah_longest_1fill(mbx_ctx *ctx, const uint32_t *wah, size_t n,
                          uint64_t *longest) {
  if (!longest || (!wah && n)) return fail(MBX_ERR_INVALID, "wah_longest_1fill: null");
  int st = n ? wah_check(ctx, wah, "wah_longest_1fill")
             : (ctx ? MBX_OK : fail(MBX_ERR_INVALID, "wah_longest_1fill: null ctx"));
  if (st) return st;
  *longest = 0;
  if (n == 0) return MBX_OK;
  if (n > kWahCnt) return fail(MBX_ERR_CAPACITY, "wah_longest_1fill: too many words");
  Scratch s(ctx);
  st = ctx->pool->alloc(256, &s.p);
  if (st) return st;
  st = launch_fill_u32(ctx->stream, s.p, 0, 2);
  if (st) return st;
  hipLaunchKernelGGL(wah_longest_1fill_kernel,
                     dim3((unsigned)std::min<size_t>(ceil_div(n, 256),
                                                     (size_t)ctx->num_cus * 16)),
                     dim3(256), 0, ctx->stream, wah, (uint32_t)n,
                     static_cast<unsigned long long *>(s.p));
  MBX_HIP(hipGetLastError());
  unsigned long long v = 0;
  MBX_HIP(hipMemcpyAsync(&v, s.p, 8, hipMemcpyDeviceToHost, ctx->stream));
  MBX_HIP(hipStreamSynchronize(ctx->stream));
  *longest = v;
  return MBX_OK;
}

}  // extern "C"
