// bins.hip -- bitmap creation for gfx950.
//
// Replaces bmpcreate (primitive.cuh:88-118) and its host wrappers create_bin /
// create_select / create_join (bmpidx_create.cu:14-136).
//
// One wave owns a block of 1024 consecutive rows (16 groups of 64).  Lane l
// reads row 64g+l of every group (coalesced 4-byte loads), so a wave64
// ballot of a per-row predicate is exactly one 64-row bitmap word: the two
// reference uint32 words of those rows, in little-endian order.  The wave
// then holds 16 words = 128 bytes per bin, lane j the line of bin j.  With
// more than 8 bins the lines go out through LDS, transposed so that each
// 16-byte store instruction writes 8 whole lines (8 lanes per line) rather
// than 16-byte pieces of 64 lines: 8 times fewer write requests to L2.  With
// up to 8 bins each lane stores its own line and the kernel keeps its LDS
// (and occupancy) for the gathers of the foreign-key path.
//
// Consecutive bins (maxs[j] == mins[j+1], the reference's bin tables,
// ssb/bmpidx-create.cu:8-30) take the radix path: each row gets a bin index,
// the wave ballots the index bits once (<= 7 ballots for 64 bins), and lane j
// rebuilds bin j's word as AND over bits of (bit set ? ballot : ~ballot).
// The cost per 64 rows is independent of the bin count up to 64 bins,
// instead of one ballot per bin.  Any other bin set takes the generic path:
// one ballot per bin, as the reference does.
#include <cstring>
#include <type_traits>

#include "devfuncs.hpp"

namespace mbx {

enum BinMode : uint32_t {
  kGeneric = 0,   // arbitrary [mins[j], maxs[j]) per bin
  kUnit = 1,      // consecutive bins of width 1: idx = v - b0
  kEqual = 2,     // consecutive bins of equal width
  kSorted = 3     // consecutive bins of any widths (binary search in LDS)
};

constexpr uint32_t kMaxSortedBounds = 2048;
constexpr uint32_t kNone = 0xffffffffu;

struct BinArgs {
  DevSrc src;
  uint32_t nrows;
  uint32_t nblocks;      // 1024-row blocks to write
  uint32_t nb;           // bins
  uint32_t mode;
  uint64_t b0, width;    // kUnit / kEqual
  double inv_width;      // kEqual
  uint32_t search_top;   // kSorted: largest power of two <= nb+1
  uint32_t tab16;        // table bytes / 16 (table copied into LDS)
  const uint64_t *mins;  // device, nb (kGeneric) / bounds nb+1 (kSorted)
  const uint64_t *maxs;  // device, nb (kGeneric)
  const void *tab;       // device, dimsz+1 bin indexes (foreign-key radix path)
  uint32_t *out;
  size_t stride_words;
};

// Value of row `row` as create_bin's fetch (bmpidx_create.cu:17-35): a
// foreign key past the dimension reads as INT64_MAX.
__device__ __forceinline__ uint64_t bin_value(const DevSrc &s, size_t row) {
  if (s.fk == nullptr) return ld_elem(s.data, s.eb, row, true);
  const uint32_t k = ld_stream(s.fk + row);
  if (k >= s.dimsz) return (uint64_t)INT64_MAX;
  return ld_elem(s.data, s.eb, k, false);
}

__device__ __forceinline__ uint32_t bin_index(const BinArgs &a, uint64_t v,
                                              const uint64_t *s_bounds) {
  if (a.mode == kUnit) {
    if (v < a.b0) return kNone;
    const uint64_t d = v - a.b0;
    return d < a.nb ? (uint32_t)d : kNone;
  }
  if (a.mode == kEqual) {
    if (v < a.b0) return kNone;
    const uint64_t d = v - a.b0;
    if (d >= (uint64_t)a.nb * a.width) return kNone;
    uint64_t q = (uint64_t)((double)d * a.inv_width);
    if (q * a.width > d) --q;
    else if ((q + 1) * a.width <= d) ++q;
    return (uint32_t)q;
  }
  // kSorted: c = #{bounds <= v}; bin = c - 1 when 1 <= c <= nb
  uint32_t pos = 0;
  for (uint32_t step = a.search_top; step > 0; step >>= 1) {
    const uint32_t np = pos + step;
    if (np <= a.nb + 1 && s_bounds[np - 1] <= v) pos = np;
  }
  return (pos >= 1 && pos <= a.nb) ? pos - 1 : kNone;
}

// Bin index of each dimension row, for a foreign-key column on the radix
// path: entry k < dimsz holds row k's bin, entry dimsz the bin of INT64_MAX
// (what a key past the dimension reads as), ~0 for no bin.  The fact pass
// then reads one 1- or 2-byte entry per row: the table of a 1.4M-row
// dimension (1.4 or 2.8 MB) stays in an XCD's 4 MB L2, where its 4-byte
// attribute column (5.6 MB) did not.
template <bool SORTED, typename T>
__global__ __launch_bounds__(256) void dim_index_kernel(BinArgs a, T *tab) {
  __shared__ uint64_t s_bounds[SORTED ? kMaxSortedBounds : 1];
  if (SORTED) {
    for (uint32_t i = threadIdx.x; i <= a.nb; i += blockDim.x)
      s_bounds[i] = a.mins[i];
    __syncthreads();
  }
  const size_t n = (size_t)a.src.dimsz + 1;
  for (size_t k = (size_t)blockIdx.x * blockDim.x + threadIdx.x; k < n;
       k += (size_t)gridDim.x * blockDim.x) {
    const uint64_t v = k < a.src.dimsz ? ld_elem(a.src.data, a.src.eb, k, true)
                                       : (uint64_t)INT64_MAX;
    const uint32_t i = bin_index(a, v, s_bounds);
    tab[k] = i == kNone ? (T)~(T)0 : (T)i;
  }
}

typedef uint64_t u64x2 __attribute__((ext_vector_type(2)));

// Writes the wave's lines: lane j holds the 16 words of bin cbase + j for
// 1024-row block blk (j < n).  Through the wave's 8 KB of LDS, stored with
// an XOR swizzle (16-byte part i of line j at slot j*8 + (i ^ (j & 7))), so
// that both the row-wise writes and the column-wise reads touch distinct
// banks; then lane l stores part l % 8 of line 8t + l / 8, t = 0..7.
__device__ __forceinline__ void store_lines(const BinArgs &a, u64x2 *tr,
                                            const uint64_t *w, uint32_t cbase,
                                            uint32_t n, uint32_t blk) {
  const uint32_t lane = lane_id();
#pragma unroll
  for (int i = 0; i < 8; ++i) {
    u64x2 x;
    x.x = w[2 * i];
    x.y = w[2 * i + 1];
    tr[lane * 8 + ((uint32_t)i ^ (lane & 7u))] = x;
  }
  wave_lds_sync();
  const uint32_t part = lane & 7u;
#pragma unroll 1
  for (uint32_t t = 0; t < 8; ++t) {
    const uint32_t line = t * 8 + (lane >> 3);
    if (line < n) {
      const u64x2 x = tr[line * 8 + (part ^ (line & 7u))];
      u64x2 *dst = reinterpret_cast<u64x2 *>(
          a.out + (size_t)(cbase + line) * a.stride_words + (size_t)blk * 32);
      // streaming store: the lines are not read back soon, and under
      // normal stores their 15 GB (1,000 bins) evicted the foreign-key bin
      // table from L2 (2.6 GB of extra HBM fetch)
      __builtin_nontemporal_store(x, dst + part);
    }
  }
  wave_lds_sync();   // the next chunk rewrites tr
}

// One lane writes a full 128-byte line (16 words of 64 rows).
__device__ __forceinline__ void store_line(uint32_t *dst, const uint64_t *w) {
  u64x2 *d = reinterpret_cast<u64x2 *>(dst);
#pragma unroll
  for (int i = 0; i < 8; ++i) {
    u64x2 x;
    x.x = w[2 * i];
    x.y = w[2 * i + 1];
    d[i] = x;   // plain stores: L2 merges the 8 parts into one full-line write
  }
}

// Transpose buffer of the TR kernels (more than 8 bins): 8 KB per wave, as
// dynamic LDS of kTransposeLds bytes.  The kernels for up to 8 bins have
// neither the buffer nor its registers (76 against 83 VGPRs: 6 waves per
// SIMD against 5 for the latency-bound foreign-key gathers).
extern __shared__ u64x2 s_tr_dyn[];
constexpr size_t kTransposeLds = 4 * 64 * 8 * sizeof(u64x2);

// Largest bin-index table copied into each workgroup's LDS (a 40,000-row
// dimension's uint8 table): random byte lookups there cost no L2 request.
constexpr size_t kLdsTableMax = 64 * 1024;

// TAB: 0 = values through bin_value; 1 / 2 = bin indexes from a.tab of
// uint8 / uint16 entries, looked up by foreign key (radix path only).
// LT: the table is first copied into dynamic LDS, after the transpose buffer.
template <bool RADIX, bool SORTED, bool TR, int TAB = 0, bool LT = false>
__global__ __launch_bounds__(256) void bins_kernel(BinArgs a) {
  static_assert(TAB == 0 || (RADIX && !SORTED), "table path is radix-only");
  static_assert(!LT || (TAB != 0 && !TR), "LDS table: table path, no transpose");
  __shared__ uint64_t s_bounds[SORTED ? kMaxSortedBounds : 1];
  if (SORTED) {
    for (uint32_t i = threadIdx.x; i <= a.nb; i += blockDim.x)
      s_bounds[i] = a.mins[i];
    __syncthreads();
  }
  u64x2 *s_tab = s_tr_dyn + (TR ? kTransposeLds / sizeof(u64x2) : 0);
  if constexpr (LT) {
    const u64x2 *src = static_cast<const u64x2 *>(a.tab);
    for (uint32_t i = threadIdx.x; i < a.tab16; i += blockDim.x) s_tab[i] = src[i];
    __syncthreads();
  }
  const uint32_t lane = lane_id();
  const uint32_t wave = blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
  const uint32_t nwaves = gridDim.x * (blockDim.x >> 6);
  const uint32_t nchunks = (a.nb + 63) / 64;

  for (uint32_t blk = wave; blk < a.nblocks; blk += nwaves) {
    const size_t row0 = (size_t)blk * 1024;
    uint64_t v[16];
    uint32_t ok = 0;
    uint32_t idx[16];
    // Every load is issued unconditionally (rows past N read row N-1) and
    // the tail is masked afterwards: a load under `if (row < N)` becomes a
    // branch whose join waits for it, which serialised the 16 key loads.
    const size_t last = (size_t)a.nrows - 1;
    if constexpr (TAB != 0) {
      using T = typename std::conditional<TAB == 1, uint8_t, uint16_t>::type;
      const T *tab = LT ? reinterpret_cast<const T *>(s_tab)
                        : static_cast<const T *>(a.tab);
      uint32_t key[16];
#pragma unroll
      for (int g = 0; g < 16; ++g)
        key[g] = ld_stream(a.src.fk + min(row0 + 64 * g + lane, last));
#pragma unroll
      for (int g = 0; g < 16; ++g) {
        const uint32_t t = tab[min(key[g], a.src.dimsz)];
        const bool live = row0 + 64 * g + lane <= last;
        idx[g] = (live && t != (uint32_t)(T)~(T)0) ? t : kNone;
      }
    } else {
      if (a.src.fk == nullptr) {
#pragma unroll
        for (int g = 0; g < 16; ++g)
          v[g] = ld_elem(a.src.data, a.src.eb, min(row0 + 64 * g + lane, last),
                         true);
      } else {
        // a key past the dimension reads as INT64_MAX (create_bin's fetch,
        // bmpidx_create.cu:17-35); its load reads row dimsz - 1
        uint32_t key[16];
#pragma unroll
        for (int g = 0; g < 16; ++g)
          key[g] = ld_stream(a.src.fk + min(row0 + 64 * g + lane, last));
        const uint32_t dl = a.src.dimsz ? a.src.dimsz - 1 : 0;
#pragma unroll
        for (int g = 0; g < 16; ++g) {
          const uint64_t x = a.src.dimsz
                                 ? ld_elem(a.src.data, a.src.eb,
                                           min(key[g], dl), false)
                                 : 0;
          v[g] = key[g] < a.src.dimsz ? x : (uint64_t)INT64_MAX;
        }
      }
#pragma unroll
      for (int g = 0; g < 16; ++g)
        if (row0 + 64 * g + lane <= last) ok |= 1u << g;
      if (RADIX) {
#pragma unroll
        for (int g = 0; g < 16; ++g)
          idx[g] = (ok >> g & 1u) ? bin_index(a, v[g], s_bounds) : kNone;
      }
    }
    for (uint32_t c = 0; c < nchunks; ++c) {
      const uint32_t cbase = c * 64;
      const uint32_t inchunk = min(64u, a.nb - cbase);
      uint64_t w[16];
      if (RADIX) {
        const uint32_t nbits = inchunk <= 1 ? 0 : 32 - __clz(inchunk - 1);
#pragma unroll
        for (int g = 0; g < 16; ++g) {
          const uint32_t rel = idx[g] - cbase;   // kNone wraps out of range
          const bool inb = rel < inchunk;
          uint64_t acc = __ballot(inb);
          for (uint32_t k = 0; k < nbits; ++k) {
            const uint64_t bk = __ballot(inb && ((rel >> k) & 1u));
            const uint64_t flip = ((lane >> k) & 1u) ? 0ull : ~0ull;
            acc &= bk ^ flip;
          }
          w[g] = acc;
        }
      } else {
#pragma unroll
        for (int g = 0; g < 16; ++g) w[g] = 0;
        for (uint32_t j = 0; j < inchunk; ++j) {
          const uint64_t lo = a.mins[cbase + j], hi = a.maxs[cbase + j];
#pragma unroll
          for (int g = 0; g < 16; ++g) {
            const bool in = ((ok >> g) & 1u) && v[g] >= lo && v[g] < hi;
            const uint64_t m = __ballot(in);
            if (lane == j) w[g] = m;
          }
        }
      }
      if (TR)
        store_lines(a, s_tr_dyn + (threadIdx.x >> 6) * 64 * 8, w, cbase,
                    inchunk, blk);
      else if (lane < inchunk)
        store_line(a.out + (size_t)(cbase + lane) * a.stride_words +
                       (size_t)blk * 32,
                   w);
    }
  }
}

// ---------------------------------------------- several attributes, one FK
// The attributes of one dimension read through one foreign-key column
// (c_region, c_nation and c_city through lo_custkey): one pass over the keys
// for all of them.  A table holds, per dimension row, the bin index of every
// attribute packed in one 32-bit entry (8 bits for fewer than 255 bins, else
// 16), so each fact row costs one key load and one table lookup however
// many attributes it feeds -- where separate passes read the keys and look
// the table up once per attribute.
constexpr int kMaxSets = 4;
struct MultiSet {
  uint32_t nb, shift, bits;   // bins; field of the entry
  uint32_t *out;              // nb bitmaps, stride_words apart
};
struct MultiArgs {
  const uint32_t *fk;
  uint32_t dimsz, nrows, nblocks, nsets;
  size_t stride_words;
  const uint32_t *tab;        // dimsz + 1 packed entries
  MultiSet set[kMaxSets];
};

template <bool SORTED>
__global__ __launch_bounds__(256) void dim_pack_kernel(BinArgs a, uint32_t shift,
                                                       uint32_t bits,
                                                       uint32_t *tab) {
  __shared__ uint64_t s_bounds[SORTED ? kMaxSortedBounds : 1];
  if (SORTED) {
    for (uint32_t i = threadIdx.x; i <= a.nb; i += blockDim.x)
      s_bounds[i] = a.mins[i];
    __syncthreads();
  }
  const uint32_t none = (1u << bits) - 1u;
  const size_t n = (size_t)a.src.dimsz + 1;
  for (size_t k = (size_t)blockIdx.x * blockDim.x + threadIdx.x; k < n;
       k += (size_t)gridDim.x * blockDim.x) {
    const uint64_t v = k < a.src.dimsz ? ld_elem(a.src.data, a.src.eb, k, true)
                                       : (uint64_t)INT64_MAX;
    const uint32_t i = bin_index(a, v, s_bounds);
    const uint32_t f = i == kNone ? none : i;
    // each attribute owns its bits: the entries start all zero and the
    // attributes' kernels run one after the other on the stream
    tab[k] |= f << shift;
  }
}

// Lines of one attribute with more than 8 bins go out through a per-wave
// LDS image of up to kScatterBins lines (128 bytes = 1024 rows each): the
// lane of each row ORs the row's bit into its bin's line (ds_or_b32), then
// the wave streams the lines out, 8 lanes per line.  The cost per 1024 rows
// is one LDS atomic per row plus a zero and a read of each line, against
// 1 + log2(64) ballots and four 64-bit ALU operations per ballot for every
// 64 rows and every 64 bins of the ballot transpose.
// Word w of line b sits at b*32 + (w ^ (b & 31)): the 32 lanes of one
// ds_or_b32, which share w, reach 32 banks when their bins differ mod 32.
constexpr uint32_t kScatterBins = 64;
constexpr size_t kScatterLds = 4 * kScatterBins * 128;

typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ void scatter_lines(const MultiSet &ms, uint32_t none,
                                              const uint32_t *ent,
                                              const bool *live, uint32_t *img,
                                              size_t stride_words, uint32_t blk) {
  const uint32_t lane = lane_id();
  u32x4 *img4 = reinterpret_cast<u32x4 *>(img);
  const uint32_t bit = 1u << (lane & 31u);
  for (uint32_t cbase = 0; cbase < ms.nb; cbase += kScatterBins) {
    const uint32_t inchunk = min(kScatterBins, ms.nb - cbase);
    const u32x4 zero = {0u, 0u, 0u, 0u};
    for (uint32_t s = lane; s < inchunk * 8; s += 64) img4[s] = zero;
    wave_lds_sync();
#pragma unroll
    for (int g = 0; g < 16; ++g) {
      const uint32_t f = (ent[g] >> ms.shift) & none;
      const uint32_t rel = f - cbase;   // none and other chunks: out of range
      if (live[g] && f != none && rel < inchunk) {
        const uint32_t w = 2u * (uint32_t)g + (lane >> 5);
        atomicOr(img + rel * 32 + (w ^ (rel & 31u)), bit);
      }
    }
    wave_lds_sync();
    for (uint32_t s = lane; s < inchunk * 8; s += 64) {
      const uint32_t line = s >> 3, part = s & 7u, sw = line & 31u;
      u32x4 v = img4[line * 8 + (part ^ (sw >> 2))];
      // logical word 4*part + i sits at position i ^ (sw & 3)
      if (sw & 1u) v = v.yxwz;
      if (sw & 2u) v = v.zwxy;
      u32x4 *dst = reinterpret_cast<u32x4 *>(
          ms.out + (size_t)(cbase + line) * stride_words + (size_t)blk * 32);
      __builtin_nontemporal_store(v, dst + part);
    }
    wave_lds_sync();   // the next chunk zeroes the image
  }
}

// SCATTER: some set has more than 8 bins (the launch gives each wave
// kScatterBins lines of dynamic LDS).
template <bool SCATTER>
__global__ __launch_bounds__(256) void bins_multi_kernel(MultiArgs a) {
  const uint32_t lane = lane_id();
  const uint32_t wave = blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
  const uint32_t nwaves = gridDim.x * (blockDim.x >> 6);
  uint32_t *img = reinterpret_cast<uint32_t *>(s_tr_dyn) +
                  (threadIdx.x >> 6) * kScatterBins * 32;
  // Loads are issued unconditionally, the tail masked after (as
  // bins_kernel), and the keys of the wave's next block are loaded while
  // this block's lines go out: a block waits for one round trip (its table
  // entries), not two.
  const size_t last = (size_t)a.nrows - 1;
  uint32_t key[16];
  if (wave < a.nblocks) {
#pragma unroll
    for (int g = 0; g < 16; ++g)
      key[g] = ld_stream(a.fk + min((size_t)wave * 1024 + 64 * g + lane, last));
  }
  for (uint32_t blk = wave; blk < a.nblocks; blk += nwaves) {
    const size_t row0 = (size_t)blk * 1024;
    uint32_t ent[16];
    bool live[16];
#pragma unroll
    for (int g = 0; g < 16; ++g) {
      live[g] = row0 + 64 * g + lane <= last;
      // a key past the dimension reads entry dimsz (the bins of INT64_MAX)
      ent[g] = a.tab[min(key[g], a.dimsz)];
    }
    if (blk + nwaves < a.nblocks) {
      const size_t next0 = (size_t)(blk + nwaves) * 1024;
#pragma unroll
      for (int g = 0; g < 16; ++g)
        key[g] = ld_stream(a.fk + min(next0 + 64 * g + lane, last));
    }
    for (uint32_t si = 0; si < a.nsets; ++si) {
      const MultiSet &ms = a.set[si];
      const uint32_t none = (1u << ms.bits) - 1u;
      if (SCATTER && ms.nb > 8) {
        scatter_lines(ms, none, ent, live, img, a.stride_words, blk);
        continue;
      }
      // a few bins (every set, without SCATTER): radix ballots, each lane
      // its own line
      const uint32_t nbits = ms.nb <= 1 ? 0 : 32 - __clz(ms.nb - 1);
      uint64_t w[16];
#pragma unroll
      for (int g = 0; g < 16; ++g) {
        const uint32_t f = (ent[g] >> ms.shift) & none;
        const bool inb = live[g] && f < ms.nb;
        uint64_t acc = __ballot(inb);
        for (uint32_t k = 0; k < nbits; ++k) {
          const uint64_t bk = __ballot(inb && ((f >> k) & 1u));
          const uint64_t flip = ((lane >> k) & 1u) ? 0ull : ~0ull;
          acc &= bk ^ flip;
        }
        w[g] = acc;
      }
      if (lane < ms.nb)
        store_line(ms.out + (size_t)lane * a.stride_words + (size_t)blk * 32, w);
    }
  }
}

static int bin_mode(const uint64_t *mins, const uint64_t *maxs, size_t nb,
                    BinArgs *a) {
  bool consecutive = true, equal = true;
  for (size_t j = 0; j < nb; ++j) {
    if (!(mins[j] < maxs[j])) consecutive = false;
    if (j + 1 < nb && maxs[j] != mins[j + 1]) consecutive = false;
    if (maxs[j] - mins[j] != maxs[0] - mins[0]) equal = false;
  }
  const uint64_t w0 = maxs[0] - mins[0];
  if (consecutive && equal && w0 == 1) {
    a->mode = kUnit;
    a->b0 = mins[0];
  } else if (consecutive && equal && w0 != 0 && w0 < (1ull << 40) &&
             (double)nb * (double)w0 < 4.0e15) {
    a->mode = kEqual;
    a->b0 = mins[0];
    a->width = w0;
    a->inv_width = 1.0 / (double)w0;
  } else if (consecutive && nb + 1 <= kMaxSortedBounds) {
    a->mode = kSorted;
    uint32_t top = 1;
    while (top * 2 <= nb + 1) top *= 2;
    a->search_top = top;
  } else {
    a->mode = kGeneric;
  }
  return a->mode;
}

// Classifies the bin set and launches the kernel.  d_tmp (nb+1 or 2nb
// uint64, device) is used for bounds the kernel must read from memory.
static int launch_bins(mbx_ctx *ctx, const mbx_colref *src, uint32_t nrows,
                       const uint64_t *mins, const uint64_t *maxs, size_t nb,
                       uint32_t *slab, size_t stride_words) {
  int st = check_src(*src, "create_bin");
  if (st) return st;
  if (nrows == 0) return fail(MBX_ERR_INVALID, "create_bin: 0 rows");
  if (src->fk && reinterpret_cast<uintptr_t>(src->fk) % 4)
    return fail(MBX_ERR_ALIGN, "create_bin: misaligned fk");
  if (nb == 0 || nb > 0x7fffffffu)
    return fail(MBX_ERR_INVALID, "create_bin: bad bin count");
  if (stride_words % 64 != 0 || stride_words < mbx_bitmap_words(nrows))
    return fail(MBX_ERR_INVALID,
                "create_bin: stride_words must be a multiple of 64 and "
                ">= ceil(N/32)");
  if (reinterpret_cast<uintptr_t>(slab) % 16)
    return fail(MBX_ERR_ALIGN, "create_bin: output not 16-byte aligned");

  BinArgs a;
  std::memset(&a, 0, sizeof(a));
  a.src.data = src->data;
  a.src.fk = src->fk;
  a.src.dimsz = src->dimsz;
  a.src.eb = src->elem_bytes;
  a.nrows = nrows;
  a.nblocks = (uint32_t)(stride_words / 32);
  a.nb = (uint32_t)nb;
  a.out = slab;
  a.stride_words = stride_words;

  bool consecutive = true, equal = true;
  for (size_t j = 0; j < nb; ++j) {
    if (!(mins[j] < maxs[j])) consecutive = false;
    if (j + 1 < nb && maxs[j] != mins[j + 1]) consecutive = false;
    if (maxs[j] - mins[j] != maxs[0] - mins[0]) equal = false;
  }
  const uint64_t w0 = maxs[0] - mins[0];
  size_t tmp_words = 0;
  if (consecutive && equal && w0 == 1) {
    a.mode = kUnit;
    a.b0 = mins[0];
  } else if (consecutive && equal && w0 != 0 && w0 < (1ull << 40) &&
             (double)nb * (double)w0 < 4.0e15) {
    a.mode = kEqual;
    a.b0 = mins[0];
    a.width = w0;
    a.inv_width = 1.0 / (double)w0;
  } else if (consecutive && nb + 1 <= kMaxSortedBounds) {
    a.mode = kSorted;
    tmp_words = nb + 1;
    uint32_t top = 1;
    while (top * 2 <= nb + 1) top *= 2;
    a.search_top = top;
  } else {
    a.mode = kGeneric;
    tmp_words = 2 * nb;
  }

  void *tmp = nullptr;
  if (tmp_words) {
    st = ctx->pool->alloc(tmp_words * 8, &tmp);
    if (st) return st;
    std::vector<uint64_t> h(tmp_words);
    if (a.mode == kSorted) {
      for (size_t j = 0; j < nb; ++j) h[j] = mins[j];
      h[nb] = maxs[nb - 1];
      a.mins = static_cast<uint64_t *>(tmp);
    } else {
      for (size_t j = 0; j < nb; ++j) h[j] = mins[j], h[nb + j] = maxs[j];
      a.mins = static_cast<uint64_t *>(tmp);
      a.maxs = static_cast<uint64_t *>(tmp) + nb;
    }
    hipError_t e = hipMemcpyAsync(tmp, h.data(), tmp_words * 8,
                                  hipMemcpyHostToDevice, ctx->stream);
    if (e != hipSuccess) {
      ctx->pool->free(tmp);
      return hip_fail(e, "create_bin: bounds upload");
    }
  }
  // Foreign-key radix path: a bin-index table over the dimension first
  // (without room for it in the pool, the values are read directly).
  int tab = 0;
  void *d_tab = nullptr;
  if (a.src.fk && a.mode != kGeneric && nb < 0xffff &&
      a.src.dimsz < 0x80000000u) {
    const int eb = nb < 0xff ? 1 : 2;
    const size_t tab_bytes = ((((size_t)a.src.dimsz + 1) * eb) + 15) & ~(size_t)15;
    if (ctx->pool->alloc(tab_bytes, &d_tab) != MBX_OK) d_tab = nullptr;
    else tab = eb;
    a.tab16 = (uint32_t)(tab_bytes / 16);
  }
  if (tab) {
    a.tab = d_tab;
    size_t tg = ceil_div((size_t)a.src.dimsz + 1, 256);
    if (tg > (size_t)ctx->num_cus * 8) tg = (size_t)ctx->num_cus * 8;
    const dim3 g((unsigned)tg), b(256);
    if (a.mode == kSorted) {
      if (tab == 1)
        hipLaunchKernelGGL((dim_index_kernel<true, uint8_t>), g, b, 0, ctx->stream, a,
                           static_cast<uint8_t *>(d_tab));
      else
        hipLaunchKernelGGL((dim_index_kernel<true, uint16_t>), g, b, 0, ctx->stream, a,
                           static_cast<uint16_t *>(d_tab));
    } else {
      if (tab == 1)
        hipLaunchKernelGGL((dim_index_kernel<false, uint8_t>), g, b, 0, ctx->stream, a,
                           static_cast<uint8_t *>(d_tab));
      else
        hipLaunchKernelGGL((dim_index_kernel<false, uint16_t>), g, b, 0, ctx->stream, a,
                           static_cast<uint16_t *>(d_tab));
    }
  }
  const bool tr = nb > 8;
  size_t lds = tr ? kTransposeLds : 0;
  size_t cap = (size_t)ctx->num_cus * 16;
  // A small table goes into LDS when the workgroup's LDS stays within the
  // device's per-block limit; the grid then holds only the resident
  // workgroups, so each copies the table once.  Not beside the transpose
  // buffer: at 72 KB per workgroup the write-bound TR kernels lose half
  // their waves (40,000-row dimension, 250 bins: 1.43 -> 1.97 ms).
  const bool lt = tab && !tr && (size_t)a.tab16 * 16 <= kLdsTableMax &&
                  (size_t)a.tab16 * 16 <= ctx->lds_per_block;
  if (lt) {
    lds = (size_t)a.tab16 * 16;
    size_t per_cu = ctx->lds_per_cu / lds;
    if (per_cu > 8) per_cu = 8;
    if (per_cu == 0) per_cu = 1;
    cap = (size_t)ctx->num_cus * per_cu;
  }
  size_t wgs = ceil_div(a.nblocks, 4);
  if (wgs > cap) wgs = cap;
  if (wgs == 0) wgs = 1;
  const dim3 g((unsigned)wgs), b(256);
  if (lt && tab == 1) {
    hipLaunchKernelGGL((bins_kernel<true, false, false, 1, true>), g, b, lds, ctx->stream, a);
  } else if (lt) {
    hipLaunchKernelGGL((bins_kernel<true, false, false, 2, true>), g, b, lds, ctx->stream, a);
  } else if (tab == 1) {
    if (tr) hipLaunchKernelGGL((bins_kernel<true, false, true, 1>), g, b, lds, ctx->stream, a);
    else hipLaunchKernelGGL((bins_kernel<true, false, false, 1>), g, b, lds, ctx->stream, a);
  } else if (tab == 2) {
    if (tr) hipLaunchKernelGGL((bins_kernel<true, false, true, 2>), g, b, lds, ctx->stream, a);
    else hipLaunchKernelGGL((bins_kernel<true, false, false, 2>), g, b, lds, ctx->stream, a);
  } else if (a.mode == kGeneric) {
    if (tr) hipLaunchKernelGGL((bins_kernel<false, false, true>), g, b, lds, ctx->stream, a);
    else hipLaunchKernelGGL((bins_kernel<false, false, false>), g, b, lds, ctx->stream, a);
  } else if (a.mode == kSorted) {
    if (tr) hipLaunchKernelGGL((bins_kernel<true, true, true>), g, b, lds, ctx->stream, a);
    else hipLaunchKernelGGL((bins_kernel<true, true, false>), g, b, lds, ctx->stream, a);
  } else {
    if (tr) hipLaunchKernelGGL((bins_kernel<true, false, true>), g, b, lds, ctx->stream, a);
    else hipLaunchKernelGGL((bins_kernel<true, false, false>), g, b, lds, ctx->stream, a);
  }
  hipError_t e = hipGetLastError();
  if (tmp || d_tab) {
    // the bounds and the table must outlive the kernels: wait first
    hipError_t e2 = hipStreamSynchronize(ctx->stream);
    if (tmp) ctx->pool->free(tmp);
    if (d_tab) ctx->pool->free(d_tab);
    if (e == hipSuccess) e = e2;
  }
  if (e != hipSuccess) return hip_fail(e, "create_bin: launch");
  return MBX_OK;
}

}  // namespace mbx

using namespace mbx;

extern "C" {

size_t mbx_bitmap_words(uint32_t nrows) { return ceil_div(nrows, 32); }

size_t mbx_bitmap_stride_words(uint32_t nrows) {
  return ceil_div(ceil_div(nrows, 32), 64) * 64;
}

int mbx_create_bin_into(mbx_ctx *ctx, const mbx_colref *src, uint32_t nrows,
                        const uint64_t *mins, const uint64_t *maxs,
                        size_t nbins, uint32_t *slab, size_t stride_words) {
  if (!ctx || !src || !mins || !maxs || !slab)
    return fail(MBX_ERR_INVALID, "create_bin_into: null argument");
  if (nrows == 0) return fail(MBX_ERR_INVALID, "create_bin_into: 0 rows");
  return launch_bins(ctx, src, nrows, mins, maxs, nbins, slab, stride_words);
}

int mbx_create_bin(mbx_ctx *ctx, const mbx_colref *src, uint32_t nrows,
                   const uint64_t *mins, const uint64_t *maxs, size_t nbins,
                   uint32_t **out, size_t *single_bytes) {
  if (!ctx || !src || !mins || !maxs || !out)
    return fail(MBX_ERR_INVALID, "create_bin: null argument");
  if (nrows == 0 || nbins == 0)
    return fail(MBX_ERR_INVALID, "create_bin: 0 rows or 0 bins");
  const size_t stride = mbx_bitmap_stride_words(nrows);
  void *slab = nullptr;
  int st = ctx->pool->alloc(stride * 4 * nbins, &slab);
  if (st) return st;
  st = launch_bins(ctx, src, nrows, mins, maxs, nbins,
                   static_cast<uint32_t *>(slab), stride);
  if (st) {
    ctx->pool->free(slab);
    return st;
  }
  for (size_t j = 0; j < nbins; ++j)
    out[j] = static_cast<uint32_t *>(slab) + j * stride;
  if (single_bytes) *single_bytes = mbx_bitmap_words(nrows) * 4;
  return MBX_OK;
}

int mbx_create_bin_multi(mbx_ctx *ctx, const uint32_t *fk, uint32_t dimsz,
                         uint32_t nrows, const mbx_binset *sets, size_t nsets,
                         size_t stride_words) {
  if (!ctx || !fk || !sets) return fail(MBX_ERR_INVALID, "create_bin_multi: null");
  if (nsets == 0 || nrows == 0)
    return fail(MBX_ERR_INVALID, "create_bin_multi: no sets or rows");
  if (reinterpret_cast<uintptr_t>(fk) % 4)
    return fail(MBX_ERR_ALIGN, "create_bin_multi: misaligned fk");
  if (stride_words % 64 != 0 || stride_words < mbx_bitmap_words(nrows))
    return fail(MBX_ERR_INVALID,
                "create_bin_multi: stride_words must be a multiple of 64 and "
                ">= ceil(N/32)");
  for (size_t i = 0; i < nsets; ++i) {
    const mbx_binset &b = sets[i];
    if (!b.data || !b.mins || !b.maxs || !b.slab || b.nbins == 0)
      return fail(MBX_ERR_INVALID, "create_bin_multi: bad set");
    if (b.elem_bytes != 1 && b.elem_bytes != 2 && b.elem_bytes != 4 &&
        b.elem_bytes != 8)
      return fail(MBX_ERR_INVALID, "create_bin_multi: elem_bytes");
    if (reinterpret_cast<uintptr_t>(b.slab) % 16)
      return fail(MBX_ERR_ALIGN, "create_bin_multi: output not 16-byte aligned");
  }
  // the packed path: radix-able bin sets whose indexes fit 32 bits together
  std::vector<BinArgs> args(nsets);
  std::vector<uint32_t> bits(nsets);
  uint32_t total = 0;
  bool packed = nsets <= (size_t)kMaxSets && dimsz < 0x80000000u;
  for (size_t i = 0; i < nsets && packed; ++i) {
    BinArgs &a = args[i];
    std::memset(&a, 0, sizeof(a));
    a.src.data = sets[i].data;
    a.src.fk = fk;
    a.src.dimsz = dimsz;
    a.src.eb = sets[i].elem_bytes;
    a.nb = sets[i].nbins;
    if (sets[i].nbins >= 0xffff ||
        bin_mode(sets[i].mins, sets[i].maxs, sets[i].nbins, &a) == kGeneric)
      packed = false;
    bits[i] = sets[i].nbins < 0xff ? 8 : 16;
    total += bits[i];
  }
  if (total > 32) packed = false;
  void *d_tab = nullptr;
  if (packed &&
      ctx->pool->alloc((((size_t)dimsz + 1) * 4 + 255) & ~(size_t)255, &d_tab) != MBX_OK)
    packed = false;
  if (!packed) {   // one pass per attribute, as create_bin_into
    for (size_t i = 0; i < nsets; ++i) {
      mbx_colref r;
      r.data = sets[i].data;
      r.fk = fk;
      r.dimsz = dimsz;
      r.elem_bytes = sets[i].elem_bytes;
      int st = launch_bins(ctx, &r, nrows, sets[i].mins, sets[i].maxs,
                           sets[i].nbins, sets[i].slab, stride_words);
      if (st) return st;
    }
    return MBX_OK;
  }
  uint32_t *tab = static_cast<uint32_t *>(d_tab);
  int st = launch_fill_u32(ctx->stream, tab, 0, (size_t)dimsz + 1);
  std::vector<void *> tmps;
  MultiArgs m;
  std::memset(&m, 0, sizeof(m));
  m.fk = fk;
  m.dimsz = dimsz;
  m.nrows = nrows;
  m.nblocks = (uint32_t)(stride_words / 32);
  m.nsets = (uint32_t)nsets;
  m.stride_words = stride_words;
  m.tab = tab;
  uint32_t shift = 0;
  size_t tg = ceil_div((size_t)dimsz + 1, 256);
  if (tg > (size_t)ctx->num_cus * 8) tg = (size_t)ctx->num_cus * 8;
  for (size_t i = 0; i < nsets && !st; ++i) {
    BinArgs &a = args[i];
    if (a.mode == kSorted) {   // bounds on the device for the search
      void *tmp = nullptr;
      st = ctx->pool->alloc((a.nb + 1) * 8, &tmp);
      if (st) break;
      tmps.push_back(tmp);
      std::vector<uint64_t> h(a.nb + 1);
      for (uint32_t j = 0; j < a.nb; ++j) h[j] = sets[i].mins[j];
      h[a.nb] = sets[i].maxs[a.nb - 1];
      hipError_t e = hipMemcpyAsync(tmp, h.data(), h.size() * 8,
                                    hipMemcpyHostToDevice, ctx->stream);
      if (e != hipSuccess) {
        st = hip_fail(e, "create_bin_multi: bounds upload");
        break;
      }
      a.mins = static_cast<uint64_t *>(tmp);
      hipLaunchKernelGGL(dim_pack_kernel<true>, dim3((unsigned)tg), dim3(256), 0,
                         ctx->stream, a, shift, bits[i], tab);
    } else {
      hipLaunchKernelGGL(dim_pack_kernel<false>, dim3((unsigned)tg), dim3(256), 0,
                         ctx->stream, a, shift, bits[i], tab);
    }
    m.set[i].nb = a.nb;
    m.set[i].shift = shift;
    m.set[i].bits = bits[i];
    m.set[i].out = sets[i].slab;
    shift += bits[i];
  }
  if (!st) {
    bool scatter = false;
    for (size_t i = 0; i < nsets; ++i) scatter |= sets[i].nbins > 8;
    // with the LDS images (32 KB per workgroup) and 113 VGPRs four
    // workgroups fit a CU: all resident, none queued
    const size_t cap = (size_t)ctx->num_cus * (scatter ? 4 : 16);
    size_t wgs = ceil_div(m.nblocks, 4);
    if (wgs > cap) wgs = cap;
    if (wgs == 0) wgs = 1;
    if (scatter)
      hipLaunchKernelGGL(bins_multi_kernel<true>, dim3((unsigned)wgs), dim3(256),
                         kScatterLds, ctx->stream, m);
    else
      hipLaunchKernelGGL(bins_multi_kernel<false>, dim3((unsigned)wgs), dim3(256),
                         0, ctx->stream, m);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) st = hip_fail(e, "create_bin_multi: launch");
  }
  // the table and bounds must outlive the kernels: wait first
  hipError_t e2 = hipStreamSynchronize(ctx->stream);
  for (void *t : tmps) ctx->pool->free(t);
  ctx->pool->free(d_tab);
  if (!st && e2 != hipSuccess) st = hip_fail(e2, "create_bin_multi");
  return st;
}

int mbx_create_join(mbx_ctx *ctx, const uint32_t *fk, const void *dimattr,
                    uint32_t elem_bytes, uint32_t nrows, uint32_t dimsz,
                    uint32_t lo, uint32_t hi, uint32_t **out) {
  if (!out) return fail(MBX_ERR_INVALID, "create_join: null out");
  *out = nullptr;
  mbx_colref src;
  src.data = dimattr;
  src.fk = fk;
  src.dimsz = dimsz;
  src.elem_bytes = elem_bytes;
  uint64_t mn = lo, mx = hi;
  if (mn >= mx) {
    // an empty range still yields an all-zero bitmap, as the reference's
    // predicate `val >= min && val < max` does
    mx = mn;
  }
  uint32_t *res[1] = {nullptr};
  if (mn == mx) {
    const size_t stride = mbx_bitmap_stride_words(nrows);
    void *slab = nullptr;
    int st = ctx ? ctx->pool->alloc(stride * 4, &slab)
                 : fail(MBX_ERR_INVALID, "null ctx");
    if (st) return st;
    st = launch_fill_u32(ctx->stream, slab, 0, stride);
    if (st) {
      ctx->pool->free(slab);
      return st;
    }
    *out = static_cast<uint32_t *>(slab);
    return MBX_OK;
  }
  int st = mbx_create_bin(ctx, &src, nrows, &mn, &mx, 1, res, nullptr);
  if (st) return st;
  *out = res[0];
  return MBX_OK;
}

int mbx_create_select(mbx_ctx *ctx, const void *srcp, uint32_t elem_bytes,
                      uint32_t nrows, uint32_t lo, uint32_t hi,
                      uint32_t **out) {
  return mbx_create_join(ctx, nullptr, srcp, elem_bytes, nrows, 0, lo, hi,
                         out);
}

}  // extern "C"
