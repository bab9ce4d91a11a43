// Device-wide primitives: exclusive scan (u32), stable LSD radix sort of
// (key, u32 value) pairs, fill. They replace del-cudarc-sys's cumsum,
// sort_by_key_u64 and array1d::fill used by the reference host layer
// (del-splat-cudarc/src/splat_gauss.rs:137-146,184-191).
//
// Radix sort: digits of at most 8 bits (the key bits spread evenly over the
// fewest passes), one (upsweep, column scan, scatter) triple per pass. A 256-thread workgroup owns a 2048-element tile; each wave64 owns 512
// consecutive elements and ranks them in rounds of 64 with a ballot-based
// digit match, keeping per-wave digit counters in LDS, so the rank of an
// element is (digit offset of earlier tiles) + (count in earlier waves of the
// tile) + (count in earlier rounds of the wave) + (matching lanes below it).
// That order is exactly element order: the sort is stable.
#include "dsplat_common.hpp"
#include "primitives.hpp"

namespace dsplat {

// ------------------------------------------------------------------ scan ----
constexpr int kScanItems = 8;
constexpr int kScanTile = kScanBlock * kScanItems;

// DPP row shifts and broadcasts: no LDS round trips.
__device__ inline uint32_t wave_inclusive_scan(uint32_t v) { return wave_incl_sum_u32(v); }

__device__ inline uint32_t load_n(uint32_t n_cap, const uint32_t* d_n) {
  if (!d_n) return n_cap;
  const uint32_t v = *d_n;
  return v < n_cap ? v : n_cap;
}

template <int ITEMS>
__global__ void __launch_bounds__(kScanBlock)
scan_reduce_kernel(uint32_t n_cap, const uint32_t* d_n, const uint32_t* __restrict__ in,
                   uint32_t* __restrict__ partial) {
  const uint32_t n = load_n(n_cap, d_n);
  const uint64_t base = uint64_t(blockIdx.x) * (kScanBlock * ITEMS);
  uint32_t s = 0;
#pragma unroll
  for (int k = 0; k < ITEMS; ++k) {
    const uint64_t i = base + uint64_t(k) * kScanBlock + threadIdx.x;
    if (i < n) s += in[i];
  }
  uint32_t tot;
  block_exclusive_scan(s, &tot);
  if (threadIdx.x == 0) partial[blockIdx.x] = tot;
}

// Single workgroup: exclusive scan of nb partial sums, in place; also writes
// the grand total to partial[nb]. 16 consecutive values per thread: one block
// scan per 4096 values.
__global__ void __launch_bounds__(kScanBlock)
scan_partials_kernel(uint32_t nb, uint32_t* __restrict__ partial, uint32_t* d_total) {
  constexpr uint32_t kPer = 16, kSpan = kScanBlock * kPer;
  uint32_t carry = 0;
  for (uint32_t base = 0; base < nb; base += kSpan) {
    const uint32_t i0 = base + threadIdx.x * kPer;
    uint32_t v[kPer];
    uint32_t sum = 0;
#pragma unroll
    for (uint32_t k = 0; k < kPer; ++k) {
      v[k] = i0 + k < nb ? partial[i0 + k] : 0u;
      sum += v[k];
    }
    uint32_t tot;
    uint32_t run = carry + block_exclusive_scan(sum, &tot);
#pragma unroll
    for (uint32_t k = 0; k < kPer; ++k) {
      if (i0 + k < nb) partial[i0 + k] = run;
      run += v[k];
    }
    carry += tot;
  }
  if (threadIdx.x == 0) {
    partial[nb] = carry;
    if (d_total) *d_total = carry;
  }
}

__global__ void set_u32_kernel(uint32_t* p, uint32_t v) { *p = v; }

__global__ void __launch_bounds__(kScanBlock)
scan_down_kernel(uint32_t n_cap, const uint32_t* d_n, const uint32_t* __restrict__ in,
                 uint32_t* __restrict__ out, const uint32_t* __restrict__ partial) {
  const uint32_t n = load_n(n_cap, d_n);
  const uint64_t base = uint64_t(blockIdx.x) * kScanTile + uint64_t(threadIdx.x) * kScanItems;
  uint32_t v[kScanItems];
  uint32_t s = 0;
#pragma unroll
  for (int k = 0; k < kScanItems; ++k) {
    const uint64_t i = base + k;
    v[k] = i < n ? in[i] : 0u;
    s += v[k];
  }
  uint32_t tot;
  uint32_t run = block_exclusive_scan(s, &tot) + partial[blockIdx.x];
#pragma unroll
  for (int k = 0; k < kScanItems; ++k) {
    const uint64_t i = base + k;
    if (i < n) out[i] = run;
    run += v[k];
  }
}

size_t scan_workspace_bytes(uint32_t n) {
  const uint32_t nb = div_up(n, kScanTile) + 1;
  return (size_t(nb) + 1) * sizeof(uint32_t) + 256;
}

int exclusive_scan(hipStream_t s, uint32_t n_cap, const uint32_t* d_n, const uint32_t* in,
                   uint32_t* out, uint32_t* ws, uint32_t* d_total) {
  if (n_cap == 0) {
    if (d_total) set_u32_kernel<<<1, 1, 0, s>>>(d_total, 0u);
    DSPLAT_HIP(hipGetLastError());
    return DSPLAT_OK;
  }
  const uint32_t nb = div_up(n_cap, kScanTile);
  scan_reduce_kernel<kScanItems><<<nb, kScanBlock, 0, s>>>(n_cap, d_n, in, ws);
  scan_partials_kernel<<<1, kScanBlock, 0, s>>>(nb, ws, d_total);
  scan_down_kernel<<<nb, kScanBlock, 0, s>>>(n_cap, d_n, in, out, ws);
  DSPLAT_HIP(hipGetLastError());
  return DSPLAT_OK;
}

// Up to this many chunk sums are scanned by one workgroup; more go through
// the three-kernel scan (one workgroup walking ~10^5 sums takes ~100 us).
constexpr uint32_t kSinglePassPartials = 16384;

__global__ void copy_word_kernel(const uint32_t* __restrict__ src, uint32_t* __restrict__ dst) {
  *dst = *src;
}

size_t scan_chunks_workspace_bytes(uint32_t n) {
  const uint32_t nb = div_up(n, kScanBlock);
  size_t b = (size_t(nb) + 2) * sizeof(uint32_t) + 256;
  if (nb > kSinglePassPartials) b = ((b + 255) & ~size_t(255)) + scan_workspace_bytes(nb);
  return b;
}

int scan_chunk_prefixes(hipStream_t s, uint32_t n, const uint32_t* in, uint32_t* ws,
                        uint32_t* d_total) {
  if (n == 0) {
    if (d_total) set_u32_kernel<<<1, 1, 0, s>>>(d_total, 0u);
    DSPLAT_HIP(hipGetLastError());
    return DSPLAT_OK;
  }
  const uint32_t nb = div_up(n, kScanBlock);
  scan_reduce_kernel<1><<<nb, kScanBlock, 0, s>>>(n, nullptr, in, ws);
  if (nb <= kSinglePassPartials) {
    scan_partials_kernel<<<1, kScanBlock, 0, s>>>(nb, ws, d_total);
  } else {
    const size_t head = ((size_t(nb) + 2) * sizeof(uint32_t) + 256 + 255) & ~size_t(255);
    uint32_t* ws2 = reinterpret_cast<uint32_t*>(reinterpret_cast<char*>(ws) + head);
    if (int rc = exclusive_scan(s, nb, nullptr, ws, ws, ws2, ws + nb)) return rc;
    if (d_total) copy_word_kernel<<<1, 1, 0, s>>>(ws + nb, d_total);
  }
  DSPLAT_HIP(hipGetLastError());
  return DSPLAT_OK;
}

// ------------------------------------------------------------ radix sort ----
// Per pass three kernels: upsweep (per-tile digit counts), column scan (one
// workgroup per digit: the prefix of the digit's counts over the tiles, and
// the digit's total), scatter (digit bases from the totals, then the stable
// in-tile rank).
constexpr int kSortBlock = 256;
// Rounds per wave: 8 (2048-element tiles); 16 (4096) for sorts of at least
// kBigSort elements, whose scatter then writes runs twice as long per digit
// (fewer partial lines) over tiles that still fill the GPU; 4 (1024) for
// sorts below kMidSort elements, which are latency-bound: more, shorter
// workgroups (the 1M-element depth sort per pass 23 -> 20 us).
constexpr int kSortItemsSmall = 4;
constexpr int kSortItems = 8;
constexpr int kSortItemsBig = 16;
constexpr uint32_t kBigSort = 8u << 20;
constexpr uint32_t kMidSort = 2u << 20;
inline int sort_items(uint32_t n) {
  return n >= kBigSort ? kSortItemsBig : n >= kMidSort ? kSortItems : kSortItemsSmall;
}
constexpr int kSortWaves = kSortBlock / 64;

struct SortLayout {
  void* kalt;
  uint32_t* valt;
  uint32_t* tot;   // [256] digit totals of the current pass
  uint32_t* cnt;   // [256][slots] per-tile digit counts, then their prefixes
  size_t total;
};

// Sort tile of workgroup b in a pass over nbv tiles. Workgroups are
// dispatched to the 8 XCDs round-robin (b mod 8), and each XCD takes a
// contiguous range of tiles: consecutive tiles then run on the same XCD at
// about the same time, so the runs that they write side by side for one
// digit (and their words of a count column) meet in one L2 and leave it as
// whole lines. With tile b on workgroup b, every digit run shared its end
// lines with a run written through another XCD's L2.
constexpr uint32_t kXcds = 8;
__host__ __device__ inline uint32_t col_stride(uint32_t nb) { return kXcds * div_up(nb, kXcds); }
__device__ inline uint32_t xcd_tile(uint32_t b, uint32_t nbv) {
  const uint32_t x = b % kXcds, q = nbv / kXcds, r = nbv % kXcds;
  return x * q + min(x, r) + b / kXcds;
}

// Sort tile of workgroup b when the nbv tiles form segments of cps tiles
// each (cps = 0: one segment): segment after segment in workgroup order, and
// within a segment (cps a multiple of 8, so b mod 8 is still the XCD) each
// XCD takes a contiguous range. The workgroups resident at once then work on
// one or two segments.
__device__ inline uint32_t seg_tile(uint32_t b, uint32_t nbv, uint32_t cps) {
  if (cps == 0) return xcd_tile(b, nbv);
  const uint32_t sg = b / cps;
  return sg * cps + xcd_tile(b - sg * cps, cps);
}

template <typename KeyT>
SortLayout sort_layout(void* base, uint32_t n) {
  const uint32_t nb = div_up(n ? n : 1, kSortBlock * sort_items(n));
  SortLayout L{};
  char* p = static_cast<char*>(base);
  size_t o = 0;
  auto take = [&](size_t bytes) {
    char* q = p ? p + o : nullptr;
    o += (bytes + 255) & ~size_t(255);
    return q;
  };
  L.kalt = take(size_t(n) * sizeof(KeyT));
  L.valt = reinterpret_cast<uint32_t*>(take(size_t(n) * 4));
  L.tot = reinterpret_cast<uint32_t*>(take(size_t(256) * kSortMaxSegs * 4));
  L.cnt = reinterpret_cast<uint32_t*>(take(size_t(256) * col_stride(nb) * 4));
  L.total = o;
  return L;
}

template <typename KeyT, int ITEMS>
__global__ void __launch_bounds__(kSortBlock)
radix_upsweep_kernel(uint32_t n_cap, const uint32_t* d_n, const KeyT* __restrict__ keys,
                     uint32_t shift, uint32_t dbits, uint32_t nb, uint32_t cps,
                     uint32_t* __restrict__ cnt) {
  constexpr int kTile = kSortBlock * ITEMS;
  __shared__ uint32_t h[256];
  const uint32_t n = load_n(n_cap, d_n);
  const uint32_t nbv = div_up(n, uint32_t(kTile));
  // workgroups past the tiles holding elements (the grid covers the
  // capacity) have nothing to count: the column scan and the scatter stop
  // at the last tile holding an element
  if (blockIdx.x >= nbv) return;  // uniform
  const uint32_t tb = seg_tile(blockIdx.x, nbv, cps);
  const uint64_t base = uint64_t(tb) * kTile;
  const int lane = threadIdx.x & 63;
  h[threadIdx.x] = 0;
  __syncthreads();
  const uint32_t mask = (1u << dbits) - 1u;
  // all loads first, then the histogram
  KeyT kk[ITEMS];
#pragma unroll
  for (int k = 0; k < ITEMS; ++k) {
    const uint64_t i = base + uint64_t(k) * kSortBlock + threadIdx.x;
    kk[k] = i < n ? keys[i] : KeyT(0);
  }
#pragma unroll
  for (int k = 0; k < ITEMS; ++k) {
    const uint64_t i = base + uint64_t(k) * kSortBlock + threadIdx.x;
    const bool valid = i < n;
    const uint32_t d = valid ? uint32_t(kk[k] >> shift) & mask : 0u;
    const uint64_t act = __ballot(valid);
    if (act == 0) continue;  // wave-uniform
    // valid lanes are a prefix of the wave, so lane 0 is one of them; a
    // digit shared by the whole wave takes one LDS add, not 64 conflicting
    const uint32_t d0 = __builtin_amdgcn_readfirstlane(d);
    if (__ballot(valid && d != d0) == 0) {
      if (lane == 0) atomicAdd(&h[d0], uint32_t(__popcll(act)));
    } else if (valid) {
      atomicAdd(&h[d], 1u);
    }
  }
  __syncthreads();
  if (threadIdx.x <= mask)
    cnt[size_t(threadIdx.x) * col_stride(nb) + tb] = h[threadIdx.x];
}

// Workgroup (d, sg): cnt[d][b] <- elements of digit d in the tiles of
// segment sg before b; tot[sg][d] <- elements of digit d in segment sg
// (segments of cps tiles; cps = 0: one segment of all tiles).
__global__ void __launch_bounds__(kSortBlock)
radix_colscan_kernel(uint32_t n_cap, const uint32_t* d_n, uint32_t nb, uint32_t tile,
                     uint32_t cps, uint32_t* __restrict__ cnt, uint32_t* __restrict__ tot) {
  const uint32_t n = load_n(n_cap, d_n);
  const uint32_t d = blockIdx.x, sg = blockIdx.y;
  const uint32_t all = div_up(n, tile);
  const uint32_t c0 = cps ? sg * cps : 0u;
  const uint32_t nbv = cps ? min(cps, all > c0 ? all - c0 : 0u) : all;
  uint32_t carry = 0;
  // the segment's column starts on a multiple of 8 words (cps is one)
  uint32_t* col = cnt + size_t(d) * col_stride(nb) + c0;
  // column values per thread and round, as two 16-byte words (the column
  // starts on a multiple of 8 words and holds col_stride(nb) >= nbv of them)
  constexpr int kPer = 8, kSpan = kSortBlock * kPer;
  for (uint32_t b0 = 0; b0 < nbv; b0 += kSpan) {
    const uint32_t i0 = b0 + threadIdx.x * kPer;
    const bool full = i0 + kPer <= nbv;
    uint32_t v[kPer], s = 0;
    if (full) {
      const uint4 a = *reinterpret_cast<const uint4*>(col + i0);
      const uint4 b = *reinterpret_cast<const uint4*>(col + i0 + 4);
      v[0] = a.x, v[1] = a.y, v[2] = a.z, v[3] = a.w, v[4] = b.x, v[5] = b.y, v[6] = b.z, v[7] = b.w;
    } else {
#pragma unroll
      for (int k = 0; k < kPer; ++k) v[k] = i0 + k < nbv ? col[i0 + k] : 0u;
    }
#pragma unroll
    for (int k = 0; k < kPer; ++k) s += v[k];
    uint32_t chunk;
    uint32_t run = carry + block_exclusive_scan(s, &chunk);
    uint32_t o[kPer];
#pragma unroll
    for (int k = 0; k < kPer; ++k) {
      o[k] = run;
      run += v[k];
    }
    if (full) {
      *reinterpret_cast<uint4*>(col + i0) = make_uint4(o[0], o[1], o[2], o[3]);
      *reinterpret_cast<uint4*>(col + i0 + 4) = make_uint4(o[4], o[5], o[6], o[7]);
    } else {
#pragma unroll
      for (int k = 0; k < kPer; ++k)
        if (i0 + k < nbv) col[i0 + k] = o[k];
    }
    carry += chunk;
  }
  if (threadIdx.x == 0) tot[size_t(sg) * 256 + d] = carry;
}

// iota: the values are the element indices (vin is not read). cdst (may be
// null): also gathers rsrc[value], a packed tile rectangle, writing it to
// rdst[pos] and its area to cdst[pos]. Without values (VALS false) the keys
// alone move: kout (may be null) receives them and vout (may be null) the
// key's low bits, key & vmask.
template <typename KeyT, int ITEMS, bool VALS>
__global__ void __launch_bounds__(kSortBlock)
radix_scatter_kernel(uint32_t n_cap, const uint32_t* d_n, const KeyT* __restrict__ kin,
                     const uint32_t* __restrict__ vin, bool iota, KeyT* __restrict__ kout,
                     uint32_t* __restrict__ vout, uint32_t shift, uint32_t dbits,
                     uint32_t nb, uint32_t cps, uint32_t seg_len,
                     const uint32_t* __restrict__ off,
                     const uint32_t* __restrict__ tot, uint32_t nbins,
                     const uint2* __restrict__ rsrc, uint32_t* __restrict__ cdst,
                     uint2* __restrict__ rdst, uint32_t vmask) {
  static_assert(kSortBlock == 256, "one thread per digit");
  constexpr int kTile = kSortBlock * ITEMS, kWaveChunk = 64 * ITEMS;
  __shared__ uint32_t wcnt[kSortWaves][256];
  __shared__ uint32_t dpre[256];
  __shared__ uint32_t dbase[256];
  __shared__ KeyT skey[kTile];
  __shared__ uint32_t sval[VALS ? kTile : 1];
  const uint32_t n = load_n(n_cap, d_n);
  const uint32_t nbv = div_up(n, uint32_t(kTile));
  if (blockIdx.x >= nbv) return;  // uniform
  const uint32_t tb = seg_tile(blockIdx.x, nbv, cps);
  const uint32_t sg = cps ? tb / cps : 0u;
  const uint64_t bstart = uint64_t(tb) * kTile;
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  // the digit total and this tile's column prefix of digit threadIdx.x,
  // fetched now so their latency hides behind the ranking
  const bool live_digit = threadIdx.x < nbins;
  const uint32_t my_tot = live_digit ? tot[size_t(sg) * 256 + threadIdx.x] : 0u;
  const uint32_t my_off =
      live_digit ? off[size_t(threadIdx.x) * col_stride(nb) + tb] : 0u;
  for (int k = threadIdx.x; k < kSortWaves * 256; k += kSortBlock) (&wcnt[0][0])[k] = 0;
  __syncthreads();
  const uint32_t mask = (1u << dbits) - 1u;
  const uint64_t base = bstart + uint64_t(w) * kWaveChunk;
  KeyT key[ITEMS];
  uint32_t val[ITEMS], rank[ITEMS], dig[ITEMS];
#pragma unroll
  for (int r = 0; r < ITEMS; ++r) {
    const uint64_t i = base + uint64_t(r) * 64 + lane;
    const bool valid = i < n;
    key[r] = valid ? kin[i] : KeyT(0);
    if (VALS) val[r] = valid ? (iota ? uint32_t(i) : vin[i]) : 0u;
  }
#pragma unroll
  for (int r = 0; r < ITEMS; ++r) {
    const uint64_t i = base + uint64_t(r) * 64 + lane;
    const bool valid = i < n;
    const uint32_t d = uint32_t(key[r] >> shift) & mask;
    dig[r] = valid ? d : 0xffffffffu;
    uint64_t peers = __ballot(valid);
    for (uint32_t b = 0; b < dbits; ++b) {
      const bool bit = (d >> b) & 1u;
      const uint64_t bb = __ballot(bit);
      peers &= bit ? bb : ~bb;
    }
    const uint32_t below = lanes_below(peers);
    uint32_t running = 0;
    if (valid) running = wcnt[w][d];
    rank[r] = running + below;
    if (valid && below == 0) wcnt[w][d] = running + uint32_t(__popcll(peers));
  }
  __syncthreads();
  // exclusive prefix of the per-wave counts over waves, per digit, and the
  // tile's digit totals
  uint32_t total = 0;
#pragma unroll
  for (int k = 0; k < kSortWaves; ++k) {
    const uint32_t c = wcnt[k][threadIdx.x];
    wcnt[k][threadIdx.x] = total;
    total += c;
  }
  // dpre[d]: elements of the tile with a smaller digit; dbase[d]: the
  // segment's start, its elements with a smaller digit, and those of digit
  // d in its earlier tiles
  uint32_t btot, all;
  dpre[threadIdx.x] = block_exclusive_scan(total, &btot);
  const uint32_t gpre = block_exclusive_scan(my_tot, &all);
  if (live_digit) dbase[threadIdx.x] = sg * seg_len + gpre + my_off;
  __syncthreads();
  // stage in LDS in tile-sorted order ...
#pragma unroll
  for (int r = 0; r < ITEMS; ++r) {
    if (dig[r] == 0xffffffffu) continue;
    const uint32_t d = dig[r];
    const uint32_t lp = dpre[d] + wcnt[w][d] + rank[r];
    skey[lp] = key[r];
    if (VALS) sval[lp] = val[r];
  }
  __syncthreads();
  // ... and write each digit's run to consecutive global addresses (the
  // gathered values loaded for all rounds before any is stored)
  const uint32_t nvalid = uint32_t(min(uint64_t(kTile), n - bstart));
  if (!VALS) {
#pragma unroll
    for (int r = 0; r < ITEMS; ++r) {
      const uint32_t i = uint32_t(r) * kSortBlock + threadIdx.x;
      if (i < nvalid) {
        const KeyT k = skey[i];
        const uint32_t d = uint32_t(k >> shift) & mask;
        const uint32_t pos = dbase[d] + (i - dpre[d]);
        if (kout) kout[pos] = k;
        if (vout) vout[pos] = uint32_t(k) & vmask;
      }
    }
    return;
  }
  uint2 gv[ITEMS];
#pragma unroll
  for (int r = 0; r < ITEMS; ++r) {
    const uint32_t i = uint32_t(r) * kSortBlock + threadIdx.x;
    gv[r] = (cdst && i < nvalid) ? rsrc[sval[i]] : make_uint2(0u, 0u);
  }
#pragma unroll
  for (int r = 0; r < ITEMS; ++r) {
    const uint32_t i = uint32_t(r) * kSortBlock + threadIdx.x;
    if (i < nvalid) {
      const KeyT k = skey[i];
      const uint32_t d = uint32_t(k >> shift) & mask;
      const uint32_t pos = dbase[d] + (i - dpre[d]);
      kout[pos] = k;
      vout[pos] = sval[i];
      if (cdst) {
        cdst[pos] = rect_area(gv[r]);
        rdst[pos] = gv[r];
      }
    }
  }
}

template <typename KeyT>
size_t sort_workspace_bytes(uint32_t n) {
  return sort_layout<KeyT>(nullptr, n).total;
}

template <typename KeyT>
int radix_sort_pairs(hipStream_t s, uint32_t n_cap, const uint32_t* d_n, KeyT* keys,
                     uint32_t* vals, uint32_t key_bits, void* ws, size_t ws_bytes,
                     bool* result_in_alt, KeyT** keys_alt_out, uint32_t** vals_alt_out,
                     bool iota_vals, const uint2* rect_src, uint32_t* cnt_dst,
                     uint2* rect_dst, uint32_t seg_len) {
  *result_in_alt = false;
  if (seg_len && (d_n || seg_len % kSortSegAlign || n_cap % seg_len ||
                  n_cap / seg_len > kSortMaxSegs))
    return dsplat_set_error(DSPLAT_ERR_ARG, "radix sort: bad segment length");
  if (ws_bytes < sort_workspace_bytes<KeyT>(n_cap))
    return dsplat_set_error(DSPLAT_ERR_CAPACITY, "radix sort: workspace too small");
  if (key_bits == 0 || key_bits > 8 * sizeof(KeyT))
    return dsplat_set_error(DSPLAT_ERR_ARG, "radix sort: key_bits out of range");
  const SortLayout L = sort_layout<KeyT>(ws, n_cap);
  KeyT* kalt = static_cast<KeyT*>(L.kalt);
  if (keys_alt_out) *keys_alt_out = kalt;
  if (vals_alt_out) *vals_alt_out = L.valt;
  if (n_cap == 0) return DSPLAT_OK;
  const int items = sort_items(n_cap);
  const uint32_t tile = kSortBlock * items;
  const uint32_t nb = div_up(n_cap, tile);
  const uint32_t cps = seg_len / tile, nseg = seg_len ? n_cap / seg_len : 1u;
  // as few passes as 8-bit digits allow, with the bits spread evenly over
  // them: fewer digits per pass give longer runs per digit in the scatter
  // (11 bits: 6 + 5 rather than 8 + 3)
  const uint32_t P = div_up(key_bits, 8u);
  KeyT* ksrc = keys; uint32_t* vsrc = vals; KeyT* kdst = kalt; uint32_t* vdst = L.valt;
  bool in_alt = false;
  uint32_t shift = 0;
  for (uint32_t p = 0; p < P; ++p) {
    const uint32_t dbits = div_up(key_bits - shift, P - p);
    const bool last = p + 1 == P;
#define DSPLAT_PASS(I_)                                                                        \
  do {                                                                                         \
    radix_upsweep_kernel<KeyT, I_><<<nb, kSortBlock, 0, s>>>(n_cap, d_n, ksrc, shift, dbits, nb, \
                                                             cps, L.cnt);                      \
    radix_colscan_kernel<<<dim3(1u << dbits, nseg), kSortBlock, 0, s>>>(n_cap, d_n, nb, tile,  \
                                                                        cps, L.cnt, L.tot);    \
    radix_scatter_kernel<KeyT, I_, true><<<nb, kSortBlock, 0, s>>>(                            \
        n_cap, d_n, ksrc, vsrc, iota_vals && p == 0, kdst, vdst, shift, dbits, nb, cps,        \
        seg_len, L.cnt, L.tot, 1u << dbits, rect_src, last ? cnt_dst : nullptr,                \
        last ? rect_dst : nullptr, 0u);                                                        \
  } while (0)
    if (items == 16)
      DSPLAT_PASS(16);
    else if (items == 8)
      DSPLAT_PASS(8);
    else
      DSPLAT_PASS(4);
#undef DSPLAT_PASS
    std::swap(ksrc, kdst);
    std::swap(vsrc, vdst);
    in_alt = !in_alt;
    shift += dbits;
  }
  DSPLAT_HIP(hipGetLastError());
  *result_in_alt = in_alt;
  return DSPLAT_OK;
}

int radix_sort_keys(hipStream_t s, uint32_t n_cap, const uint32_t* d_n, uint32_t* keys,
                    uint32_t shift0, uint32_t key_bits, uint32_t vmask, uint32_t* vout, void* ws,
                    size_t ws_bytes, bool* result_in_alt, uint32_t** keys_alt_out) {
  *result_in_alt = false;
  if (ws_bytes < sort_workspace_bytes<uint32_t>(n_cap))
    return dsplat_set_error(DSPLAT_ERR_CAPACITY, "radix sort: workspace too small");
  if (key_bits == 0 || shift0 + key_bits > 32)
    return dsplat_set_error(DSPLAT_ERR_ARG, "radix sort: key bits out of range");
  const SortLayout L = sort_layout<uint32_t>(ws, n_cap);
  uint32_t* kalt = static_cast<uint32_t*>(L.kalt);
  if (keys_alt_out) *keys_alt_out = kalt;
  if (n_cap == 0) return DSPLAT_OK;
  const int items = sort_items(n_cap);
  const uint32_t tile = kSortBlock * items;
  const uint32_t nb = div_up(n_cap, tile);
  const uint32_t P = div_up(key_bits, 8u);
  uint32_t* ksrc = keys;
  uint32_t* kdst = kalt;
  bool in_alt = false;
  uint32_t used = 0;
  for (uint32_t p = 0; p < P; ++p) {
    const uint32_t dbits = div_up(key_bits - used, P - p);
    const uint32_t shift = shift0 + used;
    uint32_t* vo = p + 1 == P ? vout : nullptr;
#define DSPLAT_KPASS(I_)                                                                        \
  do {                                                                                          \
    radix_upsweep_kernel<uint32_t, I_><<<nb, kSortBlock, 0, s>>>(n_cap, d_n, ksrc, shift,       \
                                                                 dbits, nb, 0u, L.cnt);         \
    radix_colscan_kernel<<<dim3(1u << dbits, 1), kSortBlock, 0, s>>>(n_cap, d_n, nb, tile, 0u,  \
                                                                     L.cnt, L.tot);             \
    radix_scatter_kernel<uint32_t, I_, false><<<nb, kSortBlock, 0, s>>>(                        \
        n_cap, d_n, ksrc, nullptr, false, kdst, vo, shift, dbits, nb, 0u, 0u, L.cnt, L.tot,     \
        1u << dbits, nullptr, nullptr, nullptr, vmask);                                         \
  } while (0)
    if (items == 16)
      DSPLAT_KPASS(16);
    else if (items == 8)
      DSPLAT_KPASS(8);
    else
      DSPLAT_KPASS(4);
#undef DSPLAT_KPASS
    std::swap(ksrc, kdst);
    in_alt = !in_alt;
    used += dbits;
  }
  DSPLAT_HIP(hipGetLastError());
  *result_in_alt = in_alt;
  return DSPLAT_OK;
}

template size_t sort_workspace_bytes<uint32_t>(uint32_t);
template size_t sort_workspace_bytes<uint64_t>(uint32_t);
template int radix_sort_pairs<uint32_t>(hipStream_t, uint32_t, const uint32_t*, uint32_t*,
                                        uint32_t*, uint32_t, void*, size_t, bool*, uint32_t**,
                                        uint32_t**, bool, const uint2*, uint32_t*, uint2*,
                                        uint32_t);
template int radix_sort_pairs<uint64_t>(hipStream_t, uint32_t, const uint32_t*, uint64_t*,
                                        uint32_t*, uint32_t, void*, size_t, bool*, uint64_t**,
                                        uint32_t**, bool, const uint2*, uint32_t*, uint2*,
                                        uint32_t);

// ------------------------------------------------------------------ fill ----
__global__ void fill_f32_kernel(uint32_t n, float* __restrict__ buf, float v) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) buf[i] = v;
}

int fill_f32(hipStream_t s, uint32_t n, float* buf, float v) {
  if (n == 0) return DSPLAT_OK;
  fill_f32_kernel<<<div_up(n, 256), 256, 0, s>>>(n, buf, v);
  DSPLAT_HIP(hipGetLastError());
  return DSPLAT_OK;
}

// ------------------------------------------------------ order by cost ----
constexpr uint32_t kCostBins = 2048;
constexpr uint32_t kOrderBlock = 1024;

__global__ void __launch_bounds__(kOrderBlock)
order_by_cost_kernel(uint32_t n, const uint32_t* __restrict__ cost, uint32_t* __restrict__ order) {
  __shared__ uint32_t bin[kCostBins];
  __shared__ uint32_t wsum[kOrderBlock / 64];
  for (uint32_t i = threadIdx.x; i < kCostBins; i += kOrderBlock) bin[i] = 0;
  __syncthreads();
  // bin 0 holds the largest costs
  for (uint32_t t = threadIdx.x; t < n; t += kOrderBlock)
    atomicAdd(&bin[kCostBins - 1 - min(cost[t], kCostBins - 1)], 1u);
  __syncthreads();
  // exclusive scan of the bins, two per thread
  const uint32_t a = bin[2 * threadIdx.x], b = bin[2 * threadIdx.x + 1];
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const uint32_t inc = wave_inclusive_scan(a + b);
  if (lane == 63) wsum[w] = inc;
  __syncthreads();
  uint32_t pre = 0;
  for (int k = 0; k < w; ++k) pre += wsum[k];
  const uint32_t ex = pre + inc - (a + b);
  bin[2 * threadIdx.x] = ex;
  bin[2 * threadIdx.x + 1] = ex + a;
  __syncthreads();
  for (uint32_t t = threadIdx.x; t < n; t += kOrderBlock) {
    const uint32_t pos = atomicAdd(&bin[kCostBins - 1 - min(cost[t], kCostBins - 1)], 1u);
    order[pos] = t;
  }
}

__global__ void __launch_bounds__(256) iota_u32_kernel(uint32_t n, uint32_t* __restrict__ order) {
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
    order[i] = i;
}

int order_by_cost(hipStream_t s, uint32_t n, const uint32_t* cost, uint32_t* order) {
  static_assert(kCostBins == 2 * kOrderBlock, "two bins per thread");
  if (n == 0) return DSPLAT_OK;
  order_by_cost_kernel<<<1, kOrderBlock, 0, s>>>(n, cost, order);
  DSPLAT_HIP(hipGetLastError());
  return DSPLAT_OK;
}

int iota_u32(hipStream_t s, uint32_t n, uint32_t* order) {
  if (n == 0) return DSPLAT_OK;
  iota_u32_kernel<<<min(div_up(n, 256), 1024u), 256, 0, s>>>(n, order);
  DSPLAT_HIP(hipGetLastError());
  return DSPLAT_OK;
}

}  // namespace dsplat

// ------------------------------------------------------------------ C ABI ---
using namespace dsplat;

extern "C" size_t dsplat_exclusive_scan_workspace_size(uint32_t n) {
  return scan_workspace_bytes(n);
}

extern "C" int dsplat_exclusive_scan_u32(void* stream, uint32_t n, const uint32_t* in,
                                         uint32_t* out, void* workspace,
                                         size_t workspace_bytes) {
  if (n && (!in || !out)) return dsplat_set_error(DSPLAT_ERR_ARG, "scan: null buffer");
  if (workspace_bytes < scan_workspace_bytes(n) || (n && !workspace))
    return dsplat_set_error(DSPLAT_ERR_CAPACITY, "scan: workspace too small");
  return exclusive_scan(as_stream(stream), n, nullptr, in, out,
                        static_cast<uint32_t*>(workspace), nullptr);
}

extern "C" size_t dsplat_radix_sort_workspace_size(uint32_t n) {
  return sort_workspace_bytes<uint64_t>(n);
}

extern "C" int dsplat_radix_sort_by_key_u64_u32(void* stream, uint32_t n, uint64_t* keys,
                                                uint32_t* vals, uint32_t key_bits,
                                                void* workspace, size_t workspace_bytes) {
  if (n && (!keys || !vals)) return dsplat_set_error(DSPLAT_ERR_ARG, "sort: null buffer");
  hipStream_t s = as_stream(stream);
  bool in_alt = false;
  uint64_t* kalt = nullptr;
  uint32_t* valt = nullptr;
  int rc = radix_sort_pairs<uint64_t>(s, n, nullptr, keys, vals, key_bits, workspace,
                                      workspace_bytes, &in_alt, &kalt, &valt);
  if (rc) return rc;
  if (in_alt && n) {
    DSPLAT_HIP(hipMemcpyAsync(keys, kalt, size_t(n) * 8, hipMemcpyDeviceToDevice, s));
    DSPLAT_HIP(hipMemcpyAsync(vals, valt, size_t(n) * 4, hipMemcpyDeviceToDevice, s));
  }
  return DSPLAT_OK;
}

extern "C" int dsplat_fill_f32(void* stream, uint32_t n, float* buf, float value) {
  if (n && !buf) return dsplat_set_error(DSPLAT_ERR_ARG, "fill: null buffer");
  return fill_f32(as_stream(stream), n, buf, value);
}

extern "C" int dsplat_tile_order(void* stream, uint32_t num_tile, const uint32_t* tile_cost,
                                 uint32_t* tile_order) {
  if (num_tile && (!tile_cost || !tile_order))
    return dsplat_set_error(DSPLAT_ERR_ARG, "tile_order: null buffer");
  return order_by_cost(as_stream(stream), num_tile, tile_cost, tile_order);
}
