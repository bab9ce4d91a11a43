// Per-tile binning and stable ordering of (tile, Gaussian) pairs.
//
// Reference: splat_gauss::tile2idx_idx2pnt (del-splat-cudarc/src/
// splat_gauss.rs:96-195) with kernels count_splat_in_tile / fill_index_info
// (splat_gauss.cu:111-155, tile_acceleration.h:15-86) and a 64-bit radix sort
// of keys tile<<32 | bits(max(ndc_z+1, 0)) whose ties keep fill order
// (ascending Gaussian index). The circle path (splat_circle.rs:3-94) is the
// same with boxes [x-r, y-r, x+r, y+r].
//
// Same order, less sorting: the key of a pair is (tile, depth bits, index)
// and the depth bits belong to the Gaussian. So the N Gaussians are first
// sorted stably by depth bits (N << K), pairs are generated in that order,
// and a stable sort by tile id alone (bitwidth(T-1) bits instead of 64)
// finishes the job: within a tile, pairs keep the (depth, index) order of the
// first sort. Tile ranges are then the boundaries of the sorted tile ids.
#include <mutex>

#include "binning.hpp"
#include "dsplat_common.hpp"
#include "primitives.hpp"

namespace dsplat {

// Box source: the Gaussian path reads [N,4] boxes, the circle path builds
// them from the pixel centre and radius (aabb2::from_point).
struct GaussBoxes {
  const float4* aabb;
  __device__ float4 operator()(uint32_t i) const { return aabb[i]; }
};
struct CircleBoxes {
  const float* pixxydepth;
  const float* pixrad;
  __device__ float4 operator()(uint32_t i) const {
    const float x = pixxydepth[3 * i], y = pixxydepth[3 * i + 1], r = pixrad[i];
    return make_float4(x - r, y - r, x + r, y + r);
  }
};

// cnt[i] = tiles of Gaussian i, dkey[i] = its depth key (dkey may be null);
// partial[block] = the workgroup's sum of cnt.
template <typename Boxes>
__global__ void __launch_bounds__(256)
bin_count_kernel(uint32_t n, Boxes boxes, const float* __restrict__ pixxydepth, uint32_t ts,
                 uint32_t tw, uint32_t th, uint2* __restrict__ rect,
                 uint32_t* __restrict__ dkey, uint32_t* __restrict__ partial) {
  __shared__ uint32_t wsum[4];
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  uint32_t v = 0;
  if (i < n) {
    const float4 b = boxes(i);
    const TileRect r = tile_rect(b.x, b.y, b.z, b.w, ts, tw, th);
    v = uint32_t(r.count());
    rect[i] = pack_rect(r);
    if (dkey) dkey[i] = depth_key(pixxydepth[3 * i + 2]);
  }
  if (!partial) return;  // uniform: K comes from the scan of the ordered counts
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += uint32_t(__shfl_xor(int(v), off, 64));
  if ((threadIdx.x & 63) == 0) wsum[threadIdx.x >> 6] = v;
  __syncthreads();
  if (threadIdx.x == 0) partial[blockIdx.x] = wsum[0] + wsum[1] + wsum[2] + wsum[3];
}

// *d_k = sum of the nb partial sums (one workgroup).
__global__ void __launch_bounds__(256)
sum_partials_kernel(uint32_t nb, const uint32_t* __restrict__ partial, uint32_t* __restrict__ d_k) {
  __shared__ uint32_t wsum[4];
  uint32_t v = 0;
  for (uint32_t i = threadIdx.x; i < nb; i += 256) v += partial[i];
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += uint32_t(__shfl_xor(int(v), off, 64));
  if ((threadIdx.x & 63) == 0) wsum[threadIdx.x >> 6] = v;
  __syncthreads();
  if (threadIdx.x == 0) *d_k = wsum[0] + wsum[1] + wsum[2] + wsum[3];
}

// Pairs of the j-th Gaussian in depth order from its first slot, the
// exclusive scan of the counts cnt (in depth order): the chunk prefix of the
// workgroup's 256 Gaussians plus a block scan. Tile ids ix-outer, iy-inner
// (tile_acceleration.h:29-50), value = Gaussian index. n_eff = K if K <=
// capacity else 0 (overflow: nothing is written and the later phases see no
// pairs); workgroup 0 stores it for them. The rectangles come in depth order
// from the sort (read in sequence, no gather of boxes). pnt_of_pair null:
// one packed word per pair, tile << ib | Gaussian index, in tile_of_pair.
__global__ void __launch_bounds__(kScanBlock)
bin_fill_kernel(uint32_t n, const uint32_t* __restrict__ order,
                const uint2* __restrict__ rect_sorted, const uint32_t* __restrict__ cnt,
                const uint32_t* __restrict__ chunk_prefix,
                const uint32_t* __restrict__ d_k, uint32_t capacity,
                uint32_t* __restrict__ n_eff_out, uint32_t tw,
                uint32_t* __restrict__ tile_of_pair, uint32_t* __restrict__ pnt_of_pair,
                uint32_t ib) {
  __shared__ ExpandLds lds[4];
  const uint32_t k = *d_k;
  const uint32_t n_eff = k <= capacity ? k : 0u;
  if (blockIdx.x == 0 && threadIdx.x == 0) *n_eff_out = n_eff;
  if (n_eff == 0) return;  // uniform
  const uint32_t j = blockIdx.x * blockDim.x + threadIdx.x;
  const bool valid = j < n;
  uint32_t g = 0, c = 0;
  TileRect r{0, 0, 0, 0};
  if (valid) {
    g = order[j];
    r = unpack_rect(rect_sorted[j]);  // gathered into depth order by the sort
    c = cnt[j];
  }
  uint32_t block_total;
  const uint32_t o = chunk_prefix[blockIdx.x] + block_exclusive_scan(c, &block_total);
  wave_expand_pairs(valid, o, r, 0u, tw, g, lds[threadIdx.x >> 6], tile_of_pair, pnt_of_pair,
                    ib);
}

// tile2idx[t] = first sorted position whose tile id is >= t, t in [0, T]
// (tile_ranges4: four positions per thread; the tile id of a sorted word is
// word >> ib).
__global__ void __launch_bounds__(256)
tile_ranges_kernel(const uint32_t* __restrict__ n_eff, uint32_t capacity, uint32_t num_tile,
                   const uint32_t* __restrict__ tile_sorted, uint32_t ib,
                   uint32_t* __restrict__ tile2idx) {
  const uint32_t k = min(*n_eff, capacity);
  tile_ranges4(blockIdx.x * blockDim.x + threadIdx.x, k, num_tile, tile_sorted, tile2idx, ib);
}

__global__ void __launch_bounds__(256)
copy_u32_kernel(const uint32_t* __restrict__ n_eff, uint32_t cap, const uint32_t* __restrict__ src,
                uint32_t* __restrict__ dst) {
  const uint32_t n = min(*n_eff, cap);
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
    dst[i] = src[i];
}

__global__ void set_u32_one(uint32_t* p, uint32_t v) { *p = v; }

inline uint32_t bit_width(uint32_t v) {
  uint32_t b = 0;
  while (v) { ++b; v >>= 1; }
  return b;
}

inline size_t align256(size_t v) { return (v + 255) & ~size_t(255); }

// Workspace of the per-Gaussian phases (count, order), sized by N, and of
// the per-pair phase (fill, tile sort), sized by the pair capacity.
struct PointLayout {
  uint2* rect;         // [N] tile rectangles (pack_rect)
  uint2* rect_sorted;  // [N] the same in depth order
  uint32_t* dkey;      // [N]
  uint32_t* order;     // [N] Gaussians in depth order
  void* dsort_ws;      // depth sort
  size_t dsort_bytes;
  uint32_t* off;       // [N + 1] pair counts in depth order (the fill scans them)
  uint32_t* scan_ws;
  uint32_t* misc;      // [4]: n_eff
  size_t total;
};

struct PairLayout {
  uint32_t* tilekey;   // [cap]
  void* tsort_ws;      // tile sort
  size_t tsort_bytes;
  size_t total;
};

inline PointLayout point_layout(void* base, uint32_t n) {
  PointLayout L{};
  char* p = static_cast<char*>(base);
  size_t o = 0;
  auto take = [&](size_t bytes) { char* q = p ? p + o : nullptr; o += align256(bytes); return q; };
  L.rect = reinterpret_cast<uint2*>(take(size_t(n) * 8));
  L.rect_sorted = reinterpret_cast<uint2*>(take(size_t(n) * 8));
  L.dkey = reinterpret_cast<uint32_t*>(take(size_t(n) * 4));
  L.order = reinterpret_cast<uint32_t*>(take(size_t(n) * 4));
  L.dsort_bytes = sort_workspace_bytes<uint32_t>(n);
  L.dsort_ws = take(L.dsort_bytes);
  L.off = reinterpret_cast<uint32_t*>(take((size_t(n) + 1) * 4));
  L.scan_ws = reinterpret_cast<uint32_t*>(take(scan_chunks_workspace_bytes(n)));
  L.misc = reinterpret_cast<uint32_t*>(take(16));
  L.total = o;
  return L;
}

inline PairLayout pair_layout(void* base, uint32_t cap) {
  PairLayout L{};
  char* p = static_cast<char*>(base);
  size_t o = 0;
  auto take = [&](size_t bytes) { char* q = p ? p + o : nullptr; o += align256(bytes); return q; };
  L.tilekey = reinterpret_cast<uint32_t*>(take(size_t(cap) * 4));
  L.tsort_bytes = sort_workspace_bytes<uint32_t>(cap);
  L.tsort_ws = take(L.tsort_bytes);
  L.total = o;
  return L;
}

// Phase 1: per-Gaussian tile counts and depth keys (none when pixxydepth is
// null: count only); K = sum of the counts into *d_num_idx.
template <typename Boxes>
int bin_count(hipStream_t s, uint32_t n, Boxes boxes, const float* pixxydepth, uint32_t ts,
              uint32_t tw, uint32_t th, const PointLayout& L, uint32_t* d_num_idx) {
  if (tw > 0xFFFFu || th > 0xFFFFu)  // tile rectangles are packed as 16-bit coordinates
    return dsplat_set_error(DSPLAT_ERR_ARG, "binning: more than 65535 tiles along an axis");
  if (n == 0) {
    if (d_num_idx) set_u32_one<<<1, 1, 0, s>>>(d_num_idx, 0u);
  } else {
    // per-workgroup sums go to off[] (rewritten by the order phase); with no
    // d_num_idx, K is left to the order phase's scan
    const uint32_t nb = div_up(n, 256);
    bin_count_kernel<Boxes><<<nb, 256, 0, s>>>(n, boxes, pixxydepth, ts, tw, th, L.rect,
                                               pixxydepth ? L.dkey : nullptr,
                                               d_num_idx ? L.off : nullptr);
    if (d_num_idx) sum_partials_kernel<<<1, 256, 0, s>>>(nb, L.off, d_num_idx);
  }
  DSPLAT_HIP(hipGetLastError());
  return DSPLAT_OK;
}

// Depth keys are the bits of a non-negative float: 31 significant bits.
constexpr uint32_t kDepthBits = 31;

// Phase 2: stable depth order of the Gaussians (values generated as iota by
// the first pass; the last pass also gathers the tile rectangles and their
// areas, the pair counts, into depth order),
// then the first level of the exclusive scan of the counts (each Gaussian's
// first pair slot), per 256-Gaussian chunk: the fill finishes it inside its
// workgroup. d_total (may be null) receives the scan's total, K.
inline int bin_order(hipStream_t s, uint32_t n, const PointLayout& L, uint32_t* d_total = nullptr) {
  if (n == 0) {
    if (d_total) set_u32_one<<<1, 1, 0, s>>>(d_total, 0u);
    return DSPLAT_OK;
  }
  bool in_alt = false;
  uint32_t *kalt = nullptr, *valt = nullptr;
  int rc = radix_sort_pairs<uint32_t>(s, n, nullptr, L.dkey, L.order, kDepthBits, L.dsort_ws,
                                      L.dsort_bytes, &in_alt, &kalt, &valt, true, L.rect, L.off,
                                      L.rect_sorted);
  if (rc) return rc;
  if (in_alt)  // 31 key bits take 4 passes: the result is in place
    return dsplat_set_error(DSPLAT_ERR_ARG, "binning: depth order in the alternate buffer");
  return scan_chunk_prefixes(s, n, L.off, L.scan_ws, d_total);
}

// Phase 3: pairs in depth order, stable sort by tile id, tile ranges. Writes
// nothing into idx2pnt (and empty ranges) when K > capacity.
template <typename Boxes>
int bin_fill(hipStream_t s, uint32_t n, Boxes boxes, uint32_t ts, uint32_t tw, uint32_t th,
             const PointLayout& L, const uint32_t* d_num_idx, uint32_t* tile2idx,
             uint32_t* idx2pnt, uint32_t capacity, const PairLayout& Q) {
  const uint32_t nt = tw * th;
  const uint32_t tbits = bit_width(nt - 1), ib = bit_width(n > 1 ? n - 1 : 0);
  if (capacity > 0 && nt > 1 && tbits + ib <= 32) {
    // packed pairs, tile << ib | Gaussian index: the sort moves one word
    // per pair and its last pass writes the indices into idx2pnt
    if (n == 0) {
      set_u32_one<<<1, 1, 0, s>>>(L.misc, 0u);
    } else {
      bin_fill_kernel<<<div_up(n, kScanBlock), kScanBlock, 0, s>>>(
          n, L.order, L.rect_sorted, L.off, L.scan_ws, d_num_idx, capacity, L.misc, tw,
          Q.tilekey, nullptr, ib);
    }
    bool in_alt = false;
    uint32_t* kalt = nullptr;
    const int rc = radix_sort_keys(s, capacity, L.misc, Q.tilekey, ib, tbits,
                                   ib ? (1u << ib) - 1u : 0u, idx2pnt, Q.tsort_ws, Q.tsort_bytes,
                                   &in_alt, &kalt);
    if (rc) return rc;
    const uint32_t* ksorted = in_alt ? kalt : Q.tilekey;
    if (reinterpret_cast<uintptr_t>(ksorted) & 15u)
      return dsplat_set_error(DSPLAT_ERR_ARG, "binning: sorted pairs not 16-byte aligned");
    tile_ranges_kernel<<<div_up(uint64_t(capacity) + 1, 1024), 256, 0, s>>>(
        L.misc, capacity, nt, ksorted, ib, tile2idx);
    DSPLAT_HIP(hipGetLastError());
    return DSPLAT_OK;
  }
  if (n == 0) {
    set_u32_one<<<1, 1, 0, s>>>(L.misc, 0u);
  } else {
    bin_fill_kernel<<<div_up(n, kScanBlock), kScanBlock, 0, s>>>(
        n, L.order, L.rect_sorted, L.off, L.scan_ws, d_num_idx, capacity, L.misc, tw, Q.tilekey,
        idx2pnt, 0u);
  }
  const uint32_t* tsorted = Q.tilekey;
  if (capacity > 0 && nt > 1) {
    bool in_alt = false;
    uint32_t *kalt = nullptr, *valt = nullptr;
    int rc = radix_sort_pairs<uint32_t>(s, capacity, L.misc, Q.tilekey, idx2pnt,
                                        bit_width(nt - 1), Q.tsort_ws, Q.tsort_bytes, &in_alt,
                                        &kalt, &valt);
    if (rc) return rc;
    if (in_alt) {
      copy_u32_kernel<<<1024, 256, 0, s>>>(L.misc, capacity, valt, idx2pnt);
      tsorted = kalt;
    }
  }
  if (reinterpret_cast<uintptr_t>(tsorted) & 15u)
    return dsplat_set_error(DSPLAT_ERR_ARG, "binning: sorted tile ids not 16-byte aligned");
  tile_ranges_kernel<<<div_up(uint64_t(capacity) + 1, 1024), 256, 0, s>>>(L.misc, capacity, nt,
                                                                          tsorted, 0u, tile2idx);
  DSPLAT_HIP(hipGetLastError());
  return DSPLAT_OK;
}

// All phases with one workspace [points | pairs].
template <typename Boxes>
int bin_and_sort(hipStream_t s, uint32_t n, Boxes boxes, const float* pixxydepth, uint32_t ts,
                 uint32_t tw, uint32_t th, uint32_t* tile2idx, uint32_t* idx2pnt,
                 uint32_t capacity, void* ws, size_t ws_bytes, uint32_t* d_num_idx) {
  const PointLayout L = point_layout(ws, n);
  const PairLayout Q = pair_layout(static_cast<char*>(ws) + L.total, capacity);
  if (ws_bytes < L.total + Q.total)
    return dsplat_set_error(DSPLAT_ERR_CAPACITY, "binning: workspace too small");
  // K from the scan of the depth-ordered counts (no separate sum)
  int rc = bin_count(s, n, boxes, pixxydepth, ts, tw, th, L, nullptr);
  if (!rc) rc = bin_order(s, n, L, d_num_idx);
  if (!rc) rc = bin_fill(s, n, boxes, ts, tw, th, L, d_num_idx, tile2idx, idx2pnt, capacity, Q);
  return rc;
}

size_t bin_workspace_bytes(uint32_t n, uint32_t capacity) {
  return point_layout(nullptr, n).total + pair_layout(nullptr, capacity).total;
}

void bin_count_slots(void* ws, uint32_t n, uint2** rect, uint32_t** dkey) {
  const PointLayout L = point_layout(ws, n);
  *rect = L.rect;
  *dkey = L.dkey;
}

int bin_after_count(hipStream_t s, uint32_t n, uint32_t ts, uint32_t tw, uint32_t th,
                    uint32_t* tile2idx, uint32_t* idx2pnt, uint32_t capacity, void* ws,
                    size_t ws_bytes, uint32_t* d_num_idx) {
  const PointLayout L = point_layout(ws, n);
  const PairLayout Q = pair_layout(static_cast<char*>(ws) + L.total, capacity);
  if (ws_bytes < L.total + Q.total)
    return dsplat_set_error(DSPLAT_ERR_CAPACITY, "binning: workspace too small");
  int rc = bin_order(s, n, L, d_num_idx);
  if (!rc)
    rc = bin_fill(s, n, GaussBoxes{nullptr}, ts, tw, th, L, d_num_idx, tile2idx, idx2pnt,
                  capacity, Q);
  return rc;
}

int check_tiles(uint32_t ts) {
  if (ts != 8 && ts != 16 && ts != 32)
    return dsplat_set_error(DSPLAT_ERR_ARG, "tile_size must be 8, 16 or 32");
  return DSPLAT_OK;
}

// Library-allocated idx2pnt: K is read back (splat_gauss.rs:137-147) right
// after the count, while the depth order is already queued behind it.
template <typename Boxes>
int bin_alloc(hipStream_t s, uint32_t n, Boxes boxes, const float* pixxydepth, uint32_t ts,
              uint32_t tw, uint32_t th, uint32_t* tile2idx, uint32_t** idx2pnt, uint32_t* num_idx) {
  const size_t pbytes = point_layout(nullptr, n).total;
  char* ws0 = nullptr;
  DSPLAT_HIP(hipMallocAsync(reinterpret_cast<void**>(&ws0), pbytes + 256, s));
  const PointLayout L = point_layout(ws0, n);
  uint32_t* dk = reinterpret_cast<uint32_t*>(ws0 + pbytes);
  uint32_t k = 0;
  hipEvent_t ev = nullptr;
  int rc = bin_count(s, n, boxes, pixxydepth, ts, tw, th, L, dk);
  if (!rc) {
    hipError_t e = hipMemcpyAsync(&k, dk, 4, hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipEventCreateWithFlags(&ev, hipEventDisableTiming);
    if (e == hipSuccess) e = hipEventRecord(ev, s);
    if (e != hipSuccess) rc = dsplat_set_hip_error(e, "binning: read back K");
  }
  if (!rc) rc = bin_order(s, n, L);
  if (!rc) {
    const hipError_t e = hipEventSynchronize(ev);
    if (e != hipSuccess) rc = dsplat_set_hip_error(e, "binning: read back K");
  }
  if (ev) (void)hipEventDestroy(ev);
  uint32_t* out = nullptr;
  void* ws1 = nullptr;
  const size_t qbytes = pair_layout(nullptr, k).total;
  if (!rc) {
    hipError_t e = hipMallocAsync(reinterpret_cast<void**>(&out), size_t(k ? k : 1) * 4, s);
    if (e == hipSuccess) e = hipMallocAsync(&ws1, qbytes ? qbytes : 256, s);
    if (e != hipSuccess) rc = dsplat_set_hip_error(e, "binning: allocate idx2pnt");
  }
  if (!rc)
    rc = bin_fill(s, n, boxes, ts, tw, th, L, dk, tile2idx, out, k, pair_layout(ws1, k));
  if (ws1) (void)hipFreeAsync(ws1, s);
  (void)hipFreeAsync(ws0, s);
  if (rc) {
    if (out) (void)hipFreeAsync(out, s);
    return rc;
  }
  *idx2pnt = out;
  *num_idx = k;
  return DSPLAT_OK;
}

}  // namespace dsplat

using namespace dsplat;

extern "C" size_t dsplat_tile2idx_idx2pnt_workspace_size(uint32_t num_pnt, uint32_t num_tile,
                                                         uint32_t capacity) {
  (void)num_tile;
  return point_layout(nullptr, num_pnt).total + pair_layout(nullptr, capacity).total;
}

extern "C" size_t dsplat_bin_points_workspace_size(uint32_t num_pnt) {
  return point_layout(nullptr, num_pnt).total;
}

extern "C" size_t dsplat_bin_pairs_workspace_size(uint32_t capacity) {
  return pair_layout(nullptr, capacity).total;
}

extern "C" int dsplat_gauss_count_pairs(void* stream, uint32_t img_w, uint32_t img_h,
                                        uint32_t tile_size, uint32_t num_pnt,
                                        const float* pnt2pixaabb, uint32_t* d_num_idx) {
  if (int rc = check_tiles(tile_size)) return rc;
  if (!d_num_idx || (num_pnt && !pnt2pixaabb))
    return dsplat_set_error(DSPLAT_ERR_ARG, "count_pairs: null buffer");
  if (reinterpret_cast<uintptr_t>(pnt2pixaabb) & 15u)
    return dsplat_set_error(DSPLAT_ERR_ARG, "count_pairs: aabb must be 16-byte aligned");
  hipStream_t s = as_stream(stream);
  uint32_t tw, th;
  tile_grid(img_w, img_h, tile_size, &tw, &th);
  void* ws = nullptr;
  const size_t b = point_layout(nullptr, num_pnt).total;
  DSPLAT_HIP(hipMallocAsync(&ws, b, s));
  const PointLayout L = point_layout(ws, num_pnt);
  const int rc = bin_count(s, num_pnt, GaussBoxes{reinterpret_cast<const float4*>(pnt2pixaabb)},
                           nullptr, tile_size, tw, th, L, d_num_idx);
  (void)hipFreeAsync(ws, s);
  return rc;
}

static int check_points_ws(uint32_t n, const void* ws, size_t bytes) {
  if (!ws || bytes < point_layout(nullptr, n).total)
    return dsplat_set_error(DSPLAT_ERR_CAPACITY, "binning: points workspace too small");
  return DSPLAT_OK;
}

extern "C" int dsplat_gauss_bin_count(void* stream, uint32_t img_w, uint32_t img_h,
                                      uint32_t tile_size, uint32_t num_pnt,
                                      const float* pnt2pixxydepth, const float* pnt2pixaabb,
                                      void* ws_points, size_t ws_points_bytes,
                                      uint32_t* d_num_idx) {
  if (int rc = check_tiles(tile_size)) return rc;
  if (!d_num_idx || (num_pnt && (!pnt2pixxydepth || !pnt2pixaabb)))
    return dsplat_set_error(DSPLAT_ERR_ARG, "bin_count: null buffer");
  if (reinterpret_cast<uintptr_t>(pnt2pixaabb) & 15u)
    return dsplat_set_error(DSPLAT_ERR_ARG, "bin_count: aabb must be 16-byte aligned");
  if (int rc = check_points_ws(num_pnt, ws_points, ws_points_bytes)) return rc;
  uint32_t tw, th;
  tile_grid(img_w, img_h, tile_size, &tw, &th);
  return bin_count(as_stream(stream), num_pnt,
                   GaussBoxes{reinterpret_cast<const float4*>(pnt2pixaabb)}, pnt2pixxydepth,
                   tile_size, tw, th, point_layout(ws_points, num_pnt), d_num_idx);
}

extern "C" int dsplat_bin_order(void* stream, uint32_t num_pnt, void* ws_points,
                                size_t ws_points_bytes) {
  if (int rc = check_points_ws(num_pnt, ws_points, ws_points_bytes)) return rc;
  return bin_order(as_stream(stream), num_pnt, point_layout(ws_points, num_pnt));
}

extern "C" int dsplat_gauss_bin_fill(void* stream, uint32_t img_w, uint32_t img_h,
                                     uint32_t tile_size, uint32_t num_pnt,
                                     const float* pnt2pixaabb, void* ws_points,
                                     size_t ws_points_bytes, const uint32_t* d_num_idx,
                                     uint32_t* tile2idx, uint32_t* idx2pnt, uint32_t capacity,
                                     void* ws_pairs, size_t ws_pairs_bytes) {
  if (int rc = check_tiles(tile_size)) return rc;
  if (!tile2idx || !d_num_idx || (num_pnt && !pnt2pixaabb) || (capacity && !idx2pnt))
    return dsplat_set_error(DSPLAT_ERR_ARG, "bin_fill: null buffer");
  if (int rc = check_points_ws(num_pnt, ws_points, ws_points_bytes)) return rc;
  if (!ws_pairs || ws_pairs_bytes < pair_layout(nullptr, capacity).total)
    return dsplat_set_error(DSPLAT_ERR_CAPACITY, "bin_fill: pairs workspace too small");
  uint32_t tw, th;
  tile_grid(img_w, img_h, tile_size, &tw, &th);
  return bin_fill(as_stream(stream), num_pnt,
                  GaussBoxes{reinterpret_cast<const float4*>(pnt2pixaabb)}, tile_size, tw, th,
                  point_layout(ws_points, num_pnt), d_num_idx, tile2idx, idx2pnt, capacity,
                  pair_layout(ws_pairs, capacity));
}

extern "C" int dsplat_gauss_tile2idx_idx2pnt_ws(void* stream, uint32_t img_w, uint32_t img_h,
                                                uint32_t tile_size, uint32_t num_pnt,
                                                const float* pnt2pixxydepth,
                                                const float* pnt2pixaabb, uint32_t* tile2idx,
                                                uint32_t* idx2pnt, uint32_t capacity,
                                                void* workspace, size_t workspace_bytes,
                                                uint32_t* d_num_idx) {
  if (int rc = check_tiles(tile_size)) return rc;
  if (!tile2idx || !d_num_idx || (num_pnt && (!pnt2pixxydepth || !pnt2pixaabb)) ||
      (capacity && !idx2pnt) || !workspace)
    return dsplat_set_error(DSPLAT_ERR_ARG, "tile2idx_idx2pnt: null buffer");
  if (reinterpret_cast<uintptr_t>(pnt2pixaabb) & 15u)
    return dsplat_set_error(DSPLAT_ERR_ARG, "tile2idx_idx2pnt: aabb must be 16-byte aligned");
  uint32_t tw, th;
  tile_grid(img_w, img_h, tile_size, &tw, &th);
  return bin_and_sort(as_stream(stream), num_pnt,
                      GaussBoxes{reinterpret_cast<const float4*>(pnt2pixaabb)}, pnt2pixxydepth,
                      tile_size, tw, th, tile2idx, idx2pnt, capacity, workspace, workspace_bytes,
                      d_num_idx);
}

// ------------------------------------------------- allocation-free binning ---
// The workspace of dsplat_gauss_tile2idx_idx2pnt_noalloc: a file-scope device
// array of DSPLAT_NOALLOC_SCRATCH_BYTES, one per GPU (each GPU that loads the
// library holds its copy until the process ends). Host code never passes the
// array's name to a kernel (in host code it is a stand-in, not a device
// address): the device address comes from hipGetSymbolAddress, taken once per
// device with that device current, and is kept per device.
static __device__ __align__(256) unsigned char scratch[DSPLAT_NOALLOC_SCRATCH_BYTES];

namespace {
constexpr int kMaxScratchDevices = 64;
std::mutex g_scratch_mu;
void* g_scratch[kMaxScratchDevices] = {};
}  // namespace

static int scratch_address(void** out) {
  int dev = 0;
  DSPLAT_HIP(hipGetDevice(&dev));
  if (dev < 0 || dev >= kMaxScratchDevices)
    return dsplat_set_error(DSPLAT_ERR_ARG, "noalloc binning: device index out of range");
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch[dev]) {
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p)
      return dsplat_set_hip_error(e != hipSuccess ? e : hipErrorInvalidSymbol,
                                  "noalloc binning: hipGetSymbolAddress(scratch)");
    g_scratch[dev] = p;
  }
  *out = g_scratch[dev];
  return DSPLAT_OK;
}

extern "C" int dsplat_gauss_tile2idx_idx2pnt_noalloc(void* stream, uint32_t img_w,
                                                     uint32_t img_h, uint32_t tile_size,
                                                     uint32_t num_pnt,
                                                     const float* pnt2pixxydepth,
                                                     const float* pnt2pixaabb,
                                                     uint32_t* tile2idx, uint32_t* idx2pnt,
                                                     uint32_t capacity, uint32_t* d_num_idx) {
  if (int rc = check_tiles(tile_size)) return rc;
  if (!tile2idx || !d_num_idx || (num_pnt && (!pnt2pixxydepth || !pnt2pixaabb)) ||
      (capacity && !idx2pnt))
    return dsplat_set_error(DSPLAT_ERR_ARG, "tile2idx_idx2pnt_noalloc: null buffer");
  if (reinterpret_cast<uintptr_t>(pnt2pixaabb) & 15u)
    return dsplat_set_error(DSPLAT_ERR_ARG,
                            "tile2idx_idx2pnt_noalloc: aabb must be 16-byte aligned");
  const size_t need = dsplat_tile2idx_idx2pnt_workspace_size(num_pnt, 0, capacity);
  if (need > size_t(DSPLAT_NOALLOC_SCRATCH_BYTES))
    return dsplat_set_error(DSPLAT_ERR_CAPACITY,
                            "tile2idx_idx2pnt_noalloc: this many Gaussians and pair capacity "
                            "need more workspace than the static scratch holds "
                            "(DSPLAT_NOALLOC_SCRATCH_BYTES); use dsplat_gauss_tile2idx_idx2pnt_ws");
  void* ws = nullptr;
  if (int rc = scratch_address(&ws)) return rc;
  uint32_t tw, th;
  tile_grid(img_w, img_h, tile_size, &tw, &th);
  // every kernel of the path initialises what it reads of the workspace, so
  // the previous call's contents need no clearing
  return bin_and_sort(as_stream(stream), num_pnt,
                      GaussBoxes{reinterpret_cast<const float4*>(pnt2pixaabb)}, pnt2pixxydepth,
                      tile_size, tw, th, tile2idx, idx2pnt, capacity, ws,
                      size_t(DSPLAT_NOALLOC_SCRATCH_BYTES), d_num_idx);
}

extern "C" int dsplat_gauss_tile2idx_idx2pnt(void* stream, uint32_t img_w, uint32_t img_h,
                                             uint32_t tile_size, uint32_t num_pnt,
                                             const float* pnt2pixxydepth,
                                             const float* pnt2pixaabb, uint32_t* tile2idx,
                                             uint32_t** idx2pnt, uint32_t* num_idx) {
  if (int rc = check_tiles(tile_size)) return rc;
  if (!tile2idx || !idx2pnt || !num_idx || (num_pnt && (!pnt2pixxydepth || !pnt2pixaabb)))
    return dsplat_set_error(DSPLAT_ERR_ARG, "tile2idx_idx2pnt: null buffer");
  if (reinterpret_cast<uintptr_t>(pnt2pixaabb) & 15u)
    return dsplat_set_error(DSPLAT_ERR_ARG, "tile2idx_idx2pnt: aabb must be 16-byte aligned");
  uint32_t tw, th;
  tile_grid(img_w, img_h, tile_size, &tw, &th);
  return bin_alloc(as_stream(stream), num_pnt,
                   GaussBoxes{reinterpret_cast<const float4*>(pnt2pixaabb)}, pnt2pixxydepth,
                   tile_size, tw, th, tile2idx, idx2pnt, num_idx);
}

// ------------------------------------------------------------ circle path ---

extern "C" int dsplat_circle_bin_count(void* stream, uint32_t tile_w, uint32_t tile_h,
                                       uint32_t tile_size, uint32_t num_pnt,
                                       const float* pnt2pixxydepth, const float* pnt2pixrad,
                                       void* ws_points, size_t ws_points_bytes,
                                       uint32_t* d_num_idx) {
  if (int rc = check_tiles(tile_size)) return rc;
  if (!d_num_idx || (num_pnt && (!pnt2pixxydepth || !pnt2pixrad)))
    return dsplat_set_error(DSPLAT_ERR_ARG, "circle bin_count: null buffer");
  if (int rc = check_points_ws(num_pnt, ws_points, ws_points_bytes)) return rc;
  return bin_count(as_stream(stream), num_pnt, CircleBoxes{pnt2pixxydepth, pnt2pixrad},
                   pnt2pixxydepth, tile_size, tile_w, tile_h, point_layout(ws_points, num_pnt),
                   d_num_idx);
}

extern "C" int dsplat_circle_bin_fill(void* stream, uint32_t tile_w, uint32_t tile_h,
                                      uint32_t tile_size, uint32_t num_pnt,
                                      const float* pnt2pixxydepth, const float* pnt2pixrad,
                                      void* ws_points, size_t ws_points_bytes,
                                      const uint32_t* d_num_idx, uint32_t* tile2idx,
                                      uint32_t* idx2pnt, uint32_t capacity, void* ws_pairs,
                                      size_t ws_pairs_bytes) {
  if (int rc = check_tiles(tile_size)) return rc;
  if (!tile2idx || !d_num_idx || (num_pnt && (!pnt2pixxydepth || !pnt2pixrad)) ||
      (capacity && !idx2pnt))
    return dsplat_set_error(DSPLAT_ERR_ARG, "circle bin_fill: null buffer");
  if (int rc = check_points_ws(num_pnt, ws_points, ws_points_bytes)) return rc;
  if (!ws_pairs || ws_pairs_bytes < pair_layout(nullptr, capacity).total)
    return dsplat_set_error(DSPLAT_ERR_CAPACITY, "circle bin_fill: pairs workspace too small");
  return bin_fill(as_stream(stream), num_pnt, CircleBoxes{pnt2pixxydepth, pnt2pixrad},
                  tile_size, tile_w, tile_h, point_layout(ws_points, num_pnt), d_num_idx,
                  tile2idx, idx2pnt, capacity, pair_layout(ws_pairs, capacity));
}

extern "C" int dsplat_circle_tile2idx_idx2pnt_ws(void* stream, uint32_t tile_w, uint32_t tile_h,
                                                 uint32_t tile_size, uint32_t num_pnt,
                                                 const float* pnt2pixxydepth,
                                                 const float* pnt2pixrad, uint32_t* tile2idx,
                                                 uint32_t* idx2pnt, uint32_t capacity,
                                                 void* workspace, size_t workspace_bytes,
                                                 uint32_t* d_num_idx) {
  if (int rc = check_tiles(tile_size)) return rc;
  if (!tile2idx || !d_num_idx || (num_pnt && (!pnt2pixxydepth || !pnt2pixrad)) ||
      (capacity && !idx2pnt) || !workspace)
    return dsplat_set_error(DSPLAT_ERR_ARG, "circle tile2idx_idx2pnt: null buffer");
  return bin_and_sort(as_stream(stream), num_pnt, CircleBoxes{pnt2pixxydepth, pnt2pixrad},
                      pnt2pixxydepth, tile_size, tile_w, tile_h, tile2idx, idx2pnt, capacity,
                      workspace, workspace_bytes, d_num_idx);
}

extern "C" int dsplat_circle_tile2idx_idx2pnt(void* stream, uint32_t tile_w, uint32_t tile_h,
                                              uint32_t tile_size, uint32_t num_pnt,
                                              const float* pnt2pixxydepth,
                                              const float* pnt2pixrad, uint32_t* tile2idx,
                                              uint32_t** idx2pnt, uint32_t* num_idx) {
  if (int rc = check_tiles(tile_size)) return rc;
  if (!tile2idx || !idx2pnt || !num_idx || (num_pnt && (!pnt2pixxydepth || !pnt2pixrad)))
    return dsplat_set_error(DSPLAT_ERR_ARG, "circle tile2idx_idx2pnt: null buffer");
  return bin_alloc(as_stream(stream), num_pnt, CircleBoxes{pnt2pixxydepth, pnt2pixrad},
                   pnt2pixxydepth, tile_size, tile_w, tile_h, tile2idx, idx2pnt, num_idx);
}
