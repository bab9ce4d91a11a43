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

template <typename Boxes>
__global__ void __launch_bounds__(256)
bin_count_kernel(uint32_t n, Boxes boxes, const float* __restrict__ pixxydepth, uint32_t ts,
                 uint32_t tw, uint32_t th, uint32_t* __restrict__ cnt,
                 uint32_t* __restrict__ dkey) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float4 b = boxes(i);
  const TileRect r = tile_rect(b.x, b.y, b.z, b.w, ts, tw, th);
  cnt[i] = uint32_t(r.count());
  if (dkey) dkey[i] = depth_key(pixxydepth[3 * i + 2]);
}

// cnt_sorted[j] = cnt[order[j]]
__global__ void __launch_bounds__(256)
gather_u32_kernel(uint32_t n, const uint32_t* __restrict__ order, const uint32_t* __restrict__ src,
                  uint32_t* __restrict__ dst) {
  const uint32_t j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j < n) dst[j] = src[order[j]];
}

// n_eff = K if K <= capacity else 0 (overflow: downstream kernels see nothing)
__global__ void clamp_count_kernel(const uint32_t* __restrict__ d_k, uint32_t capacity,
                                   uint32_t* __restrict__ n_eff) {
  const uint32_t k = *d_k;
  *n_eff = k <= capacity ? k : 0u;
}

// Pairs of the j-th Gaussian in depth order at off[j]: tile ids ix-outer,
// iy-inner (tile_acceleration.h:29-50), value = Gaussian index.
template <typename Boxes>
__global__ void __launch_bounds__(256)
bin_fill_kernel(uint32_t n, Boxes boxes, const uint32_t* __restrict__ order,
                const uint32_t* __restrict__ off, const uint32_t* __restrict__ n_eff,
                uint32_t ts, uint32_t tw, uint32_t th, uint32_t* __restrict__ tile_of_pair,
                uint32_t* __restrict__ pnt_of_pair) {
  __shared__ ExpandLds lds[4];
  if (*n_eff == 0) return;  // uniform: overflow, nothing is written
  const uint32_t j = blockIdx.x * blockDim.x + threadIdx.x;
  const bool valid = j < n;
  uint32_t g = 0, o = 0;
  TileRect r{0, 0, 0, 0};
  if (valid) {
    g = order[j];
    const float4 b = boxes(g);
    r = tile_rect(b.x, b.y, b.z, b.w, ts, tw, th);
    o = off[j];
  }
  wave_expand_pairs(valid, o, r, 0u, tw, g, lds[threadIdx.x >> 6], tile_of_pair, pnt_of_pair);
}

// tile2idx[t] = first sorted position whose tile id is >= t, t in [0, T].
__global__ void __launch_bounds__(256)
tile_ranges_kernel(const uint32_t* __restrict__ n_eff, uint32_t capacity, uint32_t num_tile,
                   const uint32_t* __restrict__ tile_sorted, uint32_t* __restrict__ tile2idx) {
  const uint32_t k = *n_eff;
  const uint32_t idx = blockIdx.x * blockDim.x + threadIdx.x;
  if (idx > k || idx > capacity) return;
  const uint32_t lo = idx == 0 ? 0u : tile_sorted[idx - 1] + 1u;
  const uint32_t hi = idx == k ? num_tile : tile_sorted[idx];
  for (uint32_t t = lo; t <= hi; ++t) tile2idx[t] = idx;
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

struct BinLayout {
  uint32_t* cnt;       // [N]
  uint32_t* dkey;      // [N]
  uint32_t* order;     // [N]
  void* dsort_ws;      // depth sort
  size_t dsort_bytes;
  uint32_t* off;       // [N+1]
  uint32_t* scan_ws;
  uint32_t* tilekey;   // [cap]
  void* tsort_ws;      // tile sort
  size_t tsort_bytes;
  uint32_t* misc;      // [4]: n_eff
  size_t total;
};

inline size_t align256(size_t v) { return (v + 255) & ~size_t(255); }

inline BinLayout bin_layout(void* base, uint32_t n, uint32_t cap) {
  BinLayout L{};
  char* p = static_cast<char*>(base);
  size_t o = 0;
  auto take = [&](size_t bytes) { char* q = p ? p + o : nullptr; o += align256(bytes); return q; };
  L.cnt = reinterpret_cast<uint32_t*>(take(size_t(n) * 4));
  L.dkey = reinterpret_cast<uint32_t*>(take(size_t(n) * 4));
  L.order = reinterpret_cast<uint32_t*>(take(size_t(n) * 4));
  L.dsort_bytes = sort_workspace_bytes<uint32_t>(n);
  L.dsort_ws = take(L.dsort_bytes);
  L.off = reinterpret_cast<uint32_t*>(take((size_t(n) + 1) * 4));
  L.scan_ws = reinterpret_cast<uint32_t*>(take(scan_workspace_bytes(n)));
  L.tilekey = reinterpret_cast<uint32_t*>(take(size_t(cap) * 4));
  L.tsort_bytes = sort_workspace_bytes<uint32_t>(cap);
  L.tsort_ws = take(L.tsort_bytes);
  L.misc = reinterpret_cast<uint32_t*>(take(16));
  L.total = o;
  return L;
}

template <typename Boxes>
int count_pairs(hipStream_t s, uint32_t n, Boxes boxes, uint32_t ts, uint32_t tw, uint32_t th,
                uint32_t* cnt, uint32_t* scan_ws, uint32_t* d_num_idx) {
  if (n == 0) {
    set_u32_one<<<1, 1, 0, s>>>(d_num_idx, 0u);
    DSPLAT_HIP(hipGetLastError());
    return DSPLAT_OK;
  }
  bin_count_kernel<Boxes><<<div_up(n, 256), 256, 0, s>>>(n, boxes, nullptr, ts, tw, th, cnt,
                                                         nullptr);
  // reduce through the scan (out == in is allowed; cnt is scratch here)
  int rc = exclusive_scan(s, n, nullptr, cnt, cnt, scan_ws, d_num_idx);
  if (rc) return rc;
  DSPLAT_HIP(hipGetLastError());
  return DSPLAT_OK;
}

template <typename Boxes>
int bin_and_sort(hipStream_t s, uint32_t n, Boxes boxes, const float* pixxydepth, uint32_t ts,
                 uint32_t tw, uint32_t th, uint32_t* tile2idx, uint32_t* idx2pnt,
                 uint32_t capacity, void* ws, size_t ws_bytes, uint32_t* d_num_idx) {
  const uint32_t nt = tw * th;
  BinLayout L = bin_layout(ws, n, capacity);
  if (ws_bytes < L.total)
    return dsplat_set_error(DSPLAT_ERR_CAPACITY, "binning: workspace too small");
  if (n == 0) {
    set_u32_one<<<1, 1, 0, s>>>(d_num_idx, 0u);
    set_u32_one<<<1, 1, 0, s>>>(L.misc, 0u);
  } else {
    const uint32_t nb = div_up(n, 256);
    bin_count_kernel<Boxes><<<nb, 256, 0, s>>>(n, boxes, pixxydepth, ts, tw, th, L.cnt, L.dkey);
    bool in_alt = false;
    uint32_t *kalt = nullptr, *valt = nullptr;
    // values: Gaussian indices (iota), generated by the first pass
    int rc = radix_sort_pairs<uint32_t>(s, n, nullptr, L.dkey, L.order, 31, L.dsort_ws,
                                        L.dsort_bytes, &in_alt, &kalt, &valt, true);
    if (rc) return rc;
    const uint32_t* order = in_alt ? valt : L.order;
    gather_u32_kernel<<<nb, 256, 0, s>>>(n, order, L.cnt, L.off);
    rc = exclusive_scan(s, n, nullptr, L.off, L.off, L.scan_ws, d_num_idx);
    if (rc) return rc;
    clamp_count_kernel<<<1, 1, 0, s>>>(d_num_idx, capacity, L.misc);
    bin_fill_kernel<Boxes><<<nb, 256, 0, s>>>(n, boxes, order, L.off, L.misc, ts, tw, th,
                                              L.tilekey, idx2pnt);
  }
  if (capacity > 0 && nt > 1) {
    bool in_alt = false;
    uint32_t *kalt = nullptr, *valt = nullptr;
    int rc = radix_sort_pairs<uint32_t>(s, capacity, L.misc, L.tilekey, idx2pnt,
                                        bit_width(nt - 1), L.tsort_ws, L.tsort_bytes, &in_alt,
                                        &kalt, &valt);
    if (rc) return rc;
    const uint32_t* tsorted = L.tilekey;
    if (in_alt) {
      copy_u32_kernel<<<1024, 256, 0, s>>>(L.misc, capacity, valt, idx2pnt);
      tsorted = kalt;
    }
    tile_ranges_kernel<<<div_up(uint64_t(capacity) + 1, 256), 256, 0, s>>>(L.misc, capacity, nt,
                                                                           tsorted, tile2idx);
  } else {
    tile_ranges_kernel<<<div_up(uint64_t(capacity) + 1, 256), 256, 0, s>>>(L.misc, capacity, nt,
                                                                           L.tilekey, tile2idx);
  }
  DSPLAT_HIP(hipGetLastError());
  return DSPLAT_OK;
}

int check_tiles(uint32_t ts) {
  if (ts != 8 && ts != 16 && ts != 32)
    return dsplat_set_error(DSPLAT_ERR_ARG, "tile_size must be 8, 16 or 32");
  return DSPLAT_OK;
}

template <typename Boxes>
int bin_alloc(hipStream_t s, uint32_t n, Boxes boxes, const float* pixxydepth, uint32_t ts,
              uint32_t tw, uint32_t th, uint32_t* tile2idx, uint32_t** idx2pnt, uint32_t* num_idx) {
  // K first, read back with a blocking copy as splat_gauss.rs:137-147 does
  const size_t off_scan = align256(size_t(n) * 4);
  const size_t off_k = off_scan + align256(scan_workspace_bytes(n));
  char* ws0 = nullptr;
  DSPLAT_HIP(hipMallocAsync(reinterpret_cast<void**>(&ws0), off_k + 256, s));
  uint32_t* dk = reinterpret_cast<uint32_t*>(ws0 + off_k);
  int rc = count_pairs(s, n, boxes, ts, tw, th, reinterpret_cast<uint32_t*>(ws0),
                       reinterpret_cast<uint32_t*>(ws0 + off_scan), dk);
  uint32_t k = 0;
  if (!rc) {
    hipError_t e = hipMemcpyAsync(&k, dk, 4, hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    if (e != hipSuccess) rc = dsplat_set_hip_error(e, "binning: read back K");
  }
  uint32_t* out = nullptr;
  void* ws = nullptr;
  const size_t wsb = bin_layout(nullptr, n, k).total;
  if (!rc) {
    hipError_t e = hipMallocAsync(reinterpret_cast<void**>(&out), size_t(k ? k : 1) * 4, s);
    if (e == hipSuccess) e = hipMallocAsync(&ws, wsb, s);
    if (e != hipSuccess) rc = dsplat_set_hip_error(e, "binning: allocate idx2pnt");
  }
  if (!rc) rc = bin_and_sort(s, n, boxes, pixxydepth, ts, tw, th, tile2idx, out, k, ws, wsb, dk);
  if (ws) (void)hipFreeAsync(ws, s);
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
  return bin_layout(nullptr, num_pnt, capacity).total;
}

extern "C" int dsplat_gauss_count_pairs(void* stream, uint32_t img_w, uint32_t img_h,
                                        uint32_t tile_size, uint32_t num_pnt,
                                        const float* pnt2pixaabb, uint32_t* d_num_idx) {
  if (int rc = check_tiles(tile_size)) return rc;
  if (!d_num_idx || (num_pnt && !pnt2pixaabb))
    return dsplat_set_error(DSPLAT_ERR_ARG, "count_pairs: null buffer");
  hipStream_t s = as_stream(stream);
  uint32_t tw, th;
  tile_grid(img_w, img_h, tile_size, &tw, &th);
  void* ws = nullptr;
  const size_t b = align256(size_t(num_pnt) * 4) + scan_workspace_bytes(num_pnt);
  DSPLAT_HIP(hipMallocAsync(&ws, b, s));
  char* p = static_cast<char*>(ws);
  int rc = count_pairs(s, num_pnt, GaussBoxes{reinterpret_cast<const float4*>(pnt2pixaabb)},
                       tile_size, tw, th, reinterpret_cast<uint32_t*>(p),
                       reinterpret_cast<uint32_t*>(p + align256(size_t(num_pnt) * 4)), d_num_idx);
  (void)hipFreeAsync(ws, s);
  return rc;
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
