// Batched multi-view rendering and differentiation over one shared set of
// Gaussians (no counterpart in the reference, which renders one view per
// call: del-splat-cudarc/src/splat_gauss.rs).
//
// A plan owns every device buffer for N Gaussians and up to V views, so a
// frame is a fixed sequence of kernels with no allocation and no host
// synchronisation; it can be captured into a hipGraph and replayed.
//
// Items: (view v, Gaussian i) -> item v*N + i. Per item the projection kernel
// writes the packed raster record and the binning counters in one pass. The
// binning is the single-view algorithm applied to all items at once: one
// stable sort of the V*N depth keys, pair generation with *global* tile ids
// (view tile offset + local tile), one stable sort by global tile id. Within
// a (view, tile) the order is (depth bits, Gaussian index), the reference's
// per-view order (tile_acceleration.h:15-52 + stable u64 sort).
//
// Raster and adjoint run as one persistent launch each: workgroups pull
// global tiles from an atomic counter until all tiles of all views are done,
// so views of uneven resolution share one launch without tail gaps between
// per-view grids.
//
// Colour is either one rgb per Gaussian or, after dsplat_mv_set_sh,
// spherical-harmonic coefficients seen from each view's camera centre: a
// per-Gaussian kernel stages the coefficients once and writes one colour per
// item, the raster adjoint then leaves dL/drgb per item, and a per-Gaussian
// kernel folds the views back into dL/dsh and the direction part of dL/dxyz.
#include <stdlib.h>
#include <string.h>

#include <vector>

#include "primitives.hpp"
#include "project_math.hpp"
#include "raster_core.hpp"
#include "sh_core.hpp"

namespace dsplat {

struct ViewDev {
  float m[16];
  uint32_t w, h, tw, th;
  uint32_t tile_off, pad0;
  uint64_t pix_off;
  float cam[3];  // camera centre (world), for the SH colour
  float pad1;
};

constexpr uint32_t kMaxViews = 64;
static_assert(kMaxViews <= kSortMaxSegs, "one depth-sort segment per view");
// views of at least this many Gaussians are depth-sorted one by one (below,
// the padding to whole segments costs more than the locality saves)
constexpr uint32_t kSegSortMin = 1u << 18;

// Finds the view of a global tile (views are few; the scan is uniform across
// the workgroup).
__device__ inline uint32_t view_of_tile(const ViewDev* __restrict__ views, uint32_t nview,
                                        uint32_t g) {
  uint32_t v = 0;
  while (v + 1 < nview && views[v + 1].tile_off <= g) ++v;
  return v;
}

__device__ inline float sigmoidf_(float x) { return 1.f / (1.f + __expf(-x)); }

// ---------------------------------------------------------- projection ---
template <bool kLogScale, bool kLogitOp, bool kRgbItem>
__global__ void __launch_bounds__(256)
mv_project_kernel(uint32_t n, uint32_t seg, const ViewDev* __restrict__ views, uint32_t ts,
                  const float* __restrict__ xyz, const float* __restrict__ scale,
                  const float4* __restrict__ quat, const float* __restrict__ opacity,
                  const float* __restrict__ rgb, float4* __restrict__ rec,
                  uint2* __restrict__ rect, uint32_t* __restrict__ dkey) {
  // the records leave through LDS as contiguous 1 KB stores per wave (the
  // wave's 64 items are consecutive), as in gauss_project_kernel
  __shared__ float4 srec[4 * 256];
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  const uint32_t lane = threadIdx.x & 63u, wbase = threadIdx.x & ~63u;
  const uint32_t g0 = blockIdx.x * blockDim.x + wbase;
  const uint32_t v = blockIdx.y;
  // the view's depth-sort segment holds its n items and seg - n padding
  // entries: an empty rectangle (no pairs) under the largest key
  const size_t sp = size_t(v) * seg + i;
  if (i >= n && i < seg) {
    rect[sp] = make_uint2(0u, 0u);
    dkey[sp] = 0x7FFFFFFFu;
  }
  if (g0 >= n) return;  // wave-uniform
  const ViewDev& vd = views[v];
  if (i < n) {
    const float4 q = quat[i];
    float s0 = scale[3 * i], s1 = scale[3 * i + 1], s2 = scale[3 * i + 2];
    if (kLogScale) {
      s0 = expf(s0);
      s1 = expf(s1);
      s2 = expf(s2);
    }
    ProjState p;
    project_state(vd.m, xyz[3 * i], xyz[3 * i + 1], xyz[3 * i + 2], q.x, q.y, q.z, q.w, s0, s1,
                  s2, vd.w, vd.h, p);
    float op = opacity[i];
    if (kLogitOp) op = sigmoidf_(op);
    const size_t item = size_t(v) * n + i;
    const float* col = rgb + 3 * (kRgbItem ? item : size_t(i));
    float4* mine = srec + 4 * threadIdx.x;
    mine[0] = make_float4(p.pix[0], p.pix[1], p.conic[0], p.conic[1]);
    mine[1] = make_float4(p.conic[2], op, col[0], col[1]);
    mine[2] = make_float4(col[2], p.aabb[0], p.aabb[1], p.aabb[2]);
    mine[3] = make_float4(p.aabb[3], 0.f, 0.f, 0.f);
    const TileRect tr = tile_rect(p.aabb[0], p.aabb[1], p.aabb[2], p.aabb[3], ts, vd.tw, vd.th);
    rect[sp] = pack_rect(tr);
    dkey[sp] = depth_key(p.ndcz);
  }
  wave_lds_fence();  // every lane of the wave is past its LDS writes
  const uint32_t nrec4 = 4u * min(64u, n - g0);
  float4* out = rec + 4 * (size_t(v) * n + g0);
#pragma unroll
  for (int c = 0; c < 4; ++c) {
    const uint32_t k = uint32_t(c) * 64u + lane;
    if (k < nrec4) out[k] = srec[4 * wbase + k];
  }
}

// ----------------------------------------------------------- SH colour ---
// One lane per Gaussian: its coefficient record is staged once, then one
// colour per view goes to the item's slot (rgb_item [items, 3]).
template <int K, int DEG>
__global__ void __launch_bounds__(256)
mv_sh_kernel(uint32_t n, uint32_t nview, const ViewDev* __restrict__ views,
             const float* __restrict__ xyz, const float* __restrict__ coef,
             float* __restrict__ rgb_item) {
  constexpr int R = 3 * K, RP = sh::record_stride(R), NK = (DEG + 1) * (DEG + 1);
  __shared__ __align__(16) float lds[4][64 * RP];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const uint32_t g0 = blockIdx.x * 256u + 64u * uint32_t(wave);
  if (g0 >= n) return;  // wave-uniform
  const uint32_t cnt = min(64u, n - g0);
  sh::stage_in<R, RP>(coef, g0, cnt, lds[wave], lane);
  sh::wave_lds_sync();
  const uint32_t i = g0 + uint32_t(lane);
  if (i >= n) return;
  float c[R];
  sh::load_record<R, RP>(lds[wave], lane, c);
  const float px = xyz[3 * i], py = xyz[3 * i + 1], pz = xyz[3 * i + 2];
  for (uint32_t v = 0; v < nview; ++v) {
    float d[3], inv, b[16], raw[3];
    sh::direction(views[v].cam, px, py, pz, d, &inv);
    sh::sh_basis<DEG>(d[0], d[1], d[2], b);
    sh::colour<NK>(c, b, raw);
    float* o = rgb_item + 3 * (size_t(v) * n + i);
#pragma unroll
    for (int ch = 0; ch < 3; ++ch) o[ch] = raw[ch] > 0.f ? raw[ch] : 0.f;
  }
}

// The adjoint of mv_sh_kernel summed over views: d_sh (stored) and the
// direction part of d_xyz (added to what the projection adjoint stored).
template <int K, int DEG>
__global__ void __launch_bounds__(256)
mv_sh_bwd_kernel(uint32_t n, uint32_t nview, const ViewDev* __restrict__ views,
                 const float* __restrict__ xyz, const float* __restrict__ coef,
                 const float* __restrict__ d_rgb_item, float* __restrict__ d_coef,
                 float* __restrict__ d_xyz) {
  constexpr int R = 3 * K, RP = sh::record_stride(R), NK = (DEG + 1) * (DEG + 1);
  __shared__ __align__(16) float lds[4][64 * RP];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const uint32_t g0 = blockIdx.x * 256u + 64u * uint32_t(wave);
  if (g0 >= n) return;  // wave-uniform
  const uint32_t cnt = min(64u, n - g0);
  float* L = lds[wave];
  sh::stage_in<R, RP>(coef, g0, cnt, L, lane);
  sh::wave_lds_sync();
  const uint32_t i = g0 + uint32_t(lane);
  float c[R];
  sh::load_record<R, RP>(L, lane, c);
  float ds[RP];
#pragma unroll
  for (int k = 0; k < RP; ++k) ds[k] = 0.f;
  if (i < n) {
    const float px = xyz[3 * i], py = xyz[3 * i + 1], pz = xyz[3 * i + 2];
    float ax[3] = {0.f, 0.f, 0.f};
    for (uint32_t v = 0; v < nview; ++v) {
      const float* gi = d_rgb_item + 3 * (size_t(v) * n + i);
      const float g0c = gi[0], g1c = gi[1], g2c = gi[2];
      if (g0c == 0.f && g1c == 0.f && g2c == 0.f) continue;  // not seen in this view
      float d[3], inv, b[16], raw[3];
      sh::direction(views[v].cam, px, py, pz, d, &inv);
      sh::sh_basis<DEG>(d[0], d[1], d[2], b);
      sh::colour<NK>(c, b, raw);
      const float gc[3] = {raw[0] < 0.f ? 0.f : g0c, raw[1] < 0.f ? 0.f : g1c,
                           raw[2] < 0.f ? 0.f : g2c};
#pragma unroll
      for (int k = 0; k < NK; ++k)
#pragma unroll
        for (int ch = 0; ch < 3; ++ch) ds[3 * k + ch] += b[k] * gc[ch];
      sh::direction_adjoint<DEG>(c, gc, d, inv, ax);
    }
#pragma unroll
    for (int a = 0; a < 3; ++a) d_xyz[3 * i + a] += ax[a];
  }
  // every lane has its record in registers; write d_sh back through LDS
  sh::wave_lds_sync();
#pragma unroll
  for (int q = 0; q < (R + 3) / 4; ++q)
    *reinterpret_cast<float4*>(L + lane * RP + 4 * q) =
        make_float4(ds[4 * q], ds[4 * q + 1], ds[4 * q + 2], ds[4 * q + 3]);
  sh::wave_lds_sync();
  sh::stage_out<R, RP, false>(d_coef, g0, cnt, L, lane);
}

// ------------------------------------------------------------- binning ---
__global__ void mv_clamp_kernel(uint32_t* __restrict__ misc, uint32_t capacity) {
  const uint32_t k = misc[0];
  misc[1] = k <= capacity ? k : 0u;
  misc[2] = 0u;  // forward work counter
  misc[3] = 0u;  // adjoint work counter
  misc[4] = k > capacity ? 1u : 0u;
}

// First pair slot of entry j (depth order): the chunk prefix of its
// workgroup's 256 entries plus a block scan of the counts cnt. The sorted
// values are positions in the depth sort's segments of seg entries per view
// (item g of view v at v seg + g); padding entries have no pairs.
__global__ void __launch_bounds__(kScanBlock)
mv_fill_kernel(uint32_t nitems, uint32_t n, uint32_t seg, const ViewDev* __restrict__ views,
               const uint32_t* __restrict__ order, const uint2* __restrict__ rect_sorted,
               const uint32_t* __restrict__ cnt, const uint32_t* __restrict__ chunk_prefix,
               const uint32_t* __restrict__ misc,
               uint32_t* __restrict__ tile_of_pair, uint32_t* __restrict__ item_of_pair) {
  __shared__ ExpandLds lds[4];
  if (misc[1] == 0) return;  // uniform: overflow, nothing is written
  const uint32_t j = blockIdx.x * blockDim.x + threadIdx.x;
  const bool valid = j < nitems;
  uint32_t item = 0, c = 0, tbase = 0, tw = 1;
  TileRect r{0, 0, 0, 0};
  if (valid) {
    const uint32_t pos = order[j], v = pos / seg;
    item = v * n + (pos - v * seg);
    const ViewDev& vd = views[v];
    r = unpack_rect(rect_sorted[j]);  // gathered into depth order by the sort
    c = cnt[j];
    tbase = vd.tile_off;
    tw = vd.tw;
  }
  uint32_t block_total;
  const uint32_t o = chunk_prefix[blockIdx.x] + block_exclusive_scan(c, &block_total);
  wave_expand_pairs(valid, o, r, tbase, tw, item, lds[threadIdx.x >> 6], tile_of_pair,
                    item_of_pair);
}

__global__ void __launch_bounds__(256)
mv_ranges_kernel(const uint32_t* __restrict__ misc, uint32_t capacity, uint32_t num_tile,
                 const uint32_t* __restrict__ tile_sorted, uint32_t* __restrict__ tile2idx) {
  tile_ranges4(blockIdx.x * blockDim.x + threadIdx.x, min(misc[1], capacity), num_tile,
               tile_sorted, tile2idx);
}

__global__ void __launch_bounds__(256)
mv_copy_kernel(const uint32_t* __restrict__ misc, uint32_t cap, const uint32_t* __restrict__ src,
               uint32_t* __restrict__ dst) {
  const uint32_t n = min(misc[1], cap);
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
    dst[i] = src[i];
}

__global__ void __launch_bounds__(256)
zero_f32_kernel(size_t n, float* __restrict__ p) {
  for (size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n;
       i += size_t(gridDim.x) * blockDim.x)
    p[i] = 0.f;
}

// -------------------------------------------------------------- raster ---
// Next work item of a persistent workgroup (0, 1, 2, ... across the grid).
// Wave 0 draws the ticket (lane 0 adds 1, the other lanes add 0) and all its
// lanes store it. Every branch is on the uniform wave index, so the compiler
// cannot split a wave at the loop head: with a per-lane `threadIdx.x == 0`
// fetch it may merge that branch with a per-lane one at the end of the loop
// body into an inner loop that lane 0 leaves while the other lanes never do.
__device__ inline uint32_t next_work_item(uint32_t* __restrict__ counter) {
  __shared__ uint32_t s_item;
  if (__builtin_amdgcn_readfirstlane(threadIdx.x >> 6) == 0) {
    const uint32_t t = atomicAdd(counter, threadIdx.x == 0 ? 1u : 0u);
    s_item = __builtin_amdgcn_readfirstlane(t);
  }
  __syncthreads();
  const uint32_t q = s_item;
  __syncthreads();
  return q;
}

template <int TS, int NT>
__global__ void __launch_bounds__(NT)
mv_raster_fwd_kernel(const ViewDev* __restrict__ views, uint32_t nview, uint32_t total_tiles,
                     const float4* __restrict__ rec, const uint32_t* __restrict__ tile2idx,
                     const uint32_t* __restrict__ idx2pnt, float* __restrict__ rgb,
                     float* __restrict__ alpha, uint32_t* __restrict__ nwalk,
                     float* __restrict__ tfinal, uint32_t* __restrict__ counter,
                     const uint32_t* __restrict__ tile_order, uint32_t* __restrict__ tile_cost) {
  const PackedLoader ld{rec};
  for (;;) {
    const uint32_t q = next_work_item(counter);
    if (q >= total_tiles) break;
    const uint32_t g = tile_order[q];  // longest walks first
    if (g >= total_tiles) continue;    // uniform; the plan writes only permutations
    const uint32_t v = view_of_tile(views, nview, g);
    const ViewDev& vd = views[v];
    const uint32_t local = g - vd.tile_off;
    const size_t po = vd.pix_off;
    raster_tile_fwd<TS, NT, PackedLoader>(local % vd.tw, local / vd.tw, vd.w, vd.h, tile2idx[g],
                                          tile2idx[g + 1], ld, idx2pnt, rgb + 3 * po, alpha + po,
                                          nwalk + po, tfinal + po, tile_cost, g);
  }
}

template <int TS, int NT>
__global__ void __launch_bounds__(NT)
mv_raster_bwd_kernel(const ViewDev* __restrict__ views, uint32_t nview, uint32_t n,
                     uint32_t total_tiles, const float4* __restrict__ rec,
                     const uint32_t* __restrict__ tile2idx, const uint32_t* __restrict__ idx2pnt,
                     const uint32_t* __restrict__ nwalk, const float* __restrict__ tfinal,
                     const float* __restrict__ d_img, const float* __restrict__ rgb,
                     const float* __restrict__ target, const float* __restrict__ d_alpha,
                     GradSink out, uint32_t* __restrict__ counter,
                     const uint32_t* __restrict__ tile_order) {
  for (;;) {
    const uint32_t q = next_work_item(counter);
    if (q >= total_tiles) break;
    const uint32_t g = tile_order[q];  // longest walks first
    if (g >= total_tiles) continue;    // uniform; the plan writes only permutations
    const uint32_t v = view_of_tile(views, nview, g);
    const ViewDev& vd = views[v];
    const uint32_t local = g - vd.tile_off;
    const size_t po = vd.pix_off;
    raster_tile_bwd<TS, NT>(local % vd.tw, local / vd.tw, vd.w, vd.h, tile2idx[g + 1], rec,
                            idx2pnt, nwalk + po, tfinal + po, d_img ? d_img + 3 * po : nullptr,
                            rgb + 3 * po, target ? target + 3 * po : nullptr,
                            d_alpha ? d_alpha + po : nullptr, v * n, out);
  }
}

// Projection adjoint summed over views, one thread per Gaussian. Overwrites
// d_xyz / d_scale / d_quat; with kLogitOp also turns dL/dopacity into
// dL/dlogit in place.
template <bool kLogScale, bool kLogitOp>
__global__ void __launch_bounds__(256)
mv_project_bwd_kernel(uint32_t n, uint32_t nview, const ViewDev* __restrict__ views,
                      const float* __restrict__ xyz, const float* __restrict__ scale,
                      const float4* __restrict__ quat, const float* __restrict__ opacity,
                      const float2* __restrict__ d_pixxy, const float* __restrict__ d_conic,
                      float* __restrict__ d_xyz, float* __restrict__ d_scale,
                      float4* __restrict__ d_quat, float* __restrict__ d_op) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  float sg[3] = {scale[3 * i], scale[3 * i + 1], scale[3 * i + 2]};
  if (kLogScale) {
    sg[0] = expf(sg[0]);
    sg[1] = expf(sg[1]);
    sg[2] = expf(sg[2]);
  }
  const float x3[3] = {xyz[3 * i], xyz[3 * i + 1], xyz[3 * i + 2]};
  const float4 qt = quat[i];
  float ax[3] = {0.f, 0.f, 0.f}, as[3] = {0.f, 0.f, 0.f}, aq[4] = {0.f, 0.f, 0.f, 0.f};
  for (uint32_t v = 0; v < nview; ++v) {
    const size_t item = size_t(v) * n + i;
    const float2 gp = d_pixxy[item];
    const float gcon[3] = {d_conic[3 * item], d_conic[3 * item + 1], d_conic[3 * item + 2]};
    if (gp.x == 0.f && gp.y == 0.f && gcon[0] == 0.f && gcon[1] == 0.f && gcon[2] == 0.f)
      continue;
    float dx[3], dsc[3], dq[4];
    if (!project_bwd_one(views[v].m, views[v].w, views[v].h, x3, qt, sg, kLogScale, gp, gcon, dx,
                         dsc, dq))
      continue;
#pragma unroll
    for (int k = 0; k < 3; ++k) {
      ax[k] += dx[k];
      as[k] += dsc[k];
    }
#pragma unroll
    for (int k = 0; k < 4; ++k) aq[k] += dq[k];
  }
  d_xyz[3 * i + 0] = ax[0];
  d_xyz[3 * i + 1] = ax[1];
  d_xyz[3 * i + 2] = ax[2];
  d_scale[3 * i + 0] = as[0];
  d_scale[3 * i + 1] = as[1];
  d_scale[3 * i + 2] = as[2];
  d_quat[i] = make_float4(aq[0], aq[1], aq[2], aq[3]);
  if (kLogitOp) {
    const float op = sigmoidf_(opacity[i]);
    d_op[i] *= op * (1.f - op);
  }
}

// ------------------------------------------------------------------ Adam ---
// Kingma & Ba (2015), Algorithm 1, elementwise; the step counter lives in
// device memory (incremented by adam_tick_kernel) so a captured graph can be
// replayed step after step.
__global__ void __launch_bounds__(256)
adam_kernel(size_t n, float* __restrict__ p, const float* __restrict__ g, float* __restrict__ m1,
            float* __restrict__ m2, float lr, float b1, float b2, float eps,
            const uint32_t* __restrict__ d_step) {
  const float t = float(*d_step);
  const float bc1 = 1.f - powf(b1, t);
  const float bc2 = 1.f - powf(b2, t);
  const float step_size = lr / bc1;
  const float inv_sqrt_bc2 = rsqrtf(bc2);
  for (size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n;
       i += size_t(gridDim.x) * blockDim.x) {
    const float gi = g[i];
    const float a = b1 * m1[i] + (1.f - b1) * gi;
    const float b = b2 * m2[i] + (1.f - b2) * gi * gi;
    m1[i] = a;
    m2[i] = b;
    p[i] -= step_size * a / (sqrtf(b) * inv_sqrt_bc2 + eps);
  }
}

__global__ void adam_tick_kernel(uint32_t* d_step) { *d_step += 1u; }

__global__ void set_word_kernel(uint32_t* p, uint32_t v) { *p = v; }

inline uint32_t grid_cap(size_t n) {
  const size_t b = (n + 255) / 256;
  return uint32_t(b > 8192 ? 8192 : (b ? b : 1));
}

}  // namespace dsplat

using namespace dsplat;

// ----------------------------------------------------------------- plan ---
struct dsplat_mv_plan {
  uint32_t n = 0, nview = 0, max_view = 0, ts = 16, threads = 0, capacity = 0;
  // entries per view in the depth sort: n, or n padded to whole sort
  // segments when the views are sorted each on its own
  uint32_t seg = 0;
  uint32_t max_tiles = 0, total_tiles = 0;
  uint64_t max_pix = 0, total_pix = 0;
  uint32_t persist_fwd = 0, persist_bwd = 0;
  std::vector<ViewDev> host_views;
  char* mem = nullptr;
  size_t mem_bytes = 0;
  ViewDev* d_views = nullptr;
  float4* rec = nullptr;
  // depth sort (seg entries per view): tile rectangles and keys per entry,
  // the sorted entries, and their rectangles and pair counts in depth order
  uint2 *rect = nullptr, *rect_sorted = nullptr;
  uint32_t *dkey = nullptr, *order = nullptr;
  void* dsort_ws = nullptr;
  size_t dsort_bytes = 0;
  uint32_t *off = nullptr, *scan_ws = nullptr;
  uint32_t *tilekey = nullptr, *idx2pnt = nullptr;
  void* tsort_ws = nullptr;
  size_t tsort_bytes = 0;
  uint32_t *tile2idx = nullptr, *misc = nullptr;
  float *rgb = nullptr, *alpha = nullptr, *tfinal = nullptr;
  uint32_t* nwalk = nullptr;
  float *d_pixxy = nullptr, *d_conic = nullptr;
  // per global tile: the last forward's longest walk, and the tiles ordered by
  // it (the schedule of the next adjoint and forward; raster order after
  // set_views)
  uint32_t *tile_cost = nullptr, *tile_order = nullptr;
  // SH colour (dsplat_mv_set_sh): coefficients per channel (0: plain rgb),
  // degree, per-item colours and their gradients, the coefficients the last
  // forward read, and whether every view has a camera centre
  uint32_t sh_k = 0, sh_deg = 0, fwd_sh_k = 0, fwd_sh_deg = 0;
  float *rgb_item = nullptr, *d_rgb_item = nullptr;
  const float* fwd_sh = nullptr;
  void* sh_mem = nullptr;
  bool cams_ok = false;
  // a forward over the current views and colour mode has been enqueued (or
  // replayed from a graph): the adjoint reads its lists, walks and
  // transmittances. Cleared by set_views, set_sh and a failed forward.
  bool fwd_valid = false;
  hipGraphExec_t exec = nullptr;
  hipStream_t cap_stream = nullptr;
};

static size_t a256(size_t v) { return (v + 255) & ~size_t(255); }

// Camera centre of a perspective world->ndc matrix (column-major, M_ij =
// m[i + 4 j]): the point where clip x, y and w all vanish, i.e. A c = -b over
// rows 0, 1, 3 (Cramer's rule in double). False if those rows are singular.
static bool camera_centre(const float* m, float* c) {
  double a[3][3], b[3];
  const int rows[3] = {0, 1, 3};
  for (int r = 0; r < 3; ++r) {
    for (int j = 0; j < 3; ++j) a[r][j] = m[rows[r] + 4 * j];
    b[r] = -double(m[rows[r] + 12]);
  }
  auto det3 = [](const double (&x)[3][3]) {
    return x[0][0] * (x[1][1] * x[2][2] - x[1][2] * x[2][1]) -
           x[0][1] * (x[1][0] * x[2][2] - x[1][2] * x[2][0]) +
           x[0][2] * (x[1][0] * x[2][1] - x[1][1] * x[2][0]);
  };
  const double d = det3(a);
  double scale = 0.0;
  for (int r = 0; r < 3; ++r)
    for (int j = 0; j < 3; ++j) scale = fmax(scale, fabs(a[r][j]));
  if (!(fabs(d) > 1e-12 * scale * scale * scale)) return false;
  for (int col = 0; col < 3; ++col) {
    double t[3][3];
    for (int r = 0; r < 3; ++r)
      for (int j = 0; j < 3; ++j) t[r][j] = j == col ? b[r] : a[r][j];
    c[col] = float(det3(t) / d);
  }
  return true;
}

static int fill_views(dsplat_mv_plan* P, uint32_t nview, const dsplat_view_desc* views) {
  if (nview == 0 || nview > P->max_view)
    return dsplat_set_error(DSPLAT_ERR_ARG, "multiview: view count out of range");
  std::vector<ViewDev> vd(nview);
  bool cams_ok = true;
  uint32_t toff = 0;
  uint64_t poff = 0;
  for (uint32_t v = 0; v < nview; ++v) {
    memcpy(vd[v].m, views[v].world2ndc, sizeof(vd[v].m));
    vd[v].w = views[v].img_w;
    vd[v].h = views[v].img_h;
    vd[v].tw = (views[v].img_w + P->ts - 1) / P->ts;
    vd[v].th = (views[v].img_h + P->ts - 1) / P->ts;
    vd[v].tile_off = toff;
    vd[v].pad0 = 0;
    vd[v].pix_off = poff;
    if (!camera_centre(vd[v].m, vd[v].cam)) {
      vd[v].cam[0] = vd[v].cam[1] = vd[v].cam[2] = 0.f;
      cams_ok = false;
    }
    vd[v].pad1 = 0.f;
    toff += vd[v].tw * vd[v].th;
    poff += uint64_t(views[v].img_w) * views[v].img_h;
  }
  for (uint32_t v = 0; v < nview; ++v)
    if (vd[v].tw > 0xFFFFu || vd[v].th > 0xFFFFu)  // rectangles are packed as 16-bit tiles
      return dsplat_set_error(DSPLAT_ERR_ARG, "multiview: more than 65535 tiles along an axis");
  if (toff > P->max_tiles || poff > P->max_pix)
    return dsplat_set_error(DSPLAT_ERR_CAPACITY, "multiview: views exceed the plan's pixels/tiles");
  P->host_views = vd;
  P->cams_ok = cams_ok;
  P->nview = nview;
  P->total_tiles = toff;
  P->total_pix = poff;
  return DSPLAT_OK;
}

extern "C" int dsplat_mv_plan_create(uint32_t num_pnt, uint32_t max_view, uint64_t max_pixels,
                                     uint32_t tile_size, uint32_t threads, uint32_t pair_capacity,
                                     dsplat_mv_plan** out) {
  if (!out) return dsplat_set_error(DSPLAT_ERR_ARG, "multiview: null plan pointer");
  *out = nullptr;
  if (tile_size != 8 && tile_size != 16 && tile_size != 32)
    return dsplat_set_error(DSPLAT_ERR_ARG, "tile_size must be 8, 16 or 32");
  if (max_view == 0 || max_view > kMaxViews || num_pnt == 0 || max_pixels == 0)
    return dsplat_set_error(DSPLAT_ERR_ARG, "multiview: empty or too many views / points");
  if (uint64_t(num_pnt) * max_view >= (1ull << 32))
    return dsplat_set_error(DSPLAT_ERR_ARG, "multiview: num_pnt * views must fit in 32 bits");
  dsplat_mv_plan* P = new dsplat_mv_plan();
  P->n = num_pnt;
  P->max_view = max_view;
  P->ts = tile_size;
  P->threads = threads;
  P->capacity = pair_capacity;
  P->max_pix = max_pixels;
  // each view adds at most one partial tile row and column
  const uint64_t tiles = max_pixels / (uint64_t(tile_size) * tile_size) +
                         uint64_t(max_view) * (2 * 4096 / tile_size + 1);
  P->max_tiles = uint32_t(tiles > 0xFFFFFFF0ull ? 0xFFFFFFF0ull : tiles);
  const size_t items = size_t(num_pnt) * max_view;
  // Views of at least kSegSortMin Gaussians are depth-sorted each in its own
  // segment (padded to whole segments): the sort's last pass gathers every
  // entry's tile rectangle, and within one view's segment those reads stay
  // in a slice of the rectangles small enough for the caches, where across
  // all views they were random reads over the whole array.
  const uint64_t seg = num_pnt >= kSegSortMin
                           ? (uint64_t(num_pnt) + kSortSegAlign - 1) / kSortSegAlign * kSortSegAlign
                           : num_pnt;
  if (seg * max_view > 0xFFFFFFFFull) {
    delete P;
    return dsplat_set_error(DSPLAT_ERR_CAPACITY, "multiview: too many items for the depth sort");
  }
  P->seg = uint32_t(seg);
  const size_t sitems = size_t(seg) * max_view;
  size_t o = 0;
  auto take = [&](size_t b) { size_t r = o; o += a256(b); return r; };
  const size_t o_views = take(sizeof(ViewDev) * max_view);
  const size_t o_rec = take(items * 64);
  const size_t o_rect = take(sitems * 8), o_rsorted = take(sitems * 8);
  const size_t o_dkey = take(sitems * 4), o_order = take(sitems * 4);
  P->dsort_bytes = sort_workspace_bytes<uint32_t>(uint32_t(sitems));
  const size_t o_dsort = take(P->dsort_bytes);
  const size_t o_off = take((sitems + 1) * 4);
  const size_t o_scan = take(scan_chunks_workspace_bytes(uint32_t(sitems)));
  const size_t o_tkey = take(size_t(pair_capacity) * 4 + 4);
  const size_t o_i2p = take(size_t(pair_capacity) * 4 + 4);
  P->tsort_bytes = sort_workspace_bytes<uint32_t>(pair_capacity);
  const size_t o_tsort = take(P->tsort_bytes);
  const size_t o_t2i = take((size_t(P->max_tiles) + 1) * 4);
  const size_t o_misc = take(64);
  const size_t o_rgb = take(max_pixels * 12), o_alpha = take(max_pixels * 4);
  const size_t o_nwalk = take(max_pixels * 4), o_tfinal = take(max_pixels * 4);
  const size_t o_dpix = take(items * 8), o_dcon = take(items * 12);
  const size_t o_tcost = take(size_t(P->max_tiles) * 4), o_tord = take(size_t(P->max_tiles) * 4);
  P->mem_bytes = o;
  hipError_t e = hipMalloc(reinterpret_cast<void**>(&P->mem), P->mem_bytes);
  if (e != hipSuccess) {
    delete P;
    return dsplat_set_hip_error(e, "multiview: hipMalloc plan memory");
  }
  char* b = P->mem;
  P->d_views = reinterpret_cast<ViewDev*>(b + o_views);
  P->rec = reinterpret_cast<float4*>(b + o_rec);
  P->rect = reinterpret_cast<uint2*>(b + o_rect);
  P->rect_sorted = reinterpret_cast<uint2*>(b + o_rsorted);
  P->dkey = reinterpret_cast<uint32_t*>(b + o_dkey);
  P->order = reinterpret_cast<uint32_t*>(b + o_order);
  P->dsort_ws = b + o_dsort;
  P->off = reinterpret_cast<uint32_t*>(b + o_off);
  P->scan_ws = reinterpret_cast<uint32_t*>(b + o_scan);
  P->tilekey = reinterpret_cast<uint32_t*>(b + o_tkey);
  P->idx2pnt = reinterpret_cast<uint32_t*>(b + o_i2p);
  P->tsort_ws = b + o_tsort;
  P->tile2idx = reinterpret_cast<uint32_t*>(b + o_t2i);
  P->misc = reinterpret_cast<uint32_t*>(b + o_misc);
  P->rgb = reinterpret_cast<float*>(b + o_rgb);
  P->alpha = reinterpret_cast<float*>(b + o_alpha);
  P->nwalk = reinterpret_cast<uint32_t*>(b + o_nwalk);
  P->tfinal = reinterpret_cast<float*>(b + o_tfinal);
  P->d_pixxy = reinterpret_cast<float*>(b + o_dpix);
  P->d_conic = reinterpret_cast<float*>(b + o_dcon);
  P->tile_cost = reinterpret_cast<uint32_t*>(b + o_tcost);
  P->tile_order = reinterpret_cast<uint32_t*>(b + o_tord);
  // persistent grids: resident workgroups per CU times CUs
  int dev = 0, cus = 0;
  if (hipGetDevice(&dev) != hipSuccess ||
      hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess)
    cus = 256;
  P->persist_fwd = uint32_t(cus > 0 ? cus : 256) * 8;
  P->persist_bwd = uint32_t(cus > 0 ? cus : 256) * 4;
  *out = P;
  return DSPLAT_OK;
}

extern "C" int dsplat_mv_plan_destroy(dsplat_mv_plan* P) {
  if (!P) return DSPLAT_OK;
  if (P->exec) (void)hipGraphExecDestroy(P->exec);
  if (P->cap_stream) (void)hipStreamDestroy(P->cap_stream);
  if (P->mem) (void)hipFree(P->mem);
  if (P->sh_mem) (void)hipFree(P->sh_mem);
  delete P;
  return DSPLAT_OK;
}

// The view table travels as kernel arguments (no host buffer has to outlive
// the call, and the host never waits): 16 views per launch.
constexpr uint32_t kViewsPerUpload = 16;
struct ViewUpload {
  ViewDev v[kViewsPerUpload];
};
static_assert(sizeof(ViewUpload) <= 2048, "view upload must fit the kernel arguments");

__global__ void __launch_bounds__(256)
upload_views_kernel(ViewUpload up, uint32_t count, ViewDev* __restrict__ dst) {
  constexpr uint32_t kWords = sizeof(ViewDev) / 4;
  const uint32_t* src = reinterpret_cast<const uint32_t*>(up.v);
  uint32_t* d = reinterpret_cast<uint32_t*>(dst);
  for (uint32_t k = threadIdx.x; k < count * kWords; k += blockDim.x) d[k] = src[k];
}

// A captured graph bakes in view counts, tile totals, grid sizes and the SH
// mode; changing any of them drops it (the next dsplat_mv_graph_launch then
// reports that nothing is captured instead of replaying stale launches).
static void drop_graph(dsplat_mv_plan* P) {
  if (P->exec) {
    (void)hipGraphExecDestroy(P->exec);
    P->exec = nullptr;
  }
}

extern "C" int dsplat_mv_set_views(dsplat_mv_plan* P, void* stream, uint32_t num_view,
                                   const dsplat_view_desc* views) {
  if (!P || !views) return dsplat_set_error(DSPLAT_ERR_ARG, "multiview: null plan/views");
  drop_graph(P);
  P->fwd_valid = false;
  if (int rc = fill_views(P, num_view, views)) return rc;
  hipStream_t s = as_stream(stream);
  for (uint32_t v0 = 0; v0 < num_view; v0 += kViewsPerUpload) {
    ViewUpload up{};
    const uint32_t cnt = min(kViewsPerUpload, num_view - v0);
    memcpy(up.v, P->host_views.data() + v0, sizeof(ViewDev) * cnt);
    upload_views_kernel<<<1, 256, 0, s>>>(up, cnt, P->d_views + v0);
  }
  DSPLAT_HIP(hipGetLastError());
  return iota_u32(s, P->total_tiles, P->tile_order);
}

extern "C" int dsplat_mv_set_sh(dsplat_mv_plan* P, uint32_t degree, uint32_t num_coef) {
  if (!P) return dsplat_set_error(DSPLAT_ERR_ARG, "multiview: null plan");
  drop_graph(P);
  P->fwd_valid = false;
  if (num_coef == 0) {
    P->sh_k = P->sh_deg = 0;
    return DSPLAT_OK;
  }
  if ((num_coef != 1 && num_coef != 4 && num_coef != 9 && num_coef != 16) || degree > 3 ||
      (degree + 1) * (degree + 1) > num_coef)
    return dsplat_set_error(DSPLAT_ERR_ARG,
                            "mv_set_sh: num_coef must be 0, 1, 4, 9 or 16 and degree <= 3 with "
                            "(degree+1)^2 <= num_coef");
  if (!P->sh_mem) {
    const size_t items = size_t(P->n) * P->max_view;
    char* m = nullptr;
    DSPLAT_HIP(hipMalloc(reinterpret_cast<void**>(&m), a256(items * 12) * 2));
    P->sh_mem = m;
    P->rgb_item = reinterpret_cast<float*>(m);
    P->d_rgb_item = reinterpret_cast<float*>(m + a256(items * 12));
  }
  P->sh_k = num_coef;
  P->sh_deg = degree;
  return DSPLAT_OK;
}

extern "C" int dsplat_mv_plan_info(const dsplat_mv_plan* P, uint32_t* num_view,
                                   uint64_t* total_pix, uint32_t* total_tiles,
                                   uint64_t* pix_offsets, uint32_t* tile_offsets) {
  if (!P) return dsplat_set_error(DSPLAT_ERR_ARG, "multiview: null plan");
  if (num_view) *num_view = P->nview;
  if (total_pix) *total_pix = P->total_pix;
  if (total_tiles) *total_tiles = P->total_tiles;
  for (uint32_t v = 0; v < P->nview; ++v) {
    if (pix_offsets) pix_offsets[v] = P->host_views[v].pix_off;
    if (tile_offsets) tile_offsets[v] = P->host_views[v].tile_off;
  }
  if (pix_offsets) pix_offsets[P->nview] = P->total_pix;
  if (tile_offsets) tile_offsets[P->nview] = P->total_tiles;
  return DSPLAT_OK;
}

extern "C" int dsplat_mv_outputs(const dsplat_mv_plan* P, float** rgb, float** alpha,
                                 uint32_t** nwalk, float** tfinal, uint32_t** tile2idx,
                                 uint32_t** idx2pnt, uint32_t** counters) {
  if (!P) return dsplat_set_error(DSPLAT_ERR_ARG, "multiview: null plan");
  if (rgb) *rgb = P->rgb;
  if (alpha) *alpha = P->alpha;
  if (nwalk) *nwalk = P->nwalk;
  if (tfinal) *tfinal = P->tfinal;
  if (tile2idx) *tile2idx = P->tile2idx;
  if (idx2pnt) *idx2pnt = P->idx2pnt;
  if (counters) *counters = P->misc;
  return DSPLAT_OK;
}

static uint32_t bitw(uint32_t v) {
  uint32_t b = 0;
  while (v) { ++b; v >>= 1; }
  return b;
}

template <int TS, int NT>
static void launch_mv_fwd(dsplat_mv_plan* P, hipStream_t s, float* out_rgb, float* out_alpha) {
  mv_raster_fwd_kernel<TS, NT><<<P->persist_fwd, NT, 0, s>>>(
      P->d_views, P->nview, P->total_tiles, P->rec, P->tile2idx, P->idx2pnt, out_rgb, out_alpha,
      P->nwalk, P->tfinal, P->misc + 2, P->tile_order, P->tile_cost);
}

template <int TS, int NT>
static void launch_mv_bwd(dsplat_mv_plan* P, hipStream_t s, const float* out_rgb,
                          const float* d_img, const float* target, const float* d_alpha,
                          GradSink out) {
  mv_raster_bwd_kernel<TS, NT><<<P->persist_bwd, NT, 0, s>>>(
      P->d_views, P->nview, P->n, P->total_tiles, P->rec, P->tile2idx, P->idx2pnt, P->nwalk,
      P->tfinal, d_img, out_rgb, target, d_alpha, out, P->misc + 3, P->tile_order);
}

template <int K, int DEG>
static void mv_sh_launch(dsplat_mv_plan* P, hipStream_t s, const float* xyz, const float* coef,
                         float* d_coef, float* d_xyz) {
  const uint32_t nb = div_up(P->n, 256);
  if (d_coef)
    mv_sh_bwd_kernel<K, DEG><<<nb, 256, 0, s>>>(P->n, P->nview, P->d_views, xyz, coef,
                                                P->d_rgb_item, d_coef, d_xyz);
  else
    mv_sh_kernel<K, DEG><<<nb, 256, 0, s>>>(P->n, P->nview, P->d_views, xyz, coef,
                                            P->rgb_item);
}

// SH colour of every item (d_coef null), or its adjoint (d_coef = dL/dsh,
// stored; the direction part added to d_xyz).
static void mv_sh(dsplat_mv_plan* P, hipStream_t s, const float* xyz, const float* coef,
                  float* d_coef, float* d_xyz) {
  switch (P->sh_k * 4 + P->sh_deg) {
    case 1 * 4 + 0: mv_sh_launch<1, 0>(P, s, xyz, coef, d_coef, d_xyz); break;
    case 4 * 4 + 0: mv_sh_launch<4, 0>(P, s, xyz, coef, d_coef, d_xyz); break;
    case 4 * 4 + 1: mv_sh_launch<4, 1>(P, s, xyz, coef, d_coef, d_xyz); break;
    case 9 * 4 + 0: mv_sh_launch<9, 0>(P, s, xyz, coef, d_coef, d_xyz); break;
    case 9 * 4 + 1: mv_sh_launch<9, 1>(P, s, xyz, coef, d_coef, d_xyz); break;
    case 9 * 4 + 2: mv_sh_launch<9, 2>(P, s, xyz, coef, d_coef, d_xyz); break;
    case 16 * 4 + 0: mv_sh_launch<16, 0>(P, s, xyz, coef, d_coef, d_xyz); break;
    case 16 * 4 + 1: mv_sh_launch<16, 1>(P, s, xyz, coef, d_coef, d_xyz); break;
    case 16 * 4 + 2: mv_sh_launch<16, 2>(P, s, xyz, coef, d_coef, d_xyz); break;
    default: mv_sh_launch<16, 3>(P, s, xyz, coef, d_coef, d_xyz); break;
  }
}

// Offset of the colour block in a flat parameter block: 11N, rounded up to a
// multiple of 4 floats under SH colour (the coefficients move as 16-byte
// words).
static size_t colour_offset(const dsplat_mv_plan* P) {
  const size_t o = 11 * size_t(P->n);
  return P->sh_k ? (o + 3) & ~size_t(3) : o;
}

static size_t param_count(const dsplat_mv_plan* P) {
  return colour_offset(P) + 3 * size_t(P->n) * (P->sh_k ? P->sh_k : 1);
}

static bool al16(const void* p) { return (reinterpret_cast<uintptr_t>(p) & 15u) == 0; }

// Under SH colour the flat gradient block has padding between the opacity
// block (ending at 11N) and the colour block (colour_offset): the adjoint
// never writes it, so it is zeroed here (Adam then leaves the padding's
// parameters and moments unchanged, whatever the caller allocated).
static void zero_grad_padding(const dsplat_mv_plan* P, hipStream_t s, float* grads) {
  const size_t o = 11 * size_t(P->n), coff = colour_offset(P);
  if (coff > o) zero_f32_kernel<<<1, 256, 0, s>>>(coff - o, grads + o);
}

static int mv_forward(dsplat_mv_plan* P, hipStream_t s, const float* xyz, const float* scale,
                      const float* quat, const float* opacity, const float* rgb,
                      uint32_t flags, float* out_rgb, float* out_alpha) {
  P->fwd_valid = false;
  if (!out_alpha) out_alpha = P->alpha;
  const uint32_t n = P->n, nv = P->nview;
  const bool log_scale = flags & DSPLAT_SCALE_IS_LOG, logit = flags & DSPLAT_OPACITY_IS_LOGIT;
  const uint32_t sitems = P->seg * nv;  // depth-sort entries
  const dim3 pg(div_up(P->seg, 256), nv);
  const float4* q4 = reinterpret_cast<const float4*>(quat);
  const float* col = rgb;
  if (P->sh_k) {
    if (!P->cams_ok)
      return dsplat_set_error(DSPLAT_ERR_ARG,
                              "mv_forward: SH colour needs a perspective camera in every view");
    if (!al16(rgb))
      return dsplat_set_error(DSPLAT_ERR_ARG, "mv_forward: SH coefficients not 16-byte aligned");
    mv_sh(P, s, xyz, rgb, nullptr, nullptr);
    col = P->rgb_item;
  }
  P->fwd_sh = P->sh_k ? rgb : nullptr;
  P->fwd_sh_k = P->sh_k;
  P->fwd_sh_deg = P->sh_deg;
#define DSPLAT_MVP(A, B, C)                                                                    \
  mv_project_kernel<A, B, C><<<pg, 256, 0, s>>>(n, P->seg, P->d_views, P->ts, xyz, scale, q4,  \
                                                opacity, col, P->rec, P->rect, P->dkey)
#define DSPLAT_MVP2(A, B) \
  if (P->sh_k) DSPLAT_MVP(A, B, true); else DSPLAT_MVP(A, B, false)
  if (log_scale && logit) { DSPLAT_MVP2(true, true); }
  else if (log_scale) { DSPLAT_MVP2(true, false); }
  else if (logit) { DSPLAT_MVP2(false, true); }
  else { DSPLAT_MVP2(false, false); }
#undef DSPLAT_MVP2
#undef DSPLAT_MVP
  bool in_alt = false;
  uint32_t *kalt = nullptr, *valt = nullptr;
  // values: entry positions (iota), generated by the first pass; the last
  // pass also gathers the tile rectangles and their areas (the pair counts)
  // into depth order. Large views are sorted each in its own segment: the
  // tile lists come out the same (every tile belongs to one view, and within
  // a view the order is by depth, then item, either way).
  int rc = radix_sort_pairs<uint32_t>(s, sitems, nullptr, P->dkey, P->order, 31, P->dsort_ws,
                                      P->dsort_bytes, &in_alt, &kalt, &valt, true, P->rect,
                                      P->off, P->rect_sorted, n >= kSegSortMin ? P->seg : 0u);
  if (rc) return rc;
  const uint32_t* order = in_alt ? valt : P->order;
  rc = scan_chunk_prefixes(s, sitems, P->off, P->scan_ws, P->misc);
  if (rc) return rc;
  mv_clamp_kernel<<<1, 1, 0, s>>>(P->misc, P->capacity);
  mv_fill_kernel<<<div_up(sitems, kScanBlock), kScanBlock, 0, s>>>(
      sitems, n, P->seg, P->d_views, order, P->rect_sorted, P->off, P->scan_ws, P->misc,
      P->tilekey, P->idx2pnt);
  const uint32_t* tsorted = P->tilekey;
  if (P->capacity > 0 && P->total_tiles > 1) {
    rc = radix_sort_pairs<uint32_t>(s, P->capacity, P->misc + 1, P->tilekey, P->idx2pnt,
                                    bitw(P->total_tiles - 1), P->tsort_ws, P->tsort_bytes,
                                    &in_alt, &kalt, &valt);
    if (rc) return rc;
    if (in_alt) {
      mv_copy_kernel<<<1024, 256, 0, s>>>(P->misc, P->capacity, valt, P->idx2pnt);
      tsorted = kalt;
    }
  }
  if (reinterpret_cast<uintptr_t>(tsorted) & 15u)
    return dsplat_set_error(DSPLAT_ERR_ARG, "multiview: sorted tile ids not 16-byte aligned");
  mv_ranges_kernel<<<div_up(uint64_t(P->capacity) + 1, 1024), 256, 0, s>>>(
      P->misc, P->capacity, P->total_tiles, tsorted, P->tile2idx);
  const uint32_t ts = P->ts, th = P->threads;
  if (ts == 8) launch_mv_fwd<8, 64>(P, s, out_rgb, out_alpha);
  else if (ts == 16 && th == 64) launch_mv_fwd<16, 64>(P, s, out_rgb, out_alpha);
  else if (ts == 16 && th == 128) launch_mv_fwd<16, 128>(P, s, out_rgb, out_alpha);
  else if (ts == 16) launch_mv_fwd<16, 256>(P, s, out_rgb, out_alpha);
  else if (th == 1024) launch_mv_fwd<32, 1024>(P, s, out_rgb, out_alpha);
  else if (th == 512) launch_mv_fwd<32, 512>(P, s, out_rgb, out_alpha);
  else launch_mv_fwd<32, 256>(P, s, out_rgb, out_alpha);
  DSPLAT_HIP(hipGetLastError());
  // longest walks first for the adjoint and the next forward
  rc = order_by_cost(s, P->total_tiles, P->tile_cost, P->tile_order);
  P->fwd_valid = rc == DSPLAT_OK;
  return rc;
}

static int mv_backward(dsplat_mv_plan* P, hipStream_t s, const float* xyz, const float* scale,
                       const float* quat, const float* opacity, const float* out_rgb,
                       const float* d_img, const float* target, const float* d_alpha,
                       float* d_xyz, float* d_scale, float* d_quat, float* d_op, float* d_rgb,
                       uint32_t flags) {
  const uint32_t n = P->n, nv = P->nview;
  const size_t items = size_t(n) * nv;
  if (!P->fwd_valid)
    return dsplat_set_error(DSPLAT_ERR_ARG,
                            "mv_backward: no forward over the current views and colour mode");
  if (P->sh_k && (P->fwd_sh_k != P->sh_k || P->fwd_sh_deg != P->sh_deg || !P->fwd_sh))
    return dsplat_set_error(DSPLAT_ERR_ARG,
                            "mv_backward: the last forward did not use the current SH colour");
  if (P->sh_k && !al16(d_rgb))
    return dsplat_set_error(DSPLAT_ERR_ARG, "mv_backward: d_sh not 16-byte aligned");
  set_word_kernel<<<1, 1, 0, s>>>(P->misc + 3, 0u);
  zero_f32_kernel<<<grid_cap(items * 2), 256, 0, s>>>(items * 2, P->d_pixxy);
  zero_f32_kernel<<<grid_cap(items * 3), 256, 0, s>>>(items * 3, P->d_conic);
  zero_f32_kernel<<<grid_cap(n), 256, 0, s>>>(n, d_op);
  GradSink out{P->d_pixxy, P->d_conic, d_op, d_rgb};
  if (P->sh_k) {
    zero_f32_kernel<<<grid_cap(items * 3), 256, 0, s>>>(items * 3, P->d_rgb_item);
    out.d_rgb = P->d_rgb_item;
    out.rgb_per_item = true;
  } else {
    zero_f32_kernel<<<grid_cap(size_t(n) * 3), 256, 0, s>>>(size_t(n) * 3, d_rgb);
  }
  const uint32_t ts = P->ts, th = P->threads;
  if (ts == 8) launch_mv_bwd<8, 64>(P, s, out_rgb, d_img, target, d_alpha, out);
  else if (ts == 16 && th == 64) launch_mv_bwd<16, 64>(P, s, out_rgb, d_img, target, d_alpha, out);
  else if (ts == 16 && th == 128) launch_mv_bwd<16, 128>(P, s, out_rgb, d_img, target, d_alpha, out);
  else if (ts == 16) launch_mv_bwd<16, 256>(P, s, out_rgb, d_img, target, d_alpha, out);
  else if (th == 1024) launch_mv_bwd<32, 1024>(P, s, out_rgb, d_img, target, d_alpha, out);
  else if (th == 512) launch_mv_bwd<32, 512>(P, s, out_rgb, d_img, target, d_alpha, out);
  else launch_mv_bwd<32, 256>(P, s, out_rgb, d_img, target, d_alpha, out);
  const bool log_scale = flags & DSPLAT_SCALE_IS_LOG, logit = flags & DSPLAT_OPACITY_IS_LOGIT;
  const float4* q4 = reinterpret_cast<const float4*>(quat);
  const float2* gp = reinterpret_cast<const float2*>(P->d_pixxy);
  float4* dq = reinterpret_cast<float4*>(d_quat);
#define DSPLAT_MVB(A, B)                                                                   \
  mv_project_bwd_kernel<A, B><<<div_up(n, 256), 256, 0, s>>>(n, nv, P->d_views, xyz, scale, \
                                                             q4, opacity, gp, P->d_conic,  \
                                                             d_xyz, d_scale, dq, d_op)
  if (log_scale && logit) DSPLAT_MVB(true, true);
  else if (log_scale) DSPLAT_MVB(true, false);
  else if (logit) DSPLAT_MVB(false, true);
  else DSPLAT_MVB(false, false);
#undef DSPLAT_MVB
  if (P->sh_k) mv_sh(P, s, xyz, P->fwd_sh, d_rgb, d_xyz);
  DSPLAT_HIP(hipGetLastError());
  return DSPLAT_OK;
}

extern "C" int dsplat_mv_forward(dsplat_mv_plan* P, void* stream, const float* xyz,
                                 const float* scale, const float* quat, const float* opacity,
                                 const float* rgb, uint32_t flags, float* out_rgb,
                                 float* out_alpha) {
  if (!P || !xyz || !scale || !quat || !opacity || !rgb || !out_rgb)
    return dsplat_set_error(DSPLAT_ERR_ARG, "mv_forward: null argument");
  if (P->nview == 0) return dsplat_set_error(DSPLAT_ERR_ARG, "mv_forward: no views set");
  if (!al16(quat)) return dsplat_set_error(DSPLAT_ERR_ARG, "mv_forward: quat not 16-byte aligned");
  return mv_forward(P, as_stream(stream), xyz, scale, quat, opacity, rgb, flags, out_rgb,
                    out_alpha);
}

extern "C" int dsplat_mv_backward(dsplat_mv_plan* P, void* stream, const float* xyz,
                                  const float* scale, const float* quat, const float* opacity,
                                  const float* out_rgb, const float* d_img, const float* target,
                                  const float* d_alpha, float* d_xyz, float* d_scale,
                                  float* d_quat, float* d_opacity, float* d_rgb, uint32_t flags) {
  if (!P || !xyz || !scale || !quat || !opacity || !d_xyz || !d_scale || !d_quat ||
      !d_opacity || !d_rgb || (!d_img && (!target || !out_rgb)))
    return dsplat_set_error(DSPLAT_ERR_ARG, "mv_backward: null argument");
  if (!al16(quat) || !al16(d_quat))
    return dsplat_set_error(DSPLAT_ERR_ARG, "mv_backward: quat buffers not 16-byte aligned");
  return mv_backward(P, as_stream(stream), xyz, scale, quat, opacity, out_rgb, d_img, target,
                     d_alpha, d_xyz, d_scale, d_quat, d_opacity, d_rgb, flags);
}

extern "C" int dsplat_memcpy_d2d(void* stream, void* dst, const void* src, uint64_t bytes) {
  if (bytes == 0) return DSPLAT_OK;
  if (!dst || !src) return dsplat_set_error(DSPLAT_ERR_ARG, "memcpy: null pointer");
  DSPLAT_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToDevice, as_stream(stream)));
  return DSPLAT_OK;
}

extern "C" int dsplat_adam_step(void* stream, uint64_t n, float* param, const float* grad,
                                float* m1, float* m2, float lr, float beta1, float beta2,
                                float eps, uint32_t* d_step) {
  if (n == 0) return DSPLAT_OK;
  if (!param || !grad || !m1 || !m2 || !d_step)
    return dsplat_set_error(DSPLAT_ERR_ARG, "adam: null argument");
  hipStream_t s = as_stream(stream);
  adam_tick_kernel<<<1, 1, 0, s>>>(d_step);
  adam_kernel<<<grid_cap(n), 256, 0, s>>>(n, param, grad, m1, m2, lr, beta1, beta2, eps, d_step);
  DSPLAT_HIP(hipGetLastError());
  return DSPLAT_OK;
}

// One training step over the plan's current views on a flat parameter block
// [xyz 3N | log-scale 3N | quat 4N | opacity logit N | rgb 3N]: forward, the
// adjoint of sum (rgb - target)^2 over all views, and Adam on all 14N values.
// Under SH colour the last block is the coefficients [N][K][3] starting at
// 11N rounded up to a multiple of 4 (colour_offset), and Adam runs over the
// whole block, padding included.
static int mv_train_step(dsplat_mv_plan* P, hipStream_t s, float* params, float* grads,
                         float* m1, float* m2, uint32_t* d_step, const float* targets,
                         float* out_rgb, float lr, float b1, float b2, float eps) {
  const size_t n = P->n;
  float* xyz = params;
  float* ls = params + 3 * n;
  float* q = params + 6 * n;
  float* op = params + 10 * n;
  const size_t coff = colour_offset(P), total = param_count(P);
  float* rgb = params + coff;
  const uint32_t flags = DSPLAT_SCALE_IS_LOG | DSPLAT_OPACITY_IS_LOGIT;
  int rc = mv_forward(P, s, xyz, ls, q, op, rgb, flags, out_rgb, nullptr);
  if (rc) return rc;
  rc = mv_backward(P, s, xyz, ls, q, op, out_rgb, nullptr, targets, nullptr, grads,
                   grads + 3 * n, grads + 6 * n, grads + 10 * n, grads + coff, flags);
  if (rc) return rc;
  zero_grad_padding(P, s, grads);
  adam_tick_kernel<<<1, 1, 0, s>>>(d_step);
  adam_kernel<<<grid_cap(total), 256, 0, s>>>(total, params, grads, m1, m2, lr, b1, b2, eps,
                                              d_step);
  DSPLAT_HIP(hipGetLastError());
  return DSPLAT_OK;
}

extern "C" int dsplat_mv_train_step(dsplat_mv_plan* P, void* stream, float* params, float* grads,
                                    float* m1, float* m2, uint32_t* d_step, const float* targets,
                                    float* out_rgb, float lr, float beta1, float beta2,
                                    float eps) {
  if (!P || !params || !grads || !m1 || !m2 || !d_step || !targets || !out_rgb)
    return dsplat_set_error(DSPLAT_ERR_ARG, "mv_train_step: null argument");
  if (P->nview == 0) return dsplat_set_error(DSPLAT_ERR_ARG, "mv_train_step: no views set");
  if (!al16(params + 6 * size_t(P->n)) || !al16(grads + 6 * size_t(P->n)))
    return dsplat_set_error(DSPLAT_ERR_ARG, "mv_train_step: quat block not 16-byte aligned");
  if (P->sh_k && (!al16(params) || !al16(grads)))
    return dsplat_set_error(DSPLAT_ERR_ARG, "mv_train_step: parameter block not 16-byte aligned");
  return mv_train_step(P, as_stream(stream), params, grads, m1, m2, d_step, targets, out_rgb, lr,
                       beta1, beta2, eps);
}

// Captures one forward (kind 0) or one training step (kind 1) with the given
// arguments into the plan's graph; dsplat_mv_graph_launch replays it.
extern "C" int dsplat_mv_graph_capture(dsplat_mv_plan* P, uint32_t kind, const float* xyz,
                                       const float* scale, const float* quat,
                                       const float* opacity, const float* rgb, uint32_t flags,
                                       float* out_rgb, float* out_alpha, float* params,
                                       float* grads, float* m1, float* m2, uint32_t* d_step,
                                       const float* targets, float lr, float beta1, float beta2,
                                       float eps) {
  if (!out_rgb) return dsplat_set_error(DSPLAT_ERR_ARG, "graph: null out_rgb");
  if (!P || P->nview == 0) return dsplat_set_error(DSPLAT_ERR_ARG, "graph: no views set");
  if (!P->cap_stream) DSPLAT_HIP(hipStreamCreateWithFlags(&P->cap_stream, hipStreamNonBlocking));
  if (P->exec) {
    DSPLAT_HIP(hipGraphExecDestroy(P->exec));
    P->exec = nullptr;
  }
  hipGraph_t graph = nullptr;
  // capturing runs nothing: the forward-valid state is that of the last
  // executed work (a replay sets it)
  const bool fwd_valid = P->fwd_valid;
  DSPLAT_HIP(hipStreamBeginCapture(P->cap_stream, hipStreamCaptureModeThreadLocal));
  int rc;
  if (kind == 0) {
    rc = (xyz && scale && quat && opacity && rgb)
             ? mv_forward(P, P->cap_stream, xyz, scale, quat, opacity, rgb, flags, out_rgb,
                          out_alpha)
             : dsplat_set_error(DSPLAT_ERR_ARG, "graph: null forward argument");
  } else if (kind == 2) {
    const size_t n = P->n;
    rc = (xyz && scale && quat && opacity && rgb && grads && targets)
             ? mv_forward(P, P->cap_stream, xyz, scale, quat, opacity, rgb, flags, out_rgb,
                          out_alpha)
             : dsplat_set_error(DSPLAT_ERR_ARG, "graph: null forward/backward argument");
    if (!rc)
      rc = mv_backward(P, P->cap_stream, xyz, scale, quat, opacity, out_rgb, nullptr, targets,
                       nullptr, grads, grads + 3 * n, grads + 6 * n, grads + 10 * n,
                       grads + colour_offset(P), flags);
    if (!rc) zero_grad_padding(P, P->cap_stream, grads);
  } else {
    rc = (params && grads && m1 && m2 && d_step && targets)
             ? mv_train_step(P, P->cap_stream, params, grads, m1, m2, d_step, targets, out_rgb,
                             lr, beta1, beta2, eps)
             : dsplat_set_error(DSPLAT_ERR_ARG, "graph: null training argument");
  }
  hipError_t e = hipStreamEndCapture(P->cap_stream, &graph);
  P->fwd_valid = fwd_valid;
  if (rc) {
    if (graph) (void)hipGraphDestroy(graph);
    return rc;
  }
  if (e != hipSuccess) return dsplat_set_hip_error(e, "graph: end capture");
  e = hipGraphInstantiate(&P->exec, graph, nullptr, nullptr, 0);
  (void)hipGraphDestroy(graph);
  if (e != hipSuccess) return dsplat_set_hip_error(e, "graph: instantiate");
  return DSPLAT_OK;
}

extern "C" int dsplat_mv_graph_launch(dsplat_mv_plan* P, void* stream) {
  if (!P || !P->exec) return dsplat_set_error(DSPLAT_ERR_ARG, "graph: nothing captured");
  DSPLAT_HIP(hipGraphLaunch(P->exec, as_stream(stream)));
  // every captured kind starts with a forward over the current views (a
  // change of views or colour mode drops the graph)
  P->fwd_valid = true;
  return DSPLAT_OK;
}
