// Single-view Gaussian raster and adjoint kernels: one workgroup per tile.
// The per-tile bodies (semantics of splat_gauss.cu:56-109 and the adjoint)
// are in raster_core.hpp.
#include "raster_core.hpp"

namespace dsplat {

// tile_order (may be null: raster order) maps workgroups to tiles, so the
// longest walks start first; tile_cost (may be null) receives each tile's
// longest walk.
// Minimum waves per SIMD the compiler must keep for a forward shape: the
// headline's 32x32 tile at 512 threads runs four workgroups per CU only
// within 64 VGPRs (8 waves per SIMD); other shapes are left to the compiler.
template <int TS, int NT>
constexpr int fwd_min_waves() { return TS == 32 && NT == 512 ? 8 : 1; }

template <int TS, int NT, typename Loader, bool kState>
__global__ void __launch_bounds__(NT) __attribute__((amdgpu_waves_per_eu(fwd_min_waves<TS, NT>())))
gauss_raster_fwd_kernel(uint32_t img_w, uint32_t img_h, uint32_t tw, Loader ld,
                        const uint32_t* __restrict__ tile2idx,
                        const uint32_t* __restrict__ idx2pnt, float* __restrict__ pix2rgb,
                        float* __restrict__ pix2alpha, uint32_t* __restrict__ pix2nwalk,
                        float* __restrict__ pix2tfinal, const uint32_t* __restrict__ tile_order,
                        uint32_t* __restrict__ tile_cost) {
  const uint32_t tile = tile_order ? tile_order[blockIdx.x] : blockIdx.x;
  if (tile >= gridDim.x) return;  // uniform: an order entry outside the grid is skipped
  raster_tile_fwd<TS, NT, Loader, kState>(tile % tw, tile / tw, img_w, img_h, tile2idx[tile],
                                  tile2idx[tile + 1], ld, idx2pnt, pix2rgb, pix2alpha, pix2nwalk,
                                  pix2tfinal, tile_cost, tile);
}

template <int TS, int NT>
__global__ void __launch_bounds__(NT)
gauss_raster_bwd_kernel(uint32_t img_w, uint32_t img_h, uint32_t tw,
                        const float4* __restrict__ splat2, const uint32_t* __restrict__ tile2idx,
                        const uint32_t* __restrict__ idx2pnt,
                        const uint32_t* __restrict__ pix2nwalk,
                        const float* __restrict__ pix2tfinal, const float* __restrict__ d_img,
                        const float* __restrict__ pix2rgb, const float* __restrict__ target,
                        const float* __restrict__ d_alpha, float* __restrict__ d_pixxy,
                        float* __restrict__ d_conic, float* __restrict__ d_op,
                        float* __restrict__ d_rgb, const uint32_t* __restrict__ tile_order) {
  const uint32_t tile = tile_order ? tile_order[blockIdx.x] : blockIdx.x;
  if (tile >= gridDim.x) return;  // uniform: an order entry outside the grid is skipped
  const GradSink out{d_pixxy, d_conic, d_op, d_rgb};
  raster_tile_bwd<TS, NT>(tile % tw, tile / tw, img_w, img_h, tile2idx[tile + 1], splat2,
                          idx2pnt, pix2nwalk, pix2tfinal, d_img, pix2rgb, target, d_alpha, 0u,
                          out);
}

// ----------------------------------------------------------- dispatch -----
template <typename Loader>
int launch_fwd(hipStream_t s, uint32_t w, uint32_t h, uint32_t ts, uint32_t threads,
               const Loader& ld, const uint32_t* tile2idx, const uint32_t* idx2pnt, float* rgb,
               float* alpha, uint32_t* nwalk, float* tfinal, const uint32_t* order = nullptr,
               uint32_t* cost = nullptr) {
  uint32_t tw, th;
  tile_grid(w, h, ts, &tw, &th);
  const uint32_t nt = tw * th;
  if (nt == 0) return DSPLAT_OK;
  // the adjoint's state and the tile costs only when asked for
  const bool state = nwalk || tfinal || cost;
#define DSPLAT_FWD(TS_, NT_)                                                                     \
  do {                                                                                           \
    if (state)                                                                                   \
      gauss_raster_fwd_kernel<TS_, NT_, Loader, true><<<nt, NT_, 0, s>>>(                        \
          w, h, tw, ld, tile2idx, idx2pnt, rgb, alpha, nwalk, tfinal, order, cost);                \
    else                                                                                         \
      gauss_raster_fwd_kernel<TS_, NT_, Loader, false><<<nt, NT_, 0, s>>>(                       \
          w, h, tw, ld, tile2idx, idx2pnt, rgb, alpha, nwalk, tfinal, order, cost);                \
  } while (0)
  if (ts == 8) {
    DSPLAT_FWD(8, 64);
  } else if (ts == 16) {
    if (threads == 64) DSPLAT_FWD(16, 64);
    else if (threads == 256) DSPLAT_FWD(16, 256);
    else DSPLAT_FWD(16, 128);
  } else if (ts == 32) {
    if (threads == 1024) DSPLAT_FWD(32, 1024);
    else if (threads == 256) DSPLAT_FWD(32, 256);
    else if (threads == 128) DSPLAT_FWD(32, 128);
    else DSPLAT_FWD(32, 512);
  } else {
    return dsplat_set_error(DSPLAT_ERR_ARG, "tile_size must be 8, 16 or 32");
  }
#undef DSPLAT_FWD
  DSPLAT_HIP(hipGetLastError());
  return DSPLAT_OK;
}

int launch_bwd(hipStream_t s, uint32_t w, uint32_t h, uint32_t ts, uint32_t threads,
               const float4* splat2, const uint32_t* tile2idx, const uint32_t* idx2pnt,
               const uint32_t* nwalk, const float* tfinal, const float* d_img,
               const float* d_alpha, float* d_pixxy, float* d_conic, float* d_op, float* d_rgb,
               const uint32_t* order = nullptr, const float* rgb = nullptr,
               const float* target = nullptr) {
  uint32_t tw, th;
  tile_grid(w, h, ts, &tw, &th);
  const uint32_t nt = tw * th;
  if (nt == 0) return DSPLAT_OK;
#define DSPLAT_BWD(TS_, NT_)                                                            \
  gauss_raster_bwd_kernel<TS_, NT_><<<nt, NT_, 0, s>>>(w, h, tw, splat2, tile2idx, idx2pnt, \
                                                       nwalk, tfinal, d_img, rgb, target,      \
                                                       d_alpha, d_pixxy, d_conic, d_op, d_rgb, \
                                                       order)
  if (ts == 8) {
    DSPLAT_BWD(8, 64);
  } else if (ts == 16) {
    if (threads == 128) DSPLAT_BWD(16, 128);
    else if (threads == 256) DSPLAT_BWD(16, 256);
    else DSPLAT_BWD(16, 64);
  } else if (ts == 32) {
    if (threads == 64) DSPLAT_BWD(32, 64);
    else if (threads == 128) DSPLAT_BWD(32, 128);
    else if (threads == 512) DSPLAT_BWD(32, 512);
    else if (threads == 1024) DSPLAT_BWD(32, 1024);
    else DSPLAT_BWD(32, 256);
  } else {
    return dsplat_set_error(DSPLAT_ERR_ARG, "tile_size must be 8, 16 or 32");
  }
#undef DSPLAT_BWD
  DSPLAT_HIP(hipGetLastError());
  return DSPLAT_OK;
}

}  // namespace dsplat

using namespace dsplat;

extern "C" int dsplat_gauss_rasterize_pnt2splat2(void* stream, uint32_t img_w, uint32_t img_h,
                                                 float* pix2rgb, const float* pnt2pixxydepth,
                                                 const float* pnt2pixconvinv,
                                                 const float* pnt2pixaabb,
                                                 const float* pnt2opacity, const float* pnt2rgb,
                                                 uint32_t tile_size, const uint32_t* tile2idx,
                                                 const uint32_t* idx2pnt) {
  if (!pix2rgb || !tile2idx || !pnt2pixxydepth || !pnt2pixconvinv || !pnt2pixaabb ||
      !pnt2opacity || !pnt2rgb)
    return dsplat_set_error(DSPLAT_ERR_ARG, "rasterize: null buffer");
  if (reinterpret_cast<uintptr_t>(pnt2pixaabb) & 15u)
    return dsplat_set_error(DSPLAT_ERR_ARG, "rasterize: aabb must be 16-byte aligned");
  ArrayLoader ld{pnt2pixxydepth, pnt2pixconvinv, reinterpret_cast<const float4*>(pnt2pixaabb),
                 pnt2opacity, pnt2rgb};
  return launch_fwd(as_stream(stream), img_w, img_h, tile_size, 0, ld, tile2idx, idx2pnt,
                    pix2rgb, nullptr, nullptr, nullptr);
}

extern "C" int dsplat_gauss_rasterize_fwd(void* stream, uint32_t img_w, uint32_t img_h,
                                          uint32_t tile_size, uint32_t threads,
                                          const float* pnt2splat2, const uint32_t* tile2idx,
                                          const uint32_t* idx2pnt, float* pix2rgb,
                                          float* pix2alpha, uint32_t* pix2nwalk,
                                          float* pix2tfinal) {
  if (!pix2rgb || !tile2idx || !pnt2splat2)
    return dsplat_set_error(DSPLAT_ERR_ARG, "rasterize_fwd: null buffer");
  if (reinterpret_cast<uintptr_t>(pnt2splat2) & 15u)
    return dsplat_set_error(DSPLAT_ERR_ARG, "rasterize_fwd: splat2 must be 16-byte aligned");
  PackedLoader ld{reinterpret_cast<const float4*>(pnt2splat2)};
  return launch_fwd(as_stream(stream), img_w, img_h, tile_size, threads, ld, tile2idx, idx2pnt,
                    pix2rgb, pix2alpha, pix2nwalk, pix2tfinal);
}

extern "C" int dsplat_gauss_rasterize_bwd(void* stream, uint32_t img_w, uint32_t img_h,
                                          uint32_t tile_size, uint32_t threads,
                                          const float* pnt2splat2, const uint32_t* tile2idx,
                                          const uint32_t* idx2pnt, const uint32_t* pix2nwalk,
                                          const float* pix2tfinal, const float* d_pix2rgb,
                                          const float* d_pix2alpha, float* d_pnt2pixxy,
                                          float* d_pnt2conic, float* d_pnt2opacity,
                                          float* d_pnt2rgb) {
  if (!pnt2splat2 || !tile2idx || !pix2nwalk || !pix2tfinal || !d_pix2rgb || !d_pnt2pixxy ||
      !d_pnt2conic || !d_pnt2opacity || !d_pnt2rgb)
    return dsplat_set_error(DSPLAT_ERR_ARG, "rasterize_bwd: null buffer");
  if (reinterpret_cast<uintptr_t>(pnt2splat2) & 15u)
    return dsplat_set_error(DSPLAT_ERR_ARG, "rasterize_bwd: splat2 must be 16-byte aligned");
  return launch_bwd(as_stream(stream), img_w, img_h, tile_size, threads,
                    reinterpret_cast<const float4*>(pnt2splat2), tile2idx, idx2pnt, pix2nwalk,
                    pix2tfinal, d_pix2rgb, d_pix2alpha, d_pnt2pixxy, d_pnt2conic, d_pnt2opacity,
                    d_pnt2rgb);
}

extern "C" int dsplat_gauss_rasterize_fwd_ordered(void* stream, uint32_t img_w, uint32_t img_h,
                                                  uint32_t tile_size, uint32_t threads,
                                                  const float* pnt2splat2,
                                                  const uint32_t* tile2idx,
                                                  const uint32_t* idx2pnt,
                                                  const uint32_t* tile_order, float* pix2rgb,
                                                  float* pix2alpha, uint32_t* pix2nwalk,
                                                  float* pix2tfinal, uint32_t* tile_cost) {
  if (!pix2rgb || !tile2idx || !pnt2splat2)
    return dsplat_set_error(DSPLAT_ERR_ARG, "rasterize_fwd: null buffer");
  if (reinterpret_cast<uintptr_t>(pnt2splat2) & 15u)
    return dsplat_set_error(DSPLAT_ERR_ARG, "rasterize_fwd: splat2 must be 16-byte aligned");
  PackedLoader ld{reinterpret_cast<const float4*>(pnt2splat2)};
  return launch_fwd(as_stream(stream), img_w, img_h, tile_size, threads, ld, tile2idx, idx2pnt,
                    pix2rgb, pix2alpha, pix2nwalk, pix2tfinal, tile_order, tile_cost);
}

extern "C" int dsplat_gauss_rasterize_bwd_ordered(
    void* stream, uint32_t img_w, uint32_t img_h, uint32_t tile_size, uint32_t threads,
    const float* pnt2splat2, const uint32_t* tile2idx, const uint32_t* idx2pnt,
    const uint32_t* tile_order, const uint32_t* pix2nwalk, const float* pix2tfinal,
    const float* d_pix2rgb, const float* d_pix2alpha, float* d_pnt2pixxy, float* d_pnt2conic,
    float* d_pnt2opacity, float* d_pnt2rgb) {
  if (!pnt2splat2 || !tile2idx || !pix2nwalk || !pix2tfinal || !d_pix2rgb || !d_pnt2pixxy ||
      !d_pnt2conic || !d_pnt2opacity || !d_pnt2rgb)
    return dsplat_set_error(DSPLAT_ERR_ARG, "rasterize_bwd: null buffer");
  if (reinterpret_cast<uintptr_t>(pnt2splat2) & 15u)
    return dsplat_set_error(DSPLAT_ERR_ARG, "rasterize_bwd: splat2 must be 16-byte aligned");
  return launch_bwd(as_stream(stream), img_w, img_h, tile_size, threads,
                    reinterpret_cast<const float4*>(pnt2splat2), tile2idx, idx2pnt, pix2nwalk,
                    pix2tfinal, d_pix2rgb, d_pix2alpha, d_pnt2pixxy, d_pnt2conic, d_pnt2opacity,
                    d_pnt2rgb, tile_order);
}

extern "C" int dsplat_gauss_rasterize_bwd_l2(
    void* stream, uint32_t img_w, uint32_t img_h, uint32_t tile_size, uint32_t threads,
    const float* pnt2splat2, const uint32_t* tile2idx, const uint32_t* idx2pnt,
    const uint32_t* tile_order, const uint32_t* pix2nwalk, const float* pix2tfinal,
    const float* pix2rgb, const float* target, float* d_pnt2pixxy, float* d_pnt2conic,
    float* d_pnt2opacity, float* d_pnt2rgb) {
  if (!pnt2splat2 || !tile2idx || !pix2nwalk || !pix2tfinal || !pix2rgb || !target ||
      !d_pnt2pixxy || !d_pnt2conic || !d_pnt2opacity || !d_pnt2rgb)
    return dsplat_set_error(DSPLAT_ERR_ARG, "rasterize_bwd_l2: null buffer");
  if (reinterpret_cast<uintptr_t>(pnt2splat2) & 15u)
    return dsplat_set_error(DSPLAT_ERR_ARG, "rasterize_bwd_l2: splat2 must be 16-byte aligned");
  return launch_bwd(as_stream(stream), img_w, img_h, tile_size, threads,
                    reinterpret_cast<const float4*>(pnt2splat2), tile2idx, idx2pnt, pix2nwalk,
                    pix2tfinal, nullptr, nullptr, d_pnt2pixxy, d_pnt2conic, d_pnt2opacity,
                    d_pnt2rgb, tile_order, pix2rgb, target);
}
