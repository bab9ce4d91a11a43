// Gaussian 3D->2D projection and its adjoint (single view). The math lives
// in project_math.hpp (semantics of splat_gauss.cu:10-53).
#include "binning.hpp"
#include "project_math.hpp"

namespace dsplat {

// Binning outputs of the projection (BinOut.rect null: none): each
// Gaussian's tile rectangle and depth key, as the binning's count phase
// derives them from the box and the depth.
struct BinOut {
  uint2* rect;
  uint32_t* dkey;
  uint32_t ts, tw, th;
};

template <bool kLogScale, bool kPacked>
__global__ void __launch_bounds__(256)
gauss_project_kernel(uint32_t n, const float* __restrict__ xyz, const float* __restrict__ scale,
                     const float4* __restrict__ quat, const float* __restrict__ opacity,
                     const float* __restrict__ rgb, const float* __restrict__ m,
                     uint32_t img_w, uint32_t img_h, float* __restrict__ pixxydepth,
                     float* __restrict__ conic, float4* __restrict__ aabb,
                     float4* __restrict__ splat2, float* __restrict__ zero9, BinOut bo) {
  // the packed records leave through LDS: each wave writes its 64 records
  // (4 KB) with four fully contiguous 1 KB stores instead of four strided
  // quarter-record stores (which leave every line partly written by each
  // store; measured 47 -> 39 us at 1M Gaussians)
  __shared__ float4 srec[kPacked ? 4 * 256 : 1];
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  const uint32_t lane = threadIdx.x & 63u, wbase = threadIdx.x & ~63u;
  const uint32_t g0 = blockIdx.x * blockDim.x + wbase;  // first Gaussian of the wave
  if (g0 >= n) return;  // wave-uniform
  if (zero9) {
    // the adjoint's accumulators of the wave's Gaussians (9 floats each),
    // cleared with coalesced stores
    const uint32_t z1 = 9u * min(g0 + 64u, n);
#pragma unroll
    for (uint32_t k = 0; k < 9; ++k) {
      const uint32_t z = 9u * g0 + k * 64u + lane;
      if (z < z1) zero9[z] = 0.f;
    }
  }
  if (i < n) {
    const float4 q = quat[i];
    float s0 = scale[3 * i], s1 = scale[3 * i + 1], s2 = scale[3 * i + 2];
    if (kLogScale) {
      s0 = expf(s0);
      s1 = expf(s1);
      s2 = expf(s2);
    }
    ProjState p;
    project_state(m, xyz[3 * i], xyz[3 * i + 1], xyz[3 * i + 2], q.x, q.y, q.z, q.w, s0, s1,
                  s2, img_w, img_h, p);
    if (pixxydepth) {
      pixxydepth[3 * i + 0] = p.pix[0];
      pixxydepth[3 * i + 1] = p.pix[1];
      pixxydepth[3 * i + 2] = p.ndcz;
    }
    if (conic) {  // the packed record holds the conic too
      conic[3 * i + 0] = p.conic[0];
      conic[3 * i + 1] = p.conic[1];
      conic[3 * i + 2] = p.conic[2];
    }
    if (aabb) aabb[i] = make_float4(p.aabb[0], p.aabb[1], p.aabb[2], p.aabb[3]);
    if (bo.rect) {
      bo.rect[i] = pack_rect(tile_rect(p.aabb[0], p.aabb[1], p.aabb[2], p.aabb[3], bo.ts, bo.tw,
                                       bo.th));
      bo.dkey[i] = depth_key(p.ndcz);
    }
    if (kPacked) {
      float4* mine = srec + 4 * threadIdx.x;
      mine[0] = make_float4(p.pix[0], p.pix[1], p.conic[0], p.conic[1]);
      mine[1] = make_float4(p.conic[2], opacity[i], rgb[3 * i], rgb[3 * i + 1]);
      mine[2] = make_float4(rgb[3 * i + 2], p.aabb[0], p.aabb[1], p.aabb[2]);
      mine[3] = make_float4(p.aabb[3], 0.f, 0.f, 0.f);
    }
  }
  if (kPacked) {
    // every lane of the wave is past its LDS writes: hand the records over
    wave_lds_fence();
    const uint32_t nrec4 = 4u * min(64u, n - g0);
#pragma unroll
    for (int c = 0; c < 4; ++c) {
      const uint32_t k = uint32_t(c) * 64u + lane;
      if (k < nrec4) splat2[4 * size_t(g0) + k] = srec[4 * wbase + k];
    }
  }
}

// Adjoint: the chain rule from (pixel centre, conic) into the mean, scale
// and quaternion gradients, added to them or (kOverwrite) stored.
template <bool kLogScale, bool kOverwrite>
__global__ void __launch_bounds__(256)
gauss_project_bwd_kernel(uint32_t n, const float* __restrict__ xyz,
                         const float* __restrict__ scale, const float4* __restrict__ quat,
                         const float* __restrict__ m, uint32_t img_w, uint32_t img_h,
                         const float2* __restrict__ d_pixxy, const float* __restrict__ d_conic,
                         float* __restrict__ d_xyz, float* __restrict__ d_scale,
                         float4* __restrict__ d_quat) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  float sg[3] = {scale[3 * i], scale[3 * i + 1], scale[3 * i + 2]};
  if (kLogScale) {
    sg[0] = expf(sg[0]);
    sg[1] = expf(sg[1]);
    sg[2] = expf(sg[2]);
  }
  const float x3[3] = {xyz[3 * i], xyz[3 * i + 1], xyz[3 * i + 2]};
  const float gcon[3] = {d_conic[3 * i], d_conic[3 * i + 1], d_conic[3 * i + 2]};
  float dx[3], dsc[3], dq[4];
  if (!project_bwd_one(m, img_w, img_h, x3, quat[i], sg, kLogScale, d_pixxy[i], gcon, dx, dsc,
                       dq)) {
    if (!kOverwrite) return;
    dx[0] = dx[1] = dx[2] = dsc[0] = dsc[1] = dsc[2] = 0.f;
    dq[0] = dq[1] = dq[2] = dq[3] = 0.f;
  }
  float4 q = kOverwrite ? make_float4(0.f, 0.f, 0.f, 0.f) : d_quat[i];
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    d_xyz[3 * i + c] = kOverwrite ? dx[c] : d_xyz[3 * i + c] + dx[c];
    d_scale[3 * i + c] = kOverwrite ? dsc[c] : d_scale[3 * i + c] + dsc[c];
  }
  q.x += dq[0];
  q.y += dq[1];
  q.z += dq[2];
  q.w += dq[3];
  d_quat[i] = q;
}

int launch_gauss_project(hipStream_t s, uint32_t n, const float* xyz, const float* scale,
                         const float* quat, const float* opacity, const float* rgb,
                         const float* m, uint32_t w, uint32_t h, bool log_scale,
                         float* pixxydepth, float* conic, float* aabb, float* splat2,
                         float* zero9, BinOut bo = BinOut{nullptr, nullptr, 0, 0, 0}) {
  if (n == 0) return DSPLAT_OK;
  const uint32_t nb = div_up(n, 256);
  const float4* q4 = reinterpret_cast<const float4*>(quat);
  float4* bb = reinterpret_cast<float4*>(aabb);
  float4* rec = reinterpret_cast<float4*>(splat2);
  if (splat2) {
    if (log_scale)
      gauss_project_kernel<true, true><<<nb, 256, 0, s>>>(n, xyz, scale, q4, opacity, rgb, m, w, h, pixxydepth, conic, bb, rec, zero9, bo);
    else
      gauss_project_kernel<false, true><<<nb, 256, 0, s>>>(n, xyz, scale, q4, opacity, rgb, m, w, h, pixxydepth, conic, bb, rec, zero9, bo);
  } else {
    if (log_scale)
      gauss_project_kernel<true, false><<<nb, 256, 0, s>>>(n, xyz, scale, q4, opacity, rgb, m, w, h, pixxydepth, conic, bb, rec, zero9, bo);
    else
      gauss_project_kernel<false, false><<<nb, 256, 0, s>>>(n, xyz, scale, q4, opacity, rgb, m, w, h, pixxydepth, conic, bb, rec, zero9, bo);
  }
  DSPLAT_HIP(hipGetLastError());
  return DSPLAT_OK;
}

}  // namespace dsplat

using namespace dsplat;

static bool aligned16(const void* p) { return (reinterpret_cast<uintptr_t>(p) & 15u) == 0; }

extern "C" int dsplat_gauss_pnt2splat3_to_pnt2splat2(void* stream, uint32_t num_pnt,
                                                     const float* pnt2xyz, const float* pnt2scale,
                                                     const float* pnt2quat, float* pnt2pixxydepth,
                                                     float* pnt2pixconvinv, float* pnt2pixaabb,
                                                     const float* transform_world2ndc,
                                                     uint32_t img_w, uint32_t img_h) {
  if (num_pnt && (!pnt2xyz || !pnt2scale || !pnt2quat || !pnt2pixxydepth || !pnt2pixconvinv ||
                  !pnt2pixaabb || !transform_world2ndc))
    return dsplat_set_error(DSPLAT_ERR_ARG, "gauss projection: null buffer");
  if (!aligned16(pnt2quat) || !aligned16(pnt2pixaabb))
    return dsplat_set_error(DSPLAT_ERR_ARG, "gauss projection: quat/aabb must be 16-byte aligned");
  return launch_gauss_project(as_stream(stream), num_pnt, pnt2xyz, pnt2scale, pnt2quat, nullptr,
                              nullptr, transform_world2ndc, img_w, img_h, false, pnt2pixxydepth,
                              pnt2pixconvinv, pnt2pixaabb, nullptr, nullptr);
}

extern "C" int dsplat_gauss_project_fwd(void* stream, uint32_t num_pnt, const float* pnt2xyz,
                                        const float* pnt2scale, const float* pnt2quat,
                                        const float* pnt2opacity, const float* pnt2rgb,
                                        const float* transform_world2ndc, uint32_t img_w,
                                        uint32_t img_h, uint32_t flags, float* pnt2pixxydepth,
                                        float* pnt2pixconvinv, float* pnt2pixaabb,
                                        float* pnt2splat2, float* d_grad9) {
  if (num_pnt && (!pnt2xyz || !pnt2scale || !pnt2quat || !pnt2opacity || !pnt2rgb ||
                  !pnt2pixxydepth || !pnt2pixaabb || !pnt2splat2 || !transform_world2ndc))
    return dsplat_set_error(DSPLAT_ERR_ARG, "gauss project_fwd: null buffer");
  if (!aligned16(pnt2quat) || !aligned16(pnt2pixaabb) || !aligned16(pnt2splat2))
    return dsplat_set_error(DSPLAT_ERR_ARG, "gauss project_fwd: quat/aabb/splat2 must be 16-byte aligned");
  return launch_gauss_project(as_stream(stream), num_pnt, pnt2xyz, pnt2scale, pnt2quat,
                              pnt2opacity, pnt2rgb, transform_world2ndc, img_w, img_h,
                              (flags & DSPLAT_SCALE_IS_LOG) != 0, pnt2pixxydepth, pnt2pixconvinv,
                              pnt2pixaabb, pnt2splat2, d_grad9);
}

extern "C" int dsplat_gauss_project_bin_ws(
    void* stream, uint32_t num_pnt, const float* pnt2xyz, const float* pnt2scale,
    const float* pnt2quat, const float* pnt2opacity, const float* pnt2rgb,
    const float* transform_world2ndc, uint32_t img_w, uint32_t img_h, uint32_t flags,
    float* pnt2pixxydepth, float* pnt2pixaabb, float* pnt2splat2, float* d_grad9,
    uint32_t tile_size, uint32_t* tile2idx, uint32_t* idx2pnt, uint32_t capacity,
    void* workspace, size_t workspace_bytes, uint32_t* d_num_idx) {
  if (tile_size != 8 && tile_size != 16 && tile_size != 32)
    return dsplat_set_error(DSPLAT_ERR_ARG, "gauss project_bin: tile_size must be 8, 16 or 32");
  if ((num_pnt && (!pnt2xyz || !pnt2scale || !pnt2quat || !pnt2opacity || !pnt2rgb ||
                   !pnt2splat2 || !transform_world2ndc)) ||
      !tile2idx || !d_num_idx || !workspace || (capacity && !idx2pnt))
    return dsplat_set_error(DSPLAT_ERR_ARG, "gauss project_bin: null buffer");
  if (!aligned16(pnt2quat) || !aligned16(pnt2pixaabb) || !aligned16(pnt2splat2))
    return dsplat_set_error(DSPLAT_ERR_ARG,
                            "gauss project_bin: quat/aabb/splat2 must be 16-byte aligned");
  uint32_t tw, th;
  tile_grid(img_w, img_h, tile_size, &tw, &th);
  if (tw > 0xFFFFu || th > 0xFFFFu)  // tile rectangles are packed as 16-bit coordinates
    return dsplat_set_error(DSPLAT_ERR_ARG, "gauss project_bin: more than 65535 tiles along an axis");
  if (workspace_bytes < bin_workspace_bytes(num_pnt, capacity))
    return dsplat_set_error(DSPLAT_ERR_CAPACITY, "gauss project_bin: workspace too small");
  hipStream_t s = as_stream(stream);
  BinOut bo{nullptr, nullptr, tile_size, tw, th};
  bin_count_slots(workspace, num_pnt, &bo.rect, &bo.dkey);
  const int rc = launch_gauss_project(s, num_pnt, pnt2xyz, pnt2scale, pnt2quat, pnt2opacity,
                                      pnt2rgb, transform_world2ndc, img_w, img_h,
                                      (flags & DSPLAT_SCALE_IS_LOG) != 0, pnt2pixxydepth,
                                      nullptr, pnt2pixaabb, pnt2splat2, d_grad9, bo);
  if (rc) return rc;
  return bin_after_count(s, num_pnt, tile_size, tw, th, tile2idx, idx2pnt, capacity, workspace,
                         workspace_bytes, d_num_idx);
}

extern "C" int dsplat_gauss_project_bwd(void* stream, uint32_t num_pnt, const float* pnt2xyz,
                                        const float* pnt2scale, const float* pnt2quat,
                                        const float* transform_world2ndc, uint32_t img_w,
                                        uint32_t img_h, uint32_t flags, const float* d_pnt2pixxy,
                                        const float* d_pnt2conic, float* d_pnt2xyz,
                                        float* d_pnt2scale, float* d_pnt2quat) {
  if (num_pnt == 0) return DSPLAT_OK;
  if (!pnt2xyz || !pnt2scale || !pnt2quat || !transform_world2ndc || !d_pnt2pixxy ||
      !d_pnt2conic || !d_pnt2xyz || !d_pnt2scale || !d_pnt2quat)
    return dsplat_set_error(DSPLAT_ERR_ARG, "gauss project_bwd: null buffer");
  if (!aligned16(pnt2quat) || !aligned16(d_pnt2quat) ||
      (reinterpret_cast<uintptr_t>(d_pnt2pixxy) & 7u))
    return dsplat_set_error(DSPLAT_ERR_ARG, "gauss project_bwd: misaligned quat/grad buffer");
  hipStream_t s = as_stream(stream);
  const uint32_t nb = div_up(num_pnt, 256);
  const float4* q4 = reinterpret_cast<const float4*>(pnt2quat);
  const float2* gp = reinterpret_cast<const float2*>(d_pnt2pixxy);
  float4* dq = reinterpret_cast<float4*>(d_pnt2quat);
#define DSPLAT_PBWD(L_, O_)                                                                \
  gauss_project_bwd_kernel<L_, O_><<<nb, 256, 0, s>>>(num_pnt, pnt2xyz, pnt2scale, q4,   \
                                                      transform_world2ndc, img_w, img_h, gp, \
                                                      d_pnt2conic, d_pnt2xyz, d_pnt2scale, dq)
  const bool log_scale = flags & DSPLAT_SCALE_IS_LOG, overwrite = flags & DSPLAT_GRAD_OVERWRITE;
  if (log_scale && overwrite) DSPLAT_PBWD(true, true);
  else if (log_scale) DSPLAT_PBWD(true, false);
  else if (overwrite) DSPLAT_PBWD(false, true);
  else DSPLAT_PBWD(false, false);
#undef DSPLAT_PBWD
  DSPLAT_HIP(hipGetLastError());
  return DSPLAT_OK;
}
