// Tiled front-to-back Gaussian rasterization and its adjoint.
//
// Forward semantics: rasterize_splat_using_tile (splat_gauss.cu:56-109). For
// the pixel centre t = (ix+.5, iy+.5) the tile's list is walked from its end
// (largest depth key = nearest) to its start; a splat is skipped when t lies
// outside its box; e = exp(-q/2) * opacity with q = d^T C d, d = t - centre;
// rgb += colour * T e; T *= 1 - e; alpha_sum += T e; the walk stops once
// alpha_sum > 0.999 (the reference compares the float sum with the double
// 0.999, which is the same as alpha_sum >= 0.999f: 0.999f is the smallest
// float above the double 0.999).
//
// One workgroup per tile. Records of NT list entries are staged in LDS as
// float4 quarters (broadcast reads), every thread walks them for its pixels,
// and the workgroup leaves as soon as every pixel has terminated
// (__syncthreads_and).
//
// Adjoint: every pixel walks its list back-to-front from the last entry the
// forward visited (pix2nwalk) with the transmittance recovered by division
// from pix2tlast (the transmittance in front of the last contributing
// splat; dividing by 1 - e of earlier splats is safe because the walk did not
// stop there, so T stayed above ~1e-3). Per list entry the nine gradient
// terms of all pixels are summed: across each 16-lane row with DPP, across
// rows and waves with LDS float atomics, and once per (tile, entry) into
// global memory with float atomics.
#include "dsplat_common.hpp"

namespace dsplat {

// ------------------------------------------------------------- loaders -----
// Packed records (project_fwd output), 4 x float4 per Gaussian.
struct PackedLoader {
  const float4* rec;
  __device__ void load(uint32_t p, float4& q0, float4& q1, float4& q2, float4& q3) const {
    const float4* r = rec + 4 * size_t(p);
    q0 = r[0];
    q1 = r[1];
    q2 = r[2];
    q3 = r[3];
  }
};

// The reference's separate arrays (splat_gauss.cu:61-66).
struct ArrayLoader {
  const float* pixxydepth;
  const float* conic;
  const float4* aabb;
  const float* opacity;
  const float* rgb;
  __device__ void load(uint32_t p, float4& q0, float4& q1, float4& q2, float4& q3) const {
    const float4 bb = aabb[p];
    q0 = make_float4(pixxydepth[3 * p], pixxydepth[3 * p + 1], conic[3 * p], conic[3 * p + 1]);
    q1 = make_float4(conic[3 * p + 2], opacity[p], rgb[3 * p], rgb[3 * p + 1]);
    q2 = make_float4(rgb[3 * p + 2], bb.x, bb.y, bb.z);
    q3 = make_float4(bb.w, 0.f, 0.f, 0.f);
  }
};

__device__ inline float fast_exp(float x) { return __expf(x); }

// --------------------------------------------------------------- forward ---
template <int TS, int NT, typename Loader>
__global__ void __launch_bounds__(NT)
gauss_raster_fwd_kernel(uint32_t img_w, uint32_t img_h, uint32_t tw, Loader ld,
                        const uint32_t* __restrict__ tile2idx,
                        const uint32_t* __restrict__ idx2pnt, float* __restrict__ pix2rgb,
                        float* __restrict__ pix2alpha, uint32_t* __restrict__ pix2nwalk,
                        float* __restrict__ pix2tlast) {
  constexpr int PPT = TS * TS / NT;
  __shared__ float4 s0[NT], s1[NT], s2[NT], s3[NT];
  const uint32_t tile = blockIdx.x;
  const uint32_t tx = tile % tw, ty = tile / tw;
  const uint32_t start = tile2idx[tile], end = tile2idx[tile + 1];
  const uint32_t num = end - start;
  float cr[PPT], cg[PPT], cbl[PPT], asum[PPT], tr[PPT], tl[PPT], px[PPT], py[PPT];
  uint32_t walked[PPT];
  bool done[PPT];
#pragma unroll
  for (int k = 0; k < PPT; ++k) {
    const uint32_t lp = threadIdx.x + k * NT;
    const uint32_t ix = tx * TS + lp % TS, iy = ty * TS + lp / TS;
    px[k] = float(ix) + 0.5f;
    py[k] = float(iy) + 0.5f;
    cr[k] = cg[k] = cbl[k] = asum[k] = 0.f;
    tr[k] = tl[k] = 1.f;
    walked[k] = num;
    done[k] = !(ix < img_w && iy < img_h);
  }
  for (uint32_t b = 0; b < num; b += NT) {
    const uint32_t it = b + threadIdx.x;
    if (it < num) {
      float4 q0, q1, q2, q3;
      ld.load(idx2pnt[end - 1 - it], q0, q1, q2, q3);
      s0[threadIdx.x] = q0;
      s1[threadIdx.x] = q1;
      s2[threadIdx.x] = q2;
      s3[threadIdx.x] = q3;
    }
    __syncthreads();
    const uint32_t cnt = min(uint32_t(NT), num - b);
#pragma unroll
    for (int k = 0; k < PPT; ++k) {
      if (done[k]) continue;
      for (uint32_t j = 0; j < cnt; ++j) {
        const float4 q2 = s2[j];
        const float4 q3 = s3[j];
        // inclusive box test (aabb2::is_include_point)
        if (!(px[k] >= q2.y && px[k] <= q2.w && py[k] >= q2.z && py[k] <= q3.x)) continue;
        const float4 q0 = s0[j];
        const float4 q1 = s1[j];
        const float dx = px[k] - q0.x, dy = py[k] - q0.y;
        const float q = q0.z * dx * dx + q0.w * (dx * dy + dy * dx) + q1.x * dy * dy;
        const float e = fast_exp(-0.5f * q) * q1.y;
        const float eo = tr[k] * e;
        cr[k] += q1.z * eo;
        cg[k] += q1.w * eo;
        cbl[k] += q2.x * eo;
        tl[k] = tr[k];
        tr[k] *= 1.f - e;
        asum[k] += eo;
        if (asum[k] >= 0.999f) {
          done[k] = true;
          walked[k] = b + j + 1;
          break;
        }
      }
    }
    bool all = true;
#pragma unroll
    for (int k = 0; k < PPT; ++k) all = all && done[k];
    if (__syncthreads_and(all)) break;
  }
#pragma unroll
  for (int k = 0; k < PPT; ++k) {
    const uint32_t lp = threadIdx.x + k * NT;
    const uint32_t ix = tx * TS + lp % TS, iy = ty * TS + lp / TS;
    if (ix >= img_w || iy >= img_h) continue;
    const size_t ip = size_t(iy) * img_w + ix;
    pix2rgb[3 * ip + 0] = cr[k];
    pix2rgb[3 * ip + 1] = cg[k];
    pix2rgb[3 * ip + 2] = cbl[k];
    if (pix2alpha) pix2alpha[ip] = asum[k];
    if (pix2nwalk) pix2nwalk[ip] = walked[k];
    if (pix2tlast) pix2tlast[ip] = tl[k];
  }
}

// -------------------------------------------------------------- backward ---
template <int CTRL>
__device__ inline float dpp_mov(float v) {
  return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), CTRL, 0xF, 0xF, false));
}

// Sum over each 16-lane row; every lane of the row receives the row sum.
__device__ inline float row_sum16(float v) {
  v += dpp_mov<0xB1>(v);   // quad_perm [1,0,3,2]
  v += dpp_mov<0x4E>(v);   // quad_perm [2,3,0,1]
  v += dpp_mov<0x141>(v);  // row_half_mirror
  v += dpp_mov<0x140>(v);  // row_mirror
  return v;
}

constexpr int kGrad = 9;  // dx, dy, dca, dcb, dcc, dop, dr, dg, db

template <int TS, int NT>
__global__ void __launch_bounds__(NT)
gauss_raster_bwd_kernel(uint32_t img_w, uint32_t img_h, uint32_t tw,
                        const float4* __restrict__ splat2, const uint32_t* __restrict__ tile2idx,
                        const uint32_t* __restrict__ idx2pnt,
                        const uint32_t* __restrict__ pix2nwalk,
                        const float* __restrict__ pix2tlast, const float* __restrict__ d_img,
                        const float* __restrict__ d_alpha, float* __restrict__ d_pixxy,
                        float* __restrict__ d_conic, float* __restrict__ d_op,
                        float* __restrict__ d_rgb) {
  constexpr int PPT = TS * TS / NT;
  __shared__ float4 s0[NT], s1[NT], s2[NT], s3[NT];
  __shared__ uint32_t spid[NT];
  __shared__ float acc[kGrad][NT];
  __shared__ uint32_t smax;
  const uint32_t tile = blockIdx.x;
  const uint32_t tx = tile % tw, ty = tile / tw;
  const uint32_t end = tile2idx[tile + 1];
  const int lane = threadIdx.x & 63;
  float px[PPT], py[PPT], tr[PPT], g0[PPT], g1[PPT], g2[PPT], ga[PPT];
  float b0[PPT], b1[PPT], b2[PPT], ba[PPT];
  uint32_t nw[PPT];
  bool first[PPT];
  uint32_t my_max = 0;
#pragma unroll
  for (int k = 0; k < PPT; ++k) {
    const uint32_t lp = threadIdx.x + k * NT;
    const uint32_t ix = tx * TS + lp % TS, iy = ty * TS + lp / TS;
    px[k] = float(ix) + 0.5f;
    py[k] = float(iy) + 0.5f;
    first[k] = true;
    b0[k] = b1[k] = b2[k] = ba[k] = 0.f;
    if (ix < img_w && iy < img_h) {
      const size_t ip = size_t(iy) * img_w + ix;
      nw[k] = pix2nwalk[ip];
      tr[k] = pix2tlast[ip];
      g0[k] = d_img[3 * ip];
      g1[k] = d_img[3 * ip + 1];
      g2[k] = d_img[3 * ip + 2];
      ga[k] = d_alpha ? d_alpha[ip] : 0.f;
    } else {
      nw[k] = 0;
      tr[k] = 1.f;
      g0[k] = g1[k] = g2[k] = ga[k] = 0.f;
    }
    my_max = max(my_max, nw[k]);
  }
  if (threadIdx.x == 0) smax = 0;
  __syncthreads();
  atomicMax(&smax, my_max);
  __syncthreads();
  const uint32_t maxw = smax;
  for (uint32_t hi = maxw; hi > 0;) {
    const uint32_t lo = hi > uint32_t(NT) ? hi - NT : 0u;
    const uint32_t cnt = hi - lo;
    if (threadIdx.x < cnt) {
      const uint32_t it = hi - 1 - threadIdx.x;
      const uint32_t p = idx2pnt[end - 1 - it];
      const float4* r = splat2 + 4 * size_t(p);
      s0[threadIdx.x] = r[0];
      s1[threadIdx.x] = r[1];
      s2[threadIdx.x] = r[2];
      s3[threadIdx.x] = r[3];
      spid[threadIdx.x] = p;
#pragma unroll
      for (int c = 0; c < kGrad; ++c) acc[c][threadIdx.x] = 0.f;
    }
    __syncthreads();
    for (uint32_t j = 0; j < cnt; ++j) {
      const uint32_t it = hi - 1 - j;
      const float4 q0 = s0[j], q1 = s1[j], q2 = s2[j], q3 = s3[j];
      float v[kGrad];
#pragma unroll
      for (int c = 0; c < kGrad; ++c) v[c] = 0.f;
      bool any = false;
#pragma unroll
      for (int k = 0; k < PPT; ++k) {
        const bool in = it < nw[k] && px[k] >= q2.y && px[k] <= q2.w && py[k] >= q2.z &&
                        py[k] <= q3.x;
        if (!in) continue;
        any = true;
        const float dx = px[k] - q0.x, dy = py[k] - q0.y;
        const float q = q0.z * dx * dx + q0.w * (dx * dy + dy * dx) + q1.x * dy * dy;
        const float gk = fast_exp(-0.5f * q);
        const float e = gk * q1.y;
        const float ome = 1.f - e;
        if (!first[k]) tr[k] = tr[k] / ome;
        first[k] = false;
        const float t = tr[k];
        const float inv = 1.f / ome;
        const float dl_de = g0[k] * (q1.z * t - b0[k] * inv) + g1[k] * (q1.w * t - b1[k] * inv) +
                            g2[k] * (q2.x * t - b2[k] * inv) + ga[k] * (t - ba[k] * inv);
        const float et = e * t;
        v[6] += g0[k] * et;
        v[7] += g1[k] * et;
        v[8] += g2[k] * et;
        b0[k] += q1.z * et;
        b1[k] += q1.w * et;
        b2[k] += q2.x * et;
        ba[k] += et;
        v[5] += dl_de * gk;
        const float dl_dq = dl_de * e * -0.5f;
        v[2] += dl_dq * dx * dx;
        v[3] += dl_dq * 2.f * dx * dy;
        v[4] += dl_dq * dy * dy;
        v[0] -= dl_dq * (2.f * q0.z * dx + 2.f * q0.w * dy);
        v[1] -= dl_dq * (2.f * q0.w * dx + 2.f * q1.x * dy);
      }
      if (__ballot(any) == 0) continue;  // wave-uniform
#pragma unroll
      for (int c = 0; c < kGrad; ++c) {
        const float r = row_sum16(v[c]);
        if ((lane & 15) == 0) atomicAdd(&acc[c][j], r);
      }
    }
    __syncthreads();
    if (threadIdx.x < cnt) {
      const uint32_t p = spid[threadIdx.x];
      const float a0 = acc[0][threadIdx.x], a1 = acc[1][threadIdx.x];
      if (a0 != 0.f) atomicAdd(&d_pixxy[2 * size_t(p)], a0);
      if (a1 != 0.f) atomicAdd(&d_pixxy[2 * size_t(p) + 1], a1);
#pragma unroll
      for (int c = 0; c < 3; ++c) {
        const float a = acc[2 + c][threadIdx.x];
        if (a != 0.f) atomicAdd(&d_conic[3 * size_t(p) + c], a);
      }
      const float ao = acc[5][threadIdx.x];
      if (ao != 0.f) atomicAdd(&d_op[p], ao);
#pragma unroll
      for (int c = 0; c < 3; ++c) {
        const float a = acc[6 + c][threadIdx.x];
        if (a != 0.f) atomicAdd(&d_rgb[3 * size_t(p) + c], a);
      }
    }
    __syncthreads();
    hi = lo;
  }
}

// ----------------------------------------------------------- dispatch -----
template <typename Loader>
int launch_fwd(hipStream_t s, uint32_t w, uint32_t h, uint32_t ts, uint32_t threads,
               const Loader& ld, const uint32_t* tile2idx, const uint32_t* idx2pnt, float* rgb,
               float* alpha, uint32_t* nwalk, float* tlast) {
  uint32_t tw, th;
  tile_grid(w, h, ts, &tw, &th);
  const uint32_t nt = tw * th;
  if (nt == 0) return DSPLAT_OK;
#define DSPLAT_FWD(TS_, NT_)                                                                   \
  gauss_raster_fwd_kernel<TS_, NT_, Loader><<<nt, NT_, 0, s>>>(w, h, tw, ld, tile2idx, idx2pnt, \
                                                               rgb, alpha, nwalk, tlast)
  if (ts == 8) {
    DSPLAT_FWD(8, 64);
  } else if (ts == 16) {
    if (threads == 64) DSPLAT_FWD(16, 64);
    else if (threads == 128) DSPLAT_FWD(16, 128);
    else DSPLAT_FWD(16, 256);
  } else if (ts == 32) {
    if (threads == 1024) DSPLAT_FWD(32, 1024);
    else if (threads == 512) DSPLAT_FWD(32, 512);
    else DSPLAT_FWD(32, 256);
  } else {
    return dsplat_set_error(DSPLAT_ERR_ARG, "tile_size must be 8, 16 or 32");
  }
#undef DSPLAT_FWD
  DSPLAT_HIP(hipGetLastError());
  return DSPLAT_OK;
}

int launch_bwd(hipStream_t s, uint32_t w, uint32_t h, uint32_t ts, uint32_t threads,
               const float4* splat2, const uint32_t* tile2idx, const uint32_t* idx2pnt,
               const uint32_t* nwalk, const float* tlast, const float* d_img,
               const float* d_alpha, float* d_pixxy, float* d_conic, float* d_op, float* d_rgb) {
  uint32_t tw, th;
  tile_grid(w, h, ts, &tw, &th);
  const uint32_t nt = tw * th;
  if (nt == 0) return DSPLAT_OK;
#define DSPLAT_BWD(TS_, NT_)                                                            \
  gauss_raster_bwd_kernel<TS_, NT_><<<nt, NT_, 0, s>>>(w, h, tw, splat2, tile2idx, idx2pnt, \
                                                       nwalk, tlast, d_img, d_alpha, d_pixxy, \
                                                       d_conic, d_op, d_rgb)
  if (ts == 8) {
    DSPLAT_BWD(8, 64);
  } else if (ts == 16) {
    if (threads == 64) DSPLAT_BWD(16, 64);
    else if (threads == 128) DSPLAT_BWD(16, 128);
    else DSPLAT_BWD(16, 256);
  } else if (ts == 32) {
    DSPLAT_BWD(32, 256);
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
                                          float* pix2tlast) {
  if (!pix2rgb || !tile2idx || !pnt2splat2)
    return dsplat_set_error(DSPLAT_ERR_ARG, "rasterize_fwd: null buffer");
  if (reinterpret_cast<uintptr_t>(pnt2splat2) & 15u)
    return dsplat_set_error(DSPLAT_ERR_ARG, "rasterize_fwd: splat2 must be 16-byte aligned");
  PackedLoader ld{reinterpret_cast<const float4*>(pnt2splat2)};
  return launch_fwd(as_stream(stream), img_w, img_h, tile_size, threads, ld, tile2idx, idx2pnt,
                    pix2rgb, pix2alpha, pix2nwalk, pix2tlast);
}

extern "C" int dsplat_gauss_rasterize_bwd(void* stream, uint32_t img_w, uint32_t img_h,
                                          uint32_t tile_size, uint32_t threads,
                                          const float* pnt2splat2, const uint32_t* tile2idx,
                                          const uint32_t* idx2pnt, const uint32_t* pix2nwalk,
                                          const float* pix2tlast, const float* d_pix2rgb,
                                          const float* d_pix2alpha, float* d_pnt2pixxy,
                                          float* d_pnt2conic, float* d_pnt2opacity,
                                          float* d_pnt2rgb) {
  if (!pnt2splat2 || !tile2idx || !pix2nwalk || !pix2tlast || !d_pix2rgb || !d_pnt2pixxy ||
      !d_pnt2conic || !d_pnt2opacity || !d_pnt2rgb)
    return dsplat_set_error(DSPLAT_ERR_ARG, "rasterize_bwd: null buffer");
  if (reinterpret_cast<uintptr_t>(pnt2splat2) & 15u)
    return dsplat_set_error(DSPLAT_ERR_ARG, "rasterize_bwd: splat2 must be 16-byte aligned");
  return launch_bwd(as_stream(stream), img_w, img_h, tile_size, threads,
                    reinterpret_cast<const float4*>(pnt2splat2), tile2idx, idx2pnt, pix2nwalk,
                    pix2tlast, d_pix2rgb, d_pix2alpha, d_pnt2pixxy, d_pnt2conic, d_pnt2opacity,
                    d_pnt2rgb);
}
