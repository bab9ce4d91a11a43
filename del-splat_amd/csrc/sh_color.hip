// View-dependent colour from spherical harmonics, and its adjoint
// (single view). No counterpart in the reference's kernels: its 3DGS PLY
// reader returns the 45 higher-order coefficients (io_ply.rs:196) and renders
// with the DC colour only. Basis, layout and staging: sh_core.hpp.
//
// The kernels are HBM-bound: 12 K + 24 bytes per Gaussian forward,
// 24 K + 36 (stored gradients) backward.
#include "sh_core.hpp"

namespace dsplat {
namespace {

using namespace sh;

constexpr int kShThreads = 256;
constexpr uint32_t kShWaves = kShThreads / 64;

// Both kernels run a grid of resident workgroups (sh_grid); each wave walks
// tiles of 64 Gaussians t, t + W, t + 2W, ... (W = waves in the grid). The
// 16-byte loads of the next tile's coefficient records are issued before
// the current tile is computed and written, so a wave keeps a tile's worth
// of loads in flight while it works instead of exposing the HBM latency
// once per tile.

// Loads of tile tt: the lane's mean (and colour gradient, when d_rgb is
// given) if the lane is live, and the tile's coefficient records if the
// tile is full (a tail tile is staged by single floats).
template <int R>
__device__ inline void fetch_tile(uint32_t n, uint32_t tt, int lane, const float* __restrict__ xyz,
                                  const float* __restrict__ d_rgb, const float* __restrict__ sh,
                                  float* p, float* q, word4* v) {
  const uint32_t g = tt * 64u + uint32_t(lane);
  if (g < n) {
#pragma unroll
    for (int a = 0; a < 3; ++a) {
      p[a] = xyz[3 * g + a];
      if (d_rgb) q[a] = d_rgb[3 * g + a];
    }
  }
  if (tt * 64u + 64u <= n) tile_load<R>(sh, tt * 64u, lane, v);
}

template <int K, int DEG>
__global__ void __launch_bounds__(kShThreads)
sh_to_rgb_kernel(uint32_t n, const float* __restrict__ cam, const float* __restrict__ xyz,
                 const float* __restrict__ sh, float* __restrict__ rgb) {
  constexpr int R = 3 * K, RP = record_stride(R), NK = (DEG + 1) * (DEG + 1);
  __shared__ __align__(16) float lds[kShWaves][64 * RP];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const uint32_t ntile = div_up(n, 64), step = gridDim.x * kShWaves;
  uint32_t t = blockIdx.x * kShWaves + uint32_t(wave);
  if (t >= ntile) return;  // wave-uniform
  float* L = lds[wave];
  word4 v[tile_words<R>()];
  float p[3];
  fetch_tile<R>(n, t, lane, xyz, nullptr, sh, p, nullptr, v);
  for (;;) {
    const uint32_t g0 = t * 64u, cnt = min(64u, n - g0);
    if (cnt == 64) tile_to_lds<R, RP>(v, L, lane);
    else stage_in<R, RP>(sh, g0, cnt, L, lane);
    const float px = p[0], py = p[1], pz = p[2];
    const uint32_t tn = t + step;
    if (tn < ntile) fetch_tile<R>(n, tn, lane, xyz, nullptr, sh, p, nullptr, v);
    wave_lds_sync();
    float c[R];
    load_record<R, RP>(L, lane, c);
    const uint32_t g = g0 + uint32_t(lane);
    if (g < n) {
      float d[3], inv, b[16], raw[3];
      direction(cam, px, py, pz, d, &inv);
      sh_basis<DEG>(d[0], d[1], d[2], b);
      colour<NK>(c, b, raw);
#pragma unroll
      for (int ch = 0; ch < 3; ++ch) rgb[3 * g + ch] = raw[ch] > 0.f ? raw[ch] : 0.f;
    }
    if (tn >= ntile) break;  // wave-uniform
    // every lane has read its record before the next tile lands in LDS
    wave_lds_sync();
    t = tn;
  }
}

template <int K, int DEG, bool kAdd>
__global__ void __launch_bounds__(kShThreads)
sh_to_rgb_bwd_kernel(uint32_t n, const float* __restrict__ cam, const float* __restrict__ xyz,
                     const float* __restrict__ sh, const float* __restrict__ d_rgb,
                     float* __restrict__ d_sh, float* __restrict__ d_xyz) {
  constexpr int R = 3 * K, RP = record_stride(R), NK = (DEG + 1) * (DEG + 1);
  __shared__ __align__(16) float lds[kShWaves][64 * RP];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const uint32_t ntile = div_up(n, 64), step = gridDim.x * kShWaves;
  uint32_t t = blockIdx.x * kShWaves + uint32_t(wave);
  if (t >= ntile) return;  // wave-uniform
  float* L = lds[wave];
  word4 v[tile_words<R>()];
  float p[3], q[3];
  fetch_tile<R>(n, t, lane, xyz, d_rgb, sh, p, q, v);
  for (;;) {
    const uint32_t g0 = t * 64u, cnt = min(64u, n - g0);
    if (cnt == 64) tile_to_lds<R, RP>(v, L, lane);
    else stage_in<R, RP>(sh, g0, cnt, L, lane);
    const float px = p[0], py = p[1], pz = p[2], q0 = q[0], q1 = q[1], q2 = q[2];
    const uint32_t tn = t + step;
    if (tn < ntile) fetch_tile<R>(n, tn, lane, xyz, d_rgb, sh, p, q, v);
    wave_lds_sync();
    const uint32_t g = g0 + uint32_t(lane);
    const bool live = g < n;
    float c[R];
    load_record<R, RP>(L, lane, c);
    float gc[3] = {0.f, 0.f, 0.f}, b[16];
    if (live) {
      const float qd[3] = {q0, q1, q2};
      float d[3], inv, raw[3], ax[3] = {0.f, 0.f, 0.f};
      direction(cam, px, py, pz, d, &inv);
      sh_basis<DEG>(d[0], d[1], d[2], b);
      colour<NK>(c, b, raw);
#pragma unroll
      for (int ch = 0; ch < 3; ++ch) gc[ch] = raw[ch] < 0.f ? 0.f : qd[ch];
      direction_adjoint<DEG>(c, gc, d, inv, ax);
#pragma unroll
      for (int a = 0; a < 3; ++a) d_xyz[3 * g + a] = kAdd ? d_xyz[3 * g + a] + ax[a] : ax[a];
    }
    // d_sh = Y_k g_c (0 for k >= NK and for dead lanes), through the same
    // LDS slice (every lane has read its record before any lane overwrites).
    wave_lds_sync();
    float o[RP];
#pragma unroll
    for (int k = 0; k < K; ++k)
#pragma unroll
      for (int ch = 0; ch < 3; ++ch)
        o[3 * k + ch] = (k < NK && live) ? b[k < NK ? k : 0] * gc[ch] : 0.f;
#pragma unroll
    for (int i = R; i < RP; ++i) o[i] = 0.f;
#pragma unroll
    for (int i = 0; i < (R + 3) / 4; ++i)
      *reinterpret_cast<float4*>(L + lane * RP + 4 * i) =
          make_float4(o[4 * i], o[4 * i + 1], o[4 * i + 2], o[4 * i + 3]);
    wave_lds_sync();
    stage_out<R, RP, kAdd>(d_sh, g0, cnt, L, lane);
    if (tn >= ntile) break;  // wave-uniform
    // every lane has read the gradients out of LDS before the next tile lands
    wave_lds_sync();
    t = tn;
  }
}

// Workgroups of one SH launch: as many as are resident at once on the
// device (occupancy of the kernel times CUs, queried once per kernel on the
// device current at its first launch), at most one wave per tile.
template <auto Kernel>
uint32_t sh_grid(uint32_t n) {
  static const uint32_t resident = [] {
    int dev = 0, cus = 0, per_cu = 0;
    if (hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess ||
        cus <= 0)
      cus = 256;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(
            &per_cu, reinterpret_cast<const void*>(Kernel), kShThreads, 0) !=
            hipSuccess ||
        per_cu <= 0)
      per_cu = 1;
    return uint32_t(cus) * uint32_t(per_cu);
  }();
  return min(div_up(div_up(n, 64), kShWaves), resident);
}

template <int K>
int launch_fwd(hipStream_t s, uint32_t n, uint32_t deg, const float* cam, const float* xyz,
               const float* sh, float* rgb) {
#define DSPLAT_SHF(D_) \
  sh_to_rgb_kernel<K, D_><<<sh_grid<sh_to_rgb_kernel<K, D_>>(n), kShThreads, 0, s>>>(n, cam, xyz, sh, rgb)
  switch (deg) {
    case 0: DSPLAT_SHF(0); break;
    case 1: if constexpr (K >= 4) { DSPLAT_SHF(1); break; }
            [[fallthrough]];
    case 2: if constexpr (K >= 9) { DSPLAT_SHF(2); break; }
            [[fallthrough]];
    case 3: if constexpr (K >= 16) { DSPLAT_SHF(3); break; }
            [[fallthrough]];
    default: return dsplat_set_error(DSPLAT_ERR_ARG, "sh_to_rgb: (degree+1)^2 exceeds num_coef");
  }
#undef DSPLAT_SHF
  return DSPLAT_OK;
}

template <int K, bool kAdd>
int launch_bwd(hipStream_t s, uint32_t n, uint32_t deg, const float* cam, const float* xyz,
               const float* sh, const float* d_rgb, float* d_sh, float* d_xyz) {
#define DSPLAT_SHB(D_)                                                                    \
  sh_to_rgb_bwd_kernel<K, D_, kAdd><<<sh_grid<sh_to_rgb_bwd_kernel<K, D_, kAdd>>(n), kShThreads, \
                                      0, s>>>(n, cam, xyz, sh, d_rgb, d_sh, d_xyz)
  switch (deg) {
    case 0: DSPLAT_SHB(0); break;
    case 1: if constexpr (K >= 4) { DSPLAT_SHB(1); break; }
            [[fallthrough]];
    case 2: if constexpr (K >= 9) { DSPLAT_SHB(2); break; }
            [[fallthrough]];
    case 3: if constexpr (K >= 16) { DSPLAT_SHB(3); break; }
            [[fallthrough]];
    default: return dsplat_set_error(DSPLAT_ERR_ARG, "sh_to_rgb_bwd: (degree+1)^2 exceeds num_coef");
  }
#undef DSPLAT_SHB
  return DSPLAT_OK;
}

int check_args(const char* who, uint32_t deg, uint32_t num_coef, const float* sh) {
  if (num_coef != 1 && num_coef != 4 && num_coef != 9 && num_coef != 16)
    return dsplat_set_error(DSPLAT_ERR_ARG, who);
  if (deg > 3 || (deg + 1) * (deg + 1) > num_coef) return dsplat_set_error(DSPLAT_ERR_ARG, who);
  if (reinterpret_cast<uintptr_t>(sh) & 15u) return dsplat_set_error(DSPLAT_ERR_ARG, who);
  return DSPLAT_OK;
}

}  // namespace
}  // namespace dsplat

using namespace dsplat;

extern "C" int dsplat_sh_to_rgb(void* stream, uint32_t num_pnt, uint32_t degree,
                                uint32_t num_coef, const float* cam_pos, const float* pnt2xyz,
                                const float* pnt2sh, float* pnt2rgb) {
  if (num_pnt == 0) return DSPLAT_OK;
  if (!cam_pos || !pnt2xyz || !pnt2sh || !pnt2rgb)
    return dsplat_set_error(DSPLAT_ERR_ARG, "sh_to_rgb: null buffer");
  if (check_args("sh_to_rgb: num_coef must be 1, 4, 9 or 16, degree <= 3 with (degree+1)^2 <= "
                 "num_coef, and pnt2sh 16-byte aligned", degree, num_coef, pnt2sh))
    return DSPLAT_ERR_ARG;
  hipStream_t s = as_stream(stream);
  int rc;
  switch (num_coef) {
    case 1: rc = launch_fwd<1>(s, num_pnt, degree, cam_pos, pnt2xyz, pnt2sh, pnt2rgb); break;
    case 4: rc = launch_fwd<4>(s, num_pnt, degree, cam_pos, pnt2xyz, pnt2sh, pnt2rgb); break;
    case 9: rc = launch_fwd<9>(s, num_pnt, degree, cam_pos, pnt2xyz, pnt2sh, pnt2rgb); break;
    default: rc = launch_fwd<16>(s, num_pnt, degree, cam_pos, pnt2xyz, pnt2sh, pnt2rgb); break;
  }
  if (rc) return rc;
  DSPLAT_HIP(hipGetLastError());
  return DSPLAT_OK;
}

extern "C" int dsplat_sh_to_rgb_bwd(void* stream, uint32_t num_pnt, uint32_t degree,
                                    uint32_t num_coef, const float* cam_pos,
                                    const float* pnt2xyz, const float* pnt2sh,
                                    const float* d_pnt2rgb, uint32_t flags, float* d_pnt2sh,
                                    float* d_pnt2xyz) {
  if (num_pnt == 0) return DSPLAT_OK;
  if (!cam_pos || !pnt2xyz || !pnt2sh || !d_pnt2rgb || !d_pnt2sh || !d_pnt2xyz)
    return dsplat_set_error(DSPLAT_ERR_ARG, "sh_to_rgb_bwd: null buffer");
  if (check_args("sh_to_rgb_bwd: num_coef must be 1, 4, 9 or 16, degree <= 3 with (degree+1)^2 "
                 "<= num_coef, and pnt2sh 16-byte aligned", degree, num_coef, pnt2sh) ||
      (reinterpret_cast<uintptr_t>(d_pnt2sh) & 15u))
    return dsplat_set_error(DSPLAT_ERR_ARG,
                            "sh_to_rgb_bwd: bad num_coef/degree or misaligned sh buffers");
  hipStream_t s = as_stream(stream);
  const bool add = !(flags & DSPLAT_GRAD_OVERWRITE);
  int rc;
#define DSPLAT_SHL(K_)                                                                        \
  rc = add ? launch_bwd<K_, true>(s, num_pnt, degree, cam_pos, pnt2xyz, pnt2sh, d_pnt2rgb,   \
                                  d_pnt2sh, d_pnt2xyz)                                       \
           : launch_bwd<K_, false>(s, num_pnt, degree, cam_pos, pnt2xyz, pnt2sh, d_pnt2rgb,  \
                                   d_pnt2sh, d_pnt2xyz)
  switch (num_coef) {
    case 1: DSPLAT_SHL(1); break;
    case 4: DSPLAT_SHL(4); break;
    case 9: DSPLAT_SHL(9); break;
    default: DSPLAT_SHL(16); break;
  }
#undef DSPLAT_SHL
  if (rc) return rc;
  DSPLAT_HIP(hipGetLastError());
  return DSPLAT_OK;
}
