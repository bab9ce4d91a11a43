// View-dependent colour from spherical harmonics, and its adjoint.
//
// No counterpart in the reference's kernels: its 3DGS PLY reader returns the
// 45 higher-order coefficients (io_ply.rs:196) and renders with the DC
// colour only. The basis is the real spherical harmonics of 3D Gaussian
// splatting's colour model, degree <= 3, coefficient k = l^2 + l + m:
//   Y_lm = sqrt(2) Re Y_l^|m| (m > 0), sqrt(2) Im Y_l^|m| (m < 0), Y_l^0
// of the complex harmonics with the Condon-Shortley phase, evaluated at the
// unit direction d from the camera centre to the mean. Colour channel c is
// max(sum_k Y_k(d) sh[k][c] + 1/2, 0).
//
// Layout: sh is [N][K][3] floats (K = 1, 4, 9 or 16 coefficients per
// channel, channel inner), so a Gaussian's coefficients are one contiguous
// record of R = 3K floats. Each wave stages the 64 records of its Gaussians
// between HBM and LDS with coalesced 16-byte accesses; in LDS the records
// sit at a stride RP (R rounded up to a multiple of 4, plus 4 when RP/4 is
// even) so that the 16-byte reads of one record per lane fall on distinct
// banks. The kernels are HBM-bound: 12 K + 24 bytes per Gaussian forward,
// 24 K + 36 (stored gradients) backward.
#include "dsplat_common.hpp"

namespace dsplat {
namespace {

constexpr float kSh0 = 0.28209479177387814f;   // 1/(2 sqrt(pi))
constexpr float kSh1 = 0.4886025119029199f;    // sqrt(3/(4 pi))
constexpr float kSh2a = 1.0925484305920792f;   // sqrt(15/(4 pi))
constexpr float kSh2b = 0.31539156525252005f;  // sqrt(5/(16 pi))
constexpr float kSh2c = 0.5462742152960396f;   // sqrt(15/(16 pi))
constexpr float kSh3a = 0.5900435899266435f;   // sqrt(35/(32 pi))
constexpr float kSh3b = 2.890611442640554f;    // sqrt(105/(4 pi))
constexpr float kSh3c = 0.4570457994644658f;   // sqrt(21/(32 pi))
constexpr float kSh3d = 0.3731763325901154f;   // sqrt(7/(16 pi))
constexpr float kSh3e = 1.445305721320277f;    // sqrt(105/(16 pi))

template <int DEG>
__device__ inline void sh_basis(float x, float y, float z, float* b) {
  b[0] = kSh0;
  if (DEG < 1) return;
  b[1] = -kSh1 * y;
  b[2] = kSh1 * z;
  b[3] = -kSh1 * x;
  if (DEG < 2) return;
  const float xx = x * x, yy = y * y, zz = z * z;
  b[4] = kSh2a * (x * y);
  b[5] = -kSh2a * (y * z);
  b[6] = kSh2b * (2.f * zz - xx - yy);
  b[7] = -kSh2a * (x * z);
  b[8] = kSh2c * (xx - yy);
  if (DEG < 3) return;
  b[9] = -kSh3a * (y * (3.f * xx - yy));
  b[10] = kSh3b * (x * y * z);
  b[11] = -kSh3c * (y * (4.f * zz - xx - yy));
  b[12] = kSh3d * (z * (2.f * zz - 3.f * xx - 3.f * yy));
  b[13] = -kSh3c * (x * (4.f * zz - xx - yy));
  b[14] = kSh3e * (z * (xx - yy));
  b[15] = -kSh3a * (x * (xx - 3.f * yy));
}

// sum_k w_k grad Y_k (the polynomials' gradients in R^3), w = per-coefficient
// weights sum_c sh[k][c] g_c.
template <int DEG>
__device__ inline void sh_basis_grad_dot(float x, float y, float z, const float* w, float* dd) {
  dd[0] = dd[1] = dd[2] = 0.f;
  if (DEG < 1) return;
  dd[0] -= kSh1 * w[3];
  dd[1] -= kSh1 * w[1];
  dd[2] += kSh1 * w[2];
  if (DEG < 2) return;
  const float xx = x * x, yy = y * y, zz = z * z;
  dd[0] += kSh2a * (w[4] * y - w[7] * z) + kSh2b * w[6] * (-2.f * x) + kSh2c * w[8] * (2.f * x);
  dd[1] += kSh2a * (w[4] * x - w[5] * z) + kSh2b * w[6] * (-2.f * y) - kSh2c * w[8] * (2.f * y);
  dd[2] += kSh2a * (-w[5] * y - w[7] * x) + kSh2b * w[6] * (4.f * z);
  if (DEG < 3) return;
  dd[0] += -kSh3a * w[9] * (6.f * x * y) + kSh3b * w[10] * (y * z) +
           kSh3c * w[11] * (2.f * x * y) + kSh3d * w[12] * (-6.f * x * z) -
           kSh3c * w[13] * (4.f * zz - 3.f * xx - yy) + kSh3e * w[14] * (2.f * x * z) -
           kSh3a * w[15] * (3.f * xx - 3.f * yy);
  dd[1] += -kSh3a * w[9] * (3.f * xx - 3.f * yy) + kSh3b * w[10] * (x * z) -
           kSh3c * w[11] * (4.f * zz - xx - 3.f * yy) + kSh3d * w[12] * (-6.f * y * z) +
           kSh3c * w[13] * (2.f * x * y) - kSh3e * w[14] * (2.f * y * z) +
           kSh3a * w[15] * (6.f * x * y);
  dd[2] += kSh3b * w[10] * (x * y) - kSh3c * w[11] * (8.f * y * z) +
           kSh3d * w[12] * (6.f * zz - 3.f * xx - 3.f * yy) - kSh3c * w[13] * (8.f * x * z) +
           kSh3e * w[14] * (xx - yy);
}

__host__ __device__ constexpr int record_stride(int r) {
  return ((r + 3) / 4 * 4) / 4 % 2 == 0 ? (r + 3) / 4 * 4 + 4 : (r + 3) / 4 * 4;
}

// Copies the calling wave's records [g0, g0 + cnt) (R floats each,
// contiguous from `src`) into the wave's LDS slice at stride RP. Full waves
// move 16-byte words; a tail wave moves single floats so that it never
// reads past the end of the array.
template <int R, int RP>
__device__ inline void stage_in(const float* __restrict__ src, uint32_t g0, uint32_t cnt,
                                float* __restrict__ lds, int lane) {
  const float* base = src + size_t(g0) * R;
  if (cnt == 64) {
    const float4* b4 = reinterpret_cast<const float4*>(base);
    float4 v[(16 * R + 63) / 64];
#pragma unroll
    for (int i = 0; i < (16 * R + 63) / 64; ++i) {
      const int f4 = lane + 64 * i;
      if (f4 < 16 * R) v[i] = b4[f4];
    }
#pragma unroll
    for (int i = 0; i < (16 * R + 63) / 64; ++i) {
      const int f4 = lane + 64 * i;
      if (f4 < 16 * R) {
        const int f = 4 * f4;
        if (R % 4 == 0) {
          *reinterpret_cast<float4*>(lds + (f / R) * RP + f % R) = v[i];
        } else {
          const float e[4] = {v[i].x, v[i].y, v[i].z, v[i].w};
#pragma unroll
          for (int j = 0; j < 4; ++j) lds[((f + j) / R) * RP + (f + j) % R] = e[j];
        }
      }
    }
  } else {
    const uint32_t total = cnt * R;
    for (uint32_t f = lane; f < total; f += 64) lds[(f / R) * RP + f % R] = base[f];
  }
}

// The inverse of stage_in: the wave's records from LDS to `dst`, stored or
// (kAdd) added to what `dst` holds.
template <int R, int RP, bool kAdd>
__device__ inline void stage_out(float* __restrict__ dst, uint32_t g0, uint32_t cnt,
                                 const float* __restrict__ lds, int lane) {
  float* base = dst + size_t(g0) * R;
  if (cnt == 64) {
    float4* b4 = reinterpret_cast<float4*>(base);
#pragma unroll
    for (int i = 0; i < (16 * R + 63) / 64; ++i) {
      const int f4 = lane + 64 * i;
      if (f4 < 16 * R) {
        const int f = 4 * f4;
        float4 v;
        if (R % 4 == 0) {
          v = *reinterpret_cast<const float4*>(lds + (f / R) * RP + f % R);
        } else {
          v.x = lds[((f + 0) / R) * RP + (f + 0) % R];
          v.y = lds[((f + 1) / R) * RP + (f + 1) % R];
          v.z = lds[((f + 2) / R) * RP + (f + 2) % R];
          v.w = lds[((f + 3) / R) * RP + (f + 3) % R];
        }
        if (kAdd) {
          const float4 o = b4[f4];
          v.x += o.x;
          v.y += o.y;
          v.z += o.z;
          v.w += o.w;
        }
        b4[f4] = v;
      }
    }
  } else {
    const uint32_t total = cnt * R;
    for (uint32_t f = lane; f < total; f += 64) {
      const float v = lds[(f / R) * RP + f % R];
      base[f] = kAdd ? base[f] + v : v;
    }
  }
}

// The calling lane's record from LDS into registers (16-byte reads; the
// words past R in the last read are the stride's padding).
template <int R, int RP>
__device__ inline void load_record(const float* __restrict__ lds, int lane, float* c) {
  float4 w[RP / 4];
#pragma unroll
  for (int i = 0; i < (R + 3) / 4; ++i)
    w[i] = *reinterpret_cast<const float4*>(lds + lane * RP + 4 * i);
#pragma unroll
  for (int i = 0; i < R; ++i) {
    const float4 q = w[i / 4];
    c[i] = (i % 4 == 0) ? q.x : (i % 4 == 1) ? q.y : (i % 4 == 2) ? q.z : q.w;
  }
}

__device__ inline void direction(const float* __restrict__ cam, const float* __restrict__ xyz,
                                 uint32_t g, float* d, float* inv_len) {
  const float vx = xyz[3 * g] - cam[0], vy = xyz[3 * g + 1] - cam[1],
              vz = xyz[3 * g + 2] - cam[2];
  const float inv = 1.f / sqrtf(vx * vx + vy * vy + vz * vz);
  d[0] = vx * inv;
  d[1] = vy * inv;
  d[2] = vz * inv;
  *inv_len = inv;
}

constexpr int kShThreads = 256;

template <int K, int DEG>
__global__ void __launch_bounds__(kShThreads)
sh_to_rgb_kernel(uint32_t n, const float* __restrict__ cam, const float* __restrict__ xyz,
                 const float* __restrict__ sh, float* __restrict__ rgb) {
  constexpr int R = 3 * K, RP = record_stride(R), NK = (DEG + 1) * (DEG + 1);
  __shared__ __align__(16) float lds[kShThreads / 64][64 * RP];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const uint32_t g0 = (blockIdx.x * kShThreads) + 64u * uint32_t(wave);
  if (g0 >= n) return;  // wave-uniform
  const uint32_t cnt = min(64u, n - g0);
  stage_in<R, RP>(sh, g0, cnt, lds[wave], lane);
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
  const uint32_t g = g0 + uint32_t(lane);
  if (g >= n) return;
  float c[R];
  load_record<R, RP>(lds[wave], lane, c);
  float d[3], inv, b[16];
  direction(cam, xyz, g, d, &inv);
  sh_basis<DEG>(d[0], d[1], d[2], b);
#pragma unroll
  for (int ch = 0; ch < 3; ++ch) {
    float acc = 0.f;
#pragma unroll
    for (int k = 0; k < NK; ++k) acc += b[k] * c[3 * k + ch];
    acc += 0.5f;
    rgb[3 * g + ch] = acc > 0.f ? acc : 0.f;
  }
}

template <int K, int DEG, bool kAdd>
__global__ void __launch_bounds__(kShThreads)
sh_to_rgb_bwd_kernel(uint32_t n, const float* __restrict__ cam, const float* __restrict__ xyz,
                     const float* __restrict__ sh, const float* __restrict__ d_rgb,
                     float* __restrict__ d_sh, float* __restrict__ d_xyz) {
  constexpr int R = 3 * K, RP = record_stride(R), NK = (DEG + 1) * (DEG + 1);
  __shared__ __align__(16) float lds[kShThreads / 64][64 * RP];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const uint32_t g0 = (blockIdx.x * kShThreads) + 64u * uint32_t(wave);
  if (g0 >= n) return;  // wave-uniform
  const uint32_t cnt = min(64u, n - g0);
  float* L = lds[wave];
  stage_in<R, RP>(sh, g0, cnt, L, lane);
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
  const uint32_t g = g0 + uint32_t(lane);
  const bool live = g < n;
  float c[R];
  load_record<R, RP>(L, lane, c);
  float gc[3] = {0.f, 0.f, 0.f}, b[16];
  if (live) {
    float d[3], inv;
    direction(cam, xyz, g, d, &inv);
    sh_basis<DEG>(d[0], d[1], d[2], b);
#pragma unroll
    for (int ch = 0; ch < 3; ++ch) {
      float acc = 0.f;
#pragma unroll
      for (int k = 0; k < NK; ++k) acc += b[k] * c[3 * k + ch];
      acc += 0.5f;
      gc[ch] = acc < 0.f ? 0.f : d_rgb[3 * g + ch];
    }
    float w[16], dd[3];
#pragma unroll
    for (int k = 0; k < NK; ++k) w[k] = c[3 * k] * gc[0] + c[3 * k + 1] * gc[1] + c[3 * k + 2] * gc[2];
    sh_basis_grad_dot<DEG>(d[0], d[1], d[2], w, dd);
    const float r = d[0] * dd[0] + d[1] * dd[1] + d[2] * dd[2];
#pragma unroll
    for (int a = 0; a < 3; ++a) {
      const float v = (dd[a] - d[a] * r) * inv;
      d_xyz[3 * g + a] = kAdd ? d_xyz[3 * g + a] + v : v;
    }
  }
  // d_sh = Y_k g_c (0 for k >= NK and for dead lanes), through the same
  // LDS slice (every lane has read its record before any lane overwrites).
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
  float o[RP];
#pragma unroll
  for (int k = 0; k < K; ++k)
#pragma unroll
    for (int ch = 0; ch < 3; ++ch) o[3 * k + ch] = (k < NK && live) ? b[k < NK ? k : 0] * gc[ch] : 0.f;
#pragma unroll
  for (int i = R; i < RP; ++i) o[i] = 0.f;
#pragma unroll
  for (int i = 0; i < (R + 3) / 4; ++i)
    *reinterpret_cast<float4*>(L + lane * RP + 4 * i) =
        make_float4(o[4 * i], o[4 * i + 1], o[4 * i + 2], o[4 * i + 3]);
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
  stage_out<R, RP, kAdd>(d_sh, g0, cnt, L, lane);
}

template <int K>
int launch_fwd(hipStream_t s, uint32_t n, uint32_t deg, const float* cam, const float* xyz,
               const float* sh, float* rgb) {
  const uint32_t nb = div_up(n, kShThreads);
  switch (deg) {
    case 0: sh_to_rgb_kernel<K, 0><<<nb, kShThreads, 0, s>>>(n, cam, xyz, sh, rgb); break;
    case 1: if constexpr (K >= 4) { sh_to_rgb_kernel<K, 1><<<nb, kShThreads, 0, s>>>(n, cam, xyz, sh, rgb); break; }
            [[fallthrough]];
    case 2: if constexpr (K >= 9) { sh_to_rgb_kernel<K, 2><<<nb, kShThreads, 0, s>>>(n, cam, xyz, sh, rgb); break; }
            [[fallthrough]];
    case 3: if constexpr (K >= 16) { sh_to_rgb_kernel<K, 3><<<nb, kShThreads, 0, s>>>(n, cam, xyz, sh, rgb); break; }
            [[fallthrough]];
    default: return dsplat_set_error(DSPLAT_ERR_ARG, "sh_to_rgb: (degree+1)^2 exceeds num_coef");
  }
  return DSPLAT_OK;
}

template <int K, bool kAdd>
int launch_bwd(hipStream_t s, uint32_t n, uint32_t deg, const float* cam, const float* xyz,
               const float* sh, const float* d_rgb, float* d_sh, float* d_xyz) {
  const uint32_t nb = div_up(n, kShThreads);
#define DSPLAT_SHB(D_) \
  sh_to_rgb_bwd_kernel<K, D_, kAdd><<<nb, kShThreads, 0, s>>>(n, cam, xyz, sh, d_rgb, d_sh, d_xyz)
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
