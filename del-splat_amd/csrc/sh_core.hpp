// Spherical-harmonic colour, device side: the basis, its gradient, and the
// staging of per-Gaussian coefficient records between HBM and LDS. Shared by
// the single-view kernels (sh_color.hip) and the batched plan
// (multiview.hip).
//
// Basis: the real spherical harmonics of 3D Gaussian splatting's colour
// model, degree <= 3, coefficient k = l^2 + l + m:
//   Y_lm = sqrt(2) Re Y_l^|m| (m > 0), sqrt(2) Im Y_l^|m| (m < 0), Y_l^0
// of the complex harmonics with the Condon-Shortley phase, at the unit
// direction d from the camera centre to the mean. Colour channel c is
// max(sum_k Y_k(d) sh[k][c] + 1/2, 0).
//
// Layout: sh is [N][K][3] floats (K = 1, 4, 9 or 16, channel inner), one
// contiguous record of R = 3K floats per Gaussian. A wave moves the records
// of its 64 Gaussians with coalesced 16-byte accesses; in LDS the records sit
// at a stride RP (R rounded up to a multiple of 4, plus 4 when RP/4 is even)
// so that the 16-byte reads of one record per lane fall on distinct banks.
#pragma once

#include "dsplat_common.hpp"

namespace dsplat {
namespace sh {

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

// Number of 16-byte words one lane moves for a full tile of 64 records.
template <int R>
constexpr int tile_words() { return (16 * R + 63) / 64; }

// A 16-byte word in registers. A clang vector rather than HIP's float4
// struct: struct copies become private-memory memcpys that stay in scratch
// when the words are carried across loop iterations.
typedef float word4 __attribute__((ext_vector_type(4)));

// Issues the 16-byte loads of a full tile of 64 records (R floats each,
// contiguous from `src` at record g0) into v, coalesced across the wave; the
// caller waits for them where it uses v.
template <int R>
__device__ inline void tile_load(const float* __restrict__ src, uint32_t g0, int lane,
                                 word4* v) {
  const word4* b4 = reinterpret_cast<const word4*>(src + size_t(g0) * R);
#pragma unroll
  for (int i = 0; i < tile_words<R>(); ++i) {
    const int f4 = lane + 64 * i;
    if (f4 < 16 * R) v[i] = b4[f4];
  }
}

// Writes a tile loaded by tile_load into the wave's LDS slice at stride RP.
template <int R, int RP>
__device__ inline void tile_to_lds(const word4* v, float* __restrict__ lds, int lane) {
#pragma unroll
  for (int i = 0; i < tile_words<R>(); ++i) {
    const int f4 = lane + 64 * i;
    if (f4 < 16 * R) {
      const int f = 4 * f4;
      if (R % 4 == 0) {
        *reinterpret_cast<word4*>(lds + (f / R) * RP + f % R) = v[i];
      } else {
        const float e[4] = {v[i].x, v[i].y, v[i].z, v[i].w};
#pragma unroll
        for (int j = 0; j < 4; ++j) lds[((f + j) / R) * RP + (f + j) % R] = e[j];
      }
    }
  }
}

// Copies the calling wave's records [g0, g0 + cnt) (R floats each,
// contiguous from `src`) into the wave's LDS slice at stride RP. Full waves
// move 16-byte words; a tail wave moves single floats so that it never
// reads past the end of the array.
template <int R, int RP>
__device__ inline void stage_in(const float* __restrict__ src, uint32_t g0, uint32_t cnt,
                                float* __restrict__ lds, int lane) {
  if (cnt == 64) {
    word4 v[tile_words<R>()];
    tile_load<R>(src, g0, lane, v);
    tile_to_lds<R, RP>(v, lds, lane);
  } else {
    const float* base = src + size_t(g0) * R;
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

// Unit direction from the camera centre to the mean p, and 1/|p - cam|.
__device__ inline void direction(const float* __restrict__ cam, float px, float py, float pz,
                                 float* d, float* inv_len) {
  const float vx = px - cam[0], vy = py - cam[1], vz = pz - cam[2];
  const float inv = 1.f / sqrtf(vx * vx + vy * vy + vz * vz);
  d[0] = vx * inv;
  d[1] = vy * inv;
  d[2] = vz * inv;
  *inv_len = inv;
}

// Clamped colour of one coefficient record c (R floats, channel inner) for
// the basis values b; raw[ch] keeps the unclamped sum + 1/2.
template <int NK>
__device__ inline void colour(const float* c, const float* b, float* raw) {
#pragma unroll
  for (int ch = 0; ch < 3; ++ch) {
    float acc = 0.f;
#pragma unroll
    for (int k = 0; k < NK; ++k) acc += b[k] * c[3 * k + ch];
    raw[ch] = acc + 0.5f;
  }
}

// Direction adjoint of one view: (I - d d^T) / |v| * sum_k w_k grad Y_k with
// w_k = sum_ch c[k][ch] gc[ch], added to ax.
template <int DEG>
__device__ inline void direction_adjoint(const float* c, const float* gc, const float* d,
                                         float inv, float* ax) {
  constexpr int NK = (DEG + 1) * (DEG + 1);
  float w[16], dd[3];
#pragma unroll
  for (int k = 0; k < NK; ++k) w[k] = c[3 * k] * gc[0] + c[3 * k + 1] * gc[1] + c[3 * k + 2] * gc[2];
  sh_basis_grad_dot<DEG>(d[0], d[1], d[2], w, dd);
  const float r = d[0] * dd[0] + d[1] * dd[1] + d[2] * dd[2];
#pragma unroll
  for (int a = 0; a < 3; ++a) ax[a] += (dd[a] - d[a] * r) * inv;
}

// Wave-level LDS hand-off: the wave's LDS writes are visible to its reads.
__device__ inline void wave_lds_sync() {
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

}  // namespace sh
}  // namespace dsplat
