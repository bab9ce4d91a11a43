// Device math of the Gaussian projection and its adjoint, shared by the
// single-view and the batched multi-view kernels.
//
// Forward semantics: splat_gauss.cu:10-53 (kernel) and pnt2xyz.rs:51-96 (CPU
// twin) of the reference. The del-geo helpers it calls are restated from
// their mathematical definitions:
//   transform_homogeneous  y = M (x,1), q = y[0..3) / y3
//   jacobian_transform     J_ij = (M_ij - q_i M_3j) / w
//   transform_ndc2pix      A = [W/2 0 W/2; 0 -H/2 H/2] (2x3; its third column
//                          multiplies the ndc-z row of J, as in the reference)
//   projected_spd_mat3     cov = (A J R) diag(sigma^2) (A J R)^T
//   safe_inverse_...       plain 2x2 inverse when lambda_min >= 1e-5, else
//                          eigenvalues clamped to 1e-5
//   aabb2 * 3 + translate  pix +- 3 (sqrt(c/d), sqrt(a/d)), d = ac - b^2 of
//                          the conic [a b c]
// Floating-point contraction is disabled in the forward so that every
// product and sum is rounded in the written order; the result is then
// bit-identical to a CPU evaluation of the same expression order.
#pragma once
#include "dsplat_common.hpp"

namespace dsplat {

struct ProjState {
  float pix[2], ndcz, w, q[3];
  float jac[9], wpix[6], rot[9], s2[3], wr[6];
  float cov[3], conic[3], aabb[4];
  bool culled;
};

// kFast: the divisions by w and by the determinant become one reciprocal
// each and products (the adjoint's recomputation, whose bar is a relative
// tolerance); otherwise every operation is rounded as written, the forward's
// bit-exact order.
template <bool kFast = false>
__device__ inline void project_state(const float* __restrict__ m, float x0, float x1, float x2,
                                     float qx, float qy, float qz, float qw, float s0, float s1,
                                     float s2, uint32_t img_w, uint32_t img_h, ProjState& p) {
#pragma clang fp contract(off)
  float y[4];
#pragma unroll
  for (int i = 0; i < 4; ++i) y[i] = m[i] * x0 + m[i + 4] * x1 + m[i + 8] * x2 + m[i + 12];
  p.w = y[3];
  if (kFast) {
    const float iw = 1.f / y[3];
    p.q[0] = y[0] * iw;
    p.q[1] = y[1] * iw;
    p.q[2] = y[2] * iw;
#pragma unroll
    for (int j = 0; j < 3; ++j)
#pragma unroll
      for (int i = 0; i < 3; ++i) p.jac[i + 3 * j] = (m[i + 4 * j] - p.q[i] * m[3 + 4 * j]) * iw;
  } else {
    p.q[0] = y[0] / y[3];
    p.q[1] = y[1] / y[3];
    p.q[2] = y[2] / y[3];
#pragma unroll
    for (int j = 0; j < 3; ++j)
#pragma unroll
      for (int i = 0; i < 3; ++i) p.jac[i + 3 * j] = (m[i + 4 * j] - p.q[i] * m[3 + 4 * j]) / p.w;
  }
  const float a0 = 0.5f * float(img_w), a1 = 0.f, a2 = 0.f, a3 = -0.5f * float(img_h);
  const float a4 = 0.5f * float(img_w), a5 = 0.5f * float(img_h);
  p.pix[0] = a0 * p.q[0] + a2 * p.q[1] + a4;
  p.pix[1] = a1 * p.q[0] + a3 * p.q[1] + a5;
  p.ndcz = p.q[2];
#pragma unroll
  for (int j = 0; j < 3; ++j) {
    p.wpix[0 + 2 * j] = a0 * p.jac[3 * j] + a2 * p.jac[1 + 3 * j] + a4 * p.jac[2 + 3 * j];
    p.wpix[1 + 2 * j] = a1 * p.jac[3 * j] + a3 * p.jac[1 + 3 * j] + a5 * p.jac[2 + 3 * j];
  }
  // Hamilton quaternion (x,y,z,w) -> column-major rotation
  p.rot[0] = 1.f - 2.f * (qy * qy + qz * qz);
  p.rot[1] = 2.f * (qx * qy + qw * qz);
  p.rot[2] = 2.f * (qx * qz - qw * qy);
  p.rot[3] = 2.f * (qx * qy - qw * qz);
  p.rot[4] = 1.f - 2.f * (qx * qx + qz * qz);
  p.rot[5] = 2.f * (qy * qz + qw * qx);
  p.rot[6] = 2.f * (qx * qz + qw * qy);
  p.rot[7] = 2.f * (qy * qz - qw * qx);
  p.rot[8] = 1.f - 2.f * (qx * qx + qy * qy);
  p.s2[0] = s0 * s0;
  p.s2[1] = s1 * s1;
  p.s2[2] = s2 * s2;
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    p.wr[0 + 2 * k] = p.wpix[0] * p.rot[3 * k] + p.wpix[2] * p.rot[1 + 3 * k] + p.wpix[4] * p.rot[2 + 3 * k];
    p.wr[1 + 2 * k] = p.wpix[1] * p.rot[3 * k] + p.wpix[3] * p.rot[1 + 3 * k] + p.wpix[5] * p.rot[2 + 3 * k];
  }
  const float* wr = p.wr;
  p.cov[0] = wr[0] * wr[0] * p.s2[0] + wr[2] * wr[2] * p.s2[1] + wr[4] * wr[4] * p.s2[2];
  p.cov[1] = wr[0] * wr[1] * p.s2[0] + wr[2] * wr[3] * p.s2[1] + wr[4] * wr[5] * p.s2[2];
  p.cov[2] = wr[1] * wr[1] * p.s2[0] + wr[3] * wr[3] * p.s2[1] + wr[5] * wr[5] * p.s2[2];
  {
    const float a = p.cov[0], b = p.cov[1], c = p.cov[2];
    const float h = (a - c) * 0.5f;
    const float r = sqrtf(h * h + b * b);
    if ((a + c) * 0.5f - r >= 1.0e-5f) {
      const float det = a * c - b * b;
      if (kFast) {
        const float idet = 1.f / det;
        p.conic[0] = c * idet;
        p.conic[1] = -b * idet;
        p.conic[2] = a * idet;
      } else {
        p.conic[0] = c / det;
        p.conic[1] = -b / det;
        p.conic[2] = a / det;
      }
    } else {
      float l0 = (a + c) * 0.5f + r;
      float l1 = (a + c) * 0.5f - r;
      float ux, uy;
      if (r == 0.f) { ux = 1.f; uy = 0.f; }
      else if (h >= 0.f) { ux = h + r; uy = b; }
      else { ux = b; uy = r - h; }
      const float nrm = sqrtf(ux * ux + uy * uy);
      ux /= nrm;
      uy /= nrm;
      l0 = fmaxf(l0, 1.0e-5f);
      l1 = fmaxf(l1, 1.0e-5f);
      const float i0 = 1.f / l0, i1 = 1.f / l1;
      p.conic[0] = i0 * ux * ux + i1 * uy * uy;
      p.conic[1] = i0 * ux * uy - i1 * ux * uy;
      p.conic[2] = i0 * uy * uy + i1 * ux * ux;
    }
  }
  const float det = p.conic[0] * p.conic[2] - p.conic[1] * p.conic[1];
  const float hx = sqrtf(p.conic[2] / det);
  const float hy = sqrtf(p.conic[0] / det);
  p.aabb[0] = p.pix[0] - 3.f * hx;
  p.aabb[1] = p.pix[1] - 3.f * hy;
  p.aabb[2] = p.pix[0] + 3.f * hx;
  p.aabb[3] = p.pix[1] + 3.f * hy;
  p.culled = !(p.w > 0.f);
  if (p.culled) {
    p.aabb[0] = p.aabb[1] = __builtin_inff();
    p.aabb[2] = p.aabb[3] = -__builtin_inff();
  }
}

// Adjoint of one projection: chain rule from (pixel centre, conic) to
// (mean, scale or log-scale, quaternion). Returns false for a culled
// Gaussian (no gradient).
__device__ inline bool project_bwd_one(const float* __restrict__ m, uint32_t img_w,
                                       uint32_t img_h, const float x3[3], float4 qt,
                                       const float sg[3], bool log_scale, float2 gp,
                                       const float gcon[3], float dx[3], float dsc[3],
                                       float dq[4]) {
  ProjState p;
  project_state<true>(m, x3[0], x3[1], x3[2], qt.x, qt.y, qt.z, qt.w, sg[0], sg[1], sg[2], img_w,
                      img_h, p);
  if (p.culled) return false;
  // d cov = -C G C with G the symmetric matrix of dL/dconic (off-diagonal halved)
  const float ca = p.conic[0], cb = p.conic[1], cc = p.conic[2];
  const float ga = gcon[0], gb = 0.5f * gcon[1], gc = gcon[2];
  const float m00 = ca * ga + cb * gb, m01 = ca * gb + cb * gc;
  const float m10 = cb * ga + cc * gb, m11 = cb * gb + cc * gc;
  const float s00 = -(m00 * ca + m01 * cb);
  const float s01 = -(m00 * cb + m01 * cc);
  const float s11 = -(m10 * cb + m11 * cc);
  float d_wr[6];
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    const float w0 = p.wr[2 * k], w1 = p.wr[1 + 2 * k];
    const float t0 = s00 * w0 + s01 * w1, t1 = s01 * w0 + s11 * w1;
    d_wr[2 * k] = 2.f * t0 * p.s2[k];
    d_wr[1 + 2 * k] = 2.f * t1 * p.s2[k];
    const float ds2 = w0 * t0 + w1 * t1;
    // s2 = sigma^2: d/dsigma = 2 sigma ds2, d/dlog sigma = 2 sigma^2 ds2
    dsc[k] = log_scale ? 2.f * p.s2[k] * ds2 : 2.f * sg[k] * ds2;
  }
  float d_wpix[6], g[9];
#pragma unroll
  for (int r = 0; r < 2; ++r)
#pragma unroll
    for (int j = 0; j < 3; ++j)
      d_wpix[r + 2 * j] = d_wr[r] * p.rot[j] + d_wr[r + 2] * p.rot[j + 3] + d_wr[r + 4] * p.rot[j + 6];
#pragma unroll
  for (int j = 0; j < 3; ++j)
#pragma unroll
    for (int k = 0; k < 3; ++k)
      g[j + 3 * k] = p.wpix[2 * j] * d_wr[2 * k] + p.wpix[1 + 2 * j] * d_wr[1 + 2 * k];
  const float x = qt.x, y = qt.y, z = qt.z, w = qt.w;
  // dR/d(x,y,z,w), column-major entries; the symmetric and antisymmetric
  // off-diagonal pairs of g enter together (the factors 2 and 4 are exact)
  const float p13 = g[1] + g[3], p26 = g[2] + g[6], p57 = g[5] + g[7];
  const float m13 = g[1] - g[3], m62 = g[6] - g[2], m57 = g[5] - g[7];
  dq[0] = 2.f * (p13 * y + p26 * z + m57 * w) - 4.f * x * (g[4] + g[8]);
  dq[1] = 2.f * (p13 * x + m62 * w + p57 * z) - 4.f * y * (g[0] + g[8]);
  dq[2] = 2.f * (p26 * x + p57 * y + m13 * w) - 4.f * z * (g[0] + g[4]);
  dq[3] = 2.f * (m13 * z + m62 * y + m57 * x);
  // Wpix = A J -> dL/dJ = A^T dWpix, A = [a0 0 a4; 0 a3 a5]
  const float a0 = 0.5f * float(img_w), a3 = -0.5f * float(img_h);
  const float a4 = 0.5f * float(img_w), a5 = 0.5f * float(img_h);
  float d_jac[9];
#pragma unroll
  for (int j = 0; j < 3; ++j) {
    d_jac[0 + 3 * j] = a0 * d_wpix[2 * j];
    d_jac[1 + 3 * j] = a3 * d_wpix[1 + 2 * j];
    d_jac[2 + 3 * j] = a4 * d_wpix[2 * j] + a5 * d_wpix[1 + 2 * j];
  }
  const float d_q0 = a0 * gp.x, d_q1 = a3 * gp.y;
  const float inv_w = 1.f / p.w;
  // dJ_rj/dx_k = -(J_rk M_3j + J_rj M_3k) / w, so
  // sum_rj dL/dJ_rj dJ_rj/dx_k = -(sum_r J_rk u_r + M_3k t) / w with
  // u_r = sum_j dL/dJ_rj M_3j and t = sum_rj dL/dJ_rj J_rj
  float u[3], t = 0.f;
#pragma unroll
  for (int r = 0; r < 3; ++r) {
    u[r] = d_jac[r] * m[3] + d_jac[r + 3] * m[7] + d_jac[r + 6] * m[11];
#pragma unroll
    for (int j = 0; j < 3; ++j) t += d_jac[r + 3 * j] * p.jac[r + 3 * j];
  }
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    const float sj = p.jac[3 * k] * u[0] + p.jac[1 + 3 * k] * u[1] + p.jac[2 + 3 * k] * u[2] +
                     m[3 + 4 * k] * t;
    dx[k] = d_q0 * p.jac[3 * k] + d_q1 * p.jac[1 + 3 * k] - sj * inv_w;
  }
  return true;
}

}  // namespace dsplat
