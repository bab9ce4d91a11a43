// dsplat oracle -- TEST INFRASTRUCTURE ONLY.
//
// A plain C++ CPU restatement of the del-splat hot path, used as the parity
// checker for the HIP library and as the CPU baseline of the benchmark. Only
// tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may load it;
// nothing in the product library links or calls it.
//
// Every function is templated on the real type T: the float instantiation is
// the parity oracle, the double instantiation is the twin used for central
// finite differences of the adjoint.
//
// Compiled with -ffp-contract=off so that every float product and sum is
// rounded separately, in the written order.
//
// Reference semantics restated here (paths relative to the del-splat repo):
//   projection          del-splat-cuda-kernels/cuda/splat_gauss.cu:10-53,
//                       del-splat-cpu/src/pnt2xyz.rs:39-96
//   binning + sort      del-splat-cuda-kernels/cuda/tile_acceleration.h:15-86,
//                       del-splat-cpu/src/tile_acceleration.rs:3-97
//   tiled raster        del-splat-cuda-kernels/cuda/splat_gauss.cu:56-109
//   naive raster        del-splat-cpu/src/pnt2pixxyndcz.rs:42-100
//   circle projection   del-splat-cuda-kernels/cuda/splat_circle.cu:7-41,
//                       del-splat-cpu/src/splat_sphere.rs:51-87
//   circle raster       del-splat-cuda-kernels/cuda/splat_circle.cu:93-132
//   point z-buffer      del-splat-cuda-kernels/cuda/splat_point.cu:1-74,
//                       del-splat-cpu/src/pnt2pixxyndcz.rs:1-40
//   transfer adjoint    del-splat-cpu/src/splat_gaussian2.rs:201-266 (commented
//                       prototype; the full adjoint below is new)
//   3DGS PLY reader     del-splat-cpu/src/io_ply.rs:128-227
//   SH colour           new (coefficients read at io_ply.rs:196, unused there)
//
// The math helpers of the un-vendored del-geo crates (transform_homogeneous,
// jacobian_transform, transform_ndc2pix, projected_spd_mat3,
// safe_inverse_preserve_positive_definiteness, aabb2) are restated from their
// mathematical definitions; their pins are closed-form tests in tests/.

#include <omp.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>

// Threads of the projection, binning, raster and adjoint loops (1: the
// sequential restatement, the parity oracle; more: the all-core CPU baseline,
// with identical results except the adjoint's float sums, which are reduced
// in another order).
static int g_threads = 1;

namespace {

template <typename T> struct V2 { T x, y; };

// ---------------------------------------------------------------------------
// del-geo math, restated (column-major 4x4: m[i + 4*j] = M_ij)
// ---------------------------------------------------------------------------

// y = M (x, 1); returns w = y3 and q = y0..2 / w.
template <typename T>
inline T transform_homogeneous(const T* m, const T* x, T* q) {
  T y[4];
  for (int i = 0; i < 4; ++i) {
    y[i] = m[i] * x[0] + m[i + 4] * x[1] + m[i + 8] * x[2] + m[i + 12];
  }
  q[0] = y[0] / y[3];
  q[1] = y[1] / y[3];
  q[2] = y[2] / y[3];
  return y[3];
}

// J_ij = d q_i / d x_j = (M_ij - q_i M_3j) / w, column-major 3x3 j[i + 3*k].
template <typename T>
inline void jacobian_transform(const T* m, const T* q, T w, T* jac) {
  for (int j = 0; j < 3; ++j) {
    for (int i = 0; i < 3; ++i) {
      jac[i + 3 * j] = (m[i + 4 * j] - q[i] * m[3 + 4 * j]) / w;
    }
  }
}

// ndc -> pixel affine map as a column-major 2x3 (transform_ndc2pix):
// [W/2, 0, 0, -H/2, W/2, H/2], i.e. pix = (W/2 (x+1), H/2 (1-y)).
template <typename T>
inline void ndc2pix(uint32_t w, uint32_t h, T* a) {
  a[0] = T(0.5) * T(w); a[1] = T(0);
  a[2] = T(0);          a[3] = T(-0.5) * T(h);
  a[4] = T(0.5) * T(w); a[5] = T(0.5) * T(h);
}

// Hamilton unit quaternion (x, y, z, w) -> column-major 3x3 rotation.
template <typename T>
inline void quat_to_mat3(const T* qt, T* r) {
  const T x = qt[0], y = qt[1], z = qt[2], w = qt[3];
  r[0] = T(1) - T(2) * (y * y + z * z);
  r[1] = T(2) * (x * y + w * z);
  r[2] = T(2) * (x * z - w * y);
  r[3] = T(2) * (x * y - w * z);
  r[4] = T(1) - T(2) * (x * x + z * z);
  r[5] = T(2) * (y * z + w * x);
  r[6] = T(2) * (x * z + w * y);
  r[7] = T(2) * (y * z - w * x);
  r[8] = T(1) - T(2) * (x * x + y * y);
}

// lambda_min of the symmetric 2x2 [a b; b c].
template <typename T>
inline T sym2_min_eig(T a, T b, T c) {
  const T h = (a - c) * T(0.5);
  return (a + c) * T(0.5) - std::sqrt(h * h + b * b);
}

// Inverse of SPD [a b; b c] that keeps the result positive definite: the
// plain inverse when both eigenvalues are >= eps, otherwise the eigenvalues
// are clamped to eps before inverting. Returns 1 if the clamp fired.
template <typename T>
inline int safe_inverse_spd(const T* abc, T eps, T* out) {
  const T a = abc[0], b = abc[1], c = abc[2];
  if (sym2_min_eig(a, b, c) >= eps) {
    const T det = a * c - b * b;
    out[0] = c / det;
    out[1] = -b / det;
    out[2] = a / det;
    return 0;
  }
  // eigen-decomposition of the 2x2
  const T h = (a - c) * T(0.5);
  const T r = std::sqrt(h * h + b * b);
  T l0 = (a + c) * T(0.5) + r;
  T l1 = (a + c) * T(0.5) - r;
  T ux, uy;  // eigenvector of l0
  if (r == T(0)) { ux = T(1); uy = T(0); }
  else if (h >= T(0)) { ux = h + r; uy = b; }
  else { ux = b; uy = r - h; }
  const T nrm = std::sqrt(ux * ux + uy * uy);
  ux /= nrm; uy /= nrm;
  l0 = std::max(l0, eps);
  l1 = std::max(l1, eps);
  const T i0 = T(1) / l0, i1 = T(1) / l1;
  // inverse = i0 u u^T + i1 v v^T with v = (-uy, ux)
  out[0] = i0 * ux * ux + i1 * uy * uy;
  out[1] = i0 * ux * uy - i1 * ux * uy;
  out[2] = i0 * uy * uy + i1 * ux * ux;
  return 1;
}

// ---------------------------------------------------------------------------
// Gaussian projection: splat_gauss.cu:10-53 / pnt2xyz.rs:51-96
// ---------------------------------------------------------------------------
template <typename T>
struct Proj {
  T pix[2];
  T ndcz;
  T w;
  T q[3];
  T jac[9];     // ndc jacobian (column-major 3x3)
  T wpix[6];    // world->pix jacobian (column-major 2x3), includes ndc-z row
  T rot[9];
  T s2[3];
  T wr[6];      // wpix * rot
  T cov[3];     // projected 2x2 covariance [a b c]
  T conic[3];
  T aabb[4];
  int clamped;
  bool culled;
};

template <typename T>
inline void project_one(const T* xyz, const T* quat, const T* scale,
                        const T* m, uint32_t img_w, uint32_t img_h, Proj<T>& p) {
  p.w = transform_homogeneous(m, xyz, p.q);
  jacobian_transform(m, p.q, p.w, p.jac);
  T a[6];
  ndc2pix(img_w, img_h, a);
  // pix = A (q0, q1, 1)
  p.pix[0] = a[0] * p.q[0] + a[2] * p.q[1] + a[4];
  p.pix[1] = a[1] * p.q[0] + a[3] * p.q[1] + a[5];
  p.ndcz = p.q[2];
  // world2pix = A J  (the third column of A multiplies the ndc-z row of J)
  for (int j = 0; j < 3; ++j) {
    for (int i = 0; i < 2; ++i) {
      p.wpix[i + 2 * j] = a[i] * p.jac[0 + 3 * j] + a[i + 2] * p.jac[1 + 3 * j] +
                          a[i + 4] * p.jac[2 + 3 * j];
    }
  }
  quat_to_mat3(quat, p.rot);
  for (int k = 0; k < 3; ++k) p.s2[k] = scale[k] * scale[k];
  for (int k = 0; k < 3; ++k) {
    for (int i = 0; i < 2; ++i) {
      p.wr[i + 2 * k] = p.wpix[i] * p.rot[0 + 3 * k] + p.wpix[i + 2] * p.rot[1 + 3 * k] +
                        p.wpix[i + 4] * p.rot[2 + 3 * k];
    }
  }
  // cov = (W R) diag(s^2) (W R)^T
  p.cov[0] = p.wr[0] * p.wr[0] * p.s2[0] + p.wr[2] * p.wr[2] * p.s2[1] + p.wr[4] * p.wr[4] * p.s2[2];
  p.cov[1] = p.wr[0] * p.wr[1] * p.s2[0] + p.wr[2] * p.wr[3] * p.s2[1] + p.wr[4] * p.wr[5] * p.s2[2];
  p.cov[2] = p.wr[1] * p.wr[1] * p.s2[0] + p.wr[3] * p.wr[3] * p.s2[1] + p.wr[5] * p.wr[5] * p.s2[2];
  p.clamped = safe_inverse_spd(p.cov, T(1.0e-5), p.conic);
  // aabb2 of the conic, scaled by 3 about its centre, translated to pix
  const T det = p.conic[0] * p.conic[2] - p.conic[1] * p.conic[1];
  const T hx = std::sqrt(p.conic[2] / det);
  const T hy = std::sqrt(p.conic[0] / det);
  p.aabb[0] = p.pix[0] - T(3) * hx;
  p.aabb[1] = p.pix[1] - T(3) * hy;
  p.aabb[2] = p.pix[0] + T(3) * hx;
  p.aabb[3] = p.pix[1] + T(3) * hy;
  // A Gaussian on or behind the eye plane (w <= 0) has no meaningful
  // projection; it is culled with an inverted (empty) box.
  p.culled = !(p.w > T(0));
  if (p.culled) {
    p.aabb[0] = p.aabb[1] = INFINITY;
    p.aabb[2] = p.aabb[3] = -INFINITY;
  }
}

template <typename T>
uint32_t gauss_project(uint32_t n, const T* xyz, const T* quat, const T* scale,
                       const T* m, uint32_t w, uint32_t h, T* pixxydepth, T* conic, T* aabb) {
  uint32_t nclamp = 0;
#pragma omp parallel for schedule(static) reduction(+ : nclamp) if (g_threads > 1) \
    num_threads(g_threads)
  for (uint32_t i = 0; i < n; ++i) {
    Proj<T> p;
    project_one(xyz + 3 * i, quat + 4 * i, scale + 3 * i, m, w, h, p);
    pixxydepth[3 * i + 0] = p.pix[0];
    pixxydepth[3 * i + 1] = p.pix[1];
    pixxydepth[3 * i + 2] = p.ndcz;
    for (int k = 0; k < 3; ++k) conic[3 * i + k] = p.conic[k];
    for (int k = 0; k < 4; ++k) aabb[4 * i + k] = p.aabb[k];
    nclamp += uint32_t(p.clamped);
  }
  return nclamp;
}

// ---------------------------------------------------------------------------
// Binning: tile_acceleration.h:15-86 / tile_acceleration.rs:3-77
// ---------------------------------------------------------------------------
struct TileRect { int x0, x1, y0, y1; };  // half-open, clipped to the grid

// Tile rectangle of a box: tiles floor(x0 / ts) .. floor(x1 / ts) along each
// axis, as the loops of tile_acceleration.h:27-44 run them (so a finite
// inverted box inside one tile row or column still covers that tile, and
// one spanning a tile boundary none); a NaN coordinate covers no tile.
inline TileRect tile_rect(const float* bb, uint32_t ts, uint32_t tw, uint32_t th) {
  TileRect r{0, 0, 0, 0};
  if (std::isnan(bb[0]) || std::isnan(bb[1]) || std::isnan(bb[2]) || std::isnan(bb[3])) return r;
  const float tsf = float(ts);
  // floor(x / ts) clamped to [-1, tw] before the integer conversion.
  auto fl = [&](float v, uint32_t lim) {
    float f = std::floor(v / tsf);
    f = std::max(f, -1.0f);
    f = std::min(f, float(lim));
    return int(f);
  };
  r.x0 = std::max(fl(bb[0], tw), 0);
  r.x1 = std::min(fl(bb[2], tw) + 1, int(tw));
  r.y0 = std::max(fl(bb[1], th), 0);
  r.y1 = std::min(fl(bb[3], th) + 1, int(th));
  if (r.x1 < r.x0) r.x1 = r.x0;
  if (r.y1 < r.y0) r.y1 = r.y0;
  return r;
}

inline uint32_t f32_bits(float v) { uint32_t u; std::memcpy(&u, &v, 4); return u; }

// key depth: bits of max(ndc_z + 1, 0) (tile_acceleration.h:43-46)
inline float depth_plus_one(float ndcz) {
  float zp1 = ndcz + 1.f;
  if (zp1 <= 0.f) zp1 = 0.f;
  return zp1;
}

// tile2pnt (tile_acceleration.rs:3-58) over explicit boxes. Writes tile2idx
// [T+1]; if idx2pnt != null also the stably sorted pnt list and, if keys !=
// null, the sorted 64-bit keys tile<<32 | bits(depth+1).
uint64_t tile2pnt_parallel(uint32_t n, const float* boxes, const float* pixxydepth,
                           uint32_t img_w, uint32_t img_h, uint32_t ts, uint32_t* tile2idx,
                           uint32_t* idx2pnt, uint64_t* keys);

uint64_t tile2pnt(uint32_t n, const float* boxes, const float* pixxydepth,
                  uint32_t img_w, uint32_t img_h, uint32_t ts,
                  uint32_t* tile2idx, uint32_t* idx2pnt, uint64_t* keys) {
  if (g_threads > 1)
    return tile2pnt_parallel(n, boxes, pixxydepth, img_w, img_h, ts, tile2idx, idx2pnt, keys);
  const uint32_t tw = (img_w + ts - 1) / ts, th = (img_h + ts - 1) / ts;
  const uint32_t nt = tw * th;
  std::vector<uint64_t> cnt(nt + 1, 0);
  for (uint32_t i = 0; i < n; ++i) {
    TileRect r = tile_rect(boxes + 4 * i, ts, tw, th);
    for (int ix = r.x0; ix < r.x1; ++ix)
      for (int iy = r.y0; iy < r.y1; ++iy) cnt[iy * tw + ix + 1] += 1;
  }
  for (uint32_t t = 0; t < nt; ++t) cnt[t + 1] += cnt[t];
  const uint64_t K = cnt[nt];
  if (tile2idx) for (uint32_t t = 0; t <= nt; ++t) tile2idx[t] = uint32_t(cnt[t]);
  if (!idx2pnt) return K;
  struct Ent { uint32_t pnt; uint32_t tile; float depth; };
  std::vector<Ent> ents;
  ents.reserve(K);
  for (uint32_t i = 0; i < n; ++i) {
    TileRect r = tile_rect(boxes + 4 * i, ts, tw, th);
    const float d = depth_plus_one(pixxydepth[3 * i + 2]);
    for (int ix = r.x0; ix < r.x1; ++ix)
      for (int iy = r.y0; iy < r.y1; ++iy) ents.push_back({i, uint32_t(iy * tw + ix), d});
  }
  // stable: ties keep fill order (ascending pnt); depth compared as floats
  std::stable_sort(ents.begin(), ents.end(), [](const Ent& a, const Ent& b) {
    if (a.tile != b.tile) return a.tile < b.tile;
    return a.depth < b.depth;
  });
  for (uint64_t k = 0; k < K; ++k) {
    idx2pnt[k] = ents[k].pnt;
    if (keys) keys[k] = (uint64_t(ents[k].tile) << 32) | f32_bits(ents[k].depth);
  }
  return K;
}

// The same result with g_threads threads: a stable counting sort by tile
// (each thread fills the pairs of a contiguous index range into its own slots
// of every tile, so a tile's entries stay in index order), then a stable sort
// of each tile's entries by depth -- the order of the sequential stable sort
// by (tile, depth).
uint64_t tile2pnt_parallel(uint32_t n, const float* boxes, const float* pixxydepth,
                           uint32_t img_w, uint32_t img_h, uint32_t ts, uint32_t* tile2idx,
                           uint32_t* idx2pnt, uint64_t* keys) {
  const uint32_t tw = (img_w + ts - 1) / ts, th = (img_h + ts - 1) / ts;
  const uint32_t nt = tw * th;
  const int nthr = g_threads;
  std::vector<uint64_t> cnt(size_t(nthr) * nt, 0);  // [thread][tile]
  auto range = [&](int t, uint32_t* a, uint32_t* b) {
    *a = uint32_t(uint64_t(n) * t / nthr);
    *b = uint32_t(uint64_t(n) * (t + 1) / nthr);
  };
#pragma omp parallel num_threads(nthr)
  {
    const int t = omp_get_thread_num();
    uint32_t a, b;
    range(t, &a, &b);
    uint64_t* c = cnt.data() + size_t(t) * nt;
    for (uint32_t i = a; i < b; ++i) {
      TileRect r = tile_rect(boxes + 4 * i, ts, tw, th);
      for (int ix = r.x0; ix < r.x1; ++ix)
        for (int iy = r.y0; iy < r.y1; ++iy) c[iy * tw + ix] += 1;
    }
  }
  std::vector<uint64_t> tbase(nt + 1, 0);
  for (uint32_t tile = 0; tile < nt; ++tile) {
    uint64_t sum = 0;
    for (int t = 0; t < nthr; ++t) sum += cnt[size_t(t) * nt + tile];
    tbase[tile + 1] = tbase[tile] + sum;
  }
  const uint64_t K = tbase[nt];
  if (tile2idx) for (uint32_t t = 0; t <= nt; ++t) tile2idx[t] = uint32_t(tbase[t]);
  if (!idx2pnt) return K;
  // per (thread, tile) first slot
  for (uint32_t tile = 0; tile < nt; ++tile) {
    uint64_t run = tbase[tile];
    for (int t = 0; t < nthr; ++t) {
      const uint64_t v = cnt[size_t(t) * nt + tile];
      cnt[size_t(t) * nt + tile] = run;
      run += v;
    }
  }
  struct Ent { uint32_t pnt; float depth; };
  std::vector<Ent> ents(K);
#pragma omp parallel num_threads(nthr)
  {
    const int t = omp_get_thread_num();
    uint32_t a, b;
    range(t, &a, &b);
    uint64_t* c = cnt.data() + size_t(t) * nt;
    for (uint32_t i = a; i < b; ++i) {
      TileRect r = tile_rect(boxes + 4 * i, ts, tw, th);
      const float d = depth_plus_one(pixxydepth[3 * i + 2]);
      for (int ix = r.x0; ix < r.x1; ++ix)
        for (int iy = r.y0; iy < r.y1; ++iy) ents[c[iy * tw + ix]++] = {i, d};
    }
  }
#pragma omp parallel for schedule(dynamic, 8) num_threads(nthr)
  for (int64_t tile = 0; tile < int64_t(nt); ++tile) {
    std::stable_sort(ents.begin() + int64_t(tbase[tile]), ents.begin() + int64_t(tbase[tile + 1]),
                     [](const Ent& x, const Ent& y) { return x.depth < y.depth; });
    for (uint64_t k = tbase[tile]; k < tbase[tile + 1]; ++k) {
      idx2pnt[k] = ents[k].pnt;
      if (keys) keys[k] = (uint64_t(tile) << 32) | f32_bits(ents[k].depth);
    }
  }
  return K;
}

// ---------------------------------------------------------------------------
// Tiled front-to-back raster: splat_gauss.cu:56-109
// ---------------------------------------------------------------------------
// Extra outputs beyond the reference's RGB: alpha (= the reference's
// alpha_sum), nwalk (1 + list position of the last splat that contributed
// to the pixel -- the terminating one if the walk terminated --, 0 if none)
// and tfinal (the pixel's transmittance after its walk, behind that splat; 1
// if none).
// If frozen_nwalk != null the walk length per pixel is taken from it and the
// termination test is skipped (the function the adjoint differentiates).
// exp_ulp != 0 moves every exp result that many ulps up (> 0) or down (< 0):
// a planted perturbation that measures how many pixels' termination at 0.999
// is decided by the last bit of the exponential (tests only).
// One pixel (ix, iy) of the tiled raster over its tile's list
// idx2pnt[i0 .. i0 + num), walked from the end. lim caps the walk (the
// frozen length when frozen is set, which also skips the termination test).
template <typename T>
struct PixelOut {
  T c[3];
  T alpha_sum, tfinal;
  uint32_t walked;
};

template <typename T>
inline PixelOut<T> raster_pixel(uint32_t ix, uint32_t iy, const T* pixxydepth, const T* conic,
                                const float* aabb, const T* opacity, const T* rgb,
                                const uint32_t* idx2pnt, uint32_t i0, uint32_t num,
                                uint32_t lim, bool frozen, int exp_ulp) {
  const float tx = float(ix) + 0.5f, ty = float(iy) + 0.5f;
  PixelOut<T> o{{0, 0, 0}, 0, 1, 0};
  T alpha_occu = 1;
  for (uint32_t it = 0; it < lim; ++it) {
    const uint32_t p = idx2pnt[i0 + num - 1 - it];
    const float* bb = aabb + 4 * p;
    // inclusive box test (aabb2::is_include_point)
    if (!(tx >= bb[0] && tx <= bb[2] && ty >= bb[1] && ty <= bb[3])) continue;
    const T dx = T(tx) - pixxydepth[3 * p + 0];
    const T dy = T(ty) - pixxydepth[3 * p + 1];
    const T* cn = conic + 3 * p;
    const T q = cn[0] * dx * dx + cn[1] * (dx * dy + dy * dx) + cn[2] * dy * dy;
    T g = std::exp(T(-0.5) * q);
    for (int u = 0; u < (exp_ulp < 0 ? -exp_ulp : exp_ulp); ++u)
      g = std::nextafter(g, exp_ulp > 0 ? T(INFINITY) : T(-INFINITY));
    const T e = g * opacity[p];
    const T eo = alpha_occu * e;
    o.c[0] += rgb[3 * p + 0] * eo;
    o.c[1] += rgb[3 * p + 1] * eo;
    o.c[2] += rgb[3 * p + 2] * eo;
    o.walked = it + 1;
    // 1 - e floored at 2^-24: only a splat with e within 2^-24 of 1 (or
    // above) is affected, the walk ends there, and the colour and alpha do
    // not depend on it; the stored transmittance then divides back exactly
    alpha_occu *= std::max(T(1) - e, T(0x1p-24));
    o.tfinal = alpha_occu;
    o.alpha_sum += eo;
    // the reference compares the float sum against the double 0.999
    if (!frozen && double(o.alpha_sum) > 0.999) break;
  }
  if (frozen) o.walked = lim;
  return o;
}

template <typename T>
void gauss_rasterize(uint32_t img_w, uint32_t img_h, const T* pixxydepth, const T* conic,
                     const float* aabb, const T* opacity, const T* rgb, uint32_t ts,
                     const uint32_t* tile2idx, const uint32_t* idx2pnt,
                     T* pix2rgb, T* pix2alpha, uint32_t* pix2nwalk, T* pix2tfinal,
                     const uint32_t* frozen_nwalk, int exp_ulp = 0) {
  const uint32_t tw = (img_w + ts - 1) / ts;
#pragma omp parallel for schedule(dynamic, 2) if (g_threads > 1) num_threads(g_threads)
  for (uint32_t iy = 0; iy < img_h; ++iy) {
    for (uint32_t ix = 0; ix < img_w; ++ix) {
      const uint32_t ipix = iy * img_w + ix;
      const uint32_t tile = (iy / ts) * tw + ix / ts;
      const uint32_t i0 = tile2idx[tile], num = tile2idx[tile + 1] - i0;
      const uint32_t lim = frozen_nwalk ? frozen_nwalk[ipix] : num;
      const PixelOut<T> o = raster_pixel(ix, iy, pixxydepth, conic, aabb, opacity, rgb, idx2pnt,
                                         i0, num, lim, frozen_nwalk != nullptr, exp_ulp);
      for (int k = 0; k < 3; ++k) pix2rgb[3 * ipix + k] = o.c[k];
      if (pix2alpha) pix2alpha[ipix] = o.alpha_sum;
      if (pix2nwalk) pix2nwalk[ipix] = o.walked;
      if (pix2tfinal) pix2tfinal[ipix] = o.tfinal;
    }
  }
}

// sum (rgb - target)^2 over the pixels of the tiles [tx0, tx1] x [ty0, ty1]
// (clipped to the image), every walk frozen at frozen_nwalk: the part of the
// L2 loss that a change of one Gaussian inside those tiles can move, for
// central differences of one Gaussian at a time on a large frame.
template <typename T>
T gauss_l2_loss_tiles(uint32_t img_w, uint32_t img_h, const T* pixxydepth, const T* conic,
                      const float* aabb, const T* opacity, const T* rgb, uint32_t ts,
                      const uint32_t* tile2idx, const uint32_t* idx2pnt,
                      const uint32_t* frozen_nwalk, const float* target, uint32_t tx0,
                      uint32_t tx1, uint32_t ty0, uint32_t ty1) {
  const uint32_t tw = (img_w + ts - 1) / ts;
  const uint32_t y0 = ty0 * ts, y1 = std::min((ty1 + 1) * ts, img_h);
  const uint32_t x0 = tx0 * ts, x1 = std::min((tx1 + 1) * ts, img_w);
  if (y1 <= y0 || x1 <= x0) return T(0);
  // rows in parallel, summed in row order afterwards: the same result for
  // any thread count
  std::vector<T> rows(y1 - y0, T(0));
#pragma omp parallel for schedule(dynamic, 4)
  for (int64_t r = 0; r < int64_t(y1 - y0); ++r) {
    const uint32_t iy = y0 + uint32_t(r);
    T acc = 0;
    for (uint32_t ix = x0; ix < x1; ++ix) {
      const uint32_t ipix = iy * img_w + ix;
      const uint32_t tile = (iy / ts) * tw + ix / ts;
      const uint32_t i0 = tile2idx[tile], num = tile2idx[tile + 1] - i0;
      const PixelOut<T> o = raster_pixel(ix, iy, pixxydepth, conic, aabb, opacity, rgb, idx2pnt,
                                         i0, num, frozen_nwalk[ipix], true, 0);
      for (int k = 0; k < 3; ++k) {
        const T d = o.c[k] - T(target[3 * ipix + k]);
        acc += d * d;
      }
    }
    rows[size_t(r)] = acc;
  }
  T loss = 0;
  for (const T v : rows) loss += v;
  return loss;
}

// ---------------------------------------------------------------------------
// Adjoint of the tiled raster (new; structure follows the transfer
// derivative of splat_gaussian2.rs:212-243 without its O(n^2) loop).
// Walks each pixel's list back-to-front from the last contributing splat,
// recovering the transmittance by division from tfinal: at every
// contributing splat, the last one included, T in front of it is T behind it
// divided by its 1 - e.
// Gradients accumulate (+=) into: d_pixxy[N,2], d_conic[N,3], d_opacity[N],
// d_rgb[N,3]. d_pix2alpha may be null.
// ---------------------------------------------------------------------------
template <typename T>
void gauss_rasterize_bwd(uint32_t img_w, uint32_t img_h, const T* pixxydepth, const T* conic,
                         const float* aabb, const T* opacity, const T* rgb, uint32_t ts,
                         const uint32_t* tile2idx, const uint32_t* idx2pnt,
                         const uint32_t* pix2nwalk, const T* pix2tfinal,
                         const T* d_pix2rgb, const T* d_pix2alpha,
                         T* d_pixxy, T* d_conic, T* d_opacity, T* d_rgb,
                         uint32_t row0 = 0, uint32_t row1 = ~0u) {
  const uint32_t tw = (img_w + ts - 1) / ts;
  for (uint32_t iy = row0; iy < img_h && iy < row1; ++iy) {
    for (uint32_t ix = 0; ix < img_w; ++ix) {
      const uint32_t ipix = iy * img_w + ix;
      const uint32_t tile = (iy / ts) * tw + ix / ts;
      const float tx = float(ix) + 0.5f, ty = float(iy) + 0.5f;
      const uint32_t i0 = tile2idx[tile], num = tile2idx[tile + 1] - i0;
      const T g[3] = {d_pix2rgb[3 * ipix], d_pix2rgb[3 * ipix + 1], d_pix2rgb[3 * ipix + 2]};
      const T ga = d_pix2alpha ? d_pix2alpha[ipix] : T(0);
      T trans = pix2tfinal[ipix];  // transmittance behind the current splat
      T back[3] = {0, 0, 0};      // sum_{m behind} c_m e_m T_m
      T back_a = 0;               // sum_{m behind} e_m T_m
      for (int it = int(pix2nwalk[ipix]) - 1; it >= 0; --it) {
        const uint32_t p = idx2pnt[i0 + num - 1 - uint32_t(it)];
        const float* bb = aabb + 4 * p;
        if (!(tx >= bb[0] && tx <= bb[2] && ty >= bb[1] && ty <= bb[3])) continue;
        const T dx = T(tx) - pixxydepth[3 * p + 0];
        const T dy = T(ty) - pixxydepth[3 * p + 1];
        const T* cn = conic + 3 * p;
        const T q = cn[0] * dx * dx + cn[1] * (dx * dy + dy * dx) + cn[2] * dy * dy;
        const T gk = std::exp(T(-0.5) * q);
        const T op = opacity[p];
        const T e = gk * op;
        const T one_m_e = std::max(T(1) - e, T(0x1p-24));  // floored as in the forward
        trans = trans / one_m_e;  // now in front of it
        // dL/de
        T dl_de = 0;
        for (int k = 0; k < 3; ++k) {
          const T ck = rgb[3 * p + k];
          dl_de += g[k] * (ck * trans - back[k] / one_m_e);
          d_rgb[3 * p + k] += g[k] * e * trans;
        }
        dl_de += ga * (trans - back_a / one_m_e);
        for (int k = 0; k < 3; ++k) back[k] += rgb[3 * p + k] * e * trans;
        back_a += e * trans;
        d_opacity[p] += dl_de * gk;
        const T dl_dq = dl_de * e * T(-0.5);
        d_conic[3 * p + 0] += dl_dq * dx * dx;
        d_conic[3 * p + 1] += dl_dq * T(2) * dx * dy;
        d_conic[3 * p + 2] += dl_dq * dy * dy;
        // q depends on the mean through dx = tx - mx
        d_pixxy[2 * p + 0] += dl_dq * -(T(2) * cn[0] * dx + T(2) * cn[1] * dy);
        d_pixxy[2 * p + 1] += dl_dq * -(T(2) * cn[1] * dx + T(2) * cn[2] * dy);
      }
    }
  }
}

// The adjoint with g_threads threads over groups of rows, each thread into its
// own gradient buffers (n Gaussians), summed in thread order afterwards.
template <typename T>
void gauss_rasterize_bwd_mt(uint32_t n, uint32_t img_w, uint32_t img_h, const T* pixxydepth,
                            const T* conic, const float* aabb, const T* opacity, const T* rgb,
                            uint32_t ts, const uint32_t* tile2idx, const uint32_t* idx2pnt,
                            const uint32_t* pix2nwalk, const T* pix2tfinal, const T* d_pix2rgb,
                            const T* d_pix2alpha, T* d_pixxy, T* d_conic, T* d_opacity,
                            T* d_rgb) {
  const int nthr = g_threads;
  std::vector<T> buf(size_t(nthr) * 9 * n, T(0));
#pragma omp parallel num_threads(nthr)
  {
    const int t = omp_get_thread_num();
    T* b = buf.data() + size_t(t) * 9 * n;
    // groups of 4 rows dealt round-robin (a fixed assignment: the same sums
    // on every run; interleaved, as the image centre is the heavy part)
    for (uint32_t r0 = 4u * uint32_t(t); r0 < img_h; r0 += 4u * uint32_t(nthr))
      gauss_rasterize_bwd<T>(img_w, img_h, pixxydepth, conic, aabb, opacity, rgb, ts, tile2idx,
                             idx2pnt, pix2nwalk, pix2tfinal, d_pix2rgb, d_pix2alpha, b, b + 2 * n,
                             b + 5 * n, b + 6 * n, r0, r0 + 4u);
  }
#pragma omp parallel for schedule(static) num_threads(nthr)
  for (int64_t i = 0; i < int64_t(9) * n; ++i) {
    T v = 0;
    for (int t = 0; t < nthr; ++t) v += buf[size_t(t) * 9 * n + size_t(i)];
    T* dst = i < 2 * int64_t(n) ? d_pixxy + i
             : i < 5 * int64_t(n) ? d_conic + (i - 2 * int64_t(n))
             : i < 6 * int64_t(n) ? d_opacity + (i - 5 * int64_t(n))
                                  : d_rgb + (i - 6 * int64_t(n));
    *dst += v;
  }
}

// ---------------------------------------------------------------------------
// Adjoint of the projection (new). Inputs: 2D grads from the raster adjoint.
// Outputs (overwritten): d_xyz[N,3], d_scale[N,3] (w.r.t. sigma, or w.r.t.
// log sigma when log_scale != 0), d_quat[N,4].
// ---------------------------------------------------------------------------
template <typename T>
void gauss_project_bwd(uint32_t n, const T* xyz, const T* quat, const T* scale, const T* m,
                       uint32_t img_w, uint32_t img_h, const T* d_pixxy, const T* d_conic,
                       T* d_xyz, T* d_scale, T* d_quat, int log_scale) {
  T a[6];
  ndc2pix(img_w, img_h, a);
#pragma omp parallel for schedule(static) if (g_threads > 1) num_threads(g_threads)
  for (uint32_t i = 0; i < n; ++i) {
    Proj<T> p;
    project_one(xyz + 3 * i, quat + 4 * i, scale + 3 * i, m, img_w, img_h, p);
    for (int k = 0; k < 3; ++k) d_xyz[3 * i + k] = 0;
    for (int k = 0; k < 3; ++k) d_scale[3 * i + k] = 0;
    for (int k = 0; k < 4; ++k) d_quat[4 * i + k] = 0;
    if (p.culled) continue;
    // conic = inv(cov): dL/dcov = -C G C with G the symmetric matrix of
    // dL/dconic (off-diagonal halved).
    const T ca = p.conic[0], cb = p.conic[1], cc = p.conic[2];
    const T ga = d_conic[3 * i + 0], gb = T(0.5) * d_conic[3 * i + 1], gc = d_conic[3 * i + 2];
    // CG
    const T m00 = ca * ga + cb * gb, m01 = ca * gb + cb * gc;
    const T m10 = cb * ga + cc * gb, m11 = cb * gb + cc * gc;
    // -(CG)C, symmetric
    const T s00 = -(m00 * ca + m01 * cb);
    const T s01 = -(m00 * cb + m01 * cc);
    const T s11 = -(m10 * cb + m11 * cc);
    // cov = WR S2 WR^T: dL/dWR = 2 Gs WR S2 ; dL/dS2_k = (WR^T Gs WR)_kk
    T d_wr[6];
    for (int k = 0; k < 3; ++k) {
      const T w0 = p.wr[0 + 2 * k], w1 = p.wr[1 + 2 * k];
      d_wr[0 + 2 * k] = T(2) * (s00 * w0 + s01 * w1) * p.s2[k];
      d_wr[1 + 2 * k] = T(2) * (s01 * w0 + s11 * w1) * p.s2[k];
      const T ds2 = w0 * (s00 * w0 + s01 * w1) + w1 * (s01 * w0 + s11 * w1);
      // s2 = s^2 -> d/ds = 2 s ds2 ; d/dlog s = 2 s^2 ds2
      d_scale[3 * i + k] = log_scale ? T(2) * p.s2[k] * ds2 : T(2) * scale[3 * i + k] * ds2;
    }
    // WR = Wpix R: dL/dWpix = dWR R^T ; dL/dR = Wpix^T dWR
    T d_wpix[6] = {0, 0, 0, 0, 0, 0};
    T d_rot[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
    for (int r = 0; r < 2; ++r)
      for (int j = 0; j < 3; ++j)
        for (int k = 0; k < 3; ++k) d_wpix[r + 2 * j] += d_wr[r + 2 * k] * p.rot[j + 3 * k];
    for (int j = 0; j < 3; ++j)
      for (int k = 0; k < 3; ++k)
        for (int r = 0; r < 2; ++r) d_rot[j + 3 * k] += p.wpix[r + 2 * j] * d_wr[r + 2 * k];
    // rotation -> quaternion
    {
      const T x = quat[4 * i + 0], y = quat[4 * i + 1], z = quat[4 * i + 2], w = quat[4 * i + 3];
      const T* g = d_rot;  // g[r + 3c]
      // partials of each entry of R (column-major) w.r.t. x, y, z, w
      const T dx_[9] = {0, 2 * y, 2 * z, 2 * y, -4 * x, 2 * w, 2 * z, -2 * w, -4 * x};
      const T dy_[9] = {-4 * y, 2 * x, -2 * w, 2 * x, 0, 2 * z, 2 * w, 2 * z, -4 * y};
      const T dz_[9] = {-4 * z, 2 * w, 2 * x, -2 * w, -4 * z, 2 * y, 2 * x, 2 * y, 0};
      const T dw_[9] = {0, 2 * z, -2 * y, -2 * z, 0, 2 * x, 2 * y, -2 * x, 0};
      T sx = 0, sy = 0, sz = 0, sw = 0;
      for (int k = 0; k < 9; ++k) {
        sx += g[k] * dx_[k]; sy += g[k] * dy_[k]; sz += g[k] * dz_[k]; sw += g[k] * dw_[k];
      }
      d_quat[4 * i + 0] = sx; d_quat[4 * i + 1] = sy; d_quat[4 * i + 2] = sz; d_quat[4 * i + 3] = sw;
    }
    // Wpix = A J: dL/dJ = A^T dWpix
    T d_jac[9];
    for (int j = 0; j < 3; ++j)
      for (int r = 0; r < 3; ++r)
        d_jac[r + 3 * j] = a[0 + 2 * r] * d_wpix[0 + 2 * j] + a[1 + 2 * r] * d_wpix[1 + 2 * j];
    // pix = A (q0, q1, 1): dL/dq
    const T gpx = d_pixxy[2 * i + 0], gpy = d_pixxy[2 * i + 1];
    const T d_q0 = a[0] * gpx + a[1] * gpy;
    const T d_q1 = a[2] * gpx + a[3] * gpy;
    // dq_i/dx_k = J_ik ; dJ_ij/dx_k = -(J_ik M_3j + J_ij M_3k) / w
    for (int k = 0; k < 3; ++k) {
      T s = d_q0 * p.jac[0 + 3 * k] + d_q1 * p.jac[1 + 3 * k];
      for (int j = 0; j < 3; ++j)
        for (int r = 0; r < 3; ++r) {
          const T djdx = -(p.jac[r + 3 * k] * m[3 + 4 * j] + p.jac[r + 3 * j] * m[3 + 4 * k]) / p.w;
          s += d_jac[r + 3 * j] * djdx;
        }
      d_xyz[3 * i + k] = s;
    }
  }
}

// ---------------------------------------------------------------------------
// Naive non-tiled renderer: pnt2pixxyndcz.rs:42-100 (global stable sort by
// ndc_z + 1 without clamp, ndc_z clip to (-1, 1), front-to-back).
// ---------------------------------------------------------------------------
void gauss_render_sort_depth(uint32_t n, const float* pixxydepth, const float* conic,
                             const float* aabb, const float* opacity, const float* rgb,
                             uint32_t img_w, uint32_t img_h, float* img) {
  std::vector<uint32_t> order(n);
  for (uint32_t i = 0; i < n; ++i) order[i] = i;
  std::stable_sort(order.begin(), order.end(), [&](uint32_t i, uint32_t j) {
    return pixxydepth[3 * i + 2] + 1.f < pixxydepth[3 * j + 2] + 1.f;
  });
  for (uint32_t ih = 0; ih < img_h; ++ih) {
    for (uint32_t iw = 0; iw < img_w; ++iw) {
      const float tx = float(iw) + 0.5f, ty = float(ih) + 0.5f;
      float alpha_sum = 0.f, alpha_occu = 1.f;
      float* c = img + 3 * (ih * img_w + iw);
      c[0] = c[1] = c[2] = 0.f;
      for (int64_t idx = int64_t(n) - 1; idx >= 0; --idx) {
        const uint32_t p = order[idx];
        const float z = pixxydepth[3 * p + 2];
        if (z <= -1.f || z >= 1.f) continue;
        const float* bb = aabb + 4 * p;
        if (!(tx >= bb[0] && tx <= bb[2] && ty >= bb[1] && ty <= bb[3])) continue;
        const float dx = tx - pixxydepth[3 * p], dy = ty - pixxydepth[3 * p + 1];
        const float* cn = conic + 3 * p;
        const float q = cn[0] * dx * dx + cn[1] * (dx * dy + dy * dx) + cn[2] * dy * dy;
        const float e = std::exp(-0.5f * q) * opacity[p];
        const float eo = alpha_occu * e;
        c[0] += rgb[3 * p] * eo;
        c[1] += rgb[3 * p + 1] * eo;
        c[2] += rgb[3 * p + 2] * eo;
        alpha_occu *= 1.f - e;
        alpha_sum += eo;
        if (double(alpha_sum) > 0.999) break;
      }
    }
  }
}

// ---------------------------------------------------------------------------
// Circle (sphere) projection: splat_circle.cu:7-41 / splat_sphere.rs:51-87
// ---------------------------------------------------------------------------
template <typename T>
inline bool inverse3(const T* a, T* inv) {
  // column-major 3x3 inverse by cofactors
  const T a00 = a[0], a10 = a[1], a20 = a[2], a01 = a[3], a11 = a[4], a21 = a[5],
          a02 = a[6], a12 = a[7], a22 = a[8];
  const T c00 = a11 * a22 - a12 * a21;
  const T c01 = a12 * a20 - a10 * a22;
  const T c02 = a10 * a21 - a11 * a20;
  const T det = a00 * c00 + a01 * c01 + a02 * c02;
  if (det == T(0)) return false;
  const T id = T(1) / det;
  inv[0] = c00 * id;
  inv[1] = c01 * id;
  inv[2] = c02 * id;
  inv[3] = (a02 * a21 - a01 * a22) * id;
  inv[4] = (a00 * a22 - a02 * a20) * id;
  inv[5] = (a01 * a20 - a00 * a21) * id;
  inv[6] = (a01 * a12 - a02 * a11) * id;
  inv[7] = (a02 * a10 - a00 * a12) * id;
  inv[8] = (a00 * a11 - a01 * a10) * id;
  return true;
}

void circle_project(uint32_t n, const float* xyz, float radius, const float* m,
                    uint32_t img_w, uint32_t img_h, float* pixxydepth, float* pixrad) {
  for (uint32_t i = 0; i < n; ++i) {
    float q[3], jac[9], inv[9];
    const float w = transform_homogeneous(m, xyz + 3 * i, q);
    pixxydepth[3 * i + 0] = (q[0] + 1.f) * 0.5f * float(img_w);
    pixxydepth[3 * i + 1] = (1.f - q[1]) * 0.5f * float(img_h);
    pixxydepth[3 * i + 2] = q[2];
    jacobian_transform(m, q, w, jac);
    float rad = 0.f;
    if (inverse3(jac, inv)) {
      const float nx = std::sqrt(inv[0] * inv[0] + inv[1] * inv[1] + inv[2] * inv[2]);
      const float ny = std::sqrt(inv[3] * inv[3] + inv[4] * inv[4] + inv[5] * inv[5]);
      const float rx = (1.f / nx) * 0.5f * float(img_w) * radius;
      const float ry = (1.f / ny) * 0.5f * float(img_h) * radius;
      rad = 0.5f * (rx + ry);
    }
    pixrad[i] = rad;
  }
}

// Circle painter raster: splat_circle.cu:93-132 (ascending = back to front,
// overwrite when the pixel centre is within the radius).
void circle_rasterize(uint32_t img_w, uint32_t img_h, const float* pixxydepth,
                      const float* pixrad, const float* rgb, uint32_t ts,
                      const uint32_t* tile2idx, const uint32_t* idx2pnt, float* pix2rgb) {
  const uint32_t tw = (img_w + ts - 1) / ts;
  for (uint32_t iy = 0; iy < img_h; ++iy)
    for (uint32_t ix = 0; ix < img_w; ++ix) {
      const uint32_t ipix = iy * img_w + ix;
      const uint32_t tile = (iy / ts) * tw + ix / ts;
      const float px = float(ix) + 0.5f, py = float(iy) + 0.5f;
      float* c = pix2rgb + 3 * ipix;
      c[0] = c[1] = c[2] = 0.f;
      for (uint32_t k = tile2idx[tile]; k < tile2idx[tile + 1]; ++k) {
        const uint32_t p = idx2pnt[k];
        const float dx = pixxydepth[3 * p] - px, dy = pixxydepth[3 * p + 1] - py;
        if (std::sqrt(dx * dx + dy * dy) > pixrad[p]) continue;
        c[0] = rgb[3 * p]; c[1] = rgb[3 * p + 1]; c[2] = rgb[3 * p + 2];
      }
    }
}

// ---------------------------------------------------------------------------
// Point / circle z-buffer (splat_point.cu:49-74 with the deterministic
// packed-key rule): key = bits(unit depth) << 32 | point index, the minimum
// key wins per pixel, then the winner's colour is gathered. unit depth =
// 1 - (ndc_z + 1) / 2, points with ndc_z outside [-1, 1] are clipped. With
// pixrad == null a point covers the pixel floor(x), floor(y); otherwise every
// pixel whose centre lies within pixrad of the point (circle raster test of
// splat_circle.cu:120-125). Empty pixels: key all-ones, colour 0.
// ---------------------------------------------------------------------------
inline bool point_key(const float* pxz, uint32_t i, uint64_t* key) {
  const float ndcz = pxz[2];
  if (ndcz < -1.f || ndcz > 1.f) return false;
  // splat_point.cu:61 evaluates 1.0-(ndcz + 1.)*0.5 in double
  const float u = float(1.0 - (double(ndcz) + 1.0) * 0.5);
  *key = (uint64_t(f32_bits(u)) << 32) | i;
  return true;
}

void zbuffer(uint32_t n, const float* pixxydepth, const float* pixrad, const float* rgb,
             uint32_t img_w, uint32_t img_h, uint64_t* pix2key, float* pix2rgb) {
  const size_t np = size_t(img_w) * img_h;
  for (size_t k = 0; k < np; ++k) pix2key[k] = ~uint64_t(0);
  for (uint32_t i = 0; i < n; ++i) {
    const float* pxz = pixxydepth + 3 * i;
    uint64_t key;
    if (!point_key(pxz, i, &key)) continue;
    if (!pixrad) {
      const float fx = std::floor(pxz[0]), fy = std::floor(pxz[1]);
      if (!(fx >= 0.f && fx < float(img_w) && fy >= 0.f && fy < float(img_h))) continue;
      const size_t ip = size_t(fy) * img_w + size_t(fx);
      pix2key[ip] = std::min(pix2key[ip], key);
      continue;
    }
    const float r = pixrad[i];
    if (!(r >= 0.f) || std::isnan(pxz[0]) || std::isnan(pxz[1])) continue;
    // pixel centres ix + 0.5 within [x - r, x + r], clamped to the image in
    // float before the conversion (a box off the image comes out empty)
    const float fx0 = std::ceil(pxz[0] - r - 0.5f), fx1 = std::floor(pxz[0] + r - 0.5f);
    const float fy0 = std::ceil(pxz[1] - r - 0.5f), fy1 = std::floor(pxz[1] + r - 0.5f);
    // (fmax / fmin: a NaN bound gives way to the other operand)
    const int x0 = int(std::fmin(std::fmax(fx0, 0.f), float(img_w)));
    const int x1 = int(std::fmax(std::fmin(fx1, float(img_w) - 1.f), -1.f));
    const int y0 = int(std::fmin(std::fmax(fy0, 0.f), float(img_h)));
    const int y1 = int(std::fmax(std::fmin(fy1, float(img_h) - 1.f), -1.f));
    for (int iy = y0; iy <= y1; ++iy)
      for (int ix = x0; ix <= x1; ++ix) {
        const float dx = pxz[0] - (float(ix) + 0.5f), dy = pxz[1] - (float(iy) + 0.5f);
        if (std::sqrt(dx * dx + dy * dy) > r) continue;
        const size_t ip = size_t(iy) * img_w + ix;
        pix2key[ip] = std::min(pix2key[ip], key);
      }
  }
  for (size_t k = 0; k < np; ++k) {
    const uint64_t key = pix2key[k];
    if (key == ~uint64_t(0)) { pix2rgb[3 * k] = pix2rgb[3 * k + 1] = pix2rgb[3 * k + 2] = 0.f; continue; }
    const uint32_t p = uint32_t(key & 0xffffffffu);
    pix2rgb[3 * k] = rgb[3 * p]; pix2rgb[3 * k + 1] = rgb[3 * p + 1]; pix2rgb[3 * k + 2] = rgb[3 * p + 2];
  }
}

// Point painter: render_pix_sort_depth (pnt2pixxyndcz.rs:1-40). Stable sort
// of the points by ndc_z + 1, then overwrite in that order; ndc_z outside
// (-1, 1) clipped; pixel = Rust's saturating `f32 as usize` cast. Pixels no
// point reaches are left at 0.
void render_pix_sort_depth(uint32_t n, const float* pxz, const float* rgb, uint32_t img_w,
                           uint32_t img_h, float* img) {
  std::vector<uint32_t> order(n);
  for (uint32_t i = 0; i < n; ++i) order[i] = i;
  std::stable_sort(order.begin(), order.end(), [&](uint32_t a, uint32_t b) {
    return pxz[3 * a + 2] + 1.f < pxz[3 * b + 2] + 1.f;
  });
  std::fill(img, img + size_t(img_w) * img_h * 3, 0.f);
  auto as_usize = [](float v) -> uint64_t {
    if (!(v > 0.f)) return 0;
    if (v >= 1.8446744e19f) return ~uint64_t(0);
    return uint64_t(v);
  };
  for (uint32_t p : order) {
    const float z = pxz[3 * p + 2];
    if (z <= -1.f || z >= 1.f) continue;
    const uint64_t ix = as_usize(pxz[3 * p]), iy = as_usize(pxz[3 * p + 1]);
    if (ix >= img_w || iy >= img_h) continue;
    float* c = img + 3 * (iy * img_w + ix);
    c[0] = rgb[3 * p]; c[1] = rgb[3 * p + 1]; c[2] = rgb[3 * p + 2];
  }
}

// ---------------------------------------------------------------------------
// View-dependent colour from spherical harmonics (new; the reference reads
// the 45 higher-order coefficients, io_ply.rs:196, but never evaluates them).
//
// Basis: real spherical harmonics of degree l <= 3, coefficient k = l^2+l+m,
//   Y_lm = sqrt(2) Re Y_l^|m| (m > 0), sqrt(2) Im Y_l^|m| (m < 0), Y_l^0,
// of the complex harmonics Y_l^m with the Condon-Shortley phase (the basis
// of 3D Gaussian splatting's colour model). Written as polynomials of the
// unit direction d = (x, y, z) from the camera centre to the mean:
//   k0  1/(2 sqrt(pi))
//   k1..3   c1 (-y, z, -x),                       c1 = sqrt(3/(4 pi))
//   k4..8   c2a xy, -c2a yz, c2b (2z^2-x^2-y^2), -c2a xz, c2c (x^2-y^2)
//           c2a = sqrt(15/(4 pi)), c2b = sqrt(5/(16 pi)), c2c = sqrt(15/(16 pi))
//   k9..15  -c3a y(3x^2-y^2), c3b xyz, -c3c y(4z^2-x^2-y^2),
//           c3d z(2z^2-3x^2-3y^2), -c3c x(4z^2-x^2-y^2), c3e z(x^2-y^2),
//           -c3a x(x^2-3y^2)
//           c3a = sqrt(35/(32 pi)), c3b = sqrt(105/(4 pi)), c3c = sqrt(21/(32 pi)),
//           c3d = sqrt(7/(16 pi)), c3e = sqrt(105/(16 pi))
// Colour channel c: max(sum_k Y_k(d) sh[k][c] + 1/2, 0), sh laid out
// [N][K][3] (coefficient-major, channel inner), K >= (degree+1)^2.
// ---------------------------------------------------------------------------
template <typename T>
struct ShConst {
  static T pi() { return T(3.14159265358979323846); }
  static T k0() { return T(std::sqrt(1.0 / (4.0 * 3.14159265358979323846))); }
  static T c1() { return T(std::sqrt(3.0 / (4.0 * 3.14159265358979323846))); }
  static T c2a() { return T(std::sqrt(15.0 / (4.0 * 3.14159265358979323846))); }
  static T c2b() { return T(std::sqrt(5.0 / (16.0 * 3.14159265358979323846))); }
  static T c2c() { return T(std::sqrt(15.0 / (16.0 * 3.14159265358979323846))); }
  static T c3a() { return T(std::sqrt(35.0 / (32.0 * 3.14159265358979323846))); }
  static T c3b() { return T(std::sqrt(105.0 / (4.0 * 3.14159265358979323846))); }
  static T c3c() { return T(std::sqrt(21.0 / (32.0 * 3.14159265358979323846))); }
  static T c3d() { return T(std::sqrt(7.0 / (16.0 * 3.14159265358979323846))); }
  static T c3e() { return T(std::sqrt(105.0 / (16.0 * 3.14159265358979323846))); }
};

// b[k] for k < (deg+1)^2.
template <typename T>
inline void sh_basis(T x, T y, T z, uint32_t deg, T* b) {
  using C = ShConst<T>;
  b[0] = C::k0();
  if (deg < 1) return;
  b[1] = -C::c1() * y;
  b[2] = C::c1() * z;
  b[3] = -C::c1() * x;
  if (deg < 2) return;
  const T xx = x * x, yy = y * y, zz = z * z;
  b[4] = C::c2a() * (x * y);
  b[5] = -C::c2a() * (y * z);
  b[6] = C::c2b() * (T(2) * zz - xx - yy);
  b[7] = -C::c2a() * (x * z);
  b[8] = C::c2c() * (xx - yy);
  if (deg < 3) return;
  b[9] = -C::c3a() * (y * (T(3) * xx - yy));
  b[10] = C::c3b() * (x * y * z);
  b[11] = -C::c3c() * (y * (T(4) * zz - xx - yy));
  b[12] = C::c3d() * (z * (T(2) * zz - T(3) * xx - T(3) * yy));
  b[13] = -C::c3c() * (x * (T(4) * zz - xx - yy));
  b[14] = C::c3e() * (z * (xx - yy));
  b[15] = -C::c3a() * (x * (xx - T(3) * yy));
}

// g[3k + i] = d b[k] / d d_i of the polynomials above (taken in R^3; the
// caller keeps only the part tangent to the sphere).
template <typename T>
inline void sh_basis_grad(T x, T y, T z, uint32_t deg, T* g) {
  using C = ShConst<T>;
  for (uint32_t i = 0; i < 48; ++i) g[i] = T(0);
  if (deg < 1) return;
  g[3 * 1 + 1] = -C::c1();
  g[3 * 2 + 2] = C::c1();
  g[3 * 3 + 0] = -C::c1();
  if (deg < 2) return;
  const T xx = x * x, yy = y * y, zz = z * z;
  g[3 * 4 + 0] = C::c2a() * y;          g[3 * 4 + 1] = C::c2a() * x;
  g[3 * 5 + 1] = -C::c2a() * z;         g[3 * 5 + 2] = -C::c2a() * y;
  g[3 * 6 + 0] = -T(2) * C::c2b() * x;  g[3 * 6 + 1] = -T(2) * C::c2b() * y;
  g[3 * 6 + 2] = T(4) * C::c2b() * z;
  g[3 * 7 + 0] = -C::c2a() * z;         g[3 * 7 + 2] = -C::c2a() * x;
  g[3 * 8 + 0] = T(2) * C::c2c() * x;   g[3 * 8 + 1] = -T(2) * C::c2c() * y;
  if (deg < 3) return;
  g[3 * 9 + 0] = -C::c3a() * (T(6) * x * y);
  g[3 * 9 + 1] = -C::c3a() * (T(3) * xx - T(3) * yy);
  g[3 * 10 + 0] = C::c3b() * (y * z);
  g[3 * 10 + 1] = C::c3b() * (x * z);
  g[3 * 10 + 2] = C::c3b() * (x * y);
  g[3 * 11 + 0] = -C::c3c() * (-T(2) * x * y);
  g[3 * 11 + 1] = -C::c3c() * (T(4) * zz - xx - T(3) * yy);
  g[3 * 11 + 2] = -C::c3c() * (T(8) * y * z);
  g[3 * 12 + 0] = C::c3d() * (-T(6) * x * z);
  g[3 * 12 + 1] = C::c3d() * (-T(6) * y * z);
  g[3 * 12 + 2] = C::c3d() * (T(6) * zz - T(3) * xx - T(3) * yy);
  g[3 * 13 + 0] = -C::c3c() * (T(4) * zz - T(3) * xx - yy);
  g[3 * 13 + 1] = -C::c3c() * (-T(2) * x * y);
  g[3 * 13 + 2] = -C::c3c() * (T(8) * x * z);
  g[3 * 14 + 0] = C::c3e() * (T(2) * x * z);
  g[3 * 14 + 1] = C::c3e() * (-T(2) * y * z);
  g[3 * 14 + 2] = C::c3e() * (xx - yy);
  g[3 * 15 + 0] = -C::c3a() * (T(3) * xx - T(3) * yy);
  g[3 * 15 + 1] = -C::c3a() * (-T(6) * x * y);
}

// Unit direction from the camera centre to the mean, and 1/|xyz - cam|.
template <typename T>
inline void sh_direction(const T* cam, const T* p, T* d, T* inv_len) {
  const T vx = p[0] - cam[0], vy = p[1] - cam[1], vz = p[2] - cam[2];
  const T inv = T(1) / std::sqrt(vx * vx + vy * vy + vz * vz);
  d[0] = vx * inv;
  d[1] = vy * inv;
  d[2] = vz * inv;
  *inv_len = inv;
}

template <typename T>
void sh_to_rgb(uint32_t n, uint32_t deg, uint32_t K, const T* cam, const T* xyz, const T* sh,
               T* rgb) {
  const uint32_t nk = (deg + 1) * (deg + 1);
  for (uint32_t i = 0; i < n; ++i) {
    T d[3], inv, b[16];
    sh_direction(cam, xyz + 3 * i, d, &inv);
    sh_basis(d[0], d[1], d[2], deg, b);
    const T* c = sh + size_t(i) * K * 3;
    for (uint32_t ch = 0; ch < 3; ++ch) {
      T acc = T(0);
      for (uint32_t k = 0; k < nk; ++k) acc += b[k] * c[3 * k + ch];
      acc += T(0.5);
      rgb[3 * i + ch] = acc > T(0) ? acc : T(0);
    }
  }
}

// Adjoint: d_sh[i][k][c] = Y_k g_c and d_xyz[i] = (I - d d^T) / |v| *
// sum_{k,c} sh[k][c] g_c grad Y_k, with g_c = d_rgb[c] where the colour was
// not clamped (sum + 1/2 >= 0) and 0 where it was. Stored (overwrite) or
// added. Coefficients k >= (deg+1)^2 get no gradient (left, or stored 0).
template <typename T>
void sh_to_rgb_bwd(uint32_t n, uint32_t deg, uint32_t K, const T* cam, const T* xyz,
                   const T* sh, const T* d_rgb, T* d_sh, T* d_xyz, bool overwrite) {
  const uint32_t nk = (deg + 1) * (deg + 1);
  for (uint32_t i = 0; i < n; ++i) {
    T d[3], inv, b[16], gb[48];
    sh_direction(cam, xyz + 3 * i, d, &inv);
    sh_basis(d[0], d[1], d[2], deg, b);
    sh_basis_grad(d[0], d[1], d[2], deg, gb);
    const T* c = sh + size_t(i) * K * 3;
    T g[3];
    for (uint32_t ch = 0; ch < 3; ++ch) {
      T acc = T(0);
      for (uint32_t k = 0; k < nk; ++k) acc += b[k] * c[3 * k + ch];
      acc += T(0.5);
      g[ch] = acc < T(0) ? T(0) : d_rgb[3 * i + ch];
    }
    T* ds = d_sh + size_t(i) * K * 3;
    for (uint32_t k = 0; k < K; ++k)
      for (uint32_t ch = 0; ch < 3; ++ch) {
        const T v = k < nk ? b[k] * g[ch] : T(0);
        if (overwrite) ds[3 * k + ch] = v;
        else ds[3 * k + ch] += v;
      }
    T dd[3] = {T(0), T(0), T(0)};
    for (uint32_t k = 1; k < nk; ++k) {
      const T w = c[3 * k] * g[0] + c[3 * k + 1] * g[1] + c[3 * k + 2] * g[2];
      dd[0] += w * gb[3 * k];
      dd[1] += w * gb[3 * k + 1];
      dd[2] += w * gb[3 * k + 2];
    }
    const T r = d[0] * dd[0] + d[1] * dd[1] + d[2] * dd[2];
    for (uint32_t a = 0; a < 3; ++a) {
      const T v = (dd[a] - d[a] * r) * inv;
      if (overwrite) d_xyz[3 * i + a] = v;
      else d_xyz[3 * i + a] += v;
    }
  }
}

}  // namespace

// ---------------------------------------------------------------------------
// C entry points (ctypes)
// ---------------------------------------------------------------------------
extern "C" {

uint32_t orc_gauss_project_f32(uint32_t n, const float* xyz, const float* quat, const float* scale,
                               const float* m, uint32_t w, uint32_t h, float* pixxydepth,
                               float* conic, float* aabb) {
  return gauss_project<float>(n, xyz, quat, scale, m, w, h, pixxydepth, conic, aabb);
}
uint32_t orc_gauss_project_f64(uint32_t n, const double* xyz, const double* quat,
                               const double* scale, const double* m, uint32_t w, uint32_t h,
                               double* pixxydepth, double* conic, double* aabb) {
  return gauss_project<double>(n, xyz, quat, scale, m, w, h, pixxydepth, conic, aabb);
}

// Intermediates of one projection, for closed-form tests: out[0..1] pix,
// [2] ndcz, [3] w, [4..12] jac, [13..18] wpix, [19..27] rot, [28..30] cov,
// [31..33] conic, [34..37] aabb, [38] clamped.
void orc_gauss_project_one_f64(const double* xyz, const double* quat, const double* scale,
                               const double* m, uint32_t w, uint32_t h, double* out) {
  Proj<double> p;
  project_one(xyz, quat, scale, m, w, h, p);
  out[0] = p.pix[0]; out[1] = p.pix[1]; out[2] = p.ndcz; out[3] = p.w;
  for (int k = 0; k < 9; ++k) out[4 + k] = p.jac[k];
  for (int k = 0; k < 6; ++k) out[13 + k] = p.wpix[k];
  for (int k = 0; k < 9; ++k) out[19 + k] = p.rot[k];
  for (int k = 0; k < 3; ++k) out[28 + k] = p.cov[k];
  for (int k = 0; k < 3; ++k) out[31 + k] = p.conic[k];
  for (int k = 0; k < 4; ++k) out[34 + k] = p.aabb[k];
  out[38] = p.clamped;
}

int orc_safe_inverse_spd_f64(const double* abc, double eps, double* out) {
  return safe_inverse_spd<double>(abc, eps, out);
}

uint64_t orc_tile2pnt(uint32_t n, const float* boxes, const float* pixxydepth, uint32_t w,
                      uint32_t h, uint32_t ts, uint32_t* tile2idx, uint32_t* idx2pnt,
                      uint64_t* keys) {
  return tile2pnt(n, boxes, pixxydepth, w, h, ts, tile2idx, idx2pnt, keys);
}

// circle boxes from_point(pix, r) = [x-r, y-r, x+r, y+r]
void orc_circle_boxes(uint32_t n, const float* pixxydepth, const float* pixrad, float* boxes) {
  for (uint32_t i = 0; i < n; ++i) {
    const float x = pixxydepth[3 * i], y = pixxydepth[3 * i + 1], r = pixrad[i];
    boxes[4 * i + 0] = x - r; boxes[4 * i + 1] = y - r;
    boxes[4 * i + 2] = x + r; boxes[4 * i + 3] = y + r;
  }
}

void orc_tile_rects(uint32_t n, const float* boxes, uint32_t ts, uint32_t tw, uint32_t th,
                    int32_t* rects) {
  for (uint32_t i = 0; i < n; ++i) {
    TileRect r = tile_rect(boxes + 4 * i, ts, tw, th);
    rects[4 * i] = r.x0; rects[4 * i + 1] = r.x1; rects[4 * i + 2] = r.y0; rects[4 * i + 3] = r.y1;
  }
}

void orc_gauss_rasterize_f32(uint32_t w, uint32_t h, const float* pixxydepth, const float* conic,
                             const float* aabb, const float* opacity, const float* rgb,
                             uint32_t ts, const uint32_t* tile2idx, const uint32_t* idx2pnt,
                             float* pix2rgb, float* pix2alpha, uint32_t* pix2nwalk,
                             float* pix2tfinal, const uint32_t* frozen) {
  gauss_rasterize<float>(w, h, pixxydepth, conic, aabb, opacity, rgb, ts, tile2idx, idx2pnt,
                         pix2rgb, pix2alpha, pix2nwalk, pix2tfinal, frozen);
}
void orc_gauss_rasterize_exp_ulp_f32(uint32_t w, uint32_t h, const float* pixxydepth,
                                     const float* conic, const float* aabb, const float* opacity,
                                     const float* rgb, uint32_t ts, const uint32_t* tile2idx,
                                     const uint32_t* idx2pnt, float* pix2rgb, float* pix2alpha,
                                     uint32_t* pix2nwalk, float* pix2tfinal, int exp_ulp) {
  gauss_rasterize<float>(w, h, pixxydepth, conic, aabb, opacity, rgb, ts, tile2idx, idx2pnt,
                         pix2rgb, pix2alpha, pix2nwalk, pix2tfinal, nullptr, exp_ulp);
}
double orc_gauss_l2_loss_tiles_f64(uint32_t w, uint32_t h, const double* pixxydepth,
                                   const double* conic, const float* aabb, const double* opacity,
                                   const double* rgb, uint32_t ts, const uint32_t* tile2idx,
                                   const uint32_t* idx2pnt, const uint32_t* frozen,
                                   const float* target, uint32_t tx0, uint32_t tx1, uint32_t ty0,
                                   uint32_t ty1) {
  return gauss_l2_loss_tiles<double>(w, h, pixxydepth, conic, aabb, opacity, rgb, ts, tile2idx,
                                     idx2pnt, frozen, target, tx0, tx1, ty0, ty1);
}
void orc_gauss_rasterize_f64(uint32_t w, uint32_t h, const double* pixxydepth,
                             const double* conic, const float* aabb, const double* opacity,
                             const double* rgb, uint32_t ts, const uint32_t* tile2idx,
                             const uint32_t* idx2pnt, double* pix2rgb, double* pix2alpha,
                             uint32_t* pix2nwalk, double* pix2tfinal, const uint32_t* frozen) {
  gauss_rasterize<double>(w, h, pixxydepth, conic, aabb, opacity, rgb, ts, tile2idx, idx2pnt,
                          pix2rgb, pix2alpha, pix2nwalk, pix2tfinal, frozen);
}

void orc_gauss_rasterize_bwd_f32(uint32_t w, uint32_t h, const float* pixxydepth,
                                 const float* conic, const float* aabb, const float* opacity,
                                 const float* rgb, uint32_t ts, const uint32_t* tile2idx,
                                 const uint32_t* idx2pnt, const uint32_t* nwalk,
                                 const float* tfinal, const float* d_rgbimg,
                                 const float* d_alphaimg, float* d_pixxy, float* d_conic,
                                 float* d_opacity, float* d_rgb) {
  gauss_rasterize_bwd<float>(w, h, pixxydepth, conic, aabb, opacity, rgb, ts, tile2idx, idx2pnt,
                             nwalk, tfinal, d_rgbimg, d_alphaimg, d_pixxy, d_conic, d_opacity,
                             d_rgb);
}
// Threads of the CPU loops (1 = the sequential parity oracle); returns the
// number set, clamped to [1, omp_get_max_threads()].
int orc_set_threads(int n) {
  const int mx = omp_get_max_threads();
  g_threads = n < 1 ? 1 : (n > mx ? mx : n);
  return g_threads;
}

void orc_gauss_rasterize_bwd_mt_f32(uint32_t n, uint32_t w, uint32_t h, const float* pixxydepth,
                                    const float* conic, const float* aabb, const float* opacity,
                                    const float* rgb, uint32_t ts, const uint32_t* tile2idx,
                                    const uint32_t* idx2pnt, const uint32_t* nwalk,
                                    const float* tfinal, const float* d_rgbimg,
                                    const float* d_alphaimg, float* d_pixxy, float* d_conic,
                                    float* d_opacity, float* d_rgb) {
  gauss_rasterize_bwd_mt<float>(n, w, h, pixxydepth, conic, aabb, opacity, rgb, ts, tile2idx,
                                idx2pnt, nwalk, tfinal, d_rgbimg, d_alphaimg, d_pixxy, d_conic,
                                d_opacity, d_rgb);
}

void orc_gauss_rasterize_bwd_f64(uint32_t w, uint32_t h, const double* pixxydepth,
                                 const double* conic, const float* aabb, const double* opacity,
                                 const double* rgb, uint32_t ts, const uint32_t* tile2idx,
                                 const uint32_t* idx2pnt, const uint32_t* nwalk,
                                 const double* tfinal, const double* d_rgbimg,
                                 const double* d_alphaimg, double* d_pixxy, double* d_conic,
                                 double* d_opacity, double* d_rgb) {
  gauss_rasterize_bwd<double>(w, h, pixxydepth, conic, aabb, opacity, rgb, ts, tile2idx, idx2pnt,
                              nwalk, tfinal, d_rgbimg, d_alphaimg, d_pixxy, d_conic, d_opacity,
                              d_rgb);
}

void orc_gauss_project_bwd_f32(uint32_t n, const float* xyz, const float* quat,
                               const float* scale, const float* m, uint32_t w, uint32_t h,
                               const float* d_pixxy, const float* d_conic, float* d_xyz,
                               float* d_scale, float* d_quat, int log_scale) {
  gauss_project_bwd<float>(n, xyz, quat, scale, m, w, h, d_pixxy, d_conic, d_xyz, d_scale,
                           d_quat, log_scale);
}
void orc_gauss_project_bwd_f64(uint32_t n, const double* xyz, const double* quat,
                               const double* scale, const double* m, uint32_t w, uint32_t h,
                               const double* d_pixxy, const double* d_conic, double* d_xyz,
                               double* d_scale, double* d_quat, int log_scale) {
  gauss_project_bwd<double>(n, xyz, quat, scale, m, w, h, d_pixxy, d_conic, d_xyz, d_scale,
                            d_quat, log_scale);
}

void orc_gauss_render_sort_depth(uint32_t n, const float* pixxydepth, const float* conic,
                                 const float* aabb, const float* opacity, const float* rgb,
                                 uint32_t w, uint32_t h, float* img) {
  gauss_render_sort_depth(n, pixxydepth, conic, aabb, opacity, rgb, w, h, img);
}

void orc_circle_project(uint32_t n, const float* xyz, float radius, const float* m, uint32_t w,
                        uint32_t h, float* pixxydepth, float* pixrad) {
  circle_project(n, xyz, radius, m, w, h, pixxydepth, pixrad);
}

void orc_circle_rasterize(uint32_t w, uint32_t h, const float* pixxydepth, const float* pixrad,
                          const float* rgb, uint32_t ts, const uint32_t* tile2idx,
                          const uint32_t* idx2pnt, float* pix2rgb) {
  circle_rasterize(w, h, pixxydepth, pixrad, rgb, ts, tile2idx, idx2pnt, pix2rgb);
}

void orc_zbuffer(uint32_t n, const float* pixxydepth, const float* pixrad, const float* rgb,
                 uint32_t w, uint32_t h, uint64_t* pix2key, float* pix2rgb) {
  zbuffer(n, pixxydepth, pixrad, rgb, w, h, pix2key, pix2rgb);
}

void orc_render_pix_sort_depth(uint32_t n, const float* pxz, const float* rgb, uint32_t w,
                               uint32_t h, float* img) {
  render_pix_sort_depth(n, pxz, rgb, w, h, img);
}

// Transfer of splat i along a ray of opacities a[0..n) (splat_gaussian2.rs:201-210)
// and its derivative w.r.t. a[j] (:212-243), for the revived known-answer test.
float orc_transfer(const float* a, uint32_t i) {
  float t = 1.f;
  for (uint32_t j = 0; j < i; ++j) t *= 1.f - a[j];
  return t * a[i];
}
float orc_dw_transfer(const float* a, uint32_t i, uint32_t j) {
  if (j > i) return 0.f;
  float d = 1.f;
  for (uint32_t k = 0; k <= i; ++k) {
    if (k == j && k == i) continue;
    if (k == j) d *= -1.f;
    else if (k == i) d *= a[k];
    else d *= 1.f - a[k];
  }
  return d;
}

// Adam (Kingma & Ba 2015, Algorithm 1), elementwise over n parameters.
void orc_adam(uint32_t n, float* param, const float* grad, float* m1, float* m2, float lr,
              float b1, float b2, float eps, uint32_t step) {
  const float bc1 = 1.f - std::pow(b1, float(step));
  const float bc2 = 1.f - std::pow(b2, float(step));
  for (uint32_t i = 0; i < n; ++i) {
    m1[i] = b1 * m1[i] + (1.f - b1) * grad[i];
    m2[i] = b2 * m2[i] + (1.f - b2) * grad[i] * grad[i];
    const float mh = m1[i] / bc1, vh = m2[i] / bc2;
    param[i] -= lr * mh / (std::sqrt(vh) + eps);
  }
}

void orc_sh_basis_f64(double x, double y, double z, uint32_t deg, double* b, double* grad) {
  sh_basis(x, y, z, deg, b);
  if (grad) sh_basis_grad(x, y, z, deg, grad);
}

void orc_sh_to_rgb_f32(uint32_t n, uint32_t deg, uint32_t K, const float* cam, const float* xyz,
                       const float* sh, float* rgb) {
  sh_to_rgb(n, deg, K, cam, xyz, sh, rgb);
}

void orc_sh_to_rgb_f64(uint32_t n, uint32_t deg, uint32_t K, const double* cam,
                       const double* xyz, const double* sh, double* rgb) {
  sh_to_rgb(n, deg, K, cam, xyz, sh, rgb);
}

void orc_sh_to_rgb_bwd_f32(uint32_t n, uint32_t deg, uint32_t K, const float* cam,
                           const float* xyz, const float* sh, const float* d_rgb, float* d_sh,
                           float* d_xyz, int overwrite) {
  sh_to_rgb_bwd(n, deg, K, cam, xyz, sh, d_rgb, d_sh, d_xyz, overwrite != 0);
}

void orc_sh_to_rgb_bwd_f64(uint32_t n, uint32_t deg, uint32_t K, const double* cam,
                           const double* xyz, const double* sh, const double* d_rgb,
                           double* d_sh, double* d_xyz, int overwrite) {
  sh_to_rgb_bwd(n, deg, K, cam, xyz, sh, d_rgb, d_sh, d_xyz, overwrite != 0);
}

}  // extern "C"
