// Circle (sphere) splatting and point / circle z-buffers.
//
// Circle projection: splat3_to_splat2 of splat_circle.cu:7-41 (CPU twin
// splat_sphere.rs:51-87): pixel centre ((x+1)W/2, (1-y)H/2), ndc z, and the
// pixel radius 1/2 (W/2 r / |col0(J^-1)| + H/2 r / |col1(J^-1)|) with J the
// ndc Jacobian of the world point.
// Circle raster: rasterize_splat_using_tile of splat_circle.cu:93-132, a
// painter's algorithm over the tile list in ascending key order that
// overwrites the pixel when |centre - pixel centre| <= radius. The last
// covering entry wins, so the kernel walks the list from its end and stops at
// the first covering entry.
// Point z-buffer: rasterize_zbuffer of splat_point.cu:49-74 made
// deterministic: every covered pixel keeps the minimum of the 64-bit key
// bits(unit depth) << 32 | point index (atomicMin), then a gather pass writes
// the winner's depth and colour. unit depth = 1 - (ndc_z + 1)/2 evaluated in
// double as the reference does; points with ndc_z outside [-1, 1] are
// clipped; among equal depths the smallest index wins (the reference keeps
// whichever thread arrived first).
// Point painter: render_pix_sort_depth of pnt2pixxyndcz.rs:1-40 -- a stable
// sort by ndc_z + 1 then overwrite, i.e. the maximum of bits(ndc_z + 1) << 32
// | index wins (atomicMax), ndc_z outside (-1, 1) clipped, pixel = the
// saturating float -> unsigned cast of the reference.
#include <float.h>

#include "dsplat_common.hpp"

namespace dsplat {

// ------------------------------------------------------- circle project ---
__global__ void __launch_bounds__(256)
circle_project_kernel(uint32_t n, const float* __restrict__ xyz, float* __restrict__ pixxydepth,
                      float* __restrict__ pixrad, const float* __restrict__ m, uint32_t img_w,
                      uint32_t img_h, float radius) {
#pragma clang fp contract(off)
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float x0 = xyz[3 * i], x1 = xyz[3 * i + 1], x2 = xyz[3 * i + 2];
  float y[4];
#pragma unroll
  for (int k = 0; k < 4; ++k) y[k] = m[k] * x0 + m[k + 4] * x1 + m[k + 8] * x2 + m[k + 12];
  const float w = y[3];
  const float q[3] = {y[0] / w, y[1] / w, y[2] / w};
  pixxydepth[3 * i + 0] = (q[0] + 1.f) * 0.5f * float(img_w);
  pixxydepth[3 * i + 1] = (1.f - q[1]) * 0.5f * float(img_h);
  pixxydepth[3 * i + 2] = q[2];
  float a[9];
#pragma unroll
  for (int j = 0; j < 3; ++j)
#pragma unroll
    for (int r = 0; r < 3; ++r) a[r + 3 * j] = (m[r + 4 * j] - q[r] * m[3 + 4 * j]) / w;
  // column-major 3x3 inverse by cofactors
  const float a00 = a[0], a10 = a[1], a20 = a[2], a01 = a[3], a11 = a[4], a21 = a[5],
              a02 = a[6], a12 = a[7], a22 = a[8];
  const float c00 = a11 * a22 - a12 * a21;
  const float c01 = a12 * a20 - a10 * a22;
  const float c02 = a10 * a21 - a11 * a20;
  const float det = a00 * c00 + a01 * c01 + a02 * c02;
  float rad = 0.f;
  if (det != 0.f) {
    const float id = 1.f / det;
    const float i0 = c00 * id, i1 = c01 * id, i2 = c02 * id;
    const float i3 = (a02 * a21 - a01 * a22) * id;
    const float i4 = (a00 * a22 - a02 * a20) * id;
    const float i5 = (a01 * a20 - a00 * a21) * id;
    const float nx = sqrtf(i0 * i0 + i1 * i1 + i2 * i2);
    const float ny = sqrtf(i3 * i3 + i4 * i4 + i5 * i5);
    const float rx = (1.f / nx) * 0.5f * float(img_w) * radius;
    const float ry = (1.f / ny) * 0.5f * float(img_h) * radius;
    rad = 0.5f * (rx + ry);
  }
  pixrad[i] = rad;
}

// --------------------------------------------------------- circle raster ---
template <int TS>
__global__ void __launch_bounds__(TS * TS)
circle_raster_kernel(uint32_t img_w, uint32_t img_h, uint32_t tw,
                     const uint32_t* __restrict__ tile2idx, const uint32_t* __restrict__ idx2pnt,
                     const float* __restrict__ pixxydepth, const float* __restrict__ pixrad,
                     const float* __restrict__ rgb, float* __restrict__ pix2rgb) {
#pragma clang fp contract(off)
  constexpr int NT = TS * TS;
  __shared__ float4 sc[NT];  // x, y, r, point index bits
  const uint32_t tile = blockIdx.x;
  const uint32_t tx = tile % tw, ty = tile / tw;
  const uint32_t ix = tx * TS + threadIdx.x % TS, iy = ty * TS + threadIdx.x / TS;
  const uint32_t start = tile2idx[tile], end = tile2idx[tile + 1];
  const uint32_t num = end - start;
  const float px = float(ix) + 0.5f, py = float(iy) + 0.5f;
  bool done = !(ix < img_w && iy < img_h);
  uint32_t win = 0xffffffffu;
  for (uint32_t b = 0; b < num; b += NT) {
    const uint32_t it = b + threadIdx.x;
    if (it < num) {
      const uint32_t p = idx2pnt[end - 1 - it];
      sc[threadIdx.x] = make_float4(pixxydepth[3 * p], pixxydepth[3 * p + 1], pixrad[p],
                                    __uint_as_float(p));
    }
    __syncthreads();
    const uint32_t cnt = min(uint32_t(NT), num - b);
    if (!done) {
      for (uint32_t j = 0; j < cnt; ++j) {
        const float4 c = sc[j];
        const float dx = c.x - px, dy = c.y - py;
        if (sqrtf(dx * dx + dy * dy) > c.z) continue;
        win = __float_as_uint(c.w);
        done = true;
        break;
      }
    }
    if (__syncthreads_and(done)) break;
  }
  if (ix < img_w && iy < img_h) {
    const size_t ip = size_t(iy) * img_w + ix;
    if (win != 0xffffffffu) {
      pix2rgb[3 * ip] = rgb[3 * win];
      pix2rgb[3 * ip + 1] = rgb[3 * win + 1];
      pix2rgb[3 * ip + 2] = rgb[3 * win + 2];
    } else {
      pix2rgb[3 * ip] = pix2rgb[3 * ip + 1] = pix2rgb[3 * ip + 2] = 0.f;
    }
  }
}

// ------------------------------------------------------------- z-buffers ---
__global__ void __launch_bounds__(256)
fill_u64_kernel(size_t n, unsigned long long* __restrict__ p, unsigned long long v) {
  for (size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n;
       i += size_t(gridDim.x) * blockDim.x)
    p[i] = v;
}

__device__ inline bool unit_depth_key(float ndcz, uint32_t i, unsigned long long* key) {
  if (ndcz < -1.f || ndcz > 1.f) return false;
  const float u = float(1.0 - (double(ndcz) + 1.0) * 0.5);
  *key = (static_cast<unsigned long long>(__float_as_uint(u)) << 32) | i;
  return true;
}

// Keys only decrease, so a key that is not below the pixel's current one
// skips the atomic: most do, since a pixel covered m times in random depth
// order sees about ln m successful updates. The read is a plain cached load:
// a stale value is never below the true one (it only lets a losing key
// through to the atomic, which then changes nothing).
__device__ inline void zbuffer_min(unsigned long long* p, unsigned long long key) {
  if (key < *p) atomicMin(p, key);
}

__global__ void __launch_bounds__(256)
point_zbuffer_kernel(uint32_t n, const float* __restrict__ pxz, uint32_t img_w, uint32_t img_h,
                     unsigned long long* __restrict__ keys) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  unsigned long long key;
  if (!unit_depth_key(pxz[3 * i + 2], i, &key)) return;
  const float fx = floorf(pxz[3 * i]), fy = floorf(pxz[3 * i + 1]);
  if (!(fx >= 0.f && fx < float(img_w) && fy >= 0.f && fy < float(img_h))) return;
  zbuffer_min(&keys[size_t(fy) * img_w + size_t(fx)], key);
}

// Per-wave LDS of circle_zbuffer_kernel: the 64 points' boxes, distance
// bounds and keys, three 16-byte records each.
struct ZbufLds {
  uint32_t own[64];  // per slot of the current chunk: 1 + lane whose range starts there
  uint4 box[64];     // first slot, box width, x0, y0
  float4 geo[64];    // x, y, d2 bounds (inside below lo, outside from hi)
  uint4 key[64];     // key lo, key hi, bits of r
};

// The distance test sqrtf(d2) <= r of splat_circle.cu:120-125 without the
// square root where d2 is clearly on one side: with rr = r*r rounded, d2 <=
// rr (1 - 2^-20) means d2 < r^2, so the rounded root is <= r; d2 >= rr (1 +
// 2^-20) means the root exceeds r by more than half an ulp of r, so the
// rounded root is > r. Between the two (and for infinite or tiny bounds)
// the exact test decides.
__device__ inline float2 circle_bounds(float r) {
#pragma clang fp contract(off)
  const float rr = r * r;
  if (!(rr >= 1e-30f) || rr == __builtin_inff()) return make_float2(-1.f, __builtin_inff());
  return make_float2(rr * 0.99999904632568359375f, rr * 1.00000095367431640625f);
}

__device__ inline bool circle_covers(float d2, float lo, float hi, float r) {
  if (d2 <= lo) return true;
  if (d2 >= hi && hi != __builtin_inff()) return false;
  return sqrtf(d2) <= r;
}

// Every pixel whose centre lies within the radius: centres ix + .5 in
// [x - r, x + r], then the exact distance test (circle_covers).
// Each wave takes 64 points and deals the pixels of their clipped boxes out
// 64 at a time (row-major within a box): the lanes stay busy whatever the
// radii, and neighbouring lanes touch neighbouring pixels. A slot's point is
// the last one whose range starts at or before it (each point marks its
// first slot; a prefix max spreads the marks, carried across chunks).
__global__ void __launch_bounds__(256)
circle_zbuffer_kernel(uint32_t n, const float* __restrict__ pxz, const float* __restrict__ pixrad,
                      uint32_t img_w, uint32_t img_h, unsigned long long* __restrict__ keys) {
#pragma clang fp contract(off)
  __shared__ ZbufLds lds[4];
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if ((i & ~63u) >= n) return;  // wave-uniform
  const int lane = threadIdx.x & 63;
  ZbufLds& L = lds[threadIdx.x >> 6];
  unsigned long long key = 0;
  uint32_t cnt = 0, bw = 0;
  int x0 = 0, y0 = 0;
  float x = 0.f, y = 0.f, r = 0.f;
  if (i < n && unit_depth_key(pxz[3 * i + 2], i, &key)) {
    x = pxz[3 * i];
    y = pxz[3 * i + 1];
    r = pixrad[i];
    if (r >= 0.f && x == x && y == y) {
      const float fx0 = ceilf(x - r - 0.5f), fx1 = floorf(x + r - 0.5f);
      const float fy0 = ceilf(y - r - 0.5f), fy1 = floorf(y + r - 0.5f);
      // clamped to the image in float before the conversion (a box off the
      // image comes out empty)
      x0 = int(fminf(fmaxf(fx0, 0.f), float(img_w)));
      y0 = int(fminf(fmaxf(fy0, 0.f), float(img_h)));
      const int x1 = int(fmaxf(fminf(fx1, float(img_w) - 1.f), -1.f));
      const int y1 = int(fmaxf(fminf(fy1, float(img_h) - 1.f), -1.f));
      if (x1 >= x0 && y1 >= y0) {
        bw = uint32_t(x1 - x0 + 1);
        cnt = bw * uint32_t(y1 - y0 + 1);
      }
    }
  }
  const uint32_t inc = wave_incl_sum_u32(cnt);
  const uint32_t off = inc - cnt;
  const uint32_t total = __builtin_amdgcn_readlane(inc, 63);
  const float2 bnd = circle_bounds(r);
  L.box[lane] = make_uint4(off, bw, uint32_t(x0), uint32_t(y0));
  L.geo[lane] = make_float4(x, y, bnd.x, bnd.y);
  L.key[lane] = make_uint4(uint32_t(key), uint32_t(key >> 32), __float_as_uint(r), 0u);
  uint32_t carry = 0;
  for (uint32_t c0 = 0; c0 < total; c0 += 64) {
    L.own[lane] = 0;
    wave_lds_fence();
    if (cnt > 0 && off >= c0 && off - c0 < 64u) L.own[off - c0] = uint32_t(lane) + 1u;
    wave_lds_fence();
    const uint32_t own = wave_incl_max_u32(max(L.own[lane], lane == 0 ? carry : 0u));
    carry = __builtin_amdgcn_readlane(own, 63);
    const uint32_t e = c0 + uint32_t(lane);
    if (e < total) {
      const uint32_t lo = own - 1u;
      const uint4 bx = L.box[lo];
      const float4 g = L.geo[lo];
      const uint32_t c = e - bx.x, w = bx.y;
      // c / w via the float reciprocal, corrected to the exact quotient
      // (c < 2^22: off by at most one)
      uint32_t q = uint32_t(float(c) * __builtin_amdgcn_rcpf(float(w)));
      while (q * w > c) --q;
      while ((q + 1u) * w <= c) ++q;
      const int iy = int(bx.w) + int(q), ix = int(bx.z) + int(c - q * w);
      const float dx = g.x - (float(ix) + 0.5f), dy = g.y - (float(iy) + 0.5f);
      const float d2 = dx * dx + dy * dy;
      const uint4 k = L.key[lo];
      if (circle_covers(d2, g.z, g.w, __uint_as_float(k.z)))
        zbuffer_min(&keys[size_t(iy) * img_w + ix],
                    (static_cast<unsigned long long>(k.y) << 32) | k.x);
    }
    // the next chunk rewrites own[] only after every lane has read it
    wave_lds_fence();
  }
}

__global__ void __launch_bounds__(256)
zbuffer_resolve_kernel(size_t np, const unsigned long long* __restrict__ keys,
                       const float* __restrict__ rgb, float* __restrict__ pix2unitdepth,
                       float* __restrict__ pix2rgb) {
  for (size_t k = size_t(blockIdx.x) * blockDim.x + threadIdx.x; k < np;
       k += size_t(gridDim.x) * blockDim.x) {
    const unsigned long long key = keys[k];
    if (key == ~0ull) {
      if (pix2unitdepth) pix2unitdepth[k] = FLT_MAX;
      pix2rgb[3 * k] = pix2rgb[3 * k + 1] = pix2rgb[3 * k + 2] = 0.f;
      continue;
    }
    const uint32_t p = uint32_t(key & 0xffffffffull);
    if (pix2unitdepth) pix2unitdepth[k] = __uint_as_float(uint32_t(key >> 32));
    pix2rgb[3 * k] = rgb[3 * size_t(p)];
    pix2rgb[3 * k + 1] = rgb[3 * size_t(p) + 1];
    pix2rgb[3 * k + 2] = rgb[3 * size_t(p) + 2];
  }
}

// Painter semantics of render_pix_sort_depth (pnt2pixxyndcz.rs:1-40).
__global__ void __launch_bounds__(256)
point_painter_kernel(uint32_t n, const float* __restrict__ pxz, uint32_t img_w, uint32_t img_h,
                     unsigned long long* __restrict__ keys) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float z = pxz[3 * i + 2];
  if (z <= -1.f || z >= 1.f || z != z) return;
  // Rust `f32 as usize`: truncation, negative and NaN saturate to 0
  const float x = pxz[3 * i], y = pxz[3 * i + 1];
  const float fx = x > 0.f ? truncf(x) : 0.f;
  const float fy = y > 0.f ? truncf(y) : 0.f;
  if (fx >= float(img_w) || fy >= float(img_h)) return;
  const unsigned long long key =
      (static_cast<unsigned long long>(__float_as_uint(z + 1.f)) << 32) | i;
  atomicMax(&keys[size_t(fy) * img_w + size_t(fx)], key);
}

__global__ void __launch_bounds__(256)
painter_resolve_kernel(size_t np, const unsigned long long* __restrict__ keys,
                       const float* __restrict__ rgb, float* __restrict__ pix2rgb) {
  for (size_t k = size_t(blockIdx.x) * blockDim.x + threadIdx.x; k < np;
       k += size_t(gridDim.x) * blockDim.x) {
    const unsigned long long key = keys[k];
    if (key == 0ull) {
      pix2rgb[3 * k] = pix2rgb[3 * k + 1] = pix2rgb[3 * k + 2] = 0.f;
      continue;
    }
    const uint32_t p = uint32_t(key & 0xffffffffull);
    pix2rgb[3 * k] = rgb[3 * size_t(p)];
    pix2rgb[3 * k + 1] = rgb[3 * size_t(p) + 1];
    pix2rgb[3 * k + 2] = rgb[3 * size_t(p) + 2];
  }
}

inline uint32_t grid_for(size_t n) {
  size_t b = (n + 255) / 256;
  return uint32_t(b > 4096 ? 4096 : (b ? b : 1));
}

}  // namespace dsplat

using namespace dsplat;

extern "C" int dsplat_circle_pnt2splat3_to_pnt2splat2(void* stream, uint32_t num_pnt,
                                                      const float* pnt2xyz,
                                                      float* pnt2pixxydepth, float* pnt2pixrad,
                                                      const float* transform_world2ndc,
                                                      uint32_t img_w, uint32_t img_h,
                                                      float radius) {
  if (num_pnt == 0) return DSPLAT_OK;
  if (!pnt2xyz || !pnt2pixxydepth || !pnt2pixrad || !transform_world2ndc)
    return dsplat_set_error(DSPLAT_ERR_ARG, "circle projection: null buffer");
  circle_project_kernel<<<div_up(num_pnt, 256), 256, 0, as_stream(stream)>>>(
      num_pnt, pnt2xyz, pnt2pixxydepth, pnt2pixrad, transform_world2ndc, img_w, img_h, radius);
  DSPLAT_HIP(hipGetLastError());
  return DSPLAT_OK;
}

extern "C" int dsplat_circle_splat(void* stream, uint32_t img_w, uint32_t img_h, float* pix2rgb,
                                   const float* pnt2pixxydepth, const float* pnt2pixrad,
                                   const float* pnt2rgb, uint32_t tile_size,
                                   const uint32_t* tile2idx, const uint32_t* idx2pnt) {
  if (!pix2rgb || !tile2idx || !pnt2pixxydepth || !pnt2pixrad || !pnt2rgb)
    return dsplat_set_error(DSPLAT_ERR_ARG, "circle splat: null buffer");
  uint32_t tw, th;
  tile_grid(img_w, img_h, tile_size, &tw, &th);
  const uint32_t nt = tw * th;
  if (nt == 0) return DSPLAT_OK;
  hipStream_t s = as_stream(stream);
  if (tile_size == 8)
    circle_raster_kernel<8><<<nt, 64, 0, s>>>(img_w, img_h, tw, tile2idx, idx2pnt, pnt2pixxydepth, pnt2pixrad, pnt2rgb, pix2rgb);
  else if (tile_size == 16)
    circle_raster_kernel<16><<<nt, 256, 0, s>>>(img_w, img_h, tw, tile2idx, idx2pnt, pnt2pixxydepth, pnt2pixrad, pnt2rgb, pix2rgb);
  else if (tile_size == 32)
    circle_raster_kernel<32><<<nt, 1024, 0, s>>>(img_w, img_h, tw, tile2idx, idx2pnt, pnt2pixxydepth, pnt2pixrad, pnt2rgb, pix2rgb);
  else
    return dsplat_set_error(DSPLAT_ERR_ARG, "tile_size must be 8, 16 or 32");
  DSPLAT_HIP(hipGetLastError());
  return DSPLAT_OK;
}

static int zbuffer_common(void* stream, uint32_t num_pnt, const float* pxz, const float* pixrad,
                          const float* rgb, unsigned long long* keys, float* pix2unitdepth,
                          float* pix2rgb, uint32_t w, uint32_t h) {
  if (!pxz || !rgb || !keys || !pix2rgb)
    return dsplat_set_error(DSPLAT_ERR_ARG, "zbuffer: null buffer");
  hipStream_t s = as_stream(stream);
  const size_t np = size_t(w) * h;
  if (np == 0) return DSPLAT_OK;
  fill_u64_kernel<<<grid_for(np), 256, 0, s>>>(np, keys, ~0ull);
  if (num_pnt) {
    if (pixrad)
      circle_zbuffer_kernel<<<div_up(num_pnt, 256), 256, 0, s>>>(num_pnt, pxz, pixrad, w, h, keys);
    else
      point_zbuffer_kernel<<<div_up(num_pnt, 256), 256, 0, s>>>(num_pnt, pxz, w, h, keys);
  }
  zbuffer_resolve_kernel<<<grid_for(np), 256, 0, s>>>(np, keys, rgb, pix2unitdepth, pix2rgb);
  DSPLAT_HIP(hipGetLastError());
  return DSPLAT_OK;
}

extern "C" int dsplat_point_rasterize_zbuffer(void* stream, uint32_t num_pnt,
                                              const float* pnt2pixxyndcz, const float* pnt2rgb,
                                              float* pix2unitdepth, float* pix2rgb,
                                              uint32_t img_w, uint32_t img_h,
                                              uint64_t* pix2key) {
  return zbuffer_common(stream, num_pnt, pnt2pixxyndcz, nullptr, pnt2rgb,
                        reinterpret_cast<unsigned long long*>(pix2key), pix2unitdepth, pix2rgb,
                        img_w, img_h);
}

extern "C" int dsplat_circle_rasterize_zbuffer(void* stream, uint32_t num_pnt,
                                               const float* pnt2pixxyndcz,
                                               const float* pnt2pixrad, const float* pnt2rgb,
                                               float* pix2unitdepth, float* pix2rgb,
                                               uint32_t img_w, uint32_t img_h,
                                               uint64_t* pix2key) {
  if (num_pnt && !pnt2pixrad) return dsplat_set_error(DSPLAT_ERR_ARG, "zbuffer: null radius");
  return zbuffer_common(stream, num_pnt, pnt2pixxyndcz, pnt2pixrad ? pnt2pixrad : pnt2pixxyndcz,
                        pnt2rgb, reinterpret_cast<unsigned long long*>(pix2key), pix2unitdepth,
                        pix2rgb, img_w, img_h);
}

extern "C" int dsplat_point_rasterize_pix(void* stream, uint32_t num_pnt,
                                          const float* pnt2pixxyndcz, const float* pnt2rgb,
                                          float* pix2rgb, uint32_t img_w, uint32_t img_h,
                                          uint64_t* pix2key) {
  if (!pnt2pixxyndcz || !pnt2rgb || !pix2rgb || !pix2key)
    return dsplat_set_error(DSPLAT_ERR_ARG, "rasterize_pix: null buffer");
  hipStream_t s = as_stream(stream);
  const size_t np = size_t(img_w) * img_h;
  if (np == 0) return DSPLAT_OK;
  unsigned long long* keys = reinterpret_cast<unsigned long long*>(pix2key);
  fill_u64_kernel<<<grid_for(np), 256, 0, s>>>(np, keys, 0ull);
  if (num_pnt)
    point_painter_kernel<<<div_up(num_pnt, 256), 256, 0, s>>>(num_pnt, pnt2pixxyndcz, img_w, img_h, keys);
  painter_resolve_kernel<<<grid_for(np), 256, 0, s>>>(np, keys, pnt2rgb, pix2rgb);
  DSPLAT_HIP(hipGetLastError());
  return DSPLAT_OK;
}
