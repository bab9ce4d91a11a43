// Shared device/host helpers of the dsplat HIP library (gfx950, wave64).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/dsplat.h"

#define DSPLAT_HIP(call)                                         \
  do {                                                           \
    hipError_t _e = (call);                                      \
    if (_e != hipSuccess) return dsplat_set_hip_error(_e, #call); \
  } while (0)

// Records the HIP error text for dsplat_last_error() and returns
// DSPLAT_ERR_HIP.
int dsplat_set_hip_error(hipError_t e, const char* what);
int dsplat_set_error(int code, const char* msg);

namespace dsplat {

constexpr int kWave = 64;

__host__ __device__ inline uint32_t div_up(uint64_t a, uint64_t b) {
  return uint32_t((a + b - 1) / b);
}

// Tile grid of an image (ceil, as splat_gauss.rs:59-62).
inline void tile_grid(uint32_t w, uint32_t h, uint32_t ts, uint32_t* tw, uint32_t* th) {
  *tw = (w + ts - 1) / ts;
  *th = (h + ts - 1) / ts;
}

inline hipStream_t as_stream(void* s) { return reinterpret_cast<hipStream_t>(s); }

// Number of set bits of `mask` in lanes below the calling lane.
__device__ inline uint32_t lanes_below(uint64_t mask) {
  return __builtin_amdgcn_mbcnt_hi(uint32_t(mask >> 32),
                                   __builtin_amdgcn_mbcnt_lo(uint32_t(mask), 0u));
}

__device__ inline uint32_t f32_bits(float v) { return __float_as_uint(v); }

// Depth part of the binning key: bits of max(ndc_z + 1, 0)
// (tile_acceleration.h:43-46). For non-negative floats the bit pattern
// orders like the value.
__device__ inline uint32_t depth_key(float ndcz) {
  float zp1 = ndcz + 1.f;
  if (zp1 <= 0.f) zp1 = 0.f;
  return __float_as_uint(zp1);
}

// Half-open tile rectangle of a pixel box, clipped to the tile grid: tiles
// floor(x0 / ts) .. floor(x1 / ts) along each axis, as the loops of
// tile_acceleration.h:27-44 run them (a finite inverted box inside one tile
// row or column covers that tile; one spanning a boundary covers none). A
// NaN coordinate covers no tile.
struct TileRect {
  int x0, x1, y0, y1;
  __device__ int count() const { return (x1 - x0) * (y1 - y0); }
};

__device__ inline int tile_floor(float v, float tsf, uint32_t lim) {
  float f = floorf(v / tsf);
  f = fmaxf(f, -1.f);
  f = fminf(f, float(lim));
  return int(f);
}

__device__ inline TileRect tile_rect(float x0, float y0, float x1, float y1, uint32_t ts,
                                     uint32_t tw, uint32_t th) {
  TileRect r{0, 0, 0, 0};
  if (x0 != x0 || y0 != y0 || x1 != x1 || y1 != y1) return r;
  const float tsf = float(ts);
  r.x0 = max(tile_floor(x0, tsf, tw), 0);
  r.x1 = min(tile_floor(x1, tsf, tw) + 1, int(tw));
  r.y0 = max(tile_floor(y0, tsf, th), 0);
  r.y1 = min(tile_floor(y1, tsf, th) + 1, int(th));
  if (r.x1 < r.x0) r.x1 = r.x0;
  if (r.y1 < r.y0) r.y1 = r.y0;
  return r;
}

// A tile rectangle packed into 8 bytes, x0 | y0 << 16 and x1 | y1 << 16
// (tile coordinates below 65536): what the depth sort carries along with a
// Gaussian so that the pair expansion reads rectangles in depth order
// sequentially instead of gathering boxes.
__device__ inline uint2 pack_rect(const TileRect& r) {
  return make_uint2(uint32_t(r.x0) | (uint32_t(r.y0) << 16), uint32_t(r.x1) | (uint32_t(r.y1) << 16));
}
__device__ inline TileRect unpack_rect(uint2 p) {
  return TileRect{int(p.x & 0xFFFFu), int(p.y & 0xFFFFu), int(p.x >> 16), int(p.y >> 16)};
}
__device__ inline uint32_t rect_area(uint2 p) {
  return ((p.y & 0xFFFFu) - (p.x & 0xFFFFu)) * ((p.y >> 16) - (p.x >> 16));
}

// tile2idx[t] = first sorted position whose tile id is >= t, t in [0, T],
// from k sorted words (16-byte aligned) whose tile id is word >> ib, for the
// four positions [4q, 4q + 4) of thread q (the position k stands for tile
// T): one 16-byte load and the previous word, and a store for every tile a
// boundary passes.
__device__ inline void tile_ranges4(uint32_t q, uint32_t k, uint32_t num_tile,
                                    const uint32_t* __restrict__ tile_sorted,
                                    uint32_t* __restrict__ tile2idx, uint32_t ib = 0) {
  const uint32_t i0 = 4u * q;
  if (i0 > k) return;
  uint32_t t[4];
  if (i0 + 4u <= k) {
    const uint4 v = reinterpret_cast<const uint4*>(tile_sorted)[q];
    t[0] = v.x >> ib; t[1] = v.y >> ib; t[2] = v.z >> ib; t[3] = v.w >> ib;
  } else {
#pragma unroll
    for (int e = 0; e < 4; ++e) t[e] = i0 + e < k ? tile_sorted[i0 + e] >> ib : num_tile;
  }
  uint32_t lo = i0 == 0 ? 0u : (tile_sorted[i0 - 1] >> ib) + 1u;
#pragma unroll
  for (int e = 0; e < 4; ++e) {
    const uint32_t idx = i0 + uint32_t(e);
    if (idx > k) break;
    const uint32_t hi = idx == k ? num_tile : t[e];
    for (uint32_t tt = lo; tt <= hi; ++tt) tile2idx[tt] = idx;
    lo = hi + 1u;
  }
}

// The calling wave's LDS writes are visible to its following LDS reads.
__device__ inline void wave_lds_fence() {
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Per-wave LDS slice of wave_expand_pairs. An item's fields are packed so
// that a slot reads its item with one 16-byte and one 8-byte LDS load.
struct ExpandLds {
  uint32_t own[128];  // per slot of the current chunk: 1 + lane whose range starts there
  uint4 rec[64];     // off, x0 | y0 << 16, rectangle height, value
  uint2 tt[64];      // tile id base, tile grid width
};

// Inclusive prefix max over the 64 lanes of a wave (row shifts by 1, 2, 4,
// 8, then row broadcasts 15 and 31; lanes shifted in from outside read 0).
__device__ inline uint32_t wave_incl_max_u32(uint32_t v) {
  v = max(v, uint32_t(__builtin_amdgcn_update_dpp(0, int(v), 0x111, 0xF, 0xF, true)));
  v = max(v, uint32_t(__builtin_amdgcn_update_dpp(0, int(v), 0x112, 0xF, 0xF, true)));
  v = max(v, uint32_t(__builtin_amdgcn_update_dpp(0, int(v), 0x114, 0xF, 0xF, true)));
  v = max(v, uint32_t(__builtin_amdgcn_update_dpp(0, int(v), 0x118, 0xF, 0xF, true)));
  v = max(v, uint32_t(__builtin_amdgcn_update_dpp(0, int(v), 0x142, 0xA, 0xF, false)));
  v = max(v, uint32_t(__builtin_amdgcn_update_dpp(0, int(v), 0x143, 0xC, 0xF, false)));
  return v;
}

// Inclusive prefix sum over the 64 lanes of a wave (same shifts and row
// broadcasts as wave_incl_max_u32).
__device__ inline uint32_t wave_incl_sum_u32(uint32_t v) {
  v += uint32_t(__builtin_amdgcn_update_dpp(0, int(v), 0x111, 0xF, 0xF, true));
  v += uint32_t(__builtin_amdgcn_update_dpp(0, int(v), 0x112, 0xF, 0xF, true));
  v += uint32_t(__builtin_amdgcn_update_dpp(0, int(v), 0x114, 0xF, 0xF, true));
  v += uint32_t(__builtin_amdgcn_update_dpp(0, int(v), 0x118, 0xF, 0xF, true));
  v += uint32_t(__builtin_amdgcn_update_dpp(0, int(v), 0x142, 0xA, 0xF, false));
  v += uint32_t(__builtin_amdgcn_update_dpp(0, int(v), 0x143, 0xC, 0xF, false));
  return v;
}

// Writes the (tile, value) pairs of the calling wave's 64 consecutive items:
// item l owns output slots [off, off + count) and covers the tile rectangle
// r (ix-outer, iy-inner, tile id = tbase + iy * tw + ix). Items are valid on
// a prefix of the lanes with non-decreasing offsets. The wave walks its slot
// range in chunks of 128 (lane l takes slots 2l and 2l + 1, so the stores
// cover the chunk contiguously); the item of a slot is the last one starting
// at or before it: each item with a non-empty range marks its first slot,
// and a prefix max over the chunk (carried across chunks) spreads the marks.
// val_out null: one packed word per pair, tile << ib | value, into tile_out
// (values below 2^ib).
__device__ inline void wave_expand_pairs(bool valid, uint32_t off, TileRect r, uint32_t tbase,
                                         uint32_t tw, uint32_t val, ExpandLds& L,
                                         uint32_t* __restrict__ tile_out,
                                         uint32_t* __restrict__ val_out, uint32_t ib = 0) {
  const int lane = threadIdx.x & 63;
  const uint64_t vmask = __ballot(valid);
  if (vmask == 0) return;
  const int nvalid = __popcll(vmask);
  const uint32_t cnt = valid ? uint32_t(r.count()) : 0u;
  // tile coordinates are below 65536 (the callers check the grid)
  L.rec[lane] = make_uint4(off, uint32_t(r.x0) | (uint32_t(r.y0) << 16), uint32_t(r.y1 - r.y0),
                           val);
  L.tt[lane] = make_uint2(tbase, tw);
  const uint32_t base = __builtin_amdgcn_readfirstlane(off);
  const uint32_t end = __builtin_amdgcn_readlane(off + cnt, nvalid - 1);
  uint32_t carry = 0;  // 1 + item owning the slot before the chunk
  for (uint32_t c0 = base; c0 < end; c0 += 128) {
    reinterpret_cast<uint2*>(L.own)[lane] = make_uint2(0u, 0u);
    wave_lds_fence();
    if (cnt > 0 && off >= c0 && off - c0 < 128u) L.own[off - c0] = uint32_t(lane) + 1u;
    wave_lds_fence();
    const uint2 o = reinterpret_cast<const uint2*>(L.own)[lane];
    const uint32_t o0 = lane == 0 ? max(o.x, carry) : o.x;
    // owner of slot 2l + 1: the prefix max up to lane l; of slot 2l: the
    // prefix max up to lane l - 1, with slot 2l's own mark
    const uint32_t incl = wave_incl_max_u32(max(o0, o.y));
    const uint32_t prev = uint32_t(__shfl_up(int(incl), 1, 64));
    const uint32_t own[2] = {max(lane == 0 ? 0u : prev, o0), incl};
    carry = __builtin_amdgcn_readlane(incl, 63);
#pragma unroll
    for (int k = 0; k < 2; ++k) {
      const uint32_t e = c0 + 2u * uint32_t(lane) + uint32_t(k);
      if (e < end) {
        const uint32_t lo = own[k] - 1u;
        const uint4 q = L.rec[lo];
        const uint2 t2 = L.tt[lo];
        const uint32_t c = e - q.x;
        const uint32_t h = q.z;
        // c / h via the float reciprocal, then corrected to the exact
        // quotient (off by at most one for c < 2^22; the loops make it exact)
        uint32_t qx = uint32_t(float(c) * __builtin_amdgcn_rcpf(float(h)));
        while (qx * h > c) --qx;
        while ((qx + 1u) * h <= c) ++qx;
        const uint32_t ix = (q.y & 0xFFFFu) + qx, iy = (q.y >> 16) + (c - qx * h);
        const uint32_t tile = t2.x + iy * t2.y + ix;
        if (val_out) {
          tile_out[e] = tile;
          val_out[e] = q.w;
        } else {
          tile_out[e] = (tile << ib) | q.w;
        }
      }
    }
  }
  __builtin_amdgcn_wave_barrier();
}

// Packed 2D splat record used by the raster kernels (one 64-byte line per
// Gaussian): pixel centre, conic [a b c], opacity, rgb, pixel box.
struct __align__(16) Splat2 {
  float x, y, ca, cb;
  float cc, op, r, g;
  float b, x0, y0, x1;
  float y1, pad0, pad1, pad2;
};
static_assert(sizeof(Splat2) == 64, "Splat2 must be one 64-byte line");

}  // namespace dsplat
