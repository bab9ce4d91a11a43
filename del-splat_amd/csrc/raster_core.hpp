// Per-tile bodies of the Gaussian raster and its adjoint, shared by the
// single-view kernels (one workgroup per tile) and the persistent multi-view
// kernels (workgroups pull (view, tile) items from a queue).
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
// Records of NT list entries are staged in LDS as float4 quarters (every lane
// reads the same address: broadcast), every thread walks them for its pixels,
// and the workgroup leaves the tile as soon as every pixel has terminated
// (__syncthreads_and).
//
// State for the adjoint: nwalk = 1 + list position of the last splat that
// contributed to the pixel (the terminating one if the walk terminated, 0 if
// none) and tfinal = the pixel's transmittance after its walk (behind that
// splat; 1 if none).
//
// Adjoint: every pixel walks its list back-to-front from position nwalk - 1
// with the transmittance recovered by division from tfinal (dividing by 1 - e
// of earlier splats is safe because the walk did not stop there, so T stayed
// above ~1e-3). Per list entry the nine gradient terms of a wave's pixels
// are summed across the wave with DPP, across waves through per-wave LDS
// slots, and once per (tile, entry) into global memory.
#pragma once
#include "dsplat_common.hpp"

namespace dsplat {

// Packed records, 4 x float4 per item: (x, y, a, b), (c, opacity, r, g),
// (b, x0, y0, x1), (y1, -, -, -).
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

// The splat weight exp(-q/2), q = d^T C d for d = (dx, dy) and C = [[a, b],
// [b, c]]. q is rounded exactly as the reference's CPU evaluation writes it,
// a dx dx + b (dx dy + dy dx) + c dy dy left to right with every product
// and sum rounded (splat_gauss.cu:97 / pnt2pixxyndcz.rs:85, contraction
// off): b (dx dy + dy dx) is formed as (2b) (dx dy), which rounds the same
// (dx dy + dy dx = 2 (dx dy) exactly, and scaling by 2 commutes with
// rounding). So q carries the same bits as the oracle's even where its terms
// cancel (flat, elongated splats: q is then a small difference of large
// terms, and any other order moves it), and only the exponential differs:
// 2^(q * -log2(e)/2), a few ulps from expf(-q/2). Shared by the forward and
// the adjoint, so every workgroup shape produces bit-identical pixels and the
// adjoint recomputes exactly the forward's weights.
// Opacity as the walks stage it: at most 1 - 2^-24, the largest float below
// 1 (an opacity of 1 becomes 1 - 2^-24, a relative change of 2^-24; the
// model's sigmoid opacities lie below 1 anyway). Gradients pass through.
__device__ __forceinline__ float stage_opacity(float o) { return fminf(o, 0x1.fffffep-1f); }

struct ConicQ {
  float a, b2, c;  // a, 2b, c
};

__device__ __forceinline__ ConicQ conic_q(float a, float b, float c) {
  return {a, b + b, c};
}

// exp(-q/2) where in, else 0 (2^-inf = 0: a select on the argument keeps
// the exp unpredicated), clamped to [0, 1] by the exp's own output clamp (no
// extra instruction): q >= 0 for a positive-definite conic, and the clamp
// only removes rounding's 1 + ulp where q cancels to just below 0. With the
// opacity staged at most 1 - 2^-24 this bounds e = g * opacity by 1 - 2^-24
// (rounding is monotone), so 1 - e >= 2^-24 holds for every splat. cyy = (c dy) dy, the last term of q, passed in by
// callers that share it between pixels of one row (the same roundings).
__device__ __forceinline__ float splat_weight_cyy(bool in, const ConicQ& k, float dx, float dy,
                                                  float cyy) {
#pragma clang fp contract(off)
  constexpr float kS = -0.72134752044448170368f;  // -log2(e) / 2
  const float q = ((k.a * dx) * dx + k.b2 * (dx * dy)) + cyy;
  return __builtin_amdgcn_fmed3f(__builtin_amdgcn_exp2f(in ? q * kS : -__builtin_inff()), 0.f, 1.f);
}

__device__ __forceinline__ float splat_weight(bool in, const ConicQ& k, float dx, float dy) {
#pragma clang fp contract(off)
  constexpr float kS = -0.72134752044448170368f;  // -log2(e) / 2
  const float q = ((k.a * dx) * dx + k.b2 * (dx * dy)) + (k.c * dy) * dy;
  return __builtin_amdgcn_fmed3f(__builtin_amdgcn_exp2f(in ? q * kS : -__builtin_inff()), 0.f, 1.f);
}

// Pixel -> thread mapping of a tile (forward and adjoint). The TS x TS tile
// is cut into RS x RS regions (RS = 16 when there are at least as many of
// them as waves, else 8); each of the W = NT/64 waves owns G of them
// ("groups"), and in each group every lane a QS x QS pixel quad (QS = RS/8).
// Box tests per group are wave-uniform, so a wave computes only the groups
// an entry's box reaches (and in the adjoint its partial sums cover all its
// groups before one reduction).
template <int TS, int NT>
struct GroupMap {
  static constexpr int W = NT / 64;
  static constexpr int RS = (TS / 16) * (TS / 16) > W ? 16 : 8;
  static constexpr int QS = RS / 8;
  static constexpr int Q = QS * QS;
  static constexpr int RXN = TS / RS;
  static constexpr int G = RXN * RXN / W;
  static constexpr int PPT = G * Q;
  // 8 x 8 regions, one pixel per lane: a list entry's box test for a region
  // is a 64-bit lane mask (region_lane_mask) formed in scalar registers
  static constexpr bool kLaneMask = RS == 8;
  // One pixel per lane and a wave's G regions side by side in one row of
  // regions (G divides the row): all of a lane's pixels share their pixel
  // row, so an entry's dy is the same for every group.
  static constexpr bool kRowShared = Q == 1 && G > 1 && RXN % G == 0;
  static_assert(TS % RS == 0 && G >= 1 && G * W == RXN * RXN, "unsupported tile / workgroup shape");
  __device__ static uint32_t wave() { return __builtin_amdgcn_readfirstlane(threadIdx.x >> 6); }
  __device__ static uint32_t rx0(int g) { return ((wave() * G + g) % RXN) * RS; }
  __device__ static uint32_t ry0(int g) { return ((wave() * G + g) / RXN) * RS; }
  __device__ static uint32_t lx(int k) {
    return rx0(k / Q) + (threadIdx.x & 7) * QS + uint32_t((k % Q) % QS);
  }
  __device__ static uint32_t ly(int k) {
    return ry0(k / Q) + ((threadIdx.x & 63) >> 3) * QS + uint32_t((k % Q) / QS);
  }
};

// The tile's regions whose pixel-centre rectangle [cx0, cx1] x [cy0, cy1]
// the box (bx0, by0, bx1, by1) can reach, !(bx1 < cx0 || bx0 > cx1 || by1 <
// cy0 || by0 > cy1), as bits ry * RXN + rx (GroupMap's numbering: group g of
// wave w is region w * G + g). One thread evaluates it per list entry while
// staging the entry (the test is separable into x and y); the waves then
// test a bit instead of comparing the box against their regions in every
// lane.
template <int TS, int NT>
__device__ inline uint32_t region_mask(uint32_t tx, uint32_t ty, float bx0, float by0, float bx1,
                                       float by1) {
  using GM = GroupMap<TS, NT>;
  constexpr int RXN = GM::RXN, RS = GM::RS;
  uint32_t xm = 0, ym = 0;
#pragma unroll
  for (int r = 0; r < RXN; ++r) {
    const float cx0 = float(tx * TS + uint32_t(r * RS)) + 0.5f;
    const float cy0 = float(ty * TS + uint32_t(r * RS)) + 0.5f;
    const float cx1 = cx0 + float(RS - 1), cy1 = cy0 + float(RS - 1);
    xm |= uint32_t(!(bx1 < cx0 || bx0 > cx1)) << r;
    ym |= uint32_t(!(by1 < cy0 || by0 > cy1)) << r;
  }
  uint32_t m = 0;
#pragma unroll
  for (int r = 0; r < RXN; ++r) m |= ((ym >> r) & 1u) ? xm << (r * RXN) : 0u;
  return m;
}

// The columns c in [0, TS) of a tile starting at pixel X0 whose centre
// p_c = float(X0 + c) + 0.5f satisfies b0 <= p_c <= b1, as bits (the same
// float comparisons as the per-pixel box test, answered for a whole row of
// the tile at once). The set is an interval [lo, hi]: lo is the first c
// with p_c >= b0, hi the last with p_c <= b1. Candidates from ceil / floor
// of the offsets are within one of the answer (the offsets are exact or
// within rounding of it wherever the clamp does not decide), and one
// comparison step each way makes them exact. NaN or b0 > b1: no centre.
template <int TS>
__host__ __device__ inline uint32_t axis_mask(uint32_t X0, float b0, float b1) {
  static_assert(TS <= 32, "one bit per column");
  if (!(b0 <= b1)) return 0u;
  const float base = float(X0) + 0.5f;
  int lo = int(ceilf(fminf(fmaxf(b0 - base, -1.f), float(TS))));
  lo = lo < 0 ? 0 : lo;
  if (lo > 0 && float(X0 + uint32_t(lo) - 1u) + 0.5f >= b0) --lo;
  if (lo < TS && !(float(X0 + uint32_t(lo)) + 0.5f >= b0)) ++lo;
  int hi = int(floorf(fminf(fmaxf(b1 - base, -1.f), float(TS))));
  hi = hi > TS - 1 ? TS - 1 : hi;
  if (hi < TS - 1 && float(X0 + uint32_t(hi + 1)) + 0.5f <= b1) ++hi;
  if (hi >= 0 && !(float(X0 + uint32_t(hi)) + 0.5f <= b1)) --hi;
  if (lo > hi) return 0u;
  return ((2u << uint32_t(hi - lo)) - 1u) << uint32_t(lo);
}

// Lanes of the 8 x 8 region at (rx0, ry0) (lane = 8 row + column, GroupMap's
// layout with one pixel per lane) whose pixel lies in the box, from the
// tile's column and row masks: the region's column byte repeated in every
// row whose row bit is set. Row expansion: bit i of a nibble times
// 0x00204081 lands on bit 8i (the four shifted copies do not overlap), and
// times 0xFF fills its byte.
__host__ __device__ inline uint64_t region_lane_mask(uint32_t xm, uint32_t ym, uint32_t rx0,
                                                     uint32_t ry0) {
  const uint32_t xb = (xm >> rx0) & 0xFFu, yb = (ym >> ry0) & 0xFFu;
  const uint32_t xr = xb * 0x01010101u;
  const uint32_t ylo = (((yb & 15u) * 0x00204081u) & 0x01010101u) * 0xFFu;
  const uint32_t yhi = (((yb >> 4) * 0x00204081u) & 0x01010101u) * 0xFFu;
  return (uint64_t(xr & yhi) << 32) | uint64_t(xr & ylo);
}

// Per-entry masks in LDS. Region form: one word per entry (region_mask).
// Lane-mask form: the 64-bit lane mask of every region, stored [region][B]
// (consecutive entries in consecutive slots: the staging threads write
// without bank conflicts, and a wave reads only its own regions).
template <int TS, int NT, int B>
struct EntryMasks {
  using GM = GroupMap<TS, NT>;
  static constexpr int NR = GM::RXN * GM::RXN;
  uint32_t region[GM::kLaneMask ? 1 : B];
  uint64_t lanes[GM::kLaneMask ? NR * B : 1];

  // By the staging thread of entry `slot`, for the box (bx0, by0, bx1, by1).
  __device__ void stage(uint32_t slot, uint32_t tx, uint32_t ty, float bx0, float by0, float bx1,
                        float by1) {
    if constexpr (GM::kLaneMask) {
      const uint32_t xm = axis_mask<TS>(tx * TS, bx0, bx1), ym = axis_mask<TS>(ty * TS, by0, by1);
#pragma unroll
      for (int r = 0; r < NR; ++r)
        lanes[r * B + slot] =
            region_lane_mask(xm, ym, uint32_t(r % GM::RXN) * 8u, uint32_t(r / GM::RXN) * 8u);
    } else {
      region[slot] = region_mask<TS, NT>(tx, ty, bx0, by0, bx1, by1);
    }
  }
};

// Scalar per-entry hit data of a wave: bit g of `hits` when group g can take
// the entry, and with lane masks each group's lanes inside the box.
template <int TS, int NT>
struct EntryHits {
  uint32_t hits;
  uint64_t lanes[GroupMap<TS, NT>::kLaneMask ? GroupMap<TS, NT>::G : 1];
};

// Entry j's hit data for the wave whose first group is region `first`
// (wave * G, formed once outside the entry loops).
template <int TS, int NT, int B>
__device__ __forceinline__ EntryHits<TS, NT> entry_hits(const EntryMasks<TS, NT, B>& m,
                                                        uint32_t j, uint32_t first) {
  using GM = GroupMap<TS, NT>;
  EntryHits<TS, NT> h;
  if constexpr (GM::kLaneMask) {
    h.hits = 0;
#pragma unroll
    for (int g = 0; g < GM::G; ++g) {
      const uint64_t v = m.lanes[(first + uint32_t(g)) * B + j];
      const uint32_t lo = __builtin_amdgcn_readfirstlane(uint32_t(v));
      const uint32_t hi = __builtin_amdgcn_readfirstlane(uint32_t(v >> 32));
      h.lanes[g] = (uint64_t(hi) << 32) | lo;
      h.hits |= uint32_t(h.lanes[g] != 0) << g;
    }
  } else {
    h.hits = (__builtin_amdgcn_readfirstlane(m.region[j]) >> first) & ((1u << GM::G) - 1u);
    h.lanes[0] = 0;
  }
  return h;
}

// The entries j < cnt of a batch whose box reaches one of the wave's groups
// (regions first .. first + G - 1), as a bitmask over the batch: lane l of
// the wave tests entries l and l + 64. The entry loops walk only the set
// bits, so an entry that misses all of the wave's pixels costs it nothing
// (in dense scenes most entries of a tile's list miss most of its waves).
template <int TS, int NT, int B>
__device__ __forceinline__ void wave_entry_mask(const EntryMasks<TS, NT, B>& m, uint32_t cnt,
                                                uint32_t first, uint64_t (&em)[(B + 63) / 64]) {
  using GM = GroupMap<TS, NT>;
  const uint32_t lane = threadIdx.x & 63u;
#pragma unroll
  for (int h = 0; h < (B + 63) / 64; ++h) {
    const uint32_t j = uint32_t(h) * 64u + lane;
    bool reach = false;
    if (j < cnt) {
      if constexpr (GM::kLaneMask) {
#pragma unroll
        for (int g = 0; g < GM::G; ++g) reach = reach || m.lanes[(first + uint32_t(g)) * B + j] != 0;
      } else {
        reach = ((m.region[j] >> first) & ((1u << GM::G) - 1u)) != 0u;
      }
    }
    em[h] = __ballot(reach);
  }
}

// Lane-mask form (one pixel per lane): the lanes of group g that take the
// entry are the walking lanes (a ballot) inside the box (the group's lane
// mask), combined in scalar registers. Returns false when there are none
// (wave-uniform).
template <int TS, int NT>
__device__ __forceinline__ bool lanes_take(const EntryHits<TS, NT>& h, int g, uint64_t walking,
                                           bool (&inm)[GroupMap<TS, NT>::Q]) {
  const uint64_t act = walking & h.lanes[g];
  inm[0] = __builtin_amdgcn_inverse_ballot_w64(act);
  return act != 0;
}

// Box-test form: pixel slot k of group g takes the entry when walking[k]
// and its centre lies in the box q2 / q3 (inclusive, aabb2::is_include_point;
// non-short-circuit &: plain compares and mask ands, no branches).
template <int TS, int NT>
__device__ __forceinline__ bool box_take(int g, const bool (&walking)[GroupMap<TS, NT>::PPT],
                                         const float (&px)[GroupMap<TS, NT>::PPT],
                                         const float (&py)[GroupMap<TS, NT>::PPT],
                                         const float4& q2, const float4& q3,
                                         bool (&inm)[GroupMap<TS, NT>::Q]) {
  using GM = GroupMap<TS, NT>;
  bool any = false;
#pragma unroll
  for (int qq = 0; qq < GM::Q; ++qq) {
    const int k = g * GM::Q + qq;
    inm[qq] = walking[k] & (px[k] >= q2.y) & (px[k] <= q2.w) & (py[k] >= q2.z) &
              (py[k] <= q3.x);
    any = any || inm[qq];
  }
  return __ballot(any) != 0;
}

// Forward of one tile. Image arrays are the view's base pointers (pixel
// index iy * img_w + ix); alpha / nwalk / tfinal may be null. The entry loop is
// outermost so one LDS read of a record serves all PPT pixels of a thread.
// kState: track the adjoint's state (nwalk, tfinal) and the tile cost; a
// forward asked for neither leaves the nwalk select per pixel and entry out.
template <int TS, int NT, typename Loader, bool kState = true>
__device__ __forceinline__ void raster_tile_fwd(uint32_t tx, uint32_t ty, uint32_t img_w,
                                                uint32_t img_h, uint32_t start, uint32_t end,
                                                const Loader& ld,
                                                const uint32_t* __restrict__ idx2pnt,
                                                float* __restrict__ pix2rgb,
                                                float* __restrict__ pix2alpha,
                                                uint32_t* __restrict__ pix2nwalk,
                                                float* __restrict__ pix2tfinal,
                                                uint32_t* __restrict__ tile_cost = nullptr,
                                                uint32_t cost_slot = 0) {
#pragma clang fp contract(off)
  using GM = GroupMap<TS, NT>;
  constexpr int PPT = GM::PPT, G = GM::G, Q = GM::Q;
  // batches of at most 128 list entries: walks end after ~150 entries at
  // typical densities, so larger batches would load records nobody reads
  constexpr int B = NT < 128 ? NT : 128;
  __shared__ float4 s0[B], s1[B], s2[B], s3[B];
  __shared__ EntryMasks<TS, NT, B> smask;
  const uint32_t num = end - start;
  const uint32_t gfirst = GM::wave() * uint32_t(G);
  // A pixel is live while !(alpha_sum >= 0.999f) (NaN stays live, as in the
  // reference's test); pixels outside the image start at alpha_sum = 1.
  float cr[PPT], cg[PPT], cbl[PPT], asum[PPT], tr[PPT], px[PPT], py[PPT];
  uint32_t walked[PPT];
#pragma unroll
  for (int k = 0; k < PPT; ++k) {
    const uint32_t ix = tx * TS + GM::lx(k), iy = ty * TS + GM::ly(k);
    px[k] = float(ix) + 0.5f;
    py[k] = float(iy) + 0.5f;
    cr[k] = cg[k] = cbl[k] = 0.f;
    asum[k] = (ix < img_w && iy < img_h) ? 0.f : 1.f;
    tr[k] = 1.f;
    walked[k] = 0;
  }
  // The staging threads hold the next batch's records in registers, loaded
  // while the current batch is walked, and its list indices one batch
  // earlier still (the record load depends on them).
  const bool stager = threadIdx.x < uint32_t(B);
  float4 n0, n1, n2, n3;
  uint32_t nidx = 0;
  if (stager && threadIdx.x < num) ld.load(idx2pnt[end - 1 - threadIdx.x], n0, n1, n2, n3);
  if (stager && uint32_t(B) + threadIdx.x < num) nidx = idx2pnt[end - 1 - (uint32_t(B) + threadIdx.x)];
  for (uint32_t b = 0; b < num; b += B) {
    const uint32_t it = b + threadIdx.x;
    if (stager && it < num) {
      const float4 q0 = n0, q1 = n1, q2 = n2, q3 = n3;
      // the conic with b doubled once per entry (every wave uses it)
      const ConicQ ke = conic_q(q0.z, q0.w, q1.x);
      s0[threadIdx.x] = make_float4(q0.x, q0.y, ke.a, ke.b2);
      s1[threadIdx.x] = make_float4(ke.c, stage_opacity(q1.y), q1.z, q1.w);
      s2[threadIdx.x] = q2;
      // the record's free second slot carries the entry's walk position + 1
      // (what a pixel it reaches records as nwalk): read from LDS it is a
      // vector operand, so the per-pixel select needs no scalar-to-vector move
      s3[threadIdx.x] = make_float4(q3.x, __uint_as_float(it + 1u), q3.z, q3.w);
      smask.stage(threadIdx.x, tx, ty, q2.y, q2.z, q2.w, q3.x);
    }
    __syncthreads();
    if (stager && it + uint32_t(B) < num) ld.load(nidx, n0, n1, n2, n3);
    if (stager && it + uint32_t(2 * B) < num) nidx = idx2pnt[end - 1 - (it + uint32_t(2 * B))];
    // scalar (a min formed as v_med3 would put the loop bound in a VGPR)
    const uint32_t cnt = __builtin_amdgcn_readfirstlane(min(uint32_t(B), num - b));
    // Branch-free per pixel: a splat that misses the pixel (or a finished
    // pixel) gets e = 0, which leaves every accumulator bit-identical
    // (x + 0 = x, T * 1 = T); only wave-uniform skips branch.
    // No continue / break in this loop: jumps out of a scope with locals
    // were compiled into a per-entry exit-state switch in scalar code.
    // Only the entries that reach one of the wave's groups are visited, in
    // list order (the set bits of the wave's entry mask, lowest first).
    constexpr int kEW = (B + 63) / 64;
    uint64_t em[kEW];
    wave_entry_mask<TS, NT, B>(smask, cnt, gfirst, em);
    bool batch_done = false;
#pragma unroll
    for (int h = 0; h < kEW; ++h) {
    uint64_t pending = batch_done ? 0ull : em[h];
    while (pending != 0) {
      const uint32_t j = uint32_t(h) * 64u + uint32_t(__builtin_ctzll(pending));
      pending &= pending - 1;
      // groups of this wave that the box reaches (scalar)
      const EntryHits<TS, NT> eh = entry_hits(smask, j, gfirst);
      const uint32_t hits = eh.hits;
      if (hits != 0) {
        bool live[PPT];
        uint64_t live_lanes[PPT];  // ballots of live (scalar)
        uint64_t anylive = 0;
#pragma unroll
        for (int k = 0; k < PPT; ++k) {
          live[k] = !(asum[k] >= 0.999f);
          live_lanes[k] = __ballot(!(asum[k] >= 0.999f));
          anylive |= live_lanes[k];
        }
        if (anylive == 0) {
          // every pixel of the wave is done: leave the batch
          pending = 0;
          batch_done = true;
        } else {
          const float4 q2 = s2[j];
          const float4 q3 = s3[j];
          const float4 q0 = s0[j];
          const float4 q1 = s1[j];
          const ConicQ ke{q0.z, q0.w, q1.x};  // b doubled at staging
          const uint32_t wpos = __float_as_uint(q3.y);  // b + j + 1
#pragma unroll
          for (int g = 0; g < G; ++g) {
            bool inm[Q];
            // wave-uniform: the box reaches the group and a live pixel of
            // it lies inside
            if (((hits >> g) & 1u) &&
                (GM::kLaneMask ? lanes_take<TS, NT>(eh, g, live_lanes[g], inm)
                               : box_take<TS, NT>(g, live, px, py, q2, q3, inm))) {
#pragma unroll
              for (int qq = 0; qq < Q; ++qq) {
                const int k = g * Q + qq;
                const bool in = inm[qq];
                const float dx = px[k] - q0.x, dy = py[k] - q0.y;
                const float e = splat_weight(in, ke, dx, dy) * q1.y;
                const float eo = tr[k] * e;
                cr[k] = fmaf(q1.z, eo, cr[k]);
                cg[k] = fmaf(q1.w, eo, cg[k]);
                cbl[k] = fmaf(q2.x, eo, cbl[k]);
                if constexpr (kState) walked[k] = in ? wpos : walked[k];
                // 1 - e >= 2^-24 by the staged bounds on g and opacity: T
                // is never 0, and the adjoint divides it back out
                tr[k] = tr[k] * (1.f - e);
                asum[k] = asum[k] + eo;
              }
            }
          }
        }
      }
    }
    }
    bool lane_live = false;
#pragma unroll
    for (int k = 0; k < PPT; ++k) lane_live = lane_live || !(asum[k] >= 0.999f);
    if (__syncthreads_and(!lane_live)) break;
  }
#pragma unroll
  for (int k = 0; k < PPT; ++k) {
    const uint32_t ix = tx * TS + GM::lx(k), iy = ty * TS + GM::ly(k);
    if (ix >= img_w || iy >= img_h) continue;
    const size_t ip = size_t(iy) * img_w + ix;
    pix2rgb[3 * ip + 0] = cr[k];
    pix2rgb[3 * ip + 1] = cg[k];
    pix2rgb[3 * ip + 2] = cbl[k];
    if (pix2alpha) pix2alpha[ip] = asum[k];
    if (kState && pix2nwalk) pix2nwalk[ip] = walked[k];
    if (kState && pix2tfinal) pix2tfinal[ip] = tr[k];
  }
  if (kState && tile_cost) {
    // The tile's longest walk: what its adjoint (and the next frame's
    // forward) will cost, used to schedule heavy tiles first. Every branch
    // here is on the (uniform) wave index and every lane of a wave stores
    // the same value: in a persistent loop, a per-lane `threadIdx.x == 0`
    // store here and the per-lane work fetch at the loop head were merged by
    // the compiler into an inner loop that lane 0 left while the other lanes
    // kept re-running the same tile.
    constexpr int W = NT / 64;
    __shared__ uint32_t s_wmax[W];
    uint32_t w = 0;
#pragma unroll
    for (int k = 0; k < PPT; ++k) w = max(w, walked[k]);
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) w = max(w, uint32_t(__shfl_xor(int(w), off, 64)));
    s_wmax[GM::wave()] = __builtin_amdgcn_readfirstlane(w);
    __syncthreads();
    if (GM::wave() == 0) {
      uint32_t m = s_wmax[0];
#pragma unroll
      for (int i = 1; i < W; ++i) m = max(m, s_wmax[i]);
      tile_cost[cost_slot] = m;
    }
  }
}

template <int CTRL>
__device__ inline float dpp_mov(float v) {
  return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), CTRL, 0xF, 0xF, false));
}

// v_permlane32_swap: the upper 32 lanes of a trade places with the lower 32
// lanes of b (one instruction moves both directions).
__device__ inline void swap_wave_halves(float& a, float& b) {
  const auto r = __builtin_amdgcn_permlane32_swap(__float_as_uint(a), __float_as_uint(b), false,
                                                  false);
  a = __uint_as_float(r[0]);
  b = __uint_as_float(r[1]);
}

// v_permlane16_swap: the odd 16-lane rows of a trade places with the even
// rows of b.
__device__ inline void swap_wave_rows(float& a, float& b) {
  const auto r = __builtin_amdgcn_permlane16_swap(__float_as_uint(a), __float_as_uint(b), false,
                                                  false);
  a = __uint_as_float(r[0]);
  b = __uint_as_float(r[1]);
}

// Which of the eight paired values wave_sum9_scatter leaves in lane l:
// 4 b3 + 2 b4 + b5, b_k = bit k of l.
__device__ inline int sum8_slot(int lane) {
  return ((lane >> 3) & 1) * 4 + ((lane >> 4) & 1) * 2 + ((lane >> 5) & 1);
}

// Sums of 9 values over the wave, as a reduce-scatter. Value pairs are
// exchanged between the wave halves (v_permlane32_swap: one swap and one add
// per pair leave the first value's partial sums in the lower half and the
// second's in the upper), then between 16-lane rows (v_permlane16_swap), then
// between the two 8-lane halves of each row (a row rotate by 8 with a
// select); three DPP levels finish the sums inside each 8-lane group. The
// ninth value rides along paired with itself (both halves / rows receive the
// same partial sum), so every lane ends with its total: 28 VALU
// instructions for the nine totals, no LDS. Returns the total of value
// sum8_slot(l) in `t8` (lanes with l % 8 == 0 hold distinct values) and the
// total of v[8] in `t9` (every lane).
__device__ inline void wave_sum9_scatter(const float (&v)[9], int lane, float& t8, float& t9) {
  // all swaps of a level before its adds: the in-place swap operands then
  // need no extra register copies
  float a[5], b[5];
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    a[k] = v[2 * k];
    b[k] = v[2 * k + 1];
  }
  a[4] = b[4] = v[8];
#pragma unroll
  for (int k = 0; k < 5; ++k) swap_wave_halves(a[k], b[k]);
  float s[5];
#pragma unroll
  for (int k = 0; k < 5; ++k) s[k] = a[k] + b[k];
  float c[3], d[3];
  c[0] = s[0];
  d[0] = s[1];
  c[1] = s[2];
  d[1] = s[3];
  c[2] = d[2] = s[4];
#pragma unroll
  for (int k = 0; k < 3; ++k) swap_wave_rows(c[k], d[k]);
  float t[3];
#pragma unroll
  for (int k = 0; k < 3; ++k) t[k] = c[k] + d[k];
  const bool b3 = lane & 8;
  float u = (b3 ? t[1] : t[0]) + dpp_mov<0x128>(b3 ? t[0] : t[1]);  // row_ror:8
  float w = t[2] + dpp_mov<0x128>(t[2]);
  u += dpp_mov<0xB1>(u);  // quad_perm [1,0,3,2]
  w += dpp_mov<0xB1>(w);
  u += dpp_mov<0x4E>(u);  // quad_perm [2,3,0,1]
  w += dpp_mov<0x4E>(w);
  u += dpp_mov<0x141>(u);  // row_half_mirror
  w += dpp_mov<0x141>(w);
  // Both totals are formed here, before the callers' lane-masked stores: an
  // empty asm that reads them keeps the last DPP add from being sunk into
  // the masked block (where it would not fuse with its DPP move).
  asm volatile("" ::"v"(u), "v"(w));
  t8 = u;
  t9 = w;
}

// Reduced per list entry: S_x, S_y, S_xx, S_xy, S_yy (moments of dL/dq over
// the pixel offsets), dL/dopacity, dL/drgb.
constexpr int kGrad = 9;

// Gradient sinks of the adjoint: 2D terms per list item, opacity / colour per
// Gaussian (item - gauss_base; 0 for a single view).
struct GradSink {
  float* d_pixxy;   // [items, 2]
  float* d_conic;   // [items, 3]
  float* d_op;      // [gaussians]
  float* d_rgb;     // [gaussians, 3], or [items, 3] with rgb_per_item
  bool rgb_per_item = false;  // per-view colours (SH colour of a batched plan)
};

// LDS of the adjoint, declared once per kernel by raster_tile_bwd (each
// template instance declaring its own arrays would double the footprint).
// Batches hold B = 64 list entries: 14.8 KB in all at 32x32 tiles and 256
// threads, so registers rather than LDS set the occupancy there (7 waves per
// SIMD; batches of 128 took 29.7 KB and held it to 5). With one or two
// waves, acc holds one partial-sum slot per wave (waves combine by plain
// stores); with more, one slot that the waves add into with LDS float
// atomics (per-wave slots would take so much LDS that fewer workgroups fit on
// a CU).
template <int TS, int NT>
struct BwdLds {
  static constexpr int B = 64;
  static constexpr int S = NT / 64 <= 2 ? NT / 64 : 1;
  float4 s0[B], s1[B], s2[B], s3[B];
  uint32_t spid[B];
  float acc[S][kGrad][B];
  EntryMasks<TS, NT, B> smask;
  uint32_t smax;
};

// Adjoint of one tile. dL/drgb is read from d_img, or -- when d_img is null
// -- formed in place as 2 (rgb - target) for the loss sum (rgb - target)^2.
// kAlpha: whether dL/dalpha is present (d_alpha != null).
template <int TS, int NT, bool kAlpha>
__device__ __forceinline__ void raster_tile_bwd_impl(uint32_t tx, uint32_t ty, uint32_t img_w,
                                                uint32_t img_h, uint32_t end,
                                                const float4* __restrict__ splat2,
                                                const uint32_t* __restrict__ idx2pnt,
                                                const uint32_t* __restrict__ pix2nwalk,
                                                const float* __restrict__ pix2tfinal,
                                                const float* __restrict__ d_img,
                                                const float* __restrict__ pix2rgb,
                                                const float* __restrict__ target,
                                                const float* __restrict__ d_alpha,
                                                uint32_t gauss_base, const GradSink& out,
                                                BwdLds<TS, NT>& L) {
#pragma clang fp contract(off)
  using GM = GroupMap<TS, NT>;
  constexpr int PPT = GM::PPT, G = GM::G, Q = GM::Q, W = GM::W;
  float4* s0 = L.s0;
  float4* s1 = L.s1;
  float4* s2 = L.s2;
  float4* s3 = L.s3;
  uint32_t* spid = L.spid;
  EntryMasks<TS, NT, BwdLds<TS, NT>::B>& smask = L.smask;
  uint32_t& smax = L.smax;
  const uint32_t wv = GM::wave();
  const uint32_t gfirst = wv * uint32_t(G);
  constexpr int B = BwdLds<TS, NT>::B, S = BwdLds<TS, NT>::S;
  float(&acc)[kGrad][B] = L.acc[S == 1 ? 0 : wv];
  const int lane = threadIdx.x & 63;
  float px[PPT], py[PPT], tr[PPT], g0[PPT], g1[PPT], g2[PPT], ga[PPT];
  // gb = g . (colour behind) + ga * (alpha behind): the back accumulators
  // only enter dL/de through this dot product
  float gb[PPT];
  // nw[k] = nwalk: the pixel's last contributing splat sits at position
  // nw - 1; tr starts at the pixel's final transmittance, behind that splat,
  // and every splat of the walk back divides it by its own 1 - e (at the
  // last contributor too: 1 - e is exact for e >= 1/2 and within half an ulp
  // of 1 - e otherwise, so the recovered T is within ~1 ulp)
  uint32_t nw[PPT];
  uint32_t my_max = 0;
#pragma unroll
  for (int k = 0; k < PPT; ++k) {
    const uint32_t ix = tx * TS + GM::lx(k), iy = ty * TS + GM::ly(k);
    px[k] = float(ix) + 0.5f;
    py[k] = float(iy) + 0.5f;
    gb[k] = 0.f;
    if (ix < img_w && iy < img_h) {
      const size_t ip = size_t(iy) * img_w + ix;
      nw[k] = pix2nwalk[ip];
      tr[k] = pix2tfinal[ip];
      if (d_img) {
        g0[k] = d_img[3 * ip];
        g1[k] = d_img[3 * ip + 1];
        g2[k] = d_img[3 * ip + 2];
      } else {
        g0[k] = 2.f * (pix2rgb[3 * ip] - target[3 * ip]);
        g1[k] = 2.f * (pix2rgb[3 * ip + 1] - target[3 * ip + 1]);
        g2[k] = 2.f * (pix2rgb[3 * ip + 2] - target[3 * ip + 2]);
      }
      ga[k] = d_alpha ? d_alpha[ip] : 0.f;
    } else {
      nw[k] = 0;
      tr[k] = 1.f;
      g0[k] = g1[k] = g2[k] = ga[k] = 0.f;
    }
    my_max = max(my_max, nw[k]);
  }
  // walk bound of this wave (entries beyond it touch none of its pixels)
  uint32_t wmax = my_max;
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) wmax = max(wmax, uint32_t(__shfl_xor(int(wmax), off, 64)));
  wmax = __builtin_amdgcn_readfirstlane(wmax);
  if (__builtin_amdgcn_readfirstlane(threadIdx.x >> 6) == 0) smax = 0;  // wave 0, uniform branch
  __syncthreads();
  if (lane == 0) atomicMax(&smax, wmax);
  __syncthreads();
  const uint32_t maxw = __builtin_amdgcn_readfirstlane(smax);  // scalar loop bounds
  for (uint32_t hi = maxw; hi > 0;) {
    const uint32_t lo = hi > uint32_t(B) ? hi - B : 0u;
    const uint32_t cnt = __builtin_amdgcn_readfirstlane(hi - lo);  // scalar loop bound
    if (threadIdx.x < cnt) {
      const uint32_t it = hi - 1 - threadIdx.x;
      const uint32_t p = idx2pnt[end - 1 - it];
      const float4* r = splat2 + 4 * size_t(p);
      const float4 q0 = r[0], q1 = r[1], q2 = r[2], q3 = r[3];
      // the conic with b doubled where the walk reads it, the raw conic in
      // the record's free slots for the flush
      const ConicQ ke = conic_q(q0.z, q0.w, q1.x);
      s0[threadIdx.x] = make_float4(q0.x, q0.y, ke.a, ke.b2);
      s1[threadIdx.x] = make_float4(ke.c, stage_opacity(q1.y), q1.z, q1.w);
      s2[threadIdx.x] = q2;
      s3[threadIdx.x] = make_float4(q3.x, q0.z, q0.w, q1.x);
      spid[threadIdx.x] = p;
      smask.stage(threadIdx.x, tx, ty, q2.y, q2.z, q2.w, q3.x);
    }
    // partial-sum slots start at zero (entries a wave skips stay zero)
    if (S > 1) {
      for (uint32_t j = lane; j < cnt; j += 64) {
#pragma unroll
        for (int c = 0; c < kGrad; ++c) acc[c][j] = 0.f;
      }
    } else if (threadIdx.x < cnt) {
#pragma unroll
      for (int c = 0; c < kGrad; ++c) acc[c][threadIdx.x] = 0.f;
    }
    __syncthreads();
    // the entries that reach one of the wave's groups, minus those beyond
    // every walk of the wave (slot j holds walk position hi - 1 - j, so
    // those are the slots below hi - wmax)
    constexpr int kEW = (B + 63) / 64;
    uint64_t em[kEW];
    wave_entry_mask<TS, NT, B>(smask, cnt, gfirst, em);
    const uint32_t jmin = hi > wmax ? hi - wmax : 0u;
#pragma unroll
    for (int h = 0; h < kEW; ++h) {
      const uint32_t j0 = uint32_t(h) * 64u;
      if (jmin >= j0 + 64u) em[h] = 0;
      else if (jmin > j0) em[h] &= ~0ull << (jmin - j0);
    }
#pragma unroll
    for (int h = 0; h < kEW; ++h)
    for (uint64_t pending = em[h]; pending != 0; pending &= pending - 1) {
      const uint32_t j = uint32_t(h) * 64u + uint32_t(__builtin_ctzll(pending));
      const uint32_t it = hi - 1 - j;
      // groups of this wave that the box reaches (scalar)
      const EntryHits<TS, NT> eh = entry_hits(smask, j, gfirst);
      const uint32_t hits = eh.hits;
      if (hits == 0) continue;
      const float4 q2 = s2[j], q3 = s3[j];
      const float4 q0 = s0[j], q1 = s1[j];
      const ConicQ ke{q0.z, q0.w, q1.x};  // b doubled at staging
      // one row for all of the lane's pixels: dy and (c dy) dy once per
      // entry (in the forward, hoisting them measured slower)
      const float dyr = py[0] - q0.y;
      const float cyy = (ke.c * dyr) * dyr;
      // -0 so that the first group's terms enter without an add (x + -0 = x)
      float v[kGrad];
#pragma unroll
      for (int c = 0; c < kGrad; ++c) v[c] = -0.f;
      bool touched = false;
#pragma unroll
      for (int g = 0; g < G; ++g) {
        if (!((hits >> g) & 1u)) continue;  // wave-uniform
        bool walking[PPT];
#pragma unroll
        for (int k = 0; k < PPT; ++k) walking[k] = it < nw[k];
        bool inm[Q];
        // no pixel of the group takes this entry
        if (GM::kLaneMask ? !lanes_take<TS, NT>(eh, g, __ballot(it < nw[g]), inm)
                          : !box_take<TS, NT>(g, walking, px, py, q2, q3, inm))
          continue;
        touched = true;
        // Branch-free per pixel: a splat that misses the pixel gets gk = e = 0,
        // which adds exact zeros and keeps T (T * rcp(1) = T).
#pragma unroll
        for (int qq = 0; qq < Q; ++qq) {
          const int k = g * Q + qq;
          const bool in = inm[qq];
          const float dx = px[k] - q0.x;
          const float dy = GM::kRowShared ? dyr : py[k] - q0.y;
          const float gk = GM::kRowShared ? splat_weight_cyy(in, ke, dx, dy, cyy)
                                          : splat_weight(in, ke, dx, dy);
          const float e = gk * q1.y;
          // 1 - e >= 2^-24 as in the forward: an opaque splat on a pixel
          // centre gives a large finite 1 / (1 - e), never 0 * inf
          const float inv = __builtin_amdgcn_rcpf(1.f - e);
          tr[k] = tr[k] * inv;
          const float t = tr[k];
          // dL/de = g.(c T - B/(1-e)) + ga (T - A/(1-e)) = T gc - gb/(1-e)
          float gc = fmaf(g0[k], q1.z, fmaf(g1[k], q1.w, g2[k] * q2.x));
          if (kAlpha) gc = gc + ga[k];
          const float dl_de = fmaf(t, gc, -(gb[k] * inv));
          const float et = e * t;
          v[6] = fmaf(g0[k], et, v[6]);
          v[7] = fmaf(g1[k], et, v[7]);
          v[8] = fmaf(g2[k], et, v[8]);
          gb[k] = fmaf(gc, et, gb[k]);
          v[5] = fmaf(dl_de, gk, v[5]);
          // moments of -2 dL/dq = dL/de e: the mean and conic gradients of
          // the entry are linear in (S_x, S_y, S_xx, S_xy, S_yy), formed once
          // at the flush (where the factor -1/2 is applied, exactly)
          const float dl_dq2 = dl_de * e;
          const float a = dl_dq2 * dx;
          v[0] = v[0] + a;
          v[2] = fmaf(a, dx, v[2]);
          if constexpr (GM::kRowShared) {
            // dy is the lane's for every group: the y moments follow from
            // sum w (= opacity * v[5]) and sum w dx after the groups
          } else {
            const float bq = dl_dq2 * dy;
            v[1] = v[1] + bq;
            v[3] = fmaf(a, dy, v[3]);
            v[4] = fmaf(bq, dy, v[4]);
          }
        }
      }
      if (!touched) continue;
      if constexpr (GM::kRowShared) {
        // sum of dL/de e = opacity * sum of dL/de g (e = g * opacity)
        v[1] = (v[5] * q1.y) * dyr;
        v[3] = v[0] * dyr;
        v[4] = v[1] * dyr;
      }
      float tot, tot9;
      wave_sum9_scatter(v, lane, tot, tot9);
      // lanes 0, 8, ..., 56 hold the eight paired totals, lane 1 writes the
      // ninth (every lane holds it)
      const bool slot_lane = (lane & 7) == 0;
      if (W == 1 || S > 1) {
        if (slot_lane) acc[sum8_slot(lane)][j] = tot;
        if (lane == 1) acc[8][j] = tot9;
      } else {
        // one atomic with a per-lane slot: a lane-uniform address would be
        // expanded by the compiler into a loop over the active lanes
        if (slot_lane || lane == 1)
          atomicAdd(&acc[slot_lane ? sum8_slot(lane) : 8][j], slot_lane ? tot : tot9);
      }
    }
    __syncthreads();
    if (threadIdx.x < cnt) {
      const uint32_t t = threadIdx.x;
      float sum[kGrad];
#pragma unroll
      for (int c = 0; c < kGrad; ++c) {
        sum[c] = L.acc[0][c][t];
#pragma unroll
        for (int w = 1; w < S; ++w) sum[c] += L.acc[w][c][t];
      }
      const uint32_t p = spid[t];
      const uint32_t g = p - gauss_base;
      const float ca = s3[t].y, cb = s3[t].z, cc = s3[t].w;
      // sum[0..4] are moments of -2 dL/dq. d q / d mean = -2 C d, so
      // dL/dmean = -2 C (S_x, S_y) = C (sum[0], sum[1]) (scaling by powers
      // of two is exact: the same bits as scaling every term)
      const float a0 = ca * sum[0] + cb * sum[1], a1 = cb * sum[0] + cc * sum[1];
      if (a0 != 0.f) atomicAdd(&out.d_pixxy[2 * size_t(p)], a0);
      if (a1 != 0.f) atomicAdd(&out.d_pixxy[2 * size_t(p) + 1], a1);
      // d q / d(a, b, c) = (dx^2, 2 dx dy, dy^2)
      const float dc[3] = {-0.5f * sum[2], -sum[3], -0.5f * sum[4]};
#pragma unroll
      for (int c = 0; c < 3; ++c) {
        if (dc[c] != 0.f) atomicAdd(&out.d_conic[3 * size_t(p) + c], dc[c]);
      }
      if (sum[5] != 0.f) atomicAdd(&out.d_op[g], sum[5]);
#pragma unroll
      for (int c = 0; c < 3; ++c) {
        if (sum[6 + c] != 0.f)
          atomicAdd(&out.d_rgb[3 * size_t(out.rgb_per_item ? p : g) + c], sum[6 + c]);
      }
    }
    __syncthreads();
    hi = lo;
  }
}

template <int TS, int NT>
__device__ __forceinline__ void raster_tile_bwd(uint32_t tx, uint32_t ty, uint32_t img_w,
                                                uint32_t img_h, uint32_t end,
                                                const float4* __restrict__ splat2,
                                                const uint32_t* __restrict__ idx2pnt,
                                                const uint32_t* __restrict__ pix2nwalk,
                                                const float* __restrict__ pix2tfinal,
                                                const float* __restrict__ d_img,
                                                const float* __restrict__ pix2rgb,
                                                const float* __restrict__ target,
                                                const float* __restrict__ d_alpha,
                                                uint32_t gauss_base, const GradSink& out) {
  __shared__ BwdLds<TS, NT> lds;
  if (d_alpha)
    raster_tile_bwd_impl<TS, NT, true>(tx, ty, img_w, img_h, end, splat2, idx2pnt, pix2nwalk,
                                       pix2tfinal, d_img, pix2rgb, target, d_alpha, gauss_base,
                                       out, lds);
  else
    raster_tile_bwd_impl<TS, NT, false>(tx, ty, img_w, img_h, end, splat2, idx2pnt, pix2nwalk,
                                        pix2tfinal, d_img, pix2rgb, target, d_alpha, gauss_base,
                                        out, lds);
}

}  // namespace dsplat
