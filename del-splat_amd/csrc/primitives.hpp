// Internal interface of the device-wide primitives (primitives.hip).
#pragma once
#include <hip/hip_runtime.h>
#include <stddef.h>
#include <stdint.h>

#include "dsplat_common.hpp"

namespace dsplat {

constexpr int kScanBlock = 256;

// Exclusive scan of one value per thread over a kScanBlock-thread block;
// returns the block total in *total. Every thread of the block must call it.
__device__ inline uint32_t block_exclusive_scan(uint32_t v, uint32_t* total) {
  __shared__ uint32_t wsum[kScanBlock / 64];
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const uint32_t inc = wave_incl_sum_u32(v);
  if (lane == 63) wsum[w] = inc;
  __syncthreads();
  uint32_t pre = 0, tot = 0;
#pragma unroll
  for (int k = 0; k < kScanBlock / 64; ++k) {
    const uint32_t s = wsum[k];
    if (k < w) pre += s;
    tot += s;
  }
  __syncthreads();
  *total = tot;
  return pre + inc - v;
}

// Workspace bytes of exclusive_scan for n elements.
size_t scan_workspace_bytes(uint32_t n);

// out[i] = sum_{j<i} in[j] over n = min(*d_n, n_cap) elements (d_n may be
// null: n = n_cap). in == out allowed. If d_total != null the total is
// written there. ws: scan_workspace_bytes(n_cap) bytes.
int exclusive_scan(hipStream_t s, uint32_t n_cap, const uint32_t* d_n, const uint32_t* in,
                   uint32_t* out, uint32_t* ws, uint32_t* d_total);

// The first level of an exclusive scan, left for a consumer to finish:
// ws[b] = in[0] + ... + in[256 b - 1] for the 256-element chunks b of
// in[0..n), ws[nb] = the total (also written to d_total if not null). The
// consumer finishes chunk b with one block_exclusive_scan over its 256
// elements. ws: scan_chunks_workspace_bytes(n) bytes.
size_t scan_chunks_workspace_bytes(uint32_t n);
int scan_chunk_prefixes(hipStream_t s, uint32_t n, const uint32_t* in, uint32_t* ws,
                        uint32_t* d_total);

template <typename KeyT>
size_t sort_workspace_bytes(uint32_t n);

// Stable LSD radix sort of (keys, vals) over the low key_bits bits, n =
// min(*d_n, n_cap). The result is in (keys, vals) if
// *result_in_alt is false, else in the workspace's alternate buffers
// (*keys_alt, *vals_alt). iota_vals: the values are the indices 0..n-1 and
// vals is not read (it still receives the result when *result_in_alt is
// false). rect_dst / cnt_dst (may be null): the last pass also writes
// rect_dst[i] = rect_src[sorted value i] and cnt_dst[i] its area.
// seg_len (0: one segment): the array is n_cap / seg_len segments of
// seg_len elements, each sorted on its own and left in place (d_n null,
// seg_len a multiple of kSortSegAlign, at most kSortMaxSegs segments).
// Every pass then moves elements only within their segment, so the last
// pass's gathers for a segment read the values one segment can hold.
constexpr uint32_t kSortSegAlign = 32768;
constexpr uint32_t kSortMaxSegs = 64;
template <typename KeyT>
int radix_sort_pairs(hipStream_t s, uint32_t n_cap, const uint32_t* d_n, KeyT* keys,
                     uint32_t* vals, uint32_t key_bits, void* ws, size_t ws_bytes,
                     bool* result_in_alt, KeyT** keys_alt, uint32_t** vals_alt,
                     bool iota_vals = false, const uint2* rect_src = nullptr,
                     uint32_t* cnt_dst = nullptr, uint2* rect_dst = nullptr,
                     uint32_t seg_len = 0);

// Stable LSD radix sort of n = min(*d_n, n_cap) u32 keys by their bits
// [shift0, shift0 + key_bits) alone (packed keys: no values move). The last
// pass also writes vout[i] = (sorted key i) & vmask (vout may be null). The
// sorted keys are in keys if *result_in_alt is false, else in *keys_alt.
// ws: sort_workspace_bytes<uint32_t>(n_cap) bytes.
int radix_sort_keys(hipStream_t s, uint32_t n_cap, const uint32_t* d_n, uint32_t* keys,
                    uint32_t shift0, uint32_t key_bits, uint32_t vmask, uint32_t* vout, void* ws,
                    size_t ws_bytes, bool* result_in_alt, uint32_t** keys_alt);

int fill_f32(hipStream_t s, uint32_t n, float* buf, float v);

// order[0..n) = 0..n-1 by descending cost (ties in any order): a counting
// sort over min(cost, 2047) in one workgroup. A schedule, not a result:
// kernels that take it produce the same output for any order.
int order_by_cost(hipStream_t s, uint32_t n, const uint32_t* cost, uint32_t* order);

// order[i] = i.
int iota_u32(hipStream_t s, uint32_t n, uint32_t* order);

}  // namespace dsplat
