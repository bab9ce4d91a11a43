// Internal interface of the binning phases (binning.hip) for producers that
// compute the count phase's per-Gaussian outputs themselves.
#pragma once
#include <hip/hip_runtime.h>
#include <stddef.h>
#include <stdint.h>

namespace dsplat {

// Bytes of the one-buffer binning workspace of n Gaussians and a pair
// capacity (dsplat_tile2idx_idx2pnt_workspace_size).
size_t bin_workspace_bytes(uint32_t n, uint32_t capacity);

// Where the count phase keeps each Gaussian's packed tile rectangle
// (pack_rect) and depth key (depth_key) in a workspace of n Gaussians: a
// producer that writes them runs bin_after_count instead of the whole
// binning.
void bin_count_slots(void* ws, uint32_t n, uint2** rect, uint32_t** dkey);

// The binning after its count phase (depth order, pair fill, tile sort,
// tile ranges; no host wait, K into *d_num_idx) over a workspace whose count
// outputs are in place. Tile grids up to 65535 per axis (the caller checks).
int bin_after_count(hipStream_t s, uint32_t n, uint32_t ts, uint32_t tw, uint32_t th,
                    uint32_t* tile2idx, uint32_t* idx2pnt, uint32_t capacity, void* ws,
                    size_t ws_bytes, uint32_t* d_num_idx);

}  // namespace dsplat
