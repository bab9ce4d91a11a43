/*
 * dsplat -- C ABI of the MI355X (gfx950) splatting library.
 *
 * Drop-in boundary for the del-splat hot path. Each entry point replaces one
 * function of the reference's host layer (del-splat-cudarc, Rust over the
 * CUDA driver API) or one primitive of del-cudarc-sys that layer calls; the
 * comment above each declaration cites the reference interface (file:line,
 * relative to the del-splat repository).
 *
 * Conventions
 *  - `stream` is a hipStream_t passed as void* (0 = the null stream). All work
 *    is enqueued on it asynchronously; only the functions marked SYNC block.
 *  - Every pointer argument is device memory unless marked "host".
 *  - Arrays are flat, row-major, float32 unless the type says otherwise:
 *    pnt2xyz [N,3], pnt2quat [N,4] (x,y,z,w), pnt2scale [N,3],
 *    pnt2pixxydepth [N,3] (pixel x, pixel y, ndc z), pnt2pixconvinv [N,3]
 *    (conic a,b,c of [[a,b],[b,c]]), pnt2pixaabb [N,4] (x0,y0,x1,y1),
 *    images [H,W,C] with pixel index iy*W+ix.
 *  - transform_world2ndc is the column-major 4x4 (m[i+4j] = M_ij), 16 floats
 *    in device memory, as in the reference.
 *  - Tile grid: tile_w = ceil(W/ts), tile_h = ceil(H/ts), T = tile_w*tile_h,
 *    tile index = iy_tile*tile_w + ix_tile; tile_size ts in {8, 16, 32}.
 *  - Return value: DSPLAT_OK (0) or a negative DSPLAT_ERR_*; the library never
 *    aborts. dsplat_last_error() returns the text of the last error of the
 *    calling thread.
 */
#ifndef DSPLAT_H_
#define DSPLAT_H_

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define DSPLAT_OK 0
#define DSPLAT_ERR_ARG (-1)       /* invalid argument (size, tile size, null) */
#define DSPLAT_ERR_HIP (-2)       /* a HIP runtime call failed */
#define DSPLAT_ERR_CAPACITY (-3)  /* workspace or capacity too small */

/* Text of the calling thread's last error ("" if none). */
const char* dsplat_last_error(void);

/* Library version, as major*10000 + minor*100 + patch. */
int dsplat_version(void);

/* ======================================================================
 * Gaussian path (del-splat-cudarc/src/splat_gauss.rs)
 * ====================================================================== */

/* 3D -> 2D projection of N Gaussians.
 * Replaces splat_gauss::pnt2splat3_to_pnt2splat2 (splat_gauss.rs:6-44,
 * kernel splat_gauss.cu:10-53). pnt2scale holds the standard deviations
 * sigma (covariance R diag(sigma^2) R^T). Writes pnt2pixxydepth,
 * pnt2pixconvinv (conic), pnt2pixaabb (3-sigma pixel box). */
int dsplat_gauss_pnt2splat3_to_pnt2splat2(void* stream, uint32_t num_pnt, const float* pnt2xyz,
                                          const float* pnt2scale, const float* pnt2quat,
                                          float* pnt2pixxydepth, float* pnt2pixconvinv,
                                          float* pnt2pixaabb, const float* transform_world2ndc,
                                          uint32_t img_w, uint32_t img_h);

/* Number of (tile, Gaussian) pairs K for the given boxes, written to the
 * device u32 *d_num_idx. First half of splat_gauss::tile2idx_idx2pnt
 * (splat_gauss.rs:104-147: count + scan, whose total the reference reads back
 * with a blocking copy). */
int dsplat_gauss_count_pairs(void* stream, uint32_t img_w, uint32_t img_h, uint32_t tile_size,
                             uint32_t num_pnt, const float* pnt2pixaabb, uint32_t* d_num_idx);

/* Workspace bytes needed by dsplat_gauss_tile2idx_idx2pnt_ws /
 * dsplat_circle_tile2idx_idx2pnt_ws for N Gaussians, T tiles and a pair
 * capacity. Host function, no device work. */
size_t dsplat_tile2idx_idx2pnt_workspace_size(uint32_t num_pnt, uint32_t num_tile,
                                              uint32_t capacity);

/* Binning + stable sort without host synchronisation (graph-capturable).
 * Replaces splat_gauss::tile2idx_idx2pnt (splat_gauss.rs:96-195) with caller-
 * owned outputs: tile2idx u32[T+1] (tile t owns idx2pnt[tile2idx[t] ..
 * tile2idx[t+1])), idx2pnt u32[capacity]. Within a tile the Gaussians are
 * ordered by ascending bits(max(ndc_z+1,0)), ties by ascending index -- the
 * order of the reference's stable sort of (tile<<32 | depth bits) keys.
 * K is written to the device u32 *d_num_idx. If K > capacity nothing else is
 * written: read *d_num_idx and call again with a larger capacity. */
int dsplat_gauss_tile2idx_idx2pnt_ws(void* stream, uint32_t img_w, uint32_t img_h,
                                     uint32_t tile_size, uint32_t num_pnt,
                                     const float* pnt2pixxydepth, const float* pnt2pixaabb,
                                     uint32_t* tile2idx, uint32_t* idx2pnt, uint32_t capacity,
                                     void* workspace, size_t workspace_bytes,
                                     uint32_t* d_num_idx);

/* Size of the static device scratch of dsplat_gauss_tile2idx_idx2pnt_noalloc
 * (128 MiB per GPU: e.g. 1M Gaussians with a pair capacity of 7M). */
#define DSPLAT_NOALLOC_SCRATCH_BYTES (128u << 20)

/* dsplat_gauss_tile2idx_idx2pnt_ws without a workspace argument, allocating
 * nothing: no hipMalloc, hipFree or allocating library call, for callers
 * that cannot allocate (code under hipGraph capture, tight serving loops,
 * harnesses that forbid it). No counterpart in the reference. The workspace is
 * a static device array of DSPLAT_NOALLOC_SCRATCH_BYTES per GPU, held until
 * the process ends; calls whose dsplat_tile2idx_idx2pnt_workspace_size(num_pnt,
 * -, capacity) exceeds it return DSPLAT_ERR_CAPACITY before any launch.
 * Results are those of the _ws variant, whatever an earlier call left in the
 * scratch. Every call on one GPU shares that GPU's scratch: calls that overlap
 * on one GPU (other streams or host threads, not ordered on one stream) race
 * and must not be issued. */
int dsplat_gauss_tile2idx_idx2pnt_noalloc(void* stream, uint32_t img_w, uint32_t img_h,
                                          uint32_t tile_size, uint32_t num_pnt,
                                          const float* pnt2pixxydepth, const float* pnt2pixaabb,
                                          uint32_t* tile2idx, uint32_t* idx2pnt,
                                          uint32_t capacity, uint32_t* d_num_idx);

/* The same binning split into phases, for callers that allocate idx2pnt
 * once K is known (the reference reads K back between counting and filling,
 * splat_gauss.rs:137-147), so the depth order can run while K travels to the
 * host. On one stream, in this order:
 *   dsplat_gauss_bin_count  K into d_num_idx (device u32)
 *   dsplat_bin_order        stable depth order of the Gaussians
 *   dsplat_gauss_bin_fill   tile2idx / idx2pnt, as the _ws variant with this
 *                           capacity (nothing binned when K > capacity)
 * ws_points (dsplat_bin_points_workspace_size(num_pnt) bytes) carries the
 * state from phase to phase; ws_pairs holds
 * dsplat_bin_pairs_workspace_size(capacity) bytes. */
size_t dsplat_bin_points_workspace_size(uint32_t num_pnt);
size_t dsplat_bin_pairs_workspace_size(uint32_t capacity);
int dsplat_gauss_bin_count(void* stream, uint32_t img_w, uint32_t img_h, uint32_t tile_size,
                           uint32_t num_pnt, const float* pnt2pixxydepth,
                           const float* pnt2pixaabb, void* ws_points, size_t ws_points_bytes,
                           uint32_t* d_num_idx);
int dsplat_bin_order(void* stream, uint32_t num_pnt, void* ws_points, size_t ws_points_bytes);
int dsplat_gauss_bin_fill(void* stream, uint32_t img_w, uint32_t img_h, uint32_t tile_size,
                          uint32_t num_pnt, const float* pnt2pixaabb, void* ws_points,
                          size_t ws_points_bytes, const uint32_t* d_num_idx, uint32_t* tile2idx,
                          uint32_t* idx2pnt, uint32_t capacity, void* ws_pairs,
                          size_t ws_pairs_bytes);

/* SYNC. Same result as the _ws variant with the reference's ownership: the
 * library allocates idx2pnt (exactly K entries, returned in *idx2pnt and the
 * host *num_idx; release it with dsplat_free) after a blocking read of K, as
 * splat_gauss.rs:147-159 does. tile2idx u32[T+1] is caller-owned. */
int dsplat_gauss_tile2idx_idx2pnt(void* stream, uint32_t img_w, uint32_t img_h,
                                  uint32_t tile_size, uint32_t num_pnt,
                                  const float* pnt2pixxydepth, const float* pnt2pixaabb,
                                  uint32_t* tile2idx, uint32_t** idx2pnt, uint32_t* num_idx);

/* Releases memory returned by the library (stream-ordered on `stream`). */
int dsplat_free(void* stream, void* ptr);

/* Tiled front-to-back alpha blending into pix2rgb [H,W,3] (overwritten; the
 * reference accumulates into an uninitialised buffer).
 * Replaces splat_gauss::rasterize_pnt2splat2 (splat_gauss.rs:46-94, kernel
 * splat_gauss.cu:56-109). num_tile+1 entries of tile2idx are read. */
int dsplat_gauss_rasterize_pnt2splat2(void* stream, uint32_t img_w, uint32_t img_h,
                                      float* pix2rgb, const float* pnt2pixxydepth,
                                      const float* pnt2pixconvinv, const float* pnt2pixaabb,
                                      const float* pnt2opacity, const float* pnt2rgb,
                                      uint32_t tile_size, const uint32_t* tile2idx,
                                      const uint32_t* idx2pnt);

/* ---- differentiable path (no counterpart in the reference) ---- */

#define DSPLAT_SCALE_IS_LOG 1u  /* pnt2scale holds log(sigma) */

/* Projection for the differentiable path: as
 * dsplat_gauss_pnt2splat3_to_pnt2splat2 (flags may select log-scales) and in
 * addition the packed per-Gaussian raster record pnt2splat2 [N,16] float
 * (x, y, a, b, c, opacity, r, g, b, x0, y0, x1, y1, 0, 0, 0).
 * pnt2pixconvinv may be null (the conics are then only in the records; the
 * raster kernels and the adjoints read nothing else). d_grad9 (may be null):
 * float [9N] set to zero -- the accumulators the raster adjoint adds into,
 * laid out d_pnt2pixxy [N,2] | d_pnt2conic [N,3] | d_pnt2opacity [N] |
 * d_pnt2rgb [N,3] -- so that no separate clearing pass runs before it. */
int dsplat_gauss_project_fwd(void* stream, uint32_t num_pnt, const float* pnt2xyz,
                             const float* pnt2scale, const float* pnt2quat,
                             const float* pnt2opacity, const float* pnt2rgb,
                             const float* transform_world2ndc, uint32_t img_w, uint32_t img_h,
                             uint32_t flags, float* pnt2pixxydepth, float* pnt2pixconvinv,
                             float* pnt2pixaabb, float* pnt2splat2, float* d_grad9);

/* Projection and binning of the differentiable path in one call, without a
 * host wait: dsplat_gauss_project_fwd (no conic array) followed by
 * dsplat_gauss_tile2idx_idx2pnt_ws, except that the projection kernel itself
 * derives each Gaussian's tile rectangle and depth key, which the binning's
 * count phase would compute from the boxes and depths. That saves one
 * kernel, and pnt2pixxydepth and pnt2pixaabb may be null (not written).
 * Records, lists, ranges and K (*d_num_idx) are identical to the two calls';
 * the workspace is dsplat_tile2idx_idx2pnt_workspace_size(num_pnt, 0,
 * capacity) bytes, and K > capacity leaves every list empty as there. No
 * counterpart in the reference (its binning is splat_gauss.rs:96-195, after
 * the projection of :6-44). */
int dsplat_gauss_project_bin_ws(void* stream, uint32_t num_pnt, const float* pnt2xyz,
                                const float* pnt2scale, const float* pnt2quat,
                                const float* pnt2opacity, const float* pnt2rgb,
                                const float* transform_world2ndc, uint32_t img_w,
                                uint32_t img_h, uint32_t flags, float* pnt2pixxydepth,
                                float* pnt2pixaabb, float* pnt2splat2, float* d_grad9,
                                uint32_t tile_size, uint32_t* tile2idx, uint32_t* idx2pnt,
                                uint32_t capacity, void* workspace, size_t workspace_bytes,
                                uint32_t* d_num_idx);

/* Tiled raster over packed records, with the state the adjoint needs:
 * pix2rgb [H,W,3], pix2alpha [H,W] (the reference's alpha_sum), pix2nwalk
 * u32 [H,W] (1 + walk position of the last splat that contributed to the
 * pixel -- the terminating one if the walk terminated --, 0 if none),
 * pix2tfinal [H,W] (the pixel's transmittance after its walk, behind that
 * splat; 1 if none).
 * pix2alpha / pix2nwalk / pix2tfinal may be null. `threads` selects the
 * workgroup size per tile (0 = default; ts 16: 64/128/256, default 128;
 * ts 32: 128/256/512/1024, default 512; ts 8: always 64). Every size
 * renders bit-identical pixels. */
int dsplat_gauss_rasterize_fwd(void* stream, uint32_t img_w, uint32_t img_h,
                               uint32_t tile_size, uint32_t threads, const float* pnt2splat2,
                               const uint32_t* tile2idx, const uint32_t* idx2pnt,
                               float* pix2rgb, float* pix2alpha, uint32_t* pix2nwalk,
                               float* pix2tfinal);

/* Adjoint of dsplat_gauss_rasterize_fwd. Walks every pixel's list
 * back-to-front from pix2tfinal. ADDS into d_pnt2pixxy [N,2], d_pnt2conic
 * [N,3], d_pnt2opacity [N], d_pnt2rgb [N,3]. d_pix2alpha may be null.
 * `threads` selects the workgroup size per tile (0 = default; ts 16:
 * 64/128/256, default 64; ts 32: 64/128/256/512/1024, default 256; ts 8:
 * always 64). */
int dsplat_gauss_rasterize_bwd(void* stream, uint32_t img_w, uint32_t img_h,
                               uint32_t tile_size, uint32_t threads, const float* pnt2splat2,
                               const uint32_t* tile2idx, const uint32_t* idx2pnt,
                               const uint32_t* pix2nwalk, const float* pix2tfinal,
                               const float* d_pix2rgb, const float* d_pix2alpha,
                               float* d_pnt2pixxy, float* d_pnt2conic, float* d_pnt2opacity,
                               float* d_pnt2rgb);

/* Tile schedule (no counterpart in the reference; results do not depend on
 * it). The _ordered raster variants process tile tile_order[i] as their i-th
 * workgroup (tile_order: u32[T], a permutation of 0..T-1; null = raster
 * order; an entry >= T is skipped, and a repeated entry renders its tile
 * twice, so only a permutation gives the full image).
 * dsplat_gauss_rasterize_fwd_ordered also writes tile_cost[t] = the
 * longest walk (pix2nwalk) in tile t (may be null), and dsplat_tile_order
 * turns such costs into the descending-cost order. Launching the longest
 * walks first shortens the tail of a launch: a frame's forward costs order
 * its adjoint and, as a hint, the next frame's forward. */
int dsplat_gauss_rasterize_fwd_ordered(void* stream, uint32_t img_w, uint32_t img_h,
                                       uint32_t tile_size, uint32_t threads,
                                       const float* pnt2splat2, const uint32_t* tile2idx,
                                       const uint32_t* idx2pnt, const uint32_t* tile_order,
                                       float* pix2rgb, float* pix2alpha, uint32_t* pix2nwalk,
                                       float* pix2tfinal, uint32_t* tile_cost);
int dsplat_gauss_rasterize_bwd_ordered(void* stream, uint32_t img_w, uint32_t img_h,
                                       uint32_t tile_size, uint32_t threads,
                                       const float* pnt2splat2, const uint32_t* tile2idx,
                                       const uint32_t* idx2pnt, const uint32_t* tile_order,
                                       const uint32_t* pix2nwalk, const float* pix2tfinal,
                                       const float* d_pix2rgb, const float* d_pix2alpha,
                                       float* d_pnt2pixxy, float* d_pnt2conic,
                                       float* d_pnt2opacity, float* d_pnt2rgb);
int dsplat_tile_order(void* stream, uint32_t num_tile, const uint32_t* tile_cost,
                      uint32_t* tile_order);

/* The adjoint of the L2 loss sum (pix2rgb - target)^2 over the image (pix2rgb
 * the forward's output, target [H,W,3]) fused into the raster adjoint: the
 * kernel forms dL/drgb = 2 (pix2rgb - target) per pixel, so no image-sized
 * gradient is written or read. Otherwise as dsplat_gauss_rasterize_bwd_ordered
 * without dL/dalpha (no counterpart in the reference, which has no adjoint). */
int dsplat_gauss_rasterize_bwd_l2(void* stream, uint32_t img_w, uint32_t img_h,
                                  uint32_t tile_size, uint32_t threads, const float* pnt2splat2,
                                  const uint32_t* tile2idx, const uint32_t* idx2pnt,
                                  const uint32_t* tile_order, const uint32_t* pix2nwalk,
                                  const float* pix2tfinal, const float* pix2rgb,
                                  const float* target, float* d_pnt2pixxy, float* d_pnt2conic,
                                  float* d_pnt2opacity, float* d_pnt2rgb);

/* Adjoint of the projection. ADDS into d_pnt2xyz [N,3], d_pnt2scale [N,3]
 * (w.r.t. sigma, or log sigma under DSPLAT_SCALE_IS_LOG), d_pnt2quat [N,4];
 * with DSPLAT_GRAD_OVERWRITE in flags it stores them instead (no zeroing
 * needed beforehand). */
#define DSPLAT_GRAD_OVERWRITE 4u
int dsplat_gauss_project_bwd(void* stream, uint32_t num_pnt, const float* pnt2xyz,
                             const float* pnt2scale, const float* pnt2quat,
                             const float* transform_world2ndc, uint32_t img_w, uint32_t img_h,
                             uint32_t flags, const float* d_pnt2pixxy, const float* d_pnt2conic,
                             float* d_pnt2xyz, float* d_pnt2scale, float* d_pnt2quat);

/* View-dependent colour from spherical harmonics (no counterpart in the
 * reference: io_ply.rs:196 reads the 45 higher-order coefficients of the 3DGS
 * PLY layout and the renderer never evaluates them).
 * pnt2rgb[i][c] = max(sum_k Y_k(d_i) pnt2sh[i][k][c] + 1/2, 0), with
 * d_i = normalize(pnt2xyz[i] - cam_pos) and Y_k the real spherical harmonics
 * of 3D Gaussian splatting, k = l^2 + l + m < (degree+1)^2 (degree <= 3).
 * pnt2sh is [N][num_coef][3], num_coef in {1, 4, 9, 16} with
 * (degree+1)^2 <= num_coef, 16-byte aligned; cam_pos is a device float[3]
 * (read by the kernel, so a captured graph follows later changes to it).
 * To render with it, pass pnt2rgb as the colour of the projection. */
int dsplat_sh_to_rgb(void* stream, uint32_t num_pnt, uint32_t degree, uint32_t num_coef,
                     const float* cam_pos, const float* pnt2xyz, const float* pnt2sh,
                     float* pnt2rgb);

/* Adjoint of dsplat_sh_to_rgb: d_pnt2sh [N][num_coef][3] = Y_k g_c and
 * d_pnt2xyz [N,3] = (I - d d^T)/|xyz - cam| * sum_{k,c} sh[k][c] g_c grad Y_k,
 * with g = d_pnt2rgb where the colour was not clamped and 0 where it was.
 * ADDS into both, or stores them under DSPLAT_GRAD_OVERWRITE. Coefficients
 * beyond (degree+1)^2 get a zero gradient. */
int dsplat_sh_to_rgb_bwd(void* stream, uint32_t num_pnt, uint32_t degree, uint32_t num_coef,
                         const float* cam_pos, const float* pnt2xyz, const float* pnt2sh,
                         const float* d_pnt2rgb, uint32_t flags, float* d_pnt2sh,
                         float* d_pnt2xyz);

/* ======================================================================
 * Circle path (del-splat-cudarc/src/splat_circle.rs) and point z-buffers
 * (del-splat-dlpack/src/splat_point.rs)
 * ====================================================================== */

/* Sphere -> screen circle: pnt2pixxydepth [N,3], pnt2pixrad [N] (pixels).
 * Replaces splat_circle::pnt2splat3_to_pnt2splat2 (splat_circle.rs:96-132,
 * kernel splat_circle.cu:7-41) and the GPU branch of
 * circle_splat_from_project_spheres (del-splat-dlpack/src/splat_circle.rs:
 * 51-71, whose module lookup is broken in the reference). */
int dsplat_circle_pnt2splat3_to_pnt2splat2(void* stream, uint32_t num_pnt, const float* pnt2xyz,
                                           float* pnt2pixxydepth, float* pnt2pixrad,
                                           const float* transform_world2ndc, uint32_t img_w,
                                           uint32_t img_h, float radius);

/* Circle binning, graph-capturable. Replaces splat_circle::tile2idx_idx2pnt
 * (splat_circle.rs:3-94; it takes the tile shape, and the tile size the
 * reference hard-codes to 16 is an argument here). Boxes are
 * [x-r, y-r, x+r, y+r]; otherwise as dsplat_gauss_tile2idx_idx2pnt_ws. */
int dsplat_circle_tile2idx_idx2pnt_ws(void* stream, uint32_t tile_w, uint32_t tile_h,
                                      uint32_t tile_size, uint32_t num_pnt,
                                      const float* pnt2pixxydepth, const float* pnt2pixrad,
                                      uint32_t* tile2idx, uint32_t* idx2pnt, uint32_t capacity,
                                      void* workspace, size_t workspace_bytes,
                                      uint32_t* d_num_idx);

/* SYNC. Allocating variant (release idx2pnt with dsplat_free). */
int dsplat_circle_tile2idx_idx2pnt(void* stream, uint32_t tile_w, uint32_t tile_h,
                                   uint32_t tile_size, uint32_t num_pnt,
                                   const float* pnt2pixxydepth, const float* pnt2pixrad,
                                   uint32_t* tile2idx, uint32_t** idx2pnt, uint32_t* num_idx);

/* Phases of the circle binning (see dsplat_gauss_bin_count; the order phase
 * is dsplat_bin_order). */
int dsplat_circle_bin_count(void* stream, uint32_t tile_w, uint32_t tile_h, uint32_t tile_size,
                            uint32_t num_pnt, const float* pnt2pixxydepth,
                            const float* pnt2pixrad, void* ws_points, size_t ws_points_bytes,
                            uint32_t* d_num_idx);
int dsplat_circle_bin_fill(void* stream, uint32_t tile_w, uint32_t tile_h, uint32_t tile_size,
                           uint32_t num_pnt, const float* pnt2pixxydepth,
                           const float* pnt2pixrad, void* ws_points, size_t ws_points_bytes,
                           const uint32_t* d_num_idx, uint32_t* tile2idx, uint32_t* idx2pnt,
                           uint32_t capacity, void* ws_pairs, size_t ws_pairs_bytes);

/* Tiled painter: per pixel the last list entry (ascending key order) whose
 * circle covers the pixel centre gives the colour; uncovered pixels are 0.
 * Replaces splat_circle::splat (splat_circle.rs:134-177, kernel
 * splat_circle.cu:93-132). */
int dsplat_circle_splat(void* stream, uint32_t img_w, uint32_t img_h, float* pix2rgb,
                        const float* pnt2pixxydepth, const float* pnt2pixrad,
                        const float* pnt2rgb, uint32_t tile_size, const uint32_t* tile2idx,
                        const uint32_t* idx2pnt);

/* Deterministic point z-buffer. Replaces point_splat_rasterize_pix_zbuffer's
 * GPU branch (del-splat-dlpack/src/splat_point.rs:80-103, kernel
 * splat_point.cu:49-74): the pixel floor(x), floor(y) of every point with
 * ndc_z in [-1,1] keeps min(bits(1-(ndc_z+1)/2) << 32 | index) in pix2key
 * u64 [H,W] (caller-owned scratch); then pix2unitdepth [H,W] gets the winning
 * depth (FLT_MAX where empty, as the reference's fill) and pix2rgb [H,W,3]
 * its colour (0 where empty). pix2unitdepth may be null. */
int dsplat_point_rasterize_zbuffer(void* stream, uint32_t num_pnt, const float* pnt2pixxyndcz,
                                   const float* pnt2rgb, float* pix2unitdepth, float* pix2rgb,
                                   uint32_t img_w, uint32_t img_h, uint64_t* pix2key);

/* As dsplat_point_rasterize_zbuffer, but every point covers all pixels whose
 * centre is within pnt2pixrad [N] of it (the circle test of
 * splat_circle.cu:120-125). No counterpart in the reference. */
int dsplat_circle_rasterize_zbuffer(void* stream, uint32_t num_pnt, const float* pnt2pixxyndcz,
                                    const float* pnt2pixrad, const float* pnt2rgb,
                                    float* pix2unitdepth, float* pix2rgb, uint32_t img_w,
                                    uint32_t img_h, uint64_t* pix2key);

/* Point painter on the GPU: the result of the reference CPU function
 * render_pix_sort_depth (del-splat-cpu/src/pnt2pixxyndcz.rs:1-40; the GPU
 * branch of point_splat_rasterize_pix is todo!() in the reference,
 * del-splat-dlpack/src/splat_point.rs:40). Per pixel the point with the
 * largest (ndc_z+1, index) wins, pixel = saturating cast of (x, y). Uncovered
 * pixels are 0. pix2key u64 [H,W] is caller-owned scratch. */
int dsplat_point_rasterize_pix(void* stream, uint32_t num_pnt, const float* pnt2pixxyndcz,
                               const float* pnt2rgb, float* pix2rgb, uint32_t img_w,
                               uint32_t img_h, uint64_t* pix2key);

/* ======================================================================
 * Batched multi-view rendering / differentiation and Adam (no counterpart in
 * the reference, which renders one view per call)
 * ====================================================================== */

#define DSPLAT_OPACITY_IS_LOGIT 2u  /* opacity input holds logit(opacity) */

/* One camera: column-major world->ndc 4x4 (host memory) and image size. */
typedef struct dsplat_view_desc {
  float world2ndc[16];
  uint32_t img_w, img_h;
} dsplat_view_desc;

/* Plan for N Gaussians and up to max_view views whose images total at most
 * max_pixels pixels, with tile size ts, raster workgroup size `threads` (as
 * dsplat_gauss_rasterize_fwd) and room for pair_capacity (tile, item) pairs.
 * Allocates all device memory once (hipMalloc); frames then allocate
 * nothing and never synchronise with the host. */
typedef struct dsplat_mv_plan dsplat_mv_plan;
int dsplat_mv_plan_create(uint32_t num_pnt, uint32_t max_view, uint64_t max_pixels,
                          uint32_t tile_size, uint32_t threads, uint32_t pair_capacity,
                          dsplat_mv_plan** plan);
int dsplat_mv_plan_destroy(dsplat_mv_plan* plan);

/* Sets the views of the next frames (host array of num_view, read before the
 * call returns). Asynchronous: the view table is written on `stream` by
 * kernels that take it as their arguments, so frames queued on `stream`
 * before the call still see the old views and frames after it the new ones
 * (give the plan's work one stream). Images are concatenated in view order:
 * view v's pixels start at pix_offsets[v]. Drops a captured graph (its
 * launches were sized for the old views). */
int dsplat_mv_set_views(dsplat_mv_plan* plan, void* stream, uint32_t num_view,
                        const dsplat_view_desc* views);

/* SH colour for the plan's next frames (no counterpart in the reference):
 * with num_coef in {1, 4, 9, 16} and (degree+1)^2 <= num_coef, the colour
 * argument of dsplat_mv_forward (and the colour block of the training
 * parameter block) holds spherical-harmonic coefficients [N][num_coef][3]
 * (16-byte aligned), evaluated per view from the camera centre of its
 * world2ndc as dsplat_sh_to_rgb does; the colour gradient of
 * dsplat_mv_backward is then dL/dsh [N][num_coef][3], and the direction part
 * of dL/dxyz is included. The backward reads the coefficients the last
 * forward was given. num_coef = 0 returns to one rgb per Gaussian. The first
 * call allocates 24 bytes per (view, Gaussian) item of the plan (hipMalloc).
 * Every view must be a perspective camera (its centre is solved from the
 * matrix when views are set). Drops a captured graph. */
int dsplat_mv_set_sh(dsplat_mv_plan* plan, uint32_t degree, uint32_t num_coef);

/* Host query: view count, total pixels/tiles and the per-view pixel and
 * global-tile offsets (arrays of num_view+1, may be null). */
int dsplat_mv_plan_info(const dsplat_mv_plan* plan, uint32_t* num_view, uint64_t* total_pix,
                        uint32_t* total_tiles, uint64_t* pix_offsets, uint32_t* tile_offsets);

/* Device pointers of plan-owned state: alpha [pix] (written only by a
 * forward given a null out_alpha, and by the training step), nwalk u32 [pix],
 * tfinal [pix], tile2idx u32 [tiles+1] (global tile ids), idx2pnt u32 [capacity]
 * (item = view*N + Gaussian), counters u32 [5] = (K, K if it fit else 0, -, -,
 * overflow flag). Any argument may be null. rgb is reserved (null). */
int dsplat_mv_outputs(const dsplat_mv_plan* plan, float** rgb, float** alpha, uint32_t** nwalk,
                      float** tfinal, uint32_t** tile2idx, uint32_t** idx2pnt,
                      uint32_t** counters);

/* Forward of all views: projection, binning + stable sort, persistent tiled
 * raster into out_rgb [pix,3] and out_alpha [pix] (null: plan-owned). flags:
 * DSPLAT_SCALE_IS_LOG, DSPLAT_OPACITY_IS_LOGIT. */
int dsplat_mv_forward(dsplat_mv_plan* plan, void* stream, const float* pnt2xyz,
                      const float* pnt2scale, const float* pnt2quat, const float* pnt2opacity,
                      const float* pnt2rgb, uint32_t flags, float* out_rgb, float* out_alpha);

/* Adjoint of the last forward, summed over views. dL/drgb comes from d_img
 * [pix,3] or, when d_img is null, from the loss sum (out_rgb - target)^2 with
 * target [pix,3]. d_alpha [pix] may be null. Overwrites d_pnt2xyz [N,3],
 * d_pnt2scale [N,3], d_pnt2quat [N,4], d_pnt2opacity [N], d_pnt2rgb [N,3]
 * (under SH colour: dL/dsh [N][K][3]). Returns DSPLAT_ERR_ARG (before any
 * launch) unless a forward over the current views and colour mode has been
 * enqueued since the plan was created or last given views / SH mode: a direct
 * dsplat_mv_forward, a training step or a graph replay (capturing alone runs
 * nothing). */
int dsplat_mv_backward(dsplat_mv_plan* plan, void* stream, const float* pnt2xyz,
                       const float* pnt2scale, const float* pnt2quat, const float* pnt2opacity,
                       const float* out_rgb, const float* d_img, const float* target,
                       const float* d_alpha, float* d_pnt2xyz, float* d_pnt2scale,
                       float* d_pnt2quat, float* d_pnt2opacity, float* d_pnt2rgb,
                       uint32_t flags);

/* Adam (Kingma & Ba 2015) over n values; *d_step (device u32) is incremented
 * first and gives the bias correction, so a captured graph can replay it. */
int dsplat_adam_step(void* stream, uint64_t n, float* param, const float* grad, float* m1,
                     float* m2, float lr, float beta1, float beta2, float eps, uint32_t* d_step);

/* One training step on the plan's views over the flat parameter block
 * params [14N] = [xyz 3N | log-scale 3N | quat 4N | opacity logit N | rgb 3N]
 * (N even): forward into out_rgb, adjoint of sum (out_rgb - targets)^2 into
 * grads [14N], Adam on all 14N values (moments m1, m2 [14N]). Under SH colour
 * (dsplat_mv_set_sh) the block is [xyz | log-scale | quat | opacity logit |
 * padding to a multiple of 4 floats | sh 3KN], i.e. the coefficients start at
 * 11N rounded up to a multiple of 4, and Adam covers the whole block; the
 * step zeroes the padding's gradients, so the padding parameters stay as
 * they are. */
int dsplat_mv_train_step(dsplat_mv_plan* plan, void* stream, float* params, float* grads,
                         float* m1, float* m2, uint32_t* d_step, const float* targets,
                         float* out_rgb, float lr, float beta1, float beta2, float eps);

/* Captures one forward (kind 0: the dsplat_mv_forward arguments), one
 * training step (kind 1: the dsplat_mv_train_step arguments) or one forward
 * + L2 adjoint (kind 2: forward arguments, targets, and grads laid out as the
 * training parameter block, 14N values or the SH layout, padding gradients
 * zeroed) into the plan's hipGraph;
 * dsplat_mv_graph_launch replays it on any stream. The pointers are baked
 * into the graph. */
int dsplat_mv_graph_capture(dsplat_mv_plan* plan, uint32_t kind, const float* pnt2xyz,
                            const float* pnt2scale, const float* pnt2quat,
                            const float* pnt2opacity, const float* pnt2rgb, uint32_t flags,
                            float* out_rgb, float* out_alpha, float* params, float* grads,
                            float* m1, float* m2, uint32_t* d_step, const float* targets,
                            float lr, float beta1, float beta2, float eps);
int dsplat_mv_graph_launch(dsplat_mv_plan* plan, void* stream);

/* Stream-ordered device-to-device copy of `bytes` bytes. */
int dsplat_memcpy_d2d(void* stream, void* dst, const void* src, uint64_t bytes);

/* ======================================================================
 * GPU primitives (replacing del-cudarc-sys)
 * ====================================================================== */

/* Workspace bytes of dsplat_exclusive_scan_u32 for n elements (host). */
size_t dsplat_exclusive_scan_workspace_size(uint32_t n);

/* out[i] = sum_{j<i} in[i] (u32, wrapping); in == out allowed.
 * Replaces del_cudarc_sys::cumsum::exclusive_scan (splat_gauss.rs:137-146). */
int dsplat_exclusive_scan_u32(void* stream, uint32_t n, const uint32_t* in, uint32_t* out,
                              void* workspace, size_t workspace_bytes);

/* Workspace bytes of dsplat_radix_sort_by_key_u64_u32 for n elements (host). */
size_t dsplat_radix_sort_workspace_size(uint32_t n);

/* Stable ascending sort of keys u64[n] carrying vals u32[n], in place. Only
 * the low `key_bits` bits of the keys are compared (1..64; higher bits must be
 * zero). Replaces del_cudarc_sys::sort_by_key_u64::radix_sort_by_key_u64_u32
 * (splat_gauss.rs:184-191) and radix_sort_by_key_u64 (splat_circle.rs:86-90). */
int dsplat_radix_sort_by_key_u64_u32(void* stream, uint32_t n, uint64_t* keys, uint32_t* vals,
                                     uint32_t key_bits, void* workspace,
                                     size_t workspace_bytes);

/* buf[0..n) = value. Replaces del_cudarc_sys::array1d::fill_f32
 * (del-splat-dlpack/src/splat_point.rs:86-87). */
int dsplat_fill_f32(void* stream, uint32_t n, float* buf, float value);

#ifdef __cplusplus
}
#endif

#endif /* DSPLAT_H_ */
