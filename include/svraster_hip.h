/*
 * C-ABI of the MI355X (gfx950) sparse-voxel SDF rasterizer library, libsvraster_hip.so.
 *
 * Plain pointers and sizes only.  Every pointer argument named like a tensor is DEVICE memory
 * (float32 / int64 / int32 / uint8 as stated); `stream` is a hipStream_t (NULL = the null
 * stream).  Functions return 0 on success and a nonzero code on failure, with a message
 * available from svr_last_error() (thread-local).  Inputs are read-only unless documented as
 * outputs or in-place; "zeroed by the caller" outputs are accumulated into.
 *
 * Each entry point names the reference interface it replaces (paths relative to the
 * reference's cuda/ directory).  The reference binds these through pybind11
 * (binding.cpp:28-58); here the Python module svraster_amd/_C.py binds them with ctypes.
 */
#ifndef SVRASTER_HIP_H_
#define SVRASTER_HIP_H_

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Message of the last failed call on this thread ("" when none). */
const char* svr_last_error(void);
/* ABI version of this header (bumped on any signature change). */
int svr_abi_version(void);

/* HIP-event timing of the compositor launches on the caller's stream (no counterpart in the
 * reference; for benchmarks).  kind 0 = forward compositor, 1 = backward compositor (also the
 * batched launches).  svr_timing_read synchronises on the recorded events and returns their
 * summed duration; reset != 0 clears the record. */
void svr_timing_enable(int on);
int svr_timing_read(int kind, int reset, double* total_ms, long long* launches);

/* Runtime options (no counterpart in the reference), the knobs the autotuner
 * (svraster_amd/autotune.py) sets.  "compositor_variant": 0 = pair-compacted compositor kernels
 * (default), 1 = one-entry-at-a-time kernels.  "batch_depth": entries staged in LDS per round by
 * the multi-view pair-compacted forward compositor, 64 or 128 (default).  "bwd_batch_depth": the
 * same for the pair-compacted reverse compositors (one-view and multi-view), 64 (default; five
 * workgroups per CU) or 128.  "persistent_blocks_per_cu": workgroups per CU of the multi-view
 * persistent launches, 1..32 (default 8).  None of them changes an image; gradients change only by
 * the order of float additions.  svr_get_option returns -1 for an unknown name. */
int svr_set_option(const char* name, int value);
int svr_get_option(const char* name);

/* ---- opaque state buffer sizes (the reference's raster_state.h:41-78 roles) ---------------- */
size_t svr_geom_state_bytes(int n_views, int P);
size_t svr_binning_state_bytes(long long K);
size_t svr_image_state_bytes(int n_views, long long n_pixels, long long n_tiles);

/* Replaces PREPROCESS::rasterize_preprocess (src/preprocess.cu:158-228).
 * w2c, c2w: device float[16] row-major.  out_n_dup: int32[P].  geom_buffer: device bytes of
 * svr_geom_state_bytes(1, P).  cam_mode must be CAM_PERSP (0). */
int svr_rasterize_preprocess(int W, int H, float tanfovx, float tanfovy, float cx, float cy,
                             const float* w2c, const float* c2w, int cam_mode, float near_plane,
                             const int64_t* octree_paths, const float* vox_centers,
                             const float* vox_lengths, const uint8_t* is_leaf, int P,
                             int32_t* out_n_dup, void* geom_buffer, void* stream);

/* Replaces the scan + num_rendered readback of FORWARD::rasterize_voxels_procedure
 * (src/forward.cu:881-897).  Synchronises `stream`.  K_out: number of (tile, rank) keys;
 * rank_shift_out: low rank bits that are zero in every visible voxel's octpath (a multiple of
 * 3), passed on to svr_rasterize_forward. */
int svr_rasterize_count(void* geom_buffer, int P, long long* K_out, int* rank_shift_out, void* stream);

/* Replaces FORWARD::rasterize_voxels (src/forward.cu:986-1096) after the count: duplication,
 * sort, tile ranges and compositing.  Only density_mode = SDF (1) with vox_geo_mode =
 * triinterp3 (3), cam_mode = 0 and need_distortion = 0 are supported.
 * Outputs (zeroed by the caller): out_color [3,H,W], out_depth [4,H,W] (when need_depth),
 * out_normal [3,H,W] (when need_normal), out_T [H,W], max_w [P] (may be NULL).
 * binning_buffer: svr_binning_state_bytes(K); image_buffer: svr_image_state_bytes(1, W*H, tiles). */
int svr_rasterize_forward(int vox_geo_mode, int density_mode, int W, int H, float tanfovx,
                          float tanfovy, float cx, float cy, const float* w2c, const float* c2w,
                          const float* background, int cam_mode, int need_depth,
                          int need_distortion, int need_normal, const int64_t* octree_paths,
                          const float* vox_centers, const float* vox_lengths, const float* geos,
                          const float* rgbs, int P, void* geom_buffer, long long K, int rank_shift,
                          void* binning_buffer, void* image_buffer, float* out_color,
                          float* out_depth, float* out_normal, float* out_T, float* max_w,
                          const float* s_val, void* stream);

/* Replaces BACKWARD::rasterize_voxels_backward (src/backward.cu:986-1109).
 * dL_dvox [P,12] = [dL/dgeos(8) | dL/drgbs(3) | subdivision priority(1)] and dL_ds [1] are
 * zeroed by the caller.  dL_ddepth / dL_dnormal / gt_color may be NULL when unused.
 * lambda_ascending and lambda_dist must be 0. */
int svr_rasterize_backward(long long K, int vox_geo_mode, int density_mode, int W, int H,
                           float tanfovx, float tanfovy, float cx, float cy, const float* w2c,
                           const float* c2w, const float* background, const float* vox_centers,
                           const float* vox_lengths, const float* geos, const float* rgbs, int P,
                           void* geom_buffer, void* binning_buffer, void* image_buffer,
                           const float* out_T, const float* dL_dcolor, const float* dL_ddepth,
                           const float* dL_dnormal, const float* dL_dT, float lambda_R_concen,
                           const float* gt_color, float lambda_ascending, float lambda_dist,
                           int need_depth, int need_normal, const float* s_val, float* dL_dvox,
                           float* dL_ds, void* stream);

/* Replaces SH_COMPUTE::sh_compute (src/sh_compute.cu:244-270).  idx: int64[n_idx] or NULL
 * (all P voxels).  shs: [P, M-1, 3].  rgbs [P,3] zeroed by the caller. */
int svr_sh_compute(int degree, const int64_t* idx, int n_idx, const float* vox_centers, int P,
                   const float* cam_pos, const float* sh0, const float* shs, int M, float* rgbs,
                   void* stream);

/* Replaces SH_COMPUTE::sh_compute_bw (src/sh_compute.cu:272-299).  dL_dsh0 [P,3] and
 * dL_dshs [P,M-1,3] zeroed by the caller. */
int svr_sh_compute_bw(int degree, int M, const int64_t* idx, int n_idx, const float* vox_centers,
                      int P, const float* cam_pos, const float* dL_drgbs, float* dL_dsh0,
                      float* dL_dshs, void* stream);

/* Replaces GEO_PARAMS_GATHER::gather_triinterp_geo_params (src/geo_params_gather.cu:134-152).
 * geos [P,8]: rows not in care_idx are left untouched. */
int svr_gather_triinterp_geo_params(const int64_t* vox_key, int P, const int64_t* care_idx,
                                    int n_care, const float* grid_pts, float* geos, void* stream);

/* Replaces gather_triinterp_geo_params_bw (src/geo_params_gather.cu:154-173).
 * dL_dgrid_pts [M] zeroed by the caller. */
int svr_gather_triinterp_geo_params_bw(const int64_t* vox_key, int P, const int64_t* care_idx,
                                       int n_care, const float* dL_dgeos, float* dL_dgrid_pts,
                                       void* stream);
/* The same scatter of every voxel's corner-SDF gradients, read from the batched backward's
 * dL_dvox [P,12] rows (slots 0-7) in place (no reference counterpart; saves the [P,8] copy). */
int svr_batch_gather_bw(int P, const int64_t* vox_key, const float* dL_dvox, float* dL_dgrid_pts, void* stream);

/* Replaces ADAM_STEP::unbiased_adam_step (src/adam_step.cu:106-135).  In place. */
int svr_unbiased_adam_step(int sparse, float* param, const float* grad, float* exp_avg,
                           float* exp_avg_sq, long long numel, double step, double lr,
                           double beta1, double beta2, float eps, void* stream);

/* Host-side table of the unbiased step's bias corrections for steps 1..n, out[2(t-1)] =
 * (float)(lr / (1 - beta1^t)) and out[2(t-1)+1] = (float)(1 / sqrt(1 - beta2^t)), from the same
 * double expressions as svr_unbiased_adam_step (src/adam_step.cu:117-122).  Copied to the device
 * once, it lets a captured training step advance its step count on the device. */
int svr_adam_bias_table(double lr, double beta1, double beta2, int n, float* out);

/* svr_unbiased_adam_step with the step count read from device memory (*step, 1-based; steps
 * beyond n_table use the last entry of the device copy of svr_adam_bias_table's table), so
 * the call can sit inside a captured graph whose replays each advance the step. */
int svr_sparse_adam_step_dev(int sparse, float* param, const float* grad, float* exp_avg, float* exp_avg_sq,
                             long long numel, const int32_t* step, const float* table, int n_table, float beta1,
                             float beta2, float eps, void* stream);

/* Replaces ADAM_STEP::biased_adam_step (src/adam_step.cu:137-158).  In place. */
int svr_biased_adam_step(int sparse, float* param, const float* grad, float* exp_avg,
                         float* exp_avg_sq, long long numel, float lr, float beta1, float beta2,
                         float eps, void* stream);

/* Replace GE_COMPUTE::grid_eikonal_bw (src/ge_compute.cu:120-161) and
 * LS_COMPUTE::laplacian_smoothness_bw (src/ls_compute.cu:112-155).  Accumulate into
 * grid_pts_grad [n_pts].  grid_mask: uint8[res^3]; grid_keys/grid2voxel: int32[G] sorted by
 * key; active_list: int32[A].  The per-key random offset is randv[A,3] when given, else a PCG
 * hash of (seed, key index). */
int svr_grid_eikonal_bw(const float* grid_pts, long long n_pts, const int64_t* vox_key, int N,
                        const float* grid_voxel_coord, const float* grid_voxel_size, int grid_res,
                        const uint8_t* grid_mask, const int32_t* grid_keys,
                        const int32_t* grid2voxel, int G, const int32_t* active_list, int A,
                        float weight, float vox_size_inv, unsigned int seed, const float* randv,
                        float* grid_pts_grad, void* stream);
int svr_laplacian_smoothness_bw(const float* grid_pts, long long n_pts, const int64_t* vox_key,
                                int N, const float* grid_voxel_coord,
                                const float* grid_voxel_size, int grid_res,
                                const uint8_t* grid_mask, const int32_t* grid_keys,
                                const int32_t* grid2voxel, int G, const int32_t* active_list,
                                int A, float weight, float vox_size_inv, unsigned int seed,
                                const float* randv, float* grid_pts_grad, void* stream);

/* Replaces TV_COMPUTE::total_variation_bw (src/tv_compute.cu:23-189).  grid_pts / grid_pts_grad
 * [M,C], vox_key [N,8], vox_size_inv [N] (read unless no_tv_s).  Corner i of each voxel and
 * channel receives w * (sum over its three edges of +1 where it exceeds the neighbour, else -1),
 * w = weight (* 0.01 * vox_size_inv unless no_tv_s).  With tv_sparse a voxel takes part only if a
 * corner's gradient is already non-zero; the test reads grid_pts_grad as it was before the call
 * (sparse_ws: N*C bytes of scratch), where the reference reads it while other threads write. */
int svr_total_variation_bw(const float* grid_pts, long long M, int C, const int64_t* vox_key, long long N,
                           float weight, const float* vox_size_inv, int no_tv_s, int tv_sparse,
                           uint8_t* sparse_ws, float* grid_pts_grad, void* stream);

/* Replaces VG_COMPUTE::voxel_gradient_bw (src/vg_compute.cu:23-103).  For each voxel in
 * active_list [A] (int32): g = sum_i s_i S_i / (4h) over its corner SDFs (s_i = +-1 per axis),
 * corner i receives 4 (|g|^2 - 2) weight (g . s_i) / 4 (the reference's constants). */
int svr_voxel_gradient_bw(const float* grid_pts, long long M, const int64_t* vox_key, long long N,
                          const float* vox_size_inv, const int32_t* active_list, int A, float weight,
                          float* grid_pts_grad, void* stream);

/* Replaces PL_COMPUTE::points_loss_bw (src/pl_compute.cu:14-104).  points_in_grid [Mp,3] in grid
 * cells; each point in the grid whose cell is in grid_mask takes the voxel of its cell (binary
 * search over the sorted grid_keys [G], grid2voxel [G]), and corner i of that voxel receives
 * 2 weight sdf w_i (unclamped trilinear weights w_i at the point). */
int svr_points_loss_bw(const float* points_in_grid, int Mp, const float* grid_pts, long long M,
                       const int64_t* vox_key, int N, const float* grid_voxel_coord, const float* grid_voxel_size,
                       int grid_res, const uint8_t* grid_mask, const int32_t* grid_keys, const int32_t* grid2voxel,
                       int G, float weight, float* grid_pts_grad, void* stream);

/* Cell -> grid_keys index for all res^3 cells (int32 out[res^3]): the entry the reference's
 * binary search over the sorted keys returns (src/auxiliary.h:377-386), -1 for absent cells.  Built
 * once per grid table; svr_grid_smoothness_bw_step then does one load per sample instead of a
 * search. */
int svr_grid_dense_index(const int32_t* grid_keys, int G, int grid_res, int32_t* out, void* stream);

/* svr_grid_eikonal_bw (laplacian = 0) / svr_laplacian_smoothness_bw (laplacian = 1) for a
 * captured training step: cells are looked up through dense_index (svr_grid_dense_index) and the
 * random sample offsets are drawn from seed and the device step counter *step, so every replay
 * samples new offsets as the reference's clock seed does (ls_compute.cu:43). */
int svr_grid_smoothness_bw_step(int laplacian, const float* grid_pts, long long n_pts, const int64_t* vox_key, int N,
                                const float* grid_voxel_coord, const float* grid_voxel_size, int grid_res,
                                const uint8_t* grid_mask, const int32_t* grid_keys, const int32_t* grid2voxel, int G,
                                const int32_t* dense_index, const int32_t* active_list, int A, float weight,
                                float vox_size_inv, unsigned int seed, const int32_t* step, float* grid_pts_grad,
                                void* stream);

/* LS_COMPUTE::laplacian_smoothness_bw (src/ls_compute.cu:112-155) with every table entry active
 * once (the reference trainer's randperm list with ls_drop_ratio 0, train.py:640-642, taken in
 * table order instead), for a captured training step.  The table: svr_valid_gradient_table_fill's
 * cells before any key sort, voxel by voxel, with each voxel's cell box reordered x fastest (z, y,
 * x loops).  This form requires exactly that layout with unique cells, grid_voxel_coord /
 * grid_voxel_size equal to each table voxel's integer box (power-of-two sizes, as octree voxels
 * have) and grid_mask the table's own mask.  The table voxels are numbered 0..T-1 (tvox: voxel
 * id of each); table_keys / table_tvox: cell and table voxel of entry e (int32[G]); cell_entry /
 * cell_tvox: entry and table voxel of each of the res^3 cells, -1 where absent.  The random
 * offset of entry e's key is drawn from (seed, *step, e).  Three launches: a 12-float record
 * per table voxel into rec_ws (float[12T]: box and corner SDFs), per-key coefficients into
 * coef_ws (float[3G]), then one lane per entry sums the contributions landing in its cell, the
 * lanes of a voxel are reduced across the wavefront and each voxel run adds its 8 corner sums
 * to grid_pts_grad. */
int svr_grid_laplacian_bw_table_order(const float* grid_pts, long long n_pts, const int64_t* vox_key, int N,
                                      const float* grid_voxel_coord, const float* grid_voxel_size, int grid_res,
                                      const uint8_t* grid_mask, const int32_t* table_keys,
                                      const int32_t* table_tvox, int G, const int32_t* cell_entry,
                                      const int32_t* cell_tvox, const int32_t* tvox, int T, float weight,
                                      float vox_size_inv, unsigned int seed, const int32_t* step, float* rec_ws,
                                      float* coef_ws, float* grid_pts_grad, void* stream);
/* The table-order form above for GE_COMPUTE::grid_eikonal_bw (src/ge_compute.cu:14-161) and/or
 * the Laplacian, as the reference trainer applies both each iteration (train.py:588-654, ge on
 * for iterations 0-6000, ls for 0-8000).  terms: 1 = Laplacian, 2 = eikonal, 3 = both; weights:
 * device float[2] = {eikonal weight, Laplacian weight} (each already times G / A), read by the
 * kernels, so a captured graph follows the trainer's decay schedule when the caller rewrites
 * them between replays.  Both terms use the same per-key random offset (the reference draws
 * each term's own from the clock; either way every key gets a uniform offset).  coef_ws holds
 * float[3G] for the Laplacian alone, float[7G] with the eikonal. */
int svr_grid_smoothness_bw_table_order(const float* grid_pts, long long n_pts, const int64_t* vox_key, int N,
                                       const float* grid_voxel_coord, const float* grid_voxel_size, int grid_res,
                                       const uint8_t* grid_mask, const int32_t* table_keys,
                                       const int32_t* table_tvox, int G, const int32_t* cell_entry,
                                       const int32_t* cell_tvox, const int32_t* tvox, int T, int terms,
                                       const float* weights, float vox_size_inv, unsigned int seed,
                                       const int32_t* step, float* rec_ws, float* coef_ws, float* grid_pts_grad,
                                       void* stream);

/* The voxel <-> grid-point link of octree_utils.build_grid_pts_link (src/utils/octree_utils.py:
 * 129-144, 203-222): the 8 corners of each voxel as integer coordinates at level 16, deduplicated
 * in lexicographic (x, y, z) order.  Outputs vox_key int64 [N,8] and grid_pts_key int64 [M,3]
 * (room for 8N rows); *n_grid_pts = M.  Synchronises the stream (M goes to the host).
 * ws: svr_grid_pts_link_workspace_bytes(N) bytes. */
size_t svr_grid_pts_link_workspace_bytes(long long N);
int svr_build_grid_pts_link(const int64_t* octpath, const int8_t* octlevel, long long N, void* ws,
                            int64_t* vox_key, int64_t* grid_pts_key, long long* n_grid_pts, void* stream);

/* agg_voxel_into_grid_pts (src/sparse_voxel_gears/adaptive.py:336-360): out [M,C] = mean over the
 * voxels naming each grid point of their per-corner values vox_val [N,8,C]; points no voxel
 * names get 0.  count_ws: M uint32 of scratch. */
int svr_agg_voxel_into_grid_pts(const int64_t* vox_key, long long N, const float* vox_val, int C, long long M,
                                uint32_t* count_ws, float* out, void* stream);
/* svr_agg_voxel_into_grid_pts without a workspace: the per-point counts live in a fixed device
 * array of the library (16 MiB per GPU, held until the process ends), so the call allocates
 * nothing.  No counterpart in the reference: for callers that cannot allocate device memory
 * (code under HIP graph capture, serving loops, harnesses that forbid hipMalloc).  M must be
 * at most 4194304 grid points (larger M fails with a message).  Calls that overlap on one GPU
 * share the array and race; the counts are cleared inside the call. */
int svr_agg_voxel_into_grid_pts_noalloc(const int64_t* vox_key, long long N, const float* vox_val, int C,
                                        long long M, float* out, void* stream);

/* Replace UTILS::ijk_2_octpath / octpath_2_ijk (src/utils.cu:129-189, 329-361). */
int svr_ijk_2_octpath(const int64_t* ijk, const int8_t* octlevel, int n, int64_t* octpath,
                      void* stream);
int svr_octpath_2_ijk(const int64_t* octpath, const int8_t* octlevel, int n, int64_t* ijk,
                      void* stream);

/* Replaces UTILS::compute_rd (src/utils.cu:45-73, 276-292; bound as _C.compute_rd,
 * binding.cpp, called from src/cameras.py:75-87).  out_rd: float [3, H, W], the world-space
 * direction of each pixel centre, c2w rotation of ((x + 0.5 - cx) 2 tanfovx / W,
 * (y + 0.5 - cy) 2 tanfovy / H, 1) (not normalised).  c2w: device float[16] row-major. */
int svr_compute_rd(int W, int H, float cx, float cy, float tanfovx, float tanfovy, const float* c2w,
                   float* out_rd, void* stream);
/* Replaces UTILS::depth2pts (src/utils.cu:75-104, 294-310).  depth: float [H*W];
 * out_pts: float [3, H, W] = camera position + depth * compute_rd direction. */
int svr_depth2pts(int W, int H, float cx, float cy, float tanfovx, float tanfovy, const float* c2w,
                  const float* depth, float* out_pts, void* stream);
/* Replaces UTILS::is_in_cone (src/utils.cu:23-43, 254-274).  pts: float [N,3];
 * out_is_in: bool/uint8 [N] = (z > near) and |x| <= tanfovx z and |y| <= tanfovy z in camera
 * space (w2c: device float[16] row-major). */
int svr_is_in_cone(float tanfovx, float tanfovy, float near_plane, const float* w2c, const float* pts,
                   long long N, uint8_t* out_is_in, void* stream);
/* Replaces UTILS::voxel_order_rank (src/utils.cu:107-127, 312-327).  out_ranks: int64 [8, P],
 * row q = octree_path XOR (q repeated in every 3-bit level field). */
int svr_voxel_order_rank(const int64_t* octree_paths, int P, int64_t* out_ranks, void* stream);

/* Device self-check (no counterpart in the reference): the preprocess takes 1/z of corner depths
 * in [2^-100, 2^100] as the hardware reciprocal plus one Newton step instead of the IEEE division
 * sequence, which must give the same bits.  This compares the two on every float32 significand of
 * the exponents [e_lo, e_hi] (within [-126, 127]), both signs, and adds the number of mismatches
 * to mismatches (device uint64[1], zeroed by the caller). */
int svr_selfcheck_rcp_rn(int e_lo, int e_hi, unsigned long long* mismatches, void* stream);

/* Device self-check (no counterpart in the reference): the reverse compositor recovers the
 * transmittance before an entry as T / (1 - alpha) with div_rn (correctly rounded reciprocal,
 * product, one FMA correction) instead of the IEEE division sequence, which must give the same
 * bits.  This compares the two on 256 * n_blocks random operand pairs (y exponents in [-30, 0], x
 * exponents in [-60, 0] or x = 0; n_blocks in 1..2^22, `seed` picks the sample) and adds the
 * number of mismatches to mismatches (device uint64[1], zeroed by the caller). */
int svr_selfcheck_div_rn(int n_blocks, unsigned seed, unsigned long long* mismatches, void* stream);

/* Replaces UTILS::valid_gradient_table (src/utils.cu:190-252, 362-408) in two calls: count
 * (synchronises `stream`; workspace of svr_valid_gradient_table_workspace_bytes(N)), then
 * fill.  grid_mask uint8[(2^grid_res_pow2)^3] zeroed by the caller; keys / voxel ids int32[count]
 * in voxel order (deterministic). */
size_t svr_valid_gradient_table_workspace_bytes(int N);
int svr_valid_gradient_table_count(const float* vox_center, const float* vox_size,
                                   const float* scene_center, float inside_extent,
                                   int grid_res_pow2, const uint8_t* is_leaf, int N,
                                   void* workspace, long long* count_out, void* stream);
int svr_valid_gradient_table_fill(const float* vox_center, const float* vox_size,
                                  const float* scene_center, float inside_extent,
                                  int grid_res_pow2, const uint8_t* is_leaf, int N,
                                  void* workspace, uint8_t* grid_mask, int32_t* grid_keys,
                                  int32_t* grid2voxel, void* stream);

/* ---- batched views (no counterpart in the reference: its API renders one view per call) ----
 * One workspace of svr_batch_workspace_bytes(V, P, n_pixels, n_tiles, K_cap) bytes serves V
 * views of one voxel set.  Per-pixel outputs are concatenated per view, each view [C, H, W].
 * Per-view colours: rgbs_views [V, P, 3].  No call below synchronises except
 * svr_batch_key_count; keys beyond K_cap are dropped (check with svr_batch_key_count).
 * svr_batch_setup: host arrays W, H, tanfovx, tanfovy, cx, cy, near_plane [V]; device
 * w2c, c2w [V,16] and background [3]; tile shape tile_w x tile_h = 16x16 (the reference's, and
 * the only shape whose per-tile lists match the per-view path's), 32x8 or 8x32 (same pixels
 * per tile, four 8x8 quadrants per tile in a row or a column); n_tiles counts tiles of that
 * shape.  Images and gradients do not depend on the shape. */
size_t svr_batch_workspace_bytes(int n_views, int P, long long n_pixels, long long n_tiles, long long K_cap);
int svr_batch_setup(void* ws, int n_views, int P, long long n_pixels, long long n_tiles, long long K_cap,
                    const int* W, const int* H, const float* tanfovx, const float* tanfovy,
                    const float* cx, const float* cy, const float* near_plane, const float* w2c,
                    const float* c2w, const float* background, int tile_w, int tile_h, void* stream);
/* Rank order of every (voxel, octant) pair, computed once per voxel set: order_out[8P] holds
 * voxel << 3 | octant sorted by the reference's in-tile rank octpath ^ repeat3(octant)
 * (auxiliary.h:40-59, forward.cu:749-802).  tmp: svr_batch_pair_order_bytes(P) bytes. */
size_t svr_batch_pair_order_bytes(int P);
int svr_batch_pair_order(int P, const int64_t* octree_paths, void* tmp, uint32_t* order_out, void* stream);
/* Inverse of the pair rank order (no reference counterpart; once per voxel set):
 * slots_out[pair_order[i]] = i for the 8P pairs, the rank-order slot of pair voxel << 3 | octant. */
int svr_batch_pair_slots(int P, const uint32_t* pair_order, uint32_t* slots_out, void* stream);
/* Preprocess (bbox and octant set; preprocess.cu:60-153) of every (view, voxel), voxel-major,
 * and the key count of every (voxel, octant) pair over all views, scanned in rank order.  Keys
 * cover the tiles of each (view, voxel)'s candidate rectangle -- the pixels whose centre lies in
 * the projected corners' rectangle (1/32-pixel margin) -- rather than its bbox, and a (view,
 * voxel) with no such pixel emits none: each tile's list is the reference's list less entries
 * whose ray / box test fails at every pixel of the tile, in the reference's order.  With pair_slots
 * (svr_batch_pair_slots of pair_order) the counts are stored straight at their rank-order slots;
 * with pair_slots = NULL they are gathered through pair_order in a pass of their own. */
int svr_batch_preprocess(void* ws, int n_views, int P, long long n_pixels, long long n_tiles, long long K_cap,
                         const float* vox_centers, const float* vox_lengths, const uint8_t* is_leaf,
                         const uint32_t* pair_order, const uint32_t* pair_slots, void* stream);
/* out[i] = is_leaf[i] and voxel i can still produce a pair alpha >= 1e-5 (MIN_ALPHA,
 * config.h:19) for the corner SDFs `geos` [P,8] and s: a conservative bound, so a voxel with
 * out[i] = 0 contributes to no pixel (forward.cu:303-305 skips every such pair).  Passing `out`
 * as svr_batch_preprocess's is_leaf leaves those voxels out of binning with unchanged images and
 * gradients.  cam_scale >= max |camera origin coordinate| over the views rendered.  No reference
 * counterpart (the reference bins every visible voxel). */
int svr_voxel_activity(int P, const uint8_t* is_leaf, const float* geos, const float* s_val,
                       const float* vox_centers, const float* vox_lengths, float cam_scale, uint8_t* out,
                       void* stream);
/* Voxel records of the multi-view compositors, out_records float [P,16] (16-byte aligned), per
 * voxel: centre xyz, size, the eight corner SDFs gathered from grid_pts [M] through vox_key [P,8]
 * (GatherGeoParams, geo_params_gather.cu:134-152 / renderer.py:370-406, fused here with the
 * record build: no [P,8] corner-SDF array is written), 1/size and three zeros.  With out_bin_leaf
 * [P] it also writes what svr_voxel_activity would for those corner SDFs (cull != 0) or is_leaf
 * itself (cull = 0). */
int svr_batch_voxel_records(int P, const int64_t* vox_key, const float* grid_pts, const float* vox_centers,
                            const float* vox_lengths, const uint8_t* is_leaf, const float* s_val, float cam_scale,
                            int cull, float* out_records, uint8_t* out_bin_leaf, void* stream);
/* SH colour of every (view, visible voxel) pair into rgbs_views. */
int svr_batch_sh_forward(void* ws, int n_views, int P, long long n_pixels, long long n_tiles, long long K_cap,
                         const float* vox_centers, const float* sh0, const float* shs, int degree, int M,
                         float* rgbs_views, void* stream);
/* Keys written pair by pair in rank order, a stable sort by global (view, tile) id only (each
 * tile's entries then follow the reference's order, forward.cu:749-829), tile ranges, and one
 * persistent compositor launch over all (view, tile) pairs, the entries' geometry read from
 * svr_batch_voxel_records' records.  Colour and T need no clearing; depth is zeroed by the
 * caller.  With mse_target (laid out like out_color) and out_dL_dcolor, each pixel write also
 * writes the gradient of its view's mean squared colour error, (color - target) * 2 / (3 W H):
 * the training step's loss fused into the compositor (both null: not written). */
int svr_batch_rasterize(void* ws, int n_views, int P, long long n_pixels, long long n_tiles, long long K_cap,
                        const float* vox_records, const float* rgbs_views, const float* s_val, const uint32_t* pair_order, int need_depth,
                        int need_normal, float* out_color, float* out_depth, float* out_normal, float* out_T,
                        const float* mse_target, float* out_dL_dcolor, void* stream);
/* svr_batch_rasterize in two calls: the keys, tile sort and tile ranges (svr_batch_bin, after
 * svr_batch_preprocess), then the compositor (svr_batch_composite, after svr_batch_sh_forward).  The
 * two calls between them do not touch each other's data, so a caller can run svr_batch_sh_forward on
 * a second stream beside svr_batch_bin and join before svr_batch_composite. */
int svr_batch_bin(void* ws, int n_views, int P, long long n_pixels, long long n_tiles, long long K_cap,
                  const uint32_t* pair_order, void* stream);
int svr_batch_composite(void* ws, int n_views, int P, long long n_pixels, long long n_tiles, long long K_cap,
                        const float* vox_records, const float* rgbs_views, const float* s_val, int need_depth,
                        int need_normal, float* out_color, float* out_depth, float* out_normal, float* out_T,
                        const float* mse_target, float* out_dL_dcolor, void* stream);
/* One persistent reverse compositor launch over all (view, tile) pairs (geometry from the same
 * records as the forward).  dL_dT may be NULL (no transmittance cotangent).  dL_dvox [P,12]
 * (colour slots unused), dL_drgb_views [V,P,3] and dL_ds [1] zeroed by the caller.  With
 * vox_key [P,8] and dL_dgrid_pts [M] (zeroed by the caller) the corner-SDF gradients are added
 * straight to the grid points as they are flushed -- the transpose of GatherGeoParams
 * (geo_params_gather.cu:22-72) fused into the compositor -- and dL_dvox slots 0-7 stay
 * untouched. */
int svr_batch_backward(void* ws, int n_views, int P, long long n_pixels, long long n_tiles, long long K_cap,
                       const float* vox_records, const float* rgbs_views, const float* s_val, const float* out_T,
                       const float* dL_dcolor, const float* dL_ddepth, const float* dL_dnormal,
                       const float* dL_dT, int need_depth, int need_normal, float* dL_dvox,
                       float* dL_drgb_views, float* dL_ds, const int64_t* vox_key, float* dL_dgrid_pts,
                       void* stream);
/* SH transpose summed over the views, one voxel per thread; writes every row of dL_dsh0 [P,3]
 * and dL_dshs [P,M-1,3] (M <= 16).  The entries of dL_drgb_views it consumes are reset to zero,
 * so a buffer that starts at zero is clear again for the next svr_batch_backward. */
int svr_batch_sh_backward(void* ws, int n_views, int P, long long n_pixels, long long n_tiles, long long K_cap,
                          const float* vox_centers, float* dL_drgb_views, int degree, int M,
                          float* dL_dsh0, float* dL_dshs, void* stream);
/* svr_batch_sh_backward fused with the unbiased Adam step of SparseAdam.step
 * (src/sparse_voxel_gears/optimizer.py:24-27, cuda/svraster_cuda/sparse_adam.py:33-72,
 * src/adam_step.cu:21-135) on sh0 [P,3] and shs [P,M-1,3]: the summed gradient of each element
 * updates (param, m, v) in place and is never written to memory.  sparse = 0 is the
 * reference trainer's dense step (every element, zero gradients included); sparse != 0 leaves
 * elements whose gradient is exactly 0 alone (adam_step.cu:44-45).  The bias corrections come
 * from svr_adam_bias_table tables (one per parameter group, as the groups' learning rates may
 * differ) indexed by the device step counter *step. */
int svr_batch_sh_backward_adam(void* ws, int n_views, int P, long long n_pixels, long long n_tiles,
                               long long K_cap, const float* vox_centers, float* dL_drgb_views, int degree, int M,
                               float* sh0, float* sh0_m, float* sh0_v, const float* table_sh0, float* shs,
                               float* shs_m, float* shs_v, const float* table_shs, int n_table,
                               const int32_t* step, float beta1, float beta2, float eps, int sparse,
                               void* stream);
/* Byte offsets inside the batch workspace of (for inspection, like unpack_ImageState /
 * unpack_BinningState, raster_state.cu:120-230): out[0] tile ranges uint2[n_tiles], out[1]
 * sorted values uint32[K], out[2] per-pixel contributor counts uint32[n_pixels], out[3]
 * last-contributor bound per tile uint32[n_tiles], out[4] per tile the range [x, y) of its
 * contributing entries (those the backward walks) uint2[n_tiles].  `out` holds 5 values. */
int svr_batch_state_offsets(int n_views, int P, long long n_pixels, long long n_tiles, long long K_cap,
                            long long* out);
/* Largest total key count any svr_batch_preprocess asked of this workspace since set-up or the
 * last reset (synchronises `stream`; no counterpart in the reference, whose binning state is
 * resized per call).  Keys past K_cap are not written, so a peak above K_cap means the images
 * and gradients of that step are incomplete: callers check it after eager steps and after
 * graph replays.  reset != 0 clears it (not inside a graph capture). */
int svr_batch_key_peak(void* ws, int n_views, int P, long long n_pixels, long long n_tiles, long long K_cap,
                       int reset, long long* peak_out, void* stream);
/* Total key count of the last svr_batch_preprocess (synchronises `stream`). */
int svr_batch_key_count(void* ws, int n_views, int P, long long n_pixels, long long n_tiles, long long K_cap,
                        long long* K_out, void* stream);

#ifdef __cplusplus
}
#endif

#endif /* SVRASTER_HIP_H_ */
