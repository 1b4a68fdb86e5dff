// Launch helpers of the per-voxel operators and grid losses (voxel_ops.hip, grid_loss.hip).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace svr {

hipError_t launch_sh_forward(int n, const int64_t* idx, int deg, int M, const float* centers,
                             const float* cam, const float* sh0, const float* shs, float* rgbs,
                             hipStream_t st);
hipError_t launch_sh_backward(int n, const int64_t* idx, int deg, int M, const float* centers,
                              const float* cam, const float* dL_drgb, float* dsh0, float* dshs,
                              hipStream_t st);
struct ViewDesc;
struct VoxelRecord;
hipError_t launch_sh_views_forward(int P, int V, const uint8_t* quads, const ViewDesc* views, int deg, int M,
                                   const float* centers, const float* sh0, const float* shs, float* rgbs_views,
                                   hipStream_t st);
// Unbiased Adam with bias corrections looked up by a device-resident step counter.
struct AdamTable {
  const int32_t* step;   // 1-based step count
  const float2* tab;     // n entries {lr / (1 - b1^t), 1 / sqrt(1 - b2^t)}
  int n;
  float b1, b2, eps;
};
// Adam state of the SH coefficients, for the SH backward that updates them in place.
struct ShAdam {
  float* sh0; float* m0; float* v0;
  float* shs; float* ms; float* vs;
  AdamTable t0, ts;
  int sparse;            // 1: elements whose gradient is exactly 0 are left alone
};
hipError_t launch_sh_views_backward(int P, int V, const uint8_t* quads, const ViewDesc* views, int deg, int M,
                                    const float* centers, float* drgb_views, float* dsh0, float* dshs,
                                    const ShAdam* adam, hipStream_t st);
hipError_t launch_voxel_records(int P, const int64_t* vox_key, const float* grid_pts, const float* centers,
                                const float* lengths, const uint8_t* is_leaf, const float* s_val, float cam_scale,
                                int cull, VoxelRecord* rec, uint8_t* bin_leaf, hipStream_t st);
hipError_t launch_voxel_activity(int P, const uint8_t* is_leaf, const float* geos, const float* s_val,
                                 const float* centers, const float* lengths, float cam_scale, uint8_t* out,
                                 hipStream_t st);
hipError_t launch_gather_geo(int n, const int64_t* vox_key, const int64_t* care, const float* grid_pts,
                             float* geos, hipStream_t st);
hipError_t launch_gather_geo_bw(int n, const int64_t* vox_key, const int64_t* care, const float* dgeos,
                                float* dgrid, hipStream_t st, int stride = 8);
hipError_t launch_adam(bool sparse, bool biased, long long n, float* p, const float* g, float* m, float* v,
                       float step_size, float b1, float b2, float rsqrt_bc2, float eps, hipStream_t st);
hipError_t launch_adam_table(bool sparse, long long n, float* p, const float* g, float* m, float* v,
                             const AdamTable& T, hipStream_t st);
hipError_t launch_grid_pts_mean(long long N, int C, const int64_t* vox_key, const float* vox_val, long long M,
                                uint32_t* count, float* out, hipStream_t st);
hipError_t launch_ijk_to_octpath(int n, const int64_t* ijk, const int8_t* lv, int64_t* out, hipStream_t st);
hipError_t launch_octpath_to_ijk(int n, const int64_t* path, const int8_t* lv, int64_t* ijk, hipStream_t st);

struct GridTable {
  const float* grid_pts; const int64_t* vox_key; const int32_t* keys; const int32_t* key2vox;
  const float* vox_coord; const float* vox_size; int n_keys; int res; int n_vox; long long n_pts;
  const int32_t* dense = nullptr;   // optional cell -> key index map (launch_grid_dense_index)
};
hipError_t launch_grid_loss(bool laplacian, const GridTable& G, const bool* mask, const int32_t* active, int A,
                            float weight, float inv, uint32_t seed, const int32_t* step, const float* randv,
                            float* grad, hipStream_t st);
// Table-order Laplacian (grid_loss.hip): every table entry active once, voxel by voxel.
struct GridOrderArgs {
  const float* grid_pts; const int64_t* vox_key; const float* coord; const float* size;
  const bool* mask;
  const int32_t* keys;        // [G] cell of each entry (table order, x fastest inside a voxel)
  const int32_t* tv;          // [G] compact table voxel of each entry
  const int32_t* cell_entry;  // [res^3] entry of each cell, -1 if none
  const int32_t* cell_tv;     // [res^3] compact table voxel of each cell, -1 if none
  const int32_t* tvox;        // [T] voxel id of each compact table voxel
  float* rec;                 // [T * 12] per-step records (scratch)
  float* coef;                // [3 * G], with the eikonal [7 * G], per-step coefficients (scratch)
  float* grad;
  int G, T, res;
  float weight, inv;          // Laplacian weight when `weights` is null; 1 / cell size
  uint32_t seed;
  const int32_t* step;
  int terms = 1;              // bit 0: Laplacian, bit 1: eikonal
  const float* weights = nullptr;  // device [2]: eikonal, Laplacian weights (overrides `weight`)
};
hipError_t launch_grid_laplacian_pull(const GridOrderArgs& A, hipStream_t st);
hipError_t launch_total_variation(long long N, int C, const float* grid_pts, const int64_t* vox_key, float weight,
                                  const float* vox_size_inv, bool no_tv_s, uint8_t* sparse_flags, float* grad,
                                  hipStream_t st);
hipError_t launch_voxel_gradient(int A, const int32_t* active, const float* grid_pts, const int64_t* vox_key,
                                 const float* vox_size_inv, float weight, float* grad, hipStream_t st);
hipError_t launch_points_loss(int Mp, const float* pts, const GridTable& G, const bool* mask, float weight,
                              float* grad, hipStream_t st);
hipError_t launch_grid_dense_index(long long n_cells, const int32_t* keys, int G, int32_t* out, hipStream_t st);
hipError_t launch_table_count(int N, const float* center, const float* size, const float* scene_center,
                              float extent, int res, const bool* is_leaf, uint32_t* count, hipStream_t st);
hipError_t launch_table_write(int N, const float* center, const float* size, const float* scene_center,
                              float extent, int res, const bool* is_leaf, const uint32_t* scan, bool* mask,
                              int32_t* keys, int32_t* key2vox, hipStream_t st);
// Camera utilities (camera.hip).
hipError_t launch_compute_rd(int W, int H, float cx, float cy, float tfx, float tfy, const float* c2w, float* rd,
                             hipStream_t st);
hipError_t launch_depth2pts(int W, int H, float cx, float cy, float tfx, float tfy, const float* c2w,
                            const float* depth, float* pts, hipStream_t st);
hipError_t launch_is_in_cone(long long N, float tfx, float tfy, float near_plane, const float* w2c,
                             const float* pts, uint8_t* out, hipStream_t st);
hipError_t launch_voxel_order_rank(int P, const int64_t* paths, int64_t* out, hipStream_t st);
hipError_t launch_selfcheck_rcp_rn(int e_lo, int e_hi, unsigned long long* out, hipStream_t st);
hipError_t launch_selfcheck_div_rn(int n_blocks, uint32_t seed, unsigned long long* out, hipStream_t st);

}  // namespace svr
