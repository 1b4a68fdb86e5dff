"""Grid-level regulariser gradients written straight into the corner-SDF gradient tensor,
API-compatible with the reference's cuda/svraster_cuda/grid_loss_bw.py:13-36.

grid_eikonal / laplacian_smoothness take two optional keywords beyond the reference's:
`seed` (int) selects the per-key random sample offset (default: a per-process counter, so every
call draws a new offset as the reference's clock seed does), and `randv` ([A,3] float32) gives
the offsets explicitly."""
from . import _C


def total_variation(grid_pts, vox_key, weight, vox_size_inv, no_tv_s, tv_sparse, grid_pts_grad):
    assert grid_pts.shape == grid_pts_grad.shape
    assert vox_key.ndim == 2 and vox_key.shape[1] == 8
    assert vox_key.shape[0] == vox_size_inv.numel()
    _C.total_variation_bw(grid_pts, vox_key, weight, vox_size_inv, no_tv_s, tv_sparse, grid_pts_grad)


def voxel_gradient(grid_pts, vox_key, vox_size_inv, active_list, weight, no_tv_s, tv_sparse, grid_pts_grad):
    assert grid_pts.shape == grid_pts_grad.shape
    assert vox_key.ndim == 2 and vox_key.shape[1] == 8
    _C.voxel_gradient_bw(grid_pts, vox_key, vox_size_inv, active_list, weight, no_tv_s, tv_sparse,
                         grid_pts_grad)


def grid_eikonal(grid_pts, vox_key, grid_voxel_coord, grid_voxel_size, grid_res, grid_mask, grid_keys,
                 grid2voxel, active_list, weight, vox_size_inv, no_tv_s, tv_sparse, grid_pts_grad,
                 seed=None, randv=None):
    assert grid_pts.shape == grid_pts_grad.shape
    assert vox_key.ndim == 2 and vox_key.shape[1] == 8
    _C.grid_eikonal_bw(grid_pts, vox_key, grid_voxel_coord, grid_voxel_size, grid_res, grid_mask,
                       grid_keys, grid2voxel, active_list, weight, vox_size_inv, no_tv_s, tv_sparse,
                       grid_pts_grad, seed=seed, randv=randv)


def laplacian_smoothness(grid_pts, vox_key, grid_voxel_coord, grid_voxel_size, grid_res, grid_mask,
                         grid_keys, grid2voxel, active_list, weight, vox_size_inv, no_tv_s, tv_sparse,
                         grid_pts_grad, seed=None, randv=None):
    assert grid_pts.shape == grid_pts_grad.shape
    assert vox_key.ndim == 2 and vox_key.shape[1] == 8
    _C.laplacian_smoothness_bw(grid_pts, vox_key, grid_voxel_coord, grid_voxel_size, grid_res, grid_mask,
                               grid_keys, grid2voxel, active_list, weight, vox_size_inv, no_tv_s,
                               tv_sparse, grid_pts_grad, seed=seed, randv=randv)


def points_loss(points_in_grid, grid_pts, vox_key, grid_voxel_coord, grid_voxel_size, grid_res, grid_mask,
                grid_keys, grid2voxel, weight, vox_size_inv, no_tv_s, tv_sparse, grid_pts_grad):
    assert grid_pts.shape == grid_pts_grad.shape
    assert vox_key.ndim == 2 and vox_key.shape[1] == 8
    _C.points_loss_bw(points_in_grid, grid_pts, vox_key, grid_voxel_coord, grid_voxel_size, grid_res,
                      grid_mask, grid_keys, grid2voxel, weight, vox_size_inv, no_tv_s, tv_sparse,
                      grid_pts_grad)
