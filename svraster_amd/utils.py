"""Octree and grid-table utilities, API-compatible with the reference's
cuda/svraster_cuda/utils.py:13-97."""
import torch

from . import _C


def voxel_order_rank(octree_paths):
    """The eight per-octant rendering ranks of each voxel, int64 [8, P] (utils.py:13-16)."""
    return _C.voxel_order_rank(octree_paths)


def is_in_cone(pts, cam):
    """Points inside the camera's view cone and in front of its near plane (utils.py:19-29)."""
    assert torch.is_tensor(pts)
    assert pts.device == cam.w2c.device
    assert len(pts.shape) == 2
    assert pts.shape[1] == 3
    return _C.is_in_cone(cam.tanfovx, cam.tanfovy, cam.near, cam.w2c, pts)


def compute_rd(width, height, cx, cy, tanfovx, tanfovy, c2w_matrix):
    """World-space ray direction of every pixel centre, float32 [3, H, W] (utils.py:32-34)."""
    assert torch.is_tensor(c2w_matrix)
    return _C.compute_rd(width, height, cx, cy, tanfovx, tanfovy, c2w_matrix)


def depth2pts(width, height, cx, cy, tanfovx, tanfovy, c2w_matrix, depth):
    """World point of every pixel of a depth map, float32 [3, H, W] (utils.py:37-41)."""
    assert torch.is_tensor(c2w_matrix)
    assert depth.device == c2w_matrix.device
    assert depth.numel() == width * height
    return _C.depth2pts(width, height, cx, cy, tanfovx, tanfovy, c2w_matrix, depth)


def ijk_2_octpath(ijk, octlevel):
    assert torch.is_tensor(ijk) and torch.is_tensor(octlevel)
    assert ijk.ndim == 2 and ijk.shape[1] == 3
    assert ijk.numel() == octlevel.numel() * 3
    assert ijk.dtype == torch.int64
    assert octlevel.dtype == torch.int8
    return _C.ijk_2_octpath(ijk, octlevel)


def octpath_2_ijk(octpath, octlevel):
    assert torch.is_tensor(octpath) and torch.is_tensor(octlevel)
    assert octpath.numel() == octlevel.numel()
    assert octpath.dtype == torch.int64
    assert octlevel.dtype == torch.int8
    return _C.octpath_2_ijk(octpath, octlevel)


def valid_gradient_table(vox_center, vox_size, scene_center, inside_extent, grid_res, is_leaf):
    return _C.valid_gradient_table(vox_center, vox_size, scene_center, inside_extent, grid_res, is_leaf)
