"""Minimal voxel model that drives the rasterizer exactly like the reference's callers:
`vox_fn` gathers corner SDFs and evaluates SH for the visible voxels
(src/sparse_voxel_gears/renderer.py:37-96) and `render` builds RasterSettings and calls
rasterize_voxels (renderer.py:98-186, without super-sampling or auto-exposure).  The optimizer
groups follow src/sparse_voxel_gears/optimizer.py:15-27."""
from __future__ import annotations

import torch

from . import renderer as R
from .sparse_adam import SparseAdam


class VoxelModel:
    def __init__(self, scene, vox_geo_mode="triinterp3", density_mode="sdf", white_background=True):
        dev = scene.vox_center.device
        self.scene = scene
        self.vox_geo_mode = vox_geo_mode
        self.density_mode = density_mode
        self.octpath = scene.octpath
        self.octlevel = scene.octlevel
        self.vox_center = scene.vox_center
        self.vox_size = scene.vox_size
        self.vox_size_inv = 1.0 / scene.vox_size
        self.vox_key = scene.vox_key
        self.grid_pts_key = scene.grid_pts_key
        self.is_leaf = scene.is_leaf
        self.scene_center = scene.scene_center
        self.scene_extent = scene.scene_extent
        self.outside_level = 0
        self.active_sh_degree = scene.sh_degree
        self._geo_grid_pts = scene.geo_grid_pts.clone().requires_grad_(True)
        self._sh0 = scene.sh0.clone().requires_grad_(True)
        self._shs = scene.shs.clone().requires_grad_(True)
        self._subdiv_p = torch.ones(scene.num_voxels, 1, device=dev, requires_grad=True)
        self._log_s = torch.tensor(scene.log_s, dtype=torch.float32, device=dev).requires_grad_(True)
        bg = 1.0 if white_background else 0.0
        self.bg_color = torch.full((3,), bg, dtype=torch.float32, device=dev)
        self.optimizer = None

    @property
    def num_voxels(self):
        return self.vox_center.shape[0]

    def vox_fn(self, idx, cam_pos, color_mode=None, viewdir=None):
        geos = R.GatherGeoParams.apply(self.vox_geo_mode, self.vox_key, self.vox_size_inv, idx,
                                       self._geo_grid_pts)
        deg = self.active_sh_degree
        if color_mode is not None and color_mode.startswith("sh") and len(color_mode) > 2:
            deg = int(color_mode[2])
        rgbs = R.SH_eval.apply(deg, idx, self.vox_center, cam_pos, viewdir, self._sh0, self._shs)
        return {"geos": geos, "rgbs": rgbs, "subdiv_p": self._subdiv_p, "log_s": self._log_s}

    def raster_settings(self, cam, need_depth=False, need_normal=False, track_max_w=False, **opt):
        return R.RasterSettings(
            color_mode=None, vox_geo_mode=self.vox_geo_mode, density_mode=self.density_mode,
            image_width=cam.image_width, image_height=cam.image_height,
            tanfovx=cam.tanfovx, tanfovy=cam.tanfovy, cx=cam.cx, cy=cam.cy,
            w2c_matrix=cam.w2c, c2w_matrix=cam.c2w, background=self.bg_color,
            cam_mode=cam.cam_mode, near=cam.near, need_depth=need_depth, need_normal=need_normal,
            track_max_w=track_max_w, **opt)

    def render(self, cam, need_depth=False, need_normal=False, track_max_w=False, **opt):
        rs = self.raster_settings(cam, need_depth, need_normal, track_max_w, **opt)
        color, depth, normal, T, max_w, feat = R.rasterize_voxels(
            rs, self.octpath, self.vox_center, self.vox_size, self.vox_fn, self.is_leaf)
        return {"color": color, "depth": depth, "normal": normal, "T": T, "max_w": max_w}

    def make_optimizer(self, geo_lr=0.005, sh0_lr=0.01, shs_lr=0.00025, betas=(0.1, 0.99), eps=1e-15):
        groups = [
            {"params": [self._geo_grid_pts], "lr": geo_lr, "name": "_geo_grid_pts"},
            {"params": [self._sh0], "lr": sh0_lr, "name": "_sh0"},
            {"params": [self._shs], "lr": shs_lr, "name": "_shs"},
            {"params": [self._subdiv_p], "lr": 0.0, "name": "_subdiv_p"},
            {"params": [self._log_s], "lr": 0.0, "name": "_log_s"},
        ]
        self.optimizer = SparseAdam(groups, betas=betas, eps=eps)
        return self.optimizer
