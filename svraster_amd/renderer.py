"""Differentiable sparse-voxel SDF rasterizer, API-compatible with the reference's
cuda/svraster_cuda/renderer.py (RasterSettings 16-41, rasterize_voxels 44-152,
_RasterizeVoxels 155-299, SH_eval 302-367, GatherGeoParams 370-406, mark_* 438-547).

The compute runs in libsvraster_hip.so through the `_C` mirror module; this file only holds
the autograd plumbing, argument checks and the per-view call order:
  preprocess -> visible voxel list -> vox_fn (corner gather + SH) -> s = exp(10 log_s)
  -> count/duplicate/sort/ranges/composite (forward) -> reverse composite (backward).
"""
from __future__ import annotations

from typing import NamedTuple

import torch

from . import _C
from .meta import CamModes, DensityModes, VoxGeoModes
from .render_neus import render_neus  # noqa: F401  (import compatibility)


class RasterSettings(NamedTuple):
    color_mode: str
    vox_geo_mode: str
    density_mode: str
    image_width: int
    image_height: int
    tanfovx: float
    tanfovy: float
    cx: float
    cy: float
    w2c_matrix: torch.Tensor
    c2w_matrix: torch.Tensor
    background: torch.Tensor
    cam_mode: str = "persp"
    near: float = 0.02
    need_depth: bool = False
    need_normal: bool = False
    track_max_w: bool = False
    lambda_R_concen: float = 0
    lambda_ascending: float = 0
    lambda_dist: float = 0
    gt_color: torch.Tensor = torch.empty(0)
    vox_feats: torch.Tensor = torch.empty([0, 0])
    debug: bool = False


def rasterize_voxels(raster_settings: RasterSettings, octree_paths: torch.Tensor,
                     vox_centers: torch.Tensor, vox_lengths: torch.Tensor, vox_fn,
                     is_leaf: torch.Tensor):
    """Render one view.  Returns (color, depth, normal, T, max_w, feat)."""
    rs = raster_settings
    if not isinstance(rs, RasterSettings):
        raise Exception("Expect RasterSettings as first argument.")
    if rs.vox_geo_mode not in VoxGeoModes:
        raise NotImplementedError("Unknown voxel geo mode.")
    if rs.density_mode not in DensityModes:
        raise NotImplementedError("Unknown density mode.")
    if rs.cam_mode not in CamModes:
        raise NotImplementedError("Unknow camera mode.")

    N = octree_paths.numel()
    device = octree_paths.device
    if vox_centers.shape[0] != N or vox_lengths.numel() != N:
        raise Exception("Size mismatched.")
    if vox_centers.ndim != 2 or vox_centers.shape[1] != 3:
        raise Exception("Expect vox_centers in shape [N, 3].")
    for t in (rs.w2c_matrix, rs.c2w_matrix, rs.background, vox_centers, vox_lengths):
        if t.device != device:
            raise Exception("Device mismatch.")
    if rs.vox_feats.requires_grad:
        raise NotImplementedError

    n_duplicates, geomBuffer = _C.rasterize_preprocess(
        rs.image_width, rs.image_height, rs.tanfovx, rs.tanfovy, rs.cx, rs.cy,
        rs.w2c_matrix, rs.c2w_matrix, CamModes[rs.cam_mode], rs.near,
        octree_paths, vox_centers, vox_lengths, is_leaf, rs.debug)
    in_frusts_idx = torch.where(n_duplicates > 0)[0]

    cam_pos = rs.c2w_matrix[:3, 3]
    vox_params = vox_fn(in_frusts_idx, cam_pos, rs.color_mode)
    geos = vox_params["geos"]
    rgbs = vox_params["rgbs"]
    subdiv_p = vox_params["subdiv_p"]
    s_val = torch.exp(vox_params["log_s"] * 10.0)

    if geos.shape != (N, 8):
        raise Exception(f"Expect geos in ({N}, 8) but got", geos.shape)
    if rgbs.shape[0] != N:
        raise Exception(f"Expect rgbs in ({N}, 3) but got", rgbs.shape)
    if subdiv_p.shape[0] != N:
        raise Exception(f"Expect subdiv_p in ({N}, 1) but got", subdiv_p.shape)
    for t, name in ((geos, "geos"), (rgbs, "rgbs"), (subdiv_p, "subdiv_p")):
        if t.device != device:
            raise Exception(f"Device mismatch: {name}.")
    if rs.lambda_R_concen > 0:
        gt = rs.gt_color
        if gt.ndim != 3 or tuple(gt.shape) != (3, rs.image_height, rs.image_width):
            raise Exception("Except gt_color in shape of [3, H, W]")
        if gt.device != device:
            raise Exception("Device mismatch.")

    return _RasterizeVoxels.apply(rs, geomBuffer, octree_paths, vox_centers, vox_lengths, geos,
                                  rgbs, rs.vox_feats, subdiv_p, s_val)


class _RasterizeVoxels(torch.autograd.Function):
    @staticmethod
    def forward(ctx, raster_settings, geomBuffer, octree_paths, vox_centers, vox_lengths, geos,
                rgbs, vox_feats, subdiv_p, s_val):
        rs = raster_settings
        (num_rendered, binningBuffer, imgBuffer, out_color, out_depth, out_normal, out_T, max_w,
         out_sdf0, out_feat) = _C.rasterize_voxels(
            VoxGeoModes[rs.vox_geo_mode], DensityModes[rs.density_mode], rs.image_width,
            rs.image_height, rs.tanfovx, rs.tanfovy, rs.cx, rs.cy, rs.w2c_matrix, rs.c2w_matrix,
            rs.background, CamModes[rs.cam_mode], rs.need_depth, rs.lambda_dist > 0,
            rs.need_normal, rs.track_max_w, octree_paths, vox_centers, vox_lengths, geos, rgbs,
            vox_feats, geomBuffer, rs.debug, s_val)
        ctx.raster_settings = rs
        ctx.num_rendered = num_rendered
        ctx.save_for_backward(octree_paths, vox_centers, vox_lengths, geos, rgbs, geomBuffer,
                              binningBuffer, imgBuffer, out_T, out_depth, out_normal, s_val, out_sdf0)
        ctx.mark_non_differentiable(max_w)
        return out_color, out_depth, out_normal, out_T, max_w, out_feat

    @staticmethod
    def backward(ctx, dL_dout_color, dL_dout_depth, dL_dout_normal, dL_dout_T, dL_dmax_w, dL_dout_feat):
        rs = ctx.raster_settings
        (octree_paths, vox_centers, vox_lengths, geos, rgbs, geomBuffer, binningBuffer, imgBuffer,
         out_T, out_depth, out_normal, s_val, out_sdf0) = ctx.saved_tensors
        if rs.cam_mode == "ortho":
            raise NotImplementedError("Backward pass of orthographic projection is not implemented.")
        dL_dgeos, dL_drgbs, subdiv_p_bw, dL_ds_val = _C.rasterize_voxels_backward(
            ctx.num_rendered, VoxGeoModes[rs.vox_geo_mode], DensityModes[rs.density_mode],
            rs.image_width, rs.image_height, rs.tanfovx, rs.tanfovy, rs.cx, rs.cy, rs.w2c_matrix,
            rs.c2w_matrix, rs.background, octree_paths, vox_centers, vox_lengths, geos, rgbs,
            geomBuffer, binningBuffer, imgBuffer, out_T, dL_dout_color, dL_dout_depth,
            dL_dout_normal, dL_dout_T, rs.lambda_R_concen, rs.gt_color, rs.lambda_ascending,
            rs.lambda_dist, rs.need_depth, rs.need_normal, out_depth, out_normal, rs.debug, s_val,
            out_sdf0)
        dL_drgbs = dL_drgbs.view(rgbs.shape).contiguous()
        return (None, None, None, None, None, dL_dgeos, dL_drgbs, None, subdiv_p_bw,
                dL_ds_val.view_as(s_val))


class SH_eval(torch.autograd.Function):
    @staticmethod
    def forward(ctx, active_sh_degree, idx, vox_centers, cam_pos, viewdir, sh0, shs):
        for t in (vox_centers, cam_pos, viewdir):
            if torch.is_tensor(t) and t.requires_grad:
                raise NotImplementedError
        if idx is None:
            idx = torch.empty(0, dtype=torch.int64)
        if viewdir is not None:
            vox_centers = viewdir
            cam_pos = torch.zeros_like(cam_pos)
        rgbs = _C.sh_compute(active_sh_degree, idx, vox_centers, cam_pos, sh0, shs)
        ctx.active_sh_degree = active_sh_degree
        ctx.M = 1 + shs.shape[1]
        ctx.save_for_backward(idx, vox_centers, cam_pos, rgbs)
        return rgbs

    @staticmethod
    def backward(ctx, dL_drgbs):
        idx, vox_centers, cam_pos, rgbs = ctx.saved_tensors
        dL_dsh0, dL_dshs = _C.sh_compute_bw(ctx.active_sh_degree, ctx.M, idx, vox_centers, cam_pos,
                                            rgbs, dL_drgbs)
        return None, None, None, None, None, dL_dsh0, dL_dshs


class GatherGeoParams(torch.autograd.Function):
    @staticmethod
    def forward(ctx, vox_geo_mode, vox_key, vox_size_inv, care_idx, grid_pts):
        assert vox_key.ndim == 2 and vox_key.shape[1] == 8
        assert care_idx.ndim == 1
        assert grid_pts.shape[0] == grid_pts.numel()
        if not vox_geo_mode.startswith("triinterp"):
            raise NotImplementedError(f"Unknow vox_geo_mode: {vox_geo_mode}")
        geo_params = _C.gather_triinterp_geo_params(vox_key, care_idx, grid_pts)
        ctx.vox_geo_mode = vox_geo_mode
        ctx.num_grid_pts = grid_pts.numel()
        ctx.save_for_backward(vox_key, vox_size_inv, care_idx)
        return geo_params

    @staticmethod
    def backward(ctx, dL_dgeo_params):
        vox_key, vox_size_inv, care_idx = ctx.saved_tensors
        if not ctx.vox_geo_mode.startswith("triinterp"):
            raise NotImplementedError(f"Unknow vox_geo_mode: {ctx.vox_geo_mode}")
        dL_dgrid_pts = _C.gather_triinterp_geo_params_bw(vox_key, care_idx, ctx.num_grid_pts,
                                                         dL_dgeo_params)
        return None, None, None, None, dL_dgrid_pts


class GatherFeatParams(torch.autograd.Function):
    @staticmethod
    def forward(ctx, vox_key, care_idx, grid_pts):
        return _C.gather_triinterp_feat_params(vox_key, care_idx, grid_pts)

    @staticmethod
    def backward(ctx, dL_dfeat_params):
        raise NotImplementedError


# ----------------------------------------------------------------------------------------------
# Visibility helpers built on the preprocess kernel (used at initialisation / adaptation time).
# ----------------------------------------------------------------------------------------------
def mark_n_duplicates(image_width, image_height, tanfovx, tanfovy, cx, cy, w2c_matrix, c2w_matrix,
                      near, octree_paths, vox_centers, vox_lengths, cam_mode="persp",
                      return_buffer=False, debug=False, is_leaf: torch.Tensor = None):
    if is_leaf is None:
        is_leaf = torch.ones(octree_paths.numel(), dtype=torch.uint8, device=octree_paths.device)
    n_duplicates, geomBuffer = _C.rasterize_preprocess(
        image_width, image_height, tanfovx, tanfovy, cx, cy, w2c_matrix, c2w_matrix,
        CamModes[cam_mode], near, octree_paths, vox_centers, vox_lengths, is_leaf, debug)
    if return_buffer:
        return n_duplicates, geomBuffer
    return n_duplicates


def _visible_rates(cam, octree_paths, vox_centers, vox_lengths, near):
    n_dup = mark_n_duplicates(cam.image_width, cam.image_height, cam.tanfovx, cam.tanfovy, cam.cx,
                              cam.cy, cam.w2c, cam.c2w, near, octree_paths, vox_centers, vox_lengths)
    zdist = ((vox_centers - cam.position) * cam.lookat).sum(-1)
    vis = torch.where((n_dup > 0) & (zdist > near))[0]
    pix_size = 2 * cam.tanfovx / cam.image_width
    rate = vox_lengths.reshape(-1)[vis] / (zdist[vis] * pix_size)
    return vis, rate


def mark_min_samp_rate(cameras, octree_paths, vox_centers, vox_lengths, near=0.02):
    big = 1e30
    out = torch.full([len(octree_paths)], big, dtype=torch.float32, device=vox_centers.device)
    for cam in cameras:
        vis, rate = _visible_rates(cam, octree_paths, vox_centers, vox_lengths, near)
        out[vis] = torch.minimum(out[vis], rate)
    out[out >= big] = 0
    return out


def mark_max_samp_rate(cameras, octree_paths, vox_centers, vox_lengths, near=0.02):
    out = torch.zeros([len(octree_paths)], dtype=torch.float32, device=vox_centers.device)
    for cam in cameras:
        vis, rate = _visible_rates(cam, octree_paths, vox_centers, vox_lengths, near)
        out[vis] = torch.maximum(out[vis], rate)
    return out


def mark_avg_samp_rate(cameras, octree_paths, vox_centers, vox_lengths, near=0.02):
    total = torch.zeros([len(octree_paths)], dtype=torch.float32, device=vox_centers.device)
    cnt = torch.zeros_like(total)
    for cam in cameras:
        vis, rate = _visible_rates(cam, octree_paths, vox_centers, vox_lengths, near)
        total[vis] += rate
        cnt[vis] += 1
    return total / cnt.clamp_min(1)


def mark_near(cameras, octree_paths, vox_centers, vox_lengths, near=0.2):
    out = torch.zeros([len(octree_paths)], dtype=torch.bool, device=vox_centers.device)
    for cam in cameras:
        n_dup = mark_n_duplicates(cam.image_width, cam.image_height, cam.tanfovx, cam.tanfovy,
                                  cam.cx, cam.cy, cam.w2c, cam.c2w, near, octree_paths,
                                  vox_centers, vox_lengths)
        vis = torch.where(n_dup > 0)[0]
        zdist = ((vox_centers[vis] - cam.position) * cam.lookat).sum(-1)
        out[vis] |= zdist <= near + vox_lengths.reshape(-1)[vis]
    return out
