"""Checkpoint save / load in the reference's on-disk format (src/sparse_voxel_gears/io.py:18-170),
with its optional 8-bit codebook quantization of the corner SDFs and SH coefficients.

A checkpoint is a torch-saved dict: active_sh_degree, ss, scene_center, inside_extent,
scene_extent, octpath, octlevel, _geo_grid_pts, _sh0, _shs, _log_s, is_leaf, quantized.  When
quantized, _geo_grid_pts is {index uint8, codebook float[256]} and _sh0 / _shs are lists of such
dicts, one per channel column.  Loading uses torch.load(weights_only=True) (the format holds only
tensors, numbers, lists and dicts), rebuilds centres and sizes from the paths and the
voxel <-> grid-point link on the GPU (svraster_amd.adaptive).
"""
from __future__ import annotations

import os
from typing import Dict

import torch

from . import adaptive


def save(model, path: str, quantize: bool = False) -> None:
    """io.py:18-47."""
    d = os.path.dirname(path)
    if d:
        os.makedirs(d, exist_ok=True)
    dev = model.octpath.device

    def t(x):
        return torch.as_tensor(x, dtype=torch.float32, device=dev).detach().contiguous()

    state: Dict = {
        "active_sh_degree": int(model.active_sh_degree),
        "ss": float(getattr(model, "ss", 1.0)),
        "scene_center": t(model.scene_center),
        "inside_extent": t(getattr(model, "inside_extent", model.scene_extent)),
        "scene_extent": t(model.scene_extent),
        "octpath": model.octpath.detach().contiguous(),
        "octlevel": model.octlevel.detach().contiguous(),
        "_geo_grid_pts": model._geo_grid_pts.detach().contiguous(),
        "_sh0": model._sh0.detach().contiguous(),
        "_shs": model._shs.detach().contiguous(),
        "_log_s": model._log_s.detach().contiguous(),
        "is_leaf": model.is_leaf.detach().contiguous(),
    }
    if quantize:
        quantize_state_dict(state)
    state["quantized"] = bool(quantize)
    torch.save(_to_cpu(state), path)


def _to_cpu(x):
    if torch.is_tensor(x):
        return x.cpu()
    if isinstance(x, dict):
        return {k: _to_cpu(v) for k, v in x.items()}
    if isinstance(x, list):
        return [_to_cpu(v) for v in x]
    return x


class LoadedModel:
    """The fields a loaded checkpoint provides (io.py:49-86), with the VoxelModel render path."""

    def __init__(self, state: Dict, device="cuda", white_background: bool = True,
                 vox_geo_mode: str = "triinterp3", density_mode: str = "sdf"):
        from .model import VoxelModel
        dev = torch.device(device)
        self.active_sh_degree = int(state["active_sh_degree"])
        self.ss = float(state["ss"])
        self.scene_center = state["scene_center"].to(dev)
        self.inside_extent = state["inside_extent"].to(dev)
        self.scene_extent = state["scene_extent"].to(dev)
        self.octpath = state["octpath"].to(dev)
        self.octlevel = state["octlevel"].to(dev).to(torch.int8)
        self.vox_center, self.vox_size = adaptive.octpath_decoding(self.octpath, self.octlevel, self.scene_center,
                                                                   self.scene_extent)
        self.grid_pts_key, self.vox_key = adaptive.build_grid_pts_link(self.octpath, self.octlevel)
        self._geo_grid_pts = state["_geo_grid_pts"].to(dev).requires_grad_()
        self._sh0 = state["_sh0"].to(dev).requires_grad_()
        self._shs = state["_shs"].to(dev).requires_grad_()
        self._log_s = state["_log_s"].to(dev).requires_grad_()
        self.is_leaf = state["is_leaf"].to(dev)
        N = self.octpath.shape[0]
        self._subdiv_p = torch.full([N, 1], 1.0, dtype=torch.float32, device=dev).requires_grad_()
        self.subdiv_meta = torch.zeros([N, 1], dtype=torch.float32, device=dev)
        self.vox_size_inv = 1.0 / self.vox_size
        self.outside_level = 0
        bg = 1.0 if white_background else 0.0
        self.bg_color = torch.full((3,), bg, dtype=torch.float32, device=dev)
        self.vox_geo_mode, self.density_mode = vox_geo_mode, density_mode
        self.optimizer = None
        # rendering: the drop-in path of VoxelModel, bound to these fields
        self.vox_fn = VoxelModel.vox_fn.__get__(self)
        self.raster_settings = VoxelModel.raster_settings.__get__(self)
        self.render = VoxelModel.render.__get__(self)

    @property
    def num_voxels(self):
        return int(self.octpath.shape[0])

    @property
    def sh_degree(self):
        return self.active_sh_degree


def load(path: str, device="cuda", white_background: bool = True) -> LoadedModel:
    """io.py:49-86 (the reference loads with weights_only=False; the format needs no pickled
    objects, so this loader refuses them)."""
    state = torch.load(path, map_location="cpu", weights_only=True)
    if state.get("quantized", False):
        dequantize_state_dict(state)
    return LoadedModel(state, device=device, white_background=white_background)


# ----------------------------------------------------------------------------------------------
# 8-bit codebook quantization (io.py:106-170)
# ----------------------------------------------------------------------------------------------
def quantize_state_dict(state: Dict) -> None:
    state["_geo_grid_pts"] = quantization(state["_geo_grid_pts"])
    state["_sh0"] = [quantization(v) for v in state["_sh0"].split(1, dim=1)]
    state["_shs"] = [quantization(v) for v in state["_shs"].split(1, dim=1)]


def dequantize_state_dict(state: Dict) -> None:
    state["_geo_grid_pts"] = dequantization(state["_geo_grid_pts"])
    state["_sh0"] = torch.cat([dequantization(v) for v in state["_sh0"]], dim=1)
    state["_shs"] = torch.cat([dequantization(v) for v in state["_shs"]], dim=1)


def quantization(src: torch.Tensor, max_iter: int = 10) -> Dict:
    """256-entry codebook: start from the 256 upper quantiles (257 evenly spaced order
    statistics, the first dropped), assign every value to the nearer of its two bracketing
    entries, then up to max_iter rounds of (entries = mean of their values; move a value one
    entry toward a strictly nearer neighbour); a final mean.  io.py:116-167."""
    shape = src.shape
    vals = src.flatten().contiguous().float()
    n = vals.numel()
    order = vals.argsort()
    q = (torch.linspace(0, 1, 257, device=vals.device) * (n - 1)).long().clamp_(0, n - 1)
    codebook = vals[order[q]].contiguous()
    codebook[0] = -torch.inf
    ind = torch.searchsorted(codebook, vals)
    codebook = codebook[1:]
    ind = (ind - 1).clamp_(0, 255)
    diff_l = (vals - codebook[ind - 1]).abs()
    diff_m = (vals - codebook[ind]).abs()
    ind = ind - 1 + (diff_m < diff_l)
    ind.clamp_(0, 255)
    for _ in range(max_iter):
        codebook = torch.zeros_like(codebook).index_reduce_(0, ind, vals, reduce="mean", include_self=False)
        diff_l = (vals - codebook[ind - 1]).abs()
        diff_r = (vals - codebook[(ind + 1).clamp_max(255)]).abs()
        diff_m = (vals - codebook[ind]).abs()
        upd = torch.minimum(diff_l, diff_r) < diff_m
        if int(upd.sum()) == 0:
            break
        shift = (diff_r < diff_l).long() * 2 - 1
        ind[upd] += shift[upd]
        ind.clamp_(0, 255)
    codebook = torch.zeros_like(codebook).index_reduce_(0, ind, vals, reduce="mean", include_self=False)
    return {"index": ind.reshape(shape).to(torch.uint8), "codebook": codebook}


def dequantization(qd: Dict) -> torch.Tensor:
    return qd["codebook"][qd["index"].long()]
