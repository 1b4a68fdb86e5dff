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
                 vox_geo_mode: str = "triinterp3", density_mode: str = "sdf", outside_level=None):
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
        # The reference takes outside_level from its config; the checkpoint fixes it through
        # scene_extent = inside_extent * 2^outside_level (constructor.py:25).
        self.outside_level = (outside_level_of(state["scene_extent"], state["inside_extent"])
                              if outside_level is None else int(outside_level))
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


def outside_level_of(scene_extent, inside_extent) -> int:
    """log2(scene_extent / inside_extent), which the reference's constructor makes an integer
    (constructor.py:25); raises when the ratio is not a power of two."""
    import math
    r = float(torch.as_tensor(scene_extent).reshape(-1)[0]) / float(torch.as_tensor(inside_extent).reshape(-1)[0])
    lv = int(round(math.log2(r))) if r > 0 else -1
    if lv < 0 or abs(r - 2.0 ** lv) > 1e-5 * r:
        raise ValueError(f"scene_extent / inside_extent = {r} is not a power of two")
    return lv


def load(path: str, device="cuda", white_background: bool = True) -> LoadedModel:
    """io.py:49-86 (the reference loads with weights_only=False; the format needs no pickled
    objects, so this loader refuses them)."""
    state = torch.load(path, map_location="cpu", weights_only=True)
    if state.get("quantized", False):
        dequantize_state_dict(state)
    return LoadedModel(state, device=device, white_background=white_background)


# ----------------------------------------------------------------------------------------------
# 8-bit codebook quantization.  On-disk layout of src/sparse_voxel_gears/io.py:110-114 and
# 166-168: the corner SDFs as one {index uint8 (shape of the tensor), codebook float32[256]}
# dict, sh0 and shs as lists of such dicts, one per channel column.
# ----------------------------------------------------------------------------------------------
def quantize_state_dict(state: Dict) -> None:
    state["_geo_grid_pts"] = quantization(state["_geo_grid_pts"])
    state["_sh0"] = [quantization(v) for v in state["_sh0"].split(1, dim=1)]
    state["_shs"] = [quantization(v) for v in state["_shs"].split(1, dim=1)]


def dequantize_state_dict(state: Dict) -> None:
    state["_geo_grid_pts"] = dequantization(state["_geo_grid_pts"])
    state["_sh0"] = torch.cat([dequantization(v) for v in state["_sh0"]], dim=1)
    state["_shs"] = torch.cat([dequantization(v) for v in state["_shs"]], dim=1)


N_CODES = 256


def _run_means(sv: torch.Tensor, csum: torch.Tensor, bounds: torch.Tensor) -> torch.Tensor:
    """Mean of each run sv[bounds[k]:bounds[k+1]] (float64).  An empty run takes the sorted
    value at its start (the last value past the end), which lies between its neighbours' means,
    so the codebook stays nondecreasing."""
    n = sv.numel()
    s, e = bounds[:-1], bounds[1:]
    cnt = (e - s).to(torch.float64)
    means = (csum[e] - csum[s]) / cnt.clamp_min(1.0)
    fill = sv[s.clamp_max(n - 1)].to(torch.float64)
    return torch.where(cnt > 0, means, fill)


def quantization(src: torch.Tensor, max_iter: int = 10) -> Dict:
    """Fit a 256-word scalar codebook to the values of `src` by one-dimensional k-means
    (Lloyd's algorithm) and return {index uint8 [src.shape], codebook float32 [256]}.

    In one dimension every k-means cell is an interval, so after one sort each codeword owns a
    contiguous run of the sorted values and a Lloyd step is two vector operations: the run
    boundaries are the positions of the midpoints between consecutive codewords (a value equal
    to a midpoint joins the upper codeword), and each codeword becomes the mean of its run,
    read from a float64 prefix sum.  The fit starts from 256 runs of equal count (sizes differ by
    at most one) and stops when no boundary moves or after `max_iter` steps.  Every used
    codeword is the mean of exactly the values indexed to it; the codebook is nondecreasing.
    Runs on the host: this is checkpoint IO, not part of the rendering path."""
    shape = src.shape
    vals = src.detach().reshape(-1).to("cpu", torch.float32)
    n = vals.numel()
    if n == 0:
        return {"index": torch.zeros(shape, dtype=torch.uint8), "codebook": torch.zeros(N_CODES)}
    sv, order = torch.sort(vals)
    csum = torch.zeros(n + 1, dtype=torch.float64)
    torch.cumsum(sv.to(torch.float64), 0, out=csum[1:])
    k = torch.arange(N_CODES + 1, dtype=torch.int64)
    bounds = (k * n) // N_CODES
    code = _run_means(sv, csum, bounds)
    for _ in range(max_iter):
        mid = 0.5 * (code[:-1] + code[1:])
        inner = torch.searchsorted(sv.to(torch.float64), mid, side="left")
        nb = torch.cat([bounds[:1], inner, bounds[-1:]])
        if torch.equal(nb, bounds):
            break
        bounds = nb
        code = _run_means(sv, csum, bounds)
    run_of_sorted = torch.searchsorted(bounds[1:-1], torch.arange(n, dtype=torch.int64), side="right")
    index = torch.empty(n, dtype=torch.int64)
    index[order] = run_of_sorted
    return {"index": index.reshape(shape).to(torch.uint8), "codebook": code.to(torch.float32)}


def dequantization(qd: Dict) -> torch.Tensor:
    return qd["codebook"][qd["index"].long()]
