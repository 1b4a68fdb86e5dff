"""Adaptive pruning and subdivision of the sparse voxel set, and the octree helpers they use,
following src/sparse_voxel_gears/adaptive.py:32-411 and src/utils/octree_utils.py:51-222.

The voxel <-> grid-point link (corner deduplication) and the mean of per-voxel corner values
into the grid points run as HIP kernels (svr_build_grid_pts_link: corner codes, a stable radix
sort, change flags and a scan; svr_agg_voxel_into_grid_pts); octpath -> ijk uses the library's
decoder.  Gathers, concatenations and the elementwise trilinear subdivision are device tensor
operations, as in the reference.

`pruning` and `subdividing` work on any object with the reference model's fields: the
per-voxel attributes octpath [N,1] int64, octlevel [N,1] int8, vox_center [N,3], vox_size [N,1],
is_leaf [N] (and subdiv_meta when present), the per-voxel parameters _sh0, _shs, _subdiv_p, the
grid-point parameter _geo_grid_pts [M,1], vox_key [N,8] and grid_pts_key [M,3], scene_center,
scene_extent, outside_level (default 0) and, optionally, an optimizer whose states are cleared
and renewed (adaptive.py:104-127).
"""
from __future__ import annotations

import ctypes
from typing import Optional

import torch

from . import _C, _lib

MAX_NUM_LEVELS = 16

PER_VOXEL_ATTRS = ("octpath", "octlevel", "vox_center", "vox_size", "subdiv_meta", "is_leaf")
PER_VOXEL_PARAMS = ("_sh0", "_shs", "_subdiv_p")
GRID_PTS_PARAMS = ("_geo_grid_pts",)


def _p(t):
    return None if t is None or t.numel() == 0 else ctypes.c_void_p(t.data_ptr())


def _stream(dev):
    return ctypes.c_void_p(torch.cuda.current_stream(dev).cuda_stream)


# ----------------------------------------------------------------------------------------------
# Octree helpers (octree_utils.py)
# ----------------------------------------------------------------------------------------------
def level_2_vox_size(scene_extent, octlevel):
    """scene_extent * 2^-level (octree_utils.py:51-60)."""
    ext = torch.as_tensor(scene_extent, dtype=torch.float32, device=octlevel.device)
    return torch.ldexp(ext, -octlevel.long())


def octpath_decoding(octpath, octlevel, scene_center, scene_extent):
    """Voxel centres and sizes of (octpath, octlevel) (octree_utils.py:99-126)."""
    octpath = octpath.reshape(-1, 1)
    octlevel = octlevel.reshape(-1, 1)
    sc = torch.as_tensor(scene_center, dtype=torch.float32, device=octpath.device).reshape(1, 3)
    ext = torch.as_tensor(scene_extent, dtype=torch.float32, device=octpath.device)
    vox_size = level_2_vox_size(ext, octlevel)
    ijk = _C.octpath_2_ijk(octpath, octlevel.to(torch.int8))
    vox_center = (sc - 0.5 * ext) + (ijk + 0.5) * vox_size
    return vox_center, vox_size


def gen_children(octpath, octlevel):
    """The eight children of each voxel, parent-major, octant order 0..7
    (octree_utils.py:166-200)."""
    octpath = octpath.reshape(-1, 1)
    octlevel = octlevel.reshape(-1, 1)
    lv = octlevel.long() + 1
    if int(lv.max()) > MAX_NUM_LEVELS:
        raise RuntimeError("Maximum level out of bound after subdivision.")
    children = torch.arange(8, dtype=torch.int64, device=octpath.device)
    shift = 3 * (MAX_NUM_LEVELS - lv)
    if bool((octpath & (children << shift)).any()):
        raise RuntimeError("octpath has bits below its level")
    path = (octpath | (children << shift)).reshape(-1, 1)
    level = (octlevel + 1).repeat_interleave(8, dim=0)
    return path, level


def build_grid_pts_link(octpath, octlevel):
    """(grid_pts_key [M,3] int64, vox_key [N,8] int64): the deduplicated level-16 corner
    coordinates of every voxel in lexicographic order, and each voxel's 8 corner indices
    (octree_utils.py:203-222)."""
    lib = _lib.load()
    dev = octpath.device
    N = int(octpath.numel())
    path = octpath.reshape(-1).contiguous()
    lv = octlevel.reshape(-1).to(torch.int8).contiguous()
    ws = torch.empty(max(int(lib.svr_grid_pts_link_workspace_bytes(N)), 1), dtype=torch.uint8, device=dev)
    vox_key = torch.empty(N, 8, dtype=torch.int64, device=dev)
    grid_key = torch.empty(max(8 * N, 1), 3, dtype=torch.int64, device=dev)
    m = ctypes.c_longlong(0)
    _lib.check(lib.svr_build_grid_pts_link(_p(path), _p(lv), N, _p(ws), _p(vox_key), _p(grid_key), ctypes.byref(m),
                                           _stream(dev)), "svr_build_grid_pts_link")
    return grid_key[:m.value].contiguous(), vox_key


def agg_voxel_into_grid_pts(num_grid_pts, vox_key, vox_val):
    """Mean of the per-voxel corner values [N,8,*] over the voxels sharing each grid point
    (adaptive.py:336-360); points no voxel names are 0."""
    lib = _lib.load()
    dev = vox_val.device
    N = int(vox_key.shape[0])
    ch = tuple(vox_val.shape[2:])
    C = 1
    for c in ch:
        C *= int(c)
    vals = vox_val.reshape(N, 8, C).float().contiguous()
    out = torch.empty(int(num_grid_pts), *ch, dtype=torch.float32, device=dev)
    count = torch.empty(max(int(num_grid_pts), 1), dtype=torch.int32, device=dev)
    _lib.check(lib.svr_agg_voxel_into_grid_pts(_p(vox_key.contiguous()), N, _p(vals), C, int(num_grid_pts),
                                               _p(count), _p(out), _stream(dev)), "svr_agg_voxel_into_grid_pts")
    return out


def subdivide_by_interp(vox_val):
    """Corner values of the 8 children of each voxel by trilinear interpolation of the parent's
    corners [N,8,*] -> [8N,8,*] (adaptive.py:365-411): child c's corner c keeps the parent
    corner, corners across one axis take the edge midpoint, across two the face centre, the
    opposite corner the voxel centre."""
    v = vox_val.contiguous()
    idx = torch.arange(8, dtype=torch.int64, device=v.device)
    out = torch.zeros([v.shape[0], 8, *v.shape[1:]], dtype=v.dtype, device=v.device)
    out[:, idx, idx] = v
    for b in (1, 2, 4):
        out[:, idx, idx ^ b] = 0.5 * (v + v[:, idx ^ b])
    for b1, b2 in ((1, 2), (1, 4), (2, 4)):
        out[:, idx, idx ^ (b1 | b2)] = 0.25 * (v + v[:, idx ^ b1] + v[:, idx ^ b2] + v[:, idx ^ (b1 | b2)])
    out[:, idx, idx ^ 7] = v.mean(1, keepdim=True)
    return out.reshape(v.shape[0] * 8, *v.shape[1:]).contiguous()


def mask_cat_perm(tensor, kept_idx=None, cat_tensor=None, perm=None):
    """Gather, append, permute (adaptive.py:317-333)."""
    if kept_idx is None and cat_tensor is None and perm is None:
        raise ValueError("mask_cat_perm needs at least one of kept_idx, cat_tensor, perm")
    dev = tensor.device
    if kept_idx is not None:
        tensor = tensor[kept_idx.to(dev)]
    if cat_tensor is not None:
        tensor = torch.cat([tensor, cat_tensor.to(dev)])
    if perm is not None:
        tensor = tensor[perm.to(dev)]
    return tensor.contiguous()


# ----------------------------------------------------------------------------------------------
# Pruning and subdivision (adaptive.py:32-72, 129-314)
# ----------------------------------------------------------------------------------------------
def _attrs(model):
    return [n for n in PER_VOXEL_ATTRS if getattr(model, n, None) is not None]


def _params(model, names):
    return [n for n in names if getattr(model, n, None) is not None]


def _clear_optimizer_states(model):
    opt = getattr(model, "optimizer", None)
    if opt is None:
        return
    for name in _params(model, PER_VOXEL_PARAMS) + _params(model, GRID_PTS_PARAMS):
        opt.state.pop(getattr(model, name), None)


def _renew_optimizer_states(model):
    opt = getattr(model, "optimizer", None)
    if opt is None:
        return
    lookup = {g.get("name"): i for i, g in enumerate(opt.param_groups)}
    for name in _params(model, PER_VOXEL_PARAMS) + _params(model, GRID_PTS_PARAMS):
        if name not in lookup:
            continue
        group = opt.param_groups[lookup[name]]
        group["params"][0] = getattr(model, name)
        opt.state[group["params"][0]] = {}


def _refresh_derived(model):
    if hasattr(model, "vox_size_inv"):
        model.vox_size_inv = 1.0 / model.vox_size


@torch.no_grad()
def pruning(model, prune_mask):
    """Drop the voxels of prune_mask [N]; per-voxel data are gathered, the grid points are
    rebuilt and their values are the mean over the kept voxels of the old corner values."""
    kept_idx = (~prune_mask.reshape(-1).bool()).nonzero(as_tuple=False).reshape(-1)
    _clear_optimizer_states(model)
    for name in _attrs(model):
        setattr(model, name, mask_cat_perm(getattr(model, name), kept_idx=kept_idx))
    for name in _params(model, PER_VOXEL_PARAMS):
        setattr(model, name, mask_cat_perm(getattr(model, name).detach(), kept_idx=kept_idx).requires_grad_())
    old_vox_key = model.vox_key
    model.grid_pts_key, model.vox_key = build_grid_pts_link(model.octpath, model.octlevel)
    M = model.grid_pts_key.shape[0]
    for name in _params(model, GRID_PTS_PARAMS):
        old = getattr(model, name).detach()
        vals = mask_cat_perm(old[old_vox_key], kept_idx=kept_idx)
        setattr(model, name, agg_voxel_into_grid_pts(M, model.vox_key, vals).requires_grad_())
    _refresh_derived(model)
    _renew_optimizer_states(model)


@torch.no_grad()
def subdividing(model, subdivide_mask):
    """Split the voxels of subdivide_mask [N] into their 8 children (appended after the kept
    voxels); voxels at depth >= 9 (relative to outside_level) stay as non-leaf parents.  Children
    copy the per-voxel parameters; their corner values come from trilinear interpolation of the
    parent's, and shared grid points take the mean over the voxels naming them."""
    subdivide_mask = subdivide_mask.reshape(-1).bool()
    depth_rel = model.octlevel.reshape(-1).long() - int(getattr(model, "outside_level", 0))
    parent_keep = subdivide_mask & (depth_rel >= 9)
    kept_idx = ((~subdivide_mask) | parent_keep).nonzero(as_tuple=False).reshape(-1)
    sub_idx = subdivide_mask.nonzero(as_tuple=False).reshape(-1)
    parent_idx = parent_keep.nonzero(as_tuple=False).reshape(-1)
    _clear_optimizer_states(model)

    path, level = gen_children(model.octpath[sub_idx], model.octlevel[sub_idx])
    center, size = octpath_decoding(path, level, model.scene_center, model.scene_extent)
    special = {"octpath": path.reshape(-1, *model.octpath.shape[1:]),
               "octlevel": level.to(model.octlevel.dtype).reshape(-1, *model.octlevel.shape[1:]),
               "vox_center": center, "vox_size": size.reshape(-1, *model.vox_size.shape[1:])}
    for name in _attrs(model):
        ori = getattr(model, name)
        if name in special:
            sub = special[name]
        elif name == "is_leaf":
            ori = ori.clone()
            ori[parent_idx] = 0
            sub = torch.ones_like(ori[sub_idx]).repeat_interleave(8, dim=0)
        else:
            sub = ori[sub_idx].repeat_interleave(8, dim=0)
        setattr(model, name, mask_cat_perm(ori, kept_idx=kept_idx, cat_tensor=sub))
    for name in _params(model, PER_VOXEL_PARAMS):
        ori = getattr(model, name).detach()
        sub = ori[sub_idx].repeat_interleave(8, dim=0)
        setattr(model, name, mask_cat_perm(ori, kept_idx=kept_idx, cat_tensor=sub).requires_grad_())
    old_vox_key = model.vox_key
    model.grid_pts_key, model.vox_key = build_grid_pts_link(model.octpath, model.octlevel)
    M = model.grid_pts_key.shape[0]
    for name in _params(model, GRID_PTS_PARAMS):
        old = getattr(model, name).detach()
        vox_val = old[old_vox_key]
        sub = subdivide_by_interp(vox_val[sub_idx])
        vals = mask_cat_perm(vox_val, kept_idx=kept_idx, cat_tensor=sub)
        setattr(model, name, agg_voxel_into_grid_pts(M, model.vox_key, vals).requires_grad_())
    _refresh_derived(model)
    _renew_optimizer_states(model)
