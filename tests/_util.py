"""Shared helpers for the parity tests: run the HIP path through the drop-in API and the oracle
on the same seeded inputs."""
from __future__ import annotations

import sys

import numpy as np
import torch

from oracle import binning as ob
from oracle import raster as orc
from oracle import sh as osh


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def rel_l2(a: torch.Tensor, b: torch.Tensor) -> float:
    a = a.double().cpu().reshape(-1)
    b = b.double().cpu().reshape(-1)
    return float((a - b).norm() / b.norm().clamp_min(1e-30))


def hip_raster(scene, cam, geos, rgbs, s_val, bg, need_depth=True, need_normal=True,
               track_max_w=False):
    """Run preprocess + forward through the `_C` mirror.  Returns a dict of outputs and state."""
    from svraster_amd import _C
    n_dup, geom = _C.rasterize_preprocess(cam.image_width, cam.image_height, cam.tanfovx, cam.tanfovy,
                                          cam.cx, cam.cy, cam.w2c, cam.c2w, 0, cam.near, scene.octpath,
                                          scene.vox_center, scene.vox_size, scene.is_leaf)
    out = _C.rasterize_voxels(3, 1, cam.image_width, cam.image_height, cam.tanfovx, cam.tanfovy,
                              cam.cx, cam.cy, cam.w2c, cam.c2w, bg, 0, need_depth, False, need_normal,
                              track_max_w, scene.octpath, scene.vox_center, scene.vox_size, geos, rgbs,
                              torch.empty(scene.num_voxels, 0, device=geos.device), geom, False, s_val)
    K, binning, img, color, depth, normal, T, max_w, sdf0, feat = out
    return dict(n_dup=n_dup, geom=geom, K=K, binning=binning, img=img, color=color, depth=depth,
                normal=normal, T=T, max_w=max_w)


def scene_params(scene, s_val, dev="cuda"):
    """geos [P,8], rgbs [P,3] (float32 on `dev`) from the scene's corner SDFs and SH."""
    geos = scene.geo_grid_pts.reshape(-1)[scene.vox_key].float().contiguous().to(dev)
    cam_pos = None
    return geos


def oracle_rgbs(scene, cam):
    return osh.sh_rgb(scene.sh_degree, None, scene.vox_center.cpu(), cam.c2w[:3, 3].cpu(),
                      scene.sh0.cpu().double(), scene.shs.cpu().double()).detach()


def oracle_binning(scene, cam):
    return ob.binning_for_camera(cam, scene)


def oracle_render(scene, cam, bn, geos64, rgbs64, s64, bg, window=None, need_depth=True,
                  need_normal=True, **kw):
    rc = orc.camera_from(cam)
    return orc.render(rc, bg, bn.vals, bn.ranges, bn.pre.bbox, scene.vox_center.cpu(),
                      scene.vox_size.cpu(), geos64, rgbs64, s64, need_depth=need_depth,
                      need_normal=need_normal, window=window, **kw)


def full_keys_from_vals(vals: np.ndarray, octpaths: np.ndarray, ranges: np.ndarray) -> np.ndarray:
    """Rebuild the reference's (tile << 48 | octpath ^ repeat3(octant)) keys of a sorted list."""
    vals = vals.astype(np.uint32)
    v = (vals & ((1 << 29) - 1)).astype(np.int64)
    q = (vals >> 29).astype(np.int64)
    tile = np.zeros(vals.shape[0], dtype=np.uint64)
    for t in np.nonzero(ranges[:, 1] > ranges[:, 0])[0]:
        tile[ranges[t, 0]:ranges[t, 1]] = t
    rank = octpaths.reshape(-1).astype(np.uint64)[v] ^ ob.REP3[q]
    return (tile << np.uint64(48)) | rank
