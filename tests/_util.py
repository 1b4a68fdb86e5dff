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


def centre_crop(cam, size=64):
    """(x0, y0, x1, y1) of a size x size window at the image centre, aligned to 16-pixel tiles."""
    cx, cy = cam.image_width // 2, cam.image_height // 2
    h = size // 2
    return (cx - h - cx % 16, cy - h - cy % 16, cx + h - cx % 16, cy + h - cy % 16)


def oracle_crop_grads(scene, cam, bn, window, s64: float, dC_win: torch.Tensor, grid_pts, sh0, shs, bg=(1.0, 1.0, 1.0)):
    """FP64 oracle forward + backward of one window of one view, restricted to the voxels listed
    in the window's tiles (every other voxel cannot touch those pixels).  The corner SDFs are
    gathered from the grid points and the colours evaluated from the SH coefficients of the view,
    so the gradients come back in the parameters the batched path returns:
    (out, voxel ids, grid point ids, dL/dgrid [G], dL/dsh0 [n,3], dL/dshs [n,M-1,3], dL/ds)."""
    W, H = cam.image_width, cam.image_height
    x0, y0, x1, y1 = window
    tiles_x = (W + 15) // 16
    mask29 = (1 << 29) - 1
    segs = []
    for ty in range(y0 // 16, (y1 + 15) // 16):
        for tx in range(x0 // 16, (x1 + 15) // 16):
            lo, hi = bn.ranges[ty * tiles_x + tx]
            segs.append(bn.vals[lo:hi])
    ent = np.concatenate(segs).astype(np.int64) if segs else np.zeros(0, np.int64)
    vid = np.unique(ent & mask29)
    P = scene.num_voxels
    lut = np.zeros(P, dtype=np.int64)
    lut[vid] = np.arange(vid.shape[0])
    vals = bn.vals.astype(np.int64)
    new_vals = ((vals >> 29) << 29) | lut[vals & mask29]
    vt = torch.from_numpy(vid)
    centers = scene.vox_center.cpu()[vt]
    sizes = scene.vox_size.cpu().reshape(-1)[vt]
    keys = scene.vox_key.cpu()[vt]
    gid, inv = torch.unique(keys.reshape(-1), return_inverse=True)
    grid = grid_pts.detach().reshape(-1).cpu()[gid].double().requires_grad_(True)
    geos = grid[inv.reshape(-1, 8)]
    s0 = sh0.detach().cpu()[vt].double().requires_grad_(True)
    ss = shs.detach().cpu()[vt].double().requires_grad_(True)
    rgbs = osh.sh_rgb(scene.sh_degree, None, centers, cam.c2w[:3, 3].cpu(), s0, ss)
    s = torch.tensor(float(s64), dtype=torch.float64, requires_grad=True)
    out = orc.render(orc.camera_from(cam), list(bg), new_vals, bn.ranges, bn.pre.bbox[vid], centers, sizes,
                     geos, rgbs, s, need_depth=False, need_normal=False, window=window)
    loss = (out.color * dC_win.double()).sum()
    g_grid, g_s0, g_ss, g_s = torch.autograd.grad(loss, [grid, s0, ss, s], allow_unused=True)
    z = lambda g, ref: torch.zeros_like(ref) if g is None else g
    return out, vid, gid.numpy(), z(g_grid, grid), z(g_s0, s0), z(g_ss, ss), z(g_s, s)
