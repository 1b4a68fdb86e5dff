"""Worker functions of the finite-difference check on the fd_256 scene (tests/test_oracle_fd_cfg3.py).

Loss L = <g, colour> over the whole 256x256 frame with a fixed random cotangent g.  The corner
SDFs are gathered from the grid points and the colours evaluated from the SH coefficients through
oracle.sh, so derivatives are taken with respect to the model's parameters.  A parameter only
changes the pixels of the tiles whose lists hold a voxel that uses it, so each evaluation renders
the window of those tiles; the analytic gradient is the sum of autograd over horizontal strips of
the frame.  Run in worker processes (spawned), one torch thread each.
"""
from __future__ import annotations

import numpy as np
import torch

from oracle import binning as ob
from oracle import raster as orc
from oracle import sh as osh

_S = {}


def setup():
    if _S:
        return _S
    torch.set_num_threads(1)
    from svraster_amd.synthetic import make_workload
    w = make_workload("fd_256")
    sc, cam = w.scene, w.cameras[0]
    bn = ob.binning_for_camera(cam, sc)
    g = torch.Generator().manual_seed(11)
    cot = torch.rand(3, cam.image_height, cam.image_width, generator=g, dtype=torch.float64) - 0.5
    _S.update(w=w, sc=sc, cam=cam, bn=bn, cot=cot, rc=orc.camera_from(cam),
              grid=sc.geo_grid_pts.reshape(-1).double(), sh0=sc.sh0.double(), shs=sc.shs.double(),
              s=torch.tensor(float(np.float32(w.s_val)), dtype=torch.float64))
    return _S


def _render(grid, sh0, shs, window):
    S = setup()
    sc, cam = S["sc"], S["cam"]
    geos = grid[sc.vox_key]
    rgbs = osh.sh_rgb(sc.sh_degree, None, sc.vox_center, cam.c2w[:3, 3], sh0, shs)
    out = orc.render(S["rc"], [1.0, 1.0, 1.0], S["bn"].vals, S["bn"].ranges, S["bn"].pre.bbox, sc.vox_center,
                     sc.vox_size, geos, rgbs, S["s"], need_depth=False, need_normal=False, window=window)
    x0, y0, x1, y1 = out.window
    loss = (out.color * S["cot"][:, y0:y1, x0:x1]).sum()
    return loss, out


def strip_gradient(rows):
    """Autograd gradient of the loss restricted to pixel rows [rows[0], rows[1])."""
    S = setup()
    grid = S["grid"].clone().requires_grad_(True)
    sh0 = S["sh0"].clone().requires_grad_(True)
    shs = S["shs"].clone().requires_grad_(True)
    loss, _ = _render(grid, sh0, shs, (0, rows[0], S["cam"].image_width, rows[1]))
    g = torch.autograd.grad(loss, [grid, sh0, shs], allow_unused=True)
    return [torch.zeros_like(x) if gx is None else gx for gx, x in zip(g, (grid, sh0, shs))]


def _fingerprint(out):
    """Discrete state of a render: contributor counts and every pair's contribution flag."""
    return (out.n_contrib.clone(), [pa[1].clone() for pa in out.pair_alpha])


def _same(a, b):
    return torch.equal(a[0], b[0]) and len(a[1]) == len(b[1]) and all(torch.equal(x, y) for x, y in zip(a[1], b[1]))


def central_difference(task):
    """task = (kind, index, window, h): kind 'grid' (index into the grid points) or 'sh' (flat
    index into [sh0 | shs] per voxel).  Returns (numeric derivative, smooth) where smooth says
    that the base, +h and -h renders make the same discrete decisions (contributing pairs,
    contributor counts), i.e. the finite difference does not straddle a threshold."""
    kind, index, window, h = task
    S = setup()
    grid, sh0, shs = S["grid"], S["sh0"], S["shs"]
    vals = {}
    fps = []
    for sign in (0.0, 1.0, -1.0):
        g2, s0, s1 = grid, sh0, shs
        if kind == "grid":
            g2 = grid.clone()
            g2[index] += sign * h
        else:
            v, k = divmod(index, 3 * (1 + shs.shape[1]))
            row, ch = divmod(k, 3)
            if row == 0:
                s0 = sh0.clone()
                s0[v, ch] += sign * h
            else:
                s1 = shs.clone()
                s1[v, row - 1, ch] += sign * h
        with torch.no_grad():
            loss, out = _render(g2, s0, s1, window)
        vals[sign] = float(loss)
        fps.append(_fingerprint(out))
    smooth = _same(fps[0], fps[1]) and _same(fps[0], fps[2])
    return (vals[1.0] - vals[-1.0]) / (2.0 * h), smooth
