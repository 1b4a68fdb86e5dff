"""Central finite differences (FP64, h = 1e-4) of the oracle's gradients on the fd_256 scene
(20k voxels, levels 4-7, 256x256, SH degree 1): 512 randomly chosen corner SDFs (grid points)
and 256 randomly chosen SH coefficients, each through the whole chain -- grid point gather, SH
evaluation, compositing -- against the autograd gradient of the full frame.  A difference whose
+-h renders change a discrete decision (a pair crossing the 1e-5 alpha cut, a pixel crossing
the 1e-4 early stop) measures a jump, not a derivative; those are skipped and counted (at most
5 % of the sample).  Eight worker processes, about a minute of host time."""
import multiprocessing as mp
import os

import numpy as np
import torch

from tests import _fd_cfg3 as F


def test_fd_256_corner_sdf_and_sh_gradients_match_central_differences():
    S = F.setup()
    sc, cam, bn = S["sc"], S["cam"], S["bn"]
    W, H = cam.image_width, cam.image_height
    tiles_x, tiles_y = (W + 15) // 16, (H + 15) // 16
    # voxel -> tiles whose lists hold it
    vox_tiles = {}
    mask29 = (1 << 29) - 1
    for t in range(tiles_x * tiles_y):
        lo, hi = bn.ranges[t]
        for v in np.unique(bn.vals[lo:hi].astype(np.int64) & mask29).tolist():
            vox_tiles.setdefault(v, set()).add(t)
    visible = np.array(sorted(vox_tiles))
    vkey = sc.vox_key.numpy()
    rng = np.random.default_rng(5)
    pts = np.unique(vkey[visible].reshape(-1))
    grid_idx = rng.choice(pts, size=512, replace=False)
    n_sh = 3 * (1 + sc.shs.shape[1])
    sh_idx = rng.choice(visible, size=256, replace=True) * n_sh + rng.integers(0, n_sh, size=256)
    owners = {}
    for v in visible.tolist():
        for g in vkey[v].tolist():
            owners.setdefault(g, []).append(v)

    def window(voxels):
        ts = set().union(*(vox_tiles[v] for v in voxels))
        tx = [t % tiles_x for t in ts]
        ty = [t // tiles_x for t in ts]
        return (16 * min(tx), 16 * min(ty), min(W, 16 * (max(tx) + 1)), min(H, 16 * (max(ty) + 1)))

    h = 1e-4
    tasks = [("grid", int(g), window(owners[int(g)]), h) for g in grid_idx]
    tasks += [("sh", int(i), window([int(i) // n_sh]), h) for i in sh_idx]
    strips = [(y, min(H, y + 32)) for y in range(0, H, 32)]
    ctx = mp.get_context("spawn")
    n_proc = min(8, os.cpu_count() or 1)
    with ctx.Pool(n_proc, initializer=F.setup) as pool:
        parts = pool.map(F.strip_gradient, strips)
        fd = pool.map(F.central_difference, tasks, chunksize=8)
    g_grid = sum(p[0] for p in parts)
    g_sh = torch.cat([sum(p[1] for p in parts).reshape(sc.num_voxels, 3),
                      sum(p[2] for p in parts).reshape(sc.num_voxels, -1)], 1).reshape(-1)
    ana = np.array([float(g_grid[g]) for g in grid_idx] + [float(g_sh[i]) for i in sh_idx])
    num = np.array([x[0] for x in fd])
    smooth = np.array([x[1] for x in fd])
    scale = np.abs(ana).max()
    err = np.abs(num - ana)
    ok = err <= 1e-4 * np.abs(ana) + 1e-7 * scale
    n_grid_nz = int((np.abs(ana[:512]) > 1e-6 * scale).sum())
    print(f"fd_256: {smooth.sum()} of {len(tasks)} differences smooth; worst smooth error "
          f"{(err[smooth] / (np.abs(ana[smooth]) + 1e-7 * scale)).max():.2e} relative; "
          f"{n_grid_nz} of 512 corner SDFs with a non-negligible gradient")
    assert (~smooth).mean() <= 0.05
    assert ok[smooth].all(), np.nonzero(~ok & smooth)[0][:10]
    assert n_grid_nz >= 128
