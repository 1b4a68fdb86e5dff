"""numpy FP64 restatement of the grid smoothness backward and the valid-gradient table
(test infrastructure only).

valid_gradient_table: cuda/src/utils.cu:190-252 -- voxels whose centre lies in the region,
leaves with >= 0.9 grid cells and non-leaves with 0.9..1.5 cells, rasterised into res^3 cells
[round((min - lo)/(hi - lo) res), round((max - lo)/(hi - lo) res)).  The reference appends
with an atomic counter and the caller sorts by key (octree_utils.py:293-315); here the table is
produced in voxel order and then stably sorted by key.

grid eikonal  cuda/src/ge_compute.cu:14-117
grid Laplacian cuda/src/ls_compute.cu:14-109
sampling helpers cuda/src/auxiliary.h:363-455 (neighbour validity, midpoint binary search,
clamped trilinear weights, corner scatter).  The per-key random offset is an explicit input.

Pinning (tests/test_oracle_misc.py):
  * eikonal on an exactly linear SDF f = a.x (corner values sampled from it, so trilinear
    interpolation is exact): the central difference is exact, |grad| = |a|, and the result
    must equal the closed form 2 (g - 1) w scattered with +-(a_i/|a|) 0.5 inv to the neighbour
    samples -- in particular it vanishes when |a| * 0.5 * inv = 1;
  * Laplacian on a linear field has zero residual -> zero gradient; on f = x^2 the residual is
    the constant 2 (cell units) -> closed-form magnitude;
  * both equal central finite differences of the scalar losses they differentiate
    (eikonal: the reference scatters dL/d(diff) without the 0.5 of the central difference, i.e.
    twice the derivative of w (g - 1)^2 -- restated and checked as such; Laplacian:
    w 0.5 inv^2 |r|).
"""
from __future__ import annotations

import numpy as np

DIRS6 = [(1, 0, 0), (-1, 0, 0), (0, 1, 0), (0, -1, 0), (0, 0, 1), (0, 0, -1)]


def valid_gradient_table(center, size, scene_center, extent, res_pow2, is_leaf):
    center = np.asarray(center, dtype=np.float32).reshape(-1, 3)
    size = np.asarray(size, dtype=np.float32).reshape(-1)
    is_leaf = np.asarray(is_leaf).reshape(-1).astype(bool)
    sc = np.asarray(scene_center, dtype=np.float32).reshape(3)
    ext = np.float32(extent)
    res = 1 << res_pow2
    mask = np.zeros(res ** 3, dtype=bool)
    keys, owners = [], []
    lo = sc - np.float32(0.5) * ext
    hi = sc + np.float32(0.5) * ext
    cells = (size / ext) * np.float32(res)
    for i in range(center.shape[0]):
        if (not is_leaf[i]) and (cells[i] > 1.5 or cells[i] < 0.9):
            continue
        if is_leaf[i] and cells[i] < 0.9:
            continue
        c = center[i]
        if np.any(c < lo) or np.any(c > hi):
            continue
        a = c - np.float32(0.5) * size[i]
        b = c + np.float32(0.5) * size[i]
        # roundf: half away from zero
        r0 = [int(np.sign(v) * np.floor(abs(v) + 0.5)) for v in ((a - lo) / (hi - lo) * np.float32(res))]
        r1 = [int(np.sign(v) * np.floor(abs(v) + 0.5)) for v in ((b - lo) / (hi - lo) * np.float32(res))]
        for x in range(r0[0], r1[0]):
            for y in range(r0[1], r1[1]):
                for z in range(r0[2], r1[2]):
                    k = x + res * (y + res * z)
                    mask[k] = True
                    keys.append(k)
                    owners.append(i)
    keys = np.asarray(keys, dtype=np.int64)
    owners = np.asarray(owners, dtype=np.int64)
    return mask, keys, owners


def key_search(keys, key):
    l, r = 0, len(keys) - 1
    while l <= r:
        m = (l + r) // 2
        if keys[m] == key:
            return m
        if keys[m] < key:
            l = m + 1
        else:
            r = m - 1
    return -1


def _neighbours_valid(x, y, z, res, mask):
    if x <= 0 or x >= res - 1 or y <= 0 or y >= res - 1 or z <= 0 or z >= res - 1:
        return False
    k = x + y * res + z * res * res
    return all(mask[k + d] for d in (1, -1, res, -res, res * res, -res * res))


def _sample(grid_pts, vox_key, coord, vsize, keys, owners, res, gx, gy, gz, rv):
    idx = key_search(keys, gx + gy * res + gz * res * res)
    if idx < 0:
        return 0.0, None, None
    v = int(owners[idx])
    q = (np.array([gx, gy, gz], dtype=np.float64) + rv - coord[v]) / vsize[v]
    wx, wy, wz = (1 - q[0], q[0]), (1 - q[1], q[1]), (1 - q[2], q[2])
    w = np.array([wx[i >> 2 & 1] * wy[i >> 1 & 1] * wz[i & 1] for i in range(8)])
    w = np.clip(w, 0.0, 1.0)
    sdf = float((grid_pts[vox_key[v]] * w).sum())
    return sdf, v, w


def grid_loss_grad(kind, grid_pts, vox_key, coord, vsize, res, mask, keys, owners, active, weight,
                   inv, randv):
    """Accumulated corner gradient [M] of the eikonal ("ge") or Laplacian ("ls") backward."""
    grid_pts = np.asarray(grid_pts, dtype=np.float64).reshape(-1)
    vox_key = np.asarray(vox_key, dtype=np.int64)
    coord = np.asarray(coord, dtype=np.float64).reshape(-1, 3)
    vsize = np.asarray(vsize, dtype=np.float64).reshape(-1)
    grad = np.zeros_like(grid_pts)
    for t, a in enumerate(active):
        gk = int(keys[a])
        x, y, z = gk % res, (gk // res) % res, gk // (res * res)
        if not mask[gk] or not _neighbours_valid(x, y, z, res, mask):
            continue
        rv = np.asarray(randv[t], dtype=np.float64)
        offs = DIRS6 if kind == "ge" else [(0, 0, 0)] + DIRS6
        samp = [_sample(grid_pts, vox_key, coord, vsize, keys, owners, res, x + d[0], y + d[1], z + d[2], rv)
                for d in offs]
        sd = [s[0] for s in samp]
        if kind == "ge":
            g = np.array([(sd[0] - sd[1]) * 0.5, (sd[2] - sd[3]) * 0.5, (sd[4] - sd[5]) * 0.5])
            gn = np.sqrt((g * g).sum() + 1e-6)
            dLdg = 2.0 * (gn * 0.5 * inv - 1.0) * weight
            l = dLdg * g / gn * 0.5 * inv
            coeffs = [l[0], -l[0], l[1], -l[1], l[2], -l[2]]
        else:
            r = np.array([sd[1] + sd[2] - 2 * sd[0], sd[3] + sd[4] - 2 * sd[0], sd[5] + sd[6] - 2 * sd[0]])
            lr = np.sqrt((r * r).sum() + 1e-10)
            l = r / lr * weight * 0.5 * inv * inv
            coeffs = [-2.0 * l.sum(), l[0], l[0], l[1], l[1], l[2], l[2]]
        for (s, v, w), c in zip(samp, coeffs):
            if v is None:
                continue
            np.add.at(grad, vox_key[v], c * w)
    return grad


def grid_loss_value(kind, grid_pts, vox_key, coord, vsize, res, mask, keys, owners, active, weight,
                    inv, randv):
    """The scalar each backward differentiates (see module docstring) -- for finite differences."""
    grid_pts = np.asarray(grid_pts, dtype=np.float64).reshape(-1)
    total = 0.0
    for t, a in enumerate(active):
        gk = int(keys[a])
        x, y, z = gk % res, (gk // res) % res, gk // (res * res)
        if not mask[gk] or not _neighbours_valid(x, y, z, res, mask):
            continue
        rv = np.asarray(randv[t], dtype=np.float64)
        offs = DIRS6 if kind == "ge" else [(0, 0, 0)] + DIRS6
        sd = [_sample(grid_pts, vox_key, coord, vsize, keys, owners, res, x + d[0], y + d[1], z + d[2], rv)[0]
              for d in offs]
        if kind == "ge":
            g = np.array([(sd[0] - sd[1]), (sd[2] - sd[3]), (sd[4] - sd[5])]) * 0.5
            gn = np.sqrt((g * g).sum() + 1e-6)
            total += 2.0 * weight * (gn * 0.5 * inv - 1.0) ** 2   # x2: the reference's scaling
        else:
            r = np.array([sd[1] + sd[2] - 2 * sd[0], sd[3] + sd[4] - 2 * sd[0], sd[5] + sd[6] - 2 * sd[0]])
            total += weight * 0.5 * inv * inv * np.sqrt((r * r).sum() + 1e-10)
    return total
