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
  * total variation: corner values 0..7 by hand, ties count as "not greater", and central
    finite differences of w * sum |edge differences| away from ties; tv_sparse by construction;
  * voxel gradient: exact on a linear field (zero where |g|^2 = 2) and h times the finite
    difference of w (|g|^2 - 2)^2;
  * points loss: central finite differences of w * sum sdf^2, skipped points by hand;
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


# ----------------------------------------------------------------------------------------------
# Per-voxel regularisers (tv_compute.cu:23-159, vg_compute.cu:23-75, pl_compute.cu:14-66)
# ----------------------------------------------------------------------------------------------
def tv_grad(grid_pts, vox_key, weight, vox_size_inv, no_tv_s, tv_sparse, grad_in):
    """Total variation: for voxel n, channel c, corner i gets w * sum over the three corners
    i^1, i^2, i^4 of (+1 if v_i > v_j else -1), w = weight (* 0.01 vox_size_inv[n] unless
    no_tv_s).  tv_sparse: voxels whose 8 corners all have a zero gradient in grad_in (the
    gradient before the call) are skipped.  Returns grad_in + the contributions (float64)."""
    g = np.asarray(grid_pts, dtype=np.float64)
    if g.ndim == 1:
        g = g[:, None]
    out = np.array(grad_in, dtype=np.float64).reshape(g.shape)
    before = out.copy()
    keys = np.asarray(vox_key, dtype=np.int64)
    for n in range(keys.shape[0]):
        k = keys[n]
        for c in range(g.shape[1]):
            if tv_sparse and not np.any(before[k, c] != 0.0):
                continue
            v = g[k, c]
            w = float(weight) if no_tv_s else float(weight) * 0.01 * float(vox_size_inv[n])
            for i in range(8):
                s = sum(1 if v[i] > v[i ^ b] else -1 for b in (1, 2, 4))
                out[k[i], c] += w * s
    return out


def tv_value(grid_pts, vox_key, weight, vox_size_inv, no_tv_s):
    """The scalar the TV gradient differentiates where no two corners tie: sum over voxels of
    w * (sum over the 12 edges of |v_a - v_b|)."""
    g = np.asarray(grid_pts, dtype=np.float64).reshape(-1)
    keys = np.asarray(vox_key, dtype=np.int64)
    total = 0.0
    for n in range(keys.shape[0]):
        v = g[keys[n]]
        w = float(weight) if no_tv_s else float(weight) * 0.01 * float(vox_size_inv[n])
        for i in range(8):
            for b in (1, 2, 4):
                if i & b == 0:
                    total += w * abs(v[i] - v[i | b])
    return total


def _signs(i):
    return ((1.0 if i & 4 else -1.0), (1.0 if i & 2 else -1.0), (1.0 if i & 1 else -1.0))


def vg_grad(grid_pts, vox_key, vox_size_inv, active, weight):
    """Voxel gradient: g = sum_i s_i S_i / (4h) (s_i = +-1 per axis, corner bits x=4, y=2, z=1);
    corner i gets 4 (|g|^2 - 2) weight (g . s_i) / 4."""
    gp = np.asarray(grid_pts, dtype=np.float64).reshape(-1)
    out = np.zeros_like(gp)
    keys = np.asarray(vox_key, dtype=np.int64)
    for v in np.asarray(active, dtype=np.int64):
        S = gp[keys[v]]
        inv4h = 0.25 * float(vox_size_inv[v])
        gvec = np.zeros(3)
        for i in range(8):
            gvec += np.array(_signs(i)) * S[i]
        gvec *= inv4h
        scale = 4.0 * (gvec @ gvec - 2.0) * float(weight)
        for i in range(8):
            out[keys[v][i]] += scale * (gvec @ np.array(_signs(i))) * 0.25
    return out


def vg_value(grid_pts, vox_key, vox_size_inv, active, weight):
    """weight * (|g|^2 - 2)^2 summed over the active voxels; vg_grad equals h times its gradient
    (the reference chains with 1/4 where the derivative has 1/(4h))."""
    gp = np.asarray(grid_pts, dtype=np.float64).reshape(-1)
    keys = np.asarray(vox_key, dtype=np.int64)
    total = 0.0
    for v in np.asarray(active, dtype=np.int64):
        S = gp[keys[v]]
        gvec = np.zeros(3)
        for i in range(8):
            gvec += np.array(_signs(i)) * S[i]
        gvec *= 0.25 * float(vox_size_inv[v])
        total += float(weight) * (gvec @ gvec - 2.0) ** 2
    return total


def _point_voxel(p, res, mask, keys, owners):
    ix, iy, iz = (int(np.trunc(p[0])), int(np.trunc(p[1])), int(np.trunc(p[2])))
    if not (0 <= ix < res and 0 <= iy < res and 0 <= iz < res):
        return -1
    cell = ix + iy * res + iz * res * res
    if not mask[cell]:
        return -1
    idx = key_search(keys, cell)
    return -1 if idx < 0 else int(owners[idx])


def pl_grad(points, grid_pts, vox_key, coord, vsize, res, mask, keys, owners, weight):
    """Points loss: each point (grid cells) on a valid cell takes its cell's voxel; with the
    unclamped trilinear weights w at the point, corner i gets 2 weight sdf w_i."""
    gp = np.asarray(grid_pts, dtype=np.float64).reshape(-1)
    out = np.zeros_like(gp)
    keys_v = np.asarray(vox_key, dtype=np.int64)
    for p in np.asarray(points, dtype=np.float64):
        v = _point_voxel(p, res, mask, keys, owners)
        if v < 0:
            continue
        q = (p - np.asarray(coord[v], dtype=np.float64)) / float(vsize[v])
        w = _trilinear(q)
        sdf = float(w @ gp[keys_v[v]])
        for i in range(8):
            out[keys_v[v][i]] += 2.0 * sdf * float(weight) * w[i]
    return out


def pl_value(points, grid_pts, vox_key, coord, vsize, res, mask, keys, owners, weight):
    """weight * sum over the points of sdf^2, the scalar pl_grad differentiates."""
    gp = np.asarray(grid_pts, dtype=np.float64).reshape(-1)
    keys_v = np.asarray(vox_key, dtype=np.int64)
    total = 0.0
    for p in np.asarray(points, dtype=np.float64):
        v = _point_voxel(p, res, mask, keys, owners)
        if v < 0:
            continue
        q = (p - np.asarray(coord[v], dtype=np.float64)) / float(vsize[v])
        total += float(weight) * float(_trilinear(q) @ gp[keys_v[v]]) ** 2
    return total


def _trilinear(q):
    x, y, z = q
    w = np.empty(8)
    for i in range(8):
        w[i] = (x if i & 4 else 1.0 - x) * (y if i & 2 else 1.0 - y) * (z if i & 1 else 1.0 - z)
    return w
