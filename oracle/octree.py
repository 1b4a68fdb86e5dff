"""Pure-Python restatement of the octree index codec (test infrastructure only).

cuda/src/utils.cu:129-189: octpath = sum over levels l of (i_l << 2 | j_l << 1 | k_l) << 3 l, then
shifted left by 3 (16 - level) so the path is left-aligned in 48 bits; the inverse undoes it.

Pinning (tests/test_oracle_misc.py): hand-computed values (e.g. level 1 child 5 -> 5 << 45),
round trips, and the property that the left-aligned path order equals depth-first (Morton)
order of the cells.

Topology changes (src/utils/octree_utils.py:129-222, src/sparse_voxel_gears/adaptive.py:32-411),
numpy: children of a voxel, the voxel <-> grid-point link (np.unique over the level-16 corner
rows, the library routine the reference's torch.unique(dim=0) computes), trilinear subdivision
of corner values, the mean of per-voxel corner values into grid points, and the prune /
subdivide procedures on a dict of arrays.  Pinned by: children tile their parent exactly;
subdividing the corner values of a linear field gives the field at the children's corners;
the link reproduces every voxel's corner coordinates; the mean of a constant is the constant.
"""
import numpy as np

MAX_NUM_LEVELS = 16


def ijk_to_octpath(i: int, j: int, k: int, level: int) -> int:
    path = 0
    for l in range(level):
        bits = (((i >> l) & 1) << 2) | (((j >> l) & 1) << 1) | ((k >> l) & 1)
        path |= bits << (3 * l)
    return path << (3 * (MAX_NUM_LEVELS - level))


def octpath_to_ijk(path: int, level: int):
    path >>= 3 * (MAX_NUM_LEVELS - level)
    i = j = k = 0
    for l in range(level):
        b = path & 7
        i |= ((b >> 2) & 1) << l
        j |= ((b >> 1) & 1) << l
        k |= (b & 1) << l
        path >>= 3
    return i, j, k


# ----------------------------------------------------------------------------------------------
# Topology changes
# ----------------------------------------------------------------------------------------------
def ijk_array(octpath, octlevel):
    return np.array([octpath_to_ijk(int(p), int(l)) for p, l in zip(np.ravel(octpath), np.ravel(octlevel))],
                    dtype=np.int64).reshape(-1, 3)


def gen_children(octpath, octlevel):
    """Children of each voxel, parent-major, octant 0..7 (octree_utils.py:166-200)."""
    p = np.ravel(octpath).astype(np.int64)
    lv = np.ravel(octlevel).astype(np.int64) + 1
    ch = np.arange(8, dtype=np.int64)
    path = (p[:, None] | (ch[None, :] << (3 * (MAX_NUM_LEVELS - lv))[:, None])).reshape(-1)
    return path, np.repeat(lv, 8)


def corner_coords(octpath, octlevel):
    """[N, 8, 3] level-16 integer coordinates of each voxel's corners (octree_utils.py:129-144),
    corner bits x = 4, y = 2, z = 1."""
    ijk = ijk_array(octpath, octlevel)
    sh = (MAX_NUM_LEVELS - np.ravel(octlevel).astype(np.int64))
    offs = np.array([[(c >> 2) & 1, (c >> 1) & 1, c & 1] for c in range(8)], dtype=np.int64)
    return (ijk[:, None, :] << sh[:, None, None]) + (offs[None] << sh[:, None, None])


def build_grid_pts_link(octpath, octlevel):
    """(grid_pts_key [M,3], vox_key [N,8]) -- np.unique over rows, lexicographic order."""
    rows = corner_coords(octpath, octlevel).reshape(-1, 3)
    key, inv = np.unique(rows, axis=0, return_inverse=True)
    return key, inv.reshape(-1, 8)


def subdivide_by_interp(vox_val):
    """[N, 8, ...] -> [8N, 8, ...]: child c, corner k = parent's trilinear value there
    (adaptive.py:365-411)."""
    v = np.asarray(vox_val, dtype=np.float64)
    out = np.zeros((v.shape[0], 8, 8) + v.shape[2:])
    for c in range(8):
        for k in range(8):
            d = c ^ k                       # axes on which corner k of child c leaves corner c
            bits = [b for b in (1, 2, 4) if d & b]
            subset = [c]
            for b in bits:
                subset = subset + [x ^ b for x in subset]
            out[:, c, k] = np.mean(v[:, subset], axis=1)
    return out.reshape((v.shape[0] * 8, 8) + v.shape[2:])


def agg_mean(num_grid_pts, vox_key, vox_val):
    """Mean over the voxels naming each grid point of their corner values; 0 elsewhere."""
    v = np.asarray(vox_val, dtype=np.float64)
    keys = np.ravel(vox_key)
    flat = v.reshape((keys.size,) + v.shape[2:])
    out = np.zeros((num_grid_pts,) + v.shape[2:])
    cnt = np.zeros(num_grid_pts)
    np.add.at(out, keys, flat)
    np.add.at(cnt, keys, 1.0)
    nz = cnt > 0
    out[nz] = out[nz] / cnt[nz].reshape((-1,) + (1,) * (out.ndim - 1))
    return out


def subdivide(model, mask, outside_level=0):
    """adaptive.py:46-72 on a dict with octpath, octlevel, is_leaf, per-voxel params sh0, shs,
    and grid_pts (with vox_key): returns the new dict (kept voxels first, then the children)."""
    mask = np.asarray(mask, dtype=bool)
    lv = np.ravel(model["octlevel"]).astype(np.int64)
    parent_keep = mask & ((lv - outside_level) >= 9)
    kept = np.nonzero(~mask | parent_keep)[0]
    sub = np.nonzero(mask)[0]
    path, level = gen_children(np.ravel(model["octpath"])[sub], lv[sub])
    leaf = np.array(model["is_leaf"]).copy()
    leaf[parent_keep] = 0
    out = {"octpath": np.concatenate([np.ravel(model["octpath"])[kept], path]),
           "octlevel": np.concatenate([lv[kept], level]),
           "is_leaf": np.concatenate([leaf[kept], np.ones(8 * sub.size, dtype=leaf.dtype)])}
    for name in ("sh0", "shs"):
        a = np.asarray(model[name])
        out[name] = np.concatenate([a[kept], np.repeat(a[sub], 8, axis=0)])
    gp_key, vox_key = build_grid_pts_link(out["octpath"], out["octlevel"])
    vox_val = np.asarray(model["grid_pts"], dtype=np.float64).reshape(-1)[np.asarray(model["vox_key"])]
    vals = np.concatenate([vox_val[kept], subdivide_by_interp(vox_val[sub])])
    out["grid_pts_key"], out["vox_key"] = gp_key, vox_key
    out["grid_pts"] = agg_mean(gp_key.shape[0], vox_key, vals)
    return out


def prune(model, mask):
    """adaptive.py:32-44 on the same dict layout."""
    kept = np.nonzero(~np.asarray(mask, dtype=bool))[0]
    out = {k: np.asarray(model[k])[kept] for k in ("is_leaf", "sh0", "shs")}
    out["octpath"] = np.ravel(model["octpath"])[kept]
    out["octlevel"] = np.ravel(model["octlevel"])[kept]
    gp_key, vox_key = build_grid_pts_link(out["octpath"], out["octlevel"])
    vox_val = np.asarray(model["grid_pts"], dtype=np.float64).reshape(-1)[np.asarray(model["vox_key"])][kept]
    out["grid_pts_key"], out["vox_key"] = gp_key, vox_key
    out["grid_pts"] = agg_mean(gp_key.shape[0], vox_key, vox_val)
    return out
