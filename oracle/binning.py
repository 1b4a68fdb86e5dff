"""numpy FP32 restatement of the integer stage (test infrastructure only).

Follows, operation by operation and in the same floating-point order with no fused
multiply-add:
  preprocess         cuda/src/preprocess.cu:60-153 (kernel) and 195-198 (focal length)
  tile rectangle     cuda/src/auxiliary.h:97-111
  key codec          cuda/src/auxiliary.h:40-74
  duplicateWithKeys  cuda/src/forward.cu:749-802
  sort               cuda/src/forward.cu:917-926 (stable LSD radix sort over the key bits; the
                     bit range always covers every tile-id bit, so a stable full-key sort gives
                     the same order -- here numpy's stable argsort)
  identifyTileRanges cuda/src/forward.cu:807-829 and the zero fill at 928

Pinning (tests/test_oracle_binning.py): hand-computed projections of single voxels, the key
codec's closed form, numpy stable argsort (library routine), sum(n_dup) == K, ranges that
partition [0, K), and a brute-force check that every pixel whose ray meets a voxel's box with
a matching octant finds that voxel in its tile list.
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np

BLOCK = 16
MAX_NUM_LEVELS = 16
NUM_BIT_ORDER_RANK = 3 * MAX_NUM_LEVELS
f32 = np.float32


def rep3(q: int) -> int:
    """q repeated in every 3-bit group of the 48 rank bits (auxiliary.h:27-49)."""
    v = 0
    for _ in range(MAX_NUM_LEVELS):
        v = (v << 3) | q
    return v


REP3 = np.array([rep3(q) for q in range(8)], dtype=np.uint64)


@dataclass
class Preprocessed:
    n_dup: np.ndarray        # int32 [P]
    bbox: np.ndarray         # uint32 [P,2]  (xmin<<16|ymin, xmax<<16|ymax)
    quad_bits: np.ndarray    # uint32 [P]
    tiles_x: int
    tiles_y: int
    tbox: np.ndarray = None  # uint32 [P,2]: candidate pixels (see TBOX_MARGIN), packed like bbox


# Candidate-pixel rectangle of the build's compositors (not a reference quantity): when every
# corner is in front of the near plane (depth also within [2^-100, 2^100]), the pixels whose
# centre lies in the projected corners' rectangle widened by TBOX_MARGIN, clipped to bbox
# (empty when x0 > x1 or y0 > y1); otherwise bbox itself.  In pixel-index units the rectangle
# is [bmin, bmax] (the pre-rounding bbox values), so the range is [ceil(bmin - m), floor(bmax + m)].
TBOX_MARGIN = f32(0.03125)


def focal(W: int, tanfov: float) -> np.float32:
    return f32(f32(0.5) * f32(W)) / f32(tanfov)


def preprocess(W, H, tanfovx, tanfovy, cx, cy, w2c, c2w, near,
               centers, lengths, is_leaf) -> Preprocessed:
    centers = np.asarray(centers, dtype=f32).reshape(-1, 3)
    lengths = np.asarray(lengths, dtype=f32).reshape(-1)
    is_leaf = np.asarray(is_leaf).reshape(-1)
    P = centers.shape[0]
    m = np.asarray(w2c, dtype=f32).reshape(-1)[:12]
    cw = np.asarray(c2w, dtype=f32).reshape(-1)
    ro = np.array([cw[3], cw[7], cw[11]], dtype=f32)
    near = f32(near)
    fx, fy = focal(W, tanfovx), focal(H, tanfovy)
    tiles_x, tiles_y = (W + BLOCK - 1) // BLOCK, (H + BLOCK - 1) // BLOCK

    vc = centers
    r = f32(0.5) * lengths
    rel = vc - ro[None]
    d2 = rel[:, 0] * rel[:, 0] + rel[:, 1] * rel[:, 1] + rel[:, 2] * rel[:, 2]
    alive = (is_leaf != 0) & ~(d2 < near * near)

    qbits = np.zeros(P, dtype=np.uint32)
    front = np.ones(P, dtype=bool)
    cmin = np.full((P, 2), f32(1e9), dtype=f32)
    cmax = np.full((P, 2), f32(-1e9), dtype=f32)
    for i in range(8):
        sx = f32(((i & 4) >> 2) * 2 - 1)
        sy = f32(((i & 2) >> 1) * 2 - 1)
        sz = f32((i & 1) * 2 - 1)
        wc = np.stack([vc[:, 0] + r * sx, vc[:, 1] + r * sy, vc[:, 2] + r * sz], -1)
        cam = np.stack([
            m[0] * wc[:, 0] + m[1] * wc[:, 1] + m[2] * wc[:, 2] + m[3],
            m[4] * wc[:, 0] + m[5] * wc[:, 1] + m[6] * wc[:, 2] + m[7],
            m[8] * wc[:, 0] + m[9] * wc[:, 1] + m[10] * wc[:, 2] + m[11]], -1)
        ok = ~(cam[:, 2] < near)
        front &= ok & (cam[:, 2] >= f32(2.0 ** -100)) & (cam[:, 2] <= f32(2.0 ** 100))
        with np.errstate(divide="ignore", invalid="ignore", over="ignore"):
            inv_z = f32(1.0) / cam[:, 2]
            co = np.stack([cam[:, 0] * inv_z, cam[:, 1] * inv_z], -1)
        q = (((wc[:, 0] < ro[0]).astype(np.uint32) << 2)
             | ((wc[:, 1] < ro[1]).astype(np.uint32) << 1)
             | (wc[:, 2] < ro[2]).astype(np.uint32))
        cmin = np.where(ok[:, None], np.minimum(cmin, co), cmin)
        cmax = np.where(ok[:, None], np.maximum(cmax, co), cmax)
        qbits = np.where(ok, qbits | (np.uint32(1) << q), qbits)

    cxh, cyh = f32(cx) - f32(0.5), f32(cy) - f32(0.5)
    with np.errstate(over="ignore", invalid="ignore"):
        bmin_x = np.maximum(fx * cmin[:, 0] + cxh, f32(0.0))
        bmin_y = np.maximum(fy * cmin[:, 1] + cyh, f32(0.0))
        bmax_x = np.minimum(fx * cmax[:, 0] + cxh, f32(W))
        bmax_y = np.minimum(fy * cmax[:, 1] + cyh, f32(H))
    alive &= ~((bmin_x > bmax_x) | (bmin_y > bmax_y))

    def rint_u(v):
        v = np.where(alive, v, f32(0.0))
        return np.rint(v).astype(np.int64).astype(np.uint32)

    bbox = np.zeros((P, 2), dtype=np.uint32)
    bbox[:, 0] = (rint_u(bmin_x) << np.uint32(16)) | rint_u(bmin_y)
    bbox[:, 1] = (rint_u(bmax_x) << np.uint32(16)) | rint_u(bmax_y)
    bbox[~alive] = 0
    qbits[~alive] = 0
    tmin, tmax = tile_rect(bbox, tiles_x, tiles_y)
    tiles = (1 + tmax[:, 1] - tmin[:, 1]) * (1 + tmax[:, 0] - tmin[:, 0])
    popc = np.array([bin(b).count("1") for b in range(256)], dtype=np.int64)[qbits & 0xFF]
    n_dup = np.where(alive, tiles * popc, 0).astype(np.int32)

    def tight(v, up):
        with np.errstate(over="ignore", invalid="ignore"):
            c = np.ceil(v - TBOX_MARGIN) if up else np.floor(v + TBOX_MARGIN)
        return np.where(alive, c, f32(0.0)).astype(np.int64)

    bx = [rint_u(v).astype(np.int64) for v in (bmin_x, bmin_y, bmax_x, bmax_y)]
    t0x = np.maximum(bx[0], tight(bmin_x, True))
    t0y = np.maximum(bx[1], tight(bmin_y, True))
    t1x = np.minimum(bx[2], tight(bmax_x, False))
    t1y = np.minimum(bx[3], tight(bmax_y, False))
    tbox = np.zeros((P, 2), dtype=np.uint32)
    tbox[:, 0] = ((t0x << 16) | t0y).astype(np.uint32)
    tbox[:, 1] = ((t1x << 16) | t1y).astype(np.uint32)
    tbox = np.where((alive & front)[:, None], tbox, bbox)
    return Preprocessed(n_dup, bbox, qbits, tiles_x, tiles_y, tbox)


def tbox_empty(tbox: np.ndarray) -> np.ndarray:
    return ((tbox[:, 0] >> 16) > (tbox[:, 1] >> 16)) | ((tbox[:, 0] & 0xFFFF) > (tbox[:, 1] & 0xFFFF))


def tight_binning_input(pre: Preprocessed) -> Preprocessed:
    """The preprocess output the build's multi-view binning uses: each (view, voxel) binned over
    the tiles of its candidate rectangle instead of its bbox, and left out when that rectangle
    holds no pixel (its keys could only land on pixels whose ray misses the voxel)."""
    empty = tbox_empty(pre.tbox)
    alive = (pre.n_dup > 0) & ~empty
    tmin, tmax = tile_rect(pre.tbox, pre.tiles_x, pre.tiles_y)
    tiles = (1 + tmax[:, 1] - tmin[:, 1]) * (1 + tmax[:, 0] - tmin[:, 0])
    popc = np.array([bin(b).count("1") for b in range(256)], dtype=np.int64)[pre.quad_bits & 0xFF]
    n_dup = np.where(alive, tiles * popc, 0).astype(np.int32)
    box = np.where(alive[:, None], pre.tbox, 0).astype(np.uint32)
    q = np.where(alive, pre.quad_bits, 0).astype(np.uint32)
    return Preprocessed(n_dup, box, q, pre.tiles_x, pre.tiles_y, box)


def tile_rect(bbox: np.ndarray, tiles_x: int, tiles_y: int):
    xmin = (bbox[:, 0] >> 16).astype(np.int64)
    ymin = (bbox[:, 0] & 0xFFFF).astype(np.int64)
    xmax = (bbox[:, 1] >> 16).astype(np.int64)
    ymax = (bbox[:, 1] & 0xFFFF).astype(np.int64)
    tmin = np.stack([np.clip(xmin // BLOCK, 0, tiles_x - 1), np.clip(ymin // BLOCK, 0, tiles_y - 1)], -1)
    tmax = np.stack([np.clip(xmax // BLOCK, 0, tiles_x - 1), np.clip(ymax // BLOCK, 0, tiles_y - 1)], -1)
    return tmin, tmax


def inclusive_scan(n_dup: np.ndarray) -> np.ndarray:
    return np.cumsum(n_dup.astype(np.uint64)).astype(np.uint32)


def duplicate_with_keys(pre: Preprocessed, octpaths: np.ndarray):
    """Emit (key, val) pairs in voxel order, octants ascending, tile rows then columns."""
    octpaths = np.asarray(octpaths, dtype=np.int64).reshape(-1).astype(np.uint64)
    P = octpaths.shape[0]
    bits = ((pre.quad_bits[:, None] >> np.arange(8, dtype=np.uint32)[None]) & 1).astype(bool)
    bits &= (pre.n_dup > 0)[:, None]
    v_idx, q_idx = np.nonzero(bits)                     # voxel-major, octant-minor
    tmin, tmax = tile_rect(pre.bbox, pre.tiles_x, pre.tiles_y)
    tw = (tmax[:, 0] - tmin[:, 0] + 1)[v_idx]
    th = (tmax[:, 1] - tmin[:, 1] + 1)[v_idx]
    cnt = tw * th
    K = int(cnt.sum())
    pair = np.repeat(np.arange(v_idx.shape[0]), cnt)
    start = np.concatenate([[0], np.cumsum(cnt)[:-1]]) if cnt.shape[0] else np.zeros(0, np.int64)
    t = np.arange(K, dtype=np.int64) - np.repeat(start, cnt)
    v = v_idx[pair]
    q = q_idx[pair]
    y = tmin[v, 1] + t // tw[pair]
    x = tmin[v, 0] + t % tw[pair]
    tile = (y * pre.tiles_x + x).astype(np.uint64)
    rank = octpaths[v] ^ REP3[q]
    keys = (tile << np.uint64(NUM_BIT_ORDER_RANK)) | rank
    vals = ((q.astype(np.uint32) << np.uint32(29)) | v.astype(np.uint32)).astype(np.uint32)
    assert K == int(pre.n_dup.astype(np.int64).sum())
    return keys, vals


def sort_pairs(keys: np.ndarray, vals: np.ndarray):
    order = np.argsort(keys, kind="stable")
    return keys[order], vals[order]


def tile_ranges(sorted_keys: np.ndarray, n_tiles: int) -> np.ndarray:
    ranges = np.zeros((n_tiles, 2), dtype=np.uint32)
    if sorted_keys.shape[0] == 0:
        return ranges
    tile = (sorted_keys >> np.uint64(NUM_BIT_ORDER_RANK)).astype(np.int64)
    ids = np.arange(n_tiles)
    lo = np.searchsorted(tile, ids, side="left")
    hi = np.searchsorted(tile, ids, side="right")
    nz = hi > lo
    ranges[nz, 0] = lo[nz]
    ranges[nz, 1] = hi[nz]
    return ranges


@dataclass
class Binning:
    pre: Preprocessed
    scan: np.ndarray
    keys: np.ndarray          # sorted
    vals: np.ndarray          # sorted
    ranges: np.ndarray        # [n_tiles, 2]

    @property
    def K(self) -> int:
        return int(self.keys.shape[0])


def binning(W, H, tanfovx, tanfovy, cx, cy, w2c, c2w, near, octpaths, centers, lengths,
            is_leaf) -> Binning:
    pre = preprocess(W, H, tanfovx, tanfovy, cx, cy, w2c, c2w, near, centers, lengths, is_leaf)
    scan = inclusive_scan(pre.n_dup)
    keys, vals = duplicate_with_keys(pre, octpaths)
    skeys, svals = sort_pairs(keys, vals)
    ranges = tile_ranges(skeys, pre.tiles_x * pre.tiles_y)
    return Binning(pre, scan, skeys, svals, ranges)


def binning_for_camera(cam, scene) -> Binning:
    """Convenience wrapper taking svraster_amd.synthetic camera/scene objects."""
    return binning(cam.image_width, cam.image_height, cam.tanfovx, cam.tanfovy, cam.cx, cam.cy,
                   cam.w2c.cpu().numpy(), cam.c2w.cpu().numpy(), cam.near,
                   scene.octpath.cpu().numpy(), scene.vox_center.cpu().numpy(),
                   scene.vox_size.cpu().numpy(), scene.is_leaf.cpu().numpy())
