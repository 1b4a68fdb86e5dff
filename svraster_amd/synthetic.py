"""Seeded synthetic sparse-voxel SDF scenes and pinhole camera rigs.

Scenes follow the reference's octree conventions:
  * octpath: int64, 3 bits per level, left-aligned over MAX_NUM_LEVELS=16 levels,
    x = bit2, y = bit1, z = bit0 (reference src/utils/octree_utils.py:13-22,
    cuda/src/utils.cu:129-158);
  * voxel size at level L = extent * 2^-L, centre = min + (ijk + 0.5) * size
    (octree_utils.py:51-60, 99-126);
  * corner grid points are integer coordinates at the finest level, deduplicated with a
    lexicographic unique so `vox_key[N,8]` indexes unique corners in the same order as
    `build_grid_pts_link` (octree_utils.py:129-144, 203-222);
  * cameras: c2w columns = right, down, lookat, position; w2c = inverse(c2w)
    (src/cameras.py:39-73, 136-162); cx = W/2, cy = H/2, near = 0.02.

The scene is the union of spheres (and optional axis-aligned boxes).  A cell is split while
|SDF(centre)| < cell diagonal, from `min_level` to `max_level`, closest-to-surface first under a
leaf budget (the stated voxel count of each workload).  Corner SDF values are the
analytic SDF plus N(0, 0.01 * size of the finest voxel touching that corner).
"""
from __future__ import annotations

import math
import os
from concurrent.futures import ThreadPoolExecutor
from dataclasses import dataclass, field
from typing import List, Optional

import numpy as np
import torch

MAX_NUM_LEVELS = 16


# ----------------------------------------------------------------------------------------------
# Octree helpers (pure integer arithmetic; identical formulas to the reference's utils.cu)
# ----------------------------------------------------------------------------------------------
def _spread3(v: np.ndarray) -> np.ndarray:
    """Move bit b of a 16-bit value to bit 3*b (the classic Morton "part1by2" masks)."""
    v = v & 0xFFFF
    v = (v | (v << 16)) & 0x0000FF0000FF
    v = (v | (v << 8)) & 0x00F00F00F00F
    v = (v | (v << 4)) & 0x0C30C30C30C3
    v = (v | (v << 2)) & 0x249249249249
    return v


def ijk_to_octpath_np(ijk: np.ndarray, level: np.ndarray) -> np.ndarray:
    """Morton-encode ijk at `level` into a left-aligned 48-bit octpath (utils.cu:129-158):
    bit 3l+2 / 3l+1 / 3l of the path are bit l of i / j / k."""
    ijk = ijk.astype(np.int64)
    level = level.astype(np.int64).reshape(-1)
    path = (_spread3(ijk[:, 0]) << 2) | (_spread3(ijk[:, 1]) << 1) | _spread3(ijk[:, 2])
    return path << (3 * (MAX_NUM_LEVELS - level))


def octpath_to_ijk_np(octpath: np.ndarray, level: np.ndarray) -> np.ndarray:
    """Inverse of `ijk_to_octpath_np` (utils.cu:160-189)."""
    level = level.astype(np.int64).reshape(-1)
    path = octpath.astype(np.int64).reshape(-1) >> (3 * (MAX_NUM_LEVELS - level))
    ijk = np.zeros((path.shape[0], 3), dtype=np.int64)
    for l in range(MAX_NUM_LEVELS):
        bits = path & 7
        ijk[:, 0] |= ((bits >> 2) & 1) << l
        ijk[:, 1] |= ((bits >> 1) & 1) << l
        ijk[:, 2] |= (bits & 1) << l
        path = path >> 3
    return ijk


# ----------------------------------------------------------------------------------------------
# Analytic SDF
# ----------------------------------------------------------------------------------------------
@dataclass
class SDFShapes:
    sphere_c: np.ndarray                      # [S,3] float64
    sphere_r: np.ndarray                      # [S]
    box_c: np.ndarray = field(default_factory=lambda: np.zeros((0, 3)))
    box_h: np.ndarray = field(default_factory=lambda: np.zeros((0, 3)))

    def eval(self, x: np.ndarray, chunk: int = 1 << 15) -> np.ndarray:
        """Union SDF at points x [n,3] (float64).  Chunks run on a thread pool (numpy releases
        the GIL in its loops); every chunk does the same operations, so the result does not
        depend on the thread count."""
        out = np.empty(x.shape[0], dtype=np.float64)

        def work(s):
            xs = x[s:s + chunk]
            d = np.full(xs.shape[0], np.inf)
            for c, r in zip(self.sphere_c, self.sphere_r):
                d = np.minimum(d, np.sqrt(((xs - c) ** 2).sum(-1)) - r)
            for c, h in zip(self.box_c, self.box_h):
                q = np.abs(xs - c) - h
                outside = np.sqrt((np.maximum(q, 0.0) ** 2).sum(-1))
                inside = np.minimum(q.max(-1), 0.0)
                d = np.minimum(d, outside + inside)
            out[s:s + chunk] = d

        starts = range(0, x.shape[0], chunk)
        if x.shape[0] <= 4 * chunk:
            for s in starts:
                work(s)
        else:
            with ThreadPoolExecutor(min(16, os.cpu_count() or 1)) as ex:
                list(ex.map(work, starts))
        return out


def random_shapes(rng: np.random.Generator, n_spheres: int, n_boxes: int = 0) -> SDFShapes:
    c = rng.uniform(-0.5, 0.5, size=(n_spheres, 3))
    r = rng.uniform(0.15, 0.4, size=n_spheres)
    bc = rng.uniform(-0.5, 0.5, size=(n_boxes, 3))
    bh = rng.uniform(0.05, 0.25, size=(n_boxes, 3))
    return SDFShapes(c, r, bc, bh)


# ----------------------------------------------------------------------------------------------
# Scene
# ----------------------------------------------------------------------------------------------
@dataclass
class Scene:
    octpath: torch.Tensor          # [N,1] int64
    octlevel: torch.Tensor         # [N,1] int8
    vox_center: torch.Tensor       # [N,3] f32
    vox_size: torch.Tensor         # [N,1] f32
    vox_key: torch.Tensor          # [N,8] int64
    grid_pts_key: torch.Tensor     # [M,3] int64 (finest-level integer corner coordinates)
    geo_grid_pts: torch.Tensor     # [M,1] f32 corner SDF
    sh0: torch.Tensor              # [N,3] f32
    shs: torch.Tensor              # [N,15,3] f32 (zeros beyond the active degree)
    sh_degree: int
    is_leaf: torch.Tensor          # [N] uint8
    log_s: float                   # s = exp(10 * log_s)
    scene_center: torch.Tensor     # [3]
    scene_extent: float
    shapes: SDFShapes = None

    @property
    def num_voxels(self) -> int:
        return int(self.octpath.shape[0])

    @property
    def s_val(self) -> float:
        return math.exp(10.0 * self.log_s)

    def to(self, device) -> "Scene":
        kw = {}
        for k, v in self.__dict__.items():
            kw[k] = v.to(device) if torch.is_tensor(v) else v
        return Scene(**kw)


def s_init_rule(h_min: float) -> float:
    """log_s init used by the reference trainer: s = ln(99) / (4 h_min)
    (train.py:186-195 with learning_thickness 2)."""
    s = math.log(99.0) / (4.0 * h_min)
    return math.log(s) / 10.0


def build_octree(shapes: SDFShapes, min_level: int, max_level: int,
                 extent: float = 2.0, center=(0.0, 0.0, 0.0),
                 target_voxels: Optional[int] = None,
                 keep_parents_from: Optional[int] = None):
    """Return (ijk [N,3] int64, level [N] int64, is_leaf [N] bool) ordered by (octpath, level).

    A cell at level L < max_level is a split candidate when |SDF(centre)| < its diagonal.
    With `target_voxels` set, candidates are split closest-to-surface first (smallest
    |SDF|/size, ties by position), spending half of the remaining leaf budget at each level
    and all of it at the last split, so every level up to max_level is populated and the
    leaf count lands near the target.

    With `keep_parents_from` set, a split cell at that level or deeper stays in the list as a
    non-leaf voxel, as the reference's subdivision keeps parents at depth >= 9
    (adaptive.py:60-65, 186-188); a parent sorts before its first child, which shares its
    octpath."""
    cmin = np.asarray(center, dtype=np.float64) - 0.5 * extent
    n0 = 1 << min_level
    g = np.arange(n0, dtype=np.int64)
    ijk = np.stack(np.meshgrid(g, g, g, indexing="ij"), -1).reshape(-1, 3)
    sub = np.array([[(c >> 2) & 1, (c >> 1) & 1, c & 1] for c in range(8)], dtype=np.int64)
    leaves_ijk, leaves_lv, leaves_flag = [], [], []
    n_leaves = 0
    for lv in range(min_level, max_level + 1):
        h = extent / (1 << lv)
        if lv == max_level:
            leaves_ijk.append(ijk)
            leaves_lv.append(np.full(ijk.shape[0], lv, dtype=np.int64))
            leaves_flag.append(np.ones(ijk.shape[0], dtype=bool))
            break
        ctr = cmin + (ijk.astype(np.float64) + 0.5) * h
        d = np.abs(shapes.eval(ctr)) / h
        cand = np.nonzero(d < math.sqrt(3.0))[0]
        if target_voxels is not None:
            room = max(target_voxels - (n_leaves + ijk.shape[0]), 0) // 7
            if lv < max_level - 1:
                room //= 2
            if cand.shape[0] > room:
                cand = cand[np.argsort(d[cand], kind="stable")[:room]]
                cand.sort()
        split = np.zeros(ijk.shape[0], dtype=bool)
        split[cand] = True
        keep = ~split if keep_parents_from is None or lv < keep_parents_from else \
            np.ones(ijk.shape[0], dtype=bool)
        leaves_ijk.append(ijk[keep])
        leaves_lv.append(np.full(int(keep.sum()), lv, dtype=np.int64))
        leaves_flag.append(~split[keep])
        n_leaves += int((~split).sum())
        par = ijk[split]
        ijk = ((par[:, None, :] << 1) + sub[None]).reshape(-1, 3)
    ijk = np.concatenate(leaves_ijk, 0)
    lv = np.concatenate(leaves_lv, 0)
    flag = np.concatenate(leaves_flag, 0)
    path = ijk_to_octpath_np(ijk, lv)
    order = np.lexsort((lv, path))
    return ijk[order], lv[order], flag[order]


def build_scene(n_spheres: int, min_level: int, max_level: int, sh_degree: int,
                seed: int = 0, n_boxes: int = 0, sh_init: str = "normal",
                device="cpu", extent: float = 2.0,
                target_voxels: Optional[int] = None,
                keep_parents_from: Optional[int] = None) -> Scene:
    """Build a synthetic scene.  `sh_init` = "uniform" draws sh0 ~ U[0,1] (degree 0 configs),
    "normal" draws all SH coefficients ~ N(0, 0.1)."""
    rng = np.random.default_rng(seed)
    shapes = random_shapes(rng, n_spheres, n_boxes)
    ijk, lv, leaf = build_octree(shapes, min_level, max_level, extent, target_voxels=target_voxels,
                                 keep_parents_from=keep_parents_from)
    N = ijk.shape[0]
    center = np.zeros(3)
    cmin = center - 0.5 * extent
    size = extent / (1 << lv).astype(np.float64)
    vox_center = cmin + (ijk.astype(np.float64) + 0.5) * size[:, None]
    octpath = ijk_to_octpath_np(ijk, lv)

    # Corner integer coordinates at the finest level, deduplicated lexicographically (x, y, z).
    sh = (MAX_NUM_LEVELS - lv)
    base = ijk << sh[:, None]
    offs = np.array([[(c >> 2) & 1, (c >> 1) & 1, c & 1] for c in range(8)], dtype=np.int64)
    corners = base[:, None, :] + (offs[None] << sh[:, None, None])          # [N,8,3]
    code = (corners[..., 0] << 34) | (corners[..., 1] << 17) | corners[..., 2]
    uniq, inv = np.unique(code.reshape(-1), return_inverse=True)
    vox_key = inv.reshape(N, 8).astype(np.int64)
    grid_key = np.stack([(uniq >> 34) & 0x1FFFF, (uniq >> 17) & 0x1FFFF, uniq & 0x1FFFF], -1)
    M = uniq.shape[0]
    finest = extent / (1 << MAX_NUM_LEVELS)
    grid_xyz = cmin + grid_key.astype(np.float64) * finest
    corner_cell = np.full(M, np.inf)
    np.minimum.at(corner_cell, vox_key.reshape(-1), np.repeat(size, 8))
    sdf = shapes.eval(grid_xyz) + rng.normal(0.0, 1.0, size=M) * 0.01 * corner_cell

    if sh_init == "uniform":
        sh0 = rng.uniform(0.0, 1.0, size=(N, 3))
        shs = np.zeros((N, 15, 3))
    else:
        sh0 = rng.normal(0.0, 0.1, size=(N, 3))
        shs = rng.normal(0.0, 0.1, size=(N, 15, 3))
        n_act = (sh_degree + 1) ** 2 - 1
        shs[:, n_act:] = 0.0

    h_min = float(size.min())
    f32 = torch.float32
    return Scene(
        octpath=torch.from_numpy(octpath).view(-1, 1).to(device),
        octlevel=torch.from_numpy(lv.astype(np.int8)).view(-1, 1).to(device),
        vox_center=torch.from_numpy(vox_center).to(f32).to(device),
        vox_size=torch.from_numpy(size).to(f32).view(-1, 1).to(device),
        vox_key=torch.from_numpy(vox_key).to(device),
        grid_pts_key=torch.from_numpy(grid_key).to(device),
        geo_grid_pts=torch.from_numpy(sdf).to(f32).view(-1, 1).to(device),
        sh0=torch.from_numpy(sh0).to(f32).to(device),
        shs=torch.from_numpy(shs).to(f32).to(device),
        sh_degree=sh_degree,
        is_leaf=torch.from_numpy(leaf.astype(np.uint8)).to(device),
        log_s=s_init_rule(h_min),
        scene_center=torch.zeros(3, dtype=f32, device=device),
        scene_extent=extent,
        shapes=shapes,
    )


# ----------------------------------------------------------------------------------------------
# Cameras
# ----------------------------------------------------------------------------------------------
@dataclass
class PinholeCamera:
    image_width: int
    image_height: int
    tanfovx: float
    tanfovy: float
    cx: float
    cy: float
    c2w: torch.Tensor        # [4,4] f32, row-major
    w2c: torch.Tensor        # [4,4] f32
    near: float = 0.02
    cam_mode: str = "persp"

    @property
    def position(self) -> torch.Tensor:
        return self.c2w[:3, 3]

    @property
    def lookat(self) -> torch.Tensor:
        return self.c2w[:3, 2]

    def to(self, device) -> "PinholeCamera":
        return PinholeCamera(self.image_width, self.image_height, self.tanfovx, self.tanfovy,
                             self.cx, self.cy, self.c2w.to(device), self.w2c.to(device),
                             self.near, self.cam_mode)


def look_at_camera(position: np.ndarray, W: int, H: int, fx: float, fy: float,
                   target=(0.0, 0.0, 0.0), device="cpu") -> PinholeCamera:
    pos = np.asarray(position, dtype=np.float64)
    fwd = np.asarray(target, dtype=np.float64) - pos
    fwd /= np.linalg.norm(fwd)
    up = np.array([0.0, 0.0, 1.0])
    if abs(float(fwd @ up)) > 0.99:
        up = np.array([0.0, 1.0, 0.0])
    right = np.cross(fwd, up)
    right /= np.linalg.norm(right)
    down = np.cross(fwd, right)
    c2w = np.eye(4)
    c2w[:3, 0], c2w[:3, 1], c2w[:3, 2], c2w[:3, 3] = right, down, fwd, pos
    w2c = np.linalg.inv(c2w)
    return PinholeCamera(
        image_width=W, image_height=H,
        tanfovx=float(W / (2.0 * fx)), tanfovy=float(H / (2.0 * fy)),
        cx=W * 0.5, cy=H * 0.5,
        c2w=torch.tensor(c2w, dtype=torch.float32, device=device),
        w2c=torch.tensor(w2c, dtype=torch.float32, device=device))


def random_direction(rng: np.random.Generator) -> np.ndarray:
    v = rng.normal(size=3)
    return v / np.linalg.norm(v)


def sphere_camera(seed: int, radius: float, W: int, H: int, fx: float, fy: float,
                  device="cpu") -> PinholeCamera:
    rng = np.random.default_rng(seed)
    return look_at_camera(random_direction(rng) * radius, W, H, fx, fy, device=device)


# ----------------------------------------------------------------------------------------------
# The workload configurations
# ----------------------------------------------------------------------------------------------
@dataclass
class Workload:
    name: str
    scene: Scene
    cameras: List[PinholeCamera]
    s_val: float

    @property
    def rays(self) -> int:
        return sum(c.image_width * c.image_height for c in self.cameras)


def make_workload(name: str, device="cpu", scale: Optional[float] = None) -> Workload:
    """Named workloads.
    small_128    : 3 spheres, levels 3-6 (~8k voxels), SH deg 0 U[0,1], one 128x128 view, fx=102.4
    single_1237  : 16 spheres, levels 6-11, SH deg 3 N(0,0.1), one 1237x822 view, fx=1000
    fd_256       : 4 spheres, levels 4-7, SH deg 1, one 256x256 view, fx=204.8
    multi_32     : 24 spheres, levels 6-12, SH deg 3, 32 views (8 each of 640x480, 1237x822,
                   1600x1064, 1920x1080), fx=0.8*W, radii U[2,3]
    train_16     : 48 spheres + 8 boxes, levels 7-13, SH deg 3, 16 views 1920x1080; split cells
                   at level >= 9 stay as non-leaf parents (the smoothness-loss grid table)
    """
    if name == "small_128":
        sc = build_scene(3, 3, 6, 0, seed=0, sh_init="uniform", device=device,
                         target_voxels=8_000)
        cams = [sphere_camera(0, 2.5, 128, 128, 102.4, 102.4, device=device)]
    elif name == "single_1237":
        sc = build_scene(16, 6, 11, 3, seed=0, device=device, target_voxels=2_000_000)
        cams = [sphere_camera(0, 2.5, 1237, 822, 1000.0, 1000.0, device=device)]
    elif name == "fd_256":
        sc = build_scene(4, 4, 7, 1, seed=0, device=device, target_voxels=20_000)
        cams = [sphere_camera(0, 2.5, 256, 256, 204.8, 204.8, device=device)]
    elif name == "multi_32":
        sc = build_scene(24, 6, 12, 3, seed=0, device=device, target_voxels=4_000_000)
        cams = []
        res = [(640, 480), (1237, 822), (1600, 1064), (1920, 1080)]
        v = 0
        for (W, H) in res:
            for _ in range(8):
                rng = np.random.default_rng(1000 + v)
                r = rng.uniform(2.0, 3.0)
                cams.append(look_at_camera(random_direction(rng) * r, W, H, 0.8 * W, 0.8 * W,
                                           device=device))
                v += 1
    elif name == "train_16":
        sc = build_scene(48, 7, 13, 3, seed=0, n_boxes=8, device=device,
                         target_voxels=16_000_000, keep_parents_from=9)
        cams = [sphere_camera(2000 + v, 2.5, 1920, 1080, 1536.0, 1536.0, device=device)
                for v in range(16)]
    else:
        raise KeyError(name)
    return Workload(name, sc, cams, sc.s_val)
