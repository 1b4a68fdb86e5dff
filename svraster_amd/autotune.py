"""Autotuner of the batched multi-view rasterizer.

Picks, for one workload (voxel set + camera batch), the launch configuration of the batched
path by timing its own fwd+bwd step on the GPU:

* tile shape: 16x16 (the reference's), 32x8 or 8x32 -- 256 pixels, four 8x8 wavefront
  quadrants per tile either way; the shape changes how many tiles a voxel's footprint
  overlaps (the key count) and how well a tile's entries are shared by its four wavefronts;
* LDS staging depth of the pair-compacted compositors: 64 or 128 entries per round;
* persistent grid size of the work-queue launches: workgroups per CU;
* compositor variant: pair-compacted (entry-major) or one entry at a time;
* corner-SDF gradients: flushed by the backward compositor straight into the grid points, or
  into a per-voxel buffer scattered by the gather transpose afterwards.

None of these changes an image or a gradient (tests/test_gpu_variants.py and
tests/test_gpu_autotune.py check that), so the tuner is free to pick the fastest.  The search
is coordinate descent from the defaults (tile shape, then depth, then grid size, then variant,
then the gradient route), 8 timed configurations instead of the 72 of the full product; each is
timed as the median of a few eager steps after a warm-up step.  Results are cached per workload
signature in a JSON file (SVRASTER_TUNE_CACHE, default ~/.cache/svraster_amd/tune.json).

There is no counterpart in the reference, whose kernels have fixed 16x16 tiles and 256-entry
batches (cuda/src/config.h:14-39, forward.cu:42-564).
"""
from __future__ import annotations

import json
import os
import statistics
from typing import Callable, Dict, List, Optional, Tuple

import torch

from . import _lib

TILE_SHAPES: Tuple[Tuple[int, int], ...] = ((16, 16), (32, 8), (8, 32))
BATCH_DEPTHS: Tuple[int, ...] = (128, 64)
BWD_BATCH_DEPTHS: Tuple[int, ...] = (64, 128)
BLOCKS_PER_CU: Tuple[int, ...] = (8, 4, 16)
VARIANTS: Tuple[int, ...] = (0, 1)
FUSE_GATHER_BW: Tuple[bool, ...] = (False, True)

DEFAULT = {"tile_shape": (16, 16), "batch_depth": 128, "bwd_batch_depth": 64, "persistent_blocks_per_cu": 8,
           "compositor_variant": 0, "fuse_gather_bw": False}


def apply_options(cfg: Dict, step=None) -> None:
    """Set the library options of a configuration (the tile shape belongs to the rasterizer,
    the gradient route to the step's rasterizer when one is given)."""
    _lib.set_option("batch_depth", int(cfg["batch_depth"]))
    _lib.set_option("bwd_batch_depth", int(cfg.get("bwd_batch_depth", DEFAULT["bwd_batch_depth"])))
    _lib.set_option("persistent_blocks_per_cu", int(cfg["persistent_blocks_per_cu"]))
    _lib.set_option("compositor_variant", int(cfg["compositor_variant"]))
    rast = getattr(step, "r", None)
    if rast is not None and hasattr(rast, "fuse_gather_bw"):
        rast.fuse_gather_bw = bool(cfg.get("fuse_gather_bw", False))


def workload_signature(n_voxels: int, cameras: List, sh_degree: int, need_depth: bool = False,
                       need_normal: bool = False) -> str:
    res = sorted((c.image_width, c.image_height) for c in cameras)
    counts: Dict[str, int] = {}
    for w, h in res:
        counts[f"{w}x{h}"] = counts.get(f"{w}x{h}", 0) + 1
    views = ",".join(f"{k}*{v}" for k, v in sorted(counts.items()))
    dev = torch.cuda.get_device_name(0) if torch.cuda.is_available() else "cpu"
    return f"{dev}|P={n_voxels}|views={views}|sh={sh_degree}|d={int(need_depth)}|n={int(need_normal)}"


def _cache_path() -> str:
    return os.environ.get("SVRASTER_TUNE_CACHE",
                          os.path.join(os.path.expanduser("~"), ".cache", "svraster_amd", "tune.json"))


def load_cached(signature: str) -> Optional[Dict]:
    try:
        with open(_cache_path()) as f:
            d = json.load(f)
        cfg = d.get(signature)
        if cfg is not None:
            cfg = dict(cfg)
            cfg["tile_shape"] = tuple(cfg["tile_shape"])
        return cfg
    except (OSError, ValueError):
        return None


def store_cached(signature: str, cfg: Dict) -> None:
    path = _cache_path()
    try:
        os.makedirs(os.path.dirname(path), exist_ok=True)
        d = {}
        if os.path.exists(path):
            with open(path) as f:
                d = json.load(f)
        d[signature] = {k: (list(v) if isinstance(v, tuple) else v) for k, v in cfg.items()}
        with open(path, "w") as f:
            json.dump(d, f, indent=1)
    except (OSError, ValueError):
        pass       # an unwritable cache only costs a re-tune


def time_step(step: Callable[[], None], steps: int = 3, warmup: int = 1) -> float:
    """Median milliseconds of `steps` eager calls after `warmup` calls (HIP events)."""
    for _ in range(warmup):
        step()
    times = []
    for _ in range(steps):
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        step()
        e1.record()
        torch.cuda.synchronize()
        times.append(e0.elapsed_time(e1))
    return statistics.median(times)


def autotune(make_step: Callable[[Tuple[int, int]], Callable[[], None]], signature: Optional[str] = None,
             steps: int = 3, warmup: int = 1, use_cache: bool = True, log: Optional[Callable] = None):
    """Coordinate descent over the launch configuration.

    make_step(tile_shape) builds the workload's batched step for a tile shape and returns a
    zero-argument callable running one fwd+bwd step (the previous one is dropped first, so only
    one rasterizer's buffers are alive at a time).  Returns (best config, table of
    (config, ms) rows); the best config's options are left applied."""
    if use_cache and signature is not None:
        cfg = load_cached(signature)
        if cfg is not None:
            apply_options(cfg)
            return cfg, [("cached", cfg, None)]
    table = []
    best = dict(DEFAULT)

    def run(cfg, step):
        apply_options(cfg, step)
        ms = time_step(step, steps, warmup)
        table.append((dict(cfg), ms))
        if log:
            log(f"[autotune] {cfg} -> {ms:.2f} ms")
        return ms

    # 1. tile shape (each needs its own rasterizer)
    best_ms = float("inf")
    for shape in TILE_SHAPES:
        cfg = dict(best, tile_shape=shape)
        step = make_step(shape)
        ms = run(cfg, step)
        del step
        torch.cuda.empty_cache()
        if ms < best_ms:
            best_ms, best = ms, cfg
    step = make_step(best["tile_shape"])
    # 2.-6. the launch options and the gradient route, one at a time
    for key, values in (("batch_depth", BATCH_DEPTHS), ("bwd_batch_depth", BWD_BATCH_DEPTHS),
                        ("persistent_blocks_per_cu", BLOCKS_PER_CU), ("compositor_variant", VARIANTS),
                        ("fuse_gather_bw", FUSE_GATHER_BW)):
        for val in values:
            if val == best[key]:
                continue
            cfg = dict(best, **{key: val})
            ms = run(cfg, step)
            if ms < best_ms:
                best_ms, best = ms, cfg
    apply_options(best, step)
    if signature is not None:
        store_cached(signature, best)
    return best, table
