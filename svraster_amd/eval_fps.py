"""Rendering-speed harness after render.py --eval_fps (render.py:52-141): a warm-up render, then
every view rendered once, forward only (no depth, normal or transmittance outputs), timed
wall-clock around the loop with a device synchronisation; reports views, seconds, peak memory
and frames per second.  Beside the reference's one-view-per-call loop it times the batched path
(all views in one forward: one preprocess, one sort, one persistent compositor launch).

    python -m svraster_amd.eval_fps --checkpoint model.pt [--views 32 --width 1600 --height 1064]
    python -m svraster_amd.eval_fps --workload multi_32 [--save-checkpoint model.pt --quantize]

Views for a checkpoint (there is no dataset here) are seeded look-at cameras on a sphere around
the scene centre, 1.25 scene extents away.
"""
from __future__ import annotations

import argparse
import json
import math
import sys
import time
from typing import List

import numpy as np
import torch


def eval_fps(model, cameras: List, warmup: int = 1) -> dict:
    """One-view-per-call forward rendering of every camera."""
    torch.cuda.empty_cache()
    torch.cuda.reset_peak_memory_stats()
    with torch.no_grad():
        for _ in range(warmup):
            model.render(cameras[0])
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for cam in cameras:
            out = model.render(cam)
        torch.cuda.synchronize()
        dt = time.perf_counter() - t0
    peak = torch.cuda.memory_stats()["allocated_bytes.all.peak"] / 1024 ** 3
    return {"n": len(cameras), "eps": dt, "peak_mem_gb": round(peak, 2), "fps": len(cameras) / dt,
            "resolution": list(out["color"].shape[-2:])}


def eval_fps_batched(model, cameras: List, warmup: int = 1) -> dict:
    """All views in one batched forward (svraster_amd.batched)."""
    from .batched import BatchedRasterizer
    torch.cuda.empty_cache()
    torch.cuda.reset_peak_memory_stats()
    r = BatchedRasterizer(model.octpath, model.octlevel, model.vox_center, model.vox_size, model.is_leaf,
                          model.vox_key, cameras, model.bg_color, model.active_sh_degree)
    s_val = torch.exp(model._log_s.detach().reshape(1) * 10.0)
    g, sh0, shs = model._geo_grid_pts.detach(), model._sh0.detach(), model._shs.detach()
    with torch.no_grad():
        for _ in range(warmup):
            r.forward(g, sh0, shs, s_val)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        r.forward(g, sh0, shs, s_val)
        torch.cuda.synchronize()
        dt = time.perf_counter() - t0
    peak = torch.cuda.memory_stats()["allocated_bytes.all.peak"] / 1024 ** 3
    return {"n": len(cameras), "eps": dt, "peak_mem_gb": round(peak, 2), "fps": len(cameras) / dt}


def sphere_cameras(center, extent: float, n: int, W: int, H: int, seed: int = 0):
    from .synthetic import look_at_camera, random_direction
    c = np.asarray(center, dtype=np.float64).reshape(3)
    cams = []
    for v in range(n):
        rng = np.random.default_rng(seed + v)
        d = random_direction(rng)
        cams.append(look_at_camera(c + d * 1.25 * extent, W, H, 0.8 * W, 0.8 * W, target=c, device="cuda"))
    return cams


def main(argv=None) -> int:
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--checkpoint", default=None)
    ap.add_argument("--workload", default=None, help="synthetic workload instead of a checkpoint")
    ap.add_argument("--views", type=int, default=32)
    ap.add_argument("--width", type=int, default=1600)
    ap.add_argument("--height", type=int, default=1064)
    ap.add_argument("--save-checkpoint", default=None)
    ap.add_argument("--quantize", action="store_true")
    args = ap.parse_args(argv)
    from . import io as svio
    if args.checkpoint:
        model = svio.load(args.checkpoint)
        ext = float(model.scene_extent.reshape(-1)[0])
        cams = sphere_cameras(model.scene_center.cpu().numpy(), ext, args.views, args.width, args.height)
    elif args.workload:
        from .model import VoxelModel
        from .synthetic import make_workload
        w = make_workload(args.workload, device="cuda")
        model = VoxelModel(w.scene)
        cams = w.cameras
    else:
        ap.error("give --checkpoint or --workload")
        return 2
    if args.save_checkpoint:
        svio.save(model, args.save_checkpoint, quantize=args.quantize)
    res = {"voxels": int(model.octpath.shape[0]), "per_view": eval_fps(model, cams),
           "batched": eval_fps_batched(model, cams)}
    print(json.dumps(res), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
