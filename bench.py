#!/usr/bin/env python
"""Benchmark: rasterizer fwd+bwd throughput (Mrays/s) on one MI355X.

Workload (default `multi_32`): 4M-voxel sphere-union SDF scene (24 spheres, octree levels
6-12, SH degree 3, FP32), 32 pinhole views of mixed resolution (8 each of 640x480, 1237x822,
1600x1064, 1920x1080; fx = 0.8 W), white background, MSE loss to a U[0,1] target per view.
One step = forward + backward of all 32 views (corner gather, SH, compositor, loss, reverse
compositor, SH and gather transposes), inputs resident in HBM.

Prints one JSON line.  Extra fields: `paths` (per-view vs batched numbers), `roofline` of the
dominant kernel (the slower of the forward and backward compositor launches, HIP events on the
stream they run on; both are listed under `roofline.kernels`, with the VALU instruction rate where
the committed rocprofv3 summary has it), `cpu_baseline` (the CPU oracle on this host: FP64 and
FP32 on a bounded crop of view 0, and whole frames of the CPU-sized scenes).
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

METRIC = "fwd+bwd Mrays/s on 1 MI355X, batched vs per-view, at ≤1e-4 img err; HBM GB/s"
HBM_PEAK_GBS = 8000.0


def bwd_kernel_name():
    """Name of the multi-view backward compositor instantiation the current options launch."""
    from svraster_amd import _lib
    halves = _lib.get_option("bwd_batch_depth") // 64
    return f"render_bwd_pairs_kernel<false, false, false, true, {halves}>"


def tune(make_step, sig, enabled):
    """Autotuned launch configuration of the batched path (svraster_amd/autotune.py)."""
    from svraster_amd import autotune
    if not enabled:
        autotune.apply_options(autotune.DEFAULT)
        return dict(autotune.DEFAULT), {"enabled": False}
    best, table = autotune.autotune(make_step, sig, steps=3, warmup=1, use_cache=False, log=log)
    rows = [{"config": {k: (list(v) if isinstance(v, tuple) else v) for k, v in c.items()}, "ms": round(ms, 3)}
            for c, ms in table]
    chosen = {k: (list(v) if isinstance(v, tuple) else v) for k, v in best.items()}
    return best, {"enabled": True, "chosen": chosen, "search": "coordinate descent, eager step median of 3",
                  "table": rows}


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def build_targets(cams, device):
    import torch
    g = torch.Generator(device="cpu").manual_seed(1)
    return [torch.rand(3, c.image_height, c.image_width, generator=g).to(device) for c in cams]


def per_view_step(model, cams, targets):
    for cam, tgt in zip(cams, targets):
        out = model.render(cam)
        loss = ((out["color"] - tgt) ** 2).mean()
        loss.backward()


def zero_grads(model):
    for p in (model._geo_grid_pts, model._sh0, model._shs, model._subdiv_p, model._log_s):
        p.grad = None


def time_steps(fn, steps, warmup, sync):
    for _ in range(warmup):
        fn()
    sync()
    t0 = time.perf_counter()
    for _ in range(steps):
        fn()
    sync()
    return (time.perf_counter() - t0) * 1000.0 / steps


def compositor_bytes(rast, dvox):
    """Algorithmic HBM bytes of the two multi-view compositor launches (one launch each, all views),
    per the byte model of the path's measurement plan (72 B per streamed (tile, entry) pair: value 4,
    bbox 8, centre 12, size 4, 8 corner SDFs 32, rgb 12):

    forward  72 E_f + 8 Nt + 20 R    (tile range 8; per ray colour 12, T 4, contributor count 4
                                       written; depth and normal are not requested by this step)
    backward 72 E_b + 16 Nt + 20 R + 52 Pv
                                      (per ray dL/dcolour 12, dL/dT 4, T 4 read; 13 floats of
                                       gradient per touched voxel written once)

    E_f: entries the forward streamed (up to each tile's last-contributor bound); E_b: entries the
    backward streams (those the forward handed over as contributing).  The hand-over itself --
    per contributing entry its four 8x8-block pixel masks, 32 B written by the forward and read by
    the backward -- is this design's own traffic, not algorithmic bytes; it is reported beside."""
    ranges, _, _, tile_last = rast.state_views()
    E_f = int((tile_last.long() - ranges[:, 0].long()).clamp_min(0).sum().item())
    cr = rast.contrib_ranges()
    E_b = int((cr[:, 1].long() - cr[:, 0].long()).clamp_min(0).sum().item())
    touched = int((dvox != 0).any(1).sum().item())
    Nt, R = rast.n_tiles, rast.n_pix
    fwd = 72.0 * E_f + 8.0 * Nt + 20.0 * R
    bwd = 72.0 * E_b + 16.0 * Nt + 20.0 * R + 52.0 * touched
    return fwd, bwd, {"forward_streamed_entries": E_f, "backward_streamed_entries": E_b, "tiles": Nt,
                      "rays": R, "touched_voxels": touched, "handover_bytes": int(36 * E_b)}


def profile_counters(kernel_prefix, workload):
    """Per-launch figures of the given kernel from the committed rocprofv3 --pmc summary of the
    workload's step (profiles/pmc_compositors_<workload>.json): HBM bytes (FETCH_SIZE + WRITE_SIZE
    scaled by the copy-kernel calibration) and VALU instructions.  None when absent."""
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "profiles",
                        f"pmc_compositors_{workload.replace('_', '')}.json")
    try:
        with open(path) as f:
            return json.load(f)["kernels"][kernel_prefix]
    except (OSError, KeyError, ValueError):
        return None


def cpu_threads():
    """Host threads the CPU leg uses: all cores, bounded by the job's thread allowance when one is set."""
    n = os.cpu_count() or 1
    env = os.environ.get("OMP_NUM_THREADS")
    if env and env.isdigit() and int(env) > 0:
        n = min(n, int(env))
    return n


def cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def _oracle_fwd_bwd(sc, cam, bn, s_val, dtype, window=None):
    """Oracle forward + backward (colour cotangent) of one view; returns (seconds, rays, output)."""
    import torch
    from oracle import raster as orc
    from oracle import sh as osh
    t0 = time.perf_counter()
    rgbs = osh.sh_rgb(sc.sh_degree, None, sc.vox_center, cam.c2w[:3, 3], sc.sh0.double(),
                      sc.shs.double()).detach().to(dtype).requires_grad_(True)
    geos = sc.geo_grid_pts.reshape(-1)[sc.vox_key].to(dtype).requires_grad_(True)
    s = torch.tensor(s_val, dtype=dtype, requires_grad=True)
    out = orc.render(orc.camera_from(cam), [1.0, 1.0, 1.0], bn.vals, bn.ranges, bn.pre.bbox,
                     sc.vox_center, sc.vox_size, geos, rgbs, s, need_depth=False, need_normal=False,
                     window=window, dtype=dtype)
    g = torch.rand(out.color.shape, generator=torch.Generator().manual_seed(3), dtype=torch.float64)
    orc.backward(out, geos, rgbs, s, g)
    x0, y0, x1, y1 = out.window
    return time.perf_counter() - t0, (x1 - x0) * (y1 - y0), out


def cpu_baseline(workload_name, gpu_color_view0=None, full_frames=("small_128",)):
    """The CPU oracle timed on this host: fwd+bwd (colour loss) of a 64x64 crop at the centre of
    view 0 of the workload in FP64 (the headline CPU figure) and FP32, plus whole frames of the
    small CPU-sized workloads.  Also returns the max-abs difference of the GPU colour on the crop
    vs the FP64 oracle."""
    import torch
    sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))
    from oracle import binning as ob
    from svraster_amd.synthetic import make_workload
    threads = cpu_threads()
    torch.set_num_threads(threads)
    w = make_workload(workload_name, device="cpu")
    sc, cam = w.scene, w.cameras[0]
    bn = ob.binning_for_camera(cam, sc)
    cx, cy = cam.image_width // 2, cam.image_height // 2
    x0, y0 = cx - 32 - cx % 16, cy - 32 - cy % 16
    win = (x0, y0, x0 + 64, y0 + 64)
    dt64, rays, out = _oracle_fwd_bwd(sc, cam, bn, w.s_val, torch.float64, win)
    dt32, _, _ = _oracle_fwd_bwd(sc, cam, bn, w.s_val, torch.float32, win)
    err = None
    if gpu_color_view0 is not None:
        got = gpu_color_view0[:, win[1]:win[3], win[0]:win[2]].detach().cpu().double()
        err = float((got - out.color.detach()).abs().max())
    frames = {}
    for name in full_frames:
        wf = make_workload(name, device="cpu")
        bnf = ob.binning_for_camera(wf.cameras[0], wf.scene)
        for dt, tag in ((torch.float64, "fp64"), (torch.float32, "fp32")):
            sec, r, _ = _oracle_fwd_bwd(wf.scene, wf.cameras[0], bnf, wf.s_val, dt)
            frames[f"{name} full frame {tag}"] = {"mrays_s": round(r / sec / 1e6, 6), "seconds": round(sec, 2),
                                                   "rays": r}
    return ({"value": round(rays / dt64 / 1e6, 6), "unit": "Mrays/s", "cores": threads, "kind": "port",
             "sample": f"FP64 torch oracle fwd+bwd (colour loss), 64x64 crop at the centre of view 0 of "
                       f"{workload_name}; {dt64:.1f} s",
             "fp32_value": round(rays / dt32 / 1e6, 6), "cpu_model": cpu_model(), "full_frames": frames}, err)


VALU_PEAK_PER_S = 256 * 2.4e9     # wave64 FP32 VALU instructions: 1 per cycle per CU (4 SIMD16), 2.4 GHz


def fwd_kernel_name():
    from svraster_amd import _lib
    halves = _lib.get_option("batch_depth") // 64
    return f"render_fwd_pairs_kernel<false, false, false, true, {halves}>"


def roofline_block(rast, kt, workload):
    """HBM roofline of both multi-view compositor launches (algorithmic bytes / the launch's
    HIP-event time on its own stream); the dominant (slower) one fills the top-level fields.
    Measured HBM traffic and VALU instruction counts per launch come from the committed rocprofv3
    summary when it describes the same kernel."""
    fb, bb, counts = compositor_bytes(rast, rast.dvox)
    entries = []
    for key, name, nbytes in (("render_fwd", fwd_kernel_name(), fb), ("render_bwd", bwd_kernel_name(), bb)):
        ms, n = kt[key]
        avg_s = ms / max(n, 1) / 1000.0
        achieved = nbytes / avg_s / 1e9
        prof = profile_counters(name, workload) or {}
        e = {"kernel": name + " (multi-view work queue: all views in one launch)", "bound": "hbm",
             "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
             "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": prof.get("hbm_bytes_per_launch"),
             "bytes_per_launch": int(nbytes), "avg_launch_ms": round(avg_s * 1000.0, 4)}
        if prof.get("valu_insts_per_launch"):
            rate = prof["valu_insts_per_launch"] / avg_s
            e["valu"] = {"insts_per_launch": prof["valu_insts_per_launch"], "achieved_per_s": rate,
                         "peak_per_s": VALU_PEAK_PER_S, "frac": round(rate / VALU_PEAK_PER_S, 4)}
        entries.append(e)
    dom = max(entries, key=lambda e: e["avg_launch_ms"])
    out = {k: dom[k] for k in ("bound", "achieved", "peak", "unit", "frac", "traffic", "kernel",
                               "bytes_per_launch", "avg_launch_ms")}
    if "valu" in dom:
        out["valu"] = dom["valu"]
    out["counts"] = counts
    out["kernels"] = entries
    return out


def kernel_times(lib, fn):
    """Run fn once with the library's HIP-event timing on; returns {name: (ms, launches)}."""
    import ctypes
    import torch
    from svraster_amd import _lib
    lib.svr_timing_enable(1)
    for k in (0, 1):
        lib.svr_timing_read(k, 1, ctypes.byref(ctypes.c_double()), ctypes.byref(ctypes.c_longlong()))
    fn()
    torch.cuda.synchronize()
    out = {}
    for k, name in ((0, "render_fwd"), (1, "render_bwd")):
        ms, n = ctypes.c_double(), ctypes.c_longlong()
        _lib.check(lib.svr_timing_read(k, 1, ctypes.byref(ms), ctypes.byref(n)))
        out[name] = (ms.value, n.value)
    lib.svr_timing_enable(0)
    return out


def train_iteration_bytes(it, smooth):
    """Algorithmic HBM bytes of the non-render stages of one training iteration.
    loss: colour and the scaled target read, dL/dcolour written (3 x 12 B per ray; with views of
    different sizes a per-element scale is read too, 4 x 12 B); 0 when the compositor writes
    dL/dcolour itself (fused loss, the default), whose 24 B per ray (target read, gradient
    written) are then part of the render stage.
    smooth: every grid key once (key 4, active index 4, voxel id 4), the mask and dense index of
    the res^3 grid once (1 + 4 B per cell), the voxel data of the table voxels once (corner ids
    64, grid coordinate 12, size 4), corner SDFs read once and the gradient read + written
    (12 B per corner).
    adam: corner SDFs, their moments and gradient (p, m, v read + written, g read: 28 B each)."""
    R = it.r.n_pix
    M = it.grid_pts.numel()
    out = {"loss": 0.0 if it.fuse_loss else (36.0 if it.loss_grad.scale is None else 48.0) * R, "adam": 28.0 * M}
    if smooth is not None:
        n_tab_vox = int(smooth.g2v.unique().numel())
        out["smooth"] = 12.0 * smooth.G + 5.0 * smooth.res ** 3 + 80.0 * n_tab_vox + 12.0 * M
    return out


def bench_train(args):
    """Full training iteration at scale (train_16): 16 views at 1920x1080 over the 16M-leaf scene;
    render + MSE + backward + grid eikonal (the trainer's weight at iteration 0, 1e-8) + Laplacian
    smoothness (weight 1e-3) + Adam (lr 1e-2, betas 0.9/0.99)
    in the reference trainer's mode: SparseAdam with its defaults (sparse=False, biased=False),
    i.e. every element of every group steps every iteration.  The plain per-view iteration runs
    the drop-in API the way the reference trainer calls it (one view per call, randperm active
    list, SparseAdam over the three groups); the batched iteration is one captured graph."""
    import torch
    import svraster_amd as svraster_cuda
    from svraster_amd import _lib
    from svraster_amd.batched import (BatchedRasterizer, BatchedTrainIteration, GridSmoothness,
                                      trainer_smoothness_weights)
    from svraster_amd.model import VoxelModel
    from svraster_amd.sparse_adam import SparseAdam
    from svraster_amd.synthetic import make_workload

    dev = torch.device("cuda:0")
    t0 = time.perf_counter()
    log(f"[bench] building scene {args.workload} ...")
    w = make_workload(args.workload, device=dev)
    sc, cams = w.scene, w.cameras
    n_leaf = int(sc.is_leaf.sum().item())
    log(f"[bench] scene {args.workload}: {sc.num_voxels} voxels ({n_leaf} leaves), "
        f"{sc.geo_grid_pts.shape[0]} corners, {len(cams)} views, {w.rays} rays/step "
        f"(built in {time.perf_counter() - t0:.1f} s)")
    model = VoxelModel(sc)
    targets = build_targets(cams, dev)
    sync = torch.cuda.synchronize
    lib = _lib.load()
    rays = w.rays
    weight, lr, betas, eps = 1e-3, 1e-2, (0.9, 0.99), 1e-15

    t0 = time.perf_counter()
    w_ge = trainer_smoothness_weights(0)[0]          # the reference trainer's eikonal weight at iteration 0
    smooth = GridSmoothness(sc.vox_center, sc.vox_size, sc.octlevel, sc.is_leaf, sc.vox_key, sc.scene_center,
                            sc.scene_extent, weight=weight, seed=0, eikonal_weight=w_ge)
    sync()
    log(f"[bench] grid table: res {smooth.res}, {smooth.G} keys ({time.perf_counter() - t0:.1f} s)")

    # -- plain per-view iteration (drop-in API) ---------------------------------------------------
    opt = SparseAdam([{"params": [model._geo_grid_pts]}, {"params": [model._sh0]}, {"params": [model._shs]}],
                     lr=lr, betas=betas, eps=eps)

    def pv_render():
        zero_grads(model)
        per_view_step(model, cams, targets)

    def pv_smooth():
        kw = dict(grid_pts=model._geo_grid_pts.detach(), vox_key=sc.vox_key, grid_voxel_coord=smooth.coord,
                  grid_voxel_size=smooth.size, grid_res=smooth.res, grid_mask=smooth.mask, grid_keys=smooth.keys,
                  grid2voxel=smooth.g2v, vox_size_inv=smooth.vox_size_inv, no_tv_s=True, tv_sparse=False,
                  grid_pts_grad=model._geo_grid_pts.grad)
        svraster_cuda.grid_loss_bw.grid_eikonal(
            active_list=torch.randperm(smooth.G, device=dev).to(torch.int32), weight=w_ge, **kw)
        svraster_cuda.grid_loss_bw.laplacian_smoothness(
            active_list=torch.randperm(smooth.G, device=dev).to(torch.int32), weight=weight, **kw)

    def pv_adam():
        opt.step()

    pv_stages = (("render", pv_render), ("smooth", pv_smooth), ("adam", pv_adam))

    def step_pv():
        for _, fn in pv_stages:
            fn()

    def stage_times(stages):
        ev = [torch.cuda.Event(enable_timing=True) for _ in range(len(stages) + 1)]
        ev[0].record()
        for i, (_, fn) in enumerate(stages):
            fn()
            ev[i + 1].record()
        sync()
        return {name: round(ev[i].elapsed_time(ev[i + 1]), 3) for i, (name, _) in enumerate(stages)}

    if args.profile:              # profiler runs time the batched iteration only, eagerly
        ms_pv, st_pv = float("nan"), {}
    else:
        ms_pv = time_steps(step_pv, args.steps, args.warmup, sync)
        st_pv = stage_times(pv_stages)
        log(f"[bench] per-view iteration: {ms_pv:.2f} ms; stages {st_pv}")
    del opt
    zero_grads(model)
    torch.cuda.empty_cache()

    # -- batched iteration, one captured graph ----------------------------------------------------
    from svraster_amd import autotune
    from svraster_amd.batched import BatchedTrainStep

    def make_step(shape):
        r = BatchedRasterizer(sc.octpath, sc.octlevel, sc.vox_center, sc.vox_size, sc.is_leaf, sc.vox_key,
                              cams, model.bg_color, sc.sh_degree, tile_shape=shape,
                              overlap_streams=not args.serial_streams)
        return BatchedTrainStep(r, model._geo_grid_pts.detach(), model._sh0.detach(), model._shs.detach(),
                                model._log_s.detach(), targets)

    best, tune_info = tune(make_step, autotune.workload_signature(sc.num_voxels, cams, sc.sh_degree),
                           not args.no_autotune)
    torch.cuda.empty_cache()
    t0 = time.perf_counter()
    rast = BatchedRasterizer(sc.octpath, sc.octlevel, sc.vox_center, sc.vox_size, sc.is_leaf, sc.vox_key,
                             cams, model.bg_color, sc.sh_degree, tile_shape=best["tile_shape"],
                             fuse_gather_bw=best["fuse_gather_bw"], overlap_streams=not args.serial_streams)
    it = BatchedTrainIteration(rast, model._geo_grid_pts.detach(), model._sh0.detach(), model._shs.detach(),
                               model._log_s.detach(), targets, smooth, lr=lr, betas=betas, eps=eps,
                               overlap_smooth=not args.serial_streams)
    it()
    sync()
    log(f"[bench] batched set-up {time.perf_counter() - t0:.1f} s, keys {rast.key_count()} / capacity {rast.cap}")
    st_b = {k: round(v, 3) for k, v in it.stage_times().items()}
    kt_b = kernel_times(lib, it)
    if not args.profile:
        it.capture()
    ms_b = time_steps(it, args.steps, args.warmup, sync)
    log(f"[bench] batched graph iteration: {ms_b:.2f} ms; stages (eager) {st_b}")

    nbytes = train_iteration_bytes(it, smooth)
    stage_gbs = {k: round(nbytes[k] / (st_b[k] / 1000.0) / 1e9, 1) for k in nbytes
                 if nbytes[k] > 0 and st_b.get(k, 0) > 0}
    roof = roofline_block(rast, kt_b, args.workload)
    it.check_capacity()
    value = rays / (ms_b / 1000.0) / 1e6
    result = {
        "metric": METRIC, "value": round(value, 3), "unit": "Mrays/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": round(ms_b, 3), "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f32",
        "data": "synthetic (seeded 48-sphere + 8-box SDF scene, random SH, random targets)",
        "config": {"workload": f"{args.workload}: {sc.num_voxels} voxels ({n_leaf} leaves, levels 7-13, parents "
                               f"kept from level 9), SH deg 3, {len(cams)} views 1920x1080, {rays} rays/step; "
                               f"training iteration: fwd + MSE + bwd + grid eikonal (weight {w_ge}) + Laplacian "
                               f"smoothness (weight {weight}) "
                               f"+ dense Adam (lr {lr}, betas {betas}, SparseAdam defaults)",
                   "path": "batched: " + ("eager (profiler run)" if args.profile
                                          else "one captured HIP graph per iteration")},
        "paths": {"batched_ms_per_step": round(ms_b, 3), "per_view_ms_per_step": round(ms_pv, 3),
                  "speedup_batched_vs_per_view": round(ms_pv / ms_b, 3),
                  "batched_mrays_s": round(value, 3), "per_view_mrays_s": round(rays / (ms_pv / 1000.0) / 1e6, 3),
                  "keys_per_step": rast.key_count(), "grid_keys": smooth.G, "grid_res": smooth.res},
        "stages_ms": {"batched_eager": st_b, "per_view": st_pv},
        "stages_algorithmic_bytes": {k: int(v) for k, v in nbytes.items()},
        "stages_gbs_batched": stage_gbs,
        "autotune": tune_info,
        "roofline": roof,
    }
    print(json.dumps(result), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--workload", default="multi_32")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--profile", action="store_true", help="short run for rocprofv3 (no CPU leg)")
    ap.add_argument("--serial-streams", action="store_true",
                    help="every stage on one stream (per-kernel times without concurrent kernels, for profiles)")
    ap.add_argument("--no-autotune", action="store_true", help="default launch configuration, no tuning")
    ap.add_argument("--batched-only", action="store_true",
                    help="profiling aid: skip the per-view leg (the JSON then carries no per-view numbers)")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only run a plain 1 GiB device copy (calibrates FETCH_SIZE / WRITE_SIZE counters)")
    args = ap.parse_args()
    if args.copy_calibration:
        import torch
        src = torch.ones(256 << 20, dtype=torch.float32, device="cuda")
        dst = torch.empty_like(src)
        for _ in range(3):
            torch.mul(src, 1.0, out=dst)        # elementwise kernel: reads 1 GiB, writes 1 GiB
        torch.cuda.synchronize()
        print(json.dumps({"copy_bytes_per_launch": src.numel() * 4, "launches": 3}), flush=True)
        return
    if args.gpus != 1:
        print(f"bench.py runs on exactly one GPU (got --gpus {args.gpus}); the multi-GPU path is "
              f"design-only here (one process per GPU is not built).", file=sys.stderr)
        sys.exit(2)

    if args.workload == "train_16":
        bench_train(args)
        return

    import torch
    from svraster_amd import _lib
    from svraster_amd.batched import BatchedRasterizer, BatchedTrainStep
    from svraster_amd.model import VoxelModel
    from svraster_amd.synthetic import make_workload

    dev = torch.device("cuda:0")
    t0 = time.perf_counter()
    w = make_workload(args.workload, device=dev)
    log(f"[bench] scene {args.workload}: {w.scene.num_voxels} voxels, {len(w.cameras)} views, "
        f"{w.rays} rays/step (built in {time.perf_counter() - t0:.1f} s)")
    model = VoxelModel(w.scene)
    cams = w.cameras
    targets = build_targets(cams, dev)
    sync = torch.cuda.synchronize
    lib = _lib.load()
    rays = w.rays

    # -- plain per-view path (drop-in API, one view per call) ----------------------------------
    def step_pv():
        zero_grads(model)
        per_view_step(model, cams, targets)

    if args.batched_only:
        step_pv = lambda: None
    ms_pv = time_steps(step_pv, args.steps, args.warmup, sync)
    kt_pv = kernel_times(lib, step_pv)
    if args.batched_only:
        ms_pv = float("nan")
    log(f"[bench] per-view path: {ms_pv:.2f} ms/step; compositor fwd {kt_pv['render_fwd'][0]:.2f} ms "
        f"bwd {kt_pv['render_bwd'][0]:.2f} ms per step")
    ref_colors = []
    if not args.batched_only:
        with torch.no_grad():
            for cam in cams[:: max(1, len(cams) // 4)]:
                ref_colors.append(model.render(cam)["color"].clone())
        zero_grads(model)
        per_view_step(model, cams, targets)
        sync()
        ref_grid = model._geo_grid_pts.grad.clone()
        ref_sh0 = model._sh0.grad.clone()

    # -- batched persistent path, autotuned, captured as a graph --------------------------------
    sc = w.scene
    from svraster_amd import autotune

    def make_step(shape):
        r = BatchedRasterizer(sc.octpath, sc.octlevel, sc.vox_center, sc.vox_size, sc.is_leaf, sc.vox_key,
                              cams, model.bg_color, sc.sh_degree, tile_shape=shape,
                              overlap_streams=not args.serial_streams)
        return BatchedTrainStep(r, model._geo_grid_pts.detach(), model._sh0.detach(), model._shs.detach(),
                                model._log_s.detach(), targets)

    best, tune_info = tune(make_step, autotune.workload_signature(sc.num_voxels, cams, sc.sh_degree),
                           not args.no_autotune)
    torch.cuda.empty_cache()
    bstep = make_step(best["tile_shape"])
    autotune.apply_options(best, bstep)
    rast = bstep.r
    bstep()
    sync()
    K_used = rast.key_count()
    img_err, grad_err = None, None
    if not args.batched_only:
        img_err = 0.0
        for i, v in enumerate(range(0, len(cams), max(1, len(cams) // 4))):
            img_err = max(img_err, (rast.view_slice(rast.color, 3, v) - ref_colors[i]).abs().max().item())
        rel = lambda a, b: float((a - b).double().norm() / b.double().norm().clamp_min(1e-30))
        grad_err = max(rel(bstep.grid_grad, ref_grid), rel(bstep.sh0_grad, ref_sh0))
    log(f"[bench] batched vs per-view: image max-abs {img_err}, grad rel-L2 {grad_err}, "
        f"keys {K_used} / capacity {rast.cap}")
    kt_b = kernel_times(lib, bstep)
    if not args.profile:          # profiler runs time the same step eagerly (no graph replay)
        bstep.capture()
    ms_b = time_steps(bstep, args.steps, args.warmup, sync)
    log(f"[bench] batched graph path: {ms_b:.2f} ms/step; compositor fwd {kt_b['render_fwd'][0]:.2f} ms "
        f"bwd {kt_b['render_bwd'][0]:.2f} ms")

    # -- roofline of the dominant kernel (the slower of the two compositor launches) --------------
    bstep()
    sync()
    roof = roofline_block(rast, kt_b, args.workload)
    bstep.check_capacity()

    value = rays / (ms_b / 1000.0) / 1e6
    value_pv = rays / (ms_pv / 1000.0) / 1e6
    result = {
        "metric": METRIC,
        "value": round(value, 3),
        "unit": "Mrays/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_b, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded sphere-union SDF scene, random SH, random targets)",
        "config": {"workload": f"{args.workload}: {sc.num_voxels} voxels (levels 6-12, SH deg 3), "
                               f"{len(cams)} views (8x640x480, 8x1237x822, 8x1600x1064, 8x1920x1080), "
                               f"{rays} rays/step, fwd+bwd of MSE to random targets",
                   "path": "batched: one launch per stage over all views, persistent (view, tile) queue, "
                           + ("eager (profiler run)" if args.profile else "HIP graph replay")},
        "paths": {"batched_mrays_s": round(value, 3), "batched_ms_per_step": round(ms_b, 3),
                  "per_view_mrays_s": round(value_pv, 3), "per_view_ms_per_step": round(ms_pv, 3),
                  "speedup_batched_vs_per_view": round(ms_pv / ms_b, 3),
                  "batched_vs_per_view_image_max_abs": img_err,
                  "batched_vs_per_view_grad_rel_l2": grad_err,
                  "keys_per_step": K_used},
        "roofline": roof,
        "autotune": tune_info,
        "kernels_ms_per_step": {"batched_render_fwd": round(kt_b["render_fwd"][0], 3),
                                "batched_render_bwd": round(kt_b["render_bwd"][0], 3),
                                "per_view_render_fwd": round(kt_pv["render_fwd"][0], 3),
                                "per_view_render_bwd": round(kt_pv["render_bwd"][0], 3)},
    }
    if not args.no_cpu_baseline and not args.profile:
        log("[bench] CPU oracle baseline ...")
        # whole frames of the two CPU-sized workloads (128x128 and 256x256, about 15 s together)
        frames = ("small_128", "fd_256")
        result["cpu_baseline"], crop_err = cpu_baseline(args.workload, rast.view_slice(rast.color, 3, 0), frames)
        result["paths"]["image_max_abs_vs_fp64_oracle_crop"] = crop_err
    print(json.dumps(result), flush=True)


if __name__ == "__main__":
    main()
