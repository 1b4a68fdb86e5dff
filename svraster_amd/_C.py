"""Python mirror of the reference's compiled `svraster_cuda._C` module (binding.cpp:28-72).

Same function names, argument order and return values as the reference's pybind11 binding;
each function validates its tensors, allocates the outputs with torch on the inputs' device and
calls the HIP library through its C-ABI on the current HIP stream.  Shape errors raise
RuntimeError like the reference's AT_ERROR.  Functions the reference exports but that are off
the training path raise NotImplementedError.
"""
from __future__ import annotations

import ctypes
import itertools

import torch

from . import _lib

# Constants (cuda/src/config.h:14-35, binding.cpp:60-72).
MAX_NUM_LEVELS = 16
STEP_SZ_SCALE = 100.0
VOX_TRIINTERP_MODE = 2
VOX_TRIINTERP1_MODE = 1
VOX_TRIINTERP3_MODE = 3
EXP_LINEAR_11_MODE = 0
SDF_MODE = 1
CAM_PERSP = 0
CAM_ORTHO = 10

_seed_counter = itertools.count(1)


def _ptr(t):
    if t is None:
        return None
    return ctypes.c_void_p(t.data_ptr()) if t.numel() > 0 else None


def _stream(device: torch.device):
    return ctypes.c_void_p(torch.cuda.current_stream(device).cuda_stream)


def _dev(t: torch.Tensor) -> torch.device:
    if not t.is_cuda:
        raise RuntimeError("expected a GPU tensor")
    return t.device


def _f32(t: torch.Tensor, name: str) -> torch.Tensor:
    if t.dtype != torch.float32:
        raise RuntimeError(f"{name} must be float32 (got {t.dtype})")
    return t.contiguous()


def _sync_if(debug: bool, dev):
    if debug:
        torch.cuda.synchronize(dev)


def _lib_call(name: str, *args):
    lib = _lib.load()
    _lib.check(getattr(lib, name)(*args), name)


# ----------------------------------------------------------------------------------------------
# Rasterizer
# ----------------------------------------------------------------------------------------------
def rasterize_preprocess(image_width, image_height, tan_fovx, tan_fovy, cx, cy, w2c_matrix,
                         c2w_matrix, cam_mode, near, octree_paths, vox_centers, vox_lengths,
                         is_leaf, debug=False):
    """preprocess.cu:158-228 -> (n_duplicates int32[P], geomBuffer uint8)."""
    if vox_centers.ndim != 2 or vox_centers.shape[1] != 3:
        raise RuntimeError("vox_centers must have dimensions (num_points, 3)")
    dev = _dev(vox_centers)
    P = vox_centers.shape[0]
    n_dup = torch.zeros(P, dtype=torch.int32, device=dev)
    if P == 0:
        return n_dup, torch.empty(0, dtype=torch.uint8, device=dev)
    lib = _lib.load()
    geom = torch.empty(int(lib.svr_geom_state_bytes(1, P)), dtype=torch.uint8, device=dev)
    w2c = _f32(w2c_matrix, "w2c_matrix")
    c2w = _f32(c2w_matrix, "c2w_matrix")
    cen = _f32(vox_centers, "vox_centers")
    lng = _f32(vox_lengths, "vox_lengths")
    leaf = is_leaf.to(torch.uint8).contiguous()
    paths = octree_paths.contiguous()
    if paths.dtype != torch.int64:
        raise RuntimeError("octree_paths must be int64")
    _lib_call("svr_rasterize_preprocess", int(image_width), int(image_height), float(tan_fovx),
              float(tan_fovy), float(cx), float(cy), _ptr(w2c), _ptr(c2w), int(cam_mode),
              float(near), _ptr(paths), _ptr(cen), _ptr(lng), _ptr(leaf), P, _ptr(n_dup),
              _ptr(geom), _stream(dev))
    _sync_if(debug, dev)
    return n_dup, geom


def rasterize_voxels(vox_geo_mode, density_mode, image_width, image_height, tan_fovx, tan_fovy,
                     cx, cy, w2c_matrix, c2w_matrix, background, cam_mode, need_depth,
                     need_distortion, need_normal, track_max_w, octree_paths, vox_centers,
                     vox_lengths, geos, rgbs, feats, geomBuffer, debug, s_val):
    """forward.cu:986-1096 -> (num_rendered, binningBuffer, imgBuffer, color, depth, normal, T,
    max_w, sdf0, feat)."""
    if vox_centers.ndim != 2 or vox_centers.shape[1] != 3:
        raise RuntimeError("vox_centers must have dimensions (num_points, 3)")
    if rgbs.ndim != 2 or rgbs.shape[1] != 3:
        raise RuntimeError("rgbs should be either (num_points, 3)")
    if feats.ndim != 2:
        raise RuntimeError("feats should be either (num_points, n_dim)")
    if feats.shape[1] > 0:
        raise NotImplementedError("feature rendering is not supported")
    if vox_centers.shape[0] != rgbs.shape[0]:
        raise RuntimeError("size mismatch")
    if need_distortion:
        raise NotImplementedError("distortion rendering (lambda_dist > 0) is not supported")
    dev = _dev(vox_centers)
    P = vox_centers.shape[0]
    H, W = int(image_height), int(image_width)
    fo = dict(dtype=torch.float32, device=dev)
    color = torch.zeros(3, H, W, **fo)
    depth = torch.zeros(4 if density_mode == SDF_MODE else 3, H, W, **fo) if need_depth else torch.empty(0)
    normal = torch.zeros(3, H, W, **fo) if need_normal else torch.empty(0)
    T = torch.zeros(1, H, W, **fo)
    max_w = torch.zeros(P, 1, **fo) if track_max_w else torch.empty(0)
    sdf0 = torch.full((H * W,), -1.0, **fo)
    feat = torch.zeros(0, H, W, **fo)
    lib = _lib.load()
    n_tiles = ((W + 15) // 16) * ((H + 15) // 16)
    if P == 0:
        return (0, torch.empty(0, dtype=torch.uint8, device=dev),
                torch.empty(0, dtype=torch.uint8, device=dev), color, depth, normal, T, max_w, sdf0, feat)
    K = ctypes.c_longlong(0)
    shift = ctypes.c_int(0)
    st = _stream(dev)
    _lib_call("svr_rasterize_count", _ptr(geomBuffer), P, ctypes.byref(K), ctypes.byref(shift), st)
    K = int(K.value)
    binning = torch.empty(int(lib.svr_binning_state_bytes(K)), dtype=torch.uint8, device=dev)
    img = torch.empty(int(lib.svr_image_state_bytes(1, W * H, n_tiles)), dtype=torch.uint8, device=dev)
    paths = octree_paths.contiguous()
    _lib_call("svr_rasterize_forward", int(vox_geo_mode), int(density_mode), W, H, float(tan_fovx),
              float(tan_fovy), float(cx), float(cy), _ptr(_f32(w2c_matrix, "w2c")),
              _ptr(_f32(c2w_matrix, "c2w")), _ptr(_f32(background, "background")), int(cam_mode),
              int(bool(need_depth)), 0, int(bool(need_normal)), _ptr(paths),
              _ptr(_f32(vox_centers, "vox_centers")), _ptr(_f32(vox_lengths, "vox_lengths")),
              _ptr(_f32(geos, "geos")), _ptr(_f32(rgbs, "rgbs")), P, _ptr(geomBuffer), K,
              int(shift.value), _ptr(binning), _ptr(img), _ptr(color),
              _ptr(depth) if need_depth else None, _ptr(normal) if need_normal else None, _ptr(T),
              _ptr(max_w) if track_max_w else None, _ptr(_f32(s_val, "s_val")), st)
    _sync_if(debug, dev)
    return K, binning, img, color, depth, normal, T, max_w, sdf0, feat


def rasterize_voxels_backward(R, vox_geo_mode, density_mode, image_width, image_height, tan_fovx,
                              tan_fovy, cx, cy, w2c_matrix, c2w_matrix, background, octree_paths,
                              vox_centers, vox_lengths, geos, rgbs, geomBuffer, binningBuffer,
                              imageBuffer, out_T, dL_dout_color, dL_dout_depth, dL_dout_normal,
                              dL_dout_T, lambda_R_concen, gt_color, lambda_ascending, lambda_dist,
                              need_depth, need_normal, out_D, out_N, debug, s_val, sdf0_buffer):
    """backward.cu:986-1109 -> (dL_dgeos [P,8], dL_drgbs [P,3], subdiv_p_bw [P,1], dL_ds [1])."""
    if vox_centers.ndim != 2 or vox_centers.shape[1] != 3:
        raise RuntimeError("vox_centers must have dimensions (num_points, 3)")
    if lambda_ascending > 0:
        raise NotImplementedError("lambda_ascending > 0 is not supported")
    if lambda_dist > 0:
        raise NotImplementedError("lambda_dist > 0 is not supported")
    dev = _dev(vox_centers)
    P = vox_centers.shape[0]
    if P == 0:
        e = torch.empty(0)
        return e, e, e, e
    H, W = int(image_height), int(image_width)
    dvox = torch.zeros(P, 12, dtype=torch.float32, device=dev)
    ds = torch.zeros(1, dtype=torch.float32, device=dev)
    use_depth = bool(need_depth) and dL_dout_depth is not None and dL_dout_depth.numel() > 0
    use_normal = bool(need_normal) and dL_dout_normal is not None and dL_dout_normal.numel() > 0
    use_concen = lambda_R_concen > 0
    if use_concen and (gt_color is None or tuple(gt_color.shape) != (3, H, W)):
        raise RuntimeError("Except gt_color in shape of [3, H, W]")
    dT = dL_dout_T if (dL_dout_T is not None and dL_dout_T.numel() > 0) else torch.zeros(1, H, W, device=dev)
    _lib_call("svr_rasterize_backward", int(R), int(vox_geo_mode), int(density_mode), W, H,
              float(tan_fovx), float(tan_fovy), float(cx), float(cy), _ptr(_f32(w2c_matrix, "w2c")),
              _ptr(_f32(c2w_matrix, "c2w")), _ptr(_f32(background, "background")),
              _ptr(_f32(vox_centers, "vox_centers")), _ptr(_f32(vox_lengths, "vox_lengths")),
              _ptr(_f32(geos, "geos")), _ptr(_f32(rgbs, "rgbs")), P, _ptr(geomBuffer),
              _ptr(binningBuffer), _ptr(imageBuffer), _ptr(_f32(out_T, "out_T")),
              _ptr(_f32(dL_dout_color, "dL_dout_color")),
              _ptr(_f32(dL_dout_depth, "dL_dout_depth")) if use_depth else None,
              _ptr(_f32(dL_dout_normal, "dL_dout_normal")) if use_normal else None,
              _ptr(_f32(dT, "dL_dout_T")), float(lambda_R_concen),
              _ptr(_f32(gt_color, "gt_color")) if use_concen else None, float(lambda_ascending),
              float(lambda_dist), int(use_depth), int(use_normal), _ptr(_f32(s_val, "s_val")),
              _ptr(dvox), _ptr(ds), _stream(dev))
    _sync_if(debug, dev)
    dgeo, drgb, dsub = dvox.split([8, 3, 1], 1)
    return dgeo, drgb, dsub, ds


def unpack_ImageState(image_width, image_height, imgBuffer):
    """(ranges [tiles_y, tiles_x, 2], tile_last [tiles_y, tiles_x], n_contrib [H, W]) views of
    the image state (debug helper, raster_state.cu:120-171 role)."""
    W, H = int(image_width), int(image_height)
    tx, ty = (W + 15) // 16, (H + 15) // 16
    nt = tx * ty

    off = _a256(128) + _a256(64 * 4)     # one view descriptor, counters
    ranges = imgBuffer[off:off + nt * 8].view(torch.int32).view(ty, tx, 2)
    off = _a256(off + nt * 8)
    tile_last = imgBuffer[off:off + nt * 4].view(torch.int32).view(ty, tx)
    off = _a256(off + nt * 4)
    n_contrib = imgBuffer[off:off + W * H * 4].view(torch.int32).view(H, W)
    return ranges, tile_last, n_contrib


def _a256(x):
    return (x + 255) & ~255


def unpack_GeomState(P, geomBuffer):
    """(n_dup uint32 [P], inclusive scan [P], bbox [P,2], octant bitsets uint8 [P]) views of a
    single-view geometry state (debug helper; layout of the library's geometry state)."""
    off = _a256(128) + _a256(64 * 4)
    n_dup = geomBuffer[off:off + 4 * P].view(torch.int32)
    off = _a256(off + 4 * P)
    scan = geomBuffer[off:off + 4 * P].view(torch.int32)
    off = _a256(off + 4 * P)
    bbox = geomBuffer[off:off + 8 * P].view(torch.int32).view(P, 2)
    off = _a256(off + 8 * P)
    quad = geomBuffer[off:off + P]
    return n_dup, scan, bbox, quad


def unpack_BinningState(K, binningBuffer):
    """(sorted compressed keys int64 [K], sorted values int32 [K]) views (debug helper).
    Values are octant << 29 | voxel id, as in the reference."""
    keys = binningBuffer[0:8 * K].view(torch.int64)
    off = _a256(8 * K)
    vals = binningBuffer[off:off + 4 * K].view(torch.int32)
    return keys, vals


def filter_geomState(*args, **kwargs):
    raise NotImplementedError("filter_geomState is a debug helper not provided by this library")


# ----------------------------------------------------------------------------------------------
# Voxel parameter operators
# ----------------------------------------------------------------------------------------------
def gather_triinterp_geo_params(vox_key, care_idx, grid_pts):
    """geo_params_gather.cu:134-152.  Rows outside care_idx are left uninitialised."""
    dev = _dev(grid_pts)
    P = vox_key.shape[0]
    geos = torch.empty(P, 8, dtype=torch.float32, device=dev)
    care = care_idx.to(device=dev, dtype=torch.int64).contiguous()
    _lib_call("svr_gather_triinterp_geo_params", _ptr(vox_key.contiguous()), P, _ptr(care), care.numel(),
              _ptr(_f32(grid_pts, "grid_pts")), _ptr(geos), _stream(dev))
    return geos


def gather_triinterp_geo_params_bw(vox_key, care_idx, num_grid_pts, dL_dgeo_params):
    """geo_params_gather.cu:154-173 -> dL_dgrid_pts [M,1]."""
    dev = _dev(dL_dgeo_params)
    out = torch.zeros(int(num_grid_pts), 1, dtype=torch.float32, device=dev)
    care = care_idx.to(device=dev, dtype=torch.int64).contiguous()
    _lib_call("svr_gather_triinterp_geo_params_bw", _ptr(vox_key.contiguous()), vox_key.shape[0],
              _ptr(care), care.numel(), _ptr(_f32(dL_dgeo_params, "dL_dgeo_params")), _ptr(out),
              _stream(dev))
    return out


def gather_triinterp_feat_params(*args, **kwargs):
    raise NotImplementedError("feature-grid gather is not on the SDF training path")


def gather_triinterp_feat_params_bw(*args, **kwargs):
    raise NotImplementedError("feature-grid gather is not on the SDF training path")


def sh_compute(D, indices, vox_centers, cam_pos, sh0, shs):
    """sh_compute.cu:244-270 -> rgbs [P,3] (zero outside `indices`)."""
    dev = _dev(vox_centers)
    P = vox_centers.shape[0]
    M = 1 + shs.shape[1]
    rgbs = torch.zeros(P, 3, dtype=torch.float32, device=dev)
    if P == 0:
        return rgbs
    idx = indices.to(device=dev, dtype=torch.int64).contiguous() if indices is not None else None
    n_idx = idx.numel() if idx is not None else 0
    _lib_call("svr_sh_compute", int(D), _ptr(idx) if n_idx else None, n_idx,
              _ptr(_f32(vox_centers, "vox_centers")), P,
              _ptr(_f32(cam_pos.to(dev), "cam_pos")), _ptr(_f32(sh0, "sh0")),
              _ptr(_f32(shs, "shs")) if shs.numel() else None, M, _ptr(rgbs), _stream(dev))
    return rgbs


def sh_compute_bw(D, M, indices, vox_centers, cam_pos, rgbs, dL_drgbs):
    """sh_compute.cu:272-299 -> (dL_dsh0 [P,3], dL_dshs [P,M-1,3])."""
    dev = _dev(vox_centers)
    P = vox_centers.shape[0]
    dsh0 = torch.zeros(P, 3, dtype=torch.float32, device=dev)
    dshs = torch.zeros(P, int(M) - 1, 3, dtype=torch.float32, device=dev)
    if P == 0:
        return dsh0, dshs
    idx = indices.to(device=dev, dtype=torch.int64).contiguous() if indices is not None else None
    n_idx = idx.numel() if idx is not None else 0
    _lib_call("svr_sh_compute_bw", int(D), int(M), _ptr(idx) if n_idx else None, n_idx,
              _ptr(_f32(vox_centers, "vox_centers")), P, _ptr(_f32(cam_pos.to(dev), "cam_pos")),
              _ptr(_f32(dL_drgbs, "dL_drgbs")), _ptr(dsh0), _ptr(dshs) if dshs.numel() else None,
              _stream(dev))
    return dsh0, dshs


def _inplace_ok(t: torch.Tensor, name: str):
    if not t.is_contiguous():
        raise RuntimeError(f"{name} must be contiguous for an in-place update")
    if t.dtype != torch.float32:
        raise RuntimeError(f"{name} must be float32")


def unbiased_adam_step(sparse, param, grad, exp_avg, exp_avg_sq, step, lr, beta1, beta2, eps):
    """adam_step.cu:106-135 (in place)."""
    for t, n in ((param, "param"), (exp_avg, "exp_avg"), (exp_avg_sq, "exp_avg_sq")):
        _inplace_ok(t, n)
    dev = _dev(param)
    _lib_call("svr_unbiased_adam_step", int(bool(sparse)), _ptr(param), _ptr(_f32(grad, "grad")),
              _ptr(exp_avg), _ptr(exp_avg_sq), param.numel(), float(step), float(lr),
              float(beta1), float(beta2), float(eps), _stream(dev))


def biased_adam_step(sparse, param, grad, exp_avg, exp_avg_sq, lr, beta1, beta2, eps):
    """adam_step.cu:137-158 (in place)."""
    for t, n in ((param, "param"), (exp_avg, "exp_avg"), (exp_avg_sq, "exp_avg_sq")):
        _inplace_ok(t, n)
    dev = _dev(param)
    _lib_call("svr_biased_adam_step", int(bool(sparse)), _ptr(param), _ptr(_f32(grad, "grad")),
              _ptr(exp_avg), _ptr(exp_avg_sq), param.numel(), float(lr), float(beta1),
              float(beta2), float(eps), _stream(dev))


# ----------------------------------------------------------------------------------------------
# Grid losses
# ----------------------------------------------------------------------------------------------
# Table-order plan of the last grid table seen by the grid losses (one: a plan holds two res^3
# index maps), keyed by the identity and version of the tensors that define the table.
_plan_cache = {}


def _table_key(*ts):
    return tuple((t.data_ptr(), t._version, tuple(t.shape)) for t in ts)


def _table_order_plan(grid_keys, grid2voxel, coord, size, res, vox_key):
    from .batched import TableOrderPlan
    key = _table_key(grid_keys, grid2voxel, coord, size, vox_key) + (int(res),)
    hit = _plan_cache.get("plan")
    if hit is not None and hit[0] == key:
        return hit[1]
    _plan_cache.clear()
    keys = grid_keys.to(torch.int32).contiguous()
    plan = None
    if keys.numel() > 0 and bool((keys[1:] > keys[:-1]).all()):       # sorted and unique, as train.py passes
        plan = TableOrderPlan(keys, grid2voxel.to(torch.int32).contiguous(), _f32(coord, "grid_voxel_coord"),
                              _f32(size.reshape(-1), "grid_voxel_size"), int(res), vox_key.contiguous())
        plan = plan if plan.ok else None
        if plan is not None:
            plan.step0 = torch.zeros(1, dtype=torch.int32, device=keys.device)
            plan.weights = torch.zeros(2, dtype=torch.float32, device=keys.device)
            plan.coef = None
    _plan_cache["plan"] = (key, plan)
    return plan


def _grid_loss(name, grid_pts, vox_key, grid_voxel_coord, grid_voxel_size, grid_res, grid_mask,
               grid_keys, grid2voxel, active_list, weight, vox_size_inv, grid_pts_grad, seed, randv):
    _inplace_ok(grid_pts_grad, "grid_pts_grad")
    dev = _dev(grid_pts)
    if seed is None:
        seed = next(_seed_counter)
    active = active_list.to(device=dev, dtype=torch.int32).contiguous()
    G = grid_keys.numel()
    if randv is None and active.numel() == G and G > 0:
        # Every key active once (the trainer's randperm list with drop ratio 0): the table-order
        # form gives each key the same terms as the per-key scatter, with the random sample
        # offsets drawn per table entry instead of per list position, and sums each voxel's
        # contributions before touching the gradient.
        plan = _table_order_plan(grid_keys, grid2voxel, grid_voxel_coord, grid_voxel_size, grid_res, vox_key)
        if plan is not None and int(torch.bincount(active.long(), minlength=G).max()) == 1:
            eik = name == "svr_grid_eikonal_bw"
            plan.weights.copy_(torch.tensor([float(weight) if eik else 0.0, 0.0 if eik else float(weight)]))
            n_coef = (7 if eik else 3) * G
            if plan.coef is None or plan.coef.numel() < n_coef:
                plan.coef = torch.empty(n_coef, dtype=torch.float32, device=dev)
            mask = grid_mask.to(torch.uint8).contiguous() if grid_mask.dtype != torch.uint8 else grid_mask.contiguous()
            plan.run(_f32(grid_pts, "grid_pts"), grid_pts_grad, mask, 2 if eik else 1, plan.weights,
                     float(vox_size_inv), int(seed), plan.step0, plan.coef)
            return
    rv = None
    if randv is not None:
        rv = _f32(randv.to(dev), "randv")
        if rv.numel() != 3 * active.numel():
            raise RuntimeError("randv must have shape [A, 3]")
    _lib_call(name, _ptr(_f32(grid_pts, "grid_pts")), grid_pts.numel(), _ptr(vox_key.contiguous()),
              vox_key.shape[0], _ptr(_f32(grid_voxel_coord, "grid_voxel_coord")),
              _ptr(_f32(grid_voxel_size, "grid_voxel_size")), int(grid_res),
              _ptr(grid_mask.to(torch.uint8).contiguous()) if grid_mask.dtype != torch.uint8
              else _ptr(grid_mask.contiguous()),
              _ptr(grid_keys.to(torch.int32).contiguous()), _ptr(grid2voxel.to(torch.int32).contiguous()),
              grid_keys.numel(), _ptr(active), active.numel(), float(weight), float(vox_size_inv),
              int(seed) & 0xFFFFFFFF, _ptr(rv), _ptr(grid_pts_grad), _stream(dev))


def grid_eikonal_bw(grid_pts, vox_key, grid_voxel_coord, grid_voxel_size, grid_res, grid_mask,
                    grid_keys, grid2voxel, active_list, weight, vox_size_inv, no_tv_s, tv_sparse,
                    grid_pts_grad, seed=None, randv=None):
    """ge_compute.cu:120-161 (accumulates into grid_pts_grad)."""
    _grid_loss("svr_grid_eikonal_bw", grid_pts, vox_key, grid_voxel_coord, grid_voxel_size, grid_res,
               grid_mask, grid_keys, grid2voxel, active_list, weight, vox_size_inv, grid_pts_grad, seed,
               randv)


def laplacian_smoothness_bw(grid_pts, vox_key, grid_voxel_coord, grid_voxel_size, grid_res,
                            grid_mask, grid_keys, grid2voxel, active_list, weight, vox_size_inv,
                            no_tv_s, tv_sparse, grid_pts_grad, seed=None, randv=None):
    """ls_compute.cu:112-155 (accumulates into grid_pts_grad)."""
    _grid_loss("svr_laplacian_smoothness_bw", grid_pts, vox_key, grid_voxel_coord, grid_voxel_size,
               grid_res, grid_mask, grid_keys, grid2voxel, active_list, weight, vox_size_inv,
               grid_pts_grad, seed, randv)


def total_variation_bw(grid_pts, vox_key, weight, vox_size_inv, no_tv_s, tv_sparse, grid_pts_grad):
    """tv_compute.cu:162-189 (accumulates into grid_pts_grad [M,C])."""
    _inplace_ok(grid_pts_grad, "grid_pts_grad")
    dev = _dev(grid_pts)
    g = _f32(grid_pts, "grid_pts")
    C = g.shape[1] if g.dim() > 1 else 1
    N = vox_key.shape[0]
    vsi = None if no_tv_s else _f32(vox_size_inv.reshape(-1), "vox_size_inv")
    ws = torch.empty(N * C if tv_sparse else 0, dtype=torch.uint8, device=dev)
    _lib_call("svr_total_variation_bw", _ptr(g), g.shape[0], C, _ptr(vox_key.contiguous()), N, float(weight),
              _ptr(vsi), int(bool(no_tv_s)), int(bool(tv_sparse)), _ptr(ws), _ptr(grid_pts_grad), _stream(dev))


def voxel_gradient_bw(grid_pts, vox_key, vox_size_inv, active_list, weight, no_tv_s, tv_sparse, grid_pts_grad):
    """vg_compute.cu:78-103 (accumulates into grid_pts_grad; no_tv_s and tv_sparse are unused
    there as well)."""
    _inplace_ok(grid_pts_grad, "grid_pts_grad")
    dev = _dev(grid_pts)
    active = active_list.to(device=dev, dtype=torch.int32).contiguous()
    _lib_call("svr_voxel_gradient_bw", _ptr(_f32(grid_pts, "grid_pts")), grid_pts.shape[0],
              _ptr(vox_key.contiguous()), vox_key.shape[0], _ptr(_f32(vox_size_inv.reshape(-1), "vox_size_inv")),
              _ptr(active), active.numel(), float(weight), _ptr(grid_pts_grad), _stream(dev))


def points_loss_bw(points_in_grid, grid_pts, vox_key, grid_voxel_coord, grid_voxel_size, grid_res, grid_mask,
                   grid_keys, grid2voxel, weight, vox_size_inv, no_tv_s, tv_sparse, grid_pts_grad):
    """pl_compute.cu:69-104 (accumulates into grid_pts_grad; vox_size_inv, no_tv_s and
    tv_sparse are unused there as well)."""
    _inplace_ok(grid_pts_grad, "grid_pts_grad")
    dev = _dev(grid_pts)
    pts = _f32(points_in_grid.to(dev), "points_in_grid")
    mask = grid_mask.to(torch.uint8).contiguous() if grid_mask.dtype != torch.uint8 else grid_mask.contiguous()
    _lib_call("svr_points_loss_bw", _ptr(pts), pts.shape[0], _ptr(_f32(grid_pts, "grid_pts")), grid_pts.shape[0],
              _ptr(vox_key.contiguous()), vox_key.shape[0], _ptr(_f32(grid_voxel_coord, "grid_voxel_coord")),
              _ptr(_f32(grid_voxel_size.reshape(-1), "grid_voxel_size")), int(grid_res), _ptr(mask),
              _ptr(grid_keys.to(torch.int32).contiguous()), _ptr(grid2voxel.to(torch.int32).contiguous()),
              grid_keys.numel(), float(weight), _ptr(grid_pts_grad), _stream(dev))


# ----------------------------------------------------------------------------------------------
# Octree utilities
# ----------------------------------------------------------------------------------------------
def ijk_2_octpath(ijk, octlevel):
    dev = _dev(ijk)
    n = octlevel.numel()
    out = torch.empty(n, 1, dtype=torch.int64, device=dev)
    _lib_call("svr_ijk_2_octpath", _ptr(ijk.contiguous()), _ptr(octlevel.contiguous()), n, _ptr(out),
              _stream(dev))
    return out


def octpath_2_ijk(octpath, octlevel):
    dev = _dev(octpath)
    n = octlevel.numel()
    out = torch.empty(n, 3, dtype=torch.int64, device=dev)
    _lib_call("svr_octpath_2_ijk", _ptr(octpath.contiguous()), _ptr(octlevel.contiguous()), n, _ptr(out),
              _stream(dev))
    return out


def valid_gradient_table(vox_center, vox_size, scene_center, inside_extent, grid_res_pow2, is_leaf):
    """utils.cu:362-408 -> (grid_mask bool[res^3], grid_keys int32[G], grid2voxel int32[G]);
    keys come in voxel order (deterministic)."""
    dev = _dev(vox_center)
    N = vox_center.shape[0]
    res = 1 << int(grid_res_pow2)
    mask = torch.zeros(res ** 3, dtype=torch.bool, device=dev)
    lib = _lib.load()
    ws = torch.empty(max(int(lib.svr_valid_gradient_table_workspace_bytes(N)), 1), dtype=torch.uint8, device=dev)
    cen = _f32(vox_center, "vox_center")
    size = _f32(vox_size.reshape(-1), "vox_size")
    sc = _f32(scene_center.to(dev).reshape(-1), "scene_center")
    leaf = is_leaf.reshape(-1).to(torch.uint8).contiguous()
    cnt = ctypes.c_longlong(0)
    st = _stream(dev)
    _lib_call("svr_valid_gradient_table_count", _ptr(cen), _ptr(size), _ptr(sc), float(inside_extent),
              int(grid_res_pow2), _ptr(leaf), N, _ptr(ws), ctypes.byref(cnt), st)
    G = int(cnt.value)
    keys = torch.empty(G, dtype=torch.int32, device=dev)
    g2v = torch.empty(G, dtype=torch.int32, device=dev)
    _lib_call("svr_valid_gradient_table_fill", _ptr(cen), _ptr(size), _ptr(sc), float(inside_extent),
              int(grid_res_pow2), _ptr(leaf), N, _ptr(ws), _ptr(mask), _ptr(keys), _ptr(g2v), st)
    return mask, keys, g2v


def is_in_cone(tanfovx, tanfovy, near, w2c_matrix, pts):
    """utils.cu:254-274 -> bool [N]."""
    dev = _dev(pts)
    if pts.ndim != 2 or pts.shape[1] != 3:
        raise RuntimeError("pts must have dimensions (num_points, 3)")
    N = pts.shape[0]
    out = torch.zeros(N, dtype=torch.bool, device=dev)
    if N == 0:
        return out
    _lib_call("svr_is_in_cone", float(tanfovx), float(tanfovy), float(near),
              _ptr(_f32(w2c_matrix.to(dev), "w2c_matrix")), _ptr(_f32(pts, "pts")), N, _ptr(out), _stream(dev))
    return out


def _c2w_on(c2w_matrix):
    if c2w_matrix.is_cuda:
        return c2w_matrix
    return c2w_matrix.to(torch.device("cuda", torch.cuda.current_device()))


def compute_rd(width, height, cx, cy, tanfovx, tanfovy, c2w_matrix):
    """utils.cu:276-292 -> float32 [3, height, width] world-space pixel ray directions."""
    c2w = _f32(_c2w_on(c2w_matrix), "c2w_matrix")
    dev = c2w.device
    W, H = int(width), int(height)
    rd = torch.empty(3, H, W, dtype=torch.float32, device=dev)
    _lib_call("svr_compute_rd", W, H, float(cx), float(cy), float(tanfovx), float(tanfovy), _ptr(c2w), _ptr(rd),
              _stream(dev))
    return rd


def depth2pts(width, height, cx, cy, tanfovx, tanfovy, c2w_matrix, depth):
    """utils.cu:294-310 -> float32 [3, height, width] world points of a depth map."""
    c2w = _f32(_c2w_on(c2w_matrix), "c2w_matrix")
    dev = c2w.device
    W, H = int(width), int(height)
    d = _f32(depth.to(dev), "depth")
    if d.numel() != W * H:
        raise RuntimeError("depth must hold width * height values")
    pts = torch.empty(3, H, W, dtype=torch.float32, device=dev)
    _lib_call("svr_depth2pts", W, H, float(cx), float(cy), float(tanfovx), float(tanfovy), _ptr(c2w), _ptr(d),
              _ptr(pts), _stream(dev))
    return pts


def voxel_order_rank(octree_paths):
    """utils.cu:312-327 -> int64 [8, P]: octpath XOR the octant repeated in every level field."""
    dev = _dev(octree_paths)
    paths = octree_paths.contiguous()
    if paths.dtype != torch.int64:
        raise RuntimeError("octree_paths must be int64")
    P = paths.shape[0]
    out = torch.empty(8, P, dtype=torch.int64, device=dev)
    if P > 0:
        _lib_call("svr_voxel_order_rank", _ptr(paths), P, _ptr(out), _stream(dev))
    return out
