"""ctypes binding of libsvraster_hip.so (C-ABI declared in include/svraster_hip.h).

The library is built in-tree by `python -m svraster_amd.build`.  Loading fails loudly when it
is missing: there is no fallback implementation.
"""
from __future__ import annotations

import ctypes
import os
import re

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libsvraster_hip.so")
HEADER_PATH = os.path.join(os.path.dirname(_HERE), "include", "svraster_hip.h")

P = ctypes.c_void_p
I = ctypes.c_int
U = ctypes.c_uint
F = ctypes.c_float
D = ctypes.c_double
LL = ctypes.c_longlong
SZ = ctypes.c_size_t

_SIGS = {
    "svr_last_error": (ctypes.c_char_p, []),
    "svr_abi_version": (I, []),
    "svr_timing_enable": (None, [I]),
    "svr_set_option": (I, [ctypes.c_char_p, I]),
    "svr_get_option": (I, [ctypes.c_char_p]),
    "svr_timing_read": (I, [I, I, ctypes.POINTER(D), ctypes.POINTER(LL)]),
    "svr_geom_state_bytes": (SZ, [I, I]),
    "svr_binning_state_bytes": (SZ, [LL]),
    "svr_image_state_bytes": (SZ, [I, LL, LL]),
    "svr_rasterize_preprocess": (I, [I, I, F, F, F, F, P, P, I, F, P, P, P, P, I, P, P, P]),
    "svr_rasterize_count": (I, [P, I, ctypes.POINTER(LL), ctypes.POINTER(I), P]),
    "svr_rasterize_forward": (I, [I, I, I, I, F, F, F, F, P, P, P, I, I, I, I, P, P, P, P, P, I,
                                  P, LL, I, P, P, P, P, P, P, P, P, P]),
    "svr_rasterize_backward": (I, [LL, I, I, I, I, F, F, F, F, P, P, P, P, P, P, P, I, P, P, P,
                                   P, P, P, P, P, F, P, F, F, I, I, P, P, P, P]),
    "svr_sh_compute": (I, [I, P, I, P, I, P, P, P, I, P, P]),
    "svr_sh_compute_bw": (I, [I, I, P, I, P, I, P, P, P, P, P]),
    "svr_gather_triinterp_geo_params": (I, [P, I, P, I, P, P, P]),
    "svr_gather_triinterp_geo_params_bw": (I, [P, I, P, I, P, P, P]),
    "svr_batch_gather_bw": (I, [I, P, P, P, P]),
    "svr_unbiased_adam_step": (I, [I, P, P, P, P, LL, D, D, D, D, F, P]),
    "svr_biased_adam_step": (I, [I, P, P, P, P, LL, F, F, F, F, P]),
    "svr_grid_eikonal_bw": (I, [P, LL, P, I, P, P, I, P, P, P, I, P, I, F, F, U, P, P, P]),
    "svr_laplacian_smoothness_bw": (I, [P, LL, P, I, P, P, I, P, P, P, I, P, I, F, F, U, P, P, P]),
    "svr_ijk_2_octpath": (I, [P, P, I, P, P]),
    "svr_octpath_2_ijk": (I, [P, P, I, P, P]),
    "svr_compute_rd": (I, [I, I, F, F, F, F, P, P, P]),
    "svr_depth2pts": (I, [I, I, F, F, F, F, P, P, P, P]),
    "svr_is_in_cone": (I, [F, F, F, P, P, LL, P, P]),
    "svr_voxel_order_rank": (I, [P, I, P, P]),
    "svr_selfcheck_rcp_rn": (I, [I, I, P, P]),
    "svr_selfcheck_div_rn": (I, [I, U, P, P]),
    "svr_valid_gradient_table_workspace_bytes": (SZ, [I]),
    "svr_valid_gradient_table_count": (I, [P, P, P, F, I, P, I, P, ctypes.POINTER(LL), P]),
    "svr_valid_gradient_table_fill": (I, [P, P, P, F, I, P, I, P, P, P, P, P]),
    "svr_batch_workspace_bytes": (SZ, [I, I, LL, LL, LL]),
    "svr_batch_setup": (I, [P, I, I, LL, LL, LL, P, P, P, P, P, P, P, P, P, P, I, I, P]),
    "svr_batch_pair_order_bytes": (SZ, [I]),
    "svr_batch_pair_order": (I, [I, P, P, P, P]),
    "svr_batch_pair_slots": (I, [I, P, P, P]),
    "svr_batch_preprocess": (I, [P, I, I, LL, LL, LL, P, P, P, P, P, P]),
    "svr_voxel_activity": (I, [I, P, P, P, P, P, F, P, P]),
    "svr_batch_sh_forward": (I, [P, I, I, LL, LL, LL, P, P, P, I, I, P, P]),
    "svr_batch_voxel_records": (I, [I, P, P, P, P, P, P, F, I, P, P, P]),
    "svr_batch_rasterize": (I, [P, I, I, LL, LL, LL, P, P, P, P, I, I, P, P, P, P, P, P, P]),
    "svr_batch_bin": (I, [P, I, I, LL, LL, LL, P, P]),
    "svr_batch_composite": (I, [P, I, I, LL, LL, LL, P, P, P, I, I, P, P, P, P, P, P, P]),
    "svr_batch_backward": (I, [P, I, I, LL, LL, LL, P, P, P, P, P, P, P, P, I, I, P, P, P, P, P, P]),
    "svr_batch_sh_backward": (I, [P, I, I, LL, LL, LL, P, P, I, I, P, P, P]),
    "svr_batch_sh_backward_adam": (I, [P, I, I, LL, LL, LL, P, P, I, I, P, P, P, P, P, P, P, P, I, P, F, F, F, I, P]),
    "svr_adam_bias_table": (I, [D, D, D, I, P]),
    "svr_sparse_adam_step_dev": (I, [I, P, P, P, P, LL, P, P, I, F, F, F, P]),
    "svr_grid_dense_index": (I, [P, I, I, P, P]),
    "svr_grid_pts_link_workspace_bytes": (SZ, [LL]),
    "svr_build_grid_pts_link": (I, [P, P, LL, P, P, P, P, P]),
    "svr_agg_voxel_into_grid_pts": (I, [P, LL, P, I, LL, P, P, P]),
    "svr_agg_voxel_into_grid_pts_noalloc": (I, [P, LL, P, I, LL, P, P]),
    "svr_total_variation_bw": (I, [P, LL, I, P, LL, F, P, I, I, P, P, P]),
    "svr_voxel_gradient_bw": (I, [P, LL, P, LL, P, P, I, F, P, P]),
    "svr_points_loss_bw": (I, [P, I, P, LL, P, I, P, P, I, P, P, P, I, F, P, P]),
    "svr_grid_smoothness_bw_step": (I, [I, P, LL, P, I, P, P, I, P, P, P, I, P, P, I, F, F, U, P, P, P]),
    "svr_grid_laplacian_bw_table_order": (I, [P, LL, P, I, P, P, I, P, P, P, I, P, P, P, I, F, F, U, P, P, P, P, P]),
    "svr_grid_smoothness_bw_table_order": (I, [P, LL, P, I, P, P, I, P, P, P, I, P, P, P, I, I, P, F, U, P, P, P, P,
                                               P]),
    "svr_batch_state_offsets": (I, [I, I, LL, LL, LL, P]),
    "svr_batch_key_count": (I, [P, I, I, LL, LL, LL, ctypes.POINTER(LL), P]),
    "svr_batch_key_peak": (I, [P, I, I, LL, LL, LL, I, ctypes.POINTER(LL), P]),
}

_lib = None


def header_symbols():
    """Names of every function declared in the C-ABI header."""
    with open(HEADER_PATH) as f:
        text = f.read()
    return sorted(set(re.findall(r"\b(svr_[a-z0-9_]+)\s*\(", text)))


def load():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(f"{LIB_PATH} is not built; run `python -m svraster_amd.build` first")
    lib = ctypes.CDLL(LIB_PATH)
    for name, (res, args) in _SIGS.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def set_option(name: str, value: int):
    check(load().svr_set_option(name.encode(), int(value)), f"set_option({name})")


def get_option(name: str) -> int:
    return int(load().svr_get_option(name.encode()))


def check(rc: int, what: str = ""):
    if rc != 0:
        msg = load().svr_last_error().decode(errors="replace")
        raise RuntimeError(f"{what}: {msg}" if what else msg)
