"""The allocation-free mean aggregation (svr_agg_voxel_into_grid_pts_noalloc, scratch in a fixed
device array of the library) against the workspace entry point svr_agg_voxel_into_grid_pts, two
calls in a row, the second smaller.  Values are small integers, so every sum is exact in float32
and the two entry points must agree bit for bit whatever the atomics' order."""
import ctypes

import pytest
import torch

from svraster_amd import _lib
from tests._util import say


def _p(t):
    return ctypes.c_void_p(t.data_ptr()) if t is not None and t.numel() else None


def _case(N, M, C, seed):
    g = torch.Generator().manual_seed(seed)
    key = torch.randint(0, M, (N, 8), generator=g, dtype=torch.int64).cuda()
    val = torch.randint(-8, 9, (N, 8, C), generator=g).float().cuda()
    return key, val


def _with_workspace(lib, key, val, M, C):
    out = torch.empty(M, C, device="cuda")
    ws = torch.empty(M, dtype=torch.int32, device="cuda")
    _lib.check(lib.svr_agg_voxel_into_grid_pts(_p(key), key.shape[0], _p(val), C, M, _p(ws), _p(out), None),
               "svr_agg_voxel_into_grid_pts")
    return out


def _no_alloc(lib, key, val, M, C):
    out = torch.empty(M, C, device="cuda")
    _lib.check(lib.svr_agg_voxel_into_grid_pts_noalloc(_p(key), key.shape[0], _p(val), C, M, _p(out), None),
               "svr_agg_voxel_into_grid_pts_noalloc")
    return out


@pytest.mark.gpu
def test_no_alloc_mean_aggregation_matches_workspace_entry():
    lib = _lib.load()
    for (N, M, C, seed) in ((50_000, 100_000, 3, 1), (300, 1_000, 1, 2)):
        say(f"[no-alloc] mean aggregation N={N} M={M} C={C}")
        key, val = _case(N, M, C, seed)
        ref = _with_workspace(lib, key, val, M, C)
        got = _no_alloc(lib, key, val, M, C)
        torch.cuda.synchronize()
        n_named = int((torch.bincount(key.reshape(-1), minlength=M) > 0).sum())
        say(f"  {n_named} of {M} points named; identical: {torch.equal(got, ref)}")
        assert torch.equal(got, ref)
        assert n_named < M            # some points are named by no voxel: they must read 0
    # the size limit of the fixed scratch array is enforced before any launch
    big = 1 << 22
    rc = lib.svr_agg_voxel_into_grid_pts_noalloc(None, 0, None, 1, big + 1, None, None)
    assert rc != 0 and b"at most" in lib.svr_last_error()
