"""Topology changes on the GPU: the voxel <-> grid-point link, the mean of corner values into the
grid points, subdivision and pruning, against the numpy restatement (oracle/octree.py)."""
import numpy as np
import pytest
import torch

from oracle import octree as oot
from svraster_amd import adaptive
from svraster_amd.model import VoxelModel
from svraster_amd.synthetic import build_scene, make_workload
from tests._util import say


def _np(t):
    return t.detach().cpu().numpy()


@pytest.mark.gpu
def test_grid_pts_link_matches_scene_and_oracle():
    say("[adaptive] voxel <-> grid-point link")
    sc = make_workload("fd_256", device="cuda").scene
    key, vk = adaptive.build_grid_pts_link(sc.octpath, sc.octlevel)
    assert torch.equal(vk, sc.vox_key) and torch.equal(key, sc.grid_pts_key)
    sc2 = build_scene(3, 3, 10, 1, seed=3, device="cuda", target_voxels=30_000, keep_parents_from=9)
    key2, vk2 = adaptive.build_grid_pts_link(sc2.octpath, sc2.octlevel)
    okey, ovk = oot.build_grid_pts_link(_np(sc2.octpath), _np(sc2.octlevel))
    say(f"  {sc.num_voxels} voxels -> {key.shape[0]} points; with parents: {sc2.num_voxels} voxels "
        f"({int((sc2.is_leaf == 0).sum())} non-leaf) -> {key2.shape[0]} points")
    assert np.array_equal(_np(key2), okey) and np.array_equal(_np(vk2), ovk)


@pytest.mark.gpu
def test_agg_mean_matches_oracle():
    say("[adaptive] mean of per-voxel corner values into the grid points")
    sc = make_workload("fd_256", device="cuda").scene
    M = sc.grid_pts_key.shape[0]
    g = torch.Generator().manual_seed(9)
    for shape in ((sc.num_voxels, 8), (sc.num_voxels, 8, 3)):
        val = torch.randn(*shape, generator=g).cuda()
        got = adaptive.agg_voxel_into_grid_pts(M + 5, sc.vox_key, val)
        ref = oot.agg_mean(M + 5, _np(sc.vox_key), _np(val))
        err = float(np.abs(_np(got) - ref).max())
        say(f"  shape {tuple(shape)}: max-abs {err:.2e}")
        assert err <= 1e-5 and np.all(_np(got)[M:] == 0.0)


def _model_dict(m):
    return {"octpath": _np(m.octpath).reshape(-1), "octlevel": _np(m.octlevel).reshape(-1),
            "is_leaf": _np(m.is_leaf).reshape(-1), "sh0": _np(m._sh0), "shs": _np(m._shs),
            "grid_pts": _np(m._geo_grid_pts).reshape(-1), "vox_key": _np(m.vox_key)}


def _check_against(m, ref, sc):
    assert np.array_equal(_np(m.octpath).reshape(-1), ref["octpath"])
    assert np.array_equal(_np(m.octlevel).reshape(-1).astype(np.int64), ref["octlevel"].astype(np.int64))
    assert np.array_equal(_np(m.is_leaf).reshape(-1), ref["is_leaf"])
    assert np.array_equal(_np(m.vox_key), ref["vox_key"]) and np.array_equal(_np(m.grid_pts_key), ref["grid_pts_key"])
    assert np.array_equal(_np(m._sh0), ref["sh0"]) and np.array_equal(_np(m._shs), ref["shs"])
    err = float(np.abs(_np(m._geo_grid_pts).reshape(-1) - ref["grid_pts"]).max())
    # centres and sizes from the paths, closed form
    ijk = oot.ijk_array(ref["octpath"], ref["octlevel"])
    size = sc.scene_extent * 2.0 ** (-ref["octlevel"].astype(np.float64))
    ctr = -0.5 * sc.scene_extent + (ijk + 0.5) * size[:, None]
    e_c = float(np.abs(_np(m.vox_center) - ctr).max())
    assert np.array_equal(_np(m.vox_size).reshape(-1), size.astype(np.float32)) and e_c <= 1e-6
    return err


@pytest.mark.gpu
def test_subdividing_matches_oracle_and_keeps_the_field():
    say("[adaptive] subdivision (10% of the voxels) vs the numpy restatement")
    sc = build_scene(3, 3, 10, 1, seed=3, device="cuda", target_voxels=30_000, keep_parents_from=9)
    m = VoxelModel(sc)
    m.make_optimizer()
    before = _model_dict(m)
    rng = np.random.default_rng(5)
    mask_np = (rng.uniform(size=sc.num_voxels) < 0.1) & (before["octlevel"] < 16) & (before["is_leaf"] != 0)
    say(f"  subdividing {int(mask_np.sum())} leaves, {int((mask_np & (before['octlevel'] >= 9)).sum())} of "
        f"them at level >= 9 (kept as parents)")
    old_center, old_size = m.vox_center.clone(), m.vox_size.clone()
    adaptive.subdividing(m, torch.from_numpy(mask_np).cuda())
    ref = oot.subdivide(before, mask_np)
    err = _check_against(m, ref, sc)
    say(f"  {sc.num_voxels} -> {m.num_voxels} voxels; grid points max-abs vs oracle {err:.2e}")
    assert err <= 1e-6
    assert m.optimizer.param_groups[0]["params"][0] is m._geo_grid_pts
    # inside each child the trilinear SDF equals the parent's (trilinear fields refine exactly)
    sub = np.nonzero(mask_np)[0]
    n_kept = m.num_voxels - 8 * sub.size
    child = n_kept + np.arange(8 * sub.size)
    parent = np.repeat(sub, 8)
    q = rng.uniform(size=(child.size, 3))
    old_corner = before["grid_pts"][before["vox_key"][parent]]
    new_corner = _np(m._geo_grid_pts).reshape(-1)[_np(m.vox_key)[child]]
    cc, cs = _np(m.vox_center)[child], _np(m.vox_size).reshape(-1)[child]
    pc, ps = _np(old_center)[parent], _np(old_size).reshape(-1)[parent]
    x = cc - 0.5 * cs[:, None] + q * cs[:, None]
    qp = (x - (pc - 0.5 * ps[:, None])) / ps[:, None]

    def tri(c, t):
        w = np.stack([(t[:, 0] if i & 4 else 1 - t[:, 0]) * (t[:, 1] if i & 2 else 1 - t[:, 1]) *
                      (t[:, 2] if i & 1 else 1 - t[:, 2]) for i in range(8)], 1)
        return (w * c).sum(1)

    # points shared with neighbouring subdivided parents average equal values, so the field is kept
    # up to float rounding except where a grid point is shared with a differently-valued neighbour
    diff = np.abs(tri(new_corner, q) - tri(old_corner, qp))
    say(f"  SDF inside the children vs the parent: median {np.median(diff):.2e}, max {diff.max():.2e}")
    assert np.median(diff) <= 1e-6


@pytest.mark.gpu
def test_pruning_matches_oracle():
    say("[adaptive] pruning (20% of the voxels) vs the numpy restatement")
    sc = make_workload("fd_256", device="cuda").scene
    m = VoxelModel(sc)
    before = _model_dict(m)
    rng = np.random.default_rng(6)
    mask_np = rng.uniform(size=sc.num_voxels) < 0.2
    adaptive.pruning(m, torch.from_numpy(mask_np).cuda())
    ref = oot.prune(before, mask_np)
    err = _check_against(m, ref, sc)
    say(f"  {sc.num_voxels} -> {m.num_voxels} voxels; grid points max-abs vs oracle {err:.2e}")
    assert err <= 1e-6
    # the pruned model still renders
    cam = make_workload("fd_256", device="cuda").cameras[0]
    out = m.render(cam)
    assert torch.isfinite(out["color"]).all()
