"""Integer stage on the GPU vs the FP32 oracle: n_dup, packed bboxes, octant bitsets, inclusive
scan, sorted values, reconstructed full keys and tile ranges -- all bit-exact."""
import numpy as np
import pytest
import torch

from svraster_amd import _C
from svraster_amd.synthetic import make_workload, sphere_camera, look_at_camera
from tests._util import full_keys_from_vals, hip_raster, oracle_binning, say


def _check_binning(scene, cam):
    bn = oracle_binning(scene, cam)
    geos = scene.geo_grid_pts.reshape(-1)[scene.vox_key].float().contiguous()
    rgbs = torch.full((scene.num_voxels, 3), 0.5, device="cuda")
    s = torch.tensor([10.0], device="cuda")
    bg = torch.ones(3, device="cuda")
    out = hip_raster(scene, cam, geos, rgbs, s, bg, need_depth=False, need_normal=False)
    P = scene.num_voxels
    n_dup, scan, bbox, quad = _C.unpack_GeomState(P, out["geom"])
    np.testing.assert_array_equal(out["n_dup"].cpu().numpy(), bn.pre.n_dup)
    np.testing.assert_array_equal(n_dup.cpu().numpy().astype(np.int64), bn.pre.n_dup.astype(np.int64))
    np.testing.assert_array_equal(scan.cpu().numpy().view(np.uint32), bn.scan)
    vis = bn.pre.n_dup > 0
    np.testing.assert_array_equal(bbox.cpu().numpy().view(np.uint32)[vis], bn.pre.bbox[vis])
    np.testing.assert_array_equal(quad.cpu().numpy()[vis], bn.pre.quad_bits[vis].astype(np.uint8))
    assert out["K"] == bn.K
    keys, vals = _C.unpack_BinningState(out["K"], out["binning"])
    vals = vals.cpu().numpy().view(np.uint32)
    np.testing.assert_array_equal(vals, bn.vals)
    ranges, tile_last, n_contrib = _C.unpack_ImageState(cam.image_width, cam.image_height, out["img"])
    rg = ranges.cpu().numpy().reshape(-1, 2).view(np.uint32)
    np.testing.assert_array_equal(rg, bn.ranges)
    full = full_keys_from_vals(vals, scene.octpath.cpu().numpy(), rg)
    np.testing.assert_array_equal(full, bn.keys)
    # Invariants: the sorted keys are non-decreasing and the ranges partition [0, K).
    assert np.all(np.diff(full.astype(np.float64)) >= 0) or np.all(full[1:] >= full[:-1])
    nz = rg[:, 1] > rg[:, 0]
    assert int((rg[nz, 1] - rg[nz, 0]).sum()) == bn.K
    return bn


@pytest.mark.gpu
@pytest.mark.parametrize("name", ["small_128", "fd_256"])
def test_binning_bit_exact_small(name):
    say(f"[binning] {name}")
    w = make_workload(name, device="cuda")
    _check_binning(w.scene, w.cameras[0])


@pytest.mark.gpu
def test_binning_bit_exact_single_1237_full_size():
    say("[binning] single_1237 (2M voxels, 1237x822)")
    w = make_workload("single_1237", device="cuda")
    bn = _check_binning(w.scene, w.cameras[0])
    say(f"  K={bn.K}")


@pytest.mark.gpu
@pytest.mark.parametrize("W,H", [(100, 37), (17, 300), (1, 1)])
def test_binning_ragged_image_sizes(W, H):
    say(f"[binning] ragged {W}x{H}")
    w = make_workload("small_128", device="cuda")
    cam = look_at_camera(np.array([0.3, -2.2, 1.1]), W, H, 0.9 * W, 0.9 * W, device="cuda")
    _check_binning(w.scene, cam)


@pytest.mark.gpu
def test_binning_camera_inside_scene_near_plane():
    say("[binning] camera inside the voxel cloud (near-plane clipping)")
    w = make_workload("small_128", device="cuda")
    cam = look_at_camera(np.array([0.05, 0.02, -0.1]), 96, 80, 70.0, 70.0, target=(0.4, 0.3, 0.5),
                         device="cuda")
    _check_binning(w.scene, cam)


@pytest.mark.gpu
def test_binning_all_culled():
    say("[binning] camera looking away (K = 0)")
    w = make_workload("small_128", device="cuda")
    cam = look_at_camera(np.array([0.0, 0.0, 3.0]), 64, 64, 50.0, 50.0, target=(0.0, 0.0, 9.0),
                         device="cuda")
    bn = _check_binning(w.scene, cam)
    assert bn.K == 0
