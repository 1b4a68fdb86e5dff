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


def _pair_order_gpu(octpaths: np.ndarray) -> np.ndarray:
    """Rank order of all (voxel, octant) pairs from svr_batch_pair_order (the 64-bit-key radix sort)."""
    from svraster_amd import _lib
    lib = _lib.load()
    P = octpaths.shape[0]
    op = torch.from_numpy(octpaths.astype(np.int64)).cuda()
    tmp = torch.empty(int(lib.svr_batch_pair_order_bytes(P)), dtype=torch.uint8, device="cuda")
    order = torch.empty(8 * P, dtype=torch.int32, device="cuda")
    stream = torch.cuda.current_stream().cuda_stream
    _lib.check(lib.svr_batch_pair_order(P, op.data_ptr(), tmp.data_ptr(), order.data_ptr(), stream),
               "svr_batch_pair_order")
    torch.cuda.synchronize()
    return order.cpu().numpy().view(np.uint32)


@pytest.mark.gpu
@pytest.mark.parametrize("P,kind", [(1, "random"), (511, "random"), (512, "random"), (513, "random"),
                                    (6000, "random"), (3000, "equal"), (3000, "three")])
def test_pair_rank_sort_stable(P, kind):
    """The radix sort behind the rank order equals a stable argsort of the 48-bit rank keys
    octpath ^ repeat3(octant): one partial sort tile, exactly one tile (4096 pairs), one key past
    it, several tiles; all voxels on one octpath (every digit in runs, ties resolved by pair
    index) and three distinct octpaths."""
    say(f"[binning] pair rank sort, P={P} ({kind})")
    rng = np.random.default_rng(P)
    full = rng.integers(0, 1 << 48, size=P, dtype=np.int64)
    if kind == "equal":
        octpaths = np.full(P, full[0])
    elif kind == "three":
        octpaths = full[:3][rng.integers(0, 3, size=P)]
    else:
        octpaths = full
    from oracle.binning import REP3
    keys = (octpaths.astype(np.uint64)[:, None] ^ REP3[None, :]).reshape(-1) & np.uint64((1 << 48) - 1)
    expect = np.argsort(keys, kind="stable").astype(np.uint32)
    order = _pair_order_gpu(octpaths)
    np.testing.assert_array_equal(order, expect)
    # svr_batch_pair_slots: the rank-order slot of every pair, the inverse permutation
    from svraster_amd import _lib
    lib = _lib.load()
    o = torch.from_numpy(order.view(np.int32)).cuda()
    slots = torch.full((8 * P,), -1, dtype=torch.int32, device="cuda")
    _lib.check(lib.svr_batch_pair_slots(P, o.data_ptr(), slots.data_ptr(), None), "svr_batch_pair_slots")
    torch.cuda.synchronize()
    inv = np.empty(8 * P, dtype=np.uint32)
    inv[expect] = np.arange(8 * P, dtype=np.uint32)
    np.testing.assert_array_equal(slots.cpu().numpy().view(np.uint32), inv)


@pytest.mark.gpu
def test_preprocess_reciprocal_matches_ieee_division():
    """The preprocess takes 1/z of corner depths in [2^-100, 2^100] as the hardware reciprocal plus
    one Newton step; the bboxes are bit-exact only if that equals the IEEE quotient.  Every
    float32 significand of every exponent in that range, both signs, must agree."""
    from svraster_amd import _lib
    lib = _lib.load()
    out = torch.zeros(1, dtype=torch.int64, device="cuda")
    say("[binning] rcp + Newton step vs IEEE 1/x, exponents -100..100")
    _lib.check(lib.svr_selfcheck_rcp_rn(-100, 100, out.data_ptr(), None), "selfcheck_rcp_rn")
    torch.cuda.synchronize()
    assert int(out.item()) == 0


@pytest.mark.gpu
def test_reverse_walk_division_matches_ieee_division():
    """The reverse compositor recovers T / (1 - alpha) with the correctly rounded reciprocal, a
    product and one FMA correction; the backward equals the IEEE-division form bit for bit only if
    that is the IEEE quotient.  2 x 2^28 random operand pairs must agree."""
    from svraster_amd import _lib
    lib = _lib.load()
    out = torch.zeros(1, dtype=torch.int64, device="cuda")
    for seed in (1, 2):
        say(f"[binning] FMA-corrected quotient vs IEEE x / y, 2^28 pairs, seed {seed}")
        _lib.check(lib.svr_selfcheck_div_rn(1 << 20, seed, out.data_ptr(), None), "selfcheck_div_rn")
    torch.cuda.synchronize()
    assert int(out.item()) == 0
