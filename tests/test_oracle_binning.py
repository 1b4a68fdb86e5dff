"""Pins of the FP32 integer-stage oracle (CPU only)."""
import numpy as np
import torch

from oracle import binning as ob
from oracle import raster as orc
from svraster_amd.synthetic import look_at_camera, make_workload


def _identity_camera(W=100, H=100, tanfov=0.5):
    c2w = np.eye(4, dtype=np.float32)
    return dict(W=W, H=H, tanfovx=tanfov, tanfovy=tanfov, cx=W / 2, cy=H / 2, w2c=c2w, c2w=c2w, near=0.02)


def test_key_codec_closed_form():
    assert ob.rep3(0) == 0
    assert ob.rep3(7) == (1 << 48) - 1
    assert ob.rep3(5) == int("101" * 16, 2)
    # rank = octpath ^ repeat3(q): for q = 0 the rank is the octpath itself.
    pre = ob.Preprocessed(np.array([1], np.int32), np.array([[0, 0]], np.uint32),
                          np.array([1 << 3], np.uint32), 1, 1)
    keys, vals = ob.duplicate_with_keys(pre, np.array([0b101 << 45]))
    assert keys[0] == (0b101 << 45) ^ ob.rep3(3)
    assert vals[0] == (3 << 29) | 0


def test_single_voxel_projection_by_hand():
    c = _identity_camera()
    pre = ob.preprocess(c["W"], c["H"], c["tanfovx"], c["tanfovy"], c["cx"], c["cy"], c["w2c"], c["c2w"],
                        c["near"], np.array([[0.0, 0.0, 5.0]]), np.array([1.0]), np.array([1]))
    # corners x,y = +-0.5 at depth 4.5: focal 100 -> 100 * (+-0.5/4.5) + 49.5 = 38.39 .. 60.61
    assert pre.bbox[0, 0] == (38 << 16) | 38
    assert pre.bbox[0, 1] == (61 << 16) | 61
    # x<0 and y<0 halves of the cube, all in front (z > 0): octants 0, 2, 4, 6
    assert pre.quad_bits[0] == (1 << 0) | (1 << 2) | (1 << 4) | (1 << 6)
    # tiles 2..3 x 2..3, four octants
    assert pre.n_dup[0] == 16


def test_near_sphere_and_behind_camera_are_culled():
    c = _identity_camera()
    pre = ob.preprocess(c["W"], c["H"], c["tanfovx"], c["tanfovy"], c["cx"], c["cy"], c["w2c"], c["c2w"],
                        c["near"], np.array([[0.0, 0.0, 0.01], [0.0, 0.0, -3.0], [0.0, 0.0, 3.0]]),
                        np.array([0.001, 1.0, 1.0]), np.array([1, 1, 0]))
    assert list(pre.n_dup) == [0, 0, 0]      # inside near sphere, behind camera, not a leaf


def test_sort_is_stable_and_ranges_partition():
    w = make_workload("small_128")
    bn = ob.binning_for_camera(w.cameras[0], w.scene)
    assert bn.K == int(bn.pre.n_dup.sum()) > 0
    assert np.all(bn.keys[1:] >= bn.keys[:-1])
    nz = bn.ranges[:, 1] > bn.ranges[:, 0]
    assert int((bn.ranges[nz, 1] - bn.ranges[nz, 0]).sum()) == bn.K
    starts = np.sort(bn.ranges[nz, 0])
    ends = np.sort(bn.ranges[nz, 1])
    assert starts[0] == 0 and ends[-1] == bn.K and np.all(starts[1:] == ends[:-1])
    # ties (equal keys) keep duplication order: values ascend within runs of equal keys
    eq = bn.keys[1:] == bn.keys[:-1]
    v = bn.vals.astype(np.int64)
    assert np.all((v[1:] & ((1 << 29) - 1))[eq] >= (v[:-1] & ((1 << 29) - 1))[eq])


def test_brute_force_every_hit_is_listed():
    """Every pixel whose (FP64) ray meets a voxel's box with a matching octant, for voxels fully
    in front of the camera, must find that voxel (with the pixel's octant) in its tile list."""
    w = make_workload("small_128")
    sc, cam = w.scene, w.cameras[0]
    bn = ob.binning_for_camera(cam, sc)
    rc = orc.camera_from(cam)
    rng = np.random.default_rng(0)
    px = torch.from_numpy(rng.integers(0, cam.image_width, 300))
    py = torch.from_numpy(rng.integers(0, cam.image_height, 300))
    ro, rd, _ = orc.pixel_rays(rc, px, py)
    pq = orc.ray_quadrant(rd)
    c = sc.vox_center.double()
    l = sc.vox_size.double().reshape(-1)
    checked = 0
    for p in range(px.numel()):
        a, b = orc.ray_aabb(c, l, ro, (1.0 / rd[p])[None])
        # corner octants relative to the camera contain the ray octant when the box is hit
        hit = (a < b - 1e-6) & (a > 0.05)
        for v in torch.nonzero(hit).reshape(-1).tolist():
            tile = (int(py[p]) // 16) * ((cam.image_width + 15) // 16) + int(px[p]) // 16
            lo, hi = bn.ranges[tile]
            vals = bn.vals[lo:hi]
            want = (int(pq[p]) << 29) | v
            assert want in set(vals.tolist()), (p, v)
            checked += 1
    assert checked > 100


def test_ragged_and_tiny_images():
    w = make_workload("small_128")
    for (W, H) in [(100, 37), (1, 1), (17, 300)]:
        cam = look_at_camera(np.array([0.3, -2.2, 1.1]), W, H, 0.9 * W, 0.9 * W)
        bn = ob.binning_for_camera(cam, w.scene)
        tiles_x, tiles_y = (W + 15) // 16, (H + 15) // 16
        assert bn.ranges.shape[0] == tiles_x * tiles_y
        tiles = bn.keys >> np.uint64(48)
        assert tiles.size == 0 or int(tiles.max()) < tiles_x * tiles_y


def _tbox_ring_misses(sc, cam):
    """Every pixel of a voxel's bbox outside its candidate rectangle (tbox): the FP64 ray through
    its centre misses the voxel's box.  Returns the number of such pixels checked."""
    bn = ob.binning_for_camera(cam, sc)
    pre = bn.pre
    bb = pre.bbox.astype(np.int64)
    tb = pre.tbox.astype(np.int64)
    bx0, by0, bx1, by1 = bb[:, 0] >> 16, bb[:, 0] & 0xFFFF, bb[:, 1] >> 16, bb[:, 1] & 0xFFFF
    tx0, ty0, tx1, ty1 = tb[:, 0] >> 16, tb[:, 0] & 0xFFFF, tb[:, 1] >> 16, tb[:, 1] & 0xFFFF
    alive = pre.n_dup > 0
    # inside bbox (or empty)
    assert np.all(~alive | (tx0 >= bx0) & (ty0 >= by0) & (tx1 <= bx1) & (ty1 <= by1))
    rc = orc.camera_from(cam)
    W, H = cam.image_width, cam.image_height
    c = sc.vox_center.double()
    l = sc.vox_size.double().reshape(-1)
    pv, px, py = [], [], []
    for v in np.nonzero(alive)[0]:
        for y in range(by0[v], min(by1[v], H - 1) + 1):
            for x in range(bx0[v], min(bx1[v], W - 1) + 1):
                if not (tx0[v] <= x <= tx1[v] and ty0[v] <= y <= ty1[v]):
                    pv.append(v); px.append(x); py.append(y)
    if not pv:
        return 0
    pv = torch.tensor(pv)
    ro, rd, _ = orc.pixel_rays(rc, torch.tensor(px), torch.tensor(py))
    inv = 1.0 / rd
    lo = (c[pv] - 0.5 * l[pv, None] - ro) * inv
    hi = (c[pv] + 0.5 * l[pv, None] - ro) * inv
    a = torch.minimum(lo, hi).max(dim=1).values
    b = torch.maximum(lo, hi).min(dim=1).values
    assert bool((a > b).all()), int((a <= b).sum())
    return len(pv)


def test_candidate_rectangle_holds_every_hit():
    """The candidate rectangle drops only pixels whose ray misses the box: checked pixel by pixel
    (FP64 rays) on every voxel of two views, one with many sub-pixel voxels."""
    w = make_workload("small_128")
    n1 = _tbox_ring_misses(w.scene, w.cameras[0])
    far = look_at_camera(np.array([3.5, -4.0, 2.0]), 97, 61, 60.0, 60.0)
    n2 = _tbox_ring_misses(w.scene, far)
    assert n1 > 1000 and n2 > 1000, (n1, n2)


def test_tight_binning_is_the_reference_list_filtered():
    """Binning over candidate rectangles gives each tile the reference list without the entries
    whose rectangle misses the tile (or is empty), in the reference order."""
    w = make_workload("small_128")
    for cam in (w.cameras[0], look_at_camera(np.array([3.5, -4.0, 2.0]), 97, 61, 60.0, 60.0)):
        _check_tight_lists(w.scene, cam)


def _check_tight_lists(sc, cam):
    bn = ob.binning_for_camera(cam, sc)
    tpre = ob.tight_binning_input(bn.pre)
    keys, vals = ob.duplicate_with_keys(tpre, sc.octpath.numpy())
    skeys, svals = ob.sort_pairs(keys, vals)
    rt = ob.tile_ranges(skeys, tpre.tiles_x * tpre.tiles_y)
    tb = bn.pre.tbox.astype(np.int64)
    x0, y0, x1, y1 = tb[:, 0] >> 16, tb[:, 0] & 0xFFFF, tb[:, 1] >> 16, tb[:, 1] & 0xFFFF
    cx0, cx1 = np.clip(x0 >> 4, 0, tpre.tiles_x - 1), np.clip(x1 >> 4, 0, tpre.tiles_x - 1)
    cy0, cy1 = np.clip(y0 >> 4, 0, tpre.tiles_y - 1), np.clip(y1 >> 4, 0, tpre.tiles_y - 1)
    dropped = 0
    for t in range(tpre.tiles_x * tpre.tiles_y):
        tx, ty = t % tpre.tiles_x, t // tpre.tiles_x
        lst = bn.vals[bn.ranges[t, 0]:bn.ranges[t, 1]]
        vox = (lst & ((1 << 29) - 1)).astype(np.int64)
        keep = ((x0[vox] <= x1[vox]) & (y0[vox] <= y1[vox]) & (cx0[vox] <= tx) & (cx1[vox] >= tx)
                & (cy0[vox] <= ty) & (cy1[vox] >= ty))
        assert np.array_equal(svals[rt[t, 0]:rt[t, 1]], lst[keep]), t
        dropped += int((~keep).sum())
    assert 0 < dropped < bn.K
