"""Batched multi-view path vs the per-view drop-in path, and HIP-graph replays vs eager."""
import numpy as np
import pytest
import torch

from svraster_amd.batched import BatchedRasterizer, BatchedTrainStep
from svraster_amd.model import VoxelModel
from svraster_amd.synthetic import look_at_camera, make_workload
from tests._util import rel_l2, say


def _setup(n_views=4):
    """fd_256 scene seen by n_views cameras of mixed sizes from random directions."""
    w = make_workload("fd_256", device="cuda")
    sc = w.scene
    rng = np.random.default_rng(7)
    sizes = [(128, 128), (96, 64), (200, 150), (50, 70), (256, 256), (33, 17)]
    sizes = [sizes[i % len(sizes)] for i in range(n_views)]
    cams = []
    for (W, H) in sizes:
        d = rng.normal(size=3)
        d /= np.linalg.norm(d)
        cams.append(look_at_camera(d * rng.uniform(2.0, 3.0), W, H, 0.8 * W, 0.8 * W, device="cuda"))
    g = torch.Generator().manual_seed(1)
    targets = [torch.rand(3, c.image_height, c.image_width, generator=g).cuda() for c in cams]
    return w, sc, cams, targets


def _per_view_reference(sc, cams, targets):
    m = VoxelModel(sc)
    colors = []
    for cam, tgt in zip(cams, targets):
        out = m.render(cam)
        colors.append(out["color"].detach().clone())
        ((out["color"] - tgt) ** 2).mean().backward()
    return m, colors


@pytest.mark.gpu
def test_batched_matches_per_view():
    say("[batched] 4 views of mixed size vs per-view drop-in path")
    w, sc, cams, targets = _setup()
    m, colors = _per_view_reference(sc, cams, targets)
    r = BatchedRasterizer(sc.octpath, sc.octlevel, sc.vox_center, sc.vox_size, sc.is_leaf, sc.vox_key, cams,
                          m.bg_color, sc.sh_degree)
    step = BatchedTrainStep(r, m._geo_grid_pts.detach(), m._sh0.detach(), m._shs.detach(), m._log_s.detach(),
                            targets)
    step()
    torch.cuda.synchronize()
    for v, c in enumerate(colors):
        err = (r.view_slice(r.color, 3, v) - c).abs().max().item()
        say(f"  view {v}: colour max-abs vs per-view {err:.2e}")
        assert err <= 1e-6
    e_grid = rel_l2(step.grid_grad, m._geo_grid_pts.grad)
    e_sh0 = rel_l2(step.sh0_grad, m._sh0.grad)
    e_shs = rel_l2(step.shs_grad, m._shs.grad)
    ds_ref = m._log_s.grad.item() / (10.0 * torch.exp(m._log_s.detach() * 10.0).item())
    e_s = abs(step.s_grad.item() - ds_ref) / max(abs(ds_ref), 1e-12)
    say(f"  rel-L2 grid {e_grid:.2e} sh0 {e_sh0:.2e} shs {e_shs:.2e} s {e_s:.2e}")
    assert e_grid <= 1e-5 and e_sh0 <= 1e-5 and e_shs <= 1e-5 and e_s <= 1e-4


@pytest.mark.gpu
def test_fused_mse_gradient():
    """The compositor's fused dL/dcolour equals (colour - target) * 2 / (3 W H) per view, computed in
    FP64 from its own colour output (a few FP32 roundings: 1e-6 relative), and the step's gradients
    equal the unfused step's (MseColorGrad after the forward) to FP32 rounding."""
    say("[batched] fused MSE colour gradient vs FP64 and vs the separate loss pass")
    w, sc, cams, targets = _setup()
    m = VoxelModel(sc)
    grads = []
    for fuse in (True, False):
        r = BatchedRasterizer(sc.octpath, sc.octlevel, sc.vox_center, sc.vox_size, sc.is_leaf, sc.vox_key, cams,
                              m.bg_color, sc.sh_degree)
        step = BatchedTrainStep(r, m._geo_grid_pts.detach(), m._sh0.detach(), m._shs.detach(), m._log_s.detach(),
                                targets, fuse_loss=fuse)
        step()
        torch.cuda.synchronize()
        if fuse:
            for v, tgt in enumerate(targets):
                c = r.view_slice(r.color, 3, v).double()
                want = (c - tgt.double()) * (2.0 / (3 * r.pix[v]))
                got = r.view_slice(step.dC, 3, v).double()
                err = ((got - want).abs() / want.abs().clamp_min(1e-30)).max().item()
                say(f"  view {v}: fused dL/dcolour max relative error {err:.2e}")
                assert err <= 1e-6
        grads.append((step.grid_grad.clone(), step.sh0_grad.clone(), step.s_grad.clone()))
    e = [rel_l2(a, b) for a, b in zip(grads[0], grads[1])]
    say(f"  fused vs unfused step gradients rel-L2 {e}")
    assert max(e) <= 1e-5


@pytest.mark.gpu
def test_batched_graph_replays_match_eager():
    say("[batched] graph capture: replays vs eager")
    w, sc, cams, targets = _setup()
    m = VoxelModel(sc)
    r = BatchedRasterizer(sc.octpath, sc.octlevel, sc.vox_center, sc.vox_size, sc.is_leaf, sc.vox_key, cams,
                          m.bg_color, sc.sh_degree)
    step = BatchedTrainStep(r, m._geo_grid_pts.detach(), m._sh0.detach(), m._shs.detach(), m._log_s.detach(),
                            targets)
    step()
    torch.cuda.synchronize()
    ref_color = r.color.clone()
    ref_grid = step.grid_grad.clone()
    ref_sh0 = step.sh0_grad.clone()
    step.capture()
    for k in range(3):
        step()
        torch.cuda.synchronize()
        junk = torch.randn(1000, device="cuda").sum()       # eager work between replays
        e_c = (r.color - ref_color).abs().max().item()
        e_g = rel_l2(step.grid_grad, ref_grid)
        e_s = rel_l2(step.sh0_grad, ref_sh0)
        say(f"  replay {k}: colour {e_c:.2e} grid {e_g:.2e} sh0 {e_s:.2e} ({junk.item():.1f})")
        assert e_c == 0.0 and e_g <= 1e-5 and e_s <= 1e-5


@pytest.mark.gpu
def test_batched_voxel_skip_keeps_results():
    say("[batched] leaving non-contributing voxels out of binning: same images, fewer keys")
    w, sc, cams, targets = _setup()
    m = VoxelModel(sc)
    res = {}
    for cull in (False, True):
        r = BatchedRasterizer(sc.octpath, sc.octlevel, sc.vox_center, sc.vox_size, sc.is_leaf, sc.vox_key,
                              cams, m.bg_color, sc.sh_degree, cull=cull)
        step = BatchedTrainStep(r, m._geo_grid_pts.detach(), m._sh0.detach(), m._shs.detach(),
                                m._log_s.detach(), targets)
        step()
        torch.cuda.synchronize()
        res[cull] = (r.color.clone(), r.T.clone(), step.grid_grad.clone(), step.sh0_grad.clone(),
                     step.shs_grad.clone(), step.s_grad.clone(), r.key_count())
    a, b = res[False], res[True]
    say(f"  keys {a[6]} -> {b[6]}")
    assert torch.equal(a[0], b[0]) and torch.equal(a[1], b[1])
    e = [rel_l2(x, y) for x, y in zip(b[2:5], a[2:5])]
    e_s = abs(b[5].item() - a[5].item()) / max(abs(a[5].item()), 1e-12)
    say(f"  rel-L2 grid {e[0]:.2e} sh0 {e[1]:.2e} shs {e[2]:.2e} s {e_s:.2e}")
    assert max(e) <= 1e-5 and e_s <= 1e-4
    assert b[6] < a[6]


@pytest.mark.gpu
@pytest.mark.parametrize("n_views", [4, 40])
def test_batched_tile_lists_match_per_view(n_views):
    """Rank-order binning of all views at once, over each (view, voxel)'s candidate rectangle, gives
    every tile the per-view (reference) list less the entries whose candidate rectangle (the
    oracle's) misses the tile, in the same order, and the same per-pixel last contributors.  With
    40 views each (voxel, octant) pair is seen by about 5 of them, so a workgroup's 256 pairs own
    more (pair, view) rectangles than the key emission's LDS list holds at once (1024) and it runs
    in several batches."""
    from oracle import binning as ob
    from svraster_amd import _C
    say(f"[batched] per-tile lists and contributor counts vs the per-view binning, {n_views} views")
    w, sc, cams, targets = _setup(n_views)
    m = VoxelModel(sc)
    r = BatchedRasterizer(sc.octpath, sc.octlevel, sc.vox_center, sc.vox_size, sc.is_leaf, sc.vox_key, cams,
                          m.bg_color, sc.sh_degree, cull=False)
    s_val = torch.exp(m._log_s.detach().reshape(1) * 10.0)
    r.forward(m._geo_grid_pts.detach(), m._sh0.detach(), m._shs.detach(), s_val)
    torch.cuda.synchronize()
    ranges_b, vals_b, ncon_b, _ = r.state_views()
    tile_base = 0
    dropped = 0
    for v, cam in enumerate(cams):
        W, H = cam.image_width, cam.image_height
        n_dup, geom = _C.rasterize_preprocess(W, H, cam.tanfovx, cam.tanfovy, cam.cx, cam.cy, cam.w2c, cam.c2w, 0,
                                              cam.near, sc.octpath, sc.vox_center, sc.vox_size, sc.is_leaf)
        rgbs = r.rgbs_views[v].contiguous()
        K, binning, img, *_ = _C.rasterize_voxels(
            3, 1, W, H, cam.tanfovx, cam.tanfovy, cam.cx, cam.cy, cam.w2c, cam.c2w, m.bg_color, 0, False, False,
            False, False, sc.octpath, sc.vox_center, sc.vox_size, r.geos, rgbs,
            torch.empty(sc.num_voxels, 0, device="cuda"), geom, False, s_val)
        _, vals_v = _C.unpack_BinningState(K, binning)
        rng_v, _, ncon_v = _C.unpack_ImageState(W, H, img)
        tx_n, ty_n = rng_v.shape[1], rng_v.shape[0]
        nt = tx_n * ty_n
        pre = ob.preprocess(W, H, cam.tanfovx, cam.tanfovy, cam.cx, cam.cy, cam.w2c.cpu().numpy(),
                            cam.c2w.cpu().numpy(), cam.near, sc.vox_center.cpu().numpy(), sc.vox_size.cpu().numpy(),
                            sc.is_leaf.cpu().numpy())
        tb = pre.tbox.astype(np.int64)
        x0, y0, x1, y1 = tb[:, 0] >> 16, tb[:, 0] & 0xFFFF, tb[:, 1] >> 16, tb[:, 1] & 0xFFFF
        cx0, cx1 = np.clip(x0 >> 4, 0, tx_n - 1), np.clip(x1 >> 4, 0, tx_n - 1)
        cy0, cy1 = np.clip(y0 >> 4, 0, ty_n - 1), np.clip(y1 >> 4, 0, ty_n - 1)
        rv = rng_v.reshape(nt, 2).long().cpu().numpy()
        rb = ranges_b[tile_base:tile_base + nt].long().cpu().numpy()
        vv = vals_v.cpu().numpy().view(np.uint32)
        vb = vals_b.cpu().numpy().view(np.uint32)
        nc_v = ncon_v.cpu().numpy().astype(np.int64)
        nc_b = ncon_b[int(r.pix_base[v]):int(r.pix_base[v]) + W * H].view(H, W).cpu().numpy().astype(np.int64)
        want_nc = np.zeros_like(nc_v)
        for t in range(nt):
            tx, ty = t % tx_n, t // tx_n
            lst = vv[rv[t, 0]:rv[t, 1]]
            vox = (lst & ((1 << 29) - 1)).astype(np.int64)
            keep = ((x0[vox] <= x1[vox]) & (y0[vox] <= y1[vox]) & (cx0[vox] <= tx) & (cx1[vox] >= tx)
                    & (cy0[vox] <= ty) & (cy1[vox] >= ty))
            got = vb[rb[t, 0]:rb[t, 1]]
            assert np.array_equal(got, lst[keep]), f"view {v} tile {t}: list contents/order"
            dropped += int((~keep).sum())
            # last contributor: the same entry, at its position in the shorter list
            kept_before = np.concatenate([[0], np.cumsum(keep)])
            blk = nc_v[16 * ty:16 * ty + 16, 16 * tx:16 * tx + 16]
            want_nc[16 * ty:16 * ty + 16, 16 * tx:16 * tx + 16] = kept_before[blk]
        assert np.array_equal(nc_b, want_nc), f"view {v}: contributor counts"
        say(f"  view {v} {W}x{H}: {int(K)} reference keys, lists and contributor counts as expected")
        tile_base += nt
    say(f"  {dropped} reference entries left out over all views")
    assert dropped > 0


@pytest.mark.gpu
def test_fused_gather_transpose_matches_separate_pass():
    say("[batched] corner-SDF gradients flushed straight to the grid points vs dL/dgeos + gather transpose")
    w, sc, cams, targets = _setup()
    out = []
    for fuse in (False, True):
        m = VoxelModel(sc)
        r = BatchedRasterizer(sc.octpath, sc.octlevel, sc.vox_center, sc.vox_size, sc.is_leaf, sc.vox_key, cams,
                              m.bg_color, sc.sh_degree, fuse_gather_bw=fuse)
        step = BatchedTrainStep(r, m._geo_grid_pts.detach(), m._sh0.detach(), m._shs.detach(),
                                m._log_s.detach(), targets)
        step()
        torch.cuda.synchronize()
        out.append((step.grid_grad.clone(), r.dvox[:, :8].abs().sum().item()))
    e = rel_l2(out[1][0], out[0][0])
    say(f"  grid gradient rel-L2 {e:.2e}; |dL/dgeos| left in the voxel buffer: separate {out[0][1]:.3e}, "
        f"fused {out[1][1]:.3e}")
    assert e <= 1e-6 and out[0][1] > 0 and out[1][1] == 0.0


@pytest.mark.gpu
def test_batched_key_capacity_is_enforced():
    say("[batched] key capacity: below the un-culled count is refused; an eager forward needing more raises")
    w, sc, cams, targets = _setup()
    m = VoxelModel(sc)
    r = BatchedRasterizer(sc.octpath, sc.octlevel, sc.vox_center, sc.vox_size, sc.is_leaf, sc.vox_key, cams,
                          m.bg_color, sc.sh_degree)
    with pytest.raises(ValueError):
        BatchedRasterizer(sc.octpath, sc.octlevel, sc.vox_center, sc.vox_size, sc.is_leaf, sc.vox_key, cams,
                          m.bg_color, sc.sh_degree, key_capacity=r.keys_unculled - 1)
    s_val = torch.exp(m._log_s.detach().reshape(1) * 10.0)
    r.forward(m._geo_grid_pts.detach(), m._sh0.detach(), m._shs.detach(), s_val)
    peak = r.check_capacity()
    say(f"  un-culled {r.keys_unculled} keys, capacity {r.cap}, peak {peak}")
    assert 0 < peak <= r.keys_unculled <= r.cap
    # the same voxel set with every voxel twice as large needs more keys than this capacity:
    # the eager forward stops right after the key count, before any key is emitted
    big = BatchedRasterizer(sc.octpath, sc.octlevel, sc.vox_center, sc.vox_size.clone(), sc.is_leaf, sc.vox_key,
                            cams, m.bg_color, sc.sh_degree, key_capacity=r.keys_unculled + 4096, cull=False)
    big.vox_size.mul_(2.0)
    with pytest.raises(RuntimeError, match="capacity"):
        big.forward(m._geo_grid_pts.detach(), m._sh0.detach(), m._shs.detach(), s_val)


@pytest.mark.gpu
def test_voxel_records_gather_corners_and_activity():
    """svr_batch_voxel_records: per voxel the centre, size, the corner SDFs of GatherGeoParams
    (grid_pts[vox_key]) and 1/size, bit for bit; its binning mask equals svr_voxel_activity's on the
    separately gathered corners (cull) and is_leaf itself (no cull)."""
    import ctypes
    from svraster_amd import _lib
    say("[batched] voxel records vs the corner gather and the activity test")
    w, sc, cams, targets = _setup()
    lib = _lib.load()
    P = sc.num_voxels
    m = VoxelModel(sc)
    grid = m._geo_grid_pts.detach().reshape(-1).contiguous()
    leaf = sc.is_leaf.reshape(-1).to(torch.uint8).contiguous()
    s_val = torch.tensor([float(w.s_val)], device="cuda")
    rec = torch.empty(P, 16, device="cuda")
    cam_scale = 3.0
    for cull in (1, 0):
        bl = torch.empty(P, dtype=torch.uint8, device="cuda")
        _lib.check(lib.svr_batch_voxel_records(P, sc.vox_key.data_ptr(), grid.data_ptr(), sc.vox_center.data_ptr(),
                                               sc.vox_size.data_ptr(), leaf.data_ptr(), s_val.data_ptr(),
                                               ctypes.c_float(cam_scale), cull, rec.data_ptr(), bl.data_ptr(), None),
                   "svr_batch_voxel_records")
        geos = grid[sc.vox_key]
        size = sc.vox_size.reshape(-1)
        assert torch.equal(rec[:, 0:3], sc.vox_center) and torch.equal(rec[:, 3], size)
        assert torch.equal(rec[:, 4:12], geos)
        assert torch.equal(rec[:, 12].cpu(), 1.0 / size.cpu()) and not rec[:, 13:16].any()
        if cull:
            act = torch.empty(P, dtype=torch.uint8, device="cuda")
            _lib.check(lib.svr_voxel_activity(P, leaf.data_ptr(), geos.contiguous().data_ptr(), s_val.data_ptr(),
                                              sc.vox_center.data_ptr(), sc.vox_size.data_ptr(),
                                              ctypes.c_float(cam_scale), act.data_ptr(), None), "svr_voxel_activity")
            say(f"  cull: {int(bl.sum())} of {int(leaf.sum())} leaves enter binning")
            assert torch.equal(bl, act) and 0 < int(bl.sum()) <= int(leaf.sum())
        else:
            assert torch.equal(bl, leaf)


@pytest.mark.gpu
def test_rank_order_count_stores_match_gather():
    """The preprocess stores each (voxel, octant) key count at its rank-order slot
    (svr_batch_pair_slots); passing no slots gathers the counts through the rank order in a pass of
    its own instead.  Both give the same key lists, contributor counts and images, bit for bit."""
    say("[batched] key counts stored at rank-order slots vs gathered through the order, 40 views")
    w, sc, cams, targets = _setup(40)
    m = VoxelModel(sc)
    r = BatchedRasterizer(sc.octpath, sc.octlevel, sc.vox_center, sc.vox_size, sc.is_leaf, sc.vox_key, cams,
                          m.bg_color, sc.sh_degree)
    s_val = torch.exp(m._log_s.detach().reshape(1) * 10.0)
    args = (m._geo_grid_pts.detach(), m._sh0.detach(), m._shs.detach(), s_val)
    r.forward(*args)
    torch.cuda.synchronize()
    want = [t.clone() for t in r.state_views()] + [r.color.clone()]
    slots, r.pair_slots = r.pair_slots, None
    try:
        r.forward(*args)
        torch.cuda.synchronize()
        got = list(r.state_views()) + [r.color]
    finally:
        r.pair_slots = slots
    for name, a, b in zip(("ranges", "vals", "n_contrib", "tile_last", "color"), got, want):
        assert torch.equal(a, b), name
