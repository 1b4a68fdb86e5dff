"""The pixel-major pair compositor (default; skips voxels whose corner SDFs cannot reach the
minimum alpha) against the one-entry-at-a-time compositor (no skipping): forward outputs must be
bit-identical (same per-pixel arithmetic and order, and a skipped voxel contributes to no pixel);
backward gradients may differ only by the summation order of float atomics.  The SDF sharpness s
is scaled to move many voxels across the skip bound."""
import pytest
import torch

from svraster_amd import _C, _lib
from svraster_amd.synthetic import make_workload
from tests._util import hip_raster, oracle_rgbs, rel_l2, say


def _run(name, variant, s_scale):
    _lib.set_option("compositor_variant", variant)
    try:
        w = make_workload(name, device="cuda")
        sc, cam = w.scene, w.cameras[0]
        rgbs = oracle_rgbs(sc, cam).float().cuda()
        geos = sc.geo_grid_pts.reshape(-1)[sc.vox_key].float().contiguous()
        s = torch.tensor([w.s_val * s_scale], dtype=torch.float32, device="cuda")
        bg = torch.ones(3, device="cuda")
        out = hip_raster(sc, cam, geos, rgbs, s, bg, track_max_w=True)
        H, W = cam.image_height, cam.image_width
        g = torch.Generator().manual_seed(2)
        dC = (torch.rand(3, H, W, generator=g) - 0.5).cuda()
        dD = torch.zeros(4, H, W)
        dD[0] = torch.rand(H, W, generator=g) - 0.5
        dN = (torch.rand(3, H, W, generator=g) - 0.5).cuda()
        dT = (torch.rand(1, H, W, generator=g) - 0.5).cuda()
        gt = torch.rand(3, H, W, generator=g).cuda()
        grads = _C.rasterize_voxels_backward(
            out["K"], 3, 1, W, H, cam.tanfovx, cam.tanfovy, cam.cx, cam.cy, cam.w2c, cam.c2w, bg,
            sc.octpath, sc.vox_center, sc.vox_size, geos, rgbs, out["geom"], out["binning"], out["img"],
            out["T"], dC, dD.cuda(), dN, dT, 0.01, gt, 0.0, 0.0, True, True, out["depth"], out["normal"],
            False, s, torch.empty(0))
        torch.cuda.synchronize()
        return out, grads
    finally:
        _lib.set_option("compositor_variant", 0)


@pytest.mark.gpu
@pytest.mark.parametrize("name,s_scale", [("small_128", 1.0), ("small_128", 0.25), ("small_128", 4.0),
                                          ("fd_256", 1.0), ("fd_256", 0.1), ("fd_256", 10.0),
                                          ("single_1237", 1.0)])
def test_pair_compositor_matches_uniform_loop(name, s_scale):
    say(f"[variants] {name} s x {s_scale}")
    a, ga = _run(name, 0, s_scale)
    b, gb = _run(name, 1, s_scale)
    for k in ("color", "T", "depth", "normal", "max_w"):
        assert torch.equal(a[k], b[k]), k
    _, _, na = _C.unpack_ImageState(a["color"].shape[2], a["color"].shape[1], a["img"])
    _, _, nb = _C.unpack_ImageState(b["color"].shape[2], b["color"].shape[1], b["img"])
    assert torch.equal(na, nb)
    errs = [rel_l2(x, y) for x, y in zip(ga, gb)]
    say(f"  forward bit-identical; backward rel-L2 geo {errs[0]:.1e} rgb {errs[1]:.1e} "
        f"subdiv {errs[2]:.1e} ds {errs[3]:.1e}")
    assert max(errs[:3]) <= 1e-5
    assert errs[3] <= 1e-3
