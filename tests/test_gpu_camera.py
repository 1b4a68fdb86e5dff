"""Camera utilities on the GPU (compute_rd, depth2pts, is_in_cone, voxel_order_rank) against
independent formulations: the pinhole intrinsics matrix inverted by numpy, projection of the
points back to pixel coordinates, and the closed-form octant rank codec."""
import numpy as np
import pytest
import torch

import svraster_amd as svraster_cuda
from oracle import binning as ob
from svraster_amd.synthetic import look_at_camera, sphere_camera
from tests._util import say


def _pinhole_dirs(cam, W, H, cx, cy):
    """World directions of pixel centres as R K^-1 [u, v, 1]^T (float64), K from the focal lengths
    W / (2 tanfovx), H / (2 tanfovy) and the principal point (cx, cy)."""
    fx = W / (2.0 * float(np.float32(cam.tanfovx)))
    fy = H / (2.0 * float(np.float32(cam.tanfovy)))
    K = np.array([[fx, 0.0, cx], [0.0, fy, cy], [0.0, 0.0, 1.0]])
    u, v = np.meshgrid(np.arange(W) + 0.5, np.arange(H) + 0.5)
    uv1 = np.stack([u, v, np.ones_like(u)], 0).reshape(3, -1)
    d_cam = np.linalg.inv(K) @ uv1
    R = cam.c2w.cpu().double().numpy()[:3, :3]
    return (R @ d_cam).reshape(3, H, W)


class _RefCamera:
    """The attributes src/cameras.py:75-87 reads, with its compute_rd call verbatim in form."""

    def __init__(self, cam):
        self.image_width, self.image_height = cam.image_width, cam.image_height
        self.tanfovx, self.tanfovy = np.float64(cam.tanfovx), np.float64(cam.tanfovy)
        self.cx, self.cy = cam.cx, cam.cy
        self.c2w = cam.c2w
        self.position = cam.c2w[:3, 3]

    def compute_rd(self, wh=None, cxcy=None, device=None):
        if wh is None:
            wh = (self.image_width, self.image_height)
        if cxcy is None:
            cxcy = (self.cx * wh[0] / self.image_width, self.cy * wh[1] / self.image_height)
        rd = svraster_cuda.utils.compute_rd(
            width=wh[0], height=wh[1],
            cx=cxcy[0], cy=cxcy[1],
            tanfovx=self.tanfovx, tanfovy=self.tanfovy,
            c2w_matrix=self.c2w.cuda())
        return rd.to(device if device is None else self.c2w.device)

    def depth2pts(self, depth):
        h, w = depth.shape[-2:]
        rd = self.compute_rd(wh=(w, h), device=depth.device)
        return self.position.view(3, 1, 1).to(depth.device) + rd * depth


@pytest.mark.gpu
@pytest.mark.parametrize("W,H", [(1, 1), (17, 5), (128, 128), (1237, 822)])
def test_compute_rd_matches_inverse_intrinsics(W, H):
    say(f"[camera] compute_rd {W}x{H} vs R K^-1 [u v 1]")
    cam = sphere_camera(3, 2.5, W, H, 0.9 * W, 0.8 * W, device="cuda")
    rc = _RefCamera(cam)
    rd = rc.compute_rd()
    assert rd.shape == (3, H, W) and rd.dtype == torch.float32 and rd.is_cuda
    ref = _pinhole_dirs(cam, W, H, cam.cx, cam.cy)
    err = float(np.abs(rd.cpu().double().numpy() - ref).max())
    say(f"  max-abs {err:.2e}")
    assert err <= 2e-6
    # a half-size render target (wh / cxcy rescaled as cameras.py:79-80 does)
    rd2 = rc.compute_rd(wh=(max(W // 2, 1), max(H // 2, 1)))
    W2, H2 = max(W // 2, 1), max(H // 2, 1)
    ref2 = _pinhole_dirs(cam, W2, H2, cam.cx * W2 / W, cam.cy * H2 / H)
    # the focal length of the half-size grid follows from the same tan(fov): K uses W2
    assert float(np.abs(rd2.cpu().double().numpy() - ref2).max()) <= 2e-6


@pytest.mark.gpu
def test_depth2pts_and_depth2normal_path():
    say("[camera] depth2pts: points project back to their pixels at their depth")
    W, H = 161, 97
    cam = sphere_camera(5, 2.2, W, H, 0.7 * W, 0.7 * W, device="cuda")
    g = torch.Generator().manual_seed(4)
    depth = (torch.rand(H, W, generator=g) * 2 + 0.5).cuda()
    pts = svraster_cuda.utils.depth2pts(W, H, cam.cx, cam.cy, cam.tanfovx, cam.tanfovy, cam.c2w, depth)
    assert pts.shape == (3, H, W)
    # the reference camera's own composition (position + rd * depth) agrees
    pts2 = _RefCamera(cam).depth2pts(depth[None])
    assert float((pts - pts2).abs().max()) <= 1e-6
    # world -> camera -> pixel: z equals the depth, (u, v) the pixel centre
    w2c = cam.w2c.cpu().double().numpy()
    P = pts.cpu().double().numpy().reshape(3, -1)
    pc = w2c[:3, :3] @ P + w2c[:3, 3:4]
    fx = W / (2.0 * cam.tanfovx)
    fy = H / (2.0 * cam.tanfovy)
    u = fx * pc[0] / pc[2] + cam.cx
    v = fy * pc[1] / pc[2] + cam.cy
    uu, vv = np.meshgrid(np.arange(W) + 0.5, np.arange(H) + 0.5)
    assert np.abs(pc[2] - depth.cpu().double().numpy().reshape(-1)).max() <= 1e-5
    assert np.abs(u - uu.reshape(-1)).max() <= 1e-3 and np.abs(v - vv.reshape(-1)).max() <= 1e-3
    # empty image
    e = svraster_cuda.utils.depth2pts(0, 0, 0.0, 0.0, 0.5, 0.5, cam.c2w, torch.empty(0, device="cuda"))
    assert e.shape == (3, 0, 0)


@pytest.mark.gpu
def test_is_in_cone():
    say("[camera] is_in_cone vs projection into the image rectangle")
    W, H = 200, 120
    cam = look_at_camera(np.array([0.3, -2.0, 0.7]), W, H, 150.0, 150.0, device="cuda")
    g = torch.Generator().manual_seed(8)
    pts = (torch.rand(50_000, 3, generator=g) * 6 - 3).cuda()
    got = svraster_cuda.utils.is_in_cone(pts, cam).cpu().numpy()
    assert got.dtype == np.bool_
    w2c = cam.w2c.cpu().double().numpy()
    pc = (w2c[:3, :3] @ pts.cpu().double().numpy().T + w2c[:3, 3:4])
    z = pc[2]
    u = pc[0] / np.where(z != 0, z, 1e-30) / cam.tanfovx     # in [-1, 1] inside the image
    v = pc[1] / np.where(z != 0, z, 1e-30) / cam.tanfovy
    ref = (z > cam.near) & (np.abs(u) <= 1.0) & (np.abs(v) <= 1.0)
    clear = (np.abs(z - cam.near) > 1e-5) & (np.abs(np.abs(u) - 1.0) > 1e-5) & (np.abs(np.abs(v) - 1.0) > 1e-5)
    say(f"  {int(ref.sum())} of {pts.shape[0]} inside; {int((~clear).sum())} on a boundary")
    assert np.array_equal(got[clear], ref[clear]) and ref.sum() > 100
    assert svraster_cuda.utils.is_in_cone(torch.empty(0, 3, device="cuda"), cam).numel() == 0


@pytest.mark.gpu
def test_voxel_order_rank():
    say("[camera] voxel_order_rank vs the rank codec")
    rng = np.random.default_rng(3)
    paths = rng.integers(0, 1 << 47, size=1001, dtype=np.int64)
    got = svraster_cuda.utils.voxel_order_rank(torch.from_numpy(paths).cuda()).cpu().numpy()
    assert got.shape == (8, 1001)
    for q in range(8):
        assert np.array_equal(got[q].astype(np.uint64), paths.astype(np.uint64) ^ ob.REP3[q])
    # octant 0 keeps the path; octant 7 reverses every level's child order
    assert np.array_equal(got[0], paths)
    assert svraster_cuda.utils.voxel_order_rank(torch.empty(0, dtype=torch.int64, device="cuda")).shape == (8, 0)
