"""Checkpoint round trips on the GPU and the eval-fps harness."""
import torch
import pytest

from svraster_amd import eval_fps as ef
from svraster_amd import io as svio
from svraster_amd.model import VoxelModel
from svraster_amd.synthetic import make_workload
from tests._util import say


@pytest.mark.gpu
def test_checkpoint_round_trip_renders_the_same(tmp_path):
    say("[io] save / load, plain and quantized")
    w = make_workload("fd_256", device="cuda")
    m = VoxelModel(w.scene)
    cam = w.cameras[0]
    with torch.no_grad():
        ref = m.render(cam)["color"].clone()
    p0, p1 = str(tmp_path / "plain.pt"), str(tmp_path / "quant.pt")
    svio.save(m, p0)
    svio.save(m, p1, quantize=True)
    a = svio.load(p0)
    assert torch.equal(a.vox_key, m.vox_key) and torch.equal(a.grid_pts_key, m.grid_pts_key)
    assert torch.equal(a.vox_size.reshape(-1), m.vox_size.reshape(-1))
    assert float((a.vox_center - m.vox_center).abs().max()) <= 1e-6
    with torch.no_grad():
        got = a.render(cam)["color"]
    e0 = float((got - ref).abs().max())
    b = svio.load(p1)
    with torch.no_grad():
        gq = b.render(cam)["color"]
    mse = float(((gq - ref) ** 2).mean())
    psnr = -10.0 * torch.log10(torch.tensor(mse)).item() if mse > 0 else float("inf")
    say(f"  plain: max-abs {e0:.2e}; quantized: PSNR {psnr:.1f} dB vs the original")
    assert e0 <= 1e-6 and psnr >= 30.0


@pytest.mark.gpu
def test_eval_fps_harness():
    say("[io] eval-fps harness, per-view and batched")
    w = make_workload("fd_256", device="cuda")
    m = VoxelModel(w.scene)
    cams = ef.sphere_cameras([0.0, 0.0, 0.0], 2.0, 4, 320, 240)
    pv = ef.eval_fps(m, cams)
    bt = ef.eval_fps_batched(m, cams)
    say(f"  per-view {pv['fps']:.0f} fps, batched {bt['fps']:.0f} fps")
    assert pv["n"] == 4 and pv["fps"] > 0 and bt["fps"] > 0 and pv["resolution"] == [240, 320]
