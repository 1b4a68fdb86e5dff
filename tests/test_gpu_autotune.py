"""The autotuner's knobs change speed only: every tile shape, staging depth and persistent grid
size renders the same images and gradients as the default configuration; the tuner itself
returns a configuration from its space and caches it."""
import os

import pytest
import torch

from svraster_amd import _lib, autotune
from svraster_amd.batched import BatchedRasterizer, BatchedTrainStep
from svraster_amd.model import VoxelModel
from tests._util import rel_l2, say
from tests.test_gpu_batched import _setup


def _run(sc, cams, targets, tile_shape=(16, 16)):
    m = VoxelModel(sc)
    r = BatchedRasterizer(sc.octpath, sc.octlevel, sc.vox_center, sc.vox_size, sc.is_leaf, sc.vox_key, cams,
                          m.bg_color, sc.sh_degree, tile_shape=tile_shape)
    step = BatchedTrainStep(r, m._geo_grid_pts.detach(), m._sh0.detach(), m._shs.detach(), m._log_s.detach(),
                            targets)
    step()
    torch.cuda.synchronize()
    return (r.color.clone(), r.T.clone(), step.grid_grad.clone(), step.sh0_grad.clone(), step.shs_grad.clone(),
            step.s_grad.clone(), r.key_count())


@pytest.fixture
def default_options():
    yield
    autotune.apply_options(autotune.DEFAULT)


@pytest.mark.gpu
@pytest.mark.parametrize("cfg", [
    {"tile_shape": (32, 8)}, {"tile_shape": (8, 32)}, {"batch_depth": 64}, {"bwd_batch_depth": 128},
    {"persistent_blocks_per_cu": 4}, {"tile_shape": (32, 8), "batch_depth": 64, "persistent_blocks_per_cu": 16},
])
def test_launch_configuration_keeps_results(cfg, default_options):
    say(f"[autotune] configuration {cfg} vs the default")
    w, sc, cams, targets = _setup(n_views=6)
    autotune.apply_options(autotune.DEFAULT)
    ref = _run(sc, cams, targets)
    full = dict(autotune.DEFAULT, **cfg)
    autotune.apply_options(full)
    got = _run(sc, cams, targets, tile_shape=full["tile_shape"])
    assert torch.equal(got[0], ref[0]) and torch.equal(got[1], ref[1]), "colour and T are bit-identical"
    errs = [rel_l2(a, b) for a, b in zip(got[2:6], ref[2:6])]
    say(f"  keys {got[6]} vs {ref[6]}; grad rel-L2 grid {errs[0]:.1e} sh0 {errs[1]:.1e} shs {errs[2]:.1e} "
        f"s {errs[3]:.1e}")
    assert max(errs) <= 1e-5


@pytest.mark.gpu
def test_autotune_picks_and_caches(tmp_path, default_options):
    say("[autotune] coordinate descent on a 4-view workload")
    os.environ["SVRASTER_TUNE_CACHE"] = str(tmp_path / "tune.json")
    try:
        w, sc, cams, targets = _setup()
        m = VoxelModel(sc)

        def make_step(shape):
            r = BatchedRasterizer(sc.octpath, sc.octlevel, sc.vox_center, sc.vox_size, sc.is_leaf, sc.vox_key,
                                  cams, m.bg_color, sc.sh_degree, tile_shape=shape)
            return BatchedTrainStep(r, m._geo_grid_pts.detach(), m._sh0.detach(), m._shs.detach(),
                                    m._log_s.detach(), targets)

        sig = autotune.workload_signature(sc.num_voxels, cams, sc.sh_degree)
        best, table = autotune.autotune(make_step, sig, steps=2, warmup=1, log=say)
        assert best["tile_shape"] in autotune.TILE_SHAPES
        assert best["batch_depth"] in autotune.BATCH_DEPTHS
        assert best["bwd_batch_depth"] in autotune.BWD_BATCH_DEPTHS
        assert best["persistent_blocks_per_cu"] in autotune.BLOCKS_PER_CU
        assert best["compositor_variant"] in autotune.VARIANTS
        assert best["fuse_gather_bw"] in autotune.FUSE_GATHER_BW
        assert len(table) == 9
        assert _lib.get_option("batch_depth") == best["batch_depth"]
        cached, rows = autotune.autotune(make_step, sig)
        assert cached == best and rows[0][0] == "cached"
    finally:
        del os.environ["SVRASTER_TUNE_CACHE"]
