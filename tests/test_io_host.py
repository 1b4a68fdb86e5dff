"""Checkpoint format and the 8-bit codebook quantization on the host (no GPU)."""
import types

import numpy as np
import torch

from svraster_amd import io as svio


def test_quantization_codebook_properties():
    g = torch.Generator().manual_seed(0)
    for src in (torch.randn(5000, generator=g), torch.rand(300, 1, generator=g) * 7 - 2,
                torch.randn(20, 15, generator=g) ** 3):
        q = svio.quantization(src)
        idx, cb = q["index"], q["codebook"]
        assert idx.dtype == torch.uint8 and idx.shape == src.shape and cb.shape == (256,)
        rec = svio.dequantization(q)
        # every used entry is the mean of the values assigned to it
        flat, ii = src.flatten().double(), idx.flatten().long()
        for k in torch.unique(ii).tolist():
            assert abs(float(flat[ii == k].mean()) - float(cb[k])) <= 1e-5 * (1 + abs(float(cb[k])))
        # reconstruction error: small on average; the tail buckets of a quantile codebook are wide
        span = float(src.max() - src.min())
        assert float((rec - src).abs().max()) <= span / 8
        assert float((rec - src).abs().mean()) <= span / 256
    u = torch.rand(20000, generator=g)
    rec = svio.dequantization(svio.quantization(u))
    assert float((rec - u).abs().max()) <= 1.0 / 128     # uniform data: buckets ~1/256 wide


def test_state_dict_format_round_trip(tmp_path):
    n = 50
    g = torch.Generator().manual_seed(1)
    model = types.SimpleNamespace(
        active_sh_degree=3, ss=1.0, scene_center=torch.zeros(3), scene_extent=torch.tensor(2.0),
        octpath=torch.arange(n, dtype=torch.int64).reshape(n, 1) << 30, octlevel=torch.full((n, 1), 5, dtype=torch.int8),
        _geo_grid_pts=torch.randn(200, 1, generator=g), _sh0=torch.randn(n, 3, generator=g),
        _shs=torch.randn(n, 15, 3, generator=g), _log_s=torch.tensor(0.3), is_leaf=torch.ones(n, dtype=torch.bool))
    for quantize in (False, True):
        path = tmp_path / f"m{int(quantize)}.pt"
        svio.save(model, str(path), quantize=quantize)
        st = torch.load(str(path), map_location="cpu", weights_only=True)
        assert st["quantized"] == quantize
        for k in ("active_sh_degree", "ss", "scene_center", "inside_extent", "scene_extent", "octpath", "octlevel",
                  "_geo_grid_pts", "_sh0", "_shs", "_log_s", "is_leaf"):
            assert k in st
        if quantize:
            assert isinstance(st["_sh0"], list) and len(st["_sh0"]) == 3 and len(st["_shs"]) == 15
            svio.dequantize_state_dict(st)
            assert st["_shs"].shape == (n, 15, 3)
            assert float((st["_geo_grid_pts"] - model._geo_grid_pts).abs().max()) < 0.2
        else:
            assert torch.equal(st["_shs"], model._shs) and torch.equal(st["octpath"], model.octpath)
