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


def test_quantization_is_a_lloyd_fixed_point():
    """At convergence every value sits with its nearest codeword (ties to the upper one) and
    every used codeword is the mean of its values: the two k-means optimality conditions."""
    g = torch.Generator().manual_seed(7)
    src = torch.cat([torch.randn(3000, generator=g) * 0.1, torch.rand(2000, generator=g) * 4 + 1])
    q = svio.quantization(src, max_iter=200)
    cb, idx = q["codebook"].double(), q["index"].long()
    v = src.double()
    d = (v[:, None] - cb[None, :]).abs()
    best = d.min(dim=1).values
    assert torch.all(d[torch.arange(v.numel()), idx] <= best + 1e-6)
    assert torch.all(cb[1:] >= cb[:-1])


def test_quantization_known_answers():
    # 512 distinct values: 256 runs of two, codeword k = mean of {2k, 2k+1}; already optimal
    src = torch.arange(512, dtype=torch.float32)
    q = svio.quantization(src)
    assert torch.equal(q["codebook"], torch.arange(256, dtype=torch.float32) * 2 + 0.5)
    assert torch.equal(q["index"].long(), torch.arange(512) // 2)
    # fewer values than codewords: exact reconstruction, nondecreasing codebook
    src = torch.tensor([[3.0], [-1.0], [2.5], [3.0]])
    q = svio.quantization(src)
    assert q["index"].shape == (4, 1) and q["codebook"].shape == (256,)
    assert torch.equal(svio.dequantization(q), src)
    assert torch.all(q["codebook"][1:] >= q["codebook"][:-1])
    # a constant tensor
    q = svio.quantization(torch.full((1000,), 0.25))
    assert torch.equal(svio.dequantization(q), torch.full((1000,), 0.25))


def test_quantization_not_worse_than_equal_count_start():
    g = torch.Generator().manual_seed(9)
    src = torch.randn(20000, generator=g) ** 3
    q0 = svio.quantization(src, max_iter=0)
    q1 = svio.quantization(src, max_iter=10)
    e0 = float(((svio.dequantization(q0) - src) ** 2).mean())
    e1 = float(((svio.dequantization(q1) - src) ** 2).mean())
    assert e1 <= e0


def test_dequantize_reference_layout_checkpoint(tmp_path):
    """A checkpoint written by hand in the on-disk layout (index uint8 + 256-float codebook per
    entry, sh0 / shs as per-column lists) loads with the safe loader and dequantizes to
    codebook[index]."""
    g = torch.Generator().manual_seed(2)
    n, M = 40, 90
    cb = lambda: torch.sort(torch.randn(256, generator=g)).values
    ix = lambda *s: torch.randint(0, 256, s, generator=g, dtype=torch.int64).to(torch.uint8)
    geo = {"index": ix(M, 1), "codebook": cb()}
    sh0 = [{"index": ix(n, 1), "codebook": cb()} for _ in range(3)]
    shs = [{"index": ix(n, 1, 3), "codebook": cb()} for _ in range(15)]
    state = {"active_sh_degree": 3, "ss": 1.0, "scene_center": torch.zeros(3), "inside_extent": torch.tensor(2.0),
             "scene_extent": torch.tensor(64.0), "octpath": torch.zeros(n, 1, dtype=torch.int64),
             "octlevel": torch.full((n, 1), 6, dtype=torch.int8), "_geo_grid_pts": geo, "_sh0": sh0, "_shs": shs,
             "_log_s": torch.tensor(0.5), "is_leaf": torch.ones(n, 1, dtype=torch.bool), "quantized": True}
    path = tmp_path / "ref_layout.pt"
    torch.save(state, str(path))
    st = torch.load(str(path), map_location="cpu", weights_only=True)
    svio.dequantize_state_dict(st)
    assert torch.equal(st["_geo_grid_pts"], geo["codebook"][geo["index"].long()])
    assert st["_sh0"].shape == (n, 3) and st["_shs"].shape == (n, 15, 3)
    for c in range(3):
        assert torch.equal(st["_sh0"][:, c:c + 1], sh0[c]["codebook"][sh0[c]["index"].long()])
    for k in range(15):
        assert torch.equal(st["_shs"][:, k:k + 1], shs[k]["codebook"][shs[k]["index"].long()])
    assert svio.outside_level_of(st["scene_extent"], st["inside_extent"]) == 5


def test_outside_level_grid_table():
    """Grid table level and voxel cell coordinates with outside_level = 5 (train.py:961-965):
    the inside region [-1, 1]^3 of a 64-wide scene, deepest level 5 + 7 -> 2^7 cells per axis."""
    from svraster_amd.batched import grid_table_level, grid_voxel_coords
    octlevel = torch.tensor([[6], [12], [9]], dtype=torch.int8)
    assert grid_table_level(octlevel, 5) == 7
    assert grid_table_level(octlevel, 0) == 9
    assert grid_table_level(torch.tensor([[14]], dtype=torch.int8), 5) == 9
    inside = 2.0
    # a level-12 voxel (5 outside + 7 inside) of size 2 / 2^7 at the inside region's min corner
    size = torch.tensor([inside / 2 ** 7, inside / 2 ** 2])
    center = torch.stack([torch.full((3,), -1.0) + size[0] / 2, torch.tensor([0.25, 0.25, -0.25])])
    coord, cells = grid_voxel_coords(center, size, torch.zeros(3), inside, 7)
    assert torch.equal(coord[0], torch.zeros(3)) and float(cells[0]) == 1.0
    assert torch.equal(coord[1], torch.tensor([64.0, 64.0, 32.0])) and float(cells[1]) == 32.0
    try:
        svio.outside_level_of(torch.tensor(3.0), torch.tensor(2.0))
        raise AssertionError("a non-power-of-two ratio must raise")
    except ValueError:
        pass
