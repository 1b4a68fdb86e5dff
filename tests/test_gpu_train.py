"""Training-iteration pieces on the GPU: device-step Adam vs the host-driven unbiased step, the SH
transpose fused with Adam vs the unfused pair, the captured-step grid Laplacian vs the drop-in
grid_loss_bw call, and a captured full iteration (render, MSE, backward, smoothness, Adam)
checked replay by replay against the drop-in pieces fed the same gradients."""
import ctypes

import numpy as np
import pytest
import torch

import svraster_amd as svraster_cuda
from svraster_amd import _C, _lib
from svraster_amd.batched import (BatchedRasterizer, BatchedTrainIteration, GridSmoothness, adam_bias_table,
                                  adam_table_length)
from svraster_amd.model import VoxelModel
from tests._util import rel_l2, say
from tests.test_gpu_batched import _setup


def _p(t):
    return ctypes.c_void_p(t.data_ptr())


@pytest.mark.gpu
@pytest.mark.parametrize("sparse", [True, False])
def test_device_step_adam_matches_host_step(sparse):
    say(f"[adam] device step counter vs host step, sparse={sparse}")
    lib = _lib.load()
    betas, lr, eps = (0.9, 0.99), 1e-2, 1e-15
    n_tab = adam_table_length(betas)
    tab = adam_bias_table(lr, betas, n_tab, "cuda")
    g = torch.Generator().manual_seed(3)
    n = 100_003
    p0 = torch.randn(n, generator=g).cuda()
    a = [p0.clone(), torch.zeros(n, device="cuda"), torch.zeros(n, device="cuda")]
    b = [p0.clone(), torch.zeros(n, device="cuda"), torch.zeros(n, device="cuda")]
    step = torch.zeros(1, dtype=torch.int32, device="cuda")
    for t in (1, 2, 3, 7, n_tab, n_tab + 5):
        grad = torch.randn(n, generator=g).cuda()
        grad[torch.rand(n, generator=g).cuda() < 0.3] = 0.0
        _C.unbiased_adam_step(sparse, a[0], grad, a[1], a[2], t, lr, betas[0], betas[1], eps)
        step.fill_(t)
        _lib.check(lib.svr_sparse_adam_step_dev(int(sparse), _p(b[0]), _p(grad), _p(b[1]), _p(b[2]), n, _p(step),
                                                _p(tab), n_tab, ctypes.c_float(betas[0]), ctypes.c_float(betas[1]),
                                                ctypes.c_float(eps), None))
        torch.cuda.synchronize()
        for x, y, name in zip(a, b, ("param", "exp_avg", "exp_avg_sq")):
            assert torch.equal(x, y), f"step {t}: {name} differs"


@pytest.mark.gpu
@pytest.mark.parametrize("sparse", [False, True])
def test_sh_backward_adam_matches_unfused(sparse):
    say(f"[sh] SH transpose + Adam in one launch vs svr_batch_sh_backward + unbiased_adam_step, sparse={sparse}")
    w, sc, cams, targets = _setup()
    m = VoxelModel(sc)
    r = BatchedRasterizer(sc.octpath, sc.octlevel, sc.vox_center, sc.vox_size, sc.is_leaf, sc.vox_key, cams,
                          m.bg_color, sc.sh_degree)
    s_val = torch.exp(m._log_s.detach().reshape(1) * 10.0)
    r.forward(m._geo_grid_pts.detach(), m._sh0.detach(), m._shs.detach(), s_val)
    torch.cuda.synchronize()
    lib = _lib.load()
    it_state = BatchedTrainIteration(r, m._geo_grid_pts.detach().clone(), m._sh0.detach().clone(),
                                     m._shs.detach().clone(), m._log_s.detach(), targets, smooth=None,
                                     lr=1e-2, lr_sh0=2e-3, lr_shs=5e-3, sparse=sparse)
    A = it_state.sh_adam
    ref = [m._sh0.detach().clone(), torch.zeros_like(A.m0), torch.zeros_like(A.v0),
           m._shs.detach().clone(), torch.zeros_like(A.ms), torch.zeros_like(A.vs)]
    g = torch.Generator().manual_seed(5)
    for t in (1, 2):
        drgb = torch.randn(r.drgb_views.shape, generator=g).cuda()
        if t == 2:
            # half the voxels get no gradient on the second step: the dense step still moves
            # them (decaying moments), the sparse step leaves them alone
            drgb[:, ::2] = 0.0
        dsh0 = torch.empty_like(r.dsh0)
        dshs = torch.empty_like(r.dshs)
        d1 = drgb.clone()
        _lib.check(lib.svr_batch_sh_backward(_p(r.ws), *r._dims(), _p(r.vox_center), _p(d1), r.deg, r.M, _p(dsh0),
                                             _p(dshs), None))
        _C.unbiased_adam_step(sparse, ref[0], dsh0, ref[1], ref[2], t, 2e-3, 0.9, 0.99, 1e-15)
        _C.unbiased_adam_step(sparse, ref[3], dshs, ref[4], ref[5], t, 5e-3, 0.9, 0.99, 1e-15)
        A.step.fill_(t)
        d2 = drgb.clone()
        _lib.check(lib.svr_batch_sh_backward_adam(
            _p(r.ws), *r._dims(), _p(r.vox_center), _p(d2), r.deg, r.M, _p(A.sh0), _p(A.m0), _p(A.v0), _p(A.table0),
            _p(A.shs), _p(A.ms), _p(A.vs), _p(A.tables), A.n_table, _p(A.step), ctypes.c_float(0.9),
            ctypes.c_float(0.99), ctypes.c_float(1e-15), int(sparse), None))
        torch.cuda.synchronize()
        n_upd = int((dshs != 0).sum()) + int((dsh0 != 0).sum())
        say(f"  step {t}: {n_upd} SH elements updated")
        assert n_upd > 0
        assert torch.equal(d1, d2), "both launches clear the gradient entries they consume"
        for x, y, name in zip(ref, [A.sh0, A.m0, A.v0, A.shs, A.ms, A.vs],
                              ("sh0", "sh0 m", "sh0 v", "shs", "shs m", "shs v")):
            assert torch.equal(x, y), f"step {t}: {name} differs"


def _smooth_setup():
    w, sc, cams, targets = _setup()
    sm = GridSmoothness(sc.vox_center, sc.vox_size, sc.octlevel, sc.is_leaf, sc.vox_key, sc.scene_center, 2.0,
                        weight=1e-3, seed=11)
    return w, sc, cams, targets, sm


@pytest.mark.gpu
def test_grid_dense_index():
    say("[grid] dense cell -> key index map")
    w, sc, cams, targets, sm = _smooth_setup()
    keys = sm.keys.long()
    dense = sm.dense.long()
    hit = dense >= 0
    cells = torch.arange(sm.res ** 3, device="cuda")
    assert torch.equal(keys[dense[hit]], cells[hit])
    present = torch.zeros(sm.res ** 3, dtype=torch.bool, device="cuda")
    present[keys] = True
    assert torch.equal(present, hit)
    say(f"  G={sm.G}, res={sm.res}, cells with a key {int(hit.sum())}")


@pytest.mark.gpu
def test_grid_smoothness_step_matches_api():
    say("[grid] captured-step Laplacian (dense index, device step seed) vs grid_loss_bw.laplacian_smoothness")
    w, sc, cams, targets, sm = _smooth_setup()
    step = torch.tensor([3], dtype=torch.int32, device="cuda")
    assert sm.pull, "cells of a table built from an octree are unique"
    got = torch.zeros_like(sc.geo_grid_pts)
    sm(sc.geo_grid_pts, got, step)
    push = GridSmoothness(sc.vox_center, sc.vox_size, sc.octlevel, sc.is_leaf, sc.vox_key, sc.scene_center, 2.0,
                          weight=1e-3, seed=11, pull=False)
    got_push = torch.zeros_like(sc.geo_grid_pts)
    push(sc.geo_grid_pts, got_push, step)
    ref = torch.zeros_like(sc.geo_grid_pts)
    seed = (11 + 0x9E3779B9 * 3) & 0xFFFFFFFF
    svraster_cuda.grid_loss_bw.laplacian_smoothness(
        grid_pts=sc.geo_grid_pts, vox_key=sc.vox_key, grid_voxel_coord=sm.coord, grid_voxel_size=sm.size,
        grid_res=sm.res, grid_mask=sm.mask, grid_keys=sm.keys, grid2voxel=sm.g2v, active_list=sm.active,
        weight=1e-3, vox_size_inv=sm.vox_size_inv, no_tv_s=True, tv_sparse=False, grid_pts_grad=ref, seed=seed)
    torch.cuda.synchronize()
    e = rel_l2(got, ref)
    e_push = rel_l2(got_push, ref)
    say(f"  rel-L2 two-pass {e:.2e}, per-key scatter {e_push:.2e}, |ref| {ref.norm().item():.3e}")
    assert ref.norm().item() > 0 and e <= 1e-5 and e_push <= 1e-5
    other = torch.zeros_like(sc.geo_grid_pts)
    sm(sc.geo_grid_pts, other, step + 1)
    assert not torch.equal(other, got), "a new step draws new sample offsets"


@pytest.mark.gpu
def test_grid_eikonal_and_laplacian_table_order_match_api():
    say("[grid] table-order eikonal + Laplacian (one launch pair) vs grid_loss_bw.grid_eikonal + laplacian_smoothness")
    w, sc, cams, targets = _setup()
    w_ge, w_ls = 2e-3, 1e-3
    sm = GridSmoothness(sc.vox_center, sc.vox_size, sc.octlevel, sc.is_leaf, sc.vox_key, sc.scene_center, 2.0,
                        weight=w_ls, seed=11, eikonal_weight=w_ge)
    assert sm.pull and sm.eikonal
    step = torch.tensor([5], dtype=torch.int32, device="cuda")
    got = torch.zeros_like(sc.geo_grid_pts)
    sm(sc.geo_grid_pts, got, step)
    seed = (11 + 0x9E3779B9 * 5) & 0xFFFFFFFF
    kw = dict(grid_pts=sc.geo_grid_pts, vox_key=sc.vox_key, grid_voxel_coord=sm.coord, grid_voxel_size=sm.size,
              grid_res=sm.res, grid_mask=sm.mask, grid_keys=sm.keys, grid2voxel=sm.g2v, active_list=sm.active,
              vox_size_inv=sm.vox_size_inv, no_tv_s=True, tv_sparse=False, seed=seed)
    ref = torch.zeros_like(sc.geo_grid_pts)
    svraster_cuda.grid_loss_bw.laplacian_smoothness(weight=w_ls, grid_pts_grad=ref, **kw)
    svraster_cuda.grid_loss_bw.grid_eikonal(weight=w_ge, grid_pts_grad=ref, **kw)
    eik = torch.zeros_like(sc.geo_grid_pts)
    svraster_cuda.grid_loss_bw.grid_eikonal(weight=w_ge, grid_pts_grad=eik, **kw)
    # the per-key scatter form (dense cell index, same offsets) of both terms
    push = GridSmoothness(sc.vox_center, sc.vox_size, sc.octlevel, sc.is_leaf, sc.vox_key, sc.scene_center, 2.0,
                          weight=w_ls, seed=11, eikonal_weight=w_ge, pull=False)
    got_push = torch.zeros_like(sc.geo_grid_pts)
    push(sc.geo_grid_pts, got_push, step)
    torch.cuda.synchronize()
    e = rel_l2(got, ref)
    e_push = rel_l2(got_push, ref)
    say(f"  rel-L2 {e:.2e} (drop-in calls), {e_push:.2e} (per-key scatter); |eikonal part| {eik.norm().item():.3e}, "
        f"|total| {ref.norm().item():.3e}")
    assert eik.norm().item() > 0 and e <= 1e-5 and e_push <= 1e-5
    # eikonal alone through the weights in device memory (Laplacian weight 0), and back
    sm.set_weights(laplacian=0.0)
    got2 = torch.zeros_like(sc.geo_grid_pts)
    sm(sc.geo_grid_pts, got2, step)
    torch.cuda.synchronize()
    assert rel_l2(got2, eik) <= 1e-5
    with pytest.raises(ValueError):
        GridSmoothness(sc.vox_center, sc.vox_size, sc.octlevel, sc.is_leaf, sc.vox_key, sc.scene_center, 2.0,
                       weight=w_ls).set_weights(eikonal=1.0)


def test_trainer_smoothness_weights_schedule():
    from svraster_amd.batched import trainer_smoothness_weights as tw
    assert tw(0) == (1e-8, 1e-10)
    assert tw(2500) == (1e-8 * 0.25, 1e-10 * 0.25)
    assert tw(5999) == (1e-8 * 0.0625, 1e-10 * 0.0625)
    assert tw(7000) == (0.0, 1e-10 * 0.0625)
    assert tw(9000) == (0.0, 0.0)
    ge, ls = tw(100, G=1000, ge_drop_ratio=0.5)
    assert ge == 1e-8 * 2.0 and ls == 1e-10


@pytest.mark.gpu
@pytest.mark.parametrize("sparse", [False, True])
def test_train_iteration_graph_replays(sparse):
    say(f"[train] captured iteration: each replay = drop-in render + smoothness + Adam on the same inputs, "
        f"sparse={sparse}")
    w, sc, cams, targets, sm = _smooth_setup()
    m = VoxelModel(sc)
    r = BatchedRasterizer(sc.octpath, sc.octlevel, sc.vox_center, sc.vox_size, sc.is_leaf, sc.vox_key, cams,
                          m.bg_color, sc.sh_degree)
    grid = m._geo_grid_pts.detach().clone()
    sh0, shs = m._sh0.detach().clone(), m._shs.detach().clone()
    it = BatchedTrainIteration(r, grid, sh0, shs, m._log_s.detach(), targets, smooth=sm, lr=1e-2,
                               betas=(0.9, 0.99), sparse=sparse)
    it.capture(warmup=1)
    torch.cuda.synchronize()
    for k in range(3):
        t = int(it.step.item()) + 1
        g0, gm, gv = grid.clone(), it.grid_m.clone(), it.grid_v.clone()
        s0, ss0 = sh0.clone(), shs.clone()
        it()
        torch.cuda.synchronize()
        junk = torch.randn(1000, device="cuda").sum()       # eager work between replays
        assert int(it.step.item()) == t
        # the replay rendered the parameters it started from
        mv = VoxelModel(sc)
        with torch.no_grad():
            mv._geo_grid_pts.copy_(g0)
            mv._sh0.copy_(s0)
            mv._shs.copy_(ss0)
            err = max((r.view_slice(r.color, 3, v) - mv.render(cam)["color"]).abs().max().item()
                      for v, cam in enumerate(cams))
        # Adam on the corner SDFs: exactly the host-driven step on the replay's gradient
        _C.unbiased_adam_step(sparse, g0, it.grid_grad, gm, gv, t, 1e-2, 0.9, 0.99, 1e-15)
        n_upd = int((it.grid_grad != 0).sum())
        n_sh = int((sh0 != s0).sum()) + int((shs != ss0).sum())
        say(f"  replay {k}: step {t}, colour max-abs vs drop-in render {err:.2e}, {n_upd} corner SDFs and "
            f"{n_sh} SH coefficients updated, junk {junk.item():.2f}")
        assert err <= 1e-6 and n_upd > 0 and n_sh > 0
        assert torch.equal(g0, grid) and torch.equal(gm, it.grid_m) and torch.equal(gv, it.grid_v)
