"""Batched multi-view rasterization: all views of one voxel set rendered and differentiated with
one launch per stage (preprocess over every (view, voxel) pair, one scan, one duplication, one
sort, one persistent compositor launch over a global (view, tile) work queue, one persistent
reverse launch, one view-summed SH transpose).  No stage reads a count back to the host, so a
whole step can be captured once and replayed as a HIP graph.

Results per view equal the drop-in per-view path: the same kernels run on the same per-view
data; view v's keys carry its global tile id (tile_base + local tile) so one stable sort yields
every view's reference ordering.

There is no counterpart in the reference, whose API renders one view per call
(cuda/svraster_cuda/renderer.py:44-152); the per-view semantics restated are those of
rasterize_voxels + SH_eval + GatherGeoParams.
"""
from __future__ import annotations

import ctypes
from typing import List, Optional

import numpy as np
import torch

from . import _lib

MAX_NUM_LEVELS = 16


def _p(t: Optional[torch.Tensor]):
    return None if t is None or t.numel() == 0 else ctypes.c_void_p(t.data_ptr())


class BatchedRasterizer:
    """Fixed voxel topology + fixed camera set.  Parameters (corner SDFs, SH, s) may change
    between calls; buffers are allocated once."""

    def __init__(self, octpath, octlevel, vox_center, vox_size, is_leaf, vox_key, cameras: List,
                 background: torch.Tensor, sh_degree: int, n_sh_coeffs: int = 16,
                 key_capacity: Optional[int] = None, need_depth=False, need_normal=False, cull: bool = True):
        self.dev = vox_center.device
        self.lib = _lib.load()
        self.V = len(cameras)
        self.P = int(vox_center.shape[0])
        self.octpath = octpath.reshape(-1).contiguous()
        self.vox_center = vox_center.contiguous()
        self.vox_size = vox_size.reshape(-1).contiguous()
        self.is_leaf = is_leaf.reshape(-1).to(torch.uint8).contiguous()
        self.vox_key = vox_key.contiguous()
        self.deg = int(sh_degree)
        self.M = int(n_sh_coeffs)
        self.need_depth, self.need_normal = bool(need_depth), bool(need_normal)
        # Rank order of every (voxel, octant) pair: fixed while the voxel set is.
        tmp = torch.empty(int(self.lib.svr_batch_pair_order_bytes(self.P)), dtype=torch.uint8, device=self.dev)
        self.pair_order = torch.empty(8 * self.P, dtype=torch.int32, device=self.dev)
        _lib.check(self.lib.svr_batch_pair_order(self.P, _p(self.octpath), _p(tmp), _p(self.pair_order),
                                                 self._stream()), "svr_batch_pair_order")
        del tmp
        self.Ws = [c.image_width for c in cameras]
        self.Hs = [c.image_height for c in cameras]
        self.tiles = [((w + 15) // 16) * ((h + 15) // 16) for w, h in zip(self.Ws, self.Hs)]
        self.pix = [w * h for w, h in zip(self.Ws, self.Hs)]
        self.n_tiles = int(sum(self.tiles))
        self.n_pix = int(sum(self.pix))
        self.pix_base = np.concatenate([[0], np.cumsum(self.pix)[:-1]]).astype(np.int64)
        self._host = dict(
            W=(ctypes.c_int * self.V)(*self.Ws), H=(ctypes.c_int * self.V)(*self.Hs),
            tx=(ctypes.c_float * self.V)(*[c.tanfovx for c in cameras]),
            ty=(ctypes.c_float * self.V)(*[c.tanfovy for c in cameras]),
            cx=(ctypes.c_float * self.V)(*[c.cx for c in cameras]),
            cy=(ctypes.c_float * self.V)(*[c.cy for c in cameras]),
            near=(ctypes.c_float * self.V)(*[c.near for c in cameras]))
        self.w2c = torch.stack([c.w2c for c in cameras]).float().contiguous().to(self.dev)
        self.c2w = torch.stack([c.c2w for c in cameras]).float().contiguous().to(self.dev)
        self.bg = background.float().contiguous().to(self.dev)
        # Voxels whose corner SDFs cannot reach the minimum alpha are left out of binning
        # (svr_voxel_activity); images and gradients are unchanged, the key count drops.
        self.cull = bool(cull)
        self.cam_scale = float(self.c2w[:, :3, 3].abs().max().item()) if self.V else 0.0
        self.bin_leaf = self.is_leaf.clone()
        if key_capacity is None:
            self.cap = 1 << 20
            self._alloc()
            self.preprocess()
            k = self.key_count()
            key_capacity = int(k * 1.25) + 4096
        self.cap = int(key_capacity)
        self._alloc()

    # -- buffers ---------------------------------------------------------------------------
    def _dims(self):
        return (self.V, self.P, self.n_pix, self.n_tiles, self.cap)

    def _alloc(self):
        dev, V, P = self.dev, self.V, self.P
        nbytes = int(self.lib.svr_batch_workspace_bytes(*self._dims()))
        self.ws = torch.empty(nbytes, dtype=torch.uint8, device=dev)
        st = self._stream()
        h = self._host
        _lib.check(self.lib.svr_batch_setup(_p(self.ws), *self._dims(), h["W"], h["H"], h["tx"], h["ty"], h["cx"],
                                            h["cy"], h["near"], _p(self.w2c), _p(self.c2w), _p(self.bg), st),
                   "svr_batch_setup")
        f = dict(dtype=torch.float32, device=dev)
        self.geos = torch.empty(P, 8, **f)
        self.rgbs_views = torch.zeros(V, P, 3, **f)
        self.color = torch.zeros(3 * self.n_pix, **f)
        self.T = torch.zeros(self.n_pix, **f)
        self.depth = torch.zeros(4 * self.n_pix, **f) if self.need_depth else None
        self.normal = torch.zeros(3 * self.n_pix, **f) if self.need_normal else None
        self.dvox = torch.zeros(P, 12, **f)
        self.drgb_views = torch.zeros(V, P, 3, **f)
        self.ds = torch.zeros(1, **f)
        self.dsh0 = torch.zeros(P, 3, **f)
        self.dshs = torch.zeros(P, self.M - 1, 3, **f)
        self.dgrid = None
        self.zero_T = torch.zeros(self.n_pix, **f)

    def _stream(self):
        return ctypes.c_void_p(torch.cuda.current_stream(self.dev).cuda_stream)

    def view_slice(self, buf: torch.Tensor, channels: int, v: int) -> torch.Tensor:
        """[C, H, W] view of view v inside a concatenated per-pixel buffer."""
        b = int(self.pix_base[v]) * channels
        return buf[b:b + channels * self.pix[v]].view(channels, self.Hs[v], self.Ws[v])

    # -- stages ----------------------------------------------------------------------------
    def preprocess(self, leaf: Optional[torch.Tensor] = None):
        leaf = self.is_leaf if leaf is None else leaf
        _lib.check(self.lib.svr_batch_preprocess(_p(self.ws), *self._dims(), _p(self.vox_center), _p(self.vox_size),
                                                 _p(leaf), _p(self.pair_order), self._stream()),
                   "svr_batch_preprocess")

    def state_views(self):
        """(ranges int32 [n_tiles, 2], vals int32 [K], n_contrib int32 [n_pixels], tile_last int32
        [n_tiles]) viewed inside the workspace (global tile ids, concatenated pixels)."""
        off = (ctypes.c_longlong * 4)()
        _lib.check(self.lib.svr_batch_state_offsets(*self._dims(), off), "svr_batch_state_offsets")
        K = min(self.key_count(), self.cap)
        i32 = lambda o, n: self.ws[o:o + 4 * n].view(torch.int32)
        return (i32(off[0], 2 * self.n_tiles).view(self.n_tiles, 2), i32(off[1], K), i32(off[2], self.n_pix),
                i32(off[3], self.n_tiles))

    def key_count(self) -> int:
        k = ctypes.c_longlong(0)
        _lib.check(self.lib.svr_batch_key_count(_p(self.ws), *self._dims(), ctypes.byref(k), self._stream()),
                   "svr_batch_key_count")
        return int(k.value)

    def forward(self, grid_pts: torch.Tensor, sh0: torch.Tensor, shs: torch.Tensor, s_val: torch.Tensor):
        """Render every view.  Returns the concatenated colour buffer (views [3,H,W] back to back)."""
        st = self._stream()
        _lib.check(self.lib.svr_gather_triinterp_geo_params(_p(self.vox_key), self.P, None, self.P,
                                                            _p(grid_pts), _p(self.geos), st), "gather")
        if self.cull:
            _lib.check(self.lib.svr_voxel_activity(self.P, _p(self.is_leaf), _p(self.geos), _p(s_val),
                                                   _p(self.vox_center), _p(self.vox_size),
                                                   ctypes.c_float(self.cam_scale), _p(self.bin_leaf), st),
                       "svr_voxel_activity")
            self.preprocess(self.bin_leaf)
        else:
            self.preprocess()
        _lib.check(self.lib.svr_batch_sh_forward(_p(self.ws), *self._dims(), _p(self.vox_center), _p(sh0), _p(shs),
                                                 self.deg, self.M, _p(self.rgbs_views), st), "sh_forward")
        # Every pixel of every tile is written by the compositor: colour and T need no clearing.
        if self.depth is not None:
            self.depth.zero_()
        if self.normal is not None:
            self.normal.zero_()
        self.s_val = s_val
        _lib.check(self.lib.svr_batch_rasterize(_p(self.ws), *self._dims(), _p(self.vox_center), _p(self.vox_size),
                                                _p(self.geos), _p(self.rgbs_views), _p(s_val), _p(self.pair_order),
                                                int(self.need_depth), int(self.need_normal),
                                                _p(self.color), _p(self.depth), _p(self.normal), _p(self.T), st),
                   "svr_batch_rasterize")
        return self.color

    def backward(self, dL_dcolor: torch.Tensor, n_grid_pts: int, dL_dT: Optional[torch.Tensor] = None,
                 dL_ddepth: Optional[torch.Tensor] = None, dL_dnormal: Optional[torch.Tensor] = None):
        """Gradients of all views at once: (dL/dgrid_pts [M,1], dL/dsh0, dL/dshs, dL/ds, subdiv [P])."""
        st = self._stream()
        if self.dgrid is None or self.dgrid.shape[0] != n_grid_pts:
            self.dgrid = torch.zeros(n_grid_pts, 1, dtype=torch.float32, device=self.dev)
        self.dvox.zero_()
        # drgb_views is zero here: it starts at zero and svr_batch_sh_backward clears what it reads.
        self.ds.zero_()
        self.dgrid.zero_()
        dT = self.zero_T if dL_dT is None else dL_dT
        _lib.check(self.lib.svr_batch_backward(_p(self.ws), *self._dims(), _p(self.vox_center), _p(self.vox_size),
                                               _p(self.geos), _p(self.rgbs_views), _p(self.s_val), _p(self.T),
                                               _p(dL_dcolor), _p(dL_ddepth), _p(dL_dnormal), _p(dT),
                                               int(self.need_depth), int(self.need_normal), _p(self.dvox),
                                               _p(self.drgb_views), _p(self.ds), st), "svr_batch_backward")
        _lib.check(self.lib.svr_batch_sh_backward(_p(self.ws), *self._dims(), _p(self.vox_center),
                                                  _p(self.drgb_views), self.deg, self.M, _p(self.dsh0),
                                                  _p(self.dshs), st), "sh_backward")
        dgeo = self.dvox[:, :8].contiguous()
        _lib.check(self.lib.svr_gather_triinterp_geo_params_bw(_p(self.vox_key), self.P, None, self.P, _p(dgeo),
                                                               _p(self.dgrid), st), "gather_bw")
        return self.dgrid, self.dsh0, self.dshs, self.ds, self.dvox[:, 11]


class BatchedTrainStep:
    """fwd + MSE-to-target + bwd of all views as one step, optionally captured as a HIP graph.
    Gradients land in `grid_grad`, `sh0_grad`, `shs_grad`, `s_grad`."""

    def __init__(self, rast: BatchedRasterizer, grid_pts, sh0, shs, log_s, targets: List[torch.Tensor]):
        self.r = rast
        self.grid_pts, self.sh0, self.shs, self.log_s = grid_pts, sh0, shs, log_s
        self.target = torch.cat([t.reshape(-1) for t in targets]).contiguous()
        scale = [torch.full((3 * p,), 2.0 / (3 * p), device=rast.dev) for p in rast.pix]
        self.scale = torch.cat(scale)
        self.s_val = torch.empty(1, dtype=torch.float32, device=rast.dev)
        self.dC = torch.empty_like(self.target)
        self.graph = None

    def _run(self):
        torch.exp(self.log_s.reshape(1) * 10.0, out=self.s_val)
        color = self.r.forward(self.grid_pts, self.sh0, self.shs, self.s_val)
        torch.mul(color - self.target, self.scale, out=self.dC)
        self.grid_grad, self.sh0_grad, self.shs_grad, self.s_grad, _ = self.r.backward(
            self.dC, self.grid_pts.shape[0])

    def capture(self, warmup: int = 2):
        s = torch.cuda.Stream(device=self.r.dev)
        s.wait_stream(torch.cuda.current_stream(self.r.dev))
        with torch.cuda.stream(s):
            for _ in range(warmup):
                self._run()
        torch.cuda.current_stream(self.r.dev).wait_stream(s)
        self.graph = torch.cuda.CUDAGraph()
        with torch.cuda.graph(self.graph):
            self._run()

    def __call__(self):
        if self.graph is not None:
            self.graph.replay()
        else:
            self._run()
