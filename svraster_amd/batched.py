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
                 key_capacity: Optional[int] = None, need_depth=False, need_normal=False, cull: bool = True,
                 tile_shape=(16, 16), fuse_gather_bw: bool = False, check_eager: bool = True,
                 overlap_streams: bool = True):
        self.dev = vox_center.device
        # overlap_streams: the SH colour pass runs on a second stream beside the key emission and
        # tile sort (and the SH transpose beside the corner scatter), joined before the compositor;
        # the second stream has the default priority.
        self.side = torch.cuda.Stream(device=self.dev) if overlap_streams else None
        # Eager forwards read the key peak back (one sync) and raise on overflow; graph replays
        # are checked by the caller with check_capacity().
        self.check_eager = bool(check_eager)
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
        # Corner-SDF gradients added to the grid points by the backward compositor's flush
        # instead of a [P,8] buffer scattered by a separate gather transpose.
        self.fuse_gather_bw = bool(fuse_gather_bw)
        # Rank order of every (voxel, octant) pair: fixed while the voxel set is.
        tmp = torch.empty(int(self.lib.svr_batch_pair_order_bytes(self.P)), dtype=torch.uint8, device=self.dev)
        self.pair_order = torch.empty(8 * self.P, dtype=torch.int32, device=self.dev)
        _lib.check(self.lib.svr_batch_pair_order(self.P, _p(self.octpath), _p(tmp), _p(self.pair_order),
                                                 self._stream()), "svr_batch_pair_order")
        del tmp
        # Rank-order slot of every pair (the inverse order): the preprocess stores key counts there.
        self.pair_slots = torch.empty(8 * self.P, dtype=torch.int32, device=self.dev)
        _lib.check(self.lib.svr_batch_pair_slots(self.P, _p(self.pair_order), _p(self.pair_slots), self._stream()),
                   "svr_batch_pair_slots")
        self.Ws = [c.image_width for c in cameras]
        self.Hs = [c.image_height for c in cameras]
        self.tile_w, self.tile_h = int(tile_shape[0]), int(tile_shape[1])
        tw, th = self.tile_w, self.tile_h
        self.tiles = [((w + tw - 1) // tw) * ((h + th - 1) // th) for w, h in zip(self.Ws, self.Hs)]
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
        # (the activity test of svr_batch_voxel_records); images and gradients are unchanged, the key
        # count drops.
        self.cull = bool(cull)
        self.cam_scale = float(self.c2w[:, :3, 3].abs().max().item()) if self.V else 0.0
        self.bin_leaf = self.is_leaf.clone()
        # Key count without the activity cull: a bound for every parameter state of this voxel
        # set and camera set (the cull only removes keys).
        self.cap = 1 << 20
        self._alloc()
        self.preprocess()
        self.keys_unculled = self.key_count()
        if key_capacity is None:
            key_capacity = int(self.keys_unculled * 1.25) + 4096
        elif int(key_capacity) < self.keys_unculled:
            raise ValueError(f"key_capacity {int(key_capacity)} is below the {self.keys_unculled} keys this voxel "
                             f"set emits for these cameras")
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
                                            h["cy"], h["near"], _p(self.w2c), _p(self.c2w), _p(self.bg),
                                            self.tile_w, self.tile_h, st),
                   "svr_batch_setup")
        f = dict(dtype=torch.float32, device=dev)
        # Per-voxel records of the compositors: centre, size, corner SDFs (gathered from the grid
        # points each forward), 1/size -- 64 bytes per voxel.
        self.vrec = torch.empty(P, 16, **f)
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

    @property
    def geos(self) -> torch.Tensor:
        """Corner SDFs [P,8] of the last forward (a copy out of the voxel records)."""
        return self.vrec[:, 4:12].contiguous()

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
                                                 _p(leaf), _p(self.pair_order), _p(self.pair_slots),
                                                 self._stream()),
                   "svr_batch_preprocess")

    def state_views(self):
        """(ranges int32 [n_tiles, 2], vals int32 [K], n_contrib int32 [n_pixels], tile_last int32
        [n_tiles]) viewed inside the workspace (global tile ids, concatenated pixels)."""
        off = self._state_offsets()
        K = min(self.key_count(), self.cap)
        i32 = lambda o, n: self.ws[o:o + 4 * n].view(torch.int32)
        return (i32(off[0], 2 * self.n_tiles).view(self.n_tiles, 2), i32(off[1], K), i32(off[2], self.n_pix),
                i32(off[3], self.n_tiles))

    def _state_offsets(self):
        off = (ctypes.c_longlong * 5)()
        _lib.check(self.lib.svr_batch_state_offsets(*self._dims(), off), "svr_batch_state_offsets")
        return off

    def contrib_ranges(self):
        """Per tile, the range [x, y) of its contributing entries in the list the forward wrote for
        the backward (int32 [n_tiles, 2], inside the workspace)."""
        o = self._state_offsets()[4]
        return self.ws[o:o + 8 * self.n_tiles].view(torch.int32).view(self.n_tiles, 2)

    def check_capacity(self, reset: bool = False) -> int:
        """Raise if any preprocess since set-up (or the last reset) needed more keys than the
        capacity: keys past it are not binned, so that step's images and gradients would be
        incomplete.  Returns the peak key count.  Synchronises the stream; call it after eager
        steps and after graph replays, not inside a capture."""
        k = ctypes.c_longlong(0)
        _lib.check(self.lib.svr_batch_key_peak(_p(self.ws), *self._dims(), int(bool(reset)), ctypes.byref(k),
                                               self._stream()), "svr_batch_key_peak")
        if k.value > self.cap:
            raise RuntimeError(f"batched binning needed {k.value} keys but its capacity is {self.cap}; "
                               f"rebuild the BatchedRasterizer with a larger key_capacity")
        return int(k.value)

    def key_count(self) -> int:
        k = ctypes.c_longlong(0)
        _lib.check(self.lib.svr_batch_key_count(_p(self.ws), *self._dims(), ctypes.byref(k), self._stream()),
                   "svr_batch_key_count")
        return int(k.value)

    def forward(self, grid_pts: torch.Tensor, sh0: torch.Tensor, shs: torch.Tensor, s_val: torch.Tensor,
                mse_target: Optional[torch.Tensor] = None, out_dL_dcolor: Optional[torch.Tensor] = None):
        """Render every view.  Returns the concatenated colour buffer (views [3,H,W] back to back).
        With mse_target (laid out like the colour buffer), the compositor also writes the gradient of
        every view's mean squared colour error into out_dL_dcolor."""
        if (mse_target is None) != (out_dL_dcolor is None):
            raise ValueError("mse_target and out_dL_dcolor go together")
        st = self._stream()
        _lib.check(self.lib.svr_batch_voxel_records(self.P, _p(self.vox_key), _p(grid_pts), _p(self.vox_center),
                                                    _p(self.vox_size), _p(self.is_leaf), _p(s_val),
                                                    ctypes.c_float(self.cam_scale), int(self.cull), _p(self.vrec),
                                                    _p(self.bin_leaf) if self.cull else None, st),
                   "svr_batch_voxel_records")
        self.preprocess(self.bin_leaf if self.cull else None)
        if self.check_eager and not torch.cuda.is_current_stream_capturing():
            self.check_capacity()
        def sh_forward():
            _lib.check(self.lib.svr_batch_sh_forward(_p(self.ws), *self._dims(), _p(self.vox_center), _p(sh0),
                                                     _p(shs), self.deg, self.M, _p(self.rgbs_views), self._stream()),
                       "sh_forward")

        if self.side is not None:
            main = torch.cuda.current_stream(self.dev)
            self.side.wait_stream(main)
            with torch.cuda.stream(self.side):
                sh_forward()
        else:
            sh_forward()
        _lib.check(self.lib.svr_batch_bin(_p(self.ws), *self._dims(), _p(self.pair_order), st), "svr_batch_bin")
        if self.side is not None:
            main.wait_stream(self.side)
        # Every pixel of every tile is written by the compositor: colour and T need no clearing.
        if self.depth is not None:
            self.depth.zero_()
        if self.normal is not None:
            self.normal.zero_()
        self.s_val = s_val
        _lib.check(self.lib.svr_batch_composite(_p(self.ws), *self._dims(), _p(self.vrec), _p(self.rgbs_views),
                                                _p(s_val), int(self.need_depth), int(self.need_normal),
                                                _p(self.color), _p(self.depth), _p(self.normal), _p(self.T),
                                                _p(mse_target), _p(out_dL_dcolor), st),
                   "svr_batch_composite")
        return self.color

    def backward(self, dL_dcolor: torch.Tensor, n_grid_pts: int, dL_dT: Optional[torch.Tensor] = None,
                 dL_ddepth: Optional[torch.Tensor] = None, dL_dnormal: Optional[torch.Tensor] = None,
                 sh_adam: Optional["ShAdamState"] = None):
        """Gradients of all views at once: (dL/dgrid_pts [M,1], dL/dsh0, dL/dshs, dL/ds, subdiv [P]).
        With `sh_adam`, the SH gradients are not returned (None): the SH transpose applies the
        Adam step to sh0/shs in place instead (svr_batch_sh_backward_adam)."""
        st = self._stream()
        if self.dgrid is None or self.dgrid.shape[0] != n_grid_pts:
            self.dgrid = torch.zeros(n_grid_pts, 1, dtype=torch.float32, device=self.dev)
        self.dvox.zero_()
        # drgb_views is zero here: it starts at zero and svr_batch_sh_backward clears what it reads.
        self.ds.zero_()
        self.dgrid.zero_()
        dT = dL_dT   # None: no transmittance cotangent (read as zero)
        fuse = self.fuse_gather_bw
        _lib.check(self.lib.svr_batch_backward(_p(self.ws), *self._dims(), _p(self.vrec), _p(self.rgbs_views),
                                               _p(self.s_val), _p(self.T),
                                               _p(dL_dcolor), _p(dL_ddepth), _p(dL_dnormal), _p(dT),
                                               int(self.need_depth), int(self.need_normal), _p(self.dvox),
                                               _p(self.drgb_views), _p(self.ds),
                                               _p(self.vox_key) if fuse else None, _p(self.dgrid) if fuse else None,
                                               st), "svr_batch_backward")
        def sh_backward():
            s2 = self._stream()
            if sh_adam is None:
                _lib.check(self.lib.svr_batch_sh_backward(_p(self.ws), *self._dims(), _p(self.vox_center),
                                                          _p(self.drgb_views), self.deg, self.M, _p(self.dsh0),
                                                          _p(self.dshs), s2), "sh_backward")
            else:
                a = sh_adam
                _lib.check(self.lib.svr_batch_sh_backward_adam(
                    _p(self.ws), *self._dims(), _p(self.vox_center), _p(self.drgb_views), self.deg, self.M,
                    _p(a.sh0), _p(a.m0), _p(a.v0), _p(a.table0), _p(a.shs), _p(a.ms), _p(a.vs), _p(a.tables),
                    a.n_table, _p(a.step), ctypes.c_float(a.b1), ctypes.c_float(a.b2), ctypes.c_float(a.eps),
                    int(a.sparse), s2), "sh_backward_adam")

        # The SH transpose (colour gradients) and the corner scatter (corner-SDF gradients) read
        # disjoint outputs of the reverse compositor: with a second stream they run side by side.
        if self.side is not None and not fuse:
            main = torch.cuda.current_stream(self.dev)
            self.side.wait_stream(main)
            with torch.cuda.stream(self.side):
                sh_backward()
        else:
            sh_backward()
        dsh0, dshs = (self.dsh0, self.dshs) if sh_adam is None else (None, None)
        if not fuse:
            _lib.check(self.lib.svr_batch_gather_bw(self.P, _p(self.vox_key), _p(self.dvox), _p(self.dgrid), st),
                       "svr_batch_gather_bw")
            if self.side is not None:
                main.wait_stream(self.side)
        return self.dgrid, dsh0, dshs, self.ds, self.dvox[:, 11]


class MseColorGrad:
    """dL/dcolor of the mean squared error per view, summed over views:
    (color - target) * 2 / (3 * pixels of the view), as one fused pass color * a - target * a with
    -target * a precomputed: a scalar when every view has the same size, else per-element scales."""

    def __init__(self, targets: List[torch.Tensor], pix: List[int], dev):
        self.target = torch.cat([t.reshape(-1) for t in targets]).contiguous()
        if len(set(pix)) == 1:
            self.alpha = 2.0 / (3 * pix[0])
            self.neg_ts = self.target * -self.alpha
            self.scale = None
        else:
            self.scale = torch.cat([torch.full((3 * p,), 2.0 / (3 * p), device=dev) for p in pix])
            self.neg_ts = -(self.target * self.scale)

    def __call__(self, color: torch.Tensor, out: torch.Tensor) -> None:
        if self.scale is None:
            torch.add(self.neg_ts, color, alpha=self.alpha, out=out)
        else:
            torch.addcmul(self.neg_ts, color, self.scale, out=out)


class BatchedTrainStep:
    """fwd + MSE-to-target + bwd of all views as one step, optionally captured as a HIP graph.
    Gradients land in `grid_grad`, `sh0_grad`, `shs_grad`, `s_grad`."""

    def __init__(self, rast: BatchedRasterizer, grid_pts, sh0, shs, log_s, targets: List[torch.Tensor],
                 fuse_loss: bool = True):
        self.r = rast
        # fuse_loss: the compositor writes dL/dcolor of the MSE as it writes the colour (one pass
        # fewer); otherwise MseColorGrad after the forward.
        self.fuse_loss = bool(fuse_loss)
        self.grid_pts, self.sh0, self.shs, self.log_s = grid_pts, sh0, shs, log_s
        self.loss_grad = MseColorGrad(targets, rast.pix, rast.dev)
        self.target = self.loss_grad.target
        self.s_val = torch.empty(1, dtype=torch.float32, device=rast.dev)
        self.dC = torch.empty_like(self.target)
        self.graph = None

    def _run(self):
        torch.exp(self.log_s.reshape(1) * 10.0, out=self.s_val)
        if self.fuse_loss:
            self.r.forward(self.grid_pts, self.sh0, self.shs, self.s_val, self.target, self.dC)
        else:
            color = self.r.forward(self.grid_pts, self.sh0, self.shs, self.s_val)
            self.loss_grad(color, self.dC)
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

    def check_capacity(self) -> int:
        return self.r.check_capacity()


# ----------------------------------------------------------------------------------------------
# Full training iteration: render + MSE + backward + grid Laplacian smoothness + sparse Adam
# ----------------------------------------------------------------------------------------------
def adam_bias_table(lr: float, betas, n: int, device) -> torch.Tensor:
    """Device copy of svr_adam_bias_table: float32 [n, 2] = {lr / (1 - b1^t), 1 / sqrt(1 - b2^t)}
    for t = 1..n, rounded from the same double expressions as the host-driven unbiased step."""
    lib = _lib.load()
    host = (ctypes.c_float * (2 * n))()
    _lib.check(lib.svr_adam_bias_table(ctypes.c_double(lr), ctypes.c_double(betas[0]), ctypes.c_double(betas[1]),
                                       n, host), "svr_adam_bias_table")
    return torch.tensor(list(host), dtype=torch.float32).view(n, 2).to(device)


def adam_table_length(betas) -> int:
    """Steps after which both bias corrections are exactly 1 in double (b^t below 2^-54), so the
    last table entry is exact for every later step; capped at 65536."""
    import math
    n = 1
    for b in betas:
        if 0.0 < b < 1.0:
            n = max(n, int(math.ceil(54.0 * math.log(2.0) / -math.log(b))) + 1)
    return min(n, 1 << 16)


class ShAdamState:
    """Adam moments and bias tables of sh0 / shs for the fused SH backward + Adam launch.
    sparse=False is the dense step (every element, zero gradients included)."""

    def __init__(self, sh0, shs, step: torch.Tensor, lr_sh0: float, lr_shs: float, betas, eps: float,
                 sparse: bool = False):
        self.sh0, self.shs, self.step = sh0, shs, step
        self.sparse = bool(sparse)
        self.m0, self.v0 = torch.zeros_like(sh0), torch.zeros_like(sh0)
        self.ms, self.vs = torch.zeros_like(shs), torch.zeros_like(shs)
        self.n_table = adam_table_length(betas)
        self.table0 = adam_bias_table(lr_sh0, betas, self.n_table, sh0.device)
        self.tables = adam_bias_table(lr_shs, betas, self.n_table, sh0.device)
        self.b1, self.b2, self.eps = float(betas[0]), float(betas[1]), float(eps)


def grid_table_level(octlevel, outside_level: int = 0) -> int:
    """Level of the valid-gradient grid table: min(max octree level - outside_level, 9)
    (train.py:961, 597, 637); the table has 2^level cells per axis over the inside region."""
    lv = min(int(octlevel.max().item()) - int(outside_level), 9)
    if lv < 0:
        raise ValueError(f"outside_level {outside_level} exceeds the deepest octree level")
    return lv


def trainer_smoothness_weights(iteration: int, G: int = 1, lambda_ge: float = 1e-8, ge_from: int = 0,
                               ge_until: int = 6000, ge_decay_every: int = 2000, ge_decay_mult: float = 0.25,
                               ge_drop_ratio: float = 0.0, lambda_ls: float = 1e-10, ls_from: int = 0,
                               ls_until: int = 8000, ls_decay_every: int = 2000, ls_decay_mult: float = 0.25,
                               ls_drop_ratio: float = 0.0):
    """(eikonal, Laplacian) weights the reference trainer passes at `iteration`
    (train.py:588-654; defaults from config.py:154-167): lambda * mult^min(iteration // every, 2)
    * G / K with K = int(G * (1 - drop_ratio)) inside [from, until], else 0."""
    def one(lam, lo, hi, every, mult, drop):
        if not lam or iteration < lo or iteration > hi:
            return 0.0
        K = int(G * (1.0 - float(drop)))
        return lam * mult ** min(iteration // every, 2) * (G / max(K, 1))
    return (one(lambda_ge, ge_from, ge_until, ge_decay_every, ge_decay_mult, ge_drop_ratio),
            one(lambda_ls, ls_from, ls_until, ls_decay_every, ls_decay_mult, ls_drop_ratio))


def grid_voxel_coords(vox_center, vox_size, scene_center, inside_extent: float, level: int):
    """Voxel min corners and sizes in cells of the 2^level grid over the inside region, rounded
    to integers (train.py:962-965): float32 [P,3] and [P]."""
    res = 1 << int(level)
    sc = torch.as_tensor(scene_center, device=vox_center.device).reshape(1, 3).float()
    coord = torch.round(((vox_center - vox_size.reshape(-1, 1) * 0.5) - (sc - inside_extent * 0.5))
                        / inside_extent * res).float().contiguous()
    size = torch.round(vox_size.reshape(-1) / inside_extent * res).float().contiguous()
    return coord, size


def grid_dense_index(keys: torch.Tensor, res: int) -> torch.Tensor:
    """Cell -> index into the sorted keys for all res^3 cells (-1 where absent)."""
    lib = _lib.load()
    dense = torch.empty(res ** 3, dtype=torch.int32, device=keys.device)
    _lib.check(lib.svr_grid_dense_index(_p(keys), int(keys.numel()), int(res), _p(dense),
                                        ctypes.c_void_p(torch.cuda.current_stream(keys.device).cuda_stream)),
               "svr_grid_dense_index")
    return dense


class TableOrderPlan:
    """Set-up of the table-order ("pull") form of the grid eikonal / Laplacian backward
    (svr_grid_smoothness_bw_table_order) for one valid-gradient table, given as its sorted keys
    and their voxels (the form train.py passes to grid_loss_bw).  The entries are put in the
    canonical table order -- voxel by voxel in voxel order, each voxel's cells x fastest -- and
    `ok` tells whether the table has the layout the form needs: unique cells, every voxel's
    entries exactly its integer cell box, power-of-two box sizes (octree voxels have them).
    Entry e of that order draws the random sample offset with index e."""

    def __init__(self, keys_sorted, g2v_sorted, coord, size, res: int, vox_key):
        self.lib = _lib.load()
        self.dev = keys_sorted.device
        self.res, self.G = int(res), int(keys_sorted.numel())
        self.coord, self.size, self.vox_key = coord, size, vox_key
        self.N = int(vox_key.shape[0])
        self.dense = grid_dense_index(keys_sorted, self.res)
        self.ok = self._canonical_order(keys_sorted.long(), g2v_sorted.long())
        if not self.ok:
            return
        n3 = self.res ** 3
        tk = self.table_keys.long()
        self.cell_entry = torch.full((n3,), -1, dtype=torch.int32, device=self.dev)
        self.cell_entry[tk] = torch.arange(self.G, dtype=torch.int32, device=self.dev)
        # compact numbering of the table's voxels, in table order
        tvox, table_tv = torch.unique_consecutive(self.table_vox, return_inverse=True)
        self.tvox = tvox.to(torch.int32).contiguous()
        self.table_tv = table_tv.to(torch.int32).contiguous()
        self.T = int(self.tvox.numel())
        self.cell_tv = torch.full((n3,), -1, dtype=torch.int32, device=self.dev)
        self.cell_tv[tk] = self.table_tv
        self.rec = torch.empty(12 * self.T, dtype=torch.float32, device=self.dev)

    def _canonical_order(self, k, v) -> bool:
        if self.G == 0 or int((self.dense >= 0).sum().item()) != self.G:
            return False                               # empty table or repeated cells
        res = self.res
        x, y, z = k % res, (k // res) % res, k // (res * res)
        lo = self.coord.long()[v]
        s = self.size.long()[v]
        if not bool(((s > 0) & ((s & (s - 1)) == 0)).all()):
            return False                               # sizes must be powers of two
        dx, dy, dz = x - lo[:, 0], y - lo[:, 1], z - lo[:, 2]
        if not bool(((dx >= 0) & (dx < s) & (dy >= 0) & (dy < s) & (dz >= 0) & (dz < s)).all()):
            return False
        local = (dz * s + dy) * s + dx                 # x fastest inside the voxel's box
        order = torch.argsort(v * (1 << 30) + local)
        v, local, s = v[order], local[order], s[order]
        same = v[1:] == v[:-1]
        # each voxel's entries run through local index 0 .. s^3 - 1, one each
        if int(local[0]) != 0 or not bool((local[1:][same] == local[:-1][same] + 1).all()):
            return False
        if not bool((local[1:][~same] == 0).all()):
            return False
        last = torch.cat([~same, torch.ones(1, dtype=torch.bool, device=self.dev)])
        if not bool((local[last] == s[last] ** 3 - 1).all()):
            return False
        self.table_keys = k[order].to(torch.int32).contiguous()
        self.table_vox = v.to(torch.int32).contiguous()
        return True

    def run(self, grid_pts, grid_grad, mask, terms: int, weights_dev, vox_size_inv: float, seed: int, step_dev,
            coef):
        """Accumulate the enabled terms' gradient (terms: 1 Laplacian, 2 eikonal, 3 both; weights
        in weights_dev = [eikonal, Laplacian]) into grid_grad; coef: float[3G] or float[7G]."""
        st = ctypes.c_void_p(torch.cuda.current_stream(self.dev).cuda_stream)
        _lib.check(self.lib.svr_grid_smoothness_bw_table_order(
            _p(grid_pts), grid_pts.numel(), _p(self.vox_key), self.N, _p(self.coord), _p(self.size), self.res,
            _p(mask), _p(self.table_keys), _p(self.table_tv), self.G, _p(self.cell_entry), _p(self.cell_tv),
            _p(self.tvox), self.T, int(terms), _p(weights_dev), ctypes.c_float(vox_size_inv),
            ctypes.c_uint(int(seed) & 0xFFFFFFFF), _p(step_dev), _p(self.rec), _p(coef), _p(grid_grad), st),
            "svr_grid_smoothness_bw_table_order")


class GridSmoothness:
    """Laplacian smoothness (and, with eikonal_weight != 0, grid eikonal) backward on the
    valid-gradient grid table, as the reference trainer calls them each iteration
    (train.py:588-654; table rebuilt after topology changes, train.py:959-972): grid resolution
    2^min(max level - outside_level, 9), voxel coordinates and sizes in grid cells rounded to
    integers, keys sorted (octree_utils.py:293-315), weight * G / A, every key active.

    The reference draws the active list as randperm(G)[:A] with A = G (ls_drop_ratio 0): every
    key once, in random order.  Here the keys are taken in table order (voxel by voxel), which
    only changes which random sample offset goes with which key and the order of the float sums;
    the offsets come from a hash of (seed, device step counter, position), so every replay of a
    captured step draws new ones.  When the table has its built layout (each voxel's entries its
    cell box, cells unique) the loss runs in the two-pass table-order form
    (svr_grid_laplacian_bw_table_order) that sums each voxel's contributions before touching the
    gradient; otherwise, or with pull=False, one thread per key scatters its 56 weighted terms
    with atomics as the reference kernel does, with cells found through a dense cell -> key map
    (svr_grid_dense_index) instead of a binary search per sample."""

    def __init__(self, vox_center, vox_size, octlevel, is_leaf, vox_key, scene_center, inside_extent: float,
                 weight: float, seed: int = 0, pull: bool = True, outside_level: int = 0,
                 eikonal_weight: float = 0.0):
        from . import _C
        self.lib = _lib.load()
        self.dev = vox_center.device
        self.outside_level = int(outside_level)
        self.max_level = grid_table_level(octlevel, self.outside_level)
        self.res = 1 << self.max_level
        mask, keys, g2v = _C.valid_gradient_table(vox_center, vox_size, scene_center, inside_extent,
                                                  self.max_level, is_leaf.reshape(-1).bool())
        self.keys, order = keys.sort(stable=True)
        self.g2v = g2v[order].contiguous()
        del order, keys, g2v
        self.mask = mask
        self.G = int(self.keys.numel())
        self.coord, self.size = grid_voxel_coords(vox_center, vox_size, scene_center, inside_extent,
                                                  self.max_level)
        self.vox_key = vox_key.contiguous()
        self.N = int(vox_center.shape[0])
        self.plan = TableOrderPlan(self.keys, self.g2v, self.coord, self.size, self.res, self.vox_key)
        self.dense = self.plan.dense
        self.pull = bool(pull) and self.plan.ok
        # Active list = every key once, in the canonical table order when the table has its built
        # layout, as indices into the sorted keys.
        if self.plan.ok:
            self.table_keys = self.plan.table_keys
            self.active = self.dense[self.table_keys.long()].contiguous()
        else:
            self.active = torch.arange(self.G, dtype=torch.int32, device=self.dev)
        # Term weights (each already * G / A with A = G): eikonal, Laplacian.  The eikonal term is
        # on when its weight is non-zero at construction; the kernels read the weights from device
        # memory, so set_weights() changes them for later replays of a captured step.
        self.eikonal = float(eikonal_weight) != 0.0
        self.weights = torch.tensor([float(eikonal_weight), float(weight)], dtype=torch.float32, device=self.dev)
        self.weight, self.eikonal_weight = float(weight), float(eikonal_weight)
        if self.pull:
            self.coef = torch.empty((7 if self.eikonal else 3) * self.G, dtype=torch.float32, device=self.dev)
        self.vox_size_inv = float(self.res / inside_extent)
        self.seed = int(seed) & 0xFFFFFFFF

    def _stream(self):
        return ctypes.c_void_p(torch.cuda.current_stream(self.dev).cuda_stream)

    def set_weights(self, eikonal: Optional[float] = None, laplacian: Optional[float] = None):
        """New term weights for the following calls / replays (a host-to-device copy; not inside a
        capture).  The eikonal weight can change only if the term was enabled at construction."""
        if eikonal is not None:
            if float(eikonal) != 0.0 and not self.eikonal:
                raise ValueError("the eikonal term was not enabled at construction (eikonal_weight = 0)")
            self.eikonal_weight = float(eikonal)
        if laplacian is not None:
            self.weight = float(laplacian)
        self.weights.copy_(torch.tensor([self.eikonal_weight, self.weight], dtype=torch.float32))

    def __call__(self, grid_pts: torch.Tensor, grid_grad: torch.Tensor, step: torch.Tensor):
        if self.pull:
            self.plan.run(grid_pts, grid_grad, self.mask, 3 if self.eikonal else 1, self.weights,
                          self.vox_size_inv, self.seed, step, self.coef)
        else:
            terms = ((1, self.weight),) + (((0, self.eikonal_weight),) if self.eikonal else ())
            for lap, wt in terms:
                _lib.check(self.lib.svr_grid_smoothness_bw_step(
                    lap, _p(grid_pts), grid_pts.numel(), _p(self.vox_key), self.N, _p(self.coord), _p(self.size),
                    self.res, _p(self.mask), _p(self.keys), _p(self.g2v), self.G, _p(self.dense), _p(self.active),
                    self.G, ctypes.c_float(wt), ctypes.c_float(self.vox_size_inv), ctypes.c_uint(self.seed),
                    _p(step), _p(grid_grad), self._stream()), "svr_grid_smoothness_bw_step")


class BatchedTrainIteration:
    """One training iteration over a batch of views in the reference trainer's order
    (train.py:324-740): render every view -> MSE to the targets -> backward -> grid Laplacian
    smoothness into the corner-SDF gradient -> unbiased Adam on the corner SDFs and SH.  The Adam
    step is dense by default, as the reference trainer builds SparseAdam with sparse=False
    (src/sparse_voxel_gears/optimizer.py:24): every element steps every iteration and the moments
    of untouched elements keep decaying.  sparse=True leaves elements with an exactly-zero
    gradient untouched.  The step counter lives on the device, so `capture()` records the whole
    iteration once and each replay is the next iteration.

    log_s is not updated (its learning rate is 0 in the reference's optimizer groups,
    optimizer.py:16-22); the SH Adam step runs inside the SH transpose launch."""

    def __init__(self, rast: BatchedRasterizer, grid_pts, sh0, shs, log_s, targets: List[torch.Tensor],
                 smooth: Optional[GridSmoothness], lr: float = 1e-2, betas=(0.9, 0.99), eps: float = 1e-15,
                 lr_sh0: Optional[float] = None, lr_shs: Optional[float] = None, sparse: bool = False,
                 fuse_loss: bool = True, overlap_smooth: bool = True):
        self.r = rast
        # overlap_smooth: the grid smoothness depends on the corner SDFs only, so it runs on a stream
        # of its own (default priority) into a gradient buffer of its own while the views render and
        # back-propagate, and is added to the render gradient before the Adam step.
        self.smooth_stream = torch.cuda.Stream(device=rast.dev) if (overlap_smooth and smooth is not None) else None
        self.sgrad = torch.zeros_like(grid_pts) if self.smooth_stream is not None else None
        self.sparse = bool(sparse)
        # fuse_loss: the render stage's compositor writes dL/dcolor of the MSE (the loss stage is
        # then empty); otherwise the loss stage runs MseColorGrad.
        self.fuse_loss = bool(fuse_loss)
        self.dev = rast.dev
        self.grid_pts, self.sh0, self.shs, self.log_s = grid_pts, sh0, shs, log_s
        self.loss_grad = MseColorGrad(targets, rast.pix, self.dev)
        self.target = self.loss_grad.target
        self.s_val = torch.empty(1, dtype=torch.float32, device=self.dev)
        self.dC = torch.empty_like(self.target)
        self.step = torch.zeros(1, dtype=torch.int32, device=self.dev)
        self.smooth = smooth
        self.betas, self.eps = (float(betas[0]), float(betas[1])), float(eps)
        self.sh_adam = ShAdamState(sh0, shs, self.step, lr if lr_sh0 is None else lr_sh0,
                                   lr if lr_shs is None else lr_shs, self.betas, self.eps, sparse=self.sparse)
        self.grid_m, self.grid_v = torch.zeros_like(grid_pts), torch.zeros_like(grid_pts)
        self.n_table = self.sh_adam.n_table
        self.grid_table = adam_bias_table(lr, self.betas, self.n_table, self.dev)
        self.graph = None
        self.lib = _lib.load()

    def _stream(self):
        return ctypes.c_void_p(torch.cuda.current_stream(self.dev).cuda_stream)

    def stage_render(self):
        self.step.add_(1)
        torch.exp(self.log_s.reshape(1) * 10.0, out=self.s_val)
        if self.fuse_loss:
            self.color = self.r.forward(self.grid_pts, self.sh0, self.shs, self.s_val, self.target, self.dC)
        else:
            self.color = self.r.forward(self.grid_pts, self.sh0, self.shs, self.s_val)

    def stage_loss(self):
        if not self.fuse_loss:
            self.loss_grad(self.color, self.dC)

    def stage_backward(self):
        self.grid_grad, _, _, self.s_grad, _ = self.r.backward(self.dC, self.grid_pts.shape[0],
                                                                sh_adam=self.sh_adam)

    def stage_smooth(self):
        if self.smooth is not None:
            self.smooth(self.grid_pts, self.grid_grad, self.step)

    def stage_adam(self):
        _lib.check(self.lib.svr_sparse_adam_step_dev(
            int(self.sparse), _p(self.grid_pts), _p(self.grid_grad), _p(self.grid_m), _p(self.grid_v), self.grid_pts.numel(),
            _p(self.step), _p(self.grid_table), self.n_table, ctypes.c_float(self.betas[0]),
            ctypes.c_float(self.betas[1]), ctypes.c_float(self.eps), self._stream()), "svr_sparse_adam_step_dev")

    STAGES = ("render", "loss", "backward", "smooth", "adam")

    def _run(self):
        if self.smooth_stream is None:
            for name in self.STAGES:
                getattr(self, "stage_" + name)()
            return
        # render prelude (step counter, s), then the smoothness forks off before the forward
        main = torch.cuda.current_stream(self.dev)
        self.step.add_(1)
        torch.exp(self.log_s.reshape(1) * 10.0, out=self.s_val)
        self.smooth_stream.wait_stream(main)
        with torch.cuda.stream(self.smooth_stream):
            self.sgrad.zero_()
            self.smooth(self.grid_pts, self.sgrad, self.step)
        if self.fuse_loss:
            self.color = self.r.forward(self.grid_pts, self.sh0, self.shs, self.s_val, self.target, self.dC)
        else:
            self.color = self.r.forward(self.grid_pts, self.sh0, self.shs, self.s_val)
        self.stage_loss()
        self.stage_backward()
        main.wait_stream(self.smooth_stream)
        self.grid_grad.add_(self.sgrad)
        self.stage_adam()

    def stage_times(self):
        """Milliseconds per stage of one eager iteration (stages one after another, no overlap),
        from HIP events on the current stream."""
        ev = [torch.cuda.Event(enable_timing=True) for _ in range(len(self.STAGES) + 1)]
        ev[0].record()
        for i, name in enumerate(self.STAGES):
            getattr(self, "stage_" + name)()
            ev[i + 1].record()
        torch.cuda.synchronize(self.dev)
        return {name: ev[i].elapsed_time(ev[i + 1]) for i, name in enumerate(self.STAGES)}

    def capture(self, warmup: int = 1):
        s = torch.cuda.Stream(device=self.dev)
        s.wait_stream(torch.cuda.current_stream(self.dev))
        with torch.cuda.stream(s):
            for _ in range(warmup):
                self._run()
        torch.cuda.current_stream(self.dev).wait_stream(s)
        self.graph = torch.cuda.CUDAGraph()
        with torch.cuda.graph(self.graph):
            self._run()

    def __call__(self):
        if self.graph is not None:
            self.graph.replay()
        else:
            self._run()

    def check_capacity(self) -> int:
        return self.r.check_capacity()
