"""torch FP64 restatement of the SDF compositor (test infrastructure only).

Forward semantics follow cuda/src/forward.cu:76-562 (renderCUDA with n_samp = 3, SDF mode):
  * ray: compute_ray_d (auxiliary.h:87-95), rotated by c2w and divided by |cam_rd|
    (forward.cu:124-133); octant from the signs of rd (auxiliary.h:76-79);
  * per (pixel, entry): bbox membership (auxiliary.h:113-120), octant match, ray/box slab test
    (auxiliary.h:251-264), a <= b;
  * SDF at the entry point and 3 accumulated steps (forward.cu:255-294), logistic CDF
    phi = 1/(1+exp(-s sdf)), alpha = max((phi_entry - phi_exit)/(phi_entry + 1e-6), 0),
    skipped when alpha < 1e-5 (forward.cu:302-305);
  * w = T alpha, C += w rgb, depth D += T dval with the 3-segment dval (forward.cu:345-370),
    normal N += w lin_n / |lin_n| (forward.cu:420-434), T *= 1 - alpha, stop when T < 1e-4
    (forward.cu:437-438);
  * n_contrib = 1 + index of the last contributing entry (forward.cu:230-241, 441, 447);
  * median depth: last contributor with T > 0.5 then a 16-step refinement (forward.cu:372-377,
    450-521); outputs color = C + T bg, depth ch0 = D/|cam_rd|, ch2 = Dmed/|cam_rd|,
    normal, T (forward.cu:526-561).

Backward: torch autograd through the forward above IS the oracle for
cuda/src/backward.cu:38-851, with the reference's two deliberate differences from the exact
derivative restated explicitly:
  1. dL/ds flows only through the main alpha; the three depth sub-alphas use a detached s
     (backward.cu:338-388 vs 662-670);
  2. the colour-concentration term injects 2 (lambda/(W H)) w (rgb - gt) into dL/drgb with w
     held constant (backward.cu:824-831).
subdiv_p_bw = sum over contributing pairs of |dL/dalpha * alpha| (backward.cu:843), read from the
gradient of the per-pair alpha tensor.

Pinning (tests/test_oracle_raster.py): central finite differences in FP64 of the same loss,
closed forms (constant SDF -> alpha = 0; single voxel with linear SDF along the ray; sum of
weights + T = 1; sub-alpha telescoping; the median depth of one or two voxels sharing a linear
SDF, where the refinement's transmittance telescopes and its stopping step follows from the
logit of the crossing), and the scalar per-pixel loop `render_scalar` which restates the
reference's batch/j_lst/contributor control flow literally.
"""
from __future__ import annotations

from dataclasses import dataclass
from typing import Dict, Optional, Sequence

import numpy as np
import torch

BLOCK = 16
F = np.float32
EPS_PHI = float(F(1e-6))
MIN_ALPHA = float(F(1e-5))
EARLY_STOP_T = float(F(1e-4))
THIRD = float(F(0.3333333))


def f32v(x) -> float:
    """Value of a Python float once passed through a C float parameter."""
    return float(F(x))


@dataclass
class RasterCamera:
    W: int
    H: int
    tanfovx: float
    tanfovy: float
    cx: float
    cy: float
    c2w: torch.Tensor     # [4,4]


def camera_from(cam) -> RasterCamera:
    return RasterCamera(cam.image_width, cam.image_height, f32v(cam.tanfovx), f32v(cam.tanfovy),
                        f32v(cam.cx), f32v(cam.cy), cam.c2w.detach().cpu().double())


def pixel_rays(cam: RasterCamera, px: torch.Tensor, py: torch.Tensor, dtype=torch.float64):
    """Ray origin, unit direction, |cam_rd| for integer pixel coords (float64 by default)."""
    dx = (px.to(dtype) + 0.5 - cam.cx) * 2.0 * cam.tanfovx / cam.W
    dy = (py.to(dtype) + 0.5 - cam.cy) * 2.0 * cam.tanfovy / cam.H
    cam_rd = torch.stack([dx, dy, torch.ones_like(dx)], -1)
    rd_norm = cam_rd.norm(dim=-1)
    R = cam.c2w[:3, :3].to(dtype)
    rd = (cam_rd @ R.T) / rd_norm[..., None]
    ro = cam.c2w[:3, 3].to(dtype)
    return ro, rd, rd_norm


def ray_quadrant(rd: torch.Tensor) -> torch.Tensor:
    return ((rd[..., 0] < 0).long() << 2) | ((rd[..., 1] < 0).long() << 1) | (rd[..., 2] < 0).long()


def trilinear_weights(q: torch.Tensor) -> torch.Tensor:
    """[..., 3] local coords -> [..., 8] weights, corner index = x<<2 | y<<1 | z."""
    x, y, z = q[..., 0], q[..., 1], q[..., 2]
    wx = (1 - x, x)
    wy = (1 - y, y)
    wz = (1 - z, z)
    return torch.stack([wx[i >> 2 & 1] * wy[i >> 1 & 1] * wz[i & 1] for i in range(8)], -1)


def ray_aabb(c, l, ro, rd_inv):
    r = 0.5 * l
    d = c - ro
    c0 = (d - r[..., None]) * rd_inv
    c1 = (d + r[..., None]) * rd_inv
    lo = torch.minimum(c0, c1)
    hi = torch.maximum(c0, c1)
    a = lo.max(-1).values
    b = hi.min(-1).values
    return a, b


def lin_normal(geo: torch.Tensor) -> torch.Tensor:
    g = geo
    nx = (g[..., 4] + g[..., 5] + g[..., 6] + g[..., 7]) - (g[..., 0] + g[..., 1] + g[..., 2] + g[..., 3])
    ny = (g[..., 2] + g[..., 3] + g[..., 6] + g[..., 7]) - (g[..., 0] + g[..., 1] + g[..., 4] + g[..., 5])
    nz = (g[..., 1] + g[..., 3] + g[..., 5] + g[..., 7]) - (g[..., 0] + g[..., 2] + g[..., 4] + g[..., 6])
    return torch.stack([nx, ny, nz], -1)


def sdf_samples(geo, c, l, ro, rd, a, b):
    """SDF at the entry point and the three accumulated steps: [..., 4]; plus the
    entry-point local coordinate and step."""
    l_inv = 1.0 / l
    p_entry = ro + a[..., None] * rd
    q0 = (p_entry - (c - 0.5 * l[..., None])) * l_inv[..., None]
    qs = ((b - a)[..., None] * rd / 3.0) * l_inv[..., None]
    sdfs = []
    for k in range(4):
        w = trilinear_weights(q0 + qs * k)
        sdfs.append((w * geo).sum(-1))
    return torch.stack(sdfs, -1)


def alphas_from_sdf(sdf4: torch.Tensor, s: torch.Tensor):
    """Main NeuS alpha (grad to s) and the three depth sub-alphas (s detached)."""
    phi0 = torch.sigmoid(s * sdf4[..., 0])
    phi3 = torch.sigmoid(s * sdf4[..., 3])
    raw = (phi0 - phi3) / (phi0 + EPS_PHI)
    alpha = torch.where(raw > 0, raw, torch.zeros_like(raw))
    phid = torch.sigmoid(s.detach() * sdf4)
    subs = []
    for k in range(3):
        r = (phid[..., k] - phid[..., k + 1]) / (phid[..., k] + EPS_PHI)
        subs.append(torch.where(r > 0, r, torch.zeros_like(r)))
    return alpha, torch.stack(subs, -1)


def depth_value(sub: torch.Tensor, a: torch.Tensor, b: torch.Tensor):
    step = THIRD * (b - a)
    t0 = a + 0.5 * step
    t1 = a + 1.5 * step
    t2 = a + 2.5 * step
    a0, a1, a2 = sub[..., 0], sub[..., 1], sub[..., 2]
    return a0 * t0 + (1 - a0) * a1 * t1 + (1 - a0) * (1 - a1) * a2 * t2


@dataclass
class OracleOut:
    color: torch.Tensor       # [3,h,w]
    depth: torch.Tensor       # [4,h,w]
    normal: torch.Tensor      # [3,h,w]
    T: torch.Tensor           # [1,h,w]
    n_contrib: torch.Tensor   # [h,w] int64
    max_w: torch.Tensor       # [P]
    reg_loss: torch.Tensor    # scalar: colour-concentration surrogate (0 when off)
    pair_alpha: list          # per tile: (alpha [npix, n] with grad, pc mask, vox ids, dalpha/ds)
    window: tuple             # (x0, y0, x1, y1)
    # [h,w]: how far the median depth (ch2) is from a different decision, in units of an FP32
    # evaluation's error: the smallest of |T - 0.5| / 1e-4 over the median-voxel choice and the
    # refinement steps, and |alpha - 1e-5| / 1e-6 over the hit entries entered with T > 0.5
    # (whether they count as contributors).  Below 1, FP32 rounding may flip a decision and move
    # ch2 by a refinement step or to another voxel.
    dmed_margin: Optional[torch.Tensor] = None


def render(cam: RasterCamera, bg, bin_vals: np.ndarray, bin_ranges: np.ndarray,
           bbox: np.ndarray, centers, lengths, geos, rgbs, s,
           need_depth=True, need_normal=True, window=None,
           lambda_R_concen: float = 0.0, gt_color: Optional[torch.Tensor] = None,
           dtype=torch.float64) -> OracleOut:
    """Render the pixels of `window` = (x0, y0, x1, y1) (default: whole image).

    `geos` [P,8], `rgbs` [P,3], `s` (0-dim) are tensors of `dtype` that may require grad.  The
    oracle proper runs in float64; dtype=torch.float32 runs the same restatement in single
    precision (the FP32 CPU baseline, and a measure of what FP32 arithmetic alone changes).
    The median depth refinement always runs in float64."""
    W, H = cam.W, cam.H
    if window is None:
        window = (0, 0, W, H)
    x0, y0, x1, y1 = window
    h, w = y1 - y0, x1 - x0
    dev = torch.device("cpu")
    dt = dtype
    bg = torch.as_tensor(bg, dtype=dt)
    centers = torch.as_tensor(centers).to(dt)
    lengths = torch.as_tensor(lengths).to(dt).reshape(-1)
    P = centers.shape[0]
    tiles_x = (W + BLOCK - 1) // BLOCK

    color = torch.zeros(3, h, w, dtype=dt)
    depth = torch.zeros(4, h, w, dtype=dt)
    normal = torch.zeros(3, h, w, dtype=dt)
    Tout = torch.ones(1, h, w, dtype=dt)
    n_contrib = torch.zeros(h, w, dtype=torch.long)
    dmed_margin = torch.full((h, w), float("inf"), dtype=dt)
    max_w = torch.zeros(P, dtype=dt)
    reg = torch.zeros((), dtype=dt)
    pair_alpha = []
    col_parts, dep_parts, nrm_parts, T_parts = [], [], [], []

    bx0, by0 = x0 // BLOCK, y0 // BLOCK
    bx1, by1 = (x1 + BLOCK - 1) // BLOCK, (y1 + BLOCK - 1) // BLOCK
    for ty in range(by0, by1):
        for tx in range(bx0, bx1):
            tile = ty * tiles_x + tx
            gx = torch.arange(max(tx * BLOCK, x0), min((tx + 1) * BLOCK, x1, W))
            gy = torch.arange(max(ty * BLOCK, y0), min((ty + 1) * BLOCK, y1, H))
            if gx.numel() == 0 or gy.numel() == 0:
                continue
            PY, PX = torch.meshgrid(gy, gx, indexing="ij")
            px, py = PX.reshape(-1), PY.reshape(-1)
            npix = px.numel()
            ro, rd, rd_norm = pixel_rays(cam, px, py, dt)
            pq = ray_quadrant(rd)
            lo, hi = int(bin_ranges[tile, 0]), int(bin_ranges[tile, 1])
            vals = bin_vals[lo:hi].astype(np.int64)
            n = vals.shape[0]
            Cpix = torch.zeros(npix, 3, dtype=dt)
            Dpix = torch.zeros(npix, dtype=dt)
            Npix = torch.zeros(npix, 3, dtype=dt)
            Tfin = torch.ones(npix, dtype=dt)
            ncon = torch.zeros(npix, dtype=torch.long)
            dmed = torch.zeros(npix, dtype=dt)
            margin = torch.full((npix,), float("inf"), dtype=dt)
            if n > 0:
                vid = torch.from_numpy(vals & ((1 << 29) - 1))
                vq = torch.from_numpy(vals >> 29)
                bb = torch.from_numpy(bbox[vid.numpy()].astype(np.int64))
                xmin, ymin = bb[:, 0] >> 16, bb[:, 0] & 0xFFFF
                xmax, ymax = bb[:, 1] >> 16, bb[:, 1] & 0xFFFF
                inb = ((px[:, None] >= xmin[None]) & (py[:, None] >= ymin[None])
                       & (px[:, None] <= xmax[None]) & (py[:, None] <= ymax[None]))
                quad_ok = pq[:, None] == vq[None]
                c = centers[vid][None].expand(npix, n, 3)
                l = lengths[vid][None].expand(npix, n)
                rd_inv = 1.0 / rd
                a, b = ray_aabb(c, l, ro, rd_inv[:, None, :])
                hit = inb & quad_ok & (a <= b)
                a = torch.where(hit, a, torch.zeros_like(a))
                b = torch.where(hit, b, torch.zeros_like(b))
                geo = geos[vid][None].expand(npix, n, 8)
                sdf4 = sdf_samples(geo, c, l, ro, rd[:, None, :], a, b)
                alpha, sub = alphas_from_sdf(sdf4, s)
                contrib = hit & (alpha.detach() >= MIN_ALPHA)
                a_eff = torch.where(contrib, alpha, torch.zeros_like(alpha))
                one_m = 1.0 - a_eff
                T_incl = torch.cumprod(one_m, dim=1)
                T_excl = torch.cat([torch.ones(npix, 1, dtype=dt), T_incl[:, :-1]], 1)
                processed = T_excl.detach() >= EARLY_STOP_T
                pc = contrib & processed
                pcf = pc.to(dt)
                wgt = T_excl * a_eff * pcf
                Cpix = (wgt[..., None] * rgbs[vid][None]).sum(1)
                Tfin = torch.prod(1.0 - a_eff * pcf, dim=1)
                if need_depth:
                    dval = depth_value(sub, a, b)
                    Dpix = (T_excl * pcf * dval).sum(1)
                if need_normal:
                    ln = lin_normal(geos[vid])
                    nn = ln * torch.rsqrt((ln * ln).sum(-1, keepdim=True) + 1e-15)
                    Npix = (wgt[..., None] * nn[None]).sum(1)
                idx = torch.arange(n)[None].expand(npix, n)
                last = torch.where(pc, idx + 1, torch.zeros_like(idx)).max(1).values
                ncon = last
                wd = wgt.detach()
                mw = torch.where(pc, wd, torch.zeros_like(wd)).max(0).values
                max_w.scatter_reduce_(0, vid, mw, reduce="amax")
                if lambda_R_concen > 0 and gt_color is not None:
                    gt = gt_color.to(dt)[:, py, px].T              # [npix,3]
                    diff = rgbs[vid][None] - gt[:, None, :]
                    reg = reg + lambda_R_concen / (W * H) * (wd * (diff * diff).sum(-1)).sum()
                if need_depth:
                    dmed, margin = median_depth(T_excl.detach().double(), pc, vid, centers.double(),
                                                lengths.double(), geos.detach().double(), float(s.detach()),
                                                ro.double(), rd.double(), return_margin=True)
                    dmed = dmed.to(dt)
                    margin = margin.to(dt)
                    with torch.no_grad():
                        Te = T_excl.detach()
                        m_sel = torch.where(pc, (Te - 0.5).abs(), torch.full_like(Te, float("inf"))).min(1).values
                        m_alpha = torch.where(hit & (Te > 0.5), (alpha.detach() - MIN_ALPHA).abs() / 1e-6,
                                              torch.full_like(Te, float("inf"))).min(1).values
                        margin = torch.minimum(torch.minimum(margin, m_sel) / 1e-4, m_alpha)
                with torch.no_grad():
                    p0 = torch.sigmoid(s * sdf4[..., 0])
                    p3 = torch.sigmoid(s * sdf4[..., 3])
                    dads = ((p0 * (1 - p0) * sdf4[..., 0] * (1 - alpha) - p3 * (1 - p3) * sdf4[..., 3])
                            / (p0 + EPS_PHI))
                pair_alpha.append((alpha, pc, vid, dads))
            lx, ly = px - x0, py - y0
            col_parts.append((ly, lx, Cpix + Tfin[:, None] * bg[None]))
            T_parts.append((ly, lx, Tfin))
            n_contrib[ly, lx] = ncon
            dmed_margin[ly, lx] = margin
            if need_depth:
                dep_parts.append((ly, lx, Dpix / rd_norm, dmed / rd_norm))
            if need_normal:
                nrm_parts.append((ly, lx, Npix))

    color = _assemble(color, [(ly, lx, v.T) for ly, lx, v in col_parts])
    Tout = _assemble(Tout, [(ly, lx, v[None]) for ly, lx, v in T_parts])
    if need_depth:
        depth = _assemble(depth, [(ly, lx, torch.stack([d0, torch.zeros_like(d0), d2.to(dt),
                                                         torch.zeros_like(d0)], 0))
                                  for ly, lx, d0, d2 in dep_parts])
    if need_normal:
        normal = _assemble(normal, [(ly, lx, v.T) for ly, lx, v in nrm_parts])
    return OracleOut(color, depth, normal, Tout, n_contrib, max_w, reg, pair_alpha, window,
                     dmed_margin if need_depth else None)


def _assemble(base: torch.Tensor, parts):
    """Differentiable scatter of per-tile pixel values into an image [C,h,w]."""
    C, h, w = base.shape
    flat = base.reshape(C, h * w)
    if not parts:
        return base
    idx = torch.cat([ly * w + lx for ly, lx, _ in parts])
    val = torch.cat([v.reshape(C, -1) for _, _, v in parts], 1)
    keep = torch.ones(h * w, dtype=torch.bool)
    keep[idx] = False
    out = flat * keep.to(base.dtype)[None]
    out = out.index_add(1, idx, val)
    return out.reshape(C, h, w)


@torch.no_grad()
def median_depth(T_excl, pc, vid, centers, lengths, geos, s, ro, rd, return_margin=False):
    """16-step refinement inside the last contributor with T > 0.5 (forward.cu:450-521).
    With return_margin, also the smallest |T - 0.5| of the refinement's comparisons per pixel."""
    npix, n = pc.shape
    cand = pc & (T_excl > 0.5)
    idx = torch.arange(n)[None].expand(npix, n)
    last = torch.where(cand, idx, torch.full_like(idx, -1)).max(1).values
    out = torch.zeros(npix, dtype=torch.float64)
    margin = torch.full((npix,), float("inf"), dtype=torch.float64)
    for p in range(npix):
        j = int(last[p])
        if j < 0:
            continue
        v = int(vid[j])
        Tm = float(T_excl[p, j])
        c, l = centers[v], float(lengths[v])
        g = geos[v].double()
        rdp = rd[p]
        a, b = ray_aabb(c, torch.tensor(l, dtype=torch.float64), ro, 1.0 / rdp)
        a, b = float(a), float(b)
        step = (b - a) / 16.0
        dm = a - 0.5 * step
        q = ((ro + a * rdp) - (c - 0.5 * l)) / l
        qstep = step * rdp / l
        phi_prev = 1.0 / (1.0 + np.exp(-s * float((trilinear_weights(q) * g).sum())))
        for k in range(16):
            margin[p] = min(float(margin[p]), abs(Tm - 0.5))
            if not Tm > 0.5:
                break
            dm += step
            q = q + qstep
            phi = 1.0 / (1.0 + np.exp(-s * float((trilinear_weights(q) * g).sum())))
            ak = (phi_prev - phi) / (phi_prev + EPS_PHI)
            Tm *= (1.0 - ak)
            phi_prev = phi
        out[p] = dm
    return (out, margin) if return_margin else out


def backward(out: OracleOut, geos, rgbs, s, dL_dcolor, dL_ddepth=None, dL_dnormal=None,
             dL_dT=None, return_ds_abs=False, retain_graph=False):
    """Gradients the reference's backward returns: (dL_dgeos, dL_drgbs, subdiv_p_bw, dL_ds)
    for the given output cotangents (each [C,h,w] over the rendered window).  With return_ds_abs,
    also the sum over pairs of |dL/dalpha * dalpha/ds| -- the scale of the cancelling sum dL/ds."""
    dt = out.color.dtype
    loss = (out.color * dL_dcolor.to(dt)).sum()
    if dL_ddepth is not None:
        loss = loss + (out.depth[0] * dL_ddepth.to(dt)[0]).sum()
    if dL_dnormal is not None:
        loss = loss + (out.normal * dL_dnormal.to(dt)).sum()
    if dL_dT is not None:
        loss = loss + (out.T * dL_dT.to(dt)).sum()
    loss = loss + out.reg_loss
    alphas = [pa[0] for pa in out.pair_alpha]
    gs = torch.autograd.grad(loss, [geos, rgbs, s] + alphas, allow_unused=True, retain_graph=retain_graph)
    g_geo, g_rgb, g_s = gs[0], gs[1], gs[2]
    P = geos.shape[0]
    subdiv = torch.zeros(P, dtype=dt)
    ds_abs = 0.0
    for (alpha, pc, vid, dads), ga in zip(out.pair_alpha, gs[3:]):
        if ga is None:
            continue
        pcf = pc.to(dt)
        v = (ga * alpha.detach()).abs() * pcf
        subdiv.index_add_(0, vid, v.sum(0))
        ds_abs += float(((ga * dads).abs() * pcf).sum())
    z = lambda t, ref: torch.zeros_like(ref) if t is None else t
    res = (z(g_geo, geos), z(g_rgb, rgbs), subdiv, z(g_s, s))
    return res + (ds_abs,) if return_ds_abs else res


# ----------------------------------------------------------------------------------------------
# Scalar loop: the reference's per-pixel control flow, literally (small cases only)
# ----------------------------------------------------------------------------------------------
def render_scalar(cam: RasterCamera, bg, bin_vals, bin_ranges, bbox, centers, lengths, geos,
                  rgbs, s: float, pixels: Sequence[tuple]) -> Dict[tuple, dict]:
    """Per pixel: batches of 256 entries, a j-list of bbox/octant/box hits built while the pixel
    is not done, contributor counting as forward.cu:182-448.  Returns color, T, n_contrib,
    depth0 per pixel."""
    centers = np.asarray(centers, dtype=np.float64)
    lengths = np.asarray(lengths, dtype=np.float64).reshape(-1)
    geos = np.asarray(geos, dtype=np.float64)
    rgbs = np.asarray(rgbs, dtype=np.float64)
    bg = np.asarray(bg, dtype=np.float64)
    tiles_x = (cam.W + BLOCK - 1) // BLOCK
    res = {}
    for (px, py) in pixels:
        ro, rd, rd_norm = pixel_rays(cam, torch.tensor([px]), torch.tensor([py]))
        ro, rd, rd_norm = ro.numpy(), rd[0].numpy(), float(rd_norm[0])
        rd_inv = 1.0 / rd
        pq = (int(rd[0] < 0) << 2) | (int(rd[1] < 0) << 1) | int(rd[2] < 0)
        tile = (py // BLOCK) * tiles_x + (px // BLOCK)
        lo, hi = int(bin_ranges[tile, 0]), int(bin_ranges[tile, 1])
        T, C, D = 1.0, np.zeros(3), 0.0
        done = False
        contributor, last_contributor = 0, 0
        to_do = hi - lo
        for i in range((to_do + 255) // 256):
            end_j = min(256, to_do - i * 256)
            j_lst = []
            for j in range(end_j):
                if done:
                    break
                val = int(bin_vals[lo + i * 256 + j])
                v, q = val & ((1 << 29) - 1), val >> 29
                xmn, ymn = int(bbox[v, 0]) >> 16, int(bbox[v, 0]) & 0xFFFF
                xmx, ymx = int(bbox[v, 1]) >> 16, int(bbox[v, 1]) & 0xFFFF
                if not (xmn <= px <= xmx and ymn <= py <= ymx) or q != pq:
                    continue
                r = 0.5 * lengths[v]
                d = centers[v] - ro
                c0, c1 = (d - r) * rd_inv, (d + r) * rd_inv
                a = max(np.minimum(c0, c1))
                b = min(np.maximum(c0, c1))
                if a > b:
                    continue
                j_lst.append((j, v, a, b))
            inc = 0
            for (j, v, a, b) in j_lst:
                if done:
                    break
                inc = j + 1
                l = lengths[v]
                q0 = (ro + a * rd - (centers[v] - 0.5 * l)) / l
                qs = ((b - a) * rd / 3.0) / l
                sd = [float((trilinear_weights(torch.from_numpy(q0 + qs * k)).numpy() * geos[v]).sum())
                      for k in range(4)]
                ph = [1.0 / (1.0 + np.exp(-s * x)) for x in sd]
                alpha = max((ph[0] - ph[3]) / (ph[0] + EPS_PHI), 0.0)
                if alpha < MIN_ALPHA:
                    continue
                subs = [max((ph[k] - ph[k + 1]) / (ph[k] + EPS_PHI), 0.0) for k in range(3)]
                step = THIRD * (b - a)
                t = [a + (k + 0.5) * step for k in range(3)]
                dval = (subs[0] * t[0] + (1 - subs[0]) * subs[1] * t[1]
                        + (1 - subs[0]) * (1 - subs[1]) * subs[2] * t[2])
                C = C + T * alpha * rgbs[v]
                D = D + T * dval
                T *= (1.0 - alpha)
                done = done or (T < EARLY_STOP_T)
                last_contributor = contributor + inc
            contributor += inc if done else end_j
        res[(px, py)] = dict(color=C + T * bg, T=T, n_contrib=last_contributor, depth0=D / rd_norm)
    return res
