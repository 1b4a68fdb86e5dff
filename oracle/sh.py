"""torch FP64 spherical-harmonics colour (degree <= 3) and the corner gather (test
infrastructure only).

SH follows cuda/src/sh_compute.cu:54-176: direction = normalize(voxel centre - camera position),
rgb = sum_k Y_k(dir) c_k + 0.5, clamped at 0 in the forward; the backward is the transpose of
the linear map with the clamp NOT masked (sh_compute.cu:118-121), restated here as a
straight-through clamp.  Basis: real SH in the 3DGS sign convention,
Y_1 = (-C1 y, C1 z, -C1 x), degree 2/3 as in sh_compute.cu:80-95.

Pinning (tests/test_oracle_sh.py): the basis is compared against scipy.special.sph_harm
(complex SH -> real SH with the Condon-Shortley phase), which pins every constant and sign;
orthonormality over a Lebedev-free Monte Carlo sphere sample; gradient by autograd.

Gather follows cuda/src/geo_params_gather.cu:22-72: geos[i,k] = grid_pts[vox_key[i,k]] and its
transpose (index_add) -- pinned by torch.index_select / index_add_.
"""
from __future__ import annotations

import math

import torch

C0 = 0.28209479177387814
C1 = 0.4886025119029199
C2 = [1.0925484305920792, -1.0925484305920792, 0.31539156525252005, -1.0925484305920792,
      0.5462742152960396]
C3 = [-0.5900435899266435, 2.890611442640554, -0.4570457994644658, 0.3731763325901154,
      -0.4570457994644658, 1.445305721320277, -0.5900435899266435]


def sh_basis(dirs: torch.Tensor, deg: int) -> torch.Tensor:
    """[N,3] unit directions -> [N,(deg+1)^2] basis values."""
    x, y, z = dirs[:, 0], dirs[:, 1], dirs[:, 2]
    out = [torch.full_like(x, C0)]
    if deg >= 1:
        out += [-C1 * y, C1 * z, -C1 * x]
    if deg >= 2:
        xx, yy, zz = x * x, y * y, z * z
        out += [C2[0] * x * y, C2[1] * y * z, C2[2] * (2 * zz - xx - yy), C2[3] * x * z,
                C2[4] * (xx - yy)]
    if deg >= 3:
        xx, yy, zz = x * x, y * y, z * z
        out += [C3[0] * y * (3 * xx - yy), C3[1] * x * y * z, C3[2] * y * (4 * zz - xx - yy),
                C3[3] * z * (2 * zz - 3 * xx - 3 * yy), C3[4] * x * (4 * zz - xx - yy),
                C3[5] * z * (xx - yy), C3[6] * x * (xx - 3 * yy)]
    return torch.stack(out, -1)


def sh_rgb(deg: int, idx, centers, cam_pos, sh0, shs):
    """rgbs [P,3] (zero rows outside idx); sh0/shs may require grad."""
    centers = centers.double()
    P = centers.shape[0]
    if idx is None or idx.numel() == 0:
        idx = torch.arange(P)
    idx = idx.long().cpu()
    d = centers[idx] - cam_pos.double().reshape(1, 3)
    d = d / d.norm(dim=-1, keepdim=True)
    b = sh_basis(d, deg)                                        # [n, nb]
    coef = torch.cat([sh0[idx].unsqueeze(1), shs[idx][:, : b.shape[1] - 1]], 1)   # [n, nb, 3]
    raw = (b.unsqueeze(-1) * coef).sum(1) + 0.5
    clamped = raw + (raw.clamp_min(0) - raw).detach()           # straight-through clamp
    out = torch.zeros(P, 3, dtype=raw.dtype)
    return out.index_put((idx,), clamped)


def gather(vox_key, grid_pts):
    return grid_pts.reshape(-1)[vox_key.long()]


def real_sh_from_scipy(dirs, deg):
    """Real SH values in the same ordering/sign convention from scipy's complex SH (for pinning)."""
    import numpy as np
    from scipy import special
    x, y, z = dirs[:, 0], dirs[:, 1], dirs[:, 2]
    theta = np.arccos(np.clip(z, -1, 1))        # polar
    phi = np.arctan2(y, x)                      # azimuth
    cols = []
    for l in range(deg + 1):
        for m in range(-l, l + 1):
            if hasattr(special, "sph_harm_y"):
                Ylm = special.sph_harm_y(l, abs(m), theta, phi)
            else:
                Ylm = special.sph_harm(abs(m), l, phi, theta)
            if m > 0:
                v = math.sqrt(2) * (-1) ** m * Ylm.real
            elif m < 0:
                v = math.sqrt(2) * (-1) ** m * Ylm.imag
            else:
                v = Ylm.real
            cols.append(v)
    return np.stack(cols, -1)
