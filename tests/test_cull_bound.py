"""The voxel skip bound used by the compositors and the batched binning (voxel_may_contribute in
svraster_amd/csrc/common.h): whenever it reports that a voxel cannot contribute, every float32
segment alpha through the voxel must stay below MIN_ALPHA = 1e-5 (config.h:19), including
segment ends slightly outside the voxel from rounding.  Brute force over random voxels, with the
alpha of forward.cu:303 evaluated in float32."""
import numpy as np


def may_contribute(g, s, pos_scale, l):
    """Restatement of voxel_may_contribute in float32, vectorised over voxels."""
    f = np.float32
    g = g.astype(np.float32)
    gmin, gmax = g.min(1), g.max(1)
    R = gmax - gmin
    eta = f(2e-6) * (f(pos_scale) + f(l)) / f(l)
    sl = R * (f(0.25) + f(8) * eta) + f(1e-6) * np.maximum(np.abs(gmin), np.abs(gmax))
    s = f(s)
    a, b = s * (gmax + sl), s * (gmin - sl)
    d = s * (R + f(2) * sl)
    with np.errstate(over="ignore"):
        sig = lambda x: f(1) / (f(1) + np.exp(-x))
        sa = sig(a)
        B = (-np.expm1(-d)) * sa * sig(-b) / (sa + f(1e-6))
    return ~(B < f(3e-6))


def bound_reference(g, s, pos_scale, l):
    """The same bound in float64 with the direct difference of sigmoids."""
    gmin, gmax = g.min(1).astype(np.float64), g.max(1).astype(np.float64)
    R = gmax - gmin
    eta = 2e-6 * (pos_scale + l) / l
    sl = R * (0.25 + 8.0 * eta) + 1e-6 * np.maximum(np.abs(gmin), np.abs(gmax))
    hi = 1.0 / (1.0 + np.exp(-s * (gmax + sl)))
    lo = 1.0 / (1.0 + np.exp(-s * (gmin - sl)))
    return (hi - lo) / (hi + 1e-6)


def trilerp32(g, q):
    x, y, z = q[:, 0:1], q[:, 1:2], q[:, 2:3]
    one = np.float32(1)
    c = lambda i: g[:, i:i + 1]
    c00 = c(0) * (one - z) + c(1) * z
    c01 = c(2) * (one - z) + c(3) * z
    c10 = c(4) * (one - z) + c(5) * z
    c11 = c(6) * (one - z) + c(7) * z
    c0 = c00 * (one - y) + c01 * y
    c1 = c10 * (one - y) + c11 * y
    return (c0 * (one - x) + c1 * x)[:, 0]


def alpha32(g, s, q0, q3):
    s = np.float32(s)
    phi0 = np.float32(1) / (np.float32(1) + np.exp(-(s * trilerp32(g, q0))))
    phi3 = np.float32(1) / (np.float32(1) + np.exp(-(s * trilerp32(g, q3))))
    return np.maximum((phi0 - phi3) / (phi0 + np.float32(1e-6)), np.float32(0))


def test_skipped_voxels_never_reach_min_alpha():
    rng = np.random.default_rng(0)
    n = 400_000
    l = np.float32(2.0 / 4096)
    s = np.log(99.0) / (4 * float(l))
    # Corner SDFs: a plane at a random offset (in voxel sizes) plus noise, spanning the skip bound.
    off = rng.uniform(-40, 40, size=(n, 1)) * l
    nrm = rng.normal(size=(n, 3))
    nrm /= np.linalg.norm(nrm, axis=1, keepdims=True)
    corners = np.array([[(i >> 2) & 1, (i >> 1) & 1, i & 1] for i in range(8)], dtype=np.float64) - 0.5
    g = (off + (corners @ nrm.T).T * l + rng.normal(scale=0.05, size=(n, 8)) * l).astype(np.float32)
    keep = may_contribute(g, s, 3.0, float(l))
    assert 0.05 < keep.mean() < 0.95          # the test spans both sides of the bound
    # The float32 factorised form agrees with the float64 difference of sigmoids away from the
    # threshold (it only decides near 3e-6).
    B = bound_reference(g, s, 3.0, float(l))
    assert not np.any(~keep & (B > 3.3e-6)) and not np.any(keep & (B < 2.7e-6))
    skip = ~keep
    worst = 0.0
    for _ in range(8):
        # Segment ends: random points of the voxel, pushed up to 0.01 voxel sizes outside it.
        q0 = rng.uniform(-0.01, 1.01, size=(n, 3)).astype(np.float32)
        q3 = rng.uniform(-0.01, 1.01, size=(n, 3)).astype(np.float32)
        a = alpha32(g, s, q0, q3)
        worst = max(worst, float(a[skip].max()))
        assert (a[skip] < 1e-5).all()
    # Voxels the bound keeps do produce alphas above the threshold (the bound is not vacuous).
    assert float(a[keep].max()) > 1e-2
    print(f"largest alpha of a skipped voxel {worst:.2e}")
