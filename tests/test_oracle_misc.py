"""Pins of the SH, Adam, octree and grid-loss oracles (CPU only)."""
import numpy as np
import torch

from oracle import adam as oad
from oracle import grid_loss as ogl
from oracle import octree as oot
from oracle import sh as osh


def test_sh_basis_matches_scipy_real_sh():
    rng = np.random.default_rng(0)
    d = rng.normal(size=(64, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    ours = osh.sh_basis(torch.from_numpy(d), 3).numpy()
    ref = osh.real_sh_from_scipy(d, 3)
    # 3DGS sign convention = real SH (Condon-Shortley) times (-1)^m, m = -l..l.
    sign = np.array([(-1) ** m for l in range(4) for m in range(-l, l + 1)], dtype=np.float64)
    np.testing.assert_allclose(ours, ref * sign, rtol=1e-12, atol=1e-12)


def test_sh_backward_is_transpose_without_clamp_mask():
    rng = np.random.default_rng(1)
    P = 20
    centers = torch.from_numpy(rng.normal(size=(P, 3)))
    cam = torch.tensor([3.0, 0.5, -1.0], dtype=torch.float64)
    sh0 = torch.from_numpy(rng.normal(size=(P, 3))).requires_grad_(True)
    shs = torch.from_numpy(rng.normal(size=(P, 15, 3))).requires_grad_(True)
    sh0.data[0] = -500.0        # force a clamped colour
    rgb = osh.sh_rgb(3, None, centers, cam, sh0, shs)
    assert (rgb.detach()[0] == 0).all()
    g = torch.ones(P, 3, dtype=torch.float64)
    d0, ds = torch.autograd.grad((rgb * g).sum(), [sh0, shs])
    assert torch.allclose(d0[0], torch.full((3,), osh.C0, dtype=torch.float64))   # not masked


def test_adam_matches_torch_optim():
    rng = np.random.default_rng(2)
    p0 = rng.normal(size=100)
    p = torch.tensor(p0, requires_grad=True)
    opt = torch.optim.Adam([p], lr=1e-2, betas=(0.1, 0.99), eps=1e-15)
    mine = (p0.copy(), np.zeros(100), np.zeros(100))
    for t in range(1, 6):
        g = rng.normal(size=100)
        p.grad = torch.tensor(g)
        opt.step()
        mine = oad.adam_step(mine[0], g, mine[1], mine[2], t, 1e-2, 0.1, 0.99, 1e-15)
    np.testing.assert_allclose(mine[0], p.detach().numpy(), rtol=1e-12, atol=1e-12)


def test_biased_adam_by_hand_and_via_torch_optim():
    """Biased form (adam_step.cu:137-158): hand values, and the identity that the biased step
    with lr' = lr sqrt(1-b2^t) / (1-b1^t) and eps' = eps sqrt(1-b2^t) is the bias-corrected step,
    i.e. torch.optim.Adam (a library routine)."""
    # step 1 from zero moments, eps = 0: m = 0.1 g, v = 0.01 g^2, p -= 0.1 * 0.1 g / (0.1 |g|)
    p, m, v = oad.adam_step([1.0, -2.0, 0.5], [4.0, -0.25, 0.0], [0.0] * 3, [0.0] * 3, 1, 0.1, 0.9, 0.99, 0.0,
                            biased=True, sparse=True)
    np.testing.assert_allclose(p, [0.9, -1.9, 0.5], rtol=0, atol=1e-15)
    np.testing.assert_allclose(m, [0.4, -0.025, 0.0], rtol=1e-15)
    np.testing.assert_allclose(v, [0.16, 0.000625, 0.0], rtol=1e-13)
    # step 2 with g = 2 on the first element: m = 0.9*0.4 + 0.1*2 = 0.56,
    # v = 0.99*0.16 + 0.01*4 = 0.1984, p = 0.9 - 0.1 * 0.56 / sqrt(0.1984)
    p2, m2, v2 = oad.adam_step(p[:1], [2.0], m[:1], v[:1], 2, 0.1, 0.9, 0.99, 0.0, biased=True)
    np.testing.assert_allclose(m2, [0.56], rtol=1e-15)
    np.testing.assert_allclose(v2, [0.1984], rtol=1e-14)
    np.testing.assert_allclose(p2, [0.9 - 0.1 * 0.56 / np.sqrt(0.1984)], rtol=1e-15)
    # biased with per-step rescaled lr / eps == torch.optim.Adam
    rng = np.random.default_rng(12)
    lr, b1, b2, eps = 1e-2, 0.9, 0.99, 1e-8
    p0 = rng.normal(size=50)
    pt = torch.tensor(p0, requires_grad=True)
    opt = torch.optim.Adam([pt], lr=lr, betas=(b1, b2), eps=eps)
    mine = (p0.copy(), np.zeros(50), np.zeros(50))
    for t in range(1, 8):
        g = rng.normal(size=50)
        pt.grad = torch.tensor(g)
        opt.step()
        bc1, bc2 = 1 - b1 ** t, 1 - b2 ** t
        mine = oad.adam_step(mine[0], g, mine[1], mine[2], t, lr * np.sqrt(bc2) / bc1, b1, b2, eps * np.sqrt(bc2),
                             biased=True)
    np.testing.assert_allclose(mine[0], pt.detach().numpy(), rtol=1e-12, atol=1e-12)


def test_adam_sparse_skips_zero_gradients():
    p, m, v = oad.adam_step(np.ones(4), np.array([0.0, 1.0, 0.0, -2.0]), np.full(4, 0.5), np.full(4, 0.25),
                            3, 0.1, 0.9, 0.99, 1e-8, sparse=True)
    assert p[0] == 1.0 and m[0] == 0.5 and v[0] == 0.25
    assert p[1] != 1.0


def test_octree_codec_by_hand_and_round_trip():
    assert oot.ijk_to_octpath(1, 0, 1, 1) == 0b101 << 45
    assert oot.ijk_to_octpath(0, 0, 0, 3) == 0
    assert oot.ijk_to_octpath(3, 0, 0, 2) == (0b100 << 45) | (0b100 << 42)
    rng = np.random.default_rng(3)
    for _ in range(200):
        lv = int(rng.integers(1, 17))
        i, j, k = (int(x) for x in rng.integers(0, 1 << lv, 3))
        assert oot.octpath_to_ijk(oot.ijk_to_octpath(i, j, k, lv), lv) == (i, j, k)
    # Morton order: a parent's children occupy contiguous, increasing paths
    kids = [oot.ijk_to_octpath(2 * 1 + (c >> 2), 2 * 0 + ((c >> 1) & 1), 2 * 1 + (c & 1), 2) for c in range(8)]
    assert kids == sorted(kids) and kids[0] == oot.ijk_to_octpath(1, 0, 1, 1)


def _dense_grid(n, fn, level_cells=1.0):
    """n^3 unit voxels in grid units, corner values from fn(corner xyz in grid units)."""
    ijk = np.stack(np.meshgrid(np.arange(n), np.arange(n), np.arange(n), indexing="ij"), -1).reshape(-1, 3)
    corners = np.array([[(c >> 2) & 1, (c >> 1) & 1, c & 1] for c in range(8)])
    cpos = ijk[:, None, :] + corners[None]
    code = (cpos[..., 0] * (n + 1) + cpos[..., 1]) * (n + 1) + cpos[..., 2]
    uniq, inv = np.unique(code.reshape(-1), return_inverse=True)
    vox_key = inv.reshape(-1, 8)
    pts = np.stack([uniq // ((n + 1) ** 2), (uniq // (n + 1)) % (n + 1), uniq % (n + 1)], -1).astype(np.float64)
    grid_pts = np.array([fn(p) for p in pts])
    keys = ijk[:, 0] + n * (ijk[:, 1] + n * ijk[:, 2])
    order = np.argsort(keys, kind="stable")
    mask = np.zeros(n ** 3, dtype=bool)
    mask[keys] = True
    return grid_pts, vox_key, ijk.astype(np.float64), np.ones(len(ijk)), keys[order], order, mask


def test_eikonal_linear_field_closed_form():
    n, a = 8, np.array([0.3, -0.4, 1.2])
    grid_pts, vox_key, coord, vsize, keys, owners, mask = _dense_grid(n, lambda p: float(p @ a))
    active = np.arange(len(keys))
    randv = np.full((len(keys), 3), 0.37)
    # inv chosen so that |grad| * 0.5 * inv == 1: the loss is at its minimum -> zero gradient
    inv = 2.0 / np.sqrt((a * a).sum() + 1e-6)
    g = ogl.grid_loss_grad("ge", grid_pts, vox_key, coord, vsize, n, mask, keys, owners, active, 1.0, inv, randv)
    assert np.abs(g).max() < 1e-9
    g2 = ogl.grid_loss_grad("ge", grid_pts, vox_key, coord, vsize, n, mask, keys, owners, active, 1.0, 1.0, randv)
    assert np.abs(g2).max() > 1e-3


def test_laplacian_linear_zero_and_quadratic_closed_form():
    n = 8
    lin = _dense_grid(n, lambda p: float(p @ np.array([0.2, 0.1, -0.3])))
    active = np.arange(len(lin[4]))
    randv = np.full((len(active), 3), 0.5)
    g = ogl.grid_loss_grad("ls", *lin[:2], lin[2], lin[3], n, lin[6], lin[4], lin[5], active, 1.0, 1.0, randv)
    assert np.abs(g).max() < 1e-4          # residual 0 -> gradient ~ r/sqrt(r^2+1e-10) * 0 = 0
    # quadratic along x: interpolated field is piecewise linear between corners; at cell
    # centres offset by 0.5 the second difference of x^2 sampled on the grid is exactly 2.
    quad = _dense_grid(n, lambda p: float(p[0] ** 2))
    val = ogl.grid_loss_value("ls", quad[0], quad[1], quad[2], quad[3], n, quad[6], quad[4], quad[5],
                              active, 1.0, 1.0, np.zeros((len(active), 3)))
    n_interior = (n - 2) ** 3
    assert abs(val - n_interior * 0.5 * 2.0) < 1e-3


def test_grid_losses_match_finite_differences():
    n = 7
    rng = np.random.default_rng(4)
    base = _dense_grid(n, lambda p: float(np.sin(0.7 * p[0]) + 0.3 * p[1] - 0.2 * p[2] ** 1.5))
    grid_pts = base[0] + rng.normal(scale=0.05, size=base[0].shape)
    active = np.arange(len(base[4]))
    randv = rng.uniform(0.2, 0.8, size=(len(active), 3))
    for kind in ("ge", "ls"):
        args = (base[1], base[2], base[3], n, base[6], base[4], base[5], active, 0.7, 1.3, randv)
        g = ogl.grid_loss_grad(kind, grid_pts, *args)
        for i in rng.choice(len(grid_pts), 15, replace=False):
            e = np.zeros_like(grid_pts)
            e[i] = 1e-6
            num = (ogl.grid_loss_value(kind, grid_pts + e, *args) - ogl.grid_loss_value(kind, grid_pts - e, *args)) / 2e-6
            assert abs(num - g[i]) <= 1e-5 * max(1.0, abs(num)), (kind, i, num, g[i])


def test_valid_gradient_table_rules():
    # one leaf of 2 cells, one non-leaf of 1 cell (kept), one non-leaf of 4 cells (dropped),
    # one leaf of 0.5 cells (dropped), one leaf outside the region (dropped)
    res_pow2 = 4
    cell = 2.0 / 16
    centers = np.array([[-0.5, -0.5, -0.5], [0.5 + cell / 2] * 3, [0.0, 0.0, 0.0],
                        [0.3, 0.3, 0.3], [1.5, 0.0, 0.0]], dtype=np.float32)
    sizes = np.array([2 * cell, cell, 4 * cell, cell / 2, cell], dtype=np.float32)
    leaf = np.array([1, 0, 0, 1, 1])
    mask, keys, owners = ogl.valid_gradient_table(centers, sizes, np.zeros(3), 2.0, res_pow2, leaf)
    assert sorted(set(owners.tolist())) == [0, 1]
    assert (owners == 0).sum() == 8 and (owners == 1).sum() == 1
    assert mask.sum() == 9


def _fd(f, x, idx, h=1e-6):
    xp, xm = x.copy(), x.copy()
    xp[idx] += h
    xm[idx] -= h
    return (f(xp) - f(xm)) / (2 * h)


def test_total_variation_closed_form_ties_and_finite_differences():
    # one voxel, corner values 0..7: corner i compares with i^1, i^2, i^4
    keys = np.arange(8)[None, :]
    g = ogl.tv_grad(np.arange(8.0), keys, 0.5, [1.0], True, False, np.zeros(8)).reshape(-1)
    assert np.array_equal(g, 0.5 * np.array([-3, -1, -1, 1, -1, 1, 1, 3], dtype=float))
    # ties count as "not greater": all equal -> -3 everywhere
    g = ogl.tv_grad(np.ones(8), keys, 1.0, [1.0], True, False, np.zeros(8)).reshape(-1)
    assert np.array_equal(g, -3.0 * np.ones(8))
    # no_tv_s False scales by 0.01 * vox_size_inv
    g2 = ogl.tv_grad(np.arange(8.0), keys, 0.5, [4.0], False, False, np.zeros(8)).reshape(-1)
    assert np.allclose(g2, 0.5 * 0.01 * 4.0 * np.array([-3, -1, -1, 1, -1, 1, 1, 3]))
    # random voxels sharing corners: the gradient of w * sum |edge differences|
    rng = np.random.default_rng(0)
    vk = np.array([[0, 1, 2, 3, 4, 5, 6, 7], [4, 5, 6, 7, 8, 9, 10, 11]])
    x = rng.normal(size=12)
    inv = np.array([2.0, 3.0])
    got = ogl.tv_grad(x, vk, 0.3, inv, False, False, np.zeros(12)).reshape(-1)
    fd = np.array([_fd(lambda y: ogl.tv_value(y, vk, 0.3, inv, False), x, i) for i in range(12)])
    assert np.allclose(got, fd, atol=1e-6)
    # tv_sparse: only voxels with a corner whose gradient is already non-zero
    gin = np.zeros(12)
    gin[0] = 1.0            # corner of voxel 0 only
    got = ogl.tv_grad(x, vk, 0.3, inv, True, True, gin).reshape(-1)
    ref = ogl.tv_grad(x, vk[:1], 0.3, inv[:1], True, False, gin).reshape(-1)
    assert np.array_equal(got, ref)


def test_voxel_gradient_closed_form_and_finite_differences():
    # corners of a voxel of size h sampled from a linear field a.x: g = a exactly
    h = 0.5
    a = np.array([1.0, 0.5, -0.25])
    corners = np.array([[(hx - 0.5) * h, (hy - 0.5) * h, (hz - 0.5) * h]
                        for hx in (0, 1) for hy in (0, 1) for hz in (0, 1)])   # bit order x=4, y=2, z=1
    S = corners @ a
    keys = np.arange(8)[None, :]
    g = ogl.vg_grad(S, keys, [1.0 / h], [0], 0.7)
    # |g|^2 = 1.3125: corner i gets 4 (|a|^2 - 2) w (a . s_i) / 4
    s = np.array([[1.0 if i & 4 else -1.0, 1.0 if i & 2 else -1.0, 1.0 if i & 1 else -1.0] for i in range(8)])
    assert np.allclose(g, (a @ a - 2.0) * 0.7 * (s @ a))
    a2 = a / np.linalg.norm(a) * np.sqrt(2.0)           # |g|^2 = 2: zero gradient
    assert np.allclose(ogl.vg_grad(corners @ a2, keys, [1.0 / h], [0], 0.7), 0.0, atol=1e-12)
    # h times the gradient of w (|g|^2 - 2)^2, repeated active entries counted twice
    rng = np.random.default_rng(1)
    vk = np.array([[0, 1, 2, 3, 4, 5, 6, 7], [4, 5, 6, 7, 8, 9, 10, 11]])
    x = rng.normal(size=12)
    inv = np.array([2.0, 2.0])
    act = [0, 1, 1]
    got = ogl.vg_grad(x, vk, inv, act, 0.3)
    fd = np.array([_fd(lambda y: ogl.vg_value(y, vk, inv, act, 0.3), x, i) for i in range(12)])
    assert np.allclose(got, 0.5 * fd, atol=1e-6)


def test_points_loss_finite_differences_and_skips():
    res = 4
    mask = np.zeros(res ** 3, dtype=bool)
    keys = np.array([0, 1, 4, 5, 16, 17, 20, 21])      # cells of a 2x2x2 voxel at the origin
    mask[keys] = True
    owners = np.zeros(8, dtype=np.int64)
    vk = np.arange(8)[None, :]
    coord = np.zeros((1, 3))
    vsize = np.array([2.0])
    rng = np.random.default_rng(2)
    x = rng.normal(size=8)
    pts = np.array([[0.5, 0.7, 1.2], [1.9, 0.1, 0.3], [-0.5, 0.2, 0.2],      # last: trunc(-0.5) = 0, kept
                    [3.5, 3.5, 3.5], [4.5, 0.0, 0.0], [-1.5, 0.0, 0.0]])      # masked, outside, outside
    got = ogl.pl_grad(pts, x, vk, coord, vsize, res, mask, keys, owners, 0.4)
    fd = np.array([_fd(lambda y: ogl.pl_value(pts, y, vk, coord, vsize, res, mask, keys, owners, 0.4), x, i)
                   for i in range(8)])
    assert np.allclose(got, fd, atol=1e-6)
    only3 = ogl.pl_grad(pts[:3], x, vk, coord, vsize, res, mask, keys, owners, 0.4)
    assert np.array_equal(got, only3)


def test_octree_topology_oracle_pins():
    rng = np.random.default_rng(4)
    lv = rng.integers(1, 6, size=40)
    ijk = np.stack([rng.integers(0, 1 << l, size=3) for l in lv])
    path = np.array([oot.ijk_to_octpath(int(a), int(b), int(c), int(l)) for (a, b, c), l in zip(ijk, lv)])
    # children tile the parent: their ijk at level + 1 are 2 * parent + {0,1}^3, octant order
    cp, cl = oot.gen_children(path, lv)
    cijk = oot.ijk_array(cp, cl).reshape(-1, 8, 3)
    offs = np.array([[(c >> 2) & 1, (c >> 1) & 1, c & 1] for c in range(8)])
    assert np.array_equal(cijk, 2 * ijk[:, None, :] + offs[None])
    assert np.array_equal(cl, np.repeat(lv + 1, 8))
    # the link reproduces every voxel's corner coordinates, keys unique and sorted
    key, vk = oot.build_grid_pts_link(path, lv)
    assert np.array_equal(key[vk], oot.corner_coords(path, lv))
    assert np.all(np.diff(key[:, 0] * (1 << 40) + key[:, 1] * (1 << 20) + key[:, 2]) > 0)
    # trilinear subdivision of a linear field's corner values = the field at the children's corners
    a = rng.normal(size=3)
    field = lambda x: x @ a + 0.7
    par = oot.corner_coords(path, lv).astype(np.float64)
    kids = oot.corner_coords(cp, cl).astype(np.float64)
    sub = oot.subdivide_by_interp(field(par))
    assert np.allclose(sub, field(kids), rtol=1e-12, atol=1e-9)
    # the mean of a constant is the constant; unnamed points are 0
    m = oot.agg_mean(key.shape[0] + 3, vk, np.full(vk.shape, 2.5))
    assert np.all(m[:key.shape[0]] == 2.5) and np.all(m[key.shape[0]:] == 0.0)
