"""Differentiable warp residual on the GPU (needs an MI355X).

Forward: the ZNSSD ``C = 2 (1 - ZNCC)`` per subset under a trilinear FE node field, against (a) the
oracle's ZNCC of an affine warp when the node field is affine (trilinear interpolation reproduces
it exactly), and (b) a float64 NumPy restatement (positions from the FE shape functions, samples
from the oracle's float64 interpolator) for a random field. Backward: against central finite
differences of that float64 forward (h = 1e-3 voxel for node displacements, 1e-3 for voxels of
g), and against the IC-GN solution: the warp-residual gradient vanishes where the oracle's IC-GN
converged. Tolerances are written in each test; the kernels compute in float32.
"""

from __future__ import annotations

import numpy as np
import pytest
import torch

from oracle import cpu_ref
from tests.parity import bead_case

pytestmark = pytest.mark.gpu

W = pytest.importorskip("pyaldvc_amd.warp")
K = pytest.importorskip("pyaldvc_amd.kernels")

HALF = (16, 16, 16)


def _fe(node_u, origin, spacing, pts):
    """Trilinear FE displacement at points (M, 3) [x, y, z]; edge cells extrapolate."""
    gnz, gny, gnx, _ = node_u.shape
    q = (pts - np.asarray(origin)) / np.asarray(spacing)
    n = np.array([gnx, gny, gnz])
    c = np.clip(np.floor(q).astype(np.int64), 0, n - 2)
    t = q - c
    u = np.zeros((pts.shape[0], 3))
    for k in (0, 1):
        for j in (0, 1):
            for i in (0, 1):
                N = (t[:, 0] if i else 1 - t[:, 0]) * (t[:, 1] if j else 1 - t[:, 1]) * (t[:, 2] if k else 1 - t[:, 2])
                u += N[:, None] * node_u[c[:, 2] + k, c[:, 1] + j, c[:, 0] + i]
    return u


def _offsets(half):
    hx, hy, hz = half
    dz, dy, dx = np.meshgrid(np.arange(-hz, hz + 1), np.arange(-hy, hy + 1), np.arange(-hx, hx + 1), indexing="ij")
    return np.stack([dx.ravel(), dy.ravel(), dz.ravel()], 1)


def forward_ref(f, g, centers, half, node_u, origin, spacing, mode=1):
    """float64 C per subset (NaN when a sample leaves the domain)."""
    off = _offsets(half)
    S, V = len(centers), off.shape[0]
    p = (centers[:, None, :] + off[None, :, :]).reshape(-1, 3)
    w = p + _fe(node_u, origin, spacing, p.astype(np.float64))
    gv = cpu_ref.sample_points(g, w[:, 2], w[:, 1], w[:, 0], mode).reshape(S, V)
    fv = f[p[:, 2], p[:, 1], p[:, 0]].astype(np.float64).reshape(S, V)
    fc = fv - fv.mean(axis=1, keepdims=True)
    gc = gv - gv.mean(axis=1, keepdims=True)
    ft = fc / np.sqrt(np.sum(fc * fc, axis=1, keepdims=True))
    gt = gc / np.sqrt(np.sum(gc * gc, axis=1, keepdims=True))
    return np.sum((ft - gt) ** 2, axis=1)  # NaN wherever a sample was NaN


def _lattice(shape, half, step):
    ax = [np.arange(h, n - h, step) for n, h in zip(shape[::-1], half)]
    X, Y, Z = np.meshgrid(*ax, indexing="ij")
    return np.stack([X.ravel(), Y.ravel(), Z.ravel()], 1).astype(np.int64)


@pytest.fixture(scope="module")
def pair():
    """96^3 bead pair under a smooth random displacement field on a 5^3 node grid (spacing 24)."""
    from scipy.ndimage import gaussian_filter

    from pyaldvc_amd import synthetic as syn

    shape = (96, 96, 96)
    origin, spacing = (0.0, 0.0, 0.0), (24.0, 24.0, 24.0)
    rng = np.random.default_rng(12)
    node_u = gaussian_filter(rng.normal(0, 1.5, (5, 5, 5, 3)), sigma=(1, 1, 1, 0)) * 1.5

    def disp(x, y, z):
        u = _fe(node_u, origin, spacing, np.stack([x, y, z], 1))
        return u[:, 0], u[:, 1], u[:, 2]

    f_raw, g_raw = syn.bead_pair_np(shape, 5000, 1.5, disp, seed=3, noise=0.01)
    f, g = syn.normalize_np(f_raw), syn.normalize_np(g_raw)
    ax = np.arange(20, 73, 13)  # windows stay inside the sampling domain under the field
    X, Y, Z = np.meshgrid(ax, ax, ax, indexing="ij")
    centers = np.stack([X.ravel(), Y.ravel(), Z.ravel()], 1).astype(np.int64)
    return dict(f=f, g=g, node_u=node_u, origin=origin, spacing=spacing, centers=centers)


def _t(x, dt=torch.float32):
    return torch.from_numpy(np.ascontiguousarray(x)).to(K.device(), dt)


def test_forward_affine_field_matches_oracle_zncc():
    cs = bead_case(shape=(64, 72, 80))
    A, c0 = cs["F"], np.array([1.3, -0.7, 0.9])
    origin, spacing = (-8.0, -8.0, -8.0), (40.0, 44.0, 48.0)
    gx, gy, gz = (origin[i] + spacing[i] * np.arange(3) for i in range(3))
    Z, Y, X = np.meshgrid(gz, gy, gx, indexing="ij")
    nodes = np.stack([X, Y, Z], -1)
    centre = np.array([(s - 1) / 2 for s in cs["shape"][::-1]])
    node_u = (nodes - centre) @ A.T + c0  # the bead_case field u = A (x - centre) + t
    centers = _lattice(cs["shape"], (8, 8, 8), 8)
    C, stats = W.warp_znssd_stats(node_u, cs["g"], cs["f"], centers, (8, 8, 8), origin, spacing, 1)
    C = C.cpu().numpy()
    P = np.zeros((len(centers), 12))
    P[:, :9] = A.reshape(-1)
    P[:, 9:] = (centers - centre) @ A.T + c0
    ones = np.ones((1, 1, 1), np.uint8)
    H, L, mf, bf, nv, valid = cpu_ref.precompute_nodes(centers, 8, 8, 8, cs["f"], cs["gx"], cs["gy"], cs["gz"], ones, 0.0, 1e300)
    z = cpu_ref.evaluate_zncc_parallel(centers, P, 8, 8, 8, cs["f"], ones, cs["g"], 1, mf, bf, np.ones(len(centers), bool))
    ref = 2.0 * (1.0 - z)
    assert np.array_equal(np.isnan(C), np.isnan(ref))
    ok = ~np.isnan(ref)
    assert ok.sum() > 0.5 * len(ref)
    assert np.max(np.abs(C[ok] - ref[ok])) < 2e-5


def test_forward_matches_float64(pair):
    p = pair
    C, _ = W.warp_znssd_stats(p["node_u"], p["g"], p["f"], p["centers"], HALF, p["origin"], p["spacing"], 1)
    ref = forward_ref(p["f"], p["g"], p["centers"], HALF, p["node_u"], p["origin"], p["spacing"], 1)
    C = C.cpu().numpy()
    assert np.array_equal(np.isnan(C), np.isnan(ref))
    ok = ~np.isnan(ref)
    assert ok.sum() == len(ref)
    assert np.max(np.abs(C[ok] - ref[ok])) < 2e-5


def _affected(centers, pos, reach):
    """Subsets whose window comes within `reach` of `pos` (per axis)."""
    return np.flatnonzero(np.all(np.abs(centers - np.asarray(pos)[None, :]) <= np.asarray(HALF) + reach, axis=1))


def test_backward_node_displacements_match_finite_differences(pair):
    p = pair
    rng = np.random.default_rng(1)
    wts = rng.uniform(0.5, 1.5, len(p["centers"]))
    u_eval = p["node_u"] + 0.25 * rng.standard_normal(p["node_u"].shape)  # away from the minimum
    nu = _t(u_eval).requires_grad_(True)
    C = W.warp_residual(nu, _t(p["g"]), _t(p["f"]), _t(p["centers"], torch.int64), HALF, p["origin"], p["spacing"], 1)
    L = torch.nansum(C * _t(wts, torch.float64))
    L.backward()
    grad = nu.grad.cpu().numpy()
    h = 1e-3
    for node in [(2, 2, 2), (1, 2, 3)]:
        pos = np.array(p["origin"]) + np.array(p["spacing"]) * np.array(node[::-1])
        sub = _affected(p["centers"], pos, np.array(p["spacing"]))
        for comp in range(3):
            num = []
            for sgn in (1, -1):
                u = u_eval.copy()
                u[node + (comp,)] += sgn * h
                c = forward_ref(p["f"], p["g"], p["centers"][sub], HALF, u, p["origin"], p["spacing"], 1)
                num.append(np.nansum(wts[sub] * c))
            fd = (num[0] - num[1]) / (2 * h)
            an = grad[node + (comp,)]
            assert abs(an - fd) <= 2e-3 * max(abs(fd), 1e-2), (node, comp, an, fd)


def test_backward_deformed_volume_matches_finite_differences(pair):
    p = pair
    rng = np.random.default_rng(2)
    wts = rng.uniform(0.5, 1.5, len(p["centers"]))
    g = _t(p["g"]).requires_grad_(True)
    C = W.warp_residual(_t(p["node_u"]), g, _t(p["f"]), _t(p["centers"], torch.int64), HALF, p["origin"], p["spacing"], 1)
    torch.nansum(C * _t(wts, torch.float64)).backward()
    grad = g.grad.cpu().numpy()
    h = 1e-3
    for vox in [(48, 48, 48), (30, 52, 41), (20, 61, 70)]:
        pos = np.array(vox[::-1])
        sub = _affected(p["centers"], pos, np.array([6, 6, 6]))
        num = []
        for sgn in (1, -1):
            gg = p["g"].astype(np.float64).copy()
            gg[vox] += sgn * h
            c = forward_ref(p["f"], gg, p["centers"][sub], HALF, p["node_u"], p["origin"], p["spacing"], 1)
            num.append(np.nansum(wts[sub] * c))
        fd = (num[0] - num[1]) / (2 * h)
        assert abs(grad[vox] - fd) <= 2e-3 * max(abs(fd), 1e-3), (vox, grad[vox], fd)


def test_gradient_vanishes_at_the_icgn_solution():
    """Translation-only pair: where the oracle's IC-GN converged, the warp residual is stationary."""
    cs = bead_case(shape=(64, 64, 64), F=np.zeros((3, 3)), t=(0.6, -0.4, 0.3), seed=7)
    centre = np.array([[32, 32, 32]], np.int64)
    ones = np.ones((1, 1, 1), np.uint8)
    H, L, mf, bf, nv, valid = cpu_ref.precompute_nodes(centre, *HALF, cs["f"], cs["gx"], cs["gy"], cs["gz"], ones, 0.5, 1e12)
    P0 = np.zeros((1, 12))
    Pc, it, st, z = cpu_ref.icgn_12dof_parallel(centre, P0, *HALF, cs["f"], cs["gx"], cs["gy"], cs["gz"], ones, cs["g"], 1,
                                                L, mf, bf, valid, 1e-4, 1e-5, 100, 5)  # fmt: skip
    assert st[0] == 0
    origin, spacing = (0.0, 0.0, 0.0), (63.0, 63.0, 63.0)

    def grad_at(t):
        nu = _t(np.broadcast_to(t, (2, 2, 2, 3)).copy()).requires_grad_(True)
        C = W.warp_residual(nu, _t(cs["g"]), _t(cs["f"]), _t(centre, torch.int64), HALF, origin, spacing, 1)
        C.sum().backward()
        return nu.grad.cpu().numpy().sum(axis=(0, 1, 2))  # d C / d(uniform translation)

    g_star = grad_at(Pc[0, 9:])
    g_off = grad_at(Pc[0, 9:] + np.array([0.3, -0.2, 0.25]))
    assert np.linalg.norm(g_star) < 0.05 * np.linalg.norm(g_off)


def _keys(t):
    """Keys cubic (a = -0.5) weights and derivatives of the 4 taps at fraction t (M,) -> (M, 4) each."""
    t = np.asarray(t, dtype=np.float64)
    w = np.stack([((-0.5 * t + 1.0) * t - 0.5) * t, (1.5 * t - 2.5) * t * t + 1.0, ((-1.5 * t + 2.0) * t + 0.5) * t,
                  (0.5 * t - 0.5) * t * t], 1)  # fmt: skip
    d = np.stack([(-1.5 * t + 2.0) * t - 0.5, (4.5 * t - 5.0) * t, (-4.5 * t + 4.0) * t + 0.5, (1.5 * t - 1.0) * t], 1)
    return w, d


def backward_ref(f, g, centers, half, node_u, origin, spacing, grad_C):
    """float64 gradients of L = sum_n grad_C[n] C_n, subset by subset: every sample's dC/dg scattered
    into its 64 cubic taps, and N_a dC/dg grad g(phi) into the 8 nodes of its cell."""
    off = _offsets(half)
    gg = np.zeros(g.shape)
    gu = np.zeros(node_u.shape)
    gnz, gny, gnx, _ = node_u.shape
    for n, c in enumerate(centers):
        if grad_C[n] == 0.0:
            continue
        p = c[None, :] + off
        u = _fe(node_u, origin, spacing, p.astype(np.float64))
        w = p + u
        b = np.floor(w).astype(np.int64)
        t = w - b
        (wx, dx), (wy, dy), (wz, dz) = (_keys(t[:, k]) for k in range(3))
        val = np.zeros(len(p))
        grad = np.zeros((len(p), 3))
        taps = []
        for k in range(4):
            for j in range(4):
                for i in range(4):
                    idx = (b[:, 2] + k - 1, b[:, 1] + j - 1, b[:, 0] + i - 1)
                    s = g[idx].astype(np.float64)
                    val += wx[:, i] * wy[:, j] * wz[:, k] * s
                    grad[:, 0] += dx[:, i] * wy[:, j] * wz[:, k] * s
                    grad[:, 1] += wx[:, i] * dy[:, j] * wz[:, k] * s
                    grad[:, 2] += wx[:, i] * wy[:, j] * dz[:, k] * s
                    taps.append((idx, wx[:, i] * wy[:, j] * wz[:, k]))
        fv = f[p[:, 2], p[:, 1], p[:, 0]].astype(np.float64)
        fc, gc = fv - fv.mean(), val - val.mean()
        bf, bg = np.sqrt(np.sum(fc * fc)), np.sqrt(np.sum(gc * gc))
        z = np.sum(fc * gc) / (bf * bg)
        dg = -2.0 * grad_C[n] * (fc / bf - z * gc / bg) / bg  # dL/dg at each sample
        for idx, wt in taps:
            np.add.at(gg, idx, dg * wt)
        q = (p - np.asarray(origin)) / np.asarray(spacing)
        cc = np.clip(np.floor(q).astype(np.int64), 0, np.array([gnx, gny, gnz]) - 2)
        tt = q - cc
        for k in (0, 1):
            for j in (0, 1):
                for i in (0, 1):
                    N = (tt[:, 0] if i else 1 - tt[:, 0]) * (tt[:, 1] if j else 1 - tt[:, 1]) * (tt[:, 2] if k else 1 - tt[:, 2])
                    np.add.at(gu, (cc[:, 2] + k, cc[:, 1] + j, cc[:, 0] + i), (N * dg)[:, None] * grad)
    return gu, gg


def _backward(p, grad_C):
    nu = _t(p["node_u"]).requires_grad_(True)
    g = _t(p["g"]).requires_grad_(True)
    C = W.warp_residual(nu, g, _t(p["f"]), _t(p["centers"], torch.int64), HALF, p["origin"], p["spacing"], 1)
    (C * _t(grad_C, torch.float64)).nansum().backward()
    return nu.grad.cpu().numpy(), g.grad.cpu().numpy()


def test_backward_is_bit_identical_run_to_run(pair):
    rng = np.random.default_rng(5)
    wts = rng.uniform(0.5, 1.5, len(pair["centers"]))
    a = _backward(pair, wts)
    b = _backward(pair, wts)
    assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1])


def test_backward_matches_a_dense_float64_scatter(pair):
    """Every node and every voxel of both gradients against the float64 per-subset scatter, on eight
    overlapping subsets (the others get no upstream gradient)."""
    p = pair
    grad_C = np.zeros(len(p["centers"]))
    pick = [0, 1, 5, 6, 25, 26, 30, 31]  # neighbours along x, y and z: their windows overlap
    grad_C[pick] = np.random.default_rng(6).uniform(0.5, 1.5, len(pick))
    gu, gg = _backward(p, grad_C)
    ru, rg = backward_ref(p["f"], p["g"], p["centers"], HALF, p["node_u"], p["origin"], p["spacing"], grad_C)
    assert np.max(np.abs(gu - ru)) <= 2e-4 * np.max(np.abs(ru))
    assert np.max(np.abs(gg - rg)) <= 2e-4 * np.max(np.abs(rg))
    assert np.array_equal(gg != 0, rg != 0)  # the same voxels receive a gradient
