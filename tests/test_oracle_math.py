"""The oracle against what the mathematics fixes, independently of the golden vectors.

Each check is chosen so that a plausible slip in the restatement (a dropped term, a wrong sign
or index, a transposed operand) fails it: library routines that compute the same quantity
(scipy map_coordinates, correlate1d, ndimage.label), closed forms (cubic reproduction of
quadratics, exact quadratic peaks, warp-composition identity), and brute force on tiny inputs.
"""

from __future__ import annotations

import numpy as np
import pytest
from scipy import ndimage

from oracle import cpu_ref, ncc_ref

RNG = np.random.default_rng(1234)


def _smooth_volume(shape=(20, 22, 24), seed=0):
    rng = np.random.default_rng(seed)
    return ndimage.gaussian_filter(rng.normal(size=shape), 1.5).astype(np.float32)


def test_linear_and_bspline_equal_scipy():
    vol = _smooth_volume()
    nz, ny, nx = vol.shape
    zs, ys, xs = RNG.uniform(1, nz - 3, 300), RNG.uniform(1, ny - 3, 300), RNG.uniform(1, nx - 3, 300)
    lin = cpu_ref.sample_points(vol, zs, ys, xs, 0)
    ref = ndimage.map_coordinates(vol.astype(np.float64), np.vstack([zs, ys, xs]), order=1, mode="nearest")
    assert np.max(np.abs(lin - ref)) < 1e-12
    coef = ndimage.spline_filter(vol, order=3, output=np.float32, mode="mirror")
    bs = cpu_ref.sample_points(coef, zs, ys, xs, 2)
    ref3 = ndimage.map_coordinates(vol.astype(np.float64), np.vstack([zs, ys, xs]), order=3, mode="mirror")
    assert np.max(np.abs(bs - ref3)) < 2e-5


def test_cubic_reproduces_quadratics_and_integers():
    z, y, x = np.mgrid[0:20, 0:22, 0:24].astype(np.float64)
    q = 0.1 * x**2 - 0.2 * y * z + 0.3 * y + 0.05 * z**2 + 1.0
    vol = q.astype(np.float32)
    zs, ys, xs = RNG.uniform(2, 17, 200), RNG.uniform(2, 19, 200), RNG.uniform(2, 21, 200)
    exact = 0.1 * xs**2 - 0.2 * ys * zs + 0.3 * ys + 0.05 * zs**2 + 1.0
    got = cpu_ref.sample_points(vol, zs, ys, xs, 1)
    assert np.allclose(got, exact, atol=1e-3)  # float32 storage limits the accuracy
    # interpolating schemes return the voxel value at integer positions
    zi, yi, xi = RNG.integers(1, 17, 50), RNG.integers(1, 19, 50), RNG.integers(1, 21, 50)
    for mode in (0, 1):
        v = cpu_ref.sample_points(vol, zi.astype(float), yi.astype(float), xi.astype(float), mode)
        assert np.array_equal(v, vol[zi, yi, xi].astype(np.float64))
    # the domain is 1 <= c <= n - 3 on every axis
    out = cpu_ref.sample_points(vol, np.array([0.99, 5.0, 17.0, 17.01]), np.array([5.0] * 4), np.array([5.0] * 4), 1)
    assert np.isnan(out[0]) and np.isfinite(out[1]) and np.isfinite(out[2]) and np.isnan(out[3])


def test_gradients_equal_correlate1d():
    f = _smooth_volume((18, 20, 22), seed=2)
    gx, gy, gz = cpu_ref.compute_gradients(f)
    st = np.array([-1 / 60, 3 / 20, -3 / 4, 0.0, 3 / 4, -3 / 20, 1 / 60])
    for g, ax in ((gx, 2), (gy, 1), (gz, 0)):
        ref = ndimage.correlate1d(f.astype(np.float64), st, axis=ax, mode="nearest")
        inner = (slice(3, -3),) * 3
        assert np.max(np.abs(g[inner] - ref[inner])) < 1e-6
        b = np.ones(f.shape, bool)
        b[inner] = False
        assert np.all(g[b] == 0.0)
    # exact for a cubic polynomial along x
    x = np.arange(22, dtype=np.float64)
    vol = np.broadcast_to((0.01 * x**3 - 0.2 * x**2 + x)[None, None, :], (18, 20, 22)).astype(np.float64)
    gx, _, _ = cpu_ref.compute_gradients(vol.astype(np.float32))
    d = 0.03 * x**2 - 0.4 * x + 1.0
    assert np.allclose(gx[5, 5, 3:-3], d[3:-3], atol=2e-4)


def _W(p):
    M = np.eye(4)
    M[:3, :3] = np.eye(3) + p[:9].reshape(3, 3)
    M[:3, 3] = p[9:]
    return M


def test_compose_warp_identity():
    for _ in range(10):
        P = RNG.normal(0, 0.05, 12)
        dP = RNG.normal(0, 0.02, 12)
        P1 = P.copy()
        assert cpu_ref.compose_warp_inplace(P1, dP)
        assert np.allclose(_W(P1) @ _W(dP), _W(P), atol=1e-12)
    sing = np.zeros(12)
    sing[0] = -1.0  # I + dF singular
    P = np.zeros(12)
    assert not cpu_ref.compose_warp_inplace(P, sing)


def test_cholesky12_equals_numpy():
    A = RNG.normal(size=(12, 12))
    H = A @ A.T + np.eye(12)
    ok, L = cpu_ref.cholesky12(H)
    assert ok and np.allclose(L, np.linalg.cholesky(H), atol=1e-12)
    H[4, 4] = -1.0
    ok, _ = cpu_ref.cholesky12(H)
    assert not ok


def test_precompute_equals_brute_force():
    f = _smooth_volume((24, 24, 24), seed=4)
    gx, gy, gz = cpu_ref.compute_gradients(f)
    mask = np.ones(f.shape, np.uint8)
    mask[:, :, 14:] = 0
    coords = np.array([[12, 11, 12], [10, 12, 13]], dtype=np.int64)
    half, stride = (5, 4, 3), 2
    H, L, mf, bf, nv, valid = cpu_ref.precompute_nodes(coords, *half, f, gx, gy, gz, mask, 0.2, 1e12, stride)
    for n in range(2):
        x0, y0, z0 = coords[n]
        dz, dy, dx = np.meshgrid(*(np.arange(-h, h + 1, stride) for h in half[::-1]), indexing="ij")
        dz, dy, dx = dz.ravel(), dy.ravel(), dx.ravel()
        zz, yy, xx = z0 + dz, y0 + dy, x0 + dx
        m = mask[zz, yy, xx] > 0
        g3 = [a[zz, yy, xx].astype(np.float64)[m] for a in (gx, gy, gz)]
        X, Y, Z = dx[m].astype(float), dy[m].astype(float), dz[m].astype(float)
        SD = np.column_stack([g3[0] * X, g3[0] * Y, g3[0] * Z, g3[1] * X, g3[1] * Y, g3[1] * Z,
                              g3[2] * X, g3[2] * Y, g3[2] * Z, g3[0], g3[1], g3[2]])  # fmt: skip
        fv = f[zz, yy, xx].astype(np.float64)[m]
        assert nv[n] == m.sum()
        assert np.allclose(H[n], SD.T @ SD, rtol=1e-10, atol=1e-10)
        assert np.isclose(mf[n], fv.mean()) and np.isclose(bf[n], np.sqrt(np.sum((fv - fv.mean()) ** 2)))
        assert valid[n] and np.allclose(L[n] @ L[n].T, H[n], rtol=1e-9, atol=1e-9)


def test_icgn12_recovers_exact_affine_warp():
    """An analytic volume warped exactly: IC-GN converges to the imposed affine map."""
    shape = (40, 40, 40)
    z, y, x = np.mgrid[0 : shape[0], 0 : shape[1], 0 : shape[2]].astype(np.float64)

    def field(x, y, z):
        return np.sin(0.35 * x) * np.cos(0.3 * y) + np.sin(0.28 * z + 0.2 * x) + 0.5 * np.cos(0.33 * y - 0.25 * z)

    F = np.array([[0.01, 0.005, 0.0], [-0.004, 0.008, 0.002], [0.0, 0.003, -0.006]])
    t = np.array([0.4, -0.3, 0.25])
    c0 = np.array([19.5, 19.5, 19.5])
    f = field(x, y, z).astype(np.float32)
    # g(x_def) = f(X) with x_def = X + t + F (X - c0): g(x) = f(A^-1 (x - c0 - t) + c0), A = I + F
    Ainv = np.linalg.inv(np.eye(3) + F)
    pts = np.stack([x, y, z], axis=-1) - c0 - t
    Xs = pts @ Ainv.T + c0
    g = field(Xs[..., 0], Xs[..., 1], Xs[..., 2]).astype(np.float32)
    gx, gy, gz = cpu_ref.compute_gradients(f)
    coords = np.array([[20, 20, 20], [17, 22, 19], [22, 18, 21]], dtype=np.int64)
    nomask = np.ones((1, 1, 1), np.uint8)
    H, L, mf, bf, nv, valid = cpu_ref.precompute_nodes(coords, 7, 7, 7, f, gx, gy, gz, nomask, 0.5, 1e12)
    U = t + (coords - c0) @ F.T
    P0 = np.zeros((3, 12))
    P0[:, 9:] = np.round(U)
    P, it, st, zc = cpu_ref.icgn_12dof_parallel(coords, P0, 7, 7, 7, f, gx, gy, gz, nomask, g, 1, L, mf, bf, valid, 1e-4, 1e-6, 100, 5)
    assert np.all(st == 0)
    assert np.max(np.abs(P[:, 9:] - U)) < 5e-3
    assert np.max(np.abs(P[:, :9] - F.ravel())) < 2e-3
    assert np.all(zc > 0.999)


def test_zncc_direct_equals_brute_force():
    f = _smooth_volume((20, 20, 20), seed=6)
    g = _smooth_volume((20, 20, 20), seed=7) + 0.5 * f
    coords = np.array([[9, 10, 8], [10, 9, 11]], dtype=np.int64)
    hx, hy, hz, rx, ry, rz = 3, 2, 3, 2, 3, 1
    wx0 = coords[:, 0] - hx - rx + np.array([1, 0])
    wy0 = coords[:, 1] - hy - ry
    wz0 = coords[:, 2] - hz - rz + np.array([0, 1])
    out = np.empty((2, 2 * rz + 1, 2 * ry + 1, 2 * rx + 1))
    cpu_ref.zncc_direct(f, g, coords, hx, hy, hz, rx, ry, rz, wx0, wy0, wz0, np.ones(2, bool), out)
    for n in range(2):
        x0, y0, z0 = coords[n]
        T = f[z0 - hz : z0 + hz + 1, y0 - hy : y0 + hy + 1, x0 - hx : x0 + hx + 1].astype(np.float64).ravel()
        T = T - T.mean()
        for a in range(2 * rz + 1):
            for b in range(2 * ry + 1):
                for c in range(2 * rx + 1):
                    z, y, x = wz0[n] + a, wy0[n] + b, wx0[n] + c
                    W = g[z : z + 2 * hz + 1, y : y + 2 * hy + 1, x : x + 2 * hx + 1].astype(np.float64).ravel()
                    ref = np.dot(T, W - W.mean()) / np.sqrt(np.sum(T * T) * np.sum((W - W.mean()) ** 2))
                    assert abs(out[n, a, b, c] - ref) < 1e-9


def test_subvoxel_peak_of_exact_quadratic():
    """A 27-point neighbourhood sampled from a concave quadratic gives its exact vertex."""
    for _ in range(20):
        d = RNG.uniform(-0.45, 0.45, 3)
        Hq = -np.diag(RNG.uniform(0.5, 2.0, 3))
        Hq[0, 1] = Hq[1, 0] = 0.1
        off = np.stack(np.meshgrid([-1, 0, 1], [-1, 0, 1], [-1, 0, 1], indexing="ij"), -1).reshape(27, 3)
        p = off[:, ::-1].astype(float) - d  # (x, y, z) offsets from the vertex
        vals = 0.5 * np.einsum("ni,ij,nj->n", p, Hq, p) + 3.0
        sub = ncc_ref.subvoxel_peak(vals[None, :])
        assert np.allclose(sub[0], d, atol=1e-10)


def test_flood_fill_rows_equal_ndimage_label():
    mask = np.ones((30, 30, 30), np.uint8)
    mask[:, 14, :] = 0
    mask[10:13, 18:22, 5:25] = 0
    coords = np.array([[15, 12, 15], [15, 17, 12], [8, 20, 11]], dtype=np.int64)
    h = 5
    rows, nk, nm = cpu_ref.build_split_rows(coords, np.arange(3), h, h, h, 1, mask, 0)
    for m in range(3):
        x0, y0, z0 = coords[m]
        sub = mask[z0 - h : z0 + h + 1, y0 - h : y0 + h + 1, x0 - h : x0 + h + 1] > 0
        lab, _ = ndimage.label(sub, structure=ndimage.generate_binary_structure(3, 1))
        comp = lab == lab[h, h, h] if lab[h, h, h] else np.zeros_like(sub)
        bits = np.unpackbits(rows[m], bitorder="little", count=sub.size).astype(bool)
        assert np.array_equal(bits, comp.ravel())
        assert nk[m] == comp.sum() and nm[m] == sub.sum()


@pytest.mark.parametrize("stride", [1, 2])
def test_icgn_noise_free_translation_converges_in_one_iteration_when_exact(stride):
    """Start at the exact integer translation of an integer-shifted pair: the first residual is
    zero, so the node converges at iteration 1 (norm_abs < ABS_TOL) with ZNCC 1."""
    f = _smooth_volume((30, 30, 30), seed=9)
    g = np.roll(f, shift=(1, -2, 2), axis=(0, 1, 2))
    gx, gy, gz = cpu_ref.compute_gradients(f)
    coords = np.array([[15, 15, 15]], dtype=np.int64)
    nomask = np.ones((1, 1, 1), np.uint8)
    H, L, mf, bf, nv, valid = cpu_ref.precompute_nodes(coords, 5, 5, 5, f, gx, gy, gz, nomask, 0.5, 1e12, stride)
    P0 = np.zeros((1, 12))
    P0[0, 9:] = [2, -2, 1]
    P, it, st, zc = cpu_ref.icgn_12dof_parallel(coords, P0, 5, 5, 5, f, gx, gy, gz, nomask, g, 1, L, mf, bf, valid, 1e-2, 1e-3, 100, 5, stride)
    assert st[0] == 0 and it[0] == 1 and abs(zc[0] - 1.0) < 1e-12
    assert np.array_equal(P[0], P0[0])
