"""Reference moments of the fused 12-DOF kernel: the separable pass over the distinct node
coordinates (tuning flag 32) and the per-node pass, read back from the workspace and checked
against float64 sums over each subset computed here from their definition
(0 n, 1 sum f', 2 sum f'^2, 3 + 4c + j sum grad_c q_j, 15 + 4c + j sum grad_c q_j f', with
q = (X, Y, Z, 1) the voxel's offset from the node and f' = f - pilot). Nodes: a lattice, irregular
points, duplicates, nodes whose subsets leave the volume and invalid ones."""

from __future__ import annotations

import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

HALF = (5, 6, 4)  # hx, hy, hz
ROW_FLAG = 32


def _p(t):
    return None if t is None else ctypes.c_void_p(t.data_ptr())


@pytest.fixture(scope="module")
def case():
    import torch

    from pyaldvc_amd import kernels as K

    assert torch.cuda.is_available()
    nz, ny, nx = 40, 44, 52
    rng = np.random.default_rng(3)
    f_np = (rng.standard_normal((nz, ny, nx)) * 2.0 + 7.0).astype(np.float32)
    f = torch.from_numpy(f_np).cuda()
    gx, gy, gz = K.compute_gradients(f)
    hx, hy, hz = HALF
    lat = np.stack(np.meshgrid(np.arange(hx, nx - hx, 7), np.arange(hy, ny - hy, 5), np.arange(hz, nz - hz, 6),
                               indexing="ij"), -1).reshape(-1, 3)  # fmt: skip
    odd = np.column_stack([rng.integers(hx, nx - hx, 40), rng.integers(hy, ny - hy, 40), rng.integers(hz, nz - hz, 40)])
    edge = np.array([[hx - 1, 20, 20], [20, ny - hy, 20], [20, 20, 0], [nx - 1, ny - 1, nz - 1]])
    coords = np.concatenate([lat, odd, lat[:9], edge]).astype(np.int64)
    rng.shuffle(coords)
    N = coords.shape[0]
    valid = (rng.random(N) > 0.1).astype(np.uint8)
    meanf = f_np.mean() + rng.standard_normal(N) * 0.5
    meanf[:5] = np.nan  # no pilot
    # a reference mask of 4-voxel blocks (about 70 % kept) and keep rows for a fifth of the nodes
    blocks = rng.random((nz // 4 + 1, ny // 4 + 1, nx // 4 + 1)) < 0.7
    mask_np = np.repeat(np.repeat(np.repeat(blocks, 4, 0), 4, 1), 4, 2)[:nz, :ny, :nx].astype(np.uint8)
    S = (2 * hx + 1) * (2 * hy + 1) * (2 * hz + 1)
    rows = rng.random((N // 5, S)) < 0.6
    split_index = np.full(N, -1, np.int64)
    split_index[rng.choice(N, N // 5, replace=False)] = np.arange(N // 5)
    keep = np.packbits(rows, axis=1, bitorder="little")
    return dict(f=f, gx=gx, gy=gy, gz=gz, f_np=f_np, g_np=[t.cpu().numpy() for t in (gx, gy, gz)], coords=coords,
                valid=valid, meanf=meanf, mask_np=mask_np, mask=torch.from_numpy(mask_np).cuda(), rows=rows,
                split_index_np=split_index, split_index=torch.from_numpy(split_index).cuda(),
                split_keep=torch.from_numpy(keep).cuda())  # fmt: skip


def _moments(c, flags, extra, masked=False, split=False):
    """Runs aldvc_icgn12_tuned (one iteration) and returns the moments of every node (N, 27);
    ``masked``: with the case's reference mask, ``split``: with its keep rows."""
    import torch

    from pyaldvc_amd import _lib

    lib = _lib.lib()
    nz, ny, nx = c["f"].shape
    N = c["coords"].shape[0]
    hx, hy, hz = HALF
    cd = torch.from_numpy(c["coords"]).cuda()
    P = torch.zeros((N, 12), dtype=torch.float64, device="cuda")
    L = torch.eye(12, dtype=torch.float64).reshape(1, 144).repeat(N, 1).cuda()
    meanf = torch.from_numpy(c["meanf"]).cuda()
    valid = torch.from_numpy(c["valid"]).cuda()
    n_iter = torch.zeros(N, dtype=torch.int32, device="cuda")
    status = torch.zeros(N, dtype=torch.int8, device="cuda")
    zncc = torch.zeros(N, dtype=torch.float64, device="cuda")
    base = int(lib.aldvc_icgn12_workspace_bytes(N, hx, hy, hz, 1, 1))
    ws = torch.full((base + extra,), 0xFF, dtype=torch.uint8, device="cuda")
    stream = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    rc = lib.aldvc_icgn12_tuned(
        _p(cd), N, _p(P), hx, hy, hz, 1, _p(c["f"]), _p(c["gx"]), _p(c["gy"]), _p(c["gz"]),
        _p(c["mask"]) if masked else None, _p(c["f"]), nz, ny, nx, 1, _p(L), None, None, 0.0, _p(meanf), _p(valid),
        1e-2, 1e-3, 1, 5, 1, _p(c["split_index"]) if split else None, _p(c["split_keep"]) if split else None,
        int(c["split_keep"].shape[1]) if split else 0,
        _p(n_iter), _p(status), _p(zncc), 1, 0, flags, _p(ws), base + extra, stream,
    )  # fmt: skip
    _lib.check(rc, "aldvc_icgn12_tuned")
    torch.cuda.synchronize()
    list_bytes = (256 + 8 * N + 255) // 256 * 256
    mom = ws[list_bytes : list_bytes + 8 * 32 * N].cpu().numpy().view(np.float64).reshape(N, 32)
    return mom[:, :27].copy()


def _expected(c, masked=False, split=False):
    """float64 moments and the sums of the terms' magnitudes, per node whose subset is inside, over
    its kept voxels (its keep row, else the mask, else every voxel)."""
    f, gs = c["f_np"].astype(np.float64), [g.astype(np.float64) for g in c["g_np"]]
    hx, hy, hz = HALF
    nz, ny, nx = f.shape
    Z, Y, X = np.meshgrid(np.arange(-hz, hz + 1), np.arange(-hy, hy + 1), np.arange(-hx, hx + 1), indexing="ij")
    q = [X.astype(float), Y.astype(float), Z.astype(float), np.ones(X.shape)]
    out, mag, inside = {}, {}, []
    for n, (x0, y0, z0) in enumerate(c["coords"]):
        if not (c["valid"][n] and hx <= x0 < nx - hx and hy <= y0 < ny - hy and hz <= z0 < nz - hz):
            continue
        inside.append(n)
        sl = (slice(z0 - hz, z0 + hz + 1), slice(y0 - hy, y0 + hy + 1), slice(x0 - hx, x0 + hx + 1))
        pf = float(np.float32(c["meanf"][n])) if np.isfinite(c["meanf"][n]) else 0.0
        w = np.ones(X.shape)
        if split and c["split_index_np"][n] >= 0:
            w = c["rows"][c["split_index_np"][n]].reshape(X.shape).astype(float)
        elif masked:
            w = (c["mask_np"][sl] != 0).astype(float)
        fp = f[sl] - pf
        terms = [w, w * fp, w * fp * fp] + [None] * 24
        for cc in range(3):
            g = gs[cc][sl] * w
            for j in range(4):
                terms[3 + 4 * cc + j] = g * q[j]
                terms[15 + 4 * cc + j] = g * q[j] * fp
        out[n] = np.array([t.sum() for t in terms])
        mag[n] = np.array([np.abs(t).sum() for t in terms])
    return out, mag, inside


def _check(got, exp, mag, inside, rel):
    for n in inside:
        err = np.abs(got[n] - exp[n])
        assert np.all(err <= rel * mag[n] + 1e-9), (n, np.max(err / (mag[n] + 1e-30)))


def test_row_moments_match_definition(case, monkeypatch):
    from pyaldvc_amd import kernels as K
    from pyaldvc_amd import _lib

    monkeypatch.setattr(K, "ROW_MOMENTS_MIN_READS", 0)  # this node set is below the size that takes it
    exp, mag, inside = _expected(case)
    assert len(inside) > 100
    import torch

    nz, ny, nx = case["f"].shape
    cd = torch.from_numpy(case["coords"]).cuda()
    vd = torch.from_numpy(case["valid"]).cuda()
    extra = K.row_moments_bytes(cd, vd, *HALF, 1, (nz, ny, nx), True)
    sel = case["coords"][case["valid"] != 0]
    hx, hy, hz = HALF
    ins = sel[(sel[:, 0] >= hx) & (sel[:, 0] < nx - hx) & (sel[:, 1] >= hy) & (sel[:, 1] < ny - hy)
              & (sel[:, 2] >= hz) & (sel[:, 2] < nz - hz)]  # fmt: skip
    planes = int(ins[:, 2].max() - ins[:, 2].min()) + 2 * hz + 1
    assert extra == _lib.lib().aldvc_icgn12_row_moments_bytes(
        len(np.unique(ins[:, 0])), len(np.unique(ins[:, 1])), planes, ny, nx)  # fmt: skip
    rows = _moments(case, ROW_FLAG, extra)
    _check(rows, exp, mag, inside, 1e-6)
    node = _moments(case, 0, 0)
    _check(node, exp, mag, inside, 1e-5)
    # the flag with too little room for the rectangle sums: the per-node pass, bit for bit
    small = _moments(case, ROW_FLAG, 4096 + 64)
    assert np.array_equal(small[inside], node[inside])


def test_row_moments_solve_matches_per_node(monkeypatch):
    """The whole 12-DOF solve with the separable moments against the per-node ones, on a bead
    pair under strain and rotation: same statuses and iterations, parameters within the update
    tolerance."""
    import torch

    from pyaldvc_amd import kernels as K
    from pyaldvc_amd import synthetic as syn

    shape = (72, 80, 88)
    F = np.array([[0.015, 0.01, 0.0], [0.0, -0.01, 0.005], [0.004, 0.0, 0.012]])
    disp = syn.affine(F, tuple((s - 1) / 2 for s in shape[::-1]), (1.4, -0.6, 0.8))
    f_raw, g_raw = syn.bead_pair_np(shape, int(0.008 * np.prod(shape)), 1.5, disp, seed=5)
    f = torch.from_numpy(syn.normalize_np(f_raw)).cuda()
    g = torch.from_numpy(syn.normalize_np(g_raw)).cuda()
    gx, gy, gz = K.compute_gradients(f)
    coords = syn.node_lattice(shape, (16, 16, 16), (4, 4, 4))
    cd = torch.from_numpy(coords).cuda()
    nomask = np.ones((1, 1, 1), np.uint8)
    H, L, mf, bf, nv, valid = K.precompute_nodes(cd, 8, 8, 8, f, gx, gy, gz, nomask, 0.5, 1e12)
    u, v, w = disp(*(coords[:, i].astype(float) for i in range(3)))
    P0 = np.zeros((coords.shape[0], 12))
    P0[:, 9:] = np.round(np.column_stack([u, v, w]))
    P0d = torch.from_numpy(P0).cuda()

    def solve():
        out = K.icgn_12dof_parallel(cd, P0d, 8, 8, 8, f, gx, gy, gz, nomask, g, 1, L, mf, bf, valid, 1e-4, 1e-5,
                                    50, 5, variant="fused")  # fmt: skip
        return [t.cpu().numpy() for t in out]

    P1, it1, st1, z1 = solve()
    monkeypatch.setattr(K, "ROW_MOMENTS_MIN_READS", 0)
    assert K.row_moments_bytes(cd, valid, 8, 8, 8, 1, shape, True) > 0
    P2, it2, st2, z2 = solve()
    assert np.array_equal(st1, st2)
    assert np.array_equal(it1, it2)
    ok = st1 == 0
    assert ok.sum() > 0.5 * len(st1)
    # the two moment sums round differently (double rectangle sums against float per-thread sums):
    # the solves agree to well within the stopping tolerance of the update (1e-4)
    assert np.max(np.abs(P1[ok] - P2[ok])) < 1e-4
    assert np.median(np.abs(P1[ok, 9:] - P2[ok, 9:])) < 1e-7


@pytest.mark.parametrize("masked,split", [(True, False), (False, True), (True, True)])
def test_row_moments_mask_and_keep_rows(case, masked, split):
    """With a reference mask the separable sums are mask-weighted (and count the kept voxels);
    nodes with keep rows take the per-node pass in the same call."""
    import torch

    from pyaldvc_amd import _lib

    exp, mag, inside = _expected(case, masked, split)
    assert len(inside) > 100
    nz, ny, nx = case["f"].shape
    hx, hy, hz = HALF
    ins = case["coords"][inside]
    planes = int(ins[:, 2].max() - ins[:, 2].min()) + 2 * hz + 1
    extra = int(_lib.lib().aldvc_icgn12_row_moments_bytes(len(np.unique(ins[:, 0])), len(np.unique(ins[:, 1])),
                                                          planes, ny, nx))  # fmt: skip
    rows = _moments(case, ROW_FLAG, extra, masked, split)
    _check(rows, exp, mag, inside, 1e-6)
    node = _moments(case, 0, 0, masked, split)
    _check(node, exp, mag, inside, 1e-5)
    if split:  # the nodes with keep rows take the per-node pass either way: bit for bit
        kr = [n for n in inside if case["split_index_np"][n] >= 0]
        assert len(kr) > 10
        assert np.array_equal(rows[kr], node[kr])


def _precompute(c, flags, extra, masked, split):
    """aldvc_precompute_tuned on the case's nodes: (H, L, meanf, bottomf, n_valid, valid) on the host."""
    import torch

    from pyaldvc_amd import _lib

    lib = _lib.lib()
    nz, ny, nx = c["f"].shape
    N = c["coords"].shape[0]
    hx, hy, hz = HALF
    cd = torch.from_numpy(c["coords"]).cuda()
    H = torch.full((N, 144), 7.0, dtype=torch.float64, device="cuda")
    L = torch.full((N, 144), 7.0, dtype=torch.float64, device="cuda")
    mf = torch.full((N,), 7.0, dtype=torch.float64, device="cuda")
    bf = torch.full((N,), 7.0, dtype=torch.float64, device="cuda")
    nv = torch.full((N,), -7, dtype=torch.int64, device="cuda")
    va = torch.full((N,), 7, dtype=torch.uint8, device="cuda")
    ws = torch.full((max(extra, 1),), 0xFF, dtype=torch.uint8, device="cuda")
    rc = lib.aldvc_precompute_tuned(
        _p(cd), N, hx, hy, hz, 1, _p(c["f"]), _p(c["gx"]), _p(c["gy"]), _p(c["gz"]),
        _p(c["mask"]) if masked else None, nz, ny, nx, 0.5, 1e12,
        _p(c["split_index"]) if split else None, _p(c["split_keep"]) if split else None,
        int(c["split_keep"].shape[1]) if split else 0, _p(H), _p(L), _p(mf), _p(bf), _p(nv), _p(va),
        flags, _p(ws), extra, ctypes.c_void_p(torch.cuda.current_stream().cuda_stream),
    )  # fmt: skip
    _lib.check(rc, "aldvc_precompute_tuned")
    torch.cuda.synchronize()
    return [t.cpu().numpy() for t in (H, L, mf, bf, nv, va)]


@pytest.mark.parametrize("masked,split", [(False, False), (True, False), (False, True), (True, True)])
def test_separable_precompute(case, masked, split):
    """The precompute's Hessian, mean, norm, count, validity and Cholesky factor from the separable
    pass (flag 32) against float64 sums over each node's kept voxels computed here, and against the
    per-node pass; nodes with keep rows take the per-node pass in the same call (bit for bit)."""
    from pyaldvc_amd import _lib

    f, gs = case["f_np"].astype(np.float64), [g.astype(np.float64) for g in case["g_np"]]
    hx, hy, hz = HALF
    nz, ny, nx = f.shape
    Z, Y, X = np.meshgrid(np.arange(-hz, hz + 1), np.arange(-hy, hy + 1), np.arange(-hx, hx + 1), indexing="ij")
    q = [X.astype(float), Y.astype(float), Z.astype(float), np.ones(X.shape)]
    N = case["coords"].shape[0]
    keep_node = (case["split_index_np"] >= 0) if split else np.zeros(N, bool)
    inside = [n for n, (x0, y0, z0) in enumerate(case["coords"])
              if hx <= x0 < nx - hx and hy <= y0 < ny - hy and hz <= z0 < nz - hz]  # fmt: skip
    sep = [n for n in inside if not keep_node[n]]
    ins = case["coords"][sep]
    planes = int(ins[:, 2].max() - ins[:, 2].min()) + 2 * hz + 1
    extra = int(_lib.lib().aldvc_precompute_row_bytes(len(np.unique(ins[:, 0])), len(np.unique(ins[:, 1])), planes,
                                                      ny, nx))  # fmt: skip
    rows = _precompute(case, 32, extra, masked, split)
    node = _precompute(case, 0, 0, masked, split)
    for n in inside:
        x0, y0, z0 = case["coords"][n]
        sl = (slice(z0 - hz, z0 + hz + 1), slice(y0 - hy, y0 + hy + 1), slice(x0 - hx, x0 + hx + 1))
        w = np.ones(X.shape)
        if keep_node[n]:
            w = case["rows"][case["split_index_np"][n]].reshape(X.shape).astype(float)
        elif masked:
            w = (case["mask_np"][sl] != 0).astype(float)
        g = [gs[c][sl] for c in range(3)]
        sd = [g[c] * q[j] for c in range(3) for j in range(3)] + [g[0], g[1], g[2]]
        Hd = np.array([[np.sum(w * a * b) for b in sd] for a in sd])
        Hm = np.array([[np.sum(np.abs(w * a * b)) for b in sd] for a in sd])
        cnt = int(w.sum())
        for got, rel in ((rows, 1e-6), (node, 1e-5)):
            Hg, Lg, mfg, bfg, nvg, vag = (a[n] for a in got)
            assert nvg == cnt
            assert np.all(np.abs(Hg.reshape(12, 12) - Hd) <= rel * Hm + 1e-9), n
            if cnt >= 27 and cnt >= 0.5 * X.size:
                mfd = np.sum(w * f[sl]) / cnt
                bfd = np.sqrt(max(np.sum(w * f[sl] ** 2) - cnt * mfd * mfd, 1e-30))
                assert abs(mfg - mfd) <= 1e-6 * abs(mfd) + 1e-9
                assert abs(bfg - bfd) <= 1e-5 * bfd
        assert rows[5][n] == node[5][n], n
        if rows[5][n]:
            Ld = np.linalg.cholesky(Hd)
            assert np.allclose(rows[1][n].reshape(12, 12), Ld, rtol=1e-5, atol=1e-5 * np.abs(Ld).max())
        if keep_node[n]:
            assert all(np.array_equal(a[n], b[n]) for a, b in zip(rows, node)), n
    assert sum(1 for n in inside if rows[5][n]) > 50
